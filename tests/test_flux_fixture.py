"""The convective-flux fixture (tests/golden/flux_faces.npz), CPU only.

The fixture holds SEMO_Solvers_Builder::calcFluxHAUS, SLAU2, Roe, APRoe,
RoeAMTP, Upwind, AUSMPWP_N and RoeM_N (solvers/flux.cpp) as the compiled
reference computed them on 400 face states (tests/golden/make_golden.py
flux_states).  What the mathematics fixes for any consistent numerical flux
is checked on it here: with identical sides every scheme gives the physical
flux (rho Un, rho Un u + P n, ..., rho Un Ht, rho Un Y) to 1e-13; the
central-only Roe flux is the mean of the two sides' physical fluxes
exactly; and with both sides at rest the upwind and central fluxes carry no
mass or energy (the split schemes' pressure-diffusion terms do move mass
across a resting pressure jump, by design)."""
import numpy as np
import pytest


SCHEMES = ("haus", "slau2", "roe", "aproe", "roeamtp", "upwind", "ausmpwp_n", "roem_n")


def _phys(s, nv, nsp):
    P, U, V, W, T, rho, C, Ht = (s[:, k] for k in range(8))
    Y = s[:, 8:8 + nsp]
    Un = U * nv[:, 0] + V * nv[:, 1] + W * nv[:, 2]
    cols = [rho * Un, rho * Un * U + P * nv[:, 0], rho * Un * V + P * nv[:, 1], rho * Un * W + P * nv[:, 2],
            rho * Un * Ht] + [rho * Un * Y[:, i] for i in range(nsp - 1)]
    return np.column_stack(cols)


def _split(d):
    st, nsp = d["flux_states"], int(d["flux_nsp"])
    w = 8 + nsp
    return st[:, :w], st[:, w:2 * w], st[:, 2 * w:2 * w + 3], st[:, 2 * w + 3:], nsp


@pytest.mark.parametrize("scheme", SCHEMES)
def test_consistency(scheme):
    d = np.load(__import__("os").path.join(__import__("conftest").GOLDEN, "flux_faces.npz"))
    L, R, nv, _, nsp = _split(d)
    m = L.shape[0] - 40
    F = d["flux_" + scheme][m:m + 10]
    ph = _phys(L[m:m + 10], nv[m:m + 10], nsp)
    assert np.abs(F - ph).max() <= 1e-13 * np.abs(ph).max()


def test_roe_is_the_central_flux():
    d = np.load(__import__("os").path.join(__import__("conftest").GOLDEN, "flux_faces.npz"))
    L, R, nv, _, nsp = _split(d)
    c = 0.5 * (_phys(L, nv, nsp) + _phys(R, nv, nsp))
    assert np.abs(d["flux_roe"] - c).max() <= 1e-15 * np.abs(c).max()


@pytest.mark.parametrize("scheme", ["upwind", "roe"])
def test_rest_states_carry_no_mass(scheme):
    d = np.load(__import__("os").path.join(__import__("conftest").GOLDEN, "flux_faces.npz"))
    L, R, nv, _, nsp = _split(d)
    m = L.shape[0] - 30
    F = d["flux_" + scheme][m:m + 10]
    assert np.abs(F[:, 0]).max() == 0.0 and np.abs(F[:, 4]).max() == 0.0
    assert np.abs(F[:, 1:4]).max() > 0.0
