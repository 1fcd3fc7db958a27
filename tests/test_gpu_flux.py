"""semo_flux on the GPU against the compiled reference's flux schemes.

tests/golden/flux_faces.npz holds SEMO_Solvers_Builder::calcFluxHAUS,
SLAU2, Roe, APRoe, RoeAMTP, Upwind, AUSMPWP_N and RoeM_N (solvers/flux.cpp)
on 400 face states (gas- and liquid-like sides, Mach 0 to ~3, identical
sides, sides at rest, nearly equal velocities, supersonic sides).  The
device's fluxes agree with them to 1e-12 of each component's largest
magnitude (the device's cos / pow and glibc's may differ in the last
place); the schemes without transcendental functions (SLAU2, Roe, Upwind)
to 1e-15.  Unknown schemes, a missing parameter array and nsp out of range
are refused."""
import ctypes as C

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

NAMES = {"HAUS": "haus", "SLAU2": "slau2", "Roe": "roe", "APRoe": "aproe", "RoeAMTP": "roeamtp",
         "Upwind": "upwind", "AUSMPWP_N": "ausmpwp_n", "RoeM_N": "roem_n"}


def _data():
    import os
    d = np.load(os.path.join(GOLDEN, "flux_faces.npz"))
    st, nsp = d["flux_states"], int(d["flux_nsp"])
    w = 8 + nsp
    return d, st[:, :w], st[:, w:2 * w], st[:, 2 * w:2 * w + 3], st[:, 2 * w + 3:]


def test_flux_schemes_vs_reference():
    from semo_mesh_amd.solver import GpuSolver
    d, L, R, nv, par = _data()
    g = GpuSolver(0)
    try:
        for scheme, key in NAMES.items():
            F = g.calcFlux(scheme, L, R, nv, par)
            ref = d["flux_" + key]
            err = (np.abs(F - ref) / np.abs(ref).max(axis=0)).max()
            print(scheme, err)
            assert err <= (1e-15 if scheme in ("SLAU2", "Roe", "Upwind") else 1e-12), (scheme, err)
    finally:
        g.close()


def test_flux_errors():
    from semo_mesh_amd.solver import GpuSolver, SemoError
    d, L, R, nv, par = _data()
    g = GpuSolver(0)
    try:
        with pytest.raises(SemoError):
            g.calcFlux("RoeM_N", L, R, nv, None)
        with pytest.raises(KeyError):
            g.calcFlux("Rusanov", L, R, nv, par)
        out = np.empty((L.shape[0], 16))
        p = lambda a: np.ascontiguousarray(a).ctypes.data_as(C.c_void_p)  # noqa: E731
        assert g._L.semo_flux(g._h, 9, L.shape[0], 2, p(L), p(R), p(nv), p(par), p(out)) != 0
        assert g._L.semo_flux(g._h, 1, L.shape[0], 0, p(L), p(R), p(nv), p(par), p(out)) != 0
        assert g.calcFlux("Upwind", L[:0], R[:0], nv[:0]).shape == (0, 6)
    finally:
        g.close()
