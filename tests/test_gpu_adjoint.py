"""Discrete adjoint of the SIMPLE iterations (semo_adjoint_alpha_p).

The reference has no adjoint, so nothing in it pins this path.  It is pinned
by the mathematics instead:
  * the adjoint's forward part is the forward path (same J, same final
    fields, bit for bit), and that forward with the Darcy term agrees with
    the oracle's (the oracle carries the same alpha_p term);
  * dJ/dalpha_p agrees with central finite differences of the forward map at
    32 seeded cells, along a random direction with central finite
    differences of the oracle's own forward map (CPU), and, on a 4^3 mesh,
    in every component with central differences of the oracle's forward.
Case: 8^3 jittered lid-driven cavity, rho = 1, mu = 0.01, dt = 0.05,
alpha_p ~ U[0, 10) seeded, 20 outer iterations of one time step,
J = int p dA (lid, zmax) - int p dA (bottom, zmin).  Solves run to 1e-14 so
that the finite differences see the discrete map, not solver noise.
Tolerance: 1e-5 relative (per cell, against max(|fd|, 1e-2 max |fd|)).
Step: h = 1e-4 for the per-cell differences.  With h = 1e-6 the iterative
solves' residual (relative 1e-14) already moves J by about 3e-17, which is
1e-4 of a per-cell difference quotient at this size; the difference quotient
itself is flat in h from 1e-3 to 1e-5 (central differences, O(h^2)).
"""
import numpy as np
import pytest

from semo_mesh_amd.cases import cavity_case
from semo_mesh_amd.mesh import cavity_mesh

pytestmark = pytest.mark.gpu

N_ITER = 20
W = {"zmax": 1.0, "zmin": -1.0}
H = 1e-4


def _case():
    c = cavity_case(dt=0.05)
    c.rho = (1.0, 1.0)
    c.mu = (0.01, 0.01)
    return c


@pytest.fixture(scope="module")
def setup():
    mesh = cavity_mesh(8, jitter=0.1, seed=11)
    alpha = 10.0 * np.random.default_rng(5).random(mesh.n_cells)
    return mesh, _case(), alpha


def _gpu(mesh, case, alpha, reorder=True):
    from semo_mesh_amd.solver import GpuSolver
    g = GpuSolver(0)
    g.upload_mesh(mesh, reorder=reorder)
    g.set_case(case)
    g.set_solver("fp64", tol_p=1e-14, tol_u=1e-14, maxit_p=300, maxit_u=1000, coarse_size=1000)
    g.set_field("alphaP", alpha)
    return g


def _J_gpu(mesh, case, alpha, reorder=True):
    g = _gpu(mesh, case, alpha, reorder)
    g.incompressiblePressureBased(N_ITER)
    j = g.objective(W)
    g.close()
    return j


def _J_oracle(oracle_lib, mesh, case, alpha):
    o = oracle_lib.Oracle(mesh, case)
    o.set_solver(tol=1e-14)
    o.set("alphaP", alpha)
    o.step(correctors=N_ITER)
    P = o.get("P")
    area = o.get("face_area")
    j = 0.0
    for name, w in W.items():
        b = mesh.patch(name)
        s, nf = int(mesh.patch_start[b]), int(mesh.patch_nfaces[b])
        f = np.arange(s, s + nf)
        j += w * float(np.sum(area[f] * P[mesh.owner[f]]))
    return j


def test_adjoint_forward_is_the_forward(setup):
    mesh, case, alpha = setup
    g = _gpu(mesh, case, alpha)
    J, grad = g.adjointAlphaP(N_ITER, W)
    fa = {k: g.get_field(k) for k in ("P", "U", "V", "W")}
    g.close()
    h = _gpu(mesh, case, alpha)
    h.incompressiblePressureBased(N_ITER)
    assert h.objective(W) == J
    for k in fa:
        assert np.array_equal(h.get_field(k), fa[k]), k
    h.close()
    assert np.all(np.isfinite(grad)) and np.abs(grad).max() > 0


def test_darcy_forward_matches_oracle(setup, oracle_lib):
    mesh, case, alpha = setup
    jg = _J_gpu(mesh, case, alpha, reorder=False)
    jo = _J_oracle(oracle_lib, mesh, case, alpha)
    assert abs(jg - jo) <= 1e-9 * abs(jo)


@pytest.mark.parametrize("reorder", [True, False])
def test_adjoint_vs_central_fd(setup, reorder):
    mesh, case, alpha = setup
    g = _gpu(mesh, case, alpha, reorder)
    _, grad = g.adjointAlphaP(N_ITER, W)
    g.close()
    cells = np.random.default_rng(17).choice(mesh.n_cells, 32, replace=False)
    fd = np.empty(len(cells))
    for q, c in enumerate(cells):
        e = np.zeros_like(alpha)
        e[c] = H
        fd[q] = (_J_gpu(mesh, case, alpha + e, reorder) - _J_gpu(mesh, case, alpha - e, reorder)) / (2 * H)
    scale = np.maximum(np.abs(fd), 1e-2 * np.abs(fd).max())
    err = np.abs(grad[cells] - fd) / scale
    print(f"adjoint vs FD (h={H}): max rel err {err.max():.3e}, max |dJ/dalpha| {np.abs(fd).max():.3e}")
    assert err.max() <= 1e-5


def test_adjoint_directional_vs_oracle_fd(setup, oracle_lib):
    mesh, case, alpha = setup
    g = _gpu(mesh, case, alpha)
    _, grad = g.adjointAlphaP(N_ITER, W)
    g.close()
    v = np.random.default_rng(23).standard_normal(mesh.n_cells)
    h = 1e-5
    fd = (_J_oracle(oracle_lib, mesh, case, alpha + h * v) - _J_oracle(oracle_lib, mesh, case, alpha - h * v)) / (2 * h)
    rel = abs(grad @ v - fd) / abs(fd)
    print(f"directional derivative vs oracle FD: rel err {rel:.3e}")
    assert rel <= 1e-5


def test_adjoint_full_gradient_vs_oracle_fd(oracle_lib):
    """Every component of dJ/dalpha_p against central differences of the
    oracle's forward (CPU restatement, ILU0-BiCGStab solves at 1e-14) on a
    4^3 jittered cavity: 64 cells, 128 oracle runs of 20 iterations."""
    mesh = cavity_mesh(4, jitter=0.1, seed=29)
    case = _case()
    alpha = 10.0 * np.random.default_rng(31).random(mesh.n_cells)
    g = _gpu(mesh, case, alpha, reorder=False)
    J, grad = g.adjointAlphaP(N_ITER, W)
    g.close()
    assert abs(J - _J_oracle(oracle_lib, mesh, case, alpha)) <= 1e-9 * abs(J)
    fd = np.empty(mesh.n_cells)
    for c in range(mesh.n_cells):
        e = np.zeros_like(alpha)
        e[c] = H
        fd[c] = (_J_oracle(oracle_lib, mesh, case, alpha + e) - _J_oracle(oracle_lib, mesh, case, alpha - e)) / (2 * H)
    scale = np.maximum(np.abs(fd), 1e-2 * np.abs(fd).max())
    err = np.abs(grad - fd) / scale
    print(f"adjoint vs oracle FD, all {mesh.n_cells} cells: max rel err {err.max():.3e}")
    assert err.max() <= 1e-5


def test_adjoint_rejects_what_it_does_not_differentiate(setup):
    from semo_mesh_amd._lib import SemoError
    mesh, case, alpha = setup
    c2 = _case()
    c2.sigma = 0.07
    g = _gpu(mesh, c2, alpha)
    with pytest.raises(SemoError):
        g.adjointAlphaP(2, W)
    g.close()


@pytest.mark.parametrize("n", [8, 16])
@pytest.mark.parametrize("physics", ["leg", "cavity"])
def test_adjoint_vs_cpu_adjoint(oracle_lib, n, physics):
    """The device adjoint against the CPU adjoint (the oracle's assembly and
    solves differentiated in reverse mode by operator overloading,
    oracle/liborc_ad.so) on the whole gradient, 20 outer iterations, at the
    adjoint leg's physics (rho 1, mu 0.01, dt 0.05) and the cavity
    configuration's (rho 1000, mu 1e-3, dt 0.01).  Both solve to 1e-14;
    relative L2 bar 1e-10."""
    mesh = cavity_mesh(n, jitter=0.1, seed=20261205)
    case = _case() if physics == "leg" else cavity_case()
    alpha = 10.0 * np.random.default_rng(20261205).random(mesh.n_cells)
    g = _gpu(mesh, case, alpha, reorder=True)
    J, grad = g.adjointAlphaP(N_ITER, W)
    g.close()
    o = oracle_lib.AdjointOracle(mesh, case)
    o.set_solver(1e-14, 20000)
    o.set("alphaP", alpha)
    Jc, gc = o.adjoint(N_ITER, W)
    rel = np.linalg.norm(grad - gc) / np.linalg.norm(gc)
    print(f"{physics} {n}^3: J {J:.6e} (CPU {Jc:.6e}), gradient rel L2 {rel:.3e}, "
          f"max |dJ/dalpha| {np.abs(gc).max():.3e}")
    assert abs(J - Jc) <= 1e-10 * abs(Jc)
    assert rel <= 1e-10


def test_darcy_term_definition():
    """The Darcy term in closed form, independent of the oracle: on one
    state, the momentum system assembled with alpha_p = a differs from the
    one with alpha_p = 0 by a_i V_i on the diagonal and -a_i V_i U_i
    (and V, W) on the right sides, and nowhere else."""
    from semo_mesh_amd.solver import GpuSolver
    mesh = cavity_mesh(6, jitter=0.1, seed=41)
    case = _case()
    rng = np.random.default_rng(43)
    n = mesh.n_cells
    state = {f: rng.standard_normal(n) for f in ("U", "V", "W", "P")}
    a = 10.0 * rng.random(n)
    out = []
    for alpha in (np.zeros(n), a):
        g = GpuSolver(0)
        g.upload_mesh(mesh, reorder=True)
        g.set_case(case)
        for f, v in state.items():
            g.set_field(f, v)
        g.set_field("Rho", np.full(n, case.rho[0]))     # as the first step's EOS would
        g.set_field("mu", np.full(n, case.mu[0]))
        g.set_field("alphaP", alpha)
        g.assemble("momentum")
        out.append(g.matrix("momentum"))
        vol = g.get_field("cell_volume")
        g.close()
    (r0, c0, v0, b0), (r1, c1, v1, b1) = out
    assert np.array_equal(r0, r1) and np.array_equal(c0, c1)
    # the diagonal comes first (device order, caller's cell numbers)
    assert np.array_equal(r0[:n], c0[:n]) and np.array_equal(np.sort(r0[:n]), np.arange(n))
    d0 = np.empty(n); d1 = np.empty(n)
    d0[r0[:n]] = v0[:n]; d1[r0[:n]] = v1[:n]
    assert np.allclose(d1 - d0, a * vol, rtol=1e-12, atol=1e-12 * np.abs(d0).max())
    assert np.array_equal(v0[n:], v1[n:])
    for k, f in enumerate(("U", "V", "W")):
        assert np.allclose(b1[k] - b0[k], -a * vol * state[f], rtol=1e-9, atol=1e-12 * np.abs(b0[k]).max())
