"""Block-coupled (BSR4) pressure-velocity system (semo_calc_coupled_eq).

The reference's calcCoupledEq is commented out (src/solvers/eqCoupled.cpp),
so no reference output exists: parity is unpinned by construction.  The
path is pinned by invariants instead:
  * every block equals its definition from the scalar momentum and pressure
    systems (themselves pinned to the reference) and the face geometry,
    rebuilt here independently in scipy (relative 1e-14);
  * the BSR4 product equals the product with the scalar-expanded CSR matrix;
  * the FP64 BiCGStab + FP32 block-AMG solve reaches its residual, measured
    independently in FP64 (scipy);
  * at convergence the coupled and the segregated (SIMPLE) iterations reach
    the same discrete solution of one time step.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from semo_mesh_amd.cases import cavity_case
from semo_mesh_amd.mesh import cavity_mesh

pytestmark = pytest.mark.gpu


def _case(urf=0.3):
    c = cavity_case(dt=0.05)
    c.rho = (1.0, 1.0)
    c.mu = (0.01, 0.01)
    c.urf_mom = c.urf_p = c.urf_pu = urf
    return c


def _gpu(mesh, case, reorder=True, tol=1e-12, maxit=500):
    from semo_mesh_amd.solver import GpuSolver
    g = GpuSolver(0)
    g.upload_mesh(mesh, reorder=reorder)
    g.set_case(case)
    g.set_solver("fp64", tol_p=tol, tol_u=tol, maxit_p=maxit, maxit_u=maxit)
    return g


def _expand(mesh, rb, cb, blk):
    n = mesh.n_cells
    R = (np.arange(4)[None, :, None] * n + rb[:, None, None]).repeat(4, 2)
    C = (np.arange(4)[None, None, :] * n + cb[:, None, None]).repeat(4, 1)
    return sp.csr_matrix((blk.ravel(), (R.ravel(), C.ravel())), shape=(4 * n, 4 * n))


def _expected(g, mesh, case, mom, pre):
    """The block system from its definition, in the caller's numbering."""
    n = mesh.n_cells
    rm, cm, vm, _ = mom
    rp, cp, vp, _ = pre
    area = g.get_field("face_area")
    nrm = g.get_field("face_normal").reshape(-1, 3)
    wC = g.get_field("face_wC")
    rows = [rp] + [k * n + rm for k in (1, 2, 3)]
    cols = [cp] + [k * n + cm for k in (1, 2, 3)]
    vals = [vp] + [vm] * 3
    nif = mesh.n_internal_faces
    P, N = mesh.owner[:nif].astype(np.int64), mesh.neighbour.astype(np.int64)
    A, w = area[:nif], wC[:nif]
    for k in (1, 2, 3):
        g_ = nrm[:nif, k - 1] * A
        # (row cell, column cell, value): continuity row x u'_k column and
        # momentum row k x p' column carry the same coefficient
        for (r, c, v) in ((P, P, g_ * w), (P, N, g_ * (1 - w)), (N, N, -g_ * (1 - w)), (N, P, -g_ * w)):
            rows += [r, k * n + r]; cols += [k * n + c, c]; vals += [v, v]
    for b, name in enumerate(mesh.patch_names):
        bc = case.bcs[name]
        f = np.arange(mesh.patch_start[b], mesh.patch_start[b] + mesh.patch_nfaces[b])
        Pb = mesh.owner[f].astype(np.int64)
        for k in (1, 2, 3):
            gb = nrm[f, k - 1] * area[f]
            if bc.p_type != 1:
                rows.append(k * n + Pb); cols.append(Pb); vals.append(gb)
            if bc.u_type == 0:
                rows.append(Pb); cols.append(k * n + Pb); vals.append(gb)
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(4 * n, 4 * n))


@pytest.fixture(scope="module")
def system():
    mesh = cavity_mesh(8, jitter=0.1, seed=13)
    case = _case()
    g = _gpu(mesh, case)
    g.incompressiblePressureBased(3)            # a non-trivial state
    g.assemble("momentum"); mom = g.matrix("momentum")
    g.assemble("pressure"); pre = g.matrix("pressure")
    g.assemble("coupled"); rb, cb, blk, rhs = g.coupled_matrix()
    yield mesh, case, g, mom, pre, _expand(mesh, rb, cb, blk), rhs
    g.close()


def test_coupled_blocks_equal_their_definition(system):
    mesh, case, g, mom, pre, K, rhs = system
    E = _expected(g, mesh, case, mom, pre)
    D = (K - E).tocsr()
    scale = np.abs(E.data).max()
    assert (np.abs(D.data).max() if D.nnz else 0.0) <= 1e-14 * scale
    assert np.array_equal(rhs[0], pre[3].ravel())
    assert np.array_equal(rhs[1:], mom[3])


def test_bsr4_spmv_equals_expanded_csr(system):
    mesh, case, g, mom, pre, K, rhs = system
    x = np.random.default_rng(3).standard_normal((4, mesh.n_cells))
    y = g.coupled_spmv(x)
    ref = K @ x.ravel()
    assert np.linalg.norm(y.ravel() - ref) <= 1e-14 * np.linalg.norm(ref)


def test_coupled_solve_reaches_its_residual(system):
    mesh, case, g, mom, pre, K, rhs = system
    x = g.coupled_solve(rhs)
    st = g.last_stats
    res = np.linalg.norm(rhs.ravel() - K @ x.ravel()) / np.linalg.norm(rhs.ravel())
    print(f"coupled solve: {st.iters[0]} iterations, {st.n_levels} levels, rel residual {res:.2e}")
    assert res <= 1e-11
    assert abs(res - st.rel_res[0]) <= 1e-3 * res + 1e-16


def test_coupled_multilevel_solve():
    mesh = cavity_mesh(16, jitter=0.1, seed=17)
    case = _case()
    g = _gpu(mesh, case, tol=1e-10)
    g.incompressiblePressureBased(2)
    g.assemble("coupled")
    rb, cb, blk, rhs = g.coupled_matrix()
    K = _expand(mesh, rb, cb, blk)
    x = g.coupled_solve(rhs)
    st = g.last_stats
    res = np.linalg.norm(rhs.ravel() - K @ x.ravel()) / np.linalg.norm(rhs.ravel())
    print(f"coupled 16^3: {st.iters[0]} iterations, {st.n_levels} levels, rel residual {res:.2e}")
    g.close()
    assert st.n_levels >= 2
    assert res <= 1e-9


def test_coupled_and_simple_share_the_fixed_point():
    mesh = cavity_mesh(6, jitter=0.1, seed=19)
    a = _gpu(mesh, _case(0.3))
    na = a.incompressiblePressureBased(300)
    b = _gpu(mesh, _case(0.8))
    nb = b.coupledPressureBased(60)
    print("SIMPLE norms", na[-2:], "coupled norms", nb[-3:])
    for f in ("P", "U", "V", "W"):
        fa, fb = a.get_field(f), b.get_field(f)
        rel = np.linalg.norm(fa - fb) / np.linalg.norm(fa)
        print(f, rel)
        assert rel <= 1e-9, f
    a.close()
    b.close()


def _bsr(rb, cb, blk, n):
    """The block COO as a scipy BSR matrix over interleaved (p, u, v, w)."""
    o = np.lexsort((cb, rb))
    ptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(ptr, rb[o] + 1, 1)
    return sp.bsr_matrix((blk[o], cb[o], np.cumsum(ptr)), shape=(4 * n, 4 * n), blocksize=(4, 4))


def test_channel_workload_full_size():
    """The block-coupled configuration at its full size: the 1.05M-cell
    jittered 256 x 64 x 64 channel with the seeded 2 % sinusoidal inlet
    perturbation (semo_mesh_amd.cases.channel_workload), two time steps of
    10 outer iterations.  Checked by properties that hold at any size:
      * every outer iteration's FP64 BiCGStab converges to 1e-8 within 20
        iterations (SIMPLE-type preconditioner, FP32 pressure V-cycle);
      * the BSR4 product equals scipy's BSR product of the exported blocks
        on a random vector;
      * a solve of the exported system reaches its residual, recomputed on
        the host in FP64 from the exported blocks."""
    from semo_mesh_amd.cases import channel_workload
    from semo_mesh_amd.solver import GpuSolver
    m, case, u_in, ufix = channel_workload()
    n = m.n_cells
    g = GpuSolver(0)
    g.upload_mesh(m, reorder=True)
    g.set_case(case)
    g.set_solver("mixed", tol_p=1e-8, tol_u=1e-8, maxit_p=1000, maxit_u=500)
    g.set_field("U", np.full(n, u_in))
    g.set_boundary_velocity(ufix)
    for step in range(2):
        norms = g.coupledPressureBased(10)
        st = g.solve_stats("momentum")
        print(f"step {step}: Krylov iterations sum {st['iters'][1]}, max {st['iters'][2]}, "
              f"last rel res {st['rel_res'][0]:.2e}, update norms {norms[0]:.3e} -> {norms[-1]:.3e}")
        assert st["iters"][2] <= 20
        assert st["rel_res"][0] <= 1e-8
        assert np.all(np.isfinite(norms))
    g.assemble("coupled")
    rb, cb, blk, rhs = g.coupled_matrix()
    K = _bsr(rb, cb, blk, n)
    x = np.random.default_rng(11).standard_normal((4, n))
    y = g.coupled_spmv(x)
    ref = (K @ x.T.ravel()).reshape(n, 4).T
    assert np.linalg.norm(y - ref) <= 1e-13 * np.linalg.norm(ref)
    xs = g.coupled_solve(rhs)
    b = rhs.T.ravel()
    res = np.linalg.norm(b - K @ xs.T.ravel()) / np.linalg.norm(b)
    print(f"standalone solve: {g.last_stats.iters[0]} iterations, host FP64 residual {res:.2e}")
    # the iteration stops on its recurrence residual; the true one agrees
    assert res <= 1.5e-8 and abs(res - g.last_stats.rel_res[0]) <= 1e-3 * res
    g.close()
