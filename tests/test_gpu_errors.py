"""Failure reporting through the C ABI: inputs that cannot be solved come
back as status codes, never as SEMO_OK with garbage fields.

Every case is forced from the host (non-finite or overflowing inputs); the
kernels themselves run normally and exit.
"""
import ctypes as C
from dataclasses import replace

import numpy as np
import pytest
import scipy.sparse as sp

from semo_mesh_amd import _lib
from semo_mesh_amd.cases import cavity_case
from semo_mesh_amd.mesh import cavity_mesh

pytestmark = pytest.mark.gpu


def _laplacian(n):
    """1-D Dirichlet Laplacian as COO (diagonal first, as SEMO appends)."""
    rows = list(range(n)); cols = list(range(n)); vals = [2.0] * n
    for i in range(n - 1):
        rows += [i, i + 1]; cols += [i + 1, i]; vals += [-1.0, -1.0]
    return np.array(rows), np.array(cols), np.array(vals)


@pytest.mark.parametrize("eq", ["pressure", "momentum"])
def test_solve_coo_rejects_non_finite_input(eq):
    from semo_mesh_amd.solver import GpuSolver
    g = GpuSolver(0)
    r, c, v = _laplacian(50)
    b = np.ones(50); b[7] = np.nan
    with pytest.raises(_lib.SemoError, match="SEMO_EINVAL"):
        g.solveAMGCL(eq, r, c, v, b)
    v2 = v.copy(); v2[3] = np.inf
    with pytest.raises(_lib.SemoError, match="SEMO_EINVAL"):
        g.solveAMGCL(eq, r, c, v2, np.ones(50))
    # the context still solves a sound system afterwards
    x = g.solveAMGCL(eq, r, c, v, np.ones(50))
    A = sp.coo_matrix((v, (r, c)), shape=(50, 50)).tocsr()
    assert np.linalg.norm(A @ x - 1.0) <= 1e-7 * np.sqrt(50)
    g.close()


@pytest.mark.parametrize("eq", ["pressure", "momentum"])
def test_overflowing_solve_reports_no_convergence(eq):
    """Finite inputs whose residual norm overflows (|b|^2 = inf): the Krylov
    scalars become non-finite, which must end as SEMO_ENOCONV."""
    from semo_mesh_amd.solver import GpuSolver
    g = GpuSolver(0)
    r, c, v = _laplacian(50)
    with pytest.raises(_lib.SemoError, match="SEMO_ENOCONV"):
        g.solveAMGCL(eq, r, c, v, np.full(50, 1e200))
    g.close()


def test_set_case_rejects_null_boundary_array():
    from semo_mesh_amd.solver import GpuSolver
    g = GpuSolver(0)
    g.upload_mesh(cavity_mesh(3), reorder=False)
    case = cavity_case()
    cs = _lib.CaseDesc((C.c_double * 2)(*case.rho), (C.c_double * 2)(*case.mu), case.sigma,
                       (C.c_double * 3)(*case.g), case.dt, case.urf_mom, case.urf_p, case.urf_pu,
                       case.limit, case.urf_vf, case.n_vof, 0, case.max_co, case.max_dt)
    rc = g._L.semo_set_case(g._h, C.byref(cs), None)
    assert _lib.STATUS[rc] == "SEMO_EINVAL"
    g.set_case(case)            # and a valid call still succeeds
    g.close()


def test_function_boundary_without_points_is_rejected():
    from semo_mesh_amd.cases import BC_FUNCTION
    from semo_mesh_amd.solver import GpuSolver
    g = GpuSolver(0)
    g.upload_mesh(cavity_mesh(3), reorder=False)
    case = cavity_case()
    case.bcs["xmin"] = replace(case.bcs["xmin"], vf_type=BC_FUNCTION)
    g.set_case(case)
    with pytest.raises(_lib.SemoError, match="SEMO_EINVAL_BC"):
        g.calcMomentumEqs()
    g.close()
