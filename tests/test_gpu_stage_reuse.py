"""Gradients reused across the stages of an outer iteration.

Within one time step the driver skips gradient passes whose inputs have not
changed: the pressure stage reuses the momentum stage's P and alpha
gradients, the momentum stage right after the VOF stage reuses that stage's
P, U, V, W gradients and bounds, and takes the alpha gradient from the
surface-tension source when it ran.  The solved fields do not read every
reused gradient, so the exposed gradient fields are checked here after whole
steps against the oracle, which recomputes every gradient in every
assembly (un-renumbered meshes).  Tolerance 1e-9 relative: the gradients of
solved fields carry the two sides' solver differences (both solve to
1e-13; gradU measured 3.8e-12 apart), while a stale gradient would be off by
the change of its field over a stage, O(dt) = 1e-3 here.
"""
import numpy as np
import pytest

from semo_mesh_amd.cases import cavity_case, dam_break_case
from semo_mesh_amd.mesh import cavity_mesh, tet_mesh

pytestmark = pytest.mark.gpu

GRADS = ("gradP", "gradU", "gradV", "gradW", "gradVF")


def _compare(g, o):
    for f in GRADS:
        a = g.get_field(f).ravel()
        b = o.get(f).ravel()
        scale = max(np.linalg.norm(b), 1e-300)
        rel = np.linalg.norm(a - b) / scale
        print(f, rel)
        assert rel <= 1e-9 or np.linalg.norm(b) == 0.0 and np.linalg.norm(a) == 0.0, f


def test_dam_break_step_gradients_match_oracle(oracle_lib):
    from semo_mesh_amd.solver import GpuSolver
    m = tet_mesh(4, 2, 2, hi=(4.0, 2.0, 2.0), jitter=0.1)
    nc = m.n_cells
    pts = m.points[m.cell_pts.reshape(nc, 4)].mean(axis=1)
    vf = ((pts[:, 0] < 1.2) & (pts[:, 2] < 1.0)).astype(np.float64)
    case = dam_break_case(dt=1e-3, n_vof=1, urf_vf=1.0)
    case.sigma = 0.07
    g = GpuSolver(0)
    g.upload_mesh(m, reorder=False)
    g.set_case(case)
    g.set_solver("fp64", tol_p=1e-13, tol_u=1e-13, maxit_p=500, maxit_u=500)
    g.set_field("VF", vf)
    o = oracle_lib.Oracle(m, case)
    o.set_solver(tol=1e-14)
    o.set("VF", vf)
    for _ in range(2):
        g.incompressiblePressureBased(1)
        o.step(correctors=1)
        _compare(g, o)
    g.close()


def test_cavity_step_gradients_match_oracle(oracle_lib):
    from semo_mesh_amd.solver import GpuSolver
    m = cavity_mesh(6, jitter=0.1, seed=37)
    case = cavity_case(dt=0.05)
    case.rho = (1.0, 1.0)
    case.mu = (0.01, 0.01)
    g = GpuSolver(0)
    g.upload_mesh(m, reorder=False)
    g.set_case(case)
    g.set_solver("fp64", tol_p=1e-13, tol_u=1e-13, maxit_p=500, maxit_u=500)
    o = oracle_lib.Oracle(m, case)
    o.set_solver(tol=1e-14)
    for _ in range(2):
        g.incompressiblePressureBased(2)
        o.step(correctors=2)
        _compare(g, o)
    g.close()


@pytest.mark.parametrize("name", ["bc_open_box654_2x2", "bc_vof_box654_2x1"])
def test_u_dependent_boundaries_refresh_reused_gradients(oracle_lib, name):
    """Boundaries whose face value follows the owner's velocity (P "switch",
    VF inletOutlet) change between the momentum and the pressure stage when
    the momentum update moves U across their threshold: the pressure stage
    must then recompute the P / alpha gradient it would otherwise keep.  The
    random initial fields of these fixtures move boundary cells across Mach 1
    and across the in/outflow switch within the first corrector."""
    from conftest import golden_case, load_golden
    from semo_mesh_amd.solver import GpuSolver
    d, m = load_golden(name)
    case = golden_case(d)
    nc = m.n_cells
    g = GpuSolver(0)
    g.upload_mesh(m, reorder=False)
    g.set_case(case)
    g.set_solver("fp64", tol_p=1e-13, tol_u=1e-13, maxit_p=500, maxit_u=500)
    o = oracle_lib.Oracle(m, case)
    o.set_solver(tol=1e-14)
    for k, f in enumerate(("P", "U", "V", "W", "VF")):
        g.set_field(f, d["init"][k * nc:(k + 1) * nc])
        o.set(f, d["init"][k * nc:(k + 1) * nc])
    for _ in range(2):
        g.incompressiblePressureBased(int(d["correctors"]))
        o.step(correctors=int(d["correctors"]))
        _compare(g, o)
    g.close()
