"""V-cycle replay as a HIP graph against the same V-cycle issued kernel by
kernel (SEMO_GPU_GRAPHS=0): the graph holds the same kernels with the same
arguments in the same order, so every step must agree bit for bit.  Several
replays per context, with the host's own eager work (assembly, momentum
solves, field reads, synchronisations) between them."""
import numpy as np
import pytest

from semo_mesh_amd.cases import cavity_case
from semo_mesh_amd.mesh import cavity_mesh

pytestmark = pytest.mark.gpu


def _run(monkeypatch, graphs, precision, mesh, steps=4):
    from semo_mesh_amd.solver import GpuSolver
    monkeypatch.setenv("SEMO_GPU_GRAPHS", "1" if graphs else "0")
    g = GpuSolver(0)
    g.upload_mesh(mesh, reorder=True)
    g.set_case(cavity_case())
    g.set_solver(precision, tol_p=1e-10, tol_u=1e-10, maxit_p=300, maxit_u=300)
    out = []
    for _ in range(steps):
        g.incompressiblePressureBased(1)
        out.append({f: g.get_field(f) for f in ("P", "U", "V", "W")})
    st = g.solve_stats("pressure")
    g.close()
    return out, st


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_vcycle_graph_replay_matches_eager(monkeypatch, precision):
    mesh = cavity_mesh(16, jitter=0.1, seed=29)
    a, sa = _run(monkeypatch, True, precision, mesh)
    b, sb = _run(monkeypatch, False, precision, mesh)
    assert sa["n_levels"] >= 2
    for k, (x, y) in enumerate(zip(a, b)):
        for f in x:
            assert np.array_equal(x[f], y[f]), (k, f)
