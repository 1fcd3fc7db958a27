"""Measurement helpers of the library on the GPU.

* semo_volume_integral: sum_i phi_i V_i in FP64 on the device equals the
  same sum over the downloaded field and cell volumes (math.fsum, exactly
  rounded), to 1e-14 relative; repeated calls give the same bits; a
  three-component field is refused.
* The kernel profiler: one SIMPLE step with per-launch events on gives the
  same fields as without, and tags with positive bytes for the kernels a
  step runs (assembly, gradients, Krylov, V-cycle)."""
import math

import numpy as np
import pytest

from semo_mesh_amd.cases import cavity_case, dam_break_case
from semo_mesh_amd.mesh import cavity_mesh, tet_mesh

pytestmark = pytest.mark.gpu


def test_volume_integral_of_alpha():
    from semo_mesh_amd.solver import GpuSolver, SemoError
    m = tet_mesh(6, 3, 3, hi=(4.0, 2.0, 2.0), jitter=0.1, seed=3)
    nc = m.n_cells
    pts = m.points[m.cell_pts.reshape(nc, 4)].mean(axis=1)
    vf = ((pts[:, 0] < 1.2) & (pts[:, 2] < 1.0)).astype(np.float64)
    vf += 0.25 * np.random.default_rng(1).random(nc) * (vf == 0)
    g = GpuSolver(0)
    try:
        g.upload_mesh(m, reorder=True)
        g.set_case(dam_break_case(dt=1e-4, n_vof=1, urf_vf=1.0))
        g.set_field("VF", vf)
        a, v = g.volume_integral("VF")
        vol = g.get_field("cell_volume")
        phi = g.get_field("VF")
        ref = math.fsum(phi * vol)          # products as the device forms them
        assert abs(a - ref) <= 1e-14 * abs(ref)
        assert abs(v - math.fsum(vol)) <= 1e-14 * math.fsum(vol)
        assert abs(v - 16.0) <= 1e-12 * 16.0        # the 4 x 2 x 2 box
        assert g.volume_integral("VF") == (a, v)
        with pytest.raises(SemoError):
            g.volume_integral("gradP")
    finally:
        g.close()


def test_profiled_step_matches_and_tags_the_kernels():
    from semo_mesh_amd.solver import GpuSolver
    m = cavity_mesh(16, jitter=0.1, seed=5)
    out = []
    for prof in (False, True):
        g = GpuSolver(0)
        try:
            g.upload_mesh(m, reorder=True)
            g.set_case(cavity_case())
            g.set_solver("mixed", tol_p=1e-8, tol_u=1e-8, maxit_p=1000, maxit_u=1000)
            g.incompressiblePressureBased(1)
            g.profile(prof)
            g.profile_reset()
            g.incompressiblePressureBased(2)
            out.append({f: g.get_field(f) for f in ("P", "U", "V", "W")})
            st = g.profile_stats()
        finally:
            g.close()
    for f in ("P", "U", "V", "W"):
        assert np.array_equal(out[0][f], out[1][f]), f
    for tag in ("grad_minmax", "mom_face", "mom_cell", "p_face", "p_cell", "bicg3_x", "spmv3_fp64",
                "fcg_update", "spmv_xpby_fp64", "krylov_finish", "amg_chebx_fp32_L0", "gg_face", "gg_cell",
                "p_correct", "recon_boundary"):
        assert tag in st, tag
        assert st[tag]["count"] > 0 and st[tag]["bytes"] > 0 and st[tag]["ms"] > 0, tag
