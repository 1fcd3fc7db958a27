"""semo_update_values on the GPU against the compiled reference's
updateValues (src/solvers/densityBased.cpp:366-512).

tests/golden/update_*.npz: seeded cell values, residuals (the implicit
update) and face Un on a jittered hex box (two species) and a tet box (three
species), with under-relaxation factors and limits that clip a share of the
cells.  On the mesh as given the device's cell values and face Un equal the
reference's bit for bit; renumbered (flipped faces take the other side's
weight, 1 - wC, and the negated normal) within 1e-13 of the largest Un."""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("reorder", [False, True])
@pytest.mark.parametrize("name", ["update_box654", "update_tet322_nsp3"])
def test_update_values_vs_reference(name, reorder):
    from semo_mesh_amd.solver import GpuSolver
    d, m = load_golden(name)
    nsp = int(d["upd_nsp"])
    hdr = d["upd_hdr"]
    g = GpuSolver(0)
    g.upload_mesh(m, reorder=reorder)
    try:
        cv, un = g.updateValues(hdr[:4], hdr[4:], d["upd_res"], d["upd_cells_in"], d["upd_un_in"])
    finally:
        g.close()
    ref_c = d["upd_cells"].reshape(m.n_cells, 5 + nsp)
    ref_u = d["upd_un"]
    assert np.array_equal(cv, ref_c)
    clipped = (ref_c[:, 0] == hdr[4]) | (ref_c[:, 0] == hdr[5]) | (ref_c[:, 4] == hdr[12]) | (ref_c[:, 4] == hdr[13])
    assert clipped.any() and not clipped.all()
    if reorder:
        err = np.abs(un - ref_u).max() / np.abs(ref_u).max()
        print(name, "renumbered Un rel err", err)
        assert err <= 1e-13
    else:
        assert np.array_equal(un, ref_u)
    # boundary faces keep their Un
    nif = m.n_internal_faces
    assert np.array_equal(ref_u[nif:], d["upd_un_in"][nif:])
