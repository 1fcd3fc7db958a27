"""semo_compress_volume_fraction on the GPU against SEMO's interface-
compression utility (src/utility/compressVolumeFraction.cpp), built from its
own source (oracle/_ref/semo_compressvf) and run on a jittered 8x6x5 box
with a smeared tanh layer of water: tests/golden/compressvf_box865.npz holds
the fractions it saved after 6 iterations at CFL 0.4 and the residuals it
printed (6 digits).  The oracle's restatement (oracle.oracle.
compress_volume_fraction) equals the saved fractions bit for bit (CPU test
below); the device on the mesh as given equals them bit for bit too, and
renumbered within 1e-13."""
import numpy as np
import pytest

from conftest import load_golden


def test_oracle_matches_utility(oracle_lib):
    from oracle.oracle import Oracle, compress_volume_fraction
    from semo_mesh_amd.cases import cavity_case
    d, m = load_golden("compressvf_box865")
    out, res = compress_volume_fraction(Oracle(m, cavity_case()), d["cvf_in"], float(d["cvf_cfl"]),
                                        int(d["cvf_iters"]))
    assert np.array_equal(out, d["cvf_out"])
    np.testing.assert_allclose(res, d["cvf_residuals"], rtol=2e-6)
    # the compression moved the field: some cells sharpened toward 0 or 1
    assert np.abs(out - d["cvf_in"]).max() > 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("reorder", [False, True])
def test_compress_vs_utility(reorder):
    from semo_mesh_amd.solver import GpuSolver
    d, m = load_golden("compressvf_box865")
    g = GpuSolver(0)
    g.upload_mesh(m, reorder=reorder)
    try:
        out, res = g.compressVolumeFraction(d["cvf_in"], float(d["cvf_cfl"]), int(d["cvf_iters"]))
    finally:
        g.close()
    ref = d["cvf_out"]
    np.testing.assert_allclose(res, d["cvf_residuals"], rtol=2e-6)
    if reorder:
        err = np.abs(out - ref).max()
        print("renumbered max abs err", err)
        assert err <= 1e-13
    else:
        assert np.array_equal(out, ref), np.abs(out - ref).max()
