"""semo_les on the GPU against the compiled reference's LES models.

The fixtures tests/golden/les_open_box654.npz and les_vof_tet433.npz hold
SEMO_TurbModel_LES's muT and kSGS (calcWALE, calcSmagorinsky,
calcDynamicSmagorinsky, turbulenceModels/LES.cpp) on a final state; the
device computes the same from that state's U, V, W, rho and right face
values, on the mesh as given and RCM-renumbered.  Tolerance: 1e-12 of the
field's largest magnitude -- the device's pow (ocml) and glibc's may differ
in the last place, and the dynamic model's Germano coefficient is a ratio
of sums that cancel.  Also: the velocity-doubling homogeneity of kSGS (4x)
and of the Smagorinsky muT (2x), to 1e-14, and an unknown model refused.

Whole runs with turbType LES (semo_set_les through the case's les_model;
calcCellTransport applies the model at the first step, from zero face
values, and after each VOF iteration): the device's FP64 run ends at the
reference's lesrun_* fields within 1e-7 (the reference's AMGCL tolerance),
and its muT and mu at the oracle's (same model, same steps) within 1e-9."""
import numpy as np
import pytest

from conftest import golden_case, load_golden

pytestmark = pytest.mark.gpu

FIXTURES = ("les_open_box654", "les_vof_tet433")
MODELS = {"WALE": "wale", "smagorinsky": "smagorinsky", "dynamicSmagorinsky": "dynamic"}


def _gpu(d, m, reorder):
    from semo_mesh_amd.solver import GpuSolver
    g = GpuSolver(0)
    g.upload_mesh(m, reorder=reorder)
    g.set_case(golden_case(d))
    return g


def _inputs(d):
    return (d["final_U"], d["final_V"], d["final_W"], d["les_Rho"], d["gg_faceR_U"], d["gg_faceR_V"],
            d["gg_faceR_W"])


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("reorder", [False, True])
def test_les_vs_reference(name, reorder):
    d, m = load_golden(name)
    g = _gpu(d, m, reorder)
    try:
        for model, key in MODELS.items():
            mu, k = g.calcLES(model, *_inputs(d))
            for a, b in ((mu, d[f"les_{key}_muT"]), (k, d[f"les_{key}_kSGS"])):
                err = np.abs(a - b).max() / np.abs(b).max()
                print(name, reorder, model, err)
                assert err <= 1e-12, (model, err)
    finally:
        g.close()


def test_les_homogeneity_and_errors():
    from semo_mesh_amd.solver import SemoError
    d, m = load_golden("les_vof_tet433")
    g = _gpu(d, m, True)
    try:
        U, V, W, R, fU, fV, fW = _inputs(d)
        mu, k = g.calcLES("smagorinsky", U, V, W, R, fU, fV, fW)
        mu2, k2 = g.calcLES("smagorinsky", 2 * U, 2 * V, 2 * W, R, 2 * fU, 2 * fV, 2 * fW)
        assert np.abs(mu2 - 2 * mu).max() <= 1e-14 * np.abs(mu2).max()
        assert np.abs(k2 - 4 * k).max() <= 1e-14 * np.abs(k2).max()
        with pytest.raises(ValueError):
            g.calcLES("kEpsilon", U, V, W, R, fU, fV, fW)
        with pytest.raises(SemoError):
            g._ok(g._L.semo_les(g._h, 0, *[np.ascontiguousarray(a).ctypes.data for a in (U, V, W, R, fU, fV, fW)],
                                mu.ctypes.data, k.ctypes.data))
    finally:
        g.close()


LES_RUNS = ("lesrun_wale_box654_2x2", "lesrun_smag_open_box654_2x1", "lesrun_dyn_vof_tet433_2x1")


@pytest.mark.parametrize("name", LES_RUNS)
def test_les_run_vs_reference_and_oracle(name, oracle_lib):
    d, m = load_golden(name)
    case = golden_case(d)
    nc = m.n_cells
    g = _gpu(d, m, True)
    o = oracle_lib.Oracle(m, case)
    o.set_solver(1e-14, 10000)
    try:
        g.set_solver("fp64", tol_p=1e-13, tol_u=1e-13, maxit_p=3000, maxit_u=3000)
        for k, f in enumerate(("P", "U", "V", "W", "VF")):
            g.set_field(f, d["init"][k * nc:(k + 1) * nc])
            o.set(f, d["init"][k * nc:(k + 1) * nc])
        for _ in range(int(d["steps"])):
            g.incompressiblePressureBased(int(d["correctors"]))
            o.step(int(d["correctors"]))
        vel = np.sqrt(sum(np.linalg.norm(d["final_" + f]) ** 2 for f in "UVW"))
        for f in ("P", "U", "V", "W", "VF"):
            a, b = g.get_field(f), d["final_" + f]
            scale = np.linalg.norm(b) if f in ("P", "VF") else vel
            err = np.linalg.norm(a - b) / scale
            print(name, f, err)
            assert err <= 1e-7, (f, err)
        for f in ("muT", "mu"):
            a, b = g.get_field(f), o.get(f)
            err = np.linalg.norm(a - b) / np.linalg.norm(b)
            print(name, f, err)
            assert err <= 1e-9, (f, err)
    finally:
        g.close()
