"""SEMO's LES models (src/turbulenceModels/LES.cpp) in the oracle, CPU only.

Pinned by the compiled reference: tests/golden/les_open_box654.npz and
les_vof_tet433.npz hold muT and kSGS that SEMO_TurbModel_LES::calcWALE,
calcSmagorinsky and calcDynamicSmagorinsky wrote on the final state of a
rotated open box (inletOutlet, zeroGradient, slip and surfaceNormalFixedValue
U patches: every Green-Gauss boundary branch) and of a two-phase tet box
(density from 1.2 to 1000, vertex stencils of up to ~70 cells).  The oracle
reproduces them bit for bit.  Independently of the reference, homogeneity:
doubling the velocities (cell and face values) quadruples kSGS exactly in
every model and doubles muT of Smagorinsky and the dynamic model (scaling by
2 commutes with rounding, pow included); SEMO's WALE is not homogeneous (its
SSd adds squares of S's diagonal, degree 2, to squares of Sd, degree 4,
LES.cpp:110-112), so its muT changes by another factor, which is checked
not to be 2 everywhere.  Doubling the density doubles muT of WALE and
Smagorinsky, quadruples the dynamic model's (its Germano coefficient is
itself linear in rho), and leaves kSGS alone."""
import numpy as np
import pytest

from conftest import golden_case, load_golden

FIXTURES = ("les_open_box654", "les_vof_tet433")
MODELS = {1: "wale", 2: "smagorinsky", 3: "dynamic"}


def _inputs(d):
    return (d["final_U"], d["final_V"], d["final_W"], d["les_Rho"], d["gg_faceR_U"], d["gg_faceR_V"],
            d["gg_faceR_W"])


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("model", [1, 2, 3])
def test_les_bit_exact_vs_reference(name, model, oracle_lib):
    d, m = load_golden(name)
    o = oracle_lib.Oracle(m, golden_case(d))
    mu, k = o.les(model, *_inputs(d))
    assert np.array_equal(mu, d[f"les_{MODELS[model]}_muT"])
    assert np.array_equal(k, d[f"les_{MODELS[model]}_kSGS"])


@pytest.mark.parametrize("model", [1, 2, 3])
def test_les_homogeneity(model, oracle_lib):
    d, m = load_golden("les_vof_tet433")
    o = oracle_lib.Oracle(m, golden_case(d))
    U, V, W, R, fU, fV, fW = _inputs(d)
    mu, k = o.les(model, U, V, W, R, fU, fV, fW)
    assert np.count_nonzero(mu) == m.n_cells
    # velocities x 2
    mu2, k2 = o.les(model, 2 * U, 2 * V, 2 * W, R, 2 * fU, 2 * fV, 2 * fW)
    assert np.array_equal(k2, 4 * k)
    if model == 1:
        assert not np.allclose(mu2, 2 * mu, rtol=1e-3)
    else:
        assert np.array_equal(mu2, 2 * mu)
    # density x 2
    mu3, k3 = o.les(model, U, V, W, 2 * R, fU, fV, fW)
    assert np.array_equal(mu3, (4 if model == 3 else 2) * mu) and np.array_equal(k3, k)


LES_RUNS = ("lesrun_wale_box654_2x2", "lesrun_smag_open_box654_2x1", "lesrun_dyn_vof_tet433_2x1")


@pytest.mark.parametrize("name", LES_RUNS)
def test_les_run_matches_reference(name, oracle_lib):
    d, mesh = load_golden(name)
    case = golden_case(d)
    assert case.les_model
    nc = mesh.n_cells
    vel = np.sqrt(sum(np.linalg.norm(d["final_" + f]) ** 2 for f in "UVW"))
    out = {}
    for les in (case.les_model, None):
        case.les_model = les
        o = oracle_lib.Oracle(mesh, case)
        o.set_solver(1e-14, 10000)
        for k, f in enumerate(("P", "U", "V", "W", "VF")):
            o.set(f, d["init"][k * nc:(k + 1) * nc])
        for _ in range(int(d["steps"])):
            o.step(int(d["correctors"]))
        out[les] = {f: o.get(f) for f in ("P", "U", "V", "W", "VF")}
    model = [k for k in out if k][0]
    for f in ("P", "U", "V", "W", "VF"):
        b = d["final_" + f]
        scale = np.linalg.norm(b) if f in ("P", "VF") else vel
        assert np.linalg.norm(out[model][f] - b) <= 1e-7 * scale, f
    assert np.linalg.norm(out[None]["U"] - d["final_U"]) > 1e-3 * vel
