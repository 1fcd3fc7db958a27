"""The CPU oracle against the golden vectors written by the compiled
reference (tests/golden/make_golden.py).  Geometry, stencil, least-squares
coefficients, RCM permutation, gradients, limiter bounds, assembled COO
systems and the corrected fields are compared bit for bit; linear-solve
results to the solver tolerances (AMGCL stops at 1e-9 relative)."""
import numpy as np
import pytest

from conftest import GOLDEN_CASES, golden_case, golden_mesh, load_golden

GEOM = ["face_area", "face_normal", "face_xyz", "face_wC", "face_magPN", "face_alphaF",
        "face_ePN", "face_skew", "face_dist", "cell_volume", "cell_xyz", "cell_lsq"]


def calls(d):
    out = []
    k = 0
    while True:
        m = [key for key in d if key.startswith(f"call{k}_") and key.endswith("_rows")]
        if not m:
            return out
        out.append(m[0][: -len("_rows")])
        k += 1


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_mesh_generator_reproduces_fixture_mesh(name):
    """The committed mesh arrays are what the seeded generator makes now."""
    d, mesh = load_golden(name)
    gen = golden_mesh(name)
    for a in ("points", "owner", "neighbour", "face_ptr", "face_pts", "cell_ptr", "cell_pts",
              "patch_start", "patch_nfaces"):
        assert np.array_equal(getattr(gen, a), d["mesh_" + a]), a


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_geometry_stencil_lsq_rcm_bit_exact(name, oracle_lib):
    d, mesh = load_golden(name)
    o = oracle_lib.Oracle(mesh, golden_case(d))
    for k in GEOM:
        assert np.array_equal(o.get(k), d[k]), k
    assert np.array_equal(o.get("stencil_ptr"), d["stencil_ptr"])
    assert np.array_equal(o.get("stencil_idx"), d["stencil_idx"])
    ptr, adj = o.cell_graph()
    perm, inv = oracle_lib.genrcm(ptr, adj)
    assert np.array_equal(perm, d["rcm_perm"])
    assert np.array_equal(inv, d["rcm_inv"])


def _set_state(o, d, p):
    for f in ("P", "U", "V", "W", "oldU", "oldV", "oldW", "Qm0", "Qm1", "Qm2", "VF", "Rho", "mu",
              "Qm3", "oldVF"):
        if p + f in d:
            o.set(f, d[p + f])


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_assembly_replay_bit_exact(name, oracle_lib):
    """Replays every solveAMGCL call: with the reference's fields at call time
    the oracle assembles the identical COO system and right-hand sides, and
    with the reference's solution it produces the identical updated fields."""
    d, mesh = load_golden(name)
    case = golden_case(d)
    o = oracle_lib.Oracle(mesh, case)
    cl = calls(d)
    assert len(cl) == int(d["steps"]) * (2 * int(d["correctors"]) + case.n_vof)
    o.eos()
    for i, p in enumerate(cl):
        p = p + "_"
        _set_state(o, d, p)
        if p + "dt" in d:
            if case.adjust_dt and (i == 0 or "pressure" in cl[i - 1]):
                # first call of a time step: the step's time step comes from
                # these fields (timestep.cpp:260-330)
                o.time_step(True)
                assert o.dt() == d[p + "dt"][0], p + "dt"
            o.set_dt(d[p + "dt"][0])
        o.recon()
        o.gravity()
        first_of_step = "momentum" in p and "volume_fraction" in cl[i - 1]
        if any(case.g):
            # the source is computed once per time step, after the VOF loop
            # (pressureBased.cpp:129-130): it matches a recomputation at the
            # first momentum call of the step; later calls use the stored one
            if first_of_step:
                for k in range(3):
                    assert np.array_equal(o.get(f"srcG{k}"), d[p + f"srcG{k}"]), p + f"srcG{k}"
            for k in range(3):
                o.set(f"srcG{k}", d[p + f"srcG{k}"])
        if case.sigma:
            # curvature + surface-tension source, once per step: equal to
            # round-off (the reference inverts the level-set least-squares
            # matrices by SVD, the oracle by cofactors)
            if first_of_step:
                o.surface_tension()
                for f in ("kappa", "srcS0", "srcS1", "srcS2"):
                    b = d[p + f]
                    assert np.linalg.norm(o.get(f) - b) <= 1e-13 * np.linalg.norm(b), p + f
            for f in ("kappa", "srcS0", "srcS1", "srcS2"):
                o.set(f, d[p + f])
        if "volume_fraction" in p:
            o.mstacs()
            for f in ("faceLVF", "faceRVF"):
                assert np.array_equal(o.get(f), d[p + f]), p + f
            o.assemble_vof()
            for f in ("gradP", "gradU", "gradV", "gradW"):
                assert np.array_equal(o.get(f), d[p + f]), p + f
            nb = 1
        elif "momentum" in p:
            o.assemble_momentum()
            for f in ("gradP", "gradU", "gradV", "gradW", "maxU", "minU", "maxP", "minP"):
                assert np.array_equal(o.get(f), d[p + f]), p + f
            nb = 3
        else:
            o.assemble_pressure()
            for f in ("gradP", "gradU", "gradV", "gradW"):
                assert np.array_equal(o.get(f), d[p + f]), p + f
            nb = 1
        r, c, v = o.coo()
        assert np.array_equal(r, d[p + "rows"]) and np.array_equal(c, d[p + "cols"])
        assert np.array_equal(v, d[p + "vals"]), p + "vals"
        for k in range(nb):
            assert np.array_equal(o.get(f"b{k}"), d[p + f"b{k}"]), p + f"b{k}"
        # the oracle's own solve meets the reference's solution to AMGCL's tolerance
        it, res = o.solve(nb)
        assert res < 1e-12
        for k in range(nb):
            x = o.get(f"x{k}"); xr = d[p + f"x{k}"]
            assert np.linalg.norm(x - xr) <= 1e-7 * max(np.linalg.norm(xr), 1e-300), p
            o.set(f"x{k}", xr)
        # update with the reference's solution; compare with the next call's fields
        nxt = cl[i + 1] + "_" if i + 1 < len(cl) else "final_"
        if "volume_fraction" in p:
            o.vof_update()
            o.eos()
            for f in ("VF", "Rho", "mu"):
                assert np.array_equal(o.get(f), d[nxt + f]), nxt + f
            continue
        if nb == 3:
            o.momentum_update()
        else:
            o.pressure_correct()
        if nb == 3 or nxt == "final_":
            for f in ("P", "U", "V", "W"):
                assert np.array_equal(o.get(f), d[nxt + f]), nxt + f


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_full_simple_run_matches_reference(name, oracle_lib):
    """The oracle's own run (its own linear solves) ends at the reference's
    fields to within the reference solver's relative tolerance."""
    d, mesh = load_golden(name)
    o = oracle_lib.Oracle(mesh, golden_case(d))
    nc = mesh.n_cells
    if "init" in d:
        ini = d["init"]
        for k, f in enumerate(("P", "U", "V", "W", "VF")):
            o.set(f, ini[k * nc:(k + 1) * nc])
    for _ in range(int(d["steps"])):
        o.step(int(d["correctors"]))
    # velocity components relative to the velocity field's norm (a component
    # can be pure round-off by symmetry), pressure relative to its own norm
    vel = np.sqrt(sum(np.linalg.norm(d["final_" + f]) ** 2 for f in "UVW"))
    for f in ("P", "U", "V", "W", "VF"):
        a, b = o.get(f), d["final_" + f]
        scale = np.linalg.norm(b) if f in ("P", "VF") else vel
        assert np.linalg.norm(a - b) <= 1e-7 * scale, f


@pytest.mark.parametrize("name", ["bc_open_box654_2x2", "bc_vof_box654_2x1"])
@pytest.mark.parametrize("field", ["P", "U", "V", "W", "VF"])
def test_gauss_green_bit_exact(name, field, oracle_lib):
    """calcGaussGreen (gradient.cpp:1778-1985) for each field on the final
    state of the boundary-condition fixtures, whose patches carry every type
    (fixedValue, zeroGradient and the mean-of-cell-and-face branch for noSlip,
    slip, surfaceNormalFixedValue, inletOutlet, switch, function)."""
    d, mesh = load_golden(name)
    o = oracle_lib.Oracle(mesh, golden_case(d))
    g = o.gauss_green(field, d["final_" + field], d["gg_faceR_" + field], d["gg_in_" + field])
    assert np.array_equal(g.ravel(), d["gg_out_" + field])
