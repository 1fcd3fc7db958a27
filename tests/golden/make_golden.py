"""Writes the golden vectors under tests/golden/ by running the compiled
reference (oracle/_ref/semo_ref, built from the reference checkout by
oracle/ref.mk) on small seeded meshes.

    make -f oracle/ref.mk REF=/root/reference -j8
    python tests/golden/make_golden.py [case ...]     (default: every case)

Each case stores the mesh arrays it ran on, the reference's stage outputs and
the fields after the run.  The reference itself never travels: only these
arrays are committed.
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from semo_mesh_amd.cases import (BC_FIXED_VALUE, BC_FUNCTION, BC_INLET_OUTLET, BC_NO_SLIP,  # noqa: E402
                                  BC_SLIP, BC_SURFACE_NORMAL_FIXED_VALUE, BC_SWITCH, BC_ZERO_GRADIENT,
                                  PatchBC, cavity_case, dam_break_case)
from semo_mesh_amd.mesh import cavity_mesh, hex_mesh, prism_stack_mesh, rotated, tet_mesh  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "semo_ref")


VOF_ONLY = ("_faceLVF", "_faceRVF", "_Qm3", "_oldVF", "_dt")


def run_case(name, mesh, case, steps, correctors, init=None, keep_calls=True, extra=(), arrays=None):
    """extra: more "key value" lines for the driver's case file; arrays: more
    arrays for the fixture."""
    with tempfile.TemporaryDirectory() as d:
        mesh.write_raw(os.path.join(d, "mesh.raw"))
        ip = None
        if init is not None:
            ip = os.path.join(d, "init.f64")
            np.ascontiguousarray(init, dtype=np.float64).tofile(ip)
        case.write_ref_case(os.path.join(d, "case.txt"), steps=steps,
                            correctors=correctors, init=ip)
        if extra:
            with open(os.path.join(d, "case.txt"), "a") as fh:
                fh.write("".join(ln + "\n" for ln in extra))
        if case.vf_points is not None:        # read as ./VF.csv (reconIncom.cpp:931)
            case.write_vf_csv(os.path.join(d, "VF.csv"))
        out = os.path.join(d, "out")
        r = subprocess.run([REF, os.path.join(d, "mesh.raw"), os.path.join(d, "case.txt"), out],
                           capture_output=True, text=True, timeout=600, cwd=d)
        if "SEMO_REF_DONE" not in r.stdout:
            raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
        arrs = {}
        for fn in sorted(os.listdir(out)):
            base, ext = os.path.splitext(fn)
            if ext not in (".f64", ".i32"):
                continue
            if base.startswith("call") and not keep_calls:
                continue
            if base.startswith("gg_") and not name.startswith(("bc_", "les_")):
                continue
            if base.endswith("_gradVF"):
                continue
            if (base.endswith("_kappa") or "_srcS" in base) and case.sigma == 0.0:
                continue
            if base.endswith(VOF_ONLY) and not case.n_vof:
                continue
            arrs[base] = np.fromfile(os.path.join(out, fn),
                                     dtype=np.float64 if ext == ".f64" else np.int32)
    arrs.update(mesh_arrays(mesh))
    arrs.update(dict(
        case_mu=np.array(case.mu), case_rho=np.array(case.rho), case_dt=np.array(case.dt),
        case_lid=np.array(case.bcs["zmax"].u_value if "zmax" in case.bcs else (0.0, 0.0, 0.0)),
        steps=np.array(steps),
        correctors=np.array(correctors), case_g=np.array(case.g),
        case_n_vof=np.array(case.n_vof), case_urf_vf=np.array(case.urf_vf),
        case_adjust_dt=np.array(int(case.adjust_dt)), case_max_co=np.array(case.max_co),
        case_max_dt=np.array(case.max_dt), case_sigma=np.array(case.sigma),
        # the patch table: (p, u, vf) types and (p, u0, u1, u2, vf) values
        case_bc_types=np.array([[case.bcs[nm].p_type, case.bcs[nm].u_type, case.bcs[nm].vf_type]
                                for nm in mesh.patch_names], dtype=np.int32),
        case_bc_values=np.array([[case.bcs[nm].p_value, *case.bcs[nm].u_value, case.bcs[nm].vf_value]
                                 for nm in mesh.patch_names], dtype=np.float64),
        case_urf=np.array([case.urf_mom, case.urf_p, case.urf_pu])))
    if case.vf_points is not None:
        arrs["case_vf_points"] = np.asarray(case.vf_points, dtype=np.float64)
    if case.les_model:
        arrs["case_les_model"] = np.array(case.les_model)
    if init is not None:
        arrs["init"] = np.asarray(init)
    arrs.update(arrays or {})
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrs)
    print(f"{name}: {len(arrs)} arrays, {os.path.getsize(path) / 1024:.0f} KiB")


def mesh_arrays(mesh):
    """The mesh as fixture arrays (read back by conftest.load_golden)."""
    return dict(mesh_points=mesh.points, mesh_owner=mesh.owner, mesh_neighbour=mesh.neighbour,
                mesh_face_ptr=mesh.face_ptr, mesh_face_pts=mesh.face_pts, mesh_cell_ptr=mesh.cell_ptr,
                mesh_cell_pts=mesh.cell_pts, mesh_patch_start=mesh.patch_start,
                mesh_patch_nfaces=mesh.patch_nfaces, mesh_patch_names=np.array(mesh.patch_names))


def random_init(nc, seed):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.normal(size=nc) * 10.0, rng.normal(size=nc),
                           rng.normal(size=nc), rng.normal(size=nc), np.ones(nc)])


def vof_init(mesh, seed):
    """Small random velocity and pressure, a sharp water block x < 0.5, z < 0.5."""
    nc = mesh.n_cells
    rng = np.random.default_rng(seed)
    cen = np.zeros((nc, 3))
    for i in range(nc):
        pts = mesh.points[mesh.cell_pts[mesh.cell_ptr[i]:mesh.cell_ptr[i + 1]]]
        cen[i] = pts.mean(axis=0)
    vf = ((cen[:, 0] < 0.5) & (cen[:, 2] < 0.5)).astype(np.float64)
    return np.concatenate([rng.normal(size=nc) * 10.0, rng.normal(size=nc) * 0.1,
                           rng.normal(size=nc) * 0.1, rng.normal(size=nc) * 0.1, vf])


def _cavity4():
    # jittered 4^3 cavity, random fields, one SIMPLE iteration: every stage
    m = cavity_mesh(4, jitter=0.1)
    c = cavity_case(); c.mu = (0.5, 0.5)
    run_case("cavity4_jitter_random", m, c, 1, 1, random_init(m.n_cells, 1))


def _box654():
    # non-cubic jittered box, 2 time steps x 2 correctors
    m = hex_mesh(6, 5, 4, jitter=0.1, seed=7)
    c = cavity_case(); c.mu = (0.05, 0.05); c.dt = 0.02
    run_case("box654_jitter_2x2", m, c, 2, 2, random_init(m.n_cells, 2))


def _cavity8():
    # plain cavity physics (rho 1000, mu 1e-3, dt 0.01) on 8^3, 1 iteration
    m = cavity_mesh(8)
    run_case("cavity8_plain", m, cavity_case(), 1, 1, None)


def _orifice():
    # lid columns inside the live orifice profile window (reconIncom.cpp:707-727):
    # x-face centres (i+0.5)/64 hit 0.0006*13 +- 1.6e-4 at i = 0
    m = hex_mesh(8, 4, 4, hi=(0.125, 0.0625, 0.0625), jitter=0.1, seed=11)
    c = cavity_case(); c.mu = (0.01, 0.01); c.dt = 0.001
    run_case("orifice_lid", m, c, 1, 1, random_init(m.n_cells, 3))


def _vof_box654():
    # two-phase VOF (MSTACS face values + volume-fraction equation): water
    # block in a jittered box, gravity, open top; 2 time steps x 2 correctors
    m = hex_mesh(6, 5, 4, jitter=0.1, seed=21, ref_face=False)
    c = dam_break_case(dt=2e-3, n_vof=1, urf_vf=0.8)
    run_case("vof_box654_2x2", m, c, 2, 2, vof_init(m, 4))


def _vof_tet():
    # the same two-phase physics on a jittered Kuhn-split tetrahedral box
    # with the Courant-number time step (timestep.cpp:260-330)
    m = tet_mesh(4, 3, 3, hi=(1.0, 0.75, 0.75), jitter=0.1, seed=8)
    c = dam_break_case(dt=1e-3, n_vof=2, urf_vf=1.0)
    c.adjust_dt, c.max_co, c.max_dt = True, 0.3, 1.0
    run_case("vof_tet433_2x1", m, c, 2, 1, vof_init(m, 6))


def _sigma_box():
    # surface tension: curvature of a jittered-box droplet, sigma = 0.07
    m = hex_mesh(7, 6, 6, jitter=0.1, seed=31, ref_face=False)
    c = dam_break_case(dt=1e-3, n_vof=1, urf_vf=1.0)
    c.sigma = 0.07
    nc = m.n_cells
    cen = np.stack([m.points[m.cell_pts[m.cell_ptr[i]:m.cell_ptr[i + 1]]].mean(0) for i in range(nc)])
    r = np.linalg.norm(cen - np.array([0.5, 0.5, 0.45]), axis=1)
    init = vof_init(m, 12)
    init[4 * nc:] = np.clip((0.32 - r) / 0.15 + 0.5, 0.0, 1.0)     # smeared droplet
    run_case("sigma_box766_2x1", m, c, 2, 1, init)


def _bc_open():
    # the velocity boundaries slip, surfaceNormalFixedValue and inletOutlet
    # (reconIncom.cpp:751-852, eqMomentum.cpp:841-864) and the Mach-switched
    # pressure outlet (reconIncom.cpp:618-631, eqPressure.cpp:625-632) on a
    # jittered box; rho = 1 gives the constant-density sound speed C = 1, so
    # the random velocities put boundary cells on both sides of Mach 1.  The
    # box is rotated (every boundary normal has three non-zero components)
    # and shifted to x > 1, clear of the orifice window of
    # reconIncom.cpp:707-727.
    m = rotated(hex_mesh(6, 5, 4, jitter=0.1, seed=41), seed=42, shift=(3.0, 0.0, 0.0))
    c = cavity_case(); c.rho = (1.0, 1.0); c.mu = (0.05, 0.05); c.dt = 0.02
    c.bcs["xmin"] = PatchBC(u_type=BC_INLET_OUTLET, u_value=(-0.8, 0.0, 0.0))
    c.bcs["xmax"] = PatchBC(p_type=BC_SWITCH, p_value=0.5, u_type=BC_ZERO_GRADIENT)
    c.bcs["ymin"] = PatchBC(u_type=BC_SLIP)
    c.bcs["ymax"] = PatchBC(u_type=BC_SURFACE_NORMAL_FIXED_VALUE, u_value=(0.3, 0.0, 0.0))
    run_case("bc_open_box654_2x2", m, c, 2, 2, random_init(m.n_cells, 5))


def _bc_vof():
    # two-phase: the volume-fraction "function" boundary (nearest point of
    # ./VF.csv, reconIncom.cpp:929-1019; eqVolfrac.cpp:663-665), inletOutlet
    # U and VF, a Mach-switched pressure boundary whose sound speed follows
    # the mixture (air cells: C ~ 1.1), slip and surfaceNormalFixedValue walls
    m = hex_mesh(6, 5, 4, jitter=0.1, seed=43, ref_face=False)
    c = dam_break_case(dt=2e-3, n_vof=1, urf_vf=0.8)
    c.bcs["xmin"] = PatchBC(u_type=BC_INLET_OUTLET, u_value=(-0.4, 0.0, 0.0),
                            vf_type=BC_FUNCTION)
    c.bcs["xmax"] = PatchBC(p_type=BC_SWITCH, p_value=0.0, u_type=BC_ZERO_GRADIENT,
                            vf_type=BC_INLET_OUTLET, vf_value=0.0)
    c.bcs["ymin"] = PatchBC(u_type=BC_SLIP)
    c.bcs["ymax"] = PatchBC(u_type=BC_SURFACE_NORMAL_FIXED_VALUE, u_value=(0.0, 0.0, 0.0),
                            vf_type=BC_FUNCTION)
    rng = np.random.default_rng(44)
    pts = np.column_stack([rng.random(40), rng.random((40, 3))])
    c.vf_points = pts
    init = vof_init(m, 45)
    nc = m.n_cells
    init[nc:4 * nc] *= 15.0            # speeds of O(1): some air cells above Mach 1
    run_case("bc_vof_box654_2x1", m, c, 2, 1, init)


def _les():
    """SEMO_TurbModel_LES (turbulenceModels/LES.cpp): WALE, Smagorinsky and
    dynamic Smagorinsky on the final state of (a) the rotated open box of
    bc_open (slip, surfaceNormalFixedValue, inletOutlet and zeroGradient U
    patches, so that the Green-Gauss gradients take every boundary branch)
    and (b) a two-phase jittered tet box (density varying from cell to
    cell, vertex stencils of up to about 70 cells for the dynamic model)."""
    m = rotated(hex_mesh(6, 5, 4, jitter=0.1, seed=41), seed=42, shift=(3.0, 0.0, 0.0))
    c = cavity_case(); c.rho = (1.0, 1.0); c.mu = (0.05, 0.05); c.dt = 0.02
    c.bcs["xmin"] = PatchBC(u_type=BC_INLET_OUTLET, u_value=(-0.8, 0.0, 0.0))
    c.bcs["xmax"] = PatchBC(p_value=0.5, u_type=BC_ZERO_GRADIENT)
    c.bcs["ymin"] = PatchBC(u_type=BC_SLIP)
    c.bcs["ymax"] = PatchBC(u_type=BC_SURFACE_NORMAL_FIXED_VALUE, u_value=(0.3, 0.0, 0.0))
    run_case("les_open_box654", m, c, 2, 1, random_init(m.n_cells, 51), keep_calls=False, extra=("les 1",))
    m = tet_mesh(4, 3, 3, hi=(1.0, 0.75, 0.75), jitter=0.1, seed=52)
    c = dam_break_case(dt=1e-3, n_vof=1, urf_vf=1.0)
    run_case("les_vof_tet433", m, c, 1, 1, vof_init(m, 53), keep_calls=False, extra=("les 1",))


def _les_runs():
    """Whole runs with turbType LES (calcCellTransport applies the model at
    the first step and after every VOF iteration, transport.cpp:38-97):
    WALE on a single-phase jittered box (2 steps x 2 correctors: the
    first-step muT, from zero face values, stays in mu), Smagorinsky on the
    open box, dynamic Smagorinsky on a two-phase tet box (muT recomputed
    after each VOF iteration from the reconstructed face values)."""
    m = hex_mesh(6, 5, 4, jitter=0.1, seed=7)
    c = cavity_case(); c.mu = (0.05, 0.05); c.dt = 0.02; c.les_model = "WALE"
    run_case("lesrun_wale_box654_2x2", m, c, 2, 2, random_init(m.n_cells, 61), keep_calls=False)
    m = rotated(hex_mesh(6, 5, 4, jitter=0.1, seed=41), seed=42, shift=(3.0, 0.0, 0.0))
    c = cavity_case(); c.rho = (1.0, 1.0); c.mu = (0.05, 0.05); c.dt = 0.02
    c.bcs["xmin"] = PatchBC(u_type=BC_INLET_OUTLET, u_value=(-0.8, 0.0, 0.0))
    c.bcs["xmax"] = PatchBC(p_value=0.5, u_type=BC_ZERO_GRADIENT)
    c.bcs["ymin"] = PatchBC(u_type=BC_SLIP)
    c.bcs["ymax"] = PatchBC(u_type=BC_SURFACE_NORMAL_FIXED_VALUE, u_value=(0.3, 0.0, 0.0))
    c.les_model = "smagorinsky"
    run_case("lesrun_smag_open_box654_2x1", m, c, 2, 1, random_init(m.n_cells, 62), keep_calls=False)
    m = tet_mesh(4, 3, 3, hi=(1.0, 0.75, 0.75), jitter=0.1, seed=8)
    c = dam_break_case(dt=1e-3, n_vof=2, urf_vf=1.0); c.les_model = "dynamicSmagorinsky"
    run_case("lesrun_dyn_vof_tet433_2x1", m, c, 2, 1, vof_init(m, 63), keep_calls=False)


def flux_states(n=400, nsp=2, seed=71):
    """Face states for the convective flux schemes: [n][rec] with rec =
    2 (8 + nsp) + 10 -- left (P, U, V, W, T, rho, C, Ht, Y...), right, the
    unit normal, then w1, w2, w3, fC, Uco, dt, Lch.  Gas- and liquid-like
    sides, Mach numbers from 0 to about 3, pressure ratios on both sides of
    3/4; the last rows are special cases: identical sides, both at rest,
    nearly equal velocities (|dU| < 1e-5) and supersonic sides."""
    rng = np.random.default_rng(seed)
    rec = 2 * (8 + nsp) + 10
    out = np.zeros((n, rec))

    def side(k):
        liquid = rng.random(k) < 0.4
        rho = np.where(liquid, 900 + 200 * rng.random(k), 0.8 + 0.8 * rng.random(k))
        C = np.where(liquid, 1200 + 600 * rng.random(k), 300 + 80 * rng.random(k))
        mach = 3.0 * rng.random(k) ** 2
        d = rng.normal(size=(k, 3)); d /= np.linalg.norm(d, axis=1)[:, None]
        U = d * (mach * C)[:, None]
        P = 1e5 * (0.5 + 1.5 * rng.random(k))
        T = 280 + 40 * rng.random(k)
        Ht = C * C / 0.4 + 0.5 * (U * U).sum(axis=1)
        Y = rng.random((k, nsp)); Y /= Y.sum(axis=1)[:, None]
        return np.column_stack([P, U, T, rho, C, Ht, Y])

    L, R = side(n), side(n)
    out[:, :8 + nsp] = L
    out[:, 8 + nsp:2 * (8 + nsp)] = R
    nv = rng.normal(size=(n, 3)); nv /= np.linalg.norm(nv, axis=1)[:, None]
    o = 2 * (8 + nsp)
    out[:, o:o + 3] = nv
    out[:, o + 3:o + 6] = 0.05 + 0.95 * rng.random((n, 3))          # w1, w2, w3
    out[:, o + 6] = 0.5 * (L[:, 6] + R[:, 6])                        # fC
    out[:, o + 7] = 1.0 + rng.random(n)                              # Uco
    out[:, o + 8] = 1e-4 * (1.0 + rng.random(n))                     # dt
    out[:, o + 9] = 0.05 + 0.1 * rng.random(n)                       # Lch
    m = n - 40
    out[m:m + 10, 8 + nsp:2 * (8 + nsp)] = out[m:m + 10, :8 + nsp]  # identical sides
    out[m + 10:m + 20, 1:4] = 0.0                                    # both at rest
    out[m + 10:m + 20, 8 + nsp + 1:8 + nsp + 4] = 0.0
    out[m + 20:m + 30, 8 + nsp + 1:8 + nsp + 4] = (out[m + 20:m + 30, 1:4]
                                                  + 1e-7 * rng.normal(size=(10, 3)))
    for s0 in (0, 8 + nsp):                                          # supersonic sides
        u = out[m + 30:m + 40, s0 + 1:s0 + 4]
        out[m + 30:m + 40, s0 + 1:s0 + 4] = u / np.linalg.norm(u, axis=1)[:, None] * (
            2.0 * out[m + 30:m + 40, s0 + 6])[:, None]
    return out


def _flux():
    """SEMO_Solvers_Builder::calcFluxHAUS, SLAU2, Roe, APRoe, RoeAMTP, Upwind,
    AUSMPWP_N and RoeM_N (solvers/flux.cpp) on the face states above."""
    nsp = 2
    st = flux_states(nsp=nsp)
    with tempfile.TemporaryDirectory() as d:
        fn = os.path.join(d, "faces.bin")
        with open(fn, "wb") as fh:
            np.array([st.shape[0], nsp], dtype=np.int32).tofile(fh)
            st.tofile(fh)
        out = os.path.join(d, "out")
        r = subprocess.run([REF, "-", "-", out, "flux", fn], capture_output=True, text=True, timeout=120, cwd=d)
        if "SEMO_REF_DONE" not in r.stdout:
            raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
        arrs = {"flux_states": st, "flux_nsp": np.array(nsp)}
        for fnm in sorted(os.listdir(out)):
            base, ext = os.path.splitext(fnm)
            if ext == ".f64":
                arrs[base] = np.fromfile(os.path.join(out, fnm), dtype=np.float64).reshape(st.shape[0], -1)
    path = os.path.join(HERE, "flux_faces.npz")
    np.savez_compressed(path, **arrs)
    print(f"flux_faces: {len(arrs)} arrays, {os.path.getsize(path) / 1024:.0f} KiB")


def linsol_states(mesh, nsp=2, seed=81, vol=None):
    """Cell states of SEMO's density-based solvers for the implicit updates:
    [nc][12 + 3 nsp] = (P, U, V, W, T, rho, C, Ht, dRhoDP, dHtDP, dRhoDT,
    dHtDT, Ur, dtPseudo, Y.., dRhoDY.., dHtDY..).  Ideal-gas air cells
    (R = 287, gamma = 1.4) and about a third liquid-like cells (rho ~ 1000,
    C ~ 1400), Mach up to 0.8, Ur = |U| clipped to [0.01 C, C], dtPseudo from
    a pseudo CFL of 1 to 5 on the cell size (vol: the cell volumes; the
    oracle's geometry when not given)."""
    rng = np.random.default_rng(seed)
    n = mesh.n_cells
    liquid = rng.random(n) < 0.35
    P = 1e5 * (0.8 + 0.4 * rng.random(n))
    T = 280.0 + 40.0 * rng.random(n)
    Rg, gam = 287.0, 1.4
    cp = gam * Rg / (gam - 1.0)
    rho = np.where(liquid, 1000.0 * (1.0 + 0.02 * rng.random(n)), P / (Rg * T))
    C = np.where(liquid, 1300.0 + 200.0 * rng.random(n), np.sqrt(gam * Rg * T))
    d = rng.normal(size=(n, 3)); d /= np.linalg.norm(d, axis=1)[:, None]
    u = d * (0.8 * rng.random(n) * C)[:, None]
    Ht = np.where(liquid, 4180.0, cp) * T + 0.5 * (u * u).sum(axis=1)
    drdp = np.where(liquid, 1.0 / (C * C), 1.0 / (Rg * T))
    dhdp = np.where(liquid, 1e-3 * (0.5 + rng.random(n)), 0.0)
    drdt = np.where(liquid, -0.2 - 0.2 * rng.random(n), -P / (Rg * T * T))
    dhdt = np.where(liquid, 4180.0, cp)
    Ur = np.clip(np.linalg.norm(u, axis=1), 0.01 * C, C)
    h = np.abs(np.asarray(geometry_volumes(mesh) if vol is None else vol)) ** (1.0 / 3.0)
    dtp = (1.0 + 4.0 * rng.random(n)) * h / (np.linalg.norm(u, axis=1) + C)
    Y = rng.random((n, nsp)); Y /= Y.sum(axis=1)[:, None]
    drdy = rho[:, None] * (rng.random((n, nsp - 1)) * 4.0 - 2.0)
    dhdy = 1e4 * (rng.random((n, nsp - 1)) * 2.0 - 1.0)
    return np.column_stack([P, u, T, rho, C, Ht, drdp, dhdp, drdt, dhdt, Ur, dtp, Y, drdy, dhdy])


def geometry_volumes(mesh):
    """Cell volumes from the oracle's geometry (only to size dtPseudo)."""
    from oracle.oracle import Oracle
    from semo_mesh_amd.cases import cavity_case
    return Oracle(mesh, cavity_case()).get("cell_volume")


def linsol_residuals(n, nsp=2, seed=82):
    """Residuals [nc][4 + nsp] scaled per equation (mass, momentum, energy,
    species)."""
    rng = np.random.default_rng(seed)
    scale = np.array([1.0, 100.0, 100.0, 100.0, 1e5] + [0.1] * (nsp - 1))
    return rng.normal(size=(n, 4 + nsp)) * scale


def _linsol():
    """calcLinearSolver (point-implicit), calcSingleLinearSolver and
    calcLinearSolverLUSGS (solvers/linearSolver.cpp, LU-SGS in lib_RCM's
    order) on seeded density-based states: jittered hex and tet boxes with
    two species, a hex box with one and a tet box with three."""
    dt = 1e-3
    for name, m, seed, nsp in (
            ("linsol_box654", hex_mesh(6, 5, 4, jitter=0.1, seed=91), 83, 2),
            ("linsol_tet433", tet_mesh(4, 3, 3, hi=(1.0, 0.75, 0.75), jitter=0.1, seed=92), 84, 2),
            ("linsol_box543_nsp1", hex_mesh(5, 4, 3, jitter=0.1, seed=93), 85, 1),
            ("linsol_tet322_nsp3", tet_mesh(3, 2, 2, jitter=0.1, seed=94), 86, 3)):
        st = linsol_states(m, nsp, seed)
        r0 = linsol_residuals(m.n_cells, nsp, seed + 100)
        with tempfile.TemporaryDirectory() as d:
            fn = os.path.join(d, "state.f64")
            with open(fn, "wb") as fh:
                np.ascontiguousarray(st).tofile(fh)
                np.ascontiguousarray(r0).tofile(fh)
            run_case(name, m, cavity_case(), 0, 1, keep_calls=False,
                     extra=(f"linsol {fn} {dt!r}", f"nspecies {nsp}"),
                     arrays={"linsol_state": st, "linsol_res0": r0, "linsol_dt": np.array(dt),
                             "linsol_nsp": np.array(nsp)})


def rhs_inputs(mesh, nsp=2, seed=101, vol=None):
    """calcRHS inputs: per face left / right records (P, U, V, W, T, rho, C,
    Ht, mu, Y_0..Y_{nsp-2}), per face nine velocity gradients, per cell (rho,
    U, V, W, Ht, P, dtPseudo, VF, Y_0..Y_{nsp-2}, Qn.., Qm..).  Gas- and
    liquid-like face sides (Mach 0 to 1.5), a smooth two-phase VF, Qn / Qm
    the conserved variables perturbed by a few per cent."""
    rng = np.random.default_rng(seed)
    nf, nc, N = mesh.n_faces, mesh.n_cells, 4 + nsp

    def side(k):
        liquid = rng.random(k) < 0.3
        rho = np.where(liquid, 900 + 200 * rng.random(k), 0.8 + 0.8 * rng.random(k))
        C = np.where(liquid, 1200 + 600 * rng.random(k), 300 + 80 * rng.random(k))
        d = rng.normal(size=(k, 3)); d /= np.linalg.norm(d, axis=1)[:, None]
        U = d * (1.5 * rng.random(k) ** 2 * C)[:, None]
        P = 1e5 * (0.5 + 1.5 * rng.random(k))
        T = 280 + 40 * rng.random(k)
        Ht = C * C / 0.4 + 0.5 * (U * U).sum(axis=1)
        mu = np.where(liquid, 1e-3, 1.8e-5) * (0.5 + rng.random(k))
        Y = rng.random((k, max(nsp - 1, 0)))
        return np.column_stack([P, U, T, rho, C, Ht, mu, Y])

    fl, fr = side(nf), side(nf)
    fg = rng.normal(size=(nf, 9)) * 50.0
    h = 1.0 if vol is None else np.cbrt(np.abs(vol))
    x = np.asarray(mesh.points)[np.asarray(mesh.cell_pts)[np.asarray(mesh.cell_ptr)[:-1]]]
    VF = 0.5 * (1.0 + np.tanh((0.45 - x[:, 2]) / 0.15))
    cs = side(nc)
    rho, U, Ht, P = cs[:, 5], cs[:, 1:4], cs[:, 7], cs[:, 0]
    Y = rng.random((nc, max(nsp - 1, 0)))
    Q = np.column_stack([rho, rho[:, None] * U, rho * Ht - P] + [rho[:, None] * Y])
    Qn = Q * (1.0 + 0.03 * rng.normal(size=Q.shape))
    Qm = Q * (1.0 + 0.03 * rng.normal(size=Q.shape))
    dtp = (1.0 + 4.0 * rng.random(nc)) * h * 0.1 / (np.linalg.norm(U, axis=1) + cs[:, 6])
    cells = np.column_stack([rho, U, Ht, P, dtp, VF, Y, Qn, Qm])
    assert cells.shape[1] == 8 + (nsp - 1) + 2 * N
    return fl, fr, fg, cells


def _rhs():
    """calcFluxes of SEMO's density-based solvers (solvers/RHS.cpp:98-406:
    SLAU2 and the viscous flux per face, summed into the cells in face order;
    calcRHS itself crashes in its sourceTerms) on seeded face and cell
    states: jittered hex and tet boxes with two species, a tet box with three
    (with one species calcFluxes reads fVF[0], which does not exist)."""
    dt = 1e-3
    for name, m, seed, nsp in (("rhs_box654", hex_mesh(6, 5, 4, jitter=0.1, seed=111), 121, 2),
                               ("rhs_tet433", tet_mesh(4, 3, 3, hi=(1.0, 0.75, 0.75), jitter=0.1, seed=112), 122, 2),
                               ("rhs_tet322_nsp3", tet_mesh(3, 2, 2, jitter=0.1, seed=113), 123, 3)):
        fl, fr, fg, cells = rhs_inputs(m, nsp, seed, vol=geometry_volumes(m))
        c = cavity_case(); c.g = (0.5, -1.0, -9.81); c.sigma = 0.0
        with tempfile.TemporaryDirectory() as d:
            fn = os.path.join(d, "rhs.f64")
            with open(fn, "wb") as fh:
                for a in (fl, fr, fg, cells):
                    np.ascontiguousarray(a).tofile(fh)
            run_case(name, m, c, 0, 1, keep_calls=False, extra=(f"rhs {fn} {dt!r}", f"nspecies {nsp}"),
                     arrays={"rhs_face_left": fl, "rhs_face_right": fr, "rhs_face_grad": fg, "rhs_cells": cells,
                             "rhs_dt": np.array(dt), "rhs_nsp": np.array(nsp)})


def _update():
    """updateValues of SEMO's density-based solvers (solvers/densityBased.cpp:
    366-512): under-relaxed increments with the limits, the mass fractions
    clipped to [0, 1] and closed, the face normal velocity from the weighted
    momentum increments; hex and tet boxes, two and three species, limits
    that clip a share of the cells."""
    for name, m, seed, nsp in (("update_box654", hex_mesh(6, 5, 4, jitter=0.1, seed=131), 141, 2),
                               ("update_tet322_nsp3", tet_mesh(3, 2, 2, jitter=0.1, seed=132), 142, 3)):
        rng = np.random.default_rng(seed)
        nc, nf, N = m.n_cells, m.n_faces, 4 + nsp
        hdr = np.array([0.7, 0.5, 0.6, 0.8, 9e4, 1.1e5, -30.0, 30.0, -30.0, 30.0, -30.0, 30.0, 250.0, 350.0])
        Y = rng.random((nc, nsp)); Y /= Y.sum(axis=1)[:, None]
        cv = np.column_stack([1e5 * (1.0 + 0.05 * rng.normal(size=nc)), rng.normal(size=(nc, 3)) * 20.0,
                              300.0 + 20.0 * rng.normal(size=nc), Y])
        rs = rng.normal(size=(nc, N)) * np.array([5e3, 8.0, 8.0, 8.0, 20.0] + [0.3] * (nsp - 1))
        un = rng.normal(size=nf) * 10.0
        with tempfile.TemporaryDirectory() as d:
            fn = os.path.join(d, "upd.f64")
            with open(fn, "wb") as fh:
                for a in (hdr, cv, rs, un):
                    np.ascontiguousarray(a).tofile(fh)
            run_case(name, m, cavity_case(), 0, 1, keep_calls=False, extra=(f"update {fn}", f"nspecies {nsp}"),
                     arrays={"upd_hdr": hdr, "upd_cells_in": cv, "upd_res": rs, "upd_un_in": un,
                             "upd_nsp": np.array(nsp)})


def _poly20():
    """Three stacked prisms over a jittered 20-gon: internal faces and caps
    of 20 points (above the 16 the device keeps in registers), lid-driven
    top (fixedValue p = 0, U = (1, 0, 0)), two SIMPLE steps."""
    m = prism_stack_mesh(20, 3, jitter=0.05, seed=3)
    c = cavity_case(dt=0.01)
    c.mu = (0.01, 0.01)
    c.bcs = {"bottom": PatchBC(), "side": PatchBC(),
             "top": PatchBC(p_type=BC_FIXED_VALUE, p_value=0.0, u_type=BC_FIXED_VALUE, u_value=(1.0, 0.0, 0.0))}
    run_case("poly20_stack3", m, c, 2, 1, random_init(m.n_cells, 81))


# ------------------------------------------------------------------ VTU files --
REF_TREE = "/root/reference"
SEMO_INITIAL = os.path.join(ROOT, "oracle", "_ref", "semo_initial")


def _vtu():
    """SEMO's own writer (SEMO_Mesh_Save::vtu through semo_ref's savevtu
    mode) on a seeded state of a jittered tet mesh, in its three encodings:
    ascii (17 digits), binary, zlib level 6.  Stores the written files and
    the state that went in (point levels i % 4, cell levels i % 3, cell
    groups i % 5, time 0.25, as the driver sets them)."""
    m = tet_mesh(2, 2, 1, jitter=0.1, seed=5)
    nc = m.n_cells
    init = random_init(nc, 71)
    init[4 * nc:] = np.random.default_rng(72).random(nc)
    case = cavity_case()
    case.bcs = {nm: PatchBC() for nm in m.patch_names}
    with tempfile.TemporaryDirectory() as d:
        m.write_raw(os.path.join(d, "mesh.raw"))
        ip = os.path.join(d, "init.f64")
        np.ascontiguousarray(init, dtype=np.float64).tofile(ip)
        case.write_ref_case(os.path.join(d, "case.txt"), init=ip)
        for tag, fmt, comp in (("ascii", "ascii", 0), ("binary", "binary", 0), ("zlib", "binary", 6)):
            out = os.path.join(d, "out_" + tag)
            r = subprocess.run([REF, os.path.join(d, "mesh.raw"), os.path.join(d, "case.txt"), out, "savevtu",
                                fmt, str(comp)], capture_output=True, text=True, timeout=120, cwd=d)
            if "SEMO_REF_DONE" not in r.stdout:
                raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
            with open(os.path.join(out, "save", "0.25", "plot.0.vtu")) as fh:
                text = fh.read()
            with open(os.path.join(HERE, f"vtu_{tag}_tet221.vtu"), "w") as fh:
                fh.write(text)
    np.savez_compressed(os.path.join(HERE, "vtu_tet221.npz"), init=init, **mesh_arrays(m))
    print("vtu_tet221: 3 files")


INITIAL_SPECS = {
    # name: (pressure, velocities, temperature, water) initialField bodies
    "box_cyl_box_sphere": (
        ["type boxToCell;", "value 0.5;", "box (0.2 0.1 -1) (0.7 0.8 0.55);", "value 2.5;"],
        ["type cylinderToCell;", "value (0.1 0 0);", "points (0.1 0.2 0.3) (0.9 0.7 0.6);", "radius 0.3;",
         "value (1 -2 0.5);"],
        ["type boxToCell;", "value 300;", "box (0.5 0 0) (1 1 1);", "value 350;"],
        ["type sphereToCell;", "value 0;", "points (0.4 0.5 0.45);", "radius 0.35;", "value 1;"]),
    # (velocities sphereToCell is left out: the reference's parser throws on
    # every input, initial.cpp:1077 erasing from an empty string)
    "fixed_box_fixed_cyl": (
        ["type fixedValue;", "value 101325;"],
        ["type boxToCell;", "value (0 0 0);", "box (0.3 0.3 0.3) (0.8 0.9 0.7);", "value (0 0.75 -1.25);"],
        ["type fixedValue;", "value 293.15;"],
        ["type cylinderToCell;", "value 0.25;", "points (0 0.5 0.5) (1 0.5 0.5);", "radius 0.33;", "value 0.9;"]),
    "box_fixed_box_box": (
        ["type boxToCell;", "value -1;", "box (0 0 0) (0.5 0.5 0.5);", "value 1;"],
        ["type fixedValue;", "value (3 2 1);"],
        ["type boxToCell;", "value 280;", "box (0 0.5 0) (1 1 0.5);", "value 290;"],
        ["type boxToCell;", "value 1;", "box (0 0 0.6) (1 1 1);", "value 0;"]),
}


def write_initial_file(path, body, patches):
    """A constant/<field> file in SEMO's layout: the initialField block and a
    zeroGradient boundaryField entry per patch (the utility matches the
    patch names against the mesh's)."""
    bf = "".join(f"\t{nm}\n\t{{\n\t\ttype\tzeroGradient;\n\t}}\n" for nm in patches)
    with open(path, "w") as fh:
        fh.write("initialField\n{\n" + "".join(f"\t{line}\n" for line in body) + "}\n\nboundaryField\n{\n"
                 + bf + "}\n")


def _initial():
    """SEMO's initial-field utility (src/utility/initial.cpp, built as it is
    into oracle/_ref/semo_initial) on a jittered 6^3 cube for each spec of
    INITIAL_SPECS, in a case holding the reference's sample system/ and
    constant/ files (saveFormat set to binary) with the spec's initialField
    blocks, the mesh as ./grid/0/plot.0.vtu from write_grid_vtu.  Stores the
    mesh, the specs and the fields the utility saved."""
    import shutil
    from semo_mesh_amd.vtu import read_vtu, write_grid_vtu
    if not os.path.exists(SEMO_INITIAL):
        raise SystemExit("build the reference first: make -f oracle/ref.mk REF=/root/reference")
    m = cavity_mesh(6, jitter=0.15, seed=73)
    nc = m.n_cells
    out = {}
    for name, spec in INITIAL_SPECS.items():
        with tempfile.TemporaryDirectory() as d:
            shutil.copytree(os.path.join(REF_TREE, "system"), os.path.join(d, "system"))
            shutil.copytree(os.path.join(REF_TREE, "constant"), os.path.join(d, "constant"))
            cd = os.path.join(d, "system", "controlDict")
            with open(cd) as fh:
                lines = fh.read().splitlines()
            lines = ["saveFormat = binary" if ln.startswith("saveFormat") else
                     "saveCompression = 0" if ln.startswith("saveCompression") else ln for ln in lines]
            with open(cd, "w") as fh:
                fh.write("\n".join(lines) + "\n")
            for fname, body in zip(("pressure", "velocities", "temperature", "water", "air"),
                                   spec + (["type fixedValue;", "value 0;"],)):
                write_initial_file(os.path.join(d, "constant", fname), body, m.patch_names)
            os.makedirs(os.path.join(d, "grid", "0"))
            write_grid_vtu(os.path.join(d, "grid", "0", "plot.0.vtu"), m)
            r = subprocess.run([SEMO_INITIAL], capture_output=True, text=True, timeout=120, cwd=d)
            res = os.path.join(d, "save", "0", "plot.0.vtu")
            if not os.path.exists(res):
                raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
            _, f, _ = read_vtu(res)
        out[name + "_P"] = f["pressure"]
        out[name + "_UVW"] = f["velocity"]
        out[name + "_T"] = f["temperature"]
        out[name + "_VF"] = f["volume-fraction-water"]
        out[name + "_MF"] = f["mass-fraction-water"]
        out[name + "_spec"] = np.array(["\n".join(b) for b in spec])
    np.savez_compressed(os.path.join(HERE, "initial_cube6.npz"), **out, **mesh_arrays(m))
    print("initial_cube6:", len(INITIAL_SPECS), "specs")


SEMO_COMPRESSVF = os.path.join(ROOT, "oracle", "_ref", "semo_compressvf")


def _compressvf():
    """SEMO's interface-compression utility (src/utility/compressVolumeFraction
    .cpp, built as it is into oracle/_ref/semo_compressvf) on a jittered
    8x6x5 box whose water fraction is a smeared tanh layer: CFL 0.4, 6
    iterations, in a case holding the reference's sample system/ and
    constant/ files (binary saves, start from 0), the state as ./save/0/
    plot.0.vtu.  Stores the mesh, the input and the compressed fractions."""
    import shutil
    from semo_mesh_amd.vtu import read_vtu, semo_fields, write_vtu
    m = hex_mesh(8, 6, 5, jitter=0.1, seed=151)
    x = np.asarray(m.points)[np.asarray(m.cell_pts)[np.asarray(m.cell_ptr)[:-1]]]
    vf = 0.5 * (1.0 + np.tanh((0.5 - x[:, 0] - 0.3 * x[:, 2]) / 0.2))
    nc = m.n_cells
    with tempfile.TemporaryDirectory() as d:
        shutil.copytree(os.path.join(REF_TREE, "system"), os.path.join(d, "system"))
        shutil.copytree(os.path.join(REF_TREE, "constant"), os.path.join(d, "constant"))
        cd = os.path.join(d, "system", "controlDict")
        with open(cd) as fh:
            lines = fh.read().splitlines()
        lines = ["saveFormat = binary" if ln.startswith("saveFormat") else
                 "saveCompression = 0" if ln.startswith("saveCompression") else
                 "startFrom = 0" if ln.startswith("startFrom") else ln for ln in lines]
        with open(cd, "w") as fh:
            fh.write("\n".join(lines) + "\n")
        for fname, body in (("pressure", ["type fixedValue;", "value 101325;"]),
                            ("velocities", ["type fixedValue;", "value (0 0 0);"]),
                            ("temperature", ["type fixedValue;", "value 300;"]),
                            ("water", ["type fixedValue;", "value 0;"]), ("air", ["type fixedValue;", "value 0;"])):
            write_initial_file(os.path.join(d, "constant", fname), body, m.patch_names)
        os.makedirs(os.path.join(d, "save", "0"))
        write_vtu(os.path.join(d, "save", "0", "plot.0.vtu"), m,
                  semo_fields(np.full(nc, 101325.0), np.zeros(nc), np.zeros(nc), np.zeros(nc), VF=vf,
                              T=np.full(nc, 300.0)), fmt="binary")
        r = subprocess.run([SEMO_COMPRESSVF, "-c", "0.4", "-i", "6"], capture_output=True, text=True, timeout=120,
                           cwd=d)
        outs = [os.path.join(d, "save", k, "plot.0.vtu") for k in os.listdir(os.path.join(d, "save"))
                if k.endswith("_CompressVF")]
        if not outs:
            raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
        _, f, _ = read_vtu(outs[0])
        res = [float(ln.split("residual = ")[1]) for ln in r.stdout.splitlines() if "residual = " in ln]
    np.savez_compressed(os.path.join(HERE, "compressvf_box865.npz"), cvf_in=vf, cvf_out=f["volume-fraction-water"],
                        cvf_cfl=np.array(0.4), cvf_iters=np.array(6), cvf_residuals=np.array(res), **mesh_arrays(m))
    print("compressvf_box865: residuals", res)


CASES = {"cavity4_jitter_random": _cavity4, "box654_jitter_2x2": _box654,
         "cavity8_plain": _cavity8, "orifice_lid": _orifice, "vof_box654_2x2": _vof_box654,
         "vof_tet433_2x1": _vof_tet, "sigma_box766_2x1": _sigma_box,
         "bc_open_box654_2x2": _bc_open, "bc_vof_box654_2x1": _bc_vof,
         "poly20_stack3": _poly20, "vtu_tet221": _vtu, "initial_cube6": _initial, "les": _les,
         "lesrun": _les_runs, "flux": _flux, "linsol": _linsol, "rhs": _rhs, "update": _update, "compressvf": _compressvf}


def main(names):
    if not os.path.exists(REF):
        raise SystemExit("build the reference first: make -f oracle/ref.mk REF=/root/reference")
    for nm in names or CASES:
        CASES[nm]()


if __name__ == "__main__":
    main(sys.argv[1:])
