"""Throughput of the SEMO assemble-and-solve hot path on one MI355X.

Workload (the 2.1M-cell configuration): lid-driven cavity on the unit cube,
128^3 hexahedra with every interior vertex jittered by a seeded U(-0.1h, 0.1h),
RCM-reordered, rho = 1000, mu = 1e-3, lid U = (1,0,0), dt = 0.01.  A step is
one SIMPLE iteration of SEMO's pressure-based solver: face reconstruction,
gradients, momentum assembly + three solves, pressure assembly + solve, p'
gradient and the correction, each linear solve to a relative residual of
1e-8 in FP64.  The FP64 path (FP64 AMG-PCG) and the mixed path (FP32 AMG
preconditioner inside the FP64 flexible CG) run the same steps from the same
state; ``value`` is the mixed path's Mcell-iterations/s, timed as a caller
runs the library (V-cycles replayed as HIP graphs).  A second pass over the
same number of steps brackets every kernel with HIP events (kernels issued
one by one) and gives the per-kernel figures and the roofline line.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Mcell-iter/s at 2.1M cells, mixed vs FP64 GPU at 1e-8 residual; adjoint grads/s"


def run_path(mesh, case, precision, steps, warmup, profile_kernels=True, cheb_degree=1):
    from semo_mesh_amd.solver import GpuSolver
    g = GpuSolver(0)
    t0 = time.perf_counter()
    g.upload_mesh(mesh, reorder=True)
    g.set_case(case)
    g.set_solver(precision, tol_p=1e-8, tol_u=1e-8, maxit_p=1000, maxit_u=1000, cheb_degree=cheb_degree)
    t_upload = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(warmup):                       # the first one builds the AMG hierarchy
        g.incompressiblePressureBased(1)
    t_warm = time.perf_counter() - t0
    amg_setup_ms = g.solve_stats("pressure")["setup_ms"]
    # timed pass: the library as a caller runs it (V-cycles replayed as graphs)
    t0 = time.perf_counter()
    for _ in range(steps):
        g.incompressiblePressureBased(1)
    t = time.perf_counter() - t0
    fields = {f: g.get_field(f) for f in ("P", "U", "V", "W")}
    # kernel pass: the same number of steps with a HIP event pair around every
    # kernel launch (kernels issued one by one), for the per-kernel figures
    prof, t_prof = {}, None
    if profile_kernels:
        g.profile(True)
        g.profile_reset()
        t0 = time.perf_counter()
        for _ in range(steps):
            g.incompressiblePressureBased(1)
        t_prof = time.perf_counter() - t0
        prof = g.profile_stats()
        g.profile(False)
    g.close()
    return dict(seconds=t, upload_s=t_upload, warmup_s=t_warm, amg_setup_ms=amg_setup_ms, prof=prof, fields=fields,
                seconds_kernel_pass=t_prof)


# stages of a SIMPLE iteration by kernel tag (every launch carries one)
STAGES = (
    ("pressure V-cycle", lambda t: t.startswith("amg_") and t not in ("amg_values_changed", "amg_refresh_L0")),
    ("pressure Krylov (FP64)", lambda t: t.startswith(("fcg_", "spmv_fp64", "spmv_xpby", "true_resid", "krylov_"))),
    ("momentum BiCGStab (FP64)", lambda t: t.startswith(("bicg3_", "spmv3_", "jacobi_dinv"))),
    ("assembly, gradients, corrections", lambda t: True),
)


def stage_table(prof, steps):
    """ms per step of each stage from the kernel pass's tags."""
    out = {name: 0.0 for name, _ in STAGES}
    for tag, v in prof.items():
        for name, test in STAGES:
            if test(tag):
                out[name] += v["ms"]
                break
    return {k: round(v / steps, 3) for k, v in out.items()}


def dam_break_leg(steps, warmup, nx=96):
    """Two-phase VOF dam break on the tetrahedral mesh of the 1.3M-cell
    configuration: 4 x 2 x 2 m box, nx x nx/2 x nx/2 jittered cubes split into
    6 tets each, water block x < 1.2, z < 1, gravity, surface tension 0.07
    N/m with curvature from alpha, dt = 1e-4; per step one VOF iteration
    (MSTACS + volume-fraction equation), curvature and sources, one
    momentum/pressure corrector, mixed-precision pressure solve."""
    from semo_mesh_amd.cases import dam_break_case
    from semo_mesh_amd.mesh import tet_mesh
    from semo_mesh_amd.solver import GpuSolver
    t0 = time.perf_counter()
    m = tet_mesh(nx, nx // 2, nx // 2, hi=(4.0, 2.0, 2.0), jitter=0.1)
    t_mesh = time.perf_counter() - t0
    nc = m.n_cells
    pts = m.points[m.cell_pts.reshape(nc, 4)].mean(axis=1)
    vf = ((pts[:, 0] < 1.2) & (pts[:, 2] < 1.0)).astype(np.float64)
    case = dam_break_case(dt=1e-4, n_vof=1, urf_vf=1.0)
    case.sigma = 0.07
    g = GpuSolver(0)
    g.upload_mesh(m, reorder=True)
    g.set_case(case)
    g.set_solver("mixed", tol_p=1e-8, tol_u=1e-8, maxit_p=1000, maxit_u=1000)
    g.set_field("VF", vf)
    for _ in range(warmup):
        g.incompressiblePressureBased(1)
    m0, vol = g.volume_integral("VF")          # sum alpha_i V_i, FP64 on the device
    t0 = time.perf_counter()
    for _ in range(steps):
        g.incompressiblePressureBased(1)
    t = time.perf_counter() - t0
    m1, _ = g.volume_integral("VF")
    vf1 = g.get_field("VF")
    g.close()
    return {"workload": f"VOF dam break, {nx}x{nx // 2}x{nx // 2} cubes x 6 tets ({nc} cells), "
                        "rho 1000/1.2, mu 1e-3/1.8e-5, g = -9.81 z, sigma 0.07, dt 1e-4",
            "cells": nc, "steps": steps, "warmup": warmup,
            "value": round(nc * steps / t / 1e6, 3), "unit": "Mcell-step/s",
            "ms_per_step": round(1e3 * t / steps, 3), "mesh_s": round(t_mesh, 1),
            "water_volume_m3": [m0, m1], "domain_volume_m3": vol,
            "water_volume_change_rel": abs(m1 - m0) / m0,
            "mass_check": "sum alpha_i V_i in FP64, reduced on the device, before and after the timed steps",
            "alpha_min": float(vf1.min()), "alpha_max": float(vf1.max())}


def adjoint_leg(n=64, n_iter=20, reps=3, fd_cells=32):
    """Discrete adjoint of the 262k-cell configuration as written: the
    cavity physics of the 32^3 configuration (rho 1000, mu 1e-3, dt 0.01,
    lid 1 m/s) on a jittered n^3 mesh, Darcy alpha_p ~ U[0, 10) seeded,
    J = int p dA (lid) - int p dA (bottom), n_iter checkpointed outer
    iterations of one time step (recomputed in the reverse sweep).  grads/s =
    1 / wall time of one gradient (forward + reverse sweep) on the FP64 path,
    solves to 1e-8; the mixed-precision path is reported beside it.  Check:
    central differences at fd_cells seeded cells with FP64 solves at 1e-14,
    each cell's quotient at h = 1 and h/2 (their spread bounds the truncation
    error; Richardson (4 D(h/2) - D(h)) / 3 removes the h^2 term) and at the
    configuration's h = 1e-6; and directional derivatives dJ/dalpha . d over
    three seeded whole-field directions d, against central differences at
    h = 1 and 1/2 with Richardson.  The torch.autograd path (forward +
    backward on a CUDA tensor) is timed too.  At rho 1000 the Darcy term (alpha_p <= 10) is
    a 1e-4 perturbation of the momentum diagonal 1.5 rho V / dt, so a
    per-cell dJ/dalpha_p is ~1e-12 against J ~ 1e-3: J's round-off (~1e-17)
    swamps a quotient at h = 1e-6 (it resolves ~1e-5 of it), while J is
    linear in alpha_p to ~1e-5 over h = 1."""
    from semo_mesh_amd.cases import cavity_case
    from semo_mesh_amd.mesh import cavity_mesh
    from semo_mesh_amd.solver import GpuSolver
    m = cavity_mesh(n, jitter=0.1, seed=20261205)
    case = cavity_case()
    alpha = 10.0 * np.random.default_rng(20261205).random(m.n_cells)
    W = {"zmax": 1.0, "zmin": -1.0}

    def ctx(prec, tol):
        g = GpuSolver(0)
        g.upload_mesh(m, reorder=True)
        g.set_case(case)
        g.set_solver(prec, tol_p=tol, tol_u=tol, maxit_p=1000, maxit_u=1000)
        g.set_field("alphaP", alpha)
        return g

    def reset(g):
        # every run starts from the same state (zero fields), so the time
        # levels the step keeps are the same too
        for f in ("U", "V", "W", "P"):
            g.set_field(f, np.zeros(m.n_cells))

    def timed(prec):
        g = ctx(prec, 1e-8)
        reset(g)
        g.adjointAlphaP(2, W)                       # warm-up: AMG hierarchy, work arrays
        times, ftimes = [], []
        for _ in range(reps):
            reset(g)
            t0 = time.perf_counter()
            g.adjointAlphaP(n_iter, W)
            times.append(time.perf_counter() - t0)
            reset(g)
            t0 = time.perf_counter()
            g.incompressiblePressureBased(n_iter)
            ftimes.append(time.perf_counter() - t0)
        g.close()
        return min(times), min(ftimes)

    t, t_fwd = timed("fp64")
    t_mx, t_fwd_mx = timed("mixed")
    t_torch = torch_path(m, alpha, n_iter, W, ctx, reset, reps)
    # gradient check: one context at tight tolerance, reused for every J
    g = ctx("fp64", 1e-14)
    reset(g)
    J, grad = g.adjointAlphaP(n_iter, W)

    def J_of(a):
        reset(g)
        g.set_field("alphaP", a)
        g.incompressiblePressureBased(n_iter)
        return g.objective(W)

    cells = np.random.default_rng(7).choice(m.n_cells, fd_cells, replace=False)
    h = 1.0
    rows = []
    for c in cells:
        q = []
        for step in (h, h / 2, 1e-6):
            e = np.zeros(m.n_cells)
            e[c] = step
            q.append((J_of(alpha + e) - J_of(alpha - e)) / (2 * step))
        rows.append([int(c), float(grad[c])] + [float(x) for x in q])
        if len(rows) % 8 == 0:
            print(f"[bench] adjoint fd check: {len(rows)}/{fd_cells} cells", file=sys.stderr, flush=True)
    # directional derivatives: dJ/dalpha . d against central differences of
    # J(alpha + h d) over whole-field directions d, whose J changes are large
    # against J's round-off (a per-cell change is not)
    rng = np.random.default_rng(11)
    u = rng.random(m.n_cells)
    dirs = {"U[0,1)": u, "U[-1,1)": 2.0 * rng.random(m.n_cells) - 1.0,
            "sign(dJ) U[0,1)": np.sign(grad) * rng.random(m.n_cells)}
    drows = {}
    for name, d in dirs.items():
        q = [(J_of(alpha + hh * d) - J_of(alpha - hh * d)) / (2 * hh) for hh in (1.0, 0.5)]
        gd = float(grad @ d)
        rich = (4.0 * q[1] - q[0]) / 3.0
        drows[name] = {"adjoint": gd, "fd_h1": q[0], "fd_h0.5": q[1], "rel_err_richardson": abs(gd - rich) / abs(rich),
                       "rel_err_h0.5": abs(gd - q[1]) / abs(q[1])}
    g.close()
    r = np.array(rows)
    adj, d1, d2, d6 = r[:, 1], r[:, 2], r[:, 3], r[:, 4]
    rich = (4.0 * d2 - d1) / 3.0
    scale = np.maximum(np.abs(rich), 1e-2 * np.abs(rich).max())
    return {"workload": f"discrete adjoint, jittered {n}^3 cavity ({m.n_cells} cells), rho 1000, mu 1e-3, "
                        f"dt 0.01, Darcy alpha_p ~ U[0,10), {n_iter} checkpointed SIMPLE iterations, "
                        "J = int p dA(lid) - int p dA(bottom), FP64 solves to 1e-8",
            "cells": m.n_cells, "iterations": n_iter, "value": round(1.0 / t, 3), "unit": "grads/s",
            "ms_per_gradient": round(1e3 * t, 1), "forward_ms": round(1e3 * t_fwd, 1),
            "gradient_over_forward": round(t / t_fwd, 2),
            "mixed": {"grads_per_s": round(1.0 / t_mx, 3), "ms_per_gradient": round(1e3 * t_mx, 1),
                      "forward_ms": round(1e3 * t_fwd_mx, 1)},
            "J": J,
            "fd_check": {"cells": fd_cells, "solves": "FP64 to 1e-14", "h": [h, h / 2, 1e-6],
                         "max_rel_err_richardson": float(np.max(np.abs(adj - rich) / scale)),
                         "max_rel_err_h": float(np.max(np.abs(adj - d1) / scale)),
                         "max_rel_err_h_half": float(np.max(np.abs(adj - d2) / scale)),
                         "max_rel_err_h_1e-6": float(np.max(np.abs(adj - d6) / scale)),
                         "max_rel_spread_h_vs_h_half": float(np.max(np.abs(d1 - d2) / scale)),
                         "per_cell": "[cell, adjoint, D(1), D(0.5), D(1e-6)]",
                         "rows": rows,
                         "directional": drows,
                         "max_rel_err_directional": max(v["rel_err_richardson"] for v in drows.values())},
            "torch_autograd": {"grads_per_s": round(1.0 / t_torch, 3), "ms_per_gradient": round(1e3 * t_torch, 1),
                               "path": "semo_mesh_amd.autograd.pressure_objective on a CUDA tensor, forward + "
                                       "backward, one persistent solver (FP64 to 1e-8)"}}


def torch_path(m, alpha, n_iter, W, ctx, reset, reps):
    """Wall time of one gradient through the torch.autograd Function: J and
    J.backward() on a CUDA tensor, one persistent solver."""
    import torch
    from semo_mesh_amd.autograd import pressure_objective
    g = ctx("fp64", 1e-8)
    a = torch.tensor(alpha, dtype=torch.float64, device="cuda", requires_grad=True)
    try:
        times = []
        for r in range(reps + 1):                # the first builds the AMG hierarchy
            a.grad = None
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            J = pressure_objective(a, g, n_iter, W, reset)
            J.backward()
            torch.cuda.synchronize()
            if r:
                times.append(time.perf_counter() - t0)
    finally:
        g.close()
    return min(times)


def coupled_leg(nx=256, ny=64, nz=64, n_iter=10, reps=2):
    """Block-coupled (BSR4) pressure-velocity iterations on the 1.05M-cell
    channel configuration (semo_mesh_amd.cases.channel_workload): 4 x 1 x 1
    box, nx x ny x nz jittered hexes, air (rho 1.177, mu 1.85e-5), inlet U =
    Mach 0.3 at 300 K (gamma 1.4, c = 347.2 m/s) with a seeded sinusoidal
    perturbation of 2 % amplitude (per-face fixedValue-U values), outlet
    p = 0, noSlip walls, dt at Courant 1 on the inlet velocity, uniform
    initial velocity, under-relaxation 0.7 (u) and 0.3 (p); per step n_iter
    coupled outer iterations (FP64 BiCGStab preconditioned by a SIMPLE-type
    block factorisation with an FP32 pressure V-cycle, solves to 1e-8).
    The flow is computed incompressible: the ideal-gas equation of state
    belongs to the density-based solver family, not to this path."""
    from semo_mesh_amd.cases import channel_workload
    from semo_mesh_amd.solver import GpuSolver
    m, case, u_in, ufix = channel_workload(nx, ny, nz)
    g = GpuSolver(0)
    g.upload_mesh(m, reorder=True)
    g.set_case(case)
    g.set_solver("mixed", tol_p=1e-8, tol_u=1e-8, maxit_p=1000, maxit_u=500)
    g.set_field("U", np.full(m.n_cells, u_in))  # uniform start
    g.set_boundary_velocity(ufix)
    g.coupledPressureBased(1)                   # warm-up: block pattern, pressure hierarchy
    times, its, its_max = [], [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        norms = g.coupledPressureBased(n_iter)
        times.append(time.perf_counter() - t0)
        st = g.solve_stats("momentum")
        its.append(st["iters"][1] / n_iter)
        its_max.append(st["iters"][2])
    st = g.solve_stats("momentum")
    u_mean = float(g.get_field("U").mean())
    g.close()
    t = min(times)
    return {"workload": f"block-coupled p-U iterations, jittered {nx}x{ny}x{nz} channel ({m.n_cells} cells), "
                        "inlet Mach 0.3 air with a seeded 2% sinusoidal perturbation, BSR4 BiCGStab + "
                        "SIMPLE-type block preconditioner (FP32 pressure V-cycle) to 1e-8",
            "cells": m.n_cells, "iterations_per_step": n_iter, "value": round(m.n_cells * n_iter / t / 1e6, 3),
            "unit": "Mcell-iter/s", "ms_per_iteration": round(1e3 * t / n_iter, 2),
            "krylov_iterations_per_outer": round(its[times.index(t)], 1),
            "krylov_iterations_max": max(its_max), "pressure_amg_levels": st["n_levels"],
            "last_rel_res": st["rel_res"][0], "last_update_norm": float(norms[-1]),
            "mean_u_over_inlet": round(u_mean / u_in, 4)}


def linsol_inputs(n_cells, nsp, h, seed=20261203):
    """Seeded density-based cell states [n][12 + 3 nsp] and residuals
    [n][4 + nsp] for SEMO's implicit updates: ideal-gas air cells (R 287,
    gamma 1.4) and a third liquid-like cells (rho ~ 1000, C ~ 1400), Mach up
    to 0.8, Ur = |U| clipped to [0.01 C, C], dtPseudo at a pseudo CFL of 1 to
    5 on the cell size h; residuals scaled per equation."""
    rng = np.random.default_rng(seed)
    n = n_cells
    liquid = rng.random(n) < 0.35
    P = 1e5 * (0.8 + 0.4 * rng.random(n))
    T = 280.0 + 40.0 * rng.random(n)
    Rg, gam = 287.0, 1.4
    cp = gam * Rg / (gam - 1.0)
    rho = np.where(liquid, 1000.0 * (1.0 + 0.02 * rng.random(n)), P / (Rg * T))
    C = np.where(liquid, 1300.0 + 200.0 * rng.random(n), np.sqrt(gam * Rg * T))
    d = rng.normal(size=(n, 3)); d /= np.linalg.norm(d, axis=1)[:, None]
    u = d * (0.8 * rng.random(n) * C)[:, None]
    um = np.linalg.norm(u, axis=1)
    st = np.column_stack([P, u, T, rho, C, np.where(liquid, 4180.0, cp) * T + 0.5 * um * um,
                          np.where(liquid, 1.0 / (C * C), 1.0 / (Rg * T)),
                          np.where(liquid, 1e-3 * (0.5 + rng.random(n)), 0.0),
                          np.where(liquid, -0.2 - 0.2 * rng.random(n), -P / (Rg * T * T)),
                          np.where(liquid, 4180.0, cp), np.clip(um, 0.01 * C, C),
                          (1.0 + 4.0 * rng.random(n)) * h / (um + C)])
    Y = rng.random((n, nsp)); Y /= Y.sum(axis=1)[:, None]
    st = np.column_stack([st, Y, rho[:, None] * (rng.random((n, nsp - 1)) * 4.0 - 2.0),
                          1e4 * (rng.random((n, nsp - 1)) * 2.0 - 1.0)])
    scale = np.array([1.0, 100.0, 100.0, 100.0, 1e5] + [0.1] * (nsp - 1))
    return st, rng.normal(size=(n, 4 + nsp)) * scale


def lusgs_leg(n=128, nsp=2, reps=5, cpu_n=48):
    """SEMO's density-based pseudo step on the jittered n^3 cavity mesh with
    seeded gas / liquid-like states and two species (6 equations): the
    residual (calcRHS, kernel times) and the implicit updates
    (solvers/linearSolver.cpp): LU-SGS (calcLinearSolverLUSGS, forward and
    backward sweeps in lib_RCM's order, level-scheduled) and the
    point-implicit update (calcLinearSolver), state and residuals resident
    on the device (semo_linear_solver_device); the compiled reference's
    calcLinearSolverLUSGS on a jittered cpu_n^3 sample, one core."""
    import ctypes as C
    from semo_mesh_amd.mesh import cavity_mesh
    from semo_mesh_amd.solver import GpuSolver
    m = cavity_mesh(n, jitter=0.1)
    nc = m.n_cells
    st, r0 = linsol_inputs(nc, nsp, 1.0 / n)
    g = GpuSolver(0)
    g.upload_mesh(m, reorder=True)
    perm, _ = g.permutation()
    hip = C.CDLL("libamdhip64.so.7")
    bufs = []

    def put(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert hip.hipMalloc(C.byref(p), C.c_size_t(a.nbytes)) == 0
        assert hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0
        bufs.append(p)
        return p

    ds, dr0, dr = put(st[perm]), put(r0[perm]), put(r0[perm])
    nbytes = C.c_size_t(r0.nbytes)
    out = {"workload": f"density-based implicit updates, jittered {n}^3 hex ({nc} cells), 2 species "
                       f"(6 equations), seeded gas / liquid-like states, state and residuals resident"}
    try:
        for method, key in ((2, "lusgs"), (0, "point")):
            def call():
                assert hip.hipMemcpy(dr, dr0, nbytes, 3) == 0
                assert hip.hipDeviceSynchronize() == 0
                t0 = time.perf_counter()
                g._ok(g._L.semo_linear_solver_device(g._h, method, nsp, 1e-3, ds, dr))
                return time.perf_counter() - t0
            first = call()
            times = [call() for _ in range(reps)]
            g.profile(True); g.profile_reset()
            call()
            prof = g.profile_stats()
            g.profile(False)
            t = min(times)
            leg = {"ms": round(1e3 * t, 3), "Mcell_per_s": round(nc / t / 1e6, 2),
                   "first_call_ms": round(1e3 * first, 1),
                   "kernels": {k: {"launches": v["count"], "ms": round(v["ms"], 3),
                                   "GBs": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 else None}
                               for k, v in prof.items()}}
            if key == "lusgs":
                leg["sweep_levels"] = prof.get("lusgs_sweep", {}).get("count")
                leg["note"] = ("first call levels the sweeps on the host (once per mesh); the level "
                               "kernels are latency-bound: one launch per level")
            out[key] = leg
        hout = np.empty_like(r0)
        assert hip.hipMemcpy(hout.ctypes.data_as(C.c_void_p), dr, nbytes, 2) == 0
        out["point_finite"] = bool(np.isfinite(hout).all())
        # the residual the updates consume: calcRHS (SLAU2 + viscous flux per
        # face, summed per cell, dual-time source, gravity, scaling), kernel
        # times of one call with host arrays (the transfers are not timed)
        rng = np.random.default_rng(7)
        nf = m.n_faces
        fs = np.empty((2, nf, 8 + nsp))
        for k in range(2):
            fs[k, :, 0] = 1e5 * (0.8 + 0.4 * rng.random(nf))
            fs[k, :, 1:4] = rng.normal(size=(nf, 3)) * 50.0
            fs[k, :, 4] = 280.0 + 40.0 * rng.random(nf)
            fs[k, :, 5] = fs[k, :, 0] / (287.0 * fs[k, :, 4])
            fs[k, :, 6] = np.sqrt(1.4 * 287.0 * fs[k, :, 4])
            fs[k, :, 7] = 1004.5 * fs[k, :, 4] + 0.5 * (fs[k, :, 1:4] ** 2).sum(axis=1)
            fs[k, :, 8] = 1.8e-5
            fs[k, :, 9:] = rng.random((nf, nsp - 1))
        N = 4 + nsp
        Q = np.column_stack([st[:, 5], st[:, 5:6] * st[:, 1:4], st[:, 5] * st[:, 7] - st[:, 0],
                             st[:, 5:6] * st[:, 14:13 + nsp]])
        cells = np.column_stack([st[:, 5], st[:, 1:4], st[:, 7], st[:, 0], st[:, 13], np.ones(nc),
                                 st[:, 14:13 + nsp], Q * 1.01, Q * 0.99])
        g.calcRHS(1e-3, np.array([0.0, 0.0, -9.81]), fs[0], fs[1], cells)   # first launches load the kernels
        g.profile(True); g.profile_reset()
        t0 = time.perf_counter()
        rr = g.calcRHS(1e-3, np.array([0.0, 0.0, -9.81]), fs[0], fs[1], cells)
        t_rhs = time.perf_counter() - t0
        prof = g.profile_stats()
        g.profile(False)
        out["rhs"] = {"kernels": {k: {"ms": round(v["ms"], 3),
                                      "GBs": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 else None}
                                  for k, v in prof.items() if k.startswith("rhs")},
                      "kernel_ms": round(sum(v["ms"] for k, v in prof.items() if k.startswith("rhs")), 3),
                      "call_ms_with_transfers": round(1e3 * t_rhs, 1), "finite": bool(np.isfinite(rr).all()),
                      "faces": nf}
    finally:
        for p in bufs:
            hip.hipFree(p)
        g.close()
    ref = os.path.join(ROOT, "oracle", "_ref", "semo_ref")
    if os.path.exists(ref):
        mc = cavity_mesh(cpu_n, jitter=0.1)
        sc, rc = linsol_inputs(mc.n_cells, nsp, 1.0 / cpu_n)
        from semo_mesh_amd.cases import cavity_case
        with tempfile.TemporaryDirectory() as d:
            mc.write_raw(os.path.join(d, "mesh.raw"))
            cavity_case().write_ref_case(os.path.join(d, "case.txt"), steps=0, correctors=1)
            fn = os.path.join(d, "state.f64")
            with open(fn, "wb") as fh:
                sc.tofile(fh)
                rc.tofile(fh)
            with open(os.path.join(d, "case.txt"), "a") as fh:
                fh.write(f"linsol {fn} 0.001\n")
            r = subprocess.run([ref, os.path.join(d, "mesh.raw"), os.path.join(d, "case.txt"),
                                os.path.join(d, "out"), "nodump"], capture_output=True, text=True, timeout=600)
            line = [x for x in r.stdout.splitlines() if x.startswith("SEMO_REF_DONE linsol")]
            if line:
                kv = dict(tok.split("=") for tok in line[0].split()[2:])
                out["cpu_reference"] = {
                    "lusgs_Mcell_per_s": round(mc.n_cells / float(kv["lusgs_s"]) / 1e6, 4),
                    "point_Mcell_per_s": round(mc.n_cells / float(kv["point_s"]) / 1e6, 4),
                    "cores": 1, "kind": "reference",
                    "sample": f"compiled SEMO calcLinearSolverLUSGS / calcLinearSolver on a jittered {cpu_n}^3 "
                              f"cavity ({mc.n_cells} cells), one call each: {float(kv['lusgs_s']):.2f} s / "
                              f"{float(kv['point_s']):.2f} s"}
    return out


def cpu_baseline(n=48):
    """The compiled reference (SEMO + AMGCL, 1 MPI rank, no OpenMP) on a
    bounded sample: one SIMPLE iteration of the same cavity at n^3."""
    from semo_mesh_amd.cases import cavity_case
    from semo_mesh_amd.mesh import cavity_mesh
    ref = os.path.join(ROOT, "oracle", "_ref", "semo_ref")
    m = cavity_mesh(n, jitter=0.1)
    with tempfile.TemporaryDirectory() as d:
        m.write_raw(os.path.join(d, "mesh.raw"))
        cavity_case().write_ref_case(os.path.join(d, "case.txt"), steps=1, correctors=1)
        if os.path.exists(ref):
            r = subprocess.run([ref, os.path.join(d, "mesh.raw"), os.path.join(d, "case.txt"),
                                os.path.join(d, "out"), "nodump"], capture_output=True, text=True,
                               timeout=600)
            line = [x for x in r.stdout.splitlines() if x.startswith("SEMO_REF_DONE")]
            if line:
                kv = dict(tok.split("=") for tok in line[0].split()[1:])
                run_s = float(kv["run_s"])
                return dict(value=m.n_cells / run_s / 1e6, unit="Mcell-iter/s", cores=1, kind="reference",
                            sample=f"compiled SEMO+AMGCL, 1 rank, one SIMPLE iteration of the jittered "
                                   f"{n}^3 cavity ({m.n_cells} cells), {run_s:.2f} s; mesh geometry / "
                                   f"LSQ / RCM set-up excluded",
                            extrapolated=f"the per-cell rate of the {n}^3 sample stands for the 128^3 "
                                         f"configuration (not timed there: the reference has no OpenMP, "
                                         f"and its multi-rank path needs a prebuilt METIS)")
        # fallback: the repo's CPU oracle (single thread)
        from oracle.oracle import Oracle
        o = Oracle(m, cavity_case())
        o.set_solver(1e-8, 5000)
        t0 = time.perf_counter()
        o.step(1)
        run_s = time.perf_counter() - t0
        return dict(value=m.n_cells / run_s / 1e6, unit="Mcell-iter/s", cores=1, kind="port",
                    sample=f"CPU oracle, one SIMPLE iteration of the jittered {n}^3 cavity, {run_s:.2f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-vof", action="store_true", help="skip the two-phase dam-break leg")
    ap.add_argument("--vof-steps", type=int, default=50)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the discrete-adjoint leg")
    ap.add_argument("--no-coupled", action="store_true", help="skip the block-coupled (BSR4) leg")
    ap.add_argument("--no-lusgs", action="store_true", help="skip the density-based LU-SGS leg")
    ap.add_argument("--cheb-degree", type=int, default=1, help="Chebyshev smoother degree of the pressure AMG")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run the library's streaming copy of MIB MiB (read MIB, write MIB; "
                         "1 + 3 launches): the reference kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.copy_probe:
        from semo_mesh_amd.solver import GpuSolver
        n = args.copy_probe * (1 << 20) // 8
        g = GpuSolver(0)
        ms = g.stream_copy(n, 3)
        g.close()
        print(json.dumps({"copy_probe_bytes_read": 8 * n, "copy_probe_bytes_written": 8 * n, "launches": 4,
                          "ms_per_copy": ms}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(domain decomposition over RCCL is design-only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if not args.no_adjoint:
        # the adjoint leg times the torch.autograd path: torch's libraries
        # must be in the process before libsemo_gpu.so
        import semo_mesh_amd.autograd  # noqa: F401

    from semo_mesh_amd.cases import cavity_case
    from semo_mesh_amd.mesh import cavity_mesh
    t0 = time.perf_counter()
    mesh = cavity_mesh(args.n, jitter=0.1)
    t_mesh = time.perf_counter() - t0
    case = cavity_case()
    print(f"[bench] mesh {args.n}^3: {mesh.n_cells} cells, {mesh.n_faces} faces ({t_mesh:.1f} s)",
          file=sys.stderr, flush=True)

    res64 = None
    if not args.no_fp64:
        res64 = run_path(mesh, case, "fp64", args.steps, args.warmup, cheb_degree=args.cheb_degree)
        print(f"[bench] fp64: {args.steps} steps in {res64['seconds']:.3f} s", file=sys.stderr, flush=True)
    resmx = run_path(mesh, case, "mixed", args.steps, args.warmup, cheb_degree=args.cheb_degree)
    print(f"[bench] mixed: {args.steps} steps in {resmx['seconds']:.3f} s", file=sys.stderr, flush=True)

    nc = mesh.n_cells
    value = nc * args.steps / resmx["seconds"] / 1e6
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mcell-iter/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * resmx["seconds"] / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 Krylov / f32 AMG", "data": "synthetic",
        "config": {"workload": f"jittered lid-driven cavity {args.n}^3 hex ({nc} cells), RCM, "
                               "1 SIMPLE iteration per step, solves to 1e-8",
                   "cells": nc, "faces": mesh.n_faces, "seed": 20261202},
    }
    out["setup"] = {"upload_geometry_rcm_s": round(resmx["upload_s"], 3),
                    "first_steps_s": round(resmx["warmup_s"], 3), "first_steps": args.warmup,
                    "amg_setup_ms": round(resmx["amg_setup_ms"], 1),
                    "note": "one-time: mesh upload, RCM, geometry and LSQ set-up on the device; the first "
                            "steps build the AMG hierarchy on the host (single thread) and capture the V-cycle graph"}
    if res64 is not None:
        v64 = nc * args.steps / res64["seconds"] / 1e6
        out["fp64"] = {"value": round(v64, 3), "ms_per_step": round(1e3 * res64["seconds"] / args.steps, 3)}
        out["speedup_mixed_vs_fp64"] = round(value / v64, 3)
        err = {}
        for f in ("P", "U", "V", "W"):
            a, b = resmx["fields"][f], res64["fields"][f]
            err[f] = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
        out["mixed_vs_fp64_rel_l2"] = err
        if res64["prof"] and resmx["prof"]:
            out["stages_ms_per_step"] = {"fp64": stage_table(res64["prof"], args.steps),
                                         "mixed": stage_table(resmx["prof"], args.steps),
                                         "timing": "kernel pass (HIP event pair per launch)"}
    # dominant kernel by device time inside the timed mixed run
    prof = resmx["prof"]
    if prof:
        # one kernel per tag; the coarse-level tag sums several kernels over
        # all coarse levels and is not a candidate
        single = {k: v for k, v in prof.items() if "coarse" not in k}
        name, st = max(single.items(), key=lambda kv: kv[1]["ms"])
        # HIP event pair around every launch inside the steps of the kernel
        # pass: the kernel's duration plus its dispatch gap
        avg_ms = st["ms"] / st["count"]
        per_launch = st["bytes"] / st["count"]
        achieved = per_launch / (avg_ms * 1e-3) / 1e9
        out["kernel_pass_ms_per_step"] = round(1e3 * resmx["seconds_kernel_pass"] / args.steps, 3)
        out["roofline"] = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1),
                           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                           "traffic": None, "avg_us": round(avg_ms * 1e3, 2),
                           "bytes_per_launch": per_launch, "launches": st["count"],
                           "timing": "HIP event pair per launch inside the steps"}
        out["kernels"] = {k: {"count": v["count"], "ms": round(v["ms"], 3),
                              "GB_per_step": round(v["bytes"] / args.steps / 1e9, 4),
                              "GBs": round(v["bytes"] / max(v["ms"], 1e-9) / 1e6, 1)} for k, v in prof.items()}
        # the whole step against the HBM roofline: algorithmic bytes of every
        # kernel of a step (every launch is tagged) over the timed step
        step_bytes = sum(v["bytes"] for v in prof.values()) / args.steps
        tagged_ms = sum(v["ms"] for v in prof.values()) / args.steps
        out["hbm_per_step"] = {
            "algorithmic_GB": round(step_bytes / 1e9, 3),
            "frac_over_step": round(step_bytes / (out["ms_per_step"] * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
            "kernel_ms": round(tagged_ms, 3),
            "frac_over_kernel_time": round(step_bytes / (tagged_ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
            "note": "kernel_ms: sum of the per-launch HIP event times of the kernel pass (launches issued "
                    "one by one); ms_per_step - kernel_ms is host time and launch gaps"}
        # the practical streaming ceiling: the library's own copy kernel on
        # 1 GiB (1 GiB read + 1 GiB written), average of 10 copies
        from semo_mesh_amd.solver import GpuSolver
        gc = GpuSolver(0)
        ms_copy = gc.stream_copy(1 << 27, 10)
        gc.close()
        copy_gbs = 2.0 * (1 << 30) / (ms_copy * 1e-3) / 1e9
        out["roofline"]["copy_ceiling_GBs"] = round(copy_gbs, 1)
        out["roofline"]["frac_of_copy"] = round(achieved / copy_gbs, 4)
        traffic = _measured_traffic(name)
        if traffic:
            out["roofline"]["traffic"] = traffic["bytes"]
            out["roofline"]["traffic_source"] = traffic["source"]
            out["roofline"]["hip_kernel"] = traffic["kernel"]
            out["roofline"]["rocprof_avg_us"] = traffic["trace_avg_us"]
    if not args.no_vof:
        out["vof_dam_break"] = dam_break_leg(args.vof_steps, 2)
        print(f"[bench] dam break: {out['vof_dam_break']['ms_per_step']} ms/step", file=sys.stderr, flush=True)
    if not args.no_coupled:
        out["coupled_bsr4"] = coupled_leg()
        print(f"[bench] coupled: {out['coupled_bsr4']['ms_per_iteration']} ms per iteration", file=sys.stderr,
              flush=True)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_leg()
        print(f"[bench] adjoint: {out['adjoint']['ms_per_gradient']} ms per gradient", file=sys.stderr, flush=True)
    if not args.no_lusgs:
        out["lusgs"] = lusgs_leg()
        print(f"[bench] lusgs: {out['lusgs']['lusgs']['ms']} ms per update", file=sys.stderr, flush=True)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


def _measured_traffic(tag):
    """HBM bytes per launch of a profiled kernel, from the rocprofv3
    FETCH_SIZE / WRITE_SIZE runs summarised in profiles/hbm_traffic.json
    (counters scaled by the plain-copy calibration stated there)."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as fh:
            tab = json.load(fh)
    except (OSError, ValueError):
        return None
    e = tab.get("kernels", {}).get(tag)
    if not e:
        return None
    return {"bytes": e["fetch_bytes"] + e["write_bytes"], "source": tab.get("source", path),
            "kernel": e.get("kernel"), "trace_avg_us": e.get("trace_avg_us")}


if __name__ == "__main__":
    main()
