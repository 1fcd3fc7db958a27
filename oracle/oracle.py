"""ctypes front end of the CPU oracle (oracle/semo_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the semo_mesh_amd package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_P = np.ctypeslib.ndpointer
_D = _P(dtype=np.float64, flags="C_CONTIGUOUS")
_I = _P(dtype=np.int32, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liborc.so is not built (make -f oracle/Makefile)")
        L = C.CDLL(path)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int] * 4 + [_D, _I, _I, _I, _I, _I, _I, C.c_int, _I, _I]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_bc.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int, _D, C.c_int, C.c_double]
        L.orc_set_vf_points.argtypes = [C.c_void_p, C.c_int, _D]
        L.orc_set_case.argtypes = [C.c_void_p, _D, _D, C.c_double, _D, C.c_double, C.c_double,
                                   C.c_double, C.c_double, C.c_double]
        L.orc_set_solver.argtypes = [C.c_void_p, C.c_double, C.c_int]
        L.orc_set_vof.argtypes = [C.c_void_p, C.c_double, C.c_int]
        L.orc_set_time_control.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double]
        L.orc_time_step.restype = C.c_double
        L.orc_time_step.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_dt.argtypes = [C.c_void_p, C.c_double]
        for nm in ("orc_get_dt", "orc_courant"):
            getattr(L, nm).restype = C.c_double
            getattr(L, nm).argtypes = [C.c_void_p]
        for nm in ("orc_mstacs", "orc_assemble_vof"):
            getattr(L, nm).argtypes = [C.c_void_p]
        for nm in ("orc_vof_update", "orc_vof_norm"):
            getattr(L, nm).restype = C.c_double
            getattr(L, nm).argtypes = [C.c_void_p]
        L.orc_size.restype = C.c_long
        L.orc_size.argtypes = [C.c_void_p, C.c_char_p]
        L.orc_get.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]
        L.orc_set.argtypes = [C.c_void_p, C.c_char_p, _D]
        for nm in ("orc_eos", "orc_recon", "orc_gravity", "orc_assemble_momentum",
                   "orc_assemble_pressure", "orc_surface_tension", "orc_curvature"):
            getattr(L, nm).argtypes = [C.c_void_p]
        L.orc_lsq.argtypes = [C.c_void_p, _D, _D, _D]
        L.orc_minmax.argtypes = [C.c_void_p, _D, _D, _D]
        L.orc_gauss_green.argtypes = [C.c_void_p, C.c_int, _D, _D, _D]
        L.orc_les.argtypes = [C.c_void_p, C.c_int] + [_D] * 9
        L.orc_set_les.argtypes = [C.c_void_p, C.c_int]
        L.orc_linear_solver.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, _D, _D]
        L.orc_solve.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double)]
        L.orc_momentum_update.restype = C.c_double
        L.orc_momentum_update.argtypes = [C.c_void_p]
        L.orc_pressure_correct.restype = C.c_double
        L.orc_pressure_correct.argtypes = [C.c_void_p]
        L.orc_step.argtypes = [C.c_void_p, C.c_int, _D]
        L.orc_genrcm.argtypes = [C.c_int, _I, _I, _I, _I]
        L.orc_cell_graph.argtypes = [C.c_void_p, _I, _I]
        _LIB = L
    return _LIB


_INT_ARRAYS = {"stencil_ptr", "stencil_idx", "rows", "cols"}


class Oracle:
    """One mesh + case, mirroring SEMO's mesh/controls objects."""

    def __init__(self, mesh, case=None):
        L = lib()
        self.mesh = mesh
        self._h = L.orc_create(mesh.n_points, mesh.n_cells, mesh.n_faces, mesh.n_internal_faces,
                               np.ascontiguousarray(mesh.points, dtype=np.float64),
                               np.ascontiguousarray(mesh.owner, dtype=np.int32),
                               np.ascontiguousarray(mesh.neighbour, dtype=np.int32),
                               np.ascontiguousarray(mesh.face_ptr, dtype=np.int32),
                               np.ascontiguousarray(mesh.face_pts, dtype=np.int32),
                               np.ascontiguousarray(mesh.cell_ptr, dtype=np.int32),
                               np.ascontiguousarray(mesh.cell_pts, dtype=np.int32),
                               len(mesh.patch_names),
                               np.ascontiguousarray(mesh.patch_start, dtype=np.int32),
                               np.ascontiguousarray(mesh.patch_nfaces, dtype=np.int32))
        if case is not None:
            self.set_case(case)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    def set_case(self, case):
        L = lib()
        L.orc_set_case(self._h, np.array(case.rho, dtype=np.float64), np.array(case.mu, dtype=np.float64),
                       case.sigma, np.array(case.g, dtype=np.float64), case.dt, case.urf_mom,
                       case.urf_p, case.urf_pu, case.limit)
        L.orc_set_vof(self._h, getattr(case, "urf_vf", 1.0), getattr(case, "n_vof", 0))
        L.orc_set_time_control(self._h, 1 if getattr(case, "adjust_dt", False) else 0,
                               getattr(case, "max_co", 1.0), getattr(case, "max_dt", 1.0))
        for i, nm in enumerate(self.mesh.patch_names):
            b = case.bcs[nm]
            L.orc_set_bc(self._h, i, b.p_type, b.p_value, b.u_type,
                         np.array(b.u_value, dtype=np.float64), b.vf_type, b.vf_value)
        pts = getattr(case, "vf_points", None)
        if pts is not None:
            a = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 4)
            L.orc_set_vf_points(self._h, a.shape[0], a)
        les = getattr(case, "les_model", None)
        L.orc_set_les(self._h, {"WALE": 1, "smagorinsky": 2, "dynamicSmagorinsky": 3}[les] if les else 0)

    def set_solver(self, tol=1e-13, maxit=5000):
        lib().orc_set_solver(self._h, tol, maxit)

    def get(self, name):
        L = lib()
        n = L.orc_size(self._h, name.encode())
        if n < 0:
            raise KeyError(name)
        out = np.empty(n, dtype=np.int32 if name in _INT_ARRAYS else np.float64)
        L.orc_get(self._h, name.encode(), out.ctypes.data)
        return out

    def set(self, name, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        if lib().orc_set(self._h, name.encode(), arr) != 0:
            raise KeyError(name)

    # stages
    def eos(self): lib().orc_eos(self._h)
    def recon(self): lib().orc_recon(self._h)
    def gravity(self): lib().orc_gravity(self._h)
    def surface_tension(self): lib().orc_surface_tension(self._h)
    def curvature(self): lib().orc_curvature(self._h)
    def assemble_momentum(self): lib().orc_assemble_momentum(self._h)
    def assemble_pressure(self): lib().orc_assemble_pressure(self._h)

    def lsq(self, phi, face_r):
        g = np.empty(3 * self.mesh.n_cells)
        lib().orc_lsq(self._h, np.ascontiguousarray(phi, dtype=np.float64),
                      np.ascontiguousarray(face_r, dtype=np.float64), g)
        return g.reshape(-1, 3)

    def gauss_green(self, field, phi, face_r, grad):
        sel = {"P": 0, "U": 1, "V": 2, "W": 3, "VF": 4}[field]
        g = np.ascontiguousarray(grad, dtype=np.float64).reshape(-1).copy()
        lib().orc_gauss_green(self._h, sel, np.ascontiguousarray(phi, dtype=np.float64),
                              np.ascontiguousarray(face_r, dtype=np.float64), g)
        return g.reshape(-1, 3)

    def les(self, model, U, V, W, rho, face_rU, face_rV, face_rW):
        """muT, kSGS of LES model 1 (WALE), 2 (Smagorinsky) or 3 (dynamic)."""
        n = self.mesh.n_cells
        mu = np.empty(n); k = np.empty(n)
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (U, V, W, rho, face_rU, face_rV, face_rW)]
        lib().orc_les(self._h, model, *a, mu, k)
        return mu, k

    def linear_solver(self, method, nsp, time_step, state, res):
        """SEMO's density-based implicit updates: method "point"
        (calcLinearSolver), "single" (calcSingleLinearSolver) or "lusgs"
        (calcLinearSolverLUSGS, in lib_RCM's order).  state [nc][12 + 3 nsp],
        res [nc][4 + nsp]; returns the updated residuals."""
        m = {"point": 0, "single": 1, "lusgs": 2}[method]
        r = np.ascontiguousarray(res, dtype=np.float64).copy()
        st = np.ascontiguousarray(state, dtype=np.float64)
        if lib().orc_linear_solver(self._h, m, nsp, time_step, st, r) != 0:
            raise ValueError("linear_solver: bad method or nsp")
        return r

    def minmax(self, phi):
        mx = np.empty(self.mesh.n_cells); mn = np.empty(self.mesh.n_cells)
        lib().orc_minmax(self._h, np.ascontiguousarray(phi, dtype=np.float64), mx, mn)
        return mx, mn

    def solve(self, nrhs):
        r = C.c_double(0.0)
        it = lib().orc_solve(self._h, nrhs, C.byref(r))
        return it, r.value

    def time_step(self, adjust): return lib().orc_time_step(self._h, 1 if adjust else 0)
    def dt(self): return lib().orc_get_dt(self._h)
    def set_dt(self, dt): lib().orc_set_dt(self._h, float(dt))
    def courant(self): return lib().orc_courant(self._h)
    def mstacs(self): lib().orc_mstacs(self._h)
    def assemble_vof(self): lib().orc_assemble_vof(self._h)
    def vof_update(self): return lib().orc_vof_update(self._h)
    def vof_norm(self): return lib().orc_vof_norm(self._h)
    def momentum_update(self): return lib().orc_momentum_update(self._h)
    def pressure_correct(self): return lib().orc_pressure_correct(self._h)

    def step(self, correctors=1):
        norms = np.zeros(2 * correctors)
        lib().orc_step(self._h, correctors, norms)
        return norms

    def cell_graph(self):
        m = self.mesh
        ptr = np.zeros(m.n_cells + 1, dtype=np.int32)
        adj = np.zeros(2 * m.n_internal_faces, dtype=np.int32)
        lib().orc_cell_graph(self._h, ptr, adj)
        return ptr, adj

    def coo(self):
        return self.get("rows"), self.get("cols"), self.get("vals")


def genrcm(ptr, adj):
    n = len(ptr) - 1
    perm = np.zeros(n, dtype=np.int32); inv = np.zeros(n, dtype=np.int32)
    lib().orc_genrcm(n, np.ascontiguousarray(ptr, dtype=np.int32),
                     np.ascontiguousarray(adj, dtype=np.int32), perm, inv)
    return perm, inv


_LIB_AD = None


def lib_ad():
    """liborc_ad.so: the oracle compiled with the taped reverse-mode type
    (the CPU adjoint)."""
    global _LIB_AD
    if _LIB_AD is None:
        path = os.path.join(_HERE, "liborc_ad.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liborc_ad.so is not built (make -f oracle/Makefile)")
        L = C.CDLL(path)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int] * 4 + [_D, _I, _I, _I, _I, _I, _I, C.c_int, _I, _I]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_bc.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int, _D, C.c_int, C.c_double]
        L.orc_set_case.argtypes = [C.c_void_p, _D, _D, C.c_double, _D, C.c_double, C.c_double,
                                   C.c_double, C.c_double, C.c_double]
        L.orc_set_solver.argtypes = [C.c_void_p, C.c_double, C.c_int]
        L.orc_ad_set.argtypes = [C.c_void_p, C.c_char_p, _D]
        L.orc_ad_get.argtypes = [C.c_void_p, C.c_char_p, _D]
        L.orc_ad_forward.restype = C.c_double
        L.orc_ad_forward.argtypes = [C.c_void_p, C.c_int, _D]
        L.orc_ad_adjoint.restype = C.c_double
        L.orc_ad_adjoint.argtypes = [C.c_void_p, C.c_int, _D, _D]
        _LIB_AD = L
    return _LIB_AD


class AdjointOracle:
    """The CPU adjoint: one time step of n_iter SIMPLE outer iterations of a
    single-phase case with the Darcy term alpha_p (momentum diagonal +=
    alpha_p V, right sides -= alpha_p V U), its objective J = sum_patches
    w_p sum_f A_f p_f, and dJ/dalpha_p by reverse-mode differentiation of the
    oracle's own assembly (ad.hpp)."""

    def __init__(self, mesh, case):
        L = lib_ad()
        self.mesh = mesh
        self._h = L.orc_create(mesh.n_points, mesh.n_cells, mesh.n_faces, mesh.n_internal_faces,
                               np.ascontiguousarray(mesh.points, dtype=np.float64),
                               np.ascontiguousarray(mesh.owner, dtype=np.int32),
                               np.ascontiguousarray(mesh.neighbour, dtype=np.int32),
                               np.ascontiguousarray(mesh.face_ptr, dtype=np.int32),
                               np.ascontiguousarray(mesh.face_pts, dtype=np.int32),
                               np.ascontiguousarray(mesh.cell_ptr, dtype=np.int32),
                               np.ascontiguousarray(mesh.cell_pts, dtype=np.int32),
                               len(mesh.patch_names),
                               np.ascontiguousarray(mesh.patch_start, dtype=np.int32),
                               np.ascontiguousarray(mesh.patch_nfaces, dtype=np.int32))
        L.orc_set_case(self._h, np.array(case.rho, dtype=np.float64), np.array(case.mu, dtype=np.float64),
                       case.sigma, np.array(case.g, dtype=np.float64), case.dt, case.urf_mom,
                       case.urf_p, case.urf_pu, case.limit)
        for i, nm in enumerate(mesh.patch_names):
            b = case.bcs[nm]
            L.orc_set_bc(self._h, i, b.p_type, b.p_value, b.u_type,
                         np.array(b.u_value, dtype=np.float64), b.vf_type, b.vf_value)

    def __del__(self):
        if getattr(self, "_h", None):
            lib_ad().orc_destroy(self._h)
            self._h = None

    def set_solver(self, tol=1e-14, maxit=20000):
        lib_ad().orc_set_solver(self._h, tol, maxit)

    def set(self, name, arr):
        if lib_ad().orc_ad_set(self._h, name.encode(), np.ascontiguousarray(arr, dtype=np.float64)) != 0:
            raise KeyError(name)

    def get(self, name):
        out = np.empty(self.mesh.n_cells)
        if lib_ad().orc_ad_get(self._h, name.encode(), out) != 0:
            raise KeyError(name)
        return out

    def _w(self, weights):
        return np.array([float(weights.get(nm, 0.0)) for nm in self.mesh.patch_names])

    def forward(self, n_iter, weights):
        """J after one time step of n_iter outer iterations (the state advances)."""
        return lib_ad().orc_ad_forward(self._h, n_iter, self._w(weights))

    def adjoint(self, n_iter, weights):
        """(J, dJ/dalpha_p) of one time step of n_iter outer iterations."""
        g = np.empty(self.mesh.n_cells)
        J = lib_ad().orc_ad_adjoint(self._h, n_iter, self._w(weights), g)
        return J, g


def compress_volume_fraction(o, vf, cfl, iters, face_r=None):
    """SEMO's interface-compression utility (src/utility/compressVolumeFraction
    .cpp:121-228, the flux's zero tests at :284-285) restated over the oracle's mesh and LSQ: tau = min over the
    cells of min(V^0.3, V / max face area); per iteration the LSQ gradient of
    alpha (the cell-vertex stencil, boundary faces at face_r -- zero, as the
    utility holds them after loading), the unit normal n = grad / |grad|
    (zero where |grad| = 0), at each face UnF = wC n_o . n_f + (1 - wC) n_n .
    n_f (0 at boundary faces), the upwind alpha(1 - alpha) of the side UnF
    leaves (0 when that quantity q of either side has q (1 - q) < 1e-15),
    flux = alphaF UnF area taken from the owner and given to the neighbour
    in face order; alpha += cfl tau resi / V, clipped to [0, 1].  Returns
    the fractions and each iteration's sqrt(sum update^2)."""
    vf = np.array(vf, dtype=np.float64)
    m = o.mesh
    nc, nif, nf = m.n_cells, m.n_internal_faces, m.n_faces
    V = o.get("cell_volume")
    area = o.get("face_area")
    nrm = o.get("face_normal").reshape(-1, 3)
    wC = o.get("face_wC")
    owner = np.asarray(m.owner)
    nbr = np.asarray(m.neighbour)
    fr = np.zeros(nf) if face_r is None else np.asarray(face_r, dtype=np.float64)
    maxA = np.zeros(nc)
    for f in range(nf):
        maxA[owner[f]] = max(maxA[owner[f]], area[f])
        if f < nif:
            maxA[nbr[f]] = max(maxA[nbr[f]], area[f])
    tau = 1e15
    for i in range(nc):
        tau = min(tau, min(V[i] ** 0.3, V[i] / maxA[i]))
    res_hist = []
    for _ in range(iters):
        g = o.lsq(vf, fr)
        mag = np.sqrt(g[:, 0] ** 2 + g[:, 1] ** 2 + g[:, 2] ** 2)
        nv = np.zeros_like(g)
        nz = mag != 0.0
        nv[nz] = g[nz] / mag[nz, None]
        aN = vf * (1.0 - vf)
        resi = np.zeros(nc)
        for f in range(nf):
            o_ = owner[f]
            n = nrm[f]
            unf = wC[f] * (nv[o_, 0] * n[0] + nv[o_, 1] * n[1] + nv[o_, 2] * n[2])
            aL, aR = aN[o_], 0.0
            if f < nif:
                q = nbr[f]
                unf += (1.0 - wC[f]) * (nv[q, 0] * n[0] + nv[q, 1] * n[1] + nv[q, 2] * n[2])
                aR = aN[q]
            else:
                unf = 0.0
            wl = 1.0 if unf > 0.0 else 0.0
            af = wl * aL + (1.0 - wl) * aR
            if aL * (1.0 - aL) < 1e-15:
                af = 0.0
            if aR * (1.0 - aR) < 1e-15:
                af = 0.0
            flux = af * unf * area[f]
            resi[o_] -= flux
            if f < nif:
                resi[nbr[f]] += flux
        r2 = 0.0
        for i in range(nc):
            up = cfl * tau * resi[i] / V[i]
            vf[i] += up
            r2 += up * up
            vf[i] = max(0.0, min(1.0, vf[i]))
        res_hist.append(np.sqrt(r2))
    return vf, res_hist
