"""Python front end of the MI355X assemble-and-solve path.

``GpuSolver`` mirrors the entry points of SEMO's ``SEMO_Solvers_Builder`` /
``SEMO_Utility_Math`` for the pressure-based incompressible path
(src/solvers/build.h, src/math/math.h): the same operation names, argument
meaning and return values, over one device context that keeps the mesh and
every field resident in HBM.  Arrays are numpy arrays in the caller's cell and
face numbering.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import F, COMPONENTS, FACE_FIELDS, SemoError

PREC = {"fp64": 0, "mixed": 1}


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class GpuSolver:
    def __init__(self, device: int = 0, stream: int | None = None):
        self._L = _lib.lib()
        h = C.c_void_p()
        rc = self._L.semo_ctx_create(device, C.c_void_p(stream) if stream else None, C.byref(h))
        if rc != 0:
            raise SemoError(rc, "semo_ctx_create failed")
        self._h = h
        self.mesh = None
        self.last_stats = None

    def close(self):
        if getattr(self, "_h", None):
            self._L.semo_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _ok(self, rc):
        if rc != 0:
            raise SemoError(rc, self._L.semo_ctx_last_error(self._h).decode())

    # ---- mesh / case / solver ---------------------------------------------------
    def upload_mesh(self, mesh, reorder: bool = True):
        """SEMO_Mesh_Geometric::init + setStencil + initLeastSquare, plus the
        RCM reordering of storage (lib_RCM)."""
        self._keep = [np.ascontiguousarray(mesh.points, dtype=np.float64)] + [
            np.ascontiguousarray(a, dtype=np.int32) for a in
            (mesh.owner, mesh.neighbour, mesh.face_ptr, mesh.face_pts, mesh.cell_ptr,
             mesh.cell_pts, mesh.patch_start, mesh.patch_nfaces)]
        k = self._keep
        d = _lib.MeshDesc(mesh.n_points, mesh.n_cells, mesh.n_faces, mesh.n_internal_faces,
                          len(mesh.patch_names), *[a.ctypes.data for a in k])
        self._ok(self._L.semo_mesh_upload(self._h, C.byref(d), 1 if reorder else 0))
        self.mesh = mesh
        del self._keep

    def set_case(self, case):
        m = self.mesh
        cs = _lib.CaseDesc((C.c_double * 2)(*case.rho), (C.c_double * 2)(*case.mu), case.sigma,
                           (C.c_double * 3)(*case.g), case.dt, case.urf_mom, case.urf_p,
                           case.urf_pu, case.limit, getattr(case, "urf_vf", 1.0),
                           getattr(case, "n_vof", 0), 1 if getattr(case, "adjust_dt", False) else 0,
                           getattr(case, "max_co", 1.0), getattr(case, "max_dt", 1.0))
        bcs = (_lib.PatchBC * len(m.patch_names))()
        for i, nm in enumerate(m.patch_names):
            b = case.bcs[nm]
            bcs[i] = _lib.PatchBC(b.p_type, b.p_value, b.u_type, (C.c_double * 3)(*b.u_value),
                                  b.vf_type, b.vf_value)
        self._ok(self._L.semo_set_case(self._h, C.byref(cs), bcs))
        pts = getattr(case, "vf_points", None)
        if pts is not None:
            a = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 4)
            self._ok(self._L.semo_set_vf_function_points(self._h, a.shape[0], _ptr(a)))
        les = getattr(case, "les_model", None)
        self._ok(self._L.semo_set_les(self._h, self.LES_MODELS[les] if les else 0))

    def set_solver(self, precision="fp64", tol_p=1e-8, tol_u=1e-8, maxit_p=500, maxit_u=500,
                   cheb_degree=1, coarse_size=1000):
        p = _lib.SolverParams(PREC[precision], tol_p, tol_u, maxit_p, maxit_u, cheb_degree, coarse_size)
        self._ok(self._L.semo_set_solver(self._h, C.byref(p)))

    # ---- fields ---------------------------------------------------------------------------
    def _size(self, name):
        if name == "face_Ufix":
            return 3 * (self.mesh.n_faces - self.mesh.n_internal_faces)
        n = self.mesh.n_faces if name in FACE_FIELDS else self.mesh.n_cells
        return n * COMPONENTS.get(name, 1)

    MSOLVERS = {"jacobi": 0, "amg": 1, "auto": 2}

    def set_momentum_solver(self, mode: str):
        """Preconditioner of the momentum / volume-fraction BiCGStab:
        "jacobi" (default), "amg" (a V-cycle of the system's own hierarchy)
        or "auto" (Jacobi until a solve needs more than 30 iterations)."""
        self._ok(self._L.semo_set_momentum_solver(self._h, self.MSOLVERS[mode]))

    def set_boundary_velocity(self, values):
        """Per-face values of the fixedValue-U boundaries: values[3][n_b] over
        the boundary faces in the mesh's face order (faces n_internal_faces ..
        n_faces-1); a fixedValue-U face then takes its entry instead of its
        patch's uniform value.  None clears them."""
        if values is None:
            self._ok(self._L.semo_field_set(self._h, F["face_Ufix"], None, 0))
            return
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(3, -1)
        self.set_field("face_Ufix", a)

    def set_field(self, name, values):
        a = np.ascontiguousarray(values, dtype=np.float64).ravel()
        self._ok(self._L.semo_field_set(self._h, F[name], _ptr(a), a.size))

    def get_field(self, name):
        n = self._size(name)
        a = np.empty(n, dtype=np.float64)
        self._ok(self._L.semo_field_get(self._h, F[name], _ptr(a), n))
        if name == "face_Ufix":
            return a.reshape(3, -1)
        c = COMPONENTS.get(name, 1)
        return a.reshape(-1, c) if c > 1 else a

    def permutation(self):
        n = self.mesh.n_cells
        perm = np.empty(n, dtype=np.int32); inv = np.empty(n, dtype=np.int32)
        self._ok(self._L.semo_mesh_get_perm(self._h, _ptr(perm), _ptr(inv)))
        return perm, inv

    def face_map(self):
        m = self.mesh
        fm = np.empty(m.n_faces, dtype=np.int32); fl = np.empty(m.n_faces, dtype=np.int8)
        self._ok(self._L.semo_mesh_get_face_map(self._h, _ptr(fm), _ptr(fl)))
        return fm, fl

    def connectivity(self):
        m = self.mesh
        o = np.empty(m.n_faces, dtype=np.int32); nb = np.empty(m.n_internal_faces, dtype=np.int32)
        self._ok(self._L.semo_mesh_get_connectivity(self._h, _ptr(o), _ptr(nb)))
        return o, nb

    # ---- SEMO_Solvers_Builder entry points -------------------------------------------------
    def calcMomentumEqs(self):
        n = C.c_double(); st = _lib.SolveStats()
        self._ok(self._L.semo_calc_momentum_eqs(self._h, C.byref(n), C.byref(st)))
        self.last_stats = st
        return n.value

    def calcPressureEq(self):
        n = C.c_double(); st = _lib.SolveStats()
        self._ok(self._L.semo_calc_pressure_eq(self._h, C.byref(n), C.byref(st)))
        self.last_stats = st
        return n.value

    def calcVolfracEq(self):
        """One VOF iteration: MSTACS face values, calcVolfracEq, EOS and
        transport; returns sqrt(sum dalpha^2)."""
        n = C.c_double(); st = _lib.SolveStats()
        self._ok(self._L.semo_calc_volfrac_eq(self._h, C.byref(n), C.byref(st)))
        self.last_stats = st
        return n.value

    def calcSourceGravity(self):
        self._ok(self._L.semo_calc_source_gravity(self._h))

    def calcSourceSurfaceTension(self):
        """Curvature and sigma kappa grad(alpha) from the current fields."""
        self._ok(self._L.semo_calc_source_surface_tension(self._h))
        return 0.0

    def calcTimeStep(self, adjust: bool = True):
        """calcTimeStepFromCorantNumber (adjust) / calcCorantNumberForPrint:
        returns (time step in use, Courant number)."""
        dt = C.c_double(); co = C.c_double()
        self._ok(self._L.semo_calc_time_step(self._h, 1 if adjust else 0, C.byref(dt), C.byref(co)))
        return dt.value, co.value

    def calcCoupledEq(self):
        """One block-coupled (p', u', v', w') outer iteration (the design of
        the reference's commented-out calcCoupledEq); returns the velocity
        correction norm, solve statistics in last_stats."""
        n = C.c_double(); st = _lib.SolveStats()
        self._ok(self._L.semo_calc_coupled_eq(self._h, C.byref(n), C.byref(st)))
        self.last_stats = st
        return n.value

    def coupledPressureBased(self, n_iter: int = 1):
        """One time step with n_iter block-coupled outer iterations."""
        norms = np.zeros(n_iter)
        self._ok(self._L.semo_coupled_pressure_based(self._h, n_iter, _ptr(norms)))
        return norms

    def coupled_matrix(self):
        """Block COO of the last assembled coupled system: rows, cols,
        blocks (nnzb, 4, 4), rhs (4, n) variable-major."""
        nnzb = C.c_int64(0)
        self._ok(self._L.semo_coupled_matrix_get(self._h, C.byref(nnzb), None, None, None, None))
        r = np.empty(nnzb.value, dtype=np.int32); c = np.empty(nnzb.value, dtype=np.int32)
        b = np.empty((nnzb.value, 4, 4)); rhs = np.empty((4, self.mesh.n_cells))
        self._ok(self._L.semo_coupled_matrix_get(self._h, C.byref(nnzb), _ptr(r), _ptr(c), _ptr(b), _ptr(rhs)))
        return r, c, b, rhs

    def coupled_spmv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(4, -1)
        y = np.empty_like(x)
        self._ok(self._L.semo_coupled_spmv(self._h, _ptr(x), _ptr(y)))
        return y

    def coupled_solve(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64).reshape(4, -1)
        x = np.empty_like(b); st = _lib.SolveStats()
        self._ok(self._L.semo_coupled_solve(self._h, _ptr(b), _ptr(x), C.byref(st)))
        self.last_stats = st
        return x

    def incompressiblePressureBased(self, n_correctors: int = 1):
        norms = np.zeros(2 * n_correctors)
        self._ok(self._L.semo_incompressible_pressure_based(self._h, n_correctors, _ptr(norms)))
        return norms

    # ---- discrete adjoint (no counterpart in the reference) -------------------------------
    def _patch_weights(self, weights):
        w = np.zeros(len(self.mesh.patch_names))
        for name, v in weights.items():
            w[self.mesh.patch(name)] = v
        return w

    def objective(self, weights: dict) -> float:
        """J = sum over patches of weight * sum_f A_f p_f (face pressure of
        the boundary reconstruction) for the current pressure field."""
        w = self._patch_weights(weights)
        J = C.c_double()
        self._ok(self._L.semo_objective_pressure(self._h, _ptr(w), C.byref(J)))
        return J.value

    def volume_integral(self, name: str):
        """(sum_i phi_i V_i, sum_i V_i) of a one-component cell field, in
        FP64 on the device (VF: the first phase's volume)."""
        a, v = C.c_double(), C.c_double()
        self._ok(self._L.semo_volume_integral(self._h, F[name], C.byref(a), C.byref(v)))
        return a.value, v.value

    def adjointAlphaP(self, n_iter: int, weights: dict):
        """Runs n_iter outer iterations (as incompressiblePressureBased) and
        returns (J of the final state, dJ/dalpha_p per cell) by reverse-mode
        differentiation of the discrete iterations."""
        w = self._patch_weights(weights)
        J = C.c_double()
        g = np.empty(self.mesh.n_cells)
        self._ok(self._L.semo_adjoint_alpha_p(self._h, n_iter, _ptr(w), C.byref(J), _ptr(g)))
        return J.value, g

    def solve_stats(self, eq: str):
        """Iterations / final relative residuals / AMG levels of the last
        momentum, pressure or volume-fraction solve."""
        st = _lib.SolveStats()
        self._ok(self._L.semo_get_solve_stats(self._h, {"momentum": 0, "pressure": 1, "volume_fraction": 2}[eq], C.byref(st)))
        return dict(iters=list(st.iters), rel_res=list(st.rel_res), n_levels=st.n_levels,
                    setup_ms=st.setup_ms)

    def assemble(self, eq: str):
        self._ok(self._L.semo_assemble(self._h, {"momentum": 0, "pressure": 1, "volume_fraction": 2,
                                                 "coupled": 3}[eq]))

    def matrix(self, eq: str):
        """COO (rows, cols, vals) and right-hand sides of the last assembly."""
        nnz = C.c_int64(0)
        self._ok(self._L.semo_matrix_get(self._h, C.byref(nnz), None, None, None, None))
        r = np.empty(nnz.value, dtype=np.int32); c = np.empty(nnz.value, dtype=np.int32)
        v = np.empty(nnz.value, dtype=np.float64)
        nr = 3 if eq == "momentum" else 1
        b = np.empty(nr * self.mesh.n_cells, dtype=np.float64)
        self._ok(self._L.semo_matrix_get(self._h, C.byref(nnz), _ptr(r), _ptr(c), _ptr(v), _ptr(b)))
        return r, c, v, b.reshape(nr, -1)

    # ---- SEMO_Utility_Math ------------------------------------------------------------------
    def calcLeastSquare(self, phi, face_r):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        fr = np.ascontiguousarray(face_r, dtype=np.float64)
        g = np.empty(3 * self.mesh.n_cells)
        self._ok(self._L.semo_least_square(self._h, _ptr(phi), _ptr(fr), _ptr(g)))
        return g.reshape(-1, 3)

    def least_square_noalloc(self, d_phi: int, d_face_r: int, d_grad: int):
        """semo_least_square_noalloc on device addresses (device numbering,
        boundary face values in device face order); issued on the context's
        stream, not synchronised."""
        self._ok(self._L.semo_least_square_noalloc(self._h, C.c_void_p(d_phi), C.c_void_p(d_face_r),
                                                   C.c_void_p(d_grad)))

    def calcGaussGreen(self, field, phi, face_r, gradient):
        """calcGaussGreen(mesh, cn, fn, gradient): field names cn ("P", "U",
        "V", "W" or "VF", whose patch types give the boundary values), phi the
        cell values, face_r the right face values, gradient the input
        gradient (skewness correction, extrapolation); returns the new one."""
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        fr = np.ascontiguousarray(face_r, dtype=np.float64)
        g = np.ascontiguousarray(gradient, dtype=np.float64).reshape(-1).copy()
        self._ok(self._L.semo_gauss_green(self._h, F[field], _ptr(phi), _ptr(fr), _ptr(g)))
        return g.reshape(-1, 3)

    LES_MODELS = {"WALE": 1, "smagorinsky": 2, "dynamicSmagorinsky": 3}

    def calcLES(self, model, U, V, W, rho, face_rU, face_rV, face_rW):
        """SEMO_TurbModel_LES (turbulenceModels/LES.cpp) by its
        turbulenceProperties name (transport.cpp:42-50): "WALE",
        "smagorinsky" or "dynamicSmagorinsky".  U, V, W, rho cell values,
        face_r* the right face values [n_faces]; returns (muT, kSGS)."""
        if model not in self.LES_MODELS:
            raise ValueError(f"LES model {model!r}: not one of {sorted(self.LES_MODELS)}")
        arr = [np.ascontiguousarray(a, dtype=np.float64) for a in (U, V, W, rho, face_rU, face_rV, face_rW)]
        mu = np.empty(self.mesh.n_cells); k = np.empty(self.mesh.n_cells)
        self._ok(self._L.semo_les(self._h, self.LES_MODELS[model], *[_ptr(a) for a in arr], _ptr(mu), _ptr(k)))
        return mu, k

    FLUX_SCHEMES = {"HAUS": 0, "SLAU2": 1, "Roe": 2, "APRoe": 3, "RoeAMTP": 4, "Upwind": 5, "AUSMPWP_N": 6,
                    "RoeM_N": 7}

    def calcFlux(self, scheme, left, right, normal, par=None):
        """The convective flux of SEMO's scheme calcFlux<scheme> (solvers/
        flux.cpp) per face: left / right [n, 8 + nsp] = (P, U, V, W, T, rho,
        C, Ht, Y...), normal [n, 3], par [n, 7] = (w1, w2, w3, fC, Uco, dt,
        Lch) for AUSMPWP_N and RoeM_N; returns [n, 4 + nsp]."""
        L = np.ascontiguousarray(left, dtype=np.float64)
        R = np.ascontiguousarray(right, dtype=np.float64)
        nv = np.ascontiguousarray(normal, dtype=np.float64)
        n, w = L.shape
        nsp = w - 8
        pa = None if par is None else np.ascontiguousarray(par, dtype=np.float64)
        out = np.empty((n, 4 + nsp))
        self._ok(self._L.semo_flux(self._h, self.FLUX_SCHEMES[scheme], n, nsp, _ptr(L), _ptr(R), _ptr(nv),
                                   None if pa is None else _ptr(pa), _ptr(out)))
        return out

    def calcFluxes(self, face_left, face_right, face_grad=None):
        """calcFluxes of SEMO's density-based solvers (solvers/RHS.cpp:98-406):
        SLAU2 + viscous flux per face, summed into the cells.  face_left /
        face_right [n_faces, 8 + nsp] = (P, U, V, W, T, rho, C, Ht, mu,
        Y_0..Y_{nsp-2}), face_grad [n_faces, 9] or None (zero, as SEMO's runs
        have them); returns [n, 4 + nsp]."""
        L = np.ascontiguousarray(face_left, dtype=np.float64)
        R = np.ascontiguousarray(face_right, dtype=np.float64)
        G = None if face_grad is None else np.ascontiguousarray(face_grad, dtype=np.float64)
        nsp = L.shape[1] - 8
        out = np.empty((self.mesh.n_cells, 4 + nsp))
        self._ok(self._L.semo_calc_fluxes(self._h, nsp, _ptr(L), _ptr(R), None if G is None else _ptr(G),
                                          _ptr(out)))
        return out

    def calcRHS(self, time_step, g, face_left, face_right, cells, face_grad=None, st_force=None):
        """calcRHS (RHS.cpp:5-28): calcFluxes, the dual-time source, gravity,
        the caller's surface-tension force [n, 3] (or None), the scaling by
        -dtPseudo / V.  cells [n, 8 + (nsp-1) + 2 (4 + nsp)] = (rho, U, V, W,
        Ht, P, dtPseudo, VF, Y.., Qn.., Qm..)."""
        L = np.ascontiguousarray(face_left, dtype=np.float64)
        R = np.ascontiguousarray(face_right, dtype=np.float64)
        G = None if face_grad is None else np.ascontiguousarray(face_grad, dtype=np.float64)
        cs = np.ascontiguousarray(cells, dtype=np.float64)
        S = None if st_force is None else np.ascontiguousarray(st_force, dtype=np.float64)
        gv = np.ascontiguousarray(g, dtype=np.float64)
        nsp = L.shape[1] - 8
        out = np.empty((self.mesh.n_cells, 4 + nsp))
        self._ok(self._L.semo_calc_rhs(self._h, nsp, float(time_step), _ptr(gv), _ptr(L), _ptr(R),
                                       None if G is None else _ptr(G), _ptr(cs), None if S is None else _ptr(S),
                                       _ptr(out)))
        return out

    def updateValues(self, urf, limits, residuals, cells, face_un):
        """updateValues (solvers/densityBased.cpp:366-512): returns the new
        cells [n, 5 + nsp] = (P, U, V, W, T, MF..) and face Un [n_faces].
        urf = (dualTimeURF_P, _U, _T, _MF); limits = (minP, maxP, .., maxT)."""
        r = np.ascontiguousarray(residuals, dtype=np.float64)
        cv = np.ascontiguousarray(cells, dtype=np.float64).copy()
        un = np.ascontiguousarray(face_un, dtype=np.float64).copy()
        u = np.ascontiguousarray(urf, dtype=np.float64)
        lim = np.ascontiguousarray(limits, dtype=np.float64)
        self._ok(self._L.semo_update_values(self._h, r.shape[1] - 4, _ptr(u), _ptr(lim), _ptr(r), _ptr(cv),
                                            _ptr(un)))
        return cv, un

    def compressVolumeFraction(self, vf, cfl, iters, face_r=None):
        """SEMO's interface-compression utility (src/utility/
        compressVolumeFraction.cpp:121-228): returns the compressed fractions
        and each iteration's residual."""
        v = np.ascontiguousarray(vf, dtype=np.float64).copy()
        fr = None if face_r is None else np.ascontiguousarray(face_r, dtype=np.float64)
        res = np.zeros(max(int(iters), 1))
        self._ok(self._L.semo_compress_volume_fraction(self._h, float(cfl), int(iters), _ptr(v),
                                                       None if fr is None else _ptr(fr), _ptr(res)))
        return v, res[:int(iters)]

    LINSOL_METHODS = {"point": 0, "single": 1, "lusgs": 2}

    def _linsol(self, method, time_step, state, residuals):
        st = np.ascontiguousarray(state, dtype=np.float64)
        r = np.ascontiguousarray(residuals, dtype=np.float64).copy()
        n, w = r.shape
        nsp = w - 4
        if st.shape != (n, 12 + 3 * nsp):
            raise ValueError(f"state must be [n, {12 + 3 * nsp}] for {nsp} species")
        self._ok(self._L.semo_linear_solver(self._h, self.LINSOL_METHODS[method], nsp, float(time_step),
                                            _ptr(st), _ptr(r)))
        return r

    def calcLinearSolver(self, time_step, state, residuals):
        """SEMO's point-implicit update (solvers/linearSolver.cpp:666-725):
        per cell (P + 1.5 dtPseudo / time_step T)^-1 residuals.  state [n,
        12 + 3 nsp] = (P, U, V, W, T, rho, C, Ht, dRhoDP, dHtDP, dRhoDT,
        dHtDT, Ur, dtPseudo, Y.., dRhoDY.., dHtDY..), residuals [n, 4 + nsp];
        returns the update."""
        return self._linsol("point", time_step, state, residuals)

    def calcSingleLinearSolver(self, state, residuals):
        """calcSingleLinearSolver (linearSolver.cpp:729-787): T^-1 residuals."""
        return self._linsol("single", 1.0, state, residuals)

    def calcLinearSolverLUSGS(self, time_step, state, residuals):
        """calcLinearSolverLUSGS (linearSolver.cpp:144-389): forward and
        backward sweeps in lib_RCM's order of the mesh as uploaded."""
        return self._linsol("lusgs", time_step, state, residuals)

    def setCellVarMinMax(self, phi):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        mx = np.empty(self.mesh.n_cells); mn = np.empty(self.mesh.n_cells)
        self._ok(self._L.semo_cell_min_max(self._h, _ptr(phi), _ptr(mx), _ptr(mn)))
        return mx, mn

    def solveAMGCL(self, equation, rows, cols, vals, b):
        """Seam 2: solve a COO system; b is (nrhs, n) or (n,)."""
        b = np.atleast_2d(np.ascontiguousarray(b, dtype=np.float64))
        nrhs, n = b.shape
        r = np.ascontiguousarray(rows, dtype=np.int32); c = np.ascontiguousarray(cols, dtype=np.int32)
        v = np.ascontiguousarray(vals, dtype=np.float64)
        x = np.empty_like(b)
        st = _lib.SolveStats()
        self._ok(self._L.semo_solve_coo(self._h, equation.encode(), n, len(r), _ptr(r), _ptr(c), _ptr(v),
                                        nrhs, _ptr(b), _ptr(x), C.byref(st)))
        self.last_stats = st
        return x if nrhs > 1 else x[0]

    # ---- profiling ---------------------------------------------------------------------------
    def stream_copy(self, n: int, reps: int = 10) -> float:
        """ms of one copy of n doubles by the library's copy kernel (average
        of reps)."""
        ms = C.c_double()
        self._ok(self._L.semo_stream_copy(self._h, int(n), int(reps), C.byref(ms)))
        return ms.value

    def profile(self, on=True):
        self._ok(self._L.semo_profile_enable(self._h, 1 if on else 0))

    def profile_reset(self):
        self._ok(self._L.semo_profile_reset(self._h))

    def profile_stats(self):
        out = {}
        i = 0
        name = C.create_string_buffer(64)
        cnt = C.c_int64(); ms = C.c_double(); by = C.c_double()
        while self._L.semo_profile_entry(self._h, i, name, C.byref(cnt), C.byref(ms), C.byref(by)) == 0:
            out[name.value.decode()] = dict(count=cnt.value, ms=ms.value, bytes=by.value)
            i += 1
        return out


def rcm(adj_ptr, adj):
    """Standalone GENRCM (lib_RCM): perm[new] = old, inv[old] = new."""
    L = _lib.lib()
    n = len(adj_ptr) - 1
    p = np.ascontiguousarray(adj_ptr, dtype=np.int32); a = np.ascontiguousarray(adj, dtype=np.int32)
    perm = np.empty(n, dtype=np.int32); inv = np.empty(n, dtype=np.int32)
    rc = L.semo_rcm(n, _ptr(p), _ptr(a), _ptr(perm), _ptr(inv))
    if rc != 0:
        raise SemoError(rc, "semo_rcm")
    return perm, inv
