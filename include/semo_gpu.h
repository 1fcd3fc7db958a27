/*
 * semo_gpu.h -- C ABI of the MI355X (gfx950) assemble-and-solve path of the
 * SEMO unstructured finite-volume solver.
 *
 * The library owns one device context per handle.  All arrays crossing the
 * boundary are host arrays in the caller's (original, un-permuted) cell and
 * face numbering; the library renumbers cells by reverse Cuthill-McKee on
 * upload and keeps every field resident in HBM.  Functions return a
 * semo_status; semo_ctx_last_error() gives the message of the last failure.
 *
 * Each entry point names the reference interface it replaces.  Reference
 * paths are relative to the SEMO_Mesh source tree.
 *
 * Threading: no internal threads; calls on one context are not re-entrant,
 * calls on distinct contexts are independent.  All device work of a context
 * is issued on the stream given to semo_ctx_create (0 = the library's own).
 */
#ifndef SEMO_GPU_H
#define SEMO_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    SEMO_OK = 0,
    SEMO_EINVAL = 1,          /* bad argument / inconsistent sizes              */
    SEMO_EINVAL_BC = 2,       /* unknown boundary type (eqMomentum.cpp:861-864) */
    SEMO_ESINGULAR_LSQ = 3,   /* detA == 0 in the LSQ setup (gradient.cpp:411)  */
    SEMO_ENEG_VOLUME = 4,     /* non-positive cell volume (geometric.cpp:1152)  */
    SEMO_EZERO_AREA = 5,      /* zero face area (geometric.cpp:197-202)         */
    SEMO_EZERO_PIVOT = 6,     /* singular coarse matrix / zero diagonal         */
    SEMO_ENOCONV = 7,         /* Krylov did not reach the tolerance             */
    SEMO_EHIP = 8,            /* HIP runtime error                              */
    SEMO_ESTATE = 9           /* call order (e.g. no mesh uploaded)             */
} semo_status;

typedef struct semo_ctx semo_ctx;

/* ---- context --------------------------------------------------------------- */
int semo_ctx_create(int device, void* hip_stream, semo_ctx** out);
int semo_ctx_destroy(semo_ctx* ctx);
const char* semo_ctx_last_error(const semo_ctx* ctx);
/* bytes of device memory the context holds */
int64_t semo_ctx_device_bytes(const semo_ctx* ctx);

/* ---- mesh ---------------------------------------------------------------------
 * Face-based polyhedral mesh as SEMO_Mesh_Load::vtu builds it
 * (src/load/load.cpp:352-438, src/mesh/build.h:42-175): internal faces first,
 * then boundary faces grouped in patches; face points ordered so the
 * right-hand normal points owner -> neighbour.                                */
typedef struct {
    int32_t n_points, n_cells, n_faces, n_internal_faces, n_patches;
    const double* points;         /* [n_points*3]                        */
    const int32_t* owner;         /* [n_faces]                           */
    const int32_t* neighbour;     /* [n_internal_faces]                  */
    const int32_t* face_ptr;      /* [n_faces+1]  face -> points CSR     */
    const int32_t* face_pts;
    const int32_t* cell_ptr;      /* [n_cells+1]  cell -> points CSR     */
    const int32_t* cell_pts;      /*   (SEMO cell.points order)          */
    const int32_t* patch_start;   /* [n_patches]  first face            */
    const int32_t* patch_nfaces;  /* [n_patches]                         */
} semo_mesh_desc;

/* Replaces SEMO_Mesh_Geometric::init + setStencil (src/mesh/geometric.cpp:
 * 792-1768, 1781-1856), SEMO_Utility_Math::initLeastSquare (src/math/
 * gradient.cpp:11-589) and, when reorder != 0, lib_RCM (src/math/RCM.cpp:
 * 34-210) applied to storage.                                                */
int semo_mesh_upload(semo_ctx* ctx, const semo_mesh_desc* mesh, int reorder);
/* cell permutation: perm[new] = old, inv[old] = new (lib_RCM's cell.RCM /
 * cell.invRCM, RCM.cpp:199-207); identity without reordering               */
int semo_mesh_get_perm(const semo_ctx* ctx, int32_t* perm, int32_t* inv);
/* device face numbering: face_new_to_old[new] = old; flipped[new] = 1 when the
 * device face's owner is the original neighbour                              */
int semo_mesh_get_face_map(const semo_ctx* ctx, int32_t* face_new_to_old, int8_t* flipped);
/* device owner/neighbour arrays (device numbering)                           */
int semo_mesh_get_connectivity(const semo_ctx* ctx, int32_t* owner, int32_t* neighbour);

/* Standalone GENRCM on a CSR graph (0-based): replaces lib_RCM's genrcm call
 * (src/math/RCM.cpp:175-183).  perm[new] = old, inv[old] = new.              */
int semo_rcm(int32_t n, const int32_t* adj_ptr, const int32_t* adj, int32_t* perm, int32_t* inv);

/* ---- case ------------------------------------------------------------------------
 * Boundary types (SEMO_Boundary::type strings, src/solvers/reconIncom.cpp:
 * 591-1032): */
enum { SEMO_BC_ZERO_GRADIENT = 0, SEMO_BC_FIXED_VALUE = 1, SEMO_BC_NO_SLIP = 2,
       SEMO_BC_SLIP = 3, SEMO_BC_SURFACE_NORMAL_FIXED_VALUE = 4, SEMO_BC_INLET_OUTLET = 5,
       /* p only: the owner cell's pressure where its Mach number |U|/C
          exceeds 1, else p_value; the pressure equation's boundary diagonal
          is added only above Mach 1 (reconIncom.cpp:618-631,
          eqPressure.cpp:625-632).  C is field SEMO_F_C.                      */
       SEMO_BC_SWITCH = 6,
       /* volume fraction only: per-face values of the nearest given point
          (semo_set_vf_function_points; reconIncom.cpp:929-1019)             */
       SEMO_BC_FUNCTION = 7 };

typedef struct {
    int32_t p_type;  double p_value;
    int32_t u_type;  double u_value[3];
    int32_t vf_type; double vf_value;
} semo_patch_bc;

typedef struct {
    double rho[2], mu[2];   /* constant-density phases (species[0..1])       */
    double sigma;           /* surface tension of phase 0                    */
    double g[3];            /* gravity                                       */
    double dt;              /* fixed time step                               */
    double urf_mom;         /* momVelURF  (eqMomentum.cpp:1040)              */
    double urf_p;           /* prePreURF  (eqPressure.cpp:827)               */
    double urf_pu;          /* preVelURF  (eqPressure.cpp:832)               */
    double limit;           /* |P|,|U|,|V|,|W| clip (controls.maxP ... minW) */
    double urf_vf;          /* vofVofURF  (eqVolfrac.cpp:776)                */
    int32_t n_vof;          /* nSpeciesEq: VOF iterations per time step
                               (pressureBased.cpp:104-127); 0 = one phase   */
    int32_t adjust_dt;      /* adjustTimeStep: dt from the Courant number    */
    double max_co, max_dt;  /* maxCo, maxTimeStep (timestep.cpp:260-330)     */
} semo_case;

/* Replaces the controls/species read by SEMO_Controls_Builder
 * (src/controls/build.cpp:7-623) for this path; bcs has n_patches entries.   */
int semo_set_case(semo_ctx* ctx, const semo_case* c, const semo_patch_bc* bcs);

/* Points of the volume-fraction "function" boundary: replaces the read of
 * ./VF.csv in reconIncomZeroOrder (src/solvers/reconIncom.cpp:929-1019).
 * pts[4*k .. 4*k+3] = (VF, x, y, z) of row k (the CSV's columns 3-6).  Every
 * boundary face takes the VF of the nearest point (squared distance; the first
 * of equal distances; points at squared distance >= 1e9 are never taken).
 * Call after semo_mesh_upload; n = 0 clears the points.                      */
int semo_set_vf_function_points(semo_ctx* ctx, int32_t n, const double* pts);

/* ---- linear solver configuration --------------------------------------------------
 * Replaces the hard-coded AMGCL set-ups of solveAMGCL (src/solvers/
 * solveAMGCL.cpp:938-980 pressure, 1198-1258 momentum).                       */
enum { SEMO_PREC_FP64 = 0, SEMO_PREC_MIXED = 1 };
typedef struct {
    int32_t precision;      /* SEMO_PREC_FP64 or SEMO_PREC_MIXED (FP32 AMG)  */
    double tol_p;           /* relative residual, pressure                   */
    double tol_u;           /* relative residual, momentum                   */
    int32_t maxit_p, maxit_u;
    int32_t cheb_degree;    /* Chebyshev smoother degree (pre = post)        */
    int32_t coarse_size;    /* coarsest AMG level size (dense direct solve)  */
} semo_solver_params;
int semo_set_solver(semo_ctx* ctx, const semo_solver_params* p);
/* Preconditioner of the momentum and volume-fraction BiCGStab solves.  SEMO
 * solves them with aggregation-AMG-preconditioned Krylov methods (AMGCL,
 * src/solvers/solveAMGCL.cpp:881-927, 1195-1258).  JACOBI (the default):
 * diagonal, enough at the time steps of its cases (2 iterations to 1e-8 on
 * the 128^3 cavity); AMG: one V-cycle of a smoothed-aggregation hierarchy of
 * the system itself per component (FP32 with SEMO_PREC_MIXED), built on the
 * host and refreshed on the device like the pressure's; AUTO: Jacobi until a
 * solve of that system needs more than 30 iterations, AMG from then on.    */
enum { SEMO_MSOLVER_JACOBI = 0, SEMO_MSOLVER_AMG = 1, SEMO_MSOLVER_AUTO = 2 };
int semo_set_momentum_solver(semo_ctx* ctx, int mode);

typedef struct {
    int32_t iters[4];       /* per right-hand side                           */
    double rel_res[4];      /* final ||b-Ax||/||b|| in FP64                  */
    int32_t n_levels;       /* AMG levels (pressure)                         */
    double setup_ms;        /* hierarchy set-up time of the last build       */
} semo_solve_stats;

/* ---- fields ----------------------------------------------------------------------- */
typedef enum {
    SEMO_F_P = 0, SEMO_F_U, SEMO_F_V, SEMO_F_W, SEMO_F_VF, SEMO_F_RHO, SEMO_F_MU,
    SEMO_F_OLDU, SEMO_F_OLDV, SEMO_F_OLDW, SEMO_F_QM0, SEMO_F_QM1, SEMO_F_QM2,
    SEMO_F_SRCG0, SEMO_F_SRCG1, SEMO_F_SRCG2,
    SEMO_F_MAXP, SEMO_F_MINP, SEMO_F_MAXU, SEMO_F_MINU, SEMO_F_MAXV, SEMO_F_MINV,
    SEMO_F_MAXW, SEMO_F_MINW,
    SEMO_F_ALPHA_P,                                         /* Darcy porosity   */
    SEMO_F_QM3, SEMO_F_OLDVF,                               /* alpha time levels */
    SEMO_F_KAPPA, SEMO_F_SRCS0, SEMO_F_SRCS1, SEMO_F_SRCS2,   /* curvature, sigma kappa grad(alpha) */
    SEMO_F_LS,                                              /* reinitialised level set (read) */
    SEMO_F_GRADP = 32, SEMO_F_GRADU, SEMO_F_GRADV, SEMO_F_GRADW, SEMO_F_GRADVF, /* [n*3] */
    /* cell speed of sound, computed at the first time step from the mixture
       EOS as calcCellEOSVF sets it at load (eos.cpp:160-293) and never
       updated by the incompressible iteration; read by SEMO_BC_SWITCH.
       Settable after that step.                                             */
    SEMO_F_C = 40,
    /* subgrid viscosity and kinetic energy of the LES model (semo_set_les),
       as the last calcCellTransport left them; read-only                    */
    SEMO_F_MUT = 41, SEMO_F_KSGS = 42,
    SEMO_F_X0 = 48, SEMO_F_X1, SEMO_F_X2,                  /* last solution     */
    SEMO_F_B0 = 56, SEMO_F_B1, SEMO_F_B2,                  /* last right sides  */
    /* geometry (read-only): cell arrays, then face arrays (original face order) */
    SEMO_G_CELL_VOLUME = 64, SEMO_G_CELL_XYZ, SEMO_G_CELL_LSQ,            /* n, 3n, 6n */
    SEMO_G_FACE_AREA = 80, SEMO_G_FACE_NORMAL, SEMO_G_FACE_XYZ, SEMO_G_FACE_WC,
    SEMO_G_FACE_MAGPN, SEMO_G_FACE_ALPHAF, SEMO_G_FACE_EPN, SEMO_G_FACE_SKEW,
    /* face volume fractions varL/varR[fVF] after the last VOF assembly (MSTACS
       on internal faces, boundary values on boundary faces), read-only      */
    SEMO_F_FACE_VFL = 96, SEMO_F_FACE_VFR,
    /* per-face values of fixedValue-U boundaries, [3][n_faces - n_internal_faces]
       (component-major, boundary faces in the caller's face order), settable:
       a fixedValue-U face takes these values instead of its patch's u_value
       (the form SEMO's UDF inlet profiles take, reconIncom.cpp:707-727);
       semo_field_set with n = 0 clears them.  The orifice profile of
       reconIncom.cpp:707-727 still applies on top.                          */
    SEMO_F_FACE_UFIX = 98
} semo_field;
/* host <-> device copies in the caller's numbering; n = element count       */
int semo_field_set(semo_ctx* ctx, int field, const double* host, int64_t n);
int semo_field_get(semo_ctx* ctx, int field, double* host, int64_t n);

/* ---- per-equation entry points (seam 1: src/solvers/build.h) ------------------ */
/* calcMomentumEqs (src/solvers/eqMomentum.cpp:9-1122): assemble, solve the
 * three right-hand sides, under-relax and clip U,V,W; *norm = |dU|/|U|.      */
int semo_calc_momentum_eqs(semo_ctx* ctx, double* norm, semo_solve_stats* st);
/* calcPressureEq (src/solvers/eqPressure.cpp:8-942): assemble, solve p',
 * two Green-Gauss passes, correct P and U,V,W; *norm as the reference.       */
int semo_calc_pressure_eq(semo_ctx* ctx, double* norm, semo_solve_stats* st);
/* One VOF iteration of incompressiblePressureBased (pressureBased.cpp:104-127):
 * setIncomValuesLeftRightFaceWithVfMSTACS (src/solvers/reconIncom.cpp:345-430,
 * calcMSTACS src/solvers/NVD.cpp:1098-1862), calcVolfracEq (src/solvers/
 * eqVolfrac.cpp:9-787), calcIncomCellEOSVF + calcCellTransport; *norm =
 * sqrt(sum dalpha^2) as calcVolfracEq returns.                                 */
int semo_calc_volfrac_eq(semo_ctx* ctx, double* norm, semo_solve_stats* st);
/* calcSourceGravity (src/solvers/source/gravity.cpp:7-75) from the current
 * fields, after the boundary face values are reconstructed.                   */
int semo_calc_source_gravity(semo_ctx* ctx);
/* calcSourceSurfaceTension (src/solvers/source/surfaceTension.cpp:7-229) with
 * calcCurvature (src/solvers/source/curvature.cpp:7-3537): kappa and the
 * source sigma kappa grad(alpha) from the current fields (zero for sigma = 0). */
int semo_calc_source_surface_tension(semo_ctx* ctx);
/* calcTimeStepFromCorantNumber (adjust != 0, src/solvers/timestep.cpp:260-330)
 * or calcCorantNumberForPrint (adjust == 0, timestep.cpp:202-253): *dt gets
 * the time step in use afterwards (adjusted: maxCo times the smallest
 * convective / viscous / gravity / capillary time scale, capped at maxTimeStep)
 * and *courant the Courant number.                                              */
int semo_calc_time_step(semo_ctx* ctx, int adjust, double* dt, double* courant);
/* incompressiblePressureBased (src/solvers/pressureBased.cpp:6-483) for
 * nPressureEq = 1: one time step of n_vof VOF iterations (semo_case.n_vof),
 * then n_correctors momentum + pressure iterations; norms[2*k], norms[2*k+1]
 * per corrector.                                                               */
int semo_incompressible_pressure_based(semo_ctx* ctx, int n_correctors, double* norms);

/* Block-coupled pressure-velocity iteration: replaces calcCoupledEq
 * (src/solvers/eqCoupled.cpp:12-1250, commented out in the reference; its
 * design: 4 unknowns per cell p', u', v', w', B_n = 4 at :26, variable-major
 * global index at :249-305, one solveAMGCL call at :1246).  One outer
 * iteration: boundary values, momentum and pressure assembly, the 4x4-block
 * system (momentum matrix + linearised pressure force; pressure-correction
 * matrix + velocity-correction flux), FP64 BiCGStab with an FP32 block-AMG
 * preconditioner to semo_solver_params.tol_u / maxit_u, then
 * p += urf_p p', U += urf_mom u' (clipped at limit).  *norm as
 * semo_calc_momentum_eqs.  semo_assemble(ctx, 3) assembles the block system
 * without solving.                                                            */
int semo_calc_coupled_eq(semo_ctx* ctx, double* norm, semo_solve_stats* st);

/* One time step with the block-coupled iteration in place of the
 * momentum / pressure correctors: the set-up of
 * semo_incompressible_pressure_based (time levels, dt, VOF, sources), then
 * n_iter semo_calc_coupled_eq iterations; norms[k] per iteration.  Afterwards
 * semo_get_solve_stats(ctx, 0, .) holds the last block solve, with iters[1]
 * and iters[2] the sum and the maximum of the Krylov iterations over the
 * step's outer iterations.                                                    */
int semo_coupled_pressure_based(semo_ctx* ctx, int n_iter, double* norms);

/* The last assembled block system: block COO in caller numbering (the n
 * diagonal blocks first, then the off-diagonal blocks row by row), blocks
 * row-major 4x4 over (p', u', v', w'), and the right-hand side variable-major
 * rhs[var * n_cells + cell] (eqCoupled.cpp:249-305).  Call with null arrays
 * to get *nnzb first.                                                          */
int semo_coupled_matrix_get(semo_ctx* ctx, int64_t* nnzb, int32_t* rows, int32_t* cols, double* blocks,
                            double* rhs);

/* y = A x and the solve A x = b with the last assembled block system; x, y,
 * b variable-major [4][n_cells], caller numbering.                           */
int semo_coupled_spmv(semo_ctx* ctx, const double* x, double* y);
int semo_coupled_solve(semo_ctx* ctx, const double* b, double* x, semo_solve_stats* st);

/* Discrete adjoint.  No counterpart in the reference: SEMO has no adjoint.
 * Runs n_iter outer iterations exactly as
 * semo_incompressible_pressure_based(ctx, n_iter, NULL) (single phase:
 * sigma = 0, g = 0, n_vof = 0, fixed dt, equal phase properties), with the
 * Darcy term of field SEMO_F_ALPHA_P (momentum diagonal += alpha_p V, right
 * side -= alpha_p V U; zero if never set).  Returns J of the final state,
 *   J = sum_patches patch_weight[p] * sum_{faces f of p} A_f p_f
 * (p_f: the boundary face pressure; fixedValue patches contribute their
 * value), and dJ/dalpha_p for every cell (caller numbering, n_cells
 * doubles), by reverse-mode differentiation of the discrete iterations
 * (checkpoint per iteration, recompute, transposed solves).  The context is
 * left in the forward's final state.                                         */
int semo_adjoint_alpha_p(semo_ctx* ctx, int n_iter, const double* patch_weight, double* J, double* dJ_dalpha);

/* J (as above) of the current pressure field.  No counterpart in the
 * reference; the objective of semo_adjoint_alpha_p.                          */
int semo_objective_pressure(semo_ctx* ctx, const double* patch_weight, double* J);

/* sum_i field_i V_i over the cells in FP64, reduced on the device (the
 * same value on every run); *volume (optional) gets sum_i V_i.  No
 * counterpart in the reference: the alpha mass balance of two-phase runs
 * (field SEMO_F_VF: the first phase's volume).                               */
int semo_volume_integral(semo_ctx* ctx, int field, double* integral, double* volume);

/* Statistics of the last momentum (eq 0) / pressure (eq 1) / VOF (eq 2) solve. */
int semo_get_solve_stats(const semo_ctx* ctx, int eq, semo_solve_stats* st);

/* Assemble without solving (eq 0 = momentum, 1 = pressure, 2 = volume
 * fraction incl. its MSTACS face values), for inspection.                      */
int semo_assemble(semo_ctx* ctx, int eq);
/* The last assembled system in the caller's numbering as COO.  Call with
 * rows == NULL to get *nnz.  b gets n*nrhs values (nrhs = 3 momentum, else 1). */
int semo_matrix_get(semo_ctx* ctx, int64_t* nnz, int32_t* rows, int32_t* cols, double* vals,
                    double* b);

/* ---- math (seam 3: src/math/math.h) ----------------------------------------------- */
/* calcLeastSquare(mesh,"cellVertex","1st",...) (src/math/gradient.cpp:593-1159):
 * phi[n] cell values, face_r[n_faces] right-face values (used at boundary
 * faces), grad[3n].                                                            */
int semo_least_square(semo_ctx* ctx, const double* phi, const double* face_r, double* grad);
/* semo_least_square without allocating: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops).  Device pointers in the library's device numbering
 * (semo_mesh_get_perm: cell k of the device is the caller's cell perm[k];
 * semo_mesh_get_face_map for faces): d_phi[n_cells], d_face_r the right
 * values of the boundary faces [n_faces - n_internal_faces] in device face
 * order, d_grad[3 n_cells] out.  Issued on the context's stream, not
 * synchronised.  Its scratch is one static device array per GPU for up to
 * 2097152 cells (64 MiB, held until the process ends); larger meshes get
 * SEMO_EINVAL.  Calls on one GPU share that scratch: concurrent calls on one
 * GPU (from different contexts or streams) race.                            */
int semo_least_square_noalloc(semo_ctx* ctx, const double* d_phi, const double* d_face_r, double* d_grad);
/* calcGaussGreen(mesh, cn, fn, gradient) (src/math/gradient.cpp:1778-1985):
 * field = SEMO_F_P, _U, _V, _W or _VF names cn, whose patch types and values
 * (semo_set_case) give the boundary faces' values (fixedValue: the patch
 * value; zeroGradient: phi_P + grad_P.(x_f - x_P); other types: the mean of
 * phi_P and face_r).  phi[n] cell values, face_r[n_faces] the right face
 * values varR[fn] (read at boundary faces), grad[3n] in: the gradient used
 * for the skewness correction and the extrapolation, out: the new gradient. */
int semo_gauss_green(semo_ctx* ctx, int field, const double* phi, const double* face_r, double* grad);
/* SEMO_TurbModel_LES (src/turbulenceModels/LES.cpp): calcWALE (:29-131),
 * calcSmagorinsky (:141-220) or calcDynamicSmagorinsky (:224-467), which
 * calcCellTransport calls (src/solvers/transport.cpp:39-56) for turbType LES.
 * The velocity gradients are calcGaussGreen's, two passes from zero, with the
 * U, V, W patch types of semo_set_case and the right face values face_rU,
 * face_rV, face_rW [n_faces] (varR of fU, fV, fW, read at boundary faces).
 * U, V, W, rho [n] cell values; muT, kSGS [n] out: the subgrid viscosity and
 * kinetic energy.  The dynamic model filters over each cell's vertex stencil
 * (the calcLeastSquare stencil).                                              */
enum { SEMO_LES_WALE = 1, SEMO_LES_SMAGORINSKY = 2, SEMO_LES_DYNAMIC_SMAGORINSKY = 3 };
/* Convective flux schemes of SEMO's density-based solvers
 * (src/solvers/flux.cpp): calcFluxHAUS (:5-156), calcFluxSLAU2 (:159-238, the
 * one calcFluxes uses, src/solvers/RHS.cpp:272-287), calcFluxRoe (:243-344,
 * its central part: the dissipation is commented out), calcFluxAPRoe
 * (:349-470), calcFluxRoeAMTP (:480-631), calcFluxUpwind (:649-692),
 * calcFluxAUSMPWP_N (:707-787), calcFluxRoeM_N (:795-922).  Per face:
 * left[n][8 + nsp] and right[n][8 + nsp] = (P, U, V, W, T, rho, C, Ht,
 * Y_0 .. Y_{nsp-1}), normal[n][3] (unit), par[n][7] = (w1, w2, w3, fC, Uco,
 * dt, Lch) for AUSMPWP_N and RoeM_N (may be NULL for the others);
 * flux[n][4 + nsp] out: (rho, rho u, rho v, rho w, rho Ht, rho Y_0 ..
 * rho Y_{nsp-2}) fluxes.  1 <= nsp <= 8.  Needs a context, not a mesh.       */
enum { SEMO_FLUX_HAUS = 0, SEMO_FLUX_SLAU2, SEMO_FLUX_ROE, SEMO_FLUX_APROE, SEMO_FLUX_ROEAMTP,
       SEMO_FLUX_UPWIND, SEMO_FLUX_AUSMPWP_N, SEMO_FLUX_ROEM_N };
int semo_flux(semo_ctx* ctx, int scheme, int64_t n, int nsp, const double* left, const double* right,
              const double* normal, const double* par, double* flux);
/* Implicit updates of SEMO's density-based dual-time solvers
 * (src/solvers/linearSolver.cpp), on the uploaded mesh:
 *   SEMO_LINSOL_POINT  calcLinearSolver (:666-725): per cell,
 *                      res <- (P + 1.5 dtPseudo / time_step T)^-1 res;
 *   SEMO_LINSOL_SINGLE calcSingleLinearSolver (:729-787): res <- T^-1 res;
 *   SEMO_LINSOL_LUSGS  calcLinearSolverLUSGS (:144-389): forward then
 *                      backward sweep in lib_RCM's cell order (RCM.cpp:34-
 *                      210) with the faces' upwind flux Jacobians
 *                      (getNumericalFluxJacobianUpwind, :394-640).
 * P and T are constructMatrix's (:791-890); blocks are inverted as
 * GaussSeidelSOR does (src/math/math.cpp:141-200, no pivoting).
 * state[n][12 + 3 nsp] per cell: P, U, V, W, T, rho, C, Ht, dRhoDP, dHtDP,
 * dRhoDT, dHtDT, Ur, dtPseudo, Y_0..Y_{nsp-1}, dRhoDY_0..dRhoDY_{nsp-2},
 * dHtDY_0..dHtDY_{nsp-2}; residuals[n][4 + nsp] in/out (the calcRHS
 * residuals in, the update out).  1 <= nsp <= 4; time_step is controls.
 * timeStep (not read by SEMO_LINSOL_SINGLE).  The LU-SGS order is the RCM
 * of the mesh as the caller numbers it, whether or not the upload reordered
 * the cells; the first LU-SGS call on a mesh levels the sweeps on the host. */
/* calcFluxes (src/solvers/RHS.cpp:98-406) of the density-based solvers on
 * the uploaded mesh: per face SLAU2 (calcFluxSLAU2, flux.cpp:159-238) of the
 * left and right records and the viscous flux (calcViscousFlux,
 * viscousFlux.cpp:5-67) of the face's velocity gradients, times the area,
 * summed into the cells in face order (+ owner, - neighbour).
 * face_left / face_right[n_faces][8 + nsp] = (P, U, V, W, T, rho, C, Ht, mu,
 * Y_0..Y_{nsp-2}) in SEMO's face order and orientation; face_grad[n_faces][9]
 * = (dU/dx, dU/dy, dU/dz, dV/dx, .., dW/dz) (face.var dUdx..dWdz), or NULL
 * for zero -- SEMO fills them nowhere (calcGradientFace is commented out),
 * so its runs see zero; residuals[n][4 + nsp] out.  2 <= nsp <= 8 (with one
 * species SEMO's calcFluxes reads a face volume fraction that does not
 * exist).                                                                    */
int semo_calc_fluxes(semo_ctx* ctx, int nsp, const double* face_left, const double* face_right,
                     const double* face_grad, double* residuals);
/* calcRHS (RHS.cpp:5-28): calcFluxes as above, then sourceTerms
 * (source.cpp:5-35) -- per cell the dual-time source (1.5 Q - 2 Qn + 0.5 Qm)
 * / time_step * V of Q = (rho, rho U, rho V, rho W, rho Ht - P, rho Y..),
 * minus rho g V in the momentum rows, minus V * st_force (the surface-tension
 * force -sigma kappa grad(alpha) per cell, [n][3], or NULL for none) -- and
 * the scaling by -dtPseudo / V.  cells[n][8 + (nsp-1) + 2(4 + nsp)] = (rho,
 * U, V, W, Ht, P, dtPseudo, VF, Y_0..Y_{nsp-2}, Qn_0.., Qm_0..); g[3].
 * SEMO's own sourceTerms computes the surface-tension force itself and, as
 * written, hands calcGaussGreen an empty gradient vector that it indexes;
 * here the caller supplies the force.                                        */
int semo_calc_rhs(semo_ctx* ctx, int nsp, double time_step, const double* g, const double* face_left,
                  const double* face_right, const double* face_grad, const double* cells, const double* st_force,
                  double* residuals);
/* updateValues (src/solvers/densityBased.cpp:366-512): cells[n][5 + nsp]
 * in/out = (P, U, V, W, T, MF_0..MF_{nsp-1}) += urf (dualTimeURF_P, _U, _T,
 * _MF: urf[4]) times the residuals[n][4 + nsp] (the implicit update), each
 * of P, U, V, W, T then held within limits[10] = (minP, maxP, minU, maxU,
 * minV, maxV, minW, maxW, minT, maxT), the first nsp - 1 mass fractions
 * clipped to [0, 1] and the last set to 1 - their sum; face_un[n_faces]
 * in/out (face.var Un) += dualTimeURF_U * the face-weighted (wC) momentum
 * increments . n at internal faces (boundary faces unchanged).  1 <= nsp <= 8.*/
int semo_update_values(semo_ctx* ctx, int nsp, const double* urf, const double* limits, const double* residuals,
                       double* cells, double* face_un);
/* SEMO's interface-compression utility (src/utility/compressVolumeFraction.
 * cpp:121-228) on the uploaded mesh: iters pseudo steps of CFL cfl that
 * move the volume fraction along its own unit normal, flux alpha(1 - alpha)
 * upwinded, alpha clipped to [0, 1].  vf[n] in/out; face_r[n_faces] the
 * boundary faces' values the LSQ gradient reads (NULL: zero, which is what
 * the utility holds after loading a case); residuals[iters] out (may be
 * NULL): each step's sqrt(sum of squared updates).                          */
int semo_compress_volume_fraction(semo_ctx* ctx, double cfl, int iters, double* vf, const double* face_r,
                                  double* residuals);
enum { SEMO_LINSOL_POINT = 0, SEMO_LINSOL_SINGLE = 1, SEMO_LINSOL_LUSGS = 2 };
int semo_linear_solver(semo_ctx* ctx, int method, int nsp, double time_step, const double* state,
                       double* residuals);
/* The same on device arrays in the context's cell numbering (the device
 * order semo_mesh_get_perm reports), on the context's stream, synchronised;
 * d_state and d_residuals as above.  No counterpart in the reference (it
 * keeps the state in the device's memory between the calls of a solver
 * loop).                                                                     */
int semo_linear_solver_device(semo_ctx* ctx, int method, int nsp, double time_step, const double* d_state,
                              double* d_residuals);
/* turbType / turbLESModel of constant/turbulenceProperties (src/controls/
 * build.cpp:100-115): model 0 = laminar (the default), else an SEMO_LES_*
 * model, which every calcCellTransport of the time step then applies
 * (src/solvers/transport.cpp:38-97): at the first step and after each VOF
 * iteration (src/solvers/pressureBased.cpp:17-19, 115-116) mu = the phases'
 * mixture + muT, with the velocity gradients taken from the boundary faces'
 * right values of that moment -- zero at the first step, as SEMO holds them
 * before its first reconstruction.                                          */
int semo_set_les(semo_ctx* ctx, int model);
int semo_les(semo_ctx* ctx, int model, const double* U, const double* V, const double* W, const double* rho,
             const double* face_rU, const double* face_rV, const double* face_rW, double* muT, double* kSGS);
/* setCellVarMinMax (src/solvers/cellVarMinMax.cpp:8-72)                        */
int semo_cell_min_max(semo_ctx* ctx, const double* phi, double* vmax, double* vmin);

/* ---- linear solve (seam 2: solveAMGCL, src/solvers/solveAMGCL.cpp:179-1271) ----
 * COO system in, solution out; equation = "pressure" (AMG-PCG family) or
 * "momentum" / "volume_fraction" (BiCGStab family).  Duplicates are summed.   */
int semo_solve_coo(semo_ctx* ctx, const char* equation, int32_t n, int64_t nnz,
                   const int32_t* rows, const int32_t* cols, const double* vals,
                   int32_t nrhs, const double* b, double* x, semo_solve_stats* st);

/* ---- profiling ------------------------------------------------------------------------
 * When enabled, the library brackets every launch of the named kernels with
 * HIP events on the context stream and accumulates count, device time and the
 * algorithmic HBM bytes of each launch.                                        */
int semo_profile_enable(semo_ctx* ctx, int on);
int semo_profile_reset(semo_ctx* ctx);
/* i-th profiled kernel: name (<= 63 chars), launches, total ms, total bytes     */
/* Average ms of one streaming copy of n doubles (read n, write n) by the
 * library's own copy kernel, over reps copies on the context's stream: the
 * practical HBM ceiling the kernels are compared with.  No counterpart in the
 * reference (a measurement aid).                                              */
int semo_stream_copy(semo_ctx* ctx, int64_t n, int reps, double* ms);

int semo_profile_entry(semo_ctx* ctx, int i, char* name, int64_t* count, double* ms,
                       double* bytes);

#ifdef __cplusplus
}
#endif
#endif /* SEMO_GPU_H */
