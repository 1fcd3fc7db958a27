// Block-coupled pressure-velocity solve: one 4x4 block per cell pair (BSR4).
//
// SEMO's calcCoupledEq (src/solvers/eqCoupled.cpp:12-1250) is commented out in
// the reference: only its design survives -- four unknowns per cell
// (p', u', v', w'), B_n = 4 (eqCoupled.cpp:26), global index
// startCellGlobal*B_n + n*var + i, i.e. variable-major (eqCoupled.cpp:249-305),
// a continuity row with the p' Laplacian and momentum rows with upwind
// convection plus the face-interpolated p' gradient (eqCoupled.cpp:470-640),
// one solveAMGCL call (:1246).  This file builds that system from the live
// assembly of this library, so its blocks are exactly:
//   momentum rows (k = 1..3): the momentum matrix of calcMomentumEqs on the
//     (k, k) entries, plus the linearised pressure force: for a face f with
//     owner P and neighbour N, row P gets +n_k A (wC p'_P + (1-wC) p'_N), row
//     N the opposite; a zeroGradient-p boundary face adds n_k A p'_P;
//   continuity row (0): the pressure-correction matrix of calcPressureEq on
//     (0, 0), plus the face flux of the velocity correction,
//     +A (wC n.u'_P + (1-wC) n.u'_N) in row P and the opposite in row N
//     (zeroGradient / outflowing inletOutlet boundary faces add A n.u'_P);
//   right-hand sides: the momentum and pressure right-hand sides.
// The solution updates p += urf_p p', U += urf_u u'.  At a converged outer
// iteration both right-hand sides vanish, so the coupled and the segregated
// (SIMPLE) iterations share their fixed point.
//
// Storage (device numbering): block CSR, diagonal blocks apart, each block 16
// contiguous values row-major; vectors interleaved [n][4] = (p, u, v, w).
// Solver: FP64 BiCGStab, right-preconditioned by a SIMPLE-type block
// factorisation whose Schur-complement step is one FP32 V-cycle of the scalar
// pressure hierarchy (see precond below).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>

#include "kern.hpp"

namespace semo {

struct Coupled {
    bool built = false;
    int n = 0;
    int64_t nnzb = 0;                           // off-diagonal blocks
    // fine level (device numbering), FP64
    std::vector<int> h_ptr, h_col;
    DBuf<int> ptr, col, fslot_own, fslot_nbr;   // per internal face: block index in row owner / neighbour
    DBuf<double> val, diag, rhs;                // [nnzb][16], [n][16], [n][4]
    DBuf<double> mval, mdiag, mrhs, pval, pdiag, prhs;   // scalar momentum / pressure systems
    DBuf<double> x;                             // [n][4] last solution
    // the couplings the preconditioner reads, FP32.  gc [3][nv]: per slot of
    // the scalar pattern (the sliced ELL of the pressure and momentum
    // matrices, entry pairs) the off-diagonal block's continuity-row u'
    // entries (0, 1..3); its momentum rows' p' entries (1..3, 0) are the same
    // three numbers (k_cpl_row writes both as sg n_k A w_j).  Padding slots
    // hold 0.  dd [6][n]: the diagonal block's (0, 1..3) and (1..3, 0)
    // entries.  im [n]: 1 / diag(M).
    DBuf<float> gc, dd;
    DBuf<double> im;
    DBuf<double> kv;                            // Krylov vectors
    DBuf<double> tp, zp;                        // Schur-complement right side / solution [n]
    DBuf<double> part, red, sc;                 // host-summed partials; device partials, scalars
    int last_iters = 0;
    double last_relres = 0.0;
};

void coupled_free(Coupled* p) { delete p; }

namespace {

// ---------------------------------------------------------------- assembly --
__global__ void __launch_bounds__(TPB) k_cpl_row(
    int nc, int nif, int nbf, const int* __restrict__ cf_sl, const int* __restrict__ cf_ell,
    const int* __restrict__ owner, const int* __restrict__ nbr, const int* __restrict__ bpatch,
    const semo_patch_bc* __restrict__ bcs, const double* __restrict__ nrm, const double* __restrict__ area,
    const double* __restrict__ wCv, const int* __restrict__ slot_own, const int* __restrict__ slot_nbr,
    const int* __restrict__ fso, const int* __restrict__ fsn, const double* __restrict__ mval,
    const double* __restrict__ mdiag, const double* __restrict__ pval, const double* __restrict__ pdiag,
    const double* __restrict__ mrhs, const double* __restrict__ prhs, const double* __restrict__ U,
    const double* __restrict__ V, const double* __restrict__ W, double* __restrict__ val, double* __restrict__ diag, double* __restrict__ rhs, int64_t nv,
    float* __restrict__ gc, float* __restrict__ dd, double* __restrict__ im) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double D[16];
    for (int q = 0; q < 16; ++q) D[q] = 0.0;
    double uc[3] = {U[i], V[i], W[i]};
    CELL_FACES_BEGIN(i, cf_sl, cf_ell, f, own)
        const double* n = nrm + 3 * (size_t)f;
        double A = area[f];
        if (f < nif) {
            double wC = wCv[f];
            // the block of this face's neighbour in row i: no other face
            // adds to it (one face per pair of cells), so it is formed in
            // registers and stored once, 16-byte stores
            double B[16];
            for (int q = 0; q < 16; ++q) B[q] = 0.0;
            int slot = own ? slot_own[f] : slot_nbr[f];
            double a = mval[slot];
            double q = pval[slot];
            B[0] += q;
            for (int k = 1; k < 4; ++k) B[5 * k] += a;
            double wi = own ? wC : 1.0 - wC, wj = 1.0 - wi, sg = own ? 1.0 : -1.0;
            for (int k = 1; k < 4; ++k) {
                double g = sg * n[k - 1] * A;
                D[k] += g * wi;       // continuity row, u'_k of this cell
                B[k] += g * wj;
                gc[(k - 1) * nv + slot] = (float)B[k];
                D[4 * k] += g * wi;   // momentum row k, p' of this cell
                B[4 * k] += g * wj;
            }
            double2* Bo = reinterpret_cast<double2*>(val + 16 * (size_t)(own ? fso[f] : fsn[f]));
            for (int q = 0; q < 8; ++q) Bo[q] = make_double2(B[2 * q], B[2 * q + 1]);
            continue;
        }
        semo_patch_bc bc = bcs[bpatch[f - nif]];
        if (bc.p_type != SEMO_BC_FIXED_VALUE)
            for (int k = 1; k < 4; ++k) D[4 * k] += n[k - 1] * A;
        bool flux = bc.u_type == SEMO_BC_ZERO_GRADIENT ||
                    (bc.u_type == SEMO_BC_INLET_OUTLET && uc[0] * n[0] + uc[1] * n[1] + uc[2] * n[2] > 0.0);
        if (flux)
            for (int k = 1; k < 4; ++k) D[k] += n[k - 1] * A;
    }
    D[0] += pdiag[i];
    for (int k = 1; k < 4; ++k) D[5 * k] += mdiag[i];
    double2* Do = reinterpret_cast<double2*>(diag + 16 * (size_t)i);
    for (int q = 0; q < 8; ++q) Do[q] = make_double2(D[2 * q], D[2 * q + 1]);
    size_t N = (size_t)nc;
    for (int k = 0; k < 3; ++k) {
        dd[k * N + i] = (float)D[1 + k];
        dd[(3 + k) * N + i] = (float)D[4 * (1 + k)];
    }
    im[i] = 1.0 / mdiag[i];
    double2* Ro = reinterpret_cast<double2*>(rhs + 4 * (size_t)i);
    Ro[0] = make_double2(prhs[i], mrhs[i]);
    Ro[1] = make_double2(mrhs[N + i], mrhs[2 * N + i]);
}

// ------------------------------------------------------------- block SpMV --
// Four lanes per block row: lane q owns output component q and reads row q of
// every block (4 contiguous values), so a quad reads each 16-value block as
// one contiguous run; x_j (4 values) is shared by the quad.
// OP 0: y = A x;  OP 1: y = b - A x
template <class T, int OP>
__global__ void __launch_bounds__(TPB) k_bsr(int n, const int* __restrict__ ptr, const int* __restrict__ col,
                                             const T* __restrict__ val, const T* __restrict__ diag,
                                             const T* __restrict__ x, const T* __restrict__ b, T* __restrict__ y) {
    int gid = xcd_block() * blockDim.x + threadIdx.x;
    int i = gid >> 2, q = gid & 3;
    if (i >= n) return;
    const T* d = diag + 16 * (size_t)i + 4 * q;
    const T* xi = x + 4 * (size_t)i;
    T s = d[0] * xi[0] + d[1] * xi[1] + d[2] * xi[2] + d[3] * xi[3];
    for (int e = ptr[i]; e < ptr[i + 1]; ++e) {
        const T* a = val + 16 * (size_t)e + 4 * q;
        const T* xj = x + 4 * (size_t)col[e];
        s += a[0] * xj[0] + a[1] * xj[1] + a[2] * xj[2] + a[3] * xj[3];
    }
    size_t o = 4 * (size_t)i + q;
    y[o] = OP == 0 ? s : b[o] - s;
}

// ------------------------------------------ SIMPLE-type block preconditioner --
// Block factorisation of K = [[Lp, D], [G, M]] (continuity row: the
// pressure-correction matrix Lp on p', the velocity-correction flux D on u';
// momentum rows: the pressure force G on p', the momentum matrix M on u'):
//   u* = diag(M)^-1 r_u;  p' = S^-1 (r_p - D u*);  u' = u* - diag(M)^-1 G p'
// with the Schur complement S = Lp - D diag(M)^-1 G replaced by Lp, applied
// as one FP32 V-cycle of the scalar pressure hierarchy.  (Scaling that
// approximation by 1.3, 1.6 or 2 -- the part D diag(M)^-1 G adds to smooth
// modes -- raised the BiCGStab iterations of the 1.05M-cell channel from
// 11.6 to 12.3, 14.1 and 15.2 per outer iteration.)
// The two kernels around the V-cycle walk the scalar pattern one lane per
// row (sliced ELL, entry pairs: the three coupling arrays and the columns
// are one 8-byte load per lane per pair, a wavefront's load one contiguous
// run); u* is formed where it is read (r_u 1/diag(M), the same product the
// correction forms), so no pass writes it.

// the columns of entries e, e + 1 (e even) of the scalar pattern, from
// int16 offsets column - row or from int32 columns
template <bool NARROW>
__device__ inline int2 pat_cols(const int* __restrict__ col, const int16_t* __restrict__ dcol, int e, int row) {
    if constexpr (NARROW) {
        int w = *reinterpret_cast<const int*>(dcol + e);
        return make_int2(row + (int)(int16_t)(w & 0xffff), row + (w >> 16));
    } else {
        return *reinterpret_cast<const int2*>(col + e);
    }
}

constexpr int PC_PAIRS = 2;   // entry pairs per trip: their loads, then their gathers

// tp = r_p - D u*
template <bool NARROW>
__global__ void __launch_bounds__(TPB) k_sp_schur_rhs(int n, const int* __restrict__ sl_ptr,
                                                      const int* __restrict__ col, const int16_t* __restrict__ dcol,
                                                      const float* __restrict__ gc, int64_t nv,
                                                      const float* __restrict__ dd, const double* __restrict__ r,
                                                      const double* __restrict__ im, double* __restrict__ tp) {
    int row = xcd_block() * TPB + threadIdx.x;
    int s = row >> 6, lane = row & (SL - 1);
    if (s >= (n + SL - 1) / SL) return;
    double t = 0.0;
    if (row < n) {
        double2 r01 = *reinterpret_cast<const double2*>(r + 4 * (size_t)row);
        double2 r23 = *reinterpret_cast<const double2*>(r + 4 * (size_t)row + 2);
        double m = im[row];
        t = r01.x;
        t -= (double)dd[row] * (r01.y * m) + (double)dd[(size_t)n + row] * (r23.x * m) +
             (double)dd[2 * (size_t)n + row] * (r23.y * m);
    }
    int end = sl_ptr[s + 1];
    for (int e = sl_ptr[s] + 2 * lane; e < end; e += 2 * SL * PC_PAIRS) {
        int2 j[PC_PAIRS];
        float2 a0[PC_PAIRS], a1[PC_PAIRS], a2[PC_PAIRS];
#pragma unroll
        for (int u = 0; u < PC_PAIRS; ++u) {
            int ee = e + 2 * SL * u;
            if (ee < end) {
                j[u] = pat_cols<NARROW>(col, dcol, ee, row);
                a0[u] = *reinterpret_cast<const float2*>(gc + ee);
                a1[u] = *reinterpret_cast<const float2*>(gc + nv + ee);
                a2[u] = *reinterpret_cast<const float2*>(gc + 2 * nv + ee);
            } else {
                j[u] = make_int2(-1, -1);
            }
        }
        double2 x0[PC_PAIRS], x1[PC_PAIRS], y0[PC_PAIRS], y1[PC_PAIRS];
        double m0[PC_PAIRS], m1[PC_PAIRS];
#pragma unroll
        for (int u = 0; u < PC_PAIRS; ++u)
            if (j[u].x >= 0) {
                x0[u] = *reinterpret_cast<const double2*>(r + 4 * (size_t)j[u].x);
                x1[u] = *reinterpret_cast<const double2*>(r + 4 * (size_t)j[u].x + 2);
                y0[u] = *reinterpret_cast<const double2*>(r + 4 * (size_t)j[u].y);
                y1[u] = *reinterpret_cast<const double2*>(r + 4 * (size_t)j[u].y + 2);
                m0[u] = im[j[u].x];
                m1[u] = im[j[u].y];
            }
#pragma unroll
        for (int u = 0; u < PC_PAIRS; ++u)
            if (j[u].x >= 0) {
                t -= (double)a0[u].x * (x0[u].y * m0[u]) + (double)a1[u].x * (x1[u].x * m0[u]) +
                     (double)a2[u].x * (x1[u].y * m0[u]);
                t -= (double)a0[u].y * (y0[u].y * m1[u]) + (double)a1[u].y * (y1[u].x * m1[u]) +
                     (double)a2[u].y * (y1[u].y * m1[u]);
            }
    }
    if (row < n) tp[row] = t;
}
// z_p = p', z_u = u* - diag(M)^-1 G p'
template <bool NARROW>
__global__ void __launch_bounds__(TPB) k_sp_correct(int n, const int* __restrict__ sl_ptr,
                                                    const int* __restrict__ col, const int16_t* __restrict__ dcol,
                                                    const float* __restrict__ gc, int64_t nv,
                                                    const float* __restrict__ dd, const double* __restrict__ r,
                                                    const double* __restrict__ im, const double* __restrict__ zp,
                                                    double* __restrict__ z) {
    int row = xcd_block() * TPB + threadIdx.x;
    int s = row >> 6, lane = row & (SL - 1);
    if (s >= (n + SL - 1) / SL) return;
    double p = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
    if (row < n) {
        p = zp[row];
        g0 = (double)dd[3 * (size_t)n + row] * p;
        g1 = (double)dd[4 * (size_t)n + row] * p;
        g2 = (double)dd[5 * (size_t)n + row] * p;
    }
    int end = sl_ptr[s + 1];
    for (int e = sl_ptr[s] + 2 * lane; e < end; e += 2 * SL * PC_PAIRS) {
        int2 j[PC_PAIRS];
        float2 a0[PC_PAIRS], a1[PC_PAIRS], a2[PC_PAIRS];
#pragma unroll
        for (int u = 0; u < PC_PAIRS; ++u) {
            int ee = e + 2 * SL * u;
            if (ee < end) {
                j[u] = pat_cols<NARROW>(col, dcol, ee, row);
                a0[u] = *reinterpret_cast<const float2*>(gc + ee);
                a1[u] = *reinterpret_cast<const float2*>(gc + nv + ee);
                a2[u] = *reinterpret_cast<const float2*>(gc + 2 * nv + ee);
            } else {
                j[u] = make_int2(-1, -1);
            }
        }
        double q0[PC_PAIRS], q1[PC_PAIRS];
#pragma unroll
        for (int u = 0; u < PC_PAIRS; ++u)
            if (j[u].x >= 0) { q0[u] = zp[j[u].x]; q1[u] = zp[j[u].y]; }
#pragma unroll
        for (int u = 0; u < PC_PAIRS; ++u)
            if (j[u].x >= 0) {
                g0 += (double)a0[u].x * q0[u]; g1 += (double)a1[u].x * q0[u]; g2 += (double)a2[u].x * q0[u];
                g0 += (double)a0[u].y * q1[u]; g1 += (double)a1[u].y * q1[u]; g2 += (double)a2[u].y * q1[u];
            }
    }
    if (row >= n) return;
    double m = im[row];
    double2 r01 = *reinterpret_cast<const double2*>(r + 4 * (size_t)row);
    double2 r23 = *reinterpret_cast<const double2*>(r + 4 * (size_t)row + 2);
    double* zi = z + 4 * (size_t)row;
    *reinterpret_cast<double2*>(zi) = make_double2(p, r01.y * m - g0 * m);
    *reinterpret_cast<double2*>(zi + 2) = make_double2(r23.x * m - g1 * m, r23.y * m - g2 * m);
}

// ----------------------------------------------------------- Krylov kernels --
// two dot products, block partials (fixed order): part[2b] = a.b, part[2b+1] = c.d
__global__ void __launch_bounds__(TPB) k_dot2v(size_t n, const double* __restrict__ a, const double* __restrict__ b,
                                               const double* __restrict__ c, const double* __restrict__ d,
                                               double* __restrict__ part) {
    size_t i = (size_t)xcd_block() * blockDim.x + threadIdx.x;
    double s0 = 0.0, s1 = 0.0;
    for (size_t k = i; k < n; k += (size_t)gridDim.x * blockDim.x) {
        s0 += a[k] * b[k];
        if (c) s1 += c[k] * d[k];
    }
    block_sum2(s0, s1, part);
}

// P += urf_p x0, U += urf_u x(1..3), clipped; partial sums of |dx|^2 and |x_new|^2
__global__ void __launch_bounds__(TPB) k_cpl_update(int nc, double urf_p, double urf_u, double lim,
                                                    const double* __restrict__ x, double* __restrict__ P,
                                                    double* __restrict__ U, double* __restrict__ V,
                                                    double* __restrict__ W, double* __restrict__ part) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    double nd = 0.0, np = 0.0;
    if (i < nc) {
        const double* xi = x + 4 * (size_t)i;
        double* f[4] = {P, U, V, W};
        for (int q = 0; q < 4; ++q) {
            double d = (q == 0 ? urf_p : urf_u) * xi[q];
            double v = f[q][i] + d;
            if (v <= -lim) v = -lim;
            if (v >= lim) v = lim;
            f[q][i] = v;
            nd += d * d; np += v * v;
        }
    }
    block_sum2(nd, np, part);
}

// device scalars of the block BiCGStab (zero denominators give zero steps,
// so that an iteration issued past convergence leaves x unchanged)
enum { C_RHO = 0, C_ALPHA, C_OMEGA, C_BETA, C_RR, C_BB, C_N };
enum { F_STORE_BB = 0, F_RHO, F_ALPHA, F_OMEGA, F_RR, F_RESTART };

__global__ void __launch_bounds__(TPB) k_cfinish(int nb, const double* __restrict__ part, double* __restrict__ sc,
                                                 int what) {
    double a0 = 0.0, a1 = 0.0;
    for (int i = threadIdx.x; i < nb; i += TPB) { a0 += part[2 * i]; a1 += part[2 * i + 1]; }
    __shared__ double sh[2][TPB];
    sh[0][threadIdx.x] = a0; sh[1][threadIdx.x] = a1;
    __syncthreads();
    for (int w = TPB / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) { sh[0][threadIdx.x] += sh[0][threadIdx.x + w]; sh[1][threadIdx.x] += sh[1][threadIdx.x + w]; }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    double d0 = sh[0][0], d1 = sh[1][0];
    switch (what) {
    case F_STORE_BB: sc[C_BB] = d0; sc[C_RR] = d0; sc[C_RHO] = 1.0; sc[C_ALPHA] = 1.0; sc[C_OMEGA] = 1.0; break;
    case F_RHO:      // d0 = rh.r
        sc[C_BETA] = (sc[C_RHO] != 0.0 && sc[C_OMEGA] != 0.0) ? (d0 / sc[C_RHO]) * (sc[C_ALPHA] / sc[C_OMEGA]) : 0.0;
        sc[C_RHO] = d0;
        break;
    case F_ALPHA: sc[C_ALPHA] = d0 != 0.0 ? sc[C_RHO] / d0 : 0.0; break;     // d0 = rh.v
    case F_OMEGA: sc[C_OMEGA] = d1 != 0.0 ? d0 / d1 : 0.0; break;            // d0 = t.s, d1 = t.t
    case F_RR: sc[C_RR] = d0; break;
    case F_RESTART: sc[C_RR] = d0; sc[C_RHO] = 1.0; sc[C_ALPHA] = 1.0; sc[C_OMEGA] = 1.0; break;
    default: break;
    }
}

// p = r + beta (p - omega v)
__global__ void __launch_bounds__(TPB) k_cb_p(size_t n, const double* __restrict__ sc, const double* __restrict__ r,
                                              const double* __restrict__ v, double* __restrict__ p) {
    size_t i = (size_t)xcd_block() * blockDim.x + threadIdx.x;
    if (i >= n) return;
    p[i] = r[i] + sc[C_BETA] * (p[i] - sc[C_OMEGA] * v[i]);
}
// s = r - alpha v
__global__ void __launch_bounds__(TPB) k_cb_s(size_t n, const double* __restrict__ sc, const double* __restrict__ r,
                                              const double* __restrict__ v, double* __restrict__ s) {
    size_t i = (size_t)xcd_block() * blockDim.x + threadIdx.x;
    if (i >= n) return;
    s[i] = r[i] - sc[C_ALPHA] * v[i];
}
// x += alpha ph + omega sh; r = s - omega t; partials of r.r
__global__ void __launch_bounds__(TPB) k_cb_x(size_t n, const double* __restrict__ sc, const double* __restrict__ ph,
                                              const double* __restrict__ sh, const double* __restrict__ s,
                                              const double* __restrict__ t, double* __restrict__ x,
                                              double* __restrict__ r, double* __restrict__ part) {
    size_t i = (size_t)xcd_block() * blockDim.x + threadIdx.x;
    double rr = 0.0;
    if (i < n) {
        double a = sc[C_ALPHA], w = sc[C_OMEGA];
        x[i] += a * ph[i] + w * sh[i];
        double ri = s[i] - w * t[i];
        r[i] = ri;
        rr = ri * ri;
    }
    block_sum2(rr, 0.0, part);
}
// restart from the current x (r = b - A x already formed): rh = r,
// p = v = 0; partials of r.r
__global__ void __launch_bounds__(TPB) k_cb_restart(size_t n, const double* __restrict__ r, double* __restrict__ rh,
                                                    double* __restrict__ p, double* __restrict__ v,
                                                    double* __restrict__ part) {
    size_t i = (size_t)xcd_block() * blockDim.x + threadIdx.x;
    double rr = 0.0;
    if (i < n) {
        double ri = r[i];
        rh[i] = ri; p[i] = 0.0; v[i] = 0.0;
        rr = ri * ri;
    }
    block_sum2(rr, 0.0, part);
}
// r = rh = b, x = p = v = 0; partials of b.b
__global__ void __launch_bounds__(TPB) k_cb_init(size_t n, const double* __restrict__ b, double* __restrict__ x,
                                                 double* __restrict__ r, double* __restrict__ rh,
                                                 double* __restrict__ p, double* __restrict__ v,
                                                 double* __restrict__ part) {
    size_t i = (size_t)xcd_block() * blockDim.x + threadIdx.x;
    double bb = 0.0;
    if (i < n) {
        double bi = b[i];
        r[i] = bi; rh[i] = bi; x[i] = 0.0; p[i] = 0.0; v[i] = 0.0;
        bb = bi * bi;
    }
    block_sum2(bb, 0.0, part);
}

// --------------------------------------------------------------- host side --
int nb_of(size_t n) { return (int)std::min<size_t>((n + TPB - 1) / TPB, 2048); }


double2 dot2(semo_ctx* c, Coupled& C, size_t n, const double* a, const double* b, const double* cc,
             const double* d) {
    int nb = nb_of(n);
    SEMO_LAUNCH(c, "cpl_dot", (cc ? 32.0 : 16.0) * n, k_dot2v, dim3(nb), dim3(TPB), n, a, b, cc, d, C.part.p);
    std::vector<double> h(2 * (size_t)nb);
    SEMO_HIP(hipMemcpyAsync(h.data(), C.part.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SEMO_HIP(hipStreamSynchronize(c->stream));
    double2 r{0.0, 0.0};
    for (int k = 0; k < nb; ++k) { r.x += h[2 * k]; r.y += h[2 * k + 1]; }
    return r;
}

// z = P^-1 r with the SIMPLE-type factorisation (FP64 vectors, FP32
// couplings and pressure V-cycle)
void precond(semo_ctx* c, Coupled& C, const double* r, double* z) {
    int n = C.n;
    const DevSell<double>& A = c->Amat;
    int64_t nv = (int64_t)c->pat.col.size();
    int nb = nblk((long)c->pat.nslices * SL);
    bool narrow = A.dcol.p != nullptr;
    double idx = narrow ? 2.0 : 4.0;
    if (narrow)
        SEMO_LAUNCH(c, "cpl_precond_schur_rhs", (idx + 12.0) * nv + 60.0 * n, k_sp_schur_rhs<true>, dim3(nb),
                    dim3(TPB), n, A.sl_ptr.p, A.col.p, A.dcol.p, C.gc.p, nv, C.dd.p, r, C.im.p, C.tp.p);
    else
        SEMO_LAUNCH(c, "cpl_precond_schur_rhs", (idx + 12.0) * nv + 60.0 * n, k_sp_schur_rhs<false>, dim3(nb),
                    dim3(TPB), n, A.sl_ptr.p, A.col.p, A.dcol.p, C.gc.p, nv, C.dd.p, r, C.im.p, C.tp.p);
    pressure_vcycle(c, C.tp.p, C.zp.p, true);
    if (narrow)
        SEMO_LAUNCH(c, "cpl_precond_correct", (idx + 12.0) * nv + 84.0 * n, k_sp_correct<true>, dim3(nb),
                    dim3(TPB), n, A.sl_ptr.p, A.col.p, A.dcol.p, C.gc.p, nv, C.dd.p, r, C.im.p, C.zp.p, z);
    else
        SEMO_LAUNCH(c, "cpl_precond_correct", (idx + 12.0) * nv + 84.0 * n, k_sp_correct<false>, dim3(nb),
                    dim3(TPB), n, A.sl_ptr.p, A.col.p, A.dcol.p, C.gc.p, nv, C.dd.p, r, C.im.p, C.zp.p, z);
}

void spmv64(semo_ctx* c, Coupled& C, const double* x, double* y) {
    SEMO_LAUNCH(c, "cpl_bsr_spmv", 4.0 * C.n + 132.0 * C.nnzb + 128.0 * C.n + 64.0 * C.n, (k_bsr<double, 0>),
                dim3(nblk(4 * (long)C.n)), dim3(TPB), C.n, C.ptr.p, C.col.p, C.val.p, C.diag.p, x,
                (const double*)nullptr, y);
}

// fine pattern and face -> block maps (once per mesh)
void coupled_setup(semo_ctx* c, Coupled& C) {
    int n = c->nc, nif = c->nif;
    C.n = n;
    std::vector<int> ptr(n + 1, 0);
    for (int f = 0; f < nif; ++f) { ptr[c->h_owner[f] + 1]++; ptr[c->h_nbr[f] + 1]++; }
    for (int i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
    std::vector<std::pair<int, int>> ent(ptr[n]);
    {
        std::vector<int> pos(ptr.begin(), ptr.end() - 1);
        for (int f = 0; f < nif; ++f) {
            ent[pos[c->h_owner[f]]++] = {c->h_nbr[f], f};
            ent[pos[c->h_nbr[f]]++] = {c->h_owner[f], ~f};
        }
    }
    std::vector<int> col(ptr[n]), fso(std::max(nif, 1)), fsn(std::max(nif, 1));
    for (int i = 0; i < n; ++i) {
        std::sort(ent.begin() + ptr[i], ent.begin() + ptr[i + 1]);
        for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
            col[k] = ent[k].first;
            int f = ent[k].second;
            if (f >= 0) fso[f] = k; else fsn[~f] = k;
        }
    }
    C.h_ptr = ptr; C.h_col = col;
    hipStream_t s = c->stream;
    C.ptr.upload(ptr, s); C.col.upload(col, s); C.fslot_own.upload(fso, s); C.fslot_nbr.upload(fsn, s);
    size_t nnzb = col.size();
    C.val.alloc(16 * std::max<size_t>(nnzb, 1)); C.diag.alloc(16 * (size_t)n); C.rhs.alloc(4 * (size_t)n);
    C.x.alloc(4 * (size_t)n); C.x.zero(s);
    size_t nv = c->pat.col.size();
    for (DBuf<double>* b : {&C.mval, &C.pval}) b->alloc(std::max<size_t>(nv, 1));
    for (DBuf<double>* b : {&C.mdiag, &C.pdiag, &C.prhs}) b->alloc(n);
    C.mrhs.alloc(3 * (size_t)n);
    C.kv.alloc(8 * 4 * (size_t)n);
    C.tp.alloc((size_t)n + 64); C.zp.alloc((size_t)n + 64);
    C.part.alloc(2 * 2048 + 64);
    C.red.alloc(2 * (size_t)nblk(4 * (long)n) + 64);
    C.sc.alloc(16);

    C.nnzb = (int64_t)nnzb;
    C.gc.alloc(3 * std::max<size_t>(nv, 1));
    C.gc.zero(s);                               // padding slots stay 0
    C.dd.alloc(6 * (size_t)n);
    C.im.alloc(n);
    SEMO_HIP(hipStreamSynchronize(s));
    C.built = true;
}

}  // namespace

// momentum + pressure assembly of the current state, then the block system
void assemble_coupled(semo_ctx* c) {
    for (const semo_patch_bc& b : c->bcs)
        SEMO_CHECK(b.p_type != SEMO_BC_SWITCH, SEMO_EINVAL_BC,
                   "coupled system: the Mach-switched pressure boundary has no block form here");
    if (!c->cpl) c->cpl = new Coupled();
    Coupled& C = *c->cpl;
    int n = c->nc;
    size_t nv = c->pat.col.size();
    assemble_pressure(c);
    if (!C.built) coupled_setup(c, C);
    // the Schur-complement preconditioner's hierarchy: built for the
    // pressure-correction operator, its finest level refreshed on the device
    // when the values change (the outer Krylov iteration sees the exact system)
    pressure_amg_update(c, [](const Amg&) { return true; });
    copy_d(c, C.pval.p, c->Amat.val.p, nv);
    copy_d(c, C.pdiag.p, c->Amat.diag.p, n);
    copy_d(c, C.prhs.p, c->Bvec.p, n);
    assemble_momentum(c);
    copy_d(c, C.mval.p, c->Amat.val.p, nv);
    copy_d(c, C.mdiag.p, c->Amat.diag.p, n);
    copy_d(c, C.mrhs.p, c->Bvec.p, 3 * (size_t)n);
    SEMO_LAUNCH(c, "cpl_assemble",
                cell_faces_bytes(c, 56.0) + 32.0 * c->nif + 128.0 * ((double)n + C.nnzb) + 104.0 * n +
                    24.0 * c->nif + 32.0 * n,
                k_cpl_row, dim3(nblk(n)), dim3(TPB), n, c->nif, c->nbf, c->d_cf_sl.p, c->d_cf_ell.p, c->d_owner.p,
                c->d_nbr.p, c->d_bpatch.p, c->d_bcs.p, c->f_nrm.p, c->f_area.p, c->f_wC.p, c->d_slot_own.p,
                c->d_slot_nbr.p, C.fslot_own.p, C.fslot_nbr.p, C.mval.p, C.mdiag.p, C.pval.p, C.pdiag.p, C.mrhs.p,
                C.prhs.p, c->U.p, c->V.p, c->W.p, C.val.p, C.diag.p, C.rhs.p, (int64_t)nv, C.gc.p, C.dd.p,
                C.im.p);
    c->last_eq = 3;
}

void coupled_spmv(semo_ctx* c, const double* x, double* y) {
    SEMO_CHECK(c->cpl && c->cpl->built, SEMO_ESTATE, "no coupled system assembled");
    spmv64(c, *c->cpl, x, y);
}

// right-preconditioned BiCGStab, x0 = 0, stop at ||b - A x|| <= tol ||b||.
// The Krylov scalars stay on the device; the host reads one residual per
// iteration, one iteration behind (pinned slot + event), so the stream never
// drains while it waits.  Restarted from the current iterate (true residual,
// new shadow vector) when RESTART_AFTER iterations have not halved the
// smallest residual seen since the last (re)start: in transients the
// unrestarted iteration was seen to stall for hundreds of iterations.
constexpr int RESTART_AFTER = 30;
void coupled_solve(semo_ctx* c, const double* b, double* x, semo_solve_stats* st) {
    SEMO_CHECK(c->cpl && c->cpl->built, SEMO_ESTATE, "no coupled system assembled");
    Coupled& C = *c->cpl;
    size_t N = 4 * (size_t)C.n;
    double* r = C.kv.p; double* rh = r + N; double* p = rh + N; double* v = p + N; double* ph = v + N;
    double* sv = ph + N; double* sh = sv + N; double* t = sh + N;
    double* sc = C.sc.p;
    int nbv = nblk((long)N);
    auto fin = [&](int nb, int what) {
        SEMO_LAUNCH(c, "cpl_finish", 16.0 * nb + 8.0, k_cfinish, dim3(1), dim3(TPB), nb, C.red.p, sc, what);
    };
    auto dot = [&](const double* a_, const double* b_, const double* c_, const double* d_, int what) {
        int nb = nb_of(N);
        SEMO_LAUNCH(c, "cpl_dot", (c_ ? 32.0 : 16.0) * N, k_dot2v, dim3(nb), dim3(TPB), N, a_, b_, c_, d_, C.red.p);
        fin(nb, what);
    };
    SEMO_LAUNCH(c, "cpl_bicg_init", 48.0 * N, k_cb_init, dim3(nbv), dim3(TPB), N, b, x, r, rh, p, v, C.red.p);
    fin(nbv, F_STORE_BB);
    double bb;
    SEMO_HIP(hipMemcpyAsync(&bb, sc + C_BB, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SEMO_HIP(hipStreamSynchronize(c->stream));
    int it = 0;
    if (bb > 0.0) {
        double nb = std::sqrt(bb), tol = c->sp.tol_u;
        ConvCheck cc(c, tol * tol * bb, bb);
        double best = bb;
        int since = 0;
        for (it = 1; it <= c->sp.maxit_u; ++it) {
            dot(rh, r, nullptr, nullptr, F_RHO);
            SEMO_LAUNCH(c, "cpl_bicg_p", 32.0 * N, k_cb_p, dim3(nbv), dim3(TPB), N, sc, r, v, p);
            precond(c, C, p, ph);
            spmv64(c, C, ph, v);
            dot(rh, v, nullptr, nullptr, F_ALPHA);
            SEMO_LAUNCH(c, "cpl_bicg_s", 24.0 * N, k_cb_s, dim3(nbv), dim3(TPB), N, sc, r, v, sv);
            precond(c, C, sv, sh);
            spmv64(c, C, sh, t);
            dot(t, sv, t, t, F_OMEGA);
            SEMO_LAUNCH(c, "cpl_bicg_x", 56.0 * N, k_cb_x, dim3(nbv), dim3(TPB), N, sc, ph, sh, sv, t, x, r, C.red.p);
            fin(nbv, F_RR);
            pipe_push(c, sc + C_RR, it);
            if (cc.done(it)) break;
            if (cc.r1 < 0.25 * best) { best = cc.r1; since = 0; }       // squared residuals
            else if (++since >= RESTART_AFTER) {
                SEMO_LAUNCH(c, "cpl_bsr_resid", 4.0 * C.n + 132.0 * C.nnzb + 128.0 * C.n + 64.0 * C.n + 32.0 * C.n,
                            (k_bsr<double, 1>), dim3(nblk(4 * (long)C.n)), dim3(TPB), C.n, C.ptr.p, C.col.p, C.val.p,
                            C.diag.p, x, b, r);
                SEMO_LAUNCH(c, "cpl_bicg_restart", 32.0 * N, k_cb_restart, dim3(nbv), dim3(TPB), N, r, rh, p, v,
                            C.red.p);
                fin(nbv, F_RESTART);
                since = 0;
                best = cc.r1;
            }
        }
        it = std::min(it, c->sp.maxit_u);
        // true residual
        SEMO_LAUNCH(c, "cpl_bsr_resid", 4.0 * C.n + 132.0 * C.nnzb + 128.0 * C.n + 64.0 * C.n + 32.0 * C.n,
                    (k_bsr<double, 1>), dim3(nblk(4 * (long)C.n)), dim3(TPB), C.n, C.ptr.p, C.col.p, C.val.p,
                    C.diag.p, x, b, r);
        double rr = dot2(c, C, N, r, r, nullptr, nullptr).x;
        C.last_relres = std::sqrt(rr) / nb;
    } else {
        C.last_relres = 0.0;
    }
    C.last_iters = it;
    if (st) {
        *st = semo_solve_stats{};
        st->iters[0] = it;
        st->rel_res[0] = C.last_relres;
        st->n_levels = (int)c->amg_p.lv.size();
    }
    check_converged(bb > 0.0 ? C.last_relres : bb, c->sp.tol_u, "coupled");
}

void coupled_solve_assembled(semo_ctx* c, semo_solve_stats* st) {
    SEMO_CHECK(c->cpl && c->cpl->built, SEMO_ESTATE, "no coupled system assembled");
    coupled_solve(c, c->cpl->rhs.p, c->cpl->x.p, st);
}

double coupled_update(semo_ctx* c) {
    Coupled& C = *c->cpl;
    int nb = nblk(c->nc);
    SEMO_LAUNCH(c, "cpl_update", 96.0 * c->nc + 16.0 * nb, k_cpl_update, dim3(nb), dim3(TPB), c->nc, c->cs.urf_p,
                c->cs.urf_mom, c->cs.limit, C.x.p, c->P.p, c->U.p, c->V.p, c->W.p, c->red.p);
    std::vector<double> h(2 * (size_t)nb);
    SEMO_HIP(hipMemcpyAsync(h.data(), c->red.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SEMO_HIP(hipStreamSynchronize(c->stream));
    double a = 0, b = 0;
    for (int k = 0; k < nb; ++k) { a += h[2 * k]; b += h[2 * k + 1]; }
    return b > 0.0 ? std::sqrt(a) / std::sqrt(b) : std::sqrt(a);
}

// block COO (diagonal blocks first, then row by row) and rhs, device numbering
void coupled_export(semo_ctx* c, std::vector<int>& rows, std::vector<int>& cols, std::vector<double>& blocks,
                    std::vector<double>& rhs) {
    SEMO_CHECK(c->cpl && c->cpl->built, SEMO_ESTATE, "no coupled system assembled");
    Coupled& C = *c->cpl;
    int n = C.n;
    size_t nnzb = C.h_col.size();
    std::vector<double> d(16 * (size_t)n), v(16 * nnzb);
    rhs.resize(4 * (size_t)n);
    C.diag.download(d.data(), d.size(), c->stream);
    if (nnzb) C.val.download(v.data(), v.size(), c->stream);
    C.rhs.download(rhs.data(), rhs.size(), c->stream);
    SEMO_HIP(hipStreamSynchronize(c->stream));
    rows.clear(); cols.clear(); blocks.clear();
    for (int i = 0; i < n; ++i) { rows.push_back(i); cols.push_back(i); }
    blocks = d;
    for (int i = 0; i < n; ++i)
        for (int k = C.h_ptr[i]; k < C.h_ptr[i + 1]; ++k) { rows.push_back(i); cols.push_back(C.h_col[k]); }
    blocks.insert(blocks.end(), v.begin(), v.end());
}

}  // namespace semo
