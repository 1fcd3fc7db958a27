// Linear algebra and Krylov solvers on the device.
//
// * sliced-ELL SpMV (one wavefront per 64-row slice, coalesced val/col loads),
//   FP64 and FP32 values;
// * deterministic two-stage dot products (wavefront shuffles -> per-block
//   partials -> one ordered final sum), results kept in device memory so the
//   Krylov scalars never leave the GPU except for the convergence test;
// * AMG V-cycle with Chebyshev smoothing, smoothed-aggregation transfer
//   operators and a dense FP64 coarsest solve (explicit inverse, one GEMV);
// * flexible PCG for the pressure (FP64 AMG, or FP32 AMG inside the FP64
//   Krylov loop) and Jacobi-BiCGStab for momentum.
//
// Replaces the solve half of solveAMGCL (src/solvers/solveAMGCL.cpp:179-1271)
// and AMGCL's amg/idrs/fgmres (lib/amgcl/mpi/amg.hpp:342-380,
// lib/amgcl/solver/idrs.hpp, fgmres.hpp).
#include <chrono>
#include <cmath>

#include "common.hpp"

namespace semo {

struct HCsr;
void amg_build_csr(semo_ctx* c, Amg& amg, int n, const std::vector<int>& ptr,
                   const std::vector<int>& col, const std::vector<double>& val);

namespace {

constexpr int TPB = 256;
constexpr int RB = 1024;   // max reduction blocks
inline int nblk(long n) { return (int)((n + TPB - 1) / TPB); }
inline int rblk(long n) { return (int)std::min<long>(RB, std::max<long>(1, (n + TPB - 1) / TPB)); }

// ------------------------------------------------------------------- SpMV --
// y = beta_y*y + alpha*(D x + O x) with D the diagonal (optional) and O the
// sliced-ELL off-diagonal part; MODE 0: y = A x, 1: y = b - A x, 2: y += A x
template <class T, class TX, class TY, int MODE>
__global__ void __launch_bounds__(TPB) k_spmv(int n, const int* __restrict__ sl_ptr,
                                              const int* __restrict__ col, const T* __restrict__ val,
                                              const T* __restrict__ diag, const TX* __restrict__ x,
                                              const TY* __restrict__ b, TY* __restrict__ y) {
    int row = blockIdx.x * blockDim.x + threadIdx.x;
    int s = row >> 6, lane = row & 63;
    int nsl = (n + 63) >> 6;
    if (s >= nsl) return;
    int e0 = sl_ptr[s], e1 = sl_ptr[s + 1];
    TY acc = 0;
    for (int e = e0 + lane; e < e1; e += 64) acc += (TY)val[e] * (TY)x[col[e]];
    if (row >= n) return;
    if (diag) acc += (TY)diag[row] * (TY)x[row];
    if (MODE == 0) y[row] = acc;
    else if (MODE == 1) y[row] = b[row] - acc;
    else y[row] += acc;
}

template <class T, class TX, class TY, int MODE>
void spmv(semo_ctx* c, const DevSell<T>& A, const TX* x, const TY* b, TY* y, bool with_diag = true,
          const char* tag = nullptr) {
    if (A.n == 0) return;
    hipEvent_t ev{};
    bool prof = c->prof.on && tag;
    if (prof) c->prof.begin(c->stream, ev);
    int nrows = A.nslices * 64;
    hipLaunchKernelGGL((k_spmv<T, TX, TY, MODE>), dim3(nblk(nrows)), dim3(TPB), 0, c->stream, A.n,
                       A.sl_ptr.p, A.col.p, A.val.p, with_diag ? A.diag.p : (const T*)nullptr, x, b, y);
    if (prof) {
        // algorithmic bytes: values + columns of the real entries, x and y once
        double bytes = (double)A.nnz * (sizeof(T) + 4) + (double)A.n * (sizeof(TX) + sizeof(TY) * (MODE == 0 ? 1 : 2)) +
                       4.0 * (A.nslices + 1);
        c->prof.end(c->stream, tag, ev, bytes);
    }
}

// ------------------------------------------------------------- reductions --
template <class T>
__device__ inline T wave_sum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// partial sums of a.b and c.d (c,d may be null) over a grid-stride range
template <class TA>
__global__ void __launch_bounds__(TPB) k_dot2(int n, const TA* __restrict__ a, const TA* __restrict__ b,
                                              const TA* __restrict__ cc, const TA* __restrict__ d,
                                              double* __restrict__ part) {
    double s1 = 0.0, s2 = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        s1 += (double)a[i] * (double)b[i];
        if (cc) s2 += (double)cc[i] * (double)d[i];
    }
    __shared__ double sh1[TPB / 64], sh2[TPB / 64];
    s1 = wave_sum(s1); s2 = wave_sum(s2);
    if ((threadIdx.x & 63) == 0) { sh1[threadIdx.x >> 6] = s1; sh2[threadIdx.x >> 6] = s2; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t1 = 0, t2 = 0;
        for (int w = 0; w < TPB / 64; ++w) { t1 += sh1[w]; t2 += sh2[w]; }
        part[2 * blockIdx.x] = t1; part[2 * blockIdx.x + 1] = t2;
    }
}

// ordered final sum of nb partial pairs into out[0], out[1]
__global__ void __launch_bounds__(TPB) k_dot_final(int nb, const double* __restrict__ part,
                                                   double* __restrict__ out) {
    double s1 = 0.0, s2 = 0.0;
    for (int i = threadIdx.x; i < nb; i += TPB) { s1 += part[2 * i]; s2 += part[2 * i + 1]; }
    __shared__ double sh1[TPB], sh2[TPB];
    sh1[threadIdx.x] = s1; sh2[threadIdx.x] = s2;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t1 = 0, t2 = 0;
        for (int w = 0; w < TPB; ++w) { t1 += sh1[w]; t2 += sh2[w]; }
        out[0] = t1; out[1] = t2;
    }
}

template <class TA>
void dot2(semo_ctx* c, int n, const TA* a, const TA* b, const TA* cc, const TA* d, double* out) {
    int nb = rblk(n);
    hipLaunchKernelGGL(k_dot2<TA>, dim3(nb), dim3(TPB), 0, c->stream, n, a, b, cc, d, c->red.p);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(TPB), 0, c->stream, nb, c->red.p, out);
}

// ------------------------------------------------------------ vector ops --
// scalar slots in c->scal
// (dot2 writes a pair: slot and slot+1)
enum { S_RZ = 0, S_ZRO = 1, S_RR = 2, S_PQ = 4, S_X = 6, S_RZOLD = 8, S_ALPHA = 10, S_BETA = 11,
       S_OMEGA = 12, S_TT = 14, S_TS = 15, S_RHO = 16, S_RHOOLD = 18, S_AV = 20 };

__global__ void k_cg_scalars(double* s, int what) {
    if (threadIdx.x != 0) return;
    if (what == 0) s[S_ALPHA] = s[S_RZ] / s[S_PQ];                   // alpha = rz / pAp
    else if (what == 1) {                                            // Polak-Ribiere beta
        double rz_old = s[S_RZOLD];
        s[S_BETA] = (s[S_RZ] - s[S_ZRO]) / rz_old;
    }
}

// x += alpha p ; r -= alpha q ; rold = r (before update)
__global__ void __launch_bounds__(TPB) k_cg_update(int n, const double* __restrict__ s, double* __restrict__ x,
                                                   double* __restrict__ r, double* __restrict__ rold,
                                                   const double* __restrict__ p, const double* __restrict__ q) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a = s[S_ALPHA];
    x[i] += a * p[i];
    double ri = r[i];
    rold[i] = ri;
    r[i] = ri - a * q[i];
}

__global__ void __launch_bounds__(TPB) k_xpby(int n, const double* __restrict__ s, int sidx,
                                              const double* __restrict__ z, double* __restrict__ p) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    p[i] = z[i] + s[sidx] * p[i];
}

template <class TA, class TB>
__global__ void __launch_bounds__(TPB) k_cast(int n, const TA* __restrict__ a, TB* __restrict__ b) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (TB)a[i];
}

template <class T>
__global__ void __launch_bounds__(TPB) k_scale(int n, const T* __restrict__ d, const T* __restrict__ a,
                                               T* __restrict__ b) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = d[i] * a[i];
}

__global__ void __launch_bounds__(TPB) k_zero64(int n, double* __restrict__ a) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

// BiCGStab scalar updates on the device
__global__ void k_bicg_scalars(double* s, int what) {
    if (threadIdx.x != 0) return;
    if (what == 0) {            // beta = (rho/rho_old)*(alpha/omega); first iteration: beta = 0
        double rho = s[S_RHO];
        s[S_BETA] = s[S_RHOOLD] == 0.0 ? 0.0 : (rho / s[S_RHOOLD]) * (s[S_ALPHA] / s[S_OMEGA]);
    } else if (what == 1) {     // alpha = rho / (rhat . v)
        s[S_ALPHA] = s[S_RHO] / s[S_AV];
    } else if (what == 2) {     // omega = (t.s)/(t.t)
        s[S_OMEGA] = s[S_TT] != 0.0 ? s[S_TS] / s[S_TT] : 0.0;
        s[S_RHOOLD] = s[S_RHO];
    }
}

// p = r + beta (p - omega v)
__global__ void __launch_bounds__(TPB) k_bicg_p(int n, const double* __restrict__ s, const double* __restrict__ r,
                                                const double* __restrict__ v, double* __restrict__ p) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = r[i] + s[S_BETA] * (p[i] - s[S_OMEGA] * v[i]);
}
// sv = r - alpha v
__global__ void __launch_bounds__(TPB) k_bicg_s(int n, const double* __restrict__ s, const double* __restrict__ r,
                                                const double* __restrict__ v, double* __restrict__ sv) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) sv[i] = r[i] - s[S_ALPHA] * v[i];
}
// x += alpha ph + omega sh ; r = sv - omega t
__global__ void __launch_bounds__(TPB) k_bicg_x(int n, const double* __restrict__ s, const double* __restrict__ ph,
                                                const double* __restrict__ sh, const double* __restrict__ sv,
                                                const double* __restrict__ t, double* __restrict__ x,
                                                double* __restrict__ r) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] += s[S_ALPHA] * ph[i] + s[S_OMEGA] * sh[i];
    r[i] = sv[i] - s[S_OMEGA] * t[i];
}

// ---------------------------------------------------------- Chebyshev ----
// first step from x = 0: r = D^-1 b, d = r/theta, x = d
template <class T>
__global__ void __launch_bounds__(TPB) k_cheb_init(int n, const T* __restrict__ dinv, const T* __restrict__ b,
                                                   T theta_inv, T* __restrict__ r, T* __restrict__ d,
                                                   T* __restrict__ x) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T ri = dinv[i] * b[i];
    T di = ri * theta_inv;
    r[i] = ri; d[i] = di; x[i] = di;
}
// first step with a nonzero x: r = D^-1 (b - A x) (t holds A x), d = r/theta, x += d
template <class T>
__global__ void __launch_bounds__(TPB) k_cheb_init_x(int n, const T* __restrict__ dinv, const T* __restrict__ b,
                                                     const T* __restrict__ t, T theta_inv, T* __restrict__ r,
                                                     T* __restrict__ d, T* __restrict__ x) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T ri = dinv[i] * (b[i] - t[i]);
    T di = ri * theta_inv;
    r[i] = ri; d[i] = di; x[i] += di;
}
// fused: t = A d (sliced ELL) ; r -= D^-1 t ; d' = c1 d + c2 r ; x += d'
template <class T>
__global__ void __launch_bounds__(TPB) k_cheb_step(int n, const int* __restrict__ sl_ptr, const int* __restrict__ col,
                                                   const T* __restrict__ val, const T* __restrict__ diag,
                                                   const T* __restrict__ dinv, const T* __restrict__ din,
                                                   T* __restrict__ dout, T* __restrict__ r, T* __restrict__ x,
                                                   T c1, T c2) {
    int row = blockIdx.x * blockDim.x + threadIdx.x;
    int s = row >> 6, lane = row & 63;
    if (s >= ((n + 63) >> 6)) return;
    T acc = 0;
    for (int e = sl_ptr[s] + lane; e < sl_ptr[s + 1]; e += 64) acc += val[e] * din[col[e]];
    if (row >= n) return;
    acc += diag[row] * din[row];
    T ri = r[row] - dinv[row] * acc;
    T di = c1 * din[row] + c2 * ri;
    r[row] = ri; dout[row] = di; x[row] += di;
}

// coarse dense solve: y = Ainv x (FP64), x/y of type T
template <class T>
__global__ void __launch_bounds__(TPB) k_dense_gemv(int n, const double* __restrict__ M, const T* __restrict__ x,
                                                    T* __restrict__ y) {
    int row = blockIdx.x * (TPB / 64) + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (row >= n) return;
    double acc = 0.0;
    for (int j = lane; j < n; j += 64) acc += M[(size_t)row * n + j] * (double)x[j];
    acc = wave_sum(acc);
    if (lane == 0) y[row] = (T)acc;
}

__global__ void __launch_bounds__(TPB) k_neq(size_t n, const double* __restrict__ a, const double* __restrict__ b,
                                             int* __restrict__ flag) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && a[i] != b[i]) *flag = 1;
}

// --------------------------------------------------------------- V-cycle --
template <class T> const DevSell<T>& opA(const AmgLevel& L);
template <> const DevSell<double>& opA<double>(const AmgLevel& L) { return L.A64; }
template <> const DevSell<float>& opA<float>(const AmgLevel& L) { return L.A32; }
template <class T> const DevSell<T>& opP(const AmgLevel& L);
template <> const DevSell<double>& opP<double>(const AmgLevel& L) { return L.P64; }
template <> const DevSell<float>& opP<float>(const AmgLevel& L) { return L.P32; }
template <class T> const DevSell<T>& opR(const AmgLevel& L);
template <> const DevSell<double>& opR<double>(const AmgLevel& L) { return L.R64; }
template <> const DevSell<float>& opR<float>(const AmgLevel& L) { return L.R32; }
template <class T> T* vx(AmgLevel& L);
template <> double* vx<double>(AmgLevel& L) { return L.x64.p; }
template <> float* vx<float>(AmgLevel& L) { return L.x32.p; }
template <class T> T* vb(AmgLevel& L);
template <> double* vb<double>(AmgLevel& L) { return L.b64.p; }
template <> float* vb<float>(AmgLevel& L) { return L.b32.p; }
template <class T> T* vr(AmgLevel& L);
template <> double* vr<double>(AmgLevel& L) { return L.r64.p; }
template <> float* vr<float>(AmgLevel& L) { return L.r32.p; }
template <class T> T* vt(AmgLevel& L);
template <> double* vt<double>(AmgLevel& L) { return L.t64.p; }
template <> float* vt<float>(AmgLevel& L) { return L.t32.p; }
template <class T> T* vu(AmgLevel& L);
template <> double* vu<double>(AmgLevel& L) { return L.u64.p; }
template <> float* vu<float>(AmgLevel& L) { return L.u32.p; }
template <class T> const T* vdinv(const AmgLevel& L);
template <> const double* vdinv<double>(const AmgLevel& L) { return L.dinv64.p; }
template <> const float* vdinv<float>(const AmgLevel& L) { return L.dinv32.p; }

// Chebyshev polynomial smoother in D^-1 A on [lmax/30, 1.1 lmax].
// x_zero: start from x = 0 (saves one SpMV).  dA/dB: double-buffered search
// direction, r: preconditioned residual.
template <class T>
void cheb_smooth(semo_ctx* c, AmgLevel& L, T* x, const T* b, T* dA, T* dB, T* r, int degree, bool x_zero,
                 const char* tag) {
    const DevSell<T>& A = opA<T>(L);
    int n = L.n;
    double hi = 1.1 * L.lmax, lo = hi / 30.0;
    double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo);
    double sigma = theta / delta, rho = 1.0 / sigma;
    if (x_zero) {
        hipLaunchKernelGGL(k_cheb_init<T>, dim3(nblk(n)), dim3(TPB), 0, c->stream, n, vdinv<T>(L), b,
                           (T)(1.0 / theta), r, dA, x);
    } else {
        spmv<T, T, T, 0>(c, A, x, (const T*)nullptr, dB);
        hipLaunchKernelGGL(k_cheb_init_x<T>, dim3(nblk(n)), dim3(TPB), 0, c->stream, n, vdinv<T>(L), b, dB,
                           (T)(1.0 / theta), r, dA, x);
    }
    for (int k = 1; k < degree; ++k) {
        double rho_new = 1.0 / (2.0 * sigma - rho);
        double c1 = rho_new * rho, c2 = 2.0 * rho_new / delta;
        rho = rho_new;
        hipEvent_t ev{};
        if (c->prof.on && tag) c->prof.begin(c->stream, ev);
        hipLaunchKernelGGL(k_cheb_step<T>, dim3(nblk(A.nslices * 64)), dim3(TPB), 0, c->stream, n, A.sl_ptr.p,
                           A.col.p, A.val.p, A.diag.p, vdinv<T>(L), dA, dB, r, x, (T)c1, (T)c2);
        if (c->prof.on && tag)
            c->prof.end(c->stream, tag, ev,
                        (double)A.nnz * (sizeof(T) + 4) + (double)n * sizeof(T) * 8);
        std::swap(dA, dB);
    }
}

template <class T>
void vcycle(semo_ctx* c, Amg& amg, int l, int degree) {
    AmgLevel& L = amg.lv[l];
    int n = L.n;
    T* x = vx<T>(L); T* b = vb<T>(L); T* r = vr<T>(L); T* t = vt<T>(L); T* u = vu<T>(L);
    if (l + 1 == (int)amg.lv.size()) {     // coarsest: dense direct solve in FP64
        hipLaunchKernelGGL(k_dense_gemv<T>, dim3((n + 3) / 4), dim3(TPB), 0, c->stream, n, amg.coarse_inv.p, b, x);
        return;
    }
    AmgLevel& C = amg.lv[l + 1];
    const char* tag = l == 0 ? (sizeof(T) == 8 ? "amg_cheb_fp64_L0" : "amg_cheb_fp32_L0") : nullptr;
    cheb_smooth<T>(c, L, x, b, t, u, r, degree, true, tag);
    spmv<T, T, T, 1>(c, opA<T>(L), x, b, t, true, l == 0 ? (sizeof(T) == 8 ? "amg_resid_fp64_L0" : "amg_resid_fp32_L0") : nullptr);
    spmv<T, T, T, 0>(c, opR<T>(L), t, (const T*)nullptr, vb<T>(C), false);
    vcycle<T>(c, amg, l + 1, degree);
    spmv<T, T, T, 2>(c, opP<T>(L), vx<T>(C), (const T*)nullptr, x, false);
    cheb_smooth<T>(c, L, x, b, t, u, r, degree, false, tag);
}

template <class T>
void precondition(semo_ctx* c, Amg& amg, const double* r64, double* z64) {
    AmgLevel& L0 = amg.lv[0];
    int n = L0.n;
    int deg = c->sp.cheb_degree > 0 ? c->sp.cheb_degree : 2;
    if (sizeof(T) == 8) {
        hipLaunchKernelGGL((k_cast<double, double>), dim3(nblk(n)), dim3(TPB), 0, c->stream, n, r64, L0.b64.p);
        vcycle<double>(c, amg, 0, deg);
        hipLaunchKernelGGL((k_cast<double, double>), dim3(nblk(n)), dim3(TPB), 0, c->stream, n, L0.x64.p, z64);
    } else {
        hipLaunchKernelGGL((k_cast<double, float>), dim3(nblk(n)), dim3(TPB), 0, c->stream, n, r64, L0.b32.p);
        vcycle<float>(c, amg, 0, deg);
        hipLaunchKernelGGL((k_cast<float, double>), dim3(nblk(n)), dim3(TPB), 0, c->stream, n, L0.x32.p, z64);
    }
}

double host_scalar(semo_ctx* c, int idx) {
    double v;
    SEMO_HIP(hipMemcpyAsync(&v, c->scal.p + idx, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SEMO_HIP(hipStreamSynchronize(c->stream));
    return v;
}

// flexible (Polak-Ribiere) PCG with the AMG preconditioner.  x0 = 0,
// stop at ||r|| <= tol ||b||.
int fpcg(semo_ctx* c, Amg& amg, const DevSell<double>& A, const double* b, double* x, double tol, int maxit,
         bool mixed, double* relres) {
    int n = A.n;
    size_t ln = (size_t)n + 64;
    c->kv64.ensure(8 * ln);
    double* r = c->kv64.p; double* z = r + ln; double* p = z + ln; double* q = p + ln; double* rold = q + ln;
    double* s = c->scal.p;
    hipLaunchKernelGGL(k_zero64, dim3(nblk(n)), dim3(TPB), 0, c->stream, n, x);
    hipLaunchKernelGGL((k_cast<double, double>), dim3(nblk(n)), dim3(TPB), 0, c->stream, n, b, r);
    dot2<double>(c, n, b, b, (const double*)nullptr, (const double*)nullptr, s + S_X);
    double bb = host_scalar(c, S_X);
    if (bb == 0.0) { *relres = 0.0; return 0; }
    double nb = std::sqrt(bb);
    if (mixed) precondition<float>(c, amg, r, z); else precondition<double>(c, amg, r, z);
    hipLaunchKernelGGL((k_cast<double, double>), dim3(nblk(n)), dim3(TPB), 0, c->stream, n, z, p);
    dot2<double>(c, n, r, z, (const double*)nullptr, (const double*)nullptr, s + S_RZ);
    int it = 0;
    double rr = bb;
    for (it = 1; it <= maxit; ++it) {
        spmv<double, double, double, 0>(c, A, p, (const double*)nullptr, q, true, "spmv_fp64");
        dot2<double>(c, n, p, q, (const double*)nullptr, (const double*)nullptr, s + S_PQ);
        hipLaunchKernelGGL(k_cg_scalars, dim3(1), dim3(64), 0, c->stream, s, 0);
        hipLaunchKernelGGL(k_cg_update, dim3(nblk(n)), dim3(TPB), 0, c->stream, n, s, x, r, rold, p, q);
        dot2<double>(c, n, r, r, (const double*)nullptr, (const double*)nullptr, s + S_RR);
        rr = host_scalar(c, S_RR);
        if (std::sqrt(rr) <= tol * nb) break;
        if (mixed) precondition<float>(c, amg, r, z); else precondition<double>(c, amg, r, z);
        // rz_old <- rz ; rz = r.z ; zro = z.rold
        SEMO_HIP(hipMemcpyAsync(s + S_RZOLD, s + S_RZ, sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        dot2<double>(c, n, r, z, z, rold, s + S_RZ);
        hipLaunchKernelGGL(k_cg_scalars, dim3(1), dim3(64), 0, c->stream, s, 1);
        hipLaunchKernelGGL(k_xpby, dim3(nblk(n)), dim3(TPB), 0, c->stream, n, s, (int)S_BETA, z, p);
    }
    // true residual in FP64
    spmv<double, double, double, 1>(c, A, x, b, r);
    dot2<double>(c, n, r, r, (const double*)nullptr, (const double*)nullptr, s + S_X);
    *relres = std::sqrt(host_scalar(c, S_X)) / nb;
    return std::min(it, maxit);
}

// right-preconditioned (Jacobi) BiCGStab, x0 = 0
int bicgstab(semo_ctx* c, const DevSell<double>& A, const double* dinv, const double* b, double* x, double tol,
             int maxit, double* relres) {
    int n = A.n;
    size_t ln = (size_t)n + 64;
    c->kv64.ensure(8 * ln);
    double* r = c->kv64.p; double* rh = r + ln; double* p = rh + ln; double* v = p + ln; double* sv = v + ln;
    double* t = sv + ln; double* ph = t + ln; double* sh = ph + ln;
    double* s = c->scal.p;
    hipLaunchKernelGGL(k_zero64, dim3(nblk(n)), dim3(TPB), 0, c->stream, n, x);
    hipLaunchKernelGGL((k_cast<double, double>), dim3(nblk(n)), dim3(TPB), 0, c->stream, n, b, r);
    hipLaunchKernelGGL((k_cast<double, double>), dim3(nblk(n)), dim3(TPB), 0, c->stream, n, b, rh);
    hipLaunchKernelGGL(k_zero64, dim3(nblk(n)), dim3(TPB), 0, c->stream, n, p);
    hipLaunchKernelGGL(k_zero64, dim3(nblk(n)), dim3(TPB), 0, c->stream, n, v);
    dot2<double>(c, n, b, b, (const double*)nullptr, (const double*)nullptr, s + S_X);
    double bb = host_scalar(c, S_X);
    if (bb == 0.0) { *relres = 0.0; return 0; }
    double nb = std::sqrt(bb);
    double zero4[4] = {0, 0, 1, 1};
    // S_RHOOLD = 0 marks the first iteration; alpha = omega = 1
    SEMO_HIP(hipMemcpyAsync(s + S_RHOOLD, &zero4[0], sizeof(double), hipMemcpyHostToDevice, c->stream));
    SEMO_HIP(hipMemcpyAsync(s + S_ALPHA, &zero4[2], sizeof(double), hipMemcpyHostToDevice, c->stream));
    SEMO_HIP(hipMemcpyAsync(s + S_OMEGA, &zero4[3], sizeof(double), hipMemcpyHostToDevice, c->stream));
    int it;
    for (it = 1; it <= maxit; ++it) {
        dot2<double>(c, n, rh, r, (const double*)nullptr, (const double*)nullptr, s + S_RHO);
        hipLaunchKernelGGL(k_bicg_scalars, dim3(1), dim3(64), 0, c->stream, s, 0);
        hipLaunchKernelGGL(k_bicg_p, dim3(nblk(n)), dim3(TPB), 0, c->stream, n, s, r, v, p);
        hipLaunchKernelGGL(k_scale<double>, dim3(nblk(n)), dim3(TPB), 0, c->stream, n, dinv, p, ph);
        spmv<double, double, double, 0>(c, A, ph, (const double*)nullptr, v, true, "spmv_fp64");
        dot2<double>(c, n, rh, v, (const double*)nullptr, (const double*)nullptr, s + S_AV);
        hipLaunchKernelGGL(k_bicg_scalars, dim3(1), dim3(64), 0, c->stream, s, 1);
        hipLaunchKernelGGL(k_bicg_s, dim3(nblk(n)), dim3(TPB), 0, c->stream, n, s, r, v, sv);
        hipLaunchKernelGGL(k_scale<double>, dim3(nblk(n)), dim3(TPB), 0, c->stream, n, dinv, sv, sh);
        spmv<double, double, double, 0>(c, A, sh, (const double*)nullptr, t, true, "spmv_fp64");
        dot2<double>(c, n, t, t, t, sv, s + S_TT);      // S_TT = t.t, S_TS = t.s
        hipLaunchKernelGGL(k_bicg_scalars, dim3(1), dim3(64), 0, c->stream, s, 2);
        hipLaunchKernelGGL(k_bicg_x, dim3(nblk(n)), dim3(TPB), 0, c->stream, n, s, ph, sh, sv, t, x, r);
        dot2<double>(c, n, r, r, (const double*)nullptr, (const double*)nullptr, s + S_RR);
        double rr = host_scalar(c, S_RR);
        if (!(std::sqrt(rr) > tol * nb)) break;
    }
    spmv<double, double, double, 1>(c, A, x, b, r);
    dot2<double>(c, n, r, r, (const double*)nullptr, (const double*)nullptr, s + S_X);
    *relres = std::sqrt(host_scalar(c, S_X)) / nb;
    return std::min(it, maxit);
}

__global__ void __launch_bounds__(TPB) k_dinv(int n, const double* __restrict__ d, double* __restrict__ o) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o[i] = 1.0 / d[i];
}

}  // namespace

// ---------------------------------------------------------------- uploads --
template <class T>
void sell_upload(DevSell<T>& M, const SellPattern& p, const std::vector<double>& vals,
                 const std::vector<double>& diag, hipStream_t s) {
    M.n = p.n; M.ncols = p.ncols; M.nslices = p.nslices;
    M.nnz = p.nnz_offdiag + (int64_t)diag.size();
    M.sl_ptr.upload(p.sl_ptr, s);
    M.col.upload(p.col, s);
    std::vector<T> v(vals.begin(), vals.end());
    M.val.upload(v, s);
    if (!diag.empty()) {
        std::vector<T> d(diag.begin(), diag.end());
        M.diag.upload(d, s);
    } else M.diag.release();
    SEMO_HIP(hipStreamSynchronize(s));
}
template void sell_upload<double>(DevSell<double>&, const SellPattern&, const std::vector<double>&,
                                  const std::vector<double>&, hipStream_t);
template void sell_upload<float>(DevSell<float>&, const SellPattern&, const std::vector<double>&,
                                 const std::vector<double>&, hipStream_t);

void spmv64(semo_ctx* c, const DevSell<double>& A, const double* x, double* y) {
    spmv<double, double, double, 0>(c, A, x, (const double*)nullptr, y);
}

// ----------------------------------------------------- equation solves ----
// pressure: hierarchy (re)built when the operator's values changed
void solve_pressure(semo_ctx* c, const double* b, double* x, semo_solve_stats* st) {
    Amg& amg = c->amg_p;
    bool rebuild = !amg.built;
    if (amg.built) {
        // device-side comparison with the values the hierarchy was built from
        DBuf<int>& flag = c->d_flag1;
        flag.ensure(1); flag.zero(c->stream);
        size_t nv = c->Amat.val.n;
        hipLaunchKernelGGL(k_neq, dim3((unsigned)((nv + TPB - 1) / TPB)), dim3(TPB), 0, c->stream, nv,
                           c->Amat.val.p, c->amg_key_val.p, flag.p);
        hipLaunchKernelGGL(k_neq, dim3((unsigned)((c->nc + TPB - 1) / TPB)), dim3(TPB), 0, c->stream,
                           (size_t)c->nc, c->Amat.diag.p, c->amg_key_diag.p, flag.p);
        int f = 0;
        SEMO_HIP(hipMemcpyAsync(&f, flag.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        SEMO_HIP(hipStreamSynchronize(c->stream));
        rebuild = f != 0;
    }
    if (rebuild) {
        amg.built = false;
        amg_build(c, amg, c->Amat);
        c->amg_key_val.alloc(c->Amat.val.n);
        c->amg_key_diag.alloc(c->nc);
        SEMO_HIP(hipMemcpyAsync(c->amg_key_val.p, c->Amat.val.p, c->Amat.val.bytes(), hipMemcpyDeviceToDevice, c->stream));
        SEMO_HIP(hipMemcpyAsync(c->amg_key_diag.p, c->Amat.diag.p, c->Amat.diag.bytes(), hipMemcpyDeviceToDevice, c->stream));
    }
    double rr = 0;
    int it = fpcg(c, amg, c->Amat, b, x, c->sp.tol_p, c->sp.maxit_p, c->sp.precision == SEMO_PREC_MIXED, &rr);
    if (st) {
        st->iters[0] = it; st->rel_res[0] = rr; st->n_levels = (int)amg.lv.size(); st->setup_ms = amg.setup_ms;
    }
    SEMO_CHECK(rr <= std::max(c->sp.tol_p, 1e-15) * 10.0 || it < c->sp.maxit_p, SEMO_ENOCONV,
               "pressure solve did not converge");
}

void solve_momentum(semo_ctx* c, int nrhs, semo_solve_stats* st) {
    int n = c->nc;
    DBuf<double>& dinv = c->dinv_mom;
    dinv.alloc(n + 64);
    hipLaunchKernelGGL(k_dinv, dim3(nblk(n)), dim3(TPB), 0, c->stream, n, c->Amat.diag.p, dinv.p);
    for (int k = 0; k < nrhs; ++k) {
        double rr = 0;
        int it = bicgstab(c, c->Amat, dinv.p, c->Bvec.p + (size_t)k * n, c->Xvec.p + (size_t)k * n, c->sp.tol_u,
                          c->sp.maxit_u, &rr);
        if (st) { st->iters[k] = it; st->rel_res[k] = rr; }
    }
}

// standalone COO system (seam 2)
void solve_coo(semo_ctx* c, const char* eq, int n, int64_t nnz, const int* rows, const int* cols,
               const double* vals, int nrhs, const double* b, double* x, semo_solve_stats* st) {
    SEMO_CHECK(n > 0 && nnz >= n && nrhs >= 1 && nrhs <= 4, SEMO_EINVAL, "solve_coo sizes");
    for (int64_t k = 0; k < nnz; ++k)
        SEMO_CHECK(rows[k] >= 0 && rows[k] < n && cols[k] >= 0 && cols[k] < n, SEMO_EINVAL, "COO index out of range");
    // CSR with summed duplicates, sorted columns
    std::vector<int> ptr(n + 1, 0);
    for (int64_t k = 0; k < nnz; ++k) ptr[rows[k] + 1]++;
    for (int i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
    std::vector<std::pair<int, double>> ent(nnz);
    {
        std::vector<int> pos(ptr.begin(), ptr.end() - 1);
        for (int64_t k = 0; k < nnz; ++k) ent[pos[rows[k]]++] = {cols[k], vals[k]};
    }
    std::vector<int> cptr(n + 1, 0), ccol; std::vector<double> cval;
    for (int i = 0; i < n; ++i) {
        std::sort(ent.begin() + ptr[i], ent.begin() + ptr[i + 1],
                  [](const auto& a, const auto& b2) { return a.first < b2.first; });
        for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
            if (!ccol.empty() && (int)ccol.size() > cptr[i] && ccol.back() == ent[k].first) { cval.back() += ent[k].second; continue; }
            ccol.push_back(ent[k].first); cval.push_back(ent[k].second);
        }
        cptr[i + 1] = (int)ccol.size();
    }
    // device operator
    std::vector<int> optr(n + 1, 0), ocol; std::vector<double> oval, diag(n, 0.0);
    for (int i = 0; i < n; ++i) {
        for (int k = cptr[i]; k < cptr[i + 1]; ++k) {
            if (ccol[k] == i) diag[i] = cval[k];
            else { ocol.push_back(ccol[k]); oval.push_back(cval[k]); }
        }
        optr[i + 1] = (int)ocol.size();
    }
    for (int i = 0; i < n; ++i) SEMO_CHECK(diag[i] != 0.0, SEMO_EZERO_PIVOT, "zero diagonal");
    std::vector<int> pos;
    SellPattern p = sell_from_csr(n, n, optr, ocol, &pos);
    std::vector<double> pv(p.col.size(), 0.0);
    for (size_t k = 0; k < ocol.size(); ++k) pv[pos[k]] = oval[k];
    DevSell<double> A;
    sell_upload<double>(A, p, pv, diag, c->stream);
    DBuf<double> db, dx;
    db.upload(b, (size_t)n * nrhs, c->stream);
    dx.alloc((size_t)n * nrhs);
    if (c->red.n < 2 * RB + 64) c->red.alloc(2 * RB + 64);
    if (c->scal.n < 64) c->scal.alloc(64);
    std::string e(eq ? eq : "");
    if (e == "pressure") {
        Amg amg;
        amg_build_csr(c, amg, n, cptr, ccol, cval);
        for (int k = 0; k < nrhs; ++k) {
            double rr = 0;
            int it = fpcg(c, amg, A, db.p + (size_t)k * n, dx.p + (size_t)k * n, c->sp.tol_p, c->sp.maxit_p,
                          c->sp.precision == SEMO_PREC_MIXED, &rr);
            if (st) { st->iters[k] = it; st->rel_res[k] = rr; st->n_levels = (int)amg.lv.size(); st->setup_ms = amg.setup_ms; }
        }
    } else {
        DBuf<double> dinv; dinv.alloc(n + 64);
        hipLaunchKernelGGL(k_dinv, dim3(nblk(n)), dim3(TPB), 0, c->stream, n, A.diag.p, dinv.p);
        for (int k = 0; k < nrhs; ++k) {
            double rr = 0;
            int it = bicgstab(c, A, dinv.p, db.p + (size_t)k * n, dx.p + (size_t)k * n, c->sp.tol_u, c->sp.maxit_u, &rr);
            if (st) { st->iters[k] = it; st->rel_res[k] = rr; }
        }
    }
    dx.download(x, (size_t)n * nrhs, c->stream);
    SEMO_HIP(hipStreamSynchronize(c->stream));
}

}  // namespace semo
