// Linear algebra and Krylov solvers on the device.
//
// * row kernels over two operator formats: sliced ELL (one lane per row,
//   coalesced val/col loads; short regular rows such as the fine-mesh
//   operator and the prolongation) and CSR with 4..64 lanes per row (long or
//   ragged rows: AMG restriction and Galerkin coarse operators);
// * fused Chebyshev steps (SpMV + residual/direction/iterate updates in one
//   pass) and a fused post-smoother start (A x, residual, first direction);
// * deterministic reductions fused into the kernels that produce their
//   operands: per-block partials in a fixed order, then one single-block
//   kernel sums them in a fixed order and computes the Krylov scalars, so
//   they never leave the GPU except for the convergence test.  (A
//   last-block-done scheme would need a device-scope release fence in every
//   block, i.e. an L2 write-back per block on this chip: measured ~0.3 ms per
//   launch at 2M rows, against ~5 us for the extra launch.)
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
constexpr int RB = 1024;   // grid of the grid-stride reductions
inline int nblk(long n) { return (int)((n + TPB - 1) / TPB); }
inline int rblk(long n) { return (int)std::min<long>(RB, std::max<long>(1, (n + TPB - 1) / TPB)); }

// ------------------------------------------------------------- reductions --
template <class T>
__device__ inline T wave_sum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Sums the two per-thread values over the block in a fixed order and stores
// them as this block's partials part[2b], part[2b+1].  Every thread of the
// block must call it.
__device__ inline void block_partials(double (&v)[2], double* __restrict__ part) {
    __shared__ double sh[2][TPB / 64];
    v[0] = wave_sum(v[0]); v[1] = wave_sum(v[1]);
    if ((threadIdx.x & 63) == 0) { sh[0][threadIdx.x >> 6] = v[0]; sh[1][threadIdx.x >> 6] = v[1]; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t0 = 0.0, t1 = 0.0;
        for (int w = 0; w < TPB / 64; ++w) { t0 += sh[0][w]; t1 += sh[1][w]; }
        part[2 * blockIdx.x] = t0; part[2 * blockIdx.x + 1] = t1;
    }
}

// Krylov scalars in device memory (slots of c->scal)
enum { S_RZ = 0, S_ZRO = 1, S_RR = 2, S_PQ = 4, S_X = 6, S_ALPHA = 10, S_BETA = 11, S_OMEGA = 12,
       S_TT = 14, S_TS = 15, S_RHO = 16, S_AV = 20 };

// what the last block does with the totals d[0], d[1]
enum { W_STORE = 0,      // s[slot] = d0, s[slot+1] = d1
       W_FCG_ALPHA,      // d0 = p.Ap:  alpha = rz / p.Ap
       W_FCG_BETA,       // d0 = r.z, d1 = z.r_old: Polak-Ribiere beta, rz = d0
       W_BICG_INIT,      // d0 = b.b: rho = rr = b.b, beta = 0, alpha = omega = 1
       W_BICG_ALPHA,     // d0 = rh.v: alpha = rho / rh.v
       W_BICG_OMEGA,     // d0 = t.t, d1 = t.s: omega = t.s / t.t
       W_BICG_RHO };     // d0 = r.r, d1 = rh.r: beta = (rho'/rho)(alpha/omega), rho = rho'

__device__ void finish_scalars(int what, int slot, const double (&d)[2], double* s) {
    switch (what) {
    case W_STORE: s[slot] = d[0]; s[slot + 1] = d[1]; break;
    // zero denominators (an exactly solved system) give zero steps, so that
    // an iteration issued past convergence leaves x unchanged
    case W_FCG_ALPHA: s[S_PQ] = d[0]; s[S_ALPHA] = d[0] != 0.0 ? s[S_RZ] / d[0] : 0.0; break;
    case W_FCG_BETA:
        s[S_BETA] = s[S_RZ] != 0.0 ? (d[0] - d[1]) / s[S_RZ] : 0.0;
        s[S_RZ] = d[0]; s[S_ZRO] = d[1];
        break;
    case W_BICG_INIT:
        s[S_X] = d[0]; s[S_RR] = d[0]; s[S_RHO] = d[0];
        s[S_BETA] = 0.0; s[S_ALPHA] = 1.0; s[S_OMEGA] = 1.0;
        break;
    case W_BICG_ALPHA: s[S_AV] = d[0]; s[S_ALPHA] = d[0] != 0.0 ? s[S_RHO] / d[0] : 0.0; break;
    case W_BICG_OMEGA: s[S_TT] = d[0]; s[S_TS] = d[1]; s[S_OMEGA] = d[0] != 0.0 ? d[1] / d[0] : 0.0; break;
    case W_BICG_RHO:
        s[S_RR] = d[0];
        s[S_BETA] = (s[S_RHO] != 0.0 && s[S_OMEGA] != 0.0) ? (d[1] / s[S_RHO]) * (s[S_ALPHA] / s[S_OMEGA]) : 0.0;
        s[S_RHO] = d[1];
        break;
    default: break;
    }
}

struct Red {                 // where a launch leaves its block partials
    double* part;
};

__device__ inline void reduce2(double (&v)[2], const Red& rd) { block_partials(v, rd.part); }

// ordered sum of nb partial pairs, then the scalar update `what`; one
// 1024-thread block, so each thread streams only a few partials
constexpr int FTPB = 1024;
__global__ void __launch_bounds__(FTPB) k_finish(int nb, const double* __restrict__ part, double* __restrict__ s,
                                                 int what, int slot) {
    double a0 = 0.0, a1 = 0.0;
    const double2* p2 = reinterpret_cast<const double2*>(part);
    for (int i = threadIdx.x; i < nb; i += FTPB) { double2 v = p2[i]; a0 += v.x; a1 += v.y; }
    __shared__ double sh[2][FTPB];
    sh[0][threadIdx.x] = a0; sh[1][threadIdx.x] = a1;
    __syncthreads();
    for (int w = FTPB / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double d[2] = {sh[0][0], sh[1][0]};
        finish_scalars(what, slot, d, s);
    }
}

// a.b (+ c.d) over a grid-stride range
__global__ void __launch_bounds__(TPB) k_dot(int n, const double* __restrict__ a, const double* __restrict__ b,
                                             const double* __restrict__ cc, const double* __restrict__ d, Red rd) {
    double v[2] = {0.0, 0.0};
    for (int i = xcd_block() * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        v[0] += a[i] * b[i];
        if (cc) v[1] += cc[i] * d[i];
    }
    reduce2(v, rd);
}

// ------------------------------------------------------------ row kernels --
enum { OP_Y = 0,       // y = A x
       OP_RES = 1,     // y = b - A x
       OP_ADD = 2,     // y += A x
       OP_CHEB = 3,    // Chebyshev step: x = d_k, y = d_k+1, r -= D^-1 A d_k, xo += d_k+1
       OP_CHEBX = 4,   // Chebyshev start from x != 0: r = D^-1 (b - A x), y = r c1, xo = x + y
       OP_YPREP = 6 }; // y = A x and xo = c1 D^-1 y (restriction writing the next level's b and
                       // its degree-1 smoothed start, rounded as k_cheb_init forms it)

template <class T>
struct RowArgs {
    int n;
    const int* __restrict__ ptr;    // slice offsets (sliced ELL) or row offsets (CSR)
    const int* __restrict__ col;
    const int16_t* __restrict__ dcol;   // sliced ELL: column - row (null: use col)
    const T* __restrict__ val;
    const T* __restrict__ diag;     // null: no diagonal part
    const T* __restrict__ x;        // gathered operand
    const T* __restrict__ b;
    T* __restrict__ y;
    const T* __restrict__ dinv;
    T* __restrict__ r;
    T* __restrict__ xo;
    double* __restrict__ zo;        // optional FP64 copy of the new iterate
    T c1, c2;
};

// entries per trip of the row sums: 2 pairs on sliced-ELL slices (rows of
// ~6-7 entries), 8 per lane of the CSR kernels (coarse rows of 30-60 entries
// over avg/4 lanes; measured best of 1/2/4/8 on MI355X)
constexpr int SELL_PAIRS = 2, CSR_UNR = 8;
// One lane's sum over its row of a sliced-ELL slice, entries e0, e0 + 64, ...
// below e1, in entry order: acc += val[e] * opnd(col(e)).  U entries per
// trip: their indices and values are loaded first, then their gathers are
// issued together, so the gathers' latency is paid once per trip instead of
// once per entry.  The sum is the same as the one-entry loop's, bit for bit.
// (STRIDE 64; the CSR kernels use it with STRIDE = lanes per row.)
template <int U, int STRIDE, class T, class ColF, class OpF>
__device__ inline T sell_row_sum(int e0, int e1, const T* __restrict__ val, ColF colf, OpF opnd) {
    T acc = 0;
    for (int e = e0; e < e1; e += STRIDE * U) {
        int j[U];
        T v[U], x[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            int ee = e + STRIDE * u;
            j[u] = ee < e1 ? colf(ee) : -1;
            v[u] = ee < e1 ? val[ee] : T(0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) x[u] = j[u] >= 0 ? opnd(j[u]) : T(0);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (j[u] >= 0) acc += v[u] * x[u];
    }
    return acc;
}

// One lane's sum over its row of a sliced-ELL slice (entry pairs, sell_pos):
// pair m of the row at base + 128 m + 2 lane, P pairs per trip, their values
// and columns loaded (one load each per pair) before their gathers are
// issued together.  acc += val * opnd(col) in entry order, bit for bit the
// sum of the one-entry loop.  colf2(e) gives the columns of entries e, e + 1.
template <int P, class T, class ColF2, class OpF>
__device__ inline T sell2_row_sum(int base, int end, int lane, const T* __restrict__ val, ColF2 colf2, OpF opnd) {
    using T2 = typename std::conditional<sizeof(T) == 8, double2, float2>::type;
    T acc = 0;
    for (int e = base + 2 * lane; e < end; e += 2 * SL * P) {
        int2 j[P];
        T2 v[P];
        T x0[P], x1[P];
#pragma unroll
        for (int u = 0; u < P; ++u) {
            int ee = e + 2 * SL * u;
            bool in = ee < end;
            j[u] = in ? colf2(ee) : make_int2(-1, -1);
            if (in) v[u] = *reinterpret_cast<const T2*>(val + ee);
        }
#pragma unroll
        for (int u = 0; u < P; ++u)
            if (j[u].x >= 0) { x0[u] = opnd(j[u].x); x1[u] = opnd(j[u].y); }
#pragma unroll
        for (int u = 0; u < P; ++u)
            if (j[u].x >= 0) { acc += v[u].x * x0[u]; acc += v[u].y * x1[u]; }
    }
    return acc;
}
// the two columns of entries e, e + 1 (e even) from int32 columns or from
// int16 offsets column - row
__device__ inline int2 sell2_col(const int* __restrict__ col, int e) {
    return *reinterpret_cast<const int2*>(col + e);
}
__device__ inline int2 sell2_dcol(const int16_t* __restrict__ dcol, int e, int row) {
    int w = *reinterpret_cast<const int*>(dcol + e);
    return make_int2(row + (int)(int16_t)(w & 0xffff), row + (w >> 16));
}

// G == 0: sliced ELL, one lane per row.  G > 0: CSR, G consecutive lanes per
// row, strided partial sums combined by a fixed shuffle tree (the result is
// deterministic for a given G).
// the row's own operands of the epilogue: diagonal and x (for the
// diagonal term, and x for the Chebyshev start), b, D^-1
template <class T>
struct Own {
    T d, x, b, dv;
};
template <class T, int OP>
__device__ inline Own<T> load_own(const RowArgs<T>& a, int row) {
    Own<T> o{};
    if (a.diag) o.d = a.diag[row];
    if (a.diag || OP == OP_CHEBX || OP == OP_CHEB) o.x = a.x[row];
    if constexpr (OP == OP_RES || OP == OP_CHEBX) o.b = a.b[row];
    if constexpr (OP == OP_YPREP || OP == OP_CHEB || OP == OP_CHEBX) o.dv = a.dinv[row];
    return o;
}

template <class T, int G, int OP>
__global__ void __launch_bounds__(TPB) k_row(RowArgs<T> a) {
    int gid = xcd_block() * TPB + threadIdx.x;
    int row;
    T acc = 0;
    Own<T> own{};
    if constexpr (G == 0) {
        row = gid;
        int s = row >> 6, lane = row & 63;
        if (s >= ((a.n + 63) >> 6)) return;
        // the row's own operands are requested before the gathers
        if (row < a.n) own = load_own<T, OP>(a, row);
        auto opnd = [&](int j) { return a.x[j]; };
        if (a.dcol)
            acc = sell2_row_sum<SELL_PAIRS>(a.ptr[s], a.ptr[s + 1], lane, a.val,
                                            [&](int e) { return sell2_dcol(a.dcol, e, row); }, opnd);
        else
            acc = sell2_row_sum<SELL_PAIRS>(a.ptr[s], a.ptr[s + 1], lane, a.val,
                                            [&](int e) { return sell2_col(a.col, e); }, opnd);
        if (row >= a.n) return;
    } else {
        row = gid / G;
        int sub = gid & (G - 1);
        if (row >= a.n) return;
        acc = sell_row_sum<CSR_UNR, G, T>(a.ptr[row] + sub, a.ptr[row + 1], a.val, [&](int e) { return a.col[e]; },
                                            [&](int j) { return a.x[j]; });
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) acc += __shfl_down(acc, o, G);
        if (sub) return;
        own = load_own<T, OP>(a, row);
    }
    if (a.diag) acc += own.d * own.x;
    if constexpr (OP == OP_YPREP) {
        a.y[row] = acc;
        T rj = own.dv * acc;
        a.xo[row] = rj * a.c1;
    } else if constexpr (OP == OP_Y) {
        a.y[row] = acc;
    } else if constexpr (OP == OP_RES) {
        a.y[row] = own.b - acc;
    } else if constexpr (OP == OP_ADD) {
        a.y[row] += acc;
    } else if constexpr (OP == OP_CHEB) {
        T ri = a.r[row] - own.dv * acc;
        T di = a.c1 * own.x + a.c2 * ri;
        a.r[row] = ri; a.y[row] = di;
        T xn = a.xo[row] + di;
        a.xo[row] = xn;
        if (a.zo) a.zo[row] = (double)xn;
    } else {
        // r, y and xo may be null: a degree-1 smoother reads none of r and
        // y afterwards, and at the finest level only the FP64 copy zo
        T ri = own.dv * (own.b - acc);
        T di = ri * a.c1;
        if (a.r) a.r[row] = ri;
        if (a.y) a.y[row] = di;
        T xn = own.x + di;
        if (a.xo) a.xo[row] = xn;
        if (a.zo) a.zo[row] = (double)xn;
    }
}

template <class T, int OP>
void row_launch(semo_ctx* c, const DevSell<T>& A, RowArgs<T> a, const char* tag, double bytes) {
    if (A.n == 0) return;
    a.n = A.n; a.ptr = A.sl_ptr.p; a.col = A.col.p; a.val = A.val.p; a.diag = A.diag.p;
    a.dcol = (A.lanes == 0 && A.dcol.n) ? A.dcol.p : nullptr;
    ProfScope ps(c, tag, bytes);
    long rows = A.n;
    switch (A.lanes) {
    case 0: hipLaunchKernelGGL((k_row<T, 0, OP>), dim3(nblk((long)A.nslices * 64)), dim3(TPB), 0, c->stream, a); break;
    case 4: hipLaunchKernelGGL((k_row<T, 4, OP>), dim3(nblk(rows * 4)), dim3(TPB), 0, c->stream, a); break;
    case 8: hipLaunchKernelGGL((k_row<T, 8, OP>), dim3(nblk(rows * 8)), dim3(TPB), 0, c->stream, a); break;
    case 16: hipLaunchKernelGGL((k_row<T, 16, OP>), dim3(nblk(rows * 16)), dim3(TPB), 0, c->stream, a); break;
    case 32: hipLaunchKernelGGL((k_row<T, 32, OP>), dim3(nblk(rows * 32)), dim3(TPB), 0, c->stream, a); break;
    case 64: hipLaunchKernelGGL((k_row<T, 64, OP>), dim3(nblk(rows * 64)), dim3(TPB), 0, c->stream, a); break;
    default: throw Error(SEMO_EINVAL, "operator with an unsupported lane count");
    }
}

// algorithmic bytes of one pass over A: values + columns of the real
// entries, row/slice offsets, plus `vec` vector accesses of n entries
template <class T>
double op_bytes(const DevSell<T>& A, int vec) {
    double ptrs = A.lanes ? 4.0 * (A.n + 1) : 4.0 * (A.nslices + 1);
    double idx = (A.lanes == 0 && A.dcol.n) ? 2.0 : 4.0;
    double nd = A.diag.n ? (double)A.n : 0.0;          // diagonal entries: a value, no index
    return ((double)A.nnz - nd) * (sizeof(T) + idx) + nd * sizeof(T) + ptrs + (double)A.n * sizeof(T) * vec;
}

template <class T, int OP>
void spmv(semo_ctx* c, const DevSell<T>& A, const T* x, const T* b, T* y, const char* tag) {
    RowArgs<T> a{};
    a.x = x; a.b = b; a.y = y;
    row_launch<T, OP>(c, A, a, tag, op_bytes(A, OP == OP_Y ? 2 : 3));
}

// y = A x (sliced ELL, FP64) with d0 = (a0 ? a0 : y).y and d1 = a1.y
__global__ void __launch_bounds__(TPB) k_spmv_dot(int n, const int* __restrict__ sl_ptr, const int* __restrict__ col,
                                                  const int16_t* __restrict__ dcol,
                                                  const double* __restrict__ val, const double* __restrict__ diag,
                                                  const double* __restrict__ x, double* __restrict__ y,
                                                  const double* __restrict__ a0, const double* __restrict__ a1, Red rd) {
    int row = xcd_block() * blockDim.x + threadIdx.x;
    int s = row >> 6, lane = row & 63;
    double v[2] = {0.0, 0.0};
    if (s < ((n + 63) >> 6)) {
        bool own = row < n;     // the row's own operands before the gathers
        double dr = own ? diag[row] : 0.0, xr = own ? x[row] : 0.0;
        auto opnd = [&](int j) { return x[j]; };
        double acc = dcol ? sell2_row_sum<SELL_PAIRS>(sl_ptr[s], sl_ptr[s + 1], lane, val,
                                                      [&](int e) { return sell2_dcol(dcol, e, row); }, opnd)
                          : sell2_row_sum<SELL_PAIRS>(sl_ptr[s], sl_ptr[s + 1], lane, val,
                                                      [&](int e) { return sell2_col(col, e); }, opnd);
        if (own) {
            acc += dr * xr;
            y[row] = acc;
            v[0] = (a0 ? a0[row] : acc) * acc;
            if (a1) v[1] = a1[row] * acc;
        }
    }
    reduce2(v, rd);
}

// FCG direction update fused into its product: pn = z + beta po (beta =
// s[S_BETA]), q = A pn, partials pn.q.  po and pn are distinct buffers:
// neighbours read po while the row writes pn.
__global__ void __launch_bounds__(TPB) k_spmv_xpby(int n, const int* __restrict__ sl_ptr,
                                                   const int* __restrict__ col, const int16_t* __restrict__ dcol,
                                                   const double* __restrict__ val, const double* __restrict__ diag,
                                                   const double* __restrict__ s, const double* __restrict__ z,
                                                   const double* __restrict__ po, double* __restrict__ pn,
                                                   double* __restrict__ q, Red rd) {
    int row = xcd_block() * blockDim.x + threadIdx.x;
    int sl = row >> 6, lane = row & 63;
    double v[2] = {0.0, 0.0};
    if (sl < ((n + 63) >> 6)) {
        double beta = s[S_BETA];
        // the row's own operands are requested before the gathers
        bool own = row < n;
        double zr = own ? z[row] : 0.0, por = own ? po[row] : 0.0, dr = own ? diag[row] : 0.0;
        auto opnd = [&](int j) { return z[j] + beta * po[j]; };
        double acc = dcol ? sell2_row_sum<SELL_PAIRS>(sl_ptr[sl], sl_ptr[sl + 1], lane, val,
                                                      [&](int e) { return sell2_dcol(dcol, e, row); }, opnd)
                          : sell2_row_sum<SELL_PAIRS>(sl_ptr[sl], sl_ptr[sl + 1], lane, val,
                                                      [&](int e) { return sell2_col(col, e); }, opnd);
        if (own) {
            double pr = zr + beta * por;
            pn[row] = pr;
            acc += dr * pr;
            q[row] = acc;
            v[0] = pr * acc;
        }
    }
    reduce2(v, rd);
}

// ------------------------------------------------------- Krylov vectors --
// Level-0 input of the next FP32 V-cycle, written by the kernel that
// produces the residual r (null b32: not used): b32 = (float) r and the
// degree-1 smoothed start x32 = c1 D^-1 b32, rounded as the V-cycle's own
// fused smoothing step forms them.
// (b64 / x64: the same for an FP64 V-cycle.)
struct Prep {
    float* __restrict__ b32;
    float* __restrict__ x32;
    const float* __restrict__ dinv32;
    float c1;
    double* __restrict__ b64;
    double* __restrict__ x64;
    const double* __restrict__ dinv64;
    double c1d;
};
__device__ inline void prep_store(const Prep& pp, int i, double ri) {
    if (pp.b32) {
        float bf = (float)ri;
        pp.b32[i] = bf;
        float rj = pp.dinv32[i] * bf;
        pp.x32[i] = rj * pp.c1;
    } else if (pp.b64) {
        pp.b64[i] = ri;
        double rj = pp.dinv64[i] * ri;
        pp.x64[i] = rj * pp.c1d;
    }
}

// x = 0, r = b, d0 = b.b
__global__ void __launch_bounds__(TPB) k_fcg_init(int n, const double* __restrict__ b, double* __restrict__ x,
                                                  double* __restrict__ r, Red rd, Prep pp) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (i < n) { double bi = b[i]; x[i] = 0.0; r[i] = bi; v[0] = bi * bi; prep_store(pp, i, bi); }
    reduce2(v, rd);
}

// x += alpha p ; r_old = r ; r -= alpha q ; d0 = r.r
__global__ void __launch_bounds__(TPB) k_fcg_update(int n, const double* __restrict__ s, double* __restrict__ x,
                                                    double* __restrict__ r, double* __restrict__ rold,
                                                    const double* __restrict__ p, const double* __restrict__ q,
                                                    Red rd, Prep pp) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (i < n) {
        double a = s[S_ALPHA];
        x[i] += a * p[i];
        double ri = r[i];
        rold[i] = ri;
        ri -= a * q[i];
        r[i] = ri;
        v[0] = ri * ri;
        prep_store(pp, i, ri);
    }
    reduce2(v, rd);
}

template <class TA, class TB>
__global__ void __launch_bounds__(TPB) k_cast(int n, const TA* __restrict__ a, TB* __restrict__ b) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (TB)a[i];
}

// x = 0, r = rh = b, p = v = 0, d0 = b.b
__global__ void __launch_bounds__(TPB) k_bicg_init(int n, const double* __restrict__ b, double* __restrict__ x,
                                                   double* __restrict__ r, double* __restrict__ rh,
                                                   double* __restrict__ p, double* __restrict__ v, Red rd) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    double w[2] = {0.0, 0.0};
    if (i < n) {
        double bi = b[i];
        x[i] = 0.0; r[i] = bi; rh[i] = bi; p[i] = 0.0; v[i] = 0.0;
        w[0] = bi * bi;
    }
    reduce2(w, rd);
}
// p = r + beta (p - omega v) ; ph = D^-1 p
__global__ void __launch_bounds__(TPB) k_bicg_p(int n, const double* __restrict__ s, const double* __restrict__ r,
                                                const double* __restrict__ v, const double* __restrict__ dinv,
                                                double* __restrict__ p, double* __restrict__ ph) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double pi = r[i] + s[S_BETA] * (p[i] - s[S_OMEGA] * v[i]);
    p[i] = pi; ph[i] = dinv[i] * pi;
}
// sv = r - alpha v ; sh = D^-1 sv
__global__ void __launch_bounds__(TPB) k_bicg_s(int n, const double* __restrict__ s, const double* __restrict__ r,
                                                const double* __restrict__ v, const double* __restrict__ dinv,
                                                double* __restrict__ sv, double* __restrict__ sh) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double si = r[i] - s[S_ALPHA] * v[i];
    sv[i] = si; sh[i] = dinv[i] * si;
}
// x += alpha ph + omega sh ; r = sv - omega t ; d0 = r.r, d1 = rh.r
__global__ void __launch_bounds__(TPB) k_bicg_x(int n, const double* __restrict__ s, const double* __restrict__ ph,
                                                const double* __restrict__ sh, const double* __restrict__ sv,
                                                const double* __restrict__ t, const double* __restrict__ rh,
                                                double* __restrict__ x, double* __restrict__ r, Red rd) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    double w[2] = {0.0, 0.0};
    if (i < n) {
        x[i] += s[S_ALPHA] * ph[i] + s[S_OMEGA] * sh[i];
        double ri = sv[i] - s[S_OMEGA] * t[i];
        r[i] = ri;
        w[0] = ri * ri; w[1] = rh[i] * ri;
    }
    reduce2(w, rd);
}

// ------------------------------------------ three right-hand sides at once --
// The momentum components U, V, W share one matrix: BiCGStab on the three
// systems runs in lockstep, so every matrix pass serves all three.  Vectors
// are [3][n] (component k at +k*n); the Krylov scalars of component k live
// in slots Q_* + k of the device scalar array.
enum { Q_RHO = 32, Q_ALPHA = 35, Q_OMEGA = 38, Q_BETA = 41, Q_BB = 44, Q_RR = 47, Q_MAXREL = 50 };
enum { F3_INIT = 0, F3_ALPHA, F3_OMEGA, F3_RHO, F3_STORE };

__device__ inline void block_partials6(double (&v)[6], double* __restrict__ part) {
    __shared__ double sh[6][TPB / 64];
#pragma unroll
    for (int q = 0; q < 6; ++q) v[q] = wave_sum(v[q]);
    if ((threadIdx.x & 63) == 0)
        for (int q = 0; q < 6; ++q) sh[q][threadIdx.x >> 6] = v[q];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int q = 0; q < 6; ++q) {
            double t = 0.0;
            for (int w = 0; w < TPB / 64; ++w) t += sh[q][w];
            part[6 * blockIdx.x + q] = t;
        }
}

__global__ void __launch_bounds__(1024) k_finish6(int nb, const double* __restrict__ part, double* __restrict__ s,
                                                  int what) {
    double a[6] = {0, 0, 0, 0, 0, 0};
    for (int i = threadIdx.x; i < nb; i += 1024)
        for (int q = 0; q < 6; ++q) a[q] += part[6 * i + q];
    __shared__ double sh[6][1024];
    for (int q = 0; q < 6; ++q) sh[q][threadIdx.x] = a[q];
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int q = 0; q < 6; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    double d[6];
    for (int q = 0; q < 6; ++q) d[q] = sh[q][0];
    double mx = 0.0;
    for (int k = 0; k < 3; ++k) {
        switch (what) {
        case F3_INIT:
            s[Q_BB + k] = d[k]; s[Q_RR + k] = d[k]; s[Q_RHO + k] = d[k];
            s[Q_ALPHA + k] = 1.0; s[Q_OMEGA + k] = 1.0; s[Q_BETA + k] = 0.0;
            break;
        case F3_ALPHA: s[Q_ALPHA + k] = d[k] != 0.0 ? s[Q_RHO + k] / d[k] : 0.0; break;
        case F3_OMEGA: s[Q_OMEGA + k] = d[3 + k] != 0.0 ? d[k] / d[3 + k] : 0.0; break;
        case F3_RHO: {
            s[Q_RR + k] = d[k];
            double rho = s[Q_RHO + k], om = s[Q_OMEGA + k];
            s[Q_BETA + k] = (rho != 0.0 && om != 0.0) ? (d[3 + k] / rho) * (s[Q_ALPHA + k] / om) : 0.0;
            s[Q_RHO + k] = d[3 + k];
            double bb = s[Q_BB + k];
            double rel = bb > 0.0 ? d[k] / bb : 0.0;
            mx = (rel > mx || rel != rel) ? rel : mx;
            break;
        }
        default: s[Q_RR + k] = d[k]; break;
        }
    }
    if (what == F3_RHO) s[Q_MAXREL] = mx;
}

__global__ void __launch_bounds__(TPB) k_bicg3_init(int n, const double* __restrict__ b, double* __restrict__ x,
                                                    double* __restrict__ r, double* __restrict__ rh,
                                                    double* __restrict__ p, double* __restrict__ v,
                                                    double* __restrict__ part) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    double w[6] = {0, 0, 0, 0, 0, 0};
    if (i < n)
        for (int k = 0; k < 3; ++k) {
            size_t j = (size_t)k * n + i;
            double bi = b[j];
            x[j] = 0.0; r[j] = bi; rh[j] = bi; p[j] = 0.0; v[j] = 0.0;
            w[k] = bi * bi;
        }
    block_partials6(w, part);
}
// p = r + beta (p - omega v) ; ph = D^-1 p
__global__ void __launch_bounds__(TPB) k_bicg3_p(int n, const double* __restrict__ s, const double* __restrict__ r,
                                                 const double* __restrict__ v, const double* __restrict__ dinv,
                                                 double* __restrict__ p, double* __restrict__ ph) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double di = dinv[i];
    for (int k = 0; k < 3; ++k) {
        size_t j = (size_t)k * n + i;
        double pi = r[j] + s[Q_BETA + k] * (p[j] - s[Q_OMEGA + k] * v[j]);
        p[j] = pi; ph[j] = di * pi;
    }
}
// sv = r - alpha v ; sh = D^-1 sv
__global__ void __launch_bounds__(TPB) k_bicg3_s(int n, const double* __restrict__ s, const double* __restrict__ r,
                                                 const double* __restrict__ v, const double* __restrict__ dinv,
                                                 double* __restrict__ sv, double* __restrict__ sh) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double di = dinv[i];
    for (int k = 0; k < 3; ++k) {
        size_t j = (size_t)k * n + i;
        double si = r[j] - s[Q_ALPHA + k] * v[j];
        sv[j] = si; sh[j] = di * si;
    }
}
// x += alpha ph + omega sh ; r = sv - omega t ; partials r.r (0..2), rh.r (3..5)
__global__ void __launch_bounds__(TPB) k_bicg3_x(int n, const double* __restrict__ s, const double* __restrict__ ph,
                                                 const double* __restrict__ sh, const double* __restrict__ sv,
                                                 const double* __restrict__ t, const double* __restrict__ rh,
                                                 double* __restrict__ x, double* __restrict__ r,
                                                 double* __restrict__ part) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    double w[6] = {0, 0, 0, 0, 0, 0};
    if (i < n)
        for (int k = 0; k < 3; ++k) {
            size_t j = (size_t)k * n + i;
            x[j] += s[Q_ALPHA + k] * ph[j] + s[Q_OMEGA + k] * sh[j];
            double ri = sv[j] - s[Q_OMEGA + k] * t[j];
            r[j] = ri;
            w[k] = ri * ri; w[3 + k] = rh[j] * ri;
        }
    block_partials6(w, part);
}
// y_k = A x_k on the fine sliced-ELL operator for k = 0..2, with the dots
// MODE 0: a_k.y_k (0..2);  MODE 1: y_k.a_k (0..2), y_k.y_k (3..5);
// MODE 2: y_k = a_k - A x_k and y_k.y_k (0..2)
template <int MODE>
__global__ void __launch_bounds__(TPB) k_spmv3(int n, const int* __restrict__ sl_ptr, const int* __restrict__ col,
                                               const int16_t* __restrict__ dcol, const double* __restrict__ val,
                                               const double* __restrict__ diag, const double* __restrict__ x,
                                               double* __restrict__ y, const double* __restrict__ a,
                                               double* __restrict__ part) {
    int row = xcd_block() * blockDim.x + threadIdx.x;
    int s = row >> 6, lane = row & 63;
    double w[6] = {0, 0, 0, 0, 0, 0};
    size_t N = (size_t)n;
    if (s < ((n + 63) >> 6)) {
        bool own = row < n;     // the row's own operands before the gathers
        double d = own ? diag[row] : 0.0;
        double xr0 = own ? x[row] : 0.0, xr1 = own ? x[N + row] : 0.0, xr2 = own ? x[2 * N + row] : 0.0;
        double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
        for (int e = sl_ptr[s] + 2 * lane; e < sl_ptr[s + 1]; e += 2 * SL) {     // entry pairs
            int2 j = dcol ? sell2_dcol(dcol, e, row) : sell2_col(col, e);
            double2 v = *reinterpret_cast<const double2*>(val + e);
            acc0 += v.x * x[j.x]; acc1 += v.x * x[N + j.x]; acc2 += v.x * x[2 * N + j.x];
            acc0 += v.y * x[j.y]; acc1 += v.y * x[N + j.y]; acc2 += v.y * x[2 * N + j.y];
        }
        if (own) {
            double acc[3] = {acc0 + d * xr0, acc1 + d * xr1, acc2 + d * xr2};
            for (int k = 0; k < 3; ++k) {
                size_t jj = (size_t)k * N + row;
                double yk = MODE == 2 ? a[jj] - acc[k] : acc[k];
                y[jj] = yk;
                if (MODE == 0) w[k] = a[jj] * yk;
                else if (MODE == 1) { w[k] = yk * a[jj]; w[3 + k] = yk * yk; }
                else w[k] = yk * yk;
            }
        }
    }
    block_partials6(w, part);
}

// ---------------------------------------------------------- Chebyshev ----
// start from x = 0: r = D^-1 b, d = r/theta, x = d.  bsrc (optional): FP64
// right-hand side, converted into b on the way.  zo (optional): FP64 copy of x.
template <class T>
__global__ void __launch_bounds__(TPB) k_cheb_init(int n, const T* __restrict__ dinv, T* __restrict__ b,
                                                   const double* __restrict__ bsrc, T theta_inv, T* __restrict__ r,
                                                   T* __restrict__ d, T* __restrict__ x, double* __restrict__ zo) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T bi;
    if (bsrc) { bi = (T)bsrc[i]; b[i] = bi; } else bi = b[i];
    T ri = dinv[i] * bi;
    T di = ri * theta_inv;
    r[i] = ri; d[i] = di; x[i] = di;
    if (zo) zo[i] = (double)di;
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
template <class T> DBuf<T>& bx(AmgLevel& L);
template <> DBuf<double>& bx<double>(AmgLevel& L) { return L.x64; }
template <> DBuf<float>& bx<float>(AmgLevel& L) { return L.x32; }
template <class T> DBuf<T>& bw(AmgLevel& L);
template <> DBuf<double>& bw<double>(AmgLevel& L) { return L.w64; }
template <> DBuf<float>& bw<float>(AmgLevel& L) { return L.w32; }
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

struct Tags { const char *cheb, *chebx, *resid, *restrict_, *prolong; };
Tags level_tags(int l, bool fp64) {
    if (l == 0)
        return fp64 ? Tags{"amg_cheb_fp64_L0", "amg_chebx_fp64_L0", "amg_resid_fp64_L0", "amg_restrict_fp64_L0",
                           "amg_prolong_fp64_L0"}
                    : Tags{"amg_cheb_fp32_L0", "amg_chebx_fp32_L0", "amg_resid_fp32_L0", "amg_restrict_fp32_L0",
                           "amg_prolong_fp32_L0"};
    const char* t = fp64 ? "amg_coarse_fp64" : "amg_coarse_fp32";
    return Tags{t, t, t, t, t};
}

// Chebyshev polynomial smoother in D^-1 A on [lmax/10, 1.1 lmax] for the
// iterate held in the level's x buffer.  x_zero: start from x = 0 (no SpMV
// for the start).  bsrc: FP64 right-hand side to convert into b first.  zo:
// FP64 copy of the final iterate.
// Chebyshev interval [lo, hi] of a level's smoother
inline void cheb_interval(const AmgLevel& L, double& lo, double& hi) { hi = 1.1 * L.lmax; lo = hi / 10.0; }

template <class T>
void cheb_smooth(semo_ctx* c, AmgLevel& L, T* b, int degree, bool x_zero, const double* bsrc, double* zo,
                 const Tags& tg) {
    const DevSell<T>& A = opA<T>(L);
    int n = L.n;
    T* dA = vt<T>(L); T* dB = vu<T>(L); T* r = vr<T>(L);
    double hi, lo;
    cheb_interval(L, lo, hi);
    double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo);
    double sigma = theta / delta, rho = 1.0 / sigma;
    if (x_zero) {
        SEMO_LAUNCH(c, tg.chebx,
                    (double)sizeof(T) * 5 * n + (bsrc ? 8.0 * n : 0.0) + (degree == 1 && zo ? 8.0 * n : 0.0),
                    k_cheb_init<T>, dim3(nblk(n)), dim3(TPB), n, vdinv<T>(L), b, bsrc, (T)(1.0 / theta), r, dA,
                    bx<T>(L).p, degree == 1 ? zo : nullptr);
    } else if (degree == 1) {
        // the only step: r and y are dead, and with zo (the finest level)
        // so is the T iterate, which keeps the level's x buffer in place
        RowArgs<T> a{};
        a.x = bx<T>(L).p; a.b = b; a.dinv = vdinv<T>(L); a.zo = zo; a.c1 = (T)(1.0 / theta);
        if (!zo) a.xo = bw<T>(L).p;
        row_launch<T, OP_CHEBX>(c, A, a, tg.chebx, op_bytes(A, 3) + (zo ? 8.0 * n : (double)sizeof(T) * n));
        if (!zo) std::swap(bx<T>(L), bw<T>(L));
    } else {
        RowArgs<T> a{};
        a.x = bx<T>(L).p; a.b = b; a.y = dA; a.dinv = vdinv<T>(L); a.r = r; a.xo = bw<T>(L).p;
        a.c1 = (T)(1.0 / theta);
        row_launch<T, OP_CHEBX>(c, A, a, tg.chebx, op_bytes(A, 6));
        std::swap(bx<T>(L), bw<T>(L));       // the new iterate is the level's x from here on
    }
    T* x = bx<T>(L).p;
    for (int k = 1; k < degree; ++k) {
        double rho_new = 1.0 / (2.0 * sigma - rho);
        double c1 = rho_new * rho, c2 = 2.0 * rho_new / delta;
        rho = rho_new;
        RowArgs<T> a{};
        a.x = dA; a.y = dB; a.dinv = vdinv<T>(L); a.r = r; a.xo = x; a.c1 = (T)c1; a.c2 = (T)c2;
        a.zo = k == degree - 1 ? zo : nullptr;
        row_launch<T, OP_CHEB>(c, A, a, tg.cheb, op_bytes(A, 8));
        std::swap(dA, dB);
    }
}

// one V-cycle for level l on b; the result is left in the level's x buffer
// (and, at the finest level, converted into zo)
// prepared (degree 1): b and the smoothed start x of level l were written by
// the kernel that produced b (the Krylov residual update at level 0, the
// restriction below it)
template <class T>
void vcycle(semo_ctx* c, Amg& amg, int l, int degree, const double* bsrc, double* zo, bool prepared = false) {
    AmgLevel& L = amg.lv[l];
    int n = L.n;
    T* b = vb<T>(L);
    if (l + 1 == (int)amg.lv.size()) {     // coarsest: dense direct solve in FP64
        if (bsrc) SEMO_LAUNCH(c, (sizeof(T) == 8 ? "amg_coarse_fp64" : "amg_coarse_fp32"), (8.0 + sizeof(T)) * n,
                              (k_cast<double, T>), dim3(nblk(n)), dim3(TPB), n, bsrc, b);
        SEMO_LAUNCH(c, (sizeof(T) == 8 ? "amg_coarse_fp64" : "amg_coarse_fp32"), 8.0 * n * n + 2.0 * sizeof(T) * n,
                    k_dense_gemv<T>, dim3((n + 3) / 4), dim3(TPB), n, amg.coarse_inv.p, b, bx<T>(L).p);
        if (zo) SEMO_LAUNCH(c, (sizeof(T) == 8 ? "amg_coarse_fp64" : "amg_coarse_fp32"), (8.0 + sizeof(T)) * n,
                            (k_cast<T, double>), dim3(nblk(n)), dim3(TPB), n, bx<T>(L).p, zo);
        return;
    }
    AmgLevel& C = amg.lv[l + 1];
    Tags tg = level_tags(l, sizeof(T) == 8);
    T* t = vt<T>(L);
    const DevSell<T>& A = opA<T>(L);
    if (degree == 1 && prepared) {
        spmv<T, OP_RES>(c, A, bx<T>(L).p, b, t, tg.resid);
    } else {
        cheb_smooth<T>(c, L, b, degree, true, bsrc, nullptr, tg);
        spmv<T, OP_RES>(c, A, bx<T>(L).p, b, t, tg.resid);
    }
    // a coarse level that smooths (not the coarsest) gets its b and its
    // degree-1 smoothed start from the restriction
    bool prep_next = degree == 1 && l + 2 < (int)amg.lv.size();
    if (prep_next) {
        double chi, clo;
        cheb_interval(C, clo, chi);
        RowArgs<T> a{};
        a.x = t; a.y = vb<T>(C); a.dinv = vdinv<T>(C); a.xo = bx<T>(C).p; a.c1 = (T)(1.0 / (0.5 * (chi + clo)));
        row_launch<T, OP_YPREP>(c, opR<T>(L), a, tg.restrict_,
                                op_bytes(opR<T>(L), 0) + (double)sizeof(T) * (n + 3.0 * C.n));
    } else {
        spmv<T, OP_Y>(c, opR<T>(L), t, (const T*)nullptr, vb<T>(C), tg.restrict_);
    }
    vcycle<T>(c, amg, l + 1, degree, nullptr, nullptr, prep_next);
    spmv<T, OP_ADD>(c, opP<T>(L), bx<T>(C).p, (const T*)nullptr, bx<T>(L).p, tg.prolong);
    cheb_smooth<T>(c, L, b, degree, false, nullptr, zo, tg);
}

// z = M^-1 r with one V-cycle in precision T.  The V-cycle is a fixed
// sequence of ~22 kernels over fixed buffers, so it is captured once as a HIP
// graph and replayed: one submission instead of ~22.  It holds no memset and
// no library call.  Profiling (per-kernel events) runs it eagerly.
int cheb_degree(const semo_ctx* c) { return c->sp.cheb_degree > 0 ? c->sp.cheb_degree : 2; }

template <class T>
void sig_op(std::vector<const void*>& s, const DevSell<T>& M) {
    s.push_back(M.sl_ptr.p); s.push_back(M.col.p); s.push_back(M.dcol.p); s.push_back(M.val.p);
    s.push_back(M.diag.p);
}
// every device address a V-cycle in precision T passes to its kernels
template <class T>
std::vector<const void*> vcycle_signature(const Amg& amg, const double* r64, const double* z64) {
    std::vector<const void*> s{r64, z64, amg.coarse_inv.p};
    for (const AmgLevel& L : amg.lv) {
        if (sizeof(T) == 8) {
            for (const void* p : {(const void*)L.x64.p, (const void*)L.b64.p, (const void*)L.r64.p,
                                  (const void*)L.t64.p, (const void*)L.u64.p, (const void*)L.w64.p,
                                  (const void*)L.dinv64.p})
                s.push_back(p);
        } else {
            for (const void* p : {(const void*)L.x32.p, (const void*)L.b32.p, (const void*)L.r32.p,
                                  (const void*)L.t32.p, (const void*)L.u32.p, (const void*)L.w32.p,
                                  (const void*)L.dinv32.p})
                s.push_back(p);
        }
        sig_op(s, opA<T>(L)); sig_op(s, opP<T>(L)); sig_op(s, opR<T>(L));
    }
    std::sort(s.begin(), s.end());
    return s;
}

template <class T>
void precondition(semo_ctx* c, Amg& amg, const double* r64, double* z64, bool prepared = false) {
    int deg = cheb_degree(c);
    if (!c->graphs || c->prof.on) {
        vcycle<T>(c, amg, 0, deg, r64, z64, prepared);
        return;
    }
    int prec = sizeof(T) == 4 ? 1 : 0;
    std::vector<const void*> sig = vcycle_signature<T>(amg, r64, z64);
    if (!amg.vg || amg.vg_stream != c->stream || amg.vg_sig != sig || amg.vg_deg != deg || amg.vg_prec != prec ||
        amg.vg_pre != prepared) {
        amg.drop_graph();
        hipGraph_t g = nullptr;
        SEMO_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        vcycle<T>(c, amg, 0, deg, r64, z64, prepared);
        SEMO_HIP(hipStreamEndCapture(c->stream, &g));
        hipError_t e = hipGraphInstantiate(&amg.vg, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        SEMO_HIP(e);
        amg.vg_stream = c->stream; amg.vg_deg = deg; amg.vg_prec = prec; amg.vg_pre = prepared;
        // the capture swapped some levels' x buffers (out-of-place smoothing
        // steps): the graph holds the addresses from before the swaps, the
        // same set
        amg.vg_sig = vcycle_signature<T>(amg, r64, z64);
        SEMO_CHECK(amg.vg_sig == sig, SEMO_EHIP, "V-cycle capture changed its buffer set");
    }
    SEMO_HIP(hipGraphLaunch(amg.vg, c->stream));
}

double host_scalar(semo_ctx* c, int idx) {
    double v;
    SEMO_HIP(hipMemcpyAsync(&v, c->scal.p + idx, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SEMO_HIP(hipStreamSynchronize(c->stream));
    return v;
}

Red red_of(semo_ctx* c) { return Red{c->red.p}; }


// sums the partials of the previous launch (nb blocks) and applies `what`
void finish(semo_ctx* c, int nb, int what, int slot = 0) {
    SEMO_LAUNCH(c, "krylov_finish", 16.0 * nb + 8.0, k_finish, dim3(1), dim3(FTPB), nb, c->red.p, c->scal.p, what,
                slot);
}

// grid-stride dot products a.b (+ c.d), finished as `what`
void dot_what(semo_ctx* c, int n, const double* a, const double* b, const double* cc, const double* d, int what,
              int slot = 0) {
    int nb = rblk(n);
    SEMO_LAUNCH(c, "krylov_dot", (cc ? 32.0 : 16.0) * n, k_dot, dim3(nb), dim3(TPB), n, a, b, cc, d, red_of(c));
    finish(c, nb, what, slot);
}
// ... stored into s[slot], s[slot+1]
void dot_store(semo_ctx* c, int n, const double* a, const double* b, const double* cc, const double* d, int slot) {
    dot_what(c, n, a, b, cc, d, W_STORE, slot);
}

// y = A x with dots of y (see k_spmv_dot), for the fine-level sliced-ELL operator
void spmv_dot(semo_ctx* c, const DevSell<double>& A, const double* x, double* y, const double* a0,
              const double* a1, int what) {
    int nb = nblk((long)A.nslices * 64);
    SEMO_LAUNCH(c, "spmv_fp64", op_bytes(A, 2 + (a0 ? 1 : 0) + (a1 ? 1 : 0)), k_spmv_dot, dim3(nb), dim3(TPB),
                A.n, A.sl_ptr.p, A.col.p, A.dcol.n ? A.dcol.p : nullptr, A.val.p, A.diag.p, x, y, a0, a1,
                red_of(c));
    finish(c, nb, what);
}

// flexible (Polak-Ribiere) PCG with the AMG preconditioner.  x0 = 0,
// stop at ||r|| <= tol ||b||.
int fpcg(semo_ctx* c, Amg& amg, const DevSell<double>& A, const double* b, double* x, double tol, int maxit,
         bool mixed, double* relres) {
    SEMO_CHECK(A.lanes == 0, SEMO_EINVAL, "Krylov operator must be sliced ELL");
    int n = A.n;
    size_t ln = (size_t)n + 64;
    c->kv64.ensure(8 * ln);
    double* r = c->kv64.p; double* z = r + ln; double* p = z + ln; double* q = p + ln; double* rold = q + ln;
    // mixed precision, degree-1 smoothing, at least two levels: the kernels
    // that produce r also write the FP32 V-cycle's level-0 input
    Prep pp{};
    bool prep = cheb_degree(c) == 1 && amg.lv.size() >= 2 && amg.lv[0].n == n;
    if (prep) {
        AmgLevel& L0 = amg.lv[0];
        double hi, lo;
        cheb_interval(L0, lo, hi);
        if (mixed) {
            pp.b32 = vb<float>(L0); pp.x32 = bx<float>(L0).p; pp.dinv32 = vdinv<float>(L0);
            pp.c1 = (float)(1.0 / (0.5 * (hi + lo)));
        } else {
            pp.b64 = vb<double>(L0); pp.x64 = bx<double>(L0).p; pp.dinv64 = vdinv<double>(L0);
            pp.c1d = 1.0 / (0.5 * (hi + lo));
        }
    }
    auto pre = [&]() {
        if (mixed) precondition<float>(c, amg, r, z, prep); else precondition<double>(c, amg, r, z, prep);
    };
    double prep_bytes = prep ? (mixed ? 8.0 : 16.0) * n : 0.0;     // the V-cycle's b and x written
    SEMO_LAUNCH(c, "fcg_init", 24.0 * n + prep_bytes, k_fcg_init, dim3(nblk(n)), dim3(TPB), n, b, x, r, red_of(c), pp);
    finish(c, nblk(n), W_STORE, S_X);
    double bb = host_scalar(c, S_X);
    if (bb == 0.0) { *relres = 0.0; return 0; }
    double nb = std::sqrt(bb);
    pre();
    SEMO_LAUNCH(c, "fcg_copy", 16.0 * n, (k_cast<double, double>), dim3(nblk(n)), dim3(TPB), n, z, p);
    dot_store(c, n, r, z, nullptr, nullptr, S_RZ);
    int it = 0;
    ConvCheck cc(c, tol * tol * bb, bb);
    double* p2 = rold + ln;                      // second direction buffer
    for (it = 1; it <= maxit; ++it) {
        if (it == 1) {
            spmv_dot(c, A, p, q, p, nullptr, W_FCG_ALPHA);                // q = A p, alpha
        } else {                                                          // p = z + beta p; q = A p, alpha
            int nbs = nblk((long)A.nslices * 64);
            SEMO_LAUNCH(c, "spmv_xpby_fp64", op_bytes(A, 4), k_spmv_xpby, dim3(nbs), dim3(TPB), A.n, A.sl_ptr.p,
                        A.col.p, A.dcol.n ? A.dcol.p : nullptr, A.val.p, A.diag.p, c->scal.p, z, p, p2, q,
                        red_of(c));
            finish(c, nbs, W_FCG_ALPHA);
            std::swap(p, p2);
        }
        SEMO_LAUNCH(c, "fcg_update", 56.0 * n + prep_bytes, k_fcg_update, dim3(nblk(n)), dim3(TPB), n, c->scal.p, x,
                    r, rold, p, q, red_of(c), pp);
        finish(c, nblk(n), W_STORE, S_RR);
        pipe_push(c, c->scal.p + S_RR, it);
        if (cc.done(it) || it == maxit) break;
        pre();
        dot_what(c, n, r, z, z, rold, W_FCG_BETA);                        // r.z, z.r_old -> beta
    }
    // true residual in FP64
    spmv<double, OP_RES>(c, A, x, b, r, "true_resid_fp64");
    dot_store(c, n, r, r, nullptr, nullptr, S_X);
    *relres = std::sqrt(host_scalar(c, S_X)) / nb;
    return std::min(it, maxit);
}

// right-preconditioned (Jacobi) BiCGStab, x0 = 0
// z = M^-1 r for one component of a momentum / volume-fraction solve: one
// V-cycle of `amg` in the solve's precision, through fixed staging buffers,
// so that one captured V-cycle graph serves every component and call
void stage_vcycle(semo_ctx* c, Amg& amg, const double* r, double* z) {
    size_t n = c->nc;
    double* sr = c->mom_stage.p;
    double* sz = sr + n + 64;
    copy_d(c, sr, r, n);
    if (c->sp.precision == SEMO_PREC_MIXED) precondition<float>(c, amg, sr, sz, false);
    else precondition<double>(c, amg, sr, sz, false);
    copy_d(c, z, sz, n);
}

int bicgstab(semo_ctx* c, const DevSell<double>& A, const double* dinv, const double* b, double* x, double tol,
             int maxit, double* relres, Amg* pre = nullptr) {
    SEMO_CHECK(A.lanes == 0, SEMO_EINVAL, "Krylov operator must be sliced ELL");
    int n = A.n;
    size_t ln = (size_t)n + 64;
    c->kv64.ensure(8 * ln);
    double* r = c->kv64.p; double* rh = r + ln; double* p = rh + ln; double* v = p + ln; double* sv = v + ln;
    double* t = sv + ln; double* ph = t + ln; double* sh = ph + ln;
    double* s = c->scal.p;
    SEMO_LAUNCH(c, "bicg_init", 48.0 * n, k_bicg_init, dim3(nblk(n)), dim3(TPB), n, b, x, r, rh, p, v, red_of(c));
    finish(c, nblk(n), W_BICG_INIT);
    double bb = host_scalar(c, S_X);
    if (bb == 0.0) { *relres = 0.0; return 0; }
    double nb = std::sqrt(bb);
    int it;
    ConvCheck cc(c, tol * tol * bb, bb);
    for (it = 1; it <= maxit; ++it) {
        SEMO_LAUNCH(c, "bicg_p", 48.0 * n, k_bicg_p, dim3(nblk(n)), dim3(TPB), n, s, r, v, dinv, p, ph);
        if (pre) stage_vcycle(c, *pre, p, ph);                            // ph = M^-1 p (AMG)
        spmv_dot(c, A, ph, v, rh, nullptr, W_BICG_ALPHA);                 // v = A ph, alpha
        SEMO_LAUNCH(c, "bicg_s", 40.0 * n, k_bicg_s, dim3(nblk(n)), dim3(TPB), n, s, r, v, dinv, sv, sh);
        if (pre) stage_vcycle(c, *pre, sv, sh);
        spmv_dot(c, A, sh, t, nullptr, sv, W_BICG_OMEGA);                 // t = A sh, omega
        SEMO_LAUNCH(c, "bicg_x", 64.0 * n, k_bicg_x, dim3(nblk(n)), dim3(TPB), n, s, ph, sh, sv, t, rh, x, r,
                    red_of(c));
        finish(c, nblk(n), W_BICG_RHO);
        pipe_push(c, c->scal.p + S_RR, it);
        if (cc.done(it)) break;
    }
    spmv<double, OP_RES>(c, A, x, b, r, "true_resid_fp64");
    dot_store(c, n, r, r, nullptr, nullptr, S_X);
    *relres = std::sqrt(host_scalar(c, S_X)) / nb;
    return std::min(it, maxit);
}

// BiCGStab on the three momentum components in lockstep (Jacobi, x0 = 0);
// stops when every component has ||r_k|| <= tol ||b_k||
int bicgstab3(semo_ctx* c, const DevSell<double>& A, const double* dinv, const double* b, double* x, double tol,
              int maxit, semo_solve_stats* st, Amg* pre = nullptr) {
    SEMO_CHECK(A.lanes == 0, SEMO_EINVAL, "Krylov operator must be sliced ELL");
    int n = A.n;
    size_t ln = 3 * (size_t)n + 64;
    c->kv64.ensure(8 * ln);
    double* r = c->kv64.p; double* rh = r + ln; double* p = rh + ln; double* v = p + ln; double* sv = v + ln;
    double* t = sv + ln; double* ph = t + ln; double* sh = ph + ln;
    double* s = c->scal.p;
    int nb = nblk(n), nbs = nblk((long)A.nslices * 64);
    size_t need = 6 * (size_t)std::max(nb, nbs) + 64;
    if (c->red.n < need) c->red.alloc(need);
    double* part = c->red.p;
    const int16_t* dc = A.dcol.n ? A.dcol.p : nullptr;
    auto fin = [&](int blocks, int what) {
        SEMO_LAUNCH(c, "bicg3_finish", 48.0 * blocks + 8.0, k_finish6, dim3(1), dim3(1024), blocks, part, s, what);
    };
    auto spmv3 = [&](int mode, const double* xin, double* y, const double* a) {
        // matrix once, three gathered operands, three results, three dot operands
        ProfScope ps(c, mode == 2 ? "spmv3_res_fp64" : "spmv3_fp64", op_bytes(A, 9));
        if (mode == 0)
            hipLaunchKernelGGL(k_spmv3<0>, dim3(nbs), dim3(TPB), 0, c->stream, n, A.sl_ptr.p, A.col.p, dc, A.val.p,
                               A.diag.p, xin, y, a, part);
        else if (mode == 1)
            hipLaunchKernelGGL(k_spmv3<1>, dim3(nbs), dim3(TPB), 0, c->stream, n, A.sl_ptr.p, A.col.p, dc, A.val.p,
                               A.diag.p, xin, y, a, part);
        else
            hipLaunchKernelGGL(k_spmv3<2>, dim3(nbs), dim3(TPB), 0, c->stream, n, A.sl_ptr.p, A.col.p, dc, A.val.p,
                               A.diag.p, xin, y, a, part);
    };
    SEMO_LAUNCH(c, "bicg3_init", 144.0 * n, k_bicg3_init, dim3(nb), dim3(TPB), n, b, x, r, rh, p, v, part);
    fin(nb, F3_INIT);
    double bb[3];
    SEMO_HIP(hipMemcpyAsync(bb, s + Q_BB, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SEMO_HIP(hipStreamSynchronize(c->stream));
    int it = 0;
    if (bb[0] > 0.0 || bb[1] > 0.0 || bb[2] > 0.0) {
        ConvCheck cc(c, tol * tol, 1.0);
        for (it = 1; it <= maxit; ++it) {
            SEMO_LAUNCH(c, "bicg3_p", 128.0 * n, k_bicg3_p, dim3(nb), dim3(TPB), n, s, r, v, dinv, p, ph);
            if (pre)                                               // ph_k = M^-1 p_k (AMG)
                for (int k = 0; k < 3; ++k) stage_vcycle(c, *pre, p + (size_t)k * n, ph + (size_t)k * n);
            spmv3(0, ph, v, rh);                                   // v = A ph, rh.v
            fin(nbs, F3_ALPHA);
            SEMO_LAUNCH(c, "bicg3_s", 104.0 * n, k_bicg3_s, dim3(nb), dim3(TPB), n, s, r, v, dinv, sv, sh);
            if (pre)
                for (int k = 0; k < 3; ++k) stage_vcycle(c, *pre, sv + (size_t)k * n, sh + (size_t)k * n);
            spmv3(1, sh, t, sv);                                   // t = A sh, t.s, t.t
            fin(nbs, F3_OMEGA);
            SEMO_LAUNCH(c, "bicg3_x", 192.0 * n, k_bicg3_x, dim3(nb), dim3(TPB), n, s, ph, sh, sv, t, rh, x, r, part);
            fin(nb, F3_RHO);
            pipe_push(c, s + Q_MAXREL, it);
            if (cc.done(it)) break;
        }
        it = std::min(it, maxit);
    }
    // true residuals in FP64
    spmv3(2, x, r, b);
    fin(nbs, F3_STORE);
    double rr[3];
    SEMO_HIP(hipMemcpyAsync(rr, s + Q_RR, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SEMO_HIP(hipStreamSynchronize(c->stream));
    for (int k = 0; k < 3; ++k) {
        double rel = bb[k] > 0.0 ? std::sqrt(rr[k] / bb[k]) : 0.0;
        if (st) { st->iters[k] = bb[k] > 0.0 ? it : 0; st->rel_res[k] = rel; }
        check_converged(std::isfinite(bb[k]) ? rel : bb[k], tol, "momentum");
    }
    return it;
}

// level-0 operator of the hierarchy refreshed from the assembled system
// (same pattern, new values): off-diagonals through the slot map, diagonal,
// D^-1, FP32 copies
__global__ void __launch_bounds__(TPB) k_refresh_offdiag(size_t nv, const int* __restrict__ map,
                                                         const double* __restrict__ src, double* __restrict__ v64,
                                                         float* __restrict__ v32) {
    size_t e = (size_t)xcd_block() * blockDim.x + threadIdx.x;
    if (e >= nv) return;
    int m = map[e];
    double v = m >= 0 ? src[m] : 0.0;
    v64[e] = v; v32[e] = (float)v;
}
__global__ void __launch_bounds__(TPB) k_refresh_diag(int n, const double* __restrict__ d, double* __restrict__ d64,
                                                      float* __restrict__ d32, double* __restrict__ di64,
                                                      float* __restrict__ di32) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = d[i], iv = 1.0 / v;
    d64[i] = v; d32[i] = (float)v; di64[i] = iv; di32[i] = (float)iv;
}

__global__ void __launch_bounds__(TPB) k_dinv(int n, const double* __restrict__ d, double* __restrict__ o) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i < n) o[i] = 1.0 / d[i];
}

}  // namespace

// ---------------------------------------------------------------- uploads --
template <class T>
void sell_upload(DevSell<T>& M, const SellPattern& p, const std::vector<double>& vals,
                 const std::vector<double>& diag, hipStream_t s) {
    M.n = p.n; M.ncols = p.ncols; M.nslices = p.nslices; M.lanes = 0;
    M.nnz = p.nnz_offdiag + (int64_t)diag.size();
    M.sl_ptr.upload(p.sl_ptr, s);
    M.col.upload(p.col, s);
    sell_narrow(p, M.dcol, s);
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

// column - row as int16 when every entry of a square pattern allows it;
// entries of the rows past n in the last slice (value 0) point at row n - 1.
// Leaves d empty otherwise.
void sell_narrow(const SellPattern& p, DBuf<int16_t>& d, hipStream_t s) {
    d.release();
    if (p.n != p.ncols || p.n == 0) return;
    std::vector<int16_t> h(p.col.size());
    for (int sl = 0; sl < p.nslices; ++sl)
        for (int e = p.sl_ptr[sl]; e < p.sl_ptr[sl + 1]; ++e) {
            int row = sl * SL + sell_lane(p.sl_ptr[sl], e);
            long dv = row < p.n ? (long)p.col[e] - row : (long)(p.n - 1) - row;
            if (dv < -32767 || dv > 32767) return;
            h[e] = (int16_t)dv;
        }
    d.upload(h, s);
    SEMO_HIP(hipStreamSynchronize(s));
}

template <class T>
void csr_upload(DevSell<T>& M, int n, int ncols, const std::vector<int>& ptr, const std::vector<int>& col,
                const std::vector<double>& vals, const std::vector<double>& diag, int lanes, hipStream_t s) {
    SEMO_CHECK(n >= 0 && ncols >= 1 && (int)ptr.size() == n + 1 && ptr[0] == 0 && ptr[n] == (int)col.size() &&
                   vals.size() == col.size() && (diag.empty() || (int)diag.size() == n),
               SEMO_EINVAL, "CSR operator: inconsistent arrays");
    SEMO_CHECK(lanes == 4 || lanes == 8 || lanes == 16 || lanes == 32 || lanes == 64, SEMO_EINVAL,
               "CSR operator: lanes per row");
    for (int r = 0; r < n; ++r) SEMO_CHECK(ptr[r + 1] >= ptr[r], SEMO_EINVAL, "CSR operator: row pointers");
    for (int cidx : col) SEMO_CHECK(cidx >= 0 && cidx < ncols, SEMO_EINVAL, "CSR operator: column out of range");
    M.n = n; M.ncols = ncols; M.nslices = (n + SL - 1) / SL; M.lanes = lanes;
    M.nnz = (int64_t)col.size() + (int64_t)diag.size();
    M.sl_ptr.upload(ptr, s);
    M.col.upload(col, s);
    std::vector<T> v(vals.begin(), vals.end());
    M.val.upload(v, s);
    if (!diag.empty()) {
        std::vector<T> d(diag.begin(), diag.end());
        M.diag.upload(d, s);
    } else M.diag.release();
    SEMO_HIP(hipStreamSynchronize(s));
}
template void csr_upload<double>(DevSell<double>&, int, int, const std::vector<int>&, const std::vector<int>&,
                                 const std::vector<double>&, const std::vector<double>&, int, hipStream_t);
template void csr_upload<float>(DevSell<float>&, int, int, const std::vector<int>&, const std::vector<int>&,
                                const std::vector<double>&, const std::vector<double>&, int, hipStream_t);

// partials for reductions over up to n rows (two values per block)
void reduction_alloc(semo_ctx* c, long n) {
    size_t need = 2 * (size_t)std::max(nblk(std::max<long>(n, 1) + 64), RB) + 64;
    if (c->red.n < need) c->red.alloc(need);
    if (c->scal.n < 64) c->scal.alloc(64);
}

// ----------------------------------------------------- equation solves ----
// The pressure hierarchy for the operator in c->Amat: built when missing;
// when the values changed, the finest level is refreshed on the device while
// keep_coarse says the coarse levels still serve, else rebuilt.
// the hierarchy `amg` for the operator in c->Amat, whose values it was built
// from are kept in key_val / key_diag (device copies)
void amg_update(semo_ctx* c, Amg& amg, DBuf<double>& key_val, DBuf<double>& key_diag,
                bool (*keep_coarse)(const Amg&)) {
    bool rebuild = !amg.built;
    if (amg.built) {
        // device-side comparison with the values the hierarchy was built from
        DBuf<int>& flag = c->d_flag1;
        flag.ensure(1);
        zero_d(c, flag.p, sizeof(int));
        size_t nv = c->Amat.val.n;
        SEMO_LAUNCH(c, "amg_values_changed", 16.0 * nv, k_neq, dim3((unsigned)((nv + TPB - 1) / TPB)), dim3(TPB), nv,
                    c->Amat.val.p, key_val.p, flag.p);
        SEMO_LAUNCH(c, "amg_values_changed", 16.0 * c->nc, k_neq, dim3((unsigned)((c->nc + TPB - 1) / TPB)),
                    dim3(TPB), (size_t)c->nc, c->Amat.diag.p, key_diag.p, flag.p);
        int f = 0;
        SEMO_HIP(hipMemcpyAsync(&f, flag.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        SEMO_HIP(hipStreamSynchronize(c->stream));
        rebuild = f != 0;
    }
    // New values on the same pattern (a variable-density flow): refresh the
    // finest level on the device and keep the coarse levels, as long as the
    // flexible CG converges within 1.5x (+3) the iterations it needed right
    // after the last full set-up; otherwise, or without a slot map, rebuild.
    // The outer Krylov method always uses the current operator, so this only
    // affects the preconditioner's quality, never the solution.
    bool refresh = false;
    if (rebuild && amg.built && amg.l0_map.n == amg.lv[0].A64.val.n && amg.l0_map.n > 0 && keep_coarse(amg)) {
        refresh = true;
        rebuild = false;
    }
    if (refresh) {
        AmgLevel& L0 = amg.lv[0];
        size_t nv = L0.A64.val.n;
        SEMO_LAUNCH(c, "amg_refresh_L0", 24.0 * nv, k_refresh_offdiag, dim3((unsigned)((nv + TPB - 1) / TPB)),
                    dim3(TPB), nv, amg.l0_map.p, c->Amat.val.p, L0.A64.val.p, L0.A32.val.p);
        SEMO_LAUNCH(c, "amg_refresh_L0", 32.0 * c->nc, k_refresh_diag, dim3(nblk(c->nc)), dim3(TPB), c->nc,
                    c->Amat.diag.p, L0.A64.diag.p, L0.A32.diag.p, L0.dinv64.p, L0.dinv32.p);
        copy_d(c, key_val.p, c->Amat.val.p, c->Amat.val.n);
        copy_d(c, key_diag.p, c->Amat.diag.p, c->Amat.diag.n);
        amg.fresh = false;
    }
    if (rebuild) {
        amg.built = false;
        amg_build(c, amg, c->Amat);
        key_val.alloc(c->Amat.val.n);
        key_diag.alloc(c->nc);
        SEMO_HIP(hipMemcpyAsync(key_val.p, c->Amat.val.p, c->Amat.val.bytes(), hipMemcpyDeviceToDevice, c->stream));
        SEMO_HIP(hipMemcpyAsync(key_diag.p, c->Amat.diag.p, c->Amat.diag.bytes(), hipMemcpyDeviceToDevice, c->stream));
        amg.fresh = true;
    }
}

void pressure_amg_update(semo_ctx* c, bool (*keep_coarse)(const Amg&)) {
    amg_update(c, c->amg_p, c->amg_key_val, c->amg_key_diag, keep_coarse);
}

// z = M^-1 r: one V-cycle of the pressure hierarchy, FP32 (fp32) or FP64,
// FP64 in and out (the Schur-complement step of the coupled solver)
void pressure_vcycle(semo_ctx* c, const double* r64, double* z64, bool fp32) {
    SEMO_CHECK(c->amg_p.built, SEMO_ESTATE, "pressure hierarchy not built");
    if (fp32) precondition<float>(c, c->amg_p, r64, z64, false);
    else precondition<double>(c, c->amg_p, r64, z64, false);
}

// flexible CG: the coarse levels serve while the iteration count stays
// within 1.5x (+3) of the count right after the last full set-up
bool fcg_keeps_coarse(const Amg& amg) {
    return amg.iters_ref >= 0 && amg.last_iters <= std::max((3 * amg.iters_ref) / 2, amg.iters_ref + 3);
}

// pressure: hierarchy (re)built when the operator's values changed
void solve_pressure(semo_ctx* c, const double* b, double* x, semo_solve_stats* st) {
    Amg& amg = c->amg_p;
    pressure_amg_update(c, fcg_keeps_coarse);
    reduction_alloc(c, c->nc);
    double rr = 0;
    int it = fpcg(c, amg, c->Amat, b, x, c->sp.tol_p, c->sp.maxit_p, c->sp.precision == SEMO_PREC_MIXED, &rr);
    if (amg.fresh) { amg.iters_ref = it; amg.fresh = false; }
    amg.last_iters = it;
    if (st) {
        st->iters[0] = it; st->rel_res[0] = rr; st->n_levels = (int)amg.lv.size(); st->setup_ms = amg.setup_ms;
    }
    check_converged(rr, c->sp.tol_p, "pressure");
}

// BiCGStab keeps the coarse levels of its AMG preconditioner while it
// converges within 1.5x (+3) the iterations right after the last build
bool bicg_keeps_coarse(const Amg& amg) {
    return amg.iters_ref >= 0 && amg.last_iters <= std::max((3 * amg.iters_ref) / 2, amg.iters_ref + 3);
}

// momentum (nrhs = 3, one operator, three right-hand sides in lockstep) and
// volume fraction (nrhs = 1): BiCGStab right-preconditioned by Jacobi, or by
// an AMG V-cycle of the system's own hierarchy (SEMO_MSOLVER_AMG, or
// SEMO_MSOLVER_AUTO once a Jacobi solve of that system has needed more than
// MSOLVER_AUTO_ITERS iterations), as SEMO's AMGCL solves precondition them
// with aggregation AMG (solveAMGCL.cpp:881-927, 1195-1258)
constexpr int MSOLVER_AUTO_ITERS = 30;
void solve_momentum(semo_ctx* c, int nrhs, semo_solve_stats* st) {
    int n = c->nc;
    reduction_alloc(c, n);
    DBuf<double>& dinv = c->dinv_mom;
    dinv.alloc(n + 64);
    SEMO_LAUNCH(c, "jacobi_dinv", 16.0 * n, k_dinv, dim3(nblk(n)), dim3(TPB), n, c->Amat.diag.p, dinv.p);
    bool mom = nrhs == 3;
    bool& auto_on = mom ? c->amg_mom_on : c->amg_vf_on;
    Amg* pre = nullptr;
    if (c->msolver == SEMO_MSOLVER_AMG || (c->msolver == SEMO_MSOLVER_AUTO && auto_on)) {
        Amg& amg = mom ? c->amg_m : c->amg_vf;
        amg_update(c, amg, mom ? c->amg_m_kv : c->amg_vf_kv, mom ? c->amg_m_kd : c->amg_vf_kd, bicg_keeps_coarse);
        c->mom_stage.ensure(2 * ((size_t)n + 64));
        pre = &amg;
    }
    int it = 0;
    if (nrhs == 3) {
        it = bicgstab3(c, c->Amat, dinv.p, c->Bvec.p, c->Xvec.p, c->sp.tol_u, c->sp.maxit_u, st, pre);
    } else {
        for (int k = 0; k < nrhs; ++k) {
            double rr = 0;
            int ik = bicgstab(c, c->Amat, dinv.p, c->Bvec.p + (size_t)k * n, c->Xvec.p + (size_t)k * n,
                              c->sp.tol_u, c->sp.maxit_u, &rr, pre);
            if (st) { st->iters[k] = ik; st->rel_res[k] = rr; }
            check_converged(rr, c->sp.tol_u, "momentum/volume-fraction");
            it = std::max(it, ik);
        }
    }
    if (pre) {
        if (pre->fresh) { pre->iters_ref = it; pre->fresh = false; }
        pre->last_iters = it;
        if (st) { st->n_levels = (int)pre->lv.size(); st->setup_ms = pre->setup_ms; }
    } else if (c->msolver == SEMO_MSOLVER_AUTO && it > MSOLVER_AUTO_ITERS) {
        auto_on = true;
    }
}

// standalone COO system (seam 2)
void solve_coo(semo_ctx* c, const char* eq, int n, int64_t nnz, const int* rows, const int* cols,
               const double* vals, int nrhs, const double* b, double* x, semo_solve_stats* st) {
    SEMO_CHECK(n > 0 && nnz >= n && nrhs >= 1 && nrhs <= 4, SEMO_EINVAL, "solve_coo sizes");
    for (int64_t k = 0; k < nnz; ++k) {
        SEMO_CHECK(rows[k] >= 0 && rows[k] < n && cols[k] >= 0 && cols[k] < n, SEMO_EINVAL, "COO index out of range");
        SEMO_CHECK(std::isfinite(vals[k]), SEMO_EINVAL, "COO value is not finite");
    }
    for (int64_t k = 0; k < (int64_t)n * nrhs; ++k)
        SEMO_CHECK(std::isfinite(b[k]), SEMO_EINVAL, "right-hand side is not finite");
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
    reduction_alloc(c, n);
    std::string e(eq ? eq : "");
    if (e == "pressure") {
        Amg amg;
        amg_build_csr(c, amg, n, cptr, ccol, cval);
        for (int k = 0; k < nrhs; ++k) {
            double rr = 0;
            int it = fpcg(c, amg, A, db.p + (size_t)k * n, dx.p + (size_t)k * n, c->sp.tol_p, c->sp.maxit_p,
                          c->sp.precision == SEMO_PREC_MIXED, &rr);
            if (st) { st->iters[k] = it; st->rel_res[k] = rr; st->n_levels = (int)amg.lv.size(); st->setup_ms = amg.setup_ms; }
            check_converged(rr, c->sp.tol_p, "pressure");
        }
        // the local hierarchy and its graph go out of scope here
        SEMO_HIP(hipStreamSynchronize(c->stream));
    } else {
        DBuf<double> dinv; dinv.alloc(n + 64);
        SEMO_LAUNCH(c, "jacobi_dinv", 16.0 * n, k_dinv, dim3(nblk(n)), dim3(TPB), n, A.diag.p, dinv.p);
        for (int k = 0; k < nrhs; ++k) {
            double rr = 0;
            int it = bicgstab(c, A, dinv.p, db.p + (size_t)k * n, dx.p + (size_t)k * n, c->sp.tol_u, c->sp.maxit_u, &rr);
            if (st) { st->iters[k] = it; st->rel_res[k] = rr; }
            check_converged(rr, c->sp.tol_u, "solve");
        }
    }
    dx.download(x, (size_t)n * nrhs, c->stream);
    SEMO_HIP(hipStreamSynchronize(c->stream));
}

}  // namespace semo
