// Internal declarations of libsemo_gpu (MI355X / gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/semo_gpu.h"

namespace semo {

// ---------------------------------------------------------------- errors ----
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define SEMO_HIP(call)                                                                   \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            throw ::semo::Error(SEMO_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define SEMO_CHECK(cond, code, msg)                                                      \
    do {                                                                                 \
        if (!(cond)) throw ::semo::Error((code), (msg));                                 \
    } while (0)

// device error flags raised by kernels (bit set in ctx->d_flags)
enum : int { FLAG_ZERO_AREA = 1, FLAG_NEG_VOLUME = 2, FLAG_SINGULAR_LSQ = 4,
             FLAG_BAD_FACE = 8, FLAG_ZERO_DIAG = 16 };

// ------------------------------------------------------ XCD-aware blocks ---
// Workgroups are dispatched to the 8 XCDs round-robin (block b runs on XCD
// b % 8), and each XCD has its own L2.  Remapping the block index so that
// each XCD walks one contiguous range of the grid, in order, keeps the cells
// and faces that neighbour each other (RCM order) in the same L2.  This is a
// bijection on [0, gridDim.x) for any grid size; only locality depends on the
// dispatch order, never correctness.
__device__ inline int xcd_block() {
    int nb = (int)gridDim.x, b = (int)blockIdx.x;
    int q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
    return x * q + (x < r ? x : r) + k;
}

// ------------------------------------------------------------ device buffer -
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr; n = 0;
    }
    void alloc(size_t count) {
        if (count == n && p) return;
        release();
        if (count) SEMO_HIP(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    // grow-only allocation (contents not preserved when it grows)
    void ensure(size_t count) {
        if (count > n || !p) alloc(count);
    }
    void upload(const T* h, size_t count, hipStream_t s) {
        alloc(count);
        if (count) SEMO_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
    void download(T* h, size_t count, hipStream_t s) const {
        if (count) SEMO_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
    void zero(hipStream_t s) {
        if (n) SEMO_HIP(hipMemsetAsync(p, 0, n * sizeof(T), s));
    }
    size_t bytes() const { return n * sizeof(T); }
};

// --------------------------------------------------------------- profiler ---
struct KernelStat { int64_t count = 0; double ms = 0; double bytes = 0; };
struct Profiler {
    bool on = false;
    std::map<std::string, KernelStat> stats;
    struct Pending { std::string name; hipEvent_t a, b; double bytes; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    hipEvent_t get();
    void begin(hipStream_t s, hipEvent_t& a);
    void end(hipStream_t s, const char* name, hipEvent_t a, double bytes);
    void flush();   // resolves pending events (synchronizes)
    ~Profiler();
};

// ------------------------------------------------------- sliced-ELL matrix ---
// Off-diagonal entries in slices of 64 rows (one wavefront per slice, one lane
// per row), in entry pairs inside a slice (sell_pos), plus a separate diagonal.  Padding
// entries have val = 0 and a column inside [0, ncols) (the row itself for
// square matrices, 0 for rectangular ones).
// Operators with long or ragged rows (AMG restriction and coarse levels) use
// CSR instead (lanes > 0): sl_ptr then holds the n+1 row offsets and `lanes`
// consecutive lanes of a wavefront share one row.
constexpr int SL = 64;
// lanes per cell of the one-field LSQ pass when not set: 4 from this
// average stencil length on, else 1
constexpr double LSQ_LANES_LONG_STENCIL = 40.0;
// Inside a slice the entries go in pairs: pair m (entries 2m, 2m+1 of each
// row) of the 64 rows is 128 consecutive elements, row r's two at 2r, 2r+1,
// so one 16-byte (FP64) or 8-byte (FP32) load per lane brings two entries of
// its row and a wavefront's load is one contiguous 1 KB / 512 B run.  Slice
// widths are even (a padding entry completes an odd one).
inline int sell_pos(int base, int k, int lane) { return base + (k >> 1) * (2 * SL) + 2 * lane + (k & 1); }
inline int sell_lane(int base, int e) { return ((e - base) >> 1) & (SL - 1); }
inline int sell_entry(int base, int e) { return (((e - base) / (2 * SL)) << 1) | ((e - base) & 1); }
struct SellPattern {
    int n = 0, ncols = 0, nslices = 0;
    std::vector<int> sl_ptr;    // element offset of each slice (nslices+1)
    std::vector<int> col;       // padded columns
    int64_t nnz_offdiag = 0;    // real off-diagonal entries
};
template <class T>
struct DevSell {
    int n = 0, ncols = 0, nslices = 0;
    int lanes = 0;              // 0: sliced ELL; 4..64: CSR, lanes per row
    int64_t nnz = 0;            // real entries incl. diagonal
    DBuf<int> sl_ptr, col;
    // sliced ELL of a square operator whose columns all lie within +-32767
    // of their row (RCM bandwidth): column - row as int16, read instead of
    // col (2 B per entry instead of 4); empty otherwise
    DBuf<int16_t> dcol;
    DBuf<T> val, diag;
    size_t padded() const { return col.n; }
};

// ------------------------------------------------------------ AMG level -----
struct AmgLevel {
    int n = 0;
    DevSell<double> A64;        // operator (fp64)
    DevSell<float> A32;         // operator (fp32 copy, mixed precision)
    DBuf<int> agg;              // fine row -> aggregate (coarse row), for R/P of next
    DevSell<double> P64, R64;   // smoothed prolongation / restriction (next level)
    DevSell<float> P32, R32;
    DBuf<double> dinv64; DBuf<float> dinv32;
    double lmax = 2.0;          // spectral bound of D^-1 A
    // work vectors (w: second buffer of the iterate x for out-of-place updates)
    DBuf<double> x64, b64, r64, t64, u64, w64;
    DBuf<float> x32, b32, r32, t32, u32, w32;
};

struct Amg {
    std::vector<AmgLevel> lv;
    int nc_dense = 0;
    DBuf<double> coarse_inv;    // dense inverse of the coarsest operator (fp64)
    double setup_ms = 0;
    bool built = false;
    std::vector<double> key;    // host copy of the fine values the hierarchy was built from
    // level-0 refresh: device element of the level-0 operator -> slot of the
    // assembled system (-1: padding); Krylov iterations right after the last
    // full build and of the last solve (rebuild policy)
    DBuf<int> l0_map;
    int iters_ref = -1, last_iters = -1;
    bool fresh = false;         // the last solve used a freshly built hierarchy
    // one V-cycle captured as a HIP graph (replayed per Krylov iteration):
    // valid for the hierarchy's buffers and the operands it was captured
    // with.  vg_sig lists every device address the captured kernels take
    // (operands, level vectors and operators), sorted: a replay whose current
    // list differs re-captures, so a reallocated or swapped buffer is never
    // reached through a stale graph.
    hipGraphExec_t vg = nullptr;
    hipStream_t vg_stream = nullptr;      // stream of the exec's launches
    std::vector<const void*> vg_sig;
    int vg_deg = 0, vg_prec = -1;
    bool vg_pre = false;
    // The exec's last launch may still be queued: wait for its stream before
    // the exec (and, in the callers, the buffers it references) goes away.
    void drop_graph() {
        if (vg) {
            if (vg_stream) (void)hipStreamSynchronize(vg_stream);
            (void)hipGraphExecDestroy(vg);
        }
        vg = nullptr; vg_stream = nullptr; vg_sig.clear(); vg_prec = -1;
    }
    void reset() {
        drop_graph();
        lv.clear(); nc_dense = 0; coarse_inv.release(); setup_ms = 0; built = false; key.clear();
        l0_map.release(); iters_ref = last_iters = -1; fresh = false;
    }
    Amg() = default;
    Amg(const Amg&) = delete;
    Amg& operator=(const Amg&) = delete;
    ~Amg() { drop_graph(); }
};

struct AdjBufs;                 // adjoint.hip: checkpoints and adjoint work arrays
struct Coupled;                 // coupled.hip: block (4x4) system and its AMG

}  // namespace semo

// ------------------------------------------------------------------ context --
struct semo_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    bool graphs = true;                    // replay V-cycles as HIP graphs (SEMO_GPU_GRAPHS=0: off)
    std::string err;
    semo::Profiler prof;

    // ---- mesh (device numbering) ----
    bool has_mesh = false;
    int np = 0, nc = 0, nf = 0, nif = 0, nbf = 0, npatch = 0;
    std::vector<int> perm, inv;            // cells: new->old, old->new
    std::vector<int> fperm;                // faces: new->old
    std::vector<int8_t> fflip;
    std::vector<int> patch_start, patch_nfaces;
    std::vector<int> h_owner, h_nbr;       // device numbering (host copies)
    std::vector<int> h_cf_ptr, h_cf_idx;   // cell -> faces (host copy)

    semo::DBuf<double> d_pts;
    semo::DBuf<int> d_owner, d_nbr, d_fptr, d_fpts;
    semo::DBuf<int> d_cf_ptr, d_cf_idx;    // cell -> faces, ascending
    semo::DBuf<int> d_st_ptr, d_st_idx;    // vertex stencil (CSR)
    // the same stencil as sliced ELL: 64-cell slices, entry k of cell
    // 64s+lane at st_sl[s] + 64k + lane, padding -1 after the last entry
    semo::DBuf<int> d_st_sl, d_st_ell;
    // the same slices as int16 offsets neighbour - cell, -32768 padding; empty
    // when some neighbour lies further than 32767 from its cell
    semo::DBuf<int16_t> d_st_dl;
    // per-face fixedValue-U values [3][nbf] (device face order), empty: the
    // patches' uniform values
    semo::DBuf<double> b_ufix;
    // packed gather records of the one-field LSQ ([x, y, z, phi]) and of the
    // curvature divergence ([x, y, z, n, -, -])
    semo::DBuf<double> xph, xn;
    // LES: the velocity gradients and a scratch gradient [4][3n]; the model
    // calcCellTransport applies (0: laminar), its muT and kSGS [n], zero
    // boundary right values [3][nbf] (the first step's)
    semo::DBuf<double> les_g;
    int les_model = 0;
    semo::DBuf<double> muT, kSGS, les_zero;
    // density-based LU-SGS (linearSolver.cpp): per sweep (forward, backward)
    // the level offsets and, in level order, one slot record per cell (the
    // cell and the first four (neighbour, face) pairs it reads, 3 x int4),
    // the further pairs in an overflow list; the inverted diagonal blocks
    // [n][N*N]
    bool reordered = false;
    bool ls_ready = false;
    std::vector<int> ls_lev_ptr[2];
    semo::DBuf<int4> ls_slots[2];
    semo::DBuf<int2> ls_ovf[2];
    semo::DBuf<double> ls_dinv;
    std::vector<double> h_ufix;     // as set, in the caller's numbering
    // per-face values of the volume-fraction "function" boundary [nbf]
    // (device face order), empty until semo_set_vf_function_points
    semo::DBuf<double> b_vffun;
    // cell -> faces as sliced ELL (ascending faces; ~f for a face the cell
    // is the neighbour of; padding INT_MIN) and cell -> boundary faces (CSR)
    semo::DBuf<int> d_cf_sl, d_cf_ell, d_bf_ptr, d_bf_idx;
    semo::DBuf<double> cpk;                // packed [x,y,z,P,U,V,W,VF] per cell
    semo::DBuf<int> d_bpatch;              // boundary face -> patch
    semo::DBuf<int8_t> d_fflip;            // face flipped by the renumbering (owner = original neighbour)
    semo::DBuf<int> d_flags;               // kernel error flags
    // geometry
    semo::DBuf<double> f_area, f_nrm, f_xyz, f_wC, f_magPN, f_alpha, f_ePN, f_skew, f_pf, f_nf;
    semo::DBuf<double> f_vsn, f_cc;        // scratch of the geometry pass
    semo::DBuf<double> c_vol, c_xyz, c_lsq;

    // ---- matrix pattern ----
    semo::SellPattern pat;
    std::vector<int> h_slot_own, h_slot_nbr;   // per internal face
    semo::DBuf<int> d_slot_own, d_slot_nbr, d_tslot;
    semo::DevSell<double> Amat;               // assembled system (values per equation)
    semo::DBuf<double> Bvec;                  // [3][n] right-hand sides
    semo::DBuf<double> Xvec;                  // [3][n] solutions
    semo::DBuf<double> fscr;                  // per-face scratch [8][nif]
    int last_eq = -1;
    // lanes per cell of the one-field LSQ pass (1, 2 or 4; the results are
    // the same bits): lsq_lanes from SEMO_GPU_LSQ_LANES at context creation
    // (0: by stencil length), lsq_lanes_eff the one the mesh set-up chose
    int lsq_lanes = 0, lsq_lanes_eff = 1;

    // ---- case ----
    semo_case cs{};
    std::vector<semo_patch_bc> bcs;
    semo::DBuf<semo_patch_bc> d_bcs;
    bool has_case = false;
    int iterReal = 0;
    double courant = 0.0;                  // Courant number of the last time-step pass

    // ---- fields (device numbering) ----
    semo::DBuf<double> P, U, V, W, VF, Rho, Mu, oldP, oldU, oldV, oldW, oldVF, Qm0, Qm1, Qm2,
        srcG, srcS, kappa, alphaP;
    semo::DBuf<double> Qm3;                    // alpha two time levels back (controls.Qm[3])
    // speed of sound of the mixture (cell.var[C]): computed from the volume
    // fraction at the start of the first time step (the state SEMO loads),
    // or by the first boundary reconstruction before that; csnd_user: set by
    // the caller, never recomputed
    semo::DBuf<double> Csnd;
    bool csnd_valid = false, csnd_user = false;
    semo::DBuf<double> c_maxA;                 // largest face area per cell (static)
    semo::DBuf<double> fVFL, fVFR;             // MSTACS face volume fractions [nif]
    // curvature: level set (2 buffers), pseudo-time step, sign velocity,
    // normals, gradient scratch, static normal-divergence LSQ inverse
    semo::DBuf<double> ls, ls2, tau_ls, vel0_ls, cnrm, cgrad, kinv;
    semo::DBuf<int8_t> iface;
    semo::DBuf<double> gP, gU, gV, gW, gVF;   // [3n] interleaved
    semo::DBuf<double> mxP, mnP, mxU, mnU, mxV, mnV, mxW, mnW;
    semo::DBuf<double> bP, bUL, bUR, bVF, bMu, bRho;   // boundary faces (bUL/bUR: [3][nbf])
    // bP / bVF as the last full gradient pass used them (U-dependent boundary
    // types only), and the device flags of their change
    semo::DBuf<double> bP_g, bVF_g;
    semo::DBuf<int> d_bflag;
    semo::DBuf<double> gpp, gtmp;              // p' Green-Gauss gradient
    semo::DBuf<double> red;                    // reduction partials
    semo::DBuf<double> scal;                   // device scalars

    // ---- solver ----
    semo_solver_params sp{};
    semo::Amg amg_p;
    semo::DBuf<double> amg_key_val, amg_key_diag;   // values the pressure hierarchy was built from
    semo::DBuf<double> dinv_mom;
    // momentum / volume-fraction preconditioner (SEMO_MSOLVER_*): their AMG
    // hierarchies, the values each was built from, whether AUTO switched to
    // them, and the staging vectors of their V-cycles
    int msolver = 0;
    semo::Amg amg_m, amg_vf;
    semo::DBuf<double> amg_m_kv, amg_m_kd, amg_vf_kv, amg_vf_kd, mom_stage;
    bool amg_mom_on = false, amg_vf_on = false;
    semo::DBuf<int> d_flag1;
    semo::DBuf<double> kv64;                   // Krylov work (fp64)
    semo::DBuf<float> kv32;
    semo_solve_stats last_stats_mom{}, last_stats_p{}, last_stats_vf{};
    double* pin = nullptr;                     // pinned Krylov residual slots [4]
    hipEvent_t pev[4] = {};
    semo::AdjBufs* adj = nullptr;
    semo::Coupled* cpl = nullptr;
    // this device's copy of the no-allocation entry point's scratch
    double* noalloc_xph = nullptr;
};

namespace semo {
// ------------------------------------------------------ tagged launches --
// While the context's kernel profiler is on (V-cycle graphs are then issued
// eagerly), a HIP event pair on the context's stream brackets the launches
// of a scope; the tag accumulates their count, time and algorithmic bytes
// (the compulsory HBM traffic of one launch: every operand read once, every
// result written once).
struct ProfScope {
    semo_ctx* c; const char* name; double bytes; hipEvent_t a{};
    ProfScope(semo_ctx* c_, const char* n, double b) : c(c_), name(n), bytes(b) {
        if (c->prof.on && name) c->prof.begin(c->stream, a);
    }
    ~ProfScope() { if (c->prof.on && name) c->prof.end(c->stream, name, a, bytes); }
    ProfScope(const ProfScope&) = delete;
    ProfScope& operator=(const ProfScope&) = delete;
};
// algorithmic bytes of a cell loop over the cell's faces: the face lists and,
// per face, `per_face` bytes of face data (counted once, although an
// internal face is reached from both of its cells)
inline double cell_faces_bytes(const semo_ctx* c, double per_face) {
    return 4.0 * (double)c->h_cf_idx.size() + per_face * (double)c->nf;
}
// one kernel launch on the context's stream under a profiler tag; `bytes` is
// evaluated only while the profiler is on
#define SEMO_LAUNCH(c, tag, bytes, kern, grid, block, ...)                                  \
    do {                                                                                    \
        ::semo::ProfScope semo_ps_((c), (tag), (c)->prof.on ? (double)(bytes) : 0.0);       \
        hipLaunchKernelGGL(kern, grid, block, 0, (c)->stream, __VA_ARGS__);                 \
    } while (0)
// the same with `shmem` bytes of dynamic LDS
#define SEMO_LAUNCH_LDS(c, tag, bytes, kern, grid, block, shmem, ...)                      \
    do {                                                                                    \
        ::semo::ProfScope semo_ps_((c), (tag), (c)->prof.on ? (double)(bytes) : 0.0);       \
        hipLaunchKernelGGL(kern, grid, block, (shmem), (c)->stream, __VA_ARGS__);           \
    } while (0)

// ------------------------------------------------- Krylov convergence tests --
// Iteration k's squared residual is copied into pinned slot k % 4 behind an
// event.  ConvCheck reads it only once iteration k + 1 has been issued, so the
// device has queued work while the host waits -- except near convergence:
// when the residual history predicts that iteration k meets the tolerance,
// k is read at once (a short stall instead of a wasted iteration).
inline void pipe_init(semo_ctx* c) {
    if (c->pin) return;
    SEMO_HIP(hipHostMalloc((void**)&c->pin, 4 * sizeof(double), hipHostMallocDefault));
    for (hipEvent_t& e : c->pev) SEMO_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}
inline void pipe_push(semo_ctx* c, const double* dscalar, int k) {
    SEMO_HIP(hipMemcpyAsync(c->pin + (k & 3), dscalar, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SEMO_HIP(hipEventRecord(c->pev[k & 3], c->stream));
}
inline double pipe_pop(semo_ctx* c, int k) {
    SEMO_HIP(hipEventSynchronize(c->pev[k & 3]));
    return ((volatile double*)c->pin)[k & 3];
}
// Final test of a solve, on its true FP64 relative residual: a non-finite
// residual (NaN/Inf in the system or the iterates) or one above
// max(10 tol, 1e-12) is a failure, whatever stopped the iteration.  (The
// 1e-12 floor admits the FP64 round-off level of solves asked for tighter
// tolerances.)
inline void check_converged(double rel, double tol, const char* what) {
    if (!std::isfinite(rel))
        throw Error(SEMO_ENOCONV, std::string(what) + " solve produced a non-finite residual");
    if (rel > std::max(10.0 * tol, 1e-12))
        throw Error(SEMO_ENOCONV, std::string(what) + " solve did not converge (relative residual " +
                                      std::to_string(rel) + ")");
}

struct ConvCheck {
    semo_ctx* c;
    double lim;                 // stop when the squared residual <= lim (NaN stops too)
    int read = 0;               // last iteration whose residual was read
    double r1, r2 = -1.0;       // last two residuals read (r1 starts as |b|^2)
    ConvCheck(semo_ctx* c_, double lim_, double bb) : c(c_), lim(lim_), r1(bb) { pipe_init(c_); }
    void take(int k) { double v = pipe_pop(c, k); r2 = r1; r1 = v; read = k; }
    // after iteration `it` pushed its residual: true = stop
    bool done(int it) {
        if (read < it - 1) {
            take(it - 1);
            if (!(r1 > lim)) return true;         // it - 1 converged; `it` ran one step further
        }
        double pred = r1 * (r2 > 0.0 ? std::min(r1 / r2, 1.0) : 0.1);
        if (pred <= 4.0 * lim) {
            take(it);
            return !(r1 > lim);
        }
        return false;
    }
};

// mesh.cpp
void mesh_setup(semo_ctx* c, const semo_mesh_desc* m, int reorder);
void genrcm(int n, const int* ptr, const int* adj, int* perm, int* inv);
SellPattern sell_from_csr(int n, int ncols, const std::vector<int>& ptr, const std::vector<int>& col,
                          std::vector<int>* pos_of_csr);
// geometry.hip
void geometry_compute(semo_ctx* c);
// assembly.hip
void fields_alloc(semo_ctx* c);
void eos_transport(semo_ctx* c, const double* bR);
void recon_boundary(semo_ctx* c);
void sound_speed(semo_ctx* c);
// nearest-point VF of every boundary face for the "function" boundary
void vf_function_values(semo_ctx* c, int n, const double* pts4_host);
// calcGaussGreen of one field: sel 0 = P, 1..3 = U, V, W, 4 = VF (its patch
// types and values); faceR_b: right face values of the boundary faces [nbf]
void gauss_green_standalone(semo_ctx* c, int sel, const double* phi, const double* faceR_b, const double* gin,
                            double* gout);
void rhs_device(semo_ctx* c, int nsp, bool src, double time_step, const double* g, const double* L, const double* R,
                const double* grad, const double* cells, const double* st, double* fterm, double* res);
void update_values_device(semo_ctx* c, int nsp, const double* urf, const double* d_lim, const double* res,
                          double* cells, double* un);
void compress_vf_device(semo_ctx* c, double cfl, int iters, double* vf, const double* faceR_b, double* residuals);
void linsol_device(semo_ctx* c, int method, int nsp, double time_step, const double* st, double* res);
int lusgs_levels(semo_ctx* c, int dir);
void flux_device(semo_ctx* c, int scheme, int n, int nsp, const double* L, const double* R, const double* nrm,
                 const double* par, double* flux);
void les_device(semo_ctx* c, int model, const double* U, const double* V, const double* W, const double* Rho,
                const double* bR, double* muT, double* kSGS);
void source_gravity(semo_ctx* c);
// uvw_only: gradients of U, V, W only; P and VF gradients are kept from the
// last full pass (valid when P, VF and their boundary values are unchanged)
void grad_minmax(semo_ctx* c, bool with_p_minmax, bool uvw_only = false);
// after_vof: called by the outer iteration right after its VOF stage(s), which
// left P, U, V, W unchanged: only the alpha gradient is recomputed, unless
// alpha_grad_ready (the step's surface-tension source computed it already)
void assemble_momentum(semo_ctx* c, bool after_vof = false, bool alpha_grad_ready = false);
// one field's LSQ gradient over the vertex stencil and the boundary faces
// (faceR indexed by face - foff; null: zero boundary difference)
// gate (optional device flag): the pass runs only where *gate != 0
void lsq1(semo_ctx* c, const double* phi, const double* faceR, int foff, double* g, const int* gate = nullptr);
// boundary types whose face values depend on the owner's velocity: P
// "switch" (Mach number), VF inletOutlet (flow direction)
struct BcDeps { bool p = false, vf = false; };
inline BcDeps bc_deps(const semo_ctx* c) {
    BcDeps d;
    for (const semo_patch_bc& b : c->bcs) {
        d.p = d.p || b.p_type == SEMO_BC_SWITCH;
        d.vf = d.vf || b.vf_type == SEMO_BC_INLET_OUTLET;
    }
    return d;
}
// after_momentum: called by an outer iteration right after its momentum
// stage (P and VF unchanged since), so the P gradient of that stage is reused
void assemble_pressure(semo_ctx* c, bool after_momentum = false);
void assemble_vof(semo_ctx* c);
double time_step(semo_ctx* c, int adjust);
// true when it computed the alpha gradient (into gVF): sigma != 0
bool surface_tension(semo_ctx* c);
double vof_update(semo_ctx* c);
double momentum_update(semo_ctx* c);
double pressure_correct(semo_ctx* c);
void lsq_standalone(semo_ctx* c, const double* d_phi, const double* d_faceR, double* d_grad);
// the LSQ gradient on the static scratch (semo_least_square_noalloc)
constexpr int NOALLOC_MAX_CELLS = 1 << 21;
double* noalloc_scratch(int device);
void lsq1_noalloc(semo_ctx* c, const double* phi, const double* faceR_b, double* grad);
void minmax_standalone(semo_ctx* c, const double* d_phi, double* d_mx, double* d_mn);
void copy_d(semo_ctx* c, double* dst, const double* src, size_t n);
// `bytes` bytes at p (a multiple of 4) set to zero by a kernel on the
// context's stream, tagged "fill" (a HIP memset would run untagged)
void zero_d(semo_ctx* c, void* p, size_t bytes);
double volume_integral(semo_ctx* c, const double* phi, double* total_vol);
// solver.hip
void sell_narrow(const SellPattern& p, DBuf<int16_t>& d, hipStream_t s);
template <class T> void sell_upload(DevSell<T>& M, const SellPattern& p, const std::vector<double>& vals,
                                    const std::vector<double>& diag, hipStream_t s);
template <class T> void csr_upload(DevSell<T>& M, int n, int ncols, const std::vector<int>& ptr,
                                   const std::vector<int>& col, const std::vector<double>& vals,
                                   const std::vector<double>& diag, int lanes, hipStream_t s);
void reduction_alloc(semo_ctx* c, long n);
// amg_setup.cpp / solver.hip
void amg_build(semo_ctx* c, Amg& amg, const DevSell<double>& A);
void solve_pressure(semo_ctx* c, const double* b, double* x, semo_solve_stats* st);
// pressure hierarchy for c->Amat (build / device refresh of level 0 while
// keep_coarse(amg)), and one V-cycle of it
void pressure_amg_update(semo_ctx* c, bool (*keep_coarse)(const Amg&));
void pressure_vcycle(semo_ctx* c, const double* r64, double* z64, bool fp32);
void solve_momentum(semo_ctx* c, int nrhs, semo_solve_stats* st);
std::vector<std::vector<int>> aggregate_hierarchy(int n, const std::vector<int>& ptr, const std::vector<int>& col,
                                                  const std::vector<double>& val, int min_coarse);
// coupled.hip
struct Coupled;
void coupled_free(Coupled* p);
void assemble_coupled(semo_ctx* c);
void coupled_solve(semo_ctx* c, const double* b_dev, double* x_dev, semo_solve_stats* st);
void coupled_solve_assembled(semo_ctx* c, semo_solve_stats* st);
double coupled_update(semo_ctx* c);
void coupled_spmv(semo_ctx* c, const double* x_dev, double* y_dev);
void coupled_export(semo_ctx* c, std::vector<int>& rows, std::vector<int>& cols, std::vector<double>& blocks,
                    std::vector<double>& rhs);
// capi.cpp: the forward driver's pieces
// returns whether the step's sources left the alpha gradient in gVF
bool simple_setup(semo_ctx* c);
double simple_momentum(semo_ctx* c, semo_solve_stats* st);
double simple_pressure(semo_ctx* c, semo_solve_stats* st);
// adjoint.hip
double adjoint_run(semo_ctx* c, int n_iter, const double* wpatch, double* grad_dev);
double objective(semo_ctx* c, const double* wpatch);
void adjoint_free(AdjBufs* b);
}  // namespace semo
