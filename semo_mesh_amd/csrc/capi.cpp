// C ABI of libsemo_gpu: context, mesh, fields, per-equation entry points and
// the SIMPLE driver.  Every function catches semo::Error and returns its code;
// the message is kept for semo_ctx_last_error.
#include <algorithm>
#include <cstring>
#include <functional>

#include <cstdlib>

#include "common.hpp"

namespace semo {
void solve_coo(semo_ctx* c, const char* eq, int n, int64_t nnz, const int* rows, const int* cols,
               const double* vals, int nrhs, const double* b, double* x, semo_solve_stats* st);

// -------------------------------------------------------------- profiler ---
hipEvent_t Profiler::get() {
    if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
    hipEvent_t e;
    SEMO_HIP(hipEventCreate(&e));
    return e;
}
void Profiler::begin(hipStream_t s, hipEvent_t& a) {
    a = get();
    SEMO_HIP(hipEventRecord(a, s));
}
void Profiler::end(hipStream_t s, const char* name, hipEvent_t a, double bytes) {
    hipEvent_t b = get();
    SEMO_HIP(hipEventRecord(b, s));
    pending.push_back({name, a, b, bytes});
    if (pending.size() > 4096) flush();
}
void Profiler::flush() {
    for (auto& p : pending) {
        SEMO_HIP(hipEventSynchronize(p.b));
        float ms = 0;
        SEMO_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        KernelStat& k = stats[p.name];
        k.count++; k.ms += ms; k.bytes += p.bytes;
        pool.push_back(p.a); pool.push_back(p.b);
    }
    pending.clear();
}
Profiler::~Profiler() {
    for (auto& p : pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
    for (auto e : pool) (void)hipEventDestroy(e);
}

namespace {

void need_mesh(semo_ctx* c) { SEMO_CHECK(c->has_mesh, SEMO_ESTATE, "no mesh uploaded"); }
void need_case(semo_ctx* c) {
    need_mesh(c);
    SEMO_CHECK(c->has_case, SEMO_ESTATE, "no case set (semo_set_case)");
}

// cell arrays: caller numbering <-> device numbering
void to_dev(semo_ctx* c, DBuf<double>& d, const double* h, int comps) {
    std::vector<double> t((size_t)c->nc * comps);
    for (int i = 0; i < c->nc; ++i)
        for (int k = 0; k < comps; ++k) t[(size_t)i * comps + k] = h[(size_t)c->perm[i] * comps + k];
    d.upload(t, c->stream);
    SEMO_HIP(hipStreamSynchronize(c->stream));   // t is a temporary
}
void from_dev(semo_ctx* c, const double* dptr, double* h, int comps, size_t stride_block = 0) {
    std::vector<double> t((size_t)c->nc * comps);
    SEMO_HIP(hipMemcpyAsync(t.data(), dptr, t.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SEMO_HIP(hipStreamSynchronize(c->stream));
    for (int i = 0; i < c->nc; ++i)
        for (int k = 0; k < comps; ++k) h[(size_t)c->perm[i] * comps + k] = t[(size_t)i * comps + k];
    (void)stride_block;
}
void face_from_dev(semo_ctx* c, const DBuf<double>& d, double* h, int comps, bool odd) {
    std::vector<double> t(d.n);
    d.download(t.data(), t.size(), c->stream);
    SEMO_HIP(hipStreamSynchronize(c->stream));
    for (int k = 0; k < c->nf; ++k) {
        int o = c->fperm[k];
        double sg = (odd && c->fflip[k]) ? -1.0 : 1.0;
        for (int a = 0; a < comps; ++a) h[(size_t)o * comps + a] = sg * t[(size_t)k * comps + a];
    }
}

DBuf<double>* cell_field(semo_ctx* c, int f, int& comps) {
    comps = 1;
    switch (f) {
        case SEMO_F_P: return &c->P; case SEMO_F_U: return &c->U; case SEMO_F_V: return &c->V;
        case SEMO_F_W: return &c->W; case SEMO_F_VF: return &c->VF; case SEMO_F_RHO: return &c->Rho;
        case SEMO_F_MU: return &c->Mu; case SEMO_F_OLDU: return &c->oldU; case SEMO_F_OLDV: return &c->oldV;
        case SEMO_F_OLDW: return &c->oldW; case SEMO_F_QM0: return &c->Qm0; case SEMO_F_QM1: return &c->Qm1;
        case SEMO_F_QM2: return &c->Qm2;
        case SEMO_F_MAXP: return &c->mxP; case SEMO_F_MINP: return &c->mnP; case SEMO_F_MAXU: return &c->mxU;
        case SEMO_F_MINU: return &c->mnU; case SEMO_F_MAXV: return &c->mxV; case SEMO_F_MINV: return &c->mnV;
        case SEMO_F_MAXW: return &c->mxW; case SEMO_F_MINW: return &c->mnW;
        case SEMO_F_ALPHA_P: return &c->alphaP;
        case SEMO_F_QM3: return &c->Qm3; case SEMO_F_OLDVF: return &c->oldVF;
        case SEMO_F_KAPPA: return &c->kappa; case SEMO_F_LS: return &c->ls;
        case SEMO_F_C: return &c->Csnd;
        case SEMO_F_MUT: return &c->muT; case SEMO_F_KSGS: return &c->kSGS;
        case SEMO_F_GRADP: comps = 3; return &c->gP; case SEMO_F_GRADU: comps = 3; return &c->gU;
        case SEMO_F_GRADV: comps = 3; return &c->gV; case SEMO_F_GRADW: comps = 3; return &c->gW;
        case SEMO_F_GRADVF: comps = 3; return &c->gVF;
        case SEMO_G_CELL_VOLUME: return &c->c_vol;
        case SEMO_G_CELL_XYZ: comps = 3; return &c->c_xyz;
        case SEMO_G_CELL_LSQ: comps = 6; return &c->c_lsq;
        default: return nullptr;
    }
}

int guard(semo_ctx* c, const std::function<void()>& fn) {
    try {
        fn();
        if (c) c->err.clear();
        return SEMO_OK;
    } catch (const Error& e) {
        if (c) c->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        if (c) c->err = e.what();
        return SEMO_EINVAL;
    }
}

void step_setup(semo_ctx* c) {
    // the speed of sound of the loaded state (calcCellEOSVF, main/
    // incompressible/pressureBased.cpp:79), then pressureBased.cpp:17-69:
    // first-step EOS/transport, time levels
    if (c->iterReal == 0 && !c->csnd_user) sound_speed(c);
    if (c->iterReal == 0) eos_transport(c, nullptr);
    const DBuf<double>& q0 = c->iterReal == 0 ? c->U : c->oldU;
    const DBuf<double>& q1 = c->iterReal == 0 ? c->V : c->oldV;
    const DBuf<double>& q2 = c->iterReal == 0 ? c->W : c->oldW;
    const DBuf<double>& q3 = c->iterReal == 0 ? c->VF : c->oldVF;
    copy_d(c, c->Qm0.p, q0.p, c->nc); copy_d(c, c->Qm1.p, q1.p, c->nc); copy_d(c, c->Qm2.p, q2.p, c->nc);
    copy_d(c, c->Qm3.p, q3.p, c->nc);
    copy_d(c, c->oldP.p, c->P.p, c->nc); copy_d(c, c->oldU.p, c->U.p, c->nc);
    copy_d(c, c->oldV.p, c->V.p, c->nc); copy_d(c, c->oldW.p, c->W.p, c->nc);
    copy_d(c, c->oldVF.p, c->VF.p, c->nc);
}

// after the VOF loop: face values and the body-force source of the step
// (pressureBased.cpp:128-130)
bool step_sources(semo_ctx* c) {
    recon_boundary(c);
    source_gravity(c);
    return surface_tension(c);
}

// one VOF iteration (pressureBased.cpp:104-127)
double volfrac_eq(semo_ctx* c, semo_solve_stats* st) {
    recon_boundary(c);
    assemble_vof(c);
    solve_momentum(c, 1, st);
    double nv = vof_update(c);
    eos_transport(c, c->bUR.p);
    return nv;
}

double momentum_eq(semo_ctx* c, semo_solve_stats* st, bool after_vof = false, bool alpha_grad_ready = false) {
    recon_boundary(c);
    assemble_momentum(c, after_vof, alpha_grad_ready);
    solve_momentum(c, 3, st);
    return momentum_update(c);
}

double pressure_eq(semo_ctx* c, semo_solve_stats* st, bool after_momentum = false) {
    recon_boundary(c);
    assemble_pressure(c, after_momentum);
    solve_pressure(c, c->Bvec.p, c->Xvec.p, st);
    return pressure_correct(c);
}

}  // namespace

// one time step's set-up and its outer iterations, for the adjoint driver
bool simple_setup(semo_ctx* c) {
    step_setup(c);
    c->courant = time_step(c, c->cs.adjust_dt);
    for (int k = 0; k < c->cs.n_vof; ++k) volfrac_eq(c, &c->last_stats_vf);
    return step_sources(c);
}
double simple_momentum(semo_ctx* c, semo_solve_stats* st) { return momentum_eq(c, st); }
double simple_pressure(semo_ctx* c, semo_solve_stats* st) { return pressure_eq(c, st, true); }

}  // namespace semo

using namespace semo;

extern "C" {

int semo_ctx_create(int device, void* stream, semo_ctx** out) {
    if (!out) return SEMO_EINVAL;
    *out = nullptr;
    semo_ctx* c = new semo_ctx;
    int rc = guard(c, [&] {
        int nd = 0;
        SEMO_HIP(hipGetDeviceCount(&nd));
        SEMO_CHECK(device >= 0 && device < nd, SEMO_EINVAL, "no such device");
        SEMO_HIP(hipSetDevice(device));
        c->device = device;
        if (stream) c->stream = (hipStream_t)stream;
        else { SEMO_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)); c->own_stream = true; }
        const char* gv = std::getenv("SEMO_GPU_GRAPHS");
        c->graphs = !(gv && gv[0] == '0');
        const char* lv = std::getenv("SEMO_GPU_LSQ_LANES");
        if (lv) c->lsq_lanes = std::atoi(lv);
        if (c->lsq_lanes != 1 && c->lsq_lanes != 2 && c->lsq_lanes != 4) c->lsq_lanes = 0;
        c->noalloc_xph = noalloc_scratch(device);
        c->sp.precision = SEMO_PREC_FP64; c->sp.tol_p = 1e-8; c->sp.tol_u = 1e-8;
        c->sp.maxit_p = 500; c->sp.maxit_u = 500; c->sp.cheb_degree = 1; c->sp.coarse_size = 1000;
    });
    if (rc != SEMO_OK) { delete c; return rc; }
    *out = c;
    return SEMO_OK;
}

int semo_ctx_destroy(semo_ctx* c) {
    if (!c) return SEMO_EINVAL;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    bool own = c->own_stream;
    hipStream_t s = c->stream;
    adjoint_free(c->adj);
    c->adj = nullptr;
    coupled_free(c->cpl);
    c->cpl = nullptr;
    if (c->pin) (void)hipHostFree(c->pin);
    for (hipEvent_t e : c->pev) if (e) (void)hipEventDestroy(e);
    delete c;
    if (own) (void)hipStreamDestroy(s);
    return SEMO_OK;
}

const char* semo_ctx_last_error(const semo_ctx* c) { return c ? c->err.c_str() : "null context"; }

int64_t semo_ctx_device_bytes(const semo_ctx* c) {
    if (!c) return 0;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return -1;
    return (int64_t)(total_b - free_b);
}

int semo_mesh_upload(semo_ctx* c, const semo_mesh_desc* m, int reorder) {
    if (!c) return SEMO_EINVAL;
    return guard(c, [&] {
        SEMO_HIP(hipSetDevice(c->device));
        c->has_mesh = false; c->has_case = false; c->iterReal = 0;
        c->amg_p.reset();
        c->b_ufix.release(); c->h_ufix.clear();
        c->b_vffun.release(); c->Csnd.release(); c->csnd_valid = c->csnd_user = false;
        coupled_free(c->cpl);          // mesh-bound block pattern and work arrays
        c->cpl = nullptr;
        adjoint_free(c->adj);
        c->adj = nullptr;
        mesh_setup(c, m, reorder);
        geometry_compute(c);
        fields_alloc(c);
        SEMO_HIP(hipStreamSynchronize(c->stream));
    });
}

int semo_mesh_get_perm(const semo_ctx* c, int32_t* perm, int32_t* inv) {
    if (!c || !c->has_mesh) return SEMO_ESTATE;
    if (perm) std::copy(c->perm.begin(), c->perm.end(), perm);
    if (inv) std::copy(c->inv.begin(), c->inv.end(), inv);
    return SEMO_OK;
}

int semo_mesh_get_face_map(const semo_ctx* c, int32_t* fmap, int8_t* flip) {
    if (!c || !c->has_mesh) return SEMO_ESTATE;
    if (fmap) std::copy(c->fperm.begin(), c->fperm.end(), fmap);
    if (flip) std::copy(c->fflip.begin(), c->fflip.end(), flip);
    return SEMO_OK;
}

int semo_mesh_get_connectivity(const semo_ctx* c, int32_t* owner, int32_t* nbr) {
    if (!c || !c->has_mesh) return SEMO_ESTATE;
    if (owner) std::copy(c->h_owner.begin(), c->h_owner.end(), owner);
    if (nbr) std::copy(c->h_nbr.begin(), c->h_nbr.end(), nbr);
    return SEMO_OK;
}

int semo_rcm(int32_t n, const int32_t* ptr, const int32_t* adj, int32_t* perm, int32_t* inv) {
    if (n <= 0 || !ptr || !adj || !perm || !inv) return SEMO_EINVAL;
    for (int i = 0; i < n; ++i) {
        if (ptr[i + 1] < ptr[i]) return SEMO_EINVAL;
        for (int k = ptr[i]; k < ptr[i + 1]; ++k) if (adj[k] < 0 || adj[k] >= n) return SEMO_EINVAL;
    }
    semo::genrcm(n, ptr, adj, perm, inv);
    return SEMO_OK;
}

int semo_set_case(semo_ctx* c, const semo_case* cs, const semo_patch_bc* bcs) {
    if (!c || !cs) return SEMO_EINVAL;
    return guard(c, [&] {
        need_mesh(c);
        SEMO_CHECK(cs->dt > 0 && cs->rho[0] > 0 && cs->rho[1] > 0 && cs->mu[0] > 0 && cs->mu[1] > 0,
                   SEMO_EINVAL, "case: dt, rho and mu must be positive");
        SEMO_CHECK(cs->sigma >= 0.0, SEMO_EINVAL, "case: sigma must be non-negative");
        SEMO_CHECK(cs->n_vof >= 0 && cs->urf_vf > 0.0, SEMO_EINVAL, "case: n_vof >= 0 and urf_vf > 0");
        SEMO_CHECK(!cs->adjust_dt || (cs->max_co > 0.0 && cs->max_dt > 0.0), SEMO_EINVAL,
                   "case: maxCo and maxTimeStep must be positive");
        SEMO_CHECK(bcs || c->npatch == 0, SEMO_EINVAL, "case: null boundary-condition array");
        for (int b = 0; b < c->npatch; ++b) {
            const semo_patch_bc& p = bcs[b];
            SEMO_CHECK(p.p_type == SEMO_BC_ZERO_GRADIENT || p.p_type == SEMO_BC_FIXED_VALUE ||
                           p.p_type == SEMO_BC_SWITCH, SEMO_EINVAL_BC,
                       "pressure boundary type not defined");
            SEMO_CHECK(p.u_type >= SEMO_BC_ZERO_GRADIENT && p.u_type <= SEMO_BC_INLET_OUTLET, SEMO_EINVAL_BC,
                       "velocity boundary type not defined");
            SEMO_CHECK(p.vf_type == SEMO_BC_ZERO_GRADIENT || p.vf_type == SEMO_BC_FIXED_VALUE ||
                           p.vf_type == SEMO_BC_INLET_OUTLET || p.vf_type == SEMO_BC_FUNCTION, SEMO_EINVAL_BC,
                       "volume-fraction boundary type not defined");
        }
        c->cs = *cs;
        c->bcs.assign(bcs, bcs + c->npatch);
        c->d_bcs.upload(c->bcs, c->stream);
        SEMO_HIP(hipStreamSynchronize(c->stream));
        c->has_case = true;
        c->amg_p.built = false;
    });
}

int semo_set_vf_function_points(semo_ctx* c, int32_t n, const double* pts) {
    if (!c || n < 0 || (n > 0 && !pts)) return SEMO_EINVAL;
    return guard(c, [&] {
        need_mesh(c);
        for (int64_t k = 0; k < 4 * (int64_t)n; ++k)
            SEMO_CHECK(std::isfinite(pts[k]), SEMO_EINVAL, "non-finite volume-fraction point");
        vf_function_values(c, n, pts);
    });
}

int semo_set_solver(semo_ctx* c, const semo_solver_params* p) {
    if (!c || !p) return SEMO_EINVAL;
    return guard(c, [&] {
        SEMO_CHECK(p->tol_p > 0 && p->tol_u > 0 && p->maxit_p > 0 && p->maxit_u > 0, SEMO_EINVAL, "solver params");
        SEMO_CHECK(p->precision == SEMO_PREC_FP64 || p->precision == SEMO_PREC_MIXED, SEMO_EINVAL, "precision");
        int coarse = p->coarse_size > 0 ? p->coarse_size : 1000;      // 0: the default
        bool rebuild = coarse != c->sp.coarse_size;
        c->sp = *p;
        if (c->sp.cheb_degree <= 0) c->sp.cheb_degree = 1;
        c->sp.coarse_size = coarse;
        if (rebuild) c->amg_p.built = false;
    });
}

int semo_field_set(semo_ctx* c, int f, const double* h, int64_t n) {
    if (!c || (!h && !(f == SEMO_F_FACE_UFIX && n == 0))) return SEMO_EINVAL;
    return guard(c, [&] {
        need_mesh(c);
        if (f == SEMO_F_FACE_UFIX) {      // [3][nbf], caller's boundary-face order
            if (n == 0) { c->b_ufix.release(); c->h_ufix.clear(); return; }
            size_t nb = (size_t)c->nbf;
            SEMO_CHECK(n == (int64_t)(3 * nb), SEMO_EINVAL, "field size");
            for (int64_t q = 0; q < n; ++q)
                SEMO_CHECK(std::isfinite(h[q]), SEMO_EINVAL, "non-finite boundary velocity");
            c->h_ufix.assign(h, h + n);
            std::vector<double> d(3 * nb);
            for (size_t j = 0; j < nb; ++j) {
                size_t o = (size_t)c->fperm[c->nif + j] - (size_t)c->nif;   // caller's boundary-face index
                for (int a = 0; a < 3; ++a) d[a * nb + j] = h[a * nb + o];
            }
            c->b_ufix.upload(d, c->stream);
            SEMO_HIP(hipStreamSynchronize(c->stream));
            return;
        }
        int comps;
        DBuf<double>* d = cell_field(c, f, comps);
        if (f >= SEMO_F_SRCS0 && f <= SEMO_F_SRCS2) {   // one component of the surface-tension source
            SEMO_CHECK(n == c->nc, SEMO_EINVAL, "field size");
            DBuf<double> t;
            to_dev(c, t, h, 1);
            copy_d(c, c->srcS.p + (size_t)(f - SEMO_F_SRCS0) * c->nc, t.p, c->nc);
            SEMO_HIP(hipStreamSynchronize(c->stream));
            return;
        }
        if (f >= SEMO_F_SRCG0 && f <= SEMO_F_SRCG2) {   // one component of the gravity source
            SEMO_CHECK(n == c->nc, SEMO_EINVAL, "field size");
            DBuf<double> t;
            to_dev(c, t, h, 1);
            copy_d(c, c->srcG.p + (size_t)(f - SEMO_F_SRCG0) * c->nc, t.p, c->nc);
            SEMO_HIP(hipStreamSynchronize(c->stream));
            return;
        }
        SEMO_CHECK(d && f < SEMO_G_CELL_VOLUME && f != SEMO_F_MUT && f != SEMO_F_KSGS, SEMO_EINVAL,
                   "field not settable");
        SEMO_CHECK(n == (int64_t)c->nc * comps, SEMO_EINVAL, "field size");
        to_dev(c, *d, h, comps);
        SEMO_HIP(hipStreamSynchronize(c->stream));
        if (f == SEMO_F_C) c->csnd_valid = c->csnd_user = true;
    });
}

int semo_field_get(semo_ctx* c, int f, double* h, int64_t n) {
    if (!c || !h) return SEMO_EINVAL;
    return guard(c, [&] {
        need_mesh(c);
        if (f == SEMO_F_FACE_UFIX) {
            SEMO_CHECK(!c->h_ufix.empty(), SEMO_ESTATE, "no per-face boundary velocity set");
            SEMO_CHECK(n == (int64_t)c->h_ufix.size(), SEMO_EINVAL, "field size");
            std::copy(c->h_ufix.begin(), c->h_ufix.end(), h);
            return;
        }
        if (f >= SEMO_F_SRCG0 && f <= SEMO_F_SRCG2) {
            SEMO_CHECK(n == c->nc, SEMO_EINVAL, "field size");
            from_dev(c, c->srcG.p + (size_t)(f - SEMO_F_SRCG0) * c->nc, h, 1);
            return;
        }
        if (f >= SEMO_F_SRCS0 && f <= SEMO_F_SRCS2) {
            SEMO_CHECK(n == c->nc, SEMO_EINVAL, "field size");
            from_dev(c, c->srcS.p + (size_t)(f - SEMO_F_SRCS0) * c->nc, h, 1);
            return;
        }
        if (f >= SEMO_F_X0 && f <= SEMO_F_X2) {
            SEMO_CHECK(n == c->nc, SEMO_EINVAL, "field size");
            from_dev(c, c->Xvec.p + (size_t)(f - SEMO_F_X0) * c->nc, h, 1);
            return;
        }
        if (f >= SEMO_F_B0 && f <= SEMO_F_B2) {
            SEMO_CHECK(n == c->nc, SEMO_EINVAL, "field size");
            from_dev(c, c->Bvec.p + (size_t)(f - SEMO_F_B0) * c->nc, h, 1);
            return;
        }
        if (f == SEMO_F_FACE_VFL || f == SEMO_F_FACE_VFR) {
            SEMO_CHECK(n == c->nf, SEMO_EINVAL, "field size");
            std::vector<double> l(std::max(c->nif, 1)), r(std::max(c->nif, 1)), b(std::max(c->nbf, 1));
            c->fVFL.download(l.data(), c->nif, c->stream);
            c->fVFR.download(r.data(), c->nif, c->stream);
            c->bVF.download(b.data(), c->nbf, c->stream);
            SEMO_HIP(hipStreamSynchronize(c->stream));
            bool left = f == SEMO_F_FACE_VFL;
            for (int k = 0; k < c->nf; ++k) {
                double v;
                if (k < c->nif) v = (left != (c->fflip[k] != 0)) ? l[k] : r[k];   // a flipped face swaps sides
                else v = b[k - c->nif];
                h[c->fperm[k]] = v;
            }
            return;
        }
        if (f >= SEMO_G_FACE_AREA) {
            const DBuf<double>* d = nullptr; int comps = 1; bool odd = false;
            switch (f) {
                case SEMO_G_FACE_AREA: d = &c->f_area; break;
                case SEMO_G_FACE_NORMAL: d = &c->f_nrm; comps = 3; odd = true; break;
                case SEMO_G_FACE_XYZ: d = &c->f_xyz; comps = 3; break;
                case SEMO_G_FACE_WC: d = &c->f_wC; break;
                case SEMO_G_FACE_MAGPN: d = &c->f_magPN; break;
                case SEMO_G_FACE_ALPHAF: d = &c->f_alpha; break;
                case SEMO_G_FACE_EPN: d = &c->f_ePN; comps = 3; odd = true; break;
                case SEMO_G_FACE_SKEW: d = &c->f_skew; comps = 3; break;
                default: SEMO_CHECK(false, SEMO_EINVAL, "unknown field");
            }
            SEMO_CHECK(n == (int64_t)c->nf * comps, SEMO_EINVAL, "field size");
            face_from_dev(c, *d, h, comps, odd);
            if (f == SEMO_G_FACE_WC)   // orientation of flipped faces (exact complement)
                for (int k = 0; k < c->nif; ++k) if (c->fflip[k]) h[c->fperm[k]] = 1.0 - h[c->fperm[k]];
            return;
        }
        int comps;
        DBuf<double>* d = cell_field(c, f, comps);
        SEMO_CHECK(d && d->n >= (size_t)c->nc * comps, SEMO_EINVAL, "unknown or empty field");
        SEMO_CHECK(n == (int64_t)c->nc * comps, SEMO_EINVAL, "field size");
        from_dev(c, d->p, h, comps);
    });
}

int semo_calc_momentum_eqs(semo_ctx* c, double* norm, semo_solve_stats* st) {
    if (!c) return SEMO_EINVAL;
    return guard(c, [&] {
        need_case(c);
        double nv = momentum_eq(c, &c->last_stats_mom);
        if (st) *st = c->last_stats_mom;
        if (norm) *norm = nv;
    });
}

int semo_calc_pressure_eq(semo_ctx* c, double* norm, semo_solve_stats* st) {
    if (!c) return SEMO_EINVAL;
    return guard(c, [&] {
        need_case(c);
        double nv = pressure_eq(c, &c->last_stats_p);
        if (st) *st = c->last_stats_p;
        if (norm) *norm = nv;
    });
}

int semo_calc_volfrac_eq(semo_ctx* c, double* norm, semo_solve_stats* st) {
    if (!c) return SEMO_EINVAL;
    return guard(c, [&] {
        need_case(c);
        double nv = volfrac_eq(c, &c->last_stats_vf);
        if (st) *st = c->last_stats_vf;
        if (norm) *norm = nv;
    });
}

int semo_calc_source_gravity(semo_ctx* c) {
    if (!c) return SEMO_EINVAL;
    return guard(c, [&] {
        need_case(c);
        recon_boundary(c);
        source_gravity(c);
        SEMO_HIP(hipStreamSynchronize(c->stream));
    });
}

int semo_calc_source_surface_tension(semo_ctx* c) {
    if (!c) return SEMO_EINVAL;
    return guard(c, [&] {
        need_case(c);
        recon_boundary(c);
        surface_tension(c);
        SEMO_HIP(hipStreamSynchronize(c->stream));
    });
}

int semo_calc_time_step(semo_ctx* c, int adjust, double* dt, double* courant) {
    if (!c) return SEMO_EINVAL;
    return guard(c, [&] {
        need_case(c);
        double co = time_step(c, adjust);
        c->courant = co;
        if (dt) *dt = c->cs.dt;
        if (courant) *courant = co;
    });
}

int semo_incompressible_pressure_based(semo_ctx* c, int ncorr, double* norms) {
    if (!c || ncorr < 1) return SEMO_EINVAL;
    return guard(c, [&] {
        need_case(c);
        // pressureBased.cpp:17-130: time levels, dt, VOF, sources
        bool alpha_grad = simple_setup(c);
        for (int it = 0; it < ncorr; ++it) {
            bool first = it == 0 && c->cs.n_vof > 0;
            double n0 = momentum_eq(c, &c->last_stats_mom, first, first && alpha_grad);
            double n1 = pressure_eq(c, &c->last_stats_p, true);
            if (norms) { norms[2 * it] = n0; norms[2 * it + 1] = n1; }
        }
        c->iterReal++;
    });
}

int semo_adjoint_alpha_p(semo_ctx* c, int n_iter, const double* patch_weight, double* J, double* dJ_dalpha) {
    if (!c || n_iter < 1 || !patch_weight) return SEMO_EINVAL;
    return guard(c, [&] {
        need_case(c);
        DBuf<double> g;
        g.alloc(c->nc);
        double j = adjoint_run(c, n_iter, patch_weight, g.p);
        if (J) *J = j;
        if (dJ_dalpha) from_dev(c, g.p, dJ_dalpha, 1);
    });
}

// ---- block-coupled (BSR4) system ------------------------------------------------
namespace {
// var-major [4][n] caller numbering <-> interleaved [n][4] device numbering
void blk_to_dev(semo_ctx* c, const double* h, DBuf<double>& d) {
    size_t n = c->nc;
    std::vector<double> t(4 * n);
    for (size_t i = 0; i < n; ++i)
        for (int q = 0; q < 4; ++q) t[4 * i + q] = h[q * n + c->perm[i]];
    d.upload(t, c->stream);
}
void blk_from_dev(semo_ctx* c, const double* d, double* h) {
    size_t n = c->nc;
    std::vector<double> t(4 * n);
    SEMO_HIP(hipMemcpyAsync(t.data(), d, t.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SEMO_HIP(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < n; ++i)
        for (int q = 0; q < 4; ++q) h[q * n + c->perm[i]] = t[4 * i + q];
}
}  // namespace

int semo_calc_coupled_eq(semo_ctx* c, double* norm, semo_solve_stats* st) {
    if (!c) return SEMO_EINVAL;
    return guard(c, [&] {
        need_case(c);
        recon_boundary(c);
        assemble_coupled(c);
        semo_solve_stats s{};
        coupled_solve_assembled(c, &s);
        if (st) *st = s;
        double nv = coupled_update(c);
        if (norm) *norm = nv;
    });
}

int semo_coupled_pressure_based(semo_ctx* c, int n_iter, double* norms) {
    if (!c || n_iter < 1) return SEMO_EINVAL;
    return guard(c, [&] {
        need_case(c);
        simple_setup(c);
        int sum = 0, mx = 0;
        for (int it = 0; it < n_iter; ++it) {
            recon_boundary(c);
            assemble_coupled(c);
            coupled_solve_assembled(c, &c->last_stats_mom);
            sum += c->last_stats_mom.iters[0];
            mx = std::max(mx, c->last_stats_mom.iters[0]);
            double nv = coupled_update(c);
            if (norms) norms[it] = nv;
        }
        c->last_stats_mom.iters[1] = sum;
        c->last_stats_mom.iters[2] = mx;
        c->iterReal++;
    });
}

int semo_coupled_matrix_get(semo_ctx* c, int64_t* nnzb, int32_t* rows, int32_t* cols, double* blocks,
                            double* rhs) {
    if (!c || !nnzb) return SEMO_EINVAL;
    return guard(c, [&] {
        need_mesh(c);
        std::vector<int> r, q;
        std::vector<double> b, h;
        coupled_export(c, r, q, b, h);
        int64_t nz = (int64_t)r.size();
        if (rows || cols || blocks || rhs)
            SEMO_CHECK(*nnzb >= nz, SEMO_EINVAL, "coupled matrix: buffers smaller than nnzb");
        *nnzb = nz;
        size_t n = c->nc;
        for (int64_t k = 0; k < nz; ++k) {
            if (rows) rows[k] = c->perm[r[k]];
            if (cols) cols[k] = c->perm[q[k]];
        }
        if (blocks) std::copy(b.begin(), b.end(), blocks);
        if (rhs)
            for (size_t i = 0; i < n; ++i)
                for (int v = 0; v < 4; ++v) rhs[v * n + c->perm[i]] = h[4 * i + v];
    });
}

int semo_coupled_spmv(semo_ctx* c, const double* x, double* y) {
    if (!c || !x || !y) return SEMO_EINVAL;
    return guard(c, [&] {
        need_mesh(c);
        DBuf<double> dx, dy;
        blk_to_dev(c, x, dx);
        dy.alloc(4 * (size_t)c->nc);
        coupled_spmv(c, dx.p, dy.p);
        blk_from_dev(c, dy.p, y);
    });
}

int semo_coupled_solve(semo_ctx* c, const double* b, double* x, semo_solve_stats* st) {
    if (!c || !b || !x) return SEMO_EINVAL;
    return guard(c, [&] {
        need_mesh(c);
        DBuf<double> db, dx;
        blk_to_dev(c, b, db);
        dx.alloc(4 * (size_t)c->nc);
        coupled_solve(c, db.p, dx.p, st);
        blk_from_dev(c, dx.p, x);
    });
}

int semo_objective_pressure(semo_ctx* c, const double* patch_weight, double* J) {
    if (!c || !patch_weight || !J) return SEMO_EINVAL;
    return guard(c, [&] {
        need_case(c);
        *J = objective(c, patch_weight);
    });
}

int semo_volume_integral(semo_ctx* c, int f, double* integral, double* volume) {
    if (!c || !integral) return SEMO_EINVAL;
    return guard(c, [&] {
        need_mesh(c);
        int comps;
        DBuf<double>* d = cell_field(c, f, comps);
        SEMO_CHECK(d && comps == 1 && d->n >= (size_t)c->nc, SEMO_EINVAL, "not a one-component cell field");
        reduction_alloc(c, (long)c->nc);
        *integral = volume_integral(c, d->p, volume);
    });
}

int semo_get_solve_stats(const semo_ctx* c, int eq, semo_solve_stats* st) {
    if (!c || !st || eq < 0 || eq > 2) return SEMO_EINVAL;
    *st = eq == 0 ? c->last_stats_mom : eq == 1 ? c->last_stats_p : c->last_stats_vf;
    return SEMO_OK;
}

int semo_assemble(semo_ctx* c, int eq) {
    if (!c) return SEMO_EINVAL;
    return guard(c, [&] {
        need_case(c);
        SEMO_CHECK(eq >= 0 && eq <= 3, SEMO_EINVAL,
                   "eq must be 0 (momentum), 1 (pressure), 2 (volume fraction) or 3 (coupled)");
        recon_boundary(c);
        if (eq == 0) assemble_momentum(c);
        else if (eq == 1) assemble_pressure(c);
        else if (eq == 2) assemble_vof(c);
        else assemble_coupled(c);
        SEMO_HIP(hipStreamSynchronize(c->stream));
    });
}

int semo_matrix_get(semo_ctx* c, int64_t* nnz, int32_t* rows, int32_t* cols, double* vals, double* b) {
    if (!c || !nnz) return SEMO_EINVAL;
    return guard(c, [&] {
        need_mesh(c);
        SEMO_CHECK(c->last_eq >= 0, SEMO_ESTATE, "nothing assembled");
        std::vector<double> val(c->Amat.val.n), diag(c->nc);
        c->Amat.val.download(val.data(), val.size(), c->stream);
        c->Amat.diag.download(diag.data(), diag.size(), c->stream);
        SEMO_HIP(hipStreamSynchronize(c->stream));
        int64_t m = c->nc;
        const SellPattern& p = c->pat;
        // real slots only; entries dropped by the reference (exact zeros) skipped
        std::vector<int32_t> r, cc; std::vector<double> v;
        for (int i = 0; i < c->nc; ++i) { r.push_back(c->perm[i]); cc.push_back(c->perm[i]); v.push_back(diag[i]); }
        for (int f = 0; f < c->nif; ++f) {
            int so = c->h_slot_own[f], sn = c->h_slot_nbr[f];
            int o = c->h_owner[f], n = c->h_nbr[f];
            bool drop = c->last_eq != 1;   // momentum and VOF drop |v| < DBL_MIN, pressure keeps all
            if (!(drop && val[so] == 0.0)) { r.push_back(c->perm[o]); cc.push_back(c->perm[n]); v.push_back(val[so]); }
            if (!(drop && val[sn] == 0.0)) { r.push_back(c->perm[n]); cc.push_back(c->perm[o]); v.push_back(val[sn]); }
        }
        (void)p;
        m = (int64_t)r.size();
        if (rows) {
            SEMO_CHECK(*nnz >= m, SEMO_EINVAL, "nnz buffer too small");
            std::copy(r.begin(), r.end(), rows);
            std::copy(cc.begin(), cc.end(), cols);
            std::copy(v.begin(), v.end(), vals);
            if (b) {
                int nr = c->last_eq == 0 ? 3 : 1;
                for (int k = 0; k < nr; ++k) from_dev(c, c->Bvec.p + (size_t)k * c->nc, b + (size_t)k * c->nc, 1);
            }
        }
        *nnz = m;
    });
}

int semo_least_square(semo_ctx* c, const double* phi, const double* face_r, double* grad) {
    if (!c || !phi || !face_r || !grad) return SEMO_EINVAL;
    return guard(c, [&] {
        need_mesh(c);
        DBuf<double> dphi, dfr, dg;
        to_dev(c, dphi, phi, 1);
        std::vector<double> fr(c->nf);
        for (int k = 0; k < c->nf; ++k) fr[k] = face_r[c->fperm[k]];
        dfr.upload(fr, c->stream);
        dg.alloc(3 * (size_t)c->nc);
        lsq_standalone(c, dphi.p, dfr.p, dg.p);
        from_dev(c, dg.p, grad, 3);
    });
}

int semo_gauss_green(semo_ctx* c, int field, const double* phi, const double* face_r, double* grad) {
    if (!c || !phi || !face_r || !grad) return SEMO_EINVAL;
    return guard(c, [&] {
        need_case(c);
        int sel = field == SEMO_F_P ? 0 : field == SEMO_F_U ? 1 : field == SEMO_F_V ? 2 : field == SEMO_F_W ? 3
                : field == SEMO_F_VF ? 4 : -1;
        SEMO_CHECK(sel >= 0, SEMO_EINVAL, "Green-Gauss: field must be P, U, V, W or VF");
        DBuf<double> dphi, dfr, gin, gout;
        to_dev(c, dphi, phi, 1);
        to_dev(c, gin, grad, 3);
        std::vector<double> fr(std::max(c->nbf, 1));
        for (int j = 0; j < c->nbf; ++j) fr[j] = face_r[c->fperm[c->nif + j]];
        dfr.upload(fr, c->stream);
        gout.alloc(3 * (size_t)c->nc);
        gauss_green_standalone(c, sel, dphi.p, dfr.p, gin.p, gout.p);
        from_dev(c, gout.p, grad, 3);
    });
}

int semo_flux(semo_ctx* c, int scheme, int64_t n, int nsp, const double* left, const double* right,
              const double* normal, const double* par, double* flux) {
    if (!c || n < 0 || !left || !right || !normal || !flux) return SEMO_EINVAL;
    return guard(c, [&] {
        SEMO_CHECK(scheme >= SEMO_FLUX_HAUS && scheme <= SEMO_FLUX_ROEM_N, SEMO_EINVAL, "unknown flux scheme");
        SEMO_CHECK(nsp >= 1 && nsp <= 8, SEMO_EINVAL, "flux: 1 <= nsp <= 8");
        SEMO_CHECK(n <= INT32_MAX / 64, SEMO_EINVAL, "flux: too many faces");
        bool needs_par = scheme == SEMO_FLUX_AUSMPWP_N || scheme == SEMO_FLUX_ROEM_N;
        SEMO_CHECK(par || !needs_par, SEMO_EINVAL, "flux: AUSMPWP_N and RoeM_N need par");
        if (n == 0) return;
        size_t w = 8 + (size_t)nsp;
        DBuf<double> dL, dR, dN, dP, dF;
        dL.upload(left, n * w, c->stream);
        dR.upload(right, n * w, c->stream);
        dN.upload(normal, 3 * (size_t)n, c->stream);
        if (needs_par) dP.upload(par, 7 * (size_t)n, c->stream);
        dF.alloc((size_t)n * (4 + nsp));
        flux_device(c, scheme, (int)n, nsp, dL.p, dR.p, dN.p, dP.p, dF.p);
        dF.download(flux, dF.n, c->stream);
        SEMO_HIP(hipStreamSynchronize(c->stream));
    });
}

// [n][comps] records, caller numbering -> device order, stored field-major
// ([comps][n]) for coalesced loads
static void face_to_dev_soa(semo_ctx* c, DBuf<double>& d, const double* h, int comps) {
    const size_t n = (size_t)c->nf;
    std::vector<double> t(n * comps);
    for (size_t k = 0; k < n; ++k)
        for (int a = 0; a < comps; ++a) t[(size_t)a * n + k] = h[(size_t)c->fperm[k] * comps + a];
    d.upload(t, c->stream);
    SEMO_HIP(hipStreamSynchronize(c->stream));
}
static void cell_to_dev_soa(semo_ctx* c, DBuf<double>& d, const double* h, int comps) {
    const size_t n = (size_t)c->nc;
    std::vector<double> t(n * comps);
    for (size_t i = 0; i < n; ++i)
        for (int a = 0; a < comps; ++a) t[(size_t)a * n + i] = h[(size_t)c->perm[i] * comps + a];
    d.upload(t, c->stream);
    SEMO_HIP(hipStreamSynchronize(c->stream));
}

static int rhs_common(semo_ctx* c, int nsp, bool src, double time_step, const double* g, const double* L,
                      const double* R, const double* grad, const double* cells, const double* st, double* res) {
    return guard(c, [&] {
        need_mesh(c);
        SEMO_CHECK(nsp >= 2 && nsp <= 8, SEMO_EINVAL, "calcFluxes / calcRHS: 2 <= nsp <= 8");
        SEMO_CHECK(!src || (std::isfinite(time_step) && time_step != 0.0), SEMO_EINVAL,
                   "calcRHS: time_step must be finite and non-zero");
        const int N = 4 + nsp;
        DBuf<double> dL, dR, dG, dC, dS, dF, dr;
        face_to_dev_soa(c, dL, L, 8 + nsp);
        face_to_dev_soa(c, dR, R, 8 + nsp);
        if (grad) face_to_dev_soa(c, dG, grad, 9);
        if (src) cell_to_dev_soa(c, dC, cells, 8 + (nsp - 1) + 2 * N);
        if (src && st) to_dev(c, dS, st, 3);
        dF.alloc((size_t)c->nf * N);
        dr.alloc((size_t)c->nc * N);
        rhs_device(c, nsp, src, time_step, g, dL.p, dR.p, grad ? dG.p : nullptr, src ? dC.p : nullptr,
                   (src && st) ? dS.p : nullptr, dF.p, dr.p);
        from_dev(c, dr.p, res, N);
    });
}

int semo_calc_fluxes(semo_ctx* c, int nsp, const double* L, const double* R, const double* grad, double* res) {
    if (!c || !L || !R || !res) return SEMO_EINVAL;
    return rhs_common(c, nsp, false, 1.0, nullptr, L, R, grad, nullptr, nullptr, res);
}

int semo_calc_rhs(semo_ctx* c, int nsp, double time_step, const double* g, const double* L, const double* R,
                  const double* grad, const double* cells, const double* st, double* res) {
    if (!c || !g || !L || !R || !cells || !res) return SEMO_EINVAL;
    return rhs_common(c, nsp, true, time_step, g, L, R, grad, cells, st, res);
}

int semo_update_values(semo_ctx* c, int nsp, const double* urf, const double* limits, const double* res,
                       double* cells, double* face_un) {
    if (!c || !urf || !limits || !res || !cells || !face_un) return SEMO_EINVAL;
    return guard(c, [&] {
        need_mesh(c);
        SEMO_CHECK(nsp >= 1 && nsp <= 8, SEMO_EINVAL, "updateValues: 1 <= nsp <= 8");
        DBuf<double> dr, dc, du, dl;
        to_dev(c, dr, res, 4 + nsp);
        to_dev(c, dc, cells, 5 + nsp);
        std::vector<double> u(c->nf);
        for (int k = 0; k < c->nf; ++k) u[k] = face_un[c->fperm[k]];
        du.upload(u, c->stream);
        dl.upload(limits, 10, c->stream);
        update_values_device(c, nsp, urf, dl.p, dr.p, dc.p, du.p);
        from_dev(c, dc.p, cells, 5 + nsp);
        du.download(u.data(), u.size(), c->stream);
        SEMO_HIP(hipStreamSynchronize(c->stream));
        for (int k = 0; k < c->nf; ++k) face_un[c->fperm[k]] = u[k];
    });
}

int semo_compress_volume_fraction(semo_ctx* c, double cfl, int iters, double* vf, const double* face_r,
                                  double* residuals) {
    if (!c || !vf) return SEMO_EINVAL;
    return guard(c, [&] {
        need_mesh(c);
        SEMO_CHECK(iters >= 0 && std::isfinite(cfl), SEMO_EINVAL, "compressVolumeFraction: iters >= 0, finite cfl");
        DBuf<double> dv, db;
        to_dev(c, dv, vf, 1);
        std::vector<double> br(std::max(c->nbf, 1), 0.0);
        if (face_r)
            for (int j = 0; j < c->nbf; ++j) br[j] = face_r[c->fperm[c->nif + j]];
        db.upload(br, c->stream);
        compress_vf_device(c, cfl, iters, dv.p, db.p, residuals);
        from_dev(c, dv.p, vf, 1);
    });
}

static void linsol_check(semo_ctx* c, int method, int nsp, double time_step) {
    need_mesh(c);
    SEMO_CHECK(method >= SEMO_LINSOL_POINT && method <= SEMO_LINSOL_LUSGS, SEMO_EINVAL,
               "method must be SEMO_LINSOL_POINT, _SINGLE or _LUSGS");
    SEMO_CHECK(nsp >= 1 && nsp <= 4, SEMO_EINVAL, "linear solver: 1 <= nsp <= 4");
    SEMO_CHECK(method == SEMO_LINSOL_SINGLE || (std::isfinite(time_step) && time_step != 0.0), SEMO_EINVAL,
               "linear solver: time_step must be finite and non-zero");
}

int semo_linear_solver(semo_ctx* c, int method, int nsp, double time_step, const double* state, double* res) {
    if (!c || !state || !res) return SEMO_EINVAL;
    return guard(c, [&] {
        linsol_check(c, method, nsp, time_step);
        DBuf<double> ds, dr;
        to_dev(c, ds, state, 12 + 3 * nsp);
        to_dev(c, dr, res, 4 + nsp);
        linsol_device(c, method, nsp, time_step, ds.p, dr.p);
        from_dev(c, dr.p, res, 4 + nsp);
    });
}

int semo_linear_solver_device(semo_ctx* c, int method, int nsp, double time_step, const double* d_state,
                              double* d_res) {
    if (!c || !d_state || !d_res) return SEMO_EINVAL;
    return guard(c, [&] {
        linsol_check(c, method, nsp, time_step);
        linsol_device(c, method, nsp, time_step, d_state, d_res);
        SEMO_HIP(hipStreamSynchronize(c->stream));
    });
}

int semo_set_momentum_solver(semo_ctx* c, int mode) {
    if (!c) return SEMO_EINVAL;
    return guard(c, [&] {
        SEMO_CHECK(mode >= SEMO_MSOLVER_JACOBI && mode <= SEMO_MSOLVER_AUTO, SEMO_EINVAL,
                   "momentum solver must be SEMO_MSOLVER_JACOBI, _AMG or _AUTO");
        c->msolver = mode;
        if (mode != SEMO_MSOLVER_AUTO) c->amg_mom_on = c->amg_vf_on = false;
    });
}

int semo_set_les(semo_ctx* c, int model) {
    if (!c) return SEMO_EINVAL;
    return guard(c, [&] {
        SEMO_CHECK(model >= 0 && model <= SEMO_LES_DYNAMIC_SMAGORINSKY, SEMO_EINVAL,
                   "LES model must be 0 (laminar) or an SEMO_LES_* model");
        c->les_model = model;
    });
}

int semo_les(semo_ctx* c, int model, const double* U, const double* V, const double* W, const double* rho,
             const double* face_rU, const double* face_rV, const double* face_rW, double* muT, double* kSGS) {
    if (!c || !U || !V || !W || !rho || !face_rU || !face_rV || !face_rW || !muT || !kSGS) return SEMO_EINVAL;
    return guard(c, [&] {
        need_case(c);
        SEMO_CHECK(model >= SEMO_LES_WALE && model <= SEMO_LES_DYNAMIC_SMAGORINSKY, SEMO_EINVAL,
                   "LES model must be SEMO_LES_WALE, SEMO_LES_SMAGORINSKY or SEMO_LES_DYNAMIC_SMAGORINSKY");
        DBuf<double> dU, dV, dW, dR, dbr, dm, dk;
        to_dev(c, dU, U, 1); to_dev(c, dV, V, 1); to_dev(c, dW, W, 1); to_dev(c, dR, rho, 1);
        size_t nb = (size_t)std::max(c->nbf, 1);
        std::vector<double> br(3 * nb, 0.0);
        const double* fr[3] = {face_rU, face_rV, face_rW};
        for (int k = 0; k < 3; ++k)
            for (int j = 0; j < c->nbf; ++j) br[k * nb + j] = fr[k][c->fperm[c->nif + j]];
        dbr.upload(br, c->stream);
        dm.alloc(c->nc); dk.alloc(c->nc);
        les_device(c, model, dU.p, dV.p, dW.p, dR.p, dbr.p, dm.p, dk.p);
        from_dev(c, dm.p, muT, 1);
        from_dev(c, dk.p, kSGS, 1);
    });
}

int semo_least_square_noalloc(semo_ctx* c, const double* d_phi, const double* d_face_r, double* d_grad) {
    if (!c || !d_phi || !d_face_r || !d_grad) return SEMO_EINVAL;
    return guard(c, [&] {
        need_mesh(c);
        lsq1_noalloc(c, d_phi, d_face_r, d_grad);
    });
}

int semo_cell_min_max(semo_ctx* c, const double* phi, double* vmax, double* vmin) {
    if (!c || !phi || !vmax || !vmin) return SEMO_EINVAL;
    return guard(c, [&] {
        need_mesh(c);
        DBuf<double> dphi, a, b;
        to_dev(c, dphi, phi, 1);
        a.alloc(c->nc); b.alloc(c->nc);
        minmax_standalone(c, dphi.p, a.p, b.p);
        from_dev(c, a.p, vmax, 1);
        from_dev(c, b.p, vmin, 1);
    });
}

int semo_solve_coo(semo_ctx* c, const char* eq, int32_t n, int64_t nnz, const int32_t* rows, const int32_t* cols,
                   const double* vals, int32_t nrhs, const double* b, double* x, semo_solve_stats* st) {
    if (!c || !rows || !cols || !vals || !b || !x) return SEMO_EINVAL;
    return guard(c, [&] {
        SEMO_HIP(hipSetDevice(c->device));
        solve_coo(c, eq, n, nnz, rows, cols, vals, nrhs, b, x, st);
    });
}

int semo_profile_enable(semo_ctx* c, int on) {
    if (!c) return SEMO_EINVAL;
    return guard(c, [&] { c->prof.flush(); c->prof.on = on != 0; });
}

int semo_profile_reset(semo_ctx* c) {
    if (!c) return SEMO_EINVAL;
    return guard(c, [&] { c->prof.flush(); c->prof.stats.clear(); });
}

int semo_stream_copy(semo_ctx* c, int64_t n, int reps, double* ms) {
    if (!c || n <= 0 || reps <= 0 || !ms) return SEMO_EINVAL;
    return guard(c, [&] {
        DBuf<double> a, b;
        a.alloc((size_t)n); b.alloc((size_t)n);
        zero_d(c, a.p, a.bytes());
        copy_d(c, b.p, a.p, (size_t)n);                 // first touch
        hipEvent_t e0, e1;
        SEMO_HIP(hipEventCreate(&e0));
        SEMO_HIP(hipEventCreate(&e1));
        SEMO_HIP(hipEventRecord(e0, c->stream));
        for (int r = 0; r < reps; ++r) copy_d(c, (r & 1) ? a.p : b.p, (r & 1) ? b.p : a.p, (size_t)n);
        SEMO_HIP(hipEventRecord(e1, c->stream));
        SEMO_HIP(hipEventSynchronize(e1));
        float t = 0.0f;
        SEMO_HIP(hipEventElapsedTime(&t, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms = (double)t / reps;
    });
}

int semo_profile_entry(semo_ctx* c, int i, char* name, int64_t* count, double* ms, double* bytes) {
    if (!c) return SEMO_EINVAL;
    return guard(c, [&] {
        c->prof.flush();
        SEMO_CHECK(i >= 0 && i < (int)c->prof.stats.size(), SEMO_EINVAL, "no such entry");
        auto it = c->prof.stats.begin();
        std::advance(it, i);
        if (name) { strncpy(name, it->first.c_str(), 63); name[63] = 0; }
        if (count) *count = it->second.count;
        if (ms) *ms = it->second.ms;
        if (bytes) *bytes = it->second.bytes;
    });
}

}  // extern "C"
