// Device helpers shared by the assembly and adjoint kernels: launch
// shape, case constants, the cell -> face walk, reference min/max
// semantics, block partial sums and the Rhie-Chow face velocity.
#pragma once
#include <climits>

#include "common.hpp"

namespace semo {
namespace {


constexpr int TPB = 256;
inline int nblk(long n) { return (int)((n + TPB - 1) / TPB); }

struct CaseK {
    double rho0, rho1, mu0, mu1, sigma, g0, g1, g2, dt, urf_u, urf_p, urf_pu, lim;
    int body;   // gravity or surface tension present
};

CaseK case_k(const semo_ctx* c) {
    CaseK k;
    k.rho0 = c->cs.rho[0]; k.rho1 = c->cs.rho[1]; k.mu0 = c->cs.mu[0]; k.mu1 = c->cs.mu[1];
    k.sigma = c->cs.sigma; k.g0 = c->cs.g[0]; k.g1 = c->cs.g[1]; k.g2 = c->cs.g[2];
    k.dt = c->cs.dt; k.urf_u = c->cs.urf_mom; k.urf_p = c->cs.urf_p; k.urf_pu = c->cs.urf_pu;
    k.lim = c->cs.limit;
    k.body = (k.g0 != 0.0 || k.g1 != 0.0 || k.g2 != 0.0 || k.sigma != 0.0) ? 1 : 0;
    return k;
}

// Faces of cell i in ascending order from the sliced-ELL cell -> face list
// (64-cell slices, entry k of a cell at slice base + 64k + lane; ~f marks a
// face whose neighbour the cell is; INT_MIN pads).  own: the cell owns f.
#define CELL_FACES_BEGIN(i, SL_, ELL_, f, own)                                            \
    for (int e_ = SL_[(i) >> 6] + ((i) & 63); e_ < SL_[((i) >> 6) + 1]; e_ += 64) {      \
        int code_ = ELL_[e_];                                                              \
        if (code_ == INT_MIN) break;                                                       \
        bool own = code_ >= 0;                                                             \
        int f = own ? code_ : ~code_;

// std::max / std::min semantics (first argument wins ties), as in the reference
__device__ inline double smax(double a, double b) { return (a < b) ? b : a; }
__device__ inline double smin(double a, double b) { return (b < a) ? b : a; }
__device__ inline double clip01(double a) { return smax(0.0, smin(1.0, a)); }

// block-partial sums of two quantities (deterministic two-stage reduction)
__device__ inline void block_sum2(double a, double b, double* __restrict__ part) {
    __shared__ double sa[TPB / 64], sb[TPB / 64];
    for (int o = 32; o > 0; o >>= 1) { a += __shfl_down(a, o, 64); b += __shfl_down(b, o, 64); }
    int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) { sa[w] = a; sb[w] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double x = 0, y = 0;
        for (int k = 0; k < TPB / 64; ++k) { x += sa[k]; y += sb[k]; }
        part[2 * blockIdx.x] = x; part[2 * blockIdx.x + 1] = y;
    }
}

// --------------------------------------------------------- Rhie-Chow face --
// Face-normal velocity of the assembly face loops: skew-corrected, stencil-
// clipped linear interpolation (eqMomentum.cpp:519-553) followed by
// calcInterpolVelPressure / Gravity / SurfTens (surfaceNormalVelocity.cpp:
// 125-223).
struct FaceIn {
    const double *P, *U, *V, *W, *VF, *Rho, *kappa, *srcG, *srcS;
    const double *gP, *gU, *gV, *gW;
    const double *mxU, *mnU, *mxV, *mnV, *mxW, *mnW, *mxP, *mnP;
    const double *nrm, *ePN, *skew, *wC, *alpha, *magPN, *area;
    const int *owner, *nbr;
    const int8_t* flip;
    int nc;
};

// load / store of data touched once per launch (per-face geometry, face
// scratch): non-temporal, so that it does not push the gathered cell data out
// of the L2
template <class T> __device__ inline T ldnt(const T* p) { return __builtin_nontemporal_load(p); }
template <class T> __device__ inline void stnt(T* p, T v) { __builtin_nontemporal_store(v, p); }
template <bool NT, class T> __device__ inline T ld(const T* p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT = true>
__device__ inline void ldnt3(const double* p, size_t f, double* o) {
    o[0] = ld<NT>(p + 3 * f); o[1] = ld<NT>(p + 3 * f + 1); o[2] = ld<NT>(p + 3 * f + 2);
}

// NT: the face geometry is read non-temporally (measured faster in the
// momentum face kernel, slower in the pressure face kernel)
template <bool NT = false>
__device__ inline double face_unf(const FaceIn& in, const CaseK& k, int f, int l, int r) {
    double n[3], e[3], sk[3];
    ldnt3<NT>(in.nrm, f, n); ldnt3<NT>(in.ePN, f, e); ldnt3<NT>(in.skew, f, sk);
    double wCL = ld<NT>(in.wC + f), wCR = 1.0 - wCL, alpha = ld<NT>(in.alpha + f), dPN = ld<NT>(in.magPN + f);
    double uL[3] = {in.U[l], in.V[l], in.W[l]}, uR[3] = {in.U[r], in.V[r], in.W[r]};
    const double* gl[3] = {in.gU + 3 * (size_t)l, in.gV + 3 * (size_t)l, in.gW + 3 * (size_t)l};
    const double* gr[3] = {in.gU + 3 * (size_t)r, in.gV + 3 * (size_t)r, in.gW + 3 * (size_t)r};
    for (int a = 0; a < 3; ++a)
        for (int v = 0; v < 3; ++v) { uL[v] += gl[v][a] * sk[a]; uR[v] += gr[v][a] * sk[a]; }
    uL[0] = smax(in.mnU[l], smin(in.mxU[l], uL[0])); uR[0] = smax(in.mnU[r], smin(in.mxU[r], uR[0]));
    uL[1] = smax(in.mnV[l], smin(in.mxV[l], uL[1])); uR[1] = smax(in.mnV[r], smin(in.mxV[r], uR[1]));
    uL[2] = smax(in.mnW[l], smin(in.mxW[l], uL[2])); uR[2] = smax(in.mnW[r], smin(in.mxW[r], uR[2]));
    double UnL = uL[0] * n[0] + uL[1] * n[1] + uL[2] * n[2];
    double UnR = uR[0] * n[0] + uR[1] * n[1] + uR[2] * n[2];
    double UnF = wCL * UnL + wCR * UnR;
    double RL = in.Rho[l], RR = in.Rho[r];
    double Rs = 1.0 / (wCL / RL + wCR / RR);
    double t1 = k.dt / Rs;
    const double* gPl = in.gP + 3 * (size_t)l;
    const double* gPr = in.gP + 3 * (size_t)r;
    UnF -= alpha * t1 * (in.P[r] - in.P[l]) / dPN;
    for (int a = 0; a < 3; ++a) {
        UnF -= t1 * Rs * wCL / RL * gPl[a] * (n[a] - alpha * e[a]);
        UnF -= t1 * Rs * wCR / RR * gPr[a] * (n[a] - alpha * e[a]);
    }
    for (int a = 0; a < 3; ++a) {
        UnF += t1 * Rs * wCL / RL * gPl[a] * n[a];
        UnF += t1 * Rs * wCR / RR * gPr[a] * n[a];
    }
    if (k.body) {
        double g[3] = {k.g0, k.g1, k.g2};
        size_t nc = (size_t)in.nc;
        for (int a = 0; a < 3; ++a) UnF += alpha * t1 * Rs * g[a] * alpha * e[a];
        for (int a = 0; a < 3; ++a) {
            UnF += t1 * Rs * wCL / RL * in.srcG[a * nc + l] * (n[a] - alpha * e[a]);
            UnF += t1 * Rs * wCR / RR * in.srcG[a * nc + r] * (n[a] - alpha * e[a]);
        }
        for (int a = 0; a < 3; ++a) {
            UnF -= alpha * t1 * Rs * wCL / RL * in.srcG[a * nc + l] * n[a];
            UnF -= alpha * t1 * Rs * wCR / RR * in.srcG[a * nc + r] * n[a];
        }
        UnF += alpha * t1 * k.sigma * (wCL * in.kappa[l] + wCR * in.kappa[r]) * (in.VF[r] - in.VF[l]) / dPN;
        for (int a = 0; a < 3; ++a) {
            UnF += t1 * Rs * wCL / RL * in.srcS[a * nc + l] * (n[a] - alpha * e[a]);
            UnF += t1 * Rs * wCR / RR * in.srcS[a * nc + r] * (n[a] - alpha * e[a]);
        }
        for (int a = 0; a < 3; ++a) {
            UnF -= alpha * t1 * Rs * wCL / RL * in.srcS[a * nc + l] * n[a];
            UnF -= alpha * t1 * Rs * wCR / RR * in.srcS[a * nc + r] * n[a];
        }
    }
    return UnF;
}


FaceIn face_in(semo_ctx* c) {
    FaceIn in;
    in.P = c->P.p; in.U = c->U.p; in.V = c->V.p; in.W = c->W.p; in.VF = c->VF.p; in.Rho = c->Rho.p;
    in.kappa = c->kappa.p; in.srcG = c->srcG.p; in.srcS = c->srcS.p;
    in.gP = c->gP.p; in.gU = c->gU.p; in.gV = c->gV.p; in.gW = c->gW.p;
    in.mxU = c->mxU.p; in.mnU = c->mnU.p; in.mxV = c->mxV.p; in.mnV = c->mnV.p; in.mxW = c->mxW.p;
    in.mnW = c->mnW.p; in.mxP = c->mxP.p; in.mnP = c->mnP.p;
    in.nrm = c->f_nrm.p; in.ePN = c->f_ePN.p; in.skew = c->f_skew.p; in.wC = c->f_wC.p;
    in.alpha = c->f_alpha.p; in.magPN = c->f_magPN.p; in.area = c->f_area.p;
    in.owner = c->d_owner.p; in.nbr = c->d_nbr.p; in.flip = c->d_fflip.p; in.nc = c->nc;
    return in;
}

}  // namespace
}  // namespace semo
