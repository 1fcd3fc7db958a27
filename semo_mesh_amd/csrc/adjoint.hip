// Discrete adjoint of the pressure-based SIMPLE iterations on the device.
//
// The reference has no adjoint.  This file differentiates, in reverse mode,
// exactly the discrete map that the forward path computes for a single-phase
// case with a fixed time step: n_iter outer iterations of
//   recon -> LSQ gradients + stencil bounds -> momentum assembly -> 3 solves
//   -> U += urf dU -> recon -> gradients -> pressure assembly -> p' solve
//   -> two Green-Gauss passes of p' -> p, U correction
// with a Darcy term alpha_p*V on the momentum diagonal and -alpha_p*V*U on its
// right-hand side.  The objective is J = sum over boundary faces of
// w[patch] * A_f * p_f of the final state.  Every piecewise operation (upwind
// switch, stencil clipping, min/max, value limiter) is differentiated on its
// active branch, the linear solves as exact inverses (A^-T for momentum, the
// symmetric pressure Laplacian with the same AMG-FCG).
//
// Memory: the state (U, V, W, P) at the start of every iteration is kept
// (32 B per cell per iteration); the reverse sweep recomputes one iteration
// forward from its checkpoint, then runs its adjoint.  All adjoint kernels
// are gathers over a cell's faces, stencil and boundary faces in a fixed
// order, so the gradient is deterministic (no floating-point atomics).  The
// vertex stencil is symmetric (cells sharing a point), which lets the
// transposed LSQ and min/max be written as gathers too; this is checked once.
#include <algorithm>
#include <cmath>

#include "kern.hpp"

namespace semo {
namespace {

// ------------------------------------------------------------ clip helpers --
// forward: v = smax(mn, smin(mx, raw)).  Adjoint of v into raw / mx / mn.
__device__ inline void clip_adj(double raw, double mx, double mn, double vb, double& rawb, double& mxb,
                                double& mnb) {
    double inner = smin(mx, raw);
    if (mn < inner) {
        if (raw < mx) rawb += vb; else mxb += vb;
    } else {
        mnb += vb;
    }
}

// Adjoint contributions of one cell side of a face.  side: 0 = owner (l),
// 1 = neighbour (r); the accumulators belong to that cell.
struct SideAcc {
    double u[3];        // cell velocity
    double gu[3][3];    // grad of U, V, W  [component][axis]
    double mx[3], mn[3];   // stencil bounds of U, V, W
    double p, gp[3], mxp, mnp;
};

__device__ inline void acc_zero(SideAcc& a) {
    for (int v = 0; v < 3; ++v) {
        a.u[v] = 0; a.mx[v] = 0; a.mn[v] = 0; a.gp[v] = 0;
        for (int q = 0; q < 3; ++q) a.gu[v][q] = 0;
    }
    a.p = 0; a.mxp = 0; a.mnp = 0;
}

// reverse of face_unf (kern.hpp) for the single-phase, no-body-force case:
// the UnF adjoint ub is pushed into the quantities of cell `me`
__device__ inline void face_unf_adj(const FaceIn& in, const CaseK& k, int f, int l, int r, double ub, int side,
                                    SideAcc& a) {
    const double* n = in.nrm + 3 * (size_t)f;
    const double* e = in.ePN + 3 * (size_t)f;
    const double* sk = in.skew + 3 * (size_t)f;
    double wCL = in.wC[f], wCR = 1.0 - wCL, alpha = in.alpha[f], dPN = in.magPN[f];
    double RL = in.Rho[l], RR = in.Rho[r];
    double Rs = 1.0 / (wCL / RL + wCR / RR);
    double t1 = k.dt / Rs;
    int c = side ? r : l;
    double w = side ? wCR : wCL;
    double Rc = side ? RR : RL;
    // -alpha t1 (P_r - P_l) / dPN
    a.p += (side ? -1.0 : 1.0) * alpha * t1 / dPN * ub;
    // gradient-of-p terms: -(n - alpha e) + n = alpha e
    for (int q = 0; q < 3; ++q) a.gp[q] += ub * t1 * Rs * w / Rc * alpha * e[q];
    // clipped skew-corrected velocity of this side
    double u[3] = {in.U[c], in.V[c], in.W[c]};
    const double* g[3] = {in.gU + 3 * (size_t)c, in.gV + 3 * (size_t)c, in.gW + 3 * (size_t)c};
    const double* mx[3] = {in.mxU, in.mxV, in.mxW};
    const double* mn[3] = {in.mnU, in.mnV, in.mnW};
    for (int v = 0; v < 3; ++v) {
        double raw = u[v];
        for (int q = 0; q < 3; ++q) raw += g[v][q] * sk[q];
        double rb = 0.0;
        clip_adj(raw, mx[v][c], mn[v][c], w * n[v] * ub, rb, a.mx[v], a.mn[v]);
        a.u[v] += rb;
        for (int q = 0; q < 3; ++q) a.gu[v][q] += rb * sk[q];
    }
}

// ------------------------------------------------------------ objective ----
// J contribution of each boundary face: w[patch] * A * p_f with the face
// pressure of the boundary reconstruction (reconIncom.cpp:591-620)
__global__ void __launch_bounds__(TPB) k_adj_objective(int nbf, int nif, const int* __restrict__ bpatch,
                                                       const semo_patch_bc* __restrict__ bcs,
                                                       const double* __restrict__ wpatch,
                                                       const int* __restrict__ owner,
                                                       const double* __restrict__ area,
                                                       const double* __restrict__ P, double* __restrict__ jf) {
    int j = xcd_block() * blockDim.x + threadIdx.x;
    if (j >= nbf) return;
    int f = nif + j, b = bpatch[j];
    double pf = bcs[b].p_type == SEMO_BC_FIXED_VALUE ? bcs[b].p_value : P[owner[f]];
    jf[j] = wpatch[b] * area[f] * pf;
}

// dJ/dP of the final state, gathered per cell over its boundary faces
__global__ void __launch_bounds__(TPB) k_adj_seed(int nc, int nif, const int* __restrict__ bf_ptr,
                                                  const int* __restrict__ bf_idx, const int* __restrict__ bpatch,
                                                  const semo_patch_bc* __restrict__ bcs,
                                                  const double* __restrict__ wpatch,
                                                  const double* __restrict__ area, double* __restrict__ Pb) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double s = 0.0;
    for (int q = bf_ptr[i]; q < bf_ptr[i + 1]; ++q) {
        int f = bf_idx[q], b = bpatch[f - nif];
        if (bcs[b].p_type != SEMO_BC_FIXED_VALUE) s += wpatch[b] * area[f];
    }
    Pb[i] = s;
}

// ------------------------------------------------------ p / U correction ---
// reverse of k_p_correct: P2 = clip(P + urf_p p'), U2 = clip(U1 - urf_pu dt/rho g2)
// aX: adjoint of the iteration's output state; X2: that output state (clip
// masks); writes p' adjoint, the U1 adjoint, g2 adjoint; P adjoint (start) +=
__global__ void __launch_bounds__(TPB) k_adj_pcorrect(int nc, CaseK k, const double* __restrict__ Rho,
                                                      const double* __restrict__ P2, const double* __restrict__ U2,
                                                      const double* __restrict__ V2, const double* __restrict__ W2,
                                                      const double* __restrict__ aP, const double* __restrict__ aU,
                                                      const double* __restrict__ aV, const double* __restrict__ aW,
                                                      double* __restrict__ ppb, double* __restrict__ nP,
                                                      double* __restrict__ u1U, double* __restrict__ u1V,
                                                      double* __restrict__ u1W, double* __restrict__ g2b) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double mp = fabs(P2[i]) < k.lim ? 1.0 : 0.0;
    double pb = mp * aP[i];
    ppb[i] = k.urf_p * pb;
    nP[i] += pb;
    double t1 = k.dt / Rho[i];
    double ub[3] = {fabs(U2[i]) < k.lim ? aU[i] : 0.0, fabs(V2[i]) < k.lim ? aV[i] : 0.0,
                    fabs(W2[i]) < k.lim ? aW[i] : 0.0};
    u1U[i] = ub[0]; u1V[i] = ub[1]; u1W[i] = ub[2];
    for (int a = 0; a < 3; ++a) g2b[3 * (size_t)i + a] = -k.urf_pu * t1 * ub[a];
}

// ------------------------------------------------- Green-Gauss (p') reverse --
// face value adjoint from the cell gradient adjoint (k_gg_cell, internal part)
__global__ void __launch_bounds__(TPB) k_adj_gg_face(int nif, const int* __restrict__ owner,
                                                     const int* __restrict__ nbr, const double* __restrict__ nrm,
                                                     const double* __restrict__ area, const double* __restrict__ vol,
                                                     const double* __restrict__ gb, double* __restrict__ vfb) {
    int f = xcd_block() * blockDim.x + threadIdx.x;
    if (f >= nif) return;
    int l = owner[f], r = nbr[f];
    const double* n = nrm + 3 * (size_t)f;
    double A = area[f], vl = vol[l], vr = vol[r], s = 0.0;
    for (int a = 0; a < 3; ++a) s += gb[3 * (size_t)l + a] * n[a] * A / vl - gb[3 * (size_t)r + a] * n[a] * A / vr;
    vfb[f] = s;
}

// per cell: p' adjoint += face-value weights + boundary faces; adjoint of the
// previous pass's gradient (skew correction of k_gg_face)
__global__ void __launch_bounds__(TPB) k_adj_gg_cell(
    int nc, int nif, const int* __restrict__ cf_sl, const int* __restrict__ cf_ell,
    const int* __restrict__ bpatch, const semo_patch_bc* __restrict__ bcs, const double* __restrict__ nrm,
    const double* __restrict__ area, const double* __restrict__ vol, const double* __restrict__ wCv,
    const double* __restrict__ skew, const double* __restrict__ gb, const double* __restrict__ vfb,
    double* __restrict__ ppb, double* __restrict__ gprevb) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double s = 0.0, gp[3] = {0, 0, 0};
    double vi = vol[i];
    CELL_FACES_BEGIN(i, cf_sl, cf_ell, f, own)
        if (f < nif) {
            double w = own ? wCv[f] : 1.0 - wCv[f];
            double v = vfb[f] * w;
            s += v;
            for (int a = 0; a < 3; ++a) gp[a] += v * skew[3 * (size_t)f + a];
        } else if (bcs[bpatch[f - nif]].p_type != SEMO_BC_FIXED_VALUE) {
            const double* n = nrm + 3 * (size_t)f;
            for (int a = 0; a < 3; ++a) s += gb[3 * (size_t)i + a] * n[a] * area[f] / vi;
        }
    }
    ppb[i] += s;
    if (gprevb)
        for (int a = 0; a < 3; ++a) gprevb[3 * (size_t)i + a] = gp[a];
}

// --------------------------------------------------- pressure assembly reverse --
// Bp_i = -sum_own UnF A + sum_nbr UnF A - sum_boundary UnF_b A (k_p_face /
// k_p_cell); lam = adjoint of Bp.  Gathers, per cell, the contributions of
// every face to the cell's own velocity, gradients, bounds and pressure;
// boundary-face velocity adjoints go to bULb/bURb (one owner per face).
struct StageAdj {
    double *u, *v, *w;              // velocity adjoint (+=)
    double *p;                      // pressure adjoint (+=)
    double *gu, *gv, *gw, *gp;      // gradient adjoints [3n] (=)
    double *mxu, *mnu, *mxv, *mnv, *mxw, *mnw, *mxp, *mnp;   // bound adjoints (=)
    double *bULb, *bURb, *bPb;      // boundary-face adjoints (+=)
};

__device__ inline void store_side(const StageAdj& o, int i, const SideAcc& a, int nbf_unused) {
    (void)nbf_unused;
    o.u[i] += a.u[0]; o.v[i] += a.u[1]; o.w[i] += a.u[2];
    o.p[i] += a.p;
    double* g[3] = {o.gu, o.gv, o.gw};
    for (int v = 0; v < 3; ++v)
        for (int q = 0; q < 3; ++q) g[v][3 * (size_t)i + q] = a.gu[v][q];
    for (int q = 0; q < 3; ++q) o.gp[3 * (size_t)i + q] = a.gp[q];
    o.mxu[i] = a.mx[0]; o.mnu[i] = a.mn[0]; o.mxv[i] = a.mx[1]; o.mnv[i] = a.mn[1];
    o.mxw[i] = a.mx[2]; o.mnw[i] = a.mn[2];
    if (o.mxp) { o.mxp[i] = a.mxp; o.mnp[i] = a.mnp; }
}

__global__ void __launch_bounds__(TPB) k_adj_p_cell(
    int nc, int nif, int nbf, CaseK k, FaceIn in, const int* __restrict__ cf_sl, const int* __restrict__ cf_ell,
    const int* __restrict__ bpatch, const semo_patch_bc* __restrict__ bcs, const double* __restrict__ bRho,
    const double* __restrict__ lam, StageAdj o) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    SideAcc a;
    acc_zero(a);
    double li = lam[i];
    CELL_FACES_BEGIN(i, cf_sl, cf_ell, f, own)
        double A = in.area[f];
        if (f < nif) {
            int l = in.owner[f], r = in.nbr[f];
            double ub = A * (lam[r] - lam[l]);
            face_unf_adj(in, k, f, l, r, ub, own ? 0 : 1, a);
            continue;
        }
        int j = f - nif;
        semo_patch_bc bc = bcs[bpatch[j]];
        const double* n = in.nrm + 3 * (size_t)f;
        double ub = -A * li;
        for (int v = 0; v < 3; ++v) {
            o.bULb[(size_t)v * nbf + j] += 0.5 * n[v] * ub;
            o.bURb[(size_t)v * nbf + j] += 0.5 * n[v] * ub;
        }
        if (bc.p_type == SEMO_BC_FIXED_VALUE) {
            const double* e = in.ePN + 3 * (size_t)f;
            double t1 = 1.0 / bRho[j] * k.dt, dPN = 0.5 * in.magPN[f], alpha = in.alpha[f];
            a.p += ub * alpha * t1 / dPN;
            for (int q = 0; q < 3; ++q) a.gp[q] += ub * alpha * t1 * e[q];
        }
    }
    store_side(o, i, a, nbf);
}

// --------------------------------------------------- momentum assembly reverse --
// k_mom_face + k_mom_cell reversed.  lam = A^-T dX-adjoint (3 components),
// X = the stage's solution (A^-bar = -lam X^T), B^-bar = lam.
__global__ void __launch_bounds__(TPB) k_adj_mom_cell(
    int nc, int nif, int nbf, CaseK k, FaceIn in, const int* __restrict__ cf_sl, const int* __restrict__ cf_ell,
    const int* __restrict__ bpatch, const semo_patch_bc* __restrict__ bcs, const double* __restrict__ vol,
    const double* __restrict__ Mu, const double* __restrict__ alphaP, const double* __restrict__ bUL,
    const double* __restrict__ bUR, const double* __restrict__ bP, const double* __restrict__ bMu,
    const double* __restrict__ bRho, const double* __restrict__ lam, const double* __restrict__ X,
    StageAdj o, double* __restrict__ alphab) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    size_t N = (size_t)nc;
    SideAcc a;
    acc_zero(a);
    double li[3] = {lam[i], lam[N + i], lam[2 * N + i]};
    double xi[3] = {X[i], X[N + i], X[2 * N + i]};
    double Db = -(li[0] * xi[0] + li[1] * xi[1] + li[2] * xi[2]);   // diagonal adjoint of row i
    double uc[3] = {in.U[i], in.V[i], in.W[i]};
    double R = in.Rho[i], vl = vol[i], dt = k.dt;
    // time term and Darcy source
    for (int v = 0; v < 3; ++v) a.u[v] += -R * 1.5 * vl / dt * li[v];
    if (alphaP) {
        double ap = alphaP[i] * vl;
        for (int v = 0; v < 3; ++v) a.u[v] -= ap * li[v];
        alphab[i] += vl * Db - vl * (li[0] * uc[0] + li[1] * uc[1] + li[2] * uc[2]);
    }
    CELL_FACES_BEGIN(i, cf_sl, cf_ell, f, own)
        double area = in.area[f];
        const double* n = in.nrm + 3 * (size_t)f;
        const double* e = in.ePN + 3 * (size_t)f;
        if (f < nif) {
            int l = in.owner[f], r = in.nbr[f];
            int side = own ? 0 : 1;
            const double* sk = in.skew + 3 * (size_t)f;
            double wCL = in.wC[f], wCR = 1.0 - wCL;
            double muF = 1.0 / (wCL / Mu[l] + wCR / Mu[r]);
            double RL = in.Rho[l], RR = in.Rho[r];
            double dPN = in.magPN[f], alpha = in.alpha[f];
            double UnF = face_unf(in, k, f, l, r);
            double wL = (UnF > 0.0) ? 1.0 : 0.0, wR = 1.0 - wL;
            double uL[3] = {in.U[l], in.V[l], in.W[l]}, uR[3] = {in.U[r], in.V[r], in.W[r]};
            double Fv[3] = {wL * uL[0] + wR * uR[0], wL * uL[1] + wR * uR[1], wL * uL[2] + wR * uR[2]};
            const double* gPl = in.gP + 3 * (size_t)l;
            const double* gPr = in.gP + 3 * (size_t)r;
            double PF0 = wCL * in.P[l] + wCR * in.P[r];
            for (int q = 0; q < 3; ++q) { PF0 += wCL * gPl[q] * sk[q]; PF0 += wCR * gPr[q] * sk[q]; }
            double PF1 = smax(in.mnP[l], smin(in.mxP[l], PF0));
            double PF = smax(in.mnP[r], smin(in.mxP[r], PF1));
            // adjoints of the face outputs
            double lL[3] = {lam[l], lam[N + l], lam[2 * N + l]}, lR[3] = {lam[r], lam[N + r], lam[2 * N + r]};
            double xL[3] = {X[l], X[N + l], X[2 * N + l]}, xR[3] = {X[r], X[N + r], X[2 * N + r]};
            double Xb = -(lL[0] * xL[0] + lL[1] * xL[1] + lL[2] * xL[2]);    // d_l += X
            double Yb = (lR[0] * xR[0] + lR[1] * xR[1] + lR[2] * xR[2]);     // d_r -= Y
            double aob = -(lL[0] * xR[0] + lL[1] * xR[1] + lL[2] * xR[2]);   // A(l, r)
            double anb = -(lR[0] * xL[0] + lR[1] * xL[1] + lR[2] * xL[2]);   // A(r, l)
            double UnFb = Xb * wL * RL * area + aob * wR * RL * area + Yb * wR * RR * area - anb * wL * RR * area;
            double PFb = 0.0, dfb[3];
            for (int v = 0; v < 3; ++v) {
                double s2b = -lL[v], s5b = lR[v];
                UnFb += s2b * RL * Fv[v] * area + s5b * RR * Fv[v] * area;
                double Fvb = (s2b * RL + s5b * RR) * UnF * area;
                PFb += (s2b + s5b) * n[v] * area;
                dfb[v] = -(s2b + s5b);
                double dsb = side ? -s5b : -s2b;       // dL (owner) / dR (neighbour) adjoint
                // Fv = wL uL + wR uR
                a.u[v] += (side ? wR : wL) * Fvb;
                // df = alpha mu (uR - uL)/dPN A + sum mu (wCL gl + wCR gr)(n - alpha e) A
                a.u[v] += (side ? 1.0 : -1.0) * alpha * muF / dPN * area * dfb[v];
                double wc = side ? wCR : wCL;
                for (int q = 0; q < 3; ++q)
                    a.gu[v][q] += muF * wc * (n[q] - alpha * e[q]) * area * dfb[v] + n[q] * muF * area * dsb;
            }
            // PF = clip_r(clip_l(PF0))
            {
                double PF1b = 0.0, PF0b = 0.0, d0 = 0.0, d1 = 0.0;
                if (side) clip_adj(PF1, in.mxP[r], in.mnP[r], PFb, PF1b, a.mxp, a.mnp);
                else clip_adj(PF1, in.mxP[r], in.mnP[r], PFb, PF1b, d0, d1);
                if (side) clip_adj(PF0, in.mxP[l], in.mnP[l], PF1b, PF0b, d0, d1);
                else clip_adj(PF0, in.mxP[l], in.mnP[l], PF1b, PF0b, a.mxp, a.mnp);
                double wc = side ? wCR : wCL;
                a.p += wc * PF0b;
                for (int q = 0; q < 3; ++q) a.gp[q] += wc * sk[q] * PF0b;
            }
            face_unf_adj(in, k, f, l, r, UnFb, side, a);
            continue;
        }
        // boundary face of cell i (eqMomentum.cpp:795-1008)
        int j = f - nif;
        semo_patch_bc bc = bcs[bpatch[j]];
        double A = area;
        double ul[3], ur[3];
        for (int q = 0; q < 3; ++q) { ul[q] = bUL[q * (size_t)nbf + j]; ur[q] = bUR[q * (size_t)nbf + j]; }
        double UnF = 0.5 * (ul[0] * n[0] + ul[1] * n[1] + ul[2] * n[2]) + 0.5 * (ur[0] * n[0] + ur[1] * n[1] + ur[2] * n[2]);
        double RL = bRho[j];
        double t1 = 1.0 / RL * dt;
        double dPN = 0.5 * in.magPN[f], alpha = in.alpha[f];
        if (bc.p_type == SEMO_BC_FIXED_VALUE) {
            UnF -= alpha * t1 * (bc.p_value - in.P[i]) / dPN;
            for (int q = 0; q < 3; ++q) UnF += alpha * t1 * in.gP[3 * (size_t)i + q] * e[q];
        }
        double cU = 0.0;
        if (bc.u_type == SEMO_BC_ZERO_GRADIENT) cU = 1.0;
        else if (bc.u_type == SEMO_BC_INLET_OUTLET) cU = (uc[0] * n[0] + uc[1] * n[1] + uc[2] * n[2]) > 0.0 ? 1.0 : 0.0;
        double muF = 0.5 * bMu[j] + 0.5 * bMu[j];
        double Fv[3] = {0.5 * ul[0] + 0.5 * ur[0], 0.5 * ul[1] + 0.5 * ur[1], 0.5 * ul[2] + 0.5 * ur[2]};
        double UnFb = Db * cU * RL * A, PFb = 0.0;
        for (int v = 0; v < 3; ++v) {
            UnFb -= li[v] * RL * Fv[v] * A;
            PFb -= li[v] * n[v] * A;
        }
        o.bPb[j] += PFb;
        for (int v = 0; v < 3; ++v) {
            double Fvb = -li[v] * RL * UnF * A;
            double dnb = li[v] * muF * A;
            o.bULb[(size_t)v * nbf + j] += 0.5 * Fvb + 0.5 * n[v] * UnFb;
            o.bURb[(size_t)v * nbf + j] += 0.5 * Fvb + 0.5 * n[v] * UnFb + alpha / dPN * dnb;
            a.u[v] -= alpha / dPN * dnb;
            for (int q = 0; q < 3; ++q) a.gu[v][q] += dnb * ((n[q] - alpha * e[q]) + n[q]);
        }
        if (bc.p_type == SEMO_BC_FIXED_VALUE) {
            a.p += UnFb * alpha * t1 / dPN;
            for (int q = 0; q < 3; ++q) a.gp[q] += UnFb * alpha * t1 * e[q];
        }
    }
    store_side(o, i, a, nbf);
}

// ----------------------------------------------- LSQ + stencil bounds reverse --
// arg max / arg min over {i} + stencil(i) with the forward's tie rule (the
// first maximum in stencil order, the cell itself first)
__global__ void __launch_bounds__(TPB) k_adj_argminmax(int nc, int nfld, const int* __restrict__ st_ptr,
                                                       const int* __restrict__ st_idx, const double* __restrict__ f0,
                                                       const double* __restrict__ f1, const double* __restrict__ f2,
                                                       const double* __restrict__ f3, int* __restrict__ amx,
                                                       int* __restrict__ amn) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    const double* fs[4] = {f0, f1, f2, f3};
    for (int v = 0; v < nfld; ++v) {
        const double* ph = fs[v];
        double mx = ph[i], mn = ph[i];
        int ax = i, an = i;
        for (int q = st_ptr[i]; q < st_ptr[i + 1]; ++q) {
            int j = st_idx[q];
            double pj = ph[j];
            if (mx < pj) { mx = pj; ax = j; }
            if (pj < mn) { mn = pj; an = j; }
        }
        amx[(size_t)v * nc + i] = ax;
        amn[(size_t)v * nc + i] = an;
    }
}

// h = C g-adjoint (the packed symmetric LSQ inverse is its own transpose)
__global__ void __launch_bounds__(TPB) k_adj_lsq_h(int nc, int nfld, const double* __restrict__ lsq,
                                                   double* __restrict__ g0, double* __restrict__ g1,
                                                   double* __restrict__ g2, double* __restrict__ g3) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    const double* C = lsq + 6 * (size_t)i;
    double* gs[4] = {g0, g1, g2, g3};
    for (int v = 0; v < nfld; ++v) {
        double* g = gs[v] + 3 * (size_t)i;
        double t0 = g[0], t1 = g[1], t2 = g[2];
        g[0] = C[0] * t0 + C[1] * t1 + C[2] * t2;
        g[1] = C[1] * t0 + C[3] * t1 + C[4] * t2;
        g[2] = C[2] * t0 + C[4] * t1 + C[5] * t2;
    }
}

// transposed LSQ gradient + bound routing, gathered per cell j over its
// (symmetric) stencil.  Fields 0..2 = U, V, W (boundary values bUR), field 3
// = P (boundary values bP); nfld = 3 or 4, nb = fields with bounds.
struct LsqT {
    const double *h[4];                          // C g-adjoint per field [3n]
    const double *mxb[4], *mnb[4];               // bound adjoints per field
    const int *amx, *amn;                        // [nb][n]
    double *out[4];                              // field adjoints (+=)
};

__global__ void __launch_bounds__(TPB) k_adj_lsq_t(int nc, int nif, int nbf, int nfld, int nb,
                                                   const int* __restrict__ st_ptr, const int* __restrict__ st_idx,
                                                   const int* __restrict__ bf_ptr, const int* __restrict__ bf_idx,
                                                   const double* __restrict__ cxyz, const double* __restrict__ fxyz,
                                                   LsqT t, double* __restrict__ bURb, double* __restrict__ bPb) {
    int j = xcd_block() * blockDim.x + threadIdx.x;
    if (j >= nc) return;
    double xj = cxyz[3 * (size_t)j], yj = cxyz[3 * (size_t)j + 1], zj = cxyz[3 * (size_t)j + 2];
    double acc[4] = {0, 0, 0, 0};
    double sw[3] = {0, 0, 0};     // sum of w d over the stencil and boundary faces of j (own terms)
    for (int v = 0; v < nb; ++v) {
        if (t.amx[(size_t)v * nc + j] == j) acc[v] += t.mxb[v][j];
        if (t.amn[(size_t)v * nc + j] == j) acc[v] += t.mnb[v][j];
    }
    for (int q = st_ptr[j]; q < st_ptr[j + 1]; ++q) {
        int i = st_idx[q];
        // the term of cell i's gradient that holds phi_j: d = x_j - x_i
        double dx = xj - cxyz[3 * (size_t)i], dy = yj - cxyz[3 * (size_t)i + 1], dz = zj - cxyz[3 * (size_t)i + 2];
        double w = 1.0 / (dx * dx + dy * dy + dz * dz);
        for (int v = 0; v < nfld; ++v) {
            const double* h = t.h[v] + 3 * (size_t)i;
            acc[v] += w * (dx * h[0] + dy * h[1] + dz * h[2]);
        }
        // j's own gradient: d = x_i - x_j = -(dx, dy, dz), same weight
        sw[0] -= w * dx; sw[1] -= w * dy; sw[2] -= w * dz;
        for (int v = 0; v < nb; ++v) {
            if (t.amx[(size_t)v * nc + i] == j) acc[v] += t.mxb[v][i];
            if (t.amn[(size_t)v * nc + i] == j) acc[v] += t.mnb[v][i];
        }
    }
    for (int q = bf_ptr[j]; q < bf_ptr[j + 1]; ++q) {
        int f = bf_idx[q], b = f - nif;
        double dx = fxyz[3 * (size_t)f] - xj, dy = fxyz[3 * (size_t)f + 1] - yj, dz = fxyz[3 * (size_t)f + 2] - zj;
        double w = 1.0 / (dx * dx + dy * dy + dz * dz);
        sw[0] += w * dx; sw[1] += w * dy; sw[2] += w * dz;
        for (int v = 0; v < nfld; ++v) {
            const double* h = t.h[v] + 3 * (size_t)j;
            double s = w * (dx * h[0] + dy * h[1] + dz * h[2]);
            if (v < 3) bURb[(size_t)v * nbf + b] += s;
            else bPb[b] += s;
        }
    }
    for (int v = 0; v < nfld; ++v) {
        const double* h = t.h[v] + 3 * (size_t)j;
        acc[v] -= sw[0] * h[0] + sw[1] * h[1] + sw[2] * h[2];
        t.out[v][j] += acc[v];
    }
}

// ------------------------------------------------------ boundary recon reverse --
// k_recon_boundary reversed into the owner cell (P and U, V, W)
__global__ void __launch_bounds__(TPB) k_adj_recon(int nc, int nif, int nbf, const int* __restrict__ bf_ptr,
                                                   const int* __restrict__ bf_idx, const int* __restrict__ bpatch,
                                                   const semo_patch_bc* __restrict__ bcs,
                                                   const double* __restrict__ nrm, const double* __restrict__ U,
                                                   const double* __restrict__ V, const double* __restrict__ W,
                                                   const double* __restrict__ bULb, const double* __restrict__ bURb,
                                                   const double* __restrict__ bPb, double* __restrict__ Ub,
                                                   double* __restrict__ Vb, double* __restrict__ Wb,
                                                   double* __restrict__ Pb) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double u[3] = {0, 0, 0}, p = 0.0;
    double uc[3] = {U[i], V[i], W[i]};
    for (int q = bf_ptr[i]; q < bf_ptr[i + 1]; ++q) {
        int f = bf_idx[q], j = f - nif;
        semo_patch_bc bc = bcs[bpatch[j]];
        if (bc.p_type != SEMO_BC_FIXED_VALUE) p += bPb[j];
        double s[3];
        for (int a = 0; a < 3; ++a) s[a] = bULb[(size_t)a * nbf + j] + bURb[(size_t)a * nbf + j];
        const double* n = nrm + 3 * (size_t)f;
        switch (bc.u_type) {
            case SEMO_BC_ZERO_GRADIENT:
                for (int a = 0; a < 3; ++a) u[a] += s[a];
                break;
            case SEMO_BC_SLIP: {
                double sn = s[0] * n[0] + s[1] * n[1] + s[2] * n[2];
                for (int a = 0; a < 3; ++a) u[a] += s[a] - sn * n[a];
                break;
            }
            case SEMO_BC_INLET_OUTLET:
                if (uc[0] * n[0] + uc[1] * n[1] + uc[2] * n[2] > 0.0)
                    for (int a = 0; a < 3; ++a) u[a] += s[a];
                break;
            default: break;   // fixed values: no dependence on the cell
        }
    }
    Ub[i] += u[0]; Vb[i] += u[1]; Wb[i] += u[2]; Pb[i] += p;
}

// ------------------------------------------------------ momentum update reverse --
// U1 = clip(U + urf dU): right-hand side of the transposed solves and the
// direct part of the start-state adjoint
__global__ void __launch_bounds__(TPB) k_adj_mom_update(int nc, CaseK k, const double* __restrict__ U1,
                                                        const double* __restrict__ V1,
                                                        const double* __restrict__ W1,
                                                        const double* __restrict__ u1U,
                                                        const double* __restrict__ u1V,
                                                        const double* __restrict__ u1W, double* __restrict__ B,
                                                        double* __restrict__ nU, double* __restrict__ nV,
                                                        double* __restrict__ nW) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    size_t N = (size_t)nc;
    double b0 = fabs(U1[i]) < k.lim ? u1U[i] : 0.0;
    double b1 = fabs(V1[i]) < k.lim ? u1V[i] : 0.0;
    double b2 = fabs(W1[i]) < k.lim ? u1W[i] : 0.0;
    B[i] = k.urf_u * b0; B[N + i] = k.urf_u * b1; B[2 * N + i] = k.urf_u * b2;
    nU[i] += b0; nV[i] += b1; nW[i] += b2;
}

// momentum operator transposed in place: the pattern is the symmetric cell
// graph with one slot pair per internal face
__global__ void __launch_bounds__(TPB) k_adj_transpose(int nif, const int* __restrict__ so,
                                                       const int* __restrict__ sn, const double* __restrict__ src,
                                                       double* __restrict__ dst) {
    int f = xcd_block() * blockDim.x + threadIdx.x;
    if (f >= nif) return;
    int a = so[f], b = sn[f];
    dst[a] = src[b];
    dst[b] = src[a];
}

__global__ void k_adj_copy(size_t n, double* __restrict__ d, const double* __restrict__ s) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = s[i];
}

}  // namespace

// ====================================================================== host ==
struct AdjBufs {
    DBuf<double> ck;                                   // [n_iter][4][n] start states
    DBuf<double> aU, aV, aW, aP, nU, nV, nW, nP;       // end / start adjoints
    DBuf<double> u1U, u1V, u1W, U1;                    // U1 adjoint, U1 state [3n]
    DBuf<double> ppb, g2b, g1b, vfb, lam;
    DBuf<double> gUb, gVb, gWb, gPb, bnd;              // gradient adjoints [3n]; bounds [8n]
    DBuf<double> bULb, bURb, bPb;
    DBuf<double> mg, mbnd, mval, mdiag, mX, mbUL, mbUR, mbP;   // momentum stage copies
    DBuf<double> fin;                                  // final state [4][n]
    DBuf<int> amx, amn;
    DBuf<double> wpatch, jf, alphab;
    bool stencil_ok = false;
};

namespace {

void dcopy(semo_ctx* c, double* d, const double* s, size_t n) {
    if (n) SEMO_LAUNCH(c, "adj_copy", 16.0 * n, k_adj_copy, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), n, d, s);
}

void check_stencil_symmetric(semo_ctx* c, AdjBufs& B) {
    if (B.stencil_ok) return;
    std::vector<int> ptr(c->nc + 1), idx(c->d_st_idx.n);
    SEMO_HIP(hipMemcpyAsync(ptr.data(), c->d_st_ptr.p, ptr.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    SEMO_HIP(hipMemcpyAsync(idx.data(), c->d_st_idx.p, idx.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    SEMO_HIP(hipStreamSynchronize(c->stream));
    for (int i = 0; i < c->nc; ++i)
        for (int q = ptr[i]; q < ptr[i + 1]; ++q) {
            int j = idx[q];
            SEMO_CHECK(j != i && std::find(idx.begin() + ptr[j], idx.begin() + ptr[j + 1], i) != idx.begin() + ptr[j + 1],
                       SEMO_EINVAL, "adjoint: vertex stencil is not symmetric");
        }
    B.stencil_ok = true;
}

FaceIn stage_in(semo_ctx* c, const double* U, const double* V, const double* W, const double* P, const double* g,
                const double* bnd) {
    FaceIn in = face_in(c);
    size_t n3 = 3 * (size_t)c->nc, n = c->nc;
    in.U = U; in.V = V; in.W = W; in.P = P;
    if (g) { in.gU = g; in.gV = g + n3; in.gW = g + 2 * n3; in.gP = g + 3 * n3; }
    if (bnd) {
        in.mxU = bnd; in.mnU = bnd + n; in.mxV = bnd + 2 * n; in.mnV = bnd + 3 * n;
        in.mxW = bnd + 4 * n; in.mnW = bnd + 5 * n; in.mxP = bnd + 6 * n; in.mnP = bnd + 7 * n;
    }
    return in;
}

StageAdj stage_out(semo_ctx* c, AdjBufs& B, double* u, double* v, double* w, double* p, bool with_p) {
    size_t n = c->nc;
    StageAdj o;
    o.u = u; o.v = v; o.w = w; o.p = p;
    o.gu = B.gUb.p; o.gv = B.gVb.p; o.gw = B.gWb.p; o.gp = B.gPb.p;
    double* m = B.bnd.p;
    o.mxu = m; o.mnu = m + n; o.mxv = m + 2 * n; o.mnv = m + 3 * n; o.mxw = m + 4 * n; o.mnw = m + 5 * n;
    o.mxp = with_p ? m + 6 * n : nullptr; o.mnp = with_p ? m + 7 * n : nullptr;
    o.bULb = B.bULb.p; o.bURb = B.bURb.p; o.bPb = B.bPb.p;
    return o;
}

// reverse of grad_minmax + recon_boundary of one stage: gradient / bound /
// boundary-value adjoints in B are pushed into the stage's input fields
void lsq_recon_reverse(semo_ctx* c, AdjBufs& B, const double* U, const double* V, const double* W, const double* P,
                       bool with_p, double* uB, double* vB, double* wB, double* pB) {
    int n = c->nc;
    dim3 gb(nblk(n)), tb(TPB);
    int nb = with_p ? 4 : 3;
    SEMO_LAUNCH(c, "adj_argminmax", 4.0 * n + 4.0 * (double)c->d_st_idx.n + 16.0 * nb * n, k_adj_argminmax, gb, tb, n,
                nb, c->d_st_ptr.p, c->d_st_idx.p, U, V, W, P, B.amx.p, B.amn.p);
    SEMO_LAUNCH(c, "adj_lsq_h", 48.0 * n + 192.0 * n, k_adj_lsq_h, gb, tb, n, 4, c->c_lsq.p, B.gUb.p, B.gVb.p,
                B.gWb.p, B.gPb.p);
    LsqT t;
    size_t N = n;
    t.h[0] = B.gUb.p; t.h[1] = B.gVb.p; t.h[2] = B.gWb.p; t.h[3] = B.gPb.p;
    const double* m = B.bnd.p;
    for (int v = 0; v < 4; ++v) { t.mxb[v] = m + 2 * v * N; t.mnb[v] = m + (2 * v + 1) * N; }
    t.amx = B.amx.p; t.amn = B.amn.p;
    t.out[0] = uB; t.out[1] = vB; t.out[2] = wB; t.out[3] = pB;
    SEMO_LAUNCH(c, "adj_lsq_t",
                4.0 * n + 4.0 * (double)c->d_st_idx.n + 24.0 * n + 96.0 * n + 64.0 * n + 44.0 * c->nbf, k_adj_lsq_t,
                gb, tb, n, c->nif, c->nbf, 4, nb, c->d_st_ptr.p, c->d_st_idx.p, c->d_bf_ptr.p, c->d_bf_idx.p,
                c->c_xyz.p, c->f_xyz.p, t, B.bURb.p, B.bPb.p);
    SEMO_LAUNCH(c, "adj_recon", 4.0 * n + 160.0 * c->nbf + 64.0 * n, k_adj_recon, gb, tb, n, c->nif, c->nbf,
                c->d_bf_ptr.p, c->d_bf_idx.p, c->d_bpatch.p, c->d_bcs.p, c->f_nrm.p, U, V, W, B.bULb.p, B.bURb.p,
                B.bPb.p, uB, vB, wB, pB);
}

void zero_stage(semo_ctx* c, AdjBufs& B) {
    for (DBuf<double>* b : {&B.bULb, &B.bURb, &B.bPb}) b->zero(c->stream);
}

}  // namespace

void adjoint_free(AdjBufs* b) { delete b; }

// J and dJ/dalpha_p (device numbering) of n_iter SIMPLE iterations run from
// the current state; leaves the context in the forward's final state
double adjoint_run(semo_ctx* c, int n_iter, const double* wpatch_h, double* grad_dev_out) {
    SEMO_CHECK(c->cs.sigma == 0.0 && c->cs.g[0] == 0.0 && c->cs.g[1] == 0.0 && c->cs.g[2] == 0.0, SEMO_EINVAL,
               "adjoint: surface tension and gravity are not differentiated (set sigma = 0, g = 0)");
    SEMO_CHECK(c->cs.n_vof == 0 && !c->cs.adjust_dt, SEMO_EINVAL,
               "adjoint: single phase with a fixed time step only (n_vof = 0, adjust_dt = 0)");
    SEMO_CHECK(c->cs.rho[0] == c->cs.rho[1] && c->cs.mu[0] == c->cs.mu[1], SEMO_EINVAL,
               "adjoint: both phases must have the same rho and mu");
    for (const semo_patch_bc& b : c->bcs)
        SEMO_CHECK(b.p_type != SEMO_BC_SWITCH, SEMO_EINVAL_BC,
                   "adjoint: the Mach-switched pressure boundary is not differentiated");
    if (!c->adj) c->adj = new AdjBufs();
    AdjBufs& B = *c->adj;
    check_stencil_symmetric(c, B);
    int n = c->nc, nbf = c->nbf, nif = c->nif;
    size_t N = n;
    if (c->alphaP.n != N) { c->alphaP.alloc(N); c->alphaP.zero(c->stream); }
    B.ck.ensure(4 * N * (size_t)n_iter);
    for (DBuf<double>* b : {&B.aU, &B.aV, &B.aW, &B.aP, &B.nU, &B.nV, &B.nW, &B.nP, &B.u1U, &B.u1V, &B.u1W, &B.ppb,
                            &B.lam, &B.alphab, &B.mdiag, &B.mbP})
        b->ensure(N + 64);
    for (DBuf<double>* b : {&B.U1, &B.g2b, &B.g1b, &B.gUb, &B.gVb, &B.gWb, &B.gPb, &B.mX}) b->ensure(3 * N + 64);
    B.bnd.ensure(8 * N); B.mbnd.ensure(8 * N); B.mg.ensure(12 * N); B.fin.ensure(4 * N);
    B.amx.ensure(4 * N); B.amn.ensure(4 * N);
    for (DBuf<double>* b : {&B.bULb, &B.bURb, &B.mbUL, &B.mbUR}) b->ensure(3 * (size_t)std::max(nbf, 1));
    B.bPb.ensure(std::max(nbf, 1)); B.mbP.ensure(std::max(nbf, 1)); B.jf.ensure(std::max(nbf, 1));
    B.vfb.ensure(std::max(nif, 1));
    B.mval.ensure(c->Amat.val.n);
    B.wpatch.upload(std::vector<double>(wpatch_h, wpatch_h + c->npatch), c->stream);
    CaseK k = case_k(c);
    dim3 gb(nblk(n)), tb(TPB);
    semo_solve_stats st{};

    // ---- forward with checkpoints (semo_incompressible_pressure_based) ----
    simple_setup(c);
    for (int it = 0; it < n_iter; ++it) {
        double* ck = B.ck.p + 4 * N * (size_t)it;
        dcopy(c, ck, c->U.p, N); dcopy(c, ck + N, c->V.p, N); dcopy(c, ck + 2 * N, c->W.p, N);
        dcopy(c, ck + 3 * N, c->P.p, N);
        simple_momentum(c, &c->last_stats_mom);
        simple_pressure(c, &c->last_stats_p);
    }
    c->iterReal++;
    // objective of the final state
    double J = 0.0;
    if (nbf) {
        SEMO_LAUNCH(c, "adj_objective", 40.0 * nbf, k_adj_objective, dim3(nblk(nbf)), tb, nbf, nif, c->d_bpatch.p,
                    c->d_bcs.p, B.wpatch.p, c->d_owner.p, c->f_area.p, c->P.p, B.jf.p);
        std::vector<double> h(nbf);
        SEMO_HIP(hipMemcpyAsync(h.data(), B.jf.p, nbf * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        SEMO_HIP(hipStreamSynchronize(c->stream));
        for (double v : h) J += v;
    }
    // the final state, restored after the sweep
    DBuf<double>& fin = B.fin;
    dcopy(c, fin.p, c->U.p, N); dcopy(c, fin.p + N, c->V.p, N); dcopy(c, fin.p + 2 * N, c->W.p, N);
    dcopy(c, fin.p + 3 * N, c->P.p, N);
    // ---- seed ----
    for (DBuf<double>* b : {&B.aU, &B.aV, &B.aW, &B.nU, &B.nV, &B.nW, &B.nP, &B.alphab}) b->zero(c->stream);
    SEMO_LAUNCH(c, "adj_seed", 40.0 * nbf + 8.0 * n, k_adj_seed, gb, tb, n, nif, c->d_bf_ptr.p, c->d_bf_idx.p,
                c->d_bpatch.p, c->d_bcs.p, B.wpatch.p, c->f_area.p, B.aP.p);
    // ---- reverse sweep ----
    for (int it = n_iter - 1; it >= 0; --it) {
        const double* ck = B.ck.p + 4 * N * (size_t)it;
        const double* nx = it + 1 < n_iter ? B.ck.p + 4 * N * (size_t)(it + 1) : fin.p;
        const double *U0 = ck, *V0 = ck + N, *W0 = ck + 2 * N, *P0 = ck + 3 * N;
        dcopy(c, c->U.p, U0, N); dcopy(c, c->V.p, V0, N); dcopy(c, c->W.p, W0, N); dcopy(c, c->P.p, P0, N);
        // momentum stage, recomputed
        recon_boundary(c);
        assemble_momentum(c);
        dcopy(c, B.mval.p, c->Amat.val.p, c->Amat.val.n);
        dcopy(c, B.mdiag.p, c->Amat.diag.p, N);
        dcopy(c, B.mg.p, c->gU.p, 3 * N); dcopy(c, B.mg.p + 3 * N, c->gV.p, 3 * N);
        dcopy(c, B.mg.p + 6 * N, c->gW.p, 3 * N); dcopy(c, B.mg.p + 9 * N, c->gP.p, 3 * N);
        const double* bsrc[8] = {c->mxU.p, c->mnU.p, c->mxV.p, c->mnV.p, c->mxW.p, c->mnW.p, c->mxP.p, c->mnP.p};
        for (int q = 0; q < 8; ++q) dcopy(c, B.mbnd.p + q * N, bsrc[q], N);
        if (nbf) {
            dcopy(c, B.mbUL.p, c->bUL.p, 3 * (size_t)nbf); dcopy(c, B.mbUR.p, c->bUR.p, 3 * (size_t)nbf);
            dcopy(c, B.mbP.p, c->bP.p, nbf);
        }
        solve_momentum(c, 3, &st);
        dcopy(c, B.mX.p, c->Xvec.p, 3 * N);
        momentum_update(c);
        dcopy(c, B.U1.p, c->U.p, N); dcopy(c, B.U1.p + N, c->V.p, N); dcopy(c, B.U1.p + 2 * N, c->W.p, N);
        // pressure stage, recomputed up to the assembled system
        recon_boundary(c);
        assemble_pressure(c);
        // -- reverse of the p / U correction and the two Green-Gauss passes --
        SEMO_LAUNCH(c, "adj_pcorrect", 128.0 * n, k_adj_pcorrect, gb, tb, n, k, c->Rho.p, nx + 3 * N, nx, nx + N,
                    nx + 2 * N, B.aP.p, B.aU.p, B.aV.p, B.aW.p, B.ppb.p, B.nP.p, B.u1U.p, B.u1V.p, B.u1W.p, B.g2b.p);
        for (int pass = 1; pass >= 0; --pass) {
            double* gcur = pass == 1 ? B.g2b.p : B.g1b.p;
            double* gprev = pass == 1 ? B.g1b.p : nullptr;
            if (nif)
                SEMO_LAUNCH(c, "adj_gg_face", 64.0 * nif + 32.0 * n, k_adj_gg_face, dim3(nblk(nif)), tb, nif,
                            c->d_owner.p, c->d_nbr.p, c->f_nrm.p, c->f_area.p, c->c_vol.p, gcur, B.vfb.p);
            SEMO_LAUNCH(c, "adj_gg_cell", cell_faces_bytes(c, 72.0) + 80.0 * n, k_adj_gg_cell, gb, tb, n, nif,
                        c->d_cf_sl.p, c->d_cf_ell.p, c->d_bpatch.p, c->d_bcs.p, c->f_nrm.p, c->f_area.p, c->c_vol.p,
                        c->f_wC.p, c->f_skew.p, gcur, B.vfb.p, B.ppb.p, gprev);
        }
        // pressure adjoint: the Laplacian is symmetric
        solve_pressure(c, B.ppb.p, B.lam.p, &st);
        // pressure assembly reverse into U1 / P
        zero_stage(c, B);
        {
            FaceIn in = stage_in(c, B.U1.p, B.U1.p + N, B.U1.p + 2 * N, P0, nullptr, nullptr);
            SEMO_LAUNCH(c, "adj_p_cell", cell_faces_bytes(c, 64.0) + 96.0 * n, k_adj_p_cell, gb, tb, n, nif, nbf, k,
                        in, c->d_cf_sl.p, c->d_cf_ell.p, c->d_bpatch.p, c->d_bcs.p, c->bRho.p, B.lam.p,
                        stage_out(c, B, B.u1U.p, B.u1V.p, B.u1W.p, B.nP.p, false));
        }
        lsq_recon_reverse(c, B, B.U1.p, B.U1.p + N, B.U1.p + 2 * N, P0, false, B.u1U.p, B.u1V.p, B.u1W.p, B.nP.p);
        // momentum update reverse -> transposed momentum solves
        SEMO_LAUNCH(c, "adj_mom_update", 80.0 * n, k_adj_mom_update, gb, tb, n, k, B.U1.p, B.U1.p + N, B.U1.p + 2 * N,
                    B.u1U.p, B.u1V.p, B.u1W.p, c->Bvec.p, B.nU.p, B.nV.p, B.nW.p);
        dcopy(c, c->Amat.val.p, B.mval.p, c->Amat.val.n);
        if (nif)
            SEMO_LAUNCH(c, "adj_transpose", 40.0 * nif, k_adj_transpose, dim3(nblk(nif)), tb, nif, c->d_slot_own.p,
                        c->d_slot_nbr.p, B.mval.p, c->Amat.val.p);
        dcopy(c, c->Amat.diag.p, B.mdiag.p, N);
        solve_momentum(c, 3, &st);
        // momentum assembly reverse into U / P / alpha_p
        zero_stage(c, B);
        {
            FaceIn in = stage_in(c, U0, V0, W0, P0, B.mg.p, B.mbnd.p);
            SEMO_LAUNCH(c, "adj_mom_cell", cell_faces_bytes(c, 96.0) + 200.0 * n, k_adj_mom_cell, gb, tb, n, nif, nbf,
                        k, in, c->d_cf_sl.p, c->d_cf_ell.p, c->d_bpatch.p, c->d_bcs.p, c->c_vol.p, c->Mu.p,
                        c->alphaP.p, B.mbUL.p, B.mbUR.p, B.mbP.p, c->bMu.p, c->bRho.p, c->Xvec.p, B.mX.p,
                        stage_out(c, B, B.nU.p, B.nV.p, B.nW.p, B.nP.p, true), B.alphab.p);
        }
        lsq_recon_reverse(c, B, U0, V0, W0, P0, true, B.nU.p, B.nV.p, B.nW.p, B.nP.p);
        // the start adjoint becomes the next (earlier) iteration's end adjoint
        std::swap(B.aU, B.nU); std::swap(B.aV, B.nV); std::swap(B.aW, B.nW); std::swap(B.aP, B.nP);
        for (DBuf<double>* b : {&B.nU, &B.nV, &B.nW, &B.nP}) b->zero(c->stream);
    }
    // leave the forward's final state in place
    dcopy(c, c->U.p, fin.p, N); dcopy(c, c->V.p, fin.p + N, N); dcopy(c, c->W.p, fin.p + 2 * N, N);
    dcopy(c, c->P.p, fin.p + 3 * N, N);
    if (grad_dev_out) dcopy(c, grad_dev_out, B.alphab.p, N);
    SEMO_HIP(hipStreamSynchronize(c->stream));
    return J;
}

double objective(semo_ctx* c, const double* wpatch_h) {
    if (!c->adj) c->adj = new AdjBufs();
    AdjBufs& B = *c->adj;
    int nbf = c->nbf;
    if (!nbf) return 0.0;
    B.jf.ensure(nbf);
    B.wpatch.upload(std::vector<double>(wpatch_h, wpatch_h + c->npatch), c->stream);
    SEMO_LAUNCH(c, "adj_objective", 40.0 * nbf, k_adj_objective, dim3(nblk(nbf)), dim3(TPB), nbf, c->nif,
                c->d_bpatch.p, c->d_bcs.p, B.wpatch.p, c->d_owner.p, c->f_area.p, c->P.p, B.jf.p);
    std::vector<double> h(nbf);
    SEMO_HIP(hipMemcpyAsync(h.data(), B.jf.p, nbf * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SEMO_HIP(hipStreamSynchronize(c->stream));
    double J = 0.0;
    for (double v : h) J += v;
    return J;
}

}  // namespace semo
