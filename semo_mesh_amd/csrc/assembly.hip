// Per-iteration assembly of the pressure-based incompressible path on the
// device.
//
// Layout: every cell field is a dense FP64 array in RCM order; boundary-face
// values are dense arrays over the boundary faces only (internal faces take
// zero-order owner/neighbour values, reconIncom.cpp:512-533, so they are never
// stored).  Matrix coefficients go straight into the static sliced-ELL
// pattern through precomputed face -> slot maps.
//
// Determinism: no floating-point atomics.  A face kernel computes each
// internal face's contributions once and stores them; a cell kernel then
// accumulates, for each cell, its faces in ascending face order -- the same
// per-cell order in which SEMO's face loops scatter (eqMomentum.cpp:331-1008,
// eqPressure.cpp:246-673) -- and adds the boundary faces inline.
#include <climits>
#include <mutex>

#include "common.hpp"

#include "kern.hpp"

namespace semo {
namespace {

// --------------------------------------------------------------------- EOS --
// calcIncomCellEOSVF (eos.cpp:10-151) + calcCellTransport (transport.cpp:9-98)
__global__ void __launch_bounds__(TPB) k_eos(int nc, CaseK k, double* __restrict__ VF,
                                             double* __restrict__ Rho, double* __restrict__ Mu) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double a = clip01(VF[i]);
    double v0 = clip01(a), v1 = clip01(1.0 - a);
    double r = 0.0; r += v0 * k.rho0; r += v1 * k.rho1;
    Rho[i] = r;
    VF[i] = v0;
    double m = 0.0; m += v0 * k.mu0; m += (1.0 - v0) * k.mu1;
    Mu[i] = m + 0.0;
}

// Speed of sound of the constant-density mixture, as calcCellEOSVF sets
// cell.var[C] at load (eos.cpp:160-293): every "const" phase has drho/dP =
// drho/dT = dh/dP = dh/dT = 1e8 (eos.cpp:184-192), mixed by volume (rho
// derivatives) and mass fraction (enthalpy derivatives).
__global__ void __launch_bounds__(TPB) k_sound_speed(int nc, CaseK k, const double* __restrict__ VF,
                                                     double* __restrict__ Cs) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double a = clip01(VF[i]);
    double v0 = clip01(a), v1 = clip01(1.0 - a);
    double rho = 0.0; rho += v0 * k.rho0; rho += v1 * k.rho1;
    double m0 = clip01(v0 * k.rho0 / rho), m1 = clip01(1.0 - m0);
    double drdp = 0.0, drdt = 0.0, dhdp = 0.0, dhdt = 0.0;
    drdp += v0 * 1.0e+8; drdt += v0 * 1.0e+8; dhdp += m0 * 1.0e+8; dhdt += m0 * 1.0e+8;
    drdp += v1 * 1.0e+8; drdt += v1 * 1.0e+8; dhdp += m1 * 1.0e+8; dhdt += m1 * 1.0e+8;
    double c2 = drdp + 1.0 / rho * drdt / dhdt * (1.0 - rho * dhdp);
    Cs[i] = sqrt(1.0 / c2);
}

// Mach number |U|/C of cell c, the P "switch" test (reconIncom.cpp:620-625,
// eqPressure.cpp:626-630)
__device__ inline double cell_mach(int c, const double* __restrict__ U, const double* __restrict__ V,
                                   const double* __restrict__ W, const double* __restrict__ Cs) {
    double u = U[c], v = V[c], w = W[c];
    return sqrt(u * u + v * v + w * w) / Cs[c];
}

// VF "function" boundary (reconIncom.cpp:929-1019): the VF of the nearest of
// the given points to each boundary face centroid; the first point of the
// smallest squared distance below 1e9 wins (none: 0)
__global__ void __launch_bounds__(TPB) k_vf_nearest(int nbf, int nif, const double* __restrict__ fxyz, int np,
                                                    const double* __restrict__ p4, double* __restrict__ out) {
    int j = xcd_block() * blockDim.x + threadIdx.x;
    if (j >= nbf) return;
    const double* x = fxyz + 3 * ((size_t)nif + j);
    double x0 = x[0], x1 = x[1], x2 = x[2];
    double best = 1.e9, v = 0.0;
    for (int q = 0; q < np; ++q) {
        const double* p = p4 + 4 * (size_t)q;
        double d0 = x0 - p[1], d1 = x1 - p[2], d2 = x2 - p[3];
        double dist = (d0 * d0 + d1 * d1 + d2 * d2);
        if (dist < best) { best = dist; v = p[0]; }
    }
    out[j] = v;
}

// flags[0] / flags[1]: some boundary P / VF value differs from the saved one
// (a null current array skips that comparison)
__global__ void __launch_bounds__(TPB) k_bchanged(int nbf, const double* __restrict__ bP,
                                                  const double* __restrict__ bP_g, const double* __restrict__ bVF,
                                                  const double* __restrict__ bVF_g, int* __restrict__ flags) {
    int j = xcd_block() * blockDim.x + threadIdx.x;
    if (j >= nbf) return;
    if (bP && bP[j] != bP_g[j]) flags[0] = 1;
    if (bVF && bVF[j] != bVF_g[j]) flags[1] = 1;
}
__global__ void k_clear2(int* __restrict__ flags) { flags[0] = 0; flags[1] = 0; }

// ------------------------------------------------------ boundary face values --
// reconIncomZeroOrder boundary dispatch (reconIncom.cpp:591-1032) and the face
// density of calcOtherDataIncomFromEOS (reconIncom.cpp:445-493)
__global__ void __launch_bounds__(TPB) k_recon_boundary(
    int nbf, int nif, CaseK k, const int* __restrict__ bpatch, const semo_patch_bc* __restrict__ bcs,
    const int* __restrict__ owner, const double* __restrict__ nrm, const double* __restrict__ fxyz,
    const double* __restrict__ P, const double* __restrict__ U, const double* __restrict__ V,
    const double* __restrict__ W, const double* __restrict__ VF, const double* __restrict__ Mu,
    const double* __restrict__ ufix, const double* __restrict__ Cs, const double* __restrict__ vffun,
    double* __restrict__ bP, double* __restrict__ bUL, double* __restrict__ bUR,
    double* __restrict__ bVF, double* __restrict__ bMu, double* __restrict__ bRho) {
    int j = xcd_block() * blockDim.x + threadIdx.x;
    if (j >= nbf) return;
    int f = nif + j;
    int c = owner[f];
    semo_patch_bc bc = bcs[bpatch[j]];
    const double* n = nrm + 3 * (size_t)f;
    if (bc.p_type == SEMO_BC_FIXED_VALUE) bP[j] = bc.p_value;
    else if (bc.p_type == SEMO_BC_SWITCH) bP[j] = cell_mach(c, U, V, W, Cs) > 1.0 ? P[c] : bc.p_value;
    else bP[j] = P[c];
    double uL[3], uR[3];
    double uc[3] = {U[c], V[c], W[c]};
    switch (bc.u_type) {
        case SEMO_BC_FIXED_VALUE: {
            // per-face values when given, else the patch's uniform value
            for (int a = 0; a < 3; ++a) uL[a] = uR[a] = ufix ? ufix[a * (size_t)nbf + j] : bc.u_value[a];
            // live orifice profile on fixedValue-U faces (reconIncom.cpp:707-727)
            double fx = fxyz[3 * (size_t)f];
            for (int q = 0; q < 24; ++q) {
                double cen = 0.0006 * (double)q;
                if (fx > cen - 0.00016 && fx < cen + 0.00016) {
                    double th = atan((fxyz[3 * (size_t)f + 2] - 0.001) / (fx - cen));
                    uL[0] = 0.0 * cos(th) - 0.1 * sin(th);
                    uL[1] = -0.1;
                    uL[2] = 0.0 * sin(th) + 0.1 * cos(th);
                }
            }
            break;
        }
        case SEMO_BC_ZERO_GRADIENT:
            for (int a = 0; a < 3; ++a) uL[a] = uR[a] = uc[a];
            break;
        case SEMO_BC_SLIP: {
            double nv = uc[0] * n[0] + uc[1] * n[1] + uc[2] * n[2];
            for (int a = 0; a < 3; ++a) uL[a] = uR[a] = uc[a] - nv * n[a];
            break;
        }
        case SEMO_BC_SURFACE_NORMAL_FIXED_VALUE:
            for (int a = 0; a < 3; ++a) uL[a] = uR[a] = bc.u_value[0] * n[a];
            break;
        case SEMO_BC_INLET_OUTLET: {
            double own = uc[0] * n[0] + uc[1] * n[1] + uc[2] * n[2];
            for (int a = 0; a < 3; ++a) uL[a] = uR[a] = own > 0.0 ? uc[a] : bc.u_value[0] * n[a];
            break;
        }
        default:   // noSlip
            for (int a = 0; a < 3; ++a) uL[a] = uR[a] = 0.0;
    }
    for (int a = 0; a < 3; ++a) { bUL[a * (size_t)nbf + j] = uL[a]; bUR[a * (size_t)nbf + j] = uR[a]; }
    double vf;
    if (bc.vf_type == SEMO_BC_FIXED_VALUE) vf = bc.vf_value;
    else if (bc.vf_type == SEMO_BC_FUNCTION) vf = vffun[j];
    else if (bc.vf_type == SEMO_BC_INLET_OUTLET) {
        double own = uc[0] * n[0] + uc[1] * n[1] + uc[2] * n[2];
        vf = own > 0.0 ? VF[c] : bc.vf_value;
    } else vf = VF[c];
    bVF[j] = vf;
    bMu[j] = Mu[c];
    double v0 = clip01(vf), v1 = clip01(1.0 - vf);
    double r = 0.0; r += v0 * k.rho0; r += v1 * k.rho1;
    bRho[j] = r;
}

__device__ inline double mix_rho(double a, const CaseK& k) {
    double v0 = clip01(a), v1 = clip01(1.0 - a);
    double r = 0.0; r += v0 * k.rho0; r += v1 * k.rho1;
    return r;
}

// ------------------------------------------------------------- gravity src --
// calcSourceGravity (source/gravity.cpp:7-75), gathered per cell
__global__ void __launch_bounds__(TPB) k_gravity(
    int nc, int nif, CaseK k, const int* __restrict__ cf_ptr, const int* __restrict__ cf_idx,
    const int* __restrict__ owner, const int* __restrict__ nbr, const double* __restrict__ nrm,
    const double* __restrict__ area, const double* __restrict__ wCv, const double* __restrict__ pf,
    const double* __restrict__ nfv, const double* __restrict__ vol, const double* __restrict__ U,
    const double* __restrict__ V, const double* __restrict__ W, const double* __restrict__ VF,
    const double* __restrict__ bUL, const double* __restrict__ bUR, const double* __restrict__ bVF,
    int nbf, double* __restrict__ srcG) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double s[4] = {0, 0, 0, 0};
    double vi = vol[i];
    for (int q = cf_ptr[i]; q < cf_ptr[i + 1]; ++q) {
        int f = cf_idx[q];
        const double* n = nrm + 3 * (size_t)f;
        double wCL = wCv[f], wCR = 1.0 - wCL;
        double uL[3], uR[3], aL, aR;
        if (f < nif) {
            int l = owner[f], r = nbr[f];
            uL[0] = U[l]; uL[1] = V[l]; uL[2] = W[l]; uR[0] = U[r]; uR[1] = V[r]; uR[2] = W[r];
            aL = VF[l]; aR = VF[r];
        } else {
            int j = f - nif;
            for (int a = 0; a < 3; ++a) { uL[a] = bUL[a * (size_t)nbf + j]; uR[a] = bUR[a * (size_t)nbf + j]; }
            aL = aR = bVF[j];
        }
        double UnL = uL[0] * n[0] + uL[1] * n[1] + uL[2] * n[2];
        double UnR = uR[0] * n[0] + uR[1] * n[1] + uR[2] * n[2];
        double UnF = wCL * UnL + wCR * UnR;
        double RhoF = 1.0 / (wCL / mix_rho(aL, k) + wCR / mix_rho(aR, k));
        double A = area[f];
        if (owner[f] == i) {
            double gL = 0.0;
            gL += k.g0 * pf[3 * (size_t)f]; gL += k.g1 * pf[3 * (size_t)f + 1]; gL += k.g2 * pf[3 * (size_t)f + 2];
            for (int a = 0; a < 3; ++a) s[a] += ((RhoF * gL * n[a]) * A / vi);
            s[3] += ((RhoF * gL * UnF) * A / vi);
        } else {
            double gR = 0.0;
            gR += k.g0 * nfv[3 * (size_t)f]; gR += k.g1 * nfv[3 * (size_t)f + 1]; gR += k.g2 * nfv[3 * (size_t)f + 2];
            for (int a = 0; a < 3; ++a) s[a] -= ((RhoF * gR * n[a]) * A / vi);
            s[3] -= ((RhoF * gR * UnF) * A / vi);
        }
    }
    for (int a = 0; a < 4; ++a) srcG[a * (size_t)nc + i] = s[a];
}

// ---------------------------------------------- LSQ gradients + stencil bounds --
// calcLeastSquare "cellVertex" "1st" "cell" (gradient.cpp:761-816 stencil,
// 997-1083 boundary faces, 1136-1159 apply) for P, U, V, W, VF in one pass,
// fused with setCellVarMinMax (cellVarMinMax.cpp:15-24) of U, V, W (and P).
//
// Layout: each cell's centroid and the five fields sit in one 64-byte record
// cpk[8i..8i+7] = [x,y,z,U,V,W,P,VF], so a stencil neighbour costs one cache
// line, and the three-field pass (U, V, W) reads only its first 48 bytes; the stencil is sliced ELL (lane = cell inside a 64-cell slice) so the
// index loads of a wavefront are contiguous.  Summation order is the stencil
// order, then the cell's boundary faces in ascending order (as the reference).
__global__ void __launch_bounds__(TPB) k_pack5(int nc, const double* __restrict__ P, const double* __restrict__ U,
                                               const double* __restrict__ V, const double* __restrict__ W,
                                               const double* __restrict__ VF, double* __restrict__ cpk) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double* o = cpk + 8 * (size_t)i;
    o[3] = U[i]; o[4] = V[i]; o[5] = W[i]; o[6] = P[i]; o[7] = VF[i];
}

// U, V, W only (the pressure stage: P and VF are unchanged since the last pack)
__global__ void __launch_bounds__(TPB) k_pack3(int nc, const double* __restrict__ U, const double* __restrict__ V,
                                               const double* __restrict__ W, double* __restrict__ cpk) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double* o = cpk + 8 * (size_t)i;
    o[3] = U[i]; o[4] = V[i]; o[5] = W[i];
}

__global__ void __launch_bounds__(TPB) k_pack_xyz(int nc, const double* __restrict__ cxyz, double* __restrict__ cpk) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double* o = cpk + 8 * (size_t)i;
    o[0] = cxyz[3 * (size_t)i]; o[1] = cxyz[3 * (size_t)i + 1]; o[2] = cxyz[3 * (size_t)i + 2];
}

// NF = 5: P, U, V, W, VF.  NF = 3: U, V, W only (the pressure stage of an
// outer iteration, whose P and VF are those of its momentum stage, so their
// gradients from that stage are reused bit for bit); bounds of U, V, W, and
// of P when with_p.
//
// The stencil is read as sliced ELL: int32 cell ids (IDX = int, -1 pads) or,
// when every neighbour lies within +-32767 of its cell, int16 offsets
// neighbour - cell (IDX = int16_t, -32768 pads).  The stencil indices, the
// LSQ coefficients and the outputs are touched once per launch and move with
// non-temporal loads and stores, so that they do not push the neighbour
// records, which every cell's stencil re-reads, out of the L2.
template <int NF, class IDX, int UNR = 2>
__global__ void __launch_bounds__(TPB) k_grad_minmax(
    int nc, int nif, int nbf, int with_p, const int* __restrict__ st_sl, const IDX* __restrict__ st_ell,
    const int* __restrict__ bf_ptr, const int* __restrict__ bf_idx, const double* __restrict__ cpk,
    const double* __restrict__ fxyz, const double* __restrict__ lsq,
    const double* __restrict__ bP, const double* __restrict__ bUR,
    const double* __restrict__ bVF, double* __restrict__ gP, double* __restrict__ gU,
    double* __restrict__ gV, double* __restrict__ gW, double* __restrict__ gVF,
    double* __restrict__ mxP, double* __restrict__ mnP, double* __restrict__ mxU, double* __restrict__ mnU,
    double* __restrict__ mxV, double* __restrict__ mnV, double* __restrict__ mxW, double* __restrict__ mnW) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    const double2* me = reinterpret_cast<const double2*>(cpk + 8 * (size_t)i);
    double2 m0 = me[0], m1 = me[1], m2 = me[2], m3 = NF == 5 ? me[3] : make_double2(0.0, 0.0);
    double xi = m0.x, yi = m0.y, zi = m1.x;
    double ph[5] = {m3.x, m1.y, m2.x, m2.y, m3.y};
    double t[5][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double mx[4] = {ph[0], ph[1], ph[2], ph[3]}, mn[4] = {ph[0], ph[1], ph[2], ph[3]};
    int s = i >> 6, lane = i & 63;
    int e0 = st_sl[s], e1 = st_sl[s + 1];
    // one stencil neighbour's terms, in stencil order
    auto add = [&](const double2& o0, const double2& o1, const double2& o2, const double2& o3) {
        double dx = o0.x - xi, dy = o0.y - yi, dz = o1.x - zi;
        double wk = 1.0 / (dx * dx + dy * dy + dz * dz);
        double pj[5] = {o3.x, o1.y, o2.x, o2.y, o3.y};
#pragma unroll
        for (int v = (NF == 5 ? 0 : 1); v < (NF == 5 ? 5 : 4); ++v) {
            double d = pj[v] - ph[v];
            t[v][0] += wk * dx * d; t[v][1] += wk * dy * d; t[v][2] += wk * dz * d;
        }
#pragma unroll
        for (int v = (NF == 5 ? 0 : 1); v < 4; ++v) { mx[v] = smax(mx[v], pj[v]); mn[v] = smin(mn[v], pj[v]); }
    };
    // UNR neighbours per trip, all records requested before the first is
    // summed, and the next trip's stencil indices requested before this
    // trip's sums.  Padding (-1) follows a cell's last neighbour.  (Keeping
    // the record loads a whole trip ahead measured slower: the compiler then
    // waits for every outstanding load at the top of the loop.)
    auto nbr = [&](int k) -> int {
        int v = __builtin_nontemporal_load(st_ell + k);
        if constexpr (sizeof(IDX) == 2) return v == -32768 ? -1 : i + v;
        else return v;
    };
    int e = e0 + lane;
    int jn[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) jn[u] = e + 64 * u < e1 ? nbr(e + 64 * u) : -1;
    while (jn[0] >= 0) {
        int j[UNR];
        double2 rec[UNR][4];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            j[u] = jn[u];
            if (j[u] >= 0) {
                const double2* o = reinterpret_cast<const double2*>(cpk + 8 * (size_t)j[u]);
                rec[u][0] = o[0]; rec[u][1] = o[1]; rec[u][2] = o[2];
                rec[u][3] = NF == 5 ? o[3] : make_double2(0.0, 0.0);
            }
        }
        e += 64 * UNR;
#pragma unroll
        for (int u = 0; u < UNR; ++u) jn[u] = (j[UNR - 1] >= 0 && e + 64 * u < e1) ? nbr(e + 64 * u) : -1;
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            if (j[u] < 0) break;
            add(rec[u][0], rec[u][1], rec[u][2], rec[u][3]);
        }
    }
    for (int q = bf_ptr[i]; q < bf_ptr[i + 1]; ++q) {
        int f = bf_idx[q];
        int j = f - nif;
        double dx = fxyz[3 * (size_t)f] - xi, dy = fxyz[3 * (size_t)f + 1] - yi, dz = fxyz[3 * (size_t)f + 2] - zi;
        double wk = 1.0 / (dx * dx + dy * dy + dz * dz);
        double fr[5] = {NF == 5 ? bP[j] : 0.0, bUR[j], bUR[(size_t)nbf + j], bUR[2 * (size_t)nbf + j],
                        NF == 5 ? bVF[j] : 0.0};
#pragma unroll
        for (int v = (NF == 5 ? 0 : 1); v < (NF == 5 ? 5 : 4); ++v) {
            double d = fr[v] - ph[v];
            t[v][0] += wk * dx * d; t[v][1] += wk * dy * d; t[v][2] += wk * dz * d;
        }
    }
    const double* Cp = lsq + 6 * (size_t)i;
    double C[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) C[k] = __builtin_nontemporal_load(Cp + k);
    double* out[5] = {gP, gU, gV, gW, gVF};
#pragma unroll
    for (int v = (NF == 5 ? 0 : 1); v < (NF == 5 ? 5 : 4); ++v) {
        double r0 = 0.0; r0 += C[0] * t[v][0]; r0 += C[1] * t[v][1]; r0 += C[2] * t[v][2];
        double r1 = 0.0; r1 += C[1] * t[v][0]; r1 += C[3] * t[v][1]; r1 += C[4] * t[v][2];
        double r2 = 0.0; r2 += C[2] * t[v][0]; r2 += C[4] * t[v][1]; r2 += C[5] * t[v][2];
        __builtin_nontemporal_store(r0, out[v] + 3 * (size_t)i);
        __builtin_nontemporal_store(r1, out[v] + 3 * (size_t)i + 1);
        __builtin_nontemporal_store(r2, out[v] + 3 * (size_t)i + 2);
    }
    if (NF == 5 && with_p) { __builtin_nontemporal_store(mx[0], mxP + i); __builtin_nontemporal_store(mn[0], mnP + i); }
    __builtin_nontemporal_store(mx[1], mxU + i); __builtin_nontemporal_store(mn[1], mnU + i);
    __builtin_nontemporal_store(mx[2], mxV + i); __builtin_nontemporal_store(mn[2], mnV + i);
    __builtin_nontemporal_store(mx[3], mxW + i); __builtin_nontemporal_store(mn[3], mnW + i);
}

// one field's LSQ gradient, calcLeastSquare "cellVertex" (gradient.cpp:
// 761-1159): the vertex stencil read as sliced ELL (int32 ids or int16
// offsets, as in k_grad_minmax) over packed records xph = [x, y, z, phi] of
// 32 B, two neighbours per trip, then the cell's boundary faces in face
// order.  faceR: boundary values indexed by face - foff; null: "input" mode
// with cn = -1, a zero boundary difference (gradient.cpp:1018-1022).  The
// sums run in stencil order, as the reference's.
// gate (optional): skip the pass unless *gate != 0
__global__ void __launch_bounds__(TPB) k_pack_xphi(int nc, const double* __restrict__ cxyz,
                                                   const double* __restrict__ phi, double* __restrict__ xph,
                                                   const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double2* o = reinterpret_cast<double2*>(xph + 4 * (size_t)i);
    o[0] = make_double2(cxyz[3 * (size_t)i], cxyz[3 * (size_t)i + 1]);
    o[1] = make_double2(cxyz[3 * (size_t)i + 2], phi[i]);
}

// (8 neighbours per trip measured best of 2/4/8 on the 1.3 M-tet dam break,
// whose vertex stencils are large)
template <class IDX, int UNR = 8>
__global__ void __launch_bounds__(TPB) k_lsq1(
    int nc, const int* __restrict__ st_sl, const IDX* __restrict__ st_ell, const int* __restrict__ bf_ptr,
    const int* __restrict__ bf_idx, const double* __restrict__ xph, const double* __restrict__ fxyz,
    const double* __restrict__ lsq, const double* __restrict__ faceR, int foff, double* __restrict__ g,
    const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    const double2* me = reinterpret_cast<const double2*>(xph + 4 * (size_t)i);
    double2 m0 = me[0], m1 = me[1];
    double xi = m0.x, yi = m0.y, zi = m1.x, p = m1.y;
    double t0 = 0, t1 = 0, t2 = 0;
    int s = i >> 6, lane = i & 63;
    int e0 = st_sl[s], e1 = st_sl[s + 1];
    auto nbr = [&](int k) -> int {
        int v = __builtin_nontemporal_load(st_ell + k);
        if constexpr (sizeof(IDX) == 2) return v == -32768 ? -1 : i + v;
        else return v;
    };
    // UNR neighbours per trip, records requested before the first is summed,
    // the next trip's indices before this trip's sums (padding -1 follows a
    // cell's last neighbour)
    int e = e0 + lane;
    int jn[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) jn[u] = e + 64 * u < e1 ? nbr(e + 64 * u) : -1;
    while (jn[0] >= 0) {
        int j[UNR];
        double2 r0[UNR], r1[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            j[u] = jn[u];
            if (j[u] >= 0) {
                const double2* o = reinterpret_cast<const double2*>(xph + 4 * (size_t)j[u]);
                r0[u] = o[0]; r1[u] = o[1];
            }
        }
        e += 64 * UNR;
#pragma unroll
        for (int u = 0; u < UNR; ++u) jn[u] = (j[UNR - 1] >= 0 && e + 64 * u < e1) ? nbr(e + 64 * u) : -1;
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            if (j[u] < 0) break;
            double dx = r0[u].x - xi, dy = r0[u].y - yi, dz = r1[u].x - zi;
            double wk = 1.0 / (dx * dx + dy * dy + dz * dz);
            double d = r1[u].y - p;
            t0 += wk * dx * d; t1 += wk * dy * d; t2 += wk * dz * d;
        }
    }
    for (int q = bf_ptr[i]; q < bf_ptr[i + 1]; ++q) {
        int f = bf_idx[q];
        double dx = fxyz[3 * (size_t)f] - xi, dy = fxyz[3 * (size_t)f + 1] - yi, dz = fxyz[3 * (size_t)f + 2] - zi;
        double wk = 1.0 / (dx * dx + dy * dy + dz * dz);
        double d = faceR ? faceR[f - foff] - p : 0.0;
        t0 += wk * dx * d; t1 += wk * dy * d; t2 += wk * dz * d;
    }
    const double* C = lsq + 6 * (size_t)i;
    double r0 = 0.0; r0 += C[0] * t0; r0 += C[1] * t1; r0 += C[2] * t2;
    double r1 = 0.0; r1 += C[1] * t0; r1 += C[3] * t1; r1 += C[4] * t2;
    double r2 = 0.0; r2 += C[2] * t0; r2 += C[4] * t1; r2 += C[5] * t2;
    g[3 * (size_t)i] = r0; g[3 * (size_t)i + 1] = r1; g[3 * (size_t)i + 2] = r2;
}

// The same gradient with G lanes per cell (G = 2 or 4): lane q of a cell's
// group gathers and weighs its stencil entries q, q + G, ..., so one load
// instruction brings G neighbours of each of 64 / G cells (neighbouring
// entries of a sorted stencil often share a cache line); the G terms of a
// round are then broadcast inside the group (DPP quad permutes, no LDS) and
// added on every lane in stencil order -- the sums, and so the gradient, are
// bit for bit those of k_lsq1.  cnt: the stencil's CSR row pointers (entry
// counts).
template <int CTRL>
__device__ inline double dpp_bcast(double v) {
    int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
    int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
// quad_perm control that gives every lane of a G-lane group the value of the
// group's lane SRC (G = 4: the quad; G = 2: each half of the quad)
template <int G, int SRC>
constexpr int group_bcast_ctrl() {
    return G == 4 ? SRC * 0x55 : (SRC | SRC << 2 | (2 + SRC) << 4 | (2 + SRC) << 6);
}
template <class IDX, int G, int UNR>
__global__ void __launch_bounds__(TPB) k_lsq1g(
    int nc, const int* __restrict__ st_sl, const IDX* __restrict__ st_ell, const int* __restrict__ st_ptr,
    const int* __restrict__ bf_ptr, const int* __restrict__ bf_idx, const double* __restrict__ xph,
    const double* __restrict__ fxyz, const double* __restrict__ lsq, const double* __restrict__ faceR, int foff,
    double* __restrict__ g, const int* __restrict__ gate) {
    static_assert(G == 2 || G == 4, "lanes per cell");
    if (gate && *gate == 0) return;
    int gid = xcd_block() * blockDim.x + threadIdx.x;
    int i = gid / G, q = gid & (G - 1);
    if (i >= nc) return;      // a group's lanes leave together
    const double2* me = reinterpret_cast<const double2*>(xph + 4 * (size_t)i);
    double2 m0 = me[0], m1 = me[1];
    double xi = m0.x, yi = m0.y, zi = m1.x, p = m1.y;
    double t0 = 0, t1 = 0, t2 = 0;
    int s = i >> 6, lane = i & 63;
    int e0 = st_sl[s];
    int cnt = st_ptr[i + 1] - st_ptr[i];
    for (int kb = 0; kb < cnt; kb += G * UNR) {
        int j[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            int k = kb + u * G + q;
            if (k < cnt) {
                int v = __builtin_nontemporal_load(st_ell + e0 + 64 * k + lane);
                j[u] = sizeof(IDX) == 2 ? i + v : v;
            } else {
                j[u] = -1;
            }
        }
        double2 r0[UNR], r1[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u)
            if (j[u] >= 0) {
                const double2* o = reinterpret_cast<const double2*>(xph + 4 * (size_t)j[u]);
                r0[u] = o[0]; r1[u] = o[1];
            }
        double a0[UNR], a1[UNR], a2[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            a0[u] = a1[u] = a2[u] = 0.0;
            if (j[u] >= 0) {
                double dx = r0[u].x - xi, dy = r0[u].y - yi, dz = r1[u].x - zi;
                double wk = 1.0 / (dx * dx + dy * dy + dz * dz);
                double d = r1[u].y - p;
                a0[u] = wk * dx * d; a1[u] = wk * dy * d; a2[u] = wk * dz * d;
            }
        }
        // stencil order: entry kb + u G + src is lane src's u-th term
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            int k = kb + u * G;
            double b0 = dpp_bcast<group_bcast_ctrl<G, 0>()>(a0[u]);
            double b1 = dpp_bcast<group_bcast_ctrl<G, 0>()>(a1[u]);
            double b2 = dpp_bcast<group_bcast_ctrl<G, 0>()>(a2[u]);
            if (k < cnt) { t0 += b0; t1 += b1; t2 += b2; }
            b0 = dpp_bcast<group_bcast_ctrl<G, 1>()>(a0[u]);
            b1 = dpp_bcast<group_bcast_ctrl<G, 1>()>(a1[u]);
            b2 = dpp_bcast<group_bcast_ctrl<G, 1>()>(a2[u]);
            if (k + 1 < cnt) { t0 += b0; t1 += b1; t2 += b2; }
            if constexpr (G == 4) {
                b0 = dpp_bcast<group_bcast_ctrl<G, 2>()>(a0[u]);
                b1 = dpp_bcast<group_bcast_ctrl<G, 2>()>(a1[u]);
                b2 = dpp_bcast<group_bcast_ctrl<G, 2>()>(a2[u]);
                if (k + 2 < cnt) { t0 += b0; t1 += b1; t2 += b2; }
                b0 = dpp_bcast<group_bcast_ctrl<G, 3>()>(a0[u]);
                b1 = dpp_bcast<group_bcast_ctrl<G, 3>()>(a1[u]);
                b2 = dpp_bcast<group_bcast_ctrl<G, 3>()>(a2[u]);
                if (k + 3 < cnt) { t0 += b0; t1 += b1; t2 += b2; }
            }
        }
    }
    if (q) return;
    for (int qq = bf_ptr[i]; qq < bf_ptr[i + 1]; ++qq) {
        int f = bf_idx[qq];
        double dx = fxyz[3 * (size_t)f] - xi, dy = fxyz[3 * (size_t)f + 1] - yi, dz = fxyz[3 * (size_t)f + 2] - zi;
        double wk = 1.0 / (dx * dx + dy * dy + dz * dz);
        double d = faceR ? faceR[f - foff] - p : 0.0;
        t0 += wk * dx * d; t1 += wk * dy * d; t2 += wk * dz * d;
    }
    const double* C = lsq + 6 * (size_t)i;
    double r0 = 0.0; r0 += C[0] * t0; r0 += C[1] * t1; r0 += C[2] * t2;
    double r1 = 0.0; r1 += C[1] * t0; r1 += C[3] * t1; r1 += C[4] * t2;
    double r2 = 0.0; r2 += C[2] * t0; r2 += C[4] * t1; r2 += C[5] * t2;
    g[3 * (size_t)i] = r0; g[3 * (size_t)i + 1] = r1; g[3 * (size_t)i + 2] = r2;
}

__global__ void __launch_bounds__(TPB) k_minmax1(int nc, const int* __restrict__ st_ptr,
                                                 const int* __restrict__ st_idx, const double* __restrict__ phi,
                                                 double* __restrict__ mx, double* __restrict__ mn) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double a = phi[i], b = phi[i];
    for (int q = st_ptr[i]; q < st_ptr[i + 1]; ++q) { double v = phi[st_idx[q]]; a = smax(a, v); b = smin(b, v); }
    mx[i] = a; mn[i] = b;
}

// --------------------------------------------------------------- momentum --
// internal face of calcMomentumEqs (eqMomentum.cpp:331-787): off-diagonals in
// place; diagonal and right-hand-side contributions to the per-face scratch
//   s[0] = X  (owner diag  +=), s[1] = Y (neighbour diag -=),
//   s[2..4] = owner B_k -=,   s[5..7] = neighbour B_k +=
__global__ void __launch_bounds__(TPB) k_mom_face(
    int nif, CaseK k, FaceIn in, const double* __restrict__ Mu, const int* __restrict__ slot_own,
    const int* __restrict__ slot_nbr, double* __restrict__ aval, double* __restrict__ s) {
    int f = xcd_block() * blockDim.x + threadIdx.x;
    if (f >= nif) return;
    int l = ldnt(in.owner + f), r = ldnt(in.nbr + f);
    double area = ldnt(in.area + f);
    double n[3], e[3], sk[3];
    ldnt3(in.nrm, f, n); ldnt3(in.ePN, f, e); ldnt3(in.skew, f, sk);
    double wCL = ldnt(in.wC + f), wCR = 1.0 - wCL;
    double muF = 1.0 / (wCL / Mu[l] + wCR / Mu[r]);
    double RL = in.Rho[l], RR = in.Rho[r];
    double dPN = ldnt(in.magPN + f), alpha = ldnt(in.alpha + f);
    double UnF = face_unf<true>(in, k, f, l, r);
    double wL = (UnF > 0.0) ? 1.0 : 0.0, wR = 1.0 - wL;
    double PL = in.P[l], PR = in.P[r];
    double uL[3] = {in.U[l], in.V[l], in.W[l]}, uR[3] = {in.U[r], in.V[r], in.W[r]};
    double Fv[3] = {wL * uL[0] + wR * uR[0], wL * uL[1] + wR * uR[1], wL * uL[2] + wR * uR[2]};
    const double* gPl = in.gP + 3 * (size_t)l;
    const double* gPr = in.gP + 3 * (size_t)r;
    double PF = wCL * PL + wCR * PR;
    for (int a = 0; a < 3; ++a) { PF += wCL * gPl[a] * sk[a]; PF += wCR * gPr[a] * sk[a]; }
    PF = smax(in.mnP[l], smin(in.mxP[l], PF));
    PF = smax(in.mnP[r], smin(in.mxP[r], PF));
    double X = (wL * RL * UnF * area + alpha * muF / dPN * area);
    double ao = (wR * RL * UnF * area - alpha * muF / dPN * area);
    double Y = (wR * RR * UnF * area - alpha * muF / dPN * area);
    double an = -(wL * RR * UnF * area + alpha * muF / dPN * area);
    // entries with |v| < DBL_MIN are dropped (eqMomentum.cpp:699-711)
    aval[ldnt(slot_own + f)] = fabs(ao) < 2.2250738585072014e-308 ? 0.0 : ao;
    aval[ldnt(slot_nbr + f)] = fabs(an) < 2.2250738585072014e-308 ? 0.0 : an;
    stnt(s + f, X);
    stnt(s + (size_t)nif + f, Y);
    const double* gl[3] = {in.gU + 3 * (size_t)l, in.gV + 3 * (size_t)l, in.gW + 3 * (size_t)l};
    const double* gr[3] = {in.gU + 3 * (size_t)r, in.gV + 3 * (size_t)r, in.gW + 3 * (size_t)r};
    for (int v = 0; v < 3; ++v) {
        double cL = (RL * Fv[v] * UnF + PF * n[v]) * area;
        double cR = (RR * Fv[v] * UnF + PF * n[v]) * area;
        double df = alpha * muF * (uR[v] - uL[v]) / dPN * area;
        for (int a = 0; a < 3; ++a) df += (muF * (wCL * gl[v][a] + wCR * gr[v][a]) * (n[a] - alpha * e[a]) * area);
        double dL = ((gl[v][0] * n[0] + gl[v][1] * n[1] + gl[v][2] * n[2]) * muF * area);
        double dR = ((gr[v][0] * n[0] + gr[v][1] * n[1] + gr[v][2] * n[2]) * muF * area);
        stnt(s + (size_t)(2 + v) * nif + f, (cL - df - dL));
        stnt(s + (size_t)(5 + v) * nif + f, (cR - df - dR));
    }
}

// cell part of calcMomentumEqs: time term and sources (eqMomentum.cpp:223-326),
// the cell's internal-face contributions in ascending face order, then its
// boundary faces (eqMomentum.cpp:795-1008); writes diagonal and 3 RHS.
__global__ void __launch_bounds__(TPB) k_mom_cell(
    int nc, int nif, int nbf, CaseK k, const int* __restrict__ cf_ptr, const int* __restrict__ cf_idx,
    const int* __restrict__ owner, const int* __restrict__ bpatch, const semo_patch_bc* __restrict__ bcs,
    const double* __restrict__ vol, const double* __restrict__ nrm, const double* __restrict__ ePN,
    const double* __restrict__ area, const double* __restrict__ magPN, const double* __restrict__ alphaF,
    const double* __restrict__ Rho, const double* __restrict__ P, const double* __restrict__ U,
    const double* __restrict__ V, const double* __restrict__ W, const double* __restrict__ oU,
    const double* __restrict__ oV, const double* __restrict__ oW, const double* __restrict__ q0,
    const double* __restrict__ q1, const double* __restrict__ q2, const double* __restrict__ srcG,
    const double* __restrict__ srcS, const double* __restrict__ alphaP, const double* __restrict__ gP,
    const double* __restrict__ gU, const double* __restrict__ gV, const double* __restrict__ gW,
    const double* __restrict__ bP, const double* __restrict__ bUL, const double* __restrict__ bUR,
    const double* __restrict__ bMu, const double* __restrict__ bRho, const double* __restrict__ s,
    double* __restrict__ diag, double* __restrict__ B) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double R = Rho[i], vl = vol[i], dt = k.dt;
    double d = R * 1.5 * vl / dt;
    double b[3];
    b[0] = -R * (1.5 * U[i] - 2.0 * oU[i] + 0.5 * q0[i]) * vl / dt;
    b[1] = -R * (1.5 * V[i] - 2.0 * oV[i] + 0.5 * q1[i]) * vl / dt;
    b[2] = -R * (1.5 * W[i] - 2.0 * oW[i] + 0.5 * q2[i]) * vl / dt;
    if (k.body) {
        for (int v = 0; v < 3; ++v) b[v] += vl * srcG[v * (size_t)nc + i];
        for (int v = 0; v < 3; ++v) b[v] += vl * srcS[v * (size_t)nc + i];
    }
    double uc[3] = {U[i], V[i], W[i]};
    if (alphaP) {   // Darcy porosity source -alpha_p*U*V, implicit (adjoint harness)
        double ap = alphaP[i] * vl;
        d += ap;
        for (int v = 0; v < 3; ++v) b[v] -= ap * uc[v];
    }
    const double* gc[3] = {gU + 3 * (size_t)i, gV + 3 * (size_t)i, gW + 3 * (size_t)i};
    CELL_FACES_BEGIN(i, cf_ptr, cf_idx, f, own)
        if (f < nif) {
            if (own) {
                d += s[f];
                for (int v = 0; v < 3; ++v) b[v] -= s[(size_t)(2 + v) * nif + f];
            } else {
                d -= s[(size_t)nif + f];
                for (int v = 0; v < 3; ++v) b[v] += s[(size_t)(5 + v) * nif + f];
            }
            continue;
        }
        int j = f - nif;
        semo_patch_bc bc = bcs[bpatch[j]];
        const double* n = nrm + 3 * (size_t)f;
        const double* e = ePN + 3 * (size_t)f;
        double A = area[f];
        double ul[3], ur[3];
        for (int a = 0; a < 3; ++a) { ul[a] = bUL[a * (size_t)nbf + j]; ur[a] = bUR[a * (size_t)nbf + j]; }
        double UnL = ul[0] * n[0] + ul[1] * n[1] + ul[2] * n[2];
        double UnR = ur[0] * n[0] + ur[1] * n[1] + ur[2] * n[2];
        double UnF = 0.5 * UnL + 0.5 * UnR;
        double RL = bRho[j];
        double t1 = 1.0 / RL * dt;
        double Fv[3] = {0.5 * ul[0] + 0.5 * ur[0], 0.5 * ul[1] + 0.5 * ur[1], 0.5 * ul[2] + 0.5 * ur[2]};
        double PF = 0.5 * bP[j] + 0.5 * bP[j];
        double cU = 0.0, cUd = 1.0;
        if (bc.u_type == SEMO_BC_ZERO_GRADIENT) { cU = 1.0; cUd = 0.0; }
        else if (bc.u_type == SEMO_BC_SLIP) { cU = 0.0; cUd = 0.0; }
        else if (bc.u_type == SEMO_BC_INLET_OUTLET) {
            double ownU = uc[0] * n[0] + uc[1] * n[1] + uc[2] * n[2];
            cU = ownU > 0.0 ? 1.0 : 0.0; cUd = ownU > 0.0 ? 0.0 : 1.0;
        }
        double muF = 0.5 * bMu[j] + 0.5 * bMu[j];
        double dPN = 0.5 * magPN[f], alpha = alphaF[f];
        if (bc.p_type == SEMO_BC_FIXED_VALUE) {
            UnF -= alpha * t1 * (bc.p_value - P[i]) / dPN;
            for (int a = 0; a < 3; ++a) UnF += alpha * t1 * gP[3 * (size_t)i + a] * e[a];
        }
        double dn[3];
        for (int v = 0; v < 3; ++v) { dn[v] = 0.0; dn[v] += alpha * (ur[v] - uc[v]) / dPN; }
        for (int a = 0; a < 3; ++a) {
            double tv = n[a] - alpha * e[a];
            for (int v = 0; v < 3; ++v) dn[v] += gc[v][a] * tv;
        }
        for (int a = 0; a < 3; ++a)
            for (int v = 0; v < 3; ++v) dn[v] += (gc[v][a] * n[a]);
        d += cU * (RL * UnF * A);
        d += cUd * (muF * alpha / dPN * A);
        for (int v = 0; v < 3; ++v) b[v] -= (RL * Fv[v] * UnF * A + PF * n[v] * A - muF * dn[v] * A);
    }
    diag[i] = d;
    for (int v = 0; v < 3; ++v) B[v * (size_t)nc + i] = b[v];
}

// --------------------------------------------------------------- pressure --
// internal face of calcPressureEq (eqPressure.cpp:246-564):
//   s[0] = alpha*dt/rho* /dPN*A (rho* with wCR twice, eqPressure.cpp:396)
//   s[1] = UnF*A
__global__ void __launch_bounds__(TPB) k_p_face(
    int nif, CaseK k, FaceIn in, const int* __restrict__ slot_own, const int* __restrict__ slot_nbr,
    double* __restrict__ aval, double* __restrict__ s) {
    int f = xcd_block() * blockDim.x + threadIdx.x;
    if (f >= nif) return;
    int l = in.owner[f], r = in.nbr[f];
    double area = in.area[f];
    // rho* with the neighbour weight twice (eqPressure.cpp:396-397): the
    // weight of the face's neighbour in the mesh's own orientation, i.e. this
    // side's owner weight when the renumbering flipped the face
    double wCL = in.wC[f], wCR = in.flip[f] ? wCL : 1.0 - wCL;
    double Rs = 1.0 / (wCR / in.Rho[l] + wCR / in.Rho[r]);
    double t1 = k.dt / Rs;
    double dPN = in.magPN[f], alpha = in.alpha[f];
    double UnF = face_unf(in, k, f, l, r);
    double cf = (+alpha * t1 / dPN * area);
    aval[slot_own[f]] = (-alpha * t1 / dPN * area);
    aval[slot_nbr[f]] = -(+alpha * t1 / dPN * area);
    s[f] = cf;
    s[(size_t)nif + f] = (UnF * area);
}

__global__ void __launch_bounds__(TPB) k_p_cell(
    int nc, int nif, int nbf, CaseK k, const int* __restrict__ cf_ptr, const int* __restrict__ cf_idx,
    const int* __restrict__ owner, const int* __restrict__ bpatch, const semo_patch_bc* __restrict__ bcs,
    const double* __restrict__ nrm, const double* __restrict__ ePN, const double* __restrict__ area,
    const double* __restrict__ magPN, const double* __restrict__ alphaF, const double* __restrict__ P,
    const double* __restrict__ gP, const double* __restrict__ bP, const double* __restrict__ bUL,
    const double* __restrict__ bUR, const double* __restrict__ bRho, const double* __restrict__ s,
    const double* __restrict__ U, const double* __restrict__ V, const double* __restrict__ W,
    const double* __restrict__ Cs, double* __restrict__ diag, double* __restrict__ B) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double d = 0.0, b = 0.0;
    CELL_FACES_BEGIN(i, cf_ptr, cf_idx, f, own)
        if (f < nif) {
            if (own) { d += s[f]; b -= s[(size_t)nif + f]; }
            else { d -= -s[f]; b += s[(size_t)nif + f]; }
            continue;
        }
        int j = f - nif;
        semo_patch_bc bc = bcs[bpatch[j]];
        const double* n = nrm + 3 * (size_t)f;
        const double* e = ePN + 3 * (size_t)f;
        double UnL = bUL[j] * n[0] + bUL[(size_t)nbf + j] * n[1] + bUL[2 * (size_t)nbf + j] * n[2];
        double UnR = bUR[j] * n[0] + bUR[(size_t)nbf + j] * n[1] + bUR[2 * (size_t)nbf + j] * n[2];
        double UnF = 0.5 * UnL + 0.5 * UnR;
        double t1 = 1.0 / bRho[j] * k.dt;
        double dPN = 0.5 * magPN[f], alpha = alphaF[f], A = area[f];
        double cPd = 0.0;
        if (bc.p_type == SEMO_BC_FIXED_VALUE) {
            cPd = 1.0;
            UnF -= alpha * t1 * (bP[j] - P[i]) / dPN;
            for (int a = 0; a < 3; ++a) UnF += alpha * t1 * gP[3 * (size_t)i + a] * e[a];
        } else if (bc.p_type == SEMO_BC_SWITCH) {
            cPd = cell_mach(i, U, V, W, Cs) > 1.0 ? 1.0 : 0.0;
        }
        d += cPd * (alpha * t1 / dPN * A);
        b -= (UnF * A);
    }
    diag[i] = d;
    B[i] = b;
}

// ---------------------------------------------------------- volume fraction --
// largest face area of each cell (the Courant number of calcMSTACS,
// NVD.cpp:1119-1127); static per mesh
__global__ void __launch_bounds__(TPB) k_cell_max_area(int nc, const int* __restrict__ cf_ptr,
                                                       const int* __restrict__ cf_idx,
                                                       const double* __restrict__ area, double* __restrict__ maxA) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double m = 0.0;
    CELL_FACES_BEGIN(i, cf_ptr, cf_idx, f, own)
        (void)own;
        m = smax(m, area[f]);
    }
    maxA[i] = m;
}

// calcMSTACS (NVD.cpp:1098-1862) for one internal face: the high-order
// volume fraction on the owner (L) and neighbour (R) side.  Courant numbers
// |U| dt maxA / V (NVD.cpp:1114-1135) are formed inline for the two cells.
__global__ void __launch_bounds__(TPB) k_mstacs_face(
    int nif, CaseK k, const int* __restrict__ owner, const int* __restrict__ nbr,
    const double* __restrict__ wCv, const double* __restrict__ nrm, const double* __restrict__ skew,
    const double* __restrict__ cxyz, const double* __restrict__ vol, const double* __restrict__ maxA,
    const double* __restrict__ U, const double* __restrict__ V, const double* __restrict__ W,
    const double* __restrict__ VF, const double* __restrict__ gVF, double* __restrict__ vfL_out,
    double* __restrict__ vfR_out) {
    int f = xcd_block() * blockDim.x + threadIdx.x;
    if (f >= nif) return;
    int l = owner[f], r = nbr[f];
    double wCL = wCv[f], wCR = 1.0 - wCL;
    double L1 = VF[l], R1 = VF[r];
    const double eps = 1.e-12;
    if ((L1 > 1.0 - eps || L1 < eps) && (R1 > 1.0 - eps || R1 < eps)) {   // both sides skipped
        vfL_out[f] = L1; vfR_out[f] = R1;
        return;
    }
    double gl[3], gr[3], pn[3];
    for (int a = 0; a < 3; ++a) {
        gl[a] = gVF[3 * (size_t)l + a]; gr[a] = gVF[3 * (size_t)r + a];
        pn[a] = cxyz[3 * (size_t)r + a] - cxyz[3 * (size_t)l + a];      // vecPN (geometric.cpp:1425-1427)
    }
    const double* n = nrm + 3 * (size_t)f;
    const double* sk = skew + 3 * (size_t)f;
    double ul = sqrt(U[l] * U[l] + V[l] * V[l] + W[l] * W[l]);
    double ur = sqrt(U[r] * U[r] + V[r] * V[r] + W[r] * W[r]);
    double coL = ul * k.dt * maxA[l] / vol[l];
    double coR = ur * k.dt * maxA[r] / vol[r];
    double coDD = smax(coL, coR);
    double L2 = 0.0;
    for (int a = 0; a < 3; ++a) L2 += gl[a] * pn[a];
    L2 = R1 - 2.0 * L2;
    double R2 = 0.0;
    for (int a = 0; a < 3; ++a) R2 += gr[a] * pn[a];
    R2 = L1 + 2.0 * R2;
    L2 = clip01(L2);
    R2 = clip01(R2);
    // blending weight from the interpolated interface normal (NVD.cpp:1437-1477)
    double mf[3];
    for (int a = 0; a < 3; ++a) mf[a] = wCL * gl[a] + wCR * gr[a];
    double mag = mf[0] * mf[0];
    mag += mf[1] * mf[1];
    mag += mf[2] * mf[2];
    mag = sqrt(mag);
    for (int a = 0; a < 3; ++a) mf[a] = mf[a] / (mag + 1.e-200);
    double cosT = mf[0] * n[0];
    cosT += mf[1] * n[1];
    cosT += mf[2] * n[2];
    cosT = fabs(cosT);
    double gamF = smin(pow(cosT, 4.0), 1.0);
    double out[2] = {L1, R1};
    for (int LR = 0; LR < 2; ++LR) {
        double self = LR == 0 ? L1 : R1;
        if (self > 1.0 - eps || self < eps) continue;
        double alpU = LR == 0 ? L2 : R2, alpD = self, alpA = LR == 0 ? R1 : L1;
        double tCd = (alpD - alpU) / (fabs(alpA - alpU) + 1.e-200) * (alpA > alpU ? 1.0 : -1.0);
        double side = LR == 0 ? 2.0 * wCR : 2.0 * wCL;
        // compressive branch (NVD.cpp:1602-1626)
        double gam = 0.0;
        if (tCd >= 0.0 && tCd < 1.0) {
            if (coDD > 0.0 && coDD <= 0.33) {
                double tCf = smin(1.0, tCd / coDD);
                gam = (tCf - tCd) / (1.0 - tCd);
            } else if (coDD > 0.33 && coDD <= 1.0) {
                double tCf = smin(1.0, 3.0 * tCd);
                gam = (tCf - tCd) / (1.0 - tCd);
            }
        }
        gam *= side;
        gam = clip01(gam);
        double vc = alpD + gam * (alpA - alpD);
        vc += wCL * gl[0] * sk[0]; vc += wCL * gl[1] * sk[1]; vc += wCL * gl[2] * sk[2];
        vc += wCR * gr[0] * sk[0]; vc += wCR * gr[1] * sk[1]; vc += wCR * gr[2] * sk[2];
        vc = clip01(vc);
        // diffusive branch (NVD.cpp:1741-1767)
        gam = 0.0;
        if (tCd >= 0.0 && tCd < 0.2) { double tCf = smin(1.0, 3.0 * tCd); gam = (tCf - tCd) / (1.0 - tCd); }
        else if (tCd >= 0.2 && tCd < 0.5) { double tCf = 0.5 + 0.5 * tCd; gam = (tCf - tCd) / (1.0 - tCd); }
        else if (tCd >= 0.5 && tCd < 0.8333) { double tCf = 0.375 + 0.75 * tCd; gam = (tCf - tCd) / (1.0 - tCd); }
        else if (tCd >= 0.8333 && tCd < 1.0) { double tCf = 1.0; gam = (tCf - tCd) / (1.0 - tCd); }
        gam *= side;
        gam = clip01(gam);
        double vd = alpD + gam * (alpA - alpD);
        vd += wCL * gl[0] * sk[0]; vd += wCL * gl[1] * sk[1]; vd += wCL * gl[2] * sk[2];
        vd += wCR * gr[0] * sk[0]; vd += wCR * gr[1] * sk[1]; vd += wCR * gr[2] * sk[2];
        vd = clip01(vd);
        out[LR] = gamF * vc + (1.0 - gamF) * vd;
    }
    vfL_out[f] = out[0];
    vfR_out[f] = out[1];
}

// internal face of calcVolfracEq (eqVolfrac.cpp:266-607): off-diagonals in
// place; s[0] = owner diag +=, s[1] = neighbour diag -=, s[2] = VF_f UnF A
// (owner B -=, neighbour B +=)
__global__ void __launch_bounds__(TPB) k_vof_face(
    int nif, CaseK k, FaceIn in, const double* __restrict__ vfL, const double* __restrict__ vfR,
    const int* __restrict__ slot_own, const int* __restrict__ slot_nbr, double* __restrict__ aval,
    double* __restrict__ s) {
    int f = xcd_block() * blockDim.x + threadIdx.x;
    if (f >= nif) return;
    int l = in.owner[f], r = in.nbr[f];
    double A = in.area[f];
    double UnF = face_unf(in, k, f, l, r);
    double wL = (UnF > 0.0) ? 1.0 : 0.0, wR = 1.0 - wL;
    double VFF = wL * vfL[f] + wR * vfR[f];
    double ao = (wR * UnF * A);
    double an = -(wL * UnF * A);
    // entries with |v| < DBL_MIN are dropped (eqVolfrac.cpp:579-591)
    aval[slot_own[f]] = fabs(ao) < 2.2250738585072014e-308 ? 0.0 : ao;
    aval[slot_nbr[f]] = fabs(an) < 2.2250738585072014e-308 ? 0.0 : an;
    s[f] = (wL * UnF * A);
    s[(size_t)nif + f] = (wR * UnF * A);
    s[2 * (size_t)nif + f] = (VFF * UnF * A);
}

// cell part of calcVolfracEq: time term (eqVolfrac.cpp:221-262), internal
// faces in ascending order, then boundary faces (eqVolfrac.cpp:611-703)
__global__ void __launch_bounds__(TPB) k_vof_cell(
    int nc, int nif, int nbf, CaseK k, const int* __restrict__ cf_ptr, const int* __restrict__ cf_idx,
    const int* __restrict__ bpatch, const semo_patch_bc* __restrict__ bcs, const double* __restrict__ vol,
    const double* __restrict__ nrm, const double* __restrict__ ePN, const double* __restrict__ area,
    const double* __restrict__ magPN, const double* __restrict__ alphaF, const double* __restrict__ P,
    const double* __restrict__ U, const double* __restrict__ V, const double* __restrict__ W,
    const double* __restrict__ VF, const double* __restrict__ oVF, const double* __restrict__ q3,
    const double* __restrict__ gP, const double* __restrict__ bVF, const double* __restrict__ bUL,
    const double* __restrict__ bUR, const double* __restrict__ bRho, const double* __restrict__ s,
    double* __restrict__ diag, double* __restrict__ B) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double vl = vol[i], dt = k.dt;
    double d = 1.5 * vl / dt;
    double b = -(1.5 * VF[i] - 2.0 * oVF[i] + 0.5 * q3[i]) * vl / dt;
    CELL_FACES_BEGIN(i, cf_ptr, cf_idx, f, own)
        if (f < nif) {
            if (own) { d += s[f]; b -= s[2 * (size_t)nif + f]; }
            else { d -= s[(size_t)nif + f]; b += s[2 * (size_t)nif + f]; }
            continue;
        }
        int j = f - nif;
        semo_patch_bc bc = bcs[bpatch[j]];
        const double* n = nrm + 3 * (size_t)f;
        const double* e = ePN + 3 * (size_t)f;
        double UnL = bUL[j] * n[0] + bUL[(size_t)nbf + j] * n[1] + bUL[2 * (size_t)nbf + j] * n[2];
        double UnR = bUR[j] * n[0] + bUR[(size_t)nbf + j] * n[1] + bUR[2 * (size_t)nbf + j] * n[2];
        double UnF = 0.5 * UnL + 0.5 * UnR;
        double t1 = 1.0 / bRho[j] * dt;
        double VFF = 0.5 * bVF[j] + 0.5 * bVF[j];
        double cVF = 0.0;
        if (bc.vf_type == SEMO_BC_ZERO_GRADIENT) cVF = 1.0;
        else if (bc.vf_type == SEMO_BC_INLET_OUTLET) {
            double ownU = U[i] * n[0] + V[i] * n[1] + W[i] * n[2];
            cVF = (ownU > 0.0) ? 1.0 : 0.0;
        }
        double dPN = 0.5 * magPN[f], alpha = alphaF[f], A = area[f];
        if (bc.p_type == SEMO_BC_FIXED_VALUE) {
            UnF -= alpha * t1 * (bc.p_value - P[i]) / dPN;
            for (int a = 0; a < 3; ++a) UnF += alpha * t1 * gP[3 * (size_t)i + a] * e[a];
        }
        d += cVF * (UnF * A);
        b -= (VFF * UnF * A);
    }
    diag[i] = d;
    B[i] = b;
}

// smallest of the convective, viscous, gravity and capillary time scales of
// each cell, block minima into part[2b] (timestep.cpp:202-253 with adjust = 0,
// timestep.cpp:260-330 with adjust = 1)
__global__ void __launch_bounds__(TPB) k_time_scale(
    int nc, CaseK k, int adjust, const int* __restrict__ cf_ptr, const int* __restrict__ cf_idx,
    const double* __restrict__ area, const double* __restrict__ vol, const double* __restrict__ U,
    const double* __restrict__ V, const double* __restrict__ W, const double* __restrict__ Mu,
    const double* __restrict__ Rho, double* __restrict__ part) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    double minDt = 1000000.0;
    if (i < nc) {
        double maxA = 0.0, minA = 1000000.0;
        CELL_FACES_BEGIN(i, cf_ptr, cf_idx, f, own)
            (void)own;
            maxA = smax(maxA, area[f]);
            minA = smin(minA, area[f]);
        }
        double magVel = sqrt(U[i] * U[i] + V[i] * V[i] + W[i] * W[i]);
        double convTime = vol[i] / maxA / (magVel + 1.e-200);
        minDt = smin(minDt, convTime);
        double viscTime = 0.5 * minA / (Mu[i] + 1.e-200) * Rho[i];
        minDt = smin(minDt, viscTime);
        double gravMag = sqrt(k.g0 * k.g0 + k.g1 * k.g1 + k.g2 * k.g2);
        double gravTime = sqrt(sqrt(minA) / (gravMag + 1.e-200));
        minDt = smin(minDt, gravTime);
        double stTime;
        if (adjust) {
            double minDX = vol[i] / maxA;
            double C_sigma = sqrt(6.28 * k.sigma / (k.rho0 + k.rho1) / minDX);
            stTime = minDX / (C_sigma + magVel);
        } else {
            stTime = sqrt(pow(minA, 1.5) * (1000.0 + 1.2) / (4.0 * 3.141592 * 0.0728));
        }
        minDt = smin(minDt, stTime);
    }
    __shared__ double sm[TPB / 64];
    for (int o = 32; o > 0; o >>= 1) minDt = smin(minDt, __shfl_down(minDt, o, 64));
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = minDt;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = sm[0];
        for (int w = 1; w < TPB / 64; ++w) m = smin(m, sm[w]);
        part[2 * blockIdx.x] = m;
    }
}

// VF += urf dVF, clipped to [0,1]; block partials of dVF^2 (eqVolfrac.cpp:770-783)
__global__ void __launch_bounds__(TPB) k_vof_update(int nc, double urf, const double* __restrict__ x,
                                                    double* __restrict__ VF, double* __restrict__ part) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    double s2 = 0.0;
    if (i < nc) {
        double xi = x[i];
        VF[i] = clip01(VF[i] + urf * xi);
        s2 = xi * xi;
    }
    block_sum2(s2, 0.0, part);
}

// ------------------------------------------------------------- curvature --
// calcCurvature (source/curvature.cpp:7-3537), live part: level-set
// reinitialisation of alpha - 1/2, then kappa = -div(n).

// inverse of a packed symmetric 3x3 matrix by cofactors, with the reference's
// determinant expression (curvature.cpp:2964-2979)
__device__ inline void inv_sym3(const double* v, double* o) {
    double detA = v[0] * v[3] * v[5] + v[1] * v[4] * v[2] + v[2] * v[1] * v[4] - v[0] * v[4] * v[4] -
                  v[2] * v[3] * v[2] - v[1] * v[1] * v[5];
    o[0] = (v[3] * v[5] - v[4] * v[4]) / detA;
    o[1] = (v[2] * v[4] - v[1] * v[5]) / detA;
    o[2] = (v[1] * v[4] - v[2] * v[3]) / detA;
    o[3] = (v[0] * v[5] - v[2] * v[2]) / detA;
    o[4] = (v[2] * v[1] - v[0] * v[4]) / detA;
    o[5] = (v[0] * v[3] - v[1] * v[1]) / detA;
}

// static: inverse of the unweighted least-squares matrix over the vertex
// stencil and the cell's boundary faces (curvature.cpp:2742-2981)
__global__ void __launch_bounds__(TPB) k_kappa_lsq_init(
    int nc, const int* __restrict__ st_ptr, const int* __restrict__ st_idx, const int* __restrict__ bf_ptr,
    const int* __restrict__ bf_idx, const double* __restrict__ cxyz, const double* __restrict__ fxyz,
    double* __restrict__ out) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double m[6] = {0, 0, 0, 0, 0, 0};
    const double* ci = cxyz + 3 * (size_t)i;
    for (int q = st_ptr[i]; q < st_ptr[i + 1]; ++q) {
        const double* cj = cxyz + 3 * (size_t)st_idx[q];
        double d[3] = {cj[0] - ci[0], cj[1] - ci[1], cj[2] - ci[2]};
        for (int a = 0, num = 0; a < 3; ++a)
            for (int b = a; b < 3; ++b) m[num++] += 1.0 * d[a] * d[b];
    }
    for (int q = bf_ptr[i]; q < bf_ptr[i + 1]; ++q) {
        const double* fc = fxyz + 3 * (size_t)bf_idx[q];
        double d[3] = {fc[0] - ci[0], fc[1] - ci[1], fc[2] - ci[2]};
        for (int a = 0, num = 0; a < 3; ++a)
            for (int b = a; b < 3; ++b) m[num++] += 1.0 * d[a] * d[b];
    }
    double inv[6];
    inv_sym3(m, inv);
    for (int k = 0; k < 6; ++k) out[6 * (size_t)i + k] = inv[k];
}

// interface cells (alpha - 1/2 changes sign between the cell and one of its
// face values, curvature.cpp:40-117); ls = alpha - 1/2
__global__ void __launch_bounds__(TPB) k_curv_iface(
    int nc, int nif, const int* __restrict__ cf_ptr, const int* __restrict__ cf_idx,
    const int* __restrict__ owner, const int* __restrict__ nbr, const double* __restrict__ wCv,
    const double* __restrict__ VF, int8_t* __restrict__ iface, double* __restrict__ ls) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double a = VF[i];
    int flag = 0;
    CELL_FACES_BEGIN(i, cf_ptr, cf_idx, f, own)
        (void)own;
        double fa = a;
        if (f < nif) { double w = wCv[f]; fa = w * VF[owner[f]] + (1.0 - w) * VF[nbr[f]]; }
        if ((a - 0.5) * (fa - 0.5) <= 0.0) flag = 1;
    }
    iface[i] = (int8_t)flag;
    ls[i] = a - 0.5;
}

// interface cells with |grad ls|^2 > 1 are rescaled; pseudo-time step and
// sign velocity of the reinitialisation (curvature.cpp:119-177)
__global__ void __launch_bounds__(TPB) k_curv_prep(
    int nc, const int8_t* __restrict__ iface, const double* __restrict__ g, const double* __restrict__ vol,
    const double* __restrict__ maxA, double* __restrict__ ls, double* __restrict__ tau,
    double* __restrict__ vel0) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double a = ls[i];
    if (iface[i]) {
        const double* gi = g + 3 * (size_t)i;
        double m = 0.0;
        m += gi[0] * gi[0]; m += gi[1] * gi[1]; m += gi[2] * gi[2];
        if (m > 1.0) { a = a / m; ls[i] = a; }
    }
    double v = vol[i];
    double t = v / maxA[i] / 1.0;
    tau[i] = smin(t, pow(v, 0.3));
    double q = fabs(a) * pow(v, 0.333);
    vel0[i] = a / sqrt(a * a + q * q);
}

// unit normal grad/|grad|; zero where |grad| is 0 (mode 0, curvature.cpp:
// 309-319) or not above 1e-15 (mode 1, curvature.cpp:738-750)
__global__ void __launch_bounds__(TPB) k_curv_normal(int nc, int mode, const double* __restrict__ g,
                                                     double* __restrict__ nrm) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    const double* gi = g + 3 * (size_t)i;
    double mg = sqrt(gi[0] * gi[0] + gi[1] * gi[1] + gi[2] * gi[2]);
    bool ok = mode == 0 ? mg != 0.0 : mg > 1.e-15;
    for (int k = 0; k < 3; ++k) nrm[3 * (size_t)i + k] = ok ? gi[k] / mg : 0.0;
}

// one reinitialisation step (curvature.cpp:203-722): the sign-weighted face
// least squares of ls (weights max(1e-5, s n.d) from the cell's side),
// solved per cell, then ls += CFL tau s (1 - |grad ls|) outside the
// interface cells.  Reads ls_in, writes ls_out.
__global__ void __launch_bounds__(TPB) k_curv_reinit(
    int nc, int nif, const int* __restrict__ cf_ptr, const int* __restrict__ cf_idx,
    const int* __restrict__ owner, const int* __restrict__ nbr, const double* __restrict__ cxyz,
    const double* __restrict__ fxyz, const double* __restrict__ nrm, const double* __restrict__ vel0,
    const double* __restrict__ tau, const int8_t* __restrict__ iface, const double* __restrict__ ls_in,
    double* __restrict__ ls_out) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    const double eps2 = 1.e-5, cfl = 0.1;
    double ai = ls_in[i];
    double s = vel0[i] > 0.0 ? 1.0 : -1.0;
    double n0 = nrm[3 * (size_t)i], n1 = nrm[3 * (size_t)i + 1], n2 = nrm[3 * (size_t)i + 2];
    double m[6] = {0, 0, 0, 0, 0, 0}, r[3] = {0, 0, 0};
    CELL_FACES_BEGIN(i, cf_ptr, cf_idx, f, own)
        double d[3], DVar;
        if (f < nif) {
            int o = owner[f], nb = nbr[f];
            for (int a = 0; a < 3; ++a) d[a] = cxyz[3 * (size_t)nb + a] - cxyz[3 * (size_t)o + a];
            DVar = ls_in[nb] - ls_in[o];
        } else {
            for (int a = 0; a < 3; ++a) d[a] = fxyz[3 * (size_t)f + a] - cxyz[3 * (size_t)i + a];
            DVar = ai - ai;
        }
        double wk = 0.0;
        if (own) { wk += s * n0 * (-d[0]); wk += s * n1 * (-d[1]); wk += s * n2 * (-d[2]); }
        else { wk += s * n0 * (d[0]); wk += s * n1 * (d[1]); wk += s * n2 * (d[2]); }
        wk = smax(eps2, wk);
        for (int a = 0, num = 0; a < 3; ++a)
            for (int b = a; b < 3; ++b) m[num++] += wk * d[a] * d[b];
        for (int a = 0; a < 3; ++a) r[a] += wk * d[a] * DVar;
    }
    double inv[6];
    inv_sym3(m, inv);
    double g0 = inv[0] * r[0] + inv[1] * r[1] + inv[2] * r[2];
    double g1 = inv[1] * r[0] + inv[3] * r[1] + inv[4] * r[2];
    double g2 = inv[2] * r[0] + inv[4] * r[1] + inv[5] * r[2];
    if (!iface[i]) {
        double mg = sqrt(g0 * g0 + g1 * g1 + g2 * g2);
        ai += cfl * tau[i] * vel0[i] * (1.0 - mg);
    }
    ls_out[i] = ai;
}

// kappa = -div(n) with the unweighted stencil least squares of each normal
// component (curvature.cpp:2982-3072, 3074-3077); the stencil as sliced ELL
// over packed records xn = [x, y, z, nx, ny, nz, -, -] of 64 B, two
// neighbours per trip, sums in stencil order
__global__ void __launch_bounds__(TPB) k_pack_xn(int nc, const double* __restrict__ cxyz,
                                                 const double* __restrict__ nrm, double* __restrict__ xn) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double2* o = reinterpret_cast<double2*>(xn + 8 * (size_t)i);
    o[0] = make_double2(cxyz[3 * (size_t)i], cxyz[3 * (size_t)i + 1]);
    o[1] = make_double2(cxyz[3 * (size_t)i + 2], nrm[3 * (size_t)i]);
    o[2] = make_double2(nrm[3 * (size_t)i + 1], nrm[3 * (size_t)i + 2]);
}

template <class IDX>
__global__ void __launch_bounds__(TPB) k_curv_kappa(
    int nc, const int* __restrict__ st_sl, const IDX* __restrict__ st_ell, const double* __restrict__ xn,
    const double* __restrict__ kinv, double* __restrict__ kappa) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    const double2* me = reinterpret_cast<const double2*>(xn + 8 * (size_t)i);
    double2 m0 = me[0], m1 = me[1], m2 = me[2];
    double ci[3] = {m0.x, m0.y, m1.x}, ni[3] = {m1.y, m2.x, m2.y};
    double r[3] = {0, 0, 0};
    auto add = [&](const double2& o0, const double2& o1, const double2& o2) {
        double cj[3] = {o0.x, o0.y, o1.x}, nj[3] = {o1.y, o2.x, o2.y};
        for (int a = 0; a < 3; ++a) {
            double d = cj[a] - ci[a];
            double dv = nj[a] - ni[a];
            r[a] += 1.0 * d * dv;
        }
    };
    int s = i >> 6, lane = i & 63;
    int e0 = st_sl[s], e1 = st_sl[s + 1];
    auto nbr = [&](int k) -> int {
        int v = __builtin_nontemporal_load(st_ell + k);
        if constexpr (sizeof(IDX) == 2) return v == -32768 ? -1 : i + v;
        else return v;
    };
    for (int e = e0 + lane; e < e1; e += 128) {
        int j0 = nbr(e);
        if (j0 < 0) break;
        int j1 = e + 64 < e1 ? nbr(e + 64) : -1;
        const double2* o = reinterpret_cast<const double2*>(xn + 8 * (size_t)j0);
        double2 a0 = o[0], a1 = o[1], a2 = o[2], b0, b1, b2;
        if (j1 >= 0) {
            const double2* q = reinterpret_cast<const double2*>(xn + 8 * (size_t)j1);
            b0 = q[0]; b1 = q[1]; b2 = q[2];
        }
        add(a0, a1, a2);
        if (j1 < 0) break;
        add(b0, b1, b2);
    }
    const double* v = kinv + 6 * (size_t)i;
    double g0 = v[0] * r[0] + v[1] * r[1] + v[2] * r[2];
    double g1 = v[1] * r[0] + v[3] * r[1] + v[4] * r[2];
    double g2 = v[2] * r[0] + v[4] * r[1] + v[5] * r[2];
    kappa[i] = -(g0 + g1 + g2);
}

// calcSourceSurfaceTension (source/surfaceTension.cpp:7-229): one
// calcGaussGreen pass of alpha (gradient.cpp:1778-1985) over its LSQ
// gradient g, then sigma kappa grad(alpha) and sigma kappa grad(alpha).U
__global__ void __launch_bounds__(TPB) k_st_source(
    int nc, int nif, CaseK k, const int* __restrict__ cf_ptr, const int* __restrict__ cf_idx,
    const int* __restrict__ owner, const int* __restrict__ nbr, const int* __restrict__ bpatch,
    const semo_patch_bc* __restrict__ bcs, const double* __restrict__ wCv, const double* __restrict__ skew,
    const double* __restrict__ nrm, const double* __restrict__ area, const double* __restrict__ vol,
    const double* __restrict__ fxyz, const double* __restrict__ cxyz, const double* __restrict__ VF,
    const double* __restrict__ bVF, const double* __restrict__ g, const double* __restrict__ kappa,
    const double* __restrict__ U, const double* __restrict__ V, const double* __restrict__ W,
    double* __restrict__ srcS) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double vi = vol[i];
    double gt[3] = {0, 0, 0};
    CELL_FACES_BEGIN(i, cf_ptr, cf_idx, f, own)
        const double* n = nrm + 3 * (size_t)f;
        double A = area[f];
        if (f < nif) {
            int o = owner[f], nb = nbr[f];
            double wCL = wCv[f], wCR = 1.0 - wCL;
            const double* sk = skew + 3 * (size_t)f;
            double v = wCL * VF[o] + wCR * VF[nb];
            for (int a = 0; a < 3; ++a) v += wCL * g[3 * (size_t)o + a] * sk[a];
            for (int a = 0; a < 3; ++a) v += wCR * g[3 * (size_t)nb + a] * sk[a];
            if (own) for (int a = 0; a < 3; ++a) gt[a] += v * n[a] * A / vi;
            else for (int a = 0; a < 3; ++a) gt[a] -= v * n[a] * A / vi;
            continue;
        }
        int j = f - nif;
        semo_patch_bc bc = bcs[bpatch[j]];
        double v = VF[i];
        if (bc.vf_type == SEMO_BC_FIXED_VALUE) v = bc.vf_value;
        else if (bc.vf_type == SEMO_BC_ZERO_GRADIENT) {
            for (int a = 0; a < 3; ++a) v += g[3 * (size_t)i + a] * (fxyz[3 * (size_t)f + a] - cxyz[3 * (size_t)i + a]);
        } else v = 0.5 * (VF[i] + bVF[j]);
        for (int a = 0; a < 3; ++a) gt[a] += v * n[a] * A / vi;
    }
    double s = k.sigma, kp = kappa[i];
    for (int a = 0; a < 3; ++a) srcS[a * (size_t)nc + i] = (s * kp * gt[a]);
    srcS[3 * (size_t)nc + i] = s * kp * (gt[0] * U[i] + gt[1] * V[i] + gt[2] * W[i]);
}

// ------------------------------------------- p' gradient and correction --
// eqPressure.cpp:686-809: two Green-Gauss passes of p' with skew correction
__global__ void __launch_bounds__(TPB) k_gg_face(int nif, const int* __restrict__ owner,
                                                 const int* __restrict__ nbr, const double* __restrict__ wCv,
                                                 const double* __restrict__ skew, const double* __restrict__ pp,
                                                 const double* __restrict__ gp, double* __restrict__ vF) {
    int f = xcd_block() * blockDim.x + threadIdx.x;
    if (f >= nif) return;
    int l = owner[f], r = nbr[f];
    double wCL = wCv[f], wCR = 1.0 - wCL;
    double v = wCL * pp[l] + wCR * pp[r];
    const double* sk = skew + 3 * (size_t)f;
    for (int a = 0; a < 3; ++a) { v += wCL * gp[3 * (size_t)l + a] * sk[a]; v += wCR * gp[3 * (size_t)r + a] * sk[a]; }
    vF[f] = v;
}

__global__ void __launch_bounds__(TPB) k_gg_cell(
    int nc, int nif, const int* __restrict__ cf_ptr, const int* __restrict__ cf_idx,
    const int* __restrict__ owner, const int* __restrict__ bpatch, const semo_patch_bc* __restrict__ bcs,
    const double* __restrict__ nrm, const double* __restrict__ area, const double* __restrict__ vol,
    const double* __restrict__ pp, const double* __restrict__ vF, double* __restrict__ gout) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double g[3] = {0, 0, 0};
    double vi = vol[i];
    CELL_FACES_BEGIN(i, cf_ptr, cf_idx, f, own)
        const double* n = nrm + 3 * (size_t)f;
        double A = area[f];
        if (f < nif) {
            double v = vF[f];
            if (own) { for (int a = 0; a < 3; ++a) g[a] += v * n[a] * A / vi; }
            else { for (int a = 0; a < 3; ++a) g[a] -= v * n[a] * A / vi; }
        } else {
            double v = pp[i];
            if (bcs[bpatch[f - nif]].p_type == SEMO_BC_FIXED_VALUE) v = 0.0;
            for (int a = 0; a < 3; ++a) g[a] += v * n[a] * A / vi;
        }
    }
    for (int a = 0; a < 3; ++a) gout[3 * (size_t)i + a] = g[a];
}


// calcGaussGreen (gradient.cpp:1778-1985) of one cell field, gathered per
// cell in ascending face order (the order in which the reference's face loops
// accumulate into each cell).  Internal faces: wC phi_P + (1 - wC) phi_N plus
// the skewness correction with the input gradient.  Boundary faces by the
// field's patch type: fixedValue -> the patch value, zeroGradient -> phi_P +
// grad_P . (x_f - x_P), any other type -> the mean of phi_P and the face's
// right value.  sel: 0 = P, 1..3 = U, V, W, 4 = VF.
__global__ void __launch_bounds__(TPB) k_gauss_green(
    int nc, int nif, int sel, const int* __restrict__ cf_ptr, const int* __restrict__ cf_idx,
    const int* __restrict__ owner, const int* __restrict__ nbr, const int* __restrict__ bpatch,
    const semo_patch_bc* __restrict__ bcs, const double* __restrict__ nrm, const double* __restrict__ area,
    const double* __restrict__ wCv, const double* __restrict__ skew, const double* __restrict__ fxyz,
    const double* __restrict__ cxyz, const double* __restrict__ vol, const double* __restrict__ phi,
    const double* __restrict__ faceR, const double* __restrict__ gin, double* __restrict__ gout) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double g[3] = {0, 0, 0};
    double vi = vol[i];
    CELL_FACES_BEGIN(i, cf_ptr, cf_idx, f, own)
        const double* n = nrm + 3 * (size_t)f;
        double A = area[f];
        if (f < nif) {
            int l = owner[f], r = nbr[f];
            double wCL = wCv[f], wCR = 1.0 - wCL;
            const double* sk = skew + 3 * (size_t)f;
            double v = wCL * phi[l] + wCR * phi[r];
            for (int a = 0; a < 3; ++a) v += wCL * gin[3 * (size_t)l + a] * sk[a];
            for (int a = 0; a < 3; ++a) v += wCR * gin[3 * (size_t)r + a] * sk[a];
            if (own) { for (int a = 0; a < 3; ++a) g[a] += v * n[a] * A / vi; }
            else { for (int a = 0; a < 3; ++a) g[a] -= v * n[a] * A / vi; }
        } else {
            int j = f - nif;
            semo_patch_bc bc = bcs[bpatch[j]];
            int type = sel == 0 ? bc.p_type : sel == 4 ? bc.vf_type : bc.u_type;
            double val = sel == 0 ? bc.p_value : sel == 4 ? bc.vf_value : bc.u_value[sel - 1];
            double v = phi[i];
            if (type == SEMO_BC_FIXED_VALUE) v = val;
            else if (type == SEMO_BC_ZERO_GRADIENT) {
                for (int a = 0; a < 3; ++a) v += gin[3 * (size_t)i + a] * (fxyz[3 * (size_t)f + a] - cxyz[3 * (size_t)i + a]);
            } else v = 0.5 * (phi[i] + faceR[j]);
            for (int a = 0; a < 3; ++a) g[a] += v * n[a] * A / vi;
        }
    }
    for (int a = 0; a < 3; ++a) gout[3 * (size_t)i + a] = g[a];
}

// y += x
__global__ void __launch_bounds__(TPB) k_add_to(int n, const double* __restrict__ x, double* __restrict__ y) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i < n) y[i] += x[i];
}

// ----------------------------------------------------------------- LES --
// SEMO_TurbModel_LES (turbulenceModels/LES.cpp) from the Green-Gauss
// gradients gU, gV, gW [3n] (two passes from zero, LES.cpp:46-51): the
// subgrid viscosity muT and kinetic energy kSGS.  MODEL 1: calcWALE
// (:29-131), 2: calcSmagorinsky (:141-220).  The reference's arithmetic and
// constants as written, including WALE's SSd, which adds the squares of S's
// diagonal to those of Sd's off-diagonals (:110-112).
__device__ inline double les_strain(const double* __restrict__ gU, const double* __restrict__ gV,
                                    const double* __restrict__ gW, size_t i, double (&S)[3][3]) {
    const double* a = gU + 3 * i; const double* b = gV + 3 * i; const double* c = gW + 3 * i;
    S[0][0] = a[0]; S[0][1] = 0.5 * (a[1] + b[0]); S[0][2] = 0.5 * (a[2] + c[0]);
    S[1][0] = S[0][1]; S[1][1] = b[1]; S[1][2] = 0.5 * (b[2] + c[1]);
    S[2][0] = S[0][2]; S[2][1] = S[1][2]; S[2][2] = c[2];
    return S[0][0] * S[0][0] + S[1][1] * S[1][1] + S[2][2] * S[2][2] +
           2.0 * (S[0][1] * S[0][1] + S[0][2] * S[0][2] + S[1][2] * S[1][2]);
}
template <int MODEL>
__global__ void __launch_bounds__(TPB) k_les_local(int nc, const double* __restrict__ gU,
                                                   const double* __restrict__ gV, const double* __restrict__ gW,
                                                   const double* __restrict__ vol, const double* __restrict__ Rho,
                                                   double* __restrict__ muT, double* __restrict__ kSGS) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double S[3][3];
    double SS = les_strain(gU, gV, gW, i, S);
    double Sabs = sqrt(2.0 * SS);
    double v = vol[i];
    if constexpr (MODEL == 1) {
        const double Ci = 0.007, Cw = 0.325;
        const double* a = gU + 3 * (size_t)i; const double* b = gV + 3 * (size_t)i;
        const double* c = gW + 3 * (size_t)i;
        double Sd01 = 0.5 * (a[1] * a[1] + b[0] * b[0]);
        double Sd02 = 0.5 * (a[2] * a[2] + c[0] * c[0]);
        double Sd12 = 0.5 * (b[2] * b[2] + c[1] * c[1]);
        double SSd = S[0][0] * S[0][0] + S[1][1] * S[1][1] + S[2][2] * S[2][2] +
                     2.0 * (Sd01 * Sd01 + Sd02 * Sd02 + Sd12 * Sd12);
        double Ls = Cw * pow(v, 0.333333333);
        double Ls2 = Ls * Ls;
        double t = pow(SS, 2.5) + pow(SSd, 1.25);
        muT[i] = t != 0.0 ? Rho[i] * Ls2 * pow(SSd, 1.5) / t : 0.0;
        kSGS[i] = Ci * pow(v, 0.6666666666) * (0.5 * Sabs * Sabs);
    } else {
        const double C_I = 0.007, C_R = 0.1;
        double D2 = pow(v, 0.666666666);
        muT[i] = C_R * Rho[i] * D2 * Sabs;
        kSGS[i] = C_I * D2 * (0.5 * Sabs * Sabs);
    }
}
// calcDynamicSmagorinsky (LES.cpp:224-467): box-filtered (cell plus its
// vertex stencil, volume-weighted) S, |S| S, rho u_i u_j, rho and u give the
// Germano identity's L and M and C_R = -1/2 <L M> / <M M>.  Kept as written:
// the running sums named sumV and sumU collect U and V (:330-331, :415-416).
__global__ void __launch_bounds__(TPB) k_les_dynamic(
    int nc, const int* __restrict__ st_ptr, const int* __restrict__ st_idx, const double* __restrict__ gU,
    const double* __restrict__ gV, const double* __restrict__ gW, const double* __restrict__ vol,
    const double* __restrict__ Rho, const double* __restrict__ U, const double* __restrict__ V,
    const double* __restrict__ W, double* __restrict__ muT, double* __restrict__ kSGS) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    const double C_I = 0.007;
    double S[3][3], M[3][3];
    double L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double SS = les_strain(gU, gV, gW, i, S);
    double Sabs = sqrt(2.0 * SS);
    double v = vol[i], r = Rho[i], u = U[i], vv = V[i], w = W[i];
    L[0][0] = r * u * u * v; L[0][1] = r * u * vv * v; L[0][2] = r * u * w * v;
    L[1][1] = r * vv * vv * v; L[1][2] = r * vv * w * v; L[2][2] = r * w * w * v;
    double sumVol = v, sumRho = r * v, sumV = u * v, sumU = vv * v, sumW = w * v, sumSabs = Sabs * v;
    double t062 = pow(v, 0.66666666);
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            M[j][k] = -t062 * Sabs * S[j][k] * v;
            S[j][k] *= v;
        }
    for (int q = st_ptr[i]; q < st_ptr[i + 1]; ++q) {
        int j = st_idx[q];
        double T[3][3];
        double SSs = les_strain(gU, gV, gW, j, T);
        double vs = vol[j];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) S[a][b] += T[a][b] * vs;
        double Sabss = sqrt(2.0 * SSs);
        double t06 = pow(vs, 0.66666666);
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) M[a][b] -= t06 * Sabss * T[a][b] * vs;
        sumSabs += Sabss * vs;
        double rs = Rho[j], us = U[j], vs2 = V[j], ws = W[j];
        L[0][0] += rs * us * us * vs; L[0][1] += rs * us * vs2 * vs; L[0][2] += rs * us * ws * vs;
        L[1][1] += rs * vs2 * vs2 * vs; L[1][2] += rs * vs2 * ws * vs; L[2][2] += rs * ws * ws * vs;
        sumRho += rs * vs; sumV += us * vs; sumU += vs2 * vs; sumW += ws * vs; sumVol += vs;
    }
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) { M[j][k] /= sumVol; S[j][k] /= sumVol; L[j][k] /= sumVol; }
    sumRho /= sumVol; sumV /= sumVol; sumU /= sumVol; sumW /= sumVol; sumSabs /= sumVol;
    L[0][0] -= sumRho * sumU * sumU; L[0][1] -= sumRho * sumU * sumV; L[0][2] -= sumRho * sumU * sumW;
    L[1][1] -= sumRho * sumV * sumV; L[1][2] -= sumRho * sumV * sumW; L[2][2] -= sumRho * sumW * sumW;
    L[1][0] = L[0][1]; L[2][0] = L[0][2]; L[2][1] = L[1][2];
    double sumDelta = pow(sumVol, 0.66666666666);
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) M[j][k] = sumDelta * sumSabs * S[j][k] - M[j][k];
    double s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) { s1 += L[j][k] * M[j][k]; s2 += M[j][k] * M[j][k]; }
    s1 /= 9.0; s2 /= 9.0;
    double C_R = s2 != 0.0 ? -0.5 * s1 / s2 : 0.0;
    double D2 = pow(v, 0.666666666);
    muT[i] = C_R * r * D2 * Sabs;
    kSGS[i] = C_I * D2 * (0.5 * Sabs * Sabs);
}

__device__ inline double clipv(double v, double lim) {
    if (v <= -lim) v = -lim;
    if (v >= lim) v = lim;
    return v;
}

// eqPressure.cpp:823-883
__global__ void __launch_bounds__(TPB) k_p_correct(
    int nc, CaseK k, const double* __restrict__ pp, const double* __restrict__ gp,
    const double* __restrict__ Rho, double* __restrict__ P, double* __restrict__ U,
    double* __restrict__ V, double* __restrict__ W, double* __restrict__ part) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    double nd = 0.0, np = 0.0;
    if (i < nc) {
        double x = pp[i];
        double p = P[i] + k.urf_p * x;
        double t1 = k.dt / Rho[i];
        double rU = k.urf_pu * t1 * gp[3 * (size_t)i], rV = k.urf_pu * t1 * gp[3 * (size_t)i + 1],
               rW = k.urf_pu * t1 * gp[3 * (size_t)i + 2];
        double u = U[i] - rU, v = V[i] - rV, w = W[i] - rW;
        p = clipv(p, k.lim); u = clipv(u, k.lim); v = clipv(v, k.lim); w = clipv(w, k.lim);
        P[i] = p; U[i] = u; V[i] = v; W[i] = w;
        nd = x * x + rU * rU + rV * rV + rW * rW;
        np = p * p + u * u + v * v + w * w;
    }
    block_sum2(nd, np, part);
}

// eqMomentum.cpp:1029-1085
__global__ void __launch_bounds__(TPB) k_mom_update(int nc, CaseK k, const double* __restrict__ X,
                                                    double* __restrict__ U, double* __restrict__ V,
                                                    double* __restrict__ W, double* __restrict__ part) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    double nd = 0.0, np = 0.0;
    if (i < nc) {
        double x0 = X[i], x1 = X[(size_t)nc + i], x2 = X[2 * (size_t)nc + i];
        double u = clipv(U[i] + k.urf_u * x0, k.lim);
        double v = clipv(V[i] + k.urf_u * x1, k.lim);
        double w = clipv(W[i] + k.urf_u * x2, k.lim);
        U[i] = u; V[i] = v; W[i] = w;
        nd = x0 * x0 + x1 * x1 + x2 * x2;
        np = u * u + v * v + w * w;
    }
    block_sum2(nd, np, part);
}

// block partials of sum phi_i V_i and of sum V_i
__global__ void __launch_bounds__(TPB) k_vol_integral(int nc, const double* __restrict__ phi,
                                                      const double* __restrict__ vol, double* __restrict__ part) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    double a = 0.0, v = 0.0;
    if (i < nc) { v = vol[i]; a = phi[i] * v; }
    block_sum2(a, v, part);
}

__global__ void k_zero(size_t n, uint32_t* __restrict__ d) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = 0u;
}

__global__ void k_copy(size_t n, double* __restrict__ d, const double* __restrict__ s) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = s[i];
}


double finish_ratio(semo_ctx* c, int nb) {
    std::vector<double> h(2 * (size_t)nb);
    SEMO_HIP(hipMemcpyAsync(h.data(), c->red.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SEMO_HIP(hipStreamSynchronize(c->stream));
    double a = 0, b = 0;
    for (int k = 0; k < nb; ++k) { a += h[2 * k]; b += h[2 * k + 1]; }
    return sqrt(a) / sqrt(b);
}

}  // namespace

// ---------------------------------------------------------------- drivers --
// algorithmic bytes of one launch (each operand read once, each result
// written once; face data counted once although a cell-loop kernel reaches
// an internal face from both of its cells)
namespace {
// k_lsq1: stencil indices, the neighbours' [x, y, z, phi] records, the LSQ
// coefficients, the gradient written, the boundary faces' centroid and value
double lsq1_bytes(const semo_ctx* c) {
    double n = c->nc, S = (double)c->d_st_idx.n, nb = c->nbf;
    return (c->d_st_dl.n ? 2.0 : 4.0) * S + 4.0 * (n / 64 + 1) + 32.0 * n + 48.0 * n + 24.0 * n + 4.0 * n +
           nb * (4.0 + 24.0 + 8.0);
}
}  // namespace

void fields_alloc(semo_ctx* c) {
    size_t n = c->nc, nb = c->nbf;
    for (DBuf<double>* b : {&c->P, &c->U, &c->V, &c->W, &c->VF, &c->Rho, &c->Mu, &c->oldP, &c->oldU,
                            &c->oldV, &c->oldW, &c->oldVF, &c->Qm0, &c->Qm1, &c->Qm2, &c->kappa,
                            &c->mxP, &c->mnP, &c->mxU, &c->mnU, &c->mxV, &c->mnV, &c->mxW, &c->mnW})
        { b->alloc(n); b->zero(c->stream); }
    c->srcG.alloc(4 * n); c->srcG.zero(c->stream);
    c->srcS.alloc(4 * n); c->srcS.zero(c->stream);
    for (DBuf<double>* b : {&c->gP, &c->gU, &c->gV, &c->gW, &c->gVF, &c->gpp, &c->gtmp})
        { b->alloc(3 * n); b->zero(c->stream); }
    for (DBuf<double>* b : {&c->bP, &c->bVF, &c->bMu, &c->bRho}) { b->alloc(nb); b->zero(c->stream); }
    c->bUL.alloc(3 * nb); c->bUR.alloc(3 * nb);
    c->Bvec.alloc(3 * n); c->Bvec.zero(c->stream);
    c->Xvec.alloc(3 * n); c->Xvec.zero(c->stream);
    c->fscr.alloc(8 * (size_t)std::max(c->nif, 1));
    reduction_alloc(c, (long)n);
    std::vector<double> one(n, 1.0);
    c->VF.upload(one, c->stream);
    c->oldVF.upload(one, c->stream);
    c->cpk.alloc(8 * n);
    SEMO_LAUNCH(c, "setup_pack_xyz", 88.0 * c->nc, k_pack_xyz, dim3(nblk(c->nc)), dim3(TPB), c->nc, c->c_xyz.p,
                c->cpk.p);
    c->Qm3.alloc(n); c->Qm3.zero(c->stream);
    c->fVFL.alloc(std::max(c->nif, 1)); c->fVFR.alloc(std::max(c->nif, 1));
    c->c_maxA.alloc(n);
    SEMO_LAUNCH(c, "setup_cell_max_area", cell_faces_bytes(c, 8.0) + 8.0 * c->nc, k_cell_max_area, dim3(nblk(c->nc)),
                dim3(TPB), c->nc, c->d_cf_sl.p, c->d_cf_ell.p, c->f_area.p, c->c_maxA.p);
    SEMO_HIP(hipStreamSynchronize(c->stream));   // host sources above are temporaries
}

// calcIncomCellEOSVF + calcCellTransport (eos.cpp:10-151, transport.cpp:
// 9-97): rho, the clipped volume fraction and mu of the mixture; with an LES
// model, then muT from the velocity gradients (the boundary faces' right
// values bR [3][nbf], null: zero, as before SEMO's first reconstruction)
// and mu += muT
void eos_transport(semo_ctx* c, const double* bR) {
    CaseK k = case_k(c);
    SEMO_LAUNCH(c, "eos", 32.0 * c->nc, k_eos, dim3(nblk(c->nc)), dim3(TPB), c->nc, k, c->VF.p, c->Rho.p, c->Mu.p);
    if (!c->les_model) return;
    size_t n = c->nc;
    c->muT.ensure(n); c->kSGS.ensure(n);
    if (!bR) {
        size_t nb = 3 * (size_t)std::max(c->nbf, 1);
        if (c->les_zero.n != nb) { c->les_zero.alloc(nb); zero_d(c, c->les_zero.p, c->les_zero.bytes()); }
        bR = c->les_zero.p;
    }
    les_device(c, c->les_model, c->U.p, c->V.p, c->W.p, c->Rho.p, bR, c->muT.p, c->kSGS.p);
    SEMO_LAUNCH(c, "les_add", 24.0 * n, k_add_to, dim3(nblk(c->nc)), dim3(TPB), c->nc, c->muT.p, c->Mu.p);
}

// cell.var[C] as calcCellEOSVF leaves it at load
void sound_speed(semo_ctx* c) {
    c->Csnd.alloc(c->nc);
    SEMO_LAUNCH(c, "sound_speed", 16.0 * c->nc, k_sound_speed, dim3(nblk(c->nc)), dim3(TPB), c->nc, case_k(c),
                c->VF.p, c->Csnd.p);
    c->csnd_valid = true;
}

void recon_boundary(semo_ctx* c) {
    if (!c->csnd_valid) sound_speed(c);
    if (c->nbf == 0) return;
    for (const semo_patch_bc& b : c->bcs)
        SEMO_CHECK(b.vf_type != SEMO_BC_FUNCTION || c->b_vffun.n == (size_t)c->nbf, SEMO_EINVAL_BC,
                   "volume-fraction function boundary without points (semo_set_vf_function_points)");
    CaseK k = case_k(c);
    SEMO_LAUNCH(c, "recon_boundary", 120.0 * c->nbf, k_recon_boundary, dim3(nblk(c->nbf)), dim3(TPB), c->nbf, c->nif,
                k, c->d_bpatch.p, c->d_bcs.p, c->d_owner.p, c->f_nrm.p, c->f_xyz.p, c->P.p, c->U.p, c->V.p, c->W.p,
                c->VF.p, c->Mu.p, c->b_ufix.n ? c->b_ufix.p : nullptr, c->Csnd.p, c->b_vffun.p, c->bP.p, c->bUL.p,
                c->bUR.p, c->bVF.p, c->bMu.p, c->bRho.p);
}

void gauss_green_standalone(semo_ctx* c, int sel, const double* phi, const double* faceR_b, const double* gin,
                            double* gout) {
    SEMO_LAUNCH(c, "gauss_green", cell_faces_bytes(c, 8.0 + 24.0 + 8.0 + 8.0 + 24.0) + 8.0 * c->nbf + 80.0 * c->nc,
                k_gauss_green, dim3(nblk(c->nc)), dim3(TPB), c->nc, c->nif, sel, c->d_cf_sl.p, c->d_cf_ell.p,
                c->d_owner.p, c->d_nbr.p, c->d_bpatch.p, c->d_bcs.p, c->f_nrm.p, c->f_area.p, c->f_wC.p, c->f_skew.p,
                c->f_xyz.p, c->c_xyz.p, c->c_vol.p, phi, faceR_b, gin, gout);
}

// SEMO_TurbModel_LES on device arrays (device numbering): the Green-Gauss
// gradients of U, V, W (two passes from zero, the boundary faces' right
// values bR [3][nbf] in device face order), then model 1 (WALE),
// 2 (Smagorinsky) or 3 (dynamic Smagorinsky) into muT, kSGS [n]
void les_device(semo_ctx* c, int model, const double* U, const double* V, const double* W, const double* Rho,
                const double* bR, double* muT, double* kSGS) {
    SEMO_CHECK(model >= 1 && model <= 3, SEMO_EINVAL, "LES model must be 1 (WALE), 2 (Smagorinsky) or 3 (dynamic)");
    size_t n = c->nc, nb = c->nbf;
    c->les_g.ensure(12 * n);
    double* g = c->les_g.p;              // [gU, gV, gW, scratch] x 3n
    const double* phi[3] = {U, V, W};
    for (int k = 0; k < 3; ++k) {
        double* gk = g + 3 * n * k;
        double* t = g + 9 * n;
        zero_d(c, t, 3 * n * sizeof(double));
        gauss_green_standalone(c, 1 + k, phi[k], bR + nb * k, t, gk);
        gauss_green_standalone(c, 1 + k, phi[k], bR + nb * k, gk, t);
        copy_d(c, gk, t, 3 * n);
    }
    const double *gU = g, *gV = g + 3 * n, *gW = g + 6 * n;
    dim3 gb(nblk(c->nc)), tb(TPB);
    if (model == 1)
        SEMO_LAUNCH(c, "les", 104.0 * n, k_les_local<1>, gb, tb, c->nc, gU, gV, gW, c->c_vol.p, Rho, muT, kSGS);
    else if (model == 2)
        SEMO_LAUNCH(c, "les", 104.0 * n, k_les_local<2>, gb, tb, c->nc, gU, gV, gW, c->c_vol.p, Rho, muT, kSGS);
    else
        SEMO_LAUNCH(c, "les", 4.0 * (double)c->d_st_idx.n + 4.0 * (n + 1) + 136.0 * n, k_les_dynamic, gb, tb,
                    c->nc, c->d_st_ptr.p, c->d_st_idx.p, gU, gV, gW, c->c_vol.p, Rho, U, V, W, muT, kSGS);
}

void vf_function_values(semo_ctx* c, int n, const double* pts4) {
    if (n == 0) { c->b_vffun.release(); return; }
    if (c->nbf == 0) return;
    DBuf<double> p;
    p.upload(pts4, 4 * (size_t)n, c->stream);
    c->b_vffun.alloc(c->nbf);
    SEMO_LAUNCH(c, "vf_function_values", 32.0 * c->nbf + 32.0 * n, k_vf_nearest, dim3(nblk(c->nbf)), dim3(TPB),
                c->nbf, c->nif, c->f_xyz.p, n, p.p, c->b_vffun.p);
    SEMO_HIP(hipStreamSynchronize(c->stream));   // p is freed on return
}

void source_gravity(semo_ctx* c) {
    CaseK k = case_k(c);
    if (k.g0 == 0.0 && k.g1 == 0.0 && k.g2 == 0.0) { zero_d(c, c->srcG.p, c->srcG.bytes()); return; }
    SEMO_LAUNCH(c, "gravity", cell_faces_bytes(c, 8.0 + 24.0 + 8.0 + 8.0 + 48.0) + 72.0 * c->nbf + 72.0 * c->nc,
                k_gravity, dim3(nblk(c->nc)), dim3(TPB), c->nc, c->nif, k, c->d_cf_ptr.p, c->d_cf_idx.p, c->d_owner.p,
                c->d_nbr.p, c->f_nrm.p, c->f_area.p, c->f_wC.p, c->f_pf.p, c->f_nf.p, c->c_vol.p, c->U.p, c->V.p,
                c->W.p, c->VF.p, c->bUL.p, c->bUR.p, c->bVF.p, c->nbf, c->srcG.p);
}

void grad_minmax(semo_ctx* c, bool with_p, bool uvw_only) {
    if (uvw_only) {
        SEMO_LAUNCH(c, "pack3", 48.0 * c->nc, k_pack3, dim3(nblk(c->nc)), dim3(TPB), c->nc, c->U.p, c->V.p, c->W.p,
                    c->cpk.p);
    } else {
        SEMO_LAUNCH(c, "pack5", 80.0 * c->nc, k_pack5, dim3(nblk(c->nc)), dim3(TPB), c->nc, c->P.p, c->U.p, c->V.p,
                    c->W.p, c->VF.p, c->cpk.p);
    }
    // algorithmic bytes of a k-field LSQ gradient: 4 S (stencil indices) +
    // the neighbours' records (read once: 64n) + k 24n (gradients written) +
    // 48n (coefficients), k = 5 or 3, plus the stencil bounds written (6 or 8
    // arrays) and the boundary faces
    bool narrow = c->d_st_dl.n > 0;
    double n = c->nc, S = (double)c->d_st_idx.n, k = uvw_only ? 3.0 : 5.0;
    double bytes = (narrow ? 2.0 : 4.0) * S + 64.0 * n + k * 24.0 * n + 48.0 * n + (with_p ? 64.0 : 48.0) * n +
                   (double)c->nbf * (24.0 + 8.0 * k + 4.0);
    const char* tag = uvw_only ? "grad_minmax_uvw" : "grad_minmax";
    auto go = [&](auto kern, const auto* idx) {
        SEMO_LAUNCH(c, tag, bytes, kern, dim3(nblk(c->nc)), dim3(TPB), c->nc, c->nif, c->nbf,
                    (with_p && !uvw_only) ? 1 : 0, c->d_st_sl.p, idx, c->d_bf_ptr.p, c->d_bf_idx.p, c->cpk.p,
                    c->f_xyz.p, c->c_lsq.p, c->bP.p, c->bUR.p, c->bVF.p, c->gP.p, c->gU.p, c->gV.p, c->gW.p, c->gVF.p,
                    c->mxP.p, c->mnP.p, c->mxU.p, c->mnU.p, c->mxV.p, c->mnV.p, c->mxW.p, c->mnW.p);
    };
    if (narrow) {
        if (uvw_only) go(k_grad_minmax<3, int16_t>, c->d_st_dl.p);
        else go(k_grad_minmax<5, int16_t>, c->d_st_dl.p);
    } else {
        if (uvw_only) go(k_grad_minmax<3, int>, c->d_st_ell.p);
        else go(k_grad_minmax<5, int>, c->d_st_ell.p);
    }
    // boundary values the P and VF gradients were computed with, where a
    // boundary type makes them depend on U (checked by a later U, V, W-only
    // pass, which keeps those gradients)
    if (!uvw_only && c->nbf) {
        BcDeps dep = bc_deps(c);
        if (dep.p) { c->bP_g.ensure(c->nbf); copy_d(c, c->bP_g.p, c->bP.p, c->nbf); }
        if (dep.vf) { c->bVF_g.ensure(c->nbf); copy_d(c, c->bVF_g.p, c->bVF.p, c->nbf); }
    }
}

// A U, V, W-only pass keeps the P and VF gradients of the last full pass.
// Their boundary values can have moved with U since (P "switch": the owner's
// Mach number, VF inletOutlet: the owner's outflow): a gradient whose
// boundary values changed is recomputed on the device, decided there, so
// no host round trip.  (Same kernels and summation order as the full pass.)
void refresh_u_dependent_grads(semo_ctx* c) {
    if (!c->nbf) return;
    BcDeps dep = bc_deps(c);
    if (!dep.p && !dep.vf) return;
    c->d_bflag.ensure(2);
    SEMO_LAUNCH(c, "lsq1_refresh_check", 8.0, k_clear2, dim3(1), dim3(1), c->d_bflag.p);
    SEMO_LAUNCH(c, "lsq1_refresh_check", 32.0 * c->nbf, k_bchanged, dim3(nblk(c->nbf)), dim3(TPB), c->nbf,
                dep.p ? c->bP.p : nullptr, c->bP_g.p, dep.vf ? c->bVF.p : nullptr, c->bVF_g.p, c->d_bflag.p);
    if (dep.p) lsq1(c, c->P.p, c->bP.p, c->nif, c->gP.p, c->d_bflag.p);
    if (dep.vf) lsq1(c, c->VF.p, c->bVF.p, c->nif, c->gVF.p, c->d_bflag.p + 1);
}

void assemble_momentum(semo_ctx* c, bool after_vof, bool alpha_grad_ready) {
    CaseK k = case_k(c);
    // right after the VOF stage only alpha has changed: its gradient alone is
    // recomputed; those of P, U, V, W and their bounds are the VOF stage's
    if (after_vof) {
        if (!alpha_grad_ready) lsq1(c, c->VF.p, c->bVF.p, c->nif, c->gVF.p);
        // gVF now goes with the current boundary values (the surface-tension
        // source computed its copy from the same ones)
        if (c->nbf && bc_deps(c).vf) { c->bVF_g.ensure(c->nbf); copy_d(c, c->bVF_g.p, c->bVF.p, c->nbf); }
    } else {
        grad_minmax(c, true, false);
    }
    if (c->nif) {
        double bytes = (double)c->nif * (13 * 8 + 8 + 8 + 8 * 8) + (double)c->nc * 8 * (4 + 12 + 8 + 2);
        SEMO_LAUNCH(c, "mom_face", bytes, k_mom_face, dim3(nblk(c->nif)), dim3(TPB), c->nif, k, face_in(c), c->Mu.p,
                    c->d_slot_own.p, c->d_slot_nbr.p, c->Amat.val.p, c->fscr.p);
    }
    {
        double bytes = (double)c->nc * 8 * (1 + 1 + 3 + 3 + 3 + 4 + 4) + 8.0 * 8 * c->nif + 4.0 * c->h_cf_idx.size();
        SEMO_LAUNCH(c, "mom_cell", bytes, k_mom_cell, dim3(nblk(c->nc)), dim3(TPB), c->nc, c->nif, c->nbf, k,
                    c->d_cf_sl.p, c->d_cf_ell.p, c->d_owner.p, c->d_bpatch.p, c->d_bcs.p, c->c_vol.p, c->f_nrm.p,
                    c->f_ePN.p, c->f_area.p, c->f_magPN.p, c->f_alpha.p, c->Rho.p, c->P.p, c->U.p, c->V.p, c->W.p,
                    c->oldU.p, c->oldV.p, c->oldW.p, c->Qm0.p, c->Qm1.p, c->Qm2.p, c->srcG.p, c->srcS.p,
                    c->alphaP.n ? c->alphaP.p : nullptr, c->gP.p, c->gU.p, c->gV.p, c->gW.p, c->bP.p, c->bUL.p,
                    c->bUR.p, c->bMu.p, c->bRho.p, c->fscr.p, c->Amat.diag.p, c->Bvec.p);
    }
    c->last_eq = 0;
}

void assemble_pressure(semo_ctx* c, bool after_momentum) {
    CaseK k = case_k(c);
    grad_minmax(c, false, after_momentum);
    if (after_momentum) refresh_u_dependent_grads(c);
    if (c->nif) {
        double bytes = (double)c->nif * (13 * 8 + 8 + 8 + 2 * 8) + (double)c->nc * 8 * (4 + 12 + 6 + 2);
        SEMO_LAUNCH(c, "p_face", bytes, k_p_face, dim3(nblk(c->nif)), dim3(TPB), c->nif, k, face_in(c),
                    c->d_slot_own.p, c->d_slot_nbr.p, c->Amat.val.p, c->fscr.p);
    }
    {
        SEMO_LAUNCH(c, "p_cell", (double)c->nc * 16 + 16.0 * c->nif + 4.0 * c->h_cf_idx.size(), k_p_cell,
                    dim3(nblk(c->nc)), dim3(TPB), c->nc, c->nif, c->nbf, k, c->d_cf_sl.p, c->d_cf_ell.p, c->d_owner.p,
                    c->d_bpatch.p, c->d_bcs.p, c->f_nrm.p, c->f_ePN.p, c->f_area.p, c->f_magPN.p, c->f_alpha.p,
                    c->P.p, c->gP.p, c->bP.p, c->bUL.p, c->bUR.p, c->bRho.p, c->fscr.p, c->U.p, c->V.p, c->W.p,
                    c->Csnd.p, c->Amat.diag.p, c->Bvec.p);
    }
    c->last_eq = 1;
}

// setIncomValuesLeftRightFaceWithVfMSTACS (reconIncom.cpp:345-430) and the
// calcVolfracEq assembly (eqVolfrac.cpp:9-765).  The LSQ gradient of alpha
// that MSTACS needs is the one of the five-field pass the assembly needs
// anyway (same field, same boundary values).
void assemble_vof(semo_ctx* c) {
    CaseK k = case_k(c);
    grad_minmax(c, true, false);       // with the P bounds the momentum stage after it reuses
    if (c->nif) {
        {
            SEMO_LAUNCH(c, "mstacs_face", (double)c->nif * (8 * 8 + 8 * 3 + 16) + (double)c->nc * 8 * 12,
                        k_mstacs_face, dim3(nblk(c->nif)), dim3(TPB), c->nif, k, c->d_owner.p, c->d_nbr.p, c->f_wC.p,
                        c->f_nrm.p, c->f_skew.p, c->c_xyz.p, c->c_vol.p, c->c_maxA.p, c->U.p, c->V.p, c->W.p, c->VF.p,
                        c->gVF.p, c->fVFL.p, c->fVFR.p);
        }
        double bytes = (double)c->nif * (13 * 8 + 8 + 8 + 16 + 3 * 8) + (double)c->nc * 8 * (4 + 12 + 6 + 2);
        SEMO_LAUNCH(c, "vof_face", bytes, k_vof_face, dim3(nblk(c->nif)), dim3(TPB), c->nif, k, face_in(c), c->fVFL.p,
                    c->fVFR.p, c->d_slot_own.p, c->d_slot_nbr.p, c->Amat.val.p, c->fscr.p);
    }
    {
        SEMO_LAUNCH(c, "vof_cell", (double)c->nc * 8 * 6 + 24.0 * c->nif + 4.0 * c->h_cf_idx.size(), k_vof_cell,
                    dim3(nblk(c->nc)), dim3(TPB), c->nc, c->nif, c->nbf, k, c->d_cf_sl.p, c->d_cf_ell.p,
                    c->d_bpatch.p, c->d_bcs.p, c->c_vol.p, c->f_nrm.p, c->f_ePN.p, c->f_area.p, c->f_magPN.p,
                    c->f_alpha.p, c->P.p, c->U.p, c->V.p, c->W.p, c->VF.p, c->oldVF.p, c->Qm3.p, c->gP.p, c->bVF.p,
                    c->bUL.p, c->bUR.p, c->bRho.p, c->fscr.p, c->Amat.diag.p, c->Bvec.p);
    }
    c->last_eq = 2;
}

// the stencil pass of the one-field LSQ gradient over packed records xph
void lsq1_launch(semo_ctx* c, const double* xph, const double* faceR, int foff, double* g, const int* gate) {
    int n = c->nc;
    int G = c->lsq_lanes_eff;     // 1: k_lsq1; 2, 4: k_lsq1g
    const int* sl = c->d_st_sl.p;
    const int* bp = c->d_bf_ptr.p;
    const int* bi = c->d_bf_idx.p;
    dim3 tb(TPB), gb(nblk((long)n * G));
    if (c->d_st_dl.n) {
        const int16_t* st = c->d_st_dl.p;
        if (G == 4)
            SEMO_LAUNCH(c, "lsq1", lsq1_bytes(c), (k_lsq1g<int16_t, 4, 2>), gb, tb, n, sl, st, c->d_st_ptr.p, bp, bi,
                        xph, c->f_xyz.p, c->c_lsq.p, faceR, foff, g, gate);
        else if (G == 2)
            SEMO_LAUNCH(c, "lsq1", lsq1_bytes(c), (k_lsq1g<int16_t, 2, 4>), gb, tb, n, sl, st, c->d_st_ptr.p, bp, bi,
                        xph, c->f_xyz.p, c->c_lsq.p, faceR, foff, g, gate);
        else
            SEMO_LAUNCH(c, "lsq1", lsq1_bytes(c), k_lsq1<int16_t>, gb, tb, n, sl, st, bp, bi, xph, c->f_xyz.p,
                        c->c_lsq.p, faceR, foff, g, gate);
    } else {
        const int* st = c->d_st_ell.p;
        if (G == 4)
            SEMO_LAUNCH(c, "lsq1", lsq1_bytes(c), (k_lsq1g<int, 4, 2>), gb, tb, n, sl, st, c->d_st_ptr.p, bp, bi, xph,
                        c->f_xyz.p, c->c_lsq.p, faceR, foff, g, gate);
        else if (G == 2)
            SEMO_LAUNCH(c, "lsq1", lsq1_bytes(c), (k_lsq1g<int, 2, 4>), gb, tb, n, sl, st, c->d_st_ptr.p, bp, bi, xph,
                        c->f_xyz.p, c->c_lsq.p, faceR, foff, g, gate);
        else
            SEMO_LAUNCH(c, "lsq1", lsq1_bytes(c), k_lsq1<int>, gb, tb, n, sl, st, bp, bi, xph, c->f_xyz.p, c->c_lsq.p,
                        faceR, foff, g, gate);
    }
}

// one field's LSQ gradient (k_lsq1): packs [x, y, z, phi], then the
// stencil pass; faceR indexed by face - foff, null for a zero boundary
// difference
void lsq1(semo_ctx* c, const double* phi, const double* faceR, int foff, double* g, const int* gate) {
    int n = c->nc;
    c->xph.ensure(4 * (size_t)n);
    dim3 gb(nblk(n)), tb(TPB);
    SEMO_LAUNCH(c, "lsq1_pack", 64.0 * n, k_pack_xphi, gb, tb, n, c->c_xyz.p, phi, c->xph.p, gate);
    lsq1_launch(c, c->xph.p, faceR, foff, g, gate);
}

// calcSourceSurfaceTension with calcCurvature (curvature.cpp:7-3537,
// surfaceTension.cpp:7-229).  With sigma = 0 every use of kappa vanishes, so
// the pipeline is skipped and the sources are zero.
bool surface_tension(semo_ctx* c) {
    if (c->cs.sigma == 0.0) {
        zero_d(c, c->srcS.p, c->srcS.bytes());
        zero_d(c, c->kappa.p, c->kappa.bytes());
        return false;
    }
    int n = c->nc;
    for (DBuf<double>* b : {&c->ls, &c->ls2, &c->tau_ls, &c->vel0_ls}) b->ensure(n);
    c->cnrm.ensure(3 * (size_t)n); c->cgrad.ensure(3 * (size_t)n); c->iface.ensure(n);
    dim3 gb(nblk(n)), tb(TPB);
    if (c->kinv.n != 6 * (size_t)n) {
        c->kinv.alloc(6 * (size_t)n);
        SEMO_LAUNCH(c, "setup_kappa_lsq", 4.0 * (double)c->d_st_idx.n + 24.0 * n + 48.0 * n, k_kappa_lsq_init, gb, tb,
                    n, c->d_st_ptr.p, c->d_st_idx.p, c->d_bf_ptr.p, c->d_bf_idx.p, c->c_xyz.p, c->f_xyz.p, c->kinv.p);
    }
    auto lsq_input = [&](const double* phi, double* g) { lsq1(c, phi, nullptr, 0, g); };
    SEMO_LAUNCH(c, "curv_iface", cell_faces_bytes(c, 16.0) + 17.0 * n, k_curv_iface, gb, tb, n, c->nif, c->d_cf_sl.p,
                c->d_cf_ell.p, c->d_owner.p, c->d_nbr.p, c->f_wC.p, c->VF.p, c->iface.p, c->ls.p);
    lsq_input(c->ls.p, c->cgrad.p);
    SEMO_LAUNCH(c, "curv_prep", 73.0 * n, k_curv_prep, gb, tb, n, c->iface.p, c->cgrad.p, c->c_vol.p, c->c_maxA.p,
                c->ls.p, c->tau_ls.p, c->vel0_ls.p);
    for (int it = 0; it < 4; ++it) {                       // LSReinitIterMax (curvature.cpp:16)
        lsq_input(c->ls.p, c->cgrad.p);
        SEMO_LAUNCH(c, "curv_normal", 48.0 * n, k_curv_normal, gb, tb, n, 0, c->cgrad.p, c->cnrm.p);
        SEMO_LAUNCH(c, "curv_reinit", cell_faces_bytes(c, 32.0) + 97.0 * n, k_curv_reinit, gb, tb, n, c->nif,
                    c->d_cf_sl.p, c->d_cf_ell.p, c->d_owner.p, c->d_nbr.p, c->c_xyz.p, c->f_xyz.p, c->cnrm.p,
                    c->vel0_ls.p, c->tau_ls.p, c->iface.p, c->ls.p, c->ls2.p);
        std::swap(c->ls, c->ls2);
    }
    lsq_input(c->ls.p, c->cgrad.p);
    SEMO_LAUNCH(c, "curv_normal", 48.0 * n, k_curv_normal, gb, tb, n, 1, c->cgrad.p, c->cnrm.p);
    c->xn.ensure(8 * (size_t)n);
    SEMO_LAUNCH(c, "curv_pack", 112.0 * n, k_pack_xn, gb, tb, n, c->c_xyz.p, c->cnrm.p, c->xn.p);
    if (c->d_st_dl.n)
        SEMO_LAUNCH(c, "curv_kappa", 2.0 * (double)c->d_st_idx.n + 48.0 * n + 48.0 * n + 8.0 * n,
                    k_curv_kappa<int16_t>, gb, tb, n, c->d_st_sl.p, c->d_st_dl.p, c->xn.p, c->kinv.p, c->kappa.p);
    else
        SEMO_LAUNCH(c, "curv_kappa", 4.0 * (double)c->d_st_idx.n + 48.0 * n + 48.0 * n + 8.0 * n, k_curv_kappa<int>,
                    gb, tb, n, c->d_st_sl.p, c->d_st_ell.p, c->xn.p, c->kinv.p, c->kappa.p);
    // grad(alpha): LSQ with the boundary values (into the alpha gradient of
    // the assembly, which is the same), then one Green-Gauss pass
    lsq1(c, c->VF.p, c->bVF.p, c->nif, c->gVF.p);
    SEMO_LAUNCH(c, "st_source", cell_faces_bytes(c, 8.0 + 8.0 + 24.0 + 24.0 + 8.0) + 40.0 * c->nbf + 128.0 * n,
                k_st_source, gb, tb, n, c->nif, case_k(c), c->d_cf_sl.p, c->d_cf_ell.p, c->d_owner.p, c->d_nbr.p,
                c->d_bpatch.p, c->d_bcs.p, c->f_wC.p, c->f_skew.p, c->f_nrm.p, c->f_area.p, c->c_vol.p, c->f_xyz.p,
                c->c_xyz.p, c->VF.p, c->bVF.p, c->gVF.p, c->kappa.p, c->U.p, c->V.p, c->W.p, c->srcS.p);
    return true;
}

// calcTimeStepFromCorantNumber / calcCorantNumberForPrint: returns the
// Courant number; with adjust the case's time step becomes maxCo times the
// smallest time scale, capped at maxTimeStep
double time_step(semo_ctx* c, int adjust) {
    CaseK k = case_k(c);
    int nb = nblk(c->nc);
    SEMO_LAUNCH(c, "time_scale", cell_faces_bytes(c, 8.0) + 56.0 * c->nc + 16.0 * nb, k_time_scale, dim3(nb),
                dim3(TPB), c->nc, k, adjust, c->d_cf_sl.p, c->d_cf_ell.p, c->f_area.p, c->c_vol.p, c->U.p, c->V.p,
                c->W.p, c->Mu.p, c->Rho.p, c->red.p);
    std::vector<double> h(2 * (size_t)nb);
    SEMO_HIP(hipMemcpyAsync(h.data(), c->red.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SEMO_HIP(hipStreamSynchronize(c->stream));
    double minDt = 1000000.0;
    for (int q = 0; q < nb; ++q) minDt = std::min(minDt, h[2 * q]);
    if (!adjust) return c->cs.dt / minDt;
    c->cs.dt = c->cs.max_co * minDt;
    double co = c->cs.max_co;
    if (c->cs.dt > c->cs.max_dt) {
        co = c->cs.max_co * c->cs.max_dt / c->cs.dt;
        c->cs.dt = c->cs.max_dt;
    }
    return co;
}

// VF += vofVofURF dVF, clipped; returns sqrt(sum dVF^2) (eqVolfrac.cpp:770-783)
double vof_update(semo_ctx* c) {
    int nb = nblk(c->nc);
    SEMO_LAUNCH(c, "vof_update", 24.0 * c->nc + 16.0 * nb, k_vof_update, dim3(nb), dim3(TPB), c->nc, c->cs.urf_vf,
                c->Xvec.p, c->VF.p, c->red.p);
    std::vector<double> h(2 * (size_t)nb);
    SEMO_HIP(hipMemcpyAsync(h.data(), c->red.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SEMO_HIP(hipStreamSynchronize(c->stream));
    double a = 0;
    for (int q = 0; q < nb; ++q) a += h[2 * q];
    return sqrt(a);
}

double momentum_update(semo_ctx* c) {
    CaseK k = case_k(c);
    int nb = nblk(c->nc);
    SEMO_LAUNCH(c, "mom_update", (3 * 8.0 + 3 * 16.0) * c->nc + 16.0 * nb, k_mom_update, dim3(nb), dim3(TPB), c->nc,
                k, c->Xvec.p, c->U.p, c->V.p, c->W.p, c->red.p);
    return finish_ratio(c, nb);
}

double pressure_correct(semo_ctx* c) {
    CaseK k = case_k(c);
    const double* pp = c->Xvec.p;
    zero_d(c, c->gpp.p, c->gpp.bytes());
    for (int pass = 0; pass < 2; ++pass) {
        if (c->nif)
            SEMO_LAUNCH(c, "gg_face", (8.0 + 8.0 + 24.0 + 8.0) * c->nif + 32.0 * c->nc, k_gg_face, dim3(nblk(c->nif)),
                        dim3(TPB), c->nif, c->d_owner.p, c->d_nbr.p, c->f_wC.p, c->f_skew.p, pp, c->gpp.p, c->fscr.p);
        SEMO_LAUNCH(c, "gg_cell", cell_faces_bytes(c, 24.0 + 8.0) + 8.0 * c->nif + 8.0 * c->nbf + 40.0 * c->nc,
                    k_gg_cell, dim3(nblk(c->nc)), dim3(TPB), c->nc, c->nif, c->d_cf_sl.p, c->d_cf_ell.p, c->d_owner.p,
                    c->d_bpatch.p, c->d_bcs.p, c->f_nrm.p, c->f_area.p, c->c_vol.p, pp, c->fscr.p, c->gtmp.p);
        std::swap(c->gpp, c->gtmp);
    }
    int nb = nblk(c->nc);
    SEMO_LAUNCH(c, "p_correct", (8.0 + 24.0 + 8.0 + 16.0 + 3 * 16.0) * c->nc + 16.0 * nb, k_p_correct, dim3(nb),
                dim3(TPB), c->nc, k, pp, c->gpp.p, c->Rho.p, c->P.p, c->U.p, c->V.p, c->W.p, c->red.p);
    return finish_ratio(c, nb);
}

void lsq_standalone(semo_ctx* c, const double* phi, const double* faceR, double* grad) {
    lsq1(c, phi, faceR, 0, grad);
}

// ------------------------------------------ LSQ gradient without allocation --
// Scratch of semo_least_square_noalloc: the packed [x, y, z, phi] records the
// one-field LSQ pass gathers (32 B per cell), for up to NOALLOC_MAX_CELLS
// cells: 64 MiB, which every GPU the library runs on holds from module load
// until the process ends.  Host code never passes the array's name to a
// kernel (on the host it is a stand-in, not a device address): it uses the
// address hipGetSymbolAddress gives for the current device.
static __device__ __align__(256) double scratch[4 * (size_t)NOALLOC_MAX_CELLS];

// this device's copy of scratch, looked up once per device with that device
// current (the context's creation calls it)
double* noalloc_scratch(int device) {
    static std::mutex mu;
    static double* ptr[64] = {};
    SEMO_CHECK(device >= 0 && device < 64, SEMO_EINVAL, "device index");
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[device]) {
        void* p = nullptr;
        SEMO_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ptr[device] = static_cast<double*>(p);
    }
    return ptr[device];
}

void lsq1_noalloc(semo_ctx* c, const double* phi, const double* faceR_b, double* grad) {
    SEMO_CHECK(c->nc <= NOALLOC_MAX_CELLS, SEMO_EINVAL,
               "semo_least_square_noalloc: more cells than its scratch holds (2097152)");
    SEMO_CHECK(c->noalloc_xph, SEMO_ESTATE, "no-allocation scratch not set up");
    int n = c->nc;
    dim3 gb(nblk(n)), tb(TPB);
    double* xph = c->noalloc_xph;
    SEMO_LAUNCH(c, "lsq1_pack", 64.0 * n, k_pack_xphi, gb, tb, n, c->c_xyz.p, phi, xph, (const int*)nullptr);
    lsq1_launch(c, xph, faceR_b, c->nif, grad, (const int*)nullptr);
    SEMO_HIP(hipGetLastError());
}

void minmax_standalone(semo_ctx* c, const double* phi, double* mx, double* mn) {
    SEMO_LAUNCH(c, "minmax1", 4.0 * (double)c->d_st_idx.n + 4.0 * c->nc + 24.0 * c->nc, k_minmax1, dim3(nblk(c->nc)),
                dim3(TPB), c->nc, c->d_st_ptr.p, c->d_st_idx.p, phi, mx, mn);
}

// sum phi_i V_i in FP64 (block partials on the device, summed on the host in
// block order: the same value on every run); *total_vol = sum V_i
double volume_integral(semo_ctx* c, const double* phi, double* total_vol) {
    int nb = nblk(c->nc);
    SEMO_LAUNCH(c, "volume_integral", 16.0 * c->nc + 16.0 * nb, k_vol_integral, dim3(nb), dim3(TPB), c->nc, phi,
                c->c_vol.p, c->red.p);
    std::vector<double> h(2 * (size_t)nb);
    SEMO_HIP(hipMemcpyAsync(h.data(), c->red.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SEMO_HIP(hipStreamSynchronize(c->stream));
    double a = 0.0, v = 0.0;
    for (int q = 0; q < nb; ++q) { a += h[2 * q]; v += h[2 * q + 1]; }
    if (total_vol) *total_vol = v;
    return a;
}

void zero_d(semo_ctx* c, void* p, size_t bytes) {
    size_t n = bytes / 4;
    SEMO_CHECK(n * 4 == bytes, SEMO_EINVAL, "zero_d: size not a multiple of 4 bytes");
    if (!n) return;
    SEMO_LAUNCH(c, "fill", 4.0 * n, k_zero, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), n,
                static_cast<uint32_t*>(p));
}

void copy_d(semo_ctx* c, double* dst, const double* src, size_t n) {
    if (!n) return;
    SEMO_LAUNCH(c, "copy", 16.0 * n, k_copy, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), n, dst, src);
}

}  // namespace semo
