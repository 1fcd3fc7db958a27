// Density-based implicit updates of SEMO's dual-time solvers
// (src/solvers/linearSolver.cpp) on the device:
//   - calcLinearSolver (:666-725): per cell, res <- (P + 1.5 dtau/dt T)^-1 res;
//   - calcSingleLinearSolver (:729-787): per cell, res <- T^-1 res;
//   - calcLinearSolverLUSGS (:144-389): a forward then a backward Gauss-Seidel
//     sweep over the cells in lib_RCM's order, the off-diagonal blocks being
//     the faces' upwind flux Jacobians (getNumericalFluxJacobianUpwind,
//     :394-640) times dtau / V * area.
// P and T are constructMatrix's (:791-890); the block inverse is
// GaussSeidelSOR's (src/math/math.cpp:141-200): Doolittle LU without
// pivoting, then a forward and a back substitution per unit column.  The
// arithmetic is the reference's, in its order; pow(x, 2.0) is written x * x.
//
// The sweeps are sequential by definition: cell i reads the already-updated
// residuals of its neighbours that come before it (forward) or after it
// (backward).  Their dependency graph is levelled on the host once per mesh:
// a cell's forward level is one more than the largest level of its earlier
// neighbours, so the cells of one level depend only on lower levels and are
// updated by one launch, in any order, with the sequential sweep's result.
// The diagonal block (the same in both sweeps) is formed and inverted once,
// fully parallel, and stored [n][N*N].
//
// Cell state records, [n][12 + 3 nsp] in device order: P, U, V, W, T, rho,
// C, Ht, dRhoDP, dHtDP, dRhoDT, dHtDT, Ur, dtPseudo, Y_0..Y_{nsp-1},
// dRhoDY_0..dRhoDY_{nsp-2}, dHtDY_0..dHtDY_{nsp-2}; residuals [n][4 + nsp].
#include "common.hpp"
#include "kern.hpp"

namespace semo {
namespace {

enum { sP = 0, sU, sV, sW, sT, sRHO, sC, sHT, sDRDP, sDHDP, sDRDT, sDHDT, sUR, sDTP, sY0 };

template <int NSP>
struct LS {
    static constexpr int N = 4 + NSP, K = 12 + 3 * NSP;
};

template <int NSP>
__device__ inline void load_rec(const double* __restrict__ st, int i, double (&s)[LS<NSP>::K]) {
    const double* p = st + (size_t)i * LS<NSP>::K;
#pragma unroll
    for (int k = 0; k < LS<NSP>::K; ++k) s[k] = p[k];
}

// constructMatrix (linearSolver.cpp:791-890)
template <int NSP>
__device__ inline void construct(const double (&s)[LS<NSP>::K], double (&Pm)[4 + NSP][4 + NSP],
                                 double (&Tm)[4 + NSP][4 + NSP]) {
    constexpr int N = 4 + NSP;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) Tm[i][j] = 0.0;
    const double DRDP = s[sDRDP], DHDP = s[sDHDP], DRDT = s[sDRDT], DHDT = s[sDHDT];
    const double Rho = s[sRHO], U = s[sU], V = s[sV], W = s[sW], Ht = s[sHT];
    Tm[0][0] = DRDP; Tm[1][0] = DRDP * U; Tm[2][0] = DRDP * V; Tm[3][0] = DRDP * W;
    Tm[4][0] = DRDP * Ht + Rho * DHDP - 1.0;
    Tm[1][1] = Rho; Tm[4][1] = Rho * U;
    Tm[2][2] = Rho; Tm[4][2] = Rho * V;
    Tm[3][3] = Rho; Tm[4][3] = Rho * W;
    Tm[0][4] = DRDT; Tm[1][4] = DRDT * U; Tm[2][4] = DRDT * V; Tm[3][4] = DRDT * W;
    Tm[4][4] = DRDT * Ht + Rho * DHDT;
#pragma unroll
    for (int a = 0; a < NSP - 1; ++a) {
        const int q = 5 + a;
        const double Y = s[sY0 + a], dRdY = s[sY0 + NSP + a], dHdY = s[sY0 + 2 * NSP - 1 + a];
        Tm[q][0] = DRDP * Y;
        Tm[q][4] = DRDT * Y;
#pragma unroll
        for (int b = 0; b < NSP - 1; ++b) Tm[q][5 + b] = Y * s[sY0 + NSP + b];
        Tm[q][q] += Rho;
        Tm[0][q] = dRdY; Tm[1][q] = dRdY * U; Tm[2][q] = dRdY * V; Tm[3][q] = dRdY * W;
        Tm[4][q] = dRdY * Ht + Rho * dHdY;
    }
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) Pm[i][j] = Tm[i][j];
    const double beta = 1.0 / (s[sUR] * s[sUR]) + DRDP - 1.0 / (s[sC] * s[sC]);
    Pm[0][0] = beta; Pm[1][0] = beta * U; Pm[2][0] = beta * V; Pm[3][0] = beta * W;
    Pm[4][0] = beta * Ht + Rho * DHDP - 1.0;
#pragma unroll
    for (int a = 0; a < NSP - 1; ++a) Pm[5 + a][0] = beta * s[sY0 + a];
}

// getNumericalFluxJacobianUpwind (linearSolver.cpp:394-640), side "R" when
// RIGHT, else "L": 0.5 dF/dQ of that side's state +/- 0.5 of the Roe-type
// dissipation of the face averages (none in the species rows).
template <int NSP, bool RIGHT>
__device__ inline void jacobian(const double (&L)[LS<NSP>::K], const double (&R)[LS<NSP>::K], const double n[3],
                                double (&J)[4 + NSP][4 + NSP]) {
    constexpr int N = 4 + NSP;
    const double UnL = L[sU] * n[0] + L[sV] * n[1] + L[sW] * n[2];
    const double UnR = R[sU] * n[0] + R[sV] * n[1] + R[sW] * n[2];
    const double RT = sqrt(L[sRHO] / R[sRHO]);
    const double Rhohat = 0.5 * (L[sRHO] + R[sRHO]);
    const double Chat = 0.5 * (L[sC] + R[sC]);
    const double Unhat = 0.5 * (UnL + UnR);
    const double hat[3] = {(RT * L[sU] + R[sU]) / (RT + 1.0), (RT * L[sV] + R[sV]) / (RT + 1.0),
                           (RT * L[sW] + R[sW]) / (RT + 1.0)};
    const double Hthat = (RT * L[sHT] + R[sHT]) / (RT + 1.0);
    const double fgL = RIGHT ? 0.0 : 1.0, fgR = RIGHT ? 1.0 : 0.0;
    const double vel[3] = {L[sU] * fgL + R[sU] * fgR, L[sV] * fgL + R[sV] * fgR, L[sW] * fgL + R[sW] * fgR};
    const double HtLR = L[sHT] * fgL + R[sHT] * fgR, RhoLR = L[sRHO] * fgL + R[sRHO] * fgR;
    const double DRDPLR = L[sDRDP] * fgL + R[sDRDP] * fgR, DHDPLR = L[sDHDP] * fgL + R[sDHDP] * fgR;
    const double DRDTLR = L[sDRDT] * fgL + R[sDRDT] * fgR, DHDTLR = L[sDHDT] * fgL + R[sDHDT] * fgR;
    const double UnLR = UnL * fgL + UnR * fgR;
    double YLR[NSP], DRDYLR[NSP], DHDYLR[NSP];
#pragma unroll
    for (int a = 0; a < NSP - 1; ++a) {
        YLR[a] = L[sY0 + a] * fgL + R[sY0 + a] * fgR;
        DRDYLR[a] = L[sY0 + NSP + a] * fgL + R[sY0 + NSP + a] * fgR;
        DHDYLR[a] = L[sY0 + 2 * NSP - 1 + a] * fgL + R[sY0 + 2 * NSP - 1 + a] * fgR;
    }
    // the convective part, fluxLRdot
    double F[N][N];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) F[i][j] = 0.0;
    F[0][0] = DRDPLR * UnLR;
#pragma unroll
    for (int r = 0; r < 3; ++r) F[1 + r][0] = DRDPLR * UnLR * vel[r] + n[r];
    F[4][0] = DRDPLR * UnLR * HtLR + RhoLR * UnLR * DHDPLR;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        F[0][1 + d] = RhoLR * n[d];
#pragma unroll
        for (int r = 0; r < 3; ++r)
            F[1 + r][1 + d] = r == d ? RhoLR * n[d] * vel[r] + RhoLR * UnLR : RhoLR * n[d] * vel[r];
        F[4][1 + d] = RhoLR * n[d] * HtLR + RhoLR * UnLR * vel[d];
    }
    F[0][4] = DRDTLR * UnLR;
#pragma unroll
    for (int r = 0; r < 3; ++r) F[1 + r][4] = DRDTLR * UnLR * vel[r];
    F[4][4] = DRDTLR * UnLR * HtLR + RhoLR * UnLR * DHDTLR;
#pragma unroll
    for (int a = 0; a < NSP - 1; ++a) {
        const int q = 5 + a;
        F[q][0] = DRDPLR * YLR[a] * UnLR;
#pragma unroll
        for (int d = 0; d < 3; ++d) F[q][1 + d] = RhoLR * YLR[a] * n[d];
        F[q][4] = DRDTLR * YLR[a] * UnLR;
#pragma unroll
        for (int b = 0; b < NSP - 1; ++b) F[q][5 + b] = DRDYLR[b] * YLR[a] * UnLR;
        F[q][q] += RhoLR * UnLR;
        F[0][q] = DRDYLR[a] * UnLR;
#pragma unroll
        for (int r = 0; r < 3; ++r) F[1 + r][q] = DRDYLR[a] * UnLR * vel[r];
        F[4][q] = DRDYLR[a] * UnLR * HtLR + RhoLR * UnLR * DHDYLR[a];
    }
    // the dissipation, dissimat (5 x 5)
    const double dissUp = (Chat - fabs(Unhat)) / Rhohat / Chat / Chat;
    const double dissUu = Unhat / Chat;
    const double dissPp = Unhat / Chat;
    const double dissPu = (Chat - fabs(Unhat)) * Rhohat;
    const double AD = fabs(Unhat);
    double D[5][5];
    D[0][0] = AD * DRDPLR + dissUp * Rhohat;
#pragma unroll
    for (int r = 0; r < 3; ++r) D[1 + r][0] = AD * DRDPLR * vel[r] + dissUp * Rhohat * hat[r] + dissPp * n[r];
    D[4][0] = AD * (DRDPLR * HtLR + RhoLR * DHDPLR - 1.0) + dissUp * Rhohat * Hthat + dissPp * Unhat;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        D[0][1 + c] = dissUu * n[c] * Rhohat;
#pragma unroll
        for (int r = 0; r < 3; ++r)
            D[1 + r][1 + c] = r == c ? AD * RhoLR + dissUu * n[c] * Rhohat * hat[r] + dissPu * n[c] * n[r]
                                     : dissUu * n[c] * Rhohat * hat[r] + dissPu * n[c] * n[r];
        D[4][1 + c] = AD * RhoLR * vel[c] + dissUu * n[c] * Rhohat * Hthat + dissPu * n[c] * Unhat;
    }
    D[0][4] = AD * DRDTLR;
#pragma unroll
    for (int r = 0; r < 3; ++r) D[1 + r][4] = AD * DRDTLR * vel[r];
    D[4][4] = AD * (DRDTLR * HtLR + RhoLR * DHDTLR);
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const double d = (i < 5 && j < 5) ? D[i][j] : 0.0;
            J[i][j] = RIGHT ? 0.5 * F[i][j] - 0.5 * d : 0.5 * F[i][j] + 0.5 * d;
        }
}

// GaussSeidelSOR (math.cpp:141-200), in place: A's upper triangle becomes U,
// its strict lower triangle the multipliers L (which the reference keeps in
// a separate matrix; the elimination never reads them back).
template <int N>
__device__ inline void lu_nopivot(double (&A)[N][N]) {
#pragma unroll
    for (int k = 0; k < N - 1; ++k)
#pragma unroll
        for (int i = k + 1; i < N; ++i) {
            const double coeff = A[i][k] / A[k][k];
#pragma unroll
            for (int j = k + 1; j < N; ++j) A[i][j] = A[i][j] - coeff * A[k][j];
            A[i][k] = coeff;
        }
}

// Column k of the inverse: forward substitution of the unit vector e_k
// through L, back substitution through U.
template <int N>
__device__ inline void inverse_column(const double (&A)[N][N], int k, double (&X)[N]) {
    double Dv[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        Dv[i] = i == k ? 1.0 : 0.0;
#pragma unroll
        for (int j = 0; j < i; ++j) Dv[i] = Dv[i] - A[i][j] * Dv[j];
    }
    X[N - 1] = Dv[N - 1] / A[N - 1][N - 1];
#pragma unroll
    for (int i = N - 2; i >= 0; --i) {
        X[i] = Dv[i];
#pragma unroll
        for (int j = N - 1; j >= i + 1; --j) X[i] = X[i] - A[i][j] * X[j];
        X[i] = X[i] / A[i][i];
    }
}

// calcLinearSolver (SINGLE false) / calcSingleLinearSolver (SINGLE true):
// one thread per cell; the inverse's columns are applied as they come, which
// is the reference's k-ordered sum sum_k Ainv[j][k] b[k].
template <int NSP, bool SINGLE>
__global__ __launch_bounds__(TPB) void k_linsol_point(int n, const double* __restrict__ st, double time_step,
                                                      double* __restrict__ res) {
    constexpr int N = 4 + NSP;
    const int i = xcd_block() * TPB + (int)threadIdx.x;
    if (i >= n) return;
    double s[LS<NSP>::K];
    load_rec<NSP>(st, i, s);
    double A[N][N];
    {
        double Pm[N][N], Tm[N][N];
        construct<NSP>(s, Pm, Tm);
        const double f = 1.5 * s[sDTP] / time_step;
#pragma unroll
        for (int a = 0; a < N; ++a)
#pragma unroll
            for (int c = 0; c < N; ++c) A[a][c] = SINGLE ? Tm[a][c] : Pm[a][c] + f * Tm[a][c];
    }
    double b[N], out[N];
#pragma unroll
    for (int a = 0; a < N; ++a) { b[a] = res[(size_t)i * N + a]; out[a] = 0.0; }
    lu_nopivot<N>(A);
#pragma unroll
    for (int k = 0; k < N; ++k) {
        double X[N];
        inverse_column<N>(A, k, X);
#pragma unroll
        for (int j = 0; j < N; ++j) out[j] += X[j] * b[k];
    }
#pragma unroll
    for (int a = 0; a < N; ++a) res[(size_t)i * N + a] = out[a];
}

__device__ inline void face_normal(const double* __restrict__ nrm, int f, bool own_is_owner, double n[3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) n[a] = own_is_owner ? nrm[3 * (size_t)f + a] : -nrm[3 * (size_t)f + a];
}

// LU-SGS diagonal blocks: P + 1.5 (dtau / dt) T + sum over the cell's faces
// (ascending) of dtau / V * J_L * area, inverted, stored [n][N*N].
template <int NSP>
__global__ __launch_bounds__(TPB) void k_lusgs_diag(int n, int nif, const int* __restrict__ cf_ptr,
                                                    const int* __restrict__ cf_idx, const int* __restrict__ owner,
                                                    const int* __restrict__ nbr, const double* __restrict__ nrm,
                                                    const double* __restrict__ area, const double* __restrict__ vol,
                                                    const double* __restrict__ st, double time_step,
                                                    double* __restrict__ dinv) {
    constexpr int N = 4 + NSP;
    const int i = xcd_block() * TPB + (int)threadIdx.x;
    if (i >= n) return;
    double s[LS<NSP>::K];
    load_rec<NSP>(st, i, s);
    const double rDtauPerVol = s[sDTP] / vol[i];
    double DS[N][N];
#pragma unroll
    for (int a = 0; a < N; ++a)
#pragma unroll
        for (int c = 0; c < N; ++c) DS[a][c] = 0.0;
    for (int k = cf_ptr[i]; k < cf_ptr[i + 1]; ++k) {
        const int f = cf_idx[k];
        const int o = owner[f];
        double nv[3];
        face_normal(nrm, f, o == i, nv);
        double J[N][N];
        if (f < nif) {
            double sn[LS<NSP>::K];
            load_rec<NSP>(st, o == i ? nbr[f] : o, sn);
            jacobian<NSP, false>(s, sn, nv, J);
        } else {
            jacobian<NSP, false>(s, s, nv, J);
        }
        const double A = area[f];
#pragma unroll
        for (int a = 0; a < N; ++a)
#pragma unroll
            for (int c = 0; c < N; ++c) DS[a][c] += rDtauPerVol * J[a][c] * A;
    }
    double M[N][N];
    {
        double Pm[N][N], Tm[N][N];
        construct<NSP>(s, Pm, Tm);
        const double timeFrac = s[sDTP] / time_step;
#pragma unroll
        for (int a = 0; a < N; ++a)
#pragma unroll
            for (int c = 0; c < N; ++c) {
                M[a][c] = Pm[a][c] + 1.5 * timeFrac * Tm[a][c];
                M[a][c] += DS[a][c];
            }
    }
    lu_nopivot<N>(M);
#pragma unroll
    for (int k = 0; k < N; ++k) {
        double X[N];
        inverse_column<N>(M, k, X);
#pragma unroll
        for (int j = 0; j < N; ++j) dinv[(size_t)i * N * N + j * N + k] = X[j];
    }
}

// Broadcast within each quad of lanes (DPP quad_perm): every lane gets lane
// SRC's value of its quad.
template <int SRC>
__device__ inline double quad_bcast(double v) {
    constexpr int CTRL = SRC * 0x55;
    int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
    int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// One level of a sweep, four lanes (a quad) per cell.  Slot t of the level
// holds its cell, and the cell's faces toward earlier (FWD) / later cells in
// ascending face order as (neighbour, face) pairs, the face stored as ~f when
// the cell is the face's neighbour (normal reversed): the first four in the
// slot's 48-byte record (one load round gives a lane its cell and its pair),
// the rest in an overflow list.  Lane l of the quad takes pairs l, l + 4, ...: its term
// dtau / V * (J_R r_ngb) * area; the quad adds the terms in pair order (the
// reference's sum), then forward: r <- Dinv (r - sum), backward:
// r <- r - Dinv sum, rows l and l + 4 on lane l.  Control flow is uniform
// within a quad, so every DPP source lane is active.
template <int NSP, bool FWD>
__global__ __launch_bounds__(TPB) void k_lusgs_level(int cnt, const int4* __restrict__ slots,
                                                     const int2* __restrict__ ovf,
                                                     const double* __restrict__ nrm, const double* __restrict__ area,
                                                     const double* __restrict__ vol, const double* __restrict__ st,
                                                     const double* __restrict__ dinv, double* __restrict__ res) {
    constexpr int N = 4 + NSP;
    static_assert(N <= 8, "two rows per lane");
    const int gt = (int)blockIdx.x * TPB + (int)threadIdx.x;
    const int t = gt >> 2, l = gt & 3;
    if (t >= cnt) return;
    // slot record: {cell, pair count, overflow offset, -}, then pairs 0..3
    const int4 h = slots[3 * (size_t)t];
    const int4 qq = slots[3 * (size_t)t + 1 + (l >> 1)];
    const int i = h.x, p1 = h.y;
    const int2 q0 = (l & 1) ? make_int2(qq.z, qq.w) : make_int2(qq.x, qq.y);
    const int p0 = 0;
    double s[LS<NSP>::K];
    load_rec<NSP>(st, i, s);
    const double rDtauPerVol = s[sDTP] / vol[i];
    // the cell's own residual and this lane's Dinv rows do not depend on the
    // neighbours: loaded first, in flight during the neighbour terms
    double ri[N], D0[N], D1[N];
    const double* Di = dinv + (size_t)i * N * N;
#pragma unroll
    for (int a = 0; a < N; ++a) {
        ri[a] = res[(size_t)i * N + a];
        D0[a] = Di[l * N + a];
        D1[a] = l + 4 < N ? Di[(l + 4) * N + a] : 0.0;
    }
    double LU[N];
#pragma unroll
    for (int a = 0; a < N; ++a) LU[a] = 0.0;
    for (int base = p0; base < p1; base += 4) {
        const int p = base + l;
        double c[N];
#pragma unroll
        for (int a = 0; a < N; ++a) c[a] = 0.0;
        if (p < p1) {
            const int2 q = p < 4 ? q0 : ovf[h.z + p - 4];
            const int g = q.x;
            const bool own_is_owner = q.y >= 0;
            const int f = own_is_owner ? q.y : ~q.y;
            double nv[3];
            face_normal(nrm, f, own_is_owner, nv);
            double sn[LS<NSP>::K], J[N][N], rg[N];
            load_rec<NSP>(st, g, sn);
#pragma unroll
            for (int k = 0; k < N; ++k) rg[k] = res[(size_t)g * N + k];
            const double A = area[f];
            jacobian<NSP, true>(s, sn, nv, J);
#pragma unroll
            for (int a = 0; a < N; ++a) {
                double sig = 0.0;
#pragma unroll
                for (int k = 0; k < N; ++k) sig += rg[k] * J[a][k];
                c[a] = rDtauPerVol * sig * A;
            }
        }
#pragma unroll
        for (int a = 0; a < N; ++a) {
            const double v0 = quad_bcast<0>(c[a]), v1 = quad_bcast<1>(c[a]);
            const double v2 = quad_bcast<2>(c[a]), v3 = quad_bcast<3>(c[a]);
            LU[a] += v0;
            if (base + 1 < p1) LU[a] += v1;
            if (base + 2 < p1) LU[a] += v2;
            if (base + 3 < p1) LU[a] += v3;
        }
    }
    double b[N];
#pragma unroll
    for (int a = 0; a < N; ++a) b[a] = FWD ? ri[a] - LU[a] : LU[a];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int j = l + 4 * r;
        if (j < N) {
            double u = 0.0;
#pragma unroll
            for (int k = 0; k < N; ++k) u += (r == 0 ? D0[k] : D1[k]) * b[k];
            if (FWD) res[(size_t)i * N + j] = u;
            else res[(size_t)i * N + j] = ri[j] - u;
        }
    }
}

// lib_RCM's order in device numbering (rank[i] = position of cell i) and
// the forward / backward level lists of the sweeps.
void lusgs_schedule(semo_ctx* c) {
    if (c->ls_ready) return;
    const int nc = c->nc, nif = c->nif;
    std::vector<int> rank(nc), order(nc);
    if (c->reordered) {
        std::iota(rank.begin(), rank.end(), 0);
    } else {   // lib_RCM's adjacency (RCM.cpp:79-126): internal faces in ascending order
        std::vector<int> cnt(nc + 1, 0);
        for (int f = 0; f < nif; ++f) { cnt[c->h_owner[f] + 1]++; cnt[c->h_nbr[f] + 1]++; }
        for (int i = 0; i < nc; ++i) cnt[i + 1] += cnt[i];
        std::vector<int> adj(cnt[nc]), pos(cnt.begin(), cnt.end() - 1);
        for (int f = 0; f < nif; ++f) {
            adj[pos[c->h_owner[f]]++] = c->h_nbr[f];
            adj[pos[c->h_nbr[f]]++] = c->h_owner[f];
        }
        std::vector<int> perm(nc);
        genrcm(nc, cnt.data(), adj.data(), perm.data(), rank.data());
    }
    for (int i = 0; i < nc; ++i) order[rank[i]] = i;
    for (int dir = 0; dir < 2; ++dir) {
        std::vector<int> lev(nc, 0);
        int nlev = 0;
        for (int p = 0; p < nc; ++p) {
            const int i = order[dir == 0 ? p : nc - 1 - p];
            int l = 0;
            for (int k = c->h_cf_ptr[i]; k < c->h_cf_ptr[i + 1]; ++k) {
                const int f = c->h_cf_idx[k];
                if (f >= nif) continue;
                const int g = c->h_owner[f] == i ? c->h_nbr[f] : c->h_owner[f];
                if (dir == 0 ? rank[g] < rank[i] : rank[g] > rank[i]) l = std::max(l, lev[g] + 1);
            }
            lev[i] = l;
            nlev = std::max(nlev, l + 1);
        }
        std::vector<int>& ptr = c->ls_lev_ptr[dir];
        ptr.assign(nlev + 1, 0);
        for (int i = 0; i < nc; ++i) ptr[lev[i] + 1]++;
        for (int l = 0; l < nlev; ++l) ptr[l + 1] += ptr[l];
        std::vector<int> cells(nc), fill(ptr.begin(), ptr.end() - 1);
        for (int p = 0; p < nc; ++p) {   // within a level, in sweep order
            const int i = order[dir == 0 ? p : nc - 1 - p];
            cells[fill[lev[i]]++] = i;
        }
        // per slot: {cell, pair count, overflow offset, 0} and its first four
        // (neighbour, face) pairs toward earlier (later) cells, ascending
        std::vector<int4> rec(3 * (size_t)nc);
        std::vector<int2> ovf;
        for (int t = 0; t < nc; ++t) {
            const int i = cells[t];
            int np = 0;
            int2 q[4] = {};
            const int off = (int)ovf.size();
            for (int k = c->h_cf_ptr[i]; k < c->h_cf_ptr[i + 1]; ++k) {
                const int f = c->h_cf_idx[k];
                if (f >= nif) continue;
                const bool own = c->h_owner[f] == i;
                const int g = own ? c->h_nbr[f] : c->h_owner[f];
                if (!(dir == 0 ? rank[g] < rank[i] : rank[g] > rank[i])) continue;
                const int2 pq = make_int2(g, own ? f : ~f);
                if (np < 4) q[np] = pq;
                else ovf.push_back(pq);
                ++np;
            }
            rec[3 * (size_t)t] = make_int4(i, np, off, 0);
            rec[3 * (size_t)t + 1] = make_int4(q[0].x, q[0].y, q[1].x, q[1].y);
            rec[3 * (size_t)t + 2] = make_int4(q[2].x, q[2].y, q[3].x, q[3].y);
        }
        if (ovf.empty()) ovf.push_back(make_int2(0, 0));
        c->ls_slots[dir].upload(rec, c->stream);
        c->ls_ovf[dir].upload(ovf, c->stream);
    }
    SEMO_HIP(hipStreamSynchronize(c->stream));
    c->ls_ready = true;
}

template <int NSP>
void linsol_run(semo_ctx* c, int method, double time_step, const double* st, double* res) {
    constexpr int N = 4 + NSP, K = 12 + 3 * NSP;
    const int n = c->nc;
    dim3 g(nblk(n)), b(TPB);
    if (method == SEMO_LINSOL_POINT) {
        SEMO_LAUNCH(c, "linsol_point", (double)n * (K + 2 * N) * 8.0, (k_linsol_point<NSP, false>), g, b, n, st,
                    time_step, res);
        return;
    }
    if (method == SEMO_LINSOL_SINGLE) {
        SEMO_LAUNCH(c, "linsol_single", (double)n * (K + 2 * N) * 8.0, (k_linsol_point<NSP, true>), g, b, n, st,
                    time_step, res);
        return;
    }
    lusgs_schedule(c);
    c->ls_dinv.ensure((size_t)N * N * n);
    const double sides = (double)c->h_cf_idx.size();
    SEMO_LAUNCH(c, "lusgs_diag", (double)n * (K + N * N + 3) * 8.0 + sides * (4.0 + 8.0 + 32.0 + K * 8.0),
                k_lusgs_diag<NSP>, g, b, n, c->nif, c->d_cf_ptr.p, c->d_cf_idx.p, c->d_owner.p, c->d_nbr.p,
                c->f_nrm.p, c->f_area.p, c->c_vol.p, st, time_step, c->ls_dinv.p);
    for (int dir = 0; dir < 2; ++dir) {
        const std::vector<int>& ptr = c->ls_lev_ptr[dir];
        const int nlev = (int)ptr.size() - 1;
        for (int l = 0; l < nlev; ++l) {
            const int cnt = ptr[l + 1] - ptr[l];
            if (cnt == 0) continue;
            // per cell: record, Dinv, own residual read + write, slot record;
            // per earlier (later) neighbour: its record and residual, the
            // face's normal and area, the pair (half the sides)
            const double bytes = (double)cnt * ((K + N * N + 2 * N + 1) * 8.0 + 48.0) +
                                 0.5 * (double)cnt * (sides / n) * ((K + N) * 8.0 + 40.0);
            const int4* slots = c->ls_slots[dir].p + 3 * (size_t)ptr[l];
            if (dir == 0)
                SEMO_LAUNCH(c, "lusgs_sweep", bytes, (k_lusgs_level<NSP, true>), dim3(nblk(4L * cnt)), b, cnt, slots,
                            c->ls_ovf[0].p, c->f_nrm.p, c->f_area.p, c->c_vol.p, st, c->ls_dinv.p, res);
            else
                SEMO_LAUNCH(c, "lusgs_sweep", bytes, (k_lusgs_level<NSP, false>), dim3(nblk(4L * cnt)), b, cnt,
                            slots, c->ls_ovf[1].p, c->f_nrm.p, c->f_area.p, c->c_vol.p, st, c->ls_dinv.p, res);
        }
    }
}

// updateValues (src/solvers/densityBased.cpp:366-512), per cell: the
// under-relaxed increments of P, U, V, W, T and the first nsp - 1 mass
// fractions, the limits (P, U, V, W, T: a value at or beyond a bound takes
// it), the mass fractions clipped to [0, 1] and the last one closing the sum.
// cells[n][5 + nsp] = (P, U, V, W, T, MF_0..MF_{nsp-1}); lim = (minP, maxP,
// minU, maxU, minV, maxV, minW, maxW, minT, maxT).
__global__ __launch_bounds__(TPB) void k_update_cells(int n, int nsp, double uP, double uU, double uT, double uMF,
                                                      const double* __restrict__ lim, const double* __restrict__ res,
                                                      double* __restrict__ cells) {
    const int i = xcd_block() * TPB + (int)threadIdx.x;
    if (i >= n) return;
    const int N = 4 + nsp, K = 5 + nsp;
    double* c = cells + (size_t)i * K;
    const double* r = res + (size_t)i * N;
    double v[5] = {c[0] + uP * r[0], c[1] + uU * r[1], c[2] + uU * r[2], c[3] + uU * r[3], c[4] + uT * r[4]};
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        if (v[k] <= lim[2 * k]) v[k] = lim[2 * k];
        if (v[k] >= lim[2 * k + 1]) v[k] = lim[2 * k + 1];
        c[k] = v[k];
    }
    double tmp = 0.0;
    for (int a = 0; a < nsp - 1; ++a) {
        double y = c[5 + a] + uMF * r[5 + a];
        y = y < 1.0 ? y : 1.0;      // std::min(1.0, y): (y < 1) ? y : 1
        y = 0.0 < y ? y : 0.0;      // std::max(0.0, y): (0 < y) ? y : 0
        c[5 + a] = y;
        tmp += y;
    }
    c[5 + nsp - 1] = 1.0 - tmp;
}

// per internal face: Un += URF_U * (wC r_own + (1 - wC) r_ngb) . n over the
// momentum increments, in the caller's orientation (a face the renumbering
// flipped takes the other side's weight and the negated normal)
__global__ __launch_bounds__(TPB) void k_update_un(int nif, int nsp, double uU, const int* __restrict__ owner,
                                                   const int* __restrict__ nbr, const int8_t* __restrict__ flip,
                                                   const double* __restrict__ wC, const double* __restrict__ nrm,
                                                   const double* __restrict__ res, double* __restrict__ un) {
    const int f = xcd_block() * TPB + (int)threadIdx.x;
    if (f >= nif) return;
    const int N = 4 + nsp;
    const bool fl = flip[f] != 0;
    const int o = fl ? nbr[f] : owner[f], g = fl ? owner[f] : nbr[f];
    const double wCL = fl ? 1.0 - wC[f] : wC[f], wCR = 1.0 - wCL;
    double n[3];
    for (int a = 0; a < 3; ++a) n[a] = fl ? -nrm[3 * (size_t)f + a] : nrm[3 * (size_t)f + a];
    const double* ro = res + (size_t)o * N;
    const double* rg = res + (size_t)g * N;
    un[f] += uU * ((wCL * ro[1] + wCR * rg[1]) * n[0] + (wCL * ro[2] + wCR * rg[2]) * n[1] +
                   (wCL * ro[3] + wCR * rg[3]) * n[2]);
}

}  // namespace

void update_values_device(semo_ctx* c, int nsp, const double* urf, const double* d_lim, const double* res,
                          double* cells, double* un) {
    const int N = 4 + nsp;
    SEMO_LAUNCH(c, "update_cells", (double)c->nc * (2.0 * (5 + nsp) + N) * 8.0, k_update_cells,
                dim3(nblk(c->nc)), dim3(TPB), c->nc, nsp, urf[0], urf[1], urf[2], urf[3], d_lim, res, cells);
    if (c->nif)
        SEMO_LAUNCH(c, "update_un", (double)c->nif * (8.0 + 1.0 + 8.0 + 24.0 + 16.0) + 6.0 * 8.0 * c->nif,
                    k_update_un, dim3(nblk(c->nif)), dim3(TPB), c->nif, nsp, urf[1], c->d_owner.p, c->d_nbr.p,
                    c->d_fflip.p, c->f_wC.p, c->f_nrm.p, res, un);
}

void linsol_device(semo_ctx* c, int method, int nsp, double time_step, const double* st, double* res) {
    switch (nsp) {
    case 1: linsol_run<1>(c, method, time_step, st, res); break;
    case 2: linsol_run<2>(c, method, time_step, st, res); break;
    case 3: linsol_run<3>(c, method, time_step, st, res); break;
    case 4: linsol_run<4>(c, method, time_step, st, res); break;
    default: SEMO_CHECK(false, SEMO_EINVAL, "linear solver: 1 <= nsp <= 4");
    }
}

int lusgs_levels(semo_ctx* c, int dir) {
    lusgs_schedule(c);
    return (int)c->ls_lev_ptr[dir].size() - 1;
}

}  // namespace semo
