// Convective flux schemes of SEMO's density-based solvers
// (src/solvers/flux.cpp): per face, from the left and right states
// (P, U, V, W, T, rho, C, Ht and the species fractions Y) and the unit
// normal, the flux of (rho, rho u, rho v, rho w, rho Ht, rho Y_0 ..
// rho Y_{nsp-2}).  One thread per face; the arithmetic is the reference's,
// in its order (pow(x, 2.0) written x * x, which is what it evaluates to),
// including what it does differently from the textbook forms:
//   - calcFluxRoe returns the central part only (its dissipation is
//     commented out, flux.cpp:330-340);
//   - calcFluxAPRoe and calcFluxRoeAMTP put the pressure dissipation's
//     z-momentum term into the energy slot (DP[4] = DP[0] n_z, flux.cpp:
//     456, 616), leaving z-momentum without it;
//   - calcFluxAUSMPWP_N and calcFluxRoeM_N take the sensors w1, w2, w3, the
//     face sound speed fC, a cut-off velocity, the time step and a length
//     (flux.cpp:707-923).
#include "common.hpp"
#include "kern.hpp"

namespace semo {
namespace {

// std::min / std::max as the reference calls them: on ties (including
// +0 / -0) the first argument
__device__ inline double mn(double a, double b) { return b < a ? b : a; }
__device__ inline double mx(double a, double b) { return a < b ? b : a; }

struct FSide {
    double P, U, V, W, T, rho, C, Ht;
    const double* Y;
};

// flux.cpp:924-953
__device__ inline double M_func(double M, double op, double alp) {
    if (fabs(M) > 1.0) return 0.5 * (M + op * fabs(M));
    return op * 0.25 * ((M + op) * (M + op)) + op * alp * ((M * M - 1.0) * (M * M - 1.0));
}
__device__ inline double pre_func(double M, double op, double alp) {
    if (fabs(M) > 1.0) return 0.5 * (1.0 + op * (M > 0.0 ? 1.0 : -1.0));
    return 0.25 * ((M + op) * (M + op)) * (2.0 - op * M) + op * alp * M * ((M * M - 1.0) * (M * M - 1.0));
}
// flux.cpp:635-644
__device__ inline double f_func(double M) {
    return mn(M * sqrt(4.0 + (1.0 - M * M) * (1.0 - M * M)) / (1.0 + M * M), 1.0);
}

// the flux of an upwind mass split (f1L, f1R) with the interface pressure PLR
__device__ inline void split_flux(const FSide& L, const FSide& R, const double* n, double f1L, double f1R,
                                  double PLR, int nsp, double* fl) {
    fl[0] = f1L + f1R;
    fl[1] = f1L * L.U + f1R * R.U + PLR * n[0];
    fl[2] = f1L * L.V + f1R * R.V + PLR * n[1];
    fl[3] = f1L * L.W + f1R * R.W + PLR * n[2];
    fl[4] = f1L * L.Ht + f1R * R.Ht;
    for (int i = 0; i < nsp - 1; ++i) fl[5 + i] = f1L * L.Y[i] + f1R * R.Y[i];
}
// the advection flux of one side, component k
__device__ inline double adv(const FSide& s, double Un, const double* n, int k) {
    switch (k) {
    case 0: return s.rho * Un;
    case 1: return s.rho * Un * s.U + s.P * n[0];
    case 2: return s.rho * Un * s.V + s.P * n[1];
    case 3: return s.rho * Un * s.W + s.P * n[2];
    case 4: return s.rho * Un * s.Ht;
    default: return s.rho * Un * s.Y[k - 5];
    }
}

// calcFluxHAUS (flux.cpp:5-156)
__device__ void flux_haus(const FSide& L, const FSide& R, const double* n, int nsp, double* fl) {
    double UnL = L.U * n[0] + L.V * n[1] + L.W * n[2];
    double UnR = R.U * n[0] + R.V * n[1] + R.W * n[2];
    double chat = 0.5 * (L.C + R.C);
    double ML = UnL / chat, MR = UnR / chat;
    double KLR = sqrt(0.5 * (L.U * L.U + L.V * L.V + L.W * L.W + R.U * R.U + R.V * R.V + R.W * R.W));
    double MLP = M_func(ML, 1.0, 0.0), MRM = M_func(MR, -1.0, 0.0);
    double preP = pre_func(ML, 1.0, 0.0), preM = pre_func(MR, -1.0, 0.0);
    double Mbar = (L.rho * fabs(ML) + R.rho * fabs(MR)) / (L.rho + R.rho);
    double g_c = 1.0 + mx(mn(ML, 0.0), -1.0) * mn(mx(MR, 0.0), 1.0);
    double D_L = ML + (1.0 - g_c) * fabs(ML);
    double D_R = MR - (1.0 - g_c) * fabs(MR);
    double D_rho = Mbar * g_c;
    double preLs = fabs(L.P) + 0.1 * L.rho * L.C * L.C;
    double preRs = fabs(R.P) + 0.1 * R.rho * R.C * R.C;
    double mr = mn(preLs / preRs, preRs / preLs);
    double w = 1.0 - mr * mr;
    double ps = preP * L.P + preM * R.P;
    double pr = mn(L.P / R.P, R.P / L.P);
    double pll = (3.0 / 4.0 <= pr && 1.0 > pr) ? 4.0 * pr - 3.0 : 0.0;
    double fL = fabs(ML) <= 1.0 ? (L.P / ps - 1.0) * pll * fabs(MLP) * mn(1.0, pow((fabs(UnL) / chat), 0.25)) : 0.0;
    double fR = fabs(MR) <= 1.0 ? (L.P / ps - 1.0) * pll * fabs(MRM) * mn(1.0, pow((fabs(UnR) / chat), 0.25)) : 0.0;
    double MPP = MLP + MRM;
    double MLP_AUSM = 0.5 * (MPP + fabs(MPP)), MRM_AUSM = 0.5 * (MPP - fabs(MPP));
    double MLP_SLAU = 0.5 * (D_L + D_rho), MRM_SLAU = 0.5 * (D_R - D_rho);
    double fa1 = w;
    double MLPL = fa1 * MLP_AUSM + (1.0 - fa1) * MLP_SLAU;
    double MRMR = fa1 * MRM_AUSM + (1.0 - fa1) * MRM_SLAU;
    double mdot = L.rho * chat * MLPL + R.rho * chat * MRMR -
                  0.5 * (1.0 - 0.5 * (1.0 - cos(3.141592 * mn(1.0, mx(fabs(ML), fabs(MR)))))) *
                      (1.0 - 0.5 * (1.0 + cos(3.141592 * mn(fabs(L.P / R.P), fabs(R.P / L.P))))) *
                      (R.P - L.P) / chat;
    MLP = 0.5 * (ML + fabs(ML));
    MRM = 0.5 * (MR - fabs(MR));
    double f1L, f1R;
    if (mdot >= 0.0) {
        f1L = mdot - (L.rho * chat * MRM) * (w * (1.0 + fR) - fR + fL);
        f1R = (R.rho * chat * MRM) * (w * (1.0 + fR));
    } else {
        f1L = (L.rho * chat * MLP) * (w * (1.0 + fL));
        f1R = mdot - (R.rho * chat * MLP) * (w * (1.0 + fL) - fL + fR);
    }
    double fa2 = w, gam = 0.6;
    double PLR = 0.5 * (L.P + R.P) - fa2 * (KLR / chat) * 0.5 * preP * preM * 0.5 * (L.P + R.P) / chat * (UnR - UnL) +
                 mx(0.2, gam) * (KLR / chat) * 0.5 * (L.P + R.P) * (preP + preM - 1.0) -
                 0.5 * (preP - preM) * (R.P - L.P);
    split_flux(L, R, n, f1L, f1R, PLR, nsp, fl);
}

// calcFluxSLAU2 (flux.cpp:159-238)
__device__ void flux_slau2(const FSide& L, const FSide& R, const double* n, int nsp, double* fl) {
    double UnL = L.U * n[0] + L.V * n[1] + L.W * n[2];
    double UnR = R.U * n[0] + R.V * n[1] + R.W * n[2];
    double chat = 0.5 * (L.C + R.C);
    double ML = UnL / chat, MR = UnR / chat;
    double KLR = sqrt(0.5 * (L.U * L.U + L.V * L.V + L.W * L.W + R.U * R.U + R.V * R.V + R.W * R.W));
    double preP = pre_func(ML, 1.0, 0.0), preM = pre_func(MR, -1.0, 0.0);
    double Mbar = (L.rho * fabs(ML) + R.rho * fabs(MR)) / (L.rho + R.rho);
    double Mcy = mn(1.0, KLR / chat);
    double phi_c = (1.0 - Mcy) * (1.0 - Mcy);
    double g_c = 1.0 + mx(mn(ML, 0.0), -1.0) * mn(mx(MR, 0.0), 1.0);
    double D_L = ML + (1.0 - g_c) * fabs(ML);
    double D_R = MR - (1.0 - g_c) * fabs(MR);
    double D_rho = Mbar * g_c;
    double MLPL = 0.5 * (D_L + D_rho), MRMR = 0.5 * (D_R - D_rho);
    double mdot = L.rho * chat * MLPL + R.rho * chat * MRMR - 0.5 * phi_c / chat * (R.P - L.P);
    double f1L = mdot >= 0.0 ? mdot : 0.0, f1R = mdot >= 0.0 ? 0.0 : mdot;
    double PLR = 0.5 * (L.P + R.P) - 0.5 * Mcy * preP * preM * 0.5 * (L.P + R.P) / chat * (UnR - UnL) +
                 Mcy * 0.5 * (L.P + R.P) * (preP + preM - 1.0) - 0.5 * (preP - preM) * (R.P - L.P);
    split_flux(L, R, n, f1L, f1R, PLR, nsp, fl);
}

// calcFluxRoe (flux.cpp:243-344): the central part
__device__ void flux_roe(const FSide& L, const FSide& R, const double* n, int nsp, double* fl) {
    double UnL = L.U * n[0] + L.V * n[1] + L.W * n[2];
    double UnR = R.U * n[0] + R.V * n[1] + R.W * n[2];
    for (int k = 0; k < 4 + nsp; ++k) fl[k] = 0.5 * (adv(L, UnL, n, k) + adv(R, UnR, n, k));
}

// the Roe averages of calcFluxAPRoe / calcFluxRoeAMTP
struct RoeAvg {
    double RT, uhat, vhat, what, Hthat, rhohat, chat, unhat;
};
__device__ inline RoeAvg roe_avg(const FSide& L, const FSide& R, double UnL, double UnR) {
    RoeAvg a;
    a.RT = sqrt(R.rho / L.rho);
    a.uhat = (L.U + a.RT * R.U) / (1.0 + a.RT);
    a.vhat = (L.V + a.RT * R.V) / (1.0 + a.RT);
    a.what = (L.W + a.RT * R.W) / (1.0 + a.RT);
    a.Hthat = (L.Ht + a.RT * R.Ht) / (1.0 + a.RT);
    a.rhohat = a.RT * L.rho;
    a.chat = (L.C + a.RT * R.C) / (1.0 + a.RT);
    a.unhat = (UnL + a.RT * UnR) / (1.0 + a.RT);
    return a;
}
// the jump of the conserved variables, component k
__device__ inline double dW(const FSide& L, const FSide& R, int k) {
    switch (k) {
    case 0: return R.rho - L.rho;
    case 1: return R.rho * R.U - L.rho * L.U;
    case 2: return R.rho * R.V - L.rho * L.V;
    case 3: return R.rho * R.W - L.rho * L.W;
    case 4: return R.rho * (R.Ht - R.P / R.rho) - L.rho * (L.Ht - L.P / L.rho);
    default: return R.rho * R.Y[k - 5] - L.rho * L.Y[k - 5];
    }
}
// U' times (1, uhat, vhat, what, Hthat, Yhat...), component k
__device__ inline double dUk(const RoeAvg& a, const FSide& L, const FSide& R, double d0, int k) {
    switch (k) {
    case 0: return d0;
    case 1: return d0 * a.uhat;
    case 2: return d0 * a.vhat;
    case 3: return d0 * a.what;
    case 4: return d0 * a.Hthat;
    default: return d0 * ((L.Y[k - 5] + a.RT * R.Y[k - 5]) / (1.0 + a.RT));
    }
}

// calcFluxAPRoe (flux.cpp:349-470)
__device__ void flux_aproe(const FSide& L, const FSide& R, const double* n, int nsp, double* fl) {
    double UnL = L.U * n[0] + L.V * n[1] + L.W * n[2];
    double UnR = R.U * n[0] + R.V * n[1] + R.W * n[2];
    RoeAvg a = roe_avg(L, R, UnL, UnR);
    double chat = a.chat, unhat = a.unhat, rhohat = a.rhohat;
    double xi = fabs(unhat);
    double Mcy = sqrt(0.5 * (L.U * L.U + L.V * L.V + L.W * L.W + R.U * R.U + R.V * R.V + R.W * R.W)) / chat;
    double theta = mn(mx(Mcy * Mcy, 0.01 * 0.01), 1.0);
    double Um = unhat * mn(mx(Mcy, 0.01), 1.0);
    double Cm = chat * mn(mx(Mcy, 0.01), 1.0);
    double thetad = mn(Mcy * Mcy, 1.0);
    double Ud = 0.5 * (1.0 + thetad);
    double Cd = 0.5 * sqrt(4.0 * chat * chat * thetad + ((1.0 - thetad) * (1.0 - thetad)) * unhat * unhat);
    double DU0 = (Cm - 0.5 * (1.0 - theta) * unhat * Um / Cm - theta * fabs(unhat)) * (R.P - L.P) / rhohat / theta /
                     chat / chat +
                 Ud / Cm * (UnR - UnL);
    DU0 = DU0 * rhohat;
    double DP0 = (Cd - fabs(unhat) + 0.5 * (1.0 - thetad) * unhat * Ud / Cm) * rhohat * (UnR - UnL) +
                 Ud * (R.P - L.P) / Cm;
    for (int k = 0; k < 4 + nsp; ++k) {
        double DP = k == 1 ? DP0 * n[0] : k == 2 ? DP0 * n[1] : k == 4 ? DP0 * n[2] : 0.0;
        fl[k] = 0.5 * (adv(L, UnL, n, k) + adv(R, UnR, n, k)) - 0.5 * (xi * dW(L, R, k) + DP + dUk(a, L, R, DU0, k));
    }
}

// calcFluxRoeAMTP (flux.cpp:480-631)
__device__ void flux_roeamtp(const FSide& L, const FSide& R, const double* n, int nsp, double* fl) {
    double UnL = L.U * n[0] + L.V * n[1] + L.W * n[2];
    double UnR = R.U * n[0] + R.V * n[1] + R.W * n[2];
    RoeAvg a = roe_avg(L, R, UnL, UnR);
    double chat = a.chat, unhat = a.unhat, rhohat = a.rhohat;
    double Morg = 0.5 * (sqrt(L.U * L.U + L.V * L.V + L.W * L.W) / L.C + sqrt(R.U * R.U + R.V * R.V + R.W * R.W) / R.C);
    double Mbar = fabs(unhat) / chat;
    double fef = 0.05 * chat;
    double ksqrt = sqrt((R.U - L.U) * (R.U - L.U) + (R.V - L.V) * (R.V - L.V) + (R.W - L.W) * (R.W - L.W));
    double n1x, n1y, n1z;
    if (ksqrt < 1.e-5) { n1x = n[0]; n1y = n[1]; n1z = n[2]; }
    else { n1x = (R.U - L.U) / ksqrt; n1y = (R.V - L.V) / ksqrt; n1z = (R.W - L.W) / ksqrt; }
    double n2x = -n1y * (n1x * n[1] - n1y * n[0]) - n1z * (n1x * n[2] - n1z * n[0]);
    double n2y = n1x * (n1x * n[1] - n1y * n[0]) - n1z * (n1y * n[2] - n1z * n[1]);
    double n2z = n1x * (n1x * n[2] - n1z * n[0]) + n1y * (n1y * n[2] - n1z * n[1]);
    double a1 = n1x * n[0] + n1y * n[1] + n1z * n[2];
    double a2 = n2x * n[0] + n2y * n[1] + n2z * n[2];
    double U1 = n1x * a.uhat + n1y * a.vhat + n1z * a.what;
    double U2 = n2x * a.uhat + n2y * a.vhat + n2z * a.what;
    double frr = fabs(a1 * U1) + fabs(a2 * U2);
    double fMbar8 = pow(f_func(Mbar), 8.0), fMorg8 = pow(f_func(Morg), 8.0);
    double xi = mx(fabs(0.5 * (UnR + UnL)) + 0.5 * (UnR - UnL), (1.0 - fMbar8) * fMorg8 * mn(fef, frr));
    double Und = fabs(unhat) - fMbar8 * ((unhat + chat > 0.0 ? 1.0 : -1.0) * mx(0.0, UnR - UnL) -
                                         (unhat - chat > 0.0 ? 1.0 : -1.0) * mx(0.0, UnR - UnL)) / 4.0;
    double theta = 1.0;
    double Umul = 0.5 * (1.0 + theta) * unhat;
    double Cmul = 0.5 * sqrt(4.0 * chat * chat * theta + ((1.0 - theta) * (1.0 - theta)) * unhat * unhat);
    double cpi3 = mn(fabs(L.P / R.P), fabs(R.P / L.P));
    double s1 = pow(f_func(cpi3), 8.0);
    double dUp0 = s1 * (1.0 - fMorg8) *
                  (mx(0.0, Cmul - Und) +
                   (1.0 - theta) * (Und - unhat * 0.5 / Cmul * (Umul > 0.0 ? 1.0 : -1.0) * mn(fabs(Umul), Cmul))) *
                  (R.P - L.P) / (theta * chat * chat);
    double xi_LR = 0.5 * (unhat > 0.0 ? 1.0 : -1.0) * mn(Und, chat) * (UnR - UnL) / chat;
    double Mcy = sqrt(0.5 * (L.U * L.U + L.V * L.V + L.W * L.W + R.U * R.U + R.V * R.V + R.W * R.W)) / chat;
    double phi_rer = 1.0 - s1 * (1.0 - f_func(Mcy));
    double dP0 = phi_rer * mx(0.0, chat - Und) * rhohat * (UnR - UnL) +
                 (unhat > 0.0 ? 1.0 : -1.0) * mn(Und, chat) * (R.P - L.P) / chat;
    for (int k = 0; k < 4 + nsp; ++k) {
        double dLR;
        switch (k) {
        case 0: dLR = xi_LR * (L.rho + R.rho); break;
        case 1: dLR = xi_LR * (L.rho * L.U + R.rho * R.U); break;
        case 2: dLR = xi_LR * (L.rho * L.V + R.rho * R.V); break;
        case 3: dLR = xi_LR * (L.rho * L.W + R.rho * R.W); break;
        case 4: dLR = xi_LR * (L.rho * L.Ht + R.rho * R.Ht); break;
        default: dLR = xi_LR * (L.rho * L.Y[k - 5] + R.rho * R.Y[k - 5]); break;
        }
        double dP = k == 1 ? dP0 * n[0] : k == 2 ? dP0 * n[1] : k == 4 ? dP0 * n[2] : 0.0;
        fl[k] = 0.5 * (adv(L, UnL, n, k) + adv(R, UnR, n, k)) -
                0.5 * (xi * dW(L, R, k) + dP + dUk(a, L, R, dUp0, k) + dLR);
    }
}

// calcFluxUpwind (flux.cpp:649-692)
__device__ void flux_upwind(const FSide& L, const FSide& R, const double* n, int nsp, double* fl) {
    double UnL = L.U * n[0] + L.V * n[1] + L.W * n[2];
    double UnR = R.U * n[0] + R.V * n[1] + R.W * n[2];
    double f1L, f1R;
    if (UnL + UnR >= 0.0) { f1L = L.rho * 0.5 * (UnL + UnR); f1R = 0.0; }
    else { f1L = 0.0; f1R = R.rho * 0.5 * (UnL + UnR); }
    split_flux(L, R, n, f1L, f1R, 0.5 * (L.P + R.P), nsp, fl);
}

// the low-Mach scaling of calcFluxAUSMPWP_N / calcFluxRoeM_N
struct Scale {
    double phiP, phiU, aSDST;
};
__device__ inline Scale low_mach_scale(double uhat, double vhat, double what, double chat, double Uco, double dtstep,
                                       double Lch) {
    double Mco = Uco / chat;
    double Mkhat = sqrt(uhat * uhat + vhat * vhat + what * what) / chat;
    double Mun = Lch / 3.141592 / dtstep / chat;
    double thetaTemp = mx(Mkhat, Mco);
    double thetaP = mn(1.0, mx(thetaTemp, Mun));
    double thetaU = mn(1.0, thetaTemp);
    Scale s;
    s.phiP = thetaP * (2.0 - thetaP);
    s.phiU = thetaU * (2.0 - thetaU);
    s.aSDST = 0.1875 * (-4.0 + 5.0 * s.phiU * s.phiU);
    return s;
}

// calcFluxAUSMPWP_N (flux.cpp:707-787)
__device__ void flux_ausmpwp_n(const FSide& L, const FSide& R, const double* n, int nsp, const double* par,
                               double* fl) {
    double w1 = par[0], w2 = par[1], fC = par[3], Uco = par[4], dtstep = par[5], Lch = par[6];
    double UnL = L.U * n[0] + L.V * n[1] + L.W * n[2];
    double UnR = R.U * n[0] + R.V * n[1] + R.W * n[2];
    double RT = sqrt(R.rho / L.rho);
    double rhohat = (L.rho + RT * R.rho) / (1.0 + RT);
    double uhat = (L.U + RT * R.U) / (1.0 + RT);
    double vhat = (L.V + RT * R.V) / (1.0 + RT);
    double what = (L.W + RT * R.W) / (1.0 + RT);
    double chat = fC;
    Scale sc = low_mach_scale(uhat, vhat, what, chat, Uco, dtstep, Lch);
    double ML = UnL / chat, MR = UnR / chat;
    double MLP = M_func(ML, 1.0, 0.125), preP = pre_func(ML, 1.0, sc.aSDST);
    double MRM = M_func(MR, -1.0, 0.125), preM = pre_func(MR, -1.0, sc.aSDST);
    double rhoLR = L.rho;
    double MLR = MLP + MRM;
    if (MLR < 0.0) rhoLR = R.rho;
    w1 = 1.0 - w1 * w1 * w1;
    w2 = 1.0 - w2 * w2;
    double w = mx(w1, w2);
    double fL = L.P / (rhohat * chat * chat) * (1.0 - w) * rhohat / rhoLR / sc.phiP;
    double fR = R.P / (rhohat * chat * chat) * (1.0 - w) * rhohat / rhoLR / sc.phiP;
    double MBLP, MBLM;
    if (MLR >= 0.0) { MBLP = MLP + MRM * ((1.0 - w) * (1.0 + fR) - fL); MBLM = MRM * w * (1.0 + fR); }
    else { MBLP = MLP * w * (1.0 + fL); MBLM = MRM + MLP * ((1.0 - w) * (1.0 + fL) - fR); }
    double f1L = chat * L.rho * MBLP, f1R = chat * R.rho * MBLM;
    double Ku = 0.5;
    double pu = -2.0 * Ku * preP * preM * rhohat * chat * (UnR - UnL);
    double PLR = preP * L.P + preM * R.P + pu * sc.phiU;
    split_flux(L, R, n, f1L, f1R, PLR, nsp, fl);
}

// calcFluxRoeM_N (flux.cpp:795-922)
__device__ void flux_roem_n(const FSide& L, const FSide& R, const double* n, int nsp, const double* par,
                            double* fl) {
    double w1 = par[0], w3 = par[2], fC = par[3], Uco = par[4], dtstep = par[5], Lch = par[6];
    double UnL = L.U * n[0] + L.V * n[1] + L.W * n[2];
    double UnR = R.U * n[0] + R.V * n[1] + R.W * n[2];
    double RT = sqrt(R.rho / L.rho);
    double rhohat = (L.rho + RT * R.rho) / (1.0 + RT);
    double uhat = (L.U + RT * R.U) / (1.0 + RT);
    double vhat = (L.V + RT * R.V) / (1.0 + RT);
    double what = (L.W + RT * R.W) / (1.0 + RT);
    double Hthat = (L.Ht + RT * R.Ht) / (1.0 + RT);
    double Unhat = uhat * n[0] + vhat * n[1] + what * n[2];
    double chat = fC;
    Scale sc = low_mach_scale(uhat, vhat, what, chat, Uco, dtstep, Lch);
    double Mhat = Unhat / chat;
    double h = 1.0 - w3;
    double f = 1.0;
    if (uhat * uhat + vhat * vhat + what * what != 0.0) f = pow(fabs(Mhat), h);
    double g = 1.0;
    if (Mhat != 0.0) g = pow(fabs(Mhat), 1.0 - w1);
    double b1 = mx(Unhat + chat, UnR + chat); b1 = mx(b1, 0.0);
    double b2 = mn(Unhat - chat, UnL - chat); b2 = mn(b2, 0.0);
    double Mtilde = (Mhat > 0.0 ? 1.0 : -1.0) * mn(1.0, fabs(Mhat));
    double b1v = mx(Unhat + chat * sc.phiU, UnR + chat * sc.phiU); b1v = mx(b1v, 0.0);
    double b2v = mn(Unhat - chat * sc.phiU, UnL - chat * sc.phiU); b2v = mn(b2v, 0.0);
    double Mvtilde = (Unhat / chat / sc.phiU > 0.0 ? 1.0 : -1.0) * mn(1.0, fabs(Unhat / chat / sc.phiU));
    double B0 = R.rho - L.rho - f * (R.P - L.P) / chat / chat / sc.phiP;
    for (int k = 0; k < 4 + nsp; ++k) {
        double dWd, B1, B2;
        switch (k) {
        case 0: dWd = R.rho - L.rho; B1 = B0; B2 = 0.0; break;
        case 1: dWd = R.rho * R.U - L.rho * L.U; B1 = B0 * uhat + rhohat * (R.U - L.U);
                B2 = rhohat * (UnR - UnL) * (-n[0]); break;
        case 2: dWd = R.rho * R.V - L.rho * L.V; B1 = B0 * vhat + rhohat * (R.V - L.V);
                B2 = rhohat * (UnR - UnL) * (-n[1]); break;
        case 3: dWd = R.rho * R.W - L.rho * L.W; B1 = B0 * what + rhohat * (R.W - L.W);
                B2 = rhohat * (UnR - UnL) * (-n[2]); break;
        case 4: dWd = R.rho * R.Ht - L.rho * L.Ht; B1 = B0 * Hthat + rhohat * (R.Ht - L.Ht); B2 = 0.0; break;
        default: {
            int i = k - 5;
            dWd = R.rho * R.Y[i] - L.rho * L.Y[i];
            B1 = B0 * ((L.Y[i] + RT * R.Y[i]) / (1.0 + RT)) + rhohat * (R.Y[i] - L.Y[i]);
            B2 = 0.0;
        }
        }
        fl[k] = (b1 * adv(L, UnL, n, k) - b2 * adv(R, UnR, n, k)) / (b1 - b2) +
                b1 * b2 / (b1 - b2) * (dWd - g / (1.0 + fabs(Mtilde)) * B1) -
                b1v * b2v / (b1v - b2v) * g / (1.0 + fabs(Mvtilde)) * B2;
    }
}

// one thread per face: states [n][8 + nsp] (P, U, V, W, T, rho, C, Ht, Y),
// normals [n][3], parameters [n][7] (w1, w2, w3, fC, Uco, dt, Lch; the two
// schemes that take them), fluxes [n][4 + nsp]
template <int SCHEME>
__global__ void __launch_bounds__(TPB) k_flux(int n, int nsp, const double* __restrict__ Ls,
                                              const double* __restrict__ Rs, const double* __restrict__ nrm,
                                              const double* __restrict__ par, double* __restrict__ flux) {
    int f = xcd_block() * blockDim.x + threadIdx.x;
    if (f >= n) return;
    int w = 8 + nsp;
    const double* l = Ls + (size_t)f * w;
    const double* r = Rs + (size_t)f * w;
    FSide L{l[0], l[1], l[2], l[3], l[4], l[5], l[6], l[7], l + 8};
    FSide R{r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r + 8};
    const double* nv = nrm + 3 * (size_t)f;
    double* fl = flux + (size_t)f * (4 + nsp);
    switch (SCHEME) {
    case SEMO_FLUX_HAUS: flux_haus(L, R, nv, nsp, fl); break;
    case SEMO_FLUX_SLAU2: flux_slau2(L, R, nv, nsp, fl); break;
    case SEMO_FLUX_ROE: flux_roe(L, R, nv, nsp, fl); break;
    case SEMO_FLUX_APROE: flux_aproe(L, R, nv, nsp, fl); break;
    case SEMO_FLUX_ROEAMTP: flux_roeamtp(L, R, nv, nsp, fl); break;
    case SEMO_FLUX_UPWIND: flux_upwind(L, R, nv, nsp, fl); break;
    case SEMO_FLUX_AUSMPWP_N: flux_ausmpwp_n(L, R, nv, nsp, par + 7 * (size_t)f, fl); break;
    default: flux_roem_n(L, R, nv, nsp, par + 7 * (size_t)f, fl); break;
    }
}

// calcFluxes (RHS.cpp:98-406), per face: SLAU2 of the face's left and right
// records (P, U, V, W, T, rho, C, Ht, mu, Y_0..Y_{nsp-2}; the last mass
// fraction is 1 - the sum of the others, as calcFluxes forms it) and
// calcViscousFlux (viscousFlux.cpp:5-67) of the face's velocity gradients
// (zero when grad is null); out: (cFlux - vFlux) * area, [nf][4 + nsp].  The
// records and the normal are in the caller's orientation: a face the
// renumbering flipped takes the negated normal.
__global__ void __launch_bounds__(TPB) k_rhs_face(int nf, int nsp, const double* __restrict__ Ls,
                                                  const double* __restrict__ Rs, const double* __restrict__ grad,
                                                  const double* __restrict__ nrm, const double* __restrict__ area,
                                                  const int8_t* __restrict__ flip, double* __restrict__ out) {
    int f = xcd_block() * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const int w = 8 + nsp, N = 4 + nsp;
    // records are stored field-major ([8 + nsp][nf]): coalesced loads
    double l[16], r[16];
    for (int k = 0; k < w; ++k) { l[k] = Ls[(size_t)k * nf + f]; r[k] = Rs[(size_t)k * nf + f]; }
    double YL[8], YR[8];
    double sl = 0.0, sr = 0.0;
    for (int k = 0; k < nsp - 1; ++k) { YL[k] = l[9 + k]; sl += YL[k]; YR[k] = r[9 + k]; sr += YR[k]; }
    YL[nsp - 1] = 1.0 - sl;
    YR[nsp - 1] = 1.0 - sr;
    FSide L{l[0], l[1], l[2], l[3], l[4], l[5], l[6], l[7], YL};
    FSide R{r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], YR};
    double nv[3];
    const bool fl_ = flip[f] != 0;
    for (int a = 0; a < 3; ++a) nv[a] = fl_ ? -nrm[3 * (size_t)f + a] : nrm[3 * (size_t)f + a];
    double cf[12];
    flux_slau2(L, R, nv, nsp, cf);
    // calcViscousFlux: face values at gc = 0.5, stresses from the gradients
    double g[9];
    for (int k = 0; k < 9; ++k) g[k] = grad ? grad[(size_t)k * nf + f] : 0.0;
    const double gc = 0.5;
    const double muF = gc * l[8] + (1.0 - gc) * r[8];
    const double UF = gc * l[1] + (1.0 - gc) * r[1];
    const double VF = gc * l[2] + (1.0 - gc) * r[2];
    const double WF = gc * l[3] + (1.0 - gc) * r[3];
    const double dUdx = g[0], dUdy = g[1], dUdz = g[2], dVdx = g[3], dVdy = g[4], dVdz = g[5];
    const double dWdx = g[6], dWdy = g[7], dWdz = g[8];
    const double txx = dUdx + dUdx - 2.0 / 3.0 * (dUdx + dVdy + dWdz);
    const double txy = dVdx + dUdy, txz = dWdx + dUdz;
    const double tyx = txy, tyy = dVdy + dVdy - 2.0 / 3.0 * (dUdx + dVdy + dWdz), tyz = dVdz + dWdy;
    const double tzx = txz, tzy = tyz, tzz = dWdz + dWdz - 2.0 / 3.0 * (dUdx + dVdy + dWdz);
    const double nx = nv[0], ny = nv[1], nz = nv[2];
    double vf[12];
    for (int k = 0; k < N; ++k) vf[k] = 0.0;
    vf[1] = muF * (txx * nx + txy * ny + txz * nz);
    vf[2] = muF * (tyx * nx + tyy * ny + tyz * nz);
    vf[3] = muF * (tzx * nx + tzy * ny + tzz * nz);
    vf[4] = muF * (txx * UF + txy * VF + txz * WF) * nx + muF * (tyx * UF + tyy * VF + tyz * WF) * ny +
            muF * (tzx * UF + tzy * VF + tzz * WF) * nz;
    const double A = area[f];
    for (int k = 0; k < N; ++k) out[(size_t)f * N + k] = (cf[k] - vf[k]) * A;
}

// per cell: the faces' terms in ascending face order, + at faces the cell
// owns (in the caller's orientation), - at the others (calcFluxes); with
// SRC, then sourceTerms (source.cpp:5-35: the dual-time source from Q, Qn,
// Qm, gravity, the caller's surface-tension force) and calcRHS's scaling by
// -dtPseudo / V (RHS.cpp:17-25).  cells[n][8 + (nsp-1) + 2N] = (rho, U, V,
// W, Ht, P, dtPseudo, VF, Y_0..Y_{nsp-2}, Qn.., Qm..).
template <bool SRC>
__global__ void __launch_bounds__(TPB) k_rhs_cell(int nc, int nif, int nsp, const int* __restrict__ cf_ptr,
                                                  const int* __restrict__ cf_idx, const int* __restrict__,
                                                  const int8_t* __restrict__ flip, const double* __restrict__ fterm,
                                                  const double* __restrict__ vol, const double* __restrict__ cells,
                                                  const double* __restrict__ st, double time_step, double g0,
                                                  double g1, double g2, double* __restrict__ res) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    const int N = 4 + nsp;
    double r[12];
    for (int k = 0; k < N; ++k) r[k] = 0.0;
    // the cell's faces in ascending order from the sliced-ELL lists
    // (coalesced index loads, ownership in the sign of the code)
    CELL_FACES_BEGIN(i, cf_ptr, cf_idx, f, own)
        const double* c = fterm + (size_t)f * N;
        const bool plus = f >= nif || (own != (flip[f] != 0));
        if (plus) for (int k = 0; k < N; ++k) r[k] += c[k];
        else for (int k = 0; k < N; ++k) r[k] -= c[k];
    }
    if (SRC) {
        // cell records field-major ([8 + (nsp-1) + 2N][nc]): field k of
        // cell i at cells[k * nc + i]
        const size_t n = (size_t)nc;
        auto cs = [&](int k) { return cells[(size_t)k * n + i]; };
        const double Rho = cs(0), V = vol[i];
        double Q[12];
        Q[0] = Rho; Q[1] = Rho * cs(1); Q[2] = Rho * cs(2); Q[3] = Rho * cs(3);
        Q[4] = Rho * cs(4) - cs(5);
        for (int a = 0; a < nsp - 1; ++a) Q[5 + a] = Rho * cs(8 + a);
        for (int j = 0; j < N; ++j) r[j] += (1.5 * Q[j] - 2.0 * cs(7 + nsp + j) + 0.5 * cs(7 + nsp + N + j)) / time_step * V;
        r[1] -= Rho * g0 * V;
        r[2] -= Rho * g1 * V;
        r[3] -= Rho * g2 * V;
        if (st)
            for (int a = 0; a < 3; ++a) r[1 + a] -= V * st[3 * (size_t)i + a];
        const double dtp = cs(6);
        for (int k = 0; k < N; ++k) r[k] *= -dtp / V;
    }
    for (int k = 0; k < N; ++k) res[(size_t)i * N + k] = r[k];
}

}  // namespace

void rhs_device(semo_ctx* c, int nsp, bool src, double time_step, const double* g, const double* L, const double* R,
                const double* grad, const double* cells, const double* st, double* fterm, double* res) {
    const int N = 4 + nsp, KC = 8 + (nsp - 1) + 2 * N;
    SEMO_LAUNCH(c, "rhs_face", (double)c->nf * ((2.0 * (8 + nsp) + (grad ? 9 : 0) + 4 + N) * 8.0 + 1.0),
                k_rhs_face, dim3(nblk(c->nf)), dim3(TPB), c->nf, nsp, L, R, grad, c->f_nrm.p, c->f_area.p,
                c->d_fflip.p, fterm);
    const double sides = (double)c->h_cf_idx.size();
    const double cbytes = sides * (4.0 + 4.0 + 1.0 + N * 8.0) + (double)c->nc * (N + 2 + (src ? KC + 1 : 0)) * 8.0;
    if (src)
        SEMO_LAUNCH(c, "rhs_cell", cbytes, k_rhs_cell<true>, dim3(nblk(c->nc)), dim3(TPB), c->nc, c->nif, nsp,
                    c->d_cf_sl.p, c->d_cf_ell.p, c->d_owner.p, c->d_fflip.p, fterm, c->c_vol.p, cells, st,
                    time_step, g[0], g[1], g[2], res);
    else
        SEMO_LAUNCH(c, "rhs_cell", cbytes, k_rhs_cell<false>, dim3(nblk(c->nc)), dim3(TPB), c->nc, c->nif, nsp,
                    c->d_cf_sl.p, c->d_cf_ell.p, c->d_owner.p, c->d_fflip.p, fterm, c->c_vol.p, cells, st,
                    time_step, 0.0, 0.0, 0.0, res);
}

void flux_device(semo_ctx* c, int scheme, int n, int nsp, const double* L, const double* R, const double* nrm,
                 const double* par, double* flux) {
    dim3 g(nblk(n)), b(TPB);
    double bytes = (double)n * (2.0 * (8 + nsp) + 3 + 7 + 4 + nsp) * 8.0;
    switch (scheme) {
    case SEMO_FLUX_HAUS: SEMO_LAUNCH(c, "flux", bytes, k_flux<SEMO_FLUX_HAUS>, g, b, n, nsp, L, R, nrm, par, flux); break;
    case SEMO_FLUX_SLAU2: SEMO_LAUNCH(c, "flux", bytes, k_flux<SEMO_FLUX_SLAU2>, g, b, n, nsp, L, R, nrm, par, flux); break;
    case SEMO_FLUX_ROE: SEMO_LAUNCH(c, "flux", bytes, k_flux<SEMO_FLUX_ROE>, g, b, n, nsp, L, R, nrm, par, flux); break;
    case SEMO_FLUX_APROE: SEMO_LAUNCH(c, "flux", bytes, k_flux<SEMO_FLUX_APROE>, g, b, n, nsp, L, R, nrm, par, flux); break;
    case SEMO_FLUX_ROEAMTP:
        SEMO_LAUNCH(c, "flux", bytes, k_flux<SEMO_FLUX_ROEAMTP>, g, b, n, nsp, L, R, nrm, par, flux); break;
    case SEMO_FLUX_UPWIND: SEMO_LAUNCH(c, "flux", bytes, k_flux<SEMO_FLUX_UPWIND>, g, b, n, nsp, L, R, nrm, par, flux); break;
    case SEMO_FLUX_AUSMPWP_N:
        SEMO_LAUNCH(c, "flux", bytes, k_flux<SEMO_FLUX_AUSMPWP_N>, g, b, n, nsp, L, R, nrm, par, flux); break;
    default: SEMO_LAUNCH(c, "flux", bytes, k_flux<SEMO_FLUX_ROEM_N>, g, b, n, nsp, L, R, nrm, par, flux); break;
    }
}

}  // namespace semo
