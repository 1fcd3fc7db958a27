// SEMO hot-path oracle: a CPU restatement, in FP64, of the reference's
// pressure-based incompressible assemble-and-solve path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP path is compared
// against; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// may load it.  The product (semo_mesh_amd, libsemo_gpu.so) never links it.
//
// Every function states the reference file:line it restates.  The loops keep
// the reference's per-cell accumulation order and the left-to-right grouping
// of its arithmetic, so that on the same mesh the results are bit-identical to
// the compiled reference (pinned in tests/test_oracle_golden.py against fixtures
// the compiled reference wrote).  The linear solves are this file's own
// (ILU0-preconditioned BiCGStab to a tight tolerance); SEMO uses AMGCL, so the
// solved fields agree to the solver tolerance, not bit for bit.
//
// Build: make -C oracle   (g++ -O2 -ffp-contract=off; no fast-math)
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

using std::vector;

// The arithmetic type of everything that depends on the flow state: double,
// or -- in the CPU-adjoint build (-DORC_AD, liborc_ad.so) -- the taped
// reverse-mode type of ad.hpp, which evaluates the same IEEE operations.
#ifdef ORC_AD
#include "ad.hpp"
using Real = ad::adouble;
using ad::val;
#else
using Real = double;
#endif
inline double val(double x) { return x; }

namespace {

// std::max(0.0, std::min(1.0, a)) of the reference's clipping
inline Real clamp01(const Real& a) { return std::max(Real(0.0), std::min(Real(1.0), a)); }

enum BcType { ZERO_GRADIENT = 0, FIXED_VALUE = 1, NO_SLIP = 2, SLIP = 3,
              SURFACE_NORMAL_FIXED_VALUE = 4, INLET_OUTLET = 5,
              SWITCH = 6,        // P: owner value above Mach 1, else the fixed value
              FUNCTION = 7 };    // VF: nearest point of ./VF.csv

struct Patch {
    int start, nfaces;
    int p_type; double p_val;
    int u_type; double u_val[3];
    int vf_type; double vf_val;
};

struct Oracle {
    // ---- mesh ------------------------------------------------------------
    int np = 0, nc = 0, nf = 0, nif = 0;
    vector<double> pts;
    vector<int> owner, nbr, fptr, fpts, cptr, cpts;
    vector<Patch> patches;
    vector<int> cfptr, cfidx;          // cell -> faces, ascending face id
    vector<int> face_patch;            // boundary face -> patch index
    // geometry (src/mesh/geometric.cpp:792-1497)
    vector<double> farea, fnrm, fxyz, fwC, fmagPN, falpha, fePN, fskew, fdist, fvecPF, fvecNF;
    vector<double> cvol, cxyz;
    vector<int> sptr, sidx;            // vertex stencil (geometric.cpp:1781-1841)
    vector<double> lsq;                // 6 coeffs per cell (gradient.cpp:246-430)

    // ---- case --------------------------------------------------------------
    double rho[2] = {1000, 1000}, mu[2] = {1e-3, 1e-3}, sigma = 0, g[3] = {0, 0, 0};
    double dt = 0.01, oldDt = 0.01, old2Dt = 0.01;
    double urf_u = 0.3, urf_p = 0.3, urf_pu = 0.3, lim = 1e10;
    double urf_vf = 1.0;               // vofVofURF (controls/build.cpp:475)
    int adjust_dt = 0;                 // adjustTimeStep == "yes"
    double max_co = 1.0, max_dt = 1.0; // maxCo, maxTimeStep
    double courant = 0.0;              // Courant number of the last time-step pass
    int n_vof = 0;                     // nSpeciesEq: VOF iterations per step (controls/build.cpp:409)
    int iterReal = 0;

    // ---- fields (cell) -------------------------------------------------------
    vector<Real> P, U, V, W, VF, MF, Rho, Mu;
    vector<Real> oldP, oldU, oldV, oldW, oldVF, oldRho, Qm0, Qm1, Qm2, Qm3;
    vector<Real> srcG[4], srcS[4], kappa, LS;
    // Darcy porosity alpha_p of the adjoint harness (no counterpart in the
    // reference): momentum diagonal += alpha_p V, right sides -= alpha_p V U
    vector<Real> alphaP;
    // per-face fixedValue-U values, [3][nbf] over the boundary faces in face
    // order (an inlet profile; the reference's form of one is a UDF,
    // reconIncom.cpp:707-727); used once set (ufix_on)
    vector<double> ufix;
    bool ufix_on = false;
    // cell speed of sound of the mixture EOS, set once at the first step the
    // way calcCellEOSVF sets it at load (eos.cpp:160-293); read by the P
    // "switch" boundary (reconIncom.cpp:618-631, eqPressure.cpp:625-632)
    vector<Real> Csnd;
    // VF "function" boundary values per boundary face (reconIncom.cpp:929-
    // 1019): the VF of the nearest point of the set given by orc_set_vf_points
    vector<double> vffun;
    // turbType LES (controls/build.cpp:100-115): model 1 WALE, 2 Smagorinsky,
    // 3 dynamic Smagorinsky, 0 laminar; the last calcCellTransport's muT, kSGS
    int les_model = 0;
    vector<double> muT, kSGS;
    vector<Real> gP, gU, gV, gW, gVF;    // [3n]
    vector<Real> maxP, minP, maxU, minU, maxV, minV, maxW, minW, maxVF, minVF;
    // face left/right values (src/solvers/reconIncom.cpp:501-1035)
    vector<Real> LP, RP, LU, RU, LV, RV, LW, RW, LVF, RVF, Lmu, Rmu, LRho, RRho;
    vector<double> co_cell;            // per-cell Courant number of the last MSTACS pass
    double vof_norm = 0.0;             // return value of the last calcVolfracEq

    // ---- last assembled system -------------------------------------------------
    vector<int> Arows, Acols;
    vector<Real> Avals, B0, B1, B2, X0, X1, X2;
    int last_iters = 0; double last_res = 0;
    double tol = 1e-13; int maxit = 5000;
};

// ------------------------------------------------------------------------
// geometry: calcUnitNormals_Area3dPolygon (geometric.cpp:16-204)
void face_polygon(const Oracle& o, int f, double nrm[3], double& area, double c[3],
                  double& VSn, double cc[3]) {
    int n = o.fptr[f + 1] - o.fptr[f];
    const int* p = &o.fpts[o.fptr[f]];
    auto X = [&](int k, int a) { return o.pts[3 * (size_t)p[k] + a]; };
    double avg[3];
    for (int a = 0; a < 3; ++a) {             // std::accumulate order, then / size
        double s = 0.0;
        for (int k = 0; k < n; ++k) s += X(k, a);
        avg[a] = s / (double)n;
    }
    double x_tmp[3] = {0, 0, 0}, V0old[3] = {0, 0, 0};
    double un[3] = {0, 0, 0}, ct[3] = {0, 0, 0}, area_tmp = 0, vsn = 0;
    for (int two = 0; two < 5; ++two) {
        un[0] = un[1] = un[2] = 0; ct[0] = ct[1] = ct[2] = 0; area_tmp = 0; vsn = 0;
        double V0[3];
        for (int a = 0; a < 3; ++a)
            V0[a] = two == 0 ? avg[a] : V0old[a] + 1.0 * (x_tmp[a] - V0old[a]);
        for (int a = 0; a < 3; ++a) { V0old[a] = V0[a]; x_tmp[a] = 0.0; }
        double totArea = 0.0;
        for (int i = 0; i < n; ++i) {
            int b = i, cidx = (i == n - 1) ? 0 : i + 1;
            double A[3], Bv[3];
            for (int a = 0; a < 3; ++a) { A[a] = X(b, a) - V0[a]; Bv[a] = X(cidx, a) - V0[a]; }
            double Nx = 0.5 * (A[1] * Bv[2] - A[2] * Bv[1]);
            double Ny = 0.5 * (A[2] * Bv[0] - A[0] * Bv[2]);
            double Nz = 0.5 * (A[0] * Bv[1] - A[1] * Bv[0]);
            un[0] += Nx; un[1] += Ny; un[2] += Nz;
            double ta = sqrtl(Nx * Nx + Ny * Ny + Nz * Nz);
            double r[3];
            for (int a = 0; a < 3; ++a) r[a] = (V0[a] + X(b, a) + X(cidx, a)) / 3.0;
            for (int a = 0; a < 3; ++a) x_tmp[a] += r[a] * ta;
            totArea += ta;
            vsn += V0[0] * Nx; vsn += V0[1] * Ny; vsn += V0[2] * Nz;
            double Nv[3] = {Nx, Ny, Nz};
            for (int a = 0; a < 3; ++a) {
                double s1 = V0[a] + X(b, a), s2 = X(b, a) + X(cidx, a), s3 = X(cidx, a) + V0[a];
                ct[a] += 2.0 * Nv[a] * (s1 * s1 + s2 * s2 + s3 * s3);
            }
        }
        double mag = sqrtl(un[0] * un[0] + un[1] * un[1] + un[2] * un[2]);
        un[0] /= mag; un[1] /= mag; un[2] /= mag;
        area_tmp = mag;
        for (int a = 0; a < 3; ++a) x_tmp[a] /= totArea;
    }
    for (int a = 0; a < 3; ++a) { nrm[a] = un[a]; cc[a] = ct[a]; c[a] = x_tmp[a]; }
    area = area_tmp; VSn = vsn;
}

void build_geometry(Oracle& o) {
    // cell -> faces, ascending (src/load/load.cpp:373-385)
    vector<int> cnt(o.nc + 1, 0);
    for (int f = 0; f < o.nf; ++f) { cnt[o.owner[f] + 1]++; if (f < o.nif) cnt[o.nbr[f] + 1]++; }
    for (int i = 0; i < o.nc; ++i) cnt[i + 1] += cnt[i];
    o.cfptr = cnt; o.cfidx.assign(cnt[o.nc], 0);
    vector<int> pos(cnt.begin(), cnt.end() - 1);
    for (int f = 0; f < o.nf; ++f) {
        o.cfidx[pos[o.owner[f]]++] = f;
        if (f < o.nif) o.cfidx[pos[o.nbr[f]]++] = f;
    }
    o.face_patch.assign(o.nf, -1);
    for (size_t b = 0; b < o.patches.size(); ++b)
        for (int f = o.patches[b].start; f < o.patches[b].start + o.patches[b].nfaces; ++f)
            o.face_patch[f] = (int)b;

    int nf = o.nf, nc = o.nc;
    o.farea.assign(nf, 0); o.fnrm.assign(3 * nf, 0); o.fxyz.assign(3 * nf, 0);
    o.cvol.assign(nc, 0); o.cxyz.assign(3 * nc, 0);
    // geometric.cpp:804-877: volume and centroid by face accumulation
    for (int f = 0; f < nf; ++f) {
        double nrm[3], area, c[3], VSn, cc[3];
        face_polygon(o, f, nrm, area, c, VSn, cc);
        if (!(area >= std::numeric_limits<double>::min())) { fprintf(stderr, "zero face area\n"); }
        o.farea[f] = area;
        for (int a = 0; a < 3; ++a) { o.fnrm[3 * f + a] = nrm[a]; o.fxyz[3 * f + a] = c[a]; }
        int ow = o.owner[f];
        o.cvol[ow] += VSn / 3.0;
        for (int a = 0; a < 3; ++a) o.cxyz[3 * ow + a] += cc[a];
        if (f < o.nif) {
            int nb = o.nbr[f];
            o.cvol[nb] -= VSn / 3.0;
            for (int a = 0; a < 3; ++a) o.cxyz[3 * nb + a] -= cc[a];
        }
    }
    for (int i = 0; i < nc; ++i)
        for (int a = 0; a < 3; ++a) o.cxyz[3 * i + a] *= 1.0 / (24.0 * 2.0 * o.cvol[i]);

    // geometric.cpp:1206-1227 distCells; 1297-1497 face metrics
    o.fdist.assign(3 * nf, 0); o.fwC.assign(nf, 0); o.fmagPN.assign(nf, 0); o.falpha.assign(nf, 0);
    o.fePN.assign(3 * nf, 0); o.fskew.assign(3 * nf, 0); o.fvecPF.assign(3 * nf, 0); o.fvecNF.assign(3 * nf, 0);
    for (int f = 0; f < nf; ++f) {
        const double* fc = &o.fxyz[3 * f];
        const double* n = &o.fnrm[3 * f];
        const double* cP = &o.cxyz[3 * o.owner[f]];
        if (f < o.nif) {
            const double* cN = &o.cxyz[3 * o.nbr[f]];
            for (int a = 0; a < 3; ++a) o.fdist[3 * f + a] = cN[a] - cP[a];
            double dFP[3], dFN[3], dPN3[3];
            for (int a = 0; a < 3; ++a) { dFP[a] = fc[a] - cP[a]; dFN[a] = fc[a] - cN[a]; dPN3[a] = cN[a] - cP[a]; }
            double dPN = sqrt(dPN3[0] * dPN3[0] + dPN3[1] * dPN3[1] + dPN3[2] * dPN3[2]);
            for (int a = 0; a < 3; ++a) { o.fvecPF[3 * f + a] = dFP[a]; o.fvecNF[3 * f + a] = dFN[a]; }
            double dFP_of = std::abs(dFP[0] * n[0] + dFP[1] * n[1] + dFP[2] * n[2]);
            double dFN_of = std::abs(dFN[0] * n[0] + dFN[1] * n[1] + dFN[2] * n[2]);
            double wC = dFN_of / (dFP_of + dFN_of);
            wC = (float)(round(wC * 10000000) / 10000000);          // geometric.cpp:1351
            o.fwC[f] = wC;
            o.fmagPN[f] = dPN;
            for (int a = 0; a < 3; ++a) o.fePN[3 * f + a] = dPN3[a] / dPN;
            double al = 0.0;
            al += n[0] * o.fePN[3 * f]; al += n[1] * o.fePN[3 * f + 1]; al += n[2] * o.fePN[3 * f + 2];
            o.falpha[f] = 1.0 / std::abs(al);
            double D = -(n[0] * fc[0] + n[1] * fc[1] + n[2] * fc[2]);
            double u = (n[0] * cP[0] + n[1] * cP[1] + n[2] * cP[2] + D) /
                       (n[0] * (cP[0] - cN[0]) + n[1] * (cP[1] - cN[1]) + n[2] * (cP[2] - cN[2]));
            for (int a = 0; a < 3; ++a) {
                double s = cP[a] - u * (cP[a] - cN[a]);
                o.fskew[3 * f + a] = fc[a] - s;
            }
        } else {
            for (int a = 0; a < 3; ++a) o.fdist[3 * f + a] = fc[a] - cP[a];
            o.fwC[f] = (float)(round(0.5 * 10000000) / 10000000);   // geometric.cpp:1434-1436
            double dFP[3];
            for (int a = 0; a < 3; ++a) dFP[a] = fc[a] - cP[a];
            double dFP_len = sqrt(dFP[0] * dFP[0] + dFP[1] * dFP[1] + dFP[2] * dFP[2]);
            for (int a = 0; a < 3; ++a) { o.fvecPF[3 * f + a] = dFP[a]; o.fvecNF[3 * f + a] = dFP[a]; }
            double dFP_of = std::abs(dFP[0] * n[0] + dFP[1] * n[1] + dFP[2] * n[2]);
            o.fmagPN[f] = 2.0 * dFP_of;
            for (int a = 0; a < 3; ++a) o.fePN[3 * f + a] = dFP[a] / dFP_len;
            double al = 0.0;
            al += n[0] * o.fePN[3 * f]; al += n[1] * o.fePN[3 * f + 1]; al += n[2] * o.fePN[3 * f + 2];
            o.falpha[f] = 1.0 / std::abs(al);
        }
    }

    // vertex stencil (geometric.cpp:1786-1841)
    vector<vector<int>> pst(o.np);
    for (int i = 0; i < nc; ++i)
        for (int k = o.cptr[i]; k < o.cptr[i + 1]; ++k) pst[o.cpts[k]].push_back(i);
    o.sptr.assign(nc + 1, 0); o.sidx.clear();
    vector<int> tmp;
    for (int i = 0; i < nc; ++i) {
        tmp.clear();
        for (int k = o.cptr[i]; k < o.cptr[i + 1]; ++k)
            for (int c : pst[o.cpts[k]]) {
                if (c == i) continue;
                if (std::find(tmp.begin(), tmp.end(), c) == tmp.end()) tmp.push_back(c);
            }
        o.sidx.insert(o.sidx.end(), tmp.begin(), tmp.end());
        o.sptr[i + 1] = (int)o.sidx.size();
    }

    // initLeastSquare, cell-vertex stencil, 3x3 block (gradient.cpp:246-430)
    vector<double> vs(6 * (size_t)nc, 0.0);
    auto acc = [&](int i, double dx, double dy, double dz) {
        double wk = 1.0 / (dx * dx + dy * dy + dz * dz);
        double v[3] = {dx, dy, dz};
        double* s = &vs[6 * (size_t)i];
        s[0] += wk * v[0] * v[0]; s[1] += wk * v[0] * v[1]; s[2] += wk * v[0] * v[2];
        s[3] += wk * v[1] * v[1]; s[4] += wk * v[1] * v[2]; s[5] += wk * v[2] * v[2];
    };
    for (int i = 0; i < nc; ++i)
        for (int k = o.sptr[i]; k < o.sptr[i + 1]; ++k) {
            int j = o.sidx[k];
            acc(i, o.cxyz[3 * j] - o.cxyz[3 * i], o.cxyz[3 * j + 1] - o.cxyz[3 * i + 1],
                o.cxyz[3 * j + 2] - o.cxyz[3 * i + 2]);
        }
    for (int f = o.nif; f < nf; ++f) {
        int i = o.owner[f];
        acc(i, o.fxyz[3 * f] - o.cxyz[3 * i], o.fxyz[3 * f + 1] - o.cxyz[3 * i + 1],
            o.fxyz[3 * f + 2] - o.cxyz[3 * i + 2]);
    }
    o.lsq.assign(6 * (size_t)nc, 0.0);
    for (int i = 0; i < nc; ++i) {
        const double* s = &vs[6 * (size_t)i];
        // packed 9x9 indices 0,1,2,9,10,17 -> s0..s5
        double v0 = s[0], v1 = s[1], v2 = s[2], v9 = s[3], v10 = s[4], v17 = s[5];
        double det = v0 * v9 * v17 + v1 * v10 * v2 + v2 * v1 * v10 - v0 * v10 * v10
                   - v2 * v9 * v2 - v1 * v1 * v17;
        if (det == 0.0) fprintf(stderr, "detA=0 in LSQ\n");
        double* c = &o.lsq[6 * (size_t)i];
        c[0] = (v9 * v17 - v10 * v10) / det;
        c[1] = (v2 * v10 - v1 * v17) / det;
        c[2] = (v1 * v10 - v2 * v9) / det;
        c[3] = (v0 * v17 - v2 * v2) / det;
        c[4] = (v2 * v1 - v0 * v10) / det;
        c[5] = (v0 * v9 - v1 * v1) / det;
    }
}

// ------------------------------------------------------------------------
// calcLeastSquare(mesh,"cellVertex","1st","cell",cn,fn,...) (gradient.cpp:593-1159)
// phi: cell values; bR: face right value at boundary faces (varR[fn]).
void lsq_grad(const Oracle& o, const vector<Real>& phi, const vector<Real>& faceR,
              vector<Real>& grad) {
    int nc = o.nc;
    vector<Real> gt(3 * (size_t)nc, 0.0);
    for (int i = 0; i < nc; ++i)
        for (int k = o.sptr[i]; k < o.sptr[i + 1]; ++k) {
            int j = o.sidx[k];
            double dx = o.cxyz[3 * j] - o.cxyz[3 * i], dy = o.cxyz[3 * j + 1] - o.cxyz[3 * i + 1],
                   dz = o.cxyz[3 * j + 2] - o.cxyz[3 * i + 2];
            double wk = 1.0 / (dx * dx + dy * dy + dz * dz);
            Real d = phi[j] - phi[i];
            gt[3 * i] += wk * dx * d; gt[3 * i + 1] += wk * dy * d; gt[3 * i + 2] += wk * dz * d;
        }
    for (int f = o.nif; f < o.nf; ++f) {
        int i = o.owner[f];
        Real d = faceR[f] - phi[i];
        double dx = o.fxyz[3 * f] - o.cxyz[3 * i], dy = o.fxyz[3 * f + 1] - o.cxyz[3 * i + 1],
               dz = o.fxyz[3 * f + 2] - o.cxyz[3 * i + 2];
        double wk = 1.0 / (dx * dx + dy * dy + dz * dz);
        gt[3 * i] += wk * dx * d; gt[3 * i + 1] += wk * dy * d; gt[3 * i + 2] += wk * dz * d;
    }
    grad.assign(3 * (size_t)nc, 0.0);
    for (int i = 0; i < nc; ++i) {
        const double* c = &o.lsq[6 * (size_t)i];
        const Real* t = &gt[3 * (size_t)i];
        Real r0 = 0.0; r0 += c[0] * t[0]; r0 += c[1] * t[1]; r0 += c[2] * t[2];
        Real r1 = 0.0; r1 += c[1] * t[0]; r1 += c[3] * t[1]; r1 += c[4] * t[2];
        Real r2 = 0.0; r2 += c[2] * t[0]; r2 += c[4] * t[1]; r2 += c[5] * t[2];
        grad[3 * i] = r0; grad[3 * i + 1] = r1; grad[3 * i + 2] = r2;
    }
}

// setCellVarMinMax (src/solvers/cellVarMinMax.cpp:8-24)
void minmax(const Oracle& o, const vector<Real>& phi, vector<Real>& mx, vector<Real>& mn) {
    mx.assign(o.nc, 0.0); mn.assign(o.nc, 0.0);
    for (int i = 0; i < o.nc; ++i) {
        Real a = phi[i], b = phi[i];
        for (int k = o.sptr[i]; k < o.sptr[i + 1]; ++k) {
            a = std::max(a, phi[o.sidx[k]]); b = std::min(b, phi[o.sidx[k]]);
        }
        mx[i] = a; mn[i] = b;
    }
}

// calcIncomCellEOSVF (eos.cpp:10-151) + calcCellTransport (transport.cpp:9-98)
void cell_eos_transport(Oracle& o) {
    for (int i = 0; i < o.nc; ++i) {
        Real a0 = clamp01(o.VF[i]);
        Real v0 = clamp01(a0);
        Real v1 = clamp01(1.0 - a0);
        Real r = 0.0; r += v0 * o.rho[0]; r += v1 * o.rho[1];
        o.Rho[i] = r;
        Real m0 = clamp01(v0 * o.rho[0] / r);
        o.MF[i] = m0;
        o.VF[i] = v0;
    }
    for (int i = 0; i < o.nc; ++i) {
        Real v0 = o.VF[i], v1 = 1.0 - o.VF[i];
        Real m = 0.0; m += v0 * o.mu[0]; m += v1 * o.mu[1];
        o.Mu[i] = m + 0.0;
    }
}

// Speed of sound of the two-phase constant-density mixture, as calcCellEOSVF
// sets cell.var[C] at load (eos.cpp:160-293; the const branch gives every
// phase drho/dP = drho/dT = dh/dP = dh/dT = 1e8, eos.cpp:184-192).  The
// incompressible iteration never updates it.
void cell_sound_speed(Oracle& o) {
    o.Csnd.assign(o.nc, 0.0);
    for (int i = 0; i < o.nc; ++i) {
        Real a = clamp01(o.VF[i]);
        Real vf[2] = {clamp01(a), 0.0};
        vf[1] = clamp01(1.0 - a);
        Real rho = 0.0;
        for (int k = 0; k < 2; ++k) rho += vf[k] * o.rho[k];
        Real mf[2];
        mf[0] = clamp01(vf[0] * o.rho[0] / rho);
        mf[1] = clamp01(1.0 - mf[0]);
        Real drdp = 0.0, drdt = 0.0, dhdp = 0.0, dhdt = 0.0;
        for (int k = 0; k < 2; ++k) {
            drdp += vf[k] * 1.0e+8; drdt += vf[k] * 1.0e+8;
            dhdp += mf[k] * 1.0e+8; dhdt += mf[k] * 1.0e+8;
        }
        Real c2 = drdp + 1.0 / rho * drdt / dhdt * (1.0 - rho * dhdp);
        o.Csnd[i] = sqrt(1.0 / c2);
    }
}

// cell Mach number of the P "switch" boundary (reconIncom.cpp:620-625)
Real mach_of(const Oracle& o, int c) {
    return sqrt(pow(o.U[c], 2.0) + pow(o.V[c], 2.0) + pow(o.W[c], 2.0)) / o.Csnd[c];
}

// setIncomValuesLeftRightFace -> reconIncomZeroOrder + calcOtherDataIncomFromEOS
// (reconIncom.cpp:9-67, 440-496, 501-1035)
void recon_faces(Oracle& o) {
    int nf = o.nf;
    if ((int)o.Csnd.size() != o.nc) cell_sound_speed(o);   // before the first step
    auto sz = [&](vector<Real>& v) { v.assign(nf, 0.0); };
    sz(o.LP); sz(o.RP); sz(o.LU); sz(o.RU); sz(o.LV); sz(o.RV); sz(o.LW); sz(o.RW);
    sz(o.LVF); sz(o.RVF); sz(o.Lmu); sz(o.Rmu); sz(o.LRho); sz(o.RRho);
    for (int f = 0; f < o.nif; ++f) {
        int l = o.owner[f], r = o.nbr[f];
        o.LP[f] = o.P[l]; o.LU[f] = o.U[l]; o.LV[f] = o.V[l]; o.LW[f] = o.W[l];
        o.LVF[f] = o.VF[l]; o.Lmu[f] = o.Mu[l];
        o.RP[f] = o.P[r]; o.RU[f] = o.U[r]; o.RV[f] = o.V[r]; o.RW[f] = o.W[r];
        o.RVF[f] = o.VF[r]; o.Rmu[f] = o.Mu[r];
    }
    for (const Patch& pb : o.patches) {
        for (int f = pb.start; f < pb.start + pb.nfaces; ++f) {
            int c = o.owner[f];
            const double* n = &o.fnrm[3 * f];
            // pressure (reconIncom.cpp:600-632)
            if (pb.p_type == FIXED_VALUE) o.LP[f] = pb.p_val;
            else if (pb.p_type == SWITCH) o.LP[f] = mach_of(o, c) > 1.0 ? o.P[c] : Real(pb.p_val);
            else o.LP[f] = o.P[c];
            o.RP[f] = o.LP[f];
            // velocity (reconIncom.cpp:642-852)
            if (pb.u_type == FIXED_VALUE) {
                double uf[3] = {pb.u_val[0], pb.u_val[1], pb.u_val[2]};
                if (o.ufix_on) {
                    size_t nb = (size_t)(o.nf - o.nif), j = (size_t)(f - o.nif);
                    for (int a = 0; a < 3; ++a) uf[a] = o.ufix[a * nb + j];
                }
                o.LU[f] = uf[0]; o.LV[f] = uf[1]; o.LW[f] = uf[2];
                // live orifice profile (reconIncom.cpp:707-727)
                for (int k = 0; k < 24; ++k) {
                    double cen = 0.0006 * (double)k;
                    double fx = o.fxyz[3 * f];
                    if (fx > cen - 0.00016 && fx < cen + 0.00016) {
                        double xn = fx - cen, zn = o.fxyz[3 * f + 2] - 0.001;
                        double th = atan(zn / xn);
                        o.LU[f] = 0.0 * cos(th) - 0.1 * sin(th);
                        o.LV[f] = -0.1;
                        o.LW[f] = 0.0 * sin(th) + 0.1 * cos(th);
                    }
                }
                o.RU[f] = uf[0]; o.RV[f] = uf[1]; o.RW[f] = uf[2];
            } else if (pb.u_type == ZERO_GRADIENT) {
                o.LU[f] = o.U[c]; o.LV[f] = o.V[c]; o.LW[f] = o.W[c];
                o.RU[f] = o.LU[f]; o.RV[f] = o.LV[f]; o.RW[f] = o.LW[f];
            } else if (pb.u_type == SLIP) {
                Real u = o.U[c], v = o.V[c], w = o.W[c];
                Real nv = u * n[0] + v * n[1] + w * n[2];
                o.LU[f] = u - nv * n[0]; o.LV[f] = v - nv * n[1]; o.LW[f] = w - nv * n[2];
                o.RU[f] = o.LU[f]; o.RV[f] = o.LV[f]; o.RW[f] = o.LW[f];
            } else if (pb.u_type == NO_SLIP) {
                o.LU[f] = o.LV[f] = o.LW[f] = 0.0; o.RU[f] = o.RV[f] = o.RW[f] = 0.0;
            } else if (pb.u_type == SURFACE_NORMAL_FIXED_VALUE) {
                double nv = pb.u_val[0];
                o.LU[f] = nv * n[0]; o.LV[f] = nv * n[1]; o.LW[f] = nv * n[2];
                o.RU[f] = o.LU[f]; o.RV[f] = o.LV[f]; o.RW[f] = o.LW[f];
            } else if (pb.u_type == INLET_OUTLET) {
                double nv = pb.u_val[0];
                Real own = o.U[c] * n[0] + o.V[c] * n[1] + o.W[c] * n[2];
                if (own > 0.0) { o.LU[f] = o.U[c]; o.LV[f] = o.V[c]; o.LW[f] = o.W[c]; }
                else { o.LU[f] = nv * n[0]; o.LV[f] = nv * n[1]; o.LW[f] = nv * n[2]; }
                o.RU[f] = o.LU[f]; o.RV[f] = o.LV[f]; o.RW[f] = o.LW[f];
            }
            // volume fraction (reconIncom.cpp:882-1019)
            if (pb.vf_type == FIXED_VALUE) o.LVF[f] = pb.vf_val;
            else if (pb.vf_type == FUNCTION) o.LVF[f] = o.vffun[f - o.nif];
            else if (pb.vf_type == INLET_OUTLET) {
                Real own = o.U[c] * n[0] + o.V[c] * n[1] + o.W[c] * n[2];
                o.LVF[f] = own > 0.0 ? o.VF[c] : Real(pb.vf_val);
            } else o.LVF[f] = o.VF[c];
            o.RVF[f] = o.LVF[f];
            o.Lmu[f] = o.Mu[c]; o.Rmu[f] = o.Lmu[f];             // reconIncom.cpp:1024-1028
        }
    }
    // face density from the mixture EOS (reconIncom.cpp:445-493, eos.cpp:439-489)
    for (int f = 0; f < nf; ++f) {
        for (int side = 0; side < 2; ++side) {
            Real a = side == 0 ? o.LVF[f] : o.RVF[f];
            Real v0 = clamp01(a);
            Real v1 = clamp01(1.0 - a);
            Real r = 0.0; r += v0 * o.rho[0]; r += v1 * o.rho[1];
            (side == 0 ? o.LRho : o.RRho)[f] = r;
        }
    }
}

// calcSourceGravity (src/solvers/source/gravity.cpp:7-75)
void source_gravity(Oracle& o) {
    for (int k = 0; k < 4; ++k) o.srcG[k].assign(o.nc, 0.0);
    for (int f = 0; f < o.nf; ++f) {
        double wCL = o.fwC[f], wCR = 1.0 - wCL;
        const double* n = &o.fnrm[3 * f];
        Real UnL = o.LU[f] * n[0] + o.LV[f] * n[1] + o.LW[f] * n[2];
        Real UnR = o.RU[f] * n[0] + o.RV[f] * n[1] + o.RW[f] * n[2];
        Real UnF = wCL * UnL + wCR * UnR;
        Real RhoF = 1.0 / (wCL / o.LRho[f] + wCR / o.RRho[f]);
        double gL = 0.0, gR = 0.0;
        for (int a = 0; a < 3; ++a) { gL += o.g[a] * o.fvecPF[3 * f + a]; gR += o.g[a] * o.fvecNF[3 * f + a]; }
        int l = o.owner[f];
        for (int a = 0; a < 3; ++a) o.srcG[a][l] += ((RhoF * gL * n[a]) * o.farea[f] / o.cvol[l]);
        o.srcG[3][l] += ((RhoF * gL * UnF) * o.farea[f] / o.cvol[l]);
        if (f < o.nif) {
            int r = o.nbr[f];
            for (int a = 0; a < 3; ++a) o.srcG[a][r] -= ((RhoF * gR * n[a]) * o.farea[f] / o.cvol[r]);
            o.srcG[3][r] -= ((RhoF * gR * UnF) * o.farea[f] / o.cvol[r]);
        }
    }
}

// Rhie-Chow face velocity of the assembly face loops:
// eqMomentum.cpp:519-602, eqPressure.cpp:416-501,
// surfaceNormalVelocity.cpp:125-223 (Pressure, Gravity, SurfTens).
Real face_UnF(const Oracle& o, int f) {
    int l = o.owner[f], r = o.nbr[f];
    const double* n = &o.fnrm[3 * f];
    const double* e = &o.fePN[3 * f];
    const double* sk = &o.fskew[3 * f];
    double wCL = o.fwC[f], wCR = 1.0 - wCL;
    double alpha = o.falpha[f], dPN = o.fmagPN[f];
    Real ULs = o.U[l], VLs = o.V[l], WLs = o.W[l], URs = o.U[r], VRs = o.V[r], WRs = o.W[r];
    for (int a = 0; a < 3; ++a) {
        ULs += o.gU[3 * l + a] * sk[a]; URs += o.gU[3 * r + a] * sk[a];
        VLs += o.gV[3 * l + a] * sk[a]; VRs += o.gV[3 * r + a] * sk[a];
        WLs += o.gW[3 * l + a] * sk[a]; WRs += o.gW[3 * r + a] * sk[a];
    }
    ULs = std::max(o.minU[l], std::min(o.maxU[l], ULs)); URs = std::max(o.minU[r], std::min(o.maxU[r], URs));
    VLs = std::max(o.minV[l], std::min(o.maxV[l], VLs)); VRs = std::max(o.minV[r], std::min(o.maxV[r], VRs));
    WLs = std::max(o.minW[l], std::min(o.maxW[l], WLs)); WRs = std::max(o.minW[r], std::min(o.maxW[r], WRs));
    Real UnL = ULs * n[0] + VLs * n[1] + WLs * n[2];
    Real UnR = URs * n[0] + VRs * n[1] + WRs * n[2];
    Real UnF = wCL * UnL + wCR * UnR;
    Real RhoL = o.Rho[l], RhoR = o.Rho[r];
    Real Rs = 1.0 / (wCL / RhoL + wCR / RhoR);
    Real t1 = o.dt / Rs;
    // pressure
    UnF -= alpha * t1 * (o.P[r] - o.P[l]) / dPN;
    for (int a = 0; a < 3; ++a) {
        UnF -= t1 * Rs * wCL / RhoL * o.gP[3 * l + a] * (n[a] - alpha * e[a]);
        UnF -= t1 * Rs * wCR / RhoR * o.gP[3 * r + a] * (n[a] - alpha * e[a]);
    }
    for (int a = 0; a < 3; ++a) {
        UnF += t1 * Rs * wCL / RhoL * o.gP[3 * l + a] * n[a];
        UnF += t1 * Rs * wCR / RhoR * o.gP[3 * r + a] * n[a];
    }
    // gravity
    for (int a = 0; a < 3; ++a) UnF += alpha * t1 * Rs * o.g[a] * alpha * e[a];
    for (int a = 0; a < 3; ++a) {
        UnF += t1 * Rs * wCL / RhoL * o.srcG[a][l] * (n[a] - alpha * e[a]);
        UnF += t1 * Rs * wCR / RhoR * o.srcG[a][r] * (n[a] - alpha * e[a]);
    }
    for (int a = 0; a < 3; ++a) {
        UnF -= alpha * t1 * Rs * wCL / RhoL * o.srcG[a][l] * n[a];
        UnF -= alpha * t1 * Rs * wCR / RhoR * o.srcG[a][r] * n[a];
    }
    // surface tension
    UnF += alpha * t1 * o.sigma * (wCL * o.kappa[l] + wCR * o.kappa[r]) * (o.VF[r] - o.VF[l]) / dPN;
    for (int a = 0; a < 3; ++a) {
        UnF += t1 * Rs * wCL / RhoL * o.srcS[a][l] * (n[a] - alpha * e[a]);
        UnF += t1 * Rs * wCR / RhoR * o.srcS[a][r] * (n[a] - alpha * e[a]);
    }
    for (int a = 0; a < 3; ++a) {
        UnF -= alpha * t1 * Rs * wCL / RhoL * o.srcS[a][l] * n[a];
        UnF -= alpha * t1 * Rs * wCR / RhoR * o.srcS[a][r] * n[a];
    }
    return UnF;
}

void gradients5(Oracle& o) {
    // eqMomentum.cpp:95-104 / eqPressure.cpp:96-105 (the zeroing at :57-72 is
    // overwritten by the least-squares call)
    lsq_grad(o, o.P, o.RP, o.gP);
    lsq_grad(o, o.U, o.RU, o.gU);
    lsq_grad(o, o.V, o.RV, o.gV);
    lsq_grad(o, o.W, o.RW, o.gW);
    lsq_grad(o, o.VF, o.RVF, o.gVF);
}

// calcMomentumEqs assembly (eqMomentum.cpp:9-1008)
void assemble_momentum(Oracle& o) {
    int nc = o.nc;
    minmax(o, o.P, o.maxP, o.minP);
    minmax(o, o.U, o.maxU, o.minU);
    minmax(o, o.V, o.maxV, o.minV);
    minmax(o, o.W, o.maxW, o.minW);
    gradients5(o);
    o.Arows.assign(nc, 0); o.Acols.assign(nc, 0); o.Avals.assign(nc, 0.0);
    o.B0.assign(nc, 0.0); o.B1.assign(nc, 0.0); o.B2.assign(nc, 0.0);
    double dt = o.dt;
    for (int i = 0; i < nc; ++i) {               // eqMomentum.cpp:223-326
        Real R = o.Rho[i];
        double vol = o.cvol[i];
        o.Arows[i] = i; o.Acols[i] = i;
        o.Avals[i] = R * 1.5 * vol / dt;
        o.B0[i] = -R * (1.5 * o.U[i] - 2.0 * o.oldU[i] + 0.5 * o.Qm0[i]) * vol / dt;
        o.B1[i] = -R * (1.5 * o.V[i] - 2.0 * o.oldV[i] + 0.5 * o.Qm1[i]) * vol / dt;
        o.B2[i] = -R * (1.5 * o.W[i] - 2.0 * o.oldW[i] + 0.5 * o.Qm2[i]) * vol / dt;
        o.B0[i] += vol * o.srcG[0][i]; o.B1[i] += vol * o.srcG[1][i]; o.B2[i] += vol * o.srcG[2][i];
        o.B0[i] += vol * o.srcS[0][i]; o.B1[i] += vol * o.srcS[1][i]; o.B2[i] += vol * o.srcS[2][i];
        Real ap = o.alphaP[i] * vol;
        o.Avals[i] += ap;
        o.B0[i] -= ap * o.U[i]; o.B1[i] -= ap * o.V[i]; o.B2[i] -= ap * o.W[i];
    }
    Real* B[3] = {o.B0.data(), o.B1.data(), o.B2.data()};
    for (int f = 0; f < o.nif; ++f) {            // eqMomentum.cpp:331-787
        int l = o.owner[f], r = o.nbr[f];
        double area = o.farea[f];
        const double* n = &o.fnrm[3 * f];
        const double* e = &o.fePN[3 * f];
        const double* sk = &o.fskew[3 * f];
        double wCL = o.fwC[f], wCR = 1.0 - wCL;
        Real muF = 1.0 / (wCL / o.Lmu[f] + wCR / o.Rmu[f]);
        Real RhoL = o.Rho[l], RhoR = o.Rho[r];
        double dPN = o.fmagPN[f], alpha = o.falpha[f];
        Real UnF = face_UnF(o, f);
        double wL = (UnF > 0.0) ? 1.0 : 0.0, wR = 1.0 - wL;
        Real PL = o.P[l], PR = o.P[r];
        Real PF = wCL * PL + wCR * PR;
        Real UF = wL * o.LU[f] + wR * o.RU[f];
        Real VFv = wL * o.LV[f] + wR * o.RV[f];
        Real WF = wL * o.LW[f] + wR * o.RW[f];
        PF = wCL * PL + wCR * PR;
        for (int a = 0; a < 3; ++a) {
            PF += wCL * o.gP[3 * l + a] * sk[a];
            PF += wCR * o.gP[3 * r + a] * sk[a];
        }
        PF = std::max(o.minP[l], std::min(o.maxP[l], PF));
        PF = std::max(o.minP[r], std::min(o.maxP[r], PF));
        o.Avals[l] += (wL * RhoL * UnF * area + alpha * muF / dPN * area);
        Real v = (wR * RhoL * UnF * area - alpha * muF / dPN * area);
        if (!(fabs(v) < std::numeric_limits<double>::min())) {
            o.Arows.push_back(l); o.Acols.push_back(r); o.Avals.push_back(v);
        }
        o.Avals[r] -= (wR * RhoR * UnF * area - alpha * muF / dPN * area);
        v = -(wL * RhoR * UnF * area + alpha * muF / dPN * area);
        if (!(fabs(v) < std::numeric_limits<double>::min())) {
            o.Arows.push_back(r); o.Acols.push_back(l); o.Avals.push_back(v);
        }
        Real Fv[3] = {UF, VFv, WF};
        Real uL[3] = {o.U[l], o.V[l], o.W[l]}, uR[3] = {o.U[r], o.V[r], o.W[r]};
        const Real* gL[3] = {&o.gU[3 * l], &o.gV[3 * l], &o.gW[3 * l]};
        const Real* gR[3] = {&o.gU[3 * r], &o.gV[3 * r], &o.gW[3 * r]};
        for (int k = 0; k < 3; ++k) {
            Real cL = (RhoL * Fv[k] * UnF + PF * n[k]) * area;
            Real cR = (RhoR * Fv[k] * UnF + PF * n[k]) * area;
            Real df = alpha * muF * (uR[k] - uL[k]) / dPN * area;
            for (int a = 0; a < 3; ++a)
                df += (muF * (wCL * gL[k][a] + wCR * gR[k][a]) * (n[a] - alpha * e[a]) * area);
            Real dL = ((gL[k][0] * n[0] + gL[k][1] * n[1] + gL[k][2] * n[2]) * muF * area);
            Real dR = ((gR[k][0] * n[0] + gR[k][1] * n[1] + gR[k][2] * n[2]) * muF * area);
            B[k][l] -= (cL - df - dL);
            B[k][r] += (cR - df - dR);
        }
    }
    for (const Patch& pb : o.patches) {          // eqMomentum.cpp:795-1008
        for (int f = pb.start; f < pb.start + pb.nfaces; ++f) {
            int c = o.owner[f];
            double area = o.farea[f];
            const double* n = &o.fnrm[3 * f];
            const double* e = &o.fePN[3 * f];
            Real UnL = o.LU[f] * n[0] + o.LV[f] * n[1] + o.LW[f] * n[2];
            Real UnR = o.RU[f] * n[0] + o.RV[f] * n[1] + o.RW[f] * n[2];
            Real UnF = 0.5 * UnL + 0.5 * UnR;
            Real RhoL = o.LRho[f];
            Real t1 = 1.0 / RhoL * o.dt;
            Real Fv[3] = {0.5 * o.LU[f] + 0.5 * o.RU[f], 0.5 * o.LV[f] + 0.5 * o.RV[f],
                          0.5 * o.LW[f] + 0.5 * o.RW[f]};
            Real PF = 0.5 * o.LP[f] + 0.5 * o.RP[f];
            double cU = 0, cUd = 0;
            if (pb.u_type == FIXED_VALUE || pb.u_type == NO_SLIP || pb.u_type == SURFACE_NORMAL_FIXED_VALUE) { cU = 0; cUd = 1; }
            else if (pb.u_type == ZERO_GRADIENT) { cU = 1; cUd = 0; }
            else if (pb.u_type == SLIP) { cU = 0; cUd = 0; }
            else if (pb.u_type == INLET_OUTLET) {
                Real own = o.U[c] * n[0] + o.V[c] * n[1] + o.W[c] * n[2];
                cU = own > 0.0 ? 1.0 : 0.0; cUd = own > 0.0 ? 0.0 : 1.0;
            }
            Real muF = 0.5 * o.Lmu[f] + 0.5 * o.Rmu[f];
            Real org[3][2] = {{o.U[c], o.RU[f]}, {o.V[c], o.RV[f]}, {o.W[c], o.RW[f]}};
            double dPN = 0.5 * o.fmagPN[f], alpha = o.falpha[f];
            if (pb.p_type == FIXED_VALUE) {
                UnF -= alpha * t1 * (pb.p_val - o.P[c]) / dPN;
                for (int a = 0; a < 3; ++a) UnF += alpha * t1 * o.gP[3 * c + a] * e[a];
            }
            const Real* gc[3] = {&o.gU[3 * c], &o.gV[3 * c], &o.gW[3 * c]};
            Real dn[3];
            for (int k = 0; k < 3; ++k) {
                dn[k] = 0.0;
                dn[k] += alpha * (org[k][1] - org[k][0]) / dPN;
            }
            for (int a = 0; a < 3; ++a) {
                double tv = n[a] - alpha * e[a];
                for (int k = 0; k < 3; ++k) dn[k] += gc[k][a] * tv;
            }
            for (int a = 0; a < 3; ++a)
                for (int k = 0; k < 3; ++k) dn[k] += (gc[k][a] * n[a]);
            o.Avals[c] += cU * (RhoL * UnF * area);
            o.Avals[c] += cUd * (muF * alpha / dPN * area);
            for (int k = 0; k < 3; ++k)
                B[k][c] -= (RhoL * Fv[k] * UnF * area + PF * n[k] * area - muF * dn[k] * area);
        }
    }
}

// calcPressureEq assembly (eqPressure.cpp:8-673)
void assemble_pressure(Oracle& o) {
    int nc = o.nc;
    minmax(o, o.U, o.maxU, o.minU);
    minmax(o, o.V, o.maxV, o.minV);
    minmax(o, o.W, o.maxW, o.minW);
    gradients5(o);
    o.Arows.assign(nc, 0); o.Acols.assign(nc, 0); o.Avals.assign(nc, 0.0);
    o.B0.assign(nc, 0.0);
    for (int i = 0; i < nc; ++i) { o.Arows[i] = i; o.Acols[i] = i; }
    for (int f = 0; f < o.nif; ++f) {
        int l = o.owner[f], r = o.nbr[f];
        double area = o.farea[f];
        double wCL = o.fwC[f], wCR = 1.0 - wCL;
        Real RhoL = o.Rho[l], RhoR = o.Rho[r];
        Real Rs = 1.0 / (wCR / RhoL + wCR / RhoR);       // eqPressure.cpp:396
        Real t1 = o.dt / Rs;
        double dPN = o.fmagPN[f], alpha = o.falpha[f];
        Real UnF = face_UnF(o, f);
        o.Avals[l] += (+alpha * t1 / dPN * area);
        o.Arows.push_back(l); o.Acols.push_back(r); o.Avals.push_back((-alpha * t1 / dPN * area));
        o.Avals[r] -= (-alpha * t1 / dPN * area);
        o.Arows.push_back(r); o.Acols.push_back(l); o.Avals.push_back(-(+alpha * t1 / dPN * area));
        o.B0[l] -= (UnF * area);
        o.B0[r] += (UnF * area);
    }
    for (const Patch& pb : o.patches) {          // eqPressure.cpp:570-673
        for (int f = pb.start; f < pb.start + pb.nfaces; ++f) {
            int c = o.owner[f];
            double area = o.farea[f];
            const double* n = &o.fnrm[3 * f];
            const double* e = &o.fePN[3 * f];
            Real UnL = o.LU[f] * n[0] + o.LV[f] * n[1] + o.LW[f] * n[2];
            Real UnR = o.RU[f] * n[0] + o.RV[f] * n[1] + o.RW[f] * n[2];
            Real UnF = 0.5 * UnL + 0.5 * UnR;
            Real RhoL = o.LRho[f];
            Real t1 = 1.0 / RhoL * o.dt;
            double cPd = pb.p_type == FIXED_VALUE ? 1.0 : 0.0;
            if (pb.p_type == SWITCH) cPd = mach_of(o, c) > 1.0 ? 1.0 : 0.0;     // eqPressure.cpp:625-632
            double dPN = 0.5 * o.fmagPN[f], alpha = o.falpha[f];
            if (pb.p_type == FIXED_VALUE) {
                UnF -= alpha * t1 * (o.RP[f] - o.P[c]) / dPN;
                for (int a = 0; a < 3; ++a) UnF += alpha * t1 * o.gP[3 * c + a] * e[a];
            }
            o.Avals[c] += cPd * (alpha * t1 / dPN * area);
            o.B0[c] -= (UnF * area);
        }
    }
}

#ifndef ORC_AD   // two-phase stages: not differentiated (the adjoint is single phase)
// ------------------------------------------------------------------------
// setIncomValuesLeftRightFaceWithVfMSTACS (reconIncom.cpp:345-430) with
// calcMSTACS (NVD.cpp:1098-1862), one rank: zero-order face values, the LSQ
// gradient of VF, per-cell Courant numbers, then the high-order volume
// fraction on each side of every internal face (LVF/RVF of internal faces).
void mstacs_faces(Oracle& o) {
    recon_faces(o);
    vector<double> g;
    lsq_grad(o, o.VF, o.RVF, g);                                  // reconIncom.cpp:366-369
    const double tmpCFL = 1.0, eps = 1.e-12;
    o.co_cell.assign(o.nc, 0.0);
    for (int i = 0; i < o.nc; ++i) {                              // NVD.cpp:1114-1135
        double magVel = sqrt(o.U[i] * o.U[i] + o.V[i] * o.V[i] + o.W[i] * o.W[i]);
        double maxA = 0.0;
        for (int k = o.cfptr[i]; k < o.cfptr[i + 1]; ++k) maxA = std::max(maxA, o.farea[o.cfidx[k]]);
        o.co_cell[i] = 1.0 / tmpCFL * magVel * o.dt * maxA / o.cvol[i];
    }
    for (int f = 0; f < o.nif; ++f) {                             // NVD.cpp:1166-1858
        int l = o.owner[f], r = o.nbr[f];
        double wCL = o.fwC[f], wCR = 1.0 - wCL;
        double L1 = o.VF[l], R1 = o.VF[r];
        const double* gl = &g[3 * (size_t)l];
        const double* gr = &g[3 * (size_t)r];
        const double* pn = &o.fdist[3 * (size_t)f];              // vecPN of an internal face
        const double* n = &o.fnrm[3 * (size_t)f];
        const double* sk = &o.fskew[3 * (size_t)f];
        double coDDL = o.co_cell[l], coDDR = o.co_cell[r];
        double L2 = 0.0;
        for (int a = 0; a < 3; ++a) L2 += gl[a] * pn[a];
        L2 = R1 - 2.0 * L2;
        double R2 = 0.0;
        for (int a = 0; a < 3; ++a) R2 += gr[a] * pn[a];
        R2 = L1 + 2.0 * R2;
        L2 = std::max(0.0, std::min(1.0, L2));
        R2 = std::max(0.0, std::min(1.0, R2));
        double vfL = L1, vfR = R1;
        for (int LR = 0; LR < 2; ++LR) {
            if (LR == 0 && (vfL > 1.0 - eps || vfL < eps)) continue;
            if (LR == 1 && (vfR > 1.0 - eps || vfR < eps)) continue;
            double alpU, alpD, alpA;
            double coDD = std::max(coDDL, coDDR);
            if (LR == 0) { alpU = L2; alpD = L1; alpA = R1; }
            else { alpU = R2; alpD = R1; alpA = L1; }
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
            cosT = std::abs(cosT);
            double gamF = std::min(pow(cosT, 4.0), 1.0);
            double tCd = (alpD - alpU) / (std::abs(alpA - alpU) + 1.e-200) * (alpA > alpU ? 1.0 : -1.0);
            double side = LR == 0 ? 2.0 * wCR : 2.0 * wCL;
            auto skewed = [&](double gam) {                       // NVD.cpp:1628-1671
                double v = alpD + gam * (alpA - alpD);
                v += wCL * gl[0] * sk[0]; v += wCL * gl[1] * sk[1]; v += wCL * gl[2] * sk[2];
                v += wCR * gr[0] * sk[0]; v += wCR * gr[1] * sk[1]; v += wCR * gr[2] * sk[2];
                return std::max(0.0, std::min(1.0, v));
            };
            // compressive (CICSAM-like) branch (NVD.cpp:1602-1626)
            double gam = 0.0;
            if (tCd >= 0.0 && tCd < 1.0) {
                if (coDD > 0.0 && coDD <= 0.33) {
                    double tCf = std::min(1.0, tCd / coDD);
                    gam = (tCf - tCd) / (1.0 - tCd);
                } else if (coDD > 0.33 && coDD <= 1.0) {
                    double tCf = std::min(1.0, 3.0 * tCd);
                    gam = (tCf - tCd) / (1.0 - tCd);
                }
            }
            gam *= side;
            gam = std::max(0.0, std::min(1.0, gam));
            double vfComp = skewed(gam);
            // diffusive (STACS-like) branch (NVD.cpp:1741-1767)
            gam = 0.0;
            if (tCd >= 0.0 && tCd < 0.2) { double tCf = std::min(1.0, 3.0 * tCd); gam = (tCf - tCd) / (1.0 - tCd); }
            else if (tCd >= 0.2 && tCd < 0.5) { double tCf = 0.5 + 0.5 * tCd; gam = (tCf - tCd) / (1.0 - tCd); }
            else if (tCd >= 0.5 && tCd < 0.8333) { double tCf = 0.375 + 0.75 * tCd; gam = (tCf - tCd) / (1.0 - tCd); }
            else if (tCd >= 0.8333 && tCd < 1.0) { double tCf = 1.0; gam = (tCf - tCd) / (1.0 - tCd); }
            gam *= side;
            gam = std::max(0.0, std::min(1.0, gam));
            double vfDiff = skewed(gam);
            double vf = gamF * vfComp + (1.0 - gamF) * vfDiff;
            if (LR == 0) vfL = vf; else vfR = vf;
        }
        o.LVF[f] = vfL;
        o.RVF[f] = vfR;
    }
}

// calcLeastSquare(...,"input", cn = -1, ...): boundary faces contribute a zero
// difference (gradient.cpp:1018-1022)
void lsq_grad_input(const Oracle& o, const vector<double>& phi, vector<double>& grad) {
    vector<double> fr(o.nf, 0.0);
    for (int f = o.nif; f < o.nf; ++f) fr[f] = phi[o.owner[f]];
    lsq_grad(o, phi, fr, grad);
}

// Inverse of the symmetric 3x3 matrix in packed form (a00,a01,a02,a11,a12,a22)
// by cofactors, with the reference's determinant expression (curvature.cpp:
// 2964-2979).  The level-set stage of the reference forms the same (positive
// definite) matrices' pseudo-inverse with an SVD (math.svdcmp); for a
// positive definite matrix that is the inverse, equal to round-off.
void inv_sym3(const double* v, double* out) {
    double detA = v[0] * v[3] * v[5] + v[1] * v[4] * v[2] + v[2] * v[1] * v[4] - v[0] * v[4] * v[4] -
                  v[2] * v[3] * v[2] - v[1] * v[1] * v[5];
    out[0] = (v[3] * v[5] - v[4] * v[4]) / detA;
    out[1] = (v[2] * v[4] - v[1] * v[5]) / detA;
    out[2] = (v[1] * v[4] - v[2] * v[3]) / detA;
    out[3] = (v[0] * v[5] - v[2] * v[2]) / detA;
    out[4] = (v[2] * v[1] - v[0] * v[4]) / detA;
    out[5] = (v[0] * v[3] - v[1] * v[1]) / detA;
}

// calcCurvature (source/curvature.cpp:7-3537) as live: alpha - 1/2 is
// reinitialised towards a signed distance (4 pseudo-time steps, CFL 0.1) with
// a sign-weighted face least squares, then kappa = -div(grad LS / |grad LS|)
// with an unweighted stencil least squares; both smoothing passes have zero
// iterations in the reference (curvature.cpp:18-19).
void curvature(Oracle& o) {
    int nc = o.nc, nf = o.nf, nif = o.nif;
    const int LSReinitIterMax = 4;
    const double CFL_LS = 0.1, epsilon2 = 1.e-5;
    vector<double> a = o.VF;
    // interface cells: alpha - 1/2 changes sign between a cell and one of its
    // face values (curvature.cpp:40-117)
    vector<char> iface(nc, 0);
    for (int f = 0; f < nf; ++f) {
        int l = o.owner[f];
        double wCL = o.fwC[f], wCR = 1.0 - wCL;
        double fa = f < nif ? wCL * a[l] + wCR * a[o.nbr[f]] : a[l];
        if ((a[l] - 0.5) * (fa - 0.5) <= 0.0) iface[l] = 1;
        if (f < nif && (a[o.nbr[f]] - 0.5) * (fa - 0.5) <= 0.0) iface[o.nbr[f]] = 1;
    }
    for (int i = 0; i < nc; ++i) a[i] = a[i] - 0.5;
    vector<double> g;
    lsq_grad_input(o, a, g);
    for (int i = 0; i < nc; ++i) {
        if (!iface[i]) continue;
        double m = 0.0;
        m += pow(g[3 * i], 2.0); m += pow(g[3 * i + 1], 2.0); m += pow(g[3 * i + 2], 2.0);
        if (m > 1.0) a[i] = a[i] / m;
    }
    vector<double> tau(nc), vel0(nc);
    for (int i = 0; i < nc; ++i) {                     // curvature.cpp:140-177
        double maxA = 0.0;
        for (int k = o.cfptr[i]; k < o.cfptr[i + 1]; ++k) maxA = std::max(maxA, o.farea[o.cfidx[k]]);
        tau[i] = o.cvol[i] / maxA / 1.0;
        tau[i] = std::min(tau[i], pow(o.cvol[i], 0.3));
        double v = a[i];
        vel0[i] = v / sqrt(pow(v, 2.0) + pow(std::abs(v) * pow(o.cvol[i], 0.333), 2.0));
    }
    vector<double> nrm(3 * (size_t)nc), M(6 * (size_t)nc), rhs(3 * (size_t)nc);
    for (int iter = 0; iter < LSReinitIterMax; ++iter) {   // curvature.cpp:203-722
        lsq_grad_input(o, a, g);
        for (int i = 0; i < nc; ++i) {
            double mg = sqrt(pow(g[3 * i], 2.0) + pow(g[3 * i + 1], 2.0) + pow(g[3 * i + 2], 2.0));
            for (int k = 0; k < 3; ++k) nrm[3 * i + k] = mg != 0.0 ? g[3 * i + k] / mg : 0.0;
        }
        std::fill(M.begin(), M.end(), 0.0);
        std::fill(rhs.begin(), rhs.end(), 0.0);
        for (int f = 0; f < nf; ++f) {
            const double* d = &o.fdist[3 * (size_t)f];
            int l = o.owner[f];
            double aL = a[l], aR = f < nif ? a[o.nbr[f]] : aL;
            double DVar = aR - aL;
            for (int side = 0; side < (f < nif ? 2 : 1); ++side) {
                int c = side == 0 ? l : o.nbr[f];
                double sgn = vel0[c] > 0.0 ? 1.0 : -1.0;
                double wk = 0.0;
                if (side == 0) {
                    wk += sgn * nrm[3 * c] * (-d[0]); wk += sgn * nrm[3 * c + 1] * (-d[1]);
                    wk += sgn * nrm[3 * c + 2] * (-d[2]);
                } else {
                    wk += sgn * nrm[3 * c] * (d[0]); wk += sgn * nrm[3 * c + 1] * (d[1]);
                    wk += sgn * nrm[3 * c + 2] * (d[2]);
                }
                wk = pow(1.0 * wk, 1.0);
                wk = std::max(epsilon2, wk);
                double* m = &M[6 * (size_t)c];
                for (int ii = 0, num = 0; ii < 3; ++ii)
                    for (int jj = ii; jj < 3; ++jj) m[num++] += wk * d[ii] * d[jj];
                for (int ii = 0; ii < 3; ++ii) rhs[3 * (size_t)c + ii] += wk * d[ii] * DVar;
            }
        }
        for (int i = 0; i < nc; ++i) {
            double inv[6];
            inv_sym3(&M[6 * (size_t)i], inv);
            const double* t = &rhs[3 * (size_t)i];
            double g0 = inv[0] * t[0] + inv[1] * t[1] + inv[2] * t[2];
            double g1 = inv[1] * t[0] + inv[3] * t[1] + inv[4] * t[2];
            double g2 = inv[2] * t[0] + inv[4] * t[1] + inv[5] * t[2];
            if (!iface[i]) {
                double mg = sqrt(pow(g0, 2.0) + pow(g1, 2.0) + pow(g2, 2.0));
                double resi = CFL_LS * tau[i] * vel0[i] * (1.0 - mg);
                a[i] += resi;
            }
        }
    }
    o.LS = a;
    // normals and their divergence (curvature.cpp:725-2434)
    lsq_grad_input(o, a, g);
    for (int i = 0; i < nc; ++i) {
        double mg = 0.0;
        for (int k = 0; k < 3; ++k) mg += g[3 * i + k] * g[3 * i + k];
        mg = sqrt(mg);
        for (int k = 0; k < 3; ++k) nrm[3 * i + k] = mg > 1.e-15 ? g[3 * i + k] / mg : 0.0;
    }
    std::fill(M.begin(), M.end(), 0.0);
    std::fill(rhs.begin(), rhs.end(), 0.0);
    for (int i = 0; i < nc; ++i)
        for (int k = o.sptr[i]; k < o.sptr[i + 1]; ++k) {
            int j = o.sidx[k];
            double d[3] = {o.cxyz[3 * j] - o.cxyz[3 * i], o.cxyz[3 * j + 1] - o.cxyz[3 * i + 1],
                           o.cxyz[3 * j + 2] - o.cxyz[3 * i + 2]};
            double* m = &M[6 * (size_t)i];
            for (int ii = 0, num = 0; ii < 3; ++ii)
                for (int jj = ii; jj < 3; ++jj) m[num++] += 1.0 * d[ii] * d[jj];
        }
    for (int f = nif; f < nf; ++f) {
        int i = o.owner[f];
        double d[3] = {o.fxyz[3 * f] - o.cxyz[3 * i], o.fxyz[3 * f + 1] - o.cxyz[3 * i + 1],
                       o.fxyz[3 * f + 2] - o.cxyz[3 * i + 2]};
        double* m = &M[6 * (size_t)i];
        for (int ii = 0, num = 0; ii < 3; ++ii)
            for (int jj = ii; jj < 3; ++jj) m[num++] += 1.0 * d[ii] * d[jj];
    }
    for (int i = 0; i < nc; ++i)
        for (int k = o.sptr[i]; k < o.sptr[i + 1]; ++k) {
            int j = o.sidx[k];
            for (int ii = 0; ii < 3; ++ii) {
                double d = o.cxyz[3 * j + ii] - o.cxyz[3 * i + ii];
                double dv = nrm[3 * j + ii] - nrm[3 * i + ii];
                rhs[3 * (size_t)i + ii] += 1.0 * d * dv;
            }
        }
    o.kappa.assign(nc, 0.0);
    for (int i = 0; i < nc; ++i) {
        double inv[6];
        inv_sym3(&M[6 * (size_t)i], inv);
        const double* t = &rhs[3 * (size_t)i];
        double g0 = inv[0] * t[0] + inv[1] * t[1] + inv[2] * t[2];
        double g1 = inv[1] * t[0] + inv[3] * t[1] + inv[4] * t[2];
        double g2 = inv[2] * t[0] + inv[4] * t[1] + inv[5] * t[2];
        o.kappa[i] = -(g0 + g1 + g2);
    }
}

// calcGaussGreen(mesh, cn, fn, gradient) (gradient.cpp:1778-1985) of one cell
// field, one rank: internal faces in face order (skew-corrected linear face
// value), then the boundary patches by the field's type (sel: 0 = P, 1..3 =
// U, V, W, 4 = VF); faceR: varR[fn] per face; grad in: the old gradient,
// out: the new one.
void gauss_green(Oracle& o, int sel, const double* phi, const double* faceR, double* grad) {
    vector<double> gt(3 * (size_t)o.nc, 0.0);
    for (int f = 0; f < o.nif; ++f) {
        int l = o.owner[f], r = o.nbr[f];
        double wCL = o.fwC[f], wCR = 1.0 - wCL;
        const double* sk = &o.fskew[3 * (size_t)f];
        double v = wCL * phi[l] + wCR * phi[r];
        v += wCL * grad[3 * l] * sk[0]; v += wCL * grad[3 * l + 1] * sk[1]; v += wCL * grad[3 * l + 2] * sk[2];
        v += wCR * grad[3 * r] * sk[0]; v += wCR * grad[3 * r + 1] * sk[1]; v += wCR * grad[3 * r + 2] * sk[2];
        for (int j = 0; j < 3; ++j) {
            gt[3 * l + j] += v * o.fnrm[3 * f + j] * o.farea[f] / o.cvol[l];
            gt[3 * r + j] -= v * o.fnrm[3 * f + j] * o.farea[f] / o.cvol[r];
        }
    }
    for (const Patch& pb : o.patches)
        for (int f = pb.start; f < pb.start + pb.nfaces; ++f) {
            int c = o.owner[f];
            int type = sel == 0 ? pb.p_type : sel == 4 ? pb.vf_type : pb.u_type;
            double val = sel == 0 ? pb.p_val : sel == 4 ? pb.vf_val : pb.u_val[sel - 1];
            double v = phi[c];
            if (type == FIXED_VALUE) v = val;
            else if (type == ZERO_GRADIENT) {
                v += grad[3 * c] * (o.fxyz[3 * f] - o.cxyz[3 * c]);
                v += grad[3 * c + 1] * (o.fxyz[3 * f + 1] - o.cxyz[3 * c + 1]);
                v += grad[3 * c + 2] * (o.fxyz[3 * f + 2] - o.cxyz[3 * c + 2]);
            } else v = 0.5 * (phi[c] + faceR[f]);
            for (int j = 0; j < 3; ++j) gt[3 * c + j] += v * o.fnrm[3 * f + j] * o.farea[f] / o.cvol[c];
        }
    std::copy(gt.begin(), gt.end(), grad);
}

// SEMO_TurbModel_LES (turbulenceModels/LES.cpp): the subgrid viscosity muT
// and kinetic energy kSGS of a cell from the Green-Gauss gradients of U, V, W
// (two passes from zero, LES.cpp:46-51, with the faces' right values RU, RV,
// RW [nf]).  model 1: calcWALE (:29-131), 2: calcSmagorinsky (:141-220),
// 3: calcDynamicSmagorinsky (:224-467).  pow(x, 2.0) is written x * x (the
// compiler folds it so; it is exact either way).  Kept as written: WALE's
// SSd sums the squares of S's diagonal with Sd's off-diagonals (:110-112),
// and the dynamic model's running sums sumV and sumU collect U and V
// respectively (:330-331, :415-416) while L uses sumU where U is meant.
void les(Oracle& o, int model, const double* U, const double* V, const double* W, const double* Rho,
         const double* RU, const double* RV, const double* RW, double* muT, double* kSGS) {
    int n = o.nc;
    vector<double> gu(3 * (size_t)n, 0.0), gv(3 * (size_t)n, 0.0), gw(3 * (size_t)n, 0.0);
    gauss_green(o, 1, U, RU, gu.data()); gauss_green(o, 1, U, RU, gu.data());
    gauss_green(o, 2, V, RV, gv.data()); gauss_green(o, 2, V, RV, gv.data());
    gauss_green(o, 3, W, RW, gw.data()); gauss_green(o, 3, W, RW, gw.data());
    // the strain-rate tensor S of a cell (LES.cpp:80-88) and S_ij S_ij (:106-108)
    auto strain = [&](int i, double S[3][3]) {
        const double* a = &gu[3 * (size_t)i]; const double* b = &gv[3 * (size_t)i];
        const double* c = &gw[3 * (size_t)i];
        S[0][0] = a[0]; S[0][1] = 0.5 * (a[1] + b[0]); S[0][2] = 0.5 * (a[2] + c[0]);
        S[1][0] = S[0][1]; S[1][1] = b[1]; S[1][2] = 0.5 * (b[2] + c[1]);
        S[2][0] = S[0][2]; S[2][1] = S[1][2]; S[2][2] = c[2];
        return S[0][0] * S[0][0] + S[1][1] * S[1][1] + S[2][2] * S[2][2] +
               2.0 * (S[0][1] * S[0][1] + S[0][2] * S[0][2] + S[1][2] * S[1][2]);
    };
    for (int i = 0; i < n; ++i) {
        double S[3][3];
        double SS = strain(i, S);
        double Sabs = std::sqrt(2.0 * SS);
        double vol = o.cvol[i];
        if (model == 1) {
            const double Ci = 0.007, Cw = 0.325;
            const double* a = &gu[3 * (size_t)i]; const double* b = &gv[3 * (size_t)i];
            const double* c = &gw[3 * (size_t)i];
            double Sd01 = 0.5 * (a[1] * a[1] + b[0] * b[0]);
            double Sd02 = 0.5 * (a[2] * a[2] + c[0] * c[0]);
            double Sd12 = 0.5 * (b[2] * b[2] + c[1] * c[1]);
            double SSd = S[0][0] * S[0][0] + S[1][1] * S[1][1] + S[2][2] * S[2][2] +
                         2.0 * (Sd01 * Sd01 + Sd02 * Sd02 + Sd12 * Sd12);
            double Ls = Cw * std::pow(vol, 0.333333333);
            double Ls2 = Ls * Ls;
            double t = std::pow(SS, 2.5) + std::pow(SSd, 1.25);
            muT[i] = 0.0;
            if (t != 0.0) muT[i] = Rho[i] * Ls2 * std::pow(SSd, 1.5) / t;
            kSGS[i] = Ci * std::pow(vol, 0.6666666666) * (0.5 * Sabs * Sabs);
        } else if (model == 2) {
            const double C_I = 0.007, C_R = 0.1;
            double D2 = std::pow(vol, 0.666666666);
            muT[i] = C_R * Rho[i] * D2 * Sabs;
            kSGS[i] = C_I * D2 * (0.5 * Sabs * Sabs);
        } else {
            const double C_I = 0.007;
            double L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, M[3][3];
            double r = Rho[i], u = U[i], v = V[i], w = W[i];
            L[0][0] = r * u * u * vol; L[0][1] = r * u * v * vol; L[0][2] = r * u * w * vol;
            L[1][1] = r * v * v * vol; L[1][2] = r * v * w * vol; L[2][2] = r * w * w * vol;
            double sumVol = vol, sumRho = r * vol, sumV = u * vol, sumU = v * vol, sumW = w * vol;
            double sumSabs = Sabs * vol;
            double t062 = std::pow(vol, 0.66666666);
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    M[j][k] = -t062 * Sabs * S[j][k] * vol;
                    S[j][k] *= vol;
                }
            for (int q = o.sptr[i]; q < o.sptr[i + 1]; ++q) {
                int j = o.sidx[q];
                double T[3][3];
                double SSs = strain(j, T);
                double vs = o.cvol[j];
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) S[a][b] += T[a][b] * vs;
                double Sabss = std::sqrt(2.0 * SSs);
                double t06 = std::pow(vs, 0.66666666);
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) M[a][b] -= t06 * Sabss * T[a][b] * vs;
                sumSabs += Sabss * vs;
                double rs = Rho[j], us = U[j], vv = V[j], ws = W[j];
                L[0][0] += rs * us * us * vs; L[0][1] += rs * us * vv * vs; L[0][2] += rs * us * ws * vs;
                L[1][1] += rs * vv * vv * vs; L[1][2] += rs * vv * ws * vs; L[2][2] += rs * ws * ws * vs;
                sumRho += rs * vs; sumV += us * vs; sumU += vv * vs; sumW += ws * vs; sumVol += vs;
            }
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) { M[j][k] /= sumVol; S[j][k] /= sumVol; L[j][k] /= sumVol; }
            sumRho /= sumVol; sumV /= sumVol; sumU /= sumVol; sumW /= sumVol; sumSabs /= sumVol;
            L[0][0] -= sumRho * sumU * sumU; L[0][1] -= sumRho * sumU * sumV; L[0][2] -= sumRho * sumU * sumW;
            L[1][1] -= sumRho * sumV * sumV; L[1][2] -= sumRho * sumV * sumW; L[2][2] -= sumRho * sumW * sumW;
            L[1][0] = L[0][1]; L[2][0] = L[0][2]; L[2][1] = L[1][2];
            double sumDelta = std::pow(sumVol, 0.66666666666);
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) M[j][k] = sumDelta * sumSabs * S[j][k] - M[j][k];
            double s1 = 0.0, s2 = 0.0;
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) { s1 += L[j][k] * M[j][k]; s2 += M[j][k] * M[j][k]; }
            s1 /= 9.0; s2 /= 9.0;
            double C_R = 0.0;
            if (s2 != 0.0) C_R = -0.5 * s1 / s2;
            double D2 = std::pow(vol, 0.666666666);
            muT[i] = C_R * r * D2 * Sabs;
            kSGS[i] = C_I * D2 * (0.5 * Sabs * Sabs);
        }
    }
}

// the turbulence part of calcCellTransport (transport.cpp:38-97) after the
// mixture viscosity of cell_eos_transport: muT of the LES model, mu += muT.
// first: the step's first call (pressureBased.cpp:17-19), when SEMO's faces
// still hold the zero right values they were loaded with; else the right
// values of the last reconstruction (after a VOF iteration, :115-116).
void cell_les(Oracle& o, bool first) {
    if (!o.les_model) return;
    vector<double> z;
    const double *RU, *RV, *RW;
    if (first || o.RU.empty()) { z.assign(o.nf, 0.0); RU = RV = RW = z.data(); }
    else { RU = o.RU.data(); RV = o.RV.data(); RW = o.RW.data(); }
    o.muT.assign(o.nc, 0.0); o.kSGS.assign(o.nc, 0.0);
    les(o, o.les_model, o.U.data(), o.V.data(), o.W.data(), o.Rho.data(), RU, RV, RW, o.muT.data(), o.kSGS.data());
    for (int i = 0; i < o.nc; ++i) o.Mu[i] += o.muT[i];
}

// calcSourceSurfaceTension (source/surfaceTension.cpp:7-229): curvature, then
// sigma kappa grad(alpha) with grad(alpha) = one Green-Gauss pass over the LSQ
// gradient.  With sigma = 0 every use of kappa vanishes and the pipeline is
// skipped (sources zero).
void source_surface_tension(Oracle& o) {
    for (int k = 0; k < 4; ++k) o.srcS[k].assign(o.nc, 0.0);
    if (o.sigma == 0.0) { o.kappa.assign(o.nc, 0.0); return; }
    curvature(o);
    vector<double> g;
    lsq_grad(o, o.VF, o.RVF, g);
    gauss_green(o, 4, o.VF.data(), o.RVF.data(), g.data());
    for (int i = 0; i < o.nc; ++i) {
        double s = o.sigma, kp = o.kappa[i];
        o.srcS[0][i] = (s * kp * g[3 * i]);
        o.srcS[1][i] = (s * kp * g[3 * i + 1]);
        o.srcS[2][i] = (s * kp * g[3 * i + 2]);
        o.srcS[3][i] = s * kp * (g[3 * i] * o.U[i] + g[3 * i + 1] * o.V[i] + g[3 * i + 2] * o.W[i]);
    }
}

// calcVolfracEq assembly (eqVolfrac.cpp:9-765)
void assemble_vof(Oracle& o) {
    int nc = o.nc;
    minmax(o, o.U, o.maxU, o.minU);
    minmax(o, o.V, o.maxV, o.minV);
    minmax(o, o.W, o.maxW, o.minW);
    gradients5(o);                                                // eqVolfrac.cpp:90-99
    o.Arows.assign(nc, 0); o.Acols.assign(nc, 0); o.Avals.assign(nc, 0.0);
    o.B0.assign(nc, 0.0);
    for (int i = 0; i < nc; ++i) {                                // eqVolfrac.cpp:221-262
        double vol = o.cvol[i];
        o.Arows[i] = i; o.Acols[i] = i;
        o.Avals[i] = 1.5 * vol / o.dt;
        o.B0[i] = -(1.5 * o.VF[i] - 2.0 * o.oldVF[i] + 0.5 * o.Qm3[i]) * vol / o.dt;
    }
    for (int f = 0; f < o.nif; ++f) {                             // eqVolfrac.cpp:266-607
        int l = o.owner[f], r = o.nbr[f];
        double area = o.farea[f];
        double UnF = face_UnF(o, f);
        double wL = (UnF > 0.0) ? 1.0 : 0.0, wR = 1.0 - wL;
        double VFF = wL * o.LVF[f] + wR * o.RVF[f];
        o.Avals[l] += (wL * UnF * area);
        double v = (wR * UnF * area);
        if (!(std::abs(v) < std::numeric_limits<double>::min())) {
            o.Arows.push_back(l); o.Acols.push_back(r); o.Avals.push_back(v);
        }
        o.Avals[r] -= (wR * UnF * area);
        v = -(wL * UnF * area);
        if (!(std::abs(v) < std::numeric_limits<double>::min())) {
            o.Arows.push_back(r); o.Acols.push_back(l); o.Avals.push_back(v);
        }
        o.B0[l] -= (VFF * UnF * area);
        o.B0[r] += (VFF * UnF * area);
    }
    for (const Patch& pb : o.patches) {                          // eqVolfrac.cpp:611-703
        for (int f = pb.start; f < pb.start + pb.nfaces; ++f) {
            int c = o.owner[f];
            double area = o.farea[f];
            const double* n = &o.fnrm[3 * f];
            const double* e = &o.fePN[3 * f];
            double UnL = o.LU[f] * n[0] + o.LV[f] * n[1] + o.LW[f] * n[2];
            double UnR = o.RU[f] * n[0] + o.RV[f] * n[1] + o.RW[f] * n[2];
            double UnF = 0.5 * UnL + 0.5 * UnR;
            double t1 = 1.0 / o.LRho[f] * o.dt;
            double VFF = 0.5 * o.LVF[f] + 0.5 * o.RVF[f];
            double cVF = 0.0;
            if (pb.vf_type == ZERO_GRADIENT) cVF = 1.0;
            else if (pb.vf_type == INLET_OUTLET) {
                double own = o.U[c] * n[0] + o.V[c] * n[1] + o.W[c] * n[2];
                cVF = (own > 0.0) ? 1.0 : 0.0;
            }
            double dPN = 0.5 * o.fmagPN[f], alpha = o.falpha[f];
            if (pb.p_type == FIXED_VALUE) {
                UnF -= alpha * t1 * (pb.p_val - o.P[c]) / dPN;
                for (int a = 0; a < 3; ++a) UnF += alpha * t1 * o.gP[3 * c + a] * e[a];
            }
            o.Avals[c] += cVF * (UnF * area);
            o.B0[c] -= (VFF * UnF * area);
        }
    }
}
#endif  // ORC_AD

// ------------------------------------------------------------------------
// linear solve: COO -> CSR (solveAMGCL.cpp:206-251), ILU0 + BiCGStab
struct Csr { int n; vector<int> ptr, col; vector<double> val; };

Csr coo_to_csr(int n, const vector<int>& r, const vector<int>& c, const vector<double>& v) {
    Csr A; A.n = n; A.ptr.assign(n + 1, 0);
    size_t nnz = r.size();
    for (size_t k = 0; k < nnz; ++k) A.ptr[r[k]]++;
    for (int i = 0, cs = 0; i < n; ++i) { int t = A.ptr[i]; A.ptr[i] = cs; cs += t; }
    A.ptr[n] = (int)nnz;
    A.col.assign(nnz, 0); A.val.assign(nnz, 0.0);
    for (size_t k = 0; k < nnz; ++k) { int d = A.ptr[r[k]]++; A.col[d] = c[k]; A.val[d] = v[k]; }
    for (int i = 0, last = 0; i <= n; ++i) { int t = A.ptr[i]; A.ptr[i] = last; last = t; }
    // sort each row by column (AMGCL sorts rows too, mpi/amg.hpp:225) and merge duplicates
    Csr S; S.n = n; S.ptr.assign(n + 1, 0);
    vector<std::pair<int, double>> row;
    for (int i = 0; i < n; ++i) {
        row.clear();
        for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) row.push_back({A.col[k], A.val[k]});
        std::stable_sort(row.begin(), row.end(), [](auto& a, auto& b) { return a.first < b.first; });
        for (size_t k = 0; k < row.size(); ++k) {
            if (k > 0 && row[k].first == row[k - 1].first) { S.val.back() += row[k].second; continue; }
            S.col.push_back(row[k].first); S.val.push_back(row[k].second);
        }
        S.ptr[i + 1] = (int)S.col.size();
    }
    return S;
}

void spmv(const Csr& A, const double* x, double* y) {
    for (int i = 0; i < A.n; ++i) {
        double s = 0.0;
        for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) s += A.val[k] * x[A.col[k]];
        y[i] = s;
    }
}

struct Ilu0 {
    const Csr* A; vector<double> lu; vector<int> diag;
    void factor(const Csr& M) {
        A = &M; lu = M.val; diag.assign(M.n, -1);
        for (int i = 0; i < M.n; ++i)
            for (int k = M.ptr[i]; k < M.ptr[i + 1]; ++k) if (M.col[k] == i) diag[i] = k;
        vector<int> pos(M.n, -1);
        for (int i = 0; i < M.n; ++i) {
            for (int k = M.ptr[i]; k < M.ptr[i + 1]; ++k) pos[M.col[k]] = k;
            for (int k = M.ptr[i]; k < M.ptr[i + 1]; ++k) {
                int j = M.col[k];
                if (j >= i) break;
                lu[k] /= lu[diag[j]];
                for (int kk = diag[j] + 1; kk < M.ptr[j + 1]; ++kk) {
                    int p = pos[M.col[kk]];
                    if (p >= 0) lu[p] -= lu[k] * lu[kk];
                }
            }
            for (int k = M.ptr[i]; k < M.ptr[i + 1]; ++k) pos[M.col[k]] = -1;
        }
    }
    void apply(const double* r, double* z) const {
        const Csr& M = *A;
        for (int i = 0; i < M.n; ++i) {
            double s = r[i];
            for (int k = M.ptr[i]; k < diag[i]; ++k) s -= lu[k] * z[M.col[k]];
            z[i] = s;
        }
        for (int i = M.n - 1; i >= 0; --i) {
            double s = z[i];
            for (int k = diag[i] + 1; k < M.ptr[i + 1]; ++k) s -= lu[k] * z[M.col[k]];
            z[i] = s / lu[diag[i]];
        }
    }
};

double dot(const vector<double>& a, const vector<double>& b) {
    double s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}

// right-preconditioned BiCGStab, x0 = 0, stop at ||b-Ax|| <= tol ||b||
int bicgstab(const Csr& A, const Ilu0& M, const vector<double>& b, vector<double>& x,
             double tol, int maxit, double& relres) {
    int n = A.n;
    x.assign(n, 0.0);
    double nb = sqrt(dot(b, b));
    if (nb == 0.0) { relres = 0.0; return 0; }
    vector<double> r = b, rh = b, p(n, 0), v(n, 0), s(n), t(n), ph(n), sh(n);
    double rho_old = 1, al = 1, om = 1;
    int it;
    for (it = 1; it <= maxit; ++it) {
        double rho_n = dot(rh, r);
        if (rho_n == 0.0) break;
        double be = (rho_n / rho_old) * (al / om);
        for (int i = 0; i < n; ++i) p[i] = r[i] + be * (p[i] - om * v[i]);
        M.apply(p.data(), ph.data());
        spmv(A, ph.data(), v.data());
        al = rho_n / dot(rh, v);
        for (int i = 0; i < n; ++i) s[i] = r[i] - al * v[i];
        if (sqrt(dot(s, s)) <= tol * nb) {
            for (int i = 0; i < n; ++i) x[i] += al * ph[i];
            r = s; break;
        }
        M.apply(s.data(), sh.data());
        spmv(A, sh.data(), t.data());
        om = dot(t, s) / dot(t, t);
        for (int i = 0; i < n; ++i) { x[i] += al * ph[i] + om * sh[i]; r[i] = s[i] - om * t[i]; }
        rho_old = rho_n;
        if (sqrt(dot(r, r)) <= tol * nb) break;
    }
    // true residual
    vector<double> ax(n);
    spmv(A, x.data(), ax.data());
    for (int i = 0; i < n; ++i) ax[i] = b[i] - ax[i];
    relres = sqrt(dot(ax, ax)) / nb;
    return it;
}

vector<double> values(const vector<Real>& v) {
    vector<double> d(v.size());
    for (size_t k = 0; k < v.size(); ++k) d[k] = val(v[k]);
    return d;
}

void solve_last(Oracle& o, int nrhs) {
    Csr A = coo_to_csr(o.nc, o.Arows, o.Acols, values(o.Avals));
    Ilu0 M; M.factor(A);
    vector<Real>* B[3] = {&o.B0, &o.B1, &o.B2};
    vector<Real>* X[3] = {&o.X0, &o.X1, &o.X2};
    vector<double> x[3];
    for (int k = 0; k < nrhs; ++k) {
        double rr; o.last_iters = bicgstab(A, M, values(*B[k]), x[k], o.tol, o.maxit, rr);
        o.last_res = std::max(k == 0 ? 0.0 : o.last_res, rr);
    }
#ifndef ORC_AD
    for (int k = 0; k < nrhs; ++k) X[k]->swap(x[k]);
#else
    // x = A^-1 b as one step of the tape: the reverse step solves
    // A^T lam = x_bar, then b_bar += lam and a_bar(k) -= lam[row k] x[col k]
    // for every COO entry k (the implicit-function rule of an exact solve)
    ad::Tape& T = ad::tape();
    int n = o.nc;
    for (int k = 0; k < nrhs; ++k) {
        X[k]->resize(n);
        for (int i = 0; i < n; ++i) (*X[k])[i] = ad::adouble::var(x[k][i]);
    }
    if (!T.on) return;
    struct SolveRec {
        int n, nrhs; double tol; int maxit;
        vector<int> rows, cols, aidx, bidx, xidx;
        vector<double> avals, x;
    };
    auto rec = std::make_shared<SolveRec>();
    rec->n = n; rec->nrhs = nrhs; rec->tol = o.tol; rec->maxit = o.maxit;
    rec->rows = o.Arows; rec->cols = o.Acols; rec->avals = values(o.Avals);
    rec->aidx.resize(o.Avals.size());
    for (size_t q = 0; q < o.Avals.size(); ++q) rec->aidx[q] = o.Avals[q].i;
    for (int k = 0; k < nrhs; ++k)
        for (int i = 0; i < n; ++i) {
            rec->bidx.push_back((*B[k])[i].i); rec->xidx.push_back((*X[k])[i].i); rec->x.push_back(x[k][i]);
        }
    T.push_custom([rec](vector<double>& adj) {
        int n = rec->n;
        Csr At = coo_to_csr(n, rec->cols, rec->rows, rec->avals);
        Ilu0 Mt; Mt.factor(At);
        for (int k = 0; k < rec->nrhs; ++k) {
            vector<double> xb(n), lam;
            for (int i = 0; i < n; ++i) xb[i] = adj[rec->xidx[(size_t)k * n + i]];
            double rr;
            bicgstab(At, Mt, xb, lam, rec->tol, rec->maxit, rr);
            for (int i = 0; i < n; ++i) {
                int b = rec->bidx[(size_t)k * n + i];
                if (b >= 0) adj[b] += lam[i];
            }
            const double* xk = &rec->x[(size_t)k * n];
            for (size_t q = 0; q < rec->rows.size(); ++q)
                if (rec->aidx[q] >= 0) adj[rec->aidx[q]] -= lam[rec->rows[q]] * xk[rec->cols[q]];
        }
    });
#endif
}

#ifndef ORC_AD
// VF += vofVofURF dVF, clipped to [0,1]; returns sqrt(sum dVF^2)
// (eqVolfrac.cpp:770-783)
double vof_update(Oracle& o) {
    double nd = 0.0;
    for (int i = 0; i < o.nc; ++i) {
        o.VF[i] += o.urf_vf * o.X0[i];
        o.VF[i] = std::max(0.0, std::min(1.0, o.VF[i]));
        nd += pow(o.X0[i], 2.0);
    }
    return sqrt(nd);
}

#endif  // ORC_AD

// momentum update (eqMomentum.cpp:1029-1119)
double momentum_update(Oracle& o) {
    double nd = 0, np = 0;
    for (int i = 0; i < o.nc; ++i) {
        o.U[i] += o.urf_u * o.X0[i]; o.V[i] += o.urf_u * o.X1[i]; o.W[i] += o.urf_u * o.X2[i];
        Real* q[3] = {&o.U[i], &o.V[i], &o.W[i]};
        for (int k = 0; k < 3; ++k) {
            if (*q[k] <= -o.lim) *q[k] = -o.lim;
            if (*q[k] >= o.lim) *q[k] = o.lim;
        }
        nd += pow(val(o.X0[i]), 2.0); nd += pow(val(o.X1[i]), 2.0); nd += pow(val(o.X2[i]), 2.0);
        np += pow(val(o.U[i]), 2.0); np += pow(val(o.V[i]), 2.0); np += pow(val(o.W[i]), 2.0);
    }
    return sqrt(nd) / sqrt(np);
}

// p' Green-Gauss gradient x2 and the correction (eqPressure.cpp:686-883)
double pressure_correct(Oracle& o) {
    int nc = o.nc;
    const vector<Real>& pp = o.X0;
    vector<Real> gp(3 * (size_t)nc, 0.0), tmp(3 * (size_t)nc);
    for (int pass = 0; pass < 2; ++pass) {
        std::fill(tmp.begin(), tmp.end(), Real(0.0));
        for (int f = 0; f < o.nif; ++f) {
            int l = o.owner[f], r = o.nbr[f];
            double wCL = o.fwC[f], wCR = 1.0 - wCL;
            Real vF = wCL * pp[l] + wCR * pp[r];
            const double* sk = &o.fskew[3 * f];
            for (int a = 0; a < 3; ++a) {
                vF += wCL * gp[3 * l + a] * sk[a];
                vF += wCR * gp[3 * r + a] * sk[a];
            }
            const double* n = &o.fnrm[3 * f];
            for (int a = 0; a < 3; ++a) tmp[3 * l + a] += vF * n[a] * o.farea[f] / o.cvol[l];
            for (int a = 0; a < 3; ++a) tmp[3 * r + a] -= vF * n[a] * o.farea[f] / o.cvol[r];
        }
        for (const Patch& pb : o.patches)
            for (int f = pb.start; f < pb.start + pb.nfaces; ++f) {
                int c = o.owner[f];
                Real vF = pp[c];
                if (pb.p_type == FIXED_VALUE) vF = 0.0;
                const double* n = &o.fnrm[3 * f];
                for (int a = 0; a < 3; ++a) tmp[3 * c + a] += vF * n[a] * o.farea[f] / o.cvol[c];
            }
        gp = tmp;
    }
    double nd = 0, np = 0;
    for (int i = 0; i < nc; ++i) {
        o.P[i] += o.urf_p * pp[i];
        Real t1 = o.dt / o.Rho[i];
        Real rU = o.urf_pu * t1 * gp[3 * i], rV = o.urf_pu * t1 * gp[3 * i + 1],
             rW = o.urf_pu * t1 * gp[3 * i + 2];
        o.U[i] -= rU; o.V[i] -= rV; o.W[i] -= rW;
        Real* q[4] = {&o.P[i], &o.U[i], &o.V[i], &o.W[i]};
        for (int k = 0; k < 4; ++k) {
            if (*q[k] <= -o.lim) *q[k] = -o.lim;
            if (*q[k] >= o.lim) *q[k] = o.lim;
        }
        nd += pow(val(pp[i]), 2.0); nd += pow(val(rU), 2.0); nd += pow(val(rV), 2.0); nd += pow(val(rW), 2.0);
        np += pow(val(o.P[i]), 2.0); np += pow(val(o.U[i]), 2.0); np += pow(val(o.V[i]), 2.0);
        np += pow(val(o.W[i]), 2.0);
    }
    return sqrt(nd) / sqrt(np);
}

#ifndef ORC_AD
// calcTimeStepFromCorantNumber (adjust != 0, timestep.cpp:260-330) or
// calcCorantNumberForPrint (timestep.cpp:202-253): the smallest of the
// convective, viscous, gravity and capillary time scales over all cells.
// Returns the Courant number; with adjust the time step becomes
// maxCo * that minimum, capped at maxTimeStep.
double time_step(Oracle& o, int adjust) {
    double minDt = 1000000.0;
    for (int i = 0; i < o.nc; ++i) {
        double maxA = 0.0, minA = 1000000.0;
        for (int k = o.cfptr[i]; k < o.cfptr[i + 1]; ++k) {
            maxA = std::max(maxA, o.farea[o.cfidx[k]]);
            minA = std::min(minA, o.farea[o.cfidx[k]]);
        }
        double magVel = sqrt(o.U[i] * o.U[i] + o.V[i] * o.V[i] + o.W[i] * o.W[i]);
        double convTime = o.cvol[i] / maxA / (magVel + 1.e-200);
        minDt = std::min(minDt, convTime);
        double viscTime = 0.5 * minA / (o.Mu[i] + 1.e-200) * o.Rho[i];
        minDt = std::min(minDt, viscTime);
        double gravMag = sqrt(pow(o.g[0], 2.0) + pow(o.g[1], 2.0) + pow(o.g[2], 2.0));
        double gravTime = sqrt(sqrt(minA) / (gravMag + 1.e-200));
        minDt = std::min(minDt, gravTime);
        double stTime;
        if (adjust) {
            double minDX = o.cvol[i] / maxA;
            double C_sigma = sqrt(6.28 * o.sigma / (o.rho[0] + o.rho[1]) / minDX);
            stTime = minDX / (C_sigma + magVel);
        } else {
            stTime = sqrt(pow(minA, 1.5) * (1000.0 + 1.2) / (4.0 * 3.141592 * 0.0728));
        }
        minDt = std::min(minDt, stTime);
    }
    if (!adjust) return o.dt / minDt;
    o.dt = o.max_co * minDt;
    double co = o.max_co;
    if (o.dt > o.max_dt) {
        co = o.max_co * o.max_dt / o.dt;
        o.dt = o.max_dt;
    }
    return co;
}

// one call of incompressiblePressureBased (pressureBased.cpp:6-413): time
// levels, time step, n_vof VOF iterations, body-force sources, then ncorr
// momentum + pressure iterations.
void pressure_based_step(Oracle& o, int ncorr, double* norms) {
    if (o.iterReal == 0) cell_sound_speed(o);
    if (o.iterReal == 0) { cell_eos_transport(o); cell_les(o, true); }
    if (o.iterReal == 0) { o.Qm0 = o.U; o.Qm1 = o.V; o.Qm2 = o.W; o.Qm3 = o.VF; }
    else { o.Qm0 = o.oldU; o.Qm1 = o.oldV; o.Qm2 = o.oldW; o.Qm3 = o.oldVF; }
    o.oldP = o.P; o.oldU = o.U; o.oldV = o.V; o.oldW = o.W; o.oldVF = o.VF; o.oldRho = o.Rho;
    o.courant = time_step(o, o.adjust_dt);            // pressureBased.cpp:74-90
    for (int k = 0; k < o.n_vof; ++k) {               // pressureBased.cpp:104-127
        mstacs_faces(o);
        assemble_vof(o);
        solve_last(o, 1);
        o.vof_norm = vof_update(o);
        cell_eos_transport(o);
        cell_les(o, false);
    }
    recon_faces(o);
    minmax(o, o.VF, o.maxVF, o.minVF);
    source_gravity(o);
    source_surface_tension(o);
    for (int it = 0; it < ncorr; ++it) {
        recon_faces(o);
        assemble_momentum(o);
        solve_last(o, 3);
        double n0 = momentum_update(o);
        recon_faces(o);
        assemble_pressure(o);
        solve_last(o, 1);
        double n1 = pressure_correct(o);
        if (norms) { norms[2 * it] = n0; norms[2 * it + 1] = n1; }
    }
    o.iterReal++;
}
#endif  // ORC_AD

// ------------------------------------------------------------------------
// GENRCM (George & Liu, SPARSPAK) as called by lib_RCM (src/math/RCM.cpp:34-210,
// 1075-1327, 2541-2676, 2839-2876, 3353-3777).  0-based arrays internally;
// the traversal order (component roots, pseudo-peripheral search, degree-
// ascending stable insertion of each new level segment, final reversal) is the
// published algorithm's.
struct Graph { int n; const int* ptr; const int* adj; };

void level_set(const Graph& G, int root, vector<int>& mask, int& nlev, vector<int>& lptr,
               vector<int>& lev) {
    lev.clear(); lptr.clear();
    mask[root] = 0; lev.push_back(root);
    size_t lbeg = 0, lend = 1;
    while (true) {
        lptr.push_back((int)lbeg);
        for (size_t i = lbeg; i < lend; ++i) {
            int node = lev[i];
            for (int k = G.ptr[node]; k < G.ptr[node + 1]; ++k) {
                int nb = G.adj[k];
                if (mask[nb] != 0) { lev.push_back(nb); mask[nb] = 0; }
            }
        }
        lbeg = lend; lend = lev.size();
        if (lend <= lbeg) break;
    }
    nlev = (int)lptr.size();
    lptr.push_back((int)lev.size());
    for (int v : lev) mask[v] = 1;
}

int root_find(const Graph& G, int root, vector<int>& mask) {
    int nlev; vector<int> lptr, lev;
    level_set(G, root, mask, nlev, lptr, lev);
    int cc = lptr[nlev];
    if (nlev == 1 || nlev == cc) return root;
    while (true) {
        int mindeg = cc;
        int js = lptr[nlev - 1];
        root = lev[js];
        if (js < cc - 1) {
            for (int j = js; j < cc; ++j) {
                int node = lev[j], nd = 0;
                for (int k = G.ptr[node]; k < G.ptr[node + 1]; ++k) if (mask[G.adj[k]] > 0) ++nd;
                if (nd < mindeg) { root = node; mindeg = nd; }
            }
        }
        int nlev2; vector<int> lptr2, lev2;
        level_set(G, root, mask, nlev2, lptr2, lev2);
        if (nlev2 <= nlev) break;
        nlev = nlev2; lptr = lptr2; lev = lev2;
        if (cc <= nlev) break;
    }
    return root;
}

int rcm_component(const Graph& G, int root, vector<int>& mask, int* perm, vector<int>& deg) {
    // degree(): BFS over the component, deg = number of unmasked neighbours
    vector<int> comp; comp.push_back(root);
    vector<char> seen(0);
    {
        std::vector<char>& vis = seen; vis.assign(G.n, 0); vis[root] = 1;
        for (size_t i = 0; i < comp.size(); ++i) {
            int node = comp[i], d = 0;
            for (int k = G.ptr[node]; k < G.ptr[node + 1]; ++k) {
                int nb = G.adj[k];
                if (mask[nb] != 0) { ++d; if (!vis[nb]) { vis[nb] = 1; comp.push_back(nb); } }
            }
            deg[node] = d;
        }
    }
    int cc = (int)comp.size();
    mask[root] = 0;
    perm[0] = root;
    if (cc == 1) return 1;
    int lvlend = 0, lnbr = 1;
    while (lvlend < lnbr) {
        int lbeg = lvlend; lvlend = lnbr;
        for (int i = lbeg; i < lvlend; ++i) {
            int node = perm[i];
            int fnbr = lnbr;
            for (int k = G.ptr[node]; k < G.ptr[node + 1]; ++k) {
                int nb = G.adj[k];
                if (mask[nb] != 0) { mask[nb] = 0; perm[lnbr++] = nb; }
            }
            if (lnbr - fnbr <= 1) continue;
            // stable insertion sort by degree (RCM.cpp:3549-3570); the loop
            // condition `while (fnbr < l)` of the reference never compares the
            // segment's first node, which therefore stays first
            for (int k = fnbr + 1; k < lnbr; ++k) {
                int nb = perm[k], l = k - 1;
                while (l > fnbr && deg[perm[l]] > deg[nb]) { perm[l + 1] = perm[l]; --l; }
                perm[l + 1] = nb;
            }
        }
    }
    std::reverse(perm, perm + cc);
    return cc;
}

}  // namespace

// ==========================================================================
// C interface (ctypes)
extern "C" {

void* orc_create(int np, int nc, int nf, int nif, const double* pts, const int* owner,
                 const int* nbr, const int* fptr, const int* fpts, const int* cptr,
                 const int* cpts, int npatch, const int* pstart, const int* pnf) {
    Oracle* o = new Oracle;
    o->np = np; o->nc = nc; o->nf = nf; o->nif = nif;
    o->pts.assign(pts, pts + 3 * (size_t)np);
    o->owner.assign(owner, owner + nf); o->nbr.assign(nbr, nbr + nif);
    o->fptr.assign(fptr, fptr + nf + 1); o->fpts.assign(fpts, fpts + fptr[nf]);
    o->cptr.assign(cptr, cptr + nc + 1); o->cpts.assign(cpts, cpts + cptr[nc]);
    for (int b = 0; b < npatch; ++b) {
        Patch p{}; p.start = pstart[b]; p.nfaces = pnf[b];
        p.p_type = ZERO_GRADIENT; p.u_type = NO_SLIP; p.vf_type = ZERO_GRADIENT;
        o->patches.push_back(p);
    }
    build_geometry(*o);
    auto z = [&](auto& v, double x = 0.0) { v.assign(nc, x); };
    z(o->P); z(o->U); z(o->V); z(o->W); z(o->VF, 1.0); z(o->MF, 1.0); z(o->Rho); z(o->Mu);
    z(o->Qm0); z(o->Qm1); z(o->Qm2); z(o->Qm3, 1.0);
    z(o->oldP); z(o->oldU); z(o->oldV); z(o->oldW); z(o->oldVF, 1.0); z(o->oldRho);
    for (int k = 0; k < 4; ++k) { z(o->srcG[k]); z(o->srcS[k]); }
    z(o->kappa); z(o->alphaP);
    o->ufix.assign(3 * (size_t)(o->nf - o->nif), 0.0);
    return o;
}

void orc_destroy(void* h) { delete (Oracle*)h; }

// VF "function" boundary values: for every boundary face, the VF of the
// nearest of the n points (vf, x, y, z) -- the rows of ./VF.csv, columns 3-6
// (reconIncom.cpp:929-1019): squared distance by pow(.., 2.0), first point of
// the smallest distance below 1e9 wins
void orc_set_vf_points(void* h, int n, const double* p4) {
    Oracle& o = *(Oracle*)h;
    int nb = o.nf - o.nif;
    o.vffun.assign(nb, 0.0);
    for (int j = 0; j < nb; ++j) {
        const double* x = &o.fxyz[3 * (size_t)(o.nif + j)];
        double best = 1.e9, v = 0.0;
        for (int k = 0; k < n; ++k) {
            const double* q = p4 + 4 * (size_t)k;
            double d2 = (pow(x[0] - q[1], 2.0) + pow(x[1] - q[2], 2.0) + pow(x[2] - q[3], 2.0));
            if (d2 < best) { best = d2; v = q[0]; }
        }
        o.vffun[j] = v;
    }
}

void orc_set_bc(void* h, int b, int p_type, double p_val, int u_type, const double* u_val,
                int vf_type, double vf_val) {
    Oracle& o = *(Oracle*)h;
    Patch& p = o.patches[b];
    p.p_type = p_type; p.p_val = p_val; p.u_type = u_type;
    for (int a = 0; a < 3; ++a) p.u_val[a] = u_val[a];
    p.vf_type = vf_type; p.vf_val = vf_val;
}

void orc_set_case(void* h, const double* rho, const double* mu, double sigma, const double* g,
                  double dt, double urf_u, double urf_p, double urf_pu, double lim) {
    Oracle& o = *(Oracle*)h;
    o.rho[0] = rho[0]; o.rho[1] = rho[1]; o.mu[0] = mu[0]; o.mu[1] = mu[1];
    o.sigma = sigma; for (int a = 0; a < 3; ++a) o.g[a] = g[a];
    o.dt = o.oldDt = o.old2Dt = dt;
    o.urf_u = urf_u; o.urf_p = urf_p; o.urf_pu = urf_pu; o.lim = lim;
}

// two-phase controls: vofVofURF and nSpeciesEq (VOF iterations per step)
#ifndef ORC_AD
void orc_set_les(void* h, int model) { ((Oracle*)h)->les_model = model; }
void orc_set_vof(void* h, double urf_vf, int n_vof) {
    Oracle& o = *(Oracle*)h;
    o.urf_vf = urf_vf; o.n_vof = n_vof;
}

// adjustTimeStep, maxCo, maxTimeStep (controls/build.cpp)
void orc_set_time_control(void* h, int adjust, double max_co, double max_dt) {
    Oracle& o = *(Oracle*)h;
    o.adjust_dt = adjust; o.max_co = max_co; o.max_dt = max_dt;
}
double orc_time_step(void* h, int adjust) { return time_step(*(Oracle*)h, adjust); }
double orc_get_dt(void* h) { return ((Oracle*)h)->dt; }
void orc_set_dt(void* h, double dt) { ((Oracle*)h)->dt = dt; }
double orc_courant(void* h) { return ((Oracle*)h)->courant; }

void orc_set_solver(void* h, double tol, int maxit) {
    Oracle& o = *(Oracle*)h; o.tol = tol; o.maxit = maxit;
}

static vector<double>* field(Oracle& o, const char* nm) {
    std::string s(nm);
#define F(name, var) if (s == name) return &o.var;
    F("P", P) F("U", U) F("V", V) F("W", W) F("VF", VF) F("Rho", Rho) F("mu", Mu) F("muT", muT) F("kSGS", kSGS)
    F("oldU", oldU) F("oldV", oldV) F("oldW", oldW) F("Qm0", Qm0) F("Qm1", Qm1) F("Qm2", Qm2)
    F("gradP", gP) F("gradU", gU) F("gradV", gV) F("gradW", gW) F("gradVF", gVF)
    F("maxP", maxP) F("minP", minP) F("maxU", maxU) F("minU", minU) F("maxV", maxV) F("minV", minV)
    F("maxW", maxW) F("minW", minW)
    F("srcG0", srcG[0]) F("srcG1", srcG[1]) F("srcG2", srcG[2]) F("srcG3", srcG[3])
    F("face_area", farea) F("face_normal", fnrm) F("face_xyz", fxyz) F("face_wC", fwC)
    F("face_magPN", fmagPN) F("face_alphaF", falpha) F("face_ePN", fePN) F("face_skew", fskew)
    F("face_dist", fdist) F("cell_volume", cvol) F("cell_xyz", cxyz) F("cell_lsq", lsq)
    F("b0", B0) F("b1", B1) F("b2", B2) F("x0", X0) F("x1", X1) F("x2", X2) F("vals", Avals)
    F("faceRP", RP) F("faceRU", RU) F("faceRV", RV) F("faceRW", RW) F("faceRVF", RVF)
    F("faceLU", LU) F("faceLV", LV) F("faceLW", LW) F("faceLRho", LRho) F("faceLVF", LVF)
    F("oldVF", oldVF) F("Qm3", Qm3) F("courant", co_cell) F("kappa", kappa) F("LS", LS)
    F("srcS0", srcS[0]) F("srcS1", srcS[1]) F("srcS2", srcS[2]) F("srcS3", srcS[3])
    F("alphaP", alphaP) F("ufix", ufix) F("C", Csnd)
#undef F
    return nullptr;
}

// size of a double array, -1 if unknown
long orc_size(void* h, const char* nm) {
    Oracle& o = *(Oracle*)h;
    std::string s(nm);
    if (s == "stencil_ptr") return o.sptr.size();
    if (s == "stencil_idx") return o.sidx.size();
    if (s == "rows" || s == "cols") return o.Arows.size();
    vector<double>* v = field(o, nm);
    return v ? (long)v->size() : -1;
}

int orc_get(void* h, const char* nm, void* out) {
    Oracle& o = *(Oracle*)h;
    std::string s(nm);
    if (s == "stencil_ptr") { memcpy(out, o.sptr.data(), o.sptr.size() * 4); return 0; }
    if (s == "stencil_idx") { memcpy(out, o.sidx.data(), o.sidx.size() * 4); return 0; }
    if (s == "rows") { memcpy(out, o.Arows.data(), o.Arows.size() * 4); return 0; }
    if (s == "cols") { memcpy(out, o.Acols.data(), o.Acols.size() * 4); return 0; }
    vector<double>* v = field(o, nm);
    if (!v) return -1;
    memcpy(out, v->data(), v->size() * 8);
    return 0;
}

int orc_set(void* h, const char* nm, const double* in) {
    Oracle& o = *(Oracle*)h;
    vector<double>* v = field(o, nm);
    if (!v) return -1;
    memcpy(v->data(), in, v->size() * 8);
    if (v == &o.ufix) o.ufix_on = true;
    return 0;
}

// stage entry points
void orc_eos(void* h) { cell_eos_transport(*(Oracle*)h); }
void orc_recon(void* h) { recon_faces(*(Oracle*)h); }
void orc_gravity(void* h) { source_gravity(*(Oracle*)h); }
void orc_surface_tension(void* h) { source_surface_tension(*(Oracle*)h); }
void orc_curvature(void* h) { curvature(*(Oracle*)h); }
void orc_lsq(void* h, const double* phi, const double* faceR, double* grad) {
    Oracle& o = *(Oracle*)h;
    vector<double> p(phi, phi + o.nc), fr(faceR, faceR + o.nf), gr;
    lsq_grad(o, p, fr, gr);
    memcpy(grad, gr.data(), gr.size() * 8);
}
void orc_minmax(void* h, const double* phi, double* mx, double* mn) {
    Oracle& o = *(Oracle*)h;
    vector<double> p(phi, phi + o.nc), a, b;
    minmax(o, p, a, b);
    memcpy(mx, a.data(), a.size() * 8); memcpy(mn, b.data(), b.size() * 8);
}
void orc_gauss_green(void* h, int sel, const double* phi, const double* faceR, double* grad) {
    gauss_green(*(Oracle*)h, sel, phi, faceR, grad);
}
void orc_les(void* h, int model, const double* U, const double* V, const double* W, const double* Rho,
             const double* RU, const double* RV, const double* RW, double* muT, double* kSGS) {
    les(*(Oracle*)h, model, U, V, W, Rho, RU, RV, RW, muT, kSGS);
}
void orc_assemble_momentum(void* h) { assemble_momentum(*(Oracle*)h); }
void orc_assemble_pressure(void* h) { assemble_pressure(*(Oracle*)h); }
int orc_solve(void* h, int nrhs, double* relres) {
    Oracle& o = *(Oracle*)h; solve_last(o, nrhs); if (relres) *relres = o.last_res; return o.last_iters;
}
double orc_momentum_update(void* h) { return momentum_update(*(Oracle*)h); }
double orc_pressure_correct(void* h) { return pressure_correct(*(Oracle*)h); }
void orc_mstacs(void* h) { mstacs_faces(*(Oracle*)h); }
void orc_assemble_vof(void* h) { assemble_vof(*(Oracle*)h); }
double orc_vof_update(void* h) { return vof_update(*(Oracle*)h); }
double orc_vof_norm(void* h) { return ((Oracle*)h)->vof_norm; }
void orc_step(void* h, int ncorr, double* norms) { pressure_based_step(*(Oracle*)h, ncorr, norms); }

#else  // ORC_AD: the CPU adjoint (liborc_ad.so)

void orc_set_solver(void* h, double tol, int maxit) {
    Oracle& o = *(Oracle*)h; o.tol = tol; o.maxit = maxit;
}

// P, U, V, W or alphaP from doubles / to doubles
int orc_ad_set(void* h, const char* nm, const double* in) {
    Oracle& o = *(Oracle*)h;
    std::string s(nm);
    vector<Real>* v = s == "P" ? &o.P : s == "U" ? &o.U : s == "V" ? &o.V : s == "W" ? &o.W
                    : s == "alphaP" ? &o.alphaP : nullptr;
    if (!v) return -1;
    for (int i = 0; i < o.nc; ++i) (*v)[i] = in[i];
    return 0;
}
int orc_ad_get(void* h, const char* nm, double* out) {
    Oracle& o = *(Oracle*)h;
    std::string s(nm);
    const vector<Real>* v = s == "P" ? &o.P : s == "U" ? &o.U : s == "V" ? &o.V : s == "W" ? &o.W : nullptr;
    if (!v) return -1;
    for (int i = 0; i < o.nc; ++i) out[i] = val((*v)[i]);
    return 0;
}

}  // extern "C"

namespace {
// the set-up of one time step of pressure_based_step for a single-phase case
// with a fixed time step and no body forces (the adjoint's scope)
void ad_step_setup(Oracle& o) {
    if (o.iterReal == 0) { cell_sound_speed(o); cell_eos_transport(o); }
    if (o.iterReal == 0) { o.Qm0 = o.U; o.Qm1 = o.V; o.Qm2 = o.W; o.Qm3 = o.VF; }
    else { o.Qm0 = o.oldU; o.Qm1 = o.oldV; o.Qm2 = o.oldW; o.Qm3 = o.oldVF; }
    o.oldP = o.P; o.oldU = o.U; o.oldV = o.V; o.oldW = o.W; o.oldVF = o.VF; o.oldRho = o.Rho;
    recon_faces(o);
    source_gravity(o);
    for (int k = 0; k < 4; ++k) o.srcS[k].assign(o.nc, 0.0);
    o.kappa.assign(o.nc, 0.0);
}
// one outer iteration (pressureBased.cpp:139-212 for one corrector)
void ad_iteration(Oracle& o) {
    recon_faces(o);
    assemble_momentum(o);
    solve_last(o, 3);
    momentum_update(o);
    recon_faces(o);
    assemble_pressure(o);
    solve_last(o, 1);
    pressure_correct(o);
}
// J = sum_patches w_p sum_f A_f p_f (p_f: a fixedValue patch's value, else the
// owner's pressure) and dJ/dP
double ad_objective(const Oracle& o, const double* w, vector<double>* dJdP) {
    double J = 0.0;
    if (dJdP) dJdP->assign(o.nc, 0.0);
    for (size_t b = 0; b < o.patches.size(); ++b) {
        const Patch& pb = o.patches[b];
        for (int f = pb.start; f < pb.start + pb.nfaces; ++f) {
            int c = o.owner[f];
            J += w[b] * o.farea[f] * (pb.p_type == FIXED_VALUE ? pb.p_val : val(o.P[c]));
            if (dJdP && pb.p_type != FIXED_VALUE) (*dJdP)[c] += w[b] * o.farea[f];
        }
    }
    return J;
}
}  // namespace

extern "C" {

// J after one time step of n_iter outer iterations from the current state
// (the state advances; no tape)
double orc_ad_forward(void* h, int n_iter, const double* w) {
    Oracle& o = *(Oracle*)h;
    ad::tape().on = false;
    ad_step_setup(o);
    for (int it = 0; it < n_iter; ++it) ad_iteration(o);
    o.iterReal++;
    return ad_objective(o, w, nullptr);
}

// J and dJ/dalpha_p of that step by reverse-mode differentiation: the state
// at the start of every outer iteration is kept (U, V, W, P); the sweep runs
// backwards, re-taping one iteration from its checkpoint with the checkpoint
// and alpha_p as independent variables, seeding the iteration's outputs with
// the adjoint of the state after it, and reversing the tape.  The old time
// levels are constants of the step.  Leaves the forward's final state.
double orc_ad_adjoint(void* h, int n_iter, const double* w, double* grad) {
    Oracle& o = *(Oracle*)h;
    ad::Tape& T = ad::tape();
    T.on = false; T.clear();
    int n = o.nc;
    ad_step_setup(o);
    vector<vector<double>> ck(n_iter);
    for (int it = 0; it < n_iter; ++it) {
        vector<double>& c = ck[it];
        c.resize(4 * (size_t)n);
        for (int i = 0; i < n; ++i) {
            c[i] = val(o.U[i]); c[n + i] = val(o.V[i]); c[2 * n + i] = val(o.W[i]); c[3 * n + i] = val(o.P[i]);
        }
        ad_iteration(o);
    }
    o.iterReal++;
    vector<double> fin[4] = {values(o.U), values(o.V), values(o.W), values(o.P)};
    vector<double> bar[4];
    for (int q = 0; q < 3; ++q) bar[q].assign(n, 0.0);
    double J = ad_objective(o, w, &bar[3]);
    vector<double> alpha = values(o.alphaP), abar(n, 0.0);
    vector<Real>* st[4] = {&o.U, &o.V, &o.W, &o.P};
    for (int it = n_iter - 1; it >= 0; --it) {
        T.clear(); T.on = true;
        vector<int> in[4], ain(n);
        for (int q = 0; q < 4; ++q) {
            in[q].resize(n);
            for (int i = 0; i < n; ++i) {
                (*st[q])[i] = ad::adouble::var(ck[it][(size_t)q * n + i]);
                in[q][i] = (*st[q])[i].i;
            }
        }
        for (int i = 0; i < n; ++i) { o.alphaP[i] = ad::adouble::var(alpha[i]); ain[i] = o.alphaP[i].i; }
        ad_iteration(o);
        T.on = false;
        vector<double> adj(T.next, 0.0);
        for (int q = 0; q < 4; ++q)
            for (int i = 0; i < n; ++i)
                if ((*st[q])[i].i >= 0) adj[(*st[q])[i].i] += bar[q][i];
        T.reverse(adj);
        for (int q = 0; q < 4; ++q)
            for (int i = 0; i < n; ++i) bar[q][i] = adj[in[q][i]];
        for (int i = 0; i < n; ++i) abar[i] += adj[ain[i]];
    }
    T.clear();
    for (int q = 0; q < 4; ++q)
        for (int i = 0; i < n; ++i) (*st[q])[i] = fin[q][i];
    for (int i = 0; i < n; ++i) { o.alphaP[i] = alpha[i]; grad[i] = abar[i]; }
    return J;
}

#endif  // ORC_AD

// RCM of an arbitrary graph: adjacency CSR (0-based), perm[new] = old,
// inv[old] = new (lib_RCM's cell.RCM / cell.invRCM).
void orc_genrcm(int n, const int* ptr, const int* adj, int* perm, int* inv) {
    Graph G{n, ptr, adj};
    vector<int> mask(n, 1), deg(n, 0);
    int num = 0;
    for (int i = 0; i < n; ++i) {
        if (mask[i] == 0) continue;
        int root = root_find(G, i, mask);
        num += rcm_component(G, root, mask, perm + num, deg);
        if (num >= n) break;
    }
    for (int i = 0; i < n; ++i) inv[perm[i]] = i;
}

// lib_RCM's adjacency (RCM.cpp:79-126): for each cell, its internal faces in
// ascending face order, the other cell.
void orc_cell_graph(void* h, int* ptr, int* adj) {
    Oracle& o = *(Oracle*)h;
    ptr[0] = 0; int m = 0;
    for (int i = 0; i < o.nc; ++i) {
        for (int k = o.cfptr[i]; k < o.cfptr[i + 1]; ++k) {
            int f = o.cfidx[k];
            if (f >= o.nif) continue;
            adj[m++] = o.owner[f] == i ? o.nbr[f] : o.owner[f];
        }
        ptr[i + 1] = m;
    }
}

}  // extern "C"

// ------------------------------------------------------------------------
// Density-based implicit updates (src/solvers/linearSolver.cpp).  A cell's
// state record holds, in order: P, U, V, W, T, rho, C, Ht, dRhoDP, dHtDP,
// dRhoDT, dHtDT, Ur, dtPseudo, Y_0..Y_{nsp-1}, dRhoDY_0..dRhoDY_{nsp-2},
// dHtDY_0..dHtDY_{nsp-2}: 12 + 3 nsp values.  nEq = 4 + nsp equations
// (controls/build.cpp:640).  Matrices are N x N row-major.
namespace db {
enum { P = 0, U, V, W, T, RHO, C, HT, DRDP, DHDP, DRDT, DHDT, UR, DTP, Y0 };
inline int rec(int nsp) { return 12 + 3 * nsp; }
inline double Y(const double* s, int k) { return s[Y0 + k]; }
inline double dRdY(const double* s, int nsp, int k) { return s[Y0 + nsp + k]; }
inline double dHdY(const double* s, int nsp, int k) { return s[Y0 + 2 * nsp - 1 + k]; }

// constructMatrix (linearSolver.cpp:791-890): the conservative-to-primitive
// Jacobian T and its preconditioned form P (column 0 with beta).
void construct(const double* s, int nsp, double* Pm, double* Tm) {
    const int N = 4 + nsp;
    std::fill(Tm, Tm + N * N, 0.0);
    auto t = [&](int i, int j) -> double& { return Tm[i * N + j]; };
    const double DRDP_ = s[DRDP], DHDP_ = s[DHDP], DRDT_ = s[DRDT], DHDT_ = s[DHDT];
    const double Rho = s[RHO], U_ = s[U], V_ = s[V], W_ = s[W], Ht = s[HT];
    t(0, 0) = DRDP_; t(1, 0) = DRDP_ * U_; t(2, 0) = DRDP_ * V_; t(3, 0) = DRDP_ * W_;
    t(4, 0) = DRDP_ * Ht + Rho * DHDP_ - 1.0;
    t(1, 1) = Rho; t(4, 1) = Rho * U_;
    t(2, 2) = Rho; t(4, 2) = Rho * V_;
    t(3, 3) = Rho; t(4, 3) = Rho * W_;
    t(0, 4) = DRDT_; t(1, 4) = DRDT_ * U_; t(2, 4) = DRDT_ * V_; t(3, 4) = DRDT_ * W_;
    t(4, 4) = DRDT_ * Ht + Rho * DHDT_;
    for (int a = 0; a < nsp - 1; ++a) {
        int q = 5 + a;
        t(q, 0) = DRDP_ * Y(s, a);
        t(q, 4) = DRDT_ * Y(s, a);
        for (int b = 0; b < nsp - 1; ++b) t(q, 5 + b) = Y(s, a) * dRdY(s, nsp, b);
        t(q, q) += Rho;
        t(0, q) = dRdY(s, nsp, a); t(1, q) = dRdY(s, nsp, a) * U_; t(2, q) = dRdY(s, nsp, a) * V_;
        t(3, q) = dRdY(s, nsp, a) * W_;
        t(4, q) = dRdY(s, nsp, a) * Ht + Rho * dHdY(s, nsp, a);
    }
    std::copy(Tm, Tm + N * N, Pm);
    double beta = 1.0 / std::pow(s[UR], 2.0) + DRDP_ - 1.0 / std::pow(s[C], 2.0);
    Pm[0] = beta; Pm[N] = beta * U_; Pm[2 * N] = beta * V_; Pm[3 * N] = beta * W_;
    Pm[4 * N] = beta * Ht + Rho * DHDP_ - 1.0;
    for (int a = 0; a < nsp - 1; ++a) Pm[(5 + a) * N] = beta * Y(s, a);
}

// getNumericalFluxJacobianUpwind (linearSolver.cpp:394-640): 0.5 dF/dQ of the
// chosen side (right = the "R" side) +/- 0.5 of a Roe-type dissipation matrix
// built from the face's averages; the species rows carry no dissipation.
void jacobian(const double* L, const double* R, int nsp, const double* n, bool right, double* J) {
    const int N = 4 + nsp;
    double UnL = L[U] * n[0] + L[V] * n[1] + L[W] * n[2];
    double UnR = R[U] * n[0] + R[V] * n[1] + R[W] * n[2];
    double RT = std::sqrt(L[RHO] / R[RHO]);
    double Rhohat = 0.5 * (L[RHO] + R[RHO]);
    double Chat = 0.5 * (L[C] + R[C]);
    double Unhat = 0.5 * (UnL + UnR);
    double Uhat = (RT * L[U] + R[U]) / (RT + 1.0);
    double Vhat = (RT * L[V] + R[V]) / (RT + 1.0);
    double What = (RT * L[W] + R[W]) / (RT + 1.0);
    double Hthat = (RT * L[HT] + R[HT]) / (RT + 1.0);
    double fgL = right ? 0.0 : 1.0, fgR = right ? 1.0 : 0.0;
    auto lr = [&](int k) { return L[k] * fgL + R[k] * fgR; };
    double ULR = lr(U), VLR = lr(V), WLR = lr(W), HtLR = lr(HT), RhoLR = lr(RHO);
    double DRDPLR = lr(DRDP), DHDPLR = lr(DHDP), DRDTLR = lr(DRDT), DHDTLR = lr(DHDT);
    double UnLR = UnL * fgL + UnR * fgR;
    vector<double> F(N * N, 0.0), D(N * N, 0.0), YLR(nsp), DRDYLR(nsp), DHDYLR(nsp);
    for (int a = 0; a < nsp - 1; ++a) {
        YLR[a] = Y(L, a) * fgL + Y(R, a) * fgR;
        DRDYLR[a] = dRdY(L, nsp, a) * fgL + dRdY(R, nsp, a) * fgR;
        DHDYLR[a] = dHdY(L, nsp, a) * fgL + dHdY(R, nsp, a) * fgR;
    }
    auto f = [&](int i, int j) -> double& { return F[i * N + j]; };
    f(0, 0) = DRDPLR * UnLR;
    f(1, 0) = DRDPLR * UnLR * ULR + n[0];
    f(2, 0) = DRDPLR * UnLR * VLR + n[1];
    f(3, 0) = DRDPLR * UnLR * WLR + n[2];
    f(4, 0) = DRDPLR * UnLR * HtLR + RhoLR * UnLR * DHDPLR;
    for (int d = 0; d < 3; ++d) {
        double vel[3] = {ULR, VLR, WLR};
        f(0, 1 + d) = RhoLR * n[d];
        for (int r = 0; r < 3; ++r) f(1 + r, 1 + d) = RhoLR * n[d] * vel[r];
        f(1 + d, 1 + d) = f(1 + d, 1 + d) + RhoLR * UnLR;
        f(4, 1 + d) = RhoLR * n[d] * HtLR + RhoLR * UnLR * vel[d];
    }
    f(0, 4) = DRDTLR * UnLR;
    f(1, 4) = DRDTLR * UnLR * ULR;
    f(2, 4) = DRDTLR * UnLR * VLR;
    f(3, 4) = DRDTLR * UnLR * WLR;
    f(4, 4) = DRDTLR * UnLR * HtLR + RhoLR * UnLR * DHDTLR;
    for (int a = 0; a < nsp - 1; ++a) {
        int q = 5 + a;
        f(q, 0) = DRDPLR * YLR[a] * UnLR;
        f(q, 1) = RhoLR * YLR[a] * n[0];
        f(q, 2) = RhoLR * YLR[a] * n[1];
        f(q, 3) = RhoLR * YLR[a] * n[2];
        f(q, 4) = DRDTLR * YLR[a] * UnLR;
        for (int b = 0; b < nsp - 1; ++b) f(q, 5 + b) = DRDYLR[b] * YLR[a] * UnLR;
        f(q, q) += RhoLR * UnLR;
        f(0, q) = DRDYLR[a] * UnLR;
        f(1, q) = DRDYLR[a] * UnLR * ULR;
        f(2, q) = DRDYLR[a] * UnLR * VLR;
        f(3, q) = DRDYLR[a] * UnLR * WLR;
        f(4, q) = DRDYLR[a] * UnLR * HtLR + RhoLR * UnLR * DHDYLR[a];
    }
    double dissUp = (Chat - std::fabs(Unhat)) / Rhohat / Chat / Chat;
    double dissUu = Unhat / Chat;
    double dissPp = Unhat / Chat;
    double dissPu = (Chat - std::fabs(Unhat)) * Rhohat;
    double AD = std::fabs(Unhat);
    double hat[3] = {Uhat, Vhat, What}, vel[3] = {ULR, VLR, WLR};
    auto d = [&](int i, int j) -> double& { return D[i * N + j]; };
    d(0, 0) = AD * DRDPLR + dissUp * Rhohat;
    for (int r = 0; r < 3; ++r) d(1 + r, 0) = AD * DRDPLR * vel[r] + dissUp * Rhohat * hat[r] + dissPp * n[r];
    d(4, 0) = AD * (DRDPLR * HtLR + RhoLR * DHDPLR - 1.0) + dissUp * Rhohat * Hthat + dissPp * Unhat;
    for (int c = 0; c < 3; ++c) {
        d(0, 1 + c) = dissUu * n[c] * Rhohat;
        for (int r = 0; r < 3; ++r) {
            d(1 + r, 1 + c) = r == c ? AD * RhoLR + dissUu * n[c] * Rhohat * hat[r] + dissPu * n[c] * n[r]
                                     : dissUu * n[c] * Rhohat * hat[r] + dissPu * n[c] * n[r];
        }
        d(4, 1 + c) = AD * RhoLR * vel[c] + dissUu * n[c] * Rhohat * Hthat + dissPu * n[c] * Unhat;
    }
    d(0, 4) = AD * DRDTLR;
    for (int r = 0; r < 3; ++r) d(1 + r, 4) = AD * DRDTLR * vel[r];
    d(4, 4) = AD * (DRDTLR * HtLR + RhoLR * DHDTLR);
    for (int k = 0; k < N * N; ++k) J[k] = right ? 0.5 * F[k] - 0.5 * D[k] : 0.5 * F[k] + 0.5 * D[k];
}

// GaussSeidelSOR (src/math/math.cpp:141-200): A <- inverse(A) by Doolittle
// LU without pivoting, then one forward and one back substitution per unit
// column.
void invert(double* A, int N) {
    vector<double> Lm(N * N, 0.0), Um(N * N, 0.0), B(N, 0.0), D(N, 0.0), X(N, 0.0);
    for (int k = 0; k < N - 1; ++k)
        for (int i = k + 1; i < N; ++i) {
            double coeff = A[i * N + k] / A[k * N + k];
            Lm[i * N + k] = coeff;
            for (int j = k + 1; j < N; ++j) A[i * N + j] = A[i * N + j] - coeff * A[k * N + j];
        }
    for (int i = 0; i < N; ++i) Lm[i * N + i] = 1.0;
    for (int j = 0; j < N; ++j)
        for (int i = 0; i <= j; ++i) Um[i * N + j] = A[i * N + j];
    for (int k = 0; k < N; ++k) {
        B[k] = 1.0;
        D[0] = B[0];
        for (int i = 1; i < N; ++i) {
            D[i] = B[i];
            for (int j = 0; j < i; ++j) D[i] = D[i] - Lm[i * N + j] * D[j];
        }
        X[N - 1] = D[N - 1] / Um[(N - 1) * N + N - 1];
        for (int i = N - 2; i >= 0; --i) {
            X[i] = D[i];
            for (int j = N - 1; j >= i + 1; --j) X[i] = X[i] - Um[i * N + j] * X[j];
            X[i] = X[i] / Um[i * N + i];
        }
        for (int i = 0; i < N; ++i) A[i * N + k] = X[i];
        B[k] = 0.0;
    }
}

// res <- A res, the product order of linearSolver.cpp:261-268
void apply(const double* A, int N, const double* b, double* out) {
    for (int j = 0; j < N; ++j) {
        double u = 0.0;
        for (int k = 0; k < N; ++k) u += A[j * N + k] * b[k];
        out[j] = u;
    }
}
}  // namespace db

// method 0: calcLinearSolver (linearSolver.cpp:666-725), point-implicit
// (P + 1.5 dtau/dt T) per cell; 1: calcSingleLinearSolver (:729-787), T per
// cell; 2: calcLinearSolverLUSGS (:144-389), forward then backward sweep in
// lib_RCM's order (cell.RCM / cell.invRCM) with the upwind flux Jacobians of
// the faces.  state[nc][12 + 3 nsp], res[nc][4 + nsp] in/out, both in the
// mesh's cell order.
void linear_solver(Oracle& o, int method, int nsp, double time_step, const double* state, double* res) {
    using namespace db;
    const int N = 4 + nsp, K = rec(nsp);
    vector<double> Pm(N * N), Tm(N * N), A(N * N), b(N), out(N);
    if (method == 0 || method == 1) {
        for (int i = 0; i < o.nc; ++i) {
            const double* s = state + (size_t)i * K;
            construct(s, nsp, Pm.data(), Tm.data());
            for (int a = 0; a < N; ++a)
                for (int c = 0; c < N; ++c)
                    A[a * N + c] = method == 1 ? Tm[a * N + c]
                                               : Pm[a * N + c] + 1.5 * s[DTP] / time_step * Tm[a * N + c];
            std::copy(res + (size_t)i * N, res + (size_t)(i + 1) * N, b.begin());
            invert(A.data(), N);
            apply(A.data(), N, b.data(), res + (size_t)i * N);
        }
        return;
    }
    // lib_RCM's order (RCM.cpp:34-210)
    vector<int> gptr(o.nc + 1), gadj(2 * (size_t)o.nif + 1), perm(o.nc), inv(o.nc);
    orc_cell_graph(&o, gptr.data(), gadj.data());
    orc_genrcm(o.nc, gptr.data(), gadj.data(), perm.data(), inv.data());
    vector<double> J(N * N), DS(N * N), LU(N);
    auto cell_pass = [&](int pos, bool forward) {
        int own = perm[pos];
        const double* s = state + (size_t)own * K;
        std::fill(DS.begin(), DS.end(), 0.0);
        std::fill(LU.begin(), LU.end(), 0.0);
        double rDtauPerVol = s[DTP] / o.cvol[own];
        for (int k = o.cfptr[own]; k < o.cfptr[own + 1]; ++k) {
            int f = o.cfidx[k];
            double nv[3];
            double sg = o.owner[f] == own ? 1.0 : -1.0;
            for (int a = 0; a < 3; ++a) nv[a] = sg > 0 ? o.fnrm[3 * f + a] : -o.fnrm[3 * f + a];
            if (f < o.nif) {
                int ngb = o.nbr[f] == own ? o.owner[f] : o.nbr[f];
                const double* sn = state + (size_t)ngb * K;
                if (forward ? inv[ngb] < pos : inv[ngb] > pos) {
                    jacobian(s, sn, nsp, nv, true, J.data());
                    for (int a = 0; a < N; ++a) {
                        double sig = 0.0;
                        for (int c = 0; c < N; ++c) sig += res[(size_t)ngb * N + c] * J[a * N + c];
                        LU[a] += rDtauPerVol * sig * o.farea[f];
                    }
                }
                jacobian(s, sn, nsp, nv, false, J.data());
            } else {
                jacobian(s, s, nsp, nv, false, J.data());
            }
            for (int a = 0; a < N * N; ++a) DS[a] += rDtauPerVol * J[a] * o.farea[f];
        }
        construct(s, nsp, Pm.data(), Tm.data());
        for (int a = 0; a < N; ++a)
            for (int c = 0; c < N; ++c) {
                double timeFrac = s[DTP] / time_step;
                A[a * N + c] = Pm[a * N + c] + 1.5 * timeFrac * Tm[a * N + c];
                A[a * N + c] += DS[a * N + c];
            }
        for (int a = 0; a < N; ++a) b[a] = forward ? res[(size_t)own * N + a] - LU[a] : LU[a];
        invert(A.data(), N);
        apply(A.data(), N, b.data(), out.data());
        for (int a = 0; a < N; ++a)
            res[(size_t)own * N + a] = forward ? out[a] : res[(size_t)own * N + a] - out[a];
    };
    for (int i = 0; i < o.nc; ++i) cell_pass(i, true);
    for (int i = o.nc - 1; i >= 0; --i) cell_pass(i, false);
}

extern "C" int orc_linear_solver(void* h, int method, int nsp, double time_step, const double* state,
                                 double* res) {
    if (method < 0 || method > 2 || nsp < 1 || nsp > 8) return -1;
    linear_solver(*(Oracle*)h, method, nsp, time_step, state, res);
    return 0;
}
