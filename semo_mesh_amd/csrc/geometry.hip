// One-time geometry on the device: face area / normal / centroid, cell volume
// and centroid, per-face interpolation metrics and the 3x3 least-squares
// inverse per cell.  Arithmetic follows SEMO_Mesh_Geometric (src/mesh/
// geometric.cpp) and initLeastSquare (src/math/gradient.cpp) term for term, so
// that on an un-renumbered mesh the result matches the CPU path bit for bit
// (the library is built with -ffp-contract=off).
#include "common.hpp"

namespace semo {
namespace {

constexpr int TPB = 256;
inline int nblk(long n) { return (int)((n + TPB - 1) / TPB); }

// sqrt of a double as x86 `sqrtl` then a store to double computes it: rounded
// first to the 64-bit x87 significand, then to 53 bits (ties to even).  That
// differs from the correctly rounded sqrt only when the 64-bit result is a
// midpoint between two doubles; detect that case from the exact FMA remainder
// r = x - s*s of the correctly rounded s and apply the tie-to-even choice.
__device__ inline double sqrt_x87(double x) {
    double s = sqrt(x);
    if (!(x > 0.0) || isinf(x)) return s;
    double r = fma(-s, s, x);                 // exact
    if (r == 0.0) return s;
    // neighbour of s on the side of the true root, and the midpoint offset h
    double nb = r > 0.0 ? nextafter(s, INFINITY) : nextafter(s, 0.0);
    double h = 0.5 * (nb - s);                // exact: half the local spacing, signed
    // (s+h)^2 - x = -(r - 2sh - h^2); the 64-bit rounding lands on the midpoint
    // iff |root - (s+h)| <= 2^-12 |nb - s|, i.e. |r - 2 s h - h^2| <= 2^-11 |nb-s| (s+h)
    double d = r - 2.0 * s * h - h * h;
    double tol = ldexp(fabs(nb - s), -11) * (s + h);
    if (fabs(d) > tol) return s;
    // tie between s and nb: even significand wins
    long long bits_s = __double_as_longlong(s);
    return (bits_s & 1LL) ? nb : s;
}

// calcUnitNormals_Area3dPolygon (geometric.cpp:16-204): 5 passes of a fan
// triangulation about the running area-weighted centroid.  Faces of up to
// 16 points keep their coordinates in registers; larger ones (polyhedral
// meshes) read them from the point array at every use, the same arithmetic.
template <class PtF>
__device__ inline void face_polygon(int n, PtF X, double& ar, double& vs, double* un, double* xt, double* ct) {
    double avg[3];
    for (int a = 0; a < 3; ++a) {
        double s = 0.0;
        for (int k = 0; k < n; ++k) s += X(k, a);
        avg[a] = s / (double)n;
    }
    double V0o[3] = {0, 0, 0};
    xt[0] = xt[1] = xt[2] = 0.0;
    for (int two = 0; two < 5; ++two) {
        un[0] = un[1] = un[2] = 0.0; ct[0] = ct[1] = ct[2] = 0.0; vs = 0.0;
        double V0[3];
        for (int a = 0; a < 3; ++a) V0[a] = two == 0 ? avg[a] : V0o[a] + 1.0 * (xt[a] - V0o[a]);
        for (int a = 0; a < 3; ++a) { V0o[a] = V0[a]; xt[a] = 0.0; }
        double tot = 0.0;
        for (int i = 0; i < n; ++i) {
            int c = (i == n - 1) ? 0 : i + 1;
            double Xi[3] = {X(i, 0), X(i, 1), X(i, 2)}, Xc[3] = {X(c, 0), X(c, 1), X(c, 2)};
            double A0 = Xi[0] - V0[0], A1 = Xi[1] - V0[1], A2 = Xi[2] - V0[2];
            double B0 = Xc[0] - V0[0], B1 = Xc[1] - V0[1], B2 = Xc[2] - V0[2];
            double Nv[3] = {0.5 * (A1 * B2 - A2 * B1), 0.5 * (A2 * B0 - A0 * B2), 0.5 * (A0 * B1 - A1 * B0)};
            un[0] += Nv[0]; un[1] += Nv[1]; un[2] += Nv[2];
            double ta = sqrt_x87(Nv[0] * Nv[0] + Nv[1] * Nv[1] + Nv[2] * Nv[2]);
            for (int a = 0; a < 3; ++a) xt[a] += ((V0[a] + Xi[a] + Xc[a]) / 3.0) * ta;
            tot += ta;
            vs += V0[0] * Nv[0]; vs += V0[1] * Nv[1]; vs += V0[2] * Nv[2];
            for (int a = 0; a < 3; ++a) {
                double s1 = V0[a] + Xi[a], s2 = Xi[a] + Xc[a], s3 = Xc[a] + V0[a];
                ct[a] += 2.0 * Nv[a] * (s1 * s1 + s2 * s2 + s3 * s3);
            }
        }
        double mag = sqrt_x87(un[0] * un[0] + un[1] * un[1] + un[2] * un[2]);
        un[0] /= mag; un[1] /= mag; un[2] /= mag;
        ar = mag;
        for (int a = 0; a < 3; ++a) xt[a] /= tot;
    }
}

constexpr int FACE_REG_POINTS = 16;

__global__ void __launch_bounds__(TPB) k_face_polygon(
    int nf, const int* __restrict__ fptr, const int* __restrict__ fpts,
    const double* __restrict__ pts, double* __restrict__ area, double* __restrict__ nrm,
    double* __restrict__ fxyz, double* __restrict__ vsn_out, double* __restrict__ cc_out,
    int* __restrict__ flags) {
    int f = xcd_block() * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    int b0 = fptr[f], n = fptr[f + 1] - b0;
    if (n < 3) { atomicOr(flags, FLAG_BAD_FACE); return; }
    double ar = 0.0, vs = 0.0, un[3], xt[3], ct[3];
    if (n <= FACE_REG_POINTS) {
        double R[FACE_REG_POINTS][3];
        for (int k = 0; k < n; ++k) {
            int p = fpts[b0 + k];
            R[k][0] = pts[3 * (size_t)p]; R[k][1] = pts[3 * (size_t)p + 1]; R[k][2] = pts[3 * (size_t)p + 2];
        }
        face_polygon(n, [&](int k, int a) { return R[k][a]; }, ar, vs, un, xt, ct);
    } else {
        face_polygon(n, [&](int k, int a) { return pts[3 * (size_t)fpts[b0 + k] + a]; }, ar, vs, un, xt, ct);
    }
    if (!(ar >= 2.2250738585072014e-308)) atomicOr(flags, FLAG_ZERO_AREA);
    area[f] = ar;
    vsn_out[f] = vs;
    for (int a = 0; a < 3; ++a) {
        nrm[3 * (size_t)f + a] = un[a]; fxyz[3 * (size_t)f + a] = xt[a]; cc_out[3 * (size_t)f + a] = ct[a];
    }
}

// cell volume (sum of VSn/3) and centroid (geometric.cpp:804-877), gathered
// over the cell's faces in ascending face order
__global__ void __launch_bounds__(TPB) k_cell_volume(
    int nc, int nif, const int* __restrict__ cf_ptr, const int* __restrict__ cf_idx,
    const int* __restrict__ owner, const double* __restrict__ vsn, const double* __restrict__ cc,
    double* __restrict__ vol, double* __restrict__ cxyz, int* __restrict__ flags) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double v = 0.0, x = 0.0, y = 0.0, z = 0.0;
    for (int k = cf_ptr[i]; k < cf_ptr[i + 1]; ++k) {
        int f = cf_idx[k];
        if (owner[f] == i) {
            v += vsn[f] / 3.0;
            x += cc[3 * (size_t)f]; y += cc[3 * (size_t)f + 1]; z += cc[3 * (size_t)f + 2];
        } else {
            v -= vsn[f] / 3.0;
            x -= cc[3 * (size_t)f]; y -= cc[3 * (size_t)f + 1]; z -= cc[3 * (size_t)f + 2];
        }
    }
    if (!(v >= 2.2250738585072014e-308)) atomicOr(flags, FLAG_NEG_VOLUME);
    double s = 1.0 / (24.0 * 2.0 * v);
    vol[i] = v;
    cxyz[3 * (size_t)i] = x * s; cxyz[3 * (size_t)i + 1] = y * s; cxyz[3 * (size_t)i + 2] = z * s;
}

// per-face metrics (geometric.cpp:1206-1227, 1297-1497)
__global__ void __launch_bounds__(TPB) k_face_metrics(
    int nf, int nif, const int* __restrict__ owner, const int* __restrict__ nbr,
    const int8_t* __restrict__ flip, const double* __restrict__ nrm, const double* __restrict__ fxyz,
    const double* __restrict__ cxyz, double* __restrict__ wC_o, double* __restrict__ magPN_o, double* __restrict__ alpha_o,
    double* __restrict__ ePN_o, double* __restrict__ skew_o, double* __restrict__ pf_o,
    double* __restrict__ nf_o) {
    int f = xcd_block() * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const double* n = nrm + 3 * (size_t)f;
    const double* fc = fxyz + 3 * (size_t)f;
    const double* cP = cxyz + 3 * (size_t)owner[f];
    double e[3], sk[3] = {0, 0, 0}, pf[3], nfv[3];
    double wC, magPN;
    if (f < nif) {
        const double* cN = cxyz + 3 * (size_t)nbr[f];
        double dFP[3], dFN[3], d[3];
        for (int a = 0; a < 3; ++a) { dFP[a] = fc[a] - cP[a]; dFN[a] = fc[a] - cN[a]; d[a] = cN[a] - cP[a]; }
        double dPN = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        double dFPo = fabs(dFP[0] * n[0] + dFP[1] * n[1] + dFP[2] * n[2]);
        double dFNo = fabs(dFN[0] * n[0] + dFN[1] * n[1] + dFN[2] * n[2]);
        // geometric.cpp:1349-1351, rounded through float in the mesh's own
        // orientation: for a face flipped by the renumbering the reference's
        // owner is this face's neighbour, so its weight is rounded and this
        // side gets the exact complement
        if (flip[f]) {
            wC = dFPo / (dFNo + dFPo);
            wC = 1.0 - (double)(float)(round(wC * 10000000) / 10000000);
        } else {
            wC = dFNo / (dFPo + dFNo);
            wC = (double)(float)(round(wC * 10000000) / 10000000);
        }
        magPN = dPN;
        for (int a = 0; a < 3; ++a) { e[a] = d[a] / dPN; pf[a] = dFP[a]; nfv[a] = dFN[a]; }
        double D = -(n[0] * fc[0] + n[1] * fc[1] + n[2] * fc[2]);
        double u = (n[0] * cP[0] + n[1] * cP[1] + n[2] * cP[2] + D) /
                   (n[0] * (cP[0] - cN[0]) + n[1] * (cP[1] - cN[1]) + n[2] * (cP[2] - cN[2]));
        for (int a = 0; a < 3; ++a) sk[a] = fc[a] - (cP[a] - u * (cP[a] - cN[a]));
    } else {
        double dFP[3];
        for (int a = 0; a < 3; ++a) dFP[a] = fc[a] - cP[a];
        double len = sqrt(dFP[0] * dFP[0] + dFP[1] * dFP[1] + dFP[2] * dFP[2]);
        double dFPo = fabs(dFP[0] * n[0] + dFP[1] * n[1] + dFP[2] * n[2]);
        wC = 0.5;
        magPN = 2.0 * dFPo;
        for (int a = 0; a < 3; ++a) { e[a] = dFP[a] / len; pf[a] = dFP[a]; nfv[a] = dFP[a]; }
    }
    double al = 0.0;
    al += n[0] * e[0]; al += n[1] * e[1]; al += n[2] * e[2];
    wC_o[f] = wC; magPN_o[f] = magPN; alpha_o[f] = 1.0 / fabs(al);
    for (int a = 0; a < 3; ++a) {
        ePN_o[3 * (size_t)f + a] = e[a]; skew_o[3 * (size_t)f + a] = sk[a];
        pf_o[3 * (size_t)f + a] = pf[a]; nf_o[3 * (size_t)f + a] = nfv[a];
    }
}

// initLeastSquare, cell-vertex stencil, 3x3 block (gradient.cpp:246-430):
// stencil entries first, then the cell's boundary faces in ascending order
__global__ void __launch_bounds__(TPB) k_lsq_init(
    int nc, int nif, const int* __restrict__ st_ptr, const int* __restrict__ st_idx,
    const int* __restrict__ cf_ptr, const int* __restrict__ cf_idx, const double* __restrict__ cxyz,
    const double* __restrict__ fxyz, double* __restrict__ lsq, int* __restrict__ flags) {
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0;
    double xi = cxyz[3 * (size_t)i], yi = cxyz[3 * (size_t)i + 1], zi = cxyz[3 * (size_t)i + 2];
    auto acc = [&](double dx, double dy, double dz) {
        double wk = 1.0 / (dx * dx + dy * dy + dz * dz);
        s0 += wk * dx * dx; s1 += wk * dx * dy; s2 += wk * dx * dz;
        s3 += wk * dy * dy; s4 += wk * dy * dz; s5 += wk * dz * dz;
    };
    for (int k = st_ptr[i]; k < st_ptr[i + 1]; ++k) {
        int j = st_idx[k];
        acc(cxyz[3 * (size_t)j] - xi, cxyz[3 * (size_t)j + 1] - yi, cxyz[3 * (size_t)j + 2] - zi);
    }
    for (int k = cf_ptr[i]; k < cf_ptr[i + 1]; ++k) {
        int f = cf_idx[k];
        if (f < nif) continue;
        acc(fxyz[3 * (size_t)f] - xi, fxyz[3 * (size_t)f + 1] - yi, fxyz[3 * (size_t)f + 2] - zi);
    }
    double det = s0 * s3 * s5 + s1 * s4 * s2 + s2 * s1 * s4 - s0 * s4 * s4 - s2 * s3 * s2 - s1 * s1 * s5;
    if (det == 0.0) atomicOr(flags, FLAG_SINGULAR_LSQ);
    double* c = lsq + 6 * (size_t)i;
    c[0] = (s3 * s5 - s4 * s4) / det;
    c[1] = (s2 * s4 - s1 * s5) / det;
    c[2] = (s1 * s4 - s2 * s3) / det;
    c[3] = (s0 * s5 - s2 * s2) / det;
    c[4] = (s2 * s1 - s0 * s4) / det;
    c[5] = (s0 * s3 - s1 * s1) / det;
}

}  // namespace

void geometry_compute(semo_ctx* c) {
    hipStream_t s = c->stream;
    int nf = c->nf, nc = c->nc;
    c->f_area.alloc(nf); c->f_nrm.alloc(3 * (size_t)nf); c->f_xyz.alloc(3 * (size_t)nf);
    c->f_vsn.alloc(nf); c->f_cc.alloc(3 * (size_t)nf);
    c->f_wC.alloc(nf); c->f_magPN.alloc(nf); c->f_alpha.alloc(nf); c->f_ePN.alloc(3 * (size_t)nf);
    c->f_skew.alloc(3 * (size_t)nf); c->f_pf.alloc(3 * (size_t)nf); c->f_nf.alloc(3 * (size_t)nf);
    c->c_vol.alloc(nc); c->c_xyz.alloc(3 * (size_t)nc); c->c_lsq.alloc(6 * (size_t)nc);
    SEMO_LAUNCH(c, "setup_face_polygon",
                4.0 * nf + 4.0 * (double)c->d_fpts.n + 24.0 * (double)c->d_fpts.n + 128.0 * nf, k_face_polygon,
                dim3(nblk(nf)), dim3(TPB), nf, c->d_fptr.p, c->d_fpts.p, c->d_pts.p, c->f_area.p, c->f_nrm.p,
                c->f_xyz.p, c->f_vsn.p, c->f_cc.p, c->d_flags.p);
    SEMO_LAUNCH(c, "setup_cell_volume", cell_faces_bytes(c, 32.0) + 32.0 * nc, k_cell_volume, dim3(nblk(nc)),
                dim3(TPB), nc, c->nif, c->d_cf_ptr.p, c->d_cf_idx.p, c->d_owner.p, c->f_vsn.p, c->f_cc.p, c->c_vol.p,
                c->c_xyz.p, c->d_flags.p);
    c->d_fflip.upload(c->fflip, s);
    SEMO_LAUNCH(c, "setup_face_metrics", (8.0 + 1.0 + 48.0 + 160.0) * nf + 24.0 * nc, k_face_metrics, dim3(nblk(nf)),
                dim3(TPB), nf, c->nif, c->d_owner.p, c->d_nbr.p, c->d_fflip.p, c->f_nrm.p, c->f_xyz.p, c->c_xyz.p,
                c->f_wC.p, c->f_magPN.p, c->f_alpha.p, c->f_ePN.p, c->f_skew.p, c->f_pf.p, c->f_nf.p);
    SEMO_LAUNCH(c, "setup_lsq", 4.0 * (double)c->d_st_idx.n + 24.0 * nc + 48.0 * nc, k_lsq_init, dim3(nblk(nc)),
                dim3(TPB), nc, c->nif, c->d_st_ptr.p, c->d_st_idx.p, c->d_cf_ptr.p, c->d_cf_idx.p, c->c_xyz.p,
                c->f_xyz.p, c->c_lsq.p, c->d_flags.p);
    SEMO_HIP(hipGetLastError());
    int flags = 0;
    SEMO_HIP(hipMemcpyAsync(&flags, c->d_flags.p, sizeof(int), hipMemcpyDeviceToHost, s));
    SEMO_HIP(hipStreamSynchronize(s));
    c->f_vsn.release(); c->f_cc.release();
    SEMO_CHECK(!(flags & FLAG_BAD_FACE), SEMO_EINVAL, "face with fewer than 3 points");
    SEMO_CHECK(!(flags & FLAG_ZERO_AREA), SEMO_EZERO_AREA, "zero face area");
    SEMO_CHECK(!(flags & FLAG_NEG_VOLUME), SEMO_ENEG_VOLUME, "non-positive cell volume");
    SEMO_CHECK(!(flags & FLAG_SINGULAR_LSQ), SEMO_ESINGULAR_LSQ, "singular least-squares matrix (detA = 0)");
}

}  // namespace semo
