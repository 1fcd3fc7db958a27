// SEMO's interface-compression utility (src/utility/compressVolumeFraction.cpp:
// 121-228) on the device.  Per iteration: the LSQ gradient of alpha (the
// library's calcLeastSquare), then
//   - per cell the unit normal n = grad / |grad| (zero where |grad| = 0) and
//     q = alpha (1 - alpha);
//   - per face UnF = wC n_own . n_f + (1 - wC) n_ngb . n_f (0 at boundary
//     faces), the upwind q of the side UnF leaves, zeroed when either side's
//     q has q (1 - q) < 1e-15 (the utility tests q, not alpha, there), and
//     flux = qF UnF area;
//   - per cell resi = the faces' fluxes in ascending face order (- where the
//     cell owns the face, + where it is the neighbour), alpha += cfl tau resi
//     / V, clipped to [0, 1] with std::min / std::max's rules, and the
//     update's square for the iteration's residual.
// tau (min over the cells of min(V^0.3, V / largest face area)) is computed
// on the host with the C library's pow, as the utility does.
#include "common.hpp"
#include "kern.hpp"

namespace semo {
namespace {

__global__ __launch_bounds__(TPB) void k_cvf_prep(int n, const double* __restrict__ vf, const double* __restrict__ g,
                                                  double* __restrict__ nq) {
    const int i = xcd_block() * TPB + (int)threadIdx.x;
    if (i >= n) return;
    const double g0 = g[3 * (size_t)i], g1 = g[3 * (size_t)i + 1], g2 = g[3 * (size_t)i + 2];
    const double mag = sqrt(g0 * g0 + g1 * g1 + g2 * g2);
    double n0 = 0.0, n1 = 0.0, n2 = 0.0;
    if (mag != 0.0) { n0 = g0 / mag; n1 = g1 / mag; n2 = g2 / mag; }
    const double a = vf[i];
    nq[4 * (size_t)i] = 1.0 * n0;
    nq[4 * (size_t)i + 1] = 1.0 * n1;
    nq[4 * (size_t)i + 2] = 1.0 * n2;
    nq[4 * (size_t)i + 3] = a * (1.0 - a);
}

__global__ __launch_bounds__(TPB) void k_cvf_face(int nf, int nif, const int* __restrict__ owner,
                                                  const int* __restrict__ nbr, const double* __restrict__ nrm,
                                                  const double* __restrict__ wC, const double* __restrict__ area,
                                                  const double* __restrict__ nq, double* __restrict__ flux) {
    const int f = xcd_block() * TPB + (int)threadIdx.x;
    if (f >= nf) return;
    const double n0 = nrm[3 * (size_t)f], n1 = nrm[3 * (size_t)f + 1], n2 = nrm[3 * (size_t)f + 2];
    const double wCL = wC[f], wCR = 1.0 - wCL;
    const double* L = nq + 4 * (size_t)owner[f];
    double UnF = wCL * (L[0] * n0 + L[1] * n1 + L[2] * n2);
    const double aL = L[3];
    double aR = 0.0;
    if (f < nif) {
        const double* R = nq + 4 * (size_t)nbr[f];
        UnF += wCR * (R[0] * n0 + R[1] * n1 + R[2] * n2);
        aR = R[3];
    } else {
        UnF = 0.0;
    }
    const double wl = UnF > 0.0 ? 1.0 : 0.0, wr = 1.0 - wl;
    double aF = wl * aL + wr * aR;
    if (aL * (1.0 - aL) < 1.e-15) aF = 0.0;
    if (aR * (1.0 - aR) < 1.e-15) aF = 0.0;
    flux[f] = aF * UnF * area[f];
}

__global__ __launch_bounds__(TPB) void k_cvf_update(int n, int nif, const int* __restrict__ cf_sl,
                                                    const int* __restrict__ cf_ell, const double* __restrict__ flux,
                                                    const double* __restrict__ vol, double ct,
                                                    double* __restrict__ vf, double* __restrict__ part) {
    const int i = xcd_block() * TPB + (int)threadIdx.x;
    double up2 = 0.0;
    if (i < n) {
        double resi = 0.0;
        CELL_FACES_BEGIN(i, cf_sl, cf_ell, f, own)
            if (own) resi -= flux[f];
            else resi += flux[f];
        }
        (void)nif;
        const double up = ct * resi / vol[i];
        double a = vf[i] + up;
        up2 = up * up;
        a = a < 1.0 ? a : 1.0;      // std::min(1.0, a)
        a = 0.0 < a ? a : 0.0;      // std::max(0.0, .)
        vf[i] = a;
    }
    block_sum2(up2, 0.0, part);
}

}  // namespace

void compress_vf_device(semo_ctx* c, double cfl, int iters, double* vf, const double* faceR_b, double* residuals) {
    const int n = c->nc;
    // tau: the utility's host loop (compressVolumeFraction.cpp:121-137)
    std::vector<double> vol(n), area(c->nf);
    c->c_vol.download(vol.data(), n, c->stream);
    c->f_area.download(area.data(), c->nf, c->stream);
    SEMO_HIP(hipStreamSynchronize(c->stream));
    double tau = 1.e15;
    for (int i = 0; i < n; ++i) {
        double maxA = 0.0;
        for (int k = c->h_cf_ptr[i]; k < c->h_cf_ptr[i + 1]; ++k) maxA = std::max(maxA, area[c->h_cf_idx[k]]);
        tau = std::min(tau, std::min(std::pow(vol[i], 0.3), vol[i] / maxA));
    }
    DBuf<double> g, nq, flux, part;
    g.alloc(3 * (size_t)n);
    nq.alloc(4 * (size_t)n);
    flux.alloc(std::max(c->nf, 1));
    const int nb = nblk(n);
    part.alloc(2 * (size_t)nb);
    std::vector<double> hp(2 * (size_t)nb);
    for (int it = 0; it < iters; ++it) {
        lsq1(c, vf, faceR_b, c->nif, g.p);
        SEMO_LAUNCH(c, "cvf_prep", 64.0 * n, k_cvf_prep, dim3(nb), dim3(TPB), n, vf, g.p, nq.p);
        SEMO_LAUNCH(c, "cvf_face", (double)c->nf * 72.0 + 64.0 * n, k_cvf_face, dim3(nblk(c->nf)), dim3(TPB), c->nf,
                    c->nif, c->d_owner.p, c->d_nbr.p, c->f_nrm.p, c->f_wC.p, c->f_area.p, nq.p, flux.p);
        SEMO_LAUNCH(c, "cvf_update", (double)c->nf * 16.0 + 24.0 * n, k_cvf_update, dim3(nb), dim3(TPB), n, c->nif,
                    c->d_cf_sl.p, c->d_cf_ell.p, flux.p, c->c_vol.p, cfl * tau, vf, part.p);
        part.download(hp.data(), hp.size(), c->stream);
        SEMO_HIP(hipStreamSynchronize(c->stream));
        double r2 = 0.0;
        for (int b = 0; b < nb; ++b) r2 += hp[2 * (size_t)b];
        if (residuals) residuals[it] = std::sqrt(r2);
    }
}

}  // namespace semo
