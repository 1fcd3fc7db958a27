// Host set-up of a smoothed-aggregation AMG hierarchy (symbolic and numeric),
// uploaded as sliced-ELL levels for the device V-cycle.
//
// This replaces the set-up half of AMGCL's mpi::amg<smoothed_aggregation,...>
// that solveAMGCL builds on every call (src/solvers/solveAMGCL.cpp:938-980;
// lib/amgcl/mpi/coarsening/smoothed_aggregation.hpp:100-190): strength of
// connection halved per level (AMGCL's default start is 0.08; this set-up
// starts at 0.06, which measured one to two fewer FCG iterations on the
// 128^3 cavity and the dam break), plain aggregation,
// tentative piecewise-constant prolongation smoothed by one damped-Jacobi step
// on the filtered operator, Galerkin coarse operators, down to a dense
// coarsest level.  The hierarchy is rebuilt only when the fine operator's
// values change (the pressure operator of a constant-density flow on a fixed
// mesh does not), whereas SEMO rebuilds it on every call.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>

#include <cstdlib>

#include "common.hpp"

namespace semo {

struct HCsr {
    int n = 0, m = 0;                 // rows, cols
    std::vector<int> ptr, col;
    std::vector<double> val;
};

namespace {

HCsr transpose(const HCsr& A) {
    HCsr T; T.n = A.m; T.m = A.n;
    T.ptr.assign(T.n + 1, 0);
    for (int c : A.col) T.ptr[c + 1]++;
    for (int i = 0; i < T.n; ++i) T.ptr[i + 1] += T.ptr[i];
    T.col.resize(A.col.size()); T.val.resize(A.col.size());
    std::vector<int> pos(T.ptr.begin(), T.ptr.end() - 1);
    for (int i = 0; i < A.n; ++i)
        for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
            int d = pos[A.col[k]]++;
            T.col[d] = i; T.val[d] = A.val[k];
        }
    return T;
}

// C = A * B, rows sorted by column
HCsr spgemm(const HCsr& A, const HCsr& B) {
    HCsr C; C.n = A.n; C.m = B.m;
    C.ptr.assign(A.n + 1, 0);
    std::vector<int> mark(B.m, -1);
    std::vector<double> acc(B.m, 0.0);
    std::vector<int> cols;
    for (int i = 0; i < A.n; ++i) {
        cols.clear();
        for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
            int j = A.col[k]; double a = A.val[k];
            for (int q = B.ptr[j]; q < B.ptr[j + 1]; ++q) {
                int c = B.col[q];
                if (mark[c] != i) { mark[c] = i; acc[c] = 0.0; cols.push_back(c); }
                acc[c] += a * B.val[q];
            }
        }
        std::sort(cols.begin(), cols.end());
        for (int c : cols) { C.col.push_back(c); C.val.push_back(acc[c]); }
        C.ptr[i + 1] = (int)C.col.size();
    }
    return C;
}

double diag_of(const HCsr& A, int i) {
    for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) if (A.col[k] == i) return A.val[k];
    return 0.0;
}

// spectral radius of D^-1 A by power iteration (symmetric scaling)
double lambda_max(const HCsr& A, const std::vector<double>& dinv) {
    int n = A.n;
    std::vector<double> x(n), y(n);
    for (int i = 0; i < n; ++i) x[i] = 1.0 + 0.1 * std::sin(1.0 + i);
    double lam = 1.0;
    for (int it = 0; it < 20; ++it) {
        double nx = 0;
        for (double v : x) nx += v * v;
        nx = std::sqrt(nx);
        for (double& v : x) v /= nx;
        for (int i = 0; i < n; ++i) {
            double s = 0;
            for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) s += A.val[k] * x[A.col[k]];
            y[i] = dinv[i] * s;
        }
        double num = 0;
        for (int i = 0; i < n; ++i) num += x[i] * y[i];
        lam = num;
        std::swap(x, y);
    }
    return std::fabs(lam);
}

// plain aggregation on the strength graph (AMGCL coarsening/plain_aggregates)
int aggregate(const HCsr& A, double eps, std::vector<int>& agg) {
    int n = A.n;
    std::vector<double> d(n);
    for (int i = 0; i < n; ++i) d[i] = std::fabs(diag_of(A, i));
    auto strong = [&](int i, int k) {
        int j = A.col[k];
        if (j == i) return false;
        double a = A.val[k];
        return a * a > eps * eps * d[i] * d[j];
    };
    const int UNDONE = -1, REMOVED = -2;
    agg.assign(n, UNDONE);
    for (int i = 0; i < n; ++i) {           // isolated rows are not aggregated
        bool any = false;
        for (int k = A.ptr[i]; k < A.ptr[i + 1] && !any; ++k) any = strong(i, k);
        if (!any) agg[i] = REMOVED;
    }
    int na = 0;
    // pass 1: a node whose strong neighbours are all free seeds an aggregate
    for (int i = 0; i < n; ++i) {
        if (agg[i] != UNDONE) continue;
        bool free_nb = true;
        for (int k = A.ptr[i]; k < A.ptr[i + 1] && free_nb; ++k)
            if (strong(i, k) && agg[A.col[k]] != UNDONE && agg[A.col[k]] != REMOVED) free_nb = false;
        if (!free_nb) continue;
        int id = na++;
        agg[i] = id;
        for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
            if (strong(i, k) && agg[A.col[k]] == UNDONE) agg[A.col[k]] = id;
    }
    // pass 2: leftovers join the aggregate of their strongest aggregated neighbour
    std::vector<int> tmp = agg;
    for (int i = 0; i < n; ++i) {
        if (agg[i] != UNDONE) continue;
        double best = -1; int b = -1;
        for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
            if (strong(i, k) && agg[A.col[k]] >= 0 && std::fabs(A.val[k]) > best) {
                best = std::fabs(A.val[k]); b = agg[A.col[k]];
            }
        tmp[i] = b;
    }
    agg = tmp;
    // pass 3: whatever is left forms its own aggregates with free neighbours
    for (int i = 0; i < n; ++i) {
        if (agg[i] != UNDONE) continue;
        int id = na++;
        agg[i] = id;
        for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
            if (strong(i, k) && agg[A.col[k]] == UNDONE) agg[A.col[k]] = id;
    }
    // removed (isolated) rows become singleton aggregates
    for (int i = 0; i < n; ++i) if (agg[i] == REMOVED) agg[i] = na++;
    return na;
}

void split_diag(const HCsr& A, std::vector<int>& ptr, std::vector<int>& col, std::vector<double>& val,
                std::vector<double>& diag) {
    ptr.assign(A.n + 1, 0); col.clear(); val.clear(); diag.assign(A.n, 0.0);
    for (int i = 0; i < A.n; ++i) {
        for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
            if (A.col[k] == i && A.n == A.m) { diag[i] += A.val[k]; continue; }
            col.push_back(A.col[k]); val.push_back(A.val[k]);
        }
        ptr[i + 1] = (int)col.size();
    }
}

// Device format of an operator: sliced ELL (one lane per row) for short rows
// of similar length; CSR with `lanes` lanes per row (about four entries per
// lane) for long rows (restriction, Galerkin coarse operators) or rows whose
// lengths vary so much that ELL padding would waste more than half the reads.
// (Average 10 and four entries per lane measured best against averages up
// to 64 for ELL and two or eight entries per lane.)
constexpr double ELL_MAX_AVG = 10.0, ENTRIES_PER_LANE = 4.0;
// SEMO_GPU_ELL_MAX_AVG overrides the average row length up to which sliced
// ELL is chosen (a tuning knob; the default is the measured best)
double ell_max_avg() {
    static const double v = [] {
        const char* e = std::getenv("SEMO_GPU_ELL_MAX_AVG");
        return e ? std::atof(e) : ELL_MAX_AVG;
    }();
    return v;
}
int choose_lanes(int n, const std::vector<int>& ptr, size_t padded) {
    size_t nnz = ptr.empty() ? 0 : (size_t)ptr[n];
    double avg = n ? (double)nnz / n : 0.0;
    if (avg <= ell_max_avg() && (double)padded <= 1.5 * (double)std::max<size_t>(nnz, 1)) return 0;
    int lanes = 4;
    while (lanes < 64 && ENTRIES_PER_LANE * lanes < avg) lanes *= 2;
    return lanes;
}

// epos (optional): device element index of every off-diagonal entry, in the
// order split_diag visits them
template <class T>
void upload_csr(DevSell<T>& M, const HCsr& A, bool square, hipStream_t s, std::vector<int>* epos = nullptr) {
    std::vector<int> ptr, col; std::vector<double> val, diag;
    if (square) split_diag(A, ptr, col, val, diag);
    else { ptr = A.ptr; col = A.col; val = A.val; }
    std::vector<int> pos;
    SellPattern p = sell_from_csr(A.n, A.m, ptr, col, &pos);
    int lanes = choose_lanes(A.n, ptr, p.col.size());
    if (epos) {
        epos->resize(col.size());
        for (size_t k = 0; k < col.size(); ++k) (*epos)[k] = lanes ? (int)k : pos[k];
    }
    if (lanes) {
        csr_upload<T>(M, A.n, A.m, ptr, col, val, diag, lanes, s);
    } else {
        std::vector<double> pv(p.col.size(), 0.0);
        for (size_t k = 0; k < col.size(); ++k) pv[pos[k]] = val[k];
        sell_upload<T>(M, p, pv, diag, s);
    }
    M.nnz = (int64_t)A.col.size();
}

}  // namespace

// src_slot (optional): for every entry of A0, the slot of the device system
// it came from (-1 on the diagonal); lets level 0 be refreshed on the device
void amg_setup_host(semo_ctx* c, Amg& amg, const HCsr& A0, const std::vector<int>* src_slot = nullptr) {
    amg.drop_graph();             // the level buffers are reallocated below
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = c->stream;
    amg.lv.clear();
    int coarse = c->sp.coarse_size > 0 ? c->sp.coarse_size : 1000;
    HCsr A = A0;
    double eps = 0.06;
    for (int level = 0; level < 30; ++level) {
        amg.lv.emplace_back();
        AmgLevel& L = amg.lv.back();
        L.n = A.n;
        std::vector<double> dinv(A.n);
        for (int i = 0; i < A.n; ++i) {
            double d = diag_of(A, i);
            SEMO_CHECK(d != 0.0, SEMO_EZERO_PIVOT, "zero diagonal in the AMG operator");
            dinv[i] = 1.0 / d;
        }
        std::vector<int> epos;
        upload_csr<double>(L.A64, A, true, s, level == 0 && src_slot ? &epos : nullptr);
        upload_csr<float>(L.A32, A, true, s);
        if (level == 0 && src_slot) {
            // device element -> source slot, for refreshing level 0's values
            std::vector<int> map(L.A64.val.n, -1);
            size_t k = 0;
            for (int i = 0; i < A.n; ++i)
                for (int q = A.ptr[i]; q < A.ptr[i + 1]; ++q)
                    if (!(A.col[q] == i && A.n == A.m)) map[epos[k++]] = (*src_slot)[q];
            amg.l0_map.upload(map, s);
            SEMO_HIP(hipStreamSynchronize(s));
        } else if (level == 0) {
            amg.l0_map.release();
        }
        L.dinv64.upload(dinv, s);
        std::vector<float> df(dinv.begin(), dinv.end());
        L.dinv32.upload(df, s);
        SEMO_HIP(hipStreamSynchronize(s));      // df / dinv are temporaries
        L.lmax = lambda_max(A, dinv);
        for (auto* b : {&L.x64, &L.b64, &L.r64, &L.t64, &L.u64, &L.w64}) b->alloc(A.n + 64);
        for (auto* b : {&L.x32, &L.b32, &L.r32, &L.t32, &L.u32, &L.w32}) b->alloc(A.n + 64);
        if (A.n <= coarse) break;

        std::vector<int> agg;
        int na = aggregate(A, eps, agg);
        if (na >= A.n || na == 0) break;       // no coarsening possible
        eps *= 0.5;
        // filtered operator: weak connections lumped into the diagonal
        std::vector<double> dF(A.n, 0.0);
        std::vector<double> dd(A.n);
        for (int i = 0; i < A.n; ++i) dd[i] = std::fabs(diag_of(A, i));
        HCsr AF; AF.n = A.n; AF.m = A.m; AF.ptr.assign(A.n + 1, 0);
        for (int i = 0; i < A.n; ++i) {
            double lump = 0.0;
            for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
                int j = A.col[k];
                double a = A.val[k];
                if (j == i) continue;
                if (a * a > eps * 2 * eps * 2 * dd[i] * dd[j]) { AF.col.push_back(j); AF.val.push_back(a); }
                else lump += a;
            }
            dF[i] = diag_of(A, i) + lump;
            AF.col.push_back(i); AF.val.push_back(dF[i]);
            AF.ptr[i + 1] = (int)AF.col.size();
        }
        std::vector<double> dFinv(A.n);
        for (int i = 0; i < A.n; ++i) dFinv[i] = dF[i] != 0.0 ? 1.0 / dF[i] : 0.0;
        double omega = (4.0 / 3.0) / std::max(lambda_max(AF, dFinv), 1e-30);
        // P = (I - omega D_F^-1 A_F) P_tent
        HCsr P; P.n = A.n; P.m = na; P.ptr.assign(A.n + 1, 0);
        {
            std::vector<int> mark(na, -1); std::vector<double> acc(na); std::vector<int> cols;
            for (int i = 0; i < A.n; ++i) {
                cols.clear();
                auto add = [&](int cc, double v) {
                    if (mark[cc] != i) { mark[cc] = i; acc[cc] = 0.0; cols.push_back(cc); }
                    acc[cc] += v;
                };
                add(agg[i], 1.0);
                for (int k = AF.ptr[i]; k < AF.ptr[i + 1]; ++k)
                    add(agg[AF.col[k]], -omega * dFinv[i] * AF.val[k]);
                std::sort(cols.begin(), cols.end());
                for (int cc : cols) if (acc[cc] != 0.0) { P.col.push_back(cc); P.val.push_back(acc[cc]); }
                P.ptr[i + 1] = (int)P.col.size();
            }
        }
        HCsr R = transpose(P);
        HCsr AP = spgemm(A, P);
        HCsr Ac = spgemm(R, AP);
        upload_csr<double>(L.P64, P, false, s);
        upload_csr<double>(L.R64, R, false, s);
        upload_csr<float>(L.P32, P, false, s);
        upload_csr<float>(L.R32, R, false, s);
        A = std::move(Ac);
    }
    // dense inverse of the coarsest operator (Gauss-Jordan, partial pivoting)
    {
        int n = A.n;
        std::vector<double> M((size_t)n * n, 0.0), I((size_t)n * n, 0.0);
        for (int i = 0; i < n; ++i) {
            for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) M[(size_t)i * n + A.col[k]] += A.val[k];
            I[(size_t)i * n + i] = 1.0;
        }
        for (int k = 0; k < n; ++k) {
            int pr = k;
            for (int i = k + 1; i < n; ++i)
                if (std::fabs(M[(size_t)i * n + k]) > std::fabs(M[(size_t)pr * n + k])) pr = i;
            SEMO_CHECK(std::fabs(M[(size_t)pr * n + k]) > 0.0, SEMO_EZERO_PIVOT,
                       "singular coarsest AMG operator");
            if (pr != k)
                for (int j = 0; j < n; ++j) {
                    std::swap(M[(size_t)k * n + j], M[(size_t)pr * n + j]);
                    std::swap(I[(size_t)k * n + j], I[(size_t)pr * n + j]);
                }
            double piv = 1.0 / M[(size_t)k * n + k];
            for (int j = 0; j < n; ++j) { M[(size_t)k * n + j] *= piv; I[(size_t)k * n + j] *= piv; }
            for (int i = 0; i < n; ++i) {
                if (i == k) continue;
                double f = M[(size_t)i * n + k];
                if (f == 0.0) continue;
                for (int j = 0; j < n; ++j) {
                    M[(size_t)i * n + j] -= f * M[(size_t)k * n + j];
                    I[(size_t)i * n + j] -= f * I[(size_t)k * n + j];
                }
            }
        }
        amg.nc_dense = n;
        amg.coarse_inv.upload(I, s);
    }
    SEMO_HIP(hipStreamSynchronize(s));
    amg.setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    amg.built = true;
}

// fine operator from the device sliced-ELL system
void amg_build(semo_ctx* c, Amg& amg, const DevSell<double>& Ad) {
    const SellPattern& p = c->pat;
    std::vector<double> val(p.col.size()), diag(c->nc);
    Ad.val.download(val.data(), val.size(), c->stream);
    Ad.diag.download(diag.data(), diag.size(), c->stream);
    SEMO_HIP(hipStreamSynchronize(c->stream));
    // same values as the last build -> keep the hierarchy
    if (amg.built && amg.key.size() == val.size() + diag.size() &&
        std::equal(val.begin(), val.end(), amg.key.begin()) &&
        std::equal(diag.begin(), diag.end(), amg.key.begin() + val.size()))
        return;
    HCsr A; A.n = A.m = c->nc; A.ptr.assign(c->nc + 1, 0);
    std::vector<int> slot;
    for (int s = 0; s < p.nslices; ++s) {
        int w = (p.sl_ptr[s + 1] - p.sl_ptr[s]) / SL;
        for (int lane = 0; lane < SL; ++lane) {
            int r = s * SL + lane;
            if (r >= c->nc) break;
            std::vector<std::pair<int, std::pair<double, int>>> row;
            row.push_back({r, {diag[r], -1}});
            for (int k = 0; k < w; ++k) {
                int e = sell_pos(p.sl_ptr[s], k, lane);
                if (p.col[e] == r) continue;     // padding
                row.push_back({p.col[e], {val[e], e}});
            }
            std::sort(row.begin(), row.end());
            for (auto& q : row) { A.col.push_back(q.first); A.val.push_back(q.second.first); slot.push_back(q.second.second); }
            A.ptr[r + 1] = (int)A.col.size();
        }
    }
    amg_setup_host(c, amg, A, &slot);
    amg.key = val;
    amg.key.insert(amg.key.end(), diag.begin(), diag.end());
}

// Aggregation hierarchy of a scalar operator (CSR, diagonal included) for the
// block AMG of the coupled system: per level the fine -> aggregate map, from
// plain aggregation on the strength graph (eps halved per level) and the
// unsmoothed Galerkin product P^T A P (sums of the entries between two
// aggregates) as the next level's scalar operator.  Stops when a level has
// at most min_coarse rows or aggregation no longer coarsens.
std::vector<std::vector<int>> aggregate_hierarchy(int n, const std::vector<int>& ptr, const std::vector<int>& col,
                                                  const std::vector<double>& val, int min_coarse) {
    std::vector<std::vector<int>> maps;
    HCsr A; A.n = A.m = n; A.ptr = ptr; A.col = col; A.val = val;
    double eps = 0.08;
    while (A.n > min_coarse) {
        std::vector<int> agg;
        int na = aggregate(A, eps, agg);
        if (na >= A.n || na == 0 || na > (A.n * 9) / 10) break;
        // coarse scalar operator: entry (agg[i], agg[j]) sums A(i, j)
        HCsr Ac; Ac.n = Ac.m = na; Ac.ptr.assign(na + 1, 0);
        std::vector<std::vector<int>> members(na);
        for (int i = 0; i < A.n; ++i) members[agg[i]].push_back(i);
        std::vector<int> mark(na, -1);
        std::vector<double> acc(na, 0.0);
        std::vector<int> cols;
        for (int I = 0; I < na; ++I) {
            cols.clear();
            for (int i : members[I])
                for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
                    int J = agg[A.col[k]];
                    if (mark[J] != I) { mark[J] = I; acc[J] = 0.0; cols.push_back(J); }
                    acc[J] += A.val[k];
                }
            std::sort(cols.begin(), cols.end());
            for (int J : cols) { Ac.col.push_back(J); Ac.val.push_back(acc[J]); }
            Ac.ptr[I + 1] = (int)Ac.col.size();
        }
        maps.push_back(std::move(agg));
        A = std::move(Ac);
        eps *= 0.5;
    }
    return maps;
}

// standalone system (COO in caller numbering) -> hierarchy
void amg_build_csr(semo_ctx* c, Amg& amg, int n, const std::vector<int>& ptr,
                   const std::vector<int>& col, const std::vector<double>& val) {
    HCsr A; A.n = A.m = n; A.ptr = ptr; A.col = col; A.val = val;
    amg_setup_host(c, amg, A);
    amg.key.clear();
}

}  // namespace semo
