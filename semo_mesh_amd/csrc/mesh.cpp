// Host-side mesh set-up: RCM renumbering, face renumbering, vertex stencil,
// static matrix pattern (sliced ELL) and the face -> matrix-slot maps that let
// the assembly kernels write coefficients in place.
//
// Reference: lib_RCM (src/math/RCM.cpp:34-210) computes the GENRCM ordering
// but never applies it; here it is applied to storage so that owner /
// neighbour gathers and SpMV column accesses stay inside a narrow band.
#include <algorithm>
#include <climits>
#include <cstring>
#include <functional>
#include <numeric>

#include "common.hpp"

namespace semo {

// ------------------------------------------------------------------ GENRCM --
// George & Liu's GENRCM (SPARSPAK), the algorithm lib_RCM calls
// (RCM.cpp:1214-1327 genrcm, 3585-3777 root_find, 2541-2676 level_set,
// 1075-1211 degree, 3353-3582 rcm).  For each connected component, starting
// at the lowest-numbered unvisited node: find a pseudo-peripheral root by
// repeated rooted level structures, number by BFS inserting each node's newly
// reached neighbours in ascending degree (stable), then reverse.
namespace {
struct Levels { std::vector<int> nodes, start; };

// rooted level structure over nodes with live[v] set
void rooted_levels(int n, const int* ptr, const int* adj, int root, std::vector<char>& live,
                   Levels& L) {
    L.nodes.clear(); L.start.clear();
    live[root] = 0;
    L.nodes.push_back(root);
    size_t b = 0, e = 1;
    for (;;) {
        L.start.push_back((int)b);
        for (size_t i = b; i < e; ++i) {
            int v = L.nodes[i];
            for (int k = ptr[v]; k < ptr[v + 1]; ++k)
                if (live[adj[k]]) { live[adj[k]] = 0; L.nodes.push_back(adj[k]); }
        }
        b = e; e = L.nodes.size();
        if (e == b) break;
    }
    L.start.push_back((int)L.nodes.size());
    for (int v : L.nodes) live[v] = 1;
    (void)n;
}

int pseudo_peripheral(int n, const int* ptr, const int* adj, int root, std::vector<char>& live) {
    Levels L;
    rooted_levels(n, ptr, adj, root, live, L);
    int depth = (int)L.start.size() - 1;
    int size = (int)L.nodes.size();
    if (depth == 1 || depth == size) return root;
    for (;;) {
        // lowest-degree node of the last level (first one on ties)
        int first = L.start[depth - 1];
        root = L.nodes[first];
        int best = size;
        if (first < size - 1) {
            for (int j = first; j < size; ++j) {
                int v = L.nodes[j], d = 0;
                for (int k = ptr[v]; k < ptr[v + 1]; ++k) d += live[adj[k]] ? 1 : 0;
                if (d < best) { best = d; root = v; }
            }
        }
        Levels L2;
        rooted_levels(n, ptr, adj, root, live, L2);
        int depth2 = (int)L2.start.size() - 1;
        if (depth2 <= depth) break;
        depth = depth2; L = std::move(L2);
        if (size <= depth) break;
    }
    return root;
}
}  // namespace

void genrcm(int n, const int* ptr, const int* adj, int* perm, int* inv) {
    std::vector<char> live(n, 1), seen(n, 0);
    std::vector<int> deg(n, 0), comp;
    int num = 0;
    for (int s = 0; s < n && num < n; ++s) {
        if (!live[s]) continue;
        int root = pseudo_peripheral(n, ptr, adj, s, live);
        // component and degrees inside the still-live subgraph
        comp.assign(1, root); seen[root] = 1;
        for (size_t i = 0; i < comp.size(); ++i) {
            int v = comp[i], d = 0;
            for (int k = ptr[v]; k < ptr[v + 1]; ++k) {
                int w = adj[k];
                if (!live[w]) continue;
                ++d;
                if (!seen[w]) { seen[w] = 1; comp.push_back(w); }
            }
            deg[v] = d;
        }
        int* out = perm + num;
        int cnt = (int)comp.size();
        live[root] = 0; out[0] = root;
        int head = 0, tail = 1;
        while (head < tail) {
            int lend = tail;
            for (; head < lend; ++head) {
                int v = out[head], first = tail;
                for (int k = ptr[v]; k < ptr[v + 1]; ++k) {
                    int w = adj[k];
                    if (live[w]) { live[w] = 0; out[tail++] = w; }
                }
                // stable ascending-degree insertion of the new segment; as in the
                // reference's translation of SPARSPAK (RCM.cpp:3549-3570, loop
                // `while (fnbr < l)`), the segment's first node is never compared
                // and so stays first
                for (int a = first + 1; a < tail; ++a) {
                    int w = out[a], b = a - 1;
                    while (b > first && deg[out[b]] > deg[w]) { out[b + 1] = out[b]; --b; }
                    out[b + 1] = w;
                }
            }
        }
        std::reverse(out, out + cnt);
        num += cnt;
    }
    for (int i = 0; i < n; ++i) inv[perm[i]] = i;
}

// ------------------------------------------------------------ sliced ELL ---
SellPattern sell_from_csr(int n, int ncols, const std::vector<int>& ptr, const std::vector<int>& col,
                          std::vector<int>* pos) {
    SEMO_CHECK(n >= 0 && ncols >= 1 && (int)ptr.size() == n + 1 && ptr[0] == 0 &&
                   ptr[n] == (int)col.size(), SEMO_EINVAL, "sliced ELL: inconsistent CSR");
    for (int r = 0; r < n; ++r) SEMO_CHECK(ptr[r + 1] >= ptr[r], SEMO_EINVAL, "sliced ELL: row pointers");
    for (int cidx : col) SEMO_CHECK(cidx >= 0 && cidx < ncols, SEMO_EINVAL, "sliced ELL: column out of range");
    SellPattern p;
    p.n = n;
    p.ncols = ncols;
    p.nslices = (n + SL - 1) / SL;
    p.sl_ptr.assign(p.nslices + 1, 0);
    for (int s = 0; s < p.nslices; ++s) {
        int w = 0;
        for (int r = s * SL; r < std::min(n, (s + 1) * SL); ++r) w = std::max(w, ptr[r + 1] - ptr[r]);
        w += w & 1;                            // entry pairs
        p.sl_ptr[s + 1] = p.sl_ptr[s] + w * SL;
    }
    p.col.assign(p.sl_ptr[p.nslices], 0);
    if (pos) pos->assign(col.size(), -1);
    for (int s = 0; s < p.nslices; ++s) {
        int w = (p.sl_ptr[s + 1] - p.sl_ptr[s]) / SL;
        for (int lane = 0; lane < SL; ++lane) {
            int r = s * SL + lane;
            for (int k = 0; k < w; ++k) {
                int e = sell_pos(p.sl_ptr[s], k, lane);
                if (r < n && k < ptr[r + 1] - ptr[r]) {
                    p.col[e] = col[ptr[r] + k];
                    if (pos) (*pos)[ptr[r] + k] = e;
                } else {
                    // padding (value 0) must still index inside the operand
                    // vector: the row itself for square matrices, else column 0
                    p.col[e] = (r < n && r < ncols) ? r : 0;
                }
            }
        }
    }
    p.nnz_offdiag = (int64_t)col.size();
    return p;
}

// --------------------------------------------------------------- mesh set-up
void mesh_setup(semo_ctx* c, const semo_mesh_desc* m, int reorder) {
    SEMO_CHECK(m && m->n_cells > 0 && m->n_faces > 0 && m->n_internal_faces >= 0 &&
                   m->n_internal_faces <= m->n_faces && m->n_points > 0,
               SEMO_EINVAL, "mesh sizes");
    const int nc = m->n_cells, nf = m->n_faces, nif = m->n_internal_faces, np = m->n_points;
    for (int f = 0; f < nf; ++f) {
        SEMO_CHECK(m->owner[f] >= 0 && m->owner[f] < nc, SEMO_EINVAL, "owner out of range");
        int k = m->face_ptr[f + 1] - m->face_ptr[f];
        SEMO_CHECK(k >= 3, SEMO_EINVAL, "faces need at least 3 points");
    }
    for (int f = 0; f < nif; ++f)
        SEMO_CHECK(m->neighbour[f] >= 0 && m->neighbour[f] < nc && m->neighbour[f] != m->owner[f],
                   SEMO_EINVAL, "neighbour out of range");
    for (int k = 0; k < m->face_ptr[nf]; ++k)
        SEMO_CHECK(m->face_pts[k] >= 0 && m->face_pts[k] < np, SEMO_EINVAL, "face point out of range");
    for (int k = 0; k < m->cell_ptr[nc]; ++k)
        SEMO_CHECK(m->cell_pts[k] >= 0 && m->cell_pts[k] < np, SEMO_EINVAL, "cell point out of range");
    {
        int expect = nif;
        for (int b = 0; b < m->n_patches; ++b) {
            SEMO_CHECK(m->patch_start[b] == expect, SEMO_EINVAL, "patches must tile the boundary faces in order");
            expect += m->patch_nfaces[b];
        }
        SEMO_CHECK(expect == nf, SEMO_EINVAL, "patches must cover every boundary face");
    }
    c->np = np; c->nc = nc; c->nf = nf; c->nif = nif; c->nbf = nf - nif; c->npatch = m->n_patches;
    c->reordered = reorder != 0;
    c->ls_ready = false;

    // ---- cell order: lib_RCM adjacency (RCM.cpp:79-126) -> GENRCM
    c->perm.resize(nc); c->inv.resize(nc);
    if (reorder) {
        std::vector<int> cnt(nc + 1, 0);
        for (int f = 0; f < nif; ++f) { cnt[m->owner[f] + 1]++; cnt[m->neighbour[f] + 1]++; }
        for (int i = 0; i < nc; ++i) cnt[i + 1] += cnt[i];
        std::vector<int> adj(cnt[nc]), pos(cnt.begin(), cnt.end() - 1);
        for (int f = 0; f < nif; ++f) {       // ascending face order per cell
            adj[pos[m->owner[f]]++] = m->neighbour[f];
            adj[pos[m->neighbour[f]]++] = m->owner[f];
        }
        genrcm(nc, cnt.data(), adj.data(), c->perm.data(), c->inv.data());
    } else {
        std::iota(c->perm.begin(), c->perm.end(), 0);
        std::iota(c->inv.begin(), c->inv.end(), 0);
    }
    const std::vector<int>& inv = c->inv;

    // ---- faces: internal faces in upper-triangular order of the new numbering,
    //      boundary faces per patch sorted by new owner (stable)
    std::vector<int> fo(nf), fn(nif);
    c->fflip.assign(nf, 0);
    for (int f = 0; f < nf; ++f) fo[f] = inv[m->owner[f]];
    for (int f = 0; f < nif; ++f) {
        int a = fo[f], b = inv[m->neighbour[f]];
        if (a > b) { std::swap(a, b); c->fflip[f] = 1; }
        fo[f] = a; fn[f] = b;
    }
    c->fperm.resize(nf);
    std::iota(c->fperm.begin(), c->fperm.end(), 0);
    if (reorder) {
        std::stable_sort(c->fperm.begin(), c->fperm.begin() + nif, [&](int x, int y) {
            return fo[x] != fo[y] ? fo[x] < fo[y] : fn[x] < fn[y];
        });
        for (int b = 0; b < m->n_patches; ++b) {
            auto first = c->fperm.begin() + m->patch_start[b];
            std::stable_sort(first, first + m->patch_nfaces[b], [&](int x, int y) { return fo[x] < fo[y]; });
        }
    }
    std::vector<int8_t> flip_new(nf);
    c->h_owner.resize(nf); c->h_nbr.resize(nif);
    std::vector<int> fptr(nf + 1, 0), fpts;
    fpts.reserve(m->face_ptr[nf]);
    for (int k = 0; k < nf; ++k) {
        int f = c->fperm[k];
        c->h_owner[k] = fo[f];
        if (k < nif) c->h_nbr[k] = fn[f];
        flip_new[k] = c->fflip[f];
        int b = m->face_ptr[f], e = m->face_ptr[f + 1];
        if (c->fflip[f]) for (int q = e - 1; q >= b; --q) fpts.push_back(m->face_pts[q]);
        else for (int q = b; q < e; ++q) fpts.push_back(m->face_pts[q]);
        fptr[k + 1] = (int)fpts.size();
    }
    c->fflip = flip_new;
    for (int k = 0; k < nif; ++k)
        SEMO_CHECK(c->h_owner[k] < c->h_nbr[k] || !reorder, SEMO_EINVAL, "face order");
    c->patch_start.assign(m->patch_start, m->patch_start + m->n_patches);
    c->patch_nfaces.assign(m->patch_nfaces, m->patch_nfaces + m->n_patches);

    // ---- cell -> faces, ascending (load.cpp:373-385)
    c->h_cf_ptr.assign(nc + 1, 0);
    for (int f = 0; f < nf; ++f) { c->h_cf_ptr[c->h_owner[f] + 1]++; if (f < nif) c->h_cf_ptr[c->h_nbr[f] + 1]++; }
    for (int i = 0; i < nc; ++i) c->h_cf_ptr[i + 1] += c->h_cf_ptr[i];
    c->h_cf_idx.assign(c->h_cf_ptr[nc], 0);
    {
        std::vector<int> pos(c->h_cf_ptr.begin(), c->h_cf_ptr.end() - 1);
        for (int f = 0; f < nf; ++f) {
            c->h_cf_idx[pos[c->h_owner[f]]++] = f;
            if (f < nif) c->h_cf_idx[pos[c->h_nbr[f]]++] = f;
        }
    }

    // ---- cell points in the new order; vertex stencil (geometric.cpp:1786-1841)
    std::vector<int> cptr(nc + 1, 0), cpts;
    cpts.reserve(m->cell_ptr[nc]);
    for (int i = 0; i < nc; ++i) {
        int o = c->perm[i];
        for (int q = m->cell_ptr[o]; q < m->cell_ptr[o + 1]; ++q) cpts.push_back(m->cell_pts[q]);
        cptr[i + 1] = (int)cpts.size();
    }
    std::vector<int> pc_ptr(np + 1, 0), pc_idx(cpts.size());
    for (int p : cpts) pc_ptr[p + 1]++;
    for (int p = 0; p < np; ++p) pc_ptr[p + 1] += pc_ptr[p];
    {
        std::vector<int> pos(pc_ptr.begin(), pc_ptr.end() - 1);
        for (int i = 0; i < nc; ++i)
            for (int q = cptr[i]; q < cptr[i + 1]; ++q) pc_idx[pos[cpts[q]]++] = i;  // ascending cells
    }
    std::vector<int> st_ptr(nc + 1, 0), st_idx;
    st_idx.reserve((size_t)nc * 27);
    {
        std::vector<int> seen(nc, -1);     // seen[j] == i: j is already in cell i's stencil
        for (int i = 0; i < nc; ++i) {
            for (int q = cptr[i]; q < cptr[i + 1]; ++q) {
                int p = cpts[q];
                for (int k = pc_ptr[p]; k < pc_ptr[p + 1]; ++k) {
                    int j = pc_idx[k];
                    if (j == i || seen[j] == i) continue;
                    seen[j] = i;
                    st_idx.push_back(j);
                }
            }
            st_ptr[i + 1] = (int)st_idx.size();
        }
    }

    // ---- matrix pattern: off-diagonal neighbours per row, ascending columns
    std::vector<int> rptr(nc + 1, 0);
    for (int f = 0; f < nif; ++f) { rptr[c->h_owner[f] + 1]++; rptr[c->h_nbr[f] + 1]++; }
    for (int i = 0; i < nc; ++i) rptr[i + 1] += rptr[i];
    std::vector<int> rcol(rptr[nc]), rface(rptr[nc]);
    {
        std::vector<int> pos(rptr.begin(), rptr.end() - 1);
        for (int f = 0; f < nif; ++f) {
            int o = c->h_owner[f], n = c->h_nbr[f];
            rcol[pos[o]] = n; rface[pos[o]++] = f;
            rcol[pos[n]] = o; rface[pos[n]++] = ~f;     // ~f: entry of the neighbour's row
        }
        for (int i = 0; i < nc; ++i) {
            std::vector<std::pair<int, int>> row;
            for (int k = rptr[i]; k < rptr[i + 1]; ++k) row.push_back({rcol[k], rface[k]});
            std::sort(row.begin(), row.end());
            for (size_t k = 0; k < row.size(); ++k) {
                SEMO_CHECK(k == 0 || row[k].first != row[k - 1].first, SEMO_EINVAL,
                           "two faces between the same pair of cells");
                rcol[rptr[i] + k] = row[k].first; rface[rptr[i] + k] = row[k].second;
            }
        }
    }
    std::vector<int> pos_of;
    c->pat = sell_from_csr(nc, nc, rptr, rcol, &pos_of);
    c->h_slot_own.assign(nif, -1); c->h_slot_nbr.assign(nif, -1);
    for (size_t k = 0; k < rcol.size(); ++k) {
        int f = rface[k];
        if (f >= 0) c->h_slot_own[f] = pos_of[k]; else c->h_slot_nbr[~f] = pos_of[k];
    }
    // transpose slots: the slot of (j,i) for every real slot (i,j); padding -> -1
    std::vector<int> tslot(c->pat.col.size(), -1);
    for (int f = 0; f < nif; ++f) { tslot[c->h_slot_own[f]] = c->h_slot_nbr[f]; tslot[c->h_slot_nbr[f]] = c->h_slot_own[f]; }

    // ---- boundary face -> patch
    std::vector<int> bpatch(nf - nif, 0);
    for (int b = 0; b < m->n_patches; ++b)
        for (int f = m->patch_start[b]; f < m->patch_start[b] + m->patch_nfaces[b]; ++f) bpatch[f - nif] = b;

    // ---- uploads
    hipStream_t s = c->stream;
    c->d_pts.upload(m->points, 3 * (size_t)np, s);
    c->d_owner.upload(c->h_owner, s);
    c->d_nbr.upload(c->h_nbr, s);
    c->d_fptr.upload(fptr, s);
    c->d_fpts.upload(fpts, s);
    c->d_cf_ptr.upload(c->h_cf_ptr, s);
    c->d_cf_idx.upload(c->h_cf_idx, s);
    c->d_st_ptr.upload(st_ptr, s);
    c->d_st_idx.upload(st_idx, s);
    // sliced-ELL copies of the stencil and of cell -> faces for coalesced
    // gathers (same per-cell order as the CSR lists)
    auto to_sliced = [&](const std::vector<int>& ptr, const std::vector<int>& idx, int pad,
                         const std::function<int(int, int)>& enc, std::vector<int>& sl,
                         std::vector<int>& ell) {
        int ns = (nc + SL - 1) / SL;
        sl.assign(ns + 1, 0);
        for (int sl_i = 0; sl_i < ns; ++sl_i) {
            int w = 0;
            for (int r = sl_i * SL; r < std::min(nc, (sl_i + 1) * SL); ++r) w = std::max(w, ptr[r + 1] - ptr[r]);
            sl[sl_i + 1] = sl[sl_i] + w * SL;
        }
        ell.assign(sl[ns], pad);
        for (int r = 0; r < nc; ++r) {
            int sl_i = r / SL, lane = r % SL;
            for (int k = ptr[r]; k < ptr[r + 1]; ++k) ell[sl[sl_i] + (k - ptr[r]) * SL + lane] = enc(r, idx[k]);
        }
    };
    {
        std::vector<int> sl, ell;
        to_sliced(st_ptr, st_idx, -1, [](int, int j) { return j; }, sl, ell);
        c->d_st_sl.upload(sl, s); c->d_st_ell.upload(ell, s);
        {
            std::vector<int16_t> dl(ell.size());
            bool fits = true;
            for (int sl_i = 0; sl_i + 1 < (int)sl.size() && fits; ++sl_i)
                for (int e = sl[sl_i]; e < sl[sl_i + 1]; ++e) {
                    int r = sl_i * SL + ((e - sl[sl_i]) & (SL - 1));
                    long d = ell[e] < 0 ? -32768 : (long)ell[e] - r;
                    if (ell[e] >= 0 && (d < -32767 || d > 32767)) { fits = false; break; }
                    dl[e] = (int16_t)d;
                }
            c->d_st_dl.release();
            if (fits) c->d_st_dl.upload(dl, s);
        }
        // lanes per cell of the one-field LSQ pass: the setting, or by
        // default 4 for long stencils (tets: about 68 neighbours, 9 % faster
        // on the 1.3 M-tet mesh) and 1 for short ones (hexes: 26, where 4
        // lanes measured 56 % slower)
        c->lsq_lanes_eff = c->lsq_lanes > 0 ? c->lsq_lanes
                                            : ((double)st_idx.size() >= LSQ_LANES_LONG_STENCIL * nc ? 4 : 1);
        SEMO_HIP(hipStreamSynchronize(s));
        to_sliced(c->h_cf_ptr, c->h_cf_idx, INT32_MIN,
                  [&](int r, int f) { return c->h_owner[f] == r ? f : ~f; }, sl, ell);
        c->d_cf_sl.upload(sl, s); c->d_cf_ell.upload(ell, s);
        std::vector<int> bptr(nc + 1, 0), bidx;
        for (int r = 0; r < nc; ++r) {
            for (int k = c->h_cf_ptr[r]; k < c->h_cf_ptr[r + 1]; ++k)
                if (c->h_cf_idx[k] >= nif) bidx.push_back(c->h_cf_idx[k]);
            bptr[r + 1] = (int)bidx.size();
        }
        c->d_bf_ptr.upload(bptr, s); c->d_bf_idx.upload(bidx, s);
        SEMO_HIP(hipStreamSynchronize(s));
    }
    c->d_bpatch.upload(bpatch, s);
    c->d_slot_own.upload(c->h_slot_own, s);
    c->d_slot_nbr.upload(c->h_slot_nbr, s);
    c->d_tslot.upload(tslot, s);
    c->Amat.n = nc; c->Amat.ncols = nc; c->Amat.nslices = c->pat.nslices; c->Amat.nnz = nc + c->pat.nnz_offdiag;
    c->Amat.sl_ptr.upload(c->pat.sl_ptr, s);
    c->Amat.col.upload(c->pat.col, s);
    sell_narrow(c->pat, c->Amat.dcol, s);
    c->Amat.val.alloc(c->pat.col.size());
    c->Amat.val.zero(s);
    c->Amat.diag.alloc(nc);
    c->d_flags.alloc(1);
    c->d_flags.zero(s);
    SEMO_HIP(hipStreamSynchronize(s));
    c->has_mesh = true;
}

}  // namespace semo
