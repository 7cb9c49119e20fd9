// Symbolic analysis of the multifrontal LU (host, once per grid size).
//
// Replaces the ordering + symbolic phase that Eigen's PardisoLU::analyzePattern
// (h_mec_rsf_v9.cpp:1050) runs through PARDISO/METIS every Picard iteration.
// The pattern of the coupled system depends only on (Nx, Ny): every row of a
// staggered node couples to nodes at most one step away (axis neighbours plus
// one anti-diagonal pair, h_mec_rsf_v9.cpp:761-765, :833-837), so a full row
// or column of nodes separates the grid.  Geometric nested dissection cuts the
// node rectangle through its longer side recursively; each separator is one
// dense front whose pivots are the separator's unknowns per node (6 for the
// coupled system, 4 after the Darcy condensation), ordered velocities first
// (vx, vy[, vxD, vyD], Pt, Pf) so that every pressure pivot comes after the
// velocities it couples to.  The front's update set is the
// set of ancestor nodes adjacent to its region (exact for connected regions).
#include <algorithm>
#include <cstdio>
#include <functional>

#include "hmec_internal.h"

namespace {
// position of each variable inside a node block of a front:
// coupled system (0 Pt, 1 vxs, 2 vys, 3 Pf, 4 vxD, 5 vyD), condensed (0 Pt,
// 1 vxs, 2 vys, 3 Pf)
const int kVarPos6[6] = {4, 0, 1, 5, 2, 3};
const int kVarPos4[4] = {2, 0, 1, 3};
const int kLeafNodes = 4;

struct Rect {
    int i0, i1, j0, j1;
};
}  // namespace

int hmec_var_pos(int dpn, int v) { return dpn == 4 ? kVarPos4[v] : kVarPos6[v]; }

int hmec_plan_build(int Nx1, int Ny1, const std::vector<int32_t> &rowp,
                    const std::vector<int32_t> &coli, FrontPlan &pl, int dpn) {
    if (dpn != 4 && dpn != 6) return HMEC_E_ARG;
    const int *kVarPos = dpn == 4 ? kVarPos4 : kVarPos6;
    const int nn = Nx1 * Ny1;
    const int64_t N = (int64_t)nn * dpn;
    if ((int64_t)rowp.size() != N + 1) return HMEC_E_ARG;
    // symmetrized node graph
    std::vector<std::vector<int>> adj(nn);
    for (int64_t r = 0; r < N; r++)
        for (int p = rowp[r]; p < rowp[r + 1]; p++) {
            const int a = (int)(r / dpn), b = coli[p] / dpn;
            if (a == b) continue;
            adj[a].push_back(b);
            adj[b].push_back(a);
        }
    for (auto &v : adj) {
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    // nested dissection
    pl = FrontPlan();
    pl.dpn = dpn;
    std::vector<Rect> region;
    std::function<int(Rect)> build = [&](Rect R) -> int {
        const int h = R.i1 - R.i0, w = R.j1 - R.j0;
        if (h <= 0 || w <= 0) return -1;
        const int f = (int)pl.nodes.size();
        pl.nodes.emplace_back();
        pl.children.emplace_back();
        region.push_back(R);
        std::vector<int> piv;
        std::vector<int> kids;
        if (h * w <= kLeafNodes) {
            for (int j = R.j0; j < R.j1; j++)
                for (int i = R.i0; i < R.i1; i++) piv.push_back(j * Ny1 + i);
        } else if (w >= h) {
            const int jm = (R.j0 + R.j1) / 2;
            int a = build(Rect{R.i0, R.i1, R.j0, jm});
            int b = build(Rect{R.i0, R.i1, jm + 1, R.j1});
            if (a >= 0) kids.push_back(a);
            if (b >= 0) kids.push_back(b);
            for (int i = R.i0; i < R.i1; i++) piv.push_back(jm * Ny1 + i);
        } else {
            const int im = (R.i0 + R.i1) / 2;
            int a = build(Rect{R.i0, im, R.j0, R.j1});
            int b = build(Rect{im + 1, R.i1, R.j0, R.j1});
            if (a >= 0) kids.push_back(a);
            if (b >= 0) kids.push_back(b);
            for (int j = R.j0; j < R.j1; j++) piv.push_back(j * Ny1 + im);
        }
        pl.nodes[f] = piv;
        pl.children[f] = kids;
        return f;
    };
    const int root = build(Rect{0, Ny1, 0, Nx1});
    const int nf = (int)pl.nodes.size();
    pl.nfronts = nf;
    pl.parent.assign(nf, -1);
    for (int f = 0; f < nf; f++)
        for (int k : pl.children[f]) pl.parent[k] = f;
    pl.np.assign(nf, 0);
    for (int f = 0; f < nf; f++) pl.np[f] = (int)pl.nodes[f].size();
    // postorder -> node elimination positions
    std::vector<int> post;
    post.reserve(nf);
    std::function<void(int)> dfs = [&](int f) {
        for (int k : pl.children[f]) dfs(k);
        post.push_back(f);
    };
    dfs(root);
    std::vector<int> node_pos(nn, -1);
    pl.node_front.assign(nn, -1);
    pl.node_local.assign(nn, -1);
    int cnt = 0;
    for (int f : post)
        for (int l = 0; l < pl.np[f]; l++) {
            const int v = pl.nodes[f][l];
            node_pos[v] = cnt++;
            pl.node_front[v] = f;
            pl.node_local[v] = l;
        }
    if (cnt != nn) {
        hmec_set_last_error("nested dissection did not cover the grid");
        return HMEC_E_ARG;
    }
    // update sets: ancestor nodes adjacent to the front's region
    std::vector<int> stamp(nn, -1);
    pl.nu.assign(nf, 0);
    for (int f = 0; f < nf; f++) {
        const Rect R = region[f];
        std::vector<int> upd;
        for (int j = R.j0; j < R.j1; j++)
            for (int i = R.i0; i < R.i1; i++) {
                // only border nodes of the rectangle can have outside neighbours
                if (i != R.i0 && i != R.i1 - 1 && j != R.j0 && j != R.j1 - 1) continue;
                for (int b : adj[j * Ny1 + i]) {
                    const int bi = b % Ny1, bj = b / Ny1;
                    if (bi >= R.i0 && bi < R.i1 && bj >= R.j0 && bj < R.j1) continue;
                    if (stamp[b] == f) continue;
                    stamp[b] = f;
                    upd.push_back(b);
                }
            }
        std::sort(upd.begin(), upd.end(), [&](int a, int b) { return node_pos[a] < node_pos[b]; });
        pl.nu[f] = (int)upd.size();
        pl.nodes[f].insert(pl.nodes[f].end(), upd.begin(), upd.end());
    }
    // levels
    pl.level.assign(nf, 0);
    for (int f : post) {
        int lv = 0;
        for (int k : pl.children[f]) lv = std::max(lv, pl.level[k] + 1);
        pl.level[f] = lv;
    }
    pl.nlevels = 0;
    for (int f = 0; f < nf; f++) pl.nlevels = std::max(pl.nlevels, pl.level[f] + 1);
    pl.levels.assign(pl.nlevels, {});
    for (int f : post) pl.levels[pl.level[f]].push_back(f);
    // arena offsets (level-major so a level's fronts are contiguous)
    pl.off.assign(nf, 0);
    pl.voff.assign(nf, 0);
    int64_t a = 0, va = 0;
    pl.max_front = 0;
    pl.flops = 0;
    pl.factor_nnz = 0;
    for (int lv = 0; lv < pl.nlevels; lv++)
        for (int f : pl.levels[lv]) {
            const int64_t n = (int64_t)dpn * (pl.np[f] + pl.nu[f]), p = (int64_t)dpn * pl.np[f], u = n - p;
            pl.off[f] = a;
            a += n * n;
            pl.voff[f] = va;
            va += n;
            pl.max_front = std::max<int>(pl.max_front, (int)n);
            for (int64_t k = 0; k < p; k++) {
                const double r = (double)(n - k - 1);
                pl.flops += r + 2.0 * r * r;
            }
            pl.factor_nnz += n * n - u * u;
        }
    pl.arena = a;
    pl.varena = va;
    // CSR entry -> front position.  The entry belongs to the front that
    // eliminates the earlier of its two nodes.
    const int64_t nnz = rowp[N];
    std::vector<int32_t> owner(nnz);
    std::vector<int64_t> bucket_cnt(nf + 1, 0);
    for (int64_t r = 0; r < N; r++)
        for (int p = rowp[r]; p < rowp[r + 1]; p++) {
            const int nr = (int)(r / dpn), nc = coli[p] / dpn;
            const int first = node_pos[nr] <= node_pos[nc] ? nr : nc;
            owner[p] = pl.node_front[first];
            bucket_cnt[owner[p] + 1]++;
        }
    for (int f = 0; f < nf; f++) bucket_cnt[f + 1] += bucket_cnt[f];
    std::vector<int64_t> bucket(nnz), fill(bucket_cnt.begin(), bucket_cnt.end() - 1);
    std::vector<int32_t> row_of(nnz);
    for (int64_t r = 0; r < N; r++)
        for (int p = rowp[r]; p < rowp[r + 1]; p++) {
            bucket[fill[owner[p]]++] = p;
            row_of[p] = (int32_t)r;
        }
    pl.nz_dst.assign(nnz, -1);
    // the same entries per front column, rows ascending (the front pull reads
    // a column's original entries with the column)
    pl.col_base.assign(nf + 1, 0);
    for (int f = 0; f < nf; f++) pl.col_base[f + 1] = pl.col_base[f] + (int64_t)dpn * (pl.np[f] + pl.nu[f]);
    pl.col_ptr.assign(pl.col_base[nf] + 1, 0);
    pl.col_row.assign(nnz, -1);
    pl.col_nz.assign(nnz, -1);
    std::vector<int> loc(nn, -1);
    std::vector<std::pair<int64_t, int64_t>> ent;   // (column * n + row, CSR entry) of one front
    int64_t ne = 0;
    for (int f = 0; f < nf; f++) {
        const auto &nl = pl.nodes[f];
        for (int l = 0; l < (int)nl.size(); l++) loc[nl[l]] = l;
        const int64_t n = (int64_t)dpn * (int64_t)nl.size();
        ent.clear();
        for (int64_t q = bucket_cnt[f]; q < bucket_cnt[f + 1]; q++) {
            const int64_t p = bucket[q];
            const int r = row_of[p], cc = coli[p];
            const int lr = loc[r / dpn], lc = loc[cc / dpn];
            if (lr < 0 || lc < 0) {
                hmec_set_last_error("CSR entry outside its front");
                return HMEC_E_ARG;
            }
            const int64_t rr = (int64_t)dpn * lr + kVarPos[r % dpn], ccol = (int64_t)dpn * lc + kVarPos[cc % dpn];
            pl.nz_dst[p] = pl.off[f] + ccol * n + rr;
            ent.emplace_back(ccol * n + rr, p);
        }
        std::sort(ent.begin(), ent.end());
        size_t k = 0;
        for (int64_t c = 0; c < n; c++) {
            pl.col_ptr[pl.col_base[f] + c] = (int32_t)ne;
            for (; k < ent.size() && ent[k].first / n == c; k++, ne++) {
                pl.col_row[ne] = (int32_t)(ent[k].first % n);
                pl.col_nz[ne] = (int32_t)ent[k].second;
            }
        }
        for (int l = 0; l < (int)nl.size(); l++) loc[nl[l]] = -1;
    }
    pl.col_ptr[pl.col_base[nf]] = (int32_t)ne;
    if (ne != nnz) {
        hmec_set_last_error("front column lists do not cover the pattern");
        return HMEC_E_ARG;
    }
    // extend-add maps: child update node -> parent local node
    pl.ea_off.assign(nf, -1);
    int64_t ea = 0;
    for (int f = 0; f < nf; f++) {
        pl.ea_off[f] = ea;
        ea += pl.nu[f];
    }
    pl.ea_map.assign(ea, -1);
    for (int pa = 0; pa < nf; pa++) {
        const auto &nl = pl.nodes[pa];
        for (int l = 0; l < (int)nl.size(); l++) loc[nl[l]] = l;
        for (int k : pl.children[pa]) {
            for (int a2 = 0; a2 < pl.nu[k]; a2++) {
                const int v = pl.nodes[k][pl.np[k] + a2];
                if (loc[v] < 0) {
                    hmec_set_last_error("child update node missing from parent front");
                    return HMEC_E_ARG;
                }
                pl.ea_map[pl.ea_off[k] + a2] = loc[v];
            }
        }
        for (int l = 0; l < (int)nl.size(); l++) loc[nl[l]] = -1;
    }
    if (pl.nu[root] != 0) {
        hmec_set_last_error("root front has an update set");
        return HMEC_E_ARG;
    }
    return HMEC_OK;
}
