// Host-only entry points of the symbolic phase (no GPU needed): the fixed
// CSR pattern of the coupled system and the nested-dissection front plan.
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "hmec_internal.h"

int hmec_pattern_host(const hmec_params &P, std::vector<int32_t> &rowp,
                      std::vector<int32_t> &coli, std::vector<int32_t> *slot_out,
                      std::vector<int32_t> *node_trip_out, std::vector<uint8_t> *sec_out = nullptr);
void hmec_footprint_record(int Nx, int Ny, int64_t arena);

extern "C" int hmec_pattern(const hmec_params *p, int64_t *n, int64_t *nnz, int32_t *rowp,
                            int32_t *coli) {
    if (!p) return HMEC_E_ARG;
    hmec_params P = *p;
    HMEC_TRY(hmec_params_finalize(&P));
    std::vector<int32_t> rp, ci;
    HMEC_TRY(hmec_pattern_host(P, rp, ci, nullptr, nullptr));
    if (n) *n = (int64_t)rp.size() - 1;
    if (nnz) *nnz = (int64_t)ci.size();
    if (rowp) std::memcpy(rowp, rp.data(), sizeof(int32_t) * rp.size());
    if (coli) std::memcpy(coli, ci.data(), sizeof(int32_t) * ci.size());
    return HMEC_OK;
}

// the factored (Darcy-condensed) system's pattern and front plan of a grid
static int condensed_plan(const hmec_params &P, CondPlan &cp, FrontPlan &pl) {
    std::vector<int32_t> rp, ci;
    HMEC_TRY(hmec_pattern_host(P, rp, ci, nullptr, nullptr));
    HMEC_TRY(hmec_condense_build((int64_t)(P.Nx + 1) * (P.Ny + 1), rp, ci, cp));
    return hmec_plan_build(P.Nx + 1, P.Ny + 1, cp.srowp, cp.scoli, pl, 4);
}

extern "C" int hmec_condensed_pattern(const hmec_params *p, int64_t *n, int64_t *nnz, int32_t *rowp,
                                      int32_t *coli) {
    if (!p) return HMEC_E_ARG;
    hmec_params P = *p;
    HMEC_TRY(hmec_params_finalize(&P));
    std::vector<int32_t> rp, ci;
    HMEC_TRY(hmec_pattern_host(P, rp, ci, nullptr, nullptr));
    CondPlan cp;
    HMEC_TRY(hmec_condense_build((int64_t)(P.Nx + 1) * (P.Ny + 1), rp, ci, cp));
    if (n) *n = cp.NS;
    if (nnz) *nnz = (int64_t)cp.scoli.size();
    if (rowp) std::memcpy(rowp, cp.srowp.data(), sizeof(int32_t) * cp.srowp.size());
    if (coli) std::memcpy(coli, cp.scoli.data(), sizeof(int32_t) * cp.scoli.size());
    return HMEC_OK;
}

extern "C" int hmec_plan_info(const hmec_params *p, hmec_solve_stats *st, int64_t *arena_entries,
                              int32_t *elim_order) {
    if (!p) return HMEC_E_ARG;
    hmec_params P = *p;
    HMEC_TRY(hmec_params_finalize(&P));
    CondPlan cp;
    FrontPlan pl;
    HMEC_TRY(condensed_plan(P, cp, pl));
    hmec_footprint_record(P.Nx, P.Ny, pl.arena);
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->factor_flops = pl.flops;
        st->factor_nnz = pl.factor_nnz;
        st->nfronts = pl.nfronts;
        st->nlevels = pl.nlevels;
        st->max_front = pl.max_front;
    }
    if (arena_entries) *arena_entries = pl.arena;
    if (elim_order) {
        // condensed unknowns in elimination order: fronts by level, pivots in
        // front order
        const int dpn = pl.dpn;
        int inv[6];
        for (int v = 0; v < dpn; v++) inv[hmec_var_pos(dpn, v)] = v;
        int64_t k = 0;
        for (int lv = 0; lv < pl.nlevels; lv++)
            for (int f : pl.levels[lv])
                for (int l = 0; l < pl.np[f]; l++)
                    for (int q = 0; q < dpn; q++) elim_order[k++] = pl.nodes[f][l] * dpn + inv[q];
    }
    return HMEC_OK;
}

// Per level of the front plan: number of fronts, largest front order and
// pivot count, and the dense flops of the level (arrays of nlevels entries;
// call with NULL arrays to get nlevels).
extern "C" int hmec_plan_levels(const hmec_params *p, int *nlevels, int *nfronts, int *maxn,
                                int *maxp, double *flops) {
    if (!p || !nlevels) return HMEC_E_ARG;
    hmec_params P = *p;
    HMEC_TRY(hmec_params_finalize(&P));
    CondPlan cp;
    FrontPlan pl;
    HMEC_TRY(condensed_plan(P, cp, pl));
    *nlevels = pl.nlevels;
    for (int lv = 0; lv < pl.nlevels; lv++) {
        int mn = 0, mp = 0;
        double fl = 0;
        for (int f : pl.levels[lv]) {
            const int n = pl.dpn * (pl.np[f] + pl.nu[f]), pp = pl.dpn * pl.np[f];
            mn = std::max(mn, n);
            mp = std::max(mp, pp);
            for (int k = 0; k < pp; k++) {
                const double r = (double)(n - k - 1);
                fl += r + 2.0 * r * r;
            }
        }
        if (nfronts) nfronts[lv] = (int)pl.levels[lv].size();
        if (maxn) maxn[lv] = mn;
        if (maxp) maxp[lv] = mp;
        if (flops) flops[lv] = fl;
    }
    return HMEC_OK;
}

// Device bytes a context of this grid allocates under the given precision
// policy, from the exact symbolic plan (cached per grid): markers and their
// derived streams and sort buffers, nodal fields, CSR and its triplet map,
// solver plan arrays and vectors, then the factor storage of the policy
// (FP64: front arena + inverted diagonal blocks; FP32: FP32 front arena,
// inverted blocks, two factor slots and the GMRES basis; an FP64 fallback
// allocates the FP64 arena on top).
namespace {
struct PlanSizes {
    int64_t arena = 0, varena = 0, nnz = 0, snnz = 0, ntrip = 0, sweep = 0, dinv = 0, nfronts = 0;
    int64_t swrows = 0;   // front rows (+1 per front) of the sweep's row source maps
    int64_t icopy = 0;    // pivot-block copies of the largest level (triangle inversion)
    int64_t iscr = 0;     // inversion products of the largest level and doubling stage
    int64_t pull = 0;     // bytes of the front pull's tiles and child maps
};
std::mutex g_ps_mu;
std::map<std::pair<int, int>, PlanSizes> g_ps;
}  // namespace

extern "C" int64_t hmec_member_footprint(const hmec_params *p, int precision) {
    if (!p) return -1;
    hmec_params P = *p;
    if (hmec_params_finalize(&P) != HMEC_OK) return -1;
    PlanSizes z;
    bool have = false;
    {
        std::lock_guard<std::mutex> lk(g_ps_mu);
        auto it = g_ps.find({P.Nx, P.Ny});
        if (it != g_ps.end()) {
            z = it->second;
            have = true;
        }
    }
    if (!have) {
        std::vector<int32_t> rp, ci, slot, nt;
        if (hmec_pattern_host(P, rp, ci, &slot, &nt) != HMEC_OK) return -1;
        CondPlan cp;
        if (hmec_condense_build((int64_t)(P.Nx + 1) * (P.Ny + 1), rp, ci, cp) != HMEC_OK) return -1;
        FrontPlan pl;
        if (hmec_plan_build(P.Nx + 1, P.Ny + 1, cp.srowp, cp.scoli, pl, 4) != HMEC_OK) return -1;
        z.arena = pl.arena;
        z.varena = pl.varena;
        z.nnz = (int64_t)ci.size();
        z.snnz = (int64_t)cp.scoli.size();
        z.ntrip = (int64_t)slot.size();
        z.nfronts = pl.nfronts;
        for (int f = 0; f < pl.nfronts; f++) {
            const int64_t n = pl.dpn * (pl.np[f] + pl.nu[f]), pv = pl.dpn * pl.np[f];
            z.sweep += n * ((pv + 3) & ~3) + pv * ((n + 3) & ~3);
            z.dinv += ((pv + 31) / 32) * 2 * 32 * 32;
            z.swrows += n + 1;
            const int64_t nodes = pl.np[f] + pl.nu[f];
            const int64_t tc = std::max<int64_t>(1, kPullTile * kPullTile / n);
            z.pull += nodes * 4 * (int64_t)pl.children[f].size() + (n + tc - 1) / tc * (int64_t)sizeof(PullTile);
        }
        for (int lv = 0; lv < pl.nlevels; lv++) {
            int64_t co = 0;
            int maxp = 0;
            for (int f : pl.levels[lv]) {
                const int64_t pv = pl.dpn * pl.np[f];
                co += pv * pv;
                maxp = std::max<int>(maxp, (int)pv);
            }
            z.icopy = std::max(z.icopy, co);
            for (int b = 32; b < maxp; b *= 2) {
                int64_t to = 0;
                for (int f : pl.levels[lv]) {
                    const int pf = pl.dpn * pl.np[f];
                    for (int a0 = 0; a0 + b < pf; a0 += 2 * b) to += 2LL * b * std::min(b, pf - a0 - b);
                }
                z.iscr = std::max(z.iscr, to);
            }
        }
        std::lock_guard<std::mutex> lk(g_ps_mu);
        g_ps[{P.Nx, P.Ny}] = z;
    }
    const int64_t ns = (int64_t)(P.Nx + 1) * (P.Ny + 1), N = 6 * ns;
    const int64_t M = hmec_marker_count(&P);
    int64_t b = M * 8 * (HMEC_NMARKER_FIELDS + MP_NPROP + 31) + M * 4 * (4 + 4 + 2) + 16 * ns;
    b += ns * 8 * HMEC_NFIELDS;
    b += (6 * kNumTimesteps + 4 * kSeriesBlocks) * 8 + (P.niterglobal + 1) * 8;   // series, DSYLSQ
    const int64_t NS = 4 * ns;
    b += z.nnz * (4 + 8) + (N + 1) * 4 + z.ntrip * 4 + ns * 8;            // coupled CSR, sections
    b += z.snnz * (4 + 8 + 16 + 3 * 4 + 4 + 4 * 4) + (NS + 1) * 16;    // condensed CSR, maps
    b += z.pull + z.snnz * (4 + 8);                                     // front pull plan, values
    b += 6 * N * 8 + 4 * NS * 8 + 3 * N * 4 + z.varena * 8 + z.nfronts * 256;   // vectors, plans
    b += z.swrows * (4 + 4) + 3 * z.varena * 8 + NS * 4;                // sweep row maps, slots, zf
    b += (z.icopy + z.iscr) * 8;                                        // inversion scratch
    if (precision == HMEC_PREC_FP64) {
        b += z.arena * 8 + z.dinv * 8;
    } else {
        // the FP32 factor slots (sweep form + scales each)
        b += z.arena * 4 + z.dinv * 4 + HMEC_FP32_SLOTS * (z.sweep * 4 + 2 * NS * 8) + (2 * 20 + 3) * N * 8;
        b += 3 * hmec_ctx::NGUESS * N * 8;                              // GMRES initial guesses
    }
    return b;
}
