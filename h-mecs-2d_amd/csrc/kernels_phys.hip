// Physics kernels of the per-timestep path for gfx950 (FP64, no FP
// contraction so that every expression rounds like the reference's):
//   marker preparation + cell sort, marker->node scatter with normalization,
//   VEP pre-update, P-scaling, CSR value fill, reload, stress update, RSF
//   fault solve with its reductions, Picard commit, final stresses, history,
//   and the RK4 marker advection.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>

#define HMEC_HD __host__ __device__
#include "assemble_emit.h"
#include "controller.h"
#include "dpp_sum.h"
#include "pow2_scale.h"
#include "xcd_map.h"
#include "hmec_internal.h"

// device scalar slots (ctx->d.scal)
enum {
    SC_PTSCALE = 0,
    SC_YNPL, SC_DDD, SC_DTLAP, SC_VMAX, SC_VXMAX, SC_VXMIN, SC_VYMAX, SC_VYMIN,
    SC_BANDX, SC_BANDY, SC_DTSLIP, SC_CAPPED,
    // device-side Picard / dt controller: its state and the exit flag
    SC_CDT, SC_CDTX, SC_CDTY, SC_CMAXVXY, SC_CYNLAST, SC_EXIT,
    // outcome of the linear solve's first GMRES step decided on the device
    SC_GMSTATE,
    SC_COUNT
};
static_assert(SC_PTSCALE == SC_PTSCALE_SLOT, "ptscale slot");
static_assert(SC_EXIT == SC_EXIT_SLOT && SC_GMSTATE == SC_GMSTATE_SLOT, "solver scalar slots");
static_assert(SC_COUNT <= 32, "scalar slots");

// coordinates: x(Nx), y(Ny), xvx(Nx1), yvx(Ny1), xvy(Nx1), yvy(Ny1), xp(Nx1), yp(Ny1)
struct Coords {
    const double *x, *y, *xvx, *yvx, *xvy, *yvy, *xp, *yp;
};
static Coords coords(hmec_ctx *c) {
    const double *b = c->d.coords;
    Coords k;
    k.x = b;
    k.y = k.x + c->Nx;
    k.xvx = k.y + c->Ny;
    k.yvx = k.xvx + c->Nx1;
    k.xvy = k.yvx + c->Ny1;
    k.yvy = k.xvy + c->Nx1;
    k.xp = k.yvy + c->Ny1;
    k.yp = k.xp + c->Nx1;
    return k;
}

#define LAUNCH_CHECK()                                                      \
    do {                                                                    \
        hipError_t e_ = hipGetLastError();                                  \
        if (e_ != hipSuccess) {                                             \
            hmec_set_last_error(std::string("launch: ") + hipGetErrorString(e_)); \
            return HMEC_E_HIP;                                              \
        }                                                                   \
    } while (0)

__device__ __forceinline__ int d_fix(double t) { return t < 0. ? (int)ceil(t) : (int)floor(t); }
__device__ __forceinline__ int d_check_bounds(int k, int bound) {
    if (k < 0) k = 0;
    else if (k > bound - 2) k = bound - 2;
    return k;
}
__device__ __forceinline__ double d_enforce(double v, double mn, double mx) {
    if (v < mn) v = mn;
    else if (v > mx) v = mx;
    return v;
}

// ---------------------------------------------- cell-ordered marker streams
// Per grid g, the marker values its scatter reads (x, y and the grid's
// properties) copied once per marker set into the grid's cell order, so a
// node reads each of its cells' markers as one contiguous run.  Layout: grid
// blocks at SORTED_OFF[g] * M, (2 + nprop_g) streams of M doubles each.
constexpr int SORTED_OFF[4] = {0, 12, 19, 25};
constexpr int SORTED_STREAMS = 31;
template <int G>
__global__ void k_sort_copy(const double *__restrict__ mk, const double *__restrict__ prop,
                            int64_t M, const int32_t *__restrict__ perm, double *__restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= M) return;
    const int m = perm[k];
    double *o = out + (int64_t)SORTED_OFF[G] * M;
    o[k] = mk[(int64_t)HMEC_M_XM * M + m];
    o[M + k] = mk[(int64_t)HMEC_M_YM * M + m];
#define PR(f) prop[(int64_t)(f) * M + m]
    if (G == 0) {
        const int pf[10] = {MP_RHOMM, MP_KKKMM, MP_TMARK, MP_INVG, MP_ETAMM0,
                            MP_COHT, MP_FRIT, MP_COHC, MP_FRIC, MP_DILC};
#pragma unroll
        for (int f = 0; f < 10; f++) o[(int64_t)(2 + f) * M + k] = PR(pf[f]);
    } else if (G == 1) {
        const int pf[5] = {MP_ETAMM, MP_PORM, MP_INVG, MP_ETAMM0, MP_ETASMM0};
#pragma unroll
        for (int f = 0; f < 5; f++) o[(int64_t)(2 + f) * M + k] = PR(pf[f]);
    } else {
        const int pf[4] = {MP_RHOMM, MP_INVETAD, MP_RHOFM, MP_PORM};
#pragma unroll
        for (int f = 0; f < 4; f++) o[(int64_t)(2 + f) * M + k] = PR(pf[f]);
    }
#undef PR
}

// ------------------------------------------------------------ marker prep
// Derived marker properties (h_mec_rsf_v9.cpp:389-423) and the anchor cell
// of each of the four grids (:432-433, :468-469, :493-494, :515-516).
// Cell id = i + j * (Ny - 1), i in [0, Ny-2], j in [0, Nx-2].
__global__ void k_prep(const double *__restrict__ mk, int64_t M, double *__restrict__ prop,
                       int32_t *__restrict__ cell, int32_t *__restrict__ keys, int Nx, int Ny,
                       double dx, double dy, double POR0, double kkkmin, double kkkmax,
                       double alpha, double etamin, double etamax) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
#define MK(f) mk[(int64_t)(f) * M + m]
#define PR(f) prop[(int64_t)(f) * M + m]
    const double porm = MK(HMEC_M_PORM);
    double kkkmm = MK(HMEC_M_KKKM) * pow(porm / POR0, 3.0);
    kkkmm = d_enforce(kkkmm, kkkmin, kkkmax);
    double cohescmm, cohestmm, frictcmm, dilatcmm, fricttmm, etasmm0, etamm0, etamm, rhomm;
    if (MK(HMEC_M_T_MARKER) != 0) {
        cohescmm = MK(HMEC_M_COHESCM) * (1 - porm);
        cohestmm = MK(HMEC_M_COHESTM) * (1 - porm);
        frictcmm = MK(HMEC_M_FRICTCM);
        dilatcmm = MK(HMEC_M_DILATCM);
        fricttmm = MK(HMEC_M_FRICTTM);
        etasmm0 = MK(HMEC_M_ETASM);
        etamm0 = MK(HMEC_M_ETASM) * exp(-alpha * porm);
        etamm = MK(HMEC_M_ETAM);
        rhomm = MK(HMEC_M_RHOM) * (1 - porm) + MK(HMEC_M_RHOFM) * porm;
    } else {
        cohescmm = MK(HMEC_M_COHESCM);
        cohestmm = MK(HMEC_M_COHESTM);
        frictcmm = MK(HMEC_M_FRICTCM);
        dilatcmm = 0;
        fricttmm = MK(HMEC_M_FRICTTM);
        etasmm0 = MK(HMEC_M_ETASM);
        etamm0 = MK(HMEC_M_ETASM);
        etamm = MK(HMEC_M_ETAM);
        rhomm = MK(HMEC_M_RHOM);
    }
    etamm0 = d_enforce(etamm0, etamin, etamax);
    etamm = d_enforce(etamm, etamin, etamax);
    const double etadm = MK(HMEC_M_ETAFM) / kkkmm;
    PR(MP_ETAMM) = etamm;
    PR(MP_RHOMM) = rhomm;
    PR(MP_KKKMM) = kkkmm;
    PR(MP_TMARK) = MK(HMEC_M_T_MARKER);
    PR(MP_INVG) = 1. / MK(HMEC_M_GM);
    PR(MP_ETAMM0) = etamm0;
    PR(MP_COHT) = cohestmm;
    PR(MP_FRIT) = fricttmm;
    PR(MP_COHC) = cohescmm;
    PR(MP_FRIC) = frictcmm;
    PR(MP_DILC) = dilatcmm;
    PR(MP_PORM) = porm;
    PR(MP_ETASMM0) = etasmm0;
    PR(MP_INVETAD) = 1. / etadm;
    PR(MP_RHOFM) = MK(HMEC_M_RHOFM);
    const double xm = MK(HMEC_M_XM), ym = MK(HMEC_M_YM);
    const int cy = Ny - 1;
    int j, i;
    j = d_check_bounds(d_fix(xm / dx), Nx);
    i = d_check_bounds(d_fix(ym / dy), Ny);
    cell[0 * M + m] = i + j * cy;
    j = d_check_bounds(d_fix((xm + dx / 2.) / dx), Nx);
    i = d_check_bounds(d_fix((ym + dy / 2.) / dy), Ny);
    cell[1 * M + m] = i + j * cy;
    j = d_check_bounds(d_fix((xm) / dx), Nx);
    i = d_check_bounds(d_fix((ym + dy / 2.) / dy), Ny);
    cell[2 * M + m] = i + j * cy;
    j = d_check_bounds(d_fix((xm + dx / 2.) / dx), Nx);
    i = d_check_bounds(d_fix((ym) / dy), Ny);
    cell[3 * M + m] = i + j * cy;
    keys[m] = (int32_t)m;
#undef MK
#undef PR
}

__global__ void k_iota(int32_t *v, int64_t M) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m < M) v[m] = (int32_t)m;
}

// cell_start[c] = first sorted position with key >= c (lower bound)
__global__ void k_cell_start(const int32_t *__restrict__ sorted_keys, int64_t M, int ncell,
                             int32_t *__restrict__ start) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > ncell) return;
    int64_t lo = 0, hi = M;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (sorted_keys[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    start[c] = (int32_t)lo;
}

int k_marker_prep(hmec_ctx *c) {
    const int64_t M = c->M;
    if (M <= 0) return HMEC_OK;
    const hmec_params &P = c->P;
    const int threads = 256;
    const int blocks = (int)((M + threads - 1) / threads);
    hipLaunchKernelGGL(k_prep, dim3(blocks), dim3(threads), 0, c->stream, c->d.mk, M, c->d.prop,
                       c->d.cell, c->d.keys_tmp, c->Nx, c->Ny, P.dx, P.dy, P.POR0, P.kkkmin,
                       P.kkkmax, P.alpha, P.etamin, P.etamax);
    LAUNCH_CHECK();
    const int ncell = (c->Nx - 1) * (c->Ny - 1);
    int nbits = 1;
    while ((1 << nbits) <= ncell) nbits++;
    for (int g = 0; g < 4; g++) {
        // cell ids are non-negative: sort them as unsigned keys
        uint32_t *keys_in = reinterpret_cast<uint32_t *>(c->d.cell + (int64_t)g * M);
        uint32_t *keys_out = reinterpret_cast<uint32_t *>(c->d.keys_tmp + M);   // sorted keys
        int32_t *vals_in = c->d.keys_tmp;                                      // iota
        int32_t *vals_out = c->d.perm + (int64_t)g * M;
        hipLaunchKernelGGL(k_iota, dim3(blocks), dim3(threads), 0, c->stream, vals_in, M);
        LAUNCH_CHECK();
        size_t tmp = c->d.sort_tmp_bytes;
        hipError_t e = rocprim::radix_sort_pairs(c->d.sort_tmp, tmp, keys_in, keys_out, vals_in,
                                                 vals_out, (unsigned int)M, 0, nbits, c->stream);
        if (e != hipSuccess) {
            hmec_set_last_error(std::string("radix_sort_pairs: ") + hipGetErrorString(e));
            return HMEC_E_HIP;
        }
        hipLaunchKernelGGL(k_cell_start, dim3((ncell + 1 + 255) / 256), dim3(256), 0, c->stream,
                           reinterpret_cast<const int32_t *>(keys_out), M, ncell, c->d.cell_start + (int64_t)g * (ncell + 1));
        LAUNCH_CHECK();
    }
    // cell-ordered streams of the scatter (one copy per marker set)
    hipLaunchKernelGGL(k_sort_copy<0>, dim3(blocks), dim3(threads), 0, c->stream, c->d.mk, c->d.prop,
                       M, c->d.perm, c->d.sorted);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sort_copy<1>, dim3(blocks), dim3(threads), 0, c->stream, c->d.mk, c->d.prop,
                       M, c->d.perm + M, c->d.sorted);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sort_copy<2>, dim3(blocks), dim3(threads), 0, c->stream, c->d.mk, c->d.prop,
                       M, c->d.perm + 2 * M, c->d.sorted);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sort_copy<3>, dim3(blocks), dim3(threads), 0, c->stream, c->d.mk, c->d.prop,
                       M, c->d.perm + 3 * M, c->d.sorted);
    LAUNCH_CHECK();
    c->markers_dirty = false;
    return HMEC_OK;
}

size_t hmec_sort_tmp_bytes(int64_t M, int nbits) {
    size_t tmp = 0;
    rocprim::radix_sort_pairs(nullptr, tmp, (uint32_t *)nullptr, (uint32_t *)nullptr,
                              (int32_t *)nullptr, (int32_t *)nullptr, (unsigned int)M, 0, nbits,
                              (hipStream_t)0);
    return tmp;
}

// normalized sums of one node -> its fields (skipped where no marker weight)
template <int G, int NF>
__device__ __forceinline__ void scatter_store(const double *acc, double *__restrict__ fields,
                                              int64_t ns, int64_t idx) {
    const double wt = acc[NF];
    if (!(wt > 0)) return;
#define FLD(f) fields[(int64_t)(f) * ns + idx]
    if (G == 0) {
        FLD(HMEC_F_RHO) = acc[0] / wt;
        FLD(HMEC_F_KKK) = acc[1] / wt;
        FLD(HMEC_F_TTT) = acc[2] / wt;
        FLD(HMEC_F_GGG) = 1. / (acc[3] / wt);
        FLD(HMEC_F_ETA0) = acc[4] / wt;
        FLD(HMEC_F_COHT) = acc[5] / wt;
        FLD(HMEC_F_FRIT) = acc[6] / wt;
        FLD(HMEC_F_COHC) = acc[7] / wt;
        FLD(HMEC_F_FRIC) = acc[8] / wt;
        FLD(HMEC_F_DILC) = acc[9] / wt;
    } else if (G == 1) {
        FLD(HMEC_F_ETAP) = acc[0] / wt;
        const double por = acc[1] / wt;
        FLD(HMEC_F_POR) = por;
        FLD(HMEC_F_ETAB0) = acc[4] / wt / por;
        FLD(HMEC_F_ETAP0) = acc[3] / wt;
        FLD(HMEC_F_GGGP) = 1. / (acc[2] / wt);
    } else if (G == 2) {
        FLD(HMEC_F_RHOX) = acc[0] / wt;
        FLD(HMEC_F_ETADX) = 1. / (acc[1] / wt);
        FLD(HMEC_F_RHOFX) = acc[2] / wt;
        FLD(HMEC_F_PORX) = acc[3] / wt;
    } else {
        FLD(HMEC_F_RHOY) = acc[0] / wt;
        FLD(HMEC_F_ETADY) = 1. / (acc[1] / wt);
        FLD(HMEC_F_RHOFY) = acc[2] / wt;
        FLD(HMEC_F_PORY) = acc[3] / wt;
    }
#undef FLD
}

// ----------------------------------------------------------------- scatter
// L lanes per run of R vertically adjacent nodes (i0 .. i0 + R - 1, column j)
// of grid g.  The run reads each marker of the cells around it once: cells
// (i0 - 1 .. i0 + R - 1) x (j, j - 1), in that order; a marker of cell
// (ci, cj) adds to node ci with weight (1 - dym) wx and to node ci + 1 with
// weight dym wx (wx = 1 - dxm in column j, dxm in column j - 1) where those
// nodes belong to the run.  Inside a cell, lane l takes the cell's markers
// l, l + L, ... of its cell-ordered run (the L lanes read consecutive
// addresses), and DPP row moves over the L lanes sum each node's partials —
// deterministic, no atomics.  The sums then go through the reference's
// normalization (h_mec_rsf_v9.cpp:531-583).
constexpr int SCATTER_T = 256;
constexpr int SCATTER_R = 2;   // nodes per run

template <int G, int L, int R>
__global__ void __launch_bounds__(SCATTER_T)
k_scatter_grid(const double *__restrict__ sorted, int64_t M, const int32_t *__restrict__ start,
               const double *__restrict__ cx, const double *__restrict__ cyv,
               double *__restrict__ fields, int64_t ns, int Nx, int Ny, double dx, double dy) {
    const int rows = (G == 0) ? Ny : Ny + 1;
    const int cols = (G == 0) ? Nx : Nx + 1;
    // runs in column-major order, 256 / L runs per workgroup; each XCD takes
    // a contiguous strip of them, so the runs around a cell mostly read its
    // markers through one L2
    const int rpc = (rows + R - 1) / R;                 // runs per column
    const int nrun = rpc * cols;
    const int nwg = (nrun + SCATTER_T / L - 1) / (SCATTER_T / L);
    const int g = (int)xcd_block(nwg) * (SCATTER_T / L) + threadIdx.x / L;
    const int l = threadIdx.x % L;
    if (g >= nrun) return;   // whole lane groups leave together
    const int j = g / rpc, i0 = (g % rpc) * R;
    const int cy = Ny - 1, cxn = Nx - 1;
    constexpr int NF = (G == 0) ? 10 : (G == 1) ? 5 : 4;
    double acc[R][NF + 1];
#pragma unroll
    for (int u = 0; u < R; u++)
#pragma unroll
        for (int f = 0; f <= NF; f++) acc[u][f] = 0.;
    const double *src = sorted + (int64_t)SORTED_OFF[G] * M;   // this grid's cell-ordered streams
    const double *xm = src, *ym = src + M;
#pragma unroll
    for (int dj = 0; dj < 2; dj++) {
        const int cj = j - dj;
        if (cj < 0 || cj >= cxn) continue;
        const double x0 = cx[cj];
#pragma unroll
        for (int u = -1; u < R; u++) {                  // cell row ci = i0 + u: nodes u, u + 1
            const int ci = i0 + u;
            if (ci < 0 || ci >= cy) continue;
            const int cid = ci + cj * cy;
            const double y0 = cyv[ci];
            const int end = start[cid + 1];
            for (int k = start[cid] + l; k < end; k += L) {
                const double dxm = (xm[k] - x0) / dx;
                const double dym = (ym[k] - y0) / dy;
                const double wx = dj ? dxm : (1 - dxm);
                double pv[NF];
#pragma unroll
                for (int f = 0; f < NF; f++) pv[f] = src[(int64_t)(2 + f) * M + k];
                if (u >= 0) {
                    const double w = (1 - dym) * wx;
#pragma unroll
                    for (int f = 0; f < NF; f++) acc[u < 0 ? 0 : u][f] += pv[f] * w;
                    acc[u < 0 ? 0 : u][NF] += w;
                }
                if (u + 1 < R) {
                    const double w = dym * wx;
#pragma unroll
                    for (int f = 0; f < NF; f++) acc[u + 1 < R ? u + 1 : 0][f] += pv[f] * w;
                    acc[u + 1 < R ? u + 1 : 0][NF] += w;
                }
            }
        }
    }
#pragma unroll
    for (int u = 0; u < R; u++)
#pragma unroll
        for (int f = 0; f <= NF; f++) acc[u][f] = group_sum<L>(acc[u][f]);
    if (l != 0) return;
#pragma unroll
    for (int u = 0; u < R; u++) {
        const int i = i0 + u;
        if (i >= rows) break;
        scatter_store<G, NF>(acc[u], fields, ns, i + (int64_t)j * rows);
    }
}

int k_scatter(hmec_ctx *c) {
    if (!c->have_markers) return HMEC_E_ARG;
    if (c->markers_dirty) HMEC_TRY(k_marker_prep(c));
    Coords k = coords(c);
    const int ncell = (c->Nx - 1) * (c->Ny - 1);
    const int64_t M = c->M;
    const hmec_params &P = c->P;
    constexpr int L = 16, NPB = SCATTER_T / L;
    // workgroups of a grid with rows x cols nodes in runs of SCATTER_R
    auto blocks = [&](int rows, int cols) -> unsigned {
        const int64_t runs = (int64_t)(rows + SCATTER_R - 1) / SCATTER_R * cols;
        return (unsigned)xcd_grid((runs + NPB - 1) / NPB);
    };
    hipLaunchKernelGGL((k_scatter_grid<0, L, SCATTER_R>), dim3(blocks(c->Ny, c->Nx)), dim3(SCATTER_T), 0,
                       c->stream, c->d.sorted, M, c->d.cell_start, k.x, k.y, c->d.fields, c->ns,
                       c->Nx, c->Ny, P.dx, P.dy);
    LAUNCH_CHECK();
    hipLaunchKernelGGL((k_scatter_grid<1, L, SCATTER_R>), dim3(blocks(c->Ny1, c->Nx1)), dim3(SCATTER_T), 0,
                       c->stream, c->d.sorted, M, c->d.cell_start + (ncell + 1), k.xp, k.yp,
                       c->d.fields, c->ns, c->Nx, c->Ny, P.dx, P.dy);
    LAUNCH_CHECK();
    hipLaunchKernelGGL((k_scatter_grid<2, L, SCATTER_R>), dim3(blocks(c->Ny1, c->Nx1)), dim3(SCATTER_T), 0,
                       c->stream, c->d.sorted, M, c->d.cell_start + 2 * (ncell + 1), k.xvx, k.yvx,
                       c->d.fields, c->ns, c->Nx, c->Ny, P.dx, P.dy);
    LAUNCH_CHECK();
    hipLaunchKernelGGL((k_scatter_grid<3, L, SCATTER_R>), dim3(blocks(c->Ny1, c->Nx1)), dim3(SCATTER_T), 0,
                       c->stream, c->d.sorted, M, c->d.cell_start + 3 * (ncell + 1), k.xvy, k.yvy,
                       c->d.fields, c->ns, c->Nx, c->Ny, P.dx, P.dy);
    LAUNCH_CHECK();
    return HMEC_OK;
}

// --------------------------------------------------------- initial state
// Fresh-start nodal state, h_mec_rsf_v9.cpp:283-338.
__global__ void k_init(double *fields, int64_t ns, const double *x, const double *y, int Nx,
                       int Ny, hmec_params P) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int Ny1 = Ny + 1, Nx1 = Nx + 1;
    if (t < Nx1 * Ny1) {
        fields[(int64_t)HMEC_F_PT0 * ns + t] = P.PCONF + P.PTFDIFF;
        fields[(int64_t)HMEC_F_PF0 * ns + t] = P.PCONF;
    }
    if (t >= Nx * Ny) return;
    const int i = t % Ny, j = t / Ny;
    double om0 = P.omm[0], ar = P.arsfm[1], br = P.brsfm[1], lr = P.lrsfm[0];
    if (y[i] > P.upper_block && y[i] < P.lower_block) {
        const double xj = x[j];
        const double TS_1 = P.TS_1, TS_2 = P.TS_2, TS_3 = P.TS_3, TS_4 = P.TS_4;
        om0 = P.omm[1];
        if (xj < TS_1) { br = P.brsfm[0]; ar = P.arsfm[0]; }
        if (xj >= TS_1 && xj < TS_2) {
            br = P.brsfm[0] - (P.brsfm[0] - P.brsfm[1]) * ((xj - TS_1) / (TS_2 - TS_1));
            ar = P.arsfm[0] - (P.arsfm[0] - P.arsfm[1]) * ((xj - TS_1) / (TS_2 - TS_1));
            lr = P.lrsfm[0] - (P.lrsfm[0] - P.lrsfm[1]) * ((xj - TS_1) / (TS_2 - TS_1));
        }
        if (xj >= TS_2 && xj <= TS_3) lr = P.lrsfm[1];
        if (xj > TS_3 && xj <= TS_4) {
            br = P.brsfm[1] - (P.brsfm[1] - P.brsfm[0]) * ((xj - TS_3) / (TS_4 - TS_3));
            ar = P.arsfm[1] - (P.arsfm[1] - P.arsfm[0]) * ((xj - TS_3) / (TS_4 - TS_3));
            lr = P.lrsfm[1] - (P.lrsfm[1] - P.lrsfm[0]) * ((xj - TS_3) / (TS_4 - TS_3));
        }
        if (xj > TS_4) { br = P.brsfm[0]; ar = P.arsfm[0]; }
    }
    fields[(int64_t)HMEC_F_OM0 * ns + t] = om0;
    fields[(int64_t)HMEC_F_OM * ns + t] = om0;
    fields[(int64_t)HMEC_F_ARSF * ns + t] = ar;
    fields[(int64_t)HMEC_F_BRSF * ns + t] = br;
    fields[(int64_t)HMEC_F_LRSF * ns + t] = lr;
}

int k_init_state(hmec_ctx *c) {
    Coords k = coords(c);
    const int n = c->Nx1 * c->Ny1;
    hipLaunchKernelGGL(k_init, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->d.fields, c->ns,
                       k.x, k.y, c->Nx, c->Ny, c->P);
    LAUNCH_CHECK();
    return HMEC_OK;
}

// up to four field copies in one launch (dst[q] = src[q], n[q] entries)
struct FieldCopies {
    double *dst[4];
    const double *src[4];
    int64_t n[4];
    int count;
};
__global__ void k_copy_fields(FieldCopies fc, int64_t nmax) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nmax) return;
#pragma unroll
    for (int q = 0; q < 4; q++)
        if (q < fc.count && t < fc.n[q]) fc.dst[q][t] = fc.src[q][t];
}

// fields dst[q] = src[q] (q < count <= 4) in one launch
int k_copy_fields(hmec_ctx *c, const int *dst, const int *src, int count) {
    FieldCopies fc{};
    int64_t nmax = 0;
    for (int q = 0; q < count; q++) {
        fc.dst[q] = F(c, dst[q]);
        fc.src[q] = F(c, src[q]);
        fc.n[q] = (dst[q] < HMEC_NBASIC_FIELDS) ? c->nb : c->ns;
        nmax = std::max(nmax, fc.n[q]);
    }
    fc.count = count;
    hipLaunchKernelGGL(k_copy_fields, dim3((unsigned)((nmax + 255) / 256)), dim3(256), 0, c->stream,
                       fc, nmax);
    LAUNCH_CHECK();
    return HMEC_OK;
}

int k_copy_field(hmec_ctx *c, int dst, int src) { return k_copy_fields(c, &dst, &src, 1); }

// ------------------------------------------------------------ VEP
// etamincur = dt * shearmod * 1e-4 with dt from the controller (ctl) or given
__device__ __forceinline__ double vep_floor(const double *ctl, double etamincur, double shearmod) {
    return ctl ? ctl[SC_CDT] * shearmod * 1e-4 : etamincur;
}

// basic nodes, h_mec_rsf_v9.cpp:620-636; threads past the basic nodes copy
// the pressure boundaries (copy_bounds of pt and pf, :616-617: the outer
// ring from the neighbouring interior node, columns then rows, so a corner
// takes its diagonal interior neighbour; the sources are interior nodes,
// which nothing here writes, and the VEP update reads no pressure)
__global__ void k_vep_basic(double *f, int64_t ns, int Nx, int Ny, double etamincur,
                            double dy_faultw, const double *ctl, double shearmod) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (ctl && ctl[SC_EXIT] != 0) return;
    if (t >= Nx * Ny) {
        const int Ny1 = Ny + 1, q = t - Nx * Ny;
        int i, j;
        if (q < 2 * Ny1) {
            i = q % Ny1;
            j = q < Ny1 ? 0 : Nx;
        } else if (q < 2 * Ny1 + 2 * (Nx - 1)) {
            const int q2 = q - 2 * Ny1;
            j = 1 + q2 % (Nx - 1);
            i = q2 < Nx - 1 ? 0 : Ny;
        } else {
            return;
        }
        const int si = i == 0 ? 1 : (i == Ny ? Ny - 1 : i);
        const int sj = j == 0 ? 1 : (j == Nx ? Nx - 1 : j);
        const int64_t d = i + (int64_t)j * Ny1, sidx = si + (int64_t)sj * Ny1;
        double *pt = f + (int64_t)HMEC_F_PT * ns, *pf = f + (int64_t)HMEC_F_PF * ns;
        pt[d] = pt[sidx];
        pf[d] = pf[sidx];
        return;
    }
    etamincur = vep_floor(ctl, etamincur, shearmod);
    const int i = t % Ny, j = t / Ny, Ny1 = Ny + 1;
    double *ETA = f + (int64_t)HMEC_F_ETA * ns;
    const double *ETA0 = f + (int64_t)HMEC_F_ETA0 * ns;
    double eta = ETA[t];
    if (eta < etamincur) { eta = etamincur; ETA[t] = eta; }
    if (eta < ETA0[t]) {
        const double *SXX = f + (int64_t)HMEC_F_SXX * ns, *SYY = f + (int64_t)HMEC_F_SYY * ns;
        const int64_t s = i + (int64_t)j * Ny1;
        const double SXX_temp = ((SXX[s] + SXX[s + 1]) + SXX[s + Ny1]) + SXX[s + Ny1 + 1];
        const double SYY_temp = ((SYY[s] + SYY[s + 1]) + SYY[s + Ny1]) + SYY[s + Ny1 + 1];
        const double sxy = f[(int64_t)HMEC_F_SXY * ns + t];
        const double siib = sqrt(sxy * sxy + (SXX_temp * SXX_temp + SYY_temp * SYY_temp +
                                              (SXX_temp + SYY_temp) * (SXX_temp + SYY_temp)) / 32.);
        f[(int64_t)HMEC_F_SIIB * ns + t] = siib;
        const double ietapl = (1. / eta - 1. / ETA0[t]);
        f[(int64_t)HMEC_F_IETAPLB * ns + t] = ietapl;
        f[(int64_t)HMEC_F_EIIB * ns + t] = dy_faultw * siib * ietapl / 2.;
    } else {
        f[(int64_t)HMEC_F_EIIB * ns + t] = 0;
        f[(int64_t)HMEC_F_IETAPLB * ns + t] = 0;
    }
}

// pressure nodes, h_mec_rsf_v9.cpp:639-669
__global__ void k_vep_p(double *f, int64_t ns, int Nx, int Ny, double etamincur,
                        double dy_faultw, const double *ctl, double shearmod) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (Nx - 1) * (Ny - 1) || (ctl && ctl[SC_EXIT] != 0)) return;
    etamincur = vep_floor(ctl, etamincur, shearmod);
    const int i = 1 + t % (Ny - 1), j = 1 + t / (Ny - 1), Ny1 = Ny + 1;
    const double *IE = f + (int64_t)HMEC_F_IETAPLB * ns;
    const double *YNY0 = f + (int64_t)HMEC_F_YNY0 * ns;
    const int64_t b00 = (i - 1) + (int64_t)(j - 1) * Ny, b10 = i + (int64_t)(j - 1) * Ny;
    const int64_t b01 = (i - 1) + (int64_t)j * Ny, b11 = i + (int64_t)j * Ny;
    const double IETAPL = (((IE[b00] + IE[b10]) + IE[b01]) + IE[b11]) / 4.;
    const int64_t s = i + (int64_t)j * Ny1;
    const double POR = f[(int64_t)HMEC_F_POR * ns + s];
    double etap, etab;
    if (YNY0[b00] > 0 || YNY0[b10] > 0 || YNY0[b01] > 0 || YNY0[b11] > 0) {
        etap = 1. / (1. / f[(int64_t)HMEC_F_ETAP0 * ns + s] + IETAPL);
        etab = 1. / (1. / f[(int64_t)HMEC_F_ETAB0 * ns + s] + dy_faultw * IETAPL * POR);
    } else {
        etap = f[(int64_t)HMEC_F_ETAP0 * ns + s];
        etab = f[(int64_t)HMEC_F_ETAB0 * ns + s];
    }
    if (etap < etamincur) etap = etamincur;
    if (etab * POR < etamincur) etab = etamincur / POR;
    f[(int64_t)HMEC_F_ETAP * ns + s] = etap;
    f[(int64_t)HMEC_F_ETAB * ns + s] = etab;
    f[(int64_t)HMEC_F_GGGB * ns + s] = f[(int64_t)HMEC_F_GGGP * ns + s] / POR;
    f[(int64_t)HMEC_F_DILP * ns + s] = 0;
}

// dev: dt from the device controller and the launches gated on its exit flag
// (the next Picard iteration's prologue enqueued before the host has read
// the controller's decision)
int k_vep_pre(hmec_ctx *c, double dt, bool dev) {
    const hmec_params &P = c->P;
    const double etamincur = dt * P.shearmod * 1e-4;
    const double dy_faultw = P.dy / P.faultwidth;
    const double *ctl = dev ? c->d.scal : nullptr;
    // basic nodes, then the pressure boundary ring
    const int nb = c->Nx * c->Ny + 2 * c->Ny1 + 2 * (c->Nx - 1);
    hipLaunchKernelGGL(k_vep_basic, dim3((nb + 255) / 256), dim3(256), 0, c->stream, c->d.fields,
                       c->ns, c->Nx, c->Ny, etamincur, dy_faultw, ctl, P.shearmod);
    LAUNCH_CHECK();
    const int np = (c->Nx - 1) * (c->Ny - 1);
    hipLaunchKernelGGL(k_vep_p, dim3((np + 255) / 256), dim3(256), 0, c->stream, c->d.fields,
                       c->ns, c->Nx, c->Ny, etamincur, dy_faultw, ctl, P.shearmod);
    LAUNCH_CHECK();
    return HMEC_OK;
}

// ---------------------------------------------------------- scale
// pfscale = ptscale = min(ETADX(1:Ny-1, 0:Nx-1)) * dx * (dt > 2e5 ? 1e17/dt^2 : 1)
// (h_mec_rsf_v9.cpp:671-677); a min is order independent, so exact.
// two stages: block minima, then one block over them
constexpr int SCALE_BLOCKS = 256;
__global__ void __launch_bounds__(256) k_scale_part(const double *ETADX, int Nx, int Ny,
                                                    double *part) {
    __shared__ double red[256];
    const int Ny1 = Ny + 1;
    double mn = INFINITY;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < (Ny - 1) * Nx;
         t += gridDim.x * blockDim.x) {
        const int i = 1 + t % (Ny - 1), j = t / (Ny - 1);
        mn = fmin(mn, ETADX[i + (int64_t)j * Ny1]);
    }
    red[threadIdx.x] = mn;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ void __launch_bounds__(SCALE_BLOCKS) k_scale_final(const double *part, double dx,
                                                              double dt, double *scal, int dev) {
    __shared__ double red[SCALE_BLOCKS];
    if (dev) {
        if (scal[SC_EXIT] != 0) return;
        dt = scal[SC_CDT];
    }
    red[threadIdx.x] = part[threadIdx.x];
    __syncthreads();
    for (int s = SCALE_BLOCKS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double m = red[0];
        scal[SC_PTSCALE] = dt > 2e5 ? m * dx * 1e17 / (dt * dt) : m * dx;
    }
}

// ptscale into the device scalars; *ptscale (when given) waits for it; dev:
// dt from the device controller, gated on its exit flag
int k_scale(hmec_ctx *c, double dt, double *ptscale, bool dev) {
    double *part = c->d.work;
    hipLaunchKernelGGL(k_scale_part, dim3(SCALE_BLOCKS), dim3(256), 0, c->stream,
                       F(c, HMEC_F_ETADX), c->Nx, c->Ny, part);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scale_final, dim3(1), dim3(SCALE_BLOCKS), 0, c->stream,
                       (const double *)part, c->P.dx, dt, c->d.scal, dev ? 1 : 0);
    LAUNCH_CHECK();
    if (!ptscale) return HMEC_OK;
    HMEC_HIP(hipMemcpyAsync(c->d.h_scal + SC_PTSCALE, c->d.scal + SC_PTSCALE, sizeof(double),
                            hipMemcpyDeviceToHost, c->stream));
    HMEC_HIP(hipStreamSynchronize(c->stream));
    *ptscale = c->d.h_scal[SC_PTSCALE];
    return HMEC_OK;
}

// ------------------------------------------------------- assembly
static AsmIn asm_in(hmec_ctx *c, const double *base, int64_t ns, double dt, double ptscale) {
    const hmec_params &P = c->P;
    AsmIn in;
#define FP(f) (base + (int64_t)(f) * ns)
    in.ETA = FP(HMEC_F_ETA); in.ETAP = FP(HMEC_F_ETAP); in.GGG = FP(HMEC_F_GGG);
    in.GGGP = FP(HMEC_F_GGGP); in.SXY0 = FP(HMEC_F_SXY0); in.SXX0 = FP(HMEC_F_SXX0);
    in.SYY0 = FP(HMEC_F_SYY0); in.RHOX = FP(HMEC_F_RHOX); in.RHO = FP(HMEC_F_RHO);
    in.RHOY = FP(HMEC_F_RHOY); in.VX0 = FP(HMEC_F_VX0); in.VY0 = FP(HMEC_F_VY0);
    in.POR = FP(HMEC_F_POR); in.ETAB = FP(HMEC_F_ETAB); in.GGGB = FP(HMEC_F_GGGB);
    in.PT0 = FP(HMEC_F_PT0); in.PF0 = FP(HMEC_F_PF0); in.DILP = FP(HMEC_F_DILP);
    in.RHOFX = FP(HMEC_F_RHOFX); in.ETADX = FP(HMEC_F_ETADX); in.PORX = FP(HMEC_F_PORX);
    in.VXF0 = FP(HMEC_F_VXF0); in.RHOFY = FP(HMEC_F_RHOFY); in.ETADY = FP(HMEC_F_ETADY);
    in.PORY = FP(HMEC_F_PORY); in.VYF0 = FP(HMEC_F_VYF0);
#undef FP
    in.Nx = c->Nx; in.Ny = c->Ny;
    in.dt = dt; in.dx = P.dx; in.dy = P.dy; in.ptscale = ptscale;
    in.gx = P.gx; in.gy = P.gy; in.inertia = P.inertia;
    in.BETASOLID = P.BETASOLID; in.BETAFLUID = P.BETAFLUID;
    in.bcupper = P.bcupper; in.bclower = P.bclower; in.bcvyflower = P.bcvyflower;
    in.PTFDIFF = P.PTFDIFF;
    return in;
}

struct PatternEmit {
    std::vector<int32_t> *ti, *tj;
    void add(int r, int cidx, double) { ti->push_back(r); tj->push_back(cidx); }
    void rhs(int, double) {}
};

// Fixed CSR pattern of the system (rows sorted by column, duplicates merged)
// and the triplet -> slot map, built on the host once per grid from the
// emission order (what Eigen's setFromTriplets + makeCompressed produce,
// h_mec_rsf_v9.cpp:1046-1049).
// Triplet emitter that only counts
struct CountEmit {
    int n = 0;
    void add(int, int, double) { n++; }
    void rhs(int, double) {}
};

// sec_out: per node, the first triplet of each row section relative to the
// node's first (8 bytes per node, sections 0..5 used)
int hmec_pattern_host(const hmec_params &P, std::vector<int32_t> &rowp,
                      std::vector<int32_t> &coli, std::vector<int32_t> *slot_out,
                      std::vector<int32_t> *node_trip_out, std::vector<uint8_t> *sec_out) {
    const int Nx1 = P.Nx + 1, Ny1 = P.Ny + 1;
    const int64_t ns = (int64_t)Nx1 * Ny1, N = ns * 6;
    std::vector<double> base((size_t)ns * HMEC_NFIELDS, 0.0);
    AsmIn in;
    {
        const double *b = base.data();
#define FP(f) (b + (int64_t)(f) * ns)
        in.ETA = FP(HMEC_F_ETA); in.ETAP = FP(HMEC_F_ETAP); in.GGG = FP(HMEC_F_GGG);
        in.GGGP = FP(HMEC_F_GGGP); in.SXY0 = FP(HMEC_F_SXY0); in.SXX0 = FP(HMEC_F_SXX0);
        in.SYY0 = FP(HMEC_F_SYY0); in.RHOX = FP(HMEC_F_RHOX); in.RHO = FP(HMEC_F_RHO);
        in.RHOY = FP(HMEC_F_RHOY); in.VX0 = FP(HMEC_F_VX0); in.VY0 = FP(HMEC_F_VY0);
        in.POR = FP(HMEC_F_POR); in.ETAB = FP(HMEC_F_ETAB); in.GGGB = FP(HMEC_F_GGGB);
        in.PT0 = FP(HMEC_F_PT0); in.PF0 = FP(HMEC_F_PF0); in.DILP = FP(HMEC_F_DILP);
        in.RHOFX = FP(HMEC_F_RHOFX); in.ETADX = FP(HMEC_F_ETADX); in.PORX = FP(HMEC_F_PORX);
        in.VXF0 = FP(HMEC_F_VXF0); in.RHOFY = FP(HMEC_F_RHOFY); in.ETADY = FP(HMEC_F_ETADY);
        in.PORY = FP(HMEC_F_PORY); in.VYF0 = FP(HMEC_F_VYF0);
#undef FP
        in.Nx = P.Nx; in.Ny = P.Ny;
        in.dt = 1; in.dx = P.dx; in.dy = P.dy; in.ptscale = 1;
        in.gx = P.gx; in.gy = P.gy; in.inertia = P.inertia;
        in.BETASOLID = P.BETASOLID; in.BETAFLUID = P.BETAFLUID;
        in.bcupper = P.bcupper; in.bclower = P.bclower; in.bcvyflower = P.bcvyflower;
        in.PTFDIFF = P.PTFDIFF;
    }
    std::vector<int32_t> ti, tj;
    ti.reserve(N * 8);
    tj.reserve(N * 8);
    PatternEmit e{&ti, &tj};
    std::vector<int32_t> node_trip((size_t)ns + 1, 0);
    for (int j = 0; j < Nx1; j++)
        for (int i = 0; i < Ny1; i++) {
            node_trip[(size_t)j * Ny1 + i] = (int32_t)ti.size();
            emit_node(i, j, in, e);
        }
    node_trip[(size_t)ns] = (int32_t)ti.size();
    if (sec_out) {
        sec_out->assign((size_t)ns * 8, 0);
        for (int j = 0; j < Nx1; j++)
            for (int i = 0; i < Ny1; i++) {
                const size_t t = (size_t)j * Ny1 + i;
                int off = 0;
                for (int q = 0; q < kAsmSections; q++) {
                    if (off > 255) return HMEC_E_ARG;
                    (*sec_out)[8 * t + q] = (uint8_t)off;
                    CountEmit ce;
                    emit_node(i, j, in, ce, 1 << q);
                    off += ce.n;
                }
                if (off != node_trip[t + 1] - node_trip[t]) {
                    hmec_set_last_error("assembly: row sections do not cover the node's triplets");
                    return HMEC_E_ARG;
                }
            }
    }
    const int64_t nt = (int64_t)ti.size();
    std::vector<int32_t> cnt(N + 1, 0);
    for (int64_t t = 0; t < nt; t++) cnt[ti[t] + 1]++;
    for (int64_t r = 0; r < N; r++) cnt[r + 1] += cnt[r];
    std::vector<int32_t> pos(cnt.begin(), cnt.end() - 1);
    std::vector<int64_t> ord(nt);
    for (int64_t t = 0; t < nt; t++) ord[pos[ti[t]]++] = t;
    for (int64_t r = 0; r < N; r++)
        for (int a = cnt[r] + 1; a < cnt[r + 1]; a++) {
            int64_t t = ord[a];
            int b = a - 1;
            while (b >= cnt[r] && tj[ord[b]] > tj[t]) { ord[b + 1] = ord[b]; b--; }
            ord[b + 1] = t;
        }
    rowp.assign(N + 1, 0);
    coli.clear();
    std::vector<int32_t> slot(nt);
    for (int64_t r = 0; r < N; r++) {
        rowp[r] = (int32_t)coli.size();
        int last = -1;
        for (int a = cnt[r]; a < cnt[r + 1]; a++) {
            const int64_t t = ord[a];
            if (tj[t] != last) { coli.push_back(tj[t]); last = tj[t]; }
            slot[t] = (int32_t)coli.size() - 1;
        }
    }
    rowp[N] = (int32_t)coli.size();
    if (slot_out) slot_out->swap(slot);
    if (node_trip_out) node_trip_out->swap(node_trip);
    return HMEC_OK;
}

int hmec_build_pattern(hmec_ctx *c) {
    const int64_t N = c->N;
    std::vector<int32_t> slot, node_trip;
    std::vector<uint8_t> sec;
    HMEC_TRY(hmec_pattern_host(c->P, c->h_rowp, c->h_coli, &slot, &node_trip, &sec));
    c->nnz = (int64_t)c->h_coli.size();
    const int64_t nt = (int64_t)slot.size();
    c->ntrip = nt;
    HMEC_HIP(hipMalloc(&c->d.rowp, sizeof(int32_t) * (N + 1)));
    HMEC_HIP(hipMalloc(&c->d.coli, sizeof(int32_t) * c->nnz));
    HMEC_HIP(hipMalloc(&c->d.val, sizeof(double) * c->nnz));
    HMEC_HIP(hipMalloc(&c->d.trip_slot, sizeof(int32_t) * nt));
    HMEC_HIP(hipMalloc(&c->d.node_trip, sizeof(int32_t) * node_trip.size()));
    HMEC_HIP(hipMalloc(&c->d.sec_off, sec.size()));
    HMEC_HIP(hipMemcpy(c->d.sec_off, sec.data(), sec.size(), hipMemcpyHostToDevice));
    HMEC_HIP(hipMemcpy(c->d.rowp, c->h_rowp.data(), sizeof(int32_t) * (N + 1), hipMemcpyHostToDevice));
    HMEC_HIP(hipMemcpy(c->d.coli, c->h_coli.data(), sizeof(int32_t) * c->nnz, hipMemcpyHostToDevice));
    HMEC_HIP(hipMemcpy(c->d.trip_slot, slot.data(), sizeof(int32_t) * nt, hipMemcpyHostToDevice));
    HMEC_HIP(hipMemcpy(c->d.node_trip, node_trip.data(), sizeof(int32_t) * node_trip.size(),
                       hipMemcpyHostToDevice));
    HMEC_HIP(hipMemset(c->d.val, 0, sizeof(double) * c->nnz));
    return HMEC_OK;
}

// Value emitter into the workgroup's LDS copy of its CSR range (the six rows
// of every node of the workgroup are consecutive in the CSR, and every triplet
// of a node lands in the node's own rows)
struct LdsEmit {
    double *val;            // LDS: CSR entries [base, ...)
    const int32_t *slot;
    int t, base;
    double *R;              // LDS: rows [rbase, rbase + 6 * nodes)
    int rbase;
    __device__ void add(int, int, double v) { val[slot[t++] - base] += v; }
    __device__ void rhs(int r, double v) { R[r - rbase] = v; }
};

// ASM_T nodes per workgroup, one wavefront per row section (wave s emits
// section s of the workgroup's nodes, from the section's first triplet): the
// triplets are summed in LDS in emission order (duplicates in the
// reference's triplet order, as setFromTriplets sums them; a section adds
// only to its own row, so the waves write disjoint entries), then the
// workgroup's contiguous CSR value range and its right-hand-side rows are
// stored with coalesced writes.  ptscale comes from the device scalars.
constexpr int ASM_T = 64;
constexpr int ASM_WG = ASM_T * kAsmSections;
__global__ void __launch_bounds__(ASM_WG)
k_assemble(AsmIn in, const double *__restrict__ scal, const int32_t *__restrict__ rowp,
           const int32_t *__restrict__ node_trip, const uint8_t *__restrict__ sec_off,
           const int32_t *__restrict__ slot, double *__restrict__ val, double *__restrict__ R,
           double *__restrict__ rscale, int Nx1, int Ny1, int dev) {
    extern __shared__ double lv[];
    if (dev) {
        if (scal[SC_EXIT] != 0) return;
        in.dt = scal[SC_CDT];
    }
    __shared__ double lr[6 * ASM_T];
    const int tid = threadIdx.x;
    const int sec = __builtin_amdgcn_readfirstlane(tid / ASM_T), q = tid % ASM_T;
    const int n0 = blockIdx.x * ASM_T;
    const int nn = min(ASM_T, Nx1 * Ny1 - n0);
    const int base = rowp[6 * n0], ne = rowp[6 * (n0 + nn)] - base;
    for (int p = tid; p < ne; p += ASM_WG) lv[p] = 0.;
    for (int r = tid; r < 6 * nn; r += ASM_WG) lr[r] = 0.;
    in.ptscale = scal[SC_PTSCALE];
    __syncthreads();
    if (q < nn) {
        const int t = n0 + q;
        LdsEmit e{lv, slot + node_trip[t] + sec_off[8 * t + sec], 0, base, lr, 6 * n0};
        emit_node(t % Ny1, t / Ny1, in, e, 1 << sec);
    }
    __syncthreads();
    for (int p = tid; p < ne; p += ASM_WG) val[base + p] = lv[p];
    for (int r = tid; r < 6 * nn; r += ASM_WG) R[6 * n0 + r] = lr[r];
    // power-of-two row scale of the section's row (the solver's
    // equilibration): sections vxs vys Pt vxD vyD Pf are rows 1 2 0 4 5 3
    if (q < nn) {
        constexpr int kRow[kAsmSections] = {1, 2, 0, 4, 5, 3};
        const int v = kRow[sec];
        const int64_t k = 6 * (int64_t)(n0 + q) + v;
        double m = 0;
        for (int p = rowp[k]; p < rowp[k + 1]; p++) m = fmax(m, fabs(lv[p - base]));
        rscale[k] = pow2_recip(m);
    }
}

int k_assemble(hmec_ctx *c, double dt, bool dev) {
    AsmIn in = asm_in(c, c->d.fields, c->ns, dt, 0.);
    const int n = c->Nx1 * c->Ny1;
    const int nblk = (n + ASM_T - 1) / ASM_T;
    if (c->asm_lds_entries == 0) {
        // largest CSR range of a workgroup (fixed with the pattern)
        int64_t mx = 0;
        for (int b = 0; b < nblk; b++) {
            const int n0 = b * ASM_T, n1 = std::min(n, n0 + ASM_T);
            mx = std::max<int64_t>(mx, c->h_rowp[6 * (int64_t)n1] - c->h_rowp[6 * (int64_t)n0]);
        }
        c->asm_lds_entries = mx;
    }
    const size_t lds = sizeof(double) * (size_t)c->asm_lds_entries;
    if (lds > 60 * 1024) {
        hmec_set_last_error("assembly: CSR range of a workgroup exceeds the LDS budget");
        return HMEC_E_ARG;
    }
    hipLaunchKernelGGL(k_assemble, dim3(nblk), dim3(ASM_WG), lds, c->stream, in, c->d.scal,
                       c->d.rowp, c->d.node_trip, (const uint8_t *)c->d.sec_off, c->d.trip_slot,
                       c->d.val, c->d.rhs, c->d.rscale, c->Nx1, c->Ny1, dev ? 1 : 0);
    LAUNCH_CHECK();
    c->rscale_ready = true;
    return HMEC_OK;
}

// ------------------------------------------------------- reload + stress
// h_mec_rsf_v9.cpp:1102-1127 (and :1480-1521 with ESP when final != 0).
// FROM_SOL (a Picard iteration): the same launch first reloads node t's
// unknowns from the solution vector S (:1060-1075; pressures scaled by
// ptscale from the device scalars) and the strain rates read the velocities
// from S, which holds exactly the values reloaded; otherwise (the final
// pass) from the fields.
// (gate: the device scalars when the launch waits on a solve decided on the
// device; it then does nothing while the exit flag holds)
template <bool FROM_SOL>
__global__ void k_stress(double *f, int64_t ns, int Nx, int Ny, double dx, double dy, double dt,
                         int final_pass, const double *S, const double *scal, const double *gate) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int Ny1 = Ny + 1;
    if (t >= (Nx + 1) * Ny1 || (gate && gate[SC_EXIT] == SC_EXIT_HOLD)) return;
    if (FROM_SOL) {
        const double sc = scal[SC_PTSCALE];
        const int64_t kp = (int64_t)t * 6;
        f[(int64_t)HMEC_F_PT * ns + t] = S[kp] * sc;
        f[(int64_t)HMEC_F_VXS * ns + t] = S[kp + 1];
        f[(int64_t)HMEC_F_VYS * ns + t] = S[kp + 2];
        f[(int64_t)HMEC_F_PF * ns + t] = S[kp + 3] * sc;
        f[(int64_t)HMEC_F_VXD * ns + t] = S[kp + 4];
        f[(int64_t)HMEC_F_VYD * ns + t] = S[kp + 5];
    }
    const double *vxf = f + (int64_t)HMEC_F_VXS * ns, *vyf = f + (int64_t)HMEC_F_VYS * ns;
    auto vxs = [&](int64_t q) { return FROM_SOL ? S[q * 6 + 1] : vxf[q]; };
    auto vys = [&](int64_t q) { return FROM_SOL ? S[q * 6 + 2] : vyf[q]; };
    const int i = t % Ny1, j = t / Ny1;
    if (i < Ny && j < Nx) {
        const int64_t b = i + (int64_t)j * Ny;
        const int64_t s = i + (int64_t)j * Ny1;
        const double dvxdy = (vxs(s + 1) - vxs(s)) / dy;
        const double dvydx = (vys(s + Ny1) - vys(s)) / dx;
        const double exy = .5 * (dvxdy + dvydx);
        if (final_pass) f[(int64_t)HMEC_F_ESP * ns + b] = .5 * (dvydx - dvxdy);
        f[(int64_t)HMEC_F_EXY * ns + b] = exy;
        const double eta = f[(int64_t)HMEC_F_ETA * ns + b];
        const double KXY = asm_divplus(dt * f[(int64_t)HMEC_F_GGG * ns + b], eta);
        f[(int64_t)HMEC_F_SXY * ns + b] = 2 * eta * exy * KXY + f[(int64_t)HMEC_F_SXY0 * ns + b] * (1 - KXY);
    }
    if (i >= 1 && i < Ny && j >= 1 && j < Nx) {
        const int64_t s = i + (int64_t)j * Ny1;
        const double dvx = vxs(s) - vxs(s - Ny1);
        const double dvy = vys(s) - vys(s - 1);
        const double exx = (2 * dvx / dx - dvy / dy) / 3.;
        const double eyy = (2 * dvy / dy - dvx / dx) / 3.;
        f[(int64_t)HMEC_F_EXX * ns + s] = exx;
        f[(int64_t)HMEC_F_EYY * ns + s] = eyy;
        const double etap = f[(int64_t)HMEC_F_ETAP * ns + s];
        const double KXX = asm_divplus(dt * f[(int64_t)HMEC_F_GGGP * ns + s], etap);
        f[(int64_t)HMEC_F_SXX * ns + s] = 2 * etap * exx * KXX + f[(int64_t)HMEC_F_SXX0 * ns + s] * (1 - KXX);
        f[(int64_t)HMEC_F_SYY * ns + s] = 2 * etap * eyy * KXX + f[(int64_t)HMEC_F_SYY0 * ns + s] * (1 - KXX);
    }
}

// --------------------------------------------------------- update reductions
// The rate-and-state fault solve of one fault node j (h_mec_rsf_v9.cpp:
// 1165-1308): the bisection for the node's stress, state, slip rate, yield
// stress and viscosity, and the node's dtlapusta candidate, bisection-cap
// flag, DSY^2 and yield flag.
// The bisection of the fault node's rate-and-state solve
// (h_mec_rsf_v9.cpp:1221-1243), emulated exactly by one wavefront: each round, lane q evaluates the midpoint of node q + 1
// of the complete binary tree of the next six sequential steps (the path
// bits of q + 1 below its leading one choose SIIB1 / SIIB2 := mid, as the
// sequential step does, from the round's bracket), then the lanes walk the
// tree in the sequential order with the sequential tests (cap, the sign
// rule, |DSIIB| > 1e-3).  Every midpoint and every evaluation is the
// sequential one, so SIIB4, V and the cap come out bit for bit; six steps
// cost one evaluation.
// lane l's value of v (l wave-uniform: two v_readlane, no LDS crossbar)
__device__ __forceinline__ double lane_read(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ void rsf_bisect_wave(double &SIIB1, double &SIIB2, const double DSIIB1,
                                                const double V0, const double arsf, const double prB,
                                                const double expf_, const double dx,
                                                const double eta0, const double siiel,
                                                const double G, const double dt, const int cap,
                                                double &SIIB4, double &V, int &capped) {
    constexpr int D = 6;
    const int nq = (threadIdx.x & 63) + 1;
    const int depth = 31 - __clz(nq);
    int it = 0;
    bool done = false;
    while (!done) {
        double s1 = SIIB1, s2 = SIIB2;
        for (int k = depth - 1; k >= 0; k--) {
            const double m = (s1 + s2) / 2.;
            if ((nq >> k) & 1) s1 = m;
            else s2 = m;
        }
        const double mid = (s1 + s2) / 2.;
        const double vq = 2 * V0 * sinh(fmax((mid), 0.) / arsf / prB) * expf_;
        const double eii = vq / dx / 2.;
        const double etapl = mid / 2. / eii;
        const double etavp = 1. / (1. / eta0 + 1. / etapl);
        const double dq = siiel * etavp / (G * dt + etavp) - mid;
        int node = 1;
        for (int k = 0; k < D; k++) {
            if (it >= cap) {
                capped = 1;
                done = true;
                break;
            }
            it++;
            const double m = lane_read(mid, node - 1), d = lane_read(dq, node - 1);
            SIIB4 = m;
            V = lane_read(vq, node - 1);
            const bool right = (d >= 0 && DSIIB1 >= 0) || (d <= 0 && DSIIB1 <= 0);
            if (right) SIIB1 = m;
            else SIIB2 = m;
            if (!(fabs(d) > 1e-3)) {
                done = true;
                break;
            }
            node = 2 * node + (right ? 1 : 0);
        }
    }
}

// The fault node (line_fault, j) by one wavefront (all lanes compute, lane 0
// stores); returns the new slip rate
__device__ __forceinline__ double rsf_node(double *f, int64_t ns, int Nx, int Ny, const hmec_params &P,
                                           double dt, int j, double *dsy2, int *flag, double *dtln,
                                           int *capn) {
    const bool st = (threadIdx.x & 63) == 0;
    const int Ny1 = Ny + 1;
    const int i = P.line_fault;
    const double V0 = P.V0, BETASOLID = P.BETASOLID, dx = P.dx;
    int capped = 0;
    double vout = 0;
    {
        const int64_t b = i + (int64_t)j * Ny;
        const int64_t s = i + (int64_t)j * Ny1;
#define BF(n) f[(int64_t)(n) * ns + b]
#define SB(n) (((f[(int64_t)(n) * ns + s] + f[(int64_t)(n) * ns + s + 1]) + f[(int64_t)(n) * ns + s + Ny1]) + f[(int64_t)(n) * ns + s + Ny1 + 1])
        const double arsf = BF(HMEC_F_ARSF), brsf = BF(HMEC_F_BRSF), lrsf = BF(HMEC_F_LRSF);
        const double fric = BF(HMEC_F_FRIC), eta0 = BF(HMEC_F_ETA0);
        const double sxx = SB(HMEC_F_SXX), syy = SB(HMEC_F_SYY);
        const double G = BF(HMEC_F_GGG), ETA = BF(HMEC_F_ETA), OMij = BF(HMEC_F_OM);
        const double sxy = BF(HMEC_F_SXY);
        const double SIIBij = sqrt(sxy * sxy + (sxx * sxx + syy * syy + (sxx + syy) * (sxx + syy)) / 32.);
        if (st) BF(HMEC_F_SIIB) = SIIBij;
        const double siiel = SIIBij / (ETA / (G * dt + ETA));
        const double prB_temp = (SB(HMEC_F_PT) - SB(HMEC_F_PF)) / 4.;
        double prB = prB_temp;
        if (prB < 1e3) prB = 1e3;
        const double expf_ = exp(-(brsf * OMij + fric) / arsf);
        double SIIB1 = SIIBij;
        double EIISLIP = V0 * sinh(fmax(SIIB1, 0.) / arsf / prB) * expf_ / dx;
        double ETAPL = SIIB1 / 2. / EIISLIP;
        double ETAVP = 1. / (1. / eta0 + 1. / ETAPL);
        double SIIB2 = siiel * ETAVP / (G * dt + ETAVP);
        const double DSIIB1 = SIIB2 - SIIB1;
        double V = 2 * V0 * sinh(fmax(SIIB2, 0.) / arsf / prB) * expf_;
        EIISLIP = V / dx / 2.;
        ETAPL = SIIB2 / 2. / EIISLIP;
        ETAVP = 1. / (1. / eta0 + 1. / ETAPL);
        const double DSIIB2 = siiel * ETAVP / (G * dt + ETAVP) - SIIB2;
        double SIIB4 = 0.;
        if ((DSIIB1 >= 0 && DSIIB2 <= 0) || (DSIIB1 <= 0 && DSIIB2 >= 0))
            rsf_bisect_wave(SIIB1, SIIB2, DSIIB1, V0, arsf, prB, expf_, dx, eta0, siiel, G, dt,
                            P.bisect_cap, SIIB4, V, capped);
        double om5;
        if (V * dt / lrsf > 1e-6) om5 = log(V0 / V + (exp(BF(HMEC_F_OM0)) - V0 / V) * exp(-V * dt / lrsf));
        else om5 = log(exp(BF(HMEC_F_OM0)) * (1 - V * dt / lrsf) + V0 * dt / lrsf);
        if (st) BF(HMEC_F_OM5) = om5;
        const double syield = fmax(P.syieldmin, prB_temp * arsf * asinh(V / 2. / V0 * exp((brsf * om5 + fric) / arsf)));
        double etapl = eta0 * syield / (eta0 * V + syield);
        if (st) BF(HMEC_F_SIGMAY) = syield;
        if (st) BF(HMEC_F_VSLIPB) = V;
        if (st) BF(HMEC_F_SII_FAULT) = SIIB4;
        const double g_temp = 2 * G;
        const double k = g_temp / (M_PI * (1 - ((3. / BETASOLID - g_temp) / (6. / BETASOLID + g_temp))) * dx);
        double dTETAmax;
        const double q = (k * lrsf / prB - brsf) / arsf - 1;
        if ((q * q / 4. - k * lrsf / arsf / prB) < 0)
            dTETAmax = fmin(1. - (brsf - arsf) * prB / (k * lrsf), .2);
        else
            dTETAmax = fmin(arsf * prB / (k * lrsf - (brsf - arsf) * prB), .2);
        const double dtl = dTETAmax * lrsf / V;
        int counted = 0;
        double d2 = 0;
        if (BF(HMEC_F_YNY0) > 0) {
            counted = 1;
            const double dsy = SIIBij - syield;
            if (st) BF(HMEC_F_DSY) = dsy;
            d2 = dsy * dsy;
        }
        const double A = syield / siiel;
        if (A < 1) {
            etapl = dt * G * A / (1 - A);
            if (etapl < eta0) {
                if (st) BF(HMEC_F_ETA5) = pow(etapl, 1 - P.etawt) * pow(ETA, P.etawt);
                if (st) BF(HMEC_F_YNY) = 1;
                if (!counted) {
                    counted = 1;
                    const double dsy = SIIBij - syield;
                    if (st) BF(HMEC_F_DSY) = dsy;
                    d2 = dsy * dsy;
                }
            } else {
                if (st) BF(HMEC_F_ETA5) = eta0;
            }
        } else {
            if (st) BF(HMEC_F_ETA5) = eta0;
        }
        if (st) {
            dsy2[j] = d2;
            flag[j] = counted;
            dtln[j] = dtl;
            capn[j] = capped;
        }
        vout = V;
#undef BF
#undef SB
    }
    return vout;
}

// One launch after the solve's reload and stresses.  Workgroups below nbn,
// thread t:
// - basic node t (t < Nx Ny): the previous slip rate into the block maximum
//   (Vmax before this iteration's fault solve, :1077); ETA5 = ETA0, DSY = YNY
//   = SigmaY = SII_fault = 0, OM5 = OM (:1152-1159); then the slip-rate
//   timestep candidate dx stpmax / VSLIPB (:1310-1436 inputs);
// - staggered node t: velocity extrema and fault-band maxima.
// With the fault active, the fault-row nodes are left to the workgroups from
// nbn on: one wavefront per fault node does the same for its node around the
// rate-and-state solve (rsf_node).
// Block partials (max / min, exact in any order) for k_update_final.
constexpr int UPD_T = 256;
constexpr int UPD_NQ = 8;   // Vmax, vx max / min, vy max / min, band x / y, dtslip
__global__ void __launch_bounds__(UPD_T) k_update_nodes(double *f, int64_t ns, int Nx, int Ny,
                                                        hmec_params P, double dt, int active,
                                                        const double *yvx, const double *yvy,
                                                        double dxstp, double *part, int nblk,
                                                        int nbn, double *dsy2, int *flag,
                                                        double *dtln, int *capn,
                                                        const double *gate) {
    __shared__ double red[UPD_NQ][UPD_T];
    if (gate && gate[SC_EXIT] == SC_EXIT_HOLD) return;
    const int tid = threadIdx.x, t = blockIdx.x * UPD_T + tid;
    const int Ny1 = Ny + 1, nb = Nx * Ny;
    double v[UPD_NQ] = {-INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, 0, 0, 1e30};
    double *VS = f + (int64_t)HMEC_F_VSLIPB * ns;
    auto reset = [&](int64_t q) {
        f[(int64_t)HMEC_F_ETA5 * ns + q] = f[(int64_t)HMEC_F_ETA0 * ns + q];
        f[(int64_t)HMEC_F_DSY * ns + q] = 0;
        f[(int64_t)HMEC_F_YNY * ns + q] = 0;
        f[(int64_t)HMEC_F_SIGMAY * ns + q] = 0;
        f[(int64_t)HMEC_F_SII_FAULT * ns + q] = 0;
        f[(int64_t)HMEC_F_OM5 * ns + q] = f[(int64_t)HMEC_F_OM * ns + q];
    };
    if ((int)blockIdx.x >= nbn) {
        // fault node j: one wavefront
        const int j = (blockIdx.x - nbn) * (UPD_T / 64) + (tid >> 6);
        if (j < Nx) {
            const int64_t q = P.line_fault + (int64_t)j * Ny;
            v[0] = VS[q];
            if ((tid & 63) == 0) reset(q);
            const double vs = rsf_node(f, ns, Nx, Ny, P, dt, j, dsy2, flag, dtln, capn);
            if (vs > 0) v[7] = dxstp / vs;
        }
    } else if (t < nb && !(active && t % Ny == P.line_fault)) {
        v[0] = VS[t];
        reset(t);
        const double vs = VS[t];
        if (vs > 0) v[7] = dxstp / vs;
    }
    if ((int)blockIdx.x < nbn && t < (Nx + 1) * Ny1) {
        const int i = t % Ny1;
        const double a = f[(int64_t)HMEC_F_VXS * ns + t], b = f[(int64_t)HMEC_F_VYS * ns + t];
        v[1] = a;
        v[2] = a;
        v[3] = b;
        v[4] = b;
        if (yvx[i] >= P.upper_block && yvx[i] <= P.lower_block) v[5] = fabs(a);
        if (yvy[i] >= P.upper_block && yvy[i] <= P.lower_block) v[6] = fabs(b);
    }
#pragma unroll
    for (int q = 0; q < UPD_NQ; q++) red[q][tid] = v[q];
    __syncthreads();
    for (int s = UPD_T / 2; s > 0; s >>= 1) {
        if (tid < s) {
            red[0][tid] = fmax(red[0][tid], red[0][tid + s]);
            red[1][tid] = fmax(red[1][tid], red[1][tid + s]);
            red[2][tid] = fmin(red[2][tid], red[2][tid + s]);
            red[3][tid] = fmax(red[3][tid], red[3][tid + s]);
            red[4][tid] = fmin(red[4][tid], red[4][tid + s]);
            red[5][tid] = fmax(red[5][tid], red[5][tid + s]);
            red[6][tid] = fmax(red[6][tid], red[6][tid + s]);
            red[7][tid] = fmin(red[7][tid], red[7][tid + s]);
        }
        __syncthreads();
    }
    if (tid < UPD_NQ) part[tid * nblk + blockIdx.x] = red[tid][0];
}

// The device-side Picard / dt controller on this iteration's scalars and the
// controller state in the device scalars, DSYLSQ in its own array; writes the
// new state and the exit flag (ynlast counts on when the loop goes on, :1455).
__device__ void controller_device(const hmec_params &P, int iterstep, double *scal, double *dsylsq) {
    hmec_iter_scalars sc;
    sc.ynpl = (int)scal[SC_YNPL];
    sc.ddd = scal[SC_DDD];
    sc.dtlapusta = scal[SC_DTLAP];
    sc.Vmax = scal[SC_VMAX];
    sc.vxs_max = scal[SC_VXMAX];
    sc.vxs_min = scal[SC_VXMIN];
    sc.vys_max = scal[SC_VYMAX];
    sc.vys_min = scal[SC_VYMIN];
    sc.maxvxs_band = scal[SC_BANDX];
    sc.maxvys_band = scal[SC_BANDY];
    sc.dtslip = scal[SC_DTSLIP];
    sc.bisect_capped = 0;
    CtlState s{scal[SC_CDT], scal[SC_CDTX], scal[SC_CDTY], sc.dtlapusta, scal[SC_CMAXVXY],
               (int)scal[SC_CYNLAST]};
    const bool stop = controller_eval(P, iterstep, sc, dsylsq, s);
    if (!stop) s.ynlast++;
    scal[SC_CDT] = s.dt;
    scal[SC_CDTX] = s.dtx;
    scal[SC_CDTY] = s.dty;
    scal[SC_CMAXVXY] = s.maxvxy;
    scal[SC_CYNLAST] = s.ynlast;
    scal[SC_EXIT] = stop ? 1.0 : 0.0;
}

// One workgroup: the update's scalars from k_update_nodes' block partials
// (Vmax, velocity extrema, band maxima, dtslip), the minimum of the fault
// nodes' dtlapusta candidates, the bisection caps, and ynpl / ddd summed
// sequentially in node order like the reference (from LDS chunks; an
// uncounted node has dsy2 = +0, and adding +0 to the non-negative sum is
// exact); then, for iterstep >= 0, the controller on them.
constexpr int UPF_T = 512;
__global__ void __launch_bounds__(UPF_T) k_update_final(int Nx, int active, double *scal,
                                                        const double *dsy2, const int *flag,
                                                        const double *dtln, const int *capn,
                                                        const double *part, int nblk, int iterstep,
                                                        hmec_params P, double *dsylsq, int gated) {
    if (gated && scal[SC_EXIT] == SC_EXIT_HOLD) return;
    __shared__ double red[UPD_NQ + 1][UPF_T];
    __shared__ int cnt[UPF_T];
    const int tid = threadIdx.x;
    // the partials' eight reductions and the dtlapusta minimum in one tree
    double v[UPD_NQ + 1] = {-INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, 0, 0, 1e30, 1e7};
    for (int b = tid; b < nblk; b += UPF_T) {
        v[0] = fmax(v[0], part[0 * nblk + b]);
        v[1] = fmax(v[1], part[1 * nblk + b]);
        v[2] = fmin(v[2], part[2 * nblk + b]);
        v[3] = fmax(v[3], part[3 * nblk + b]);
        v[4] = fmin(v[4], part[4 * nblk + b]);
        v[5] = fmax(v[5], part[5 * nblk + b]);
        v[6] = fmax(v[6], part[6 * nblk + b]);
        v[7] = fmin(v[7], part[7 * nblk + b]);
    }
    int cap = 0;
    if (active)
        for (int j = tid; j < Nx; j += UPF_T) {
            v[8] = fmin(v[8], dtln[j]);
            cap += capn[j];
        }
#pragma unroll
    for (int q = 0; q <= UPD_NQ; q++) red[q][tid] = v[q];
    cnt[tid] = cap;
    __syncthreads();
    for (int s = UPF_T / 2; s > 0; s >>= 1) {
        if (tid < s) {
            red[0][tid] = fmax(red[0][tid], red[0][tid + s]);
            red[1][tid] = fmax(red[1][tid], red[1][tid + s]);
            red[2][tid] = fmin(red[2][tid], red[2][tid + s]);
            red[3][tid] = fmax(red[3][tid], red[3][tid + s]);
            red[4][tid] = fmin(red[4][tid], red[4][tid + s]);
            red[5][tid] = fmax(red[5][tid], red[5][tid + s]);
            red[6][tid] = fmax(red[6][tid], red[6][tid + s]);
            red[7][tid] = fmin(red[7][tid], red[7][tid + s]);
            red[8][tid] = fmin(red[8][tid], red[8][tid + s]);
            cnt[tid] += cnt[tid + s];
        }
        __syncthreads();
    }
    double out[UPD_NQ];
#pragma unroll
    for (int q = 0; q < UPD_NQ; q++) out[q] = red[q][0];
    const double dtlap = red[8][0];
    const int capped = cnt[0];
    double *seq = red[0];   // reused below for the ordered sums
    int ynpl = 0;
    double ddd = 0;
    if (active) {
        for (int j0 = 0; j0 < Nx; j0 += UPF_T) {
            __syncthreads();
            if (j0 + tid < Nx) {
                seq[tid] = dsy2[j0 + tid];
                cnt[tid] = flag[j0 + tid];
            }
            __syncthreads();
            if (tid == 0) {
                const int n = min(UPF_T, Nx - j0);
                for (int q = 0; q < n; q++) {
                    ddd += seq[q];
                    ynpl += cnt[q];
                }
            }
        }
    }
    if (tid == 0) {
        scal[SC_VMAX] = out[0];
        scal[SC_VXMAX] = out[1];
        scal[SC_VXMIN] = out[2];
        scal[SC_VYMAX] = out[3];
        scal[SC_VYMIN] = out[4];
        scal[SC_BANDX] = out[5];
        scal[SC_BANDY] = out[6];
        scal[SC_DTSLIP] = out[7];
        scal[SC_DTLAP] = dtlap;
        scal[SC_CAPPED] = capped;
        scal[SC_YNPL] = ynpl;
        scal[SC_DDD] = ddd;
        if (iterstep >= 0) controller_device(P, iterstep, scal, dsylsq);
    }
}

// reload, stresses, fault solve and the update's reductions (three launches);
// iterstep >= 0: the controller as well; gated: every launch does nothing
// while the exit flag holds (a solve whose outcome the host decides)
int k_update_launch(hmec_ctx *c, double dt, int iterstep, bool gated) {
    const double *gate = gated ? (const double *)c->d.scal : nullptr;
    const hmec_params &P = c->P;
    const int n = c->Nx1 * c->Ny1;
    hipLaunchKernelGGL(k_stress<true>, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->d.fields,
                       c->ns, c->Nx, c->Ny, P.dx, P.dy, dt, 0, (const double *)c->d.sol,
                       (const double *)c->d.scal, gate);
    LAUNCH_CHECK();
    const int active = (c->timestep > P.tyield) ? 1 : 0;
    double *dsy2 = c->d.work2;
    int *flag = c->d.counters + 16;
    double *dtln = c->d.work2 + c->Nx;
    int *capn = flag + c->Nx;
    double *part = c->d.work;          // block partials (free between solves)
    const int nbn = (n + UPD_T - 1) / UPD_T;
    const int nblk = nbn + (active ? (c->Nx + UPD_T / 64 - 1) / (UPD_T / 64) : 0);
    Coords k = coords(c);
    hipLaunchKernelGGL(k_update_nodes, dim3(nblk), dim3(UPD_T), 0, c->stream, c->d.fields, c->ns, c->Nx,
                       c->Ny, P, dt, active, k.yvx, k.yvy, P.dx * P.stpmax, part, nblk, nbn, dsy2,
                       flag, dtln, capn, gate);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_update_final, dim3(1), dim3(UPF_T), 0, c->stream, c->Nx, active, c->d.scal,
                       (const double *)dsy2, (const int *)flag, (const double *)dtln, (const int *)capn,
                       (const double *)part, nblk, iterstep, P, c->d.dsylsq, gated ? 1 : 0);
    LAUNCH_CHECK();
    return HMEC_OK;
}

// the device scalars back to the host: the copy and an event after it are
// enqueued now (work enqueued later keeps the device busy), k_scalars_wait
// waits for the event alone and unpacks
int k_scalars_enqueue(hmec_ctx *c) {
    HMEC_HIP(hipMemcpyAsync(c->d.h_scal, c->d.scal, sizeof(double) * SC_COUNT,
                            hipMemcpyDeviceToHost, c->stream));
    if (!c->ev_scal) HMEC_HIP(hipEventCreateWithFlags(&c->ev_scal, hipEventDisableTiming));
    HMEC_HIP(hipEventRecord(c->ev_scal, c->stream));
    return HMEC_OK;
}

int k_scalars_wait(hmec_ctx *c, hmec_iter_scalars *out) {
    HMEC_HIP(hipEventSynchronize(c->ev_scal));
    const double *h = c->d.h_scal;
    out->ynpl = (int)h[SC_YNPL];
    out->ddd = h[SC_DDD];
    out->dtlapusta = h[SC_DTLAP];
    out->Vmax = h[SC_VMAX];
    out->vxs_max = h[SC_VXMAX];
    out->vxs_min = h[SC_VXMIN];
    out->vys_max = h[SC_VYMAX];
    out->vys_min = h[SC_VYMIN];
    out->maxvxs_band = h[SC_BANDX];
    out->maxvys_band = h[SC_BANDY];
    out->dtslip = h[SC_DTSLIP];
    out->bisect_capped = (int)h[SC_CAPPED];
    return HMEC_OK;
}

int k_update(hmec_ctx *c, double dt, hmec_iter_scalars *out) {
    HMEC_TRY(k_update_launch(c, dt, -1, false));
    HMEC_TRY(k_scalars_enqueue(c));
    return k_scalars_wait(c, out);
}

// the device controller's outcome after k_update_read: its state into the
// context, true when the exit test fired
bool k_controller_read(hmec_ctx *c) {
    const double *h = c->d.h_scal;
    c->dt = h[SC_CDT];
    c->dtx = h[SC_CDTX];
    c->dty = h[SC_CDTY];
    c->maxvxy = h[SC_CMAXVXY];
    c->ynlast = (int)h[SC_CYNLAST];
    return h[SC_EXIT] != 0;
}

// ------------------------------------------------ Picard commit
// ETA = ETA0 when ynpl == 0 (h_mec_rsf_v9.cpp:1315-1317); when not
// converged: ETA = max(min(ETA5, ETA0), etamin), ETA50 = ETA, YNY0 = YNY,
// OM = OM5 (:1440-1448).
// (ctl: ynpl and the exit decision from the device scalars instead)
__global__ void k_commit(double *f, int64_t ns, int nb, int ynpl, int converged, double etamin,
                         const double *ctl) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nb || (ctl && ctl[SC_EXIT] == SC_EXIT_HOLD)) return;
    if (ctl) {
        ynpl = (int)ctl[SC_YNPL];
        converged = ctl[SC_EXIT] != 0;
    }
    double *ETA = f + (int64_t)HMEC_F_ETA * ns;
    const double eta0 = f[(int64_t)HMEC_F_ETA0 * ns + t];
    if (ynpl == 0) ETA[t] = eta0;
    if (!converged) {
        const double e = fmax(fmin(f[(int64_t)HMEC_F_ETA5 * ns + t], eta0), etamin);
        ETA[t] = e;
        f[(int64_t)HMEC_F_ETA50 * ns + t] = e;
        f[(int64_t)HMEC_F_YNY0 * ns + t] = f[(int64_t)HMEC_F_YNY * ns + t];
        f[(int64_t)HMEC_F_OM * ns + t] = f[(int64_t)HMEC_F_OM5 * ns + t];
    }
}

int k_iter_commit(hmec_ctx *c, int ynpl, int converged) {
    const int nb = c->Nx * c->Ny;
    hipLaunchKernelGGL(k_commit, dim3((nb + 255) / 256), dim3(256), 0, c->stream, c->d.fields,
                       c->ns, nb, ynpl, converged, c->P.etamin, (const double *)nullptr);
    LAUNCH_CHECK();
    return HMEC_OK;
}

// ----------------------------------- device-side Picard / dt controller
// One thread: controller_device on the scalars the caller placed.
__global__ void k_controller(hmec_params P, int iterstep, double *scal, double *dsylsq) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    controller_device(P, iterstep, scal, dsylsq);
}

// the iteration epilogue with the device controller's decision (the
// controller ran at the end of k_update_launch)
int k_controller_commit(hmec_ctx *c) {
    const int nb = c->Nx * c->Ny;
    hipLaunchKernelGGL(k_commit, dim3((nb + 255) / 256), dim3(256), 0, c->stream, c->d.fields,
                       c->ns, nb, 0, 0, c->P.etamin, (const double *)c->d.scal);
    LAUNCH_CHECK();
    return HMEC_OK;
}

// the exit flag back to "go on" after the host decided a held solve (a
// one-thread kernel: no memset in the stream)
__global__ void k_exit_release(double *scal) {
    if (threadIdx.x == 0 && blockIdx.x == 0) scal[SC_EXIT] = 0;
}
int k_exit_release(hmec_ctx *c) {
    hipLaunchKernelGGL(k_exit_release, dim3(1), dim3(64), 0, c->stream, c->d.scal);
    LAUNCH_CHECK();
    return HMEC_OK;
}

// the controller's starting state of a timestep: dt (after the step's growth
// rule), ynlast = 0, no exit
int k_controller_begin(hmec_ctx *c) {
    double *h = c->d.h_scal;
    h[SC_CDT] = c->dt;
    h[SC_CDTX] = c->dtx;
    h[SC_CDTY] = c->dty;
    h[SC_CMAXVXY] = c->maxvxy;
    h[SC_CYNLAST] = 0;
    h[SC_EXIT] = 0;
    HMEC_HIP(hipMemcpyAsync(c->d.scal + SC_CDT, h + SC_CDT, sizeof(double) * (SC_EXIT - SC_CDT + 1),
                            hipMemcpyHostToDevice, c->stream));
    return HMEC_OK;
}

// One controller evaluation on caller-given scalars (stage-level callers):
// the scalars and the context's controller state up, k_controller, the
// outcome back into the context.  ptscale's slot is left alone.
int k_controller_stage(hmec_ctx *c, const hmec_iter_scalars &sc, int iterstep, bool *stop) {
    HMEC_HIP(hipStreamSynchronize(c->stream));   // no copy into the host mirror in flight
    double *h = c->d.h_scal;
    h[SC_YNPL] = sc.ynpl;
    h[SC_DDD] = sc.ddd;
    h[SC_DTLAP] = sc.dtlapusta;
    h[SC_VMAX] = sc.Vmax;
    h[SC_VXMAX] = sc.vxs_max;
    h[SC_VXMIN] = sc.vxs_min;
    h[SC_VYMAX] = sc.vys_max;
    h[SC_VYMIN] = sc.vys_min;
    h[SC_BANDX] = sc.maxvxs_band;
    h[SC_BANDY] = sc.maxvys_band;
    h[SC_DTSLIP] = sc.dtslip;
    h[SC_CAPPED] = sc.bisect_capped;
    h[SC_CDT] = c->dt;
    h[SC_CDTX] = c->dtx;
    h[SC_CDTY] = c->dty;
    h[SC_CMAXVXY] = c->maxvxy;
    h[SC_CYNLAST] = c->ynlast;
    h[SC_EXIT] = 0;
    const size_t bytes = sizeof(double) * (SC_COUNT - SC_YNPL);
    HMEC_HIP(hipMemcpyAsync(c->d.scal + SC_YNPL, h + SC_YNPL, bytes, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_controller, dim3(1), dim3(64), 0, c->stream, c->P, iterstep, c->d.scal,
                       c->d.dsylsq);
    LAUNCH_CHECK();
    HMEC_HIP(hipMemcpyAsync(h + SC_YNPL, c->d.scal + SC_YNPL, bytes, hipMemcpyDeviceToHost, c->stream));
    HMEC_HIP(hipStreamSynchronize(c->stream));
    c->dtlapusta = sc.dtlapusta;
    *stop = k_controller_read(c);
    return HMEC_OK;
}

// ------------------------------------------------ diagnostics (outputs)
// Pressure cells i in [1, Ny), j in [1, Nx) of the final stress state,
// h_mec_rsf_v9.cpp:1522-1558: EII, EIIVP, SII (the xy terms as sums of
// squares over the four surrounding basic nodes, in the reference's block
// order (i-1,j-1), (i-1,j), (i,j-1), (i,j)), DIS (its xy term summed in the
// order (i,j), (i-1,j), (i,j-1), (i-1,j-1)), pt_ave / pf_ave (mean with the
// cell below, the last row its own value), VIS_COMP and EL_DECOM.  The
// fields outside those cells stay 0.
__global__ void k_diagnostics(double *f, int64_t ns, int Nx, int Ny, double dt, double betasolid) {
    const int Ny1 = Ny + 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (Nx - 1) * (Ny - 1)) return;
    const int i = 1 + t % (Ny - 1), j = 1 + t / (Ny - 1);
    const int64_t s = i + (int64_t)j * Ny1;
    const int64_t b00 = (i - 1) + (int64_t)(j - 1) * Ny, b01 = (i - 1) + (int64_t)j * Ny;
    const int64_t b10 = i + (int64_t)(j - 1) * Ny, b11 = i + (int64_t)j * Ny;
#define FB(n, k) f[(int64_t)(n) * ns + (k)]
#define FS(n) f[(int64_t)(n) * ns + s]
    const double sxx = FS(HMEC_F_SXX), syy = FS(HMEC_F_SYY), etap = FS(HMEC_F_ETAP);
    const double exx = FS(HMEC_F_EXX);
    const double e00 = FB(HMEC_F_EXY, b00), e01 = FB(HMEC_F_EXY, b01);
    const double e10 = FB(HMEC_F_EXY, b10), e11 = FB(HMEC_F_EXY, b11);
    const double s00 = FB(HMEC_F_SXY, b00), s01 = FB(HMEC_F_SXY, b01);
    const double s10 = FB(HMEC_F_SXY, b10), s11 = FB(HMEC_F_SXY, b11);
    const double n00 = FB(HMEC_F_ETA, b00), n01 = FB(HMEC_F_ETA, b01);
    const double n10 = FB(HMEC_F_ETA, b10), n11 = FB(HMEC_F_ETA, b11);
    const double q00 = s00 / n00, q01 = s01 / n01, q10 = s10 / n10, q11 = s11 / n11;
    const double exy2 = ((e00 * e00 + e01 * e01) + e10 * e10) + e11 * e11;
    const double sxy2 = ((s00 * s00 + s01 * s01) + s10 * s10) + s11 * s11;
    const double q2 = ((q00 * q00 + q01 * q01) + q10 * q10) + q11 * q11;
    FS(HMEC_F_EII) = sqrt(exx * exx + exy2 / 4.);
    const double rx = sxx / (2 * etap), ry = syy / (2 * etap);
    FS(HMEC_F_EIIVP) = sqrt(.5 * (rx * rx + ry * ry) + q2 / 16.);
    FS(HMEC_F_SII) = sqrt(.5 * (sxx * sxx + syy * syy) + sxy2 / 4.);
    const double disxy = (((s11 * s11 / n11 + s01 * s01 / n01) + s10 * s10 / n10) + s00 * s00 / n00) / 2.;
    FS(HMEC_F_DIS) = (sxx * sxx / (2 * etap) + syy * syy / (2 * etap)) + disxy;
    const double *pt = f + (int64_t)HMEC_F_PT * ns, *pf = f + (int64_t)HMEC_F_PF * ns;
    const double *pt0 = f + (int64_t)HMEC_F_PT0 * ns, *pf0 = f + (int64_t)HMEC_F_PF0 * ns;
    double pta, pfa, pt0a, pf0a;
    if (i < Ny - 1) {
        pta = (pt[s] + pt[s + 1]) / 2.;
        pfa = (pf[s] + pf[s + 1]) / 2.;
        pt0a = (pt0[s] + pt0[s + 1]) / 2.;
        pf0a = (pf0[s] + pf0[s + 1]) / 2.;
    } else {
        pta = pt[s];
        pfa = pf[s];
        pt0a = pt0[s];
        pf0a = pf0[s];
    }
    FS(HMEC_F_PT_AVE) = pta;
    FS(HMEC_F_PF_AVE) = pfa;
    const double por = FS(HMEC_F_POR);
    FS(HMEC_F_VIS_COMP) = (pta - pfa) / (FS(HMEC_F_ETAB) * (1 - por));
    const double betadrained = (1 / FS(HMEC_F_GGGB) + betasolid) / (1 - por);
    const double kbw = 1 - betasolid / betadrained;
    FS(HMEC_F_EL_DECOM) = betadrained * (((pta - pt0a) - kbw * pfa) + kbw * pf0a) / dt;
#undef FB
#undef FS
}

int k_final_stress(hmec_ctx *c, double dt) {
    const int n = c->Nx1 * c->Ny1;
    hipLaunchKernelGGL(k_stress<false>, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->d.fields,
                       c->ns, c->Nx, c->Ny, c->P.dx, c->P.dy, dt, 1, (const double *)nullptr,
                       (const double *)nullptr, (const double *)nullptr);
    LAUNCH_CHECK();
    const int nc = (c->Nx - 1) * (c->Ny - 1);
    if (nc > 0) {
        hipLaunchKernelGGL(k_diagnostics, dim3((nc + 255) / 256), dim3(256), 0, c->stream, c->d.fields,
                           c->ns, c->Nx, c->Ny, dt, c->P.BETASOLID);
        LAUNCH_CHECK();
    }
    return HMEC_OK;
}

// ------------------------------------------------------ fault-row streams
// Row `line` of the fields the EVO_*.txt streams print
// (h_mec_rsf_v9.cpp:1751-1800), one launch: series k is field f[k] (minus
// field sub[k] when sub[k] >= 0: pt - pf) on a grid of rows[k] x cols[k],
// written to out[k * ld + j].
__global__ void k_fault_rows(const double *f, int64_t ns, int line, FaultRows a, int ld, double *out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (k >= FaultRows::NSERIES || j >= a.cols[k]) return;
    const int64_t o = line + (int64_t)j * a.rows[k];
    double v = f[(int64_t)a.f[k] * ns + o];
    if (a.sub[k] >= 0) v -= f[(int64_t)a.sub[k] * ns + o];
    out[(int64_t)k * ld + j] = v;
}

int k_fault_rows_gather(hmec_ctx *c, const FaultRows &a, double *dev_out) {
    const int ld = c->Nx1;
    hipLaunchKernelGGL(k_fault_rows, dim3((unsigned)((ld + 255) / 256), FaultRows::NSERIES), dim3(256), 0,
                       c->stream, c->d.fields, c->ns, c->P.line_fault, a, ld, dev_out);
    LAUNCH_CHECK();
    return HMEC_OK;
}

// ------------------------------------------------------ history
// h_mec_rsf_v9.cpp:1672-1726
__global__ void k_history(double *f, int64_t ns, int nstag, int nb) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nstag) return;
#define FF(n) f[(int64_t)(n) * ns + t]
    const double vx = FF(HMEC_F_VXS), vy = FF(HMEC_F_VYS);
    FF(HMEC_F_VX0) = vx;
    FF(HMEC_F_VY0) = vy;
    if (FF(HMEC_F_PORX) > 0) FF(HMEC_F_VXF0) = vx + FF(HMEC_F_VXD) / FF(HMEC_F_PORX);
    if (FF(HMEC_F_PORY) > 0) FF(HMEC_F_VYF0) = vy + FF(HMEC_F_VYD) / FF(HMEC_F_PORY);
    FF(HMEC_F_SXX0) = FF(HMEC_F_SXX);
    FF(HMEC_F_SYY0) = FF(HMEC_F_SYY);
    if (t < nb) FF(HMEC_F_SXY0) = FF(HMEC_F_SXY);
    const double pt = FF(HMEC_F_PT), pf = FF(HMEC_F_PF);
    FF(HMEC_F_PTF0) = pt - pf;
    FF(HMEC_F_PT0) = pt;
    FF(HMEC_F_PF0) = pf;
#undef FF
}

int k_history(hmec_ctx *c) {
    const int n = c->Nx1 * c->Ny1, nb = c->Nx * c->Ny;
    hipLaunchKernelGGL(k_history, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->d.fields,
                       c->ns, n, nb);
    LAUNCH_CHECK();
    return HMEC_OK;
}

// Per-step series of h_mec_rsf_v9.cpp:1661-1690: timesumcur, dtcur and the
// extrema of vxs over the nodes i < Ny, j < Nx with RHOX > 2000 and i != 0,
// of vys over the same range with RHOY > 2000 and j != 0 (start values -1e30
// and 1e30), stored at entry t of the six series (length nt).  Two stages:
// kSeriesB workgroups stride over the nodes (consecutive lanes on consecutive
// i, one column after another) and leave 4 partial extrema each in part;
// one workgroup reduces them.  Extrema are exact in any order.
constexpr int kSeriesT = 256;
__device__ __forceinline__ void series_tree(double (*red)[kSeriesT], int tid) {
    __syncthreads();
    for (int s = kSeriesT / 2; s > 0; s >>= 1) {
        if (tid < s) {
            red[0][tid] = fmax(red[0][tid], red[0][tid + s]);
            red[1][tid] = fmin(red[1][tid], red[1][tid + s]);
            red[2][tid] = fmax(red[2][tid], red[2][tid + s]);
            red[3][tid] = fmin(red[3][tid], red[3][tid + s]);
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kSeriesT) k_series_part(const double *f, int64_t ns, int Nx, int Ny,
                                                          double *part) {
    __shared__ double red[4][kSeriesT];
    const int Ny1 = Ny + 1, tid = threadIdx.x;
    const double *vxs = f + (int64_t)HMEC_F_VXS * ns, *vys = f + (int64_t)HMEC_F_VYS * ns;
    const double *rhox = f + (int64_t)HMEC_F_RHOX * ns, *rhoy = f + (int64_t)HMEC_F_RHOY * ns;
    double mx = -1e30, nx = 1e30, my = -1e30, ny = 1e30;
    const int n = Nx * Ny;
    for (int q = blockIdx.x * kSeriesT + tid; q < n; q += gridDim.x * kSeriesT) {
        const int i = q % Ny, j = q / Ny;
        const int64_t s = i + (int64_t)j * Ny1;
        if (rhox[s] > 2000 && i != 0) {
            mx = fmax(mx, vxs[s]);
            nx = fmin(nx, vxs[s]);
        }
        if (rhoy[s] > 2000 && j != 0) {
            my = fmax(my, vys[s]);
            ny = fmin(ny, vys[s]);
        }
    }
    red[0][tid] = mx; red[1][tid] = nx; red[2][tid] = my; red[3][tid] = ny;
    series_tree(red, tid);
    if (tid < 4) part[tid * gridDim.x + blockIdx.x] = red[tid][0];
}

__global__ void __launch_bounds__(kSeriesT) k_series_final(const double *part, int nparts,
                                                           double *series, int64_t nt, int64_t t,
                                                           double timesum, double dt) {
    __shared__ double red[4][kSeriesT];
    const int tid = threadIdx.x;
    double v[4] = {-1e30, 1e30, -1e30, 1e30};
    for (int b = tid; b < nparts; b += kSeriesT) {
        v[0] = fmax(v[0], part[b]);
        v[1] = fmin(v[1], part[nparts + b]);
        v[2] = fmax(v[2], part[2 * nparts + b]);
        v[3] = fmin(v[3], part[3 * nparts + b]);
    }
    for (int q = 0; q < 4; q++) red[q][tid] = v[q];
    series_tree(red, tid);
    if (tid == 0) {
        series[t] = timesum;
        series[nt + t] = dt;
        for (int q = 0; q < 4; q++) series[(2 + q) * nt + t] = red[q][0];
    }
}

int k_series_step(hmec_ctx *c, double timesum, double dt) {
    const int64_t t = c->timestep - 1;
    if (t < 0 || t >= kNumTimesteps) return HMEC_OK;     // beyond the reference's series length
    const int n = c->Nx * c->Ny;
    const int nb = std::max(1, std::min(kSeriesBlocks, (n + kSeriesT - 1) / kSeriesT));
    double *part = c->d.series + 6 * kNumTimesteps;
    hipLaunchKernelGGL(k_series_part, dim3(nb), dim3(kSeriesT), 0, c->stream, c->d.fields, c->ns,
                       c->Nx, c->Ny, part);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_series_final, dim3(1), dim3(kSeriesT), 0, c->stream, (const double *)part,
                       nb, c->d.series, kNumTimesteps, t, timesum, dt);
    LAUNCH_CHECK();
    return HMEC_OK;
}

// ---------------------------------------------------- advection
// Node -> marker bilinear gather of vx, vy (vx / vy grids) and spin (basic
// grid) with the reference's RK stages (h_mec_rsf_v9.cpp:1563-1656: stage 3
// samples stage 2's position), stress rotation and periodic recycling in x.
__global__ void k_advect(double *mk, int64_t M, const double *__restrict__ f, int64_t ns, Coords k,
                         int Nx, int Ny, double dx, double dy, double dt, double xsize) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    const int Ny1 = Ny + 1;
    double *xm = mk + (int64_t)HMEC_M_XM * M, *ym = mk + (int64_t)HMEC_M_YM * M;
    const double *vxs = f + (int64_t)HMEC_F_VXS * ns, *vys = f + (int64_t)HMEC_F_VYS * ns;
    const double *ESP = f + (int64_t)HMEC_F_ESP * ns;
    const double xold = xm[m], yold = ym[m];
    double X = xold, Y = yold;
    double vxm[4], vym[4], spm0 = 0;
    for (int rk = 0; rk < 4; rk++) {
        int j = d_check_bounds(d_fix(X / dx), Nx);
        int i = d_check_bounds(d_fix((Y + dy / 2.) / dy), Ny);
        double dxm = (X - k.xvx[j]) / dx, dym = (Y - k.yvx[i]) / dy;
        int64_t s = i + (int64_t)j * Ny1;
        vxm[rk] = vxs[s] * (1 - dxm) * (1 - dym) + vxs[s + 1] * (1 - dxm) * dym +
                  vxs[s + Ny1] * dxm * (1 - dym) + vxs[s + Ny1 + 1] * dxm * dym;
        j = d_check_bounds(d_fix((X + dx / 2.) / dx), Nx);
        i = d_check_bounds(d_fix(Y / dy), Ny);
        dxm = (X - k.xvy[j]) / dx;
        dym = (Y - k.yvy[i]) / dy;
        s = i + (int64_t)j * Ny1;
        vym[rk] = vys[s] * (1 - dxm) * (1 - dym) + vys[s + 1] * (1 - dxm) * dym +
                  vys[s + Ny1] * dxm * (1 - dym) + vys[s + Ny1 + 1] * dxm * dym;
        if (rk == 0) {
            j = d_check_bounds(d_fix((X) / dx), Nx);
            i = d_check_bounds(d_fix((Y) / dy), Ny);
            dxm = (X - k.x[j]) / dx;
            dym = (Y - k.y[i]) / dy;
            const int64_t b = i + (int64_t)j * Ny;
            spm0 = ESP[b] * (1 - dxm) * (1 - dym) + ESP[b + 1] * (1 - dxm) * dym +
                   ESP[b + Ny] * dxm * (1 - dym) + ESP[b + Ny + 1] * dxm * dym;
        }
        if (rk < 2) {
            X = xold + vxm[rk] * dt / 2.;
            Y = yold + vym[rk] * dt / 2.;
        } else if (rk == 3) {
            X = xold + vxm[rk] * dt;
            Y = yold + vym[rk] * dt;
        }
    }
    const double dspeff = spm0 * dt;
    double *sxxm = mk + (int64_t)HMEC_M_SXXM * M, *syym = mk + (int64_t)HMEC_M_SYYM * M,
           *sxym = mk + (int64_t)HMEC_M_SXYM * M;
    const double msxxold = sxxm[m], msyyold = syym[m], msxyold = sxym[m];
    const double s2 = sin(2 * dspeff), c2 = cos(2 * dspeff), cs = cos(dspeff), sn = sin(dspeff);
    sxym[m] = .5 * (msxxold - msyyold) * s2 + msxyold * c2;
    sxxm[m] = msxxold * (cs * cs) + msyyold * (sn * sn) - msxyold * s2;
    syym[m] = msxxold * (sn * sn) + msyyold * (cs * cs) + msxyold * s2;
    double xn = xold + (vxm[0] + 2 * vxm[1] + 2 * vxm[2] + vxm[3]) / 6. * dt;
    const double yn = yold + (vym[0] + 2 * vym[1] + 2 * vym[2] + vym[3]) / 6. * dt;
    if (xn < 0) xn = xn + xsize;
    if (xn > xsize) xn = xn - xsize;
    xm[m] = xn;
    ym[m] = yn;
}

int k_advect(hmec_ctx *c, double dt) {
    if (!c->have_markers || c->M == 0) return HMEC_OK;
    const int64_t M = c->M;
    hipLaunchKernelGGL(k_advect, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, c->stream,
                       c->d.mk, M, c->d.fields, c->ns, coords(c), c->Nx, c->Ny, c->P.dx, c->P.dy,
                       dt, c->P.xsize);
    LAUNCH_CHECK();
    c->markers_dirty = true;
    return HMEC_OK;
}
