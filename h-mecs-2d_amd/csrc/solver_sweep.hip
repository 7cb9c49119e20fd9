// Sweep form of the multifrontal factor: the triangular solves of
// PardisoLU::solve (h_mec_rsf_v9.cpp:1052) as one matrix-vector product per
// front and direction.
//
// For a front with n = p + u rows, factored pivot block L11 U11 (unit lower /
// upper), L21 (u x p) and U12 (p x u), the substitutions of one front are
//   forward   y1 = L11^-1 b1,      c = b2 - L21 y1          (c -> parent)
//   backward  x1 = U11^-1 (y1 - U12 x2)
// With W = L21 L11^-1 and V = U11^-1 U12 (the off-diagonal blocks of the
// inverses of the unit-extended n x n triangles) both become ONE product
// whose input is complete before the front starts:
//   [y1; c] = [b1; b2] + Fm b1      Fm = [strict_lower(L11^-1); -W]   n x p
//   x1      = Bm [y1; x2]           Bm = [U11^-1 | -V]                 p x n
// Fm and Bm of every front are formed once per factorization by one MFMA
// launch that reads the triangle inverses (kept in place in the pivot block
// by the factorization), L21 and U12, and writes a separate sweep arena with
// both blocks ROW-major, rows padded to 16 bytes: every row is one contiguous
// dot product, so a sweep parallelizes over rows without splitting K (no
// partial sums, no inter-workgroup synchronization) and reads every stored
// nonzero entry once with 16-byte vector loads.
//
// A sweep is one launch per level and direction; a workgroup owns a block of
// rows of one front, loads the front's input vector into LDS once, and a
// group of L lanes (L chosen per level from the row length) forms each row's
// dot product, summed by fixed DPP lane moves.  Forward, every input is
// contiguous and in front order, so a level's loads depend only on its task:
// the right-hand side at the fronts' pivots arrives in front order (zf,
// written by the operator that forms it), and a child stores each
// contribution row into one of two per-row slots of its parent's input (slot
// 0: the first child, in plan order, that reaches the row; a slot no child
// fills stays 0), which the parent adds.  Backward, a front reads its update
// rows' values from z through their dofs (pushing them into the children's
// inputs instead measured slower: its extra stores cost more than the index
// load they save).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "hmec_internal.h"
#include "mfma_traits.h"
#include "dpp_sum.h"


#define SW_LAUNCH_CHECK()                                                   \
    do {                                                                    \
        hipError_t e_ = hipGetLastError();                                  \
        if (e_ != hipSuccess) {                                             \
            hmec_set_last_error(std::string("launch: ") + hipGetErrorString(e_)); \
            return HMEC_E_HIP;                                              \
        }                                                                   \
    } while (0)

namespace {

constexpr int SW_T = 256;      // threads per sweep workgroup (64 on small-front levels)
constexpr int SW_RPG = 2;      // rows in flight per lane group (64 VGPRs: 8 waves per SIMD)
constexpr int SW_PASSES = 4;   // most row passes per full-size workgroup
constexpr int SW_FILL = 1024;  // workgroups a level keeps at least (4 per CU)
constexpr int SW_SMALL = 256;  // levels whose fronts have at most this many rows: one-wave workgroups
constexpr int SW_TILE = 64;    // sweep-form GEMM tile
constexpr int SW_TK = 16;      // its K chunk

struct SwFront {
    int64_t aoff;    // front offset in the factor arena (n x n, ld n)
    int64_t soff;    // Fm offset in the sweep arena (n rows x ldf)
    int64_t boff;    // Bm offset in the sweep arena (p rows x ldb)
    int64_t voff;    // front vector offset in the per-front vector arena
    int64_t dofo;    // offset of the front's global dof list
    int64_t goff;    // offset of the front's gather row pointers (n + 1)
    int n, p;
    int ldf, ldb;    // padded row lengths (multiples of 4 entries)
};

// one workgroup's rows of one front, with what the kernel needs of the
// front (no second dependent load for a front descriptor)
struct SwTask {
    int64_t moff;    // Fm (forward) / Bm (backward) offset in the sweep arena
    int64_t voff;    // front vector offset in the per-front vector arena
    int64_t goff;    // first row source record of the front
    int n, p, ld;    // front order, pivots, padded row length of the block
    int r0, nr;      // row block
    int kids;        // the front has children (forward: contributions to gather)
};

// Per front row (indexed from the front's goff): zidx, the row's dof in the
// condensed system; fdest, where a contribution row (r >= p) stores its value
// in the parent's slots (-1: no parent).

struct SwTile {
    int f;
    int kind;        // 0: -W tile, 1: -V tile, 2: pivot-block copy tile
    int ti, tj;
};

template <typename T>
struct SwMem {       // one member of a (batched) sweep-form launch
    const T *arena;  // factor arena
    T *sa;           // sweep arena
};

// ------------------------------------------------------------ sweep form
template <typename T>
__global__ void __launch_bounds__(256) k_sweep_form(const SwTile *tiles, const SwFront *fronts,
                                                    const SwMem<T> *mem) {
    const T *const arena = mem[blockIdx.z].arena;
    T *const sa = mem[blockIdx.z].sa;
    typedef typename Mfma<T>::acc_t acc_t;
    __shared__ T As[SW_TK][SW_TILE + 4];
    __shared__ T Bs[SW_TK][SW_TILE + 4];
    const SwTile tl = tiles[blockIdx.x];
    const SwFront F = fronts[tl.f];
    const int n = F.n, p = F.p, u = n - p;
    const T *A = arena + F.aoff;
    T *Fm = sa + F.soff;
    T *Bm = sa + F.boff;
    const int64_t ldf = F.ldf, ldb = F.ldb;
    const int tid = threadIdx.x;
    if (tl.kind == 2) {
        // strict lower part of L11^-1 -> Fm rows < p; U11^-1 -> Bm columns < p
        const int r0 = tl.ti * SW_TILE, c0 = tl.tj * SW_TILE;
        for (int e = tid; e < SW_TILE * SW_TILE; e += blockDim.x) {
            const int r = r0 + e % SW_TILE, k = c0 + e / SW_TILE;
            if (r >= p || k >= p) continue;
            const T v = A[(int64_t)k * n + r];
            Fm[(int64_t)r * ldf + k] = r > k ? v : (T)0;
            Bm[(int64_t)r * ldb + k] = r <= k ? v : (T)0;
        }
        return;
    }
    // kind 0: C = -(L21 * L11^-1)   rows of L21, K from the tile's first column
    // kind 1: C = -(U11^-1 * U12)   rows of U11^-1, K from the tile's first row
    const bool w = tl.kind == 0;
    const int m = w ? min(SW_TILE, u - tl.ti * SW_TILE) : min(SW_TILE, p - tl.ti * SW_TILE);
    const int nn = w ? min(SW_TILE, p - tl.tj * SW_TILE) : min(SW_TILE, u - tl.tj * SW_TILE);
    const int kb = w ? tl.tj * SW_TILE : tl.ti * SW_TILE;
    const int rbase = tl.ti * SW_TILE, cbase = tl.tj * SW_TILE;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = (wave & 1) * 32, wc = (wave >> 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    acc_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = acc_t{0, 0, 0, 0};
    // the next K chunk is loaded into registers while the current one is
    // multiplied from LDS
    constexpr int PER = SW_TK * SW_TILE / 256;
    T ra[PER], rb[PER];
    auto fetch = [&](int kk) {
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int e = tid + q * 256;
            // A operand, row index fastest (column-major source)
            const int r = e % SW_TILE, k = e / SW_TILE, kg = kk + k;
            T va = (T)0;
            if (r < m && kg < p) {
                if (w) {
                    va = A[(int64_t)kg * n + p + rbase + r];                    // L21[r, kg]
                } else {
                    const int i = rbase + r;
                    va = kg >= i ? A[(int64_t)kg * n + i] : (T)0;                // U11^-1[i, kg]
                }
            }
            ra[q] = va;
            // B operand, K index fastest
            const int cc = e / SW_TK, k2 = e % SW_TK, kg2 = kk + k2;
            T vb = (T)0;
            if (cc < nn && kg2 < p) {
                if (w) {
                    const int j = cbase + cc;                                    // L11^-1[kg2, j]
                    vb = kg2 > j ? A[(int64_t)j * n + kg2] : (kg2 == j ? (T)1 : (T)0);
                } else {
                    vb = A[(int64_t)(p + cbase + cc) * n + kg2];                 // U12[kg2, cc]
                }
            }
            rb[q] = vb;
        }
    };
    if (kb < p) fetch(kb);
    for (int kk = kb; kk < p; kk += SW_TK) {
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int e = tid + q * 256;
            As[e / SW_TILE][e % SW_TILE] = ra[q];
            Bs[e % SW_TK][e / SW_TK] = rb[q];
        }
        __syncthreads();
        if (kk + SW_TK < p) fetch(kk + SW_TK);
#pragma unroll
        for (int k4 = 0; k4 < SW_TK; k4 += 4) {
            const T a0 = As[k4 + lk][wr + li], a1 = As[k4 + lk][wr + 16 + li];
            const T b0 = Bs[k4 + lk][wc + li], b1 = Bs[k4 + lk][wc + 16 + li];
            acc[0][0] = Mfma<T>::mma(a0, b0, acc[0][0]);
            acc[0][1] = Mfma<T>::mma(a0, b1, acc[0][1]);
            acc[1][0] = Mfma<T>::mma(a1, b0, acc[1][0]);
            acc[1][1] = Mfma<T>::mma(a1, b1, acc[1][1]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = wr + a * 16 + Mfma<T>::row(lane, r), col = wc + b * 16 + li;
                if (row >= m || col >= nn) continue;
                if (w) Fm[(int64_t)(p + rbase + row) * ldf + cbase + col] = -acc[a][b][r];
                else Bm[(int64_t)(rbase + row) * ldb + p + cbase + col] = -acc[a][b][r];
            }
}

// ------------------------------------------------------------ sweep
template <typename T> struct Vec16;
template <> struct Vec16<float> {
    typedef float4 type;
    static constexpr int W = 4;
    static __device__ __forceinline__ double dot(const type &v, const double *x) {
        return (double)v.x * x[0] + (double)v.y * x[1] + (double)v.z * x[2] + (double)v.w * x[3];
    }
};
template <> struct Vec16<double> {
    typedef double2 type;
    static constexpr int W = 2;
    static __device__ __forceinline__ double dot(const type &v, const double *x) {
        return v.x * x[0] + v.y * x[1];
    }
};

// DIR 0 (forward, leaves -> root): x = b1 gathered from z (scaled right-hand
// side at the front's pivot dofs) and the children's contributions;
// out[r] = b[r] + Fm[r,:] x -> per-front vector (y1 rows < p, contribution c
// rows >= p).
// DIR 1 (backward, root -> leaves): x = [y1; z(update dofs)],
// out[r] = Bm[r,:] x -> z at the front's pivot dofs.
// Dynamic LDS: xs (K entries padded to a multiple of 4, zero padded) then,
// forward only, the gathered b of the block's rows.
// Row descriptors of one pass of a lane group: RB rows q0, q0 + NG, ...,
// with the store targets of the rows (loaded with the descriptors, so the
// stores do not wait on an index load)
template <typename T, int DIR, int RB, int NG>
struct SwPass {
    typedef typename Vec16<T>::type VT;
    int cb[RB], ce[RB];
    const VT *row[RB];
    int cmin, cmax;
    int zdst[RB];        // backward: dof of the row
    int fdst[RB];        // forward: parent slot entry of a contribution row
    __device__ __forceinline__ void setup(const SwTask &t, const T *base, int q0,
                                          const int32_t *zi, const int32_t *fd) {
        cmin = 1 << 30;
        cmax = 0;
#pragma unroll
        for (int u = 0; u < RB; u++) {
            const int q = q0 + u * NG;
            const int r = t.r0 + q;
            fdst[u] = -1;
            zdst[u] = 0;
            if (q < t.nr) {
                if (DIR == 0) {
                    cb[u] = 0;
                    ce[u] = r < t.p ? r : t.p;      // Fm rows < p: strictly lower
                    if (r >= t.p) fdst[u] = fd[r];
                } else {
                    cb[u] = r & ~3;                 // Bm: upper in the pivot columns
                    ce[u] = t.n;
                    zdst[u] = zi[r];
                }
            } else {
                cb[u] = 0;
                ce[u] = 0;
            }
            row[u] = reinterpret_cast<const VT *>(base + (int64_t)(q < t.nr ? r : t.r0) * t.ld);
            if (ce[u] > cb[u]) {
                cmin = min(cmin, cb[u]);
                cmax = max(cmax, ce[u]);
            }
        }
    }
    // the rows' vectors at column c (zero outside a row's own [cb, ce))
    __device__ __forceinline__ void load(int c, VT *v) const {
#pragma unroll
        for (int u = 0; u < RB; u++)
            v[u] = (c >= cb[u] && c < ce[u]) ? row[u][c / Vec16<T>::W] : VT{};
    }
};

// Sweep state shared by the level launches of a context
struct SwVec {
    const int32_t *zidx;
    const int32_t *fdest;
    const double *zf;    // right-hand side at the fronts' pivots, front order
    double *z;           // backward: the ancestors' values in, the solution out (condensed dofs)
    double *wbuf;        // forward y1 (pivot rows) of every front
    double *xin;         // forward contribution slots: [0, varena) and [varena, 2 varena);
                         // slots no child fills stay 0
    int64_t varena;
};

// DIR 0 (forward, leaves -> root): x = b1 (right-hand side at the front's
// pivots plus the children's contribution slots); out[r] = b[r] + Fm[r,:] x:
// rows < p (y1) -> wbuf, rows >= p (the contribution c) -> the parent's slot.
// DIR 1 (backward, root -> leaves): x = [y1; x2] (x2: the ancestors' values
// at the update rows' dofs); out[r] = Bm[r,:] x -> z at the row's dof.
// Dynamic LDS: xs (K entries padded to a multiple of 4, zero padded) then,
// forward only, the b of the block's rows.  The first pass's first column
// chunk is loaded before the input, so the matrix loads overlap it.
template <typename T, int DIR, int L, int NT>
__global__ void __launch_bounds__(NT) k_sweep(const SwTask *tasks, const T *sa, SwVec sv) {
    extern __shared__ double xs[];
    typedef Vec16<T> V;
    typedef typename V::type VT;
    constexpr int NG = NT / L;
    constexpr int RB = SW_RPG;                     // rows in flight per lane group
    constexpr int STEP = V::W * L;
    const SwTask t = tasks[blockIdx.x];
    const int tid = threadIdx.x;
    const int lane = tid % L, grp = tid / L;
    const T *base = sa + t.moff;
    const int32_t *zi = sv.zidx + t.goff;
    const int32_t *fd = sv.fdest + t.goff;
    SwPass<T, DIR, RB, NG> ps;
    ps.setup(t, base, grp, zi, fd);
    int c = (ps.cmin < ps.cmax ? ps.cmin : 0) + V::W * lane;
    VT v0[RB], v1[RB];
    ps.load(c, v0);
    ps.load(c + STEP, v1);

    const double *x0 = sv.xin + t.voff, *x1 = sv.xin + sv.varena + t.voff;
    const int K = DIR == 0 ? t.p : t.n;
    const int Kp = (K + 3) & ~3;
    for (int k = tid; k < Kp; k += NT) {
        double v = 0.0;
        if (k < K) {
            if (DIR == 0) {
                // a slot no child fills holds 0 (set once with the plan)
                v = sv.zf[t.voff + k];
                if (t.kids) v = (v + x0[k]) + x1[k];
            } else {
                v = k < t.p ? sv.wbuf[t.voff + k] : sv.z[zi[k]];
            }
        }
        xs[k] = v;
    }
    double *bi = xs + Kp;      // forward: b of rows >= p of this block
    if (DIR == 0) {
        for (int q = tid; q < t.nr; q += NT) {
            const int r = t.r0 + q;
            double v = 0.0;
            if (r >= t.p && t.kids) v = (v + x0[r]) + x1[r];
            bi[q] = v;
        }
    }
    __syncthreads();
    for (int q0 = grp; q0 < t.nr; q0 += NG * RB) {
        if (q0 != grp) {
            ps.setup(t, base, q0, zi, fd);
            c = (ps.cmin < ps.cmax ? ps.cmin : 0) + V::W * lane;
            ps.load(c, v0);
            ps.load(c + STEP, v1);
        }
        double acc[RB];
#pragma unroll
        for (int u = 0; u < RB; u++) acc[u] = 0.0;
        // each lane walks the rows' common column range; a row only loads the
        // vectors of its own [cb, ce) range (cb, ce, STEP multiples of 4)
        if (c + STEP < ps.cmax) {
#pragma unroll
            for (int u = 0; u < RB; u++) acc[u] += V::dot(v0[u], xs + c) + V::dot(v1[u], xs + c + STEP);
            c += 2 * STEP;
            for (; c + STEP < ps.cmax; c += 2 * STEP) {
                ps.load(c, v0);
                ps.load(c + STEP, v1);
#pragma unroll
                for (int u = 0; u < RB; u++)
                    acc[u] += V::dot(v0[u], xs + c) + V::dot(v1[u], xs + c + STEP);
            }
            if (c < ps.cmax) {
                ps.load(c, v0);
#pragma unroll
                for (int u = 0; u < RB; u++) acc[u] += V::dot(v0[u], xs + c);
            }
        } else if (c < ps.cmax) {
#pragma unroll
            for (int u = 0; u < RB; u++) acc[u] += V::dot(v0[u], xs + c);
        }
        // each row's sum in the lane group's last lane
#pragma unroll
        for (int u = 0; u < RB; u++) acc[u] = group_total<L>(acc[u]);
        if (lane == L - 1) {
#pragma unroll
            for (int u = 0; u < RB; u++) {
                const int q = q0 + u * NG;
                if (q >= t.nr) continue;
                const int r = t.r0 + q;
                if (DIR == 0) {
                    const double b = r < t.p ? xs[r] : bi[q];
                    const double out = b + acc[u];
                    if (r < t.p) sv.wbuf[t.voff + r] = out;
                    else if (ps.fdst[u] >= 0) sv.xin[ps.fdst[u]] = out;
                } else {
                    sv.z[ps.zdst[u]] = acc[u];
                }
            }
        }
    }
}

// zf[fpos[r]] = z[r]: the right-hand side in front order (when the operator
// that formed z did not write it)
__global__ void k_front_rhs(const double *z, const int32_t *fpos, int64_t NS, double *zf) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < NS) zf[fpos[r]] = z[r];
}

// ------------------------------------------------------------ host state
struct SweepDev {
    bool ready = false;
    SwFront *fronts = nullptr;
    std::vector<SwFront> h_fronts;
    int32_t *zidx = nullptr;
    int32_t *fdest = nullptr;
    int32_t *fpos = nullptr;                     // condensed dof -> its pivot row in front order
    double *zf = nullptr;                        // right-hand side in front order
    double *xin = nullptr;                       // forward contribution slots (2 x varena)
    bool zf_given = false;                       // the caller wrote zf for the next solve
    SwTile *tiles = nullptr;
    int ntiles = 0;
    std::vector<SwTask *> d_fwd, d_bwd;
    std::vector<int> n_fwd, n_bwd;
    std::vector<int> l_fwd, l_bwd;               // lanes per row
    std::vector<int> t_fwd, t_bwd;               // threads per workgroup
    std::vector<int> lds_fwd, lds_bwd;           // dynamic LDS bytes
    std::vector<double> bytes_fwd, bytes_bwd;    // stored entries read per level
    int64_t sa_entries = 0;
    void *sa[2] = {nullptr, nullptr};            // [1]: FP64 sweep arena (when formed)
    int built = -1;                              // precision of the current sweep form
    // FP32 factor cache: each slot keeps the sweep form and the equilibration
    // scales of one FP32 factorization of the condensed system and the dt it
    // was computed at
    struct Slot {
        void *sa = nullptr;
        double *rs = nullptr, *cs = nullptr;
        double dt = 0;
        int64_t used = 0;
        bool valid = false;
    };
    std::vector<Slot> slots = std::vector<Slot>(HMEC_FP32_SLOTS);
    int cur = 0;                                 // slot of the active factor
    int target = 0;                              // slot the next FP32 factorization fills
    int64_t clock = 0;
    double form_flops = 0;
    void *mem = nullptr;                         // member table of the last form launch
    int64_t mem_cap = 0;
    std::vector<unsigned char> mem_host;
};

std::mutex g_sw_mu;
std::map<hmec_ctx *, SweepDev *> g_sw;

SweepDev *swdev(hmec_ctx *c) {
    std::lock_guard<std::mutex> lk(g_sw_mu);
    auto it = g_sw.find(c);
    if (it != g_sw.end()) return it->second;
    SweepDev *s = new SweepDev();
    g_sw[c] = s;
    return s;
}

template <typename V>
int upload(V **dst, const std::vector<V> &v) {
    HMEC_HIP(hipMalloc(dst, sizeof(V) * std::max<size_t>(1, v.size())));
    if (!v.empty()) HMEC_HIP(hipMemcpy(*dst, v.data(), sizeof(V) * v.size(), hipMemcpyHostToDevice));
    return HMEC_OK;
}

// row-block tasks of one level and direction; lanes per row from the
// level's largest row length
void level_tasks(const std::vector<SwFront> &fr, const std::vector<int> &fronts,
                 const std::vector<char> &kids, int dir, int vw, std::vector<SwTask> &tasks,
                 int &lanes, int &nt, int &lds, double &entries) {
    tasks.clear();
    entries = 0;
    int kmax = 0, mmax = 0;
    for (int f : fronts) {
        kmax = std::max(kmax, dir == 0 ? fr[f].p : fr[f].n);
        mmax = std::max(mmax, dir == 0 ? fr[f].n : fr[f].p);
    }
    lanes = 4;
    while (lanes < 64 && lanes * vw * 2 < kmax) lanes *= 2;
    // small fronts: one-wave workgroups, so a CU keeps 4x more fronts (and
    // their dependent gather/load chains) in flight
    nt = (mmax <= SW_SMALL && lanes <= 16) ? 64 : SW_T;
    // a workgroup may take several passes of rows so that the gather
    // of the front's input vector into LDS is shared by more rows, as long as
    // the level keeps enough workgroups to fill the device
    const int rows1 = (nt / lanes) * SW_RPG;
    int64_t blocks1 = 0;
    for (int f : fronts) blocks1 += ((dir == 0 ? fr[f].n : fr[f].p) + rows1 - 1) / rows1;
    const int passes = (int)std::max<int64_t>(1, std::min<int64_t>(SW_PASSES, blocks1 / SW_FILL));
    const int rows_per_block = rows1 * passes;
    int nrmax = 0;
    for (int f : fronts) {
        const int n = fr[f].n, p = fr[f].p;
        const int m = dir == 0 ? n : p;
        for (int r0 = 0; r0 < m; r0 += rows_per_block) {
            const int nr = std::min(rows_per_block, m - r0);
            const SwFront &F = fr[f];
            tasks.push_back(SwTask{dir == 0 ? F.soff : F.boff, F.voff, F.goff, F.n, F.p,
                                   dir == 0 ? F.ldf : F.ldb, r0, nr, kids[f] ? 1 : 0});
            nrmax = std::max(nrmax, nr);
        }
        // stored nonzero entries read: Fm = strict lower p x p + u x p,
        // Bm = upper p x p + p x u
        entries += (double)p * (p + (dir == 0 ? -1 : 1)) / 2 + (double)(n - p) * p;
    }
    const int kp = (kmax + 3) & ~3;
    lds = (int)sizeof(double) * (kp + (dir == 0 ? nrmax : 0));
}

}  // namespace

// Per-context sweep plan: fronts, gather map, tasks per level, form tiles.
static int sweep_setup(hmec_ctx *c, SweepDev *s) {
    const FrontPlan &pl = c->plan;
    const DevPlan &dp = c->dp;
    const int nf = pl.nfronts;
    const int dpn = pl.dpn;
    s->h_fronts.resize(nf);
    int64_t soff = 0, goff = 0;
    for (int f = 0; f < nf; f++) {
        const int n = dpn * (pl.np[f] + pl.nu[f]), p = dpn * pl.np[f];
        const int ldf = (p + 3) & ~3, ldb = (n + 3) & ~3;
        const int64_t fo = soff, bo = soff + (int64_t)n * ldf;
        s->h_fronts[f] = SwFront{pl.off[f], fo, bo, pl.voff[f], dp.h_front_dof_off[f], goff, n, p,
                                 ldf, ldb};
        soff = bo + (int64_t)p * ldb;
        goff += n + 1;
    }
    s->sa_entries = soff;
    if (pl.varena > INT32_MAX) {
        hmec_set_last_error("sweep: vector arena exceeds 32-bit gather indices");
        return HMEC_E_ARG;
    }
    // row maps: the row's dof; the forward contribution slots (a parent row
    // takes at most two children's rows, slot 0 for the first child in plan
    // order) and where each child row stores into them; the children's
    // backward input entries each row feeds
    int inv[6];
    for (int v = 0; v < dpn; v++) inv[hmec_var_pos(dpn, v)] = v;
    const int64_t va = pl.varena;
    std::vector<int32_t> zidx(goff), fdest(goff, -1), fpos(c->cp.NS, -1);
    std::vector<uint8_t> fmask(goff, 0);
    for (int f = 0; f < nf; f++) {
        const SwFront &F = s->h_fronts[f];
        for (int r = 0; r < F.n; r++) {
            const int node = pl.nodes[f][r / dpn];
            zidx[F.goff + r] = (int32_t)(node * dpn + inv[r % dpn]);
        }
        for (int r = 0; r < F.p; r++) {
            const int32_t d = zidx[F.goff + r];
            if (d < 0 || d >= c->cp.NS || fpos[d] >= 0) {
                hmec_set_last_error("sweep: a condensed dof is not the pivot of exactly one front");
                return HMEC_E_ARG;
            }
            fpos[d] = (int32_t)(F.voff + r);
        }
        for (int ch : pl.children[f]) {
            const SwFront &Ch = s->h_fronts[ch];
            const int cn = dpn * (pl.np[ch] + pl.nu[ch]), cp = dpn * pl.np[ch];
            const int32_t *map = pl.ea_map.data() + pl.ea_off[ch];
            for (int a = 0; a < cn - cp; a++) {
                const int R = map[a / dpn] * dpn + a % dpn;   // parent row of child row cp + a
                uint8_t &m = fmask[F.goff + R];
                int slot;
                if (!(m & 1)) slot = 0;
                else if (!(m & 2)) slot = 1;
                else {
                    hmec_set_last_error("sweep: a front row receives from more than two children");
                    return HMEC_E_ARG;
                }
                m |= (uint8_t)(1 << slot);
                fdest[Ch.goff + cp + a] = (int32_t)(slot * va + F.voff + R);
            }
        }
    }
    for (int64_t d = 0; d < c->cp.NS; d++)
        if (fpos[d] < 0) {
            hmec_set_last_error("sweep: a condensed dof is the pivot of no front");
            return HMEC_E_ARG;
        }
    if (2 * va > INT32_MAX) {
        hmec_set_last_error("sweep: contribution slots exceed 32-bit indices");
        return HMEC_E_ARG;
    }
    // sweep-form tiles
    std::vector<SwTile> tiles;
    double flops = 0;
    for (int f = 0; f < nf; f++) {
        const int n = s->h_fronts[f].n, p = s->h_fronts[f].p, u = n - p;
        const int tp = (p + SW_TILE - 1) / SW_TILE, tu = (u + SW_TILE - 1) / SW_TILE;
        for (int i = 0; i < tp; i++)
            for (int j = 0; j < tp; j++) tiles.push_back(SwTile{f, 2, i, j});
        for (int i = 0; i < tu; i++)
            for (int j = 0; j < tp; j++) tiles.push_back(SwTile{f, 0, i, j});
        for (int i = 0; i < tp; i++)
            for (int j = 0; j < tu; j++) tiles.push_back(SwTile{f, 1, i, j});
        flops += 2.0 * (double)u * p * p;     // two triangular products, u p^2 each
    }
    s->form_flops = flops;
    s->ntiles = (int)tiles.size();
    HMEC_TRY(upload(&s->fronts, s->h_fronts));
    HMEC_TRY(upload(&s->zidx, zidx));
    HMEC_TRY(upload(&s->fdest, fdest));
    HMEC_TRY(upload(&s->fpos, fpos));
    HMEC_HIP(hipMalloc(&s->zf, sizeof(double) * std::max<int64_t>(1, va)));
    HMEC_HIP(hipMalloc(&s->xin, sizeof(double) * std::max<int64_t>(1, 2 * va)));
    HMEC_HIP(hipMemset(s->xin, 0, sizeof(double) * std::max<int64_t>(1, 2 * va)));
    HMEC_TRY(upload(&s->tiles, tiles));
    // tasks per level (the LDS need is the same for both precisions)
    s->d_fwd.assign(pl.nlevels, nullptr);
    s->d_bwd.assign(pl.nlevels, nullptr);
    s->n_fwd.assign(pl.nlevels, 0);
    s->n_bwd.assign(pl.nlevels, 0);
    s->l_fwd.assign(pl.nlevels, 0);
    s->l_bwd.assign(pl.nlevels, 0);
    s->t_fwd.assign(pl.nlevels, 0);
    s->t_bwd.assign(pl.nlevels, 0);
    s->lds_fwd.assign(pl.nlevels, 0);
    s->lds_bwd.assign(pl.nlevels, 0);
    s->bytes_fwd.assign(pl.nlevels, 0);
    s->bytes_bwd.assign(pl.nlevels, 0);
    std::vector<SwTask> tk;
    std::vector<char> kids(nf, 0);
    for (int f = 0; f < nf; f++) kids[f] = pl.children[f].empty() ? 0 : 1;
    for (int lv = 0; lv < pl.nlevels; lv++) {
        for (int dir = 0; dir < 2; dir++) {
            int lanes, nt, lds;
            double by;
            level_tasks(s->h_fronts, pl.levels[lv], kids, dir, 4, tk, lanes, nt, lds, by);
            if (lds > 64 * 1024) {
                hmec_set_last_error("sweep: front too large for the LDS input vector");
                return HMEC_E_ARG;
            }
            HMEC_TRY(upload(dir == 0 ? &s->d_fwd[lv] : &s->d_bwd[lv], tk));
            (dir == 0 ? s->n_fwd : s->n_bwd)[lv] = (int)tk.size();
            (dir == 0 ? s->l_fwd : s->l_bwd)[lv] = lanes;
            (dir == 0 ? s->t_fwd : s->t_bwd)[lv] = nt;
            (dir == 0 ? s->lds_fwd : s->lds_bwd)[lv] = lds;
            (dir == 0 ? s->bytes_fwd : s->bytes_bwd)[lv] = by;
        }
    }
    s->ready = true;
    return HMEC_OK;
}

template <typename T>
static int form_batch(hmec_ctx *const *cs, int n, int q, SweepDev *s0) {
    std::vector<SwMem<T>> h(n);
    for (int k = 0; k < n; k++) {
        SweepDev *sk = swdev(cs[k]);
        const T *ar = q ? (const T *)(const void *)cs[k]->d.arena : (const T *)(const void *)cs[k]->d.arena32;
        T *dst = q ? (T *)sk->sa[1] : (T *)sk->slots[sk->target].sa;
        h[k] = SwMem<T>{ar, dst};
    }
    hipStream_t st = cs[0]->stream;
    const size_t bytes = sizeof(SwMem<T>) * n;
    std::vector<unsigned char> hb(bytes);
    std::memcpy(hb.data(), h.data(), bytes);
    if (s0->mem_host != hb) {
        HMEC_HIP(hipStreamSynchronize(st));
        if (s0->mem_cap < (int64_t)bytes) {
            hipFree(s0->mem);
            s0->mem = nullptr;
            HMEC_HIP(hipMalloc(&s0->mem, bytes));
            s0->mem_cap = (int64_t)bytes;
        }
        HMEC_HIP(hipMemcpy(s0->mem, hb.data(), bytes, hipMemcpyHostToDevice));
        s0->mem_host = hb;
    }
    if (s0->ntiles > 0) {
        hipLaunchKernelGGL(k_sweep_form<T>, dim3(s0->ntiles, 1, n), dim3(256), 0, st, s0->tiles,
                           s0->fronts, (const SwMem<T> *)s0->mem);
        SW_LAUNCH_CHECK();
    }
    return HMEC_OK;
}

// Form Fm / Bm of every front of n same-size contexts (one launch) from the
// factors in their arenas of the given precision (triangle inverses in
// place, L21, U12).  Runs on the first context's stream.
int sv_sweep_build_batch(hmec_ctx *const *cs, int n, int precision) {
    const int q = precision == HMEC_PREC_FP64 ? 1 : 0;
    const size_t es = q ? sizeof(double) : sizeof(float);
    auto alloc_arena = [&](void **pp) -> int {
        if (*pp) return HMEC_OK;
        if (hipMalloc(pp, es * std::max<int64_t>(1, swdev(cs[0])->sa_entries)) != hipSuccess) {
            hipGetLastError();
            hmec_set_last_error("sweep arena allocation failed");
            return HMEC_E_OOM;
        }
        // row padding and the unused triangles are never written: zero once
        HMEC_HIP(hipMemset(*pp, 0, es * std::max<int64_t>(1, swdev(cs[0])->sa_entries)));
        return HMEC_OK;
    };
    for (int k = 0; k < n; k++) {
        SweepDev *s = swdev(cs[k]);
        if (!s->ready) HMEC_TRY(sweep_setup(cs[k], s));
        if (q) {
            HMEC_TRY(alloc_arena(&s->sa[1]));
        } else {
            SweepDev::Slot &sl = s->slots[s->target];
            HMEC_TRY(alloc_arena(&sl.sa));
            if (!sl.rs) HMEC_HIP(hipMalloc(&sl.rs, sizeof(double) * cs[k]->cp.NS));
            if (!sl.cs) HMEC_HIP(hipMalloc(&sl.cs, sizeof(double) * cs[k]->cp.NS));
            sl.valid = false;
        }
    }
    SweepDev *s0 = swdev(cs[0]);
    kt_begin(cs[0], "sweep_form");
    if (q) HMEC_TRY(form_batch<double>(cs, n, q, s0));
    else HMEC_TRY(form_batch<float>(cs, n, q, s0));
    kt_end(cs[0], "sweep_form", s0->form_flops * n);
    for (int k = 0; k < n; k++) {
        SweepDev *s = swdev(cs[k]);
        s->built = precision;
        if (!q) {
            // keep this factorization's scales with its slot (same stream as
            // the factorization that computed them)
            SweepDev::Slot &sl = s->slots[s->target];
            const size_t nb = sizeof(double) * cs[k]->cp.NS;
            HMEC_HIP(hipMemcpyAsync(sl.rs, cs[k]->dc.rs, nb, hipMemcpyDeviceToDevice, cs[0]->stream));
            HMEC_HIP(hipMemcpyAsync(sl.cs, cs[k]->dc.cs, nb, hipMemcpyDeviceToDevice, cs[0]->stream));
            sl.valid = true;
            sl.dt = cs[k]->dt;
            sl.used = ++s->clock;
            s->cur = s->target;
        }
    }
    return HMEC_OK;
}

// ---- FP32 factor cache (slots) ----
int sv_slots_set(hmec_ctx *c, int n) {
    if (n < 1 || n > 8) return HMEC_E_ARG;
    SweepDev *s = swdev(c);
    for (size_t k = n; k < s->slots.size(); k++) {
        hipFree(s->slots[k].sa);
        hipFree(s->slots[k].rs);
        hipFree(s->slots[k].cs);
    }
    s->slots.resize(n);
    if (s->cur >= n) s->cur = 0;
    if (s->target >= n) s->target = 0;
    if (!s->slots[s->cur].valid && s->built == HMEC_PREC_FP32) s->built = -1;
    return HMEC_OK;
}

int sv_slots_count(hmec_ctx *c) { return (int)swdev(c)->slots.size(); }

// valid slot with the dt nearest to dt (|log| distance), -1 if none
int sv_slot_nearest(hmec_ctx *c, double dt, double *dist) {
    SweepDev *s = swdev(c);
    int best = -1;
    double bd = 0;
    for (int k = 0; k < (int)s->slots.size(); k++) {
        const SweepDev::Slot &sl = s->slots[k];
        if (!sl.valid) continue;
        const double d = std::fabs(std::log(dt / sl.dt));
        if (best < 0 || d < bd) {
            best = k;
            bd = d;
        }
    }
    if (dist) *dist = bd;
    return best;
}

// an invalid slot, else the least recently used one other than `keep`
int sv_slot_victim(hmec_ctx *c, int keep) {
    SweepDev *s = swdev(c);
    int v = -1;
    for (int k = 0; k < (int)s->slots.size(); k++) {
        if (k == keep && s->slots.size() > 1) continue;
        if (!s->slots[k].valid) return k;
        if (v < 0 || s->slots[k].used < s->slots[v].used) v = k;
    }
    return v < 0 ? 0 : v;
}

void sv_slot_target(hmec_ctx *c, int k) { swdev(c)->target = k; }

// slot k holds no factor, or one no solve has used in the last n activations
bool sv_slot_idle(hmec_ctx *c, int k, int64_t n) {
    SweepDev *s = swdev(c);
    if (k < 0 || k >= (int)s->slots.size()) return false;
    return !s->slots[k].valid || s->clock - s->slots[k].used >= n;
}

// make slot k the active FP32 factor (its scales are read in place, see
// sv_active_scales)
int sv_slot_activate(hmec_ctx *c, int k) {
    SweepDev *s = swdev(c);
    if (k < 0 || k >= (int)s->slots.size() || !s->slots[k].valid) return HMEC_E_ARG;
    SweepDev::Slot &sl = s->slots[k];
    s->cur = k;
    s->built = HMEC_PREC_FP32;
    sl.used = ++s->clock;
    c->factor_prec = HMEC_PREC_FP32;
    c->factor_valid = true;
    c->factor_dt = sl.dt;
    return HMEC_OK;
}

bool sv_sweep_ready(hmec_ctx *c, int precision) {
    SweepDev *s = swdev(c);
    if (s->built != precision) return false;
    return precision != HMEC_PREC_FP32 || s->slots[s->cur].valid;
}

// The condensed row / column equilibration of the active factor: the FP32
// slot's own copy while that slot's sweep form solves, else the scales the
// last factorization computed
void sv_active_scales(hmec_ctx *c, const double **rs, const double **cs) {
    SweepDev *s = swdev(c);
    if (c->factor_prec == HMEC_PREC_FP32 && sv_sweep_ready(c, HMEC_PREC_FP32)) {
        *rs = s->slots[s->cur].rs;
        *cs = s->slots[s->cur].cs;
    } else {
        *rs = c->dc.rs;
        *cs = c->dc.cs;
    }
}

// an FP32 factorization is about to overwrite the target slot; an FP64 one
// leaves the FP32 slots alone
void sv_sweep_invalidate(hmec_ctx *c, int precision) {
    SweepDev *s = swdev(c);
    if (precision == HMEC_PREC_FP32) s->slots[s->target].valid = false;
    s->built = -1;
}

template <typename T, int DIR>
static int launch_sweep(hmec_ctx *c, SweepDev *s, int lv, const T *sa, double *z) {
    const int lanes = DIR == 0 ? s->l_fwd[lv] : s->l_bwd[lv];
    const int nth = DIR == 0 ? s->t_fwd[lv] : s->t_bwd[lv];
    const int n = DIR == 0 ? s->n_fwd[lv] : s->n_bwd[lv];
    const int lds = DIR == 0 ? s->lds_fwd[lv] : s->lds_bwd[lv];
    const SwTask *tk = DIR == 0 ? s->d_fwd[lv] : s->d_bwd[lv];
    hipStream_t st = c->stream;
    const SwVec sv{s->zidx, s->fdest, s->zf, z, c->d.vec, s->xin, c->plan.varena};
#define SW_GO(LN, NTH)                                                                          \
    hipLaunchKernelGGL((k_sweep<T, DIR, LN, NTH>), dim3(n), dim3(NTH), lds, st, tk, sa, sv)
    if (nth == 64) {
        switch (lanes) {
            case 4: SW_GO(4, 64); break;
            case 8: SW_GO(8, 64); break;
            default: SW_GO(16, 64); break;
        }
    } else {
        switch (lanes) {
            case 4: SW_GO(4, SW_T); break;
            case 8: SW_GO(8, SW_T); break;
            case 16: SW_GO(16, SW_T); break;
            case 32: SW_GO(32, SW_T); break;
            default: SW_GO(64, SW_T); break;
        }
    }
#undef SW_GO
    SW_LAUNCH_CHECK();
    return HMEC_OK;
}

// Where the operator forming the next sweep's right-hand side may also store
// it in front order (zf[fpos[r]] = z[r]); null when the sweep form of that
// precision is not the active solver.  The next sv_sweep_solve then skips its
// own reordering.
void sv_sweep_rhs_target(hmec_ctx *c, int precision, double **zf, const int32_t **fpos) {
    *zf = nullptr;
    *fpos = nullptr;
    if (!sv_sweep_ready(c, precision)) return;
    SweepDev *s = swdev(c);
    *zf = s->zf;
    *fpos = s->fpos;
}
void sv_sweep_rhs_given(hmec_ctx *c) { swdev(c)->zf_given = true; }

// z (scaled right-hand side) -> solution of the scaled system
int sv_sweep_solve(hmec_ctx *c, double *z) {
    SweepDev *s = swdev(c);
    if (s->built < 0) return HMEC_E_ARG;
    const FrontPlan &pl = c->plan;
    const bool q = s->built == HMEC_PREC_FP64;
    const double es = q ? 8.0 : 4.0;
    // one timing interval per direction (the level launches run back to
    // back): per-launch averages = interval / launches, bytes summed
    double bf = 0, bb = 0;
    for (int lv = 0; lv < pl.nlevels; lv++) {
        bf += s->bytes_fwd[lv] * es;
        bb += s->bytes_bwd[lv] * es;
    }
    const double *sa64 = (const double *)s->sa[1];
    const float *sa32 = q ? nullptr : (const float *)s->slots[s->cur].sa;
    if (!s->zf_given) {
        const int64_t NS = c->cp.NS;
        hipLaunchKernelGGL(k_front_rhs, dim3((unsigned)((NS + 255) / 256)), dim3(256), 0, c->stream,
                           (const double *)z, (const int32_t *)s->fpos, NS, s->zf);
        SW_LAUNCH_CHECK();
    }
    s->zf_given = false;
    kt_begin(c, "sweep_fwd");
    for (int lv = 0; lv < pl.nlevels; lv++) {
        if (q) HMEC_TRY((launch_sweep<double, 0>(c, s, lv, sa64, z)));
        else HMEC_TRY((launch_sweep<float, 0>(c, s, lv, sa32, z)));
    }
    kt_end(c, "sweep_fwd", bf, pl.nlevels);
    kt_begin(c, "sweep_bwd");
    for (int lv = pl.nlevels - 1; lv >= 0; lv--) {
        if (q) HMEC_TRY((launch_sweep<double, 1>(c, s, lv, sa64, z)));
        else HMEC_TRY((launch_sweep<float, 1>(c, s, lv, sa32, z)));
    }
    kt_end(c, "sweep_bwd", bb, pl.nlevels);
    return HMEC_OK;
}

void sv_sweep_free(hmec_ctx *c) {
    SweepDev *s = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_sw_mu);
        auto it = g_sw.find(c);
        if (it == g_sw.end()) return;
        s = it->second;
        g_sw.erase(it);
    }
    hipFree(s->fronts);
    hipFree(s->zidx);
    hipFree(s->fdest);
    hipFree(s->fpos);
    hipFree(s->zf);
    hipFree(s->xin);
    hipFree(s->tiles);
    for (auto *p : s->d_fwd) hipFree(p);
    for (auto *p : s->d_bwd) hipFree(p);
    hipFree(s->sa[0]);
    hipFree(s->sa[1]);
    for (auto &sl : s->slots) {
        hipFree(sl.sa);
        hipFree(sl.rs);
        hipFree(sl.cs);
    }
    hipFree(s->mem);
    delete s;
}

int64_t sv_sweep_bytes(hmec_ctx *c, int precision) {
    SweepDev *s = swdev(c);
    return s->sa_entries * (precision == HMEC_PREC_FP64 ? 8 : 4);
}

