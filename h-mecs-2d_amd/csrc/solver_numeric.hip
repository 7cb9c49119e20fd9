// Numeric multifrontal LU and solve on gfx950.
//
// Replaces PardisoLU::factorize / solve (h_mec_rsf_v9.cpp:1051-1052).
// The matrix is first equilibrated with power-of-two row and column scales
// (exact in floating point), then assembled into dense column-major fronts.
// Fronts are processed level by level (all fronts of a level in one batched
// launch sequence):
//   extend-add of the children's contribution blocks,
//   blocked right-looking LU of the pivot columns / rows with static pivots
//     (a pivot below tau * max|a| is perturbed to +-tau, as PARDISO does with
//     iparm[9] = 13), restricted to the pivot rows / columns,
//   one Schur-complement GEMM  CB -= L21 * U12  with K = #pivots.
// The factor stays in the front arena for the triangular solves, which walk
// the tree leaves->root (forward, contribution vectors extend-added like the
// factor) and root->leaves (backward).  FP64 iterative refinement on the
// original CSR closes the loop; the same kernels run with FP32 fronts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "hmec_internal.h"
#include "mfma_traits.h"
#include "dpp_sum.h"
#include "pow2_scale.h"
#include "xcd_map.h"


namespace {

constexpr int NB = 32;      // panel width
constexpr int TM = 64;      // GEMM tile (rows / cols)
constexpr int TK = 16;      // GEMM K chunk
constexpr int PR = 256;     // rows / columns per panel-solve workgroup

#define LAUNCH_CHECK()                                                      \
    do {                                                                    \
        hipError_t e_ = hipGetLastError();                                  \
        if (e_ != hipSuccess) {                                             \
            hmec_set_last_error(std::string("launch: ") + hipGetErrorString(e_)); \
            return HMEC_E_HIP;                                              \
        }                                                                   \
    } while (0)

// Per-member device pointers of a (possibly batched) factorization: every
// dense kernel of the factorization takes the table and blockIdx.z selects
// the member, so same-size contexts share one launch per kernel (they share
// the symbolic plan and its task arrays).
template <typename T>
struct FacPtrs {
    T *arena;              // front arena
    T *dinv;               // inverted NB x NB diagonal blocks
    T *dscr;               // factored diagonal blocks of the current panel
    T *icopy;              // pivot-block copies of the current level (inversion)
    T *iscr;               // inversion products of the current level and stage
    int *npert;            // perturbed-pivot counter
    const double *val;     // CSR values
    const double *rs, *cs; // equilibration scales
    T *aval;               // the equilibrated original entries in pull-tile order
};

// ------------------------------------------------------------ equilibration

// max that keeps a NaN (a NaN anywhere must show in the result)
__device__ __forceinline__ double nan_max(double m, double v) { return (v > m || v != v) ? v : m; }

// (bad: optional counter of rows / columns with no finite non-zero entry,
// which no pivot perturbation can make solvable)
__global__ void k_row_scale(const int32_t *rowp, const double *val, int64_t N, double *rs,
                            int *bad) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= N) return;
    double m = 0;
    for (int p = rowp[r]; p < rowp[r + 1]; p++) m = nan_max(m, fabs(val[p]));
    if (bad && !(m > 0 && isfinite(m))) atomicAdd(bad, 1);
    rs[r] = pow2_recip(m);
}

__global__ void k_col_scale(const int32_t *cptr, const int32_t *cidx, const int32_t *crow,
                            const double *val, const double *rs, int64_t N, double *cs, int *bad) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    double m = 0;
    for (int q = cptr[c]; q < cptr[c + 1]; q++) m = nan_max(m, fabs(val[cidx[q]] * rs[crow[q]]));
    if (bad && !(m > 0 && isfinite(m))) atomicAdd(bad, 1);
    cs[c] = pow2_recip(m);
}

// ------------------------------------------------------------ lane broadcasts
// wave-uniform broadcast of lane k's value
__device__ __forceinline__ double bcast(double v, int k) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), k);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float bcast(float v, int k) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

// Value of v in lane l of the wavefront (l uniform)
__device__ __forceinline__ double lane_value(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float lane_value(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// ------------------------------------------------------------ front pull
// Forms every front of a level in one launch, replacing a zero fill of the
// arena, the scatter of the original entries and one extend-add launch per
// child slot: entry (r, c) of the front is
//   ((A(r, c) + CB_0(r', c')) + CB_1(r'', c''))
// where A is the equilibrated original entry (0 if none) and CB_q the
// contribution block of the child in slot q at the rows / columns the
// parent's local nodes map to (a term is left out where the child has no
// such node).  The additions run in that fixed order, so each entry is the
// value of zeroing, storing A and adding the two children in slot order.
// One workgroup per tile of at most 64 x 64 entries (a small front is one
// tile): the tile's original entries are placed in an LDS tile, the
// children's values are loaded meanwhile, and every column of the tile is
// stored as 64 consecutive rows.
// PB entries of a thread: the children's values are all loaded before any
// entry is stored (the parent never overlaps its children's blocks)
template <typename T, int PB>
__device__ __forceinline__ void pull_batch(const PullTile &t, unsigned e0, unsigned E, unsigned nr,
                                           const int32_t *__restrict__ im0,
                                           const int32_t *__restrict__ im1,
                                           const T *__restrict__ ch0, const T *__restrict__ ch1,
                                           const T *A, T *__restrict__ out) {
    constexpr int D = 4;                    // unknowns per node (condensed system)
    int r[PB], c[PB], a0[PB], a1[PB];
    T v0[PB], v1[PB];
#pragma unroll
    for (int j = 0; j < PB; j++) {
        const unsigned e = e0 + 256 * j;
        const unsigned cc = e / nr, rr = e - cc * nr;
        r[j] = t.r0 + (int)rr;
        c[j] = t.c0 + (int)cc;
        a0[j] = a1[j] = -1;
        if (e < E && t.nch > 0) {
            const int ar = im0[r[j] / D], ac = im0[c[j] / D];
            if (ar >= 0 && ac >= 0) a0[j] = (t.cp[0] + ac * D + c[j] % D) * t.cn[0] + t.cp[0] + ar * D + r[j] % D;
        }
        if (e < E && t.nch > 1) {
            const int ar = im1[r[j] / D], ac = im1[c[j] / D];
            if (ar >= 0 && ac >= 0) a1[j] = (t.cp[1] + ac * D + c[j] % D) * t.cn[1] + t.cp[1] + ar * D + r[j] % D;
        }
    }
#pragma unroll
    for (int j = 0; j < PB; j++) {
        v0[j] = a0[j] >= 0 ? ch0[a0[j]] : (T)0;
        v1[j] = a1[j] >= 0 ? ch1[a1[j]] : (T)0;
    }
#pragma unroll
    for (int j = 0; j < PB; j++) {
        const unsigned e = e0 + 256 * j;
        if (e >= E) break;
        T acc = A[e];
        if (a0[j] >= 0) acc += v0[j];
        if (a1[j] >= 0) acc += v1[j];
        out[e] = acc;
    }
}

// The equilibrated original entries of the factored matrix, in pull-tile
// order (once per factorization, before the levels)
template <typename T>
__global__ void k_pull_values(const ColEnt *ent, int64_t nnz, const FacPtrs<T> *mem) {
    const FacPtrs<T> M = mem[blockIdx.z];
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    const ColEnt x = ent[e];
    M.aval[e] = (T)(M.val[x.nz] * M.rs[x.grow] * M.cs[x.gcol]);
}

template <typename T>
__global__ void __launch_bounds__(256) k_front_pull(const PullTile *tiles, const int32_t *imap,
                                                    const int32_t *rc, const FacPtrs<T> *mem) {
    __shared__ T A[kPullTile * kPullTile];  // the tile, column-major with leading dimension nr
    const FacPtrs<T> M = mem[blockIdx.z];
    const PullTile t = tiles[blockIdx.x];
    const int tid = threadIdx.x;
    const unsigned E = (unsigned)(t.nr * t.nc), nr = (unsigned)t.nr;
    // the tile's original entries (at most a few per thread) are loaded
    // before the zero fill's barrier
    constexpr int EPT = 4;
    int xr[EPT];
    T xv[EPT];
#pragma unroll
    for (int j = 0; j < EPT; j++) {
        const int e = tid + 256 * j;
        xr[j] = e < t.ne ? rc[t.e0 + e] : -1;
        xv[j] = e < t.ne ? M.aval[t.e0 + e] : (T)0;
    }
    for (unsigned e = tid; e < E; e += 256) A[e] = (T)0;
    __syncthreads();
    for (int base = 0;; base += 256 * EPT) {
#pragma unroll
        for (int j = 0; j < EPT; j++)
            if (xr[j] >= 0) A[(xr[j] >> 12) * nr + (xr[j] & 4095)] = xv[j];
        if (base + 256 * EPT >= t.ne) break;
#pragma unroll
        for (int j = 0; j < EPT; j++) {
            const int e = base + 256 * EPT + tid + 256 * j;
            xr[j] = e < t.ne ? rc[t.e0 + e] : -1;
            xv[j] = e < t.ne ? M.aval[t.e0 + e] : (T)0;
        }
    }
    __syncthreads();
    // consecutive threads take consecutive rows of a column (a tile is one
    // contiguous piece of the front)
    const int32_t *im0 = imap + t.imoff[0], *im1 = imap + t.imoff[1];
    const T *ch0 = M.arena + t.coff[0], *ch1 = M.arena + t.coff[1];
    T *out = M.arena + t.off + (int64_t)t.c0 * t.n;
    constexpr int PB = 8;
    for (unsigned e0 = tid; e0 < E; e0 += 256 * PB)
        pull_batch<T, PB>(t, e0, E, nr, im0, im1, ch0, ch1, A, out);
}

// ------------------------------------------------------------ panel
// One wavefront: in-place LU (no row exchanges, static-pivot perturbation
// below tau on the first kb pivots) of the KB x KB block D in LDS
// (identity-padded beyond kb, which the steps leave as it is).  Lane r keeps
// row r in registers; step k broadcasts the pivot row from lane k, and every
// lane below it forms its multiplier and updates its row.  Rows at or above
// the pivot take a zero multiplier instead of a select, so each updated entry
// is one fused multiply-add (the rows above keep their values: the pivot
// row's entries are finite once the pivot is).
template <typename T, int KB = NB>
__device__ __forceinline__ void factor_diag_wave(T (*D)[KB + 1], int kb, double tau, int *npert,
                                                 bool count) {
    const int lane = threadIdx.x & 63, r = lane & (KB - 1);
    T a[KB];
#pragma unroll
    for (int c = 0; c < KB; c++) a[c] = D[r][c];
#pragma unroll
    for (int k = 0; k < KB; k++) {
        if (r == k && k < kb && !(fabs((double)a[k]) >= tau)) {
            // a non-finite pivot cannot be repaired: counted apart (npert[1])
            if (count && lane == k) atomicAdd(isfinite((double)a[k]) ? npert : npert + 1, 1);
            a[k] = (a[k] < (T)0) ? (T)(-tau) : (T)tau;
        }
        const T d = lane_value(a[k], k);
        const T l = (r > k) ? a[k] / d : (T)0;
#pragma unroll
        for (int c = k + 1; c < KB; c++) a[c] = fma(-l, lane_value(a[c], k), a[c]);
        if (r > k) a[k] = l;
    }
    if (lane < KB) {
#pragma unroll
        for (int c = 0; c < KB; c++) D[r][c] = a[c];
    }
}

// One wavefront: inverse of the upper (U, with diagonal) or unit-lower (L)
// triangle of the factored KB x KB block D (TRANS: its transpose stored).
// Lane c keeps row c of D (broadcast by row) and column c of the inverse in
// registers.
template <typename T, bool UPPER, int KB = NB, bool TRANS = false>
__device__ __forceinline__ void invert_tri_wave(T (*D)[KB + 1], T (*Inv)[KB + 1]) {
    const int lane = threadIdx.x & 63, c = lane & (KB - 1);
    T a[KB], x[KB];
#pragma unroll
    for (int q = 0; q < KB; q++) a[q] = D[c][q];
    if (UPPER) {
#pragma unroll
        for (int r = KB - 1; r >= 0; r--) {
            T s = (r == c) ? (T)1 : (T)0;
#pragma unroll
            for (int q = r + 1; q < KB; q++) s = fma(-lane_value(a[q], r), x[q], s);
            x[r] = (r <= c) ? s / lane_value(a[r], r) : (T)0;
        }
    } else {
#pragma unroll
        for (int r = 0; r < KB; r++) {
            T s = (r == c) ? (T)1 : (T)0;
#pragma unroll
            for (int q = 0; q < r; q++) s = fma(-lane_value(a[q], r), x[q], s);
            x[r] = (r >= c) ? s : (T)0;
        }
    }
    if (lane < KB) {
#pragma unroll
        for (int r = 0; r < KB; r++) {
            if (TRANS) Inv[c][r] = x[r];   // Inv holds the transpose
            else Inv[r][c] = x[r];
        }
    }
}

// A front with at most KB pivots (one panel) and at most kSoloMaxN rows,
// factored completely by one wavefront, for the levels of small fronts (in
// place of the panel, trailing update, Schur-complement and inversion
// launches): the KB x KB pivot block is factored and both its triangles
// inverted as in the panel kernel, and the inverses replace it (the layout
// the solves read); then
//   U12 := L11^-1 A12 (lane per column),   L21 := A21 U11^-1 (lane per row),
//   CB  := A22 - L21 U12 (lane per row; U12 staged in LDS by column chunks
//                         in the space of the factored block).
constexpr int kSoloMaxN = 128;
constexpr int kSoloMaxP = 16;   // wider pivot blocks factor faster in the panel sequence
template <typename T, int KB>
__global__ void __launch_bounds__(64) k_front_solo(const FrontTask *tasks, const FacPtrs<T> *mem,
                                                   double tau) {
    constexpr int YC = 2 * KB * (KB + 1) / KB >= 64 ? 64 : 32;   // U12 columns per chunk
    T *const arena = mem[blockIdx.z].arena;
    int *const npert = mem[blockIdx.z].npert;
    __shared__ T DL[2][KB][KB + 1];     // factored block, then L11^-1; later U12 chunks
    __shared__ T IuT[KB][KB + 1];       // U11^-1, transposed
    T (*const D)[KB + 1] = DL[0];
    T (*const Il)[KB + 1] = DL[1];
    T (*const Yt)[KB] = reinterpret_cast<T (*)[KB]>(&DL[0][0][0]);   // Yt[col][q] = U12[q][col]
    const FrontTask tk = tasks[blockIdx.y];
    const int kb = tk.p;
    const int n = tk.n;
    if (kb <= 0 || kb > KB || n > kSoloMaxN) return;
    T *const F = arena + tk.off;
    const int lane = threadIdx.x;
    for (int e = lane; e < KB * KB; e += 64) {
        const int r = e % KB, c = e / KB;
        D[r][c] = (r < kb && c < kb) ? F[(int64_t)c * n + r] : (r == c ? (T)1 : (T)0);
    }
    __syncthreads();
    factor_diag_wave<T, KB>(D, kb, tau, npert, true);
    __syncthreads();
    invert_tri_wave<T, true, KB, true>(D, IuT);
    invert_tri_wave<T, false, KB>(D, Il);
    __syncthreads();
    for (int e = lane; e < kb * kb; e += 64) {
        const int r = e % kb, c = e / kb;
        F[(int64_t)c * n + r] = r > c ? Il[r][c] : IuT[c][r];
    }
    T a[KB];
    // U12 := L11^-1 A12
    for (int col = kb + lane; col < n; col += 64) {
        T *Fc = F + (int64_t)col * n;
#pragma unroll
        for (int q = 0; q < KB; q++) a[q] = q < kb ? Fc[q] : (T)0;
        for (int r = 0; r < kb; r++) {
            T s = (T)0;
#pragma unroll
            for (int q = 0; q < KB; q++) s = fma(Il[r][q], a[q], s);
            Fc[r] = s;
        }
    }
    // L21 := A21 U11^-1 for this lane's rows (at most two: n <= 128)
    T x0[KB], x1[KB];
    const int r0 = kb + lane, r1 = kb + 64 + lane;
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int r = h ? r1 : r0;
        T *x = h ? x1 : x0;
#pragma unroll
        for (int q = 0; q < KB; q++) a[q] = (r < n && q < kb) ? F[(int64_t)q * n + r] : (T)0;
#pragma unroll
        for (int c = 0; c < KB; c++) {
            T s = (T)0;
#pragma unroll
            for (int q = 0; q < KB; q++) s = fma(a[q], IuT[c][q], s);
            x[c] = s;
            if (r < n && c < kb) F[(int64_t)c * n + r] = s;
        }
    }
    // CB := A22 - L21 U12, U12 read back by chunks of YC columns into LDS
    for (int c0 = kb; c0 < n; c0 += YC) {
        const int nc = min(YC, n - c0);
        __syncthreads();
        for (int e = lane; e < nc * KB; e += 64) {
            const int q = e % KB, cc = e / KB;
            Yt[cc][q] = q < kb ? F[(int64_t)(c0 + cc) * n + q] : (T)0;
        }
        __syncthreads();
        for (int cc = 0; cc < nc; cc++) {
            T *Fc = F + (int64_t)(c0 + cc) * n;
            T s0 = r0 < n ? Fc[r0] : (T)0, s1 = r1 < n ? Fc[r1] : (T)0;
#pragma unroll
            for (int q = 0; q < KB; q++) {
                const T y = Yt[cc][q];
                s0 = fma(-x0[q], y, s0);
                s1 = fma(-x1[q], y, s1);
            }
            if (r0 < n) Fc[r0] = s0;
            if (r1 < n) Fc[r1] = s1;
        }
    }
}

// Panel k0 of every front of a level (no row exchanges):
//   every workgroup loads the kb x kb diagonal block D and factors it with
//   one wavefront, then inverts both triangles (one wavefront each);
//   workgroup 0 stores the factored D in dscr (copied into the front by the
//   following update launch, so no workgroup of this launch reads a
//   half-written D) and the inverses in dinv; row workgroups form
//   X = A U_D^-1 for rows [k0+kb, n), column workgroups Y = L_D^-1 A for
//   columns [k0+kb, n) (products with the inverted triangles: no serial chain
//   per row).
// WG < PR: narrower workgroups for levels whose fronts all have at most WG
// trailing rows at this panel (one workgroup per front, more fronts resident
// per CU).  With one wavefront (WG = 64) both inversions run on it in turn,
// and L_D^-1 overwrites D once D is stored.
template <typename T, int WG>
__global__ void __launch_bounds__(WG) k_panel(const FrontTask *tasks, int k0, const FacPtrs<T> *mem,
                                              double tau) {
    T *const arena = mem[blockIdx.z].arena;
    T *const dscr = mem[blockIdx.z].dscr;
    T *const dinv = mem[blockIdx.z].dinv;
    int *const npert = mem[blockIdx.z].npert;
    __shared__ T D[NB][NB + 1];
    __shared__ T Iu[NB][NB + 1];      // U_D^-1 (upper)
    __shared__ T IlS[WG > 64 ? NB : 1][NB + 1];
    T (*const Il)[NB + 1] = WG > 64 ? IlS : D;   // L_D^-1 (unit lower)
    const FrontTask tk = tasks[blockIdx.y];
    if (tk.p <= k0) return;
    const int kb = min(NB, tk.p - k0);
    const int64_t n = tk.n;
    const int rest = tk.n - k0 - kb;
    const int nch = (rest + PR - 1) / PR;
    // a front whose trailing rows fit one workgroup is done by workgroup 0
    // alone (one diagonal factorization instead of three)
    const bool solo = WG < PR || nch <= 1;
    if ((int)blockIdx.x >= (solo ? 1 : 1 + 2 * nch)) return;
    T *F = arena + tk.off;
    const int tid = threadIdx.x, wave = tid >> 6;
    for (int e = tid; e < NB * NB; e += blockDim.x) {
        const int r = e % NB, c = e / NB;
        D[r][c] = (r < kb && c < kb) ? F[(int64_t)(k0 + c) * n + k0 + r] : (r == c ? (T)1 : (T)0);
    }
    __syncthreads();
    if (wave == 0) factor_diag_wave<T>(D, kb, tau, npert, blockIdx.x == 0);
    __syncthreads();
    // factored D for the update launch, and both inverted triangles of this
    // diagonal block for the solves (block k0/NB of the front: [0] = L^-1
    // unit lower, [1] = U^-1 upper)
    T *ds = dscr + (int64_t)blockIdx.y * NB * NB;
    T *dst = dinv + tk.doff + (int64_t)(k0 / NB) * 2 * NB * NB;
    if (WG == 64) {
        for (int e = tid; e < NB * NB; e += blockDim.x) ds[e] = D[e % NB][e / NB];
        invert_tri_wave<T, true>(D, Iu);
        invert_tri_wave<T, false>(D, Il);
    } else {
        if (wave == 0) invert_tri_wave<T, true>(D, Iu);
        else if (wave == 1) invert_tri_wave<T, false>(D, Il);
    }
    __syncthreads();
    if (blockIdx.x == 0) {
        for (int e = tid; e < NB * NB; e += blockDim.x) {
            const int r = e % NB, c = e / NB;
            if (WG > 64) ds[e] = D[r][c];
            dst[e] = Il[r][c];
            dst[NB * NB + e] = Iu[r][c];
        }
        if (!solo) return;
    }
    const bool cols = solo || (int)blockIdx.x > nch;
    const bool rows = solo || !cols;
    T a[NB];
    if (cols) {
        // Y = L_D^-1 A on columns [k0+kb, n)
        const int col = k0 + kb + (solo ? 0 : (blockIdx.x - 1 - nch) * PR) + tid;
        if (col < tk.n) {
            T *Fc = F + (int64_t)col * n + k0;
#pragma unroll
            for (int q = 0; q < NB; q++) a[q] = (q < kb) ? Fc[q] : (T)0;
            for (int r = 0; r < kb; r++) {
                T s = (T)0;
#pragma unroll
                for (int q = 0; q < NB; q++) s = fma(Il[r][q], a[q], s);
                Fc[r] = s;
            }
        }
    }
    if (rows) {
        // X = A U_D^-1 on rows [k0+kb, n)
        const int r = k0 + kb + (solo ? 0 : (blockIdx.x - 1) * PR) + tid;
        if (r < tk.n) {
#pragma unroll
            for (int q = 0; q < NB; q++) a[q] = (q < kb) ? F[(int64_t)(k0 + q) * n + r] : (T)0;
            for (int c = 0; c < kb; c++) {
                T s = (T)0;
#pragma unroll
                for (int q = 0; q < NB; q++) s = fma(a[q], Iu[q][c], s);
                F[(int64_t)(k0 + c) * n + r] = s;
            }
        }
    }
}

// ------------------------------------------------------------ GEMM tile
// C(m x nn) -= A(m x K) * B(K x nn); column-major with leading dimension ld.
// 64 x 64 output tile per workgroup of 4 wavefronts (2 x 2 of 32 x 32), each
// wavefront a 2 x 2 grid of 16 x 16 MFMA blocks (v_mfma_*_16x16x4: lane l
// supplies A(l%16, l/16) and B(l/16, l%16); element r of the result holds
// D(Mfma<T>::row(l, r), l%16): row 4*(l/16)+r for f32, 4*r+l/16 for f64, as
// measured on gfx950).  K is streamed through LDS in chunks of TK.
template <typename T>
__device__ __forceinline__ void gemm_tile_sub(T *C, const T *A, const T *B, int64_t ld, int m,
                                              int nn, int K) {
    typedef typename Mfma<T>::acc_t acc_t;
    __shared__ T As[TK][TM + 4];
    __shared__ T Bs[TK][TM + 4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = (wave & 1) * 32, wc = (wave >> 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    constexpr int PER = TK * TM / 256;       // operand elements per thread and chunk
    acc_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = acc_t{0, 0, 0, 0};
    // the next K chunk is loaded into registers while the current one is
    // multiplied from LDS (the global loads overlap the MFMAs)
    T ra[PER], rb[PER];
    auto fetch = [&](int kk) {
        const int kc = min(TK, K - kk);
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int e = threadIdx.x + q * 256;
            const int r = e % TM, k = e / TM;
            ra[q] = (k < kc && r < m) ? A[(int64_t)(kk + k) * ld + r] : (T)0;
            const int cc = e / TK, k2 = e % TK;
            rb[q] = (k2 < kc && cc < nn) ? B[(int64_t)cc * ld + kk + k2] : (T)0;
        }
    };
    auto stage = [&]() {
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int e = threadIdx.x + q * 256;
            As[e / TM][e % TM] = ra[q];
            Bs[e % TK][e / TK] = rb[q];
        }
    };
    if (K > 0) fetch(0);
    for (int kk = 0; kk < K; kk += TK) {
        stage();
        __syncthreads();
        if (kk + TK < K) fetch(kk + TK);
#pragma unroll
        for (int k4 = 0; k4 < TK; k4 += 4) {
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
                if (row < m && col < nn) C[(int64_t)col * ld + row] -= acc[a][b][r];
            }
}

// Trailing updates of a pair of panels P0 = [k0, k0 + kb0) and
// P1 = [k0 + kb0, k0 + kb0 + kb1) (kb1 = 0 when the front has no second
// panel there), restricted to rows < p or columns < p (the contribution block
// is updated once per level by k_schur):
//   MODE 0 (after P0's panel): P0's factored diagonal block back into the
//     front, and P0's rank-kb0 update of P1's column and row strips only, so
//     P1 can be factored;
//   MODE 1 (after P1's panel): P1's factored diagonal block back, and the
//     rank-(kb0 + kb1) update of everything after P1 in one pass.
// A front without P1 gets the plain rank-kb0 update from MODE 1.
template <typename T, int MODE>
__global__ void __launch_bounds__(256) k_update(const FrontTask *tasks, int k0,
                                                const FacPtrs<T> *mem) {
    T *const arena = mem[blockIdx.z].arena;
    const T *const dscr = mem[blockIdx.z].dscr;
    const FrontTask tk = tasks[blockIdx.y];
    if (tk.p <= k0) return;
    const int kb0 = min(NB, tk.p - k0);
    const int kb1 = max(0, min(NB, tk.p - k0 - NB));
    T *F = arena + tk.off;
    const int64_t n = tk.n;
    if (blockIdx.x == 0 && (MODE == 0 || kb1 > 0)) {
        // the factored diagonal block of the panel just factored
        const T *ds = dscr + (int64_t)blockIdx.y * NB * NB;
        const int d0 = MODE == 0 ? k0 : k0 + kb0, kb = MODE == 0 ? kb0 : kb1;
        for (int e = threadIdx.x; e < kb * kb; e += blockDim.x) {
            const int r = e % kb, c = e / kb;
            F[(int64_t)(d0 + c) * n + d0 + r] = ds[c * NB + r];
        }
    }
    const int t = blockIdx.x;
    int r0, m, c0, nn, K;
    if (MODE == 0) {
        if (kb1 == 0) return;
        const int s1 = k0 + kb0, e1 = s1 + kb1;
        const int tr = (tk.n - s1 + TM - 1) / TM;   // P1's column strip, rows [s1, n)
        const int tc = (tk.n - e1 + TM - 1) / TM;   // P1's row strip, columns [e1, n)
        if (t < tr) {
            r0 = s1 + t * TM; m = min(TM, tk.n - r0); c0 = s1; nn = kb1;
        } else if (t < tr + tc) {
            r0 = s1; m = kb1; c0 = e1 + (t - tr) * TM; nn = min(TM, tk.n - c0);
        } else {
            return;
        }
        K = kb0;
    } else {
        const int s = k0 + kb0 + kb1;             // first trailing index
        const int tp = (tk.p - s + TM - 1) / TM;  // pivot-part tiles
        const int u = tk.n - tk.p;
        const int tu = (u + TM - 1) / TM;
        int rt, ct;
        if (t < tp * (tp + tu)) {
            rt = t / (tp + tu);
            ct = t % (tp + tu);
        } else {
            const int t2 = t - tp * (tp + tu);
            if (t2 >= tu * tp) return;
            rt = tp + t2 / tp;
            ct = t2 % tp;
        }
        if (rt < tp) { r0 = s + rt * TM; m = min(TM, tk.p - r0); }
        else { r0 = tk.p + (rt - tp) * TM; m = min(TM, tk.n - r0); }
        if (ct < tp) { c0 = s + ct * TM; nn = min(TM, tk.p - c0); }
        else { c0 = tk.p + (ct - tp) * TM; nn = min(TM, tk.n - c0); }
        K = kb0 + kb1;
    }
    gemm_tile_sub<T>(F + (int64_t)c0 * n + r0, F + (int64_t)k0 * n + r0, F + (int64_t)c0 * n + k0,
                     n, m, nn, K);
}

// Schur complement CB -= L21 U12 (K = p), 64 x 64 MFMA tiles
template <typename T>
__global__ void __launch_bounds__(256) k_schur(const FrontTask *tasks, const FacPtrs<T> *mem) {
    T *const arena = mem[blockIdx.z].arena;
    const FrontTask tk = tasks[blockIdx.y];
    const int u = tk.n - tk.p;
    if (u <= 0 || tk.p <= 0) return;
    const int tu = (u + TM - 1) / TM;
    const int t = blockIdx.x;
    if (t >= tu * tu) return;
    const int rt = t % tu, ct = t / tu;
    const int r0 = tk.p + rt * TM, c0 = tk.p + ct * TM;
    const int m = min(TM, tk.n - r0), nn = min(TM, tk.n - c0);
    T *F = arena + tk.off;
    const int64_t n = tk.n;
    gemm_tile_sub<T>(F + (int64_t)c0 * n + r0, F + r0, F + (int64_t)c0 * n, n, m, nn, tk.p);
}

// ------------------------------------------------------------ triangle inverses
// After a level's panels, L11 (unit lower) and U11 (upper) of every front are
// replaced in place by their inverses, so the solves are plain GEMVs.
// Blocked, right to left / bottom to top: with X the already inverted
// trailing block and the original triangles read from a copy,
//   L: X[r0:p, blk] = -X[r0:p, r0:p] L[r0:p, blk] inv(L_blk)
//   U: Y[blk, r0:p] = -inv(U_blk) U[blk, r0:p] Y[r0:p, r0:p]
// (inv(L_blk), inv(U_blk) are the panel's NB x NB inverted diagonal blocks).
template <typename T>
__global__ void k_pivot_copy(const FrontTask *tasks, const FacPtrs<T> *mem) {
    const T *const arena = mem[blockIdx.z].arena;
    T *const copy = mem[blockIdx.z].icopy;
    const FrontTask tk = tasks[blockIdx.y];
    const int p = tk.p;
    const int64_t tot = (int64_t)p * p;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(e % p), c = (int)(e / p);
        copy[tk.coff + e] = arena[tk.off + (int64_t)c * tk.n + r];
    }
}

// ------------------------------------------------------------ triangle inversion
// The inverses of the factored pivot block's triangles, in place (strict
// lower part: L11^-1 without its unit diagonal; upper part with the
// diagonal: U11^-1), by recursive doubling over the NB-wide diagonal blocks.
// With the inverses over two adjacent pivot ranges A and B (B after A) known,
//   (L^-1)_BA = -L_BB^-1 (L_BA L_AA^-1),   (U^-1)_AB = -U_AA^-1 (U_AB U_BB^-1).
// Stage t pairs ranges of NB 2^t pivots, so a front of p pivots needs
// ceil(log2(p / NB)) stages; a stage is two batched MFMA tile launches (the
// parenthesized products into scratch, then the outer products in place).
// L_BA and U_AB come from the pivot-block copy.
struct InvTile {
    int f;           // front (index in the level's task list)
    int kind;        // 0: lower (output |B| x |A|), 1: upper (output |A| x |B|)
    int a0, na, nb;  // A = [a0, a0 + na), B = [a0 + na, a0 + na + nb)
    int ti, tj;      // TM x TM output tile
    int pad;
    int64_t toff;    // the pair's product block in the level scratch
};

// NB x NB diagonal-block inverses (from the panels) into the pivot block
template <typename T>
__global__ void __launch_bounds__(256) k_inv_diag(const FrontTask *tasks, const FacPtrs<T> *mem) {
    T *const arena = mem[blockIdx.z].arena;
    const T *const dinv = mem[blockIdx.z].dinv;
    const FrontTask tk = tasks[blockIdx.y];
    const int c0 = blockIdx.x * NB;
    if (c0 >= tk.p) return;
    const int jw = min(NB, tk.p - c0);
    const T *Dl = dinv + tk.doff + (int64_t)blockIdx.x * 2 * NB * NB, *Du = Dl + NB * NB;
    T *F = arena + tk.off;
    for (int e = threadIdx.x; e < jw * jw; e += blockDim.x) {
        const int r = e % jw, c = e / jw;
        F[(int64_t)(c0 + c) * tk.n + c0 + r] = r > c ? Dl[c * NB + r] : Du[c * NB + r];
    }
}

// PASS 1: S = L_BA L_AA^-1 (lower) / U_AB U_BB^-1 (upper)
// PASS 2: (L^-1)_BA = -L_BB^-1 S (lower) / (U^-1)_AB = -U_AA^-1 S (upper)
// One TM x TM output tile per workgroup of 4 wavefronts (2 x 2 of 32 x 32,
// each a 2 x 2 grid of 16 x 16 MFMA blocks); K in chunks of TK through LDS,
// restricted to where the triangular factor is nonzero.
template <typename T, int PASS>
__global__ void __launch_bounds__(256) k_inv_gemm(const InvTile *tiles, const FrontTask *tasks,
                                                  const FacPtrs<T> *mem) {
    typedef typename Mfma<T>::acc_t acc_t;
    __shared__ T As[TK][TM + 4];
    __shared__ T Bs[TK][TM + 4];
    const InvTile t = tiles[blockIdx.x];
    const FrontTask tk = tasks[t.f];
    T *const F = mem[blockIdx.z].arena + tk.off;
    const T *const Cp = mem[blockIdx.z].icopy + tk.coff;
    T *const S = mem[blockIdx.z].iscr + t.toff;
    const int64_t n = tk.n, p = tk.p;
    const int a0 = t.a0, b0 = t.a0 + t.na, na = t.na, nb = t.nb;
    const bool lower = t.kind == 0;
    const int m = lower ? nb : na, nn = lower ? na : nb;
    const int i0 = t.ti * TM, j0 = t.tj * TM;
    int kmin, kmax;
    if (PASS == 1) {
        kmin = lower ? j0 : 0;                       // L_AA^-1[k, j]: k >= j
        kmax = lower ? na : min(nb, j0 + TM);        // U_BB^-1[k, j]: k <= j
    } else {
        kmin = lower ? 0 : i0;                       // U_AA^-1[i, k]: k >= i
        kmax = lower ? min(nb, i0 + TM) : na;        // L_BB^-1[i, k]: k <= i
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = (wave & 1) * 32, wc = (wave >> 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    acc_t acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) acc[x][y] = acc_t{0, 0, 0, 0};
    // the next K chunk is loaded into registers while the current one is
    // multiplied from LDS
    constexpr int PER = TK * TM / 256;
    T ra[PER], rb[PER];
    auto fetch = [&](int kk) {
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int e = threadIdx.x + q * 256;
            // left operand (i, k), consecutive threads along i
            {
                const int r = e % TM, kl = e / TM;
                const int i = i0 + r, k = kk + kl;
                T v = (T)0;
                if (i < m && k < kmax) {
                    if (PASS == 1) {
                        v = lower ? Cp[(int64_t)(a0 + k) * p + b0 + i]      // L_BA
                                  : Cp[(int64_t)(b0 + k) * p + a0 + i];     // U_AB
                    } else if (lower) {                                   // L_BB^-1 (unit)
                        v = k < i ? F[(int64_t)(b0 + k) * n + b0 + i] : (k == i ? (T)1 : (T)0);
                    } else {                                              // U_AA^-1
                        v = k >= i ? F[(int64_t)(a0 + k) * n + a0 + i] : (T)0;
                    }
                }
                ra[q] = v;
            }
            // right operand (k, j), consecutive threads along k
            {
                const int c = e / TK, kl = e % TK;
                const int j = j0 + c, k = kk + kl;
                T v = (T)0;
                if (j < nn && k < kmax) {
                    if (PASS == 1) {
                        if (lower)                                        // L_AA^-1 (unit)
                            v = k > j ? F[(int64_t)(a0 + j) * n + a0 + k] : (k == j ? (T)1 : (T)0);
                        else                                              // U_BB^-1
                            v = k <= j ? F[(int64_t)(b0 + j) * n + b0 + k] : (T)0;
                    } else {
                        v = S[(int64_t)j * (lower ? nb : na) + k];        // the PASS 1 product
                    }
                }
                rb[q] = v;
            }
        }
    };
    if (kmin < kmax) fetch(kmin);
    for (int kk = kmin; kk < kmax; kk += TK) {
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int e = threadIdx.x + q * 256;
            As[e / TM][e % TM] = ra[q];
            Bs[e % TK][e / TK] = rb[q];
        }
        __syncthreads();
        if (kk + TK < kmax) fetch(kk + TK);
#pragma unroll
        for (int k4 = 0; k4 < TK; k4 += 4) {
            const T x0 = As[k4 + lk][wr + li], x1 = As[k4 + lk][wr + 16 + li];
            const T y0 = Bs[k4 + lk][wc + li], y1 = Bs[k4 + lk][wc + 16 + li];
            acc[0][0] = Mfma<T>::mma(x0, y0, acc[0][0]);
            acc[0][1] = Mfma<T>::mma(x0, y1, acc[0][1]);
            acc[1][0] = Mfma<T>::mma(x1, y0, acc[1][0]);
            acc[1][1] = Mfma<T>::mma(x1, y1, acc[1][1]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int i = i0 + wr + x * 16 + Mfma<T>::row(lane, q), j = j0 + wc + y * 16 + li;
                if (i >= m || j >= nn) continue;
                const T v = acc[x][y][q];
                if (PASS == 1) S[(int64_t)j * m + i] = v;
                else if (lower) F[(int64_t)(a0 + j) * n + b0 + i] = -v;
                else F[(int64_t)(b0 + j) * n + a0 + i] = -v;
            }
}

// ------------------------------------------------------------ solves
// With L11 and U11 replaced by their inverses every sweep is a sequence of
// parallel GEMVs per level (HBM bound):
//  forward, leaves -> root:  k_fwd_gather  w = [z(pivots); 0] + children's cv
//                            y  = L11^-1 w_p        (MODE 2)      -> z, wbuf
//                            cv = w_u - L21 y       (MODE 0)      -> wbuf
//  backward, root -> leaves: t  = y - U12 x_u       (MODE 1)      -> wbuf
//                            x_p = U11^-1 t         (MODE 3)      -> z
// Each GEMV is split over row and column chunks; the partial sums are
// reduced in a fixed order (deterministic).
constexpr int GR = 256;    // rows per GEMV workgroup
constexpr int GC = 256;    // columns per GEMV workgroup

__global__ void __launch_bounds__(256) k_fwd_gather(const FrontTask *tasks, const int *fronts,
                                                    const int32_t *dofs, const int64_t *dof_off,
                                                    const int64_t *voff, const int *child_ptr,
                                                    const int *child_list, const int64_t *ea_off,
                                                    const int32_t *ea_map, const FrontTask *all_tasks,
                                                    const double *z, double *wbuf, int dpn) {
    const int f = fronts[blockIdx.x];
    const FrontTask tk = tasks[blockIdx.x];
    const int n = tk.n, p = tk.p;
    const int32_t *dof = dofs + dof_off[f];
    double *w = wbuf + voff[f];
    for (int l = threadIdx.x; l < n; l += blockDim.x) w[l] = (l < p) ? z[dof[l]] : 0.0;
    __syncthreads();
    for (int q = child_ptr[f]; q < child_ptr[f + 1]; q++) {
        const int ch = child_list[q];
        const FrontTask ct = all_tasks[ch];
        const int u = ct.n - ct.p;
        const double *cvc = wbuf + voff[ch] + ct.p;
        const int32_t *map = ea_map + ea_off[ch];
        for (int a = threadIdx.x; a < u; a += blockDim.x) w[map[a / dpn] * dpn + a % dpn] += cvc[a];
        __syncthreads();
    }
}

// MODE 0: L21 (rows u @p, cols p @0, x = wbuf[0:p])
// MODE 1: U12 (rows p @0, cols u @p, x = z[dofs[p:n]])
// MODE 2: L11^-1 unit lower (rows p, cols p, x = wbuf[0:p])
// MODE 3: U11^-1 upper (rows p, cols p, x = wbuf[0:p])
template <typename T, int MODE>
__global__ void __launch_bounds__(GR) k_gemv_part(const FrontTask *tasks, const int *fronts,
                                                  const int32_t *dofs, const int64_t *dof_off,
                                                  const int64_t *voff, const T *arena,
                                                  const double *z, const double *wbuf,
                                                  double *part) {
    __shared__ double xs[GC];
    const int f = fronts[blockIdx.y];
    const FrontTask tk = tasks[blockIdx.y];
    const int n = tk.n, p = tk.p, u = n - p;
    const int rows = (MODE == 0) ? u : p, cols = (MODE == 1) ? u : p;
    if (rows <= 0 || cols <= 0) return;
    const int nrc = (rows + GR - 1) / GR, ncc = (cols + GC - 1) / GC;
    if ((int)blockIdx.x >= nrc * ncc) return;
    const int rc = blockIdx.x % nrc, cc = blockIdx.x / nrc;
    const int c0 = cc * GC, cn = min(GC, cols - c0);
    const int r0 = rc * GR;
    // triangular blocks: skip tiles entirely outside the triangle
    if (MODE == 2 && c0 > r0 + GR - 1) return;
    if (MODE == 3 && c0 + cn - 1 < r0) return;
    const int tid = threadIdx.x;
    if (tid < cn) {
        if (MODE == 1) xs[tid] = z[dofs[dof_off[f] + p + c0 + tid]];
        else xs[tid] = wbuf[voff[f] + c0 + tid];
    }
    __syncthreads();
    const int r = r0 + tid;
    if (r >= rows) return;
    const int row = (MODE == 0) ? p + r : r, col0 = (MODE == 1) ? p + c0 : c0;
    const T *M = arena + tk.off + (int64_t)col0 * n + row;
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    // column range inside the triangle
    int cb = 0, ce = cn;
    if (MODE == 2) ce = max(0, min(cn, r - c0));     // strictly below the diagonal
    if (MODE == 3) cb = min(cn, max(0, r - c0));     // on and above the diagonal
    int c = cb;
    for (; c + 8 <= ce; c += 8) {
        T l[8];
#pragma unroll
        for (int q = 0; q < 8; q++) l[q] = M[(int64_t)(c + q) * n];
        s0 += (double)l[0] * xs[c] + (double)l[4] * xs[c + 4];
        s1 += (double)l[1] * xs[c + 1] + (double)l[5] * xs[c + 5];
        s2 += (double)l[2] * xs[c + 2] + (double)l[6] * xs[c + 6];
        s3 += (double)l[3] * xs[c + 3] + (double)l[7] * xs[c + 7];
    }
    for (; c < ce; c++) s0 += (double)M[(int64_t)c * n] * xs[c];
    double s = (s0 + s1) + (s2 + s3);
    // unit diagonal of L11^-1 (the stored diagonal belongs to U11^-1)
    if (MODE == 2 && r >= c0 && r < c0 + cn) s += xs[r - c0];
    part[tk.poff + (int64_t)cc * rows + r] = s;
}

template <int MODE>
__global__ void __launch_bounds__(256) k_gemv_reduce(const FrontTask *tasks, const int *fronts,
                                                     const int32_t *dofs, const int64_t *dof_off,
                                                     const int64_t *voff, double *z, double *wbuf,
                                                     const double *part) {
    const int f = fronts[blockIdx.y];
    const FrontTask tk = tasks[blockIdx.y];
    const int n = tk.n, p = tk.p, u = n - p;
    const int rows = (MODE == 0) ? u : p, cols = (MODE == 1) ? u : p;
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= rows) return;
    const int ncc = (cols + GC - 1) / GC;
    const int r0 = (r / GR) * GR;
    double s = 0;
    for (int cc = 0; cc < ncc; cc++) {
        const int c0 = cc * GC;
        if (MODE == 2 && c0 > r0 + GR - 1) break;                     // skipped tiles
        if (MODE == 3 && c0 + min(GC, cols - c0) - 1 < r0) continue;
        s += part[tk.poff + (int64_t)cc * rows + r];
    }
    double *w = wbuf + voff[f];
    const int32_t *dof = dofs + dof_off[f];
    if (MODE == 0) w[p + r] -= s;
    else if (MODE == 1) w[r] = z[dof[r]] - s;
    else if (MODE == 2) { w[r] = s; z[dof[r]] = s; }
    else z[dof[r]] = s;
}

__global__ void k_scale_vec(const double *a, const double *s, double *out, int64_t N) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < N) out[t] = a[t] * s[t];
}
// x += c * z  (first == 1: x = c * z)
__global__ void k_axpy_scaled(double *x, const double *cs, const double *z, int64_t N, int first) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const double d = cs[t] * z[t];
    x[t] = first ? d : x[t] + d;
}
// r = b - A x
// CSR products over a block of SPMV_ROWS rows per workgroup: the threads
// form the products of the block's contiguous entry range (coalesced val /
// coli loads, LDS chunks of SPMV_CHUNK entries), then each thread sums its
// row's products in entry order.  MODE 0: y = rs * (A x); MODE 1: y = b - A x
// and, when y2 is given, y2 = rs * y.
constexpr int SPMV_ROWS = 256;
constexpr int SPMV_CHUNK = 2048;
// gated launches (a solve's first GMRES step decided on the device) run only
// when the step's state in the device scalars is this one (GM_COMBINE)
constexpr int GM_GATE_COMBINE = 2;
constexpr int GM_GATE_DONE = 5, GM_GATE_RESTART = 6;   // the closing check's outcomes
// MODE 2: y = rs * (A x) and, per workgroup, the partial dot products of its
// rows of y with V_0..V_{nv-1} (columns ldv apart) -> part[block * nq + i],
// nq = nv, or nv + 1 with y . y as the last when sq is set.
constexpr int GM_MAXV = 32;
template <int MODE>
__global__ void __launch_bounds__(SPMV_ROWS)
k_csr_stream(const int32_t *__restrict__ rowp, const int32_t *__restrict__ coli,
             const double *__restrict__ val, const double *__restrict__ x,
             const double *__restrict__ rs, const double *__restrict__ b, double *__restrict__ y,
             double *__restrict__ y2, int64_t N, const double *__restrict__ V = nullptr,
             int64_t ldv = 0, int nv = 0, double *__restrict__ part = nullptr,
             const double *__restrict__ gate = nullptr, int sq = 0) {
    __shared__ double prod[SPMV_CHUNK];
    __shared__ int32_t rp[SPMV_ROWS + 1];
    if (gate && gate[SC_GMSTATE_SLOT] != GM_GATE_COMBINE) return;
    // row blocks in contiguous ranges per XCD (the x entries a block reads
    // overlap its neighbours')
    const int64_t nblk = (N + SPMV_ROWS - 1) / SPMV_ROWS;
    const int64_t blk = xcd_block(nblk);
    if (blk >= nblk) return;
    const int64_t r0 = blk * SPMV_ROWS;
    const int nr = (int)min((int64_t)SPMV_ROWS, N - r0);
    const int t = threadIdx.x;
    for (int k = t; k <= nr; k += SPMV_ROWS) rp[k] = rowp[r0 + k];
    __syncthreads();
    const int e0 = rp[0], e1 = rp[nr];
    const int mb = t < nr ? rp[t] : 0, me = t < nr ? rp[t + 1] : 0;
    double s = 0;
    for (int c0 = e0; c0 < e1; c0 += SPMV_CHUNK) {
        const int c1 = min(e1, c0 + SPMV_CHUNK);
        for (int p = c0 + t; p < c1; p += SPMV_ROWS) prod[p - c0] = val[p] * x[coli[p]];
        __syncthreads();
        const int pe = min(me, c1);
        for (int p = max(mb, c0); p < pe; p++) s += prod[p - c0];
        __syncthreads();
    }
    if (MODE == 2) {
        __shared__ double red[SPMV_ROWS / 64][GM_MAXV];
        double yv = 0.0;
        if (t < nr) {
            yv = rs[r0 + t] * s;
            y[r0 + t] = yv;
        }
        const int lane = t & 63, wv = t >> 6, nq = nv + (sq ? 1 : 0);
        for (int i = 0; i < nq; i++) {
            const double d = group_total<64>(
                t < nr ? yv * (i < nv ? V[(int64_t)i * ldv + r0 + t] : yv) : 0.0);
            if (lane == 63) red[wv][i] = d;
        }
        __syncthreads();
        if (t < nq) {
            double d = 0.0;
#pragma unroll
            for (int q = 0; q < SPMV_ROWS / 64; q++) d += red[q][t];
            part[blk * nq + t] = d;
        }
        return;
    }
    if (MODE == 1 && part) {
        // the residual's norms in the same pass: per workgroup |y|inf,
        // |b|inf, |rs y|inf, |rs b|inf and (rs y).(rs y) -> part[5 blk + q]
        // (nan_max: a NaN anywhere reaches the result)
        __shared__ double red[5][4];
        double m[5] = {0, 0, 0, 0, 0};
        if (t < nr) {
            const int64_t r = r0 + t;
            const double v = b[r] - s, w = rs[r] * v;
            y[r] = v;
            y2[r] = w;
            m[0] = fabs(v);
            m[1] = fabs(b[r]);
            m[2] = fabs(w);
            m[3] = fabs(b[r] * rs[r]);
            m[4] = w * w;
        }
        const auto mx = [](double u, double v) { return nan_max(u, v); };
        const auto add = [](double u, double v) { return u + v; };
#pragma unroll
        for (int q = 0; q < 5; q++) {
            const double v = q == 4 ? block_total256(m[q], add, red[q]) : block_total256(m[q], mx, red[q]);
            if (t == 0) part[5 * blk + q] = v;
        }
        return;
    }
    if (t >= nr) return;
    const int64_t r = r0 + t;
    if (MODE == 0) {
        y[r] = rs[r] * s;
    } else {
        const double v = b[r] - s;
        y[r] = v;
        if (y2) y2[r] = rs[r] * v;
    }
}
// y = b - A x (y2 = rs * y when given)
static void launch_residual(hmec_ctx *c, const double *x, double *y, double *y2, const double *rs,
                            const double *gate = nullptr) {
    const int64_t N = c->N;
    hipLaunchKernelGGL(k_csr_stream<1>, dim3((unsigned)xcd_grid((N + SPMV_ROWS - 1) / SPMV_ROWS)),
                       dim3(SPMV_ROWS), 0, c->stream, c->d.rowp, c->d.coli, c->d.val, x, rs,
                       c->d.rhs, y, y2, N, (const double *)nullptr, (int64_t)0, 0, (double *)nullptr,
                       gate);
}
// y = rs * (A x) with the partial dot products of y with V_0..V_{nv-1}
static void launch_spmv_rs_dots(hmec_ctx *c, const double *x, double *y, const double *V,
                                int nv, double *part, int sq = 0) {
    const int64_t N = c->N;
    hipLaunchKernelGGL(k_csr_stream<2>, dim3((unsigned)xcd_grid((N + SPMV_ROWS - 1) / SPMV_ROWS)),
                       dim3(SPMV_ROWS), 0, c->stream, c->d.rowp, c->d.coli, c->d.val, x,
                       c->d.rscale, nullptr, y, nullptr, N, V, N, nv, part,
                       (const double *)nullptr, sq);
}
// y = rs * (A x)
static void launch_spmv_rs(hmec_ctx *c, const double *x, double *y) {
    const int64_t N = c->N;
    hipLaunchKernelGGL(k_csr_stream<0>, dim3((unsigned)xcd_grid((N + SPMV_ROWS - 1) / SPMV_ROWS)),
                       dim3(SPMV_ROWS), 0, c->stream, c->d.rowp, c->d.coli, c->d.val, x,
                       c->d.rscale, nullptr, y, nullptr, N);
}
// out[0] = max|a|, out[1] = max|b| in two stages (block maxima, then one
// block over them); a NaN anywhere propagates to the result
constexpr int NRM_BLOCKS = 256;
// Inf-norms of NV vectors (x[0..NV)) in two stages: block maxima, then one
// block over them; a NaN anywhere propagates to the result
template <int NV>
struct NormArgs {
    const double *x[NV];
};
// (SQ: one more output, out[NV] = x[2] . x[2], summed in a fixed order)
template <int NV, bool SQ = false>
__global__ void __launch_bounds__(256) k_norms_part(NormArgs<NV> a, int64_t N, double *part,
                                                    const double *gate = nullptr) {
    constexpr int NO = NV + (SQ ? 1 : 0);
    __shared__ double r[NO][4];
    if (gate && gate[SC_GMSTATE_SLOT] != GM_GATE_COMBINE) return;
    double m[NO];
#pragma unroll
    for (int q = 0; q < NO; q++) m[q] = 0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < N;
         t += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int q = 0; q < NV; q++) m[q] = nan_max(m[q], fabs(a.x[q][t]));
        if (SQ) m[NV] += a.x[2][t] * a.x[2][t];
    }
    const auto mx = [](double u, double v) { return nan_max(u, v); };
    const auto add = [](double u, double v) { return u + v; };
#pragma unroll
    for (int q = 0; q < NO; q++) {
        const double v = (SQ && q == NV) ? block_total256(m[q], add, r[q]) : block_total256(m[q], mx, r[q]);
        if (threadIdx.x == 0) part[NO * blockIdx.x + q] = v;
    }
}
// (cnt: optional, out[NO] = cnt[0] and out[NO + 1] = cnt[1], so the counts of
// perturbed and of unrepairable pivots come back with the norms)
template <int NV, bool SQ = false>
// (vscal: the closing check of a solve's first GMRES step decided on the
// device: both relative residuals of out against vtol give GM_DONE, else
// GM_RESTART with the exit flag held; the state also at *vstate)
__global__ void __launch_bounds__(256) k_norms_final(const double *part, int nb, double *out,
                                                     const int *cnt, const double *gate = nullptr,
                                                     double vtol = 0, double *vscal = nullptr,
                                                     double *vstate = nullptr) {
    constexpr int NO = NV + (SQ ? 1 : 0);
    __shared__ double r[NO][4];
    if (gate && gate[SC_GMSTATE_SLOT] != GM_GATE_COMBINE) return;
    double m[NO];
#pragma unroll
    for (int q = 0; q < NO; q++) m[q] = 0;
    for (int k = threadIdx.x; k < nb; k += blockDim.x) {
#pragma unroll
        for (int q = 0; q < NV; q++) m[q] = nan_max(m[q], part[NO * k + q]);
        if (SQ) m[NV] += part[NO * k + NV];
    }
    const auto mx = [](double u, double v) { return nan_max(u, v); };
    const auto add = [](double u, double v) { return u + v; };
#pragma unroll
    for (int q = 0; q < NO; q++) {
        const double v = (SQ && q == NV) ? block_total256(m[q], add, r[q]) : block_total256(m[q], mx, r[q]);
        if (threadIdx.x == 0) out[q] = v;
    }
    if (threadIdx.x == 0) {
        if (cnt) {
            out[NO] = (double)cnt[0];
            out[NO + 1] = (double)cnt[1];
        }
        if (vscal) {
            const double rel_u = out[1] > 0 ? out[0] / out[1] : out[0];
            const double rel_s = out[3] > 0 ? out[2] / out[3] : out[2];
            const int st = (rel_u <= vtol && rel_s <= vtol) ? GM_GATE_DONE : GM_GATE_RESTART;
            vscal[SC_GMSTATE_SLOT] = st;
            *vstate = st;
            if (st == GM_GATE_RESTART) vscal[SC_EXIT_SLOT] = SC_EXIT_HOLD;
        }
    }
}


// ------------------------------------------------------------ Darcy condensation
// E = -A_DD^-1 A_DR of the current matrix (3 entries per Darcy unknown)
// (also the Darcy coefficients the preconditioner applications read,
// dcoef[3d..3d+3) = a_dd, a_nn, a_dn, gathered once per matrix)
// (bad: optional counter of Darcy diagonals that are zero or non-finite, the
// pivots of this block elimination, which no perturbation repairs)
// One launch, three independent index ranges: E and the Darcy coefficients
// per Darcy unknown d < ND; the A_RD coefficients of the condensed
// right-hand side (rcoef[t] = val[raslot[t]], t < nt); and, when given,
// out = a * sc over Nv entries (the equilibrated right-hand side of a solve)
__global__ void k_darcy_e(const double *val, int64_t ND, const int32_t *ddiag,
                          const int32_t *dnbdiag, const int32_t *dnbe, const int32_t *des,
                          double *e, double *dcoef, int *bad, const int32_t *raslot, int64_t nt,
                          double *rcoef, const double *va, const double *vsc, double *vout,
                          int64_t Nv) {
    int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= ND) {
        d -= ND;
        if (d < nt) {
            rcoef[d] = val[raslot[d]];
            return;
        }
        d -= nt;
        if (d < Nv) vout[d] = va[d] * vsc[d];
        return;
    }
    const double add = val[ddiag[d]];
    if (bad && !(fabs(add) > 0 && isfinite(add))) atomicAdd(bad, 1);
    dcoef[3 * d] = add;
    dcoef[3 * d + 1] = dnbe[d] < 0 ? 0.0 : val[dnbdiag[d]];
    dcoef[3 * d + 2] = dnbe[d] < 0 ? 0.0 : val[dnbe[d]];
    if (dnbe[d] < 0) {
        for (int k = 0; k < 3; k++) {
            const int32_t p = des[3 * d + k];
            e[3 * d + k] = p >= 0 ? -val[p] / add : 0.0;
        }
    } else {
        // boundary row d -> n: E[d,k] = -(a_dn / a_dd) E[n,k]
        const double r = val[dnbe[d]] / add, ann = val[dnbdiag[d]];
        for (int k = 0; k < 3; k++) {
            const int32_t p = des[3 * d + k];
            e[3 * d + k] = p >= 0 ? -r * (-val[p] / ann) : 0.0;
        }
    }
}
// S = A_RR + A_RD E, one thread per condensed row, terms in a fixed order
__global__ void k_condense_matrix(const double *val, int64_t NS, const int32_t *srowp,
                                  const int32_t *ssrc, const int32_t *tptr, const int32_t *taslot,
                                  const int32_t *tdk, const int32_t *tsslot, const double *e,
                                  double *sval) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= NS) return;
    for (int j = srowp[r]; j < srowp[r + 1]; j++) sval[j] = ssrc[j] >= 0 ? val[ssrc[j]] : 0.0;
    for (int t = tptr[r]; t < tptr[r + 1]; t++) sval[tsslot[t]] += val[taslot[t]] * e[tdk[t]];
}
// gv = A_DD^-1 (v_D / rs)   (rs: optional exact power-of-two row scaling)
// A GMRES cycle's first basis vector formed where the preconditioner reads
// it: v = a src with a = 1 / sqrt(rr[0]) (0 when rr[0] <= 0), each thread
// writing the entry of the unknown it owns into vout (the Darcy unknowns in
// k_darcy_gv, the others in k_cond_rhs: together every entry once)
struct VScale {
    const double *src = nullptr, *rr = nullptr;
    double *vout = nullptr;
};
__device__ __forceinline__ double vscale_a(const VScale &sc) {
    const double q = *sc.rr;
    return q > 0 ? 1.0 / sqrt(q) : 0.0;
}
__global__ void k_darcy_gv(const double *v, const double *rs, int64_t ND, const int32_t *dglob,
                           const int32_t *dnb, const int32_t *dnbe, const double *dcoef,
                           double *gv, VScale sc) {
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= ND) return;
    const int64_t k = dglob[d];
    const double a = sc.src ? vscale_a(sc) : 0.0;
    double vk;
    if (sc.src) {
        vk = a * sc.src[k];
        sc.vout[k] = vk;
    } else {
        vk = v[k];
    }
    const double vb = rs ? vk / rs[k] : vk;
    if (dnbe[d] < 0) {
        gv[d] = vb / dcoef[3 * d];
    } else {
        const int64_t n = dnb[d];
        const double vnr = sc.src ? a * sc.src[n] : v[n];
        const double vn = rs ? vnr / rs[n] : vnr;
        const double gn = vn / dcoef[3 * d + 1];
        gv[d] = (vb - dcoef[3 * d + 2] * gn) / dcoef[3 * d];
    }
}
// rcoef[t] = the A_RD entry of right-hand-side term t (once per matrix)
// vs = rsS * (v_R / rs - A_RD gv)   (condensed right-hand side, equilibrated)
// (zf, when given: the values go to the sweep's front order, zf[fpos[r]],
// instead; the sweep then writes its solution into vs)
__global__ void k_cond_rhs(const double *v, const double *rs, const double *rcoef, int64_t NS,
                           const int32_t *full_of_s, const int32_t *rptr, const int32_t *rd,
                           const double *gv, const double *rsS, double *vs, double *zf,
                           const int32_t *fpos, VScale sc) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= NS) return;
    const int64_t k = full_of_s[r];
    double vk;
    if (sc.src) {
        vk = vscale_a(sc) * sc.src[k];
        sc.vout[k] = vk;
    } else {
        vk = v[k];
    }
    double s = rs ? vk / rs[k] : vk;
    for (int t = rptr[r]; t < rptr[r + 1]; t++) s -= rcoef[t] * gv[rd[t]];
    const double o = rsS[r] * s;
    if (zf) zf[fpos[r]] = o;
    else vs[r] = o;
}
// z_R = csS * vs, z_D = gv + E z_R
__global__ void k_expand(const double *vs, const double *csS, const double *gv, const double *e,
                         int64_t N, const int32_t *s_of_full, const int32_t *d_of_full,
                         const int32_t *dec, double *z) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const int32_t si = s_of_full[k];
    if (si >= 0) {
        z[k] = csS[si] * vs[si];
        return;
    }
    const int64_t d = d_of_full[k];
    double s = gv[d];
    for (int q = 0; q < 3; q++) {
        const int32_t cidx = dec[3 * d + q];
        if (cidx >= 0) s += e[3 * d + q] * (csS[cidx] * vs[cidx]);
    }
    z[k] = s;
}
// y = a u + b v
__global__ void k_lincomb(double *y, double a, const double *u, double b, const double *v, int64_t N) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < N) y[t] = a * u[t] + b * v[t];
}
// x = z (first) or x += z
__global__ void k_add(double *x, const double *z, int64_t N, int first) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < N) x[t] = first ? z[t] : x[t] + z[t];
}

// ------------------------------------------------------------ GMRES helpers
constexpr int RB = 512;      // blocks of the two-stage reductions
// y = A x
__global__ void k_spmv(const int32_t *rowp, const int32_t *coli, const double *val,
                       const double *x, double *y, int64_t N) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    double s = 0;
    for (int p = rowp[t]; p < rowp[t + 1]; p++) s += val[p] * x[coli[p]];
    y[t] = s;
}
// y = Dr (A x)   (row-equilibrated operator)
// partial[b * nv + i] = sum over block b's slice of V_i . w   (nv vectors)
// (self: one more output, partial[.. + nv] = w . w)
__global__ void __launch_bounds__(256) k_dots_part(const double *V, int64_t ldv, int nv,
                                                   const double *w, int64_t N, double *partial,
                                                   int self = 0) {
    __shared__ double red[256];
    const int64_t chunk = (N + gridDim.x - 1) / gridDim.x;
    const int64_t b0 = (int64_t)blockIdx.x * chunk, b1 = min(N, b0 + chunk);
    const int no = nv + self;
    for (int i = 0; i < no; i++) {
        const double *v = i < nv ? V + (int64_t)i * ldv : w;
        double s = 0;
        for (int64_t t = b0 + threadIdx.x; t < b1; t += blockDim.x) s += v[t] * w[t];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int k = 128; k > 0; k >>= 1) {
            if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * no + i] = red[0];
        __syncthreads();
    }
}
// out[i] = sum_b partial[b * nv + i]: one workgroup per vector, fixed
// order (strided per thread, DPP wavefront sums, then the four wavefronts)
__global__ void __launch_bounds__(256) k_dots_final(const double *partial, int nb, int nv,
                                                    double *out) {
    __shared__ double red[4];
    const int i = blockIdx.x;
    double s = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) s += partial[(int64_t)b * nv + i];
    s = block_total256(s, [](double u, double v) { return u + v; }, red);
    if (threadIdx.x == 0) out[i] = s;
}
// coefficients passed by value (no upload before the kernel)
struct GmCoef {
    double v[32];
};
// w -= sum_i h_i V_i, coefficients from device memory (no host round trip)
// First Gram-Schmidt update with the second pass's dot products:
// w' = w - V h (h = h[0..nv), nv <= NVMAX), and per workgroup the partial
// sums of V_i . w' (i < nv) and w' . w' -> part[block * (nv + 1) + i].
// Grid-stride (RB workgroups): an element's basis values are loaded once.
template <int NVMAX>
__global__ void __launch_bounds__(256) k_gs_update_dots(const double *V, int64_t ldv, int nv,
                                                        const double *h, double *w, int64_t N,
                                                        double *part) {
    __shared__ double hs[NVMAX];
    __shared__ double red[4][NVMAX + 1];
    if ((int)threadIdx.x < nv) hs[threadIdx.x] = h[threadIdx.x];
    __syncthreads();
    double acc[NVMAX + 1];
#pragma unroll
    for (int i = 0; i <= NVMAX; i++) acc[i] = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < N;
         t += (int64_t)gridDim.x * blockDim.x) {
        double vv[NVMAX];
        double s = w[t];
#pragma unroll
        for (int i = 0; i < NVMAX; i++) {
            vv[i] = i < nv ? V[(int64_t)i * ldv + t] : 0.0;
            if (i < nv) s -= hs[i] * vv[i];
        }
        w[t] = s;
#pragma unroll
        for (int i = 0; i < NVMAX; i++) acc[i] += vv[i] * s;
        acc[NVMAX] += s * s;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i <= NVMAX; i++) {
        if (i < nv || i == NVMAX) {
            const double d = group_total<64>(acc[i]);
            if (lane == 63) red[wv][i] = d;
        }
    }
    __syncthreads();
    if ((int)threadIdx.x <= nv) {
        const int i = (int)threadIdx.x < nv ? (int)threadIdx.x : NVMAX;   // w'.w' last
        double d = 0.0;
#pragma unroll
        for (int q = 0; q < 4; q++) d += red[q][i];
        part[(int64_t)blockIdx.x * (nv + 1) + threadIdx.x] = d;
    }
}

// Second Gram-Schmidt pass and normalization: w'' = w' - V h (h = h[0..nv),
// ||w'||^2 = h[nv]); ||w''||^2 = ||w'||^2 - ||h||^2 (V orthonormal, h the
// small second-pass coefficients) -> *ww; y = w'' / ||w''|| when y is given.
__global__ void k_gs_update_norm(const double *V, int64_t ldv, int nv, const double *h,
                                 const double *w, double *y, double *ww, int64_t N) {
    __shared__ double hs[33];
    if ((int)threadIdx.x <= nv) hs[threadIdx.x] = h[threadIdx.x];
    __syncthreads();
    double n2 = hs[nv];
    for (int i = 0; i < nv; i++) n2 -= hs[i] * hs[i];
    n2 = n2 > 0 ? n2 : 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) ww[0] = n2;
    if (!y) return;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    double s = w[t];
    for (int i = 0; i < nv; i++) s -= hs[i] * V[(int64_t)i * ldv + t];
    const double n = sqrt(n2);
    y[t] = n > 0 ? s * (1.0 / n) : 0.0;
}
__global__ void k_gs_update(const double *V, int64_t ldv, int nv, const double *h, double *w,
                            int64_t N) {
    __shared__ double hs[32];
    if ((int)threadIdx.x < nv) hs[threadIdx.x] = h[threadIdx.x];
    __syncthreads();
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    double s = w[t];
    for (int i = 0; i < nv; i++) s -= hs[i] * V[(int64_t)i * ldv + t];
    w[t] = s;
}
// y = a * x
__global__ void k_vscale(const double *x, double a, double *y, int64_t N) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < N) y[t] = a * x[t];
}
// y = x / sqrt(*rr) with the squared norm read on the device (zero when
// *rr is not positive)
__global__ void k_vscale_rr(const double *x, const double *rr, double *y, int64_t N) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const double q = *rr;
    const double a = q > 0 ? 1.0 / sqrt(q) : 0.0;
    if (t < N) y[t] = a * x[t];
}
// x += sum_j c_j Z_j
__global__ void k_combine(const double *Z, int64_t ldz, int nz, GmCoef c, double *x,
                          int64_t N) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    double s = x[t];
    for (int j = 0; j < nz; j++) s += c.v[j] * Z[(int64_t)j * ldz + t];
    x[t] = s;
}
}  // namespace

// ------------------------------------------------------------ host side
struct SolverDev {
    int *fronts = nullptr;        // per level list of front ids (flattened)
    std::vector<int64_t> lvl_fronts_off;
    FrontTask *all_tasks = nullptr;
    int *child_ptr = nullptr, *child_list = nullptr;
    int64_t *voff = nullptr, *ea_off = nullptr;
    double *norms = nullptr;
    double *h_norms = nullptr;
    double *dscr = nullptr;        // factored diagonal blocks of a panel step
    double *dinv64 = nullptr;      // inverted diagonal blocks (FP64 factor)
    float *dinv32 = nullptr;       // inverted diagonal blocks (FP32 factor)
    double *part = nullptr;        // GEMV partial sums of one level
    double *icopy = nullptr;       // pivot-block copies of one level (inversion)
    double *iscr = nullptr;        // inversion products of one level and stage
    double *aval = nullptr;        // equilibrated original entries (pull-tile order)
    std::vector<std::vector<InvTile *>> inv_tiles;   // per level, per doubling stage
    std::vector<std::vector<int>> inv_ntiles;
    int64_t dinv_size = 0;
    std::vector<int> gemv_fwd_grid, gemv_bwd_grid, gemv_tri_grid;   // per level grid.x
    std::vector<double> schur_flops, fwdupd_entries, bwdupd_entries;   // per level work
    int max_level_fronts = 0;
    std::vector<int> maxu;         // per level max update size
    std::vector<int> maxt_fwd, maxt_bwd;   // per level row-chunk counts
    double *nrm_part = nullptr;    // block maxima of the residual norms
    double *rsd = nullptr, *bs = nullptr;   // FP64 refinement: Dr (b - A x) and Dr b
    void *mem = nullptr;           // member pointer table of the last factorization
    int64_t mem_cap = 0;           // bytes
    std::vector<unsigned char> mem_host;   // its host copy
};

static SolverDev *sdev(hmec_ctx *c);

#include <map>
#include <mutex>
static std::mutex g_sdev_mu;
static std::map<hmec_ctx *, SolverDev *> g_sdev;
static SolverDev *sdev(hmec_ctx *c) {
    std::lock_guard<std::mutex> lk(g_sdev_mu);
    auto it = g_sdev.find(c);
    if (it != g_sdev.end()) return it->second;
    SolverDev *s = new SolverDev();
    g_sdev[c] = s;
    return s;
}

int sv_free_device(hmec_ctx *c) {
    SolverDev *s = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_sdev_mu);
        auto it = g_sdev.find(c);
        if (it != g_sdev.end()) {
            s = it->second;
            g_sdev.erase(it);
        }
    }
    DevPlan &dp = c->dp;
    for (auto p : dp.d_tasks) hipFree(p);
    for (auto p : dp.d_pull_t) hipFree(p);
    dp.d_pull_t.clear();
    dp.d_tasks.clear();
    {
        CondDev &q = c->dc;
        for (void *ptr : {(void *)q.srowp, (void *)q.scoli, (void *)q.ssrc, (void *)q.sval,
                          (void *)q.s_of_full, (void *)q.d_of_full, (void *)q.full_of_s,
                          (void *)q.dglob, (void *)q.ddiag, (void *)q.dnb, (void *)q.dnbdiag,
                          (void *)q.dnbe, (void *)q.des, (void *)q.dec, (void *)q.rptr,
                          (void *)q.raslot, (void *)q.rd, (void *)q.tptr, (void *)q.taslot,
                          (void *)q.tdk, (void *)q.tsslot, (void *)q.e, (void *)q.gv, (void *)q.vs,
                          (void *)q.rs, (void *)q.cs, (void *)q.dcoef, (void *)q.rcoef})
            hipFree(ptr);
        q = CondDev();
    }
    hipFree(dp.d_ea_map);
    hipFree(dp.d_imap);
    hipFree(dp.d_pull_ent);
    hipFree(dp.d_pull_rc);
    hipFree(dp.d_csc_ptr);
    hipFree(dp.d_csc_idx);
    hipFree(dp.d_csc_row);
    hipFree(dp.d_front_dofs);
    hipFree(dp.d_front_dof_off);
    dp = DevPlan();
    if (s) {
        hipFree(s->rsd);
        hipFree(s->bs);
        hipFree(s->fronts);
        hipFree(s->all_tasks);
        hipFree(s->child_ptr);
        hipFree(s->child_list);
        hipFree(s->voff);
        hipFree(s->ea_off);
        hipFree(s->norms);
        hipFree(s->nrm_part);
        hipFree(s->dscr);
        hipFree(s->dinv64);
        hipFree(s->dinv32);
        hipFree(s->part);
        hipFree(s->icopy);
        hipFree(s->iscr);
        hipFree(s->aval);
        for (auto &v : s->inv_tiles)
            for (InvTile *q : v) hipFree(q);
        s->inv_tiles.clear();
        s->inv_ntiles.clear();
        hipFree(s->mem);
        if (s->h_norms) hipHostFree(s->h_norms);
        delete s;
    }
    hipFree(c->d.arena);
    hipFree(c->d.arena32);
    hipFree(c->d.vec);
    c->d.arena = nullptr;
    c->d.arena32 = nullptr;
    c->d.vec = nullptr;
    return HMEC_OK;
}

template <typename T>
static int upload(T **dst, const std::vector<T> &v) {
    HMEC_HIP(hipMalloc(dst, sizeof(T) * std::max<size_t>(1, v.size())));
    if (!v.empty()) HMEC_HIP(hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return HMEC_OK;
}

int sv_setup_device(hmec_ctx *c) {
    const FrontPlan &pl = c->plan;
    DevPlan &dp = c->dp;
    SolverDev *s = sdev(c);
    // the factored system is the Darcy-condensed one
    const CondPlan &cp = c->cp;
    const std::vector<int32_t> &srp = cp.srowp, &sci = cp.scoli;
    const int64_t N = cp.NS, nnz = (int64_t)sci.size();
    const int dpn = pl.dpn;
    dp.nlevels = pl.nlevels;
    // all tasks (indexed by front id)
    std::vector<FrontTask> all(pl.nfronts);
    int64_t doff = 0;
    for (int f = 0; f < pl.nfronts; f++) {
        const int p = dpn * pl.np[f];
        all[f] = FrontTask{pl.off[f], dpn * (pl.np[f] + pl.nu[f]), p, doff, 0, 0};
        doff += (int64_t)((p + NB - 1) / NB) * 2 * NB * NB;
    }
    s->dinv_size = doff;
    // per level partial-sum offsets of the split GEMVs (scratch reused by levels)
    int64_t part_max = 0;
    s->gemv_fwd_grid.assign(pl.nlevels, 0);
    s->gemv_bwd_grid.assign(pl.nlevels, 0);
    s->gemv_tri_grid.assign(pl.nlevels, 0);
    int64_t copy_max = 0;
    for (int lv = 0; lv < pl.nlevels; lv++) {
        int64_t po = 0, co = 0;
        for (int f : pl.levels[lv]) {
            const int n = all[f].n, p = all[f].p, u = n - p;
            all[f].poff = po;
            all[f].coff = co;
            co += (int64_t)p * p;
            po += (int64_t)std::max({(p + GC - 1) / GC * u, (u + GC - 1) / GC * p, (p + GC - 1) / GC * p});
            if (u > 0 && p > 0) {
                s->gemv_fwd_grid[lv] = std::max(s->gemv_fwd_grid[lv], ((u + GR - 1) / GR) * ((p + GC - 1) / GC));
                s->gemv_bwd_grid[lv] = std::max(s->gemv_bwd_grid[lv], ((p + GR - 1) / GR) * ((u + GC - 1) / GC));
            }
            s->gemv_tri_grid[lv] = std::max(s->gemv_tri_grid[lv], ((p + GR - 1) / GR) * ((p + GC - 1) / GC));
        }
        part_max = std::max(part_max, po);
        copy_max = std::max(copy_max, co);
    }
    HMEC_HIP(hipMalloc(&s->icopy, sizeof(double) * std::max<int64_t>(1, copy_max)));
    // triangle-inversion tiles per level and doubling stage: pairs of
    // adjacent ranges of NB 2^t pivots (the last range of a front may be
    // shorter), products placed one after another in the level scratch
    {
        int64_t scr_max = 0;
        s->inv_tiles.assign(pl.nlevels, {});
        s->inv_ntiles.assign(pl.nlevels, {});
        for (int lv = 0; lv < pl.nlevels; lv++) {
            int maxp = 0;
            for (int f : pl.levels[lv]) maxp = std::max(maxp, all[f].p);
            for (int b = NB; b < maxp; b *= 2) {
                std::vector<InvTile> tl;
                int64_t to = 0;
                int fi = 0;
                for (int f : pl.levels[lv]) {
                    const int pf = all[f].p;
                    for (int a0 = 0; a0 + b < pf; a0 += 2 * b) {
                        const int na = b, nb = std::min(b, pf - a0 - b);
                        for (int kind = 0; kind < 2; kind++) {
                            const int m = kind == 0 ? nb : na, nn = kind == 0 ? na : nb;
                            for (int ti = 0; ti < (m + TM - 1) / TM; ti++)
                                for (int tj = 0; tj < (nn + TM - 1) / TM; tj++)
                                    tl.push_back(InvTile{fi, kind, a0, na, nb, ti, tj, 0, to});
                            to += (int64_t)na * nb;
                        }
                    }
                    fi++;
                }
                scr_max = std::max(scr_max, to);
                InvTile *d = nullptr;
                HMEC_HIP(hipMalloc(&d, sizeof(InvTile) * std::max<size_t>(1, tl.size())));
                if (!tl.empty())
                    HMEC_HIP(hipMemcpy(d, tl.data(), sizeof(InvTile) * tl.size(), hipMemcpyHostToDevice));
                s->inv_tiles[lv].push_back(d);
                s->inv_ntiles[lv].push_back((int)tl.size());
            }
        }
        HMEC_HIP(hipMalloc(&s->iscr, sizeof(double) * std::max<int64_t>(1, scr_max)));
    }
    HMEC_HIP(hipMalloc(&s->part, sizeof(double) * std::max<int64_t>(1, part_max)));
    HMEC_TRY(upload(&s->all_tasks, all));
    std::vector<int> flist;
    dp.h_tasks.assign(pl.nlevels, {});
    dp.maxp.assign(pl.nlevels, 0);
    dp.maxn.assign(pl.nlevels, 0);
    s->lvl_fronts_off.assign(pl.nlevels + 1, 0);
    // the pull of each level: its fronts with their children's blocks and the
    // parent-node -> child-update-node maps, its tiles of at most kPullTile
    // x kPullTile entries, and each tile's original entries
    std::vector<int32_t> imap;
    std::vector<std::vector<PullTile>> hpt(pl.nlevels);
    std::vector<ColEnt> pent;
    std::vector<int32_t> row_of(nnz);
    for (int64_t r = 0; r < N; r++)
        for (int p = srp[r]; p < srp[r + 1]; p++) row_of[p] = (int32_t)r;
    for (int lv = 0; lv < pl.nlevels; lv++) {
        s->lvl_fronts_off[lv] = (int64_t)flist.size();
        for (int f : pl.levels[lv]) {
            flist.push_back(f);
            dp.h_tasks[lv].push_back(all[f]);
            dp.maxp[lv] = std::max(dp.maxp[lv], all[f].p);
            dp.maxn[lv] = std::max(dp.maxn[lv], all[f].n);
            const auto &kids = pl.children[f];
            if (kids.size() > 2) {
                hmec_set_last_error("front with more than two children");
                return HMEC_E_ARG;
            }
            const int nodes = pl.np[f] + pl.nu[f];
            const int n = all[f].n;
            PullTile pt{};
            pt.off = all[f].off;
            pt.n = n;
            pt.nch = (int)kids.size();
            for (int q = 0; q < (int)kids.size(); q++) {
                const int k = kids[q];
                pt.coff[q] = all[k].off;
                pt.cn[q] = all[k].n;
                pt.cp[q] = all[k].p;
                pt.imoff[q] = (int32_t)imap.size();
                imap.resize(imap.size() + nodes, -1);
                for (int a = 0; a < pl.nu[k]; a++) {
                    const int loc = pl.ea_map[pl.ea_off[k] + a];
                    if (loc < 0 || loc >= nodes) {
                        hmec_set_last_error("extend-add map outside the parent front");
                        return HMEC_E_ARG;
                    }
                    imap[pt.imoff[q] + loc] = a;
                }
            }
            if (imap.size() > INT32_MAX) {
                hmec_set_last_error("pull maps exceed 32-bit offsets");
                return HMEC_E_ARG;
            }
            // tiles of whole columns (kPullTile^2 entries at most: a tile is
            // one contiguous piece of the front); the column lists (rows
            // ascending) give each tile's entries in order
            const int tcols = std::max(1, kPullTile * kPullTile / n);
            if (n > kPullTile * kPullTile) {
                hmec_set_last_error("front order above the pull tile");
                return HMEC_E_ARG;
            }
            for (int c0 = 0; c0 < n; c0 += tcols) {
                PullTile t = pt;
                t.r0 = 0;
                t.c0 = c0;
                t.nr = n;
                t.nc = std::min(tcols, n - c0);
                t.e0 = (int64_t)pent.size();
                for (int c = t.c0; c < t.c0 + t.nc; c++) {
                    const int64_t cb = pl.col_base[f] + c;
                    for (int e = pl.col_ptr[cb]; e < pl.col_ptr[cb + 1]; e++) {
                        const int row = pl.col_row[e];
                        const int32_t p = pl.col_nz[e];
                        pent.push_back(ColEnt{row | ((c - t.c0) << 12), p, row_of[p], sci[p]});
                    }
                }
                t.ne = (int)((int64_t)pent.size() - t.e0);
                hpt[lv].push_back(t);
            }
        }
    }
    if ((int64_t)pent.size() != nnz) {
        hmec_set_last_error("pull tiles do not cover the pattern");
        return HMEC_E_ARG;
    }
    if (dpn != 4) {
        hmec_set_last_error("the front pull assumes 4 unknowns per node");
        return HMEC_E_ARG;
    }
    s->lvl_fronts_off[pl.nlevels] = (int64_t)flist.size();
    s->max_level_fronts = 0;
    s->maxu.assign(pl.nlevels, 0);
    s->maxt_fwd.assign(pl.nlevels, 0);
    s->maxt_bwd.assign(pl.nlevels, 0);
    s->schur_flops.assign(pl.nlevels, 0);
    s->fwdupd_entries.assign(pl.nlevels, 0);
    s->bwdupd_entries.assign(pl.nlevels, 0);
    for (int lv = 0; lv < pl.nlevels; lv++) {
        for (const auto &t : dp.h_tasks[lv]) {
            const double u = t.n - t.p, pp = t.p;
            s->schur_flops[lv] += 2.0 * u * u * pp;
            s->fwdupd_entries[lv] += u * pp;
            s->bwdupd_entries[lv] += u * pp;
        }
        s->max_level_fronts = std::max<int>(s->max_level_fronts, (int)pl.levels[lv].size());
        for (const auto &t : dp.h_tasks[lv]) {
            s->maxu[lv] = std::max(s->maxu[lv], t.n - t.p);
            s->maxt_fwd[lv] = std::max(s->maxt_fwd[lv], (t.n - t.p + 255) / 256);
            s->maxt_bwd[lv] = std::max(s->maxt_bwd[lv], (t.p + 255) / 256);
        }
    }
    HMEC_HIP(hipMalloc(&s->dscr, sizeof(double) * NB * NB * std::max(1, s->max_level_fronts)));
    HMEC_TRY(upload(&s->fronts, flist));
    dp.d_tasks.assign(pl.nlevels, nullptr);
    dp.d_pull_t.assign(pl.nlevels, nullptr);
    dp.n_pull_t.assign(pl.nlevels, 0);
    for (int lv = 0; lv < pl.nlevels; lv++) {
        HMEC_TRY(upload(&dp.d_tasks[lv], dp.h_tasks[lv]));
        HMEC_TRY(upload(&dp.d_pull_t[lv], hpt[lv]));
        dp.n_pull_t[lv] = (int)hpt[lv].size();
    }
    HMEC_TRY(upload(&dp.d_imap, imap));
    HMEC_TRY(upload(&dp.d_pull_ent, pent));
    {
        std::vector<int32_t> prc(pent.size());
        for (size_t e = 0; e < pent.size(); e++) prc[e] = pent[e].rc;
        HMEC_TRY(upload(&dp.d_pull_rc, prc));
    }
    HMEC_HIP(hipMalloc(&s->aval, sizeof(double) * std::max<int64_t>(1, nnz)));
    // children lists
    std::vector<int> cptr(pl.nfronts + 1, 0), clist;
    for (int f = 0; f < pl.nfronts; f++) {
        cptr[f] = (int)clist.size();
        for (int k : pl.children[f]) clist.push_back(k);
    }
    cptr[pl.nfronts] = (int)clist.size();
    HMEC_TRY(upload(&s->child_ptr, cptr));
    HMEC_TRY(upload(&s->child_list, clist));
    HMEC_TRY(upload(&s->voff, pl.voff));
    HMEC_TRY(upload(&s->ea_off, pl.ea_off));
    HMEC_TRY(upload(&dp.d_ea_map, pl.ea_map));
    // the transposed pattern
    std::vector<int32_t> cptr2(N + 1, 0), cidx(nnz), crow(nnz);
    for (int64_t p = 0; p < nnz; p++) cptr2[sci[p] + 1]++;
    for (int64_t k = 0; k < N; k++) cptr2[k + 1] += cptr2[k];
    std::vector<int32_t> fill(cptr2.begin(), cptr2.end() - 1);
    for (int64_t r = 0; r < N; r++)
        for (int p = srp[r]; p < srp[r + 1]; p++) {
            const int q = fill[sci[p]]++;
            cidx[q] = p;
            crow[q] = (int32_t)r;
        }
    HMEC_TRY(upload(&dp.d_csc_ptr, cptr2));
    HMEC_TRY(upload(&dp.d_csc_idx, cidx));
    HMEC_TRY(upload(&dp.d_csc_row, crow));
    // per front global dofs
    std::vector<int32_t> dofs;
    dp.h_front_dof_off.assign(pl.nfronts + 1, 0);
    int inv[6];
    for (int v = 0; v < dpn; v++) inv[hmec_var_pos(dpn, v)] = v;
    for (int f = 0; f < pl.nfronts; f++) {
        dp.h_front_dof_off[f] = (int64_t)dofs.size();
        for (int node : pl.nodes[f])
            for (int q = 0; q < dpn; q++) dofs.push_back(node * dpn + inv[q]);
    }
    dp.h_front_dof_off[pl.nfronts] = (int64_t)dofs.size();
    HMEC_TRY(upload(&dp.d_front_dofs, dofs));
    HMEC_TRY(upload(&dp.d_front_dof_off, dp.h_front_dof_off));
    // condensation structures and the condensed system's vectors
    CondDev &q = c->dc;
    q.snnz = nnz;
    HMEC_TRY(upload(&q.srowp, cp.srowp));
    HMEC_TRY(upload(&q.scoli, cp.scoli));
    HMEC_TRY(upload(&q.ssrc, cp.ssrc));
    HMEC_TRY(upload(&q.s_of_full, cp.s_of_full));
    HMEC_TRY(upload(&q.d_of_full, cp.d_of_full));
    HMEC_TRY(upload(&q.full_of_s, cp.full_of_s));
    HMEC_TRY(upload(&q.dglob, cp.dglob));
    HMEC_TRY(upload(&q.ddiag, cp.ddiag));
    HMEC_TRY(upload(&q.dnb, cp.dnb));
    HMEC_TRY(upload(&q.dnbdiag, cp.dnbdiag));
    HMEC_TRY(upload(&q.dnbe, cp.dnbe));
    HMEC_TRY(upload(&q.des, cp.des));
    HMEC_TRY(upload(&q.dec, cp.dec));
    HMEC_TRY(upload(&q.rptr, cp.rptr));
    HMEC_TRY(upload(&q.raslot, cp.raslot));
    HMEC_TRY(upload(&q.rd, cp.rd));
    HMEC_TRY(upload(&q.tptr, cp.tptr));
    HMEC_TRY(upload(&q.taslot, cp.taslot));
    HMEC_TRY(upload(&q.tdk, cp.tdk));
    HMEC_TRY(upload(&q.tsslot, cp.tsslot));
    HMEC_HIP(hipMalloc(&q.sval, sizeof(double) * std::max<int64_t>(1, nnz)));
    HMEC_HIP(hipMalloc(&q.e, sizeof(double) * 3 * std::max<int64_t>(1, cp.ND)));
    HMEC_HIP(hipMalloc(&q.dcoef, sizeof(double) * 3 * std::max<int64_t>(1, cp.ND)));
    HMEC_HIP(hipMalloc(&q.rcoef, sizeof(double) * std::max<int64_t>(1, cp.rptr.empty() ? 1 : (int64_t)cp.rptr.back())));
    HMEC_HIP(hipMalloc(&q.gv, sizeof(double) * std::max<int64_t>(1, cp.ND)));
    HMEC_HIP(hipMalloc(&q.vs, sizeof(double) * std::max<int64_t>(1, N)));
    HMEC_HIP(hipMalloc(&q.rs, sizeof(double) * std::max<int64_t>(1, N)));
    HMEC_HIP(hipMalloc(&q.cs, sizeof(double) * std::max<int64_t>(1, N)));
    // arenas
    HMEC_HIP(hipMalloc(&c->d.vec, sizeof(double) * std::max<int64_t>(1, pl.varena)));
    HMEC_HIP(hipMalloc(&s->norms, sizeof(double) * 8));
    HMEC_HIP(hipMalloc(&s->nrm_part, sizeof(double) * 5 * NRM_BLOCKS));
    HMEC_HIP(hipHostMalloc(&s->h_norms, sizeof(double) * 8, hipHostMallocDefault));
    return HMEC_OK;
}

static int ensure_arena(hmec_ctx *c, int precision) {
    const int64_t n = c->plan.arena;
    SolverDev *s = sdev(c);
    if (precision == HMEC_PREC_FP64 && !s->dinv64)
        HMEC_HIP(hipMalloc(&s->dinv64, sizeof(double) * std::max<int64_t>(1, s->dinv_size)));
    if (precision == HMEC_PREC_FP32 && !s->dinv32)
        HMEC_HIP(hipMalloc(&s->dinv32, sizeof(float) * std::max<int64_t>(1, s->dinv_size)));
    if (precision == HMEC_PREC_FP64 && !c->d.arena) {
        if (hipMalloc(&c->d.arena, sizeof(double) * n) != hipSuccess) {
            hipGetLastError();
            hmec_set_last_error("front arena (FP64) allocation failed");
            return HMEC_E_OOM;
        }
    }
    if (precision == HMEC_PREC_FP32 && !c->d.arena32) {
        if (hipMalloc(&c->d.arena32, sizeof(float) * n) != hipSuccess) {
            hipGetLastError();
            hmec_set_last_error("front arena (FP32) allocation failed");
            return HMEC_E_OOM;
        }
    }
    return HMEC_OK;
}

template <typename T>
static int factor_levels(hmec_ctx *const *cs, int B, const FacPtrs<T> *mem, hipStream_t st,
                         double tau) {
    hmec_ctx *c = cs[0];                   // plan and task arrays shared by the batch
    const FrontPlan &pl = c->plan;
    DevPlan &dp = c->dp;
    SolverDev *s = sdev(c);
    const unsigned Bz = (unsigned)B;
    const int64_t snnz = c->dc.snnz;
    hipLaunchKernelGGL(k_pull_values<T>, dim3((unsigned)((snnz + 255) / 256), 1, Bz), dim3(256), 0, st,
                       dp.d_pull_ent, snnz, mem);
    LAUNCH_CHECK();
    for (int lv = 0; lv < pl.nlevels; lv++) {
        const int nfr = (int)dp.h_tasks[lv].size();
        // the level's fronts: original entries plus the children's blocks
        const int npt = dp.n_pull_t[lv];
        if (npt > 0) {
            hipLaunchKernelGGL(k_front_pull<T>, dim3((unsigned)npt, 1, Bz), dim3(256), 0, st,
                               dp.d_pull_t[lv], dp.d_imap, dp.d_pull_rc, mem);
            LAUNCH_CHECK();
        }
        // a level of fronts with at most one panel of pivots: one wavefront
        // factors each front completely
        const int maxp = dp.maxp[lv];
        if (maxp <= kSoloMaxP && dp.maxn[lv] <= kSoloMaxN) {
            hipLaunchKernelGGL((k_front_solo<T, kSoloMaxP>), dim3(1, nfr, Bz), dim3(64), 0, st,
                               dp.d_tasks[lv], mem, tau);
            LAUNCH_CHECK();
            continue;
        }
        // pivot panels, in pairs: panel, narrow update of the next panel's
        // strips, next panel, one update of the rest for both
        auto panel = [&](int k0) -> int {
            int gx = 1, rest = 0;
            for (const auto &t : dp.h_tasks[lv]) {
                if (t.p <= k0) continue;
                const int kb = std::min(NB, t.p - k0);
                const int nch = (t.n - k0 - kb + PR - 1) / PR;
                gx = std::max(gx, nch <= 1 ? 1 : 1 + 2 * nch);
                rest = std::max(rest, t.n - k0 - kb);
            }
            // the narrow shapes take one workgroup per front and need every
            // front's trailing rows within the workgroup
            if (rest <= 64)
                hipLaunchKernelGGL((k_panel<T, 64>), dim3(1, nfr, Bz), dim3(64), 0, st, dp.d_tasks[lv],
                                   k0, mem, tau);
            else if (rest <= 128)
                hipLaunchKernelGGL((k_panel<T, 128>), dim3(1, nfr, Bz), dim3(128), 0, st,
                                   dp.d_tasks[lv], k0, mem, tau);
            else
                hipLaunchKernelGGL((k_panel<T, PR>), dim3(gx, nfr, Bz), dim3(PR), 0, st, dp.d_tasks[lv],
                                   k0, mem, tau);
            LAUNCH_CHECK();
            return HMEC_OK;
        };
        for (int k0 = 0; k0 < maxp; k0 += 2 * NB) {
            HMEC_TRY(panel(k0));
            int maxn = 1, maxw = 1;
            bool second = false;
            for (const auto &t : dp.h_tasks[lv]) {
                if (t.p <= k0) continue;
                const int kb0 = std::min(NB, t.p - k0);
                const int kb1 = std::max(0, std::min(NB, t.p - k0 - NB));
                if (kb1 > 0) {
                    second = true;
                    const int s1 = k0 + kb0, e1 = s1 + kb1;
                    maxn = std::max(maxn, (t.n - s1 + TM - 1) / TM + (t.n - e1 + TM - 1) / TM);
                }
                const int sw = k0 + kb0 + kb1;
                const int tp = (t.p - sw + TM - 1) / TM;
                const int tu = (t.n - t.p + TM - 1) / TM;
                maxw = std::max(maxw, tp * (tp + tu) + tu * tp);
            }
            hipLaunchKernelGGL((k_update<T, 0>), dim3(maxn, nfr, Bz), dim3(256), 0, st, dp.d_tasks[lv], k0,
                               mem);
            LAUNCH_CHECK();
            if (second) HMEC_TRY(panel(k0 + NB));
            hipLaunchKernelGGL((k_update<T, 1>), dim3(maxw, nfr, Bz), dim3(256), 0, st, dp.d_tasks[lv], k0,
                               mem);
            LAUNCH_CHECK();
        }
        int maxt = 0;
        for (const auto &t : dp.h_tasks[lv]) {
            const int tu = (t.n - t.p + TM - 1) / TM;
            maxt = std::max(maxt, tu * tu);
        }
        if (maxt > 0) {
            kt_begin(c, "schur");
            hipLaunchKernelGGL(k_schur<T>, dim3(maxt, nfr, Bz), dim3(256), 0, st, dp.d_tasks[lv], mem);
            LAUNCH_CHECK();
            kt_end(c, "schur", s->schur_flops[lv] * B);
        }
        // L11, U11 -> inverses (pivot block only: disjoint from L21, U12, CB)
        hipLaunchKernelGGL(k_pivot_copy<T>, dim3(std::max(1, std::min(1024, maxp * maxp / 256 + 1)), nfr, Bz),
                           dim3(256), 0, st, dp.d_tasks[lv], mem);
        LAUNCH_CHECK();
        const int nblk = (maxp + NB - 1) / NB;
        hipLaunchKernelGGL(k_inv_diag<T>, dim3(std::max(1, nblk), nfr, Bz), dim3(256), 0, st,
                           dp.d_tasks[lv], mem);
        LAUNCH_CHECK();
        for (size_t t = 0; t < s->inv_tiles[lv].size(); t++) {
            const int nt = s->inv_ntiles[lv][t];
            if (nt == 0) continue;
            hipLaunchKernelGGL((k_inv_gemm<T, 1>), dim3(nt, 1, Bz), dim3(256), 0, st,
                               s->inv_tiles[lv][t], dp.d_tasks[lv], mem);
            LAUNCH_CHECK();
            hipLaunchKernelGGL((k_inv_gemm<T, 2>), dim3(nt, 1, Bz), dim3(256), 0, st,
                               s->inv_tiles[lv][t], dp.d_tasks[lv], mem);
            LAUNCH_CHECK();
        }
    }
    return HMEC_OK;
}

static int darcy_e(hmec_ctx *c, hipStream_t st, int *bad = nullptr, const double *va = nullptr,
                   const double *vsc = nullptr, double *vout = nullptr);

// Numeric factorization of n same-size contexts' current matrices with one
// launch per kernel (n = 1: a single context).  The batch runs on the first
// context's stream; the other members' streams are joined before and after
// with events, so each member's own stream order is kept.
int sv_factorize_batch(hmec_ctx *const *cs, int n, int precision) {
    if (!cs || n <= 0) return HMEC_E_ARG;
    hmec_ctx *c0 = cs[0];
    for (int k = 0; k < n; k++) {
        if (!cs[k]->analyzed || cs[k]->Nx != c0->Nx || cs[k]->Ny != c0->Ny ||
            cs[k]->plan.arena != c0->plan.arena || cs[k]->device != c0->device)
            return HMEC_E_ARG;
        HMEC_TRY(ensure_arena(cs[k], precision));
    }
    hipStream_t st = c0->stream;
    std::vector<hipEvent_t> evs;
    auto join = [&](hipStream_t from, hipStream_t to) -> int {
        hipEvent_t e;
        HMEC_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        HMEC_HIP(hipEventRecord(e, from));
        HMEC_HIP(hipStreamWaitEvent(to, e, 0));
        evs.push_back(e);
        return HMEC_OK;
    };
    for (int k = 1; k < n; k++) HMEC_TRY(join(cs[k]->stream, st));
    // per member (small kernels): Darcy condensation S = A_RR + A_RD E of the
    // current matrix, power-of-two equilibration of S, counters cleared
    for (int k = 0; k < n; k++) {
        hmec_ctx *c = cs[k];
        const CondPlan &cp = c->cp;
        CondDev &q = c->dc;
        const int64_t NS = cp.NS;
        HMEC_HIP(hipMemsetAsync(c->d.counters, 0, sizeof(int) * 4, st));
        HMEC_TRY(darcy_e(c, st, c->d.counters + 1));
        hipLaunchKernelGGL(k_condense_matrix, dim3((unsigned)((NS + 255) / 256)), dim3(256), 0, st,
                           c->d.val, NS, q.srowp, q.ssrc, q.tptr, q.taslot, q.tdk, q.tsslot, q.e,
                           q.sval);
        LAUNCH_CHECK();
        hipLaunchKernelGGL(k_row_scale, dim3((unsigned)((NS + 255) / 256)), dim3(256), 0, st,
                           c0->dc.srowp, q.sval, NS, q.rs, c->d.counters + 1);
        LAUNCH_CHECK();
        hipLaunchKernelGGL(k_col_scale, dim3((unsigned)((NS + 255) / 256)), dim3(256), 0, st,
                           c0->dp.d_csc_ptr, c0->dp.d_csc_idx, c0->dp.d_csc_row, q.sval, q.rs, NS,
                           q.cs, c->d.counters + 1);
        LAUNCH_CHECK();
        sv_sweep_invalidate(c, precision);
    }
    // member pointer table (uploaded with the stream's order)
    SolverDev *s0 = sdev(c0);
    const bool fp64 = precision == HMEC_PREC_FP64;
    const size_t psz = fp64 ? sizeof(FacPtrs<double>) : sizeof(FacPtrs<float>);
    std::vector<unsigned char> host(psz * n);
    for (int k = 0; k < n; k++) {
        hmec_ctx *c = cs[k];
        SolverDev *s = sdev(c);
        if (fp64) {
            FacPtrs<double> f{c->d.arena, s->dinv64, s->dscr, s->icopy, s->iscr, c->d.counters,
                              c->dc.sval, c->dc.rs, c->dc.cs, s->aval};
            std::memcpy(host.data() + psz * k, &f, psz);
        } else {
            FacPtrs<float> f{c->d.arena32, s->dinv32, reinterpret_cast<float *>(s->dscr),
                             reinterpret_cast<float *>(s->icopy), reinterpret_cast<float *>(s->iscr),
                             c->d.counters, c->dc.sval, c->dc.rs, c->dc.cs,
                             reinterpret_cast<float *>(s->aval)};
            std::memcpy(host.data() + psz * k, &f, psz);
        }
    }
    // the table only changes with the member set or the precision: upload it
    // then, after the stream has drained any launch that still reads it
    if (s0->mem_host != host) {
        HMEC_HIP(hipStreamSynchronize(st));
        if (s0->mem_cap < (int64_t)host.size()) {
            hipFree(s0->mem);
            s0->mem = nullptr;
            HMEC_HIP(hipMalloc(&s0->mem, host.size()));
            s0->mem_cap = (int64_t)host.size();
        }
        HMEC_HIP(hipMemcpy(s0->mem, host.data(), host.size(), hipMemcpyHostToDevice));
        s0->mem_host = host;
    }
    if (fp64)
        HMEC_TRY(factor_levels<double>(cs, n, (const FacPtrs<double> *)s0->mem, st, 1e-13));
    else
        HMEC_TRY(factor_levels<float>(cs, n, (const FacPtrs<float> *)s0->mem, st, 1e-6));
    // FP32 factors are reused over many solves: form the one-GEMV-per-front
    // sweep form once
    if (!fp64) HMEC_TRY(sv_sweep_build_batch(cs, n, precision));
    for (int k = 0; k < n; k++) {
        cs[k]->factor_prec = precision;
        cs[k]->factor_valid = true;
        cs[k]->stale_solves = 0;
    }
    for (int k = 1; k < n; k++) HMEC_TRY(join(st, cs[k]->stream));
    for (hipEvent_t e : evs) hipEventDestroy(e);
    return HMEC_OK;
}

int sv_factorize(hmec_ctx *c, int precision) { return sv_factorize_batch(&c, 1, precision); }

template <typename T>
static int tri_solve(hmec_ctx *c, const T *arena, double *z) {
    const FrontPlan &pl = c->plan;
    DevPlan &dp = c->dp;
    SolverDev *s = sdev(c);
    hipStream_t st = c->stream;
    double *wbuf = c->d.vec;
    // one timing interval per direction (levels run back to back)
    kt_begin(c, "fwd");
    for (int lv = 0; lv < pl.nlevels; lv++) {
        const int nfr = (int)dp.h_tasks[lv].size();
        const int *fr = s->fronts + s->lvl_fronts_off[lv];
        const int bs = dp.maxn[lv] <= 128 ? 64 : 256;
        hipLaunchKernelGGL(k_fwd_gather, dim3(nfr), dim3(bs), 0, st, dp.d_tasks[lv], fr,
                           dp.d_front_dofs, dp.d_front_dof_off, s->voff, s->child_ptr, s->child_list,
                           s->ea_off, dp.d_ea_map, s->all_tasks, z, wbuf, pl.dpn);
        LAUNCH_CHECK();
        hipLaunchKernelGGL((k_gemv_part<T, 2>), dim3(s->gemv_tri_grid[lv], nfr), dim3(GR), 0, st,
                           dp.d_tasks[lv], fr, dp.d_front_dofs, dp.d_front_dof_off, s->voff, arena, z,
                           wbuf, s->part);
        LAUNCH_CHECK();
        hipLaunchKernelGGL(k_gemv_reduce<2>, dim3((dp.maxp[lv] + 255) / 256, nfr), dim3(256), 0, st,
                           dp.d_tasks[lv], fr, dp.d_front_dofs, dp.d_front_dof_off, s->voff, z, wbuf,
                           s->part);
        LAUNCH_CHECK();
        if (s->gemv_fwd_grid[lv] > 0) {
            hipLaunchKernelGGL((k_gemv_part<T, 0>), dim3(s->gemv_fwd_grid[lv], nfr), dim3(GR), 0, st,
                               dp.d_tasks[lv], fr, dp.d_front_dofs, dp.d_front_dof_off, s->voff,
                               arena, z, wbuf, s->part);
            LAUNCH_CHECK();
            hipLaunchKernelGGL(k_gemv_reduce<0>, dim3((s->maxu[lv] + 255) / 256, nfr), dim3(256), 0,
                               st, dp.d_tasks[lv], fr, dp.d_front_dofs, dp.d_front_dof_off, s->voff, z,
                               wbuf, s->part);
            LAUNCH_CHECK();
        }
    }
    kt_end(c, "fwd", 0, pl.nlevels);
    kt_begin(c, "bwd");
    for (int lv = pl.nlevels - 1; lv >= 0; lv--) {
        const int nfr = (int)dp.h_tasks[lv].size();
        const int *fr = s->fronts + s->lvl_fronts_off[lv];
        if (s->gemv_bwd_grid[lv] > 0) {
            hipLaunchKernelGGL((k_gemv_part<T, 1>), dim3(s->gemv_bwd_grid[lv], nfr), dim3(GR), 0, st,
                               dp.d_tasks[lv], fr, dp.d_front_dofs, dp.d_front_dof_off, s->voff,
                               arena, z, wbuf, s->part);
            LAUNCH_CHECK();
        }
        // t = y - U12 x_u (a plain copy of y where the front has no update set)
        hipLaunchKernelGGL(k_gemv_reduce<1>, dim3((dp.maxp[lv] + 255) / 256, nfr), dim3(256), 0, st,
                           dp.d_tasks[lv], fr, dp.d_front_dofs, dp.d_front_dof_off, s->voff, z, wbuf,
                           s->part);
        LAUNCH_CHECK();
        hipLaunchKernelGGL((k_gemv_part<T, 3>), dim3(s->gemv_tri_grid[lv], nfr), dim3(GR), 0, st,
                           dp.d_tasks[lv], fr, dp.d_front_dofs, dp.d_front_dof_off, s->voff, arena, z,
                           wbuf, s->part);
        LAUNCH_CHECK();
        hipLaunchKernelGGL(k_gemv_reduce<3>, dim3((dp.maxp[lv] + 255) / 256, nfr), dim3(256), 0, st,
                           dp.d_tasks[lv], fr, dp.d_front_dofs, dp.d_front_dof_off, s->voff, z, wbuf,
                           s->part);
        LAUNCH_CHECK();
    }
    kt_end(c, "bwd", 0, pl.nlevels);
    return HMEC_OK;
}

// x = A^-1 b by the factor + iterative refinement on the FP64 CSR.
// tol: target ||b - A x||inf / ||b||inf; max_ir: refinement sweeps after the
// first solve.  Returns HMEC_E_NOCONV when the target is not met.
// out[0..1] = max|a|, max|b|, out[2] = perturbed pivots of the current factor
// (device), two launches on the context stream
static int launch_norms(hmec_ctx *c, const double *a, const double *b, double *out) {
    SolverDev *s = sdev(c);
    NormArgs<2> na{{a, b}};
    hipLaunchKernelGGL(k_norms_part<2>, dim3(NRM_BLOCKS), dim3(256), 0, c->stream, na, c->N, s->nrm_part);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_norms_final<2>, dim3(1), dim3(256), 0, c->stream, s->nrm_part, NRM_BLOCKS, out,
                       (const int *)c->d.counters);
    LAUNCH_CHECK();
    return HMEC_OK;
}
// out[0..4) = inf-norms of a, b, c2, d, out[4] = c2 . c2 in one pass, and
// out[5] = perturbed pivots of the current factor
static int launch_norms4(hmec_ctx *c, const double *a, const double *b, const double *c2,
                         const double *d, double *out, const double *gate = nullptr) {
    SolverDev *s = sdev(c);
    NormArgs<4> na{{a, b, c2, d}};
    hipLaunchKernelGGL((k_norms_part<4, true>), dim3(NRM_BLOCKS), dim3(256), 0, c->stream, na, c->N,
                       s->nrm_part, gate);
    LAUNCH_CHECK();
    hipLaunchKernelGGL((k_norms_final<4, true>), dim3(1), dim3(256), 0, c->stream, s->nrm_part,
                       NRM_BLOCKS, out, (const int *)c->d.counters, gate);
    LAUNCH_CHECK();
    return HMEC_OK;
}

// vs <- S^-1 vs (equilibrated condensed vector) with the current factor: the
// sweep form when it has been formed for this factor, else the level-by-level
// inverted-triangle solves
static int solve_condensed(hmec_ctx *c, double *vs) {
    if (sv_sweep_ready(c, c->factor_prec)) return sv_sweep_solve(c, vs);
    if (c->factor_prec == HMEC_PREC_FP64) return tri_solve<double>(c, c->d.arena, vs);
    return tri_solve<float>(c, c->d.arena32, vs);
}

// E = -A_DD^-1 A_DR of the current matrix
// E, the Darcy coefficients and the A_RD right-hand-side coefficients of the
// current matrix; va / vsc given: also vout = va * vsc over N entries
static int darcy_e(hmec_ctx *c, hipStream_t st, int *bad, const double *va, const double *vsc,
                   double *vout) {
    const CondPlan &cp = c->cp;
    CondDev &q = c->dc;
    const int64_t nt = cp.rptr.empty() ? 0 : cp.rptr.back();
    const int64_t Nv = va ? c->N : 0;
    const int64_t tot = cp.ND + nt + Nv;
    if (tot == 0) return HMEC_OK;
    hipLaunchKernelGGL(k_darcy_e, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, c->d.val,
                       cp.ND, q.ddiag, q.dnbdiag, q.dnbe, q.des, q.e, q.dcoef, bad, q.raslot, nt, q.rcoef,
                       va, vsc, vout, Nv);
    LAUNCH_CHECK();
    return HMEC_OK;
}

// z = A^-1 (v / rs) for the full coupled system by block elimination of the
// Darcy unknowns around the condensed factor (rs: optional row scaling of
// the operator, exact powers of two); E must be the current matrix's
// (sc: v is formed from sc.src on the way in and written to sc.vout)
static int precond_apply(hmec_ctx *c, const double *v, const double *rs, double *z,
                         VScale sc = VScale()) {
    const CondPlan &cp = c->cp;
    CondDev &q = c->dc;
    hipStream_t st = c->stream;
    hipLaunchKernelGGL(k_darcy_gv, dim3((unsigned)((cp.ND + 255) / 256)), dim3(256), 0, st, v, rs,
                       cp.ND, q.dglob, q.dnb, q.dnbe, q.dcoef, q.gv, sc);
    LAUNCH_CHECK();
    double *zf = nullptr;
    const int32_t *fpos = nullptr;
    sv_sweep_rhs_target(c, c->factor_prec, &zf, &fpos);
    const double *srs, *scs;
    sv_active_scales(c, &srs, &scs);
    hipLaunchKernelGGL(k_cond_rhs, dim3((unsigned)((cp.NS + 255) / 256)), dim3(256), 0, st, v, rs,
                       q.rcoef, cp.NS, q.full_of_s, q.rptr, q.rd, q.gv, srs, q.vs, zf, fpos, sc);
    LAUNCH_CHECK();
    if (zf) sv_sweep_rhs_given(c);
    HMEC_TRY(solve_condensed(c, q.vs));
    hipLaunchKernelGGL(k_expand, dim3((unsigned)((c->N + 255) / 256)), dim3(256), 0, st, q.vs, scs,
                       q.gv, q.e, c->N, q.s_of_full, q.d_of_full, q.dec, z);
    LAUNCH_CHECK();
    return HMEC_OK;
}

// The refinement's work vectors rsd = Dr (b - A x) and bs = Dr b (N doubles
// each) are passed in: the context's own (allocated on first use) or the
// allocation-free entry point's scratch array.
static int sv_solve_impl(hmec_ctx *c, int max_ir, double tol, hmec_solve_stats *out, double *rsd,
                         double *bs) {
    if (!c->factor_valid) return HMEC_E_ARG;
    SolverDev *s = sdev(c);
    const int64_t N = c->N;
    hipStream_t st = c->stream;
    const unsigned g = (unsigned)((N + 255) / 256);
    double *x = c->d.sol, *r = c->d.work, *z = c->d.work2;
    // the operator's power-of-two row scales Dr (the assembly forms them with
    // the values) and Dr b: the refinement also watches the row-equilibrated
    // residual, in which the small-magnitude rows (Darcy, continuity) count
    // like the momentum rows
    if (!c->rscale_ready) {
        hipLaunchKernelGGL(k_row_scale, dim3(g), dim3(256), 0, st, c->d.rowp, c->d.val, N, c->d.rscale,
                           nullptr);
        LAUNCH_CHECK();
        c->rscale_ready = true;
    }
    hipLaunchKernelGGL(k_scale_vec, dim3(g), dim3(256), 0, st, c->d.rhs, c->d.rscale, bs, N);
    LAUNCH_CHECK();
    // r = b
    HMEC_HIP(hipMemcpyAsync(r, c->d.rhs, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
    HMEC_TRY(darcy_e(c, st));
    double res = 0, res_s = 0, prev_s = INFINITY;
    int it = 0, npert = 0, nbad = 0;
    // sweeps until both the unscaled and the equilibrated residual meet tol,
    // the equilibrated one stops halving (rounding level), or max_ir
    for (;; it++) {
        HMEC_TRY(precond_apply(c, r, nullptr, z));
        hipLaunchKernelGGL(k_add, dim3(g), dim3(256), 0, st, x, z, N, it == 0 ? 1 : 0);
        LAUNCH_CHECK();
        launch_residual(c, x, r, rsd, c->d.rscale);
        LAUNCH_CHECK();
        HMEC_TRY(launch_norms4(c, r, c->d.rhs, rsd, bs, s->norms));
        HMEC_HIP(hipMemcpyAsync(s->h_norms, s->norms, sizeof(double) * 7, hipMemcpyDeviceToHost, st));
        HMEC_HIP(hipStreamSynchronize(st));
        const double *h = s->h_norms;
        res = h[1] > 0 ? h[0] / h[1] : h[0];
        res_s = h[3] > 0 ? h[2] / h[3] : h[2];
        npert = (int)h[5];
        nbad = (int)h[6];
        if (nbad > 0 || (res <= tol && res_s <= tol) || it >= max_ir) break;
        if (it > 0 && !(res_s <= 0.5 * prev_s)) break;
        prev_s = res_s;
    }
    c->last_ir = it;
    c->last_res = res;
    c->last_perturbed = npert;
    if (out) {
        out->ir_iterations = it;
        out->rel_residual = res;
        out->perturbed_pivots = npert;
        out->factor_flops = c->plan.flops;
        out->factor_nnz = c->plan.factor_nnz;
        out->nfronts = c->plan.nfronts;
        out->nlevels = c->plan.nlevels;
        out->max_front = c->plan.max_front;
    }
    c->stale_solves++;
    if (nbad > 0) return HMEC_E_PIVOT;
    return (res <= tol) ? HMEC_OK : HMEC_E_NOCONV;
}

int sv_solve(hmec_ctx *c, int max_ir, double tol, hmec_solve_stats *out) {
    if (!c->factor_valid) return HMEC_E_ARG;
    SolverDev *s = sdev(c);
    if (!s->rsd) {
        HMEC_HIP(hipMalloc(&s->rsd, sizeof(double) * c->N));
        HMEC_HIP(hipMalloc(&s->bs, sizeof(double) * c->N));
    }
    return sv_solve_impl(c, max_ir, tol, out, s->rsd, s->bs);
}

// ------------------------------------------------ allocation-free solve
// hmec_solve_noalloc keeps the refinement's two work vectors in this
// file-scope device array instead of allocating them on first use: room for
// kNoallocMaxN unknowns (801x201 has 972,024), 2 x 8 MiB.  Every device the
// library runs on holds its own copy until the process ends; host code reaches
// it only through the address hipGetSymbolAddress returns for the current
// device (the array's name in host code is a stand-in, not that address).
// Both vectors are fully written before they are read on every call, so
// nothing needs clearing; calls that overlap on one device share the array and
// race.
constexpr int64_t kNoallocMaxN = int64_t(1) << 20;
static __device__ __align__(256) double scratch[2 * kNoallocMaxN];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};

// the scratch array's address on the current device (looked up once per device)
int scratch_address(int device, double **out) {
    if (device < 0 || device >= kMaxDevices) return HMEC_E_ARG;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        HMEC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            hmec_set_last_error("hipGetSymbolAddress(scratch) returned no address");
            return HMEC_E_HIP;
        }
        g_scratch[device] = static_cast<double *>(p);
    }
    *out = g_scratch[device];
    return HMEC_OK;
}
}  // namespace

extern "C" int hmec_solve_noalloc(hmec_ctx *c, hmec_solve_stats *st) {
    if (!c) return HMEC_E_ARG;
    if (c->N > kNoallocMaxN) {
        hmec_set_last_error("hmec_solve_noalloc: " + std::to_string(c->N) +
                            " unknowns, above the " + std::to_string(kNoallocMaxN) +
                            " its scratch array holds");
        return HMEC_E_ARG;
    }
    if (!c->factor_valid) return HMEC_E_ARG;
    HMEC_HIP(hipSetDevice(c->device));
    double *sp = nullptr;
    HMEC_TRY(scratch_address(c->device, &sp));
    const bool fp32 = c->factor_prec == HMEC_PREC_FP32;
    int rc = sv_solve_impl(c, fp32 ? 30 : c->fp64_max_ir, fp32 ? 1e-12 : c->fp64_tol, st, sp,
                           sp + kNoallocMaxN);
    if (rc == HMEC_E_NOCONV && !fp32 && c->last_res <= kAcceptResidual) rc = HMEC_OK;
    return rc;
}

int sv_batch_factorize(hmec_ctx *const *cs, int n, int precision) {
    return sv_factorize_batch(cs, n, precision);
}


// ---------------------------------------------------------------- GMRES
// Right-preconditioned restarted GMRES in FP64 on the original CSR, with the
// (possibly stale) factor as preconditioner: M^-1 v = Dc LU^-1 Dr v.  The
// initial guess is the context's last solution (warm start across Picard
// iterations and timesteps).  Classical Gram-Schmidt with one
// re-orthogonalization; every reduction is two-stage in a fixed order.
struct GmresBuf {
    double *V = nullptr, *Z = nullptr, *w = nullptr, *part = nullptr, *hdev = nullptr;
    double *bs = nullptr, *rsd = nullptr;   // Dr b, Dr (b - A x)
    double *hhost = nullptr;
    hipEvent_t ev_first = nullptr;          // the device-decided first step's outcome on the host
    int m = 0;
    int64_t N = 0;
};
static std::mutex g_gm_mu;
static std::map<hmec_ctx *, GmresBuf> g_gm;

void sv_gmres_free(hmec_ctx *c) {
    std::lock_guard<std::mutex> lk(g_gm_mu);
    auto it = g_gm.find(c);
    if (it == g_gm.end()) return;
    GmresBuf &g = it->second;
    hipFree(g.V); hipFree(g.Z); hipFree(g.w); hipFree(g.part); hipFree(g.hdev);
    hipFree(g.bs); hipFree(g.rsd);
    if (g.hhost) hipHostFree(g.hhost);
    if (g.ev_first) hipEventDestroy(g.ev_first);
    g_gm.erase(it);
}

static int gm_buf(hmec_ctx *c, int m, GmresBuf **out) {
    std::lock_guard<std::mutex> lk(g_gm_mu);
    GmresBuf &g = g_gm[c];
    if (g.m < m || g.N != c->N) {
        hipFree(g.V); hipFree(g.Z); hipFree(g.w); hipFree(g.part); hipFree(g.hdev);
        hipFree(g.bs); hipFree(g.rsd);
        if (g.hhost) hipHostFree(g.hhost);
        g = GmresBuf();
        g.m = m;
        g.N = c->N;
        HMEC_HIP(hipMalloc(&g.V, sizeof(double) * (m + 1) * c->N));
        HMEC_HIP(hipMalloc(&g.Z, sizeof(double) * m * c->N));
        HMEC_HIP(hipMalloc(&g.w, sizeof(double) * c->N));
        HMEC_HIP(hipMalloc(&g.bs, sizeof(double) * c->N));
        HMEC_HIP(hipMalloc(&g.rsd, sizeof(double) * c->N));
        const int64_t nbs = std::max<int64_t>(RB, (c->N + SPMV_ROWS - 1) / SPMV_ROWS);
        HMEC_HIP(hipMalloc(&g.part, sizeof(double) * nbs * (m + 2)));
        // three Hessenberg columns, the residual norms of a cycle start, the
        // closing residual norms and the first step's coefficient (device path)
        HMEC_HIP(hipMalloc(&g.hdev, sizeof(double) * (3 * (m + 2) + 24)));
        HMEC_HIP(hipHostMalloc(&g.hhost, sizeof(double) * (3 * (m + 2) + 24), hipHostMallocDefault));
    }
    *out = &g;
    return HMEC_OK;
}

// dot products of w with V_0..V_{nv-1} -> device out[0..nv)
static int gm_dots_dev(hmec_ctx *c, GmresBuf &g, const double *V, int nv, const double *w,
                       double *out) {
    hipStream_t st = c->stream;
    hipLaunchKernelGGL(k_dots_part, dim3(RB), dim3(256), 0, st, V, c->N, nv, w, c->N, g.part);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_dots_final, dim3(nv), dim3(256), 0, st, g.part, RB, nv, out);
    LAUNCH_CHECK();
    return HMEC_OK;
}

// dot products of w with V_0..V_{nv-1} -> host h[0..nv)
static int gm_dots(hmec_ctx *c, GmresBuf &g, const double *V, int nv, const double *w,
                   double *h) {
    hipStream_t st = c->stream;
    hipLaunchKernelGGL(k_dots_part, dim3(RB), dim3(256), 0, st, V, c->N, nv, w, c->N, g.part);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_dots_final, dim3(nv), dim3(256), 0, st, g.part, RB, nv, g.hdev);
    LAUNCH_CHECK();
    HMEC_HIP(hipMemcpyAsync(g.hhost, g.hdev, sizeof(double) * nv, hipMemcpyDeviceToHost, st));
    HMEC_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < nv; i++) h[i] = g.hhost[i];
    return HMEC_OK;
}

// z = A^-1 Dr^-1 v: right preconditioner of the row-equilibrated operator Dr A
static int gm_precond(hmec_ctx *c, const double *v, double *zout, VScale sc = VScale()) {
    return precond_apply(c, v, c->d.rscale, zout, sc);
}

// Relative inf-norm residuals of x: unscaled ||b - A x|| / ||b|| (the
// reported one) and row-equilibrated ||Dr (b - A x)|| / ||Dr b||; leaves
// Dr (b - A x) in g.rsd (the GMRES residual)
// (the residual kernel forms the norms' block partials with the residual,
// one more launch reduces them)
static int gm_residual_launch(hmec_ctx *c, GmresBuf &g, const double *x, double *out,
                              const double *gate = nullptr, double vtol = 0,
                              double *vstate = nullptr) {
    const int64_t N = c->N;
    const int nblk = (int)((N + SPMV_ROWS - 1) / SPMV_ROWS);
    hipLaunchKernelGGL(k_csr_stream<1>, dim3((unsigned)xcd_grid(nblk)), dim3(SPMV_ROWS), 0, c->stream,
                       c->d.rowp, c->d.coli, c->d.val, x, c->d.rscale, c->d.rhs, c->d.work, g.rsd, N,
                       (const double *)nullptr, (int64_t)0, 0, g.part, gate);
    LAUNCH_CHECK();
    hipLaunchKernelGGL((k_norms_final<4, true>), dim3(1), dim3(256), 0, c->stream, (const double *)g.part,
                       nblk, out, (const int *)c->d.counters, gate, vtol,
                       vstate ? c->d.scal : (double *)nullptr, vstate);
    LAUNCH_CHECK();
    return HMEC_OK;
}
// (h[5], h[6]: perturbed and unrepairable pivots of the current factor)
constexpr int GM_NRM = 7;   // residual norms of launch_norms4 (with the pivot counts)
static void gm_residual_read(const double *h, double *rel_u, double *rel_s, double *rr, int *npert,
                             int *nbad) {
    *rel_u = h[1] > 0 ? h[0] / h[1] : h[0];
    *rel_s = h[3] > 0 ? h[2] / h[3] : h[2];
    *rr = h[4];
    *npert = (int)h[5];
    *nbad = (int)h[6];
}
static int gm_true_residual(hmec_ctx *c, GmresBuf &g, const double *x, double *rel_u,
                            double *rel_s, double *rr, int *npert, int *nbad) {
    SolverDev *s = sdev(c);
    hipStream_t st = c->stream;
    HMEC_TRY(gm_residual_launch(c, g, x, s->norms));
    HMEC_HIP(hipMemcpyAsync(s->h_norms, s->norms, sizeof(double) * GM_NRM, hipMemcpyDeviceToHost, st));
    HMEC_HIP(hipStreamSynchronize(st));
    gm_residual_read(s->h_norms, rel_u, rel_s, rr, npert, nbad);
    return HMEC_OK;
}

// GMRES on the row-equilibrated system Dr A x = Dr b (every row weighted by
// its own scale, so small-magnitude rows such as the Darcy equations are
// resolved like the others), right-preconditioned by the FP32 factor.
// Converged when both the reported ||b - A x||_inf <= tol ||b||_inf and the
// equilibrated ||Dr (b - A x)||_inf <= tol ||Dr b||_inf hold.  Each restart
// cycle starts from the true residual; its Arnoldi loop stops once the
// estimated 2-norm has dropped by the factor the inf-norms still need (with
// a margin of 2).
// y = a u + b v + d w
__global__ void k_lincomb3(double *y, double a, const double *u, double b, const double *v, double d,
                           const double *w, int64_t N) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < N) y[t] = (a * u[t] + b * v[t]) + d * w[t];
}
int sv_lincomb3(hmec_ctx *c, double *y, double a, const double *u, double b, const double *v,
                double d, const double *w) {
    hipLaunchKernelGGL(k_lincomb3, dim3((unsigned)((c->N + 255) / 256)), dim3(256), 0, c->stream, y, a, u,
                       b, v, d, w, c->N);
    LAUNCH_CHECK();
    return HMEC_OK;
}
int sv_lincomb(hmec_ctx *c, double *y, double a, const double *u, double b, const double *v) {
    hipLaunchKernelGGL(k_lincomb, dim3((unsigned)((c->N + 255) / 256)), dim3(256), 0, c->stream, y, a, u, b, v,
                       c->N);
    LAUNCH_CHECK();
    return HMEC_OK;
}

// The first Arnoldi step of a solve decided on the device (hmec_step's
// Picard loop enqueues the rest of the iteration behind it without waiting):
//   GM_LEAD_OK   the start already meets the tolerance, x is kept;
//   GM_COMBINE   the step's residual estimate meets its target: x += y0 Z0,
//                then the closing residual's norm combine checks the true
//                residual;
//   GM_CONTINUE  more Arnoldi steps are needed (the host continues the cycle);
//   GM_HOLD      an unrepairable pivot or a zero residual norm: the host
//                decides, as sv_gmres does;
//   GM_DONE      verified: both residuals meet the tolerance;
//   GM_RESTART   the true residual misses it: the host starts the next cycle.
// GM_CONTINUE, GM_HOLD and GM_RESTART set the exit flag to hold, so the update
// and the next prologue enqueued behind the solve do nothing until the host
// has finished the solve and released the flag.
enum GmState { GM_NONE = 0, GM_LEAD_OK = 1, GM_COMBINE = 2, GM_CONTINUE = 3, GM_HOLD = 4,
               GM_DONE = 5, GM_RESTART = 6 };
static_assert(GM_LEAD_OK == kGmLeadOk && GM_CONTINUE == kGmContinue && GM_HOLD == kGmHold &&
                  GM_DONE == kGmDone && GM_RESTART == kGmRestart && GM_COMBINE == GM_GATE_COMBINE &&
                  GM_DONE == GM_GATE_DONE && GM_RESTART == GM_GATE_RESTART,
              "GmState values");

// nrm: |r|inf, |b|inf, |Dr r|inf, |Dr b|inf, |Dr r|2^2, perturbed pivots, unrepairable pivots
__device__ __forceinline__ void gm_rel(const double *nrm, double *rel_u, double *rel_s) {
    *rel_u = nrm[1] > 0 ? nrm[0] / nrm[1] : nrm[0];
    *rel_s = nrm[3] > 0 ? nrm[2] / nrm[3] : nrm[2];
}
// h: the first Hessenberg column (first and second Gram-Schmidt passes, the
// squared norm of the new vector) M2 apart, then the start's norms at 3 M2;
// y0 at 3 M2 + 16; the GmState at 3 M2 + 20 (beside the device scalars' copy,
// so one copy of h brings the step's whole outcome to the host)
// The first step's decision from h (the same in every thread that takes it)
__device__ __forceinline__ int gm_step0_eval(const double *h, int M2, double tol, double *y0) {
    const double *nrm = h + 3 * M2;
    double relinf, rel_s;
    gm_rel(nrm, &relinf, &rel_s);
    // (the larger as std::max takes it: a NaN first operand stays)
    const double worst = relinf < rel_s ? rel_s : relinf, beta = sqrt(nrm[4]);
    *y0 = 0.0;
    if ((int)nrm[6] > 0) return GM_HOLD;
    // a NaN norm lands here and is left to the host (not converged)
    if (!(worst > tol)) return (relinf <= tol && rel_s <= tol) ? GM_LEAD_OK : GM_HOLD;
    if (!(beta > 0)) return GM_HOLD;
    // from w . V_0 and w . w of the operator product alone (no Gram-Schmidt
    // pass): |w'|^2 = w . w - (w . V_0)^2.  Its cancellation only blurs the
    // estimate below ~1e-8 |w|; the true residual of every combined step is
    // checked after it.
    const double target = beta * (tol / worst) * 0.5;
    const double h0 = h[0], hn = sqrt(fmax(h[1] - h0 * h0, 0.0));
    const double den = hypot(h0, hn);
    const double cs = den > 0 ? h0 / den : 1.0, sn = den > 0 ? hn / den : 0.0;
    const double gv1 = -sn * beta, gv0 = cs * beta;
    if (fabs(gv1) <= target || !(hn > 0)) {
        *y0 = gv0 / den;
        return GM_COMBINE;
    }
    return GM_CONTINUE;
}
// The first step's decision and, when it meets its target, x += y0 Z0: every
// thread evaluates the decision from h (uniform loads), thread 0 of block 0
// records it (state in the device scalars and at h[3 M2 + 20], y0 at
// h[3 M2 + 16], the exit flag held for CONTINUE / HOLD).
__global__ void k_gm_step0_combine(double *h, int M2, double tol, double *scal, const double *Z0,
                                   double *x, int64_t N) {
    double y0;
    const int st = gm_step0_eval(h, M2, tol, &y0);
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) {
        scal[SC_GMSTATE_SLOT] = st;
        h[3 * M2 + 20] = st;
        if (st == GM_COMBINE) h[3 * M2 + 16] = y0;
        if (st == GM_CONTINUE || st == GM_HOLD) scal[SC_EXIT_SLOT] = SC_EXIT_HOLD;
    }
    if (t >= N || st != GM_COMBINE) return;
    double v = x[t];
    v += y0 * Z0[t];
    x[t] = v;
}
// operator scaling, E and Dr b of the current matrix (once per solve)
static int gm_setup(hmec_ctx *c, GmresBuf &g) {
    hipStream_t st = c->stream;
    const int64_t N = c->N;
    if (!c->rscale_ready) {   // the assembly forms it with the values
        hipLaunchKernelGGL(k_row_scale, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, c->d.rowp,
                           c->d.val, N, c->d.rscale, nullptr);
        LAUNCH_CHECK();
        c->rscale_ready = true;
    }
    return darcy_e(c, st, nullptr, c->d.rhs, c->d.rscale, g.bs);   // and bs = Dr b
}

// One Arnoldi step j of the current cycle (V_j given): Z_j = M^-1 V_j, the
// twice-orthogonalized A Z_j with its coefficients left on the device, the
// next basis vector, and the Hessenberg column copied to the host (with the
// cycle start's norms when nrm is set; no copy when copy is unset); not
// waited for.  Without gs only Z_j, w = A Z_j and its first-pass dot
// products with V_0..V_j plus w . w (at h[j + 1]) are formed: the first step
// of a solve decided on the device (k_gm_step0_combine) needs no more, and gm_gs
// completes the step when the cycle goes on.
static int gm_gs(hmec_ctx *c, GmresBuf &g, int j, int m, bool nrm, bool copy);
// (scaled: V_0 = Dr r / ||Dr r||_2 is formed by the preconditioner's first
// kernels from g.rsd and the start's squared norm)
static int gm_arnoldi(hmec_ctx *c, GmresBuf &g, int j, int m, bool nrm, bool copy = true,
                      bool gs = true, bool scaled = false) {
    hipStream_t st = c->stream;
    const int64_t N = c->N;
    const int M2 = m + 2;
    double *Zj = g.Z + (int64_t)j * N;
    VScale sc;
    if (scaled) {
        sc.src = g.rsd;
        sc.rr = g.hdev + 3 * M2 + 4;
        sc.vout = g.V + (int64_t)j * N;
    }
    HMEC_TRY(gm_precond(c, g.V + (int64_t)j * N, Zj, sc));
    // classical Gram-Schmidt, twice, coefficients kept on the device;
    // one copy of the Hessenberg column per iteration.  The first
    // pass's dot products come with the operator product, the second
    // pass's with ||w'||^2, and its update writes the next basis vector.
    double *h1d = g.hdev;
    const unsigned nbs = (unsigned)((N + SPMV_ROWS - 1) / SPMV_ROWS);
    const int nq = j + 1 + (gs ? 0 : 1);
    launch_spmv_rs_dots(c, Zj, g.w, g.V, j + 1, g.part, gs ? 0 : 1);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_dots_final, dim3(nq), dim3(256), 0, st, g.part, (int)nbs, nq, h1d);
    LAUNCH_CHECK();
    if (!gs) return HMEC_OK;
    return gm_gs(c, g, j, m, nrm, copy);
}
// The Gram-Schmidt passes of Arnoldi step j after its operator product and
// first-pass dot products (gm_arnoldi)
static int gm_gs(hmec_ctx *c, GmresBuf &g, int j, int m, bool nrm, bool copy) {
    hipStream_t st = c->stream;
    const int64_t N = c->N;
    const unsigned gN = (unsigned)((N + 255) / 256);
    const int M2 = m + 2;
    double *h1d = g.hdev, *h2d = g.hdev + M2, *wwd = g.hdev + 2 * M2;
#define GS_GO(NVM)                                                                          \
    hipLaunchKernelGGL(k_gs_update_dots<NVM>, dim3(RB), dim3(256), 0, st, g.V, N, j + 1, h1d, g.w, \
                       N, g.part)
    if (j + 1 <= 4) GS_GO(4);
    else if (j + 1 <= 8) GS_GO(8);
    else if (j + 1 <= 16) GS_GO(16);
    else GS_GO(GM_MAXV);
#undef GS_GO
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_dots_final, dim3(j + 2), dim3(256), 0, st, g.part, RB, j + 2, h2d);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_gs_update_norm, dim3(gN), dim3(256), 0, st, g.V, N, j + 1, h2d, g.w,
                       j + 1 < m ? g.V + (int64_t)(j + 1) * N : nullptr, wwd, N);
    LAUNCH_CHECK();
    if (copy)
        HMEC_HIP(hipMemcpyAsync(g.hhost, g.hdev, sizeof(double) * (3 * M2 + (nrm ? GM_NRM : 0)),
                                hipMemcpyDeviceToHost, st));
    return HMEC_OK;
}

constexpr int GM_M = 20;   // restart length (<= 31: GmCoef and the fused dots hold the coefficients)
static_assert(GM_M + 1 <= GM_MAXV, "restart length above the fused dot-product width");

// Enqueues a solve's start and its first Arnoldi step with the step's
// decision on the device (GmState, in the device scalars); nothing waits.
// The host reads the outcome with the scalars and completes the solve with
// sv_gmres(..., resume) when the state asks for it.
int sv_gmres_first(hmec_ctx *c, double tol) {
    if (!c->factor_valid) return HMEC_E_ARG;
    GmresBuf *gp;
    HMEC_TRY(gm_buf(c, GM_M, &gp));
    GmresBuf &g = *gp;
    const int64_t N = c->N;
    const unsigned gN = (unsigned)((N + 255) / 256);
    hipStream_t st = c->stream;
    const int M2 = GM_M + 2;
    double *nrm_d = g.hdev + 3 * M2, *nrm2_d = g.hdev + 3 * M2 + 8;
    HMEC_TRY(gm_setup(c, g));
    HMEC_TRY(gm_residual_launch(c, g, c->d.sol, nrm_d));
    HMEC_TRY(gm_arnoldi(c, g, 0, GM_M, true, false, false, true));
    hipLaunchKernelGGL(k_gm_step0_combine, dim3(gN), dim3(256), 0, st, g.hdev, M2, tol, c->d.scal,
                       (const double *)g.Z, c->d.sol, N);
    LAUNCH_CHECK();
    // only after a combine: the true residual, checked against tol
    HMEC_TRY(gm_residual_launch(c, g, c->d.sol, nrm2_d, c->d.scal, tol, g.hdev + 3 * M2 + 20));
    // the column, both residuals' norms, y0 and the state in one copy
    HMEC_HIP(hipMemcpyAsync(g.hhost, g.hdev, sizeof(double) * (3 * M2 + 21), hipMemcpyDeviceToHost, st));
    if (!g.ev_first) HMEC_HIP(hipEventCreateWithFlags(&g.ev_first, hipEventDisableTiming));
    HMEC_HIP(hipEventRecord(g.ev_first, st));
    return HMEC_OK;
}

// Waits for the outcome of sv_gmres_first (before the work enqueued behind
// it has run) and returns its GmState
int sv_gmres_first_state(hmec_ctx *c, int *state) {
    GmresBuf *gp;
    HMEC_TRY(gm_buf(c, GM_M, &gp));
    if (!gp->ev_first) return HMEC_E_ARG;
    HMEC_HIP(hipEventSynchronize(gp->ev_first));
    *state = (int)gp->hhost[3 * (GM_M + 2) + 20];
    return HMEC_OK;
}

// The outcome of sv_gmres_first for a state that needs no host work
// (GM_LEAD_OK, GM_DONE), after the host has waited for the scalars
int sv_gmres_first_result(hmec_ctx *c, int state, hmec_solve_stats *out) {
    GmresBuf *gp;
    HMEC_TRY(gm_buf(c, GM_M, &gp));
    const int M2 = GM_M + 2;
    const double *h = gp->hhost + 3 * M2 + (state == GM_DONE ? 8 : 0);
    double relinf, rel_s, rr;
    int npert, nbad;
    gm_residual_read(h, &relinf, &rel_s, &rr, &npert, &nbad);
    const int it = state == GM_DONE ? 1 : 0;
    c->last_ir = it;
    c->last_res = relinf;
    if (out) {
        out->ir_iterations = it;
        out->rel_residual = relinf;
        out->perturbed_pivots = npert;
        out->factor_flops = c->plan.flops;
        out->factor_nnz = c->plan.factor_nnz;
        out->nfronts = c->plan.nfronts;
        out->nlevels = c->plan.nlevels;
        out->max_front = c->plan.max_front;
    }
    c->stale_solves++;
    return HMEC_OK;
}

// resume: 0 a fresh solve; GM_CONTINUE / GM_RESTART / GM_HOLD the solve
// sv_gmres_first left to the host (its setup done; CONTINUE: the first
// step's operator product and the start's norms on the device;
// RESTART: one iteration done, the next cycle starts from the true residual)
int sv_gmres(hmec_ctx *c, int max_it, double tol, hmec_solve_stats *out, int resume) {
    if (!c->factor_valid) return HMEC_E_ARG;
    const int m = GM_M;
    GmresBuf *gp;
    HMEC_TRY(gm_buf(c, m, &gp));
    GmresBuf &g = *gp;
    const int64_t N = c->N;
    const unsigned gN = (unsigned)((N + 255) / 256);
    hipStream_t st = c->stream;
    double *x = c->d.sol;
    double *r = g.rsd;
    int total_it = resume == GM_RESTART ? 1 : 0, npert = 0, nbad = 0;
    double relinf = 0, last_rel_s = 0;
    std::vector<double> H((m + 1) * m), cs(m), sn(m), gv(m + 1), h(m + 2), h2(m + 2), y(m);
    if (resume == 0) HMEC_TRY(gm_setup(c, g));
    // The first cycle does not wait for its starting residual: the first
    // basis vector is scaled on the device and the residual norms come back
    // with the first Hessenberg column.  A start that already meets the
    // tolerance discards that iteration (x is not touched).
    const int M2 = m + 2;
    double *const nrm_d = g.hdev + 3 * M2;
    bool lead = resume != GM_RESTART && total_it < max_it, done = false;
    bool have0 = resume == GM_CONTINUE;   // the first column is already on the host
    if (have0) {
        // sv_gmres_first formed the first step's operator product only: its
        // Gram-Schmidt passes and the next basis vector now
        HMEC_TRY(gm_gs(c, g, 0, m, true, true));
        HMEC_HIP(hipStreamSynchronize(st));
    }
    for (;;) {
        double rel_s = 0, rr = 0;   // rr = ||Dr (b - A x)||_2^2 from the same pass
        double target = 0;
        std::fill(gv.begin(), gv.end(), 0.0);
        if (lead) {
            if (!have0) {
                HMEC_TRY(gm_residual_launch(c, g, x, nrm_d));
                hipLaunchKernelGGL(k_vscale_rr, dim3(gN), dim3(256), 0, st, r, nrm_d + 4, g.V, N);
                LAUNCH_CHECK();
            }
        } else {
            HMEC_TRY(gm_true_residual(c, g, x, &relinf, &rel_s, &rr, &npert, &nbad));
            last_rel_s = rel_s;
            const double worst = std::max(relinf, rel_s);
            if (nbad > 0 || !(worst > tol) || total_it >= max_it) break;
            const double beta = std::sqrt(rr);
            if (!(beta > 0)) break;
            target = beta * (tol / worst) * 0.5;
            hipLaunchKernelGGL(k_vscale, dim3(gN), dim3(256), 0, st, r, 1.0 / beta, g.V, N);
            LAUNCH_CHECK();
            gv[0] = beta;
        }
        int j = 0;
        for (; j < m && total_it < max_it; j++) {
            if (!(have0 && j == 0)) {
                HMEC_TRY(gm_arnoldi(c, g, j, m, lead && j == 0));
                HMEC_HIP(hipStreamSynchronize(st));
            }
            const bool first = lead && j == 0;
            if (first) {
                gm_residual_read(g.hhost + 3 * M2, &relinf, &rel_s, &rr, &npert, &nbad);
                last_rel_s = rel_s;
                const double worst = std::max(relinf, rel_s);
                const double beta = std::sqrt(rr);
                if (nbad > 0 || !(worst > tol) || !(beta > 0)) {
                    done = true;
                    break;
                }
                target = beta * (tol / worst) * 0.5;
                gv[0] = beta;
            }
            for (int i = 0; i <= j; i++) h[i] = g.hhost[i] + g.hhost[M2 + i];
            const double ww = g.hhost[2 * M2];
            const double hn = std::sqrt(ww);
            h[j + 1] = hn;
            total_it++;
            // previous rotations, then the new one
            for (int i = 0; i < j; i++) {
                const double t1 = cs[i] * h[i] + sn[i] * h[i + 1];
                h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1];
                h[i] = t1;
            }
            const double den = std::hypot(h[j], h[j + 1]);
            cs[j] = den > 0 ? h[j] / den : 1.0;
            sn[j] = den > 0 ? h[j + 1] / den : 0.0;
            h[j] = den;
            h[j + 1] = 0;
            gv[j + 1] = -sn[j] * gv[j];
            gv[j] = cs[j] * gv[j];
            for (int i = 0; i <= j; i++) H[i + j * (m + 1)] = h[i];
            if (std::fabs(gv[j + 1]) <= target || !(hn > 0)) {
                j++;
                break;
            }
        }
        lead = false;
        have0 = false;
        if (done || j == 0) break;
        // y = R^-1 g ; x += Z y
        for (int i = j - 1; i >= 0; i--) {
            double t1 = gv[i];
            for (int k = i + 1; k < j; k++) t1 -= H[i + k * (m + 1)] * y[k];
            y[i] = t1 / H[i + i * (m + 1)];
        }
        GmCoef cy;
        for (int i = 0; i < j; i++) cy.v[i] = y[i];
        hipLaunchKernelGGL(k_combine, dim3(gN), dim3(256), 0, st, g.Z, N, j, cy, x, N);
        LAUNCH_CHECK();
    }
    c->last_ir = total_it;
    c->last_res = relinf;
    if (out) {
        out->ir_iterations = total_it;
        out->rel_residual = relinf;
        out->perturbed_pivots = npert;
        out->factor_flops = c->plan.flops;
        out->factor_nnz = c->plan.factor_nnz;
        out->nfronts = c->plan.nfronts;
        out->nlevels = c->plan.nlevels;
        out->max_front = c->plan.max_front;
    }
    c->stale_solves++;
    if (nbad > 0) return HMEC_E_PIVOT;
    return (relinf <= tol && last_rel_s <= tol) ? HMEC_OK : HMEC_E_NOCONV;
}
