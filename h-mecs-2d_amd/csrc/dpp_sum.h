// Lane-group sums of doubles with DPP lane moves (register to register),
// instead of ds_bpermute, which goes through the LDS crossbar once per
// 32-bit half.  Groups are aligned runs of L lanes of a 64-lane wavefront.
#pragma once
#include <hip/hip_runtime.h>

// DPP move of a double (its two 32-bit halves): lane i takes the value of
// the lane CTRL selects; rows outside ROWS keep 0.
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ double dpp_f64(double x) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, ROWS, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, ROWS, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Sum over each aligned group of L lanes (L = 1, 2, 4, 8, 16), left in every
// lane of the group: row_mirror (16), row_half_mirror (8), then the quad
// permutations [2,3,0,1] and [1,0,3,2].  Every lane of the group must be
// active.
template <int L>
__device__ __forceinline__ double group_sum(double x) {
    static_assert(L == 1 || L == 2 || L == 4 || L == 8 || L == 16, "lane group");
    if (L >= 16) x += dpp_f64<0x140>(x);
    if (L >= 8) x += dpp_f64<0x141>(x);
    if (L >= 4) x += dpp_f64<0x4E>(x);
    if (L >= 2) x += dpp_f64<0xB1>(x);
    return x;
}

// Sum over each aligned group of L lanes (L = 1 ... 64), valid in the
// group's last lane (lane L - 1 of the group): the sums of the 16-lane rows,
// then row_bcast:15 (lane 15 of a row into rows 1 and 3) and row_bcast:31
// (lane 31 into rows 2 and 3).  Every lane of the group must be active.
template <int L>
__device__ __forceinline__ double group_total(double x) {
    static_assert(L >= 1 && L <= 64 && (L & (L - 1)) == 0, "lane group");
    x = group_sum<(L < 16 ? L : 16)>(x);
    if (L >= 32) x += dpp_f64<0x142, 0xA>(x);
    if (L >= 64) x += dpp_f64<0x143, 0xC>(x);
    return x;
}

// The same with any commutative, associative operation op (e.g. a maximum):
// lanes of rows outside a row_bcast's mask combine with 0, so op(x, 0) must
// leave x unchanged on the values reduced (sums; maxima of values >= 0).
template <int L, class Op>
__device__ __forceinline__ double group_total(double x, Op op) {
    static_assert(L >= 1 && L <= 64 && (L & (L - 1)) == 0, "lane group");
    if (L >= 16) x = op(x, dpp_f64<0x140>(x));
    if (L >= 8) x = op(x, dpp_f64<0x141>(x));
    if (L >= 4) x = op(x, dpp_f64<0x4E>(x));
    if (L >= 2) x = op(x, dpp_f64<0xB1>(x));
    if (L >= 32) x = op(x, dpp_f64<0x142, 0xA>(x));
    if (L >= 64) x = op(x, dpp_f64<0x143, 0xC>(x));
    return x;
}

// Reduction over a 256-thread workgroup: each wavefront's total (in its lane
// 63) through 4 LDS slots, combined in the fixed order ((w0, w1), (w2, w3));
// the result is returned to every thread.  sh: 4 doubles of LDS; every
// thread of the workgroup must call it.
template <class Op>
__device__ __forceinline__ double block_total256(double x, Op op, double *sh) {
    x = group_total<64>(x, op);
    if ((threadIdx.x & 63) == 63) sh[threadIdx.x >> 6] = x;
    __syncthreads();
    const double r = op(op(sh[0], sh[1]), op(sh[2], sh[3]));
    __syncthreads();
    return r;
}
