// Block placement across the XCDs.
//
// Workgroup b of a launch runs on XCD b % 8 (the dispatcher's round-robin
// order; measured with the XCC_ID register on an idle device).  Placement is
// a locality hint only: no result depends on it.  xcd_block gives each XCD a
// contiguous range of the n logical blocks of a launch, so blocks that read
// the same data (neighbouring nodes' marker cells, overlapping ranges of the
// operator's x) share one XCD's L2 instead of each XCD fetching them from
// HBM.  Launch xcd_grid(n) workgroups; a workgroup whose xcd_block(n) >= n
// has nothing to do.
#pragma once
#include <hip/hip_runtime.h>

constexpr int HMEC_XCDS = 8;

inline int64_t xcd_grid(int64_t n) { return (n + HMEC_XCDS - 1) / HMEC_XCDS * HMEC_XCDS; }

__device__ __forceinline__ int64_t xcd_block(int64_t n) {
    const int64_t chunk = (n + HMEC_XCDS - 1) / HMEC_XCDS;
    return (int64_t)(blockIdx.x % HMEC_XCDS) * chunk + (int64_t)(blockIdx.x / HMEC_XCDS);
}
