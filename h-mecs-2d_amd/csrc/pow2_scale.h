// Power-of-two equilibration factor shared by the assembly and the solver.
#pragma once
#include <hip/hip_runtime.h>

// 2^-e with m in [2^(e-1), 2^e): a power of two near 1/m (exact scaling);
// 1 for zero or non-finite m
__device__ __forceinline__ double pow2_recip(double m) {
    if (!(m > 0) || !isfinite(m)) return 1.0;
    int e;
    frexp(m, &e);
    return ldexp(1.0, -e + 1);
}
