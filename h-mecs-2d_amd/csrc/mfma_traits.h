// gfx950 MFMA 16x16x4 wrappers shared by the dense front kernels.
//
// v_mfma_*_16x16x4: lane l supplies A(l%16, l/16) and B(l/16, l%16); element
// r of the 4-wide result holds D(row(l, r), l%16) with row = 4*(l/16)+r for
// f32 and 4*r+l/16 for f64 (layouts measured on gfx950 with an in-bounds
// probe of both builtins).
#pragma once
#include <hip/hip_runtime.h>

typedef double hmec_v4d __attribute__((ext_vector_type(4)));
typedef float hmec_v4f __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef hmec_v4d acc_t;
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * r + (lane >> 4); }
};
template <> struct Mfma<float> {
    typedef hmec_v4f acc_t;
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};
