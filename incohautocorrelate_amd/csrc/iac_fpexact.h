// iac_fpexact.h -- correctly rounded float square root and division by a bin
// width, in fewer instructions than the general library forms, for the AC1D
// pair kernels (each pair takes one |q| = sqrtf(...) and one q / dq); and the
// reference's x86 float -> int conversion of per-event sums.
#pragma once

#include <hip/hip_runtime.h>

namespace iac {

// sqrtf(x) correctly rounded for x == +-0, x >= 2^-96, +inf and NaN: the raw
// v_sqrt_f32 (within 1 ulp there), then the neighbour whose square brackets x.
// s - 1 ulp is taken when x <= (s - ulp) s and s + 1 ulp when x > (s + ulp) s,
// the residuals computed exactly by fma.
__device__ __forceinline__ float sqrt_rn_normal(float x)
{
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sd = __uint_as_float(__float_as_uint(s) - 1u);
    const float su = __uint_as_float(__float_as_uint(s) + 1u);
    const float rd = __builtin_fmaf(-sd, s, x);
    const float ru = __builtin_fmaf(-su, s, x);
    return ru > 0.f ? su : (rd <= 0.f ? sd : s);
}

// sqrtf(x) for every x: a wave holding any nonzero x < 2^-96 (negative inputs,
// whose subnormals the raw instruction flushes to -0, and squared distances of
// nearly coincident pixels, in practice never) takes the library sqrtf on a
// branch of its own; the empty volatile asm keeps it from being if-converted.
__device__ __forceinline__ float sqrt_rn(float x)
{
    if (__builtin_amdgcn_ballot_w64(x != 0.f && x < 0x1p-96f) != 0) {
        asm volatile("");
        return sqrtf(x);
    }
    return sqrt_rn_normal(x);
}

// (int)x as the reference's x86-64 build converts it (cvttss2si): truncation toward zero, and
// 0x80000000 ("integer indefinite") for NaN and for values outside the int range, where the
// GPU's conversion would saturate.
__device__ __forceinline__ int cvtt_i32(float x)
{
    return (x >= -0x1p31f && x < 0x1p31f) ? (int)x : (-2147483647 - 1);
}

}  // namespace iac
