// iac_serial_sum.h -- ArrayOperators::Sum<float> (IAC/ArrayOperators.h:260-268) of one frame per
// wave: the float sum in pixel order, bit for bit.  The 64 lanes load 64 consecutive pixels
// (one 256-byte access) and v_readlane feeds them, in order, into a running sum that every lane
// holds; the frame's chain of adds stays serial, but each frame has a SIMD of its own.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace iac {

constexpr uint32_t kSumWave = 64;

// value(j, v): what pixel j of value v adds to the sum.  Called by all 64 lanes of a one-wave
// block; returns the frame's sum in every lane.
template <class V>
__device__ __forceinline__ float serial_sum_wave(const float *__restrict__ fr, uint32_t npix, V value)
{
    const uint32_t lane = threadIdx.x;
    float s = 0.0f;
    uint32_t j0 = 0;
    for (; j0 + kSumWave <= npix; j0 += kSumWave) {
        const int v = __float_as_int(value(j0 + lane, fr[j0 + lane]));
#pragma unroll
        for (int i = 0; i < (int)kSumWave; i++) s += __int_as_float(__builtin_amdgcn_readlane(v, i));
    }
    if (j0 < npix) {
        const int v = __float_as_int(j0 + lane < npix ? value(j0 + lane, fr[j0 + lane]) : 0.0f);
        for (uint32_t i = 0; i < npix - j0; i++) s += __int_as_float(__builtin_amdgcn_readlane(v, (int)i));
    }
    return s;
}

}  // namespace iac
