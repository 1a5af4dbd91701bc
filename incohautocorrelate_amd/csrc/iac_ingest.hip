// iac_ingest.hip -- the averaged-intensity pass over a stack of dense frames
// (Detector::LoadAndAverageIntensity, IAC/Detector.cpp:687-790) on gfx950, with
// the reference's float arithmetic: pixel mask multiply, threshold or photon
// discretisation, per-pixel accumulation over frames in frame order, the
// 1/N scale, and the per-event MeanIntensity / PhotonCount sums.
#include <algorithm>
#include <cmath>

#include "iac_fpexact.h"
#include "iac_internal.h"
#include "iac_serial_sum.h"

#pragma clang fp contract(off)

namespace {

// ArrayOperators::DiscretizeToPhotons, IAC/ArrayOperators.h:138-148
__device__ __forceinline__ float discretize(float I, float T, float S) { return ceilf((I - T) / S) * (float)(I >= T); }

// The per-frame value tmpIntensity[j] after masking and the branch of
// Detector.cpp:718-760 (the value the event sums see).
__device__ __forceinline__ float frame_value(float I, int m, bool has_mask, float T, float S)
{
    float v = has_mask ? I * (float)m : I;
    if (S > 0) return v >= T ? discretize(v, T, S) : 0.0f;
    if (T > 0 && v < T) return 0.0f;
    return v;
}

// One thread per pixel, frames in order: adds to the float sum (ParAdd with
// threshold) or to the long photon tally; calls over consecutive frame stacks
// continue the same per-pixel sequence.
__global__ void __launch_bounds__(256)
    accumulate_pixels_kernel(const float *__restrict__ frames, uint32_t nframes, uint32_t npix,
                             const int *__restrict__ mask, float T, float S, float *__restrict__ sum,
                             long long *__restrict__ tally)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= npix) return;
    const int m = mask ? mask[j] : 1;
    if (S > 0) {
        long long acc = tally[j];
        for (uint32_t f = 0; f < nframes; f++) {
            const float v = frame_value(frames[(size_t)f * npix + j], m, mask != nullptr, T, S);
            if (v != 0.0f) acc += (long long)floor((double)v + 0.5);
        }
        tally[j] = acc;
    } else {
        float acc = sum[j];
        for (uint32_t f = 0; f < nframes; f++) {
            const float v = frame_value(frames[(size_t)f * npix + j], m, mask != nullptr, T, S);
            if (v >= T) acc += v;
        }
        sum[j] = acc;
    }
}

// Intensity = (float)tally (discretised) or the float sum, x (1.0f / N)
__global__ void finish_kernel(uint32_t npix, uint32_t ntotal, float S, const float *__restrict__ sum,
                              const long long *__restrict__ tally, float *__restrict__ avg)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= npix) return;
    const float v = S > 0 ? (float)tally[j] : sum[j];
    avg[j] = v * (1.0f / (float)ntotal);
}

// One wave per frame: ArrayOperators::Sum<float> (IAC/ArrayOperators.h:260-268) of the frame's
// values, a serial float sum in pixel order; MeanIntensity = sum / P and, with photon
// discretisation, PhotonCount = (int) sum.
__global__ void __launch_bounds__(iac::kSumWave)
    frame_sums_kernel(const float *__restrict__ frames, uint32_t npix, const int *__restrict__ mask, float T, float S,
                      float *__restrict__ mean_intensity, int *__restrict__ photon_count)
{
    const uint32_t f = blockIdx.x;
    const float s = iac::serial_sum_wave(frames + (size_t)f * npix, npix, [=](uint32_t j, float v) {
        return frame_value(v, mask ? mask[j] : 1, mask != nullptr, T, S);
    });
    if (threadIdx.x != 0) return;
    if (mean_intensity) mean_intensity[f] = s / (float)npix;
    if (photon_count && S > 0) photon_count[f] = iac::cvtt_i32(s);
}

}  // namespace

extern "C" {

int iac_intensity_accumulate(iac_ctx *ctx, const float *d_frames, uint32_t nframes, uint32_t npix, const int *d_mask,
                             float threshold, float step, float *d_sum, int64_t *d_tally, float *d_mean_intensity,
                             int *d_photon_count)
{
    IAC_TRY(iac::bind(ctx));
    if (!d_frames || npix == 0) return iac::fail(IAC_ERR_INVALID_ARG, "bad averaging request");
    if (step > 0 ? !d_tally : !d_sum) return iac::fail(IAC_ERR_INVALID_ARG, "missing accumulator");
    if (nframes == 0) return IAC_OK;
    hipLaunchKernelGGL(accumulate_pixels_kernel, dim3((npix + 255) / 256), dim3(256), 0, ctx->stream, d_frames,
                       nframes, npix, d_mask, threshold, step, d_sum, reinterpret_cast<long long *>(d_tally));
    IAC_HIP(hipGetLastError());
    if (d_mean_intensity || d_photon_count) {
        if (nframes > 0x7fffffffu) return iac::fail(IAC_ERR_LIMIT, "too many frames in one stack");
        hipLaunchKernelGGL(frame_sums_kernel, dim3(nframes), dim3(iac::kSumWave), 0, ctx->stream, d_frames, npix,
                           d_mask, threshold, step, d_mean_intensity, d_photon_count);
        IAC_HIP(hipGetLastError());
    }
    return IAC_OK;
}

int iac_intensity_finish(iac_ctx *ctx, uint32_t npix, uint32_t ntotal_frames, float step, const float *d_sum,
                         const int64_t *d_tally, float *d_avg)
{
    IAC_TRY(iac::bind(ctx));
    if (!d_avg || npix == 0) return iac::fail(IAC_ERR_INVALID_ARG, "bad averaging request");
    if (ntotal_frames == 0) return iac::fail(IAC_ERR_PRECONDITION, "event list is empty, nothing to average");
    if (step > 0 ? !d_tally : !d_sum) return iac::fail(IAC_ERR_INVALID_ARG, "missing accumulator");
    hipLaunchKernelGGL(finish_kernel, dim3((npix + 255) / 256), dim3(256), 0, ctx->stream, npix, ntotal_frames, step,
                       d_sum, reinterpret_cast<const long long *>(d_tally), d_avg);
    IAC_HIP(hipGetLastError());
    return IAC_OK;
}

int iac_average_intensity(iac_ctx *ctx, const float *d_frames, uint32_t nframes, uint32_t npix, const int *d_mask,
                          float threshold, float step, float *d_avg, float *d_mean_intensity, int *d_photon_count)
{
    IAC_TRY(iac::bind(ctx));
    if (!d_frames || !d_avg || npix == 0) return iac::fail(IAC_ERR_INVALID_ARG, "bad averaging request");
    if (nframes == 0) return iac::fail(IAC_ERR_PRECONDITION, "event list is empty, nothing to average");
    void *acc = nullptr;
    IAC_TRY(iac::scratch(ctx, 0, (size_t)npix * 8, &acc));
    IAC_HIP(hipMemsetAsync(acc, 0, (size_t)npix * 8, ctx->stream));
    float *sum = step > 0 ? nullptr : static_cast<float *>(acc);
    int64_t *tally = step > 0 ? static_cast<int64_t *>(acc) : nullptr;
    IAC_TRY(iac_intensity_accumulate(ctx, d_frames, nframes, npix, d_mask, threshold, step, sum, tally,
                                     d_mean_intensity, d_photon_count));
    return iac_intensity_finish(ctx, npix, nframes, step, sum, tally, d_avg);
}

}  // extern "C"
