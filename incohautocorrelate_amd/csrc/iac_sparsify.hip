// iac_sparsify.hip -- frame sparsification on gfx950: the mask multiply
// (ArrayOperators::ParMultiplyElementwise, IAC/ArrayOperators.h:120-127), the
// hit test and photon discretisation of Detector::CreateSparseHitList
// (IAC/Detector.cpp:926-958, ArrayOperators::DiscretizeToPhotons :138-148),
// producing {kx, ky, kz, value} records in pixel order per frame.
#include <algorithm>
#include <cmath>

#include "iac_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ bool hit(const float *frame, const int *mask, uint32_t p, float T, float &I)
{
    I = frame[p];
    if (mask) I = I * (float)mask[p];
    return I >= T && I > 0;
}

__device__ __forceinline__ float discretize(float I, float T, float S)
{
    return S > 0 ? ceilf((I - T) / S) * (float)(I >= T) : I;
}

// count[f] = kept pixels of frame f (one block per frame)
__global__ void __launch_bounds__(kBlock) count_kernel(const float *__restrict__ frames, const int *__restrict__ mask,
                                                       uint32_t npix, float T, uint64_t *__restrict__ count)
{
    __shared__ uint32_t part[kBlock / 64];
    const float *fr = frames + (size_t)blockIdx.x * npix;
    uint32_t c = 0;
    for (uint32_t p = threadIdx.x; p < npix; p += kBlock) {
        float I;
        c += hit(fr, mask, p, T, I) ? 1u : 0u;
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kBlock / 64; w++) t += part[w];
        count[blockIdx.x] = t;
    }
}

// exclusive scan count[0..n) -> off[0..n]
__global__ void __launch_bounds__(1024) scan_kernel(const uint64_t *__restrict__ count, uint32_t n,
                                                    uint64_t *__restrict__ off)
{
    __shared__ uint64_t part[1024];
    const uint32_t t = threadIdx.x, per = (n + 1023) / 1024;
    const uint32_t f0 = min(n, t * per), f1 = min(n, f0 + per);
    uint64_t s = 0;
    for (uint32_t f = f0; f < f1; f++) s += count[f];
    part[t] = s;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1) {
        const uint64_t v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint64_t run = t ? part[t - 1] : 0;
    for (uint32_t f = f0; f < f1; f++) {
        off[f] = run;
        run += count[f];
    }
    if (t == 1023) off[n] = part[1023];
}

// stable compaction of frame blockIdx.x's hits into records starting at off[f]
__global__ void __launch_bounds__(kBlock)
    compact_kernel(const float *__restrict__ frames, const int *__restrict__ mask, const float *__restrict__ kmap,
                   uint32_t npix, float T, float S, const uint64_t *__restrict__ off, float4 *__restrict__ rec)
{
    __shared__ uint32_t wcount[kBlock / 64];
    __shared__ uint32_t base_s;
    const float *fr = frames + (size_t)blockIdx.x * npix;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) base_s = 0;
    __syncthreads();
    for (uint32_t p0 = 0; p0 < npix; p0 += kBlock) {
        const uint32_t p = p0 + threadIdx.x;
        float I = 0.0f;
        const bool h = p < npix && hit(fr, mask, p, T, I);
        const unsigned long long bal = __ballot(h);
        const uint32_t before = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wcount[wave] = (uint32_t)__popcll(bal);
        __syncthreads();
        uint32_t wbase = base_s;
        for (int w = 0; w < wave; w++) wbase += wcount[w];
        if (h) {
            const uint64_t o = off[blockIdx.x] + wbase + before;
            rec[o] = make_float4(kmap[3 * (size_t)p], kmap[3 * (size_t)p + 1], kmap[3 * (size_t)p + 2],
                                 discretize(I, T, S));
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t t = 0;
            for (int w = 0; w < kBlock / 64; w++) t += wcount[w];
            base_s += t;
        }
        __syncthreads();
    }
}

}  // namespace

extern "C" {

int iac_sparse_hit_list(iac_ctx *ctx, const float *d_kmap, uint32_t npix, const float *d_frames, uint32_t nframes,
                        const int *d_mask, float threshold, float step, float *d_rec4, uint64_t *d_frame_off,
                        uint64_t *n_total)
{
    IAC_TRY(iac::bind(ctx));
    if (!d_kmap || !d_frames || !d_rec4 || !d_frame_off || !n_total || npix == 0)
        return iac::fail(IAC_ERR_INVALID_ARG, "null input");
    *n_total = 0;
    if (nframes == 0) return IAC_OK;
    void *cnt = nullptr;
    IAC_TRY(iac::scratch(ctx, 1, (size_t)nframes * sizeof(uint64_t), &cnt));
    hipLaunchKernelGGL(count_kernel, dim3(nframes), dim3(kBlock), 0, ctx->stream, d_frames, d_mask, npix, threshold,
                       static_cast<uint64_t *>(cnt));
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, ctx->stream, static_cast<const uint64_t *>(cnt), nframes,
                       d_frame_off);
    IAC_TRY(iac::timer_begin(ctx));
    hipLaunchKernelGGL(compact_kernel, dim3(nframes), dim3(kBlock), 0, ctx->stream, d_frames, d_mask, d_kmap, npix,
                       threshold, step, d_frame_off, reinterpret_cast<float4 *>(d_rec4));
    IAC_HIP(hipGetLastError());
    IAC_TRY(iac::timer_end(ctx));
    uint64_t n = 0;
    IAC_HIP(hipMemcpyAsync(&n, d_frame_off + nframes, sizeof(n), hipMemcpyDeviceToHost, ctx->stream));
    IAC_HIP(hipStreamSynchronize(ctx->stream));
    *n_total = n;
    return IAC_OK;
}

}  // extern "C"
