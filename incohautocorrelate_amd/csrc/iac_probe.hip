// iac_probe.hip -- measurement kernels: the device-memory ceilings the pair
// engine is compared against.  uint64 atomic adds spread uniformly over a mesh,
// concentrated on a hot region (the occupied centre slab of a sparse run), or
// into an LDS box (the engine's privatised path), and a plain streaming copy.
// Also the exhaustive self-check of the short correctly rounded square root.
#include <algorithm>

#include "iac_fpexact.h"
#include "iac_internal.h"

namespace {

__device__ __forceinline__ uint32_t mix32(uint64_t x)
{
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return (uint32_t)x;
}

// mode 0: uniform over nvox; mode 1: a hot window of `hot` voxels at the middle;
// mode 3: runs of 8 consecutive voxels (one 64-byte segment) at uniform random
// segments, lanes 8k..8k+7 of a wave on one run
__global__ void __launch_bounds__(256)
    global_atomic_kernel(unsigned long long *__restrict__ mesh, uint64_t nvox, uint64_t n, int mode, uint64_t hot)
{
    const uint64_t base = mode == 1 ? (nvox - hot) / 2 : 0;
    const uint64_t span = mode == 1 ? hot : nvox;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        atomicAdd(mesh + (mode == 3 ? (mix32(i >> 3) % (nvox >> 3)) * 8 + (i & 7) : base + mix32(i) % span), 1ull);
}

// LDS box of 4096 uint64 (T = unsigned long long) or double (the AC1D histograms' adds)
// per block, flushed once
template <typename T>
__global__ void __launch_bounds__(256)
    lds_atomic_kernel(unsigned long long *__restrict__ mesh, uint64_t nvox, uint64_t n_per_block)
{
    __shared__ T box[4096];
    for (int v = threadIdx.x; v < 4096; v += 256) box[v] = T(0);
    __syncthreads();
    const uint64_t i0 = (uint64_t)blockIdx.x * n_per_block;
    for (uint64_t i = threadIdx.x; i < n_per_block; i += 256) atomicAdd(&box[mix32(i0 + i) & 4095u], T(1));
    __syncthreads();
    const uint64_t off = ((uint64_t)blockIdx.x * 4096) % (nvox - 4096 + 1);
    for (int v = threadIdx.x; v < 4096; v += 256)
        if (box[v] != T(0)) atomicAdd(mesh + off + v, (unsigned long long)box[v]);
}

__global__ void __launch_bounds__(256) copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// every float bit pattern in [lo, hi): iac::sqrt_rn against the library sqrtf,
// bit for bit (NaN results compared as NaN)
__global__ void __launch_bounds__(256) sqrt_check_kernel(uint64_t lo, uint64_t hi, unsigned long long *bad)
{
    unsigned long long nbad = 0;
    for (uint64_t u = lo + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < hi;
         u += (uint64_t)gridDim.x * blockDim.x) {
        const float x = __uint_as_float((uint32_t)u);
        const float a = iac::sqrt_rn(x), b = sqrtf(x);
        if (__float_as_uint(a) != __float_as_uint(b) && !(a != a && b != b)) nbad++;
    }
    if (nbad) atomicAdd(bad, nbad);
}

}  // namespace

extern "C" {

int iac_selfcheck_sqrt(iac_ctx *ctx, uint64_t lo, uint64_t hi, uint64_t *mismatches)
{
    IAC_TRY(iac::bind(ctx));
    if (!mismatches || lo > hi || hi > (1ull << 32)) return iac::fail(IAC_ERR_INVALID_ARG, "bad sqrt check range");
    void *d = nullptr;
    IAC_TRY(iac::scratch(ctx, 0, sizeof(uint64_t), &d));
    IAC_HIP(hipMemsetAsync(d, 0, sizeof(uint64_t), ctx->stream));
    if (hi > lo) {
        const unsigned blocks = (unsigned)std::max(1, ctx->num_cus * 16);
        hipLaunchKernelGGL(sqrt_check_kernel, dim3(blocks), dim3(256), 0, ctx->stream, lo, hi,
                           static_cast<unsigned long long *>(d));
        IAC_HIP(hipGetLastError());
    }
    IAC_HIP(hipMemcpyAsync(mismatches, d, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    IAC_HIP(hipStreamSynchronize(ctx->stream));
    return IAC_OK;
}

int iac_probe_atomics(iac_ctx *ctx, uint64_t *d_mesh, uint64_t nvox, uint64_t n_updates, int mode, uint64_t hot,
                      float *ms)
{
    IAC_TRY(iac::bind(ctx));
    if (!d_mesh || !ms || nvox < 4096 || n_updates == 0) return iac::fail(IAC_ERR_INVALID_ARG, "bad probe");
    if (mode < 0 || mode > 4) return iac::fail(IAC_ERR_INVALID_ARG, "probe mode is 0 .. 4");
    if (mode == 1 && (hot == 0 || hot > nvox)) return iac::fail(IAC_ERR_INVALID_ARG, "bad hot window");
    hipEvent_t a, b;
    IAC_HIP(hipEventCreate(&a));
    IAC_HIP(hipEventCreate(&b));
    const unsigned blocks = (unsigned)std::max(1, ctx->num_cus * 16);
    IAC_HIP(hipEventRecord(a, ctx->stream));
    if (mode == 0 || mode == 1 || mode == 3) {
        hipLaunchKernelGGL(global_atomic_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                           reinterpret_cast<unsigned long long *>(d_mesh), nvox, n_updates, mode, hot);
    } else {
        const uint64_t per = std::max<uint64_t>(1, n_updates / blocks);
        if (mode == 2)
            hipLaunchKernelGGL(lds_atomic_kernel<unsigned long long>, dim3(blocks), dim3(256), 0, ctx->stream,
                               reinterpret_cast<unsigned long long *>(d_mesh), nvox, per);
        else
            hipLaunchKernelGGL(lds_atomic_kernel<double>, dim3(blocks), dim3(256), 0, ctx->stream,
                               reinterpret_cast<unsigned long long *>(d_mesh), nvox, per);
    }
    IAC_HIP(hipGetLastError());
    IAC_HIP(hipEventRecord(b, ctx->stream));
    IAC_HIP(hipEventSynchronize(b));
    IAC_HIP(hipEventElapsedTime(ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return IAC_OK;
}

int iac_probe_copy(iac_ctx *ctx, const void *d_src, void *d_dst, size_t bytes, float *ms)
{
    IAC_TRY(iac::bind(ctx));
    if (!d_src || !d_dst || !ms || bytes < 16 || bytes % 16) return iac::fail(IAC_ERR_INVALID_ARG, "bad copy probe");
    hipEvent_t a, b;
    IAC_HIP(hipEventCreate(&a));
    IAC_HIP(hipEventCreate(&b));
    const unsigned blocks = (unsigned)std::max(1, ctx->num_cus * 8);
    IAC_HIP(hipEventRecord(a, ctx->stream));
    hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, ctx->stream, static_cast<const uint4 *>(d_src),
                       static_cast<uint4 *>(d_dst), bytes / 16);
    IAC_HIP(hipGetLastError());
    IAC_HIP(hipEventRecord(b, ctx->stream));
    IAC_HIP(hipEventSynchronize(b));
    IAC_HIP(hipEventElapsedTime(ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return IAC_OK;
}

}  // extern "C"
