// iac_ac1d.hip -- AC1D: angular-averaged (radial) pair histograms on gfx950.
//
//   iac_ac1d_cq:   AutoCorr_CQ_AV (IAC/kernel.cl:656-764), every ordered pixel
//                  pair of one intensity image, uint64 fixed point per bin.
//   iac_ac1d_acuw: Calculate_AC_UW_Mapped's CPU branch (IAC/AC1D.cpp:256-293),
//                  photon pairs j <= k of each frame weighted 2 (self pairs
//                  too, as the reference does), accumulated in double.
//   iac_ac1d_calc_ac: AC1D::CalcAC (IAC/AC1D.cpp:482-493).
//
// |q| and its bin are the same for (a,b) and (b,a), so the C(q) kernel bins
// each unordered pair once with twice its value.  Every block keeps a private
// histogram of all bins in LDS across many 64x64 pair tiles (persistent grid)
// and flushes it once.  Integer photon counts in nearest-neighbour mode make
// every AC_uw value an integer-valued float: those runs accumulate in uint64
// and convert at the end, which reproduces the reference's double sums bit
// for bit (all partial sums are integers below 2^53).
#include <algorithm>
#include <cmath>
#include <mutex>
#include <numeric>
#include <type_traits>
#include <vector>

#include "iac_fpexact.h"
#include "iac_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr int kTile = 64;
constexpr int kBlock = 256;
constexpr int kBPerThread = kTile / (kBlock / kTile);  // 16
constexpr int kMaxBins = 6144;                         // 48 KiB of uint64 / double in LDS

__device__ __forceinline__ uint64_t tri_tiles(uint64_t n)
{
    const uint64_t nt = (n + kTile - 1) / kTile;
    return nt * (nt + 1) / 2;
}

__device__ __forceinline__ void decode_tri(uint64_t l, uint32_t nt, uint32_t &ta, uint32_t &tb)
{
    const double b = 2.0 * nt + 1.0;
    int ri = (int)floor((b - sqrt(b * b - 8.0 * (double)l)) * 0.5);
    ri = max(0, min(ri, (int)nt - 1));
    auto S = [nt](int rr) -> uint64_t { return (uint64_t)rr * nt - (uint64_t)rr * (uint64_t)(rr - 1) / 2; };
    while (ri > 0 && S(ri) > l) ri--;
    while (ri + 1 < (int)nt && S(ri + 1) <= l) ri++;
    ta = (uint32_t)ri;
    tb = (uint32_t)(ri + (l - S(ri)));
}

// |k_a - k_b| as the reference computes it: a float expression of repeated
// differences, then a correctly rounded float sqrt.
__device__ __forceinline__ float qnorm(const float4 &a, const float4 &b)
{
    return iac::sqrt_rn((a.x - b.x) * (a.x - b.x) + (a.y - b.y) * (a.y - b.y) + (a.z - b.z) * (a.z - b.z));
}

// q / dq rounded to float.  kFast: q * RN(1/dq) in double, then rounded to
// float.  The exact quotient of two floats is never within 2^-49 (relative) of a
// float rounding midpoint (a midpoint m has an odd 25-bit significand, so q - m dq
// is a nonzero multiple of a power of two at least q 2^-49), and the double
// product errs by at most 2^-52, so both round to the same float.  The host takes
// this path only while every quotient of a binned pair is 0 or a normal float.
template <bool kFast>
__device__ __forceinline__ float div_dq(float q, float dq, double inv_dq)
{
    return kFast ? (float)((double)q * inv_dq) : q / dq;
}

// (unsigned long long)(long long)x for x >= 0 as the reference truncates its
// weighted values.  Values below 2^32 (the usual case: the multiplicator scales
// I_i I_j / <I>^2 to about 1e5..1e6) take the one-instruction u32 conversion;
// a wave with any larger (or NaN) value takes the full 64-bit conversion on a
// branch of its own (the empty volatile asm keeps it from being if-converted).
__device__ __forceinline__ unsigned long long trunc_u64(double x)
{
    if (__builtin_amdgcn_ballot_w64(!(x < 4294967296.0)) != 0) {
        asm volatile("");
        return (unsigned long long)(long long)x;
    }
    return (uint32_t)x;
}

__device__ __forceinline__ void trunc_u64x2(double x1, double x2, unsigned long long &v1, unsigned long long &v2)
{
    if (__builtin_amdgcn_ballot_w64(!(x1 < 4294967296.0 && x2 < 4294967296.0)) != 0) {
        asm volatile("");
        v1 = (unsigned long long)(long long)x1;
        v2 = (unsigned long long)(long long)x2;
    } else {
        v1 = (uint32_t)x1;
        v2 = (uint32_t)x2;
    }
}

// AutoCorr_CQ_AV over one image: records are {k, I} of every pixel.
template <bool kFastDiv, bool kNoDrop>
__global__ void __launch_bounds__(kBlock)
    ac1d_cq_kernel(const float4 *__restrict__ rec, uint32_t n, uint32_t size, float max_q, float dq, double inv_dq,
                   int interp, double mult, unsigned long long *__restrict__ hist,
                   unsigned long long *__restrict__ dropped)
{
    // dynamic LDS: the block's histogram (size bins), then the staged b tile
    extern __shared__ float4 smem[];
    unsigned long long *h = reinterpret_cast<unsigned long long *>(smem);
    float4 *sb = smem + (size + 1) / 2;
    for (uint32_t i = threadIdx.x; i < size; i += kBlock) h[i] = 0ull;
    unsigned long long drop = 0;
    const uint64_t total = tri_tiles(n);
    const uint32_t nt = (n + kTile - 1) / kTile;
    const int lane_a = threadIdx.x & (kTile - 1), bgrp = threadIdx.x / kTile;
    // a contiguous range of tiles per block, row-major over the triangle (ta <= tb)
    const uint64_t t0 = total * blockIdx.x / gridDim.x, t1 = total * (blockIdx.x + 1) / gridDim.x;
    uint32_t ta = 0, tb = 0;
    if (t0 < t1) decode_tri(t0, nt, ta, tb);
    for (uint64_t t = t0; t < t1; t++) {
        if (t > t0 && ++tb == nt) {
            ta++;
            tb = ta;
        }
        __syncthreads();
        if (threadIdx.x < kTile) {
            const uint32_t b = tb * kTile + threadIdx.x;
            sb[threadIdx.x] = b < n ? rec[b] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        __syncthreads();
        const uint32_t a = ta * kTile + lane_a;
        if (a >= n) continue;
        const float4 ra = rec[a];
        // pairs j >= i of the image with I_i I_j >= 1e-37 and |q| <= Max_Q.  An edge tile (the
        // diagonal, or the ragged last column) tests b; an interior one holds 64 x 64 distinct pairs.
        auto tile = [&](auto edge_c) {
            constexpr bool kEdge = decltype(edge_c)::value;
#pragma unroll 4
            for (int jj = 0; jj < kBPerThread; jj++) {
                const int bl = bgrp * kBPerThread + jj;
                const uint32_t b = tb * kTile + bl;
                const float4 rb = sb[bl];
                const float w = ra.w * rb.w;
                float q = qnorm(ra, rb);
                if (kEdge && (b >= n || (ta == tb && b < a))) continue;
                if (w < 1e-37f || q > max_q) continue;
                const double val = (double)w;
                q = div_dq<kFastDiv>(q, dq, inv_dq);
                const unsigned sh = kEdge ? (a != b) : 1u;  // off-diagonal pairs count twice
                if (interp == 0) {
                    const uint32_t sc = (uint32_t)floor((double)q + 0.5);
                    const unsigned long long v = trunc_u64(val * mult);
                    if (kNoDrop) atomicAdd(&h[min(sc, size - 1)], v << sh);
                    else if (sc < size) atomicAdd(&h[sc], v << sh);
                    else drop += 1u << sh;
                } else {
                    const uint32_t sc1 = (uint32_t)floorf(q), sc2 = sc1 + 1;
                    const double sep = (double)(q - floorf(q));
                    const double x1 = (val * (1 - sep)) * mult, x2 = (val * sep) * mult;
                    unsigned long long v1, v2;
                    trunc_u64x2(x1, x2, v1, v2);
                    if (kNoDrop) {
                        const uint32_t s1 = min(sc1, size - 2);
                        atomicAdd(&h[s1], v1 << sh);
                        atomicAdd(&h[s1 + 1], v2 << sh);
                    } else {
                        if (sc1 < size) atomicAdd(&h[sc1], v1 << sh); else drop += 1u << sh;
                        if (sc2 < size) atomicAdd(&h[sc2], v2 << sh); else drop += 1u << sh;
                    }
                }
            }
        };
        if (ta == tb || tb + 1 == nt) tile(std::true_type{}); else tile(std::false_type{});
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < size; i += kBlock)
        if (h[i]) atomicAdd(hist + i, h[i]);
    if (drop) atomicAdd(dropped, drop);
}

// Exclusive scan of per-frame triangle tile counts (diagonal included).
__global__ void __launch_bounds__(1024) tile_scan_kernel(const uint64_t *__restrict__ frame_off, uint32_t nframes,
                                                         uint64_t *__restrict__ tile_off)
{
    __shared__ uint64_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (nframes + 1023) / 1024;
    const uint32_t f0 = min(nframes, t * per), f1 = min(nframes, f0 + per);
    uint64_t s = 0;
    for (uint32_t f = f0; f < f1; f++) s += tri_tiles(frame_off[f + 1] - frame_off[f]);
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
        tile_off[f] = run;
        run += tri_tiles(frame_off[f + 1] - frame_off[f]);
    }
    if (t == 1023) tile_off[nframes] = part[1023];
}

__global__ void gather_kernel(const float *__restrict__ kmap, uint32_t npix, const uint32_t *__restrict__ pix,
                              const uint16_t *__restrict__ cnt, uint64_t n, float4 *__restrict__ rec,
                              unsigned long long *invalid)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t p = pix[i];
        if (p < npix) {
            rec[i] = make_float4(kmap[3 * (size_t)p], kmap[3 * (size_t)p + 1], kmap[3 * (size_t)p + 2], (float)cnt[i]);
        } else {
            rec[i] = make_float4(0.f, 0.f, 0.f, 0.f);
            atomicAdd(invalid, 1ull);
        }
    }
}

// Record i holds pixel (i * stride) mod npix, stride coprime to npix: the 64
// records of a tile lie far apart on the detector, so the 64 lanes of a wave
// pairing them with one b pixel bin into different |q| and seldom hit the same
// histogram word.  The integer histogram does not depend on the pixel order.
__global__ void gather_image_kernel(const float *__restrict__ kmap, const float *__restrict__ I, uint32_t npix,
                                    uint32_t stride, float4 *__restrict__ rec)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += gridDim.x * blockDim.x) {
        const size_t p = (size_t)i * stride % npix;
        rec[i] = make_float4(kmap[3 * p], kmap[3 * p + 1], kmap[3 * p + 2], I[p]);
    }
}

// AC_uw: photon pairs j <= k of every frame.  kExact: integer-valued weights
// accumulate in uint64 (nearest neighbour); otherwise double.
template <bool kExact, bool kFastDiv, bool kNoDrop>
__global__ void __launch_bounds__(kBlock)
    ac1d_acuw_kernel(const float4 *__restrict__ rec, const uint64_t *__restrict__ frame_off,
                     const uint64_t *__restrict__ tile_off, uint32_t nframes, uint32_t size, float max_q, float dq,
                     double inv_dq, int interp, float self_factor, unsigned long long *__restrict__ hist_u,
                     double *__restrict__ hist_d, unsigned long long *__restrict__ dropped)
{
    // dynamic LDS: the block's histogram (size bins of 8 bytes), then the staged b tile
    extern __shared__ float4 smem[];
    unsigned long long *hu = reinterpret_cast<unsigned long long *>(smem);
    double *hd = reinterpret_cast<double *>(smem);
    float4 *sb = smem + (size + 1) / 2;
    for (uint32_t i = threadIdx.x; i < size; i += kBlock) {
        if (kExact) hu[i] = 0ull; else hd[i] = 0.0;
    }
    unsigned long long drop = 0;
    const uint64_t total = tile_off[nframes];
    const int lane_a = threadIdx.x & (kTile - 1), bgrp = threadIdx.x / kTile;
    // a contiguous range of tiles per block: the frame is searched once, then advanced
    const uint64_t t0 = total * blockIdx.x / gridDim.x, t1 = total * (blockIdx.x + 1) / gridDim.x;
    uint32_t lo = 0;
    {
        uint32_t hi = nframes;  // the frame of t0: largest f with tile_off[f] <= t0
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (tile_off[mid] <= t0) lo = mid; else hi = mid;
        }
    }
    uint32_t ta = 0, tb = 0, cur = 0xffffffffu;  // tile (ta, tb) of frame cur
    for (uint64_t t = t0; t < t1; t++) {
        while (tile_off[lo + 1] <= t) lo++;  // frames without tiles are skipped
        const uint64_t base = frame_off[lo];
        const uint32_t n = (uint32_t)(frame_off[lo + 1] - base);
        const uint32_t nt = (n + kTile - 1) / kTile;
        if (lo != cur) {  // first tile of the block in this frame
            decode_tri(t - tile_off[lo], nt, ta, tb);
            cur = lo;
        } else if (++tb == nt) {  // next tile, row-major over the triangle
            ta++;
            tb = ta;
        }
        __syncthreads();
        if (threadIdx.x < kTile) {
            const uint32_t b = tb * kTile + threadIdx.x;
            sb[threadIdx.x] = b < n ? rec[base + b] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        __syncthreads();
        const uint32_t a = ta * kTile + lane_a;
        if (a >= n) continue;
        const float4 ra = rec[base + a];
        // an edge tile (the diagonal, or the frame's ragged last column) tests b
        auto tile = [&](auto edge_c) {
            constexpr bool kEdge = decltype(edge_c)::value;
#pragma unroll 4
            for (int jj = 0; jj < kBPerThread; jj++) {
                const int bl = bgrp * kBPerThread + jj;
                const uint32_t b = tb * kTile + bl;
                const float4 rb = sb[bl];
                float q = qnorm(ra, rb);
                if (kEdge && (b >= n || (ta == tb && b < a))) continue;
                if (q > max_q) continue;
                q = div_dq<kFastDiv>(q, dq, inv_dq);
                const float f2 = (kEdge && a == b) ? self_factor : 2.0f;
                if (interp == 0) {
                    uint32_t sc = (uint32_t)floor((double)q + 0.5);
                    const float v = f2 * (ra.w * rb.w);
                    if (kNoDrop || sc < size) {
                        if (kNoDrop) sc = min(sc, size - 1);
                        if (kExact) atomicAdd(&hu[sc], (unsigned long long)v);
                        else atomicAdd(&hd[sc], (double)v);
                    } else {
                        drop++;
                    }
                } else {
                    const uint32_t sc1 = (uint32_t)floorf(q), sc2 = sc1 + 1;
                    const float sep = q - (float)sc1;
                    const float v1 = f2 * ((ra.w * rb.w) * (1 - sep));
                    const float v2 = f2 * ((ra.w * rb.w) * sep);
                    if (kNoDrop) {
                        const uint32_t s1 = min(sc1, size - 2);
                        atomicAdd(&hd[s1], (double)v1);
                        atomicAdd(&hd[s1 + 1], (double)v2);
                    } else {
                        if (sc1 < size) atomicAdd(&hd[sc1], (double)v1); else drop++;
                        if (sc2 < size) atomicAdd(&hd[sc2], (double)v2); else drop++;
                    }
                }
            }
        };
        if (ta == tb || tb + 1 == nt) tile(std::true_type{}); else tile(std::false_type{});
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < size; i += kBlock) {
        if (kExact) {
            if (hu[i]) atomicAdd(hist_u + i, hu[i]);
        } else {
            if (hd[i] != 0.0) atomicAdd(hist_d + i, hd[i]);
        }
    }
    if (drop) atomicAdd(dropped, drop);
}

__global__ void add_u64_as_f64_kernel(const unsigned long long *__restrict__ in, uint32_t n, double *__restrict__ out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] += (double)in[i];
}

__global__ void zero_u64_kernel(unsigned long long *p, uint32_t n)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0ull;
}

__global__ void calc_ac_kernel(const double *__restrict__ ac_uw, const double *__restrict__ cq, uint32_t n,
                               double *__restrict__ ac)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double v = ac_uw[i] / cq[i];
        if (isnan(v)) v = 0;
        ac[i] = v;
    }
}

unsigned persistent_blocks(iac_ctx *ctx) { return (unsigned)std::max(1, ctx->num_cus * 4); }

// div_dq<true> is exact while dq is a normal float, every binned |q| <= max_q
// keeps q / dq <= 2^126, and the smallest nonzero |q| (2^-75, above the square
// root of the smallest subnormal) keeps q / dq >= 2^-126.
bool fast_div_ok(float max_q, float dq)
{
    return std::isnormal(dq) && dq > 0.f && dq <= 0x1p51f && (double)max_q <= (double)dq * 0x1p126;
}

// Every binned |q| <= Max_Q has q / dq <= fl(Max_Q / dq) (rounding is monotone); below
// size - 1 no bin (nearest, or both linear ones) can leave the histogram, so the
// kernels skip the range test (AC1D::Initialize's dq = Max_Q / (Size - 2) always
// passes).  They still clamp, which changes nothing then and keeps a NaN k in range.
bool no_drop(float max_q, float dq, uint32_t size)
{
    const float r = max_q / dq;
    return size >= 2 && r >= 0.f && r < (float)(size - 1);
}
// dynamic LDS of the AC_uw kernel: the histogram, then one staged tile of records
size_t acuw_lds(uint32_t size) { return (size_t)(size + 1) / 2 * sizeof(float4) + kTile * sizeof(float4); }

}  // namespace

namespace {

// Work memory of iac_ac1d_acuw_records_noalloc: a file-scope device array (one copy per
// GPU, held until the process ends) sized for kNoAllocMaxFrames frames and kMaxBins bins.
// Host code never passes `scratch` itself to a kernel (its host-side name is not a device
// address): noalloc_scratch looks up each device's copy once, with that device current.
constexpr uint32_t kNoAllocMaxFrames = 65536;
constexpr size_t kNoAllocScratchBytes = (size_t)(kNoAllocMaxFrames + 1) * 8 + (size_t)kMaxBins * 8;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_noalloc_mutex;
std::vector<void *> g_noalloc_ptr;  // per device: the address of its copy of `scratch`

// The device address of `device`'s copy of scratch (the caller has made that device current).
int noalloc_scratch(int device, void **out)
{
    std::lock_guard<std::mutex> g(g_noalloc_mutex);
    if ((size_t)device >= g_noalloc_ptr.size()) g_noalloc_ptr.resize((size_t)device + 1, nullptr);
    if (!g_noalloc_ptr[device]) {
        void *p = nullptr;
        IAC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_noalloc_ptr[device] = p;
    }
    *out = g_noalloc_ptr[device];
    return IAC_OK;
}

int acuw_check(uint32_t size, int interp, int self_pair_factor)
{
    if (size == 0 || size > (uint32_t)kMaxBins) return iac::fail(IAC_ERR_LIMIT, "AC1D size must be 1..6144 bins");
    if (interp != IAC_NEAREST && interp != IAC_LINEAR) return iac::fail(IAC_ERR_INVALID_ARG, "bad interpolation");
    if (self_pair_factor != 1 && self_pair_factor != 2)
        return iac::fail(IAC_ERR_INVALID_ARG, "self_pair_factor is 1 or 2");
    return IAC_OK;
}

// bytes of work memory of acuw_run: the frames' tile offsets, then (exact bins) the uint64 histogram
size_t acuw_work_bytes(uint32_t nframes, uint32_t size)
{
    return (size_t)(nframes + 1) * sizeof(uint64_t) + (size_t)size * 8;
}

// Shared tail of the AC_uw entry points: frame tiles, then the histogram
// kernel (uint64 bins when every weight is an integer, else double bins), with
// acuw_work_bytes of work memory at s1 (every word is written before it is read).
int acuw_run(iac_ctx *ctx, void *s1, const float4 *rec, const uint64_t *d_frame_off, uint32_t nframes, uint32_t size,
             float max_q, float dq, int interp, int self_pair_factor, bool exact, double *d_hist, bool timed)
{
    IAC_TRY(acuw_check(size, interp, self_pair_factor));
    auto *toff = static_cast<uint64_t *>(s1);
    auto *hu = reinterpret_cast<unsigned long long *>(static_cast<char *>(s1) + (size_t)(nframes + 1) * 8);
    hipLaunchKernelGGL(tile_scan_kernel, dim3(1), dim3(1024), 0, ctx->stream, d_frame_off, nframes, toff);
    IAC_HIP(hipGetLastError());
    const float sf = (float)self_pair_factor;
    const bool fdiv = fast_div_ok(max_q, dq), nd = no_drop(max_q, dq, size);
    const double inv_dq = 1.0 / (double)dq;
    auto kern = exact ? (fdiv ? (nd ? ac1d_acuw_kernel<true, true, true> : ac1d_acuw_kernel<true, true, false>)
                              : (nd ? ac1d_acuw_kernel<true, false, true> : ac1d_acuw_kernel<true, false, false>))
                      : (fdiv ? (nd ? ac1d_acuw_kernel<false, true, true> : ac1d_acuw_kernel<false, true, false>)
                              : (nd ? ac1d_acuw_kernel<false, false, true> : ac1d_acuw_kernel<false, false, false>));
    if (timed) IAC_TRY(iac::timer_begin(ctx));
    if (exact) hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(256), 0, ctx->stream, hu, size);
    hipLaunchKernelGGL(kern, dim3(persistent_blocks(ctx)), dim3(kBlock), acuw_lds(size), ctx->stream, rec,
                       d_frame_off, toff, nframes, size, max_q, dq, inv_dq, interp, sf, hu, d_hist, ctx->d_counters);
    IAC_HIP(hipGetLastError());
    if (timed) IAC_TRY(iac::timer_end(ctx));
    if (exact) {
        hipLaunchKernelGGL(add_u64_as_f64_kernel, dim3(std::max(1u, (size + 255) / 256)), dim3(256), 0, ctx->stream,
                           hu, size, d_hist);
        IAC_HIP(hipGetLastError());
    }
    return IAC_OK;
}

}  // namespace

extern "C" {


int iac_ac1d_shape(const float max_q[3], uint32_t size, float q_zoom, float *max_q_out, float *dq_out)
{
    if (!max_q || !max_q_out || !dq_out || size < 3) return iac::fail(IAC_ERR_INVALID_ARG, "bad AC1D shape request");
    const float m = std::sqrt(max_q[0] * max_q[0] + max_q[1] * max_q[1] + max_q[2] * max_q[2]) / q_zoom;
    *max_q_out = m;
    *dq_out = m / ((float)(size - 2));
    return IAC_OK;
}

int iac_ac1d_cq_multiplicator(const float *intensity, uint32_t npix, double *mult)
{
    if (!intensity || !mult || npix == 0) return iac::fail(IAC_ERR_INVALID_ARG, "null input");
    double t = 0;
    for (uint32_t i = 0; i < npix; i++) t += intensity[i];
    const float mean = (float)(t / npix);
    double m = 1e-10;
    for (; 1 > mean * mean * m;) m *= 10;
    m = m * 1e5;
    if (m > 1e18) m = 1e18;
    *mult = m;
    return IAC_OK;
}

int iac_ac1d_cq(iac_ctx *ctx, const float *d_kmap, uint32_t npix, const float *d_intensity, uint32_t size,
                float max_q, float dq, int interp, double mult, uint64_t *d_hist)
{
    IAC_TRY(iac::bind(ctx));
    if (!d_kmap || !d_intensity || !d_hist || npix == 0) return iac::fail(IAC_ERR_INVALID_ARG, "null input");
    if (size == 0 || size > (uint32_t)kMaxBins) return iac::fail(IAC_ERR_LIMIT, "AC1D size must be 1..6144 bins");
    if (interp != IAC_NEAREST && interp != IAC_LINEAR) return iac::fail(IAC_ERR_INVALID_ARG, "bad interpolation");
    void *rec = nullptr;
    IAC_TRY(iac::scratch(ctx, 0, (size_t)npix * sizeof(float4), &rec));
    uint32_t stride = npix / kTile + 1;  // about npix / 64, coprime to npix
    while (std::gcd(stride, npix) != 1) stride++;
    hipLaunchKernelGGL(gather_image_kernel, dim3(std::min<uint32_t>(4096, (npix + 255) / 256)), dim3(256), 0,
                       ctx->stream, d_kmap, d_intensity, npix, stride, static_cast<float4 *>(rec));
    IAC_HIP(hipGetLastError());
    IAC_TRY(iac::timer_begin(ctx));
    const bool fdiv = fast_div_ok(max_q, dq), nd = no_drop(max_q, dq, size);
    auto kern = fdiv ? (nd ? ac1d_cq_kernel<true, true> : ac1d_cq_kernel<true, false>)
                     : (nd ? ac1d_cq_kernel<false, true> : ac1d_cq_kernel<false, false>);
    hipLaunchKernelGGL(kern, dim3(persistent_blocks(ctx)), dim3(kBlock), acuw_lds(size), ctx->stream,
                       static_cast<const float4 *>(rec), npix, size, max_q, dq, 1.0 / (double)dq, interp, mult,
                       reinterpret_cast<unsigned long long *>(d_hist), ctx->d_counters);
    IAC_HIP(hipGetLastError());
    IAC_TRY(iac::timer_end(ctx));
    return IAC_OK;
}

int iac_ac1d_acuw(iac_ctx *ctx, const float *d_kmap, uint32_t npix, const uint32_t *d_pix, const uint16_t *d_cnt,
                  const uint64_t *d_frame_off, uint32_t nframes, uint64_t n_total, uint32_t size, float max_q,
                  float dq, int interp, int self_pair_factor, double *d_hist)
{
    IAC_TRY(iac::bind(ctx));
    if (!d_kmap || !d_frame_off || !d_hist || npix == 0) return iac::fail(IAC_ERR_INVALID_ARG, "null input");
    if (nframes == 0 || n_total == 0) return IAC_OK;
    if (!d_pix || !d_cnt) return iac::fail(IAC_ERR_PRECONDITION, "no photon lists");
    void *s0 = nullptr;
    IAC_TRY(iac::scratch(ctx, 0, n_total * sizeof(float4), &s0));
    auto *rec = static_cast<float4 *>(s0);
    const unsigned gb = (unsigned)std::min<uint64_t>(65535, (n_total + 255) / 256);
    hipLaunchKernelGGL(gather_kernel, dim3(gb), dim3(256), 0, ctx->stream, d_kmap, npix, d_pix, d_cnt, n_total, rec,
                       ctx->d_counters + 1);
    IAC_HIP(hipGetLastError());
    // uint16 counts: every weight is an integer-valued float, exact in uint64 bins
    IAC_TRY(acuw_check(size, interp, self_pair_factor));
    void *s1 = nullptr;
    IAC_TRY(iac::scratch(ctx, 1, acuw_work_bytes(nframes, size), &s1));
    return acuw_run(ctx, s1, rec, d_frame_off, nframes, size, max_q, dq, interp, self_pair_factor,
                    interp == IAC_NEAREST, d_hist, true);
}

int iac_ac1d_acuw_records(iac_ctx *ctx, const float *d_rec4, const uint64_t *d_frame_off, uint32_t nframes,
                          uint64_t n_total, uint32_t size, float max_q, float dq, int interp, int self_pair_factor,
                          double *d_hist)
{
    IAC_TRY(iac::bind(ctx));
    if (!d_frame_off || !d_hist) return iac::fail(IAC_ERR_INVALID_ARG, "null input");
    if (nframes == 0 || n_total == 0) return IAC_OK;
    if (!d_rec4) return iac::fail(IAC_ERR_PRECONDITION, "no hit-list records");
    IAC_TRY(acuw_check(size, interp, self_pair_factor));
    void *s1 = nullptr;
    IAC_TRY(iac::scratch(ctx, 1, acuw_work_bytes(nframes, size), &s1));
    return acuw_run(ctx, s1, reinterpret_cast<const float4 *>(d_rec4), d_frame_off, nframes, size, max_q, dq, interp,
                    self_pair_factor, false, d_hist, true);
}

int iac_ac1d_acuw_records_noalloc(iac_ctx *ctx, const float *d_rec4, const uint64_t *d_frame_off, uint32_t nframes,
                                  uint64_t n_total, uint32_t size, float max_q, float dq, int interp,
                                  int self_pair_factor, double *d_hist)
{
    IAC_TRY(iac::bind(ctx));
    if (!d_frame_off || !d_hist) return iac::fail(IAC_ERR_INVALID_ARG, "null input");
    if (nframes > kNoAllocMaxFrames)
        return iac::fail(IAC_ERR_LIMIT, "iac_ac1d_acuw_records_noalloc: at most 65536 frames per call");
    IAC_TRY(acuw_check(size, interp, self_pair_factor));
    if (nframes == 0 || n_total == 0) return IAC_OK;
    if (!d_rec4) return iac::fail(IAC_ERR_PRECONDITION, "no hit-list records");
    void *s1 = nullptr;
    IAC_TRY(noalloc_scratch(ctx->device, &s1));
    return acuw_run(ctx, s1, reinterpret_cast<const float4 *>(d_rec4), d_frame_off, nframes, size, max_q, dq, interp,
                    self_pair_factor, false, d_hist, false);
}

int iac_ac1d_calc_ac(iac_ctx *ctx, const double *d_ac_uw, const double *d_cq, uint32_t size, double *d_ac)
{
    IAC_TRY(iac::bind(ctx));
    if (!d_ac_uw || !d_cq || !d_ac) return iac::fail(IAC_ERR_INVALID_ARG, "null pointer");
    if (size == 0) return IAC_OK;
    hipLaunchKernelGGL(calc_ac_kernel, dim3(std::max(1u, (size + 255) / 256)), dim3(256), 0, ctx->stream, d_ac_uw,
                       d_cq, size, d_ac);
    IAC_HIP(hipGetLastError());
    return IAC_OK;
}

}  // extern "C"
