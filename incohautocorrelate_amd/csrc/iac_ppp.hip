// iac_ppp.hip -- pattern pre-processing in front of the path: the per-frame chain of
// PPP::ProcessData_PF_LAP (IAC/PPP.cpp:386-524) over a stack of dense frames on gfx950.
// Pixel mask, offset and gain correction, then the largest-adjacent-pixel (LAP) photon
// finder per detector panel, and each event's PhotonCount / MeanIntensity, with the
// reference's float arithmetic bit for bit.
//
// One thread owns one pixel across a group of frames: its mask, offset and gain (and
// those of its four panel neighbours) are read once per group, and each frame costs one
// 4-byte read and one 4-byte write of HBM; the neighbour reads of a frame hit the lines
// the adjacent threads and row blocks fetch at the same time.
#include <algorithm>
#include <cstring>
#include <vector>

#include "iac_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr uint32_t kMaxPanels = 128;
constexpr uint32_t kFrameGroup = 16;
constexpr int kBlock = 256;
constexpr int kBatch = 2;

// PPP::DetectorPanel (IAC/PPP.h:14-18), sorted by first pixel: a panel is `fs` pixels wide
// (the fast-scan direction) and size / fs rows tall.
struct PanelTable {
    uint32_t n;
    uint32_t first[kMaxPanels];
    uint32_t fs[kMaxPanels];
    uint32_t size[kMaxPanels];
};

struct Corr {
    int m;
    float o, g;
};

// Detector::ApplyPixelMask (IAC/Detector.cpp:662-665: Intensity *= (float)mask), then
// PPP::GainCorrection (IAC/PPP.cpp:781-803): Intensity + (-offset), divided by the gain
// where (value >= 0 or negatives allowed) and gain > 0 and mask == 1, else 0.
template <bool kMask, bool kGain>
__device__ __forceinline__ float corrected(float v, Corr c, bool allow_neg)
{
    if (kMask) v = v * (float)c.m;
    if (kGain) {
        v = v + (-c.o);
        v = ((v >= 0.0f || allow_neg) && c.g > 0.0f && c.m == 1) ? v / c.g : 0.0f;
    }
    return v;
}

// IAC/PPP.cpp:29-37: whole photons floorf(I / ADU) and the remaining fraction, both 0
// where I <= 0 (the reference leaves its zero-initialised arrays there).
__device__ __forceinline__ float fraction(float v, float adu, float *whole)
{
    if (v <= 0.0f) {
        *whole = 0.0f;
        return 0.0f;
    }
    const float r = v / adu;
    *whole = floorf(r);
    return r - *whole;
}

template <bool kMask, bool kGain>
__device__ __forceinline__ Corr load_corr(const int *mask, const float *offset, const float *gain, uint32_t j)
{
    Corr c{1, 0.0f, 0.0f};
    if (kMask) c.m = mask[j];
    if (kGain) {
        c.o = offset[j];
        c.g = gain[j];
    }
    return c;
}

// grid: x over pixels, y over groups of kFrameGroup frames
template <bool kMask, bool kGain>
__global__ void __launch_bounds__(kBlock)
    lap_kernel(const float *__restrict__ in, float *__restrict__ out, uint32_t nframes, uint32_t npix,
               const int *__restrict__ mask, const float *__restrict__ offset, const float *__restrict__ gain,
               int allow_neg, int lap, PanelTable T, float adu, float seed, float comb)
{
    const uint32_t ind = blockIdx.x * kBlock + threadIdx.x;
    if (ind >= npix) return;
    const uint32_t f0 = blockIdx.y * kFrameGroup, f1 = min(nframes, f0 + kFrameGroup);
    const bool an = allow_neg != 0;
    const Corr c = load_corr<kMask, kGain>(mask, offset, gain, ind);

    // the panel holding this pixel: the last one starting at or before it
    int p = -1;
    if (lap && T.n) {
        uint32_t lo = 0, hi = T.n;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (T.first[mid] <= ind) lo = mid; else hi = mid;
        }
        if (T.first[lo] <= ind && ind - T.first[lo] < T.size[lo]) p = (int)lo;
    }
    if (p < 0) {
        for (uint32_t f = f0; f < f1; f++) {
            const size_t o = (size_t)f * npix + ind;
            out[o] = corrected<kMask, kGain>(in[o], c, an);
        }
        return;
    }
    // neighbours inside the panel, IAC/PPP.cpp:57-76; the upper one only for i > fs
    const uint32_t i = ind - T.first[p], fs = T.fs[p];
    const bool hl = i % fs != 0, hr = (i + 1) % fs != 0, hu = i > fs, hd = (uint64_t)i + fs < T.size[p];
    Corr cl{}, cr{}, cu{}, cd{};
    if (hl) cl = load_corr<kMask, kGain>(mask, offset, gain, ind - 1);
    if (hr) cr = load_corr<kMask, kGain>(mask, offset, gain, ind + 1);
    if (hu) cu = load_corr<kMask, kGain>(mask, offset, gain, ind - fs);
    if (hd) cd = load_corr<kMask, kGain>(mask, offset, gain, ind + fs);
    // kBatch frames at a time: their pixel loads first, then the neighbour loads of the
    // frames whose fraction reaches the seed threshold, so each batch waits on memory twice
    for (uint32_t fb = f0; fb < f1; fb += kBatch) {
        float whole[kBatch], frag[kBatch], x[kBatch][4];
#pragma unroll
        for (int k = 0; k < kBatch; k++) {
            frag[k] = 0.0f;
            whole[k] = 0.0f;
            if (fb + k < f1) frag[k] = fraction(corrected<kMask, kGain>(in[(size_t)(fb + k) * npix + ind], c, an),
                                                adu, &whole[k]);
        }
#pragma unroll
        for (int k = 0; k < kBatch; k++) {
            x[k][0] = x[k][1] = x[k][2] = x[k][3] = 0.0f;
            if (fb + k < f1 && frag[k] >= seed) {
                const float *fr = in + (size_t)(fb + k) * npix;
                if (hl) x[k][0] = fr[ind - 1];
                if (hr) x[k][1] = fr[ind + 1];
                if (hu) x[k][2] = fr[ind - fs];
                if (hd) x[k][3] = fr[ind + fs];
            }
        }
#pragma unroll
        for (int k = 0; k < kBatch; k++) {
            if (fb + k >= f1) break;
            float r = whole[k];
            if (frag[k] >= seed) {
                // IAC/PPP.cpp:53-87: the largest neighbouring fraction, then one more photon
                // when fraction + LAP reaches the combined threshold
                float best = 0.0f, w, y;
                if (hl) { y = fraction(corrected<kMask, kGain>(x[k][0], cl, an), adu, &w); if (y > best) best = y; }
                if (hr) { y = fraction(corrected<kMask, kGain>(x[k][1], cr, an), adu, &w); if (y > best) best = y; }
                if (hu) { y = fraction(corrected<kMask, kGain>(x[k][2], cu, an), adu, &w); if (y > best) best = y; }
                if (hd) { y = fraction(corrected<kMask, kGain>(x[k][3], cd, an), adu, &w); if (y > best) best = y; }
                if (frag[k] + best >= comb) r = whole[k] + 1.0f;
            }
            out[(size_t)(fb + k) * npix + ind] = r;
        }
    }
}

// Per-event photon counts.  ArrayOperators::Sum<float> (IAC/ArrayOperators.h:260-268) is a
// serial float sum.  When every value of a frame is a whole number and the sum of their
// magnitudes is at most 2^24, every partial sum of that chain is an integer of at most 2^24
// and therefore exact, so the serial result is the exact integer sum, whatever the order:
// pass 1 sums each frame in parallel (int64 sum, magnitude sum, a flag for values that are
// not whole or exceed 2^24); pass 2 keeps those sums where they are exact and sums the
// remaining frames serially.
struct FrameAcc {
    long long sum;
    unsigned long long mag;
    unsigned int bad;
    unsigned int pad;
};

__global__ void __launch_bounds__(kBlock)
    frame_int_sums_kernel(const float *__restrict__ frames, uint32_t nframes, uint32_t npix,
                          FrameAcc *__restrict__ acc)
{
    __shared__ long long ss[kBlock / 64];
    __shared__ unsigned long long sm[kBlock / 64];
    __shared__ unsigned int sb[kBlock / 64];
    const int w = threadIdx.x / 64;
    for (uint32_t f = blockIdx.y; f < nframes; f += gridDim.y) {
        const float *fr = frames + (size_t)f * npix;
        long long s = 0;
        unsigned long long m = 0;
        unsigned int bad = 0;
        for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < npix; j += gridDim.x * kBlock) {
            const float v = fr[j];
            const float a = fabsf(v);
            if (!(a <= 0x1p24f) || v != rintf(v)) {
                bad = 1;
            } else {
                s += (long long)v;
                m += (unsigned long long)a;
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            s += __shfl_xor(s, o);
            m += __shfl_xor(m, o);
            bad |= __shfl_xor(bad, o);
        }
        if ((threadIdx.x & 63) == 0) {
            ss[w] = s;
            sm[w] = m;
            sb[w] = bad;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int k = 1; k < kBlock / 64; k++) {
                s += ss[k];
                m += sm[k];
                bad |= sb[k];
            }
            atomicAdd(reinterpret_cast<unsigned long long *>(&acc[f].sum), (unsigned long long)s);
            atomicAdd(&acc[f].mag, m);
            if (bad) atomicOr(&acc[f].bad, 1u);
        }
        __syncthreads();  // the per-wave slots are reused by the next frame
    }
}

__global__ void zero_acc_kernel(FrameAcc *acc, uint32_t n)
{
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < n) acc[f] = FrameAcc{0, 0, 0, 0};
}

// Pass 2, one thread per frame: PhotonCount = (int) of the frame sum and MeanIntensity =
// (float)PhotonCount / (float)npix (IAC/PPP.cpp:490-491).
template <bool kVec4>
__global__ void __launch_bounds__(64)
    photon_count_kernel(const float *__restrict__ frames, uint32_t nframes, uint32_t npix,
                        const FrameAcc *__restrict__ acc, int *__restrict__ count, float *__restrict__ mean)
{
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nframes) return;
    const float *fr = frames + (size_t)f * npix;
    float s = 0.0f;
    if (!acc[f].bad && acc[f].mag <= (1ull << 24)) {
        s = (float)acc[f].sum;
    } else if (kVec4) {
        const float4 *v = reinterpret_cast<const float4 *>(fr);
        for (uint32_t j = 0; j < npix / 4; j++) {
            const float4 x = v[j];
            s += x.x;
            s += x.y;
            s += x.z;
            s += x.w;
        }
    } else {
        for (uint32_t j = 0; j < npix; j++) s += fr[j];
    }
    const int pc = (int)s;
    if (count) count[f] = pc;
    if (mean) mean[f] = (float)pc / (float)npix;
}

}  // namespace

extern "C" {

int iac_ppp_lap(iac_ctx *ctx, const float *d_in, float *d_out, uint32_t nframes, uint32_t npix, const int *d_mask,
                const float *d_offset, const float *d_gain, const iac_panel *panels, uint32_t npanels,
                float adu_per_photon, float seed_threshold, float combined_threshold, int gain_only,
                int *d_photon_count, float *d_mean_intensity)
{
    IAC_TRY(iac::bind(ctx));
    if (npix == 0 || (nframes && (!d_in || !d_out))) return iac::fail(IAC_ERR_INVALID_ARG, "bad LAP request");
    if ((d_offset == nullptr) != (d_gain == nullptr))
        return iac::fail(IAC_ERR_INVALID_ARG, "gain correction needs both the offset and the gain map");
    if (npanels && !panels) return iac::fail(IAC_ERR_INVALID_ARG, "null panel list");
    if (npanels > kMaxPanels) return iac::fail(IAC_ERR_LIMIT, "at most 128 detector panels");
    PanelTable T;
    std::memset(&T, 0, sizeof(T));
    if (!gain_only) {
        // the reference sizes its per-pixel arrays by the panels' total pixel count
        // (IAC/PPP.cpp:420-426) and indexes them with each panel's pixels
        uint64_t full = 0;
        std::vector<iac_panel> sorted(panels, panels + npanels);
        for (const iac_panel &q : sorted) {
            if (q.fs == 0 || q.ss == 0) return iac::fail(IAC_ERR_INVALID_ARG, "empty detector panel");
            full += (uint64_t)q.fs * q.ss;
        }
        if (full > npix) return iac::fail(IAC_ERR_SHAPE, "detector panels hold more pixels than a frame");
        std::sort(sorted.begin(), sorted.end(),
                  [](const iac_panel &a, const iac_panel &b) { return a.first_ind < b.first_ind; });
        uint64_t end_prev = 0;
        for (uint32_t k = 0; k < npanels; k++) {
            const iac_panel &q = sorted[k];
            const uint64_t end = (uint64_t)q.first_ind + (uint64_t)q.fs * q.ss;
            if (end > full) return iac::fail(IAC_ERR_SHAPE, "a detector panel reaches past the panels' pixel count");
            if (k && q.first_ind < end_prev) return iac::fail(IAC_ERR_INVALID_ARG, "detector panels overlap");
            end_prev = end;
            T.first[k] = q.first_ind;
            T.fs[k] = q.fs;
            T.size[k] = (uint32_t)((uint64_t)q.fs * q.ss);
        }
        T.n = npanels;
    }
    if (nframes == 0) return IAC_OK;
    if ((uint64_t)nframes * npix > ((uint64_t)1 << 40)) return iac::fail(IAC_ERR_LIMIT, "frame stack too large");
    const size_t bytes = (size_t)nframes * npix * sizeof(float);
    const float *src = d_in;
    if (d_in == d_out) {
        // in place, like the reference: stage the input so neighbour reads see the
        // corrected-but-unfound values of every pixel
        void *s = nullptr;
        IAC_TRY(iac::scratch(ctx, 0, bytes, &s));
        IAC_HIP(hipMemcpyAsync(s, d_in, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        src = static_cast<const float *>(s);
    }
    const bool hm = d_mask != nullptr, hg = d_gain != nullptr;
    auto kern = hm ? (hg ? lap_kernel<true, true> : lap_kernel<true, false>)
                   : (hg ? lap_kernel<false, true> : lap_kernel<false, false>);
    const dim3 grid((npix + kBlock - 1) / kBlock, (nframes + kFrameGroup - 1) / kFrameGroup);
    if (grid.y > 65535) return iac::fail(IAC_ERR_LIMIT, "at most 1,048,560 frames per call");
    IAC_TRY(iac::timer_begin(ctx));
    hipLaunchKernelGGL(kern, grid, dim3(kBlock), 0, ctx->stream, src, d_out, nframes, npix, d_mask, d_offset, d_gain,
                       gain_only ? 1 : 0, gain_only ? 0 : 1, T, adu_per_photon, seed_threshold, combined_threshold);
    IAC_HIP(hipGetLastError());
    IAC_TRY(iac::timer_end(ctx));
    if (d_photon_count || d_mean_intensity) {
        void *s = nullptr;
        IAC_TRY(iac::scratch(ctx, 1, (size_t)nframes * sizeof(FrameAcc), &s));
        auto *acc = static_cast<FrameAcc *>(s);
        hipLaunchKernelGGL(zero_acc_kernel, dim3((nframes + 255) / 256), dim3(256), 0, ctx->stream, acc, nframes);
        IAC_HIP(hipGetLastError());
        // about 16 values per thread
        const uint32_t bx = std::max<uint32_t>(1, std::min<uint32_t>(1024, (npix + kBlock * 16 - 1) / (kBlock * 16)));
        hipLaunchKernelGGL(frame_int_sums_kernel, dim3(bx, std::min<uint32_t>(nframes, 65535)), dim3(kBlock), 0,
                           ctx->stream, d_out, nframes, npix, acc);
        IAC_HIP(hipGetLastError());
        const bool v4 = npix % 4 == 0 && reinterpret_cast<uintptr_t>(d_out) % 16 == 0;
        hipLaunchKernelGGL(v4 ? photon_count_kernel<true> : photon_count_kernel<false>, dim3((nframes + 63) / 64),
                           dim3(64), 0, ctx->stream, d_out, nframes, npix, acc, d_photon_count, d_mean_intensity);
        IAC_HIP(hipGetLastError());
    }
    return IAC_OK;
}

}  // extern "C"
