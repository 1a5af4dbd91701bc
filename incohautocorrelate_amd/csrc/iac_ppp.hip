// iac_ppp.hip -- pattern pre-processing in front of the path: the per-frame chain of
// PPP::ProcessData_PF_LAP (IAC/PPP.cpp:386-524) over a stack of dense frames on gfx950.
// Pixel mask, offset and gain correction, then the largest-adjacent-pixel (LAP) photon
// finder per detector panel, and each event's PhotonCount / MeanIntensity, with the
// reference's float arithmetic bit for bit.
//
// A panel is walked row by row: one wave owns a strip of 62 lanes (two columns per lane
// where the panels allow 8-byte accesses; its two outer lanes hold the neighbouring
// columns) over 64 rows and two frames, keeping the rows above and below in registers and
// taking the left and right neighbours from the lane's other column or the adjacent lanes,
// so each pixel of a frame is read about 1.06 times and written once.  The mask, offset and
// gain of a tile are read by its waves for 16 frame pairs in a row, from their XCD's L2.
#include <algorithm>
#include <cstring>
#include <type_traits>
#include <vector>

#include "iac_fpexact.h"
#include "iac_internal.h"
#include "iac_serial_sum.h"

#pragma clang fp contract(off)

namespace {

constexpr uint32_t kMaxPanels = 128;
constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr uint32_t kCols = kWave - 2;  // columns a wave writes (lanes 1..62)
constexpr uint32_t kRows = 64;         // rows a wave walks
constexpr uint32_t kF = 2;             // frames a wave carries
constexpr uint32_t kFrameGroup = 16;   // frames per thread of the elementwise kernel
static_assert(kRows <= 64, "per-lane whole-photon sums of a tile must fit an unsigned int");

// PPP::DetectorPanel (IAC/PPP.h:14-18), sorted by first pixel: a panel is `fs` pixels wide
// (the fast-scan direction) and `ss` rows tall; tile0[p] is the first wave tile of panel p
// (strips of kCols columns times segments of kRows rows), tile0[n] the tiles per frame group.
struct PanelTable {
    uint32_t n;
    uint32_t first[kMaxPanels];
    uint32_t fs[kMaxPanels];
    uint32_t ss[kMaxPanels];
    uint32_t tile0[kMaxPanels + 1];
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

// Per-event photon counts.  ArrayOperators::Sum<float> (IAC/ArrayOperators.h:260-268) is a
// serial float sum.  When every value of a frame is a whole number and the sum of their
// magnitudes is at most 2^24, every partial sum of that chain is an integer of at most 2^24
// and therefore exact, so the serial result is the exact integer sum, whatever the order.
// The kernels that write a frame add its values to an int64 sum and a magnitude sum, and
// flag values that are not whole or exceed 2^24; photon_count_kernel keeps those sums where
// they are exact and sums the remaining frames serially.
struct FrameAcc {
    long long sum;
    unsigned long long mag;
    unsigned int bad;
    unsigned int pad;
};

// adds one value to a lane's partial sums (at most 32 values of at most 2^24 per lane
// between flushes, so 32-bit partials cannot overflow)
__device__ __forceinline__ void tally(float v, int *s, unsigned int *m, bool *bad)
{
    const float a = fabsf(v);
    if (!(a <= 0x1p24f) || v != rintf(v)) {
        *bad = true;
    } else {
        *s += (int)v;
        *m += (unsigned int)a;
    }
}

// The finder's output inside a panel is floorf(I / ADU) or one more, a whole number; for
// values in [0, 2^24] its magnitude is itself, and anything else (negative with a negative
// ADU, above 2^24, NaN) is flagged.  2 * kRows values of at most 2^24 fit an unsigned int.
__device__ __forceinline__ void tally_whole(float v, unsigned int *s, bool *bad)
{
    if (v >= 0.0f && v <= 0x1p24f) *s += (unsigned int)v;
    else *bad = true;
}

// The finder's per-lane sums of frame f (at most kRows whole values each) into acc[f]; a
// lane above 2^24 already puts the frame past 2^24 (serial path), so it is flagged and
// dropped and the 64-lane sum stays below 2^31.  Every lane of the wave calls it.
__device__ __forceinline__ void flush_whole(FrameAcc *acc, uint32_t f, unsigned int su, bool bad)
{
    int s = (int)su;
    if (su > (1u << 24)) {
        bad = true;
        s = 0;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    const bool anybad = __any(bad);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        atomicAdd(reinterpret_cast<unsigned long long *>(&acc[f].sum), (unsigned long long)(long long)s);
        atomicAdd(&acc[f].mag, (unsigned long long)s);
        if (anybad) atomicOr(&acc[f].bad, 1u);
    }
}

// the wave's partial sums of frame f into acc[f]; every lane of the wave calls it
__device__ __forceinline__ void flush_tally(FrameAcc *acc, uint32_t f, int s32, unsigned int m32, bool bad)
{
    long long s = s32;
    unsigned long long m = m32;
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o);
        m += __shfl_xor(m, o);
    }
    const bool anybad = __any(bad);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        atomicAdd(reinterpret_cast<unsigned long long *>(&acc[f].sum), (unsigned long long)s);
        atomicAdd(&acc[f].mag, m);
        if (anybad) atomicOr(&acc[f].bad, 1u);
    }
}

// Pixels outside every panel (and every pixel in gain-only mode): the corrected value.
// grid: x over pixels [j0, npix), y over groups of kFrameGroup frames.
template <bool kMask, bool kGain>
__global__ void __launch_bounds__(kBlock)
    correct_kernel(const float *__restrict__ in, float *__restrict__ out, uint32_t nframes, uint32_t npix, uint32_t j0,
                   const int *__restrict__ mask, const float *__restrict__ offset, const float *__restrict__ gain,
                   int allow_neg, FrameAcc *__restrict__ acc)
{
    const uint32_t j = j0 + blockIdx.x * kBlock + threadIdx.x;
    const bool act = j < npix;  // every lane stays for the wave sums
    const Corr c = act ? load_corr<kMask, kGain>(mask, offset, gain, j) : Corr{1, 0.0f, 0.0f};
    for (uint32_t f = blockIdx.y * kFrameGroup; f < nframes; f += gridDim.y * kFrameGroup) {
        const uint32_t f1 = min(nframes, f + kFrameGroup);
        for (uint32_t g = f; g < f1; g++) {
            const size_t o = (size_t)g * npix + j;
            float v = 0.0f;
            if (act) {
                v = corrected<kMask, kGain>(in[o], c, allow_neg != 0);
                out[o] = v;
            }
            if (acc) {
                int s = 0;
                unsigned int m = 0;
                bool bad = false;
                tally(v, &s, &m, &bad);
                flush_tally(acc, g, s, m, bad);
            }
        }
    }
}

// kW consecutive values from p: one 8-byte access for kW = 2 (p 8-byte aligned)
template <int kW, typename T>
__device__ __forceinline__ void load_w(const T *p, T *v)
{
    if constexpr (kW == 2) {
        using V = typename std::conditional<std::is_same<T, int>::value, int2, float2>::type;
        const V x = *reinterpret_cast<const V *>(p);
        v[0] = x.x;
        v[1] = x.y;
    } else {
        v[0] = *p;
    }
}

template <bool kMask, bool kGain, int kW>
__device__ __forceinline__ void load_corr_w(const int *mask, const float *offset, const float *gain, size_t j,
                                            Corr *c)
{
    int m[kW];
    float o[kW], g[kW];
    for (int w = 0; w < kW; w++) {
        m[w] = 1;
        o[w] = g[w] = 0.0f;
    }
    if (kMask) load_w<kW>(mask + j, m);
    if (kGain) {
        load_w<kW>(offset + j, o);
        load_w<kW>(gain + j, g);
    }
    for (int w = 0; w < kW; w++) c[w] = Corr{m[w], o[w], g[w]};
}

// The LAP finder over the panels, one wave per tile (a strip of kCols * kW columns and kRows
// rows of one panel) and group of kF frames.  Lane l holds the kW columns from
// strip * kCols * kW + (l - 1) * kW (kW = 2 only where every panel starts and ends on even
// pixels, so the pairs are 8-byte aligned).  Grid: ntiles * ngrp blocks for ngrp frame groups
// (each block then steps by ngrp groups), a multiple of G.  Block order: chunks of G groups,
// in each chunk tile after tile, the G groups of a tile together.  Blocks go to the 8 XCDs
// round robin; the remap hands each XCD a contiguous range of that order, so a tile's G groups
// run on one XCD and read the tile's mask, offset and gain rows from its L2 (a frame group's
// whole detector at a time keeps the frame reads spread over the HBM channels; all groups of
// one tile at once would read one offset of hundreds of frames, 4 MB apart).
// IAC/PPP.cpp:53-87: where a pixel's fraction reaches the seed threshold, the largest
// fraction among its left / right / upper / lower neighbours inside the panel (the upper one
// only for i > fs, i = row * fs + column) is added and one more photon counted when the sum
// reaches the combined threshold.
template <bool kMask, bool kGain, int kW>
__global__ void __launch_bounds__(kWave)
    lap_rows_kernel(const float *__restrict__ in, float *__restrict__ out, uint32_t nframes, uint32_t npix,
                    const int *__restrict__ mask, const float *__restrict__ offset, const float *__restrict__ gain,
                    PanelTable T, float adu, float seed, float comb, FrameAcc *__restrict__ acc, uint32_t ngrp,
                    uint32_t G)
{
    const uint32_t nb = gridDim.x, b = blockIdx.x, xcd = b % 8, local = b / 8;
    const uint32_t q8 = nb / 8, r8 = nb % 8;
    const uint32_t lin = xcd < r8 ? xcd * (q8 + 1) + local : r8 * (q8 + 1) + (xcd - r8) * q8 + local;
    const uint32_t ntiles = T.tile0[T.n], chunk = lin / (ntiles * G), rem = lin % (ntiles * G);
    const uint32_t tile = rem / G, grp = chunk * G + rem % G;
    uint32_t lo = 0, hi = T.n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (T.tile0[mid] <= tile) lo = mid; else hi = mid;
    }
    const uint32_t fs = T.fs[lo], ss = T.ss[lo], base = T.first[lo];
    const uint32_t cols = kCols * kW;
    const uint32_t nstrip = (fs + cols - 1) / cols, t = tile - T.tile0[lo];
    const uint32_t strip = t % nstrip, seg = t / nstrip;
    const int lane = threadIdx.x;
    const int64_t c = (int64_t)strip * cols + (int64_t)(lane - 1) * kW;  // first column of the lane
    const bool valid = c >= 0 && c < (int64_t)fs;  // with kW = 2, fs is even: both columns
    const bool owner = valid && lane >= 1 && lane <= (int)kCols;
    const uint32_t cc = valid ? (uint32_t)c : 0u;
    const uint32_t r0 = seg * kRows, r1 = min(ss, r0 + kRows);

    for (uint32_t fg = grp * kF; fg < nframes; fg += ngrp * kF) {
        const uint32_t nf = min(kF, nframes - fg);
        // fractions (and whole photons) of row r for the kF frames; 0 off the panel
        auto load_row = [&](uint32_t r, float (*frag)[kW], float (*whole)[kW]) {
            const size_t j = (size_t)base + (size_t)r * fs + cc;
            Corr k[kW];
            for (int w = 0; w < kW; w++) k[w] = Corr{1, 0.0f, 0.0f};
            if (valid) load_corr_w<kMask, kGain, kW>(mask, offset, gain, j, k);
            float v[kF][kW];
#pragma unroll
            for (uint32_t q = 0; q < kF; q++) {
                for (int w = 0; w < kW; w++) v[q][w] = 0.0f;
                if (valid && q < nf) load_w<kW>(in + (size_t)(fg + q) * npix + j, v[q]);
            }
#pragma unroll
            for (uint32_t q = 0; q < kF; q++)
#pragma unroll
                for (int w = 0; w < kW; w++) {
                    frag[q][w] = 0.0f;
                    whole[q][w] = 0.0f;
                    if (valid && q < nf)
                        frag[q][w] = fraction(corrected<kMask, kGain>(v[q][w], k[w], false), adu, &whole[q][w]);
                }
        };
        float fu[kF][kW], fc[kF][kW], wc[kF][kW], fd[kF][kW], wd[kF][kW];
        unsigned int ts[kF];
        bool tb[kF];
#pragma unroll
        for (uint32_t q = 0; q < kF; q++) {
            ts[q] = 0;
            tb[q] = false;
        }
        if (r0 > 0) {
            load_row(r0 - 1, fu, wd);
        } else {
#pragma unroll
            for (uint32_t q = 0; q < kF; q++)
                for (int w = 0; w < kW; w++) fu[q][w] = 0.0f;
        }
        load_row(r0, fc, wc);
        for (uint32_t r = r0; r < r1; r++) {
            const bool hd = r + 1 < ss;
            if (hd) {
                load_row(r + 1, fd, wd);
            } else {
#pragma unroll
                for (uint32_t q = 0; q < kF; q++)
                    for (int w = 0; w < kW; w++) fd[q][w] = wd[q][w] = 0.0f;
            }
            const size_t orow = (size_t)base + (size_t)r * fs + cc;
#pragma unroll
            for (uint32_t q = 0; q < kF; q++) {
                // the column left of the lane's first and right of its last, from the next lanes
                const float lft = __shfl_up(fc[q][kW - 1], 1), rgt = __shfl_down(fc[q][0], 1);
                float res[kW];
#pragma unroll
                for (int w = 0; w < kW; w++) {
                    const int64_t cw = c + w;
                    const float left = w == 0 ? lft : fc[q][w - 1];
                    const float right = w == kW - 1 ? rgt : fc[q][w + 1];
                    const bool hl = cw > 0, hr = cw + 1 < (int64_t)fs, hu = r >= 2 || (r == 1 && cw >= 1);
                    res[w] = wc[q][w];
                    if (fc[q][w] >= seed) {
                        float best = 0.0f;
                        if (hl && left > best) best = left;
                        if (hr && right > best) best = right;
                        if (hu && fu[q][w] > best) best = fu[q][w];
                        if (hd && fd[q][w] > best) best = fd[q][w];
                        if (fc[q][w] + best >= comb) res[w] = wc[q][w] + 1.0f;
                    }
                }
                if (owner && q < nf) {
                    float *o = out + (size_t)(fg + q) * npix + orow;
                    if constexpr (kW == 2) *reinterpret_cast<float2 *>(o) = make_float2(res[0], res[1]);
                    else o[0] = res[0];
#pragma unroll
                    for (int w = 0; w < kW; w++) tally_whole(res[w], &ts[q], &tb[q]);
                }
            }
#pragma unroll
            for (uint32_t q = 0; q < kF; q++)
#pragma unroll
                for (int w = 0; w < kW; w++) {
                    fu[q][w] = fc[q][w];
                    fc[q][w] = fd[q][w];
                    wc[q][w] = wd[q][w];
                }
        }
        if (acc) {
#pragma unroll
            for (uint32_t q = 0; q < kF; q++)
                if (q < nf) flush_whole(acc, fg + q, ts[q], tb[q]);
        }
    }
}

__global__ void zero_acc_kernel(FrameAcc *acc, uint32_t n)
{
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < n) acc[f] = FrameAcc{0, 0, 0, 0};
}

// Pass 2: PhotonCount = (int) of the frame sum and MeanIntensity = (float)PhotonCount /
// (float)npix (IAC/PPP.cpp:490-491); the exact integer sum where it holds, else the serial float
// sum.  One wave per frame.
__global__ void __launch_bounds__(iac::kSumWave)
    photon_count_kernel(const float *__restrict__ frames, uint32_t npix, const FrameAcc *__restrict__ acc,
                        int *__restrict__ count, float *__restrict__ mean)
{
    const uint32_t f = blockIdx.x;
    const bool exact = !acc[f].bad && acc[f].mag <= (1ull << 24);
    const float s = exact ? (float)acc[f].sum
                          : iac::serial_sum_wave(frames + (size_t)f * npix, npix, [](uint32_t, float v) { return v; });
    if (threadIdx.x != 0) return;
    const int pc = iac::cvtt_i32(s);
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
    uint64_t full = 0;  // pixels [0, full) lie in the panels
    if (!gain_only) {
        // the reference sizes its per-pixel arrays by the panels' total pixel count
        // (IAC/PPP.cpp:420-426) and indexes them with each panel's pixels
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
            T.ss[k] = q.ss;
        }
        T.n = npanels;
    }
    if (nframes == 0) return IAC_OK;
    if ((uint64_t)nframes * npix > ((uint64_t)1 << 40) || nframes > 0x7fffffffu)
        return iac::fail(IAC_ERR_LIMIT, "frame stack too large");
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
    FrameAcc *acc = nullptr;
    if (d_photon_count || d_mean_intensity) {
        void *s = nullptr;
        IAC_TRY(iac::scratch(ctx, 1, (size_t)nframes * sizeof(FrameAcc), &s));
        acc = static_cast<FrameAcc *>(s);
        hipLaunchKernelGGL(zero_acc_kernel, dim3((nframes + 255) / 256), dim3(256), 0, ctx->stream, acc, nframes);
        IAC_HIP(hipGetLastError());
    }
    IAC_TRY(iac::timer_begin(ctx));
    if (T.n) {
        // two columns per lane (8-byte accesses) where every panel starts and ends on an even
        // pixel of an even-sized frame and every buffer is 8-byte aligned
        auto a8 = [](const void *p) { return reinterpret_cast<uintptr_t>(p) % 8 == 0; };
        bool pairs = npix % 2 == 0 && a8(src) && a8(d_out) && a8(d_mask) && a8(d_offset) && a8(d_gain);
        for (uint32_t k = 0; k < T.n; k++) pairs = pairs && T.first[k] % 2 == 0 && T.fs[k] % 2 == 0;
        const uint32_t cols = kCols * (pairs ? 2 : 1);
        uint64_t tiles = 0;
        for (uint32_t k = 0; k < T.n; k++) {
            T.tile0[k] = (uint32_t)tiles;
            tiles += (uint64_t)((T.fs[k] + cols - 1) / cols) * ((T.ss[k] + kRows - 1) / kRows);
            if (tiles > 0x7fffffffu) return iac::fail(IAC_ERR_LIMIT, "detector panels too large");
        }
        T.tile0[T.n] = (uint32_t)tiles;
        auto kern = pairs ? (hm ? (hg ? lap_rows_kernel<true, true, 2> : lap_rows_kernel<true, false, 2>)
                                : (hg ? lap_rows_kernel<false, true, 2> : lap_rows_kernel<false, false, 2>))
                          : (hm ? (hg ? lap_rows_kernel<true, true, 1> : lap_rows_kernel<true, false, 1>)
                                : (hg ? lap_rows_kernel<false, true, 1> : lap_rows_kernel<false, false, 1>));
        // every frame group its own block while the grid stays below 2^31 blocks; groups in
        // chunks of G (blocks of a padded chunk past the last frame do nothing)
        const uint32_t ntiles = T.tile0[T.n];
        const uint64_t nfg = (nframes + kF - 1) / kF;
        const uint32_t G = (uint32_t)std::min<uint64_t>(ctx->lap_groups ? ctx->lap_groups : 16, nfg);
        const uint64_t cap = 0x7fffffffu / ((uint64_t)ntiles * G);  // chunks per grid
        if (cap == 0) return iac::fail(IAC_ERR_LIMIT, "detector panels too large for IAC_OPT_LAP_GROUPS");
        const uint32_t ngrp = (uint32_t)(std::min<uint64_t>((nfg + G - 1) / G, cap) * G);
        hipLaunchKernelGGL(kern, dim3(ntiles * ngrp), dim3(kWave), 0, ctx->stream, src, d_out, nframes, npix, d_mask,
                           d_offset, d_gain, T, adu_per_photon, seed_threshold, combined_threshold, acc, ngrp, G);
        IAC_HIP(hipGetLastError());
    }
    if (full < npix) {
        auto kern = hm ? (hg ? correct_kernel<true, true> : correct_kernel<true, false>)
                       : (hg ? correct_kernel<false, true> : correct_kernel<false, false>);
        const uint32_t rest = npix - (uint32_t)full;
        const dim3 grid((rest + kBlock - 1) / kBlock, std::min<uint32_t>(65535, (nframes + kFrameGroup - 1) / kFrameGroup));
        hipLaunchKernelGGL(kern, grid, dim3(kBlock), 0, ctx->stream, src, d_out, nframes, npix, (uint32_t)full, d_mask,
                           d_offset, d_gain, gain_only ? 1 : 0, acc);
        IAC_HIP(hipGetLastError());
    }
    IAC_TRY(iac::timer_end(ctx));
    if (acc) {
        hipLaunchKernelGGL(photon_count_kernel, dim3(nframes), dim3(iac::kSumWave), 0, ctx->stream, d_out, npix, acc,
                           d_photon_count, d_mean_intensity);
        IAC_HIP(hipGetLastError());
    }
    return IAC_OK;
}

}  // extern "C"
