// iac_sim.hip -- the emitter speckle simulator (SimulateCrystal, IAC/kernel.cl:768-859,
// driven by Simulator::SimulatePart, IAC/Simulator.cpp:662-944) and its
// post-processing (rescale + Poisson, IAC/Simulator.cpp:870-888) on gfx950.
//
// Two ways to the field psi(p, f) = sum_e exp(i (k_p . r_e + phi_{e,f})):
//   direct FP32: the reference kernel's arithmetic, one thread per (pixel, frame),
//     emitters staged through LDS.  The plain baseline.
//   complex GEMM: psi = A B with A[p,e] = exp(i k_p . r_e) (phase argument and its
//     reduction in FP64, once per geometry: the plan) and B[e,f] = exp(i phi_{e,f}).
//     With C + iS = A and c + is = B:  Re psi = C c - S s,  Im psi = C s + S c.
//     Every operand is stored split, x = hi + lo in bf16 or fp16, and each product
//     is hi*hi + hi*lo + lo*hi on v_mfma_f32_32x32x16_{bf16,f16} with FP32
//     accumulation, which keeps ~16 bits per element (the dropped lo*lo is ~2^-16 of
//     a term).  One kernel does all of it and forms |psi| (reference, kernel.cl:844)
//     or |psi|^2 from the Re / Im accumulators in registers: the FP32 field never
//     reaches memory.
//
// MFMA operand layout (gfx950, 32x32x16, 64 lanes): lane l supplies row l % 32 of A
// and column l % 32 of B, each with k = 8 (l / 32) .. 8 (l / 32) + 7 as 8 16-bit
// values (16 bytes); accumulator register r of lane l is D[8 (r / 4) + 4 (l / 32) +
// r % 4][l % 32].  The plan and the phase operand are written in exactly that order,
// one 1 KB "fragment" per (32 rows or columns) x (16 emitters) x (cos or sin) x (hi or
// lo), so every operand load is one contiguous 16-byte-per-lane global load.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "iac_internal.h"
#include "iac_rng.h"

#pragma clang fp contract(off)

namespace {

constexpr double kPi = 3.14159265358979323846;

// GEMM tiling: a block of 4 waves owns 256 rows x 128 columns (frame x mode); wave w
// owns rows 128 (w & 1) .. +127 and columns 64 (w >> 1) .. +63, i.e. 4 x 2 tiles of
// 32 x 32, each with a Re and an Im accumulator (256 accumulator registers per lane).
constexpr uint32_t kRowsPerBlock = 256, kColsPerBlock = 128;
constexpr uint32_t kWaveRowTiles = 4, kWaveColTiles = 2;
constexpr uint32_t kEmittersPerStep = 16;  // the MFMA's k
constexpr uint32_t kFragU4 = 64;            // uint4 per fragment (1 KB)

}  // namespace

struct iac_sim_plan {
    iac_ctx *ctx = nullptr;
    int mode = 0;
    uint32_t npix = 0, nsub = 0, rows = 0, nem = 0;
    uint32_t rows_pad = 0;  // multiple of kRowsPerBlock
    uint32_t nkc = 0;       // emitter steps of 16, an even count (emitters zero-padded)
    void *d_x = nullptr;    // fragments [rows_pad / 32][nkc][cos, sin][hi (, lo)]
};

namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

// per row (pixel p, subpixel s): k = 2 pi / lambda * pos / |pos| in double,
// pos = p + v * v_step + u * u_step (Simulator.cpp:742-760, kernel.cl:823-832)
__global__ void k_rows_kernel(const float *__restrict__ pixmap, uint32_t npix, int susa, double3 us, double3 vs,
                              double kk, double *__restrict__ krow)
{
    const int side = 2 * susa + 1;
    const uint32_t nsub = (uint32_t)(side * side);
    const uint64_t rows = (uint64_t)npix * nsub;
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < rows; r += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t p = (uint32_t)(r / nsub), s = (uint32_t)(r % nsub);
        const int v = (int)(s / side) - susa, u = (int)(s % side) - susa;
        const double x = (double)pixmap[3 * (size_t)p] + v * vs.x + u * us.x;
        const double y = (double)pixmap[3 * (size_t)p + 1] + v * vs.y + u * us.y;
        const double z = (double)pixmap[3 * (size_t)p + 2] + v * vs.z + u * us.z;
        const double nrm = kk / sqrt(x * x + y * y + z * z);
        krow[3 * r] = x * nrm;
        krow[3 * r + 1] = y * nrm;
        krow[3 * r + 2] = z * nrm;
    }
}

template <typename T>
__device__ __forceinline__ uint16_t bits16(T v)
{
    return __builtin_bit_cast(uint16_t, v);
}

// x = hi + lo, hi = T(x), lo = T(x - hi) (through float, as the values are |x| <= 1)
template <typename T>
__device__ __forceinline__ void split16(double x, uint16_t &hi, uint16_t &lo)
{
    const T h = (T)(float)x;
    hi = bits16(h);
    lo = bits16((T)(float)(x - (double)(float)h));
}

__device__ __forceinline__ uint4 pack8(const uint16_t *v)
{
    return make_uint4(v[0] | ((uint32_t)v[1] << 16), v[2] | ((uint32_t)v[3] << 16), v[4] | ((uint32_t)v[5] << 16),
                      v[6] | ((uint32_t)v[7] << 16));
}

// One thread per (row tile t, emitter step kc, lane): the 8 entries of that lane's slot
// in the four fragments (cos hi, cos lo, sin hi, sin lo; hi only when nh == 1) of
// X[row][e] = exp(i k_row . r_e), row = 32 t + lane % 32, e = 16 kc + 8 (lane / 32) + j.
// Padded rows and emitters are zero (they add nothing to any sum).
template <typename T>
__global__ void phase_matrix_kernel(const double *__restrict__ krow, uint32_t rows, uint32_t rows_pad,
                                    const double *__restrict__ em, uint32_t nem, uint32_t nkc, int nh,
                                    uint4 *__restrict__ x)
{
    const uint64_t slots = (uint64_t)(rows_pad / 32) * nkc * 64;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < slots;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t lane = (uint32_t)(i % 64);
        const uint64_t tk = i / 64;  // t * nkc + kc
        const uint32_t kc = (uint32_t)(tk % nkc), t = (uint32_t)(tk / nkc);
        const uint32_t row = 32 * t + lane % 32;
        uint16_t ch[8], cl[8], sh[8], sl[8];
        double kx = 0, ky = 0, kz = 0;
        if (row < rows) {
            kx = krow[3 * (size_t)row];
            ky = krow[3 * (size_t)row + 1];
            kz = krow[3 * (size_t)row + 2];
        }
        for (int j = 0; j < 8; j++) {
            const uint32_t e = kEmittersPerStep * kc + 8 * (lane / 32) + j;
            double sn = 0.0, cs = 0.0;
            if (row < rows && e < nem) {
                const double arg = kx * em[3 * (size_t)e] + ky * em[3 * (size_t)e + 1] + kz * em[3 * (size_t)e + 2];
                sincos(arg, &sn, &cs);
            }
            split16<T>(cs, ch[j], cl[j]);
            split16<T>(sn, sh[j], sl[j]);
        }
        uint4 *f = x + tk * (2 * nh) * kFragU4 + lane;
        f[0] = pack8(ch);
        if (nh == 2) {
            f[kFragU4] = pack8(cl);
            f[2 * kFragU4] = pack8(sh);
            f[3 * kFragU4] = pack8(sl);
        } else {
            f[kFragU4] = pack8(sh);
        }
    }
}

// The phase operand, one thread per (column tile u, emitter step kc, lane): fragments
// (cos hi, cos lo, sin hi, sin lo) of B[e][col] = exp(i phi[col][e]) for col = 32 u +
// lane % 32 (col = frame * modes + mode), e = 16 kc + 8 (lane / 32) + j.
template <typename T>
__global__ void phase_vector_kernel(const double *__restrict__ phases, uint32_t ncols, uint32_t ncols_pad,
                                    uint32_t nem, uint32_t nkc, int nh, uint4 *__restrict__ y)
{
    const uint64_t slots = (uint64_t)(ncols_pad / 32) * nkc * 64;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < slots;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t lane = (uint32_t)(i % 64);
        const uint64_t uk = i / 64;
        const uint32_t kc = (uint32_t)(uk % nkc), u = (uint32_t)(uk / nkc);
        const uint32_t col = 32 * u + lane % 32;
        uint16_t ch[8], cl[8], sh[8], sl[8];
        for (int j = 0; j < 8; j++) {
            const uint32_t e = kEmittersPerStep * kc + 8 * (lane / 32) + j;
            double sn = 0.0, cs = 0.0;
            if (col < ncols && e < nem) sincos(phases[(size_t)col * nem + e], &sn, &cs);
            split16<T>(cs, ch[j], cl[j]);
            split16<T>(sn, sh[j], sl[j]);
        }
        uint4 *f = y + uk * (2 * nh) * kFragU4 + lane;
        f[0] = pack8(ch);
        if (nh == 2) {
            f[kFragU4] = pack8(cl);
            f[2 * kFragU4] = pack8(sh);
            f[3 * kFragU4] = pack8(sl);
        } else {
            f[kFragU4] = pack8(sh);
        }
    }
}

template <bool kBf16>
__device__ __forceinline__ f32x16_t mfma(const uint4 &a, const uint4 &b, f32x16_t c)
{
    if constexpr (kBf16)
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, a),
                                                      __builtin_bit_cast(bf16x8_t, b), c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8_t, a), __builtin_bit_cast(f16x8_t, b),
                                                     c, 0, 0, 0);
}

__device__ __forceinline__ uint4 negate16(const uint4 &v)
{
    constexpr uint32_t s = 0x80008000u;  // sign bit of both 16-bit halves (bf16 and fp16 alike)
    return make_uint4(v.x ^ s, v.y ^ s, v.z ^ s, v.w ^ s);
}

// One emitter step's operands of a wave: A fragments per row tile [cos, sin][hi, lo],
// B fragments per column tile [cos, sin][hi, lo].
template <int NH>
struct StepFrags {
    uint4 a[kWaveRowTiles][2][NH];
    uint4 b[kWaveColTiles][2][NH];
};

template <int NH>
__device__ __forceinline__ void load_step(StepFrags<NH> &s, const uint4 *__restrict__ xa, const uint4 *__restrict__ yb,
                                          size_t row_tile_stride, size_t col_tile_stride, uint32_t kc)
{
#pragma unroll
    for (uint32_t i = 0; i < kWaveRowTiles; i++) {
        const uint4 *p = xa + i * row_tile_stride + (size_t)kc * (2 * NH * kFragU4);
#pragma unroll
        for (int q = 0; q < 2; q++)
#pragma unroll
            for (int h = 0; h < NH; h++) s.a[i][q][h] = p[(q * NH + h) * kFragU4];
    }
#pragma unroll
    for (uint32_t j = 0; j < kWaveColTiles; j++) {
        const uint4 *p = yb + j * col_tile_stride + (size_t)kc * (2 * NH * kFragU4);
#pragma unroll
        for (int q = 0; q < 2; q++)
#pragma unroll
            for (int h = 0; h < NH; h++) s.b[j][q][h] = p[(q * NH + h) * kFragU4];
    }
}

// Re += C c - S s, Im += C s + S c for every (row tile, column tile) of the wave; with
// NH == 2 each product is hi hi + hi lo + lo hi.  Terms outermost, tiles innermost:
// consecutive MFMAs write different accumulators.
template <bool kBf16, int NH>
__device__ __forceinline__ void step_mfma(const StepFrags<NH> &s, f32x16_t (&re)[kWaveRowTiles][kWaveColTiles],
                                          f32x16_t (&im)[kWaveRowTiles][kWaveColTiles])
{
    constexpr int nterms = NH == 2 ? 3 : 1;
    constexpr int ah[3] = {0, 0, 1}, bh[3] = {0, 1, 0};
    uint4 ns[kWaveColTiles][NH];
#pragma unroll
    for (uint32_t j = 0; j < kWaveColTiles; j++)
#pragma unroll
        for (int h = 0; h < NH; h++) ns[j][h] = negate16(s.b[j][1][h]);
#pragma unroll
    for (int t = 0; t < nterms; t++) {
#pragma unroll
        for (uint32_t i = 0; i < kWaveRowTiles; i++)
#pragma unroll
            for (uint32_t j = 0; j < kWaveColTiles; j++) re[i][j] = mfma<kBf16>(s.a[i][0][ah[t]], s.b[j][0][bh[t]], re[i][j]);
#pragma unroll
        for (uint32_t i = 0; i < kWaveRowTiles; i++)
#pragma unroll
            for (uint32_t j = 0; j < kWaveColTiles; j++) im[i][j] = mfma<kBf16>(s.a[i][0][ah[t]], s.b[j][1][bh[t]], im[i][j]);
#pragma unroll
        for (uint32_t i = 0; i < kWaveRowTiles; i++)
#pragma unroll
            for (uint32_t j = 0; j < kWaveColTiles; j++) re[i][j] = mfma<kBf16>(s.a[i][1][ah[t]], ns[j][bh[t]], re[i][j]);
#pragma unroll
        for (uint32_t i = 0; i < kWaveRowTiles; i++)
#pragma unroll
            for (uint32_t j = 0; j < kWaveColTiles; j++) im[i][j] = mfma<kBf16>(s.a[i][1][ah[t]], s.b[j][0][bh[t]], im[i][j]);
    }
}

// Block index -> (row block, column block).  Blocks go to the 8 XCDs round robin; the
// remap hands each XCD a contiguous range of the (row block major) tile order, so the
// column blocks of one row block run together on one XCD and read its rows of the plan
// from that XCD's L2 once.
__device__ __forceinline__ void block_tile(uint32_t nblocks, uint32_t ncb, uint32_t &rb, uint32_t &cb)
{
    const uint32_t b = blockIdx.x, xcd = b % 8, local = b / 8;
    const uint32_t q = nblocks / 8, r = nblocks % 8;
    const uint32_t lin = xcd < r ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
    rb = lin / ncb;
    cb = lin % ncb;
}

// The fused simulator GEMM.  Grid: nrb x ncb blocks of 256 threads (one wave per SIMD).
// kDirect (one subpixel, one mode): out[f][p] = |psi| or |psi|^2 straight from the
// accumulators (float4 stores when vec4: npix % 4 == 0 and a 16-byte aligned out);
// otherwise the magnitude of every (row, column) goes to
// w[col][row - row0] for the subpixel / mode reduction (w_rows rows per column).
template <bool kBf16, int NH, bool kDirect>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
    sim_gemm_kernel(const uint4 *__restrict__ x, const uint4 *__restrict__ y, uint32_t nkc, uint32_t rb0,
                    uint32_t nrb, uint32_t ncb, uint32_t rows, uint32_t ncols, int kind, float *__restrict__ out,
                    uint32_t npix, int vec4, float *__restrict__ w, uint32_t w_rows)
{
    uint32_t rb, cb;
    block_tile(nrb * ncb, ncb, rb, cb);
    rb += rb0;
    const uint32_t wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const uint32_t t0 = rb * (kRowsPerBlock / 32) + (wave & 1) * kWaveRowTiles;
    const uint32_t u0 = cb * (kColsPerBlock / 32) + (wave >> 1) * kWaveColTiles;
    const size_t tile_stride = (size_t)nkc * 2 * NH * kFragU4;
    const uint4 *xa = x + (size_t)t0 * tile_stride + lane;
    const uint4 *yb = y + (size_t)u0 * tile_stride + lane;

    f32x16_t re[kWaveRowTiles][kWaveColTiles], im[kWaveRowTiles][kWaveColTiles];
#pragma unroll
    for (uint32_t i = 0; i < kWaveRowTiles; i++)
#pragma unroll
        for (uint32_t j = 0; j < kWaveColTiles; j++) {
            re[i][j] = f32x16_t{};
            im[i][j] = f32x16_t{};
        }

    // two steps in flight: the next step's fragments load while this one's MFMAs run.
    // nkc is even; the last iteration reloads its own step instead of reading past the
    // end, so the loop body has no branch (no copies of the accumulators).
    StepFrags<NH> s0, s1;
    load_step<NH>(s0, xa, yb, tile_stride, tile_stride, 0);
    // The scheduling barriers keep each step's loads ahead of the previous step's MFMAs
    // (left alone, the scheduler sinks the loads to their first use).
    for (uint32_t kc = 0; kc < nkc; kc += 2) {
        load_step<NH>(s1, xa, yb, tile_stride, tile_stride, kc + 1);
        __builtin_amdgcn_sched_barrier(0);
        step_mfma<kBf16, NH>(s0, re, im);
        __builtin_amdgcn_sched_barrier(0);
        load_step<NH>(s0, xa, yb, tile_stride, tile_stride, kc + 2 < nkc ? kc + 2 : kc);
        __builtin_amdgcn_sched_barrier(0);
        step_mfma<kBf16, NH>(s1, re, im);
        __builtin_amdgcn_sched_barrier(0);
    }

    // epilogue: lane holds column 32 u + lane % 32 and rows 32 t + 8 g + 4 (lane / 32) + r
#pragma unroll
    for (uint32_t i = 0; i < kWaveRowTiles; i++) {
#pragma unroll
        for (uint32_t j = 0; j < kWaveColTiles; j++) {
            const uint32_t col = 32 * (u0 + j) + lane % 32;
            if (col >= ncols) continue;
#pragma unroll
            for (uint32_t g = 0; g < 4; g++) {
                const uint32_t row = 32 * (t0 + i) + 8 * g + 4 * (lane / 32);
                float v[4];
#pragma unroll
                for (uint32_t r = 0; r < 4; r++) {
                    const float a = re[i][j][4 * g + r], b = im[i][j][4 * g + r];
                    const float a2 = a * a + b * b;
                    v[r] = kind == 0 ? sqrtf(a2) : a2;
                }
                if (kDirect) {
                    float *o = out + (size_t)col * npix + row;
                    if (vec4 && row + 3 < rows) {
                        *reinterpret_cast<float4 *>(o) = make_float4(v[0], v[1], v[2], v[3]);
                    } else {
#pragma unroll
                        for (uint32_t r = 0; r < 4; r++)
                            if (row + r < rows) o[r] = v[r];
                    }
                } else {
                    float *o = w + (size_t)col * w_rows + (row - rb0 * kRowsPerBlock);
                    *reinterpret_cast<float4 *>(o) = make_float4(v[0], v[1], v[2], v[3]);
                }
            }
        }
    }
}

// out[f][p] = sum over the modes m of (sum over the subpixels s of w[f modes + m][row]) /
// nsub, rows p nsub + s relative to the chunk's first pixel p0; sums in order.
__global__ void subpixel_mode_reduce_kernel(const float *__restrict__ w, uint32_t w_rows, uint32_t p0, uint32_t np,
                                            uint32_t nsub, uint32_t nframes, uint32_t modes,
                                            float *__restrict__ out, uint32_t npix)
{
    const uint32_t pl = blockIdx.x * blockDim.x + threadIdx.x;
    if (pl >= np) return;
    for (uint32_t f = blockIdx.y; f < nframes; f += gridDim.y) {
        float acc = 0.0f;
        for (uint32_t m = 0; m < modes; m++) {
            const float *wc = w + (size_t)(f * modes + m) * w_rows + (size_t)pl * nsub;
            float t = 0.0f;
            for (uint32_t s = 0; s < nsub; s++) t += wc[s];
            acc += t / (float)nsub;
        }
        out[(size_t)f * npix + p0 + pl] = acc;
    }
}

// SimulateCrystal restated in FP32 (kernel.cl:806-856), one thread per
// (pixel, frame); emitters and the mode's phases staged through LDS; the
// modes of a frame add up as SimulatePart's ParAdd does (Simulator.cpp:862).
__global__ void __launch_bounds__(256)
    direct_fp32_kernel(const float *__restrict__ pixmap, uint32_t npix, int susa, float3 us, float3 vs,
                       float wavelength, const float *__restrict__ em, uint32_t nem, const float *__restrict__ phases,
                       uint32_t modes, int kind, float *__restrict__ out)
{
    __shared__ float4 se[512];
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x, f = blockIdx.y;
    const bool live = p < npix;
    float px = 0, py = 0, pz = 0;
    if (live) {
        px = pixmap[3 * (size_t)p];
        py = pixmap[3 * (size_t)p + 1];
        pz = pixmap[3 * (size_t)p + 2];
    }
    float result = 0.0f;
    for (uint32_t m = 0; m < modes; m++) {
        const float *ph = phases + ((size_t)f * modes + m) * nem;
        float total = 0.0f;
        for (int v = -susa; v <= susa; v++) {
            for (int u = -susa; u <= susa; u++) {
                const float cx = px + (v * vs.x) + (u * us.x);
                const float cy = py + (v * vs.y) + (u * us.y);
                const float cz = pz + (v * vs.z) + (u * us.z);
                // 1.0 / sqrt(...) and 2.0 * PI / lambda are double expressions in OpenCL C
                float nk = (float)(1.0 / (double)sqrtf(cx * cx + cy * cy + cz * cz));
                nk = (float)((double)nk * ((2.0 * kPi) / (double)wavelength));
                const float kx = cx * nk, ky = cy * nk, kz = cz * nk;
                float re = 0.0f, im = 0.0f;
                for (uint32_t e0 = 0; e0 < nem; e0 += 512) {
                    __syncthreads();
                    for (uint32_t i = threadIdx.x; i < 512 && e0 + i < nem; i += blockDim.x) {
                        const size_t e = e0 + i;
                        se[i] = make_float4(em[3 * e], em[3 * e + 1], em[3 * e + 2], ph[e]);
                    }
                    __syncthreads();
                    const uint32_t ne = min(512u, nem - e0);
                    for (uint32_t i = 0; i < ne; i++) {
                        const float4 q = se[i];
                        float arg = 0.0f;
                        arg += kx * q.x;
                        arg += ky * q.y;
                        arg += kz * q.z;
                        arg += q.w;
                        float sn, cs;
                        sincosf(arg, &sn, &cs);
                        re += cs;
                        im += sn;
                    }
                }
                const float a2 = re * re + im * im;
                total += kind == 0 ? sqrtf(a2) : a2;
            }
        }
        result += total / ((float)((2 * susa + 1) * (2 * susa + 1)));
    }
    if (live) out[(size_t)f * npix + p] = result;
}

// SimulatePart's post-processing (Simulator.cpp:868-888): the frame sum in double
// over 1024-pixel chunks in order (one thread per chunk), then over the chunks in
// order; x (float)(n_expected / sum); Poisson draw from stream 0; white noise
// (2u - 1) * adu_noise in double from stream 1; x value_per_photon in float.
// kVec4: every frame starts 16-byte aligned (npix % 4 == 0, aligned base), so a chunk
// is read as float4 and its four values added in order: the same sum, a quarter of the
// loads (each lane's loads walk its own 4 KB chunk, so every wave load touches 64 lines)
template <bool kVec4>
__global__ void __launch_bounds__(256)
    chunk_sums_kernel(const float *__restrict__ frames, uint32_t npix, uint32_t nch, double *__restrict__ part)
{
    const uint32_t f = blockIdx.y, c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nch) return;
    const float *I = frames + (size_t)f * npix;
    double s = 0.0;
    const uint32_t e = min(npix, (c + 1) * 1024);
    uint32_t l = c * 1024;
    if (kVec4) {
        const float4 *I4 = reinterpret_cast<const float4 *>(I);
        for (; l + 4 <= e; l += 4) {
            const float4 v = I4[l / 4];
            s += v.x;
            s += v.y;
            s += v.z;
            s += v.w;
        }
    }
    for (; l < e; l++) s += I[l];
    part[(size_t)f * nch + c] = s;
}

// the frame's factor n_expected / (sum of its chunk sums, in order)
__global__ void frame_factor_kernel(const double *__restrict__ part, uint32_t nch, uint32_t nframes,
                                    double n_expected, float *__restrict__ fac)
{
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nframes) return;
    double integ = 0.0;
    for (uint32_t c = 0; c < nch; c++) integ += part[(size_t)f * nch + c];
    fac[f] = (float)(n_expected / integ);
}

constexpr uint32_t kPostSlice = 4096;  // pixels per post-processing block

__global__ void __launch_bounds__(256)
    postprocess_kernel(float *__restrict__ frames, uint32_t npix, const float *__restrict__ fac_of, int do_poisson,
                       float adu_noise, float value_per_photon, uint64_t seed, uint32_t frame0)
{
    const uint32_t f = blockIdx.y;
    const float fac = fac_of[f];
    float *I = frames + (size_t)f * npix;
    const uint32_t l1 = min(npix, (blockIdx.x + 1) * kPostSlice);
    for (uint32_t l = blockIdx.x * kPostSlice + threadIdx.x; l < l1; l += blockDim.x) {
        float v = I[l] * fac;
        double u[2];
        if (do_poisson) {
            iac::uniform2(seed, frame0 + f, l, 0, u);
            v = (float)iac::poisson_sample((double)v, u);
        }
        if (adu_noise != 0.0f) {
            iac::uniform2(seed, frame0 + f, l, 1, u);
            v = (float)((double)v + (2.0 * u[0] - 1.0) * (double)adu_noise);
        }
        I[l] = v * value_per_photon;
    }
}

bool split_mode(int mode) { return mode == IAC_SIM_BF16X3 || mode == IAC_SIM_FP16X3; }
bool bf16_mode(int mode) { return mode == IAC_SIM_BF16X3 || mode == IAC_SIM_BF16; }

unsigned grid_for(uint64_t threads)
{
    return (unsigned)std::min<uint64_t>((threads + 255) / 256, 1u << 20);
}

template <bool kBf16, int NH>
void launch_gemm(hipStream_t st, bool direct, const uint4 *x, const uint4 *y, uint32_t nkc, uint32_t rb0, uint32_t nrb,
                 uint32_t ncb, uint32_t rows, uint32_t ncols, int kind, float *out, uint32_t npix, bool vec4, float *w,
                 uint32_t w_rows)
{
    const dim3 grid(nrb * ncb), block(256);
    if (direct)
        hipLaunchKernelGGL((sim_gemm_kernel<kBf16, NH, true>), grid, block, 0, st, x, y, nkc, rb0, nrb, ncb, rows,
                           ncols, kind, out, npix, vec4 ? 1 : 0, w, w_rows);
    else
        hipLaunchKernelGGL((sim_gemm_kernel<kBf16, NH, false>), grid, block, 0, st, x, y, nkc, rb0, nrb, ncb, rows,
                           ncols, kind, out, npix, 0, w, w_rows);
}

void launch_gemm_mode(int mode, hipStream_t st, bool direct, const uint4 *x, const uint4 *y, uint32_t nkc,
                      uint32_t rb0, uint32_t nrb, uint32_t ncb, uint32_t rows, uint32_t ncols, int kind, float *out,
                      uint32_t npix, bool vec4, float *w, uint32_t w_rows)
{
    switch (mode) {
    case IAC_SIM_BF16X3:
        launch_gemm<true, 2>(st, direct, x, y, nkc, rb0, nrb, ncb, rows, ncols, kind, out, npix, vec4, w, w_rows);
        break;
    case IAC_SIM_FP16X3:
        launch_gemm<false, 2>(st, direct, x, y, nkc, rb0, nrb, ncb, rows, ncols, kind, out, npix, vec4, w, w_rows);
        break;
    case IAC_SIM_BF16:
        launch_gemm<true, 1>(st, direct, x, y, nkc, rb0, nrb, ncb, rows, ncols, kind, out, npix, vec4, w, w_rows);
        break;
    default:
        launch_gemm<false, 1>(st, direct, x, y, nkc, rb0, nrb, ncb, rows, ncols, kind, out, npix, vec4, w, w_rows);
        break;
    }
}

}  // namespace

namespace iac {
void free_sim_resources(iac_ctx *) {}
}  // namespace iac

extern "C" {

int iac_sim_plan_create(iac_ctx *ctx, const float *d_pixmap, uint32_t npix, int subsampling, const double u_step[3],
                        const double v_step[3], double wavelength, const double *d_emitters, uint32_t nem, int mode,
                        iac_sim_plan **out)
{
    IAC_TRY(iac::bind(ctx));
    if (!out || !d_pixmap || !d_emitters || npix == 0 || nem == 0 || !u_step || !v_step || !(wavelength > 0))
        return iac::fail(IAC_ERR_INVALID_ARG, "bad simulator plan request");
    if (subsampling < 0 || subsampling > 8) return iac::fail(IAC_ERR_LIMIT, "subsampling is 0..8");
    if (nem > (1u << 29)) return iac::fail(IAC_ERR_LIMIT, "at most 2^29 emitters");
    if (mode < IAC_SIM_BF16X3 || mode > IAC_SIM_FP16)
        return iac::fail(IAC_ERR_INVALID_ARG, "plan modes are the GEMM modes (IAC_SIM_BF16X3..IAC_SIM_FP16)");
    *out = nullptr;
    const uint32_t nsub = (uint32_t)((2 * subsampling + 1) * (2 * subsampling + 1));
    const uint64_t rows = (uint64_t)npix * nsub;
    const uint64_t rows_pad = (rows + kRowsPerBlock - 1) / kRowsPerBlock * kRowsPerBlock;
    if (rows_pad >= (1ull << 31)) return iac::fail(IAC_ERR_LIMIT, "pixels x subpixels must stay below 2^31");
    auto *pl = new iac_sim_plan();
    pl->ctx = ctx;
    pl->mode = mode;
    pl->npix = npix;
    pl->nsub = nsub;
    pl->rows = (uint32_t)rows;
    pl->rows_pad = (uint32_t)rows_pad;
    pl->nem = nem;
    pl->nkc = (nem + 2 * kEmittersPerStep - 1) / (2 * kEmittersPerStep) * 2;  // even: the GEMM loop takes steps in pairs
    const int nh = split_mode(mode) ? 2 : 1;
    const size_t bytes = rows_pad * (size_t)pl->nkc * kEmittersPerStep * 2 * nh * 2;
    double *krow = nullptr;
    hipError_t e = hipMalloc(&pl->d_x, bytes);
    if (e == hipSuccess) e = hipMalloc(&krow, rows * 3 * sizeof(double));
    if (e != hipSuccess) {
        if (pl->d_x) (void)hipFree(pl->d_x);
        if (krow) (void)hipFree(krow);
        delete pl;
        return iac::hip_fail(e, "iac_sim_plan_create allocation", __FILE__, __LINE__);
    }
    const double kk = (2.0 * kPi) / wavelength;
    hipLaunchKernelGGL(k_rows_kernel, dim3((unsigned)std::min<uint64_t>(65535, (rows + 255) / 256)), dim3(256), 0,
                       ctx->stream, d_pixmap, npix, subsampling, make_double3(u_step[0], u_step[1], u_step[2]),
                       make_double3(v_step[0], v_step[1], v_step[2]), kk, krow);
    const uint64_t slots = (rows_pad / 32) * (uint64_t)pl->nkc * 64;
    if (bf16_mode(mode))
        hipLaunchKernelGGL(phase_matrix_kernel<__bf16>, dim3(grid_for(slots)), dim3(256), 0, ctx->stream, krow,
                           pl->rows, pl->rows_pad, d_emitters, nem, pl->nkc, nh, static_cast<uint4 *>(pl->d_x));
    else
        hipLaunchKernelGGL(phase_matrix_kernel<_Float16>, dim3(grid_for(slots)), dim3(256), 0, ctx->stream, krow,
                           pl->rows, pl->rows_pad, d_emitters, nem, pl->nkc, nh, static_cast<uint4 *>(pl->d_x));
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(krow);
    if (e != hipSuccess) {
        (void)hipFree(pl->d_x);
        delete pl;
        return iac::hip_fail(e, "iac_sim_plan_create", __FILE__, __LINE__);
    }
    *out = pl;
    return IAC_OK;
}

int iac_sim_plan_destroy(iac_sim_plan *plan)
{
    if (!plan) return IAC_OK;
    (void)hipSetDevice(plan->ctx->device);
    (void)hipStreamSynchronize(plan->ctx->stream);
    if (plan->d_x) (void)hipFree(plan->d_x);
    delete plan;
    return IAC_OK;
}

int iac_sim_run(iac_ctx *ctx, iac_sim_plan *plan, const double *d_phases, uint32_t nframes, uint32_t modes,
                int out_kind, float *d_out)
{
    IAC_TRY(iac::bind(ctx));
    if (!plan || plan->ctx != ctx || !d_phases || !d_out || modes == 0)
        return iac::fail(IAC_ERR_INVALID_ARG, "bad simulator run");
    if (out_kind != IAC_SIM_AMPLITUDE && out_kind != IAC_SIM_INTENSITY)
        return iac::fail(IAC_ERR_INVALID_ARG, "out_kind is IAC_SIM_AMPLITUDE or IAC_SIM_INTENSITY");
    if (nframes == 0) return IAC_OK;
    const uint64_t cols = (uint64_t)nframes * modes;
    if (cols > (1u << 24)) return iac::fail(IAC_ERR_LIMIT, "at most 2^24 frames x modes per call");
    const uint32_t ncols = (uint32_t)cols;
    const uint32_t ncols_pad = (ncols + kColsPerBlock - 1) / kColsPerBlock * kColsPerBlock;
    const uint32_t ncb = ncols_pad / kColsPerBlock;
    const int nh = split_mode(plan->mode) ? 2 : 1;
    const bool bf = bf16_mode(plan->mode);
    // the phase operand (scratch slot 0)
    void *yb = nullptr;
    const size_t ybytes = (size_t)ncols_pad * plan->nkc * kEmittersPerStep * 2 * nh * 2;
    IAC_TRY(iac::scratch(ctx, 0, ybytes, &yb));
    const uint64_t yslots = (uint64_t)(ncols_pad / 32) * plan->nkc * 64;
    if (bf)
        hipLaunchKernelGGL(phase_vector_kernel<__bf16>, dim3(grid_for(yslots)), dim3(256), 0, ctx->stream, d_phases,
                           ncols, ncols_pad, plan->nem, plan->nkc, nh, static_cast<uint4 *>(yb));
    else
        hipLaunchKernelGGL(phase_vector_kernel<_Float16>, dim3(grid_for(yslots)), dim3(256), 0, ctx->stream,
                           d_phases, ncols, ncols_pad, plan->nem, plan->nkc, nh, static_cast<uint4 *>(yb));
    IAC_HIP(hipGetLastError());
    const uint4 *x = static_cast<const uint4 *>(plan->d_x), *y = static_cast<const uint4 *>(yb);
    const uint32_t nrb_all = plan->rows_pad / kRowsPerBlock;
    const bool direct = plan->nsub == 1 && modes == 1;
    if (direct) {
        // one launch; the magnitudes go straight to d_out (float4 stores need npix % 4 == 0
        // and a 16-byte aligned output)
        const bool vec4 = plan->npix % 4 == 0 && reinterpret_cast<uintptr_t>(d_out) % 16 == 0;
        const uint32_t nrb_step = std::max<uint32_t>(1, (1u << 30) / ncb);  // blocks per launch < 2^31
        IAC_TRY(iac::timer_begin(ctx));
        for (uint32_t rb0 = 0; rb0 < nrb_all; rb0 += nrb_step) {
            launch_gemm_mode(plan->mode, ctx->stream, true, x, y, plan->nkc, rb0, std::min(nrb_step, nrb_all - rb0),
                             ncb, plan->rows, ncols, out_kind, d_out, plan->npix, vec4, nullptr, 0);
            IAC_HIP(hipGetLastError());
        }
        IAC_TRY(iac::timer_end(ctx));
        return IAC_OK;
    }
    // subpixels or modes: magnitudes per (row, column) into w (scratch slot 1), chunks of
    // whole 256-pixel groups (a multiple of 256 rows), then the ordered reduction
    uint64_t chunk_pix = ctx->sim_chunk_rows ? ctx->sim_chunk_rows / plan->nsub
                                             : (1ull << 30) / (4ull * ncols_pad * plan->nsub);
    chunk_pix = std::max<uint64_t>(kRowsPerBlock, chunk_pix / kRowsPerBlock * kRowsPerBlock);
    const uint64_t chunk_rows_max = std::min<uint64_t>(chunk_pix * plan->nsub, plan->rows_pad);
    void *wb = nullptr;
    IAC_TRY(iac::scratch(ctx, 1, chunk_rows_max * ncols_pad * sizeof(float), &wb));
    IAC_TRY(iac::timer_begin(ctx));
    for (uint64_t p0 = 0; p0 < plan->npix; p0 += chunk_pix) {
        const uint64_t np = std::min<uint64_t>(chunk_pix, plan->npix - p0);
        const uint64_t r0 = p0 * plan->nsub;  // a multiple of 256
        const uint64_t r1 = std::min<uint64_t>(plan->rows_pad, (p0 + np) * plan->nsub + kRowsPerBlock - 1) /
                            kRowsPerBlock * kRowsPerBlock;
        const uint32_t rb0 = (uint32_t)(r0 / kRowsPerBlock), nrb = (uint32_t)((r1 - r0 + kRowsPerBlock - 1) / kRowsPerBlock);
        const uint32_t w_rows = nrb * kRowsPerBlock;
        launch_gemm_mode(plan->mode, ctx->stream, false, x, y, plan->nkc, rb0, nrb, ncb, plan->rows, ncols, out_kind,
                         nullptr, plan->npix, false, static_cast<float *>(wb), w_rows);
        IAC_HIP(hipGetLastError());
        hipLaunchKernelGGL(subpixel_mode_reduce_kernel, dim3((unsigned)((np + 255) / 256), std::min(nframes, 65535u)),
                           dim3(256), 0,
                           ctx->stream, static_cast<const float *>(wb), w_rows, (uint32_t)p0, (uint32_t)np, plan->nsub,
                           nframes, modes, d_out, plan->npix);
        IAC_HIP(hipGetLastError());
    }
    IAC_TRY(iac::timer_end(ctx));
    return IAC_OK;
}

int iac_sim_direct_fp32(iac_ctx *ctx, const float *d_pixmap, uint32_t npix, int subsampling, const float u_step[3],
                        const float v_step[3], float wavelength, const float *d_emitters, uint32_t nem,
                        const float *d_phases, uint32_t nframes, uint32_t modes, int out_kind, float *d_out)
{
    IAC_TRY(iac::bind(ctx));
    if (!d_pixmap || !d_emitters || !d_phases || !d_out || !u_step || !v_step || npix == 0 || modes == 0 ||
        !(wavelength > 0))
        return iac::fail(IAC_ERR_INVALID_ARG, "bad simulator request");
    if (subsampling < 0 || subsampling > 8) return iac::fail(IAC_ERR_LIMIT, "subsampling is 0..8");
    if (out_kind != IAC_SIM_AMPLITUDE && out_kind != IAC_SIM_INTENSITY)
        return iac::fail(IAC_ERR_INVALID_ARG, "out_kind is IAC_SIM_AMPLITUDE or IAC_SIM_INTENSITY");
    if (nframes == 0) return IAC_OK;
    if (nframes > 65535) return iac::fail(IAC_ERR_LIMIT, "at most 65535 frames per call");
    IAC_TRY(iac::timer_begin(ctx));
    hipLaunchKernelGGL(direct_fp32_kernel, dim3((npix + 255) / 256, nframes), dim3(256), 0, ctx->stream, d_pixmap,
                       npix, subsampling, make_float3(u_step[0], u_step[1], u_step[2]),
                       make_float3(v_step[0], v_step[1], v_step[2]), wavelength, d_emitters, nem, d_phases, modes,
                       out_kind, d_out);
    IAC_HIP(hipGetLastError());
    IAC_TRY(iac::timer_end(ctx));
    return IAC_OK;
}

int iac_sim_postprocess(iac_ctx *ctx, float *d_frames, uint32_t npix, uint32_t nframes, double n_expected,
                        int do_poisson, float adu_noise, float value_per_photon, uint64_t seed, uint32_t frame0)
{
    IAC_TRY(iac::bind(ctx));
    if (!d_frames || npix == 0) return iac::fail(IAC_ERR_INVALID_ARG, "bad post-processing request");
    if (npix > 4096u * 1024u) return iac::fail(IAC_ERR_LIMIT, "at most 4 Mi pixels per frame");
    if (nframes == 0) return IAC_OK;
    if (nframes > 65535) return iac::fail(IAC_ERR_LIMIT, "at most 65535 frames per call");
    // chunk sums (1024 pixels each, in order), the frame factor over them in order, then
    // rescale + Poisson + noise over 4096-pixel slices
    const uint32_t nch = (npix + 1023) / 1024;
    void *sp = nullptr;
    IAC_TRY(iac::scratch(ctx, 2, (size_t)nframes * nch * sizeof(double) + (size_t)nframes * sizeof(float), &sp));
    double *part = static_cast<double *>(sp);
    float *fac = reinterpret_cast<float *>(part + (size_t)nframes * nch);
    const bool vec4 = npix % 4 == 0 && reinterpret_cast<uintptr_t>(d_frames) % 16 == 0;
    hipLaunchKernelGGL(vec4 ? chunk_sums_kernel<true> : chunk_sums_kernel<false>, dim3((nch + 255) / 256, nframes),
                       dim3(256), 0, ctx->stream, d_frames, npix, nch, part);
    hipLaunchKernelGGL(frame_factor_kernel, dim3((nframes + 255) / 256), dim3(256), 0, ctx->stream, part, nch,
                       nframes, n_expected, fac);
    hipLaunchKernelGGL(postprocess_kernel, dim3((npix + kPostSlice - 1) / kPostSlice, nframes), dim3(256), 0,
                       ctx->stream, d_frames, npix, fac, do_poisson, adu_noise, value_per_photon, seed, frame0);
    IAC_HIP(hipGetLastError());
    return IAC_OK;
}

}  // extern "C"
