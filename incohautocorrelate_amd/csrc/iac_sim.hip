// iac_sim.hip -- the emitter speckle simulator (SimulateCrystal, IAC/kernel.cl:768-859,
// driven by Simulator::SimulatePart, IAC/Simulator.cpp:662-944) and its
// post-processing (rescale + Poisson, IAC/Simulator.cpp:870-888) on gfx950.
//
// Two ways to the field psi(p, f) = sum_e exp(i (k_p . r_e + phi_{e,f})):
//   direct FP32: the reference kernel's arithmetic, one thread per (pixel, frame),
//     emitters staged through LDS.  The plain baseline.
//   complex GEMM: psi = A B with A[p,e] = exp(i k_p . r_e) (phase argument and its
//     reduction in FP64, once per geometry) and B[e,f] = exp(i phi_{e,f}).  As one
//     real GEMM Z = X Y: X = [cos | sin] (rows x 2E), Y[f] = [cos phi | -sin phi],
//     Y[F+f] = [sin phi | cos phi], Z = [Re psi | Im psi].  X and Y are stored split,
//     x = hi + lo in bf16 or fp16; Z = Xh Yh + Xh Yl + Xl Yh on MFMA with FP32
//     accumulation recovers ~16 bits per element (the dropped Xl Yl is ~2^-16 of a term).
// The epilogue forms |psi| (reference, kernel.cl:844) or |psi|^2, averaged over subpixels.
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "iac_internal.h"
#include "iac_rng.h"

#pragma clang fp contract(off)

struct iac_sim_plan {
    iac_ctx *ctx = nullptr;
    int mode = 0;
    uint32_t npix = 0, nsub = 0, rows = 0, nem = 0;
    void *d_xhi = nullptr, *d_xlo = nullptr;  // [rows][2 nem] 16-bit; lo only for split modes
};

namespace {

constexpr double kPi = 3.14159265358979323846;

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
__device__ __forceinline__ void split16(double x, T &hi, T &lo)
{
    hi = (T)(float)x;
    lo = (T)(float)(x - (double)(float)hi);
}

// X[row][e] = cos(k_row . r_e), X[row][E + e] = sin(k_row . r_e), split hi/lo
template <typename T>
__global__ void phase_matrix_kernel(const double *__restrict__ krow, uint64_t rows, const double *__restrict__ em,
                                    uint32_t nem, T *__restrict__ xhi, T *__restrict__ xlo)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nem) return;
    const double ex = em[3 * (size_t)e], ey = em[3 * (size_t)e + 1], ez = em[3 * (size_t)e + 2];
    const size_t K = 2 * (size_t)nem;
    for (uint64_t r = blockIdx.y; r < rows; r += gridDim.y) {
        const double arg = krow[3 * r] * ex + krow[3 * r + 1] * ey + krow[3 * r + 2] * ez;
        double sn, cs;
        sincos(arg, &sn, &cs);
        T h, l;
        split16(cs, h, l);
        xhi[r * K + e] = h;
        if (xlo) xlo[r * K + e] = l;
        split16(sn, h, l);
        xhi[r * K + nem + e] = h;
        if (xlo) xlo[r * K + nem + e] = l;
    }
}

// Y rows: n = c -> [cos phi | -sin phi], n = C + c -> [sin phi | cos phi], for the
// C = frames x modes phase columns c; lo rows only for split modes (ylo non-null).
template <typename T>
__global__ void phase_vector_kernel(const double *__restrict__ phases, uint32_t ncols, uint32_t nem,
                                    T *__restrict__ yhi, T *__restrict__ ylo)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nem) return;
    const size_t K = 2 * (size_t)nem;
    for (uint32_t c = blockIdx.y; c < ncols; c += gridDim.y) {
        double sn, cs;
        sincos(phases[(size_t)c * nem + e], &sn, &cs);
        const size_t r0 = (size_t)c * K, r1 = (size_t)(ncols + c) * K;
        T h, l;
        split16(cs, h, l);
        yhi[r0 + e] = h;
        yhi[r1 + nem + e] = h;
        if (ylo) {
            ylo[r0 + e] = l;
            ylo[r1 + nem + e] = l;
        }
        split16(-sn, h, l);
        yhi[r0 + nem + e] = h;
        if (ylo) ylo[r0 + nem + e] = l;
        split16(sn, h, l);
        yhi[r1 + e] = h;
        if (ylo) ylo[r1 + e] = l;
    }
}

// out[f][p] = sum over the modes m of the subpixel mean of |psi| (kind 0) or
// |psi|^2 (kind 1) of column f * modes + m; Z is [rows][2C], C = frames x modes
// (Re in column c, Im in C + c).  32 x 32 tile transposed through LDS.
__global__ void epilogue_kernel(const float *__restrict__ Z, uint32_t rows0, uint32_t nrows, uint32_t nsub,
                                uint32_t nframes, uint32_t modes, int kind, float *__restrict__ out, uint32_t npix)
{
    __shared__ float tile[32][33];
    const uint32_t f0 = blockIdx.x * 32, p0 = blockIdx.y * 32;  // pixel index relative to rows0 / nsub
    const uint32_t tx = threadIdx.x, ty = threadIdx.y;          // 32 x 8
    const uint32_t C = nframes * modes, C2 = 2 * C;
    for (uint32_t j = ty; j < 32; j += 8) {
        const uint32_t pl = p0 + j, f = f0 + tx;
        float acc = 0.0f;
        if (pl * nsub < nrows && f < nframes) {
            for (uint32_t m = 0; m < modes; m++) {
                const uint32_t c = f * modes + m;
                float t = 0.0f;
                for (uint32_t s = 0; s < nsub; s++) {
                    const size_t row = (size_t)pl * nsub + s;
                    const float re = Z[row * C2 + c], im = Z[row * C2 + C + c];
                    const float a2 = re * re + im * im;
                    t += kind == 0 ? sqrtf(a2) : a2;
                }
                acc += t / (float)nsub;
            }
        }
        tile[j][tx] = acc;
    }
    __syncthreads();
    for (uint32_t j = ty; j < 32; j += 8) {
        const uint32_t f = f0 + j, pl = p0 + tx;
        const uint32_t p = rows0 / nsub + pl;
        if (f < nframes && pl * nsub < nrows && p < npix) out[(size_t)f * npix + p] = tile[tx][j];
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

int rocblas_handle_for(iac_ctx *ctx, rocblas_handle *h)
{
    if (!ctx->rocblas) {
        rocblas_handle nh;
        if (rocblas_create_handle(&nh) != rocblas_status_success)
            return iac::fail(IAC_ERR_HIP, "rocblas_create_handle failed");
        ctx->rocblas = nh;
    }
    *h = (rocblas_handle)ctx->rocblas;
    if (rocblas_set_stream(*h, ctx->stream) != rocblas_status_success)
        return iac::fail(IAC_ERR_HIP, "rocblas_set_stream failed");
    return IAC_OK;
}

}  // namespace

namespace iac {
void free_sim_resources(iac_ctx *ctx)
{
    if (ctx->rocblas) {
        (void)rocblas_destroy_handle((rocblas_handle)ctx->rocblas);
        ctx->rocblas = nullptr;
    }
}
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
    auto *pl = new iac_sim_plan();
    pl->ctx = ctx;
    pl->mode = mode;
    pl->npix = npix;
    pl->nsub = (uint32_t)((2 * subsampling + 1) * (2 * subsampling + 1));
    const uint64_t rows = (uint64_t)npix * pl->nsub;
    if (rows >= (1ull << 31)) {
        delete pl;
        return iac::fail(IAC_ERR_LIMIT, "pixels x subpixels must stay below 2^31");
    }
    pl->rows = (uint32_t)rows;
    pl->nem = nem;
    const size_t K = 2 * (size_t)nem;
    const size_t bytes = rows * K * 2;
    double *krow = nullptr;
    hipError_t e = hipMalloc(&pl->d_xhi, bytes);
    if (e == hipSuccess && split_mode(mode)) e = hipMalloc(&pl->d_xlo, bytes);
    if (e == hipSuccess) e = hipMalloc(&krow, rows * 3 * sizeof(double));
    if (e != hipSuccess) {
        if (pl->d_xhi) (void)hipFree(pl->d_xhi);
        if (pl->d_xlo) (void)hipFree(pl->d_xlo);
        if (krow) (void)hipFree(krow);
        delete pl;
        return iac::hip_fail(e, "iac_sim_plan_create allocation", __FILE__, __LINE__);
    }
    const double kk = (2.0 * kPi) / wavelength;
    hipLaunchKernelGGL(k_rows_kernel, dim3((unsigned)std::min<uint64_t>(65535, (rows + 255) / 256)), dim3(256), 0,
                       ctx->stream, d_pixmap, npix, subsampling, make_double3(u_step[0], u_step[1], u_step[2]),
                       make_double3(v_step[0], v_step[1], v_step[2]), kk, krow);
    const dim3 grid((nem + 255) / 256, (unsigned)std::min<uint64_t>(65535, rows));
    if (bf16_mode(mode))
        hipLaunchKernelGGL(phase_matrix_kernel<__bf16>, grid, dim3(256), 0, ctx->stream, krow, rows, d_emitters, nem,
                           (__bf16 *)pl->d_xhi, (__bf16 *)pl->d_xlo);
    else
        hipLaunchKernelGGL(phase_matrix_kernel<_Float16>, grid, dim3(256), 0, ctx->stream, krow, rows, d_emitters,
                           nem, (_Float16 *)pl->d_xhi, (_Float16 *)pl->d_xlo);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(krow);
    if (e != hipSuccess) {
        (void)hipFree(pl->d_xhi);
        (void)hipFree(pl->d_xlo);
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
    if (plan->d_xhi) (void)hipFree(plan->d_xhi);
    if (plan->d_xlo) (void)hipFree(plan->d_xlo);
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
    const uint32_t E = plan->nem;
    const uint64_t N = 2 * cols;
    const size_t K = 2 * (size_t)E;
    const bool split = split_mode(plan->mode), bf = bf16_mode(plan->mode);
    // Y hi and lo (scratch slot 0), a Z chunk (slot 1)
    void *yb = nullptr, *zb = nullptr;
    const size_t ybytes = (size_t)N * K * 2;
    IAC_TRY(iac::scratch(ctx, 0, (split ? 2 : 1) * ybytes, &yb));
    void *yhi = yb, *ylo = split ? static_cast<char *>(yb) + ybytes : nullptr;
    // rows per GEMM chunk: a Z chunk of at most 1 GiB holding whole pixels
    uint64_t chunk = ctx->sim_chunk_rows ? ctx->sim_chunk_rows : (1ull << 30) / (4ull * N);
    chunk = std::max<uint64_t>(plan->nsub, chunk);
    chunk = std::min<uint64_t>(chunk, plan->rows);
    chunk = std::min<uint64_t>(chunk, 65535ull * 32 * plan->nsub);  // epilogue grid.y bound
    chunk -= chunk % plan->nsub;
    IAC_TRY(iac::scratch(ctx, 1, chunk * N * sizeof(float), &zb));
    rocblas_handle h;
    IAC_TRY(rocblas_handle_for(ctx, &h));
    const dim3 yg((E + 255) / 256, (unsigned)std::min<uint64_t>(cols, 65535));
    if (bf)
        hipLaunchKernelGGL(phase_vector_kernel<__bf16>, yg, dim3(256), 0, ctx->stream, d_phases, (uint32_t)cols, E,
                           (__bf16 *)yhi, (__bf16 *)ylo);
    else
        hipLaunchKernelGGL(phase_vector_kernel<_Float16>, yg, dim3(256), 0, ctx->stream, d_phases, (uint32_t)cols, E,
                           (_Float16 *)yhi, (_Float16 *)ylo);
    IAC_HIP(hipGetLastError());
    const rocblas_datatype dt = bf ? rocblas_datatype_bf16_r : rocblas_datatype_f16_r;
    const float one = 1.0f, zero = 0.0f;
    IAC_TRY(iac::timer_begin(ctx));
    for (uint64_t r0 = 0; r0 < plan->rows; r0 += chunk) {
        const uint64_t nr = std::min<uint64_t>(chunk, plan->rows - r0);
        const char *xh = static_cast<const char *>(plan->d_xhi) + r0 * K * 2;
        const char *xl = split ? static_cast<const char *>(plan->d_xlo) + r0 * K * 2 : nullptr;
        // column major: Z^T (N x nr) = Y^T (N x K) * X^T (K x nr); the lo x lo term is dropped
        struct Term {
            const void *y, *x;
            const float *beta;
        } terms[3] = {{yhi, xh, &zero}, {ylo, xh, &one}, {yhi, xl, &one}};
        const int nterms = split ? 3 : 1;
        for (int t = 0; t < nterms; t++) {
            rocblas_status st = rocblas_gemm_ex(
                h, rocblas_operation_transpose, rocblas_operation_none, (rocblas_int)N, (rocblas_int)nr,
                (rocblas_int)K, &one, terms[t].y, dt, (rocblas_int)K, terms[t].x, dt, (rocblas_int)K, terms[t].beta,
                zb, rocblas_datatype_f32_r, (rocblas_int)N, zb, rocblas_datatype_f32_r, (rocblas_int)N,
                rocblas_datatype_f32_r, rocblas_gemm_algo_standard, 0, 0);
            if (st != rocblas_status_success)
                return iac::fail(IAC_ERR_HIP, std::string("rocblas_gemm_ex: ") + rocblas_status_to_string(st));
        }
        const dim3 eg((nframes + 31) / 32, (unsigned)((nr / plan->nsub + 31) / 32));
        hipLaunchKernelGGL(epilogue_kernel, eg, dim3(32, 8), 0, ctx->stream, static_cast<const float *>(zb),
                           (uint32_t)r0, (uint32_t)nr, plan->nsub, nframes, modes, out_kind, d_out, plan->npix);
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
