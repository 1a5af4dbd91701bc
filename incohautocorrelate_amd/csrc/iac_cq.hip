// iac_cq.hip -- C(q): dense pixel-pair normalisation mesh, rotation merge and
// the final AC/C(q) division, on gfx950.
//
//   iac_cq_small / iac_cq_dense_frames: AutoCorr_CQ_small (IAC/kernel.cl:27-184)
//     through the LDS cell-box pair engine (one record per pixel).
//   iac_merge_cq: Merge_CQ (IAC/kernel.cl:189-344), one thread per small-mesh
//     voxel and event slice.
//   iac_merge_ac_cq: RunIAC::Merge_ACandCQ (IAC/RunIAC.cpp:120-158).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "iac_binning.h"
#include "iac_internal.h"
#include "iac_pairs.h"

#pragma clang fp contract(off)

namespace {

// records {k[p], I[f][p]} of dense frames; off[f] = f * npix
__global__ void gather_dense_kernel(const float *__restrict__ kmap, const float *__restrict__ frames,
                                    uint32_t npix, uint32_t nframes, float4 *__restrict__ rec,
                                    uint64_t *__restrict__ off)
{
    const uint64_t n = (uint64_t)npix * nframes;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t p = (uint32_t)(i % npix);
        rec[i] = make_float4(kmap[3 * (size_t)p], kmap[3 * (size_t)p + 1], kmap[3 * (size_t)p + 2], frames[i]);
    }
    for (uint64_t f = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; f <= nframes; f += (uint64_t)gridDim.x * blockDim.x)
        off[f] = f * npix;
}

__global__ void u64_to_f64_kernel(const unsigned long long *__restrict__ in, size_t n, double mult,
                                  double *__restrict__ out)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = (double)in[i] / mult;
}

// Merge_CQ: small voxel `ind` (blockIdx.x * blockDim.x + threadIdx.x) rotated
// by the events [e0, e1) of this block row.  Row-major rotation, nearest or
// linear binning into the cubic big mesh whose slow-scan index also uses the
// AB centre (kernel.cl:287).
__global__ void merge_cq_kernel(const double *__restrict__ small, size_t small_stride, int sab, int sc, int ka,
                                int kb, int kc, const float *__restrict__ rw, uint32_t nev, uint32_t ev_per_row,
                                int interp, double mult, unsigned long long *__restrict__ big,
                                unsigned long long *__restrict__ dropped)
{
    const size_t nsmall = (size_t)sab * sab * sc;
    const size_t ind = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (ind >= nsmall) return;
    const uint32_t e0 = blockIdx.y * ev_per_row, e1 = min(nev, e0 + ev_per_row);
    const int cab = (sab - 1) / 2, cc = (sc - 1) / 2;
    const int ss_in = (int)(ind / ((size_t)sab * sab));
    const int ms_in = (int)((ind - (size_t)ss_in * sab * sab) / sab);
    const int fs_in = (int)(ind - (size_t)ss_in * sab * sab - (size_t)ms_in * sab);
    float q_in[3];
    q_in[ka] = (float)(fs_in - cab);
    q_in[kb] = (float)(ms_in - cab);
    q_in[kc] = (float)(ss_in - cc);
    auto add = [&](int fs, int ms, int ss, unsigned long long v) {
        if ((unsigned)fs < (unsigned)sab && (unsigned)ms < (unsigned)sab && (unsigned)ss < (unsigned)sab) {
            if (v) atomicAdd(big + (size_t)fs + (size_t)ms * sab + (size_t)ss * sab * sab, v);
        } else {
            atomicAdd(dropped, 1ull);
        }
    };
    for (uint32_t i = e0; i < e1; i++) {
        const double val_in = small[small_stride * i + ind];  // stride 0: one mesh for all events
        if (val_in < (double)1e-37f) continue;
        const float *R = rw + 10 * (size_t)i;
        const float q0 = R[0] * q_in[0] + R[1] * q_in[1] + R[2] * q_in[2];
        const float q1 = R[3] * q_in[0] + R[4] * q_in[1] + R[5] * q_in[2];
        const float q2 = R[6] * q_in[0] + R[7] * q_in[1] + R[8] * q_in[2];
        const double w = (double)R[9];
        if (interp == 0) {
            add((int)(roundf(q0) + (float)cab), (int)(roundf(q1) + (float)cab), (int)(roundf(q2) + (float)cab),
                (unsigned long long)((val_in * w) * mult));
        } else {
            const double sfs = (double)(1.0f - (q0 - floorf(q0)));
            const double sms = (double)(1.0f - (q1 - floorf(q1)));
            const double sss = (double)(1.0f - (q2 - floorf(q2)));
            const int fs = (int)(floorf(q0) + (float)cab);
            const int ms = (int)(floorf(q1) + (float)cab);
            const int ss = (int)(floorf(q2) + (float)cab);
            const double vmw = (val_in * mult) * w;
            add(fs + 0, ms + 0, ss + 0, (unsigned long long)(vmw * ((sfs * sms) * sss)));
            add(fs + 0, ms + 0, ss + 1, (unsigned long long)(vmw * ((sfs * sms) * (1 - sss))));
            add(fs + 1, ms + 0, ss + 0, (unsigned long long)(vmw * (((1 - sfs) * sms) * sss)));
            add(fs + 1, ms + 0, ss + 1, (unsigned long long)(vmw * (((1 - sfs) * sms) * (1 - sss))));
            add(fs + 0, ms + 1, ss + 0, (unsigned long long)(vmw * ((sfs * (1 - sms)) * sss)));
            add(fs + 0, ms + 1, ss + 1, (unsigned long long)(vmw * ((sfs * (1 - sms)) * (1 - sss))));
            add(fs + 1, ms + 1, ss + 0, (unsigned long long)(vmw * (((1 - sfs) * (1 - sms)) * sss)));
            add(fs + 1, ms + 1, ss + 1, (unsigned long long)(vmw * (((1 - sfs) * (1 - sms)) * (1 - sss))));
        }
    }
}

// Merge_ACandCQ: AC = IntToFloat(Mesh) (a float) / CQ in double, 0 where CQ <= 0
// or the quotient is NaN.
__global__ void merge_ac_cq_kernel(const unsigned long long *__restrict__ ac, const double *__restrict__ cq,
                                   size_t n, double *__restrict__ out)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float a = ((float)ac[i] * 0.001f) - 0.0f;
        const double c = cq[i];
        double o = c <= 0 ? 0.0 : (double)a / c;
        if (isnan(o)) o = 0.0;
        out[i] = o;
    }
}

int check_small(const iac_mesh_shape *s)
{
    if (!s) return iac::fail(IAC_ERR_INVALID_ARG, "null mesh shape");
    if (s->size_ab % 2 == 0 || s->size_c % 2 == 0 || s->size_ab < 5)
        return iac::fail(IAC_ERR_EVEN_MESH, "Meshsize must be odd for each dimension");
    if (s->k_a > 2 || s->k_b > 2 || s->k_c > 2 || s->k_a == s->k_b || s->k_a == s->k_c || s->k_b == s->k_c)
        return iac::fail(IAC_ERR_INVALID_ARG, "mesh axis alignment must be a permutation of 0,1,2");
    return IAC_OK;
}

unsigned blocks_for(size_t n) { return (unsigned)std::max<size_t>(1, std::min<size_t>(65535, (n + 255) / 256)); }

int cq_frames(iac_ctx *ctx, const float *d_kmap, uint32_t npix, const float *d_frames, uint32_t nframes,
              const uint32_t *h_group_off, uint32_t ngroups, const iac_mesh_shape *shape, int interp, double mult,
              uint64_t *d_meshes)
{
    IAC_TRY(iac::bind(ctx));
    IAC_TRY(check_small(shape));
    if (!d_kmap || !d_frames || !d_meshes || npix == 0) return iac::fail(IAC_ERR_INVALID_ARG, "null input");
    if (interp != IAC_NEAREST && interp != IAC_LINEAR) return iac::fail(IAC_ERR_INVALID_ARG, "bad interpolation");
    if (!(mult > 0)) return iac::fail(IAC_ERR_INVALID_ARG, "multiplicator must be > 0");
    if (nframes == 0) return IAC_OK;
    const uint64_t n = (uint64_t)npix * nframes;
    // records, then the frame offsets, in scratch slot 0
    void *rec = nullptr;
    IAC_TRY(iac::scratch(ctx, 0, n * sizeof(float4) + (nframes + 1) * sizeof(uint64_t), &rec));
    void *off = static_cast<char *>(rec) + n * sizeof(float4);
    hipLaunchKernelGGL(gather_dense_kernel, dim3(blocks_for(n)), dim3(256), 0, ctx->stream, d_kmap, d_frames,
                       npix, nframes, static_cast<float4 *>(rec), static_cast<uint64_t *>(off));
    IAC_HIP(hipGetLastError());
    iac::PairRun R;
    std::memset(&R.spec, 0, sizeof(R.spec));
    R.spec.kind = iac::kKindCq;
    R.spec.ka = (int)shape->k_a;
    R.spec.kb = (int)shape->k_b;
    R.spec.kc = (int)shape->k_c;
    R.spec.interp = interp;
    R.spec.mult = mult;
    R.spec.P.cut_sq = iac::cut_sq(shape->max_q);
    R.spec.P.vox = (float)shape->voxel_per_dq;
    R.spec.P.c0 = (int)(shape->size_ab - 1) / 2;
    R.spec.P.c1 = (int)(shape->size_ab - 1) / 2;
    R.spec.P.c2 = (int)(shape->size_c - 1) / 2;
    R.spec.P.sab = (int)shape->size_ab;
    R.spec.P.sc = (int)shape->size_c;
    R.spec.P.rot_identity = 1;
    R.ordered = interp == IAC_LINEAR;
    R.d_kmap = d_kmap;
    R.npix = npix;
    R.rec = static_cast<const float4 *>(rec);
    R.d_frame_off = static_cast<const uint64_t *>(off);
    R.nframes = nframes;
    R.n_total = n;
    R.h_group_off = h_group_off;
    R.ngroups = ngroups;
    R.d_mesh = d_meshes;
    return iac::pairs_run(ctx, R);
}

}  // namespace

extern "C" {

int iac_cq_small_multiplicator(const float *intensity, uint32_t npix, double *mult)
{
    if (!intensity || !mult || npix == 0) return iac::fail(IAC_ERR_INVALID_ARG, "null input");
    double t = 0;
    for (uint32_t i = 0; i < npix; i++) t += intensity[i];
    t = t / npix;
    const float mean = (float)t;
    if (mean <= 0)
        return iac::fail(IAC_ERR_PRECONDITION, "No intensity > 0 found in averaged intensity");
    double m = 1e-10;
    for (; 1 > mean * mean * m;) m *= 10;
    *mult = m * 100000;
    return IAC_OK;
}

int iac_cq_small(iac_ctx *ctx, const float *d_kmap, uint32_t npix, const float *d_intensity,
                 const iac_mesh_shape *small_shape, int interp, double mult, uint64_t *d_mesh)
{
    const uint32_t goff[2] = {0, 1};
    return cq_frames(ctx, d_kmap, npix, d_intensity, 1, goff, 1, small_shape, interp, mult, d_mesh);
}

int iac_cq_dense_frames(iac_ctx *ctx, const float *d_kmap, uint32_t npix, const float *d_frames,
                        uint32_t nframes, const uint32_t *h_group_off, uint32_t ngroups,
                        const iac_mesh_shape *shape, int interp, double mult, uint64_t *d_meshes)
{
    if (!h_group_off || ngroups == 0) return iac::fail(IAC_ERR_INVALID_ARG, "no frame groups");
    if (h_group_off[0] != 0 || h_group_off[ngroups] != nframes)
        return iac::fail(IAC_ERR_SHAPE, "frame groups must cover [0, nframes)");
    for (uint32_t g = 0; g < ngroups; g++)
        if (h_group_off[g + 1] < h_group_off[g]) return iac::fail(IAC_ERR_SHAPE, "frame groups must be sorted");
    return cq_frames(ctx, d_kmap, npix, d_frames, nframes, h_group_off, ngroups, shape, interp, mult, d_meshes);
}

int iac_u64_to_f64(iac_ctx *ctx, const uint64_t *d_in, size_t n, double mult, double *d_out)
{
    IAC_TRY(iac::bind(ctx));
    if (!d_in || !d_out) return iac::fail(IAC_ERR_INVALID_ARG, "null pointer");
    if (n == 0) return IAC_OK;
    hipLaunchKernelGGL(u64_to_f64_kernel, dim3(blocks_for(n)), dim3(256), 0, ctx->stream,
                       reinterpret_cast<const unsigned long long *>(d_in), n, mult, d_out);
    IAC_HIP(hipGetLastError());
    return IAC_OK;
}

int iac_merge_prepare(const float *rot9, const float *mean_intensity, uint32_t nev, const double *small_cq,
                      size_t nsmall, float *rw10, double *mult_out)
{
    if (!rot9 || !mean_intensity || !small_cq || !rw10 || !mult_out || nev == 0 || nsmall == 0)
        return iac::fail(IAC_ERR_INVALID_ARG, "null input");
    double max_weight = 0.0;
    for (uint32_t i = 0; i < nev; i++) {
        for (int k = 0; k < 9; k++) rw10[10 * (size_t)i + k] = rot9[9 * (size_t)i + k];
        rw10[10 * (size_t)i + 9] = (float)((double)mean_intensity[i] * (double)mean_intensity[i]);
        if (mean_intensity[i] > max_weight) max_weight = mean_intensity[i];
    }
    double t = 0;
    for (size_t i = 0; i < nsmall; i++) t += small_cq[i];
    const double mean_cq = t / nsmall;
    double m = 1e-12;
    for (; 1 > max_weight * max_weight * mean_cq * m;) m *= 10;
    uint64_t oom = 1;
    for (; oom < (uint64_t)nev;) oom *= 10;
    m = m / oom;
    m *= 1e10;
    if (m > 0.6) m = std::round(m);
    if (m < 1 && m > 0.1) m = 1;
    if (m < 1e-12) m = 1e-12;
    for (uint32_t i = 0; i < nev; i++) {
        float red = (float)m;
        if (m > 0.6) red = std::round((float)m);
        float w = rw10[10 * (size_t)i + 9];
        w *= red;
        w = std::floor((float)((double)w + 0.5));
        rw10[10 * (size_t)i + 9] = w / red;
    }
    *mult_out = m;
    return IAC_OK;
}

static int merge_launch(iac_ctx *ctx, const double *d_small_cq, bool per_event, const iac_mesh_shape *small_shape,
                        const float *d_rw10, uint32_t nev, int interp, double mult, uint64_t *d_big)
{
    IAC_TRY(iac::bind(ctx));
    IAC_TRY(check_small(small_shape));
    if (!d_small_cq || !d_rw10 || !d_big) return iac::fail(IAC_ERR_INVALID_ARG, "null pointer");
    if (interp != IAC_NEAREST && interp != IAC_LINEAR) return iac::fail(IAC_ERR_INVALID_ARG, "bad interpolation");
    if (nev == 0) return IAC_OK;
    const size_t nsmall = (size_t)small_shape->size_ab * small_shape->size_ab * small_shape->size_c;
    const uint32_t per_row = 16;
    const dim3 grid((unsigned)((nsmall + 255) / 256), (nev + per_row - 1) / per_row);
    IAC_TRY(iac::timer_begin(ctx));
    hipLaunchKernelGGL(merge_cq_kernel, grid, dim3(256), 0, ctx->stream, d_small_cq, per_event ? nsmall : (size_t)0,
                       (int)small_shape->size_ab,
                       (int)small_shape->size_c, (int)small_shape->k_a, (int)small_shape->k_b,
                       (int)small_shape->k_c, d_rw10, nev, per_row, interp, mult,
                       reinterpret_cast<unsigned long long *>(d_big), ctx->d_counters);
    IAC_HIP(hipGetLastError());
    IAC_TRY(iac::timer_end(ctx));
    return IAC_OK;
}

int iac_merge_cq(iac_ctx *ctx, const double *d_small_cq, const iac_mesh_shape *small_shape, const float *d_rw10,
                 uint32_t nev, int interp, double mult, uint64_t *d_big)
{
    return merge_launch(ctx, d_small_cq, false, small_shape, d_rw10, nev, interp, mult, d_big);
}

int iac_merge_cq_multi(iac_ctx *ctx, const double *d_small_meshes, const iac_mesh_shape *small_shape,
                       const float *d_rw10, uint32_t nev, int interp, double mult, uint64_t *d_big)
{
    return merge_launch(ctx, d_small_meshes, true, small_shape, d_rw10, nev, interp, mult, d_big);
}

int iac_merge_ac_cq(iac_ctx *ctx, const uint64_t *d_ac, const double *d_cq, size_t n, double *d_out)
{
    IAC_TRY(iac::bind(ctx));
    if (!d_ac || !d_cq || !d_out) return iac::fail(IAC_ERR_INVALID_ARG, "null pointer");
    if (n == 0) return IAC_OK;
    hipLaunchKernelGGL(merge_ac_cq_kernel, dim3(blocks_for(n)), dim3(256), 0, ctx->stream,
                       reinterpret_cast<const unsigned long long *>(d_ac), d_cq, n, d_out);
    IAC_HIP(hipGetLastError());
    return IAC_OK;
}

}  // extern "C"
