// iac_ac3d.hip -- 3D AC_uw pair accumulation (sparse photon lists) on gfx950.
//
// Work decomposition: each frame's n photon records are cut into 64-record
// tiles and the upper triangle of its n x n pair matrix into 64 x 64 tiles.
// A persistent grid walks the concatenated tile list of all frames (uneven
// frames cost only their own tiles).  Each 256-thread block stages the
// b-tile in LDS and every thread owns one a-record and 16 b-records.
//
// With nearest-neighbour second-level binning the ordered pair (b,a) lands on
// the point reflection of (a,b)'s voxel with the same value (iac_binning.h),
// so only unordered pairs are binned and the mirrored voxel receives the same
// add; the self pair lands on the centre once.  Linear modes bin both orders.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "iac_binning.h"
#include "iac_internal.h"
#include "iac_pairs.h"

#pragma clang fp contract(off)

namespace {

constexpr int kTile = 64;
constexpr int kBlock = 256;
constexpr int kBPerThread = kTile / (kBlock / kTile);  // 16

struct Rot9 {
    float m[9];
};

struct GlobalSink {
    unsigned long long *mesh;
    unsigned long long *dropped;
    int sab, sc;
    __device__ __forceinline__ void add(int fs, int ms, int ss, uint32_t v)
    {
        if ((unsigned)fs < (unsigned)sab && (unsigned)ms < (unsigned)sab && (unsigned)ss < (unsigned)sc) {
            if (v) atomicAdd(mesh + (size_t)fs + (size_t)ms * sab + (size_t)ss * sab * sab,
                             (unsigned long long)v);
        } else {
            atomicAdd(dropped, 1ull);
        }
    }
};

__device__ __forceinline__ uint64_t tri_tiles(uint64_t n)
{
    const uint64_t nt = (n + kTile - 1) / kTile;
    return nt * (nt + 1) / 2;
}

// Row-major upper-triangle tile l -> (ta, tb), ta <= tb.  Row r starts at
// S(r) = r*nt - r*(r-1)/2.
__device__ __forceinline__ void decode_tri(uint64_t l, uint32_t nt, uint32_t &ta, uint32_t &tb)
{
    const double b = 2.0 * nt + 1.0;
    double r = floor((b - sqrt(b * b - 8.0 * (double)l)) * 0.5);
    int ri = (int)r;
    if (ri < 0) ri = 0;
    if (ri > (int)nt - 1) ri = (int)nt - 1;
    auto S = [nt](int rr) -> uint64_t {
        return (uint64_t)rr * nt - (uint64_t)rr * (uint64_t)(rr - 1) / 2;
    };
    while (ri > 0 && S(ri) > l) ri--;
    while (ri + 1 < (int)nt && S(ri + 1) <= l) ri++;
    ta = (uint32_t)ri;
    tb = (uint32_t)(ri + (l - S(ri)));
}

// Photon records {k, count} of all frames.  A pixel index outside the k-map
// is counted in invalid[0] and becomes a zero-valued record at k = 0.
__global__ void gather_records_kernel(const float *__restrict__ kmap, uint32_t npix,
                                      const uint32_t *__restrict__ pix,
                                      const uint16_t *__restrict__ cnt, uint64_t n,
                                      float4 *__restrict__ rec, unsigned long long *invalid)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t p = pix[i];
        if (p < npix) {
            rec[i] = make_float4(kmap[3 * (size_t)p], kmap[3 * (size_t)p + 1], kmap[3 * (size_t)p + 2],
                                 (float)cnt[i]);
        } else {
            rec[i] = make_float4(0.f, 0.f, 0.f, 0.f);
            atomicAdd(invalid, 1ull);
        }
    }
}

// Exclusive scan of per-frame triangle tile counts into tile_off[0..nframes].
__global__ void __launch_bounds__(1024) tile_scan_kernel(const uint64_t *__restrict__ frame_off,
                                                         uint32_t nframes,
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
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        uint64_t v = t >= off ? part[t - off] : 0;
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

// kSym: nearest-neighbour second level (bin unordered pairs, add mirror).
template <bool kSym>
__global__ void __launch_bounds__(kBlock)
    ac3d_tiles_kernel(const float4 *__restrict__ rec, const uint64_t *__restrict__ frame_off,
                      const uint64_t *__restrict__ tile_off, uint32_t nframes, uint64_t single_n,
                      const float *__restrict__ rots, Rot9 single_rot, iac::BinParams P, unsigned long long *__restrict__ mesh,
                      unsigned long long *__restrict__ dropped)
{
    __shared__ float4 sb[kTile];
    const bool multi = frame_off != nullptr;
    const uint64_t total = multi ? tile_off[nframes] : tri_tiles(single_n);
    const int lane_a = threadIdx.x & (kTile - 1);
    const int bgrp = threadIdx.x / kTile;
    GlobalSink sink{mesh, dropped, P.sab, P.sc};
    const int S = P.sab;
    const size_t last = (size_t)S * S * P.sc - 1;  // mirror(idx) = last - idx (cubic mesh)

    for (uint64_t t = blockIdx.x; t < total; t += gridDim.x) {
        uint32_t f = 0;
        uint64_t base = 0, n = single_n, l = t;
        if (multi) {
            uint32_t lo = 0, hi = nframes;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (tile_off[mid] <= t) lo = mid; else hi = mid;
            }
            f = lo;
            base = frame_off[f];
            n = frame_off[f + 1] - base;
            l = t - tile_off[f];
        }
        const uint32_t nt = (uint32_t)((n + kTile - 1) / kTile);
        uint32_t ta, tb;
        decode_tri(l, nt, ta, tb);
        float R[9];
        const float *src = (multi && rots) ? rots + 9 * (size_t)f : nullptr;
#pragma unroll
        for (int i = 0; i < 9; i++) R[i] = src ? src[i] : single_rot.m[i];
        iac::BinParams Q = P;
        Q.rot_identity = (R[0] == 1.0f && R[1] == 0.0f && R[2] == 0.0f && R[3] == 0.0f &&
                          R[4] == 1.0f && R[5] == 0.0f && R[6] == 0.0f && R[7] == 0.0f &&
                          R[8] == 1.0f);

        __syncthreads();
        if (threadIdx.x < kTile) {
            const uint64_t b = (uint64_t)tb * kTile + threadIdx.x;
            sb[threadIdx.x] = b < n ? rec[base + b] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        __syncthreads();

        const uint64_t a = (uint64_t)ta * kTile + lane_a;
        if (a < n) {
            const float4 ra = rec[base + a];
            const int j0 = bgrp * kBPerThread;
            for (int jj = 0; jj < kBPerThread; jj++) {
                const int bl = j0 + jj;
                const uint64_t b = (uint64_t)tb * kTile + bl;
                if (b >= n) break;
                if (ta == tb && b < a) continue;
                const float4 rb = sb[bl];
                if (kSym) {
                    int dx, dy, dz;
                    if (!iac::bin_nn(ra.x, ra.y, ra.z, rb.x, rb.y, rb.z, Q, R, dx, dy, dz)) continue;
                    const uint32_t v = iac::pair_val(ra.w * rb.w);
                    const int fs = dx + Q.c0, ms = dy + Q.c1, ss = dz + Q.c2;
                    if ((unsigned)fs < (unsigned)S && (unsigned)ms < (unsigned)S &&
                        (unsigned)ss < (unsigned)Q.sc) {
                        if (v) {
                            const size_t idx = (size_t)fs + (size_t)ms * S + (size_t)ss * S * S;
                            atomicAdd(mesh + idx, (unsigned long long)v);
                            if (a != b) atomicAdd(mesh + (last - idx), (unsigned long long)v);
                        }
                    } else {
                        atomicAdd(dropped, a != b ? 2ull : 1ull);
                    }
                } else {
                    const float v = ra.w * rb.w;
                    iac::add_q_entry(sink, ra.x, ra.y, ra.z, rb.x, rb.y, rb.z, v, Q, R);
                    if (a != b) iac::add_q_entry(sink, rb.x, rb.y, rb.z, ra.x, ra.y, ra.z, v, Q, R);
                }
            }
        }
    }
}

int make_params(const iac_mesh_shape *s, int interp1, int interp2, int double_map,
                iac::BinParams *P)
{
    if (!s) return iac::fail(IAC_ERR_INVALID_ARG, "null mesh shape");
    if ((interp1 != IAC_NEAREST && interp1 != IAC_LINEAR) ||
        (interp2 != IAC_NEAREST && interp2 != IAC_LINEAR))
        return iac::fail(IAC_ERR_INVALID_ARG, "interpolation mode must be 0 or 1");
    if (s->size_ab % 2 == 0 || s->size_c % 2 == 0 || s->size_ab < 5)
        return iac::fail(IAC_ERR_EVEN_MESH, "Meshsize must be odd for each dimension");
    if (s->size_c != s->size_ab)
        return iac::fail(IAC_ERR_SHAPE, "AC_uw mesh must be cubic (CreateBigMeshForDetector)");
    P->cut_sq = iac::cut_sq(s->max_q);
    P->vox = (float)s->voxel_per_dq;
    P->c0 = (int)s->center[0];
    P->c1 = (int)s->center[1];
    P->c2 = (int)s->center[2];
    P->sab = (int)s->size_ab;
    P->sc = (int)s->size_c;
    P->interp1 = interp1;
    P->interp2 = interp2;
    P->double_map = double_map ? 1 : 0;
    P->rot_identity = 1;
    P->dval = 0;
    return IAC_OK;
}

bool symmetric_mode(const iac::BinParams &P)
{
    return P.interp2 == IAC_NEAREST && (!P.double_map || P.interp1 == IAC_NEAREST);
}

// The LDS engine's binning (identity rotation).  A double map with a nearest first
// level hands the second level an integral voxel: its linear split has one corner,
// of weight 1, the voxel itself (Atomic_Add_rotated_q_Entry, ACMesh.cpp:337-384,
// with Sep = 1), valued FloatToInt((double)v).  That is nearest-neighbour binning
// with the double fixed point, point symmetric like the nearest modes.
iac::BinParams engine_params(const iac::BinParams &P)
{
    iac::BinParams E = P;
    E.rot_identity = 1;
    if (P.double_map && P.interp1 == IAC_NEAREST && P.interp2 == IAC_LINEAR) {
        E.interp2 = IAC_NEAREST;
        E.dval = 1;
    }
    return E;
}

int launch_tiles(iac_ctx *ctx, const float4 *rec, const uint64_t *frame_off, const uint64_t *tile_off,
                 uint32_t nframes, uint64_t single_n, const float *rots, const Rot9 &single_rot,
                 const iac::BinParams &P, uint64_t *mesh)
{
    uint64_t blocks = (uint64_t)ctx->num_cus * 8;
    if (!frame_off) {
        const uint64_t nt = (single_n + kTile - 1) / kTile;
        blocks = std::min<uint64_t>(blocks, std::max<uint64_t>(1, nt * (nt + 1) / 2));
    }
    auto *m = reinterpret_cast<unsigned long long *>(mesh);
    IAC_TRY(iac::timer_begin(ctx));
    if (symmetric_mode(P))
        hipLaunchKernelGGL(ac3d_tiles_kernel<true>, dim3((unsigned)blocks), dim3(kBlock), 0, ctx->stream,
                           rec, frame_off, tile_off, nframes, single_n, rots, single_rot, P, m,
                           ctx->d_counters);
    else
        hipLaunchKernelGGL(ac3d_tiles_kernel<false>, dim3((unsigned)blocks), dim3(kBlock), 0,
                           ctx->stream, rec, frame_off, tile_off, nframes, single_n, rots, single_rot,
                           P, m, ctx->d_counters);
    IAC_HIP(hipGetLastError());
    IAC_TRY(iac::timer_end(ctx));
    return IAC_OK;
}

// Per-frame rotations on the LDS engine.  With a double map and a nearest second
// level a pair adds, per first-level voxel c (one, or the eight linear corners),
// FloatToInt(its value) at round(R_f c): c carries the pair, the rotation acts on c
// alone.  So the frames bin under the identity and each box voxel is rotated when
// it flushes.  Taken when no rotated voxel can leave the mesh (the tile kernel
// counts such updates pair by pair): a first-level voxel is at most
// r1 = Max_Q vox (1 + 1e-6) + sqrt(3)/2 long (sqrt(3) for linear corners), its
// rotated component i at most r1 |R_i|_2 plus the float rounding of the row product.
bool rotations_on_engine(const iac::BinParams &P, const std::vector<float> &h, uint32_t nframes)
{
    if (!(P.double_map && P.interp2 == IAC_NEAREST)) return false;
    const double r1 = std::sqrt((double)P.cut_sq) * (double)P.vox * (1.0 + 1e-6) +
                      (P.interp1 == IAC_NEAREST ? 0.8660254038 : 1.7320508076);
    const int lim = std::min({P.c0, P.c1, P.c2, P.sab - 1 - P.c0, P.sab - 1 - P.c1, P.sc - 1 - P.c2});
    for (uint32_t f = 0; f < nframes; f++) {
        for (int i = 0; i < 3; i++) {
            double l1 = 0, l2 = 0;
            for (int j = 0; j < 3; j++) {
                const double a = h[9 * (size_t)f + 3 * i + j];
                if (!std::isfinite(a)) return false;
                l1 += std::fabs(a);
                l2 += a * a;
            }
            if (!(r1 * std::sqrt(l2) + 4e-7 * r1 * l1 < lim + 0.5)) return false;
        }
    }
    return true;
}

// Frames of records through the LDS cell-box engine (the identity, or per-frame
// rotations applied at the box flush) or the tile kernel.
int batch_records(iac_ctx *ctx, const float *d_kmap, uint32_t npix, const float4 *rec,
                  const uint64_t *d_frame_off, uint32_t nframes, uint64_t n_total, const float *d_rots,
                  const iac::BinParams &P, uint64_t *d_mesh)
{
    bool engine_ok = n_total < (1ull << 32) && ctx->ac3d_kernel != 1;
    std::vector<float> hrot;
    if (engine_ok && d_rots) {
        hrot.resize(9 * (size_t)nframes);
        IAC_HIP(hipMemcpyAsync(hrot.data(), d_rots, hrot.size() * sizeof(float), hipMemcpyDeviceToHost,
                               ctx->stream));
        IAC_HIP(hipStreamSynchronize(ctx->stream));
        engine_ok = rotations_on_engine(P, hrot, nframes);
        bool all_identity = true;
        for (uint32_t f = 0; f < nframes && all_identity; f++)
            for (int i = 0; i < 9; i++)
                if (hrot[9 * (size_t)f + i] != ((i % 4) == 0 ? 1.0f : 0.0f)) all_identity = false;
        if (engine_ok && all_identity) {
            d_rots = nullptr;
            hrot.clear();
        }
    }
    if (ctx->ac3d_kernel == 2 && !engine_ok)
        return iac::fail(IAC_ERR_PRECONDITION,
                         "the LDS engine needs < 2^32 photons and, for per-frame rotations, a double map with "
                         "nearest-neighbour levels whose rotated voxels stay in the mesh");
    if (engine_ok) {
        iac::PairRun R;
        std::memset(&R.spec, 0, sizeof(R.spec));
        R.spec.kind = iac::kKindAc3d;
        for (int i = 0; i < 9; i++) R.spec.R[i] = (i % 4) == 0 ? 1.0f : 0.0f;
        R.spec.P = engine_params(P);
        R.ordered = !symmetric_mode(R.spec.P);
        R.d_kmap = d_kmap;
        R.npix = npix;
        R.rec = rec;
        R.d_frame_off = d_frame_off;
        R.nframes = nframes;
        R.n_total = n_total;
        const uint32_t goff[2] = {0, nframes};
        R.h_group_off = goff;
        R.ngroups = 1;
        R.d_mesh = d_mesh;
        if (d_rots) {
            R.d_rots = d_rots;
            R.h_rots = hrot.data();
        }
        return iac::pairs_run(ctx, R);
    }
    void *toff = nullptr;
    IAC_TRY(iac::scratch(ctx, 1, (size_t)(nframes + 1) * sizeof(uint64_t), &toff));
    hipLaunchKernelGGL(tile_scan_kernel, dim3(1), dim3(1024), 0, ctx->stream, d_frame_off, nframes,
                       static_cast<uint64_t *>(toff));
    IAC_HIP(hipGetLastError());
    const Rot9 ident{{1, 0, 0, 0, 1, 0, 0, 0, 1}};
    return launch_tiles(ctx, rec, d_frame_off, static_cast<const uint64_t *>(toff), nframes, 0, d_rots, ident, P,
                        d_mesh);
}

__global__ void zero_u64_kernel(unsigned long long *__restrict__ p, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

}  // namespace

// A prepared frame batch on the LDS engine (identity rotation): the plan, chunks
// and buffers are made once; a launch enqueues gather + sort + pairs (+ mirror)
// without host synchronisation or allocation, directly or by replaying a graph.
struct iac_ac3d_batch {
    iac_ctx *ctx = nullptr;
    const float *d_kmap = nullptr;
    uint32_t npix = 0;
    const uint32_t *d_pix = nullptr;
    const uint16_t *d_cnt = nullptr;
    uint64_t n_total = 0;
    size_t nvox = 0;
    float4 *rec = nullptr;
    iac::PairBatch *pb = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t *graph_mesh = nullptr;
    int graph_zero = -1;
};

namespace {

int batch_enqueue(iac_ac3d_batch *b, hipStream_t st, uint64_t *d_mesh, int zero_mesh)
{
    if (zero_mesh) {
        const unsigned zb = (unsigned)std::max<size_t>(1, std::min<size_t>(4096, (b->nvox + 255) / 256));
        hipLaunchKernelGGL(zero_u64_kernel, dim3(zb), dim3(256), 0, st,
                           reinterpret_cast<unsigned long long *>(d_mesh), b->nvox);
    }
    const unsigned gblocks = (unsigned)std::min<uint64_t>((b->n_total + 255) / 256, 65535);
    hipLaunchKernelGGL(gather_records_kernel, dim3(gblocks), dim3(256), 0, st, b->d_kmap, b->npix, b->d_pix,
                       b->d_cnt, b->n_total, b->rec, b->ctx->d_counters + 1);
    IAC_HIP(hipGetLastError());
    return iac::pairs_enqueue(b->pb, st, d_mesh);
}

void drop_graph(iac_ac3d_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_mesh = nullptr;
    b->graph_zero = -1;
}

}  // namespace

extern "C" {

int iac_ac3d_batch_create(iac_ctx *ctx, const float *d_kmap, uint32_t npix, const uint32_t *d_pix,
                          const uint16_t *d_cnt, const uint64_t *d_frame_off, uint32_t nframes, uint64_t n_total,
                          const iac_mesh_shape *shape, int interp1, int interp2, int double_map,
                          iac_ac3d_batch **out)
{
    IAC_TRY(iac::bind(ctx));
    if (!out || !d_frame_off || !d_kmap || npix == 0 || !d_pix || !d_cnt)
        return iac::fail(IAC_ERR_INVALID_ARG, "null input of a batch");
    *out = nullptr;
    if (nframes == 0 || n_total == 0) return iac::fail(IAC_ERR_INVALID_ARG, "a batch needs frames and photons");
    if (n_total >= (1ull << 32)) return iac::fail(IAC_ERR_LIMIT, "a batch holds at most 2^32 - 1 photons");
    iac::BinParams P;
    IAC_TRY(make_params(shape, interp1, interp2, double_map, &P));
    auto *b = new iac_ac3d_batch();
    b->ctx = ctx;
    b->d_kmap = d_kmap;
    b->npix = npix;
    b->d_pix = d_pix;
    b->d_cnt = d_cnt;
    b->n_total = n_total;
    b->nvox = (size_t)P.sab * P.sab * P.sc;
    hipError_t e = hipMalloc(&b->rec, n_total * sizeof(float4));
    if (e != hipSuccess) {
        delete b;
        return iac::hip_fail(e, "iac_ac3d_batch_create allocation", __FILE__, __LINE__);
    }
    iac::PairRun R;
    std::memset(&R.spec, 0, sizeof(R.spec));
    R.spec.kind = iac::kKindAc3d;
    for (int i = 0; i < 9; i++) R.spec.R[i] = (i % 4) == 0 ? 1.0f : 0.0f;
    R.spec.P = engine_params(P);
    R.ordered = !symmetric_mode(R.spec.P);
    R.d_kmap = d_kmap;
    R.npix = npix;
    R.rec = b->rec;
    R.d_frame_off = d_frame_off;
    R.nframes = nframes;
    R.n_total = n_total;
    const uint32_t goff[2] = {0, nframes};
    R.h_group_off = goff;
    R.ngroups = 1;
    R.d_mesh = nullptr;
    const int rc = iac::pairs_prepare(ctx, R, &b->pb);
    if (rc != IAC_OK) {
        (void)hipFree(b->rec);
        delete b;
        return rc;
    }
    *out = b;
    return IAC_OK;
}

int iac_ac3d_batch_launch(iac_ctx *ctx, iac_ac3d_batch *batch, uint64_t *d_mesh, int zero_mesh)
{
    IAC_TRY(iac::bind(ctx));
    if (!batch || batch->ctx != ctx || !d_mesh) return iac::fail(IAC_ERR_INVALID_ARG, "bad batch launch");
    // timing brackets the pair kernel of an eager launch, the whole graph of a replay
    if (batch->exec && batch->graph_mesh == d_mesh && batch->graph_zero == (zero_mesh ? 1 : 0)) {
        IAC_TRY(iac::timer_begin(ctx));
        IAC_HIP(hipGraphLaunch(batch->exec, ctx->stream));
        IAC_TRY(iac::timer_end(ctx));
    } else {
        IAC_TRY(batch_enqueue(batch, ctx->stream, d_mesh, zero_mesh));
    }
    return IAC_OK;
}

int iac_ac3d_batch_capture(iac_ctx *ctx, iac_ac3d_batch *batch, uint64_t *d_mesh, int zero_mesh)
{
    IAC_TRY(iac::bind(ctx));
    if (!batch || batch->ctx != ctx || !d_mesh) return iac::fail(IAC_ERR_INVALID_ARG, "bad batch capture");
    drop_graph(batch);
    // capture on a private stream of default priority (the null stream cannot be captured)
    hipStream_t cs = nullptr;
    IAC_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    int rc = IAC_OK;
    if (e == hipSuccess) rc = batch_enqueue(batch, cs, d_mesh, zero_mesh);
    hipGraph_t g = nullptr;
    const hipError_t e2 = e == hipSuccess ? hipStreamEndCapture(cs, &g) : e;
    (void)hipStreamDestroy(cs);
    if (rc != IAC_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e2 != hipSuccess) return iac::hip_fail(e2, "batch stream capture", __FILE__, __LINE__);
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return iac::hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
    }
    batch->graph = g;
    batch->exec = x;
    batch->graph_mesh = d_mesh;
    batch->graph_zero = zero_mesh ? 1 : 0;
    return IAC_OK;
}

int iac_ac3d_batch_destroy(iac_ac3d_batch *batch)
{
    if (!batch) return IAC_OK;
    (void)hipSetDevice(batch->ctx->device);
    (void)hipStreamSynchronize(batch->ctx->stream);
    drop_graph(batch);
    iac::pairs_batch_free(batch->pb);
    if (batch->rec) (void)hipFree(batch->rec);
    delete batch;
    return IAC_OK;
}

int iac_ac3d_sparse(iac_ctx *ctx, const float *d_hitlist4, uint32_t n, const float rot[9],
                    const iac_mesh_shape *shape, int interp1, int interp2, int double_map,
                    uint64_t *d_mesh)
{
    IAC_TRY(iac::bind(ctx));
    if (!d_mesh || !rot) return iac::fail(IAC_ERR_INVALID_ARG, "null mesh or rotation");
    iac::BinParams P;
    IAC_TRY(make_params(shape, interp1, interp2, double_map, &P));
    if (n == 0) return IAC_OK;
    if (!d_hitlist4) return iac::fail(IAC_ERR_PRECONDITION, "no sparse hit list");
    Rot9 R;
    for (int i = 0; i < 9; i++) R.m[i] = rot[i];
    return launch_tiles(ctx, reinterpret_cast<const float4 *>(d_hitlist4), nullptr, nullptr, 0, n,
                        nullptr, R, P, d_mesh);
}

// ---- autotuner ---------------------------------------------------------------------

}  // extern "C"

namespace {

struct TunePoint {
    int pair_mode, cell_vox, box_cap, frame_chunk;
};
// Candidates around the measured defaults: nearest-neighbour binning (one LDS add per pair,
// symmetric: unordered pairs and a mirror pass) and linear binning (eight corner adds per
// ordered pair, 4096-voxel boxes); walks 1 = lanes on consecutive pairs, 2 = a pair range
// per thread.
const TunePoint kTuneNearest[] = {{1, 6, 1920, 32}, {1, 5, 1920, 32}, {1, 4, 1920, 32}, {1, 6, 1920, 64},
                                  {1, 8, 4096, 32}, {1, 4, 1024, 32}, {2, 4, 1920, 32}};
const TunePoint kTuneLinear[] = {{1, 5, 4096, 32}, {1, 4, 4096, 32}, {1, 6, 4096, 32},
                                 {1, 4, 1920, 32}, {2, 4, 4096, 32}, {1, 5, 4096, 64}};

void apply_point(iac_ctx *ctx, const iac_tune_result &r)
{
    ctx->pair_mode = r.pair_mode;
    ctx->cell_vox = r.cell_vox;
    ctx->cell_vox_v = r.cell_vox_v;
    ctx->box_cap = r.box_cap;
    ctx->task_region = r.task_region;
    ctx->frame_chunk = (uint32_t)r.frame_chunk;
}

bool same_shape(const iac_mesh_shape &a, const iac_mesh_shape &b) { return std::memcmp(&a, &b, sizeof(a)) == 0; }

}  // namespace

extern "C" {

int iac_ac3d_tune(iac_ctx *ctx, const float *d_kmap, uint32_t npix, const uint32_t *d_pix, const uint16_t *d_cnt,
                  const uint64_t *d_frame_off, uint32_t nframes, const iac_mesh_shape *shape, int interp1,
                  int interp2, int double_map, uint32_t prefix_frames, int force_retune, iac_tune_result *best,
                  iac_tune_result *all, uint32_t max_all, uint32_t *n_all)
{
    IAC_TRY(iac::bind(ctx));
    if (!d_kmap || npix == 0 || !d_pix || !d_cnt || !d_frame_off || !best)
        return iac::fail(IAC_ERR_INVALID_ARG, "null tuner argument");
    if (nframes == 0) return iac::fail(IAC_ERR_INVALID_ARG, "the tuner needs frames");
    iac::BinParams P;
    IAC_TRY(make_params(shape, interp1, interp2, double_map, &P));
    if (n_all) *n_all = 0;
    for (const auto &e : ctx->tune_cache)
        if (!force_retune && e.kmap == d_kmap && e.npix == npix && same_shape(e.shape, *shape) &&
            e.interp1 == interp1 && e.interp2 == interp2 && e.double_map == (double_map ? 1 : 0)) {
            *best = e.best;
            best->ms = -1.0f;
            apply_point(ctx, e.best);
            return IAC_OK;
        }
    const uint32_t nf = std::min(nframes, prefix_frames ? prefix_frames : 1000u);
    uint64_t n = 0;
    IAC_HIP(hipMemcpyAsync(&n, d_frame_off + nf, sizeof(n), hipMemcpyDeviceToHost, ctx->stream));
    IAC_HIP(hipStreamSynchronize(ctx->stream));
    if (n == 0) return iac::fail(IAC_ERR_PRECONDITION, "no photons in the tuning prefix");
    if (n >= (1ull << 32)) return iac::fail(IAC_ERR_LIMIT, "tuning prefix above 2^32 - 1 photons");
    const bool linear = !symmetric_mode(engine_params(P));
    const TunePoint *pts = linear ? kTuneLinear : kTuneNearest;
    const size_t npts = linear ? sizeof(kTuneLinear) / sizeof(kTuneLinear[0])
                               : sizeof(kTuneNearest) / sizeof(kTuneNearest[0]);
    const size_t nvox = (size_t)P.sab * P.sab * P.sc;
    uint64_t *mesh = nullptr;
    IAC_HIP(hipMalloc(&mesh, nvox * sizeof(uint64_t)));
    // the caller's timing switch and kernel choice are restored on every exit
    const bool timing = ctx->timing;
    const int kernel = ctx->ac3d_kernel;
    ctx->ac3d_kernel = 2;
    int rc = IAC_OK;
    iac_tune_result win{};
    win.ms = -1.0f;
    uint32_t count = 0;
    for (size_t k = 0; k < npts && rc == IAC_OK; k++) {
        iac_tune_result r{pts[k].pair_mode, pts[k].cell_vox, pts[k].cell_vox, pts[k].box_cap, 0, pts[k].frame_chunk,
                          0.0f};
        apply_point(ctx, r);
        float t_best = -1.0f;
        for (int rep = 0; rep < 3 && rc == IAC_OK; rep++) {  // the first run plans
            ctx->timing = rep > 0;
            ctx->ev_used = 0;
            rc = iac_ac3d_sparse_batch(ctx, d_kmap, npix, d_pix, d_cnt, d_frame_off, nf, n, nullptr, shape, interp1,
                                       interp2, double_map, mesh);
            if (rc == IAC_OK && rep > 0) {
                double ms = 0;
                uint64_t launches = 0;
                rc = iac_ctx_timing(ctx, &ms, &launches, 1);
                if (rc == IAC_OK && (t_best < 0 || ms < t_best)) t_best = (float)ms;
            }
        }
        if (rc != IAC_OK) break;
        r.ms = t_best;
        if (all && count < max_all) all[count] = r;
        count++;
        if (win.ms < 0 || r.ms < win.ms) win = r;
    }
    ctx->timing = timing;
    ctx->ev_used = 0;  // pending kernel timings of the caller are discarded
    ctx->ac3d_kernel = kernel;
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(mesh);
    if (rc != IAC_OK) return rc;
    if (n_all) *n_all = std::min(count, max_all);
    apply_point(ctx, win);
    *best = win;
    iac_ctx::TuneEntry e{d_kmap, npix, *shape, interp1, interp2, double_map ? 1 : 0, win};
    ctx->tune_cache.push_back(e);
    return IAC_OK;
}

// ---- frame queue: the per-event loop at batch speed --------------------------------

}  // extern "C"

// Two slots of frames: while one slot's uploads and engine run are in flight, pushes fill
// the other.  A slot's uploads run on the queue's copy stream after the last engine run
// that read the slot (event done), and its engine run on the context's stream after them
// (event copied), so the uploads of one flush overlap the previous flush's pair pass.
struct QueueSlot {
    // device: records | offsets | rotations | tile offsets (tile kernel fallback)
    float4 *d_rec = nullptr;
    uint64_t *d_off = nullptr;
    float *d_rot = nullptr;
    uint64_t *d_toff = nullptr;
    // pinned host staging of the same (records only for host pushes)
    float4 *h_rec = nullptr;
    uint64_t *h_off = nullptr;
    float *h_rot = nullptr;
    hipEvent_t copied = nullptr;  // this slot's last uploads are done
    hipEvent_t done = nullptr;    // the last engine run that read this slot is done
    bool uploading = false;       // an upload from the host staging may still be in flight
};

struct iac_ac3d_queue {
    iac_ctx *ctx = nullptr;
    const float *d_kmap = nullptr;
    uint32_t npix = 0;
    iac::BinParams P{};
    uint64_t *d_mesh = nullptr;
    uint64_t cap_rec = 0;
    uint32_t cap_frames = 0;
    void *dmem = nullptr;  // both slots' device buffers
    void *hmem = nullptr;  // both slots' pinned staging
    QueueSlot slot[2];
    int cur = 0;                     // the slot pushes fill
    hipStream_t copy_st = nullptr;   // uploads (default priority)
    // the current slot's buffers
    float4 *d_rec = nullptr;
    uint64_t *d_off = nullptr;
    float *d_rot = nullptr;
    uint64_t *d_toff = nullptr;
    float4 *h_rec = nullptr;
    uint64_t *h_off = nullptr;
    float *h_rot = nullptr;
    // pending frames
    uint32_t nframes = 0;
    uint64_t nrec = 0;
    bool any_rot = false;
    int source = 0;  // 0 none yet, 1 host records, 2 device records
    iac::PairQueue *pq = nullptr;
};

namespace {

// Waits for both streams, then frees everything the queue holds.
void queue_release(iac_ac3d_queue *q)
{
    (void)hipStreamSynchronize(q->ctx->stream);
    if (q->copy_st) (void)hipStreamSynchronize(q->copy_st);
    iac::pairs_queue_free(q->pq);
    if (q->dmem) (void)hipFree(q->dmem);
    if (q->hmem) (void)hipHostFree(q->hmem);
    for (QueueSlot &S : q->slot) {
        if (S.copied) (void)hipEventDestroy(S.copied);
        if (S.done) (void)hipEventDestroy(S.done);
    }
    if (q->copy_st) (void)hipStreamDestroy(q->copy_st);
    delete q;
}

void queue_point(iac_ac3d_queue *q)
{
    const QueueSlot &S = q->slot[q->cur];
    q->d_rec = S.d_rec;
    q->d_off = S.d_off;
    q->d_rot = S.d_rot;
    q->d_toff = S.d_toff;
    q->h_rec = S.h_rec;
    q->h_off = S.h_off;
    q->h_rot = S.h_rot;
}

// Waits (once per flush of a slot) until the slot's last uploads have read its staging.
int queue_claim_staging(iac_ac3d_queue *q)
{
    QueueSlot &S = q->slot[q->cur];
    if (S.uploading) {
        IAC_HIP(hipEventSynchronize(S.copied));
        S.uploading = false;
    }
    return IAC_OK;
}

// The pair pass over one flushed slot (engine, or the tile kernel where the engine does not apply).
int queue_run(iac_ac3d_queue *q, QueueSlot &S, uint32_t n, uint64_t nrec, bool rot)
{
    iac_ctx *ctx = q->ctx;
    hipStream_t st = ctx->stream;
    // the LDS engine takes the identity, or per-frame rotations whose voxels stay in the mesh
    std::vector<float> hrot;
    bool engine = q->pq != nullptr;
    if (engine && rot) {
        hrot.assign(S.h_rot, S.h_rot + (size_t)n * 9);
        engine = rotations_on_engine(q->P, hrot, n);
    }
    if (engine) {
        iac::PairRun R;
        std::memset(&R.spec, 0, sizeof(R.spec));
        R.rec = S.d_rec;
        R.d_frame_off = S.d_off;
        R.nframes = n;
        R.n_total = nrec;
        R.d_mesh = q->d_mesh;
        if (rot) {
            R.d_rots = S.d_rot;
            R.h_rots = hrot.data();
        }
        return iac::pairs_queue_run(q->pq, R, S.h_off);
    }
    hipLaunchKernelGGL(tile_scan_kernel, dim3(1), dim3(1024), 0, st, S.d_off, n, S.d_toff);
    IAC_HIP(hipGetLastError());
    const Rot9 ident{{1, 0, 0, 0, 1, 0, 0, 0, 1}};
    return launch_tiles(ctx, S.d_rec, S.d_off, S.d_toff, n, 0, rot ? S.d_rot : nullptr, ident, q->P, q->d_mesh);
}

int queue_flush(iac_ac3d_queue *q)
{
    if (q->nframes == 0) return IAC_OK;
    iac_ctx *ctx = q->ctx;
    hipStream_t st = ctx->stream;
    const uint32_t n = q->nframes;
    const uint64_t nrec = q->nrec;
    QueueSlot &S = q->slot[q->cur];
    // uploads on the copy stream once the slot's previous engine run is done, the pair pass
    // on the context's stream once they are (device records came in on that stream already)
    IAC_HIP(hipStreamWaitEvent(q->copy_st, S.done, 0));
    if (q->source == 1 && nrec)
        IAC_HIP(hipMemcpyAsync(S.d_rec, S.h_rec, nrec * sizeof(float4), hipMemcpyHostToDevice, q->copy_st));
    IAC_HIP(hipMemcpyAsync(S.d_off, S.h_off, (n + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, q->copy_st));
    bool rot = q->any_rot;
    if (rot)
        IAC_HIP(hipMemcpyAsync(S.d_rot, S.h_rot, (size_t)n * 9 * sizeof(float), hipMemcpyHostToDevice, q->copy_st));
    IAC_HIP(hipEventRecord(S.copied, q->copy_st));
    IAC_HIP(hipStreamWaitEvent(st, S.copied, 0));
    S.uploading = true;
    q->nframes = 0;
    q->nrec = 0;
    q->any_rot = false;
    q->source = 0;
    // pushes go to the other slot from here on; this slot's run is recorded as done below
    q->cur ^= 1;
    queue_point(q);
    int rc = IAC_OK;
    if (nrec) rc = queue_run(q, S, n, nrec, rot);
    const hipError_t e = hipEventRecord(S.done, st);
    if (rc == IAC_OK && e != hipSuccess) rc = iac::hip_fail(e, "queue flush", __FILE__, __LINE__);
    return rc;
}

// Room for one more frame of n records (flushing first when full); writes its rotation.
int queue_begin_frame(iac_ac3d_queue *q, uint32_t n, const float rot[9], int source)
{
    if (n > q->cap_rec) return iac::fail(IAC_ERR_LIMIT, "frame has more records than the queue holds");
    if (q->nframes == q->cap_frames || q->nrec + n > q->cap_rec || (q->source && q->source != source))
        IAC_TRY(queue_flush(q));
    IAC_TRY(queue_claim_staging(q));
    const uint32_t f = q->nframes;
    if (f == 0) q->h_off[0] = 0;
    float *r = q->h_rot + 9 * (size_t)f;
    bool ident = true;
    for (int i = 0; i < 9; i++) {
        r[i] = rot[i];
        ident = ident && rot[i] == ((i % 4) == 0 ? 1.0f : 0.0f);
    }
    q->any_rot = q->any_rot || !ident;
    q->source = source;
    return IAC_OK;
}

void queue_end_frame(iac_ac3d_queue *q, uint32_t n)
{
    q->nrec += n;
    q->nframes++;
    q->h_off[q->nframes] = q->nrec;
}

}  // namespace

extern "C" {

int iac_ac3d_queue_create(iac_ctx *ctx, const float *d_kmap, uint32_t npix, const iac_mesh_shape *shape,
                          int interp1, int interp2, int double_map, uint64_t *d_mesh, uint64_t max_records,
                          uint32_t max_frames, iac_ac3d_queue **out)
{
    IAC_TRY(iac::bind(ctx));
    if (!out || !d_kmap || npix == 0 || !d_mesh) return iac::fail(IAC_ERR_INVALID_ARG, "null queue argument");
    *out = nullptr;
    if (max_records == 0) max_records = 1ull << 22;
    if (max_frames == 0) max_frames = 4096;
    if (max_records >= (1ull << 32)) return iac::fail(IAC_ERR_LIMIT, "a queue holds at most 2^32 - 1 records");
    if (max_frames > (1u << 24)) return iac::fail(IAC_ERR_LIMIT, "a queue holds at most 2^24 frames");
    iac::BinParams P;
    IAC_TRY(make_params(shape, interp1, interp2, double_map, &P));
    auto *q = new iac_ac3d_queue();
    q->ctx = ctx;
    q->d_kmap = d_kmap;
    q->npix = npix;
    q->P = P;
    q->d_mesh = d_mesh;
    q->cap_rec = max_records;
    q->cap_frames = max_frames;
    auto up = [](size_t x) { return (x + 255) / 256 * 256; };
    const size_t brec = up(max_records * sizeof(float4)), boff = up(((size_t)max_frames + 1) * sizeof(uint64_t)),
                 brot = up((size_t)max_frames * 9 * sizeof(float));
    const size_t dslot = brec + 2 * boff + brot, hslot = brec + boff + brot;
    hipError_t e = hipMalloc(&q->dmem, 2 * dslot);
    if (e == hipSuccess) e = hipHostMalloc(&q->hmem, 2 * hslot, hipHostMallocDefault);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&q->copy_st, hipStreamNonBlocking);
    for (int k = 0; k < 2 && e == hipSuccess; k++) {
        e = hipEventCreateWithFlags(&q->slot[k].copied, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&q->slot[k].done, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventRecord(q->slot[k].done, ctx->stream);
    }
    int rc = IAC_OK;
    if (e != hipSuccess) rc = iac::hip_fail(e, "iac_ac3d_queue_create allocation", __FILE__, __LINE__);
    if (rc == IAC_OK) {
        for (int k = 0; k < 2; k++) {
            QueueSlot &S = q->slot[k];
            char *d = static_cast<char *>(q->dmem) + k * dslot;
            S.d_rec = reinterpret_cast<float4 *>(d);
            S.d_off = reinterpret_cast<uint64_t *>(d + brec);
            S.d_rot = reinterpret_cast<float *>(d + brec + boff);
            S.d_toff = reinterpret_cast<uint64_t *>(d + brec + boff + brot);
            char *h = static_cast<char *>(q->hmem) + k * hslot;
            S.h_rec = reinterpret_cast<float4 *>(h);
            S.h_off = reinterpret_cast<uint64_t *>(h + brec);
            S.h_rot = reinterpret_cast<float *>(h + brec + boff);
            S.h_off[0] = 0;
        }
        queue_point(q);
        // the LDS engine, unless the context asks for the tile kernel
        if (ctx->ac3d_kernel != 1) {
            iac::PairRun R;
            std::memset(&R.spec, 0, sizeof(R.spec));
            R.spec.kind = iac::kKindAc3d;
            for (int i = 0; i < 9; i++) R.spec.R[i] = (i % 4) == 0 ? 1.0f : 0.0f;
            R.spec.P = engine_params(P);
            R.ordered = !symmetric_mode(R.spec.P);
            R.d_kmap = d_kmap;
            R.npix = npix;
            rc = iac::pairs_queue_create(ctx, R, max_records, max_frames, &q->pq);
        }
    }
    if (rc != IAC_OK) {
        queue_release(q);
        return rc;
    }
    *out = q;
    return IAC_OK;
}

int iac_ac3d_queue_push(iac_ac3d_queue *q, const float *h_hitlist4, uint32_t n, const float rot[9])
{
    if (!q) return iac::fail(IAC_ERR_INVALID_ARG, "null queue");
    IAC_TRY(iac::bind(q->ctx));
    if (!rot || (n && !h_hitlist4)) return iac::fail(IAC_ERR_INVALID_ARG, "null records or rotation");
    IAC_TRY(queue_begin_frame(q, n, rot, 1));
    if (n) std::memcpy(q->h_rec + q->nrec, h_hitlist4, (size_t)n * sizeof(float4));
    queue_end_frame(q, n);
    return IAC_OK;
}

int iac_ac3d_queue_push_device(iac_ac3d_queue *q, const float *d_hitlist4, uint32_t n, const float rot[9])
{
    if (!q) return iac::fail(IAC_ERR_INVALID_ARG, "null queue");
    IAC_TRY(iac::bind(q->ctx));
    if (!rot || (n && !d_hitlist4)) return iac::fail(IAC_ERR_INVALID_ARG, "null records or rotation");
    IAC_TRY(queue_begin_frame(q, n, rot, 2));
    if (n)
        IAC_HIP(hipMemcpyAsync(q->d_rec + q->nrec, d_hitlist4, (size_t)n * sizeof(float4), hipMemcpyDeviceToDevice,
                               q->ctx->stream));
    queue_end_frame(q, n);
    return IAC_OK;
}

int iac_ac3d_queue_flush(iac_ac3d_queue *q)
{
    if (!q) return iac::fail(IAC_ERR_INVALID_ARG, "null queue");
    IAC_TRY(iac::bind(q->ctx));
    return queue_flush(q);
}

int iac_ac3d_queue_pending(const iac_ac3d_queue *q, uint32_t *frames, uint64_t *records)
{
    if (!q) return iac::fail(IAC_ERR_INVALID_ARG, "null queue");
    if (frames) *frames = q->nframes;
    if (records) *records = q->nrec;
    return IAC_OK;
}

int iac_ac3d_queue_destroy(iac_ac3d_queue *q)
{
    if (!q) return IAC_OK;
    int rc = iac::bind(q->ctx);
    if (rc == IAC_OK) rc = queue_flush(q);
    queue_release(q);
    return rc;
}

int iac_ac3d_sparse_batch(iac_ctx *ctx, const float *d_kmap, uint32_t npix, const uint32_t *d_pix,
                          const uint16_t *d_cnt, const uint64_t *d_frame_off, uint32_t nframes,
                          uint64_t n_total, const float *d_rots, const iac_mesh_shape *shape,
                          int interp1, int interp2, int double_map, uint64_t *d_mesh)
{
    IAC_TRY(iac::bind(ctx));
    if (!d_mesh || !d_frame_off || !d_kmap || npix == 0)
        return iac::fail(IAC_ERR_INVALID_ARG, "null mesh, offsets or k-map");
    iac::BinParams P;
    IAC_TRY(make_params(shape, interp1, interp2, double_map, &P));
    if (nframes == 0 || n_total == 0) return IAC_OK;
    if (!d_pix || !d_cnt) return iac::fail(IAC_ERR_PRECONDITION, "no photon lists");
    void *rec = nullptr;
    IAC_TRY(iac::scratch(ctx, 0, n_total * sizeof(float4), &rec));
    const unsigned gblocks = (unsigned)std::min<uint64_t>((n_total + 255) / 256, 65535);
    hipLaunchKernelGGL(gather_records_kernel, dim3(gblocks), dim3(256), 0, ctx->stream, d_kmap, npix,
                       d_pix, d_cnt, n_total, static_cast<float4 *>(rec), ctx->d_counters + 1);
    IAC_HIP(hipGetLastError());
    return batch_records(ctx, d_kmap, npix, static_cast<const float4 *>(rec), d_frame_off, nframes, n_total,
                         d_rots, P, d_mesh);
}

int iac_ac3d_records_batch(iac_ctx *ctx, const float *d_kmap, uint32_t npix, const float *d_rec4,
                           const uint64_t *d_frame_off, uint32_t nframes, uint64_t n_total,
                           const float *d_rots, const iac_mesh_shape *shape, int interp1, int interp2,
                           int double_map, uint64_t *d_mesh)
{
    IAC_TRY(iac::bind(ctx));
    if (!d_mesh || !d_frame_off || !d_kmap || npix == 0)
        return iac::fail(IAC_ERR_INVALID_ARG, "null mesh, offsets or k-map");
    iac::BinParams P;
    IAC_TRY(make_params(shape, interp1, interp2, double_map, &P));
    if (nframes == 0 || n_total == 0) return IAC_OK;
    if (!d_rec4) return iac::fail(IAC_ERR_PRECONDITION, "no hit-list records");
    return batch_records(ctx, d_kmap, npix, reinterpret_cast<const float4 *>(d_rec4), d_frame_off, nframes,
                         n_total, d_rots, P, d_mesh);
}

}  // extern "C"
