// iac_ac3d_lds.hip -- LDS-privatised sparse 3D AC_uw for nearest-neighbour
// second-level binning and frames sharing the identity rotation.
//
// Pipeline of one batch call:
//   1. k-extents of the k-map -> cell grid over the two widest k axes
//      (cell side `cell_vox` voxels); per-cell k extents; host plan (cached).
//   2. cell_sort: per frame, a counting sort of the photon records by cell
//      (one block per frame) -> cell-sorted records + cell offsets.
//   3. pairs: one block per (cell offset delta, frame chunk).  Every unordered
//      photon pair {a,b} with cell(a) - cell(b) = delta in the positive half
//      plane (plus distinct pairs inside one cell for delta = 0) is binned
//      exactly once.  All such pairs land in a small voxel box fixed by the
//      cell extents, privatised in LDS across the chunk's frames and flushed
//      to the half-mesh H with one atomic per nonzero voxel.  A pair that
//      falls outside the box (never for a flat detector) adds to H directly,
//      so the result does not depend on the box being tight.
//   4. self pairs (q = 0 -> centre voxel) summed over all records.
//   5. symmetrize: mesh[v] += H[v] + H[mirror v]; mesh[centre] += self.
// Step 5 is exact because the (b,a) pair of every (a,b) lands on the point
// reflection with the same value (iac_binning.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "iac_binning.h"
#include "iac_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr int kBlock = 256;
constexpr int kItemsPerThread = 4;
constexpr int kGroup = kBlock * kItemsPerThread;  // items per scan group
constexpr int kBoxMax = 4096;                      // LDS box capacity (uint64 voxels, 32 KiB)
constexpr int kMaxCells = 4096;                    // NC_u * NC_v

struct CellGeom {
    int au, av;          // k axes spanned by the cell grid
    float u0, v0;        // grid origin
    float inv_w;         // 1 / cell width (k units)
    int ncu, ncv;        // cells per axis
};

struct Task {
    int du, dv;          // cell offset delta (u, v)
    int iu0, iv0;        // first cell i whose partner i - delta is on the grid
    int nu, nv;          // extent of that cell range
    int bx0, by0, bz0;   // LDS box origin (voxel indices)
    int bx, by, bz;      // LDS box extent; bx == 0 -> no box, every add goes to H
};

__device__ __forceinline__ uint32_t fkey(float f)
{
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__host__ __device__ __forceinline__ float unkey(uint32_t k)
{
    uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

__device__ __forceinline__ int cell_of(const CellGeom &g, float ku, float kv)
{
    int cu = (int)floorf((ku - g.u0) * g.inv_w);
    int cv = (int)floorf((kv - g.v0) * g.inv_w);
    cu = min(max(cu, 0), g.ncu - 1);
    cv = min(max(cv, 0), g.ncv - 1);
    return cv * g.ncu + cu;
}

__device__ __forceinline__ float comp(const float4 &r, int a) { return a == 0 ? r.x : (a == 1 ? r.y : r.z); }

__global__ void fill_u32_kernel(uint32_t *p, uint32_t v, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void fill_u64_kernel(unsigned long long *p, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// per-cell extents start empty: min keys at the top, max keys at the bottom
__global__ void init_cell_ext_kernel(uint32_t *ext, int ncell)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncell)
        for (int d = 0; d < 3; d++) {
            ext[6 * c + d] = 0xffffffffu;
            ext[6 * c + 3 + d] = 0u;
        }
}

// min/max of each k component: ext[0..2] = min keys, ext[3..5] = max keys
__global__ void kext_kernel(const float *__restrict__ kmap, uint32_t npix, uint32_t *ext)
{
    uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += gridDim.x * blockDim.x) {
        for (int d = 0; d < 3; d++) {
            const uint32_t k = fkey(kmap[3 * (size_t)i + d]);
            mn[d] = min(mn[d], k);
            mx[d] = max(mx[d], k);
        }
    }
    for (int d = 0; d < 3; d++) {
        atomicMin(ext + d, mn[d]);
        atomicMax(ext + 3 + d, mx[d]);
    }
}

// per-cell k extents: ext[6c + d] = min key of component d, ext[6c + 3 + d] = max key
__global__ void cell_ext_kernel(const float *__restrict__ kmap, uint32_t npix, CellGeom g, uint32_t *ext)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += gridDim.x * blockDim.x) {
        const float k[3] = {kmap[3 * (size_t)i], kmap[3 * (size_t)i + 1], kmap[3 * (size_t)i + 2]};
        const int c = cell_of(g, k[g.au], k[g.av]);
        for (int d = 0; d < 3; d++) {
            atomicMin(ext + 6 * c + d, fkey(k[d]));
            atomicMax(ext + 6 * c + 3 + d, fkey(k[d]));
        }
    }
}

// One block per frame: counting sort of the frame's records by cell.
// cs[f*(ncell+1) + c] = first sorted position of cell c relative to off[f].
__global__ void __launch_bounds__(kBlock)
    cell_sort_kernel(const float4 *__restrict__ rec, const uint64_t *__restrict__ off, CellGeom g,
                     float4 *__restrict__ srec, uint32_t *__restrict__ cs)
{
    __shared__ uint32_t cnt[kMaxCells];
    __shared__ uint32_t part[kBlock];
    const uint32_t f = blockIdx.x;
    const int ncell = g.ncu * g.ncv;
    const uint64_t base = off[f];
    const uint32_t n = (uint32_t)(off[f + 1] - base);
    for (int c = threadIdx.x; c < ncell; c += kBlock) cnt[c] = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += kBlock) {
        const float4 r = rec[base + i];
        atomicAdd(&cnt[cell_of(g, comp(r, g.au), comp(r, g.av))], 1u);
    }
    __syncthreads();
    // exclusive scan of cnt[0..ncell): per-thread chunk, then a block scan of the chunk sums
    const int per = (ncell + kBlock - 1) / kBlock;
    const int c0 = min(ncell, (int)threadIdx.x * per), c1 = min(ncell, c0 + per);
    uint32_t s = 0;
    for (int c = c0; c < c1; c++) s += cnt[c];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < kBlock; o <<= 1) {
        const uint32_t v = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
    uint32_t *csf = cs + (size_t)f * (ncell + 1);
    for (int c = c0; c < c1; c++) {
        const uint32_t v = cnt[c];
        csf[c] = run;
        cnt[c] = run;  // becomes the scatter cursor
        run += v;
    }
    if (threadIdx.x == 0) csf[ncell] = n;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += kBlock) {
        const float4 r = rec[base + i];
        const uint32_t pos = atomicAdd(&cnt[cell_of(g, comp(r, g.au), comp(r, g.av))], 1u);
        srec[base + pos] = r;
    }
}

// Sum over all records of the self pair's value (q = 0 lands on the centre voxel).
__global__ void self_sum_kernel(const float4 *__restrict__ rec, uint64_t n, unsigned long long *out)
{
    unsigned long long s = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const float c = rec[i].w;
        s += iac::pair_val(c * c);
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(out, s);
}

// mesh[v] += H[v] + H[last - v]; the centre also receives the self-pair sum.
__global__ void symmetrize_kernel(const unsigned long long *__restrict__ H, unsigned long long *__restrict__ mesh,
                                  size_t nvox, size_t centre, const unsigned long long *self)
{
    const size_t last = nvox - 1;
    for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < nvox; v += (size_t)gridDim.x * blockDim.x) {
        unsigned long long add = H[v] + H[last - v];
        if (v == centre) add += *self;
        if (add) mesh[v] += add;
    }
}

struct PairSink {
    unsigned long long *box;  // LDS
    unsigned long long *H;
    int bx0, by0, bz0, bx, by, bz;
    int S, SC;
    __device__ __forceinline__ void add(int fs, int ms, int ss, uint32_t v, unsigned long long *dropped)
    {
        const unsigned lx = (unsigned)(fs - bx0), ly = (unsigned)(ms - by0), lz = (unsigned)(ss - bz0);
        if (lx < (unsigned)bx && ly < (unsigned)by && lz < (unsigned)bz) {
            atomicAdd(box + lx + (size_t)bx * (ly + (size_t)by * lz), (unsigned long long)v);
        } else if ((unsigned)fs < (unsigned)S && (unsigned)ms < (unsigned)S && (unsigned)ss < (unsigned)SC) {
            atomicAdd(H + (size_t)fs + (size_t)ms * S + (size_t)ss * S * S, (unsigned long long)v);
        } else {
            atomicAdd(dropped, 2ull);  // the pair and its mirror
        }
    }
};

__global__ void __launch_bounds__(kBlock)
    pairs_lds_kernel(const float4 *__restrict__ srec, const uint64_t *__restrict__ off,
                     const uint32_t *__restrict__ cs, uint32_t nframes, uint32_t fchunk,
                     const Task *__restrict__ tasks, uint32_t ntasks, CellGeom g, iac::BinParams P,
                     unsigned long long *__restrict__ H, unsigned long long *__restrict__ dropped)
{
    __shared__ unsigned long long box[kBoxMax];
    __shared__ uint4 item[kGroup];        // {a base, na, b base, nb}
    __shared__ uint32_t pre[kGroup + 1];  // exclusive prefix of pair counts
    __shared__ uint32_t wsum[kBlock];

    const uint32_t tix = blockIdx.x % ntasks;
    const uint32_t chunk = blockIdx.x / ntasks;
    const Task T = tasks[tix];
    const uint32_t f0 = chunk * fchunk;
    const uint32_t f1 = min(nframes, f0 + fchunk);
    const int ncell = g.ncu * g.ncv;
    const bool tri = (T.du == 0 && T.dv == 0);
    const float R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};

    PairSink sink{box, H, T.bx0, T.by0, T.bz0, T.bx, T.by, T.bz, P.sab, P.sc};
    const int vbox = T.bx * T.by * T.bz;
    for (int v = threadIdx.x; v < vbox; v += kBlock) box[v] = 0ull;
    __syncthreads();

    const uint32_t cells = (uint32_t)T.nu * (uint32_t)T.nv;
    const uint64_t nitems = (uint64_t)(f1 > f0 ? f1 - f0 : 0) * cells;
    for (uint64_t ub = 0; ub < nitems; ub += kGroup) {
        // phase 1: item descriptors and pair counts (4 items per thread, contiguous)
        uint32_t mycnt[kItemsPerThread];
        uint32_t tsum = 0;
#pragma unroll
        for (int k = 0; k < kItemsPerThread; k++) {
            const uint32_t li = threadIdx.x * kItemsPerThread + k;
            const uint64_t u = ub + li;
            uint4 d = make_uint4(0, 0, 0, 0);
            uint32_t c = 0;
            if (u < nitems) {
                const uint32_t f = f0 + (uint32_t)(u / cells);
                const uint32_t rem = (uint32_t)(u % cells);
                const int iv = T.iv0 + (int)(rem / T.nu), iu = T.iu0 + (int)(rem % T.nu);
                const int ci = iv * g.ncu + iu;
                const int cj = (iv - T.dv) * g.ncu + (iu - T.du);
                const uint32_t *csf = cs + (size_t)f * (ncell + 1);
                const uint32_t base = (uint32_t)off[f];
                const uint32_t a0 = csf[ci], na = csf[ci + 1] - a0;
                const uint32_t b0 = csf[cj], nb = csf[cj + 1] - b0;
                d = make_uint4(base + a0, na, base + b0, nb);
                c = tri ? na * (na - 1) / 2 : na * nb;
            }
            item[li] = d;
            mycnt[k] = c;
            tsum += c;
        }
        wsum[threadIdx.x] = tsum;
        __syncthreads();
        for (int o = 1; o < kBlock; o <<= 1) {
            const uint32_t v = threadIdx.x >= (unsigned)o ? wsum[threadIdx.x - o] : 0u;
            __syncthreads();
            wsum[threadIdx.x] += v;
            __syncthreads();
        }
        {
            uint32_t run = threadIdx.x ? wsum[threadIdx.x - 1] : 0u;
#pragma unroll
            for (int k = 0; k < kItemsPerThread; k++) {
                pre[threadIdx.x * kItemsPerThread + k] = run;
                run += mycnt[k];
            }
            if (threadIdx.x == kBlock - 1) pre[kGroup] = run;
        }
        __syncthreads();

        // phase 2: thread t bins the contiguous pair range [p0, p1)
        const uint32_t total = pre[kGroup];
        const uint32_t p0 = (uint32_t)(((uint64_t)total * threadIdx.x) / kBlock);
        const uint32_t p1 = (uint32_t)(((uint64_t)total * (threadIdx.x + 1)) / kBlock);
        if (p0 < p1) {
            // the item holding p0: largest u with pre[u] <= p0
            int lo = 0, hi = kGroup;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (pre[mid] <= p0) lo = mid; else hi = mid;
            }
            int u = lo;
            uint4 d = item[u];
            uint32_t r = p0 - pre[u];
            uint32_t ia, ib;
            if (tri) {
                ia = 0;
                while (r >= d.y - 1 - ia) {
                    r -= d.y - 1 - ia;
                    ia++;
                }
                ib = ia + 1 + r;
            } else {
                ia = r / d.w;
                ib = r % d.w;
            }
            float4 ra = srec[d.x + ia];
            for (uint32_t p = p0;;) {
                const float4 rb = srec[d.z + ib];
                int dx, dy, dz;
                if (iac::bin_nn(ra.x, ra.y, ra.z, rb.x, rb.y, rb.z, P, R, dx, dy, dz)) {
                    const uint32_t v = iac::pair_val(ra.w * rb.w);
                    if (v) sink.add(dx + P.c0, dy + P.c1, dz + P.c2, v, dropped);
                }
                if (++p == p1) break;
                // next pair; one exists because p < p1 <= total
                ib++;
                if (ib >= (tri ? d.y : d.w)) {
                    ia++;
                    ib = tri ? ia + 1 : 0;
                    if ((tri ? ib : ia) >= d.y) {  // item exhausted: next item with pairs
                        do {
                            u++;
                        } while (pre[u + 1] == pre[u]);
                        d = item[u];
                        ia = 0;
                        ib = tri ? 1 : 0;
                    }
                    ra = srec[d.x + ia];
                }
            }
        }
        __syncthreads();
    }

    // flush the LDS box into H
    for (int v = threadIdx.x; v < vbox; v += kBlock) {
        const unsigned long long x = box[v];
        if (x) {
            const int lx = v % T.bx, rest = v / T.bx, ly = rest % T.by, lz = rest / T.by;
            const size_t gi = (size_t)(T.bx0 + lx) + (size_t)(T.by0 + ly) * P.sab +
                              (size_t)(T.bz0 + lz) * P.sab * P.sab;
            atomicAdd(H + gi, x);
        }
    }
}

// ---------------------------------------------------------------------------
// host planning

}  // namespace

struct iac_ac3d_plan {
    bool valid = false;
    const float *kmap = nullptr;
    uint32_t npix = 0;
    iac_mesh_shape shape{};
    int cell_vox = 0;
    CellGeom g{};
    std::vector<uint32_t> cell_ext;  // 6 keys per cell
    std::vector<Task> tasks;
    Task *d_tasks = nullptr;
};

namespace {
using Plan = iac_ac3d_plan;

// Conservative interval of round(x) for x in [lo, hi] (x in voxel units).
inline void round_range(double lo, double hi, int &a, int &b)
{
    a = (int)std::floor(lo) - 1;
    b = (int)std::ceil(hi) + 1;
}

void build_tasks(Plan &pl, const iac::BinParams &P)
{
    const CellGeom &g = pl.g;
    const int nc = g.ncu * g.ncv;
    std::vector<char> has(nc, 0);
    std::vector<double> kmn(3 * nc), kmx(3 * nc);
    for (int c = 0; c < nc; c++) {
        const uint32_t *e = &pl.cell_ext[6 * c];
        if (e[0] == 0xffffffffu) continue;  // no pixel in this cell
        has[c] = 1;
        for (int d = 0; d < 3; d++) {
            kmn[3 * c + d] = unkey(e[d]);
            kmx[3 * c + d] = unkey(e[3 + d]);
        }
    }
    const double vox = P.vox;
    const double maxq = std::sqrt((double)P.cut_sq);
    pl.tasks.clear();
    for (int dv = 0; dv < g.ncv; dv++) {
        for (int du = -(g.ncu - 1); du < g.ncu; du++) {
            if (dv == 0 && du < 0) continue;  // positive half plane of offsets
            Task t{};
            t.du = du;
            t.dv = dv;
            t.iu0 = std::max(0, du);
            t.iv0 = std::max(0, dv);
            t.nu = std::min(g.ncu, g.ncu + du) - t.iu0;
            t.nv = std::min(g.ncv, g.ncv + dv) - t.iv0;
            if (t.nu <= 0 || t.nv <= 0) continue;
            double qlo[3] = {1e30, 1e30, 1e30}, qhi[3] = {-1e30, -1e30, -1e30};
            bool any = false;
            for (int iv = t.iv0; iv < t.iv0 + t.nv; iv++) {
                for (int iu = t.iu0; iu < t.iu0 + t.nu; iu++) {
                    const int ci = iv * g.ncu + iu, cj = (iv - dv) * g.ncu + (iu - du);
                    if (!has[ci] || !has[cj]) continue;
                    any = true;
                    for (int d = 0; d < 3; d++) {
                        qlo[d] = std::min(qlo[d], kmn[3 * ci + d] - kmx[3 * cj + d]);
                        qhi[d] = std::max(qhi[d], kmx[3 * ci + d] - kmn[3 * cj + d]);
                    }
                }
            }
            if (!any) continue;
            // every pair of this task is cut if the box lies beyond Max_Q
            double dist2 = 0;
            for (int d = 0; d < 3; d++) {
                const double m = qlo[d] > 0 ? qlo[d] : (qhi[d] < 0 ? -qhi[d] : 0.0);
                dist2 += m * m;
            }
            if (std::sqrt(dist2) > maxq * (1.0 + 1e-5) + 1e-7) continue;
            int lo[3], hi[3];
            for (int d = 0; d < 3; d++) {
                round_range(qlo[d] * vox, qhi[d] * vox, lo[d], hi[d]);
                const int c = d == 0 ? P.c0 : (d == 1 ? P.c1 : P.c2);
                const int n = d == 2 ? P.sc : P.sab;
                lo[d] = std::max(0, lo[d] + c);
                hi[d] = std::min(n - 1, hi[d] + c);
            }
            const long vol = (long)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
            if (vol > 0 && vol <= kBoxMax) {
                t.bx0 = lo[0];
                t.by0 = lo[1];
                t.bz0 = lo[2];
                t.bx = hi[0] - lo[0] + 1;
                t.by = hi[1] - lo[1] + 1;
                t.bz = hi[2] - lo[2] + 1;
            }
            pl.tasks.push_back(t);
        }
    }
}

}  // namespace

namespace iac {

void free_ac3d_plan(iac_ctx *ctx)
{
    if (!ctx->ac3d_plan) return;
    if (ctx->ac3d_plan->d_tasks) (void)hipFree(ctx->ac3d_plan->d_tasks);
    delete ctx->ac3d_plan;
    ctx->ac3d_plan = nullptr;
}

// Entry used by iac_ac3d_sparse_batch for nearest-neighbour second-level
// binning with identity rotation.  rec: gathered records (scratch slot 0).
int ac3d_batch_lds(iac_ctx *ctx, const float *d_kmap, uint32_t npix, const float4 *rec,
                   const uint64_t *d_frame_off, uint32_t nframes, uint64_t n_total,
                   const iac_mesh_shape *shape, const iac::BinParams &P, uint64_t *d_mesh,
                   int cell_vox, uint32_t fchunk)
{
    hipStream_t st = ctx->stream;
    if (!ctx->ac3d_plan) ctx->ac3d_plan = new iac_ac3d_plan();
    Plan &pl = *ctx->ac3d_plan;
    if (cell_vox <= 0) cell_vox = 4;
    if (fchunk == 0) fchunk = 128;

    // ---- plan: cell grid from the k extents, per-cell extents, tasks (cached)
    void *ext = nullptr;
    IAC_TRY(scratch(ctx, 3, (6 + 6 * kMaxCells) * sizeof(uint32_t), &ext));
    uint32_t *d_ext = static_cast<uint32_t *>(ext);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, st, d_ext, 0xffffffffu, (size_t)3);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, st, d_ext + 3, 0u, (size_t)3);
    const unsigned eb = std::min<unsigned>(1024, (npix + 255) / 256);
    hipLaunchKernelGGL(kext_kernel, dim3(eb), dim3(256), 0, st, d_kmap, npix, d_ext);
    IAC_HIP(hipGetLastError());
    uint32_t hext[6];
    IAC_HIP(hipMemcpyAsync(hext, d_ext, sizeof(hext), hipMemcpyDeviceToHost, st));
    IAC_HIP(hipStreamSynchronize(st));
    float kmin[3], kmax[3], ext3[3];
    for (int d = 0; d < 3; d++) {
        kmin[d] = unkey(hext[d]);
        kmax[d] = unkey(hext[3 + d]);
        ext3[d] = kmax[d] - kmin[d];
    }
    CellGeom g{};
    // the two widest axes span the grid
    int order[3] = {0, 1, 2};
    std::sort(order, order + 3, [&](int a, int b) { return ext3[a] > ext3[b]; });
    g.au = std::min(order[0], order[1]);
    g.av = std::max(order[0], order[1]);
    double w = (double)cell_vox / (double)P.vox;
    auto ncells = [&](int a) { return std::max(1, (int)std::ceil(ext3[a] / w * (1 + 1e-6))); };
    while ((long)ncells(g.au) * ncells(g.av) > kMaxCells) w *= 1.25;
    g.inv_w = (float)(1.0 / w);
    g.u0 = kmin[g.au];
    g.v0 = kmin[g.av];
    g.ncu = ncells(g.au);
    g.ncv = ncells(g.av);
    const int nc = g.ncu * g.ncv;
    {
        hipLaunchKernelGGL(init_cell_ext_kernel, dim3((nc + 255) / 256), dim3(256), 0, st, d_ext + 6, nc);
        hipLaunchKernelGGL(cell_ext_kernel, dim3(eb), dim3(256), 0, st, d_kmap, npix, g, d_ext + 6);
        IAC_HIP(hipGetLastError());
        std::vector<uint32_t> cext(6 * (size_t)nc);
        IAC_HIP(hipMemcpyAsync(cext.data(), d_ext + 6, cext.size() * 4, hipMemcpyDeviceToHost, st));
        IAC_HIP(hipStreamSynchronize(st));
        const bool same = pl.valid && pl.kmap == d_kmap && pl.npix == npix && pl.cell_vox == cell_vox &&
                          std::memcmp(&pl.shape, shape, sizeof(*shape)) == 0 &&
                          std::memcmp(&pl.g, &g, sizeof(g)) == 0 && pl.cell_ext == cext;
        if (!same) {
            pl.valid = false;
            pl.kmap = d_kmap;
            pl.npix = npix;
            pl.shape = *shape;
            pl.cell_vox = cell_vox;
            pl.g = g;
            pl.cell_ext.swap(cext);
            build_tasks(pl, P);
            if (pl.d_tasks) IAC_HIP(hipFree(pl.d_tasks));
            pl.d_tasks = nullptr;
            if (!pl.tasks.empty()) {
                IAC_HIP(hipMalloc(&pl.d_tasks, pl.tasks.size() * sizeof(Task)));
                IAC_HIP(hipMemcpy(pl.d_tasks, pl.tasks.data(), pl.tasks.size() * sizeof(Task),
                                  hipMemcpyHostToDevice));
            }
            pl.valid = true;
        }
    }

    // ---- buffers: sorted records (slot 1), cell offsets (slot 2), half mesh H (own)
    const size_t nvox = (size_t)P.sab * P.sab * P.sc;
    void *srec = nullptr, *csb = nullptr, *hb = nullptr;
    IAC_TRY(scratch(ctx, 1, n_total * sizeof(float4) + (size_t)nframes * (nc + 1) * sizeof(uint32_t), &srec));
    csb = static_cast<char *>(srec) + n_total * sizeof(float4);
    IAC_TRY(scratch(ctx, 2, nvox * sizeof(unsigned long long), &hb));
    auto *H = static_cast<unsigned long long *>(hb);
    const unsigned zb = (unsigned)std::min<size_t>(4096, (nvox + 255) / 256);
    hipLaunchKernelGGL(fill_u64_kernel, dim3(zb), dim3(256), 0, st, H, nvox);
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, st, ctx->d_counters + 2, (size_t)1);

    hipLaunchKernelGGL(cell_sort_kernel, dim3(nframes), dim3(kBlock), 0, st, rec, d_frame_off, g,
                       static_cast<float4 *>(srec), static_cast<uint32_t *>(csb));
    IAC_HIP(hipGetLastError());
    const unsigned sb = (unsigned)std::min<uint64_t>(2048, (n_total + 255) / 256);
    hipLaunchKernelGGL(self_sum_kernel, dim3(sb), dim3(256), 0, st, rec, n_total, ctx->d_counters + 2);
    IAC_HIP(hipGetLastError());

    const uint32_t ntasks = (uint32_t)pl.tasks.size();
    const uint32_t nchunks = (nframes + fchunk - 1) / fchunk;
    if (ntasks) {
        IAC_TRY(timer_begin(ctx));
        hipLaunchKernelGGL(pairs_lds_kernel, dim3(ntasks * nchunks), dim3(kBlock), 0, st,
                           static_cast<const float4 *>(srec), d_frame_off,
                           static_cast<const uint32_t *>(csb), nframes, fchunk, pl.d_tasks, ntasks, g, P,
                           H, ctx->d_counters);
        IAC_HIP(hipGetLastError());
        IAC_TRY(timer_end(ctx));
    }
    const size_t centre = (size_t)P.c0 + (size_t)P.c1 * P.sab + (size_t)P.c2 * P.sab * P.sab;
    hipLaunchKernelGGL(symmetrize_kernel, dim3(zb), dim3(256), 0, st, H,
                       reinterpret_cast<unsigned long long *>(d_mesh), nvox, centre, ctx->d_counters + 2);
    IAC_HIP(hipGetLastError());
    return IAC_OK;
}

}  // namespace iac
