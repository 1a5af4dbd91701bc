// iac_pairs.hip -- the LDS cell-box pair engine shared by the 3D AC_uw path
// (Detector::AutoCorrelateSparseList) and the dense C(q) path
// (AutoCorr_CQ_small, IAC/kernel.cl:27-184).
//
// Input: photon or pixel records {kx, ky, kz, value} of nframes frames; frames
// form groups (frame range -> one output mesh each).  One call:
//   1. plan (cached): the k-map's extents define a cell grid over its two
//      widest k axes (cell side `cell_vox` voxels); per-cell k extents bound
//      the voxel box reached by the pairs of every cell offset delta.
//   2. cell_sort: per frame, counting sort of its records by cell, dropping
//      zero-valued records (they add nothing in any mode), and the record count
//      of its fullest cell; in symmetric mode the frame's self-pair value is
//      summed for its group.
//   3. pairs: one block per (delta [, region of cells], chunk of frames of one
//      group).  The block privatises delta's voxel box in LDS across the chunk
//      and bins
//        symmetric modes: each unordered pair of distinct records once
//          (deltas of the positive half plane; within-cell pairs for 0),
//        ordered modes: every ordered pair (all deltas; self pairs in 0),
//      flushing the box with one atomic per nonzero voxel.  A pair outside
//      the box adds to global memory directly, except on in-box tasks of the
//      AC_uw specialised path, whose exact box provably holds every pair.
//      Work inside a block goes by groups of (frame, cell) items: the items'
//      pair counts are scanned (32-bit), then AC_uw runs walk the group's pairs
//      with the lanes of a wave on consecutive pairs (coalesced record loads)
//      and C(q) runs give each thread a contiguous pair range (either on
//      request).  A chunk with a frame whose fullest cell could push a group
//      past 2^32 pairs bins its items one after another instead.
//   4. symmetric modes: mesh[v] += H[v] + H[mirror v], centre += self.
// Nearest-neighbour binning is odd in q (iac_binning.h): the (b,a) pair lands
// on the point reflection of (a,b) with the same value, so step 4 restores
// the ordered-pair mesh exactly.  Linear modes are not odd and run ordered.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "iac_binning.h"
#include "iac_internal.h"
#include "iac_pairs.h"

#pragma clang fp contract(off)

namespace iac {
namespace {

constexpr int kBlock = 256;
// LDS box capacity (uint64 voxels): 4096 (32 KiB, 1024 items per scan group, 3 blocks
// per CU), 1920 (15 KiB, 512 items: 25.9 KB in all, so 6 blocks per CU) or 1024 (8 KiB,
// 512 items); chosen per plan.  1920 holds the tight boxes of 6-voxel cells (at most
// 1,859 voxels on a 1024^2 detector).
constexpr int kBoxMax = 4096;
constexpr int kBoxSmall = 1920;
constexpr int kBoxTiny = 1024;
// waves per SIMD the specialised-path kernels are compiled for: 6 keeps them within
// 85 VGPRs (no spills) so that 6 blocks fit a CU
template <bool FAST, int BOX>
struct Occupancy {
    static constexpr int blocks = FAST && BOX != kBoxMax ? 6 : 1;
};
template <int BOX>
struct BoxCfg {
    static constexpr int group = BOX == kBoxMax ? 1024 : 512;  // items per scan group
    static constexpr int ipt = group / kBlock;                  // items per thread
};
constexpr int kMaxCells = 4096;                    // NC_u * NC_v
// blocks per pair-kernel launch (tasks x chunks): 2^22 blocks of 256 threads keep
// the grid's work-item count far below 2^32; longer runs launch in chunk slices
constexpr uint32_t kMaxGridBlocks = 1u << 22;

struct CellGeom {
    int au, av;    // k axes spanned by the cell grid
    float u0, v0;  // grid origin
    float inv_wu;  // 1 / cell width along u (k units)
    float inv_wv;  // 1 / cell width along v
    int ncu, ncv;  // cells per axis
};

struct Task {
    int du, dv;         // cell offset delta (u, v)
    int iu0, iv0;       // first cell i whose partner i - delta is on the grid
    int nu, nv;         // extent of that cell range
    int bx0, by0, bz0;  // LDS box origin (voxel indices)
    int bx, by, bz;     // LDS box extent; bx == 0 -> no box, every add goes to global
    int inbox;          // the box lies inside the mesh: every nearest-neighbour pair lands in it
    int allin;          // the task's q box lies inside the cut sphere: no pair fails the cut
};

struct Chunk {
    uint32_t f0, f1, group;
};

__device__ __forceinline__ uint32_t fkey(float f)
{
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

inline float unkey(uint32_t k)
{
    uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

__device__ __forceinline__ int cell_of(const CellGeom &g, float ku, float kv)
{
    int cu = (int)floorf((ku - g.u0) * g.inv_wu);
    int cv = (int)floorf((kv - g.v0) * g.inv_wv);
    cu = min(max(cu, 0), g.ncu - 1);
    cv = min(max(cv, 0), g.ncv - 1);
    return cv * g.ncu + cu;
}

__device__ __forceinline__ float comp(const float4 &r, int a) { return a == 0 ? r.x : (a == 1 ? r.y : r.z); }

// q = u / d, r = u % d for u < 2^24, d >= 1, from the float reciprocal inv = 1.0f / d:
// the estimate is within one of the quotient, one correction step makes it exact.
__device__ __forceinline__ void divmod24(uint32_t u, uint32_t d, float inv, uint32_t &q, uint32_t &r)
{
    uint32_t t = (uint32_t)((float)u * inv);
    int32_t m = (int32_t)(u - __umul24(t, d));
    if (m < 0) {
        t--;
        m += (int32_t)d;
    } else if (m >= (int32_t)d) {
        t++;
        m -= (int32_t)d;
    }
    q = t;
    r = (uint32_t)m;
}

// (int)roundf(x) (half away from zero) for |x| < 2^23: x + copysign(0.49999997f, x)
// rounded to nearest, truncated.  Equal to roundf for every float in that range
// (checked exhaustively over all 2.5e9 of them); 3 instructions instead of 7.
__device__ __forceinline__ int round_away(float x)
{
    return (int)(x + copysignf(0.49999997f, x));
}

// pair_val with the integer fast path in a 24-bit multiply
__device__ __forceinline__ uint32_t pair_val_small(float p)
{
    if (p >= 0.0f && p < 21054.0f && p == truncf(p)) return __umul24((uint32_t)p, 1000u);
    return pair_val(p);
}

// AC_uw with a double map, a linear first level and the identity rotation: a pair's
// eight first-level corners are its voxels.  A linear second level over such an
// integral corner keeps one corner of weight 1 (Sep = 1), valued FloatToInt(double)
// of the corner's float value; a nearest one takes FloatToInt(float).
__host__ __device__ __forceinline__ bool lin_first_identity(const BinParams &P)
{
    return P.double_map && P.interp1 == 1 && P.rot_identity;
}

// the value of a symmetric AC_uw pair: FloatToInt(float), or FloatToInt(double)
// for a linear second level over an integral first-level voxel (BinParams::dval)
__device__ __forceinline__ uint32_t sym_val(const BinParams &P, float p)
{
    return P.dval ? pair_val_d((double)p) : pair_val_small(p);
}

// ---- binning policies -------------------------------------------------------

// C(q) fixed point of AutoCorr_CQ_small (kernel.cl:93-101): (ulong)((double)(Ia*Ib) * M),
// pairs below 1e-37f skipped.
__device__ __forceinline__ bool cq_value(const PairSpec &S, float p, double &val)
{
    val = (double)p;
    return !(val < (double)1e-37f);
}

// Symmetric (nearest-neighbour) binning of one pair: voxel + value.
template <int KIND>
__device__ __forceinline__ bool bin_sym(const PairSpec &S, const float4 &a, const float4 &b, int &fs,
                                        int &ms, int &ss, uint64_t &v)
{
    if (KIND == kKindAc3d) {
        int dx, dy, dz;
        if (!bin_nn(a.x, a.y, a.z, b.x, b.y, b.z, S.P, S.R, dx, dy, dz)) return false;
        v = sym_val(S.P, a.w * b.w);
        fs = dx + S.P.c0;
        ms = dy + S.P.c1;
        ss = dz + S.P.c2;
        return true;
    } else {
        double val;
        if (!cq_value(S, a.w * b.w, val)) return false;
        float q[3] = {a.x - b.x, a.y - b.y, a.z - b.z};
        const float s = (q[0] * q[0] + q[1] * q[1]) + q[2] * q[2];
        if (!(s < S.P.cut_sq)) return false;
        q[0] = q[0] * S.P.vox;
        q[1] = q[1] * S.P.vox;
        q[2] = q[2] * S.P.vox;
        fs = (int)(roundf(q[S.ka]) + (float)S.P.c0);
        ms = (int)(roundf(q[S.kb]) + (float)S.P.c1);
        ss = (int)(roundf(q[S.kc]) + (float)S.P.c2);
        v = (uint64_t)(val * S.mult);
        return true;
    }
}

template <int KIND>
__device__ __forceinline__ uint64_t self_value(const PairSpec &S, float c)
{
    if (KIND == kKindAc3d) return sym_val(S.P, c * c);
    double val;
    return cq_value(S, c * c, val) ? (uint64_t)(val * S.mult) : 0ull;
}

// Ordered (linear-mode) binning of one pair into a sink.
template <int KIND, class Sink>
__device__ __forceinline__ void bin_ordered(const PairSpec &S, const float4 &a, const float4 &b, Sink &sink)
{
    if (KIND == kKindAc3d) {
        add_q_entry(sink, a.x, a.y, a.z, b.x, b.y, b.z, a.w * b.w, S.P, S.R);
    } else {
        double val;
        if (!cq_value(S, a.w * b.w, val)) return;
        float q[3] = {a.x - b.x, a.y - b.y, a.z - b.z};
        const float s = (q[0] * q[0] + q[1] * q[1]) + q[2] * q[2];
        if (!(s < S.P.cut_sq)) return;
        q[0] = q[0] * S.P.vox;
        q[1] = q[1] * S.P.vox;
        q[2] = q[2] * S.P.vox;
        if (S.interp == 0) {
            sink.add((int)(roundf(q[S.ka]) + (float)S.P.c0), (int)(roundf(q[S.kb]) + (float)S.P.c1),
                     (int)(roundf(q[S.kc]) + (float)S.P.c2), (uint64_t)(val * S.mult));
            return;
        }
        // kernel.cl:136-181: `1 - (q - floor(q))` in float, weights in double
        const float qa = q[S.ka], qb = q[S.kb], qc = q[S.kc];
        const double sfs = (double)(1.0f - (qa - floorf(qa)));
        const double sms = (double)(1.0f - (qb - floorf(qb)));
        const double sss = (double)(1.0f - (qc - floorf(qc)));
        const int fs = (int)(floorf(qa) + (float)S.P.c0);
        const int ms = (int)(floorf(qb) + (float)S.P.c1);
        const int ss = (int)(floorf(qc) + (float)S.P.c2);
        const double vm = val * S.mult;
        sink.add(fs + 0, ms + 0, ss + 0, (uint64_t)(vm * ((sfs * sms) * sss)));
        sink.add(fs + 0, ms + 0, ss + 1, (uint64_t)(vm * ((sfs * sms) * (1 - sss))));
        sink.add(fs + 1, ms + 0, ss + 0, (uint64_t)(vm * (((1 - sfs) * sms) * sss)));
        sink.add(fs + 1, ms + 0, ss + 1, (uint64_t)(vm * (((1 - sfs) * sms) * (1 - sss))));
        sink.add(fs + 0, ms + 1, ss + 0, (uint64_t)(vm * ((sfs * (1 - sms)) * sss)));
        sink.add(fs + 0, ms + 1, ss + 1, (uint64_t)(vm * ((sfs * (1 - sms)) * (1 - sss))));
        sink.add(fs + 1, ms + 1, ss + 0, (uint64_t)(vm * (((1 - sfs) * (1 - sms)) * sss)));
        sink.add(fs + 1, ms + 1, ss + 1, (uint64_t)(vm * (((1 - sfs) * (1 - sms)) * (1 - sss))));
    }
}

// ---- small kernels -------------------------------------------------------------

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

// min/max of each k component: ext[0..2] min keys, ext[3..5] max keys; reduced over
// the wave by shuffles, then one atomic per wave and component
__global__ void __launch_bounds__(256) kext_kernel(const float *__restrict__ kmap, uint32_t npix, uint32_t *ext)
{
    uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += gridDim.x * blockDim.x)
        for (int d = 0; d < 3; d++) {
            const uint32_t k = fkey(kmap[3 * (size_t)i + d]);
            mn[d] = min(mn[d], k);
            mx[d] = max(mx[d], k);
        }
    for (int d = 0; d < 3; d++)
        for (int o = 32; o > 0; o >>= 1) {
            mn[d] = min(mn[d], (uint32_t)__shfl_xor(mn[d], o, 64));
            mx[d] = max(mx[d], (uint32_t)__shfl_xor(mx[d], o, 64));
        }
    if ((threadIdx.x & 63u) == 0)
        for (int d = 0; d < 3; d++) {
            atomicMin(ext + d, mn[d]);
            atomicMax(ext + 3 + d, mx[d]);
        }
}

// per-cell k extents: ext[6c + d] min key of component d, ext[6c + 3 + d] max key.
// LDS: the block's extents of every cell (6 ncell keys, dynamic), merged into ext
// with one atomic per touched cell and key.
__global__ void __launch_bounds__(256)
    cell_ext_lds_kernel(const float *__restrict__ kmap, uint32_t npix, CellGeom g, uint32_t *ext)
{
    extern __shared__ uint32_t sext[];
    const int ncell = g.ncu * g.ncv;
    for (int c = threadIdx.x; c < ncell; c += blockDim.x)
        for (int d = 0; d < 3; d++) {
            sext[6 * c + d] = 0xffffffffu;
            sext[6 * c + 3 + d] = 0u;
        }
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += gridDim.x * blockDim.x) {
        const float k[3] = {kmap[3 * (size_t)i], kmap[3 * (size_t)i + 1], kmap[3 * (size_t)i + 2]};
        const int c = cell_of(g, k[g.au], k[g.av]);
        for (int d = 0; d < 3; d++) {
            atomicMin(sext + 6 * c + d, fkey(k[d]));
            atomicMax(sext + 6 * c + 3 + d, fkey(k[d]));
        }
    }
    __syncthreads();
    for (int c = threadIdx.x; c < ncell; c += blockDim.x)
        if (sext[6 * c] != 0xffffffffu)
            for (int d = 0; d < 3; d++) {
                atomicMin(ext + 6 * c + d, sext[6 * c + d]);
                atomicMax(ext + 6 * c + 3 + d, sext[6 * c + 3 + d]);
            }
}

// the same with global atomics only (grids too large for the LDS copy)
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

// One block per frame: counting sort of the frame's nonzero records by cell.
// cs[f*(ncell+1) + c] = first sorted position of cell c relative to off[f];
// cs[f*(ncell+1) + ncell] = number of kept records; cmax[f] = records of the
// frame's fullest cell.  In symmetric mode the frame's self-pair values are added
// to self[group of f].
template <int KIND, bool ORD>
__global__ void __launch_bounds__(kBlock)
    cell_sort_kernel(const float4 *__restrict__ rec, const uint64_t *__restrict__ off, CellGeom g,
                     float4 *__restrict__ srec, uint32_t *__restrict__ cs, PairSpec S,
                     const uint32_t *__restrict__ group_off, uint32_t ngroups,
                     unsigned long long *__restrict__ self, uint32_t *__restrict__ wflag,
                     uint32_t *__restrict__ cmax)
{
    __shared__ uint32_t cnt[kMaxCells];
    __shared__ uint32_t part[kBlock];
    __shared__ unsigned long long sred[kBlock / 64];
    __shared__ uint32_t smax[kBlock / 64];
    const uint32_t f = blockIdx.x;
    const int ncell = g.ncu * g.ncv;
    const uint64_t base = off[f];
    const uint32_t n = (uint32_t)(off[f + 1] - base);
    for (int c = threadIdx.x; c < ncell; c += kBlock) cnt[c] = 0;
    __syncthreads();
    unsigned long long sv = 0;
    bool small = true;  // every kept weight an integer in [1, 129]
    for (uint32_t i = threadIdx.x; i < n; i += kBlock) {
        const float4 r = rec[base + i];
        if (r.w == 0.0f) continue;
        atomicAdd(&cnt[cell_of(g, comp(r, g.au), comp(r, g.av))], 1u);
        if (!ORD) sv += self_value<KIND>(S, r.w);
        small = small && r.w >= 1.0f && r.w <= 129.0f && r.w == truncf(r.w);
    }
    if (!small) wflag[0] = 0u;  // the run's flag starts at 1
    if (!ORD) {
        for (int o = 32; o > 0; o >>= 1) sv += __shfl_down(sv, o, 64);
        if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = sv;
    }
    __syncthreads();
    if (!ORD && threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < kBlock / 64; w++) t += sred[w];
        if (t) {
            uint32_t lo = 0, hi = ngroups;  // group of f: largest gi with group_off[gi] <= f
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (group_off[mid] <= f) lo = mid; else hi = mid;
            }
            atomicAdd(self + lo, t);
        }
    }
    // exclusive scan of cnt[0..ncell)
    const int per = (ncell + kBlock - 1) / kBlock;
    const int c0 = min(ncell, (int)threadIdx.x * per), c1 = min(ncell, c0 + per);
    uint32_t s = 0, m = 0;
    for (int c = c0; c < c1; c++) {
        s += cnt[c];
        m = max(m, cnt[c]);
    }
    part[threadIdx.x] = s;
    for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_down(m, o, 64));
    if ((threadIdx.x & 63) == 0) smax[threadIdx.x >> 6] = m;
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
    if (threadIdx.x == kBlock - 1) csf[ncell] = part[kBlock - 1];
    if (threadIdx.x == 0) cmax[f] = max(max(smax[0], smax[1]), max(smax[2], smax[3]));
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += kBlock) {
        const float4 r = rec[base + i];
        if (r.w == 0.0f) continue;
        const uint32_t pos = atomicAdd(&cnt[cell_of(g, comp(r, g.au), comp(r, g.av))], 1u);
        srec[base + pos] = r;
    }
}

// mesh[g][v] += H[g][v] + H[g][last - v]; the centre also receives self[g].
__global__ void symmetrize_kernel(const unsigned long long *__restrict__ H, unsigned long long *__restrict__ mesh,
                                  size_t nvox, uint32_t ngroups, size_t centre,
                                  const unsigned long long *__restrict__ self)
{
    const size_t last = nvox - 1, total = nvox * ngroups;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const size_t gi = i / nvox, v = i - gi * nvox;
        const unsigned long long *h = H + gi * nvox;
        unsigned long long add = h[v] + h[last - v];
        if (v == centre) add += self[gi];
        if (add) mesh[i] += add;
    }
}

// A first-level voxel (mesh coordinates of the identity binning) under the frame's
// rotation: Atomic_Add_rotated_q_Entry's nearest second level (ACMesh.cpp:319-336)
// applied to the float voxel offsets, which equal the first level's roundf(q vox).
__device__ __forceinline__ void rotate_voxel(const float *r, int c0, int c1, int c2, int &fs, int &ms, int &ss)
{
    float q0 = (float)(fs - c0), q1 = (float)(ms - c1), q2 = (float)(ss - c2);
    rotate3(q0, q1, q2, r);
    fs = (int)roundf(q0) + c0;
    ms = (int)roundf(q1) + c1;
    ss = (int)roundf(q2) + c2;
}

struct BoxSink {
    unsigned long long *box;  // LDS
    unsigned long long *out;  // global mesh (group's H or mesh)
    unsigned long long *dropped;
    int bx0, by0, bz0, bx, by, bz;
    int S, SC;
    unsigned long long drop_weight;
    const float *rot;  // the chunk's rotation of first-level voxels, or null
    int c0, c1, c2;
    // an add outside the LDS box: the mesh voxel under the chunk's rotation, or the dropped
    // count when it leaves the mesh.  Target and value by selection, one atomic: nested
    // divergent branches here kept values live across the whole group loop (spills).
    __device__ __forceinline__ void add_outside(int fs, int ms, int ss, uint64_t v)
    {
        if (rot) rotate_voxel(rot, c0, c1, c2, fs, ms, ss);
        const bool ok = (unsigned)fs < (unsigned)S && (unsigned)ms < (unsigned)S && (unsigned)ss < (unsigned)SC;
        const size_t idx = ok ? (size_t)fs + (size_t)ms * S + (size_t)ss * S * S : 0;
        atomicAdd(ok ? out + idx : dropped, ok ? (unsigned long long)v : drop_weight);
    }
    __device__ __forceinline__ void add(int fs, int ms, int ss, uint64_t v)
    {
        if (!v) return;
        const unsigned lx = (unsigned)(fs - bx0), ly = (unsigned)(ms - by0), lz = (unsigned)(ss - bz0);
        if (lx < (unsigned)bx && ly < (unsigned)by && lz < (unsigned)bz)
            atomicAdd(box + lx + (size_t)bx * (ly + (size_t)by * lz), (unsigned long long)v);
        else
            add_outside(fs, ms, ss, v);
    }
};

// Pairs [p0, p0 + left) of a group on the specialised path (AC_uw, symmetric
// nearest-neighbour binning, identity rotation), records read from `src` (the
// LDS stage, or the sorted records in global memory).  Software pipelined: the
// next pair's records are requested before this pair is binned.  INBOX: every
// pair lands in the LDS box -- its voxel range covers the rounded q range of the
// task and lies inside the mesh -- so there is no bounds test and the box offset
// is two 24-bit multiply-adds on byte strides; otherwise the bounds test and the
// global sink.  SMALL: every weight of the run is an integer in [1, 129], so the
// fixed-point value 1000 wa wb (below 2^24) is exactly (1000 wa) * wb in float.
template <bool SMALL, bool INBOX>
__device__ __forceinline__ void fast_pairs(const float4 *__restrict__ src, const uint4 *item, const uint32_t *pre,
                                           int u, uint4 d, uint32_t ia, uint32_t ib, uint32_t left, bool tri,
                                           const PairSpec &S, unsigned long long *box, const Task &T,
                                           BoxSink &sink)
{
    const int fox = S.P.c0 - T.bx0, foy = S.P.c1 - T.by0, foz = S.P.c2 - T.bz0;
    const int bx8 = 8 * T.bx, bxby8 = 8 * T.bx * T.by;
    const int c8 = 8 * fox + bx8 * foy + bxby8 * foz;
    const float vox = S.P.vox, cut_sq = S.P.cut_sq;
    float4 ra = src[d.x + ia];
    float4 rb = src[d.z + ib];
    uint32_t bi = d.z + ib, bend = d.z + (tri ? d.y : d.w);
    char *const boxb = reinterpret_cast<char *>(box);
    for (;;) {
        const float4 ca = ra, cb = rb;
        if (left > 1) {
            if (++bi == bend) {  // row of a done: next a, or the next item with pairs
                ia++;
                uint32_t nb = tri ? ia + 1 : 0;
                if ((tri ? nb : ia) >= d.y) {
                    do {
                        u++;
                    } while (pre[u + 1] == pre[u]);
                    d = item[u];
                    ia = 0;
                    nb = tri ? 1 : 0;
                }
                ra = src[d.x + ia];
                bi = d.z + nb;
                bend = d.z + (tri ? d.y : d.w);
            }
            rb = src[bi];
        }
        asm volatile("" ::"v"(cb.w));  // one 16-byte load
        const float q0 = ca.x - cb.x, q1 = ca.y - cb.y, q2 = ca.z - cb.z;
        const float sq = (q0 * q0 + q1 * q1) + q2 * q2;
        if (sq < cut_sq) {
            const int dx = round_away(q0 * vox);
            const int dy = round_away(q1 * vox);
            const int dz = round_away(q2 * vox);
            const uint32_t v = SMALL ? (uint32_t)((ca.w * 1000.0f) * cb.w) : pair_val_small(ca.w * cb.w);
            if (INBOX) {
                const int o = __mul24(dz, bxby8) + (__mul24(dy, bx8) + (dx * 8 + c8));
                atomicAdd(reinterpret_cast<unsigned long long *>(boxb + o), (unsigned long long)v);
            } else {
                const uint32_t lx = (uint32_t)(dx + fox), ly = (uint32_t)(dy + foy), lz = (uint32_t)(dz + foz);
                if (lx < (uint32_t)T.bx && ly < (uint32_t)T.by && lz < (uint32_t)T.bz)
                    atomicAdd(box + lx + __umul24(ly, (uint32_t)T.bx) + __umul24(lz, (uint32_t)(T.bx * T.by)),
                              (unsigned long long)v);
                else
                    sink.add(dx + S.P.c0, dy + S.P.c1, dz + S.P.c2, v);
            }
        }
        if (--left == 0) break;
    }
}

template <bool INBOX>
__device__ __forceinline__ void fast_pairs_w(bool small, const float4 *__restrict__ src, const uint4 *item,
                                             const uint32_t *pre, int u, uint4 d, uint32_t ia, uint32_t ib,
                                             uint32_t left, bool tri, const PairSpec &S, unsigned long long *box,
                                             const Task &T, BoxSink &sink)
{
    if (small)
        fast_pairs<true, INBOX>(src, item, pre, u, d, ia, ib, left, tri, S, box, T, sink);
    else
        fast_pairs<false, INBOX>(src, item, pre, u, d, ia, ib, left, tri, S, box, T, sink);
}

// Pair r (row-major, ia < ib) of the distinct unordered pairs of n records.
__device__ __forceinline__ void tri_decode(uint32_t r, uint32_t n, uint32_t &ia, uint32_t &ib)
{
    // rows before row a: S(a) = a (2n - a - 1) / 2; a = the largest with S(a) <= r
    const double b = 2.0 * n - 1.0;
    uint32_t a = (uint32_t)fmax(0.0, floor((b - sqrt(fmax(0.0, b * b - 8.0 * (double)r))) * 0.5));
    auto S = [n](uint64_t x) { return x * (2ull * n - x - 1) / 2; };
    while (a > 0 && S(a) > r) a--;
    while (S(a + 1) <= r) a++;
    ia = a;
    ib = a + 1 + (uint32_t)(r - S(a));
}

// Phase 2 of a group on the specialised path with lanes on consecutive pairs:
// wave w walks the pair range [w0, w1) 64 at a time, lane l on pair P + l, so the
// records a wave loads lie in the few rows those 64 pairs span rather than in
// 64 unrelated rows (one or two cache lines per load instead of up to 64).
//
// Item lookup: the group's items are compacted (every item has pairs, so item
// starts are distinct) and followed by 64 sentinel starts (~0).  The starts in
// [P, P + 64) tag their positions with P in a per-wave LDS row; a lane's item is
// the item before the step plus the number of tagged positions up to the lane
// (ballot + mbcnt, the count's base and the next step's item in scalar registers).
//
// Division: a non-triangle item carries h = RN(1 / (2 nb)) in place of na and -nb
// in place of nb, and ia = trunc(f32(2r + 1) * h) is exactly r / nb for every
// r < 2^22: (r + 1/2) / nb lies at least 1 / (2 nb) from the integers, and the two
// roundings move it by at most (r + 1/2) / nb * 2^-23.  A group of more than 2^22
// pairs takes a correction step (below 2^24) or divides in integers.
//
// Records are read at 32-bit byte offsets from the chunk's first record, one step
// ahead of the binning.  Lanes past the range read a real pair (their position
// clamped to the group's last pair) and do not bin.
// a volatile LDS word: cross-lane tags stay LDS accesses, in program order
typedef volatile __attribute__((address_space(3))) uint32_t lds_tag_t;

// The binning: FASTBIN, the AC_uw specialised inline path (SMALL and INBOX apply
// there); otherwise bin_ordered for ordered runs and bin_sym for symmetric ones.
// EXACT: the group divides by h (non-triangle, at most 2^22 pairs); otherwise the
// triangle decode, or the reciprocal 2h with one correction step below 2^24 pairs,
// or an integer division.  CUT: FASTBIN tests the cut (false on an all-in task).
template <int KIND, bool ORD, bool FASTBIN, bool SMALL, bool INBOX, bool EXACT, bool CUT = true>
__device__ __forceinline__ void fast_walk(const float4 *__restrict__ src, const uint4 *item, const uint32_t *pre,
                                          uint32_t u0, uint32_t w0, uint32_t w1, uint32_t total, bool tri,
                                          const PairSpec &S, unsigned long long *box, const Task &T, BoxSink &sink,
                                          lds_tag_t *fl)
{
    const uint32_t lane = threadIdx.x & 63u;
    // wave-uniform values in scalar registers
    w0 = __builtin_amdgcn_readfirstlane(w0);
    w1 = __builtin_amdgcn_readfirstlane(w1);
    const uint32_t last = __builtin_amdgcn_readfirstlane(total) - 1u;
    uint32_t ub = __builtin_amdgcn_readfirstlane(u0);  // the item before the step's first start
    const int fox = S.P.c0 - T.bx0, foy = S.P.c1 - T.by0, foz = S.P.c2 - T.bz0;
    const int bx8 = 8 * T.bx, bxby8 = 8 * T.bx * T.by;
    const int c8 = 8 * fox + bx8 * foy + bxby8 * foz;
    // C(q): byte stride of the mesh axis that record component i bins along (k_a -> fs, k_b -> ms, k_c -> ss)
    const int cst0 = S.ka == 0 ? 8 : (S.kb == 0 ? bx8 : bxby8);
    const int cst1 = S.ka == 1 ? 8 : (S.kb == 1 ? bx8 : bxby8);
    const int cst2 = S.ka == 2 ? 8 : (S.kb == 2 ? bx8 : bxby8);
    const float vox = S.P.vox, cut_sq = S.P.cut_sq;
    const char *const srcb = reinterpret_cast<const char *>(src);
    char *const boxb = reinterpret_cast<char *>(box);

    fl[lane] = ~0u;  // untagged: a tag is a pair index, below 2^32 - 1
    // the item of pair P + lane; ub moves to the item of pair P + 63
    auto item_at = [&](uint32_t P) -> uint32_t {
        const uint32_t t = pre[ub + 1u + lane] - P;
        if (t < 64u) fl[t] = P;
        const uint64_t starts = __ballot(fl[lane] == P);
        const uint32_t base = ub + (uint32_t)(starts & 1ull);
        ub += (uint32_t)__popcll(starts);
        return __builtin_amdgcn_mbcnt_hi((uint32_t)(starts >> 33),
                                         __builtin_amdgcn_mbcnt_lo((uint32_t)(starts >> 1), base));
    };
    // the descriptor and first pair of the item of pair P + lane, a step before its decode
    struct Ahead {
        uint4 d;
        uint32_t pu;
    };
    auto lookup = [&](uint32_t P) -> Ahead {
        const uint32_t u = item_at(P);
        return Ahead{item[u], pre[u]};
    };
    // byte offsets of the two records of pair min(P + lane, last), which item A holds
    auto decode = [&](const Ahead &A, uint32_t P, uint32_t &oa, uint32_t &ob) {
        const uint4 d = A.d;
        const uint32_t r = min(P + lane, last) - A.pu;
        uint32_t ia, ib;
        if (EXACT) {
            // d.w = -nb: b's record index d.z + r - ia nb in one 24-bit multiply-add
            // (ia, nb <= r < 2^22)
            ia = (uint32_t)((float)(2u * r + 1u) * __uint_as_float(d.y));
            oa = (d.x + ia) << 4;
            ob = (uint32_t)(__mul24((int)ia, (int)d.w) + (int)(d.z + r)) << 4;
            return;
        }
        if (tri) {
            tri_decode(r, d.y, ia, ib);
        } else if (last < (1u << 24)) {
            divmod24(r, 0u - d.w, 2.0f * __uint_as_float(d.y), ia, ib);  // 2h = RN(1 / nb)
        } else {
            const uint32_t nb = 0u - d.w;
            ia = r / nb;
            ib = r - ia * nb;
        }
        oa = (d.x + ia) << 4;
        ob = (d.z + ib) << 4;
    };
    auto load = [&](uint32_t o) { return *reinterpret_cast<const float4 *>(srcb + o); };

    auto bin = [&](const float4 &ca, const float4 &cb, bool active) {
        if (FASTBIN) {
            asm volatile("" ::"v"(cb.w), "v"(ca.w));  // 16-byte loads
            const float q0 = ca.x - cb.x, q1 = ca.y - cb.y, q2 = ca.z - cb.z;
            const float sq = (q0 * q0 + q1 * q1) + q2 * q2;
            const bool in = active && (!CUT || sq < cut_sq);
            if (INBOX && SMALL) {
                // every pair of an in-box task (lanes past the range read pairs of the same task)
                // lands in the box: branch-free, a lane outside the range or the cut adds 0, so
                // the step is one basic block and the next step's item lookup interleaves with it
                const int dx = round_away(q0 * vox);
                const int dy = round_away(q1 * vox);
                const int dz = round_away(q2 * vox);
                const uint32_t v = (uint32_t)((ca.w * 1000.0f) * cb.w);
                const int o = __mul24(dz, bxby8) + (__mul24(dy, bx8) + (dx * 8 + c8));
                atomicAdd(reinterpret_cast<unsigned long long *>(boxb + o), (unsigned long long)(in ? v : 0u));
                return;
            }
            if (in) {
                const int dx = round_away(q0 * vox);
                const int dy = round_away(q1 * vox);
                const int dz = round_away(q2 * vox);
                const uint32_t v = SMALL ? (uint32_t)((ca.w * 1000.0f) * cb.w) : pair_val_small(ca.w * cb.w);
                if (INBOX) {
                    const int o = __mul24(dz, bxby8) + (__mul24(dy, bx8) + (dx * 8 + c8));
                    atomicAdd(reinterpret_cast<unsigned long long *>(boxb + o), (unsigned long long)v);
                } else {
                    const uint32_t lx = (uint32_t)(dx + fox), ly = (uint32_t)(dy + foy), lz = (uint32_t)(dz + foz);
                    if (lx < (uint32_t)T.bx && ly < (uint32_t)T.by && lz < (uint32_t)T.bz)
                        atomicAdd(box + lx + __umul24(ly, (uint32_t)T.bx) + __umul24(lz, (uint32_t)(T.bx * T.by)),
                                  (unsigned long long)v);
                    else if (v)
                        sink.add_outside(dx + S.P.c0, dy + S.P.c1, dz + S.P.c2, v);
                }
            }
            return;
        }
        if (!active) return;
        if (ORD && KIND == kKindAc3d && INBOX) {
            // a linear first level (identity) on a task whose box holds every corner: add_q_entry's
            // eight corners at one base byte offset plus constant strides, their values
            // ((v wx) wy) wz from a product tree in the same order
            const float q0 = ca.x - cb.x, q1 = ca.y - cb.y, q2 = ca.z - cb.z;
            const float sq = (q0 * q0 + q1 * q1) + q2 * q2;
            if (sq < cut_sq) {
                const float x0 = q0 * vox, x1 = q1 * vox, x2 = q2 * vox;
                const float f0 = floorf(x0), f1 = floorf(x1), f2 = floorf(x2);
                const double s0 = 1.0 - (double)(x0 - f0), s1 = 1.0 - (double)(x1 - f1);
                const double s2 = 1.0 - (double)(x2 - f2);
                const double v = (double)(ca.w * cb.w);
                const double vx[2] = {v * s0, v * (1 - s0)};
                const int o = __mul24((int)f2, bxby8) + (__mul24((int)f1, bx8) + ((int)f0 * 8 + c8));
                auto corners = [&](auto dval_c) {
                    constexpr bool kDouble = decltype(dval_c)::value;  // linear second level
#pragma unroll
                    for (int c = 0; c < 8; c++) {
                        const int ex = c & 1, ey = (c >> 1) & 1, ez = (c >> 2) & 1;
                        const double w = (vx[ex] * (ey ? 1 - s1 : s1)) * (ez ? 1 - s2 : s2);
                        const uint32_t u = kDouble ? pair_val_d((double)(float)w) : pair_val((float)w);
                        if (u)
                            atomicAdd(
                                reinterpret_cast<unsigned long long *>(boxb + o + ex * 8 + ey * bx8 + ez * bxby8),
                                (unsigned long long)u);
                    }
                };
                if (S.P.interp2) corners(std::true_type{}); else corners(std::false_type{});
            }
        } else if (ORD) {
            // ordered runs (linear second level): every corner through the checked box
            bin_ordered<KIND>(S, ca, cb, sink);
        } else if (KIND == kKindCq && INBOX) {
            // C(q) of a task whose box holds every voxel its pairs reach: the bin of
            // bin_sym (its value and axis order: round(q[ka] vox) + c0 is round_away
            // of the same product plus c0) as one byte offset, component i scaled by
            // the box stride of the mesh axis it maps to
            asm volatile("" ::"v"(cb.w), "v"(ca.w));  // 16-byte loads
            const float w = ca.w * cb.w;
            const float q0 = ca.x - cb.x, q1 = ca.y - cb.y, q2 = ca.z - cb.z;
            const float sq = (q0 * q0 + q1 * q1) + q2 * q2;
            if (sq < cut_sq && !(w < 1e-37f)) {
                const int o = __mul24(round_away(q2 * vox), cst2) +
                              (__mul24(round_away(q1 * vox), cst1) + (__mul24(round_away(q0 * vox), cst0) + c8));
                atomicAdd(reinterpret_cast<unsigned long long *>(boxb + o),
                          (unsigned long long)(uint64_t)((double)w * S.mult));
            }
        } else {
            // C(q) through the checked box, or the general symmetric binning
            int fs, ms, ss;
            uint64_t v;
            if (bin_sym<KIND>(S, ca, cb, fs, ms, ss, v)) sink.add(fs, ms, ss, v);
        }
    };

    // two steps per iteration, the next step's records in flight while one bins.  The
    // look-ahead is unconditional (past the range it reads the group's last pair, or the
    // next wave's), and so is the second step (past the range no lane bins): a load
    // behind a branch, or a second exit, would make the binning wait for the loads in
    // flight too.  The scheduling barriers keep the binning's first use of its records
    // after the issue of the next step's loads.
    // lane of step P in range: with the exact division every pair index stays below
    // 2^22 + 192, so P + lane cannot wrap (one compare on the decode's own sum)
    auto active = [&](uint32_t P) { return EXACT ? P + lane < w1 : lane < (w1 > P ? w1 - P : 0u); };
    uint32_t oa, ob;
    decode(lookup(w0), w0, oa, ob);
    float4 a0 = load(oa), b0 = load(ob);
    Ahead A = lookup(w0 + 64u);
    for (uint32_t P = w0;; P += 128u) {
        decode(A, P + 64u, oa, ob);
        const float4 a1 = load(oa), b1 = load(ob);
        __builtin_amdgcn_sched_barrier(0);
        A = lookup(P + 128u);
        bin(a0, b0, active(P));
        decode(A, P + 128u, oa, ob);
        a0 = load(oa);
        b0 = load(ob);
        __builtin_amdgcn_sched_barrier(0);
        A = lookup(P + 192u);
        bin(a1, b1, active(P + 64u));
        if (P + 128u >= w1) break;
    }
}

template <int KIND, bool ORD, bool FAST>
__device__ __forceinline__ void fast_walk_w(bool small, bool inbox, bool exact, const float4 *__restrict__ src,
                                            const uint4 *item, const uint32_t *pre, uint32_t u0, uint32_t w0,
                                            uint32_t w1, uint32_t total, bool tri, const PairSpec &S,
                                            unsigned long long *box, const Task &T, BoxSink &sink, lds_tag_t *fl)
{
    // the triangle task and groups of more than 2^22 pairs: one general instance (the
    // checked box and the general value), so the rare decodes add one walk loop, not four
    if (!exact) {
        fast_walk<KIND, ORD, FAST, false, false, false>(src, item, pre, u0, w0, w1, total, tri, S, box, T, sink, fl);
        return;
    }
    auto go = [&](auto fastbin_c, auto small_c, auto inbox_c) {
        constexpr bool kFastbin = decltype(fastbin_c)::value, kSmall = decltype(small_c)::value;
        constexpr bool kInbox = decltype(inbox_c)::value;
        fast_walk<KIND, ORD, kFastbin, kSmall, kInbox, true>(src, item, pre, u0, w0, w1, total, tri, S, box, T, sink,
                                                             fl);
    };
    using T_ = std::true_type;
    using F_ = std::false_type;
    if (!FAST) {
        if (((KIND == kKindCq && !ORD && S.interp == 0) || (KIND == kKindAc3d && ORD && lin_first_identity(S.P))) &&
            inbox)
            go(F_{}, F_{}, T_{});
        else
            go(F_{}, F_{}, F_{});
    } else if (inbox) {
        if (small && T.allin)  // the headline case: no cut test
            fast_walk<KIND, ORD, true, true, true, true, false>(src, item, pre, u0, w0, w1, total, tri, S, box, T,
                                                                sink, fl);
        else if (small) go(T_{}, T_{}, T_{});
        else go(T_{}, F_{}, T_{});
    } else {
        if (small) go(T_{}, T_{}, F_{});
        else go(T_{}, F_{}, F_{});
    }
}

// PF: 0 lanes on consecutive pairs (the walk), 1 a pair range per thread, pair by
// pair, 4 or 8 a range per thread in batches of PF pairs
template <int KIND, bool ORD, int PF, bool FAST, int BOX>
__global__ void __launch_bounds__(kBlock, (Occupancy<FAST, BOX>::blocks))
    pairs_kernel(const float4 *__restrict__ srec, const uint64_t *__restrict__ off,
                 const uint32_t *__restrict__ cs, const Chunk *__restrict__ chunks,
                 const Task *__restrict__ tasks, uint32_t ntasks, CellGeom g, PairSpec S,
                 unsigned long long *__restrict__ target, size_t nvox,
                 unsigned long long *__restrict__ dropped, const uint32_t *__restrict__ wflag,
                 const float *__restrict__ rots, const uint32_t *__restrict__ cmax)
{
    constexpr int kGroup = BoxCfg<BOX>::group;
    constexpr int kItemsPerThread = BoxCfg<BOX>::ipt;
    __shared__ unsigned long long box[BOX];
    __shared__ uint4 item[kGroup];         // {a base, na (walk: h), b base, nb}, bases from the chunk's first record
    __shared__ uint32_t pre[kGroup + 64];  // exclusive prefix of pair counts (walk: then 64 sentinels)
    __shared__ uint32_t wtot[kBlock / 64];
    __shared__ uint32_t wnz[kBlock / 64];         // walk: items with pairs per wave
    __shared__ uint32_t wtags[kBlock / 64][64];   // walk: item-start tags per wave
    __shared__ uint32_t wfirst[kBlock / 64];      // walk: the item holding each wave's first pair
    __shared__ uint32_t s_total;                  // walk: pairs of the group
    __shared__ uint32_t s_heavy[kBlock / 64];     // per wave: the chunk holds an item of 2^32 / kGroup pairs or more

    const Task T = tasks[blockIdx.x % ntasks];
    const Chunk CH = chunks[blockIdx.x / ntasks];
    // the chunk's records: item bases count from its first record, so record byte
    // offsets stay below 2^32 (a chunk holds far fewer than 2^28 records)
    const uint64_t rec0 = off[CH.f0];
    const float4 *__restrict__ crec = srec + rec0;
    const int ncell = g.ncu * g.ncv;
    const bool zero_delta = (T.du == 0 && T.dv == 0);
    const bool tri = !ORD && zero_delta;  // distinct unordered pairs within a cell

    // per-frame rotations: a chunk holds frames of one rotation, applied to first-level voxels
    const float *crot = rots ? rots + 9 * (size_t)CH.f0 : nullptr;
    BoxSink sink{box, target + (size_t)CH.group * nvox, dropped, T.bx0, T.by0, T.bz0, T.bx, T.by, T.bz,
                 S.P.sab, S.P.sc, ORD ? 1ull : 2ull, crot, S.P.c0, S.P.c1, S.P.c2};
    const int vbox = T.bx * T.by * T.bz;
    const int fox = S.P.c0 - T.bx0, foy = S.P.c1 - T.by0, foz = S.P.c2 - T.bz0;  // fast path offsets
    const uint32_t fbxby = (uint32_t)(T.bx * T.by);
    // in-box tasks: the AC_uw specialised path, or symmetric nearest-neighbour C(q) (tight boxes)
    const bool inbox = (FAST || (KIND == kKindCq && !ORD && S.interp == 0) ||
                        (KIND == kKindAc3d && ORD && lin_first_identity(S.P))) &&
                       T.inbox != 0;
    const bool small = FAST && wflag[0] != 0u;
    // in-box C(q) on a pair range per thread: byte strides of the box axis each record
    // component bins along, and the centre's byte offset in the box
    const int cbx8 = 8 * T.bx, cbxby8 = 8 * T.bx * T.by;
    const int cc8 = 8 * fox + cbx8 * foy + cbxby8 * foz;
    const int cst0 = S.ka == 0 ? 8 : (S.kb == 0 ? cbx8 : cbxby8);
    const int cst1 = S.ka == 1 ? 8 : (S.kb == 1 ? cbx8 : cbxby8);
    const int cst2 = S.ka == 2 ? 8 : (S.kb == 2 ? cbx8 : cbxby8);
    char *const cboxb = reinterpret_cast<char *>(box);
    // pair counts and prefixes are 32-bit: a group of kGroup items each holding fewer
    // than 2^32 / kGroup pairs stays below 2^32.  A frame whose fullest cell holds M
    // records has items of at most M^2 pairs.
    constexpr uint64_t kHeavy = (1ull << 32) / kGroup;
    bool hv = false;
    for (uint32_t f = CH.f0 + threadIdx.x; f < CH.f1; f += kBlock) hv = hv || (uint64_t)cmax[f] * cmax[f] >= kHeavy;
    const bool whv = __ballot(hv) != 0ull;
    if ((threadIdx.x & 63u) == 0) s_heavy[threadIdx.x >> 6] = whv;
    for (int v = threadIdx.x; v < vbox; v += kBlock) box[v] = 0ull;
    __syncthreads();
    const bool heavy = (s_heavy[0] | s_heavy[1] | s_heavy[2] | s_heavy[3]) != 0u;

    // items (frame, cell) of the chunk: fewer than 2^24 (at most 4096 frames x 4096 cells),
    // so quotients come from a float reciprocal with one correction step
    const uint32_t cells = (uint32_t)T.nu * (uint32_t)T.nv;
    const uint32_t nitems = (CH.f1 - CH.f0) * cells;
    const float inv_cells = 1.0f / (float)cells, inv_nu = 1.0f / (float)T.nu;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    // lanes on consecutive pairs (every mode unless a pair range per thread is asked for)
    constexpr bool walking = PF == 0;
    if (heavy) {
        // items one after another, each a grid of rows along its shorter side (triangle
        // items: row a, columns b > a) with the block's threads along a row and four
        // record loads in flight per thread; general binning through the checked box
        for (uint32_t u = 0; u < nitems; u++) {
            uint32_t fr, rem, row, col;
            divmod24(u, cells, inv_cells, fr, rem);
            divmod24(rem, (uint32_t)T.nu, inv_nu, row, col);
            const uint32_t f = CH.f0 + fr;
            const int iv = T.iv0 + (int)row, iu = T.iu0 + (int)col;
            const int ci = iv * g.ncu + iu;
            const int cj = (iv - T.dv) * g.ncu + (iu - T.du);
            const uint32_t *csf = cs + (size_t)f * (ncell + 1);
            const uint32_t base = (uint32_t)(off[f] - rec0);
            const uint32_t a0 = base + csf[ci], na = csf[ci + 1] - csf[ci];
            const uint32_t b0 = base + csf[cj], nb = csf[cj + 1] - csf[cj];
            const bool by_a = tri || na <= nb;  // rows are a records, columns b records
            const uint32_t nrow = by_a ? na : nb, ncol = by_a ? nb : na;
            const uint32_t rbase = by_a ? a0 : b0, cbase = by_a ? b0 : a0;
            if (ncol == 0) continue;
            for (uint32_t r = 0; r < nrow; r++) {
                const float4 rr = crec[rbase + r];
                for (uint32_t c = (tri ? r + 1 : 0) + threadIdx.x; c < ncol; c += 4 * kBlock) {
                    float4 rc[4];
#pragma unroll
                    for (int k = 0; k < 4; k++) rc[k] = crec[cbase + min(c + k * kBlock, ncol - 1)];
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        if (c + k * kBlock >= ncol) break;
                        const float4 &ra = by_a ? rr : rc[k];
                        const float4 &rb = by_a ? rc[k] : rr;
                        if (ORD) {
                            bin_ordered<KIND>(S, ra, rb, sink);
                        } else {
                            int fs, ms, ss;
                            uint64_t v;
                            if (bin_sym<KIND>(S, ra, rb, fs, ms, ss, v)) sink.add(fs, ms, ss, v);
                        }
                    }
                }
            }
        }
        __syncthreads();  // the box is complete before the flush
    }
    for (uint32_t ub = 0; ub < (heavy ? 0u : nitems); ub += kGroup) {
        // phase 1: item descriptors and pair counts (contiguous items per thread)
        uint32_t mycnt[kItemsPerThread];
        uint4 myd[kItemsPerThread];
        uint32_t tsum = 0, nz = 0;
#pragma unroll
        for (int k = 0; k < kItemsPerThread; k++) {
            const uint32_t li = threadIdx.x * kItemsPerThread + k;
            const uint32_t u = ub + li;
            uint4 d = make_uint4(0, 0, 0, 0);
            uint32_t c = 0;
            if (u < nitems) {
                uint32_t fr, rem, row, col;
                divmod24(u, cells, inv_cells, fr, rem);
                divmod24(rem, (uint32_t)T.nu, inv_nu, row, col);
                const uint32_t f = CH.f0 + fr;
                const int iv = T.iv0 + (int)row, iu = T.iu0 + (int)col;
                const int ci = iv * g.ncu + iu;
                const int cj = (iv - T.dv) * g.ncu + (iu - T.du);
                const uint32_t *csf = cs + (size_t)f * (ncell + 1);
                const uint32_t base = (uint32_t)(off[f] - rec0);
                const uint32_t a0 = csf[ci], na = csf[ci + 1] - a0;
                const uint32_t b0 = csf[cj], nb = csf[cj + 1] - b0;
                d = make_uint4(base + a0, na, base + b0, nb);
                c = tri ? na * (na - 1) / 2 : na * nb;
                // the walk decodes (ia, ib) of a non-triangle item with h = RN(1 / (2 nb)) in
                // place of na (2 nb is exact in float: nb < 2^24) and -nb in place of nb
                if (walking && !tri && c) {
                    d.y = __float_as_uint(1.0f / (float)(2u * nb));
                    d.w = 0u - nb;
                }
            }
            if (!walking) item[li] = d;
            myd[k] = d;
            mycnt[k] = c;
            tsum += c;
            nz += c ? 1u : 0u;
        }
        // block exclusive scan of the per-thread sums: within each wave by shuffles,
        // then across the four waves through LDS (two barriers)
        uint32_t x = tsum, z = nz;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            const uint32_t yz = walking ? __shfl_up(z, o, 64) : 0u;
            if (lane >= (uint32_t)o) {
                x += y;
                z += yz;
            }
        }
        if (lane == 63) {
            wtot[wave] = x;
            wnz[wave] = z;
        }
        __syncthreads();
        if (walking) {
            // items with pairs only, in order: item starts become distinct
            uint32_t run = x - tsum, rz = z - nz, tot = 0, nzt = 0;
            for (uint32_t w = 0; w < kBlock / 64; w++) {
                if (w < wave) {
                    run += wtot[w];
                    rz += wnz[w];
                }
                tot += wtot[w];
                nzt += wnz[w];
            }
#pragma unroll
            for (int k = 0; k < kItemsPerThread; k++) {
                if (mycnt[k]) {
                    // the item holding the first pair of waves 1..3 (the walk splits the
                    // group's pairs in quarters)
                    for (uint32_t w = 1; w < kBlock / 64; w++) {
                        const uint32_t first = (uint32_t)(((uint64_t)tot * w) / (kBlock / 64));
                        if (first >= run && first - run < mycnt[k]) wfirst[w] = rz;
                    }
                    item[rz] = myd[k];
                    pre[rz] = run;
                    rz++;
                    run += mycnt[k];
                }
            }
            if (threadIdx.x < 64) pre[nzt + threadIdx.x] = ~0u;  // sentinel starts
            if (threadIdx.x == 0) {
                wfirst[0] = 0;
                s_total = tot;
            }
        } else {
            uint32_t run = x - tsum;
            for (uint32_t w = 0; w < wave; w++) run += wtot[w];
#pragma unroll
            for (int k = 0; k < kItemsPerThread; k++) {
                pre[threadIdx.x * kItemsPerThread + k] = run;
                run += mycnt[k];
            }
            if (threadIdx.x == kBlock - 1) pre[kGroup] = run;
        }
        __syncthreads();

        // phase 2: thread t bins the contiguous pair range [p0, p1).  PF == 1: pair
        // by pair, one load each.  PF > 1: PF pairs at a time with an integer-only
        // enumeration, so the 2 PF record loads of a batch are in flight together.
        const uint32_t total = walking ? s_total : pre[kGroup];
        if (walking) {
            // waves on quarters of the group's pairs, lanes on consecutive pairs
            const uint32_t w0 = (uint32_t)(((uint64_t)total * wave) / (kBlock / 64));
            const uint32_t w1 = (uint32_t)(((uint64_t)total * (wave + 1)) / (kBlock / 64));
            if (w0 < w1)
                fast_walk_w<KIND, ORD, FAST>(small, inbox, !tri && total <= (1u << 22), crec, item, pre, wfirst[wave],
                                             w0, w1, total, tri, S, box, T, sink, (lds_tag_t *)wtags[wave]);
            __syncthreads();
            continue;
        }
        const uint32_t p0 = (uint32_t)(((uint64_t)total * threadIdx.x) / kBlock);
        const uint32_t p1 = (uint32_t)(((uint64_t)total * (threadIdx.x + 1)) / kBlock);
        if (p0 < p1) {
            int lo = 0, hi = kGroup;  // the item holding p0: largest u with pre[u] <= p0
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
            if (FAST && inbox) {
                fast_pairs_w<true>(small, crec, item, pre, u, d, ia, ib, p1 - p0, tri, S, box, T, sink);
            } else if (PF == 1) {
                // one record load per pair: the a record stays in a register along its row
                float4 ra = crec[d.x + ia];
                // in-box C(q): consecutive pairs of a row fall in runs on one voxel, so
                // a run's values add in a register and reach the box in one atomic
                int co = -1;
                unsigned long long cacc = 0;
                for (uint32_t p = p0;;) {
                    const float4 rb = crec[d.z + ib];
                    if (KIND == kKindCq && !ORD && inbox) {
                        asm volatile("" ::"v"(rb.w));  // one 16-byte load
                        const float w = ra.w * rb.w;
                        const float q0 = ra.x - rb.x, q1 = ra.y - rb.y, q2 = ra.z - rb.z;
                        const float sq = (q0 * q0 + q1 * q1) + q2 * q2;
                        if (sq < S.P.cut_sq && !(w < 1e-37f)) {
                            const int o = __mul24(round_away(q2 * S.P.vox), cst2) +
                                          (__mul24(round_away(q1 * S.P.vox), cst1) +
                                           (__mul24(round_away(q0 * S.P.vox), cst0) + cc8));
                            const unsigned long long v = (uint64_t)((double)w * S.mult);
                            if (o == co) {
                                cacc += v;
                            } else {
                                if (co >= 0) atomicAdd(reinterpret_cast<unsigned long long *>(cboxb + co), cacc);
                                co = o;
                                cacc = v;
                            }
                        }
                    } else if (FAST) {
                        // identity rotation, nearest neighbour: round(round(x)) == round(x), so every
                        // symmetric mode bins at (int)roundf(q * vox); box index in 24-bit multiplies
                        asm volatile("" ::"v"(rb.w));  // one 16-byte load, not xyz now and w later
                        const float q0 = ra.x - rb.x, q1 = ra.y - rb.y, q2 = ra.z - rb.z;
                        const float sq = (q0 * q0 + q1 * q1) + q2 * q2;
                        if (sq < S.P.cut_sq) {
                            const int dx = round_away(q0 * S.P.vox);
                            const int dy = round_away(q1 * S.P.vox);
                            const int dz = round_away(q2 * S.P.vox);
                            const uint64_t v = pair_val_small(ra.w * rb.w);
                            const uint32_t lx = (uint32_t)(dx + fox), ly = (uint32_t)(dy + foy),
                                           lz = (uint32_t)(dz + foz);
                            if (lx < (uint32_t)T.bx && ly < (uint32_t)T.by && lz < (uint32_t)T.bz)
                                atomicAdd(box + lx + __umul24(ly, (uint32_t)T.bx) + __umul24(lz, fbxby),
                                          (unsigned long long)v);
                            else
                                sink.add(dx + S.P.c0, dy + S.P.c1, dz + S.P.c2, v);
                        }
                    } else if (ORD) {
                        bin_ordered<KIND>(S, ra, rb, sink);
                    } else {
                        int fs, ms, ss;
                        uint64_t v;
                        if (bin_sym<KIND>(S, ra, rb, fs, ms, ss, v)) sink.add(fs, ms, ss, v);
                    }
                    if (++p == p1) break;
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
                        ra = crec[d.x + ia];
                    }
                }
                if (co >= 0) atomicAdd(reinterpret_cast<unsigned long long *>(cboxb + co), cacc);
            } else
            for (uint32_t p = p0; p < p1;) {
                const uint32_t nb = min((uint32_t)PF, p1 - p);
                float4 A[PF > 1 ? PF : 1], B[PF > 1 ? PF : 1];
#pragma unroll
                for (int k = 0; k < PF; k++) {
                    if ((uint32_t)k < nb) {
                        A[k] = crec[d.x + ia];
                        B[k] = crec[d.z + ib];
                        if (p + k + 1 < p1) {  // a next pair exists: advance the enumeration
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
                            }
                        }
                    }
                }
#pragma unroll
                for (int k = 0; k < PF; k++) {
                    if ((uint32_t)k < nb) {
                        if (ORD) {
                            bin_ordered<KIND>(S, A[k], B[k], sink);
                        } else {
                            int fs, ms, ss;
                            uint64_t v;
                            if (bin_sym<KIND>(S, A[k], B[k], fs, ms, ss, v)) sink.add(fs, ms, ss, v);
                        }
                    }
                }
                p += nb;
            }
        }
        __syncthreads();
    }

    // flush the LDS box (under the chunk's rotation, voxel by voxel: every pair of a box
    // voxel shares its first-level voxel, so the rotated second level bins them alike)
    unsigned long long *out = target + (size_t)CH.group * nvox;
    float rr[9];
    int ax = 0;  // the box axis consecutive lanes walk
    if (crot) {
        for (int i = 0; i < 9; i++) rr[i] = crot[i];
        // the axis whose rotated image moves most along the mesh's fast axis, so that
        // neighbouring lanes add to neighbouring voxels (one L2 request per 64-byte segment
        // a wave instruction touches)
        const float a0 = fabsf(rr[0]), a1 = fabsf(rr[1]), a2 = fabsf(rr[2]);
        ax = a1 > a0 ? (a2 > a1 ? 2 : 1) : (a2 > a0 ? 2 : 0);
    }
    const int e0 = ax == 0 ? T.bx : (ax == 1 ? T.by : T.bz), e1 = ax == 0 ? T.by : T.bx;
    for (int v = threadIdx.x; v < vbox; v += kBlock) {
        const int i0 = v % e0, rest = v / e0, i1 = rest % e1, i2 = rest / e1;
        const int lx = ax == 0 ? i0 : i1, ly = ax == 1 ? i0 : (ax == 0 ? i1 : i2), lz = ax == 2 ? i0 : i2;
        const unsigned long long x = box[lx + T.bx * (ly + T.by * lz)];
        if (x) {
            int fs = T.bx0 + lx, ms = T.by0 + ly, ss = T.bz0 + lz;
            if (crot) {
                rotate_voxel(rr, S.P.c0, S.P.c1, S.P.c2, fs, ms, ss);
                if ((unsigned)fs >= (unsigned)S.P.sab || (unsigned)ms >= (unsigned)S.P.sab ||
                    (unsigned)ss >= (unsigned)S.P.sc) {
                    atomicAdd(dropped, 1ull);  // excluded by the caller's bound check
                    continue;
                }
            }
            atomicAdd(out + (size_t)fs + (size_t)ms * S.P.sab + (size_t)ss * S.P.sab * S.P.sab, x);
        }
    }
}

// ---- host planning ---------------------------------------------------------------

// The specialised pair path: AC_uw, symmetric (nearest-neighbour) binning, identity
// rotation; round_away needs |q * vox| < 2^23 and the cut bounds |q| by Max_Q.
inline bool fast_path(const PairSpec &S, bool ordered)
{
    // (its inline value is FloatToInt(float): the double-valued symmetric mode takes bin_sym)
    return S.kind == kKindAc3d && !ordered && S.P.rot_identity && !S.P.dval &&
           std::sqrt((double)S.P.cut_sq) * S.P.vox < 4194304.0;
}

// symmetric nearest-neighbour C(q): its voxels are round-half-away of the same
// fl(fl(ka - kb) * vox) products, so the tight boxes (and in-box tasks) apply
inline bool cq_tight(const PairSpec &S, bool ordered)
{
    return S.kind == kKindCq && !ordered && S.interp == 0 && std::sqrt((double)S.P.cut_sq) * S.P.vox < 4194304.0;
}

// interval of round(x) / of floor(x)+{0,1} for x in [lo, hi]
inline void nn_range(double &lo, double &hi)
{
    lo = std::floor(lo);
    hi = std::ceil(hi);
}
inline void lin_range(double &lo, double &hi)
{
    lo = std::floor(lo);
    hi = std::ceil(hi) + 1;
}

// Conservative voxel box (per mesh axis, relative to the centre) of the pairs
// whose q lies in [qlo, qhi].
void voxel_box(const PairSpec &S, const double qlo[3], const double qhi[3], double lo[3], double hi[3])
{
    const double vox = S.P.vox;
    double l[3], h[3];
    for (int d = 0; d < 3; d++) {
        l[d] = qlo[d] * vox;
        h[d] = qhi[d] * vox;
    }
    if (S.kind == kKindAc3d) {
        if (S.P.double_map) {
            for (int d = 0; d < 3; d++) (S.P.interp1 ? lin_range : nn_range)(l[d], h[d]);
        }
        if (!S.P.rot_identity) {
            double rl[3], rh[3];
            for (int i = 0; i < 3; i++) {
                rl[i] = rh[i] = 0;
                for (int j = 0; j < 3; j++) {
                    const double a = S.R[3 * i + j] * l[j], b = S.R[3 * i + j] * h[j];
                    rl[i] += std::min(a, b);
                    rh[i] += std::max(a, b);
                }
            }
            for (int d = 0; d < 3; d++) {
                l[d] = rl[d];
                h[d] = rh[d];
            }
        }
        for (int d = 0; d < 3; d++) {
            (S.P.interp2 ? lin_range : nn_range)(l[d], h[d]);
            lo[d] = l[d];
            hi[d] = h[d];
        }
    } else {
        const int ax[3] = {S.ka, S.kb, S.kc};
        for (int m = 0; m < 3; m++) {
            double a = l[ax[m]], b = h[ax[m]];
            (S.interp ? lin_range : nn_range)(a, b);
            lo[m] = a;
            hi[m] = b;
        }
    }
    for (int d = 0; d < 3; d++) {  // one voxel of slack for float rounding
        lo[d] -= 1;
        hi[d] += 1;
    }
}

}  // namespace
}  // namespace iac

struct iac_pair_plan {
    bool valid = false;
    // key
    const float *kmap = nullptr;
    uint32_t npix = 0;
    int cell_vox = 0;
    int box_cap = 0;  // LDS box capacity the tasks were sized for
    int region = 0;   // cells per task region side (0: whole grid)
    iac::PairSpec spec{};
    iac::CellGeom g{};
    std::vector<uint32_t> cell_ext;  // 6 keys per cell
    // product
    std::vector<iac::Task> tasks;
    iac::Task *d_tasks = nullptr;
    int pins = 0;  // frame queues holding this plan (never evicted while > 0)
};

namespace iac {
namespace {

bool same_spec(const PairSpec &a, const PairSpec &b)
{
    return std::memcmp(&a, &b, sizeof(PairSpec)) == 0;
}

// Tasks: one per cell offset delta and region of first cells i.  A region of
// `region` x `region` cells (0: the whole grid) narrows the q range of its pairs
// -- q along the beam grows with the distance from the centre -- so larger cells
// still fit their voxel box.
void build_tasks(iac_pair_plan &pl, bool ordered)
{
    const CellGeom &g = pl.g;
    const PairSpec &S = pl.spec;
    const int nc = g.ncu * g.ncv;
    std::vector<char> has(nc, 0);
    std::vector<double> kmn(3 * (size_t)nc), kmx(3 * (size_t)nc);
    for (int c = 0; c < nc; c++) {
        const uint32_t *e = &pl.cell_ext[6 * (size_t)c];
        if (e[0] == 0xffffffffu) continue;  // no pixel in this cell
        has[c] = 1;
        for (int d = 0; d < 3; d++) {
            kmn[3 * c + d] = unkey(e[d]);
            kmx[3 * c + d] = unkey(e[3 + d]);
        }
    }
    const double maxq = std::sqrt((double)S.P.cut_sq);
    const int cen[3] = {S.P.c0, S.P.c1, S.P.c2};
    const int dim[3] = {S.P.sab, S.P.sab, S.P.sc};
    const int reg = pl.region > 0 ? pl.region : std::max(g.ncu, g.ncv);
    const bool fast = fast_path(S, ordered), tight_cq = cq_tight(S, ordered);
    const int cax[3] = {S.ka, S.kb, S.kc};  // C(q): the record component each mesh axis bins
    pl.tasks.clear();
    for (int dv = ordered ? -(g.ncv - 1) : 0; dv < g.ncv; dv++) {
        for (int du = -(g.ncu - 1); du < g.ncu; du++) {
            if (!ordered && dv == 0 && du < 0) continue;  // positive half plane
            const int iu0 = std::max(0, du), iv0 = std::max(0, dv);
            const int nu = std::min(g.ncu, g.ncu + du) - iu0, nv = std::min(g.ncv, g.ncv + dv) - iv0;
            if (nu <= 0 || nv <= 0) continue;
            for (int rv = iv0; rv < iv0 + nv; rv += reg) {
                for (int ru = iu0; ru < iu0 + nu; ru += reg) {
                    Task t{};
                    t.du = du;
                    t.dv = dv;
                    t.iu0 = ru;
                    t.iv0 = rv;
                    t.nu = std::min(reg, iu0 + nu - ru);
                    t.nv = std::min(reg, iv0 + nv - rv);
                    double qlo[3] = {1e30, 1e30, 1e30}, qhi[3] = {-1e30, -1e30, -1e30};
                    bool any = false;
                    for (int iv = t.iv0; iv < t.iv0 + t.nv; iv++)
                        for (int iu = t.iu0; iu < t.iu0 + t.nu; iu++) {
                            const int ci = iv * g.ncu + iu, cj = (iv - dv) * g.ncu + (iu - du);
                            if (!has[ci] || !has[cj]) continue;
                            any = true;
                            for (int d = 0; d < 3; d++) {
                                qlo[d] = std::min(qlo[d], kmn[3 * ci + d] - kmx[3 * cj + d]);
                                qhi[d] = std::max(qhi[d], kmx[3 * ci + d] - kmn[3 * cj + d]);
                            }
                        }
                    if (!any) continue;
                    double dist2 = 0;  // all pairs of the task are cut if the q box lies beyond Max_Q
                    for (int d = 0; d < 3; d++) {
                        const double m = qlo[d] > 0 ? qlo[d] : (qhi[d] < 0 ? -qhi[d] : 0.0);
                        dist2 += m * m;
                    }
                    if (std::sqrt(dist2) > maxq * (1.0 + 1e-5) + 1e-7) continue;
                    // the farthest corner of the q box inside the cut, with margin for the float
                    // rounding of the kernel's q and |q|^2 (a relative 3.1e-7 at most)
                    double far2 = 0;
                    for (int d = 0; d < 3; d++) far2 += std::max(qlo[d] * qlo[d], qhi[d] * qhi[d]);
                    t.allin = far2 * (1.0 + 1e-5) < (double)S.P.cut_sq ? 1 : 0;
                    double lo[3], hi[3];
                    if (fast || tight_cq) {
                        // the pair bins at round_away(fl(fl(ka - kb) * vox)): within eps of x in [l, h]
                        // (two float roundings, |x| < 2^22), so round-half-away lands in
                        // [ceil(l - eps - 1/2), floor(h + eps + 1/2)]
                        for (int d = 0; d < 3; d++) {
                            const int c = tight_cq ? cax[d] : d;
                            const double l = qlo[c] * S.P.vox, h = qhi[c] * S.P.vox;
                            const double eps = 1e-3 + 1e-6 * std::max(std::fabs(l), std::fabs(h));
                            lo[d] = std::ceil(l - eps - 0.5);
                            hi[d] = std::floor(h + eps + 0.5);
                        }
                    } else {
                        voxel_box(S, qlo, qhi, lo, hi);
                    }
                    int ilo[3], ihi[3];
                    for (int d = 0; d < 3; d++) {
                        ilo[d] = (int)std::max(0.0, lo[d] + cen[d]);
                        ihi[d] = (int)std::min((double)dim[d] - 1, hi[d] + cen[d]);
                    }
                    const long vol = (long)(ihi[0] - ilo[0] + 1) * (ihi[1] - ilo[1] + 1) * (ihi[2] - ilo[2] + 1);
                    if (ihi[0] >= ilo[0] && ihi[1] >= ilo[1] && ihi[2] >= ilo[2] && vol <= pl.box_cap) {
                        t.inbox = 1;
                        for (int d = 0; d < 3; d++)
                            if (lo[d] + cen[d] < 0.0 || hi[d] + cen[d] > (double)dim[d] - 1) t.inbox = 0;
                        if (S.kind == kKindCq && !tight_cq) t.inbox = 0;  // in-box C(q) bins need the tight boxes
                        t.bx0 = ilo[0];
                        t.by0 = ilo[1];
                        t.bz0 = ilo[2];
                        t.bx = ihi[0] - ilo[0] + 1;
                        t.by = ihi[1] - ilo[1] + 1;
                        t.bz = ihi[2] - ilo[2] + 1;
                    }
                    pl.tasks.push_back(t);
                }
            }
        }
    }
}

iac_pair_plan *find_plan(iac_ctx *ctx, const float *kmap, uint32_t npix, int cell_vox, int box_cap, int region,
                         const PairSpec &S)
{
    for (auto *p : ctx->pair_plans)
        if (p->valid && p->kmap == kmap && p->npix == npix && p->cell_vox == cell_vox && p->box_cap == box_cap &&
            p->region == region && same_spec(p->spec, S))
            return p;
    return nullptr;
}

// Cell grid of the k-map (two host round trips for the extents) and the task
// table, from the context's cache when the geometry matches.
int get_plan(iac_ctx *ctx, hipStream_t st, const PairRun &R, int cell_vox, int box_cap, int region,
             iac_pair_plan **out)
{
    const PairSpec &S = R.spec;
    void *ext = nullptr;
    IAC_TRY(scratch(ctx, 3, (6 + 6 * (size_t)kMaxCells) * sizeof(uint32_t), &ext));
    uint32_t *d_ext = static_cast<uint32_t *>(ext);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, st, d_ext, 0xffffffffu, (size_t)3);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, st, d_ext + 3, 0u, (size_t)3);
    const unsigned eb = std::max(1u, std::min<unsigned>(1024, (R.npix + 255) / 256));
    // few blocks: the six extent words take one atomic per wave, all on the same addresses
    hipLaunchKernelGGL(kext_kernel, dim3(std::min(eb, 64u)), dim3(256), 0, st, R.d_kmap, R.npix, d_ext);
    IAC_HIP(hipGetLastError());
    uint32_t hext[6];
    IAC_HIP(hipMemcpyAsync(hext, d_ext, sizeof(hext), hipMemcpyDeviceToHost, st));
    IAC_HIP(hipStreamSynchronize(st));
    float kmin[3], ext3[3];
    for (int d = 0; d < 3; d++) {
        kmin[d] = unkey(hext[d]);
        ext3[d] = unkey(hext[3 + d]) - kmin[d];
    }
    CellGeom g{};
    int order[3] = {0, 1, 2};
    std::sort(order, order + 3, [&](int a, int b) { return ext3[a] > ext3[b]; });
    g.au = std::min(order[0], order[1]);
    g.av = std::max(order[0], order[1]);
    // cell_vox = u side + 1000 * v side (voxels); cells may be longer along u
    double wu = (double)(cell_vox % 1000) / (double)S.P.vox;
    double wv = (double)(cell_vox / 1000 ? cell_vox / 1000 : cell_vox % 1000) / (double)S.P.vox;
    auto ncells = [&](int a, double w) { return std::max(1, (int)std::ceil(ext3[a] / w * (1 + 1e-6))); };
    while ((long)ncells(g.au, wu) * ncells(g.av, wv) > kMaxCells) {
        wu *= 1.25;
        wv *= 1.25;
    }
    g.inv_wu = (float)(1.0 / wu);
    g.inv_wv = (float)(1.0 / wv);
    g.u0 = kmin[g.au];
    g.v0 = kmin[g.av];
    g.ncu = ncells(g.au, wu);
    g.ncv = ncells(g.av, wv);
    const int nc = g.ncu * g.ncv;
    hipLaunchKernelGGL(init_cell_ext_kernel, dim3((nc + 255) / 256), dim3(256), 0, st, d_ext + 6, nc);
    const size_t lds = 6 * sizeof(uint32_t) * (size_t)nc;
    if (lds <= 48 * 1024) {
        const unsigned lb = std::max(1u, std::min<unsigned>(512, (R.npix + 4095) / 4096));
        hipLaunchKernelGGL(cell_ext_lds_kernel, dim3(lb), dim3(256), lds, st, R.d_kmap, R.npix, g, d_ext + 6);
    } else {
        hipLaunchKernelGGL(cell_ext_kernel, dim3(eb), dim3(256), 0, st, R.d_kmap, R.npix, g, d_ext + 6);
    }
    IAC_HIP(hipGetLastError());
    std::vector<uint32_t> cext(6 * (size_t)nc);
    IAC_HIP(hipMemcpyAsync(cext.data(), d_ext + 6, cext.size() * 4, hipMemcpyDeviceToHost, st));
    IAC_HIP(hipStreamSynchronize(st));
    iac_pair_plan *pl = find_plan(ctx, R.d_kmap, R.npix, cell_vox, box_cap, region, S);
    if (pl && (std::memcmp(&pl->g, &g, sizeof(g)) != 0 || pl->cell_ext != cext)) pl = nullptr;
    if (!pl) {
        // reuse the oldest unpinned slot beyond four cached plans
        if (ctx->pair_plans.size() >= 4) {
            for (auto it = ctx->pair_plans.begin(); it != ctx->pair_plans.end(); ++it) {
                if ((*it)->pins > 0) continue;
                if ((*it)->d_tasks) IAC_HIP(hipFree((*it)->d_tasks));
                delete *it;
                ctx->pair_plans.erase(it);
                break;
            }
        }
        pl = new iac_pair_plan();
        ctx->pair_plans.push_back(pl);
        pl->kmap = R.d_kmap;
        pl->npix = R.npix;
        pl->cell_vox = cell_vox;
        pl->box_cap = box_cap;
        pl->region = region;
        pl->spec = S;
        pl->g = g;
        pl->cell_ext.swap(cext);
        build_tasks(*pl, R.ordered);
        if (pl->tasks.size() > kMaxGridBlocks)
            return fail(IAC_ERR_LIMIT, "pair engine: more than 2^22 tasks (raise the task region or the cell size)");
        if (!pl->tasks.empty()) {
            IAC_HIP(hipMalloc(&pl->d_tasks, pl->tasks.size() * sizeof(Task)));
            IAC_HIP(hipMemcpy(pl->d_tasks, pl->tasks.data(), pl->tasks.size() * sizeof(Task),
                              hipMemcpyHostToDevice));
        }
        pl->valid = true;
    }
    *out = pl;
    return IAC_OK;
}

// Frame chunks inside each group, balanced by pair work: a chunk closes once its
// frames' n^2 reach fchunk times the mean per frame, or at 4 fchunk frames.
std::vector<Chunk> make_chunks(const uint64_t *off, uint32_t nframes, const uint32_t *group_off, uint32_t ngroups,
                               uint32_t fchunk, const float *h_rots = nullptr)
{
    double total = 0;
    for (uint32_t f = 0; f < nframes; f++) {
        const double n = (double)(off[f + 1] - off[f]);
        total += n * n;
    }
    // fchunk mean frames of pair work, at most 2^28 pairs: dense frames (C(q) images hold
    // P^2 pairs) become chunks of one frame, so a run still spreads over many blocks
    const double budget = std::max(1.0, std::min(total / std::max(1u, nframes) * fchunk, 268435456.0));
    const uint32_t fmax = std::min<uint32_t>(4 * fchunk, 4096);  // items per chunk < 2^24 (pairs_kernel)
    std::vector<Chunk> chunks;
    for (uint32_t gi = 0; gi < ngroups; gi++) {
        uint32_t f0 = group_off[gi];
        double acc = 0;
        for (uint32_t f = group_off[gi]; f < group_off[gi + 1]; f++) {
            const double n = (double)(off[f + 1] - off[f]);
            acc += n * n;
            // per-frame rotations: a chunk ends where the next frame's rotation differs
            const bool rot_break = h_rots && f + 1 < group_off[gi + 1] &&
                                   std::memcmp(h_rots + 9 * (size_t)f0, h_rots + 9 * (size_t)(f + 1), 36) != 0;
            if (acc >= budget || f + 1 - f0 >= fmax || rot_break) {
                chunks.push_back(Chunk{f0, f + 1, gi});
                f0 = f + 1;
                acc = 0;
            }
        }
        if (f0 < group_off[gi + 1]) chunks.push_back(Chunk{f0, group_off[gi + 1], gi});
    }
    return chunks;
}

// The pair kernel reads a chunk's records at 32-bit byte offsets from its first record: every
// chunk spans fewer than 2^28 records (the chunk budget keeps a chunk's other frames to a few
// million records, so only a single frame of ~2^28 records can break this).
int check_chunks(const uint64_t *off, const std::vector<Chunk> &chunks)
{
    for (const Chunk &c : chunks)
        if (off[c.f1] - off[c.f0] >= (1ull << 28))
            return fail(IAC_ERR_LIMIT, "pair engine: a frame chunk of 2^28 or more records (split the frame)");
    return IAC_OK;
}

struct Bufs {
    float4 *srec;
    uint32_t *cs;
    unsigned long long *H;  // symmetric modes: ngroups half-mesh accumulators
    uint32_t *d_goff;
    unsigned long long *d_self;
    uint32_t *d_flag;  // 1 while every weight of the run is a small integer (after d_self)
    Chunk *d_chunks;
    uint32_t nchunks;
    const Task *d_tasks;
    uint32_t ntasks;
    CellGeom g;
    int box_cap;
};

// per frame, the records of its fullest cell: after the frames' cell starts
uint32_t *cell_max(const PairRun &R, const Bufs &b)
{
    return b.cs + (size_t)R.nframes * (b.g.ncu * b.g.ncv + 1);
}

template <int K, bool O, int BOX>
void launch_pairs_box(int pf, bool fast, int walk, dim3 grid, hipStream_t st, const PairRun &R, const Bufs &b,
                      unsigned long long *target, size_t nvox, unsigned long long *dropped)
{
    auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, grid, dim3(kBlock), 0, st, b.srec, R.d_frame_off, b.cs, b.d_chunks, b.d_tasks,
                           b.ntasks, b.g, R.spec, target, nvox, dropped, b.d_flag, R.d_rots, cell_max(R, b));
    };
    constexpr bool kFast = K == kKindAc3d && !O;
    if ((O || pf == 1) && fast) {
        if (walk) go(pairs_kernel<K, O, 0, kFast, BOX>);
        else go(pairs_kernel<K, O, 1, kFast, BOX>);
    } else if (O || pf == 1 || BOX != kBoxMax) {
        if (walk) go(pairs_kernel<K, O, 0, false, BOX>);
        else go(pairs_kernel<K, O, 1, false, BOX>);
    } else if (pf == 4) {
        go(pairs_kernel<K, O, O ? 1 : 4, false, BOX>);
    } else {
        go(pairs_kernel<K, O, O ? 1 : 8, false, BOX>);
    }
}

// walk: lanes on consecutive pairs (1) or a contiguous pair range per thread (0)
template <int K, bool O>
void launch_pairs(int pf, int walk, dim3 grid, hipStream_t st, const PairRun &R, const Bufs &b,
                  unsigned long long *target, size_t nvox, unsigned long long *dropped)
{
    const bool fast = fast_path(R.spec, O);
    if (b.box_cap == kBoxSmall)
        launch_pairs_box<K, O, kBoxSmall>(pf, fast, walk, grid, st, R, b, target, nvox, dropped);
    else if (b.box_cap == kBoxTiny)
        launch_pairs_box<K, O, kBoxTiny>(pf, fast, walk, grid, st, R, b, target, nvox, dropped);
    else
        launch_pairs_box<K, O, kBoxMax>(pf, fast, walk, grid, st, R, b, target, nvox, dropped);
}

// Everything after planning, enqueue only (no host synchronisation, no
// allocation, fills by kernel): capturable into a graph.
bool walks(iac_ctx *ctx, const PairRun &R);

int enqueue(iac_ctx *ctx, hipStream_t st, const PairRun &R, const Bufs &b, bool timed)
{
    const PairSpec &S = R.spec;
    const bool ordered = R.ordered;
    const size_t nvox = (size_t)S.P.sab * S.P.sab * S.P.sc;
    const unsigned zb = (unsigned)std::max<size_t>(1, std::min<size_t>(4096, (nvox * R.ngroups + 255) / 256));
    if (!ordered) {
        hipLaunchKernelGGL(fill_u64_kernel, dim3(zb), dim3(256), 0, st, b.H, nvox * R.ngroups);
        hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(256), 0, st, b.d_self, (size_t)R.ngroups);
    }
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, st, b.d_flag, 1u, (size_t)1);
    if (S.kind == kKindAc3d) {
        if (ordered)
            hipLaunchKernelGGL((cell_sort_kernel<kKindAc3d, true>), dim3(R.nframes), dim3(kBlock), 0, st, R.rec,
                               R.d_frame_off, b.g, b.srec, b.cs, S, b.d_goff, R.ngroups, b.d_self, b.d_flag, cell_max(R, b));
        else
            hipLaunchKernelGGL((cell_sort_kernel<kKindAc3d, false>), dim3(R.nframes), dim3(kBlock), 0, st, R.rec,
                               R.d_frame_off, b.g, b.srec, b.cs, S, b.d_goff, R.ngroups, b.d_self, b.d_flag, cell_max(R, b));
    } else {
        if (ordered)
            hipLaunchKernelGGL((cell_sort_kernel<kKindCq, true>), dim3(R.nframes), dim3(kBlock), 0, st, R.rec,
                               R.d_frame_off, b.g, b.srec, b.cs, S, b.d_goff, R.ngroups, b.d_self, b.d_flag, cell_max(R, b));
        else
            hipLaunchKernelGGL((cell_sort_kernel<kKindCq, false>), dim3(R.nframes), dim3(kBlock), 0, st, R.rec,
                               R.d_frame_off, b.g, b.srec, b.cs, S, b.d_goff, R.ngroups, b.d_self, b.d_flag, cell_max(R, b));
    }
    IAC_HIP(hipGetLastError());
    auto *target = ordered ? reinterpret_cast<unsigned long long *>(R.d_mesh) : b.H;
    if (b.ntasks && b.nchunks) {
        if (timed) IAC_TRY(timer_begin(ctx));
        // ordered (linear) binning holds eight corner adds per pair in registers: no batching there
        const int pf = ordered || ctx->pair_prefetch == 0 || ctx->pair_prefetch == 1 ? 1 : (ctx->pair_prefetch == 4 ? 4 : 8);
        // block = chunk * ntasks + task; whole chunks per launch, the slice's chunks by pointer
        const int walk = walks(ctx, R) ? 1 : 0;
        const uint32_t per = std::max<uint32_t>(1, kMaxGridBlocks / b.ntasks);
        for (uint32_t c0 = 0; c0 < b.nchunks; c0 += per) {
            Bufs bs = b;
            bs.d_chunks = b.d_chunks + c0;
            const dim3 grid(b.ntasks * std::min(per, b.nchunks - c0));
            if (S.kind == kKindAc3d) {
                if (ordered) launch_pairs<kKindAc3d, true>(pf, walk, grid, st, R, bs, target, nvox, ctx->d_counters);
                else launch_pairs<kKindAc3d, false>(pf, walk, grid, st, R, bs, target, nvox, ctx->d_counters);
            } else {
                if (ordered) launch_pairs<kKindCq, true>(pf, walk, grid, st, R, bs, target, nvox, ctx->d_counters);
                else launch_pairs<kKindCq, false>(pf, walk, grid, st, R, bs, target, nvox, ctx->d_counters);
            }
            IAC_HIP(hipGetLastError());
        }
        if (timed) IAC_TRY(timer_end(ctx));
    }
    if (!ordered) {
        const size_t centre = (size_t)S.P.c0 + (size_t)S.P.c1 * S.P.sab + (size_t)S.P.c2 * S.P.sab * S.P.sab;
        hipLaunchKernelGGL(symmetrize_kernel, dim3(zb), dim3(256), 0, st, b.H,
                           reinterpret_cast<unsigned long long *>(R.d_mesh), nvox, R.ngroups, centre, b.d_self);
        IAC_HIP(hipGetLastError());
    }
    return IAC_OK;
}

// Defaults from the tuner (bench.py sweeps and the ordered-mode measurements):
//   AC_uw nearest neighbour, identity rotation: lanes on consecutive pairs, 6-voxel cells,
//     1920-voxel boxes (1024^2 sparse and lognormal workloads);
//   AC_uw linear (ordered): lanes on consecutive pairs, 5-voxel cells, 4096-voxel boxes
//     (64 LDS adds per pair with a linear first level: LDS-atomic bound);
//   C(q) nearest neighbour: lanes on consecutive pairs, 3-voxel cells, 1920-voxel boxes,
//     4 x 4-cell task regions (dense 256^2 images);
//   C(q) linear: a pair range per thread, 4-voxel cells, 4096-voxel boxes, regions of 4.
bool walks(iac_ctx *ctx, const PairRun &R)
{
    if (ctx->pair_mode == 1) return true;
    if (ctx->pair_mode == 2) return false;
    // C(q) of dense frames: a pair range per thread, whose runs of one voxel add in a register
    return R.spec.kind != kKindCq;
}
int cell_vox_of(iac_ctx *ctx, const PairRun &R)
{
    int dflt = 4;
    if (R.spec.kind == kKindCq && !R.ordered) dflt = 3;
    else if (walks(ctx, R)) dflt = R.ordered ? 5 : (fast_path(R.spec, R.ordered) ? 6 : 4);
    const int u = ctx->cell_vox > 0 ? ctx->cell_vox : dflt;
    return u + 1000 * (ctx->cell_vox_v > 0 ? ctx->cell_vox_v : u);
}
int box_cap_of(iac_ctx *ctx, const PairRun &R)
{
    if (ctx->box_cap == kBoxMax || ctx->box_cap == kBoxTiny) return ctx->box_cap;
    if (ctx->box_cap == kBoxSmall || ctx->box_cap == 2048) return kBoxSmall;
    return R.ordered ? kBoxMax : kBoxSmall;
}
int region_of(iac_ctx *ctx, const PairRun &R)
{
    if (ctx->task_region > 0) return ctx->task_region;
    if (R.spec.kind != kKindCq) return 0;
    return R.ordered ? 4 : 2;
}
uint32_t fchunk_of(iac_ctx *ctx) { return ctx->frame_chunk > 0 ? ctx->frame_chunk : 32; }

// byte sizes of the per-run buffers
struct Sizes {
    size_t srec, cs, h, goff, self, chunks;
};
Sizes sizes_of(const PairRun &R, int nc, size_t nchunks)
{
    const size_t nvox = (size_t)R.spec.P.sab * R.spec.P.sab * R.spec.P.sc;
    auto up = [](size_t x) { return (x + 255) / 256 * 256; };
    return Sizes{up(R.n_total * sizeof(float4)),
                 up((size_t)R.nframes * (nc + 2) * sizeof(uint32_t)),  // cell starts, then the fullest cells
                 R.ordered ? 0 : nvox * R.ngroups * sizeof(unsigned long long),
                 up((R.ngroups + 1) * sizeof(uint32_t)),
                 up(R.ngroups * sizeof(unsigned long long) + sizeof(uint32_t)),
                 up(std::max<size_t>(1, nchunks) * sizeof(Chunk))};
}

}  // namespace

void free_pair_plans(iac_ctx *ctx)
{
    for (auto *p : ctx->pair_plans) {
        if (p->d_tasks) (void)hipFree(p->d_tasks);
        delete p;
    }
    ctx->pair_plans.clear();
}

int pairs_run(iac_ctx *ctx, const PairRun &R)
{
    hipStream_t st = ctx->stream;
    if (R.n_total >= (1ull << 32)) return fail(IAC_ERR_LIMIT, "pair engine: at most 2^32 - 1 records per call");
    iac_pair_plan *pl = nullptr;
    IAC_TRY(get_plan(ctx, st, R, cell_vox_of(ctx, R), box_cap_of(ctx, R), region_of(ctx, R), &pl));
    std::vector<uint64_t> off(R.nframes + 1);
    IAC_HIP(hipMemcpyAsync(off.data(), R.d_frame_off, off.size() * 8, hipMemcpyDeviceToHost, st));
    IAC_HIP(hipStreamSynchronize(st));
    const std::vector<Chunk> chunks = make_chunks(off.data(), R.nframes, R.h_group_off, R.ngroups, fchunk_of(ctx), R.h_rots);
    IAC_TRY(check_chunks(off.data(), chunks));
    const int nc = pl->g.ncu * pl->g.ncv;
    const Sizes z = sizes_of(R, nc, chunks.size());
    // scratch: sorted records + cell offsets (slot 1), H + group offsets + self + chunks (slot 2)
    void *s1 = nullptr, *s2 = nullptr;
    IAC_TRY(scratch(ctx, 1, z.srec + z.cs, &s1));
    IAC_TRY(scratch(ctx, 2, z.h + z.goff + z.self + z.chunks, &s2));
    Bufs b{};
    b.srec = static_cast<float4 *>(s1);
    b.cs = reinterpret_cast<uint32_t *>(static_cast<char *>(s1) + z.srec);
    char *p2 = static_cast<char *>(s2);
    b.H = reinterpret_cast<unsigned long long *>(p2);
    b.d_goff = reinterpret_cast<uint32_t *>(p2 + z.h);
    b.d_self = reinterpret_cast<unsigned long long *>(p2 + z.h + z.goff);
    b.d_flag = reinterpret_cast<uint32_t *>(b.d_self + R.ngroups);
    b.d_chunks = reinterpret_cast<Chunk *>(p2 + z.h + z.goff + z.self);
    b.nchunks = (uint32_t)chunks.size();
    b.d_tasks = pl->d_tasks;
    b.ntasks = (uint32_t)pl->tasks.size();
    b.g = pl->g;
    b.box_cap = pl->box_cap;
    IAC_HIP(hipMemcpyAsync(b.d_goff, R.h_group_off, (R.ngroups + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    if (!chunks.empty())
        IAC_HIP(hipMemcpyAsync(b.d_chunks, chunks.data(), chunks.size() * sizeof(Chunk), hipMemcpyHostToDevice, st));
    IAC_HIP(hipStreamSynchronize(st));  // the host arrays above are released on return
    return enqueue(ctx, st, R, b, true);
}

// ---- prepared runs -----------------------------------------------------------

struct PairBatch {
    iac_ctx *ctx = nullptr;
    PairRun R{};
    std::vector<uint32_t> group_off;
    void *mem = nullptr;     // srec | cs | H | goff | self | chunks | tasks
    Bufs b{};
};

int pairs_prepare(iac_ctx *ctx, const PairRun &R0, PairBatch **out)
{
    *out = nullptr;
    hipStream_t st = ctx->stream;
    if (R0.n_total >= (1ull << 32)) return fail(IAC_ERR_LIMIT, "pair engine: at most 2^32 - 1 records per call");
    iac_pair_plan *pl = nullptr;
    IAC_TRY(get_plan(ctx, st, R0, cell_vox_of(ctx, R0), box_cap_of(ctx, R0), region_of(ctx, R0), &pl));
    std::vector<uint64_t> off(R0.nframes + 1);
    IAC_HIP(hipMemcpyAsync(off.data(), R0.d_frame_off, off.size() * 8, hipMemcpyDeviceToHost, st));
    IAC_HIP(hipStreamSynchronize(st));
    const std::vector<Chunk> chunks = make_chunks(off.data(), R0.nframes, R0.h_group_off, R0.ngroups, fchunk_of(ctx));
    IAC_TRY(check_chunks(off.data(), chunks));
    auto *B = new PairBatch();
    B->ctx = ctx;
    B->R = R0;
    B->group_off.assign(R0.h_group_off, R0.h_group_off + R0.ngroups + 1);
    B->R.h_group_off = B->group_off.data();
    const int nc = pl->g.ncu * pl->g.ncv;
    const Sizes z = sizes_of(B->R, nc, chunks.size());
    const size_t tbytes = std::max<size_t>(1, pl->tasks.size()) * sizeof(Task);
    const size_t total = z.srec + z.cs + z.h + z.goff + z.self + z.chunks + tbytes;
    hipError_t e = hipMalloc(&B->mem, total);
    if (e != hipSuccess) {
        delete B;
        return hip_fail(e, "pairs_prepare allocation", __FILE__, __LINE__);
    }
    char *p = static_cast<char *>(B->mem);
    Bufs &b = B->b;
    b.srec = reinterpret_cast<float4 *>(p);
    b.cs = reinterpret_cast<uint32_t *>(p + z.srec);
    b.H = reinterpret_cast<unsigned long long *>(p + z.srec + z.cs);
    b.d_goff = reinterpret_cast<uint32_t *>(p + z.srec + z.cs + z.h);
    b.d_self = reinterpret_cast<unsigned long long *>(p + z.srec + z.cs + z.h + z.goff);
    b.d_flag = reinterpret_cast<uint32_t *>(b.d_self + R0.ngroups);
    b.d_chunks = reinterpret_cast<Chunk *>(p + z.srec + z.cs + z.h + z.goff + z.self);
    Task *tasks = reinterpret_cast<Task *>(p + z.srec + z.cs + z.h + z.goff + z.self + z.chunks);
    b.nchunks = (uint32_t)chunks.size();
    b.d_tasks = tasks;
    b.ntasks = (uint32_t)pl->tasks.size();
    b.g = pl->g;
    b.box_cap = pl->box_cap;
    e = hipMemcpyAsync(b.d_goff, B->group_off.data(), B->group_off.size() * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && !chunks.empty())
        e = hipMemcpyAsync(b.d_chunks, chunks.data(), chunks.size() * sizeof(Chunk), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && !pl->tasks.empty())
        e = hipMemcpyAsync(tasks, pl->tasks.data(), pl->tasks.size() * sizeof(Task), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        (void)hipFree(B->mem);
        delete B;
        return hip_fail(e, "pairs_prepare upload", __FILE__, __LINE__);
    }
    *out = B;
    return IAC_OK;
}

int pairs_enqueue(PairBatch *B, hipStream_t st, uint64_t *d_mesh)
{
    PairRun R = B->R;
    R.d_mesh = d_mesh;
    // the pair kernel is timed on the context's own stream (not inside a capture)
    return enqueue(B->ctx, st, R, B->b, st == B->ctx->stream);
}

void pairs_batch_free(PairBatch *B)
{
    if (!B) return;
    if (B->mem) (void)hipFree(B->mem);
    delete B;
}

// ---- frame queues --------------------------------------------------------------

struct PairQueue {
    iac_ctx *ctx = nullptr;
    PairRun R0{};
    iac_pair_plan *pl = nullptr;
    uint64_t max_records = 0;
    uint32_t max_frames = 0;
    void *mem = nullptr;        // srec | cs | H | goff | self | chunks (device)
    Bufs b{};
    Chunk *h_chunks = nullptr;  // pinned: the chunk table and group offsets of the last run
    uint32_t *h_goff = nullptr;
    hipEvent_t staged = nullptr;  // the last run's uploads from h_chunks / h_goff are done
};

int pairs_queue_create(iac_ctx *ctx, const PairRun &R0, uint64_t max_records, uint32_t max_frames, PairQueue **out)
{
    *out = nullptr;
    if (max_records == 0 || max_records >= (1ull << 32) || max_frames == 0)
        return fail(IAC_ERR_LIMIT, "pair engine queue: 1 .. 2^32 - 1 records and at least one frame");
    iac_pair_plan *pl = nullptr;
    IAC_TRY(get_plan(ctx, ctx->stream, R0, cell_vox_of(ctx, R0), box_cap_of(ctx, R0), region_of(ctx, R0), &pl));
    auto *Q = new PairQueue();
    Q->ctx = ctx;
    Q->R0 = R0;
    Q->R0.ngroups = 1;
    Q->R0.n_total = max_records;
    Q->R0.nframes = max_frames;
    Q->max_records = max_records;
    Q->max_frames = max_frames;
    const int nc = pl->g.ncu * pl->g.ncv;
    const Sizes z = sizes_of(Q->R0, nc, max_frames);
    hipError_t e = hipMalloc(&Q->mem, z.srec + z.cs + z.h + z.goff + z.self + z.chunks);
    if (e == hipSuccess)
        e = hipHostMalloc((void **)&Q->h_chunks, (size_t)max_frames * sizeof(Chunk) + 64, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&Q->staged, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(Q->staged, ctx->stream);
    if (e != hipSuccess) {
        if (Q->mem) (void)hipFree(Q->mem);
        if (Q->h_chunks) (void)hipHostFree(Q->h_chunks);
        if (Q->staged) (void)hipEventDestroy(Q->staged);
        delete Q;
        return hip_fail(e, "pair engine queue allocation", __FILE__, __LINE__);
    }
    Q->h_goff = reinterpret_cast<uint32_t *>(Q->h_chunks + max_frames);
    char *p = static_cast<char *>(Q->mem);
    Bufs &b = Q->b;
    b.srec = reinterpret_cast<float4 *>(p);
    b.cs = reinterpret_cast<uint32_t *>(p + z.srec);
    b.H = reinterpret_cast<unsigned long long *>(p + z.srec + z.cs);
    b.d_goff = reinterpret_cast<uint32_t *>(p + z.srec + z.cs + z.h);
    b.d_self = reinterpret_cast<unsigned long long *>(p + z.srec + z.cs + z.h + z.goff);
    b.d_flag = reinterpret_cast<uint32_t *>(b.d_self + 1);
    b.d_chunks = reinterpret_cast<Chunk *>(p + z.srec + z.cs + z.h + z.goff + z.self);
    b.d_tasks = pl->d_tasks;
    b.ntasks = (uint32_t)pl->tasks.size();
    b.g = pl->g;
    b.box_cap = pl->box_cap;
    pl->pins++;
    Q->pl = pl;
    *out = Q;
    return IAC_OK;
}

int pairs_queue_run(PairQueue *Q, const PairRun &R, const uint64_t *h_off)
{
    if (R.nframes == 0 || R.n_total == 0) return IAC_OK;
    if (R.nframes > Q->max_frames || R.n_total > Q->max_records)
        return fail(IAC_ERR_LIMIT, "pair engine queue: run larger than the queue");
    hipStream_t st = Q->ctx->stream;
    const uint32_t goff[2] = {0, R.nframes};
    const std::vector<Chunk> chunks = make_chunks(h_off, R.nframes, goff, 1, fchunk_of(Q->ctx), R.h_rots);
    IAC_TRY(check_chunks(h_off, chunks));
    // the previous run's uploads read the pinned table: wait for them before rewriting it
    IAC_HIP(hipEventSynchronize(Q->staged));
    std::copy(chunks.begin(), chunks.end(), Q->h_chunks);
    Q->h_goff[0] = 0;
    Q->h_goff[1] = R.nframes;
    IAC_HIP(hipMemcpyAsync(Q->b.d_goff, Q->h_goff, 2 * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    if (!chunks.empty())
        IAC_HIP(hipMemcpyAsync(Q->b.d_chunks, Q->h_chunks, chunks.size() * sizeof(Chunk), hipMemcpyHostToDevice, st));
    IAC_HIP(hipEventRecord(Q->staged, st));
    PairRun RR = Q->R0;
    RR.rec = R.rec;
    RR.d_frame_off = R.d_frame_off;
    RR.nframes = R.nframes;
    RR.n_total = R.n_total;
    RR.h_group_off = goff;
    RR.ngroups = 1;
    RR.d_mesh = R.d_mesh;
    RR.d_rots = R.d_rots;
    RR.h_rots = R.h_rots;
    Bufs b = Q->b;
    b.nchunks = (uint32_t)chunks.size();
    return enqueue(Q->ctx, st, RR, b, true);
}

void pairs_queue_free(PairQueue *Q)
{
    if (!Q) return;
    (void)hipStreamSynchronize(Q->ctx->stream);
    if (Q->pl) Q->pl->pins--;
    if (Q->mem) (void)hipFree(Q->mem);
    if (Q->h_chunks) (void)hipHostFree(Q->h_chunks);
    if (Q->staged) (void)hipEventDestroy(Q->staged);
    delete Q;
}

}  // namespace iac
