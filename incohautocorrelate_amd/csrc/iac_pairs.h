// iac_pairs.h -- interface of the LDS cell-box pair engine (iac_pairs.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "iac_binning.h"

struct iac_ctx;

namespace iac {

enum { kKindAc3d = 0, kKindCq = 1 };

// What one pair contributes.  Trivially copyable; zero it before filling (it
// is compared bytewise as a plan-cache key).
struct PairSpec {
    int kind;        // kKindAc3d: ACMesh binning, kKindCq: AutoCorr_CQ_small binning
    int ka, kb, kc;  // C(q): k components feeding fs, ms, ss (MeshShape k_A, k_B, k_C)
    int interp;      // C(q): interpolation (0 nearest, 1 linear)
    int reserved;
    double mult;     // C(q): fixed-point multiplicator, value (uint64)((double)(Ia*Ib) * mult)
    float R[9];      // AC_uw: rotation shared by all frames (row major)
    BinParams P;     // cut, voxel scale, centres, sizes, AC_uw interpolation and double map
};

struct PairRun {
    PairSpec spec;
    bool ordered;                 // linear modes: every ordered pair, no mirror pass
    const float *d_kmap;          // [npix][3], defines the cell grid
    uint32_t npix;
    const float4 *rec;            // records of all frames
    const uint64_t *d_frame_off;  // [nframes + 1]
    uint32_t nframes;
    uint64_t n_total;
    const uint32_t *h_group_off;  // host [ngroups + 1]: frames of group g are [off[g], off[g+1])
    uint32_t ngroups;
    uint64_t *d_mesh;             // ngroups meshes of size_ab^2 * size_c uint64, accumulated into
    // AC_uw with a double map and nearest-neighbour levels: per-frame row-major rotations
    // ([nframes][9], device, and the same on the host for chunking), applied to the
    // first-level voxel when a box flushes; the pairs bin under the identity (spec.R).
    // The caller has checked that no rotated voxel leaves the mesh.
    const float *d_rots = nullptr;
    const float *h_rots = nullptr;
};

int pairs_run(iac_ctx *ctx, const PairRun &run);

// Prepared runs: the plan, the frame chunks and every buffer made once (run.rec,
// run.d_frame_off and the kmap are kept by pointer; their contents may change
// between launches as long as sizes and the frame count stay); pairs_enqueue
// then only enqueues kernels (no synchronisation, no allocation, no memset),
// so it can be captured into a graph.  run.d_mesh is ignored; each enqueue names it.
struct PairBatch;
int pairs_prepare(iac_ctx *ctx, const PairRun &run, PairBatch **out);
int pairs_enqueue(PairBatch *batch, hipStream_t stream, uint64_t *d_mesh);
void pairs_batch_free(PairBatch *batch);

// Frame queues: buffers for up to max_records records and max_frames frames made once,
// the plan of R0 (spec, ordered, k-map) pinned in the context's cache.  pairs_queue_run
// bins the frames of run (rec, d_frame_off, nframes, n_total, d_rots / h_rots, d_mesh)
// given their offsets on the host too (h_off, nframes + 1): it builds the chunk table on
// the host, uploads it from pinned memory and enqueues, with no allocation and no host
// synchronisation beyond waiting for the previous run's upload of that table.
struct PairQueue;
int pairs_queue_create(iac_ctx *ctx, const PairRun &R0, uint64_t max_records, uint32_t max_frames, PairQueue **out);
int pairs_queue_run(PairQueue *queue, const PairRun &run, const uint64_t *h_off);
void pairs_queue_free(PairQueue *queue);

}  // namespace iac
