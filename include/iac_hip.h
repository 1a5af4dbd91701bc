/*
 * iac_hip.h -- C ABI of the MI355X (gfx950) pixel-pair autocorrelation path.
 *
 * Plain C: no C++ or torch types cross this boundary.  Device pointers
 * (prefix d_) are HIP device addresses owned by the caller unless stated;
 * host pointers (prefix h_ or none) are read during the call only.  Every
 * entry point returns IAC_OK (0) or a negative iac error code; iac_strerror()
 * names the code and iac_last_error() gives the detail of the last failure on
 * the calling thread.  An iac_ctx is bound to one device and one HIP stream and
 * must be used by one host thread at a time; work is enqueued on the stream
 * and results are visible after iac_ctx_sync() (or any blocking copy).
 *
 * The entry points replace the OpenCL dispatch blocks of the reference host
 * methods (paths relative to the reference root, IAC = IncohAutoCorrelate/).
 */
#ifndef IAC_HIP_H
#define IAC_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IAC_ABI_VERSION 1

/* error codes */
#define IAC_OK 0
#define IAC_ERR_INVALID_ARG (-1)  /* null pointer, zero size, bad enum */
#define IAC_ERR_EVEN_MESH (-2)    /* IAC/ACMesh.cpp:76-81 throws std::invalid_argument */
#define IAC_ERR_NO_DEVICE (-3)    /* IAC/Settings.cpp:348-360 (no usable device) */
#define IAC_ERR_HIP (-4)          /* a HIP runtime call failed (checkErr, IAC/Settings.cpp:322-329) */
#define IAC_ERR_OOM (-5)          /* device allocation failed */
#define IAC_ERR_SHAPE (-6)        /* shape mismatch, e.g. IAC/Detector.cpp:1324-1329 */
#define IAC_ERR_PRECONDITION (-7) /* e.g. no hit list / no event, IAC/Detector.cpp:1012-1023 */
#define IAC_ERR_LIMIT (-8)        /* a size exceeds a documented maximum */

/* Settings::Interpolation, IAC/Settings.h:158 */
#define IAC_NEAREST 0
#define IAC_LINEAR 1

/* ACMesh::MeshShape, IAC/ACMesh.h:21-29 (same fields, same meaning) */
typedef struct {
    uint32_t k_a, k_b, k_c;   /* axis alignment, k_a largest extent */
    uint32_t size_ab, size_c; /* edge sizes including the 2 padding voxels */
    uint32_t center[3];       /* voxel holding q = 0 */
    float dq_per_voxel;
    uint32_t voxel_per_dq;
    float max_q;
} iac_mesh_shape;

typedef struct iac_ctx iac_ctx;

const char *iac_strerror(int code);
const char *iac_last_error(void);
int iac_abi_version(void);

/* ---- device, context, memory --------------------------------------------------- */
/* Replaces Settings::SetUp_OpenCL device discovery, IAC/Settings.cpp:331-430. */
int iac_device_count(int *count);
/* Replaces Settings::OCL_ReserveDevice + the per-call cl::CommandQueue,
 * IAC/Settings.cpp:432-455, IAC/Detector.cpp:1074-1090.  Creates a context on
 * `device` with a stream of default priority. */
int iac_ctx_create(int device, iac_ctx **out);
/* Replaces Settings::OCL_FreeDevice, IAC/Settings.cpp:457-470. */
int iac_ctx_destroy(iac_ctx *ctx);
/* Enqueue on an external HIP stream (hipStream_t) instead of the context's
 * own; NULL selects the device's default (null) stream.
 * iac_ctx_use_own_stream() returns to the context's own stream. */
int iac_ctx_set_stream(iac_ctx *ctx, void *hip_stream);
int iac_ctx_use_own_stream(iac_ctx *ctx);
void *iac_ctx_get_stream(iac_ctx *ctx);
int iac_ctx_sync(iac_ctx *ctx);
/* Replace the CL_MEM_COPY_HOST_PTR buffers and blocking enqueueReadBuffer of
 * every reference launch site (e.g. IAC/Detector.cpp:1100-1140). */
int iac_dev_alloc(iac_ctx *ctx, size_t bytes, void **d_ptr);
int iac_dev_free(iac_ctx *ctx, void *d_ptr);
int iac_memcpy_h2d(iac_ctx *ctx, void *d_dst, const void *h_src, size_t bytes);
int iac_memcpy_d2h(iac_ctx *ctx, void *h_dst, const void *d_src, size_t bytes);
int iac_memset_zero(iac_ctx *ctx, void *d_ptr, size_t bytes);
/* Page-locked host memory for frame staging (the reference stages through pageable
 * new[] arrays, e.g. IAC/Detector.cpp:1093-1110): copies from it run at full link speed. */
int iac_pinned_alloc(size_t bytes, void **h_ptr);
int iac_pinned_free(void *h_ptr);
/* Kernel timing for measurement: while enabled, the dominant kernel of every
 * call is bracketed by HIP events on the context's stream.  iac_ctx_timing()
 * synchronizes, returns the summed milliseconds and the number of timed
 * launches since the last reset, and resets when `reset` is nonzero. */
int iac_ctx_enable_timing(iac_ctx *ctx, int enable);
int iac_ctx_timing(iac_ctx *ctx, double *ms, uint64_t *launches, int reset);
/* Context options (tuning and kernel selection for measurement and tests). */
#define IAC_OPT_AC3D_KERNEL 1 /* 0 automatic, 1 global-atomic tile kernel, 2 LDS cell-box kernel */
#define IAC_OPT_CELL_VOX 2    /* LDS kernel: cell side (along the widest k axis) in voxels, 0 = default (6 for
                                 AC_uw nearest neighbour, 5 for AC_uw linear, 3 for C(q) nearest neighbour, 4
                                 for C(q) linear) */
#define IAC_OPT_FRAME_CHUNK 3 /* LDS kernel: mean frames per block, at most 2^28 pairs of work, 0 = default (32) */
#define IAC_OPT_PAIR_PREFETCH 5 /* LDS kernel: record pairs loaded per batch, 1, 4 or 8 (0 = 1) */
#define IAC_OPT_BOX_CAP 6 /* LDS kernel: box capacity in voxels, 1024, 1920 (2048 selects it too) or 4096
                               (0 = 1920 for nearest neighbour, 4096 for linear) */
#define IAC_OPT_CELL_VOX_V 7 /* LDS kernel: cell side along the second k axis, 0 = as IAC_OPT_CELL_VOX */
#define IAC_OPT_TASK_REGION 8 /* LDS kernel: first cells per task region side, 0 = default (2 for C(q) nearest
                                   neighbour, 4 for C(q) linear, else the whole grid) */
#define IAC_OPT_PAIR_MODE 9 /* LDS kernel: 0 default (1 for AC_uw, 2 for C(q)), 1 lanes on consecutive pairs,
                                2 a contiguous pair range per thread */
#define IAC_OPT_SIM_CHUNK_ROWS 4 /* simulator with subpixels or modes: detector rows (pixel x subpixel) per GEMM
                                    launch, rounded to whole groups of 256 pixels, 0 = auto (1 GiB of FP32
                                    magnitudes); one subpixel and one mode need no chunks */
#define IAC_OPT_LAP_GROUPS 10 /* LAP finder: two-frame groups a tile's waves take in a row of the block order
                                 (the XCD's L2 keeps the tile's mask, offset and gain rows for them),
                                 0 = default (16) */
int iac_ctx_set_option(iac_ctx *ctx, int option, int64_t value);
/* Updates whose voxel fell outside the mesh since the context was created.
 * The reference writes such updates out of bounds; this path drops them. */
int iac_dropped_updates(iac_ctx *ctx, uint64_t *count);

/* ---- host geometry (the reference computes these on the host as well) --------- */
/* Simulator::GeneratePixelMap, IAC/Simulator.cpp:495-519; pixmap is [size_a][size_b][3]. */
int iac_generate_pixel_map(int size_a, int size_b, float pixel_size, const float center[3],
                           const float vec_a[3], const float vec_b[3], float *pixmap);
/* Detector::Calc_kMap, IAC/Detector.cpp:310-361 */
int iac_calc_kmap(const float *pixmap, uint32_t npix, float *kmap, float max_k[3],
                  float min_k[3], float max_q[3]);
/* ACMesh::CreateBigMeshForDetector, IAC/ACMesh.cpp:74-106 (and the identical
 * CreateBigMesh_CofQ_ForDetector, :107-139) */
int iac_mesh_shape_big(const float max_q[3], int edge_size, float q_zoom, iac_mesh_shape *out);
/* ACMesh::CreateSmallMesh_CofQ_ForDetector, IAC/ACMesh.cpp:140-206 */
int iac_mesh_shape_small_cofq(const float max_q[3], int perp_size, float q_zoom,
                              iac_mesh_shape *out);

/* ---- 3D pair accumulation (AC_uw) -------------------------------------------- */
/* Detector::AutoCorrelateSparseList, IAC/Detector.cpp:1010-1181, with the CPU
 * branch's arithmetic (ACMesh::Atomic_Add_q_Entry, IAC/ACMesh.cpp:209-385):
 * every ordered pair (a,b) of the hit list, the self pair once, is binned and
 * `(unsigned)FloatToInt(v_a*v_b)` is added to the uint64 mesh (accumulating).
 * d_hitlist4 is the reference's SparseHitList layout {kx,ky,kz,value} x n
 * (IAC/Detector.h:120).  rot is the event's row-major RotMatrix
 * (IAC/Settings.h:107), host memory.  d_mesh is size_ab^2*size_c uint64. */
int iac_ac3d_sparse(iac_ctx *ctx, const float *d_hitlist4, uint32_t n, const float rot[9],
                    const iac_mesh_shape *shape, int interp1, int interp2, int double_map,
                    uint64_t *d_mesh);
/* The RunIAC frame loop (IAC/RunIAC.cpp:524-550) in one call: frame f holds
 * the photon entries [d_frame_off[f], d_frame_off[f+1]) of (pixel index,
 * count) lists d_pix/d_cnt; k vectors are gathered from d_kmap ([npix][3]).
 * d_rots is [nframes][9] floats on the device or NULL for the identity.
 * n_total = d_frame_off[nframes] (host copy, sizes the scratch).  The mesh is
 * identical to calling iac_ac3d_sparse once per frame.  Runs on the LDS
 * cell-box engine; per-frame rotations too when the double map has a nearest-
 * neighbour second level and no rotated voxel can leave the mesh (the frames bin
 * under the identity and each box voxel is rotated as it flushes), otherwise
 * on the global-atomic tile kernel. */
int iac_ac3d_sparse_batch(iac_ctx *ctx, const float *d_kmap, uint32_t npix,
                          const uint32_t *d_pix, const uint16_t *d_cnt,
                          const uint64_t *d_frame_off, uint32_t nframes, uint64_t n_total,
                          const float *d_rots, const iac_mesh_shape *shape, int interp1,
                          int interp2, int double_map, uint64_t *d_mesh);

/* The same frame loop over ready-made hit-list records {kx,ky,kz,value}
 * (d_rec4, e.g. from iac_sparse_hit_list), frame f owning records
 * [d_frame_off[f], d_frame_off[f+1]). */
int iac_ac3d_records_batch(iac_ctx *ctx, const float *d_kmap, uint32_t npix, const float *d_rec4,
                           const uint64_t *d_frame_off, uint32_t nframes, uint64_t n_total,
                           const float *d_rots, const iac_mesh_shape *shape, int interp1, int interp2,
                           int double_map, uint64_t *d_mesh);

/* A prepared frame batch for repeated passes (identity rotation, the LDS engine):
 * the cell plan, the pair-work-balanced frame chunks and all buffers are made once;
 * iac_ac3d_batch_launch then enqueues gather + cell sort + pair pass (+ mirror) with
 * no host synchronisation and no allocation.  iac_ac3d_batch_capture records that
 * sequence into a HIP graph for (d_mesh, zero_mesh); later launches with the same pair
 * replay the graph.  d_kmap / d_pix / d_cnt / d_frame_off are kept by pointer and must
 * outlive the batch; the photon lists may change between launches with the frame count
 * and n_total unchanged.  zero_mesh
 * clears d_mesh first (by kernel).  No counterpart in the reference: it is the
 * frame loop of IAC/RunIAC.cpp:524-550 as one replayable unit. */
typedef struct iac_ac3d_batch iac_ac3d_batch;
int iac_ac3d_batch_create(iac_ctx *ctx, const float *d_kmap, uint32_t npix, const uint32_t *d_pix,
                          const uint16_t *d_cnt, const uint64_t *d_frame_off, uint32_t nframes, uint64_t n_total,
                          const iac_mesh_shape *shape, int interp1, int interp2, int double_map,
                          iac_ac3d_batch **out);
int iac_ac3d_batch_launch(iac_ctx *ctx, iac_ac3d_batch *batch, uint64_t *d_mesh, int zero_mesh);
int iac_ac3d_batch_capture(iac_ctx *ctx, iac_ac3d_batch *batch, uint64_t *d_mesh, int zero_mesh);
int iac_ac3d_batch_destroy(iac_ac3d_batch *batch);

/* The engine's autotuner: sweeps the LDS cell-box engine's settings (pair walk, cell side,
 * box capacity, frames per block) on the first prefix_frames frames of a batch (0: 1,000)
 * with the given binning modes, timing the pair kernel of each with HIP events (best of two
 * runs after a planning run, into a scratch mesh the call allocates and frees), and makes
 * the fastest the context's options (IAC_OPT_PAIR_MODE, IAC_OPT_CELL_VOX(_V), IAC_OPT_BOX_CAP,
 * IAC_OPT_FRAME_CHUNK, IAC_OPT_TASK_REGION), which every later engine call of the context
 * uses.  The winner is cached in the context per (k-map, mesh shape, modes): a second call
 * with the same key applies it without timing (force_retune = 0).  *best gets the winner;
 * up to max_all candidate results go to all (may be NULL), their count to *n_all.  Pending
 * kernel timings of the context (iac_ctx_timing) are discarded.  No counterpart in the
 * reference (it has no tunable engine). */
typedef struct {
    int32_t pair_mode, cell_vox, cell_vox_v, box_cap, task_region, frame_chunk;
    float ms; /* pair-kernel milliseconds on the prefix (-1: taken from the cache) */
} iac_tune_result;
int iac_ac3d_tune(iac_ctx *ctx, const float *d_kmap, uint32_t npix, const uint32_t *d_pix, const uint16_t *d_cnt,
                  const uint64_t *d_frame_off, uint32_t nframes, const iac_mesh_shape *shape, int interp1,
                  int interp2, int double_map, uint32_t prefix_frames, int force_retune, iac_tune_result *best,
                  iac_tune_result *all, uint32_t max_all, uint32_t *n_all);

/* A frame queue: the per-event call of the reference's frame loop,
 * Detector::AutoCorrelateSparseList once per event (IAC/RunIAC.cpp:524-550,
 * IAC/Detector.cpp:1010-1181), at the speed of the batched pass.  Each push appends one
 * frame's hit list {kx,ky,kz,value} x n and its row-major rotation to the queue (host
 * records are copied into pinned staging; device records by one device-to-device copy on
 * the context's stream); nothing is binned until the queue fills (max_records records or
 * max_frames frames; 0 selects 2^22 and 4096) or is flushed.  A flush uploads the pending
 * frames and enqueues one pass of the LDS cell-box engine (per-frame rotations included
 * when the second level is nearest-neighbour and every rotated voxel stays in the mesh,
 * else the tile kernel), with no allocation.  Two slots (pinned staging and device records,
 * 2 x 16 max_records bytes each side) alternate: a flush's uploads run on the queue's own
 * stream and overlap the previous flush's pass, and the host waits only when it returns to
 * a slot whose uploads are still running.  The mesh equals calling iac_ac3d_sparse once
 * per frame.  d_kmap ([npix][3]) sets the engine's cell grid (records may hold any k);
 * d_kmap and d_mesh must outlive the queue, and the queue must be destroyed before its
 * context.  After iac_ac3d_queue_flush the mesh is complete in stream order;
 * iac_ac3d_queue_destroy flushes and synchronises. */
typedef struct iac_ac3d_queue iac_ac3d_queue;
int iac_ac3d_queue_create(iac_ctx *ctx, const float *d_kmap, uint32_t npix, const iac_mesh_shape *shape,
                          int interp1, int interp2, int double_map, uint64_t *d_mesh, uint64_t max_records,
                          uint32_t max_frames, iac_ac3d_queue **out);
int iac_ac3d_queue_push(iac_ac3d_queue *queue, const float *h_hitlist4, uint32_t n, const float rot[9]);
int iac_ac3d_queue_push_device(iac_ac3d_queue *queue, const float *d_hitlist4, uint32_t n, const float rot[9]);
int iac_ac3d_queue_flush(iac_ac3d_queue *queue);
int iac_ac3d_queue_pending(const iac_ac3d_queue *queue, uint32_t *frames, uint64_t *records);
int iac_ac3d_queue_destroy(iac_ac3d_queue *queue);

/* ---- sparsification ------------------------------------------------------------- */
/* Detector::CreateSparseHitList(Threshold, PhotonSamplingStep) after the pixel-mask
 * multiply, IAC/Detector.cpp:926-958, IAC/RunIAC.cpp:530-536, for nframes dense frames
 * ([nframes][npix] floats): every pixel with I*mask >= threshold and > 0 becomes a
 * record {k, ceilf((I*mask - threshold)/step)} (step <= 0 keeps I*mask), in pixel
 * order.  d_mask is [npix] ints or NULL (all ones); d_rec4 needs room for
 * nframes*npix records; d_frame_off gets nframes+1 offsets; *n_total the count. */
int iac_sparse_hit_list(iac_ctx *ctx, const float *d_kmap, uint32_t npix, const float *d_frames,
                        uint32_t nframes, const int *d_mask, float threshold, float step,
                        float *d_rec4, uint64_t *d_frame_off, uint64_t *n_total);

/* ---- averaged intensity -------------------------------------------------------- */
/* Detector::LoadAndAverageIntensity(Events, Threshold, PhotonSamplingStep, Lower, Upper,
 * true), IAC/Detector.cpp:687-790, over nframes dense frames d_frames ([nframes][npix]
 * floats, already the events [Lower, Upper)): mask multiply (d_mask [npix] ints or NULL);
 * step > 0: photon discretisation, per-pixel long tally; else values >= threshold summed
 * in float, frames in order; then x (1.0f / nframes) into d_avg [npix].  Per event
 * (optional, device arrays of nframes): MeanIntensity = serial float sum / npix and,
 * when step > 0, PhotonCount = (int) sum. */
int iac_average_intensity(iac_ctx *ctx, const float *d_frames, uint32_t nframes, uint32_t npix, const int *d_mask,
                          float threshold, float step, float *d_avg, float *d_mean_intensity, int *d_photon_count);
/* The same over consecutive frame stacks: each call continues the per-pixel sequence in
 * d_sum (float, step <= 0) or d_tally (int64, step > 0), both [npix], zeroed by the caller
 * before the first stack; iac_intensity_finish then writes d_avg for ntotal_frames frames. */
int iac_intensity_accumulate(iac_ctx *ctx, const float *d_frames, uint32_t nframes, uint32_t npix, const int *d_mask,
                             float threshold, float step, float *d_sum, int64_t *d_tally, float *d_mean_intensity,
                             int *d_photon_count);
int iac_intensity_finish(iac_ctx *ctx, uint32_t npix, uint32_t ntotal_frames, float step, const float *d_sum,
                         const int64_t *d_tally, float *d_avg);

/* ---- pattern pre-processing: largest-adjacent-pixel photon finder --------------- */
/* PPP::DetectorPanel, IAC/PPP.h:14-18 (Scans = {fs, ss}): pixel first_ind + row*fs + col. */
typedef struct {
    uint32_t first_ind;
    uint32_t fs;
    uint32_t ss;
} iac_panel;
/* The per-frame chain of PPP::ProcessData_PF_LAP, IAC/PPP.cpp:469-495, over nframes dense
 * frames ([nframes][npix] floats; d_out may equal d_in, then the call stages the input in
 * its scratch memory; other overlaps are not allowed):
 *  - Detector::ApplyPixelMask (IAC/Detector.cpp:662-665) when d_mask ([npix] ints) is given;
 *  - PPP::GainCorrection (IAC/PPP.cpp:781-803) when d_offset and d_gain ([npix] floats, the
 *    gain file's maps) are given, with AllowNegativeValues = gain_only as the reference
 *    passes it;
 *  - unless gain_only, PPP::PhotonFinder_LargestAdjacentPixel (IAC/PPP.cpp:18-95) over the
 *    npanels panels (host array, at most 128, disjoint, inside the first sum(fs*ss) pixels
 *    of the frame, as the reference's arrays are sized); pixels outside every panel keep
 *    their corrected value;
 *  - per event (optional, device arrays of nframes): PhotonCount = (int) serial float sum
 *    of the frame and MeanIntensity = PhotonCount / npix (IAC/PPP.cpp:490-491).
 * Replaces the in-place PPP::PhotonFinder_LargestAdjacentPixel(float *Intensity, ...)
 * (IAC/PPP.h:90) and the frame loop of PPP::ProcessData_PF_LAP (IAC/PPP.h:95); file reads
 * and writes stay with the caller. */
int iac_ppp_lap(iac_ctx *ctx, const float *d_in, float *d_out, uint32_t nframes, uint32_t npix, const int *d_mask,
                const float *d_offset, const float *d_gain, const iac_panel *panels, uint32_t npanels,
                float adu_per_photon, float seed_threshold, float combined_threshold, int gain_only,
                int *d_photon_count, float *d_mean_intensity);

/* ---- AC1D: angular-averaged correlation ---------------------------------------- */
/* AC1D::Initialize(Det, ArraySize, QZoom), IAC/AC1D.cpp:53-58 (host). */
int iac_ac1d_shape(const float max_q[3], uint32_t size, float q_zoom, float *max_q_out, float *dq_out);
/* AC1D::Calculate_CQ multiplicator, IAC/AC1D.cpp:78-90 (host). */
int iac_ac1d_cq_multiplicator(const float *intensity, uint32_t npix, double *mult);
/* AC1D::Calculate_CQ / AutoCorr_CQ_AV, IAC/AC1D.cpp:61-195, IAC/kernel.cl:656-764: every
 * ordered pixel pair adds (long)(Val*mult) (nearest) or its linear split to the uint64
 * radial histogram d_hist (size bins, accumulating); the reference's double C(q)
 * is d_hist / mult.  size <= 6144. */
int iac_ac1d_cq(iac_ctx *ctx, const float *d_kmap, uint32_t npix, const float *d_intensity,
                uint32_t size, float max_q, float dq, int interp, double mult, uint64_t *d_hist);
/* AC1D::Calculate_AC_UW_MR, CPU branch of Calculate_AC_UW_Mapped, IAC/AC1D.cpp:256-293:
 * photon pairs j <= k of every frame add 2*(c_j*c_k) (or its linear split) to the
 * double histogram d_hist (accumulating).  self_pair_factor 2 weights self pairs
 * like the reference; 1 counts them once. */
int iac_ac1d_acuw(iac_ctx *ctx, const float *d_kmap, uint32_t npix, const uint32_t *d_pix,
                  const uint16_t *d_cnt, const uint64_t *d_frame_off, uint32_t nframes,
                  uint64_t n_total, uint32_t size, float max_q, float dq, int interp,
                  int self_pair_factor, double *d_hist);
/* The same over ready-made hit-list records {kx,ky,kz,value} (e.g. from
 * iac_sparse_hit_list with the evaluation's Photonisation, IAC/AC1D.cpp:246-249);
 * frame f owns records [d_frame_off[f], d_frame_off[f+1]). */
int iac_ac1d_acuw_records(iac_ctx *ctx, const float *d_rec4, const uint64_t *d_frame_off, uint32_t nframes,
                          uint64_t n_total, uint32_t size, float max_q, float dq, int interp, int self_pair_factor,
                          double *d_hist);
/* iac_ac1d_acuw_records without any allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and result as
 * iac_ac1d_acuw_records, at most 65,536 frames per call (IAC_ERR_LIMIT beyond).  Its work
 * memory is a static device array of 560 KiB per GPU, held until the process ends and
 * shared by every call on that GPU: calls that overlap on one GPU (several contexts or
 * streams at once) race; calls in one stream do not.  Not timed by iac_ctx_enable_timing. */
int iac_ac1d_acuw_records_noalloc(iac_ctx *ctx, const float *d_rec4, const uint64_t *d_frame_off, uint32_t nframes,
                                  uint64_t n_total, uint32_t size, float max_q, float dq, int interp,
                                  int self_pair_factor, double *d_hist);
/* AC1D::CalcAC, IAC/AC1D.cpp:482-493: ac = ac_uw / cq, NaN -> 0. */
int iac_ac1d_calc_ac(iac_ctx *ctx, const double *d_ac_uw, const double *d_cq, uint32_t size,
                     double *d_ac);

/* ---- C(q): detector-pair normalisation mesh ------------------------------------ */
/* Multiplicator of Detector::AutoCorrelate_CofQ_SmallMesh, IAC/Detector.cpp:1215-1227:
 * the smallest 10^k >= 1e-10 with <I>^2 * 10^k >= 1, times 1e5 (host). */
int iac_cq_small_multiplicator(const float *intensity, uint32_t npix, double *mult);
/* Detector::AutoCorrelate_CofQ_SmallMesh / AutoCorr_CQ_small, IAC/Detector.cpp:1184-1315,
 * IAC/kernel.cl:27-184: every ordered pixel pair of d_intensity (npix floats)
 * adds (uint64)((double)(I_a*I_b) * mult) (pairs below 1e-37 skipped) to the
 * uint64 small mesh d_mesh (size_ab^2*size_c, axes k_a/k_b/k_c), accumulating. */
int iac_cq_small(iac_ctx *ctx, const float *d_kmap, uint32_t npix, const float *d_intensity,
                 const iac_mesh_shape *small_shape, int interp, double mult, uint64_t *d_mesh);
/* The same binning for nframes dense frames d_frames ([nframes][npix]); frames
 * [h_group_off[g], h_group_off[g+1]) accumulate into mesh g of d_meshes
 * (ngroups consecutive meshes) -- e.g. one group per crystal orientation. */
int iac_cq_dense_frames(iac_ctx *ctx, const float *d_kmap, uint32_t npix, const float *d_frames,
                        uint32_t nframes, const uint32_t *h_group_off, uint32_t ngroups,
                        const iac_mesh_shape *shape, int interp, double mult, uint64_t *d_meshes);
/* d_out[i] = (double)d_in[i] / mult (IAC/Detector.cpp:1304-1308, 1525-1534). */
int iac_u64_to_f64(iac_ctx *ctx, const uint64_t *d_in, size_t n, double mult, double *d_out);
/* Detector::Merge_smallCofQ host preparation, IAC/Detector.cpp:1339-1427: rw10 gets
 * the 9 rotation entries and the weight <I>^2 (quantised to the Multiplicator)
 * per event; *mult_out gets the Multiplicator (host). */
int iac_merge_prepare(const float *rot9, const float *mean_intensity, uint32_t nev,
                      const double *small_cq, size_t nsmall, float *rw10, double *mult_out);
/* Merge_CQ, IAC/kernel.cl:189-344: every small-mesh voxel (double C(q), device)
 * rotated by each event's matrix and weighted, into the cubic uint64 big mesh
 * d_big (size_ab^3), accumulating. */
int iac_merge_cq(iac_ctx *ctx, const double *d_small_cq, const iac_mesh_shape *small_shape,
                 const float *d_rw10, uint32_t nev, int interp, double mult, uint64_t *d_big);
/* Merge_CQ with one small mesh per event: d_small_meshes holds nev consecutive
 * double meshes (e.g. one dense C(q) per crystal orientation); event i rotates
 * mesh i.  Otherwise as iac_merge_cq. */
int iac_merge_cq_multi(iac_ctx *ctx, const double *d_small_meshes, const iac_mesh_shape *small_shape,
                       const float *d_rw10, uint32_t nev, int interp, double mult, uint64_t *d_big);
/* RunIAC::Merge_ACandCQ, IAC/RunIAC.cpp:120-158: out = IntToFloat(ac) / cq, 0 where
 * cq <= 0 or the quotient is NaN. */
int iac_merge_ac_cq(iac_ctx *ctx, const uint64_t *d_ac, const double *d_cq, size_t n, double *d_out);

/* ---------------- simulator: IAC/Simulator.cpp:662-944, IAC/kernel.cl:768-859 ---------------- */
/* psi(p, f) = sum_e exp(i (k_p . r_e + phi_{e,f})), k_p = 2 pi / lambda * p / |p|, over the
 * (2 SuSa + 1)^2 subpixels p + v v_step + u u_step of each pixel.  Phases are
 * [nframes * modes][nem] (mode-minor: row f * modes + m); the output [nframes][npix]
 * sums, over the modes of a frame, the subpixel mean of |psi| (IAC_SIM_AMPLITUDE, what
 * SimulateCrystal returns, kernel.cl:844) or |psi|^2 (IAC_SIM_INTENSITY). */
#define IAC_SIM_FP32_DIRECT 0 /* the reference kernel's FP32 arithmetic, iac_sim_direct_fp32 */
#define IAC_SIM_BF16X3 1      /* complex GEMM, bf16 hi + lo split, 3 MFMA products, FP32 accumulate */
#define IAC_SIM_FP16X3 2      /* the same with fp16 hi + lo */
#define IAC_SIM_BF16 3        /* one bf16 product (about 2^-9 per element; below the 1e-3 bar) */
#define IAC_SIM_FP16 4        /* one fp16 product */
#define IAC_SIM_AMPLITUDE 0
#define IAC_SIM_INTENSITY 1
typedef struct iac_sim_plan iac_sim_plan;
/* The emitter-geometry half of the GEMM: X[p s][e] = exp(i k_{p s} . r_e) with the phase
 * argument in FP64, stored as 16-bit hi (+ lo) of [cos | sin] (npix (2SuSa+1)^2 x 2 nem x
 * 2 or 4 bytes of device memory, kept until iac_sim_plan_destroy).  d_pixmap: float
 * [npix][3] (Detector::PixelMap), d_emitters: double [nem][3]; u_step / v_step as
 * SimulatePart computes them (Simulator.cpp:742-760); mode IAC_SIM_BF16X3..IAC_SIM_FP16.
 * Replaces the per-frame upload and the k / argument work of SimulateCrystal. */
int iac_sim_plan_create(iac_ctx *ctx, const float *d_pixmap, uint32_t npix, int subsampling, const double u_step[3],
                        const double v_step[3], double wavelength, const double *d_emitters, uint32_t nem, int mode,
                        iac_sim_plan **out);
int iac_sim_plan_destroy(iac_sim_plan *plan);
/* The SimulatePart frame loop (Simulator.cpp:694-862) for nframes frames at once:
 * d_phases double [nframes * modes][nem] -> d_out float [nframes][npix]. */
int iac_sim_run(iac_ctx *ctx, iac_sim_plan *plan, const double *d_phases, uint32_t nframes, uint32_t modes,
                int out_kind, float *d_out);
/* SimulateCrystal as the reference computes it, in FP32 (float positions and phases,
 * k recomputed per emitter), for nframes frames: d_emitters float [nem][3], d_phases
 * float [nframes * modes][nem] -> d_out float [nframes][npix].  The plain baseline. */
int iac_sim_direct_fp32(iac_ctx *ctx, const float *d_pixmap, uint32_t npix, int subsampling, const float u_step[3],
                        const float v_step[3], float wavelength, const float *d_emitters, uint32_t nem,
                        const float *d_phases, uint32_t nframes, uint32_t modes, int out_kind, float *d_out);
/* SimulatePart's post-processing (Simulator.cpp:868-888) per frame, in place: scale to
 * n_expected photons (double sum, float factor), Poisson draw (ScalarPoissonSampling,
 * ArrayOperators.h:25-55) if do_poisson, white noise (2u - 1) adu_noise if adu_noise != 0
 * (ArrayOperators.h:219-226), times value_per_photon.  The random numbers come from
 * Philox4x32-10 keyed by seed with counter (pixel, frame0 + f, tag, stream): reproducible
 * and independent of launch order (the reference's mt19937_64 seeded from
 * random_device is neither). */
int iac_sim_postprocess(iac_ctx *ctx, float *d_frames, uint32_t npix, uint32_t nframes, double n_expected,
                        int do_poisson, float adu_noise, float value_per_photon, uint64_t seed, uint32_t frame0);

/* ---- measurement probes (no counterpart in the reference) ------------------------ */
/* n_updates uint64 atomic adds of 1 into d_mesh (nvox >= 4096): mode 0 uniform over the
 * mesh, mode 1 over a window of `hot` voxels in its middle, mode 2 into a 4096-voxel LDS
 * box per block flushed once, mode 3 in runs of 8 consecutive voxels (one 64-byte segment
 * per 8 lanes) at uniform random segments, mode 4 as mode 2 with double adds of 1.0 in the
 * box (flushed as integers).  *ms = kernel time (HIP events on the context's stream). */
int iac_probe_atomics(iac_ctx *ctx, uint64_t *d_mesh, uint64_t nvox, uint64_t n_updates, int mode, uint64_t hot,
                      float *ms);
/* A plain 16-byte-vector copy of `bytes` (a multiple of 16); *ms = kernel time. */
int iac_probe_copy(iac_ctx *ctx, const void *d_src, void *d_dst, size_t bytes, float *ms);
/* Self-check of the short correctly rounded square root the AC1D kernels use for |q|:
 * every float bit pattern u in [lo, hi) (hi <= 2^32) against the library sqrtf; *mismatches
 * = the count of results that differ bit for bit.  Synchronous. */
int iac_selfcheck_sqrt(iac_ctx *ctx, uint64_t lo, uint64_t hi, uint64_t *mismatches);

#ifdef __cplusplus
}
#endif
#endif
