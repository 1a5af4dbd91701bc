// iac_internal.h -- context, error plumbing and launch helpers shared by the
// HIP translation units of libiac_hip.so.  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/iac_hip.h"

struct iac_pair_plan;  // cell grid and task table of the LDS pair engine

struct iac_ctx {
    int device = 0;
    int num_cus = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    // device counters: [0] dropped updates, [1..7] reserved
    unsigned long long *d_counters = nullptr;
    // growable scratch arenas (device); each call may reuse them
    void *scratch[4] = {nullptr, nullptr, nullptr, nullptr};
    size_t scratch_bytes[4] = {0, 0, 0, 0};
    // measurement: event pairs around the dominant kernel of each call
    bool timing = false;
    std::vector<hipEvent_t> ev_start, ev_stop;
    size_t ev_used = 0;
    std::vector<iac_pair_plan *> pair_plans;  // small cache, oldest first
    // options (IAC_OPT_*)
    int ac3d_kernel = 0;
    int cell_vox = 0;     // pair engine: cell side along u in voxels (0 = per kind and walk)
    int cell_vox_v = 0;   // pair engine: cell side along v (0 = as along u)
    int task_region = 0;  // pair engine: cells per task region side (0 = per kind)
    int pair_mode = 0;    // pair engine: 0/1 lanes on consecutive pairs, 2 a pair range per thread
    uint32_t frame_chunk = 0;
    uint32_t lap_groups = 0;      // LAP finder: two-frame groups per tile in a row of the block order (0: 16)
    uint64_t sim_chunk_rows = 0;  // simulator: rows (pixel x subpixel) per GEMM launch with subpixels or modes
    int box_cap = 0;        // pair engine: LDS box voxels, 1024, 1920 or 4096 (0 = per kind)
    int pair_prefetch = 0;  // pair engine: pairs per load batch (1, 4 or 8; 0 = 1)
    // autotuner winners per (k-map, mesh shape, binning modes)
    struct TuneEntry {
        const float *kmap;
        uint32_t npix;
        iac_mesh_shape shape;
        int interp1, interp2, double_map;
        iac_tune_result best;
    };
    std::vector<TuneEntry> tune_cache;
};

namespace iac {

// Thread-local detail of the last failure (iac_last_error()).
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what, const char *file, int line);

// Makes ctx->device current on the calling thread.
int bind(iac_ctx *ctx);
// Ensures scratch arena `slot` holds at least `bytes`; returns its pointer in *out.
int scratch(iac_ctx *ctx, int slot, size_t bytes, void **out);

// Brackets the dominant kernel of a call when timing is enabled.
int timer_begin(iac_ctx *ctx);
int timer_end(iac_ctx *ctx);

void free_pair_plans(iac_ctx *ctx);
void free_sim_resources(iac_ctx *ctx);

// Smallest float s >= 0 with sqrtf(s) > max_q (correctly rounded sqrtf).
// `sqrtf(s) > max_q` is then exactly `s >= cut_sq(max_q)` for s >= 0.
float cut_sq(float max_q);

}  // namespace iac

#define IAC_HIP(call)                                                              \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) return iac::hip_fail(e_, #call, __FILE__, __LINE__); \
    } while (0)

#define IAC_TRY(call)            \
    do {                         \
        int r_ = (call);         \
        if (r_ != IAC_OK) return r_; \
    } while (0)
