// iac_ctx.hip -- errors, device contexts, device memory and the host-side
// geometry of the C ABI (include/iac_hip.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "iac_internal.h"

#pragma clang fp contract(off)

namespace iac {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg)
{
    set_error(msg);
    return code;
}

int hip_fail(hipError_t e, const char *what, const char *file, int line)
{
    std::string m = std::string(what) + " failed: " + hipGetErrorString(e) + " (" + file + ":" +
                    std::to_string(line) + ")";
    set_error(m);
    return e == hipErrorOutOfMemory ? IAC_ERR_OOM : IAC_ERR_HIP;
}

int bind(iac_ctx *ctx)
{
    if (!ctx) return fail(IAC_ERR_INVALID_ARG, "null context");
    IAC_HIP(hipSetDevice(ctx->device));
    return IAC_OK;
}

int scratch(iac_ctx *ctx, int slot, size_t bytes, void **out)
{
    if (slot < 0 || slot >= 4) return fail(IAC_ERR_INVALID_ARG, "bad scratch slot");
    if (bytes == 0) bytes = 256;
    if (ctx->scratch_bytes[slot] < bytes) {
        if (ctx->scratch[slot]) {
            IAC_HIP(hipStreamSynchronize(ctx->stream));
            IAC_HIP(hipFree(ctx->scratch[slot]));
            ctx->scratch[slot] = nullptr;
            ctx->scratch_bytes[slot] = 0;
        }
        size_t want = bytes + bytes / 4;  // grow with headroom
        IAC_HIP(hipMalloc(&ctx->scratch[slot], want));
        ctx->scratch_bytes[slot] = want;
    }
    *out = ctx->scratch[slot];
    return IAC_OK;
}

int timer_begin(iac_ctx *ctx)
{
    if (!ctx->timing) return IAC_OK;
    if (ctx->ev_used == ctx->ev_start.size()) {
        hipEvent_t a, b;
        IAC_HIP(hipEventCreate(&a));
        IAC_HIP(hipEventCreate(&b));
        ctx->ev_start.push_back(a);
        ctx->ev_stop.push_back(b);
    }
    IAC_HIP(hipEventRecord(ctx->ev_start[ctx->ev_used], ctx->stream));
    return IAC_OK;
}

int timer_end(iac_ctx *ctx)
{
    if (!ctx->timing) return IAC_OK;
    IAC_HIP(hipEventRecord(ctx->ev_stop[ctx->ev_used], ctx->stream));
    ctx->ev_used++;
    return IAC_OK;
}

float cut_sq(float max_q)
{
    if (!(max_q >= 0.0f) || std::isinf(max_q)) return max_q >= 0.0f ? INFINITY : 0.0f;
    float t = max_q * max_q;
    // walk down while the predecessor still passes, then up until t passes
    while (t > 0.0f && std::sqrt(std::nextafter(t, 0.0f)) > max_q) t = std::nextafter(t, 0.0f);
    while (!(std::sqrt(t) > max_q)) t = std::nextafter(t, INFINITY);
    return t;
}

}  // namespace iac

using iac::fail;

extern "C" {

const char *iac_strerror(int code)
{
    switch (code) {
    case IAC_OK: return "ok";
    case IAC_ERR_INVALID_ARG: return "invalid argument";
    case IAC_ERR_EVEN_MESH: return "mesh size must be odd for each dimension";
    case IAC_ERR_NO_DEVICE: return "no usable HIP device";
    case IAC_ERR_HIP: return "HIP runtime error";
    case IAC_ERR_OOM: return "device out of memory";
    case IAC_ERR_SHAPE: return "shape mismatch";
    case IAC_ERR_PRECONDITION: return "precondition not met";
    case IAC_ERR_LIMIT: return "size exceeds a documented limit";
    default: return "unknown error";
    }
}

const char *iac_last_error(void) { return iac::g_last_error.c_str(); }

int iac_abi_version(void) { return IAC_ABI_VERSION; }

int iac_device_count(int *count)
{
    if (!count) return fail(IAC_ERR_INVALID_ARG, "null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return IAC_OK;
}

int iac_ctx_create(int device, iac_ctx **out)
{
    if (!out) return fail(IAC_ERR_INVALID_ARG, "null output");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return fail(IAC_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= n) return fail(IAC_ERR_NO_DEVICE, "device index out of range");
    iac_ctx *c = new iac_ctx();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&c->d_counters, 8 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(c->d_counters, 0, 8 * sizeof(unsigned long long));
    hipDeviceProp_t prop;
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) {
        int r = iac::hip_fail(e, "iac_ctx_create", __FILE__, __LINE__);
        if (c->d_counters) (void)hipFree(c->d_counters);
        if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
        delete c;
        return r;
    }
    c->num_cus = prop.multiProcessorCount;
    c->stream = c->own_stream;
    *out = c;
    return IAC_OK;
}

int iac_ctx_destroy(iac_ctx *ctx)
{
    if (!ctx) return IAC_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (int i = 0; i < 4; i++)
        if (ctx->scratch[i]) (void)hipFree(ctx->scratch[i]);
    iac::free_pair_plans(ctx);
    iac::free_sim_resources(ctx);
    for (auto e : ctx->ev_start) (void)hipEventDestroy(e);
    for (auto e : ctx->ev_stop) (void)hipEventDestroy(e);
    if (ctx->d_counters) (void)hipFree(ctx->d_counters);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return IAC_OK;
}

int iac_ctx_set_stream(iac_ctx *ctx, void *hip_stream)
{
    IAC_TRY(iac::bind(ctx));
    ctx->stream = (hipStream_t)hip_stream;  // NULL is the device's default (null) stream
    return IAC_OK;
}

int iac_ctx_use_own_stream(iac_ctx *ctx)
{
    IAC_TRY(iac::bind(ctx));
    ctx->stream = ctx->own_stream;
    return IAC_OK;
}

void *iac_ctx_get_stream(iac_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int iac_ctx_sync(iac_ctx *ctx)
{
    IAC_TRY(iac::bind(ctx));
    IAC_HIP(hipStreamSynchronize(ctx->stream));
    return IAC_OK;
}

int iac_dev_alloc(iac_ctx *ctx, size_t bytes, void **d_ptr)
{
    if (!d_ptr || bytes == 0) return fail(IAC_ERR_INVALID_ARG, "bad allocation request");
    IAC_TRY(iac::bind(ctx));
    IAC_HIP(hipMalloc(d_ptr, bytes));
    return IAC_OK;
}

int iac_dev_free(iac_ctx *ctx, void *d_ptr)
{
    IAC_TRY(iac::bind(ctx));
    if (d_ptr) {
        IAC_HIP(hipStreamSynchronize(ctx->stream));
        IAC_HIP(hipFree(d_ptr));
    }
    return IAC_OK;
}

int iac_pinned_alloc(size_t bytes, void **h_ptr)
{
    if (!h_ptr || bytes == 0) return fail(IAC_ERR_INVALID_ARG, "null pointer or zero size");
    *h_ptr = nullptr;
    IAC_HIP(hipHostMalloc(h_ptr, bytes, hipHostMallocDefault));
    return IAC_OK;
}

int iac_pinned_free(void *h_ptr)
{
    if (!h_ptr) return IAC_OK;
    IAC_HIP(hipHostFree(h_ptr));
    return IAC_OK;
}

int iac_memcpy_h2d(iac_ctx *ctx, void *d_dst, const void *h_src, size_t bytes)
{
    IAC_TRY(iac::bind(ctx));
    if (bytes == 0) return IAC_OK;
    if (!d_dst || !h_src) return fail(IAC_ERR_INVALID_ARG, "null pointer");
    IAC_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
    IAC_HIP(hipStreamSynchronize(ctx->stream));
    return IAC_OK;
}

int iac_memcpy_d2h(iac_ctx *ctx, void *h_dst, const void *d_src, size_t bytes)
{
    IAC_TRY(iac::bind(ctx));
    if (bytes == 0) return IAC_OK;
    if (!h_dst || !d_src) return fail(IAC_ERR_INVALID_ARG, "null pointer");
    IAC_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    IAC_HIP(hipStreamSynchronize(ctx->stream));
    return IAC_OK;
}

int iac_memset_zero(iac_ctx *ctx, void *d_ptr, size_t bytes)
{
    IAC_TRY(iac::bind(ctx));
    if (bytes == 0) return IAC_OK;
    if (!d_ptr) return fail(IAC_ERR_INVALID_ARG, "null pointer");
    IAC_HIP(hipMemsetAsync(d_ptr, 0, bytes, ctx->stream));
    return IAC_OK;
}

int iac_ctx_enable_timing(iac_ctx *ctx, int enable)
{
    IAC_TRY(iac::bind(ctx));
    ctx->timing = enable != 0;
    return IAC_OK;
}

int iac_ctx_timing(iac_ctx *ctx, double *ms, uint64_t *launches, int reset)
{
    if (!ms || !launches) return fail(IAC_ERR_INVALID_ARG, "null output");
    IAC_TRY(iac::bind(ctx));
    double total = 0.0;
    for (size_t i = 0; i < ctx->ev_used; i++) {
        IAC_HIP(hipEventSynchronize(ctx->ev_stop[i]));
        float t = 0.0f;
        IAC_HIP(hipEventElapsedTime(&t, ctx->ev_start[i], ctx->ev_stop[i]));
        total += t;
    }
    *ms = total;
    *launches = ctx->ev_used;
    if (reset) ctx->ev_used = 0;
    return IAC_OK;
}

int iac_ctx_set_option(iac_ctx *ctx, int option, int64_t value)
{
    IAC_TRY(iac::bind(ctx));
    switch (option) {
    case IAC_OPT_AC3D_KERNEL:
        if (value < 0 || value > 2) return fail(IAC_ERR_INVALID_ARG, "IAC_OPT_AC3D_KERNEL is 0, 1 or 2");
        ctx->ac3d_kernel = (int)value;
        return IAC_OK;
    case IAC_OPT_CELL_VOX:
        if (value < 0 || value > 64) return fail(IAC_ERR_INVALID_ARG, "IAC_OPT_CELL_VOX is 0..64");
        ctx->cell_vox = (int)value;
        return IAC_OK;
    case IAC_OPT_FRAME_CHUNK:
        if (value < 0 || value > (1 << 20)) return fail(IAC_ERR_INVALID_ARG, "IAC_OPT_FRAME_CHUNK is 0..2^20");
        ctx->frame_chunk = (uint32_t)value;
        return IAC_OK;
    case IAC_OPT_CELL_VOX_V:
        if (value < 0 || value > 64) return fail(IAC_ERR_INVALID_ARG, "IAC_OPT_CELL_VOX_V is 0..64");
        ctx->cell_vox_v = (int)value;
        return IAC_OK;
    case IAC_OPT_TASK_REGION:
        if (value < 0 || value > 4096) return fail(IAC_ERR_INVALID_ARG, "IAC_OPT_TASK_REGION is 0..4096");
        ctx->task_region = (int)value;
        return IAC_OK;
    case IAC_OPT_PAIR_MODE:
        if (value < 0 || value > 2) return fail(IAC_ERR_INVALID_ARG, "IAC_OPT_PAIR_MODE is 0, 1 or 2");
        ctx->pair_mode = (int)value;
        return IAC_OK;
    case IAC_OPT_BOX_CAP:
        if (value != 0 && value != 1024 && value != 1920 && value != 2048 && value != 4096)
            return fail(IAC_ERR_INVALID_ARG, "IAC_OPT_BOX_CAP is 0, 1024, 1920 (or 2048, the same) or 4096");
        ctx->box_cap = (int)value;
        return IAC_OK;
    case IAC_OPT_PAIR_PREFETCH:
        if (value != 0 && value != 1 && value != 4 && value != 8)
            return fail(IAC_ERR_INVALID_ARG, "IAC_OPT_PAIR_PREFETCH is 0, 1, 4 or 8");
        ctx->pair_prefetch = (int)value;
        return IAC_OK;
    case IAC_OPT_LAP_GROUPS:
        if (value < 0 || value > 65536) return fail(IAC_ERR_INVALID_ARG, "IAC_OPT_LAP_GROUPS is 0..65536");
        ctx->lap_groups = (uint32_t)value;
        return IAC_OK;
    case IAC_OPT_SIM_CHUNK_ROWS:
        if (value < 0 || value > (1ll << 31)) return fail(IAC_ERR_INVALID_ARG, "IAC_OPT_SIM_CHUNK_ROWS is 0..2^31");
        ctx->sim_chunk_rows = (uint64_t)value;
        return IAC_OK;
    default:
        return fail(IAC_ERR_INVALID_ARG, "unknown option");
    }
}

int iac_dropped_updates(iac_ctx *ctx, uint64_t *count)
{
    if (!count) return fail(IAC_ERR_INVALID_ARG, "null count");
    IAC_TRY(iac::bind(ctx));
    unsigned long long v = 0;
    IAC_HIP(hipMemcpyAsync(&v, ctx->d_counters, sizeof(v), hipMemcpyDeviceToHost, ctx->stream));
    IAC_HIP(hipStreamSynchronize(ctx->stream));
    *count = v;
    return IAC_OK;
}

// ---- host geometry ---------------------------------------------------------

// Simulator::GeneratePixelMap, IAC/Simulator.cpp:495-519.  Each element is
// Center + (VecA*PS)*(a - 0.5*SizeA) + (VecB*PS)*(b - 0.5*SizeB): float
// products widened by the 0.5 literal, summed in double, stored as float.
int iac_generate_pixel_map(int size_a, int size_b, float pixel_size, const float center[3],
                           const float vec_a[3], const float vec_b[3], float *pixmap)
{
    if (size_a <= 0 || size_b <= 0 || !center || !vec_a || !vec_b || !pixmap)
        return fail(IAC_ERR_INVALID_ARG, "iac_generate_pixel_map: bad arguments");
    float va[3] = {vec_a[0], vec_a[1], vec_a[2]};
    float vb[3] = {vec_b[0], vec_b[1], vec_b[2]};
    float na = std::sqrt(va[0] * va[0] + va[1] * va[1] + va[2] * va[2]);
    float nb = std::sqrt(vb[0] * vb[0] + vb[1] * vb[1] + vb[2] * vb[2]);
    for (int d = 0; d < 3; d++) {
        va[d] = va[d] / na;
        vb[d] = vb[d] / nb;
    }
    for (int a = 0; a < size_a; a++) {
        const double da = (double)a - 0.5 * size_a;
        for (int b = 0; b < size_b; b++) {
            const double db = (double)b - 0.5 * size_b;
            float *px = pixmap + 3 * ((size_t)a * size_b + b);
            for (int d = 0; d < 3; d++) {
                const float ca = va[d] * pixel_size;
                const float cb = vb[d] * pixel_size;
                px[d] = (float)((double)center[d] + (double)ca * da + (double)cb * db);
            }
        }
    }
    return IAC_OK;
}

// Detector::Calc_kMap, IAC/Detector.cpp:310-361
int iac_calc_kmap(const float *pixmap, uint32_t npix, float *kmap, float max_k[3], float min_k[3],
                  float max_q[3])
{
    if (!pixmap || !kmap || !max_k || !min_k || !max_q || npix == 0)
        return fail(IAC_ERR_INVALID_ARG, "iac_calc_kmap: bad arguments");
    for (int d = 0; d < 3; d++) {
        max_k[d] = -100.0f;
        min_k[d] = 100.0f;
    }
    for (uint32_t i = 0; i < npix; i++) {
        float x = pixmap[3 * (size_t)i], y = pixmap[3 * (size_t)i + 1], z = pixmap[3 * (size_t)i + 2];
        float r = std::sqrt(x * x + y * y + z * z);
        x = x / r;
        y = y / r;
        z = z / r;
        kmap[3 * (size_t)i] = x;
        kmap[3 * (size_t)i + 1] = y;
        kmap[3 * (size_t)i + 2] = z;
        if (x > max_k[0]) max_k[0] = x;
        if (y > max_k[1]) max_k[1] = y;
        if (z > max_k[2]) max_k[2] = z;
        if (x < min_k[0]) min_k[0] = x;
        if (y < min_k[1]) min_k[1] = y;
        if (z < min_k[2]) min_k[2] = z;
    }
    for (int d = 0; d < 3; d++) max_q[d] = std::fabs(max_k[d] - min_k[d]);
    return IAC_OK;
}

static void finish_shape(iac_mesh_shape *s, float maxq)
{
    s->center[0] = (s->size_ab - 1) / 2;
    s->center[1] = (s->size_ab - 1) / 2;
    s->center[2] = (s->size_c - 1) / 2;
    s->max_q = maxq;
    const uint32_t den = ((s->size_ab - 1) / 2) - 2;
    const float dq = maxq / (float)den;
    s->voxel_per_dq = (uint32_t)std::floor(1 / dq);
    s->dq_per_voxel = 1.0f / ((float)s->voxel_per_dq);
}

// ACMesh::CreateBigMeshForDetector, IAC/ACMesh.cpp:74-106
int iac_mesh_shape_big(const float max_q[3], int edge_size, float q_zoom, iac_mesh_shape *out)
{
    if (!max_q || !out || edge_size < 5) return fail(IAC_ERR_INVALID_ARG, "bad mesh request");
    if (edge_size % 2 == 0) return fail(IAC_ERR_EVEN_MESH, "Meshsize must be odd for each dimension");
    std::memset(out, 0, sizeof(*out));
    out->k_a = 0;
    out->k_b = 1;
    out->k_c = 2;
    out->size_ab = (uint32_t)(edge_size + 2);
    out->size_c = out->size_ab;
    const float mq = std::max(std::max(max_q[0], max_q[1]), max_q[2]);
    finish_shape(out, mq / q_zoom);
    return IAC_OK;
}

// ACMesh::CreateSmallMesh_CofQ_ForDetector, IAC/ACMesh.cpp:140-206
int iac_mesh_shape_small_cofq(const float max_q[3], int perp_size, float q_zoom,
                              iac_mesh_shape *out)
{
    if (!max_q || !out || perp_size < 5) return fail(IAC_ERR_INVALID_ARG, "bad mesh request");
    if (perp_size % 2 == 0) return fail(IAC_ERR_EVEN_MESH, "PerpSize must be odd");
    std::memset(out, 0, sizeof(*out));
    uint32_t ka = 0, kb = 1, kc = 2;
    if (max_q[kb] > max_q[ka]) std::swap(ka, kb);
    if (max_q[kc] > max_q[kb]) std::swap(kc, kb);
    if (max_q[kb] > max_q[ka]) std::swap(ka, kb);
    out->k_a = ka;
    out->k_b = kb;
    out->k_c = kc;
    out->size_ab = (uint32_t)(perp_size + 2);
    if (q_zoom == 1) {
        const float ratio = (max_q[kc] / max_q[ka]) * (float)perp_size;
        out->size_c = (uint32_t)(int)std::floor((double)ratio + 2.5);
    } else {
        out->size_c = out->size_ab;
    }
    if (out->size_c % 2 == 0) out->size_c += 1;
    finish_shape(out, max_q[ka] / q_zoom);
    return IAC_OK;
}

}  // extern "C"
