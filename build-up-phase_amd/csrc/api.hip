// api.hip — the C ABI declared in include/rtx.h.
//
// Host-side validation mirrors the valid-usage rules the reference relies on (main.cpp:726-812, 848-945,
// 1349-1355): vertex format R32G32B32_SFLOAT, uint32 indices, caller-owned storage at least as large as the
// size query, 256-B aligned addresses.  Errors come back as rtx_result codes with a message in
// rtx_last_error_string; nothing throws across the ABI.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "build.h"
#include "noalloc.h"
#include "nrc.h"
#include "pathtrace.h"
#include "rtx_internal.h"
#include "trace.h"
#include "util.h"

struct rtx_context_t {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    rtx::DeviceState* state = nullptr;
    uint2* gstack = nullptr;
    uint64_t gstack_entries = 0;
    int cu_count = 256;
    rtx_trace_config cfg{256, 4, 12, RTX_TRACE_RAY_CHUNK_AUTO, 48, 4, 8, 0};
    uint32_t last_stack = 0;   // stack entries per lane of the last traversal launch (error message)
    uint32_t render_form = RTX_RENDER_FORM_NONE;
    bool stats = false;
    bool used_noalloc = false;  // rtx_sync also reads the device's no-alloc error flags
    void* pt_scratch = nullptr;
    uint64_t pt_scratch_bytes = 0;
    void* ms_buf = nullptr;  // plain renders: per-path state of several samples per launch
    uint64_t ms_bytes = 0;
    std::vector<rtx_nrc_t*> caches;  // live caches; detached when the context is destroyed first
    // batched renders: uploaded views, the graph of the last recorded call and the word it reads sample_offset from
    void* views_buf = nullptr;
    uint64_t views_bytes = 0;
    hipStream_t capture_stream = nullptr;
    hipGraphExec_t graph = nullptr;
    uint64_t graph_key = 0;
    hipGraphExec_t nrc_graph = nullptr;  // the last recorded NRC frame (rtx_cmd_render_nrc with the graph bit)
    uint64_t nrc_graph_key = 0;
    uint32_t* sample_word = nullptr;
    // training wavefront of NRC frames: side stream, fork / join events, its ray counter and spill stack
    hipStream_t side_stream = nullptr;
    hipEvent_t side_fork = nullptr, side_join = nullptr;
    rtx::DeviceState* side_state = nullptr;
    uint2* side_gstack = nullptr;
    uint64_t side_gstack_entries = 0;
    // set once a call has been recorded by a capture of the context stream: the graph may hold the context's
    // buffers, so buffers replaced by larger ones are kept (retired) until the context is destroyed
    bool recorded = false;
    std::vector<void*> retired;
};

struct rtx_nrc_t {
    rtx_context ctx = nullptr;
    rtx::NrcState st;
    void* train_buf = nullptr;  // training samples of rtx_cmd_render_nrc
    uint64_t train_buf_bytes = 0;
};

namespace {

constexpr uint32_t kMaxPrims = (1u << 30) - 1u;
// NRC training paths: the pixel-path launch leaves kSideReserve 256-thread blocks per CU free for them, and their
// launch has kSideBlocks 256-thread blocks per CU: the blocks beyond the reserve start as pixel-path blocks retire
// (the traversal hands out paths from an atomic counter, so late blocks only take what is left).
constexpr uint32_t kSideReserve = 1u;
constexpr uint32_t kSideBlocks = 2u;
// Samples per whole-path launch of a plain render: up to 16, and at most 2^25 paths per launch (the per-path state
// takes 84 B per pixel and sample: 2.8 GB at 1920x1080 x 16).  At 1080p, 64 spp: 4 / 8 / 16 samples per launch
// 303 / 289 / 282 ms against 324 ms for one.
constexpr uint32_t kSampleViews = 16u;
constexpr uint64_t kSampleViewPaths = 1ull << 25;

// splitmix64 (Steele, Lea & Flood 2014): deterministic weight initialisation
uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void nrc_init_params(uint32_t seed, std::vector<float>& p) {
    p.assign(RTX_NRC_PARAM_COUNT, 0.0f);
    uint64_t st = seed;
    const int in_dim[6] = {96, 64, 64, 64, 64, 64}, out_dim[6] = {64, 64, 64, 64, 64, 16};
    const int fan_in[6] = {81, 64, 64, 64, 64, 64};
    size_t off = 0;
    for (int l = 0; l < 6; ++l) {
        const double bound = std::sqrt(6.0 / fan_in[l]);
        for (int r = 0; r < out_dim[l]; ++r)
            for (int c = 0; c < in_dim[l]; ++c) {
                const double u = (double)(splitmix64(st) >> 11) * (1.0 / 9007199254740992.0);
                const bool used = c < fan_in[l] && (l < 5 || r < 3);
                p[off + (size_t)r * in_dim[l] + c] = used ? (float)((2.0 * u - 1.0) * bound) : 0.0f;
            }
        off += (size_t)out_dim[l] * in_dim[l] + out_dim[l];  // biases stay 0
    }
}

int fail(rtx_context ctx, int code, const char* fmt, ...) __attribute__((format(printf, 3, 4)));
int fail(rtx_context ctx, int code, const char* fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->last_error = buf;
    }
    return code;
}

int hip_fail(rtx_context ctx, hipError_t e, const char* where) {
    return fail(ctx, RTX_ERROR_DEVICE_LOST, "%s: %s", where, hipGetErrorString(e));
}

bool aligned(const void* p, uint64_t a) { return ((uintptr_t)p % a) == 0; }

int check_geometries(rtx_context ctx, const rtx_blas_geometry* g, uint32_t count, uint64_t* total) {
    if (count > 0 && !g) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "geometries is NULL");
    uint64_t n = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const rtx_blas_geometry& x = g[i];
        if (x.primitive_count == 0) continue;
        if (x.geometry_type == RTX_GEOMETRY_TYPE_AABBS) {
            if (!x.vertex_data_device)
                return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "geometry %u: AABB data is NULL", i);
            if (x.vertex_stride < 24 || (x.vertex_stride & 7))
                return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "geometry %u: AABB stride must be >= 24 and a multiple of 8",
                            i);
            if (x.primitive_offset & 7)
                return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "geometry %u: AABB primitive_offset must be a multiple of 8",
                            i);
            n += x.primitive_count;
            continue;
        }
        if (x.geometry_type != RTX_GEOMETRY_TYPE_TRIANGLES)
            return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "geometry %u: unknown geometry_type %u", i, x.geometry_type);
        if (!x.vertex_data_device || !x.index_data_device)
            return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "geometry %u: vertex or index data is NULL", i);
        if (x.index_type != RTX_INDEX_TYPE_UINT32)
            return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "geometry %u: only RTX_INDEX_TYPE_UINT32 is supported", i);
        if (x.vertex_stride < 12 || (x.vertex_stride & 3))
            return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "geometry %u: vertex_stride must be >= 12 and a multiple of 4",
                        i);
        if (x.primitive_offset & 3)
            return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "geometry %u: primitive_offset must be a multiple of 4", i);
        if (x.transform_data_device && (x.transform_offset & 15))
            return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "geometry %u: transform_offset must be a multiple of 16", i);
        n += x.primitive_count;
    }
    if (n > kMaxPrims) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "too many primitives (%llu)", (unsigned long long)n);
    *total = n;
    return RTX_SUCCESS;
}

// ---- TLAS properties the host needs before a launch -------------------------------------------------------------
// The traversal variant (and whether renders can use whole-path launches) follows the traced TLAS's content bits,
// and the stack size its depth.  Both live in the TLAS header on the device.  Every TLAS build (and AS copy)
// enqueues a 16-byte read-back of the header words into pinned host memory behind the build and records an event; a
// launch on that address reads them once the event has completed, waiting for it only if the build is still running
// (the launch would wait for it on the stream anyway).  Slots are process-wide (a TLAS built by one context can be
// traced by another) and reused least recently first.
constexpr uint32_t kPropSlots = 256;
constexpr uint32_t kPropWords = 4;  // AsHeader words 8..11: max_depth, max_total_depth, build_error, content
static_assert(offsetof(rtx::AsHeader, max_depth) == 32 && offsetof(rtx::AsHeader, content) == 44, "header words");

struct PropSlot {
    uint64_t address = 0;
    uint32_t* host = nullptr;   // kPropWords pinned words
    hipEvent_t ready = nullptr;
    hipStream_t stream = nullptr;  // the stream the read-back was enqueued on
    bool done = false;             // the read-back is known to have completed (host words valid)
    int device = -1;
    uint64_t used = 0;
};
std::mutex g_prop_mu;
PropSlot g_prop[kPropSlots];
uint32_t* g_prop_host = nullptr;
uint64_t g_prop_clock = 0;

struct TlasProps {
    uint32_t content;  // kContent* bits
    uint32_t depth;    // BVH8 levels, TLAS + deepest BLAS (0: unknown)
    bool known;
};
// what a launch assumes when the properties cannot be read without breaking a graph capture: every program kind and
// the default stack (recorded launches then let the device choose the traversal variant where they can)
constexpr TlasProps kUnknownProps = {rtx::kContentNonOpaque | rtx::kContentAabb, 0, false};

bool stream_capturing(hipStream_t s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}

// Grow a context-owned device buffer to `need` bytes (no-op when it is large enough).  A call being recorded into
// a graph cannot allocate (allocations are not stream work): the buffer must have been sized by an earlier direct
// call of the same shape.  A replaced buffer is freed, or retired when a recorded graph may still use it.
int grow_buffer(rtx_context ctx, void** buf, uint64_t* have, uint64_t need, const char* what) {
    if (need <= *have) return RTX_SUCCESS;
    if (stream_capturing(ctx->stream))
        return fail(ctx, RTX_ERROR_OUT_OF_DEVICE_MEMORY,
                    "%s must grow to %llu bytes while the call is being recorded: make the same call once before "
                    "recording it", what, (unsigned long long)need);
    if (*buf) {
        if (ctx->recorded)
            ctx->retired.push_back(*buf);
        else
            (void)hipFree(*buf);
    }
    *buf = nullptr;
    *have = 0;
    hipError_t e = hipMalloc(buf, need);
    if (e != hipSuccess) {
        *buf = nullptr;
        return fail(ctx, RTX_ERROR_OUT_OF_DEVICE_MEMORY, "%s: %s", what, hipGetErrorString(e));
    }
    *have = need;
    return RTX_SUCCESS;
}

// A host array that a call being recorded must carry inside the graph (rtx::upload_host_array's limit).
int check_recordable(rtx_context ctx, uint64_t bytes, const char* what) {
    if (bytes <= rtx::kArgUploadMax || !stream_capturing(ctx->stream)) return RTX_SUCCESS;
    return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "a recorded call carries at most %llu bytes of %s (%llu given)",
                (unsigned long long)rtx::kArgUploadMax, what, (unsigned long long)bytes);
}

// After enqueuing a build or copy of the structure at `address` on the context stream.
void note_as_written(rtx_context ctx, uint64_t address) {
    std::lock_guard<std::mutex> lock(g_prop_mu);
    PropSlot* slot = nullptr;
    for (PropSlot& p : g_prop)
        if (p.address == address) slot = &p;
    if (stream_capturing(ctx->stream)) {  // no read-back inside a capture: forget the address instead
        if (slot) slot->address = 0;
        return;
    }
    if (!g_prop_host) {
        if (hipHostMalloc((void**)&g_prop_host, kPropSlots * kPropWords * sizeof(uint32_t), hipHostMallocDefault) !=
            hipSuccess) {
            g_prop_host = nullptr;
            if (slot) slot->address = 0;
            return;
        }
        for (uint32_t i = 0; i < kPropSlots; ++i) g_prop[i].host = g_prop_host + i * kPropWords;
    }
    if (!slot) {
        slot = &g_prop[0];
        for (PropSlot& p : g_prop)
            if (p.used < slot->used) slot = &p;
    }
    // a read-back still in flight into this slot finishes before the slot is reused
    if (slot->ready) (void)hipEventSynchronize(slot->ready);
    if (slot->ready && slot->device != ctx->device) {
        (void)hipEventDestroy(slot->ready);
        slot->ready = nullptr;
    }
    if (!slot->ready && hipEventCreateWithFlags(&slot->ready, hipEventDisableTiming) != hipSuccess) {
        slot->ready = nullptr;
        slot->address = 0;
        return;
    }
    slot->device = ctx->device;
    if (hipMemcpyAsync(slot->host, (const uint8_t*)(uintptr_t)address + 32, kPropWords * sizeof(uint32_t),
                       hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipEventRecord(slot->ready, ctx->stream) != hipSuccess) {
        slot->address = 0;
        return;
    }
    slot->address = address;
    slot->stream = ctx->stream;
    slot->done = false;
    slot->used = ++g_prop_clock;
}

// After the context stream has drained: its read-backs are complete.
void note_stream_idle(rtx_context ctx) {
    std::lock_guard<std::mutex> lock(g_prop_mu);
    for (PropSlot& p : g_prop)
        if (p.address && p.device == ctx->device && p.stream == ctx->stream) p.done = true;
}

// Properties of the TLAS at `address` for a launch on the context stream.
TlasProps tlas_props(rtx_context ctx, uint64_t address) {
    const bool capturing = stream_capturing(ctx->stream);
    {
        std::lock_guard<std::mutex> lock(g_prop_mu);
        for (PropSlot& p : g_prop) {
            if (p.address != address) continue;
            if (!p.done) {
                // inside a capture, querying the event would invalidate the capture: the properties are known only
                // if the read-back was seen complete before (rtx_sync or an earlier launch)
                if (capturing) return kUnknownProps;
                if (hipEventSynchronize(p.ready) != hipSuccess) break;
                p.done = true;
            }
            p.used = ++g_prop_clock;
            return TlasProps{p.host[3], p.host[1], true};
        }
    }
    // not recorded (built inside a capture, or the slot was reused): read the header after the context's work
    if (capturing) return kUnknownProps;
    uint32_t w[kPropWords];
    if (hipStreamSynchronize(ctx->stream) != hipSuccess ||
        hipMemcpy(w, (const uint8_t*)(uintptr_t)address + 32, sizeof w, hipMemcpyDeviceToHost) != hipSuccess)
        return kUnknownProps;
    return TlasProps{w[3], w[1], true};
}

// Traversal stack entries per lane: 2 per BVH8 level of the deepest TLAS -> BLAS path + 4 (each level's node step
// pushes at most the rest of its node group and one leaf group; entering a BLAS pushes the TLAS's pending groups),
// or the configured limit; 76 when the depth is unknown.
uint32_t stack_entries(rtx_context ctx, uint32_t depth) {
    if (ctx->cfg.stack_limit) return ctx->cfg.stack_limit;
    return depth ? 2u * depth + 4u : 76u;
}

// Spill entries per lane that a launch with this LDS depth needs when the TLAS depth is unknown (a call recorded
// into a graph cannot read it).  Spill areas are sized for at least that, so that recording a call never has to
// grow them.
uint32_t unknown_depth_gdepth(rtx_context ctx, const rtx::TraceLaunch& L) {
    const uint32_t total = stack_entries(ctx, 0);
    return total - std::min(L.lds_depth, total);
}

// Every traversal and render call passes here, so this is also where a recording of the context stream is noted.
int ensure_gstack(rtx_context ctx, const rtx::TraceLaunch& L) {
    if (stream_capturing(ctx->stream)) ctx->recorded = true;
    ctx->last_stack = L.lds_depth + L.gdepth;
    uint64_t bytes = ctx->gstack_entries * sizeof(uint2);
    const uint64_t per_lane = std::max(L.gdepth, unknown_depth_gdepth(ctx, L));
    const int r = grow_buffer(ctx, (void**)&ctx->gstack, &bytes, (uint64_t)L.grid * L.block_size * per_lane *
                                                                      sizeof(uint2), "stack spill area");
    ctx->gstack_entries = bytes / sizeof(uint2);
    return r;
}

// Side stream, events, ray counter and spill stack of the NRC training wavefront (created on first use), and its
// launch shape: 256 threads per CU of the 128-VGPR traversal variant, which fits beside three blocks of the render
// wavefront's on every SIMD.
int side_chain(rtx_context ctx, const rtx::TraceLaunch& L, rtx::SideChain& side) {
    hipError_t e = hipSuccess;
    if ((!ctx->side_stream || !ctx->side_fork || !ctx->side_join || !ctx->side_state) && stream_capturing(ctx->stream))
        return fail(ctx, RTX_ERROR_OUT_OF_DEVICE_MEMORY,
                    "the NRC side stream is created by the first NRC frame: make one before recording frames");
    if (!ctx->side_stream) e = hipStreamCreateWithFlags(&ctx->side_stream, hipStreamNonBlocking);
    if (e == hipSuccess && !ctx->side_fork) e = hipEventCreateWithFlags(&ctx->side_fork, hipEventDisableTiming);
    if (e == hipSuccess && !ctx->side_join) e = hipEventCreateWithFlags(&ctx->side_join, hipEventDisableTiming);
    if (e == hipSuccess && !ctx->side_state) {
        e = hipMalloc(&ctx->side_state, sizeof(rtx::DeviceState));
        if (e == hipSuccess) e = hipMemset(ctx->side_state, 0, sizeof(rtx::DeviceState));
        if (e != hipSuccess && ctx->side_state) {
            (void)hipFree(ctx->side_state);
            ctx->side_state = nullptr;
        }
    }
    if (e != hipSuccess) return fail(ctx, RTX_ERROR_OUT_OF_DEVICE_MEMORY, "nrc side stream: %s", hipGetErrorString(e));
    rtx::TraceLaunch S = L;
    S.blocks_per_cu = kSideBlocks * 256u / L.block_size > 0 ? kSideBlocks * 256u / L.block_size : 1u;
    S.grid = (uint32_t)ctx->cu_count * S.blocks_per_cu;
    S.waves = 4;
    uint64_t bytes = ctx->side_gstack_entries * sizeof(uint2);
    const uint64_t per_lane = std::max(S.gdepth, unknown_depth_gdepth(ctx, S));
    const int r = grow_buffer(ctx, (void**)&ctx->side_gstack, &bytes,
                              (uint64_t)S.grid * S.block_size * per_lane * sizeof(uint2), "nrc spill stack");
    ctx->side_gstack_entries = bytes / sizeof(uint2);
    if (r) return r;
    side.stream = ctx->side_stream;
    side.fork = ctx->side_fork;
    side.join = ctx->side_join;
    side.state = ctx->side_state;
    side.gstack = ctx->side_gstack;
    side.launch = S;
    return RTX_SUCCESS;
}

rtx::TraceLaunch launch_shape(rtx_context ctx, uint32_t depth) {
    rtx::TraceLaunch L;
    L.block_size = ctx->cfg.block_size;
    L.grid = (uint32_t)ctx->cu_count * ctx->cfg.blocks_per_cu;
    L.blocks_per_cu = ctx->cfg.blocks_per_cu;
    const uint32_t total = stack_entries(ctx, depth);
    L.lds_depth = std::min(ctx->cfg.lds_stack_depth, total);
    L.gdepth = total - L.lds_depth;
    L.ray_chunk = ctx->cfg.ray_chunk;
    L.refill_lanes = ctx->cfg.refill_lanes;
    L.refill_interval = ctx->cfg.refill_interval;
    L.leaf_lanes = ctx->cfg.leaf_lanes;
    const uint32_t waves = (L.block_size / 64u) * L.blocks_per_cu / 4u;  // per SIMD
    L.waves = waves >= 5u ? 5u : (waves == 4u ? 4u : 0u);
    return L;
}

int check_tlas(rtx_context ctx, uint64_t tlas) {
    if (tlas == 0 || (tlas & 255u)) return fail(ctx, RTX_ERROR_INVALID_AS, "TLAS address 0x%llx is not a TLAS",
                                               (unsigned long long)tlas);
    return RTX_SUCCESS;
}

}  // namespace

extern "C" {

int rtx_create_context(int device, void* hip_stream, rtx_context* out_ctx) {
    if (!out_ctx) return RTX_ERROR_INVALID_ARGUMENT;
    *out_ctx = nullptr;
    rtx_context ctx = new (std::nothrow) rtx_context_t();
    if (!ctx) return RTX_ERROR_OUT_OF_HOST_MEMORY;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
        delete ctx;
        return RTX_ERROR_INITIALIZATION_FAILED;
    }
    ctx->device = device;
    ctx->stream = (hipStream_t)hip_stream;  // NULL: the default stream
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        ctx->cu_count = cus;
    e = hipMalloc(&ctx->state, sizeof(rtx::DeviceState));
    if (e == hipSuccess) e = hipMemset(ctx->state, 0, sizeof(rtx::DeviceState));
    // the device's scratch array of rtx_cmd_trace_ray_buffer_noalloc (looked up once per device)
    if (e == hipSuccess) e = rtx::noalloc_setup(device, nullptr);
    if (e != hipSuccess) {
        if (ctx->state) (void)hipFree(ctx->state);
        delete ctx;
        return RTX_ERROR_INITIALIZATION_FAILED;
    }
    *out_ctx = ctx;
    return RTX_SUCCESS;
}

int rtx_destroy_context(rtx_context ctx) {
    if (!ctx) return RTX_ERROR_INVALID_ARGUMENT;
    (void)hipStreamSynchronize(ctx->stream);
    for (rtx_nrc n : ctx->caches) n->ctx = nullptr;  // they can still be destroyed, nothing else
    if (ctx->side_stream) (void)hipStreamSynchronize(ctx->side_stream);
    if (ctx->side_gstack) (void)hipFree(ctx->side_gstack);
    if (ctx->side_state) (void)hipFree(ctx->side_state);
    if (ctx->side_fork) (void)hipEventDestroy(ctx->side_fork);
    if (ctx->side_join) (void)hipEventDestroy(ctx->side_join);
    if (ctx->side_stream) (void)hipStreamDestroy(ctx->side_stream);
    if (ctx->graph) (void)hipGraphExecDestroy(ctx->graph);
    if (ctx->nrc_graph) (void)hipGraphExecDestroy(ctx->nrc_graph);
    if (ctx->capture_stream) (void)hipStreamDestroy(ctx->capture_stream);
    if (ctx->sample_word) (void)hipFree(ctx->sample_word);
    if (ctx->views_buf) (void)hipFree(ctx->views_buf);
    if (ctx->gstack) (void)hipFree(ctx->gstack);
    if (ctx->pt_scratch) (void)hipFree(ctx->pt_scratch);
    if (ctx->ms_buf) (void)hipFree(ctx->ms_buf);
    for (void* p : ctx->retired) (void)hipFree(p);
    if (ctx->state) (void)hipFree(ctx->state);
    delete ctx;
    return RTX_SUCCESS;
}

const char* rtx_last_error_string(rtx_context ctx) { return ctx ? ctx->last_error.c_str() : "no context"; }

void* rtx_get_stream(rtx_context ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int rtx_sync(rtx_context ctx) {
    if (!ctx) return RTX_ERROR_INVALID_ARGUMENT;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "rtx_sync");
    note_stream_idle(ctx);
    uint32_t flags = 0;
    e = hipMemcpy(&flags, &ctx->state->error_flags, 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(ctx, e, "rtx_sync flags");
    uint32_t zero = 0;
    if (flags) (void)hipMemcpy(&ctx->state->error_flags, &zero, 4, hipMemcpyHostToDevice);
    // flags raised by rtx_cmd_trace_ray_buffer_noalloc launches on this device (their state lives in the device's
    // scratch array)
    if (rtx::DeviceState* ns = ctx->used_noalloc ? rtx::noalloc_state(ctx->device) : nullptr) {
        uint32_t nf = 0;
        if (hipMemcpy(&nf, &ns->error_flags, 4, hipMemcpyDeviceToHost) == hipSuccess && nf) {
            (void)hipMemcpy(&ns->error_flags, &zero, 4, hipMemcpyHostToDevice);
            flags |= nf;
        }
    }
    if (flags) {
        if (flags & rtx::kErrStackOverflow)
            return fail(ctx, RTX_ERROR_STACK_OVERFLOW, "traversal stack overflow (more than %u entries per lane)",
                        ctx->last_stack);
        if (flags & rtx::kErrSortTimeout) return fail(ctx, RTX_ERROR_DEVICE_LOST, "radix sort look-back timed out");
        if (flags & rtx::kErrBuildDepth)
            return fail(ctx, RTX_ERROR_INVALID_AS, "BVH deeper than the builder's collapse level bound");
        return fail(ctx, RTX_ERROR_INVALID_AS, "device reported an invalid acceleration structure");
    }
    return RTX_SUCCESS;
}

int rtx_get_blas_build_sizes(rtx_context ctx, const rtx_blas_geometry* geometries, uint32_t geometry_count,
                             uint32_t build_flags, rtx_as_build_sizes* out_sizes) {
    if (!ctx || !out_sizes) return RTX_ERROR_INVALID_ARGUMENT;
    uint64_t n = 0;
    int r = check_geometries(ctx, geometries, geometry_count, &n);
    if (r) return r;
    const bool k64 = (build_flags & RTX_BUILD_MORTON63_BIT) != 0;
    const bool upd = (build_flags & RTX_BUILD_ALLOW_UPDATE_BIT) != 0;
    out_sizes->acceleration_structure_size = rtx::blas_as_bytes((uint32_t)n, upd);
    out_sizes->build_scratch_size = rtx::blas_scratch_bytes((uint32_t)n, k64, geometry_count);
    out_sizes->update_scratch_size = upd ? rtx::blas_scratch_bytes((uint32_t)n, false, geometry_count) : 0;
    return RTX_SUCCESS;
}

int rtx_get_tlas_build_sizes(rtx_context ctx, uint32_t instance_count, uint32_t build_flags,
                             rtx_as_build_sizes* out_sizes) {
    if (!ctx || !out_sizes) return RTX_ERROR_INVALID_ARGUMENT;
    if (instance_count > (1u << 24)) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "instance_count above 2^24");
    const bool k64 = (build_flags & RTX_BUILD_MORTON63_BIT) != 0;
    out_sizes->acceleration_structure_size = rtx::tlas_as_bytes(instance_count);
    out_sizes->build_scratch_size = rtx::tlas_scratch_bytes(instance_count, k64);
    out_sizes->update_scratch_size = 0;
    return RTX_SUCCESS;
}

int rtx_cmd_build_blas(rtx_context ctx, const rtx_blas_geometry* geometries, uint32_t geometry_count,
                       uint32_t build_flags, void* dst_device, uint64_t dst_size, void* scratch_device,
                       uint64_t scratch_size, uint64_t* out_address) {
    if (!ctx) return RTX_ERROR_INVALID_ARGUMENT;
    uint64_t n = 0;
    int r = check_geometries(ctx, geometries, geometry_count, &n);
    if (r) return r;
    const bool k64 = (build_flags & RTX_BUILD_MORTON63_BIT) != 0;
    if (!dst_device || !aligned(dst_device, 256) || !scratch_device || !aligned(scratch_device, 256))
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "dst and scratch must be non-NULL and 256-B aligned");
    if (dst_size < rtx::blas_as_bytes((uint32_t)n, (build_flags & RTX_BUILD_ALLOW_UPDATE_BIT) != 0))
        return fail(ctx, RTX_ERROR_OUT_OF_DEVICE_MEMORY, "dst_size %llu below the size query",
                    (unsigned long long)dst_size);
    if (scratch_size < rtx::blas_scratch_bytes((uint32_t)n, k64, geometry_count))
        return fail(ctx, RTX_ERROR_OUT_OF_DEVICE_MEMORY, "scratch_size %llu below the size query",
                    (unsigned long long)scratch_size);
    if ((r = check_recordable(ctx, (uint64_t)geometry_count * sizeof(rtx::GeomDesc), "geometry descriptors")))
        return r;
    std::vector<rtx::GeomDesc> gd(geometry_count ? geometry_count : 1);
    uint32_t base = 0;
    for (uint32_t i = 0; i < geometry_count; ++i) {
        const rtx_blas_geometry& x = geometries[i];
        rtx::GeomDesc& d = gd[i];
        d.vtx = (const uint8_t*)x.vertex_data_device;
        d.idx = (const uint32_t*)((const uint8_t*)x.index_data_device + x.primitive_offset);
        d.xform = x.transform_data_device ? (const float*)((const uint8_t*)x.transform_data_device + x.transform_offset)
                                          : nullptr;
        d.stride = x.vertex_stride;
        d.prim_count = x.primitive_count;
        d.first_vertex = x.first_vertex;
        d.prim_base = base;
        d.max_vertex = x.max_vertex;
        d.geom_local = i;
        d.flags = x.flags;
        d.type = x.geometry_type;
        if (x.geometry_type == RTX_GEOMETRY_TYPE_AABBS) {
            d.vtx = (const uint8_t*)x.vertex_data_device + x.primitive_offset;
            d.idx = nullptr;
            d.xform = nullptr;
        }
        d.blas = 0;
        base += x.primitive_count;
    }
    hipError_t e = rtx::build_blas(gd.data(), geometry_count, (uint32_t)n, build_flags, dst_device, scratch_device,
                                   &ctx->state->error_flags, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "rtx_cmd_build_blas");
    if (out_address) *out_address = (uint64_t)(uintptr_t)dst_device;
    return RTX_SUCCESS;
}

static int check_batch(rtx_context ctx, const rtx_blas_build_info* infos, uint32_t count, uint32_t flags,
                       uint64_t* total, uint32_t* ngeom) {
    if (count == 0 || !infos) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "no BLAS infos");
    if (count > (1u << 20)) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "more than 2^20 BLASes in one batch");
    if (flags & RTX_BUILD_MORTON63_BIT)
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "batched builds use 30-bit Morton codes");
    if (flags & RTX_BUILD_ALLOW_UPDATE_BIT)
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "batched builds do not keep hierarchies for updates");
    uint64_t n = 0, g = 0;
    for (uint32_t b = 0; b < count; ++b) {
        uint64_t nb = 0;
        int r = check_geometries(ctx, infos[b].geometries, infos[b].geometry_count, &nb);
        if (r) return r;
        n += nb;
        g += infos[b].geometry_count;
    }
    if (n > kMaxPrims) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "too many primitives (%llu)", (unsigned long long)n);
    if (g > (1u << 24)) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "too many geometries");
    *total = n;
    *ngeom = (uint32_t)g;
    return RTX_SUCCESS;
}

int rtx_get_blas_batch_build_sizes(rtx_context ctx, const rtx_blas_build_info* infos, uint32_t count,
                                   uint32_t build_flags, uint64_t* as_sizes, uint64_t* scratch_size) {
    if (!ctx || !scratch_size) return RTX_ERROR_INVALID_ARGUMENT;
    uint64_t n = 0;
    uint32_t ng = 0;
    int r = check_batch(ctx, infos, count, build_flags, &n, &ng);
    if (r) return r;
    if (as_sizes) {
        for (uint32_t b = 0; b < count; ++b) {
            uint64_t nb = 0;
            (void)check_geometries(ctx, infos[b].geometries, infos[b].geometry_count, &nb);
            as_sizes[b] = rtx::blas_as_bytes((uint32_t)nb);
        }
    }
    *scratch_size = rtx::blas_batch_scratch_bytes((uint32_t)n, ng, count);
    return RTX_SUCCESS;
}

int rtx_cmd_build_blas_batch(rtx_context ctx, const rtx_blas_build_info* infos, uint32_t count, uint32_t build_flags,
                             void* scratch_device, uint64_t scratch_size, uint64_t* out_addresses) {
    if (!ctx) return RTX_ERROR_INVALID_ARGUMENT;
    uint64_t n = 0;
    uint32_t ng = 0;
    int r = check_batch(ctx, infos, count, build_flags, &n, &ng);
    if (r) return r;
    if (!scratch_device || !aligned(scratch_device, 256))
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "scratch must be non-NULL and 256-B aligned");
    if (scratch_size < rtx::blas_batch_scratch_bytes((uint32_t)n, ng, count))
        return fail(ctx, RTX_ERROR_OUT_OF_DEVICE_MEMORY, "scratch_size %llu below the size query",
                    (unsigned long long)scratch_size);
    if ((r = check_recordable(ctx, (uint64_t)ng * sizeof(rtx::GeomDesc), "geometry descriptors")) ||
        (r = check_recordable(ctx, (uint64_t)count * sizeof(void*), "BLAS addresses")))
        return r;
    std::vector<rtx::GeomDesc> gd(ng ? ng : 1);
    std::vector<uint2> binfo(count);
    std::vector<rtx::AsHeader*> dst(count);
    uint32_t base = 0, gi = 0;
    for (uint32_t b = 0; b < count; ++b) {
        const rtx_blas_build_info& in = infos[b];
        uint64_t nb = 0;
        (void)check_geometries(ctx, in.geometries, in.geometry_count, &nb);
        if (!in.dst_device || !aligned(in.dst_device, 256))
            return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "BLAS %u: dst must be non-NULL and 256-B aligned", b);
        if (in.dst_size < rtx::blas_as_bytes((uint32_t)nb))
            return fail(ctx, RTX_ERROR_OUT_OF_DEVICE_MEMORY, "BLAS %u: dst_size below the size query", b);
        for (uint32_t i = 0; i < in.geometry_count; ++i, ++gi) {
            const rtx_blas_geometry& x = in.geometries[i];
            rtx::GeomDesc& d = gd[gi];
            d.vtx = (const uint8_t*)x.vertex_data_device;
            d.idx = (const uint32_t*)((const uint8_t*)x.index_data_device + x.primitive_offset);
            d.xform = x.transform_data_device
                          ? (const float*)((const uint8_t*)x.transform_data_device + x.transform_offset)
                          : nullptr;
            d.stride = x.vertex_stride;
            d.prim_count = x.primitive_count;
            d.first_vertex = x.first_vertex;
            d.prim_base = base;
            d.max_vertex = x.max_vertex;
            d.geom_local = i;
        d.flags = x.flags;
        d.type = x.geometry_type;
        if (x.geometry_type == RTX_GEOMETRY_TYPE_AABBS) {
            d.vtx = (const uint8_t*)x.vertex_data_device + x.primitive_offset;
            d.idx = nullptr;
            d.xform = nullptr;
        }
            d.blas = b;
            base += x.primitive_count;
        }
        binfo[b] = make_uint2((uint32_t)nb, in.geometry_count);
        dst[b] = (rtx::AsHeader*)in.dst_device;
    }
    hipError_t e = rtx::build_blas_batch(gd.data(), ng, binfo.data(), dst.data(), count, (uint32_t)n, build_flags,
                                         scratch_device, &ctx->state->error_flags, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "rtx_cmd_build_blas_batch");
    // the host arrays above are read by asynchronous copies from pageable memory, which complete before return
    if (out_addresses)
        for (uint32_t b = 0; b < count; ++b) out_addresses[b] = (uint64_t)(uintptr_t)infos[b].dst_device;
    return RTX_SUCCESS;
}

int rtx_cmd_update_blas(rtx_context ctx, const rtx_blas_geometry* geometries, uint32_t geometry_count,
                        uint64_t blas_address, void* scratch_device, uint64_t scratch_size) {
    if (!ctx) return RTX_ERROR_INVALID_ARGUMENT;
    uint64_t n = 0;
    int r = check_geometries(ctx, geometries, geometry_count, &n);
    if (r) return r;
    if (!blas_address || (blas_address & 255) || !scratch_device || !aligned(scratch_device, 256))
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "BLAS and scratch must be non-NULL and 256-B aligned");
    if (stream_capturing(ctx->stream))  // the synchronisation below would invalidate the caller's capture
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "an update reads the structure's header and cannot be recorded");
    if ((r = rtx_sync(ctx))) return r;
    rtx::AsHeader h;
    hipError_t e = hipMemcpy(&h, (const void*)(uintptr_t)blas_address, sizeof h, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(ctx, e, "rtx_cmd_update_blas header");
    if (h.magic != rtx::kMagicBlas) return fail(ctx, RTX_ERROR_INVALID_AS, "not a BLAS");
    if (!(h.flags & RTX_BUILD_ALLOW_UPDATE_BIT) || (n > 0 && h.update_offset == 0))
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "the BLAS was not built with RTX_BUILD_ALLOW_UPDATE_BIT");
    if (h.prim_count != n || h.geometry_count != geometry_count)
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "an update needs the build's geometry and triangle counts");
    uint32_t content = 0;
    for (uint32_t i = 0; i < geometry_count; ++i)
        if (geometries[i].primitive_count)
            content |= ((geometries[i].flags & RTX_GEOMETRY_OPAQUE_BIT) ? 0u : rtx::kContentNonOpaque) |
                       (geometries[i].geometry_type == RTX_GEOMETRY_TYPE_AABBS ? rtx::kContentAabb : 0u);
    if (content != h.content)  // TLASes built over this BLAS chose their traversal from these bits
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "an update keeps the build's geometry types and opacity flags");
    if (scratch_size < rtx::blas_scratch_bytes((uint32_t)n, false, geometry_count))
        return fail(ctx, RTX_ERROR_OUT_OF_DEVICE_MEMORY, "scratch_size below update_scratch_size");
    std::vector<rtx::GeomDesc> gd(geometry_count ? geometry_count : 1);
    uint32_t base = 0;
    for (uint32_t i = 0; i < geometry_count; ++i) {
        const rtx_blas_geometry& x = geometries[i];
        rtx::GeomDesc& d = gd[i];
        d.vtx = (const uint8_t*)x.vertex_data_device;
        d.idx = (const uint32_t*)((const uint8_t*)x.index_data_device + x.primitive_offset);
        d.xform = x.transform_data_device ? (const float*)((const uint8_t*)x.transform_data_device + x.transform_offset)
                                          : nullptr;
        d.stride = x.vertex_stride;
        d.prim_count = x.primitive_count;
        d.first_vertex = x.first_vertex;
        d.prim_base = base;
        d.max_vertex = x.max_vertex;
        d.geom_local = i;
        d.flags = x.flags;
        d.type = x.geometry_type;
        if (x.geometry_type == RTX_GEOMETRY_TYPE_AABBS) {
            d.vtx = (const uint8_t*)x.vertex_data_device + x.primitive_offset;
            d.idx = nullptr;
            d.xform = nullptr;
        }
        d.blas = 0;
        base += x.primitive_count;
    }
    e = rtx::update_blas(gd.data(), geometry_count, (uint32_t)n, h.flags, (void*)(uintptr_t)blas_address,
                         scratch_device, &ctx->state->error_flags, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "rtx_cmd_update_blas");
    return RTX_SUCCESS;
}

int rtx_cmd_build_tlas(rtx_context ctx, const rtx_instance* instances_device, uint32_t instance_count,
                       uint32_t build_flags, void* dst_device, uint64_t dst_size, void* scratch_device,
                       uint64_t scratch_size, uint64_t* out_address) {
    if (!ctx) return RTX_ERROR_INVALID_ARGUMENT;
    const bool k64 = (build_flags & RTX_BUILD_MORTON63_BIT) != 0;
    if (instance_count > 0 && !instances_device)
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "instances is NULL");
    if (instance_count > (1u << 24)) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "instance_count above 2^24");
    if (!dst_device || !aligned(dst_device, 256) || !scratch_device || !aligned(scratch_device, 256))
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "dst and scratch must be non-NULL and 256-B aligned");
    if (dst_size < rtx::tlas_as_bytes(instance_count))
        return fail(ctx, RTX_ERROR_OUT_OF_DEVICE_MEMORY, "dst_size below the size query");
    if (scratch_size < rtx::tlas_scratch_bytes(instance_count, k64))
        return fail(ctx, RTX_ERROR_OUT_OF_DEVICE_MEMORY, "scratch_size below the size query");
    hipError_t e = rtx::build_tlas(instances_device, instance_count, build_flags, dst_device, scratch_device,
                                   &ctx->state->error_flags, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "rtx_cmd_build_tlas");
    note_as_written(ctx, (uint64_t)(uintptr_t)dst_device);
    if (out_address) *out_address = (uint64_t)(uintptr_t)dst_device;
    return RTX_SUCCESS;
}

int rtx_get_as_info(rtx_context ctx, uint64_t as_address, rtx_as_info* out_info) {
    if (!ctx || !out_info) return RTX_ERROR_INVALID_ARGUMENT;
    if (as_address == 0 || (as_address & 255u)) return fail(ctx, RTX_ERROR_INVALID_AS, "bad AS address");
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "rtx_get_as_info");
    rtx::AsHeader h;
    e = hipMemcpy(&h, (const void*)(uintptr_t)as_address, sizeof(h), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(ctx, e, "rtx_get_as_info copy");
    if (h.magic != rtx::kMagicBlas && h.magic != rtx::kMagicTlas)
        return fail(ctx, RTX_ERROR_INVALID_AS, "no acceleration structure at 0x%llx", (unsigned long long)as_address);
    out_info->type = h.type;
    out_info->primitive_count = h.prim_count;
    out_info->node_count = h.node_count;
    out_info->slot_count = h.slot_count;
    out_info->max_depth = h.type ? h.max_total_depth : h.max_depth;
    out_info->geometry_count = h.geometry_count;
    for (int a = 0; a < 3; ++a) {
        out_info->bounds[a] = h.lo[a];
        out_info->bounds[3 + a] = h.hi[a];
    }
    out_info->sah_cost = h.sah_cost;
    out_info->content_flags = h.content;
    out_info->reserved = 0;
    return RTX_SUCCESS;
}

int rtx_cmd_trace_rays(rtx_context ctx, uint64_t tlas_address, const rtx_sbt* sbt, const rtx_camera* camera,
                       const rtx_trace_params* params, uint32_t width, uint32_t height, float* film_rgba32f_device) {
    if (!ctx) return RTX_ERROR_INVALID_ARGUMENT;
    int r = check_tlas(ctx, tlas_address);
    if (r) return r;
    if (!sbt || !camera || !film_rgba32f_device)
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "sbt, camera and film must be non-NULL");
    if (sbt->hit_record_count > 0 && !sbt->hit_records_device)
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "sbt hit records are NULL");
    if ((uint64_t)width * height > 0xFFFFFFFFull)
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "launch size above 2^32 rays");
    rtx_trace_params p = params ? *params : rtx_trace_params{RTX_RAY_FLAG_OPAQUE, 0xFFu, 0.0f, 100.0f};
    rtx::TraceRequest q{};
    q.tlas = (const void*)(uintptr_t)tlas_address;
    q.mode = 2;
    q.n = width * height;
    q.cull_mask = p.cull_mask & 0xFFu;
    q.ray_flags = p.ray_flags;
    // raygen aspect terms (main.cpp:1038-1039) in the GLSL operation order, tan evaluated in FP64
    const double rad = (double)camera->y_fov_degree * (3.14159265358979323846 / 180.0);
    const float ay = (float)std::tan(rad * 0.5);
    const float t = ay * (float)width;
    q.ay = ay;
    q.ax = t / (float)height;
    for (int k = 0; k < 3; ++k) {
        q.cam[k] = camera->position[k];
        q.miss[k] = sbt->miss_color[k];
    }
    q.width = width;
    q.height = height;
    q.tmin = p.tmin;
    q.tmax = p.tmax;
    q.records = sbt->hit_records_device;
    q.nrec = sbt->hit_record_count;
    q.sbt_offset = sbt->sbt_record_offset;
    q.sbt_stride = sbt->sbt_record_stride;
    q.film = film_rgba32f_device;
    q.state = ctx->state;
    q.stats = ctx->stats;
    const TlasProps tp = tlas_props(ctx, tlas_address);
    q.content = tp.content;
    q.guard = tp.known ? 0u : 3u;  // recorded before the content could be read back: the device picks the variant
    rtx::TraceLaunch L = launch_shape(ctx, tp.depth);
    if ((r = ensure_gstack(ctx, L))) return r;
    q.gstack = ctx->gstack;
    hipError_t e = rtx::trace_rays_launch(q, L, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "rtx_cmd_trace_rays");
    return RTX_SUCCESS;
}

int rtx_cmd_trace_ray_buffer(rtx_context ctx, uint64_t tlas_address, const rtx_ray* rays_device, uint32_t ray_count,
                             uint32_t mode, uint32_t ray_flags, uint32_t cull_mask, rtx_hit* hits_device,
                             uint32_t* occluded_device) {
    if (!ctx) return RTX_ERROR_INVALID_ARGUMENT;
    int r = check_tlas(ctx, tlas_address);
    if (r) return r;
    if (mode != RTX_QUERY_CLOSEST_HIT && mode != RTX_QUERY_ANY_HIT)
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "mode must be RTX_QUERY_CLOSEST_HIT or RTX_QUERY_ANY_HIT");
    if (ray_count > 0 && !rays_device) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "rays is NULL");
    if (ray_count > 0 && mode == RTX_QUERY_CLOSEST_HIT && !hits_device)
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "hits is NULL");
    if (ray_count > 0 && mode == RTX_QUERY_ANY_HIT && !occluded_device)
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "occluded is NULL");
    if (!aligned(rays_device, 16)) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "rays must be 16-B aligned");
    rtx::TraceRequest q{};
    q.tlas = (const void*)(uintptr_t)tlas_address;
    q.mode = mode;
    q.n = ray_count;
    q.cull_mask = cull_mask & 0xFFu;
    q.ray_flags = ray_flags;
    q.rays = rays_device;
    q.hits = hits_device;
    q.occluded = occluded_device;
    q.state = ctx->state;
    q.stats = ctx->stats;
    const TlasProps tp = tlas_props(ctx, tlas_address);
    q.content = tp.content;
    q.guard = tp.known ? 0u : 3u;  // recorded before the content could be read back: the device picks the variant
    rtx::TraceLaunch L = launch_shape(ctx, tp.depth);
    if ((r = ensure_gstack(ctx, L))) return r;
    q.gstack = ctx->gstack;
    hipError_t e = rtx::trace_rays_launch(q, L, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "rtx_cmd_trace_ray_buffer");
    return RTX_SUCCESS;
}

int rtx_cmd_trace_ray_buffer_noalloc(rtx_context ctx, uint64_t tlas_address, const rtx_ray* rays_device,
                                     uint32_t ray_count, uint32_t mode, uint32_t ray_flags, uint32_t cull_mask,
                                     rtx_hit* hits_device, uint32_t* occluded_device) {
    if (!ctx) return RTX_ERROR_INVALID_ARGUMENT;
    int r = check_tlas(ctx, tlas_address);
    if (r) return r;
    if (mode != RTX_QUERY_CLOSEST_HIT && mode != RTX_QUERY_ANY_HIT)
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "mode must be RTX_QUERY_CLOSEST_HIT or RTX_QUERY_ANY_HIT");
    if (ray_count > 0 && !rays_device) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "rays is NULL");
    if (ray_count > 0 && mode == RTX_QUERY_CLOSEST_HIT && !hits_device)
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "hits is NULL");
    if (ray_count > 0 && mode == RTX_QUERY_ANY_HIT && !occluded_device)
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "occluded is NULL");
    if (!aligned(rays_device, 16)) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "rays must be 16-B aligned");
    rtx::TraceRequest q{};
    q.tlas = (const void*)(uintptr_t)tlas_address;
    q.mode = mode;
    q.n = ray_count;
    q.cull_mask = cull_mask & 0xFFu;
    q.ray_flags = ray_flags;
    q.rays = rays_device;
    q.hits = hits_device;
    q.occluded = occluded_device;
    q.stats = false;  // the traversal counters live in the context's state, which this variant does not use
    const TlasProps tp = tlas_props(ctx, tlas_address);
    q.content = tp.content;
    q.guard = tp.known ? 0u : 3u;
    const rtx::TraceLaunch L = launch_shape(ctx, tp.depth);
    ctx->last_stack = L.lds_depth + L.gdepth;
    uint64_t need = 0;
    ctx->used_noalloc = true;
    const hipError_t e = rtx::noalloc_trace(ctx->device, q, L, ctx->stream, &need);
    if (e == hipErrorNotInitialized)
        return fail(ctx, RTX_ERROR_INITIALIZATION_FAILED, "no scratch array on device %d", ctx->device);
    if (e == hipErrorOutOfMemory)
        return fail(ctx, RTX_ERROR_OUT_OF_DEVICE_MEMORY,
                    "the launch needs %llu bytes of scratch (%u stack entries per lane in the global spill area), the "
                    "no-alloc entry point has %llu: use fewer blocks per CU, more LDS stack entries or a shallower TLAS",
                    (unsigned long long)need, L.gdepth, (unsigned long long)rtx::noalloc_scratch_bytes());
    if (e != hipSuccess) return hip_fail(ctx, e, "rtx_cmd_trace_ray_buffer_noalloc");
    return RTX_SUCCESS;
}

static int render_setup(rtx_context ctx, uint64_t tlas_address, const rtx_sbt* sbt, const rtx_camera* camera,
                        const rtx_render_params* params, bool have_out, rtx::PtArgs& a, rtx::TraceLaunch& L,
                        uint32_t extra_paths = 0, uint32_t extra_cdf = 0) {
    int r = check_tlas(ctx, tlas_address);
    if (r) return r;
    if (!sbt || !camera || !params) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "sbt, camera, params must be set");
    const rtx_render_params& p = *params;
    if (p.width == 0 || p.height == 0 || (uint64_t)p.width * p.height > (1ull << 28))
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "film size must be 1..2^28 pixels");
    if (p.max_depth == 0 || p.max_depth > 64) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "max_depth in 1..64");
    if (p.light_triangle_count > 0 && !p.light_triangles_device)
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "light triangles are NULL");
    if (p.light_triangle_count > (1u << 20)) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "too many lights");
    if (!have_out) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "film and accumulator are both NULL");
    if (sbt->hit_record_count > 0 && !sbt->hit_records_device)
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "sbt hit records are NULL");
    const uint32_t npix = p.width * p.height;
    if (stream_capturing(ctx->stream)) ctx->recorded = true;
    const uint64_t need = rtx::pt_scratch_bytes(npix + extra_paths, p.max_depth, p.light_triangle_count + extra_cdf);
    if ((r = grow_buffer(ctx, &ctx->pt_scratch, &ctx->pt_scratch_bytes, need, "path tracer queues"))) return r;
    a = rtx::PtArgs{};
    rtx::pt_layout(a, ctx->pt_scratch, npix + extra_paths, p.max_depth, p.light_triangle_count + extra_cdf);
    a.tlas = (const void*)(uintptr_t)tlas_address;
    a.width = p.width;
    a.height = p.height;
    a.npix = npix;
    a.npaths = npix;
    a.max_depth = p.max_depth;
    a.rr_depth = p.rr_depth;
    a.seed = p.seed;
    a.cache_depth = 2;
    for (int k = 0; k < 3; ++k) a.cam[k] = camera->position[k];
    const double rad = (double)camera->y_fov_degree * (3.14159265358979323846 / 180.0);
    a.ay = (float)std::tan(rad * 0.5);
    const float t = a.ay * (float)p.width;
    a.ax = t / (float)p.height;
    a.ray_offset = p.ray_offset;
    a.records = sbt->hit_records_device;
    a.nrec = sbt->hit_record_count;
    a.sbt_offset = sbt->sbt_record_offset;
    a.sbt_stride = sbt->sbt_record_stride;
    a.lights = p.light_triangles_device;
    a.nlights = p.light_triangle_count;
    a.state = ctx->state;
    const TlasProps tp = tlas_props(ctx, tlas_address);
    a.content = tp.content;
    a.dual = tp.known ? 0u : 1u;  // recorded before the content could be read back: both forms, the device picks
    L = launch_shape(ctx, tp.depth);
    if ((r = ensure_gstack(ctx, L))) return r;
    a.gstack = ctx->gstack;
    ctx->render_form = rtx::pt_dual_forms(a) ? RTX_RENDER_FORM_DEVICE_CHOSEN
                       : rtx::pt_whole_paths(a) ? RTX_RENDER_FORM_WHOLE_PATH : RTX_RENDER_FORM_WAVEFRONT;
    return RTX_SUCCESS;
}

int rtx_cmd_film_to_rgba8(rtx_context ctx, const float* film_rgba32f_device, uint32_t width, uint32_t height,
                          uint32_t flags, uint8_t* out_rgba8_device) {
    if (!ctx) return RTX_ERROR_INVALID_ARGUMENT;
    const uint64_t npix = (uint64_t)width * height;
    if (npix > 0 && (!film_rgba32f_device || !out_rgba8_device))
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "film or output is NULL");
    if (!aligned(film_rgba32f_device, 16) || !aligned(out_rgba8_device, 4))
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "film must be 16-B and output 4-B aligned");
    if (flags & ~(RTX_FILM_SRGB_BIT | RTX_FILM_SWAP_RB_BIT | RTX_FILM_OPAQUE_BIT))
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "unknown film flags");
    hipError_t e = rtx::film_to_rgba8(film_rgba32f_device, npix, flags, out_rgba8_device, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "rtx_cmd_film_to_rgba8");
    return RTX_SUCCESS;
}

int rtx_get_as_compacted_size(rtx_context ctx, uint64_t as_address, uint64_t* out_bytes) {
    if (!ctx || !out_bytes) return RTX_ERROR_INVALID_ARGUMENT;
    int r = rtx_sync(ctx);
    if (r) return r;
    rtx::AsHeader h;
    hipError_t e = hipMemcpy(&h, (const void*)(uintptr_t)as_address, sizeof h, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(ctx, e, "rtx_get_as_compacted_size");
    if (h.magic != rtx::kMagicBlas && h.magic != rtx::kMagicTlas)
        return fail(ctx, RTX_ERROR_INVALID_AS, "not an acceleration structure");
    *out_bytes = rtx::compacted_bytes(h);
    return RTX_SUCCESS;
}

int rtx_cmd_copy_as(rtx_context ctx, uint64_t src_address, uint32_t mode, void* dst_device, uint64_t dst_size,
                    uint64_t* out_address) {
    if (!ctx) return RTX_ERROR_INVALID_ARGUMENT;
    if (mode != RTX_COPY_MODE_CLONE && mode != RTX_COPY_MODE_COMPACT)
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "mode must be CLONE or COMPACT");
    if (!src_address || (src_address & 255) || !dst_device || !aligned(dst_device, 256))
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "source and destination must be 256-B aligned");
    hipError_t e = rtx::copy_as((const void*)(uintptr_t)src_address, dst_device, dst_size,
                                mode == RTX_COPY_MODE_COMPACT, &ctx->state->error_flags, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "rtx_cmd_copy_as");
    note_as_written(ctx, (uint64_t)(uintptr_t)dst_device);
    if (out_address) *out_address = (uint64_t)(uintptr_t)dst_device;
    return RTX_SUCCESS;
}

int rtx_cmd_render(rtx_context ctx, uint64_t tlas_address, const rtx_sbt* sbt, const rtx_camera* camera,
                   const rtx_render_params* params, float* film_rgba32f_device, double* accum_rgb_device) {
    if (!ctx) return RTX_ERROR_INVALID_ARGUMENT;
    rtx::PtArgs a;
    rtx::TraceLaunch L;
    int r = render_setup(ctx, tlas_address, sbt, camera, params, film_rgba32f_device || accum_rgb_device, a, L);
    if (r) return r;
    // whole-path renders of several samples run up to kSampleViews samples per launch (one launch tail per group)
    const uint32_t S = (uint32_t)std::min<uint64_t>(
        std::min<uint32_t>(params->spp, kSampleViews), std::max<uint64_t>(1, kSampleViewPaths / std::max(1u, a.npix)));
    if (S > 1 && rtx::pt_whole_paths(a) && !rtx::pt_dual_forms(a)) {
        if ((r = grow_buffer(ctx, &ctx->ms_buf, &ctx->ms_bytes, rtx::pt_sample_views_bytes(a.npix, S),
                             "multi-sample path state")))
            return r;
        a.ms_buf = ctx->ms_buf;
        a.ms_max = S;
    }
    hipError_t e = rtx::pt_render(a, L, params->spp, params->sample_offset, film_rgba32f_device, accum_rgb_device,
                                  ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "rtx_cmd_render");
    return RTX_SUCCESS;
}

namespace {
__global__ void set_word_kernel(uint32_t* p, uint32_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

uint64_t fnv1a(uint64_t h, const void* data, size_t n) {
    const uint8_t* p = (const uint8_t*)data;
    for (size_t i = 0; i < n; ++i) h = (h ^ p[i]) * 1099511628211ull;
    return h;
}

// The word recorded launches read sample_offset from, and the private stream calls are captured on.
int graph_resources(rtx_context ctx) {
    if (!ctx->sample_word && hipMalloc(&ctx->sample_word, 256) != hipSuccess) {
        ctx->sample_word = nullptr;
        return fail(ctx, RTX_ERROR_OUT_OF_DEVICE_MEMORY, "graph sample word");
    }
    hipError_t e;
    if (!ctx->capture_stream && (e = hipStreamCreateWithFlags(&ctx->capture_stream, hipStreamNonBlocking)))
        return hip_fail(ctx, e, "capture stream");
    return RTX_SUCCESS;
}

// Record record(stream) on the capture stream into *exec (replacing the previous graph).
extern "C++" template <typename Record>
int capture_graph(rtx_context ctx, hipGraphExec_t* exec, Record record, const char* who) {
    if (*exec) {
        (void)hipGraphExecDestroy(*exec);
        *exec = nullptr;
    }
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(ctx->capture_stream, hipStreamCaptureModeThreadLocal);
    if (e) return hip_fail(ctx, e, who);
    const hipError_t le = record(ctx->capture_stream);
    e = hipStreamEndCapture(ctx->capture_stream, &g);
    if (le != hipSuccess || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail(ctx, le != hipSuccess ? le : e, who);
    }
    e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        *exec = nullptr;
        return hip_fail(ctx, e, who);
    }
    return RTX_SUCCESS;
}

// Replay: the sample offset into the word the recorded launches read, then the graph, on the context stream.
int replay_graph(rtx_context ctx, hipGraphExec_t exec, uint32_t sample_offset, const char* who) {
    set_word_kernel<<<1, 64, 0, ctx->stream>>>(ctx->sample_word, sample_offset);
    hipError_t e;
    if ((e = hipGetLastError()) || (e = hipGraphLaunch(exec, ctx->stream))) return hip_fail(ctx, e, who);
    return RTX_SUCCESS;
}
}  // namespace

int rtx_cmd_render_batch(rtx_context ctx, uint64_t tlas_address, const rtx_sbt* sbt, const rtx_render_view* views,
                         uint32_t view_count, const rtx_render_params* params, float* films_device,
                         double* accum_rgb_device) {
    if (!ctx) return RTX_ERROR_INVALID_ARGUMENT;
    if (!views || view_count == 0 || !params) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "no views or params");
    if (view_count > (1u << 16)) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "more than 65536 views");
    uint64_t npix = 0;
    std::vector<rtx::BatchView> bv(view_count);
    for (uint32_t v = 0; v < view_count; ++v) {
        const rtx_render_view& w = views[v];
        if (w.width == 0 || w.height == 0) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "view %u: empty film", v);
        if ((uint64_t)w.light_first + w.light_count > params->light_triangle_count)
            return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "view %u: lights out of range", v);
        for (int k = 0; k < 3; ++k)
            if (!(w.clip_min[k] <= w.clip_max[k]))
                return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "view %u: empty clip box", v);
        rtx::BatchView& b = bv[v];
        for (int k = 0; k < 3; ++k) {
            b.cam[k] = w.camera.position[k];
            b.lo[k] = w.clip_min[k];
            b.hi[k] = w.clip_max[k];
        }
        const double rad = (double)w.camera.y_fov_degree * (3.14159265358979323846 / 180.0);
        b.ay = (float)std::tan(rad * 0.5);
        const float t = b.ay * (float)w.width;
        b.ax = t / (float)w.height;
        b.width = w.width;
        b.height = w.height;
        b.pix_base = (uint32_t)npix;
        b.light_first = w.light_first;
        b.light_count = w.light_count;
        b.cdf_first = w.light_first + v;
        npix += (uint64_t)w.width * w.height;
        if (npix > (1ull << 28)) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "more than 2^28 pixels in a batch");
    }
    // render_setup sizes the queues for the whole batch (one CDF entry more per view); the views carry the cameras
    rtx_render_params p = *params;
    p.width = (uint32_t)npix;
    p.height = 1;
    static const rtx_camera kViewsCarryCameras = {{0.0f, 0.0f, 0.0f}, 60.0f};
    rtx::PtArgs a;
    rtx::TraceLaunch L;
    int r = render_setup(ctx, tlas_address, sbt, &kViewsCarryCameras, &p, films_device || accum_rgb_device, a, L, 0,
                         view_count - 1);
    if (r) return r;
    // two view slots: the recorded graph's views, and those of direct calls (which must not disturb the graph)
    const uint64_t vbytes = view_count * sizeof(rtx::BatchView);
    if ((r = check_recordable(ctx, vbytes, "batch views"))) return r;
    const uint64_t vslot = (vbytes + 255) & ~255ull;
    if (2 * vslot > ctx->views_bytes) {
        ctx->graph_key = 0;
        if ((r = grow_buffer(ctx, &ctx->views_buf, &ctx->views_bytes, 2 * vslot, "batch views"))) return r;
    }
    // a call that is itself being recorded (a capture of the context stream) is recorded as a direct call
    const bool graph = (params->flags & RTX_RENDER_GRAPH_BIT) != 0 && !stream_capturing(ctx->stream);
    void* vdst = graph ? ctx->views_buf : (void*)((uint8_t*)ctx->views_buf + ctx->views_bytes / 2);
    a.views = (const rtx::BatchView*)vdst;
    a.nviews = view_count;
    // key of the recorded sequence: everything the launches read, except sample_offset
    uint64_t key = 1469598103934665603ull;
    key = fnv1a(key, &tlas_address, sizeof tlas_address);
    key = fnv1a(key, sbt, sizeof *sbt);
    key = fnv1a(key, bv.data(), vbytes);
    rtx_render_params kp = *params;
    kp.sample_offset = 0;
    key = fnv1a(key, &kp, sizeof kp);
    key = fnv1a(key, &films_device, sizeof films_device);
    key = fnv1a(key, &accum_rgb_device, sizeof accum_rgb_device);
    key = fnv1a(key, &ctx->cfg, sizeof ctx->cfg);
    // the launch shape and traversal variant the recorded launches bake in (TLAS content, stack, integrator form)
    key = fnv1a(key, &L, sizeof L);
    key = fnv1a(key, &a.content, sizeof a.content);
    key = fnv1a(key, &ctx->render_form, sizeof ctx->render_form);
    key = fnv1a(key, &ctx->pt_scratch, sizeof ctx->pt_scratch);
    key = fnv1a(key, &ctx->gstack, sizeof ctx->gstack);
    key = fnv1a(key, &ctx->views_buf, sizeof ctx->views_buf);
    hipError_t e;
    if (!graph) {
        if ((e = rtx::upload_host_array(vdst, bv.data(), vbytes, ctx->stream)))
            return hip_fail(ctx, e, "rtx_cmd_render_batch views");
        e = rtx::pt_render_batch(a, L, params->spp, params->sample_offset, films_device, accum_rgb_device,
                                 ctx->stream);
        if (e != hipSuccess) return hip_fail(ctx, e, "rtx_cmd_render_batch");
        return RTX_SUCCESS;
    }
    if ((r = graph_resources(ctx))) return r;
    if (!ctx->graph || ctx->graph_key != key) {
        if ((e = hipMemcpyAsync(vdst, bv.data(), vbytes, hipMemcpyHostToDevice, ctx->stream)))
            return hip_fail(ctx, e, "rtx_cmd_render_batch views");
        ctx->graph_key = 0;
        a.sample_dev = ctx->sample_word;
        if ((r = capture_graph(ctx, &ctx->graph, [&](hipStream_t cs) {
                 return rtx::pt_render_batch(a, L, params->spp, 0u, films_device, accum_rgb_device, cs);
             }, "rtx_cmd_render_batch capture")))
            return r;
        ctx->graph_key = key;
    }
    return replay_graph(ctx, ctx->graph, params->sample_offset, "rtx_cmd_render_batch launch");
}

int rtx_cmd_render_nrc(rtx_context ctx, uint64_t tlas_address, const rtx_sbt* sbt, const rtx_camera* camera,
                       const rtx_render_params* params, rtx_nrc nrc, uint32_t train_paths,
                       float* film_rgba32f_device, double* accum_rgb_device) {
    if (!ctx || !nrc) return RTX_ERROR_INVALID_ARGUMENT;
    if (nrc->ctx != ctx) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "the cache belongs to another context");
    rtx::PtArgs a;
    rtx::TraceLaunch L;
    if (params && (uint64_t)params->width * params->height + train_paths > (1ull << 28))
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "pixels + train_paths above 2^28");
    const bool train_only = params && params->spp == 0;
    if (train_only && train_paths == 0)
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "a frame needs spp > 0 or train_paths > 0");
    int r = render_setup(ctx, tlas_address, sbt, camera, params, film_rgba32f_device || accum_rgb_device || train_only,
                         a, L, train_paths);
    if (r) return r;
    if (a.dual) {  // NRC frames record one form: the wavefront where the TLAS content is not known
        a.dual = 0;
        ctx->render_form = RTX_RENDER_FORM_WAVEFRONT;
    }
    if (params->max_depth <= nrc->st.cache_vertex)
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "max_depth must exceed the cache vertex");
    a.cache_depth = nrc->st.cache_vertex;
    const uint32_t rec_depths = rtx::nrc_record_depths(params->max_depth, a.cache_depth, nrc->st.train_suffix);
    const uint64_t need = rtx::pt_train_bytes(train_paths, rec_depths);
    if ((r = grow_buffer(ctx, &nrc->train_buf, &nrc->train_buf_bytes, need, "nrc training buffers"))) return r;
    rtx::SideChain side{};
    if ((r = side_chain(ctx, L, side))) return r;
    // the pixel paths leave kSideReserve 256-thread blocks per CU to the training paths
    const uint32_t reserve = kSideReserve * 256u / L.block_size > 0 ? kSideReserve * 256u / L.block_size : 1u;
    if (L.blocks_per_cu > reserve) {
        L.blocks_per_cu -= reserve;
        L.grid = (uint32_t)ctx->cu_count * L.blocks_per_cu;
    }
    if (!(params->flags & RTX_RENDER_GRAPH_BIT) || stream_capturing(ctx->stream)) {
        hipError_t e = rtx::pt_render_nrc(a, L, side, nrc->st, train_paths, nrc->st.train_suffix, nrc->train_buf,
                                          params->spp, params->sample_offset, film_rgba32f_device, accum_rgb_device,
                                          ctx->stream);
        if (e != hipSuccess) return hip_fail(ctx, e, "rtx_cmd_render_nrc");
        return RTX_SUCCESS;
    }
    // the frame recorded once into a graph (both streams: the training wavefront's side stream joins the capture
    // through the fork / join events) and replayed; the cache's parameters, moments and step counter live in device
    // memory, so replays train on; sample_offset is read from the context's sample word
    uint64_t key = 1469598103934665603ull;
    key = fnv1a(key, &tlas_address, sizeof tlas_address);
    key = fnv1a(key, sbt, sizeof *sbt);
    key = fnv1a(key, camera, sizeof *camera);
    rtx_render_params kp = *params;
    kp.sample_offset = 0;
    key = fnv1a(key, &kp, sizeof kp);
    key = fnv1a(key, &nrc, sizeof nrc);
    key = fnv1a(key, &nrc->train_buf, sizeof nrc->train_buf);
    key = fnv1a(key, &train_paths, sizeof train_paths);
    key = fnv1a(key, &film_rgba32f_device, sizeof film_rgba32f_device);
    key = fnv1a(key, &accum_rgb_device, sizeof accum_rgb_device);
    key = fnv1a(key, &L, sizeof L);
    key = fnv1a(key, &side.launch, sizeof side.launch);
    key = fnv1a(key, &side.gstack, sizeof side.gstack);
    key = fnv1a(key, &a.content, sizeof a.content);
    key = fnv1a(key, &ctx->render_form, sizeof ctx->render_form);
    key = fnv1a(key, &ctx->pt_scratch, sizeof ctx->pt_scratch);
    key = fnv1a(key, &ctx->gstack, sizeof ctx->gstack);
    if ((r = graph_resources(ctx))) return r;
    if (!ctx->nrc_graph || ctx->nrc_graph_key != key) {
        ctx->nrc_graph_key = 0;
        a.sample_dev = ctx->sample_word;
        if ((r = capture_graph(ctx, &ctx->nrc_graph, [&](hipStream_t cs) {
                 return rtx::pt_render_nrc(a, L, side, nrc->st, train_paths, nrc->st.train_suffix, nrc->train_buf,
                                           params->spp, 0u, film_rgba32f_device, accum_rgb_device, cs);
             }, "rtx_cmd_render_nrc capture")))
            return r;
        ctx->nrc_graph_key = key;
    }
    return replay_graph(ctx, ctx->nrc_graph, params->sample_offset, "rtx_cmd_render_nrc launch");
}

int rtx_get_render_ray_count(rtx_context ctx, uint64_t* out_rays) {
    if (!ctx || !out_rays) return RTX_ERROR_INVALID_ARGUMENT;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "rtx_get_render_ray_count");
    rtx::DeviceState st;
    e = hipMemcpy(&st, ctx->state, sizeof(st), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(ctx, e, "rtx_get_render_ray_count copy");
    *out_rays = st.pt_rays;
    return RTX_SUCCESS;
}

int rtx_get_render_form(rtx_context ctx, uint32_t* out_form) {
    if (!ctx || !out_form) return RTX_ERROR_INVALID_ARGUMENT;
    *out_form = ctx->render_form;
    return RTX_SUCCESS;
}

int rtx_nrc_create(rtx_context ctx, const rtx_nrc_config* cfg, rtx_nrc* out_nrc) {
    if (!ctx || !cfg || !out_nrc) return RTX_ERROR_INVALID_ARGUMENT;
    *out_nrc = nullptr;
    if (cfg->batch_size == 0 || (cfg->batch_size % 64) || cfg->batch_size > (1u << 24))
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "batch_size must be a positive multiple of 64 up to 2^24");
    rtx_nrc n = new (std::nothrow) rtx_nrc_t();
    if (!n) return RTX_ERROR_OUT_OF_HOST_MEMORY;
    n->ctx = ctx;
    rtx::NrcState& st = n->st;
    st.batch_size = cfg->batch_size;
    st.max_blocks = cfg->batch_size / 64;
    st.cu_count = (uint32_t)ctx->cu_count;
    st.lr = cfg->learning_rate;
    st.beta1 = cfg->beta1;
    st.beta2 = cfg->beta2;
    st.eps = cfg->epsilon;
    st.cache_vertex = cfg->cache_vertex ? cfg->cache_vertex : 1u;
    st.train_suffix = cfg->train_suffix;
    if (st.train_suffix > 62) {
        delete n;
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "train_suffix above 62");
    }
    if (st.cache_vertex > 62) {
        delete n;
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "cache_vertex above 62");
    }
    for (int a = 0; a < 3; ++a) {
        st.lo[a] = cfg->scene_min[a];
        st.hi[a] = cfg->scene_max[a];
    }
    const uint64_t pbytes = RTX_NRC_PARAM_COUNT * sizeof(float);
    hipError_t e = hipMalloc(&st.params, pbytes);
    if (!e) e = hipMalloc(&st.m, pbytes);
    if (!e) e = hipMalloc(&st.v, pbytes);
    if (!e) e = hipMalloc(&st.fwd, rtx::nrc_fragment_bytes());
    st.bwd = nullptr;
    if (!e) e = hipMalloc(&st.partials, (uint64_t)st.max_blocks * pbytes);
    if (!e) e = hipMalloc(&st.loss, sizeof(float) * 4);
    if (!e) e = hipMalloc(&st.loss_part, sizeof(float) * st.max_blocks);
    if (!e) e = hipMalloc(&st.steps_dev, 256);
    if (e) {
        rtx_nrc_destroy(n);
        return fail(ctx, RTX_ERROR_OUT_OF_DEVICE_MEMORY, "nrc buffers: %s", hipGetErrorString(e));
    }
    st.bwd = (uint8_t*)st.fwd + 46 * 64 * 16;
    std::vector<float> p;
    nrc_init_params(cfg->seed, p);
    std::vector<float> zeros(RTX_NRC_PARAM_COUNT, 0.0f);
    e = hipMemcpy(st.params, p.data(), pbytes, hipMemcpyHostToDevice);
    if (!e) e = hipMemcpy(st.m, zeros.data(), pbytes, hipMemcpyHostToDevice);
    if (!e) e = hipMemcpy(st.v, zeros.data(), pbytes, hipMemcpyHostToDevice);
    if (!e) e = hipMemcpy(st.loss, zeros.data(), sizeof(float) * 4, hipMemcpyHostToDevice);
    if (!e) e = hipMemcpy(st.steps_dev, zeros.data(), sizeof(uint32_t), hipMemcpyHostToDevice);
    if (!e) e = rtx::nrc_pack(st, ctx->stream);
    if (!e) e = hipStreamSynchronize(ctx->stream);
    if (e) {
        rtx_nrc_destroy(n);
        return hip_fail(ctx, e, "rtx_nrc_create");
    }
    ctx->caches.push_back(n);
    *out_nrc = n;
    return RTX_SUCCESS;
}

int rtx_nrc_destroy(rtx_nrc n) {
    if (!n) return RTX_ERROR_INVALID_ARGUMENT;
    if (n->ctx) {
        (void)hipStreamSynchronize(n->ctx->stream);
        auto& v = n->ctx->caches;
        v.erase(std::remove(v.begin(), v.end(), n), v.end());
    }
    rtx::NrcState& st = n->st;
    if (st.params) (void)hipFree(st.params);
    if (st.m) (void)hipFree(st.m);
    if (st.v) (void)hipFree(st.v);
    if (st.fwd) (void)hipFree(st.fwd);
    if (st.partials) (void)hipFree(st.partials);
    if (st.loss) (void)hipFree(st.loss);
    if (st.loss_part) (void)hipFree(st.loss_part);
    if (st.steps_dev) (void)hipFree(st.steps_dev);
    if (n->train_buf) (void)hipFree(n->train_buf);
    delete n;
    return RTX_SUCCESS;
}

int rtx_nrc_cmd_infer(rtx_nrc n, const rtx_nrc_query* queries_device, uint32_t count, float* out_rgb_device) {
    if (!n || !n->ctx) return RTX_ERROR_INVALID_ARGUMENT;  // NULL, or its context was destroyed
    if (count > 0 && (!queries_device || !out_rgb_device))
        return fail(n->ctx, RTX_ERROR_INVALID_ARGUMENT, "NULL buffer");
    if (!aligned(queries_device, 16)) return fail(n->ctx, RTX_ERROR_INVALID_ARGUMENT, "queries must be 16-B aligned");
    hipError_t e = rtx::nrc_infer(n->st, queries_device, count, out_rgb_device, n->ctx->stream);
    if (e) return hip_fail(n->ctx, e, "rtx_nrc_cmd_infer");
    return RTX_SUCCESS;
}

int rtx_nrc_cmd_train(rtx_nrc n, const rtx_nrc_sample* samples_device, uint32_t count) {
    if (!n || !n->ctx) return RTX_ERROR_INVALID_ARGUMENT;  // NULL, or its context was destroyed
    if (count > 0 && !samples_device) return fail(n->ctx, RTX_ERROR_INVALID_ARGUMENT, "NULL buffer");
    if (!aligned(samples_device, 16)) return fail(n->ctx, RTX_ERROR_INVALID_ARGUMENT, "samples must be 16-B aligned");
    for (uint64_t off = 0; off < count; off += n->st.batch_size) {
        const uint32_t m = (uint32_t)std::min<uint64_t>(n->st.batch_size, count - off);
        hipError_t e = rtx::nrc_step(n->st, samples_device + off, m, nullptr, nullptr, n->ctx->stream);
        if (e) return hip_fail(n->ctx, e, "rtx_nrc_cmd_train");
    }
    return RTX_SUCCESS;
}

int rtx_nrc_debug_gradient(rtx_nrc n, const rtx_nrc_sample* samples_device, uint32_t count, float* grad_device,
                           float* loss_device) {
    if (!n || !n->ctx) return RTX_ERROR_INVALID_ARGUMENT;  // NULL, or its context was destroyed
    if (!grad_device || !loss_device) return RTX_ERROR_INVALID_ARGUMENT;
    if (count == 0 || count > n->st.batch_size)
        return fail(n->ctx, RTX_ERROR_INVALID_ARGUMENT, "count must be in 1..batch_size");
    hipError_t e = rtx::nrc_step(n->st, samples_device, count, grad_device, loss_device, n->ctx->stream);
    if (e) return hip_fail(n->ctx, e, "rtx_nrc_debug_gradient");
    return RTX_SUCCESS;
}

int rtx_nrc_get_params(rtx_nrc n, float* params_host) {
    if (!n || !n->ctx) return RTX_ERROR_INVALID_ARGUMENT;  // NULL, or its context was destroyed
    if (!params_host) return RTX_ERROR_INVALID_ARGUMENT;
    hipError_t e = hipStreamSynchronize(n->ctx->stream);
    if (!e) e = hipMemcpy(params_host, n->st.params, RTX_NRC_PARAM_COUNT * sizeof(float), hipMemcpyDeviceToHost);
    if (e) return hip_fail(n->ctx, e, "rtx_nrc_get_params");
    return RTX_SUCCESS;
}

int rtx_nrc_set_params(rtx_nrc n, const float* params_host) {
    if (!n || !n->ctx) return RTX_ERROR_INVALID_ARGUMENT;  // NULL, or its context was destroyed
    if (!params_host) return RTX_ERROR_INVALID_ARGUMENT;
    hipError_t e = hipStreamSynchronize(n->ctx->stream);
    if (!e) e = hipMemcpy(n->st.params, params_host, RTX_NRC_PARAM_COUNT * sizeof(float), hipMemcpyHostToDevice);
    if (!e) e = rtx::nrc_pack(n->st, n->ctx->stream);
    if (e) return hip_fail(n->ctx, e, "rtx_nrc_set_params");
    return RTX_SUCCESS;
}

int rtx_nrc_get_status(rtx_nrc n, uint32_t* steps, float* last_loss) {
    if (!n || !n->ctx) return RTX_ERROR_INVALID_ARGUMENT;  // NULL, or its context was destroyed
    hipError_t e = hipStreamSynchronize(n->ctx->stream);
    float l = 0.0f;
    uint32_t k = 0;
    if (!e) e = hipMemcpy(&l, n->st.loss, sizeof(float), hipMemcpyDeviceToHost);
    if (!e) e = hipMemcpy(&k, n->st.steps_dev, sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (e) return hip_fail(n->ctx, e, "rtx_nrc_get_status");
    if (steps) *steps = k;
    if (last_loss) *last_loss = l;
    return RTX_SUCCESS;
}

int rtx_set_trace_stats(rtx_context ctx, int enable) {
    if (!ctx) return RTX_ERROR_INVALID_ARGUMENT;
    ctx->stats = enable != 0;
    return RTX_SUCCESS;
}

int rtx_get_trace_stats(rtx_context ctx, rtx_trace_stats* out_stats) {
    if (!ctx || !out_stats) return RTX_ERROR_INVALID_ARGUMENT;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "rtx_get_trace_stats");
    rtx::DeviceState st;
    e = hipMemcpy(&st, ctx->state, sizeof(st), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(ctx, e, "rtx_get_trace_stats copy");
    out_stats->rays = st.stat_rays;
    out_stats->node_fetches = st.stat_nodes;
    out_stats->triangle_tests = st.stat_tris;
    out_stats->instance_entries = st.stat_insts;
    return RTX_SUCCESS;
}

int rtx_set_trace_config(rtx_context ctx, const rtx_trace_config* cfg) {
    if (!ctx || !cfg) return RTX_ERROR_INVALID_ARGUMENT;
    rtx_trace_config c = ctx->cfg;
    if (cfg->block_size) c.block_size = cfg->block_size;
    if (cfg->blocks_per_cu) c.blocks_per_cu = cfg->blocks_per_cu;
    if (cfg->lds_stack_depth) c.lds_stack_depth = cfg->lds_stack_depth;
    if (cfg->ray_chunk) c.ray_chunk = cfg->ray_chunk;
    if (cfg->refill_lanes) c.refill_lanes = cfg->refill_lanes;
    if (cfg->refill_interval) c.refill_interval = cfg->refill_interval;
    if (cfg->leaf_lanes) c.leaf_lanes = cfg->leaf_lanes;
    if (cfg->stack_limit) c.stack_limit = cfg->stack_limit == RTX_TRACE_STACK_LIMIT_DERIVED ? 0u : cfg->stack_limit;
    if (c.stack_limit > 4096) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "stack_limit above 4096 entries");
    if (c.leaf_lanes < 1 || c.leaf_lanes > 64)
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "leaf_lanes must be in [1, 64]");
    if (c.ray_chunk != RTX_TRACE_RAY_CHUNK_AUTO && (c.ray_chunk < 64 || c.ray_chunk > 4096 || (c.ray_chunk & 63)))
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "ray_chunk must be a multiple of 64 up to 4096, or AUTO");
    if (c.refill_lanes < 1 || c.refill_lanes > 64)
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "refill_lanes must be in [1, 64]");
    if (c.refill_interval < 1 || c.refill_interval > 64)
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "refill_interval must be in [1, 64]");
    if (c.block_size < 64 || c.block_size > 256 || (c.block_size & 63))
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "block_size must be 64, 128, 192 or 256");
    if (c.blocks_per_cu < 1 || c.blocks_per_cu > 32)
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "blocks_per_cu must be in [1, 32]");
    // per lane: the stack entries (8 B each) and the world ray (24 B)
    if (c.lds_stack_depth < 1 || ((uint64_t)c.lds_stack_depth * 8 + 24) * c.block_size > 64 * 1024)
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "LDS stack above 64 KiB per block");
    ctx->cfg = c;
    return RTX_SUCCESS;
}

int rtx_get_trace_config(rtx_context ctx, rtx_trace_config* out_cfg) {
    if (!ctx || !out_cfg) return RTX_ERROR_INVALID_ARGUMENT;
    *out_cfg = ctx->cfg;
    return RTX_SUCCESS;
}

// Coordinate descent over the traversal knobs, each candidate timed with HIP events on the context stream; a
// candidate replaces the incumbent only when it is faster by more than timing noise (1.5%).
constexpr float kTuneGain = 0.985f;

extern "C++" template <typename Launch>
static int autotune_core(rtx_context ctx, const char* who, uint32_t reps, Launch launch, rtx_trace_config* out_cfg,
                         float* out_ms) {
    if (reps == 0 || reps > 64) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "%s: reps must be in [1, 64]", who);
    hipEvent_t ev[2];
    hipError_t e = hipEventCreate(&ev[0]);
    if (e == hipSuccess) {
        e = hipEventCreate(&ev[1]);
        if (e != hipSuccess) (void)hipEventDestroy(ev[0]);
    }
    if (e != hipSuccess) return hip_fail(ctx, e, who);
    const rtx_trace_config start = ctx->cfg;
    // time one candidate; +inf when the configuration is rejected
    auto time_cfg = [&](const rtx_trace_config& c, float& ms) -> int {
        ms = INFINITY;
        const rtx_trace_config keep = ctx->cfg;
        if (rtx_set_trace_config(ctx, &c) != RTX_SUCCESS) {
            ctx->cfg = keep;
            ctx->last_error.clear();
            return RTX_SUCCESS;
        }
        int rr = launch();
        if (rr) return rr;
        hipError_t he = hipEventRecord(ev[0], ctx->stream);
        for (uint32_t i = 0; he == hipSuccess && i < reps; ++i)
            if ((rr = launch())) return rr;
        if (he == hipSuccess) he = hipEventRecord(ev[1], ctx->stream);
        if (he == hipSuccess) he = hipEventSynchronize(ev[1]);
        float el = 0.0f;
        if (he == hipSuccess) he = hipEventElapsedTime(&el, ev[0], ev[1]);
        if (he != hipSuccess) return hip_fail(ctx, he, who);
        ms = el / (float)reps;
        return rtx_sync(ctx);
    };
    static const uint32_t shapes[][2] = {{256, 2}, {256, 3}, {256, 4}, {128, 6},  {128, 8},
                                         {64, 12}, {64, 16}, {256, 5}, {128, 10}, {64, 20}};
    static const uint32_t refills[][2] = {{40, 16}, {32, 8}, {40, 4}, {48, 4}, {48, 2}, {56, 2}, {56, 4}, {60, 1}};
    static const uint32_t leafs[] = {1, 4, 8, 16, 24, 32, 48};
    static const uint32_t chunks[] = {RTX_TRACE_RAY_CHUNK_AUTO, 64, 128, 256, 512};
    static const uint32_t depths[] = {4, 6, 8, 12, 16};
    rtx_trace_config best = start;
    float best_ms = INFINITY;
    int r = time_cfg(best, best_ms);
    auto consider = [&](const rtx_trace_config& c) {
        float ms;
        if (r == RTX_SUCCESS && (r = time_cfg(c, ms)) == RTX_SUCCESS && ms < kTuneGain * best_ms) {
            best_ms = ms;
            best = c;
        }
    };
    // knob groups in order; each keeps the best values found so far for the other knobs
    for (int pass = 0; pass < 2 && r == RTX_SUCCESS; ++pass) {
        for (auto& sh : shapes) {
            rtx_trace_config c = best;
            c.block_size = sh[0];
            c.blocks_per_cu = sh[1];
            consider(c);
        }
        for (auto& rf : refills) {
            rtx_trace_config c = best;
            c.refill_lanes = rf[0];
            c.refill_interval = rf[1];
            consider(c);
        }
        for (uint32_t v : leafs) {
            rtx_trace_config c = best;
            c.leaf_lanes = v;
            consider(c);
        }
        for (uint32_t v : chunks) {
            rtx_trace_config c = best;
            c.ray_chunk = v;
            consider(c);
        }
        for (uint32_t v : depths) {
            rtx_trace_config c = best;
            c.lds_stack_depth = v;
            consider(c);
        }
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (r != RTX_SUCCESS) {
        ctx->cfg = start;
        return r;
    }
    ctx->cfg = best;
    if (out_cfg) *out_cfg = best;
    if (out_ms) *out_ms = best_ms;
    return RTX_SUCCESS;
}

int rtx_autotune_trace(rtx_context ctx, uint64_t tlas_address, const rtx_ray* rays_device, uint32_t ray_count,
                       uint32_t mode, void* results_device, uint32_t reps, rtx_trace_config* out_cfg,
                       float* out_ms_per_launch) {
    if (!ctx) return RTX_ERROR_INVALID_ARGUMENT;
    if (ray_count == 0 || !rays_device || !results_device)
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "autotuning needs a non-empty ray buffer and a result buffer");
    rtx_hit* hits = mode == RTX_QUERY_CLOSEST_HIT ? (rtx_hit*)results_device : nullptr;
    uint32_t* occl = mode == RTX_QUERY_ANY_HIT ? (uint32_t*)results_device : nullptr;
    return autotune_core(
        ctx, "rtx_autotune_trace", reps,
        [&]() {
            return rtx_cmd_trace_ray_buffer(ctx, tlas_address, rays_device, ray_count, mode, 0u, 0xFFu, hits, occl);
        },
        out_cfg, out_ms_per_launch);
}

int rtx_autotune_render_batch(rtx_context ctx, uint64_t tlas_address, const rtx_sbt* sbt,
                              const rtx_render_view* views, uint32_t view_count, const rtx_render_params* params,
                              float* films_device, uint32_t reps, rtx_trace_config* out_cfg,
                              float* out_ms_per_launch) {
    if (!ctx) return RTX_ERROR_INVALID_ARGUMENT;
    if (!films_device || !params) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "autotuning needs films and params");
    rtx_render_params p = *params;
    p.flags &= ~RTX_RENDER_GRAPH_BIT;
    return autotune_core(
        ctx, "rtx_autotune_render_batch", reps,
        [&]() { return rtx_cmd_render_batch(ctx, tlas_address, sbt, views, view_count, &p, films_device, nullptr); },
        out_cfg, out_ms_per_launch);
}

int rtx_autotune_render(rtx_context ctx, uint64_t tlas_address, const rtx_sbt* sbt, const rtx_camera* camera,
                        const rtx_render_params* params, float* film_rgba32f_device, uint32_t reps,
                        rtx_trace_config* out_cfg, float* out_ms_per_launch) {
    if (!ctx) return RTX_ERROR_INVALID_ARGUMENT;
    if (!film_rgba32f_device) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "autotuning needs a film");
    return autotune_core(
        ctx, "rtx_autotune_render", reps,
        [&]() { return rtx_cmd_render(ctx, tlas_address, sbt, camera, params, film_rgba32f_device, nullptr); },
        out_cfg, out_ms_per_launch);
}

int rtx_debug_lbvh(rtx_context ctx, const float* boxes_device, uint32_t n, uint32_t morton_bits, void* codes_device,
                   uint32_t* order_device, int32_t* left_device, int32_t* right_device) {
    if (!ctx) return RTX_ERROR_INVALID_ARGUMENT;
    if (morton_bits != 30 && morton_bits != 63)
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "morton_bits must be 30 or 63");
    if (n > kMaxPrims) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "n too large");
    if (n > 0 && (!boxes_device || !codes_device || !order_device))
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "NULL buffer");
    if (n > 1 && (!left_device || !right_device)) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "NULL buffer");
    hipError_t e = rtx::debug_lbvh(boxes_device, n, morton_bits == 63, codes_device, order_device, left_device,
                                   right_device, &ctx->state->error_flags, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "rtx_debug_lbvh");
    return rtx_sync(ctx);
}

int rtx_debug_treelet(rtx_context ctx, const float* boxes_device, uint32_t n, uint32_t morton_bits, uint32_t rounds,
                      uint32_t* order_device, int32_t* left_device, int32_t* right_device, float* cost_device,
                      uint32_t* height_device) {
    if (!ctx) return RTX_ERROR_INVALID_ARGUMENT;
    if (morton_bits != 30 && morton_bits != 63)
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "morton_bits must be 30 or 63");
    if (rounds > 3) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "rounds must be 0..3");
    if (n < 2 || n > kMaxPrims) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "n must be in [2, 2^30]");
    if (!boxes_device || !order_device || !left_device || !right_device || !cost_device || !height_device)
        return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "NULL buffer");
    hipError_t e = rtx::debug_treelet(boxes_device, n, morton_bits == 63, rounds, order_device, left_device,
                                      right_device, cost_device, height_device, &ctx->state->error_flags, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "rtx_debug_treelet");
    return rtx_sync(ctx);
}

int rtx_debug_sort(rtx_context ctx, void* keys_device, uint32_t* values_device, uint32_t n, uint32_t key_bytes) {
    if (!ctx) return RTX_ERROR_INVALID_ARGUMENT;
    if (key_bytes != 4 && key_bytes != 8) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "key_bytes must be 4 or 8");
    if (n > kMaxPrims) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "n too large");
    if (n > 0 && (!keys_device || !values_device)) return fail(ctx, RTX_ERROR_INVALID_ARGUMENT, "NULL buffer");
    hipError_t e = rtx::debug_sort(keys_device, values_device, n, (int)key_bytes, &ctx->state->error_flags, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "rtx_debug_sort");
    return rtx_sync(ctx);
}

// Name -> entry point, as vkGetDeviceProcAddr resolves the reference's KHR functions (main.cpp:25-33, 137-145).
void* rtx_get_proc_addr(const char* name) {
    struct Proc {
        const char* name;
        void* fn;
    };
#define RTX_PROC(f) {#f, (void*)&f}
    static const Proc kProcs[] = {
        RTX_PROC(rtx_create_context),
        RTX_PROC(rtx_destroy_context),
        RTX_PROC(rtx_last_error_string),
        RTX_PROC(rtx_sync),
        RTX_PROC(rtx_get_stream),
        RTX_PROC(rtx_get_blas_build_sizes),
        RTX_PROC(rtx_get_tlas_build_sizes),
        RTX_PROC(rtx_cmd_build_blas),
        RTX_PROC(rtx_get_blas_batch_build_sizes),
        RTX_PROC(rtx_cmd_build_blas_batch),
        RTX_PROC(rtx_cmd_update_blas),
        RTX_PROC(rtx_cmd_build_tlas),
        RTX_PROC(rtx_get_as_info),
        RTX_PROC(rtx_get_as_compacted_size),
        RTX_PROC(rtx_cmd_copy_as),
        RTX_PROC(rtx_obj_load),
        RTX_PROC(rtx_obj_parse),
        RTX_PROC(rtx_obj_free),
        RTX_PROC(rtx_cmd_trace_rays),
        RTX_PROC(rtx_cmd_trace_ray_buffer),
        RTX_PROC(rtx_cmd_trace_ray_buffer_noalloc),
        RTX_PROC(rtx_set_trace_stats),
        RTX_PROC(rtx_get_trace_stats),
        RTX_PROC(rtx_set_trace_config),
        RTX_PROC(rtx_get_trace_config),
        RTX_PROC(rtx_autotune_trace),
        RTX_PROC(rtx_cmd_film_to_rgba8),
        RTX_PROC(rtx_cmd_render),
        RTX_PROC(rtx_cmd_render_batch),
        RTX_PROC(rtx_autotune_render_batch),
        RTX_PROC(rtx_autotune_render),
        RTX_PROC(rtx_get_render_ray_count),
        RTX_PROC(rtx_get_render_form),
        RTX_PROC(rtx_nrc_create),
        RTX_PROC(rtx_nrc_destroy),
        RTX_PROC(rtx_nrc_cmd_infer),
        RTX_PROC(rtx_nrc_cmd_train),
        RTX_PROC(rtx_nrc_get_params),
        RTX_PROC(rtx_nrc_set_params),
        RTX_PROC(rtx_nrc_get_status),
        RTX_PROC(rtx_cmd_render_nrc),
        RTX_PROC(rtx_debug_lbvh),
        RTX_PROC(rtx_debug_treelet),
        RTX_PROC(rtx_debug_sort),
        RTX_PROC(rtx_nrc_debug_gradient),
        RTX_PROC(rtx_get_proc_addr),
        // reference calls that one librtx function replaces (with librtx's signature)
        {"vkCmdTraceRaysKHR", (void*)&rtx_cmd_trace_rays},
        {"vkCmdCopyAccelerationStructureKHR", (void*)&rtx_cmd_copy_as},
        {"vkQueueWaitIdle", (void*)&rtx_sync},
    };
#undef RTX_PROC
    if (!name) return nullptr;
    for (const Proc& p : kProcs)
        if (std::strcmp(p.name, name) == 0) return p.fn;
    return nullptr;
}

}  // extern "C"
