// trace.h — launch interface of the persistent traversal kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rtx_internal.h"

namespace rtx {

struct PtArgs;

struct TraceRequest {
    const void* tlas;
    uint32_t mode;  // 0 closest (ray buffer), 1 any (ray buffer), 2 raygen into a film, 3 path hits,
                    // 4 path hits for `rays` + occlusion for `rays2` in one launch,
                    // 5 whole training paths of an NRC frame (n paths, pt / pt_sample),
                    // 6 whole pixel paths of one sample of a plain render,
                    // 7 whole pixel paths of an NRC frame up to their cache query
    uint32_t n;
    const uint32_t* n_ptr;  // device ray count (overrides n when set)
    const rtx_ray* rays2;   // mode 4: the any-hit rays and their device count
    const uint32_t* n2_ptr;
    PathHit* path_hits;
    uint32_t cull_mask;
    uint32_t ray_flags;  // RTX_RAY_FLAG_* (0 for the integrator's queries)
    const rtx_ray* rays;
    rtx_hit* hits;
    uint32_t* occluded;
    // raygen
    float cam[3];
    float ax, ay;
    uint32_t width, height;
    float tmin, tmax;
    const rtx_hit_record* records;
    uint32_t nrec, sbt_offset, sbt_stride;
    float miss[3];
    float* film;
    // runtime
    DeviceState* state;
    uint2* gstack;
    bool stats;
    const PtArgs* pt;  // mode 5: the path tracer's arguments
    uint32_t pt_sample;
    uint32_t content;  // AsHeader::content of the TLAS (kContent* bits): selects the traversal variant
    // Device-side variant choice, for calls recorded into a graph before the TLAS's content can be read back:
    // 1 = the launch runs only if the TLAS, read on the device, needs no FLAGS traversal for ray_flags; 2 = only if
    // it does; 3 (ray buffers and raygen) = both variants are launched, each guarded, so the device runs one.
    uint32_t guard = 0;
};

struct TraceLaunch {
    uint32_t grid;
    uint32_t block_size;
    uint32_t blocks_per_cu;
    uint32_t lds_depth;  // stack entries per lane in LDS
    uint32_t gdepth;     // further entries per lane in the global spill area
    uint32_t ray_chunk;        // ray ids per queue fetch (< 64: automatic)
    uint32_t refill_lanes;     // refill a wave once fewer lanes than this are busy (1..64)
    uint32_t refill_interval;  // traversal steps between refill checks (1..64)
    uint32_t leaf_lanes;       // pending leaf groups in a wave before the leaf step runs (1..64)
    uint32_t waves;            // VGPR-bounded kernel variant: 0 (unbounded), 4 (128 VGPRs) or 5 (96 VGPRs) waves per SIMD
};

hipError_t trace_rays_launch(const TraceRequest& req, const TraceLaunch& launch, hipStream_t stream);
// Whether rays with these flags through a TLAS of this content need the FLAGS traversal variant (opacity, any-hit and
// intersection programs, ray-flag culls); modes 5-7 (whole paths in one launch) exist only in the default variant, so
// they serve a TLAS for which trace_needs_flags(0, content) is false.
bool trace_needs_flags(uint32_t ray_flags, uint32_t content);

// Ray flags that need the FLAGS traversal variant, and the variant choice itself (host and device).
constexpr uint32_t kFlagsVariantRayFlags =
    RTX_RAY_FLAG_TERMINATE_ON_FIRST_HIT | RTX_RAY_FLAG_SKIP_CLOSEST_HIT_SHADER |
    RTX_RAY_FLAG_CULL_BACK_FACING_TRIANGLES | RTX_RAY_FLAG_CULL_FRONT_FACING_TRIANGLES | RTX_RAY_FLAG_CULL_OPAQUE |
    RTX_RAY_FLAG_CULL_NO_OPAQUE | RTX_RAY_FLAG_NO_OPAQUE | RTX_RAY_FLAG_SKIP_TRIANGLES | RTX_RAY_FLAG_SKIP_AABBS;
__host__ __device__ __forceinline__ bool needs_flags_variant(uint32_t ray_flags, uint32_t content) {
    return (ray_flags & kFlagsVariantRayFlags) || (content & kContentAabb) ||
           (!(ray_flags & RTX_RAY_FLAG_OPAQUE) && (content & kContentNonOpaque));
}
// The choice on the device, from the TLAS header: whether a launch guarded with `guard` (1 / 2 as
// TraceRequest::guard) is the one to run.
__device__ __forceinline__ bool guard_passes(const void* tlas, uint32_t ray_flags, uint32_t guard) {
    return needs_flags_variant(ray_flags, ((const AsHeader*)tlas)->content) == (guard == 2u);
}

}  // namespace rtx
