// pathtrace.h — wavefront path tracer launch interface.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rtx_internal.h"
#include "trace.h"

namespace rtx {

struct PtArgs {
    const void* tlas;
    uint32_t width, height, npix;
    uint32_t max_depth, rr_depth, seed;
    float cam[3];
    float ax, ay;
    float ray_offset;
    const rtx_hit_record* records;
    uint32_t nrec, sbt_offset, sbt_stride;
    const float* lights;
    uint32_t nlights;
    DeviceState* state;
    uint2* gstack;
    // scratch (pt_layout)
    rtx_ray* ext_rays[2];
    uint32_t* ext_pid[2];
    PathHit* hits;
    rtx_ray* sh_rays;
    uint32_t* sh_pid;
    float4* sh_contrib;
    uint32_t* occl;
    float4* thr;
    float4* rad;
    float4* accum;
    uint32_t* counts;  // per depth d: [3d] extension rays, [3d+1] shadow rays
    float* cdf;
};

uint64_t pt_scratch_bytes(uint32_t npix, uint32_t max_depth, uint32_t nlights);
void pt_layout(PtArgs& a, void* scratch, uint32_t npix, uint32_t max_depth, uint32_t nlights);
hipError_t pt_render(PtArgs a, const TraceLaunch& launch, uint32_t spp, uint32_t sample_offset, float* film,
                     double* accum64, hipStream_t stream);

}  // namespace rtx
