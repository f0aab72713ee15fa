// pathtrace.h — wavefront path tracer launch interface.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nrc.h"
#include "rtx_internal.h"
#include "trace.h"

namespace rtx {

enum PtMode : uint32_t { kPtPlain = 0, kPtCache = 2 };

// One view of a batched render (many uneven scenes in one wavefront): its camera, film, light range, the box its
// rays are clipped to (the scenes of a batch sit in disjoint boxes of one TLAS) and the offset added to the pass's
// sample index for its paths (a plain render runs several samples in one launch as views of one camera, each with
// its own offset and an unbounded box).
struct BatchView {
    float cam[3];
    float ax, ay;
    uint32_t width, height;
    uint32_t pix_base;     // first path slot / film pixel of the view
    uint32_t light_first;  // light triangles [light_first, light_first + light_count)
    uint32_t light_count;
    uint32_t cdf_first;    // its light CDF: cdf[cdf_first .. cdf_first + light_count]
    float lo[3], hi[3];
    uint32_t sample_offset;
};

struct PtArgs {
    const void* tlas;
    uint32_t width, height, npix;
    uint32_t npaths;       // path slots of this pass: npix render paths + ntrain training paths
    uint32_t first_path;   // this wavefront traces paths [first_path, npaths); queue slots count from first_path
    uint32_t ntrain;       // training paths (slots npix .. npix + ntrain - 1; cache mode)
    uint32_t mode;         // PtMode
    uint32_t cache_depth;  // vertex at which cache-mode render paths query the cache
    uint32_t train_end;    // last vertex of training paths
    uint32_t tail;         // 1: training paths query the cache at train_end (self-training)
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
    uint32_t* counts;  // per depth d: [3d] extension rays, [3d+1] shadow rays; [2] cache queries of the pass
    uint32_t* train_counts;  // the same for the training wavefront of an NRC frame
    float* cdf;
    rtx_nrc_query* cq;  // cache queries
    uint32_t* cq_pid;
    float4* rec;        // training vertex records, [(d-1) * ntrain + (pid - npix)] x 5 float4
    // batched render (views != nullptr)
    const BatchView* views;
    uint32_t nviews;
    uint32_t* path_view;       // view of each path slot
    const uint32_t* sample_dev;  // when set, added to every pass's sample index (graph replays)
    uint32_t content;            // AsHeader::content of the TLAS (traversal variant, whole-path availability)
    // Calls recorded before the TLAS's content can be read back: dual = 1 records a plain sample both as the
    // whole-path launch (guarded to run only for opaque-triangle content) and as the wavefront (guard = 2: its
    // camera stage runs only for other content, so its later stages find empty queues), and the device runs one.
    uint32_t dual;
    uint32_t guard;
    // Plain renders of several samples: each whole-path launch runs up to ms_max samples as views of the camera,
    // their per-path state in ms_buf (pt_sample_views_bytes), so that a launch's tail of long paths is paid once per
    // ms_max samples instead of once per sample.
    void* ms_buf;
    uint32_t ms_max;
};

// Whether the render's samples run as whole-path launches (trace modes 6 / 7, and 5 for full-length training paths):
// for a TLAS whose content the default traversal variant serves, unless RTX_PATH_WAVEFRONT=1 asks for the per-vertex
// wavefront (both give bit-identical films).
bool pt_whole_paths(const PtArgs& a);
// Bytes of PtArgs::ms_buf for a film of npix pixels and up to `samples` samples per launch.
uint64_t pt_sample_views_bytes(uint32_t npix, uint32_t samples);
// Whether a plain render records both forms (PtArgs::dual, unless RTX_PATH_WAVEFRONT=1 forces the wavefront).
bool pt_dual_forms(const PtArgs& a);

uint64_t pt_scratch_bytes(uint32_t npix, uint32_t max_depth, uint32_t nlights);
void pt_layout(PtArgs& a, void* scratch, uint32_t npix, uint32_t max_depth, uint32_t nlights);
hipError_t pt_render(PtArgs a, const TraceLaunch& launch, uint32_t spp, uint32_t sample_offset, float* film,
                     double* accum64, hipStream_t stream);
// batched render: a.views / a.nviews set, a.npix = total pixels, films concatenated in view order; no host
// synchronisation, so the whole call can be captured into a graph
hipError_t pt_render_batch(PtArgs a, const TraceLaunch& launch, uint32_t spp, uint32_t sample_offset, float* films,
                           double* accum64, hipStream_t stream);
uint64_t pt_train_bytes(uint32_t train_paths, uint32_t record_depths);

// Resources of the training wavefront of an NRC frame, which runs on its own stream beside the render paths: the
// persistent traversal of each wavefront needs its own ray counter (DeviceState) and spill stack.
struct SideChain {
    hipStream_t stream;
    hipEvent_t fork, join;
    DeviceState* state;
    uint2* gstack;
    TraceLaunch launch;  // launch shape of the training wavefront's traversals
};
// vertices per training path that become samples
uint32_t nrc_record_depths(uint32_t max_depth, uint32_t cache_depth, uint32_t train_suffix);
// one frame: spp cache-terminated passes, the first carrying train_paths training paths (self-training when
// train_suffix > 0) as a second wavefront on side.stream, then the Adam steps over the device-resident sample count
// (no host synchronisation: the frame can be captured into a graph)
hipError_t pt_render_nrc(PtArgs a, const TraceLaunch& launch, const SideChain& side, NrcState& nrc,
                         uint32_t train_paths, uint32_t train_suffix, void* train_buf, uint32_t spp,
                         uint32_t sample_offset, float* film, double* accum64, hipStream_t stream);

}  // namespace rtx
