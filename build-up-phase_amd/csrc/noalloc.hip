// noalloc.hip — rtx_cmd_trace_ray_buffer_noalloc: ray-buffer queries that allocate nothing.
//
// rtx_cmd_trace_ray_buffer allocates its traversal-stack spill area on the first call of a launch shape (and again
// when a deeper structure or a larger shape needs more).  Callers that cannot allocate - code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc - use this variant: its ray counter, error flags and
// stack spill live in `scratch`, a file-scope device array of kScratchBytes on every GPU the library runs on (held
// until the process ends).  Host code never uses the array's name as an address: rtx_create_context looks up each
// device's copy once with hipGetSymbolAddress (that device current) and the launches take that pointer.  Every call
// on one GPU shares its array, which keeps the last call's contents; the ray counter is cleared by a kernel before
// each launch and the error flags by rtx_sync, so calls that overlap on one GPU (two streams, two contexts) race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "noalloc.h"
#include "rtx_internal.h"
#include "trace.h"

namespace rtx {

namespace {

// 48 MiB: a DeviceState and the stack spill of 262,144 lanes (4 blocks of 256 on 256 CUs) x 24 entries of 8 B,
// i.e. TLAS + BLAS depth up to 16 BVH8 levels with 12 stack entries per lane in LDS.
constexpr uint64_t kScratchBytes = 48ull << 20;
constexpr uint64_t kStateBytes = 256;
static_assert(sizeof(DeviceState) <= kStateBytes, "device state fits its slot");

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint8_t* g_ptr[kMaxDevices] = {};  // each device's copy of `scratch`

__global__ void clear_state_kernel(uint32_t* w, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) w[i] = 0u;
}

}  // namespace

uint64_t noalloc_scratch_bytes() { return kScratchBytes; }

hipError_t noalloc_setup(int device, hipStream_t stream) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_ptr[device]) return hipSuccess;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    // the device state starts at zero (ray counter, error flags, statistics)
    clear_state_kernel<<<1, 64, 0, stream>>>((uint32_t*)p, (uint32_t)(kStateBytes / 4));
    if ((e = hipGetLastError()) || (e = hipStreamSynchronize(stream))) return e;
    g_ptr[device] = (uint8_t*)p;
    return hipSuccess;
}

DeviceState* noalloc_state(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_mu);
    return (DeviceState*)g_ptr[device];
}

hipError_t noalloc_trace(int device, TraceRequest q, const TraceLaunch& L, hipStream_t stream, uint64_t* need_bytes) {
    uint8_t* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (device >= 0 && device < kMaxDevices) base = g_ptr[device];
    }
    if (!base) return hipErrorNotInitialized;
    const uint64_t need = kStateBytes + (uint64_t)L.grid * L.block_size * L.gdepth * sizeof(uint2);
    *need_bytes = need;
    if (need > kScratchBytes) return hipErrorOutOfMemory;
    q.state = (DeviceState*)base;
    q.gstack = (uint2*)(base + kStateBytes);
    return trace_rays_launch(q, L, stream);
}

}  // namespace rtx
