// noalloc.h — ray-buffer queries whose working memory is a file-scope device array (no allocation per call).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rtx_internal.h"
#include "trace.h"

namespace rtx {

// Size of each device's scratch array (device state + traversal-stack spill).
uint64_t noalloc_scratch_bytes();
// Once per device, with that device current: look up the device's copy of the array and clear its device state.
hipError_t noalloc_setup(int device, hipStream_t stream);
// The device state inside the array (ray counter, error flags), or nullptr before noalloc_setup.
DeviceState* noalloc_state(int device);
// Launch q with its state and stack spill in the array; hipErrorOutOfMemory (and *need_bytes) when the launch
// shape's spill area does not fit.
hipError_t noalloc_trace(int device, TraceRequest q, const TraceLaunch& launch, hipStream_t stream,
                         uint64_t* need_bytes);

}  // namespace rtx
