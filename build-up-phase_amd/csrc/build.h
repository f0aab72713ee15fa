// build.h — acceleration-structure builders (BLAS over triangles, TLAS over instances).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rtx_internal.h"

namespace rtx {

// Host -> device copy of a small host array (build descriptors, batch views) enqueued on `stream`.  Outside a
// stream capture it is hipMemcpyAsync.  While the stream is being captured into a graph, the bytes travel as the
// arguments of copy launches, which the graph keeps, so a replay does not read host memory that the recording
// call has since released; up to kArgUploadMax bytes (multiples of 4 B), hipErrorNotSupported above.
constexpr uint64_t kArgUploadMax = 64 * 1024;
hipError_t upload_host_array(void* dst_device, const void* src_host, uint64_t bytes, hipStream_t stream);

uint64_t blas_as_bytes(uint32_t prim_count, bool allow_update = false);
uint64_t tlas_as_bytes(uint32_t instance_count);
uint64_t blas_scratch_bytes(uint32_t prim_count, bool morton63, uint32_t geometry_count);
uint64_t tlas_scratch_bytes(uint32_t instance_count, bool morton63);

// Enqueue a BLAS build of prim_count triangles described by geometry_count host descriptors.
hipError_t build_blas(const GeomDesc* geoms_host, uint32_t geometry_count, uint32_t prim_count, uint32_t flags,
                      void* dst, void* scratch, uint32_t* err_flags, hipStream_t stream);
// Enqueue the builds of nbatch BLASes in one pass over all their triangles.  geoms_host: all geometries in BLAS
// order (GeomDesc.blas / geom_local / prim_base over the concatenation); binfo_host[b] = (triangles, geometries);
// dst_host[b]: the BLAS storage (>= blas_as_bytes of its triangle count).  30-bit Morton codes only.
uint64_t blas_batch_scratch_bytes(uint32_t prim_count, uint32_t geometry_count, uint32_t nbatch);
hipError_t build_blas_batch(const GeomDesc* geoms_host, uint32_t geometry_count, const uint2* binfo_host,
                            AsHeader* const* dst_host, uint32_t nbatch, uint32_t prim_count, uint32_t flags,
                            void* scratch, uint32_t* err_flags, hipStream_t stream);
// UPDATE mode for a BLAS built with RTX_BUILD_ALLOW_UPDATE_BIT (same geometry and triangle counts; scratch of
// blas_scratch_bytes).
hipError_t update_blas(const GeomDesc* geoms_host, uint32_t geometry_count, uint32_t prim_count, uint32_t flags,
                       void* dst, void* scratch, uint32_t* err_flags, hipStream_t stream);
// Enqueue a TLAS build over instance_count device-resident rtx_instance records.
hipError_t build_tlas(const rtx_instance* instances_device, uint32_t instance_count, uint32_t flags, void* dst,
                      void* scratch, uint32_t* err_flags, hipStream_t stream);

// Verification hooks: Morton -> sort -> Karras on caller boxes; the bare radix sort.  Both block.
hipError_t debug_lbvh(const float* boxes, uint32_t n, bool morton63, void* codes_out, uint32_t* order_out,
                      int32_t* left_out, int32_t* right_out, uint32_t* err_flags, hipStream_t stream);
hipError_t debug_treelet(const float* boxes, uint32_t n, bool morton63, uint32_t rounds, uint32_t* order_out,
                         int32_t* left_out, int32_t* right_out, float* cost_out, uint32_t* hgt_out,
                         uint32_t* err_flags, hipStream_t s);
hipError_t debug_sort(void* keys, uint32_t* vals, uint32_t n, int key_bytes, uint32_t* err_flags, hipStream_t stream);

}  // namespace rtx
