// util.h — film conversion and acceleration-structure copies.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rtx_internal.h"

namespace rtx {

constexpr uint32_t kFilmSrgb = 1u;
constexpr uint32_t kFilmSwapRb = 2u;
constexpr uint32_t kFilmOpaque = 4u;

hipError_t film_to_rgba8(const float* film, uint64_t npix, uint32_t flags, uint8_t* out, hipStream_t stream);
// CLONE (compact = false: same capacity and layout) or COMPACT copy of the structure at src into dst
hipError_t copy_as(const void* src, void* dst, uint64_t dst_size, bool compact, uint32_t* err_flags,
                   hipStream_t stream);
// bytes of the compacted copy of a structure with this header
uint64_t compacted_bytes(const AsHeader& h);

}  // namespace rtx
