// util.hip — film conversion and acceleration-structure copies (the steps next to the traced film and the built
// structures in the reference's flow).
//
// Film: the reference stores vec4(hitValue, 0.0) (main.cpp:1054) into an R8G8B8A8_UNORM storage image through a
// view whose R and B are swizzled (main.cpp:952-974) and copies its bytes to the B8G8R8A8 swapchain image
// (vkCmdCopyImage, main.cpp:1357-1389).  rtx_cmd_film_to_rgba8 produces those bytes from an RGBA32F film:
// UNORM quantisation round(clamp(x, 0, 1) * 255), optional B/R swap, optional sRGB encoding, alpha from the film
// or opaque.  One thread per pixel: 16 B read, 4 B written (HBM-bound).
//
// Copies: vkCmdCopyAccelerationStructureKHR in CLONE and COMPACT modes.  The copy kernel reads the source header
// on the device (no host round trip): a compacted structure keeps the header and the used slots and moves a
// TLAS's instance table right after them.
#include <cmath>

#include "rtx_internal.h"
#include "util.h"

namespace rtx {

namespace {

__device__ __forceinline__ uint32_t unorm8(float x) {
    if (!(x > 0.0f)) return 0u;  // negatives and NaN
    if (x >= 1.0f) return 255u;
    return (uint32_t)(x * 255.0f + 0.5f);
}

__device__ __forceinline__ float srgb_encode(float x) {
    if (!(x > 0.0f)) return 0.0f;
    if (x >= 1.0f) return 1.0f;
    return x <= 0.0031308f ? 12.92f * x : 1.055f * powf(x, 1.0f / 2.4f) - 0.055f;
}

__global__ __launch_bounds__(256) void film_to_rgba8_kernel(const float4* __restrict__ film, uint64_t npix,
                                                            uint32_t flags, uint32_t* __restrict__ out) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= npix) return;
    float4 c = film[i];
    if (flags & kFilmSrgb) {
        c.x = srgb_encode(c.x);
        c.y = srgb_encode(c.y);
        c.z = srgb_encode(c.z);
    }
    uint32_t r = unorm8(c.x), g = unorm8(c.y), b = unorm8(c.z);
    const uint32_t a = (flags & kFilmOpaque) ? 255u : unorm8(c.w);
    if (flags & kFilmSwapRb) {
        const uint32_t t = r;
        r = b;
        b = t;
    }
    out[i] = r | (g << 8) | (b << 16) | (a << 24);
}

// bytes of the source structure that a copy writes, and where a TLAS's instance table goes
__device__ __forceinline__ void copy_extent(const AsHeader* h, bool compact, uint64_t& slot_bytes,
                                            uint64_t& table_src, uint64_t& table_dst, uint64_t& table_bytes) {
    const uint64_t nslots = compact ? h->slot_count : h->capacity_slots;
    slot_bytes = kHeaderBytes + nslots * kSlotBytes;
    table_bytes = h->type == 1 ? (uint64_t)h->prim_count * sizeof(InstanceInfo) : 0;
    table_src = h->inst_table_offset;
    table_dst = compact && table_bytes ? align_up(slot_bytes, 256) : h->inst_table_offset;
}

__global__ __launch_bounds__(256) void copy_as_kernel(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst,
                                                      uint64_t dst_size, int compact, uint32_t* err) {
    const AsHeader* h = (const AsHeader*)src;
    if (h->magic != kMagicBlas && h->magic != kMagicTlas) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(err, kErrBadAs);
        return;
    }
    uint64_t slot_bytes, table_src, table_dst, table_bytes;
    copy_extent(h, compact != 0, slot_bytes, table_src, table_dst, table_bytes);
    if ((table_bytes ? table_dst + table_bytes : slot_bytes) > dst_size) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(err, kErrBadAs);
        return;
    }
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    for (uint64_t w = t0; w < slot_bytes / 16; w += stride) ((uint4*)dst)[w] = ((const uint4*)src)[w];
    for (uint64_t w = t0; w < table_bytes / 16; w += stride)
        ((uint4*)(dst + table_dst))[w] = ((const uint4*)(src + table_src))[w];
}

// After the copy: the new capacity and table offset (a separate launch, so no thread of the copy can overwrite it)
__global__ void copy_as_patch_kernel(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, uint64_t dst_size,
                                     int compact) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const AsHeader* h = (const AsHeader*)src;
    if (h->magic != kMagicBlas && h->magic != kMagicTlas) return;
    uint64_t slot_bytes, table_src, table_dst, table_bytes;
    copy_extent(h, compact != 0, slot_bytes, table_src, table_dst, table_bytes);
    if ((table_bytes ? table_dst + table_bytes : slot_bytes) > dst_size) return;
    AsHeader* d = (AsHeader*)dst;
    if (compact) d->capacity_slots = h->slot_count;
    d->inst_table_offset = table_bytes ? table_dst : h->inst_table_offset;
}

}  // namespace

hipError_t film_to_rgba8(const float* film, uint64_t npix, uint32_t flags, uint8_t* out, hipStream_t s) {
    if (npix == 0) return hipSuccess;
    film_to_rgba8_kernel<<<(unsigned)((npix + 255) / 256), 256, 0, s>>>((const float4*)film, npix, flags,
                                                                          (uint32_t*)out);
    return hipGetLastError();
}

hipError_t copy_as(const void* src, void* dst, uint64_t dst_size, bool compact, uint32_t* err, hipStream_t s) {
    copy_as_kernel<<<1024, 256, 0, s>>>((const uint8_t*)src, (uint8_t*)dst, dst_size, compact ? 1 : 0, err);
    copy_as_patch_kernel<<<1, 64, 0, s>>>((const uint8_t*)src, (uint8_t*)dst, dst_size, compact ? 1 : 0);
    return hipGetLastError();
}

uint64_t compacted_bytes(const AsHeader& h) {
    const uint64_t slot_bytes = kHeaderBytes + (uint64_t)h.slot_count * kSlotBytes;
    if (h.type != 1) return align_up(slot_bytes, 256);
    return align_up(align_up(slot_bytes, 256) + (uint64_t)h.prim_count * sizeof(InstanceInfo), 256);
}

}  // namespace rtx
