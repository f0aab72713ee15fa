// rtx_internal.h — memory layout of acceleration structures and shared device helpers (gfx950 only).
//
// Acceleration-structure storage (caller-owned, 256-B aligned):
//
//   [0, 256)        AsHeader
//   [256, ...)      64-B slots.  Slot 0 is the root node.  A slot holds either
//                     Node8    — compressed 8-wide node, children quantised to 8 bits per plane
//                     TriSlot  — one triangle or AABB primitive (BLAS leaf)
//                     InstSlot — one instance reference (TLAS leaf)
//                   The valid children of a node occupy consecutive slots starting at Node8::base, in the
//                   node's child-position order (position k is the k-th set bit of vmask).
//   [inst_table_offset, ...)   TLAS only: InstanceInfo[instance_count] (128 B each)
//
// Node8 (64 B):  quantisation origin p is stored as 3 x uint16 relative to the AS root box
//   p_a = fma(pq_a, qscale_a, root_lo_a)
// and each child plane as an 8-bit multiple of a power of two 2^(e_a - 127):
//   plane = fma(q, 2^(e_a-127), p_a)
// The builder chooses pq, e and q so that the decoded planes (evaluated with exactly these FP32 operations)
// enclose the child boxes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rtx.h"

namespace rtx {

constexpr uint32_t kMagicBlas = 0x53414C42u;  // "BLAS"
constexpr uint32_t kMagicTlas = 0x53414C54u;  // "TLAS"
constexpr uint32_t kVersion = 1u;
constexpr uint32_t kHeaderBytes = 256u;
constexpr uint32_t kSlotBytes = 64u;
constexpr uint32_t kNoHit = 0xFFFFFFFFu;

struct AsHeader {
    uint32_t magic;
    uint32_t version;
    uint32_t type;  // 0 BLAS, 1 TLAS
    uint32_t flags;
    uint32_t prim_count;
    uint32_t geometry_count;
    uint32_t node_count;
    uint32_t slot_count;
    uint32_t max_depth;        // BVH8 levels (root = 1)
    uint32_t max_total_depth;  // TLAS: own depth + deepest referenced BLAS
    uint32_t build_error;      // non-zero when a build step hit a device-side limit
    uint32_t content;          // kContent* bits: what rays through this structure can meet
    float lo[4];      // root box
    float hi[4];
    float qscale[4];  // step of the 16-bit node origins
    uint64_t slots_offset;
    uint64_t inst_table_offset;
    uint64_t capacity_slots;
    double sah_cost;
    uint64_t update_offset;  // BLAS built to allow updates: offset of the kept hierarchy (0 otherwise)
    uint8_t reserved[256 - 136];
};
static_assert(sizeof(AsHeader) == 256, "AsHeader must be 256 B");
// AsHeader::content.  A BLAS: the union over its geometries with primitives; a TLAS: the union over the BLASes its
// valid instances reference, plus kContentNonOpaque when an instance has FORCE_NO_OPAQUE.  Traversal of a
// structure with neither bit runs the opaque-triangle variant.
constexpr uint32_t kContentNonOpaque = RTX_AS_CONTENT_NON_OPAQUE_BIT;  // geometry without RTX_GEOMETRY_OPAQUE_BIT
constexpr uint32_t kContentAabb = RTX_AS_CONTENT_AABBS_BIT;            // AABB primitives (intersection programs)

struct Node8 {
    uint16_t pq[3];
    uint8_t e[3];
    uint8_t imask;  // child position k is an internal node
    uint8_t vmask;  // child position k exists
    uint8_t pad;
    uint32_t base;  // slot of the first valid child
    uint8_t qlo[3][8];
    uint8_t qhi[3][8];
};
static_assert(sizeof(Node8) == 64, "Node8 must be 64 B");

struct TriSlot {
    float v[9];  // v0, v1, v2 after the geometry transform
    uint32_t prim;
    uint32_t geom;
    uint32_t bits;  // kTriValid | kTriNonOpaque | kTriAabb
    uint32_t pad[4];
};
constexpr uint32_t kTriValid = 1u;
constexpr uint32_t kTriNonOpaque = 2u;  // its geometry lacks RTX_GEOMETRY_OPAQUE_BIT
constexpr uint32_t kTriAabb = 4u;       // an AABB primitive: v[0..2] = min, v[3..5] = max, v[6..8] = 0
static_assert(sizeof(TriSlot) == 64, "TriSlot must be 64 B");

struct InstSlot {
    float inv[12];  // world -> object (FP32 copy)
    uint64_t blas;
    uint32_t instance_id;
    uint32_t mask;  // instance mask (bits 0-7) | instance flags << 8
};
static_assert(sizeof(InstSlot) == 64, "InstSlot must be 64 B");

struct InstanceInfo {
    double inv[12];   // world -> object, computed in FP64 from the FP32 object -> world transform
    uint32_t custom;  // instanceCustomIndex (24 bit)
    uint32_t mask;
    uint32_t sbt_offset;
    uint32_t flags;
    uint64_t blas;
    uint64_t pad;
};
static_assert(sizeof(InstanceInfo) == 128, "InstanceInfo must be 128 B");

// Per-geometry build descriptor (device copy of rtx_blas_geometry with resolved pointers).
struct GeomDesc {
    const uint8_t* vtx;
    const uint32_t* idx;    // already offset by primitive_offset
    const float* xform;     // already offset by transform_offset, or nullptr
    uint64_t stride;
    uint32_t prim_count;
    uint32_t first_vertex;
    uint32_t prim_base;     // prefix sum of primitive counts
    uint32_t max_vertex;
    uint32_t geom_local;    // geometry index within its BLAS (gl_GeometryIndexEXT)
    uint32_t blas;          // BLAS of a batched build (0 otherwise)
    uint32_t flags;         // RTX_GEOMETRY_*_BIT
    uint32_t type;          // RTX_GEOMETRY_TYPE_*
};

// Triangle record in build scratch (48 B): transformed vertices + IDs.
struct BuildTri {
    float v[9];
    uint32_t prim;
    uint32_t geom;
    uint32_t valid;
};
static_assert(sizeof(BuildTri) == 48, "BuildTri must be 48 B");

// Closest hit as the path tracer consumes it (48 B): world-space unit geometric normal and SBT record index.
struct PathHit {
    float t, u, v;
    uint32_t inst, prim, geom;
    float n[3];
    uint32_t rec;  // 0xFFFFFFFF on a miss
    uint32_t pad[2];
};
static_assert(sizeof(PathHit) == 48, "PathHit must be 48 B");

// Context-owned device state.
struct DeviceState {
    uint32_t ray_counter;
    uint32_t error_flags;  // bit 0: stack overflow, bit 1: sort look-back timeout, bit 2: bad AS
    uint32_t pad[2];
    unsigned long long stat_nodes;
    unsigned long long stat_tris;
    unsigned long long stat_insts;
    unsigned long long stat_rays;
    unsigned long long pt_rays;  // rays traced by the path tracer (extension + shadow)
};

constexpr uint32_t kErrStackOverflow = 1u;
constexpr uint32_t kErrSortTimeout = 2u;
constexpr uint32_t kErrBadAs = 4u;
constexpr uint32_t kErrBuildDepth = 8u;

// ---- small device helpers ---------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t float_to_ordered(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ordered_to_float(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

__host__ __device__ inline uint32_t spread10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__host__ __device__ inline uint64_t spread21(uint64_t v) {
    v &= 0x1FFFFFull;
    v = (v | (v << 32)) & 0x1F00000000FFFFull;
    v = (v | (v << 16)) & 0x1F0000FF0000FFull;
    v = (v | (v << 8)) & 0x100F00F00F00F00Full;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

__host__ __device__ inline uint64_t align_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

}  // namespace rtx
