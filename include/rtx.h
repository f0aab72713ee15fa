/*
 * rtx.h — C ABI of the MI355X (gfx950) ray-tracing core.
 *
 * The entry points follow the Vulkan KHR ray-tracing flow of
 * vulkan-raytraced-triangle/main.cpp one for one:
 *
 *   vkGetAccelerationStructureBuildSizesKHR (main.cpp:757, 893) -> rtx_get_blas_build_sizes / rtx_get_tlas_build_sizes
 *   vkCreateAccelerationStructureKHR + ...DeviceAddressKHR (main.cpp:776-785, 912-919)
 *                                                             -> folded into rtx_cmd_build_* (caller owns the storage,
 *                                                                the call returns the 64-bit AS address)
 *   vkCmdBuildAccelerationStructuresKHR (main.cpp:808, 932)   -> rtx_cmd_build_blas / rtx_cmd_build_tlas
 *   vkCreateRayTracingPipelinesKHR + SBT (main.cpp:1093-1172, 1264-1320)
 *                                                             -> compiled-in programs + rtx_hit_record table (rtx_sbt)
 *   vkCmdTraceRaysKHR (main.cpp:1349-1355)                    -> rtx_cmd_trace_rays
 *   vkQueueWaitIdle (main.cpp:820, 942)                       -> rtx_sync
 *
 * plus ray-buffer queries (closest hit / any hit) and the path-tracing integrator.
 *
 * Conventions
 *  - Every pointer named *_device is a HIP device pointer; everything else is host memory.
 *  - Return values follow VkResult: 0 = success, negative = error.  No C++ exception crosses the ABI.
 *  - An rtx_context is externally synchronised (like a VkQueue).  All rtx_cmd_* calls are ordered on the
 *    context's HIP stream and return before the GPU work finishes; rtx_sync waits for it.
 *  - Recording (the command-buffer analogue, main.cpp:789-812 and 1340-1356): rtx_cmd_build_blas,
 *    rtx_cmd_build_blas_batch, rtx_cmd_build_tlas, rtx_cmd_copy_as, rtx_cmd_trace_rays, rtx_cmd_trace_ray_buffer(
 *    _noalloc), rtx_cmd_render(_batch, _nrc) and rtx_cmd_film_to_rgba8 may be recorded by a HIP stream capture of
 *    the context's stream (hipStreamBeginCapture ... hipStreamEndCapture, or torch.cuda.graph on that stream);
 *    replaying the graph redoes the recorded work on the device data as it is at replay time.  Host arrays passed
 *    to a recorded call (build descriptors, batch views: up to 64 KiB each) are copied into the graph.  A
 *    recorded call cannot allocate: make each call once outside the capture first, so the context's internal
 *    buffers have their size (a call that would grow one fails with RTX_ERROR_OUT_OF_DEVICE_MEMORY).  A TLAS
 *    built and synchronised before the capture is traced with the traversal its content allows; for one built
 *    inside it, traces and plain / batched renders record both traversal variants (and both integrator forms),
 *    each guarded by the content the device reads at replay, and NRC frames take the general traversal and the
 *    wavefront; all give the same results.  RTX_RENDER_GRAPH_BIT is ignored inside a capture.
 *    rtx_cmd_update_blas and the rtx_get_* queries synchronise and cannot be recorded.
 *  - Only gfx950 code objects are built.
 */
#ifndef RTX_H_
#define RTX_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RTX_API_VERSION 1u

/* ---- result codes (VkResult values where one exists) ---------------------------------------- */
typedef enum rtx_result {
    RTX_SUCCESS = 0,
    RTX_ERROR_OUT_OF_HOST_MEMORY = -1,      /* VK_ERROR_OUT_OF_HOST_MEMORY */
    RTX_ERROR_OUT_OF_DEVICE_MEMORY = -2,    /* VK_ERROR_OUT_OF_DEVICE_MEMORY: storage/scratch below the size query */
    RTX_ERROR_INITIALIZATION_FAILED = -3,   /* VK_ERROR_INITIALIZATION_FAILED */
    RTX_ERROR_DEVICE_LOST = -4,             /* VK_ERROR_DEVICE_LOST: a HIP call failed */
    RTX_ERROR_INVALID_ARGUMENT = -1000,     /* bad pointer, size, count or flag */
    RTX_ERROR_STACK_OVERFLOW = -1001,       /* traversal stack exceeded the built depth bound */
    RTX_ERROR_INVALID_AS = -1002,           /* address is not an acceleration structure of this library */
    RTX_ERROR_FORMAT = -1003                /* malformed input file (rtx_obj_*) */
} rtx_result;

typedef struct rtx_context_t* rtx_context;

/* Create a context on HIP device `device`.  `hip_stream` is the hipStream_t to record on; NULL means the
 * device's default (null) stream, as in HIP itself. */
int rtx_create_context(int device, void* hip_stream, rtx_context* out_ctx);
int rtx_destroy_context(rtx_context ctx);
/* Last error message of this context (never NULL; "" when none). */
const char* rtx_last_error_string(rtx_context ctx);
/* vkQueueWaitIdle (main.cpp:820, 942): wait for all work recorded on the context's stream, then report
 * device-side error flags raised by that work (e.g. RTX_ERROR_STACK_OVERFLOW). */
int rtx_sync(rtx_context ctx);
/* The HIP stream the context records on. */
void* rtx_get_stream(rtx_context ctx);

/* ---- acceleration-structure inputs ------------------------------------------------------------ */

/* VkGeometryFlagBitsKHR / VkBuildAccelerationStructureFlagBitsKHR subset */
#define RTX_GEOMETRY_OPAQUE_BIT 0x1u                       /* main.cpp:741; without it the geometry's candidate
                                                              hits run the record's any-hit program */
#define RTX_BUILD_ALLOW_UPDATE_BIT 0x1u                    /* BLAS keeps its hierarchy for rtx_cmd_update_blas */
#define RTX_BUILD_ALLOW_COMPACTION_BIT 0x2u                /* accepted; every structure can be compacted */
#define RTX_BUILD_PREFER_FAST_TRACE_BIT 0x4u               /* main.cpp:751, 887 */
#define RTX_BUILD_PREFER_FAST_BUILD_BIT 0x8u               /* no treelet refinement unless
                                                              RTX_BUILD_TREELET_ROUNDS asks for it */
#define RTX_BUILD_MORTON63_BIT 0x10000u                    /* 63-bit Morton keys (default: 30-bit) */
#define RTX_BUILD_NO_TREELET_BIT 0x20000u                  /* skip SAH treelet refinement */
/* Treelet refinement rounds in bits 18-19 (0 = default: 2, or none with PREFER_FAST_BUILD; 1..3 rounds); round r
 * restructures nodes with at least 7 * 2^r primitives. */
#define RTX_BUILD_TREELET_ROUNDS_SHIFT 18u
#define RTX_BUILD_TREELET_ROUNDS(r) (((uint32_t)(r) & 3u) << RTX_BUILD_TREELET_ROUNDS_SHIFT)
/* Children per BVH node in bits 20-22: 0 = 8 (default), 2..7 caps the SAH-optimal wide collapse at that many (the
 * same 64-B node, fewer children: shallower per-node work, deeper trees).  A TLAS and its BLASes may differ. */
#define RTX_BUILD_MAX_WIDTH_SHIFT 20u
#define RTX_BUILD_MAX_WIDTH(w) (((uint32_t)(w) & 7u) << RTX_BUILD_MAX_WIDTH_SHIFT)

#define RTX_INDEX_TYPE_UINT32 1u                           /* VK_INDEX_TYPE_UINT32, main.cpp:736 */

/* One triangle geometry: VkAccelerationStructureGeometryTrianglesDataKHR (main.cpp:726-742) merged with its
 * VkAccelerationStructureBuildRangeInfoKHR (main.cpp:791-804).  Vertex format is R32G32B32_SFLOAT
 * (main.cpp:732).  Triangle p of the geometry uses vertices
 *   index_data[primitive_offset/4 + 3p + k] + first_vertex,  k = 0,1,2
 * each fetched at vertex_data + vertex_index * vertex_stride bytes.  When transform_data_device is not NULL
 * the row-major 3x4 matrix at transform_data_device + transform_offset bytes is applied to the vertices at
 * build time (VkTransformMatrixKHR, main.cpp:684-695, 795-804). */
typedef struct rtx_blas_geometry {
    const float* vertex_data_device;
    uint64_t vertex_stride;           /* bytes, >= 12, multiple of 4 */
    uint32_t max_vertex;              /* highest vertex index the indices may reference */
    uint32_t index_type;              /* RTX_INDEX_TYPE_UINT32 */
    const uint32_t* index_data_device;
    const float* transform_data_device; /* nullable */
    uint32_t primitive_count;
    uint32_t primitive_offset;        /* bytes into index data, multiple of 4 */
    uint32_t first_vertex;
    uint32_t transform_offset;        /* bytes into transform data, multiple of 16 */
    uint32_t flags;                   /* RTX_GEOMETRY_*_BIT */
    uint32_t geometry_type;           /* RTX_GEOMETRY_TYPE_* (VkGeometryTypeKHR) */
} rtx_blas_geometry;

/* VkGeometryTypeKHR.  An AABB geometry (VkAccelerationStructureGeometryAabbsDataKHR) reads primitive p's box
 * (VkAabbPositionsKHR: min x y z, max x y z as 6 floats) at vertex_data_device + primitive_offset + p *
 * vertex_stride bytes (stride a multiple of 8, >= 24; index and transform fields unused); a box with a NaN or
 * min > max is inactive.  Its candidate hits come from the intersection program of the hit record
 * (RTX_HIT_GROUP), RTX_INTERSECTION_BOX when the call has no hit-record table. */
#define RTX_GEOMETRY_TYPE_TRIANGLES 0u
#define RTX_GEOMETRY_TYPE_AABBS 1u

/* VkAccelerationStructureInstanceKHR, byte-for-byte (64 B; main.cpp:848-858).  transform is the
 * object-to-world row-major 3x4; blas_address is a value returned by rtx_cmd_build_blas. */
typedef struct rtx_instance {
    float transform[3][4];
    uint32_t custom_index_and_mask;          /* instanceCustomIndex:24 | mask:8 (mask in the high byte) */
    uint32_t sbt_offset_and_flags;           /* instanceShaderBindingTableRecordOffset:24 | flags:8 */
    uint64_t blas_address;
} rtx_instance;

#define RTX_INSTANCE_TRIANGLE_FACING_CULL_DISABLE_BIT 0x1u /* VK_GEOMETRY_INSTANCE_..., main.cpp:852 */
#define RTX_INSTANCE_TRIANGLE_FLIP_FACING_BIT 0x2u
#define RTX_INSTANCE_FORCE_OPAQUE_BIT 0x4u                 /* the instance's triangles count as opaque */
#define RTX_INSTANCE_FORCE_NO_OPAQUE_BIT 0x8u              /* ... as non-opaque */

/* VkAccelerationStructureBuildSizesInfoKHR */
typedef struct rtx_as_build_sizes {
    uint64_t acceleration_structure_size;
    uint64_t build_scratch_size;
    uint64_t update_scratch_size;
} rtx_as_build_sizes;

/* vkGetAccelerationStructureBuildSizesKHR for a BLAS (main.cpp:757-762).  Pure query. */
int rtx_get_blas_build_sizes(rtx_context ctx, const rtx_blas_geometry* geometries, uint32_t geometry_count,
                             uint32_t build_flags, rtx_as_build_sizes* out_sizes);
/* vkGetAccelerationStructureBuildSizesKHR for a TLAS (main.cpp:893-898).  Pure query. */
int rtx_get_tlas_build_sizes(rtx_context ctx, uint32_t instance_count, uint32_t build_flags,
                             rtx_as_build_sizes* out_sizes);

/* vkCreateAccelerationStructureKHR + vkCmdBuildAccelerationStructuresKHR + vkGetAccelerationStructureDevice-
 * AddressKHR for a BLAS (main.cpp:776-812).  dst_device (>= acceleration_structure_size bytes, 256-B aligned)
 * and scratch_device (>= build_scratch_size bytes, 256-B aligned) are caller-owned.  Inputs and scratch may
 * be released once the stream has passed this call (main.cpp:823-830).  *out_address receives the BLAS
 * address to put into rtx_instance.blas_address. */
int rtx_cmd_build_blas(rtx_context ctx, const rtx_blas_geometry* geometries, uint32_t geometry_count,
                       uint32_t build_flags, void* dst_device, uint64_t dst_size, void* scratch_device,
                       uint64_t scratch_size, uint64_t* out_address);

/* One BLAS of a batched build: its geometries (host array) and its caller-owned storage. */
typedef struct rtx_blas_build_info {
    const rtx_blas_geometry* geometries;
    uint32_t geometry_count;
    uint32_t reserved;
    void* dst_device;                 /* >= its acceleration_structure_size, 256-B aligned */
    uint64_t dst_size;
} rtx_blas_build_info;

/* vkCmdBuildAccelerationStructuresKHR with infoCount > 1 (main.cpp:808 passes one): every build step runs once over
 * the triangles of all `count` BLASes, so small BLASes no longer leave the GPU idle.  Sizes: as_sizes[b] (may be
 * NULL) per BLAS, *scratch_size for the whole batch.  30-bit Morton codes only (RTX_BUILD_MORTON63_BIT is
 * rejected).  out_addresses[b] receives BLAS b's address.  The structures are the same as single builds would
 * give, except that equal Morton codes are ordered by the batch-wide triangle index. */
int rtx_get_blas_batch_build_sizes(rtx_context ctx, const rtx_blas_build_info* infos, uint32_t count,
                                   uint32_t build_flags, uint64_t* as_sizes, uint64_t* scratch_size);
int rtx_cmd_build_blas_batch(rtx_context ctx, const rtx_blas_build_info* infos, uint32_t count, uint32_t build_flags,
                             void* scratch_device, uint64_t scratch_size, uint64_t* out_addresses);

/* vkCmdBuildAccelerationStructuresKHR in UPDATE mode (not used by the reference) for a BLAS built with
 * RTX_BUILD_ALLOW_UPDATE_BIT (whose size query then includes the kept hierarchy, ~20 B per triangle): the same
 * geometries with the same triangle counts and new vertex positions or transforms.  The binary hierarchy and
 * triangle order of the build are kept (no re-sort, no treelet restructuring), boxes are refitted and the
 * 8-wide nodes re-emitted in place, so traversal stays exact while quality decays with large deformations.
 * scratch >= update_scratch_size of the size query.  Reads the structure's header (one synchronisation). */
int rtx_cmd_update_blas(rtx_context ctx, const rtx_blas_geometry* geometries, uint32_t geometry_count,
                        uint64_t blas_address, void* scratch_device, uint64_t scratch_size);

/* Same for a TLAS over instance_count rtx_instance records in device memory (main.cpp:912-945). */
int rtx_cmd_build_tlas(rtx_context ctx, const rtx_instance* instances_device, uint32_t instance_count,
                       uint32_t build_flags, void* dst_device, uint64_t dst_size, void* scratch_device,
                       uint64_t scratch_size, uint64_t* out_address);

/* Statistics written by a build (read from device memory by the host; blocks until the build is done). */
typedef struct rtx_as_info {
    uint32_t type;                 /* 0 = BLAS, 1 = TLAS */
    uint32_t primitive_count;      /* triangles or instances */
    uint32_t node_count;           /* 64-B BVH8 nodes */
    uint32_t slot_count;           /* 64-B slots (nodes + leaves) */
    uint32_t max_depth;            /* BVH8 levels */
    uint32_t geometry_count;
    float bounds[6];               /* lo xyz, hi xyz */
    double sah_cost;               /* SAH cost of the binary hierarchy before collapse: sum of internal-node
                                      surface areas / root surface area (0 for fewer than 2 primitives) */
    uint32_t content_flags;        /* RTX_AS_CONTENT_* bits */
    uint32_t reserved;
} rtx_as_info;
/* What rays through a structure can meet, recorded at build time.  A BLAS: its geometries with primitives; a TLAS:
 * the BLASes its instances reference, and NON_OPAQUE as well when an instance has RTX_INSTANCE_FORCE_NO_OPAQUE_BIT.
 * Traces and renders of a TLAS with neither bit run the opaque-triangle traversal (and whole-path renders); the
 * others resolve opacity, any-hit and intersection programs.  The choice is made per call from the traced TLAS. */
#define RTX_AS_CONTENT_NON_OPAQUE_BIT 0x1u  /* geometry built without RTX_GEOMETRY_OPAQUE_BIT */
#define RTX_AS_CONTENT_AABBS_BIT 0x2u       /* AABB geometry */
int rtx_get_as_info(rtx_context ctx, uint64_t as_address, rtx_as_info* out_info);

/* vkCmdWriteAccelerationStructuresPropertiesKHR(COMPACTED_SIZE) + vkCmdCopyAccelerationStructureKHR (not used by
 * the reference, which builds with PREFER_FAST_TRACE only, main.cpp:751).  The compacted size keeps the header
 * and the used 64-B slots (plus a TLAS's instance table); it blocks until the build is done.  A copy is enqueued
 * on the stream and reads the source header on the device.  A TLAS refers to BLASes by address: after
 * compacting a BLAS, build the TLAS again with the new address (as in Vulkan). */
#define RTX_COPY_MODE_CLONE 0u
#define RTX_COPY_MODE_COMPACT 1u
int rtx_get_as_compacted_size(rtx_context ctx, uint64_t as_address, uint64_t* out_bytes);
int rtx_cmd_copy_as(rtx_context ctx, uint64_t src_address, uint32_t mode, void* dst_device, uint64_t dst_size,
                    uint64_t* out_address);

/* ---- OBJ meshes (host only; the reference's next step: "load an .obj file -> build the acceleration structure",
 * vulkan-raytracing-basic/README.md:225-226) ----------------------------------------------------------------
 * Reads `v` and `f` statements (face corners v, v/vt, v//vn, v/vt/vn; 1-based or negative relative indices;
 * polygons fan-triangulated as (v0, vk, vk+1)); `o` / `g` start a new geometry.  All geometries share the vertex
 * array: geometry g is triangles [geometry_first_triangle[g], geometry_first_triangle[g + 1]), i.e. an
 * rtx_blas_geometry with primitive_offset = 12 * first and primitive_count = first[g + 1] - first[g].  Arrays are
 * host memory owned by the mesh until rtx_obj_free.  Malformed input returns RTX_ERROR_FORMAT. */
typedef struct rtx_obj_mesh {
    uint32_t vertex_count;
    uint32_t triangle_count;
    uint32_t geometry_count;
    uint32_t reserved;
    float* vertices;                     /* 3 floats per vertex */
    uint32_t* indices;                   /* 3 per triangle */
    uint32_t* geometry_first_triangle;   /* geometry_count + 1 entries */
} rtx_obj_mesh;
int rtx_obj_load(const char* path, rtx_obj_mesh* out_mesh);
int rtx_obj_parse(const char* text, size_t length, rtx_obj_mesh* out_mesh);
void rtx_obj_free(rtx_obj_mesh* mesh);

/* ---- shader binding table ----------------------------------------------------------------------- */

/* Compiled-in hit programs (the reference compiles GLSL at run time, shader_module.h:11-74).
 * rtx_hit_record.program = RTX_HIT_GROUP(closest-hit program, intersection program). */
#define RTX_PROGRAM_RECORD_COLOR 0u      /* hitValue = record color */
#define RTX_PROGRAM_BARYCENTRIC 1u       /* hitValue = (1-u-v, u, v) */
#define RTX_PROGRAM_MATCH_BARYCENTRIC 2u /* main.cpp:1080-1091: barycentric colour when gl_PrimitiveID,
                                            gl_InstanceID, gl_InstanceCustomIndexEXT and gl_GeometryIndexEXT
                                            all equal the record's match_* fields, else the record color */
#define RTX_PROGRAM_LAMBERTIAN 3u        /* path tracer: albedo = color, emission = emission */

/* Compiled-in intersection programs for AABB geometry (bits 8-15 of rtx_hit_record.program).  Each reports the
 * first intersection with t in [tmin, tmax] of the ray with a surface inside the primitive's box, and (u, v):
 *   BOX     the box's surface; (u, v) = the hit point's position across the hit face, along the two other axes in
 *           cyclic order (x -> y, z; y -> z, x; z -> x, y), each in [0, 1] from the box's min to max corner;
 *   SPHERE  the sphere centred in the box with radius half the smallest extent; (u, v) = ((n.x + 1) / 2,
 *           (n.y + 1) / 2) for the outward unit normal n in object space. */
#define RTX_INTERSECTION_BOX 0u
#define RTX_INTERSECTION_SPHERE 1u
#define RTX_HIT_GROUP(closest, intersection) ((uint32_t)(closest) | ((uint32_t)(intersection) << 8))

/* Compiled-in any-hit programs (rtx_hit_record.any_hit: program in bits 0-7, parameter in bits 8-31), run on the
 * non-opaque candidate hits of the record's geometry.  (u, v) are the candidate's hitAttributeEXT barycentrics,
 * computed in FP64 exactly as reported for a closest hit. */
#define RTX_ANY_HIT_ACCEPT 0u  /* accept every candidate (a hit group without an any-hit shader) */
#define RTX_ANY_HIT_CUTOUT 1u  /* ignoreIntersectionEXT() where floor(n u) + floor(n v) is odd, n = parameter
                                  (0 counts as 1): a checkerboard of holes over the triangle */
#define RTX_ANY_HIT_PROGRAM(p, param) ((uint32_t)(p) | ((uint32_t)(param) << 8))

/* One hit-group record (main.cpp:1307-1317: handle + float color[3]); 48 B. */
typedef struct rtx_hit_record {
    uint32_t program;
    float color[3];
    float emission[3];
    int32_t match_primitive;
    int32_t match_instance;
    int32_t match_custom_index;
    int32_t match_geometry;
    uint32_t any_hit; /* RTX_ANY_HIT_PROGRAM(program, parameter) */
} rtx_hit_record;

/* Hit-group table + miss program.  Record index for a hit (Vulkan spec 40.3.1, main.cpp:1252-1263):
 *   instance sbt offset + geometry index * sbt_record_stride + sbt_record_offset */
typedef struct rtx_sbt {
    const rtx_hit_record* hit_records_device;
    uint32_t hit_record_count;
    uint32_t sbt_record_offset;    /* traceRayEXT sbtRecordOffset (main.cpp:1049: 0) */
    uint32_t sbt_record_stride;    /* traceRayEXT sbtRecordStride (main.cpp:1049: 1) */
    uint32_t reserved;
    float miss_color[3];           /* miss program (main.cpp:1063-1066: (0, 0, 0.2)) */
    float miss_alpha_reserved;
} rtx_sbt;

/* Camera uniform block of the raygen shader (std140, 16 B; main.cpp:1003-1006, value {0,0,10,60} :1015). */
typedef struct rtx_camera {
    float position[3];
    float y_fov_degree;
} rtx_camera;

/* traceRayEXT arguments used by the raygen program (main.cpp:1047-1052). */
/* traceRayEXT ray flags (gl_RayFlags*EXT values; the reference passes gl_RayFlagsOpaqueEXT, main.cpp:1048).
 * Opacity of a candidate hit (Vulkan order): the geometry's RTX_GEOMETRY_OPAQUE_BIT, overridden by the instance's
 * FORCE_OPAQUE / FORCE_NO_OPAQUE, overridden by the ray's OPAQUE / NO_OPAQUE.  CULL_OPAQUE / CULL_NO_OPAQUE drop
 * candidates of that opacity.  A non-opaque candidate runs the any-hit program of its hit record (when the call
 * has a hit-record table), which accepts it or ignores it.  Facing: a triangle is front-facing when
 * v0, v1, v2 appear clockwise seen from the ray origin (dot(d, (v1 - v0) x (v2 - v0)) > 0 in object space);
 * RTX_INSTANCE_TRIANGLE_FLIP_FACING_BIT makes counter-clockwise front-facing and
 * RTX_INSTANCE_TRIANGLE_FACING_CULL_DISABLE_BIT exempts the instance from the CULL_*_FACING flags.
 * TERMINATE_ON_FIRST_HIT reports the first accepted hit, not necessarily the closest. */
#define RTX_RAY_FLAG_OPAQUE 0x1u                 /* gl_RayFlagsOpaqueEXT */
#define RTX_RAY_FLAG_NO_OPAQUE 0x2u
#define RTX_RAY_FLAG_TERMINATE_ON_FIRST_HIT 0x4u /* gl_RayFlagsTerminateOnFirstHitEXT */
#define RTX_RAY_FLAG_SKIP_CLOSEST_HIT_SHADER 0x8u
#define RTX_RAY_FLAG_CULL_BACK_FACING_TRIANGLES 0x10u
#define RTX_RAY_FLAG_CULL_FRONT_FACING_TRIANGLES 0x20u
#define RTX_RAY_FLAG_CULL_OPAQUE 0x40u
#define RTX_RAY_FLAG_CULL_NO_OPAQUE 0x80u
#define RTX_RAY_FLAG_SKIP_TRIANGLES 0x100u
#define RTX_RAY_FLAG_SKIP_AABBS 0x200u          /* gl_RayFlagsSkipAABBEXT */
typedef struct rtx_trace_params {
    uint32_t ray_flags;    /* RTX_RAY_FLAG_* */
    uint32_t cull_mask;    /* 8-bit, ANDed with the instance mask */
    float tmin;
    float tmax;
} rtx_trace_params;

/* vkCmdTraceRaysKHR(width, height, 1) with the reference raygen program (main.cpp:1033-1055):
 *   aspect_y = tan(radians(fov) / 2), aspect_x = aspect_y * width / height
 *   ndc = (launchID + 0.5) / launchSize * 2 - 1
 *   dir = (ndc.x * aspect_x, -ndc.y * aspect_y, -1)   (not normalised)
 * and stores vec4(hitValue, 0) into film_rgba32f_device[(y * width + x) * 4 + c].
 * params may be NULL: then tmin 0, tmax 100, opaque, cull mask 0xFF (main.cpp:1047-1052). */
int rtx_cmd_trace_rays(rtx_context ctx, uint64_t tlas_address, const rtx_sbt* sbt, const rtx_camera* camera,
                       const rtx_trace_params* params, uint32_t width, uint32_t height,
                       float* film_rgba32f_device);

/* ---- ray-buffer queries (ray-query form of traceRayEXT) ------------------------------------------- */

/* 32 B */
typedef struct rtx_ray {
    float origin[3];
    float tmin;
    float direction[3];
    float tmax;
} rtx_ray;

/* 24 B.  instance_id == 0xFFFFFFFF marks a miss.  (u, v) are the barycentrics of the Vulkan
 * hitAttributeEXT: P = (1-u-v) v0 + u v1 + v v2. */
typedef struct rtx_hit {
    float t;
    float u;
    float v;
    uint32_t primitive_id;    /* gl_PrimitiveID: index within the geometry */
    uint32_t geometry_index;  /* gl_GeometryIndexEXT */
    uint32_t instance_id;     /* gl_InstanceID: index in the TLAS instance array */
} rtx_hit;

#define RTX_QUERY_CLOSEST_HIT 0u
#define RTX_QUERY_ANY_HIT 1u

/* Trace ray_count rays.  mode RTX_QUERY_CLOSEST_HIT writes hits_device[i] (rtx_hit); RTX_QUERY_ANY_HIT writes
 * occluded_device[i] = 1 if any opaque triangle lies in [tmin, tmax], else 0.  The unused output may be NULL.
 * ray_flags: RTX_RAY_FLAG_* for every ray; cull_mask is ANDed with each instance mask (0xFF in the reference). */
int rtx_cmd_trace_ray_buffer(rtx_context ctx, uint64_t tlas_address, const rtx_ray* rays_device,
                             uint32_t ray_count, uint32_t mode, uint32_t ray_flags, uint32_t cull_mask,
                             rtx_hit* hits_device, uint32_t* occluded_device);

/* rtx_cmd_trace_ray_buffer that allocates nothing.  No counterpart in the reference: for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc), where
 * rtx_cmd_trace_ray_buffer allocates its traversal-stack spill area on first use of a launch shape.  The ray counter,
 * error flags and stack spill live in a 48 MiB array that the library keeps on each device until the process ends:
 * enough for the default launch shape (4 blocks of 256 per CU) on a TLAS whose depth with its deepest BLAS is at most
 * 16 BVH levels; a launch needing more returns RTX_ERROR_OUT_OF_DEVICE_MEMORY (with the byte count in
 * rtx_last_error_string).  Every call on one GPU shares that array: calls that overlap on one GPU (several streams
 * or contexts) race.  Results equal rtx_cmd_trace_ray_buffer's; traversal statistics are not collected; errors are
 * reported by rtx_sync.  Recorded into a graph, it needs a TLAS whose depth the host read back before the capture:
 * for a TLAS built in the same recording the stack of an unknown depth (76 entries) exceeds the array. */
int rtx_cmd_trace_ray_buffer_noalloc(rtx_context ctx, uint64_t tlas_address, const rtx_ray* rays_device,
                                     uint32_t ray_count, uint32_t mode, uint32_t ray_flags, uint32_t cull_mask,
                                     rtx_hit* hits_device, uint32_t* occluded_device);

/* Per-ray traversal counters of the last rtx_cmd_trace_ray_buffer call issued with stats enabled:
 * total 64-B node fetches, triangle tests and instance entries summed over all rays. */
typedef struct rtx_trace_stats {
    uint64_t rays;
    uint64_t node_fetches;
    uint64_t triangle_tests;
    uint64_t instance_entries;
} rtx_trace_stats;
int rtx_set_trace_stats(rtx_context ctx, int enable);
int rtx_get_trace_stats(rtx_context ctx, rtx_trace_stats* out_stats);

/* Traversal kernel launch shape and scheduling (the autotuner's knobs).  0 fields keep the current value. */
typedef struct rtx_trace_config {
    uint32_t block_size;      /* threads per block: 64, 128, 192 or 256 */
    uint32_t blocks_per_cu;   /* persistent blocks per compute unit (1..32) */
    uint32_t lds_stack_depth; /* stack entries held in LDS per lane before spilling to global memory; with the
                                 lane's world ray (24 B) at most 64 KiB of LDS per block: (8 depth + 24) block */
    uint32_t ray_chunk;       /* ray ids a wave takes from the global queue per fetch: a multiple of 64 up to
                                 4096, or RTX_TRACE_RAY_CHUNK_AUTO (64, or 256 when every wave gets >= 8 chunks) */
    uint32_t refill_lanes;    /* a wave fetches new rays once fewer than this many of its 64 lanes are busy */
    uint32_t refill_interval; /* traversal steps between two refill checks (1..64) */
    uint32_t leaf_lanes;      /* postponed leaves: a wave runs its leaf step (BLAS entry, triangle tests) once this
                                 many lanes hold a pending leaf group or no lane can traverse (1..64; 1 = at once) */
    uint32_t stack_limit;     /* traversal stack entries per lane, LDS and global spill together: 0 (default) sizes the
                                 stack from the traced TLAS's depth (2 entries per BVH level + 4, enough for every
                                 ray); with a smaller value a ray that needs more ends there (reporting the closest
                                 hit found so far, or a miss) and rtx_sync reports RTX_ERROR_STACK_OVERFLOW.  Not an autotuner knob; rtx_set_trace_config keeps it
                                 when the field is 0, so RTX_TRACE_STACK_LIMIT_DERIVED restores the default. */
} rtx_trace_config;
#define RTX_TRACE_STACK_LIMIT_DERIVED 0xFFFFFFFFu
#define RTX_TRACE_RAY_CHUNK_AUTO 1u
int rtx_set_trace_config(rtx_context ctx, const rtx_trace_config* cfg);
int rtx_get_trace_config(rtx_context ctx, rtx_trace_config* out_cfg);

/* Project autotuner for the traversal kernel (no counterpart in the reference, whose traversal is the driver's).
 * Times candidate configurations on the caller's ray buffer with HIP events (coordinate descent, two passes:
 * block size x blocks per CU, refill policy, postponed-leaf threshold, ray chunk, LDS stack depth; `reps` timed
 * launches each; a change
 * must gain more than 1.5%), keeps the fastest on the context and returns it with its time per launch.  Tune on
 * the workload itself: a few milliseconds of rays per launch at least.  results_device receives the query results
 * (rtx_hit[] for closest hit, uint32[] for any hit).  Blocks until done. */
int rtx_autotune_trace(rtx_context ctx, uint64_t tlas_address, const rtx_ray* rays_device, uint32_t ray_count,
                       uint32_t query_mode, void* results_device, uint32_t reps, rtx_trace_config* out_cfg,
                       float* out_ms_per_launch);

/* ---- film conversion -------------------------------------------------------------------------------------
 * RGBA32F film -> 8-bit RGBA bytes: q = round(clamp(x, 0, 1) * 255) per channel (UNORM), optionally after the sRGB
 * transfer function, with B and R swapped, or with alpha 255.  RTX_FILM_REFERENCE_BYTES gives the bytes the
 * reference shows: its R8G8B8A8_UNORM out image is written through an R/B-swizzled view (main.cpp:952-974) and
 * copied byte for byte to the B8G8R8A8 swapchain image (main.cpp:1357-1389). */
#define RTX_FILM_SRGB_BIT 0x1u
#define RTX_FILM_SWAP_RB_BIT 0x2u
#define RTX_FILM_OPAQUE_BIT 0x4u
#define RTX_FILM_REFERENCE_BYTES RTX_FILM_SWAP_RB_BIT
int rtx_cmd_film_to_rgba8(rtx_context ctx, const float* film_rgba32f_device, uint32_t width, uint32_t height,
                          uint32_t flags, uint8_t* out_rgba8_device);

/* ---- path tracing: wavefront integrator (an extension of the raygen / closest-hit flow) ----------------------
 *
 * Camera rays follow the reference raygen model (main.cpp:1033-1052) with the pixel position jittered inside the
 * pixel.  Surfaces are Lambertian (hit records with RTX_PROGRAM_LAMBERTIAN: albedo = color, emitted radiance =
 * emission on the side the geometric normal (v1-v0)x(v2-v0) points to).  Estimator per pixel sample:
 *   vertex 0 adds the emission it sees; every vertex d < max_depth adds next-event estimation toward one point
 *   chosen on the light triangles (triangle by area, point uniform); paths continue by cosine-weighted sampling,
 *   with Russian roulette p = min(0.95, max(throughput)) when continuing from vertex d >= rr_depth.
 * Random numbers are counter-based: pcg4d(x, y, sample, seed + 2d) and pcg4d(x, y, sample, seed + 2d + 1) per
 * vertex, pcg4d(x, y, sample, seed ^ 0xA511E9B3) for the pixel jitter. */
typedef struct rtx_render_params {
    uint32_t width;
    uint32_t height;
    uint32_t spp;                 /* samples per pixel rendered by this call */
    uint32_t max_depth;           /* path vertices, camera hit included (8) */
    uint32_t rr_depth;            /* Russian roulette when continuing from vertex index >= rr_depth (3) */
    uint32_t seed;                /* PCG seed (1234) */
    uint32_t sample_offset;       /* index of this call's first sample (progressive rendering) */
    uint32_t flags;               /* RTX_RENDER_*_BIT (rtx_cmd_render_batch, rtx_cmd_render_nrc), else 0 */
    const float* light_triangles_device; /* 12 floats per light: v0, v1, v2 (world), emitted radiance rgb */
    uint32_t light_triangle_count;
    float ray_offset;             /* secondary-ray origin offset along the normal, scene units (1e-4) */
} rtx_render_params;

/* Render params->spp samples per pixel.  film_rgba32f_device receives the mean radiance of these samples
 * (alpha 1).  accum_rgb_device (nullable) receives += the per-pixel radiance sums in FP64 (3 doubles per pixel),
 * for progressive FP64-accumulated references. */
int rtx_cmd_render(rtx_context ctx, uint64_t tlas_address, const rtx_sbt* sbt, const rtx_camera* camera,
                   const rtx_render_params* params, float* film_rgba32f_device, double* accum_rgb_device);

/* Rays traced by rtx_cmd_render since the context was created (camera, extension and shadow rays). */
/* ---- batched path tracing (many uneven scenes in one wavefront; no counterpart in the reference) --------------
 *
 * Many views of one TLAS rendered by one wavefront: every stage (camera rays, extension traversal, shading, shadow
 * traversal, gathering) is one launch over the paths of all views, whatever their resolutions.  Scenes of
 * different content share the TLAS as a scene atlas: each scene's instances sit in their own box, and every ray
 * of a view is clipped to that view's box, so it only ever meets its own scene. */
typedef struct rtx_render_view {
    rtx_camera camera;            /* world space, raygen model of main.cpp:1033-1052 */
    uint32_t width, height;
    uint32_t light_first;         /* the view's lights: params->light_triangles_device[light_first ..] */
    uint32_t light_count;
    float clip_min[3];            /* the view's scene box */
    float clip_max[3];
} rtx_render_view;

/* params: as rtx_render_params of rtx_cmd_render with width / height unused.  films: the views' RGBA32F films
 * concatenated in view order (sum of width x height pixels); accum (optional): FP64 rgb, same order.
 * With RTX_RENDER_GRAPH_BIT the call's launch sequence is recorded into a hipGraph the first time and replayed by
 * later calls with identical arguments apart from sample_offset (which the graph reads from device memory); the
 * graph is launched on the context stream.  A call with other arguments, a TLAS of other content (re)built at the
 * same address, or another traversal configuration records a new graph.  rtx_cmd_render_nrc takes the bit too. */
#define RTX_RENDER_GRAPH_BIT 0x1u
int rtx_cmd_render_batch(rtx_context ctx, uint64_t tlas_address, const rtx_sbt* sbt, const rtx_render_view* views,
                         uint32_t view_count, const rtx_render_params* params, float* films_device,
                         double* accum_rgb_device);
/* rtx_autotune_render for batched renders (times rtx_cmd_render_batch calls without the graph). */
int rtx_autotune_render_batch(rtx_context ctx, uint64_t tlas_address, const rtx_sbt* sbt,
                              const rtx_render_view* views, uint32_t view_count, const rtx_render_params* params,
                              float* films_device, uint32_t reps, rtx_trace_config* out_cfg,
                              float* out_ms_per_launch);

/* rtx_autotune_trace for the path tracer: times `reps` rtx_cmd_render calls with `params` per candidate
 * configuration (film written, no FP64 accumulation), keeps the fastest.  Blocks until done.  Tune with the
 * sample count of the renders to come, or with a few samples where that is many: a render of several samples
 * traces up to 16 of them in one launch, whose best shape can differ from that of a one-sample launch. */
int rtx_autotune_render(rtx_context ctx, uint64_t tlas_address, const rtx_sbt* sbt, const rtx_camera* camera,
                        const rtx_render_params* params, float* film_rgba32f_device, uint32_t reps,
                        rtx_trace_config* out_cfg, float* out_ms_per_launch);

int rtx_get_render_ray_count(rtx_context ctx, uint64_t* out_rays);

/* Form of the integrator the last rtx_cmd_render / _batch / _nrc call of the context enqueued (diagnostic; both forms
 * give bit-identical films): WHOLE_PATH = each sample's paths traced, shaded and continued in one traversal launch;
 * WAVEFRONT = per-vertex launches (shade, fused shadow + extension trace, gather), used for TLASes whose content
 * needs opacity or intersection programs, or when the environment variable RTX_PATH_WAVEFRONT=1 asks for it;
 * DEVICE_CHOSEN = a plain or batched render recorded by a stream capture before the TLAS's content could be read
 * back (e.g. a TLAS built in the same recording): both forms are recorded, each guarded by the content bits the
 * device reads from the TLAS header at replay, so that one of them runs (NRC frames take the wavefront there). */
#define RTX_RENDER_FORM_NONE 0u
#define RTX_RENDER_FORM_WHOLE_PATH 1u
#define RTX_RENDER_FORM_WAVEFRONT 2u
#define RTX_RENDER_FORM_DEVICE_CHOSEN 3u
int rtx_get_render_form(rtx_context ctx, uint32_t* out_form);

/* ---- neural radiance cache (Muller et al. 2021) --------------------------------------------------------------
 *
 * MLP: 96 inputs (81 used) -> 5 hidden layers of 64 (ReLU) -> 16 outputs (RGB in the first 3), BF16 weights and
 * activations on MFMA with FP32 accumulation, FP32 master weights, Adam.  Input encoding of a query: position
 * normalised to the config box, 12 frequencies per axis sin(2^k pi p), cos(2^k pi p) (72 features), normal (3),
 * direction (3), albedo (3).  Loss: relative L2, sum over RGB of (y - t)^2 / (y^2 + 0.01), mean over the batch.
 * Parameter vector (RTX_NRC_PARAM_COUNT floats): W1[64][96] b1[64] W2[64][64] b2 ... W5 b5 W6[16][64] b6[16]. */
#define RTX_NRC_PARAM_COUNT 23888u

typedef struct rtx_nrc_t* rtx_nrc;

typedef struct rtx_nrc_config {
    uint32_t seed;            /* weight initialisation seed (He-uniform weights, zero biases) */
    uint32_t batch_size;      /* samples per Adam step, multiple of 64 */
    float learning_rate;      /* 1e-3 */
    float beta1, beta2, epsilon;
    float scene_min[3];       /* position normalisation box */
    float scene_max[3];
    uint32_t cache_vertex;    /* rtx_cmd_render_nrc: path vertex (camera hit = 0) replaced by the cache query;
                                 0 selects 1, i.e. paths scatter at the primary hit and end at their 2nd bounce */
    uint32_t train_suffix;    /* training paths: 0 = full length (unbiased targets); k > 0 = self-training: they
                                 stop at vertex cache_vertex + k and take the cache's prediction there as the rest
                                 of their radiance */
} rtx_nrc_config;

/* 48 B: a path vertex as the cache sees it; direction points from the vertex toward the previous vertex */
typedef struct rtx_nrc_query {
    float position[3];
    float normal[3];
    float direction[3];
    float albedo[3];
} rtx_nrc_query;

/* 64 B: a training sample, query + target radiance */
typedef struct rtx_nrc_sample {
    rtx_nrc_query query;
    float target[3];
    float reserved;
} rtx_nrc_sample;

/* A cache belongs to the context it was created on.  Destroying that context first leaves the cache usable only
 * for rtx_nrc_destroy (every other call then returns RTX_ERROR_INVALID_ARGUMENT). */
int rtx_nrc_create(rtx_context ctx, const rtx_nrc_config* cfg, rtx_nrc* out_nrc);
int rtx_nrc_destroy(rtx_nrc nrc);
/* radiance predictions (3 floats per query, clamped at 0) */
int rtx_nrc_cmd_infer(rtx_nrc nrc, const rtx_nrc_query* queries_device, uint32_t count, float* out_rgb_device);
/* ceil(count / batch_size) Adam steps over consecutive batches of samples_device */
int rtx_nrc_cmd_train(rtx_nrc nrc, const rtx_nrc_sample* samples_device, uint32_t count);
/* FP32 master parameters (checkpoint / restore); host buffers of RTX_NRC_PARAM_COUNT floats */
int rtx_nrc_get_params(rtx_nrc nrc, float* params_host);
int rtx_nrc_set_params(rtx_nrc nrc, const float* params_host);
/* Adam steps taken so far and the mean loss of the last step */
int rtx_nrc_get_status(rtx_nrc nrc, uint32_t* steps, float* last_loss);

/* One frame of the cache-accelerated integrator (no host synchronisation: the Adam steps read the frame's sample count
 * from device memory; with RTX_RENDER_GRAPH_BIT in params->flags the frame is recorded into a hipGraph once and
 * replayed, as for rtx_cmd_render_batch).  Every pixel traces params->spp paths that stop at vertex
 * cfg->cache_vertex (instead of scattering there) and add throughput x the cache's prediction.  The first of these
 * passes also carries train_paths training paths from random pixels (sample index params->sample_offset) through
 * the same queues: each gives one training sample per vertex after the first (target: the path's radiance from
 * that vertex on), either at full length or, with cfg->train_suffix > 0, ending in a cache query of their own
 * (self-training; those tail queries are answered by the cache as it was before this frame).  Order within the
 * frame: the training paths' samples take one Adam step per batch_size samples (interleaved batches), and then the
 * pixel paths' cache queries of every pass are answered by the updated cache.  film / accum as in rtx_cmd_render.
 * params->spp == 0 makes a frame that only trains: its training paths and Adam steps exactly as in a rendered frame
 * (the cache ends bit-identical), no pixel paths; film and accum are not touched and may be NULL (train_paths must
 * then be > 0).  It is for the first frames of a run, whose images an accumulating caller would discard. */
int rtx_cmd_render_nrc(rtx_context ctx, uint64_t tlas_address, const rtx_sbt* sbt, const rtx_camera* camera,
                       const rtx_render_params* params, rtx_nrc nrc, uint32_t train_paths,
                       float* film_rgba32f_device, double* accum_rgb_device);

/* ---- verification hooks (no counterpart in the reference: they expose intermediate LBVH build steps so that
 * parity tests can compare them with the CPU oracle) ---------------------------------------------------------- */

/* Run the builder's Morton / sort / hierarchy steps on n primitive boxes (device, 6 floats each: lo xyz, hi xyz)
 * and copy out (device buffers): sorted codes (uint32 for morton_bits 30, uint64 for 63), the sorted primitive
 * order (uint32), and for the n-1 internal nodes of the binary radix tree the left/right child codes
 * (c >= 0 internal node, c < 0 leaf at sorted position -1-c).  Blocks until done; allocates its own scratch. */
int rtx_debug_lbvh(rtx_context ctx, const float* boxes_device, uint32_t n, uint32_t morton_bits, void* codes_device,
                   uint32_t* order_device, int32_t* left_device, int32_t* right_device);

/* rtx_debug_lbvh followed by the build's refit and SAH treelet refinement (rounds 0..3; 0 = refit only), on
 * n >= 2 boxes: copies out the sorted primitive order, the refined left/right child codes and, per internal node,
 * the subtree cost (float: sum of internal-node surface areas) and height (uint32).  Blocks until done. */
int rtx_debug_treelet(rtx_context ctx, const float* boxes_device, uint32_t n, uint32_t morton_bits, uint32_t rounds,
                      uint32_t* order_device, int32_t* left_device, int32_t* right_device, float* cost_device,
                      uint32_t* height_device);

/* Sort n (key, value) pairs with the builder's radix sort: keys uint32 (key_bytes 4) or uint64 (8); values are
 * replaced by the stable sorted permutation.  Blocks until done; allocates its own scratch. */
int rtx_debug_sort(rtx_context ctx, void* keys_device, uint32_t* values_device, uint32_t n, uint32_t key_bytes);

/* Gradient of the batch loss with respect to the FP32 parameters (RTX_NRC_PARAM_COUNT floats, device) and the
 * loss itself (1 float, device), without updating the cache.  No counterpart in the reference: verification of
 * the backward kernels. */
int rtx_nrc_debug_gradient(rtx_nrc nrc, const rtx_nrc_sample* samples_device, uint32_t count, float* grad_device,
                           float* loss_device);

/* ---- entry-point lookup ---------------------------------------------------------------------------------------- */

/* Address of the entry point `name`, or NULL: every rtx_* function of this header by its own name (as
 * vkGetDeviceProcAddr resolves the reference's KHR entry points, main.cpp:25-33, 137-145), and the reference calls
 * that a single librtx function replaces - "vkCmdTraceRaysKHR" -> rtx_cmd_trace_rays,
 * "vkCmdCopyAccelerationStructureKHR" -> rtx_cmd_copy_as, "vkQueueWaitIdle" -> rtx_sync - with librtx's
 * signature, not Vulkan's.  Calls that map to several librtx functions (the build-size query and the build, which
 * librtx splits by BLAS / TLAS) resolve to NULL under their Vulkan names. */
void* rtx_get_proc_addr(const char* name);

#ifdef __cplusplus
}
#endif
#endif /* RTX_H_ */
