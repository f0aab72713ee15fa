// raytraced_triangle.cpp — the reference sample's flow (vulkan-raytraced-triangle/main.cpp) on librtx.
//
// Same scene, same calls in the same order, with the Vulkan KHR entry points replaced by the rtx C ABI:
//   createBLAS  (main.cpp:674-831)  -> rtx_get_blas_build_sizes + rtx_cmd_build_blas
//   createTLAS  (main.cpp:833-949)  -> rtx_get_tlas_build_sizes + rtx_cmd_build_tlas
//   createUniformBuffer / SBT       -> rtx_camera + rtx_hit_record table
//   render (main.cpp:1322-1423)     -> rtx_cmd_trace_rays into an RGBA32F film, written out as a PPM
// No window and no swapchain: the film is converted with the reference's R/B-swizzled view semantics removed
// (plain RGB) and clamped to 8 bits.
//
// Build:  hipcc -O2 -std=c++17 -I../../include raytraced_triangle.cpp -L.. -lrtx -Wl,-rpath,'$ORIGIN/..'
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "rtx.h"

#define CHECK_RTX(ctx, x)                                                                        \
    do {                                                                                         \
        int r_ = (x);                                                                            \
        if (r_ != RTX_SUCCESS) {                                                                 \
            std::fprintf(stderr, "%s failed (%d): %s\n", #x, r_, rtx_last_error_string(ctx));    \
            std::exit(1);                                                                        \
        }                                                                                        \
    } while (0)

#define CHECK_HIP(x)                                                                             \
    do {                                                                                         \
        hipError_t e_ = (x);                                                                     \
        if (e_ != hipSuccess) {                                                                  \
            std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));                  \
            std::exit(1);                                                                        \
        }                                                                                        \
    } while (0)

static const uint32_t WIDTH = 1200;  // main.cpp:13
static const uint32_t HEIGHT = 800;  // main.cpp:14

template <typename T>
static T* upload(const T* src, size_t count) {
    T* d = nullptr;
    CHECK_HIP(hipMalloc(&d, count * sizeof(T)));
    CHECK_HIP(hipMemcpy(d, src, count * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

static void* device_alloc(uint64_t bytes) {
    void* p = nullptr;
    CHECK_HIP(hipMalloc(&p, bytes));  // hipMalloc returns 256-B aligned memory
    return p;
}

int main(int argc, char** argv) {
    const char* out_path = argc > 1 ? argv[1] : "raytraced_triangle.ppm";
    rtx_context ctx;
    if (rtx_create_context(0, nullptr, &ctx) != RTX_SUCCESS) {
        std::fprintf(stderr, "no GPU\n");
        return 1;
    }

    // ---- createBLAS (main.cpp:674-831) ----
    const float vertices[4][3] = {{-1.0f, -1.0f, 0.0f}, {1.0f, -1.0f, 0.0f}, {1.0f, 1.0f, 0.0f}, {-1.0f, 1.0f, 0.0f}};
    const uint32_t indices[] = {0, 1, 3, 1, 2, 3};
    const float geo_transforms[2][12] = {{1, 0, 0, -2, 0, 1, 0, 0, 0, 0, 1, 0}, {1, 0, 0, 2, 0, 1, 0, 0, 0, 0, 1, 0}};
    float* d_vertices = upload(&vertices[0][0], 12);
    uint32_t* d_indices = upload(indices, 6);
    float* d_xforms = upload(&geo_transforms[0][0], 24);

    rtx_blas_geometry geometry0{};
    geometry0.vertex_data_device = d_vertices;
    geometry0.vertex_stride = sizeof(vertices[0]);
    geometry0.max_vertex = 3;
    geometry0.index_type = RTX_INDEX_TYPE_UINT32;
    geometry0.index_data_device = d_indices;
    geometry0.transform_data_device = d_xforms;
    geometry0.primitive_count = 2;
    geometry0.flags = RTX_GEOMETRY_OPAQUE_BIT;
    rtx_blas_geometry geometries[2] = {geometry0, geometry0};
    geometries[1].transform_offset = sizeof(geo_transforms[0]);  // main.cpp:801-804

    rtx_as_build_sizes blas_sizes;
    CHECK_RTX(ctx, rtx_get_blas_build_sizes(ctx, geometries, 2, RTX_BUILD_PREFER_FAST_TRACE_BIT, &blas_sizes));
    void* blas_mem = device_alloc(blas_sizes.acceleration_structure_size);
    void* blas_scratch = device_alloc(blas_sizes.build_scratch_size);
    uint64_t blas_address = 0;
    CHECK_RTX(ctx, rtx_cmd_build_blas(ctx, geometries, 2, RTX_BUILD_PREFER_FAST_TRACE_BIT, blas_mem,
                                      blas_sizes.acceleration_structure_size, blas_scratch,
                                      blas_sizes.build_scratch_size, &blas_address));
    CHECK_RTX(ctx, rtx_sync(ctx));  // vkQueueWaitIdle (main.cpp:820)
    CHECK_HIP(hipFree(blas_scratch));
    CHECK_HIP(hipFree(d_vertices));
    CHECK_HIP(hipFree(d_indices));
    CHECK_HIP(hipFree(d_xforms));

    // ---- createTLAS (main.cpp:833-949) ----
    rtx_instance instance0{};
    instance0.custom_index_and_mask = 100u | (0xFFu << 24);
    instance0.sbt_offset_and_flags = 0u | (RTX_INSTANCE_TRIANGLE_FACING_CULL_DISABLE_BIT << 24);
    instance0.blas_address = blas_address;
    rtx_instance instance_data[2] = {instance0, instance0};
    const float ins_transforms[2][12] = {{1, 0, 0, 0, 0, 1, 0, 2, 0, 0, 1, 0}, {1, 0, 0, 0, 0, 1, 0, -2, 0, 0, 1, 0}};
    for (int i = 0; i < 2; ++i)
        for (int k = 0; k < 12; ++k) instance_data[i].transform[k / 4][k % 4] = ins_transforms[i][k];
    instance_data[1].sbt_offset_and_flags = 2u | (RTX_INSTANCE_TRIANGLE_FACING_CULL_DISABLE_BIT << 24);  // :858
    rtx_instance* d_instances = upload(instance_data, 2);

    rtx_as_build_sizes tlas_sizes;
    CHECK_RTX(ctx, rtx_get_tlas_build_sizes(ctx, 2, RTX_BUILD_PREFER_FAST_TRACE_BIT, &tlas_sizes));
    void* tlas_mem = device_alloc(tlas_sizes.acceleration_structure_size);
    void* tlas_scratch = device_alloc(tlas_sizes.build_scratch_size);
    uint64_t tlas_address = 0;
    CHECK_RTX(ctx, rtx_cmd_build_tlas(ctx, d_instances, 2, RTX_BUILD_PREFER_FAST_TRACE_BIT, tlas_mem,
                                      tlas_sizes.acceleration_structure_size, tlas_scratch,
                                      tlas_sizes.build_scratch_size, &tlas_address));
    CHECK_RTX(ctx, rtx_sync(ctx));  // main.cpp:942
    CHECK_HIP(hipFree(tlas_scratch));
    CHECK_HIP(hipFree(d_instances));

    // ---- createOutImage / createUniformBuffer / createShaderBindingTable (main.cpp:951-1017, 1264-1320) ----
    float* d_film = (float*)device_alloc((uint64_t)WIDTH * HEIGHT * 16);
    const rtx_camera camera = {{0.0f, 0.0f, 10.0f}, 60.0f};  // main.cpp:1015
    rtx_hit_record records[4] = {};
    const float colors[4][3] = {{0.6f, 0.1f, 0.2f}, {0.1f, 0.8f, 0.4f}, {0.9f, 0.7f, 0.1f}, {0.3f, 0.6f, 0.9f}};
    for (int i = 0; i < 4; ++i) {
        records[i].program = RTX_PROGRAM_MATCH_BARYCENTRIC;  // the closest-hit shader, main.cpp:1080-1091
        for (int c = 0; c < 3; ++c) records[i].color[c] = colors[i][c];
        records[i].match_primitive = 1;
        records[i].match_instance = 1;
        records[i].match_custom_index = 100;
        records[i].match_geometry = 1;
    }
    rtx_hit_record* d_records = upload(records, 4);
    rtx_sbt sbt{};
    sbt.hit_records_device = d_records;
    sbt.hit_record_count = 4;
    sbt.sbt_record_offset = 0;
    sbt.sbt_record_stride = 1;
    sbt.miss_color[2] = 0.2f;  // main.cpp:1065

    // ---- render (main.cpp:1322-1423): one frame ----
    CHECK_RTX(ctx, rtx_cmd_trace_rays(ctx, tlas_address, &sbt, &camera, nullptr, WIDTH, HEIGHT, d_film));
    CHECK_RTX(ctx, rtx_sync(ctx));

    std::vector<float> film((size_t)WIDTH * HEIGHT * 4);
    CHECK_HIP(hipMemcpy(film.data(), d_film, film.size() * sizeof(float), hipMemcpyDeviceToHost));
    FILE* f = std::fopen(out_path, "wb");
    if (!f) return 1;
    std::fprintf(f, "P6\n%u %u\n255\n", WIDTH, HEIGHT);
    for (size_t i = 0; i < (size_t)WIDTH * HEIGHT; ++i)
        for (int c = 0; c < 3; ++c) {
            float v = film[4 * i + c];
            v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
            std::fputc((int)(v * 255.0f + 0.5f), f);
        }
    std::fclose(f);
    std::printf("wrote %s (%ux%u)\n", out_path, WIDTH, HEIGHT);

    CHECK_HIP(hipFree(d_film));
    CHECK_HIP(hipFree(d_records));
    CHECK_HIP(hipFree(tlas_mem));
    CHECK_HIP(hipFree(blas_mem));
    rtx_destroy_context(ctx);
    return 0;
}
