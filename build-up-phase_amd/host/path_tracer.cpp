// path_tracer.cpp — path tracing on librtx from C++: an OBJ scene with one Lambertian material per OBJ group,
// rendered by the plain path tracer (rtx_cmd_render), by the same frame recorded into a HIP graph (BLAS build, TLAS
// build and render captured from the context's stream and replayed) and by the neural radiance cache renderer
// (rtx_cmd_render_nrc), the same flow the reference's render loop (main.cpp:1322-1423) drives with
// vkCmdTraceRaysKHR and its recorded command buffers.
//
// Usage: path_tracer scene.obj materials.txt out_prefix [width height spp nrc_frames]
//   materials.txt: "camera x y z fov", then one line per OBJ group, "r g b er eg eb": albedo and emitted radiance.
//   The triangles of emitting groups are the lights (next-event estimation picks them by area).
//   Defaults: 1920 x 1080, 64 spp, 64 NRC frames (65,536 training paths each, the first 2 frames not averaged).
// Writes out_prefix_plain.f32 and out_prefix_nrc.f32 (RGBA32F films, row-major, width x height x 4 floats) and
// an 8-bit .ppm of each (linear radiance clamped to [0, 1]).
//
// Build:  make -C build-up-phase_amd samples
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
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

template <typename T>
static T* upload(const T* src, size_t count) {
    T* d = nullptr;
    CHECK_HIP(hipMalloc(&d, std::max<size_t>(count, 1) * sizeof(T)));
    if (count) CHECK_HIP(hipMemcpy(d, src, count * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

static void* device_alloc(uint64_t bytes) {
    void* p = nullptr;
    CHECK_HIP(hipMalloc(&p, bytes));
    return p;
}

static void write_film(const std::string& prefix, const std::vector<float>& rgba, uint32_t w, uint32_t h) {
    FILE* f = std::fopen((prefix + ".f32").c_str(), "wb");
    if (!f || std::fwrite(rgba.data(), sizeof(float), rgba.size(), f) != rgba.size()) {
        std::fprintf(stderr, "cannot write %s.f32\n", prefix.c_str());
        std::exit(1);
    }
    std::fclose(f);
    f = std::fopen((prefix + ".ppm").c_str(), "wb");
    if (!f) {
        std::fprintf(stderr, "cannot write %s.ppm\n", prefix.c_str());
        std::exit(1);
    }
    std::fprintf(f, "P6\n%u %u\n255\n", w, h);
    std::vector<unsigned char> row(3 * (size_t)w);
    for (uint32_t y = 0; y < h; ++y) {
        for (uint32_t x = 0; x < w; ++x)
            for (int c = 0; c < 3; ++c) {
                const float v = std::min(1.0f, std::max(0.0f, rgba[4 * ((size_t)y * w + x) + c]));
                row[3 * x + c] = (unsigned char)(v * 255.0f + 0.5f);
            }
        std::fwrite(row.data(), 1, row.size(), f);
    }
    std::fclose(f);
}

// mean of the FP64 per-pixel sums over `samples` samples, as an RGBA32F film (alpha 1)
static std::vector<float> resolve(const double* d_accum, uint32_t w, uint32_t h, uint32_t samples) {
    std::vector<double> acc(3 * (size_t)w * h);
    CHECK_HIP(hipMemcpy(acc.data(), d_accum, acc.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<float> film(4 * (size_t)w * h);
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        for (int c = 0; c < 3; ++c) film[4 * i + c] = (float)(acc[3 * i + c] / (double)samples);
        film[4 * i + 3] = 1.0f;
    }
    return film;
}

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s scene.obj materials.txt out_prefix [width height spp nrc_frames]\n", argv[0]);
        return 2;
    }
    const std::string prefix = argv[3];
    const uint32_t W = argc > 4 ? (uint32_t)std::atoi(argv[4]) : 1920u;
    const uint32_t H = argc > 5 ? (uint32_t)std::atoi(argv[5]) : 1080u;
    const uint32_t SPP = argc > 6 ? (uint32_t)std::atoi(argv[6]) : 64u;
    const uint32_t FRAMES = argc > 7 ? (uint32_t)std::atoi(argv[7]) : 64u;
    const uint32_t kWarmup = 2;  // first NRC frames: training only, nothing averaged (the cache starts untrained)
    if (W == 0 || H == 0 || SPP == 0 || FRAMES <= kWarmup) {
        std::fprintf(stderr, "width, height and spp must be positive and nrc_frames above %u\n", kWarmup);
        return 2;
    }

    // a stream of the program's own (default priority): a graph cannot be captured from the null stream
    hipStream_t stream;
    if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) {
        std::fprintf(stderr, "no GPU\n");
        return 1;
    }
    rtx_context ctx;
    if (rtx_create_context(0, stream, &ctx) != RTX_SUCCESS) {
        std::fprintf(stderr, "no GPU\n");
        return 1;
    }

    // ---- scene: OBJ geometry (one BLAS geometry per group), materials, lights ----
    rtx_obj_mesh mesh{};
    CHECK_RTX(ctx, rtx_obj_load(argv[1], &mesh));
    FILE* mf = std::fopen(argv[2], "r");
    if (!mf) {
        std::fprintf(stderr, "cannot open %s\n", argv[2]);
        return 1;
    }
    rtx_camera camera{};
    if (std::fscanf(mf, " camera %f %f %f %f", &camera.position[0], &camera.position[1], &camera.position[2],
                    &camera.y_fov_degree) != 4) {
        std::fprintf(stderr, "%s: expected \"camera x y z fov\"\n", argv[2]);
        return 1;
    }
    std::vector<rtx_hit_record> records(mesh.geometry_count);
    for (uint32_t g = 0; g < mesh.geometry_count; ++g) {
        rtx_hit_record& r = records[g];
        r = rtx_hit_record{};
        r.program = RTX_PROGRAM_LAMBERTIAN;
        if (std::fscanf(mf, " %f %f %f %f %f %f", &r.color[0], &r.color[1], &r.color[2], &r.emission[0],
                        &r.emission[1], &r.emission[2]) != 6) {
            std::fprintf(stderr, "%s: expected a material line for each of the %u OBJ groups\n", argv[2],
                         mesh.geometry_count);
            return 1;
        }
    }
    std::fclose(mf);
    std::vector<float> lights;  // 12 floats per emitting triangle: v0, v1, v2, emitted radiance
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (uint32_t v = 0; v < mesh.vertex_count; ++v)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], mesh.vertices[3 * v + a]);
            hi[a] = std::max(hi[a], mesh.vertices[3 * v + a]);
        }
    for (uint32_t g = 0; g < mesh.geometry_count; ++g) {
        const rtx_hit_record& r = records[g];
        if (r.emission[0] <= 0.0f && r.emission[1] <= 0.0f && r.emission[2] <= 0.0f) continue;
        for (uint32_t t = mesh.geometry_first_triangle[g]; t < mesh.geometry_first_triangle[g + 1]; ++t) {
            for (int k = 0; k < 3; ++k)
                for (int a = 0; a < 3; ++a) lights.push_back(mesh.vertices[3 * mesh.indices[3 * t + k] + a]);
            for (int c = 0; c < 3; ++c) lights.push_back(r.emission[c]);
        }
    }

    // ---- BLAS over the OBJ groups (createBLAS, main.cpp:674-831) ----
    float* d_vertices = upload(mesh.vertices, 3 * (size_t)mesh.vertex_count);
    uint32_t* d_indices = upload(mesh.indices, 3 * (size_t)mesh.triangle_count);
    std::vector<rtx_blas_geometry> geoms(mesh.geometry_count);
    for (uint32_t g = 0; g < mesh.geometry_count; ++g) {
        rtx_blas_geometry& gd = geoms[g];
        gd = rtx_blas_geometry{};
        gd.vertex_data_device = d_vertices;
        gd.vertex_stride = 12;
        gd.max_vertex = mesh.vertex_count ? mesh.vertex_count - 1 : 0;
        gd.index_type = RTX_INDEX_TYPE_UINT32;
        gd.index_data_device = d_indices;
        gd.primitive_count = mesh.geometry_first_triangle[g + 1] - mesh.geometry_first_triangle[g];
        gd.primitive_offset = 12u * mesh.geometry_first_triangle[g];
        gd.flags = RTX_GEOMETRY_OPAQUE_BIT;
    }
    rtx_as_build_sizes bs;
    CHECK_RTX(ctx, rtx_get_blas_build_sizes(ctx, geoms.data(), mesh.geometry_count, RTX_BUILD_PREFER_FAST_TRACE_BIT,
                                            &bs));
    void* blas_mem = device_alloc(bs.acceleration_structure_size);
    void* blas_scratch = device_alloc(bs.build_scratch_size);
    uint64_t blas_address = 0;
    CHECK_RTX(ctx, rtx_cmd_build_blas(ctx, geoms.data(), mesh.geometry_count, RTX_BUILD_PREFER_FAST_TRACE_BIT,
                                      blas_mem, bs.acceleration_structure_size, blas_scratch, bs.build_scratch_size,
                                      &blas_address));
    CHECK_RTX(ctx, rtx_sync(ctx));

    // ---- TLAS: one identity instance (createTLAS, main.cpp:833-949) ----
    rtx_instance inst{};
    inst.transform[0][0] = inst.transform[1][1] = inst.transform[2][2] = 1.0f;
    inst.custom_index_and_mask = 0xFFu << 24;
    inst.blas_address = blas_address;
    rtx_instance* d_inst = upload(&inst, 1);
    rtx_as_build_sizes ts;
    CHECK_RTX(ctx, rtx_get_tlas_build_sizes(ctx, 1, RTX_BUILD_PREFER_FAST_TRACE_BIT, &ts));
    void* tlas_mem = device_alloc(ts.acceleration_structure_size);
    void* tlas_scratch = device_alloc(ts.build_scratch_size);
    uint64_t tlas_address = 0;
    CHECK_RTX(ctx, rtx_cmd_build_tlas(ctx, d_inst, 1, RTX_BUILD_PREFER_FAST_TRACE_BIT, tlas_mem,
                                      ts.acceleration_structure_size, tlas_scratch, ts.build_scratch_size,
                                      &tlas_address));
    CHECK_RTX(ctx, rtx_sync(ctx));

    // ---- hit records (black environment), lights, films ----
    rtx_hit_record* d_records = upload(records.data(), records.size());
    rtx_sbt sbt{};
    sbt.hit_records_device = d_records;
    sbt.hit_record_count = (uint32_t)records.size();
    sbt.sbt_record_stride = 1;
    float* d_lights = upload(lights.data(), lights.size());
    float* d_film = (float*)device_alloc((uint64_t)W * H * 16);
    double* d_accum = (double*)device_alloc((uint64_t)W * H * 24);

    rtx_render_params p{};
    p.width = W;
    p.height = H;
    p.max_depth = 8;
    p.rr_depth = 3;
    p.seed = 1234;
    p.light_triangles_device = d_lights;
    p.light_triangle_count = (uint32_t)(lights.size() / 12);
    p.ray_offset = 1e-4f;

    hipEvent_t e0, e1;
    CHECK_HIP(hipEventCreate(&e0));
    CHECK_HIP(hipEventCreate(&e1));
    float ms = 0.0f;

    // ---- plain path tracing: SPP samples per pixel in one call ----
    CHECK_HIP(hipMemset(d_accum, 0, (size_t)W * H * 24));
    p.spp = SPP;
    p.sample_offset = 0;
    CHECK_HIP(hipEventRecord(e0, (hipStream_t)rtx_get_stream(ctx)));
    CHECK_RTX(ctx, rtx_cmd_render(ctx, tlas_address, &sbt, &camera, &p, d_film, d_accum));
    CHECK_HIP(hipEventRecord(e1, (hipStream_t)rtx_get_stream(ctx)));
    CHECK_RTX(ctx, rtx_sync(ctx));
    CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
    std::printf("plain: %u spp at %ux%u in %.2f ms (%u triangles, %zu lights)\n", SPP, W, H, ms, mesh.triangle_count,
                lights.size() / 12);
    write_film(prefix + "_plain", resolve(d_accum, W, H, SPP), W, H);
    std::vector<float> film_direct((size_t)W * H * 4), film_recorded((size_t)W * H * 4);
    CHECK_HIP(hipMemcpy(film_direct.data(), d_film, film_direct.size() * 4, hipMemcpyDeviceToHost));

    // ---- the same frame recorded: BLAS build, TLAS build and render captured from the context's stream into a
    //      HIP graph (the command-buffer form of main.cpp:789-812, 930-942, 1340-1356) and replayed; every replay
    //      rebuilds both structures from the device data, and the render of a TLAS built in the same recording is
    //      recorded in both integrator forms, of which the device runs the one its content allows ----
    float* d_film_rec = (float*)device_alloc((uint64_t)W * H * 16);
    hipGraph_t graph;
    hipGraphExec_t exec;
    CHECK_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    CHECK_RTX(ctx, rtx_cmd_build_blas(ctx, geoms.data(), mesh.geometry_count, RTX_BUILD_PREFER_FAST_TRACE_BIT,
                                      blas_mem, bs.acceleration_structure_size, blas_scratch, bs.build_scratch_size,
                                      &blas_address));
    CHECK_RTX(ctx, rtx_cmd_build_tlas(ctx, d_inst, 1, RTX_BUILD_PREFER_FAST_TRACE_BIT, tlas_mem,
                                      ts.acceleration_structure_size, tlas_scratch, ts.build_scratch_size,
                                      &tlas_address));
    CHECK_RTX(ctx, rtx_cmd_render(ctx, tlas_address, &sbt, &camera, &p, d_film_rec, nullptr));
    CHECK_HIP(hipStreamEndCapture(stream, &graph));
    CHECK_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    CHECK_HIP(hipGraphLaunch(exec, stream));  // first replay (uploads the graph)
    CHECK_RTX(ctx, rtx_sync(ctx));
    const uint32_t kReplays = 3;
    CHECK_HIP(hipEventRecord(e0, stream));
    for (uint32_t k = 0; k < kReplays; ++k) CHECK_HIP(hipGraphLaunch(exec, stream));
    CHECK_HIP(hipEventRecord(e1, stream));
    CHECK_RTX(ctx, rtx_sync(ctx));
    CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
    CHECK_HIP(hipMemcpy(film_recorded.data(), d_film_rec, film_recorded.size() * 4, hipMemcpyDeviceToHost));
    const bool same = std::memcmp(film_direct.data(), film_recorded.data(), film_direct.size() * 4) == 0;
    std::printf("recorded: BLAS + TLAS build + %u-spp render replayed in %.2f ms per frame; film equals the direct "
                "call's: %s\n", SPP, ms / (float)kReplays, same ? "yes" : "no");
    CHECK_HIP(hipGraphExecDestroy(exec));
    CHECK_HIP(hipGraphDestroy(graph));
    CHECK_HIP(hipFree(d_film_rec));
    if (!same) return 1;

    // ---- neural radiance cache: one sample per frame, 65,536 training paths per frame ----
    rtx_nrc_config c{};
    c.seed = 1;
    c.batch_size = 16384;
    c.learning_rate = 1e-3f;
    c.beta1 = 0.9f;
    c.beta2 = 0.99f;
    c.epsilon = 1e-8f;
    for (int a = 0; a < 3; ++a) {
        c.scene_min[a] = lo[a];
        c.scene_max[a] = hi[a];
    }
    c.cache_vertex = 1;
    rtx_nrc nrc;
    CHECK_RTX(ctx, rtx_nrc_create(ctx, &c, &nrc));
    CHECK_HIP(hipMemset(d_accum, 0, (size_t)W * H * 24));
    CHECK_HIP(hipEventRecord(e0, (hipStream_t)rtx_get_stream(ctx)));
    for (uint32_t frame = 0; frame < FRAMES; ++frame) {
        p.sample_offset = frame;
        // the first frames only train the fresh cache (no pixel paths); the others render one sample each
        p.spp = frame >= kWarmup ? 1 : 0;
        CHECK_RTX(ctx, rtx_cmd_render_nrc(ctx, tlas_address, &sbt, &camera, &p, nrc, 65536, d_film,
                                          frame >= kWarmup ? d_accum : nullptr));
    }
    CHECK_HIP(hipEventRecord(e1, (hipStream_t)rtx_get_stream(ctx)));
    CHECK_RTX(ctx, rtx_sync(ctx));
    CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
    std::printf("nrc: %u frames in %.2f ms (%.3f ms per frame)\n", FRAMES, ms, ms / (float)FRAMES);
    write_film(prefix + "_nrc", resolve(d_accum, W, H, FRAMES - kWarmup), W, H);
    CHECK_RTX(ctx, rtx_nrc_destroy(nrc));

    CHECK_HIP(hipFree(d_film));
    CHECK_HIP(hipFree(d_accum));
    CHECK_HIP(hipFree(d_lights));
    CHECK_HIP(hipFree(d_records));
    CHECK_HIP(hipFree(tlas_mem));
    CHECK_HIP(hipFree(tlas_scratch));
    CHECK_HIP(hipFree(d_inst));
    CHECK_HIP(hipFree(blas_mem));
    CHECK_HIP(hipFree(blas_scratch));
    CHECK_HIP(hipFree(d_vertices));
    CHECK_HIP(hipFree(d_indices));
    rtx_obj_free(&mesh);
    CHECK_RTX(ctx, rtx_destroy_context(ctx));
    CHECK_HIP(hipStreamDestroy(stream));
    return 0;
}
