// pathtrace.hip — wavefront path tracer (Laine, Karras & Aila 2013, "Megakernels Considered Harmful") for gfx950,
// plain and with a neural radiance cache (Muller et al. 2021).
//
// One pass traces one path per path slot.  Per path vertex (depth d):
//   shade   emission seen at d = 0, next-event estimation toward one light point (shadow queue), Russian
//           roulette, cosine-weighted continuation (extension queue for d + 1); queue pushes are aggregated per
//           256-thread block (one atomic per queue per block)
//   trace   one persistent launch over the shadow queue of d (any hit) and the extension queue of d + 1
//           (closest hit with path data); counts are read from device memory                [trace.hip]
//   gather  unoccluded shadow contributions are added to their path's radiance
// Every path owns its slot in a pass, so no accumulation needs atomics and results do not depend on queue order.
//
// Cache mode (one pass per frame, Muller et al. 2021): path slots [0, npix) are render paths that stop at vertex
// `cache_depth` and query the cache there; slots [npix, npix + ntrain) are training paths from random pixels that
// share the same queues, record at every vertex d >= 1 the cache query, the throughput and the radiance gathered
// so far, and either run to full length or, with a training suffix, stop at vertex train_end and query the cache
// themselves (self-training).  All queries of the pass are inferred in one launch; every path adds throughput x
// prediction; then each record becomes a training sample whose target is (final radiance - radiance before the
// vertex) / throughput, i.e. the radiance the path carried from that vertex toward the previous one.  Samples are
// compacted with a deterministic block scan.
#include <algorithm>
#include <cstdlib>
#include <cmath>

#include "nrc.h"
#include "pathtrace.h"
#include "rtx_internal.h"
#include "pt_shade.h"
#include "sampling.h"

namespace rtx {

namespace {

// Block-wide compaction of up to 3 queues: one atomic per queue per 256-thread block.  Every thread of the block
// must call it.  Returns each thread's slot (or 0xFFFFFFFF) in o[q] and the block totals in tot[q].
__device__ __forceinline__ void block_push3(uint32_t* c0, uint32_t* c1, uint32_t* c2, const bool p[3], uint32_t o[3],
                                            uint32_t tot[3]) {
    __shared__ uint32_t s_cnt[3][4];
    __shared__ uint32_t s_base[3];
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint64_t m[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        m[q] = __ballot(p[q]);
        if (lane == 0) s_cnt[q][w] = (uint32_t)__popcll(m[q]);
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int q = threadIdx.x;
        const uint32_t t = s_cnt[q][0] + s_cnt[q][1] + s_cnt[q][2] + s_cnt[q][3];
        uint32_t* c = q == 0 ? c0 : (q == 1 ? c1 : c2);
        s_base[q] = (t && c) ? atomicAdd(c, t) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        uint32_t off = s_base[q];
        for (uint32_t k = 0; k < w; ++k) off += s_cnt[q][k];
        o[q] = p[q] ? off + (uint32_t)__popcll(m[q] & lt) : 0xFFFFFFFFu;
        tot[q] = s_cnt[q][0] + s_cnt[q][1] + s_cnt[q][2] + s_cnt[q][3];
    }
}

// counts[] slot of the training-tail query queue (past the per-depth extension / shadow counts)
__host__ __device__ __forceinline__ uint32_t tail_counter(const PtArgs& a) { return 3u * a.max_depth + 2u; }

// Per-view light CDFs of a batched render (one thread per view).
__global__ void light_cdf_batch_kernel(const BatchView* __restrict__ views, uint32_t nviews,
                                       const float* __restrict__ lights, float* __restrict__ cdf) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nviews) return;
    const BatchView bv = views[v];
    float acc = 0.0f;
    for (uint32_t i = 0; i < bv.light_count; ++i) {
        const float* L = lights + 12 * (uint64_t)(bv.light_first + i);
        const float e1x = L[3] - L[0], e1y = L[4] - L[1], e1z = L[5] - L[2];
        const float e2x = L[6] - L[0], e2y = L[7] - L[1], e2z = L[8] - L[2];
        const float cx = e1y * e2z - e1z * e2y, cy = e1z * e2x - e1x * e2z, cz = e1x * e2y - e1y * e2x;
        acc += 0.5f * sqrtf(cx * cx + cy * cy + cz * cz);
        cdf[bv.cdf_first + i] = acc;
    }
    cdf[bv.cdf_first + bv.light_count] = acc;
}

__global__ void light_cdf_kernel(const float* __restrict__ lights, uint32_t n, float* __restrict__ cdf) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float acc = 0.0f;
    for (uint32_t i = 0; i < n; ++i) {
        const float* L = lights + 12 * (uint64_t)i;
        const float e1x = L[3] - L[0], e1y = L[4] - L[1], e1z = L[5] - L[2];
        const float e2x = L[6] - L[0], e2y = L[7] - L[1], e2z = L[8] - L[2];
        const float cx = e1y * e2z - e1z * e2y, cy = e1z * e2x - e1x * e2z, cz = e1x * e2y - e1y * e2x;
        acc += 0.5f * sqrtf(cx * cx + cy * cy + cz * cz);
        cdf[i] = acc;
    }
    cdf[n] = acc;
}

// Camera rays of paths [first_path, npaths): path i enters the extension queue at slot i - first_path.
__global__ __launch_bounds__(256) void camera_kernel(PtArgs a, uint32_t sample) {
    if (a.guard && !guard_passes(a.tlas, 0u, a.guard)) return;  // the other form runs this sample
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t i = a.first_path + slot;
    if (slot == 0) a.counts[0] = a.npaths - a.first_path;
    if (i >= a.npaths) return;
    sample = pass_sample(a, sample);
    rtx_ray r;
    pt_camera_ray(a, i, sample, r);
    a.ext_rays[0][slot] = r;
    a.ext_pid[0][slot] = i;
    a.thr[i] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    a.rad[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// `cur` selects the extension queue pair with a conditional, never as a runtime index into the kernel argument
// (which would copy the whole argument block to per-lane scratch memory)
template <int MODE>
__global__ __launch_bounds__(256) void shade_kernel(PtArgs a, uint32_t depth, uint32_t sample, int cur) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t count = a.counts[3 * depth];
    sample = pass_sample(a, sample);
    ShadeOut so;
    uint32_t pid = 0;
    if (q < count) {
        pid = (cur ? a.ext_pid[1] : a.ext_pid[0])[q];
        const rtx_ray r = (cur ? a.ext_rays[1] : a.ext_rays[0])[q];
        const PathHit h = a.hits[q];
        shade_vertex<MODE>(a, pid, r, h, depth, sample, so);
    }
    const bool push_sh = so.push_sh, push_ext = so.push_ext, push_cq = so.push_cq;
    const rtx_ray& ext = so.ext;
    const rtx_ray& sh = so.sh;
    const rtx_nrc_query& cq = so.cq;
    const float4 contrib = so.contrib, thr_new = so.thr_new;
    const bool pr[3] = {push_sh, push_ext, push_cq};
    uint32_t o[3], tot[3];
    // render paths queue their cache queries at cache_depth (slots from 0), training tails at train_end (slots from
    // npix); the two never share a depth
    const bool tail_depth = MODE == kPtCache && depth != a.cache_depth;
    block_push3(&a.counts[3 * depth + 1], &a.counts[3 * (depth + 1)],
                MODE == kPtCache ? &a.counts[tail_depth ? tail_counter(a) : 2] : nullptr, pr, o, tot);
    if (tail_depth && push_cq) o[2] += a.npix;
    if (push_sh) {
        a.sh_rays[o[0]] = sh;
        a.sh_pid[o[0]] = pid;
        a.sh_contrib[o[0]] = contrib;
    }
    if (push_ext) {
        (cur ? a.ext_rays[0] : a.ext_rays[1])[o[1]] = ext;
        (cur ? a.ext_pid[0] : a.ext_pid[1])[o[1]] = pid;
        a.thr[pid] = thr_new;
    }
    if (MODE == kPtCache && push_cq) {
        a.cq[o[2]] = cq;
        a.cq_pid[o[2]] = pid;
    }
    if (threadIdx.x == 0 && (tot[0] + tot[1]) > 0) atomicAdd(&a.state->pt_rays, (unsigned long long)(tot[0] + tot[1]));
}

__global__ __launch_bounds__(256) void shadow_gather_kernel(PtArgs a, uint32_t depth) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.counts[3 * depth + 1]) return;
    if (a.occl[q]) return;
    const uint32_t pid = a.sh_pid[q];
    const float4 c = a.sh_contrib[q];
    float4 r = a.rad[pid];
    r.x += c.x;
    r.y += c.y;
    r.z += c.z;
    a.rad[pid] = r;
}

__global__ __launch_bounds__(256) void accumulate_kernel(PtArgs a, double* __restrict__ acc64) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.npix) return;
    const float4 r = a.rad[i];
    float4 s = a.accum[i];
    s.x += r.x;
    s.y += r.y;
    s.z += r.z;
    a.accum[i] = s;
    if (acc64) {
        acc64[3 * (uint64_t)i] += (double)r.x;
        acc64[3 * (uint64_t)i + 1] += (double)r.y;
        acc64[3 * (uint64_t)i + 2] += (double)r.z;
    }
}

// The views of a multi-sample plain launch: view k = the render's camera and film, pixel paths from k * npix, the
// pass's sample index + k, every light, and an unbounded clip box (the exit distance is +inf, as for the plain
// render's rays).
__global__ void sample_views_kernel(PtArgs a, uint32_t nviews, BatchView* __restrict__ views) {
    const uint32_t k = threadIdx.x;
    if (blockIdx.x != 0 || k >= nviews) return;
    BatchView v;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        v.cam[c] = a.cam[c];
        v.lo[c] = -INFINITY;
        v.hi[c] = INFINITY;
    }
    v.ax = a.ax;
    v.ay = a.ay;
    v.width = a.width;
    v.height = a.height;
    v.pix_base = k * a.npix;
    v.light_first = 0;
    v.light_count = a.nlights;
    v.cdf_first = 0;
    v.sample_offset = k;
    views[k] = v;
}

// accumulate_kernel for `nviews` samples at once: pixel i adds the radiance of its path in each view, in view
// (= sample) order, the same additions in the same order as one accumulate_kernel per sample.
__global__ __launch_bounds__(256) void accumulate_samples_kernel(PtArgs a, const float4* __restrict__ rad,
                                                                 uint32_t nviews, double* __restrict__ acc64) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.npix) return;
    float4 s = a.accum[i];
    double d[3] = {0.0, 0.0, 0.0};
    if (acc64) {
#pragma unroll
        for (int c = 0; c < 3; ++c) d[c] = acc64[3 * (uint64_t)i + c];
    }
    for (uint32_t k = 0; k < nviews; ++k) {
        const float4 r = rad[(uint64_t)k * a.npix + i];
        s.x += r.x;
        s.y += r.y;
        s.z += r.z;
        d[0] += (double)r.x;
        d[1] += (double)r.y;
        d[2] += (double)r.z;
    }
    a.accum[i] = s;
    if (acc64) {
#pragma unroll
        for (int c = 0; c < 3; ++c) acc64[3 * (uint64_t)i + c] = d[c];
    }
}

__global__ __launch_bounds__(256) void resolve_kernel(PtArgs a, float inv_spp, float* __restrict__ film) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.npix) return;
    const float4 s = a.accum[i];
    *(float4*)(film + 4 * (uint64_t)i) = make_float4(s.x * inv_spp, s.y * inv_spp, s.z * inv_spp, 1.0f);
}

__global__ void fill_words_kernel(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void count_camera_rays_kernel(DeviceState* st, uint32_t n, const void* tlas, uint32_t guard) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && (!guard || guard_passes(tlas, 0u, guard)))
        atomicAdd(&st->pt_rays, (unsigned long long)n);
}

// Folds the training wavefront's ray count and error flags into the context's state and clears them.
__global__ void merge_state_kernel(DeviceState* dst, DeviceState* src) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    atomicAdd(&dst->pt_rays, src->pt_rays);
    atomicOr(&dst->error_flags, src->error_flags);
    src->pt_rays = 0;
    src->error_flags = 0;
}

// ---- training samples from vertex records: block count -> scan -> scatter (deterministic order) ----------------
__device__ __forceinline__ bool record_valid(const PtArgs& a, uint64_t r) { return a.rec[r * 5 + 4].z > 0.5f; }

__global__ __launch_bounds__(256) void sample_count_kernel(PtArgs a, uint64_t nrec, uint32_t* __restrict__ block_counts) {
    const uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const bool v = r < nrec && record_valid(a, r);
    const uint64_t m = __ballot(v);
    __shared__ uint32_t wc[4];
    if ((threadIdx.x & 63u) == 0) wc[threadIdx.x >> 6] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) block_counts[blockIdx.x] = wc[0] + wc[1] + wc[2] + wc[3];
}

__global__ __launch_bounds__(1024) void sample_scan_kernel(uint32_t* __restrict__ block_counts, uint32_t nblocks,
                                                           uint32_t* __restrict__ total) {
    __shared__ uint32_t s[1024];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nblocks; base += 1024) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < nblocks ? block_counts[i] : 0u;
        s[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t off = 1; off < 1024; off <<= 1) {
            const uint32_t t = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nblocks) block_counts[i] = carry + s[threadIdx.x] - v;
        carry += s[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(256) void sample_scatter_kernel(PtArgs a, uint64_t nrec,
                                                             const uint32_t* __restrict__ block_offsets,
                                                             rtx_nrc_sample* __restrict__ out) {
    const uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const bool v = r < nrec && record_valid(a, r);
    const uint64_t m = __ballot(v);
    __shared__ uint32_t wc[4];
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    if (lane == 0) wc[w] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t off = block_offsets[blockIdx.x];
    for (uint32_t k = 0; k < w; ++k) off += wc[k];
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    if (!v) return;
    off += (uint32_t)__popcll(m & lt);
    const uint32_t pid = a.npix + (uint32_t)(r % a.ntrain);
    const float4* rp = a.rec + r * 5;
    const float4 q0 = rp[0], q1 = rp[1], q2 = rp[2], q3 = rp[3], q4 = rp[4];
    const float4 fin = a.rad[pid];
    const float pre[3] = {q3.w, q4.x, q4.y};
    const float thr[3] = {q3.x, q3.y, q3.z};
    const float total[3] = {fin.x, fin.y, fin.z};
    rtx_nrc_sample smp;
    smp.query.position[0] = q0.x;
    smp.query.position[1] = q0.y;
    smp.query.position[2] = q0.z;
    smp.query.normal[0] = q0.w;
    smp.query.normal[1] = q1.x;
    smp.query.normal[2] = q1.y;
    smp.query.direction[0] = q1.z;
    smp.query.direction[1] = q1.w;
    smp.query.direction[2] = q2.x;
    smp.query.albedo[0] = q2.y;
    smp.query.albedo[1] = q2.z;
    smp.query.albedo[2] = q2.w;
#pragma unroll
    for (int c = 0; c < 3; ++c) smp.target[c] = thr[c] > 0.0f ? fmaxf(0.0f, total[c] - pre[c]) / thr[c] : 0.0f;
    smp.reserved = 0.0f;
    out[off] = smp;
}

}  // namespace

static bool wavefront_forced() {
    const char* e = std::getenv("RTX_PATH_WAVEFRONT");
    return e && e[0] == '1';
}

bool pt_whole_paths(const PtArgs& a) { return !wavefront_forced() && !trace_needs_flags(0u, a.content); }

bool pt_dual_forms(const PtArgs& a) { return a.dual && !wavefront_forced(); }

uint64_t pt_scratch_bytes(uint32_t npix, uint32_t max_depth, uint32_t nlights) {
    uint64_t n = npix ? npix : 1;
    uint64_t b = 0;
    b += 2 * align_up(n * sizeof(rtx_ray), 256) + 2 * align_up(n * 4, 256);  // extension queues
    b += align_up(n * sizeof(PathHit), 256);                                  // hits
    b += align_up(n * sizeof(rtx_ray), 256) + align_up(n * 4, 256) + align_up(n * 16, 256);  // shadow queue
    b += align_up(n * 4, 256);                                                // occluded
    b += 3 * align_up(n * 16, 256);                                           // thr, rad, accum
    b += 2 * align_up((3ull * (max_depth + 2)) * 4, 256);                      // counts (two wavefronts)
    b += align_up((nlights + 1ull) * 4, 256);                                 // light cdf
    b += align_up(n * 4, 256);                                                // path view
    b += align_up(n * sizeof(rtx_nrc_query), 256) + align_up(n * 4, 256);  // cache queries
    return b;
}

void pt_layout(PtArgs& a, void* scratch, uint32_t npix, uint32_t max_depth, uint32_t nlights) {
    uint8_t* p = (uint8_t*)scratch;
    uint64_t n = npix ? npix : 1;
    auto take = [&](uint64_t bytes) {
        uint8_t* r = p;
        p += align_up(bytes, 256);
        return r;
    };
    a.ext_rays[0] = (rtx_ray*)take(n * sizeof(rtx_ray));
    a.ext_rays[1] = (rtx_ray*)take(n * sizeof(rtx_ray));
    a.ext_pid[0] = (uint32_t*)take(n * 4);
    a.ext_pid[1] = (uint32_t*)take(n * 4);
    a.hits = (PathHit*)take(n * sizeof(PathHit));
    a.sh_rays = (rtx_ray*)take(n * sizeof(rtx_ray));
    a.sh_pid = (uint32_t*)take(n * 4);
    a.sh_contrib = (float4*)take(n * 16);
    a.occl = (uint32_t*)take(n * 4);
    a.thr = (float4*)take(n * 16);
    a.rad = (float4*)take(n * 16);
    a.accum = (float4*)take(n * 16);
    a.counts = (uint32_t*)take(3ull * (max_depth + 2) * 4);
    a.train_counts = (uint32_t*)take(3ull * (max_depth + 2) * 4);
    a.cdf = (float*)take((nlights + 1ull) * 4);
    a.path_view = (uint32_t*)take(n * 4);
    a.cq = (rtx_nrc_query*)take(n * sizeof(rtx_nrc_query));
    a.cq_pid = (uint32_t*)take(n * 4);
}

uint64_t pt_train_bytes(uint32_t train_paths, uint32_t record_depths) {
    const uint64_t nrec = (uint64_t)train_paths * (record_depths ? record_depths : 1);
    const uint64_t nblk = (nrec + 255) / 256;
    return align_up(nrec * 80, 256) + align_up(nrec * sizeof(rtx_nrc_sample), 256) + align_up(nblk * 4, 256) + 256;
}

// One pass: vertices 0 .. last_depth (cache queries stay queued for the caller).
static hipError_t run_pass(PtArgs& a, const TraceLaunch& L, uint32_t sample, uint32_t last_depth, hipStream_t s) {
    const uint32_t n = a.npaths - a.first_path;
    const unsigned gb = (n + 255) / 256;
    const uint64_t ncounts = 3ull * (a.max_depth + 2);
    fill_words_kernel<<<1, 64, 0, s>>>(a.counts, ncounts, 0u);
    camera_kernel<<<gb, 256, 0, s>>>(a, sample);
    count_camera_rays_kernel<<<1, 64, 0, s>>>(a.state, n, a.tlas, a.guard);
    hipError_t e;
    // vertex 0's extension rays (camera rays) alone; afterwards vertex d's shadow rays and vertex d + 1's extension
    // rays, both produced by shade(d), go through one fused launch
    TraceRequest t{};
    t.tlas = a.tlas;
    t.cull_mask = 0xFFu;
    t.path_hits = a.hits;
    t.records = a.records;  // any-hit programs of non-opaque geometry
    t.nrec = a.nrec;
    t.sbt_offset = a.sbt_offset;
    t.sbt_stride = a.sbt_stride;
    t.state = a.state;
    t.gstack = a.gstack;
    t.content = a.content;
    t.mode = 3;
    t.n_ptr = a.counts;
    t.rays = a.ext_rays[0];
    if ((e = trace_rays_launch(t, L, s))) return e;
    int cur = 0;
    for (uint32_t d = 0; d <= last_depth; ++d) {
        if (a.mode == kPtCache)
            shade_kernel<kPtCache><<<gb, 256, 0, s>>>(a, d, sample, cur);
        else
            shade_kernel<kPtPlain><<<gb, 256, 0, s>>>(a, d, sample, cur);
        t.mode = 4;
        t.n_ptr = a.counts + 3 * (d + 1);
        t.rays = a.ext_rays[cur ^ 1];
        t.n2_ptr = a.counts + 3 * d + 1;
        t.rays2 = a.sh_rays;
        t.occluded = a.occl;
        if ((e = trace_rays_launch(t, L, s))) return e;
        shadow_gather_kernel<<<gb, 256, 0, s>>>(a, d);
        cur ^= 1;
    }
    return hipGetLastError();
}

// Cache predictions for one query queue, added to the paths' radiance by the inference kernel itself.
static hipError_t infer_gather(PtArgs& a, const NrcState& nrc, const uint32_t* count, uint32_t base, uint32_t cap,
                               hipStream_t s) {
    InferGather g;
    g.pid = a.cq_pid + base;
    g.thr = a.thr;
    g.rad = a.rad;
    return nrc_infer(nrc, a.cq + base, cap, nullptr, s, count, &g);
}

// One sample of a plain render: every pixel path whole in one traversal launch (each lane traces, shades and
// continues its own paths), or the per-vertex wavefront where that launch is not available.
static hipError_t render_sample(PtArgs& a, const TraceLaunch& L, uint32_t sample, hipStream_t s) {
    const bool dual = pt_dual_forms(a);
    if (!dual && !pt_whole_paths(a)) return run_pass(a, L, sample, a.max_depth - 1, s);
    TraceRequest tr{};
    tr.tlas = a.tlas;
    tr.mode = 6;
    tr.n = a.npix;
    tr.cull_mask = 0xFFu;
    tr.records = a.records;
    tr.nrec = a.nrec;
    tr.sbt_offset = a.sbt_offset;
    tr.sbt_stride = a.sbt_stride;
    tr.state = a.state;
    tr.gstack = a.gstack;
    tr.pt = &a;
    tr.pt_sample = sample;
    tr.content = dual ? 0u : a.content;
    tr.guard = dual ? 1u : 0u;
    hipError_t e = trace_rays_launch(tr, L, s);
    if (e || !dual) return e;
    PtArgs w = a;
    w.guard = 2;
    return run_pass(w, L, sample, a.max_depth - 1, s);
}

// The pixel paths of one NRC sample up to their cache queries (queued from a.counts[2]): whole paths in one
// traversal launch, or the per-vertex wavefront where that launch is not available.
static hipError_t render_cache_sample(PtArgs& a, const TraceLaunch& L, uint32_t sample, hipStream_t s) {
    if (!pt_whole_paths(a)) return run_pass(a, L, sample, a.cache_depth, s);
    fill_words_kernel<<<1, 64, 0, s>>>(a.counts, 3ull * (a.max_depth + 2), 0u);
    TraceRequest tr{};
    tr.tlas = a.tlas;
    tr.mode = 7;
    tr.n = a.npix;
    tr.cull_mask = 0xFFu;
    tr.records = a.records;
    tr.nrec = a.nrec;
    tr.sbt_offset = a.sbt_offset;
    tr.sbt_stride = a.sbt_stride;
    tr.state = a.state;
    tr.gstack = a.gstack;
    tr.pt = &a;
    tr.pt_sample = sample;
    tr.content = a.content;
    return trace_rays_launch(tr, L, s);
}

// Per path slot of a multi-sample launch, every array the whole-path traversal indexes by path: throughput and
// radiance, the pending continuation ray and shadow contribution (which wait in the path's slot), and its view.
uint64_t pt_sample_views_bytes(uint32_t npix, uint32_t samples) {
    const uint64_t n = (uint64_t)npix * samples;
    return align_up(samples * sizeof(BatchView), 256) + 3 * align_up(n * 16, 256) + align_up(n * sizeof(rtx_ray), 256) +
           align_up(n * 4, 256);
}

hipError_t pt_render(PtArgs a, const TraceLaunch& L, uint32_t spp, uint32_t sample_offset, float* film,
                     double* acc64, hipStream_t s) {
    const unsigned gb = (a.npix + 255) / 256;
    a.mode = kPtPlain;
    a.npaths = a.npix;
    a.ntrain = 0;
    if (a.nlights) light_cdf_kernel<<<1, 64, 0, s>>>(a.lights, a.nlights, a.cdf);
    fill_words_kernel<<<gb, 256, 0, s>>>((uint32_t*)a.accum, 4ull * a.npix, 0u);
    hipError_t e;
    const uint32_t S = std::min(spp, a.ms_max);
    if (a.ms_buf && !a.views && S > 1 && pt_whole_paths(a) && !pt_dual_forms(a)) {
        // several samples per whole-path launch, as views of the camera; their per-path state in ms_buf
        uint8_t* p = (uint8_t*)a.ms_buf;
        BatchView* views = (BatchView*)p;
        p += align_up(S * sizeof(BatchView), 256);
        const uint64_t n = (uint64_t)a.npix * S;
        PtArgs m = a;
        m.thr = (float4*)p;
        p += align_up(n * 16, 256);
        m.rad = (float4*)p;
        p += align_up(n * 16, 256);
        m.sh_contrib = (float4*)p;
        p += align_up(n * 16, 256);
        m.ext_rays[0] = (rtx_ray*)p;
        p += align_up(n * sizeof(rtx_ray), 256);
        m.path_view = (uint32_t*)p;
        m.views = views;
        sample_views_kernel<<<1, 64, 0, s>>>(a, S, views);
        for (uint32_t si = 0; si < spp; si += S) {
            const uint32_t k = std::min(S, spp - si);
            m.nviews = k;
            m.npix = a.npix * k;
            m.npaths = m.npix;
            if ((e = render_sample(m, L, sample_offset + si, s))) return e;
            accumulate_samples_kernel<<<gb, 256, 0, s>>>(a, m.rad, k, acc64);
        }
        if (film) resolve_kernel<<<gb, 256, 0, s>>>(a, 1.0f / (float)spp, film);
        return hipGetLastError();
    }
    for (uint32_t si = 0; si < spp; ++si) {
        if ((e = render_sample(a, L, sample_offset + si, s))) return e;
        accumulate_kernel<<<gb, 256, 0, s>>>(a, acc64);
    }
    if (film) resolve_kernel<<<gb, 256, 0, s>>>(a, spp ? 1.0f / (float)spp : 0.0f, film);
    return hipGetLastError();
}

hipError_t pt_render_batch(PtArgs a, const TraceLaunch& L, uint32_t spp, uint32_t sample_offset, float* films,
                           double* acc64, hipStream_t s) {
    const unsigned gb = (a.npix + 255) / 256;
    a.mode = kPtPlain;
    a.npaths = a.npix;
    a.ntrain = 0;
    light_cdf_batch_kernel<<<(a.nviews + 63) / 64, 64, 0, s>>>(a.views, a.nviews, a.lights, a.cdf);
    fill_words_kernel<<<gb, 256, 0, s>>>((uint32_t*)a.accum, 4ull * a.npix, 0u);
    hipError_t e;
    for (uint32_t si = 0; si < spp; ++si) {
        if ((e = render_sample(a, L, sample_offset + si, s))) return e;
        accumulate_kernel<<<gb, 256, 0, s>>>(a, acc64);
    }
    if (films) resolve_kernel<<<gb, 256, 0, s>>>(a, spp ? 1.0f / (float)spp : 0.0f, films);
    return hipGetLastError();
}

hipError_t pt_render_nrc(PtArgs a, const TraceLaunch& L, const SideChain& side, NrcState& nrc,
                         uint32_t train_paths, uint32_t train_suffix, void* train_buf, uint32_t spp,
                         uint32_t sample_offset, float* film, double* acc64, hipStream_t s) {
    const unsigned gb = (a.npix + 255) / 256;
    hipError_t e;
    a.dual = 0;  // NRC frames take one form (the wavefront where the content is not known)
    // spp == 0: a frame that only trains (its training paths and Adam steps, exactly as in a rendered frame; no
    // pixel paths, film and accumulator untouched)
    const bool train_only = spp == 0;
    if (a.nlights) light_cdf_kernel<<<1, 64, 0, s>>>(a.lights, a.nlights, a.cdf);
    if (!train_only) fill_words_kernel<<<gb, 256, 0, s>>>((uint32_t*)a.accum, 4ull * a.npix, 0u);
    a.mode = kPtCache;
    a.tail = train_suffix > 0 ? 1u : 0u;
    a.train_end = a.tail ? std::min(a.max_depth - 1, a.cache_depth + train_suffix) : a.max_depth - 1;
    a.first_path = 0;
    a.npaths = a.npix;
    const uint32_t rec_depths = nrc_record_depths(a.max_depth, a.cache_depth, train_suffix);
    for (uint32_t si = 0; si < (train_only ? 1u : spp); ++si) {
        // the first sample carries the frame's training paths
        const uint32_t ntrain = si == 0 ? train_paths : 0u;
        a.ntrain = ntrain;
        if (ntrain == 0) {
            if ((e = render_cache_sample(a, L, sample_offset + si, s))) return e;
            if ((e = infer_gather(a, nrc, a.counts + 2, 0, a.npix, s))) return e;
            accumulate_kernel<<<gb, 256, 0, s>>>(a, acc64);
            continue;
        }
        const uint64_t nrec = (uint64_t)ntrain * rec_depths;
        const uint64_t nblk = (nrec + 255) / 256;
        uint8_t* p = (uint8_t*)train_buf;
        a.rec = (float4*)p;
        p += align_up(nrec * 80, 256);
        rtx_nrc_sample* samples = (rtx_nrc_sample*)p;
        p += align_up(nrec * sizeof(rtx_nrc_sample), 256);
        uint32_t* block_counts = (uint32_t*)p;
        p += align_up(nblk * 4, 256);
        uint32_t* total = (uint32_t*)p;
        // Training paths [npix, npix + ntrain) form a second wavefront on the side stream: its queues are the
        // slots from npix of the pass's queue buffers, its counters train_counts, its traversals count rays on
        // side.state.  Path slots, queue slots, cache-query slots (tails from npix) and records of the two
        // wavefronts are disjoint, and every path draws its random numbers from (pixel or path index, sample,
        // depth), so the frame's results do not depend on how the two interleave.
        PtArgs t = a;
        t.first_path = a.npix;
        t.npaths = a.npix + ntrain;
        for (int k = 0; k < 2; ++k) {
            t.ext_rays[k] = a.ext_rays[k] + a.npix;
            t.ext_pid[k] = a.ext_pid[k] + a.npix;
        }
        t.hits = a.hits + a.npix;
        t.sh_rays = a.sh_rays + a.npix;
        t.sh_pid = a.sh_pid + a.npix;
        t.sh_contrib = a.sh_contrib + a.npix;
        t.occl = a.occl + a.npix;
        t.counts = a.train_counts;
        t.state = side.state;
        t.gstack = side.gstack;
        if ((e = hipEventRecord(side.fork, s))) return e;
        if ((e = hipStreamWaitEvent(side.stream, side.fork, 0))) return e;
        // render paths up to their cache vertex (a few long launches, enqueued first so that the GPU has them
        // while the host enqueues the training wavefront's many short ones)
        if (!train_only && (e = render_cache_sample(a, L, sample_offset + si, s))) return e;
        fill_words_kernel<<<1024, 256, 0, side.stream>>>((uint32_t*)a.rec, nrec * 20, 0u);
        if (!t.tail && pt_whole_paths(t)) {
            // the whole training wavefront in one traversal launch (each lane shades and continues its own path)
            TraceRequest tr{};
            tr.tlas = t.tlas;
            tr.mode = 5;
            tr.n = ntrain;
            tr.cull_mask = 0xFFu;
            tr.records = t.records;
            tr.nrec = t.nrec;
            tr.sbt_offset = t.sbt_offset;
            tr.sbt_stride = t.sbt_stride;
            tr.state = t.state;
            tr.gstack = t.gstack;
            tr.pt = &t;
            tr.pt_sample = sample_offset + si;
            tr.content = t.content;
            if ((e = trace_rays_launch(tr, side.launch, side.stream))) return e;
        } else if ((e = run_pass(t, side.launch, sample_offset + si, t.train_end, side.stream))) {
            return e;
        }
        // training tails (self-training) are predicted with the cache before this frame's Adam steps
        if (t.tail && (e = infer_gather(t, nrc, t.counts + tail_counter(t), a.npix, ntrain, side.stream))) return e;
        // the samples and the Adam steps follow the training paths on the side stream, beside the render paths'
        // launch if that is still running; then the render queries with the updated cache
        sample_count_kernel<<<(unsigned)nblk, 256, 0, side.stream>>>(a, nrec, block_counts);
        sample_scan_kernel<<<1, 1024, 0, side.stream>>>(block_counts, (uint32_t)nblk, total);
        sample_scatter_kernel<<<(unsigned)nblk, 256, 0, side.stream>>>(a, nrec, block_counts, samples);
        // interleaved batches over the device-resident sample count (batch b takes samples b, b + nb, b + 2 nb, ...:
        // all depths in every batch): as many steps as the largest count needs, the surplus ones return at once, so
        // the frame has no host round trip
        const uint32_t max_batches = (uint32_t)((nrec + nrc.batch_size - 1) / nrc.batch_size);
        if ((e = nrc_steps_device(nrc, samples, total, max_batches, side.stream))) return e;
        merge_state_kernel<<<1, 64, 0, side.stream>>>(a.state, side.state);
        if ((e = hipEventRecord(side.join, side.stream))) return e;
        if ((e = hipStreamWaitEvent(s, side.join, 0))) return e;
        if (train_only) break;
        if ((e = infer_gather(a, nrc, a.counts + 2, 0, a.npix, s))) return e;
        accumulate_kernel<<<gb, 256, 0, s>>>(a, acc64);
    }
    if (film && !train_only) resolve_kernel<<<gb, 256, 0, s>>>(a, 1.0f / (float)spp, film);
    return hipGetLastError();
}

uint32_t nrc_record_depths(uint32_t max_depth, uint32_t cache_depth, uint32_t train_suffix) {
    // records at vertices 1 .. train_end - 1 (suffix) or 1 .. max_depth - 1 (full length)
    if (train_suffix > 0) {
        const uint32_t end = std::min(max_depth - 1, cache_depth + train_suffix);
        return end > 1 ? end - 1 : 1;
    }
    return max_depth > 1 ? max_depth - 1 : 1;
}

}  // namespace rtx
