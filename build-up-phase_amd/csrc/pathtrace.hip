// pathtrace.hip — wavefront path tracer (Laine, Karras & Aila 2013, "Megakernels Considered Harmful") for gfx950,
// plain and with a neural radiance cache (Muller et al. 2021).
//
// One pass traces one path per path slot.  Per path vertex (depth d):
//   extend  persistent traversal over the extension queue (ray count read from device memory)  [trace.hip]
//   shade   emission seen at d = 0, next-event estimation toward one light point (shadow queue), Russian
//           roulette, cosine-weighted continuation (extension queue for d + 1); queues are compacted with one
//           64-lane ballot + one atomic per wave
//   shadow  any-hit traversal over the shadow queue                                               [trace.hip]
//   gather  unoccluded shadow contributions are added to their path's radiance
// Every path owns its slot in a pass, so no accumulation needs atomics and results do not depend on queue order.
//
// Cache mode: render paths stop at vertex `cache_depth`, emit a cache query there, and add throughput x the
// cache's prediction.  Training mode: paths start at random pixels and record, at every vertex d >= 1, the cache
// query, the throughput and the radiance gathered so far; after the pass the record becomes a training sample
// whose target is (final radiance - radiance before the vertex) / throughput, i.e. the radiance the path carried
// from that vertex toward the previous one.  Samples are compacted with a deterministic block scan.
#include <cmath>

#include "nrc.h"
#include "pathtrace.h"
#include "rtx_internal.h"
#include "sampling.h"

namespace rtx {

namespace {

constexpr uint32_t kTrainKey = 0x5EEDu;

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

__device__ __forceinline__ void path_xy(const PtArgs& a, uint32_t pid, uint32_t& x, uint32_t& y) {
    if (a.mode == kPtTrain) {
        x = pid;
        y = 0x7FFFFFFFu;
    } else {
        x = pid % a.width;
        y = pid / a.width;
    }
}

__global__ __launch_bounds__(256) void camera_kernel(PtArgs a, uint32_t sample) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) a.counts[0] = a.npaths;
    if (i >= a.npaths) return;
    uint32_t px, py;
    uint4 j;
    if (a.mode == kPtTrain) {
        const uint4 h = pcg4d(i, sample, a.seed, kTrainKey);
        const uint32_t pix = h.x % a.npix;
        px = pix % a.width;
        py = pix / a.width;
        j = h;
        j.x = h.y;
        j.y = h.z;
    } else {
        px = i % a.width;
        py = i / a.width;
        j = pcg4d(px, py, sample, a.seed ^ kJitterKey);
    }
    const float sx = ((float)px + unit_float(j.x)) / (float)a.width;
    const float sy = ((float)py + unit_float(j.y)) / (float)a.height;
    const float nx = sx * 2.0f - 1.0f;
    const float ny = sy * 2.0f - 1.0f;
    rtx_ray r;
    r.origin[0] = a.cam[0];
    r.origin[1] = a.cam[1];
    r.origin[2] = a.cam[2];
    r.tmin = 0.0f;
    r.direction[0] = nx * a.ax;
    r.direction[1] = -(ny * a.ay);
    r.direction[2] = -1.0f;
    r.tmax = INFINITY;
    a.ext_rays[0][i] = r;
    a.ext_pid[0][i] = i;
    a.thr[i] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    a.rad[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

template <int MODE>
__global__ __launch_bounds__(256) void shade_kernel(PtArgs a, uint32_t depth, uint32_t sample, int cur) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t count = a.counts[3 * depth];
    const bool cache_vertex = MODE == kPtCache && depth == a.cache_depth;
    bool push_ext = false, push_sh = false, push_cq = false;
    rtx_ray ext{}, sh{};
    rtx_nrc_query cq{};
    float4 contrib = make_float4(0, 0, 0, 0);
    float4 thr_new = make_float4(0, 0, 0, 0);
    uint32_t pid = 0;
    if (q < count) {
        pid = a.ext_pid[cur][q];
        const rtx_ray r = a.ext_rays[cur][q];
        const PathHit h = a.hits[q];
        if (h.rec != 0xFFFFFFFFu && h.rec < a.nrec) {
            const rtx_hit_record rec = a.records[h.rec];
            const float alb[3] = {rec.color[0], rec.color[1], rec.color[2]};
            const bool lambert = rec.program == RTX_PROGRAM_LAMBERTIAN;
            const float d[3] = {r.direction[0], r.direction[1], r.direction[2]};
            const float dn = h.n[0] * d[0] + h.n[1] * d[1] + h.n[2] * d[2];
            const bool front = dn < 0.0f;
            const float4 thr = a.thr[pid];
            if (depth == 0 && front && lambert) {
                float4 rd = a.rad[pid];
                rd.x += thr.x * rec.emission[0];
                rd.y += thr.y * rec.emission[1];
                rd.z += thr.z * rec.emission[2];
                a.rad[pid] = rd;
            }
            const float nf[3] = {front ? h.n[0] : -h.n[0], front ? h.n[1] : -h.n[1], front ? h.n[2] : -h.n[2]};
            float P[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) P[k] = fmaf(h.t, d[k], r.origin[k]);
            float org[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) org[k] = fmaf(a.ray_offset, nf[k], P[k]);
            const bool reflective = alb[0] > 0.0f || alb[1] > 0.0f || alb[2] > 0.0f;
            if (MODE != kPtPlain && (cache_vertex || (MODE == kPtTrain && depth >= 1))) {
                const float il = 1.0f / sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    cq.position[k] = P[k];
                    cq.normal[k] = nf[k];
                    cq.direction[k] = -d[k] * il;
                    cq.albedo[k] = alb[k];
                }
            }
            if (MODE == kPtTrain && depth >= 1) {
                // record the vertex before its own next-event contribution is gathered
                const float4 rd = a.rad[pid];
                float4* rp = a.rec + ((uint64_t)(depth - 1) * a.npaths + pid) * 5;
                rp[0] = make_float4(cq.position[0], cq.position[1], cq.position[2], cq.normal[0]);
                rp[1] = make_float4(cq.normal[1], cq.normal[2], cq.direction[0], cq.direction[1]);
                rp[2] = make_float4(cq.direction[2], cq.albedo[0], cq.albedo[1], cq.albedo[2]);
                rp[3] = make_float4(thr.x, thr.y, thr.z, rd.x);
                rp[4] = make_float4(rd.y, rd.z, 1.0f, 0.0f);
            }
            if (cache_vertex) {
                push_cq = reflective;
            } else {
                uint32_t x, y;
                path_xy(a, pid, x, y);
                const uint4 ra = pcg4d(x, y, sample, a.seed + 2u * depth);
                const uint4 rb = pcg4d(x, y, sample, a.seed + 2u * depth + 1u);
                if (reflective && a.nlights > 0) {
                    // next-event estimation: light triangle by area, point uniform on it
                    const float total = a.cdf[a.nlights];
                    const float ut = unit_float(ra.x) * total;
                    uint32_t lo = 0, hi = a.nlights - 1;
                    while (lo < hi) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (a.cdf[mid] > ut) hi = mid; else lo = mid + 1;
                    }
                    const float* L = a.lights + 12 * (uint64_t)lo;
                    const float su = sqrtf(unit_float(ra.y));
                    const float b0 = 1.0f - su, b1 = unit_float(ra.z) * su, b2 = 1.0f - b0 - b1;
                    float lp[3], e1[3], e2[3];
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        lp[k] = b0 * L[k] + b1 * L[3 + k] + b2 * L[6 + k];
                        e1[k] = L[3 + k] - L[k];
                        e2[k] = L[6 + k] - L[k];
                    }
                    const float nl[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                                         e1[0] * e2[1] - e1[1] * e2[0]};
                    const float nll = sqrtf(nl[0] * nl[0] + nl[1] * nl[1] + nl[2] * nl[2]);
                    const float wi[3] = {lp[0] - org[0], lp[1] - org[1], lp[2] - org[2]};
                    const float d2 = wi[0] * wi[0] + wi[1] * wi[1] + wi[2] * wi[2];
                    if (nll > 0.0f && d2 > 0.0f) {
                        const float inv_d = 1.0f / sqrtf(d2);
                        const float cx = (nf[0] * wi[0] + nf[1] * wi[1] + nf[2] * wi[2]) * inv_d;
                        const float cl = -(nl[0] * wi[0] + nl[1] * wi[1] + nl[2] * wi[2]) * inv_d / nll;
                        if (cx > 0.0f && cl > 0.0f) {
                            const float g = cx * cl / d2 * total * 0.31830988618379067154f;
                            contrib.x = thr.x * alb[0] * L[9] * g;
                            contrib.y = thr.y * alb[1] * L[10] * g;
                            contrib.z = thr.z * alb[2] * L[11] * g;
                            push_sh = contrib.x > 0.0f || contrib.y > 0.0f || contrib.z > 0.0f;
                            sh.origin[0] = org[0];
                            sh.origin[1] = org[1];
                            sh.origin[2] = org[2];
                            sh.tmin = 0.0f;
                            sh.direction[0] = wi[0];
                            sh.direction[1] = wi[1];
                            sh.direction[2] = wi[2];
                            sh.tmax = 1.0f - 1e-4f;
                        }
                    }
                }
                if (reflective && depth + 1 < a.max_depth) {
                    thr_new = make_float4(thr.x * alb[0], thr.y * alb[1], thr.z * alb[2], 0.0f);
                    bool alive = true;
                    if (depth >= a.rr_depth) {
                        const float p = fminf(0.95f, fmaxf(thr_new.x, fmaxf(thr_new.y, thr_new.z)));
                        if (unit_float(ra.w) >= p) {
                            alive = false;
                        } else {
                            thr_new.x /= p;
                            thr_new.y /= p;
                            thr_new.z /= p;
                        }
                    }
                    if (alive) {
                        float dir[3];
                        cosine_dir(nf, unit_float(rb.x), unit_float(rb.y), dir);
                        ext.origin[0] = org[0];
                        ext.origin[1] = org[1];
                        ext.origin[2] = org[2];
                        ext.tmin = 0.0f;
                        ext.direction[0] = dir[0];
                        ext.direction[1] = dir[1];
                        ext.direction[2] = dir[2];
                        ext.tmax = INFINITY;
                        push_ext = true;
                    }
                }
            }
        }
    }
    const bool pr[3] = {push_sh, push_ext, push_cq};
    uint32_t o[3], tot[3];
    block_push3(&a.counts[3 * depth + 1], &a.counts[3 * (depth + 1)],
                MODE == kPtCache ? &a.counts[3 * (depth + 1) + 2] : nullptr, pr, o, tot);
    if (push_sh) {
        a.sh_rays[o[0]] = sh;
        a.sh_pid[o[0]] = pid;
        a.sh_contrib[o[0]] = contrib;
    }
    if (push_ext) {
        a.ext_rays[cur ^ 1][o[1]] = ext;
        a.ext_pid[cur ^ 1][o[1]] = pid;
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

__global__ __launch_bounds__(256) void cache_gather_kernel(PtArgs a, uint32_t depth) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.counts[3 * (depth + 1) + 2]) return;
    const uint32_t pid = a.cq_pid[q];
    const float4 t = a.thr[pid];
    float4 r = a.rad[pid];
    r.x += t.x * a.cq_pred[3 * (uint64_t)q];
    r.y += t.y * a.cq_pred[3 * (uint64_t)q + 1];
    r.z += t.z * a.cq_pred[3 * (uint64_t)q + 2];
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

__global__ void count_camera_rays_kernel(DeviceState* st, uint32_t n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) st->pt_rays += n;
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
    const uint32_t pid = (uint32_t)(r % a.npaths);
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

uint64_t pt_scratch_bytes(uint32_t npix, uint32_t max_depth, uint32_t nlights) {
    uint64_t n = npix ? npix : 1;
    uint64_t b = 0;
    b += 2 * align_up(n * sizeof(rtx_ray), 256) + 2 * align_up(n * 4, 256);  // extension queues
    b += align_up(n * sizeof(PathHit), 256);                                  // hits
    b += align_up(n * sizeof(rtx_ray), 256) + align_up(n * 4, 256) + align_up(n * 16, 256);  // shadow queue
    b += align_up(n * 4, 256);                                                // occluded
    b += 3 * align_up(n * 16, 256);                                           // thr, rad, accum
    b += align_up((3ull * (max_depth + 2)) * 4, 256);                          // counts
    b += align_up((nlights + 1ull) * 4, 256);                                 // light cdf
    b += align_up(n * sizeof(rtx_nrc_query), 256) + align_up(n * 4, 256) + align_up(n * 12, 256);  // cache queries
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
    a.cdf = (float*)take((nlights + 1ull) * 4);
    a.cq = (rtx_nrc_query*)take(n * sizeof(rtx_nrc_query));
    a.cq_pid = (uint32_t*)take(n * 4);
    a.cq_pred = (float*)take(n * 12);
}

uint64_t pt_train_bytes(uint32_t train_paths, uint32_t max_depth) {
    const uint64_t nrec = (uint64_t)train_paths * (max_depth > 1 ? max_depth - 1 : 1);
    const uint64_t nblk = (nrec + 255) / 256;
    return align_up(nrec * 80, 256) + align_up(nrec * sizeof(rtx_nrc_sample), 256) + align_up(nblk * 4, 256) + 256;
}

// One pass: vertices 0 .. last_depth.  In cache mode the cache is queried at last_depth == a.cache_depth.
static hipError_t run_pass(PtArgs& a, const TraceLaunch& L, uint32_t sample, uint32_t last_depth,
                           const NrcState* nrc, hipStream_t s) {
    const unsigned gb = (a.npaths + 255) / 256;
    const uint64_t ncounts = 3ull * (a.max_depth + 2);
    fill_words_kernel<<<1, 64, 0, s>>>(a.counts, ncounts, 0u);
    camera_kernel<<<gb, 256, 0, s>>>(a, sample);
    count_camera_rays_kernel<<<1, 64, 0, s>>>(a.state, a.npaths);
    int cur = 0;
    hipError_t e;
    for (uint32_t d = 0; d <= last_depth; ++d) {
        TraceRequest ext{};
        ext.tlas = a.tlas;
        ext.mode = 3;
        ext.n_ptr = a.counts + 3 * d;
        ext.cull_mask = 0xFFu;
        ext.rays = a.ext_rays[cur];
        ext.path_hits = a.hits;
        ext.sbt_offset = a.sbt_offset;
        ext.sbt_stride = a.sbt_stride;
        ext.state = a.state;
        ext.gstack = a.gstack;
        if ((e = trace_rays_launch(ext, L, s))) return e;
        const bool cache_here = a.mode == kPtCache && d == a.cache_depth;
        if (a.mode == kPtTrain)
            shade_kernel<kPtTrain><<<gb, 256, 0, s>>>(a, d, sample, cur);
        else if (a.mode == kPtCache)
            shade_kernel<kPtCache><<<gb, 256, 0, s>>>(a, d, sample, cur);
        else
            shade_kernel<kPtPlain><<<gb, 256, 0, s>>>(a, d, sample, cur);
        if (cache_here) {
            if ((e = nrc_infer(*nrc, a.cq, a.npaths, a.cq_pred, s, a.counts + 3 * (d + 1) + 2))) return e;
            cache_gather_kernel<<<gb, 256, 0, s>>>(a, d);
            break;
        }
        TraceRequest sh{};
        sh.tlas = a.tlas;
        sh.mode = 1;
        sh.n_ptr = a.counts + 3 * d + 1;
        sh.cull_mask = 0xFFu;
        sh.rays = a.sh_rays;
        sh.occluded = a.occl;
        sh.state = a.state;
        sh.gstack = a.gstack;
        if ((e = trace_rays_launch(sh, L, s))) return e;
        shadow_gather_kernel<<<gb, 256, 0, s>>>(a, d);
        cur ^= 1;
    }
    return hipGetLastError();
}

hipError_t pt_render(PtArgs a, const TraceLaunch& L, uint32_t spp, uint32_t sample_offset, float* film,
                     double* acc64, hipStream_t s) {
    const unsigned gb = (a.npix + 255) / 256;
    a.mode = kPtPlain;
    a.npaths = a.npix;
    if (a.nlights) light_cdf_kernel<<<1, 64, 0, s>>>(a.lights, a.nlights, a.cdf);
    fill_words_kernel<<<gb, 256, 0, s>>>((uint32_t*)a.accum, 4ull * a.npix, 0u);
    hipError_t e;
    for (uint32_t si = 0; si < spp; ++si) {
        if ((e = run_pass(a, L, sample_offset + si, a.max_depth - 1, nullptr, s))) return e;
        accumulate_kernel<<<gb, 256, 0, s>>>(a, acc64);
    }
    if (film) resolve_kernel<<<gb, 256, 0, s>>>(a, spp ? 1.0f / (float)spp : 0.0f, film);
    return hipGetLastError();
}

hipError_t pt_render_nrc(PtArgs a, const TraceLaunch& L, NrcState& nrc, uint32_t train_paths, void* train_buf,
                         uint32_t spp, uint32_t sample_offset, float* film, double* acc64, uint32_t* host_count,
                         hipStream_t s) {
    const unsigned gb = (a.npix + 255) / 256;
    hipError_t e;
    if (a.nlights) light_cdf_kernel<<<1, 64, 0, s>>>(a.lights, a.nlights, a.cdf);
    if (train_paths > 0) {
        // training paths: full length, vertex records
        PtArgs t = a;
        t.mode = kPtTrain;
        t.npaths = train_paths;
        const uint64_t nrec = (uint64_t)train_paths * (a.max_depth - 1);
        const uint64_t nblk = (nrec + 255) / 256;
        uint8_t* p = (uint8_t*)train_buf;
        t.rec = (float4*)p;
        p += align_up(nrec * 80, 256);
        rtx_nrc_sample* samples = (rtx_nrc_sample*)p;
        p += align_up(nrec * sizeof(rtx_nrc_sample), 256);
        uint32_t* block_counts = (uint32_t*)p;
        p += align_up(nblk * 4, 256);
        uint32_t* total = (uint32_t*)p;
        fill_words_kernel<<<1024, 256, 0, s>>>((uint32_t*)t.rec, nrec * 20, 0u);
        if ((e = run_pass(t, L, sample_offset, a.max_depth - 1, nullptr, s))) return e;
        sample_count_kernel<<<(unsigned)nblk, 256, 0, s>>>(t, nrec, block_counts);
        sample_scan_kernel<<<1, 1024, 0, s>>>(block_counts, (uint32_t)nblk, total);
        sample_scatter_kernel<<<(unsigned)nblk, 256, 0, s>>>(t, nrec, block_counts, samples);
        if ((e = hipMemcpyAsync(host_count, total, 4, hipMemcpyDeviceToHost, s))) return e;
        if ((e = hipStreamSynchronize(s))) return e;
        const uint32_t count = *host_count;
        if (count > 0) {
            // interleaved batches: batch b takes samples b, b + nb, b + 2 nb, ... (all depths in every batch)
            const uint32_t nb = (count + nrc.batch_size - 1) / nrc.batch_size;
            for (uint32_t b = 0; b < nb; ++b) {
                const uint32_t nbatch = (count - b + nb - 1) / nb;
                if ((e = nrc_step(nrc, samples, nbatch, nullptr, nullptr, s, nb, b))) return e;
            }
        }
    }
    // render paths terminated into the cache
    a.mode = kPtCache;
    a.npaths = a.npix;
    fill_words_kernel<<<gb, 256, 0, s>>>((uint32_t*)a.accum, 4ull * a.npix, 0u);
    for (uint32_t si = 0; si < spp; ++si) {
        if ((e = run_pass(a, L, sample_offset + si, a.cache_depth, &nrc, s))) return e;
        accumulate_kernel<<<gb, 256, 0, s>>>(a, acc64);
    }
    if (film) resolve_kernel<<<gb, 256, 0, s>>>(a, spp ? 1.0f / (float)spp : 0.0f, film);
    return hipGetLastError();
}

}  // namespace rtx
