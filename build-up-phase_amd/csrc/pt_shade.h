// pt_shade.h — per-path steps of the path tracer shared by the wavefront kernels (pathtrace.hip) and the
// training-path traversal (trace.hip): camera rays and the shading of one path vertex.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pathtrace.h"
#include "rtx_internal.h"
#include "sampling.h"

namespace rtx {

constexpr uint32_t kTrainKey = 0x5EEDu;

__device__ __forceinline__ uint32_t view_of(const PtArgs& a, uint32_t pid) {
    uint32_t lo = 0, hi = a.nviews - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (a.views[mid].pix_base <= pid) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Distance along d from o (inside the box) to the exit of the view's clip box.
__device__ __forceinline__ float clip_exit(const BatchView& v, const float* o, const float* d) {
    float t = INFINITY;
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (d[k] != 0.0f) t = fminf(t, ((d[k] > 0.0f ? v.hi[k] : v.lo[k]) - o[k]) / d[k]);
    return fmaxf(t, 0.0f);
}

__device__ __forceinline__ uint32_t pass_sample(const PtArgs& a, uint32_t sample) {
    return a.sample_dev ? sample + *a.sample_dev : sample;
}

__device__ __forceinline__ void path_xy(const PtArgs& a, uint32_t pid, uint32_t& x, uint32_t& y) {
    if (a.views) {  // batched render: pixel of the path's view
        const BatchView& v = a.views[a.path_view[pid]];
        const uint32_t local = pid - v.pix_base;
        x = local % v.width;
        y = local / v.width;
    } else if (pid >= a.npix) {  // training path
        x = pid - a.npix;
        y = 0x7FFFFFFFu;
    } else {
        x = pid % a.width;
        y = pid / a.width;
    }
}

// Camera ray of path i: pixel paths jitter within their pixel (of their view, in a batched render: the path's
// view is looked up and stored), training paths start at a random pixel.
__device__ __forceinline__ void pt_camera_ray(const PtArgs& a, uint32_t i, uint32_t sample, rtx_ray& r) {
    uint32_t px, py;
    uint4 j;
    if (a.views) {  // batched render
        const uint32_t vi = view_of(a, i);
        a.path_view[i] = vi;
        const BatchView& v = a.views[vi];
        const uint32_t local = i - v.pix_base;
        px = local % v.width;
        py = local / v.width;
        j = pcg4d(px, py, sample + v.sample_offset, a.seed ^ kJitterKey);
        const float sx = ((float)px + unit_float(j.x)) / (float)v.width;
        const float sy = ((float)py + unit_float(j.y)) / (float)v.height;
        const float nx = sx * 2.0f - 1.0f;
        const float ny = sy * 2.0f - 1.0f;
        r.origin[0] = v.cam[0];
        r.origin[1] = v.cam[1];
        r.origin[2] = v.cam[2];
        r.tmin = 0.0f;
        r.direction[0] = nx * v.ax;
        r.direction[1] = -(ny * v.ay);
        r.direction[2] = -1.0f;
        r.tmax = clip_exit(v, r.origin, r.direction);
        return;
    }
    if (i >= a.npix) {  // training path: a random pixel
        const uint4 h = pcg4d(i - a.npix, sample, a.seed, kTrainKey);
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
    r.origin[0] = a.cam[0];
    r.origin[1] = a.cam[1];
    r.origin[2] = a.cam[2];
    r.tmin = 0.0f;
    r.direction[0] = nx * a.ax;
    r.direction[1] = -(ny * a.ay);
    r.direction[2] = -1.0f;
    r.tmax = INFINITY;
}

// Result of shading one path vertex: the shadow ray toward a light point with its contribution, the continuation
// ray with its new throughput, or (cache mode) a cache query.
struct ShadeOut {
    bool push_ext = false, push_sh = false, push_cq = false;
    rtx_ray ext{}, sh{};
    rtx_nrc_query cq{};
    float4 contrib = make_float4(0, 0, 0, 0);
    float4 thr_new = make_float4(0, 0, 0, 0);
};

// Shades vertex `depth` of path pid, reached by ray r with hit h: adds the emission seen at depth 0 to the path's
// radiance, writes a training path's vertex record, and returns the next-event shadow ray, the continuation and
// its throughput, or the cache query (the caller queues them).
template <int MODE>
__device__ __forceinline__ void shade_vertex(const PtArgs& a, uint32_t pid, const rtx_ray& r, const PathHit& h,
                                             uint32_t depth, uint32_t sample, ShadeOut& o) {
    const bool train = MODE == kPtCache && pid >= a.npix;
    // the vertex where this path queries the cache instead of scattering
    const bool query_here = MODE == kPtCache && ((!train && depth == a.cache_depth) ||
                                                 (train && a.tail && depth == a.train_end));
    const bool record = train && depth >= 1 && !query_here;
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
        if (query_here || record) {
            const float il = 1.0f / sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                o.cq.position[k] = P[k];
                o.cq.normal[k] = nf[k];
                o.cq.direction[k] = -d[k] * il;
                o.cq.albedo[k] = alb[k];
            }
        }
        if (record) {
            // record the vertex before its own next-event contribution is gathered
            const float4 rd = a.rad[pid];
            float4* rp = a.rec + ((uint64_t)(depth - 1) * a.ntrain + (pid - a.npix)) * 5;
            rp[0] = make_float4(o.cq.position[0], o.cq.position[1], o.cq.position[2], o.cq.normal[0]);
            rp[1] = make_float4(o.cq.normal[1], o.cq.normal[2], o.cq.direction[0], o.cq.direction[1]);
            rp[2] = make_float4(o.cq.direction[2], o.cq.albedo[0], o.cq.albedo[1], o.cq.albedo[2]);
            rp[3] = make_float4(thr.x, thr.y, thr.z, rd.x);
            rp[4] = make_float4(rd.y, rd.z, 1.0f, 0.0f);
        }
        if (query_here) {
            o.push_cq = reflective;
        } else {
            uint32_t x, y;
            path_xy(a, pid, x, y);
            const BatchView* bv = a.views ? a.views + a.path_view[pid] : nullptr;
            const uint32_t smp = bv ? sample + bv->sample_offset : sample;
            const uint4 ra = pcg4d(x, y, smp, a.seed + 2u * depth);
            const uint4 rb = pcg4d(x, y, smp, a.seed + 2u * depth + 1u);
            const uint32_t nl = bv ? bv->light_count : a.nlights;
            const float* cdf = bv ? a.cdf + bv->cdf_first : a.cdf;
            const float* lights = bv ? a.lights + 12 * (uint64_t)bv->light_first : a.lights;
            if (reflective && nl > 0) {
                // next-event estimation: light triangle by area, point uniform on it
                const float total = cdf[nl];
                const float ut = unit_float(ra.x) * total;
                uint32_t lo = 0, hi = nl - 1;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (cdf[mid] > ut) hi = mid; else lo = mid + 1;
                }
                const float* L = lights + 12 * (uint64_t)lo;
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
                        o.contrib.x = thr.x * alb[0] * L[9] * g;
                        o.contrib.y = thr.y * alb[1] * L[10] * g;
                        o.contrib.z = thr.z * alb[2] * L[11] * g;
                        o.push_sh = o.contrib.x > 0.0f || o.contrib.y > 0.0f || o.contrib.z > 0.0f;
                        o.sh.origin[0] = org[0];
                        o.sh.origin[1] = org[1];
                        o.sh.origin[2] = org[2];
                        o.sh.tmin = 0.0f;
                        o.sh.direction[0] = wi[0];
                        o.sh.direction[1] = wi[1];
                        o.sh.direction[2] = wi[2];
                        o.sh.tmax = 1.0f - 1e-4f;
                    }
                }
            }
            if (reflective && depth + 1 < a.max_depth && (!train || !a.tail || depth < a.train_end)) {
                o.thr_new = make_float4(thr.x * alb[0], thr.y * alb[1], thr.z * alb[2], 0.0f);
                bool alive = true;
                if (depth >= a.rr_depth) {
                    const float p = fminf(0.95f, fmaxf(o.thr_new.x, fmaxf(o.thr_new.y, o.thr_new.z)));
                    if (unit_float(ra.w) >= p) {
                        alive = false;
                    } else {
                        o.thr_new.x /= p;
                        o.thr_new.y /= p;
                        o.thr_new.z /= p;
                    }
                }
                if (alive) {
                    float dir[3];
                    cosine_dir(nf, unit_float(rb.x), unit_float(rb.y), dir);
                    o.ext.origin[0] = org[0];
                    o.ext.origin[1] = org[1];
                    o.ext.origin[2] = org[2];
                    o.ext.tmin = 0.0f;
                    o.ext.direction[0] = dir[0];
                    o.ext.direction[1] = dir[1];
                    o.ext.direction[2] = dir[2];
                    o.ext.tmax = bv ? clip_exit(*bv, o.ext.origin, o.ext.direction) : INFINITY;
                    o.push_ext = true;
                }
            }
        }
    }
}

}  // namespace rtx
