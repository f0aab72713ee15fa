// build.hip — LBVH builder for BLAS (triangles) and TLAS (instances) on gfx950.
//
// Replaces the driver build behind vkCmdBuildAccelerationStructuresKHR (main.cpp:808 for the BLAS,
// main.cpp:932 for the TLAS).  Pipeline, all on the context stream:
//   prep      one thread per primitive: fetch + transform triangle vertices (BLAS) or read the 64-B instance
//             record, invert its transform in FP64 and bound the BLAS root box in world space (TLAS);
//             writes a 64-B leaf record, the primitive box, and the centroid bounds (ordered-int atomics)
//   morton    30-bit (uint32) or 63-bit (uint64) Morton code of each box centroid
//   sort      stable onesweep radix sort of (code, primitive index)                         [sort.hip]
//   karras    binary radix tree over the sorted codes (Karras 2012), one thread per internal node
//   refit     bottom-up boxes: the second thread to reach a node merges its children (atomic arrival count)
//   collapse  level-synchronous top-down conversion to 8-wide nodes: each node adopts up to 8 children by
//             repeatedly opening the child with the largest surface area, places them by octant, quantises
//             their boxes to 8 bits per plane and allocates their slots with one atomic add.
// Karras trees over 30-bit codes with 32-bit index tie-breaks are at most 64 levels deep (63-bit codes: 96),
// so the host enqueues that many collapse levels; a level whose frontier is empty returns at once.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "build.h"
#include "rtx_internal.h"
#include "sort.h"

namespace rtx {

namespace {

constexpr int kCollapseBlocks = 1024;
constexpr int kMaxLevels32 = 66;
constexpr int kMaxLevels64 = 98;

struct Box {
    float lo[3], hi[3];
};

__device__ __forceinline__ float box_area(const float* b) {
    float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    if (!(dx >= 0.0f) || !(dy >= 0.0f) || !(dz >= 0.0f)) return 0.0f;
    return dx * dy + dy * dz + dz * dx;
}

// Geometry transform at build time: x' = ((m00*x + m01*y) + m02*z) + m03 with separate roundings
// (this file is compiled with -ffp-contract=off).
__device__ __forceinline__ void xform_vertex(const float* m, const float* v, float* o) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        float a = m[4 * r + 0] * v[0];
        float b = m[4 * r + 1] * v[1];
        float c = m[4 * r + 2] * v[2];
        float s = a + b;
        s = s + c;
        s = s + m[4 * r + 3];
        o[r] = s;
    }
}

__device__ void reduce_bounds(const float* c, bool valid, uint32_t* cb) {
    // block-wide min/max of centroids through LDS, then one atomic per block and axis
    __shared__ uint32_t s[6];
    if (threadIdx.x < 6) s[threadIdx.x] = (threadIdx.x < 3) ? 0xFFFFFFFFu : 0u;
    __syncthreads();
    if (valid) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            uint32_t o = float_to_ordered(c[a]);
            atomicMin(&s[a], o);
            atomicMax(&s[3 + a], o);
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        if (s[threadIdx.x] != 0xFFFFFFFFu) atomicMin(&cb[threadIdx.x], s[threadIdx.x]);
    } else if (threadIdx.x < 6) {
        if (s[threadIdx.x] != 0u) atomicMax(&cb[threadIdx.x], s[threadIdx.x]);
    }
}

// Batched builds: per-BLAS centroid bounds (6 words per BLAS).  Threads of the block's first BLAS reduce through
// LDS; the few threads of a block that straddles a BLAS boundary use global atomics directly.
__device__ void reduce_bounds_seg(const float* c, bool valid, uint32_t blas, uint32_t* cb) {
    __shared__ uint32_t s[6];
    __shared__ uint32_t sb;
    if (threadIdx.x == 0) sb = blas;
    if (threadIdx.x < 6) s[threadIdx.x] = (threadIdx.x < 3) ? 0xFFFFFFFFu : 0u;
    __syncthreads();
    const uint32_t b0 = sb;
    if (valid) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const uint32_t o = float_to_ordered(c[a]);
            if (blas == b0) {
                atomicMin(&s[a], o);
                atomicMax(&s[3 + a], o);
            } else {
                atomicMin(&cb[6 * (uint64_t)blas + a], o);
                atomicMax(&cb[6 * (uint64_t)blas + 3 + a], o);
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        if (s[threadIdx.x] != 0xFFFFFFFFu) atomicMin(&cb[6 * (uint64_t)b0 + threadIdx.x], s[threadIdx.x]);
    } else if (threadIdx.x < 6) {
        if (s[threadIdx.x] != 0u) atomicMax(&cb[6 * (uint64_t)b0 + threadIdx.x], s[threadIdx.x]);
    }
}

// prim_blas != nullptr: batched build (per-primitive BLAS id out, per-BLAS centroid bounds).
__global__ __launch_bounds__(256) void prep_tris_kernel(const GeomDesc* __restrict__ geoms, int ngeom, uint32_t n,
                                                        TriSlot* __restrict__ leaves, float* __restrict__ boxes,
                                                        uint32_t* __restrict__ cbounds,
                                                        uint32_t* __restrict__ prim_blas) {
    uint32_t blas = 0;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = false;
    float cen[3] = {0, 0, 0};
    if (i < n) {
        int lo = 0, hi = ngeom - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (geoms[mid].prim_base <= i) lo = mid; else hi = mid - 1;
        }
        const GeomDesc g = geoms[lo];
        uint32_t p = i - g.prim_base;
        blas = g.blas;
        if (prim_blas) prim_blas[i] = blas;
        TriSlot t;
        valid = true;
        if (g.type == RTX_GEOMETRY_TYPE_AABBS) {
            // VkAabbPositionsKHR: min xyz, max xyz; inactive when NaN or min > max
            const float* src = (const float*)(g.vtx + (uint64_t)p * g.stride);
#pragma unroll
            for (int k = 0; k < 6; ++k) t.v[k] = src[k];
            t.v[6] = t.v[7] = t.v[8] = 0.0f;
#pragma unroll
            for (int a = 0; a < 3; ++a) valid = valid && isfinite(t.v[a]) && isfinite(t.v[3 + a]) && t.v[a] <= t.v[3 + a];
        } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            uint32_t vi = g.idx[3 * (uint64_t)p + k];
            float v[3];
            if (vi > g.max_vertex) {
                valid = false;
                v[0] = v[1] = v[2] = 0.0f;
            } else {
                const float* src = (const float*)(g.vtx + (uint64_t)(vi + g.first_vertex) * g.stride);
                v[0] = src[0];
                v[1] = src[1];
                v[2] = src[2];
            }
            if (g.xform) {
                float o[3];
                xform_vertex(g.xform, v, o);
                t.v[3 * k] = o[0];
                t.v[3 * k + 1] = o[1];
                t.v[3 * k + 2] = o[2];
            } else {
                t.v[3 * k] = v[0];
                t.v[3 * k + 1] = v[1];
                t.v[3 * k + 2] = v[2];
            }
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) valid = valid && isfinite(t.v[k]);
        }
        t.prim = p;
        t.geom = g.geom_local;
        t.bits = (valid ? kTriValid : 0u) | ((g.flags & RTX_GEOMETRY_OPAQUE_BIT) ? 0u : kTriNonOpaque) |
                 (g.type == RTX_GEOMETRY_TYPE_AABBS ? kTriAabb : 0u);
        t.pad[0] = t.pad[1] = t.pad[2] = t.pad[3] = 0u;
        leaves[i] = t;
        float* b = boxes + 6 * (uint64_t)i;
        if (valid) {
            const bool aabb = g.type == RTX_GEOMETRY_TYPE_AABBS;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                float l = aabb ? t.v[a] : fminf(t.v[a], fminf(t.v[3 + a], t.v[6 + a]));
                float h = aabb ? t.v[3 + a] : fmaxf(t.v[a], fmaxf(t.v[3 + a], t.v[6 + a]));
                b[a] = l;
                b[3 + a] = h;
                float s = l + h;
                cen[a] = s * 0.5f;
            }
        } else {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                b[a] = INFINITY;
                b[3 + a] = -INFINITY;
            }
        }
    }
    if (prim_blas)
        reduce_bounds_seg(cen, valid, blas, cbounds);
    else
        reduce_bounds(cen, valid, cbounds);
}

__device__ void invert3x4_d(const double* m, double* inv) {
    double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], i = m[10];
    double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    double det = a * A + b * B + c * C;
    double r = 1.0 / det;
    inv[0] = A * r;
    inv[1] = -(b * i - c * h) * r;
    inv[2] = (b * f - c * e) * r;
    inv[4] = B * r;
    inv[5] = (a * i - c * g) * r;
    inv[6] = -(a * f - c * d) * r;
    inv[8] = C * r;
    inv[9] = -(a * h - b * g) * r;
    inv[10] = (a * e - b * d) * r;
#pragma unroll
    for (int row = 0; row < 3; ++row)
        inv[4 * row + 3] = -(inv[4 * row + 0] * m[3] + inv[4 * row + 1] * m[7] + inv[4 * row + 2] * m[11]);
}

__device__ __forceinline__ float round_down_f(double x) {
    float f = (float)x;
    if ((double)f > x) f = nextafterf(f, -INFINITY);
    return f;
}
__device__ __forceinline__ float round_up_f(double x) {
    float f = (float)x;
    if ((double)f < x) f = nextafterf(f, INFINITY);
    return f;
}

// World box of an instance: the union of its BLAS's second-level boxes (the root's leaf children, and the
// children of its internal children), each decoded from the 8-bit planes with the FP32 operations the traversal
// uses (so it encloses what they enclose) and transformed corner by corner in FP64.  For a rotated instance this
// is much tighter than the transformed root box, and every ray that misses it skips the instance's FP64 entry.
// A BLAS whose nodes do not check out falls back to the root box.
__device__ void add_world_box(const double* m, const float* blo, const float* bhi, double* lo, double* hi) {
    for (int c = 0; c < 8; ++c) {
        const double p[3] = {(double)((c & 1) ? bhi[0] : blo[0]), (double)((c & 2) ? bhi[1] : blo[1]),
                             (double)((c & 4) ? bhi[2] : blo[2])};
        for (int a = 0; a < 3; ++a) {
            const double w = m[4 * a] * p[0] + m[4 * a + 1] * p[1] + m[4 * a + 2] * p[2] + m[4 * a + 3];
            lo[a] = fmin(lo[a], w);
            hi[a] = fmax(hi[a], w);
        }
    }
}

__device__ void decode_child_box(const Node8* nd, int k, const float* rlo, const float* qs, float* blo, float* bhi) {
    for (int a = 0; a < 3; ++a) {
        const float p = fmaf((float)nd->pq[a], qs[a], rlo[a]);
        const float sc = __uint_as_float((uint32_t)nd->e[a] << 23);
        blo[a] = fmaf((float)nd->qlo[a][k], sc, p);
        bhi[a] = fmaf((float)nd->qhi[a][k], sc, p);
    }
}

__device__ void instance_world_box(const AsHeader* h, const double* m, double* lo, double* hi) {
    const uint8_t* slots = (const uint8_t*)h + kHeaderBytes;
    const Node8* root = (const Node8*)slots;
    const float rlo[3] = {h->lo[0], h->lo[1], h->lo[2]}, qs[3] = {h->qscale[0], h->qscale[1], h->qscale[2]};
    const uint32_t nslots = h->slot_count;
    bool ok = root->vmask != 0 && root->base < nslots;
    for (int k = 0; k < 8 && ok; ++k) {
        if (!((root->vmask >> k) & 1u) || !((root->imask >> k) & 1u)) continue;
        const uint32_t cs = root->base + (uint32_t)__popc(root->vmask & ((1u << k) - 1u));
        ok = cs < nslots && ((const Node8*)(slots + (uint64_t)cs * kSlotBytes))->vmask != 0;
    }
    if (!ok) {
        add_world_box(m, h->lo, h->hi, lo, hi);
        return;
    }
    for (int k = 0; k < 8; ++k) {
        if (!((root->vmask >> k) & 1u)) continue;
        float blo[3], bhi[3];
        if ((root->imask >> k) & 1u) {
            const uint32_t cs = root->base + (uint32_t)__popc(root->vmask & ((1u << k) - 1u));
            const Node8* c = (const Node8*)(slots + (uint64_t)cs * kSlotBytes);
            for (int j = 0; j < 8; ++j) {
                if (!((c->vmask >> j) & 1u)) continue;
                decode_child_box(c, j, rlo, qs, blo, bhi);
                add_world_box(m, blo, bhi, lo, hi);
            }
        } else {
            decode_child_box(root, k, rlo, qs, blo, bhi);
            add_world_box(m, blo, bhi, lo, hi);
        }
    }
}

__global__ __launch_bounds__(256) void prep_instances_kernel(const rtx_instance* __restrict__ inst, uint32_t n,
                                                             InstSlot* __restrict__ leaves,
                                                             InstanceInfo* __restrict__ table,
                                                             float* __restrict__ boxes,
                                                             uint32_t* __restrict__ cbounds,
                                                             uint32_t* __restrict__ max_blas_depth,
                                                             uint32_t* __restrict__ content_out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = false;
    uint32_t content = 0;
    float cen[3] = {0, 0, 0};
    if (i < n) {
        rtx_instance r = inst[i];
        double m[12];
        for (int k = 0; k < 12; ++k) m[k] = (double)(&r.transform[0][0])[k];
        double inv[12];
        invert3x4_d(m, inv);
        InstanceInfo info;
        for (int k = 0; k < 12; ++k) info.inv[k] = inv[k];
        info.custom = r.custom_index_and_mask & 0xFFFFFFu;
        info.mask = r.custom_index_and_mask >> 24;
        info.sbt_offset = r.sbt_offset_and_flags & 0xFFFFFFu;
        info.flags = r.sbt_offset_and_flags >> 24;
        info.blas = r.blas_address;
        info.pad = 0;
        bool ok = r.blas_address != 0 && (r.blas_address & 255u) == 0 && isfinite(inv[0]) && isfinite(inv[5]) &&
                  isfinite(inv[10]);
        InstSlot s;
        for (int k = 0; k < 12; ++k) s.inv[k] = (float)inv[k];
        s.blas = r.blas_address;
        s.instance_id = i;
        s.mask = info.mask | (info.flags << 8);
        float* b = boxes + 6 * (uint64_t)i;
        if (ok) {
            const AsHeader* h = (const AsHeader*)r.blas_address;
            ok = h->magic == kMagicBlas && h->prim_count > 0 && h->slot_count > 0;
            if (ok) {
                double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
                instance_world_box(h, m, lo, hi);
                double ext = 0;
                for (int a = 0; a < 3; ++a) ext = fmax(ext, fmax(fabs(lo[a]), fabs(hi[a])) + (hi[a] - lo[a]));
                double pad = ext * 0x1p-20;
                for (int a = 0; a < 3; ++a) {
                    b[a] = round_down_f(lo[a] - pad);
                    b[3 + a] = round_up_f(hi[a] + pad);
                    float s2 = b[a] + b[3 + a];
                    cen[a] = s2 * 0.5f;
                }
                atomicMax(max_blas_depth, h->max_depth);
                content = h->content | ((info.flags & RTX_INSTANCE_FORCE_NO_OPAQUE_BIT) ? kContentNonOpaque : 0u);
                valid = true;
            }
        }
        if (!valid) {
            for (int a = 0; a < 3; ++a) {
                b[a] = INFINITY;
                b[3 + a] = -INFINITY;
            }
        }
        leaves[i] = s;
        table[i] = info;
    }
    // one atomic per wave for the content bits
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) content |= __shfl_xor(content, off);
    if ((threadIdx.x & 63u) == 0 && content) atomicOr(content_out, content);
    reduce_bounds(cen, valid, cbounds);
}

__global__ __launch_bounds__(256) void box_bounds_kernel(const float* __restrict__ boxes, uint32_t n,
                                                         uint32_t* __restrict__ cbounds) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float cen[3] = {0, 0, 0};
    bool valid = false;
    if (i < n) {
        const float* b = boxes + 6 * (uint64_t)i;
        valid = b[0] <= b[3] && b[1] <= b[4] && b[2] <= b[5];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float s = b[a] + b[3 + a];
            cen[a] = s * 0.5f;
        }
    }
    reduce_bounds(cen, valid, cbounds);
}

__device__ __forceinline__ uint32_t quant_axis(float c, float cmin, float scale, uint32_t maxq) {
    float f = (c - cmin) * scale;
    if (!(f > 0.0f)) return 0u;
    if (f >= (float)maxq) return maxq;
    return (uint32_t)f;
}

template <typename K>
__global__ __launch_bounds__(256) void morton_kernel(const float* __restrict__ boxes, uint32_t n,
                                                     const uint32_t* __restrict__ cbounds, K* __restrict__ keys) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    constexpr bool k64 = sizeof(K) == 8;
    const float cells = k64 ? 2097152.0f : 1024.0f;
    const uint32_t maxq = k64 ? 2097151u : 1023u;
    const float* b = boxes + 6 * (uint64_t)i;
    uint32_t q[3];
    bool valid = b[0] <= b[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float cmin = ordered_to_float(cbounds[a]);
        float cmax = ordered_to_float(cbounds[3 + a]);
        float ext = cmax - cmin;
        float scale = ext > 0.0f ? cells / ext : 0.0f;
        float s = b[a] + b[3 + a];
        float c = s * 0.5f;
        q[a] = quant_axis(c, cmin, scale, maxq);
    }
    K code;
    if (k64)
        code = (K)((spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]));
    else
        code = (K)((spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]));
    keys[i] = valid ? code : (K)~(K)0;
}

// Batched builds: key = BLAS id << 30 | 30-bit Morton code over that BLAS's centroid bounds, so that a stable
// sort groups the BLASes and every BLAS becomes one subtree of the binary radix tree.
__global__ __launch_bounds__(256) void morton_batch_kernel(const float* __restrict__ boxes, uint32_t n,
                                                           const uint32_t* __restrict__ cbounds,
                                                           const uint32_t* __restrict__ prim_blas,
                                                           uint64_t* __restrict__ keys) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t blas = prim_blas[i];
    const uint32_t* cb = cbounds + 6 * (uint64_t)blas;
    const float* b = boxes + 6 * (uint64_t)i;
    uint32_t q[3];
    const bool valid = b[0] <= b[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float cmin = ordered_to_float(cb[a]);
        float cmax = ordered_to_float(cb[3 + a]);
        float ext = cmax - cmin;
        float scale = ext > 0.0f ? 1024.0f / ext : 0.0f;
        float s = b[a] + b[3 + a];
        float c = s * 0.5f;
        q[a] = quant_axis(c, cmin, scale, 1023u);
    }
    const uint32_t code = valid ? ((spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2])) : 0x3FFFFFFFu;
    keys[i] = ((uint64_t)blas << 30) | code;
}

constexpr uint32_t kMixedBlas = 0xFFFFFFFFu;

template <typename K>
__device__ __forceinline__ int karras_delta(const K* keys, int64_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= n) return -1;
    K a = keys[i], b = keys[j];
    constexpr int kb = (int)sizeof(K) * 8;
    if (a == b) return kb + __clz((uint32_t)(i ^ j));
    if (sizeof(K) == 8) return __clzll((unsigned long long)(a ^ b));
    return __clz((uint32_t)(a ^ b));
}

// node_blas != nullptr (batched builds): the BLAS that owns all primitives of internal node i, or kMixedBlas.
template <typename K>
__global__ __launch_bounds__(256) void karras_kernel(const K* __restrict__ keys, uint32_t n, int32_t* __restrict__ left,
                                                     int32_t* __restrict__ right, int32_t* __restrict__ parent,
                                                     uint32_t* __restrict__ node_blas = nullptr) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i == 0) parent[0] = -1;
    if (i >= (int64_t)n - 1) return;
    int64_t N = n;
    int d = (karras_delta(keys, N, i, i + 1) - karras_delta(keys, N, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = karras_delta(keys, N, i, i - d);
    int64_t lmax = 2;
    while (karras_delta(keys, N, i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
        if (karras_delta(keys, N, i, i + (l + t) * d) > dmin) l += t;
    int64_t j = i + l * d;
    int dnode = karras_delta(keys, N, i, j);
    int64_t s = 0;
    int64_t t = l;
    do {
        t = (t + 1) / 2;
        if (s + t <= l && karras_delta(keys, N, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int64_t gamma = i + s * d + (d < 0 ? -1 : 0);
    int64_t lo = i < j ? i : j, hi = i < j ? j : i;
    int32_t lc = (lo == gamma) ? (int32_t)(-1 - gamma) : (int32_t)gamma;
    int32_t rc = (hi == gamma + 1) ? (int32_t)(-1 - (gamma + 1)) : (int32_t)(gamma + 1);
    left[i] = lc;
    right[i] = rc;
    parent[lc >= 0 ? lc : (N - 1) + (-1 - lc)] = (int32_t)i;
    parent[rc >= 0 ? rc : (N - 1) + (-1 - rc)] = (int32_t)i;
    if (node_blas) {
        const uint64_t blo = (uint64_t)keys[lo] >> 30, bhi = (uint64_t)keys[hi] >> 30;
        node_blas[i] = blo == bhi ? (uint32_t)blo : kMixedBlas;
    }
}

// Batched builds: the root of BLAS b is its topmost node whose primitives all belong to b (a leaf for a
// one-triangle BLAS).  Runs over the 2n-1 nodes before the refit; the refinement never moves these roots.
__global__ __launch_bounds__(256) void blas_roots_kernel(uint32_t n, const uint64_t* __restrict__ keys,
                                                         const int32_t* __restrict__ parent,
                                                         const uint32_t* __restrict__ node_blas,
                                                         int32_t* __restrict__ roots) {
    const uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (k >= 2ull * n - 1) return;
    uint32_t b;
    int32_t code;
    if (k < (uint64_t)n - 1) {
        b = node_blas[k];
        code = (int32_t)k;
    } else {
        const uint64_t pos = k - (n - 1);
        b = (uint32_t)(keys[pos] >> 30);
        code = -1 - (int32_t)pos;
    }
    if (b == kMixedBlas) return;
    const int32_t p = parent[k];
    if (p < 0 || node_blas[p] == kMixedBlas) roots[b] = code;
}

__device__ __forceinline__ float load_coherent(const float* p) {
    return __uint_as_float(__hip_atomic_load((const uint32_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
__device__ __forceinline__ void store_coherent(float* p, float v) {
    __hip_atomic_store((uint32_t*)p, __float_as_uint(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t load_coherent_u32(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_coherent_u32(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int32_t load_coherent_i32(const int32_t* p) {
    return (int32_t)load_coherent_u32((const uint32_t*)p);
}
__device__ __forceinline__ void store_coherent_i32(int32_t* p, int32_t v) { store_coherent_u32((uint32_t*)p, (uint32_t)v); }

// ---- bottom-up refit with SAH treelet refinement ------------------------------------------------------------
//
// Node numbering: internal nodes 0..n-2, leaves n-1+pos (sorted position); a child code c >= 0 is internal node
// c, c < 0 the leaf at position -1-c.  Per internal node the pass keeps the box, the SAH cost of its subtree
// (sum of internal-node surface areas, the only term a restructuring can change), the primitive count and the
// height (leaves: cost 0, count 1, height 0).  The node cost is area + (cost(left) + cost(right)).
//
// The second thread to reach a node (atomic arrival count) merges its children.  When a node with at least
// min_prims primitives is complete, the whole wave restructures the treelet below it (Karras & Aila 2013,
// "Fast Parallel Construction of High-Quality Bounding Volume Hierarchies"):
//   formation  start from the node's two children and open the internal treelet leaf with the largest surface
//              area (lowest slot on ties) until there are 7 leaves; the node and the 5 opened nodes are the
//              treelet's 6 internal nodes
//   DP         over the 127 leaf subsets S in order of size: cost(S) = area(S) + min over splits {P, S\P} of
//              (cost(P) + cost(S\P)), P containing the lowest leaf of S; ties go to the smallest P
//   rebuild    only on a strict cost decrease and when the treelet height does not grow (so the whole tree
//              keeps the depth bound the collapse relies on): breadth-first from S = all leaves, internal
//              nodes reused in formation order.
// A treelet covers only nodes of its root's finished subtree and is processed after that subtree is final, so
// the result does not depend on the order in which waves finish.  Cross-wave hand-off uses agent-scope atomic
// loads and stores (coherent across XCDs without L2 write-back fences): a wave waits for its stores to
// complete (s_waitcnt) before it bumps a parent's arrival count.
constexpr int kTreeletLeaves = 7;
constexpr int kTreeletSubsets = 127;

// Subsets of the 7 treelet leaves ordered by size, then value (size k starts at entry 0, 7, 28, 63, 98, 119, 126
// for k = 1..7).
__constant__ uint8_t kSubsetsBySize[127] = {
    1, 2, 4, 8, 16, 32, 64, 3, 5, 6, 9, 10, 12, 17, 18, 20, 24, 33, 34, 36, 40, 48, 65, 66, 68, 72, 80, 96, 7, 11,
    13, 14, 19, 21, 22, 25, 26, 28, 35, 37, 38, 41, 42, 44, 49, 50, 52, 56, 67, 69, 70, 73, 74, 76, 81, 82, 84, 88,
    97, 98, 100, 104, 112, 15, 23, 27, 29, 30, 39, 43, 45, 46, 51, 53, 54, 57, 58, 60, 71, 75, 77, 78, 83, 85, 86,
    89, 90, 92, 99, 101, 102, 105, 106, 108, 113, 114, 116, 120, 31, 47, 55, 59, 61, 62, 79, 87, 91, 93, 94, 103,
    107, 109, 110, 115, 117, 118, 121, 122, 124, 63, 95, 111, 119, 123, 125, 126, 127};

struct TreeletLds {
    float area[128];
    float cost[128];
    uint32_t part[128];
    float box[kTreeletLeaves][6];
    float lcost[kTreeletLeaves];
    uint32_t lcnt[kTreeletLeaves], lhgt[kTreeletLeaves];
    int32_t leaf[kTreeletLeaves];
    int32_t pool[kTreeletLeaves - 1];
    // rebuild list: subset, node, child codes, child list index (-1: treelet leaf), height
    uint32_t ls[kTreeletLeaves - 1];
    int32_t lx[kTreeletLeaves - 1], lc[kTreeletLeaves - 1][2], li[kTreeletLeaves - 1][2];
    uint32_t lh[kTreeletLeaves - 1];
    uint32_t commit;
};

__device__ __forceinline__ uint64_t node_index(uint32_t n, int32_t c) {
    return c >= 0 ? (uint64_t)c : (uint64_t)n - 1 + (uint64_t)(-1 - c);
}

__device__ __forceinline__ float coherent_area(const float* nboxes, uint64_t i) {
    float b[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) b[k] = load_coherent(nboxes + 6 * i + k);
    return box_area(b);
}

// Treelet formation for root R, by one lane: open the largest-area internal treelet leaf until there are 7
// leaves.  Writes the leaves and the 6 internal nodes (R first, in opening order) to the lane's staging row;
// false when fewer than 7 primitives lie below R (nothing to restructure).  The lanes of a wave form their
// treelets at the same time: their roots were completed in the same step of the climb, so none is an ancestor
// of another and their subtrees - everything a formation reads and a restructuring writes - are disjoint.
struct TreeletForm {
    int32_t leaf[kTreeletLeaves];
    int32_t pool[kTreeletLeaves - 1];
};

__device__ bool treelet_form(int32_t R, const int32_t* left, const int32_t* right, const float* nboxes,
                             TreeletForm& out) {
    int32_t lf[kTreeletLeaves];
    float la[kTreeletLeaves];
    int32_t pool[kTreeletLeaves - 1];
    lf[0] = load_coherent_i32(left + R);
    lf[1] = load_coherent_i32(right + R);
    la[0] = lf[0] >= 0 ? coherent_area(nboxes, (uint64_t)lf[0]) : 0.0f;
    la[1] = lf[1] >= 0 ? coherent_area(nboxes, (uint64_t)lf[1]) : 0.0f;
    pool[0] = R;
#pragma unroll
    for (int k = 2; k < kTreeletLeaves; ++k) {
        int best = -1;
        float ba = -1.0f;
#pragma unroll
        for (int j = 0; j < k; ++j)
            if (lf[j] >= 0 && la[j] > ba) {
                ba = la[j];
                best = j;
            }
        if (best < 0) return false;
        int32_t c = 0;
#pragma unroll
        for (int j = 0; j < k; ++j)
            if (j == best) c = lf[j];
        pool[k - 1] = c;
        const int32_t l = load_coherent_i32(left + c), r = load_coherent_i32(right + c);
        const float al = l >= 0 ? coherent_area(nboxes, (uint64_t)l) : 0.0f;
        const float ar = r >= 0 ? coherent_area(nboxes, (uint64_t)r) : 0.0f;
#pragma unroll
        for (int j = 0; j < k; ++j)
            if (j == best) {
                lf[j] = l;
                la[j] = al;
            }
        lf[k] = r;
        la[k] = ar;
    }
#pragma unroll
    for (int j = 0; j < kTreeletLeaves; ++j) out.leaf[j] = lf[j];
#pragma unroll
    for (int j = 0; j < kTreeletLeaves - 1; ++j) out.pool[j] = pool[j];
    return true;
}

// DP step for the treelet-leaf subsets of size K.  A subset S's splits {P, S\P} with P containing S's lowest
// leaf are P = low | Q for the proper subsets Q of S\low (2^(K-1) - 1 of them).  G consecutive lanes (G a power
// of two, G x C(7, K) <= 64) share a subset, lane r taking the Q of index r, r + G, ... (the index's bits placed
// on S\low's leaves); the group then takes the minimum (cost, P), with (inf, 0) while no finite split was seen -
// the same result as one lane scanning every split, whatever the order.
template <int K, int G>
__device__ __forceinline__ void treelet_dp_size(TreeletLds& T, uint32_t lane) {
    constexpr uint32_t kOffset[kTreeletLeaves] = {0, 0, 7, 28, 63, 98, 119};
    constexpr uint32_t kCount[kTreeletLeaves] = {0, 7, 21, 35, 35, 21, 7};
    constexpr uint32_t kSplits = (1u << (K - 1)) - 1u;
    constexpr uint32_t kPerLane = (kSplits + G - 1) / G;
    static_assert((G & (G - 1)) == 0 && G * kCount[K] <= 64u, "one wave per subset size");
    const uint32_t sub = lane / G, r = lane % G;
    uint32_t S = 0;
    float best = INFINITY;
    uint32_t bp = 0;
    if (sub < kCount[K]) {
        S = kSubsetsBySize[kOffset[K] + sub];
        const uint32_t low = S & (0u - S), rest = S ^ low;
#pragma unroll
        for (uint32_t t = 0; t < kPerLane; ++t) {
            const uint32_t i = r + t * G;
            if (i < kSplits) {
                uint32_t q = 0, m = rest;
#pragma unroll
                for (int b = 0; b < K - 1; ++b) {
                    const uint32_t bit = m & (0u - m);
                    if (i & (1u << b)) q |= bit;
                    m ^= bit;
                }
                const uint32_t P = low | q;
                const float c = T.cost[P] + T.cost[S ^ P];
                if (c < best || (c == best && P < bp)) {
                    best = c;
                    bp = P;
                }
            }
        }
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        const float oc = __shfl_xor(best, off);
        const uint32_t op = (uint32_t)__shfl_xor((int)bp, off);
        if (oc < best || (oc == best && op < bp)) {
            best = oc;
            bp = op;
        }
    }
    if (sub < kCount[K] && r == 0) {
        T.cost[S] = T.area[S] + best;
        T.part[S] = bp;
    }
    __syncthreads();
}

// Whole-wave treelet restructuring below internal node R (uniform arguments) from its formation, which lane
// `owner` holds in `f` (formed = true).  Returns R's cost and height afterwards through new_cost / new_h.
__device__ void treelet_optimize(TreeletLds& T, uint32_t n, int32_t R, float old_cost, uint32_t old_h,
                                 uint32_t owner, bool formed, const TreeletForm& f, int32_t* left, int32_t* right,
                                 int32_t* parent, float* nboxes, float* cost, uint32_t* cnt, uint32_t* hgt,
                                 float& new_cost, uint32_t& new_h) {
    const uint32_t lane = threadIdx.x;
    __syncthreads();
    if (!formed) {  // fewer than 7 primitives below R: nothing to restructure
        new_cost = old_cost;
        new_h = old_h;
        return;
    }
    if (lane == owner) {
#pragma unroll
        for (int j = 0; j < kTreeletLeaves; ++j) T.leaf[j] = f.leaf[j];
#pragma unroll
        for (int j = 0; j < kTreeletLeaves - 1; ++j) T.pool[j] = f.pool[j];
    }
    __syncthreads();
    if (lane < (uint32_t)kTreeletLeaves) {
        const int32_t c = T.leaf[lane];
        const uint64_t bi = node_index(n, c);
#pragma unroll
        for (int k = 0; k < 6; ++k) T.box[lane][k] = load_coherent(nboxes + 6 * bi + k);
        T.lcost[lane] = c >= 0 ? load_coherent(cost + c) : 0.0f;
        T.lcnt[lane] = c >= 0 ? load_coherent_u32(cnt + c) : 1u;
        T.lhgt[lane] = c >= 0 ? load_coherent_u32(hgt + c) : 0u;
    }
    __syncthreads();
    // surface area of every subset's union box; singleton costs
    for (uint32_t S = lane + 1; S <= (uint32_t)kTreeletSubsets; S += 64) {
        float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
#pragma unroll
        for (int j = 0; j < kTreeletLeaves; ++j)
            if (S & (1u << j)) {
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    b[a] = fminf(b[a], T.box[j][a]);
                    b[3 + a] = fmaxf(b[3 + a], T.box[j][3 + a]);
                }
            }
        T.area[S] = box_area(b);
        if (__popc(S) == 1) {
            T.cost[S] = T.lcost[__ffs(S) - 1];
            T.part[S] = 0;
        }
    }
    __syncthreads();
    // subsets of size 2..6, each size after the smaller ones
    treelet_dp_size<2, 1>(T, lane);
    treelet_dp_size<3, 1>(T, lane);
    treelet_dp_size<4, 1>(T, lane);
    treelet_dp_size<5, 2>(T, lane);
    treelet_dp_size<6, 8>(T, lane);
    // all 7 leaves: the 63 splits containing leaf 0, one per lane, then a (cost, split) minimum over the wave
    float c7 = INFINITY;
    uint32_t p7 = 0xFFFFFFFFu;
    if (lane < 63u) {
        p7 = 1u | (lane << 1);
        c7 = T.cost[p7] + T.cost[127u ^ p7];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const float oc = __shfl_xor(c7, off);
        const uint32_t op = (uint32_t)__shfl_xor((int)p7, off);
        if (oc < c7 || (oc == c7 && op < p7)) {
            c7 = oc;
            p7 = op;
        }
    }
    const float total = T.area[127] + c7;
    new_cost = old_cost;
    new_h = old_h;
    if (!(total < old_cost)) return;
    if (lane == 0) {
        T.part[127] = p7;
        T.ls[0] = 127u;
        T.lx[0] = R;
        uint32_t cntl = 1, next = 1;
        bool ok = true;
        for (uint32_t i = 0; i < cntl && ok; ++i) {
            const uint32_t S = T.ls[i];
            const uint32_t P = T.part[S];
            const uint32_t C[2] = {P, S ^ P};
            if (P == 0u || (P & S) != P || P == S) {  // no finite split was found
                ok = false;
                break;
            }
            for (int h = 0; h < 2; ++h) {
                if (__popc(C[h]) == 1) {
                    T.lc[i][h] = T.leaf[__ffs(C[h]) - 1];
                    T.li[i][h] = -1 - (__ffs(C[h]) - 1);
                } else {
                    const int32_t x = T.pool[next++];
                    T.ls[cntl] = C[h];
                    T.lx[cntl] = x;
                    T.lc[i][h] = x;
                    T.li[i][h] = (int32_t)cntl;
                    ++cntl;
                }
            }
        }
        for (int i = kTreeletLeaves - 2; ok && i >= 0; --i) {
            uint32_t hc[2];
            for (int h = 0; h < 2; ++h) {
                const int32_t q = T.li[i][h];
                hc[h] = q >= 0 ? T.lh[q] : T.lhgt[-1 - q];
            }
            T.lh[i] = 1u + (hc[0] > hc[1] ? hc[0] : hc[1]);
        }
        T.commit = ok && T.lh[0] <= old_h ? 1u : 0u;
    }
    __syncthreads();
    if (!T.commit) return;
    if (lane < (uint32_t)(kTreeletLeaves - 1)) {
        const int32_t x = T.lx[lane];
        const uint32_t S = T.ls[lane];
        store_coherent_i32(left + x, T.lc[lane][0]);
        store_coherent_i32(right + x, T.lc[lane][1]);
        store_coherent_i32(parent + node_index(n, T.lc[lane][0]), x);
        store_coherent_i32(parent + node_index(n, T.lc[lane][1]), x);
        if (lane > 0) {  // the root's box is unchanged; its cost and height are published by its owner
            float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
            uint32_t np = 0;
            for (int j = 0; j < kTreeletLeaves; ++j)
                if (S & (1u << j)) {
                    for (int a = 0; a < 3; ++a) {
                        b[a] = fminf(b[a], T.box[j][a]);
                        b[3 + a] = fmaxf(b[3 + a], T.box[j][3 + a]);
                    }
                    np += T.lcnt[j];
                }
#pragma unroll
            for (int k = 0; k < 6; ++k) store_coherent(nboxes + 6 * (uint64_t)x + k, b[k]);
            store_coherent(cost + x, T.cost[S]);
            store_coherent_u32(cnt + x, np);
            store_coherent_u32(hgt + x, T.lh[lane]);
        }
    }
    new_cost = total;
    new_h = T.lh[0];
}

// Treelet refinement rounds: default kTreeletRounds, none under PREFER_FAST_BUILD, RTX_BUILD_TREELET_ROUNDS_*
// overrides both, NO_TREELET disables; round r restructures nodes with at least kTreeletMinPrims[r] primitives.
// (Each round costs ~17 ms on 20M triangles and buys 0.5-1.5% of closest-hit speed on the query scenes.)
constexpr int kTreeletRounds = 2;
constexpr uint32_t kTreeletMinPrims[3] = {7, 14, 28};
int treelet_rounds(uint32_t flags, uint32_t n) {
    if ((flags & RTX_BUILD_NO_TREELET_BIT) || n < (uint32_t)kTreeletLeaves) return 0;
    const uint32_t r = (flags >> RTX_BUILD_TREELET_ROUNDS_SHIFT) & 3u;
    if (r) return (int)r;
    return (flags & RTX_BUILD_PREFER_FAST_BUILD_BIT) ? 0 : kTreeletRounds;
}

// One wave per block.  min_prims = 0xFFFFFFFF: plain refit.
__global__ __launch_bounds__(64) void refit_refine_kernel(uint32_t n, int32_t* left, int32_t* right, int32_t* parent,
                                                          const uint32_t* __restrict__ order,
                                                          const float* __restrict__ prim_boxes, float* nboxes,
                                                          float* cost, uint32_t* cnt, uint32_t* hgt,
                                                          uint32_t* arrivals, uint32_t min_prims,
                                                          const uint32_t* __restrict__ node_blas) {
    __shared__ TreeletLds T;
    const uint32_t lane = threadIdx.x;
    const uint32_t pos = blockIdx.x * 64u + lane;
    bool active = pos < n;
    float cur[6] = {0, 0, 0, 0, 0, 0};
    float ccost = 0.0f;
    uint32_t ccnt = 1, chgt = 0;
    uint64_t self = 0;
    int32_t node = -1;
    if (active) {
        const float* pb = prim_boxes + 6 * (uint64_t)order[pos];
#pragma unroll
        for (int k = 0; k < 6; ++k) cur[k] = pb[k];
        self = (uint64_t)n - 1 + pos;
#pragma unroll
        for (int k = 0; k < 6; ++k) store_coherent(nboxes + 6 * self + k, cur[k]);
        node = load_coherent_i32(parent + self);
    }
    while (__ballot(active)) {
        bool own = false;
        if (active) {
            if (node < 0) {
                active = false;
            } else {
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                __builtin_amdgcn_s_waitcnt(0);
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                const uint32_t old = __hip_atomic_fetch_add(&arrivals[node], 1u, __ATOMIC_RELAXED,
                                                            __HIP_MEMORY_SCOPE_AGENT);
                if (old == 0u) {
                    active = false;
                } else {
                    own = true;
                    const int32_t l = load_coherent_i32(left + node), r = load_coherent_i32(right + node);
                    const uint64_t li = node_index(n, l), ri = node_index(n, r);
                    const bool self_left = li == self;
                    const int32_t sc = self_left ? r : l;
                    const uint64_t sib = self_left ? ri : li;
                    const float scost = sc >= 0 ? load_coherent(cost + sc) : 0.0f;
                    const uint32_t scnt = sc >= 0 ? load_coherent_u32(cnt + sc) : 1u;
                    const uint32_t shgt = sc >= 0 ? load_coherent_u32(hgt + sc) : 0u;
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        cur[a] = fminf(cur[a], load_coherent(nboxes + 6 * sib + a));
                        cur[3 + a] = fmaxf(cur[3 + a], load_coherent(nboxes + 6 * sib + 3 + a));
                    }
                    const float cl = self_left ? ccost : scost, cr = self_left ? scost : ccost;
                    ccost = box_area(cur) + (cl + cr);
                    ccnt += scnt;
                    chgt = 1u + (chgt > shgt ? chgt : shgt);
                }
            }
        }
        // batched builds never restructure across BLASes
        const bool root = own && ccnt >= min_prims && (!node_blas || node_blas[node] != kMixedBlas);
        uint64_t need = __ballot(root);
        // every lane with a treelet root forms its treelet at once (the dependent child / box loads overlap), then
        // the wave restructures them one after another
        TreeletForm form;
        const bool formed = root && treelet_form(node, left, right, nboxes, form);
        const uint64_t formed_mask = __ballot(formed);
        while (need) {
            const int L = __ffsll((unsigned long long)need) - 1;
            need &= need - 1;
            const int32_t R = __shfl(node, L);
            const float rc = __shfl(ccost, L);
            const uint32_t rh = (uint32_t)__shfl((int)chgt, L);
            float nc;
            uint32_t nh;
            treelet_optimize(T, n, R, rc, rh, (uint32_t)L, ((formed_mask >> L) & 1ull) != 0ull, form, left, right,
                             parent, nboxes, cost, cnt, hgt, nc, nh);
            if ((int)lane == L) {
                ccost = nc;
                chgt = nh;
            }
        }
        if (own) {
#pragma unroll
            for (int k = 0; k < 6; ++k) store_coherent(nboxes + 6 * (uint64_t)node + k, cur[k]);
            store_coherent(cost + node, ccost);
            store_coherent_u32(cnt + node, ccnt);
            store_coherent_u32(hgt + node, chgt);
            self = (uint64_t)node;
            node = load_coherent_i32(parent + node);
        }
    }
}

// ---- costs of the 8-wide collapse ------------------------------------------------------------------------------
//
// The SAH-optimal collapse of the binary tree into 8-wide nodes (Ylitie, Karras & Laine 2017, section 4), for
// this layout's leaves (one primitive per child slot).  wc[8x + j - 1] = C(x, j), the least cost of the subtree of
// x spread over at most j child slots of its parent:
//   leaf:      C(x, j) = c_leaf * A(x)
//   internal:  D(x, j) = min over k in [1, j-1] of C(l, k) + C(r, j - k)        (x dissolved into j slots)
//              C(x, 1) = A(x) + D(x, W)                                         (x becomes a wide node)
//              C(x, j) = min(C(x, j - 1), D(x, j)),  j >= 2
// computed bottom-up (second arrival merges, as the refit).  The collapse then takes, for each wide node x, the
// children that D(x, W) picks.  W is the build's node width (8, or 2..7 from RTX_BUILD_MAX_WIDTH): narrower trees
// use the same 64-B node with fewer children.
__device__ __forceinline__ void wide_table(uint32_t n, int32_t c, const float* nboxes, const float* wc, float c_leaf,
                                           float* t) {
    if (c < 0) {
        float b[6];
        const uint64_t bi = (uint64_t)n - 1 + (uint64_t)(-1 - c);
#pragma unroll
        for (int k = 0; k < 6; ++k) b[k] = load_coherent(nboxes + 6 * bi + k);
        const float a = c_leaf * box_area(b);
#pragma unroll
        for (int j = 0; j < 8; ++j) t[j] = a;
    } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) t[j] = load_coherent(wc + 8 * (uint64_t)c + j);
    }
}

// D(x, j) from the children's tables: the least cost and the left child's slot count k.
__device__ __forceinline__ float wide_split(const float* tl, const float* tr, int j, int& best_k) {
    float best = INFINITY;
    best_k = 1;
    for (int k = 1; k < j; ++k) {
        const float c = tl[k - 1] + tr[j - k - 1];
        if (c < best) {
            best = c;
            best_k = k;
        }
    }
    return best;
}

__global__ __launch_bounds__(256) void wide_cost_kernel(uint32_t n, const int32_t* __restrict__ left,
                                                        const int32_t* __restrict__ right,
                                                        const int32_t* __restrict__ parent, const float* nboxes,
                                                        float* wc, uint32_t* __restrict__ wdec, uint32_t* arrivals,
                                                        float c_leaf, int width) {
    const uint32_t pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= n) return;
    int32_t node = load_coherent_i32(parent + (uint64_t)n - 1 + pos);
    while (node >= 0) {
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        const uint32_t old =
            __hip_atomic_fetch_add(&arrivals[node], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old == 0u) return;  // the sibling's thread merges
        float tl[8], tr[8], b[6];
        wide_table(n, load_coherent_i32(left + node), nboxes, wc, c_leaf, tl);
        wide_table(n, load_coherent_i32(right + node), nboxes, wc, c_leaf, tr);
#pragma unroll
        for (int k = 0; k < 6; ++k) b[k] = load_coherent(nboxes + 6 * (uint64_t)node + k);
        float c[8];
        int k;
        c[0] = box_area(b) + wide_split(tl, tr, width, k);
        uint32_t dec = 0;
        for (int j = 2; j <= 8; ++j) {
            const float d = wide_split(tl, tr, j, k);
            dec |= ((uint32_t)k | (d < c[j - 2] ? 8u : 0u)) << (4 * (j - 2));
            c[j - 1] = fminf(c[j - 2], d);
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) store_coherent(wc + 8 * (uint64_t)node + j, c[j]);
        wdec[node] = dec;  // read by later launches only
        node = load_coherent_i32(parent + node);
    }
}

// Header setup after refit: root box and the 16-bit origin quantisation step.
__global__ void header_kernel(AsHeader* h, uint32_t type, uint32_t flags, uint32_t n, uint32_t ngeom,
                              const float* nboxes, const float* cost, uint64_t capacity_slots, uint64_t table_offset,
                              uint32_t* level_counts, uint2* tasks0) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    h->magic = type ? kMagicTlas : kMagicBlas;
    h->version = kVersion;
    h->type = type;
    h->flags = flags;
    h->prim_count = n;
    h->geometry_count = ngeom;
    h->node_count = 0;
    h->slot_count = 0;
    h->max_depth = 0;
    h->max_total_depth = 0;
    h->build_error = 0;
    h->content = 0;
    h->slots_offset = kHeaderBytes;
    h->inst_table_offset = table_offset;
    h->capacity_slots = capacity_slots;
    h->sah_cost = 0.0;
    h->update_offset = 0;
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    if (n > 0) {
        const float* rb = (n == 1) ? nboxes + 0 : nboxes;  // root: internal 0, or leaf 0 when n == 1
        for (int a = 0; a < 3; ++a) {
            lo[a] = rb[a];
            hi[a] = rb[3 + a];
        }
    }
    bool empty = !(lo[0] <= hi[0]) || n == 0;
    if (n > 1) {
        const float ra = box_area(nboxes);
        h->sah_cost = ra > 0.0f ? (double)cost[0] / (double)ra : 0.0;
    }
    for (int a = 0; a < 3; ++a) {
        if (empty) {
            lo[a] = 0.0f;
            hi[a] = 0.0f;
        }
        h->lo[a] = lo[a];
        h->hi[a] = hi[a];
        float ext = hi[a] - lo[a];
        h->qscale[a] = ext > 0.0f ? ext / 65535.0f : 0.0f;
    }
    h->lo[3] = h->hi[3] = h->qscale[3] = 0.0f;
    if (empty) {
        level_counts[0] = 0;
        return;
    }
    h->slot_count = 1;  // slot 0 = root
    level_counts[0] = 1;
    // root task: BVH2 root is internal node 0, or leaf 0 (encoded -1) when n == 1
    tasks0[0] = make_uint2(n == 1 ? 0xFFFFFFFFu : 0u, 0u);
}

__device__ __forceinline__ float exp_scale(uint32_t e) { return __uint_as_float(e << 23); }

// Quantisation frame of one axis: origin p (decoded from pq) and exponent byte e with fma(255, 2^e, p) >= hi.
__device__ void make_frame(float lo, float hi, float root_lo, float qscale, uint16_t& pq_out, float& p_out,
                           uint32_t& e_out) {
    uint32_t pq = 0;
    if (qscale > 0.0f) {
        float f = floorf((lo - root_lo) / qscale);
        pq = f <= 0.0f ? 0u : (f >= 65535.0f ? 65535u : (uint32_t)f);
        while (pq > 0 && fmaf((float)pq, qscale, root_lo) > lo) --pq;
    }
    float p = fmaf((float)pq, qscale, root_lo);
    uint32_t e = 1;  // 2^-126
    float x = (hi - p) / 255.0f;
    if (x > 0.0f) {
        int k;
        frexpf(x, &k);  // x = m * 2^k, m in [0.5, 1)
        int ee = k - 1 + 127;
        if (ee < 1) ee = 1;
        if (ee > 254) ee = 254;
        e = (uint32_t)ee;
        while (e < 254 && fmaf(255.0f, exp_scale(e), p) < hi) ++e;
    }
    pq_out = (uint16_t)pq;
    p_out = p;
    e_out = e;
}

__device__ __forceinline__ void quant_child(float clo, float chi, float p, uint32_t e, uint8_t& qlo, uint8_t& qhi) {
    float s = exp_scale(e);
    float fl = floorf((clo - p) / s);
    int ql = fl <= 0.0f ? 0 : (fl >= 255.0f ? 255 : (int)fl);
    while (ql > 0 && fmaf((float)ql, s, p) > clo) --ql;
    float fh = ceilf((chi - p) / s);
    int qh = fh <= 0.0f ? 0 : (fh >= 255.0f ? 255 : (int)fh);
    while (qh < 255 && fmaf((float)qh, s, p) < chi) ++qh;
    if (qh < ql) qh = ql;
    qlo = (uint8_t)ql;
    qhi = (uint8_t)qh;
}

// atomicAdd(addr, val), val < 16, for the active lanes of a wave whose addresses may differ (batched builds: one
// header per task): one atomic per distinct address, lanes of one address get consecutive ranges in lane order.
// Prefixes come from ballots of val's four bit planes, so lanes that are not active contribute nothing.  Every
// active lane of the wave must call it.
__device__ __forceinline__ uint32_t grouped_atomic_add(uint32_t* addr, uint32_t val) {
    const uint32_t lane = __lane_id();
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const unsigned long long a = (unsigned long long)(uintptr_t)addr;
    uint64_t todo = __ballot(1);
    uint32_t out = 0;
    while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const unsigned long long la = __shfl(a, leader);
        const bool mine = a == la;
        uint32_t prefix = 0, total = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const uint64_t m = __ballot(mine && ((val >> b) & 1u));
            prefix += (uint32_t)__popcll(m & below) << b;
            total += (uint32_t)__popcll(m) << b;
        }
        uint32_t base = 0;
        if ((int)lane == leader) base = atomicAdd(addr, total);
        base = __shfl(base, leader);
        if (mine) out = base + prefix;
        todo &= ~__ballot(mine);
    }
    return out;
}

// atomicMax(addr, val) with a wave-uniform val, one atomic per distinct address of the active lanes.
__device__ __forceinline__ void grouped_atomic_max(uint32_t* addr, uint32_t val) {
    const unsigned long long a = (unsigned long long)(uintptr_t)addr;
    uint64_t todo = __ballot(1);
    while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const unsigned long long la = __shfl(a, leader);
        const uint64_t grp = __ballot(a == la) & todo;
        if ((int)__lane_id() == leader) atomicMax(addr, val);
        todo &= ~grp;
    }
}

// One collapse level.  Task = (BVH2 node code, destination slot); a negative node code is a leaf that is the
// root itself (one-primitive structure).  Batched builds (BATCHED, as_ptr) carry each task's BLAS in a parallel
// array and write into that BLAS's structure.  Single builds keep the header pointer a kernel argument: the
// per-node atomics on it then have a wave-uniform address, which the compiler combines into one atomic per wave
// (a per-task pointer would send 64 times as many atomics to the same few words).
template <bool BATCHED>
__global__ __launch_bounds__(256) void collapse_level_kernel(uint32_t n, const int32_t* __restrict__ left,
                                                             const int32_t* __restrict__ right,
                                                             const uint32_t* __restrict__ order,
                                                             const float* __restrict__ nboxes,
                                                             const uint4* __restrict__ leaf_src,
                                                             AsHeader* __restrict__ h1, uint4* __restrict__ slots1,
                                                             uint32_t* __restrict__ level_counts, int level,
                                                             const uint2* __restrict__ tasks_in,
                                                             uint2* __restrict__ tasks_out,
                                                             AsHeader* const* __restrict__ as_ptr,
                                                             const uint32_t* __restrict__ tblas_in,
                                                             uint32_t* __restrict__ tblas_out,
                                                             const uint32_t* __restrict__ wdec, int width) {
    const uint32_t count = level_counts[level];
    if (count == 0) return;
    for (uint32_t ti = blockIdx.x * blockDim.x + threadIdx.x; ti < count; ti += gridDim.x * blockDim.x) {
        uint2 task = tasks_in[ti];
        const uint32_t tb = BATCHED ? tblas_in[ti] : 0u;
        AsHeader* h = BATCHED ? as_ptr[tb] : h1;
        uint4* slots = BATCHED ? (uint4*)((uint8_t*)h + kHeaderBytes) : slots1;
        const float rlo[3] = {h->lo[0], h->lo[1], h->lo[2]};
        const float qs[3] = {h->qscale[0], h->qscale[1], h->qscale[2]};
        int32_t ch[8];
        int cnt;
        if ((int32_t)task.x < 0) {
            ch[0] = (int32_t)task.x;
            cnt = 1;
        } else if (wdec) {
            // the children D(task, W) picks: a stack of (subtree, slots) pairs whose slots add up to W; each visited
            // node's split and "fewer slots" choices for every slot count come from the cost pass (wdec)
            int32_t sn[8];
            int si[8];
            int sp = 0;
            sn[sp] = (int32_t)task.x;
            si[sp++] = width;
            bool root = true;  // the task node itself is always split
            cnt = 0;
            while (sp > 0) {
                --sp;
                const int32_t x = sn[sp];
                const int j = si[sp];
                if (x < 0 || (j == 1 && !root)) {
                    ch[cnt++] = x;
                    continue;
                }
                const int32_t l = left[x], r = right[x];
                const uint32_t dj = wdec[x] >> (4 * (j - 2));
                const int k = (int)(dj & 7u);
                if (!root && !(dj & 8u)) {
                    sn[sp] = x;  // fewer slots cost no more
                    si[sp++] = j - 1;
                } else {
                    sn[sp] = l;
                    si[sp++] = k;
                    sn[sp] = r;
                    si[sp++] = j - k;
                }
                root = false;
            }
        } else {
            ch[0] = left[task.x];
            ch[1] = right[task.x];
            cnt = 2;
        }
        // greedy (no cost table): open the internal child with the largest surface area until W children
        while (!wdec && cnt < width) {
            int best = -1;
            float best_a = -1.0f;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                if (j < cnt && ch[j] >= 0) {
                    float a = box_area(nboxes + 6 * (uint64_t)ch[j]);
                    if (a > best_a) {
                        best_a = a;
                        best = j;
                    }
                }
            }
            if (best < 0) break;
            int32_t c = 0;
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (j == best) c = ch[j];
            int32_t l = left[c], r = right[c];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                if (j == best) ch[j] = l;
                if (j == cnt) ch[j] = r;
            }
            ++cnt;
        }
        // child boxes; drop empty ones
        float cb[8][6];
        bool cv[8];
        float nlo[3] = {INFINITY, INFINITY, INFINITY}, nhi[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            cv[j] = false;
            if (j < cnt) {
                int32_t c = ch[j];
                uint64_t bi = c >= 0 ? (uint64_t)c : (uint64_t)n - 1 + (uint64_t)(-1 - c);
#pragma unroll
                for (int k = 0; k < 6; ++k) cb[j][k] = nboxes[6 * bi + k];
                cv[j] = cb[j][0] <= cb[j][3] && cb[j][1] <= cb[j][4] && cb[j][2] <= cb[j][5];
                if (cv[j]) {
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        nlo[a] = fminf(nlo[a], cb[j][a]);
                        nhi[a] = fmaxf(nhi[a], cb[j][3 + a]);
                    }
                }
            }
        }
        // octant placement: greedy assignment of children to positions by centroid direction
        int pos_of[8];
        uint32_t used = 0;
        float nc[3] = {0.5f * (nlo[0] + nhi[0]), 0.5f * (nlo[1] + nhi[1]), 0.5f * (nlo[2] + nhi[2])};
#pragma unroll
        for (int j = 0; j < 8; ++j) pos_of[j] = -1;
        for (int it = 0; it < 8; ++it) {
            float best = -INFINITY;
            int bj = -1, bp = -1;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                if (!cv[j] || pos_of[j] >= 0) continue;
                float dx = 0.5f * (cb[j][0] + cb[j][3]) - nc[0];
                float dy = 0.5f * (cb[j][1] + cb[j][4]) - nc[1];
                float dz = 0.5f * (cb[j][2] + cb[j][5]) - nc[2];
#pragma unroll
                for (int p = 0; p < 8; ++p) {
                    if (used & (1u << p)) continue;
                    float sc = ((p & 1) ? dx : -dx) + ((p & 2) ? dy : -dy) + ((p & 4) ? dz : -dz);
                    if (sc > best) {
                        best = sc;
                        bj = j;
                        bp = p;
                    }
                }
            }
            if (bj < 0) break;
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (j == bj) pos_of[j] = bp;
            used |= 1u << bp;
        }
        // node frame; the 64-B node is assembled in registers (child bytes placed by shifts: a per-child byte
        // index into a local array would put the node in private memory)
        float p[3];
        uint32_t e[3];
        uint16_t pq[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) make_frame(nlo[a], nhi[a], rlo[a], qs[a], pq[a], p[a], e[a]);
        uint64_t qlo[3] = {0, 0, 0}, qhi[3] = {0, 0, 0};
        uint32_t imask = 0;
        const uint32_t nvalid = (uint32_t)__popc(used);
        const uint32_t base = BATCHED ? grouped_atomic_add(&h->slot_count, nvalid) : atomicAdd(&h->slot_count, nvalid);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if (!cv[j]) continue;
            int k = pos_of[j];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                uint8_t ql, qh;
                quant_child(cb[j][a], cb[j][3 + a], p[a], e[a], ql, qh);
                qlo[a] |= (uint64_t)ql << (8 * k);
                qhi[a] |= (uint64_t)qh << (8 * k);
            }
            uint32_t slot = base + (uint32_t)__popc(used & ((1u << k) - 1u));
            int32_t c = ch[j];
            if (c >= 0) {
                imask |= 1u << k;
                uint32_t o = atomicAdd(&level_counts[level + 1], 1u);
                tasks_out[o] = make_uint2((uint32_t)c, slot);
                if (BATCHED) tblas_out[o] = tb;
            } else {
                uint32_t prim = order[-1 - c];
                const uint4* src = leaf_src + 4 * (uint64_t)prim;
                uint4* dst = slots + 4 * (uint64_t)slot;
                dst[0] = src[0];
                dst[1] = src[1];
                dst[2] = src[2];
                dst[3] = src[3];
            }
        }
        // Node8 layout: pq[3] (bytes 0-5), e[3] (6-8), imask (9), vmask (10), pad (11), base (12-15),
        // qlo[3][8] (16-39), qhi[3][8] (40-63)
        uint4* ndst = slots + 4 * (uint64_t)task.y;
        ndst[0] = make_uint4((uint32_t)pq[0] | (uint32_t)pq[1] << 16, (uint32_t)pq[2] | e[0] << 16 | e[1] << 24,
                             e[2] | imask << 8 | used << 16, base);
        ndst[1] = make_uint4((uint32_t)qlo[0], (uint32_t)(qlo[0] >> 32), (uint32_t)qlo[1], (uint32_t)(qlo[1] >> 32));
        ndst[2] = make_uint4((uint32_t)qlo[2], (uint32_t)(qlo[2] >> 32), (uint32_t)qhi[0], (uint32_t)(qhi[0] >> 32));
        ndst[3] = make_uint4((uint32_t)qhi[1], (uint32_t)(qhi[1] >> 32), (uint32_t)qhi[2], (uint32_t)(qhi[2] >> 32));
        if (BATCHED) {
            grouped_atomic_add(&h->node_count, 1u);
            grouped_atomic_max(&h->max_depth, (uint32_t)(level + 1));
        } else {
            atomicAdd(&h->node_count, 1u);
            atomicMax(&h->max_depth, (uint32_t)(level + 1));
        }
    }
}

// Batched builds: one thread per BLAS writes its header (as header_kernel) and queues its root task.
__global__ __launch_bounds__(64) void batch_header_kernel(AsHeader* const* __restrict__ as_ptr,
                                                          const uint2* __restrict__ binfo, uint32_t nbatch,
                                                          const int32_t* __restrict__ roots, uint32_t n,
                                                          uint32_t flags, const float* __restrict__ nboxes,
                                                          const float* __restrict__ cost,
                                                          uint32_t* __restrict__ level_counts,
                                                          uint2* __restrict__ tasks0, uint32_t* __restrict__ tblas0) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbatch) return;
    AsHeader* h = as_ptr[b];
    const uint32_t np = binfo[b].x;
    h->magic = kMagicBlas;
    h->version = kVersion;
    h->type = 0;
    h->flags = flags;
    h->prim_count = np;
    h->geometry_count = binfo[b].y;
    h->node_count = 0;
    h->slot_count = 0;
    h->max_depth = 0;
    h->max_total_depth = 0;
    h->build_error = 0;
    h->content = 0;
    h->slots_offset = kHeaderBytes;
    h->inst_table_offset = 0;
    h->capacity_slots = 2ull * (np ? np : 1) + 1;
    h->sah_cost = 0.0;
    h->update_offset = 0;
    const int32_t root = np ? roots[b] : 0x7FFFFFFF;
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    bool empty = root == 0x7FFFFFFF;
    if (!empty) {
        const float* rb = nboxes + 6 * (root >= 0 ? (uint64_t)root : (uint64_t)n - 1 + (uint64_t)(-1 - root));
        for (int a = 0; a < 3; ++a) {
            lo[a] = rb[a];
            hi[a] = rb[3 + a];
        }
        empty = !(lo[0] <= hi[0]);
        if (root >= 0) {
            const float ra = box_area(rb);
            h->sah_cost = ra > 0.0f ? (double)cost[root] / (double)ra : 0.0;
        }
    }
    for (int a = 0; a < 3; ++a) {
        if (empty) {
            lo[a] = 0.0f;
            hi[a] = 0.0f;
        }
        h->lo[a] = lo[a];
        h->hi[a] = hi[a];
        const float ext = hi[a] - lo[a];
        h->qscale[a] = ext > 0.0f ? ext / 65535.0f : 0.0f;
    }
    h->lo[3] = h->hi[3] = h->qscale[3] = 0.0f;
    if (empty) return;
    h->slot_count = 1;
    const uint32_t o = atomicAdd(&level_counts[0], 1u);
    tasks0[o] = make_uint2((uint32_t)root, 0u);
    tblas0[o] = b;
}

__global__ void batch_finish_kernel(AsHeader* const* __restrict__ as_ptr, uint32_t nbatch) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nbatch) as_ptr[b]->max_total_depth = as_ptr[b]->max_depth;
}

__global__ void fill_bounds_kernel(uint32_t* __restrict__ cb, uint32_t nbatch) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 6 * nbatch) cb[i] = (i % 6) < 3 ? 0xFFFFFFFFu : 0u;
}

// Tasks left after the last enqueued level mean the binary tree was deeper than the level bound.
__global__ void collapse_check_kernel(AsHeader* h, const uint32_t* level_counts, int max_levels, uint32_t* err) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && level_counts[max_levels] != 0u) {
        if (h) h->build_error = 1u;
        atomicOr(err, kErrBuildDepth);
    }
}

__global__ void finish_tlas_kernel(AsHeader* h, const uint32_t* max_blas_depth) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        h->max_total_depth = h->max_depth + max_blas_depth[0];
        h->content = max_blas_depth[1];
    }
}

// BLAS content bits from its geometries (one thread per geometry; the header's word was zeroed by the header
// kernel): into `single`, or into as_ptr[geometry's BLAS] for batched builds.
__global__ void content_kernel(const GeomDesc* __restrict__ geoms, uint32_t ngeom, AsHeader* single,
                               AsHeader* const* __restrict__ as_ptr) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngeom) return;
    const GeomDesc d = geoms[g];
    if (d.prim_count == 0) return;
    const uint32_t c = ((d.flags & RTX_GEOMETRY_OPAQUE_BIT) ? 0u : kContentNonOpaque) |
                       (d.type == RTX_GEOMETRY_TYPE_AABBS ? kContentAabb : 0u);
    if (c) atomicOr(&(single ? single : as_ptr[d.blas])->content, c);
}
__global__ void finish_blas_kernel(AsHeader* h) {
    if (threadIdx.x == 0 && blockIdx.x == 0) h->max_total_depth = h->max_depth;
}

__global__ void fill_kernel(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// One piece of a host array carried by value as a launch argument (a recorded launch keeps its arguments).
constexpr uint32_t kArgBlobWords = 256;
struct ArgBlob {
    uint32_t w[kArgBlobWords];
};
__global__ void arg_blob_kernel(ArgBlob b, uint32_t* __restrict__ dst, uint32_t words) {
    if (threadIdx.x < words) dst[threadIdx.x] = b.w[threadIdx.x];
}

// ---- scratch layout --------------------------------------------------------------------------------------------
struct Layout {
    uint64_t leaves, boxes, keys, vals, sort, left, right, parent, nboxes, arrivals, cost, cnt, hgt, wcost, wdec,
        tasks0, tasks1, levels, cbounds, misc, geoms, total;
    // batched builds only
    uint64_t pblas = 0, nblas = 0, roots = 0, asptr = 0, binfo = 0, tblas0 = 0, tblas1 = 0, bcb = 0;
};

Layout make_layout(uint32_t n, int key_bytes, uint32_t ngeom, uint32_t nbatch = 0) {
    Layout L;
    uint64_t o = 0;
    auto take = [&](uint64_t bytes) {
        uint64_t at = o;
        o += align_up(bytes ? bytes : 1, 256);
        return at;
    };
    uint64_t nn = n ? n : 1;
    L.leaves = take(nn * 64);
    L.boxes = take(nn * 24);
    L.keys = take(nn * key_bytes);
    L.vals = take(nn * 4);
    L.sort = take(sort_scratch_bytes(n, key_bytes));
    L.left = take(nn * 4);
    L.right = take(nn * 4);
    L.parent = take(2 * nn * 4);
    L.nboxes = take(2 * nn * 24);
    L.arrivals = take(nn * 4);
    L.cost = take(nn * 4);
    L.cnt = take(nn * 4);
    L.hgt = take(nn * 4);
    L.wcost = take(nn * 32);
    L.wdec = take(nn * 4);
    L.tasks0 = take(nn * 8);
    L.tasks1 = take(nn * 8);
    L.levels = take(128 * 4);
    L.cbounds = take(8 * 4);
    L.misc = take(64);
    L.geoms = take((uint64_t)(ngeom ? ngeom : 1) * sizeof(GeomDesc));
    if (nbatch) {
        L.pblas = take(nn * 4);
        L.nblas = take(nn * 4);
        L.roots = take((uint64_t)nbatch * 4);
        L.asptr = take((uint64_t)nbatch * 8);
        L.binfo = take((uint64_t)nbatch * 8);
        L.tblas0 = take(nn * 4);
        L.tblas1 = take(nn * 4);
        L.bcb = take((uint64_t)nbatch * 24);
    }
    L.total = o;
    return L;
}

uint64_t as_bytes(uint32_t n, bool tlas, bool update = false) {
    uint64_t slots = 2ull * (n ? n : 1) + 1;
    uint64_t b = kHeaderBytes + slots * kSlotBytes;
    if (tlas) b = align_up(b, 256) + (uint64_t)(n ? n : 1) * sizeof(InstanceInfo);
    if (update) b = align_up(b, 256) + (uint64_t)(n ? n : 1) * 4 * 5;  // order, left, right, parent (2n)
    return align_up(b, 256);
}

// Kept hierarchy of an updatable BLAS: order[n], left[n-1], right[n-1], parent[2n-1] (slots of 4 B).
uint64_t update_offset(uint32_t n) { return align_up(kHeaderBytes + (2ull * (n ? n : 1) + 1) * kSlotBytes, 256); }

__global__ void set_update_offset_kernel(AsHeader* h, uint64_t off) {
    if (threadIdx.x == 0 && blockIdx.x == 0) h->update_offset = off;
}

// Leaf cost of the wide collapse, relative to a node visit (cost 1): a triangle test (~80 VALU operations
// against the node test's ~250) or an instance entry (FP64 ray transform and a BLAS root visit).
constexpr float kLeafCostTri = 0.3f;
constexpr float kLeafCostInstance = 1.0f;

// Refit with treelet refinement; each further round restructures the tree the previous one left.  Then the costs
// of the 8-wide collapse (wcost) over the final tree.
// Children per wide node: RTX_BUILD_MAX_WIDTH(w) caps it at w (2..7), 0 = 8.
int node_width(uint32_t flags) {
    const uint32_t w = (flags >> RTX_BUILD_MAX_WIDTH_SHIFT) & 7u;
    return w == 0u ? 8 : (w < 2u ? 2 : (int)w);
}

void refit_passes(uint32_t n, uint32_t flags, uint8_t* scratch, const Layout& L, const float* prim_boxes,
                  hipStream_t s, const uint32_t* node_blas = nullptr, float c_leaf = kLeafCostTri) {
    const int rounds = treelet_rounds(flags, n);
    uint32_t* arrivals = (uint32_t*)(scratch + L.arrivals);
    for (int round = 0; round < (rounds ? rounds : 1); ++round) {
        if (n > 1) fill_kernel<<<256, 256, 0, s>>>(arrivals, n, 0u);
        refit_refine_kernel<<<(n + 63) / 64, 64, 0, s>>>(
            n, (int32_t*)(scratch + L.left), (int32_t*)(scratch + L.right), (int32_t*)(scratch + L.parent),
            (const uint32_t*)(scratch + L.vals), prim_boxes, (float*)(scratch + L.nboxes), (float*)(scratch + L.cost),
            (uint32_t*)(scratch + L.cnt), (uint32_t*)(scratch + L.hgt), arrivals,
            rounds ? kTreeletMinPrims[round] : 0xFFFFFFFFu, node_blas);
    }
    if (n > 1) {
        fill_kernel<<<256, 256, 0, s>>>(arrivals, n, 0u);
        wide_cost_kernel<<<(n + 255) / 256, 256, 0, s>>>(
            n, (const int32_t*)(scratch + L.left), (const int32_t*)(scratch + L.right),
            (const int32_t*)(scratch + L.parent), (const float*)(scratch + L.nboxes), (float*)(scratch + L.wcost),
            (uint32_t*)(scratch + L.wdec), arrivals, c_leaf, node_width(flags));
    }
}

// Common LBVH tail: morton -> sort -> karras -> refit -> header -> collapse.
hipError_t lbvh_common(uint32_t n, bool k64, uint32_t type, uint32_t flags, uint32_t ngeom, uint8_t* scratch,
                       const Layout& L, uint8_t* dst, uint64_t table_offset, uint32_t* err_flags, hipStream_t s) {
    AsHeader* h = (AsHeader*)dst;
    uint4* slots = (uint4*)(dst + kHeaderBytes);
    float* boxes = (float*)(scratch + L.boxes);
    uint32_t* cb = (uint32_t*)(scratch + L.cbounds);
    uint32_t* vals = (uint32_t*)(scratch + L.vals);
    int32_t* left = (int32_t*)(scratch + L.left);
    int32_t* right = (int32_t*)(scratch + L.right);
    int32_t* parent = (int32_t*)(scratch + L.parent);
    float* nboxes = (float*)(scratch + L.nboxes);
    float* cost = (float*)(scratch + L.cost);
    uint32_t* levels = (uint32_t*)(scratch + L.levels);
    uint2* t0 = (uint2*)(scratch + L.tasks0);
    uint2* t1 = (uint2*)(scratch + L.tasks1);
    const unsigned gb = (unsigned)((n + 255) / 256);
    hipError_t err;
    if (n > 0) {
        if (k64) {
            morton_kernel<uint64_t><<<gb, 256, 0, s>>>(boxes, n, cb, (uint64_t*)(scratch + L.keys));
            if ((err = sort_pairs_u64((uint64_t*)(scratch + L.keys), vals, n, scratch + L.sort, err_flags, s)))
                return err;
        } else {
            morton_kernel<uint32_t><<<gb, 256, 0, s>>>(boxes, n, cb, (uint32_t*)(scratch + L.keys));
            if ((err = sort_pairs_u32((uint32_t*)(scratch + L.keys), vals, n, scratch + L.sort, err_flags, s)))
                return err;
        }
        if (n > 1) {
            if (k64)
                karras_kernel<uint64_t><<<gb, 256, 0, s>>>((uint64_t*)(scratch + L.keys), n, left, right, parent);
            else
                karras_kernel<uint32_t><<<gb, 256, 0, s>>>((uint32_t*)(scratch + L.keys), n, left, right, parent);
        } else {
            fill_kernel<<<1, 64, 0, s>>>((uint32_t*)parent, 1, 0xFFFFFFFFu);
        }
        refit_passes(n, flags, scratch, L, boxes, s, nullptr, type == 1 ? kLeafCostInstance : kLeafCostTri);
    }
    fill_kernel<<<1, 128, 0, s>>>(levels, 128, 0u);
    // root box: internal node 0, or the single leaf (stored at index n-1 == 0)
    header_kernel<<<1, 64, 0, s>>>(h, type, flags, n, ngeom, nboxes, cost, 2ull * (n ? n : 1) + 1, table_offset, levels, t0);
    if (n > 0) {
        const int max_levels = k64 ? kMaxLevels64 : kMaxLevels32;
        for (int lv = 0; lv < max_levels; ++lv) {
            uint2* tin = (lv & 1) ? t1 : t0;
            uint2* tout = (lv & 1) ? t0 : t1;
            collapse_level_kernel<false><<<kCollapseBlocks, 256, 0, s>>>(n, left, right, vals, nboxes,
                                                                  (const uint4*)(scratch + L.leaves), h, slots, levels,
                                                                  lv, tin, tout, nullptr, nullptr, nullptr,
                                                                  n > 1 ? (const uint32_t*)(scratch + L.wdec) : nullptr,
                                                                  node_width(flags));
        }
        collapse_check_kernel<<<1, 64, 0, s>>>(h, levels, max_levels, err_flags);
    }
    return hipGetLastError();
}

}  // namespace

namespace {
// Morton / sort / Karras on caller boxes, optionally followed by the refit + refinement passes (rounds >= 0).
hipError_t debug_tree(const float* boxes, uint32_t n, bool k64, int rounds, void* codes_out, uint32_t* order_out,
                      int32_t* left_out, int32_t* right_out, float* cost_out, uint32_t* hgt_out, uint32_t* err_flags,
                      hipStream_t s) {
    Layout L = make_layout(n, k64 ? 8 : 4, 0);
    uint8_t* sc = nullptr;
    hipError_t e = hipMalloc(&sc, L.total);
    if (e) return e;
    uint32_t* cb = (uint32_t*)(sc + L.cbounds);
    fill_kernel<<<1, 64, 0, s>>>(cb, 3, 0xFFFFFFFFu);
    fill_kernel<<<1, 64, 0, s>>>(cb + 3, 3, 0u);
    const unsigned gb = (unsigned)((n + 255) / 256);
    uint32_t* vals = (uint32_t*)(sc + L.vals);
    if (n > 0) {
        box_bounds_kernel<<<gb, 256, 0, s>>>(boxes, n, cb);
        if (k64) {
            morton_kernel<uint64_t><<<gb, 256, 0, s>>>(boxes, n, cb, (uint64_t*)(sc + L.keys));
            e = sort_pairs_u64((uint64_t*)(sc + L.keys), vals, n, sc + L.sort, err_flags, s);
            if (!e && n > 1)
                karras_kernel<uint64_t><<<gb, 256, 0, s>>>((uint64_t*)(sc + L.keys), n, (int32_t*)(sc + L.left),
                                                           (int32_t*)(sc + L.right), (int32_t*)(sc + L.parent));
        } else {
            morton_kernel<uint32_t><<<gb, 256, 0, s>>>(boxes, n, cb, (uint32_t*)(sc + L.keys));
            e = sort_pairs_u32((uint32_t*)(sc + L.keys), vals, n, sc + L.sort, err_flags, s);
            if (!e && n > 1)
                karras_kernel<uint32_t><<<gb, 256, 0, s>>>((uint32_t*)(sc + L.keys), n, (int32_t*)(sc + L.left),
                                                           (int32_t*)(sc + L.right), (int32_t*)(sc + L.parent));
        }
    }
    if (!e && rounds >= 0 && n > 1)
        refit_passes(n, rounds ? RTX_BUILD_TREELET_ROUNDS(rounds) : RTX_BUILD_NO_TREELET_BIT, sc, L, boxes, s);
    const uint64_t m = n > 1 ? (uint64_t)(n - 1) * 4 : 0;
    if (!e && n > 0 && codes_out)
        e = hipMemcpyAsync(codes_out, sc + L.keys, (uint64_t)n * (k64 ? 8 : 4), hipMemcpyDeviceToDevice, s);
    if (!e && n > 0) e = hipMemcpyAsync(order_out, vals, (uint64_t)n * 4, hipMemcpyDeviceToDevice, s);
    if (!e && m) e = hipMemcpyAsync(left_out, sc + L.left, m, hipMemcpyDeviceToDevice, s);
    if (!e && m) e = hipMemcpyAsync(right_out, sc + L.right, m, hipMemcpyDeviceToDevice, s);
    if (!e && m && cost_out) e = hipMemcpyAsync(cost_out, sc + L.cost, m, hipMemcpyDeviceToDevice, s);
    if (!e && m && hgt_out) e = hipMemcpyAsync(hgt_out, sc + L.hgt, m, hipMemcpyDeviceToDevice, s);
    hipError_t e2 = hipStreamSynchronize(s);
    (void)hipFree(sc);
    return e ? e : e2;
}
}  // namespace

hipError_t debug_lbvh(const float* boxes, uint32_t n, bool k64, void* codes_out, uint32_t* order_out,
                      int32_t* left_out, int32_t* right_out, uint32_t* err_flags, hipStream_t s) {
    return debug_tree(boxes, n, k64, -1, codes_out, order_out, left_out, right_out, nullptr, nullptr, err_flags, s);
}

hipError_t debug_treelet(const float* boxes, uint32_t n, bool k64, uint32_t rounds, uint32_t* order_out,
                         int32_t* left_out, int32_t* right_out, float* cost_out, uint32_t* hgt_out,
                         uint32_t* err_flags, hipStream_t s) {
    return debug_tree(boxes, n, k64, (int)rounds, nullptr, order_out, left_out, right_out, cost_out, hgt_out,
                      err_flags, s);
}

hipError_t debug_sort(void* keys, uint32_t* vals, uint32_t n, int key_bytes, uint32_t* err_flags, hipStream_t s) {
    uint8_t* sc = nullptr;
    hipError_t e = hipMalloc(&sc, sort_scratch_bytes(n, key_bytes) + 256);
    if (e) return e;
    if (key_bytes == 8)
        e = sort_pairs_u64((uint64_t*)keys, vals, n, sc, err_flags, s);
    else
        e = sort_pairs_u32((uint32_t*)keys, vals, n, sc, err_flags, s);
    hipError_t e2 = hipStreamSynchronize(s);
    (void)hipFree(sc);
    return e ? e : e2;
}

hipError_t upload_host_array(void* dst, const void* src, uint64_t bytes, hipStream_t s) {
    if (bytes == 0) return hipSuccess;
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    hipError_t e = hipStreamIsCapturing(s, &st);
    if (e != hipSuccess) return e;
    if (st == hipStreamCaptureStatusNone) return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
    if (bytes > kArgUploadMax || (bytes & 3) || ((uintptr_t)dst & 3)) return hipErrorNotSupported;
    const uint8_t* in = (const uint8_t*)src;
    for (uint64_t at = 0; at < bytes; at += 4ull * kArgBlobWords) {
        const uint32_t words = (uint32_t)std::min<uint64_t>(kArgBlobWords, (bytes - at) / 4);
        ArgBlob b{};
        std::memcpy(b.w, in + at, 4ull * words);
        arg_blob_kernel<<<1, kArgBlobWords, 0, s>>>(b, (uint32_t*)((uint8_t*)dst + at), words);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

uint64_t blas_as_bytes(uint32_t n, bool update) { return as_bytes(n, false, update); }
uint64_t tlas_as_bytes(uint32_t n) { return as_bytes(n, true); }
uint64_t blas_scratch_bytes(uint32_t n, bool k64, uint32_t ngeom) { return make_layout(n, k64 ? 8 : 4, ngeom).total; }
uint64_t tlas_scratch_bytes(uint32_t n, bool k64) { return make_layout(n, k64 ? 8 : 4, 0).total; }
uint64_t blas_batch_scratch_bytes(uint32_t n, uint32_t ngeom, uint32_t nbatch) {
    return make_layout(n, 8, ngeom, nbatch ? nbatch : 1).total;
}

hipError_t build_blas(const GeomDesc* geoms_host, uint32_t ngeom, uint32_t n, uint32_t flags, void* dst,
                      void* scratch, uint32_t* err_flags, hipStream_t s) {
    const bool k64 = (flags & RTX_BUILD_MORTON63_BIT) != 0;
    Layout L = make_layout(n, k64 ? 8 : 4, ngeom);
    uint8_t* sc = (uint8_t*)scratch;
    hipError_t err;
    GeomDesc* gd = (GeomDesc*)(sc + L.geoms);
    if (ngeom) {
        if ((err = upload_host_array(gd, geoms_host, ngeom * sizeof(GeomDesc), s))) return err;
    }
    uint32_t* cb = (uint32_t*)(sc + L.cbounds);
    fill_kernel<<<1, 64, 0, s>>>(cb, 3, 0xFFFFFFFFu);
    fill_kernel<<<1, 64, 0, s>>>(cb + 3, 3, 0u);
    if (n > 0)
        prep_tris_kernel<<<(n + 255) / 256, 256, 0, s>>>(gd, (int)ngeom, n, (TriSlot*)(sc + L.leaves),
                                                         (float*)(sc + L.boxes), cb, nullptr);
    if ((err = lbvh_common(n, k64, 0, flags, ngeom, sc, L, (uint8_t*)dst, 0, err_flags, s))) return err;
    if (ngeom) content_kernel<<<(ngeom + 255) / 256, 256, 0, s>>>(gd, ngeom, (AsHeader*)dst, nullptr);
    finish_blas_kernel<<<1, 64, 0, s>>>((AsHeader*)dst);
    if ((flags & RTX_BUILD_ALLOW_UPDATE_BIT) && n > 0) {
        // keep the final hierarchy (after treelet refinement) and the triangle order for updates
        const uint64_t off = update_offset(n);
        uint8_t* u = (uint8_t*)dst + off;
        if ((err = hipMemcpyAsync(u, sc + L.vals, 4ull * n, hipMemcpyDeviceToDevice, s))) return err;
        if (n > 1) {
            if ((err = hipMemcpyAsync(u + 4ull * n, sc + L.left, 4ull * (n - 1), hipMemcpyDeviceToDevice, s)))
                return err;
            if ((err = hipMemcpyAsync(u + 4ull * (2 * n - 1), sc + L.right, 4ull * (n - 1), hipMemcpyDeviceToDevice,
                                      s)))
                return err;
        }
        if ((err = hipMemcpyAsync(u + 4ull * (3 * n - 2), sc + L.parent, 4ull * (2 * n - 1), hipMemcpyDeviceToDevice,
                                  s)))
            return err;
        set_update_offset_kernel<<<1, 64, 0, s>>>((AsHeader*)dst, off);
    }
    return hipGetLastError();
}

// UPDATE mode: new triangle positions, the kept order and hierarchy, plain refit, header and collapse again.
hipError_t update_blas(const GeomDesc* geoms_host, uint32_t ngeom, uint32_t n, uint32_t flags, void* dst,
                       void* scratch, uint32_t* err_flags, hipStream_t s) {
    Layout L = make_layout(n, 4, ngeom);
    uint8_t* sc = (uint8_t*)scratch;
    hipError_t err;
    GeomDesc* gd = (GeomDesc*)(sc + L.geoms);
    if (ngeom && (err = upload_host_array(gd, geoms_host, ngeom * sizeof(GeomDesc), s)))
        return err;
    uint32_t* cb = (uint32_t*)(sc + L.cbounds);
    fill_kernel<<<1, 64, 0, s>>>(cb, 3, 0xFFFFFFFFu);
    fill_kernel<<<1, 64, 0, s>>>(cb + 3, 3, 0u);
    AsHeader* h = (AsHeader*)dst;
    const uint64_t off = update_offset(n);
    const uint8_t* u = (const uint8_t*)dst + off;
    if (n > 0) {
        prep_tris_kernel<<<(n + 255) / 256, 256, 0, s>>>(gd, (int)ngeom, n, (TriSlot*)(sc + L.leaves),
                                                         (float*)(sc + L.boxes), cb, nullptr);
        if ((err = hipMemcpyAsync(sc + L.vals, u, 4ull * n, hipMemcpyDeviceToDevice, s))) return err;
        if (n > 1) {
            if ((err = hipMemcpyAsync(sc + L.left, u + 4ull * n, 4ull * (n - 1), hipMemcpyDeviceToDevice, s)))
                return err;
            if ((err = hipMemcpyAsync(sc + L.right, u + 4ull * (2 * n - 1), 4ull * (n - 1), hipMemcpyDeviceToDevice,
                                      s)))
                return err;
        }
        if ((err = hipMemcpyAsync(sc + L.parent, u + 4ull * (3 * n - 2), 4ull * (2 * n - 1),
                                  hipMemcpyDeviceToDevice, s)))
            return err;
        refit_passes(n, flags | RTX_BUILD_NO_TREELET_BIT, sc, L, (const float*)(sc + L.boxes), s);
    }
    uint32_t* levels = (uint32_t*)(sc + L.levels);
    uint2* t0 = (uint2*)(sc + L.tasks0);
    uint2* t1 = (uint2*)(sc + L.tasks1);
    fill_kernel<<<1, 128, 0, s>>>(levels, 128, 0u);
    header_kernel<<<1, 64, 0, s>>>(h, 0, flags, n, ngeom, (const float*)(sc + L.nboxes), (const float*)(sc + L.cost),
                                   2ull * (n ? n : 1) + 1, 0, levels, t0);
    if (n > 0) {
        for (int lv = 0; lv < kMaxLevels32; ++lv) {
            uint2* tin = (lv & 1) ? t1 : t0;
            uint2* tout = (lv & 1) ? t0 : t1;
            collapse_level_kernel<false><<<kCollapseBlocks, 256, 0, s>>>(
                n, (const int32_t*)(sc + L.left), (const int32_t*)(sc + L.right), (const uint32_t*)(sc + L.vals),
                (const float*)(sc + L.nboxes), (const uint4*)(sc + L.leaves), h, (uint4*)((uint8_t*)dst + kHeaderBytes),
                levels, lv, tin, tout, nullptr, nullptr, nullptr, n > 1 ? (const uint32_t*)(sc + L.wdec) : nullptr,
                node_width(flags));
        }
        collapse_check_kernel<<<1, 64, 0, s>>>(h, levels, kMaxLevels32, err_flags);
    }
    if (ngeom) content_kernel<<<(ngeom + 255) / 256, 256, 0, s>>>(gd, ngeom, h, nullptr);
    finish_blas_kernel<<<1, 64, 0, s>>>(h);
    set_update_offset_kernel<<<1, 64, 0, s>>>(h, off);
    return hipGetLastError();
}

hipError_t build_tlas(const rtx_instance* inst, uint32_t n, uint32_t flags, void* dst, void* scratch,
                      uint32_t* err_flags, hipStream_t s) {
    const bool k64 = (flags & RTX_BUILD_MORTON63_BIT) != 0;
    Layout L = make_layout(n, k64 ? 8 : 4, 0);
    uint8_t* sc = (uint8_t*)scratch;
    uint64_t table_offset = align_up(kHeaderBytes + (2ull * (n ? n : 1) + 1) * kSlotBytes, 256);
    uint32_t* cb = (uint32_t*)(sc + L.cbounds);
    uint32_t* maxd = (uint32_t*)(sc + L.misc);
    fill_kernel<<<1, 64, 0, s>>>(cb, 3, 0xFFFFFFFFu);
    fill_kernel<<<1, 64, 0, s>>>(cb + 3, 3, 0u);
    fill_kernel<<<1, 64, 0, s>>>(maxd, 2, 0u);  // deepest BLAS, content bits
    if (n > 0)
        prep_instances_kernel<<<(n + 255) / 256, 256, 0, s>>>(inst, n, (InstSlot*)(sc + L.leaves),
                                                              (InstanceInfo*)((uint8_t*)dst + table_offset),
                                                              (float*)(sc + L.boxes), cb, maxd, maxd + 1);
    hipError_t err;
    if ((err = lbvh_common(n, k64, 1, flags, 0, sc, L, (uint8_t*)dst, table_offset, err_flags, s))) return err;
    finish_tlas_kernel<<<1, 64, 0, s>>>((AsHeader*)dst, maxd);
    return hipGetLastError();
}

// Batched BLAS build: every step runs once over the primitives of all BLASes.  Keys are BLAS id << 30 | Morton
// code, so after the stable sort each BLAS is one contiguous range and one subtree of the single Karras tree; the
// part of the tree above the BLAS roots is never collapsed, refined or stored.
hipError_t build_blas_batch(const GeomDesc* geoms_host, uint32_t ngeom, const uint2* binfo_host,
                            AsHeader* const* dst_host, uint32_t nbatch, uint32_t n, uint32_t flags, void* scratch,
                            uint32_t* err_flags, hipStream_t s) {
    Layout L = make_layout(n, 8, ngeom, nbatch);
    uint8_t* sc = (uint8_t*)scratch;
    hipError_t err;
    GeomDesc* gd = (GeomDesc*)(sc + L.geoms);
    AsHeader** as_ptr = (AsHeader**)(sc + L.asptr);
    uint2* binfo = (uint2*)(sc + L.binfo);
    if (ngeom && (err = upload_host_array(gd, geoms_host, ngeom * sizeof(GeomDesc), s)))
        return err;
    if ((err = upload_host_array(as_ptr, dst_host, nbatch * sizeof(AsHeader*), s))) return err;
    if ((err = upload_host_array(binfo, binfo_host, nbatch * sizeof(uint2), s))) return err;
    uint32_t* cb = (uint32_t*)(sc + L.bcb);
    uint32_t* pblas = (uint32_t*)(sc + L.pblas);
    uint32_t* nblas = (uint32_t*)(sc + L.nblas);
    int32_t* roots = (int32_t*)(sc + L.roots);
    uint64_t* keys = (uint64_t*)(sc + L.keys);
    uint32_t* vals = (uint32_t*)(sc + L.vals);
    int32_t* left = (int32_t*)(sc + L.left);
    int32_t* right = (int32_t*)(sc + L.right);
    int32_t* parent = (int32_t*)(sc + L.parent);
    float* nboxes = (float*)(sc + L.nboxes);
    float* boxes = (float*)(sc + L.boxes);
    uint32_t* levels = (uint32_t*)(sc + L.levels);
    const unsigned gb = (unsigned)((n + 255) / 256);
    fill_bounds_kernel<<<(6 * nbatch + 255) / 256, 256, 0, s>>>(cb, nbatch);
    fill_kernel<<<(nbatch + 255) / 256, 256, 0, s>>>((uint32_t*)roots, nbatch, 0x7FFFFFFFu);
    if (n > 0) {
        prep_tris_kernel<<<gb, 256, 0, s>>>(gd, (int)ngeom, n, (TriSlot*)(sc + L.leaves), boxes, cb, pblas);
        morton_batch_kernel<<<gb, 256, 0, s>>>(boxes, n, cb, pblas, keys);
        if ((err = sort_pairs_u64(keys, vals, n, sc + L.sort, err_flags, s))) return err;
        if (n > 1) {
            karras_kernel<uint64_t><<<gb, 256, 0, s>>>(keys, n, left, right, parent, nblas);
        } else {
            fill_kernel<<<1, 64, 0, s>>>((uint32_t*)parent, 1, 0xFFFFFFFFu);
        }
        blas_roots_kernel<<<(unsigned)((2ull * n - 1 + 255) / 256), 256, 0, s>>>(n, keys, parent, nblas, roots);
        refit_passes(n, flags, sc, L, boxes, s, n > 1 ? nblas : nullptr);
    }
    fill_kernel<<<1, 128, 0, s>>>(levels, 128, 0u);
    batch_header_kernel<<<(nbatch + 63) / 64, 64, 0, s>>>(as_ptr, binfo, nbatch, roots, n, flags, nboxes,
                                                          (const float*)(sc + L.cost), levels,
                                                          (uint2*)(sc + L.tasks0), (uint32_t*)(sc + L.tblas0));
    if (n > 0) {
        for (int lv = 0; lv < kMaxLevels32; ++lv) {
            uint2* tin = (uint2*)(sc + ((lv & 1) ? L.tasks1 : L.tasks0));
            uint2* tout = (uint2*)(sc + ((lv & 1) ? L.tasks0 : L.tasks1));
            uint32_t* bin = (uint32_t*)(sc + ((lv & 1) ? L.tblas1 : L.tblas0));
            uint32_t* bout = (uint32_t*)(sc + ((lv & 1) ? L.tblas0 : L.tblas1));
            collapse_level_kernel<true><<<kCollapseBlocks, 256, 0, s>>>(n, left, right, vals, nboxes,
                                                                  (const uint4*)(sc + L.leaves), nullptr, nullptr,
                                                                  levels, lv, tin, tout, as_ptr, bin, bout,
                                                                  n > 1 ? (const uint32_t*)(sc + L.wdec) : nullptr,
                                                                  node_width(flags));
        }
        collapse_check_kernel<<<1, 64, 0, s>>>(nullptr, levels, kMaxLevels32, err_flags);
    }
    if (ngeom) content_kernel<<<(ngeom + 255) / 256, 256, 0, s>>>(gd, ngeom, nullptr, as_ptr);
    batch_finish_kernel<<<(nbatch + 63) / 64, 64, 0, s>>>(as_ptr, nbatch);
    return hipGetLastError();
}

}  // namespace rtx
