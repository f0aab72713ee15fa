// cpu_baseline.cpp — the reference's ray tracing on CPU cores, timed by bench.py as its cpu_baseline.
//
// BASELINE ONLY: bench.py's cpu_baseline leg loads it to time the CPU path beside the GPU; tests/ check that its
// answers agree with the FP64 oracle.  Nothing of the product links it.
//
// The reference traces with the Vulkan driver (vkCmdTraceRaysKHR, main.cpp:1349-1355).  Its CPU form would be Mesa's
// lavapipe, which this image lacks, so this file restates the same work the way a CPU driver runs it, in FP32:
//   - raygen / closest-hit / miss of main.cpp:1033-1091 (the same ray model and colours as rtx_cmd_trace_rays);
//   - acceleration structures: a binary BVH per BLAS and one over the TLAS instances, built top-down with binned SAH
//     (16 bins along the widest centroid axis, leaves of at most 4 triangles), FP32 boxes;
//   - traversal: near-child-first stack traversal, ray transformed into object space with the FP32 inverse instance
//     transform, watertight ray-triangle test (Woop, Benthin & Wald 2013, FP64 recomputation of exact-zero edge
//     functions), closest hit or any hit;
//   - the path-tracing estimator of rtx_cmd_render (Lambertian, next-event estimation toward one light point, Russian
//     roulette, cosine continuation, pcg4d counter-based random numbers), in FP32.
// Rays are spread over std::thread workers in chunks from an atomic counter.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <thread>
#include <vector>

namespace {

constexpr int kBins = 16;
constexpr uint32_t kLeafMax = 4;

struct Node {
    float lo[3], hi[3];
    uint32_t first;  // internal: left child (right = first + 1); leaf: first primitive
    uint32_t count;  // 0: internal
};

struct Tri {
    float v0[3], v1[3], v2[3];
    uint32_t prim, geom;
};

struct Bvh {
    std::vector<Node> nodes;
    std::vector<uint32_t> order;  // leaf primitive order
};

struct Blas {
    Bvh bvh;
    std::vector<Tri> tris;  // in leaf order
};

struct Inst {
    float inv[12];   // world -> object
    float xf[12];    // object -> world
    int blas;
    uint32_t custom, sbt_offset, mask;
};

struct Tlas {
    Bvh bvh;
    std::vector<Inst> inst;  // in leaf order
    std::vector<uint32_t> id;  // instance index of each leaf slot
};

std::vector<std::unique_ptr<Blas>> g_blas;
std::vector<std::unique_ptr<Tlas>> g_tlas;

inline float area(const float* lo, const float* hi) {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return dx < 0 ? 0.0f : 2.0f * (dx * dy + dy * dz + dz * dx);
}

// Top-down binned-SAH build over primitive boxes (6 floats each).
void build_bvh(const std::vector<float>& boxes, Bvh& out) {
    const uint32_t n = (uint32_t)(boxes.size() / 6);
    out.nodes.clear();
    out.order.resize(n);
    for (uint32_t i = 0; i < n; ++i) out.order[i] = i;
    std::vector<float> cen(3 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) cen[3 * i + a] = 0.5f * (boxes[6 * i + a] + boxes[6 * i + 3 + a]);
    out.nodes.reserve(2 * (size_t)std::max(n, 1u));
    out.nodes.push_back(Node{});
    struct Task {
        uint32_t node, begin, end;
    };
    std::vector<Task> stack{{0, 0, n}};
    while (!stack.empty()) {
        const Task t = stack.back();
        stack.pop_back();
        Node nd;
        float clo[3], chi[3];
        for (int a = 0; a < 3; ++a) {
            nd.lo[a] = clo[a] = std::numeric_limits<float>::infinity();
            nd.hi[a] = chi[a] = -std::numeric_limits<float>::infinity();
        }
        for (uint32_t i = t.begin; i < t.end; ++i) {
            const uint32_t p = out.order[i];
            for (int a = 0; a < 3; ++a) {
                nd.lo[a] = std::min(nd.lo[a], boxes[6 * p + a]);
                nd.hi[a] = std::max(nd.hi[a], boxes[6 * p + 3 + a]);
                clo[a] = std::min(clo[a], cen[3 * p + a]);
                chi[a] = std::max(chi[a], cen[3 * p + a]);
            }
        }
        const uint32_t cnt = t.end - t.begin;
        nd.first = t.begin;
        nd.count = cnt;
        if (cnt <= kLeafMax) {
            out.nodes[t.node] = nd;
            continue;
        }
        int axis = 0;
        for (int a = 1; a < 3; ++a)
            if (chi[a] - clo[a] > chi[axis] - clo[axis]) axis = a;
        const float ext = chi[axis] - clo[axis];
        uint32_t mid = t.begin + cnt / 2;
        if (ext > 0.0f) {
            uint32_t bc[kBins] = {};
            float blo[kBins][3], bhi[kBins][3];
            for (int b = 0; b < kBins; ++b)
                for (int a = 0; a < 3; ++a) {
                    blo[b][a] = std::numeric_limits<float>::infinity();
                    bhi[b][a] = -std::numeric_limits<float>::infinity();
                }
            const float scale = (float)kBins / ext;
            auto bin_of = [&](uint32_t p) {
                const int b = (int)((cen[3 * p + axis] - clo[axis]) * scale);
                return std::min(std::max(b, 0), kBins - 1);
            };
            for (uint32_t i = t.begin; i < t.end; ++i) {
                const uint32_t p = out.order[i];
                const int b = bin_of(p);
                ++bc[b];
                for (int a = 0; a < 3; ++a) {
                    blo[b][a] = std::min(blo[b][a], boxes[6 * p + a]);
                    bhi[b][a] = std::max(bhi[b][a], boxes[6 * p + 3 + a]);
                }
            }
            // sweep: cost of splitting after bin s = A(left) n(left) + A(right) n(right)
            float rarea[kBins];
            uint32_t rcnt[kBins];
            float lo[3], hi[3];
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::numeric_limits<float>::infinity();
                hi[a] = -std::numeric_limits<float>::infinity();
            }
            uint32_t c = 0;
            for (int b = kBins - 1; b > 0; --b) {
                for (int a = 0; a < 3; ++a) {
                    lo[a] = std::min(lo[a], blo[b][a]);
                    hi[a] = std::max(hi[a], bhi[b][a]);
                }
                c += bc[b];
                rarea[b] = area(lo, hi);
                rcnt[b] = c;
            }
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::numeric_limits<float>::infinity();
                hi[a] = -std::numeric_limits<float>::infinity();
            }
            c = 0;
            float best = std::numeric_limits<float>::infinity();
            int best_s = -1;
            for (int s = 0; s < kBins - 1; ++s) {
                for (int a = 0; a < 3; ++a) {
                    lo[a] = std::min(lo[a], blo[s][a]);
                    hi[a] = std::max(hi[a], bhi[s][a]);
                }
                c += bc[s];
                if (c == 0 || rcnt[s + 1] == 0) continue;
                const float cost = area(lo, hi) * (float)c + rarea[s + 1] * (float)rcnt[s + 1];
                if (cost < best) {
                    best = cost;
                    best_s = s;
                }
            }
            if (best_s >= 0) {
                auto it = std::partition(out.order.begin() + t.begin, out.order.begin() + t.end,
                                         [&](uint32_t p) { return bin_of(p) <= best_s; });
                mid = (uint32_t)(it - out.order.begin());
            }
        }
        if (mid == t.begin || mid == t.end) mid = t.begin + cnt / 2;
        const uint32_t left = (uint32_t)out.nodes.size();
        out.nodes.push_back(Node{});
        out.nodes.push_back(Node{});
        nd.first = left;
        nd.count = 0;
        out.nodes[t.node] = nd;
        stack.push_back({left + 1, mid, t.end});
        stack.push_back({left, t.begin, mid});
    }
}

// Ray in one space with the watertight test's shear (Woop, Benthin & Wald 2013).
struct Ray {
    float o[3], d[3], inv[3];
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

inline void setup_ray(Ray& r) {
    for (int a = 0; a < 3; ++a) r.inv[a] = 1.0f / r.d[a];
    const float ax = std::fabs(r.d[0]), ay = std::fabs(r.d[1]), az = std::fabs(r.d[2]);
    r.kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    r.kx = r.kz == 2 ? 0 : r.kz + 1;
    r.ky = r.kx == 2 ? 0 : r.kx + 1;
    if (r.d[r.kz] < 0.0f) std::swap(r.kx, r.ky);
    r.Sx = r.d[r.kx] / r.d[r.kz];
    r.Sy = r.d[r.ky] / r.d[r.kz];
    r.Sz = 1.0f / r.d[r.kz];
}

// Slab test against [tmin, tmax]; the far distance is widened by 2 ulps-ish (1 + 2^-22) so that FP32 rounding never
// culls a box the ray touches.
inline bool hit_box(const Node& n, const Ray& r, float tmin, float tmax, float& tnear) {
    float t0 = tmin, t1 = tmax;
    for (int a = 0; a < 3; ++a) {
        float ta = (n.lo[a] - r.o[a]) * r.inv[a], tb = (n.hi[a] - r.o[a]) * r.inv[a];
        if (ta > tb) std::swap(ta, tb);
        tb *= 1.0000003f;
        t0 = ta > t0 ? ta : t0;
        t1 = tb < t1 ? tb : t1;
        if (!(t0 <= t1)) return false;
    }
    tnear = t0;
    return true;
}

inline bool hit_tri(const Tri& tr, const Ray& r, float tmin, float tmax, float& t, float& u, float& v) {
    float A[3], B[3], C[3];
    for (int a = 0; a < 3; ++a) {
        A[a] = tr.v0[a] - r.o[a];
        B[a] = tr.v1[a] - r.o[a];
        C[a] = tr.v2[a] - r.o[a];
    }
    const float Ax = A[r.kx] - r.Sx * A[r.kz], Ay = A[r.ky] - r.Sy * A[r.kz];
    const float Bx = B[r.kx] - r.Sx * B[r.kz], By = B[r.ky] - r.Sy * B[r.kz];
    const float Cx = C[r.kx] - r.Sx * C[r.kz], Cy = C[r.ky] - r.Sy * C[r.kz];
    float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * By - (double)Cy * Bx);
        V = (float)((double)Ax * Cy - (double)Ay * Cx);
        W = (float)((double)Bx * Ay - (double)By * Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    const float Az = r.Sz * A[r.kz], Bz = r.Sz * B[r.kz], Cz = r.Sz * C[r.kz];
    float T = U * Az + V * Bz + W * Cz;
    if (det < 0.0f) {
        T = -T;
        det = -det;
        U = -U;
        V = -V;
        W = -W;
    }
    if (T < tmin * det || T > tmax * det) return false;
    const float rd = 1.0f / det;
    t = T * rd;
    u = V * rd;
    v = W * rd;
    return true;
}

struct Hit {
    float t = 0, u = 0, v = 0;
    uint32_t prim = 0, geom = 0, inst = 0xFFFFFFFFu;
    const Inst* in = nullptr;
    const Tri* tri = nullptr;
};

// Closest hit (any_hit: the first one) of world ray (o, d) in [tmin, tmax].
bool trace(const Tlas& tl, const float* o, const float* d, float tmin, float tmax, bool any_hit, Hit& h) {
    if (tl.bvh.nodes.empty() || tl.inst.empty()) return false;
    Ray w;
    for (int a = 0; a < 3; ++a) {
        w.o[a] = o[a];
        w.d[a] = d[a];
        w.inv[a] = 1.0f / d[a];
    }
    float tbest = tmax;
    bool found = false;
    uint32_t stack[64];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const Node& n = tl.bvh.nodes[stack[--sp]];
        float tn;
        if (!hit_box(n, w, tmin, tbest, tn)) continue;
        if (n.count == 0) {
            const Node& l = tl.bvh.nodes[n.first];
            const Node& r = tl.bvh.nodes[n.first + 1];
            float tl_, tr_;
            const bool hl = hit_box(l, w, tmin, tbest, tl_), hr = hit_box(r, w, tmin, tbest, tr_);
            if (hl && hr) {
                const bool lf = tl_ <= tr_;
                stack[sp++] = lf ? n.first + 1 : n.first;
                stack[sp++] = lf ? n.first : n.first + 1;
            } else if (hl) {
                stack[sp++] = n.first;
            } else if (hr) {
                stack[sp++] = n.first + 1;
            }
            continue;
        }
        for (uint32_t k = 0; k < n.count; ++k) {
            const Inst& in = tl.inst[n.first + k];
            if ((in.mask & 0xFFu) == 0) continue;
            const Blas& b = *g_blas[in.blas];
            if (b.bvh.nodes.empty()) continue;
            Ray r;
            for (int a = 0; a < 3; ++a) {
                r.o[a] = in.inv[4 * a] * o[0] + in.inv[4 * a + 1] * o[1] + in.inv[4 * a + 2] * o[2] + in.inv[4 * a + 3];
                r.d[a] = in.inv[4 * a] * d[0] + in.inv[4 * a + 1] * d[1] + in.inv[4 * a + 2] * d[2];
            }
            setup_ray(r);
            uint32_t bs[64];
            int bsp = 0;
            bs[bsp++] = 0;
            while (bsp) {
                const Node& m = b.bvh.nodes[bs[--bsp]];
                float tm;
                if (!hit_box(m, r, tmin, tbest, tm)) continue;
                if (m.count == 0) {
                    const Node& l = b.bvh.nodes[m.first];
                    const Node& rr = b.bvh.nodes[m.first + 1];
                    float a0, a1;
                    const bool hl = hit_box(l, r, tmin, tbest, a0), hr = hit_box(rr, r, tmin, tbest, a1);
                    if (hl && hr) {
                        const bool lf = a0 <= a1;
                        bs[bsp++] = lf ? m.first + 1 : m.first;
                        bs[bsp++] = lf ? m.first : m.first + 1;
                    } else if (hl) {
                        bs[bsp++] = m.first;
                    } else if (hr) {
                        bs[bsp++] = m.first + 1;
                    }
                    continue;
                }
                for (uint32_t j = 0; j < m.count; ++j) {
                    const Tri& tr = b.tris[m.first + j];
                    float t, u, v;
                    if (!hit_tri(tr, r, tmin, tbest, t, u, v)) continue;
                    found = true;
                    tbest = t;
                    h.t = t;
                    h.u = u;
                    h.v = v;
                    h.prim = tr.prim;
                    h.geom = tr.geom;
                    h.inst = tl.id[n.first + k];
                    h.in = &in;
                    h.tri = &tr;
                    if (any_hit) return true;
                }
            }
        }
    }
    return found;
}

template <typename F>
void parallel_for(int64_t n, int64_t chunk, int threads, F f) {
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::atomic<int64_t> next{0};
    auto work = [&]() {
        for (;;) {
            const int64_t s = next.fetch_add(chunk);
            if (s >= n) break;
            const int64_t e = std::min(n, s + chunk);
            for (int64_t i = s; i < e; ++i) f(i);
        }
    };
    std::vector<std::thread> pool;
    for (int k = 1; k < threads; ++k) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
}

void invert3x4(const double* m, double* inv) {
    const double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], i = m[10];
    const double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    const double det = a * A + b * B + c * C;
    const double r = det != 0.0 ? 1.0 / det : 0.0;
    const double R[9] = {A * r, -(b * i - c * h) * r, (b * f - c * e) * r,
                         B * r, (a * i - c * g) * r, -(a * f - c * d) * r,
                         C * r, -(a * h - b * g) * r, (a * e - b * d) * r};
    for (int row = 0; row < 3; ++row) {
        for (int col = 0; col < 3; ++col) inv[4 * row + col] = R[3 * row + col];
        inv[4 * row + 3] = -(R[3 * row] * m[3] + R[3 * row + 1] * m[7] + R[3 * row + 2] * m[11]);
    }
}

uint32_t pcg4d_x(uint32_t x, uint32_t y, uint32_t z, uint32_t w, uint32_t* out) {
    x = x * 1664525u + 1013904223u;
    y = y * 1664525u + 1013904223u;
    z = z * 1664525u + 1013904223u;
    w = w * 1664525u + 1013904223u;
    x += y * w;
    y += z * x;
    z += x * y;
    w += y * z;
    x ^= x >> 16;
    y ^= y >> 16;
    z ^= z >> 16;
    w ^= w >> 16;
    x += y * w;
    y += z * x;
    z += x * y;
    w += y * z;
    out[0] = x;
    out[1] = y;
    out[2] = z;
    out[3] = w;
    return x;
}
inline float unit(uint32_t v) { return (float)(v >> 8) * (1.0f / 16777216.0f); }

// rtx_hit_record (48 B): program, color[3], emission[3], match[4], any_hit
struct Record {
    uint32_t program;
    float color[3];
    float emission[3];
    int32_t match[4];
    uint32_t any_hit;
};
static_assert(sizeof(Record) == 48, "record layout");

// A BLAS over geometries [g0, g1) of the per-geometry arrays (vertex / index pointers, triangle counts, optional
// 3x4 row-major transform applied to the vertices in the FP32 order ((m0 x + m1 y) + m2 z) + m3).
std::unique_ptr<Blas> make_blas(int g0, int g1, const uint64_t* vptr, const uint64_t* iptr, const uint32_t* counts,
                                const uint64_t* xptr) {
    std::vector<Tri> tris;
    for (int g = g0; g < g1; ++g) {
        const float* V = (const float*)(uintptr_t)vptr[g];
        const uint32_t* I = (const uint32_t*)(uintptr_t)iptr[g];
        const float* M = (const float*)(uintptr_t)xptr[g];
        for (uint32_t p = 0; p < counts[g]; ++p) {
            Tri t;
            float* dst[3] = {t.v0, t.v1, t.v2};
            for (int c = 0; c < 3; ++c) {
                const float* s = V + 3 * (size_t)I[3 * p + c];
                for (int a = 0; a < 3; ++a)
                    dst[c][a] = M ? ((M[4 * a] * s[0] + M[4 * a + 1] * s[1]) + M[4 * a + 2] * s[2]) + M[4 * a + 3]
                                  : s[a];
            }
            t.prim = p;
            t.geom = (uint32_t)(g - g0);
            tris.push_back(t);
        }
    }
    auto b = std::make_unique<Blas>();
    std::vector<float> boxes(6 * tris.size());
    for (size_t i = 0; i < tris.size(); ++i)
        for (int a = 0; a < 3; ++a) {
            boxes[6 * i + a] = std::min({tris[i].v0[a], tris[i].v1[a], tris[i].v2[a]});
            boxes[6 * i + 3 + a] = std::max({tris[i].v0[a], tris[i].v1[a], tris[i].v2[a]});
        }
    build_bvh(boxes, b->bvh);
    b->tris.resize(tris.size());
    for (size_t i = 0; i < tris.size(); ++i) b->tris[i] = tris[b->bvh.order[i]];
    return b;
}

}  // namespace

extern "C" {

void cb_reset() {
    g_blas.clear();
    g_tlas.clear();
}

// One BLAS over ngeom triangle geometries (vertex / index pointers, triangle counts, optional 3x4 transforms);
// returns its handle.
int cb_blas_create(int ngeom, const uint64_t* vptr, const uint64_t* iptr, const uint32_t* counts,
                   const uint64_t* xptr) {
    g_blas.push_back(make_blas(0, ngeom, vptr, iptr, counts, xptr));
    return (int)g_blas.size() - 1;
}

// nblas BLASes at once, one thread each: BLAS k over geometries [geom_first[k], geom_first[k + 1]) of the arrays.
void cb_blas_create_many(int nblas, const int* geom_first, const uint64_t* vptr, const uint64_t* iptr,
                         const uint32_t* counts, const uint64_t* xptr, int* handles) {
    std::vector<std::unique_ptr<Blas>> built(nblas);
    std::vector<std::thread> pool;
    for (int k = 0; k < nblas; ++k)
        pool.emplace_back([&, k]() { built[k] = make_blas(geom_first[k], geom_first[k + 1], vptr, iptr, counts, xptr); });
    for (auto& th : pool) th.join();
    for (int k = 0; k < nblas; ++k) {
        g_blas.push_back(std::move(built[k]));
        handles[k] = (int)g_blas.size() - 1;
    }
}

// TLAS over n 64-B instance records (VkAccelerationStructureInstanceKHR layout, main.cpp:848-858) whose address
// field holds a cb BLAS handle; returns its handle.
int cb_tlas_create(int n, const void* records) {
    auto tl = std::make_unique<Tlas>();
    std::vector<Inst> inst;
    std::vector<float> boxes;
    const uint8_t* rp = (const uint8_t*)records;
    for (int i = 0; i < n; ++i) {
        float xf[12];
        uint32_t cm, sf;
        uint64_t addr;
        std::memcpy(xf, rp + 64 * (size_t)i, 48);
        std::memcpy(&cm, rp + 64 * (size_t)i + 48, 4);
        std::memcpy(&sf, rp + 64 * (size_t)i + 52, 4);
        std::memcpy(&addr, rp + 64 * (size_t)i + 56, 8);
        Inst in;
        double m[12], inv[12];
        for (int k = 0; k < 12; ++k) {
            m[k] = xf[k];
            in.xf[k] = xf[k];
        }
        invert3x4(m, inv);
        for (int k = 0; k < 12; ++k) in.inv[k] = (float)inv[k];
        in.blas = (int)addr;
        in.custom = cm & 0xFFFFFFu;
        in.mask = cm >> 24;
        in.sbt_offset = sf & 0xFFFFFFu;
        const Blas& b = *g_blas[in.blas];
        if (b.bvh.nodes.empty()) continue;
        const Node& r = b.bvh.nodes[0];
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int c = 0; c < 8; ++c) {
            const double p[3] = {(c & 1) ? r.hi[0] : r.lo[0], (c & 2) ? r.hi[1] : r.lo[1], (c & 4) ? r.hi[2] : r.lo[2]};
            for (int a = 0; a < 3; ++a) {
                const double w = m[4 * a] * p[0] + m[4 * a + 1] * p[1] + m[4 * a + 2] * p[2] + m[4 * a + 3];
                lo[a] = std::min(lo[a], w);
                hi[a] = std::max(hi[a], w);
            }
        }
        for (int a = 0; a < 3; ++a) boxes.push_back(std::nextafter((float)lo[a], -INFINITY));
        for (int a = 0; a < 3; ++a) boxes.push_back(std::nextafter((float)hi[a], INFINITY));
        inst.push_back(in);
        tl->id.push_back((uint32_t)i);
    }
    build_bvh(boxes, tl->bvh);
    tl->inst.resize(inst.size());
    std::vector<uint32_t> id(inst.size());
    for (size_t i = 0; i < inst.size(); ++i) {
        tl->inst[i] = inst[tl->bvh.order[i]];
        id[i] = tl->id[tl->bvh.order[i]];
    }
    tl->id = id;
    g_tlas.push_back(std::move(tl));
    return (int)g_tlas.size() - 1;
}

// Closest-hit (any_hit = 0) or any-hit queries of n rays (32-B records: ox oy oz tmin dx dy dz tmax).  Closest hit
// writes hits[6 i ..] = t, u, v (float bits), prim, geom, inst (0xFFFFFFFF miss) like rtx_hit; any hit writes
// occluded[i].
void cb_trace(int tlas, const float* rays, int64_t n, int any_hit, uint32_t* hits, uint32_t* occluded, int threads) {
    const Tlas& tl = *g_tlas[tlas];
    parallel_for(n, 1024, threads, [&](int64_t i) {
        const float* r = rays + 8 * i;
        Hit h;
        const bool f = trace(tl, r, r + 4, r[3], r[7], any_hit != 0, h);
        if (any_hit) {
            occluded[i] = f ? 1u : 0u;
            return;
        }
        uint32_t* o = hits + 6 * i;
        std::memcpy(o, &h.t, 4);
        std::memcpy(o + 1, &h.u, 4);
        std::memcpy(o + 2, &h.v, 4);
        o[3] = f ? h.prim : 0u;
        o[4] = f ? h.geom : 0u;
        o[5] = f ? h.inst : 0xFFFFFFFFu;
    });
}

// vkCmdTraceRaysKHR(W, H) with the reference raygen / closest-hit / miss (main.cpp:1033-1091) into an RGBA32F film:
// aspect_y = tan(radians(fov) / 2) (FP64, rounded), aspect_x = aspect_y * W / H, ndc = (id + 0.5) / size * 2 - 1,
// dir = (ndc.x ax, -ndc.y ay, -1); record = instance SBT offset + geometry x stride + offset; programs as
// rtx_hit_record.program (record colour, barycentric, barycentric when the match fields all agree); film value
// vec4(hitValue, 0).
void cb_trace_rays(int tlas, const void* records, uint32_t nrec, uint32_t sbt_offset, uint32_t sbt_stride,
                   const float* miss, const float* cam, uint32_t W, uint32_t H, float tmin, float tmax, float* film,
                   int threads) {
    const Tlas& tl = *g_tlas[tlas];
    const Record* rec = (const Record*)records;
    const double rad = (double)cam[3] * (3.14159265358979323846 / 180.0);
    const float ay = (float)std::tan(rad * 0.5);
    const float ax = ay * (float)W / (float)H;
    parallel_for((int64_t)W * H, 1024, threads, [&](int64_t p) {
        const uint32_t x = (uint32_t)(p % W), y = (uint32_t)(p / W);
        const float nx = ((float)x + 0.5f) / (float)W * 2.0f - 1.0f;
        const float ny = ((float)y + 0.5f) / (float)H * 2.0f - 1.0f;
        const float d[3] = {nx * ax, -(ny * ay), -1.0f};
        Hit h;
        float c[3] = {miss[0], miss[1], miss[2]};
        if (trace(tl, cam, d, tmin, tmax, false, h)) {
            const uint32_t r = h.in->sbt_offset + h.geom * sbt_stride + sbt_offset;
            if (r < nrec) {
                const Record& R = rec[r];
                bool bary = R.program == 1;
                if (R.program == 2)
                    bary = (int32_t)h.prim == R.match[0] && (int32_t)h.inst == R.match[1] &&
                           (int32_t)h.in->custom == R.match[2] && (int32_t)h.geom == R.match[3];
                for (int k = 0; k < 3; ++k) c[k] = R.color[k];
                if (bary) {
                    c[0] = 1.0f - h.u - h.v;
                    c[1] = h.u;
                    c[2] = h.v;
                }
            } else {
                c[0] = c[1] = c[2] = 0.0f;
            }
        }
        float* o = film + 4 * p;
        o[0] = c[0];
        o[1] = c[1];
        o[2] = c[2];
        o[3] = 0.0f;
    });
}

// The path-tracing estimator of rtx_cmd_render in FP32 (pixel jitter pcg4d(x, y, s, seed ^ 0xA511E9B3); per vertex
// d pcg4d(x, y, s, seed + 2d) and (..., seed + 2d + 1); emission seen at vertex 0, NEE toward one light point chosen by
// area, Russian roulette when continuing from d >= rr_depth, cosine-weighted continuation).  Adds each pixel's
// radiance sums to accum (3 floats per pixel); returns the rays traced (camera, extension and shadow rays).
uint64_t cb_render(int tlas, const void* records, uint32_t nrec, const float* cam, uint32_t W, uint32_t H,
                   uint32_t spp, uint32_t sample_offset, uint32_t max_depth, uint32_t rr_depth, uint32_t seed,
                   const float* lights, uint32_t nlights, float ray_offset, float* accum, int threads) {
    const Tlas& tl = *g_tlas[tlas];
    const Record* rec = (const Record*)records;
    const double radn = (double)cam[3] * (3.14159265358979323846 / 180.0);
    const float ay = (float)std::tan(radn * 0.5);
    const float ax = ay * (float)W / (float)H;
    std::vector<float> cdf(nlights + 1, 0.0f);
    float acc = 0.0f;
    for (uint32_t i = 0; i < nlights; ++i) {
        const float* L = lights + 12 * i;
        const float e1[3] = {L[3] - L[0], L[4] - L[1], L[5] - L[2]}, e2[3] = {L[6] - L[0], L[7] - L[1], L[8] - L[2]};
        const float c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        acc += 0.5f * std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        cdf[i] = acc;
    }
    cdf[nlights] = acc;
    std::atomic<uint64_t> rays{0};
    parallel_for((int64_t)W * H, 64, threads, [&](int64_t p) {
        const uint32_t x = (uint32_t)(p % W), y = (uint32_t)(p / W);
        uint64_t nr = 0;
        float sum[3] = {0, 0, 0};
        for (uint32_t si = 0; si < spp; ++si) {
            const uint32_t s = sample_offset + si;
            uint32_t j[4], ra[4], rb[4];
            pcg4d_x(x, y, s, seed ^ 0xA511E9B3u, j);
            const float nx = ((float)x + unit(j[0])) / (float)W * 2.0f - 1.0f;
            const float ny = ((float)y + unit(j[1])) / (float)H * 2.0f - 1.0f;
            float o[3] = {cam[0], cam[1], cam[2]};
            float d[3] = {nx * ax, -(ny * ay), -1.0f};
            float thr[3] = {1, 1, 1}, rad[3] = {0, 0, 0};
            ++nr;
            for (uint32_t depth = 0; depth < max_depth; ++depth) {
                Hit h;
                if (!trace(tl, o, d, 0.0f, INFINITY, false, h)) break;
                const uint32_t ri = h.in->sbt_offset + h.geom;
                if (ri >= nrec) break;
                const Record& R = rec[ri];
                // world geometric normal: object normal through the inverse transpose of the instance transform
                const Tri& tr = *h.tri;
                const float e1[3] = {tr.v1[0] - tr.v0[0], tr.v1[1] - tr.v0[1], tr.v1[2] - tr.v0[2]};
                const float e2[3] = {tr.v2[0] - tr.v0[0], tr.v2[1] - tr.v0[1], tr.v2[2] - tr.v0[2]};
                const float no[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                                     e1[0] * e2[1] - e1[1] * e2[0]};
                float n[3];
                for (int c = 0; c < 3; ++c)
                    n[c] = h.in->inv[c] * no[0] + h.in->inv[4 + c] * no[1] + h.in->inv[8 + c] * no[2];
                const float nl = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
                for (int c = 0; c < 3; ++c) n[c] = nl > 0 ? n[c] / nl : 0.0f;
                const bool front = n[0] * d[0] + n[1] * d[1] + n[2] * d[2] < 0.0f;
                const bool lambert = R.program == 3;
                if (depth == 0 && front && lambert)
                    for (int c = 0; c < 3; ++c) rad[c] += thr[c] * R.emission[c];
                const float nf[3] = {front ? n[0] : -n[0], front ? n[1] : -n[1], front ? n[2] : -n[2]};
                float org[3];
                for (int c = 0; c < 3; ++c) org[c] = o[c] + h.t * d[c] + ray_offset * nf[c];
                const float* alb = R.color;
                const bool reflective = alb[0] > 0 || alb[1] > 0 || alb[2] > 0;
                if (!reflective) break;
                pcg4d_x(x, y, s, seed + 2u * depth, ra);
                pcg4d_x(x, y, s, seed + 2u * depth + 1u, rb);
                if (nlights > 0) {
                    const float total = cdf[nlights];
                    const float ut = unit(ra[0]) * total;
                    uint32_t lo = 0, hi = nlights - 1;
                    while (lo < hi) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (cdf[mid] > ut) hi = mid; else lo = mid + 1;
                    }
                    const float* L = lights + 12 * lo;
                    const float su = std::sqrt(unit(ra[1]));
                    const float b0 = 1.0f - su, b1 = unit(ra[2]) * su, b2 = 1.0f - b0 - b1;
                    float lp[3], f1[3], f2[3];
                    for (int c = 0; c < 3; ++c) {
                        lp[c] = b0 * L[c] + b1 * L[3 + c] + b2 * L[6 + c];
                        f1[c] = L[3 + c] - L[c];
                        f2[c] = L[6 + c] - L[c];
                    }
                    const float ln[3] = {f1[1] * f2[2] - f1[2] * f2[1], f1[2] * f2[0] - f1[0] * f2[2],
                                         f1[0] * f2[1] - f1[1] * f2[0]};
                    const float lnl = std::sqrt(ln[0] * ln[0] + ln[1] * ln[1] + ln[2] * ln[2]);
                    const float wi[3] = {lp[0] - org[0], lp[1] - org[1], lp[2] - org[2]};
                    const float d2 = wi[0] * wi[0] + wi[1] * wi[1] + wi[2] * wi[2];
                    if (lnl > 0 && d2 > 0) {
                        const float id = 1.0f / std::sqrt(d2);
                        const float cx = (nf[0] * wi[0] + nf[1] * wi[1] + nf[2] * wi[2]) * id;
                        const float cl = -(ln[0] * wi[0] + ln[1] * wi[1] + ln[2] * wi[2]) * id / lnl;
                        if (cx > 0 && cl > 0) {
                            const float g = cx * cl / d2 * total * 0.31830988618379067154f;
                            const float ct[3] = {thr[0] * alb[0] * L[9] * g, thr[1] * alb[1] * L[10] * g,
                                                 thr[2] * alb[2] * L[11] * g};
                            if (ct[0] > 0 || ct[1] > 0 || ct[2] > 0) {
                                ++nr;
                                Hit sh;
                                if (!trace(tl, org, wi, 0.0f, 1.0f - 1e-4f, true, sh))
                                    for (int c = 0; c < 3; ++c) rad[c] += ct[c];
                            }
                        }
                    }
                }
                if (depth + 1 >= max_depth) break;
                float tn[3] = {thr[0] * alb[0], thr[1] * alb[1], thr[2] * alb[2]};
                if (depth >= rr_depth) {
                    const float pr = std::min(0.95f, std::max(tn[0], std::max(tn[1], tn[2])));
                    if (unit(ra[3]) >= pr) break;
                    for (int c = 0; c < 3; ++c) tn[c] /= pr;
                }
                // cosine-weighted continuation in the basis of Duff et al. 2017
                const float sign = std::copysign(1.0f, nf[2]);
                const float a = -1.0f / (sign + nf[2]);
                const float bb = nf[0] * nf[1] * a;
                const float t[3] = {1.0f + sign * nf[0] * nf[0] * a, sign * bb, -sign * nf[0]};
                const float b[3] = {bb, sign + nf[1] * nf[1] * a, -nf[1]};
                const float u1 = unit(rb[0]), u2 = unit(rb[1]);
                const float r = std::sqrt(u1), phi = 6.28318530717958647692f * u2;
                const float lx = r * std::cos(phi), ly = r * std::sin(phi), lz = std::sqrt(std::max(0.0f, 1.0f - u1));
                for (int c = 0; c < 3; ++c) {
                    d[c] = lx * t[c] + ly * b[c] + lz * nf[c];
                    o[c] = org[c];
                    thr[c] = tn[c];
                }
                ++nr;
            }
            for (int c = 0; c < 3; ++c) sum[c] += rad[c];
        }
        for (int c = 0; c < 3; ++c) accum[3 * p + c] += sum[c];
        rays.fetch_add(nr, std::memory_order_relaxed);
    });
    return rays.load();
}

}  // extern "C"
