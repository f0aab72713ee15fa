// rtx_oracle.cpp — FP64 CPU oracle for the ray-tracing hot path.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this
// library, as the checker; the product (build-up-phase_amd/librtx.so) never links or calls it.
//
// What it restates (reference = vulkan-raytraced-triangle/main.cpp):
//   * raygen (main.cpp:1033-1055): pinhole camera, unnormalised direction, tmin 0 / tmax 100, film = vec4(hit, 0)
//   * miss (main.cpp:1063-1066) and closest-hit (main.cpp:1080-1091) programs, SBT record indexing
//     instanceSBTOffset + geometryIndex * stride + offset (main.cpp:1252-1263, 1307-1317)
//   * the driver's traversal semantics (traceRayEXT call site main.cpp:1047-1052; the driver itself is closed):
//     closest opaque triangle with t in [tmin, tmax], per-geometry 3x4 transforms applied at build time
//     (main.cpp:684-695, 795-804), per-instance object-to-world 3x4 transforms (main.cpp:835-858), instance
//     mask & cull mask, barycentrics (u, v) with P = (1-u-v) v0 + u v1 + v v2, IDs gl_PrimitiveID /
//     gl_GeometryIndexEXT / gl_InstanceID / gl_InstanceCustomIndexEXT.
//     Intersection is evaluated in FP64 (Woop/Benthin/Wald 2013 edge functions in a sheared frame).
//   * LBVH build steps: 30/63-bit Morton codes of triangle-box centroids, stable key sort, Karras 2012
//     hierarchy emission.
//   * PCG32 (O'Neill 2014) and the pcg4d hash (Jarzynski & Olano 2020).
//
// Epsilon-ties.  The GPU decides hits in FP32; this oracle decides in FP64 and, for every ray, reports whether
// the FP64 answer is robust against FP32 rounding.  In object space the GPU works on a ray whose origin has
// been moved to the BLAS bounding box, so all FP32 position errors are bounded by
//     delta = 2^-19 * (R + D)       R = max |coordinate| of the BLAS box, D = its diagonal
// (32 ulps of the largest magnitude involved).  Positions are compared in the sheared 2D frame of the
// watertight test, where a 3D error of size e moves a point by at most 3e.  A ray is an epsilon-tie when
//   - a candidate triangle's 2D point lies within 3*delta of one of its edges, or
//   - two candidates' hit distances differ by less than the sum of their t-uncertainties, or
//   - the closest hit lies within its t-uncertainty of tmin or tmax.
// Everything else must match the GPU bit-exactly in hit IDs.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace {

struct d3 {
    double x, y, z;
};
inline d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline d3 operator*(d3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline d3 cross(d3 a, d3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double comp(d3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
inline double len(d3 a) { return std::sqrt(dot(a, a)); }

constexpr double kEps32 = 5.9604644775390625e-08;  // 2^-24
constexpr double kDeltaScale = 1.0 / 524288.0;     // 2^-19

// ---------------------------------------------------------------------------------------------------------
// Geometry transform at build time.  Defined FP32 operation order (no contraction):
//   x' = ((m00*x + m01*y) + m02*z) + m03
// The GPU builder evaluates the same expression, so transformed vertices are bit-identical.
// ---------------------------------------------------------------------------------------------------------
inline void xform_f32(const float* m, const float* v, float* o) {
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

struct Tri {
    float v[9];  // triangle vertices, or an AABB primitive's min (v[0..2]) and max (v[3..5])
    uint32_t prim, geom;
    bool opaque;  // its geometry has VK_GEOMETRY_OPAQUE_BIT
    bool aabb;    // an AABB primitive (VK_GEOMETRY_TYPE_AABBS_KHR)
};

struct BNode {
    double lo[3], hi[3];
    int32_t left;   // internal: left child index; leaf: -1
    int32_t right;  // internal: right child; leaf: first triangle index
    int32_t count;  // leaf: triangle count; internal: 0
};

struct Blas {
    std::vector<Tri> tris;
    std::vector<uint32_t> order;  // BVH leaf order -> tris index
    std::vector<BNode> nodes;
    double lo[3], hi[3];
    double delta;  // FP32 position uncertainty (object space)
};

struct Inst {
    double m[12];     // object -> world
    double minv[12];  // world -> object
    uint32_t custom, mask, sbt_offset, flags;
    int blas;
    bool valid;
};

struct Tlas {
    std::vector<Inst> inst;
};

std::vector<Blas*> g_blas;
std::vector<Tlas*> g_tlas;

void tri_box(const Tri& t, double* lo, double* hi) {
    if (t.aabb) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = t.v[a];
            hi[a] = t.v[3 + a];
        }
        return;
    }
    for (int a = 0; a < 3; ++a) {
        double x0 = t.v[a], x1 = t.v[3 + a], x2 = t.v[6 + a];
        lo[a] = std::min(x0, std::min(x1, x2));
        hi[a] = std::max(x0, std::max(x1, x2));
    }
}

// Median-split binary BVH, leaf size <= 4.  Boxes are exact FP64 bounds; the traversal inflates them.
int build_rec(Blas& b, std::vector<double>& cen, uint32_t* idx, int n, int depth) {
    int me = (int)b.nodes.size();
    b.nodes.push_back(BNode{});
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    double clo[3] = {1e300, 1e300, 1e300}, chi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < n; ++i) {
        double tl[3], th[3];
        tri_box(b.tris[idx[i]], tl, th);
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], tl[a]);
            hi[a] = std::max(hi[a], th[a]);
            clo[a] = std::min(clo[a], cen[3 * idx[i] + a]);
            chi[a] = std::max(chi[a], cen[3 * idx[i] + a]);
        }
    }
    for (int a = 0; a < 3; ++a) {
        b.nodes[me].lo[a] = lo[a];
        b.nodes[me].hi[a] = hi[a];
    }
    if (n <= 4 || depth > 60) {
        b.nodes[me].left = -1;
        b.nodes[me].right = (int)(idx - b.order.data());
        b.nodes[me].count = n;
        return me;
    }
    int ax = 0;
    double ext = chi[0] - clo[0];
    for (int a = 1; a < 3; ++a)
        if (chi[a] - clo[a] > ext) {
            ext = chi[a] - clo[a];
            ax = a;
        }
    int mid = n / 2;
    std::nth_element(idx, idx + mid, idx + n, [&](uint32_t p, uint32_t q) {
        double cp = cen[3 * p + ax], cq = cen[3 * q + ax];
        return cp < cq || (cp == cq && p < q);
    });
    int l = build_rec(b, cen, idx, mid, depth + 1);
    int r = build_rec(b, cen, idx + mid, n - mid, depth + 1);
    b.nodes[me].left = l;
    b.nodes[me].right = r;
    b.nodes[me].count = 0;
    return me;
}

void invert3x4(const double* m, double* inv) {
    // inverse of [A | t]: [A^-1 | -A^-1 t], A^-1 by cofactors (FP64)
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
    for (int row = 0; row < 3; ++row)
        inv[4 * row + 3] = -(inv[4 * row + 0] * m[3] + inv[4 * row + 1] * m[7] + inv[4 * row + 2] * m[11]);
}

// ---------------------------------------------------------------------------------------------------------
// FP64 ray/triangle evaluation in the sheared frame of the watertight test.
// ---------------------------------------------------------------------------------------------------------
struct RayFrame {
    d3 o, d;
    int kx, ky, kz;
    double Sx, Sy, Sz;
    double dlen;
};

RayFrame make_frame(d3 o, d3 d) {
    RayFrame f;
    f.o = o;
    f.d = d;
    double ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    int kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    if (comp(d, kz) < 0) std::swap(kx, ky);
    f.kx = kx;
    f.ky = ky;
    f.kz = kz;
    f.Sx = comp(d, kx) / comp(d, kz);
    f.Sy = comp(d, ky) / comp(d, kz);
    f.Sz = 1.0 / comp(d, kz);
    f.dlen = len(d);
    return f;
}

struct Cand {
    bool inside;    // FP64 says hit (all margins >= 0, det != 0)
    bool near;      // within 3*delta of an edge
    double margin;  // signed min distance to the edges in the 2D frame (>0 inside)
    double t, u, v;
    double tunc;    // t uncertainty
};

// Returns false when the triangle is far from the ray (margin < -3 delta) or degenerate/parallel.
bool eval_tri(const RayFrame& f, const Tri& tr, double delta, Cand& c) {
    d3 v0{tr.v[0], tr.v[1], tr.v[2]}, v1{tr.v[3], tr.v[4], tr.v[5]}, v2{tr.v[6], tr.v[7], tr.v[8]};
    d3 A = v0 - f.o, B = v1 - f.o, C = v2 - f.o;
    double Ax = comp(A, f.kx) - f.Sx * comp(A, f.kz), Ay = comp(A, f.ky) - f.Sy * comp(A, f.kz);
    double Bx = comp(B, f.kx) - f.Sx * comp(B, f.kz), By = comp(B, f.ky) - f.Sy * comp(B, f.kz);
    double Cx = comp(C, f.kx) - f.Sx * comp(C, f.kz), Cy = comp(C, f.ky) - f.Sy * comp(C, f.kz);
    double U = Cx * By - Cy * Bx;
    double V = Ax * Cy - Ay * Cx;
    double W = Bx * Ay - By * Ax;
    double det = U + V + W;
    if (det == 0.0) return false;
    double s = det > 0 ? 1.0 : -1.0;
    double lbc = std::hypot(Bx - Cx, By - Cy), lca = std::hypot(Cx - Ax, Cy - Ay), lab = std::hypot(Ax - Bx, Ay - By);
    if (lbc == 0 || lca == 0 || lab == 0) return false;
    double mU = s * U / lbc, mV = s * V / lca, mW = s * W / lab;
    double m = std::min(mU, std::min(mV, mW));
    double band = 3.0 * delta;
    if (m < -band) return false;
    double Az = f.Sz * comp(A, f.kz), Bz = f.Sz * comp(B, f.kz), Cz = f.Sz * comp(C, f.kz);
    double T = U * Az + V * Bz + W * Cz;
    c.t = T / det;
    c.u = V / det;
    c.v = W / det;
    c.margin = m;
    c.inside = m >= 0;
    c.near = std::fabs(m) < band;
    // t uncertainty: a position error delta becomes delta / sin(theta) along the ray
    d3 n = cross(v1 - v0, v2 - v0);
    double sn = std::fabs(dot(n, f.d)) / (len(n) * f.dlen);
    sn = std::max(sn, 1e-12);
    c.tunc = (3.0 * delta / sn) / f.dlen + 64.0 * kEps32 * std::fabs(c.t);
    return true;
}

// Intersection programs of AABB primitives (rtx.h RTX_INTERSECTION_*), FP64: the first intersection with t in
// [tmin, tmax] of o + t d with the program's surface, (u, v) and the outward object-space normal.  BOX: the box
// surface, (u, v) across the hit face along the next two axes in cyclic order; SPHERE: the sphere centred in the
// box with radius half the smallest extent, (u, v) = ((n.x + 1) / 2, (n.y + 1) / 2).  `tie` reports decisions
// FP32 could make differently: a grazing sphere, a hit within 1e-5 (relative) of a box face's edge, a box corner.
bool proc_hit(uint32_t prog, const Tri& tr, const d3& o, const d3& d, double tmin, double tmax, double& t, double& u,
              double& v, double* n, bool& tie) {
    const double lo[3] = {tr.v[0], tr.v[1], tr.v[2]}, hi[3] = {tr.v[3], tr.v[4], tr.v[5]};
    const double oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
    tie = false;
    if (prog == 1) {
        double c[3], oc[3], r = std::min(hi[0] - lo[0], std::min(hi[1] - lo[1], hi[2] - lo[2])) * 0.5;
        for (int k = 0; k < 3; ++k) {
            c[k] = (lo[k] + hi[k]) * 0.5;
            oc[k] = oo[k] - c[k];
        }
        const double A = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
        const double B = oc[0] * dd[0] + oc[1] * dd[1] + oc[2] * dd[2];
        const double C = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - r * r;
        const double disc = B * B - A * C;
        if (!(A > 0) || !(r > 0)) return false;
        if (std::fabs(disc) <= 1e-4 * A * r * r) tie = true;  // grazing
        if (disc < 0) return false;
        const double sq = std::sqrt(disc);
        double tt = (-B - sq) / A;
        if (std::fabs(tt - tmin) <= 1e-5 * std::max(std::fabs(tt), std::fabs(tmin)) + 1e-7) tie = true;  // root at tmin
        if (tt < tmin) tt = (-B + sq) / A;
        if (!(tt >= tmin && tt <= tmax)) return false;
        for (int k = 0; k < 3; ++k) n[k] = (oo[k] + tt * dd[k] - c[k]) / r;
        t = tt;
        u = (n[0] + 1) * 0.5;
        v = (n[1] + 1) * 0.5;
        return true;
    }
    double tn = -INFINITY, tf = INFINITY;
    int an = 0, af = 0;
    double ts_near[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < 3; ++k) {
        if (dd[k] == 0) {
            if (oo[k] < lo[k] || oo[k] > hi[k]) return false;
            continue;
        }
        double t1 = (lo[k] - oo[k]) / dd[k], t2 = (hi[k] - oo[k]) / dd[k];
        if (t1 > t2) std::swap(t1, t2);
        ts_near[k] = t1;
        if (t1 > tn) {
            tn = t1;
            an = k;
        }
        if (t2 < tf) {
            tf = t2;
            af = k;
        }
    }
    if (!(tn <= tf)) {
        if (tn <= tf + 1e-5 * (std::fabs(tn) + std::fabs(tf))) tie = true;
        return false;
    }
    if (std::fabs(tn - tmin) <= 1e-5 * std::max(std::fabs(tn), std::fabs(tmin)) + 1e-7) tie = true;  // entry at tmin
    int ax;
    if (tn >= tmin && tn <= tmax) {
        t = tn;
        ax = an;
    } else if (tf >= tmin && tf <= tmax) {
        t = tf;
        ax = af;
    } else {
        return false;
    }
    const int a1 = (ax + 1) % 3, a2 = (ax + 2) % 3;
    const double p1 = oo[a1] + t * dd[a1], p2 = oo[a2] + t * dd[a2];
    u = hi[a1] > lo[a1] ? (p1 - lo[a1]) / (hi[a1] - lo[a1]) : 0;
    v = hi[a2] > lo[a2] ? (p2 - lo[a2]) / (hi[a2] - lo[a2]) : 0;
    const double e = 1e-5;
    if (u < e || u > 1 - e || v < e || v > 1 - e) tie = true;  // near an edge of the hit face
    const double sgn = (t == tn) ? (dd[ax] > 0 ? -1.0 : 1.0) : (dd[ax] > 0 ? 1.0 : -1.0);
    n[0] = ax == 0 ? sgn : 0;
    n[1] = ax == 1 ? sgn : 0;
    n[2] = ax == 2 ? sgn : 0;
    (void)ts_near;
    return true;
}

struct HitOut {
    double t, u, v;
    uint32_t prim, geom, inst, custom;
    int32_t status;  // 0 robust miss, 1 robust hit, 2 epsilon-tie
    uint32_t n_cand;
};

struct Collector {
    // best inside candidate + bookkeeping for tie detection
    bool have = false;
    double t = 0, u = 0, v = 0, tunc = 0;
    uint32_t prim = 0, geom = 0, inst = 0;
    const Tri* tri = nullptr;
    bool tie = false;
    // list of "possibly hit" candidates (inside or near) for t-tie checks
    struct C {
        double t, tunc;
        bool inside, near;
        uint32_t inst, geom, prim;
    };
    std::vector<C> cands;
};

// traceRayEXT facing culls (gl_RayFlagsCull{Back,Front}FacingTrianglesEXT) with the instance flags
// TRIANGLE_FACING_CULL_DISABLE (0x1) and TRIANGLE_FLIP_FACING (0x2): front-facing = v0, v1, v2 clockwise seen from
// the ray origin, dot(d, (v1 - v0) x (v2 - v0)) > 0 in object space.  Returns 0 keep, 1 cull, 2 too close to
// edge-on to tell in FP32 (an epsilon-tie).
constexpr uint32_t kCullBack = 0x10u, kCullFront = 0x20u;
int facing_cull(const RayFrame& f, const Tri& tr, uint32_t ray_flags, uint32_t inst_flags) {
    if (!(ray_flags & (kCullBack | kCullFront)) || (inst_flags & 0x1u)) return 0;
    d3 v0{tr.v[0], tr.v[1], tr.v[2]}, v1{tr.v[3], tr.v[4], tr.v[5]}, v2{tr.v[6], tr.v[7], tr.v[8]};
    d3 n = cross(v1 - v0, v2 - v0);
    double fd = dot(f.d, n);
    if (std::fabs(fd) <= 1e-5 * len(n) * f.dlen) return 2;
    bool front = fd > 0;
    if (inst_flags & 0x2u) front = !front;
    return (front ? (ray_flags & kCullFront) : (ray_flags & kCullBack)) ? 1 : 0;
}

// Hit-record table for any-hit programs: the any_hit word of each 48-B record (rtx_hit_record layout, offset 44)
// and the ray's SBT offset / stride; record of a candidate = instance SBT offset + geometry x stride + offset.
struct AnyHitTable {
    const uint8_t* records = nullptr;
    uint32_t nrec = 0, sbt_offset = 0, sbt_stride = 1;
    uint32_t field(uint32_t inst_sbt, uint32_t geom, int offset) const {
        const uint32_t r = inst_sbt + geom * sbt_stride + sbt_offset;
        if (!records || r >= nrec) return 0;
        uint32_t w;
        std::memcpy(&w, records + 48 * (size_t)r + offset, 4);
        return w;
    }
    uint32_t word(uint32_t inst_sbt, uint32_t geom) const { return field(inst_sbt, geom, 44); }  // any_hit
    // intersection program of an AABB primitive: bits 8-15 of the record's program (BOX without a table)
    uint32_t intersection(uint32_t inst_sbt, uint32_t geom) const { return (field(inst_sbt, geom, 0) >> 8) & 255u; }
};

// Candidate filter of traceRayEXT for non-opaque geometry.  Opacity: the geometry's OPAQUE bit, overridden by
// the instance's FORCE_OPAQUE (0x4) / FORCE_NO_OPAQUE (0x8), overridden by the ray's OPAQUE (0x1) / NO_OPAQUE (0x2);
// CULL_OPAQUE (0x40) / CULL_NO_OPAQUE (0x80) drop candidates of that opacity; a non-opaque candidate runs its
// record's any-hit program.  Program 1 (cutout, parameter n = word >> 8, 0 -> 1) ignores the candidate where
// floor(n u) + floor(n v) is odd.  Returns 0 drop, 1 keep, 2 keep as an epsilon-tie (n u or n v within 2^-16 of
// an integer: the GPU decides on FP32 barycentrics).
int any_hit_filter(const Tri& tr, uint32_t inst_flags, uint32_t inst_sbt, uint32_t ray_flags, double u, double v,
                   const AnyHitTable* ah) {
    bool opaque = tr.opaque;
    if (inst_flags & 0x4u) opaque = true;
    else if (inst_flags & 0x8u) opaque = false;
    if (ray_flags & 0x1u) opaque = true;
    else if (ray_flags & 0x2u) opaque = false;
    if ((opaque && (ray_flags & 0x40u)) || (!opaque && (ray_flags & 0x80u))) return 0;
    if (opaque || !ah) return 1;
    const uint32_t w = ah->word(inst_sbt, tr.geom);
    if ((w & 255u) != 1u) return 1;
    const double n = (w >> 8) ? (double)(w >> 8) : 1.0;
    const double fu = u * n, fv = v * n;
    const double band = 0x1p-16;
    const bool tie = std::fabs(fu - std::nearbyint(fu)) < band || std::fabs(fv - std::nearbyint(fv)) < band;
    const bool keep = ((int64_t)std::floor(fu) + (int64_t)std::floor(fv)) % 2 == 0;
    return tie ? 2 : (keep ? 1 : 0);
}

void trace_blas(const Blas& b, const RayFrame& f, double tmin, double tmax, uint32_t inst_id, Collector& col,
                uint32_t ray_flags = 0, uint32_t inst_flags = 0, uint32_t inst_sbt = 0,
                const AnyHitTable* ah = nullptr) {
    if (b.nodes.empty()) return;
    // boxes inflated by 4*delta so that every near-edge candidate is visited
    double infl = 4.0 * b.delta;
    double inv[3] = {1.0 / f.d.x, 1.0 / f.d.y, 1.0 / f.d.z};
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const BNode& n = b.nodes[stack[--sp]];
        double t0 = -std::numeric_limits<double>::infinity(), t1 = std::numeric_limits<double>::infinity();
        bool miss = false;
        for (int a = 0; a < 3; ++a) {
            double o = comp(f.o, a), lo = n.lo[a] - infl, hi = n.hi[a] + infl;
            double da = comp(f.d, a);
            if (da == 0.0) {
                if (o < lo || o > hi) miss = true;
                continue;
            }
            double ta = (lo - o) * inv[a], tb = (hi - o) * inv[a];
            if (ta > tb) std::swap(ta, tb);
            t0 = std::max(t0, ta);
            t1 = std::min(t1, tb);
        }
        if (miss) continue;
        double slack = 1e-9 * (std::fabs(t0) + std::fabs(t1)) + 1e-300;
        if (t0 > t1 + slack) continue;
        // t window with a generous margin for tie detection
        double tw = (tmax < std::numeric_limits<double>::infinity()) ? tmax * (1 + 1e-4) + 1e-6 : tmax;
        if (t1 < tmin - std::fabs(tmin) * 1e-4 - 1e-6 || t0 > tw) continue;
        if (n.left >= 0) {
            stack[sp++] = n.left;
            stack[sp++] = n.right;
            continue;
        }
        for (int k = 0; k < n.count; ++k) {
            const Tri& tr = b.tris[b.order[n.right + k]];
            Cand c;
            if (tr.aabb) {
                if (ray_flags & 0x200u) continue;  // SKIP_AABBS
                const uint32_t prog = ah ? ah->intersection(inst_sbt, tr.geom) : 0u;
                double nn[3];
                bool tie = false;
                const bool hit = proc_hit(prog, tr, f.o, f.d, tmin, tmax, c.t, c.u, c.v, nn, tie);
                if (!hit && !tie) continue;
                c.inside = hit;
                c.near = tie;
                c.margin = 0;
                c.tunc = 1e-5 * std::fabs(c.t) + 1e-7;
                if (!hit) c.t = tmin;  // a grazing miss: a tie candidate at an unknown distance
            } else {
                if (ray_flags & 0x100u) continue;  // SKIP_TRIANGLES
                if (!eval_tri(f, tr, b.delta, c)) continue;
                const int cull = facing_cull(f, tr, ray_flags, inst_flags);
                if (cull == 1) continue;
                if (cull == 2) c.near = true;
            }
            const int keep = any_hit_filter(tr, inst_flags, inst_sbt, ray_flags, c.u, c.v, ah);
            if (keep == 0) continue;
            if (keep == 2) {  // may be ignored on the GPU: a tie candidate that cannot win robustly
                c.near = true;
            }
            // candidates relevant to [tmin, tmax] (with uncertainty)
            if (c.t + c.tunc < tmin || c.t - c.tunc > tmax) continue;
            if (!c.inside && !c.near) continue;
            col.cands.push_back({c.t, c.tunc, c.inside, c.near, inst_id, tr.geom, tr.prim});
            if (c.inside && c.t >= tmin && c.t <= tmax) {
                bool better = !col.have || c.t < col.t ||
                              (c.t == col.t && (inst_id < col.inst || (inst_id == col.inst &&
                                                                       (tr.geom < col.geom || (tr.geom == col.geom && tr.prim < col.prim)))));
                if (better) {
                    col.have = true;
                    col.t = c.t;
                    col.u = c.u;
                    col.v = c.v;
                    col.tunc = c.tunc;
                    col.prim = tr.prim;
                    col.geom = tr.geom;
                    col.inst = inst_id;
                    col.tri = &tr;
                }
            }
        }
    }
}

// ray_flags: gl_RayFlags*EXT (facing and opacity culls, SKIP_TRIANGLES); ah: any-hit programs (may be NULL).
void trace_one(const Tlas& tl, const float* ray, uint32_t cull, int any_hit, HitOut& out, uint32_t ray_flags = 0,
               const AnyHitTable* ah = nullptr, Collector* keep = nullptr) {
    d3 o{ray[0], ray[1], ray[2]}, d{ray[4], ray[5], ray[6]};
    double tmin = ray[3], tmax = ray[7];
    Collector local;
    Collector& col = keep ? *keep : local;
    const bool skip_all = ((ray_flags & 0x100u) && (ray_flags & 0x200u)) ||
                          ((ray_flags & 0x40u) && (ray_flags & (0x1u | 0x80u))) ||
                          ((ray_flags & 0x80u) && (ray_flags & 0x2u));
    for (size_t i = 0; i < tl.inst.size() && !skip_all; ++i) {
        const Inst& in = tl.inst[i];
        if (!in.valid || ((in.mask & cull) == 0)) continue;
        const Blas& b = *g_blas[in.blas];
        const double* M = in.minv;
        d3 oo{M[0] * o.x + M[1] * o.y + M[2] * o.z + M[3], M[4] * o.x + M[5] * o.y + M[6] * o.z + M[7],
              M[8] * o.x + M[9] * o.y + M[10] * o.z + M[11]};
        d3 dd{M[0] * d.x + M[1] * d.y + M[2] * d.z, M[4] * d.x + M[5] * d.y + M[6] * d.z,
              M[8] * d.x + M[9] * d.y + M[10] * d.z};
        if (dd.x == 0 && dd.y == 0 && dd.z == 0) continue;
        RayFrame f = make_frame(oo, dd);
        trace_blas(b, f, tmin, tmax, (uint32_t)i, col, ray_flags, in.flags, in.sbt_offset, ah);
    }
    out.n_cand = (uint32_t)col.cands.size();
    out.custom = 0;
    // Every entry of col.cands is either inside the triangle or within 3*delta of an edge, and its t lies in
    // [tmin, tmax] up to its t-uncertainty.
    if (any_hit) {
        bool robust = false;
        for (auto& c : col.cands)
            if (c.inside && !c.near && c.t - c.tunc >= tmin && c.t + c.tunc <= tmax) robust = true;
        out.status = robust ? 1 : (col.cands.empty() ? 0 : 2);
    }
    if (!col.have) {
        out.t = 0;
        out.u = out.v = 0;
        out.prim = out.geom = 0;
        out.inst = 0xFFFFFFFFu;
        if (!any_hit) out.status = col.cands.empty() ? 0 : 2;
        return;
    }
    out.t = col.t;
    out.u = col.u;
    out.v = col.v;
    out.prim = col.prim;
    out.geom = col.geom;
    out.inst = col.inst;
    out.custom = tl.inst[col.inst].custom;
    if (any_hit) return;
    bool tie = col.t - col.tunc < tmin || col.t + col.tunc > tmax;
    int close = 0;
    for (auto& c : col.cands) {
        // a near-edge candidate that could be in front of (or level with) the winner
        if (c.near && c.t - c.tunc <= col.t + col.tunc) tie = true;
        // another hit whose distance is indistinguishable from the winner's
        if (c.inside && std::fabs(c.t - col.t) <= c.tunc + col.tunc) ++close;
    }
    if (close > 1) tie = true;
    out.status = tie ? 2 : 1;
}

// ---------------------------------------------------------------------------------------------------------
// Morton codes (LBVH step 1).  Defined FP32 arithmetic, restated bit-for-bit by the GPU builder.
// ---------------------------------------------------------------------------------------------------------
inline uint32_t quant(float c, float cmin, float scale, uint32_t maxq) {
    float f = (c - cmin) * scale;
    if (!(f > 0.0f)) return 0;  // also catches NaN
    if (f >= (float)maxq) return maxq;
    return (uint32_t)f;
}
inline uint32_t spread10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
inline uint64_t spread21(uint64_t v) {
    v &= 0x1FFFFFull;
    v = (v | (v << 32)) & 0x1F00000000FFFFull;
    v = (v | (v << 16)) & 0x1F0000FF0000FFull;
    v = (v | (v << 8)) & 0x100F00F00F00F00Full;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

inline int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }
inline int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }

// ---------------------------------------------------------------------------------------------------------
// PCG (O'Neill 2014 pcg32; Jarzynski & Olano 2020 pcg4d)
// ---------------------------------------------------------------------------------------------------------
struct Pcg32 {
    uint64_t state, inc;
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
    }
    void seed(uint64_t initstate, uint64_t initseq) {
        state = 0u;
        inc = (initseq << 1u) | 1u;
        next();
        state += initstate;
        next();
    }
};

}  // namespace

extern "C" {

// ---- scene ------------------------------------------------------------------------------------------------
// Create a BLAS from ngeom triangle geometries.  vtx[g]: FP32 xyz vertices (tightly packed), idx[g]: uint32
// triangles (3 per primitive), xform[g]: row-major 3x4 or NULL.
int oracle_blas_create(int ngeom, const float* const* vtx, const uint32_t* const* idx, const uint32_t* prim_count,
                       const float* const* xform, const uint32_t* geom_flags) {
    Blas* b = new Blas();
    for (int g = 0; g < ngeom; ++g) {
        if (geom_flags && (geom_flags[g] & 0x100u)) {  // AABB geometry: 6 floats per primitive in vtx[g]
            for (uint32_t p = 0; p < prim_count[g]; ++p) {
                Tri t;
                std::memcpy(t.v, vtx[g] + 6 * (size_t)p, 24);
                t.v[6] = t.v[7] = t.v[8] = 0.0f;
                bool ok = true;
                for (int a = 0; a < 3; ++a) ok = ok && std::isfinite(t.v[a]) && std::isfinite(t.v[3 + a]) && t.v[a] <= t.v[3 + a];
                if (!ok) continue;  // inactive
                t.prim = p;
                t.geom = (uint32_t)g;
                t.opaque = (geom_flags[g] & 0x1u) != 0;
                t.aabb = true;
                b->tris.push_back(t);
            }
            continue;
        }
        for (uint32_t p = 0; p < prim_count[g]; ++p) {
            Tri t;
            for (int k = 0; k < 3; ++k) {
                const float* v = vtx[g] + 3 * (size_t)idx[g][3 * (size_t)p + k];
                if (xform[g])
                    xform_f32(xform[g], v, &t.v[3 * k]);
                else
                    std::memcpy(&t.v[3 * k], v, 12);
            }
            // inactive triangle (Vulkan spec "Inactive Primitives and Instances": a NaN X component makes a
            // triangle inactive; other NaN / Inf components are undefined, and the build treats them the same way):
            // never hit, but it keeps its primitive index, so the triangles after it keep theirs
            bool finite = true;
            for (int k = 0; k < 9; ++k) finite = finite && std::isfinite(t.v[k]);
            if (!finite) continue;
            t.prim = p;
            t.geom = (uint32_t)g;
            t.opaque = !geom_flags || (geom_flags[g] & 0x1u);
            t.aabb = false;
            b->tris.push_back(t);
        }
    }
    size_t n = b->tris.size();
    for (int a = 0; a < 3; ++a) {
        b->lo[a] = 1e300;
        b->hi[a] = -1e300;
    }
    std::vector<double> cen(3 * n);
    for (size_t i = 0; i < n; ++i) {
        double lo[3], hi[3];
        tri_box(b->tris[i], lo, hi);
        for (int a = 0; a < 3; ++a) {
            cen[3 * i + a] = 0.5 * (lo[a] + hi[a]);
            b->lo[a] = std::min(b->lo[a], lo[a]);
            b->hi[a] = std::max(b->hi[a], hi[a]);
        }
    }
    if (n) {
        double R = 0, D2 = 0;
        for (int a = 0; a < 3; ++a) {
            R = std::max(R, std::max(std::fabs(b->lo[a]), std::fabs(b->hi[a])));
            D2 += (b->hi[a] - b->lo[a]) * (b->hi[a] - b->lo[a]);
        }
        b->delta = kDeltaScale * (R + std::sqrt(D2));
        b->order.resize(n);
        for (size_t i = 0; i < n; ++i) b->order[i] = (uint32_t)i;
        b->nodes.reserve(n);
        build_rec(*b, cen, b->order.data(), (int)n, 0);
    } else {
        b->delta = 0;
    }
    g_blas.push_back(b);
    return (int)g_blas.size() - 1;
}

// Instances: 64-B records in the VkAccelerationStructureInstanceKHR layout (main.cpp:848-858), except that
// the last 8 bytes hold an oracle BLAS handle instead of a device address.
int oracle_tlas_create(int ninst, const void* records) {
    Tlas* t = new Tlas();
    const uint8_t* p = (const uint8_t*)records;
    for (int i = 0; i < ninst; ++i, p += 64) {
        float m[12];
        std::memcpy(m, p, 48);
        uint32_t cm, sf;
        uint64_t h;
        std::memcpy(&cm, p + 48, 4);
        std::memcpy(&sf, p + 52, 4);
        std::memcpy(&h, p + 56, 8);
        Inst in;
        for (int k = 0; k < 12; ++k) in.m[k] = m[k];
        invert3x4(in.m, in.minv);
        in.custom = cm & 0xFFFFFFu;
        in.mask = cm >> 24;
        in.sbt_offset = sf & 0xFFFFFFu;
        in.flags = sf >> 24;
        in.blas = (int)h;
        in.valid = (h < g_blas.size()) && std::isfinite(in.minv[0]);
        t->inst.push_back(in);
    }
    g_tlas.push_back(t);
    return (int)g_tlas.size() - 1;
}

void oracle_reset() {
    for (auto* b : g_blas) delete b;
    for (auto* t : g_tlas) delete t;
    g_blas.clear();
    g_tlas.clear();
}

double oracle_blas_delta(int blas) { return g_blas[blas]->delta; }

// Trace n rays (32-B records: ox oy oz tmin dx dy dz tmax).  Outputs (any may be NULL):
//   t_uv[3n] doubles, ids[4n] = prim, geom, inst (0xFFFFFFFF miss), custom index, status[n] (0 miss, 1 hit,
//   2 epsilon-tie).  threads <= 0 uses hardware concurrency.
void oracle_trace(int tlas, const float* rays, int64_t n, int any_hit, uint32_t cull_mask, uint32_t ray_flags,
                  const void* records, uint32_t nrec, uint32_t sbt_offset, uint32_t sbt_stride, double* t_uv,
                  uint32_t* ids, int32_t* status, int threads) {
    const Tlas& tl = *g_tlas[tlas];
    AnyHitTable ahs;
    ahs.records = (const uint8_t*)records;
    ahs.nrec = nrec;
    ahs.sbt_offset = sbt_offset;
    ahs.sbt_stride = sbt_stride;
    const AnyHitTable* ah = records ? &ahs : nullptr;
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::atomic<int64_t> next{0};
    auto work = [&]() {
        for (;;) {
            int64_t s = next.fetch_add(256);
            if (s >= n) break;
            int64_t e = std::min<int64_t>(n, s + 256);
            for (int64_t i = s; i < e; ++i) {
                HitOut h;
                trace_one(tl, rays + 8 * i, cull_mask, any_hit, h, ray_flags, ah);
                if (t_uv) {
                    t_uv[3 * i] = h.t;
                    t_uv[3 * i + 1] = h.u;
                    t_uv[3 * i + 2] = h.v;
                }
                if (ids) {
                    ids[4 * i] = h.prim;
                    ids[4 * i + 1] = h.geom;
                    ids[4 * i + 2] = h.inst;
                    ids[4 * i + 3] = h.custom;
                }
                if (status) status[i] = h.status;
            }
        }
    };
    std::vector<std::thread> pool;
    for (int k = 1; k < threads; ++k) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
}

// Candidate sets of epsilon-tie rays: every triangle the FP64 test finds inside, or within its FP32 uncertainty
// of an edge, with t in [tmin, tmax] up to its t-uncertainty - the triangles an FP32 traversal may report.  For ray
// i: count[i] candidates (all of them, also beyond max_c), the first min(count, max_c) in ascending t with
// ids[3 (i max_c + k) ..] = inst, geom, prim; t / tunc[i max_c + k]; flags = 1 inside | 2 near an edge.
void oracle_trace_candidates(int tlas, const float* rays, int64_t n, uint32_t cull_mask, int max_c, uint32_t* count,
                             uint32_t* ids, double* t, double* tunc, uint32_t* flags, int threads) {
    const Tlas& tl = *g_tlas[tlas];
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::atomic<int64_t> next{0};
    auto work = [&]() {
        for (;;) {
            int64_t s = next.fetch_add(64);
            if (s >= n) break;
            int64_t e = std::min<int64_t>(n, s + 64);
            for (int64_t i = s; i < e; ++i) {
                HitOut h;
                Collector col;
                trace_one(tl, rays + 8 * i, cull_mask, 0, h, 0u, nullptr, &col);
                std::sort(col.cands.begin(), col.cands.end(),
                          [](const Collector::C& a, const Collector::C& b) { return a.t < b.t; });
                count[i] = (uint32_t)col.cands.size();
                for (int k = 0; k < max_c && k < (int)col.cands.size(); ++k) {
                    const Collector::C& c = col.cands[k];
                    const int64_t j = i * max_c + k;
                    ids[3 * j] = c.inst;
                    ids[3 * j + 1] = c.geom;
                    ids[3 * j + 2] = c.prim;
                    t[j] = c.t;
                    tunc[j] = c.tunc;
                    flags[j] = (c.inside ? 1u : 0u) | (c.near ? 2u : 0u);
                }
            }
        }
    };
    std::vector<std::thread> pool;
    for (int k = 1; k < threads; ++k) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
}

// Brute-force variant of oracle_trace for small scenes: every triangle of every instance, no BVH.
void oracle_trace_brute(int tlas, const float* rays, int64_t n, uint32_t cull_mask, double* t_uv, uint32_t* ids,
                        int32_t* status) {
    const Tlas& tl = *g_tlas[tlas];
    for (int64_t r = 0; r < n; ++r) {
        const float* ray = rays + 8 * r;
        d3 o{ray[0], ray[1], ray[2]}, d{ray[4], ray[5], ray[6]};
        double tmin = ray[3], tmax = ray[7];
        bool have = false;
        double bt = 0, bu = 0, bv = 0;
        uint32_t bp = 0, bg = 0, bi = 0xFFFFFFFFu;
        bool tie = false;
        for (size_t i = 0; i < tl.inst.size(); ++i) {
            const Inst& in = tl.inst[i];
            if (!in.valid || ((in.mask & cull_mask) == 0)) continue;
            const Blas& b = *g_blas[in.blas];
            const double* M = in.minv;
            d3 oo{M[0] * o.x + M[1] * o.y + M[2] * o.z + M[3], M[4] * o.x + M[5] * o.y + M[6] * o.z + M[7],
                  M[8] * o.x + M[9] * o.y + M[10] * o.z + M[11]};
            d3 dd{M[0] * d.x + M[1] * d.y + M[2] * d.z, M[4] * d.x + M[5] * d.y + M[6] * d.z,
                  M[8] * d.x + M[9] * d.y + M[10] * d.z};
            RayFrame f = make_frame(oo, dd);
            for (const Tri& tr : b.tris) {
                Cand c;
                if (!eval_tri(f, tr, b.delta, c)) continue;
                if (c.near && c.t + c.tunc >= tmin && c.t - c.tunc <= tmax) tie = true;
                if (c.inside && c.t >= tmin && c.t <= tmax) {
                    bool better = !have || c.t < bt;
                    if (better) {
                        have = true;
                        bt = c.t;
                        bu = c.u;
                        bv = c.v;
                        bp = tr.prim;
                        bg = tr.geom;
                        bi = (uint32_t)i;
                    }
                }
            }
        }
        if (t_uv) {
            t_uv[3 * r] = bt;
            t_uv[3 * r + 1] = bu;
            t_uv[3 * r + 2] = bv;
        }
        if (ids) {
            ids[4 * r] = bp;
            ids[4 * r + 1] = bg;
            ids[4 * r + 2] = bi;
            ids[4 * r + 3] = have ? tl.inst[bi].custom : 0;
        }
        if (status) status[r] = tie ? 2 : (have ? 1 : 0);
    }
}

// ---- raygen (main.cpp:1033-1055) ------------------------------------------------------------------------
// aspect terms are computed once per launch in FP32, following the GLSL expression order:
//   aspect_y = tan(radians(fov) * 0.5)   (evaluated in FP64 and rounded: GLSL leaves tan's precision open)
//   aspect_x = aspect_y * float(W) / float(H)
void oracle_aspect(float fov_deg, uint32_t w, uint32_t h, float* ax, float* ay) {
    double rad = (double)fov_deg * (3.14159265358979323846 / 180.0);
    float a_y = (float)std::tan(rad * 0.5);
    float t = a_y * (float)w;
    *ay = a_y;
    *ax = t / (float)h;
}

// Primary rays of a width x height launch, row-major, 32-B records.
void oracle_raygen(const float* cam, uint32_t w, uint32_t h, float tmin, float tmax, float* rays) {
    float ax, ay;
    oracle_aspect(cam[3], w, h, &ax, &ay);
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            float sx = ((float)x + 0.5f) / (float)w;
            float sy = ((float)y + 0.5f) / (float)h;
            float nx = sx * 2.0f - 1.0f;
            float ny = sy * 2.0f - 1.0f;
            float* r = rays + 8 * ((size_t)y * w + x);
            r[0] = cam[0];
            r[1] = cam[1];
            r[2] = cam[2];
            r[3] = tmin;
            r[4] = nx * ax;
            r[5] = -(ny * ay);
            r[6] = -1.0f;
            r[7] = tmax;
        }
}

// closest-hit / miss programs (main.cpp:1063-1066, 1080-1091) applied to oracle hits.
// records: 12 floats each (48 B rtx_hit_record layout: program, color[3], emission[3], match[4], reserved).
void oracle_shade(int tlas, int64_t n, const double* t_uv, const uint32_t* ids, const void* records,
                  uint32_t nrec, uint32_t sbt_offset, uint32_t sbt_stride, const float* miss, float* film) {
    const Tlas& tl = *g_tlas[tlas];
    const uint8_t* rec = (const uint8_t*)records;
    for (int64_t i = 0; i < n; ++i) {
        float* px = film + 4 * i;
        px[3] = 0.0f;
        uint32_t inst = ids[4 * i + 2];
        if (inst == 0xFFFFFFFFu) {
            px[0] = miss[0];
            px[1] = miss[1];
            px[2] = miss[2];
            continue;
        }
        uint32_t prim = ids[4 * i], geom = ids[4 * i + 1];
        uint32_t r = tl.inst[inst].sbt_offset + geom * sbt_stride + sbt_offset;
        if (r >= nrec) {
            px[0] = px[1] = px[2] = 0.0f;
            continue;
        }
        const uint8_t* R = rec + 48 * (size_t)r;
        uint32_t prog;
        float col[3];
        int32_t match[4];
        std::memcpy(&prog, R, 4);
        std::memcpy(col, R + 4, 12);
        std::memcpy(match, R + 28, 16);
        float u = (float)t_uv[3 * i + 1], v = (float)t_uv[3 * i + 2];
        bool bary = prog == 1u;
        if (prog == 2u)
            bary = (int32_t)prim == match[0] && (int32_t)inst == match[1] &&
                   (int32_t)tl.inst[inst].custom == match[2] && (int32_t)geom == match[3];
        if (bary) {
            px[0] = 1.0f - u - v;
            px[1] = u;
            px[2] = v;
        } else {
            px[0] = col[0];
            px[1] = col[1];
            px[2] = col[2];
        }
    }
}

// ---- LBVH steps -------------------------------------------------------------------------------------------
// Triangle boxes and centroids after the build-time transform; centroid = (lo + hi) * 0.5 in FP32.
void oracle_centroids(int ngeom, const float* const* vtx, const uint32_t* const* idx, const uint32_t* prim_count,
                      const float* const* xform, float* cen) {
    size_t o = 0;
    for (int g = 0; g < ngeom; ++g)
        for (uint32_t p = 0; p < prim_count[g]; ++p, ++o) {
            float v[9];
            for (int k = 0; k < 3; ++k) {
                const float* s = vtx[g] + 3 * (size_t)idx[g][3 * (size_t)p + k];
                if (xform[g])
                    xform_f32(xform[g], s, &v[3 * k]);
                else
                    std::memcpy(&v[3 * k], s, 12);
            }
            for (int a = 0; a < 3; ++a) {
                float lo = std::fmin(v[a], std::fmin(v[3 + a], v[6 + a]));
                float hi = std::fmax(v[a], std::fmax(v[3 + a], v[6 + a]));
                float s = lo + hi;
                cen[3 * o + a] = s * 0.5f;
            }
        }
}

// Morton codes over the centroid bounds.  bits = 30 (uint32 out) or 63 (uint64 out).
// scale_a = float(2^k) / (cmax_a - cmin_a) (0 for a flat axis); q = trunc((c - cmin) * scale), clamped.
void oracle_morton(const float* cen, int64_t n, int bits, void* out) {
    float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            cmin[a] = std::fmin(cmin[a], cen[3 * i + a]);
            cmax[a] = std::fmax(cmax[a], cen[3 * i + a]);
        }
    float cells = bits == 30 ? 1024.0f : 2097152.0f;
    uint32_t maxq = bits == 30 ? 1023u : 2097151u;
    float scale[3];
    for (int a = 0; a < 3; ++a) {
        float ext = cmax[a] - cmin[a];
        scale[a] = ext > 0.0f ? cells / ext : 0.0f;
    }
    for (int64_t i = 0; i < n; ++i) {
        uint32_t q[3];
        for (int a = 0; a < 3; ++a) q[a] = quant(cen[3 * i + a], cmin[a], scale[a], maxq);
        if (bits == 30)
            ((uint32_t*)out)[i] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
        else
            ((uint64_t*)out)[i] = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
    }
}

// Karras 2012 hierarchy over n sorted keys (key_bits 32 or 64; keys passed as uint64).  Internal node i
// (0..n-2) gets children left[i], right[i]; a child value c >= 0 is an internal node, c < 0 encodes leaf
// (-1 - sorted position).  Equal keys are disambiguated by position: delta = key_bits + clz32(i ^ j).
void oracle_karras(const uint64_t* keys, int64_t n, int key_bits, int32_t* left, int32_t* right) {
    auto delta = [&](int64_t i, int64_t j) -> int {
        if (j < 0 || j >= n) return -1;
        uint64_t a = keys[i], b = keys[j];
        if (a == b) return key_bits + clz32((uint32_t)(i ^ j));
        return key_bits == 32 ? clz32((uint32_t)(a ^ b)) : clz64(a ^ b);
    };
    for (int64_t i = 0; i + 1 < n; ++i) {
        int d = (delta(i, i + 1) - delta(i, i - 1)) >= 0 ? 1 : -1;
        int dmin = delta(i, i - d);
        int64_t lmax = 2;
        while (delta(i, i + lmax * d) > dmin) lmax *= 2;
        int64_t l = 0;
        for (int64_t t = lmax / 2; t >= 1; t /= 2)
            if (delta(i, i + (l + t) * d) > dmin) l += t;
        int64_t j = i + l * d;
        int dnode = delta(i, j);
        int64_t s = 0;
        int64_t span = l;
        // binary search for the split: largest s with delta(i, i + s*d) > dnode
        int64_t t = span;
        do {
            t = (t + 1) / 2;
            if (s + t < span + 1 && delta(i, i + (s + t) * d) > dnode) s += t;
        } while (t > 1);
        int64_t gamma = i + s * d + std::min(d, 0);
        int64_t lo = std::min(i, j), hi = std::max(i, j);
        left[i] = (lo == gamma) ? (int32_t)(-1 - gamma) : (int32_t)gamma;
        right[i] = (hi == gamma + 1) ? (int32_t)(-1 - (gamma + 1)) : (int32_t)(gamma + 1);
    }
}

// ---- SAH treelet refinement (restates refit_refine_kernel, build-up-phase_amd/csrc/build.hip) -----------------
// FP32 throughout, same operation order as the GPU: surface area dx*dy + dy*dz + dz*dx (0 for an empty box),
// node cost = area + (cost(left) + cost(right)) with leaf cost 0.  Nodes are processed in postorder; because a
// treelet only touches its root's finished subtree, every children-first order (the GPU's included) gives the
// same tree.
static inline float tl_area(const float* b) {
    float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    if (!(dx >= 0.0f) || !(dy >= 0.0f) || !(dz >= 0.0f)) return 0.0f;
    return dx * dy + dy * dz + dz * dx;
}

// DP over the 127 subsets of 7 treelet leaves: cost[S] = area(S) + min over splits {P, S\P}, P containing the
// lowest leaf of S, of (cost[P] + cost[S\P]); ties go to the smallest P.  cost/part have 128 entries.
void oracle_treelet_dp(const float* boxes, const float* leaf_cost, float* area, float* cost, uint32_t* part) {
    for (uint32_t S = 1; S < 128; ++S) {
        float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        for (int j = 0; j < 7; ++j)
            if (S & (1u << j))
                for (int a = 0; a < 3; ++a) {
                    b[a] = std::fmin(b[a], boxes[6 * j + a]);
                    b[3 + a] = std::fmax(b[3 + a], boxes[6 * j + 3 + a]);
                }
        area[S] = tl_area(b);
        if (__builtin_popcount(S) == 1) {
            cost[S] = leaf_cost[__builtin_ctz(S)];
            part[S] = 0;
        }
    }
    for (int k = 2; k <= 7; ++k)
        for (uint32_t S = 1; S < 128; ++S) {
            if (__builtin_popcount(S) != k) continue;
            const uint32_t low = S & (0u - S);
            float best = INFINITY;
            uint32_t bp = 0;
            for (uint32_t P = 1; P < S; ++P) {
                if ((P & S) != P || !(P & low)) continue;
                const float c = cost[P] + cost[S ^ P];
                if (c < best || (c == best && P < bp)) {
                    best = c;
                    bp = P;
                }
            }
            cost[S] = area[S] + best;
            part[S] = bp;
        }
}

namespace {
struct TreeletRefiner {
    uint32_t n;
    int32_t* L;
    int32_t* R;
    std::vector<float> box;   // (2n-1) x 6
    std::vector<float> cost;  // n-1
    std::vector<uint32_t> cnt, hgt;
    std::vector<int32_t> parent;
    size_t idx(int32_t c) const { return c >= 0 ? (size_t)c : (size_t)n - 1 + (size_t)(-1 - c); }
    float area_of(int32_t c) const { return tl_area(&box[6 * idx(c)]); }

    void optimize(int32_t X) {
        int32_t lf[7];
        float la[7];
        int32_t pool[6];
        lf[0] = L[X];
        lf[1] = R[X];
        la[0] = lf[0] >= 0 ? area_of(lf[0]) : 0.0f;
        la[1] = lf[1] >= 0 ? area_of(lf[1]) : 0.0f;
        pool[0] = X;
        for (int k = 2; k < 7; ++k) {
            int best = -1;
            float ba = -1.0f;
            for (int j = 0; j < k; ++j)
                if (lf[j] >= 0 && la[j] > ba) {
                    ba = la[j];
                    best = j;
                }
            if (best < 0) return;
            const int32_t c = lf[best];
            pool[k - 1] = c;
            const int32_t l = L[c], r = R[c];
            lf[best] = l;
            la[best] = l >= 0 ? area_of(l) : 0.0f;
            lf[k] = r;
            la[k] = r >= 0 ? area_of(r) : 0.0f;
        }
        float tb[42], lc[7], area[128], dc[128];
        uint32_t part[128], lcnt[7], lh[7];
        for (int j = 0; j < 7; ++j) {
            std::memcpy(&tb[6 * j], &box[6 * idx(lf[j])], 24);
            lc[j] = lf[j] >= 0 ? cost[lf[j]] : 0.0f;
            lcnt[j] = lf[j] >= 0 ? cnt[lf[j]] : 1u;
            lh[j] = lf[j] >= 0 ? hgt[lf[j]] : 0u;
        }
        oracle_treelet_dp(tb, lc, area, dc, part);
        const float total = dc[127];
        if (!(total < cost[X])) return;
        // breadth-first rebuild from all leaves; internal nodes reused in formation order
        uint32_t ls[6];
        int32_t lx[6], lcode[6][2], li[6][2];
        uint32_t lhh[6];
        ls[0] = 127u;
        lx[0] = X;
        uint32_t cntl = 1, next = 1;
        for (uint32_t i = 0; i < cntl; ++i) {
            const uint32_t P = part[ls[i]];
            if (P == 0u || (P & ls[i]) != P || P == ls[i]) return;
            const uint32_t C[2] = {P, ls[i] ^ P};
            for (int h = 0; h < 2; ++h) {
                if (__builtin_popcount(C[h]) == 1) {
                    lcode[i][h] = lf[__builtin_ctz(C[h])];
                    li[i][h] = -1 - __builtin_ctz(C[h]);
                } else {
                    const int32_t x = pool[next++];
                    ls[cntl] = C[h];
                    lx[cntl] = x;
                    lcode[i][h] = x;
                    li[i][h] = (int32_t)cntl;
                    ++cntl;
                }
            }
        }
        for (int i = 5; i >= 0; --i) {
            uint32_t hc[2];
            for (int h = 0; h < 2; ++h) hc[h] = li[i][h] >= 0 ? lhh[li[i][h]] : lh[-1 - li[i][h]];
            lhh[i] = 1u + std::max(hc[0], hc[1]);
        }
        if (lhh[0] > hgt[X]) return;
        for (int i = 0; i < 6; ++i) {
            const int32_t x = lx[i];
            L[x] = lcode[i][0];
            R[x] = lcode[i][1];
            parent[idx(lcode[i][0])] = x;
            parent[idx(lcode[i][1])] = x;
            if (i > 0) {
                float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
                uint32_t np = 0;
                for (int j = 0; j < 7; ++j)
                    if (ls[i] & (1u << j)) {
                        for (int a = 0; a < 3; ++a) {
                            b[a] = std::fmin(b[a], tb[6 * j + a]);
                            b[3 + a] = std::fmax(b[3 + a], tb[6 * j + 3 + a]);
                        }
                        np += lcnt[j];
                    }
                std::memcpy(&box[6 * (size_t)x], b, 24);
                cost[x] = dc[ls[i]];
                cnt[x] = np;
                hgt[x] = lhh[i];
            }
        }
        cost[X] = total;
        hgt[X] = lhh[0];
    }

    void pass(uint32_t min_prims) {
        // postorder of the current tree
        std::vector<int32_t> post;
        post.reserve(n - 1);
        std::vector<std::pair<int32_t, int>> st;
        st.push_back({0, 0});
        while (!st.empty()) {
            auto& [x, state] = st.back();
            if (state < 2) {
                const int32_t c = state == 0 ? L[x] : R[x];
                ++state;
                if (c >= 0) st.push_back({c, 0});
            } else {
                post.push_back(x);
                st.pop_back();
            }
        }
        for (int32_t x : post) {
            const int32_t l = L[x], r = R[x];
            float* b = &box[6 * (size_t)x];
            const float* bl = &box[6 * idx(l)];
            const float* br = &box[6 * idx(r)];
            for (int a = 0; a < 3; ++a) {
                b[a] = std::fmin(bl[a], br[a]);
                b[3 + a] = std::fmax(bl[3 + a], br[3 + a]);
            }
            const float cl = l >= 0 ? cost[l] : 0.0f, cr = r >= 0 ? cost[r] : 0.0f;
            cost[x] = tl_area(b) + (cl + cr);
            cnt[x] = (l >= 0 ? cnt[l] : 1u) + (r >= 0 ? cnt[r] : 1u);
            hgt[x] = 1u + std::max(l >= 0 ? hgt[l] : 0u, r >= 0 ? hgt[r] : 0u);
            if (cnt[x] >= min_prims) optimize(x);
        }
    }
};
}  // namespace

// Refit + treelet refinement of a binary radix tree (n >= 2): left/right (n-1 entries, updated in place), leaf
// boxes in sorted position order (n x 6).  rounds = 0: refit only; round r restructures nodes with at least
// 7 * 2^r primitives.  Outputs per-internal-node subtree cost and height, and cost(root) / area(root).
void oracle_treelet_refine(uint32_t n, int32_t* left, int32_t* right, const float* leaf_boxes, uint32_t rounds,
                           float* cost_out, uint32_t* hgt_out, double* sah_out) {
    TreeletRefiner t;
    t.n = n;
    t.L = left;
    t.R = right;
    t.box.assign(6 * (2 * (size_t)n - 1), 0.0f);
    std::memcpy(&t.box[6 * ((size_t)n - 1)], leaf_boxes, 24 * (size_t)n);
    t.cost.assign(n - 1, 0.0f);
    t.cnt.assign(n - 1, 0u);
    t.hgt.assign(n - 1, 0u);
    t.parent.assign(2 * (size_t)n - 1, -1);
    for (uint32_t i = 0; i + 1 < n; ++i) {
        t.parent[t.idx(left[i])] = (int32_t)i;
        t.parent[t.idx(right[i])] = (int32_t)i;
    }
    const uint32_t passes = (rounds && n >= 7) ? rounds : 1;
    for (uint32_t r = 0; r < passes; ++r) t.pass((rounds && n >= 7) ? (7u << r) : 0xFFFFFFFFu);
    std::memcpy(cost_out, t.cost.data(), 4 * ((size_t)n - 1));
    std::memcpy(hgt_out, t.hgt.data(), 4 * ((size_t)n - 1));
    const float ra = tl_area(&t.box[0]);
    *sah_out = ra > 0.0f ? (double)t.cost[0] / (double)ra : 0.0;
}

// ---- PCG ----------------------------------------------------------------------------------------------------
void oracle_pcg32(uint64_t initstate, uint64_t initseq, int n, uint32_t* out) {
    Pcg32 r;
    r.seed(initstate, initseq);
    for (int i = 0; i < n; ++i) out[i] = r.next();
}

// pcg4d (Jarzynski & Olano 2020, "Hash Functions for GPU Rendering", JCGT 9(3)).
void oracle_pcg4d(const uint32_t* in, int64_t n, uint32_t* out) {
    for (int64_t i = 0; i < n; ++i) {
        uint32_t x = in[4 * i], y = in[4 * i + 1], z = in[4 * i + 2], w = in[4 * i + 3];
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
        out[4 * i] = x;
        out[4 * i + 1] = y;
        out[4 * i + 2] = z;
        out[4 * i + 3] = w;
    }
}

// ---- path tracer (the estimator of rtx_cmd_render, evaluated in FP64) --------------------------------------
}  // extern "C"

namespace {

struct PHit {
    bool hit = false;
    double t = 0;
    uint32_t inst = 0, geom = 0, prim = 0;
    double n[3] = {0, 0, 0};
};

// Closest opaque hit of the world ray (o, d) in [tmin, tmax] (FP64, no tie bookkeeping needed here).
PHit pt_closest(const Tlas& tl, const double* o, const double* d, double tmin, double tmax,
                const AnyHitTable* ah) {
    Collector col;
    for (size_t i = 0; i < tl.inst.size(); ++i) {
        const Inst& in = tl.inst[i];
        if (!in.valid) continue;
        const double* M = in.minv;
        d3 oo{M[0] * o[0] + M[1] * o[1] + M[2] * o[2] + M[3], M[4] * o[0] + M[5] * o[1] + M[6] * o[2] + M[7],
              M[8] * o[0] + M[9] * o[1] + M[10] * o[2] + M[11]};
        d3 dd{M[0] * d[0] + M[1] * d[1] + M[2] * d[2], M[4] * d[0] + M[5] * d[1] + M[6] * d[2],
              M[8] * d[0] + M[9] * d[1] + M[10] * d[2]};
        if (dd.x == 0 && dd.y == 0 && dd.z == 0) continue;
        RayFrame f = make_frame(oo, dd);
        trace_blas(*g_blas[in.blas], f, tmin, tmax, (uint32_t)i, col, 0, in.flags, in.sbt_offset, ah);
    }
    PHit h;
    if (!col.have) return h;
    h.hit = true;
    h.t = col.t;
    h.inst = col.inst;
    h.geom = col.geom;
    h.prim = col.prim;
    const Tri& tr = *col.tri;
    const double* M = tl.inst[col.inst].minv;
    double no[3];
    if (tr.aabb) {  // the intersection program's normal at the hit, from the object-space ray
        d3 oo{M[0] * o[0] + M[1] * o[1] + M[2] * o[2] + M[3], M[4] * o[0] + M[5] * o[1] + M[6] * o[2] + M[7],
              M[8] * o[0] + M[9] * o[1] + M[10] * o[2] + M[11]};
        d3 dd{M[0] * d[0] + M[1] * d[1] + M[2] * d[2], M[4] * d[0] + M[5] * d[1] + M[6] * d[2],
              M[8] * d[0] + M[9] * d[1] + M[10] * d[2]};
        double t, u, v;
        bool tie;
        const uint32_t prog = ah ? ah->intersection(tl.inst[col.inst].sbt_offset, tr.geom) : 0u;
        if (!proc_hit(prog, tr, oo, dd, tmin, tmax, t, u, v, no, tie)) no[0] = no[1] = no[2] = 0;
    } else {
        double e1[3], e2[3];
        for (int a = 0; a < 3; ++a) {
            e1[a] = (double)tr.v[3 + a] - tr.v[a];
            e2[a] = (double)tr.v[6 + a] - tr.v[a];
        }
        no[0] = e1[1] * e2[2] - e1[2] * e2[1];
        no[1] = e1[2] * e2[0] - e1[0] * e2[2];
        no[2] = e1[0] * e2[1] - e1[1] * e2[0];
    }
    double nw[3];
    for (int c = 0; c < 3; ++c) nw[c] = M[c] * no[0] + M[4 + c] * no[1] + M[8 + c] * no[2];
    double l = std::sqrt(nw[0] * nw[0] + nw[1] * nw[1] + nw[2] * nw[2]);
    for (int c = 0; c < 3; ++c) h.n[c] = l > 0 ? nw[c] / l : 0.0;
    return h;
}

inline double ufl(uint32_t v) { return (double)(v >> 8) * (1.0 / 16777216.0); }

void pcg4d_one(uint32_t x, uint32_t y, uint32_t z, uint32_t w, uint32_t* o) {
    uint32_t in[4] = {x, y, z, w};
    oracle_pcg4d(in, 1, o);
}

}  // namespace

extern "C" {

// FP64 restatement of rtx_cmd_render for pixels [0, W*H): adds each sample's radiance to accum (3 doubles per
// pixel).  records: rtx_hit_record array (48 B each); lights: 12 floats per triangle.
void oracle_render(int tlas, const void* records, uint32_t nrec, const float* cam, uint32_t W, uint32_t H,
                   uint32_t spp, uint32_t sample_offset, uint32_t max_depth, uint32_t rr_depth, uint32_t seed,
                   const float* lights, uint32_t nlights, float ray_offset, double* accum, int threads) {
    const Tlas& tl = *g_tlas[tlas];
    float axf, ayf;
    oracle_aspect(cam[3], W, H, &axf, &ayf);
    const double ax = axf, ay = ayf;
    std::vector<double> cdf(nlights + 1, 0.0);
    double acc = 0;
    for (uint32_t i = 0; i < nlights; ++i) {
        const float* L = lights + 12 * i;
        double e1[3] = {(double)L[3] - L[0], (double)L[4] - L[1], (double)L[5] - L[2]};
        double e2[3] = {(double)L[6] - L[0], (double)L[7] - L[1], (double)L[8] - L[2]};
        double c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        acc += 0.5 * std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        cdf[i] = acc;
    }
    const double total = acc;
    const uint8_t* rec = (const uint8_t*)records;
    AnyHitTable aht;  // any-hit programs of non-opaque geometry (SBT offset 0, stride 1)
    aht.records = rec;
    aht.nrec = nrec;
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::atomic<int64_t> next{0};
    const int64_t npix = (int64_t)W * H;
    auto work = [&]() {
        for (;;) {
            int64_t p = next.fetch_add(1);
            if (p >= npix) break;
            uint32_t x = (uint32_t)(p % W), y = (uint32_t)(p / W);
            for (uint32_t si = 0; si < spp; ++si) {
                uint32_t sample = sample_offset + si;
                uint32_t j[4];
                pcg4d_one(x, y, sample, seed ^ 0xA511E9B3u, j);
                double sx = ((double)x + ufl(j[0])) / W, sy = ((double)y + ufl(j[1])) / H;
                double o[3] = {cam[0], cam[1], cam[2]};
                double d[3] = {(sx * 2 - 1) * ax, -((sy * 2 - 1) * ay), -1.0};
                double thr[3] = {1, 1, 1}, rad[3] = {0, 0, 0};
                for (uint32_t depth = 0; depth < max_depth; ++depth) {
                    PHit h = pt_closest(tl, o, d, 0.0, INFINITY, &aht);
                    if (!h.hit) break;
                    uint32_t ri = tl.inst[h.inst].sbt_offset + h.geom;
                    if (ri >= nrec) break;
                    const uint8_t* R = rec + 48 * (size_t)ri;
                    uint32_t prog;
                    float colf[3], emf[3];
                    std::memcpy(&prog, R, 4);
                    std::memcpy(colf, R + 4, 12);
                    std::memcpy(emf, R + 16, 12);
                    const bool lambert = prog == 3u;
                    double dn = h.n[0] * d[0] + h.n[1] * d[1] + h.n[2] * d[2];
                    bool front = dn < 0;
                    if (depth == 0 && front && lambert)
                        for (int c = 0; c < 3; ++c) rad[c] += thr[c] * emf[c];
                    double nf[3];
                    for (int c = 0; c < 3; ++c) nf[c] = front ? h.n[c] : -h.n[c];
                    double P[3], org[3];
                    for (int c = 0; c < 3; ++c) {
                        P[c] = o[c] + h.t * d[c];
                        org[c] = P[c] + (double)ray_offset * nf[c];
                    }
                    bool reflective = colf[0] > 0 || colf[1] > 0 || colf[2] > 0;
                    if (!reflective) break;
                    uint32_t ra[4], rb[4];
                    pcg4d_one(x, y, sample, seed + 2u * depth, ra);
                    pcg4d_one(x, y, sample, seed + 2u * depth + 1u, rb);
                    if (nlights > 0) {
                        double ut = ufl(ra[0]) * total;
                        uint32_t li = 0;
                        while (li + 1 < nlights && cdf[li] <= ut) ++li;
                        const float* L = lights + 12 * li;
                        double su = std::sqrt(ufl(ra[1]));
                        double b0 = 1 - su, b1 = ufl(ra[2]) * su, b2 = 1 - b0 - b1;
                        double lp[3], e1[3], e2[3];
                        for (int c = 0; c < 3; ++c) {
                            lp[c] = b0 * L[c] + b1 * L[3 + c] + b2 * L[6 + c];
                            e1[c] = (double)L[3 + c] - L[c];
                            e2[c] = (double)L[6 + c] - L[c];
                        }
                        double nl[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                                        e1[0] * e2[1] - e1[1] * e2[0]};
                        double nll = std::sqrt(nl[0] * nl[0] + nl[1] * nl[1] + nl[2] * nl[2]);
                        double wi[3] = {lp[0] - org[0], lp[1] - org[1], lp[2] - org[2]};
                        double d2 = wi[0] * wi[0] + wi[1] * wi[1] + wi[2] * wi[2];
                        if (nll > 0 && d2 > 0) {
                            double id = 1.0 / std::sqrt(d2);
                            double cx = (nf[0] * wi[0] + nf[1] * wi[1] + nf[2] * wi[2]) * id;
                            double cl = -(nl[0] * wi[0] + nl[1] * wi[1] + nl[2] * wi[2]) * id / nll;
                            if (cx > 0 && cl > 0) {
                                PHit sh = pt_closest(tl, org, wi, 0.0, 1.0 - 1e-4, &aht);
                                if (!sh.hit) {
                                    double g = cx * cl / d2 * total / 3.14159265358979323846;
                                    for (int c = 0; c < 3; ++c) rad[c] += thr[c] * colf[c] * L[9 + c] * g;
                                }
                            }
                        }
                    }
                    if (depth + 1 >= max_depth) break;
                    double tn[3] = {thr[0] * colf[0], thr[1] * colf[1], thr[2] * colf[2]};
                    if (depth >= rr_depth) {
                        double pr = std::min(0.95, std::max(tn[0], std::max(tn[1], tn[2])));
                        if (ufl(ra[3]) >= pr) break;
                        for (int c = 0; c < 3; ++c) tn[c] /= pr;
                    }
                    for (int c = 0; c < 3; ++c) thr[c] = tn[c];
                    // cosine-weighted direction around nf, basis of Duff et al. 2017
                    double sign = std::copysign(1.0, nf[2]);
                    double a = -1.0 / (sign + nf[2]);
                    double bb = nf[0] * nf[1] * a;
                    double t[3] = {1.0 + sign * nf[0] * nf[0] * a, sign * bb, -sign * nf[0]};
                    double b[3] = {bb, sign + nf[1] * nf[1] * a, -nf[1]};
                    double u1 = ufl(rb[0]), u2 = ufl(rb[1]);
                    double r = std::sqrt(u1), phi = 2 * 3.14159265358979323846 * u2;
                    double lx = r * std::cos(phi), ly = r * std::sin(phi), lz = std::sqrt(std::max(0.0, 1 - u1));
                    for (int c = 0; c < 3; ++c) {
                        d[c] = lx * t[c] + ly * b[c] + lz * nf[c];
                        o[c] = org[c];
                    }
                }
                for (int c = 0; c < 3; ++c) accum[3 * p + c] += rad[c];
            }
        }
    };
    std::vector<std::thread> pool;
    for (int k = 1; k < threads; ++k) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
}

}  // extern "C"

// ---- neural radiance cache MLP (restatement of the GPU network, FP64) -----------------------------------------
namespace {

constexpr int kNrcParams = 23888;
const int kIn[6] = {96, 64, 64, 64, 64, 64};
const int kOut[6] = {64, 64, 64, 64, 64, 16};
int nrc_woff(int l) { return l == 0 ? 0 : 6208 + (l - 1) * 4160; }
int nrc_boff(int l) { return nrc_woff(l) + kOut[l] * kIn[l]; }

// IEEE binary32 -> bfloat16 -> binary32, round to nearest even
float bf16r(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) return x;
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

// sin(pi x) and cos(pi x) with the argument reduced exactly to the nearest multiple of 1/2 (x = n/2 + f,
// |f| <= 1/4), so that multiples of 1/2 give exactly 0 and +-1, as the GPU's sincospif does.
void sincospi_d(double x, double* s, double* c) {
    double r = std::fmod(x, 2.0);
    if (r < 0.0) r += 2.0;
    const double n = std::nearbyint(2.0 * r);
    const double f = r - 0.5 * n;  // exact
    const double sf = std::sin(3.14159265358979323846 * f), cf = std::cos(3.14159265358979323846 * f);
    switch ((int)n & 3) {
        case 0: *s = sf; *c = cf; break;
        case 1: *s = cf; *c = -sf; break;
        case 2: *s = -sf; *c = -cf; break;
        default: *s = -cf; *c = sf; break;
    }
}
double sinpi_d(double x) {
    double s, c;
    sincospi_d(x, &s, &c);
    return s;
}
double cospi_d(double x) {
    double s, c;
    sincospi_d(x, &s, &c);
    return c;
}

// features of one query (12 floats: position, normal, direction, albedo)
void nrc_features(const float* q, const float* lo, const float* hi, double* f) {
    float pn[3];
    for (int a = 0; a < 3; ++a) {
        float e = hi[a] - lo[a];
        float inv = e > 0.0f ? 1.0f / e : 0.0f;
        float d = q[a] - lo[a];
        pn[a] = d * inv;
    }
    for (int i = 0; i < 96; ++i) {
        double v = 0.0;
        if (i < 72) {
            int a = i / 24, r = i % 24, k = r >> 1;
            double x = std::ldexp((double)pn[a], k);
            v = (r & 1) ? cospi_d(x) : sinpi_d(x);
        } else if (i < 81) {
            v = q[3 + (i - 72)];
        }
        f[i] = v;
    }
}

}  // namespace

extern "C" {

// Forward + backward of the cache MLP over n samples (16 floats each: query[12], target[3], pad).  With bf16 != 0
// the GPU's rounding points are emulated (weights, features, activations and layer gradients rounded to bf16,
// FP64 accumulation).  Outputs (nullable): out[n*3] predictions (before clamping), grad[23888] = d(mean loss)/dp,
// loss[1] = mean relative L2 loss.
void oracle_nrc(const float* params, const float* samples, int n, const float* lo, const float* hi, int bf16,
                double* out, double* grad, double* loss) {
    auto R = [&](double x) -> double { return bf16 ? (double)bf16r((float)x) : x; };
    std::vector<double> W(kNrcParams);
    for (int i = 0; i < kNrcParams; ++i) W[i] = bf16 ? (double)bf16r(params[i]) : (double)params[i];
    // biases are used in FP32 on the GPU
    for (int l = 0; l < 6; ++l)
        for (int r = 0; r < kOut[l]; ++r) W[nrc_boff(l) + r] = params[nrc_boff(l) + r];
    if (grad)
        for (int i = 0; i < kNrcParams; ++i) grad[i] = 0.0;
    double total = 0.0;
    std::vector<double> act[7];
    for (int b = 0; b < n; ++b) {
        const float* sm = samples + 16 * b;
        act[0].assign(96, 0.0);
        nrc_features(sm, lo, hi, act[0].data());
        for (int i = 0; i < 96; ++i) act[0][i] = R(act[0][i]);
        for (int l = 0; l < 6; ++l) {
            act[l + 1].assign(kOut[l], 0.0);
            for (int r = 0; r < kOut[l]; ++r) {
                double acc = W[nrc_boff(l) + r];
                for (int c = 0; c < kIn[l]; ++c) acc += W[nrc_woff(l) + r * kIn[l] + c] * act[l][c];
                act[l + 1][r] = l < 5 ? R(std::max(0.0, acc)) : acc;
            }
        }
        std::vector<double> g(16, 0.0);
        for (int i = 0; i < 3; ++i) {
            double y = act[6][i], t = sm[12 + i];
            double den = y * y + 0.01;
            total += (y - t) * (y - t) / den;
            g[i] = 2.0 * (y - t) / den / n;
            if (out) out[3 * b + i] = y;
        }
        if (!grad) continue;
        for (int l = 5; l >= 0; --l) {
            for (int r = 0; r < kOut[l]; ++r) g[r] = R(g[r]);
            for (int r = 0; r < kOut[l]; ++r) {
                grad[nrc_boff(l) + r] += g[r];
                for (int c = 0; c < kIn[l]; ++c) grad[nrc_woff(l) + r * kIn[l] + c] += g[r] * act[l][c];
            }
            if (l == 0) break;
            std::vector<double> gp(kIn[l], 0.0);
            for (int c = 0; c < kIn[l]; ++c) {
                double acc = 0.0;
                for (int r = 0; r < kOut[l]; ++r) acc += W[nrc_woff(l) + r * kIn[l] + c] * g[r];
                gp[c] = act[l][c] > 0.0 ? acc : 0.0;
            }
            g = gp;
        }
    }
    if (loss) *loss = total / n;
}

// One Adam step on FP32 parameters (restatement of the update: bias-corrected moments).
void oracle_adam(float* params, float* m, float* v, const double* grad, int n, float lr, float b1, float b2,
                 float eps, int step) {
    double c1 = 1.0 / (1.0 - std::pow((double)b1, step)), c2 = 1.0 / (1.0 - std::pow((double)b2, step));
    for (int i = 0; i < n; ++i) {
        double g = grad[i];
        double mm = b1 * (double)m[i] + (1.0 - b1) * g;
        double vv = b2 * (double)v[i] + (1.0 - b2) * g * g;
        m[i] = (float)mm;
        v[i] = (float)vv;
        params[i] = (float)((double)params[i] - lr * (mm * c1) / (std::sqrt(vv * c2) + eps));
    }
}

}  // extern "C"
