// trace.hip — persistent BVH8 traversal for gfx950: ray buffers (closest / any hit) and the reference raygen.
//
// Replaces the driver traversal behind traceRayEXT (call site main.cpp:1047-1052) and the raygen / miss /
// closest-hit programs (main.cpp:1033-1091).
//
// Per lane (wave64; one persistent grid; a wave takes ray ids from an atomic counter in chunks and refills its
// idle lanes once fewer than refill_lanes of 64 are busy, checked every refill_interval steps):
//   * world ray (FP32) -> TLAS: nodes are tested with a conservative quantised slab test, instance leaves
//     enter their BLAS;
//   * BLAS entry (FP64): the ray is transformed by the FP64 inverse instance transform and its origin moved
//     onto the BLAS root box (t_e = entry distance), then rounded to FP32.  All FP32 work inside the BLAS
//     happens on coordinates of the size of the BLAS box, not of the distance to the camera;
//   * triangles: watertight test (Woop, Benthin & Wald 2013) with the shear written as three row vectors, so
//     no per-lane component selection is needed per triangle; exact-zero edge functions are recomputed in FP64;
//   * traversal order: children are stored by octant; the group mask is permuted by the ray's sign octant so
//     that lowest-bit-first visits near children first (Ylitie, Karras & Laine 2017 node groups);
//   * stack: per-lane entries (slot base, child mask, valid mask, flags) in LDS, spilling to a global
//     per-thread area past the LDS depth; overflow beyond that raises kErrStackOverflow;
//   * closest hit: t, u, v are recomputed in FP64 (Moller-Trumbore) for the winning triangle;
//   * the FLAGS variant adds the traceRayEXT ray flags, triangle opacity (geometry / instance / ray) and the
//     any-hit programs of non-opaque candidates; the host uses it only where they can change a result;
//   * a lane's ended ray is finished at the wave's next refill point, together with the wave's other ended rays
//     (result writes and their FP64 refinement run with many lanes instead of one or two); in the whole-path
//     modes (training paths, pixel paths) finishing means shading the vertex and continuing the lane's path
//     with its next ray, so one launch traces whole paths.
#include <cmath>

#include "pt_shade.h"
#include "rtx_internal.h"
#include "trace.h"

namespace rtx {

namespace {

constexpr float kErrScale = 0x1p-20f;
constexpr float kTinyDir = 0x1p-80f;
constexpr uint32_t kEntryLeaf = 1u << 16;
constexpr uint32_t kEntryTlas = 1u << 17;
// Triangles a lane tests per leaf step; the rest of its leaf group waits for a later leaf step, so lanes with
// many candidates do not hold the wave in a long loop with few lanes active (2 measured best: +2.5% closest hit
// coherent, +2.7% path tracing against testing the whole group).
constexpr uint32_t kTrisPerStep = 2u;
constexpr uint32_t kMaxSteps = 1u << 22;  // per ray; a valid BVH never needs more than ~2 steps per slot

// traceRayEXT ray flags (gl_RayFlags*EXT values) and instance flags (VkGeometryInstanceFlagBitsKHR)
constexpr uint32_t kRayTerminateOnFirstHit = RTX_RAY_FLAG_TERMINATE_ON_FIRST_HIT;
constexpr uint32_t kRaySkipClosestHit = RTX_RAY_FLAG_SKIP_CLOSEST_HIT_SHADER;
constexpr uint32_t kRayCullBack = RTX_RAY_FLAG_CULL_BACK_FACING_TRIANGLES;
constexpr uint32_t kRayCullFront = RTX_RAY_FLAG_CULL_FRONT_FACING_TRIANGLES;
constexpr uint32_t kRayCullOpaque = RTX_RAY_FLAG_CULL_OPAQUE;
constexpr uint32_t kRaySkipTriangles = RTX_RAY_FLAG_SKIP_TRIANGLES;
constexpr uint32_t kRayOpaque = RTX_RAY_FLAG_OPAQUE;
constexpr uint32_t kRayNoOpaque = RTX_RAY_FLAG_NO_OPAQUE;
constexpr uint32_t kRayCullNoOpaque = RTX_RAY_FLAG_CULL_NO_OPAQUE;
constexpr uint32_t kRaySkipAabbs = RTX_RAY_FLAG_SKIP_AABBS;
// FLAGS variant: the current instance register holds the instance id in bits 0-29 and the instance's
// FORCE_OPAQUE (bit 30) / FORCE_NO_OPAQUE (bit 31) flags.
constexpr uint32_t kInstIdMask = 0x3FFFFFFFu;


__device__ __forceinline__ uint32_t oct_permute(uint32_t m, uint32_t oct) {
    if (oct & 1u) m = ((m & 0x55u) << 1) | ((m >> 1) & 0x55u);
    if (oct & 2u) m = ((m & 0x33u) << 2) | ((m >> 2) & 0x33u);
    if (oct & 4u) m = ((m & 0x0Fu) << 4) | ((m >> 4) & 0x0Fu);
    return m;
}

// 1/d for the slab tests, |d| clamped to at least 2^-80.  v_rcp_f32 (1 ulp) on d * 2^-32 keeps the operand and
// the result in the normal range for every |d| in [2^-80, 2^128); the node test's error margin (2^-20 relative
// per term) covers the approximate reciprocal.
__device__ __forceinline__ float safe_rcp(float d) {
    float a = fabsf(d) < kTinyDir ? copysignf(kTinyDir, d) : d;
    return __builtin_amdgcn_rcpf(a * 0x1p-32f) * 0x1p-32f;
}

__device__ __forceinline__ float byte_f(uint32_t w, int b) {
    return (float)((w >> (8 * b)) & 255u);
}

struct Space {
    float ox, oy, oz;
    float dx, dy, dz;
    float ix, iy, iz;
    uint32_t oct;
    float toff;
};

__device__ __forceinline__ void setup_space(Space& s) {
    s.ix = safe_rcp(s.dx);
    s.iy = safe_rcp(s.dy);
    s.iz = safe_rcp(s.dz);
    s.oct = (s.dx < 0.0f ? 1u : 0u) | (s.dy < 0.0f ? 2u : 0u) | (s.dz < 0.0f ? 4u : 0u);
}

// Facing culls.  A triangle is front-facing when v0, v1, v2 appear clockwise seen from the ray origin, i.e.
// dot(d, (v1 - v0) x (v2 - v0)) > 0 in object space; an instance with TRIANGLE_FLIP_FACING makes counter-
// clockwise front-facing, and TRIANGLE_FACING_CULL_DISABLE (the reference's instances, main.cpp:852) exempts it.
__device__ __forceinline__ bool facing_culled(const uint4 a, const uint4 b, const uint4 c, const Space& s,
                                              uint32_t ray_flags, uint32_t inst_flags) {
    if (inst_flags & RTX_INSTANCE_TRIANGLE_FACING_CULL_DISABLE_BIT) return false;
    const float e1x = __uint_as_float(a.w) - __uint_as_float(a.x);
    const float e1y = __uint_as_float(b.x) - __uint_as_float(a.y);
    const float e1z = __uint_as_float(b.y) - __uint_as_float(a.z);
    const float e2x = __uint_as_float(b.z) - __uint_as_float(a.x);
    const float e2y = __uint_as_float(b.w) - __uint_as_float(a.y);
    const float e2z = __uint_as_float(c.x) - __uint_as_float(a.z);
    const float nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
    bool front = s.dx * nx + s.dy * ny + s.dz * nz > 0.0f;
    if (inst_flags & RTX_INSTANCE_TRIANGLE_FLIP_FACING_BIT) front = !front;
    return front ? (ray_flags & kRayCullFront) != 0 : (ray_flags & kRayCullBack) != 0;
}

// Quantised 8-child slab test.  Returns the child-position hit mask.
__device__ __forceinline__ uint32_t test_node(const uint4 w0, const uint4 w1, const uint4 w2, const uint4 w3,
                                              const Space& s, const float* rlo, const float* qs, float smin,
                                              float smax) {
    const float pqx = (float)(w0.x & 0xFFFFu), pqy = (float)(w0.x >> 16), pqz = (float)(w0.y & 0xFFFFu);
    const uint32_t ex = (w0.y >> 16) & 255u, ey = w0.y >> 24, ez = w0.z & 255u;
    const uint32_t vmask = (w0.z >> 16) & 255u;
    const float px = fmaf(pqx, qs[0], rlo[0]);
    const float py = fmaf(pqy, qs[1], rlo[1]);
    const float pz = fmaf(pqz, qs[2], rlo[2]);
    const float sx = __uint_as_float(ex << 23), sy = __uint_as_float(ey << 23), sz = __uint_as_float(ez << 23);
    const float Ax = (px - s.ox) * s.ix, Ay = (py - s.oy) * s.iy, Az = (pz - s.oz) * s.iz;
    const float Sx = sx * s.ix, Sy = sy * s.iy, Sz = sz * s.iz;
    const float ex_ = (fabsf(Ax) + 255.0f * fabsf(Sx) + fabsf(px * s.ix)) * kErrScale;
    const float ey_ = (fabsf(Ay) + 255.0f * fabsf(Sy) + fabsf(py * s.iy)) * kErrScale;
    const float ez_ = (fabsf(Az) + 255.0f * fabsf(Sz) + fabsf(pz * s.iz)) * kErrScale;
    const float Anx = Ax - ex_, Afx = Ax + ex_;
    const float Any = Ay - ey_, Afy = Ay + ey_;
    const float Anz = Az - ez_, Afz = Az + ez_;
    // byte arrays: qlo x = (w1.x, w1.y), qlo y = (w1.z, w1.w), qlo z = (w2.x, w2.y),
    //              qhi x = (w2.z, w2.w), qhi y = (w3.x, w3.y), qhi z = (w3.z, w3.w)
    const bool nx = s.ix < 0.0f, ny = s.iy < 0.0f, nz = s.iz < 0.0f;
    const uint32_t nxa = nx ? w2.z : w1.x, nxb = nx ? w2.w : w1.y, fxa = nx ? w1.x : w2.z, fxb = nx ? w1.y : w2.w;
    const uint32_t nya = ny ? w3.x : w1.z, nyb = ny ? w3.y : w1.w, fya = ny ? w1.z : w3.x, fyb = ny ? w1.w : w3.y;
    const uint32_t nza = nz ? w3.z : w2.x, nzb = nz ? w3.w : w2.y, fza = nz ? w2.x : w3.z, fzb = nz ? w2.y : w3.w;
    uint32_t hit = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int b = k & 3;
        const uint32_t wnx = k < 4 ? nxa : nxb, wfx = k < 4 ? fxa : fxb;
        const uint32_t wny = k < 4 ? nya : nyb, wfy = k < 4 ? fya : fyb;
        const uint32_t wnz = k < 4 ? nza : nzb, wfz = k < 4 ? fza : fzb;
        const float tnx = fmaf(byte_f(wnx, b), Sx, Anx);
        const float tny = fmaf(byte_f(wny, b), Sy, Any);
        const float tnz = fmaf(byte_f(wnz, b), Sz, Anz);
        const float tfx = fmaf(byte_f(wfx, b), Sx, Afx);
        const float tfy = fmaf(byte_f(wfy, b), Sy, Afy);
        const float tfz = fmaf(byte_f(wfz, b), Sz, Afz);
        const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, smin));
        const float tf = fminf(fminf(tfx, tfy), fminf(tfz, smax));
        hit |= (tn <= tf) ? (1u << k) : 0u;
    }
    return hit & vmask;
}

struct Shear {
    float rx0, rx1, rx2, ry0, ry1, ry2, rz0, rz1, rz2;
};

__device__ __forceinline__ void setup_shear(const Space& s, Shear& sh) {
    const float ax = fabsf(s.dx), ay = fabsf(s.dy), az = fabsf(s.dz);
    int kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    const float dkz = kz == 0 ? s.dx : (kz == 1 ? s.dy : s.dz);
    if (dkz < 0.0f) {
        int t = kx;
        kx = ky;
        ky = t;
    }
    const float dkx = kx == 0 ? s.dx : (kx == 1 ? s.dy : s.dz);
    const float dky = ky == 0 ? s.dx : (ky == 1 ? s.dy : s.dz);
    const float Sx = dkx / dkz, Sy = dky / dkz, Sz = 1.0f / dkz;
    sh.rx0 = kx == 0 ? 1.0f : (kz == 0 ? -Sx : 0.0f);
    sh.rx1 = kx == 1 ? 1.0f : (kz == 1 ? -Sx : 0.0f);
    sh.rx2 = kx == 2 ? 1.0f : (kz == 2 ? -Sx : 0.0f);
    sh.ry0 = ky == 0 ? 1.0f : (kz == 0 ? -Sy : 0.0f);
    sh.ry1 = ky == 1 ? 1.0f : (kz == 1 ? -Sy : 0.0f);
    sh.ry2 = ky == 2 ? 1.0f : (kz == 2 ? -Sy : 0.0f);
    sh.rz0 = kz == 0 ? Sz : 0.0f;
    sh.rz1 = kz == 1 ? Sz : 0.0f;
    sh.rz2 = kz == 2 ? Sz : 0.0f;
}

// Watertight test in the current (rebased object) space.  Returns true and the parametric distance s when
// the triangle is hit with s in [smin, smax].
__device__ __forceinline__ bool test_triangle(const uint4 a, const uint4 b, const uint4 c, const Space& s,
                                              const Shear& sh, float smin, float smax, float& t_out, float& u_out,
                                              float& v_out) {
    const float v0x = __uint_as_float(a.x) - s.ox, v0y = __uint_as_float(a.y) - s.oy, v0z = __uint_as_float(a.z) - s.oz;
    const float v1x = __uint_as_float(a.w) - s.ox, v1y = __uint_as_float(b.x) - s.oy, v1z = __uint_as_float(b.y) - s.oz;
    const float v2x = __uint_as_float(b.z) - s.ox, v2y = __uint_as_float(b.w) - s.oy, v2z = __uint_as_float(c.x) - s.oz;
    const float Ax = fmaf(v0z, sh.rx2, fmaf(v0y, sh.rx1, v0x * sh.rx0));
    const float Ay = fmaf(v0z, sh.ry2, fmaf(v0y, sh.ry1, v0x * sh.ry0));
    const float Az = fmaf(v0z, sh.rz2, fmaf(v0y, sh.rz1, v0x * sh.rz0));
    const float Bx = fmaf(v1z, sh.rx2, fmaf(v1y, sh.rx1, v1x * sh.rx0));
    const float By = fmaf(v1z, sh.ry2, fmaf(v1y, sh.ry1, v1x * sh.ry0));
    const float Bz = fmaf(v1z, sh.rz2, fmaf(v1y, sh.rz1, v1x * sh.rz0));
    const float Cx = fmaf(v2z, sh.rx2, fmaf(v2y, sh.rx1, v2x * sh.rx0));
    const float Cy = fmaf(v2z, sh.ry2, fmaf(v2y, sh.ry1, v2x * sh.ry0));
    const float Cz = fmaf(v2z, sh.rz2, fmaf(v2y, sh.rz1, v2x * sh.rz0));
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    float T = U * Az + V * Bz + W * Cz;
    const float adet = fabsf(det);
    const float Ts = det < 0.0f ? -T : T;
    if (Ts < smin * adet || Ts > smax * adet) return false;
    t_out = Ts / adet;
    const float rdet = 1.0f / det;
    u_out = V * rdet;
    v_out = W * rdet;
    return true;
}

struct TraceArgs {
    const uint8_t* tlas;
    const rtx_ray* rays;
    const uint32_t* n_ptr;  // when set, the ray count is read from device memory (queues)
    const rtx_ray* rays2;   // kModeFused: the any-hit queue and its device count
    const uint32_t* n2_ptr;
    PathHit* path_hits;
    uint32_t n;
    uint32_t cull_mask;
    rtx_hit* hits;
    uint32_t* occluded;
    // raygen
    float cam[3];
    float ax, ay;
    uint32_t width, height;
    float tmin, tmax;
    const rtx_hit_record* records;
    uint32_t nrec, sbt_offset, sbt_stride;
    float miss[3];
    float* film;
    // runtime
    DeviceState* state;
    uint2* gstack;
    uint32_t lds_depth;
    uint32_t gdepth;
    uint32_t ray_chunk;        // ray ids per queue fetch (< 64: automatic)
    uint32_t refill_lanes;     // refill a wave once fewer lanes than this are busy
    uint32_t refill_interval;  // steps between refill checks
    uint32_t leaf_lanes;       // pending leaf groups needed before a leaf step runs
    uint32_t ray_flags;        // RTX_RAY_FLAG_* of the launch
    // kModeTrain / kModePixelPaths: the path tracer's state, the first path id (npix for training paths, 0 for
    // pixel paths) and the pass's sample index
    PtArgs pt;
    uint32_t pt_first;
    uint32_t pt_sample;
};

// kModeFused: one launch over two queues of the wavefront integrator, rays [0, nA) of `rays` as kModePath
// (extension rays, path hits) and rays [nA, nA + nB) of `rays2` as kModeAny (shadow rays, occlusion flags).
// kModeTrain / kModePixelPaths: whole paths in one launch - the NRC frame's training paths, or one sample of a
// plain render's pixel paths.  A lane takes a path index from the ray pool, traces its camera ray, and whenever
// one of its rays ends (the wave handles its lanes' ended rays together at its next refill) shades the vertex
// (pathtrace.hip's per-vertex step) and continues the same path with the next-event shadow ray and then the
// continuation ray, until the path ends; then it takes a new path.
enum Mode {
    kModeClosest = 0,
    kModeAny = 1,
    kModeRaygen = 2,
    kModePath = 3,
    kModeFused = 4,
    kModeTrain = 5,
    kModePixelPaths = 6,
    kModePixelCache = 7  // the pixel paths of an NRC frame: whole paths up to their cache query
};

// FP64 recomputation of (t, u, v) for the winning triangle (Moller & Trumbore 1997) in object space.
__device__ void refine_hit(const double* inv, const float* wo, const float* wd, const TriSlot* tri, float& t,
                           float& u, float& v) {
    double o[3], d[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        o[r] = inv[4 * r] * wo[0] + inv[4 * r + 1] * wo[1] + inv[4 * r + 2] * wo[2] + inv[4 * r + 3];
        d[r] = inv[4 * r] * wd[0] + inv[4 * r + 1] * wd[1] + inv[4 * r + 2] * wd[2];
    }
    double v0[3], e1[3], e2[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        v0[a] = tri->v[a];
        e1[a] = (double)tri->v[3 + a] - v0[a];
        e2[a] = (double)tri->v[6 + a] - v0[a];
    }
    double p[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
    double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
    double inv_det = 1.0 / det;
    double sv[3] = {o[0] - v0[0], o[1] - v0[1], o[2] - v0[2]};
    double uu = (sv[0] * p[0] + sv[1] * p[1] + sv[2] * p[2]) * inv_det;
    double q[3] = {sv[1] * e1[2] - sv[2] * e1[1], sv[2] * e1[0] - sv[0] * e1[2], sv[0] * e1[1] - sv[1] * e1[0]};
    double vv = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * inv_det;
    double tt = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv_det;
    t = (float)tt;
    u = (float)uu;
    v = (float)vv;
}

// Intersection programs of AABB primitives (box lo = a.xyz, hi = a.w, b.x, b.y).  The first intersection with t
// in [tmin, tmax] of the ray o + t d with the program's surface, its (u, v) and the outward normal n (object
// space, unnormalised for boxes).  T is float for the traversal's candidate test (in the current rebased space)
// and double for the reported hit.
// FP32 world-space unit geometric normal of a triangle (the path tracer's shading normal; the instance inverse is
// rounded to FP32).
__device__ __forceinline__ void world_normal_f32(const double* inv, const TriSlot* tri, float* n) {
    float e1[3], e2[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        e1[a] = tri->v[3 + a] - tri->v[a];
        e2[a] = tri->v[6 + a] - tri->v[a];
    }
    const float no[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    float nw[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) nw[c] = (float)inv[c] * no[0] + (float)inv[4 + c] * no[1] + (float)inv[8 + c] * no[2];
    const float l = sqrtf(nw[0] * nw[0] + nw[1] * nw[1] + nw[2] * nw[2]);
    const float r = l > 0.0f ? 1.0f / l : 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) n[c] = nw[c] * r;
}

template <typename T>
__device__ __forceinline__ bool procedural_hit(uint32_t prog, const T* lo, const T* hi, const T* o, const T* d,
                                               T tmin, T tmax, T& t_out, T& u_out, T& v_out, T* n) {
    if (prog == RTX_INTERSECTION_SPHERE) {
        T c[3], oc[3];
        T r = hi[0] - lo[0];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            c[k] = (lo[k] + hi[k]) * T(0.5);
            oc[k] = o[k] - c[k];
            r = k ? (hi[k] - lo[k] < r ? hi[k] - lo[k] : r) : r;
        }
        r *= T(0.5);
        const T A = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        const T B = oc[0] * d[0] + oc[1] * d[1] + oc[2] * d[2];
        const T C = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - r * r;
        const T disc = B * B - A * C;
        if (!(disc >= T(0)) || !(A > T(0)) || !(r > T(0))) return false;
        const T sq = sqrt(disc);
        T t = (-B - sq) / A;
        if (t < tmin) t = (-B + sq) / A;
        if (!(t >= tmin && t <= tmax)) return false;
#pragma unroll
        for (int k = 0; k < 3; ++k) n[k] = (o[k] + t * d[k] - c[k]) / r;
        t_out = t;
        u_out = (n[0] + T(1)) * T(0.5);
        v_out = (n[1] + T(1)) * T(0.5);
        return true;
    }
    // RTX_INTERSECTION_BOX: slab entry, or the exit when the entry lies before tmin
    T tn = -INFINITY, tf = INFINITY;
    int an = 0, af = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (d[k] == T(0)) {
            if (o[k] < lo[k] || o[k] > hi[k]) return false;
            continue;
        }
        const T i = T(1) / d[k];
        T t1 = (lo[k] - o[k]) * i, t2 = (hi[k] - o[k]) * i;
        if (t1 > t2) {
            const T x = t1;
            t1 = t2;
            t2 = x;
        }
        if (t1 > tn) {
            tn = t1;
            an = k;
        }
        if (t2 < tf) {
            tf = t2;
            af = k;
        }
    }
    if (!(tn <= tf)) return false;
    T t;
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
    const int a1 = ax == 2 ? 0 : ax + 1, a2 = ax == 0 ? 2 : ax - 1;
    const T p1 = (a1 == 0 ? o[0] : (a1 == 1 ? o[1] : o[2])) + t * (a1 == 0 ? d[0] : (a1 == 1 ? d[1] : d[2]));
    const T p2 = (a2 == 0 ? o[0] : (a2 == 1 ? o[1] : o[2])) + t * (a2 == 0 ? d[0] : (a2 == 1 ? d[1] : d[2]));
    const T l1 = a1 == 0 ? lo[0] : (a1 == 1 ? lo[1] : lo[2]), h1 = a1 == 0 ? hi[0] : (a1 == 1 ? hi[1] : hi[2]);
    const T l2 = a2 == 0 ? lo[0] : (a2 == 1 ? lo[1] : lo[2]), h2 = a2 == 0 ? hi[0] : (a2 == 1 ? hi[1] : hi[2]);
    t_out = t;
    u_out = h1 > l1 ? (p1 - l1) / (h1 - l1) : T(0);
    v_out = h2 > l2 ? (p2 - l2) / (h2 - l2) : T(0);
    // outward normal of the hit face: -d's sign along ax at the entry, +d's at the exit
    const T dax = ax == 0 ? d[0] : (ax == 1 ? d[1] : d[2]);
    const T sgn = (t == tn) ? (dax > T(0) ? T(-1) : T(1)) : (dax > T(0) ? T(1) : T(-1));
    n[0] = ax == 0 ? sgn : T(0);
    n[1] = ax == 1 ? sgn : T(0);
    n[2] = ax == 2 ? sgn : T(0);
    return true;
}

// Intersection program of an AABB primitive's hit record (bits 8-15 of its program), RTX_INTERSECTION_BOX without
// a hit-record table.
__device__ __forceinline__ uint32_t intersection_program(const TraceArgs& args, const InstanceInfo* info,
                                                         uint32_t geom) {
    if (!args.records) return RTX_INTERSECTION_BOX;
    const uint32_t rec = info->sbt_offset + geom * args.sbt_stride + args.sbt_offset;
    return rec < args.nrec ? (args.records[rec].program >> 8) & 255u : RTX_INTERSECTION_BOX;
}

// FP64 (t, u, v) and world-space unit normal of an AABB primitive hit of the world ray (wo, wd) under the
// instance transform with inverse inv; false when FP64 finds no intersection at or after tmin (then the caller
// keeps the FP32 values).
__device__ bool refine_procedural(const double* inv, const float* wo, const float* wd, const TriSlot* tri,
                                  uint32_t prog, float tmin, float& t, float& u, float& v, float* nw) {
    double o[3], d[3], lo[3], hi[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        o[r] = inv[4 * r] * wo[0] + inv[4 * r + 1] * wo[1] + inv[4 * r + 2] * wo[2] + inv[4 * r + 3];
        d[r] = inv[4 * r] * wd[0] + inv[4 * r + 1] * wd[1] + inv[4 * r + 2] * wd[2];
        lo[r] = tri->v[r];
        hi[r] = tri->v[3 + r];
    }
    double tt, uu, vv, n[3];
    if (!procedural_hit<double>(prog, lo, hi, o, d, (double)tmin, (double)INFINITY, tt, uu, vv, n)) return false;
    t = (float)tt;
    u = (float)uu;
    v = (float)vv;
    if (nw) {
        double w[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) w[c] = inv[c] * n[0] + inv[4 + c] * n[1] + inv[8 + c] * n[2];
        const double l = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        const double rl = l > 0.0 ? 1.0 / l : 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) nw[c] = (float)(w[c] * rl);
    }
    return true;
}

// The any-hit program of a non-opaque candidate (its hit record: instance SBT offset + geometry index x stride +
// ray SBT offset, as for closest hits) on the candidate's barycentrics (u, v) from the watertight test (FP32):
// true to accept it.  Without a hit-record table, or with a record index outside it, the candidate is accepted
// (a hit group without an any-hit shader).
__device__ __forceinline__ bool any_hit_accepts(const TraceArgs& args, const InstanceInfo* info, uint32_t geom,
                                                float u, float v) {
    if (!args.records) return true;
    const uint32_t rec = info->sbt_offset + geom * args.sbt_stride + args.sbt_offset;
    if (rec >= args.nrec) return true;
    const uint32_t ah = args.records[rec].any_hit;
    if ((ah & 255u) != RTX_ANY_HIT_CUTOUT) return true;
    const float n = (ah >> 8) ? (float)(ah >> 8) : 1.0f;
    const int cell = (int)floorf(u * n) + (int)floorf(v * n);
    return (cell & 1) == 0;
}

// Opacity of a candidate (geometry bit, overridden by the instance's FORCE flags, overridden by the ray's OPAQUE /
// NO_OPAQUE), the CULL_OPAQUE / CULL_NO_OPAQUE culls, then the any-hit program of a non-opaque candidate.  Used by
// the FLAGS variant only; the default variant treats every triangle as opaque and runs only where that holds.
__device__ __forceinline__ bool candidate_accepted(const TraceArgs& args, uint32_t ray_flags, uint32_t tri_bits,
                                                   uint32_t geom, uint32_t cur_inst, const InstanceInfo* table,
                                                   float u, float v) {
    bool opaque = !(tri_bits & kTriNonOpaque);
    const uint32_t force = cur_inst >> 30;
    if (force) opaque = (force & 1u) != 0;
    if (ray_flags & kRayOpaque) opaque = true;
    else if (ray_flags & kRayNoOpaque) opaque = false;
    if ((opaque && (ray_flags & kRayCullOpaque)) || (!opaque && (ray_flags & kRayCullNoOpaque))) return false;
    if (opaque) return true;
    return any_hit_accepts(args, table + (cur_inst & kInstIdMask), geom, u, v);
}

template <int MODE, bool STATS, bool FLAGS>
__device__ __forceinline__ void trace_body(const TraceArgs& args) {
    const uint32_t ray_flags = FLAGS ? args.ray_flags : 0u;  // the default variant compiles the flags out
    // dynamic LDS: the per-lane stack (lds_depth entries per lane); the FLAGS variant and the path-tracer modes
    // add the lane's world ray (6 floats per lane), which they read back where they need it (leaving a BLAS, the
    // final FP64 refinement) instead of keeping it in registers (measured: spill-free at 128 VGPRs and 1% faster
    // path tracing); the other variants keep it in wo / wd
    constexpr bool kTrain = MODE == kModeTrain || MODE == kModePixelPaths || MODE == kModePixelCache;  // whole paths
    constexpr int kShadeMode = MODE == kModePixelPaths ? (int)kPtPlain : (int)kPtCache;
    constexpr bool kRayLds = FLAGS || MODE == kModePath || MODE == kModeFused || kTrain;
    extern __shared__ uint2 lds_stack[];
    float* const lds_ray = (float*)(lds_stack + args.lds_depth * blockDim.x) + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint64_t gtid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const AsHeader* th = (const AsHeader*)args.tlas;
    const uint8_t* tslots = args.tlas + kHeaderBytes;
    const InstanceInfo* table = (const InstanceInfo*)(args.tlas + th->inst_table_offset);
    const float trlo[3] = {th->lo[0], th->lo[1], th->lo[2]};
    const float tqs[3] = {th->qscale[0], th->qscale[1], th->qscale[2]};
    const bool tlas_empty = th->slot_count == 0;
    const uint32_t tlas_nslots = th->slot_count;
    const uint32_t tlas_ninst = th->prim_count;
    const uint32_t D = args.lds_depth;
    const uint32_t GD = args.gdepth;
    // ray flags under which no triangle can be accepted
    const bool cull_all = ((ray_flags & kRaySkipTriangles) && (ray_flags & kRaySkipAabbs)) ||
                          ((ray_flags & kRayCullOpaque) && (ray_flags & (kRayOpaque | kRayCullNoOpaque))) ||
                          ((ray_flags & kRayCullNoOpaque) && (ray_flags & kRayNoOpaque));
    const uint32_t nA = args.n_ptr ? *args.n_ptr : args.n;
    const uint32_t nrays = MODE == kModeFused ? nA + *args.n2_ptr : nA;

    uint64_t st_nodes = 0, st_tris = 0, st_insts = 0, st_rays = 0;

    bool active = false, exhausted = false, finished = false;
    uint32_t ray_id = 0;
    float wo[3], wd[3], wtmin = 0.0f, tc = 0.0f;
    Space s;
    Shear sh;
    const uint8_t* cslots = tslots;
    float crlo[3], cqs[3];
    bool in_blas = false;
    uint32_t cur_inst = 0;
    uint32_t cur_nslots = tlas_nslots;
    uint32_t steps = 0;
    uint32_t ng_base = 0, ng_bits = 0;  // bits: ordered mask | vmask << 8
    uint32_t lg_base = 0, lg_bits = 0;  // bits: child-position mask | vmask << 8
    uint32_t sp = 0;
    uint32_t hit_inst = kNoHit, hit_slot = 0;
    bool occl = false;
    bool shadow_ray = MODE == kModeAny;  // this lane's ray is an any-hit query
    uint32_t pdepth = 0;       // kModeTrain: vertex of the path that the current ray leaves
    bool ext_pending = false;  // kModeTrain: the continuation ray waits behind the shadow ray
    uint32_t pt_rays = 0;      // kModeTrain: rays traced (path-tracer ray count)
    const uint32_t pt_sample = kTrain ? pass_sample(args.pt, args.pt_sample) : 0u;

    // a new ray (wo, wd, wtmin, tc set): world space at the TLAS root
    auto begin_ray = [&]() {
        if (kRayLds) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                lds_ray[k * blockDim.x] = wo[k];
                lds_ray[(3 + k) * blockDim.x] = wd[k];
            }
        }
        s.ox = wo[0];
        s.oy = wo[1];
        s.oz = wo[2];
        s.dx = wd[0];
        s.dy = wd[1];
        s.dz = wd[2];
        s.toff = 0.0f;
        setup_space(s);
        cslots = tslots;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            crlo[a] = trlo[a];
            cqs[a] = tqs[a];
        }
        in_blas = false;
        cur_nslots = tlas_nslots;
        steps = 0;
        ng_base = 0;
        // the root as a one-child group: child position 0 sits at ordered bit (0 ^ oct)
        ng_bits = tlas_empty || cull_all ? 0u : ((1u << s.oct) | (1u << 8));
        lg_bits = 0;
        sp = 0;
        hit_inst = kNoHit;
        occl = false;
        if (STATS) ++st_rays;
    };
    auto set_ray = [&](const rtx_ray& r) {
        wo[0] = r.origin[0];
        wo[1] = r.origin[1];
        wo[2] = r.origin[2];
        wtmin = r.tmin;
        wd[0] = r.direction[0];
        wd[1] = r.direction[1];
        wd[2] = r.direction[2];
        tc = r.tmax;
    };

    // Wave-uniform pool of ray ids [pool_next, pool_end): one atomic fetches `chunk` consecutive ids (256 when the
    // launch has at least 8 rounds of work for every wave, else 64 so that small launches still spread over all
    // waves).
    const uint32_t waves_total = gridDim.x * (blockDim.x >> 6);
    const uint32_t chunk = args.ray_chunk >= 64u ? args.ray_chunk
                           : ((uint64_t)nrays >= (uint64_t)waves_total * 64u * 8u ? 256u : 64u);
    uint32_t pool_next = 0, pool_end = 0;
    for (;;) {
        if (kTrain && finished) {
            // continue the lane's path: gather a shadow ray's contribution, or shade the vertex the continuation
            // ray reached; then trace the next ray of the path, or let the lane take a new path.  The pending
            // continuation ray and the shadow contribution wait in the path's slot (ray_id) of the queue buffers.
            const uint32_t pid = args.pt_first + ray_id;
            bool cont = false;
            if (shadow_ray) {
                if (!occl) {
                    const float4 c = args.pt.sh_contrib[ray_id];
                    float4 r = args.pt.rad[pid];
                    r.x += c.x;
                    r.y += c.y;
                    r.z += c.z;
                    args.pt.rad[pid] = r;
                }
                shadow_ray = false;
                if (ext_pending) {
                    set_ray(args.pt.ext_rays[0][ray_id]);
                    ++pdepth;
                    cont = true;
                }
            } else {
                PathHit ph;
                ph.t = 0.0f;
                ph.u = 0.0f;
                ph.v = 0.0f;
                ph.inst = hit_inst;
                ph.prim = 0;
                ph.geom = 0;
                ph.n[0] = ph.n[1] = ph.n[2] = 0.0f;
                ph.rec = 0xFFFFFFFFu;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    wo[k] = lds_ray[k * blockDim.x];
                    wd[k] = lds_ray[(3 + k) * blockDim.x];
                }
                if (hit_inst != kNoHit) {
                    const InstanceInfo* info = table + hit_inst;
                    double M[12];
#pragma unroll
                    for (int kk = 0; kk < 12; ++kk) M[kk] = info->inv[kk];
                    const TriSlot* tri =
                        (const TriSlot*)((const uint8_t*)info->blas + kHeaderBytes + (uint64_t)hit_slot * kSlotBytes);
                    ph.t = tc;
                    ph.prim = tri->prim;
                    ph.geom = tri->geom;
                    world_normal_f32(M, tri, ph.n);
                    ph.rec = info->sbt_offset + ph.geom * args.sbt_stride + args.sbt_offset;
                }
                rtx_ray cr;
                cr.origin[0] = wo[0];
                cr.origin[1] = wo[1];
                cr.origin[2] = wo[2];
                cr.tmin = wtmin;
                cr.direction[0] = wd[0];
                cr.direction[1] = wd[1];
                cr.direction[2] = wd[2];
                cr.tmax = tc;
                ShadeOut so;
                shade_vertex<kShadeMode>(args.pt, pid, cr, ph, pdepth, pt_sample, so);
                pt_rays += (so.push_sh ? 1u : 0u) + (so.push_ext ? 1u : 0u);
                if (so.push_ext) {
                    args.pt.thr[pid] = so.thr_new;
                    args.pt.ext_rays[0][ray_id] = so.ext;
                }
                ext_pending = so.push_ext;
                if (MODE == kModePixelCache) {
                    // the path ends in a cache query: one slot per query, one atomic per wave
                    const uint64_t qm = __ballot(so.push_cq);
                    if (qm) {
                        const uint32_t leader = (uint32_t)(__ffsll((long long)qm) - 1);
                        uint32_t base = 0;
                        if (lane == leader) base = atomicAdd(&args.pt.counts[2], (uint32_t)__popcll(qm));
                        base = __shfl(base, (int)leader);
                        if (so.push_cq) {
                            const uint32_t slot = base + (uint32_t)__popcll(qm & lt_mask);
                            args.pt.cq[slot] = so.cq;
                            args.pt.cq_pid[slot] = pid;
                        }
                    }
                }
                if (so.push_sh) {
                    args.pt.sh_contrib[ray_id] = so.contrib;
                    set_ray(so.sh);
                    shadow_ray = true;
                    cont = true;
                } else if (so.push_ext) {
                    set_ray(so.ext);
                    ++pdepth;
                    cont = true;
                }
            }
            if (cont) {
                begin_ray();
                active = true;
            }
            finished = false;
        }
        if (!kTrain && finished) {
            // write the result of ray_id (lanes whose rays ended since the last refill, in one pass)
            if (MODE == kModeAny || (MODE == kModeFused && shadow_ray)) {
                args.occluded[MODE == kModeFused ? ray_id - nA : ray_id] = occl ? 1u : 0u;
            } else {
                float t = 0.0f, u = 0.0f, v = 0.0f;
                uint32_t prim = 0, geom = 0;
                const InstanceInfo* info = nullptr;
                bool proc = false;  // an AABB primitive's hit: normal from its intersection program
                float pn[3] = {0.0f, 0.0f, 0.0f};
                if (hit_inst != kNoHit) {
                    info = table + hit_inst;
                    double M[12];
#pragma unroll
                    for (int kk = 0; kk < 12; ++kk) M[kk] = info->inv[kk];
                    const TriSlot* tri =
                        (const TriSlot*)((const uint8_t*)info->blas + kHeaderBytes + (uint64_t)hit_slot * kSlotBytes);
                    if (kRayLds) {
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            wo[k] = lds_ray[k * blockDim.x];
                            wd[k] = lds_ray[(3 + k) * blockDim.x];
                        }
                    }
                    if (MODE == kModePath || MODE == kModeFused) {
                        // the path tracer shades with t and the normal only: the traversal's FP32 t (u, v
                        // are not needed); procedural normals below
                        t = tc;
                        proc = FLAGS && (tri->bits & kTriAabb);
                        if (proc && !refine_procedural(M, wo, wd, tri, intersection_program(args, info, tri->geom),
                                                       wtmin, t, u, v, pn)) {
                            const float l = sqrtf(wd[0] * wd[0] + wd[1] * wd[1] + wd[2] * wd[2]);
#pragma unroll
                            for (int k = 0; k < 3; ++k) pn[k] = l > 0.0f ? -wd[k] / l : 0.0f;
                        }
                    } else if (FLAGS && (tri->bits & kTriAabb)) {
                        // keep the traversal's FP32 t if FP64 finds no intersection (the normal then faces
                        // the ray)
                        proc = true;
                        t = tc;
                        if (!refine_procedural(M, wo, wd, tri, intersection_program(args, info, tri->geom),
                                               wtmin, t, u, v,
                                               (MODE == kModePath || MODE == kModeFused) ? pn : nullptr)) {
                            const float l = sqrtf(wd[0] * wd[0] + wd[1] * wd[1] + wd[2] * wd[2]);
#pragma unroll
                            for (int k = 0; k < 3; ++k) pn[k] = l > 0.0f ? -wd[k] / l : 0.0f;
                        }
                    } else {
                        refine_hit(M, wo, wd, tri, t, u, v);
                    }
                    prim = tri->prim;
                    geom = tri->geom;
                }
                if (MODE == kModePath || MODE == kModeFused) {
                    PathHit ph;
                    ph.t = t;
                    ph.u = u;
                    ph.v = v;
                    ph.inst = hit_inst;
                    ph.prim = prim;
                    ph.geom = geom;
                    ph.n[0] = ph.n[1] = ph.n[2] = 0.0f;
                    ph.rec = 0xFFFFFFFFu;
                    if (hit_inst != kNoHit) {
                        double M[12];
#pragma unroll
                        for (int kk = 0; kk < 12; ++kk) M[kk] = info->inv[kk];
                        const TriSlot* tri = (const TriSlot*)((const uint8_t*)info->blas + kHeaderBytes +
                                                              (uint64_t)hit_slot * kSlotBytes);
                        if (FLAGS && proc) {
                            ph.n[0] = pn[0];
                            ph.n[1] = pn[1];
                            ph.n[2] = pn[2];
                        } else {
                            world_normal_f32(M, tri, ph.n);
                        }
                        ph.rec = info->sbt_offset + geom * args.sbt_stride + args.sbt_offset;
                    }
                    args.path_hits[ray_id] = ph;
                } else if (MODE == kModeClosest) {
                    uint32_t* hp = (uint32_t*)(args.hits + ray_id);
                    hp[0] = __float_as_uint(t);
                    hp[1] = __float_as_uint(u);
                    hp[2] = __float_as_uint(v);
                    hp[3] = prim;
                    hp[4] = geom;
                    hp[5] = hit_inst;
                } else {
                    float r = args.miss[0], g = args.miss[1], bl = args.miss[2];
                    if (hit_inst != kNoHit && (ray_flags & kRaySkipClosestHit)) {
                        r = g = bl = 0.0f;  // the closest-hit program does not run: the payload keeps its 0
                    } else if (hit_inst != kNoHit) {
                        const uint32_t rec = info->sbt_offset + geom * args.sbt_stride + args.sbt_offset;
                        if (rec < args.nrec) {
                            const rtx_hit_record hr = args.records[rec];
                            bool bary = hr.program == RTX_PROGRAM_BARYCENTRIC;
                            if (hr.program == RTX_PROGRAM_MATCH_BARYCENTRIC)
                                bary = (int32_t)prim == hr.match_primitive && (int32_t)hit_inst == hr.match_instance &&
                                       (int32_t)info->custom == hr.match_custom_index &&
                                       (int32_t)geom == hr.match_geometry;
                            if (bary) {
                                r = 1.0f - u - v;
                                g = u;
                                bl = v;
                            } else {
                                r = hr.color[0];
                                g = hr.color[1];
                                bl = hr.color[2];
                            }
                        } else {
                            r = g = bl = 0.0f;
                        }
                    }
                    *(float4*)(args.film + 4 * (uint64_t)ray_id) = make_float4(r, g, bl, 0.0f);
                }
            }
            finished = false;
        }
        const uint64_t need = __ballot(!active && !exhausted);
        if (need) {
            const uint32_t want = (uint32_t)__popcll(need);
            const uint32_t rank = (uint32_t)__popcll(need & lt_mask);
            const uint32_t avail = pool_end - pool_next;
            const uint32_t served = avail < want ? avail : want;
            uint32_t id = rank < served ? pool_next + rank : 0xFFFFFFFFu;
            pool_next += served;
            if (served < want && pool_end < nrays) {
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(&args.state->ray_counter, chunk);
                base = __shfl(base, 0);
                pool_next = base;
                pool_end = base + chunk;
                if (rank >= served) id = pool_next + (rank - served);
                pool_next += want - served;
            }
            if (!active && !exhausted) {
                if (id < nrays) {
                    ray_id = id;
                    active = true;
                    if (MODE == kModeRaygen) {
                        const uint32_t x = id % args.width, y = id / args.width;
                        const float sx = ((float)x + 0.5f) / (float)args.width;
                        const float sy = ((float)y + 0.5f) / (float)args.height;
                        const float nx = sx * 2.0f - 1.0f;
                        const float ny = sy * 2.0f - 1.0f;
                        wo[0] = args.cam[0];
                        wo[1] = args.cam[1];
                        wo[2] = args.cam[2];
                        wd[0] = nx * args.ax;
                        wd[1] = -(ny * args.ay);
                        wd[2] = -1.0f;
                        wtmin = args.tmin;
                        tc = args.tmax;
                    } else if (kTrain) {
                        const uint32_t pid = args.pt_first + id;
                        rtx_ray r;
                        pt_camera_ray(args.pt, pid, pt_sample, r);
                        args.pt.thr[pid] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
                        args.pt.rad[pid] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                        pdepth = 0;
                        shadow_ray = false;
                        ext_pending = false;
                        ++pt_rays;
                        set_ray(r);
                    } else {
                        if (MODE == kModeFused) shadow_ray = id >= nA;
                        const float4* rp = (const float4*)(MODE == kModeFused && id >= nA ? args.rays2 + (id - nA)
                                                                                          : args.rays + id);
                        const float4 r0 = rp[0], r1 = rp[1];
                        wo[0] = r0.x;
                        wo[1] = r0.y;
                        wo[2] = r0.z;
                        wtmin = r0.w;
                        wd[0] = r1.x;
                        wd[1] = r1.y;
                        wd[2] = r1.z;
                        tc = r1.w;
                    }
                    begin_ray();
                } else {
                    exhausted = true;
                }
            }
        }
        if (!__any(active)) break;
        const bool wave_exhausted = __any(exhausted);
        // kModeTrain: back to the refill point (where ended rays continue their paths) once a quarter of the lanes
        // that start this round (64 - refill_lanes of 64) have ended their rays, at least one
        uint32_t train_floor = 0;
        if (kTrain) {
            const uint32_t n0 = (uint32_t)__popcll(__ballot(active));
            const uint32_t q = n0 * (64u - args.refill_lanes) / 64u;
            train_floor = n0 - (q > 0u ? q : 1u);
        }

        int iter = 0;
        while (active) {
            bool done = false;
            if (++steps > kMaxSteps) {
                atomicOr(&args.state->error_flags, kErrBadAs);
                lg_bits = 0;
                ng_bits = 0;
                sp = 0;
            }
            // a lane without pending work pops its next group and processes it in this same step
            if (!(lg_bits & 255u) && !(ng_bits & 255u)) {
                if (sp > 0) {
                    --sp;
                    const uint2 e =
                        sp < D ? lds_stack[sp * blockDim.x + threadIdx.x] : args.gstack[gtid * GD + (sp - D)];
                    if ((e.y & kEntryTlas) && in_blas) {
                        in_blas = false;
                        if (kRayLds) {
#pragma unroll
                            for (int k = 0; k < 3; ++k) {
                                wo[k] = lds_ray[k * blockDim.x];
                                wd[k] = lds_ray[(3 + k) * blockDim.x];
                            }
                        }
                        s.ox = wo[0];
                        s.oy = wo[1];
                        s.oz = wo[2];
                        s.dx = wd[0];
                        s.dy = wd[1];
                        s.dz = wd[2];
                        s.toff = 0.0f;
                        setup_space(s);
                        cslots = tslots;
                        cur_nslots = tlas_nslots;
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            crlo[a] = trlo[a];
                            cqs[a] = tqs[a];
                        }
                    }
                    if (e.y & kEntryLeaf) {
                        lg_base = e.x;
                        lg_bits = e.y & 0xFFFFu;
                    } else {
                        ng_base = e.x;
                        ng_bits = e.y & 0xFFFFu;
                    }
                } else {
                    done = true;
                }
            }
            // Postponed leaves: the leaf step (instance entry in FP64, or triangle tests) runs only once enough lanes
            // of the wave hold a pending leaf group, or when no lane can traverse; until then lanes with a pending
            // leaf keep traversing their node group, or wait.
            const bool has_leaf = !done && (lg_bits & 255u) != 0u;
            const uint64_t leaf_lanes = __ballot(has_leaf);
            const uint64_t node_lanes = __ballot(!done && (ng_bits & 255u) != 0u);
            const bool leaf_phase = (uint32_t)__popcll(leaf_lanes) >= args.leaf_lanes || node_lanes == 0ull;
            if (done) {
            } else if (has_leaf && leaf_phase) {
                if (!in_blas) {
                    // instance leaf: enter its BLAS
                    const uint32_t m = lg_bits & 255u;
                    const uint32_t k = (uint32_t)(__ffs(m) - 1);
                    const uint32_t vm = lg_bits >> 8;
                    lg_bits &= ~(1u << k);
                    const uint32_t slot = lg_base + (uint32_t)__popc(vm & ((1u << k) - 1u));
                    uint4 q = make_uint4(0u, 0u, 0u, 0u);
                    if (slot < tlas_nslots) {
                        q = *(const uint4*)(tslots + (uint64_t)slot * kSlotBytes + 48);
                        if (q.z >= tlas_ninst) q.w = 0u;
                    } else {
                        atomicOr(&args.state->error_flags, kErrBadAs);
                    }
                    const uint64_t blas = (uint64_t)q.x | ((uint64_t)q.y << 32);
                    const uint32_t iid = q.z, imask = q.w;
                    if ((imask & args.cull_mask & 255u) != 0u) {
                        if (STATS) ++st_insts;
                        // push the remaining TLAS work
                        if (lg_bits & 255u) {
                            const uint2 e = make_uint2(lg_base, (lg_bits & 0xFFFFu) | kEntryLeaf | kEntryTlas);
                            if (sp < D) lds_stack[sp * blockDim.x + threadIdx.x] = e;
                            else if (sp < D + GD) args.gstack[gtid * GD + (sp - D)] = e;
                            ++sp;
                        }
                        if (ng_bits & 255u) {
                            const uint2 e = make_uint2(ng_base, (ng_bits & 0xFFFFu) | kEntryTlas);
                            if (sp < D) lds_stack[sp * blockDim.x + threadIdx.x] = e;
                            else if (sp < D + GD) args.gstack[gtid * GD + (sp - D)] = e;
                            ++sp;
                        }
                        lg_bits = 0;
                        ng_bits = 0;
                        if (sp > D + GD) {
                            atomicOr(&args.state->error_flags, kErrStackOverflow);
                            done = true;
                        } else {
                            const InstanceInfo* info = table + iid;
                            const AsHeader* bh = (const AsHeader*)blas;
                            double M[12];
#pragma unroll
                            for (int kk = 0; kk < 12; ++kk) M[kk] = info->inv[kk];
                            if (kRayLds) {  // in the TLAS the current space is world space
                                wo[0] = s.ox;
                                wo[1] = s.oy;
                                wo[2] = s.oz;
                                wd[0] = s.dx;
                                wd[1] = s.dy;
                                wd[2] = s.dz;
                            }
                            double od[3], dd[3];
#pragma unroll
                            for (int r = 0; r < 3; ++r) {
                                od[r] = M[4 * r] * wo[0] + M[4 * r + 1] * wo[1] + M[4 * r + 2] * wo[2] + M[4 * r + 3];
                                dd[r] = M[4 * r] * wd[0] + M[4 * r + 1] * wd[1] + M[4 * r + 2] * wd[2];
                            }
                            const float4 blo = *(const float4*)bh->lo;
                            const float4 bhi = *(const float4*)bh->hi;
                            const float4 bqs = *(const float4*)bh->qscale;
                            // entry distance of the padded BLAS box, in FP32: it only chooses how far the origin
                            // moves (hits between tmin and te stay in the window, whose lower end becomes
                            // tmin - te), so its rounding costs nothing but keeps the FP64 divisions off this path
                            const float lo3[3] = {blo.x, blo.y, blo.z}, hi3[3] = {bhi.x, bhi.y, bhi.z};
                            float tn = -INFINITY;
#pragma unroll
                            for (int a = 0; a < 3; ++a) {
                                const float pad = (hi3[a] - lo3[a]) * (1.0f / 64.0f) + 1e-30f;
                                const float da = (float)dd[a], oa = (float)od[a];
                                if (fabsf(da) > 1e-30f) {
                                    const float r = __builtin_amdgcn_rcpf(da);
                                    const float t1 = (lo3[a] - pad - oa) * r, t2 = (hi3[a] + pad - oa) * r;
                                    tn = fmaxf(tn, fminf(t1, t2));
                                }
                            }
                            const double te = fmax((double)wtmin, (double)tn);
                            s.ox = (float)(od[0] + te * dd[0]);
                            s.oy = (float)(od[1] + te * dd[1]);
                            s.oz = (float)(od[2] + te * dd[2]);
                            s.dx = (float)dd[0];
                            s.dy = (float)dd[1];
                            s.dz = (float)dd[2];
                            s.toff = (float)te;
                            setup_space(s);
                            setup_shear(s, sh);
                            cslots = (const uint8_t*)blas + kHeaderBytes;
                            crlo[0] = blo.x;
                            crlo[1] = blo.y;
                            crlo[2] = blo.z;
                            cqs[0] = bqs.x;
                            cqs[1] = bqs.y;
                            cqs[2] = bqs.z;
                            in_blas = true;
                            cur_inst = FLAGS ? iid | (((imask >> 10) & 3u) << 30) : iid;  // InstSlot mask | flags << 8
                            cur_nslots = bh->slot_count;
                            ng_base = 0;
                            ng_bits = cur_nslots ? ((1u << s.oct) | (1u << 8)) : 0u;
                        }
                    }
                } else {
                    // triangles of the leaf group
                    const uint32_t vm = lg_bits >> 8;
                    uint32_t m = lg_bits & 255u;
                    lg_bits = 0;
                    const float smin = wtmin - s.toff;
                    uint32_t budget = kTrisPerStep;
                    while (m) {
                        if (budget-- == 0u) {  // the rest waits for a later leaf step
                            lg_bits = m | (vm << 8);
                            break;
                        }
                        const uint32_t k = (uint32_t)(__ffs(m) - 1);
                        m &= m - 1u;
                        const uint32_t slot = lg_base + (uint32_t)__popc(vm & ((1u << k) - 1u));
                        if (slot >= cur_nslots) {
                            atomicOr(&args.state->error_flags, kErrBadAs);
                            break;
                        }
                        const uint4* tp = (const uint4*)(cslots + (uint64_t)slot * kSlotBytes);
                        const uint4 a = tp[0], b = tp[1], c = tp[2];
                        if (STATS) ++st_tris;
                        float th, bu, bv;
                        const float smax = tc - s.toff;
                        bool cand;
                        if (FLAGS && (c.w & kTriAabb)) {
                            // AABB primitive: the intersection program of its hit record
                            const float lo[3] = {__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z)};
                            const float hi[3] = {__uint_as_float(a.w), __uint_as_float(b.x), __uint_as_float(b.y)};
                            const float ro[3] = {s.ox, s.oy, s.oz}, rd[3] = {s.dx, s.dy, s.dz};
                            float pn[3];
                            cand = !(ray_flags & kRaySkipAabbs) &&
                                   procedural_hit<float>(
                                       intersection_program(args, table + (cur_inst & kInstIdMask), c.z), lo, hi, ro,
                                       rd, smin, smax, th, bu, bv, pn);
                        } else {
                            cand = !(ray_flags & kRaySkipTriangles) &&
                                   test_triangle(a, b, c, s, sh, smin, smax, th, bu, bv) &&
                                   !(ray_flags & (kRayCullBack | kRayCullFront) &&
                                     facing_culled(a, b, c, s, ray_flags, table[cur_inst & kInstIdMask].flags));
                        }
                        if (cand &&
                            (!FLAGS || candidate_accepted(args, ray_flags, c.w, c.z, cur_inst, table, bu, bv))) {
                            if (MODE == kModeAny || ((MODE == kModeFused || kTrain) && shadow_ray)) {
                                occl = true;
                                done = true;
                                break;
                            }
                            // exact ties in t go to the smaller (instance, slot), so the closest hit does not
                            // depend on the order in which a lane meets its candidates (which the wave's leaf
                            // scheduling can change)
                            const float tn = s.toff + th;
                            const uint32_t ci = cur_inst & kInstIdMask;
                            if (!(tn == tc && (ci > hit_inst || (ci == hit_inst && slot > hit_slot)))) {
                                tc = tn;
                                hit_inst = ci;
                                hit_slot = slot;
                            }
                            if (ray_flags & kRayTerminateOnFirstHit) {
                                done = true;
                                break;
                            }
                        }
                    }
                }
            } else if (ng_bits & 255u) {
                const uint32_t om = ng_bits & 255u;
                const uint32_t j = (uint32_t)(__ffs(om) - 1);
                const uint32_t k = j ^ s.oct;
                const uint32_t vm = (ng_bits >> 8) & 255u;
                const uint32_t slot = ng_base + (uint32_t)__popc(vm & ((1u << k) - 1u));
                ng_bits &= ~(1u << j);
                if (ng_bits & 255u) {
                    const uint2 e = make_uint2(ng_base, (ng_bits & 0xFFFFu) | (in_blas ? 0u : kEntryTlas));
                    if (sp < D) lds_stack[sp * blockDim.x + threadIdx.x] = e;
                    else if (sp < D + GD) args.gstack[gtid * GD + (sp - D)] = e;
                    ++sp;
                    if (sp > D + GD) {
                        atomicOr(&args.state->error_flags, kErrStackOverflow);
                        done = true;
                    }
                }
                if (!done && slot >= cur_nslots) {
                    atomicOr(&args.state->error_flags, kErrBadAs);
                    ng_bits = 0;
                    lg_bits = 0;
                    done = sp == 0;
                } else if (!done) {
                    const uint4* np = (const uint4*)(cslots + (uint64_t)slot * kSlotBytes);
                    const uint4 w0 = np[0], w1 = np[1], w2 = np[2], w3 = np[3];
                    if (STATS) ++st_nodes;
                    const float smin = wtmin - s.toff;
                    const float smax = tc - s.toff;
                    const uint32_t hm = test_node(w0, w1, w2, w3, s, crlo, cqs, smin, smax);
                    const uint32_t imask = (w0.z >> 8) & 255u;
                    const uint32_t vmask = (w0.z >> 16) & 255u;
                    ng_base = w0.w;
                    ng_bits = oct_permute(hm & imask, s.oct) | (vmask << 8);
                    const uint32_t lg_new = (hm & ~imask) | (vmask << 8);
                    if (!(lg_bits & 255u)) {
                        lg_base = w0.w;
                        lg_bits = lg_new;
                    } else if (lg_new & 255u) {
                        // a leaf group is already pending: stack the new one
                        const uint2 e =
                            make_uint2(w0.w, (lg_new & 0xFFFFu) | kEntryLeaf | (in_blas ? 0u : kEntryTlas));
                        if (sp < D) lds_stack[sp * blockDim.x + threadIdx.x] = e;
                        else if (sp < D + GD) args.gstack[gtid * GD + (sp - D)] = e;
                        ++sp;
                        if (sp > D + GD) {
                            atomicOr(&args.state->error_flags, kErrStackOverflow);
                            done = true;
                        }
                    }
                }
            }

            if (done) {
                // the result is written at the next refill, together with the wave's other finished lanes
                finished = true;
                active = false;
                break;
            }
            if (++iter >= (int)args.refill_interval) {
                iter = 0;
                if (kTrain ? (uint32_t)__popcll(__ballot(true)) <= train_floor
                           : (!wave_exhausted && __popcll(__ballot(true)) < args.refill_lanes))
                    break;
            }
        }
    }

    if (kTrain && pt_rays) atomicAdd(&args.state->pt_rays, (unsigned long long)pt_rays);
    if (STATS) {
        atomicAdd(&args.state->stat_nodes, (unsigned long long)st_nodes);
        atomicAdd(&args.state->stat_tris, (unsigned long long)st_tris);
        atomicAdd(&args.state->stat_insts, (unsigned long long)st_insts);
        atomicAdd(&args.state->stat_rays, (unsigned long long)st_rays);
    }
}

// Three occupancy variants of the same traversal: unconstrained registers, at most 128 VGPRs (four waves per
// SIMD) and at most 96 VGPRs (five waves per SIMD), chosen by the waves per SIMD that the launch shape asks for.
template <int MODE, bool STATS, bool FLAGS>
__global__ __launch_bounds__(256) void trace_kernel(TraceArgs args) {
    trace_body<MODE, STATS, FLAGS>(args);
}
template <int MODE, bool STATS, bool FLAGS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void trace_kernel_w4(TraceArgs args) {
    trace_body<MODE, STATS, FLAGS>(args);
}

template <int MODE, bool STATS, bool FLAGS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void trace_kernel_w5(TraceArgs args) {
    trace_body<MODE, STATS, FLAGS>(args);
}

__global__ void reset_counter_kernel(DeviceState* st, int clear_stats) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        st->ray_counter = 0;
        if (clear_stats) {
            st->stat_nodes = 0;
            st->stat_tris = 0;
            st->stat_insts = 0;
            st->stat_rays = 0;
        }
    }
}

template <int MODE, bool STATS, bool FLAGS>
hipError_t launch_variant(const TraceArgs& a, const TraceLaunch& L, hipStream_t s) {
    // the stack, and the lanes' world rays where the kernel keeps them in LDS
    constexpr bool kRayLds = FLAGS || MODE == kModePath || MODE == kModeFused || MODE == kModeTrain ||
                             MODE == kModePixelPaths || MODE == kModePixelCache;
    const size_t lds =
        (size_t)L.lds_depth * L.block_size * sizeof(uint2) + (kRayLds ? (size_t)L.block_size * 6 * sizeof(float) : 0);
    if (L.waves >= 5)
        trace_kernel_w5<MODE, STATS, FLAGS><<<L.grid, L.block_size, lds, s>>>(a);
    else if (L.waves == 4)
        trace_kernel_w4<MODE, STATS, FLAGS><<<L.grid, L.block_size, lds, s>>>(a);
    else
        trace_kernel<MODE, STATS, FLAGS><<<L.grid, L.block_size, lds, s>>>(a);
    return hipGetLastError();
}

// ray flags that need the FLAGS variant
constexpr uint32_t kActiveRayFlags = kRayTerminateOnFirstHit | kRaySkipClosestHit | kRayCullBack | kRayCullFront |
                                     kRayCullOpaque | kRayCullNoOpaque | kRayNoOpaque | kRaySkipTriangles |
                                     kRaySkipAabbs;

template <int MODE, bool STATS>
hipError_t launch(const TraceArgs& a, const TraceLaunch& L, uint32_t content, hipStream_t s) {
    if (trace_needs_flags(a.ray_flags, content)) return launch_variant<MODE, STATS, true>(a, L, s);
    return launch_variant<MODE, STATS, false>(a, L, s);
}

}  // namespace

// The FLAGS variant resolves opacity and runs any-hit programs (needed for non-opaque content unless the ray is
// OPAQUE), runs intersection programs (AABB content) and applies the ray-flag culls; the default variant treats
// every primitive as an opaque triangle.
static_assert(kFlagsVariantRayFlags == kActiveRayFlags, "one set of ray flags selects the FLAGS variant");
bool trace_needs_flags(uint32_t ray_flags, uint32_t content) { return needs_flags_variant(ray_flags, content); }

namespace {
// A guarded launch that the device does not choose takes no ray: its rays come from the ray counter, which is set
// past every ray id (no wrap-around: fetches add at most a few chunks per wave).
constexpr uint32_t kGuardSkip = 0xC0000000u;
__global__ void guard_counter_kernel(const void* tlas, uint32_t ray_flags, uint32_t guard, DeviceState* st) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && !guard_passes(tlas, ray_flags, guard)) st->ray_counter = kGuardSkip;
}
}  // namespace

hipError_t trace_rays_launch(const TraceRequest& r, const TraceLaunch& L, hipStream_t s) {
    TraceArgs a{};
    a.tlas = (const uint8_t*)r.tlas;
    a.rays = r.rays;
    a.n_ptr = r.n_ptr;
    a.rays2 = r.rays2;
    a.n2_ptr = r.n2_ptr;
    a.path_hits = r.path_hits;
    a.n = r.n;
    a.cull_mask = r.cull_mask;
    a.hits = r.hits;
    a.occluded = r.occluded;
    for (int k = 0; k < 3; ++k) {
        a.cam[k] = r.cam[k];
        a.miss[k] = r.miss[k];
    }
    a.ax = r.ax;
    a.ay = r.ay;
    a.width = r.width;
    a.height = r.height;
    a.tmin = r.tmin;
    a.tmax = r.tmax;
    a.records = r.records;
    a.nrec = r.nrec;
    a.sbt_offset = r.sbt_offset;
    a.sbt_stride = r.sbt_stride;
    a.film = r.film;
    a.state = r.state;
    a.gstack = r.gstack;
    a.lds_depth = L.lds_depth;
    a.gdepth = L.gdepth;
    a.ray_chunk = L.ray_chunk;
    a.refill_lanes = L.refill_lanes;
    a.refill_interval = L.refill_interval;
    a.leaf_lanes = L.leaf_lanes;
    a.ray_flags = r.ray_flags;
    if (r.mode >= 5 && r.mode <= 7) {
        // whole paths: opaque triangle geometry and no ray flags only (the default traversal variant)
        if (!r.pt || trace_needs_flags(0u, r.content)) return hipErrorInvalidValue;
        a.pt = *r.pt;
        a.pt_first = r.mode == 5 ? r.pt->npix : 0u;
        a.pt_sample = r.pt_sample;
    }
    if (r.n == 0 && !r.n_ptr) {
        reset_counter_kernel<<<1, 64, 0, s>>>(r.state, r.stats ? 1 : 0);
        return hipGetLastError();
    }
    if (r.guard == 3u && r.mode <= 2 && !(r.ray_flags & kActiveRayFlags)) {
        // both variants, each guarded by the TLAS content the device reads: one of them takes every ray
        hipError_t e;
        for (uint32_t g = 1; g <= 2; ++g) {
            reset_counter_kernel<<<1, 64, 0, s>>>(r.state, r.stats && g == 1 ? 1 : 0);
            guard_counter_kernel<<<1, 64, 0, s>>>(r.tlas, r.ray_flags, g, r.state);
            const uint32_t c = g == 1 ? 0u : (kContentAabb | kContentNonOpaque);
            switch (r.mode) {
                case 0:
                    e = r.stats ? launch<kModeClosest, true>(a, L, c, s) : launch<kModeClosest, false>(a, L, c, s);
                    break;
                case 1: e = r.stats ? launch<kModeAny, true>(a, L, c, s) : launch<kModeAny, false>(a, L, c, s); break;
                default:
                    e = r.stats ? launch<kModeRaygen, true>(a, L, c, s) : launch<kModeRaygen, false>(a, L, c, s);
            }
            if (e) return e;
        }
        return hipSuccess;
    }
    reset_counter_kernel<<<1, 64, 0, s>>>(r.state, r.stats ? 1 : 0);
    if (r.guard == 1u || r.guard == 2u) guard_counter_kernel<<<1, 64, 0, s>>>(r.tlas, r.ray_flags, r.guard, r.state);
    switch (r.mode) {
        case 0:
            return r.stats ? launch<kModeClosest, true>(a, L, r.content, s) : launch<kModeClosest, false>(a, L, r.content, s);
        case 1: return r.stats ? launch<kModeAny, true>(a, L, r.content, s) : launch<kModeAny, false>(a, L, r.content, s);
        case 3: return launch<kModePath, false>(a, L, r.content, s);
        case 4: return launch<kModeFused, false>(a, L, r.content, s);
        case 5: return launch_variant<kModeTrain, false, false>(a, L, s);
        case 6: return launch_variant<kModePixelPaths, false, false>(a, L, s);
        case 7: return launch_variant<kModePixelCache, false, false>(a, L, s);
        default:
            return r.stats ? launch<kModeRaygen, true>(a, L, r.content, s) : launch<kModeRaygen, false>(a, L, r.content, s);
    }
}

}  // namespace rtx
