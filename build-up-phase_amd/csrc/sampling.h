// sampling.h — counter-based random numbers and Lambertian sampling shared by the integrators.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rtx {

// pcg4d: Jarzynski & Olano 2020, "Hash Functions for GPU Rendering" (JCGT 9(3)).
__host__ __device__ inline uint4 pcg4d(uint32_t x, uint32_t y, uint32_t z, uint32_t w) {
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
    return make_uint4(x, y, z, w);
}

// 24-bit uniform in [0, 1)
__host__ __device__ inline float unit_float(uint32_t v) { return (float)(v >> 8) * (1.0f / 16777216.0f); }

constexpr uint32_t kJitterKey = 0xA511E9B3u;

// Orthonormal basis around a unit vector n (Duff et al. 2017, "Building an Orthonormal Basis, Revisited").
__device__ inline void onb(const float* n, float* t, float* b) {
    const float sign = copysignf(1.0f, n[2]);
    const float a = -1.0f / (sign + n[2]);
    const float bb = n[0] * n[1] * a;
    t[0] = 1.0f + sign * n[0] * n[0] * a;
    t[1] = sign * bb;
    t[2] = -sign * n[0];
    b[0] = bb;
    b[1] = sign + n[1] * n[1] * a;
    b[2] = -n[1];
}

// Cosine-weighted direction around n: r = sqrt(u1), phi = 2 pi u2, local (r cos phi, r sin phi, sqrt(1 - u1)).
__device__ inline void cosine_dir(const float* n, float u1, float u2, float* d) {
    float t[3], b[3];
    onb(n, t, b);
    const float r = sqrtf(u1);
    const float phi = 6.28318530717958647692f * u2;
    const float lx = r * cosf(phi), ly = r * sinf(phi), lz = sqrtf(fmaxf(0.0f, 1.0f - u1));
#pragma unroll
    for (int a = 0; a < 3; ++a) d[a] = lx * t[a] + ly * b[a] + lz * n[a];
}

}  // namespace rtx
