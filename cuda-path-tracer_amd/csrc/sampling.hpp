// Sampling for the progressive bounce loop.  The random numbers are a
// counter-based hash of (pixel, pass, bounce, seed), so a path's numbers do
// not depend on the order in which wavefront queues visit it.
#pragma once

#include "devmath.hpp"

namespace pt {

struct U4 {
    uint32_t x, y, z, w;
};

// pcg4d (Jarzynski & Olano, "Hash Functions for GPU Rendering", JCGT 9(3), 2020).
__device__ __forceinline__ U4 pcg4d(U4 v) {
    v.x = v.x * 1664525u + 1013904223u;
    v.y = v.y * 1664525u + 1013904223u;
    v.z = v.z * 1664525u + 1013904223u;
    v.w = v.w * 1664525u + 1013904223u;
    v.x += v.y * v.w;
    v.y += v.z * v.x;
    v.z += v.x * v.y;
    v.w += v.y * v.z;
    v.x ^= v.x >> 16;
    v.y ^= v.y >> 16;
    v.z ^= v.z >> 16;
    v.w ^= v.w >> 16;
    v.x += v.y * v.w;
    v.y += v.z * v.x;
    v.z += v.x * v.y;
    v.w += v.y * v.z;
    return v;
}

// Top 24 bits as a float in [0, 1).
__device__ __forceinline__ float u24(uint32_t x) {
    return static_cast<float>(x >> 8) * (1.0f / 16777216.0f);
}

// sin and cos of (pi/2) f for f in [0, 0.5]: Taylor series in x = f pi/2
// (|x| <= pi/4, truncation error below 2e-9) in Horner form.
__device__ __forceinline__ void sincos_quarter(float f, float& s, float& c) {
    const float x = f * 1.57079632679489662f;
    const float x2 = x * x;
    float ps = -1.0f / 39916800.0f;
    ps = ps * x2 + 1.0f / 362880.0f;
    ps = ps * x2 + -1.0f / 5040.0f;
    ps = ps * x2 + 1.0f / 120.0f;
    ps = ps * x2 + -1.0f / 6.0f;
    s = x + x * (x2 * ps);
    float pc = 1.0f / 479001600.0f;
    pc = pc * x2 + -1.0f / 3628800.0f;
    pc = pc * x2 + 1.0f / 40320.0f;
    pc = pc * x2 + -1.0f / 720.0f;
    pc = pc * x2 + 1.0f / 24.0f;
    pc = pc * x2 + -0.5f;
    c = 1.0f + x2 * pc;
}

// sin and cos of 2 pi u, u in [0, 1): exact quarter-turn split of 4u.
__device__ __forceinline__ void sincos2pi(float u, float& s, float& c) {
    const float q4 = u * 4.0f;
    const float qf = __builtin_floorf(q4);
    const int q = static_cast<int>(qf) & 3;
    const float f = q4 - qf;
    float a, b;
    if (f <= 0.5f) {
        sincos_quarter(f, a, b);
    } else {
        sincos_quarter(1.0f - f, b, a);
    }
    if (q == 0) { s = a; c = b; }
    else if (q == 1) { s = b; c = -a; }
    else if (q == 2) { s = -a; c = -b; }
    else { s = -b; c = a; }
}

// x^y for x >= 0, y > 0 (the Blinn-Phong specular power): 2^(y log2 x) with
// log2 of the mantissa m in [sqrt(1/2), sqrt(2)) from atanh((m-1)/(m+1))
// (odd series to s^11) and 2^f, |f| <= 1/2, from exp's Taylor series to
// degree 9; results below 2^-125 are 0.  The same float operations run in
// the CPU oracle, so both give the same bits.
__device__ __forceinline__ float pow_pos(float x, float y) {
    if (!(x > 0.0f)) return 0.0f;
    if (x > kFltMax) return x;
    int e = 0;
    if (x < 1.17549435e-38f) {
        x = x * 16777216.0f;  // exact: subnormal -> normal
        e = -24;
    }
    const uint32_t bits = __float_as_uint(x);
    e += static_cast<int>((bits >> 23) & 0xffu) - 127;
    float m = __uint_as_float((bits & 0x7fffffu) | 0x3f800000u);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e += 1;
    }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    float p = 1.0f / 11.0f;
    p = p * s2 + 1.0f / 9.0f;
    p = p * s2 + 1.0f / 7.0f;
    p = p * s2 + 1.0f / 5.0f;
    p = p * s2 + 1.0f / 3.0f;
    p = p * s2 + 1.0f;
    const float log2m = (2.88539008f * s) * p;  // 2 / ln 2
    const float t = y * static_cast<float>(e) + y * log2m;
    if (t < -125.0f) return 0.0f;
    if (t > 127.0f) return __uint_as_float(0x7f800000u);
    const float n = __builtin_rintf(t);
    const float f = (t - n) * 0.693147181f;  // ln 2
    float q = 1.0f / 362880.0f;
    q = q * f + 1.0f / 40320.0f;
    q = q * f + 1.0f / 5040.0f;
    q = q * f + 1.0f / 720.0f;
    q = q * f + 1.0f / 120.0f;
    q = q * f + 1.0f / 24.0f;
    q = q * f + 1.0f / 6.0f;
    q = q * f + 0.5f;
    q = q * f + 1.0f;
    q = q * f + 1.0f;
    return q * __uint_as_float(static_cast<uint32_t>(static_cast<int>(n) + 127) << 23);
}

// toHemisphere (reference math_util.h:56-67).
__device__ __forceinline__ V3 to_hemisphere(V3 v, V3 n) {
    V3 x = (__builtin_fabsf(n.x) < kFltEps) ? cross(n, v3(1.0f, 0.0f, 0.0f))
                                            : cross(n, v3(0.0f, 1.0f, 0.0f));
    x = normalize(x);
    const V3 y = normalize(cross(n, x));
    return add(add(scale(x, v.x), scale(y, v.y)), scale(n, v.z));
}

}  // namespace pt
