// Device float helpers with the reference's glm semantics.  Every function
// here is compiled with -ffp-contract=off and correctly rounded divide/sqrt,
// so each operation rounds once, in the written order -- the same values the
// oracle computes on the host.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pt {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;
constexpr float kFltMax = 3.402823466e+38f;
constexpr float kFltEps = 1.192092896e-07f;
// Largest float below FLT_MAX: "t <= kMinTInit" is the set "t < FLT_MAX" the
// reference accepts with its FLT_MAX sentinel (trace.h:97,119,134).
constexpr float kMinTInit = 3.4028232635611926e+38f;

// glm::min / glm::max scalar (glm/detail/func_common.inl:15-28).
__device__ __forceinline__ float gmin(float x, float y) { return x < y ? x : y; }
__device__ __forceinline__ float gmax(float x, float y) { return x > y ? x : y; }

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 mul(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 scale(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 neg(V3 a) { return V3{-a.x, -a.y, -a.z}; }
// glm::cross (glm/detail/func_geometric.inl:74-85).
__device__ __forceinline__ V3 cross(V3 x, V3 y) {
    return V3{x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
}
// glm::dot for vec3: products, then a left-to-right sum.
__device__ __forceinline__ float dot(V3 a, V3 b) {
    float x = a.x * b.x;
    float y = a.y * b.y;
    float z = a.z * b.z;
    return (x + y) + z;
}
// glm::normalize = v * (1 / sqrt(dot(v, v))).
__device__ __forceinline__ V3 normalize(V3 v) {
    float s = 1.0f / __builtin_sqrtf(dot(v, v));
    return scale(v, s);
}
// glm::mix = x + a * (y - x).
__device__ __forceinline__ V3 mix(V3 x, V3 y, float a) { return add(x, scale(sub(y, x), a)); }

// Moller-Trumbore (reference intersection.h:28-59) on a triangle stored as
// p0 and the edges e1 = p1 - p0, e2 = p2 - p0 (the same float subtractions
// the reference performs inside the test).  Returns t, or +inf on a miss.
__device__ __forceinline__ float mt_test(V3 o, V3 d, V3 p0, V3 e1, V3 e2, float& u, float& v) {
    V3 q = cross(d, e2);
    float det = dot(e1, q);
    if (det > -kFltEps && det < kFltEps) return __builtin_huge_valf();
    float inv = 1.0f / det;
    V3 s = sub(o, p0);
    u = inv * dot(s, q);
    if (u < 0.0f) return __builtin_huge_valf();
    V3 r = cross(s, e1);
    v = inv * dot(d, r);
    if (v < 0.0f || (u + v) > 1.0f) return __builtin_huge_valf();
    float t = inv * dot(e2, r);
    return t >= 0.0f ? t : __builtin_huge_valf();
}

// Slab test (reference intersection.h:106-117) with dInv = 1/d per ray.
// Box given as (minx, maxx), (miny, maxy), (minz, maxz).  Returns t_near, or
// +inf on a miss.
__device__ __forceinline__ float slab_test(V3 o, V3 dinv, float minx, float maxx, float miny,
                                           float maxy, float minz, float maxz) {
    float tx0 = (minx - o.x) * dinv.x, tx1 = (maxx - o.x) * dinv.x;
    float ty0 = (miny - o.y) * dinv.y, ty1 = (maxy - o.y) * dinv.y;
    float tz0 = (minz - o.z) * dinv.z, tz1 = (maxz - o.z) * dinv.z;
    float n = gmax(gmax(gmin(tx0, tx1), gmin(ty0, ty1)), gmin(tz0, tz1));
    float f = gmin(gmin(gmax(tx0, tx1), gmax(ty0, ty1)), gmax(tz0, tz1));
    return (n <= f && f > 0.0f) ? n : __builtin_huge_valf();
}

// The same slab test for rays whose dInv has no infinite component: then no
// NaN can arise, and hardware min/max give the values glm's ternaries give
// (they can differ only in the sign of a zero, which no later comparison sees).
__device__ __forceinline__ float slab_test_finite(V3 o, V3 dinv, float minx, float maxx,
                                                  float miny, float maxy, float minz,
                                                  float maxz) {
    float tx0 = (minx - o.x) * dinv.x, tx1 = (maxx - o.x) * dinv.x;
    float ty0 = (miny - o.y) * dinv.y, ty1 = (maxy - o.y) * dinv.y;
    float tz0 = (minz - o.z) * dinv.z, tz1 = (maxz - o.z) * dinv.z;
    float n = __builtin_fmaxf(__builtin_fmaxf(__builtin_fminf(tx0, tx1), __builtin_fminf(ty0, ty1)),
                              __builtin_fminf(tz0, tz1));
    float f = __builtin_fminf(__builtin_fminf(__builtin_fmaxf(tx0, tx1), __builtin_fmaxf(ty0, ty1)),
                              __builtin_fmaxf(tz0, tz1));
    return (n <= f && f > 0.0f) ? n : __builtin_huge_valf();
}

}  // namespace pt
