// CPU oracle: restatement of the reference's LBVH build, closest-hit traversal
// and progressive bounce loop.  Test infrastructure only (see oracle.h).
// Built with -O2 -ffp-contract=off and SSE2 float arithmetic so that every
// float operation rounds exactly once, in the order written here.
#include "oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

namespace {

// glm scalar min/max (third_party/glm/glm/detail/func_common.inl:15-28):
// NaN in either operand yields the second operand.
inline float gmin(float x, float y) { return x < y ? x : y; }
inline float gmax(float x, float y) { return x > y ? x : y; }

struct V3 {
    float x, y, z;
};
inline V3 vld(const float* p) { return {p[0], p[1], p[2]}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 scale(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 neg(V3 a) { return {-a.x, -a.y, -a.z}; }
// glm::cross (func_geometric.inl:74-85).
inline V3 cross(V3 x, V3 y) {
    return {x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
}
// glm::dot for vec3 (func_geometric.inl:54-61): products, then left-to-right sum.
inline float dot(V3 a, V3 b) {
    V3 m = mul(a, b);
    return (m.x + m.y) + m.z;
}
// glm::normalize = v * inversesqrt(dot(v, v)), inversesqrt = 1 / sqrt
// (func_geometric.inl:88-96, func_exponential.inl:130-133).
inline V3 normalize(V3 v) {
    float s = 1.0f / std::sqrt(dot(v, v));
    return scale(v, s);
}
// glm::mix(x, y, a) = x + a * (y - x) (func_common.inl:143-151).
inline V3 mix(V3 x, V3 y, float a) { return add(x, scale(sub(y, x), a)); }

constexpr uint32_t kInvalid = 0xFFFFFFFFu;
// "No hit yet" bound: the reference starts at FLT_MAX with strict tests
// (trace.h:97,119,134); t <= 0x7f7ffffe is the same set as t < FLT_MAX.
const float kMinTInit = std::nextafter(FLT_MAX, 0.0f);
const float kInf = INFINITY;

// Cast like CUDA's cvt.rzi.s32.f32: truncation, saturating, NaN -> 0.
inline int32_t f2i_sat(float x) {
    if (!(x == x)) return 0;
    if (x >= 2147483648.0f) return INT32_MAX;
    if (x <= -2147483648.0f) return INT32_MIN;
    return static_cast<int32_t>(x);
}

inline int clz32(uint32_t x) { return x == 0 ? 32 : __builtin_clz(x); }

struct Tri {
    V3 v0, v1, v2;
    uint32_t id;
};

inline Tri load_tri(const pt_triangle& t) { return {vld(t.v0), vld(t.v1), vld(t.v2), t.id}; }

// Moller-Trumbore (intersection.h:28-59).  Returns true on hit.
inline bool mt(V3 o, V3 d, V3 p0, V3 p1, V3 p2, float& u, float& v, float& t) {
    V3 e1 = sub(p1, p0);
    V3 e2 = sub(p2, p0);
    V3 q = cross(d, e2);
    float det = dot(e1, q);
    if (det > -FLT_EPSILON && det < FLT_EPSILON) return false;
    float inv = 1.0f / det;
    V3 s = sub(o, p0);
    u = inv * dot(s, q);
    if (u < 0.0f) return false;
    V3 r = cross(s, e1);
    v = inv * dot(d, r);
    if (v < 0.0f || (u + v) > 1.0f) return false;
    t = inv * dot(e2, r);
    return t >= 0.0f;
}

// Slab test (intersection.h:106-117) with dInv = 1 / d precomputed per ray
// (the same values: the division does not depend on the box).
inline bool slab(V3 o, V3 dinv, const float* b, float& tnear) {
    V3 tmin = mul(sub(vld(b), o), dinv);
    V3 tmax = mul(sub(vld(b + 3), o), dinv);
    V3 t1{gmin(tmin.x, tmax.x), gmin(tmin.y, tmax.y), gmin(tmin.z, tmax.z)};
    V3 t2{gmax(tmin.x, tmax.x), gmax(tmin.y, tmax.y), gmax(tmin.z, tmax.z)};
    tnear = gmax(gmax(t1.x, t1.y), t1.z);
    float tfar = gmin(gmin(t2.x, t2.y), t2.z);
    return tnear <= tfar && tfar > 0.0f;
}

struct Hit {
    float t = FLT_MAX;
    uint32_t sorted = kInvalid;
    uint32_t id = kInvalid;
    float u = 0.0f, v = 0.0f;
};

struct Counters {
    uint64_t iters = 0, boxes = 0, tris = 0;
    uint64_t depth = 0;  // deepest deferred-node stack
};

struct Bvh {
    const pt_triangle* tris;
    const uint32_t* perm;
    const uint32_t* children;
    const float* boxes;
    uint32_t n;
};

// Closest hit, reference traversal order (trace.h:73-169).  Differences, all
// for the lowest-id tie rule: culling keeps children with t == best
// (trace.h:119 drops them), the far child is kept when t == best (:134), and
// a triangle updates the best hit when it is nearer or equally near with a
// lower original id (:143,:152 update only when nearer).
Hit trace_one(const Bvh& bvh, V3 o, V3 d, Counters* cnt) {
    Hit best;
    float minT = kMinTInit;
    const uint32_t n = bvh.n;
    if (n == 0) return best;
    auto tri_test = [&](uint32_t c, float& h, float& hu, float& hv) {
        Tri tr = load_tri(bvh.tris[bvh.perm[c]]);
        float u, v, t;
        if (cnt) cnt->tris++;
        if (mt(o, d, tr.v0, tr.v1, tr.v2, u, v, t)) {
            h = t;
            hu = u;
            hv = v;
        } else {
            h = kInf;
        }
    };
    auto update = [&](uint32_t c, float h, float hu, float hv) {
        uint32_t id = bvh.tris[bvh.perm[c]].id;
        if (h < minT || (h == minT && id < best.id)) {
            minT = h;
            best.t = h;
            best.sorted = c;
            best.id = id;
            best.u = hu;
            best.v = hv;
        }
    };
    if (n == 1) {
        float h, hu, hv;
        tri_test(0, h, hu, hv);
        if (h <= minT) update(0, h, hu, hv);
        return best;
    }
    V3 dinv{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;  // root = internal node 0 (trace.h:94)
    while (sp > 0) {
        uint32_t node = stack[--sp];
        for (;;) {
            if (cnt) cnt->iters++;
            uint32_t w = bvh.children[node];
            uint32_t ci[2] = {w >> 2, (w >> 2) + 1};
            bool internal[2] = {(w & 1u) != 0, (w & 2u) != 0};
            float h[2], hu[2] = {0, 0}, hv[2] = {0, 0};
            for (int k = 0; k < 2; ++k) {
                if (internal[k]) {
                    float tn;
                    if (cnt) cnt->boxes++;
                    h[k] = slab(o, dinv, bvh.boxes + 6 * static_cast<size_t>(ci[k]), tn) ? tn : kInf;
                } else {
                    tri_test(ci[k], h[k], hu[k], hv[k]);
                }
            }
            if (!(h[0] <= minT) && !(h[1] <= minT)) break;
            int cl = (h[1] < h[0]) ? 1 : 0;
            int fa = 1 - cl;
            if (internal[cl]) {
                node = ci[cl];
                if (h[fa] <= minT) {
                    if (internal[fa]) {
                        // The depth bound (< 64 for < 2^30 leaves) keeps sp < 128.
                        if (sp >= 128) std::abort();
                        stack[sp++] = ci[fa];
                        if (cnt && static_cast<uint64_t>(sp) > cnt->depth) cnt->depth = static_cast<uint64_t>(sp);
                    } else {
                        update(ci[fa], h[fa], hu[fa], hv[fa]);
                    }
                }
            } else {
                update(ci[cl], h[cl], hu[cl], hv[cl]);
                // Far child at exactly the same distance: still a candidate.
                if (h[fa] <= minT) {
                    if (internal[fa]) {
                        node = ci[fa];
                        continue;
                    }
                    update(ci[fa], h[fa], hu[fa], hv[fa]);
                }
                break;
            }
        }
    }
    return best;
}

// ---- sampling (the build's own, identical in the HIP renderer) ------------

// pcg4d hash (Jarzynski & Olano, "Hash Functions for GPU Rendering", JCGT 2020).
inline void pcg4d(uint32_t v[4]) {
    for (int k = 0; k < 4; ++k) v[k] = v[k] * 1664525u + 1013904223u;
    v[0] += v[1] * v[3];
    v[1] += v[2] * v[0];
    v[2] += v[0] * v[1];
    v[3] += v[1] * v[2];
    for (int k = 0; k < 4; ++k) v[k] ^= v[k] >> 16;
    v[0] += v[1] * v[3];
    v[1] += v[2] * v[0];
    v[2] += v[0] * v[1];
    v[3] += v[1] * v[2];
}
inline float u24(uint32_t x) { return static_cast<float>(x >> 8) * (1.0f / 16777216.0f); }

// sin and cos of (pi/2) * f for f in [0, 0.5]: Taylor series in x = f*pi/2
// (|x| <= pi/4; truncation error < 2e-9), evaluated in Horner form.
inline void sincos_quarter(float f, float& s, float& c) {
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
// sin and cos of 2*pi*u for u in [0, 1): exact quadrant split of 4u, then the
// quarter-turn series on the part nearer to a multiple of pi/2.
inline void sincos2pi(float u, float& s, float& c) {
    float q4 = u * 4.0f;
    float qf = std::floor(q4);
    int q = static_cast<int>(qf) & 3;
    float f = q4 - qf;  // exact
    float a, b;        // a = sin(pi/2 f), b = cos(pi/2 f)
    if (f <= 0.5f) {
        sincos_quarter(f, a, b);
    } else {
        sincos_quarter(1.0f - f, b, a);  // sin(pi/2 f) = cos(pi/2 (1-f))
    }
    switch (q) {
        case 0: s = a; c = b; break;
        case 1: s = b; c = -a; break;
        case 2: s = -a; c = -b; break;
        default: s = -b; c = a; break;
    }
}

// toHemisphere (math_util.h:56-67).
inline V3 to_hemisphere(V3 v, V3 n) {
    V3 x = (std::fabs(n.x) < FLT_EPSILON) ? cross(n, V3{1.0f, 0.0f, 0.0f})
                                          : cross(n, V3{0.0f, 1.0f, 0.0f});
    x = normalize(x);
    V3 y = normalize(cross(n, x));
    return add(add(scale(x, v.x), scale(y, v.y)), scale(n, v.z));
}

struct ShadeInfo {
    V3 normal, albedo, emission;
};

// Surface detail of a scene (pt_triangle_detail per original id, textures);
// empty: pt_triangle_info alone.
struct Detail {
    const pt_triangle_detail* det = nullptr;
    const pt_texture* tex = nullptr;
    uint32_t ntex = 0;
};

// tex2D on the reference's texture objects (Scene.cu:66-73: normalised
// coordinates, wrap, bilinear, level 0; texels c / 255), with fp32 weights:
// x = s * width - 0.5, y = t * height - 0.5, glm::mix along x then y.
V3 tex_sample(const pt_texture& tx, float s, float t) {
    const uint32_t W = tx.width, H = tx.height;
    float x = s * static_cast<float>(W) - 0.5f;
    float y = t * static_cast<float>(H) - 0.5f;
    if (!(std::fabs(x) < 1e9f)) x = 0.0f;
    if (!(std::fabs(y) < 1e9f)) y = 0.0f;
    const float fx = std::floor(x), fy = std::floor(y);
    const float a = x - fx, b = y - fy;
    int64_t i0 = static_cast<int64_t>(fx) % static_cast<int64_t>(W);
    int64_t j0 = static_cast<int64_t>(fy) % static_cast<int64_t>(H);
    if (i0 < 0) i0 += W;
    if (j0 < 0) j0 += H;
    const int64_t i1 = (i0 + 1) % W, j1 = (j0 + 1) % H;
    auto texel = [&](int64_t i, int64_t j) {
        const uint8_t* c = tx.rgba + 4 * (static_cast<size_t>(j) * W + static_cast<size_t>(i));
        return V3{static_cast<float>(c[0]) / 255.0f, static_cast<float>(c[1]) / 255.0f,
                  static_cast<float>(c[2]) / 255.0f};
    };
    const V3 lo = mix(texel(i0, j0), texel(i1, j0), a);
    const V3 hi = mix(texel(i0, j1), texel(i1, j1), a);
    return mix(lo, hi, b);
}

// barycentricLerp (math_util.h:19-25): v0 * (1 - u - v) + v1 * u + v2 * v.
inline V3 bary3(const float (*v)[3], float w, float u, float t) {
    return add(add(scale(vld(v[0]), w), scale(vld(v[1]), u)), scale(vld(v[2]), t));
}

// extractInfo (cuda_path_tracer.cu:115-138): interpolated, normalised vertex
// normal (TriangleInfo.h:44-47 via barycentricLerp, math_util.h:19-25) and
// the albedo as given; with surface detail, the albedo texture at the
// interpolated uv or normalize(interpolated vertex colour) (computeColor,
// TriangleInfo.h:39-42), and the normal map through the interpolated tangent
// frame (unpackNormal, cuda_path_tracer.cu:85-104).
inline ShadeInfo extract(const pt_triangle_info& in, uint32_t id, const Detail& dt, float u, float v) {
    float w = (1.0f - u) - v;
    V3 nrm = normalize(add(add(scale(vld(in.n0), w), scale(vld(in.n1), u)), scale(vld(in.n2), v)));
    V3 albedo = vld(in.albedo);
    if (dt.det) {
        const pt_triangle_detail& d = dt.det[id];
        const float s = ((d.uv[0][0] * w) + (d.uv[1][0] * u)) + (d.uv[2][0] * v);
        const float t = ((d.uv[0][1] * w) + (d.uv[1][1] * u)) + (d.uv[2][1] * v);
        albedo = d.albedo_tex != PT_NO_TEXTURE ? tex_sample(dt.tex[d.albedo_tex], s, t)
                                               : normalize(bary3(d.color, w, u, v));
        if (d.normal_tex != PT_NO_TEXTURE) {
            const V3 ns = tex_sample(dt.tex[d.normal_tex], s, t);
            const float nx = 2.0f * ns.x - 1.0f, ny = 2.0f * ns.y - 1.0f, nz = 2.0f * ns.z - 1.0f;
            const V3 tg = normalize(bary3(d.tangent, w, u, v));
            const V3 bt = normalize(bary3(d.bitangent, w, u, v));
            nrm = normalize(add(add(scale(tg, nx), scale(bt, ny)), scale(nrm, nz)));
        }
    }
    return {nrm, albedo, vld(in.emission)};
}

inline bool emissive(V3 e) { return ((e.x + e.y) + e.z) > 1e-7f; }  // CUDA_EPSILON

inline float bits_to_float(uint32_t b) {
    float f;
    std::memcpy(&f, &b, sizeof f);
    return f;
}

// x^y for x >= 0, y > 0 (the Blinn-Phong specular power; the build's own
// function, not libm's pow): 2^(y log2 x), log2 of the mantissa m in
// [sqrt(1/2), sqrt(2)) from atanh((m-1)/(m+1)) (odd series to s^11), 2^f for
// |f| <= 1/2 from exp's Taylor series to degree 9; below 2^-125 the result is
// 0.  Same float operations as the kernels.
float pow_pos(float x, float y) {
    if (!(x > 0.0f)) return 0.0f;
    if (x > FLT_MAX) return x;
    int e = 0;
    if (x < 1.17549435e-38f) {
        x = x * 16777216.0f;
        e = -24;
    }
    uint32_t bits;
    std::memcpy(&bits, &x, sizeof bits);
    e += static_cast<int>((bits >> 23) & 0xffu) - 127;
    float m = bits_to_float((bits & 0x7fffffu) | 0x3f800000u);
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
    const float log2m = (2.88539008f * s) * p;
    const float t = y * static_cast<float>(e) + y * log2m;
    if (t < -125.0f) return 0.0f;
    if (t > 127.0f) return bits_to_float(0x7f800000u);
    const float n = std::rint(t);
    const float f = (t - n) * 0.693147181f;
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
    return q * bits_to_float(static_cast<uint32_t>(static_cast<int>(n) + 127) << 23);
}

// computeL (LightAlgorithms.h:12-20): shadow ray from pos - dir * 1e-3
// (CUDA_EPSILON3) towards the light, a closest-hit trace (trace.h:45-50);
// in shadow iff t < FLT_MAX.
V3 light_L(const Bvh& bvh, V3 pos, const pt_dir_light& lt) {
    const V3 ld = vld(lt.direction);
    const Hit h = trace_one(bvh, sub(pos, scale(ld, 1e-3f)), neg(ld), nullptr);
    if (h.t < FLT_MAX) return V3{0.0f, 0.0f, 0.0f};
    return scale(vld(lt.color), lt.intensity);
}

// blinnPhongBRDF (cuda_path_tracer.cu:36-50).
V3 blinn_phong(V3 Id, V3 kd, V3 Is, V3 ks, V3 n, V3 l, V3 hw, float shininess) {
    const float df = gmax(dot(n, l), 0.0f);
    const V3 diffuse = scale(mul(Id, kd), df);
    V3 specular{0.0f, 0.0f, 0.0f};
    if (shininess > 0.0f && df > 0.0f) specular = scale(mul(Is, ks), pow_pos(gmax(dot(n, hw), 0.0f), shininess));
    return add(diffuse, specular);
}

// Primary-hit shading of a non-emissive surface (cuda_path_tracer.cu:201-214):
// L sums the lights seen so far, each light adds Blinn-Phong with that L.
V3 primary_lights(const Bvh& bvh, V3 pos, V3 d, V3 nrm, V3 albedo, float spec, const pt_dir_light* lights,
                  uint32_t nlights) {
    V3 Lc{0.0f, 0.0f, 0.0f}, color{0.0f, 0.0f, 0.0f};
    for (uint32_t k = 0; k < nlights; ++k) {
        const V3 ld = vld(lights[k].direction);
        Lc = add(Lc, light_L(bvh, pos, lights[k]));
        const V3 hw = scale(sub(neg(d), ld), 0.5f);
        color = add(color, blinn_phong(Lc, albedo, Lc, V3{0.5f, 0.5f, 0.5f}, nrm, neg(ld), hw, spec));
    }
    return color;
}

// Bounce-hit light term of a non-emissive surface (cuda_path_tracer.cu:314-321).
V3 bounce_lights(const Bvh& bvh, V3 pos, V3 nrm, const pt_dir_light* lights, uint32_t nlights) {
    V3 L{0.0f, 0.0f, 0.0f};
    for (uint32_t k = 0; k < nlights; ++k) {
        const V3 ld = vld(lights[k].direction);
        L = add(L, scale(light_L(bvh, pos, lights[k]), gmax(dot(neg(ld), nrm), 0.0f)));
    }
    return L;
}

}  // namespace

extern "C" {

uint32_t or_encode30(uint32_t x, uint32_t y, uint32_t z) {
    // Spread the low 10 bits of each coordinate to every third bit
    // (morton.h:84-102): x -> bits 0,3,6..., y -> 1,4,..., z -> 2,5,...
    auto spread = [](uint32_t a) {
        a &= 0x3FFu;
        a = (a | (a << 16)) & 0x030000FFu;
        a = (a | (a << 8)) & 0x0300F00Fu;
        a = (a | (a << 4)) & 0x030C30C3u;
        a = (a | (a << 2)) & 0x09249249u;
        return a;
    };
    return spread(x) | (spread(y) << 1) | (spread(z) << 2);
}

uint32_t or_quantize(float x, uint32_t levels) {
    // quantize (math_util.h:10-13): clamp(int(x * n), 0, n - 1).
    int32_t q = f2i_sat(x * static_cast<float>(levels));
    int32_t hi = static_cast<int32_t>(levels - 1);
    q = q < hi ? q : hi;
    q = q > 0 ? q : 0;
    return static_cast<uint32_t>(q);
}

void or_morton(const pt_triangle* tris, uint32_t n, const float* bmin, const float* bmax,
               uint32_t* keys) {
    // morton_encoder (BVHBuilder.cu:12-32): invScale = 1 / (max - min) on the host.
    float inv[3], mn[3];
    for (int k = 0; k < 3; ++k) {
        mn[k] = bmin[k];
        inv[k] = 1.0f / (bmax[k] - bmin[k]);
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < static_cast<int64_t>(n); ++i) {
        const pt_triangle& t = tris[i];
        uint32_t q[3];
        for (int k = 0; k < 3; ++k) {
            float c = ((t.v0[k] + t.v1[k]) + t.v2[k]) / 3.0f;
            q[k] = or_quantize((c - mn[k]) * inv[k], 1024u);
        }
        keys[i] = or_encode30(q[0], q[1], q[2]);
    }
}

void or_sort(const uint32_t* keys, uint32_t n, uint32_t* sorted_keys, uint32_t* perm) {
    std::vector<uint32_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0u);
    std::stable_sort(idx.begin(), idx.end(),
                     [keys](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
    for (uint32_t k = 0; k < n; ++k) {
        perm[k] = idx[k];
        sorted_keys[k] = keys[idx[k]];
    }
}

void or_radix_tree(const uint32_t* keys, uint32_t n, uint32_t* children,
                   uint32_t* parents_leaf, uint32_t* parents_internal) {
    if (n < 2) {
        if (n == 1) parents_leaf[0] = kInvalid;
        return;
    }
    const int64_t N = n;
    // delta (BVHBuilder.cu:116-133): longest common prefix of the keys,
    // augmented with the index when keys are equal.
    auto delta = [&](int64_t i, int64_t j) -> int {
        if (j < 0 || j >= N) return -1;
        uint32_t ki = keys[i], kj = keys[j];
        if (ki == kj) return 32 + clz32(static_cast<uint32_t>(i ^ j));
        return clz32(ki ^ kj);
    };
    parents_internal[0] = kInvalid;  // BVHBuilder.cu:224
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N - 1; ++i) {
        // Direction, range and split (BVHBuilder.cu:135-166).
        int dr = delta(i, i + 1) - delta(i, i - 1);
        int64_t d = (dr > 0) - (dr < 0);
        int dmin = delta(i, i - d);
        int64_t lmax = 2;
        while (delta(i, i + lmax * d) > dmin) lmax *= 2;
        int64_t l = 0;
        while (lmax > 1) {
            lmax /= 2;
            if (delta(i, i + (l + lmax) * d) > dmin) l += lmax;
        }
        int64_t j = i + l * d;
        int dnode = delta(i, j);
        int64_t s = 0;
        while (l > 1) {
            l = (l + 1) / 2;
            if (delta(i, i + (s + l) * d) > dnode) s += l;
        }
        s = i + s * d + std::min<int64_t>(d, 0);
        // Children and parents (BVHBuilder.cu:168-195).
        uint32_t type = 0;
        if (std::min(i, j) != s) {
            type |= 1u;
            parents_internal[s] = static_cast<uint32_t>(i);
        } else {
            parents_leaf[s] = static_cast<uint32_t>(i);
        }
        if (std::max(i, j) != s + 1) {
            type |= 2u;
            parents_internal[s + 1] = static_cast<uint32_t>(i);
        } else {
            parents_leaf[s + 1] = static_cast<uint32_t>(i);
        }
        children[i] = (static_cast<uint32_t>(s) << 2) | type;
    }
}

void or_refit(const pt_triangle* tris, const uint32_t* perm, const uint32_t* children,
              const uint32_t* parents_leaf, const uint32_t* parents_internal, uint32_t n,
              float* boxes) {
    if (n < 2) return;
    // compute_internal_node_bbox_kernel (BVHBuilder.cu:54-91) run leaf by leaf:
    // the first visitor of a node stops, the second unites the child boxes.
    // min/max is exact, so the visiting order does not change the result.
    std::vector<uint8_t> visits(n - 1, 0);
    auto leaf_box = [&](uint32_t c, float* b) {
        // BBox() then unite({v0, v1, v2}) (BBox.cpp:4-6, 14-24).
        const pt_triangle& t = tris[perm[c]];
        float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        const float* vs[3] = {t.v0, t.v1, t.v2};
        for (int v = 0; v < 3; ++v)
            for (int k = 0; k < 3; ++k) {
                mn[k] = gmin(mn[k], vs[v][k]);
                mx[k] = gmax(mx[k], vs[v][k]);
            }
        for (int k = 0; k < 3; ++k) {
            b[k] = mn[k];
            b[3 + k] = mx[k];
        }
    };
    for (uint32_t leaf = 0; leaf < n; ++leaf) {
        uint32_t p = parents_leaf[leaf];
        while (p != kInvalid) {
            if (visits[p]++ == 0) break;
            uint32_t w = children[p];
            uint32_t l = w >> 2, r = l + 1;
            float lb[6], rb[6];
            if (w & 1u) std::memcpy(lb, boxes + 6 * static_cast<size_t>(l), sizeof lb);
            else leaf_box(l, lb);
            if (w & 2u) std::memcpy(rb, boxes + 6 * static_cast<size_t>(r), sizeof rb);
            else leaf_box(r, rb);
            float* out = boxes + 6 * static_cast<size_t>(p);
            for (int k = 0; k < 3; ++k) {  // leftBBox.unite(rightBBox) (BBox.cpp:30-36)
                out[k] = gmin(lb[k], rb[k]);
                out[3 + k] = gmax(lb[3 + k], rb[3 + k]);
            }
            p = parents_internal[p];
        }
    }
}

int or_build(const pt_triangle* tris, uint32_t n, const float* bmin, const float* bmax,
             uint32_t* sorted_keys, uint32_t* perm, uint32_t* children, uint32_t* parents_leaf,
             uint32_t* parents_internal, float* boxes) {
    std::vector<uint32_t> keys(n);
    or_morton(tris, n, bmin, bmax, keys.data());
    or_sort(keys.data(), n, sorted_keys, perm);
    or_radix_tree(sorted_keys, n, children, parents_leaf, parents_internal);
    or_refit(tris, perm, children, parents_leaf, parents_internal, n, boxes);
    return 0;
}

int or_ray_triangle(const float* o, const float* d, const float* p0, const float* p1,
                    const float* p2, float* u, float* v, float* t) {
    float uu = 0, vv = 0, tt = 0;
    bool hit = mt(vld(o), vld(d), vld(p0), vld(p1), vld(p2), uu, vv, tt);
    *u = uu;
    *v = vv;
    *t = tt;
    return hit ? 1 : 0;
}

int or_ray_aabb(const float* o, const float* d, const float* bmin, const float* bmax, float* t) {
    V3 dd = vld(d);
    V3 dinv{1.0f / dd.x, 1.0f / dd.y, 1.0f / dd.z};
    float b[6] = {bmin[0], bmin[1], bmin[2], bmax[0], bmax[1], bmax[2]};
    return slab(vld(o), dinv, b, *t) ? 1 : 0;
}

void or_trace(const pt_triangle* tris, const uint32_t* perm, const uint32_t* children,
              const float* boxes, uint32_t n, const float* rays, uint32_t nrays, uint32_t* id,
              float* t, float* uv, uint64_t* counters, int threads) {
    Bvh bvh{tris, perm, children, boxes, n};
    uint64_t it = 0, bx = 0, tr = 0, dp = 0;
    if (threads <= 0) threads = 1;
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads) reduction(+ : it, bx, tr) reduction(max : dp)
    for (int64_t i = 0; i < static_cast<int64_t>(nrays); ++i) {
        const float* r = rays + 6 * i;
        Counters c;
        Hit h = trace_one(bvh, vld(r), vld(r + 3), counters ? &c : nullptr);
        if (id) id[i] = h.id;
        if (t) t[i] = h.t;
        if (uv) {
            uv[2 * i] = h.u;
            uv[2 * i + 1] = h.v;
        }
        it += c.iters;
        bx += c.boxes;
        tr += c.tris;
        dp = std::max(dp, c.depth);
    }
    if (counters) {
        counters[0] = it;
        counters[1] = bx;
        counters[2] = tr;
        counters[3] = dp;
    }
}

void or_brute(const pt_triangle* tris, uint32_t n, const float* rays, uint32_t nrays,
              uint32_t* id, float* t, float* uv, int threads) {
    if (threads <= 0) threads = 1;
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
    for (int64_t i = 0; i < static_cast<int64_t>(nrays); ++i) {
        V3 o = vld(rays + 6 * i), d = vld(rays + 6 * i + 3);
        Hit best;
        float minT = kMinTInit;
        for (uint32_t k = 0; k < n; ++k) {
            Tri tr = load_tri(tris[k]);
            float u, v, tt;
            if (!mt(o, d, tr.v0, tr.v1, tr.v2, u, v, tt)) continue;
            if (tt < minT || (tt == minT && tr.id < best.id)) {
                minT = tt;
                best.t = tt;
                best.id = tr.id;
                best.u = u;
                best.v = v;
            }
        }
        if (id) id[i] = best.id;
        if (t) t[i] = best.t;
        if (uv) {
            uv[2 * i] = best.u;
            uv[2 * i + 1] = best.v;
        }
    }
}

void or_pcg4d(const uint32_t* in, uint32_t* out) {
    uint32_t v[4] = {in[0], in[1], in[2], in[3]};
    pcg4d(v);
    for (int k = 0; k < 4; ++k) out[k] = v[k];
}

void or_sincos2pi(float u, float* s, float* c) { sincos2pi(u, *s, *c); }

void or_to_hemisphere(const float* v, const float* n, float* out) {
    V3 r = to_hemisphere(vld(v), vld(n));
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

void or_render_pass(const pt_triangle* tris, const uint32_t* perm, const uint32_t* children,
                    const float* boxes, uint32_t n, const pt_triangle_info* infos,
                    const pt_settings* st, int w, int h, const float* dirs, const float* eye,
                    uint32_t pass, uint64_t seed, double* accum, uint64_t* seg_counts,
                    int threads) {
    or_render_pass_lit(tris, perm, children, boxes, n, infos, nullptr, 0, st, w, h, dirs, eye, pass, seed,
                       accum, seg_counts, threads);
}

void or_render_pass_lit(const pt_triangle* tris, const uint32_t* perm, const uint32_t* children,
                        const float* boxes, uint32_t n, const pt_triangle_info* infos,
                        const pt_dir_light* lights, uint32_t nlights, const pt_settings* st, int w,
                        int h, const float* dirs, const float* eye, uint32_t pass, uint64_t seed,
                        double* accum, uint64_t* seg_counts, int threads) {
    or_render_pass_detail(tris, perm, children, boxes, n, infos, nullptr, nullptr, 0, lights, nlights, st, w,
                          h, dirs, eye, pass, seed, accum, seg_counts, threads);
}

namespace {
// One pass (or_render_pass_detail); stats (may be NULL) receives the traced
// segments, their node visits, box tests and triangle tests (the
// reference-order traversal's counts), and the segments that hit.
void render_pass_impl(const pt_triangle* tris, const uint32_t* perm, const uint32_t* children,
                      const float* boxes, uint32_t n, const pt_triangle_info* infos, const Detail& dt,
                      const pt_dir_light* lights, uint32_t nlights, const pt_settings* st, int w, int h,
                      const float* dirs, const float* eye, uint32_t pass, uint64_t seed, double* accum,
                      uint64_t* stats, int threads) {
    Bvh bvh{tris, perm, children, boxes, n};
    const V3 d00 = vld(dirs), d10 = vld(dirs + 3), d11 = vld(dirs + 6), d01 = vld(dirs + 9);
    const V3 e = vld(eye);
    const uint32_t seedw = static_cast<uint32_t>(seed) ^ static_cast<uint32_t>(seed >> 32);
    const uint32_t bounces = st->bounces;
    const uint32_t rr_start = st->rr_start;
    const float intensity = st->indirect_intensity;
    const bool strat = st->stratified != 0;
    uint64_t segs = 0, iters = 0, nbox = 0, ntri = 0, hits = 0;
    if (threads <= 0) threads = 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads) reduction(+ : segs, iters, nbox, ntri, hits)
    for (int64_t pix = 0; pix < static_cast<int64_t>(w) * h; ++pix) {
        const int x = static_cast<int>(pix % w), y = static_cast<int>(pix / w);
        Counters cnt;
        // Primary ray (cuda_path_tracer.cu:404-407, 376-379): bilinear blend of
        // the corner directions at the pixel corner (x/w, y/h), not renormalised.
        const float tx = static_cast<float>(x) / static_cast<float>(w);
        const float ty = static_cast<float>(y) / static_cast<float>(h);
        V3 d = mix(mix(d00, d10, tx), mix(d01, d11, tx), ty);
        V3 o = e;
        V3 accumc{0.0f, 0.0f, 0.0f};
        ++segs;
        Hit hit = trace_one(bvh, o, d, stats ? &cnt : nullptr);
        hits += hit.id != kInvalid;
        if (hit.id != kInvalid) {
            // blinnPhongShadingShadowTrace (cuda_path_tracer.cu:176-199) with no
            // directional lights: emission or zero.
            V3 pos = add(scale(d, hit.t), o);
            ShadeInfo si = extract(infos[hit.id], hit.id, dt, hit.u, hit.v);
            V3 albedo = si.albedo;
            V3 nrm = si.normal;
            if (emissive(si.emission)) {
                accumc = si.emission;
                albedo = V3{1.0f, 1.0f, 1.0f};
            } else if (nlights) {
                accumc = primary_lights(bvh, pos, d, nrm, albedo, infos[hit.id].specularity, lights, nlights);
            }
            V3 absorbed = albedo;
            // Indirect bounces (cuda_path_tracer.cu:267-330).
            for (uint32_t i = 0; i < bounces; ++i) {
                uint32_t rv[4] = {static_cast<uint32_t>(pix), pass, i, seedw};
                pcg4d(rv);
                float u1 = u24(rv[0]), u2 = u24(rv[1]);
                if (i >= rr_start) {
                    // Russian roulette (build addition): survive with
                    // p = max albedo component of the current vertex.
                    float p = gmax(gmax(albedo.x, albedo.y), albedo.z);
                    p = gmin(p, 1.0f);
                    if (!(u24(rv[2]) < p)) break;
                    absorbed = V3{absorbed.x / p, absorbed.y / p, absorbed.z / p};
                }
                o = add(pos, scale(nrm, 1e-5f));  // CUDA_EPSILON5
                if (strat && i == 0) {
                    // 8x8 strata, stratum = pass % 64 for every pixel (:273-280).
                    uint32_t p1 = pass % 64u;
                    u1 = (static_cast<float>(p1 % 8u) + u1) / 8.0f;
                    u2 = (static_cast<float>(p1 / 8u) + u2) / 8.0f;
                }
                // Cosine-weighted hemisphere sample (random.h:95-102).
                float r = std::sqrt(u1);
                float sn, cs;
                sincos2pi(u2, sn, cs);
                V3 local{r * cs, r * sn, std::sqrt(1.0f - u1)};
                d = to_hemisphere(local, nrm);
                ++segs;
                hit = trace_one(bvh, o, d, stats ? &cnt : nullptr);
                if (hit.id == kInvalid) break;
                ++hits;
                pos = add(scale(d, hit.t), o);
                si = extract(infos[hit.id], hit.id, dt, hit.u, hit.v);
                nrm = si.normal;
                albedo = si.albedo;
                V3 L{0.0f, 0.0f, 0.0f};
                if (emissive(si.emission)) {
                    float c = dot(neg(d), nrm);
                    c = gmax(c, 0.0f);
                    L = scale(si.emission, c);
                    albedo = V3{1.0f, 1.0f, 1.0f};
                } else if (nlights) {
                    L = bounce_lights(bvh, pos, nrm, lights, nlights);
                }
                absorbed = mul(absorbed, albedo);
                accumc = add(accumc, scale(mul(absorbed, L), intensity));
            }
        }
        double* a = accum + 3 * pix;
        a[0] += static_cast<double>(accumc.x);
        a[1] += static_cast<double>(accumc.y);
        a[2] += static_cast<double>(accumc.z);
        iters += cnt.iters;
        nbox += cnt.boxes;
        ntri += cnt.tris;
    }
    if (stats) {
        stats[0] = segs;
        stats[1] = iters;
        stats[2] = nbox;
        stats[3] = ntri;
        stats[4] = hits;
    }
}
}  // namespace

void or_render_pass_detail(const pt_triangle* tris, const uint32_t* perm, const uint32_t* children,
                           const float* boxes, uint32_t n, const pt_triangle_info* infos,
                           const pt_triangle_detail* detail, const pt_texture* textures,
                           uint32_t ntextures, const pt_dir_light* lights, uint32_t nlights,
                           const pt_settings* st, int w, int h, const float* dirs, const float* eye,
                           uint32_t pass, uint64_t seed, double* accum, uint64_t* seg_counts,
                           int threads) {
    uint64_t stats[5];
    render_pass_impl(tris, perm, children, boxes, n, infos, Detail{detail, textures, ntextures}, lights,
                     nlights, st, w, h, dirs, eye, pass, seed, accum, seg_counts ? stats : nullptr, threads);
    if (seg_counts) *seg_counts = stats[0];
}

void or_render_pass_stats(const pt_triangle* tris, const uint32_t* perm, const uint32_t* children,
                          const float* boxes, uint32_t n, const pt_triangle_info* infos,
                          const pt_settings* st, int w, int h, const float* dirs, const float* eye,
                          uint32_t pass, uint64_t seed, double* accum, uint64_t* stats, int threads) {
    render_pass_impl(tris, perm, children, boxes, n, infos, Detail{}, nullptr, 0, st, w, h, dirs, eye, pass,
                     seed, accum, stats, threads);
}

// Learned-cache frame (the build's own addition; no reference counterpart).
// Pixel paths follow or_render_pass to their second vertex x2 and add
// throughput * cache(x2), the cache evaluated in FP64 by or_mlp_forward with
// the given parameters.  Training path k starts on pixel
// (ox + 4 (c % cw), oy + 4 (c / cw)), c = floor(k * cw * ch / ntrain),
// (ox, oy) = (pass % 4, (pass / 4) % 4) clamped to the image; it draws from
// the stream seedw ^ 0x5bd1e995 keyed by k, records x2 (position, normal,
// incoming direction, albedo) in train_raw[12k..] and gathers the suffix
// radiance past x2 (throughput reset to 1 there) for extra_bounces bounces
// into train_target[3k..]; NaN when it never reaches x2.
void or_render_cache_pass(const pt_triangle* tris, const uint32_t* perm, const uint32_t* children,
                          const float* boxes, uint32_t n, const pt_triangle_info* infos,
                          const pt_dir_light* lights, uint32_t nlights,
                          const pt_settings* st, int w, int h, const float* dirs, const float* eye,
                          uint32_t pass, uint64_t seed, const float* params, const float* cbmin,
                          const float* cbmax, uint32_t ntrain, uint32_t extra_bounces, double* accum,
                          float* train_raw, float* train_target, int threads) {
    or_render_cache_pass_detail(tris, perm, children, boxes, n, infos, nullptr, nullptr, 0, lights, nlights,
                                st, w, h, dirs, eye, pass, seed, params, cbmin, cbmax, ntrain,
                                extra_bounces, accum, train_raw, train_target, threads);
}

void or_render_cache_pass_detail(const pt_triangle* tris, const uint32_t* perm, const uint32_t* children,
                                 const float* boxes, uint32_t n, const pt_triangle_info* infos,
                                 const pt_triangle_detail* detail, const pt_texture* textures,
                                 uint32_t ntextures, const pt_dir_light* lights, uint32_t nlights,
                                 const pt_settings* st, int w, int h, const float* dirs, const float* eye,
                                 uint32_t pass, uint64_t seed, const float* params, const float* cbmin,
                                 const float* cbmax, uint32_t ntrain, uint32_t extra_bounces,
                                 double* accum, float* train_raw, float* train_target, int threads) {
    const Detail dt{detail, textures, ntextures};
    Bvh bvh{tris, perm, children, boxes, n};
    const V3 d00 = vld(dirs), d10 = vld(dirs + 3), d11 = vld(dirs + 6), d01 = vld(dirs + 9);
    const V3 e = vld(eye);
    const uint32_t seedw = static_cast<uint32_t>(seed) ^ static_cast<uint32_t>(seed >> 32);
    const uint32_t rr_start = st->rr_start;
    const float intensity = st->indirect_intensity;
    const bool strat = st->stratified != 0;
    if (threads <= 0) threads = 1;
    const int64_t npix = static_cast<int64_t>(w) * h;
    const uint32_t ox = std::min(pass % 4u, static_cast<uint32_t>(w) - 1u);
    const uint32_t oy = std::min((pass / 4u) % 4u, static_cast<uint32_t>(h) - 1u);
    const uint32_t cw = (static_cast<uint32_t>(w) - ox + 3u) / 4u;
    const uint32_t ch = (static_cast<uint32_t>(h) - oy + 3u) / 4u;

    // One path of the cache frame; returns whether it reached x2.  levels:
    // hits to take (2 for a pixel path).  train: reset at x2 and continue.
    auto walk = [&](int x, int y, uint32_t key, uint32_t sw, uint32_t levels, bool train, float* raw,
                    V3& absorbed_out, V3& acc_out) -> bool {
        const float tx = static_cast<float>(x) / static_cast<float>(w);
        const float ty = static_cast<float>(y) / static_cast<float>(h);
        V3 d = mix(mix(d00, d10, tx), mix(d01, d11, tx), ty);
        V3 o = e;
        V3 acc{0.0f, 0.0f, 0.0f};
        bool reached = false;
        Hit hit = trace_one(bvh, o, d, nullptr);
        if (hit.id == kInvalid) return false;
        V3 pos = add(scale(d, hit.t), o);
        ShadeInfo si = extract(infos[hit.id], hit.id, dt, hit.u, hit.v);
        V3 albedo = si.albedo;
        V3 nrm = si.normal;
        if (emissive(si.emission)) {
            acc = si.emission;
            albedo = V3{1.0f, 1.0f, 1.0f};
        } else if (nlights) {
            acc = primary_lights(bvh, pos, d, nrm, albedo, infos[hit.id].specularity, lights, nlights);
        }
        V3 absorbed = albedo;
        for (uint32_t i = 0; i + 1 < levels; ++i) {
            uint32_t rv[4] = {key, pass, i, sw};
            pcg4d(rv);
            float u1 = u24(rv[0]), u2 = u24(rv[1]);
            if (i >= rr_start) {
                float p = gmax(gmax(albedo.x, albedo.y), albedo.z);
                p = gmin(p, 1.0f);
                if (!(u24(rv[2]) < p)) break;
                absorbed = V3{absorbed.x / p, absorbed.y / p, absorbed.z / p};
            }
            o = add(pos, scale(nrm, 1e-5f));
            if (strat && i == 0) {
                uint32_t p1 = pass % 64u;
                u1 = (static_cast<float>(p1 % 8u) + u1) / 8.0f;
                u2 = (static_cast<float>(p1 / 8u) + u2) / 8.0f;
            }
            float r = std::sqrt(u1);
            float sn, cs;
            sincos2pi(u2, sn, cs);
            V3 local{r * cs, r * sn, std::sqrt(1.0f - u1)};
            d = to_hemisphere(local, nrm);
            hit = trace_one(bvh, o, d, nullptr);
            if (hit.id == kInvalid) break;
            pos = add(scale(d, hit.t), o);
            si = extract(infos[hit.id], hit.id, dt, hit.u, hit.v);
            nrm = si.normal;
            albedo = si.albedo;
            V3 L{0.0f, 0.0f, 0.0f};
            if (emissive(si.emission)) {
                float c = dot(neg(d), nrm);
                c = gmax(c, 0.0f);
                L = scale(si.emission, c);
                albedo = V3{1.0f, 1.0f, 1.0f};
            } else if (nlights) {
                L = bounce_lights(bvh, pos, nrm, lights, nlights);
            }
            absorbed = mul(absorbed, albedo);
            acc = add(acc, scale(mul(absorbed, L), intensity));
            if (i == 0) {
                const float f[12] = {pos.x, pos.y, pos.z, nrm.x, nrm.y, nrm.z,
                                     d.x,   d.y,   d.z,   albedo.x, albedo.y, albedo.z};
                std::memcpy(raw, f, sizeof f);
                reached = true;
                if (train) {
                    absorbed = V3{1.0f, 1.0f, 1.0f};
                    acc = V3{0.0f, 0.0f, 0.0f};
                }
            }
        }
        absorbed_out = absorbed;
        acc_out = acc;
        return reached;
    };

    std::vector<float> qraw(static_cast<size_t>(npix) * 12, 0.0f);
    std::vector<float> qabs(static_cast<size_t>(npix) * 3), qacc(static_cast<size_t>(npix) * 3);
    std::vector<uint8_t> qvalid(static_cast<size_t>(npix));
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
    for (int64_t pix = 0; pix < npix; ++pix) {
        V3 a{0, 0, 0}, c{0, 0, 0};
        qvalid[pix] = walk(static_cast<int>(pix % w), static_cast<int>(pix / w), static_cast<uint32_t>(pix), seedw,
                           2u, false, &qraw[12 * pix], a, c);
        qabs[3 * pix] = a.x;
        qabs[3 * pix + 1] = a.y;
        qabs[3 * pix + 2] = a.z;
        qacc[3 * pix] = c.x;
        qacc[3 * pix + 1] = c.y;
        qacc[3 * pix + 2] = c.z;
    }
    std::vector<double> pred(static_cast<size_t>(npix) * 3);
    or_mlp_forward(params, qraw.data(), static_cast<uint32_t>(npix), cbmin, cbmax, pred.data());
    for (int64_t pix = 0; pix < npix; ++pix) {
        for (int k = 0; k < 3; ++k) {
            float v = qacc[3 * pix + k];
            if (qvalid[pix]) v = v + qabs[3 * pix + k] * static_cast<float>(pred[3 * pix + k]);
            accum[3 * pix + k] += static_cast<double>(v);
        }
    }
    const uint32_t tsw = seedw ^ 0x5bd1e995u;
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
    for (int64_t k = 0; k < static_cast<int64_t>(ntrain); ++k) {
        const uint64_t c = (static_cast<uint64_t>(k) * (static_cast<uint64_t>(cw) * ch)) / ntrain;
        const int x = static_cast<int>(ox + 4u * static_cast<uint32_t>(c % cw));
        const int y = static_cast<int>(oy + 4u * static_cast<uint32_t>(c / cw));
        V3 a{0, 0, 0}, acc{0, 0, 0};
        float* raw = train_raw + 12 * k;
        for (int j = 0; j < 12; ++j) raw[j] = 0.0f;
        const bool ok = walk(x, y, static_cast<uint32_t>(k), tsw, 2u + extra_bounces, true, raw, a, acc);
        float* t = train_target + 3 * k;
        if (ok) {
            t[0] = acc.x;
            t[1] = acc.y;
            t[2] = acc.z;
        } else {
            t[0] = t[1] = t[2] = std::numeric_limits<float>::quiet_NaN();
        }
    }
}

float or_pow_pos(float x, float y) { return pow_pos(x, y); }

void or_tex_sample(const pt_texture* tex, float s, float t, float* rgb) {
    const V3 c = tex_sample(*tex, s, t);
    rgb[0] = c.x;
    rgb[1] = c.y;
    rgb[2] = c.z;
}

void or_shade_detail(const pt_triangle_info* info, const pt_triangle_detail* detail,
                     const pt_texture* textures, uint32_t ntextures, float u, float v, float* normal,
                     float* albedo) {
    const Detail dt{detail, textures, ntextures};
    const ShadeInfo si = extract(*info, 0, dt, u, v);
    normal[0] = si.normal.x;
    normal[1] = si.normal.y;
    normal[2] = si.normal.z;
    albedo[0] = si.albedo.x;
    albedo[1] = si.albedo.y;
    albedo[2] = si.albedo.z;
}

// a14: PathTracingApp::screenToRay (PathTracingApp.cpp:217-242) for the main
// camera component.  View matrix rows right, up, look with the position
// folded in (CameraComponent::updateViewMatrix, CameraComponent.cpp:97-135);
// glm::perspective with -1..1 depth, tanHalfFovy computed in float from the
// value as passed (third_party/glm/glm/gtc/matrix_transform.inl:248-300):
// perspectiveRH (left_handed = 0, glm's default, w = -z) or perspectiveLH
// (left_handed = 1, GLM_FORCE_LEFT_HANDED, w = +z); the jitter translate
// times viewProj, inverted; screenToNDC (CameraComponent.cpp:149-155); start
// and end points unprojected and divided by w; direction normalize(end -
// start).  The matrix work is in double with a general 4x4 inverse
// (Gauss-Jordan, partial pivoting), the reference's is in float: the
// directions agree to a few float ulps, not bit for bit.
void or_screen_to_ray(const float* rot, const float* pos, float fovy, float screen_w, float screen_h,
                      float znear, float zfar, float jx, float jy, const float* p, int left_handed,
                      float* dir) {
    double V[4][4] = {}, P[4][4] = {}, T[4][4] = {}, M[4][4] = {}, TP[4][4] = {};
    for (int r = 0; r < 3; ++r) {
        double d = 0.0;
        for (int c = 0; c < 3; ++c) {
            V[r][c] = rot[3 * r + c];  // row r = rotation column r (right, up, look)
            d += static_cast<double>(pos[c]) * rot[3 * r + c];
        }
        V[r][3] = -d;
    }
    V[3][3] = 1.0;
    const float tan_half = std::tan(fovy / 2.0f);
    const double aspect = static_cast<double>(screen_w) / screen_h;
    const double n = znear, f = zfar;
    P[0][0] = 1.0 / (aspect * tan_half);
    P[1][1] = 1.0 / tan_half;
    P[2][2] = (left_handed ? 1.0 : -1.0) * (f + n) / (f - n);
    P[2][3] = -(2.0 * f * n) / (f - n);
    P[3][2] = left_handed ? 1.0 : -1.0;
    for (int k = 0; k < 4; ++k) T[k][k] = 1.0;
    T[0][3] = jx;
    T[1][3] = jy;
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c)
            for (int k = 0; k < 4; ++k) TP[r][c] += T[r][k] * P[k][c];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c)
            for (int k = 0; k < 4; ++k) M[r][c] += TP[r][k] * V[k][c];
    // inverse(M): Gauss-Jordan on [M | I].
    double A[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 8; ++c) A[r][c] = c < 4 ? M[r][c] : (c - 4 == r ? 1.0 : 0.0);
    for (int c = 0; c < 4; ++c) {
        int piv = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(A[r][c]) > std::fabs(A[piv][c])) piv = r;
        for (int k = 0; k < 8; ++k) std::swap(A[c][k], A[piv][k]);
        const double s = 1.0 / A[c][c];
        for (int k = 0; k < 8; ++k) A[c][k] *= s;
        for (int r = 0; r < 4; ++r) {
            if (r == c) continue;
            const double m = A[r][c];
            for (int k = 0; k < 8; ++k) A[r][k] -= m * A[c][k];
        }
    }
    const double ndc[3] = {p[0] / static_cast<double>(screen_w) * 2.0 - 1.0,
                           p[1] / static_cast<double>(screen_h) * 2.0 - 1.0,
                           (p[2] - n) / (f - n) * 2.0 - 1.0};
    auto unproject = [&](double z, double* out) {
        const double v[4] = {ndc[0], ndc[1], z, 1.0};
        double w[4] = {0.0, 0.0, 0.0, 0.0};
        for (int r = 0; r < 4; ++r)
            for (int k = 0; k < 4; ++k) w[r] += A[r][4 + k] * v[k];
        for (int r = 0; r < 3; ++r) out[r] = w[r] / w[3];
    };
    double s[3], e[3];
    unproject(ndc[2], s);
    unproject(1.0, e);
    const double d[3] = {e[0] - s[0], e[1] - s[1], e[2] - s[2]};
    const double len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int k = 0; k < 3; ++k) dir[k] = static_cast<float>(d[k] / len);
}

}  // extern "C"
