// Seeded synthetic workloads (host only).  Replaces the reference app's scene
// extraction (PathTracingApp.cpp:31-124) and corner-ray setup
// (PathTracingApp.cpp:217-242) with deterministic generators.
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "pt_scenes.h"

namespace {

// SplitMix64 (Steele, Lea, Flood, OOPSLA 2014).
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // 24-bit uniform float in [0, 1).
    float u01() { return static_cast<float>(next() >> 40) * (1.0f / 16777216.0f); }
    // 53-bit uniform double in [0, 1).
    double d01() { return static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0); }
    double gauss() {  // Box-Muller
        double u1 = d01();
        double u2 = d01();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

uint64_t mix64(uint64_t a, uint64_t b) {
    SplitMix64 r(a ^ (b * 0xD1B54A32D192ED03ull));
    return r.next();
}

// glm::min / glm::max scalar semantics (third_party/glm/glm/detail/func_common.inl:15-28).
inline float gmin(float x, float y) { return x < y ? x : y; }
inline float gmax(float x, float y) { return x > y ? x : y; }

struct V3 {
    float x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
inline float dot(V3 a, V3 b) {
    float x = a.x * b.x;
    float y = a.y * b.y;
    float z = a.z * b.z;
    return (x + y) + z;
}
inline V3 normalize(V3 v) {
    float s = 1.0f / std::sqrt(dot(v, v));
    return {v.x * s, v.y * s, v.z * s};
}

void set_tri(pt_triangle* t, V3 a, V3 b, V3 c, uint32_t id) {
    t->v0[0] = a.x; t->v0[1] = a.y; t->v0[2] = a.z;
    t->v1[0] = b.x; t->v1[1] = b.y; t->v1[2] = b.z;
    t->v2[0] = c.x; t->v2[1] = c.y; t->v2[2] = c.z;
    t->id = id;
}

// Flat shading record: the reference falls back to the geometric normal for
// every vertex when a mesh has no normals (PathTracingApp.cpp:74-77).
void set_info(pt_triangle_info* info, const pt_triangle* t, float ar, float ag, float ab,
              float e) {
    V3 v0{t->v0[0], t->v0[1], t->v0[2]};
    V3 v1{t->v1[0], t->v1[1], t->v1[2]};
    V3 v2{t->v2[0], t->v2[1], t->v2[2]};
    V3 n = normalize(cross(sub(v1, v0), sub(v2, v0)));
    if (!(n.x == n.x) || !(n.y == n.y) || !(n.z == n.z)) n = {0.0f, 1.0f, 0.0f};  // zero-area
    for (int k = 0; k < 3; ++k) {
        info->n0[k] = (&n.x)[k];
        info->n1[k] = (&n.x)[k];
        info->n2[k] = (&n.x)[k];
    }
    info->albedo[0] = ar;
    info->albedo[1] = ag;
    info->albedo[2] = ab;
    info->emission[0] = e;
    info->emission[1] = e;
    info->emission[2] = e;
    info->specularity = 0.0f;
}

// Axis-aligned quad at height y over [x0,x1] x [z0,z1], two triangles whose
// geometric normal points -y (down).
uint32_t emit_quad_down(pt_triangle* tris, pt_triangle_info* infos, uint32_t at, uint32_t id,
                        float x0, float x1, float y, float z0, float z1, float albedo,
                        float emission) {
    V3 a{x0, y, z0}, b{x1, y, z0}, c{x1, y, z1}, d{x0, y, z1};
    set_tri(&tris[at], a, b, c, id);
    set_tri(&tris[at + 1], a, c, d, id + 1);
    if (infos) {
        set_info(&infos[id], &tris[at], albedo, albedo, albedo, emission);
        set_info(&infos[id + 1], &tris[at + 1], albedo, albedo, albedo, emission);
    }
    return 2;
}

// 3D value noise: lattice values in [-1,1] from a hash, smoothstep-weighted
// trilinear interpolation.
double lattice(int64_t x, int64_t y, int64_t z, uint64_t seed) {
    uint64_t h = mix64(seed, static_cast<uint64_t>(x) * 0x8CB92BA72F3D8DD7ull ^
                                 static_cast<uint64_t>(y) * 0x9E6C63D0676A9A99ull ^
                                 static_cast<uint64_t>(z) * 0xA24BAED4963EE407ull);
    return static_cast<double>(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}
double value_noise(double x, double y, double z, uint64_t seed) {
    double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int64_t ix = static_cast<int64_t>(fx), iy = static_cast<int64_t>(fy), iz = static_cast<int64_t>(fz);
    double tx = x - fx, ty = y - fy, tz = z - fz;
    tx = tx * tx * (3.0 - 2.0 * tx);
    ty = ty * ty * (3.0 - 2.0 * ty);
    tz = tz * tz * (3.0 - 2.0 * tz);
    double c[2][2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int d = 0; d < 2; ++d) c[a][b][d] = lattice(ix + a, iy + b, iz + d, seed);
    double x00 = c[0][0][0] + tx * (c[1][0][0] - c[0][0][0]);
    double x10 = c[0][1][0] + tx * (c[1][1][0] - c[0][1][0]);
    double x01 = c[0][0][1] + tx * (c[1][0][1] - c[0][0][1]);
    double x11 = c[0][1][1] + tx * (c[1][1][1] - c[0][1][1]);
    double y0 = x00 + ty * (x10 - x00);
    double y1 = x01 + ty * (x11 - x01);
    return y0 + tz * (y1 - y0);
}
double fbm6(double x, double y, double z, uint64_t seed) {
    double sum = 0.0, amp = 1.0, norm = 0.0, f = 2.0;
    for (int o = 0; o < 6; ++o) {
        sum += amp * value_noise(x * f, y * f, z * f, seed + static_cast<uint64_t>(o) * 7919u);
        norm += amp;
        amp *= 0.5;
        f *= 2.0;
    }
    return sum / norm;
}

}  // namespace

extern "C" {

int pt_scene_bounds(const pt_triangle* tris, uint32_t n, float* bmin, float* bmax) {
    if (!bmin || !bmax || (n && !tris)) return PT_E_INVALID_ARG;
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t i = 0; i < n; ++i) {
        // Triangle::bbox() = BBox::from({v0, v1, v2}) then enclosing.unite(bbox).
        for (int k = 0; k < 3; ++k) {
            float lo = FLT_MAX, hi = -FLT_MAX;
            lo = gmin(lo, tris[i].v0[k]); hi = gmax(hi, tris[i].v0[k]);
            lo = gmin(lo, tris[i].v1[k]); hi = gmax(hi, tris[i].v1[k]);
            lo = gmin(lo, tris[i].v2[k]); hi = gmax(hi, tris[i].v2[k]);
            mn[k] = gmin(mn[k], lo);
            mx[k] = gmax(mx[k], hi);
        }
    }
    for (int k = 0; k < 3; ++k) {
        bmin[k] = mn[k];
        bmax[k] = mx[k];
    }
    return static_cast<int>(n);
}

int pt_gen_soup(uint64_t seed, uint32_t n, int with_emitter, uint32_t first_id,
                pt_triangle* tris, pt_triangle_info* infos) {
    if (!tris) return PT_E_INVALID_ARG;
    SplitMix64 rng(seed);
    for (uint32_t i = 0; i < n; ++i) {
        float cx = rng.u01(), cy = rng.u01(), cz = rng.u01();
        V3 v[3];
        for (int k = 0; k < 3; ++k) {
            float ox = (rng.u01() * 2.0f - 1.0f) * 0.01f;
            float oy = (rng.u01() * 2.0f - 1.0f) * 0.01f;
            float oz = (rng.u01() * 2.0f - 1.0f) * 0.01f;
            v[k] = {cx + ox, cy + oy, cz + oz};
        }
        set_tri(&tris[i], v[0], v[1], v[2], first_id + i);
        if (infos) set_info(&infos[i], &tris[i], 0.7f, 0.7f, 0.7f, 0.0f);
    }
    uint32_t count = n;
    if (with_emitter) {
        // infos are indexed by id - first_id in this generator's output.
        pt_triangle_info* base = infos;
        V3 a{0.3f, 1.2f, 0.3f}, b{0.7f, 1.2f, 0.3f}, c{0.7f, 1.2f, 0.7f}, d{0.3f, 1.2f, 0.7f};
        set_tri(&tris[n], a, b, c, first_id + n);
        set_tri(&tris[n + 1], a, c, d, first_id + n + 1);
        if (base) {
            set_info(&base[n], &tris[n], 0.7f, 0.7f, 0.7f, 10.0f);
            set_info(&base[n + 1], &tris[n + 1], 0.7f, 0.7f, 0.7f, 10.0f);
        }
        count += 2;
    }
    return static_cast<int>(count);
}

int pt_gen_sphere(uint64_t seed, uint32_t bands, int with_room, pt_triangle* tris,
                  pt_triangle_info* infos) {
    if (!tris || bands < 2 || bands > 20000) return PT_E_INVALID_ARG;
    const uint32_t B = bands;
    const double pi = 3.14159265358979323846;
    // Vertex grid: (B + 1) latitude rows (poles included) x B longitudes.
    V3* grid = new V3[static_cast<size_t>(B + 1) * B];
    for (uint32_t i = 0; i <= B; ++i) {
        double th = pi * static_cast<double>(i) / static_cast<double>(B);
        for (uint32_t j = 0; j < B; ++j) {
            double ph = 2.0 * pi * static_cast<double>(j) / static_cast<double>(B);
            double px, py, pz;
            if (i == 0) {
                px = 0.0; py = 1.0; pz = 0.0;
            } else if (i == B) {
                px = 0.0; py = -1.0; pz = 0.0;
            } else {
                px = std::sin(th) * std::cos(ph);
                py = std::cos(th);
                pz = std::sin(th) * std::sin(ph);
            }
            double r = 1.0 + 0.15 * fbm6(px, py, pz, seed);
            grid[static_cast<size_t>(i) * B + j] = {static_cast<float>(r * px),
                                                    static_cast<float>(r * py),
                                                    static_cast<float>(r * pz)};
        }
    }
    SplitMix64 rng(seed ^ 0xA5A5A5A5u);
    uint32_t t = 0;
    for (uint32_t i = 0; i < B; ++i) {
        for (uint32_t j = 0; j < B; ++j) {
            uint32_t j1 = (j + 1) % B;
            V3 a = grid[static_cast<size_t>(i) * B + j];
            V3 b = grid[static_cast<size_t>(i + 1) * B + j];
            V3 c = grid[static_cast<size_t>(i + 1) * B + j1];
            V3 d = grid[static_cast<size_t>(i) * B + j1];
            // Outward winding: cross(v1 - v0, v2 - v0) points away from the centre.
            set_tri(&tris[t], a, c, b, t);
            set_tri(&tris[t + 1], a, d, c, t + 1);
            for (int k = 0; k < 2; ++k) {
                float ar = 0.3f + 0.6f * rng.u01();
                float ag = 0.3f + 0.6f * rng.u01();
                float ab = 0.3f + 0.6f * rng.u01();
                if (infos) set_info(&infos[t + k], &tris[t + k], ar, ag, ab, 0.0f);
            }
            t += 2;
        }
    }
    delete[] grid;
    if (with_room) {
        const float s = 2.0f;
        // Corners of the room box [-2,2]^3.
        V3 p[8] = {{-s, -s, -s}, {s, -s, -s}, {s, s, -s}, {-s, s, -s},
                   {-s, -s, s},  {s, -s, s},  {s, s, s},  {-s, s, s}};
        // Faces as quads (a, b, c, d) wound so that the normal points inward.
        const int q[6][4] = {
            {0, 4, 5, 1},  // floor y=-2, normal +y
            {3, 2, 6, 7},  // ceiling y=+2, normal -y
            {0, 3, 7, 4},  // x=-2, normal +x
            {1, 5, 6, 2},  // x=+2, normal -x
            {0, 1, 2, 3},  // z=-2, normal +z
            {4, 7, 6, 5},  // z=+2, normal -z
        };
        for (int f = 0; f < 6; ++f) {
            V3 a = p[q[f][0]], b = p[q[f][1]], c = p[q[f][2]], d = p[q[f][3]];
            set_tri(&tris[t], a, b, c, t);
            set_tri(&tris[t + 1], a, c, d, t + 1);
            if (infos) {
                set_info(&infos[t], &tris[t], 0.7f, 0.7f, 0.7f, 0.0f);
                set_info(&infos[t + 1], &tris[t + 1], 0.7f, 0.7f, 0.7f, 0.0f);
            }
            t += 2;
        }
        t += emit_quad_down(tris, infos, t, t, -0.5f, 0.5f, 1.98f, -0.5f, 0.5f, 0.7f, 10.0f);
    }
    return static_cast<int>(t);
}

int pt_gen_hairball(uint64_t seed, uint32_t strands, uint32_t steps, int with_light,
                    pt_triangle* tris, pt_triangle_info* infos) {
    if (!tris) return PT_E_INVALID_ARG;
    SplitMix64 rng(seed);
    const double hw = 0.001, len = 0.005, sigma = 0.3;
    uint32_t t = 0;
    for (uint32_t s = 0; s < strands; ++s) {
        // Root uniform on the unit sphere; initial direction = outward normal.
        double z = 2.0 * rng.d01() - 1.0;
        double ph = 6.283185307179586 * rng.d01();
        double rr = std::sqrt(1.0 - z * z);
        double p[3] = {rr * std::cos(ph), rr * std::sin(ph), z};
        double d[3] = {p[0], p[1], p[2]};
        for (uint32_t k = 0; k < steps; ++k) {
            // Rotate d by N(0, sigma) about a random axis perpendicular to d.
            double ref[3] = {0.0, 1.0, 0.0};
            if (std::fabs(d[1]) > 0.9) { ref[0] = 1.0; ref[1] = 0.0; }
            double u[3] = {d[1] * ref[2] - d[2] * ref[1], d[2] * ref[0] - d[0] * ref[2],
                           d[0] * ref[1] - d[1] * ref[0]};
            double un = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            for (double& c : u) c /= un;
            double v[3] = {d[1] * u[2] - d[2] * u[1], d[2] * u[0] - d[0] * u[2],
                           d[0] * u[1] - d[1] * u[0]};
            double az = 6.283185307179586 * rng.d01();
            double ang = sigma * rng.gauss();
            double ax[3];
            for (int c = 0; c < 3; ++c) ax[c] = std::cos(az) * u[c] + std::sin(az) * v[c];
            double ca = std::cos(ang), sa = std::sin(ang);
            double nd[3];
            for (int c = 0; c < 3; ++c) nd[c] = ca * d[c] + sa * ax[c];
            double nn = std::sqrt(nd[0] * nd[0] + nd[1] * nd[1] + nd[2] * nd[2]);
            for (int c = 0; c < 3; ++c) d[c] = nd[c] / nn;
            double q[3] = {p[0] + len * d[0], p[1] + len * d[1], p[2] + len * d[2]};
            // Camera-independent ribbon: side = normalize(cross(d, ref)).
            double sd[3] = {d[1] * ref[2] - d[2] * ref[1], d[2] * ref[0] - d[0] * ref[2],
                            d[0] * ref[1] - d[1] * ref[0]};
            double sn = std::sqrt(sd[0] * sd[0] + sd[1] * sd[1] + sd[2] * sd[2]);
            for (double& c : sd) c = c / sn * hw;
            V3 a0{(float)(p[0] - sd[0]), (float)(p[1] - sd[1]), (float)(p[2] - sd[2])};
            V3 a1{(float)(p[0] + sd[0]), (float)(p[1] + sd[1]), (float)(p[2] + sd[2])};
            V3 b0{(float)(q[0] - sd[0]), (float)(q[1] - sd[1]), (float)(q[2] - sd[2])};
            V3 b1{(float)(q[0] + sd[0]), (float)(q[1] + sd[1]), (float)(q[2] + sd[2])};
            set_tri(&tris[t], a0, a1, b1, t);
            set_tri(&tris[t + 1], a0, b1, b0, t + 1);
            if (infos) {
                set_info(&infos[t], &tris[t], 0.7f, 0.7f, 0.7f, 0.0f);
                set_info(&infos[t + 1], &tris[t + 1], 0.7f, 0.7f, 0.7f, 0.0f);
            }
            t += 2;
            for (int c = 0; c < 3; ++c) p[c] = q[c];
        }
    }
    if (with_light) t += emit_quad_down(tris, infos, t, t, -1.0f, 1.0f, 2.5f, -1.0f, 1.0f, 0.7f, 10.0f);
    return static_cast<int>(t);
}

int pt_gen_shell_rays(uint64_t seed, const float* bmin, const float* bmax, float enlarge,
                      uint32_t n, float* rays) {
    if (!bmin || !bmax || !rays) return PT_E_INVALID_ARG;
    float c[3], h[3];
    for (int k = 0; k < 3; ++k) {
        c[k] = 0.5f * (bmin[k] + bmax[k]);
        h[k] = 0.5f * (bmax[k] - bmin[k]) * enlarge;
    }
    // Face areas of the enlarged box: pairs of faces normal to x, y, z.
    double ax = 4.0 * h[1] * h[2], ay = 4.0 * h[0] * h[2], az = 4.0 * h[0] * h[1];
    double tot = 2.0 * (ax + ay + az);
    SplitMix64 rng(seed);
    for (uint32_t i = 0; i < n; ++i) {
        double pick = rng.d01() * tot;
        int axis, side;
        if (pick < 2.0 * ax) { axis = 0; side = pick < ax ? 0 : 1; }
        else if (pick < 2.0 * (ax + ay)) { axis = 1; side = (pick - 2.0 * ax) < ay ? 0 : 1; }
        else { axis = 2; side = (pick - 2.0 * (ax + ay)) < az ? 0 : 1; }
        float o[3];
        for (int k = 0; k < 3; ++k) o[k] = c[k] + h[k] * (rng.u01() * 2.0f - 1.0f);
        o[axis] = c[axis] + (side ? h[axis] : -h[axis]);
        float z = rng.u01() * 2.0f - 1.0f;
        float ph = 6.2831853f * rng.u01();
        float r = std::sqrt(gmax(0.0f, 1.0f - z * z));
        float* ray = rays + 6 * static_cast<size_t>(i);
        ray[0] = o[0]; ray[1] = o[1]; ray[2] = o[2];
        ray[3] = r * std::cos(ph);
        ray[4] = r * std::sin(ph);
        ray[5] = z;
    }
    return static_cast<int>(n);
}

int pt_gen_batch_counts(uint64_t seed, uint32_t nscenes, uint32_t lo, uint32_t hi,
                        uint32_t* counts) {
    if (!counts || lo == 0 || hi < lo) return PT_E_INVALID_ARG;
    SplitMix64 rng(seed);
    double llo = std::log(static_cast<double>(lo)), lhi = std::log(static_cast<double>(hi));
    for (uint32_t i = 0; i < nscenes; ++i) {
        double v = std::exp(llo + (lhi - llo) * rng.d01());
        uint32_t c = static_cast<uint32_t>(v);
        if (c < lo) c = lo;
        if (c > hi) c = hi;
        counts[i] = c;
    }
    return static_cast<int>(nscenes);
}

int pt_camera_corners_basis(const float* right, const float* up, const float* look, float fovy,
                            float aspect, float jx, float jy, int left_handed, float* dirs) {
    if (!right || !up || !look || !dirs) return PT_E_INVALID_ARG;
    // screenToRay's corner directions: the unprojected NDC corners (shifted by
    // the jitter translation) lie on +-look + ndc.x * aspect * t * right +
    // ndc.y * t * up, t = tan(fovy / 2) as glm::perspective computes it; the
    // left-handed projection looks along +look, the right-handed along -look.
    const float sgn = left_handed ? 1.0f : -1.0f;
    const V3 r{right[0], right[1], right[2]};
    const V3 u{up[0], up[1], up[2]};
    const V3 b{sgn * look[0], sgn * look[1], sgn * look[2]};
    const float th = std::tan(fovy / 2.0f);
    const float cx[4] = {-1.0f, 1.0f, 1.0f, -1.0f};
    const float cy[4] = {-1.0f, -1.0f, 1.0f, 1.0f};
    for (int k = 0; k < 4; ++k) {
        const float sx = (cx[k] - jx) * th * aspect;
        const float sy = (cy[k] - jy) * th;
        V3 d{b.x + sx * r.x + sy * u.x, b.y + sx * r.y + sy * u.y, b.z + sx * r.z + sy * u.z};
        d = normalize(d);
        dirs[3 * k + 0] = d.x;
        dirs[3 * k + 1] = d.y;
        dirs[3 * k + 2] = d.z;
    }
    return 4;
}

int pt_camera_corners(const float* eye, const float* target, const float* up, float fovy_deg,
                      float aspect, float jx, float jy, float* dirs) {
    if (!eye || !target || !up || !dirs) return PT_E_INVALID_ARG;
    // Transform::lookAt (Transform.cpp:78-87): look = normalize(target - eye),
    // right = normalize(up x look), up' = look x right; seen through the
    // left-handed projection, so the rays leave towards the target.
    const V3 e{eye[0], eye[1], eye[2]};
    const V3 f = normalize(sub(V3{target[0], target[1], target[2]}, e));
    const V3 r = normalize(cross(V3{up[0], up[1], up[2]}, f));
    const V3 u = cross(f, r);
    const float rr[3] = {r.x, r.y, r.z}, uu[3] = {u.x, u.y, u.z}, ll[3] = {f.x, f.y, f.z};
    return pt_camera_corners_basis(rr, uu, ll, fovy_deg * 3.14159265358979f / 180.0f, aspect, jx, jy, 1, dirs);
}

int pt_pass_jitter(uint64_t seed, uint32_t pass, int w, int h, float* jx, float* jy) {
    if (!jx || !jy || w <= 0 || h <= 0) return PT_E_INVALID_ARG;
    SplitMix64 rng(mix64(seed, pass));
    float wi = 2.0f / static_cast<float>(w);
    float hi = 2.0f / static_cast<float>(h);
    *jx = -wi + 2.0f * wi * rng.u01();
    *jy = -hi + 2.0f * hi * rng.u01();
    return 0;
}

}  // extern "C"
