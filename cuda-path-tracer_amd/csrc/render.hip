// Progressive unidirectional path tracer as a wavefront pipeline
// (reference: raytrace_kernel + blinnPhongShadingPathTrace,
// cuda_path_tracer.cu:245-333, 385-421, which run one megakernel thread per
// pixel and one launch per 1-spp pass).  One frame renders P passes at once
// (path = pass * pixels + pixel) so small images still fill the GPU:
//   k_generate     primary rays of every path into queue 0, counters reset
//   k_extend(L)    persistent closest hit for every ray of queue L
//   k_shade(L)     shade the hits, append surviving paths to queue L+1
//                  (workgroup-aggregated atomics), L = 0 .. bounces
//   k_accumulate   per pixel, FP64 sums += each pass's radiance, in pass order
// Queue lengths live on the device, so the host never waits inside a frame,
// and a frame's kernel sequence is captured once as a hipGraph and replayed.
//
// Cache frames (learned radiance cache, no reference counterpart): every
// pixel's path stops at its second vertex x2 (after the first bounce) and adds
// throughput * cache(x2); a set of training paths, started on 1/16 of the
// pixels, records x2's features and extends past it to give the cache's
// target there (the radiance the suffix brings back, per unit throughput).
// Per-pass values (corner rays, eye, pass index, seed) are read from device
// memory, so replays only need a fresh copy of that small block.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <type_traits>

#include "context.hpp"
#include "sampling.hpp"
#include "traverse.hpp"

namespace pt {
namespace {

// Fixed for a captured frame.
struct FrameConst {
    int w, h;
    uint32_t npix;
    uint32_t npasses;
    uint32_t bounces;
    uint32_t rr_start;
    float intensity;
    uint32_t stratified;
    uint32_t nlights;  // directional lights (shadow rays at every vertex)
    uint32_t ntex;     // textures of the surface detail
    const pt_triangle_detail* detail;  // per original id, or null: pt_triangle_info alone
    const uint32_t* texels;            // RGBA8 texels of all textures
    const uint4* texinfo;              // per texture: first texel, width, height, 0
    uint32_t* status;                  // the scene's status word (PT_STATUS_* flags)
};

// Structure-of-arrays queue of rays (origin, direction, path index).
struct RayQueue {
    float *ox, *oy, *oz, *dx, *dy, *dz;
    uint32_t* path;
};
struct HitQueue {
    float* t;
    uint32_t* id;
    float *u, *v;
};
struct PathState {
    float *ar, *ag, *ab;  // throughput ("absorbed", cuda_path_tracer.cu:259)
    float *cr, *cg, *cb;  // radiance ("accumulatedColor")
};

// Cache-frame buffers: queries of the pixel paths (slot = pixel) and the
// training batch (slot = training path).  Raw samples are 12 floats:
// position, shading normal, incoming direction, albedo.
struct CacheIO {
    float* qraw;        // [npix][12]
    float* qout;        // [npix][3] cache predictions
    uint32_t* qvalid;   // path reached x2
    float* traw;        // [ntrain][12]
    float* ttgt;        // [ntrain][3] targets (NaN: masked)
    uint32_t* tvalid;
    uint32_t ntrain;        // training paths
    uint32_t train_levels;  // queue levels of a training path (2 + extra bounces)
};

// Training paths draw from their own random stream.
constexpr uint32_t kTrainSalt = 0x5bd1e995u;

constexpr int kThreads = 256;
// k_shade workgroups: its queue append is one atomic per workgroup pass, and
// one counter address sustains only ~90 M returning atomics/s, so a wide
// workgroup (16 waves) keeps a 1M-path level's appends at ~1K atomics.
constexpr int kShadeThreads = 1024;
// Counter slots a cache frame's training paths report their first two
// levels in (their later levels go to the level counters): the frame's own
// levels 0 and 1 count pixel paths only, and run at the same time.
constexpr uint32_t kTrainLevel0Slot = 62, kTrainLevel1Slot = 63;
constexpr uint32_t kMaxLevels = 64;  // ray queue blocks: at most 61 path levels (60 bounces)
// A plain frame's paths may be split into up to kMaxChains chains: contiguous
// ranges of the first level's slots, each traced level by level on a stream
// of its own, with its own queue lengths (words 0..63 of the chain's counter
// block) and ray queue blocks (the kMaxLevels blocks after them).  Every level
// ends with its longest rays on a few lanes; with several chains one chain's
// tail overlaps the other chains' work.  A chain keeps its paths in its own
// slot range at every level (paths only drop out), so the queues are shared.
constexpr uint32_t kMaxChains = 8;
static_assert(sizeof(Renderer::chain_st) / sizeof(hipStream_t) == kMaxChains - 1, "chain streams");
constexpr size_t kChainWords = 64 + static_cast<size_t>(kMaxLevels) * kQueueWords;
constexpr uint32_t kMinChainPaths = 16384;  // smaller chains are merged

// First level-0 slot of chain c of n (chain c holds [chain_begin(c), chain_begin(c + 1))).
__host__ __device__ __forceinline__ uint32_t chain_begin(uint32_t npaths, uint32_t c, uint32_t n) {
    return static_cast<uint32_t>((static_cast<uint64_t>(npaths) * c) / n);
}

// Queue slot -> pixel for primary rays: 8x8 tiles in bands of 8 rows (the
// last band and the last tile of a band may be smaller), so the 64 lanes of
// a wave start on one tile -- coherent rays -- rather than on a 64-pixel row
// strip.  A bijection of [0, w*h); paths keep their pixel as identity, so
// the order changes no result.
__device__ __forceinline__ uint32_t tiled_pixel(uint32_t i, uint32_t w, uint32_t h) {
    const uint32_t y0 = 8u * (i / (8u * w));
    const uint32_t hb = min(8u, h - y0);
    const uint32_t r = i - y0 * w;
    const uint32_t tx = r / (8u * hb);
    const uint32_t c = r - tx * 8u * hb;
    const uint32_t tw = min(8u, w - 8u * tx);
    return (y0 + c / tw) * w + 8u * tx + c % tw;
}

// counters: the first of `nchains` chain counter blocks (kChainWords words
// apart; the grid has at least nchains workgroups).
__global__ __launch_bounds__(kThreads) void k_generate(const FrameConst fc,
                                                       const PassParams* __restrict__ pp,
                                                       RayQueue q, PathState ps,
                                                       uint32_t* __restrict__ counters,
                                                       uint32_t nlevels, uint32_t nchains) {
    const uint32_t npaths = fc.npix * fc.npasses;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid == 0) {
        for (uint32_t c = 0; c < nchains; ++c) {
            uint32_t* cc = counters + c * kChainWords;
            cc[0] = chain_begin(npaths, c + 1, nchains) - chain_begin(npaths, c, nchains);
            for (uint32_t k = 1; k <= nlevels; ++k) cc[k] = 0u;
        }
        counters[kTrainLevel0Slot] = 0u;
        counters[kTrainLevel1Slot] = 0u;
    }
    if (blockIdx.x < nchains) zero_queue(counters + blockIdx.x * kChainWords + 64, nlevels);
    for (uint32_t slot = gid; slot < npaths; slot += gridDim.x * blockDim.x) {
        const uint32_t pass = slot / fc.npix;
        const PassParams& P = pp[pass];
        const uint32_t pix = tiled_pixel(slot - pass * fc.npix, static_cast<uint32_t>(fc.w),
                                         static_cast<uint32_t>(fc.h));
        const uint32_t p = pass * fc.npix + pix;
        const int x = static_cast<int>(pix % static_cast<uint32_t>(fc.w));
        const int y = static_cast<int>(pix / static_cast<uint32_t>(fc.w));
        // Pixel-corner blend of the corner directions (cuda_path_tracer.cu:404-407).
        const float tx = static_cast<float>(x) / static_cast<float>(fc.w);
        const float ty = static_cast<float>(y) / static_cast<float>(fc.h);
        const V3 d00 = v3(P.dirs[0], P.dirs[1], P.dirs[2]);
        const V3 d10 = v3(P.dirs[3], P.dirs[4], P.dirs[5]);
        const V3 d11 = v3(P.dirs[6], P.dirs[7], P.dirs[8]);
        const V3 d01 = v3(P.dirs[9], P.dirs[10], P.dirs[11]);
        const V3 d = mix(mix(d00, d10, tx), mix(d01, d11, tx), ty);
        q.ox[slot] = P.eye[0];
        q.oy[slot] = P.eye[1];
        q.oz[slot] = P.eye[2];
        q.dx[slot] = d.x;
        q.dy[slot] = d.y;
        q.dz[slot] = d.z;
        q.path[slot] = p;
        ps.cr[p] = 0.0f;
        ps.cg[p] = 0.0f;
        ps.cb[p] = 0.0f;
    }
}

// Closest hit for every ray of queue `level` (count on the device): the
// persistent traversal, rays handed out from the level's ray queue block.
// Each path traces the scene of its pass (root internal node off[scene]).
template <int LDS>
__global__ __launch_bounds__(kPersistThreads) void k_extend(const FrameConst fc,
                                                            const PassParams* __restrict__ pp,
                                                            const float4* __restrict__ nodes,
                                                            const float4* __restrict__ tris,
                                                            const uint32_t* __restrict__ off,
                                                            RayQueue q, HitQueue hq,
                                                            const uint32_t* __restrict__ counters,
                                                            uint32_t* __restrict__ ray_counters,
                                                            uint32_t level, uint32_t refill, uint32_t pool) {
    __shared__ uint32_t s_stack[LDS][kPersistThreads];
    const uint32_t count = counters[level];
    TraceCount cnt{0, 0, 0, 0};
    auto load = [&](uint32_t i, V3& o, V3& d, uint32_t& single) -> uint32_t {
        o = v3(q.ox[i], q.oy[i], q.oz[i]);
        d = v3(q.dx[i], q.dy[i], q.dz[i]);
        const uint32_t sc = pp[min(q.path[i] / fc.npix, fc.npasses - 1u)].scene;
        const uint32_t b = off[sc], e = off[sc + 1];
        if (e - b >= 2) return b;
        if (e - b == 1) single = b;
        return kInvalid;
    };
    auto store = [&](uint32_t i, const TraceHit& h) {
        hq.t[i] = h.t;
        hq.id[i] = h.id;
        hq.u[i] = h.u;
        hq.v[i] = h.v;
    };
    persistent_trace<false, LDS>(nodes, tris, count, ray_counters + level * kQueueWords, fc.status, refill, pool,
                                 s_stack, load, store, cnt);
}

__device__ __forceinline__ V3 ld3(const float* p) { return v3(p[0], p[1], p[2]); }

// Light reaching `pos` from a directional light (computeL, LightAlgorithms.h:
// 12-20): a shadow ray from pos - dir * 1e-3 towards the light; any hit
// (closest-hit t < FLT_MAX, trace.h:45-50) shadows it.  The any-hit search
// answers the same question and stops at its first hit.
template <bool LDSS = false>
__device__ __forceinline__ V3 light_L(const float4* __restrict__ nodes, const float4* __restrict__ tris,
                                      uint32_t b, uint32_t e, V3 pos, const pt_dir_light& lt, uint32_t* status,
                                      const ShadowStack& ss = ShadowStack{}) {
    const V3 ld = ld3(lt.direction);
    if (any_hit_in<LDSS>(nodes, tris, b, e, sub(pos, scale(ld, 1e-3f)), neg(ld), status, ss))
        return v3(0.0f, 0.0f, 0.0f);
    return scale(ld3(lt.color), lt.intensity);
}

// blinnPhongBRDF (cuda_path_tracer.cu:36-50).
__device__ __forceinline__ V3 blinn_phong(V3 Id, V3 kd, V3 Is, V3 ks, V3 n, V3 l, V3 hw, float shininess) {
    const float df = gmax(dot(n, l), 0.0f);
    const V3 diffuse = scale(mul(Id, kd), df);
    V3 specular = v3(0.0f, 0.0f, 0.0f);
    if (shininess > 0.0f && df > 0.0f) specular = scale(mul(Is, ks), pow_pos(gmax(dot(n, hw), 0.0f), shininess));
    return add(diffuse, specular);
}

// Shading data of a hit (cuda_path_tracer.cu:190-191, 298-300): position,
// interpolated normal, albedo, emission.
struct HitShade {
    V3 pos, nrm, albedo, e;
    float spec;
    bool emissive;
};

// barycentricLerp (math_util.h:19-25): v0 * (1 - u - v) + v1 * u + v2 * v.
__device__ __forceinline__ V3 bary3(const float (*v)[3], float bw, float bu, float bv) {
    return add(add(scale(ld3(v[0]), bw), scale(ld3(v[1]), bu)), scale(ld3(v[2]), bv));
}

// One texel as c / 255 per channel (rgb).
__device__ __forceinline__ V3 texel(const uint32_t* __restrict__ texels, uint32_t k) {
    const uint32_t c = texels[k];
    return v3(static_cast<float>(c & 0xFFu) / 255.0f, static_cast<float>((c >> 8) & 0xFFu) / 255.0f,
              static_cast<float>((c >> 16) & 0xFFu) / 255.0f);
}

// tex2D on a normalised-coordinate, wrapping, bilinear texture at level 0
// (Scene.cu:66-73): x = s * width - 0.5, y = t * height - 0.5, the four
// texels around (x, y) blended with fp32 weights, glm::mix order
// (p + w * (q - p)), first along x then along y.  Coordinates that are not
// finite or beyond 1e9 texels read as 0.
__device__ __forceinline__ V3 tex_sample(const uint32_t* __restrict__ texels, uint4 info, float s, float t) {
    const uint32_t W = info.y, H = info.z;
    float x = s * static_cast<float>(W) - 0.5f;
    float y = t * static_cast<float>(H) - 0.5f;
    if (!(__builtin_fabsf(x) < 1e9f)) x = 0.0f;
    if (!(__builtin_fabsf(y) < 1e9f)) y = 0.0f;
    const float fx = __builtin_floorf(x), fy = __builtin_floorf(y);
    const float a = x - fx, b = y - fy;
    int i0 = static_cast<int>(fx) % static_cast<int>(W), j0 = static_cast<int>(fy) % static_cast<int>(H);
    if (i0 < 0) i0 += static_cast<int>(W);
    if (j0 < 0) j0 += static_cast<int>(H);
    const uint32_t i1 = static_cast<uint32_t>(i0) + 1u == W ? 0u : static_cast<uint32_t>(i0) + 1u;
    const uint32_t j1 = static_cast<uint32_t>(j0) + 1u == H ? 0u : static_cast<uint32_t>(j0) + 1u;
    const uint32_t r0 = info.x + static_cast<uint32_t>(j0) * W, r1 = info.x + j1 * W;
    const V3 lo = mix(texel(texels, r0 + static_cast<uint32_t>(i0)), texel(texels, r0 + i1), a);
    const V3 hi = mix(texel(texels, r1 + static_cast<uint32_t>(i0)), texel(texels, r1 + i1), a);
    return mix(lo, hi, b);
}

// DETAIL: the scene has surface detail (fc.detail); kept out of the other
// kernels, whose registers it would otherwise size.
template <bool DETAIL>
__device__ __forceinline__ HitShade hit_shade(const FrameConst& fc, const pt_triangle_info* __restrict__ infos,
                                              uint32_t id, float t, float bu, float bv, V3 o, V3 d) {
    const pt_triangle_info& in = infos[id];
    HitShade hs;
    hs.pos = add(scale(d, t), o);
    const float bw = (1.0f - bu) - bv;
    hs.nrm = normalize(add(add(scale(ld3(in.n0), bw), scale(ld3(in.n1), bu)), scale(ld3(in.n2), bv)));
    if (DETAIL) {
        // extractInfo / unpackNormal (cuda_path_tracer.cu:85-138).
        const pt_triangle_detail& dt = fc.detail[id];
        const float s = ((dt.uv[0][0] * bw) + (dt.uv[1][0] * bu)) + (dt.uv[2][0] * bv);
        const float tc = ((dt.uv[0][1] * bw) + (dt.uv[1][1] * bu)) + (dt.uv[2][1] * bv);
        hs.albedo = dt.albedo_tex != PT_NO_TEXTURE ? tex_sample(fc.texels, fc.texinfo[dt.albedo_tex], s, tc)
                                                   : normalize(bary3(dt.color, bw, bu, bv));
        if (dt.normal_tex != PT_NO_TEXTURE) {
            const V3 ns = tex_sample(fc.texels, fc.texinfo[dt.normal_tex], s, tc);
            const float nx = 2.0f * ns.x - 1.0f, ny = 2.0f * ns.y - 1.0f, nz = 2.0f * ns.z - 1.0f;
            const V3 tg = normalize(bary3(dt.tangent, bw, bu, bv));
            const V3 bt = normalize(bary3(dt.bitangent, bw, bu, bv));
            hs.nrm = normalize(add(add(scale(tg, nx), scale(bt, ny)), scale(hs.nrm, nz)));
        }
    } else {
        hs.albedo = ld3(in.albedo);
    }
    hs.e = ld3(in.emission);
    hs.spec = in.specularity;
    hs.emissive = ((hs.e.x + hs.e.y) + hs.e.z) > 1e-7f;
    return hs;
}

// Bounce hit (level >= 1, cuda_path_tracer.cu:298-325): emitter cosine term
// (albedo 1 on emitters) or the directional lights' term behind shadow rays;
// throughput *= albedo, radiance += throughput * L * intensity.  LIGHTS:
// whether the scene has directional lights (kept out of the no-light kernels,
// whose registers and stack it would otherwise size).
template <bool LIGHTS, bool LDSS = false>
__device__ __forceinline__ void bounce_hit(HitShade& hs, V3 d, const FrameConst& fc,
                                           const pt_dir_light* __restrict__ lights,
                                           const float4* __restrict__ nodes, const float4* __restrict__ tris,
                                           uint32_t sb, uint32_t se, V3& absorbed, V3& acc,
                                           const ShadowStack& ss = ShadowStack{}) {
    V3 L = v3(0.0f, 0.0f, 0.0f);
    if (hs.emissive) {
        const float c = gmax(dot(neg(d), hs.nrm), 0.0f);
        L = scale(hs.e, c);
        hs.albedo = v3(1.0f, 1.0f, 1.0f);
    } else if (LIGHTS) {
        for (uint32_t k = 0; k < fc.nlights; ++k) {
            const V3 ld = ld3(lights[k].direction);
            L = add(L, scale(light_L<LDSS>(nodes, tris, sb, se, hs.pos, lights[k], fc.status, ss),
                             gmax(dot(neg(ld), hs.nrm), 0.0f)));
        }
    }
    absorbed = mul(absorbed, hs.albedo);
    acc = add(acc, scale(mul(absorbed, L), fc.intensity));
}

// Next segment of bounce b (cuda_path_tracer.cu:267-290): Russian roulette
// from rr_start (survive with p = max albedo, throughput / p), stratified
// first bounce, cosine-weighted direction about the shading normal.  Returns
// whether the path continues; no/nd are the new origin and direction.
__device__ __forceinline__ bool next_bounce(uint32_t key, uint32_t pass, uint32_t b, uint32_t seedw,
                                            const FrameConst& fc, const HitShade& hs, V3& absorbed, V3& no,
                                            V3& nd) {
    const U4 rv = pcg4d(U4{key, pass, b, seedw});
    float u1 = u24(rv.x), u2 = u24(rv.y);
    if (b >= fc.rr_start) {
        float p = gmax(gmax(hs.albedo.x, hs.albedo.y), hs.albedo.z);
        p = gmin(p, 1.0f);
        if (!(u24(rv.z) < p)) return false;
        absorbed = v3(absorbed.x / p, absorbed.y / p, absorbed.z / p);
    }
    no = add(hs.pos, scale(hs.nrm, 1e-5f));
    if (fc.stratified != 0u && b == 0) {
        // 8x8 strata, stratum = pass % 64 (cuda_path_tracer.cu:273-280).
        const uint32_t p1 = pass % 64u;
        u1 = (static_cast<float>(p1 % 8u) + u1) / 8.0f;
        u2 = (static_cast<float>(p1 / 8u) + u2) / 8.0f;
    }
    // Cosine-weighted hemisphere sample (random.h:95-102).
    const float r = __builtin_sqrtf(u1);
    float sn, cs;
    sincos2pi(u2, sn, cs);
    const V3 local = v3(r * cs, r * sn, __builtin_sqrtf(1.0f - u1));
    nd = to_hemisphere(local, hs.nrm);
    return true;
}

// Primary hit (level 0): emission, else Blinn-Phong from the directional
// lights with shadow rays, the light term L summed over the lights seen so
// far (cuda_path_tracer.cu:194-214, 258-259); the path's throughput starts
// at the hit's albedo (1 on emitters).
template <bool LIGHTS, bool LDSS = false>
__device__ __forceinline__ void primary_hit(HitShade& hs, V3 d, const FrameConst& fc,
                                            const pt_dir_light* __restrict__ lights,
                                            const float4* __restrict__ nodes, const float4* __restrict__ tris,
                                            uint32_t sb, uint32_t se, V3& absorbed, V3& acc,
                                            const ShadowStack& ss = ShadowStack{}) {
    acc = v3(0.0f, 0.0f, 0.0f);
    if (hs.emissive) {
        acc = hs.e;
        hs.albedo = v3(1.0f, 1.0f, 1.0f);
    } else if (LIGHTS) {
        V3 Lc = v3(0.0f, 0.0f, 0.0f);
        for (uint32_t k = 0; k < fc.nlights; ++k) {
            const V3 ld = ld3(lights[k].direction);
            Lc = add(Lc, light_L<LDSS>(nodes, tris, sb, se, hs.pos, lights[k], fc.status, ss));
            const V3 hw = scale(sub(neg(d), ld), 0.5f);
            acc = add(acc, blinn_phong(Lc, hs.albedo, Lc, v3(0.5f, 0.5f, 0.5f), hs.nrm, neg(ld), hw, hs.spec));
        }
    }
    absorbed = hs.albedo;
}

// A cache frame's second vertex x2: a pixel path's cache query, or a
// training sample whose suffix radiance is gathered from here with
// throughput 1.
__device__ __forceinline__ void record_cache_vertex(const CacheIO& io, bool train, uint32_t pix, uint32_t path,
                                                    const HitShade& hs, V3 d, V3& absorbed, V3& acc) {
    float* f = train ? io.traw + 12 * static_cast<size_t>(pix) : io.qraw + 12 * static_cast<size_t>(path);
    f[0] = hs.pos.x; f[1] = hs.pos.y; f[2] = hs.pos.z;
    f[3] = hs.nrm.x; f[4] = hs.nrm.y; f[5] = hs.nrm.z;
    f[6] = d.x;      f[7] = d.y;      f[8] = d.z;
    f[9] = hs.albedo.x; f[10] = hs.albedo.y; f[11] = hs.albedo.z;
    if (train) {
        io.tvalid[pix] = 1u;
        absorbed = v3(1.0f, 1.0f, 1.0f);
        acc = v3(0.0f, 0.0f, 0.0f);
    } else {
        io.qvalid[path] = 1u;
    }
}

// CACHE: a cache frame (one pass; paths >= npix are training paths).
// LIGHTS: the scene has directional lights.
template <bool CACHE, bool LIGHTS, bool DETAIL>
__global__ __launch_bounds__(kShadeThreads) void k_shade(const FrameConst fc,
                                                    const PassParams* __restrict__ pp,
                                                    const pt_triangle_info* __restrict__ infos,
                                                    RayQueue qin, HitQueue hq, RayQueue qout,
                                                    PathState ps, uint32_t* __restrict__ counters,
                                                    uint32_t level, CacheIO io,
                                                    const pt_dir_light* __restrict__ lights,
                                                    const float4* __restrict__ nodes,
                                                    const float4* __restrict__ tris,
                                                    const uint32_t* __restrict__ off) {
    __shared__ uint32_t s_off[kShadeThreads / 64];  // per-wave offsets in the workgroup's append
    __shared__ uint32_t s_base;                      // the workgroup's first slot in qout
    const uint32_t count = counters[level];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const uint64_t lt = (1ull << lane) - 1ull;
    const uint32_t stride = gridDim.x * blockDim.x;
    // Workgroup-uniform trip count so every lane reaches the ballots and
    // barriers below.
    for (uint32_t base = blockIdx.x * blockDim.x; base < count; base += stride) {
        const uint32_t i = base + threadIdx.x;
        bool cont = false;
        V3 no{0.0f, 0.0f, 0.0f}, nd{0.0f, 0.0f, 0.0f};
        uint32_t path = 0;
        if (i < count && hq.id[i] != kInvalid) {
            path = qin.path[i];
            const bool train = CACHE && path >= fc.npix;
            const PassParams& P = pp[CACHE ? 0u : path / fc.npix];
            // RNG key: pixel (training paths: their index, own stream).
            const uint32_t pix = train ? path - fc.npix : path % fc.npix;
            const uint32_t seedw = train ? (P.seedw ^ kTrainSalt) : P.seedw;
            const V3 o = v3(qin.ox[i], qin.oy[i], qin.oz[i]);
            const V3 d = v3(qin.dx[i], qin.dy[i], qin.dz[i]);
            HitShade hs = hit_shade<DETAIL>(fc, infos, hq.id[i], hq.t[i], hq.u[i], hq.v[i], o, d);
            const uint32_t sb = off[P.scene], se = off[P.scene + 1];
            V3 absorbed, acc;
            if (level == 0) {
                primary_hit<LIGHTS>(hs, d, fc, lights, nodes, tris, sb, se, absorbed, acc);
            } else {
                absorbed = v3(ps.ar[path], ps.ag[path], ps.ab[path]);
                acc = v3(ps.cr[path], ps.cg[path], ps.cb[path]);
                bounce_hit<LIGHTS>(hs, d, fc, lights, nodes, tris, sb, se, absorbed, acc);
            }
            if (CACHE && level == 1) record_cache_vertex(io, train, pix, path, hs, d, absorbed, acc);
            // Next bounce index = level (bounce loop, cuda_path_tracer.cu:267).
            // Cache frames: pixel paths end at x2, training paths after
            // train_levels - 1 bounces.
            const uint32_t b = level;
            const uint32_t limit = CACHE ? (train ? io.train_levels - 1u : 1u) : fc.bounces;
            if (b < limit) cont = next_bounce(pix, P.pass, b, seedw, fc, hs, absorbed, no, nd);
            ps.ar[path] = absorbed.x;
            ps.ag[path] = absorbed.y;
            ps.ab[path] = absorbed.z;
            ps.cr[path] = acc.x;
            ps.cg[path] = acc.y;
            ps.cb[path] = acc.z;
        }
        // Append the continuing paths to qout: one atomic per workgroup,
        // waves in order inside it, lanes in order inside a wave.
        const uint64_t m = __ballot(cont);
        if (lane == 0) s_off[wave] = static_cast<uint32_t>(__popcll(m));
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t total = 0;
            for (uint32_t k = 0; k < kShadeThreads / 64; ++k) {
                const uint32_t c = s_off[k];
                s_off[k] = total;
                total += c;
            }
            s_base = total ? atomicAdd(counters + level + 1, total) : 0u;
        }
        __syncthreads();
        // (The next pass rewrites s_off and s_base only after its first
        // barrier, which every wave reaches after its reads here.)
        if (cont) {
            const uint32_t s = s_base + s_off[wave] + static_cast<uint32_t>(__popcll(m & lt));
            qout.ox[s] = no.x;
            qout.oy[s] = no.y;
            qout.oz[s] = no.z;
            qout.dx[s] = nd.x;
            qout.dy[s] = nd.y;
            qout.dz[s] = nd.z;
            qout.path[s] = path;
        }
    }
}

// Hand-out of an index range in pools (persistent_trace's protocol).
//
// Wave-uniform state over [0, n): pools of `pool` indices from slice
// counters of a queue block (see persistent_trace).  next_pool() leaves a
// non-empty pool in [next, end) or sets `drained` once every slice is used
// up; each call either takes a pool or adds a used-up slice to `seen`, so
// it returns after at most nq + 1 atomics.
struct Handout {
    uint32_t* queue;
    uint32_t nq, per, extra, pool, q;
    uint64_t seen;
    uint32_t next, end;
    bool drained;

    __device__ __forceinline__ void init(uint32_t* qb, uint32_t n, uint32_t pool_size, uint32_t wave_id) {
        queue = qb;
        nq = queue_slices(n);
        per = n / nq;
        extra = n % nq;
        pool = pool_size;
        q = wave_id % nq;
        const uint64_t all = nq == 64u ? ~0ull : (1ull << nq) - 1ull;
        seen = ~all;
        next = end = 0;
        drained = n == 0;
    }

    __device__ __forceinline__ void next_pool() {
        const int lane = threadIdx.x & 63;
        uint64_t* mask = reinterpret_cast<uint64_t*>(queue + kQueueMask);
        while (!drained && next >= end) {
            const uint32_t beg = q * per + min(q, extra);
            const uint32_t len = per + (q < extra ? 1u : 0u);
            uint32_t base = 0;
            if (lane == 0) {
                base = atomicAdd(queue + q * kQueueStride, pool);
                if (base < len && len - base <= pool) atomicOr(mask, 1ull << q);
            }
            base = __builtin_amdgcn_readfirstlane(base);
            if (base < len) {
                next = beg + base;
                end = beg + base + min(pool, len - base);
                return;
            }
            seen |= 1ull << q;
            uint32_t lo = 0, hi = 0;
            if (lane == 0) {
                const uint64_t m = __hip_atomic_load(mask, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                lo = static_cast<uint32_t>(m);
                hi = static_cast<uint32_t>(m >> 32);
            }
            lo = __builtin_amdgcn_readfirstlane(lo);
            hi = __builtin_amdgcn_readfirstlane(hi);
            const uint64_t open = ~(seen | (static_cast<uint64_t>(hi) << 32) | lo);
            if (open == 0ull) {
                drained = true;
            } else {
                const uint64_t above = q < 63u ? open & (~0ull << (q + 1u)) : 0ull;
                q = static_cast<uint32_t>(__builtin_ctzll(above ? above : open));
            }
        }
    }
};

// Path-persistent frame (render.paths): one launch traces every path of a
// frame from its primary ray to its end, one lane per path at a time.  A
// lane advances its path's current segment by one node visit per loop
// iteration (persistent_trace's visit); a lane whose segment has ended waits
// until `shade_min` lanes of its wave wait (or none is traversing), then those
// lanes shade together -- primary_hit / bounce_hit / next_bounce, the
// wavefront's functions, so every path computes the same bits -- and either
// start their next segment or end the path (radiance into the path state for
// k_accumulate).  Lanes without a path take new ones (pools of path slots,
// persistent_trace's hand-out) when fewer than `refill` are busy.  There are
// no bounce levels, so no frame-wide wait for each level's longest ray: a
// frame waits once, for its longest path.  Segments per bounce are counted
// per workgroup in LDS and added to counters[b] at the end.
enum : uint32_t { kLaneIdle = 0u, kLaneTrace = 1u, kLaneShade = 2u };

__global__ __launch_bounds__(64) void k_paths_setup(uint32_t* __restrict__ counters, uint32_t* __restrict__ queue,
                                                    uint32_t nlevels) {
    for (uint32_t k = threadIdx.x; k <= nlevels && k < kTrainLevel0Slot; k += blockDim.x) counters[k] = 0u;
    if (threadIdx.x == 0) {
        counters[kTrainLevel0Slot] = 0u;
        counters[kTrainLevel1Slot] = 0u;
    }
    zero_queue(queue, 1u);
}

// At most 80 VGPRs (6 waves per SIMD): the plain variant fits without
// spilling, the surface-detail variants too (109 VGPRs uncapped), and the lit
// variants spill a little in their shading, which costs no more than the
// waves it buys.
// CACHE: the pixel paths of a cache frame (one pass, pixel = path), each
// ending in the cache at its second hit (record_cache_vertex) with its
// throughput in ps.ar/ag/ab -- the same values the levels k_extend /
// k_shade<true> leave.
template <int LDS, bool LIGHTS, bool DETAIL, bool CACHE>
__global__ __launch_bounds__(kPersistThreads, 6) void k_paths(const FrameConst fc, const PassParams* __restrict__ pp,
                                                           const pt_triangle_info* __restrict__ infos,
                                                           const float4* __restrict__ nodes,
                                                           const float4* __restrict__ tris,
                                                           const uint32_t* __restrict__ off,
                                                           const pt_dir_light* __restrict__ lights, PathState ps,
                                                           uint32_t* __restrict__ counters,
                                                           uint32_t* __restrict__ queue, uint32_t refill,
                                                           uint32_t pool_large, uint32_t shade_min, CacheIO io) {
    __shared__ uint32_t s_stack[LDS][kPersistThreads];
    __shared__ uint32_t s_level[kMaxLevels];
    // Plain frames keep each lane's path throughput and radiance in LDS
    // while it traverses: the registers saved take the kernel to 6 waves per
    // SIMD.  (With lights or surface detail the kernel stays at 4 waves per
    // SIMD either way, and the LDS round trips only cost.)
    constexpr bool kCarryLds = !LIGHTS && !DETAIL;
    __shared__ float s_carry[kCarryLds ? 6 : 1][kPersistThreads];
    for (uint32_t k = threadIdx.x; k < kMaxLevels; k += blockDim.x) s_level[k] = 0u;
    __syncthreads();
    const uint32_t npaths = fc.npix * fc.npasses;
    const uint32_t wave_id = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / 64u) + threadIdx.x / 64u);
    const int lane = threadIdx.x & 63;
    // Waves past one per kRayPoolSmall paths skip the loop (but reach the
    // final barrier).
    if (wave_id == 0u || static_cast<uint64_t>(wave_id) * kRayPoolSmall < npaths) {
        uint32_t spill[kStackMax - LDS];
        TraceCount cnt{0, 0, 0, 0};
        Handout ho;
        ho.init(queue, npaths, npaths / 4u >= gridDim.x * blockDim.x ? pool_large : kRayPoolSmall, wave_id);
        RayState r;
        uint32_t state = kLaneIdle;
        uint32_t path = 0, bounce = 0;  // the path's pass and pixel follow from its index
        uint32_t root = kInvalid;
        V3 absorbed_r = v3(1.0f, 1.0f, 1.0f), acc_r = v3(0.0f, 0.0f, 0.0f);
        const int tid = threadIdx.x;
        const uint32_t w = static_cast<uint32_t>(fc.w), h = static_cast<uint32_t>(fc.h);
        for (;;) {
            // New paths for idle lanes.
            uint64_t busy = __ballot(state != kLaneIdle);
            if (!ho.drained && static_cast<uint32_t>(__popcll(busy)) < refill) {
                uint64_t idle = ~busy;
                while (idle != 0ull) {
                    ho.next_pool();
                    if (ho.drained) break;
                    const uint32_t avail = ho.end - ho.next;
                    const uint32_t rank = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(idle >> 32),
                                                                    __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(idle), 0u));
                    const bool take = ((idle >> lane) & 1ull) && rank < avail;
                    const uint32_t took = min(static_cast<uint32_t>(__popcll(idle)), avail);
                    if (take) {
                        const uint32_t slot = ho.next + rank;
                        const uint32_t pass = slot / fc.npix;
                        const uint32_t pix = tiled_pixel(slot - pass * fc.npix, w, h);
                        path = pass * fc.npix + pix;
                        const PassParams& P = pp[pass];
                        const float tx = static_cast<float>(pix % w) / static_cast<float>(fc.w);
                        const float ty = static_cast<float>(pix / w) / static_cast<float>(fc.h);
                        const V3 d00 = v3(P.dirs[0], P.dirs[1], P.dirs[2]);
                        const V3 d10 = v3(P.dirs[3], P.dirs[4], P.dirs[5]);
                        const V3 d11 = v3(P.dirs[6], P.dirs[7], P.dirs[8]);
                        const V3 d01 = v3(P.dirs[9], P.dirs[10], P.dirs[11]);
                        r.o = v3(P.eye[0], P.eye[1], P.eye[2]);
                        r.d = mix(mix(d00, d10, tx), mix(d01, d11, tx), ty);
                        const uint32_t sb = off[P.scene], se = off[P.scene + 1];
                        root = se - sb >= 2 ? sb : kInvalid;
                        bounce = 0;
                        if constexpr (kCarryLds) {
                            s_carry[0][tid] = 1.0f; s_carry[1][tid] = 1.0f; s_carry[2][tid] = 1.0f;
                            s_carry[3][tid] = 0.0f; s_carry[4][tid] = 0.0f; s_carry[5][tid] = 0.0f;
                        } else {
                            absorbed_r = v3(1.0f, 1.0f, 1.0f);
                            acc_r = v3(0.0f, 0.0f, 0.0f);
                        }
                        init_ray(r, root);
                        state = kLaneTrace;
                    }
                    ho.next += took;
                    idle = ~__ballot(state != kLaneIdle);
                }
            }
            busy = __ballot(state != kLaneIdle);
            if (busy == 0ull && ho.drained) break;
            // Count the segments that start this iteration (bounce of each).
            // (Counted when a lane enters kLaneTrace: at take and at continue.)
            // One node visit for every traversing lane; a scene with fewer
            // than two triangles answers at once.
            if (state == kLaneTrace) {
                if (root == kInvalid) {
                    r.minT = kMinTInit;
                    r.best = kInvalid;
                    // A scene of fewer than two triangles: its one triangle, if any.
                    const uint32_t psc = pp[path / fc.npix].scene;
                    const uint32_t single = off[psc + 1] - off[psc] == 1u ? off[psc] : kInvalid;
                    if (single != kInvalid) {
                        float u = 0.0f, v = 0.0f;
                        uint32_t id;
                        const float t = tri_hit(tris, single, r.o, r.d, u, v, id);
                        if (t <= kMinTInit) {
                            r.minT = t; r.best = id; r.bu = u; r.bv = v;
                        }
                    }
                    state = kLaneShade;
                } else if (!(r.finite ? visit_step<false, LDS, true>(nodes, tris, r, s_stack, spill, cnt, fc.status)
                                      : visit_step<false, LDS, false>(nodes, tris, r, s_stack, spill, cnt, fc.status))) {
                    state = kLaneShade;
                }
            }
            const uint64_t shading = __ballot(state == kLaneShade);
            const uint64_t tracing = __ballot(state == kLaneTrace);
            if (shading != 0ull && (static_cast<uint32_t>(__popcll(shading)) >= shade_min || tracing == 0ull)) {
                if (state == kLaneShade) {
                    atomicAdd(&s_level[bounce], 1u);  // this segment, now complete
                    bool cont = false;
                    V3 absorbed = absorbed_r, acc = acc_r;
                    if constexpr (kCarryLds) {
                        absorbed = v3(s_carry[0][tid], s_carry[1][tid], s_carry[2][tid]);
                        acc = v3(s_carry[3][tid], s_carry[4][tid], s_carry[5][tid]);
                    }
                    if (r.best != kInvalid) {
                        const uint32_t pass = path / fc.npix;
                        const uint32_t pix = path - pass * fc.npix;
                        const PassParams& P = pp[pass];
                        HitShade hs = hit_shade<DETAIL>(fc, infos, r.best, r.minT, r.bu, r.bv, r.o, r.d);
                        const uint32_t sb = off[P.scene], se = off[P.scene + 1];
                        // Shadow rays keep their deferred nodes in this lane's
                        // (now empty) traversal stack.
                        ShadowStack ss;
                        ss.lds = &s_stack[0][threadIdx.x];
                        ss.stride = kPersistThreads;
                        ss.depth = LDS;
                        ss.spill = spill;
                        if (bounce == 0) {
                            primary_hit<LIGHTS, true>(hs, r.d, fc, lights, nodes, tris, sb, se, absorbed, acc, ss);
                        } else {
                            bounce_hit<LIGHTS, true>(hs, r.d, fc, lights, nodes, tris, sb, se, absorbed, acc, ss);
                        }
                        if (CACHE && bounce == 1) record_cache_vertex(io, false, pix, path, hs, r.d, absorbed, acc);
                        V3 no{0.0f, 0.0f, 0.0f}, nd{0.0f, 0.0f, 0.0f};
                        const uint32_t limit = CACHE ? 1u : fc.bounces;
                        if (bounce < limit) cont = next_bounce(pix, P.pass, bounce, P.seedw, fc, hs, absorbed, no, nd);
                        if (cont) {
                            ++bounce;
                            r.o = no;
                            r.d = nd;
                            init_ray(r, root);
                            state = kLaneTrace;
                            if constexpr (kCarryLds) {
                                s_carry[0][tid] = absorbed.x; s_carry[1][tid] = absorbed.y; s_carry[2][tid] = absorbed.z;
                                s_carry[3][tid] = acc.x; s_carry[4][tid] = acc.y; s_carry[5][tid] = acc.z;
                            } else {
                                absorbed_r = absorbed;
                                acc_r = acc;
                            }
                        }
                    }
                    if (!cont) {
                        ps.cr[path] = acc.x;
                        ps.cg[path] = acc.y;
                        ps.cb[path] = acc.z;
                        if constexpr (CACHE) {  // the weight of the cache's answer (k_cache_resolve)
                            ps.ar[path] = absorbed.x;
                            ps.ag[path] = absorbed.y;
                            ps.ab[path] = absorbed.z;
                        }
                        state = kLaneIdle;
                    }
                }
            }
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < kMaxLevels && k < kTrainLevel0Slot; k += blockDim.x)
        if (s_level[k]) atomicAdd(counters + k, s_level[k]);
}

// Cache frames: the training paths, one lane per path from its primary ray
// to its end, on a stream of their own beside the pixel paths' two levels,
// the inference and the resolve (the per-level launches of a few thousand
// rays are latency bound -- each level waits for its slowest ray -- and one
// lane per path waits only for the slowest path).  Traversal, shading and
// sampling are k_extend / k_shade's (primary_hit, bounce_hit,
// record_cache_vertex, next_bounce), so the training batches are the same
// bits as when these paths ran through the levels.
constexpr int kSuffixThreads = 64;
constexpr int kTrainLds = 16;  // traversal stack entries per lane kept in LDS

template <bool DETAIL>
__global__ __launch_bounds__(kSuffixThreads) void k_train_paths(const FrameConst fc,
                                                                const PassParams* __restrict__ pp,
                                                                const pt_triangle_info* __restrict__ infos,
                                                                RayQueue q, PathState ps,
                                                                uint32_t* __restrict__ counters, CacheIO io,
                                                                const pt_dir_light* __restrict__ lights,
                                                                const float4* __restrict__ nodes,
                                                                const float4* __restrict__ tris,
                                                                const uint32_t* __restrict__ off) {
    __shared__ uint32_t s_tstack[kTrainLds][kSuffixThreads];
    uint32_t spill[kStackMax - kTrainLds];
    const PassParams& P = pp[0];
    const uint32_t seedw = P.seedw ^ kTrainSalt;
    const uint32_t sb = off[P.scene], se = off[P.scene + 1];
    for (uint32_t key = blockIdx.x * blockDim.x + threadIdx.x; key < io.ntrain; key += gridDim.x * blockDim.x) {
        const uint32_t path = fc.npix + key;  // its slot in the first level's queue
        V3 o = v3(q.ox[path], q.oy[path], q.oz[path]);
        V3 d = v3(q.dx[path], q.dy[path], q.dz[path]);
        V3 absorbed = v3(1.0f, 1.0f, 1.0f);
        V3 acc = v3(0.0f, 0.0f, 0.0f);
        atomicAdd(counters + kTrainLevel0Slot, 1u);
        for (uint32_t b = 0;; ++b) {
            TraceHit h;
            if (se - sb < 2) {
                h = trace_ray_in(nodes, tris, sb, se, o, d, fc.status);
            } else {
                // The reference-order walk with its first kTrainLds deferred
                // nodes in LDS (a lone wave per SIMD waits out every stack
                // read; scratch would make each pop a memory round trip).
                RayState r;
                r.o = o;
                r.d = d;
                init_ray(r, sb);
                TraceCount cnt{0, 0, 0, 0};
                if (r.finite) {
                    while (visit_step<false, kTrainLds, true>(nodes, tris, r, s_tstack, spill, cnt, fc.status)) {
                    }
                } else {
                    while (visit_step<false, kTrainLds, false>(nodes, tris, r, s_tstack, spill, cnt, fc.status)) {
                    }
                }
                h = final_hit(r);
            }
            if (h.id == kInvalid) break;
            HitShade hs = hit_shade<DETAIL>(fc, infos, h.id, h.t, h.u, h.v, o, d);
            if (b == 0) {
                if (fc.nlights)
                    primary_hit<true>(hs, d, fc, lights, nodes, tris, sb, se, absorbed, acc);
                else
                    primary_hit<false>(hs, d, fc, lights, nodes, tris, sb, se, absorbed, acc);
            } else if (fc.nlights) {
                bounce_hit<true>(hs, d, fc, lights, nodes, tris, sb, se, absorbed, acc);
            } else {
                bounce_hit<false>(hs, d, fc, lights, nodes, tris, sb, se, absorbed, acc);
            }
            if (b == 1) record_cache_vertex(io, true, key, path, hs, d, absorbed, acc);
            if (!(b < io.train_levels - 1u)) break;
            V3 no, nd;
            if (!next_bounce(key, P.pass, b, seedw, fc, hs, absorbed, no, nd)) break;
            atomicAdd(counters + (b == 0 ? kTrainLevel1Slot : b + 1), 1u);
            o = no;
            d = nd;
        }
        ps.ar[path] = absorbed.x;
        ps.ag[path] = absorbed.y;
        ps.ab[path] = absorbed.z;
        ps.cr[path] = acc.x;
        ps.cg[path] = acc.y;
        ps.cb[path] = acc.z;
    }
}

// FP64 accumulation, pass by pass in order -- the same additions as one
// accumulate per 1-spp pass.
__global__ __launch_bounds__(kThreads) void k_accumulate(const FrameConst fc, PathState ps,
                                                         double* __restrict__ accum) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < fc.npix; p += gridDim.x * blockDim.x) {
        double r = accum[3 * static_cast<size_t>(p) + 0];
        double g = accum[3 * static_cast<size_t>(p) + 1];
        double b = accum[3 * static_cast<size_t>(p) + 2];
        for (uint32_t k = 0; k < fc.npasses; ++k) {
            const size_t s = static_cast<size_t>(k) * fc.npix + p;
            r += static_cast<double>(ps.cr[s]);
            g += static_cast<double>(ps.cg[s]);
            b += static_cast<double>(ps.cb[s]);
        }
        accum[3 * static_cast<size_t>(p) + 0] = r;
        accum[3 * static_cast<size_t>(p) + 1] = g;
        accum[3 * static_cast<size_t>(p) + 2] = b;
    }
}

// Cache frame: pixel paths (0 .. npix) as k_generate with one pass, then
// training paths on one cell of every 4x4 pixel block, the cell (ox, oy)
// stepping through all 16 with the pass index; training path k takes block
// floor(k * nblocks / ntrain), so the paths spread evenly over the image.
__global__ __launch_bounds__(kThreads) void k_generate_cache(const FrameConst fc,
                                                             const PassParams* __restrict__ pp,
                                                             RayQueue q, PathState ps,
                                                             uint32_t* __restrict__ counters,
                                                             uint32_t* __restrict__ ray_counters,
                                                             uint32_t nlevels, CacheIO io, uint32_t count0) {
    const uint32_t npaths = fc.npix + io.ntrain;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid == 0) {
        // The pixel paths' first-level count: npix when the levels trace
        // them, 0 when a path-persistent launch does and counts its own
        // segments; the training paths' rays (slots npix ..) are read by
        // k_train_paths.
        counters[0] = count0;
        for (uint32_t k = 1; k <= nlevels; ++k) counters[k] = 0u;
        counters[kTrainLevel0Slot] = 0u;
        counters[kTrainLevel1Slot] = 0u;
    }
    if (blockIdx.x == 0) zero_queue(ray_counters, nlevels);
    const PassParams& P = pp[0];
    const uint32_t w = static_cast<uint32_t>(fc.w), h = static_cast<uint32_t>(fc.h);
    const uint32_t ox = min(P.pass % 4u, w - 1u), oy = min((P.pass / 4u) % 4u, h - 1u);
    const uint32_t cw = (w - ox + 3u) / 4u, ch = (h - oy + 3u) / 4u;
    const uint64_t ncell = static_cast<uint64_t>(cw) * ch;
    for (uint32_t p = gid; p < npaths; p += gridDim.x * blockDim.x) {
        uint32_t x, y;
        if (p < fc.npix) {
            x = p % w;
            y = p / w;
            io.qvalid[p] = 0u;
        } else {
            const uint32_t k = p - fc.npix;
            const uint32_t c = static_cast<uint32_t>((static_cast<uint64_t>(k) * ncell) / io.ntrain);
            x = ox + 4u * (c % cw);
            y = oy + 4u * (c / cw);
            io.tvalid[k] = 0u;
        }
        const float tx = static_cast<float>(x) / static_cast<float>(fc.w);
        const float ty = static_cast<float>(y) / static_cast<float>(fc.h);
        const V3 d00 = v3(P.dirs[0], P.dirs[1], P.dirs[2]);
        const V3 d10 = v3(P.dirs[3], P.dirs[4], P.dirs[5]);
        const V3 d11 = v3(P.dirs[6], P.dirs[7], P.dirs[8]);
        const V3 d01 = v3(P.dirs[9], P.dirs[10], P.dirs[11]);
        const V3 d = mix(mix(d00, d10, tx), mix(d01, d11, tx), ty);
        q.ox[p] = P.eye[0];
        q.oy[p] = P.eye[1];
        q.oz[p] = P.eye[2];
        q.dx[p] = d.x;
        q.dy[p] = d.y;
        q.dz[p] = d.z;
        q.path[p] = p;
        ps.cr[p] = 0.0f;
        ps.cg[p] = 0.0f;
        ps.cb[p] = 0.0f;
    }
}

// Pixel paths that reached x2: radiance += throughput * cache(x2).
__global__ __launch_bounds__(kThreads) void k_cache_resolve(const FrameConst fc, CacheIO io, PathState ps) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < fc.npix; p += gridDim.x * blockDim.x) {
        if (!io.qvalid[p]) continue;
        const float* o = io.qout + 3 * static_cast<size_t>(p);
        ps.cr[p] = ps.cr[p] + ps.ar[p] * o[0];
        ps.cg[p] = ps.cg[p] + ps.ag[p] * o[1];
        ps.cb[p] = ps.cb[p] + ps.ab[p] * o[2];
    }
}

// Training targets: the suffix radiance of training paths that reached x2;
// NaN (masked) for the others.
__global__ __launch_bounds__(kThreads) void k_train_targets(const FrameConst fc, CacheIO io, PathState ps) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < io.ntrain; k += gridDim.x * blockDim.x) {
        const size_t p = static_cast<size_t>(fc.npix) + k;
        float* t = io.ttgt + 3 * static_cast<size_t>(k);
        if (io.tvalid[k]) {
            t[0] = ps.cr[p];
            t[1] = ps.cg[p];
            t[2] = ps.cb[p];
        } else {
            t[0] = t[1] = t[2] = __builtin_nanf("");
        }
    }
}

__global__ __launch_bounds__(kThreads) void k_zero_f64(double* __restrict__ p, size_t n) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        p[i] = 0.0;
}

struct Layout {
    RayQueue q[2];
    HitQueue hq;
    PathState ps;
};

constexpr int kQueueArrays = 2 * 7 + 4 + 6;

size_t slab_bytes(size_t npaths) { return (npaths * 4 + 255) & ~size_t(255); }

Layout carve(char* base, size_t npaths) {
    const size_t a = slab_bytes(npaths);
    auto take = [&]() {
        char* p = base;
        base += a;
        return p;
    };
    Layout L;
    for (int k = 0; k < 2; ++k) {
        L.q[k].ox = reinterpret_cast<float*>(take());
        L.q[k].oy = reinterpret_cast<float*>(take());
        L.q[k].oz = reinterpret_cast<float*>(take());
        L.q[k].dx = reinterpret_cast<float*>(take());
        L.q[k].dy = reinterpret_cast<float*>(take());
        L.q[k].dz = reinterpret_cast<float*>(take());
        L.q[k].path = reinterpret_cast<uint32_t*>(take());
    }
    L.hq.t = reinterpret_cast<float*>(take());
    L.hq.id = reinterpret_cast<uint32_t*>(take());
    L.hq.u = reinterpret_cast<float*>(take());
    L.hq.v = reinterpret_cast<float*>(take());
    L.ps.ar = reinterpret_cast<float*>(take());
    L.ps.ag = reinterpret_cast<float*>(take());
    L.ps.ab = reinterpret_cast<float*>(take());
    L.ps.cr = reinterpret_cast<float*>(take());
    L.ps.cg = reinterpret_cast<float*>(take());
    L.ps.cb = reinterpret_cast<float*>(take());
    return L;
}

uint32_t grid_for(size_t work) {
    size_t b = (work + kThreads - 1) / kThreads;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    return static_cast<uint32_t>(b);
}

// k_shade's grid: one workgroup pass covers kShadeThreads paths.
uint32_t shade_grid_for(size_t work) {
    size_t b = (work + kShadeThreads - 1) / kShadeThreads;
    if (b > 2048) b = 2048;
    if (b < 1) b = 1;
    return static_cast<uint32_t>(b);
}

// Everything a captured frame's kernel arguments depend on.
void set_detail(FrameConst& fc, const DeviceScene& scene) {
    fc.ntex = scene.ntex;
    fc.detail = scene.ndetail ? scene.detail.as<pt_triangle_detail>() : nullptr;
    fc.texels = scene.ntex ? scene.texels.as<uint32_t>() : nullptr;
    fc.texinfo = scene.ntex ? scene.texinfo.as<uint4>() : nullptr;
}

struct GraphKey {
    FrameConst fc;
    const void* nodes;
    const void* tris;
    const void* infos;
    const void* off;
    const void* queues;
    const void* params;
    const void* accum;
    const void* counters;
    const void* lights;
    uint32_t blocks, refill, lds, pool, chains, paths, shade_min;
    hipStream_t stream;
};

RayQueue shifted(const RayQueue& q, uint32_t k) {
    return RayQueue{q.ox + k, q.oy + k, q.oz + k, q.dx + k, q.dy + k, q.dz + k, q.path + k};
}
HitQueue shifted(const HitQueue& h, uint32_t k) { return HitQueue{h.t + k, h.id + k, h.u + k, h.v + k}; }

// Chains of a plain frame of npaths paths (render.chains, at most one per
// kMinChainPaths paths, at least one); 0: a path-persistent frame
// (render.paths).
uint32_t frame_chains(size_t npaths) {
    if (render_tuning().paths) return 0u;
    const uint32_t want = std::min(kMaxChains, std::max(1u, render_tuning().chains));
    const size_t fit = std::max<size_t>(1, npaths / kMinChainPaths);
    return static_cast<uint32_t>(std::min<size_t>(want, fit));
}

// Chain streams and their fork/join events (default priority), created
// outside any capture before the first frame that uses them.
void ensure_chain_streams(Renderer& r, uint32_t nchains) {
    if (nchains > 1 && !r.chain_fork) PT_HIP(hipEventCreateWithFlags(&r.chain_fork, hipEventDisableTiming));
    for (uint32_t c = 1; c < nchains; ++c) {
        if (!r.chain_st[c - 1]) PT_HIP(hipStreamCreateWithFlags(&r.chain_st[c - 1], hipStreamNonBlocking));
        if (!r.chain_join[c - 1]) PT_HIP(hipEventCreateWithFlags(&r.chain_join[c - 1], hipEventDisableTiming));
    }
}

template <int LDS>
auto paths_kernel(const FrameConst& fc) {
    return fc.detail ? (fc.nlights ? k_paths<LDS, true, true, false> : k_paths<LDS, false, true, false>)
                     : (fc.nlights ? k_paths<LDS, true, false, false> : k_paths<LDS, false, false, false>);
}

template <int LDS>
auto cache_paths_kernel(const FrameConst& fc) {
    return fc.detail ? (fc.nlights ? k_paths<LDS, true, true, true> : k_paths<LDS, false, true, true>)
                     : (fc.nlights ? k_paths<LDS, true, false, true> : k_paths<LDS, false, false, true>);
}

void launch_frame(Renderer& r, const DeviceScene& scene, const FrameConst& fc, uint32_t nchains, hipStream_t st) {
    const size_t npaths = static_cast<size_t>(fc.npix) * fc.npasses;
    Layout L = carve(r.queues.as<char>(), npaths);
    uint32_t* cnt = r.counters.as<uint32_t>();
    const PassParams* pp = r.params.as<PassParams>();
    const TraceTuning& tune = trace_tuning();
    const uint32_t nlev = fc.bounces + 1;
    const uint32_t grid = grid_for(npaths);
    const uint32_t refill = std::min(64u, std::max(1u, tune.refill));
    if (nchains == 0) {  // path-persistent frame
        hipLaunchKernelGGL(k_paths_setup, dim3(1), dim3(64), 0, st, cnt, cnt + 64, nlev);
        const uint32_t pgrid = std::min(std::max(1u, tune.blocks), grid);
        const uint32_t pool = pool_for(tune.pool, static_cast<uint32_t>(npaths), pgrid);
        auto kern = tune.lds <= 8 ? paths_kernel<8>(fc) : paths_kernel<16>(fc);
        hipLaunchKernelGGL(kern, dim3(pgrid), dim3(kPersistThreads), 0, st, fc, pp,
                           scene.infos.as<pt_triangle_info>(), scene.bvh.nodes.as<float4>(),
                           scene.bvh.tris.as<float4>(), scene.off.as<uint32_t>(), scene.lights.as<pt_dir_light>(),
                           L.ps, cnt, cnt + 64, refill, pool,
                           std::min(64u, std::max(1u, render_tuning().shade_min)), CacheIO{});
        hipLaunchKernelGGL(k_accumulate, dim3(grid_for(fc.npix)), dim3(kThreads), 0, st, fc, L.ps,
                           r.accum.as<double>());
        return;
    }
    auto extend = tune.lds <= 8 ? k_extend<8> : (tune.lds <= 16 ? k_extend<16> : k_extend<32>);
    auto shade = fc.detail ? (fc.nlights ? k_shade<false, true, true> : k_shade<false, false, true>)
                           : (fc.nlights ? k_shade<false, true, false> : k_shade<false, false, false>);
    hipLaunchKernelGGL(k_generate, dim3(std::max(grid, nchains)), dim3(kThreads), 0, st, fc, pp, L.q[0], L.ps, cnt,
                       nlev, nchains);
    if (nchains > 1) {
        PT_HIP(hipEventRecord(r.chain_fork, st));
        for (uint32_t c = 1; c < nchains; ++c) PT_HIP(hipStreamWaitEvent(r.chain_st[c - 1], r.chain_fork, 0));
    }
    const uint32_t np32 = static_cast<uint32_t>(npaths);
    for (uint32_t c = 0; c < nchains; ++c) {
        hipStream_t cs = c == 0 ? st : r.chain_st[c - 1];
        const uint32_t b = chain_begin(np32, c, nchains), len = chain_begin(np32, c + 1, nchains) - b;
        const RayQueue q[2] = {shifted(L.q[0], b), shifted(L.q[1], b)};
        const HitQueue hq = shifted(L.hq, b);
        uint32_t* cc = cnt + c * kChainWords;
        const uint32_t pgrid = std::min(std::max(1u, tune.blocks), grid_for(len));
        const uint32_t pool = pool_for(tune.pool, len, pgrid);
        for (uint32_t lvl = 0; lvl < nlev; ++lvl) {
            hipLaunchKernelGGL(extend, dim3(pgrid), dim3(kPersistThreads), 0, cs, fc, pp,
                               scene.bvh.nodes.as<float4>(), scene.bvh.tris.as<float4>(),
                               scene.off.as<uint32_t>(), q[lvl & 1], hq, cc, cc + 64, lvl, refill, pool);
            hipLaunchKernelGGL(shade, dim3(shade_grid_for(len)), dim3(kShadeThreads), 0, cs, fc, pp,
                               scene.infos.as<pt_triangle_info>(), q[lvl & 1], hq, q[(lvl + 1) & 1], L.ps, cc, lvl,
                               CacheIO{}, scene.lights.as<pt_dir_light>(), scene.bvh.nodes.as<float4>(),
                               scene.bvh.tris.as<float4>(), scene.off.as<uint32_t>());
        }
        if (c > 0) {
            PT_HIP(hipEventRecord(r.chain_join[c - 1], cs));
            PT_HIP(hipStreamWaitEvent(st, r.chain_join[c - 1], 0));
        }
    }
    hipLaunchKernelGGL(k_accumulate, dim3(grid_for(fc.npix)), dim3(kThreads), 0, st, fc, L.ps,
                       r.accum.as<double>());
}

}  // namespace

Renderer::~Renderer() {
    if (graph_exec) (void)hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec));
    if (cache_graph_exec) (void)hipGraphExecDestroy(static_cast<hipGraphExec_t>(cache_graph_exec));
    for (auto& e : ring_ev)
        if (e) (void)hipEventDestroy(e);
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    if (chain_fork) (void)hipEventDestroy(chain_fork);
    for (auto& e : chain_join)
        if (e) (void)hipEventDestroy(e);
    for (auto& c : chain_st)
        if (c) (void)hipStreamDestroy(c);
    if (join_ev) (void)hipEventDestroy(join_ev);
    if (side) (void)hipStreamDestroy(side);
}

void render_resize(Renderer& r, int w, int h, hipStream_t st) {
    if (w <= 0 || h <= 0) throw Error(PT_E_INVALID_ARG, "render size must be positive");
    if (static_cast<uint64_t>(w) * static_cast<uint64_t>(h) > (1ull << 28))
        throw Error(PT_E_INVALID_ARG, "at most 2^28 pixels");
    if (r.w == w && r.h == h) return;
    const size_t np = static_cast<size_t>(w) * h;
    r.w = w;
    r.h = h;
    r.accum.reserve(np * 3 * sizeof(double));
    // Per chain: queue lengths (64 words), then one ray queue block per level.
    r.counters.reserve(kMaxChains * kChainWords * sizeof(uint32_t));
    render_reset(r, st);
}

namespace {
// One thread per pixel: rgba = (float(sum * (1 / spp)), 1), the value
// pt_read_image computes on the host (same double product, same rounding).
__global__ __launch_bounds__(kThreads) void k_display(const double* __restrict__ accum, int w,
                                                      uint32_t spp, char* __restrict__ out,
                                                      uint32_t ow, uint32_t oh, size_t pitch) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= ow || y >= oh) return;
    const double inv = spp ? 1.0 / static_cast<double>(spp) : 0.0;
    const double* s = accum + 3 * (static_cast<size_t>(y) * w + x);
    float4* row = reinterpret_cast<float4*>(out + static_cast<size_t>(y) * pitch);
    row[x] = make_float4(static_cast<float>(s[0] * inv), static_cast<float>(s[1] * inv),
                         static_cast<float>(s[2] * inv), 1.0f);
}
}  // namespace

void render_display(const Renderer& r, void* out, uint32_t out_w, uint32_t out_h, size_t pitch,
                    hipStream_t st) {
    const uint32_t w = std::min(out_w, static_cast<uint32_t>(r.w));
    const uint32_t h = std::min(out_h, static_cast<uint32_t>(r.h));
    if (w == 0 || h == 0) return;
    hipLaunchKernelGGL(k_display, dim3((w + kThreads - 1) / kThreads, h), dim3(kThreads), 0, st,
                       r.accum.as<double>(), r.w, r.spp, static_cast<char*>(out), w, h, pitch);
    PT_HIP(hipGetLastError());
}

void render_reset(Renderer& r, hipStream_t st) {
    r.spp = 0;
    r.gen_seconds = -1.0;
    if (r.w <= 0) return;
    const size_t n = static_cast<size_t>(r.w) * r.h * 3;
    hipLaunchKernelGGL(k_zero_f64, dim3(grid_for(n)), dim3(kThreads), 0, st, r.accum.as<double>(), n);
    PT_HIP(hipGetLastError());
}

namespace {

// Per-pass block: through a ring of pinned slots, so a slot is rewritten only
// after its previous copy has left.
void upload_pass_params(Renderer& r, uint32_t npasses, const float* dirs, const uint64_t* seeds,
                        const uint32_t* scenes, hipStream_t st) {
    r.params.reserve(sizeof(PassParams) * kMaxPasses);
    r.params_host.reserve(sizeof(PassParams) * kMaxPasses * kParamRing);
    if (!r.ring_ev[0])
        for (auto& e : r.ring_ev) PT_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    const int slot = r.ring_next;
    r.ring_next = (r.ring_next + 1) % kParamRing;
    if (r.ring_used[slot]) PT_HIP(hipEventSynchronize(r.ring_ev[slot]));
    PassParams* hp = r.params_host.as<PassParams>() + static_cast<size_t>(slot) * kMaxPasses;
    for (uint32_t k = 0; k < npasses; ++k) {
        std::memcpy(hp[k].dirs, dirs ? dirs + 12 * k : r.dirs, sizeof hp[k].dirs);
        std::memcpy(hp[k].eye, r.eye, sizeof hp[k].eye);
        hp[k].pass = r.spp + k;
        hp[k].seedw = static_cast<uint32_t>(seeds[k]) ^ static_cast<uint32_t>(seeds[k] >> 32);
        hp[k].scene = scenes ? scenes[k] : 0u;
    }
    PT_HIP(hipMemcpyAsync(r.params.as<void>(), hp, sizeof(PassParams) * npasses, hipMemcpyHostToDevice, st));
    PT_HIP(hipEventRecord(r.ring_ev[slot], st));
    r.ring_used[slot] = true;
}

// Capture `launch` into the graph slot (exec, key) unless the key matches
// the last capture, then launch the graph.
template <typename Key, size_t Cap, typename Launch>
void replay(void*& exec, char (&key_store)[Cap], const Key& key, hipStream_t st, Launch launch) {
    static_assert(std::is_trivially_copyable<Key>::value, "graph key");
    static_assert(sizeof(Key) <= Cap, "graph key storage too small for the key");
    if (!exec || std::memcmp(key_store, &key, sizeof key) != 0) {
        if (exec) {
            PT_HIP(hipStreamSynchronize(st));
            PT_HIP(hipGraphExecDestroy(static_cast<hipGraphExec_t>(exec)));
            exec = nullptr;
        }
        hipGraph_t g = nullptr;
        PT_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            launch();
        } catch (...) {
            // End the capture so the stream is usable again; drop the partial graph.
            hipGraph_t partial = nullptr;
            (void)hipStreamEndCapture(st, &partial);
            if (partial) (void)hipGraphDestroy(partial);
            (void)hipGetLastError();
            throw;
        }
        const hipError_t le = hipGetLastError();
        const hipError_t ce = hipStreamEndCapture(st, &g);
        if (le != hipSuccess || ce != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            if (le != hipSuccess) throw Error(PT_E_HIP, std::string("frame capture: ") + hipGetErrorString(le));
            throw Error(PT_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
        }
        hipGraphExec_t ex = nullptr;
        const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) throw Error(PT_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
        exec = ex;
        std::memcpy(key_store, &key, sizeof key);
    }
    PT_HIP(hipGraphLaunch(static_cast<hipGraphExec_t>(exec), st));
}

}  // namespace

void render_passes(Renderer& r, const DeviceScene& scene, uint32_t npasses, const float* dirs,
                   const uint64_t* seeds, const uint32_t* scenes, hipStream_t st) {
    if (!scene.ready) throw Error(PT_E_NO_SCENE, "no scene uploaded");
    if (r.settings.bounces > 60) throw Error(PT_E_INVALID_ARG, "at most 60 bounces");
    if (npasses == 0 || npasses > kMaxPasses) throw Error(PT_E_INVALID_ARG, "1 .. 1024 passes per call");
    if (scenes)
        for (uint32_t k = 0; k < npasses; ++k)
            if (scenes[k] >= scene.nscenes()) throw Error(PT_E_INVALID_ARG, "scene index out of range");
    const size_t npix = static_cast<size_t>(r.w) * r.h;
    if (npix * npasses >= (1ull << 31)) throw Error(PT_E_INVALID_ARG, "too many paths in one frame");
    r.queues.reserve(slab_bytes(npix * npasses) * kQueueArrays);
    upload_pass_params(r, npasses, dirs, seeds, scenes, st);

    FrameConst fc;
    fc.w = r.w;
    fc.h = r.h;
    fc.npix = static_cast<uint32_t>(npix);
    fc.npasses = npasses;
    fc.bounces = r.settings.bounces;
    fc.rr_start = r.settings.rr_start;
    fc.intensity = r.settings.indirect_intensity;
    fc.stratified = r.settings.stratified;
    fc.nlights = scene.nlights;
    fc.status = scene.status.as<uint32_t>();
    set_detail(fc, scene);

    const uint32_t nchains = frame_chains(npix * npasses);
    ensure_chain_streams(r, std::max(1u, nchains));
    if (!render_tuning().graphs) {
        launch_frame(r, scene, fc, nchains, st);
        PT_HIP(hipGetLastError());
    } else {
        GraphKey key;
        std::memset(&key, 0, sizeof key);
        key.fc = fc;
        key.nodes = scene.bvh.nodes.as<void>();
        key.tris = scene.bvh.tris.as<void>();
        key.infos = scene.infos.as<void>();
        key.off = scene.off.as<void>();
        key.queues = r.queues.as<void>();
        key.params = r.params.as<void>();
        key.accum = r.accum.as<void>();
        key.counters = r.counters.as<void>();
        key.lights = scene.lights.as<void>();
        key.blocks = trace_tuning().blocks;
        key.refill = trace_tuning().refill;
        key.lds = trace_tuning().lds;
        key.pool = trace_tuning().pool;
        key.chains = nchains;
        key.paths = render_tuning().paths;
        key.shade_min = render_tuning().shade_min;
        key.stream = st;
        replay(r.graph_exec, r.graph_key, key, st,
               [&] { launch_frame(r, scene, fc, nchains, st); });
    }
    r.spp += npasses;
    r.last_levels = fc.bounces + 1;
    r.last_chains = std::max(1u, nchains);
}

namespace {

struct CacheFrame {
    FrameConst fc;
    CacheIO io;
    uint32_t steps;   // Adam steps over the training batch
    bool train;
};

void launch_cache_frame(Renderer& r, const DeviceScene& scene, RadianceCache& cache, const CacheFrame& cf,
                        hipStream_t st) {
    const FrameConst& fc = cf.fc;
    const CacheIO& io = cf.io;
    const size_t npaths = static_cast<size_t>(fc.npix) + io.ntrain;
    Layout L = carve(r.queues.as<char>(), npaths);
    uint32_t* cnt = r.counters.as<uint32_t>();
    uint32_t* rcnt = cnt + 64;
    const PassParams* pp = r.params.as<PassParams>();
    const TraceTuning& tune = trace_tuning();
    const uint32_t nlev = io.ntrain ? std::max(2u, io.train_levels) : 2u;
    const uint32_t grid = grid_for(npaths);
    const uint32_t pgrid = std::min(std::max(1u, tune.blocks), grid);
    const uint32_t refill = std::min(64u, std::max(1u, tune.refill));
    const uint32_t pool = pool_for(tune.pool, static_cast<uint32_t>(std::min<size_t>(npaths, 0xffffffffu)), pgrid);
    auto extend = tune.lds <= 8 ? k_extend<8> : (tune.lds <= 16 ? k_extend<16> : k_extend<32>);
    auto shade = fc.detail ? (fc.nlights ? k_shade<true, true, true> : k_shade<true, false, true>)
                           : (fc.nlights ? k_shade<true, true, false> : k_shade<true, false, false>);
    auto train_paths = fc.detail ? k_train_paths<true> : k_train_paths<false>;
    const bool paths = render_tuning().paths != 0;
    hipLaunchKernelGGL(k_generate_cache, dim3(grid), dim3(kThreads), 0, st, fc, pp, L.q[0], L.ps, cnt, rcnt,
                       nlev, io, paths ? 0u : fc.npix);
    // The training paths run on the side stream from here, beside the pixel
    // paths (two levels, or one path-persistent launch), inference, resolve
    // and accumulation: they read the first level's queue slots from npix on
    // and write only training-path state, the training batch and counter
    // slots the pixel paths do not read.
    // The join comes before the training targets and the Adam steps, which
    // change the weights the inference reads.
    if (io.ntrain) {
        PT_HIP(hipEventRecord(r.fork_ev, st));
        PT_HIP(hipStreamWaitEvent(r.side, r.fork_ev, 0));
        const uint32_t sgrid = std::max(1u, std::min(4096u, (io.ntrain + kSuffixThreads - 1) / kSuffixThreads));
        hipLaunchKernelGGL(train_paths, dim3(sgrid), dim3(kSuffixThreads), 0, r.side, fc, pp,
                           scene.infos.as<pt_triangle_info>(), L.q[0], L.ps, cnt, io,
                           scene.lights.as<pt_dir_light>(), scene.bvh.nodes.as<float4>(),
                           scene.bvh.tris.as<float4>(), scene.off.as<uint32_t>());
        PT_HIP(hipEventRecord(r.join_ev, r.side));
    }
    if (paths) {
        // The pixel paths in one path-persistent launch (render.paths), each
        // lane carrying a path from its primary ray to the cache.
        const uint32_t ppgrid = std::min(std::max(1u, tune.blocks), grid_for(fc.npix));
        const uint32_t ppool = pool_for(tune.pool, fc.npix, ppgrid);
        auto kern = tune.lds <= 8 ? cache_paths_kernel<8>(fc) : cache_paths_kernel<16>(fc);
        hipLaunchKernelGGL(kern, dim3(ppgrid), dim3(kPersistThreads), 0, st, fc, pp,
                           scene.infos.as<pt_triangle_info>(), scene.bvh.nodes.as<float4>(),
                           scene.bvh.tris.as<float4>(), scene.off.as<uint32_t>(), scene.lights.as<pt_dir_light>(),
                           L.ps, cnt, rcnt, refill, ppool, std::min(64u, std::max(1u, render_tuning().shade_min)),
                           io);
    } else {
        // Levels 0 and 1 of the pixel paths (a pixel path ends in the cache at x2).
        for (uint32_t lvl = 0; lvl < 2; ++lvl) {
            hipLaunchKernelGGL(extend, dim3(pgrid), dim3(kPersistThreads), 0, st, fc, pp,
                               scene.bvh.nodes.as<float4>(), scene.bvh.tris.as<float4>(),
                               scene.off.as<uint32_t>(), L.q[lvl & 1], L.hq, cnt, rcnt, lvl, refill, pool);
            hipLaunchKernelGGL(shade, dim3(shade_grid_for(npaths)), dim3(kShadeThreads), 0, st, fc, pp,
                               scene.infos.as<pt_triangle_info>(), L.q[lvl & 1], L.hq,
                               L.q[(lvl + 1) & 1], L.ps, cnt, lvl, io, scene.lights.as<pt_dir_light>(),
                               scene.bvh.nodes.as<float4>(), scene.bvh.tris.as<float4>(),
                               scene.off.as<uint32_t>());
        }
    }
    cache.infer(io.qraw, fc.npix, io.qout, st);
    hipLaunchKernelGGL(k_cache_resolve, dim3(grid_for(fc.npix)), dim3(kThreads), 0, st, fc, io, L.ps);
    hipLaunchKernelGGL(k_accumulate, dim3(grid_for(fc.npix)), dim3(kThreads), 0, st, fc, L.ps,
                       r.accum.as<double>());
    if (io.ntrain) PT_HIP(hipStreamWaitEvent(st, r.join_ev, 0));
    if (io.ntrain) {
        hipLaunchKernelGGL(k_train_targets, dim3(grid_for(io.ntrain)), dim3(kThreads), 0, st, fc, io, L.ps);
        if (cf.train) {
            const uint32_t chunk = (io.ntrain + cf.steps - 1) / cf.steps;
            for (uint32_t b = 0; b < io.ntrain; b += chunk) {
                const uint32_t n = std::min(chunk, io.ntrain - b);
                cache.launch_train(io.traw + 12 * static_cast<size_t>(b), io.ttgt + 3 * static_cast<size_t>(b),
                                   n, nullptr, nullptr, st);
            }
        }
    }
}

struct CacheGraphKey {
    FrameConst fc;
    CacheIO io;
    uint32_t steps, train, blocks, refill, lds, pool, paths, shade_min;
    const void* ptrs[13];
    float box[6], adam[4];
    hipStream_t stream;
};

}  // namespace

void render_cache_frame(Renderer& r, const DeviceScene& scene, RadianceCache& cache, uint64_t seed,
                        const pt_cache_settings& cs, hipStream_t st) {
    if (!scene.ready) throw Error(PT_E_NO_SCENE, "no scene uploaded");
    if (!cache.ready) throw Error(PT_E_INVALID_ARG, "radiance cache not initialised");
    if (cs.extra_bounces > 58) throw Error(PT_E_INVALID_ARG, "at most 58 extra bounces");
    if (cs.train_steps < 1 || cs.train_steps > 64) throw Error(PT_E_INVALID_ARG, "1 .. 64 training steps");
    if (cs.flags & ~static_cast<uint32_t>(PT_CACHE_TRAIN)) throw Error(PT_E_INVALID_ARG, "unknown cache flag");
    const size_t npix = static_cast<size_t>(r.w) * r.h;
    if (cs.train_samples > npix) throw Error(PT_E_INVALID_ARG, "more training paths than pixels");
    if (npix + cs.train_samples >= (1ull << 31)) throw Error(PT_E_INVALID_ARG, "too many paths in one frame");
    const uint32_t ntrain = cs.train_samples;

    CacheFrame cf;
    cf.fc.w = r.w;
    cf.fc.h = r.h;
    cf.fc.npix = static_cast<uint32_t>(npix);
    cf.fc.npasses = 1;
    cf.fc.bounces = 1;
    cf.fc.rr_start = r.settings.rr_start;
    cf.fc.intensity = r.settings.indirect_intensity;
    cf.fc.stratified = r.settings.stratified;
    cf.fc.nlights = scene.nlights;
    cf.fc.status = scene.status.as<uint32_t>();
    set_detail(cf.fc, scene);
    cf.steps = cs.train_steps;
    cf.train = (cs.flags & PT_CACHE_TRAIN) != 0;

    // Allocation first: nothing below may allocate inside a capture.
    r.queues.reserve(slab_bytes(npix + ntrain) * kQueueArrays);
    auto bytes = [](size_t n) { return (n * 4 + 255) & ~size_t(255); };
    r.cache_io.reserve(bytes(npix * 12) + bytes(npix * 3) + bytes(npix) + bytes(size_t(ntrain) * 12) +
                       bytes(size_t(ntrain) * 3) + bytes(ntrain));
    char* base = r.cache_io.as<char>();
    auto take = [&](size_t n) {
        char* p = base;
        base += bytes(n);
        return p;
    };
    cf.io.qraw = reinterpret_cast<float*>(take(npix * 12));
    cf.io.qout = reinterpret_cast<float*>(take(npix * 3));
    cf.io.qvalid = reinterpret_cast<uint32_t*>(take(npix));
    cf.io.traw = reinterpret_cast<float*>(take(size_t(ntrain) * 12));
    cf.io.ttgt = reinterpret_cast<float*>(take(size_t(ntrain) * 3));
    cf.io.tvalid = reinterpret_cast<uint32_t*>(take(ntrain));
    cf.io.ntrain = ntrain;
    cf.io.train_levels = 2u + cs.extra_bounces;
    if (ntrain && cf.train) cache.prepare_train((ntrain + cf.steps - 1) / cf.steps);
    upload_pass_params(r, 1, nullptr, &seed, nullptr, st);
    if (!r.side || !r.fork_ev || !r.join_ev) {
        if (!r.side) PT_HIP(hipStreamCreateWithFlags(&r.side, hipStreamNonBlocking));
        if (!r.fork_ev) PT_HIP(hipEventCreateWithFlags(&r.fork_ev, hipEventDisableTiming));
        if (!r.join_ev) PT_HIP(hipEventCreateWithFlags(&r.join_ev, hipEventDisableTiming));
    }

    if (!render_tuning().graphs) {
        launch_cache_frame(r, scene, cache, cf, st);
        PT_HIP(hipGetLastError());
    } else {
        CacheGraphKey key;
        std::memset(&key, 0, sizeof key);
        key.fc = cf.fc;
        key.io = cf.io;
        key.steps = cf.steps;
        key.train = cf.train;
        key.blocks = trace_tuning().blocks;
        key.refill = trace_tuning().refill;
        key.lds = trace_tuning().lds;
        key.pool = trace_tuning().pool;
        key.paths = render_tuning().paths;
        key.shade_min = render_tuning().shade_min;
        const void* ptrs[13] = {scene.bvh.nodes.as<void>(), scene.bvh.tris.as<void>(), scene.infos.as<void>(),
                                scene.off.as<void>(),       r.queues.as<void>(),     r.params.as<void>(),
                                r.accum.as<void>(),         r.counters.as<void>(),   cache.master.as<void>(),
                                cache.f32.as<void>(),       cache.part.as<void>(),   cache.loss.as<void>(),
                                scene.lights.as<void>()};
        std::memcpy(key.ptrs, ptrs, sizeof ptrs);
        for (int k = 0; k < 3; ++k) {
            key.box[k] = cache.box_min[k];
            key.box[3 + k] = cache.box_inv[k];
        }
        key.adam[0] = cache.lr;
        key.adam[1] = cache.beta1;
        key.adam[2] = cache.beta2;
        key.adam[3] = cache.eps;
        key.stream = st;
        replay(r.cache_graph_exec, r.cache_graph_key, key, st,
               [&] { launch_cache_frame(r, scene, cache, cf, st); });
    }
    r.spp += 1;
    r.last_levels = ntrain ? std::max(2u, cf.io.train_levels) : 2u;
    r.last_chains = 1;
    r.last_train = ntrain;
}

void render_last_queries(Renderer& r, float* raw, float* pred, uint32_t* valid, hipStream_t st) {
    if (r.last_train == 0 && r.cache_io.capacity() == 0) throw Error(PT_E_INVALID_ARG, "no cache frame rendered");
    const size_t npix = static_cast<size_t>(r.w) * r.h;
    auto bytes = [](size_t k) { return (k * 4 + 255) & ~size_t(255); };
    const char* base = r.cache_io.as<char>();
    if (raw) PT_HIP(hipMemcpyAsync(raw, base, npix * 48, hipMemcpyDeviceToHost, st));
    base += bytes(npix * 12);
    if (pred) PT_HIP(hipMemcpyAsync(pred, base, npix * 12, hipMemcpyDeviceToHost, st));
    base += bytes(npix * 3);
    if (valid) PT_HIP(hipMemcpyAsync(valid, base, npix * 4, hipMemcpyDeviceToHost, st));
    PT_HIP(hipStreamSynchronize(st));
}

void render_last_batch(Renderer& r, float* raw, float* target, uint32_t n, hipStream_t st) {
    if (n > r.last_train) throw Error(PT_E_INVALID_ARG, "more samples than the last cache frame's batch");
    if (n == 0) return;
    const size_t npix = static_cast<size_t>(r.w) * r.h;
    auto bytes = [](size_t k) { return (k * 4 + 255) & ~size_t(255); };
    const char* base = r.cache_io.as<char>() + bytes(npix * 12) + bytes(npix * 3) + bytes(npix);
    if (raw) PT_HIP(hipMemcpyAsync(raw, base, size_t(n) * 48, hipMemcpyDeviceToHost, st));
    base += bytes(size_t(r.last_train) * 12);
    if (target) PT_HIP(hipMemcpyAsync(target, base, size_t(n) * 12, hipMemcpyDeviceToHost, st));
    PT_HIP(hipStreamSynchronize(st));
}

uint64_t render_last_rays(Renderer& r, hipStream_t st) {
    if (r.last_levels == 0) return 0;
    // Levels 0 .. last_levels-1 plus the cache frame's training paths at
    // their first two levels (slots 62 and 63, zeroed by every frame).
    // A plain frame's chains each count their own levels.
    uint32_t h[kMaxChains][64];
    const uint32_t nch = std::max(1u, std::min(kMaxChains, r.last_chains));
    for (uint32_t c = 0; c < nch; ++c)
        PT_HIP(hipMemcpyAsync(h[c], r.counters.as<uint32_t>() + c * kChainWords, 64 * sizeof(uint32_t),
                              hipMemcpyDeviceToHost, st));
    PT_HIP(hipStreamSynchronize(st));
    uint64_t total = 0;
    for (uint32_t c = 0; c < nch; ++c)
        for (uint32_t k = 0; k < std::min(r.last_levels, kTrainLevel0Slot); ++k) total += h[c][k];
    return total + h[0][kTrainLevel0Slot] + h[0][kTrainLevel1Slot];
}

RenderTuning& render_tuning() {
    static RenderTuning t;
    return t;
}

}  // namespace pt
