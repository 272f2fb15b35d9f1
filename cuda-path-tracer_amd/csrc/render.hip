// Progressive unidirectional path tracer as a wavefront pipeline
// (reference: raytrace_kernel + blinnPhongShadingPathTrace,
// cuda_path_tracer.cu:245-333, 385-421, which run one megakernel thread per
// pixel and one launch per 1-spp pass).  One frame renders P passes at once
// (path = pass * pixels + pixel) so small images still fill the GPU:
//   k_generate     primary rays of every path into queue 0, counters reset
//   k_extend(L)    persistent closest hit for every ray of queue L
//   k_shade(L)     shade the hits, append surviving paths to queue L+1
//                  (wave-aggregated atomics), L = 0 .. bounces
//   k_accumulate   per pixel, FP64 sums += each pass's radiance, in pass order
// Queue lengths live on the device, so the host never waits inside a frame,
// and a frame's kernel sequence is captured once as a hipGraph and replayed.
// Per-pass values (corner rays, eye, pass index, seed) are read from device
// memory, so replays only need a fresh copy of that small block.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

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

constexpr int kThreads = 256;

__global__ __launch_bounds__(kThreads) void k_generate(const FrameConst fc,
                                                       const PassParams* __restrict__ pp,
                                                       RayQueue q, PathState ps,
                                                       uint32_t* __restrict__ counters,
                                                       uint32_t* __restrict__ ray_counters,
                                                       uint32_t nlevels) {
    const uint32_t npaths = fc.npix * fc.npasses;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid == 0) {
        counters[0] = npaths;
        for (uint32_t k = 1; k <= nlevels; ++k) counters[k] = 0u;
        for (uint32_t k = 0; k <= nlevels; ++k) ray_counters[k] = 0u;
    }
    for (uint32_t p = gid; p < npaths; p += gridDim.x * blockDim.x) {
        const PassParams& P = pp[p / fc.npix];
        const uint32_t pix = p % fc.npix;
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

// Closest hit for every ray of queue `level` (count on the device): the
// persistent traversal, rays handed out from ray_counters[level].
// Each path traces the scene of its pass (root internal node off[scene]).
__global__ __launch_bounds__(kPersistThreads) void k_extend(const FrameConst fc,
                                                            const PassParams* __restrict__ pp,
                                                            const float4* __restrict__ nodes,
                                                            const float4* __restrict__ tris,
                                                            const uint32_t* __restrict__ off,
                                                            RayQueue q, HitQueue hq,
                                                            const uint32_t* __restrict__ counters,
                                                            uint32_t* __restrict__ ray_counters,
                                                            uint32_t level, uint32_t refill) {
    __shared__ uint32_t s_stack[kLdsStack][kPersistThreads];
    const uint32_t count = counters[level];
    TraceCount cnt{0, 0, 0};
    auto load = [&](uint32_t i, V3& o, V3& d, uint32_t& single) -> uint32_t {
        o = v3(q.ox[i], q.oy[i], q.oz[i]);
        d = v3(q.dx[i], q.dy[i], q.dz[i]);
        const uint32_t sc = pp[q.path[i] / fc.npix].scene;
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
    persistent_trace<false>(nodes, tris, count, ray_counters + level, refill, s_stack, load, store, cnt);
}

__device__ __forceinline__ V3 ld3(const float* p) { return v3(p[0], p[1], p[2]); }

__global__ __launch_bounds__(kThreads) void k_shade(const FrameConst fc,
                                                    const PassParams* __restrict__ pp,
                                                    const pt_triangle_info* __restrict__ infos,
                                                    RayQueue qin, HitQueue hq, RayQueue qout,
                                                    PathState ps, uint32_t* __restrict__ counters,
                                                    uint32_t level) {
    const uint32_t count = counters[level];
    const int lane = threadIdx.x & 63;
    const uint64_t lt = (1ull << lane) - 1ull;
    const uint32_t stride = gridDim.x * blockDim.x;
    // Wave-uniform trip count so every lane reaches the ballots below.
    for (uint32_t base = blockIdx.x * blockDim.x; base < count; base += stride) {
        const uint32_t i = base + threadIdx.x;
        bool cont = false;
        V3 no{0.0f, 0.0f, 0.0f}, nd{0.0f, 0.0f, 0.0f};
        uint32_t path = 0;
        if (i < count && hq.id[i] != kInvalid) {
            path = qin.path[i];
            const PassParams& P = pp[path / fc.npix];
            const uint32_t pix = path % fc.npix;
            const V3 o = v3(qin.ox[i], qin.oy[i], qin.oz[i]);
            const V3 d = v3(qin.dx[i], qin.dy[i], qin.dz[i]);
            const float t = hq.t[i];
            const pt_triangle_info& in = infos[hq.id[i]];
            // Hit point and shading data (cuda_path_tracer.cu:190-191, 298-300).
            const V3 pos = add(scale(d, t), o);
            const float bu = hq.u[i], bv = hq.v[i];
            const float bw = (1.0f - bu) - bv;
            const V3 nrm = normalize(add(add(scale(ld3(in.n0), bw), scale(ld3(in.n1), bu)),
                                         scale(ld3(in.n2), bv)));
            V3 albedo = ld3(in.albedo);
            const V3 e = ld3(in.emission);
            const bool emissive = ((e.x + e.y) + e.z) > 1e-7f;
            V3 absorbed, acc;
            if (level == 0) {
                // Primary hit: emission or (no directional lights) zero
                // (cuda_path_tracer.cu:194-199, 258-259).
                acc = v3(0.0f, 0.0f, 0.0f);
                if (emissive) {
                    acc = e;
                    albedo = v3(1.0f, 1.0f, 1.0f);
                }
                absorbed = albedo;
            } else {
                absorbed = v3(ps.ar[path], ps.ag[path], ps.ab[path]);
                acc = v3(ps.cr[path], ps.cg[path], ps.cb[path]);
                V3 L = v3(0.0f, 0.0f, 0.0f);
                if (emissive) {
                    // Emitter cosine (cuda_path_tracer.cu:304-312).
                    const float c = gmax(dot(neg(d), nrm), 0.0f);
                    L = scale(e, c);
                    albedo = v3(1.0f, 1.0f, 1.0f);
                }
                absorbed = mul(absorbed, albedo);
                acc = add(acc, scale(mul(absorbed, L), fc.intensity));
            }
            // Next bounce index = level (bounce loop, cuda_path_tracer.cu:267).
            const uint32_t b = level;
            if (b < fc.bounces) {
                const U4 rv = pcg4d(U4{pix, P.pass, b, P.seedw});
                float u1 = u24(rv.x), u2 = u24(rv.y);
                cont = true;
                if (b >= fc.rr_start) {
                    float p = gmax(gmax(albedo.x, albedo.y), albedo.z);
                    p = gmin(p, 1.0f);
                    if (!(u24(rv.z) < p)) {
                        cont = false;
                    } else {
                        absorbed = v3(absorbed.x / p, absorbed.y / p, absorbed.z / p);
                    }
                }
                if (cont) {
                    no = add(pos, scale(nrm, 1e-5f));
                    if (fc.stratified != 0u && b == 0) {
                        // 8x8 strata, stratum = pass % 64 (cuda_path_tracer.cu:273-280).
                        const uint32_t p1 = P.pass % 64u;
                        u1 = (static_cast<float>(p1 % 8u) + u1) / 8.0f;
                        u2 = (static_cast<float>(p1 / 8u) + u2) / 8.0f;
                    }
                    // Cosine-weighted hemisphere sample (random.h:95-102).
                    const float r = __builtin_sqrtf(u1);
                    float sn, cs;
                    sincos2pi(u2, sn, cs);
                    const V3 local = v3(r * cs, r * sn, __builtin_sqrtf(1.0f - u1));
                    nd = to_hemisphere(local, nrm);
                }
            }
            ps.ar[path] = absorbed.x;
            ps.ag[path] = absorbed.y;
            ps.ab[path] = absorbed.z;
            ps.cr[path] = acc.x;
            ps.cg[path] = acc.y;
            ps.cb[path] = acc.z;
        }
        const uint64_t m = __ballot(cont);
        if (m != 0ull) {
            const int leader = __ffsll(static_cast<long long>(m)) - 1;
            uint32_t wbase = 0;
            if (lane == leader) wbase = atomicAdd(counters + level + 1, static_cast<uint32_t>(__popcll(m)));
            wbase = __shfl(wbase, leader, 64);
            if (cont) {
                const uint32_t s = wbase + static_cast<uint32_t>(__popcll(m & lt));
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

// Everything a captured frame's kernel arguments depend on.
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
    uint32_t blocks, refill;
    hipStream_t stream;
};

void launch_frame(Renderer& r, const DeviceScene& scene, const FrameConst& fc, hipStream_t st) {
    const size_t npaths = static_cast<size_t>(fc.npix) * fc.npasses;
    Layout L = carve(r.queues.as<char>(), npaths);
    uint32_t* cnt = r.counters.as<uint32_t>();
    uint32_t* rcnt = cnt + 64;
    const PassParams* pp = r.params.as<PassParams>();
    const TraceTuning& tune = trace_tuning();
    const uint32_t nlev = fc.bounces + 1;
    const uint32_t grid = grid_for(npaths);
    const uint32_t pgrid = std::min(std::max(1u, tune.blocks), grid);
    const uint32_t refill = std::min(64u, std::max(1u, tune.refill));
    hipLaunchKernelGGL(k_generate, dim3(grid), dim3(kThreads), 0, st, fc, pp, L.q[0], L.ps, cnt, rcnt, nlev);
    for (uint32_t lvl = 0; lvl < nlev; ++lvl) {
        hipLaunchKernelGGL(k_extend, dim3(pgrid), dim3(kPersistThreads), 0, st, fc, pp,
                           scene.bvh.nodes.as<float4>(), scene.bvh.tris.as<float4>(),
                           scene.off.as<uint32_t>(), L.q[lvl & 1], L.hq, cnt, rcnt, lvl, refill);
        hipLaunchKernelGGL(k_shade, dim3(grid), dim3(kThreads), 0, st, fc, pp,
                           scene.infos.as<pt_triangle_info>(), L.q[lvl & 1], L.hq,
                           L.q[(lvl + 1) & 1], L.ps, cnt, lvl);
    }
    hipLaunchKernelGGL(k_accumulate, dim3(grid_for(fc.npix)), dim3(kThreads), 0, st, fc, L.ps,
                       r.accum.as<double>());
}

}  // namespace

Renderer::~Renderer() {
    if (graph_exec) (void)hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec));
    for (auto& e : ring_ev)
        if (e) (void)hipEventDestroy(e);
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
    r.counters.reserve(128 * sizeof(uint32_t));  // queue lengths, then ray hand-out counters
    render_reset(r, st);
}

void render_reset(Renderer& r, hipStream_t st) {
    r.spp = 0;
    if (r.w <= 0) return;
    const size_t n = static_cast<size_t>(r.w) * r.h * 3;
    hipLaunchKernelGGL(k_zero_f64, dim3(grid_for(n)), dim3(kThreads), 0, st, r.accum.as<double>(), n);
    PT_HIP(hipGetLastError());
}

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
    r.params.reserve(sizeof(PassParams) * kMaxPasses);
    r.params_host.reserve(sizeof(PassParams) * kMaxPasses * kParamRing);
    if (!r.ring_ev[0])
        for (auto& e : r.ring_ev) PT_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));

    // Per-pass block: through a ring of pinned slots, so a slot is rewritten
    // only after its previous copy has left.
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

    FrameConst fc;
    fc.w = r.w;
    fc.h = r.h;
    fc.npix = static_cast<uint32_t>(npix);
    fc.npasses = npasses;
    fc.bounces = r.settings.bounces;
    fc.rr_start = r.settings.rr_start;
    fc.intensity = r.settings.indirect_intensity;
    fc.stratified = r.settings.stratified;

    if (!render_tuning().graphs) {
        launch_frame(r, scene, fc, st);
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
        key.blocks = trace_tuning().blocks;
        key.refill = trace_tuning().refill;
        key.stream = st;
        static_assert(sizeof(GraphKey) <= sizeof(r.graph_key), "graph key storage");
        if (!r.graph_exec || std::memcmp(r.graph_key, &key, sizeof key) != 0) {
            if (r.graph_exec) {
                PT_HIP(hipStreamSynchronize(st));
                PT_HIP(hipGraphExecDestroy(static_cast<hipGraphExec_t>(r.graph_exec)));
                r.graph_exec = nullptr;
            }
            hipGraph_t g = nullptr;
            PT_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            launch_frame(r, scene, fc, st);
            const hipError_t le = hipGetLastError();
            const hipError_t ce = hipStreamEndCapture(st, &g);
            if (le != hipSuccess) throw Error(PT_E_HIP, std::string("frame capture: ") + hipGetErrorString(le));
            if (ce != hipSuccess) throw Error(PT_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
            hipGraphExec_t ex = nullptr;
            const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) throw Error(PT_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
            r.graph_exec = ex;
            std::memcpy(r.graph_key, &key, sizeof key);
        }
        PT_HIP(hipGraphLaunch(static_cast<hipGraphExec_t>(r.graph_exec), st));
    }
    r.spp += npasses;
    r.last_levels = fc.bounces + 1;
}

uint64_t render_last_rays(Renderer& r, hipStream_t st) {
    if (r.last_levels == 0) return 0;
    uint32_t h[64];
    PT_HIP(hipMemcpyAsync(h, r.counters.as<void>(), r.last_levels * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    PT_HIP(hipStreamSynchronize(st));
    uint64_t total = 0;
    for (uint32_t k = 0; k < r.last_levels; ++k) total += h[k];
    return total;
}

RenderTuning& render_tuning() {
    static RenderTuning t;
    return t;
}

}  // namespace pt
