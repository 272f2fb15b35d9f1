// Progressive unidirectional path tracer as a wavefront pipeline
// (reference: raytrace_kernel + blinnPhongShadingPathTrace,
// cuda_path_tracer.cu:245-333, 385-421, which run one megakernel thread per
// pixel).  Per pass:
//   k_generate     primary rays for every pixel, queue 0, counters reset
//   k_extend(L)    closest hit for every ray of queue L
//   k_shade(L)     shade the hits, append surviving paths to queue L+1
//                  (wave-aggregated atomics), L = 0 .. bounces
//   k_accumulate   FP64 per-pixel sums += this pass's path radiance
// The queue lengths live on the device, so the host never waits inside a
// pass; the kernels loop over their queue with a fixed grid.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "context.hpp"
#include "sampling.hpp"
#include "traverse.hpp"

namespace pt {
namespace {

struct FrameParams {
    float dirs[12];
    float eye[3];
    uint32_t pass;
    uint32_t seedw;
    int w, h;
    uint32_t bounces;
    uint32_t rr_start;
    float intensity;
    uint32_t stratified;
};

// Structure-of-arrays queue of rays (origin, direction, pixel).
struct RayQueue {
    float *ox, *oy, *oz, *dx, *dy, *dz;
    uint32_t* pix;
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

__global__ __launch_bounds__(kThreads) void k_generate(const FrameParams fp_,
                                                       RayQueue q, PathState ps,
                                                       uint32_t* __restrict__ counters,
                                                       uint32_t* __restrict__ ray_counters,
                                                       uint32_t nlevels) {
    const FrameParams* fp = &fp_;
    const int w = fp->w, h = fp->h;
    const uint32_t npix = static_cast<uint32_t>(w) * static_cast<uint32_t>(h);
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid == 0) {
        counters[0] = npix;
        for (uint32_t k = 1; k <= nlevels; ++k) counters[k] = 0u;
        for (uint32_t k = 0; k <= nlevels; ++k) ray_counters[k] = 0u;
    }
    const V3 d00 = v3(fp->dirs[0], fp->dirs[1], fp->dirs[2]);
    const V3 d10 = v3(fp->dirs[3], fp->dirs[4], fp->dirs[5]);
    const V3 d11 = v3(fp->dirs[6], fp->dirs[7], fp->dirs[8]);
    const V3 d01 = v3(fp->dirs[9], fp->dirs[10], fp->dirs[11]);
    for (uint32_t p = gid; p < npix; p += gridDim.x * blockDim.x) {
        const int x = static_cast<int>(p % static_cast<uint32_t>(w));
        const int y = static_cast<int>(p / static_cast<uint32_t>(w));
        // Pixel-corner blend of the corner directions (cuda_path_tracer.cu:404-407).
        const float tx = static_cast<float>(x) / static_cast<float>(w);
        const float ty = static_cast<float>(y) / static_cast<float>(h);
        const V3 d = mix(mix(d00, d10, tx), mix(d01, d11, tx), ty);
        q.ox[p] = fp->eye[0];
        q.oy[p] = fp->eye[1];
        q.oz[p] = fp->eye[2];
        q.dx[p] = d.x;
        q.dy[p] = d.y;
        q.dz[p] = d.z;
        q.pix[p] = p;
        ps.cr[p] = 0.0f;
        ps.cg[p] = 0.0f;
        ps.cb[p] = 0.0f;
    }
}

// Closest hit for every ray of queue `level` (count on the device): the
// persistent traversal, rays handed out from ray_counters[level].
__global__ __launch_bounds__(kPersistThreads) void k_extend(const float4* __restrict__ nodes,
                                                            const float4* __restrict__ tris,
                                                            uint32_t ntri, RayQueue q, HitQueue hq,
                                                            const uint32_t* __restrict__ counters,
                                                            uint32_t* __restrict__ ray_counters,
                                                            uint32_t level, uint32_t refill) {
    __shared__ uint32_t s_stack[kLdsStack][kPersistThreads];
    const uint32_t count = counters[level];
    TraceCount cnt{0, 0, 0};
    auto load = [&](uint32_t i, V3& o, V3& d) {
        o = v3(q.ox[i], q.oy[i], q.oz[i]);
        d = v3(q.dx[i], q.dy[i], q.dz[i]);
    };
    auto store = [&](uint32_t i, const TraceHit& h) {
        hq.t[i] = h.t;
        hq.id[i] = h.id;
        hq.u[i] = h.u;
        hq.v[i] = h.v;
    };
    if (ntri < 2) {
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
            V3 o, d;
            load(i, o, d);
            store(i, trace_ray<false>(nodes, tris, ntri, o, d, cnt));
        }
        return;
    }
    persistent_trace<false>(nodes, tris, count, ray_counters + level, refill, s_stack, load, store, cnt);
}

__device__ __forceinline__ V3 ld3(const float* p) { return v3(p[0], p[1], p[2]); }

__global__ __launch_bounds__(kThreads) void k_shade(const FrameParams fp_,
                                                    const pt_triangle_info* __restrict__ infos,
                                                    RayQueue qin, HitQueue hq, RayQueue qout,
                                                    PathState ps, uint32_t* __restrict__ counters,
                                                    uint32_t level) {
    const FrameParams* fp = &fp_;
    const uint32_t count = counters[level];
    const uint32_t bounces = fp->bounces;
    const uint32_t rr_start = fp->rr_start;
    const float intensity = fp->intensity;
    const int lane = threadIdx.x & 63;
    const uint64_t lt = (1ull << lane) - 1ull;
    const uint32_t stride = gridDim.x * blockDim.x;
    // Wave-uniform trip count so every lane reaches the ballots below.
    for (uint32_t base = blockIdx.x * blockDim.x; base < count; base += stride) {
        const uint32_t i = base + threadIdx.x;
        bool cont = false;
        V3 no{0.0f, 0.0f, 0.0f}, nd{0.0f, 0.0f, 0.0f};
        uint32_t pix = 0;
        if (i < count && hq.id[i] != kInvalid) {
            pix = qin.pix[i];
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
                absorbed = v3(ps.ar[pix], ps.ag[pix], ps.ab[pix]);
                acc = v3(ps.cr[pix], ps.cg[pix], ps.cb[pix]);
                V3 L = v3(0.0f, 0.0f, 0.0f);
                if (emissive) {
                    // Emitter cosine (cuda_path_tracer.cu:304-312).
                    const float c = gmax(dot(neg(d), nrm), 0.0f);
                    L = scale(e, c);
                    albedo = v3(1.0f, 1.0f, 1.0f);
                }
                absorbed = mul(absorbed, albedo);
                acc = add(acc, scale(mul(absorbed, L), intensity));
            }
            // Next bounce index = level (bounce loop, cuda_path_tracer.cu:267).
            const uint32_t b = level;
            if (b < bounces) {
                const U4 rv = pcg4d(U4{pix, fp->pass, b, fp->seedw});
                float u1 = u24(rv.x), u2 = u24(rv.y);
                cont = true;
                if (b >= rr_start) {
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
                    if (fp->stratified != 0u && b == 0) {
                        const uint32_t p1 = fp->pass % 64u;
                        u1 = (static_cast<float>(p1 % 8u) + u1) / 8.0f;
                        u2 = (static_cast<float>(p1 / 8u) + u2) / 8.0f;
                    }
                    const float r = __builtin_sqrtf(u1);
                    float sn, cs;
                    sincos2pi(u2, sn, cs);
                    const V3 local = v3(r * cs, r * sn, __builtin_sqrtf(1.0f - u1));
                    nd = to_hemisphere(local, nrm);
                }
            }
            ps.ar[pix] = absorbed.x;
            ps.ag[pix] = absorbed.y;
            ps.ab[pix] = absorbed.z;
            ps.cr[pix] = acc.x;
            ps.cg[pix] = acc.y;
            ps.cb[pix] = acc.z;
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
                qout.pix[s] = pix;
            }
        }
    }
}

__global__ __launch_bounds__(kThreads) void k_accumulate(const FrameParams fp_,
                                                         PathState ps, double* __restrict__ accum) {
    const FrameParams* fp = &fp_;
    const uint32_t npix = static_cast<uint32_t>(fp->w) * static_cast<uint32_t>(fp->h);
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += gridDim.x * blockDim.x) {
        accum[3 * static_cast<size_t>(p) + 0] += static_cast<double>(ps.cr[p]);
        accum[3 * static_cast<size_t>(p) + 1] += static_cast<double>(ps.cg[p]);
        accum[3 * static_cast<size_t>(p) + 2] += static_cast<double>(ps.cb[p]);
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

Layout carve(Renderer& r) {
    const size_t np = static_cast<size_t>(r.w) * r.h;
    char* base = r.queues.as<char>();
    auto take = [&](size_t bytes) {
        char* p = base;
        base += (bytes + 255) & ~size_t(255);
        return p;
    };
    Layout L;
    for (int k = 0; k < 2; ++k) {
        L.q[k].ox = reinterpret_cast<float*>(take(np * 4));
        L.q[k].oy = reinterpret_cast<float*>(take(np * 4));
        L.q[k].oz = reinterpret_cast<float*>(take(np * 4));
        L.q[k].dx = reinterpret_cast<float*>(take(np * 4));
        L.q[k].dy = reinterpret_cast<float*>(take(np * 4));
        L.q[k].dz = reinterpret_cast<float*>(take(np * 4));
        L.q[k].pix = reinterpret_cast<uint32_t*>(take(np * 4));
    }
    L.hq.t = reinterpret_cast<float*>(take(np * 4));
    L.hq.id = reinterpret_cast<uint32_t*>(take(np * 4));
    L.hq.u = reinterpret_cast<float*>(take(np * 4));
    L.hq.v = reinterpret_cast<float*>(take(np * 4));
    L.ps.ar = reinterpret_cast<float*>(take(np * 4));
    L.ps.ag = reinterpret_cast<float*>(take(np * 4));
    L.ps.ab = reinterpret_cast<float*>(take(np * 4));
    L.ps.cr = reinterpret_cast<float*>(take(np * 4));
    L.ps.cg = reinterpret_cast<float*>(take(np * 4));
    L.ps.cb = reinterpret_cast<float*>(take(np * 4));
    return L;
}

size_t queue_bytes(size_t np) {
    const size_t a = ((np * 4 + 255) & ~size_t(255));
    return a * (2 * 7 + 4 + 6);
}

uint32_t grid_for(size_t work) {
    size_t b = (work + kThreads - 1) / kThreads;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    return static_cast<uint32_t>(b);
}

}  // namespace

void render_resize(Renderer& r, int w, int h, hipStream_t st) {
    if (w <= 0 || h <= 0) throw Error(PT_E_INVALID_ARG, "render size must be positive");
    if (r.w == w && r.h == h) return;
    const size_t np = static_cast<size_t>(w) * h;
    r.w = w;
    r.h = h;
    r.accum.reserve(np * 3 * sizeof(double));
    r.queues.reserve(queue_bytes(np));
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

void render_pass(Renderer& r, const DeviceScene& scene, uint64_t seed, hipStream_t st) {
    if (!scene.ready) throw Error(PT_E_NO_SCENE, "no scene uploaded");
    if (r.settings.bounces > 60) throw Error(PT_E_INVALID_ARG, "at most 60 bounces");
    const size_t np = static_cast<size_t>(r.w) * r.h;
    FrameParams frame;
    FrameParams* fh = &frame;
    std::memcpy(fh->dirs, r.dirs, sizeof fh->dirs);
    std::memcpy(fh->eye, r.eye, sizeof fh->eye);
    fh->pass = r.spp;
    fh->seedw = static_cast<uint32_t>(seed) ^ static_cast<uint32_t>(seed >> 32);
    fh->w = r.w;
    fh->h = r.h;
    fh->bounces = r.settings.bounces;
    fh->rr_start = r.settings.rr_start;
    fh->intensity = r.settings.indirect_intensity;
    fh->stratified = r.settings.stratified;
    const FrameParams& fd = frame;

    Layout L = carve(r);
    uint32_t* cnt = r.counters.as<uint32_t>();
    uint32_t* rcnt = cnt + 64;
    const TraceTuning& tune = trace_tuning();
    const uint32_t pgrid = std::max(1u, tune.blocks);
    const uint32_t refill = std::min(64u, std::max(1u, tune.refill));
    const uint32_t nlev = r.settings.bounces + 1;
    const uint32_t grid = grid_for(np);
    hipLaunchKernelGGL(k_generate, dim3(grid), dim3(kThreads), 0, st, fd, L.q[0], L.ps, cnt, rcnt, nlev);
    for (uint32_t lvl = 0; lvl < nlev; ++lvl) {
        hipLaunchKernelGGL(k_extend, dim3(std::min(pgrid, grid)), dim3(kPersistThreads), 0, st,
                           scene.bvh.nodes.as<float4>(), scene.bvh.tris.as<float4>(), scene.bvh.n,
                           L.q[lvl & 1], L.hq, cnt, rcnt, lvl, refill);
        hipLaunchKernelGGL(k_shade, dim3(grid), dim3(kThreads), 0, st, fd,
                           scene.infos.as<pt_triangle_info>(), L.q[lvl & 1], L.hq,
                           L.q[(lvl + 1) & 1], L.ps, cnt, lvl);
    }
    hipLaunchKernelGGL(k_accumulate, dim3(grid), dim3(kThreads), 0, st, fd, L.ps, r.accum.as<double>());
    PT_HIP(hipGetLastError());
    r.spp += 1;
}

}  // namespace pt
