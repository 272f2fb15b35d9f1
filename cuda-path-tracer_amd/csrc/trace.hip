// Batch closest-hit queries (the reference's CUDA::trace, trace.h:73-169,
// exposed as a ray-batch entry point).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "context.hpp"
#include "traverse.hpp"

// Scratch of the allocation-free trace entry point: one ray queue block
// (slice counters and used-up mask, 16,640 bytes).  Every GPU the library
// runs on holds its own copy until the process ends.  In host code this name
// is only a handle for hipGetSymbolAddress: trace_scratch_setup resolves each
// device's copy once, and kernels receive that address.
static __device__ __align__(256) uint32_t scratch[pt::kQueueWords];

namespace pt {
namespace {

constexpr int kTraceThreads = 256;

// Wave-reduces the per-lane counts: one atomic per wave and counter
// (sums for visits, boxes and triangles, the maximum for the stack depth).
__device__ __forceinline__ void add_counts(const TraceCount& cnt, unsigned long long* counters) {
    unsigned long long a = cnt.iters, b = cnt.boxes, c = cnt.tris;
    uint32_t dmax = cnt.depth;
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off, 64);
        b += __shfl_down(b, off, 64);
        c += __shfl_down(c, off, 64);
        dmax = max(dmax, static_cast<uint32_t>(__shfl_down(static_cast<int>(dmax), off, 64)));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(counters + 0, a);
        atomicAdd(counters + 1, b);
        atomicAdd(counters + 2, c);
        atomicMax(counters + 3, static_cast<unsigned long long>(dmax));
    }
}

template <bool COUNT>
__global__ __launch_bounds__(kTraceThreads) void k_trace(
    const float4* __restrict__ nodes, const float4* __restrict__ tris, uint32_t ntri,
    const float* __restrict__ rays, uint32_t nrays, uint32_t* __restrict__ out_id,
    float* __restrict__ out_t, float2* __restrict__ out_uv,
    unsigned long long* __restrict__ counters, uint32_t* __restrict__ status) {
    TraceCount cnt{0, 0, 0, 0};
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nrays; i += stride) {
        const float* r = rays + 6 * static_cast<size_t>(i);
        const V3 o = v3(r[0], r[1], r[2]);
        const V3 d = v3(r[3], r[4], r[5]);
        const TraceHit h = trace_ray<COUNT>(nodes, tris, ntri, o, d, cnt, status);
        if (out_id) out_id[i] = h.id;
        if (out_t) out_t[i] = h.t;
        if (out_uv) out_uv[i] = make_float2(h.u, h.v);
    }
    if (COUNT) add_counts(cnt, counters);
}

// Persistent closest-hit kernel over an AoS ray batch (ntri >= 2); FILTER:
// the root filter (persistent_trace_filtered), without counting.
template <bool COUNT, int LDS, bool FILTER>
__global__ __launch_bounds__(kPersistThreads) void k_trace_persistent(
    const float4* __restrict__ nodes, const float4* __restrict__ tris, const float* __restrict__ rays,
    uint32_t nrays, uint32_t* __restrict__ out_id, float* __restrict__ out_t,
    float2* __restrict__ out_uv, uint32_t* __restrict__ ray_counter, uint32_t refill, uint32_t pool,
    unsigned long long* __restrict__ counters, uint32_t* __restrict__ status) {
    static_assert(!(COUNT && FILTER), "the filter's repeated root visit would change the visit counts");
    __shared__ uint32_t s_stack[LDS][kPersistThreads];
    auto load = [&](uint32_t i, V3& o, V3& d, uint32_t&) -> uint32_t {
        const float* rp = rays + 6 * static_cast<size_t>(i);
        o = v3(rp[0], rp[1], rp[2]);
        d = v3(rp[3], rp[4], rp[5]);
        return 0u;
    };
    auto store = [&](uint32_t i, const TraceHit& h) {
        if (out_id) out_id[i] = h.id;
        if (out_t) out_t[i] = h.t;
        if (out_uv) out_uv[i] = make_float2(h.u, h.v);
    };
    if constexpr (FILTER) {
        __shared__ uint32_t s_surv[kPersistThreads / 64][kSurvCap];
        persistent_trace_filtered<LDS>(nodes, tris, nrays, ray_counter, status, refill, pool, s_stack,
                                       s_surv[threadIdx.x / 64], load, store);
    } else {
        TraceCount cnt{0, 0, 0, 0};
        persistent_trace<COUNT, LDS>(nodes, tris, nrays, ray_counter, status, refill, pool, s_stack, load, store,
                                     cnt);
        if (COUNT) add_counts(cnt, counters);
    }
}

// Persistent closest hit over rays tagged with a scene of a batch: the ray
// starts at its scene's root (internal node off[s]).
template <int LDS>
__global__ __launch_bounds__(kPersistThreads) void k_trace_scenes(
    const float4* __restrict__ nodes, const float4* __restrict__ tris, const uint32_t* __restrict__ off,
    uint32_t nscenes, const float* __restrict__ rays, const uint32_t* __restrict__ scene, uint32_t nrays,
    uint32_t* __restrict__ out_id, float* __restrict__ out_t, float2* __restrict__ out_uv,
    uint32_t* __restrict__ ray_counter, uint32_t refill, uint32_t pool, uint32_t* __restrict__ status) {
    __shared__ uint32_t s_stack[LDS][kPersistThreads];
    TraceCount cnt{0, 0, 0, 0};
    persistent_trace<false, LDS>(
        nodes, tris, nrays, ray_counter, status, refill, pool, s_stack,
        [&](uint32_t i, V3& o, V3& d, uint32_t& single) -> uint32_t {
            const float* rp = rays + 6 * static_cast<size_t>(i);
            o = v3(rp[0], rp[1], rp[2]);
            d = v3(rp[3], rp[4], rp[5]);
            const uint32_t sc = scene[i];
            if (sc >= nscenes) return kInvalid;  // unknown scene: a miss
            const uint32_t b = off[sc], e = off[sc + 1];
            if (e - b >= 2) return b;
            if (e - b == 1) single = b;
            return kInvalid;
        },
        [&](uint32_t i, const TraceHit& h) {
            if (out_id) out_id[i] = h.id;
            if (out_t) out_t[i] = h.t;
            if (out_uv) out_uv[i] = make_float2(h.u, h.v);
        },
        cnt);
}

__global__ __launch_bounds__(64) void k_zero_queue(uint32_t* queue) { zero_queue(queue, 1u); }

template <bool COUNT>
void launch_trace(const DeviceScene& s, uint32_t* queue, const float* d_rays, uint32_t n, uint32_t* d_id,
                  float* d_t, float* d_uv, unsigned long long* d_counters, hipStream_t st) {
    if (n == 0) return;
    uint32_t* status = s.status.as<uint32_t>();
    if (s.bvh.n < 2) {
        const uint32_t blocks = static_cast<uint32_t>((static_cast<uint64_t>(n) + kTraceThreads - 1) /
                                                      kTraceThreads);
        hipLaunchKernelGGL(k_trace<COUNT>, dim3(blocks), dim3(kTraceThreads), 0, st,
                           s.bvh.nodes.as<float4>(), s.bvh.tris.as<float4>(), s.bvh.n, d_rays, n,
                           d_id, d_t, reinterpret_cast<float2*>(d_uv), d_counters, status);
        PT_HIP(hipGetLastError());
        return;
    }
    const TraceTuning& tune = trace_tuning();
    hipLaunchKernelGGL(k_zero_queue, dim3(1), dim3(64), 0, st, queue);
    const uint64_t want = (static_cast<uint64_t>(n) + kPersistThreads - 1) / kPersistThreads;
    const uint32_t blocks =
        static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(want, tune.blocks)));
    const bool filter = !COUNT && tune.filter != 0;
    auto kern = filter ? (tune.lds <= 8 ? k_trace_persistent<false, 8, true>
                                        : (tune.lds <= 16 ? k_trace_persistent<false, 16, true>
                                                          : k_trace_persistent<false, 32, true>))
                       : (tune.lds <= 8 ? k_trace_persistent<COUNT, 8, false>
                                        : (tune.lds <= 16 ? k_trace_persistent<COUNT, 16, false>
                                                          : k_trace_persistent<COUNT, 32, false>));
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(kPersistThreads), 0, st,
                       s.bvh.nodes.as<float4>(), s.bvh.tris.as<float4>(), d_rays, n, d_id, d_t,
                       reinterpret_cast<float2*>(d_uv), queue, std::min(64u, std::max(1u, tune.refill)),
                       pool_for(tune.pool, n, blocks), d_counters, status);
    PT_HIP(hipGetLastError());
}

}  // namespace

size_t trace_queue_bytes() { return kQueueWords * sizeof(uint32_t); }

namespace {
std::mutex g_scratch_mu;
std::vector<uint32_t*> g_scratch;  // per device: address of its `scratch`, or null
}  // namespace

void trace_scratch_setup(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0) throw Error(PT_E_INVALID_ARG, "negative device index");
    if (g_scratch.size() <= static_cast<size_t>(device)) g_scratch.resize(static_cast<size_t>(device) + 1, nullptr);
    if (g_scratch[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
        throw Error(PT_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = static_cast<uint32_t*>(p);
}

uint32_t* trace_scratch(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || static_cast<size_t>(device) >= g_scratch.size()) return nullptr;
    return g_scratch[device];
}

TraceTuning& trace_tuning() {
    static TraceTuning t = [] {
        TraceTuning v;
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            cus > 0)
            v.blocks = static_cast<uint32_t>(cus) * 8u;
        return v;
    }();
    return t;
}

void trace_closest_scenes(const DeviceScene& s, uint32_t* queue, const float* d_rays, const uint32_t* d_scene,
                          uint32_t n, uint32_t* d_id, float* d_t, float* d_uv, hipStream_t st) {
    if (n == 0) return;
    const TraceTuning& tune = trace_tuning();
    hipLaunchKernelGGL(k_zero_queue, dim3(1), dim3(64), 0, st, queue);
    const uint64_t want = (static_cast<uint64_t>(n) + kPersistThreads - 1) / kPersistThreads;
    const uint32_t blocks =
        static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(want, tune.blocks)));
    auto kern = tune.lds <= 8 ? k_trace_scenes<8> : (tune.lds <= 16 ? k_trace_scenes<16> : k_trace_scenes<32>);
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(kPersistThreads), 0, st,
                       s.bvh.nodes.as<float4>(), s.bvh.tris.as<float4>(), s.off.as<uint32_t>(),
                       s.nscenes(), d_rays, d_scene, n, d_id, d_t, reinterpret_cast<float2*>(d_uv),
                       queue, std::min(64u, std::max(1u, tune.refill)), pool_for(tune.pool, n, blocks),
                       s.status.as<uint32_t>());
    PT_HIP(hipGetLastError());
}

void trace_closest(const DeviceScene& s, uint32_t* queue, const float* d_rays, uint32_t n, uint32_t* d_id,
                   float* d_t, float* d_uv, hipStream_t st) {
    launch_trace<false>(s, queue, d_rays, n, d_id, d_t, d_uv, nullptr, st);
}

void trace_closest_counted(const DeviceScene& s, uint32_t* queue, const float* d_rays, uint32_t n,
                           uint32_t* d_id, float* d_t, float* d_uv, unsigned long long* d_counters,
                           hipStream_t st) {
    launch_trace<true>(s, queue, d_rays, n, d_id, d_t, d_uv, d_counters, st);
}

}  // namespace pt
