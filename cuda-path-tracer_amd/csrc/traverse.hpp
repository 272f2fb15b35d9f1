// Closest-hit traversal over the 64-byte node records (reference: CUDA::trace,
// trace.h:73-169).  Visiting order is the reference's: test both children,
// descend into the nearer (left on ties), defer the farther internal child on
// a stack.  For the lowest-id tie rule the culling keeps children at exactly
// the best distance, and a triangle replaces the best hit when nearer, or
// equally near with a lower original id.
#pragma once

#include <algorithm>

#include "devmath.hpp"
#include "pt_api.h"

namespace pt {

struct TraceHit {
    float t;        // FLT_MAX on a miss
    uint32_t id;    // original triangle id, kInvalid on a miss
    float u, v;     // barycentrics of v1, v2
};

// Measurement counts of one lane: node visits, box tests, triangle tests and
// the deepest deferred-node stack reached.
struct TraceCount {
    uint32_t iters, boxes, tris, depth;
};

// Stack bound: every deferred node is the sibling of a distinct ancestor of
// the current node, and a radix tree over < 2^30 augmented 62-bit keys is at
// most 63 internal levels deep.  A push past the bound is not made; it sets
// kStatusStackOverflow in the caller's status word (which the host reports
// as PT_E_INTERNAL), so a violated bound cannot pass silently.
constexpr int kStackMax = 64;
constexpr uint32_t kStatusStackOverflow = PT_STATUS_STACK_OVERFLOW;

__device__ __forceinline__ void raise_status(uint32_t* status, uint32_t bit) {
    if (status) atomicOr(status, bit);
}

template <bool COUNT, bool FINITE>
__device__ __forceinline__ float child_box(V3 o, V3 dinv, float a, float b, float c, float d,
                                           float e, float f) {
    return FINITE ? slab_test_finite(o, dinv, a, b, c, d, e, f) : slab_test(o, dinv, a, b, c, d, e, f);
}

__device__ __forceinline__ float tri_hit(const float4* __restrict__ tris, uint32_t k, V3 o, V3 d,
                                         float& u, float& v, uint32_t& id) {
    float4 a = tris[3 * static_cast<size_t>(k)];
    const float4 b = tris[3 * static_cast<size_t>(k) + 1];
    const float4 c = tris[3 * static_cast<size_t>(k) + 2];
    // Keep p0's load beside the edges' (the compiler would otherwise sink it
    // below the determinant test, a second memory round trip).
    asm volatile("" : "+v"(a.x), "+v"(a.y), "+v"(a.z));
    id = __float_as_uint(a.w);
    return mt_test(o, d, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(c.x, c.y, c.z), u, v);
}

template <bool COUNT, bool FINITE>
__device__ __forceinline__ TraceHit traverse_impl(const float4* __restrict__ nodes,
                                                  const float4* __restrict__ tris, V3 o, V3 d,
                                                  V3 dinv, TraceCount& cnt, uint32_t* status,
                                                  uint32_t root = 0) {
    float minT = kMinTInit;
    uint32_t best = kInvalid;
    float bu = 0.0f, bv = 0.0f;
    uint32_t stack[kStackMax];
    int sp = 0;
    uint32_t node = root;
    for (;;) {
        for (;;) {
            const float4* rec = nodes + 4 * static_cast<size_t>(node);
            const float4 r0 = rec[0], r1 = rec[1], r2 = rec[2], r3 = rec[3];
            const uint32_t w = __float_as_uint(r3.x);
            const uint32_t l = w >> 2;
            const bool in0 = (w & 1u) != 0, in1 = (w & 2u) != 0;
            float h0, h1, u0 = 0.0f, v0 = 0.0f, u1 = 0.0f, v1 = 0.0f;
            uint32_t id0 = kInvalid, id1 = kInvalid;
            if (COUNT) cnt.iters++;
            // Both child boxes for every lane (a leaf child's value is
            // discarded), then the leaf children's triangle tests in one loop.
            h0 = child_box<COUNT, FINITE>(o, dinv, r0.x, r0.y, r0.z, r0.w, r1.x, r1.y);
            h1 = child_box<COUNT, FINITE>(o, dinv, r1.z, r1.w, r2.x, r2.y, r2.z, r2.w);
            if (COUNT) cnt.boxes += static_cast<uint32_t>(in0) + static_cast<uint32_t>(in1);
            const uint32_t nleaf = static_cast<uint32_t>(!in0) + static_cast<uint32_t>(!in1);
            for (uint32_t j = 0; j < nleaf; ++j) {
                const bool first = j == 0 && !in0;
                float u, v;
                uint32_t id;
                const float h = tri_hit(tris, first ? l : l + 1, o, d, u, v, id);
                if (COUNT) cnt.tris++;
                if (first) {
                    h0 = h; u0 = u; v0 = v; id0 = id;
                } else {
                    h1 = h; u1 = u; v1 = v; id1 = id;
                }
            }
            if (!(h0 <= minT) && !(h1 <= minT)) break;
            const bool sw = h1 < h0;  // left child wins ties (trace.h:125)
            const float hc = sw ? h1 : h0, hf = sw ? h0 : h1;
            const uint32_t cc = sw ? l + 1 : l, cf = sw ? l : l + 1;
            const bool ic = sw ? in1 : in0, jf = sw ? in0 : in1;
            const float uc = sw ? u1 : u0, vc = sw ? v1 : v0, uf = sw ? u0 : u1, vf = sw ? v0 : v1;
            const uint32_t idc = sw ? id1 : id0, idf = sw ? id0 : id1;
            if (ic) {
                node = cc;
                if (hf <= minT) {
                    if (jf) {
                        if (sp < kStackMax) {
                            stack[sp++] = cf;
                            if (COUNT) cnt.depth = max(cnt.depth, static_cast<uint32_t>(sp));
                        } else {
                            raise_status(status, kStatusStackOverflow);
                        }
                    } else if (hf < minT || idf < best) {
                        minT = hf; best = idf; bu = uf; bv = vf;
                    }
                }
                continue;
            }
            if (hc < minT || idc < best) {
                minT = hc; best = idc; bu = uc; bv = vc;
            }
            if (hf <= minT) {
                if (jf) {
                    node = cf;
                    continue;
                }
                if (hf < minT || idf < best) {
                    minT = hf; best = idf; bu = uf; bv = vf;
                }
            }
            break;
        }
        if (sp == 0) break;
        node = stack[--sp];
    }
    TraceHit hit;
    hit.id = best;
    hit.t = best == kInvalid ? kFltMax : minT;
    hit.u = best == kInvalid ? 0.0f : bu;
    hit.v = best == kInvalid ? 0.0f : bv;
    return hit;
}

// Entry for one ray over a BVH of ntri triangles.
template <bool COUNT>
__device__ __forceinline__ TraceHit trace_ray(const float4* __restrict__ nodes,
                                              const float4* __restrict__ tris, uint32_t ntri, V3 o,
                                              V3 d, TraceCount& cnt, uint32_t* status) {
    TraceHit miss{kFltMax, kInvalid, 0.0f, 0.0f};
    if (ntri == 0) return miss;
    if (ntri == 1) {
        float u = 0.0f, v = 0.0f;
        uint32_t id;
        if (COUNT) cnt.tris++;
        const float h = tri_hit(tris, 0, o, d, u, v, id);
        if (h <= kMinTInit) return TraceHit{h, id, u, v};
        return miss;
    }
    const V3 dinv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    const bool finite = __builtin_fabsf(dinv.x) <= kFltMax && __builtin_fabsf(dinv.y) <= kFltMax &&
                        __builtin_fabsf(dinv.z) <= kFltMax;
    if (finite) return traverse_impl<COUNT, true>(nodes, tris, o, d, dinv, cnt, status);
    return traverse_impl<COUNT, false>(nodes, tris, o, d, dinv, cnt, status);
}

// One ray in the scene whose triangles (and internal nodes) occupy slots
// [b, e) of a batched BVH; a single scene has b = 0.
__device__ __forceinline__ TraceHit trace_ray_in(const float4* __restrict__ nodes,
                                                 const float4* __restrict__ tris, uint32_t b, uint32_t e,
                                                 V3 o, V3 d, uint32_t* status) {
    TraceHit miss{kFltMax, kInvalid, 0.0f, 0.0f};
    TraceCount cnt{0, 0, 0, 0};
    if (e - b == 0) return miss;
    if (e - b == 1) {
        float u = 0.0f, v = 0.0f;
        uint32_t id;
        const float h = tri_hit(tris, b, o, d, u, v, id);
        if (h <= kMinTInit) return TraceHit{h, id, u, v};
        return miss;
    }
    const V3 dinv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    const bool finite = __builtin_fabsf(dinv.x) <= kFltMax && __builtin_fabsf(dinv.y) <= kFltMax &&
                        __builtin_fabsf(dinv.z) <= kFltMax;
    if (finite) return traverse_impl<false, true>(nodes, tris, o, d, dinv, cnt, status, b);
    return traverse_impl<false, false>(nodes, tris, o, d, dinv, cnt, status, b);
}

// Where a lane keeps a shadow ray's deferred nodes: its column of an LDS
// stack (entry k at lds[k * stride], the first `depth` entries) and a
// private array for the rest; lds == nullptr: a private array throughout.
struct ShadowStack {
    uint32_t* lds = nullptr;
    uint32_t stride = 0;
    int depth = 0;
    uint32_t* spill = nullptr;
};

// Whether a ray hits any triangle of the scene in [b, e) that the closest-hit
// search would accept (the shadow-ray question of computeL,
// LightAlgorithms.h:12-20: closest-hit t < FLT_MAX).  Until its first hit
// the closest-hit search culls nothing (its bound is still the initial
// one), so it visits the same boxes as this search does and finds a hit
// exactly when this one does; this one stops at the first.  LDSS: the
// deferred nodes go to `ss` (a lane whose own traversal stack is empty
// lends it to its shadow rays) instead of a private array.
template <bool LDSS = false>
__device__ __forceinline__ bool any_hit_in(const float4* __restrict__ nodes, const float4* __restrict__ tris,
                                          uint32_t b, uint32_t e, V3 o, V3 d, uint32_t* status,
                                          const ShadowStack& ss = ShadowStack{}) {
    float u, v;
    uint32_t id;
    if (e - b == 0) return false;
    if (e - b == 1) return tri_hit(tris, b, o, d, u, v, id) <= kMinTInit;
    const V3 dinv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    uint32_t stack[LDSS ? 1 : kStackMax];
    int sp = 0;
    uint32_t node = b;
    for (;;) {
        const float4* rec = nodes + 4 * static_cast<size_t>(node);
        const float4 r0 = rec[0], r1 = rec[1], r2 = rec[2], r3 = rec[3];
        const uint32_t w = __float_as_uint(r3.x);
        const uint32_t l = w >> 2;
        const bool in0 = (w & 1u) != 0, in1 = (w & 2u) != 0;
        const float h0 = in0 ? slab_test(o, dinv, r0.x, r0.y, r0.z, r0.w, r1.x, r1.y)
                             : tri_hit(tris, l, o, d, u, v, id);
        const float h1 = in1 ? slab_test(o, dinv, r1.z, r1.w, r2.x, r2.y, r2.z, r2.w)
                             : tri_hit(tris, l + 1, o, d, u, v, id);
        if ((!in0 && h0 <= kMinTInit) || (!in1 && h1 <= kMinTInit)) return true;
        const bool go0 = in0 && h0 <= kMinTInit, go1 = in1 && h1 <= kMinTInit;
        if (go0 && go1) {
            if constexpr (LDSS) {
                if (sp < ss.depth) ss.lds[static_cast<uint32_t>(sp) * ss.stride] = l + 1;
                else if (sp < kStackMax) ss.spill[sp - ss.depth] = l + 1;
                else raise_status(status, kStatusStackOverflow);
                if (sp < kStackMax) ++sp;
            } else {
                if (sp < kStackMax) stack[sp++] = l + 1;
                else raise_status(status, kStatusStackOverflow);
            }
            node = l;
        } else if (go0 || go1) {
            node = go0 ? l : l + 1;
        } else {
            if (sp == 0) return false;
            --sp;
            if constexpr (LDSS) node = sp < ss.depth ? ss.lds[static_cast<uint32_t>(sp) * ss.stride] : ss.spill[sp - ss.depth];
            else node = stack[sp];
        }
    }
}

// ---- persistent traversal -------------------------------------------------
//
// One lane = one ray at a time, advanced one node visit per loop iteration;
// a lane whose ray has finished takes the next ray from its wave's pool, so
// long rays do not hold a whole wave's lanes idle.  The visiting order per ray
// is exactly traverse_impl's.  The deferred-node stack keeps its first
// kLdsStack entries in LDS ([depth][thread] so a wave's lanes hit distinct
// banks) and the rest in a private array (scratch), touched only by deep rays.

constexpr int kLdsStack = 16;  // default LDS depth (tuned: 8, 16 or 32 entries)
constexpr int kPersistThreads = 256;

struct RayState {
    V3 o, d, dinv;
    float minT;
    uint32_t best;
    float bu, bv;
    uint32_t node;
    int sp;
    bool finite;
};

template <bool COUNT, int LDS, bool FINITE, int COLS = kPersistThreads>
__device__ __forceinline__ bool visit_step(const float4* __restrict__ nodes,
                                           const float4* __restrict__ tris, RayState& r,
                                           uint32_t (*s_stack)[COLS],
                                           uint32_t* spill, TraceCount& cnt, uint32_t* status) {
    const float4* rec = nodes + 4 * static_cast<size_t>(r.node);
    const float4 r0 = rec[0], r1 = rec[1], r2 = rec[2], r3 = rec[3];
    const uint32_t w = __float_as_uint(r3.x);
    const uint32_t l = w >> 2;
    const bool in0 = (w & 1u) != 0, in1 = (w & 2u) != 0;
    if (COUNT) cnt.iters++;
    // Both child boxes in one code path for every lane (a leaf child's slot
    // holds no box; its value is discarded).
    const float h0 = child_box<COUNT, FINITE>(r.o, r.dinv, r0.x, r0.y, r0.z, r0.w, r1.x, r1.y);
    const float h1 = child_box<COUNT, FINITE>(r.o, r.dinv, r1.z, r1.w, r2.x, r2.y, r2.z, r2.w);
    if (COUNT) cnt.boxes += static_cast<uint32_t>(in0) + static_cast<uint32_t>(in1);
    // Leaf children first.  The reference orders a node's two children by
    // distance and lets a leaf replace the best hit when nearer, or equally
    // near with a lower id (trace.h:125-160): over the leaves of one node
    // that is the lexicographic minimum of (t, id), whatever their order.
    // Taking it before deciding on the internal children visits the same
    // nodes: an internal child nearer than a leaf stays within the updated
    // best distance, and one farther than a leaf is culled against the
    // distance the leaf set, as in the reference.
    const uint32_t nleaf = static_cast<uint32_t>(!in0) + static_cast<uint32_t>(!in1);
    for (uint32_t j = 0; j < nleaf; ++j) {
        float u, v;
        uint32_t id;
        const float t = tri_hit(tris, (j == 0 && !in0) ? l : l + 1, r.o, r.d, u, v, id);
        if (COUNT) cnt.tris++;
        if (t < r.minT || (t == r.minT && id < r.best)) {
            r.minT = t; r.best = id; r.bu = u; r.bv = v;
        }
    }
    const bool g0 = in0 && h0 <= r.minT, g1 = in1 && h1 <= r.minT;
    if (g0 || g1) {
        const bool sw = g1 && (!g0 || h1 < h0);  // left child wins ties (trace.h:125)
        r.node = sw ? l + 1 : l;
        if (g0 && g1) {
            const uint32_t far = sw ? l : l + 1;
            const int sp = r.sp;
            if (sp < LDS) {
                s_stack[sp][threadIdx.x] = far;
                r.sp = sp + 1;
            } else if (sp < kStackMax) {
                spill[sp - LDS] = far;
                r.sp = sp + 1;
            } else {
                raise_status(status, kStatusStackOverflow);
            }
            if (COUNT) cnt.depth = max(cnt.depth, static_cast<uint32_t>(r.sp));
        }
        return true;
    }
    if (r.sp == 0) return false;
    const int sp = r.sp - 1;
    // The LDS entry is read by every lane and passed through an empty asm
    // statement, so the compiler keeps the LDS and scratch reads apart (a
    // select of the two addresses would become one slower flat load).
    uint32_t next = s_stack[sp < LDS ? sp : LDS - 1][threadIdx.x];
    asm volatile("" : "+v"(next));
    if (sp >= LDS) next = spill[sp - LDS];
    r.node = next;
    r.sp = sp;
    return true;
}

// Ray hand-out.  The rays [0, nrays) are split into nq contiguous slices
// (nq <= 64, about one per 16K rays), each with its own counter 256 bytes
// from the next: one address sustains only about 90 M returning atomics/s on
// MI355X, so a single counter caps how small the pools can be, and large
// pools leave waves unevenly loaded at the end.  A wave starts on slice
// (wave index mod nq), takes pools from it and, once it is used up, moves on
// to the next slice that is neither marked in a 64-bit mask of used-up
// slices nor found used up by its own probes.  A slice is marked by the one
// wave whose pool reaches its end, after every ray of it has been handed out;
// the mask only saves probes.  A wave stops when every slice is marked or
// found used up.  Slices hold at least 16K rays, so none is empty.  Counters
// and mask live in one queue block of kQueueWords words that the caller
// zeroes (zero_queue) before the launch.
constexpr uint32_t kQueueSlices = 64;
constexpr uint32_t kQueueStride = 64;  // words between counters (256 bytes)
constexpr uint32_t kQueueMask = kQueueSlices * kQueueStride;  // word offset of the mask
constexpr uint32_t kQueueWords = kQueueMask + 64;
constexpr uint32_t kSliceRays = 16384;

__device__ __forceinline__ uint32_t queue_slices(uint32_t nrays) {
    return min(kQueueSlices, max(1u, nrays / kSliceRays));
}


// Zeroes the counters and the mask of `levels` consecutive queue blocks with
// the threads of one workgroup.
__device__ __forceinline__ void zero_queue(uint32_t* __restrict__ queue, uint32_t levels) {
    const uint32_t per = kQueueSlices + 2;
    for (uint32_t i = threadIdx.x; i < levels * per; i += blockDim.x) {
        const uint32_t lvl = i / per, k = i % per;
        queue[lvl * kQueueWords + (k < kQueueSlices ? k * kQueueStride : kQueueMask + (k - kQueueSlices))] = 0u;
    }
}

// Persistent closest-hit loop over rays [0, nrays), handed out in per-wave
// pools from the queue block `queue` (see above): pool_large rays (tuned,
// 256 by default) when there are at least four per lane of the grid, else 64
// so that a short queue still spreads over every wave.  Idle lanes are
// refilled when fewer than `refill` (1..64) lanes of the wave are busy.
// load(i, o, d) fetches ray i; store(i, hit) writes its result.  load returns
// the ray's root internal node, or kInvalid for a scene without one (fewer
// than two triangles); then `single` names its one triangle (or kInvalid) and
// the ray is answered at once without entering the loop.
constexpr uint32_t kRayPoolSmall = 64;

// The tuned pool size a launch of `blocks` workgroups over at most `nmax` rays
// may use: every wave adds at most one pool to a slice counter past the
// slice's end before the slice is marked, so nmax + (waves + 1) * pool must
// stay below 2^32.
inline uint32_t pool_for(uint32_t pool, uint32_t nmax, uint32_t blocks) {
    const uint64_t waves = static_cast<uint64_t>(blocks) * (kPersistThreads / 64) + 1;
    const uint64_t room = (0xffffffffull - nmax) / waves;
    return static_cast<uint32_t>(std::max<uint64_t>(kRayPoolSmall, std::min<uint64_t>(pool, room)));
}

template <bool COUNT, int LDS, class Load, class Store>
__device__ __forceinline__ void persistent_trace(const float4* __restrict__ nodes,
                                                 const float4* __restrict__ tris, uint32_t nrays,
                                                 uint32_t* __restrict__ queue, uint32_t* status,
                                                 uint32_t refill, uint32_t pool_large,
                                                 uint32_t (*s_stack)[kPersistThreads], Load load,
                                                 Store store, TraceCount& cnt) {
    // Waves past one per kRayPoolSmall rays leave at once: every wave that
    // stays drains the queue with the others, and on a short queue (a late
    // render level, a cache frame's training levels) the extra waves would
    // only add returning atomics on its few slice counters.
    const uint32_t wave_id = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / 64u) + threadIdx.x / 64u);
    if (wave_id > 0u && static_cast<uint64_t>(wave_id) * kRayPoolSmall >= nrays) return;
    uint32_t spill[kStackMax - LDS];
    const int lane = threadIdx.x & 63;
    const uint64_t lt = (1ull << lane) - 1ull;
    RayState r;
    bool active = false;
    uint32_t ray = 0;
    uint32_t pool_next = 0, pool_end = 0;  // wave-uniform
    bool drained = false;                  // wave-uniform: no rays left to hand out
    const uint32_t pool =
        nrays / 4u >= gridDim.x * blockDim.x ? pool_large : kRayPoolSmall;
    const uint32_t nq = queue_slices(nrays);
    const uint32_t per = nrays / nq, extra = nrays % nq;  // slice q: per rays, +1 if q < extra
    const uint64_t all = nq == 64u ? ~0ull : (1ull << nq) - 1ull;
    uint64_t* mask = reinterpret_cast<uint64_t*>(queue + kQueueMask);
    // Wave-uniform state is read through readfirstlane so that it lives in
    // scalar registers (the compiler cannot tell that threadIdx.x / 64 is
    // uniform over a wave).
    uint32_t q = wave_id % nq;
    uint64_t seen = ~all;  // wave-uniform: slices this wave found used up
    bool used_up = false;  // wave-uniform: the refill loop found slice q used up
    for (;;) {
        const uint64_t busy = __ballot(active);
        if (!drained && static_cast<uint32_t>(__popcll(busy)) < refill) {
            uint64_t idle = ~busy;
            while (idle != 0ull) {
                if (pool_next >= pool_end) {
                    const uint32_t beg = q * per + min(q, extra);
                    const uint32_t len = per + (q < extra ? 1u : 0u);
                    uint32_t base = 0;
                    if (lane == 0) {
                        base = atomicAdd(queue + q * kQueueStride, pool);
                        // Only the pool that reaches the slice's end marks it:
                        // one atomic per slice (every wave marking the slices
                        // it finds used up would serialise on the mask).
                        if (base < len && len - base <= pool) atomicOr(mask, 1ull << q);
                    }
                    base = __builtin_amdgcn_readfirstlane(base);
                    if (base >= len) {
                        used_up = true;
                        break;
                    }
                    pool_next = beg + base;
                    pool_end = beg + base + min(pool, len - base);
                }
                const uint32_t avail = pool_end - pool_next;
                const uint32_t rank = static_cast<uint32_t>(__popcll(idle & lt));
                const bool take = ((idle >> lane) & 1ull) && rank < avail;
                const uint32_t took = min(static_cast<uint32_t>(__popcll(idle)), avail);
                if (take) {
                    ray = pool_next + rank;
                    uint32_t single = kInvalid;
                    const uint32_t root = load(ray, r.o, r.d, single);
                    if (root == kInvalid) {
                        TraceHit h{kFltMax, kInvalid, 0.0f, 0.0f};
                        if (single != kInvalid) {
                            float u = 0.0f, v = 0.0f;
                            uint32_t id;
                            if (COUNT) cnt.tris++;
                            const float t = tri_hit(tris, single, r.o, r.d, u, v, id);
                            if (t <= kMinTInit) h = TraceHit{t, id, u, v};
                        }
                        store(ray, h);
                    } else {
                        r.dinv = v3(1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z);
                        r.finite = __builtin_fabsf(r.dinv.x) <= kFltMax &&
                                   __builtin_fabsf(r.dinv.y) <= kFltMax &&
                                   __builtin_fabsf(r.dinv.z) <= kFltMax;
                        r.minT = kMinTInit;
                        r.best = kInvalid;
                        r.bu = 0.0f;
                        r.bv = 0.0f;
                        r.node = root;
                        r.sp = 0;
                        active = true;
                    }
                }
                pool_next += took;
                idle = ~__ballot(active);
            }
            if (used_up) {
                // Slice q is used up: move to the next slice not known to be
                // used up (kept out of the refill loop above, which stays
                // short).  Every visit here adds a slice to `seen`, so after
                // at most nq of them the wave has a slice with rays or knows
                // that all are used up.
                used_up = false;
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
        // With no busy lane the wave stops once no slice has rays left; until
        // then the next pass probes the slice just chosen.
        if (__ballot(active) == 0ull && drained) break;
        if (active && !(r.finite ? visit_step<COUNT, LDS, true>(nodes, tris, r, s_stack, spill, cnt, status)
                                 : visit_step<COUNT, LDS, false>(nodes, tris, r, s_stack, spill, cnt, status))) {
            TraceHit h;
            h.id = r.best;
            h.t = r.best == kInvalid ? kFltMax : r.minT;
            h.u = r.best == kInvalid ? 0.0f : r.bu;
            h.v = r.best == kInvalid ? 0.0f : r.bv;
            store(ray, h);
            active = false;
        }
    }
}

// ---- root filter ---------------------------------------------------------
//
// A batch of rays that start outside the scene's box (the closest-hit
// microbench's shell rays) mostly leaves after the root visit.  In the
// persistent loop above such a ray takes a lane through a refill and one
// visit while its neighbours walk deep paths, so most of the wave's issue
// slots go to lanes that do nothing.  The filtered variant first runs the
// root visit of 64 fresh rays at once on every lane of the wave (busy lanes
// keep their own ray's state in other registers), answers the rays that end
// there, and queues the rest ("survivors") in a per-wave LDS ring; idle
// lanes take survivors from the ring and trace them from the root as above.
// Results are the same bits: a ray that ends at the root visit ends exactly
// as the full traversal would, and a survivor is traced in full.

constexpr uint32_t kSurvCap = 128;  // survivor ring entries per wave

// The root visit alone (visit_step's decisions without the stack): leaf
// children fold into the best hit; returns whether an internal child is to
// be visited.
template <bool FINITE>
__device__ __forceinline__ bool root_visit(const float4* __restrict__ nodes, const float4* __restrict__ tris,
                                           RayState& r) {
    const float4* rec = nodes + 4 * static_cast<size_t>(r.node);
    const float4 r0 = rec[0], r1 = rec[1], r2 = rec[2], r3 = rec[3];
    const uint32_t w = __float_as_uint(r3.x);
    const uint32_t l = w >> 2;
    const bool in0 = (w & 1u) != 0, in1 = (w & 2u) != 0;
    const float h0 = child_box<false, FINITE>(r.o, r.dinv, r0.x, r0.y, r0.z, r0.w, r1.x, r1.y);
    const float h1 = child_box<false, FINITE>(r.o, r.dinv, r1.z, r1.w, r2.x, r2.y, r2.z, r2.w);
    const uint32_t nleaf = static_cast<uint32_t>(!in0) + static_cast<uint32_t>(!in1);
    for (uint32_t j = 0; j < nleaf; ++j) {
        float u, v;
        uint32_t id;
        const float t = tri_hit(tris, (j == 0 && !in0) ? l : l + 1, r.o, r.d, u, v, id);
        if (t < r.minT || (t == r.minT && id < r.best)) {
            r.minT = t; r.best = id; r.bu = u; r.bv = v;
        }
    }
    return (in0 && h0 <= r.minT) || (in1 && h1 <= r.minT);
}

__device__ __forceinline__ void init_ray(RayState& r, uint32_t root) {
    r.dinv = v3(1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z);
    r.finite = __builtin_fabsf(r.dinv.x) <= kFltMax && __builtin_fabsf(r.dinv.y) <= kFltMax &&
               __builtin_fabsf(r.dinv.z) <= kFltMax;
    r.minT = kMinTInit;
    r.best = kInvalid;
    r.bu = 0.0f;
    r.bv = 0.0f;
    r.node = root;
    r.sp = 0;
}

__device__ __forceinline__ TraceHit final_hit(const RayState& r) {
    TraceHit h;
    h.id = r.best;
    h.t = r.best == kInvalid ? kFltMax : r.minT;
    h.u = r.best == kInvalid ? 0.0f : r.bu;
    h.v = r.best == kInvalid ? 0.0f : r.bv;
    return h;
}

// persistent_trace with the root filter (no counting).  s_surv: this wave's
// kSurvCap-entry ring in LDS.  The hand-out protocol (pools from slice
// counters, used-up mask, `seen`) is persistent_trace's; a wave stops once
// every slice is used up, its ring is empty and no lane is busy.
template <int LDS, class Load, class Store>
__device__ __forceinline__ void persistent_trace_filtered(const float4* __restrict__ nodes,
                                                          const float4* __restrict__ tris, uint32_t nrays,
                                                          uint32_t* __restrict__ queue, uint32_t* status,
                                                          uint32_t refill, uint32_t pool_large,
                                                          uint32_t (*s_stack)[kPersistThreads],
                                                          uint32_t* __restrict__ s_surv, Load load, Store store) {
    const uint32_t wave_id = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / 64u) + threadIdx.x / 64u);
    if (wave_id > 0u && static_cast<uint64_t>(wave_id) * kRayPoolSmall >= nrays) return;
    uint32_t spill[kStackMax - LDS];
    TraceCount cnt{0, 0, 0, 0};
    const int lane = threadIdx.x & 63;
    const uint64_t lt = (1ull << lane) - 1ull;
    RayState r;
    bool active = false;
    uint32_t ray = 0;
    uint32_t pool_next = 0, pool_end = 0;  // wave-uniform
    bool drained = false;                  // wave-uniform: no rays left to hand out
    uint32_t qhead = 0, qtail = 0;         // wave-uniform: survivor ring positions
    const uint32_t pool =
        nrays / 4u >= gridDim.x * blockDim.x ? pool_large : kRayPoolSmall;
    const uint32_t nq = queue_slices(nrays);
    const uint32_t per = nrays / nq, extra = nrays % nq;
    const uint64_t all = nq == 64u ? ~0ull : (1ull << nq) - 1ull;
    uint64_t* mask = reinterpret_cast<uint64_t*>(queue + kQueueMask);
    uint32_t q = wave_id % nq;
    uint64_t seen = ~all;
    for (;;) {
        const uint32_t nbusy = static_cast<uint32_t>(__popcll(__ballot(active)));
        if (nbusy < refill && (!drained || qtail != qhead)) {
            // Filter rounds until the ring can fill every idle lane (or the
            // rays run out).  Each pass either consumes pool rays, takes a
            // new pool, or finds a slice used up (at most nq times).
            while (!drained && qtail - qhead < 64u - nbusy && qtail - qhead <= kSurvCap - 64u) {
                if (pool_next >= pool_end) {
                    const uint32_t beg = q * per + min(q, extra);
                    const uint32_t len = per + (q < extra ? 1u : 0u);
                    uint32_t base = 0;
                    if (lane == 0) {
                        base = atomicAdd(queue + q * kQueueStride, pool);
                        if (base < len && len - base <= pool) atomicOr(mask, 1ull << q);
                    }
                    base = __builtin_amdgcn_readfirstlane(base);
                    if (base >= len) {
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
                        continue;
                    }
                    pool_next = beg + base;
                    pool_end = beg + base + min(pool, len - base);
                }
                const uint32_t avail = min(64u, pool_end - pool_next);
                const uint32_t fr = pool_next + static_cast<uint32_t>(lane);
                bool surv = false;
                if (static_cast<uint32_t>(lane) < avail) {
                    RayState f;
                    uint32_t single = kInvalid;
                    const uint32_t root = load(fr, f.o, f.d, single);
                    if (root == kInvalid) {
                        TraceHit h{kFltMax, kInvalid, 0.0f, 0.0f};
                        if (single != kInvalid) {
                            float u = 0.0f, v = 0.0f;
                            uint32_t id;
                            const float t = tri_hit(tris, single, f.o, f.d, u, v, id);
                            if (t <= kMinTInit) h = TraceHit{t, id, u, v};
                        }
                        store(fr, h);
                    } else {
                        init_ray(f, root);
                        surv = f.finite ? root_visit<true>(nodes, tris, f) : root_visit<false>(nodes, tris, f);
                        if (!surv) store(fr, final_hit(f));
                    }
                }
                const uint64_t sm = __ballot(surv);
                if (surv) s_surv[(qtail + static_cast<uint32_t>(__popcll(sm & lt))) & (kSurvCap - 1u)] = fr;
                qtail += static_cast<uint32_t>(__popcll(sm));
                pool_next += avail;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // Idle lanes take survivors, in ring order.
            const uint64_t idle = ~__ballot(active);
            const uint32_t qn = qtail - qhead;
            const uint32_t rank = static_cast<uint32_t>(__popcll(idle & lt));
            if (((idle >> lane) & 1ull) && rank < qn) {
                ray = s_surv[(qhead + rank) & (kSurvCap - 1u)];
                uint32_t single = kInvalid;
                const uint32_t root = load(ray, r.o, r.d, single);
                init_ray(r, root);
                active = true;
            }
            qhead += min(static_cast<uint32_t>(__popcll(idle)), qn);
        }
        if (__ballot(active) == 0ull && drained && qtail == qhead) break;
        if (active && !(r.finite ? visit_step<false, LDS, true>(nodes, tris, r, s_stack, spill, cnt, status)
                                 : visit_step<false, LDS, false>(nodes, tris, r, s_stack, spill, cnt, status))) {
            store(ray, final_hit(r));
            active = false;
        }
    }
}

}  // namespace pt
