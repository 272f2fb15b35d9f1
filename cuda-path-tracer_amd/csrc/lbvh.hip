// LBVH construction on gfx950 (reference: CUDA::BVHBuilder::build,
// BVHBuilder.cu:207-269, after Karras 2012), for one scene or for a batch of
// scenes at once.  Scenes are contiguous segments [off[s], off[s+1]) of the
// triangle array; every array below is indexed by global leaf slot, and the
// internal nodes of scene s are numbered off[s] + local index (slot
// off[s+1]-1 is unused), so each node's index lies inside its own leaf range.
// Child and parent indices are global; the parity dump converts to local.
//
//   k_morton        30-bit Morton key (the scene's box) + index + scene id,
//                   and the first sort pass's per-tile digit counts
//   k_sort_hist     \
//   k_sort_rowscan   } LSD radix sort, 10-bit digits, stable: 3 passes over
//   k_sort_scatter  /  the 30-bit key, then 1-2 over the scene id when
//                      batched (wave64 ballot match + LDS staging per
//                      4096-key tile)
//   k_radix_tree    Karras split per internal node, parents, counter reset,
//                   list of nodes whose leaf range crosses a refit block
//   k_refit_local   sorted triangle records + bottom-up AABBs inside each
//                   1024-leaf block (LDS counters/boxes, workgroup scope);
//                   extra workgroups stage the level-1 nodes' links
//   k_refit_mid     nodes inside each 65536-leaf block, one workgroup each;
//                   one more workgroup stages the level-2 nodes' links
//   k_refit_top     the remaining nodes, one workgroup
// Each node's traversal record is written whole by the thread that
// completes the node's box.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "context.hpp"
#include "devmath.hpp"
#include "traverse.hpp"

namespace pt {
namespace {

constexpr int kSortTile = 4096;                                // keys per sort tile
constexpr int kDigitBits = 10;                                 // 3 passes over a 30-bit key
constexpr uint32_t kRadix = 1u << kDigitBits;                  // 1024 digits
constexpr int kScatterThreads = 512;                           // scatter: 8 keys per thread, 8 waves per tile
// Refit levels.  A node's level is the smallest k whose block of sorted
// leaves contains the node's whole leaf range:
//   level 0: inside a 1024-leaf block    -> k_refit_local (LDS, one workgroup per block)
//   level 1: inside a 65536-leaf block   -> k_refit_mid (one workgroup per block)
//   level 2: anything larger             -> k_refit_top (one workgroup)
// A node's children have a level no higher than its own, so each phase only
// needs the previous phases' boxes.  Levels are kept in bits 30-31 of the
// node's refit counter.
constexpr int kRefitBlock = 1024;
constexpr uint32_t kMidBlock = 65536;
// Upper-level node lists.  A level-1 node goes to the list of the 1024-leaf
// block holding the start of its leaf range, a level-2 node to that of the
// 65536-leaf block; either way its range crosses that block's end, and a
// boundary is crossed only by the <= 64 ancestors of its two leaves' lowest
// common ancestor, so a list never holds more than kListCap nodes.  (Lists
// per small block keep the appends' atomics spread over many counters.)
// One 65536-leaf block's 64 level-1 lists together hold at most kMidCap.
constexpr uint32_t kListCap = 64;
constexpr uint32_t kMidCap = 64 * kListCap;
constexpr uint32_t kMaxMidBlocks = (1u << 30) / kMidBlock;  // builds hold < 2^30 triangles
constexpr uint32_t kLevelShift = 30;
constexpr uint32_t kCountMask = (1u << kLevelShift) - 1u;

// ---- Morton keys ------------------------------------------------------------

__device__ __forceinline__ uint32_t spread10(uint32_t a) {
    a &= 0x3FFu;
    a = (a | (a << 16)) & 0x030000FFu;
    a = (a | (a << 8)) & 0x0300F00Fu;
    a = (a | (a << 4)) & 0x030C30C3u;
    a = (a | (a << 2)) & 0x09249249u;
    return a;
}

// quantize (math_util.h:10-13) with CUDA's saturating float->int conversion.
__device__ __forceinline__ uint32_t quantize1024(float x) {
    float s = x * 1024.0f;
    if (!(s == s)) return 0u;
    if (s >= 1023.0f) return 1023u;
    if (s < 1.0f) return 0u;
    return static_cast<uint32_t>(static_cast<int32_t>(s));
}

// Scene of leaf slot i: the last s with off[s] <= i (off has nscenes + 1 entries).
__device__ __forceinline__ uint32_t scene_of_slot(const uint32_t* __restrict__ off, uint32_t nscenes,
                                                  uint32_t i) {
    uint32_t lo = 0, hi = nscenes - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// morton_encoder (BVHBuilder.cu:12-32): centroid ((v0 + v1) + v2) / 3, mapped
// into the scene's enclosing box with invScale computed on the host
// (box[6*s..] = min.xyz, invScale.xyz).  One workgroup per 4096-key sort
// tile, which also counts the tile's first-pass digits (key bits 0-9) into
// its row of tile_hist -- the first sort pass then starts at the scan.  Also
// zeroes the refit list counters.
constexpr int kMortonThreads = 1024;

__global__ __launch_bounds__(kMortonThreads) void k_morton(const pt_triangle* __restrict__ tris, uint32_t n,
                                                           const uint32_t* __restrict__ off, uint32_t nscenes,
                                                           const float* __restrict__ box,
                                                           uint32_t* __restrict__ keys,
                                                           uint32_t* __restrict__ vals,
                                                           uint32_t* __restrict__ scene_of,
                                                           uint32_t* __restrict__ tile_hist,
                                                           uint32_t* __restrict__ span_count,
                                                           uint32_t n_span_counts) {
    __shared__ uint32_t h[kRadix];
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = blockIdx.x * kMortonThreads + tid; i < n_span_counts; i += gridDim.x * kMortonThreads)
        span_count[i] = 0u;
    for (uint32_t d = tid; d < kRadix; d += kMortonThreads) h[d] = 0u;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
#pragma unroll
    for (int r = 0; r < kSortTile / kMortonThreads; ++r) {
        const uint32_t i = base + r * kMortonThreads + tid;
        if (i >= n) break;
        const uint32_t sc = scene_of_slot(off, nscenes, i);
        const float* bx = box + 6 * static_cast<size_t>(sc);
        const float* t = reinterpret_cast<const float*>(tris + i);
        float cx = ((t[0] + t[3]) + t[6]) / 3.0f;
        float cy = ((t[1] + t[4]) + t[7]) / 3.0f;
        float cz = ((t[2] + t[5]) + t[8]) / 3.0f;
        uint32_t qx = quantize1024((cx - bx[0]) * bx[3]);
        uint32_t qy = quantize1024((cy - bx[1]) * bx[4]);
        uint32_t qz = quantize1024((cz - bx[2]) * bx[5]);
        const uint32_t key = spread10(qx) | (spread10(qy) << 1) | (spread10(qz) << 2);
        keys[i] = key;
        vals[i] = i;
        scene_of[i] = sc;
        atomicAdd(&h[key & (kRadix - 1u)], 1u);
    }
    __syncthreads();
    for (uint32_t d = tid; d < kRadix; d += kMortonThreads)
        tile_hist[static_cast<size_t>(blockIdx.x) * kRadix + d] = h[d];
}

// ---- radix sort ---------------------------------------------------------------

// Digit of an element: bits [shift, shift+10) of its key, or -- in the
// batched build's last passes -- of the scene id of its original slot.
template <bool BY_SCENE>
__device__ __forceinline__ uint32_t sort_digit(uint32_t k, uint32_t v,
                                               const uint32_t* __restrict__ scene_of, int shift) {
    return ((BY_SCENE ? scene_of[v] : k) >> shift) & (kRadix - 1u);
}

// Per-tile digit histogram: tile_hist[tile * kRadix + digit] (one contiguous
// row per tile).
constexpr int kHistThreads = 1024;

template <bool BY_SCENE>
__global__ __launch_bounds__(kHistThreads) void k_sort_hist(const uint32_t* __restrict__ keys,
                                                            const uint32_t* __restrict__ vals,
                                                            const uint32_t* __restrict__ scene_of,
                                                            uint32_t n, int shift,
                                                            uint32_t num_tiles,
                                                            uint32_t* __restrict__ tile_hist) {
    __shared__ uint32_t h[kRadix];
    const int tid = threadIdx.x;
    for (int d = tid; d < static_cast<int>(kRadix); d += kHistThreads) h[d] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
#pragma unroll
    for (int r = 0; r < kSortTile / kHistThreads; ++r) {
        uint32_t i = base + r * kHistThreads + tid;
        if (i < n) atomicAdd(&h[sort_digit<BY_SCENE>(keys[i], BY_SCENE ? vals[i] : 0u, scene_of, shift)], 1u);
    }
    __syncthreads();
    for (int d = tid; d < static_cast<int>(kRadix); d += kHistThreads)
        tile_hist[static_cast<size_t>(blockIdx.x) * kRadix + d] = h[d];
}

// Exclusive scan of each digit's counts over the tiles (in place) and the
// digit's total.  One workgroup per 64 digits: lane = digit, so every load
// and store of a tile's 64 counts is one contiguous 256-byte access; the 16
// waves split the tiles into 16 ranges, and each wave's range starts at the
// sum of the earlier ranges.
constexpr int kScanDigits = 64;
constexpr int kScanWaves = 16;

__global__ __launch_bounds__(kScanWaves * 64) void k_sort_rowscan(uint32_t* __restrict__ tile_hist,
                                                                  uint32_t num_tiles,
                                                                  uint32_t* __restrict__ digit_total) {
    __shared__ uint32_t s_part[kScanWaves][kScanDigits];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const uint32_t d = blockIdx.x * kScanDigits + lane;
    const uint32_t per = (num_tiles + kScanWaves - 1) / kScanWaves;
    const uint32_t t0 = min(num_tiles, wave * per), t1 = min(num_tiles, t0 + per);
    // Loads in batches of kScanBatch, issued together (the counts come from
    // other XCDs' writes: each load is a far round trip).  The first batch
    // stays in registers for the second sweep (up to 256 tiles, 2^20 keys,
    // that is all of it).
    constexpr uint32_t kScanBatch = 16;
    uint32_t v0[kScanBatch];
#pragma unroll
    for (uint32_t k = 0; k < kScanBatch; ++k)
        v0[k] = t0 + k < t1 ? tile_hist[static_cast<size_t>(t0 + k) * kRadix + d] : 0u;
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < kScanBatch; ++k) sum += v0[k];
    for (uint32_t t = t0 + kScanBatch; t < t1; t += kScanBatch) {
        uint32_t v[kScanBatch];
#pragma unroll
        for (uint32_t k = 0; k < kScanBatch; ++k)
            v[k] = t + k < t1 ? tile_hist[static_cast<size_t>(t + k) * kRadix + d] : 0u;
#pragma unroll
        for (uint32_t k = 0; k < kScanBatch; ++k) sum += v[k];
    }
    s_part[wave][lane] = sum;
    __syncthreads();
    uint32_t run = 0;
    for (int w = 0; w < wave; ++w) run += s_part[w][lane];
#pragma unroll
    for (uint32_t k = 0; k < kScanBatch; ++k) {
        if (t0 + k < t1) tile_hist[static_cast<size_t>(t0 + k) * kRadix + d] = run;
        run += v0[k];
    }
    for (uint32_t t = t0 + kScanBatch; t < t1; t += kScanBatch) {
        uint32_t v[kScanBatch];
#pragma unroll
        for (uint32_t k = 0; k < kScanBatch; ++k)
            v[k] = t + k < t1 ? tile_hist[static_cast<size_t>(t + k) * kRadix + d] : 0u;
#pragma unroll
        for (uint32_t k = 0; k < kScanBatch; ++k) {
            if (t + k < t1) tile_hist[static_cast<size_t>(t + k) * kRadix + d] = run;
            run += v[k];
        }
    }
    if (wave == kScanWaves - 1) digit_total[d] = run;
}

// Block-wide exclusive scan of one value per thread over WAVES waves;
// `total` receives the block sum.  Contains two barriers.
template <int WAVES>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_w, uint32_t& total) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_w[wave] = x;
    __syncthreads();
    uint32_t pre = 0, sum = 0;
#pragma unroll
    for (int k = 0; k < WAVES; ++k) {
        const uint32_t w = s_w[k];
        pre += (k < wave) ? w : 0u;
        sum += w;
    }
    total = sum;
    __syncthreads();
    return pre + x - v;
}

// Stable scatter of one 4096-key tile by THREADS threads.  Each wave ranks
// 64 keys at a time against the keys of its chunk so far: lanes holding the
// same digit are found with ten ballots, their order is the popcount of
// lower peer lanes, and a per-wave LDS counter per digit carries the count
// between rounds.  Keys are then staged in LDS in digit order and written
// out in runs.  Thread t owns digits DPT*t .. DPT*t+DPT-1 in the per-digit
// scans.
template <bool BY_SCENE, int THREADS>
__global__ __launch_bounds__(THREADS) void k_sort_scatter(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
    const uint32_t* __restrict__ scene_of, uint32_t n, int shift,
    uint32_t num_tiles, const uint32_t* __restrict__ tile_hist,
    const uint32_t* __restrict__ digit_total) {
    constexpr int WAVES = THREADS / 64;
    constexpr int KPT = kSortTile / THREADS;  // keys per thread
    constexpr int DPT = kRadix / THREADS;     // digits per thread
    static_assert(KPT * THREADS == kSortTile && DPT * THREADS == static_cast<int>(kRadix), "tile shape");
    __shared__ uint32_t s_wc[WAVES][kRadix];
    __shared__ uint32_t s_dstart[kRadix];
    __shared__ uint32_t s_gbase[kRadix];
    __shared__ uint32_t s_keys[kSortTile];
    __shared__ uint32_t s_vals[kSortTile];
    __shared__ uint32_t s_w[WAVES];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
#pragma unroll
    for (int k = 0; k < WAVES; ++k)
#pragma unroll
        for (int j = 0; j < DPT; ++j) s_wc[k][tid + j * THREADS] = 0;
    __syncthreads();

    const uint32_t tile = blockIdx.x;
    const uint32_t base = tile * kSortTile;
    const uint32_t count = min(static_cast<uint32_t>(kSortTile), n - base);
    const uint64_t lt = (1ull << lane) - 1ull;

    uint32_t key[KPT], val[KPT], rk[KPT];
    uint32_t dig[KPT];
#pragma unroll
    for (int r = 0; r < KPT; ++r) {
        const uint32_t li = wave * (KPT * 64) + r * 64 + lane;
        const bool valid = li < count;
        key[r] = valid ? keys_in[base + li] : 0u;
        val[r] = valid ? vals_in[base + li] : 0u;
    }
#pragma unroll
    for (int r = 0; r < KPT; ++r) {
        const uint32_t li = wave * (KPT * 64) + r * 64 + lane;
        const bool valid = li < count;
        const uint32_t d = valid ? sort_digit<BY_SCENE>(key[r], val[r], scene_of, shift) : 0u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kDigitBits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t rank = static_cast<uint32_t>(__popcll(peers & lt));
        uint32_t old = 0;
        if (valid) {
            old = s_wc[wave][d];
            if (rank == 0) s_wc[wave][d] = old + static_cast<uint32_t>(__popcll(peers));
        }
        dig[r] = d;
        rk[r] = old + rank;
    }
    __syncthreads();

    // Digits DPT*t ..: offsets of each wave's run, tile-local start, global base.
    uint32_t tot[DPT], gt[DPT];
    uint32_t tsum = 0, gsum = 0;
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
        const int d = DPT * tid + j;
        uint32_t run = 0;
#pragma unroll
        for (int k = 0; k < WAVES; ++k) {  // each digit's words are read and rewritten by its own thread only
            const uint32_t c = s_wc[k][d];
            s_wc[k][d] = run;
            run += c;
        }
        tot[j] = run;
        gt[j] = digit_total[d];
        tsum += tot[j];
        gsum += gt[j];
    }
    uint32_t tmp;
    uint32_t dstart = block_excl_scan<WAVES>(tsum, s_w, tmp);
    uint32_t gpre = block_excl_scan<WAVES>(gsum, s_w, tmp);
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
        const int d = DPT * tid + j;
        s_dstart[d] = dstart;
        s_gbase[d] = gpre + tile_hist[static_cast<size_t>(tile) * kRadix + d];
        dstart += tot[j];
        gpre += gt[j];
    }
    __syncthreads();

#pragma unroll
    for (int r = 0; r < KPT; ++r) {
        const uint32_t li = wave * (KPT * 64) + r * 64 + lane;
        if (li < count) {
            const uint32_t d = dig[r];
            const uint32_t pos = s_dstart[d] + s_wc[wave][d] + rk[r];
            s_keys[pos] = key[r];
            s_vals[pos] = val[r];
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < count; i += THREADS) {
        const uint32_t k = s_keys[i];
        const uint32_t v = s_vals[i];
        const uint32_t d = sort_digit<BY_SCENE>(k, v, scene_of, shift);
        const uint32_t g = s_gbase[d] + (i - s_dstart[d]);
        keys_out[g] = k;
        vals_out[g] = v;
    }
}

// ---- radix tree -----------------------------------------------------------------

// delta (BVHBuilder.cu:116-133) of key i and the key x places from it
// (forward or backward): common prefix length, augmented with the index
// bits when the keys are equal; -1 outside [0, n).  Indices are 32-bit (a
// scene holds < 2^30 keys, so an offset clamped to 2^30 is already out of
// range).
__device__ __forceinline__ int delta_at(const uint32_t* __restrict__ keys, uint32_t n, uint32_t i, uint32_t ki,
                                        bool fwd, uint32_t x) {
    x = min(x, 1u << 30);
    const int32_t j = fwd ? static_cast<int32_t>(i + x) : static_cast<int32_t>(i) - static_cast<int32_t>(x);
    if (j < 0 || static_cast<uint32_t>(j) >= n) return -1;
    const uint32_t kj = keys[j];
    if (ki == kj) return 32 + __clz(static_cast<int>(i ^ static_cast<uint32_t>(j)));
    return __clz(static_cast<int>(ki ^ kj));
}

// radix_tree_construction_kernel (BVHBuilder.cu:135-196), one thread per
// internal node, on its scene's segment of sorted keys with local indices
// (the index augmentation of delta uses them, as in a single-scene build).
// Also initialises the node's refit counter with its refit level (from its
// global leaf range) and appends level-1 nodes to their 1024-leaf block's
// list (span_list[block * kListCap ...], count in span_count[block]) and
// level-2 nodes to their 65536-leaf block's (span_top[...], count in
// span_count[n_lists1 + block]), with each listed node's list and position
// in slots[node].  The node's traversal record
// is written whole by the refit, once its box is known: a 16-byte store into
// each 64-byte record here would cost a partial-line write per node.
__global__ __launch_bounds__(256) void k_radix_tree(const uint32_t* __restrict__ keys, uint32_t n,
                                                    const uint32_t* __restrict__ off,
                                                    const uint32_t* __restrict__ scene_of,
                                                    uint32_t* __restrict__ children,
                                                    uint32_t* __restrict__ parents_leaf,
                                                    uint32_t* __restrict__ parents_internal,
                                                    uint32_t* __restrict__ counters,
                                                    float4* __restrict__ nodes,
                                                    uint32_t* __restrict__ span_list,
                                                    uint32_t* __restrict__ span_top,
                                                    uint32_t* __restrict__ span_count,
                                                    uint32_t n_lists1,
                                                    uint32_t* __restrict__ slots,
                                                    uint32_t* __restrict__ status) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const uint32_t sc = scene_of[g];
    const uint32_t base = off[sc];
    const int64_t N = off[sc + 1] - base;
    const int64_t i = g - base;
    if (i >= N - 1) return;  // unused slot of the scene (or a 1-triangle scene)
    const uint32_t* sk = keys + base;
    const uint32_t n_loc = static_cast<uint32_t>(N), il = static_cast<uint32_t>(i);
    const uint32_t ki = sk[il];
    // Direction, the range's far end j and the split (Karras 2012, Alg. 2;
    // the reference's exponential, binary and ceil-halving searches), on
    // 32-bit offsets from i.
    const int dp = delta_at(sk, n_loc, il, ki, true, 1), dm = delta_at(sk, n_loc, il, ki, false, 1);
    const bool fwd = dp > dm;  // never equal: augmented keys are distinct
    const int dmin = fwd ? dm : dp;
    uint32_t lmax = 2;
    while (delta_at(sk, n_loc, il, ki, fwd, lmax) > dmin) lmax *= 2;  // a probe at 2^30 fails: no overflow
    uint32_t l = 0;
    while (lmax > 1) {
        lmax /= 2;
        if (delta_at(sk, n_loc, il, ki, fwd, l + lmax) > dmin) l += lmax;
    }
    const int dnode = delta_at(sk, n_loc, il, ki, fwd, l);
    uint32_t sx = 0, r = l;
    while (r > 1) {
        r = (r + 1) / 2;
        if (delta_at(sk, n_loc, il, ki, fwd, sx + r) > dnode) sx += r;
    }
    const int64_t d = fwd ? 1 : -1;
    const int64_t j = i + static_cast<int64_t>(l) * d;
    const int64_t s = i + static_cast<int64_t>(sx) * d + (d < 0 ? d : 0);
    uint32_t type = 0;
    if ((i < j ? i : j) != s) {
        type |= 1u;
        parents_internal[base + s] = g;
    } else {
        parents_leaf[base + s] = g;
    }
    if ((i > j ? i : j) != s + 1) {
        type |= 2u;
        parents_internal[base + s + 1] = g;
    } else {
        parents_leaf[base + s + 1] = g;
    }
    children[g] = ((base + static_cast<uint32_t>(s)) << 2) | type;
    const int64_t lo = base + (i < j ? i : j), hi = base + (i > j ? i : j);
    uint32_t level = 0;
    if ((lo / kRefitBlock) != (hi / kRefitBlock)) level = ((lo / kMidBlock) != (hi / kMidBlock)) ? 2u : 1u;
    counters[g] = level << kLevelShift;
    if (level != 0) {
        const uint32_t blk = static_cast<uint32_t>(level == 1 ? lo / kRefitBlock : lo / kMidBlock);
        const uint32_t slot = atomicAdd(span_count + (level == 1 ? 0u : n_lists1) + blk, 1u);
        if (slot < kListCap) {
            (level == 1 ? span_list : span_top)[static_cast<size_t>(blk) * kListCap + slot] = g;
            // The list within its 65536-leaf block (level 1) or overall (level 2), and the place in it.
            slots[g] = ((level == 1 ? blk % (kMidBlock / kRefitBlock) : blk) << 6) | slot;
        } else {
            raise_status(status, PT_STATUS_REFIT_OVERFLOW);  // bound argued above
        }
    }
    if (i == 0) parents_internal[g] = kInvalid;  // BVHBuilder.cu:224
}

// ---- refit ------------------------------------------------------------------------

struct Box {
    float mnx, mxx, mny, mxy, mnz, mxz;
};

__device__ __forceinline__ void store_slot(float4* nodes, uint32_t p, bool left, const Box& b) {
    float4* rec = nodes + 4 * static_cast<size_t>(p);
    if (left) {
        rec[0] = make_float4(b.mnx, b.mxx, b.mny, b.mxy);
        reinterpret_cast<float2*>(rec + 1)[0] = make_float2(b.mnz, b.mxz);
    } else {
        reinterpret_cast<float2*>(rec + 1)[1] = make_float2(b.mnx, b.mxx);
        rec[2] = make_float4(b.mny, b.mxy, b.mnz, b.mxz);
    }
}

__device__ __forceinline__ Box unite(const Box& a, const Box& b) {
    // leftBBox.unite(rightBBox) (BVHBuilder.cu:85, BBox.cpp:30-36).
    Box o;
    o.mnx = gmin(a.mnx, b.mnx);
    o.mxx = gmax(a.mxx, b.mxx);
    o.mny = gmin(a.mny, b.mny);
    o.mxy = gmax(a.mxy, b.mxy);
    o.mnz = gmin(a.mnz, b.mnz);
    o.mxz = gmax(a.mxz, b.mxz);
    return o;
}

// Node p's whole traversal record: both child boxes and the child word, in
// one 64-byte line.
__device__ __forceinline__ void store_record(float4* nodes, uint32_t p, const Box& lb, const Box& rb, uint32_t w) {
    float4* rec = nodes + 4 * static_cast<size_t>(p);
    rec[0] = make_float4(lb.mnx, lb.mxx, lb.mny, lb.mxy);
    rec[1] = make_float4(lb.mnz, lb.mxz, rb.mnx, rb.mxx);
    rec[2] = make_float4(rb.mny, rb.mxy, rb.mnz, rb.mxz);
    rec[3] = make_float4(__uint_as_float(w), 0.0f, 0.0f, 0.0f);
}

__device__ __forceinline__ void store_root(float* root_box, const Box& b) {
    root_box[0] = b.mnx; root_box[1] = b.mny; root_box[2] = b.mnz;
    root_box[3] = b.mxx; root_box[4] = b.mxy; root_box[5] = b.mxz;
}

// LDS box array in structure-of-arrays form (one float plane per bound).
struct LdsBoxes {
    float v[6][kRefitBlock];
    __device__ __forceinline__ void put(int k, const Box& b) {
        v[0][k] = b.mnx; v[1][k] = b.mxx; v[2][k] = b.mny;
        v[3][k] = b.mxy; v[4][k] = b.mnz; v[5][k] = b.mxz;
    }
    __device__ __forceinline__ Box get(int k) const {
        return Box{v[0][k], v[1][k], v[2][k], v[3][k], v[4][k], v[5][k]};
    }
};

// Refit of the level-L nodes listed in `list` (count at *count), all inside
// one workgroup so the counter hand-off needs only workgroup-scope ordering.
// Every child box of a node sits in the node's own record: lower-level
// children were stored by earlier phases, same-level children are stored here
// before their parent's counter is incremented.  Each lower-level child of a
// listed node is one "arrival"; the thread completing a node climbs on while
// the parent is of the same level.
__device__ __forceinline__ void refit_level(const uint32_t* __restrict__ children,
                                            const uint32_t* __restrict__ scene_of,
                                            const uint32_t* __restrict__ parents_internal,
                                            uint32_t* __restrict__ counters,
                                            float4* __restrict__ nodes,
                                            float* __restrict__ root_box,
                                            const uint32_t* list, uint32_t count, uint32_t level) {
    for (uint32_t idx = threadIdx.x; idx < count; idx += blockDim.x) {
        const uint32_t s = list[idx];
        const uint32_t w = children[s];
        const uint32_t l = w >> 2, r = l + 1;
        int arrivals = 0;
        if (!(w & 1u) || (counters[l] >> kLevelShift) < level) ++arrivals;
        if (!(w & 2u) || (counters[r] >> kLevelShift) < level) ++arrivals;
        for (int a = 0; a < arrivals; ++a) {
            uint32_t p = s;
            for (;;) {
                const uint32_t old = __hip_atomic_fetch_add(counters + p, 1u, __ATOMIC_ACQ_REL,
                                                            __HIP_MEMORY_SCOPE_WORKGROUP);
                if ((old & kCountMask) == 0u) break;
                const float4* rec = nodes + 4 * static_cast<size_t>(p);
                const float4 r0 = rec[0], r1 = rec[1], r2 = rec[2];
                const Box lb{r0.x, r0.y, r0.z, r0.w, r1.x, r1.y};
                const Box rb{r1.z, r1.w, r2.x, r2.y, r2.z, r2.w};
                const Box b = unite(lb, rb);
                nodes[4 * static_cast<size_t>(p) + 3] = make_float4(__uint_as_float(children[p]), 0.0f, 0.0f, 0.0f);
                const uint32_t pp = parents_internal[p];
                if (pp == kInvalid) {
                    store_root(root_box + 6 * static_cast<size_t>(scene_of[p]), b);
                    break;
                }
                store_slot(nodes, pp, (children[pp] >> 2) == p, b);
                if ((counters[pp] >> kLevelShift) != level) break;
                p = pp;
            }
        }
    }
}

// The same refit with the list staged in LDS (lists of up to kStageCap
// nodes): every listed node's record (the child boxes the earlier phases
// stored), its parent and its count of same-level children are gathered
// first, so the climb itself touches only LDS.  Each completed node's
// record is written whole; a box whose parent belongs to a later phase also
// goes into that parent's record.
constexpr uint32_t kStageCap = 1024;
constexpr uint32_t kUpOutside = 0xffffffffu;  // parent handled by a later phase, or none

struct LevelStage {
    float box[12][kStageCap];  // left child box, then right, in record order
    uint32_t node[kStageCap];  // the listed node
    uint32_t word[kStageCap];  // its child word
    uint32_t par[kStageCap];   // its parent (kInvalid at a scene root)
    uint32_t up[kStageCap];    // parent's slot << 1 | right side, or kUpOutside
    uint32_t wait[kStageCap];  // same-level children not yet complete
};

// A listed node's links: its child word, parent, the parent's slot in the
// same list (or kUpOutside) and its number of same-level children.
struct Links {
    uint32_t word, par, up, wait;
};

__device__ __forceinline__ Links node_links(uint32_t s, const uint32_t* __restrict__ children,
                                            const uint32_t* __restrict__ parents_internal,
                                            const uint32_t* __restrict__ counters,
                                            const uint32_t* __restrict__ slots, const uint32_t* list_start,
                                            uint32_t level) {
    Links L;
    L.word = children[s];
    const uint32_t l = L.word >> 2, r = l + 1;
    L.wait = static_cast<uint32_t>((L.word & 1u) && (counters[l] >> kLevelShift) == level) +
             static_cast<uint32_t>((L.word & 2u) && (counters[r] >> kLevelShift) == level);
    L.par = parents_internal[s];
    L.up = kUpOutside;
    if (L.par != kInvalid && (counters[L.par] >> kLevelShift) == level) {
        const uint32_t sl = slots[L.par];
        L.up = ((list_start[sl >> 6] + (sl & 63u)) << 1) | static_cast<uint32_t>((children[L.par] >> 2) != s);
    }
    return L;
}

__device__ __forceinline__ void stage_links(LevelStage& S, uint32_t idx, uint32_t s, const Links& L) {
    S.node[idx] = s;
    S.word[idx] = L.word;
    S.par[idx] = L.par;
    S.up[idx] = L.up;
    S.wait[idx] = L.wait;
}

// A listed node's two child boxes, from its record (stored by earlier phases).
__device__ __forceinline__ void stage_boxes(LevelStage& S, uint32_t idx, uint32_t s,
                                            const float4* __restrict__ nodes) {
    const float4* rec = nodes + 4 * static_cast<size_t>(s);
    const float4 r0 = rec[0], r1 = rec[1], r2 = rec[2];
    const float v[12] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w, r2.x, r2.y, r2.z, r2.w};
    for (int k = 0; k < 12; ++k) S.box[k][idx] = v[k];
}

// The climb over a staged list (after a barrier): every node with no
// same-level child starts one; the thread completing a node's second
// same-level child goes on to the node.
__device__ __forceinline__ void climb_staged(LevelStage& S, const uint32_t* __restrict__ children,
                                             const uint32_t* __restrict__ scene_of, float4* __restrict__ nodes,
                                             float* __restrict__ root_box, uint32_t count) {
    for (uint32_t first = threadIdx.x; first < count; first += blockDim.x) {
        if (S.wait[first] != 0u) continue;
        uint32_t idx = first;
        for (;;) {
            const Box lb{S.box[0][idx], S.box[1][idx], S.box[2][idx], S.box[3][idx], S.box[4][idx], S.box[5][idx]};
            const Box rb{S.box[6][idx], S.box[7][idx], S.box[8][idx], S.box[9][idx], S.box[10][idx], S.box[11][idx]};
            const Box b = unite(lb, rb);
            const uint32_t s = S.node[idx];
            store_record(nodes, s, lb, rb, S.word[idx]);
            const uint32_t pp = S.par[idx];
            if (pp == kInvalid) {
                store_root(root_box + 6 * static_cast<size_t>(scene_of[s]), b);
                break;
            }
            const uint32_t up = S.up[idx];
            if (up == kUpOutside) {
                store_slot(nodes, pp, (children[pp] >> 2) == s, b);
                break;
            }
            const uint32_t ps = up >> 1;
            const int o = (up & 1u) ? 6 : 0;
            S.box[o + 0][ps] = b.mnx; S.box[o + 1][ps] = b.mxx; S.box[o + 2][ps] = b.mny;
            S.box[o + 3][ps] = b.mxy; S.box[o + 4][ps] = b.mnz; S.box[o + 5][ps] = b.mxz;
            const uint32_t old = __hip_atomic_fetch_sub(&S.wait[ps], 1u, __ATOMIC_ACQ_REL,
                                                        __HIP_MEMORY_SCOPE_WORKGROUP);
            if (old != 1u) break;  // the other same-level child completes the parent
            idx = ps;
        }
    }
}

__device__ __forceinline__ void refit_level_staged(LevelStage& S, const uint32_t* __restrict__ children,
                                                   const uint32_t* __restrict__ scene_of,
                                                   const uint32_t* __restrict__ parents_internal,
                                                   const uint32_t* __restrict__ counters,
                                                   const uint32_t* __restrict__ slots, const uint32_t* list_start,
                                                   float4* __restrict__ nodes, float* __restrict__ root_box,
                                                   const uint32_t* list, uint32_t count, uint32_t level) {
    for (uint32_t idx = threadIdx.x; idx < count; idx += blockDim.x) {
        const uint32_t s = list[idx];
        stage_boxes(S, idx, s, nodes);
        stage_links(S, idx, s, node_links(s, children, parents_internal, counters, slots, list_start, level));
    }
    __syncthreads();
    climb_staged(S, children, scene_of, nodes, root_box, count);
}

// The lists of one level, concatenated in LDS-held order: list k starts at
// start[k] (exclusive prefix of the counts, one thread per list).
__device__ __forceinline__ uint32_t concat_lists(const uint32_t* __restrict__ counts, uint32_t nlists,
                                                 const uint32_t* __restrict__ lists, uint32_t* start,
                                                 uint32_t* out, uint32_t* status) {
    __shared__ uint32_t s_carry;
    __shared__ uint32_t s_w[16];
    if (threadIdx.x == 0) s_carry = 0u;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t b0 = 0; b0 < nlists; b0 += blockDim.x) {
        const uint32_t k = b0 + threadIdx.x;
        uint32_t c = 0;
        if (k < nlists) {
            c = counts[k];
            if (c > kListCap) {  // counted past a full list (flagged by the append)
                raise_status(status, PT_STATUS_REFIT_OVERFLOW);
                c = kListCap;
            }
        }
        uint32_t x = c;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) s_w[wave] = x;
        __syncthreads();
        uint32_t pre = s_carry;
        for (int w = 0; w < wave; ++w) pre += s_w[w];
        if (k < nlists) start[k] = pre + x - c;
        __syncthreads();
        if (threadIdx.x == blockDim.x - 1) s_carry = pre + x;
        __syncthreads();
    }
    const uint32_t total = s_carry;
    if (threadIdx.x == 0) start[nlists] = total;
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < nlists * kListCap; e += blockDim.x) {
        const uint32_t k = e / kListCap, j = e % kListCap;
        if (j < start[k + 1] - start[k]) out[start[k] + j] = lists[static_cast<size_t>(k) * kListCap + j];
    }
    __syncthreads();
    return total;
}

// Level-1 links prepared by k_refit_local's last workgroups, per 65536-leaf
// block: [0] the block's list length, then per listed node the node, its
// child word, parent, parent's slot and same-level child count (kStageCap
// each).
constexpr uint32_t kMidLinksWords = 1 + 5 * kStageCap;
static_assert(sizeof(LdsBoxes) >= 4 * kMidCap, "level-1 list staging fits in a box array");

// Refit, phase 1 (compute_internal_node_bbox_kernel, BVHBuilder.cu:54-91):
// one thread per sorted leaf, one workgroup per block of kRefitBlock leaves.
// The leaf writes its sorted triangle record (the reference instead permutes
// the 40-byte triangles inside the sort), then climbs: the first thread to
// reach a node stops, the second unites the two child boxes.  Every node
// below the spanning ones has its whole leaf range -- hence both children --
// inside the block, so counters and boxes live in LDS and the hand-off needs
// only workgroup-scope ordering.  The thread completing a node writes the
// node's whole traversal record (both child boxes, leaves' included, and the
// child word); a box whose parent is finished by a later phase is also
// stored into that parent's record.  A node inside the block has its index
// inside its leaf range, so the block's nodes are the indices [base, base +
// kRefitBlock): their child words, parents and levels are loaded into LDS
// once, coalesced, and the climb reads no global memory.  The launch's last
// mid_blocks workgroups instead concatenate each 65536-leaf block's level-1
// lists and stage their links (children, parents, counters and slots, none
// of which this launch writes) into mid_links for k_refit_mid.
__global__ __launch_bounds__(kRefitBlock) void k_refit_local(
    const pt_triangle* __restrict__ in_tris, const uint32_t* __restrict__ perm,
    const uint32_t* __restrict__ off, const uint32_t* __restrict__ scene_of,
    const uint32_t* __restrict__ children, uint32_t* __restrict__ parents_leaf,
    const uint32_t* __restrict__ parents_internal, const uint32_t* __restrict__ counters,
    float4* __restrict__ nodes, float4* __restrict__ tri_recs, float* __restrict__ root_box,
    uint32_t n, const uint32_t* __restrict__ span_list, const uint32_t* __restrict__ span_count,
    const uint32_t* __restrict__ slots, uint32_t stage_max, uint32_t* __restrict__ mid_links,
    uint32_t* __restrict__ status) {
    __shared__ uint32_t s_cnt[kRefitBlock];
    __shared__ uint32_t s_word[kRefitBlock];   // child word of node base + j
    __shared__ uint32_t s_par[kRefitBlock];    // its parent
    __shared__ uint32_t s_local[kRefitBlock];  // 1: refit here (level 0)
    __shared__ LdsBoxes s_leaf;
    __shared__ LdsBoxes s_node;
    const int tid = threadIdx.x;
    const uint32_t n_lists1 = (n + kRefitBlock - 1) / kRefitBlock;
    if (blockIdx.x >= n_lists1) {  // level-1 link staging for 65536-leaf block m
        constexpr uint32_t kLists = kMidBlock / kRefitBlock;
        const uint32_t m = blockIdx.x - n_lists1;
        uint32_t* s_list = reinterpret_cast<uint32_t*>(&s_leaf);   // kMidCap words
        uint32_t* s_start = reinterpret_cast<uint32_t*>(&s_node);  // kLists + 1 words
        const uint32_t k0 = m * kLists;
        const uint32_t count = concat_lists(span_count + k0, min(kLists, n_lists1 - k0),
                                            span_list + static_cast<size_t>(k0) * kListCap, s_start, s_list, status);
        uint32_t* out = mid_links + static_cast<size_t>(m) * kMidLinksWords;
        if (tid == 0) out[0] = count;
        if (count > stage_max) return;
        for (uint32_t idx = tid; idx < count; idx += kRefitBlock) {
            const uint32_t sn = s_list[idx];
            const Links L = node_links(sn, children, parents_internal, counters, slots, s_start, 1u);
            out[1 + idx] = sn;
            out[1 + kStageCap + idx] = L.word;
            out[1 + 2 * kStageCap + idx] = L.par;
            out[1 + 3 * kStageCap + idx] = L.up;
            out[1 + 4 * kStageCap + idx] = L.wait;
        }
        return;
    }
    const uint32_t base = blockIdx.x * kRefitBlock;
    const uint32_t k = base + tid;
    s_cnt[tid] = 0u;
    if (k < n) {
        s_word[tid] = children[k];
        s_par[tid] = parents_internal[k];
        s_local[tid] = (counters[k] >> kLevelShift) == 0u ? 1u : 0u;
    } else {
        s_local[tid] = 0u;
    }
    __syncthreads();
    if (k >= n) return;
    // Node q is refit by this block: inside [base, base + kRefitBlock) and of level 0.
    auto local = [&](uint32_t q) { return q - base < static_cast<uint32_t>(kRefitBlock) && s_local[q - base]; };
    const pt_triangle* t = in_tris + perm[k];
    const V3 v0 = v3(t->v0[0], t->v0[1], t->v0[2]);
    const V3 v1 = v3(t->v1[0], t->v1[1], t->v1[2]);
    const V3 v2 = v3(t->v2[0], t->v2[1], t->v2[2]);
    const uint32_t id = t->id;
    const V3 e1 = sub(v1, v0), e2 = sub(v2, v0);
    tri_recs[3 * static_cast<size_t>(k) + 0] = make_float4(v0.x, v0.y, v0.z, __uint_as_float(id));
    tri_recs[3 * static_cast<size_t>(k) + 1] = make_float4(e1.x, e1.y, e1.z, 0.0f);
    tri_recs[3 * static_cast<size_t>(k) + 2] = make_float4(e2.x, e2.y, e2.z, 0.0f);
    // BBox() then unite({v0, v1, v2}) (BBox.cpp:4-6, 14-24).
    Box b;
    b.mnx = gmin(gmin(gmin(kFltMax, v0.x), v1.x), v2.x);
    b.mny = gmin(gmin(gmin(kFltMax, v0.y), v1.y), v2.y);
    b.mnz = gmin(gmin(gmin(kFltMax, v0.z), v1.z), v2.z);
    b.mxx = gmax(gmax(gmax(-kFltMax, v0.x), v1.x), v2.x);
    b.mxy = gmax(gmax(gmax(-kFltMax, v0.y), v1.y), v2.y);
    b.mxz = gmax(gmax(gmax(-kFltMax, v0.z), v1.z), v2.z);
    const uint32_t sc = scene_of[k];
    if (off[sc + 1] - off[sc] == 1) {  // one-triangle scene: no internal node
        parents_leaf[k] = kInvalid;
        store_root(root_box + 6 * static_cast<size_t>(sc), b);
        return;
    }
    s_leaf.put(tid, b);
    uint32_t p = parents_leaf[k];
    if (!local(p)) {
        store_slot(nodes, p, (children[p] >> 2) == k, b);  // finished by a later phase
        return;
    }
    for (;;) {
        const uint32_t old = __hip_atomic_fetch_add(s_cnt + (p - base), 1u, __ATOMIC_ACQ_REL,
                                                    __HIP_MEMORY_SCOPE_WORKGROUP);
        if (old == 0u) return;
        const uint32_t w = s_word[p - base];
        const uint32_t l = w >> 2, r = l + 1;
        const Box lb = (w & 1u) ? s_node.get(l - base) : s_leaf.get(l - base);
        const Box rb = (w & 2u) ? s_node.get(r - base) : s_leaf.get(r - base);
        b = unite(lb, rb);
        s_node.put(p - base, b);
        store_record(nodes, p, lb, rb, w);
        const uint32_t pp = s_par[p - base];
        if (pp == kInvalid) {
            store_root(root_box + 6 * static_cast<size_t>(scene_of[p]), b);
            return;
        }
        if (!local(pp)) {  // finished by a later phase, from its record
            store_slot(nodes, pp, (children[pp] >> 2) == p, b);
            return;
        }
        p = pp;
    }
}

// The level-2 stage prepared by k_refit_mid's last workgroup for k_refit_top:
// [0] the list's length, then per listed node its child word, parent,
// parent's slot and same-level child count (kStageCap each).
constexpr uint32_t kTopLinksWords = 1 + 4 * kStageCap;

// Level-1 nodes: one workgroup per 65536-leaf block (its 64 lists).  With
// level-2 nodes present, one more workgroup concatenates their lists into
// top_list and stages their links (children, parents, counters and slots,
// none of which this launch writes) into top_links, so that k_refit_top
// only has to gather the child boxes.
__global__ __launch_bounds__(1024) void k_refit_mid(const uint32_t* __restrict__ children,
                                                    const uint32_t* __restrict__ scene_of,
                                                    const uint32_t* __restrict__ parents_internal,
                                                    uint32_t* __restrict__ counters,
                                                    float4* __restrict__ nodes,
                                                    float* __restrict__ root_box,
                                                    const uint32_t* __restrict__ span_list,
                                                    const uint32_t* __restrict__ span_count, uint32_t n_lists1,
                                                    const uint32_t* __restrict__ span_top, uint32_t n_lists2,
                                                    uint32_t* __restrict__ top_list, uint32_t* __restrict__ top_links,
                                                    const uint32_t* __restrict__ mid_links,
                                                    const uint32_t* __restrict__ slots, uint32_t stage_max,
                                                    uint32_t* __restrict__ status) {
    constexpr uint32_t kLists = kMidBlock / kRefitBlock;
    __shared__ LevelStage S;
    __shared__ uint32_t s_buf[kMaxMidBlocks + 1 > kMidCap + kLists + 1 ? kMaxMidBlocks + 1 : kMidCap + kLists + 1];
    const uint32_t n_mid = (n_lists1 + kLists - 1) / kLists;
    if (blockIdx.x == n_mid) {  // the level-2 preparation
        uint32_t* s_start = s_buf;
        const uint32_t count = concat_lists(span_count + n_lists1, n_lists2, span_top, s_start, top_list, status);
        if (threadIdx.x == 0) top_links[0] = count;
        if (count > stage_max) return;
        for (uint32_t idx = threadIdx.x; idx < count; idx += blockDim.x) {
            const Links L = node_links(top_list[idx], children, parents_internal, counters, slots, s_start, 2u);
            top_links[1 + idx] = L.word;
            top_links[1 + kStageCap + idx] = L.par;
            top_links[1 + 2 * kStageCap + idx] = L.up;
            top_links[1 + 3 * kStageCap + idx] = L.wait;
        }
        return;
    }
    const uint32_t* links = mid_links + static_cast<size_t>(blockIdx.x) * kMidLinksWords;
    const uint32_t staged = links[0];
    if (staged <= stage_max) {
        for (uint32_t idx = threadIdx.x; idx < staged; idx += blockDim.x) {
            const uint32_t sn = links[1 + idx];
            S.node[idx] = sn;
            S.word[idx] = links[1 + kStageCap + idx];
            S.par[idx] = links[1 + 2 * kStageCap + idx];
            S.up[idx] = links[1 + 3 * kStageCap + idx];
            S.wait[idx] = links[1 + 4 * kStageCap + idx];
            stage_boxes(S, idx, sn, nodes);
        }
        __syncthreads();
        climb_staged(S, children, scene_of, nodes, root_box, staged);
        return;
    }
    uint32_t* s_list = s_buf;
    uint32_t* s_start = s_buf + kMidCap;
    const uint32_t k0 = blockIdx.x * kLists;
    const uint32_t count = concat_lists(span_count + k0, min(kLists, n_lists1 - k0),
                                        span_list + static_cast<size_t>(k0) * kListCap, s_start, s_list, status);
    refit_level(children, scene_of, parents_internal, counters, nodes, root_box, s_list, count, 1u);
}

// Level-2 nodes: one workgroup, over the list and links k_refit_mid prepared.
__global__ __launch_bounds__(1024) void k_refit_top(const uint32_t* __restrict__ children,
                                                    const uint32_t* __restrict__ scene_of,
                                                    const uint32_t* __restrict__ parents_internal,
                                                    uint32_t* __restrict__ counters,
                                                    float4* __restrict__ nodes,
                                                    float* __restrict__ root_box,
                                                    const uint32_t* __restrict__ top_list,
                                                    const uint32_t* __restrict__ top_links, uint32_t stage_max) {
    __shared__ LevelStage S;
    const uint32_t count = top_links[0];
    if (count <= stage_max) {
        for (uint32_t idx = threadIdx.x; idx < count; idx += blockDim.x) {
            const uint32_t s = top_list[idx];
            S.node[idx] = s;
            S.word[idx] = top_links[1 + idx];
            S.par[idx] = top_links[1 + kStageCap + idx];
            S.up[idx] = top_links[1 + 2 * kStageCap + idx];
            S.wait[idx] = top_links[1 + 3 * kStageCap + idx];
            stage_boxes(S, idx, s, nodes);
        }
        __syncthreads();
        climb_staged(S, children, scene_of, nodes, root_box, count);
    } else {
        refit_level(children, scene_of, parents_internal, counters, nodes, root_box, top_list, count, 2u);
    }
}

// Parity dump of scene [base, base + n): internal node boxes in the
// reference's BBox layout (min.xyz, max.xyz), read back from the parents'
// records, plus the child and parent words converted to local indices.
__global__ __launch_bounds__(256) void k_dump_scene(const float4* __restrict__ nodes,
                                                    const uint32_t* __restrict__ children,
                                                    const uint32_t* __restrict__ parents_leaf,
                                                    const uint32_t* __restrict__ parents_internal,
                                                    const float* __restrict__ root_box, uint32_t base,
                                                    uint32_t n, float* __restrict__ out_boxes,
                                                    uint32_t* __restrict__ out_children,
                                                    uint32_t* __restrict__ out_pl,
                                                    uint32_t* __restrict__ out_pi) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) {
        const uint32_t p = parents_leaf[base + c];
        out_pl[c] = p == kInvalid ? kInvalid : p - base;
    }
    if (c + 1 >= n) return;
    const uint32_t g = base + c;
    out_children[c] = children[g] - (base << 2);
    const uint32_t pg = parents_internal[g];
    out_pi[c] = pg == kInvalid ? kInvalid : pg - base;
    float* o = out_boxes + 6 * static_cast<size_t>(c);
    if (pg == kInvalid) {
        for (int k = 0; k < 6; ++k) o[k] = root_box[k];
        return;
    }
    const float4* rec = nodes + 4 * static_cast<size_t>(pg);
    const float4 r0 = rec[0], r1 = rec[1], r2 = rec[2];
    if ((children[pg] >> 2) == g) {
        o[0] = r0.x; o[1] = r0.z; o[2] = r1.x; o[3] = r0.y; o[4] = r0.w; o[5] = r1.y;
    } else {
        o[0] = r1.z; o[1] = r2.x; o[2] = r2.z; o[3] = r1.w; o[4] = r2.y; o[5] = r2.w;
    }
}

__global__ __launch_bounds__(256) void k_sub_u32(const uint32_t* __restrict__ in, uint32_t sub,
                                                 uint32_t n, uint32_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i] - sub;
}

inline uint32_t div_up(uint64_t a, uint64_t b) { return static_cast<uint32_t>((a + b - 1) / b); }

}  // namespace

BuildTuning& build_tuning() {
    static BuildTuning t;
    return t;
}

void build_lbvh(DeviceScene& s, hipStream_t st) {
    DeviceBvh& b = s.bvh;
    const uint32_t n = s.n;
    const uint32_t nscenes = static_cast<uint32_t>(s.off_host.size()) - 1;
    b.n = n;
    if (n == 0) return;
    if (n >= (1u << 30)) throw Error(PT_E_INVALID_ARG, "LBVH supports fewer than 2^30 triangles");
    const size_t n4 = static_cast<size_t>(n) * 4;
    b.keys.reserve(n4);
    b.vals.reserve(n4);
    b.keys_alt.reserve(n4);
    b.vals_alt.reserve(n4);
    b.scene_of.reserve(n4);
    b.children.reserve(n4);
    b.parents_leaf.reserve(n4);
    b.parents_internal.reserve(n4);
    b.counters.reserve(n4);
    b.nodes.reserve(static_cast<size_t>(n) * 64);
    b.tris.reserve(static_cast<size_t>(n) * 48);
    b.root_box.reserve(static_cast<size_t>(nscenes) * 6 * sizeof(float));
    const uint32_t num_tiles = div_up(n, kSortTile);
    b.tile_hist.reserve(static_cast<size_t>(num_tiles) * kRadix * 4);
    b.digit_total.reserve(kRadix * 4);
    const uint32_t mid_blocks = div_up(n, kMidBlock);
    const uint32_t n_lists1 = div_up(n, kRefitBlock);
    b.span_list.reserve(static_cast<size_t>(n_lists1) * kListCap * 4);
    b.span_top.reserve(static_cast<size_t>(mid_blocks) * kListCap * 4 * 2);  // lists, then their concatenation
    b.span_count.reserve((static_cast<size_t>(n_lists1) + mid_blocks) * 4);
    b.top_links.reserve(kTopLinksWords * 4);
    b.mid_links.reserve(static_cast<size_t>(mid_blocks) * kMidLinksWords * 4);
    b.slots.reserve(n4);

    uint32_t* kin = b.keys.as<uint32_t>();
    uint32_t* vin = b.vals.as<uint32_t>();
    hipLaunchKernelGGL(k_morton, dim3(num_tiles), dim3(kMortonThreads), 0, st, s.tris.as<pt_triangle>(), n,
                       s.off.as<uint32_t>(), nscenes, s.morton_box.as<float>(), kin, vin,
                       b.scene_of.as<uint32_t>(), b.tile_hist.as<uint32_t>(), b.span_count.as<uint32_t>(),
                       n_lists1 + mid_blocks);
    PT_HIP(hipGetLastError());

    if (n >= 2) {
        uint32_t* kout = b.keys_alt.as<uint32_t>();
        uint32_t* vout = b.vals_alt.as<uint32_t>();
        // 3 passes over the 30-bit keys, then the scene id's 10-bit digits
        // (stable: each scene keeps its key order, ties in input order).
        constexpr int kKeyPasses = (30 + kDigitBits - 1) / kDigitBits;
        int scene_passes = 0;
        for (uint32_t v = nscenes - 1; v != 0; v >>= kDigitBits) ++scene_passes;
        const int passes = kKeyPasses + scene_passes;
        for (int pass = 0; pass < passes; ++pass) {
            const bool by_scene = pass >= kKeyPasses;
            const int shift = kDigitBits * (by_scene ? pass - kKeyPasses : pass);
            // The first pass's counts come from k_morton.
            if (by_scene) {
                hipLaunchKernelGGL(k_sort_hist<true>, dim3(num_tiles), dim3(kHistThreads), 0, st, kin, vin,
                                   b.scene_of.as<uint32_t>(), n, shift, num_tiles, b.tile_hist.as<uint32_t>());
            } else if (pass > 0) {
                hipLaunchKernelGGL(k_sort_hist<false>, dim3(num_tiles), dim3(kHistThreads), 0, st, kin, vin,
                                   b.scene_of.as<uint32_t>(), n, shift, num_tiles, b.tile_hist.as<uint32_t>());
            }
            hipLaunchKernelGGL(k_sort_rowscan, dim3(kRadix / kScanDigits), dim3(kScanWaves * 64), 0, st,
                               b.tile_hist.as<uint32_t>(), num_tiles, b.digit_total.as<uint32_t>());
            if (by_scene) {
                hipLaunchKernelGGL((k_sort_scatter<true, kScatterThreads>), dim3(num_tiles), dim3(kScatterThreads), 0, st, kin, vin,
                                   kout, vout, b.scene_of.as<uint32_t>(), n, shift, num_tiles,
                                   b.tile_hist.as<uint32_t>(), b.digit_total.as<uint32_t>());
            } else {
                hipLaunchKernelGGL((k_sort_scatter<false, kScatterThreads>), dim3(num_tiles), dim3(kScatterThreads), 0, st, kin, vin,
                                   kout, vout, b.scene_of.as<uint32_t>(), n, shift, num_tiles,
                                   b.tile_hist.as<uint32_t>(), b.digit_total.as<uint32_t>());
            }
            std::swap(kin, kout);
            std::swap(vin, vout);
        }
        if (passes % 2) {
            // Odd pass count: the result sits in the alternate buffers.
            swap(b.keys, b.keys_alt);
            swap(b.vals, b.vals_alt);
        }
        PT_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_radix_tree, dim3(div_up(n, 256)), dim3(256), 0, st, b.keys.as<uint32_t>(), n,
                           s.off.as<uint32_t>(), b.scene_of.as<uint32_t>(), b.children.as<uint32_t>(),
                           b.parents_leaf.as<uint32_t>(), b.parents_internal.as<uint32_t>(),
                           b.counters.as<uint32_t>(), b.nodes.as<float4>(), b.span_list.as<uint32_t>(),
                           b.span_top.as<uint32_t>(), b.span_count.as<uint32_t>(), n_lists1,
                           b.slots.as<uint32_t>(), s.status.as<uint32_t>());
        PT_HIP(hipGetLastError());
    }
    const uint32_t stage_max = build_tuning().staged ? kStageCap : 0u;
    // With level-1 nodes present, mid_blocks more workgroups stage their links.
    const uint32_t mid_prep = n > static_cast<uint32_t>(kRefitBlock) ? mid_blocks : 0u;
    hipLaunchKernelGGL(k_refit_local, dim3(n_lists1 + mid_prep), dim3(kRefitBlock), 0, st,
                       s.tris.as<pt_triangle>(), b.vals.as<uint32_t>(), s.off.as<uint32_t>(),
                       b.scene_of.as<uint32_t>(), b.children.as<uint32_t>(),
                       b.parents_leaf.as<uint32_t>(), b.parents_internal.as<uint32_t>(),
                       b.counters.as<uint32_t>(), b.nodes.as<float4>(), b.tris.as<float4>(),
                       b.root_box.as<float>(), n, b.span_list.as<uint32_t>(), b.span_count.as<uint32_t>(),
                       b.slots.as<uint32_t>(), stage_max, b.mid_links.as<uint32_t>(), s.status.as<uint32_t>());
    if (n > static_cast<uint32_t>(kRefitBlock)) {
        // One more workgroup prepares the level-2 refit when there is one.
        const uint32_t top = n > kMidBlock ? 1u : 0u;
        uint32_t* top_list = b.span_top.as<uint32_t>() + static_cast<size_t>(mid_blocks) * kListCap;
        hipLaunchKernelGGL(k_refit_mid, dim3(mid_blocks + top), dim3(1024), 0, st, b.children.as<uint32_t>(),
                           b.scene_of.as<uint32_t>(), b.parents_internal.as<uint32_t>(),
                           b.counters.as<uint32_t>(), b.nodes.as<float4>(), b.root_box.as<float>(),
                           b.span_list.as<uint32_t>(), b.span_count.as<uint32_t>(), n_lists1,
                           b.span_top.as<uint32_t>(), mid_blocks, top_list, b.top_links.as<uint32_t>(),
                           b.mid_links.as<uint32_t>(), b.slots.as<uint32_t>(), stage_max, s.status.as<uint32_t>());
        if (top) {
            hipLaunchKernelGGL(k_refit_top, dim3(1), dim3(1024), 0, st, b.children.as<uint32_t>(),
                               b.scene_of.as<uint32_t>(), b.parents_internal.as<uint32_t>(),
                               b.counters.as<uint32_t>(), b.nodes.as<float4>(), b.root_box.as<float>(), top_list,
                               b.top_links.as<uint32_t>(), stage_max);
        }
    }
    PT_HIP(hipGetLastError());
}

void dump_bvh(DeviceScene& s, uint32_t scene, hipStream_t st, uint32_t* morton, uint32_t* perm,
              uint32_t* children, uint32_t* parents_leaf, uint32_t* parents_internal,
              float* bboxes) {
    const DeviceBvh& b = s.bvh;
    if (scene + 1 >= s.off_host.size()) throw Error(PT_E_INVALID_ARG, "scene index out of range");
    const uint32_t base = s.off_host[scene];
    const uint32_t n = s.off_host[scene + 1] - base;
    if (n == 0) return;
    const size_t m = n - 1;
    DevBuf tmp;
    tmp.reserve(static_cast<size_t>(n) * 4 * 4 + m * 24 + 256);
    uint32_t* t_perm = tmp.as<uint32_t>();
    uint32_t* t_ch = t_perm + n;
    uint32_t* t_pl = t_ch + n;
    uint32_t* t_pi = t_pl + n;
    float* t_box = reinterpret_cast<float*>(t_pi + n);
    if (morton)
        PT_HIP(hipMemcpyAsync(morton, b.keys.as<uint32_t>() + base, n * 4ull, hipMemcpyDeviceToHost, st));
    hipLaunchKernelGGL(k_sub_u32, dim3(div_up(n, 256)), dim3(256), 0, st, b.vals.as<uint32_t>() + base,
                       base, n, t_perm);
    hipLaunchKernelGGL(k_dump_scene, dim3(div_up(n, 256)), dim3(256), 0, st, b.nodes.as<float4>(),
                       b.children.as<uint32_t>(), b.parents_leaf.as<uint32_t>(),
                       b.parents_internal.as<uint32_t>(), b.root_box.as<float>() + 6 * static_cast<size_t>(scene),
                       base, n, t_box, t_ch, t_pl, t_pi);
    PT_HIP(hipGetLastError());
    if (perm) PT_HIP(hipMemcpyAsync(perm, t_perm, n * 4ull, hipMemcpyDeviceToHost, st));
    if (parents_leaf) PT_HIP(hipMemcpyAsync(parents_leaf, t_pl, n * 4ull, hipMemcpyDeviceToHost, st));
    if (m > 0) {
        if (children) PT_HIP(hipMemcpyAsync(children, t_ch, m * 4, hipMemcpyDeviceToHost, st));
        if (parents_internal) PT_HIP(hipMemcpyAsync(parents_internal, t_pi, m * 4, hipMemcpyDeviceToHost, st));
        if (bboxes) PT_HIP(hipMemcpyAsync(bboxes, t_box, m * 24, hipMemcpyDeviceToHost, st));
    }
    PT_HIP(hipStreamSynchronize(st));
}

}  // namespace pt
