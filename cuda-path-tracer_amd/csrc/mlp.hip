// Radiance-cache MLP on MFMA (no reference counterpart: the reference has no
// learned component).  96 inputs -> 4 x (64, ReLU) -> 3 outputs with exp;
// fp16 weights and activations as MFMA operands, fp32 accumulation, fp32
// master weights, Adam.
//
// Layout trick: every layer computes Z^T = W^T X^T with
// v_mfma_f32_16x16x16f16, activations as the B operand.  A 16x16 D fragment
// (lane l: rows 4(l/16)+e, column l%16) is then exactly the B fragment of the
// next layer's k-block (lane l: k = 4(l/16)+e, column l%16), so a sample
// block's activations stay in registers from the encoding to the output.
// Weights are pre-packed in A-fragment order (8 bytes per lane per 16x16
// block) and staged in LDS.
//
// Inference uses gfx950's doubled-K v_mfma_f32_16x16x32_f16 with its own
// fragment packing (below, kperm) and four 16-sample blocks per wave.
//
// Training step (deterministic: no floating-point atomics), two launches:
//   k_mlp_train      per workgroup, 64-sample rounds (one 16-sample block per
//                    wave): forward, relative-L2 loss, backward deltas, then
//                    the layer inputs and deltas go to LDS feature-major and
//                    the four waves accumulate dW = act^T delta (MFMA, K =
//                    samples) and db = sum delta over the workgroup's
//                    samples; one fp32 partial per workgroup
//   k_mlp_adam_pack  sums the partials in workgroup order, Adam on the fp32
//                    masters, and writes each updated weight into its fp16
//                    fragments (forward, backward and inference layouts)
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "context.hpp"
#include "sampling.hpp"

namespace pt {
namespace {

typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 96, kHid = 64, kOut = 3, kLayers = 5;
__host__ __device__ constexpr int dim(int l) { return l == 0 ? kIn : (l == kLayers ? kOut : kHid); }
__host__ __device__ constexpr int w_off(int l) {
    return l == 0 ? 0 : w_off(l - 1) + dim(l - 1) * dim(l) + dim(l);
}
__host__ __device__ constexpr int b_off(int l) { return w_off(l) + dim(l) * dim(l + 1); }
constexpr int kParams = w_off(kLayers);  // 18,883
static_assert(kParams == 18883, "parameter count");
constexpr float kRelEps = 0.01f;

// 16-row/col blocks per layer: rows = output features (forward A), k = inputs.
__host__ __device__ constexpr int rb(int l) { return (dim(l + 1) + 15) / 16; }
__host__ __device__ constexpr int fb(int l) { return (dim(l) + 15) / 16; }
__host__ __device__ constexpr int fwd_base(int l) { return l == 0 ? 0 : fwd_base(l - 1) + rb(l - 1) * fb(l - 1); }
constexpr int kFwdFrags = fwd_base(kLayers);  // 76
// Backward A fragments (delta_{l-1} = W_l delta_l) for l = 1..4: rows = inputs, k = outputs.
__host__ __device__ constexpr int bwd_base(int l) { return l <= 1 ? 0 : bwd_base(l - 1) + fb(l - 1) * rb(l - 1); }
constexpr int kBwdFrags = bwd_base(kLayers);  // 52
constexpr int kFragHalves = 64 * 4;

// Inference fragments for gfx950's v_mfma_f32_16x16x32_f16 (A: lane L holds
// row L%16, k-elements 8(L/16)+j).  Two 16x16 D fragments of one layer --
// output rows 32q..32q+15 and 32q+16..32q+31 -- concatenated lane by lane
// (lane L: rows 4(L/16)+e of each) are the next layer's B operand for k-block
// q when the k index is permuted: element j of lane group g is feature
// kperm(q, g, j).  The weights are packed with the same permutation, so
// activations still never leave the registers.
__host__ __device__ constexpr int kperm(int q, int g, int j) { return 32 * q + (j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4)); }
__host__ __device__ constexpr int kb32(int l) { return (dim(l) + 31) / 32; }
__host__ __device__ constexpr int f32_base(int l) { return l == 0 ? 0 : f32_base(l - 1) + rb(l - 1) * kb32(l - 1); }
constexpr int kF32Frags = f32_base(kLayers);  // 12 + 3 * 8 + 2 = 38
static_assert(kF32Frags == 38, "16x16x32 fragment count");
constexpr int kF32Halves = 64 * 8;

constexpr int kThreads = 256;

__device__ __forceinline__ floatx4 mfma(half4 a, half4 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ floatx4 mfma32(half8 a, half8 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// Masters -> 16x16x32 forward fragment (l, r, q), lane L, element j:
// W_l[kperm(q, L/16, j)][16r + L%16], zero outside the layer's shape.
__device__ __forceinline__ void pack32_frag(const float* __restrict__ master, _Float16* __restrict__ f32,
                                            int frag, int lane) {
    int l = 0;
    while (l + 1 < kLayers && frag >= f32_base(l + 1)) ++l;
    const int k = frag - f32_base(l);
    const int r = k / kb32(l), q = k % kb32(l);
    const int ni = dim(l), no = dim(l + 1);
    const float* W = master + w_off(l);
    const int o = 16 * r + lane % 16;
    _Float16* out = f32 + static_cast<size_t>(frag) * kF32Halves + lane * 8;
    for (int j = 0; j < 8; ++j) {
        const int i = kperm(q, lane / 16, j);
        out[j] = (i < ni && o < no) ? static_cast<_Float16>(W[i * no + o]) : static_cast<_Float16>(0.0f);
    }
}

// Masters -> fp16 fragments.  Forward fragment (l, r, f), lane L, element e:
// W_l[16f + 4(L/16) + e][16r + L%16]; backward fragment (l, r, f):
// W_l[16r + L%16][16f + 4(L/16) + e]; zero outside the layer's shape; then
// the 16x16x32 inference fragments (pack32_frag).  Also advances the
// device-side Adam step counter by `advance` (one thread).
__global__ __launch_bounds__(kThreads) void k_mlp_pack(const float* __restrict__ master,
                                                       _Float16* __restrict__ fwd,
                                                       _Float16* __restrict__ bwd,
                                                       _Float16* __restrict__ f32,
                                                       uint32_t* __restrict__ step, uint32_t advance) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0 && advance) step[0] += advance;
    const int frag = t / 64, lane = t % 64;
    if (frag >= kFwdFrags + kBwdFrags + kF32Frags) return;
    if (frag >= kFwdFrags + kBwdFrags) {
        pack32_frag(master, f32, frag - kFwdFrags - kBwdFrags, lane);
        return;
    }
    int l = 0, r = 0, f = 0;
    bool forward = frag < kFwdFrags;
    if (forward) {
        while (l + 1 < kLayers && frag >= fwd_base(l + 1)) ++l;
        const int k = frag - fwd_base(l);
        r = k / fb(l);
        f = k % fb(l);
    } else {
        const int fr = frag - kFwdFrags;
        l = 1;
        while (l + 1 < kLayers && fr >= bwd_base(l + 1)) ++l;
        const int k = fr - bwd_base(l);
        r = k / rb(l);
        f = k % rb(l);
    }
    const int ni = dim(l), no = dim(l + 1);
    const float* W = master + w_off(l);
    _Float16* out = (forward ? fwd + frag * kFragHalves : bwd + (frag - kFwdFrags) * kFragHalves) + lane * 4;
    for (int e = 0; e < 4; ++e) {
        int i, o;
        if (forward) {
            i = 16 * f + 4 * (lane / 16) + e;
            o = 16 * r + lane % 16;
        } else {
            i = 16 * r + lane % 16;
            o = 16 * f + 4 * (lane / 16) + e;
        }
        out[e] = (i < ni && o < no) ? static_cast<_Float16>(W[i * no + o]) : static_cast<_Float16>(0.0f);
    }
}

__device__ __forceinline__ half4 to_half4(floatx4 v) {
    half4 h;
    h.x = static_cast<_Float16>(v.x);
    h.y = static_cast<_Float16>(v.y);
    h.z = static_cast<_Float16>(v.z);
    h.w = static_cast<_Float16>(v.w);
    return h;
}

struct SceneBox {
    float mn[3], inv_ext[3];
};

// Input features idx and idx + 1 (idx even, 0..94) of one sample from its
// 12 raw floats: below 72 the (sin, cos) pair of 2^k pi p for one axis and
// frequency -- sin/cos(2^k pi p) = sin/cos(2 pi frac(2^(k-1) p)), the
// hardware sine and cosine taking revolutions -- then the normal, incoming
// direction and albedo, then zero padding.
__device__ __forceinline__ void feature_pair(int idx, const float* p01, const float* raw, float& f0, float& f1) {
    if (idx < 72) {
        const int a = idx / 24, k = (idx % 24) / 2;
        const float rev = ldexpf(p01[a], k - 1);
        const float fr = rev - floorf(rev);
        f0 = __builtin_amdgcn_sinf(fr);
        f1 = __builtin_amdgcn_cosf(fr);
    } else {
        f0 = idx < 81 ? raw[3 + idx - 72] : 0.0f;
        f1 = idx + 1 < 81 ? raw[3 + idx + 1 - 72] : 0.0f;
    }
}

// Lane's view of one 16-sample block: the input fragments X[f] (k-block f).
__device__ __forceinline__ void encode_block(const float* __restrict__ raw, uint32_t n, uint32_t s,
                                             const SceneBox& box, half4* X) {
    const int lane = threadIdx.x & 63;
    float rw[12];
    if (s < n) {
        for (int k = 0; k < 12; ++k) rw[k] = raw[12 * static_cast<size_t>(s) + k];
    } else {
        for (int k = 0; k < 12; ++k) rw[k] = 0.0f;
    }
    float p01[3];
    for (int a = 0; a < 3; ++a) p01[a] = (rw[a] - box.mn[a]) * box.inv_ext[a];
#pragma unroll
    for (int f = 0; f < fb(0); ++f) {
        float a, b, c, d;
        feature_pair(16 * f + 4 * (lane / 16) + 0, p01, rw, a, b);
        feature_pair(16 * f + 4 * (lane / 16) + 2, p01, rw, c, d);
        X[f] = to_half4(floatx4{a, b, c, d});
    }
}

__device__ __forceinline__ half4 lds_frag(const uint2* __restrict__ frags, int idx) {
    const uint2 u = frags[idx * 64 + (threadIdx.x & 63)];
    half4 h;
    __builtin_memcpy(&h, &u, sizeof h);
    return h;
}

__device__ __forceinline__ floatx4 bias4(const float* __restrict__ b, int r, int nvalid) {
    const int lane = threadIdx.x & 63;
    floatx4 v;
    const int o = 16 * r + 4 * (lane / 16);
    v.x = o + 0 < nvalid ? b[o + 0] : 0.0f;
    v.y = o + 1 < nvalid ? b[o + 1] : 0.0f;
    v.z = o + 2 < nvalid ? b[o + 2] : 0.0f;
    v.w = o + 3 < nvalid ? b[o + 3] : 0.0f;
    return v;
}

__device__ __forceinline__ floatx4 relu4(floatx4 v) {
    v.x = v.x > 0.0f ? v.x : 0.0f;
    v.y = v.y > 0.0f ? v.y : 0.0f;
    v.z = v.z > 0.0f ? v.z : 0.0f;
    v.w = v.w > 0.0f ? v.w : 0.0f;
    return v;
}

// Forward of one block.  H[l][r] = fp16 input fragments of layer l (l = 1..4);
// returns the output pre-activation fragment (rows 0..2 valid on lanes 0..15).
__device__ __forceinline__ floatx4 forward_block(const uint2* __restrict__ frags,
                                                 const float* __restrict__ master, const half4* X,
                                                 half4 (*H)[4]) {
    // The four row blocks' accumulators advance together (independent MFMA
    // chains); each still sums its k-blocks in order.
    {
        floatx4 acc[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[r] = bias4(master + b_off(0), r, kHid);
#pragma unroll
        for (int f = 0; f < fb(0); ++f)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[r] = mfma(lds_frag(frags, fwd_base(0) + r * fb(0) + f), X[f], acc[r]);
#pragma unroll
        for (int r = 0; r < 4; ++r) H[1][r] = to_half4(relu4(acc[r]));
    }
#pragma unroll
    for (int l = 1; l < 4; ++l) {
        floatx4 acc[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[r] = bias4(master + b_off(l), r, kHid);
#pragma unroll
        for (int f = 0; f < 4; ++f)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[r] = mfma(lds_frag(frags, fwd_base(l) + r * 4 + f), H[l][f], acc[r]);
#pragma unroll
        for (int r = 0; r < 4; ++r) H[l + 1][r] = to_half4(relu4(acc[r]));
    }
    floatx4 z = bias4(master + b_off(4), 0, kOut);
#pragma unroll
    for (int f = 0; f < 4; ++f) z = mfma(lds_frag(frags, fwd_base(4) + f), H[4][f], z);
    return z;
}

// Global -> LDS copy of `count` elements by the workgroup, eight loads in
// flight per thread before their stores (a load-store loop would wait out
// one memory round trip per element).
template <class T>
__device__ __forceinline__ void stage(T* dst, const T* __restrict__ src, int count) {
    constexpr int kBatch = 8;
    const int step = static_cast<int>(blockDim.x);
    for (int base = static_cast<int>(threadIdx.x); base < count; base += kBatch * step) {
        T v[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const int k = base + j * step;
            if (k < count) v[j] = src[k];
        }
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const int k = base + j * step;
            if (k < count) dst[k] = v[j];
        }
    }
}

__device__ __forceinline__ void stage_frags(uint2* dst, const uint2* __restrict__ src, int nfrags) {
    stage(dst, src, nfrags * 64);
    __syncthreads();
}

// Input features kperm(q, g, 0..7) of one sample for the 16x16x32 B operand
// (lane group g): pairs kperm(q, g, j), +1 for even j.
__device__ __forceinline__ half8 encode32(int q, int g, const float* p01, const float* raw) {
    half8 x;
#pragma unroll
    for (int j = 0; j < 8; j += 2) {
        float f0, f1;
        feature_pair(kperm(q, g, j), p01, raw, f0, f1);
        x[j] = static_cast<_Float16>(f0);
        x[j + 1] = static_cast<_Float16>(f1);
    }
    return x;
}

__device__ __forceinline__ half8 lds_frag32(const uint4* __restrict__ frags, int idx) {
    const uint4 u = frags[idx * 64 + (threadIdx.x & 63)];
    half8 h;
    __builtin_memcpy(&h, &u, sizeof h);
    return h;
}

// Next layer's B operand for k-block q from the D fragments of row blocks
// 2q and 2q + 1 (ReLU, fp16).
__device__ __forceinline__ half8 join_relu(floatx4 a, floatx4 b) {
    half8 h;
    h[0] = static_cast<_Float16>(a.x > 0.0f ? a.x : 0.0f);
    h[1] = static_cast<_Float16>(a.y > 0.0f ? a.y : 0.0f);
    h[2] = static_cast<_Float16>(a.z > 0.0f ? a.z : 0.0f);
    h[3] = static_cast<_Float16>(a.w > 0.0f ? a.w : 0.0f);
    h[4] = static_cast<_Float16>(b.x > 0.0f ? b.x : 0.0f);
    h[5] = static_cast<_Float16>(b.y > 0.0f ? b.y : 0.0f);
    h[6] = static_cast<_Float16>(b.z > 0.0f ? b.z : 0.0f);
    h[7] = static_cast<_Float16>(b.w > 0.0f ? b.w : 0.0f);
    return h;
}

// Inference: out[3s + c] = exp(z_c).  A wave runs kInferBlocks 16-sample
// blocks at a time so each weight fragment read from LDS feeds kInferBlocks
// MFMAs.
constexpr int kInferBlocks = 4;

__global__ __launch_bounds__(kThreads) void k_mlp_infer(const uint4* __restrict__ f32,
                                                        const float* __restrict__ master,
                                                        const float* __restrict__ raw, uint32_t n,
                                                        SceneBox box, float* __restrict__ out) {
    __shared__ uint4 s_frag[kF32Frags * 64];
    __shared__ float s_bias[kLayers][kHid];
    stage(s_frag, f32, kF32Frags * 64);
    for (int k = threadIdx.x; k < kLayers * kHid; k += blockDim.x) {
        const int l = k / kHid, o = k % kHid;
        s_bias[l][o] = o < dim(l + 1) ? master[b_off(l) + o] : 0.0f;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, g = lane / 16;
    const uint32_t wave = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const uint32_t nwaves = gridDim.x * (blockDim.x / 64);
    auto bias = [&](int l, int r) {
        const int o = 16 * r + 4 * g;
        floatx4 v;
        v.x = s_bias[l][o + 0];
        v.y = s_bias[l][o + 1];
        v.z = s_bias[l][o + 2];
        v.w = s_bias[l][o + 3];
        return v;
    };
    constexpr uint32_t kSpan = 16 * kInferBlocks;
    const bool aligned = (reinterpret_cast<uintptr_t>(raw) & 15u) == 0u;  // 48-byte samples, 16-byte loads
    for (uint32_t base = wave * kSpan; base < n; base += nwaves * kSpan) {
        // Weight fragments are read from LDS where used: kept from one
        // iteration to the next they would take 152 VGPRs (one wave per SIMD).
        asm volatile("" ::: "memory");
        half8 X[kInferBlocks][3];
#pragma unroll
        for (int b = 0; b < kInferBlocks; ++b) {
            const uint32_t s = base + 16 * b + lane % 16;
            float rw[12];
            if (s < n && aligned) {
                const float4* rp = reinterpret_cast<const float4*>(raw + 12 * static_cast<size_t>(s));
                const float4 r0 = rp[0], r1 = rp[1], r2 = rp[2];
                rw[0] = r0.x; rw[1] = r0.y; rw[2] = r0.z; rw[3] = r0.w;
                rw[4] = r1.x; rw[5] = r1.y; rw[6] = r1.z; rw[7] = r1.w;
                rw[8] = r2.x; rw[9] = r2.y; rw[10] = r2.z; rw[11] = r2.w;
            } else if (s < n) {
                for (int k = 0; k < 12; ++k) rw[k] = raw[12 * static_cast<size_t>(s) + k];
            } else {
                for (int k = 0; k < 12; ++k) rw[k] = 0.0f;
            }
            float p01[3];
            for (int a = 0; a < 3; ++a) p01[a] = (rw[a] - box.mn[a]) * box.inv_ext[a];
#pragma unroll
            for (int q = 0; q < 3; ++q) X[b][q] = encode32(q, g, p01, rw);
        }
        // Layer 0: 96 -> 64.
        half8 H[kInferBlocks][2];
#pragma unroll
        for (int rp = 0; rp < 2; ++rp) {
            floatx4 acc[kInferBlocks][2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const floatx4 bv = bias(0, 2 * rp + h);
#pragma unroll
                for (int b = 0; b < kInferBlocks; ++b) acc[b][h] = bv;
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const half8 a = lds_frag32(s_frag, f32_base(0) + (2 * rp + h) * 3 + q);
#pragma unroll
                    for (int b = 0; b < kInferBlocks; ++b) acc[b][h] = mfma32(a, X[b][q], acc[b][h]);
                }
            }
#pragma unroll
            for (int b = 0; b < kInferBlocks; ++b) H[b][rp] = join_relu(acc[b][0], acc[b][1]);
        }
        // Layers 1..3: 64 -> 64.
#pragma unroll
        for (int l = 1; l < 4; ++l) {
            half8 Hn[kInferBlocks][2];
#pragma unroll
            for (int rp = 0; rp < 2; ++rp) {
                floatx4 acc[kInferBlocks][2];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const floatx4 bv = bias(l, 2 * rp + h);
#pragma unroll
                    for (int b = 0; b < kInferBlocks; ++b) acc[b][h] = bv;
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        const half8 a = lds_frag32(s_frag, f32_base(l) + (2 * rp + h) * 2 + q);
#pragma unroll
                        for (int b = 0; b < kInferBlocks; ++b) acc[b][h] = mfma32(a, H[b][q], acc[b][h]);
                    }
                }
#pragma unroll
                for (int b = 0; b < kInferBlocks; ++b) Hn[b][rp] = join_relu(acc[b][0], acc[b][1]);
            }
#pragma unroll
            for (int b = 0; b < kInferBlocks; ++b) {
                H[b][0] = Hn[b][0];
                H[b][1] = Hn[b][1];
            }
        }
        // Layer 4: 64 -> 3 (rows 0..2 of the first row block), exp.
        floatx4 z[kInferBlocks];
        const floatx4 b4 = bias(4, 0);
#pragma unroll
        for (int b = 0; b < kInferBlocks; ++b) z[b] = b4;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const half8 a = lds_frag32(s_frag, f32_base(4) + q);
#pragma unroll
            for (int b = 0; b < kInferBlocks; ++b) z[b] = mfma32(a, H[b][q], z[b]);
        }
#pragma unroll
        for (int b = 0; b < kInferBlocks; ++b) {
            const uint32_t s = base + 16 * b + lane % 16;
            if (lane < 16 && s < n) {
                out[3 * static_cast<size_t>(s) + 0] = expf(z[b].x);
                out[3 * static_cast<size_t>(s) + 1] = expf(z[b].y);
                out[3 * static_cast<size_t>(s) + 2] = expf(z[b].z);
            }
        }
    }
}

// Training rounds: 64 samples per workgroup round, one 16-sample block per
// wave.  Layer inputs act_l (act0: 96 rows, act1..4: 64) and deltas
// delta_l = dL/dz_l scaled by n (delta0..3: 64 rows, delta4: 16 rows with
// rows 3..15 zero) go to LDS feature-major, [row][sample] with a padded row.
constexpr int kTrainChunk = 64;
constexpr int kActRows = kIn + 4 * kHid;   // 352
constexpr int kDelRows = 4 * kHid + 16;    // 272
constexpr int kRowLd = kTrainChunk + 4;    // 68 halves: rows 34 banks apart
__host__ __device__ constexpr int act_row(int l) { return l == 0 ? 0 : kIn + kHid * (l - 1); }
__host__ __device__ constexpr int del_row(int l) { return kHid * l; }
constexpr int kWTiles = fwd_base(kLayers);  // 76 weight tiles (16 x 16), 19 per wave
constexpr int kTilesPerWave = kWTiles / 4;
static_assert(kWTiles % 4 == 0, "weight tiles split evenly over four waves");
constexpr uint32_t kMaxTrainBlocks = 256;  // workgroups (gradient partials) of one step

// Deltas are gradients of n * loss (a static loss scale of n keeps them well
// inside fp16's normal range); they are clamped to fp16's finite range so an
// outlier sample cannot turn a whole step into inf/NaN.
constexpr float kHalfMax = 65504.0f;
__device__ __forceinline__ float sat(float x) {
    return x > kHalfMax ? kHalfMax : (x < -kHalfMax ? -kHalfMax : x);
}

__device__ __forceinline__ half4 mask4(floatx4 d, half4 h) {
    floatx4 m;
    m.x = h.x > static_cast<_Float16>(0.0f) ? sat(d.x) : 0.0f;
    m.y = h.y > static_cast<_Float16>(0.0f) ? sat(d.y) : 0.0f;
    m.z = h.z > static_cast<_Float16>(0.0f) ? sat(d.z) : 0.0f;
    m.w = h.w > static_cast<_Float16>(0.0f) ? sat(d.w) : 0.0f;
    return to_half4(m);
}

// A 16x16 D-layout fragment (lane: rows 4(l/16)+e, column = sample l%16) into
// LDS rows row0 .. row0+15, columns col0 .. col0+15.
__device__ __forceinline__ void lds_put(_Float16 (*rows)[kRowLd], int row0, int col0, half4 v) {
    const int lane = threadIdx.x & 63;
    const int r = row0 + 4 * (lane / 16), c = col0 + lane % 16;
    rows[r + 0][c] = v.x;
    rows[r + 1][c] = v.y;
    rows[r + 2][c] = v.z;
    rows[r + 3][c] = v.w;
}

// MFMA operand from LDS: lane l reads row row0 + l%16, samples k0 + 4(l/16) .. +3.
__device__ __forceinline__ half4 lds_get(const _Float16 (*rows)[kRowLd], int row0, int k0) {
    const int lane = threadIdx.x & 63;
    half4 h;
    __builtin_memcpy(&h, &rows[row0 + lane % 16][k0 + 4 * (lane / 16)], sizeof h);
    return h;
}

// One training step's gradient partials.  Workgroup b handles rounds b,
// b + gridDim.x, ... and writes part[b][param] (weights and biases, fp32).
// loss[s] = the sample's relative-L2 loss summed over channels.  A sample
// whose first target channel is NaN is masked: zero loss, zero deltas (it
// still counts in the mean's n).  Thread 0 of workgroup 0 advances the Adam
// step counter by `advance` (read by k_mlp_adam_pack after this launch).
__global__ __launch_bounds__(kThreads) void k_mlp_train(
    const uint2* __restrict__ fwd, const uint2* __restrict__ bwd, const float* __restrict__ master,
    const float* __restrict__ raw, const float* __restrict__ target, uint32_t n, SceneBox box,
    float* __restrict__ part, float* __restrict__ loss, uint32_t* __restrict__ step, uint32_t advance) {
    __shared__ uint2 s_fwd[kFwdFrags * 64];
    __shared__ uint2 s_bwd[kBwdFrags * 64];
    __shared__ _Float16 s_act[kActRows][kRowLd];
    __shared__ _Float16 s_del[kDelRows][kRowLd];
    if (blockIdx.x == 0 && threadIdx.x == 0 && advance) step[0] += advance;
    stage(s_bwd, bwd, kBwdFrags * 64);
    stage_frags(s_fwd, fwd, kFwdFrags);
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x / 64;
    floatx4 acc[kTilesPerWave];
#pragma unroll
    for (int t = 0; t < kTilesPerWave; ++t) acc[t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    // This wave's dW tiles: LDS rows of their layer inputs and deltas
    // (wave-uniform, computed once).
    uint32_t arow[kTilesPerWave], drow[kTilesPerWave];
#pragma unroll
    for (int t = 0; t < kTilesPerWave; ++t) {
        const int task = wave + 4 * t;
        int l = 0;
        while (l + 1 < kLayers && task >= fwd_base(l + 1)) ++l;
        const int k = task - fwd_base(l);
        const int ri = k % fb(l), co = k / fb(l);
        arow[t] = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(act_row(l) + 16 * ri));
        drow[t] = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>((l == 4 ? kHid * 4 : del_row(l)) + 16 * co));
    }
    // Bias sums: thread t owns delta row t (layers 0..3) and, for t < 3,
    // output bias t (delta4 row t).
    float bsum = 0.0f, bsum4 = 0.0f;
    for (uint32_t round = blockIdx.x; round * kTrainChunk < n; round += gridDim.x) {
        const uint32_t s = round * kTrainChunk + 16 * wave + lane % 16;
        const int col0 = 16 * wave;
        half4 X[fb(0)];
        half4 H[5][4];
        encode_block(raw, n, s, box, X);
        const floatx4 z = forward_block(s_fwd, master, X, H);
        for (int f = 0; f < fb(0); ++f) lds_put(s_act, act_row(0) + 16 * f, col0, X[f]);
        for (int l = 1; l < 5; ++l)
            for (int r = 0; r < 4; ++r) lds_put(s_act, act_row(l) + 16 * r, col0, H[l][r]);
        // Output delta: d/dz of (e^z - t)^2 / (sg(e^z)^2 + eps), per sample.
        floatx4 d4 = {0.0f, 0.0f, 0.0f, 0.0f};
        if (lane < 16 && s < n && target[3 * static_cast<size_t>(s)] == target[3 * static_cast<size_t>(s)]) {
            float lsum = 0.0f;
            float zz[3] = {z.x, z.y, z.z};
            float dd[3];
            for (int c = 0; c < 3; ++c) {
                const float y = expf(zz[c]);
                const float t = target[3 * static_cast<size_t>(s) + c];
                const float den = y * y + kRelEps;
                lsum += (y - t) * (y - t) / den;
                const float g = 2.0f * (y - t) / den * y;
                dd[c] = g == g ? sat(g) : 0.0f;  // NaN (inf/inf) -> no update
            }
            d4.x = dd[0];
            d4.y = dd[1];
            d4.z = dd[2];
            if (loss) loss[s] = lsum;
        } else if (lane < 16 && s < n && loss) {
            loss[s] = 0.0f;
        }
        const half4 D4 = to_half4(d4);
        lds_put(s_del, del_row(4), col0, D4);
        // delta_3 = W4 delta_4 (masked by ReLU of layer 3's output = H[4]).
        half4 Dn[4];
        for (int r = 0; r < 4; ++r) {
            floatx4 a = {0.0f, 0.0f, 0.0f, 0.0f};
            a = mfma(lds_frag(s_bwd, bwd_base(4) + r), D4, a);
            Dn[r] = mask4(a, H[4][r]);
            lds_put(s_del, del_row(3) + 16 * r, col0, Dn[r]);
        }
        for (int l = 3; l >= 1; --l) {
            floatx4 a[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) a[r] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int f = 0; f < 4; ++f)
#pragma unroll
                for (int r = 0; r < 4; ++r) a[r] = mfma(lds_frag(s_bwd, bwd_base(l) + r * 4 + f), Dn[f], a[r]);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                Dn[r] = mask4(a[r], H[l][r]);
                lds_put(s_del, del_row(l - 1) + 16 * r, col0, Dn[r]);
            }
        }
        __syncthreads();
        // dW tiles of this wave over the round's 64 samples (K = 4 x 16).
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
#pragma unroll
            for (int t = 0; t < kTilesPerWave; ++t) {
                const half4 a = lds_get(s_act, static_cast<int>(arow[t]), 16 * kb);
                const half4 b = lds_get(s_del, static_cast<int>(drow[t]), 16 * kb);
                acc[t] = mfma(a, b, acc[t]);
            }
        }
        {
            const int t = threadIdx.x;
            float a = 0.0f, a4 = 0.0f;
            for (int j = 0; j < kTrainChunk; ++j) {
                a += static_cast<float>(s_del[t][j]);
                if (t < kOut) a4 += static_cast<float>(s_del[kHid * 4 + t][j]);
            }
            bsum += a;
            bsum4 += a4;
        }
        __syncthreads();  // the next round rewrites the LDS rows
    }
    float* out = part + static_cast<size_t>(blockIdx.x) * kParams;
#pragma unroll
    for (int t = 0; t < kTilesPerWave; ++t) {
        const int task = wave + 4 * t;
        int l = 0;
        while (l + 1 < kLayers && task >= fwd_base(l + 1)) ++l;
        const int k = task - fwd_base(l);
        const int ri = k % fb(l), co = k / fb(l);
        const int ni = dim(l), no = dim(l + 1);
        // D lane: column (output) 16co + lane%16, rows (inputs) 16ri + 4(lane/16) + e.
        const int o = 16 * co + lane % 16;
        for (int e = 0; e < 4; ++e) {
            const int i = 16 * ri + 4 * (lane / 16) + e;
            if (i < ni && o < no) out[w_off(l) + i * no + o] = acc[t][e];
        }
    }
    {
        const int t = threadIdx.x;
        out[b_off(t / kHid) + t % kHid] = bsum;
        if (t < kOut) out[b_off(4) + t] = bsum4;
    }
}

// Adam (Kingma & Ba 2015) on the fp32 masters, one thread per parameter; the
// gradient is the sum of the workgroup partials in workgroup order divided
// by n (the loss is a mean over n).  The bias corrections use step t =
// step[0] (k_mlp_train advanced it), read on the device so a captured frame
// replays with the right t.  The updated weight is written into its fp16
// fragments: forward (16x16x16, layer input as k), backward (16x16x16, layer
// output as k) and inference (16x16x32, permuted k).  grad_out: only the
// gradient, no update.
// One workgroup per 64 parameters: the four waves sum the partials
// s = w, w + 4, w + 8, ... of their parameter (coalesced rows), then the
// first wave adds the four sums in wave order -- a fixed order, so the
// gradient is the same bits on every call.
constexpr int kAdamParams = 64;

__global__ __launch_bounds__(kThreads) void k_mlp_adam_pack(float* __restrict__ master, float* __restrict__ m,
                                                            float* __restrict__ v, const float* __restrict__ part,
                                                            uint32_t nparts, float inv_n, float lr, float b1,
                                                            float b2, float eps, const uint32_t* __restrict__ step,
                                                            float* __restrict__ grad_out, _Float16* __restrict__ fwd,
                                                            _Float16* __restrict__ bwd, _Float16* __restrict__ f32) {
    __shared__ float s_sum[kThreads / kAdamParams][kAdamParams];
    const int lane = threadIdx.x % kAdamParams, wave = threadIdx.x / kAdamParams;
    const int p = blockIdx.x * kAdamParams + lane;
    constexpr int kWaves = kThreads / kAdamParams;
    float acc = 0.0f;
    if (p < kParams) {
        uint32_t s = static_cast<uint32_t>(wave);
#pragma unroll 8
        for (; s < nparts; s += kWaves) acc += part[static_cast<size_t>(s) * kParams + p];
    }
    s_sum[wave][lane] = acc;
    __syncthreads();
    if (wave != 0 || p >= kParams) return;
    float g = ((s_sum[0][lane] + s_sum[1][lane]) + s_sum[2][lane]) + s_sum[3][lane];
    static_assert(kWaves == 4, "four partial sums per parameter");
    g *= inv_n;
    if (grad_out) {
        grad_out[p] = g;
        return;
    }
    const float t = static_cast<float>(step[0]);
    const float c1 = 1.0f - powf(b1, t);
    const float c2 = 1.0f - powf(b2, t);
    const float mk = b1 * m[p] + (1.0f - b1) * g;
    const float vk = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mk;
    v[p] = vk;
    const float w = master[p] - lr * (mk / c1) / (sqrtf(vk / c2) + eps);
    master[p] = w;
    int l = 0;
    while (l + 1 < kLayers && p >= w_off(l + 1)) ++l;
    if (p >= b_off(l)) return;  // a bias: no fragment
    const int no = dim(l + 1);
    const int i = (p - w_off(l)) / no, o = (p - w_off(l)) % no;
    const _Float16 h = static_cast<_Float16>(w);
    // Forward fragment (l, o/16, i/16): lane 16((i%16)/4) + o%16, element i%4.
    fwd[(fwd_base(l) + (o / 16) * fb(l) + i / 16) * kFragHalves + (16 * ((i % 16) / 4) + o % 16) * 4 + i % 4] = h;
    // Backward fragment (l, i/16, o/16), l >= 1: lane 16((o%16)/4) + i%16, element o%4.
    if (l >= 1)
        bwd[(bwd_base(l) + (i / 16) * rb(l) + o / 16) * kFragHalves + (16 * ((o % 16) / 4) + i % 16) * 4 + o % 4] = h;
    // Inference fragment (l, o/16, i/32): k index i - 32q = kperm's position.
    const int q = i / 32, ip = i - 32 * q;
    const int g4 = ip < 16 ? ip / 4 : (ip - 16) / 4;
    const int j = ip < 16 ? ip % 4 : 4 + (ip - 16) % 4;
    f32[(f32_base(l) + (o / 16) * kb32(l) + q) * kF32Halves + (16 * g4 + o % 16) * 8 + j] = h;
}

uint32_t blocks_for(uint64_t work, uint32_t per) { return static_cast<uint32_t>((work + per - 1) / per); }


}  // namespace

int RadianceCache::param_count() { return kParams; }

void RadianceCache::set_box(const float* bmin, const float* bmax) {
    for (int k = 0; k < 3; ++k) {
        box_min[k] = bmin[k];
        box_inv[k] = 1.0f / (bmax[k] - bmin[k]);
    }
}

void RadianceCache::init(uint64_t seed, hipStream_t st) {
    // He-uniform weights for the ReLU layers, a small output layer, zero
    // biases; drawn on the host from SplitMix64.
    std::vector<float> h(kParams, 0.0f);
    uint64_t s = seed;
    auto u01 = [&]() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return static_cast<float>(z >> 40) * (1.0f / 16777216.0f);
    };
    for (int l = 0; l < kLayers; ++l) {
        const int ni = dim(l), no = dim(l + 1);
        const float a = (l + 1 < kLayers ? std::sqrt(6.0f / ni) : 0.1f * std::sqrt(6.0f / ni));
        for (int k = 0; k < ni * no; ++k) h[w_off(l) + k] = (2.0f * u01() - 1.0f) * a;
    }
    set_params(h.data(), st);
}

void RadianceCache::set_params(const float* host, hipStream_t st) {
    master.reserve(kParams * sizeof(float));
    m.reserve(kParams * sizeof(float));
    v.reserve(kParams * sizeof(float));
    fwd.reserve(static_cast<size_t>(kFwdFrags) * kFragHalves * 2);
    f32.reserve(static_cast<size_t>(kF32Frags) * kF32Halves * 2);
    bwd.reserve(static_cast<size_t>(kBwdFrags) * kFragHalves * 2);
    dstep.reserve(sizeof(uint32_t));
    PT_HIP(hipMemcpyAsync(master.as<void>(), host, kParams * sizeof(float), hipMemcpyHostToDevice, st));
    PT_HIP(hipMemsetAsync(m.as<void>(), 0, kParams * sizeof(float), st));
    PT_HIP(hipMemsetAsync(v.as<void>(), 0, kParams * sizeof(float), st));
    PT_HIP(hipMemsetAsync(dstep.as<void>(), 0, sizeof(uint32_t), st));
    PT_HIP(hipStreamSynchronize(st));
    ready = true;
    pack(st, 0);
}

void RadianceCache::get_params(float* host, hipStream_t st) {
    if (!ready) throw Error(PT_E_INVALID_ARG, "radiance cache not initialised");
    PT_HIP(hipMemcpyAsync(host, master.as<void>(), kParams * sizeof(float), hipMemcpyDeviceToHost, st));
    PT_HIP(hipStreamSynchronize(st));
}

void RadianceCache::pack(hipStream_t st, uint32_t advance) {
    hipLaunchKernelGGL(k_mlp_pack, dim3(blocks_for((kFwdFrags + kBwdFrags + kF32Frags) * 64, kThreads)),
                       dim3(kThreads), 0, st, master.as<float>(), fwd.as<_Float16>(), bwd.as<_Float16>(),
                       f32.as<_Float16>(), dstep.as<uint32_t>(), advance);
    PT_HIP(hipGetLastError());
}

void RadianceCache::infer(const float* d_raw, uint32_t n, float* d_out, hipStream_t st) {
    if (!ready) throw Error(PT_E_INVALID_ARG, "radiance cache not initialised");
    if (n == 0) return;
    SceneBox b;
    for (int k = 0; k < 3; ++k) {
        b.mn[k] = box_min[k];
        b.inv_ext[k] = box_inv[k];
    }
    const uint32_t blocks = std::min<uint32_t>(blocks_for(n, 4 * 16 * kInferBlocks), 2048);
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(kThreads), 0, st, f32.as<uint4>(), master.as<float>(),
                       d_raw, n, b, d_out);
    PT_HIP(hipGetLastError());
}

// Workgroups (gradient partials) of a training step over n samples.
uint32_t train_blocks(uint32_t n) {
    return std::max<uint32_t>(1u, std::min<uint32_t>(kMaxTrainBlocks, blocks_for(n, kTrainChunk)));
}

void RadianceCache::prepare_train(uint32_t n) {
    part.reserve(static_cast<size_t>(train_blocks(n)) * kParams * sizeof(float));
    loss.reserve((static_cast<size_t>(n) + 16) * sizeof(float));
}

void RadianceCache::launch_train(const float* d_raw, const float* d_target, uint32_t n, float* d_loss,
                                 float* d_grad_out, hipStream_t st) {
    const uint32_t nb = train_blocks(n);
    SceneBox b;
    for (int k = 0; k < 3; ++k) {
        b.mn[k] = box_min[k];
        b.inv_ext[k] = box_inv[k];
    }
    hipLaunchKernelGGL(k_mlp_train, dim3(nb), dim3(kThreads), 0, st, fwd.as<uint2>(), bwd.as<uint2>(),
                       master.as<float>(), d_raw, d_target, n, b, part.as<float>(),
                       d_loss ? d_loss : loss.as<float>(), dstep.as<uint32_t>(), d_grad_out ? 0u : 1u);
    hipLaunchKernelGGL(k_mlp_adam_pack, dim3(blocks_for(kParams, kAdamParams)), dim3(kThreads), 0, st,
                       master.as<float>(), m.as<float>(), v.as<float>(), part.as<float>(), nb,
                       1.0f / static_cast<float>(n), lr, beta1, beta2, eps, dstep.as<uint32_t>(), d_grad_out,
                       fwd.as<_Float16>(), bwd.as<_Float16>(), f32.as<_Float16>());
    PT_HIP(hipGetLastError());
}

void RadianceCache::train(const float* d_raw, const float* d_target, uint32_t n, float* d_loss,
                          float* d_grad_out, hipStream_t st) {
    if (!ready) throw Error(PT_E_INVALID_ARG, "radiance cache not initialised");
    if (n == 0) return;
    prepare_train(n);
    launch_train(d_raw, d_target, n, d_loss, d_grad_out, st);
}

}  // namespace pt
