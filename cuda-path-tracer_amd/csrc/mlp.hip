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
// Training step (deterministic: no floating-point atomics):
//   k_mlp_fwd_bwd   per 16-sample block: forward, relative-L2 loss, backward
//                   deltas; stores layer inputs and deltas, feature-major fp16
//   k_mlp_wgrad     dW = act^T delta and db = sum delta as split-K MFMA GEMMs,
//                   one partial per sample slice
//   k_mlp_adam      sums the partials in slice order, Adam on the masters
//   k_mlp_pack      masters -> fp16 fragments for the next step
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "context.hpp"
#include "sampling.hpp"

namespace pt {
namespace {

typedef _Float16 half4 __attribute__((ext_vector_type(4)));
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

constexpr int kThreads = 256;

__device__ __forceinline__ floatx4 mfma(half4 a, half4 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}

// Masters -> fp16 fragments.  Forward fragment (l, r, f), lane L, element e:
// W_l[16f + 4(L/16) + e][16r + L%16]; backward fragment (l, r, f):
// W_l[16r + L%16][16f + 4(L/16) + e]; zero outside the layer's shape.
// Also advances the device-side Adam step counter by `advance` (one thread).
__global__ __launch_bounds__(kThreads) void k_mlp_pack(const float* __restrict__ master,
                                                       _Float16* __restrict__ fwd,
                                                       _Float16* __restrict__ bwd,
                                                       uint32_t* __restrict__ step, uint32_t advance) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0 && advance) step[0] += advance;
    const int frag = t / 64, lane = t % 64;
    if (frag >= kFwdFrags + kBwdFrags) return;
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

// Input feature `idx` (0..95) of one sample from its 12 raw floats.
__device__ __forceinline__ float feature(int idx, const float* p01, const float* raw) {
    if (idx < 72) {
        const int a = idx / 24, k = (idx % 24) / 2;
        // sin/cos(2^k pi p) = sin/cos(2 pi frac(2^(k-1) p)); the scaling is exact.
        const float rev = ldexpf(p01[a], k - 1);
        const float fr = rev - floorf(rev);
        float s, c;
        sincos2pi(fr, s, c);
        return (idx & 1) ? c : s;
    }
    if (idx < 81) return raw[3 + idx - 72];
    return 0.0f;
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
        floatx4 v;
        v.x = feature(16 * f + 4 * (lane / 16) + 0, p01, rw);
        v.y = feature(16 * f + 4 * (lane / 16) + 1, p01, rw);
        v.z = feature(16 * f + 4 * (lane / 16) + 2, p01, rw);
        v.w = feature(16 * f + 4 * (lane / 16) + 3, p01, rw);
        X[f] = to_half4(v);
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
    for (int r = 0; r < 4; ++r) {
        floatx4 acc = bias4(master + b_off(0), r, kHid);
#pragma unroll
        for (int f = 0; f < fb(0); ++f) acc = mfma(lds_frag(frags, fwd_base(0) + r * fb(0) + f), X[f], acc);
        H[1][r] = to_half4(relu4(acc));
    }
#pragma unroll
    for (int l = 1; l < 4; ++l) {
        for (int r = 0; r < 4; ++r) {
            floatx4 acc = bias4(master + b_off(l), r, kHid);
#pragma unroll
            for (int f = 0; f < 4; ++f) acc = mfma(lds_frag(frags, fwd_base(l) + r * 4 + f), H[l][f], acc);
            H[l + 1][r] = to_half4(relu4(acc));
        }
    }
    floatx4 z = bias4(master + b_off(4), 0, kOut);
#pragma unroll
    for (int f = 0; f < 4; ++f) z = mfma(lds_frag(frags, fwd_base(4) + f), H[4][f], z);
    return z;
}

__device__ __forceinline__ void stage_frags(uint2* dst, const uint2* __restrict__ src, int nfrags) {
    for (int k = threadIdx.x; k < nfrags * 64; k += blockDim.x) dst[k] = src[k];
    __syncthreads();
}

// Inference: out[3s + c] = exp(z_c).
__global__ __launch_bounds__(kThreads) void k_mlp_infer(const uint2* __restrict__ fwd,
                                                        const float* __restrict__ master,
                                                        const float* __restrict__ raw, uint32_t n,
                                                        SceneBox box, float* __restrict__ out) {
    __shared__ uint2 s_frag[kFwdFrags * 64];
    stage_frags(s_frag, fwd, kFwdFrags);
    const int lane = threadIdx.x & 63;
    const uint32_t wave = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const uint32_t nwaves = gridDim.x * (blockDim.x / 64);
    for (uint32_t blk = wave; blk * 16 < n; blk += nwaves) {
        const uint32_t s = blk * 16 + lane % 16;
        half4 X[fb(0)];
        half4 H[5][4];
        encode_block(raw, n, s, box, X);
        const floatx4 z = forward_block(s_frag, master, X, H);
        if (lane < 16 && s < n) {
            out[3 * static_cast<size_t>(s) + 0] = expf(z.x);
            out[3 * static_cast<size_t>(s) + 1] = expf(z.y);
            out[3 * static_cast<size_t>(s) + 2] = expf(z.z);
        }
    }
}

// Store a 16x16 fragment (lane: rows 4(l/16)+e, column = sample l%16) into a
// feature-major fp16 array [feature][ld] at feature row0, sample s0.
__device__ __forceinline__ void store_fm(_Float16* __restrict__ a, size_t ld, int row0, uint32_t s,
                                         uint32_t n, half4 v, int nrows) {
    const int lane = threadIdx.x & 63;
    if (s >= n) return;
    const int r = row0 + 4 * (lane / 16);
    if (r + 0 < nrows) a[static_cast<size_t>(r + 0) * ld + s] = v.x;
    if (r + 1 < nrows) a[static_cast<size_t>(r + 1) * ld + s] = v.y;
    if (r + 2 < nrows) a[static_cast<size_t>(r + 2) * ld + s] = v.z;
    if (r + 3 < nrows) a[static_cast<size_t>(r + 3) * ld + s] = v.w;
}

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

// Forward + loss + backward per 16-sample block.  Stores, feature-major with
// leading dimension ld (>= n, multiple of 16): act_l = input of layer l
// (act0: 96 rows, act1..4: 64) and delta_l = dL/dz_l scaled by n
// (delta0..3: 64 rows, delta4: 16 rows with rows 3..15 zero).  loss[s] is the
// sample's sum over channels.  A sample whose first target channel is NaN is
// masked: zero loss, zero deltas (it still counts in the mean's n).
__global__ __launch_bounds__(kThreads) void k_mlp_fwd_bwd(
    const uint2* __restrict__ fwd, const uint2* __restrict__ bwd, const float* __restrict__ master,
    const float* __restrict__ raw, const float* __restrict__ target, uint32_t n, size_t ld,
    SceneBox box, _Float16* __restrict__ acts, _Float16* __restrict__ deltas,
    float* __restrict__ loss) {
    __shared__ uint2 s_fwd[kFwdFrags * 64];
    __shared__ uint2 s_bwd[kBwdFrags * 64];
    for (int k = threadIdx.x; k < kBwdFrags * 64; k += blockDim.x) s_bwd[k] = bwd[k];
    stage_frags(s_fwd, fwd, kFwdFrags);
    const int lane = threadIdx.x & 63;
    const uint32_t wave = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const uint32_t nwaves = gridDim.x * (blockDim.x / 64);
    _Float16* act[5];
    act[0] = acts;
    for (int l = 1; l < 5; ++l) act[l] = acts + (kIn + 64 * (l - 1)) * ld;
    _Float16* dl[5];
    for (int l = 0; l < 5; ++l) dl[l] = deltas + 64 * l * ld;
    for (uint32_t blk = wave; blk * 16 < n; blk += nwaves) {
        const uint32_t s = blk * 16 + lane % 16;
        half4 X[fb(0)];
        half4 H[5][4];
        encode_block(raw, n, s, box, X);
        const floatx4 z = forward_block(s_fwd, master, X, H);
        for (int f = 0; f < fb(0); ++f) store_fm(act[0], ld, 16 * f, s, n, X[f], kIn);
        for (int l = 1; l < 5; ++l)
            for (int r = 0; r < 4; ++r) store_fm(act[l], ld, 16 * r, s, n, H[l][r], kHid);
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
            loss[s] = lsum;
        } else if (lane < 16 && s < n) {
            loss[s] = 0.0f;
        }
        half4 D[4];
        D[0] = to_half4(d4);
        store_fm(dl[4], ld, 0, s, n, D[0], 16);
        // delta_3 = W4 delta_4 (masked by ReLU of layer 3's output = H[4]).
        half4 Dn[4];
        for (int r = 0; r < 4; ++r) {
            floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = mfma(lds_frag(s_bwd, bwd_base(4) + r), D[0], acc);
            Dn[r] = mask4(acc, H[4][r]);
            store_fm(dl[3], ld, 16 * r, s, n, Dn[r], kHid);
        }
        for (int l = 3; l >= 1; --l) {
            half4 Dp[4];
            for (int r = 0; r < 4; ++r) {
                floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int f = 0; f < 4; ++f) acc = mfma(lds_frag(s_bwd, bwd_base(l) + r * 4 + f), Dn[f], acc);
                Dp[r] = mask4(acc, H[l][r]);
                store_fm(dl[l - 1], ld, 16 * r, s, n, Dp[r], kHid);
            }
            for (int r = 0; r < 4; ++r) Dn[r] = Dp[r];
        }
    }
}

// Split-K weight and bias gradients.  Task (layer l, row block r, column
// block c, slice) computes dW_l[16r..][16c..] over samples
// [slice * slice_len, ...) with one MFMA per 16 samples; bias tasks sum the
// deltas.  Each parameter is written once per slice into part[slice][param].
struct GradTask {
    int layer, r, c;  // c = -1: bias block r of the layer
};

__global__ __launch_bounds__(kThreads) void k_mlp_wgrad(const _Float16* __restrict__ acts,
                                                        const _Float16* __restrict__ deltas,
                                                        uint32_t n, size_t ld, uint32_t slice_len,
                                                        uint32_t nslices, float* __restrict__ part) {
    const int lane = threadIdx.x & 63;
    const uint32_t wave = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    // Task list: weight blocks of every layer, then bias blocks.
    constexpr int kWTasks = fwd_base(kLayers);  // fb(l) * rb(l) per layer
    constexpr int kBTasks = 4 * 4 + 1;
    const uint32_t ntasks = (kWTasks + kBTasks) * nslices;
    if (wave >= ntasks) return;
    const uint32_t slice = wave % nslices;
    int task = static_cast<int>(wave / nslices);
    const uint32_t s_begin = slice * slice_len;
    const uint32_t s_end = min(n, s_begin + slice_len);
    float* out = part + static_cast<size_t>(slice) * kParams;
    const _Float16* act[5];
    act[0] = acts;
    for (int l = 1; l < 5; ++l) act[l] = acts + (kIn + 64 * (l - 1)) * ld;
    const _Float16* dl[5];
    for (int l = 0; l < 5; ++l) dl[l] = deltas + 64 * l * ld;
    if (task < kWTasks) {
        int l = 0;
        while (l + 1 < kLayers && task >= fwd_base(l + 1)) ++l;
        const int k = task - fwd_base(l);
        const int r = k % fb(l);  // rows: input features
        const int c = k / fb(l);  // columns: output features
        const int ni = dim(l), no = dim(l + 1);
        floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        const _Float16* arow = act[l] + static_cast<size_t>(16 * r + lane % 16) * ld;
        const _Float16* drow = dl[l] + static_cast<size_t>(16 * c + lane % 16) * ld;
        const bool arow_ok = 16 * r + lane % 16 < ni;
        const bool drow_ok = 16 * c + lane % 16 < (l == 4 ? 16 : no);
        for (uint32_t s0 = s_begin; s0 < s_end; s0 += 16) {
            const uint32_t s = s0 + 4 * (lane / 16);
            half4 a = {0, 0, 0, 0}, b = {0, 0, 0, 0};
            if (arow_ok) {
                for (int e = 0; e < 4; ++e)
                    if (s + e < s_end) a[e] = arow[s + e];
            }
            if (drow_ok) {
                for (int e = 0; e < 4; ++e)
                    if (s + e < s_end) b[e] = drow[s + e];
            }
            acc = mfma(a, b, acc);
        }
        // D lane: column (output) lane % 16, rows (inputs) 4(lane/16) + e.
        const int o = 16 * c + lane % 16;
        for (int e = 0; e < 4; ++e) {
            const int i = 16 * r + 4 * (lane / 16) + e;
            if (i < ni && o < no) out[w_off(l) + i * no + o] = acc[e];
        }
        return;
    }
    task -= kWTasks;
    const int l = task / 4, r = task % 4;
    const int no = dim(l + 1);
    const int o = 16 * r + lane % 16;
    float sum = 0.0f;
    if (o < no) {
        const _Float16* drow = dl[l] + static_cast<size_t>(o) * ld;
        for (uint32_t s = s_begin + lane / 16; s < s_end; s += 4) sum += static_cast<float>(drow[s]);
    }
    sum += __shfl_xor(sum, 16, 64);
    sum += __shfl_xor(sum, 32, 64);
    if (lane < 16 && o < no) out[b_off(l) + o] = sum;
}

// Adam (Kingma & Ba 2015) on the fp32 masters; the gradient is the sum of the
// slice partials in slice order divided by n (the loss is a mean over n).
// The bias corrections use step t = step[0] + 1, read on the device so a
// captured frame replays with the right t (k_mlp_pack advances it).
__global__ __launch_bounds__(kThreads) void k_mlp_adam(float* __restrict__ master, float* __restrict__ m,
                                                       float* __restrict__ v, const float* __restrict__ part,
                                                       uint32_t nslices, float inv_n, float lr, float b1,
                                                       float b2, float eps, const uint32_t* __restrict__ step,
                                                       float* __restrict__ grad_out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kParams) return;
    float g = 0.0f;
    for (uint32_t s = 0; s < nslices; ++s) g += part[static_cast<size_t>(s) * kParams + p];
    g *= inv_n;
    if (grad_out) {
        grad_out[p] = g;
        return;
    }
    const float t = static_cast<float>(step[0] + 1u);
    const float c1 = 1.0f - powf(b1, t);
    const float c2 = 1.0f - powf(b2, t);
    const float mk = b1 * m[p] + (1.0f - b1) * g;
    const float vk = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mk;
    v[p] = vk;
    master[p] = master[p] - lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

uint32_t blocks_for(uint64_t work, uint32_t per) { return static_cast<uint32_t>((work + per - 1) / per); }

constexpr uint32_t kSliceLen = 1024;  // samples per weight-gradient partial

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
    hipLaunchKernelGGL(k_mlp_pack, dim3(blocks_for((kFwdFrags + kBwdFrags) * 64, kThreads)), dim3(kThreads), 0,
                       st, master.as<float>(), fwd.as<_Float16>(), bwd.as<_Float16>(), dstep.as<uint32_t>(),
                       advance);
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
    const uint32_t blocks = std::min<uint32_t>(blocks_for(n, 64), 4096);
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(kThreads), 0, st, fwd.as<uint2>(), master.as<float>(),
                       d_raw, n, b, d_out);
    PT_HIP(hipGetLastError());
}

void RadianceCache::prepare_train(uint32_t n) {
    const size_t ld = (static_cast<size_t>(n) + 15) & ~size_t(15);
    acts.reserve(ld * (kIn + 4 * kHid) * 2);
    deltas.reserve(ld * (5 * kHid) * 2);
    part.reserve(static_cast<size_t>(blocks_for(n, kSliceLen)) * kParams * sizeof(float));
    loss.reserve(ld * sizeof(float));
}

void RadianceCache::launch_train(const float* d_raw, const float* d_target, uint32_t n, float* d_loss,
                                 float* d_grad_out, hipStream_t st) {
    const size_t ld = (static_cast<size_t>(n) + 15) & ~size_t(15);
    const uint32_t nslices = blocks_for(n, kSliceLen);
    SceneBox b;
    for (int k = 0; k < 3; ++k) {
        b.mn[k] = box_min[k];
        b.inv_ext[k] = box_inv[k];
    }
    hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(std::min<uint32_t>(blocks_for(n, 64), 2048)), dim3(kThreads), 0, st,
                       fwd.as<uint2>(), bwd.as<uint2>(), master.as<float>(), d_raw, d_target, n, ld, b,
                       acts.as<_Float16>(), deltas.as<_Float16>(), d_loss ? d_loss : loss.as<float>());
    constexpr int kTasks = fwd_base(kLayers) + 4 * 4 + 1;
    hipLaunchKernelGGL(k_mlp_wgrad, dim3(blocks_for(static_cast<uint64_t>(kTasks) * nslices, kThreads / 64)),
                       dim3(kThreads), 0, st, acts.as<_Float16>(), deltas.as<_Float16>(), n, ld, kSliceLen,
                       nslices, part.as<float>());
    hipLaunchKernelGGL(k_mlp_adam, dim3(blocks_for(kParams, kThreads)), dim3(kThreads), 0, st, master.as<float>(),
                       m.as<float>(), v.as<float>(), part.as<float>(), nslices, 1.0f / static_cast<float>(n), lr,
                       beta1, beta2, eps, dstep.as<uint32_t>(), d_grad_out);
    if (!d_grad_out) pack(st, 1);
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
