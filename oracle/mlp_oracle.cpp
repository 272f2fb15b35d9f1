// CPU FP64 oracle of the radiance-cache MLP (test infrastructure only).
// The reference has no learned component; this is the build's own network,
// restated in double precision so the HIP kernels (fp16 MFMA operands, fp32
// accumulation) can be checked against it and finite differences can check
// this backward pass.
//
// Network: 96 inputs -> 4 x (64, ReLU) -> 3 outputs with exp.
// Inputs per sample from 12 raw floats (position, normal, view direction,
// albedo): position mapped into [0,1]^3 of the scene box, then
//   enc[a*24 + 2k]     = sin(2^k * pi * p_a)
//   enc[a*24 + 2k + 1] = cos(2^k * pi * p_a),   a = 0..2, k = 0..11
// followed by normal (3), view (3), albedo (3) and 15 zeros.
// Parameters (flat fp32): W0[96][64] b0[64] W1[64][64] b1 W2 b2 W3 b3
// W4[64][3] b4[3]; W[in][out].
// Loss (relative L2): (1/N) sum_s sum_c (y - t)^2 / (sg(y)^2 + 0.01), where
// sg() is not differentiated.
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

constexpr int kIn = 96, kHid = 64, kOut = 3, kLayers = 5;
constexpr int kDims[kLayers + 1] = {kIn, kHid, kHid, kHid, kHid, kOut};
constexpr double kRelEps = 0.01;

size_t w_off(int l) {
    size_t o = 0;
    for (int k = 0; k < l; ++k) o += static_cast<size_t>(kDims[k]) * kDims[k + 1] + kDims[k + 1];
    return o;
}
size_t b_off(int l) { return w_off(l) + static_cast<size_t>(kDims[l]) * kDims[l + 1]; }

// The position is mapped in fp32 as the kernels do, p = (raw - bmin) *
// (1 / (bmax - bmin)); the encoding is then evaluated exactly from that p.
void encode(const float* raw, const float* bmin, const float* bmax, double* x) {
    const double pi = 3.14159265358979323846;
    for (int a = 0; a < 3; ++a) {
        const float inv = 1.0f / (bmax[a] - bmin[a]);
        const double p = static_cast<double>((raw[a] - bmin[a]) * inv);
        for (int k = 0; k < 12; ++k) {
            const double arg = std::ldexp(pi * p, k);
            x[a * 24 + 2 * k] = std::sin(arg);
            x[a * 24 + 2 * k + 1] = std::cos(arg);
        }
    }
    for (int k = 0; k < 9; ++k) x[72 + k] = raw[3 + k];
    for (int k = 81; k < kIn; ++k) x[k] = 0.0;
}

// Forward of one sample; acts[l] = layer-l input activations (l = 0..4),
// z = output pre-activation.
void forward(const float* P, const double* x, std::vector<std::vector<double>>& acts, double* z) {
    acts[0].assign(x, x + kIn);
    for (int l = 0; l < kLayers; ++l) {
        const int ni = kDims[l], no = kDims[l + 1];
        const float* W = P + w_off(l);
        const float* b = P + b_off(l);
        std::vector<double> y(no);
        for (int o = 0; o < no; ++o) {
            double s = b[o];
            for (int i = 0; i < ni; ++i) s += acts[l][i] * static_cast<double>(W[static_cast<size_t>(i) * no + o]);
            y[o] = s;
        }
        if (l + 1 < kLayers) {
            for (double& v : y) v = v > 0.0 ? v : 0.0;
            acts[l + 1] = y;
        } else {
            for (int o = 0; o < no; ++o) z[o] = y[o];
        }
    }
}

}  // namespace

extern "C" {

int or_mlp_param_count() { return static_cast<int>(w_off(kLayers)); }

void or_mlp_encode(const float* raw, const float* bmin, const float* bmax, double* x) {
    encode(raw, bmin, bmax, x);
}

void or_mlp_forward(const float* P, const float* raw, uint32_t n, const float* bmin,
                    const float* bmax, double* out) {
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < static_cast<int64_t>(n); ++s) {
        std::vector<std::vector<double>> acts(kLayers);
        double x[kIn], z[kOut];
        encode(raw + 12 * s, bmin, bmax, x);
        forward(P, x, acts, z);
        for (int c = 0; c < kOut; ++c) out[3 * s + c] = std::exp(z[c]);
    }
}

// Loss and its gradient with respect to every parameter (same flat layout).
double or_mlp_loss_grad(const float* P, const float* raw, const float* target, uint32_t n,
                        const float* bmin, const float* bmax, double* grad) {
    const size_t np = w_off(kLayers);
    for (size_t k = 0; k < np; ++k) grad[k] = 0.0;
    double loss = 0.0;
    for (uint32_t s = 0; s < n; ++s) {
        std::vector<std::vector<double>> acts(kLayers);
        double x[kIn], z[kOut];
        encode(raw + 12 * static_cast<size_t>(s), bmin, bmax, x);
        forward(P, x, acts, z);
        // delta of the output pre-activation: d/dz of (e^z - t)^2 / (sg(e^z)^2 + eps)
        std::vector<double> delta(kOut);
        for (int c = 0; c < kOut; ++c) {
            const double y = std::exp(z[c]);
            const double t = target[3 * static_cast<size_t>(s) + c];
            const double den = y * y + kRelEps;
            loss += (y - t) * (y - t) / den;
            delta[c] = 2.0 * (y - t) / den * y / n;
        }
        for (int l = kLayers - 1; l >= 0; --l) {
            const int ni = kDims[l], no = kDims[l + 1];
            const float* W = P + w_off(l);
            double* gW = grad + w_off(l);
            double* gb = grad + b_off(l);
            for (int o = 0; o < no; ++o) gb[o] += delta[o];
            for (int i = 0; i < ni; ++i)
                for (int o = 0; o < no; ++o) gW[static_cast<size_t>(i) * no + o] += acts[l][i] * delta[o];
            if (l == 0) break;
            std::vector<double> prev(ni, 0.0);
            for (int i = 0; i < ni; ++i) {
                if (!(acts[l][i] > 0.0)) continue;  // ReLU'(z) with z > 0  <=>  a > 0
                double s2 = 0.0;
                for (int o = 0; o < no; ++o) s2 += static_cast<double>(W[static_cast<size_t>(i) * no + o]) * delta[o];
                prev[i] = s2;
            }
            delta.swap(prev);
        }
    }
    return loss / n;
}

// One Adam step (Kingma & Ba 2015) on fp32 master weights, with the bias
// correction of step t (1-based).
void or_adam_step(float* w, float* m, float* v, const double* g, uint32_t count, uint32_t t,
                  float lr, float b1, float b2, float eps) {
    const double c1 = 1.0 - std::pow(static_cast<double>(b1), t);
    const double c2 = 1.0 - std::pow(static_cast<double>(b2), t);
    for (uint32_t k = 0; k < count; ++k) {
        const double mk = b1 * static_cast<double>(m[k]) + (1.0 - b1) * g[k];
        const double vk = b2 * static_cast<double>(v[k]) + (1.0 - b2) * g[k] * g[k];
        m[k] = static_cast<float>(mk);
        v[k] = static_cast<float>(vk);
        w[k] = static_cast<float>(w[k] - lr * (mk / c1) / (std::sqrt(vk / c2) + eps));
    }
}

}  // extern "C"
