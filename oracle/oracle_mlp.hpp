// Oracle of the neural shadow-visibility MLP (test infrastructure; no reference counterpart:
// the network is an extension of the hair path, defined here and in the HIP kernels).
//
// Network: input (p, d) with p normalized to [-1, 1] by the scene box, 6 dims x 8 frequencies
// x {sin, cos}(2^k pi x) = 96 features -> 4 x (64, ReLU) -> 1 logit -> sigmoid = predicted
// visibility p; its Bernoulli variance p (1 - p) gates the exact fallback.
// Parameters (FP32, flat): W1[64][96] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64][64] b4[64]
// w5[64] b5.
// MlpRef<T, BF16>: BF16 = true mirrors the GPU kernels (features, weights, hidden activations
// and back-propagated deltas rounded to bf16 before each matrix product, FP32 accumulation,
// output layer in FP32); BF16 = false with T = double is the exact function that the
// finite-difference checks pin.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace orc {

constexpr int kMlpIn = 6, kMlpFreq = 8, kMlpEnc = 96, kMlpHid = 64;
constexpr int kMlpOffW1 = 0, kMlpOffB1 = kMlpOffW1 + kMlpHid * kMlpEnc;
constexpr int kMlpOffW2 = kMlpOffB1 + kMlpHid, kMlpOffB2 = kMlpOffW2 + kMlpHid * kMlpHid;
constexpr int kMlpOffW3 = kMlpOffB2 + kMlpHid, kMlpOffB3 = kMlpOffW3 + kMlpHid * kMlpHid;
constexpr int kMlpOffW4 = kMlpOffB3 + kMlpHid, kMlpOffB4 = kMlpOffW4 + kMlpHid * kMlpHid;
constexpr int kMlpOffW5 = kMlpOffB4 + kMlpHid, kMlpOffB5 = kMlpOffW5 + kMlpHid;
constexpr int kMlpParams = kMlpOffB5 + 1;  // 18753

// float -> bf16 -> float, round to nearest even (finite inputs).
inline float bf16_round(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

// sin(pi v) / cos(pi v) with exact argument reduction (v - 2 round(v / 2) is exact).
inline double sinpi_d(double v) { return std::sin(M_PI * (v - 2.0 * std::nearbyint(0.5 * v))); }
inline double cospi_d(double v) { return std::cos(M_PI * (v - 2.0 * std::nearbyint(0.5 * v))); }

// The kernels' FP32 evaluation of sin(pi v), cos(pi v) (|v| < 2^22), operation for operation:
// v = f + n / 2 with n = rint(2 v), |f| <= 1/4 (exact); odd / even polynomials in f whose
// coefficients are a least-squares minimax fit on [0, 1/4] rounded to FP32; quadrant n mod 4.
// Its accuracy against sin / cos in double is pinned by tests/test_mlp_oracle.py.
inline void sincospi_f32(float v, float& s, float& c) {
    const float n = std::nearbyint(2.0f * v);  // round half to even, as v_rndne_f32
    const float f = std::fma(-0.5f, n, v);
    const float u = f * f;
    float ps = std::fma(u, -0.587813f, 2.549629f);
    ps = std::fma(u, ps, -5.1677036f);
    ps = std::fma(u, ps, 3.1415925f);
    const float sf = f * ps;
    float pc = std::fma(u, 0.23142953f, -1.3350585f);
    pc = std::fma(u, pc, 4.0587077f);
    pc = std::fma(u, pc, -4.934802f);
    const float cf = std::fma(u, pc, 1.0f);
    const int q = int(n) & 3;
    const float a = (q & 1) ? cf : sf, b = (q & 1) ? sf : cf;
    s = (q & 2) ? -a : a;
    c = ((q + 1) & 2) ? -b : b;
}

template <class T, bool BF16>
struct MlpRef {
    static T rd(T x) {
        if constexpr (BF16)
            return T(bf16_round(float(x)));
        else
            return x;
    }

    // Feature f = (dim * 8 + k) * 2 + s: s = 0 sin, 1 cos of pi 2^k x_dim.
    static void encode(const float* sample, const float center[3], const float inv_half[3], T enc[kMlpEnc]) {
        T x[kMlpIn];
        for (int i = 0; i < 3; ++i) x[i] = T((sample[i] - center[i]) * inv_half[i]);
        for (int i = 0; i < 3; ++i) x[3 + i] = T(sample[4 + i]);
        for (int i = 0; i < kMlpIn; ++i)
            for (int k = 0; k < kMlpFreq; ++k) {
                T s, c;
                if constexpr (BF16) {
                    // the kernels' FP32 features: the polynomial at even k, the double-angle
                    // formulas from frequency k - 1 at odd k
                    float s0, c0;
                    sincospi_f32(std::ldexp(float(x[i]), k & ~1), s0, c0);
                    if (k & 1) {
                        const float s1 = 2.0f * s0 * c0, c1 = (c0 - s0) * (c0 + s0);
                        s0 = s1;
                        c0 = c1;
                    }
                    s = T(s0);
                    c = T(c0);
                } else {
                    const double v = std::ldexp(double(x[i]), k);
                    s = T(sinpi_d(v));
                    c = T(cospi_d(v));
                }
                enc[(i * kMlpFreq + k) * 2 + 0] = rd(s);
                enc[(i * kMlpFreq + k) * 2 + 1] = rd(c);
            }
    }

    struct Acts {
        T enc[kMlpEnc];
        T h[4][kMlpHid];  // post-ReLU (unrounded)
        T z5, p;
    };

    static void forward(const std::vector<T>& P, const float* sample, const float center[3], const float inv_half[3],
                        Acts& a) {
        encode(sample, center, inv_half, a.enc);
        const int in_dim[4] = {kMlpEnc, kMlpHid, kMlpHid, kMlpHid};
        const int offW[4] = {kMlpOffW1, kMlpOffW2, kMlpOffW3, kMlpOffW4};
        const int offB[4] = {kMlpOffB1, kMlpOffB2, kMlpOffB3, kMlpOffB4};
        for (int L = 0; L < 4; ++L) {
            T xin[kMlpEnc];
            for (int i = 0; i < in_dim[L]; ++i) xin[i] = L == 0 ? a.enc[i] : rd(a.h[L - 1][i]);
            for (int o = 0; o < kMlpHid; ++o) {
                T acc = T(0);
                for (int i = 0; i < in_dim[L]; ++i) acc += rd(P[offW[L] + o * in_dim[L] + i]) * xin[i];
                acc += P[offB[L] + o];
                a.h[L][o] = acc > T(0) ? acc : T(0);
            }
        }
        T z = T(0);
        for (int k = 0; k < kMlpHid; ++k) z += P[kMlpOffW5 + k] * a.h[3][k];
        z += P[kMlpOffB5];
        a.z5 = z;
        a.p = T(1) / (T(1) + std::exp(-z));
    }

    // Mean binary cross-entropy over n samples and its gradient (added into g).
    static double grad(const std::vector<T>& P, const float* samples, const float* targets, size_t n,
                       const float center[3], const float inv_half[3], std::vector<T>& g) {
        g.assign(kMlpParams, T(0));
        double loss = 0.0;
        const int in_dim[4] = {kMlpEnc, kMlpHid, kMlpHid, kMlpHid};
        const int offW[4] = {kMlpOffW1, kMlpOffW2, kMlpOffW3, kMlpOffW4};
        const int offB[4] = {kMlpOffB1, kMlpOffB2, kMlpOffB3, kMlpOffB4};
        const T inv_n = T(1) / T(n);
        for (size_t s = 0; s < n; ++s) {
            Acts a;
            forward(P, samples + s * 8, center, inv_half, a);
            const T y = T(targets[s]);
            // log-sigmoid forms of the BCE
            const double z = double(a.z5);
            const double lp = z >= 0 ? -std::log1p(std::exp(-z)) : z - std::log1p(std::exp(z));
            const double lq = z >= 0 ? -z - std::log1p(std::exp(-z)) : -std::log1p(std::exp(z));
            loss -= double(y) * lp + (1.0 - double(y)) * lq;
            const T dz5 = (a.p - y) * inv_n;
            for (int k = 0; k < kMlpHid; ++k) g[kMlpOffW5 + k] += dz5 * a.h[3][k];
            g[kMlpOffB5] += dz5;
            T dh[kMlpHid];
            for (int k = 0; k < kMlpHid; ++k) dh[k] = dz5 * P[kMlpOffW5 + k];
            for (int L = 3; L >= 0; --L) {
                T dz[kMlpHid];
                for (int o = 0; o < kMlpHid; ++o) dz[o] = rd(a.h[L][o] > T(0) ? dh[o] : T(0));
                const int ni = in_dim[L];
                for (int o = 0; o < kMlpHid; ++o) {
                    g[offB[L] + o] += dz[o];
                    for (int i = 0; i < ni; ++i) {
                        const T xi = L == 0 ? a.enc[i] : rd(a.h[L - 1][i]);
                        g[offW[L] + o * ni + i] += dz[o] * xi;
                    }
                }
                if (L > 0) {
                    for (int i = 0; i < kMlpHid; ++i) {
                        T acc = T(0);
                        for (int o = 0; o < kMlpHid; ++o) acc += rd(P[offW[L] + o * ni + i]) * dz[o];
                        dh[i] = acc;
                    }
                }
            }
        }
        return loss / double(n);
    }
};

}  // namespace orc
