// Neural shadow-visibility MLP on gfx950: BF16 MFMA forward, backward and fused inference with
// FP32 accumulation and FP32 master weights (Adam). No reference counterpart: the reference
// has no shadows; this is the learned form of the shadow ray of the hair path.
//
// Network: (p, d) -> p normalized by the scene box -> 6 dims x 8 frequencies x {sin, cos}
// (pi 2^k x) = 96 features -> 4 x (64, ReLU) -> logit -> sigmoid = visibility. Parameters,
// FP32, flat: W1[64][96] b1 W2[64][64] b2 W3 b3 W4 b4 w5[64] b5 (18,753 values).
//
// MFMA mapping (gfx950 v_mfma_f32_16x16x32_bf16, K = 32): a layer computes D = W X^T for 16
// outputs x 16 samples. Lane l = 16 g + c holds D[4 g + j][c] (j = 0..3), i.e. outputs
// 4g..4g+3 of sample c. The B operand of a K = 32 step holds 8 K rows per lane (8g..8g+7 of
// column c); a step covers two 16-row input tiles t0, t1, and physical row 8g + i stands for
// input 16 t0 + 4g + i (i < 4) or 16 t1 + 4g + i - 4 (i >= 4). With that order of K, the
// outputs one lane holds of two output tiles are exactly its B operand of the next layer: the
// forward pass and the back-propagated deltas (dX = W^T dZ) chain from layer to layer in
// registers, and the weight fragments are packed in the same K order. Weight gradients
// (dW = dZ X^T, K = samples) take samples in plain order and are staged through LDS.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mlp.hpp"

namespace vk {
namespace mlp_detail {
namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef short s8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
typedef float f2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef short s2 __attribute__((ext_vector_type(2)));

constexpr int kEnc = 96, kHid = 64;
constexpr int kOffW1 = 0, kOffB1 = 6144, kOffW2 = 6208, kOffB2 = 10304, kOffW3 = 10368, kOffB3 = 14464;
constexpr int kOffW4 = 14528, kOffB4 = 18624, kOffW5 = 18688, kOffB5 = 18752;
// packed bf16 fragments: [frag][64 lanes][8]; forward W1 (4 output tiles x 3 K steps), W2..W4
// (4 x 2 each), then the transposed W2..W4 (4 input tiles x 2 K steps each) for back-propagation
constexpr int kFragW1 = 0, kFragW2 = 12, kFragW3 = 20, kFragW4 = 28, kFragT2 = 36, kFragT3 = 44, kFragT4 = 52;
constexpr int kFragsFwd = 36, kFragsAll = 60;
constexpr int kSmall = 4 * kHid + kHid + 1;  // biases b1..b4, w5, b5 (FP32)

__device__ __forceinline__ uint32_t bf16_bits(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ __forceinline__ float bf16_val(float x) { return __uint_as_float(bf16_bits(x) << 16); }
// Two floats to bf16, round to nearest even, a in the low half (gfx950's v_cvt_pk_bf16_f32:
// the same bits as bf16_bits for finite values).
__device__ __forceinline__ s2 pack2(float a, float b) {
    return __builtin_bit_cast(s2, __builtin_convertvector(f2{a, b}, bf2));
}
__device__ __forceinline__ s4 pack4(float a, float b, float c, float d) {
    return __builtin_bit_cast(s4, __builtin_convertvector(f4{a, b, c, d}, bf4));
}
// ReLU of four accumulators as bf16: max(x, 0) then rounding equals rounding then max, and a
// negative bf16 (sign bit set) is a negative int16, so the packed integer max clears it.
__device__ __forceinline__ s4 relu_pack4(f4 a) {
    const s2 lo = __builtin_elementwise_max(pack2(a[0], a[1]), s2{0, 0});
    const s2 hi = __builtin_elementwise_max(pack2(a[2], a[3]), s2{0, 0});
    return __builtin_shufflevector(lo, hi, 0, 1, 2, 3);
}
// max(x, 0) of a finite FP32 value as one integer max: the bit patterns of non-negative floats
// are non-negative integers in the same order, those of negative floats (and -0) negative.
__device__ __forceinline__ float relu_bits(float x) { return __int_as_float(max(__float_as_int(x), 0)); }
__device__ __forceinline__ s8 cat(s4 a, s4 b) { return __builtin_shufflevector(a, b, 0, 1, 2, 3, 4, 5, 6, 7); }
__device__ __forceinline__ f4 mfma(s8 a, s8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a), __builtin_bit_cast(bf8, b), c, 0, 0, 0);
}

// Source index in the flat parameters of packed element (frag, lane, i): lane 16 g + c, K
// step p over the 16-row tiles 2p, 2p + 1, physical K row 8 g + i = logical row k below.
__device__ __forceinline__ int packed_source(int frag, int lane, int i) {
    const int g = lane >> 4, c = lane & 15;
    auto krow = [&](int p) { return i < 4 ? 32 * p + 4 * g + i : 32 * p + 16 + 4 * g + (i - 4); };
    if (frag < kFragW2) {  // W1 [m][p]: row = output 16 m + c, column = input feature k
        const int m = frag / 3, p = frag % 3;
        return kOffW1 + (16 * m + c) * kEnc + krow(p);
    }
    if (frag < kFragT2) {  // W2..W4 [m][p]
        const int L = (frag - kFragW2) / 8, q = (frag - kFragW2) % 8, m = q / 2, p = q % 2;
        const int off = L == 0 ? kOffW2 : (L == 1 ? kOffW3 : kOffW4);
        return off + (16 * m + c) * kHid + krow(p);
    }
    // transposed W2..W4 [t][p]: row = input 16 t + c, K = output k
    const int L = (frag - kFragT2) / 8, q = (frag - kFragT2) % 8, t = q / 2, p = q % 2;
    const int off = L == 0 ? kOffW2 : (L == 1 ? kOffW3 : kOffW4);
    return off + krow(p) * kHid + 16 * t + c;
}

__global__ void k_pack(const float* __restrict__ master, uint16_t* __restrict__ packed) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;  // element
    if (e >= kFragsAll * 512) return;
    const int frag = e >> 9, lane = (e >> 3) & 63, i = e & 7;
    packed[e] = uint16_t(bf16_bits(master[packed_source(frag, lane, i)]));
}

// sin(pi v) and cos(pi v) for |v| < 2^22: v = f + n / 2 with n = rint(2 v) and |f| <= 1/4,
// both exact; odd and even polynomials in f (coefficients of a least-squares minimax fit on
// [0, 1/4] rounded to FP32: absolute error below 9e-8 for sin and 6e-8 for cos, far below the
// bf16 rounding of the features), then the quadrant n mod 4. The oracle evaluates the same
// operations, so the features agree bit for bit.
constexpr float kSinPi1 = 3.1415925f, kSinPi3 = -5.1677036f, kSinPi5 = 2.549629f, kSinPi7 = -0.587813f;
constexpr float kCosPi2 = -4.934802f, kCosPi4 = 4.0587077f, kCosPi6 = -1.3350585f, kCosPi8 = 0.23142953f;
__device__ __forceinline__ void sincospi(float v, float& s, float& c) {
    const float n = rintf(2.0f * v);
    const float f = fmaf(-0.5f, n, v);
    const float u = f * f;
    float ps = fmaf(u, kSinPi7, kSinPi5);
    ps = fmaf(u, ps, kSinPi3);
    ps = fmaf(u, ps, kSinPi1);
    const float sf = f * ps;
    float pc = fmaf(u, kCosPi8, kCosPi6);
    pc = fmaf(u, pc, kCosPi4);
    pc = fmaf(u, pc, kCosPi2);
    const float cf = fmaf(u, pc, 1.0f);
    const int q = int(n) & 3;  // sin(pi f + q pi / 2), cos(pi f + q pi / 2): swap, then sign flips
    const float a = (q & 1) ? cf : sf, b = (q & 1) ? sf : cf;
    s = __uint_as_float(__float_as_uint(a) ^ (uint32_t(q & 2) << 30));
    c = __uint_as_float(__float_as_uint(b) ^ (uint32_t((q + 1) & 2) << 30));
}

// Feature f = 16 t + 4 g + j of the lane's sample for input dim t: k = 2 g + j / 2,
// sin for even j, cos for odd j (feature (dim * 8 + k) * 2 + s). The odd frequency comes from
// the even one by the double-angle formulas (a few ulp in FP32), which halves the polynomial
// evaluations.
__device__ __forceinline__ void encode(const float x[6], int g, s4 enc[6], int t0 = 0, int t1 = 6) {
#pragma unroll
    for (int t = t0; t < t1; ++t) {
        float s0, c0;
        sincospi(ldexpf(x[t], 2 * g), s0, c0);
        const float s1 = 2.0f * s0 * c0, c1 = (c0 - s0) * (c0 + s0);
        enc[t] = pack4(s0, c0, s1, c1);
    }
}

__device__ __forceinline__ void load_sample(const float4* s, float4 c, float4 ih, float x[6]) {
    const float4 a = s[0], b = s[1];
    x[0] = (a.x - c.x) * ih.x;
    x[1] = (a.y - c.y) * ih.y;
    x[2] = (a.z - c.z) * ih.z;
    x[3] = b.x;
    x[4] = b.y;
    x[5] = b.z;
}

// Forward of 16 samples (lane l: sample l & 15); h[L][m] = post-ReLU FP32 outputs 16 m + 4 g + j.
__device__ __forceinline__ float forward16(const s8* __restrict__ W, const float* __restrict__ small, const s4 enc[6],
                                           int lane, f4 h[4][4]) {
    const int g = lane >> 4;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int p = 0; p < 3; ++p) acc = mfma(W[(kFragW1 + m * 3 + p) * 64 + lane], cat(enc[2 * p], enc[2 * p + 1]), acc);
#pragma unroll
        for (int j = 0; j < 4; ++j) h[0][m][j] = fmaxf(acc[j] + small[16 * m + 4 * g + j], 0.0f);
    }
#pragma unroll
    for (int L = 1; L < 4; ++L) {
        s4 xin[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) xin[t] = pack4(h[L - 1][t][0], h[L - 1][t][1], h[L - 1][t][2], h[L - 1][t][3]);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int p = 0; p < 2; ++p)
                acc = mfma(W[(kFragW2 + (L - 1) * 8 + m * 2 + p) * 64 + lane], cat(xin[2 * p], xin[2 * p + 1]), acc);
#pragma unroll
            for (int j = 0; j < 4; ++j) h[L][m][j] = fmaxf(acc[j] + small[kHid * L + 16 * m + 4 * g + j], 0.0f);
        }
    }
    float z = 0.0f;
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int j = 0; j < 4; ++j) z += small[4 * kHid + 16 * m + 4 * g + j] * h[3][m][j];
    z += __shfl_xor(z, 16);
    z += __shfl_xor(z, 32);
    return z + small[4 * kHid + kHid];
}

// Inference forward of 32 samples, two 16-sample tiles u (lane l: samples l & 15 of each
// tile): every weight fragment read from LDS serves both tiles. The bias is the accumulator's
// start value, hidden outputs go to bf16 through relu_pack4, and the output layer runs in FP32.
// z[u] is the logit of the lane's sample of tile u (in all four lane groups).
__device__ __forceinline__ void forward32(const s8* __restrict__ W, const float* __restrict__ small,
                                          const s4 enc[2][6], int lane, float z[2]) {
    const int g = lane >> 4;
    s4 xin[2][4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const f4 b = *reinterpret_cast<const f4*>(small + 16 * m + 4 * g);
        f4 a0 = b, a1 = b;
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            const s8 w = W[(kFragW1 + m * 3 + p) * 64 + lane];
            a0 = mfma(w, cat(enc[0][2 * p], enc[0][2 * p + 1]), a0);
            a1 = mfma(w, cat(enc[1][2 * p], enc[1][2 * p + 1]), a1);
        }
        xin[0][m] = relu_pack4(a0);
        xin[1][m] = relu_pack4(a1);
    }
#pragma unroll
    for (int L = 1; L < 3; ++L) {
        s4 nx[2][4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const f4 b = *reinterpret_cast<const f4*>(small + kHid * L + 16 * m + 4 * g);
            f4 a0 = b, a1 = b;
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const s8 w = W[(kFragW2 + (L - 1) * 8 + m * 2 + p) * 64 + lane];
                a0 = mfma(w, cat(xin[0][2 * p], xin[0][2 * p + 1]), a0);
                a1 = mfma(w, cat(xin[1][2 * p], xin[1][2 * p + 1]), a1);
            }
            nx[0][m] = relu_pack4(a0);
            nx[1][m] = relu_pack4(a1);
        }
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int m = 0; m < 4; ++m) xin[u][m] = nx[u][m];
    }
    float zz[2] = {0.0f, 0.0f};
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const f4 b = *reinterpret_cast<const f4*>(small + 3 * kHid + 16 * m + 4 * g);
        const f4 w5 = *reinterpret_cast<const f4*>(small + 4 * kHid + 16 * m + 4 * g);
        f4 a0 = b, a1 = b;
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const s8 w = W[(kFragW4 + m * 2 + p) * 64 + lane];
            a0 = mfma(w, cat(xin[0][2 * p], xin[0][2 * p + 1]), a0);
            a1 = mfma(w, cat(xin[1][2 * p], xin[1][2 * p + 1]), a1);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            zz[0] = fmaf(w5[j], relu_bits(a0[j]), zz[0]);
            zz[1] = fmaf(w5[j], relu_bits(a1[j]), zz[1]);
        }
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        zz[u] += __shfl_xor(zz[u], 16);
        zz[u] += __shfl_xor(zz[u], 32);
        z[u] = zz[u] + small[4 * kHid + kHid];
    }
}

// Loads sample s of the inference input (zeros past n) and encodes it for lane group g.
__device__ __forceinline__ void encode_sample(const float4* __restrict__ samples, uint32_t s, uint32_t n, float4 c,
                                              float4 ih, int g, s4 enc[6]) {
    float x[6] = {0, 0, 0, 0, 0, 0};
    if (s < n) load_sample(samples + size_t(s) * 2, c, ih, x);
    encode(x, g, enc);
}

__device__ __forceinline__ float sigmoid(float z) { return 1.0f / (1.0f + expf(-z)); }
// Inference form: the hardware exp2 and reciprocal (a few ulp, far below the network's error).
__device__ __forceinline__ float sigmoid_fast(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

// Stages the forward fragments and the FP32 small parameters into LDS (all threads).
__device__ __forceinline__ void stage_weights(const uint16_t* __restrict__ packed, const float* __restrict__ master,
                                              s8* lw, float* lsmall, int nfrags) {
    const s8* src = reinterpret_cast<const s8*>(packed);
    for (int i = threadIdx.x; i < nfrags * 64; i += blockDim.x) lw[i] = src[i];
    for (int i = threadIdx.x; i < kSmall; i += blockDim.x) {
        int s;
        if (i < 4 * kHid) {
            const int L = i / kHid, o = i % kHid;
            s = (L == 0 ? kOffB1 : L == 1 ? kOffB2 : L == 2 ? kOffB3 : kOffB4) + o;
        } else {
            s = kOffW5 + (i - 4 * kHid);
        }
        lsmall[i] = master[s];
    }
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_infer(MlpDev m, const float4* __restrict__ samples, uint32_t n,
                                               float* __restrict__ out) {
    __shared__ s8 lw[kFragsFwd * 64];
    __shared__ __align__(16) float lsmall[kSmall];
    stage_weights(m.packed, m.master, lw, lsmall, kFragsFwd);
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nwaves = (gridDim.x * blockDim.x) >> 6;
    const float4 c = make_float4(m.center[0], m.center[1], m.center[2], 0.0f);
    const float4 ih = make_float4(m.inv_half[0], m.inv_half[1], m.inv_half[2], 0.0f);
    for (uint32_t base = wave * 32; base < n; base += nwaves * 32) {
        asm volatile("" ::: "memory");  // weight fragments are re-read from LDS in every pass
        s4 enc[2][6];
        encode_sample(samples, base + uint32_t(lane & 15), n, c, ih, g, enc[0]);
        encode_sample(samples, base + 16u + uint32_t(lane & 15), n, c, ih, g, enc[1]);
        float z[2];
        forward32(lw, lsmall, enc, lane, z);
        // lane groups 0 and 1 write tiles 0 and 1
        const uint32_t s = base + uint32_t(16 * g + (lane & 15));
        if (g < 2 && s < n) out[s] = sigmoid_fast(g == 0 ? z[0] : z[1]);
    }
}

// Render form: for each queued shadow pixel, the MLP's visibility p at (hit point, light
// direction). Confident predictions (p (1 - p) <= threshold) are written as the pixel's
// visibility; the rest are appended to the exact-fallback queue.
//
// A lane's inputs of one 32-query pass: the origins and directions of its sample column in
// the two tiles (paired: one origin, the two light samples' directions) and the pixel id of
// the query it owns. They are loaded one pass ahead, so their memory latency overlaps the
// previous pass's network.
struct GateIn {
    float4 o[2], d[2];
    uint32_t pix;
};

__device__ __forceinline__ void gate_load(const MlpShadowParams& sp, bool paired, uint32_t base, uint32_t n, int lane,
                                          GateIn& in) {
    const int g = lane >> 4;
    const uint32_t c = uint32_t(lane & 15);
    const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    in.o[0] = in.o[1] = in.d[0] = in.d[1] = z;
    if (paired) {
        const uint32_t q0 = base + 2u * c;
        if (q0 < n) {
            in.o[0] = in.o[1] = sp.origins[q0];
            in.d[0] = sp.dirs[q0];
        }
        if (q0 + 1u < n) in.d[1] = sp.dirs[q0 + 1u];
    } else {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const uint32_t q = base + 16u * uint32_t(u) + c;
            if (q < n) {
                in.o[u] = sp.origins[q];
                if (sp.dirs)
                    in.d[u] = sp.dirs[q];
                else
                    in.d[u] = make_float4(sp.dir[0], sp.dir[1], sp.dir[2], 0.0f);
            }
        }
    }
    // lane groups 0 and 1 own the queries of tiles 0 and 1
    const uint32_t q = paired ? base + 2u * c + uint32_t(g) : base + 16u * uint32_t(g) + c;
    in.pix = 0u;
    if (g < 2 && q < n) in.pix = sp.queue[q];
}

__device__ __forceinline__ void gate_features(const float4& o, const float4& d, float4 c, float4 ih, float x[6]) {
    x[0] = (o.x - c.x) * ih.x;
    x[1] = (o.y - c.y) * ih.y;
    x[2] = (o.z - c.z) * ih.z;
    x[3] = d.x;
    x[4] = d.y;
    x[5] = d.z;
}

// One 1024-thread block per CU (16 wavefronts, four per SIMD at the kernel's registers): one
// copy of the weight fragments in LDS per CU instead of four, a quarter of the staging reads.
constexpr int kGateWaves = 16;

// kPaired: queue entries 2k, 2k + 1 are the two light samples of one hit (area light, two
// samples per hit), so tile u of a pass holds light sample u of hits base / 2 + c.
template <bool kPaired>
__global__ __launch_bounds__(kGateWaves * 64) void k_shadow_gate(MlpDev m, MlpShadowParams sp) {
    __shared__ s8 lw[kFragsFwd * 64];
    __shared__ __align__(16) float lsmall[kSmall];
    __shared__ uint32_t s_fb[kGateWaves][128];  // per-wave fallback pixels, flushed 64+ at a time
    __shared__ uint32_t s_fbn[kGateWaves], s_fbbase;
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const uint32_t n = *sp.count;
    constexpr bool paired = kPaired;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nwaves = (gridDim.x * blockDim.x) >> 6;
    GateIn cur;
    if (wave * 32u < n) gate_load(sp, paired, wave * 32u, n, lane, cur);
    stage_weights(m.packed, m.master, lw, lsmall, kFragsFwd);
    uint32_t* fb = s_fb[threadIdx.x >> 6];
    uint32_t fbn = 0;  // wave-uniform
    const float4 c = make_float4(m.center[0], m.center[1], m.center[2], 0.0f);
    const float4 ih = make_float4(m.inv_half[0], m.inv_half[1], m.inv_half[2], 0.0f);
    for (uint32_t base = wave * 32; base < n; base += nwaves * 32) {
        // a compiler-only fence: the weight fragments are read from LDS in every pass rather
        // than hoisted out of the loop into ~150 registers, which would leave one wave per SIMD
        asm volatile("" ::: "memory");
        const GateIn in = cur;
        const uint32_t next = base + nwaves * 32;
        if (next < n) gate_load(sp, paired, next, n, lane, cur);
        s4 enc[2][6];
        float x[6];
        gate_features(in.o[0], in.d[0], c, ih, x);
        encode(x, g, enc[0]);
        if (paired) {
            // tile u holds light sample u of hits base / 2 + c: one origin, its features once
#pragma unroll
            for (int t = 0; t < 3; ++t) enc[1][t] = enc[0][t];
            x[3] = in.d[1].x;
            x[4] = in.d[1].y;
            x[5] = in.d[1].z;
            encode(x, g, enc[1], 3, 6);
        } else {
            gate_features(in.o[1], in.d[1], c, ih, x);
            encode(x, g, enc[1]);
        }
        float z[2];
        forward32(lw, lsmall, enc, lane, z);
        const uint32_t q = paired ? base + 2u * uint32_t(lane & 15) + uint32_t(g) : base + uint32_t(16 * g + (lane & 15));
        const bool own = g < 2 && q < n;
        const float p = sigmoid_fast(g == 0 ? z[0] : z[1]);
        const uint32_t pix = in.pix;
        const bool fallback = own && p * (1.0f - p) > sp.threshold;
        if (own && !fallback) sp.vis[pix] = p;
        const uint64_t fm = __ballot(fallback);
        if (fallback) fb[fbn + __popcll(fm & ((1ull << lane) - 1ull))] = pix;
        fbn += uint32_t(__popcll(fm));  // <= 32 per pass: at most 63 + 32 < 128 before a flush
        if (fbn >= 64u) {
            uint32_t b = 0;
            if (lane == 0) b = atomicAdd(sp.fallback_count, fbn);
            b = __shfl(b, 0);
            for (uint32_t i = uint32_t(lane); i < fbn; i += 64u) sp.fallback_queue[b + i] = fb[i];
            fbn = 0;
        }
    }
    // The waves' last partial buffers go out with one atomic per block: all waves end at about
    // the same time, and one returning atomic per wave on the single queue counter serializes
    // (about 4k of them took ~16 us with a few thousand fallback queries).
    const uint32_t w = threadIdx.x >> 6;
    if (lane == 0) s_fbn[w] = fbn;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (int k = 0; k < kGateWaves; ++k) total += s_fbn[k];
        s_fbbase = total ? atomicAdd(sp.fallback_count, total) : 0u;
    }
    __syncthreads();
    uint32_t b = s_fbbase;
    for (uint32_t k = 0; k < w; ++k) b += s_fbn[k];
    for (uint32_t i = uint32_t(lane); i < fbn; i += 64u) sp.fallback_queue[b + i] = fb[i];
}

// Forward + backward of the batch; each block accumulates its share of the gradient and
// writes it to partial[block][param]. Wave w of a 64-sample chunk runs 16 samples through the
// network and back, staging bf16 inputs X_L and deltas dZ_L per layer in LDS; then wave w
// accumulates dW of layer w + 1 over the chunk's 64 samples with MFMAs (K = samples).
struct TrainLds {
    alignas(16) s8 w[kFragsAll * 64];       // 61,440 B
    alignas(16) float small[kSmall];
    alignas(16) uint16_t x1[kEnc * 64];     // [feature][sample]
    alignas(16) uint16_t x[3][kHid * 64];   // h1, h2, h3 as inputs of layers 2..4
    alignas(16) uint16_t dz[4][kHid * 64];  // [output][sample]
    alignas(16) float db[4 * kHid + kHid + 1];  // bias, w5 and b5 gradients
    float loss;
};

__global__ __launch_bounds__(256) void k_train(MlpDev m, MlpTrainParams tp) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    TrainLds& S = *reinterpret_cast<TrainLds*>(lds_raw);
    {
        const s8* src = reinterpret_cast<const s8*>(m.packed);
        for (int i = threadIdx.x; i < kFragsAll * 64; i += blockDim.x) S.w[i] = src[i];
        for (int i = threadIdx.x; i < kSmall; i += blockDim.x) {
            int s;
            if (i < 4 * kHid) {
                const int L = i / kHid, o = i % kHid;
                s = (L == 0 ? kOffB1 : L == 1 ? kOffB2 : L == 2 ? kOffB3 : kOffB4) + o;
            } else {
                s = kOffW5 + (i - 4 * kHid);
            }
            S.small[i] = m.master[s];
        }
        for (int i = threadIdx.x; i < 4 * kHid + kHid + 1; i += blockDim.x) S.db[i] = 0.0f;
        if (threadIdx.x == 0) S.loss = 0.0f;
        __syncthreads();
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, cs = lane & 15;
    const int col = 16 * w + cs;  // this lane's sample column within the chunk
    const float4 c = make_float4(m.center[0], m.center[1], m.center[2], 0.0f);
    const float4 ih = make_float4(m.inv_half[0], m.inv_half[1], m.inv_half[2], 0.0f);
    const float inv_b = 1.0f / float(tp.n);
    // dW accumulators of this wave's layer: D[o = 16 mo + 4 g + j][i = 16 ti + cs]
    f4 acc[24];
#pragma unroll
    for (int i = 0; i < 24; ++i) acc[i] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    const uint32_t nchunks = (tp.n + 63u) / 64u;
    for (uint32_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
        const uint32_t s = chunk * 64u + uint32_t(col);
        const bool valid = s < tp.n;
        float x[6] = {0, 0, 0, 0, 0, 0};
        float y = 0.0f;
        if (valid) {
            // epoch order: sample (first + s) of the permutation i -> (i mul + add) mod total
            const uint32_t src = uint32_t(((uint64_t(tp.first) + s) * tp.perm_mul + tp.perm_add) % tp.total);
            load_sample(tp.samples + size_t(src) * 2, c, ih, x);
            y = tp.targets[src];
        }
        s4 enc[6];
        encode(x, g, enc);
        f4 h[4][4];
        const float z = forward16(S.w, S.small, enc, lane, h);
        const float p = sigmoid(z);
        const float dz5 = valid ? (p - y) * inv_b : 0.0f;
        if (valid && g == 0) {
            const float lp = z >= 0.0f ? -log1pf(expf(-z)) : z - log1pf(expf(z));
            const float lq = z >= 0.0f ? -z - log1pf(expf(-z)) : -log1pf(expf(z));
            atomicAdd(&S.loss, -(y * lp + (1.0f - y) * lq));
        }
        // stage inputs: X_1 = features, X_2..4 = bf16(h1..h3)
#pragma unroll
        for (int t = 0; t < 6; ++t) {
            S.x1[(16 * t + 4 * g + 0) * 64 + col] = uint16_t(enc[t].x);
            S.x1[(16 * t + 4 * g + 1) * 64 + col] = uint16_t(enc[t].y);
            S.x1[(16 * t + 4 * g + 2) * 64 + col] = uint16_t(enc[t].z);
            S.x1[(16 * t + 4 * g + 3) * 64 + col] = uint16_t(enc[t].w);
        }
#pragma unroll
        for (int L = 0; L < 3; ++L)
#pragma unroll
            for (int mm = 0; mm < 4; ++mm)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    S.x[L][(16 * mm + 4 * g + j) * 64 + col] = uint16_t(bf16_bits(h[L][mm][j]));
        // output layer gradients and the delta entering layer 4
        f4 dh[4];
#pragma unroll
        for (int mm = 0; mm < 4; ++mm)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int k = 16 * mm + 4 * g + j;
                if (dz5 != 0.0f) atomicAdd(&S.db[4 * kHid + k], dz5 * h[3][mm][j]);
                dh[mm][j] = dz5 * S.small[4 * kHid + k];
            }
        if (g == 0 && dz5 != 0.0f) atomicAdd(&S.db[5 * kHid], dz5);
        // back-propagation through layers 4..1 (0-based L = 3..0)
#pragma unroll
        for (int L = 3; L >= 0; --L) {
            s4 dzb[4];
#pragma unroll
            for (int mm = 0; mm < 4; ++mm) {
                float d[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    d[j] = bf16_val(h[L][mm][j] > 0.0f ? dh[mm][j] : 0.0f);
                    const int o = 16 * mm + 4 * g + j;
                    S.dz[L][o * 64 + col] = uint16_t(bf16_bits(d[j]));
                    if (d[j] != 0.0f) atomicAdd(&S.db[kHid * L + o], d[j]);
                }
                dzb[mm] = pack4(d[0], d[1], d[2], d[3]);
            }
            if (L > 0) {
                const int tb = L == 1 ? kFragT2 : (L == 2 ? kFragT3 : kFragT4);
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    f4 a = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                    for (int p = 0; p < 2; ++p)
                        a = mfma(S.w[(tb + t * 2 + p) * 64 + lane], cat(dzb[2 * p], dzb[2 * p + 1]), a);
                    dh[t] = a;
                }
            }
        }
        __syncthreads();
        // dW of layer w + 1 over the chunk: A = dZ tile (rows = outputs), B = X tile (cols = inputs)
        const uint16_t* dz = S.dz[w];
        const uint16_t* xs = w == 0 ? S.x1 : S.x[w - 1];
        const int ntiles = w == 0 ? 6 : 4;
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
#pragma unroll
            for (int mo = 0; mo < 4; ++mo) {
                const s8 a = *reinterpret_cast<const s8*>(dz + (16 * mo + cs) * 64 + 32 * ks + 8 * g);
#pragma unroll
                for (int ti = 0; ti < 6; ++ti) {
                    if (ti < ntiles) {
                        const s8 b = *reinterpret_cast<const s8*>(xs + (16 * ti + cs) * 64 + 32 * ks + 8 * g);
                        acc[mo * 6 + ti] = mfma(a, b, acc[mo * 6 + ti]);
                    }
                }
            }
        }
        __syncthreads();
    }
    // write this block's partial gradient
    float* part = tp.partial + size_t(blockIdx.x) * kMlpParamCount;
    const int in_dim = w == 0 ? kEnc : kHid;
    const int offw = w == 0 ? kOffW1 : (w == 1 ? kOffW2 : (w == 2 ? kOffW3 : kOffW4));
    const int ntiles = w == 0 ? 6 : 4;
#pragma unroll
    for (int mo = 0; mo < 4; ++mo)
#pragma unroll
        for (int ti = 0; ti < 6; ++ti)
            if (ti < ntiles)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    part[offw + (16 * mo + 4 * g + j) * in_dim + 16 * ti + cs] = acc[mo * 6 + ti][j];
    __syncthreads();
    for (int i = threadIdx.x; i < 4 * kHid + kHid + 1; i += blockDim.x) {
        int d;
        if (i < 4 * kHid) {
            const int L = i / kHid, o = i % kHid;
            d = (L == 0 ? kOffB1 : L == 1 ? kOffB2 : L == 2 ? kOffB3 : kOffB4) + o;
        } else if (i < 5 * kHid) {
            d = kOffW5 + (i - 4 * kHid);
        } else {
            d = kOffB5;
        }
        part[d] = S.db[i];
    }
    if (threadIdx.x == 0) atomicAdd(tp.loss, S.loss);
}

// Sums the block partials; either returns the gradient or applies one Adam step. A block
// covers 64 parameters with 16 wavefronts: wavefront r sums the partial rows b = r mod 16
// (coalesced 256-byte rows), and the 16 sums meet in LDS in a fixed order, so the result does
// not depend on timing.
constexpr int kReduceParams = 64, kReduceRows = 16;
__global__ __launch_bounds__(kReduceParams * kReduceRows) void k_reduce_adam(MlpDev m, const float* __restrict__ partial,
                                                                             uint32_t nblocks, float* grad_out,
                                                                             AdamParams ap) {
    __shared__ float s_sum[kReduceRows][kReduceParams];
    const int lane = threadIdx.x & (kReduceParams - 1), r = threadIdx.x / kReduceParams;
    const int p = blockIdx.x * kReduceParams + lane;
    float part = 0.0f;
    if (p < kMlpParamCount)
        for (uint32_t b = uint32_t(r); b < nblocks; b += kReduceRows) part += partial[size_t(b) * kMlpParamCount + p];
    s_sum[r][lane] = part;
    __syncthreads();
    if (r != 0 || p >= kMlpParamCount) return;
    float g = 0.0f;
    for (int k = 0; k < kReduceRows; ++k) g += s_sum[k][lane];
    if (grad_out) {
        grad_out[p] = g;
        return;
    }
    const float m1 = ap.beta1 * m.adam_m[p] + (1.0f - ap.beta1) * g;
    const float m2 = ap.beta2 * m.adam_v[p] + (1.0f - ap.beta2) * g * g;
    m.adam_m[p] = m1;
    m.adam_v[p] = m2;
    const float mh = m1 / ap.bias1, vh = m2 / ap.bias2;
    m.master[p] -= ap.lr * mh / (sqrtf(vh) + ap.eps);
}

// Training samples from traced shadow rays: (origin, 0, direction, 0), target 1 = unoccluded.
__global__ void k_collect(const float4* __restrict__ rays, const HitRec* __restrict__ hits, uint32_t n,
                          float4* __restrict__ samples, float* __restrict__ targets) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 a = rays[size_t(i) * 2], b = rays[size_t(i) * 2 + 1];
    samples[size_t(i) * 2] = make_float4(a.x, a.y, a.z, 0.0f);
    samples[size_t(i) * 2 + 1] = make_float4(b.x, b.y, b.z, 0.0f);
    targets[i] = hits[i].prim_id == 0xFFFFFFFFu ? 1.0f : 0.0f;
}

}  // namespace
}  // namespace mlp_detail

size_t mlp_train_lds_bytes() { return sizeof(mlp_detail::TrainLds); }

hipError_t launch_mlp_pack(const MlpDev& m, hipStream_t st) {
    hipLaunchKernelGGL(mlp_detail::k_pack, dim3((mlp_detail::kFragsAll * 512 + 255) / 256), dim3(256), 0, st, m.master,
                       m.packed);
    return hipGetLastError();
}

hipError_t launch_mlp_infer(const MlpDev& m, const float* samples, uint32_t n, float* out, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const uint32_t groups = (n + 31u) / 32u;
    const uint32_t blocks = std::min<uint32_t>((groups + 3u) / 4u, 2048u);
    hipLaunchKernelGGL(mlp_detail::k_infer, dim3(blocks), dim3(256), 0, st, m,
                       reinterpret_cast<const float4*>(samples), n, out);
    return hipGetLastError();
}

hipError_t launch_mlp_shadow_gate(const MlpDev& m, const MlpShadowParams& sp, uint32_t max_count, hipStream_t st) {
    if (max_count == 0) return hipSuccess;
    // one resident block per CU: the four waves of a SIMD overlap one another's encoding
    // (VALU) with their matrix products (MFMA); each wavefront walks many 32-sample groups
    constexpr uint32_t W = mlp_detail::kGateWaves;
    int dev = 0, cus = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const uint32_t groups = (max_count + 31u) / 32u;
    const uint32_t blocks = std::max(1u, std::min<uint32_t>((groups + W - 1u) / W, uint32_t(std::max(cus, 1))));
    if (sp.paired && sp.dirs)
        hipLaunchKernelGGL(mlp_detail::k_shadow_gate<true>, dim3(blocks), dim3(W * 64), 0, st, m, sp);
    else
        hipLaunchKernelGGL(mlp_detail::k_shadow_gate<false>, dim3(blocks), dim3(W * 64), 0, st, m, sp);
    return hipGetLastError();
}

hipError_t launch_mlp_train_step(const MlpDev& m, const MlpTrainParams& tp, uint32_t nblocks, float* grad_out,
                                 const AdamParams& ap, hipStream_t st) {
    hipLaunchKernelGGL(mlp_detail::k_train, dim3(nblocks), dim3(256), mlp_train_lds_bytes(), st, m, tp);
    hipLaunchKernelGGL(mlp_detail::k_reduce_adam,
                       dim3((kMlpParamCount + mlp_detail::kReduceParams - 1) / mlp_detail::kReduceParams),
                       dim3(mlp_detail::kReduceParams * mlp_detail::kReduceRows), 0, st, m, tp.partial, nblocks, grad_out,
                       ap);
    if (!grad_out) return launch_mlp_pack(m, st);
    return hipGetLastError();
}

hipError_t launch_mlp_collect(const float* rays, const HitRec* hits, uint32_t n, float* samples, float* targets,
                              hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(mlp_detail::k_collect, dim3((n + 255) / 256), dim3(256), 0, st,
                       reinterpret_cast<const float4*>(rays), hits, n, reinterpret_cast<float4*>(samples), targets);
    return hipGetLastError();
}

hipError_t mlp_enable_train_lds() {
    return hipFuncSetAttribute(reinterpret_cast<const void*>(mlp_detail::k_train),
                               hipFuncAttributeMaxDynamicSharedMemorySize, int(mlp_train_lds_bytes()));
}

}  // namespace vk
