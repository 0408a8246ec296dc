// Double-word FP32 arithmetic on gfx950: a value is hi + lo, two floats with |lo| <= ulp(hi)/2
// (about 48 significant bits), built from FP32 adds and multiplies and the correctly rounded
// fmaf only (no FP64).
//  - error-free transforms: 2Sum (Knuth), Fast2Sum (Dekker), 2Prod with an FMA;
//  - add: AccurateDWPlusDW, mul: DWTimesDW3 (Joldes, Muller, Popescu, "Tight and rigorous
//    error bounds for basic building blocks of double-word arithmetic", ACM TOMS 44(2), 2017,
//    Algorithms 6 and 12);
//  - div: q1 = xh / yh corrected once by (x - q1 y)_h / yh (the "sloppy" division of the
//    double-double library of Hida, Li and Bailey); sqrt: s = sqrt(xh) corrected by
//    (x - s^2) / (2 s) with s^2 formed exactly.
// With these, the templated Prhi iteration (intersect.hpp) runs in double-word precision.
// Everything lives in vk::dw and is found by argument-dependent lookup, so an unqualified
// fabs / sqrt on a float elsewhere in vk still resolves to the float function.
#pragma once
#include <hip/hip_runtime.h>

#include "device_math.hpp"

namespace vk {
namespace dw {

struct ff {
    float hi, lo;
    ff() = default;
    __host__ __device__ __forceinline__ ff(float h) : hi(h), lo(0.0f) {}  // NOLINT: converts like a float
    __host__ __device__ __forceinline__ ff(int v) : hi(float(v)), lo(0.0f) {}
    __host__ __device__ __forceinline__ ff(float h, float l) : hi(h), lo(l) {}
};

__host__ __device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
    s = a + b;
    const float a1 = s - b;
    const float b1 = s - a1;
    e = (a - a1) + (b - b1);
}
__host__ __device__ __forceinline__ void fast_two_sum(float a, float b, float& s, float& e) {
    s = a + b;
    e = b - (s - a);
}
__host__ __device__ __forceinline__ void two_prod(float a, float b, float& p, float& e) {
    p = a * b;
    e = fmaf(a, b, -p);
}

__host__ __device__ __forceinline__ ff operator+(ff x, ff y) {
    float sh, sl, th, tl, vh, vl, zh, zl;
    two_sum(x.hi, y.hi, sh, sl);
    two_sum(x.lo, y.lo, th, tl);
    const float c = sl + th;
    fast_two_sum(sh, c, vh, vl);
    const float w = tl + vl;
    fast_two_sum(vh, w, zh, zl);
    return ff(zh, zl);
}
__host__ __device__ __forceinline__ ff operator-(ff x) { return ff(-x.hi, -x.lo); }
__host__ __device__ __forceinline__ ff operator-(ff x, ff y) { return x + (-y); }
__host__ __device__ __forceinline__ ff operator*(ff x, ff y) {
    float ch, cl1, zh, zl;
    two_prod(x.hi, y.hi, ch, cl1);
    const float tl0 = x.lo * y.lo;
    const float tl1 = fmaf(x.hi, y.lo, tl0);
    const float cl2 = fmaf(x.lo, y.hi, tl1);
    const float cl3 = cl1 + cl2;
    fast_two_sum(ch, cl3, zh, zl);
    return ff(zh, zl);
}
__host__ __device__ __forceinline__ ff operator/(ff x, ff y) {
    const float q1 = x.hi / y.hi;
    const ff r = x - y * ff(q1);
    const float q2 = r.hi / y.hi;
    float zh, zl;
    fast_two_sum(q1, q2, zh, zl);
    return ff(zh, zl);
}
__host__ __device__ __forceinline__ ff sqrt(ff x) {
    if (!(x.hi > 0.0f)) return ff(sqrtf(x.hi));
    const float s = sqrtf(x.hi);
    float p, e;
    two_prod(s, s, p, e);
    const float r = ((x.hi - p) - e) + x.lo;
    const float corr = r / (2.0f * s);
    float zh, zl;
    fast_two_sum(s, corr, zh, zl);
    return ff(zh, zl);
}
__host__ __device__ __forceinline__ ff& operator+=(ff& x, ff y) { return x = x + y; }
__host__ __device__ __forceinline__ bool operator<(ff x, ff y) { return x.hi < y.hi || (x.hi == y.hi && x.lo < y.lo); }
__host__ __device__ __forceinline__ bool operator>(ff x, ff y) { return y < x; }
__host__ __device__ __forceinline__ bool operator<=(ff x, ff y) {
    return x.hi < y.hi || (x.hi == y.hi && x.lo <= y.lo);
}
__host__ __device__ __forceinline__ ff fabs(ff x) { return x.hi < 0.0f || (x.hi == 0.0f && x.lo < 0.0f) ? -x : x; }

struct ff3 {
    ff x, y, z;
};
__host__ __device__ __forceinline__ ff3 add(ff3 a, ff3 b) { return ff3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ __forceinline__ ff3 sub(ff3 a, ff3 b) { return ff3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ __forceinline__ ff3 scl(ff s, ff3 a) { return ff3{s * a.x, s * a.y, s * a.z}; }
__host__ __device__ __forceinline__ ff3 scr(ff3 a, ff s) { return ff3{a.x * s, a.y * s, a.z * s}; }
__host__ __device__ __forceinline__ ff dot3(ff3 a, ff3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ __forceinline__ ff3 cross3(ff3 a, ff3 b) {
    return ff3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__host__ __device__ __forceinline__ ff3 nrm3(ff3 a) { return scr(a, ff(1) / sqrt(dot3(a, a))); }
__host__ __device__ __forceinline__ ff3 toff(f3 a) { return ff3{ff(a.x), ff(a.y), ff(a.z)}; }

}  // namespace dw
using dw::ff;
using dw::ff3;
}  // namespace vk
