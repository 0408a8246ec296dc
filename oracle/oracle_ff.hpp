// Double-word FP32 arithmetic for the oracle (test infrastructure only): a value is the
// unevaluated sum hi + lo of two floats with |lo| <= ulp(hi) / 2, about 48 significant bits,
// computed with FP32 operations and correctly rounded fmaf only.
//  - 2Sum (Knuth), Fast2Sum (Dekker), 2Prod with an FMA: the error-free transforms.
//  - add: AccurateDWPlusDW, mul: DWTimesDW3 (Joldes, Muller, Popescu, "Tight and rigorous error
//    bounds for basic building blocks of double-word arithmetic", ACM TOMS 44(2), 2017,
//    Algorithms 6 and 12).
//  - div: one long-division correction, q1 = xh / yh, q2 = (x - q1 y)_h / yh (the double-double
//    library of Hida, Li and Bailey, "sloppy" division); sqrt: s = sqrt(xh) corrected by
//    (x - s^2) / (2 s) with s^2 formed exactly.
// The device code (vkhrt_amd/csrc/ffloat.hpp) runs the same algorithms, so both round alike.
// Pinned against exact rational arithmetic in tests/test_ff_oracle.py.
// The type and its functions live in orc::dw and are found by argument-dependent lookup, so
// an unqualified sqrt / fabs on a float elsewhere in orc still resolves to the float function.
#pragma once
#include <cmath>

namespace orc {
namespace dw {

struct FF {
    float hi = 0.0f, lo = 0.0f;
    FF() = default;
    FF(float h) : hi(h), lo(0.0f) {}  // NOLINT: implicit like a float
    FF(int v) : hi(float(v)), lo(0.0f) {}
    FF(double v) : hi(float(v)), lo(float(v - double(float(v)))) {}
    FF(float h, float l) : hi(h), lo(l) {}
    explicit operator double() const { return double(hi) + double(lo); }
    explicit operator float() const { return hi; }
};

inline void two_sum(float a, float b, float& s, float& e) {
    s = a + b;
    const float a1 = s - b;
    const float b1 = s - a1;
    e = (a - a1) + (b - b1);
}
inline void fast_two_sum(float a, float b, float& s, float& e) {
    s = a + b;
    e = b - (s - a);
}
inline void two_prod(float a, float b, float& p, float& e) {
    p = a * b;
    e = std::fmaf(a, b, -p);
}

inline FF operator+(FF x, FF y) {
    float sh, sl, th, tl, vh, vl, zh, zl;
    two_sum(x.hi, y.hi, sh, sl);
    two_sum(x.lo, y.lo, th, tl);
    const float c = sl + th;
    fast_two_sum(sh, c, vh, vl);
    const float w = tl + vl;
    fast_two_sum(vh, w, zh, zl);
    return FF(zh, zl);
}
inline FF operator-(FF x) { return FF(-x.hi, -x.lo); }
inline FF operator-(FF x, FF y) { return x + (-y); }
inline FF operator*(FF x, FF y) {
    float ch, cl1, zh, zl;
    two_prod(x.hi, y.hi, ch, cl1);
    const float tl0 = x.lo * y.lo;
    const float tl1 = std::fmaf(x.hi, y.lo, tl0);
    const float cl2 = std::fmaf(x.lo, y.hi, tl1);
    const float cl3 = cl1 + cl2;
    fast_two_sum(ch, cl3, zh, zl);
    return FF(zh, zl);
}
inline FF operator/(FF x, FF y) {
    const float q1 = x.hi / y.hi;
    const FF r = x - y * FF(q1);
    const float q2 = r.hi / y.hi;
    float zh, zl;
    fast_two_sum(q1, q2, zh, zl);
    return FF(zh, zl);
}
inline FF sqrt(FF x) {
    if (!(x.hi > 0.0f)) return FF(std::sqrt(x.hi));
    const float s = std::sqrt(x.hi);
    float p, e;
    two_prod(s, s, p, e);
    const float r = ((x.hi - p) - e) + x.lo;
    const float corr = r / (2.0f * s);
    float zh, zl;
    fast_two_sum(s, corr, zh, zl);
    return FF(zh, zl);
}
inline FF fabs(FF x) { return x.hi < 0.0f || (x.hi == 0.0f && x.lo < 0.0f) ? -x : x; }
inline FF& operator+=(FF& x, FF y) { return x = x + y; }

inline bool operator<(FF x, FF y) { return x.hi < y.hi || (x.hi == y.hi && x.lo < y.lo); }
inline bool operator>(FF x, FF y) { return y < x; }
inline bool operator<=(FF x, FF y) { return x.hi < y.hi || (x.hi == y.hi && x.lo <= y.lo); }
inline bool operator>=(FF x, FF y) { return y <= x; }

}  // namespace dw
using dw::FF;
}  // namespace orc
