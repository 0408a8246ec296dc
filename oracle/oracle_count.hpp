// Operation-counting float for the oracle (test infrastructure only): a value computes exactly
// as a float and every arithmetic operation on it increments a counter, so running an oracle
// template with T = CountF counts the FP32 operations that code performs on its inputs.
// Classes: add/sub, mul, div, sqrt, min/max/abs/compare (not FLOPs, reported apart).
// FLOP = add + mul + div + sqrt (a fused multiply-add would count as 2; the oracle has none).
// The counters are per thread. Found by argument-dependent lookup like the double-word FF.
#pragma once
#include <cmath>
#include <cstdint>

namespace orc {
namespace cnt {

struct Counts {
    uint64_t add = 0, mul = 0, div = 0, sqrt = 0, cmp = 0;
    uint64_t flops() const { return add + mul + div + sqrt; }
};
inline Counts& counts() {
    static thread_local Counts c;
    return c;
}

struct CountF {
    float v = 0.0f;
    CountF() = default;
    CountF(float x) : v(x) {}  // NOLINT: implicit like a float
    CountF(int x) : v(float(x)) {}
    CountF(double x) : v(float(x)) {}
    explicit operator float() const { return v; }
    explicit operator double() const { return double(v); }
};

inline CountF operator+(CountF a, CountF b) {
    ++counts().add;
    return CountF(a.v + b.v);
}
inline CountF operator-(CountF a, CountF b) {
    ++counts().add;
    return CountF(a.v - b.v);
}
inline CountF operator-(CountF a) { return CountF(-a.v); }  // a sign flip folds into the consumer
inline CountF operator*(CountF a, CountF b) {
    ++counts().mul;
    return CountF(a.v * b.v);
}
inline CountF operator/(CountF a, CountF b) {
    ++counts().div;
    return CountF(a.v / b.v);
}
inline CountF& operator+=(CountF& a, CountF b) { return a = a + b; }
inline CountF& operator-=(CountF& a, CountF b) { return a = a - b; }
inline CountF& operator*=(CountF& a, CountF b) { return a = a * b; }
inline CountF sqrt(CountF a) {
    ++counts().sqrt;
    return CountF(std::sqrt(a.v));
}
inline CountF fabs(CountF a) {
    ++counts().cmp;
    return CountF(std::fabs(a.v));
}
inline bool operator<(CountF a, CountF b) {
    ++counts().cmp;
    return a.v < b.v;
}
inline bool operator>(CountF a, CountF b) {
    ++counts().cmp;
    return a.v > b.v;
}
inline bool operator<=(CountF a, CountF b) {
    ++counts().cmp;
    return a.v <= b.v;
}
inline bool operator>=(CountF a, CountF b) {
    ++counts().cmp;
    return a.v >= b.v;
}
inline bool operator==(CountF a, CountF b) {
    ++counts().cmp;
    return a.v == b.v;
}
inline bool operator!=(CountF a, CountF b) {
    ++counts().cmp;
    return a.v != b.v;
}

}  // namespace cnt
using cnt::CountF;
}  // namespace orc
