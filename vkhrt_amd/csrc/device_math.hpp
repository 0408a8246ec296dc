// Device math for the vkhrt hair path on gfx950.
//
// Every helper spells out its IEEE operation order so FP32 results are reproducible on the
// host: GLSL's vec3 expressions are evaluated component-wise, left to right, and no
// multiply-add is fused in code compiled with -ffp-contract=off (the intersection code).
// Division and square root are the correctly rounded HIP defaults.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vk {

struct f3 {
    float x, y, z;
};
__host__ __device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__host__ __device__ __forceinline__ f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ __forceinline__ f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ __forceinline__ f3 scl(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
__host__ __device__ __forceinline__ f3 scr(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__host__ __device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return f3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__host__ __device__ __forceinline__ float len3(f3 a) { return sqrtf(dot3(a, a)); }
// glm normalize: v * (1 / sqrt(dot(v, v)))
__host__ __device__ __forceinline__ f3 nrm3(f3 a) { return scr(a, 1.0f / sqrtf(dot3(a, a))); }
// GLSL min / max (NaN in x propagates like the GLSL definitions)
__host__ __device__ __forceinline__ float gmin(float x, float y) { return y < x ? y : x; }
__host__ __device__ __forceinline__ float gmax(float x, float y) { return x < y ? y : x; }
__host__ __device__ __forceinline__ f3 ld3(const float* p) { return f3{p[0], p[1], p[2]}; }

struct d3 {
    double x, y, z;
};
__host__ __device__ __forceinline__ d3 add(d3 a, d3 b) { return d3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ __forceinline__ d3 sub(d3 a, d3 b) { return d3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ __forceinline__ d3 scl(double s, d3 a) { return d3{s * a.x, s * a.y, s * a.z}; }
__host__ __device__ __forceinline__ d3 scr(d3 a, double s) { return d3{a.x * s, a.y * s, a.z * s}; }
__host__ __device__ __forceinline__ double dot3(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ __forceinline__ d3 cross3(d3 a, d3 b) {
    return d3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__host__ __device__ __forceinline__ d3 nrm3(d3 a) { return scr(a, 1.0 / sqrt(dot3(a, a))); }
__host__ __device__ __forceinline__ d3 tod(f3 a) { return d3{double(a.x), double(a.y), double(a.z)}; }

// Rows 0..2 of inverse(mat4(c0, c1, c2, c3)) for columns c0..c2 = (e, 0) and c3 = (o, 1), by
// Laplace expansion over 2x2 minors (D. Eberly, "The Laplace Expansion Theorem", 2008),
// adjugate times 1/det. The full general formula is evaluated (no algebraic shortcut for
// the zero bottom row) so the result is bit-identical to a generic GLSL-style inverse.
// out[r*4 + c] = B[r][c]; the fourth row is never needed (TransformCurve drops w).
template <class R, class V>
__host__ __device__ __forceinline__ void rcc_rows(V e1, V e2, V e3, V o, R out[12]) {
    // A[r][c] = column c, row r.
    const R A00 = e1.x, A10 = e1.y, A20 = e1.z, A30 = R(0);
    const R A01 = e2.x, A11 = e2.y, A21 = e2.z, A31 = R(0);
    const R A02 = e3.x, A12 = e3.y, A22 = e3.z, A32 = R(0);
    const R A03 = o.x, A13 = o.y, A23 = o.z, A33 = R(1);
    R s0 = A00 * A11 - A10 * A01;
    R s1 = A00 * A12 - A10 * A02;
    R s2 = A00 * A13 - A10 * A03;
    R s3 = A01 * A12 - A11 * A02;
    R s4 = A01 * A13 - A11 * A03;
    R s5 = A02 * A13 - A12 * A03;
    R c5 = A22 * A33 - A32 * A23;
    R c4 = A21 * A33 - A31 * A23;
    R c3 = A21 * A32 - A31 * A22;
    R c2 = A20 * A33 - A30 * A23;
    R c1 = A20 * A32 - A30 * A22;
    R c0 = A20 * A31 - A30 * A21;
    R det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
    R id = R(1) / det;
    out[0] = (A11 * c5 - A12 * c4 + A13 * c3) * id;
    out[1] = (-A01 * c5 + A02 * c4 - A03 * c3) * id;
    out[2] = (A31 * s5 - A32 * s4 + A33 * s3) * id;
    out[3] = (-A21 * s5 + A22 * s4 - A23 * s3) * id;
    out[4] = (-A10 * c5 + A12 * c2 - A13 * c1) * id;
    out[5] = (A00 * c5 - A02 * c2 + A03 * c1) * id;
    out[6] = (-A30 * s5 + A32 * s2 - A33 * s1) * id;
    out[7] = (A20 * s5 - A22 * s2 + A23 * s1) * id;
    out[8] = (A10 * c4 - A11 * c2 + A13 * c0) * id;
    out[9] = (-A00 * c4 + A01 * c2 - A03 * c0) * id;
    out[10] = (A30 * s4 - A31 * s2 + A33 * s0) * id;
    out[11] = (-A20 * s4 + A21 * s2 - A23 * s0) * id;
}

// ray.glsl:13-17 basis: v from the larger of |w.x|, |w.y|; u = cross(v, w).
template <class V>
__host__ __device__ __forceinline__ void onb(V w, V& u, V& v) {
    if (fabs(w.x) > fabs(w.y)) {
        V t;
        t.x = -w.z;
        t.y = 0;
        t.z = w.x;
        v = nrm3(t);
    } else {
        V t;
        t.x = 0;
        t.y = w.z;
        t.z = -w.y;
        v = nrm3(t);
    }
    u = cross3(v, w);
}

}  // namespace vk
