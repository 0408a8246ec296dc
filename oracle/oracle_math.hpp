// Oracle for the vkhrt hair ray-tracing path: CPU restatement of the reference's
// GLSL device code and host geometry processing.
//
// THIS IS TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker or the CPU baseline. The
// product library (vkhrt_amd/) never includes or links anything from oracle/.
//
// Every function is templated on the real type T so the same code runs as the FP32
// transliteration (T=float, the reference's arithmetic) and as the FP64 error
// reference (T=double). Operation order follows the GLSL text left to right; the
// oracle is compiled with -ffp-contract=off so no multiply-add is fused.
//
// Parity pins (tests/test_oracle_*.py): every function below is checked against a
// closed form, an invariant or brute force; see the per-function notes.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace orc {

// sqrt / fabs found by argument-dependent lookup, so a user scalar type (the double-word FF
// of oracle_ff.hpp) runs the same templates; for float and double these are std's.
template <class T>
inline T osqrt(T x) {
    using std::sqrt;
    return sqrt(x);
}
template <class T>
inline T ofabs(T x) {
    using std::fabs;
    return fabs(x);
}

template <class T>
struct V3 {
    T x, y, z;
};
template <class T>
inline V3<T> mk(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T>
inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T>
inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T>
inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T>
inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T>
inline V3<T> operator-(V3<T> a, T s) { return {a.x - s, a.y - s, a.z - s}; }
template <class T>
inline V3<T> operator+(V3<T> a, T s) { return {a.x + s, a.y + s, a.z + s}; }
// GLSL dot: x*x' + y*y' + z*z', summed left to right.
template <class T>
inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// GLSL/glm cross.
template <class T>
inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <class T>
inline T length(V3<T> a) { return osqrt(dot(a, a)); }
// glm::normalize: v * inversesqrt(dot(v, v)), inversesqrt(x) = 1 / sqrt(x).
template <class T>
inline V3<T> normalize(V3<T> a) { return a * (T(1) / osqrt(dot(a, a))); }
// GLSL min/max definitions (min(x,y) = y < x ? y : x; max(x,y) = x < y ? y : x).
template <class T>
inline T gmin(T x, T y) { return y < x ? y : x; }
template <class T>
inline T gmax(T x, T y) { return x < y ? y : x; }
template <class T>
inline V3<T> vmin(V3<T> a, V3<T> b) { return {gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)}; }
template <class T>
inline V3<T> vmax(V3<T> a, V3<T> b) { return {gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)}; }
template <class T, class U>
inline V3<T> cvt(V3<U> a) { return {T(a.x), T(a.y), T(a.z)}; }

// Column-major 4x4 (GLSL / glm convention): m[c][r].
template <class T>
struct M4 {
    T m[4][4];
};
// GLSL mat4 * vec4: r[i] = m[0][i]*v.x + m[1][i]*v.y + m[2][i]*v.z + m[3][i]*v.w.
template <class T>
inline void mul4(const M4<T>& M, const T v[4], T out[4]) {
    for (int i = 0; i < 4; ++i) out[i] = M.m[0][i] * v[0] + M.m[1][i] * v[1] + M.m[2][i] * v[2] + M.m[3][i] * v[3];
}
template <class T>
inline V3<T> xform_point(const M4<T>& M, V3<T> p) {
    T v[4] = {p.x, p.y, p.z, T(1)}, o[4];
    mul4(M, v, o);
    return {o[0], o[1], o[2]};
}

// GLSL inverse(mat4). Cofactor inverse by Laplace expansion over 2x2 minors of the top
// and bottom row pairs (D. Eberly, "The Laplace Expansion Theorem: Computing the
// Determinants and Inverses of Matrices", 2008), adjugate scaled by 1/det.
// A[r][c] = m[c][r]. Pinned by: A * inverse(A) = I on random matrices, and inverse of
// an orthonormal frame equals its transpose within ulps (tests/test_oracle_math.py).
template <class T>
inline M4<T> inverse(const M4<T>& M) {
#define A_(r, c) M.m[c][r]
    T s0 = A_(0, 0) * A_(1, 1) - A_(1, 0) * A_(0, 1);
    T s1 = A_(0, 0) * A_(1, 2) - A_(1, 0) * A_(0, 2);
    T s2 = A_(0, 0) * A_(1, 3) - A_(1, 0) * A_(0, 3);
    T s3 = A_(0, 1) * A_(1, 2) - A_(1, 1) * A_(0, 2);
    T s4 = A_(0, 1) * A_(1, 3) - A_(1, 1) * A_(0, 3);
    T s5 = A_(0, 2) * A_(1, 3) - A_(1, 2) * A_(0, 3);
    T c5 = A_(2, 2) * A_(3, 3) - A_(3, 2) * A_(2, 3);
    T c4 = A_(2, 1) * A_(3, 3) - A_(3, 1) * A_(2, 3);
    T c3 = A_(2, 1) * A_(3, 2) - A_(3, 1) * A_(2, 2);
    T c2 = A_(2, 0) * A_(3, 3) - A_(3, 0) * A_(2, 3);
    T c1 = A_(2, 0) * A_(3, 2) - A_(3, 0) * A_(2, 2);
    T c0 = A_(2, 0) * A_(3, 1) - A_(3, 0) * A_(2, 1);
    T det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
    T id = T(1) / det;
    T B[4][4];
    B[0][0] = (A_(1, 1) * c5 - A_(1, 2) * c4 + A_(1, 3) * c3) * id;
    B[0][1] = (-A_(0, 1) * c5 + A_(0, 2) * c4 - A_(0, 3) * c3) * id;
    B[0][2] = (A_(3, 1) * s5 - A_(3, 2) * s4 + A_(3, 3) * s3) * id;
    B[0][3] = (-A_(2, 1) * s5 + A_(2, 2) * s4 - A_(2, 3) * s3) * id;
    B[1][0] = (-A_(1, 0) * c5 + A_(1, 2) * c2 - A_(1, 3) * c1) * id;
    B[1][1] = (A_(0, 0) * c5 - A_(0, 2) * c2 + A_(0, 3) * c1) * id;
    B[1][2] = (-A_(3, 0) * s5 + A_(3, 2) * s2 - A_(3, 3) * s1) * id;
    B[1][3] = (A_(2, 0) * s5 - A_(2, 2) * s2 + A_(2, 3) * s1) * id;
    B[2][0] = (A_(1, 0) * c4 - A_(1, 1) * c2 + A_(1, 3) * c0) * id;
    B[2][1] = (-A_(0, 0) * c4 + A_(0, 1) * c2 - A_(0, 3) * c0) * id;
    B[2][2] = (A_(3, 0) * s4 - A_(3, 1) * s2 + A_(3, 3) * s0) * id;
    B[2][3] = (-A_(2, 0) * s4 + A_(2, 1) * s2 - A_(2, 3) * s0) * id;
    B[3][0] = (-A_(1, 0) * c3 + A_(1, 1) * c1 - A_(1, 2) * c0) * id;
    B[3][1] = (A_(0, 0) * c3 - A_(0, 1) * c1 + A_(0, 2) * c0) * id;
    B[3][2] = (-A_(3, 0) * s3 + A_(3, 1) * s1 - A_(3, 2) * s0) * id;
    B[3][3] = (A_(2, 0) * s3 - A_(2, 1) * s1 + A_(2, 2) * s0) * id;
#undef A_
    M4<T> R;
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) R.m[c][r] = B[r][c];
    return R;
}

// ---------------------------------------------------------------- ray.glsl
template <class T>
struct Ray {
    V3<T> origin, direction;
};

// ray.glsl:13-17 MakeOrthonormalBasis. Pinned by orthonormality checks.
template <class T>
inline void MakeOrthonormalBasis(V3<T>& u, V3<T>& v, V3<T> w) {
    v = ofabs(w.x) > ofabs(w.y) ? normalize(mk(-w.z, T(0), w.x)) : normalize(mk(T(0), w.z, -w.y));
    u = cross(v, w);
}

// ray.glsl:19-33 CreateRCCMatrix = inverse(mat4(e1, e2, e3, origin)).
// Pinned by M*o = 0 and M*(o + d) = (0, 0, |d|) within ulps.
template <class T>
inline M4<T> CreateRCCMatrix(const Ray<T>& ray) {
    V3<T> e1, e2, e3 = normalize(ray.direction);
    MakeOrthonormalBasis(e1, e2, e3);
    M4<T> R;
    T cols[4][4] = {{e1.x, e1.y, e1.z, T(0)}, {e2.x, e2.y, e2.z, T(0)}, {e3.x, e3.y, e3.z, T(0)},
                    {ray.origin.x, ray.origin.y, ray.origin.z, T(1)}};
    std::memcpy(R.m, cols, sizeof(cols));
    return inverse(R);
}

// ---------------------------------------------------------------- curve.glsl
template <class T>
struct Curve {
    V3<T> start, cp1, cp2, end;
};

// curve.glsl:9-21 SampleCurvePoint (Bernstein form). Pinned by de Casteljau in FP64.
template <class T>
inline V3<T> SampleCurvePoint(const Curve<T>& c, T t) {
    T u = T(1) - t;
    T tt = t * t;
    T uu = u * u;
    T uuu = uu * u;
    T ttt = tt * t;
    return uuu * c.start + T(3) * uu * t * c.cp1 + T(3) * u * tt * c.cp2 + ttt * c.end;
}
// Scalar Bernstein for per-control-point radii (same weights and order).
template <class T>
inline T SampleCurveScalar(const T r[4], T t) {
    T u = T(1) - t;
    T tt = t * t;
    T uu = u * u;
    T uuu = uu * u;
    T ttt = tt * t;
    return uuu * r[0] + T(3) * uu * t * r[1] + T(3) * u * tt * r[2] + ttt * r[3];
}

// curve.glsl:23-31 SampleCurveAxis (derivative). Pinned by FP64 central differences.
template <class T>
inline V3<T> SampleCurveAxis(const Curve<T>& c, T t) {
    T u = T(1) - t;
    return T(-3) * u * u * c.start + T(3) * (T(3) * t * t - T(4) * t + T(1)) * c.cp1 +
           T(3) * (T(2) - T(3) * t) * t * c.cp2 + T(3) * t * t * c.end;
}
template <class T>
inline T SampleCurveScalarDeriv(const T r[4], T t) {
    T u = T(1) - t;
    return T(-3) * u * u * r[0] + T(3) * (T(3) * t * t - T(4) * t + T(1)) * r[1] +
           T(3) * (T(2) - T(3) * t) * t * r[2] + T(3) * t * t * r[3];
}

// curve.glsl:33-42 TransformCurve.
template <class T>
inline Curve<T> TransformCurve(const Curve<T>& c, const M4<T>& M) {
    return {xform_point(M, c.start), xform_point(M, c.cp1), xform_point(M, c.cp2), xform_point(M, c.end)};
}

// curve.glsl:44-47 CurveDistanceToCylinder (point-to-chord-line distance).
template <class T>
inline T CurveDistanceToCylinder(const Curve<T>& c, V3<T> p) {
    return length(cross(p - c.start, p - c.end)) / length(c.end - c.start);
}

// ---------------------------------------------------------------- cone.glsl
template <class T>
struct Cone {
    V3<T> center;
    T radius;
    V3<T> axis;
    T slant;
};
template <class T>
struct RCI {
    bool isIntersection;
    T s, dt, dp, dc, sp;
    T a, gscale;  // coefficient a ~ |axis|^2 (dist^2 - r^2) and |axis|^2 r^2 (grazing test)
};

// cone.glsl:21-62 RayConeIntersectRCC. With slant = 0 it reduces to ray vs infinite
// cylinder about the tangent line (pinned analytically in tests).
template <class T>
inline RCI<T> RayConeIntersectRCC(const Cone<T>& cone) {
    T r2 = cone.radius * cone.radius;
    T drr = cone.radius * cone.slant;
    T ddd = cone.axis.x * cone.axis.x + cone.axis.y * cone.axis.y;
    T dp = cone.center.x * cone.center.x + cone.center.y * cone.center.y;
    T cdd = cone.center.x * cone.axis.x + cone.center.y * cone.axis.y;
    T cxd = cone.center.x * cone.axis.y - cone.center.y * cone.axis.x;
    T c = ddd;
    T b = cone.axis.z * (drr - cdd);
    T cdz2 = cone.axis.z * cone.axis.z;
    ddd += cdz2;
    T a = T(2) * drr * cdd + cxd * cxd - ddd * r2 + dp * cdz2;
    T det = b * b - a * c;
    RCI<T> r;
    r.a = a;
    r.gscale = ddd * r2;
    r.isIntersection = det > T(0);
    r.s = (b - (r.isIntersection ? osqrt(det) : T(0))) / c;
    r.dt = (r.s * cone.axis.z - cdd) / ddd;
    r.dc = r.s * r.s + dp;
    r.sp = cdd / cone.axis.z;
    r.dp = dp + r.sp * r.sp;
    return r;
}

// ---------------------------------------------------------------- cylinder.glsl
// cylinder.glsl:8-46 RayCylinderIntersect (bool; sign of t ignored, as in the GLSL).
// Pinned by textbook capped-cylinder cases.
template <class T>
inline bool RayCylinderIntersect(const Ray<T>& ray, V3<T> p0, V3<T> p1, T radius) {
    V3<T> ba = p1 - p0;
    V3<T> oc = ray.origin - p0;
    T baba = dot(ba, ba);
    T bard = dot(ba, ray.direction);
    T baoc = dot(ba, oc);
    T k2 = baba - bard * bard;
    T k1 = baba * dot(oc, ray.direction) - baoc * bard;
    T k0 = baba * dot(oc, oc) - baoc * baoc - radius * radius * baba;
    T h = k1 * k1 - k2 * k0;
    if (h < T(0)) return false;
    h = osqrt(h);
    T t = (-k1 - h) / k2;
    T y = baoc + t * bard;
    if (y > T(0) && y < baba) return true;
    t = ((y < T(0) ? T(0) : baba) - baoc) / bard;
    if (ofabs(k1 + k2 * t) < h) return true;
    return false;
}

// ---------------------------------------------------------------- hair_intersection.rint
// Radius model. The reference hard-codes radius 0.02 and slant 0
// (hair_intersection.rint:18,62). Per-control-point radii (tapered strands) are an
// extension: radius(t) = Bernstein(r0..r3, t), slant = d radius / dt.
enum RadiusMode { RADIUS_CONSTANT = 0, RADIUS_PER_POINT = 1 };

template <class T>
struct PrhiOut {
    T t;            // 0 = no intersection (hair_intersection.rint:17,146)
    T u;            // curve parameter at convergence
    V3<T> n_ref;    // normalize(hitPosition - curvePosition) in world space (:74-76)
    V3<T> n_rcc;    // same vector evaluated in the ray-centric frame, then rotated back
    int iters;      // total cone iterations executed (both sides)
    bool polished;  // root came from the FP64 polish step
};

struct PrhiOpts {
    int radius_mode = RADIUS_CONSTANT;
    float radius = 0.02f;   // constant radius (reference: 0.02)
    bool fp64_polish = false;
    bool cull = true;       // the early-out cylinder test (:24-33); off to re-run a known hit
};

// FP64 continuation of a side that ran all iterations without converging while grazing.
// Same update rule as the FP32 loop, evaluated in double on a curve transformed with a
// double RCC matrix. Extension with no reference counterpart.
// Trigger: the FP32 side used all 8 iterations, its final t is in [0, 1], |dt| < 0.05 and
// the ray grazes the tube at the last iterate: |a| <= 1e-2 |axis|^2 r^2, where the cone
// coefficient a ~ |axis|^2 (d^2 - r^2) and d is the ray's distance to the cone axis (the
// ray is within ~0.5% of the radius of tangent; there the FP32 iteration converges slowly).
template <class T>
struct PolishState {
    double t, told, dt1, dt2;
};

template <class T>
PrhiOut<T> Prhi(const Ray<T>& ray, const Curve<T>& curve, const T radii[4], const PrhiOpts& o);

// Shared iteration body, returns true on convergence. Used for both FP32 and the FP64
// polish so the update rule is written once (hair_intersection.rint:46-116).
template <class T>
inline bool prhi_side_iterate(const Curve<T>& rcc, const T radii[4], int radius_mode, T const_r, T& t, T& told,
                              T& dt1, T& dt2, int max_iter, int first_i, int& iters, T& s_out, T& u_out,
                              Cone<T>& cone_out, bool& exhausted, T& last_dt, bool& grazing) {
    exhausted = false;
    for (int i = first_i; i < first_i + max_iter; ++i) {
        Cone<T> cone;
        cone.center = SampleCurvePoint(rcc, t);
        if (radius_mode == RADIUS_CONSTANT) {
            cone.radius = const_r;
            cone.slant = T(0);
        } else {
            cone.radius = SampleCurveScalar(radii, t);
            cone.slant = SampleCurveScalarDeriv(radii, t);
        }
        cone.axis = SampleCurveAxis(rcc, t);
        RCI<T> rci = RayConeIntersectRCC(cone);
        ++iters;
        grazing = ofabs(rci.a) <= T(1e-2f) * rci.gscale;
        if (rci.isIntersection && ofabs(rci.dt) < T(5e-5f)) {
            s_out = rci.s;  // relative to cone.center.z
            u_out = t;
            cone_out = cone;
            return true;
        }
        rci.dt = gmin(rci.dt, T(0.5f));
        rci.dt = gmax(rci.dt, T(-0.5f));
        dt1 = dt2;
        dt2 = rci.dt;
        last_dt = rci.dt;
        if (dt1 * dt2 < T(0)) {
            T tnext = T(0);
            if ((i & 3) == 0)
                tnext = T(0.5f) * (told + t);
            else
                tnext = (dt2 * told - dt1 * t) / (dt2 - dt1);
            told = t;
            t = tnext;
        } else {
            told = t;
            t += rci.dt;
        }
        if (t < T(0) || t > T(1)) return false;
    }
    exhausted = true;
    return false;
}

template <class T>
PrhiOut<T> Prhi(const Ray<T>& ray, const Curve<T>& curve, const T radii[4], const PrhiOpts& o) {
    PrhiOut<T> out{};
    out.t = T(0);
    T result = T(0);
    const T const_r = T(o.radius);
    T rpad = const_r;
    if (o.radius_mode == RADIUS_PER_POINT) rpad = gmax(gmax(radii[0], radii[1]), gmax(radii[2], radii[3]));

    if (o.cull) {
        T rmax = CurveDistanceToCylinder(curve, SampleCurvePoint(curve, T(0.5f)));
        rmax += rpad;
        if (!RayCylinderIntersect(ray, curve.start, curve.end, rmax)) return out;
    }

    M4<T> M = CreateRCCMatrix(ray);
    Curve<T> rcc = TransformCurve(curve, M);
    const V3<T> rayDirection = mk(T(0), T(0), T(1));
    V3<T> curveDirection = normalize(rcc.end - rcc.start);
    T tStart = dot(curveDirection, rayDirection) > T(0) ? T(0) : T(1);

    for (int side = 0; side < 2; ++side) {
        T t = tStart, told = T(0), dt1 = T(0), dt2 = T(0);
        T s_rel = T(0), u = T(0), last_dt = T(0);
        Cone<T> cone{};
        bool exhausted = false, grazing = false;
        bool conv = prhi_side_iterate(rcc, radii, o.radius_mode, const_r, t, told, dt1, dt2, 8, 0, out.iters, s_rel,
                                      u, cone, exhausted, last_dt, grazing);
        if (conv) {
            T s = s_rel + cone.center.z;
            result = s;
            out.t = s;
            out.u = u;
            V3<T> hitPosition = ray.origin + result * ray.direction;
            V3<T> curvePosition = SampleCurvePoint(curve, u);
            out.n_ref = normalize(hitPosition - curvePosition);
            // Same vector in ray-centric coordinates: hit = (0,0,s), centre = cone.center,
            // so hit - centre = (-cx, -cy, s_rel); rotate back with e1, e2, e3.
            V3<T> e1, e2, e3 = normalize(ray.direction);
            MakeOrthonormalBasis(e1, e2, e3);
            V3<T> nr = mk(-cone.center.x, -cone.center.y, s_rel);
            out.n_rcc = normalize(nr.x * e1 + nr.y * e2 + nr.z * e3);
        } else if (o.fp64_polish && exhausted && t >= T(0) && t <= T(1) && ofabs(last_dt) < T(0.05f) && grazing) {
            // FP64 polish of a grazing side (extension).
            Ray<double> rd{cvt<double>(ray.origin), cvt<double>(ray.direction)};
            Curve<double> cd{cvt<double>(curve.start), cvt<double>(curve.cp1), cvt<double>(curve.cp2),
                             cvt<double>(curve.end)};
            double rr[4] = {double(radii[0]), double(radii[1]), double(radii[2]), double(radii[3])};
            M4<double> Md = CreateRCCMatrix(rd);
            Curve<double> rccd = TransformCurve(cd, Md);
            double td = double(t), toldd = double(told), dt1d = double(dt1), dt2d = double(dt2);
            double sd = 0, ud = 0, lastd = 0;
            Cone<double> coned{};
            bool exd = false, grd = false;
            int it2 = 0;
            bool convd = prhi_side_iterate(rccd, rr, o.radius_mode, double(const_r), td, toldd, dt1d, dt2d, 8, 8,
                                           it2, sd, ud, coned, exd, lastd, grd);
            out.iters += it2;
            if (convd) {
                double s = sd + coned.center.z;
                result = T(s);
                out.t = T(s);
                out.u = T(ud);
                out.polished = true;
                V3<double> hit = rd.origin + s * rd.direction;
                V3<double> cp = SampleCurvePoint(cd, ud);
                out.n_ref = cvt<T>(normalize(hit - cp));
                V3<double> e1, e2, e3 = normalize(rd.direction);
                MakeOrthonormalBasis(e1, e2, e3);
                V3<double> nr = mk(-coned.center.x, -coned.center.y, sd);
                out.n_rcc = cvt<T>(normalize(nr.x * e1 + nr.y * e2 + nr.z * e3));
            }
        }
        if (result > T(0))
            break;
        else
            tStart = T(1) - tStart;
    }
    if (!(result > T(0))) out.t = T(0);
    return out;
}

// ---------------------------------------------------------------- shading / debug
// shading.glsl:1-10 Shade(normal) = |n . (0,-1,0)| * (0.4,0.2,0.1) + 0.3.
// The shadowed form multiplies the direct term by the light visibility (extension).
template <class T>
inline V3<T> Shade(V3<T> n, T visibility) {
    const V3<T> lightColor = mk(T(0.4f), T(0.2f), T(0.1f));
    const V3<T> lightDirection = mk(T(0), T(-1), T(0));
    const V3<T> ambientColor = mk(T(0.3f), T(0.3f), T(0.3f));
    T ndl = ofabs(dot(n, lightDirection));
    V3<T> color = (ndl * visibility) * lightColor;
    return color + ambientColor;
}

// debug.glsl:1-7 GetDebugColor(primitiveID) = palette[primitiveID % 6].
inline V3<float> GetDebugColor(int32_t primitiveID) {
    static const float colors[6][3] = {{1.0f, 0.0f, 0.3f}, {0.8f, 0.2f, 0.3f}, {0.6f, 0.4f, 0.3f},
                                       {0.4f, 0.6f, 0.3f}, {0.2f, 0.8f, 0.3f}, {0.0f, 1.0f, 0.3f}};
    int idx = primitiveID % 6;
    return mk(colors[idx][0], colors[idx][1], colors[idx][2]);
}

}  // namespace orc
