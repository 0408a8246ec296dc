// Primitive intersection on gfx950: phantom ray-hair (cubic Bezier), linear swept
// spheres and DOTS triangles. Included by trace.hip only, which is compiled with
// -ffp-contract=off so these FP32 sequences round exactly like the GLSL text they follow.
//
// Phantom: hair_intersection.rint:15-130 (Prhi), cone.glsl:21-62, curve.glsl:9-47,
// cylinder.glsl:8-46, ray.glsl:13-33. The ray-centric frame (basis and the rows of
// inverse(mat4(e1, e2, e3, o))) depends only on the ray, so it is computed once per ray
// (RayFrame) instead of once per ray x segment as in the GLSL; the arithmetic is the same.
#pragma once
#include "device_math.hpp"
#include "ffloat.hpp"

namespace vk {

enum RadiusModeDev { kRadiusConstant = 0, kRadiusPerPoint = 1 };

struct RayFrame {
    f3 o, d;      // ray in the primitive's space
    float M[12];  // rows 0..2 of the RCC matrix
};

__device__ __forceinline__ void make_frame(f3 o, f3 d, RayFrame& F) {
    F.o = o;
    F.d = d;
    f3 e1, e2, e3 = nrm3(d);
    onb(e3, e1, e2);
    rcc_rows<float>(e1, e2, e3, o, F.M);
}

// (M * vec4(p, 1)).xyz with GLSL mat*vec summation order.
__device__ __forceinline__ f3 rcc_xform(const float* M, f3 p) {
    return f3{M[0] * p.x + M[1] * p.y + M[2] * p.z + M[3] * 1.0f, M[4] * p.x + M[5] * p.y + M[6] * p.z + M[7] * 1.0f,
              M[8] * p.x + M[9] * p.y + M[10] * p.z + M[11] * 1.0f};
}

// curve.glsl:9-21 Bernstein point.
template <class V, class R>
__device__ __forceinline__ V bez_point(V p0, V p1, V p2, V p3, R t) {
    R u = R(1) - t;
    R tt = t * t;
    R uu = u * u;
    R uuu = uu * u;
    R ttt = tt * t;
    R w1 = R(3) * uu * t;
    R w2 = R(3) * u * tt;
    V r;
    r.x = uuu * p0.x + w1 * p1.x + w2 * p2.x + ttt * p3.x;
    r.y = uuu * p0.y + w1 * p1.y + w2 * p2.y + ttt * p3.y;
    r.z = uuu * p0.z + w1 * p1.z + w2 * p2.z + ttt * p3.z;
    return r;
}
template <class R>
__device__ __forceinline__ R bez_scalar(R a, R b, R c, R d, R t) {
    R u = R(1) - t;
    R tt = t * t;
    R uu = u * u;
    R uuu = uu * u;
    R ttt = tt * t;
    return uuu * a + R(3) * uu * t * b + R(3) * u * tt * c + ttt * d;
}
// curve.glsl:23-31 derivative.
template <class V, class R>
__device__ __forceinline__ V bez_axis(V p0, V p1, V p2, V p3, R t) {
    R u = R(1) - t;
    R w0 = R(-3) * u * u;
    R w1 = R(3) * (R(3) * t * t - R(4) * t + R(1));
    R w2 = R(3) * (R(2) - R(3) * t) * t;
    R w3 = R(3) * t * t;
    V r;
    r.x = w0 * p0.x + w1 * p1.x + w2 * p2.x + w3 * p3.x;
    r.y = w0 * p0.y + w1 * p1.y + w2 * p2.y + w3 * p3.y;
    r.z = w0 * p0.z + w1 * p1.z + w2 * p2.z + w3 * p3.z;
    return r;
}
template <class R>
__device__ __forceinline__ R bez_scalar_axis(R a, R b, R c, R d, R t) {
    R u = R(1) - t;
    return R(-3) * u * u * a + R(3) * (R(3) * t * t - R(4) * t + R(1)) * b + R(3) * (R(2) - R(3) * t) * t * c +
           R(3) * t * t * d;
}

// cone.glsl:21-62 for the quantities Prhi uses (isIntersection, s, dt).
template <class R>
__device__ __forceinline__ void cone_rcc(R cx, R cy, R ax, R ay, R az, R radius, R slant, bool& isect, R& s, R& dt,
                                         bool& grazing) {
    R r2 = radius * radius;
    R drr = radius * slant;
    R ddd = ax * ax + ay * ay;
    R dp = cx * cx + cy * cy;
    R cdd = cx * ax + cy * ay;
    R cxd = cx * ay - cy * ax;
    R c = ddd;
    R b = az * (drr - cdd);
    R cdz2 = az * az;
    ddd += cdz2;
    R a = R(2) * drr * cdd + cxd * cxd - ddd * r2 + dp * cdz2;
    R det = b * b - a * c;
    grazing = fabs(a) <= R(1e-2f) * (ddd * r2);
    isect = det > R(0);
    s = (b - (isect ? sqrt(det) : R(0))) / c;
    dt = (s * az - cdd) / ddd;
}

// cylinder.glsl:8-46 (bool; sign of t ignored as in the GLSL).
__device__ __forceinline__ bool ray_cylinder(f3 o, f3 d, f3 p0, f3 p1, float radius) {
    f3 ba = sub(p1, p0);
    f3 oc = sub(o, p0);
    float baba = dot3(ba, ba);
    float bard = dot3(ba, d);
    float baoc = dot3(ba, oc);
    float k2 = baba - bard * bard;
    float k1 = baba * dot3(oc, d) - baoc * bard;
    float k0 = baba * dot3(oc, oc) - baoc * baoc - radius * radius * baba;
    float h = k1 * k1 - k2 * k0;
    if (h < 0.0f) return false;
    h = sqrtf(h);
    float t = (-k1 - h) / k2;
    float y = baoc + t * bard;
    if (y > 0.0f && y < baba) return true;
    t = ((y < 0.0f ? 0.0f : baba) - baoc) / bard;
    return fabsf(k1 + k2 * t) < h;
}

struct PhantomHit {
    float t;      // distance along the normalized ray direction (hair_intersection.rint:70-71)
    float u;      // curve parameter at convergence
    float ncx, ncy, nsz;  // hit - curve(u) in the ray-centric frame
    bool polished;
    bool deferred;  // kPolishDefer: this curve needs the FP64 continuation (evaluated later)
};

// FP64 polish handling inside Prhi: off, evaluated in place, or deferred (the FP32 pass
// stops at the first side that would enter the FP64 continuation and flags the curve; a
// later kernel re-runs the full Prhi with kPolishInline for flagged ray/curve pairs).
enum PolishMode { kPolishOff = 0, kPolishInline = 1, kPolishDefer = 2 };

// One side of the Prhi iteration (hair_intersection.rint:46-116) on an RCC-space curve.
template <class V, class R>
__device__ __forceinline__ int prhi_iterate(V q0, V q1, V q2, V q3, const R rr[4], int radius_mode, R const_r, R& t,
                                            R& told, R& dt1, R& dt2, int i0, int n, R& s_rel, R& cx, R& cy, R& cz,
                                            R& last_dt) {
    // returns 1 converged, 0 left [0,1], 2 exhausted, 3 exhausted while grazing
    bool last_grazing = false;
    for (int i = i0; i < i0 + n; ++i) {
        V c = bez_point(q0, q1, q2, q3, t);
        V ax = bez_axis(q0, q1, q2, q3, t);
        R radius, slant;
        if (radius_mode == kRadiusConstant) {
            radius = const_r;
            slant = R(0);
        } else {
            radius = bez_scalar(rr[0], rr[1], rr[2], rr[3], t);
            slant = bez_scalar_axis(rr[0], rr[1], rr[2], rr[3], t);
        }
        bool isect, grazing;
        R s, dt;
        cone_rcc(c.x, c.y, ax.x, ax.y, ax.z, radius, slant, isect, s, dt, grazing);
        last_grazing = grazing;
        if (isect && fabs(dt) < R(5e-5f)) {
            s_rel = s;
            cx = c.x;
            cy = c.y;
            cz = c.z;
            return 1;
        }
        dt = dt > R(0.5f) ? R(0.5f) : dt;    // GLSL min(dt, 0.5)
        dt = dt < R(-0.5f) ? R(-0.5f) : dt;  // GLSL max(dt, -0.5)
        dt1 = dt2;
        dt2 = dt;
        last_dt = dt;
        if (dt1 * dt2 < R(0)) {
            R tnext;
            if ((i & 3) == 0)
                tnext = R(0.5f) * (told + t);
            else
                tnext = (dt2 * told - dt1 * t) / (dt2 - dt1);
            told = t;
            t = tnext;
        } else {
            told = t;
            t += dt;
        }
        if (t < R(0) || t > R(1)) return 0;
    }
    return last_grazing ? 3 : 2;
}

// FP64 continuation of a grazing side (extension, no reference counterpart): the FP32 side
// used all 8 iterations, its final t is in [0, 1], |dt| < 0.05 and the ray grazes the tube
// at the last iterate (|a| <= 1e-2 |axis|^2 r^2, a ~ |axis|^2 (d^2 - r^2) with d the ray's
// distance to the cone axis). The same update rule runs 8 more iterations in double on the
// curve transformed by a double RCC matrix.
__device__ __forceinline__ bool prhi_polish_f64(f3 of, f3 df, const float4* cv, int radius_mode, float const_r,
                                                float t32, float told32, float dt1_32, float dt2_32, PhantomHit& out) {
    d3 o = tod(of), d = tod(df);
    d3 e1, e2, e3 = nrm3(d);
    onb(e3, e1, e2);
    double M[12];
    rcc_rows<double>(e1, e2, e3, o, M);
    d3 q[4];
    double rr[4];
    for (int k = 0; k < 4; ++k) {
        d3 p = d3{double(cv[k].x), double(cv[k].y), double(cv[k].z)};
        q[k] = d3{M[0] * p.x + M[1] * p.y + M[2] * p.z + M[3] * 1.0, M[4] * p.x + M[5] * p.y + M[6] * p.z + M[7] * 1.0,
                  M[8] * p.x + M[9] * p.y + M[10] * p.z + M[11] * 1.0};
        rr[k] = double(cv[k].w);
    }
    double t = t32, told = told32, dt1 = dt1_32, dt2 = dt2_32, s_rel = 0, cx = 0, cy = 0, cz = 0, last = 0;
    int st = prhi_iterate(q[0], q[1], q[2], q[3], rr, radius_mode, double(const_r), t, told, dt1, dt2, 8, 8, s_rel, cx,
                          cy, cz, last);
    if (st != 1) return false;
    const double s = s_rel + cz;
    out.t = float(s);
    out.u = float(t);
    // normal in world space from the double frame, rounded once
    d3 nr = d3{-cx, -cy, s_rel};
    d3 n = nrm3(add(add(scl(nr.x, e1), scl(nr.y, e2)), scl(nr.z, e3)));
    out.ncx = float(n.x);
    out.ncy = float(n.y);
    out.nsz = float(n.z);
    out.polished = true;
    return true;
}

// hair_intersection.rint:15-130 Prhi. Returns true with out.t > 0 on an intersection.
// RMODE (constant / per-point radius) and POLISH are compile-time so each traversal kernel
// carries only the arithmetic it uses and no device-function calls.
// `rmax` is the curve's early-out cylinder radius (:21-22), precomputed per curve at build
// time with the same arithmetic. CULL = false skips that test for callers that ran it already.
template <int RMODE, int PMODE, bool CULL = true>
__device__ __forceinline__ bool prhi(const RayFrame& F, const float4* cv, float rmax, float const_r, PhantomHit& out) {
    out.deferred = false;
    constexpr int radius_mode = RMODE;
    const f3 p0 = f3{cv[0].x, cv[0].y, cv[0].z}, p1 = f3{cv[1].x, cv[1].y, cv[1].z};
    const f3 p2 = f3{cv[2].x, cv[2].y, cv[2].z}, p3 = f3{cv[3].x, cv[3].y, cv[3].z};
    // early out (:24-33)
    if (CULL && !ray_cylinder(F.o, F.d, p0, p3, rmax)) return false;
    // ray-centric curve (:38-39)
    f3 q0 = rcc_xform(F.M, p0), q1 = rcc_xform(F.M, p1), q2 = rcc_xform(F.M, p2), q3 = rcc_xform(F.M, p3);
    // start end (:42-44): dot(normalize(q3 - q0), (0,0,1)) > 0
    f3 cd = nrm3(sub(q3, q0));
    float dz = cd.x * 0.0f + cd.y * 0.0f + cd.z * 1.0f;
    float tStart = dz > 0.0f ? 0.0f : 1.0f;
    const float rr[4] = {cv[0].w, cv[1].w, cv[2].w, cv[3].w};
    for (int side = 0; side < 2; ++side) {
        float t = tStart, told = 0.0f, dt1 = 0.0f, dt2 = 0.0f, s_rel = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f, last = 0.0f;
        int st = prhi_iterate(q0, q1, q2, q3, rr, radius_mode, const_r, t, told, dt1, dt2, 0, 8, s_rel, cx, cy, cz, last);
        if (st == 1) {
            float s = s_rel + cz;
            if (s > 0.0f) {
                out.t = s;
                out.u = t;
                out.ncx = -cx;
                out.ncy = -cy;
                out.nsz = s_rel;
                out.polished = false;
                return true;
            }
        } else if (PMODE != kPolishOff && st == 3 && t >= 0.0f && t <= 1.0f && fabsf(last) < 0.05f) {
            if (PMODE == kPolishDefer) {
                out.deferred = true;
                return false;
            }
            PhantomHit ph;
            if (prhi_polish_f64(F.o, F.d, cv, radius_mode, const_r, t, told, dt1, dt2, ph) && ph.t > 0.0f) {
                out = ph;
                out.deferred = false;
                return true;
            }
        }
        tStart = 1.0f - tStart;
    }
    return false;
}

// Final-hit re-run of Prhi in double-word FP32 (extension, ffloat.hpp): the iteration of the
// FP64 evaluation from the start, without the early-out test, for a ray / curve pair the FP32
// pass already chose. On convergence with s > 0 it returns t, u and the unit normal (ray-
// centric vector rotated back), rounded once, in `out` (marked polished: the normal is final).
__device__ __forceinline__ bool prhi_ff(f3 of, f3 df, const float4* cv, int radius_mode, float const_r,
                                        PhantomHit& out) {
    const ff3 o = dw::toff(of), d = dw::toff(df);
    ff3 e1, e2, e3 = nrm3(d);
    onb(e3, e1, e2);
    ff M[12];
    rcc_rows<ff>(e1, e2, e3, o, M);
    ff3 q[4];
    ff rr[4];
    for (int k = 0; k < 4; ++k) {
        const ff3 p = ff3{ff(cv[k].x), ff(cv[k].y), ff(cv[k].z)};
        q[k] = ff3{M[0] * p.x + M[1] * p.y + M[2] * p.z + M[3] * ff(1), M[4] * p.x + M[5] * p.y + M[6] * p.z + M[7] * ff(1),
                   M[8] * p.x + M[9] * p.y + M[10] * p.z + M[11] * ff(1)};
        rr[k] = ff(cv[k].w);
    }
    const ff3 cd = nrm3(sub(q[3], q[0]));
    const ff dz = cd.x * ff(0) + cd.y * ff(0) + cd.z * ff(1);
    ff tStart = dz > ff(0) ? ff(0) : ff(1);
    for (int side = 0; side < 2; ++side) {
        ff t = tStart, told = ff(0), dt1 = ff(0), dt2 = ff(0), s_rel = ff(0), cx = ff(0), cy = ff(0), cz = ff(0),
           last = ff(0);
        const int st = prhi_iterate(q[0], q[1], q[2], q[3], rr, radius_mode, ff(const_r), t, told, dt1, dt2, 0, 8, s_rel,
                                    cx, cy, cz, last);
        if (st == 1) {
            const ff s = s_rel + cz;
            if (s > ff(0)) {
                const ff3 n = nrm3(add(add(scl(-cx, e1), scl(-cy, e2)), scl(s_rel, e3)));
                out.t = s.hi;
                out.u = t.hi;
                out.ncx = n.x.hi;
                out.ncy = n.y.hi;
                out.nsz = n.z.hi;
                out.polished = true;
                out.deferred = false;
                return true;
            }
        }
        tStart = ff(1) - tStart;
    }
    return false;
}

// World-space unit normal of a phantom hit: rotate the ray-centric vector back with the
// basis of the ray (normalize(n.x e1 + n.y e2 + n.z e3)). Polished hits carry it already.
__device__ __forceinline__ f3 phantom_normal(f3 d, const PhantomHit& h) {
    if (h.polished) return f3{h.ncx, h.ncy, h.nsz};
    f3 e1, e2, e3 = nrm3(d);
    onb(e3, e1, e2);
    return nrm3(add(add(scl(h.ncx, e1), scl(h.ncy, e2)), scl(h.nsz, e3)));
}

// ------------------------------------------------------------ linear swept spheres
// Entry point of the ray into the convex hull of spheres (a, ra), (b, rb); the start cap is
// left out when chained to the previous piece. Defined by this build (the reference's LSS
// intersection runs in the NVIDIA RT core); the oracle restates the same definition.
struct LssHit {
    float t, u;
    f3 n;
};
__device__ __forceinline__ bool lss_intersect(f3 o, f3 d, f3 a, float ra, f3 b, float rb, bool cap_a, float tmin,
                                              float tmax, LssHit& out) {
    const float dd = dot3(d, d);
    const float ta = dot3(sub(a, o), d) / dd;
    const f3 os = add(o, scl(ta, d));
    const f3 w = sub(os, a);
    const f3 ba = sub(b, a);
    const float l2 = dot3(ba, ba);
    const float dr = ra - rb;
    const float k = l2 - dr * dr;
    const float wd = dot3(w, d), ww = dot3(w, w), wb = dot3(w, ba), db = dot3(d, ba);
    float best = INFINITY, bestY = 0.0f;
    int which = -1;
    const bool sphere_only = !(k > 0.0f);
    if (!sphere_only) {
        const float q = ra * l2 - dr * wb;
        const float A = k * (l2 * dd - db * db) - dr * dr * db * db;
        const float B = k * (l2 * wd - wb * db) + q * dr * db;
        const float C = k * (l2 * ww - wb * wb) - q * q;
        const float disc = B * B - A * C;
        if (disc >= 0.0f && A != 0.0f) {
            const float sq = sqrtf(disc);
            const float roots[2] = {(-B - sq) / A, (-B + sq) / A};
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                float ts = roots[i];
                float Y = wb + ts * db;
                if (!(Y >= ra * dr && Y <= l2 + rb * dr)) continue;
                f3 p = add(w, scl(ts, d));
                f3 radial = sub(p, scl(Y / l2, ba));
                float rl = len3(radial);
                float nx_l = sqrtf(k);
                float dn = nx_l * dot3(d, radial) / rl + dr * db / sqrtf(l2);
                if (!(dn < 0.0f)) continue;
                if (ts < best) {
                    best = ts;
                    which = 0;
                    bestY = Y;
                }
            }
        }
    }
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        if (s == 0 && !cap_a && !sphere_only) continue;
        if (sphere_only && ((s == 0) != (ra >= rb))) continue;
        f3 c = s == 0 ? a : b;
        float r = s == 0 ? ra : rb;
        f3 wc = sub(os, c);
        float bq = dot3(wc, d);
        float cq = dot3(wc, wc) - r * r;
        float disc = bq * bq - dd * cq;
        if (disc < 0.0f) continue;
        float sq = sqrtf(disc);
        float ts = (-bq - sq) / dd;
        f3 p = add(w, scl(ts, d));
        float Y = dot3(p, ba);
        if (!sphere_only) {
            if (s == 0 && !(Y <= ra * dr)) continue;
            if (s == 1 && !(Y >= l2 + rb * dr)) continue;
        }
        if (ts < best) {
            best = ts;
            which = 1 + s;
            bestY = Y;
        }
    }
    if (which < 0) return false;
    float t = ta + best;
    if (!(t >= tmin && t <= tmax)) return false;
    float u;
    if (which == 0)
        u = (bestY - ra * dr) / k;
    else
        u = which == 1 ? 0.0f : 1.0f;
    if (sphere_only) u = ra >= rb ? 0.0f : 1.0f;
    out.t = t;
    out.u = u;
    f3 hit = add(w, scl(best, d));
    f3 ctr = scl(u, ba);
    out.n = nrm3(sub(hit, ctr));
    return true;
}

// ------------------------------------------------------------ DOTS triangles
// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013), no facing cull. The paper
// re-evaluates zero edge functions in double; here they are re-evaluated in FP32 by Kahan's
// fused multiply-add 2x2 determinant, whose relative error is at most 2u (Jeannerod, Louvet,
// Muller, Math. Comp. 2013): a nonzero exact value keeps its sign and an exact zero stays
// zero. Per-ray shear constants come from TriRay.
struct TriRay {
    int kx, ky, kz;
    float Sx, Sy, Sz;
};
__device__ __forceinline__ float cmp3(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
__device__ __forceinline__ TriRay make_tri_ray(f3 d) {
    TriRay R;
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    if (cmp3(d, kz) < 0.0f) {
        int t = kx;
        kx = ky;
        ky = t;
    }
    R.kx = kx;
    R.ky = ky;
    R.kz = kz;
    R.Sx = cmp3(d, kx) / cmp3(d, kz);
    R.Sy = cmp3(d, ky) / cmp3(d, kz);
    R.Sz = 1.0f / cmp3(d, kz);
    return R;
}
// a*b - c*d: w = c*d rounded, e = w - c*d exactly, f = a*b - w rounded once; f + e.
__device__ __forceinline__ float det2_kahan(float a, float b, float c, float d) {
    const float w = c * d;
    const float e = fmaf(-c, d, w);
    const float f = fmaf(a, b, -w);
    return f + e;
}
__device__ __forceinline__ bool tri_intersect(const TriRay& R, f3 o, f3 v0, f3 v1, f3 v2, float tmin, float tmax,
                                              float& t, float& b1, float& b2) {
    f3 A = sub(v0, o), B = sub(v1, o), C = sub(v2, o);
    float Ax = cmp3(A, R.kx) - R.Sx * cmp3(A, R.kz);
    float Ay = cmp3(A, R.ky) - R.Sy * cmp3(A, R.kz);
    float Bx = cmp3(B, R.kx) - R.Sx * cmp3(B, R.kz);
    float By = cmp3(B, R.ky) - R.Sy * cmp3(B, R.kz);
    float Cx = cmp3(C, R.kx) - R.Sx * cmp3(C, R.kz);
    float Cy = cmp3(C, R.ky) - R.Sy * cmp3(C, R.kz);
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = det2_kahan(Cx, By, Cy, Bx);
        V = det2_kahan(Ax, Cy, Ay, Cx);
        W = det2_kahan(Bx, Ay, By, Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = R.Sz * cmp3(A, R.kz);
    float Bz = R.Sz * cmp3(B, R.kz);
    float Cz = R.Sz * cmp3(C, R.kz);
    float T = U * Az + V * Bz + W * Cz;
    float rcp = 1.0f / det;
    float tt = T * rcp;
    if (!(tt >= tmin && tt <= tmax)) return false;
    t = tt;
    b1 = V * rcp;
    b2 = W * rcp;
    return true;
}

// DOTS piece (geometry_processor.cpp:221-271): record = (start, r0), (end, r1), (s, 0), (t, 0).
// Triangle (face, tri): tri 0 = (start+v r0, end-v r1, end+v r1), tri 1 = (start+v r0,
// start-v r0, end-v r1) with v = s (face 0) or t (face 1).
__device__ __forceinline__ void dots_tri(const float4* P, int face, int tri, f3& v0, f3& v1, f3& v2) {
    f3 st = f3{P[0].x, P[0].y, P[0].z}, en = f3{P[1].x, P[1].y, P[1].z};
    float r0 = P[0].w, r1 = P[1].w;
    f3 v = face == 0 ? f3{P[2].x, P[2].y, P[2].z} : f3{P[3].x, P[3].y, P[3].z};
    f3 a = scr(v, r0), b = scr(v, r1);
    if (tri == 0) {
        v0 = add(st, a);
        v1 = sub(en, b);
        v2 = add(en, b);
    } else {
        v0 = add(st, a);
        v1 = sub(st, a);
        v2 = sub(en, b);
    }
}

}  // namespace vk
