// Oracle (test infrastructure only; see oracle_math.hpp header): geometry processing,
// LSS / DOTS intersection, camera and the seeded strand generator.
#pragma once
#include <vector>

#include "oracle_math.hpp"

namespace orc {

// ---------------------------------------------------------------- LSS (round cone)
// Linear swept sphere between a (radius ra) and b (radius rb): the convex hull of the two
// spheres. The reference leaves LSS intersection to the NVIDIA RT core
// (triangle_closest_hit.rchit:43-58 only unpacks the hit), so this definition is the
// build's own: the ray's entry point into the hull (front-facing boundary point with the
// smallest t), accepted in [tmin, tmax]. With chained end caps the start sphere belongs to
// the previous piece of the strand and is left out (cap_a = false). The computation is
// done from the ray point closest to a (origin shift) to avoid cancellation.
// u = parameter of the touching sphere centre along a->b; normal = normalize(hit - mix(a,b,u))
// (triangle_closest_hit.rchit:50-55). Pinned by analytic capsule cases in tests.
template <class T>
struct LssOut {
    T t, u;
    V3<T> n;
};

template <class T>
inline bool IntersectLSS(V3<T> o, V3<T> d, V3<T> a, T ra, V3<T> b, T rb, bool cap_a, T tmin, T tmax, LssOut<T>& out) {
    const T dd = dot(d, d);
    // shift origin to the ray point nearest to a
    const T ta = dot(a - o, d) / dd;
    const V3<T> os = o + ta * d;
    const V3<T> w = os - a;  // small vector
    const V3<T> ba = b - a;
    const T l2 = dot(ba, ba);
    const T dr = ra - rb;
    const T k = l2 - dr * dr;
    const T wd = dot(w, d), ww = dot(w, w), wb = dot(w, ba), db = dot(d, ba);
    T best = T(INFINITY);
    int which = -1;  // 0 lateral, 1 cap a, 2 cap b
    T bestY = T(0);
    bool sphere_only = !(k > T(0));
    if (!sphere_only) {
        const T q = ra * l2 - dr * wb;
        const T A = k * (l2 * dd - db * db) - dr * dr * db * db;
        const T B = k * (l2 * wd - wb * db) + q * dr * db;
        const T C = k * (l2 * ww - wb * wb) - q * q;
        const T disc = B * B - A * C;
        if (disc >= T(0) && A != T(0)) {
            const T sq = std::sqrt(disc);
            T roots[2] = {(-B - sq) / A, (-B + sq) / A};
            for (int i = 0; i < 2; ++i) {
                T ts = roots[i];
                T Y = wb + ts * db;
                if (!(Y >= ra * dr && Y <= l2 + rb * dr)) continue;
                // outward normal ~ nx * radial_unit + ny * axis_unit; facing test uses its sign only
                V3<T> p = w + ts * d;
                V3<T> radial = p - (Y / l2) * ba;
                T rl = length(radial);
                T nx_l = std::sqrt(k);  // nx * l
                // n * l = nx_l * radial/|radial| + dr * ba/l ... compare dot(d, n) < 0
                T dn = nx_l * dot(d, radial) / rl + dr * db / std::sqrt(l2);
                if (!(dn < T(0))) continue;
                if (ts < best) {
                    best = ts;
                    which = 0;
                    bestY = Y;
                }
            }
        }
    }
    // sphere caps
    for (int s = 0; s < 2; ++s) {
        if (s == 0 && !cap_a && !sphere_only) continue;
        V3<T> c = s == 0 ? a : b;
        T r = s == 0 ? ra : rb;
        if (sphere_only) {
            // one sphere contains the other: only the larger one is the hull
            if ((s == 0) != (ra >= rb)) continue;
        }
        V3<T> wc = os - c;
        T bq = dot(wc, d);
        T cq = dot(wc, wc) - r * r;
        T disc = bq * bq - dd * cq;
        if (disc < T(0)) continue;
        T sq = std::sqrt(disc);
        T ts = (-bq - sq) / dd;  // entry root of the sphere
        V3<T> p = w + ts * d;
        T Y = dot(p, ba);
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
    T t = ta + best;
    if (!(t >= tmin && t <= tmax)) return false;
    T u;
    if (which == 0)
        u = (bestY - ra * dr) / k;
    else
        u = which == 1 ? T(0) : T(1);
    if (sphere_only) u = ra >= rb ? T(0) : T(1);
    out.t = t;
    out.u = u;
    V3<T> hit = w + best * d;  // hit - a
    V3<T> ctr = u * ba;        // mix(a, b, u) - a
    out.n = normalize(hit - ctr);
    return true;
}

// ---------------------------------------------------------------- DOTS triangles
// Watertight ray-triangle intersection (S. Woop, C. Benthin, I. Wald, "Watertight
// Ray/Triangle Intersection", JCGT 2(1), 2013), no facing cull
// (top_level_acceleration_structure.cpp:28), double-precision fallback for zero edge
// functions as in the paper. Barycentrics (b1, b2) are the weights of v1 and v2 as in
// triangle_closest_hit.rchit:33.
template <class T>
struct TriOut {
    T t, b1, b2;
};

inline int max_dim(V3<float> v) {
    float ax = std::fabs(v.x), ay = std::fabs(v.y), az = std::fabs(v.z);
    return ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
}
template <class T>
inline T comp(V3<T> v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// a*b - c*d in FP32 by Kahan's algorithm with fused multiply-adds (relative error <= 2u,
// Jeannerod, Louvet and Muller, Math. Comp. 2013): w = c*d rounded, e = w - c*d exactly,
// f = a*b - w rounded once, result f + e. A nonzero exact value keeps its sign and an exact
// zero stays zero, which is what the watertight test's re-evaluation of a zero edge function
// needs (the paper re-evaluates in double; this build keeps the triangle test in FP32).
inline float Det2Kahan(float a, float b, float c, float d) {
    const float w = c * d;
    const float e = std::fma(-c, d, w);
    const float f = std::fma(a, b, -w);
    return f + e;
}

// Watertight ray/triangle (Woop, Benthin, Wald, JCGT 2013), no facing cull; zero edge
// functions are re-evaluated by Det2Kahan.
inline bool IntersectTriangle(V3<float> o, V3<float> d, V3<float> v0, V3<float> v1, V3<float> v2, float tmin,
                              float tmax, TriOut<float>& out) {
    int kz = max_dim(d);
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    if (comp(d, kz) < 0.0f) {
        int tmp = kx;
        kx = ky;
        ky = tmp;
    }
    float Sx = comp(d, kx) / comp(d, kz);
    float Sy = comp(d, ky) / comp(d, kz);
    float Sz = 1.0f / comp(d, kz);
    V3<float> A = v0 - o, B = v1 - o, C = v2 - o;
    float Ax = comp(A, kx) - Sx * comp(A, kz);
    float Ay = comp(A, ky) - Sy * comp(A, kz);
    float Bx = comp(B, kx) - Sx * comp(B, kz);
    float By = comp(B, ky) - Sy * comp(B, kz);
    float Cx = comp(C, kx) - Sx * comp(C, kz);
    float Cy = comp(C, ky) - Sy * comp(C, kz);
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = Det2Kahan(Cx, By, Cy, Bx);
        V = Det2Kahan(Ax, Cy, Ay, Cx);
        W = Det2Kahan(Bx, Ay, By, Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = Sz * comp(A, kz);
    float Bz = Sz * comp(B, kz);
    float Cz = Sz * comp(C, kz);
    float Tt = U * Az + V * Bz + W * Cz;
    float rcp = 1.0f / det;
    float t = Tt * rcp;
    if (!(t >= tmin && t <= tmax)) return false;
    out.t = t;
    out.b1 = V * rcp;
    out.b2 = W * rcp;
    return true;
}

// geometry_processor.cpp:201-213 PerpStark and :216-219 BuildFrame: Stark's branch-free
// perpendicular (JGT 14(1), 2009), u crossed with the axis of its smallest-magnitude component
// (x only if strictly smallest, then y if strictly below z, else z), normalized. For IEEE floats
// with gradual underflow, a < b decides exactly as the reference's (a - b) < 0.
inline V3<float> PerpStark(V3<float> u) {
    const float ax = std::fabs(u.x), ay = std::fabs(u.y), az = std::fabs(u.z);
    const bool along_x = ax < ay && ax < az;
    const bool along_y = !along_x && ay < az;
    return normalize(cross(u, mk(along_x ? 1.0f : 0.0f, along_y ? 1.0f : 0.0f, along_x || along_y ? 0.0f : 1.0f)));
}

// DOTS piece: two orthogonal quads through the line (geometry_processor.cpp:221-271).
// Quad f in {0,1} uses frame vector v[f] (v[0] = PerpStark(fwd), v[1] = cross(fwd, v[0]));
// triangles (start+v r0, end-v r1, end+v r1) and (start+v r0, start-v r0, end-v r1).
// The reference uses one radius for both ends; per-end radii are the tapered extension.
struct DotsPiece {
    V3<float> start, end, s, t;
    float r0, r1;
};
inline DotsPiece MakeDotsPiece(V3<float> start, V3<float> end, float r0, float r1) {
    DotsPiece p;
    p.start = start;
    p.end = end;
    p.r0 = r0;
    p.r1 = r1;
    V3<float> fwd = normalize(end - start);
    p.s = PerpStark(fwd);
    p.t = cross(fwd, p.s);
    return p;
}
inline void DotsTriangle(const DotsPiece& p, int face, int tri, V3<float>& v0, V3<float>& v1, V3<float>& v2) {
    V3<float> v = face == 0 ? p.s : p.t;
    if (tri == 0) {
        v0 = p.start + v * p.r0;
        v1 = p.end - v * p.r1;
        v2 = p.end + v * p.r1;
    } else {
        v0 = p.start + v * p.r0;
        v1 = p.start - v * p.r0;
        v2 = p.end - v * p.r1;
    }
}

// ---------------------------------------------------------------- geometry processing
struct LineR {
    V3<float> start, end;
    float r0, r1;
};

// geometry_processor.cpp:123-156 GenerateCurves (Catmull-Rom -> Bezier, tension 1).
// Neighbours are reused only when exactly connected (end == start, :140-141). Radii follow
// the same rule (extension). Output per curve: 4 x (xyz, r).
inline void GenerateCurves(const std::vector<LineR>& lines, std::vector<float>& out, float tension = 1.0f) {
    size_t n = lines.size();
    out.assign(n * 16, 0.0f);
    if (n == 0) return;
    auto eq = [](V3<float> a, V3<float> b) { return a.x == b.x && a.y == b.y && a.z == b.z; };
    const float k = tension / 6.0f;
    for (size_t i = 0; i < n; ++i) {
        // linesToProcess = [first] + lines + [last]; index j = i + 1
        const LineR& cur = lines[i];
        const LineR& prv = i == 0 ? lines[0] : lines[i - 1];
        const LineR& nxt = i + 1 == n ? lines[n - 1] : lines[i + 1];
        const LineR& P = eq(prv.end, cur.start) ? prv : cur;
        const LineR& N = eq(cur.end, nxt.start) ? nxt : cur;
        V3<float> p0 = P.start, p1 = cur.start, p2 = cur.end, p3 = N.end;
        float q0 = P.r0, q1 = cur.r0, q2 = cur.r1, q3 = N.r1;
        V3<float> c1 = p1 + (p2 - p0) * k;
        V3<float> c2 = p2 - (p3 - p1) * k;
        float rc1 = q1 + (q2 - q0) * k;
        float rc2 = q2 - (q3 - q1) * k;
        float* o = &out[i * 16];
        float vals[16] = {p1.x, p1.y, p1.z, q1, c1.x, c1.y, c1.z, rc1, c2.x, c2.y, c2.z, rc2, p2.x, p2.y, p2.z, q2};
        std::memcpy(o, vals, sizeof(vals));
    }
}

// geometry_processor.cpp:422-436 GenerateAABBs with the pad radius. The per-point
// radius extension pads by the largest control-point radius.
inline void CurveAABB(const float* c, bool per_point, float radius, float mn[3], float mx[3]) {
    float r = radius;
    if (per_point) r = gmax(gmax(c[3], c[7]), gmax(c[11], c[15]));
    for (int k = 0; k < 3; ++k) {
        mn[k] = gmin(gmin(c[k], c[12 + k]), gmin(c[4 + k], c[8 + k])) - r;
        mx[k] = gmax(gmax(c[k], c[12 + k]), gmax(c[4 + k], c[8 + k])) + r;
    }
}

// Split each cubic into `pieces` linear pieces sampled at t = k/pieces with the Bernstein
// form of Curve::Sample (model.cpp:20-30); radii sampled the same way.
inline void CurvesToLines(const std::vector<float>& curves, int pieces, std::vector<LineR>& lines) {
    size_t n = curves.size() / 16;
    lines.resize(n * pieces);
    for (size_t i = 0; i < n; ++i) {
        const float* c = &curves[i * 16];
        Curve<float> cv{mk(c[0], c[1], c[2]), mk(c[4], c[5], c[6]), mk(c[8], c[9], c[10]), mk(c[12], c[13], c[14])};
        float rr[4] = {c[3], c[7], c[11], c[15]};
        for (int k = 0; k < pieces; ++k) {
            float t0 = float(k) / float(pieces), t1 = float(k + 1) / float(pieces);
            LineR& L = lines[i * pieces + k];
            L.start = SampleCurvePoint(cv, t0);
            L.end = SampleCurvePoint(cv, t1);
            L.r0 = SampleCurveScalar(rr, t0);
            L.r1 = SampleCurveScalar(rr, t1);
        }
    }
}

// ---------------------------------------------------------------- camera
// fly_camera.cpp:25-35 ViewMatrix = glm::lookAt(eye, eye+front, up) (right-handed) and
// ProjectionMatrix = perspectiveRH_ZO(fov, aspect, n, f) with [1][1] *= -1;
// renderer.cpp:189-195 uploads their inverses.
inline M4<float> LookAtRH(V3<float> eye, V3<float> center, V3<float> up) {
    V3<float> f = normalize(center - eye);
    V3<float> s = normalize(cross(f, up));
    V3<float> u = cross(s, f);
    M4<float> R{};
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) R.m[c][r] = c == r ? 1.0f : 0.0f;
    R.m[0][0] = s.x;
    R.m[1][0] = s.y;
    R.m[2][0] = s.z;
    R.m[0][1] = u.x;
    R.m[1][1] = u.y;
    R.m[2][1] = u.z;
    R.m[0][2] = -f.x;
    R.m[1][2] = -f.y;
    R.m[2][2] = -f.z;
    R.m[3][0] = -dot(s, eye);
    R.m[3][1] = -dot(u, eye);
    R.m[3][2] = dot(f, eye);
    return R;
}
inline M4<float> PerspectiveRH_ZO_FlipY(float tanHalfFovy, float aspect, float zNear, float zFar) {
    M4<float> R{};
    R.m[0][0] = 1.0f / (aspect * tanHalfFovy);
    R.m[1][1] = 1.0f / tanHalfFovy;
    R.m[2][2] = zFar / (zNear - zFar);
    R.m[2][3] = -1.0f;
    R.m[3][2] = -(zFar * zNear) / (zFar - zNear);
    R.m[1][1] *= -1.0f;
    return R;
}

// ray_gen.rgen:18-28 primary ray for launch id (x, y) of (W, H).
inline void RayGen(const M4<float>& viewInv, const M4<float>& projInv, uint32_t x, uint32_t y, uint32_t W, uint32_t H,
                   V3<float>& o, V3<float>& d) {
    float pcx = float(x) + 0.5f, pcy = float(y) + 0.5f;
    float ux = pcx / float(W), uy = pcy / float(H);
    float dx = ux * 2.0f - 1.0f, dy = uy * 2.0f - 1.0f;
    float v0[4] = {0, 0, 0, 1}, org[4];
    mul4(viewInv, v0, org);
    float v1[4] = {dx, dy, 1, 1}, tgt[4];
    mul4(projInv, v1, tgt);
    V3<float> tn = normalize(mk(tgt[0], tgt[1], tgt[2]));
    float v2[4] = {tn.x, tn.y, tn.z, 0}, dir[4];
    mul4(viewInv, v2, dir);
    o = mk(org[0], org[1], org[2]);
    d = mk(dir[0], dir[1], dir[2]);
}

// ---------------------------------------------------------------- seeded strand generator
// Frozen synthetic hair (no reference counterpart; the reference loads assets):
// roots uniform on the upper hemisphere (y >= 0) of a radius-R sphere by rejection
// sampling of the cube; polyline of nseg+1 vertices with step `step`; per step
// d <- normalize(d + gravity*(0,-1,0) + U[-curl, curl]^3) then p += step*d; radius tapers
// linearly r_root -> r_tip. PRNG: SplitMix64 per strand, uniforms from the top 24 bits.
// Each strand's stream starts at mix64(seed * C + strand): hashed starting points, so the
// streams of neighbouring strands are not shifted copies of each other.
inline uint64_t Mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    float uniform() { return float(next() >> 40) * (1.0f / 16777216.0f); }
};

struct HairParams {
    float sphere_radius = 100.0f, step = 1.5f, gravity = 0.3f, curl = 0.2f, r_root = 0.04f, r_tip = 0.01f;
};

// Writes (nseg+1) vertices of (x, y, z, r) for strand `strand`.
inline void GenerateStrand(uint64_t seed, uint64_t strand, int nseg, const HairParams& hp, float* out) {
    SplitMix64 rng{Mix64(seed * 0xD1B54A32D192ED03ULL + strand)};
    V3<float> n;
    for (;;) {
        float x = rng.uniform() * 2.0f - 1.0f;
        float y = rng.uniform() * 2.0f - 1.0f;
        float z = rng.uniform() * 2.0f - 1.0f;
        y = std::fabs(y);
        float l2 = x * x + y * y + z * z;
        if (l2 > 1.0f || l2 < 1e-4f) continue;
        float l = std::sqrt(l2);
        n = mk(x / l, y / l, z / l);
        break;
    }
    V3<float> p = n * hp.sphere_radius;
    V3<float> d = n;
    const V3<float> g = mk(0.0f, -hp.gravity, 0.0f);
    for (int k = 0; k <= nseg; ++k) {
        if (k > 0) {
            float nx = (rng.uniform() * 2.0f - 1.0f) * hp.curl;
            float ny = (rng.uniform() * 2.0f - 1.0f) * hp.curl;
            float nz = (rng.uniform() * 2.0f - 1.0f) * hp.curl;
            d = normalize(d + g + mk(nx, ny, nz));
            p = p + d * hp.step;
        }
        float r = hp.r_root - (hp.r_root - hp.r_tip) * (float(k) / float(nseg));
        out[k * 4 + 0] = p.x;
        out[k * 4 + 1] = p.y;
        out[k * 4 + 2] = p.z;
        out[k * 4 + 3] = r;
    }
}

}  // namespace orc
