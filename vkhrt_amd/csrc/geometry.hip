// Geometry front-end on the GPU (geometry_processor.cpp:45-67, 123-156, 201-298, 422-436)
// and the seeded synthetic strand generator. One thread per strand / line / curve.
// Compiled with -ffp-contract=off so results round exactly like the host definitions.
#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "geometry.hpp"

namespace vk {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ uint64_t splitmix_next(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ float uni(uint64_t& s) { return float(splitmix_next(s) >> 40) * (1.0f / 16777216.0f); }
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Synthetic hair: roots uniform on the upper hemisphere (y >= 0) of a sphere by rejection
// sampling of the cube (y folded to |y|); per step d <- normalize(d + (0,-gravity,0) +
// U[-curl,curl]^3), p += step * d; radius tapers linearly r_root -> r_tip.
// PRNG: SplitMix64 per strand starting at mix64(seed * C + strand) (hashed, so neighbouring
// strands do not share shifted streams); uniforms from the top 24 bits.
__global__ void k_strands(uint64_t seed, uint32_t nstrands, uint32_t nseg, StrandParams hp, float4* __restrict__ out) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nstrands) return;
    uint64_t st = mix64(seed * 0xD1B54A32D192ED03ull + uint64_t(s));
    f3 n;
    for (;;) {
        float x = uni(st) * 2.0f - 1.0f;
        float y = uni(st) * 2.0f - 1.0f;
        float z = uni(st) * 2.0f - 1.0f;
        y = fabsf(y);
        float l2 = x * x + y * y + z * z;
        if (l2 > 1.0f || l2 < 1e-4f) continue;
        float l = sqrtf(l2);
        n = f3{x / l, y / l, z / l};
        break;
    }
    f3 p = scr(n, hp.sphere_radius);
    f3 d = n;
    const f3 g = f3{0.0f, -hp.gravity, 0.0f};
    float4* o = out + size_t(s) * (nseg + 1);
    for (uint32_t k = 0; k <= nseg; ++k) {
        if (k > 0) {
            float nx = (uni(st) * 2.0f - 1.0f) * hp.curl;
            float ny = (uni(st) * 2.0f - 1.0f) * hp.curl;
            float nz = (uni(st) * 2.0f - 1.0f) * hp.curl;
            d = nrm3(add(add(d, g), f3{nx, ny, nz}));
            p = add(p, scr(d, hp.step));
        }
        float r = hp.r_root - (hp.r_root - hp.r_tip) * (float(k) / float(nseg));
        o[k] = make_float4(p.x, p.y, p.z, r);
    }
}

// GenerateLines (geometry_processor.cpp:45-67) over a strip index buffer: line k of strand s
// joins vertices k and k+1. Record: (start.xyz, r0), (end.xyz, r1).
__global__ void k_strand_lines(const float4* __restrict__ v, uint32_t nstrands, uint32_t nseg, float4* __restrict__ lines) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nstrands * nseg) return;
    const uint32_t s = i / nseg, k = i % nseg;
    lines[size_t(i) * 2] = v[size_t(s) * (nseg + 1) + k];
    lines[size_t(i) * 2 + 1] = v[size_t(s) * (nseg + 1) + k + 1];
}

__device__ __forceinline__ bool same(float4 a, float4 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

// GenerateCurves (geometry_processor.cpp:123-156), tension 1: cp1 = p1 + (p2 - p0) * (1/6),
// cp2 = p2 - (p3 - p1) * (1/6); neighbours used only when exactly connected (:140-141),
// the first and last lines duplicated (:129-133). Radii follow the same rule.
__global__ void k_curves(const float4* __restrict__ L, uint32_t n, float tension, float4* __restrict__ C) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 cs = L[size_t(i) * 2], ce = L[size_t(i) * 2 + 1];
    const uint32_t ip = i == 0 ? 0 : i - 1, in = i + 1 == n ? n - 1 : i + 1;
    const float4 ps = L[size_t(ip) * 2], pe = L[size_t(ip) * 2 + 1];
    const float4 ns = L[size_t(in) * 2], ne = L[size_t(in) * 2 + 1];
    const bool pc = same(pe, cs), nc = same(ce, ns);
    const float4 p0 = pc ? ps : cs, p1 = cs, p2 = ce, p3 = nc ? ne : ce;
    const float k = tension / 6.0f;
    float4 c1, c2;
    c1.x = p1.x + (p2.x - p0.x) * k;
    c1.y = p1.y + (p2.y - p0.y) * k;
    c1.z = p1.z + (p2.z - p0.z) * k;
    c1.w = p1.w + (p2.w - p0.w) * k;
    c2.x = p2.x - (p3.x - p1.x) * k;
    c2.y = p2.y - (p3.y - p1.y) * k;
    c2.z = p2.z - (p3.z - p1.z) * k;
    c2.w = p2.w - (p3.w - p1.w) * k;
    C[size_t(i) * 4 + 0] = p1;
    C[size_t(i) * 4 + 1] = c1;
    C[size_t(i) * 4 + 2] = c2;
    C[size_t(i) * 4 + 3] = p2;
}

__device__ __forceinline__ float4 bez4(const float4* c, float t) {
    float u = 1.0f - t;
    float tt = t * t;
    float uu = u * u;
    float uuu = uu * u;
    float ttt = tt * t;
    float w1 = 3.0f * uu * t, w2 = 3.0f * u * tt;
    float4 r;
    r.x = uuu * c[0].x + w1 * c[1].x + w2 * c[2].x + ttt * c[3].x;
    r.y = uuu * c[0].y + w1 * c[1].y + w2 * c[2].y + ttt * c[3].y;
    r.z = uuu * c[0].z + w1 * c[1].z + w2 * c[2].z + ttt * c[3].z;
    r.w = uuu * c[0].w + w1 * c[1].w + w2 * c[2].w + ttt * c[3].w;
    return r;
}

// Each cubic -> `pieces` lines sampled at t = k / pieces (Curve::Sample, model.cpp:20-30).
__global__ void k_curve_lines(const float4* __restrict__ C, uint32_t n, uint32_t pieces, float4* __restrict__ L) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * pieces) return;
    const uint32_t c = i / pieces, k = i % pieces;
    float4 cv[4] = {C[size_t(c) * 4], C[size_t(c) * 4 + 1], C[size_t(c) * 4 + 2], C[size_t(c) * 4 + 3]};
    const float t0 = float(k) / float(pieces), t1 = float(k + 1) / float(pieces);
    L[size_t(i) * 2] = bez4(cv, t0);
    L[size_t(i) * 2 + 1] = bez4(cv, t1);
}

// GenerateLinearSweptSpheres (geometry_processor.cpp:273-298): one capsule per line with
// radius max(r, 0.001) (:285). constant_radius >= 0 replaces the per-end radii.
__global__ void k_lss(const float4* __restrict__ L, uint32_t n, float constant_radius, float4* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 a = L[size_t(i) * 2], b = L[size_t(i) * 2 + 1];
    float r0 = constant_radius >= 0.0f ? constant_radius : a.w;
    float r1 = constant_radius >= 0.0f ? constant_radius : b.w;
    a.w = gmax(r0, 0.001f);
    b.w = gmax(r1, 0.001f);
    out[size_t(i) * 2] = a;
    out[size_t(i) * 2 + 1] = b;
}

// Branch-free perpendicular of M. M. Stark ("Efficient construction of perpendicular vectors
// without branching", Journal of Graphics, GPU, and Game Tools 14(1), 2009), the frame vector of
// the DOTS quads (PerpStark, geometry_processor.cpp:201-213): u crossed with the coordinate axis of
// u's smallest-magnitude component, x only if strictly smallest, then y if strictly below z, else
// z; normalized. The order of the tests fixes which axis a tie picks.
__device__ __forceinline__ f3 perp_stark(f3 u) {
    const float ax = fabsf(u.x), ay = fabsf(u.y), az = fabsf(u.z);
    const bool along_x = ax < ay && ax < az;
    const bool along_y = !along_x && ay < az;
    const f3 axis = f3{along_x ? 1.0f : 0.0f, along_y ? 1.0f : 0.0f, along_x || along_y ? 0.0f : 1.0f};
    return nrm3(cross3(u, axis));
}

// DOTS piece per line (geometry_processor.cpp:221-271): frame s = PerpStark(fwd),
// t = cross(fwd, s). Record (start, r0), (end, r1), (s, 0), (t, 0).
__global__ void k_dots(const float4* __restrict__ L, uint32_t n, float constant_radius, float4* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 a = L[size_t(i) * 2], b = L[size_t(i) * 2 + 1];
    f3 st = f3{a.x, a.y, a.z}, en = f3{b.x, b.y, b.z};
    f3 fwd = nrm3(sub(en, st));
    f3 s = perp_stark(fwd);
    f3 t = cross3(fwd, s);
    a.w = constant_radius >= 0.0f ? constant_radius : a.w;
    b.w = constant_radius >= 0.0f ? constant_radius : b.w;
    out[size_t(i) * 4] = a;
    out[size_t(i) * 4 + 1] = b;
    out[size_t(i) * 4 + 2] = make_float4(s.x, s.y, s.z, 0.0f);
    out[size_t(i) * 4 + 3] = make_float4(t.x, t.y, t.z, 0.0f);
}

// LSS BLAS records: start cap dropped (radius sign bit set) when chained to the previous
// capsule, i.e. previous end == this start exactly.
__global__ void k_lss_chain(const float4* __restrict__ in, uint32_t n, float4* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 a = in[size_t(i) * 2], b = in[size_t(i) * 2 + 1];
    bool chained = false;
    if (i > 0) {
        float4 pb = in[size_t(i - 1) * 2 + 1];
        chained = pb.x == a.x && pb.y == a.y && pb.z == a.z;
    }
    a.w = chained ? -fabsf(a.w) : fabsf(a.w);
    if (chained && a.w == 0.0f) a.w = -0.0f;
    out[size_t(i) * 2] = a;
    out[size_t(i) * 2 + 1] = b;
}

inline dim3 grid(uint64_t n) { return dim3(uint32_t((n + kBlock - 1) / kBlock)); }

// Indexed mesh -> triangle records: 3 x float4 vertex positions per triangle and its 3 vertex
// normals (the reference's Vertex is position, normal, texCoord = 8 floats; Indices are u32
// triples, triangle_closest_hit.rchit:19-29). An index past the vertex count is replaced by 0
// and counted in `bad`, so no read leaves the vertex buffer.
__global__ void k_tri_records(const float* __restrict__ verts, uint32_t nverts, const uint32_t* __restrict__ idx,
                              uint32_t ntri, float4* __restrict__ recs, float4* __restrict__ normals,
                              uint32_t* __restrict__ bad) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntri) return;
    for (int k = 0; k < 3; ++k) {
        uint32_t v = idx[size_t(t) * 3 + k];
        if (v >= nverts) {
            atomicAdd(bad, 1u);
            v = 0;
        }
        const float* p = verts + size_t(v) * 8;
        recs[size_t(t) * 3 + k] = make_float4(p[0], p[1], p[2], 0.0f);
        normals[size_t(t) * 3 + k] = make_float4(p[3], p[4], p[5], 0.0f);
    }
}

// GenerateLines (geometry_processor.cpp:45-67) for an indexed line mesh: line i joins the
// positions of vertices idx[2i] and idx[2i + 1], both indices mesh-local (the reference's
// firstVertex is added once, not again as at :61-62), with radius r at both ends. An index past
// the vertex count is replaced by 0 and counted in `bad`.
__global__ void k_mesh_lines(const float* __restrict__ verts, uint32_t nverts, const uint32_t* __restrict__ idx,
                             uint32_t n, float r, float4* __restrict__ lines, uint32_t* __restrict__ bad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t miss = 0;
    for (int k = 0; k < 2; ++k) {
        uint32_t v = idx[size_t(i) * 2 + k];
        if (v >= nverts) {
            ++miss;
            v = 0;
        }
        const float* p = verts + size_t(v) * 8;
        lines[size_t(i) * 2 + k] = make_float4(p[0], p[1], p[2], r);
    }
    if (miss) atomicAdd(bad, miss);
}

// Inverse of the leaf order: slot_of[prim_ids[s]] = s.
__global__ void k_slot_of(const uint32_t* __restrict__ prim_ids, uint32_t n, uint32_t* __restrict__ slot_of) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t p = prim_ids[s];
    if (p < n) slot_of[p] = s;
}

}  // namespace

hipError_t gen_slot_of(const uint32_t* prim_ids, uint32_t n, uint32_t* slot_of, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_slot_of, grid(n), dim3(kBlock), 0, st, prim_ids, n, slot_of);
    return hipGetLastError();
}

hipError_t gen_mesh_lines(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t n, float radius,
                          float4* lines, uint32_t* bad, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mesh_lines, grid(n), dim3(kBlock), 0, st, verts, nverts, idx, n, radius, lines, bad);
    return hipGetLastError();
}

hipError_t gen_tri_records(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t ntri, float4* recs,
                           float4* normals, uint32_t* bad, hipStream_t st) {
    if (ntri == 0) return hipSuccess;
    hipLaunchKernelGGL(k_tri_records, dim3((ntri + 255) / 256), dim3(256), 0, st, verts, nverts, idx, ntri, recs, normals,
                       bad);
    return hipGetLastError();
}

hipError_t gen_strands(uint64_t seed, uint32_t nstrands, uint32_t nseg, const StrandParams& hp, float4* out,
                       hipStream_t st) {
    if (nstrands == 0) return hipSuccess;
    hipLaunchKernelGGL(k_strands, grid(nstrands), dim3(kBlock), 0, st, seed, nstrands, nseg, hp, out);
    return hipGetLastError();
}
hipError_t gen_strand_lines(const float4* v, uint32_t nstrands, uint32_t nseg, float4* lines, hipStream_t st) {
    if (uint64_t(nstrands) * nseg == 0) return hipSuccess;
    hipLaunchKernelGGL(k_strand_lines, grid(uint64_t(nstrands) * nseg), dim3(kBlock), 0, st, v, nstrands, nseg, lines);
    return hipGetLastError();
}
hipError_t gen_curves(const float4* lines, uint32_t n, float tension, float4* curves, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_curves, grid(n), dim3(kBlock), 0, st, lines, n, tension, curves);
    return hipGetLastError();
}
hipError_t gen_curve_lines(const float4* curves, uint32_t n, uint32_t pieces, float4* lines, hipStream_t st) {
    if (uint64_t(n) * pieces == 0) return hipSuccess;
    hipLaunchKernelGGL(k_curve_lines, grid(uint64_t(n) * pieces), dim3(kBlock), 0, st, curves, n, pieces, lines);
    return hipGetLastError();
}
hipError_t gen_lss(const float4* lines, uint32_t n, float constant_radius, float4* caps, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_lss, grid(n), dim3(kBlock), 0, st, lines, n, constant_radius, caps);
    return hipGetLastError();
}
hipError_t gen_dots(const float4* lines, uint32_t n, float constant_radius, float4* pieces, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_dots, grid(n), dim3(kBlock), 0, st, lines, n, constant_radius, pieces);
    return hipGetLastError();
}
hipError_t lss_chain(const float4* caps, uint32_t n, float4* out, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_lss_chain, grid(n), dim3(kBlock), 0, st, caps, n, out);
    return hipGetLastError();
}

}  // namespace vk
