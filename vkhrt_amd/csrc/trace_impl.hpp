// Traversal and shading kernels for the traceRayEXT path (ray_gen.rgen:16-48,
// hair_intersection.rint:132-150, hair_closest_hit.rchit, triangle_closest_hit.rchit,
// shading.glsl, debug.glsl).
//
// One ray per lane. The 8-wide BVH is walked front-to-back with node groups (child base,
// internal-child mask, remaining hit bits permuted by the ray octant) kept on a short
// per-lane stack in LDS (sized from the tree depth by the host). Closest-hit semantics: the
// smallest t in [tMin, tMax]; equal t resolve to the lower (instance, primitive). Shadow
// rays stop at the first accepted hit.
//
// Every kernel is specialized at compile time on the geometry variant (mode, radius model,
// FP64 polish), the any-hit flag and the LDS stack depth, so it carries no runtime mode
// branches and makes no device-function calls. The variants are instantiated in
// trace_*.hip. Compiled with -ffp-contract=off: intersector arithmetic rounds like the GLSL.
#pragma once
#include <hip/hip_runtime.h>

#include "build.hpp"
#include "intersect.hpp"
#include "mlp.hpp"

namespace vk {

// Geometry variant of a traversal kernel.
template <int MODE_, int RMODE_, bool POLISH_>
struct Geo {
    static constexpr bool kMixed = false;
    static constexpr int kMode = MODE_, kRMode = RMODE_;
    static constexpr bool kPolish = POLISH_;
};
// Instances of different modes in one TLAS: runtime dispatch per instance.
struct GeoMixed {
    static constexpr bool kMixed = true;
    static constexpr int kMode = -1, kRMode = 0;
    static constexpr bool kPolish = false;
};

namespace trace_detail {
namespace {

constexpr int kTraceBlock = 256;

struct Best {
    float t;
    uint32_t inst;
    uint32_t pid;  // original primitive id (DOTS: triangle id)
    uint32_t slot; // leaf slot (phantom/LSS) or piece slot (DOTS)
    float u, a, b, c;  // mode-specific hit data
    uint32_t extra;    // polished flag / DOTS triangle sub-index
};

// Merge key of a hit: t (a non-negative float, so its bits order like the value), then the
// hit's global primitive id gbase(instance) + primitive. Global ids order like (instance,
// primitive) pairs, so equal t resolve to the lower pair. 0xFFFFFFFF (above every global id;
// the TLAS build checks that) marks no hit.
__device__ __forceinline__ unsigned long long pack_key(float t, uint32_t gid) {
    return (static_cast<unsigned long long>(__float_as_uint(t)) << 32) | gid;
}
__device__ __forceinline__ unsigned long long best_key(const TraceParams& tp, const Best& b) {
    return (static_cast<unsigned long long>(__float_as_uint(b.t)) << 32) |
           (b.inst == 0xFFFFFFFFu ? 0xFFFFFFFFu : tp.inst[b.inst].blas.gbase + b.pid);
}
// Instance of a global primitive id: the last instance whose gbase is <= gid.
__device__ __forceinline__ uint32_t gid_instance(const InstanceDev* inst, uint32_t ninst, uint32_t gid) {
    uint32_t lo = 0, hi = ninst;  // gbase(lo) <= gid < gbase(hi)
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (inst[mid].blas.gbase <= gid)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__device__ __forceinline__ bool better(float t, uint32_t inst, uint32_t pid, const Best& B) {
    if (t < B.t) return true;
    if (t == B.t) return inst < B.inst || (inst == B.inst && pid < B.pid);
    return false;
}

// 1/d with zero components replaced by +-1e-20: the quantized slab form q*(s*inv) + (p-o)*inv
// would otherwise produce inf - inf = NaN for axis-parallel rays and drop that slab.
__device__ __forceinline__ float safe_rcp(float x) { return 1.0f / (fabsf(x) < 1e-20f ? copysignf(1e-20f, x) : x); }
__device__ __forceinline__ f3 safe_inv(f3 d) { return f3{safe_rcp(d.x), safe_rcp(d.y), safe_rcp(d.z)}; }

// The same address as a global-memory pointer: loads through it are global_load, not flat
// loads (a flat load also counts against the LDS counter, so waiting for an LDS access would
// wait for it too). The BLAS arrays are device allocations.
// (The host pass only checks these device bodies; it keeps the generic pointer.)
template <class T>
__device__ __forceinline__ auto gmem(const T* p) {
#if defined(__HIP_DEVICE_COMPILE__)
    return (const __attribute__((address_space(1))) T*)(p);
#else
    return p;
#endif
}

__device__ __forceinline__ float qbyte(uint32_t lo, uint32_t hi, int s) {
    uint32_t w = s < 4 ? lo : hi;
    return float((w >> (8 * (s & 3))) & 0xffu);
}

// Per-lane stack of node groups in LDS (stride kTraceBlock entries). A group is pushed only
// when descending one level, so a tree of depth D never needs more than D entries; the host
// picks STACK >= depth and a push past STACK reports an error instead of spilling.
// group: x = child_base, y = imask | (remaining hit bits in octant order) << 8
// Traverses one BLAS. Returns false on stack overflow.
// Per-ray traversal statistics (nodes visited, primitive records tested).
struct TravStats {
    uint32_t nodes = 0, prims = 0;
};

enum StepResult { kStepMore = 0, kStepDone = 1, kStepHit = 2, kStepError = 3 };

// Resumable traversal of one BLAS by one lane: begin() sets up the ray, each step() visits
// one 8-wide node (and tests its leaf primitives) and selects the next node. Used whole by
// traverse() and one step per iteration by the persistent kernels.
// Tests the 8 children of one node against [tmin, tcur]: `hits` gets the internal children
// hit (bit j ^ octant for slot j, so lowest-bit-first is near-to-far), `pmask` the leaf
// records hit (relative to the node's prim base).
__device__ __forceinline__ void node_visit(const WideNode* nodes, uint32_t node, f3 o, f3 inv, uint32_t oct,
                                           float tmin, float tcur, uint32_t& hits, uint32_t& pmask,
                                           uint32_t& child_base, uint32_t& imask, uint32_t& pbase) {
    const bool nx = oct & 1u, ny = oct & 2u, nz = oct & 4u;
    const auto* np = gmem(reinterpret_cast<const uint4*>(nodes + node));
    const uint4 n0 = np[0], n1 = np[1], n2 = np[2], n3 = np[3], n4 = np[4];
    const float sx = __uint_as_float((n0.w & 0xffu) << 23) * inv.x;
    const float sy = __uint_as_float(((n0.w >> 8) & 0xffu) << 23) * inv.y;
    const float sz = __uint_as_float(((n0.w >> 16) & 0xffu) << 23) * inv.z;
    const float ox = (__uint_as_float(n0.x) - o.x) * inv.x;
    const float oy = (__uint_as_float(n0.y) - o.y) * inv.y;
    const float oz = (__uint_as_float(n0.z) - o.z) * inv.z;
    imask = n0.w >> 24;
    child_base = n1.x;
    pbase = n1.y;
    // near/far plane bytes per axis by ray direction sign
    const uint32_t nxl = nx ? n3.z : n2.x, nxh = nx ? n3.w : n2.y;
    const uint32_t fxl = nx ? n2.x : n3.z, fxh = nx ? n2.y : n3.w;
    const uint32_t nyl = ny ? n4.x : n2.z, nyh = ny ? n4.y : n2.w;
    const uint32_t fyl = ny ? n2.z : n4.x, fyh = ny ? n2.w : n4.y;
    const uint32_t nzl = nz ? n4.z : n3.x, nzh = nz ? n4.w : n3.y;
    const uint32_t fzl = nz ? n3.x : n4.z, fzh = nz ? n3.y : n4.w;
    // The slab test only has to be conservative (the quantized boxes round outward and tf is
    // widened by 8 ulp), so it uses fused multiply-adds. Each slot contributes one bit of a
    // slot-order hit mask; the metadata is decoded afterwards for the slots hit only.
    uint32_t lm = 0;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const uint32_t meta = ((s < 4 ? n1.z : n1.w) >> (8 * (s & 3))) & 0xffu;
        const float t0x = fmaf(qbyte(nxl, nxh, s), sx, ox), t1x = fmaf(qbyte(fxl, fxh, s), sx, ox);
        const float t0y = fmaf(qbyte(nyl, nyh, s), sy, oy), t1y = fmaf(qbyte(fyl, fyh, s), sy, oy);
        const float t0z = fmaf(qbyte(nzl, nzh, s), sz, oz), t1z = fmaf(qbyte(fzl, fzh, s), sz, oz);
        const float tn = fmaxf(fmaxf(t0x, t0y), fmaxf(t0z, tmin));
        const float tf = fminf(fminf(t1x, t1y), fminf(t1z, tcur)) * 1.00000095f;
        lm |= (meta != 0 && tn <= tf) ? (1u << s) : 0u;
    }
    // internal children to octant order: bit j of `hits` is slot j ^ oct, so the lowest set
    // bit is the nearest child; three conditional swaps apply the XOR to the bit indices
    uint32_t x = lm & imask;
    x = (oct & 1u) ? (((x & 0x55u) << 1) | ((x >> 1) & 0x55u)) : x;
    x = (oct & 2u) ? (((x & 0x33u) << 2) | ((x >> 2) & 0x33u)) : x;
    x = (oct & 4u) ? (((x & 0x0Fu) << 4) | ((x >> 4) & 0x0Fu)) : x;
    hits = x;
    // leaf children hit: their record ranges (count << 5 | first in the metadata byte)
    pmask = 0;
    for (uint32_t lf = lm & ~imask; lf; lf &= lf - 1u) {
        const uint32_t s = uint32_t(__ffs(lf) - 1);
        const uint32_t meta = ((s < 4u ? n1.z : n1.w) >> (8u * (s & 3u))) & 0xffu;
        pmask |= ((1u << (meta >> 5)) - 1u) << (meta & 31u);
    }
}

template <int MODE, int RMODE, bool POLISH, bool ANY, int STACK, bool STATS = false>
struct Walker {
    f3 o, d, inv;
    uint32_t oct;
    RayFrame F;
    TriRay TR;
    uint2 cur;  // current group: x = child_base, y = imask | remaining hit bits (octant order) << 8
    uint32_t node;
    int sp;
    bool have_node;

    __device__ __forceinline__ void begin(f3 o_, f3 d_) {
        o = o_;
        d = d_;
        inv = safe_inv(d);
        oct = uint32_t(signbit(d.x)) | (uint32_t(signbit(d.y)) << 1) | (uint32_t(signbit(d.z)) << 2);
        if (MODE == kPhantom) make_frame(o, d, F);
        if (MODE == kDots || MODE == kTriangles) TR = make_tri_ray(d);
        cur = make_uint2(0u, 0u);
        node = 0;
        sp = 0;
        have_node = true;
    }

    __device__ __forceinline__ int step(const BlasDev& B, float tmin, uint32_t inst, Best& best, uint2* lds,
                                        TravStats* st, const TraceParams& tp, uint32_t ray_id) {
        if (have_node) {
            if (node >= B.nnodes) return kStepError;  // malformed structure: report, never read out of bounds
            if (STATS) ++st->nodes;
            uint32_t hits, pmask, child_base, imask, pbase;
            node_visit(B.nodes, node, o, inv, oct, tmin, best.t, hits, pmask, child_base, imask, pbase);
            // leaf primitives of this node
            if (pmask && pbase + (32u - __clz(pmask)) > B.nprims) return kStepError;
            while (pmask) {
                const uint32_t k = __ffs(pmask) - 1;
                pmask &= pmask - 1;
                const uint32_t slot = pbase + k;
                if (STATS) ++st->prims;
                if (MODE == kPhantom) {
                    const float4* pr = B.prims + size_t(slot) * 4;
                    float4 cv[4] = {pr[0], pr[1], pr[2], pr[3]};
                    PhantomHit ph;
                    const bool hit = prhi<RMODE, POLISH ? kPolishDefer : kPolishOff>(F, cv, B.aux[slot], B.radius, ph);
                    if (POLISH && ph.deferred) {
                        // this curve needs the FP64 continuation: queue (ray, instance, slot)
                        const uint32_t q = atomicAdd(tp.defer_count, 1u);
                        if (q < tp.defer_cap)
                            tp.defer[q] = make_uint4(ray_id, inst, slot, 0u);
                        else
                            atomicAdd(&tp.status[1], 1u);
                    }
                    if (hit && ph.t >= tmin && ph.t <= best.t) {
                        const uint32_t pid = B.prim_ids[slot];
                        if (better(ph.t, inst, pid, best)) {
                            best.t = ph.t;
                            best.inst = inst;
                            best.pid = pid;
                            best.slot = slot;
                            best.u = ph.u;
                            best.a = ph.ncx;
                            best.b = ph.ncy;
                            best.c = ph.nsz;
                            best.extra = ph.polished ? 1u : 0u;
                            if (ANY) return kStepHit;
                        }
                    }
                } else if (MODE == kLss) {
                    const float4* pr = B.prims + size_t(slot) * 2;
                    const float4 a = pr[0], b = pr[1];
                    LssHit lh;
                    if (lss_intersect(o, d, f3{a.x, a.y, a.z}, fabsf(a.w), f3{b.x, b.y, b.z}, b.w, !signbit(a.w), tmin,
                                      best.t, lh)) {
                        const uint32_t pid = B.prim_ids[slot];
                        if (better(lh.t, inst, pid, best)) {
                            best.t = lh.t;
                            best.inst = inst;
                            best.pid = pid;
                            best.slot = slot;
                            best.u = lh.u;
                            best.a = lh.n.x;
                            best.b = lh.n.y;
                            best.c = lh.n.z;
                            best.extra = 0;
                            if (ANY) return kStepHit;
                        }
                    }
                } else if (MODE == kTriangles) {
                    const float4* pr = B.prims + size_t(slot) * 3;
                    const float4 a = pr[0], b = pr[1], c = pr[2];
                    float t, b1, b2;
                    if (tri_intersect(TR, o, f3{a.x, a.y, a.z}, f3{b.x, b.y, b.z}, f3{c.x, c.y, c.z}, tmin, best.t, t,
                                      b1, b2)) {
                        const uint32_t pid = B.prim_ids[slot];
                        if (better(t, inst, pid, best)) {
                            best.t = t;
                            best.inst = inst;
                            best.pid = pid;
                            best.slot = slot;
                            best.u = b1;
                            best.a = b2;
                            best.extra = 0;
                            if (ANY) return kStepHit;
                        }
                    }
                } else {
                    const float4* pr = B.prims + size_t(slot) * 4;
                    float4 P[4] = {pr[0], pr[1], pr[2], pr[3]};
                    uint32_t piece = 0xFFFFFFFFu;
#pragma unroll
                    for (int tri = 0; tri < 4; ++tri) {
                        f3 v0, v1, v2;
                        dots_tri(P, tri >> 1, tri & 1, v0, v1, v2);
                        float t, b1, b2;
                        if (tri_intersect(TR, o, v0, v1, v2, tmin, best.t, t, b1, b2)) {
                            if (piece == 0xFFFFFFFFu) piece = B.prim_ids[slot];
                            const uint32_t pid = piece * 4u + uint32_t(tri);
                            if (better(t, inst, pid, best)) {
                                best.t = t;
                                best.inst = inst;
                                best.pid = pid;
                                best.slot = slot;
                                best.u = b1;
                                best.a = b2;
                                best.extra = uint32_t(tri);
                                if (ANY) return kStepHit;
                            }
                        }
                    }
                }
            }
            have_node = false;
            if (hits) {
                const uint2 g = make_uint2(child_base, imask | (hits << 8));
                if (cur.y >> 8) {
                    if (sp >= STACK) return kStepError;
                    lds[sp * kTraceBlock] = cur;
                    ++sp;
                }
                cur = g;
            }
        }
        if ((cur.y >> 8) == 0) {
            if (sp == 0) return kStepDone;
            --sp;
            cur = lds[sp * kTraceBlock];
        }
        // next child of the current group
        uint32_t hb = cur.y >> 8;
        const uint32_t j = __ffs(hb) - 1;
        hb &= hb - 1;
        const uint32_t slot = j ^ oct;
        const uint32_t im = cur.y & 0xffu;
        node = cur.x + __popc(im & ((1u << slot) - 1u));
        cur.y = im | (hb << 8);
        have_node = true;
        return kStepMore;
    }
};

// Whole traversal of one BLAS. Returns false on a malformed structure or stack overflow.
template <int MODE, int RMODE, bool POLISH, bool ANY, int STACK, bool STATS = false>
__device__ __forceinline__ bool traverse(const BlasDev& B, f3 o, f3 d, float tmin, uint32_t inst, Best& best,
                                         uint2* lds, bool& done, const TraceParams& tp, uint32_t ray_id,
                                         TravStats* st = nullptr) {
    if (B.nprims == 0) return true;
    Walker<MODE, RMODE, POLISH, ANY, STACK, STATS> W;
    W.begin(o, d);
    for (;;) {
        const int r = W.step(B, tmin, inst, best, lds, st, tp, ray_id);
        if (r == kStepMore) continue;
        if (r == kStepHit) done = true;
        return r != kStepError;
    }
}

__device__ __forceinline__ f3 xform_pt(const float* m /*col-major 4x4*/, f3 p) {
    return f3{m[0] * p.x + m[4] * p.y + m[8] * p.z + m[12] * 1.0f, m[1] * p.x + m[5] * p.y + m[9] * p.z + m[13] * 1.0f,
              m[2] * p.x + m[6] * p.y + m[10] * p.z + m[14] * 1.0f};
}
__device__ __forceinline__ f3 xform_dir(const float* m, f3 p) {
    return f3{m[0] * p.x + m[4] * p.y + m[8] * p.z + m[12] * 0.0f, m[1] * p.x + m[5] * p.y + m[9] * p.z + m[13] * 0.0f,
              m[2] * p.x + m[6] * p.y + m[10] * p.z + m[14] * 0.0f};
}

__device__ __forceinline__ bool ray_box(f3 o, f3 d, const float* lo, const float* hi, float tmin, float tmax) {
    float t0 = tmin, t1 = tmax;
    const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float inv = safe_rcp(dd[k]);
        float a = (lo[k] - oo[k]) * inv, b = (hi[k] - oo[k]) * inv;
        t0 = fmaxf(t0, fminf(a, b));
        t1 = fminf(t1, fmaxf(a, b));
    }
    return t0 <= t1 * 1.00000095f;
}

// Traces one world-space ray through all instances. Fills `best`; returns false on overflow.
// Voxel instances (BLASType::eVoxels): a 3D DDA (Amanatides, Woo, Eurographics 1987) over the
// occupancy grid from the ray's entry into the grid box. The hit is the first filled cell the
// walk enters within [tmin, tcur], at its entry t (never before the box entry), with the normal
// of the face crossed; a ray that starts inside a filled cell takes the face across its major
// axis. Cell boundaries are origin + i * size in float, like the reference's voxel boxes
// (GetVoxelWorldPosition, geometry_processor.cpp:438-449); the primitive id is the cell's
// linear index (its AABB index there, gl_PrimitiveID in voxel_closest_hit.rchit:20).
// Empty space is crossed a 16x16x16 block or a 4x4x4 brick at a time, with the result of the
// cell-by-cell walk. That walk visits the face crossings of the three axes in the order of the
// key (t of the face, axis), each axis's t being non-decreasing along its steps: it leaves an
// aligned S-cell block at the least key among the block's three exit faces, and each other axis
// has by then taken exactly its crossings with smaller keys (at most S - 1 inside the block). An
// empty block (its bit clear) is left in one move to that state: the cell, the entry t and the
// face are the walk's own.
template <bool ANY>
__device__ __forceinline__ void voxel_walk(const BlasDev& B, f3 o, f3 d, float tmin, uint32_t inst, Best& best,
                                           bool& done) {
    const int R[3] = {B.vres[0], B.vres[1], B.vres[2]};
    if (R[0] <= 0 || R[1] <= 0 || R[2] <= 0) return;
    const float O[3] = {B.vorigin[0], B.vorigin[1], B.vorigin[2]};
    const float c = B.vcell;
    const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
    float inv[3];
    float t0 = tmin, t1 = best.t;
    int ea = -1;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float hi = O[k] + float(R[k]) * c;
        if (dd[k] == 0.0f) {
            if (oo[k] < O[k] || oo[k] > hi) return;
            inv[k] = 0.0f;
            continue;
        }
        inv[k] = 1.0f / dd[k];
        const float ta = (O[k] - oo[k]) * inv[k], tb = (hi - oo[k]) * inv[k];
        const float lo = ta < tb ? ta : tb, up = ta < tb ? tb : ta;
        if (lo > t0) {
            t0 = lo;
            ea = k;
        }
        if (up < t1) t1 = up;
    }
    if (!(t0 <= t1)) return;
    int i[3], step[3];
    float tn[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float f = floorf((oo[k] + dd[k] * t0 - O[k]) / c);
        i[k] = f < 0.0f ? 0 : f > float(R[k] - 1) ? R[k] - 1 : int(f);
        step[k] = dd[k] > 0.0f ? 1 : dd[k] < 0.0f ? -1 : 0;
        tn[k] = step[k] == 0 ? INFINITY : (O[k] + float(i[k] + (step[k] > 0 ? 1 : 0)) * c - oo[k]) * inv[k];
    }
    const uint32_t rx = uint32_t(R[0]), ry = uint32_t(R[1]);
    const uint32_t bx = (rx + 3u) >> 2, by = (ry + 3u) >> 2, kx = (rx + 15u) >> 4, ky = (ry + 15u) >> 4;
    float te = t0;
    for (int it = 0, lim = R[0] + R[1] + R[2] + 3; it < lim; ++it) {
        const uint32_t idx = uint32_t(i[0]) + uint32_t(i[1]) * rx + uint32_t(i[2]) * (rx * ry);
        int S = 1;  // the empty aligned block around the cell: 16, 4, or 1 (not empty)
        if (B.vbricks && B.vblocks) {
            const uint32_t ki = (uint32_t(i[0]) >> 4) + (uint32_t(i[1]) >> 4) * kx + (uint32_t(i[2]) >> 4) * (kx * ky);
            if (!((B.vblocks[ki >> 5] >> (ki & 31u)) & 1u)) {
                S = 16;
            } else {
                const uint32_t bi =
                    (uint32_t(i[0]) >> 2) + (uint32_t(i[1]) >> 2) * bx + (uint32_t(i[2]) >> 2) * (bx * by);
                if (!((B.vbricks[bi >> 5] >> (bi & 31u)) & 1u)) S = 4;
            }
        }
        if (S > 1) {
            float fe = INFINITY;
            int a = -1, ja = 0;
            bool nan = false;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                if (step[k] == 0) continue;
                const int j = step[k] > 0 ? min(i[k] | (S - 1), R[k] - 1) : (i[k] & ~(S - 1));
                const float f = (O[k] + float(j + (step[k] > 0 ? 1 : 0)) * c - oo[k]) * inv[k];
                nan = nan || f != f;
                if (a < 0 || f < fe) {
                    fe = f;
                    a = k;
                    ja = j;
                }
            }
            if (!nan && a >= 0) {
                if (!(fe <= t1)) return;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    if (k == a || step[k] == 0) continue;
                    for (int s = 0; s < S - 1 && (tn[k] < fe || (tn[k] == fe && k < a)); ++s) {
                        i[k] += step[k];
                        tn[k] = (O[k] + float(i[k] + (step[k] > 0 ? 1 : 0)) * c - oo[k]) * inv[k];
                    }
                }
                te = fe > t0 ? fe : t0;
                ea = a;
                i[a] = ja + step[a];
                if (i[a] < 0 || i[a] >= R[a]) return;
                tn[a] = (O[a] + float(i[a] + (step[a] > 0 ? 1 : 0)) * c - oo[a]) * inv[a];
                continue;
            }
        }
        if ((B.voxels[idx >> 5] >> (idx & 31u)) & 1u) {
            if (te < best.t || (te == best.t && (inst < best.inst || (inst == best.inst && idx < best.pid)))) {
                int ax = ea;
                if (ax < 0) {
                    ax = 0;
                    if (fabsf(dd[1]) > fabsf(dd[ax])) ax = 1;
                    if (fabsf(dd[2]) > fabsf(dd[ax])) ax = 2;
                }
                const float sgn = dd[ax] > 0.0f ? -1.0f : 1.0f;
                best.t = te;
                best.inst = inst;
                best.pid = idx;
                best.slot = idx;
                best.u = 0.0f;
                best.a = ax == 0 ? sgn : 0.0f;
                best.b = ax == 1 ? sgn : 0.0f;
                best.c = ax == 2 ? sgn : 0.0f;
                best.extra = 0;
                if (ANY) done = true;
            }
            return;
        }
        int a = 0;
        if (tn[1] < tn[a]) a = 1;
        if (tn[2] < tn[a]) a = 2;
        if (!(tn[a] <= t1)) return;
        te = tn[a] > t0 ? tn[a] : t0;
        ea = a;
        i[a] += step[a];
        if (i[a] < 0 || i[a] >= R[a]) return;
        tn[a] = (O[a] + float(i[a] + (step[a] > 0 ? 1 : 0)) * c - oo[a]) * inv[a];
    }
}

template <class G, bool ANY, int STACK, bool STATS = false>
__device__ __forceinline__ bool trace_world(const TraceParams& tp, f3 o, f3 d, float tmin, float tmax, Best& best,
                                            uint2* lds, uint32_t ray_id, TravStats* st = nullptr) {
    best.t = tmax;
    best.inst = 0xFFFFFFFFu;
    best.pid = 0xFFFFFFFFu;
    best.slot = 0xFFFFFFFFu;
    best.u = best.a = best.b = best.c = 0.0f;
    best.extra = 0;
    bool ok = true;
    for (uint32_t i = 0; i < tp.ninst; ++i) {
        const InstanceDev& I = tp.inst[i];
        if (!ray_box(o, d, I.lo, I.hi, tmin, best.t)) continue;
        f3 oo = o, dd = d;
        if (!I.identity) {
            oo = xform_pt(I.inv, o);
            dd = xform_dir(I.inv, d);
        }
        bool done = false;
        bool r;
        if constexpr (G::kMixed) {
            const bool pol = (tp.flags & kTraceFp64Polish) != 0;
            if (I.blas.mode == kLss) {
                r = traverse<kLss, 0, false, ANY, STACK>(I.blas, oo, dd, tmin, i, best, lds, done, tp, ray_id);
            } else if (I.blas.mode == kDots) {
                r = traverse<kDots, 0, false, ANY, STACK>(I.blas, oo, dd, tmin, i, best, lds, done, tp, ray_id);
            } else if (I.blas.mode == kTriangles) {
                r = traverse<kTriangles, 0, false, ANY, STACK>(I.blas, oo, dd, tmin, i, best, lds, done, tp, ray_id);
            } else if (I.blas.mode == kVoxels) {
                voxel_walk<ANY>(I.blas, oo, dd, tmin, i, best, done);
                r = true;
            } else if (I.blas.radius_mode == kRadiusPerPoint) {
                r = pol ? traverse<kPhantom, kRadiusPerPoint, true, ANY, STACK>(I.blas, oo, dd, tmin, i, best, lds, done, tp, ray_id)
                        : traverse<kPhantom, kRadiusPerPoint, false, ANY, STACK>(I.blas, oo, dd, tmin, i, best, lds, done, tp, ray_id);
            } else {
                r = pol ? traverse<kPhantom, kRadiusConstant, true, ANY, STACK>(I.blas, oo, dd, tmin, i, best, lds, done, tp, ray_id)
                        : traverse<kPhantom, kRadiusConstant, false, ANY, STACK>(I.blas, oo, dd, tmin, i, best, lds, done, tp, ray_id);
            }
        } else {
            r = traverse<G::kMode, G::kRMode, G::kPolish, ANY, STACK, STATS>(I.blas, oo, dd, tmin, i, best, lds, done,
                                                                             tp, ray_id, st);
        }
        ok = ok && r;
        if (ANY && done) break;
    }
    return ok;
}

// SplitMix64 output function (Steele, Lea, Flood, OOPSLA 2014).
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Uniform sample in [0, 1) with 24 random bits: number k of slot b of item p (a path's depth,
// a hit's light sample).
__device__ __forceinline__ float path_rand(unsigned long long seed, uint32_t p, uint32_t b, uint32_t k) {
    const unsigned long long key =
        (static_cast<unsigned long long>(p) << 24) | (static_cast<unsigned long long>(b) << 16) | k;
    return float(uint32_t(mix64((seed ^ key) + 0x9E3779B97F4A7C15ull) >> 40)) * 0x1p-24f;
}

// Tangents t1, t2 of the unit vector n (Duff et al., "Building an Orthonormal Basis,
// Revisited", JCGT 6(1), 2017).
__device__ __forceinline__ void onb_duff(f3 n, f3& t1, f3& t2) {
    const float sg = copysignf(1.0f, n.z);
    const float a = -1.0f / (sg + n.z);
    const float bb = n.x * n.y * a;
    t1 = f3{1.0f + sg * n.x * n.x * a, sg * bb, -sg * n.x};
    t2 = f3{bb, sg + n.y * n.y * a, -n.y};
}

// Shadow ray of light sample s of pixel `pix` from the hit point o: direction and tmax.
__device__ __forceinline__ void light_ray(const AreaLight& L, uint32_t pix, uint32_t s, f3 o, f3& d, float& tmax) {
    if (!L.enabled) {
        d = f3{0.0f, 1.0f, 0.0f};
        tmax = 10000.0f;
        return;
    }
    const float a = 2.0f * path_rand(L.seed, pix, s, 0) - 1.0f;
    const float b = 2.0f * path_rand(L.seed, pix, s, 1) - 1.0f;
    const f3 y{L.c[0] + a * L.u[0] + b * L.v[0], L.c[1] + a * L.u[1] + b * L.v[1], L.c[2] + a * L.u[2] + b * L.v[2]};
    const f3 w = sub(y, o);
    d = nrm3(w);
    tmax = sqrtf(dot3(w, w));
}

// Resolves the hit record of `best` (normal in world space).
__device__ __forceinline__ void finish_hit(const TraceParams& tp, f3 o, f3 d, const Best& best, HitRec& h) {
    if (best.inst == 0xFFFFFFFFu) {
        h.t = 0.0f;
        h.prim_id = 0xFFFFFFFFu;
        h.inst_id = 0xFFFFFFFFu;
        h.u = 0.0f;
        h.n[0] = h.n[1] = h.n[2] = 0.0f;
        h.v = 0.0f;
        return;
    }
    const InstanceDev& I = tp.inst[best.inst];
    f3 dd = d;
    if (!I.identity) dd = xform_dir(I.inv, d);
    f3 n;
    float v = 0.0f;
    if (I.blas.mode == kPhantom) {
        PhantomHit ph;
        ph.ncx = best.a;
        ph.ncy = best.b;
        ph.nsz = best.c;
        ph.polished = best.extra != 0;
        n = phantom_normal(dd, ph);
        v = best.extra ? 1.0f : 0.0f;
    } else if (I.blas.mode == kLss || I.blas.mode == kVoxels) {
        n = f3{best.a, best.b, best.c};
    } else if (I.blas.mode == kTriangles) {
        // triangle_closest_hit.rchit:31-36: normalize(n0 b0 + n1 b1 + n2 b2), b = (1 - u - v, u, v);
        // the geometric normal when the vertex normals cancel
        const float4* vn = I.blas.normals + size_t(best.pid) * 3;
        const float4 n0 = vn[0], n1 = vn[1], n2 = vn[2];
        const float b0 = 1.0f - best.u - best.a;
        const f3 s{n0.x * b0 + n1.x * best.u + n2.x * best.a, n0.y * b0 + n1.y * best.u + n2.y * best.a,
                   n0.z * b0 + n1.z * best.u + n2.z * best.a};
        if (dot3(s, s) > 0.0f) {
            n = nrm3(s);
        } else {
            const float4* pr = I.blas.prims + size_t(best.slot) * 3;
            const float4 a = pr[0], b = pr[1], c = pr[2];
            n = nrm3(cross3(f3{b.x - a.x, b.y - a.y, b.z - a.z}, f3{c.x - a.x, c.y - a.y, c.z - a.z}));
        }
        v = best.a;
    } else {
        const float4* pr = I.blas.prims + size_t(best.slot) * 4;
        float4 P[4] = {pr[0], pr[1], pr[2], pr[3]};
        f3 v0, v1, v2;
        dots_tri(P, int(best.extra) >> 1, int(best.extra) & 1, v0, v1, v2);
        n = nrm3(cross3(sub(v1, v0), sub(v2, v0)));
        v = best.a;
    }
    if (!I.identity) {
        const float* x = I.xf;
        n = nrm3(f3{x[0] * n.x + x[1] * n.y + x[2] * n.z, x[4] * n.x + x[5] * n.y + x[6] * n.z,
                    x[8] * n.x + x[9] * n.y + x[10] * n.z});
    }
    h.t = best.t;
    h.prim_id = best.pid;
    h.inst_id = best.inst;
    h.u = best.u;
    h.n[0] = n.x;
    h.n[1] = n.y;
    h.n[2] = n.z;
    h.v = v;
}

template <class G, bool ANY, int STACK, bool STATS>
__global__ __launch_bounds__(kTraceBlock) void k_trace_rays(TraceParams tp, const float* __restrict__ rays, uint32_t n,
                                                            HitRec* __restrict__ hits, uint2* __restrict__ stats) {
    __shared__ uint2 stk[STACK * kTraceBlock];
    const uint32_t i = blockIdx.x * kTraceBlock + threadIdx.x;
    if (i >= n) return;
    const float4 r0 = reinterpret_cast<const float4*>(rays)[size_t(i) * 2];
    const float4 r1 = reinterpret_cast<const float4*>(rays)[size_t(i) * 2 + 1];
    const f3 o = f3{r0.x, r0.y, r0.z}, d = f3{r1.x, r1.y, r1.z};
    Best best;
    TravStats ts;
    const bool ok = trace_world<G, ANY, STACK, STATS>(tp, o, d, r0.w, r1.w, best, stk + threadIdx.x, i, &ts);
    if (STATS) stats[i] = make_uint2(ts.nodes, ts.prims);
    if (tp.keys) tp.keys[i] = best_key(tp, best);
    if (!ok) atomicAdd(&tp.status[0], 1u);
    HitRec h;
    finish_hit(tp, o, d, best, h);
    hits[i] = h;
}

// n / d for n < 2^24 from a float reciprocal, corrected to the exact quotient.
__device__ __forceinline__ uint32_t udiv_small(uint32_t n, uint32_t d, float inv_d) {
    uint32_t q = uint32_t(float(n) * inv_d);
    if (q * d > n)
        --q;
    else if ((q + 1u) * d <= n)
        ++q;
    return q;
}

// Work item r of a frame -> pixel (x, y); false for tile slots outside the image.
__device__ __forceinline__ bool work_pixel(const Camera& c, uint32_t r, uint32_t& x, uint32_t& y) {
    if (c.order == 0) {
        y = r < (1u << 24) ? udiv_small(r, c.width, c.inv_w) : r / c.width;
        x = r - y * c.width;
        return y < c.height;
    }
    if (c.order == 1) {
        const uint32_t tile = r >> 6, within = r & 63u;
        const uint32_t ty = tile < (1u << 24) ? udiv_small(tile, c.tiles_x, c.inv_tiles_x) : tile / c.tiles_x;
        x = (tile - ty * c.tiles_x) * 8u + (within & 7u);
        y = ty * 8u + (within >> 3);
        return x < c.width && y < c.height;
    }
    // 64x64 tiles (4096 items) in row-major order, each 64 8x8 tiles in row-major order
    const uint32_t big = r >> 12, within = r & 4095u, small = within >> 6;
    const uint32_t by = big < (1u << 24) ? udiv_small(big, c.tiles_x, c.inv_tiles_x) : big / c.tiles_x;
    x = (big - by * c.tiles_x) * 64u + (small & 7u) * 8u + (within & 7u);
    y = by * 64u + (small >> 3) * 8u + ((within >> 3) & 7u);
    return x < c.width && y < c.height;
}

// ray_gen.rgen:18-24 with GLSL mat4*vec4 order.
__device__ __forceinline__ void primary_ray(const Camera& c, uint32_t x, uint32_t y, f3& o, f3& d) {
    const float pcx = float(x) + 0.5f, pcy = float(y) + 0.5f;
    const float ux = pcx / float(c.width), uy = pcy / float(c.height);
    const float dx = ux * 2.0f - 1.0f, dy = uy * 2.0f - 1.0f;
    const float* V = c.view_inv;
    const float* P = c.proj_inv;
    o = f3{V[0] * 0.0f + V[4] * 0.0f + V[8] * 0.0f + V[12] * 1.0f, V[1] * 0.0f + V[5] * 0.0f + V[9] * 0.0f + V[13] * 1.0f,
           V[2] * 0.0f + V[6] * 0.0f + V[10] * 0.0f + V[14] * 1.0f};
    f3 tg = f3{P[0] * dx + P[4] * dy + P[8] * 1.0f + P[12] * 1.0f, P[1] * dx + P[5] * dy + P[9] * 1.0f + P[13] * 1.0f,
               P[2] * dx + P[6] * dy + P[10] * 1.0f + P[14] * 1.0f};
    f3 tn = nrm3(tg);
    d = f3{V[0] * tn.x + V[4] * tn.y + V[8] * tn.z + V[12] * 0.0f, V[1] * tn.x + V[5] * tn.y + V[9] * tn.z + V[13] * 0.0f,
           V[2] * tn.x + V[6] * tn.y + V[10] * tn.z + V[14] * 0.0f};
}

// shading.glsl:1-10 with the direct term scaled by light visibility (1 = lit).
__device__ __forceinline__ f3 shade(f3 n, float vis) {
    const float ndl = fabsf(n.x * 0.0f + n.y * -1.0f + n.z * 0.0f);
    const float k = ndl * vis;
    return f3{k * 0.4f + 0.3f, k * 0.2f + 0.3f, k * 0.1f + 0.3f};
}

// debug.glsl:1-7
__device__ __forceinline__ f3 debug_color(uint32_t prim) {
    const int idx = int(prim) % 6;
    const float r[6] = {1.0f, 0.8f, 0.6f, 0.4f, 0.2f, 0.0f};
    const float g[6] = {0.0f, 0.2f, 0.4f, 0.6f, 0.8f, 1.0f};
    return f3{r[idx], g[idx], 0.3f};
}

// miss.rmiss:17-38: the inverted ray direction -> equirectangular uv (DirectionToUV), a
// bilinear RGBA32F lookup with repeat addressing (texel centres at (i + 0.5) / size), then
// 1 - exp(-c exposure) and the 1/2.2 gamma.
__device__ __forceinline__ f3 env_color(const EnvMap& E, f3 d) {
    const f3 v = nrm3(f3{-d.x, -d.y, -d.z});
    const float gamma = asinf(v.y);
    const float theta = atan2f(v.x, -v.z);
    const float u = theta * 0.3183098861837f * 0.5f + 0.5f;
    const float t = gamma * 0.3183098861837f + 0.5f;
    const float x = u * float(E.width) - 0.5f, y = t * float(E.height) - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float ax = x - fx, ay = y - fy;
    auto wrap = [](int i, int n) {
        const int r = i % n;
        return r < 0 ? r + n : r;
    };
    const int w = int(E.width), h = int(E.height);
    const int x0 = wrap(int(fx), w), x1 = wrap(int(fx) + 1, w), y0 = wrap(int(fy), h), y1 = wrap(int(fy) + 1, h);
    const float4 c00 = E.texels[size_t(y0) * w + x0], c10 = E.texels[size_t(y0) * w + x1];
    const float4 c01 = E.texels[size_t(y1) * w + x0], c11 = E.texels[size_t(y1) * w + x1];
    auto lerp = [](float a, float b, float s) { return a + (b - a) * s; };
    const float r = lerp(lerp(c00.x, c10.x, ax), lerp(c01.x, c11.x, ax), ay);
    const float g = lerp(lerp(c00.y, c10.y, ax), lerp(c01.y, c11.y, ax), ay);
    const float b = lerp(lerp(c00.z, c10.z, ax), lerp(c01.z, c11.z, ax), ay);
    const float k = 1.0f / 2.2f;
    return f3{powf(1.0f - expf(-r * E.exposure), k), powf(1.0f - expf(-g * E.exposure), k),
              powf(1.0f - expf(-b * E.exposure), k)};
}

// RGBA32F (alpha 1) or RGBA8 (clamped to [0, 1], rounded) pixel store.
__device__ __forceinline__ void store_rgba(void* image, uint32_t format, uint32_t pix, f3 c) {
    if (format == 0) {
        reinterpret_cast<float4*>(image)[pix] = make_float4(c.x, c.y, c.z, 1.0f);
    } else {
        auto q = [](float v) { return uint32_t(fminf(fmaxf(v, 0.0f), 1.0f) * 255.0f + 0.5f); };
        reinterpret_cast<uint32_t*>(image)[pix] = q(c.x) | (q(c.y) << 8) | (q(c.z) << 16) | (255u << 24);
    }
}
__device__ __forceinline__ void store_px(const RenderParams& rp, uint32_t pix, f3 c) {
    store_rgba(rp.image, rp.out_format, pix, c);
}

template <class G, int STACK>
__global__ __launch_bounds__(kTraceBlock) void k_render_primary(RenderParams rp) {
    __shared__ uint2 stk[STACK * kTraceBlock];
    const uint32_t npix = rp.cam.width * rp.cam.height;
    const uint32_t pix = blockIdx.x * kTraceBlock + threadIdx.x;
    bool hit = false;
    if (pix < npix) {
        const uint32_t x = pix % rp.cam.width, y = pix / rp.cam.width;
        f3 o, d;
        primary_ray(rp.cam, x, y, o, d);
        Best best;
        bool ok = trace_world<G, false, STACK>(rp.tp, o, d, 0.001f, 10000.0f, best, stk + threadIdx.x, pix);
        if (!ok) atomicAdd(&rp.tp.status[0], 1u);
        HitRec h;
        finish_hit(rp.tp, o, d, best, h);
        hit = h.prim_id != 0xFFFFFFFFu;
        if (rp.prim_out) rp.prim_out[pix] = h.prim_id;
        const f3 n = f3{h.n[0], h.n[1], h.n[2]};
        if (!hit) {
            f3 c{0.0f, 0.0f, 0.0f};
            if (rp.env.texels) {
                f3 o, d;
                primary_ray(rp.cam, x, y, o, d);
                c = env_color(rp.env, d);
            }
            store_px(rp, pix, c);
        } else if (rp.debug_palette) {
            store_px(rp, pix, debug_color(h.prim_id));
            hit = false;
        } else if (!rp.shadows) {
            store_px(rp, pix, shade(n, 1.0f));
            hit = false;
        } else {
            rp.hitbuf[pix] = make_float4(n.x, n.y, n.z, h.t);
        }
    }
    // wave-aggregated append of shadow work
    const uint64_t m = __ballot(hit);
    if (m) {
        const uint32_t lane = threadIdx.x & 63;
        const uint32_t leader = __ffsll((unsigned long long)m) - 1;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&rp.counters[0], uint32_t(__popcll(m)));
        base = __shfl(base, int(leader));
        if (hit) rp.queue[base + __popcll(m & ((1ull << lane) - 1ull))] = pix;
    }
}

template <class G, int STACK>
__global__ __launch_bounds__(kTraceBlock) void k_render_shadow(RenderParams rp) {
    __shared__ uint2 stk[STACK * kTraceBlock];
    const uint32_t qi = blockIdx.x * kTraceBlock + threadIdx.x;
    if (qi >= rp.counters[0]) return;
    const uint32_t pix = rp.queue[qi];
    const uint32_t x = pix % rp.cam.width, y = pix / rp.cam.width;
    f3 o, d;
    primary_ray(rp.cam, x, y, o, d);
    const float4 hb = rp.hitbuf[pix];
    float vis;
    if (rp.light_vis) {
        vis = rp.light_vis[pix];
    } else {
        const f3 hp = add(o, scl(hb.w, d));
        Best best;
        bool ok = trace_world<G, true, STACK>(rp.tp, hp, f3{0.0f, 1.0f, 0.0f}, 0.001f, 10000.0f, best, stk + threadIdx.x,
                                              qi);
        if (!ok) atomicAdd(&rp.tp.status[0], 1u);
        vis = best.inst == 0xFFFFFFFFu ? 1.0f : 0.0f;
    }
    store_px(rp, pix, shade(f3{hb.x, hb.y, hb.z}, vis));
}

// ------------------------------------------------------------------ persistent lanes
// How the persistent loops reach the BLAS fields of a step's instance: through the lane's
// instance index in global memory (many instances), once into scalar registers (one), or
// staged per block in LDS (up to kFewInstances).
enum InstanceMode { kInstMany = 0, kInstOne = 1, kInstFew = 2 };
constexpr uint32_t kFewInstances = 16;
inline int instance_mode(uint32_t ninst) {
    return ninst == 1 ? kInstOne : ninst <= kFewInstances ? kInstFew : kInstMany;
}

// A copy of a value every lane of the wave reads alike, word by word through readfirstlane,
// so that it is held in scalar registers.
template <class T>
__device__ __forceinline__ T wave_uniform(const T& v) {
    static_assert(sizeof(T) % 4 == 0, "whole 32-bit words");
    T out;
    const uint32_t* src = reinterpret_cast<const uint32_t*>(&v);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
    for (uint32_t i = 0; i < sizeof(T) / 4; ++i) dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
    return out;
}

// Hands work items [0, total) to the waves of a persistent grid. Without chunks one counter
// serves the grid. With chunks (a power of two) the items are cut into runs of `chunk`
// consecutive items; run k belongs to XCD k mod 8 and is handed out by that XCD's own counter
// (xcd[kXcdStride * (k mod 8)], one 128-byte line each), so the rays of one image region meet
// the BVH nodes they share in a single XCD's L2 (each of the 8 XCDs has its own). A wave whose
// XCD's counter has run dry moves on to the next XCD's counter, round the eight, so no XCD
// idles while another has rays left. The claim that empties a counter marks it in the block's
// `dry` mask (LDS), and the block's other waves skip it without an atomic: counter atomics to
// one line are serialized, and a grid of 4096 waves probing 7 empty counters each would cost
// more than the whole tail. All members are wave-uniform.
struct WorkClaim {
    uint32_t* counter;
    uint32_t* xcd;  // null: one counter
    uint32_t* dry;  // LDS, bit c: counter c is empty (chunked only)
    uint32_t total, shift;
    uint32_t cx, left;  // counter in use; XCD counters not yet found dry
    bool exhausted;     // no item left anywhere

    __device__ __forceinline__ WorkClaim(uint32_t* c, uint32_t* x, uint32_t chunk, uint32_t n, uint32_t* d)
        : counter(c), xcd(chunk ? x : nullptr), dry(d), total(n), shift(chunk ? uint32_t(__ffs(chunk) - 1) : 0u),
          cx(__builtin_amdgcn_s_getreg((3 << 11) | 20) & 7u),  // hwreg(HW_REG_XCC_ID, 0, 4)
          left(8u), exhausted(false) {}

    // The idle lanes of the wave (mask `idle`, the same in every lane) claim one item each;
    // returns the calling lane's item for `rank` = its rank among the idle lanes, or a value
    // >= total when none was left for it.
    __device__ __forceinline__ uint32_t claim(uint64_t idle, uint32_t rank) {
        const uint32_t n = __popcll(idle);
        const uint32_t leader = __ffsll((unsigned long long)idle) - 1;
        const uint32_t lane = threadIdx.x & 63u;
        uint32_t base = 0;
        if (!xcd) {
            if (lane == leader) base = atomicAdd(counter, n);
            base = __shfl(base, int(leader));
            if (base + n >= total) exhausted = true;
            return base + rank;
        }
        // skip the counters the block already found empty
        const uint32_t known = __shfl(*reinterpret_cast<volatile uint32_t*>(dry), 0);
        while (left != 0u && ((known >> cx) & 1u)) {
            cx = (cx + 1u) & 7u;
            --left;
        }
        if (left == 0u) {
            exhausted = true;
            return 0xFFFFFFFFu;
        }
        const uint64_t chunk = 1ull << shift;
        const uint64_t runs = (uint64_t(total) + chunk - 1) >> shift;
        const uint64_t own = runs > cx ? (runs - 1 - cx) / 8 + 1 : 0;  // runs of XCD cx
        uint64_t lim = own << shift;                                   // their items
        if (runs != 0 && (runs - 1) % 8 == cx) lim -= (runs << shift) - total;  // the last run is short
        if (lane == leader) base = atomicAdd(xcd + kXcdStride * cx, n);
        base = __shfl(base, int(leader));
        const uint64_t i = uint64_t(base) + rank;
        uint32_t r = 0xFFFFFFFFu;
        if (i < lim) r = uint32_t(((((i >> shift) << 3) + cx) << shift) + (i & (chunk - 1)));
        if (uint64_t(base) + n >= lim) {
            if (lane == leader) atomicOr(dry, 1u << cx);
            cx = (cx + 1u) & 7u;
            if (--left == 0u) exhausted = true;
        }
        return r;
    }
};

// Persistent wavefronts with per-lane refill: every iteration each active lane performs one
// walker step; lanes whose ray finished take new work from a global counter (one atomic per
// wave per refill), so a wavefront is not held back by its slowest ray. src(r, id, o, d, tmin,
// tmax) produces work item r and names its ray `id` (the index results and deferred-polish
// entries use); snk(id, o, d, best) consumes the result (called by the lanes that finish in
// the same iteration, so it may use wave ballots).
template <class G, bool ANY, int STACK, int IM, class Src, class Snk>
__device__ __forceinline__ void persistent_generic(const TraceParams& tp, WorkClaim& work, uint2* lds, Src&& src,
                                                   Snk&& snk) {
    BlasDev b1{};  // one instance: as in persistent_compact (a few: read per lane here)
    if constexpr (IM == kInstOne) b1 = wave_uniform(tp.inst[0].blas);
    static_assert(!G::kMixed, "persistent kernels serve single-mode TLASes");
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t lt = (1ull << lane) - 1ull;
    Walker<G::kMode, G::kRMode, G::kPolish, ANY, STACK> W;
    Best best;
    f3 ow = f3{0, 0, 0}, dw = f3{0, 0, 0};
    float tmin = 0.0f, tmax = 0.0f;
    uint32_t idx = 0, inst = 0;
    bool active = false, need_inst = false;
    for (;;) {
        if (!work.exhausted) {
            const uint64_t idle = __ballot(!active);
            if (idle && (__popcll(idle) >= tp.refill_min || idle == __ballot(true))) {
                const uint32_t r = work.claim(idle, __popcll(idle & lt));
                if (!active) {
                    if (r < work.total) {
                        src(r, idx, ow, dw, tmin, tmax);
                        best.t = tmax;
                        best.inst = 0xFFFFFFFFu;
                        best.pid = 0xFFFFFFFFu;
                        best.slot = 0xFFFFFFFFu;
                        best.u = best.a = best.b = best.c = 0.0f;
                        best.extra = 0;
                        inst = 0;
                        need_inst = true;
                        active = true;
                    }
                }
            }
        }
        if (!__ballot(active)) break;
        bool finish = false;
        if (active && need_inst) {
            // next instance whose world box the ray overlaps (top_level_acceleration_structure.cpp:23-43)
            while (inst < tp.ninst) {
                const InstanceDev& I = tp.inst[inst];
                if (I.blas.nprims != 0 && ray_box(ow, dw, I.lo, I.hi, tmin, best.t)) break;
                ++inst;
            }
            if (inst >= tp.ninst) {
                finish = true;
            } else {
                const InstanceDev& I = tp.inst[inst];
                f3 oo = ow, dd = dw;
                if (!I.identity) {
                    oo = xform_pt(I.inv, ow);
                    dd = xform_dir(I.inv, dw);
                }
                W.begin(oo, dd);
                need_inst = false;
            }
        }
        // up to tp.steps_per_iter traversal steps before the refill and finish bookkeeping again
        const uint32_t nsteps = tp.steps_per_iter > 1u ? tp.steps_per_iter : 1u;
        for (uint32_t si = 0; si < nsteps; ++si) {
            const bool walk = active && !need_inst && !finish;
            if (si > 0u && !__ballot(walk)) break;
            if (walk) {
                const int r = W.step(IM == kInstOne ? b1 : tp.inst[inst].blas, tmin, inst, best, lds, nullptr, tp, idx);
                if (r == kStepError) {
                    atomicAdd(&tp.status[0], 1u);
                    finish = true;
                } else if (r == kStepHit) {
                    finish = true;
                } else if (r == kStepDone) {
                    ++inst;
                    need_inst = true;
                }
            }
        }
        if (active && finish) {
            snk(idx, ow, dw, best);
            active = false;
        }
    }
}

// ------------------------------------------------------------------ compacted phantom lanes
// Phantom persistent loop with wave-level compaction of the intersection work. Lanes run one
// traversal step per iteration: a node test, or one leaf record through the capped-cylinder
// early-out (hair_intersection.rint:24-33). A (lane, leaf slot) pair that survives the
// early-out is appended to a per-wave queue in LDS. When the queue holds 64 pairs, or no
// lane can traverse, every lane of the wave takes one pair and runs the rest of Prhi on the
// owner's ray (read with cross-lane permutes), so the iteration loop runs on full
// wavefronts instead of on the few lanes that happen to sit at a leaf. Results merge into the
// owner's LDS slot by a 64-bit atomicMin on (t, instance, primitive): that minimum is unique,
// so the closest hit and its tie rule do not depend on the processing order. A lane does not
// move to its next instance, finish or take a new ray while pairs of its ray are queued.
constexpr uint32_t kCompactQueue = 128;    // per wave: < 64 left after a round + <= 64 appended per step
constexpr uint32_t kNoHitLow = 0xFFFFFFFFu;
constexpr uint32_t kCompactIterationCap = 1u << 24;

template <int STACK>
struct CWalker {
    f3 o, d, inv;
    uint32_t oct;
    uint2 cur;  // current group: x = child_base, y = imask | remaining hit bits (octant order) << 8
    uint32_t node, pbase, pmask;
    int sp, sb;  // the stack is a deque: push / pop at sp, other lanes take entries at sb
    bool have_node;

    __device__ __forceinline__ void begin(f3 o_, f3 d_) {
        o = o_;
        d = d_;
        inv = safe_inv(d);
        oct = uint32_t(signbit(d.x)) | (uint32_t(signbit(d.y)) << 1) | (uint32_t(signbit(d.z)) << 2);
        cur = make_uint2(0u, 0u);
        node = 0;
        pbase = 0;
        pmask = 0;
        sp = sb = 0;
        have_node = true;
    }

    // Continues another lane's ray from one of its stacked node groups.
    __device__ __forceinline__ void adopt(f3 o_, f3 d_, f3 inv_, uint32_t oct_, uint2 group) {
        o = o_;
        d = d_;
        inv = inv_;
        oct = oct_;
        cur = group;
        node = 0;
        pbase = 0;
        pmask = 0;
        sp = sb = 0;
        have_node = false;
    }

    // One step: one pending leaf record through the early-out, or one node test; then, when
    // no record of the node is pending, the next node is selected. `cand` receives the leaf
    // slot that passed the early-out (0xFFFFFFFF if none); it is valid with any return value.
    __device__ __forceinline__ int step(const BlasDev& B, float tmin, float tcur, uint2* lds, uint32_t& cand) {
        cand = 0xFFFFFFFFu;
        if (pmask) {
            const uint32_t slot = pbase + uint32_t(__ffs(pmask) - 1);
            pmask &= pmask - 1;
            const auto* pr = gmem(B.prims + size_t(slot) * 4);
            const float4 a = pr[0], b = pr[3];
            if (ray_cylinder(o, d, f3{a.x, a.y, a.z}, f3{b.x, b.y, b.z}, gmem(B.aux)[slot])) cand = slot;
        } else if (have_node) {
            if (node >= B.nnodes) return kStepError;  // malformed structure: report, never read out of bounds
            uint32_t hits, pm, child_base, imask, pb;
            node_visit(B.nodes, node, o, inv, oct, tmin, tcur, hits, pm, child_base, imask, pb);
            if (pm && pb + (32u - __clz(pm)) > B.nprims) return kStepError;
            pbase = pb;
            pmask = pm;
            have_node = false;
            if (hits) {
                const uint2 g = make_uint2(child_base, imask | (hits << 8));
                if (cur.y >> 8) {
                    if (sp >= STACK) return kStepError;
                    lds[sp * kTraceBlock] = cur;
                    ++sp;
                }
                cur = g;
            }
        }
        if (pmask || have_node) return kStepMore;
        if ((cur.y >> 8) == 0) {
            if (sp == sb) return kStepDone;
            --sp;
            cur = lds[sp * kTraceBlock];
        }
        uint32_t hb = cur.y >> 8;
        const uint32_t j = __ffs(hb) - 1;
        hb &= hb - 1;
        const uint32_t slot = j ^ oct;
        const uint32_t im = cur.y & 0xffu;
        node = cur.x + __popc(im & ((1u << slot) - 1u));
        cur.y = im | (hb << 8);
        have_node = true;
        return kStepMore;
    }
};

template <class G, bool ANY, int STACK, int IM, class Src, class Snk>
__device__ __forceinline__ void persistent_compact(const TraceParams& tp, WorkClaim& work, uint2* lds, Src&& src,
                                                   Snk&& snk) {
    // One instance: its BLAS fields are read once here, into scalar registers, instead of per
    // step and lane. A few instances: the block stages all their BLAS fields in LDS, so a
    // step reads them from there instead of through two dependent global loads.
    BlasDev b1{};
    const BlasDev* few = nullptr;
    if constexpr (IM == kInstOne) b1 = wave_uniform(tp.inst[0].blas);
    if constexpr (IM == kInstFew) {
        __shared__ BlasDev s_blas[kFewInstances];
        for (uint32_t i = threadIdx.x; i < tp.ninst && i < kFewInstances; i += blockDim.x) s_blas[i] = tp.inst[i].blas;
        __syncthreads();
        few = s_blas;
    }
    auto blas = [&](uint32_t i) -> const BlasDev& {
        if constexpr (IM == kInstOne)
            return b1;
        else if constexpr (IM == kInstFew)
            return few[i];
        else
            return tp.inst[i].blas;
    };
    static_assert(G::kMode == kPhantom, "compacted lanes serve the phantom intersector");
    __shared__ uint32_t s_queue[(kTraceBlock / 64) * kCompactQueue];
    __shared__ unsigned long long s_key[kTraceBlock];  // best (t, instance, primitive) of the lane's ray
    __shared__ float4 s_pay[kTraceBlock];              // its u and ray-centric normal vector
    __shared__ uint32_t s_pol[kTraceBlock];            // its polished flag
    __shared__ uint32_t s_cnt[kTraceBlock];            // pairs of the lane's ray still queued
    __shared__ uint32_t s_walk[kTraceBlock];           // other lanes walking subtrees of the lane's ray
    __shared__ uint32_t s_xfer[kTraceBlock];           // donor lane by rank (subtree hand-over)
    // kTrav: walking its own ray; kHelp: walking a subtree of the ray of lane `home`; kWait: own
    // walk done, waiting for queued pairs and helpers of its ray
    enum { kIdle = 0, kInst = 1, kTrav = 2, kWait = 3, kHelp = 4 };
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wbase = threadIdx.x & ~63u;
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t* queue = s_queue + (threadIdx.x >> 6) * kCompactQueue;
    CWalker<STACK> W;
    W.o = W.d = f3{0.0f, 0.0f, 0.0f};
    f3 ow = f3{0, 0, 0}, dw = f3{0, 0, 0};
    float tmin = 0.0f, tmax = 0.0f, tcur = 0.0f;
    uint32_t idx = 0, inst = 0, home = lane;
    int state = kIdle;
    bool ray_done = false;
    uint32_t qn = 0;         // wave-uniform queue length
#ifdef VKHRT_WORK_COUNTERS
    unsigned long long wc[14] = {};  // WorkCounter sums of this wave (wave-uniform)
    unsigned long long t_it = clock64(), t_mid = t_it;
    wc[kWcWaves] = 1;
#define VK_WC(i, v) (wc[i] += (unsigned long long)(v))
#define VK_WC_MARK() (t_mid = clock64())
#define VK_WC_END(kind)                                   \
    do {                                                  \
        const unsigned long long t_now = clock64();       \
        wc[kWcCyclesOther] += t_mid - t_it;               \
        wc[kind] += t_now - t_mid;                        \
        t_it = t_mid = t_now;                             \
    } while (0)
#else
#define VK_WC(i, v) ((void)0)
#define VK_WC_MARK() ((void)0)
#define VK_WC_END(kind) ((void)0)
#endif
    // Every wave leaves the loop: a normal frame takes ~10^4 iterations per wave; past the cap
    // the wave reports an error (status[0]) and stops, leaving its rays unwritten.
    for (uint32_t guard = 0;; ++guard) {
        VK_WC(kWcIterations, 1);
        if (guard >= kCompactIterationCap) {
            if (lane == 0) atomicAdd(&tp.status[0], 1u);
            break;
        }
        // (a) a waiting lane whose pairs are all merged moves to its next instance or finishes
        if (state == kWait && s_cnt[threadIdx.x] == 0u && s_walk[threadIdx.x] == 0u) {
            if (ray_done) {
                state = kInst;
                inst = tp.ninst;  // finish below
            } else {
                ++inst;
                state = kInst;
            }
        }
        // (b) refill idle lanes from the work counters
        if (!work.exhausted) {
            const uint64_t idle = __ballot(state == kIdle);
            if (idle && (__popcll(idle) >= tp.refill_min || idle == __ballot(true))) {
                const uint32_t r = work.claim(idle, __popcll(idle & lt));
                if (state == kIdle) {
                    if (r < work.total) {
                        src(r, idx, ow, dw, tmin, tmax);
                        tcur = tmax;
                        s_key[threadIdx.x] =
                            (static_cast<unsigned long long>(__float_as_uint(tmax)) << 32) | kNoHitLow;
                        s_pol[threadIdx.x] = 0u;
                        s_cnt[threadIdx.x] = 0u;
                        s_walk[threadIdx.x] = 0u;
                        home = lane;
                        inst = 0;
                        ray_done = false;
                        state = kInst;
                    }
                }
            }
        }
        // (c) next instance whose world box the ray overlaps (top_level_acceleration_structure.cpp:23-43)
#ifdef VKHRT_WORK_COUNTERS
        bool finished = false;
#endif
        if (state == kInst) {
            while (inst < tp.ninst) {
                const InstanceDev& I = tp.inst[inst];
                if (I.blas.nprims != 0 && ray_box(ow, dw, I.lo, I.hi, tmin, tcur)) break;
                ++inst;
            }
            if (inst >= tp.ninst) {
                const unsigned long long kk = s_key[threadIdx.x];
                const uint32_t lo = uint32_t(kk);
                Best best;
                best.slot = 0xFFFFFFFFu;
                if (lo == kNoHitLow) {
                    best.t = tmax;
                    best.inst = best.pid = 0xFFFFFFFFu;
                    best.u = best.a = best.b = best.c = 0.0f;
                    best.extra = 0u;
                } else {
                    const float4 pay = s_pay[threadIdx.x];
                    best.t = __uint_as_float(uint32_t(kk >> 32));
                    if constexpr (IM == kInstOne) {
                        best.inst = 0;
                    } else if constexpr (IM == kInstFew) {
                        uint32_t k = 0;
                        for (uint32_t j = 1; j < tp.ninst; ++j) k = few[j].gbase <= lo ? j : k;
                        best.inst = k;
                    } else {
                        best.inst = gid_instance(tp.inst, tp.ninst, lo);
                    }
                    best.pid = lo - blas(best.inst).gbase;
                    best.u = pay.x;
                    best.a = pay.y;
                    best.b = pay.z;
                    best.c = pay.w;
                    best.extra = s_pol[threadIdx.x];
                }
                snk(idx, ow, dw, best);
                state = kIdle;
#ifdef VKHRT_WORK_COUNTERS
                finished = true;
#endif
            } else {
                const InstanceDev& I = tp.inst[inst];
                f3 oo = ow, dd = dw;
                if (!I.identity) {
                    oo = xform_pt(I.inv, ow);
                    dd = xform_dir(I.inv, dw);
                }
                W.begin(oo, dd);
                state = kTrav;
            }
        }
#ifdef VKHRT_WORK_COUNTERS
        VK_WC(kWcRays, __popcll(__ballot(finished)));
#endif
        // (c') once the counter is exhausted, idle lanes take the bottom (largest) stacked node
        // group of a walking lane and continue that ray from there, so the last long rays of
        // the frame are walked by many lanes instead of one
        if (work.exhausted) {
            const bool give = (state == kTrav || state == kHelp) && W.sp > W.sb;
            const uint64_t dm = __ballot(give);
            const uint64_t im = __ballot(state == kIdle);
            if (dm && im) {
                const uint32_t n = min(uint32_t(__popcll(dm)), uint32_t(__popcll(im)));
                const uint32_t kd = __popcll(dm & lt), ki = __popcll(im & lt);
                if (give && kd < n) s_xfer[wbase + kd] = lane;
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                __builtin_amdgcn_wave_barrier();
                const bool take = state == kIdle && ki < n;
                const int from = take ? int(s_xfer[wbase + ki]) : int(lane);
                const f3 go = f3{__shfl(W.o.x, from), __shfl(W.o.y, from), __shfl(W.o.z, from)};
                const f3 gd = f3{__shfl(W.d.x, from), __shfl(W.d.y, from), __shfl(W.d.z, from)};
                const f3 gi = f3{__shfl(W.inv.x, from), __shfl(W.inv.y, from), __shfl(W.inv.z, from)};
                const uint32_t goct = __shfl(W.oct, from), gsb = uint32_t(__shfl(W.sb, from));
                const uint32_t ghome = __shfl(home, from), ginst = __shfl(inst, from);
                const float gtmin = __shfl(tmin, from);
                if (take) {
                    const uint2 grp = lds[(from - int(lane)) + int(gsb) * kTraceBlock];
                    W.adopt(go, gd, gi, goct, grp);
                    home = ghome;
                    inst = ginst;
                    tmin = gtmin;
                    tcur = __uint_as_float(uint32_t(s_key[wbase + home] >> 32));
                    atomicAdd(&s_walk[wbase + home], 1u);
                    state = kHelp;
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                __builtin_amdgcn_wave_barrier();
                if (give && kd < n) ++W.sb;  // the group now belongs to the helper
                VK_WC(kWcHandovers, n);
            }
        }
        // lanes that finished in (c) take new rays in the next iteration: leave only when the
        // counter is exhausted and no lane holds a ray (then no pair is queued either)
        if (!__ballot(state != kIdle)) {
            if (work.exhausted) break;
            continue;
        }
        // (d) intersection round on up to 64 queued pairs, one per lane
        VK_WC_MARK();
        if (qn >= 64u || (qn != 0u && !__ballot(state == kTrav || state == kHelp))) {
            const uint32_t m = qn < 64u ? qn : 64u;
            VK_WC(kWcRounds, 1);
            VK_WC(kWcRoundLanes, m);
            const uint32_t e = lane < m ? queue[qn - m + lane] : 0u;
            const int owner = int(e >> kCompactSlotBits);
            const uint32_t slot = e & ((1u << kCompactSlotBits) - 1u);
            const f3 po = f3{__shfl(W.o.x, owner), __shfl(W.o.y, owner), __shfl(W.o.z, owner)};
            const f3 pd = f3{__shfl(W.d.x, owner), __shfl(W.d.y, owner), __shfl(W.d.z, owner)};
            const float ptmin = __shfl(tmin, owner), ptmax = __shfl(tmax, owner);
            const uint32_t pinst = __shfl(inst, owner), pidx = __shfl(idx, owner);
            const uint32_t os = wbase + uint32_t(owner);
            unsigned long long k = ~0ull;
            PhantomHit ph;
            bool hit = false;
            if (lane < m) {
                const bool skip = ANY && uint32_t(s_key[os]) != kNoHitLow;  // already occluded
                if (!skip) {
                    const BlasDev& B = blas(pinst);
                    RayFrame F;
                    make_frame(po, pd, F);
                    const auto* pr = gmem(B.prims + size_t(slot) * 4);
                    float4 cv[4] = {pr[0], pr[1], pr[2], pr[3]};
                    hit = prhi<G::kRMode, G::kPolish ? kPolishDefer : kPolishOff, false>(F, cv, 0.0f, B.radius, ph);
                    if (G::kPolish && ph.deferred) {
                        // this curve needs the FP64 continuation: queue (ray, instance, slot)
                        const uint32_t q = atomicAdd(tp.defer_count, 1u);
                        if (q < tp.defer_cap)
                            tp.defer[q] = make_uint4(pidx, pinst, slot, 0u);
                        else
                            atomicAdd(&tp.status[1], 1u);
                    }
                    hit = hit && ph.t >= ptmin && ph.t <= ptmax;
                    if (hit) {
                        k = pack_key(ph.t, B.gbase + B.prim_ids[slot]);
                        atomicMin(&s_key[os], k);
                    }
                }
                atomicSub(&s_cnt[os], 1u);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
            if (hit && s_key[os] == k) {  // the unique winner for this owner writes the payload
                s_pay[os] = make_float4(ph.u, ph.ncx, ph.ncy, ph.nsz);
                s_pol[os] = ph.polished ? 1u : 0u;
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
            qn -= m;
            if (state != kIdle) {
                const unsigned long long kk = s_key[wbase + home];
                tcur = __uint_as_float(uint32_t(kk >> 32));
                if (ANY && state != kHelp && uint32_t(kk) != kNoHitLow && !ray_done) {
                    ray_done = true;
                    state = kWait;
                }
            }
            VK_WC_END(kWcCyclesRounds);
            continue;
        }
        // (e) traversal steps, one per traversing lane each; survivors of the early-out are queued.
        // tp.steps_per_iter > 1 runs up to that many steps before the bookkeeping above comes
        // round again, while the queue can take a full step of pairs and some lane still walks.
        // (Stopping early once refill_min lanes have stopped walking measured no better.)
        const uint32_t nsteps = tp.steps_per_iter > 1u ? tp.steps_per_iter : 1u;
        for (uint32_t si = 0; si < nsteps; ++si) {
            if (si > 0u && (qn >= 64u || !__ballot(state == kTrav || state == kHelp))) break;
#ifdef VKHRT_WORK_COUNTERS
            {
                const bool wk = state == kTrav || state == kHelp;
                VK_WC(kWcNodeTests, __popcll(__ballot(wk && W.pmask == 0u && W.have_node)));
                VK_WC(kWcLeafTests, __popcll(__ballot(wk && W.pmask != 0u)));
                VK_WC(kWcStepLanes, __popcll(__ballot(wk)));
            }
#endif
            uint32_t cand = 0xFFFFFFFFu;
            if (state == kHelp && ANY && uint32_t(s_key[wbase + home]) != kNoHitLow) {
                atomicSub(&s_walk[wbase + home], 1u);  // the ray is occluded already
                state = kIdle;
            } else if (state == kTrav || state == kHelp) {
                const int r = W.step(blas(inst), tmin, tcur, lds, cand);
                if (r == kStepError) atomicAdd(&tp.status[0], 1u);
                if (state == kHelp) {
                    if (r == kStepError || r == kStepDone) {
                        atomicSub(&s_walk[wbase + home], 1u);
                        state = kIdle;
                    }
                } else if (r == kStepError) {
                    ray_done = true;
                    state = kWait;
                } else if (r == kStepDone) {
                    state = kWait;
                }
            }
            const bool app = cand != 0xFFFFFFFFu;
            const uint64_t am = __ballot(app);
            if (app) {
                queue[qn + __popcll(am & lt)] = (home << kCompactSlotBits) | cand;
                atomicAdd(&s_cnt[wbase + home], 1u);
            }
            qn += __popcll(am);
            VK_WC(kWcSteps, 1);
            VK_WC(kWcPairs, __popcll(am));
            VK_WC_END(kWcCyclesSteps);
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
    }
#ifdef VKHRT_WORK_COUNTERS
    if (lane == 0 && tp.work_counts) {
        const uint32_t base = ANY ? kWorkAnyHit : 0u;
        for (int i = 0; i < 14; ++i)
            if (wc[i]) atomicAdd(&tp.work_counts[base + i], wc[i]);
    }
#endif
#undef VK_WC
#undef VK_WC_MARK
#undef VK_WC_END
}

// Persistent loop for one geometry variant: compacted intersection rounds for phantom curves
// (the expensive, iterative test), the per-lane loop for LSS and DOTS.
template <class G, bool ANY, int STACK, int IM = kInstMany, class Src, class Snk>
__device__ __forceinline__ void persistent_lanes(const TraceParams& tp, uint32_t total, uint32_t* counter, uint2* lds,
                                                 Src&& src, Snk&& snk, uint32_t* xcd = nullptr) {
    // an empty queue (uniform over the grid) ends the launch without touching the counter
    if (total == 0) return;
    const uint32_t chunk = xcd ? tp.xcd_chunk : 0u;
    __shared__ uint32_t s_dry;
    if (chunk) {
        if (threadIdx.x == 0) s_dry = 0u;
        __syncthreads();
    }
    WorkClaim work(counter, xcd, chunk, total, &s_dry);
    if constexpr (G::kMode == kPhantom)
        persistent_compact<G, ANY, STACK, IM>(tp, work, lds, src, snk);
    else
        persistent_generic<G, ANY, STACK, IM>(tp, work, lds, src, snk);
}

template <class G, bool ANY, int STACK, int IM>
__global__ __launch_bounds__(kTraceBlock) void k_trace_rays_persistent(TraceParams tp, const float* __restrict__ rays,
                                                                       uint32_t n, HitRec* __restrict__ hits,
                                                                       uint32_t* __restrict__ counter) {
    __shared__ uint2 stk[STACK * kTraceBlock];
    persistent_lanes<G, ANY, STACK, IM>(
        tp, n, counter, stk + threadIdx.x,
        [&](uint32_t i, uint32_t& id, f3& o, f3& d, float& tmin, float& tmax) {
            id = i;
            const float4 r0 = reinterpret_cast<const float4*>(rays)[size_t(i) * 2];
            const float4 r1 = reinterpret_cast<const float4*>(rays)[size_t(i) * 2 + 1];
            o = f3{r0.x, r0.y, r0.z};
            d = f3{r1.x, r1.y, r1.z};
            tmin = r0.w;
            tmax = r1.w;
        },
        [&](uint32_t i, f3 o, f3 d, const Best& best) {
            HitRec h;
            finish_hit(tp, o, d, best, h);
            hits[i] = h;
            if (G::kPolish) tp.keys[i] = best_key(tp, best);
        },
        tp.xcd_work);
}

// ------------------------------------------------------------------ deferred FP64 polish

// Ray sources for candidate evaluation: the ray id stored with a candidate -> world ray.
struct SrcRays {
    const float4* rays;
    __device__ __forceinline__ void get(uint32_t i, f3& o, f3& d, float& tmin, float& tmax) const {
        const float4 r0 = rays[size_t(i) * 2], r1 = rays[size_t(i) * 2 + 1];
        o = f3{r0.x, r0.y, r0.z};
        d = f3{r1.x, r1.y, r1.z};
        tmin = r0.w;
        tmax = r1.w;
    }
};
struct SrcCamera {
    Camera cam;
    __device__ __forceinline__ void get(uint32_t pix, f3& o, f3& d, float& tmin, float& tmax) const {
        primary_ray(cam, pix % cam.width, pix / cam.width, o, d);
        tmin = 0.001f;
        tmax = 10000.0f;
    }
};
// Double-word re-run of each final phantom hit (VKHRT_*_ACCURATE_NORMALS): the ray from `src`,
// in the instance's object space, against the hit's curve record; t, u and the normal are
// replaced when the re-run converges. Misses, other modes and FP64-polished roots are kept.
template <class Src>
__device__ __forceinline__ bool accurate_refine(const TraceParams& tp, const Src& src, uint32_t i, HitRec& h) {
    if (h.prim_id == 0xFFFFFFFFu || h.v != 0.0f || h.inst_id >= tp.ninst) return false;
    const InstanceDev& I = tp.inst[h.inst_id];
    if (I.blas.mode != kPhantom || !I.blas.slot_of || h.prim_id >= I.blas.nprims) return false;
    f3 o, d;
    float tmin, tmax;
    src.get(i, o, d, tmin, tmax);
    if (!I.identity) {
        o = xform_pt(I.inv, o);
        d = xform_dir(I.inv, d);
    }
    const uint32_t slot = I.blas.slot_of[h.prim_id];
    if (slot >= I.blas.nprims) return false;
    PhantomHit ph;
    if (!prhi_ff(o, d, I.blas.prims + size_t(slot) * 4, int(I.blas.radius_mode), I.blas.radius, ph)) return false;
    f3 nn{ph.ncx, ph.ncy, ph.nsz};
    if (!I.identity) {
        const float* x = I.xf;
        nn = nrm3(f3{x[0] * nn.x + x[1] * nn.y + x[2] * nn.z, x[4] * nn.x + x[5] * nn.y + x[6] * nn.z,
                     x[8] * nn.x + x[9] * nn.y + x[10] * nn.z});
    }
    h.t = ph.t;
    h.u = ph.u;
    h.n[0] = nn.x;
    h.n[1] = nn.y;
    h.n[2] = nn.z;
    return true;
}
// The double-word re-run of ray i's final hit (not FP64-polished, phantom): t, u and the
// normal are replaced; records it does not apply to are left as they are.
template <class Src>
__device__ __forceinline__ void accurate_one(const TraceParams& tp, const Src& src, uint32_t i, HitRec* hits) {
    HitRec h = hits[i];
    if (accurate_refine(tp, src, i, h)) hits[i] = h;
}
template <class Src>
__global__ __launch_bounds__(kTraceBlock) void k_accurate_hits(TraceParams tp, Src src, HitRec* __restrict__ hits,
                                                               uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) accurate_one(tp, src, i, hits);
}

// Shadow queue entries are ids pixel * samples + light sample.
struct SrcShadow {
    Camera cam;
    const uint32_t* queue;
    const float4* hitbuf;
    AreaLight light;
    __device__ __forceinline__ void get(uint32_t qi, f3& o, f3& d, float& tmin, float& tmax) const {
        const uint32_t id = queue[qi];
        const uint32_t pix = light.samples == 1 ? id : id / light.samples;
        f3 po, pd;
        primary_ray(cam, pix % cam.width, pix / cam.width, po, pd);
        o = add(po, scl(hitbuf[pix].w, pd));
        light_ray(light, pix, id - pix * light.samples, o, d, tmax);
        tmin = 0.001f;
    }
};

// Full Prhi (FP32 sides and the FP64 continuation) for one queued (ray, instance, slot).
__device__ __forceinline__ bool eval_candidate(const TraceParams& tp, uint4 c, f3 ow, f3 dw, float tmin, float tmax,
                                               HitRec& h, unsigned long long& key) {
    const InstanceDev& I = tp.inst[c.y];
    f3 oo = ow, dd = dw;
    if (!I.identity) {
        oo = xform_pt(I.inv, ow);
        dd = xform_dir(I.inv, dw);
    }
    RayFrame F;
    make_frame(oo, dd, F);
    const float4* pr = I.blas.prims + size_t(c.z) * 4;
    float4 cv[4] = {pr[0], pr[1], pr[2], pr[3]};
    PhantomHit ph;
    const bool hit = I.blas.radius_mode == kRadiusPerPoint
                         ? prhi<kRadiusPerPoint, kPolishInline>(F, cv, I.blas.aux[c.z], I.blas.radius, ph)
                         : prhi<kRadiusConstant, kPolishInline>(F, cv, I.blas.aux[c.z], I.blas.radius, ph);
    if (!hit || !(ph.t >= tmin && ph.t <= tmax)) return false;
    const uint32_t pid = I.blas.prim_ids[c.z];
    key = pack_key(ph.t, I.blas.gbase + pid);
    f3 n = phantom_normal(dd, ph);
    if (!I.identity) {
        const float* x = I.xf;
        n = nrm3(f3{x[0] * n.x + x[1] * n.y + x[2] * n.z, x[4] * n.x + x[5] * n.y + x[6] * n.z,
                    x[8] * n.x + x[9] * n.y + x[10] * n.z});
    }
    h.t = ph.t;
    h.prim_id = pid;
    h.inst_id = c.y;
    h.u = ph.u;
    h.n[0] = n.x;
    h.n[1] = n.y;
    h.n[2] = n.z;
    h.v = ph.polished ? 1.0f : 0.0f;
    return true;
}

// Closest-hit merge: the candidate's key competes with the ray's FP32 result by atomicMin.
template <class Src>
__global__ void k_polish_closest(TraceParams tp, Src src, HitRec* __restrict__ cand_hits,
                                 unsigned long long* __restrict__ cand_keys) {
    const uint32_t n = min(*tp.defer_count, tp.defer_cap);
    for (uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x; ci < n; ci += gridDim.x * blockDim.x) {
        const uint4 c = tp.defer[ci];
        f3 o, d;
        float tmin, tmax;
        src.get(c.x, o, d, tmin, tmax);
        HitRec h;
        unsigned long long key = ~0ull;
        if (eval_candidate(tp, c, o, d, tmin, tmax, h, key)) {
            atomicMin(&tp.keys[c.x], key);
            cand_hits[ci] = h;
        }
        cand_keys[ci] = key;
    }
}

// Accurate render with the FP64 polish: the closest-hit merge of the deferred candidates
// (k_polish_closest's work) and the double-word re-run of the queued primary hits in one
// launch, so the candidates' long FP64 chains run beside the re-run instead of before it. The
// first `polish_blocks` blocks take the candidates, the others the queue. The queue was
// gathered before the merge, from the FP32 hits: a queued hit that a polished candidate later
// replaces is re-run in vain (k_polish_fixup_accurate overwrites it), and a candidate that
// wins unpolished (its FP64 side failed and the other side converged in FP32) is re-run there,
// so every final unpolished phantom hit gets the re-run, as when the merge comes first.
// Held to four waves per SIMD (128 registers, the re-run's occupancy on its own): without the
// bound the FP64 candidate path and the double-word re-run share one allocation of 162.
template <class Src>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(4)))
void k_accurate_polish(TraceParams tp, Src src, HitRec* __restrict__ hits, const uint32_t* __restrict__ queue,
                       const uint32_t* __restrict__ count, HitRec* __restrict__ cand_hits,
                       unsigned long long* __restrict__ cand_keys, uint32_t polish_blocks) {
    if (blockIdx.x < polish_blocks) {
        const uint32_t n = min(*tp.defer_count, tp.defer_cap);
        for (uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x; ci < n; ci += polish_blocks * blockDim.x) {
            const uint4 c = tp.defer[ci];
            f3 o, d;
            float tmin, tmax;
            src.get(c.x, o, d, tmin, tmax);
            HitRec h;
            unsigned long long key = ~0ull;
            if (eval_candidate(tp, c, o, d, tmin, tmax, h, key)) {
                atomicMin(&tp.keys[c.x], key);
                cand_hits[ci] = h;
            }
            cand_keys[ci] = key;
        }
        return;
    }
    const uint32_t qi = (blockIdx.x - polish_blocks) * blockDim.x + threadIdx.x;
    if (qi >= *count) return;
    accurate_one(tp, src, queue[qi], hits);
}

// The candidate that won the merge writes its hit record.
__global__ void k_polish_fixup(TraceParams tp, const HitRec* __restrict__ cand_hits,
                               const unsigned long long* __restrict__ cand_keys, HitRec* __restrict__ out) {
    const uint32_t n = min(*tp.defer_count, tp.defer_cap);
    for (uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x; ci < n; ci += gridDim.x * blockDim.x) {
        const unsigned long long k = cand_keys[ci];
        const uint4 c = tp.defer[ci];
        if (k != ~0ull && k == tp.keys[c.x]) out[c.x] = cand_hits[ci];
    }
}

// k_polish_fixup after k_accurate_polish: a winner that the FP64 continuation did not polish
// gets the double-word re-run the queued hits had.
template <class Src>
__global__ void k_polish_fixup_accurate(TraceParams tp, Src src, const HitRec* __restrict__ cand_hits,
                                        const unsigned long long* __restrict__ cand_keys, HitRec* __restrict__ out) {
    const uint32_t n = min(*tp.defer_count, tp.defer_cap);
    for (uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x; ci < n; ci += gridDim.x * blockDim.x) {
        const unsigned long long k = cand_keys[ci];
        const uint4 c = tp.defer[ci];
        if (k == ~0ull || k != tp.keys[c.x]) continue;
        HitRec h = cand_hits[ci];
        accurate_refine(tp, src, c.x, h);  // unpolished winners only (h.v == 0)
        out[c.x] = h;
    }
}

// Shadow rays: any polished hit in range occludes. A ray the traversal already found occluded
// (an FP32 hit after the candidate was deferred) needs no FP64 continuation.
__global__ void k_polish_shadow(TraceParams tp, SrcShadow src, float* __restrict__ vis) {
    const uint32_t n = min(*tp.defer_count, tp.defer_cap);
    for (uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x; ci < n; ci += gridDim.x * blockDim.x) {
        const uint4 c = tp.defer[ci];
        const uint32_t id = src.queue[c.x];
        if (vis[id] == 0.0f) continue;
        f3 o, d;
        float tmin, tmax;
        src.get(c.x, o, d, tmin, tmax);
        HitRec h;
        unsigned long long key;
        if (eval_candidate(tp, c, o, d, tmin, tmax, h, key)) vis[id] = 0.0f;
    }
}

// Persistent render, pass 1: primary rays (FP32) -> per-pixel hit record (and the merge key
// when grazing roots are deferred to the FP64 polish).
// Ordered append from a whole block with one atomic: every thread whose `flag` is set gets a
// slot; slots follow thread order within the block (one counter update per block instead of
// one per wavefront keeps the single queue counter from serializing large grids). Called by
// all threads of the block.
// With m > 1 each flagged thread gets m consecutive slots (the returned one is the first).
constexpr int kAppendBlock = 1024;
__device__ __forceinline__ uint32_t block_append(bool flag, uint32_t* counter, uint32_t m = 1) {
    __shared__ uint32_t s_wave[kAppendBlock / 64 + 1];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t bal = __ballot(flag);
    if (lane == 0) s_wave[wave] = uint32_t(__popcll(bal));
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t nw = (blockDim.x + 63u) / 64u;
        uint32_t sum = 0;
        for (uint32_t w = 0; w < nw; ++w) {
            const uint32_t c = s_wave[w];
            s_wave[w] = sum;
            sum += c;
        }
        s_wave[kAppendBlock / 64] = sum ? atomicAdd(counter, sum * m) : 0u;
    }
    __syncthreads();
    const uint32_t slot = s_wave[kAppendBlock / 64] + (s_wave[wave] + uint32_t(__popcll(bal & ((1ull << lane) - 1ull)))) * m;
    __syncthreads();
    return slot;
}

// Ordered append of kAppendItems work items per thread with one atomic per block: item i of
// thread t is work index blockIdx.x * kAppendBlock * kAppendItems + i * kAppendBlock + t, and
// the slots follow that work order. Every kernel of a frame that appends this way ends with all
// its blocks updating the one queue counter at about the same time; device-wide atomics on one
// address serialize, so fewer, larger blocks shorten the kernel. Called by all threads of the
// block (kAppendBlock threads).
constexpr int kAppendItems = 4;
__device__ __forceinline__ void block_append_items(const bool (&flag)[kAppendItems], uint32_t* counter, uint32_t m,
                                                   uint32_t (&slot)[kAppendItems]) {
    constexpr uint32_t NW = kAppendBlock / 64;
    static_assert(kAppendItems * NW <= 64, "one wavefront scans the per-wave counts");
    __shared__ uint32_t s_cnt[kAppendItems * NW + 1];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    uint64_t bal[kAppendItems];
#pragma unroll
    for (int i = 0; i < kAppendItems; ++i) {
        bal[i] = __ballot(flag[i]);
        if (lane == 0) s_cnt[i * NW + wave] = uint32_t(__popcll(bal[i]));
    }
    __syncthreads();
    if (wave == 0) {
        // exclusive scan of the counts in work order (item-major, then wavefront)
        const uint32_t c = lane < kAppendItems * NW ? s_cnt[lane] : 0u;
        uint32_t v = c;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t u = __shfl_up(v, off);
            if (lane >= uint32_t(off)) v += u;
        }
        const uint32_t sum = __shfl(v, 63);
        uint32_t base = 0;
        if (lane == 0 && sum) base = atomicAdd(counter, sum * m);
        base = __shfl(base, 0);
        if (lane < kAppendItems * NW) s_cnt[lane] = v - c;
        if (lane == 0) s_cnt[kAppendItems * NW] = base;
    }
    __syncthreads();
    const uint32_t base = s_cnt[kAppendItems * NW];
#pragma unroll
    for (int i = 0; i < kAppendItems; ++i)
        slot[i] = base + (s_cnt[i * NW + wave] + uint32_t(__popcll(bal[i] & below))) * m;
    __syncthreads();
}

// Persistent render, pass 0: the primary rays that enter an instance box are appended (in
// work order) to the live queue, so the persistent lanes only take rays that enter the scene,
// and marked in the live mask (bit r & 63 of word r >> 6 for work index r). The others are
// misses: classify reads the mask and writes their pixels without a hit record.
__global__ __launch_bounds__(kAppendBlock) void k_render_live(RenderParams rp) {
    bool live[kAppendItems];
    uint32_t pixs[kAppendItems];
#pragma unroll
    for (int k = 0; k < kAppendItems; ++k) {
        const uint32_t r = (blockIdx.x * kAppendItems + k) * kAppendBlock + threadIdx.x;
        uint32_t x = 0, y = 0;
        const bool valid = r < rp.cam.work && work_pixel(rp.cam, r, x, y);
        const uint32_t pix = y * rp.cam.width + x;
        bool lv = false;
        if (valid) {
            f3 o, d;
            primary_ray(rp.cam, x, y, o, d);
            for (uint32_t i = 0; i < rp.tp.ninst && !lv; ++i) {
                const InstanceDev& I = rp.tp.inst[i];
                lv = I.blas.nprims != 0 && ray_box(o, d, I.lo, I.hi, 0.001f, 10000.0f);
            }
        }
        // a wavefront's lanes hold 64 consecutive work indexes, the first a multiple of 64
        const uint64_t bits = __ballot(lv);
        if ((threadIdx.x & 63u) == 0u && r < rp.cam.work) rp.live_mask[r >> 6] = bits;
        live[k] = lv;
        pixs[k] = pix;
    }
    uint32_t slot[kAppendItems];
    block_append_items(live, &rp.counters[6], 1u, slot);
#pragma unroll
    for (int k = 0; k < kAppendItems; ++k)
        if (live[k]) rp.live[slot[k]] = pixs[k];
}

template <class G, int STACK, int IM>
__global__ __launch_bounds__(kTraceBlock) void k_render_primary_defer(RenderParams rp) {
    __shared__ uint2 stk[STACK * kTraceBlock];
    persistent_lanes<G, false, STACK, IM>(
        rp.tp, rp.counters[6], rp.counters + 1, stk + threadIdx.x,
        [&](uint32_t r, uint32_t& pix, f3& o, f3& d, float& tmin, float& tmax) {
            pix = rp.live[r];
            const uint32_t y = pix < (1u << 24) ? udiv_small(pix, rp.cam.width, rp.cam.inv_w) : pix / rp.cam.width;
            primary_ray(rp.cam, pix - y * rp.cam.width, y, o, d);
            tmin = 0.001f;
            tmax = 10000.0f;
        },
        [&](uint32_t pix, f3 o, f3 d, const Best& best) {
            HitRec h;
            finish_hit(rp.tp, o, d, best, h);
            rp.hitrec[pix] = h;
            if (G::kPolish) rp.tp.keys[pix] = best_key(rp.tp, best);
        },
        rp.counters + kXcdCounters);
}

// Pass 2 (after the optional polish merge): shade misses / palette / unshadowed pixels, queue
// the rest for shadow rays.
// Walks the pixels in the primary pass's work order, so the shadow queue keeps its coherence.
__global__ __launch_bounds__(kAppendBlock) void k_render_classify(RenderParams rp) {
    // the primary pass's deferred candidates are merged: the shadow pass starts a new list
    if (blockIdx.x == 0 && threadIdx.x == 0 && rp.tp.defer) *rp.tp.defer_count = 0u;
    bool hits[kAppendItems];
    uint32_t xs[kAppendItems], ys[kAppendItems];
    float hts[kAppendItems];
#pragma unroll
    for (int k = 0; k < kAppendItems; ++k) {
        const uint32_t r = (blockIdx.x * kAppendItems + k) * kAppendBlock + threadIdx.x;
        uint32_t x = 0, y = 0;
        const bool valid = r < rp.cam.work && work_pixel(rp.cam, r, x, y);
        const uint32_t pix = y * rp.cam.width + x;
        bool hit = false;
        float ht = 0.0f;
        // the same word for the whole wavefront (64 consecutive work indexes)
        const uint64_t mask = r < rp.cam.work ? rp.live_mask[r >> 6] : 0ull;
        if (valid) {
            HitRec h;
            if ((mask >> (r & 63u)) & 1ull) {
                h = rp.hitrec[pix];
            } else {  // the ray misses every instance box
                h.t = 0.0f;
                h.prim_id = h.inst_id = 0xFFFFFFFFu;
                h.u = h.v = 0.0f;
                h.n[0] = h.n[1] = h.n[2] = 0.0f;
            }
            hit = h.prim_id != 0xFFFFFFFFu;
            ht = h.t;
            if (rp.prim_out) rp.prim_out[pix] = h.prim_id;
            const f3 n = f3{h.n[0], h.n[1], h.n[2]};
            if (!hit) {
                f3 c{0.0f, 0.0f, 0.0f};
                if (rp.env.texels) {
                    f3 o, d;
                    primary_ray(rp.cam, x, y, o, d);
                    c = env_color(rp.env, d);
                }
                store_px(rp, pix, c);
            } else if (rp.debug_palette) {
                store_px(rp, pix, debug_color(h.prim_id));
                hit = false;
            } else if (!rp.shadows) {
                store_px(rp, pix, shade(n, 1.0f));
                hit = false;
            } else {
                rp.hitbuf[pix] = make_float4(n.x, n.y, n.z, h.t);
            }
        }
        hits[k] = hit;
        xs[k] = x;
        ys[k] = y;
        hts[k] = ht;
    }
    const uint32_t S = rp.light.samples;
    uint32_t slots[kAppendItems];
    block_append_items(hits, &rp.counters[0], S, slots);
#pragma unroll
    for (int k = 0; k < kAppendItems; ++k) {
        if (!hits[k]) continue;
        const uint32_t x = xs[k], y = ys[k], pix = y * rp.cam.width + x, slot = slots[k];
        for (uint32_t j = 0; j < S; ++j) rp.queue[slot + j] = pix * S + j;
        if (rp.mlp) {
            // neural form: the MLP's inputs of each queued sample, as SrcShadow::get forms the
            // shadow ray (origin on the primary hit, direction toward the light sample)
            f3 po, pd;
            primary_ray(rp.cam, x, y, po, pd);
            const f3 o = add(po, scl(hts[k], pd));
            for (uint32_t j = 0; j < S; ++j) {
                f3 d;
                float tmax;
                light_ray(rp.light, pix, j, o, d, tmax);
                rp.origins[slot + j] = make_float4(o.x, o.y, o.z, 0.0f);
                if (rp.dirs) rp.dirs[slot + j] = make_float4(d.x, d.y, d.z, 0.0f);
            }
        }
    }
}

// Pass 3: shadow rays (FP32, any hit) -> per-pixel visibility.
template <class G, int STACK, int IM>
__global__ __launch_bounds__(kTraceBlock) void k_render_shadow_defer(RenderParams rp, const uint32_t* queue,
                                                                     const uint32_t* count, uint32_t* work,
                                                                     uint32_t skip_upto) {
    __shared__ uint2 stk[STACK * kTraceBlock];
    const uint32_t total = *count;
    if (total <= skip_upto) return;  // a short queue is walked by k_render_shadow_groups
    const SrcShadow src{rp.cam, queue, rp.hitbuf, rp.light};
    persistent_lanes<G, true, STACK, IM>(
        rp.tp, total, work, stk + threadIdx.x,
        [&](uint32_t qi, uint32_t& id, f3& o, f3& d, float& tmin, float& tmax) {
            id = qi;
            src.get(qi, o, d, tmin, tmax);
        },
        [&](uint32_t qi, f3 o, f3 d, const Best& best) {
            rp.vis[queue[qi]] = best.inst == 0xFFFFFFFFu ? 1.0f : 0.0f;
        },
        rp.counters + kXcdCounters + 8 * kXcdStride);
}

// Accurate hits of the render: the primary phantom hits (not FP64-polished) gathered in work
// order from the live rays, so the double-word re-run fills whole wavefronts.
__global__ __launch_bounds__(kAppendBlock) void k_accurate_gather(RenderParams rp) {
    bool take[kAppendItems];
    uint32_t pixs[kAppendItems];
    const uint32_t nlive = rp.counters[6];
#pragma unroll
    for (int k = 0; k < kAppendItems; ++k) {
        const uint32_t r = (blockIdx.x * kAppendItems + k) * kAppendBlock + threadIdx.x;
        take[k] = false;
        pixs[k] = 0u;
        if (r < nlive) {
            const uint32_t pix = rp.live[r];
            const HitRec& h = rp.hitrec[pix];
            take[k] = h.prim_id != 0xFFFFFFFFu && h.v == 0.0f && h.inst_id < rp.tp.ninst &&
                      rp.tp.inst[h.inst_id].blas.mode == kPhantom;
            pixs[k] = pix;
        }
    }
    uint32_t slot[kAppendItems];
    block_append_items(take, &rp.counters[7], 1u, slot);
#pragma unroll
    for (int k = 0; k < kAppendItems; ++k)
        if (take[k]) rp.queue[slot[k]] = pixs[k];
}

template <class Src>
__global__ __launch_bounds__(kTraceBlock) void k_accurate_queued(TraceParams tp, Src src, HitRec* __restrict__ hits,
                                                                 const uint32_t* __restrict__ queue,
                                                                 const uint32_t* __restrict__ count) {
    const uint32_t qi = blockIdx.x * blockDim.x + threadIdx.x;
    if (qi >= *count) return;
    accurate_one(tp, src, queue[qi], hits);
}

// ------------------------------------------------------------------ lane groups per shadow ray
// Any-hit traversal of a short shadow queue (the neural form's exact fallback). Persistent lanes
// take one ray each and hide memory latency behind the other rays of a long queue; with a few
// thousand rays the frame waits for the slowest ray, whose walk is a chain of dependent loads.
// Here a group of LPR lanes walks one ray, 64 / LPR rays per wavefront. Each iteration a group
// pops up to LPR nodes of its ray from its stack in LDS and takes up to LPR pending leaf records
// from its record queue in LDS, one per lane, and issues the loads of both together: an
// iteration costs one memory round trip. Node tests push the internal children they hit and
// append their leaf records; records run the early-out test, and the (group, record) pairs
// that pass wait in a queue of the wavefront. Prhi runs in rounds over that queue, one pair per
// lane at full wave width, when 64 pairs wait or no group can step on without their results (on
// the 5k-ray fallback: 59 us, against 60, 71 and 67 us for rounds from 32, 16 and 8 pairs); a
// grazing root runs the FP64 continuation at once (kPolishInline: the few thousand walks have
// the latency to spare, and the FP64 polish launch after them finds nothing to do). Without the
// queue every group paid for any group's Prhi call: a wavefront of 8 rays ran Prhi in most
// iterations. A walk ends at the first hit in [tmin, tmax] (occluded; its waiting pairs are
// dropped) or when stack, record queue and waiting pairs are empty in every instance the ray
// enters (visible); the group then takes its next ray. The visibility equals the persistent
// kernel's for every ray. On the 16M-segment area-light frame's fallback (5k rays) the walk
// takes 62 us with 16 lanes per ray (the default) and 86-100 us with 8 (deferred Prhi results
// let occluded walks run longer); a whole wavefront per ray without the pair queue took 84-97 us.
//
// Nodes are popped only while the record queue has room for all they can append (24 records
// per node: 8 slots of up to 3) and the stack has room for their children (a popped node adds
// at most 7 entries). Past kStack - kGrpReserve a group pops one node at a time, a depth-first
// walk whose stack grows by at most 7 entries per level, so the reserve covers trees 32 levels
// deep; past kStack the walk reports an overflow.
constexpr uint32_t kGrpReserve = 7u * 32u + 8u;
constexpr uint32_t kGrpNodeRecords = 24;  // leaf records one node can append
constexpr uint32_t kGrpGrid = 8192;       // one-wave blocks of the launch
constexpr uint32_t kGrpRound = 64;        // waiting pairs that start a Prhi round (a full wavefront)
constexpr uint32_t kGrpSlotBits = 26;     // pair = group << 26 | leaf slot (phantom BLASes stay below 2^26 records)
// Value of lane (l - k) for lane l, for the prefix sums over a lane group of LPR lanes (the
// caller uses it only where l - k lies in l's group). A group of 8 or 16 lanes lies inside one
// 16-lane row, so its shifts go through the DPP row shift (no LDS round trip; lanes whose
// source is outside the row read 0). A row shift does not cross rows, so groups of 32 or 64
// lanes, whose lanes take sources from the row below, use the LDS permute for every shift.
template <int LPR>
__device__ __forceinline__ uint32_t group_shift_up(uint32_t v, int k) {
    if (LPR > 16) return uint32_t(__shfl_up(int(v), k));
    switch (k) {
        case 1: return uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x111, 0xF, 0xF, true));  // row_shr:1
        case 2: return uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x112, 0xF, 0xF, true));  // row_shr:2
        case 4: return uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x114, 0xF, 0xF, true));  // row_shr:4
        case 8: return uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x118, 0xF, 0xF, true));  // row_shr:8
        default: return uint32_t(__shfl_up(int(v), k));
    }
}
// The value of the last lane of lane group `grp` (LPR lanes), in every lane: one scalar
// read per group and a select, no LDS round trip.
template <int LPR>
__device__ __forceinline__ uint32_t group_last(uint32_t v, uint32_t grp) {
    uint32_t r = 0;
#pragma unroll
    for (int g = 0; g < 64 / LPR; ++g) {
        const uint32_t t = uint32_t(__builtin_amdgcn_readlane(int(v), g * LPR + LPR - 1));
        r = grp == uint32_t(g) ? t : r;
    }
    return r;
}
template <int LPR>
struct GroupShape {
    static constexpr uint32_t kGroups = 64u / LPR;  // rays per wavefront
    static constexpr uint32_t kStack = LPR >= 32 ? 1024u : (LPR == 16 ? 512u : 384u);
    static constexpr uint32_t kLeaves = LPR >= 32 ? 512u : (LPR == 16 ? 512u : 256u);
};
// A group's ray as the Prhi rounds read it: frame, interval and the instance's records.
struct GroupRay {
    RayFrame F;
    float tmin, tmax, radius;
    const float4* prims;
};

template <class G, int IM, int LPR>
__global__ __launch_bounds__(64) void k_render_shadow_groups(RenderParams rp, const uint32_t* queue, const uint32_t* count,
                                                             uint32_t limit) {
    static_assert(G::kMode == kPhantom && !G::kMixed, "lane-group shadows serve phantom TLASes");
    static_assert(LPR == 8 || LPR == 16 || LPR == 32 || LPR == 64, "lanes per ray: 8, 16, 32 or 64");
    using S = GroupShape<LPR>;
    constexpr uint32_t NG = S::kGroups;
    __shared__ uint32_t s_stack[NG][S::kStack];
    __shared__ uint32_t s_leaf[NG][S::kLeaves];
    __shared__ uint32_t s_pairs[128];  // waiting (group, slot) pairs: fewer than 64 + 64
    __shared__ GroupRay s_ray[NG];
    __shared__ uint32_t s_hit[NG], s_done[NG];
    const uint32_t total = *count;
    if (total == 0 || total > limit) return;
    const TraceParams& tp = rp.tp;
    const uint32_t lane = threadIdx.x, grp = lane / LPR, sub = lane % LPR, last = grp * LPR + LPR - 1u;
    const uint64_t gmask = (LPR == 64 ? ~0ull : ((1ull << LPR) - 1ull)) << (grp * LPR);
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t* stk = s_stack[grp];
    uint32_t* lf = s_leaf[grp];
    const SrcShadow src{rp.cam, queue, rp.hitbuf, rp.light};
    const uint32_t qstride = gridDim.x * NG;
    if (lane < NG) s_hit[lane] = s_done[lane] = 0u;
#ifdef VKHRT_WORK_COUNTERS
    unsigned long long wc[7] = {};
    unsigned long long t_ray = 0, t_max = 0;
    // per wavefront: cycles in Prhi rounds, rounds, loop iterations, cycles of the whole loop
    unsigned long long c_round = 0, n_round = 0, n_iter = 0;
    const unsigned long long t_loop = clock64();
#endif
    // group state: the same value in every lane of the group
    uint32_t qi = blockIdx.x * NG + grp;
    bool have_ray = false, walking = false, occluded = false, error = false;
    uint32_t inst = 0, sp = 0, lh = 0, lt = 0, pend = 0;
    f3 ow{}, dw{}, o{}, d{}, inv{};
    float tmin = 0.0f, tmax = 0.0f;
    uint32_t oct = 0;
    BlasDev B{};
    uint32_t npairs = 0;  // wave-uniform
    while (true) {
        // per group: take a ray, then enter its next instance (or finish the ray)
        if (!have_ray && qi < total) {
            src.get(qi, ow, dw, tmin, tmax);
            have_ray = true;
            inst = 0;
            occluded = error = false;
#ifdef VKHRT_WORK_COUNTERS
            t_ray = clock64();
            if (sub == 0) ++wc[0];
#endif
        }
        while (have_ray && !walking) {
            if (inst < tp.ninst && !occluded && !error) {
                const InstanceDev& I = tp.inst[inst];
                if (I.blas.nprims == 0 || !ray_box(ow, dw, I.lo, I.hi, tmin, tmax)) {
                    ++inst;
                    continue;
                }
                B = IM == kInstOne ? wave_uniform(I.blas) : I.blas;
                o = ow;
                d = dw;
                if (!I.identity) {
                    o = xform_pt(I.inv, ow);
                    d = xform_dir(I.inv, dw);
                }
                inv = safe_inv(d);
                oct = uint32_t(signbit(d.x)) | (uint32_t(signbit(d.y)) << 1) | (uint32_t(signbit(d.z)) << 2);
                if (sub == 0) {
                    GroupRay& R = s_ray[grp];
                    make_frame(o, d, R.F);
                    R.tmin = tmin;
                    R.tmax = tmax;
                    R.radius = B.radius;
                    R.prims = B.prims;
                    stk[0] = 0u;  // the root
                }
                sp = 1;
                lh = lt = 0;
                pend = 0;
                walking = true;
            } else {
                if (error && sub == 0) atomicAdd(&tp.status[0], 1u);
                if (sub == 0) rp.vis[queue[qi]] = occluded ? 0.0f : 1.0f;
#ifdef VKHRT_WORK_COUNTERS
                if (sub == 0) {
                    const unsigned long long c = clock64() - t_ray;
                    wc[6] += c;
                    t_max = c > t_max ? c : t_max;
                }
#endif
                have_ray = false;
                qi += qstride;
                if (qi < total) {
                    src.get(qi, ow, dw, tmin, tmax);
                    have_ray = true;
                    inst = 0;
                    occluded = error = false;
#ifdef VKHRT_WORK_COUNTERS
                    t_ray = clock64();
                    if (sub == 0) ++wc[0];
#endif
                }
            }
        }
        if (__ballot(walking) == 0) break;
#ifdef VKHRT_WORK_COUNTERS
        ++n_iter;
#endif
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // one step of every walking group: its nodes (as many as its record queue and stack can
        // take) and its records
        uint32_t m = 0, nl = 0;
        if (walking) {
            const uint32_t by_stack = sp + kGrpReserve <= S::kStack ? (S::kStack - kGrpReserve - sp) / 7u : 0u;
            const uint32_t by_leaves = (S::kLeaves - lt) / kGrpNodeRecords;
            m = min(min(sp, uint32_t(LPR)), min(by_stack > 1u ? by_stack : 1u, by_leaves));
            nl = min(lt - lh, uint32_t(LPR));
        }
        const uint32_t node = sub < m ? stk[sp - m + sub] : 0xFFFFFFFFu;
        const uint32_t slot = sub < nl ? lf[lh + sub] : 0xFFFFFFFFu;
        if (walking) {
            sp -= m;
            lh += nl;
            if (lh == lt) lh = lt = 0;
        }
        // issue the early-out's loads (the record's end points and radius) before the node test
        float4 e0{}, e3{};
        float rmax = 0.0f;
        if (slot != 0xFFFFFFFFu) {
            const auto* pr = gmem(B.prims + size_t(slot) * 4);
            e0 = pr[0];
            e3 = pr[3];
            rmax = gmem(B.aux)[slot];
        }
        uint32_t hits = 0, pmask = 0, child_base = 0, imask = 0, pbase = 0;
        bool bad = false;
        if (node != 0xFFFFFFFFu) {
            if (node < B.nnodes) {
                node_visit(B.nodes, node, o, inv, oct, tmin, tmax, hits, pmask, child_base, imask, pbase);
                if (pmask && pbase + (32u - __clz(pmask)) > B.nprims) {
                    bad = true;
                    hits = pmask = 0;
                }
            } else {
                bad = true;
            }
        }
        const bool pass = slot != 0xFFFFFFFFu && ray_cylinder(o, d, f3{e0.x, e0.y, e0.z}, f3{e3.x, e3.y, e3.z}, rmax);
        // the pairs that passed join the wavefront's queue (npairs < 64 here, so they fit)
        const uint64_t pb = __ballot(pass);
        if (pass) s_pairs[npairs + uint32_t(__popcll(pb & below))] = (grp << kGrpSlotBits) | slot;
        npairs += uint32_t(__popcll(pb));
        const uint64_t eb = __ballot(bad);
#ifdef VKHRT_WORK_COUNTERS
        if (walking && sub == 0) {
            ++wc[1];
            wc[2] += m;
            wc[3] += nl;
            wc[4] += __popcll(pb & gmask);
            wc[5] += nl ? 1u : 0u;
        }
#endif
        // push the internal children hit (octant bit j -> slot j ^ oct -> child index) and
        // append the leaf records hit (inclusive prefix sums over the group's lanes)
        const uint32_t nc = __popc(hits), np = __popc(pmask);
        uint32_t cpre = nc, ppre = np;
#pragma unroll
        for (int k = 1; k < LPR; k <<= 1) {
            const uint32_t x = group_shift_up<LPR>(cpre, k), y = group_shift_up<LPR>(ppre, k);
            if (int(sub) >= k) {
                cpre += x;
                ppre += y;
            }
        }
        const uint32_t ctot = group_last<LPR>(cpre, grp), ptot = group_last<LPR>(ppre, grp);
        if (walking) {
            pend += uint32_t(__popcll(pb & gmask));
            error = (eb & gmask) != 0 || sp + ctot > S::kStack || lt + ptot > S::kLeaves;
            if (!error) {
                uint32_t w = sp + cpre - nc;
                for (uint32_t h = hits; h; h &= h - 1u) {
                    const uint32_t j = __ffs(h) - 1, sl = j ^ oct;
                    stk[w++] = child_base + __popc(imask & ((1u << sl) - 1u));
                }
                w = lt + ppre - np;
                for (uint32_t pm = pmask; pm; pm &= pm - 1u) lf[w++] = pbase + uint32_t(__ffs(pm) - 1);
                sp += ctot;
                lt += ptot;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // a Prhi round when enough pairs wait, or when no walking group can step on without them
        // (every group with pairs waiting has an empty stack and record queue)
        const bool can_step = walking && !error && (sp > 0 || lh < lt);
        if (npairs >= kGrpRound || (npairs > 0 && __ballot(can_step) == 0)) {
#ifdef VKHRT_WORK_COUNTERS
            const unsigned long long t_round = clock64();
#endif
            const uint32_t k = npairs < 64u ? npairs : 64u;
            if (lane < k) {
                const uint32_t e = s_pairs[lane];
                const uint32_t g = e >> kGrpSlotBits;
                const uint32_t sl = e & ((1u << kGrpSlotBits) - 1u);
                const GroupRay& R = s_ray[g];
                const auto* pr = gmem(R.prims + size_t(sl) * 4);
                float4 cv[4] = {pr[0], pr[1], pr[2], pr[3]};
                PhantomHit ph;
                const bool hit = prhi<G::kRMode, G::kPolish ? kPolishInline : kPolishOff, false>(R.F, cv, 0.0f,
                                                                                                 R.radius, ph) &&
                                 ph.t >= R.tmin && ph.t <= R.tmax;
                if (hit) s_hit[g] = 1u;
                atomicAdd(&s_done[g], 1u);
            }
            // the pairs not taken move to the front ([64, npairs) -> [0, npairs - 64): no overlap)
            const uint32_t rest = npairs - k;
            const uint32_t moved = lane < rest ? s_pairs[64u + lane] : 0u;
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
            if (lane < rest) s_pairs[lane] = moved;
            npairs = rest;
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
            if (walking) {
                occluded = occluded || s_hit[grp] != 0u;
                pend -= s_done[grp];
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
            if (lane < NG) s_hit[lane] = s_done[lane] = 0u;
#ifdef VKHRT_WORK_COUNTERS
            c_round += clock64() - t_round;
            ++n_round;
#endif
        }
        // an occluded or failed walk drops its waiting pairs (the others keep their order)
        const bool drop = walking && (occluded || error);
        if (__ballot(drop) != 0) {
            if (lane < NG) s_hit[lane] = 0u;
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
            if (drop && sub == 0) s_hit[grp] = 1u;  // groups whose pairs go
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
            uint32_t kept_total = 0;
            for (uint32_t base = 0; base < npairs; base += 64u) {
                const uint32_t i = base + lane;
                const uint32_t e = i < npairs ? s_pairs[i] : 0u;
                const bool keep = i < npairs && s_hit[e >> kGrpSlotBits] == 0u;
                const uint64_t kb = __ballot(keep);
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                __builtin_amdgcn_wave_barrier();
                if (keep) s_pairs[kept_total + uint32_t(__popcll(kb & below))] = e;
                kept_total += uint32_t(__popcll(kb));
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                __builtin_amdgcn_wave_barrier();
            }
            npairs = kept_total;
            if (lane < NG) s_hit[lane] = 0u;
            if (drop) pend = 0;
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
        // the instance is done when the ray is occluded, an error stopped it, or nothing is left
        if (walking && (occluded || error || (sp == 0 && lh == lt && pend == 0))) {
            walking = false;
            ++inst;
        }
    }
#ifdef VKHRT_WORK_COUNTERS
    for (int i = 0; i < 7; ++i) {
        unsigned long long v = wc[i];
        for (int k = 32; k >= 1; k >>= 1) v += __shfl_xor(v, k);
        wc[i] = v;
    }
    for (int k = 32; k >= 1; k >>= 1) {
        const unsigned long long v = __shfl_xor(t_max, k);
        t_max = v > t_max ? v : t_max;
    }
    if (lane == 0 && tp.work_counts) {
        for (int i = 0; i < 7; ++i)
            if (wc[i]) atomicAdd(&tp.work_counts[kWorkWave + i], wc[i]);
        atomicMax(&tp.work_counts[kWorkWave + 7], t_max);
        if (n_iter) {
            atomicAdd(&tp.work_counts[kWorkWave + 8], c_round);
            atomicAdd(&tp.work_counts[kWorkWave + 9], n_round);
            atomicAdd(&tp.work_counts[kWorkWave + 10], n_iter);
            atomicAdd(&tp.work_counts[kWorkWave + 11], clock64() - t_loop);
            atomicAdd(&tp.work_counts[kWorkWave + 12], 1ull);
        }
    }
#endif
}

// The same two traversal passes for TLASes that mix geometry modes: one thread per queued
// ray, trace_world's per-instance dispatch (with the FP64 polish candidates of phantom
// instances deferred like the persistent kernels'); the queue lengths are read on the device.
template <class G, int STACK>
__global__ __launch_bounds__(kTraceBlock) void k_render_primary_flat(RenderParams rp) {
    __shared__ uint2 stk[STACK * kTraceBlock];
    const uint32_t r = blockIdx.x * kTraceBlock + threadIdx.x;
    if (r >= rp.counters[6]) return;
    const uint32_t pix = rp.live[r];
    const uint32_t y = pix < (1u << 24) ? udiv_small(pix, rp.cam.width, rp.cam.inv_w) : pix / rp.cam.width;
    f3 o, d;
    primary_ray(rp.cam, pix - y * rp.cam.width, y, o, d);
    Best best;
    if (!trace_world<G, false, STACK>(rp.tp, o, d, 0.001f, 10000.0f, best, stk + threadIdx.x, pix))
        atomicAdd(&rp.tp.status[0], 1u);
    HitRec h;
    finish_hit(rp.tp, o, d, best, h);
    rp.hitrec[pix] = h;
    if (rp.tp.keys) rp.tp.keys[pix] = best_key(rp.tp, best);
}

template <class G, int STACK>
__global__ __launch_bounds__(kTraceBlock) void k_render_shadow_flat(RenderParams rp, const uint32_t* queue,
                                                                    const uint32_t* count) {
    __shared__ uint2 stk[STACK * kTraceBlock];
    const uint32_t qi = blockIdx.x * kTraceBlock + threadIdx.x;
    if (qi >= *count) return;
    f3 o, d;
    float tmin, tmax;
    SrcShadow{rp.cam, queue, rp.hitbuf, rp.light}.get(qi, o, d, tmin, tmax);
    Best best;
    if (!trace_world<G, true, STACK>(rp.tp, o, d, tmin, tmax, best, stk + threadIdx.x, qi))
        atomicAdd(&rp.tp.status[0], 1u);
    rp.vis[queue[qi]] = best.inst == 0xFFFFFFFFu ? 1.0f : 0.0f;
}

// Shades the queued hit pixels with their light visibility: the mean over the pixel's light
// samples (ids pixel * S + s; the entry of sample 0 shades the pixel). A caller's per-pixel
// visibility (light_visibility) comes with S = 1.
__global__ void k_render_apply_visibility(RenderParams rp, const float* __restrict__ vis) {
    const uint32_t qi = blockIdx.x * blockDim.x + threadIdx.x;
    if (qi >= rp.counters[0]) return;
    const uint32_t S = rp.light.samples;
    const uint32_t id = rp.queue[qi];
    const uint32_t pix = S == 1 ? id : id / S;
    if (id != pix * S) return;
    float v = vis[id];
    for (uint32_t k = 1; k < S; ++k) v = v + vis[id + k];
    if (S > 1) v = v / float(S);
    const float4 hb = rp.hitbuf[pix];
    store_px(rp, pix, shade(f3{hb.x, hb.y, hb.z}, v));
}

__global__ void k_zero(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

}  // namespace
}  // namespace trace_detail

// Helpers shared with the auxiliary kernels in trace.hip.
namespace trace_detail_pub {
__device__ __forceinline__ void primary_ray_pub(const Camera& c, uint32_t x, uint32_t y, f3& o, f3& d) {
    trace_detail::primary_ray(c, x, y, o, d);
}
__device__ __forceinline__ f3 xform_pt_pub(const float* m, f3 p) { return trace_detail::xform_pt(m, p); }
__device__ __forceinline__ f3 xform_dir_pub(const float* m, f3 p) { return trace_detail::xform_dir(m, p); }
}  // namespace trace_detail_pub

// Grid of a persistent kernel: enough 256-thread blocks to fill every CU at the occupancy
// the LDS stack allows (4 wavefronts per block, at most 8 blocks per CU).
// `blocks_per_cu` (> 0) caps the blocks per CU below that occupancy limit.
template <class K>
inline dim3 persistent_grid(K kernel, uint32_t blocks_per_cu = 0) {
    int dev = 0, cus = 0, per_cu = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kernel),
                                                 trace_detail::kTraceBlock, 0);
    if (per_cu < 1) per_cu = 1;
    if (blocks_per_cu > 0 && int(blocks_per_cu) < per_cu) per_cu = int(blocks_per_cu);
    if (cus < 1) cus = 1;
    return dim3(uint32_t(cus * per_cu));
}

template <class G, int STACK>
hipError_t launch_trace_pass(const TraceParams& tp, const float* rays, uint32_t n, HitRec* hits, dim3 grid, dim3 block,
                             bool any, hipStream_t st);

// Launchers for one geometry variant and stack depth.
template <class G, int STACK>
hipError_t launch_trace_variant(const TraceParams& tp, const float* rays, uint32_t n, HitRec* hits, uint2* stats,
                                hipStream_t st) {
    using namespace trace_detail;
    if (n == 0) return hipSuccess;
    const dim3 grid((n + kTraceBlock - 1) / kTraceBlock), block(kTraceBlock);
    const bool any = (tp.flags & kTraceAnyHit) != 0;
    if (stats) {
        if constexpr (!G::kMixed && !G::kPolish) {
            if (any)
                hipLaunchKernelGGL((k_trace_rays<G, true, STACK, true>), grid, block, 0, st, tp, rays, n, hits, stats);
            else
                hipLaunchKernelGGL((k_trace_rays<G, false, STACK, true>), grid, block, 0, st, tp, rays, n, hits, stats);
            return hipGetLastError();
        }
        return hipErrorNotSupported;
    }
    if (tp.defer) hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, st, tp.defer_count, 1u);
    hipError_t e = launch_trace_pass<G, STACK>(tp, rays, n, hits, grid, block, any, st);
    if (e == hipSuccess && tp.defer) {
        const dim3 pgrid(1024);
        hipLaunchKernelGGL(k_polish_closest<SrcRays>, pgrid, block, 0, st, tp,
                           SrcRays{reinterpret_cast<const float4*>(rays)}, tp.cand_hits, tp.cand_keys);
        hipLaunchKernelGGL(k_polish_fixup, pgrid, block, 0, st, tp, tp.cand_hits, tp.cand_keys, hits);
        e = hipGetLastError();
    }
    return e;
}

template <class G, int STACK>
hipError_t launch_trace_pass(const TraceParams& tp, const float* rays, uint32_t n, HitRec* hits, dim3 grid, dim3 block,
                             bool any, hipStream_t st) {
    using namespace trace_detail;
    uint2* stats = nullptr;
    if constexpr (!G::kMixed) {
        if (tp.work) {
            hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, st, tp.work, 1u);
            if (tp.xcd_chunk && tp.xcd_work)
                hipLaunchKernelGGL(k_zero, dim3(8 * kXcdStride / 64), dim3(64), 0, st, tp.xcd_work, 8u * kXcdStride);
            // one-instance TLASes take kernels that read the BLAS fields once
            const bool one = instance_mode(tp.ninst) == kInstOne;
            if (any && one)
                hipLaunchKernelGGL((k_trace_rays_persistent<G, true, STACK, kInstOne>),
                                   persistent_grid(k_trace_rays_persistent<G, true, STACK, kInstOne>, tp.blocks_per_cu), block, 0,
                                   st, tp, rays, n, hits, tp.work);
            else if (any)
                hipLaunchKernelGGL((k_trace_rays_persistent<G, true, STACK, kInstMany>),
                                   persistent_grid(k_trace_rays_persistent<G, true, STACK, kInstMany>, tp.blocks_per_cu), block, 0,
                                   st, tp, rays, n, hits, tp.work);
            else if (one)
                hipLaunchKernelGGL((k_trace_rays_persistent<G, false, STACK, kInstOne>),
                                   persistent_grid(k_trace_rays_persistent<G, false, STACK, kInstOne>, tp.blocks_per_cu), block,
                                   0, st, tp, rays, n, hits, tp.work);
            else
                hipLaunchKernelGGL((k_trace_rays_persistent<G, false, STACK, kInstMany>),
                                   persistent_grid(k_trace_rays_persistent<G, false, STACK, kInstMany>, tp.blocks_per_cu), block,
                                   0, st, tp, rays, n, hits, tp.work);
            return hipGetLastError();
        }
    }
    if (any)
        hipLaunchKernelGGL((k_trace_rays<G, true, STACK, false>), grid, block, 0, st, tp, rays, n, hits, stats);
    else
        hipLaunchKernelGGL((k_trace_rays<G, false, STACK, false>), grid, block, 0, st, tp, rays, n, hits, stats);
    return hipGetLastError();
}

// The neural fallback's lane-group walker: rp.group_lanes (8, 16, 32 or 64; 16 otherwise) lanes per ray.
template <class G>
hipError_t launch_shadow_groups(const RenderParams& rp, const uint32_t* squeue, const uint32_t* scount, uint32_t limit,
                                int im, hipStream_t st) {
    using namespace trace_detail;
    const dim3 grid(kGrpGrid), block(64);
#define VK_GROUPS(L)                                                                                               \
    if (im == kInstOne)                                                                                            \
        hipLaunchKernelGGL((k_render_shadow_groups<G, kInstOne, L>), grid, block, 0, st, rp, squeue, scount, limit); \
    else                                                                                                           \
        hipLaunchKernelGGL((k_render_shadow_groups<G, kInstMany, L>), grid, block, 0, st, rp, squeue, scount, limit);
    switch (rp.group_lanes) {
        case 8: VK_GROUPS(8) break;
        case 32: VK_GROUPS(32) break;
        case 64: VK_GROUPS(64) break;
        default: VK_GROUPS(16) break;
    }
#undef VK_GROUPS
    return hipGetLastError();
}

template <class G, int STACK>
hipError_t launch_render_variant(const RenderParams& rp, hipStream_t st, hipEvent_t* ev) {
    using namespace trace_detail;
    const uint32_t npix = rp.cam.width * rp.cam.height;
    if (npix == 0) return hipSuccess;
    const dim3 grid((npix + kTraceBlock - 1) / kTraceBlock), block(kTraceBlock);
    // the shadow queue holds up to npix * samples ids
    const dim3 sgrid(uint32_t((uint64_t(npix) * rp.light.samples + kTraceBlock - 1) / kTraceBlock));
    // the frame's counters, and the per-XCD counters of its two traversal launches
    const uint32_t zero_words = rp.tp.xcd_chunk ? kXcdCounters + 16u * kXcdStride : 8u;
    hipLaunchKernelGGL(k_zero, dim3((zero_words + 255) / 256), dim3(256), 0, st, rp.counters, zero_words);
    {
        if (rp.tp.work) {
            // Deferred pipeline: the traversal kernels only record hits (their sinks stay
            // small because a lane finishing any ray makes the whole wave run the sink);
            // shading, the shadow queue and the optional deferred FP64 polish run as passes.
            // A TLAS of one geometry mode traverses with persistent lanes, a mixed TLAS with one
            // thread per queued ray.
            const TraceParams& tp = rp.tp;
            const dim3 pgrid(1024);
            hipLaunchKernelGGL(k_render_live, dim3((rp.cam.work + kAppendBlock * kAppendItems - 1) / (kAppendBlock * kAppendItems)), dim3(kAppendBlock),
                               0, st, rp);
            // the events bracket the traversal kernels alone (the per-kernel timings)
            if (ev) hipEventRecord(ev[0], st);
            // one-instance TLASes take kernels that read the BLAS fields once, TLASes of up to
            // kFewInstances instances kernels that stage them in LDS
            const int im = instance_mode(tp.ninst);
            if constexpr (G::kMixed)
                hipLaunchKernelGGL((k_render_primary_flat<G, STACK>), grid, block, 0, st, rp);
            else if (im == kInstOne)
                hipLaunchKernelGGL((k_render_primary_defer<G, STACK, kInstOne>),
                                   persistent_grid(k_render_primary_defer<G, STACK, kInstOne>, tp.blocks_per_cu), block, 0,
                                   st, rp);
            else if (im == kInstFew)
                hipLaunchKernelGGL((k_render_primary_defer<G, STACK, kInstFew>),
                                   persistent_grid(k_render_primary_defer<G, STACK, kInstFew>, tp.blocks_per_cu), block, 0,
                                   st, rp);
            else
                hipLaunchKernelGGL((k_render_primary_defer<G, STACK, kInstMany>),
                                   persistent_grid(k_render_primary_defer<G, STACK, kInstMany>, tp.blocks_per_cu), block,
                                   0, st, rp);
            if (ev) hipEventRecord(ev[1], st);
            const dim3 agrid((npix + kAppendBlock * kAppendItems - 1) / (kAppendBlock * kAppendItems));
            if (rp.accurate && tp.defer) {
                // the FP64 merge beside the double-word re-run (one block per CU for the
                // candidates), then the winners' records
                constexpr uint32_t kPolishBlocks = 256;
                hipLaunchKernelGGL(k_accurate_gather, agrid, dim3(kAppendBlock), 0, st, rp);
                hipLaunchKernelGGL(k_accurate_polish<SrcCamera>, dim3(grid.x + kPolishBlocks), block, 0, st, tp,
                                   SrcCamera{rp.cam}, rp.hitrec, rp.queue, rp.counters + 7, tp.cand_hits, tp.cand_keys,
                                   kPolishBlocks);
                hipLaunchKernelGGL(k_polish_fixup_accurate<SrcCamera>, pgrid, block, 0, st, tp, SrcCamera{rp.cam},
                                   tp.cand_hits, tp.cand_keys, rp.hitrec);
            } else {
                if (tp.defer) {
                    hipLaunchKernelGGL(k_polish_closest<SrcCamera>, pgrid, block, 0, st, tp, SrcCamera{rp.cam},
                                       tp.cand_hits, tp.cand_keys);
                    hipLaunchKernelGGL(k_polish_fixup, pgrid, block, 0, st, tp, tp.cand_hits, tp.cand_keys, rp.hitrec);
                }
                if (rp.accurate) {
                    hipLaunchKernelGGL(k_accurate_gather, agrid, dim3(kAppendBlock), 0, st, rp);
                    hipLaunchKernelGGL(k_accurate_queued<SrcCamera>, grid, block, 0, st, tp, SrcCamera{rp.cam},
                                       rp.hitrec, rp.queue, rp.counters + 7);
                }
            }
            hipLaunchKernelGGL(k_render_classify, dim3((rp.cam.work + kAppendBlock * kAppendItems - 1) / (kAppendBlock * kAppendItems)),
                               dim3(kAppendBlock), 0, st, rp);
            bool shadow_timed = false;
            if (rp.shadows && !rp.debug_palette) {
                if (rp.light_vis) {
                    hipLaunchKernelGGL(k_render_apply_visibility, grid, block, 0, st, rp, rp.light_vis);
                } else {
                    // exact: every queued pixel traces its shadow ray; neural: the MLP answers the
                    // confident ones and the rest (fallback queue, counters[5]) trace it
                    const uint32_t* squeue = rp.queue;
                    const uint32_t* scount = rp.counters;
                    if (ev) hipEventRecord(ev[2], st);
                    if (rp.mlp) {
                        MlpShadowParams sp{rp.queue,          rp.counters,      rp.origins,
                                           {0.0f, 1.0f, 0.0f}, rp.mlp_threshold, rp.vis,
                                           rp.fallback_queue, rp.counters + 5, rp.dirs,
                                           rp.light.samples == 2u ? 1u : 0u};
                        const hipError_t e = launch_mlp_shadow_gate(*rp.mlp, sp, npix * rp.light.samples, st);
                        if (e != hipSuccess) return e;
                        squeue = rp.fallback_queue;
                        scount = rp.counters + 5;
                    }
                    // the neural form's fallback queue is short: a queue of up to rp.wave_limit
                    // rays is walked one ray per wavefront, a longer one by the persistent lanes
                    // (both kernels read the length on the device; one of them returns at once)
                    uint32_t skip = 0;
                    if constexpr (!G::kMixed && G::kMode == kPhantom) {
                        if (rp.mlp && rp.wave_limit) {
                            skip = rp.wave_limit;
                            const hipError_t e = launch_shadow_groups<G>(rp, squeue, scount, skip, im, st);
                            if (e != hipSuccess) return e;
                        }
                    }
                    if constexpr (G::kMixed)
                        hipLaunchKernelGGL((k_render_shadow_flat<G, STACK>), sgrid, block, 0, st, rp, squeue, scount);
                    else if (im == kInstOne)
                        hipLaunchKernelGGL((k_render_shadow_defer<G, STACK, kInstOne>),
                                           persistent_grid(k_render_shadow_defer<G, STACK, kInstOne>, tp.blocks_per_cu),
                                           block, 0, st, rp, squeue, scount, rp.counters + 2, skip);
                    else if (im == kInstFew)
                        hipLaunchKernelGGL((k_render_shadow_defer<G, STACK, kInstFew>),
                                           persistent_grid(k_render_shadow_defer<G, STACK, kInstFew>, tp.blocks_per_cu),
                                           block, 0, st, rp, squeue, scount, rp.counters + 2, skip);
                    else
                        hipLaunchKernelGGL((k_render_shadow_defer<G, STACK, kInstMany>),
                                           persistent_grid(k_render_shadow_defer<G, STACK, kInstMany>, tp.blocks_per_cu),
                                           block, 0, st, rp, squeue, scount, rp.counters + 2, skip);
                    if (ev) hipEventRecord(ev[3], st);
                    shadow_timed = true;
                    if (tp.defer)
                        hipLaunchKernelGGL(k_polish_shadow, pgrid, block, 0, st, tp,
                                           SrcShadow{rp.cam, squeue, rp.hitbuf, rp.light}, rp.vis);
                    hipLaunchKernelGGL(k_render_apply_visibility, sgrid, block, 0, st, rp, rp.vis);
                }
            }
            if (ev && !shadow_timed) {
                hipEventRecord(ev[2], st);
                hipEventRecord(ev[3], st);
            }
            return hipGetLastError();
        }
    }
    // one thread per pixel (mixed-mode TLASes, or persistence switched off)
    if (ev) hipEventRecord(ev[0], st);
    hipLaunchKernelGGL((k_render_primary<G, STACK>), grid, block, 0, st, rp);
    if (ev) hipEventRecord(ev[1], st);
    if (ev) hipEventRecord(ev[2], st);
    if (rp.shadows && !rp.debug_palette) {
        if (rp.light_vis)
            hipLaunchKernelGGL(k_render_apply_visibility, grid, block, 0, st, rp, rp.light_vis);
        else
            hipLaunchKernelGGL((k_render_shadow<G, STACK>), grid, block, 0, st, rp);
    }
    if (ev) hipEventRecord(ev[3], st);
    return hipGetLastError();
}

}  // namespace vk

#include "batch_impl.hpp"

namespace vk {

// Explicit instantiation helpers for the trace_*.hip translation units.
#define VKHRT_INSTANTIATE_TRACE(G)                                                                              \
    template hipError_t launch_batch_variant<G, 8>(const BatchParams&, hipStream_t);                            \
    template hipError_t launch_batch_variant<G, 12>(const BatchParams&, hipStream_t);                           \
    template hipError_t launch_batch_variant<G, 16>(const BatchParams&, hipStream_t);                           \
    template hipError_t launch_batch_variant<G, 32>(const BatchParams&, hipStream_t);                           \
    template hipError_t launch_trace_variant<G, 8>(const TraceParams&, const float*, uint32_t, HitRec*, uint2*, \
                                                   hipStream_t);                                                \
    template hipError_t launch_trace_variant<G, 16>(const TraceParams&, const float*, uint32_t, HitRec*, uint2*,\
                                                    hipStream_t);                                               \
    template hipError_t launch_trace_variant<G, 12>(const TraceParams&, const float*, uint32_t, HitRec*, uint2*,\
                                                    hipStream_t);                                               \
    template hipError_t launch_render_variant<G, 8>(const RenderParams&, hipStream_t, hipEvent_t*);             \
    template hipError_t launch_render_variant<G, 16>(const RenderParams&, hipStream_t, hipEvent_t*);            \
    template hipError_t launch_render_variant<G, 12>(const RenderParams&, hipStream_t, hipEvent_t*);            \
    template hipError_t launch_trace_variant<G, 32>(const TraceParams&, const float*, uint32_t, HitRec*, uint2*,\
                                                    hipStream_t);                                               \
    template hipError_t launch_render_variant<G, 32>(const RenderParams&, hipStream_t, hipEvent_t*);

}  // namespace vk
