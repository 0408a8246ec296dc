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

__device__ __forceinline__ bool better(float t, uint32_t inst, uint32_t pid, const Best& B) {
    if (t < B.t) return true;
    if (t == B.t) return inst < B.inst || (inst == B.inst && pid < B.pid);
    return false;
}

// 1/d with zero components replaced by +-1e-20: the quantized slab form q*(s*inv) + (p-o)*inv
// would otherwise produce inf - inf = NaN for axis-parallel rays and drop that slab.
__device__ __forceinline__ float safe_rcp(float x) { return 1.0f / (fabsf(x) < 1e-20f ? copysignf(1e-20f, x) : x); }
__device__ __forceinline__ f3 safe_inv(f3 d) { return f3{safe_rcp(d.x), safe_rcp(d.y), safe_rcp(d.z)}; }

__device__ __forceinline__ float qbyte(uint32_t lo, uint32_t hi, int s) {
    uint32_t w = s < 4 ? lo : hi;
    return float((w >> (8 * (s & 3))) & 0xffu);
}

// Per-lane stack of node groups in LDS (stride kTraceBlock entries). A group is pushed
// only when descending one level, so a tree of depth D never needs more than D entries;
// the host picks STACK >= depth and a push past STACK reports an error instead of spilling.
template <int STACK>
struct LaneStack {
    uint2* lds;
    int sp;
    __device__ __forceinline__ bool push(uint2 g) {
        if (sp >= STACK) return false;
        lds[sp * kTraceBlock] = g;
        ++sp;
        return true;
    }
    __device__ __forceinline__ uint2 top() const { return lds[(sp - 1) * kTraceBlock]; }
};

// group: x = child_base, y = imask | (remaining hit bits in octant order) << 8
// Traverses one BLAS. Returns false on stack overflow.
// Per-ray traversal statistics (nodes visited, primitive records tested).
struct TravStats {
    uint32_t nodes = 0, prims = 0;
};

template <int MODE, int RMODE, bool POLISH, bool ANY, int STACK, bool STATS = false>
__device__ __forceinline__ bool traverse(const BlasDev& B, f3 o, f3 d, float tmin, uint32_t inst, Best& best,
                                         uint2* lds, bool& done, TravStats* st = nullptr) {
    if (B.nprims == 0) return true;
    const f3 inv = safe_inv(d);
    const bool nx = signbit(d.x), ny = signbit(d.y), nz = signbit(d.z);
    const uint32_t oct = uint32_t(nx) | (uint32_t(ny) << 1) | (uint32_t(nz) << 2);
    RayFrame F;
    TriRay TR;
    if (MODE == kPhantom) make_frame(o, d, F);
    if (MODE == kDots) TR = make_tri_ray(d);

    LaneStack<STACK> S;
    S.lds = lds;
    S.sp = 0;
    uint2 cur = make_uint2(0u, 0u);  // current group (empty)
    uint32_t node = 0;
    bool have_node = true;
    for (;;) {
        if (have_node) {
            if (node >= B.nnodes) return false;  // malformed structure: report, never read out of bounds
            if (STATS) ++st->nodes;
            const uint4* np = reinterpret_cast<const uint4*>(B.nodes + node);
            const uint4 n0 = np[0], n1 = np[1], n2 = np[2], n3 = np[3], n4 = np[4];
            const float sx = __uint_as_float((n0.w & 0xffu) << 23) * inv.x;
            const float sy = __uint_as_float(((n0.w >> 8) & 0xffu) << 23) * inv.y;
            const float sz = __uint_as_float(((n0.w >> 16) & 0xffu) << 23) * inv.z;
            const float ox = (__uint_as_float(n0.x) - o.x) * inv.x;
            const float oy = (__uint_as_float(n0.y) - o.y) * inv.y;
            const float oz = (__uint_as_float(n0.z) - o.z) * inv.z;
            const uint32_t imask = n0.w >> 24;
            // near/far plane bytes per axis by ray direction sign
            const uint32_t nxl = nx ? n3.z : n2.x, nxh = nx ? n3.w : n2.y;  // near x
            const uint32_t fxl = nx ? n2.x : n3.z, fxh = nx ? n2.y : n3.w;  // far x
            const uint32_t nyl = ny ? n4.x : n2.z, nyh = ny ? n4.y : n2.w;
            const uint32_t fyl = ny ? n2.z : n4.x, fyh = ny ? n2.w : n4.y;
            const uint32_t nzl = nz ? n4.z : n3.x, nzh = nz ? n4.w : n3.y;
            const uint32_t fzl = nz ? n3.x : n4.z, fzh = nz ? n3.y : n4.w;
            const float tcur = ANY ? best.t : best.t;
            uint32_t hits = 0, pmask = 0;
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const uint32_t meta = ((s < 4 ? n1.z : n1.w) >> (8 * (s & 3))) & 0xffu;
                float t0x = qbyte(nxl, nxh, s) * sx + ox, t1x = qbyte(fxl, fxh, s) * sx + ox;
                float t0y = qbyte(nyl, nyh, s) * sy + oy, t1y = qbyte(fyl, fyh, s) * sy + oy;
                float t0z = qbyte(nzl, nzh, s) * sz + oz, t1z = qbyte(fzl, fzh, s) * sz + oz;
                float tn = fmaxf(fmaxf(t0x, t0y), fmaxf(t0z, tmin));
                float tf = fminf(fminf(t1x, t1y), fminf(t1z, tcur)) * 1.00000095f;
                bool hit = meta != 0 && tn <= tf;
                if (hit) {
                    if (meta & 0x80u)
                        hits |= 1u << (uint32_t(s) ^ oct);
                    else
                        pmask |= ((1u << (meta >> 5)) - 1u) << (meta & 31u);
                }
            }
            // leaf primitives of this node
            const uint32_t pbase = n1.y;
            if (pmask && pbase + (32u - __clz(pmask)) > B.nprims) return false;
            while (pmask) {
                const uint32_t k = __ffs(pmask) - 1;
                pmask &= pmask - 1;
                const uint32_t slot = pbase + k;
                if (STATS) ++st->prims;
                if (MODE == kPhantom) {
                    const float4* pr = B.prims + size_t(slot) * 4;
                    float4 cv[4] = {pr[0], pr[1], pr[2], pr[3]};
                    PhantomHit ph;
                    if (prhi<RMODE, POLISH>(F, cv, B.radius, ph) && ph.t >= tmin && ph.t <= best.t) {
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
                            if (ANY) {
                                done = true;
                                return true;
                            }
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
                            if (ANY) {
                                done = true;
                                return true;
                            }
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
                                if (ANY) {
                                    done = true;
                                    return true;
                                }
                            }
                        }
                    }
                }
            }
            have_node = false;
            if (hits) {
                const uint2 g = make_uint2(n1.x, imask | (hits << 8));
                if (cur.y >> 8) {
                    if (!S.push(cur)) return false;
                }
                cur = g;
            }
        }
        if ((cur.y >> 8) == 0) {
            if (S.sp == 0) break;
            cur = S.top();
            --S.sp;
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
    }
    return true;
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
template <class G, bool ANY, int STACK, bool STATS = false>
__device__ __forceinline__ bool trace_world(const TraceParams& tp, f3 o, f3 d, float tmin, float tmax, Best& best,
                                            uint2* lds, TravStats* st = nullptr) {
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
                r = traverse<kLss, 0, false, ANY, STACK>(I.blas, oo, dd, tmin, i, best, lds, done);
            } else if (I.blas.mode == kDots) {
                r = traverse<kDots, 0, false, ANY, STACK>(I.blas, oo, dd, tmin, i, best, lds, done);
            } else if (I.blas.radius_mode == kRadiusPerPoint) {
                r = pol ? traverse<kPhantom, kRadiusPerPoint, true, ANY, STACK>(I.blas, oo, dd, tmin, i, best, lds, done)
                        : traverse<kPhantom, kRadiusPerPoint, false, ANY, STACK>(I.blas, oo, dd, tmin, i, best, lds, done);
            } else {
                r = pol ? traverse<kPhantom, kRadiusConstant, true, ANY, STACK>(I.blas, oo, dd, tmin, i, best, lds, done)
                        : traverse<kPhantom, kRadiusConstant, false, ANY, STACK>(I.blas, oo, dd, tmin, i, best, lds, done);
            }
        } else {
            r = traverse<G::kMode, G::kRMode, G::kPolish, ANY, STACK, STATS>(I.blas, oo, dd, tmin, i, best, lds, done,
                                                                             st);
        }
        ok = ok && r;
        if (ANY && done) break;
    }
    return ok;
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
    } else if (I.blas.mode == kLss) {
        n = f3{best.a, best.b, best.c};
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
    const bool ok = trace_world<G, ANY, STACK, STATS>(tp, o, d, r0.w, r1.w, best, stk + threadIdx.x, &ts);
    if (STATS) stats[i] = make_uint2(ts.nodes, ts.prims);
    if (!ok) atomicAdd(&tp.status[0], 1u);
    HitRec h;
    finish_hit(tp, o, d, best, h);
    hits[i] = h;
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

__device__ __forceinline__ void store_px(const RenderParams& rp, uint32_t pix, f3 c) {
    if (rp.out_format == 0) {
        reinterpret_cast<float4*>(rp.image)[pix] = make_float4(c.x, c.y, c.z, 1.0f);
    } else {
        auto q = [](float v) { return uint32_t(fminf(fmaxf(v, 0.0f), 1.0f) * 255.0f + 0.5f); };
        reinterpret_cast<uint32_t*>(rp.image)[pix] = q(c.x) | (q(c.y) << 8) | (q(c.z) << 16) | (255u << 24);
    }
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
        bool ok = trace_world<G, false, STACK>(rp.tp, o, d, 0.001f, 10000.0f, best, stk + threadIdx.x);
        if (!ok) atomicAdd(&rp.tp.status[0], 1u);
        HitRec h;
        finish_hit(rp.tp, o, d, best, h);
        hit = h.prim_id != 0xFFFFFFFFu;
        if (rp.prim_out) rp.prim_out[pix] = h.prim_id;
        const f3 n = f3{h.n[0], h.n[1], h.n[2]};
        if (!hit) {
            store_px(rp, pix, f3{0.0f, 0.0f, 0.0f});
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
        bool ok = trace_world<G, true, STACK>(rp.tp, hp, f3{0.0f, 1.0f, 0.0f}, 0.001f, 10000.0f, best, stk + threadIdx.x);
        if (!ok) atomicAdd(&rp.tp.status[0], 1u);
        vis = best.inst == 0xFFFFFFFFu ? 1.0f : 0.0f;
    }
    store_px(rp, pix, shade(f3{hb.x, hb.y, hb.z}, vis));
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
    if (any)
        hipLaunchKernelGGL((k_trace_rays<G, true, STACK, false>), grid, block, 0, st, tp, rays, n, hits, stats);
    else
        hipLaunchKernelGGL((k_trace_rays<G, false, STACK, false>), grid, block, 0, st, tp, rays, n, hits, stats);
    return hipGetLastError();
}

template <class G, int STACK>
hipError_t launch_render_variant(const RenderParams& rp, hipStream_t st, hipEvent_t* ev) {
    using namespace trace_detail;
    const uint32_t npix = rp.cam.width * rp.cam.height;
    if (npix == 0) return hipSuccess;
    const dim3 grid((npix + kTraceBlock - 1) / kTraceBlock), block(kTraceBlock);
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, st, rp.counters, 4u);
    if (ev) hipEventRecord(ev[0], st);
    hipLaunchKernelGGL((k_render_primary<G, STACK>), grid, block, 0, st, rp);
    if (ev) hipEventRecord(ev[1], st);
    if (ev) hipEventRecord(ev[2], st);
    if (rp.shadows && !rp.debug_palette) hipLaunchKernelGGL((k_render_shadow<G, STACK>), grid, block, 0, st, rp);
    if (ev) hipEventRecord(ev[3], st);
    return hipGetLastError();
}

// Explicit instantiation helpers for the trace_*.hip translation units.
#define VKHRT_INSTANTIATE_TRACE(G)                                                                              \
    template hipError_t launch_trace_variant<G, 8>(const TraceParams&, const float*, uint32_t, HitRec*, uint2*, \
                                                   hipStream_t);                                                \
    template hipError_t launch_trace_variant<G, 16>(const TraceParams&, const float*, uint32_t, HitRec*, uint2*,\
                                                    hipStream_t);                                               \
    template hipError_t launch_trace_variant<G, 24>(const TraceParams&, const float*, uint32_t, HitRec*, uint2*,\
                                                    hipStream_t);                                               \
    template hipError_t launch_render_variant<G, 8>(const RenderParams&, hipStream_t, hipEvent_t*);             \
    template hipError_t launch_render_variant<G, 16>(const RenderParams&, hipStream_t, hipEvent_t*);            \
    template hipError_t launch_render_variant<G, 24>(const RenderParams&, hipStream_t, hipEvent_t*);            \
    template hipError_t launch_trace_variant<G, 32>(const TraceParams&, const float*, uint32_t, HitRec*, uint2*,\
                                                    hipStream_t);                                               \
    template hipError_t launch_render_variant<G, 32>(const RenderParams&, hipStream_t, hipEvent_t*);

}  // namespace vk
