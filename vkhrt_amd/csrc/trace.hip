// Dispatch of the traversal kernels to their compile-time variants (trace_*.hip).
#include "trace_impl.hpp"

namespace vk {

#define VKHRT_DECLARE_VARIANT(G)                                                                                   \
    extern template hipError_t launch_trace_variant<G, 8>(const TraceParams&, const float*, uint32_t, HitRec*,    \
                                                          uint2*, hipStream_t);                                   \
    extern template hipError_t launch_trace_variant<G, 16>(const TraceParams&, const float*, uint32_t, HitRec*,   \
                                                           uint2*, hipStream_t);                                  \
    extern template hipError_t launch_trace_variant<G, 12>(const TraceParams&, const float*, uint32_t, HitRec*,   \
                                                           uint2*, hipStream_t);                                  \
    extern template hipError_t launch_render_variant<G, 8>(const RenderParams&, hipStream_t, hipEvent_t*);        \
    extern template hipError_t launch_render_variant<G, 16>(const RenderParams&, hipStream_t, hipEvent_t*);       \
    extern template hipError_t launch_render_variant<G, 12>(const RenderParams&, hipStream_t, hipEvent_t*);       \
    extern template hipError_t launch_trace_variant<G, 32>(const TraceParams&, const float*, uint32_t, HitRec*,   \
                                                           uint2*, hipStream_t);                                  \
    extern template hipError_t launch_render_variant<G, 32>(const RenderParams&, hipStream_t, hipEvent_t*);        \
    extern template hipError_t launch_batch_variant<G, 8>(const BatchParams&, hipStream_t);                       \
    extern template hipError_t launch_batch_variant<G, 12>(const BatchParams&, hipStream_t);                      \
    extern template hipError_t launch_batch_variant<G, 16>(const BatchParams&, hipStream_t);                      \
    extern template hipError_t launch_batch_variant<G, 32>(const BatchParams&, hipStream_t);

using PhConst = Geo<kPhantom, kRadiusConstant, false>;
using PhConstPolish = Geo<kPhantom, kRadiusConstant, true>;
using PhPoint = Geo<kPhantom, kRadiusPerPoint, false>;
using PhPointPolish = Geo<kPhantom, kRadiusPerPoint, true>;
using LssG = Geo<kLss, 0, false>;
using DotsG = Geo<kDots, 0, false>;
using TrisG = Geo<kTriangles, 0, false>;
VKHRT_DECLARE_VARIANT(PhConst)
VKHRT_DECLARE_VARIANT(PhConstPolish)
VKHRT_DECLARE_VARIANT(PhPoint)
VKHRT_DECLARE_VARIANT(PhPointPolish)
VKHRT_DECLARE_VARIANT(LssG)
VKHRT_DECLARE_VARIANT(DotsG)
VKHRT_DECLARE_VARIANT(TrisG)
VKHRT_DECLARE_VARIANT(GeoMixed)

namespace {
template <class G>
hipError_t trace_stack(const TraceParams& tp, const float* rays, uint32_t n, HitRec* hits, uint2* stats,
                       int stack_kind, hipStream_t st) {
    if (stack_kind == 0) return launch_trace_variant<G, 8>(tp, rays, n, hits, stats, st);
    if (stack_kind == 1) return launch_trace_variant<G, 12>(tp, rays, n, hits, stats, st);
    if (stack_kind == 2) return launch_trace_variant<G, 16>(tp, rays, n, hits, stats, st);
    return launch_trace_variant<G, 32>(tp, rays, n, hits, stats, st);
}
template <class G>
hipError_t render_stack(const RenderParams& rp, int stack_kind, hipStream_t st, hipEvent_t* ev) {
    if (stack_kind == 0) return launch_render_variant<G, 8>(rp, st, ev);
    if (stack_kind == 1) return launch_render_variant<G, 12>(rp, st, ev);
    if (stack_kind == 2) return launch_render_variant<G, 16>(rp, st, ev);
    return launch_render_variant<G, 32>(rp, st, ev);
}
template <class G>
hipError_t batch_stack(const BatchParams& bp, int stack_kind, hipStream_t st) {
    if (stack_kind == 0) return launch_batch_variant<G, 8>(bp, st);
    if (stack_kind == 1) return launch_batch_variant<G, 12>(bp, st);
    if (stack_kind == 2) return launch_batch_variant<G, 16>(bp, st);
    return launch_batch_variant<G, 32>(bp, st);
}
}  // namespace

hipError_t launch_batch(const BatchParams& bp, int variant, int stack_kind, hipStream_t st) {
    switch (variant) {
        case kVarPhConst: return batch_stack<PhConst>(bp, stack_kind, st);
        case kVarPhConstPolish: return batch_stack<PhConstPolish>(bp, stack_kind, st);
        case kVarPhPoint: return batch_stack<PhPoint>(bp, stack_kind, st);
        case kVarPhPointPolish: return batch_stack<PhPointPolish>(bp, stack_kind, st);
        case kVarLss: return batch_stack<LssG>(bp, stack_kind, st);
        case kVarDots: return batch_stack<DotsG>(bp, stack_kind, st);
        case kVarTris: return batch_stack<TrisG>(bp, stack_kind, st);
        default: return batch_stack<GeoMixed>(bp, stack_kind, st);
    }
}

hipError_t launch_trace_rays(const TraceParams& tp, const float* rays, uint32_t n, HitRec* hits, uint2* stats,
                             int variant, int stack_kind, hipStream_t st) {
    switch (variant) {
        case kVarPhConst: return trace_stack<PhConst>(tp, rays, n, hits, stats, stack_kind, st);
        case kVarPhConstPolish: return trace_stack<PhConstPolish>(tp, rays, n, hits, stats, stack_kind, st);
        case kVarPhPoint: return trace_stack<PhPoint>(tp, rays, n, hits, stats, stack_kind, st);
        case kVarPhPointPolish: return trace_stack<PhPointPolish>(tp, rays, n, hits, stats, stack_kind, st);
        case kVarLss: return trace_stack<LssG>(tp, rays, n, hits, stats, stack_kind, st);
        case kVarDots: return trace_stack<DotsG>(tp, rays, n, hits, stats, stack_kind, st);
        case kVarTris: return trace_stack<TrisG>(tp, rays, n, hits, stats, stack_kind, st);
        default: return trace_stack<GeoMixed>(tp, rays, n, hits, stats, stack_kind, st);
    }
}

hipError_t launch_render(const RenderParams& rp, int variant, int stack_kind, hipStream_t st, hipEvent_t* ev) {
    switch (variant) {
        case kVarPhConst: return render_stack<PhConst>(rp, stack_kind, st, ev);
        case kVarPhConstPolish: return render_stack<PhConstPolish>(rp, stack_kind, st, ev);
        case kVarPhPoint: return render_stack<PhPoint>(rp, stack_kind, st, ev);
        case kVarPhPointPolish: return render_stack<PhPointPolish>(rp, stack_kind, st, ev);
        case kVarLss: return render_stack<LssG>(rp, stack_kind, st, ev);
        case kVarDots: return render_stack<DotsG>(rp, stack_kind, st, ev);
        case kVarTris: return render_stack<TrisG>(rp, stack_kind, st, ev);
        default: return render_stack<GeoMixed>(rp, stack_kind, st, ev);
    }
}

// ------------------------------------------------------------------ auxiliary kernels
namespace {

using trace_detail_pub::primary_ray_pub;

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ray_gen.rgen:18-28 as a standalone ray generator: one vkhrt_ray per launch id.
__global__ void k_camera_rays(Camera cam, float tmin, float tmax, float4* __restrict__ rays) {
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= cam.width * cam.height) return;
    f3 o, d;
    primary_ray_pub(cam, pix % cam.width, pix / cam.width, o, d);
    rays[size_t(pix) * 2] = make_float4(o.x, o.y, o.z, tmin);
    rays[size_t(pix) * 2 + 1] = make_float4(d.x, d.y, d.z, tmax);
}

// Training rays of the shadow MLP: a point on a phantom fibre's surface (record, curve
// parameter t and angle about the axis uniform; radius as the intersector's per-point model)
// toward a light point. Random numbers: key (sample i, slot 1) for the surface, the light's
// own (i, 0) pair, and one 64-bit draw for the record.
__global__ void k_surface_shadow_rays(SurfaceParams sp, uint32_t count, float4* __restrict__ rays) {
    using namespace trace_detail;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const unsigned long long r = mix64((sp.seed ^ (static_cast<unsigned long long>(i) << 24)) + 0x9E3779B97F4A7C15ull +
                                       0x632BE59BD9B4E019ull);
    const uint32_t g = uint32_t(__umul64hi(r, sp.prefix[sp.ninst]));
    uint32_t k = 0;
    while (k + 1 < sp.ninst && sp.prefix[k + 1] <= g) ++k;
    const InstanceDev& I = sp.inst[k];
    const float4* cv = I.blas.prims + size_t(g - sp.prefix[k]) * 4;
    const float4 c0 = cv[0], c1 = cv[1], c2 = cv[2], c3 = cv[3];
    const f3 p0{c0.x, c0.y, c0.z}, p1{c1.x, c1.y, c1.z}, p2{c2.x, c2.y, c2.z}, p3{c3.x, c3.y, c3.z};
    const float t = path_rand(sp.seed, i, 1, 0);
    const float turn = path_rand(sp.seed, i, 1, 1);
    const f3 P = bez_point(p0, p1, p2, p3, t);
    const float rad = I.blas.radius_mode == kRadiusPerPoint ? bez_scalar(c0.w, c1.w, c2.w, c3.w, t) : I.blas.radius;
    f3 t1, t2;
    onb_duff(nrm3(bez_axis(p0, p1, p2, p3, t)), t1, t2);
    float sn, cs;
    sincospif(2.0f * turn, &sn, &cs);
    f3 x = add(P, scl(rad, add(scl(cs, t1), scl(sn, t2))));
    if (!I.identity) {
        const float* m = I.xf;
        x = f3{m[0] * x.x + m[1] * x.y + m[2] * x.z + m[3], m[4] * x.x + m[5] * x.y + m[6] * x.z + m[7],
               m[8] * x.x + m[9] * x.y + m[10] * x.z + m[11]};
    }
    f3 d;
    float tmax;
    light_ray(sp.light, i, 0, x, d, tmax);
    rays[size_t(i) * 2] = make_float4(x.x, x.y, x.z, 0.001f);
    rays[size_t(i) * 2 + 1] = make_float4(d.x, d.y, d.z, tmax);
}

// Shadow rays toward `dir` from every hit: origin = o + t d (FP32, as the render kernels).
// Compacted with wave-aggregated atomics; src[i] records the source ray of each output.
__global__ void k_shadow_rays(const float4* __restrict__ rays, const HitRec* __restrict__ hits, uint32_t n, float3 dir,
                              float tmin, float tmax, float4* __restrict__ out, uint32_t* __restrict__ src,
                              uint32_t* __restrict__ counter) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool hit = false;
    float4 r0 = make_float4(0, 0, 0, 0), r1 = r0;
    float t = 0.0f;
    if (i < n) {
        hit = hits[i].prim_id != 0xFFFFFFFFu;
        r0 = rays[size_t(i) * 2];
        r1 = rays[size_t(i) * 2 + 1];
        t = hits[i].t;
    }
    const uint64_t m = __ballot(hit);
    if (!m) return;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, uint32_t(__popcll(m)));
    base = __shfl(base, int(leader));
    if (hit) {
        const uint32_t k = base + __popcll(m & ((1ull << lane) - 1ull));
        out[size_t(k) * 2] = make_float4(r0.x + t * r1.x, r0.y + t * r1.y, r0.z + t * r1.z, tmin);
        out[size_t(k) * 2 + 1] = make_float4(dir.x, dir.y, dir.z, tmax);
        if (src) src[k] = i;
    }
}

// Exact GenerateAABBs-style box of leaf record `slot` (no pad). DOTS: triangle `tri`.
__device__ __forceinline__ void exact_box(const BlasDev& B, uint32_t slot, int tri, float lo[3], float hi[3]) {
    if (B.mode == kPhantom) {
        const float4* c = B.prims + size_t(slot) * 4;
        const float4 p0 = c[0], p1 = c[1], p2 = c[2], p3 = c[3];
        float r = B.radius;
        if (B.radius_mode == kRadiusPerPoint) r = gmax(gmax(p0.w, p1.w), gmax(p2.w, p3.w));
        const float xs[4] = {p0.x, p1.x, p2.x, p3.x}, ys[4] = {p0.y, p1.y, p2.y, p3.y}, zs[4] = {p0.z, p1.z, p2.z, p3.z};
        const float* a[3] = {xs, ys, zs};
        for (int k = 0; k < 3; ++k) {
            lo[k] = gmin(gmin(a[k][0], a[k][3]), gmin(a[k][1], a[k][2])) - r;
            hi[k] = gmax(gmax(a[k][0], a[k][3]), gmax(a[k][1], a[k][2])) + r;
        }
    } else if (B.mode == kLss) {
        const float4 a = B.prims[size_t(slot) * 2], b = B.prims[size_t(slot) * 2 + 1];
        const float ra = fabsf(a.w), rb = b.w;
        const float ax[3] = {a.x, a.y, a.z}, bx[3] = {b.x, b.y, b.z};
        for (int k = 0; k < 3; ++k) {
            lo[k] = gmin(ax[k] - ra, bx[k] - rb);
            hi[k] = gmax(ax[k] + ra, bx[k] + rb);
        }
    } else if (B.mode == kTriangles) {
        const float4* pr = B.prims + size_t(slot) * 3;
        const float4 v[3] = {pr[0], pr[1], pr[2]};
        const float a[3] = {v[0].x, v[0].y, v[0].z}, b[3] = {v[1].x, v[1].y, v[1].z}, c[3] = {v[2].x, v[2].y, v[2].z};
        for (int k = 0; k < 3; ++k) {
            lo[k] = gmin(a[k], gmin(b[k], c[k]));
            hi[k] = gmax(a[k], gmax(b[k], c[k]));
        }
    } else {
        const float4* pr = B.prims + size_t(slot) * 4;
        float4 P[4] = {pr[0], pr[1], pr[2], pr[3]};
        f3 v0, v1, v2;
        dots_tri(P, tri >> 1, tri & 1, v0, v1, v2);
        const float a[3] = {v0.x, v0.y, v0.z}, b[3] = {v1.x, v1.y, v1.z}, c[3] = {v2.x, v2.y, v2.z};
        for (int k = 0; k < 3; ++k) {
            lo[k] = gmin(a[k], gmin(b[k], c[k]));
            hi[k] = gmax(a[k], gmax(b[k], c[k]));
        }
    }
}

// Slab test with explicit NaN handling (origin on a slab plane with a zero direction
// component): same operation order as the oracle's candidate count.
__device__ __forceinline__ bool slab_exact(const float lo[3], const float hi[3], f3 o, f3 inv, float tmin, float tmax) {
    float t0 = tmin, t1 = tmax;
    const float oo[3] = {o.x, o.y, o.z}, iv[3] = {inv.x, inv.y, inv.z};
    for (int k = 0; k < 3; ++k) {
        float a = (lo[k] - oo[k]) * iv[k];
        float c = (hi[k] - oo[k]) * iv[k];
        if (isnan(a) || isnan(c)) {
            a = -INFINITY;
            c = INFINITY;
        }
        const float l = c < a ? c : a, h = a < c ? c : a;
        t0 = t0 < l ? l : t0;
        const float hh = h * (1.0f + 4e-7f);
        t1 = hh < t1 ? hh : t1;
    }
    return t0 <= t1;
}

// N*_prim: primitives whose exact box overlaps [tmin, t_end] (t_end = hit t, or tmax on a
// miss), summed over instances. Visits the BVH without intersection tests.
__global__ void k_count_candidates(TraceParams tp, const float4* __restrict__ rays, const HitRec* __restrict__ hits,
                                   uint32_t n, uint32_t* __restrict__ counts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 r0 = rays[size_t(i) * 2], r1 = rays[size_t(i) * 2 + 1];
    const float tend = hits[i].prim_id != 0xFFFFFFFFu ? hits[i].t : r1.w;
    uint32_t count = 0;
    for (uint32_t in = 0; in < tp.ninst; ++in) {
        const InstanceDev& I = tp.inst[in];
        const BlasDev& B = I.blas;
        if (B.nprims == 0) continue;
        f3 o = f3{r0.x, r0.y, r0.z}, d = f3{r1.x, r1.y, r1.z};
        if (!I.identity) {
            o = trace_detail_pub::xform_pt_pub(I.inv, o);
            d = trace_detail_pub::xform_dir_pub(I.inv, d);
        }
        const f3 inv = f3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
        uint32_t stack[64];
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            const uint32_t node = stack[--sp];
            if (node >= B.nnodes) break;
            const WideNode N = B.nodes[node];
            const uint32_t imask = N.e_imask >> 24;
            float sc[3];
            for (int k = 0; k < 3; ++k) sc[k] = __uint_as_float(((N.e_imask >> (8 * k)) & 0xffu) << 23);
            const float org[3] = {N.px, N.py, N.pz};
            const uint32_t* ql[3] = {N.qlo_x, N.qlo_y, N.qlo_z};
            const uint32_t* qh[3] = {N.qhi_x, N.qhi_y, N.qhi_z};
            uint32_t rank = 0;
            for (int s = 0; s < 8; ++s) {
                const uint32_t meta = (N.meta[s >> 2] >> (8 * (s & 3))) & 0xffu;
                if (!meta) continue;
                float lo[3], hi[3];
                for (int k = 0; k < 3; ++k) {
                    const float a = float((ql[k][s >> 2] >> (8 * (s & 3))) & 0xffu);
                    const float b = float((qh[k][s >> 2] >> (8 * (s & 3))) & 0xffu);
                    const float m = fmaxf(fabsf(org[k]), fabsf(org[k] + 255.0f * sc[k]));
                    lo[k] = org[k] + a * sc[k] - (1e-4f + 1e-6f * m);
                    hi[k] = org[k] + b * sc[k] + (1e-4f + 1e-6f * m);
                }
                const bool isint = (meta & 0x80u) != 0;
                const uint32_t child = N.child_base + rank;
                if (isint) ++rank;
                if (!slab_exact(lo, hi, o, inv, r0.w, tend)) continue;
                if (isint) {
                    if (sp < 64) stack[sp++] = child;
                    else atomicAdd(&tp.status[0], 1u);
                } else {
                    const uint32_t cnt = meta >> 5, off = meta & 31u;
                    for (uint32_t j = 0; j < cnt; ++j) {
                        const uint32_t slot = N.prim_base + off + j;
                        if (slot >= B.nprims) continue;
                        const int ntri = B.mode == kDots ? 4 : 1;
                        for (int tri = 0; tri < ntri; ++tri) {
                            float plo[3], phi[3];
                            exact_box(B, slot, tri, plo, phi);
                            if (slab_exact(plo, phi, o, inv, r0.w, tend)) ++count;
                        }
                    }
                }
            }
        }
    }
    counts[i] = count;
}

}  // namespace

hipError_t launch_surface_shadow_rays(const SurfaceParams& sp, uint32_t count, float* rays, hipStream_t st) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_surface_shadow_rays, dim3((count + 255) / 256), dim3(256), 0, st, sp, count,
                       reinterpret_cast<float4*>(rays));
    return hipGetLastError();
}

hipError_t launch_camera_rays(const Camera& cam, float tmin, float tmax, float* rays, hipStream_t st) {
    const uint32_t n = cam.width * cam.height;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_camera_rays, dim3((n + 255) / 256), dim3(256), 0, st, cam, tmin, tmax,
                       reinterpret_cast<float4*>(rays));
    return hipGetLastError();
}

hipError_t launch_shadow_rays(const float* rays, const HitRec* hits, uint32_t n, const float dir[3], float tmin,
                              float tmax, float* out, uint32_t* src, uint32_t* counter, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, counter, 1u);  // fill kernel, safe under capture
    if (n == 0) return hipGetLastError();
    hipLaunchKernelGGL(k_shadow_rays, dim3((n + 255) / 256), dim3(256), 0, st, reinterpret_cast<const float4*>(rays), hits,
                       n, make_float3(dir[0], dir[1], dir[2]), tmin, tmax, reinterpret_cast<float4*>(out), src, counter);
    return hipGetLastError();
}

hipError_t launch_accurate_hits(const TraceParams& tp, const float* rays, uint32_t n, HitRec* hits, hipStream_t st) {
    using namespace trace_detail;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_accurate_hits<SrcRays>, dim3((n + kTraceBlock - 1) / kTraceBlock), dim3(kTraceBlock), 0, st, tp,
                       SrcRays{reinterpret_cast<const float4*>(rays)}, hits, n);
    return hipGetLastError();
}

hipError_t launch_count_candidates(const TraceParams& tp, const float* rays, const HitRec* hits, uint32_t n,
                                   uint32_t* counts, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_count_candidates, dim3((n + 255) / 256), dim3(256), 0, st, tp,
                       reinterpret_cast<const float4*>(rays), hits, n, counts);
    return hipGetLastError();
}

}  // namespace vk
