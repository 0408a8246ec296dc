// Dispatch of the traversal kernels to their compile-time variants (trace_*.hip).
#include "trace_impl.hpp"

namespace vk {

#define VKHRT_DECLARE_VARIANT(G)                                                                                   \
    extern template hipError_t launch_trace_variant<G, 8>(const TraceParams&, const float*, uint32_t, HitRec*,    \
                                                          hipStream_t);                                           \
    extern template hipError_t launch_trace_variant<G, 16>(const TraceParams&, const float*, uint32_t, HitRec*,   \
                                                           hipStream_t);                                          \
    extern template hipError_t launch_trace_variant<G, 24>(const TraceParams&, const float*, uint32_t, HitRec*,   \
                                                           hipStream_t);                                          \
    extern template hipError_t launch_render_variant<G, 8>(const RenderParams&, hipStream_t, hipEvent_t*);        \
    extern template hipError_t launch_render_variant<G, 16>(const RenderParams&, hipStream_t, hipEvent_t*);       \
    extern template hipError_t launch_render_variant<G, 24>(const RenderParams&, hipStream_t, hipEvent_t*);

using PhConst = Geo<kPhantom, kRadiusConstant, false>;
using PhConstPolish = Geo<kPhantom, kRadiusConstant, true>;
using PhPoint = Geo<kPhantom, kRadiusPerPoint, false>;
using PhPointPolish = Geo<kPhantom, kRadiusPerPoint, true>;
using LssG = Geo<kLss, 0, false>;
using DotsG = Geo<kDots, 0, false>;
VKHRT_DECLARE_VARIANT(PhConst)
VKHRT_DECLARE_VARIANT(PhConstPolish)
VKHRT_DECLARE_VARIANT(PhPoint)
VKHRT_DECLARE_VARIANT(PhPointPolish)
VKHRT_DECLARE_VARIANT(LssG)
VKHRT_DECLARE_VARIANT(DotsG)
VKHRT_DECLARE_VARIANT(GeoMixed)

namespace {
template <class G>
hipError_t trace_stack(const TraceParams& tp, const float* rays, uint32_t n, HitRec* hits, int stack_kind,
                       hipStream_t st) {
    if (stack_kind == 0) return launch_trace_variant<G, 8>(tp, rays, n, hits, st);
    if (stack_kind == 1) return launch_trace_variant<G, 16>(tp, rays, n, hits, st);
    return launch_trace_variant<G, 24>(tp, rays, n, hits, st);
}
template <class G>
hipError_t render_stack(const RenderParams& rp, int stack_kind, hipStream_t st, hipEvent_t* ev) {
    if (stack_kind == 0) return launch_render_variant<G, 8>(rp, st, ev);
    if (stack_kind == 1) return launch_render_variant<G, 16>(rp, st, ev);
    return launch_render_variant<G, 24>(rp, st, ev);
}
}  // namespace

hipError_t launch_trace_rays(const TraceParams& tp, const float* rays, uint32_t n, HitRec* hits, int variant,
                             int stack_kind, hipStream_t st) {
    switch (variant) {
        case kVarPhConst: return trace_stack<PhConst>(tp, rays, n, hits, stack_kind, st);
        case kVarPhConstPolish: return trace_stack<PhConstPolish>(tp, rays, n, hits, stack_kind, st);
        case kVarPhPoint: return trace_stack<PhPoint>(tp, rays, n, hits, stack_kind, st);
        case kVarPhPointPolish: return trace_stack<PhPointPolish>(tp, rays, n, hits, stack_kind, st);
        case kVarLss: return trace_stack<LssG>(tp, rays, n, hits, stack_kind, st);
        case kVarDots: return trace_stack<DotsG>(tp, rays, n, hits, stack_kind, st);
        default: return trace_stack<GeoMixed>(tp, rays, n, hits, stack_kind, st);
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
        default: return render_stack<GeoMixed>(rp, stack_kind, st, ev);
    }
}

}  // namespace vk
