// Batched path tracing over many views of one TLAS (an extension of the traceRaysKHR path:
// the reference renders one view per traceRaysKHR, renderer.cpp:156-166, and shades with
// Shade() alone, shading.glsl:1-10).
//
// Every pixel of every view is one path. The frame is a fixed sequence of kernels over a
// global path queue, so views of any resolution and instances of any size share every
// launch:
//   primary rays (all views) -> for depth b = 0..bounces:
//     closest-hit trace of the queue (one persistent launch) [-> FP64 polish merge]
//     [-> material histogram + scatter] -> shade: environment on a miss, the light
//     contribution and shadow ray of a hit, the next bounce ray
//     -> shadow trace (one persistent launch) [-> FP64 polish] -> add visible light
//   -> resolve every view's image.
// The kernel sequence does not depend on what the rays hit (queue lengths are read on the
// device), so it can be captured once as a graph and replayed.
//
// Path rule (fixed; the oracle restates it): light from direction L = (0, 1, 0) with
// radiance 1 (the reversed lightDirection of shading.glsl:4); at a hit with normal n, the
// facing normal nf = n flipped against the ray, albedo a = tint[material] * Shade(n) (the
// reference's colour with full light), light contribution T * a * (nf . L) when nf . L > 0
// and its shadow ray is unoccluded; the bounce direction is cosine-distributed about nf and
// the throughput becomes T * a (a Lambertian lobe whose 1/pi cancels the sampling density);
// a path that leaves the scene adds T * env. Sampling uses only +, -, *, /, sqrt, so the
// oracle reproduces every bit.
#pragma once

namespace vk {
namespace trace_detail {
namespace {

__host__ __device__ __forceinline__ uint32_t bc(uint32_t depth, uint32_t word) { return depth * kBatchDepthWords + word; }

constexpr uint32_t kDiskTries = 16;

// Cosine-distributed unit direction about the unit normal n: a uniform point of the unit
// disk by rejection (at most kDiskTries pairs; the centre if all fail) lifted to the
// hemisphere (Malley's method), in the orthonormal basis of onb_duff.
__device__ __forceinline__ f3 cosine_dir(f3 n, unsigned long long seed, uint32_t p, uint32_t b) {
    float x = 0.0f, y = 0.0f, r2 = 0.0f;
    for (uint32_t k = 0; k < kDiskTries; ++k) {
        const float cx = 2.0f * path_rand(seed, p, b, 2 * k) - 1.0f;
        const float cy = 2.0f * path_rand(seed, p, b, 2 * k + 1) - 1.0f;
        const float q = cx * cx + cy * cy;
        if (q < 1.0f) {
            x = cx;
            y = cy;
            r2 = q;
            break;
        }
    }
    const float z = sqrtf(1.0f - r2);
    f3 t1, t2;
    onb_duff(n, t1, t2);
    const f3 w{(x * t1.x + y * t2.x) + z * n.x, (x * t1.y + y * t2.y) + z * n.y, (x * t1.z + y * t2.z) + z * n.z};
    return nrm3(w);
}

// View of path p: the last view whose first path is <= p.
__device__ __forceinline__ uint32_t find_view(const BatchView* v, uint32_t nviews, uint32_t p) {
    uint32_t lo = 0, hi = nviews - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (v[mid].first <= p)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Queue of depth b: the ray-sorted copy for bounce depths when rays are sorted.
__host__ __device__ __forceinline__ uint32_t* depth_queue(const BatchParams& bp, uint32_t b) {
    return ((bp.sort & kBatchSortRays) && b > 0) ? bp.qsort : bp.q[b & 1u];
}

// Sort bin of a ray: direction octant, then the Morton code of its origin's cell.
__device__ __forceinline__ uint32_t ray_bin(const BatchParams& bp, float4 r0, float4 r1) {
    const uint32_t oct = (r1.x < 0.0f ? 1u : 0u) | (r1.y < 0.0f ? 2u : 0u) | (r1.z < 0.0f ? 4u : 0u);
    const float o[3] = {r0.x, r0.y, r0.z};
    uint32_t c[3];
    for (int k = 0; k < 3; ++k) {
        const float u = (o[k] - bp.scene_lo[k]) * bp.cell_scale[k];
        const uint32_t top = (1u << bp.cell_bits[k]) - 1u;
        c[k] = u > 0.0f ? min(uint32_t(u), top) : 0u;
    }
    uint32_t code = 0;
    for (int bit = 8; bit >= 0; --bit)
        for (int k = 0; k < 3; ++k)
            if (uint32_t(bit) < bp.cell_bits[k]) code = (code << 1) | ((c[k] >> bit) & 1u);
    return (oct << 9) | code;
}

// Ray sort of the depth-b queue (b >= 1), pass 1: bin histogram.
__global__ __launch_bounds__(kAppendBlock) void k_batch_ray_hist(BatchParams bp, uint32_t b) {
    __shared__ uint32_t s_h[kRayBins];
    for (uint32_t i = threadIdx.x; i < kRayBins; i += blockDim.x) s_h[i] = 0u;
    __syncthreads();
    const uint32_t n = bp.counters[bc(b, kBcQueue)];
    const uint32_t* q = bp.q[b & 1u];
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        const size_t p = q[s];
        atomicAdd(&s_h[ray_bin(bp, bp.ray[2 * p], bp.ray[2 * p + 1])], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kRayBins; i += blockDim.x)
        if (s_h[i]) atomicAdd(&bp.bins[i], s_h[i]);
}

// Pass 2 (one block): histogram -> exclusive bin offsets, in place.
__global__ __launch_bounds__(kAppendBlock) void k_batch_ray_scan(BatchParams bp) {
    __shared__ uint32_t s_part[kAppendBlock];
    constexpr uint32_t per = kRayBins / kAppendBlock;
    uint32_t v[per], sum = 0;
    for (uint32_t j = 0; j < per; ++j) {
        v[j] = bp.bins[threadIdx.x * per + j];
        sum += v[j];
    }
    s_part[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < kAppendBlock; off <<= 1) {  // inclusive scan of the thread sums
        const uint32_t add = threadIdx.x >= off ? s_part[threadIdx.x - off] : 0u;
        __syncthreads();
        s_part[threadIdx.x] += add;
        __syncthreads();
    }
    uint32_t run = s_part[threadIdx.x] - sum;
    for (uint32_t j = 0; j < per; ++j) {
        bp.bins[threadIdx.x * per + j] = run;
        run += v[j];
    }
}

// Pass 3: scatter the queue into bin order (bin offset + the block's reservation in the bin
// + rank within the block).
__global__ __launch_bounds__(kAppendBlock) void k_batch_ray_scatter(BatchParams bp, uint32_t b) {
    __shared__ uint32_t s_cnt[kRayBins];
    __shared__ uint32_t s_base[kRayBins];
    const uint32_t n = bp.counters[bc(b, kBcQueue)];
    const uint32_t* q = bp.q[b & 1u];
    uint32_t* cursor = bp.bins + kRayBins;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        for (uint32_t i = threadIdx.x; i < kRayBins; i += blockDim.x) s_cnt[i] = 0u;
        __syncthreads();
        const uint32_t s = base + threadIdx.x;
        uint32_t bin = 0, rank = 0, p = 0;
        if (s < n) {
            p = q[s];
            bin = ray_bin(bp, bp.ray[2 * size_t(p)], bp.ray[2 * size_t(p) + 1]);
            rank = atomicAdd(&s_cnt[bin], 1u);
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < kRayBins; i += blockDim.x)
            s_base[i] = s_cnt[i] ? atomicAdd(&cursor[i], s_cnt[i]) : 0u;
        __syncthreads();
        if (s < n) bp.qsort[bp.bins[bin] + s_base[bin] + rank] = p;
        __syncthreads();
    }
}

// Primary ray of every path; paths that miss every instance box add the environment and end.
__global__ __launch_bounds__(kAppendBlock) void k_batch_primary(BatchParams bp) {
    const BatchFrame& F = *bp.frame;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    bool live = false;
    if (p < bp.total) {
        const BatchView& V = bp.views[find_view(bp.views, F.nviews, p)];
        const uint32_t local = p - V.first;
        const uint32_t y = local / V.cam.width, x = local - y * V.cam.width;
        f3 o, d;
        primary_ray(V.cam, x, y, o, d);
        bp.ray[2 * size_t(p)] = make_float4(o.x, o.y, o.z, 0.001f);
        bp.ray[2 * size_t(p) + 1] = make_float4(d.x, d.y, d.z, 10000.0f);
        bp.thr[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
        for (uint32_t i = 0; i < bp.tp.ninst && !live; ++i) {
            const InstanceDev& I = bp.tp.inst[i];
            live = I.blas.nprims != 0 && ray_box(o, d, I.lo, I.hi, 0.001f, 10000.0f);
        }
        float4 L = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (!live) {
            L.x = L.x + 1.0f * F.env[0];
            L.y = L.y + 1.0f * F.env[1];
            L.z = L.z + 1.0f * F.env[2];
            if (V.prim_out) V.prim_out[local] = 0xFFFFFFFFu;
        }
        bp.rad[p] = L;
    }
    const uint32_t slot = block_append(live, &bp.counters[bc(0, kBcQueue)]);
    if (live) bp.q[0][slot] = p;
}

// Closest hit of the depth-b queue: hit record per queue slot (and the merge key when
// grazing roots are deferred to the FP64 polish).
template <class G, int STACK, int IM>
__global__ __launch_bounds__(kTraceBlock) void k_batch_trace(BatchParams bp, uint32_t b) {
    __shared__ uint2 stk[STACK * kTraceBlock];
    const uint32_t* q = depth_queue(bp, b);
    persistent_lanes<G, false, STACK, IM>(
        bp.tp, bp.counters[bc(b, kBcQueue)], bp.counters + bc(b, kBcWork), stk + threadIdx.x,
        [&](uint32_t qi, uint32_t& id, f3& o, f3& d, float& tmin, float& tmax) {
            id = qi;
            const size_t p = q[qi];
            const float4 r0 = bp.ray[2 * p], r1 = bp.ray[2 * p + 1];
            o = f3{r0.x, r0.y, r0.z};
            d = f3{r1.x, r1.y, r1.z};
            tmin = r0.w;
            tmax = r1.w;
        },
        [&](uint32_t qi, f3 o, f3 d, const Best& best) {
            HitRec h;
            finish_hit(bp.tp, o, d, best, h);
            bp.hits[qi] = h;
            if (G::kPolish) bp.tp.keys[qi] = best_key(bp.tp, best);
        },
        bp.counters + batch_xcd(b, 0));
}

// Mixed-mode TLASes (per-instance geometry dispatch): one lane per queue slot, the queue length
// read on the device; the grid covers every path.
template <class G, int STACK>
__global__ __launch_bounds__(kTraceBlock) void k_batch_trace_flat(BatchParams bp, uint32_t b) {
    __shared__ uint2 stk[STACK * kTraceBlock];
    const uint32_t qi = blockIdx.x * kTraceBlock + threadIdx.x;
    if (qi >= bp.counters[bc(b, kBcQueue)]) return;
    const size_t p = depth_queue(bp, b)[qi];
    const float4 r0 = bp.ray[2 * p], r1 = bp.ray[2 * p + 1];
    const f3 o{r0.x, r0.y, r0.z}, d{r1.x, r1.y, r1.z};
    Best best;
    if (!trace_world<G, false, STACK>(bp.tp, o, d, r0.w, r1.w, best, stk + threadIdx.x, qi))
        atomicAdd(&bp.tp.status[0], 1u);
    HitRec h;
    finish_hit(bp.tp, o, d, best, h);
    bp.hits[qi] = h;
    if (bp.tp.keys) bp.tp.keys[qi] = best_key(bp.tp, best);
}

// Candidate sources of the polish kernels: queue slot -> world ray.
struct SrcBatch {
    const uint32_t* q;
    const float4* ray;
    __device__ __forceinline__ void get(uint32_t qi, f3& o, f3& d, float& tmin, float& tmax) const {
        const size_t p = q[qi];
        const float4 r0 = ray[2 * p], r1 = ray[2 * p + 1];
        o = f3{r0.x, r0.y, r0.z};
        d = f3{r1.x, r1.y, r1.z};
        tmin = r0.w;
        tmax = r1.w;
    }
};
struct SrcBatchShadow {
    const uint32_t* sq;
    const float4* ray;
    __device__ __forceinline__ void get(uint32_t qi, f3& o, f3& d, float& tmin, float& tmax) const {
        const float4 r0 = ray[2 * size_t(sq[qi])];
        o = f3{r0.x, r0.y, r0.z};
        d = f3{0.0f, 1.0f, 0.0f};
        tmin = 0.001f;
        tmax = 10000.0f;
    }
};

// Shadow rays whose deferred grazing candidate the FP64 polish confirms are occluded.
__global__ void k_batch_polish_shadow(TraceParams tp, SrcBatchShadow src, float4* __restrict__ nee) {
    const uint32_t n = min(*tp.defer_count, tp.defer_cap);
    for (uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x; ci < n; ci += gridDim.x * blockDim.x) {
        const uint4 c = tp.defer[ci];
        const uint32_t p = src.sq[c.x];
        if (nee[p].w == 0.0f) continue;  // already occluded by an FP32 hit of the traversal
        f3 o, d;
        float tmin, tmax;
        src.get(c.x, o, d, tmin, tmax);
        HitRec h;
        unsigned long long key;
        if (eval_candidate(tp, c, o, d, tmin, tmax, h, key)) nee[p].w = 0.0f;
    }
}

// Material bin of a hit (misses last).
__device__ __forceinline__ uint32_t material_bin(const TraceParams& tp, const HitRec& h) {
    return h.inst_id == 0xFFFFFFFFu ? kBatchMaxMaterials : min(tp.inst[h.inst_id].material, kBatchMaxMaterials - 1u);
}

// Material histogram of the depth-b hits (one global atomic per block and bin).
__global__ __launch_bounds__(kAppendBlock) void k_batch_material_hist(BatchParams bp, uint32_t b) {
    __shared__ uint32_t s_h[kBatchMaxMaterials + 1];
    if (threadIdx.x <= kBatchMaxMaterials) s_h[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t n = bp.counters[bc(b, kBcQueue)];
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x)
        atomicAdd(&s_h[material_bin(bp.tp, bp.hits[s])], 1u);
    __syncthreads();
    if (threadIdx.x <= kBatchMaxMaterials && s_h[threadIdx.x])
        atomicAdd(&bp.counters[kBatchMatBase + 64 * b + threadIdx.x], s_h[threadIdx.x]);
}

// Queue slots of depth b in material order: bin offset (prefix of the histogram) + the
// block's reservation in the bin + the slot's rank within the block.
__global__ __launch_bounds__(kAppendBlock) void k_batch_material_scatter(BatchParams bp, uint32_t b) {
    __shared__ uint32_t s_pre[kBatchMaxMaterials + 1], s_cnt[kBatchMaxMaterials + 1], s_base[kBatchMaxMaterials + 1];
    const uint32_t* hist = bp.counters + kBatchMatBase + 64 * b;
    uint32_t* cursor = bp.counters + kBatchMatBase + 64 * b + 32;
    if (threadIdx.x == 0) {
        uint32_t sum = 0;
        for (uint32_t k = 0; k <= kBatchMaxMaterials; ++k) {
            s_pre[k] = sum;
            sum += hist[k];
        }
    }
    const uint32_t n = bp.counters[bc(b, kBcQueue)];
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        if (threadIdx.x <= kBatchMaxMaterials) s_cnt[threadIdx.x] = 0u;
        __syncthreads();
        const uint32_t s = base + threadIdx.x;
        uint32_t bin = 0, rank = 0;
        if (s < n) {
            bin = material_bin(bp.tp, bp.hits[s]);
            rank = atomicAdd(&s_cnt[bin], 1u);
        }
        __syncthreads();
        if (threadIdx.x <= kBatchMaxMaterials)
            s_base[threadIdx.x] = s_cnt[threadIdx.x] ? atomicAdd(&cursor[threadIdx.x], s_cnt[threadIdx.x]) : 0u;
        __syncthreads();
        if (s < n) bp.order[s_pre[bin] + s_base[bin] + rank] = s;
        __syncthreads();
    }
}

// Shades the depth-b queue (in material order when sorted): environment for misses; for
// hits the light contribution + shadow ray and, below the bounce limit, the next ray.
__global__ __launch_bounds__(kAppendBlock) void k_batch_shade(BatchParams bp, uint32_t b) {
    const BatchFrame& F = *bp.frame;
    const uint32_t n = bp.counters[bc(b, kBcQueue)];
    const uint32_t* q = depth_queue(bp, b);
    uint32_t* qn = bp.q[(b + 1u) & 1u];
    const bool more = b < bp.bounces;
    // the trip count is uniform per block (block_append needs every thread)
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const uint32_t s = base + threadIdx.x;
        bool shadow = false, bounce = false;
        uint32_t p = 0;
        if (s < n) {
            const uint32_t qi = (bp.sort & kBatchSortMaterial) ? bp.order[s] : s;
            p = q[qi];
            const HitRec h = bp.hits[qi];
            const float4 T = bp.thr[p];
            if (b == 0) {
                const BatchView& V = bp.views[find_view(bp.views, F.nviews, p)];
                if (V.prim_out) V.prim_out[p - V.first] = h.prim_id;
            }
            if (h.prim_id == 0xFFFFFFFFu) {
                float4 L = bp.rad[p];
                L.x = L.x + T.x * F.env[0];
                L.y = L.y + T.y * F.env[1];
                L.z = L.z + T.z * F.env[2];
                bp.rad[p] = L;
            } else {
                const float4 r0 = bp.ray[2 * size_t(p)], r1 = bp.ray[2 * size_t(p) + 1];
                const f3 d{r1.x, r1.y, r1.z};
                const f3 x = add(f3{r0.x, r0.y, r0.z}, scl(h.t, d));
                const f3 nn{h.n[0], h.n[1], h.n[2]};
                const f3 nf = dot3(nn, d) > 0.0f ? f3{-nn.x, -nn.y, -nn.z} : nn;
                const f3 s0 = shade(nn, 1.0f);
                const float* tint = F.albedo[min(bp.tp.inst[h.inst_id].material, kBatchMaxMaterials - 1u)];
                const f3 a{tint[0] * s0.x, tint[1] * s0.y, tint[2] * s0.z};
                const float c = nf.y;  // nf . (0, 1, 0)
                bp.ray[2 * size_t(p)] = make_float4(x.x, x.y, x.z, 0.001f);
                if (c > 0.0f) {
                    shadow = true;
                    bp.nee[p] = make_float4(T.x * a.x * c, T.y * a.y * c, T.z * a.z * c, 0.0f);
                }
                if (more) {
                    bounce = true;
                    const f3 w = cosine_dir(nf, F.seed, p, b);
                    bp.thr[p] = make_float4(T.x * a.x, T.y * a.y, T.z * a.z, 0.0f);
                    bp.ray[2 * size_t(p) + 1] = make_float4(w.x, w.y, w.z, 10000.0f);
                }
            }
        }
        const uint32_t ss = block_append(shadow, &bp.counters[bc(b, kBcShadow)]);
        if (shadow) bp.squeue[ss] = p;
        const uint32_t bs = block_append(bounce, &bp.counters[bc(b + 1u, kBcQueue)]);
        if (bounce) qn[bs] = p;
    }
}

// Shadow rays of depth b (any hit): visibility into nee.w.
template <class G, int STACK, int IM>
__global__ __launch_bounds__(kTraceBlock) void k_batch_shadow(BatchParams bp, uint32_t b) {
    __shared__ uint2 stk[STACK * kTraceBlock];
    const SrcBatchShadow src{bp.squeue, bp.ray};
    persistent_lanes<G, true, STACK, IM>(
        bp.tp, bp.counters[bc(b, kBcShadow)], bp.counters + bc(b, kBcShadowWork), stk + threadIdx.x,
        [&](uint32_t qi, uint32_t& id, f3& o, f3& d, float& tmin, float& tmax) {
            id = qi;
            src.get(qi, o, d, tmin, tmax);
        },
        [&](uint32_t qi, f3, f3, const Best& best) {
            bp.nee[bp.squeue[qi]].w = best.inst == 0xFFFFFFFFu ? 1.0f : 0.0f;
        },
        bp.counters + batch_xcd(b, 1));
}

template <class G, int STACK>
__global__ __launch_bounds__(kTraceBlock) void k_batch_shadow_flat(BatchParams bp, uint32_t b) {
    __shared__ uint2 stk[STACK * kTraceBlock];
    const uint32_t qi = blockIdx.x * kTraceBlock + threadIdx.x;
    if (qi >= bp.counters[bc(b, kBcShadow)]) return;
    f3 o, d;
    float tmin, tmax;
    SrcBatchShadow{bp.squeue, bp.ray}.get(qi, o, d, tmin, tmax);
    Best best;
    if (!trace_world<G, true, STACK>(bp.tp, o, d, tmin, tmax, best, stk + threadIdx.x, qi))
        atomicAdd(&bp.tp.status[0], 1u);
    bp.nee[bp.squeue[qi]].w = best.inst == 0xFFFFFFFFu ? 1.0f : 0.0f;
}

// Adds the light contribution of the unoccluded shadow rays of depth b.
__global__ void k_batch_apply(BatchParams bp, uint32_t b) {
    const uint32_t n = bp.counters[bc(b, kBcShadow)];
    for (uint32_t qi = blockIdx.x * blockDim.x + threadIdx.x; qi < n; qi += gridDim.x * blockDim.x) {
        const uint32_t p = bp.squeue[qi];
        const float4 v = bp.nee[p];
        if (v.w != 0.0f) {
            float4 L = bp.rad[p];
            L.x = L.x + v.x;
            L.y = L.y + v.y;
            L.z = L.z + v.z;
            bp.rad[p] = L;
        }
    }
}

// Radiance of every path -> its view's image.
__global__ void k_batch_resolve(BatchParams bp) {
    const BatchFrame& F = *bp.frame;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= bp.total) return;
    const BatchView& V = bp.views[find_view(bp.views, F.nviews, p)];
    const float4 L = bp.rad[p];
    store_rgba(V.image, F.out_format, p - V.first, f3{L.x, L.y, L.z});
}

}  // namespace
}  // namespace trace_detail

template <class G, int STACK>
hipError_t launch_batch_variant(const BatchParams& bp0, hipStream_t st) {
    using namespace trace_detail;
    {
        const uint32_t total = bp0.total;
        if (total == 0) return hipSuccess;
        const dim3 ab(kAppendBlock), tb(kTraceBlock);
        const uint32_t all_blocks = (total + kAppendBlock - 1) / kAppendBlock;
        const dim3 all(all_blocks), strided(std::min<uint32_t>(all_blocks, 2048u)), pgrid(1024);
        hipLaunchKernelGGL(k_zero, dim3((kBatchCounterWords + 255) / 256), dim3(256), 0, st, bp0.counters,
                           kBatchCounterWords);
        const int im = instance_mode(bp0.tp.ninst);  // how the traversal reaches the BLAS fields
        hipLaunchKernelGGL(k_batch_primary, all, ab, 0, st, bp0);
        for (uint32_t b = 0; b <= bp0.bounces; ++b) {
            BatchParams bp = bp0;
            if ((bp.sort & kBatchSortRays) && b > 0) {
                hipLaunchKernelGGL(k_zero, dim3(2 * kRayBins / 256), dim3(256), 0, st, bp.bins, 2 * kRayBins);
                hipLaunchKernelGGL(k_batch_ray_hist, strided, ab, 0, st, bp, b);
                hipLaunchKernelGGL(k_batch_ray_scan, dim3(1), ab, 0, st, bp);
                hipLaunchKernelGGL(k_batch_ray_scatter, strided, ab, 0, st, bp, b);
            }
            bp.tp.defer_count = bp0.counters + bc(b, kBcDefer);
            if constexpr (G::kMixed)
                hipLaunchKernelGGL((k_batch_trace_flat<G, STACK>), dim3((total + kTraceBlock - 1) / kTraceBlock), tb, 0,
                                   st, bp, b);
            else if (im == kInstOne)
                hipLaunchKernelGGL((k_batch_trace<G, STACK, kInstOne>),
                                   persistent_grid(k_batch_trace<G, STACK, kInstOne>, bp.tp.blocks_per_cu), tb, 0, st, bp, b);
            else if (im == kInstFew)
                hipLaunchKernelGGL((k_batch_trace<G, STACK, kInstFew>),
                                   persistent_grid(k_batch_trace<G, STACK, kInstFew>, bp.tp.blocks_per_cu), tb, 0, st, bp, b);
            else
                hipLaunchKernelGGL((k_batch_trace<G, STACK, kInstMany>),
                                   persistent_grid(k_batch_trace<G, STACK, kInstMany>, bp.tp.blocks_per_cu), tb, 0, st, bp,
                                   b);
            if (bp.tp.defer) {
                const SrcBatch src{depth_queue(bp, b), bp.ray};
                hipLaunchKernelGGL(k_polish_closest<SrcBatch>, pgrid, tb, 0, st, bp.tp, src, bp.tp.cand_hits,
                                   bp.tp.cand_keys);
                hipLaunchKernelGGL(k_polish_fixup, pgrid, tb, 0, st, bp.tp, bp.tp.cand_hits, bp.tp.cand_keys, bp.hits);
            }
            if (bp.sort & kBatchSortMaterial) {
                hipLaunchKernelGGL(k_batch_material_hist, strided, ab, 0, st, bp, b);
                hipLaunchKernelGGL(k_batch_material_scatter, strided, ab, 0, st, bp, b);
            }
            hipLaunchKernelGGL(k_batch_shade, strided, ab, 0, st, bp, b);
            bp.tp.defer_count = bp0.counters + bc(b, kBcDeferShadow);
            if constexpr (G::kMixed)
                hipLaunchKernelGGL((k_batch_shadow_flat<G, STACK>), dim3((total + kTraceBlock - 1) / kTraceBlock), tb, 0,
                                   st, bp, b);
            else if (im == kInstOne)
                hipLaunchKernelGGL((k_batch_shadow<G, STACK, kInstOne>),
                                   persistent_grid(k_batch_shadow<G, STACK, kInstOne>, bp.tp.blocks_per_cu), tb, 0, st, bp,
                                   b);
            else if (im == kInstFew)
                hipLaunchKernelGGL((k_batch_shadow<G, STACK, kInstFew>),
                                   persistent_grid(k_batch_shadow<G, STACK, kInstFew>, bp.tp.blocks_per_cu), tb, 0, st, bp,
                                   b);
            else
                hipLaunchKernelGGL((k_batch_shadow<G, STACK, kInstMany>),
                                   persistent_grid(k_batch_shadow<G, STACK, kInstMany>, bp.tp.blocks_per_cu), tb, 0, st,
                                   bp, b);
            if (bp.tp.defer)
                hipLaunchKernelGGL(k_batch_polish_shadow, pgrid, tb, 0, st, bp.tp, SrcBatchShadow{bp.squeue, bp.ray},
                                   bp.nee);
            hipLaunchKernelGGL(k_batch_apply, strided, tb, 0, st, bp, b);
        }
        hipLaunchKernelGGL(k_batch_resolve, dim3((total + kTraceBlock - 1) / kTraceBlock), tb, 0, st, bp0);
        return hipGetLastError();
    }
}

}  // namespace vk
