// Host-side entry points of the device code (C++ linkage, used by api.cpp only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "internal.hpp"

namespace vk {

struct BuiltBvh {
    WideNode* nodes = nullptr;
    uint32_t nnodes = 0;
    uint32_t* prim_ids = nullptr;
    float4* prims = nullptr;  // primitive records in leaf order
    float* aux = nullptr;     // phantom: early-out cylinder radius per leaf slot
    uint32_t depth = 0;       // wide levels
};

// Builds the 8-wide BVH over `n` primitive records (device pointer) on stream `st`.
// `width` (2..8) caps the children per wide node. `collapse` 0: greedy largest-area opening;
// c >= 1: SAH-optimal collapse with node cost 1 and primitive cost c / 4.
hipError_t build_bvh(int mode, const float4* prims_in, uint32_t n, int radius_mode, float radius, int leaf_max,
                     int width, int collapse, hipStream_t st, BuiltBvh* out, std::string& err);

struct TraceParams {
    const InstanceDev* inst;
    uint32_t ninst;
    uint32_t flags;
    uint32_t* status;  // [0] stack overflow count
    uint32_t* work;    // persistent kernels: global work counter (null = one thread per ray)
    uint32_t refill_min;  // persistent kernels: idle lanes needed before a wave fetches new rays
    // deferred FP64 polish (phantom POLISH variants): (ray, instance, slot) candidates
    uint4* defer;
    uint32_t* defer_count;
    uint32_t defer_cap;
    unsigned long long* keys;  // per ray (t bits << 32 | global primitive id) for the merge
    HitRec* cand_hits;         // per candidate
    unsigned long long* cand_keys;
    uint32_t blocks_per_cu;    // persistent grids: 256-thread blocks per CU, 0 = as many as fit
    uint32_t xcd_chunk;        // persistent grids: 0 one work counter; a power of two: rays handed
                               // out per XCD in runs of this many (each launch brings 8 counters)
    uint32_t* xcd_work;        // the 8 per-XCD counters of vkhrt_trace_rays's persistent launch
    unsigned long long* work_counts;  // counting build (VKHRT_WORK_COUNTERS): kWorkCounters words, else null
    uint32_t steps_per_iter;   // compacted phantom loop: traversal steps between bookkeeping passes (0 = 1)
};

// Work counters of the counting build (compiled with VKHRT_WORK_COUNTERS; the product build
// leaves them out). Per traversal kind (closest hit from 0, any hit from kWorkAnyHit; the lane-
// group shadow walker from kWorkWave), wave-level sums over the persistent loop:
enum WorkCounter {
    kWcIterations = 0,   // loop iterations of all waves
    kWcRounds = 1,       // intersection rounds (iterations that run Prhi on the queued pairs)
    kWcSteps = 2,        // traversal-step iterations
    kWcNodeTests = 3,    // lanes that tested a node in a step
    kWcLeafTests = 4,    // lanes that ran a leaf record's early-out test in a step
    kWcPairs = 5,        // (ray, record) pairs that passed the early-out (Prhi calls)
    kWcRoundLanes = 6,   // lanes busy in intersection rounds (sum of the pairs per round)
    kWcCyclesRounds = 7, // shader clock cycles spent in rounds, summed over waves
    kWcCyclesSteps = 8,  // ... in traversal steps
    kWcCyclesOther = 9,  // ... in refill, instance selection, tail hand-over and sinks
    kWcWaves = 10,       // waves that ran the loop
    kWcRays = 11,        // rays finished
    kWcHandovers = 12,   // stacked node groups handed to idle lanes (tail splitting)
    kWcStepLanes = 13,   // lanes walking (node test or leaf test) summed over step iterations
};
constexpr uint32_t kWorkAnyHit = 16, kWorkWave = 32, kWorkCounters = 48;
// lane-group shadow walker: kWorkWave + 0 rays, 1 group iterations, 2 node tests, 3 leaf records tested,
// 4 Prhi calls, 5 iterations that ran records, 6 cycles from a ray's start to its end (summed),
// 7 the largest of those; over the wavefronts that walked: 8 cycles in Prhi rounds, 9 rounds,
// 10 loop iterations, 11 cycles of the whole loop, 12 wavefronts

struct Camera {
    float view_inv[16];
    float proj_inv[16];
    uint32_t width, height;
    uint32_t order;    // work item -> pixel: 0 row-major, 1 8x8 tiles (one tile per wavefront),
                       // 2 64x64 tiles made of 8x8 tiles in row-major order
    uint32_t tiles_x;  // tiles per image row: ceil(width / 8) for order 1, ceil(width / 64) for 2
    uint32_t work;     // work items of a frame (orders 1, 2: whole tiles, some outside the image)
    float inv_w, inv_tiles_x;  // 1 / width, 1 / tiles_x for the index division
};

struct MlpDev;

// Allocation-free trace (vkhrt_trace_rays_noalloc): the layout of its per-device scratch array.
// [0, 4 KiB) counters: word 0 the work counter, word 1 the polish candidate count, words
// 32.. the 8 per-XCD work counters (kXcdStride apart); then, for traces with FP64 polish of up
// to kNoAllocMaxPolishRays rays, the merge key per ray and a candidate queue with room for one
// candidate per ray (queue entry, candidate hit and candidate key).
constexpr uint32_t kNoAllocMaxPolishRays = 1u << 18;
constexpr size_t kNoAllocCounterBytes = 4096;
constexpr size_t kNoAllocKeysOffset = kNoAllocCounterBytes;
constexpr size_t kNoAllocDeferOffset = kNoAllocKeysOffset + size_t(kNoAllocMaxPolishRays) * 8;
constexpr size_t kNoAllocCandHitsOffset = kNoAllocDeferOffset + size_t(kNoAllocMaxPolishRays) * 16;
constexpr size_t kNoAllocCandKeysOffset = kNoAllocCandHitsOffset + size_t(kNoAllocMaxPolishRays) * 32;
constexpr size_t kNoAllocScratchBytes = kNoAllocCandKeysOffset + size_t(kNoAllocMaxPolishRays) * 8;
// Device address of the scratch array on `device` (the current device), looked up on the
// first call for that device.
hipError_t noalloc_scratch_lookup(int device, void** out);

// Compacted phantom lanes queue (lane, leaf slot) pairs as lane << kCompactSlotBits | slot, so
// they serve phantom BLASes of fewer than 2^kCompactSlotBits records.
constexpr uint32_t kCompactSlotBits = 26;

// Context work counters (RenderParams::counters): words [0, 8) per frame; from kXcdCounters
// the per-XCD counters of the render's primary and shadow traversal launches and of
// vkhrt_trace_rays's launch, 8 each, kXcdStride words (one 128-byte line) apart.
constexpr uint32_t kXcdCounters = 64;
constexpr uint32_t kXcdStride = 32;
constexpr uint32_t kRenderCounterWords = kXcdCounters + 24 * kXcdStride;

// Equirectangular RGBA32F environment map of the miss shader (miss.rmiss), sampled bilinearly
// with repeat addressing like the reference's default sampler; texels == null: misses are black.
struct EnvMap {
    const float4* texels;
    uint32_t width, height;
    float exposure;
};

// Light of the render's shadow rays: the directional light toward (0, 1, 0) (one ray per
// hit), or a parallelogram c + a u + b v (a, b in [-1, 1]) sampled `samples` times per hit at
// points fixed by (seed, pixel, sample).
struct AreaLight {
    float c[3], u[3], v[3];
    uint32_t samples;  // shadow rays per hit pixel (1 with the directional light)
    uint32_t enabled;  // 0: directional
    unsigned long long seed;
};

struct RenderParams {
    TraceParams tp;
    Camera cam;
    uint32_t shadows;     // 0: Shade(n) only, 1: exact shadow ray
    uint32_t out_format;  // 0: RGBA32F, 1: RGBA8
    uint32_t debug_palette;
    void* image;
    uint32_t* prim_out;   // optional per-pixel primitive id
    float4* hitbuf;       // per-pixel (n.xyz, t) scratch
    uint32_t* queue;      // shadow queue (pixel indices)
    uint32_t* counters;   // [0] queue length; from [kXcdCounters]: 8 + 8 per-XCD work counters of the
                          // primary and shadow traversal launches, kXcdStride words apart
    float* light_vis;     // optional per-pixel visibility override (neural form), may be null
    // deferred-polish pipeline scratch
    HitRec* hitrec;       // per-pixel primary hit
    float* vis;           // per-pixel shadow visibility
    // neural shadow form (optional): the MLP gates the shadow queue, uncertain pixels fall back
    const MlpDev* mlp;         // host-side descriptor (read by the launch code only), null = exact
    float mlp_threshold;       // exact fallback when p (1 - p) > threshold
    uint32_t wave_limit;       // fallback queues up to this long are walked by lane groups
    uint32_t group_lanes;      // lanes per fallback ray of that walk: 8, 16, 32 or 64
    float4* origins;           // shadow-ray origin per queue entry
    uint32_t* fallback_queue;  // pixels left to the exact shadow ray ([5] of counters = length)
    uint32_t* live;            // persistent pipeline: pixels whose primary ray enters an instance box
                               // ([6] of counters = length); the rest are misses
    uint64_t* live_mask;       // the same as bits over the camera's work order (bit r & 63 of word r >> 6)
    AreaLight light;           // shadow queue entries are ids pixel * light.samples + sample
    EnvMap env;                // miss colour
    float4* dirs;              // neural form with an area light: shadow-ray direction per queue entry
    uint32_t accurate;         // re-evaluate primary phantom hits in double-word FP32
};

// Compile-time geometry variants of the traversal kernels. A TLAS whose instances share
// one mode uses the matching specialized kernels; mixed TLASes use kVarMixed.
enum TraceVariant {
    kVarPhConst = 0,
    kVarPhConstPolish = 1,
    kVarPhPoint = 2,
    kVarPhPointPolish = 3,
    kVarLss = 4,
    kVarDots = 5,
    kVarMixed = 6,
    kVarTris = 7
};

// ---- batched path tracing: many views of one TLAS in one wavefront pipeline
constexpr uint32_t kBatchMaxBounces = 8;
constexpr uint32_t kBatchMaxMaterials = 16;
// Counter words (zeroed at the start of a frame): per depth b = 0..kBatchMaxBounces, 8 words
// from kBatchDepthWords * b; the material histogram and cursors of depth b (64 words) from
// kBatchMatBase + 64 * b.
constexpr uint32_t kBatchDepthWords = 8;
enum BatchCounter {
    kBcQueue = 0,        // paths whose ray of depth b is traced (depth 0: those entering an instance box)
    kBcWork = 1,         // persistent work counter of that trace
    kBcShadow = 2,       // shadow rays of depth b
    kBcShadowWork = 3,
    kBcDefer = 4,        // deferred FP64 polish candidates of the closest-hit trace
    kBcDeferShadow = 5,  // ... of the shadow trace
};
constexpr uint32_t kBatchMatBase = 128;
// then the per-XCD work counters of the closest-hit (s = 0) and shadow (s = 1) traces of depth b:
// 8 counters kXcdStride words apart from batch_xcd(b, s)
constexpr uint32_t kBatchXcdBase = kBatchMatBase + 64 * (kBatchMaxBounces + 1);
constexpr uint32_t kBatchCounterWords = kBatchXcdBase + (kBatchMaxBounces + 1) * 2 * 8 * kXcdStride;
__host__ __device__ constexpr uint32_t batch_xcd(uint32_t b, uint32_t s) {
    return kBatchXcdBase + (2 * b + s) * 8 * kXcdStride;
}

// One view of a batch (device table): its camera (width/height set, row-major work order),
// its first path index and its output buffers.
struct BatchView {
    Camera cam;
    uint32_t first;
    uint32_t pad_;
    void* image;
    uint32_t* prim_out;  // optional: primary hit primitive per pixel
};

// Values of one frame that the kernels read from device memory, so that a captured graph
// replays with new ones.
struct BatchFrame {
    unsigned long long seed;
    float env[3];      // radiance of rays that leave the scene
    uint32_t nviews;
    uint32_t out_format;  // 0: RGBA32F, 1: RGBA8
    uint32_t pad_;
    float albedo[kBatchMaxMaterials][3];  // per-material tint of the albedo
};

// Kernel arguments: buffers (fixed while a captured graph is reused) and shape.
struct BatchParams {
    TraceParams tp;
    const BatchView* views;
    const BatchFrame* frame;
    uint32_t total;    // paths (pixels of all views)
    uint32_t bounces;  // indirect bounces per path
    uint32_t sort;     // kBatchSortMaterial | kBatchSortRays
    uint32_t pad_;
    float4* ray;       // per path, 2 x float4: (origin, tmin), (direction, tmax)
    float4* thr;       // per path: throughput rgb
    float4* rad;       // per path: radiance rgb
    float4* nee;       // per path: light contribution of the pending shadow ray, w = its visibility
    HitRec* hits;      // per queue slot of the current depth
    uint32_t* q[2];    // path queues of even and odd depths
    uint32_t* order;   // queue slots in material order
    uint32_t* squeue;  // paths with a shadow ray at the current depth
    uint32_t* counters;  // kBatchCounterWords
    // ray sort of the bounce queues: bin = direction octant << 9 | Morton cell of the origin in
    // the scene box (cell_bits per axis, 9 in total)
    uint32_t* qsort;     // depth >= 1 queue in bin order
    uint32_t* bins;      // kRayBins histogram (then bin offsets) + kRayBins cursors
    float scene_lo[3];
    float cell_scale[3];  // cells per unit length
    uint32_t cell_bits[3];
};
enum BatchSort { kBatchSortMaterial = 1, kBatchSortRays = 2 };
constexpr uint32_t kRayBins = 4096;

hipError_t launch_trace_rays(const TraceParams& tp, const float* rays, uint32_t n, HitRec* hits, uint2* stats,
                             int variant, int stack_kind, hipStream_t st);

// Shadow rays from points on phantom-curve surfaces: record g uniform over all instances'
// records (prefix[i] = records before instance i, prefix[ninst] = total), curve parameter and
// angle about the fibre uniform; toward a point of `light` (or along (0, 1, 0)).
constexpr uint32_t kSurfaceMaxInstances = 64;
struct SurfaceParams {
    const InstanceDev* inst;
    uint32_t ninst;
    uint32_t prefix[kSurfaceMaxInstances + 1];
    AreaLight light;
    unsigned long long seed;
};
hipError_t launch_surface_shadow_rays(const SurfaceParams& sp, uint32_t count, float* rays, hipStream_t st);

hipError_t launch_batch(const BatchParams& bp, int variant, int stack_kind, hipStream_t st);
hipError_t launch_render(const RenderParams& rp, int variant, int stack_kind, hipStream_t st,
                         hipEvent_t* ev /*[4] or null*/);

hipError_t launch_camera_rays(const Camera& cam, float tmin, float tmax, float* rays, hipStream_t st);
hipError_t launch_shadow_rays(const float* rays, const HitRec* hits, uint32_t n, const float dir[3], float tmin,
                              float tmax, float* out, uint32_t* src, uint32_t* counter, hipStream_t st);
// Double-word re-run of the final phantom hit of every ray (k_accurate_hits).
hipError_t launch_accurate_hits(const TraceParams& tp, const float* rays, uint32_t n, HitRec* hits, hipStream_t st);
hipError_t launch_count_candidates(const TraceParams& tp, const float* rays, const HitRec* hits, uint32_t n,
                                   uint32_t* counts, hipStream_t st);

template <class G, int STACK>
hipError_t launch_trace_variant(const TraceParams& tp, const float* rays, uint32_t n, HitRec* hits, uint2* stats,
                                hipStream_t st);
template <class G, int STACK>
hipError_t launch_render_variant(const RenderParams& rp, hipStream_t st, hipEvent_t* ev);
template <class G, int STACK>
hipError_t launch_batch_variant(const BatchParams& bp, hipStream_t st);

}  // namespace vk
