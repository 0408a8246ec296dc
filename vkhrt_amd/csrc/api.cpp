// C-ABI of the vkhrt hair path (include/vkhrt.h): contexts, BLAS/TLAS build, tracing,
// rendering, geometry front-end and profiling. Host C++ over hand-written HIP kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <atomic>
#include <map>
#include <numeric>
#include <new>
#include <string>
#include <vector>

#include "vkhrt.h"
#include "asset.hpp"
#include "build.hpp"
#include "geometry.hpp"
#include "internal.hpp"
#include "mlp.hpp"
#include "voxel.hpp"

namespace {

struct PendingTiming {
    std::string name;
    hipEvent_t a, b;
};

}  // namespace

struct vkhrt_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int stack_kind = 1;
    int leaf_size = 1;
    int bvh_width = 8;
    int bvh_collapse = 0;  // greedy largest-area collapse (fewer node visits per ray on hair than SAH)
    bool persistent = true;
    uint32_t refill_min = 16;
    uint32_t ray_order = 0;  // render work order: 0 row-major, 1 8x8 tiles, 2 64x64 tiles of 8x8 tiles
    uint32_t xcd_chunk = 4096;  // render traversal: rays handed out per XCD in runs of this many (0: one counter)
    uint32_t blocks_per_cu = 0;  // persistent grids: 256-thread blocks per CU (0 = occupancy limit)
    uint32_t shadow_wave_limit = 65536;  // neural fallback queues up to this long: lane groups per ray
    uint32_t shadow_group_lanes = 16;    // lanes per fallback ray of that walk
    uint32_t steps_per_iter = 3;         // compacted phantom loop: traversal steps per bookkeeping pass
    bool profiling = false;
    std::vector<PendingTiming> pending;
    std::map<std::string, std::pair<double, uint32_t>> stats;
    uint32_t* status = nullptr;  // device counters
    void* noalloc = nullptr;     // device address of this GPU's allocation-free trace scratch
    unsigned long long* work_counts = nullptr;  // counting build: the kernels' work counters
    // render scratch
    float4* hitbuf = nullptr;
    uint32_t* queue = nullptr;
    uint32_t* counters = nullptr;
    size_t cap_pixels = 0;
    size_t cap_work = 0;  // work items the live mask covers
    uint64_t last_pixels = 0;
    bool last_shadows = false;
    hipStream_t last_stream = nullptr;
    vk::HitRec* hitrec = nullptr;  // per pixel (deferred-polish render)
    float* vis = nullptr;          // per pixel
    uint32_t* live = nullptr;      // per pixel: primary rays that enter an instance box
    uint64_t* live_mask = nullptr; // the same as one bit per pixel in the camera's work order
    // deferred FP64 polish
    uint4* defer = nullptr;
    vk::HitRec* cand_hits = nullptr;
    unsigned long long* cand_keys = nullptr;
    uint32_t defer_cap = 0;
    unsigned long long* keys = nullptr;
    size_t keys_cap = 0;
    // neural shadow form
    float4* origins = nullptr;
    float4* dirs = nullptr;
    uint32_t* fallback = nullptr;
    size_t cap_neural = 0;
    bool last_neural = false;
    // batched path tracing: per-path state, queues, counters and its own polish arrays
    // (a captured graph holds these addresses; `batch_gen` changes whenever they move)
    float4* b_ray = nullptr;
    float4* b_thr = nullptr;
    float4* b_rad = nullptr;
    float4* b_nee = nullptr;
    vk::HitRec* b_hits = nullptr;
    uint32_t* b_q0 = nullptr;
    uint32_t* b_q1 = nullptr;
    uint32_t* b_order = nullptr;
    uint32_t* b_squeue = nullptr;
    uint32_t* b_qsort = nullptr;
    uint32_t* b_bins = nullptr;
    uint32_t* b_counters = nullptr;
    unsigned long long* b_keys = nullptr;
    uint4* b_defer = nullptr;
    vk::HitRec* b_cand_hits = nullptr;
    unsigned long long* b_cand_keys = nullptr;
    size_t b_cap = 0;
    vk::BatchView* b_views = nullptr;  // device view table
    uint32_t b_views_cap = 0;
    vk::BatchFrame* b_frame = nullptr;
    uint8_t* b_host = nullptr;         // pinned staging of the view table + frame values
    size_t b_host_cap = 0;
    hipEvent_t b_upload = nullptr;     // last upload out of the staging buffer
    uint64_t batch_gen = 0;
    // captured frame graph and what it was captured for
    hipStream_t capture_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<uint64_t> graph_key;
    uint32_t graph_captures = 0;
    // last batch (stats)
    uint64_t last_batch_paths = 0;
    uint32_t last_batch_bounces = 0;
    hipStream_t last_batch_stream = nullptr;
};

struct vkhrt_shadow_mlp {
    vkhrt_context* ctx = nullptr;
    vk::MlpDev dev{};
    float* partial = nullptr;  // [blocks][params] gradient partials
    uint32_t partial_blocks = 0;
    float* loss = nullptr;     // device scalar
    uint32_t step = 0;         // Adam steps taken
};

struct vkhrt_blas {
    vkhrt_context* ctx = nullptr;
    float4* normals = nullptr;  // triangle meshes: 3 vertex normals per triangle
    uint32_t* slot_of = nullptr;  // phantom: leaf slot per original primitive id
    uint32_t mode = 0;
    uint32_t radius_mode = 0;
    float radius = 0.02f;
    uint32_t material_index = 0;
    bool has_xf = false;
    float xf[12];
    vk::BuiltBvh bvh;
    uint32_t records = 0;
    float lo[3], hi[3];  // object-space bounds
    vkhrt_blas_info info{};
    uint32_t* voxels = nullptr;  // voxel BLAS: occupancy bits of the traced level
    uint32_t* vbricks = nullptr;  // voxel BLAS: one bit per 4x4x4 brick of the traced level
    uint32_t* vblocks = nullptr;  // voxel BLAS: one bit per 16x16x16 block of the traced level
    float vorigin[3] = {0, 0, 0};
    float vcell = 0.0f;
    int32_t vres[3] = {0, 0, 0};
};

struct vkhrt_tlas {
    vkhrt_context* ctx = nullptr;
    vk::InstanceDev* dev = nullptr;
    uint32_t count = 0;
    int variant = vk::kVarMixed;  // kernel variant without FP64 polish
    uint32_t max_depth = 0;       // deepest BLAS (8-wide levels)
    bool polish_ok = true;        // merge keys and compacted-queue entries can encode every (instance, slot)
    uint64_t serial = 0;          // unique per TLAS built in this process (graph reuse key)
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};  // world bounds of all instances
    std::vector<uint32_t> records;  // leaf records per instance
    std::vector<uint32_t> modes;    // geometry mode per instance
};

namespace {
// LDS stack entries per lane for a stack kind.
constexpr uint32_t kStackEntries[4] = {8, 12, 16, 32};

// Stack kind for a TLAS: the context's setting, raised until it holds the deepest tree.
int stack_kind_for(const vkhrt_context* c, const vkhrt_tlas* t) {
    int k = c->stack_kind;
    while (k < 3 && kStackEntries[k] < t->max_depth) ++k;
    return k;
}

// Work items of a frame and the constants of the work item -> pixel division.
void set_work_order(vk::Camera& cm, uint32_t order) {
    cm.order = order;
    const uint32_t side = order == 2 ? 64u : 8u;
    cm.tiles_x = (cm.width + side - 1u) / side;
    const uint32_t tiles_y = (cm.height + side - 1u) / side;
    cm.work = order ? cm.tiles_x * tiles_y * side * side : cm.width * cm.height;
    cm.inv_w = cm.width ? 1.0f / float(cm.width) : 0.0f;
    cm.inv_tiles_x = cm.tiles_x ? 1.0f / float(cm.tiles_x) : 0.0f;
}

bool is_phantom(const vkhrt_tlas* t) { return t->variant == vk::kVarPhConst || t->variant == vk::kVarPhPoint; }

// Whether any instance is a phantom-curve BLAS (the FP64 polish has work).
bool has_phantom(const vkhrt_tlas* t) {
    for (uint32_t m : t->modes)
        if (m == VKHRT_GEOMETRY_PHANTOM_CURVES) return true;
    return false;
}

// Primitive ids a BLAS's hits report: one per record, four triangles per DOTS piece, one per
// grid cell of a voxel BLAS.
uint64_t blas_id_range(const vkhrt_blas* b) {
    if (b->mode == VKHRT_GEOMETRY_DOTS) return uint64_t(b->records) * 4u;
    if (b->mode == VKHRT_GEOMETRY_VOXELS) return uint64_t(uint32_t(b->vres[0])) * uint32_t(b->vres[1]) * uint32_t(b->vres[2]);
    return b->records;
}

// Kernel variant for a TLAS and the polish flag.
int trace_variant(const vkhrt_tlas* t, bool polish) {
    int v = t->variant;
    if (polish && (v == vk::kVarPhConst || v == vk::kVarPhPoint)) ++v;
    return v;
}
}  // namespace

namespace {

int fail(vkhrt_context* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

int hip_fail(vkhrt_context* c, hipError_t e, const char* what) {
    int code = e == hipErrorOutOfMemory ? VKHRT_ERROR_OUT_OF_MEMORY : VKHRT_ERROR_DEVICE;
    return fail(c, code, std::string(what) + ": " + hipGetErrorString(e));
}

// NULL selects HIP's null stream, as for any HIP API.
hipStream_t pick(vkhrt_context*, void* s) { return reinterpret_cast<hipStream_t>(s); }

struct Timer {
    vkhrt_context* c;
    hipStream_t s;
    const char* name;
    hipEvent_t a = nullptr, b = nullptr;
    Timer(vkhrt_context* c_, hipStream_t s_, const char* n) : c(c_), s(s_), name(n) {
        if (c->profiling && name) {  // null name: not timed (the allocation-free trace creates no events)
            hipEventCreate(&a);
            hipEventCreate(&b);
            hipEventRecord(a, s);
        }
    }
    void stop() {
        if (a) {
            hipEventRecord(b, s);
            c->pending.push_back(PendingTiming{name, a, b});
            a = b = nullptr;
        }
    }
    ~Timer() { stop(); }
};

void resolve_timings(vkhrt_context* c) {
    for (auto& p : c->pending) {
        hipEventSynchronize(p.b);
        float ms = 0.0f;
        hipEventElapsedTime(&ms, p.a, p.b);
        auto& s = c->stats[p.name];
        s.first += ms;
        s.second += 1;
        hipEventDestroy(p.a);
        hipEventDestroy(p.b);
    }
    c->pending.clear();
}

// ---- host 4x4 helpers (column-major m[c*4 + r], GLSL/glm conventions)
void inverse4(const float* m, float* out) {
#define A_(r, c) m[(c)*4 + (r)]
    float s0 = A_(0, 0) * A_(1, 1) - A_(1, 0) * A_(0, 1);
    float s1 = A_(0, 0) * A_(1, 2) - A_(1, 0) * A_(0, 2);
    float s2 = A_(0, 0) * A_(1, 3) - A_(1, 0) * A_(0, 3);
    float s3 = A_(0, 1) * A_(1, 2) - A_(1, 1) * A_(0, 2);
    float s4 = A_(0, 1) * A_(1, 3) - A_(1, 1) * A_(0, 3);
    float s5 = A_(0, 2) * A_(1, 3) - A_(1, 2) * A_(0, 3);
    float c5 = A_(2, 2) * A_(3, 3) - A_(3, 2) * A_(2, 3);
    float c4 = A_(2, 1) * A_(3, 3) - A_(3, 1) * A_(2, 3);
    float c3 = A_(2, 1) * A_(3, 2) - A_(3, 1) * A_(2, 2);
    float c2 = A_(2, 0) * A_(3, 3) - A_(3, 0) * A_(2, 3);
    float c1 = A_(2, 0) * A_(3, 2) - A_(3, 0) * A_(2, 2);
    float c0 = A_(2, 0) * A_(3, 1) - A_(3, 0) * A_(2, 1);
    float det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
    float id = 1.0f / det;
    float B[4][4];
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
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) out[c * 4 + r] = B[r][c];
}

struct H3 {
    float x, y, z;
};
H3 h_sub(H3 a, H3 b) { return H3{a.x - b.x, a.y - b.y, a.z - b.z}; }
float h_dot(H3 a, H3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
H3 h_cross(H3 a, H3 b) { return H3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
H3 h_norm(H3 a) {
    float k = 1.0f / std::sqrt(h_dot(a, a));
    return H3{a.x * k, a.y * k, a.z * k};
}

// Decodes the root node's child boxes into object-space bounds.
void root_bounds(const vk::WideNode& N, float lo[3], float hi[3]) {
    for (int k = 0; k < 3; ++k) {
        lo[k] = INFINITY;
        hi[k] = -INFINITY;
    }
    const float org[3] = {N.px, N.py, N.pz};
    const uint32_t* ql[3] = {N.qlo_x, N.qlo_y, N.qlo_z};
    const uint32_t* qh[3] = {N.qhi_x, N.qhi_y, N.qhi_z};
    for (int s = 0; s < 8; ++s) {
        uint32_t meta = (N.meta[s >> 2] >> (8 * (s & 3))) & 0xffu;
        if (!meta) continue;
        for (int k = 0; k < 3; ++k) {
            uint32_t e = (N.e_imask >> (8 * k)) & 0xffu;
            float sc;
            uint32_t bits = e << 23;
            std::memcpy(&sc, &bits, 4);
            float a = float((ql[k][s >> 2] >> (8 * (s & 3))) & 0xffu);
            float b = float((qh[k][s >> 2] >> (8 * (s & 3))) & 0xffu);
            lo[k] = std::min(lo[k], org[k] + a * sc);
            hi[k] = std::max(hi[k], org[k] + b * sc);
        }
    }
    for (int k = 0; k < 3; ++k) {  // absorb decode rounding
        float m = std::max(std::fabs(lo[k]), std::fabs(hi[k]));
        lo[k] -= 1e-4f + 1e-6f * m;
        hi[k] += 1e-4f + 1e-6f * m;
    }
}

// npix: pixels times light samples; nwork: work items of the frame's camera in any work order.
int ensure_render_scratch(vkhrt_context* c, size_t npix, size_t nwork) {
    if (npix <= c->cap_pixels && nwork <= c->cap_work) return VKHRT_SUCCESS;
    npix = std::max(npix, c->cap_pixels);
    nwork = std::max(nwork, c->cap_work);
    hipFree(c->hitbuf);
    hipFree(c->queue);
    hipFree(c->hitrec);
    hipFree(c->vis);
    hipFree(c->live);
    hipFree(c->live_mask);
    c->live = nullptr;
    c->live_mask = nullptr;
    c->hitbuf = nullptr;
    c->queue = nullptr;
    c->hitrec = nullptr;
    c->vis = nullptr;
    c->cap_pixels = 0;
    c->cap_work = 0;
    hipError_t e = hipMalloc(&c->hitbuf, npix * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc(&c->queue, npix * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&c->hitrec, npix * sizeof(vk::HitRec));
    if (e == hipSuccess) e = hipMalloc(&c->vis, npix * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&c->live, npix * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&c->live_mask, (nwork + 63) / 64 * sizeof(uint64_t));
    if (e != hipSuccess) return hip_fail(c, e, "render scratch");
    c->cap_pixels = npix;
    c->cap_work = nwork;
    return VKHRT_SUCCESS;
}

// Deferred-polish scratch for `nrays` rays: merge keys per ray and a candidate queue with
// room for one candidate per ray (at least 64Ki).
int ensure_polish_scratch(vkhrt_context* c, size_t nrays) {
    hipError_t e = hipSuccess;
    if (nrays > c->keys_cap) {
        hipFree(c->keys);
        c->keys = nullptr;
        c->keys_cap = 0;
        e = hipMalloc(&c->keys, nrays * sizeof(unsigned long long));
        if (e != hipSuccess) return hip_fail(c, e, "polish keys");
        c->keys_cap = nrays;
    }
    const size_t cap = std::max<size_t>(nrays, 65536);
    if (cap > c->defer_cap) {
        hipFree(c->defer);
        hipFree(c->cand_hits);
        hipFree(c->cand_keys);
        c->defer = nullptr;
        c->cand_hits = nullptr;
        c->cand_keys = nullptr;
        c->defer_cap = 0;
        e = hipMalloc(&c->defer, cap * sizeof(uint4));
        if (e == hipSuccess) e = hipMalloc(&c->cand_hits, cap * sizeof(vk::HitRec));
        if (e == hipSuccess) e = hipMalloc(&c->cand_keys, cap * sizeof(unsigned long long));
        if (e != hipSuccess) return hip_fail(c, e, "polish queue");
        c->defer_cap = uint32_t(cap);
    }
    return VKHRT_SUCCESS;
}

int ensure_neural_scratch(vkhrt_context* c, size_t npix) {
    if (npix <= c->cap_neural) return VKHRT_SUCCESS;
    hipFree(c->origins);
    hipFree(c->fallback);
    hipFree(c->dirs);
    c->origins = nullptr;
    c->fallback = nullptr;
    c->dirs = nullptr;
    c->cap_neural = 0;
    hipError_t e = hipMalloc(&c->origins, npix * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc(&c->fallback, npix * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&c->dirs, npix * sizeof(float4));
    if (e != hipSuccess) return hip_fail(c, e, "neural shadow scratch");
    c->cap_neural = npix;
    return VKHRT_SUCCESS;
}

std::atomic<uint64_t> g_tlas_serial{0};

void free_batch_scratch(vkhrt_context* c) {
    void* bufs[] = {c->b_ray,    c->b_thr,    c->b_rad,   c->b_nee,      c->b_hits,  c->b_q0,
                    c->b_q1,     c->b_order,  c->b_squeue, c->b_qsort,   c->b_bins,  c->b_counters,
                    c->b_keys,   c->b_defer,  c->b_cand_hits, c->b_cand_keys};
    for (void* b : bufs) hipFree(b);
    c->b_ray = c->b_thr = c->b_rad = c->b_nee = nullptr;
    c->b_hits = c->b_cand_hits = nullptr;
    c->b_q0 = c->b_q1 = c->b_order = c->b_squeue = c->b_qsort = c->b_bins = c->b_counters = nullptr;
    c->b_keys = c->b_cand_keys = nullptr;
    c->b_defer = nullptr;
    c->b_cap = 0;
}

// Per-path batch state for `paths` paths (the polish candidate queue holds one candidate per
// path per depth, at least 64Ki).
int ensure_batch_scratch(vkhrt_context* c, size_t paths) {
    if (paths <= c->b_cap && c->b_counters) return VKHRT_SUCCESS;
    free_batch_scratch(c);
    ++c->batch_gen;
    const size_t cand = std::max<size_t>(paths, 65536);
    hipError_t e = hipMalloc(&c->b_ray, paths * 2 * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc(&c->b_thr, paths * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc(&c->b_rad, paths * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc(&c->b_nee, paths * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc(&c->b_hits, paths * sizeof(vk::HitRec));
    if (e == hipSuccess) e = hipMalloc(&c->b_q0, paths * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&c->b_q1, paths * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&c->b_order, paths * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&c->b_squeue, paths * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&c->b_qsort, paths * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&c->b_bins, 2 * vk::kRayBins * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&c->b_counters, vk::kBatchCounterWords * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&c->b_keys, paths * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&c->b_defer, cand * sizeof(uint4));
    if (e == hipSuccess) e = hipMalloc(&c->b_cand_hits, cand * sizeof(vk::HitRec));
    if (e == hipSuccess) e = hipMalloc(&c->b_cand_keys, cand * sizeof(unsigned long long));
    if (e != hipSuccess) {
        free_batch_scratch(c);
        return hip_fail(c, e, "batch scratch");
    }
    c->b_cap = paths;
    return VKHRT_SUCCESS;
}

// Device view table (+ frame values) and its pinned staging buffer.
int ensure_batch_tables(vkhrt_context* c, uint32_t nviews) {
    if (!c->b_frame) {
        hipError_t e = hipMalloc(&c->b_frame, sizeof(vk::BatchFrame));
        if (e == hipSuccess && !c->b_upload) e = hipEventCreateWithFlags(&c->b_upload, hipEventDisableTiming);
        if (e != hipSuccess) return hip_fail(c, e, "batch frame");
    }
    if (nviews > c->b_views_cap) {
        hipFree(c->b_views);
        c->b_views = nullptr;
        c->b_views_cap = 0;
        ++c->batch_gen;
        hipError_t e = hipMalloc(&c->b_views, size_t(nviews) * sizeof(vk::BatchView));
        if (e != hipSuccess) return hip_fail(c, e, "batch views");
        c->b_views_cap = nviews;
    }
    const size_t need = size_t(nviews) * sizeof(vk::BatchView) + sizeof(vk::BatchFrame);
    if (need > c->b_host_cap) {
        if (c->b_upload) hipEventSynchronize(c->b_upload);
        hipHostFree(c->b_host);
        c->b_host = nullptr;
        c->b_host_cap = 0;
        hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&c->b_host), need, hipHostMallocDefault);
        if (e != hipSuccess) return hip_fail(c, e, "batch staging");
        c->b_host_cap = need;
    }
    return VKHRT_SUCCESS;
}

const float kZeroF = 0.0f;

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void set_polish(vkhrt_context* c, vk::TraceParams& tp) {
    tp.defer = c->defer;
    tp.defer_count = c->counters + 4;
    tp.defer_cap = c->defer_cap;
    tp.keys = c->keys;
    tp.cand_hits = c->cand_hits;
    tp.cand_keys = c->cand_keys;
}

}  // namespace

extern "C" {

int vkhrt_abi_version(void) { return VKHRT_ABI_VERSION; }

int vkhrt_context_create(int device, vkhrt_context** out) {
    if (!out) return VKHRT_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return VKHRT_ERROR_NO_DEVICE;
    if (device < 0 || device >= n) return VKHRT_ERROR_INVALID_ARGUMENT;
    if (hipSetDevice(device) != hipSuccess) return VKHRT_ERROR_DEVICE;
    vkhrt_context* c = new (std::nothrow) vkhrt_context();
    if (!c) return VKHRT_ERROR_OUT_OF_MEMORY;
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&c->status, 16);
    if (e == hipSuccess) e = hipMalloc(&c->counters, vk::kRenderCounterWords * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(c->status, 0, 16);
#ifdef VKHRT_WORK_COUNTERS
    if (e == hipSuccess) e = hipMalloc(&c->work_counts, vk::kWorkCounters * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(c->work_counts, 0, vk::kWorkCounters * sizeof(unsigned long long));
#endif
    if (e != hipSuccess) {
        vkhrt_context_destroy(c);
        return VKHRT_ERROR_DEVICE;
    }
    // the allocation-free trace's scratch array on this device (left null if the lookup fails;
    // vkhrt_trace_rays_noalloc then reports it)
    if (vk::noalloc_scratch_lookup(device, &c->noalloc) != hipSuccess) c->noalloc = nullptr;
    *out = c;
    return VKHRT_SUCCESS;
}

void vkhrt_context_destroy(vkhrt_context* c) {
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (auto& p : c->pending) {
        hipEventDestroy(p.a);
        hipEventDestroy(p.b);
    }
    hipFree(c->status);
    hipFree(c->work_counts);
    hipFree(c->counters);
    hipFree(c->hitbuf);
    hipFree(c->queue);
    hipFree(c->hitrec);
    hipFree(c->vis);
    hipFree(c->live);
    hipFree(c->live_mask);
    hipFree(c->defer);
    hipFree(c->cand_hits);
    hipFree(c->cand_keys);
    hipFree(c->keys);
    hipFree(c->origins);
    hipFree(c->fallback);
    hipFree(c->dirs);
    free_batch_scratch(c);
    hipFree(c->b_views);
    hipFree(c->b_frame);
    if (c->b_upload) hipEventSynchronize(c->b_upload);
    if (c->b_upload) hipEventDestroy(c->b_upload);
    hipHostFree(c->b_host);
    if (c->graph_exec) hipGraphExecDestroy(c->graph_exec);
    if (c->graph) hipGraphDestroy(c->graph);
    if (c->capture_stream) hipStreamDestroy(c->capture_stream);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

const char* vkhrt_last_error(const vkhrt_context* c) { return c ? c->err.c_str() : "null context"; }

int vkhrt_synchronize(vkhrt_context* c, void* stream) {
    if (!c) return VKHRT_ERROR_INVALID_ARGUMENT;
    hipError_t e = hipStreamSynchronize(pick(c, stream));
    if (e != hipSuccess) return hip_fail(c, e, "synchronize");
    uint32_t st[4];
    e = hipMemcpy(st, c->status, 16, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(c, e, "status readback");
    if (st[0] || st[1]) {
        hipMemset(c->status, 0, 16);
        if (st[0])
            return fail(c, VKHRT_ERROR_STACK_OVERFLOW,
                        "traversal stack overflow on " + std::to_string(st[0]) + " rays (results incomplete)");
        return fail(c, VKHRT_ERROR_OUT_OF_MEMORY,
                    "FP64 polish queue overflow: " + std::to_string(st[1]) + " candidates dropped (results incomplete)");
    }
    return VKHRT_SUCCESS;
}

int vkhrt_set_option(vkhrt_context* c, int option, int value) {
    if (!c) return VKHRT_ERROR_INVALID_ARGUMENT;
    if (option == VKHRT_OPTION_STACK_KIND) {
        if (value < 0 || value > 3) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "stack kind must be 0..3");
        c->stack_kind = value;
        return VKHRT_SUCCESS;
    }
    if (option == VKHRT_OPTION_PERSISTENT) {
        if (value < 0 || value > 1) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "persistent must be 0 or 1");
        c->persistent = value != 0;
        return VKHRT_SUCCESS;
    }
    if (option == VKHRT_OPTION_REFILL_MIN) {
        if (value < 1 || value > 64) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "refill minimum must be 1..64");
        c->refill_min = uint32_t(value);
        return VKHRT_SUCCESS;
    }
    if (option == VKHRT_OPTION_RAY_ORDER) {
        if (value < 0 || value > 2) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "ray order must be 0, 1 or 2");
        c->ray_order = uint32_t(value);
        return VKHRT_SUCCESS;
    }
    if (option == VKHRT_OPTION_BLOCKS_PER_CU) {
        if (value < 0 || value > 32) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "blocks per CU must be 0..32");
        c->blocks_per_cu = uint32_t(value);
        return VKHRT_SUCCESS;
    }
    if (option == VKHRT_OPTION_XCD_CHUNK) {
        if (value != 0 && (value < 64 || value > (1 << 20) || (value & (value - 1)) != 0))
            return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "XCD chunk must be 0 or a power of two in 64..2^20");
        c->xcd_chunk = uint32_t(value);
        return VKHRT_SUCCESS;
    }
    if (option == VKHRT_OPTION_STEPS_PER_ITERATION) {
        if (value < 1 || value > 8) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "steps per iteration must be 1..8");
        c->steps_per_iter = uint32_t(value);
        return VKHRT_SUCCESS;
    }
    if (option == VKHRT_OPTION_SHADOW_GROUP_LANES) {
        if (value != 8 && value != 16 && value != 32 && value != 64)
            return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "shadow group lanes must be 8, 16, 32 or 64");
        c->shadow_group_lanes = uint32_t(value);
        return VKHRT_SUCCESS;
    }
    if (option == VKHRT_OPTION_SHADOW_WAVE_LIMIT) {
        if (value < 0 || value > (1 << 20))
            return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "shadow wave limit must be 0..2^20");
        c->shadow_wave_limit = uint32_t(value);
        return VKHRT_SUCCESS;
    }
    if (option == VKHRT_OPTION_BVH_COLLAPSE) {
        if (value < 0 || value > 8) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "BVH collapse must be 0..8");
        c->bvh_collapse = value;
        return VKHRT_SUCCESS;
    }
    if (option == VKHRT_OPTION_BVH_WIDTH) {
        if (value < 4 || value > 8) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "BVH width must be 4..8");
        c->bvh_width = value;
        return VKHRT_SUCCESS;
    }
    if (option == VKHRT_OPTION_LEAF_SIZE) {
        if (value < 1 || value > 3) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "leaf size must be 1..3");
        c->leaf_size = value;
        return VKHRT_SUCCESS;
    }
    return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "unknown option");
}

namespace {

// Voxel BLAS (BLASType::eVoxels): ProcessHairVoxels' grid and fill (geometry_processor.cpp:
// 690-737, GenerateVoxelMesh :515-602) on the device, then `voxel_lod` OR-reductions.
int build_voxels(vkhrt_context* c, const vkhrt_blas_desc* d, hipStream_t st, vkhrt_blas* b) {
    if (!(d->voxel_size > 0.0f) || !std::isfinite(d->voxel_size))
        return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "voxel size must be finite and > 0");
    if (!(d->radius >= 0.0f) || !std::isfinite(d->radius))
        return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "voxel radius must be finite and >= 0");
    if (d->voxel_lod > 15) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "voxel level of detail must be 0..15");
    if (d->flags & VKHRT_BLAS_PER_POINT_RADIUS)
        return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "voxels take one radius, not per-point radii");
    const size_t line_bytes = size_t(d->count) * 8 * sizeof(float);
    float4* lines = nullptr;
    bool own = false;
    uint32_t* bits = nullptr;
    uint32_t* bricks = nullptr;
    uint32_t* blocks = nullptr;
    uint32_t* scratch = nullptr;  // [0] overflow, [1] filled count, [2..7] bounds keys
    int rc = VKHRT_SUCCESS;
    hipError_t e = hipMalloc(&scratch, 8 * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemsetAsync(scratch, 0, 8 * sizeof(uint32_t), st);
    if (e == hipSuccess && d->count) {
        if (d->flags & VKHRT_BLAS_HOST_POINTERS) {
            e = hipMalloc(&lines, line_bytes);
            if (e == hipSuccess) e = hipMemcpyAsync(lines, d->primitives, line_bytes, hipMemcpyHostToDevice, st);
            own = true;
        } else {
            lines = reinterpret_cast<float4*>(const_cast<float*>(d->primitives));
        }
    }
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    if (e == hipSuccess) {
        if (d->flags & VKHRT_BLAS_HAS_BOUNDS) {
            std::memcpy(lo, d->bounds_min, sizeof(lo));
            std::memcpy(hi, d->bounds_max, sizeof(hi));
        } else if (d->count) {
            int32_t keys[6];
            e = vk::voxel_line_bounds(lines, d->count, reinterpret_cast<int32_t*>(scratch + 2), st);
            if (e == hipSuccess) e = hipMemcpyAsync(keys, scratch + 2, sizeof(keys), hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e == hipSuccess) vk::voxel_decode_bounds(keys, lo, hi);
        }
    }
    vk::VoxelGrid g{};
    int32_t res[3] = {0, 0, 0};
    float cell = d->voxel_size;
    uint32_t filled = 0;
    if (e != hipSuccess) {
        rc = hip_fail(c, e, "voxel input");
        goto done;
    }
    if (!vk::voxel_grid_params(lo, hi, d->voxel_size, g)) {
        rc = fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "voxel grid: bounds not finite, or 2^31 cells or more");
        goto done;
    }
    std::memcpy(res, g.res, sizeof(res));
    if (g.count) {
        uint32_t overflow = 0;
        const size_t words = (size_t(g.count) + 31) / 32;
        e = hipMalloc(&bits, words * sizeof(uint32_t));
        if (e == hipSuccess) e = hipMemsetAsync(bits, 0, words * sizeof(uint32_t), st);
        if (e == hipSuccess)
            e = vk::voxelize_lines(lines, d->count, g, d->radius, (d->flags & VKHRT_BLAS_VOXEL_REFERENCE_SWEEP) != 0,
                                   bits, scratch, st);
        if (e == hipSuccess) e = hipMemcpyAsync(&overflow, scratch, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            rc = hip_fail(c, e, "voxelize");
            goto done;
        }
        if (overflow) {
            rc = fail(c, VKHRT_ERROR_INVALID_ARGUMENT,
                      std::to_string(overflow) + " lines cover more than 65536 voxel cells each (voxel size too small)");
            goto done;
        }
        for (uint32_t l = 0; l < d->voxel_lod; ++l) {
            const int32_t cres[3] = {(res[0] + 1) / 2, (res[1] + 1) / 2, (res[2] + 1) / 2};
            const size_t cwords = (size_t(cres[0]) * size_t(cres[1]) * size_t(cres[2]) + 31) / 32;
            uint32_t* coarse = nullptr;
            e = hipMalloc(&coarse, cwords * sizeof(uint32_t));
            if (e == hipSuccess) e = vk::voxel_downsample(bits, res, coarse, cres, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e != hipSuccess) {
                hipFree(coarse);
                rc = hip_fail(c, e, "voxel level of detail");
                goto done;
            }
            hipFree(bits);
            bits = coarse;
            std::memcpy(res, cres, sizeof(res));
            cell = cell * 2.0f;
        }
        const uint32_t nwords = uint32_t((uint64_t(res[0]) * uint64_t(res[1]) * uint64_t(res[2]) + 31) / 32);
        e = vk::voxel_count(bits, nwords, scratch + 1, st);
        if (e == hipSuccess) e = hipMemcpyAsync(&filled, scratch + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            rc = hip_fail(c, e, "voxel count");
            goto done;
        }
        // Empty-space levels: pairs of 2x downsamplings give one bit per 4x4x4 cells (brick
        // (X, Y, Z) covers cells 4X..4X+3 per axis, clipped to the grid), then per 16x16x16.
        int32_t hres[3], bres[3], h2res[3], kres[3];
        for (int k = 0; k < 3; ++k) {
            hres[k] = (res[k] + 1) / 2;
            bres[k] = (hres[k] + 1) / 2;
            h2res[k] = (bres[k] + 1) / 2;
            kres[k] = (h2res[k] + 1) / 2;
        }
        const size_t hwords = (size_t(hres[0]) * size_t(hres[1]) * size_t(hres[2]) + 31) / 32;
        const size_t bwords = (size_t(bres[0]) * size_t(bres[1]) * size_t(bres[2]) + 31) / 32;
        const size_t kwords = (size_t(kres[0]) * size_t(kres[1]) * size_t(kres[2]) + 31) / 32;
        uint32_t* half = nullptr;
        e = hipMalloc(&half, hwords * sizeof(uint32_t));
        if (e == hipSuccess) e = hipMalloc(&bricks, bwords * sizeof(uint32_t));
        if (e == hipSuccess) e = hipMalloc(&blocks, kwords * sizeof(uint32_t));
        if (e == hipSuccess) e = vk::voxel_downsample(bits, res, half, hres, st);
        if (e == hipSuccess) e = vk::voxel_downsample(half, hres, bricks, bres, st);
        if (e == hipSuccess) e = vk::voxel_downsample(bricks, bres, half, h2res, st);
        if (e == hipSuccess) e = vk::voxel_downsample(half, h2res, blocks, kres, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        hipFree(half);
        if (e != hipSuccess) {
            rc = hip_fail(c, e, "voxel bricks");
            goto done;
        }
        b->info.device_bytes = uint64_t(nwords) * 4u + uint64_t(bwords) * 4u + uint64_t(kwords) * 4u;
    }
    b->voxels = bits;
    bits = nullptr;
    b->vbricks = bricks;
    bricks = nullptr;
    b->vblocks = blocks;
    blocks = nullptr;
    b->records = filled;
    b->vcell = cell;
    std::memcpy(b->vres, res, sizeof(res));
    std::memcpy(b->vorigin, g.origin, sizeof(b->vorigin));
    {
        // object bounds: the grid box, padded like the BVH boxes
        float m = 0.0f;
        for (int k = 0; k < 3; ++k) {
            b->lo[k] = g.origin[k];
            b->hi[k] = g.origin[k] + float(res[k]) * cell;
            m = std::max(m, std::max(std::fabs(b->lo[k]), std::fabs(b->hi[k])));
        }
        for (int k = 0; k < 3; ++k) {
            b->lo[k] -= 1e-4f + 4e-7f * m;
            b->hi[k] += 1e-4f + 4e-7f * m;
        }
    }
done:
    hipStreamSynchronize(st);
    if (own) hipFree(lines);
    hipFree(bits);
    hipFree(bricks);
    hipFree(blocks);
    hipFree(scratch);
    return rc;
}

}  // namespace

int vkhrt_blas_build(vkhrt_context* c, const vkhrt_blas_desc* d, void* stream, vkhrt_blas** out) {
    if (!c || !d || !out) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (d->mode > VKHRT_GEOMETRY_VOXELS) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "unknown geometry mode");
    if (d->count && !d->primitives) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "null primitives");
    const bool mesh = d->mode == VKHRT_GEOMETRY_TRIANGLES;
    if (mesh && d->count && (!d->indices || d->vertex_count == 0))
        return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "a triangle mesh needs indices and vertices");
    if (d->count > (1u << 28)) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "too many primitives");
    if (d->leaf_size > 3) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "leaf size must be 0..3");
    const bool per_point = (d->flags & VKHRT_BLAS_PER_POINT_RADIUS) != 0;
    if (!per_point && !(d->radius >= 0.0f)) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "radius must be >= 0");
    hipSetDevice(c->device);
    hipStream_t st = pick(c, stream);
    if (d->mode == VKHRT_GEOMETRY_VOXELS) {
        vkhrt_blas* vb = new (std::nothrow) vkhrt_blas();
        if (!vb) return fail(c, VKHRT_ERROR_OUT_OF_MEMORY, "blas");
        vb->ctx = c;
        vb->mode = d->mode;
        vb->radius = d->radius;
        vb->material_index = d->material_index;
        vb->has_xf = (d->flags & VKHRT_BLAS_HAS_TRANSFORM) != 0;
        std::memcpy(vb->xf, d->transform, sizeof(vb->xf));
        const auto t0 = std::chrono::steady_clock::now();
        const int rc = build_voxels(c, d, st, vb);
        if (rc != VKHRT_SUCCESS) {
            vkhrt_blas_destroy(vb);
            return rc;
        }
        vb->info.mode = d->mode;
        vb->info.primitive_count = vb->info.leaf_records = vb->records;
        vb->info.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        *out = vb;
        return VKHRT_SUCCESS;
    }
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipEventRecord(e0, st);
    const uint32_t in_floats = d->mode == VKHRT_GEOMETRY_PHANTOM_CURVES ? 16u : 8u;
    const size_t in_bytes = (mesh ? size_t(d->vertex_count) : size_t(d->count)) * in_floats * sizeof(float);
    const size_t idx_bytes = mesh ? size_t(d->count) * 3 * sizeof(uint32_t) : 0;
    float4* input = nullptr;
    float4* records = nullptr;
    uint32_t* indices = nullptr;
    uint32_t* bad = nullptr;
    bool own_input = false;
    vkhrt_blas* b = new (std::nothrow) vkhrt_blas();
    if (!b) return fail(c, VKHRT_ERROR_OUT_OF_MEMORY, "blas");
    b->ctx = c;
    b->mode = d->mode;
    b->radius_mode = per_point ? 1u : 0u;
    b->radius = d->radius;
    b->material_index = d->material_index;
    b->has_xf = (d->flags & VKHRT_BLAS_HAS_TRANSFORM) != 0;
    std::memcpy(b->xf, d->transform, sizeof(b->xf));
    b->records = d->count;
    hipError_t e = hipSuccess;
    std::string err;
    int rc = VKHRT_SUCCESS;
    if (d->count) {
        if (d->flags & VKHRT_BLAS_HOST_POINTERS) {
            e = hipMalloc(&input, in_bytes);
            if (e == hipSuccess) e = hipMemcpyAsync(input, d->primitives, in_bytes, hipMemcpyHostToDevice, st);
            if (e == hipSuccess && mesh) e = hipMalloc(&indices, idx_bytes);
            if (e == hipSuccess && mesh) e = hipMemcpyAsync(indices, d->indices, idx_bytes, hipMemcpyHostToDevice, st);
            own_input = true;
        } else {
            input = reinterpret_cast<float4*>(const_cast<float*>(d->primitives));
            indices = const_cast<uint32_t*>(d->indices);
        }
        if (e != hipSuccess) {
            rc = hip_fail(c, e, "blas upload");
            goto out;
        }
        if (d->mode == VKHRT_GEOMETRY_PHANTOM_CURVES) {
            records = input;
        } else if (d->mode == VKHRT_GEOMETRY_LSS) {
            e = hipMalloc(&records, size_t(d->count) * 2 * sizeof(float4));
            if (e == hipSuccess) e = vk::lss_chain(input, d->count, records, st);
        } else if (mesh) {
            uint32_t nbad = 0;
            e = hipMalloc(&records, size_t(d->count) * 3 * sizeof(float4));
            if (e == hipSuccess) e = hipMalloc(&b->normals, size_t(d->count) * 3 * sizeof(float4));
            if (e == hipSuccess) e = hipMalloc(&bad, sizeof(uint32_t));
            if (e == hipSuccess) e = hipMemsetAsync(bad, 0, sizeof(uint32_t), st);
            if (e == hipSuccess)
                e = vk::gen_tri_records(reinterpret_cast<const float*>(input), d->vertex_count, indices, d->count,
                                        records, b->normals, bad, st);
            if (e == hipSuccess) e = hipMemcpyAsync(&nbad, bad, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e == hipSuccess && nbad) {
                rc = fail(c, VKHRT_ERROR_INVALID_ARGUMENT,
                          std::to_string(nbad) + " triangle indices are not below the vertex count");
                goto out;
            }
        } else {
            e = hipMalloc(&records, size_t(d->count) * 4 * sizeof(float4));
            if (e == hipSuccess) e = vk::gen_dots(input, d->count, per_point ? -1.0f : d->radius, records, st);
        }
        if (e != hipSuccess) {
            rc = hip_fail(c, e, "blas records");
            goto out;
        }
        {
            Timer tm(c, st, "blas_build");
            e = vk::build_bvh(int(d->mode), records, d->count, int(b->radius_mode), d->radius,
                              d->leaf_size ? int(d->leaf_size) : c->leaf_size, c->bvh_width, c->bvh_collapse, st,
                              &b->bvh, err);
        }
        if (e != hipSuccess) {
            rc = fail(c, e == hipErrorOutOfMemory ? VKHRT_ERROR_OUT_OF_MEMORY : VKHRT_ERROR_DEVICE, "bvh build: " + err);
            goto out;
        }
        vk::WideNode root;
        e = hipMemcpyAsync(&root, b->bvh.nodes, sizeof(root), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            rc = hip_fail(c, e, "root readback");
            goto out;
        }
        root_bounds(root, b->lo, b->hi);
        if (d->mode == VKHRT_GEOMETRY_PHANTOM_CURVES) {  // for the double-word re-run of final hits
            e = hipMalloc(&b->slot_of, size_t(d->count) * sizeof(uint32_t));
            if (e == hipSuccess) e = vk::gen_slot_of(b->bvh.prim_ids, d->count, b->slot_of, st);
            if (e != hipSuccess) {
                rc = hip_fail(c, e, "slot map");
                goto out;
            }
        }
    } else {
        for (int k = 0; k < 3; ++k) {
            b->lo[k] = INFINITY;
            b->hi[k] = -INFINITY;
        }
    }
    hipEventRecord(e1, st);
    hipEventSynchronize(e1);
    {
        float ms = 0.0f;
        hipEventElapsedTime(&ms, e0, e1);
        b->info.build_ms = ms;
    }
    b->info.mode = d->mode;
    b->info.leaf_records = d->count;
    b->info.primitive_count = d->mode == VKHRT_GEOMETRY_DOTS ? d->count * 4u : d->count;
    b->info.node_count = b->bvh.nnodes;
    b->info.depth = b->bvh.depth;
    b->info.device_bytes = uint64_t(b->bvh.nnodes) * sizeof(vk::WideNode) + uint64_t(d->count) * 4u +
                           uint64_t(d->count) * (d->mode == VKHRT_GEOMETRY_LSS ? 32u : mesh ? 96u : 64u) +
                           (d->mode == VKHRT_GEOMETRY_PHANTOM_CURVES ? uint64_t(d->count) * 4u : 0u);
out:
    if (own_input) {
        hipFree(input);
        hipFree(indices);
    }
    hipFree(bad);
    if (records && records != input) hipFree(records);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc != VKHRT_SUCCESS) {
        vkhrt_blas_destroy(b);
        return rc;
    }
    *out = b;
    return VKHRT_SUCCESS;
}

int vkhrt_blas_voxel_grid_get(const vkhrt_blas* b, vkhrt_voxel_grid* out) {
    if (!b || !out || b->mode != VKHRT_GEOMETRY_VOXELS) return VKHRT_ERROR_INVALID_ARGUMENT;
    std::memcpy(out->res, b->vres, sizeof(out->res));
    std::memcpy(out->origin, b->vorigin, sizeof(out->origin));
    out->cell = b->vcell;
    out->words = (uint64_t(b->vres[0]) * uint64_t(b->vres[1]) * uint64_t(b->vres[2]) + 31) / 32;
    return VKHRT_SUCCESS;
}

int vkhrt_blas_voxels_read(vkhrt_context* c, const vkhrt_blas* b, uint32_t* words, uint64_t count) {
    if (!c || !b || (count && !words)) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "null argument");
    if (b->mode != VKHRT_GEOMETRY_VOXELS) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "not a voxel BLAS");
    const uint64_t have = (uint64_t(b->vres[0]) * uint64_t(b->vres[1]) * uint64_t(b->vres[2]) + 31) / 32;
    if (count > have) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "more words than the grid holds");
    if (count == 0) return VKHRT_SUCCESS;
    hipSetDevice(c->device);
    hipError_t e = hipMemcpy(words, b->voxels, count * sizeof(uint32_t), hipMemcpyDeviceToHost);
    return e == hipSuccess ? VKHRT_SUCCESS : hip_fail(c, e, "voxel readback");
}

int vkhrt_blas_info_get(const vkhrt_blas* b, vkhrt_blas_info* out) {
    if (!b || !out) return VKHRT_ERROR_INVALID_ARGUMENT;
    *out = b->info;
    return VKHRT_SUCCESS;
}

void vkhrt_blas_destroy(vkhrt_blas* b) {
    if (!b) return;
    hipFree(b->bvh.nodes);
    hipFree(b->bvh.prim_ids);
    hipFree(b->bvh.prims);
    hipFree(b->bvh.aux);
    hipFree(b->normals);
    hipFree(b->voxels);
    hipFree(b->vbricks);
    hipFree(b->vblocks);
    hipFree(b->slot_of);
    delete b;
}

int vkhrt_tlas_build(vkhrt_context* c, const vkhrt_blas* const* blases, uint32_t count, void* stream,
                     vkhrt_tlas** out) {
    if (!c || !out || (count && !blases)) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    std::vector<vk::InstanceDev> inst(count);
    uint64_t gid_total = 0;  // global primitive ids of the merge keys, instance by instance
    for (uint32_t i = 0; i < count; ++i) {
        const vkhrt_blas* b = blases[i];
        if (!b) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "null blas");
        vk::InstanceDev& I = inst[i];
        std::memset(&I, 0, sizeof(I));
        I.blas.gbase = uint32_t(std::min<uint64_t>(gid_total, 0xFFFFFFFFull));
        gid_total += blas_id_range(b);
        I.blas.prims = b->bvh.prims;
        I.blas.aux = b->bvh.aux;
        I.blas.prim_ids = b->bvh.prim_ids;
        I.blas.nodes = b->bvh.nodes;
        I.blas.nprims = b->records;
        I.blas.nnodes = b->bvh.nnodes;
        I.blas.mode = b->mode;
        I.blas.radius_mode = b->radius_mode;
        I.blas.radius = b->radius;
        I.blas.normals = b->normals;
        I.blas.slot_of = b->slot_of;
        I.blas.voxels = b->voxels;
        I.blas.vbricks = b->vbricks;
        I.blas.vblocks = b->vblocks;
        std::memcpy(I.blas.vorigin, b->vorigin, sizeof(I.blas.vorigin));
        I.blas.vcell = b->vcell;
        std::memcpy(I.blas.vres, b->vres, sizeof(I.blas.vres));
        I.material = b->material_index;
        const float ident[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
        const float* xf = b->has_xf ? b->xf : ident;
        std::memcpy(I.xf, xf, sizeof(I.xf));
        I.identity = std::memcmp(xf, ident, sizeof(ident)) == 0 ? 1u : 0u;
        float M[16] = {xf[0], xf[4], xf[8], 0, xf[1], xf[5], xf[9], 0, xf[2], xf[6], xf[10], 0, xf[3], xf[7], xf[11], 1};
        inverse4(M, I.inv);
        // world bounds of the transformed object box
        for (int k = 0; k < 3; ++k) {
            I.lo[k] = INFINITY;
            I.hi[k] = -INFINITY;
        }
        if (b->records) {
            for (int corner = 0; corner < 8; ++corner) {
                float p[3] = {(corner & 1) ? b->hi[0] : b->lo[0], (corner & 2) ? b->hi[1] : b->lo[1],
                              (corner & 4) ? b->hi[2] : b->lo[2]};
                for (int r = 0; r < 3; ++r) {
                    float w = xf[r * 4 + 0] * p[0] + xf[r * 4 + 1] * p[1] + xf[r * 4 + 2] * p[2] + xf[r * 4 + 3];
                    I.lo[r] = std::min(I.lo[r], w);
                    I.hi[r] = std::max(I.hi[r], w);
                }
            }
            if (!I.identity)
                for (int r = 0; r < 3; ++r) {
                    float m = std::max(std::fabs(I.lo[r]), std::fabs(I.hi[r]));
                    I.lo[r] -= 1e-4f + 1e-5f * m;
                    I.hi[r] += 1e-4f + 1e-5f * m;
                }
        }
    }
    vkhrt_tlas* t = new (std::nothrow) vkhrt_tlas();
    if (!t) return fail(c, VKHRT_ERROR_OUT_OF_MEMORY, "tlas");
    t->ctx = c;
    t->count = count;
    t->serial = ++g_tlas_serial;
    for (uint32_t i = 0; i < count; ++i) {
        t->records.push_back(blases[i]->records);
        t->modes.push_back(blases[i]->mode);
    }
    for (int k = 0; k < 3; ++k) {
        t->lo[k] = INFINITY;
        t->hi[k] = -INFINITY;
    }
    for (const vk::InstanceDev& I : inst)
        for (int k = 0; k < 3; ++k)
            if (I.lo[k] <= I.hi[k]) {
                t->lo[k] = std::min(t->lo[k], I.lo[k]);
                t->hi[k] = std::max(t->hi[k], I.hi[k]);
            }
    for (int k = 0; k < 3; ++k)
        if (!(t->lo[k] <= t->hi[k])) t->lo[k] = t->hi[k] = 0.0f;
    {
        int v = -1;
        for (uint32_t i = 0; i < count; ++i) {
            const vkhrt_blas* b = blases[i];
            int vi = b->mode == VKHRT_GEOMETRY_VOXELS      ? vk::kVarMixed  // per-instance dispatch walks the grid
                     : b->mode == VKHRT_GEOMETRY_LSS       ? vk::kVarLss
                     : b->mode == VKHRT_GEOMETRY_DOTS      ? vk::kVarDots
                     : b->mode == VKHRT_GEOMETRY_TRIANGLES ? vk::kVarTris
                     : b->radius_mode                      ? vk::kVarPhPoint
                                                           : vk::kVarPhConst;
            v = (v < 0 || v == vi) ? vi : vk::kVarMixed;
            t->max_depth = std::max(t->max_depth, b->bvh.depth);
            // the compacted phantom lanes queue (lane, leaf slot) pairs in 32 bits
            if (b->mode == VKHRT_GEOMETRY_PHANTOM_CURVES && b->records >= (1u << vk::kCompactSlotBits))
                t->polish_ok = false;
        }
        if (gid_total >= 0xFFFFFFFFull) t->polish_ok = false;  // merge keys need every global id below 2^32 - 1
        if (t->max_depth > kStackEntries[3]) {
            vkhrt_tlas_destroy(t);
            return fail(c, VKHRT_ERROR_UNSUPPORTED, "BVH deeper than the largest traversal stack (32 levels)");
        }
        t->variant = v < 0 ? vk::kVarPhConst : v;
    }
    if (count) {
        hipStream_t st = pick(c, stream);
        hipError_t e = hipMalloc(&t->dev, sizeof(vk::InstanceDev) * count);
        if (e == hipSuccess)
            e = hipMemcpyAsync(t->dev, inst.data(), sizeof(vk::InstanceDev) * count, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            vkhrt_tlas_destroy(t);
            return hip_fail(c, e, "tlas upload");
        }
    }
    *out = t;
    return VKHRT_SUCCESS;
}

void vkhrt_tlas_destroy(vkhrt_tlas* t) {
    if (!t) return;
    hipFree(t->dev);
    delete t;
}

}  // extern "C"

namespace {

// Scratch of one trace call: the work counter, the 8 per-XCD counters, and (with FP64 polish)
// the merge keys and candidate queue.
struct TraceScratch {
    uint32_t* work;
    uint32_t* xcd_work;
    uint32_t* defer_count;
    uint4* defer;
    uint32_t defer_cap;
    unsigned long long* keys;
    vk::HitRec* cand_hits;
    unsigned long long* cand_keys;
};

int check_trace_args(vkhrt_context* c, const vkhrt_tlas* t, const vkhrt_ray* rays, uint32_t count, vkhrt_hit* hits,
                     uint32_t flags) {
    if (!c || !t) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "null argument");
    if (count && (!rays || !hits)) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "null rays or hits");
    if (flags & ~(VKHRT_TRACE_TERMINATE_ON_FIRST_HIT | VKHRT_TRACE_FP64_POLISH | VKHRT_TRACE_ACCURATE_NORMALS))
        return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "unknown trace flags");
    if ((flags & VKHRT_TRACE_FP64_POLISH) && !t->polish_ok)
        return fail(c, VKHRT_ERROR_UNSUPPORTED,
                    "FP64 polish needs phantom BLASes of < 2^26 records and < 2^32 - 1 primitive ids per TLAS");
    return VKHRT_SUCCESS;
}

// The trace launches of vkhrt_trace_rays and vkhrt_trace_rays_noalloc (arguments checked).
int trace_launch(vkhrt_context* c, const vkhrt_tlas* t, const vkhrt_ray* rays, uint32_t count, vkhrt_hit* hits,
                 uint32_t flags, hipStream_t st, const TraceScratch& s, bool timed) {
    const bool accurate = (flags & VKHRT_TRACE_ACCURATE_NORMALS) && !(flags & VKHRT_TRACE_TERMINATE_ON_FIRST_HIT);
    flags &= ~uint32_t(VKHRT_TRACE_ACCURATE_NORMALS);
    // persistent phantom kernels compact intersection work through keys that need polish_ok
    const bool persistent = c->persistent && (t->polish_ok || !is_phantom(t));
    vk::TraceParams tp{t->dev, t->count, flags, c->status, persistent ? s.work : nullptr, c->refill_min};
    tp.work_counts = c->work_counts;
    tp.blocks_per_cu = c->blocks_per_cu;
    tp.xcd_chunk = c->xcd_chunk;
    tp.steps_per_iter = c->steps_per_iter;
    tp.xcd_work = s.xcd_work;
    const bool polish = (flags & VKHRT_TRACE_FP64_POLISH) != 0;
    if (polish) {
        tp.defer = s.defer;
        tp.defer_count = s.defer_count;
        tp.defer_cap = s.defer_cap;
        tp.keys = s.keys;
        tp.cand_hits = s.cand_hits;
        tp.cand_keys = s.cand_keys;
        if (!tp.work) tp.work = s.work;  // the deferred pipeline runs on persistent kernels
    }
    Timer tm(c, st, timed ? "trace_rays" : nullptr);
    hipError_t e = vk::launch_trace_rays(tp, reinterpret_cast<const float*>(rays), count,
                                         reinterpret_cast<vk::HitRec*>(hits), nullptr, trace_variant(t, polish),
                                         stack_kind_for(c, t), st);
    if (e == hipSuccess && accurate)
        e = vk::launch_accurate_hits(tp, reinterpret_cast<const float*>(rays), count, reinterpret_cast<vk::HitRec*>(hits),
                                     st);
    tm.stop();
    if (e != hipSuccess) return hip_fail(c, e, "trace launch");
    return VKHRT_SUCCESS;
}

}  // namespace

extern "C" {

int vkhrt_trace_rays(vkhrt_context* c, const vkhrt_tlas* t, const vkhrt_ray* rays, uint32_t count, vkhrt_hit* hits,
                     uint32_t flags, void* stream) {
    int rc = check_trace_args(c, t, rays, count, hits, flags);
    if (rc != VKHRT_SUCCESS) return rc;
    if (flags & VKHRT_TRACE_FP64_POLISH) {
        rc = ensure_polish_scratch(c, count);
        if (rc != VKHRT_SUCCESS) return rc;
    }
    const TraceScratch s{c->counters + 3, c->counters + vk::kXcdCounters + 16 * vk::kXcdStride, c->counters + 4,
                         c->defer, c->defer_cap, c->keys, c->cand_hits, c->cand_keys};
    return trace_launch(c, t, rays, count, hits, flags, pick(c, stream), s, true);
}

int vkhrt_trace_rays_noalloc(vkhrt_context* c, const vkhrt_tlas* t, const vkhrt_ray* rays, uint32_t count,
                             vkhrt_hit* hits, uint32_t flags, void* stream) {
    int rc = check_trace_args(c, t, rays, count, hits, flags);
    if (rc != VKHRT_SUCCESS) return rc;
    if (!c->noalloc) return fail(c, VKHRT_ERROR_DEVICE, "allocation-free trace: no scratch address on this device");
    if ((flags & VKHRT_TRACE_FP64_POLISH) && count > vk::kNoAllocMaxPolishRays)
        return fail(c, VKHRT_ERROR_INVALID_ARGUMENT,
                    "allocation-free trace with FP64 polish takes at most " + std::to_string(vk::kNoAllocMaxPolishRays) +
                        " rays per call (asked for " + std::to_string(count) + ")");
    auto* base = static_cast<unsigned char*>(c->noalloc);
    auto* ctr = reinterpret_cast<uint32_t*>(base);
    const TraceScratch s{ctr, ctr + 32, ctr + 1, reinterpret_cast<uint4*>(base + vk::kNoAllocDeferOffset),
                         vk::kNoAllocMaxPolishRays, reinterpret_cast<unsigned long long*>(base + vk::kNoAllocKeysOffset),
                         reinterpret_cast<vk::HitRec*>(base + vk::kNoAllocCandHitsOffset),
                         reinterpret_cast<unsigned long long*>(base + vk::kNoAllocCandKeysOffset)};
    return trace_launch(c, t, rays, count, hits, flags, pick(c, stream), s, false);
}

int vkhrt_render(vkhrt_context* c, const vkhrt_tlas* t, const vkhrt_camera* cam, const vkhrt_render_desc* d,
                 void* stream) {
    if (!c || !t || !cam || !d) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "null argument");
    if (!d->image) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "null image");
    if (d->width == 0 || d->height == 0 || uint64_t(d->width) * d->height > (1ull << 28))
        return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "bad image size");
    const uint32_t known = VKHRT_RENDER_SHADOWS | VKHRT_RENDER_DEBUG_PALETTE | VKHRT_RENDER_RGBA8 |
                           VKHRT_RENDER_FP64_POLISH | VKHRT_RENDER_NEURAL_SHADOWS | VKHRT_RENDER_ACCURATE_NORMALS;
    if (d->flags & ~known) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "unknown render flags");
    const size_t npix = size_t(d->width) * d->height;
    const vkhrt_area_light* al = d->light_visibility ? nullptr : d->area_light;
    const uint32_t S = al ? al->samples : 1u;
    if (al) {
        if (S < 1 || S > 16) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "area light samples must be 1..16");
        if (!c->persistent || (is_phantom(t) && !t->polish_ok))
            return fail(c, VKHRT_ERROR_UNSUPPORTED, "area lights need the deferred render pipeline (persistent option on)");
        if (npix * S > (size_t(1) << 31)) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "too many shadow rays");
    }
    // the largest work count of the camera's orders: whole 64 x 64 tiles
    const size_t nwork = size_t((d->width + 63u) / 64u) * 64u * size_t((d->height + 63u) / 64u) * 64u;
    int rc = ensure_render_scratch(c, npix * S, nwork);
    if (rc != VKHRT_SUCCESS) return rc;
    hipStream_t st = pick(c, stream);
    vk::RenderParams rp{};
    if (d->environment) {
        const vkhrt_environment_map* em = d->environment;
        if (!em->texels || em->width == 0 || em->height == 0 || !(em->exposure >= 0.0f))
            return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "bad environment map");
        rp.env = vk::EnvMap{reinterpret_cast<const float4*>(em->texels), em->width, em->height, em->exposure};
    }
    rp.light.samples = S;
    if (al) {
        rp.light.enabled = 1;
        std::memcpy(rp.light.c, al->center, 12);
        std::memcpy(rp.light.u, al->edge_u, 12);
        std::memcpy(rp.light.v, al->edge_v, 12);
        rp.light.seed = al->seed;
    }
    rp.tp = vk::TraceParams{t->dev, t->count, (d->flags & VKHRT_RENDER_FP64_POLISH) ? uint32_t(vk::kTraceFp64Polish) : 0u,
                            c->status, c->persistent && (t->polish_ok || !is_phantom(t)) ? c->counters + 1 : nullptr,
                            c->refill_min};
    rp.tp.blocks_per_cu = c->blocks_per_cu;
    rp.tp.work_counts = c->work_counts;
    std::memcpy(rp.cam.view_inv, cam->view_inverse, 64);
    std::memcpy(rp.cam.proj_inv, cam->proj_inverse, 64);
    rp.cam.width = d->width;
    rp.cam.height = d->height;
    set_work_order(rp.cam, c->ray_order);
    rp.shadows = (d->flags & VKHRT_RENDER_SHADOWS) ? 1u : 0u;
    rp.out_format = (d->flags & VKHRT_RENDER_RGBA8) ? 1u : 0u;
    rp.debug_palette = (d->flags & VKHRT_RENDER_DEBUG_PALETTE) ? 1u : 0u;
    rp.image = d->image;
    rp.prim_out = d->primitive_ids;
    rp.hitbuf = c->hitbuf;
    rp.queue = c->queue;
    rp.counters = c->counters;
    rp.tp.xcd_chunk = c->xcd_chunk;
    rp.tp.steps_per_iter = c->steps_per_iter;
    rp.light_vis = const_cast<float*>(d->light_visibility);
    rp.hitrec = c->hitrec;
    rp.vis = c->vis;
    rp.live = c->live;
    rp.live_mask = c->live_mask;
    if (d->flags & VKHRT_RENDER_FP64_POLISH) {
        if (!t->polish_ok)
            return fail(c, VKHRT_ERROR_UNSUPPORTED,
                        "render FP64 polish needs phantom BLASes of < 2^26 records and < 2^32 - 1 primitive ids per TLAS");
        rc = ensure_polish_scratch(c, npix * S);
        if (rc != VKHRT_SUCCESS) return rc;
        set_polish(c, rp.tp);
        if (!rp.tp.work) rp.tp.work = c->counters + 1;
    }
    if (d->flags & VKHRT_RENDER_ACCURATE_NORMALS) {
        if (!rp.tp.work)
            return fail(c, VKHRT_ERROR_UNSUPPORTED, "accurate normals need the deferred render pipeline (persistent option on)");
        rp.accurate = 1;
    }
    const bool neural = (d->flags & VKHRT_RENDER_NEURAL_SHADOWS) && rp.shadows && !rp.debug_palette &&
                        !d->light_visibility;
    if (neural) {
        if (!d->shadow_mlp) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "neural shadows need shadow_mlp");
        if (d->shadow_mlp->ctx != c) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "shadow_mlp of another context");
        if (!rp.tp.work)
            return fail(c, VKHRT_ERROR_UNSUPPORTED, "neural shadows need the deferred render pipeline (persistent option on)");
        if (!(d->shadow_variance_threshold >= 0.0f))
            return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "variance threshold must be >= 0");
        rc = ensure_neural_scratch(c, npix * S);
        if (rc != VKHRT_SUCCESS) return rc;
        rp.mlp = &d->shadow_mlp->dev;
        if (al) rp.dirs = c->dirs;
        rp.mlp_threshold = d->shadow_variance_threshold;
        rp.wave_limit = c->shadow_wave_limit;
        rp.group_lanes = c->shadow_group_lanes;
        rp.origins = c->origins;
        rp.fallback_queue = c->fallback;
    }
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    if (c->profiling)
        for (auto& x : ev) hipEventCreate(&x);
    hipError_t e = vk::launch_render(rp, trace_variant(t, (d->flags & VKHRT_RENDER_FP64_POLISH) != 0),
                                     stack_kind_for(c, t), st, c->profiling ? ev : nullptr);
    if (c->profiling) {
        c->pending.push_back(PendingTiming{"render_primary", ev[0], ev[1]});
        c->pending.push_back(PendingTiming{"render_shadow", ev[2], ev[3]});
    }
    if (e != hipSuccess) return hip_fail(c, e, "render launch");
    c->last_pixels = npix;
    c->last_shadows = rp.shadows && !rp.debug_palette && !d->light_visibility;
    c->last_neural = neural;
    c->last_stream = st;
    return VKHRT_SUCCESS;
}

int vkhrt_generate_surface_shadow_rays(vkhrt_context* c, const vkhrt_tlas* t, const vkhrt_area_light* light,
                                       uint32_t count, uint64_t seed, vkhrt_ray* rays, void* stream) {
    if (!c || !t || (count && !rays)) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "null argument");
    if (t->count == 0 || t->count > vk::kSurfaceMaxInstances)
        return fail(c, VKHRT_ERROR_UNSUPPORTED, "surface sampling needs 1..64 instances");
    vk::SurfaceParams sp{};
    sp.inst = t->dev;
    sp.ninst = t->count;
    uint64_t total = 0;
    for (uint32_t i = 0; i < t->count; ++i) {
        if (t->modes[i] != VKHRT_GEOMETRY_PHANTOM_CURVES)
            return fail(c, VKHRT_ERROR_UNSUPPORTED, "surface sampling needs phantom-curve instances");
        sp.prefix[i] = uint32_t(total);
        total += t->records[i];
    }
    if (total == 0 || total >= (1ull << 32)) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "no curve records");
    sp.prefix[t->count] = uint32_t(total);
    sp.light.samples = 1;
    sp.light.seed = seed;
    if (light) {
        sp.light.enabled = 1;
        std::memcpy(sp.light.c, light->center, 12);
        std::memcpy(sp.light.u, light->edge_u, 12);
        std::memcpy(sp.light.v, light->edge_v, 12);
    }
    sp.seed = seed;
    const hipError_t e = vk::launch_surface_shadow_rays(sp, count, reinterpret_cast<float*>(rays), pick(c, stream));
    return e == hipSuccess ? VKHRT_SUCCESS : hip_fail(c, e, "surface shadow rays");
}

int vkhrt_render_stats_get(vkhrt_context* c, vkhrt_render_stats* out) {
    if (!c || !out) return VKHRT_ERROR_INVALID_ARGUMENT;
    *out = vkhrt_render_stats{c->last_pixels, 0, 0};
    if (c->last_pixels && c->last_shadows) {
        uint32_t q[6] = {0, 0, 0, 0, 0, 0};
        hipError_t e = hipStreamSynchronize(c->last_stream);
        if (e == hipSuccess) e = hipMemcpy(q, c->counters, sizeof(q), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(c, e, "render stats");
        out->shadow_queries = q[0];
        out->shadow_rays_traced = c->last_neural ? q[5] : q[0];
    }
    return VKHRT_SUCCESS;
}

int vkhrt_render_batch(vkhrt_context* c, const vkhrt_tlas* t, const vkhrt_view* views, uint32_t view_count,
                       const vkhrt_batch_desc* d, void* stream) {
    if (!c || !t || !d || !views) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "null argument");
    if (view_count == 0 || view_count > 65536) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "view count must be 1..65536");
    const uint32_t known = VKHRT_BATCH_RGBA8 | VKHRT_BATCH_FP64_POLISH | VKHRT_BATCH_SORT_MATERIALS |
                           VKHRT_BATCH_SORT_RAYS | VKHRT_BATCH_GRAPH;
    if (d->flags & ~known) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "unknown batch flags");
    if (d->bounces > VKHRT_BATCH_MAX_BOUNCES) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "bounces above the maximum");
    if (d->material_count > VKHRT_BATCH_MAX_MATERIALS || (d->material_count && !d->material_albedo))
        return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "bad material table");
    if ((is_phantom(t) || (has_phantom(t) && (d->flags & VKHRT_BATCH_FP64_POLISH))) && !t->polish_ok)
        return fail(c, VKHRT_ERROR_UNSUPPORTED,
                    "batched phantom tracing needs phantom BLASes of < 2^26 records and < 2^32 - 1 primitive ids per TLAS");
    uint64_t total = 0;
    for (uint32_t i = 0; i < view_count; ++i) {
        if (!views[i].image) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "null view image");
        if (views[i].width == 0 || views[i].height == 0)
            return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "empty view");
        total += uint64_t(views[i].width) * views[i].height;
        if (total > (1ull << 31)) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "more than 2^31 pixels in a batch");
    }
    int rc = ensure_batch_scratch(c, total);
    if (rc == VKHRT_SUCCESS) rc = ensure_batch_tables(c, view_count);
    if (rc != VKHRT_SUCCESS) return rc;
    hipStream_t st = pick(c, stream);

    // view table and frame values, staged in pinned memory once the previous upload is done
    hipError_t e = hipEventSynchronize(c->b_upload);
    if (e != hipSuccess) return hip_fail(c, e, "batch staging");
    auto* hv = reinterpret_cast<vk::BatchView*>(c->b_host);
    uint32_t first = 0;
    for (uint32_t i = 0; i < view_count; ++i) {
        vk::BatchView& v = hv[i];
        std::memset(&v, 0, sizeof(v));
        std::memcpy(v.cam.view_inv, views[i].camera.view_inverse, 64);
        std::memcpy(v.cam.proj_inv, views[i].camera.proj_inverse, 64);
        v.cam.width = views[i].width;
        v.cam.height = views[i].height;
        set_work_order(v.cam, 0);
        v.first = first;
        v.image = views[i].image;
        v.prim_out = views[i].primitive_ids;
        first += views[i].width * views[i].height;
    }
    auto* hf = reinterpret_cast<vk::BatchFrame*>(c->b_host + size_t(view_count) * sizeof(vk::BatchView));
    std::memset(hf, 0, sizeof(*hf));
    hf->seed = d->seed;
    for (int k = 0; k < 3; ++k) hf->env[k] = d->environment[k];
    hf->nviews = view_count;
    hf->out_format = (d->flags & VKHRT_BATCH_RGBA8) ? 1u : 0u;
    for (uint32_t m = 0; m < vk::kBatchMaxMaterials; ++m)
        for (int k = 0; k < 3; ++k) hf->albedo[m][k] = m < d->material_count ? d->material_albedo[m * 3 + k] : 1.0f;
    e = hipMemcpyAsync(c->b_views, hv, size_t(view_count) * sizeof(vk::BatchView), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(c->b_frame, hf, sizeof(*hf), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipEventRecord(c->b_upload, st);
    if (e != hipSuccess) return hip_fail(c, e, "batch upload");

    const bool polish = (d->flags & VKHRT_BATCH_FP64_POLISH) && has_phantom(t);
    vk::BatchParams bp{};
    bp.tp = vk::TraceParams{t->dev, t->count, polish ? uint32_t(vk::kTraceFp64Polish) : 0u, c->status,
                            c->b_counters, c->refill_min};
    bp.tp.blocks_per_cu = c->blocks_per_cu;
    bp.tp.xcd_chunk = c->xcd_chunk;
    bp.tp.steps_per_iter = c->steps_per_iter;
    if (polish) {
        bp.tp.defer = c->b_defer;
        bp.tp.defer_cap = uint32_t(std::max<size_t>(c->b_cap, 65536));
        bp.tp.keys = c->b_keys;
        bp.tp.cand_hits = c->b_cand_hits;
        bp.tp.cand_keys = c->b_cand_keys;
    }
    bp.views = c->b_views;
    bp.frame = c->b_frame;
    bp.total = uint32_t(total);
    bp.bounces = d->bounces;
    bp.sort = ((d->flags & VKHRT_BATCH_SORT_MATERIALS) ? uint32_t(vk::kBatchSortMaterial) : 0u) |
              ((d->flags & VKHRT_BATCH_SORT_RAYS) ? uint32_t(vk::kBatchSortRays) : 0u);
    bp.qsort = c->b_qsort;
    bp.bins = c->b_bins;
    // cell grid of the ray sort: 9 bits over the scene box, split toward the longest cell edge
    for (int k = 0; k < 3; ++k) {
        bp.scene_lo[k] = t->lo[k];
        bp.cell_bits[k] = 0;
    }
    for (int i = 0; i < 9; ++i) {
        int best = 0;
        float edge = -1.0f;
        for (int k = 0; k < 3; ++k) {
            const float e = std::ldexp(std::max(t->hi[k] - t->lo[k], 0.0f), -int(bp.cell_bits[k]));
            if (e > edge) {
                edge = e;
                best = k;
            }
        }
        ++bp.cell_bits[best];
    }
    for (int k = 0; k < 3; ++k) {
        const float ext = t->hi[k] - t->lo[k];
        bp.cell_scale[k] = ext > 0.0f ? std::ldexp(1.0f, int(bp.cell_bits[k])) / ext : 0.0f;
    }
    bp.ray = c->b_ray;
    bp.thr = c->b_thr;
    bp.rad = c->b_rad;
    bp.nee = c->b_nee;
    bp.hits = c->b_hits;
    bp.q[0] = c->b_q0;
    bp.q[1] = c->b_q1;
    bp.order = c->b_order;
    bp.squeue = c->b_squeue;
    bp.counters = c->b_counters;
    const int variant = trace_variant(t, polish), stack = stack_kind_for(c, t);

    hipEvent_t ev[2] = {nullptr, nullptr};
    if (c->profiling) {
        for (auto& x : ev) hipEventCreate(&x);
        hipEventRecord(ev[0], st);
    }
    if (d->flags & VKHRT_BATCH_GRAPH) {
        const std::vector<uint64_t> key = {t->serial, uint64_t(variant), uint64_t(stack), total,         d->bounces,
                                           bp.sort,   polish,            c->refill_min,   c->blocks_per_cu, c->batch_gen,
                                           c->xcd_chunk, c->steps_per_iter};
        if (!c->graph_exec || c->graph_key != key) {
            if (c->graph_exec) hipGraphExecDestroy(c->graph_exec);
            if (c->graph) hipGraphDestroy(c->graph);
            c->graph_exec = nullptr;
            c->graph = nullptr;
            c->graph_key.clear();
            if (!c->capture_stream) e = hipStreamCreateWithFlags(&c->capture_stream, hipStreamNonBlocking);
            if (e == hipSuccess) e = hipStreamBeginCapture(c->capture_stream, hipStreamCaptureModeRelaxed);
            if (e != hipSuccess) return hip_fail(c, e, "batch graph capture");
            const hipError_t le = vk::launch_batch(bp, variant, stack, c->capture_stream);
            hipGraph_t g = nullptr;
            e = hipStreamEndCapture(c->capture_stream, &g);
            if (le != hipSuccess || e != hipSuccess) {
                if (g) hipGraphDestroy(g);
                return hip_fail(c, le != hipSuccess ? le : e, "batch graph capture");
            }
            e = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                hipGraphDestroy(g);
                c->graph_exec = nullptr;
                return hip_fail(c, e, "batch graph instantiate");
            }
            c->graph = g;
            c->graph_key = key;
            ++c->graph_captures;
        }
        e = hipGraphLaunch(c->graph_exec, st);
    } else {
        e = vk::launch_batch(bp, variant, stack, st);
    }
    if (c->profiling) {
        hipEventRecord(ev[1], st);
        c->pending.push_back(PendingTiming{"render_batch", ev[0], ev[1]});
    }
    if (e != hipSuccess) return hip_fail(c, e, "batch launch");
    c->last_batch_paths = total;
    c->last_batch_bounces = d->bounces;
    c->last_batch_stream = st;
    return VKHRT_SUCCESS;
}

int vkhrt_batch_stats_get(vkhrt_context* c, vkhrt_batch_stats* out) {
    if (!c || !out) return VKHRT_ERROR_INVALID_ARGUMENT;
    *out = vkhrt_batch_stats{c->last_batch_paths, 0, 0, c->graph_captures, 0};
    if (!c->last_batch_paths) return VKHRT_SUCCESS;
    uint32_t w[vk::kBatchDepthWords * (vk::kBatchMaxBounces + 2)];
    hipError_t e = hipStreamSynchronize(c->last_batch_stream);
    if (e == hipSuccess) e = hipMemcpy(w, c->b_counters, sizeof(w), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(c, e, "batch stats");
    for (uint32_t b = 0; b <= c->last_batch_bounces; ++b) {
        if (b > 0) out->bounce_rays += w[b * vk::kBatchDepthWords + vk::kBcQueue];
        out->shadow_rays += w[b * vk::kBatchDepthWords + vk::kBcShadow];
    }
    return VKHRT_SUCCESS;
}

}  // extern "C"

namespace {

// Median time (ms) of `repeats` runs of `work` on `st` after one warm-up run.
template <class F>
int time_workload(vkhrt_context* c, hipStream_t st, uint32_t repeats, F&& work, float& ms) {
    int rc = work();
    if (rc != VKHRT_SUCCESS) return rc;
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    std::vector<float> t;
    for (uint32_t i = 0; i < repeats && e == hipSuccess && rc == VKHRT_SUCCESS; ++i) {
        e = hipEventRecord(a, st);
        rc = work();
        if (e == hipSuccess) e = hipEventRecord(b, st);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        float x = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&x, a, b);
        t.push_back(x);
    }
    if (a) hipEventDestroy(a);
    if (b) hipEventDestroy(b);
    if (rc != VKHRT_SUCCESS) return rc;
    if (e != hipSuccess) return hip_fail(c, e, "autotune timing");
    std::nth_element(t.begin(), t.begin() + t.size() / 2, t.end());
    ms = t[t.size() / 2];
    return VKHRT_SUCCESS;
}

struct Knob {
    int* value;
    std::vector<int> candidates;
};

// Coordinate descent: two sweeps over the knobs, each value tried with the others fixed;
// `apply` pushes the knob values into the context before a trial. On a failed trial the
// starting values are restored.
template <class Apply, class F>
int coordinate_descent(vkhrt_context* c, hipStream_t st, uint32_t repeats, std::vector<Knob>& knobs, Apply&& apply,
                       F&& work, vkhrt_tuning* out) {
    std::vector<int> start;
    for (const Knob& k : knobs) start.push_back(*k.value);
    auto restore = [&] {
        for (size_t i = 0; i < knobs.size(); ++i) *knobs[i].value = start[i];
        apply();
    };
    apply();
    float best = 0.0f;
    int rc = time_workload(c, st, repeats, work, best);
    if (rc != VKHRT_SUCCESS) {
        restore();
        return rc;
    }
    out->baseline_ms = best;
    out->trials = 1;
    for (int pass = 0; pass < 2; ++pass) {
        for (Knob& k : knobs) {
            const int keep = *k.value;
            int chosen = keep;
            for (int v : k.candidates) {
                if (v == keep) continue;
                *k.value = v;
                apply();
                float ms = 0.0f;
                rc = time_workload(c, st, repeats, work, ms);
                if (rc != VKHRT_SUCCESS) {
                    restore();
                    return rc;
                }
                ++out->trials;
                if (ms < best) {
                    best = ms;
                    chosen = v;
                }
            }
            *k.value = chosen;
            apply();
        }
    }
    out->best_ms = best;
    return VKHRT_SUCCESS;
}

void fill_tuning(const vkhrt_context* c, vkhrt_tuning* out) {
    out->stack_kind = c->stack_kind;
    out->blocks_per_cu = int32_t(c->blocks_per_cu);
    out->refill_min = int32_t(c->refill_min);
    out->ray_order = int32_t(c->ray_order);
    out->xcd_chunk = int32_t(c->xcd_chunk);
    out->bvh_width = c->bvh_width;
    out->bvh_collapse = c->bvh_collapse;
    out->leaf_size = c->leaf_size;
    out->steps_per_iteration = int32_t(c->steps_per_iter);
}

// Stack kinds that can hold the TLAS's deepest tree (smaller ones are raised to the same).
std::vector<int> stack_candidates(const vkhrt_tlas* t) {
    std::vector<int> v;
    for (int k = 0; k < 4; ++k)
        if (kStackEntries[k] >= t->max_depth || k == 3) v.push_back(k);
    return v;
}

}  // namespace

extern "C" {

int vkhrt_autotune_render(vkhrt_context* c, const vkhrt_tlas* t, const vkhrt_camera* cam, const vkhrt_render_desc* d,
                          uint32_t repeats, vkhrt_tuning* out, void* stream) {
    if (!c || !t || !cam || !d || !out) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "null argument");
    if (repeats < 1 || repeats > 100) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "repeats must be 1..100");
    *out = vkhrt_tuning{};
    hipStream_t st = pick(c, stream);
    int stack = c->stack_kind, bpc = int(c->blocks_per_cu), refill = int(c->refill_min), order = int(c->ray_order);
    int chunk = int(c->xcd_chunk), steps = int(c->steps_per_iter);
    std::vector<Knob> knobs = {{&stack, stack_candidates(t)},
                               {&bpc, {0, 6, 4, 3, 2}},
                               {&refill, {4, 8, 16, 32, 48}},
                               {&order, {0, 1, 2}},
                               {&chunk, {0, 512, 4096, 16384}},
                               {&steps, {1, 2, 3, 4, 6}}};
    auto apply = [&] {
        c->stack_kind = stack;
        c->blocks_per_cu = uint32_t(bpc);
        c->refill_min = uint32_t(refill);
        c->ray_order = uint32_t(order);
        c->xcd_chunk = uint32_t(chunk);
        c->steps_per_iter = uint32_t(steps);
    };
    const int rc = coordinate_descent(c, st, repeats, knobs, apply, [&] { return vkhrt_render(c, t, cam, d, stream); },
                                      out);
    fill_tuning(c, out);
    return rc;
}

int vkhrt_autotune_batch(vkhrt_context* c, const vkhrt_tlas* t, const vkhrt_view* views, uint32_t view_count,
                         const vkhrt_batch_desc* d, uint32_t repeats, vkhrt_tuning* out, void* stream) {
    if (!c || !t || !views || !d || !out) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "null argument");
    if (repeats < 1 || repeats > 100) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "repeats must be 1..100");
    *out = vkhrt_tuning{};
    hipStream_t st = pick(c, stream);
    const uint32_t sort_bits = VKHRT_BATCH_SORT_MATERIALS | VKHRT_BATCH_SORT_RAYS;
    vkhrt_batch_desc dd = *d;
    int stack = c->stack_kind, bpc = int(c->blocks_per_cu), refill = int(c->refill_min);
    int sort = int(d->flags & sort_bits), chunk = int(c->xcd_chunk), steps = int(c->steps_per_iter);
    std::vector<Knob> knobs = {{&stack, stack_candidates(t)},
                               {&bpc, {0, 6, 4, 3, 2}},
                               {&refill, {4, 8, 16, 32, 48}},
                               {&sort, {0, int(VKHRT_BATCH_SORT_RAYS), int(VKHRT_BATCH_SORT_MATERIALS), int(sort_bits)}},
                               {&chunk, {0, 512, 4096, 16384}},
                               {&steps, {1, 2, 3, 4, 6}}};
    auto apply = [&] {
        c->stack_kind = stack;
        c->blocks_per_cu = uint32_t(bpc);
        c->refill_min = uint32_t(refill);
        c->xcd_chunk = uint32_t(chunk);
        c->steps_per_iter = uint32_t(steps);
        dd.flags = (d->flags & ~sort_bits) | uint32_t(sort);
    };
    const int rc = coordinate_descent(
        c, st, repeats, knobs, apply, [&] { return vkhrt_render_batch(c, t, views, view_count, &dd, stream); }, out);
    fill_tuning(c, out);
    out->batch_flags = uint32_t(sort);
    return rc;
}

int vkhrt_autotune_blas(vkhrt_context* c, const vkhrt_blas_desc* bd, const vkhrt_camera* cam,
                        const vkhrt_render_desc* d, uint32_t repeats, vkhrt_tuning* out, void* stream) {
    if (!c || !bd || !cam || !d || !out) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "null argument");
    if (repeats < 1 || repeats > 100) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "repeats must be 1..100");
    *out = vkhrt_tuning{};
    hipStream_t st = pick(c, stream);
    const int width0 = c->bvh_width, leaf0 = c->leaf_size, coll0 = c->bvh_collapse;
    // one trial: build with (collapse, width, leaf), render, time
    auto trial = [&](int coll, int width, int leaf, float& ms) {
        c->bvh_collapse = coll;
        c->bvh_width = width;
        vkhrt_blas_desc b = *bd;
        b.leaf_size = uint32_t(leaf);
        vkhrt_blas* blas = nullptr;
        vkhrt_tlas* tlas = nullptr;
        int rc = vkhrt_blas_build(c, &b, stream, &blas);
        if (rc == VKHRT_SUCCESS) rc = vkhrt_tlas_build(c, &blas, 1, stream, &tlas);
        if (rc == VKHRT_SUCCESS)
            rc = time_workload(c, st, repeats, [&] { return vkhrt_render(c, tlas, cam, d, stream); }, ms);
        if (rc == VKHRT_SUCCESS) rc = vkhrt_synchronize(c, stream);
        if (tlas) vkhrt_tlas_destroy(tlas);
        if (blas) vkhrt_blas_destroy(blas);
        return rc;
    };
    float best = 0.0f;
    int rc = trial(coll0, width0, leaf0, best);
    if (rc != VKHRT_SUCCESS) {
        c->bvh_width = width0;
        c->bvh_collapse = coll0;
        return rc;
    }
    out->baseline_ms = best;
    out->trials = 1;
    int bw = width0, bl = leaf0, bc = coll0;
    for (int coll : {0, 2, 4})
        for (int width : {4, 6, 8})
            for (int leaf : {1, 2, 3}) {
                if (coll == coll0 && width == width0 && leaf == leaf0) continue;
                float ms = 0.0f;
                rc = trial(coll, width, leaf, ms);
                if (rc != VKHRT_SUCCESS) {
                    c->bvh_width = width0;
                    c->leaf_size = leaf0;
                    c->bvh_collapse = coll0;
                    return rc;
                }
                ++out->trials;
                if (ms < best) {
                    best = ms;
                    bw = width;
                    bl = leaf;
                    bc = coll;
                }
            }
    c->bvh_width = bw;
    c->leaf_size = bl;
    c->bvh_collapse = bc;
    out->best_ms = best;
    fill_tuning(c, out);
    return VKHRT_SUCCESS;
}

int vkhrt_trace_stats(vkhrt_context* c, const vkhrt_tlas* t, const vkhrt_ray* rays, uint32_t count, vkhrt_hit* hits,
                      uint32_t flags, uint32_t* stats, void* stream) {
    if (!c || !t || (count && (!rays || !hits || !stats))) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "bad argument");
    if (flags & ~VKHRT_TRACE_TERMINATE_ON_FIRST_HIT)
        return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "trace_stats supports only the any-hit flag");
    if (t->variant == vk::kVarMixed) return fail(c, VKHRT_ERROR_UNSUPPORTED, "trace_stats needs a single-mode TLAS");
    vk::TraceParams tp{t->dev, t->count, flags, c->status, nullptr, 1u};
    hipError_t e = vk::launch_trace_rays(tp, reinterpret_cast<const float*>(rays), count,
                                         reinterpret_cast<vk::HitRec*>(hits), reinterpret_cast<uint2*>(stats),
                                         t->variant, stack_kind_for(c, t), pick(c, stream));
    return e == hipSuccess ? VKHRT_SUCCESS : hip_fail(c, e, "trace stats");
}

int vkhrt_render_ray_counts(vkhrt_context* c, uint64_t* primary, uint64_t* shadow) {
    if (!c || !primary || !shadow) return VKHRT_ERROR_INVALID_ARGUMENT;
    *primary = c->last_pixels;
    *shadow = 0;
    if (c->last_pixels && c->last_shadows) {
        uint32_t q = 0;
        hipError_t e = hipStreamSynchronize(c->last_stream);
        if (e == hipSuccess) e = hipMemcpy(&q, c->counters + (c->last_neural ? 5 : 0), 4, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(c, e, "render counts");
        *shadow = q;
    }
    return VKHRT_SUCCESS;
}

int vkhrt_generate_camera_rays(vkhrt_context* c, const vkhrt_camera* cam, uint32_t width, uint32_t height, float tmin,
                               float tmax, vkhrt_ray* rays, void* stream) {
    if (!c || !cam || (uint64_t(width) * height && !rays)) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "bad argument");
    vk::Camera cm{};
    std::memcpy(cm.view_inv, cam->view_inverse, 64);
    std::memcpy(cm.proj_inv, cam->proj_inverse, 64);
    cm.width = width;
    cm.height = height;
    set_work_order(cm, 0);
    hipError_t e = vk::launch_camera_rays(cm, tmin, tmax, reinterpret_cast<float*>(rays), pick(c, stream));
    return e == hipSuccess ? VKHRT_SUCCESS : hip_fail(c, e, "camera rays");
}

int vkhrt_generate_shadow_rays(vkhrt_context* c, const vkhrt_ray* rays, const vkhrt_hit* hits, uint32_t count,
                               const float direction[3], float tmin, float tmax, vkhrt_ray* out, uint32_t* source,
                               uint32_t* count_out, void* stream) {
    if (!c || !direction || !count_out || (count && (!rays || !hits || !out)))
        return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "bad argument");
    hipStream_t st = pick(c, stream);
    hipError_t e = vk::launch_shadow_rays(reinterpret_cast<const float*>(rays), reinterpret_cast<const vk::HitRec*>(hits), count,
                                   direction, tmin, tmax, reinterpret_cast<float*>(out), source, count_out, st);
    return e == hipSuccess ? VKHRT_SUCCESS : hip_fail(c, e, "shadow rays");
}

int vkhrt_count_candidates(vkhrt_context* c, const vkhrt_tlas* t, const vkhrt_ray* rays, const vkhrt_hit* hits,
                           uint32_t count, uint32_t* counts, void* stream) {
    if (!c || !t || (count && (!rays || !hits || !counts))) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "bad argument");
    vk::TraceParams tp{t->dev, t->count, 0u, c->status, nullptr, 1u};
    hipError_t e = vk::launch_count_candidates(tp, reinterpret_cast<const float*>(rays),
                                               reinterpret_cast<const vk::HitRec*>(hits), count, counts, pick(c, stream));
    return e == hipSuccess ? VKHRT_SUCCESS : hip_fail(c, e, "count candidates");
}

int vkhrt_camera_look_at(const float eye[3], const float center[3], const float up[3], float tan_half_fovy,
                         float aspect, float znear, float zfar, vkhrt_camera* out) {
    if (!eye || !center || !up || !out) return VKHRT_ERROR_INVALID_ARGUMENT;
    if (!(tan_half_fovy > 0.0f) || !(aspect > 0.0f) || !(zfar != znear)) return VKHRT_ERROR_INVALID_ARGUMENT;
    // glm::lookAtRH
    H3 E{eye[0], eye[1], eye[2]}, C{center[0], center[1], center[2]}, U{up[0], up[1], up[2]};
    H3 f = h_norm(h_sub(C, E));
    H3 s = h_norm(h_cross(f, U));
    H3 u = h_cross(s, f);
    float V[16] = {0};
    V[0] = s.x;
    V[4] = s.y;
    V[8] = s.z;
    V[1] = u.x;
    V[5] = u.y;
    V[9] = u.z;
    V[2] = -f.x;
    V[6] = -f.y;
    V[10] = -f.z;
    V[12] = -h_dot(s, E);
    V[13] = -h_dot(u, E);
    V[14] = h_dot(f, E);
    V[15] = 1.0f;
    // glm::perspectiveRH_ZO, then projection[1][1] *= -1 (fly_camera.cpp:32-33)
    float P[16] = {0};
    P[0] = 1.0f / (aspect * tan_half_fovy);
    P[5] = 1.0f / tan_half_fovy;
    P[10] = zfar / (znear - zfar);
    P[11] = -1.0f;
    P[14] = -(zfar * znear) / (zfar - znear);
    P[5] *= -1.0f;
    inverse4(V, out->view_inverse);
    inverse4(P, out->proj_inverse);
    return VKHRT_SUCCESS;
}

// ---- geometry front-end
int vkhrt_generate_strands(vkhrt_context* c, uint64_t seed, uint32_t nstrands, uint32_t nseg,
                           const vkhrt_strand_params* p, float* vertices, void* stream) {
    if (!c || (nstrands && !vertices) || nseg == 0) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "bad argument");
    vk::StrandParams hp{100.0f, 1.5f, 0.3f, 0.2f, 0.04f, 0.01f};
    if (p) hp = vk::StrandParams{p->sphere_radius, p->step, p->gravity, p->curl, p->r_root, p->r_tip};
    hipError_t e = vk::gen_strands(seed, nstrands, nseg, hp, reinterpret_cast<float4*>(vertices), pick(c, stream));
    return e == hipSuccess ? VKHRT_SUCCESS : hip_fail(c, e, "generate strands");
}
int vkhrt_generate_lines(vkhrt_context* c, const float* v, uint32_t nstrands, uint32_t nseg, float* lines,
                         void* stream) {
    if (!c || (nstrands && (!v || !lines))) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "bad argument");
    hipError_t e = vk::gen_strand_lines(reinterpret_cast<const float4*>(v), nstrands, nseg,
                                        reinterpret_cast<float4*>(lines), pick(c, stream));
    return e == hipSuccess ? VKHRT_SUCCESS : hip_fail(c, e, "generate lines");
}
int vkhrt_generate_curves(vkhrt_context* c, const float* lines, uint32_t n, float tension, float* curves,
                          void* stream) {
    if (!c || (n && (!lines || !curves))) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "bad argument");
    hipError_t e = vk::gen_curves(reinterpret_cast<const float4*>(lines), n, tension, reinterpret_cast<float4*>(curves),
                                  pick(c, stream));
    return e == hipSuccess ? VKHRT_SUCCESS : hip_fail(c, e, "generate curves");
}
int vkhrt_curves_to_lines(vkhrt_context* c, const float* curves, uint32_t n, uint32_t pieces, float* lines,
                          void* stream) {
    if (!c || pieces == 0 || (n && (!curves || !lines))) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "bad argument");
    hipError_t e = vk::gen_curve_lines(reinterpret_cast<const float4*>(curves), n, pieces,
                                       reinterpret_cast<float4*>(lines), pick(c, stream));
    return e == hipSuccess ? VKHRT_SUCCESS : hip_fail(c, e, "curves to lines");
}
int vkhrt_generate_lss(vkhrt_context* c, const float* lines, uint32_t n, float constant_radius, float* caps,
                       void* stream) {
    if (!c || (n && (!lines || !caps))) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "bad argument");
    hipError_t e = vk::gen_lss(reinterpret_cast<const float4*>(lines), n, constant_radius,
                               reinterpret_cast<float4*>(caps), pick(c, stream));
    return e == hipSuccess ? VKHRT_SUCCESS : hip_fail(c, e, "generate lss");
}
int vkhrt_generate_dots(vkhrt_context* c, const float* lines, uint32_t n, float constant_radius, float* pieces,
                        void* stream) {
    if (!c || (n && (!lines || !pieces))) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "bad argument");
    hipError_t e = vk::gen_dots(reinterpret_cast<const float4*>(lines), n, constant_radius,
                                reinterpret_cast<float4*>(pieces), pick(c, stream));
    return e == hipSuccess ? VKHRT_SUCCESS : hip_fail(c, e, "generate dots");
}
int vkhrt_generate_mesh_lines(vkhrt_context* c, const float* v, uint32_t nv, const uint32_t* idx, uint32_t n,
                              float radius, float* lines, void* stream) {
    if (!c || (n && (!v || !idx || !lines))) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "bad argument");
    if (n == 0) return VKHRT_SUCCESS;
    if (nv == 0) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "line indices given without vertices");
    hipSetDevice(c->device);
    hipStream_t st = pick(c, stream);
    uint32_t* bad = nullptr;
    uint32_t nbad = 0;
    hipError_t e = hipMalloc(&bad, sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemsetAsync(bad, 0, sizeof(uint32_t), st);
    if (e == hipSuccess) e = vk::gen_mesh_lines(v, nv, idx, n, radius, reinterpret_cast<float4*>(lines), bad, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&nbad, bad, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    hipFree(bad);
    if (e != hipSuccess) return hip_fail(c, e, "generate mesh lines");
    if (nbad) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, std::to_string(nbad) + " line indices are not below the vertex count");
    return VKHRT_SUCCESS;
}

// ---- assets
int vkhrt_asset_build(vkhrt_context* c, const vkhrt_asset* a, const vkhrt_asset_build_desc* d, void* stream,
                      vkhrt_blas** blases, uint32_t capacity, uint32_t* count) {
    if (!c || !a || !d) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "null argument");
    const uint32_t n = uint32_t(a->instances.size());
    if (count) *count = n;
    if (!blases) return capacity == 0 ? VKHRT_SUCCESS : fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "null blases");
    if (capacity < n)
        return fail(c, VKHRT_ERROR_INVALID_ARGUMENT,
                    "capacity " + std::to_string(capacity) + " is below the asset's " + std::to_string(n) + " instances");
    if (d->hair_mode > VKHRT_GEOMETRY_VOXELS || d->hair_mode == VKHRT_GEOMETRY_TRIANGLES)
        return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "hair mode must be phantom curves, LSS, DOTS or voxels");
    if (d->voxel_flags & ~uint32_t(VKHRT_BLAS_VOXEL_REFERENCE_SWEEP))
        return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "unknown voxel flags");
    if (!(d->radius >= 0.0f) || !std::isfinite(d->radius))
        return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "radius must be finite and >= 0");
    if (!std::isfinite(d->tension)) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "tension must be finite");
    if (d->leaf_size > 3) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "leaf size must be 0..3");
    for (uint32_t i = 0; i < n; ++i) blases[i] = nullptr;
    hipSetDevice(c->device);
    hipStream_t st = pick(c, stream);
    // device records per mesh, generated at its first instance and shared by the others
    struct Gen {
        float* verts = nullptr;
        uint32_t* idx = nullptr;
        float* lines = nullptr;
        float* recs = nullptr;
        uint32_t count = 0;
        bool done = false;
    };
    std::vector<Gen> gen(a->meshes.size());
    int rc = VKHRT_SUCCESS;
    for (uint32_t i = 0; i < n && rc == VKHRT_SUCCESS; ++i) {
        const vk::AssetInstance& I = a->instances[i];
        const vk::AssetMesh& M = a->meshes[I.mesh];
        const bool hair = M.primitive == VKHRT_PRIMITIVE_LINES;
        const uint32_t nv = uint32_t(M.vertices.size() / 8), ni = uint32_t(M.indices.size());
        Gen& g = gen[I.mesh];
        if (!g.done) {
            g.done = true;
            g.count = hair ? ni / 2 : ni / 3;
            if (g.count) {
                hipError_t e = hipMalloc(&g.verts, M.vertices.size() * sizeof(float));
                if (e == hipSuccess)
                    e = hipMemcpyAsync(g.verts, M.vertices.data(), M.vertices.size() * sizeof(float),
                                       hipMemcpyHostToDevice, st);
                if (e == hipSuccess) e = hipMalloc(&g.idx, size_t(ni) * sizeof(uint32_t));
                if (e == hipSuccess)
                    e = hipMemcpyAsync(g.idx, M.indices.data(), size_t(ni) * sizeof(uint32_t), hipMemcpyHostToDevice, st);
                if (e == hipSuccess && hair) e = hipMalloc(&g.lines, size_t(g.count) * 8 * sizeof(float));
                if (e != hipSuccess) {
                    rc = hip_fail(c, e, "asset upload");
                    break;
                }
                if (!hair) {
                    g.recs = g.verts;
                } else {
                    rc = vkhrt_generate_mesh_lines(c, g.verts, nv, g.idx, g.count, d->radius, g.lines, stream);
                    if (rc == VKHRT_SUCCESS &&
                        (d->hair_mode == VKHRT_GEOMETRY_DOTS || d->hair_mode == VKHRT_GEOMETRY_VOXELS)) {
                        g.recs = g.lines;
                    } else if (rc == VKHRT_SUCCESS) {
                        const bool curves = d->hair_mode == VKHRT_GEOMETRY_PHANTOM_CURVES;
                        e = hipMalloc(&g.recs, size_t(g.count) * (curves ? 16 : 8) * sizeof(float));
                        if (e != hipSuccess) rc = hip_fail(c, e, "asset records");
                        else if (curves) rc = vkhrt_generate_curves(c, g.lines, g.count, d->tension, g.recs, stream);
                        else rc = vkhrt_generate_lss(c, g.lines, g.count, d->radius, g.recs, stream);
                    }
                }
                if (rc != VKHRT_SUCCESS) break;
            }
        }
        vkhrt_blas_desc bd{};
        bd.mode = hair ? d->hair_mode : uint32_t(VKHRT_GEOMETRY_TRIANGLES);
        bd.flags = VKHRT_BLAS_HAS_TRANSFORM;
        bd.primitives = g.recs;
        bd.count = g.count;
        bd.radius = d->radius;
        bd.material_index = M.material >= 0 ? uint32_t(M.material) : d->default_material;
        bd.leaf_size = d->leaf_size;
        const float* W = a->nodes[I.node].world;
        for (int r = 0; r < 3; ++r)
            for (int k = 0; k < 4; ++k) bd.transform[r * 4 + k] = W[k * 4 + r];
        if (!hair) {
            bd.indices = g.idx;
            bd.vertex_count = nv;
        } else if (d->hair_mode == VKHRT_GEOMETRY_VOXELS) {
            // ProcessHairVoxels: the grid over the mesh's bounding box (geometry_processor.cpp:716)
            bd.flags |= VKHRT_BLAS_HAS_BOUNDS | d->voxel_flags;
            bd.voxel_size = d->voxel_size;
            bd.voxel_lod = d->voxel_lod;
            std::memcpy(bd.bounds_min, M.lo, sizeof(bd.bounds_min));
            std::memcpy(bd.bounds_max, M.hi, sizeof(bd.bounds_max));
        }
        rc = vkhrt_blas_build(c, &bd, stream, &blases[i]);
        if (rc != VKHRT_SUCCESS) c->err = "asset instance " + std::to_string(i) + ": " + c->err;
    }
    hipStreamSynchronize(st);
    for (Gen& g : gen) {
        if (g.recs != g.verts && g.recs != g.lines) hipFree(g.recs);
        hipFree(g.verts);
        hipFree(g.idx);
        hipFree(g.lines);
    }
    if (rc != VKHRT_SUCCESS)
        for (uint32_t i = 0; i < n; ++i) {
            vkhrt_blas_destroy(blases[i]);
            blases[i] = nullptr;
        }
    return rc;
}

// ---- profiling
// ---- neural shadow visibility
int vkhrt_shadow_mlp_create(vkhrt_context* c, const vkhrt_shadow_mlp_desc* d, vkhrt_shadow_mlp** out) {
    if (!c || !d || !out) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    float center[3], inv_half[3];
    for (int i = 0; i < 3; ++i) {
        const float lo = d->scene_min[i], hi = d->scene_max[i];
        if (!(hi > lo)) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "empty scene box");
        center[i] = 0.5f * (lo + hi);
        inv_half[i] = 2.0f / (hi - lo);
    }
    hipError_t e = vk::mlp_enable_train_lds();
    if (e != hipSuccess) return hip_fail(c, e, "MLP training kernel LDS");
    vkhrt_shadow_mlp* m = new (std::nothrow) vkhrt_shadow_mlp();
    if (!m) return VKHRT_ERROR_OUT_OF_MEMORY;
    m->ctx = c;
    std::memcpy(m->dev.center, center, sizeof(center));
    std::memcpy(m->dev.inv_half, inv_half, sizeof(inv_half));
    const size_t P = vk::kMlpParamCount;
    e = hipMalloc(&m->dev.master, P * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&m->dev.adam_m, P * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&m->dev.adam_v, P * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&m->dev.packed, 120 * 256 * sizeof(uint16_t));
    if (e == hipSuccess) e = hipMalloc(&m->loss, sizeof(float));
    if (e != hipSuccess) {
        vkhrt_shadow_mlp_destroy(m);
        return hip_fail(c, e, "MLP buffers");
    }
    // He-uniform weights, zero biases
    std::vector<float> w(P, 0.0f);
    uint64_t st = d->seed * 0xD1B54A32D192ED03ull + 0x5DEECE66Dull;
    auto fill = [&](int off, int fan_in, int count) {
        const float lim = std::sqrt(6.0f / float(fan_in));
        for (int i = 0; i < count; ++i)
            w[off + i] = lim * (2.0f * float(splitmix(st) >> 40) * (1.0f / 16777216.0f) - 1.0f);
    };
    fill(0, 96, 64 * 96);
    fill(6208, 64, 64 * 64);
    fill(10368, 64, 64 * 64);
    fill(14528, 64, 64 * 64);
    fill(18688, 64, 64);
    int rc = vkhrt_shadow_mlp_set_params(c, m, w.data());
    if (rc != VKHRT_SUCCESS) {
        vkhrt_shadow_mlp_destroy(m);
        return rc;
    }
    *out = m;
    return VKHRT_SUCCESS;
}

void vkhrt_shadow_mlp_destroy(vkhrt_shadow_mlp* m) {
    if (!m) return;
    hipFree(m->dev.master);
    hipFree(m->dev.adam_m);
    hipFree(m->dev.adam_v);
    hipFree(m->dev.packed);
    hipFree(m->partial);
    hipFree(m->loss);
    delete m;
}

int vkhrt_shadow_mlp_get_params(vkhrt_context* c, const vkhrt_shadow_mlp* m, float* params) {
    if (!c || !m || !params) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "null argument");
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(params, m->dev.master, vk::kMlpParamCount * sizeof(float), hipMemcpyDeviceToHost);
    return e == hipSuccess ? VKHRT_SUCCESS : hip_fail(c, e, "MLP params");
}

// Sets the FP32 parameters, clears the Adam state and rebuilds the bf16 fragments.
int vkhrt_shadow_mlp_set_params(vkhrt_context* c, vkhrt_shadow_mlp* m, const float* params) {
    if (!c || !m || !params) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "null argument");
    const size_t bytes = vk::kMlpParamCount * sizeof(float);
    std::vector<float> zero(vk::kMlpParamCount, 0.0f);
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(m->dev.master, params, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(m->dev.adam_m, zero.data(), bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(m->dev.adam_v, zero.data(), bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = vk::launch_mlp_pack(m->dev, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    m->step = 0;
    return e == hipSuccess ? VKHRT_SUCCESS : hip_fail(c, e, "MLP params");
}

int vkhrt_shadow_mlp_collect(vkhrt_context* c, const vkhrt_tlas* t, const vkhrt_ray* rays, uint32_t count,
                             float* samples, float* targets, void* stream) {
    if (!c || !t || (count && (!rays || !samples || !targets))) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "bad argument");
    if (count == 0) return VKHRT_SUCCESS;
    hipStream_t st = pick(c, stream);
    vk::HitRec* hits = nullptr;
    hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&hits), size_t(count) * sizeof(vk::HitRec), st);
    if (e != hipSuccess) return hip_fail(c, e, "collect scratch");
    int rc = vkhrt_trace_rays(c, t, rays, count, reinterpret_cast<vkhrt_hit*>(hits), VKHRT_TRACE_TERMINATE_ON_FIRST_HIT,
                              stream);
    if (rc == VKHRT_SUCCESS) {
        e = vk::launch_mlp_collect(reinterpret_cast<const float*>(rays), hits, count, samples, targets, st);
        if (e != hipSuccess) rc = hip_fail(c, e, "collect");
    }
    hipFreeAsync(hits, st);
    return rc;
}

namespace {
// One pass of gradient steps over `count` samples (or the full-batch gradient when `grad`).
int mlp_run(vkhrt_context* c, vkhrt_shadow_mlp* m, const float* samples, const float* targets, uint32_t count,
            uint32_t batch, uint32_t epochs, float lr, uint64_t seed, float* grad, float* mean_loss, hipStream_t st) {
    int dev = 0, cus = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const uint32_t step_n = std::min(batch, count);
    const uint32_t blocks = std::max(1u, std::min<uint32_t>(uint32_t(std::max(cus, 1)), (step_n + 63u) / 64u));
    if (blocks > m->partial_blocks) {
        hipFree(m->partial);
        m->partial = nullptr;
        m->partial_blocks = 0;
        hipError_t e = hipMalloc(&m->partial, size_t(blocks) * vk::kMlpParamCount * sizeof(float));
        if (e != hipSuccess) return hip_fail(c, e, "MLP partials");
        m->partial_blocks = blocks;
    }
    uint64_t rs = seed * 0x9E3779B97F4A7C15ull + 1;
    hipError_t e = hipSuccess;
    for (uint32_t ep = 0; ep < epochs && e == hipSuccess; ++ep) {
        vk::MlpTrainParams tp{};
        tp.samples = reinterpret_cast<const float4*>(samples);
        tp.targets = targets;
        tp.total = count;
        tp.partial = m->partial;
        tp.loss = m->loss;
        // epoch order i -> (i mul + add) mod count with gcd(mul, count) = 1; identity for the gradient
        tp.perm_mul = 1;
        tp.perm_add = 0;
        if (!grad && count > 1) {
            uint64_t mul = (splitmix(rs) % count) | 1u;
            while (std::gcd<uint64_t, uint64_t>(mul, count) != 1) mul += 2;
            tp.perm_mul = mul % count;
            tp.perm_add = splitmix(rs) % count;
        }
        e = hipMemcpyAsync(m->loss, &kZeroF, sizeof(float), hipMemcpyHostToDevice, st);
        for (uint32_t first = 0; first < count && e == hipSuccess; first += step_n) {
            tp.first = first;
            tp.n = std::min(step_n, count - first);
            const uint32_t nb = std::max(1u, std::min(blocks, (tp.n + 63u) / 64u));
            vk::AdamParams ap{lr, 0.9f, 0.999f, 1e-8f, 1.0f, 1.0f};
            if (!grad) {
                ++m->step;
                ap.bias1 = 1.0f - std::pow(0.9f, float(m->step));
                ap.bias2 = 1.0f - std::pow(0.999f, float(m->step));
            }
            e = vk::launch_mlp_train_step(m->dev, tp, nb, grad, ap, st);
            if (grad) break;
        }
    }
    if (e != hipSuccess) return hip_fail(c, e, "MLP training");
    if (mean_loss) {
        float l = 0.0f;
        e = hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(c, e, "MLP loss");
        *mean_loss = l / float(count);
    }
    return VKHRT_SUCCESS;
}
}  // namespace

int vkhrt_shadow_mlp_train(vkhrt_context* c, vkhrt_shadow_mlp* m, const float* samples, const float* targets,
                           uint32_t count, const vkhrt_shadow_mlp_train_desc* d, float* mean_loss, void* stream) {
    if (!c || !m || !d || (count && (!samples || !targets))) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "bad argument");
    if (count == 0 || d->epochs == 0) return VKHRT_SUCCESS;
    if (d->batch == 0 || !(d->learning_rate > 0.0f)) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "bad training setup");
    return mlp_run(c, m, samples, targets, count, d->batch, d->epochs, d->learning_rate, d->seed, nullptr, mean_loss,
                   pick(c, stream));
}

int vkhrt_shadow_mlp_gradient(vkhrt_context* c, vkhrt_shadow_mlp* m, const float* samples, const float* targets,
                              uint32_t count, float* gradient, float* mean_loss, void* stream) {
    if (!c || !m || !gradient || !count || !samples || !targets)
        return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "bad argument");
    return mlp_run(c, m, samples, targets, count, count, 1, 0.0f, 0, gradient, mean_loss, pick(c, stream));
}

int vkhrt_shadow_mlp_infer(vkhrt_context* c, const vkhrt_shadow_mlp* m, const float* samples, uint32_t count,
                           float* visibility, void* stream) {
    if (!c || !m || (count && (!samples || !visibility))) return fail(c, VKHRT_ERROR_INVALID_ARGUMENT, "bad argument");
    hipError_t e = vk::launch_mlp_infer(m->dev, samples, count, visibility, pick(c, stream));
    return e == hipSuccess ? VKHRT_SUCCESS : hip_fail(c, e, "MLP inference");
}

int vkhrt_profiling_enable(vkhrt_context* c, int enable) {
    if (!c) return VKHRT_ERROR_INVALID_ARGUMENT;
    resolve_timings(c);
    c->profiling = enable != 0;
    c->stats.clear();
    return VKHRT_SUCCESS;
}

int vkhrt_kernel_stats(vkhrt_context* c, const char* name, double* total_ms, uint32_t* launches) {
    if (!c || !name) return VKHRT_ERROR_INVALID_ARGUMENT;
    resolve_timings(c);
    auto it = c->stats.find(name);
    if (total_ms) *total_ms = it == c->stats.end() ? 0.0 : it->second.first;
    if (launches) *launches = it == c->stats.end() ? 0u : it->second.second;
    return VKHRT_SUCCESS;
}

int vkhrt_work_counters(vkhrt_context* c, uint64_t* out, uint32_t count, int reset) {
    if (!c || (count && !out)) return VKHRT_ERROR_INVALID_ARGUMENT;
    if (!c->work_counts) return fail(c, VKHRT_ERROR_UNSUPPORTED, "work counters exist in the counting build only");
    const uint32_t n = std::min<uint32_t>(count, vk::kWorkCounters);
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess && n) e = hipMemcpy(out, c->work_counts, n * sizeof(uint64_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess && reset) e = hipMemset(c->work_counts, 0, vk::kWorkCounters * sizeof(unsigned long long));
    return e == hipSuccess ? VKHRT_SUCCESS : hip_fail(c, e, "work counters");
}

int vkhrt_status_counters(vkhrt_context* c, uint32_t* overflows) {
    if (!c || !overflows) return VKHRT_ERROR_INVALID_ARGUMENT;
    uint32_t st[4];
    hipError_t e = hipMemcpy(st, c->status, 16, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(c, e, "status readback");
    *overflows = st[0];
    return VKHRT_SUCCESS;
}

}  // extern "C"
