/*
 * vkhrt C-ABI: MI355X (gfx950) drop-in for vkhrt's traceRayEXT hair path.
 *
 * Reference interfaces replaced (mmzala/vkhrt, file:line):
 *   BottomLevelAccelerationStructure(const BLASInput&, ...)
 *       include/bottom_level_acceleration_structure.hpp:34, BLASInput :19-29, BLASType :12-17
 *       -> vkhrt_blas_build / vkhrt_blas_destroy
 *   TopLevelAccelerationStructure(const std::vector<BottomLevelAccelerationStructure>&, ...)
 *       include/top_level_acceleration_structure.hpp:12; instances
 *       source/top_level_acceleration_structure.cpp:23-43 (customIndex = i, mask 0xFF)
 *       -> vkhrt_tlas_build / vkhrt_tlas_destroy
 *   traceRayEXT(tlas, gl_RayFlagsOpaqueEXT, 0xFF, 0, 0, 0, o, tMin, d, tMax, 0)
 *       shaders/ray_gen.rgen:26-45, reportIntersectionEXT shaders/hair_intersection.rint:146-149
 *       -> vkhrt_trace_rays
 *   Renderer::RecordRayTracingCommands -> traceRaysKHR(W, H, 1) source/renderer.cpp:156-166
 *       with ray_gen.rgen + hair_closest_hit.rchit / triangle_closest_hit.rchit + miss
 *       -> vkhrt_render
 *   CameraUniformData{viewInverse, projInverse} include/resources/camera_resource.hpp:7-11,
 *   FlyCamera::ViewMatrix/ProjectionMatrix source/fly_camera.cpp:25-35
 *       -> vkhrt_camera_look_at
 *   ProcessHairCurves / GenerateCurves / GenerateLinearSweptSpheres /
 *   GenerateDisjointOrthogonalTriangleStrips source/resources/model/geometry_processor.cpp:
 *   123-156, 221-298, 604-654 -> vkhrt_generate_* (run on the GPU)
 *   ModelLoader::LoadFromFile source/resources/model/model_loader.cpp:274-312, Node::GetWorldMatrix
 *   source/resources/model/model.cpp:6-18 -> vkhrt_asset_* (host glTF 2.0 reader);
 *   ProcessHair* + Renderer::InitializeBLAS (renderer.cpp:694-728) -> vkhrt_asset_build
 *
 * Conventions: all functions return a vkhrt_status (0 = success, negative = error, no
 * aborts, no exceptions across the ABI); vkhrt_last_error() describes the last failure of
 * a context. Buffers are device pointers unless a flag says otherwise. `stream` is a
 * hipStream_t (NULL = the HIP null stream). A context is not internally locked: use one
 * per host thread. Calls on one context share its scratch memory (counters, queues, hit and
 * polish buffers): issue them on one stream, or give each stream its own context.
 */
#ifndef VKHRT_H
#define VKHRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VKHRT_ABI_VERSION 5

typedef struct vkhrt_context vkhrt_context;
typedef struct vkhrt_blas vkhrt_blas;
typedef struct vkhrt_tlas vkhrt_tlas;
typedef struct vkhrt_shadow_mlp vkhrt_shadow_mlp;

typedef enum vkhrt_status {
    VKHRT_SUCCESS = 0,
    VKHRT_ERROR_INVALID_ARGUMENT = -1,
    VKHRT_ERROR_OUT_OF_MEMORY = -2,
    VKHRT_ERROR_DEVICE = -3,
    VKHRT_ERROR_UNSUPPORTED = -4,
    VKHRT_ERROR_STACK_OVERFLOW = -5,
    VKHRT_ERROR_NO_DEVICE = -6
} vkhrt_status;

/* Geometry modes: BLASType::eHair with AABB geometry + hair_intersection.rint (phantom),
 * VK_NV_ray_tracing_linear_swept_spheres (LSS), or DOTS triangles
 * (source/resources/model/model_loader.cpp:334 picks LSS or DOTS; phantom is wired here). */
typedef enum vkhrt_geometry_mode {
    VKHRT_GEOMETRY_PHANTOM_CURVES = 0,
    VKHRT_GEOMETRY_LSS = 1,
    VKHRT_GEOMETRY_DOTS = 2,
    VKHRT_GEOMETRY_TRIANGLES = 3, /* indexed triangle mesh (BLASType::eMesh, renderer.cpp:543-582) */
    VKHRT_GEOMETRY_VOXELS = 4     /* voxelized hair (BLASType::eVoxels, renderer.cpp:620-651) */
} vkhrt_geometry_mode;

/* vkhrt_blas_desc.flags */
#define VKHRT_BLAS_HOST_POINTERS (1u << 0)    /* `primitives` is host memory */
#define VKHRT_BLAS_PER_POINT_RADIUS (1u << 1) /* use per-point radii (.w) instead of `radius` */
#define VKHRT_BLAS_HAS_TRANSFORM (1u << 2)    /* `transform` is valid (else identity) */
#define VKHRT_BLAS_HAS_BOUNDS (1u << 3)       /* VOXELS: grid over bounds_min..bounds_max (else the lines' box) */
#define VKHRT_BLAS_VOXEL_REFERENCE_SWEEP (1u << 4) /* VOXELS: the reference's fill loop, see below */

/* Mirrors BLASInput (bottom_level_acceleration_structure.hpp:19-29).
 * primitives, `count` records of:
 *   PHANTOM_CURVES: 16 floats = start, cp1, cp2, end, each (x, y, z, radius) -- the 48-byte
 *                   Curve of include/resources/model/model.hpp:32-38 plus one radius per point
 *   LSS:            8 floats = (a.xyz, ra, b.xyz, rb); list indexing, chained end caps
 *                   (renderer.cpp:664-674): a start cap equal to the previous end is shared
 *   DOTS:           8 floats = line (start.xyz, r0, end.xyz, r1); 2 orthogonal quads
 *                   (4 triangles) per line are generated; primitive id = 4 * line + triangle
 *   TRIANGLES:      `primitives` holds vertex_count vertices of 8 floats (position, normal,
 *                   texCoord: the Vertex of model.hpp:44-49) and `indices` 3 u32 per triangle;
 *                   count = triangles; the hit reports the interpolated vertex normal
 *                   (triangle_closest_hit.rchit:31-36) and barycentrics u, v
 *   VOXELS:         8 floats = line (start.xyz, -, end.xyz, -), voxelized as capsules of radius
 *                   `radius` into a grid of `voxel_size` cells over the bounds
 *                   (GenerateVoxelMesh, geometry_processor.cpp:515-602; the bounds' extent is
 *                   rounded up to whole cells, the resolution truncated as there). By default
 *                   every cell of each sweep box is filled; with VKHRT_BLAS_VOXEL_REFERENCE_SWEEP
 *                   only the cells at the box's minimum x, with the index wrapping like the
 *                   reference's (:585-593). `voxel_lod` L traces the grid OR-reduced L times
 *                   (cells of voxel_size * 2^L). A hit is the first filled cell the ray enters,
 *                   at its entry t, primitive id = the cell's linear index x + y rx + z rx ry
 *                   at that level, normal = the face crossed
 * radius: constant radius used when VKHRT_BLAS_PER_POINT_RADIUS is not set (reference 0.02).
 * transform: row-major 3x4 world-from-object (BLASInput.transform, Node::GetWorldMatrix). */
typedef struct vkhrt_blas_desc {
    uint32_t mode;
    uint32_t flags;
    const float* primitives;
    uint32_t count;
    float radius;
    uint32_t material_index;
    uint32_t leaf_size; /* primitives per leaf child, 1..3; 0 = default (1) */
    float transform[12];
    const uint32_t* indices; /* TRIANGLES: 3 x count vertex indices (u32) */
    uint32_t vertex_count;   /* TRIANGLES: vertices in `primitives` */
    float voxel_size;        /* VOXELS: cell edge (0.1 in the reference) */
    uint32_t voxel_lod;      /* VOXELS: level of detail traced, 0..15 */
    float bounds_min[3];     /* VOXELS with VKHRT_BLAS_HAS_BOUNDS: the mesh bounds */
    float bounds_max[3];
} vkhrt_blas_desc;

typedef struct vkhrt_blas_info {
    uint32_t mode;
    uint32_t primitive_count; /* reported primitives (DOTS: triangles) */
    uint32_t leaf_records;    /* curves / capsules / DOTS pieces / filled voxels */
    uint32_t node_count;      /* 8-wide nodes */
    uint32_t depth;           /* 8-wide levels */
    uint64_t device_bytes;
    double build_ms;
} vkhrt_blas_info;

/* Ray record (32 B): origin, tmin, direction, tmax. */
typedef struct vkhrt_ray {
    float origin[3];
    float tmin;
    float direction[3];
    float tmax;
} vkhrt_ray;

/* Hit record (32 B). Miss: primitive_id = instance_id = 0xFFFFFFFF, t = 0.
 * t: phantom reports the distance along the normalized direction like
 *    hair_intersection.rint:70-71 (use unit directions); LSS/DOTS the ray parameter.
 * u: phantom curve parameter / LSS segment parameter / DOTS barycentric of vertex 1.
 * v: DOTS barycentric of vertex 2; phantom 1 if the root came from the FP64 polish.
 * normal: world-space unit normal (phantom: hit - curve(u), hair_intersection.rint:74-76;
 *    LSS: hit - mix(a, b, u), triangle_closest_hit.rchit:50-55; DOTS: geometric normal). */
typedef struct vkhrt_hit {
    float t;
    uint32_t primitive_id;
    uint32_t instance_id;
    float u;
    float normal[3];
    float v;
} vkhrt_hit;

/* vkhrt_trace_rays flags */
#define VKHRT_TRACE_TERMINATE_ON_FIRST_HIT (1u << 0) /* shadow / any-hit rays */
#define VKHRT_TRACE_FP64_POLISH (1u << 3)            /* FP64 Newton polish of grazing phantom roots */
#define VKHRT_TRACE_ACCURATE_NORMALS (1u << 5)       /* phantom hits: t, u and normal of the final hit
                                                        re-evaluated in double-word FP32 (the FP64
                                                        intersector's iteration; no FP64 used) */

/* Column-major inverse view / projection (CameraUniformData). */
typedef struct vkhrt_camera {
    float view_inverse[16];
    float proj_inverse[16];
} vkhrt_camera;

/* vkhrt_render_desc.flags */
#define VKHRT_RENDER_SHADOWS (1u << 0)       /* one shadow ray per hit toward (0, 1, 0) */
#define VKHRT_RENDER_DEBUG_PALETTE (1u << 1) /* debug.glsl palette by primitive id */
#define VKHRT_RENDER_RGBA8 (1u << 2)         /* image is RGBA8 (else RGBA32F) */
#define VKHRT_RENDER_FP64_POLISH (1u << 3)
#define VKHRT_RENDER_ACCURATE_NORMALS (1u << 5) /* primary phantom hits re-evaluated as with
                                                   VKHRT_TRACE_ACCURATE_NORMALS (persistent
                                                   kernels, single-mode TLAS) */
#define VKHRT_RENDER_NEURAL_SHADOWS (1u << 4) /* shadow visibility from `shadow_mlp`; rays whose
                                                 predicted variance p (1 - p) exceeds
                                                 `shadow_variance_threshold` trace exactly */

/* Area light (extension; the reference lights with the fixed direction of shading.glsl:4): the
 * parallelogram center + a edge_u + b edge_v, a, b in [-1, 1]. Each hit traces `samples`
 * shadow rays toward points of it fixed by (seed, pixel, sample) (tmax = distance to the
 * point); the pixel's visibility is their mean. */
typedef struct vkhrt_area_light {
    float center[3];
    float edge_u[3];
    float edge_v[3];
    uint32_t samples; /* 1..16 */
    uint64_t seed;
} vkhrt_area_light;

/* Environment map of the miss shader (miss.rmiss:17-38): an equirectangular RGBA32F image
 * (row-major, `width` x `height` texels, device memory) sampled bilinearly with repeat
 * addressing at DirectionToUV(normalize(-direction)), then 1 - exp(-c exposure) and the 1/2.2
 * gamma. Without one, primary misses are black. */
typedef struct vkhrt_environment_map {
    const float* texels;
    uint32_t width;
    uint32_t height;
    float exposure; /* 1 in the reference */
} vkhrt_environment_map;

typedef struct vkhrt_render_desc {
    uint32_t width;
    uint32_t height;
    uint32_t flags;
    void* image;                    /* device, width * height pixels */
    uint32_t* primitive_ids;        /* device, optional (0xFFFFFFFF = miss) */
    const float* light_visibility;  /* device, optional per-pixel light visibility in [0, 1]
                                       replacing the shadow ray */
    const vkhrt_shadow_mlp* shadow_mlp; /* with VKHRT_RENDER_NEURAL_SHADOWS */
    float shadow_variance_threshold;    /* 0 = every ray exact, 0.25 = none */
    const vkhrt_area_light* area_light; /* host, optional: shadow rays toward an area light instead
                                           of the direction (0, 1, 0); needs a TLAS of one
                                           geometry mode */
    const vkhrt_environment_map* environment; /* host, optional: the miss colour of primary rays */
} vkhrt_render_desc;

/* ---- context */
int vkhrt_context_create(int device, vkhrt_context** out);
void vkhrt_context_destroy(vkhrt_context* ctx);
const char* vkhrt_last_error(const vkhrt_context* ctx);
int vkhrt_synchronize(vkhrt_context* ctx, void* stream);
int vkhrt_abi_version(void);

/* Options (autotuned parameters). */
#define VKHRT_OPTION_STACK_KIND 1 /* minimum LDS stack entries per lane: 0 -> 8, 1 -> 12, 2 -> 16,
                                     3 -> 32; raised automatically to the deepest BVH */
#define VKHRT_OPTION_LEAF_SIZE 2  /* default leaf size for builds (1..3) */
#define VKHRT_OPTION_PERSISTENT 3 /* 1 (default): persistent wavefronts with per-lane ray refill;
                                     0: one thread per ray */
#define VKHRT_OPTION_REFILL_MIN 4 /* persistent: idle lanes (1..64) a wavefront waits for before
                                     fetching new rays together (default 16) */
#define VKHRT_OPTION_BLOCKS_PER_CU 6 /* persistent traversal grids: 256-thread blocks (4 wavefronts)
                                       per CU, 0 (default) = as many as the LDS stack allows */
#define VKHRT_OPTION_BVH_WIDTH 7     /* children per node for later builds, 4..8 (default 8; nodes
                                       keep the 8-slot layout) */
#define VKHRT_OPTION_BVH_COLLAPSE 9  /* how later builds collapse the binary tree into wide nodes: 0 greedy
                                       (open the largest-area child until the width is reached; default),
                                       c in 1..8 SAH-optimal (node cost 1, primitive cost c / 4) */
#define VKHRT_OPTION_RAY_ORDER 5   /* vkhrt_render work order: 0 row-major (default), 1 8x8 pixel
                                     tiles, one tile per wavefront, 2 64x64 tiles of 8x8 tiles */
#define VKHRT_OPTION_XCD_CHUNK 8   /* persistent traversal (vkhrt_render, vkhrt_render_batch,
                                     vkhrt_trace_rays): 0 one work counter for the grid; a power
                                     of two in 64..2^20 (default 4096): the rays are handed out in
                                     runs of this many, run k by the counter of XCD k mod 8 (each
                                     XCD's L2 then serves its own image regions) */
#define VKHRT_OPTION_SHADOW_WAVE_LIMIT 10 /* neural shadow form: a fallback queue of up to this many
                                            rays (0..2^20, default 65536) is walked by lane groups, one
                                            ray per group, a longer one by the persistent lanes; 0 =
                                            always the persistent lanes. Results do not depend on it */
#define VKHRT_OPTION_SHADOW_GROUP_LANES 11 /* lanes per ray of that walk: 8, 16 (default), 32 or 64 */
#define VKHRT_OPTION_STEPS_PER_ITERATION 12 /* persistent phantom traversal: 1..8 traversal steps per
                                              pass over the lanes' refill and finish bookkeeping
                                              (default 3) */
int vkhrt_set_option(vkhrt_context* ctx, int option, int value);

/* ---- acceleration structures */
int vkhrt_blas_build(vkhrt_context* ctx, const vkhrt_blas_desc* desc, void* stream, vkhrt_blas** out);
int vkhrt_blas_info_get(const vkhrt_blas* blas, vkhrt_blas_info* out);
void vkhrt_blas_destroy(vkhrt_blas* blas);
/* Voxel BLAS: the traced grid (cells per axis, origin, cell edge, occupancy words). */
typedef struct vkhrt_voxel_grid {
    int32_t res[3];
    float origin[3];
    float cell;
    uint64_t words; /* u32 words of occupancy, (res product + 31) / 32 */
} vkhrt_voxel_grid;
int vkhrt_blas_voxel_grid_get(const vkhrt_blas* blas, vkhrt_voxel_grid* out);
/* Copies `words` occupancy words (bit i of word w = cell 32 w + i, x fastest) to host memory;
 * synchronous. */
int vkhrt_blas_voxels_read(vkhrt_context* ctx, const vkhrt_blas* blas, uint32_t* words, uint64_t count);
/* Instance i = blases[i] with its transform; customIndex = i. */
int vkhrt_tlas_build(vkhrt_context* ctx, const vkhrt_blas* const* blases, uint32_t count, void* stream,
                     vkhrt_tlas** out);
void vkhrt_tlas_destroy(vkhrt_tlas* tlas);

/* ---- tracing */
int vkhrt_trace_rays(vkhrt_context* ctx, const vkhrt_tlas* tlas, const vkhrt_ray* rays, uint32_t count,
                     vkhrt_hit* hits, uint32_t flags, void* stream);
/* vkhrt_trace_rays without any device allocation: no counterpart in the reference. For callers
 * that cannot allocate (HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, flags and results as vkhrt_trace_rays; its counters and FP64
 * polish scratch live in one fixed device array per GPU (16,781,312 bytes, held until the process
 * ends). With VKHRT_TRACE_FP64_POLISH, count is at most 262,144 (VKHRT_ERROR_INVALID_ARGUMENT
 * above); without it, any count. Every call on one GPU shares that array: calls that overlap
 * on one GPU (on different streams, or from different host threads) race and must be ordered
 * by the caller. Errors found on the device are reported by vkhrt_synchronize. */
int vkhrt_trace_rays_noalloc(vkhrt_context* ctx, const vkhrt_tlas* tlas, const vkhrt_ray* rays, uint32_t count,
                             vkhrt_hit* hits, uint32_t flags, void* stream);
int vkhrt_render(vkhrt_context* ctx, const vkhrt_tlas* tlas, const vkhrt_camera* camera,
                 const vkhrt_render_desc* desc, void* stream);
/* glm::lookAt + perspectiveRH_ZO with the Vulkan Y flip, inverted (host function). */
int vkhrt_camera_look_at(const float eye[3], const float center[3], const float up[3], float tan_half_fovy,
                         float aspect, float znear, float zfar, vkhrt_camera* out);

/* ---- geometry front-end (device buffers, float4 records) */
typedef struct vkhrt_strand_params {
    float sphere_radius; /* 100 */
    float step;          /* 1.5 */
    float gravity;       /* 0.3 */
    float curl;          /* 0.2 */
    float r_root;        /* 0.04 */
    float r_tip;         /* 0.01 */
} vkhrt_strand_params;
/* Seeded synthetic strands: nstrands x (nseg + 1) vertices of (x, y, z, radius). */
int vkhrt_generate_strands(vkhrt_context* ctx, uint64_t seed, uint32_t nstrands, uint32_t nseg,
                           const vkhrt_strand_params* params, float* vertices, void* stream);
/* Strand polylines -> lines (start.xyz, r0, end.xyz, r1), GenerateLines. */
int vkhrt_generate_lines(vkhrt_context* ctx, const float* vertices, uint32_t nstrands, uint32_t nseg, float* lines,
                         void* stream);
/* GenerateCurves: lines -> curves (16 floats each), tension 1 in the reference. */
int vkhrt_generate_curves(vkhrt_context* ctx, const float* lines, uint32_t count, float tension, float* curves,
                          void* stream);
/* Cubic -> `pieces` lines per curve sampled at t = k / pieces. */
int vkhrt_curves_to_lines(vkhrt_context* ctx, const float* curves, uint32_t count, uint32_t pieces, float* lines,
                          void* stream);
/* GenerateLinearSweptSpheres: lines -> capsules (8 floats), radius max(r, 0.001);
 * constant_radius < 0 keeps per-end radii. */
int vkhrt_generate_lss(vkhrt_context* ctx, const float* lines, uint32_t count, float constant_radius, float* capsules,
                       void* stream);
/* GenerateDisjointOrthogonalTriangleStrips frames: lines -> pieces (16 floats:
 * start, r0, end, r1, s, 0, t, 0). constant_radius < 0 keeps per-end radii. */
int vkhrt_generate_dots(vkhrt_context* ctx, const float* lines, uint32_t count, float constant_radius, float* pieces,
                        void* stream);

/* ---- auxiliary ray work */
/* ray_gen.rgen as a standalone generator: width * height rays (row-major launch ids). */
int vkhrt_generate_camera_rays(vkhrt_context* ctx, const vkhrt_camera* camera, uint32_t width, uint32_t height,
                               float tmin, float tmax, vkhrt_ray* rays, void* stream);
/* One ray toward `direction` from every hit (origin = o + t d); compacted. `count_out` and
 * `source` (index of the source ray, may be NULL) are device pointers. */
int vkhrt_generate_shadow_rays(vkhrt_context* ctx, const vkhrt_ray* rays, const vkhrt_hit* hits, uint32_t count,
                               const float direction[3], float tmin, float tmax, vkhrt_ray* out, uint32_t* source,
                               uint32_t* count_out, void* stream);
/* Per ray, the number of primitives whose exact bounding box (GenerateAABBs,
 * geometry_processor.cpp:422-436) overlaps [tmin, t_end], t_end = hit t or tmax on a miss. */
int vkhrt_count_candidates(vkhrt_context* ctx, const vkhrt_tlas* tlas, const vkhrt_ray* rays, const vkhrt_hit* hits,
                           uint32_t count, uint32_t* counts, void* stream);
/* vkhrt_trace_rays plus per-ray traversal statistics: stats[2i] = 8-wide nodes visited,
 * stats[2i+1] = primitive records tested (single-mode TLAS, no FP64 polish). */
int vkhrt_trace_stats(vkhrt_context* ctx, const vkhrt_tlas* tlas, const vkhrt_ray* rays, uint32_t count,
                      vkhrt_hit* hits, uint32_t flags, uint32_t* stats, void* stream);
/* Rays of the last vkhrt_render on this context (synchronizes its stream). */
int vkhrt_render_ray_counts(vkhrt_context* ctx, uint64_t* primary, uint64_t* shadow);

/* Rays of the last vkhrt_render: primary rays, shadow queries (hit pixels that needed a light
 * visibility) and shadow rays traced exactly (all queries in the exact form, the fallback in
 * the neural form). Synchronizes the render's stream. */
typedef struct vkhrt_render_stats {
    uint64_t primary_rays;
    uint64_t shadow_queries;
    uint64_t shadow_rays_traced;
} vkhrt_render_stats;
int vkhrt_render_stats_get(vkhrt_context* ctx, vkhrt_render_stats* out);

/* ---- batched path tracing (extension; no reference counterpart: the reference renders one
 * view per traceRaysKHR, source/renderer.cpp:156-166)
 * Every pixel of every view is one path through the TLAS; all views go through the same
 * launches (one persistent traversal launch per depth for the whole batch, over a global
 * path queue). Path rule: light from (0, 1, 0) with radiance 1; at a hit with normal n
 * (facing normal nf = n turned against the ray), albedo a = material_albedo[instance
 * material] * Shade(n) (shading.glsl:1-10 with full light); light contribution
 * T * a * (nf . (0,1,0)) when positive and its shadow ray (origin = hit, tmin 0.001) is
 * unoccluded; up to `bounces` cosine-distributed bounce rays about nf with throughput
 * T * a; a ray that leaves the scene adds T * environment. Random numbers are a fixed
 * function of (seed, path, depth), so a frame is deterministic. */
#define VKHRT_BATCH_MAX_BOUNCES 8
#define VKHRT_BATCH_MAX_MATERIALS 16 /* material indices above 15 use entry 15 */
#define VKHRT_BATCH_RGBA8 (1u << 2)          /* images are RGBA8 (else RGBA32F radiance) */
#define VKHRT_BATCH_FP64_POLISH (1u << 3)    /* FP64 Newton polish of grazing phantom roots */
#define VKHRT_BATCH_SORT_MATERIALS (1u << 5) /* shade each depth in material order */
#define VKHRT_BATCH_SORT_RAYS (1u << 7)      /* reorder each bounce depth's rays by direction octant
                                                and origin cell (Morton order in the scene box)
                                                before tracing */
#define VKHRT_BATCH_GRAPH (1u << 6)          /* run the frame as a captured HIP graph; the graph
                                                is captured on the first call and replayed while
                                                the TLAS, total pixel count, bounces, flags and
                                                options stay the same (cameras, view sizes,
                                                output buffers, seed and colours may change) */
typedef struct vkhrt_view {
    vkhrt_camera camera;
    uint32_t width;
    uint32_t height;
    void* image;             /* device, width * height pixels */
    uint32_t* primitive_ids; /* device, optional: primary hit primitive per pixel */
} vkhrt_view;
typedef struct vkhrt_batch_desc {
    uint32_t flags;
    uint32_t bounces;                /* 0..VKHRT_BATCH_MAX_BOUNCES (0 = direct light only) */
    uint64_t seed;                   /* sampling seed (e.g. the frame number) */
    float environment[3];            /* radiance of rays leaving the scene */
    const float* material_albedo;    /* host, optional: material_count x 3 tints (default 1) */
    uint32_t material_count;         /* <= VKHRT_BATCH_MAX_MATERIALS */
} vkhrt_batch_desc;
/* Renders `view_count` views (at most 65536, 2^31 pixels in total) of `tlas`. A TLAS of one
 * geometry mode runs the persistent traversal kernels; a mixed one (e.g. hair and triangle
 * meshes from one asset) one lane per path and no FP64 polish. */
int vkhrt_render_batch(vkhrt_context* ctx, const vkhrt_tlas* tlas, const vkhrt_view* views, uint32_t view_count,
                       const vkhrt_batch_desc* desc, void* stream);
/* Rays of the last vkhrt_render_batch (synchronizes its stream): one primary ray per path,
 * the bounce rays and the shadow rays traced. */
typedef struct vkhrt_batch_stats {
    uint64_t paths;
    uint64_t bounce_rays;
    uint64_t shadow_rays;
    uint32_t graph_captures; /* graphs captured on this context so far */
    uint32_t reserved;
} vkhrt_batch_stats;
int vkhrt_batch_stats_get(vkhrt_context* ctx, vkhrt_batch_stats* out);

/* ---- autotuner (extension; the reference's driver builds and traces with fixed settings)
 * Times the caller's own workload under candidate settings (coordinate descent from the
 * current ones: each knob is swept with the others fixed, twice) and leaves the fastest in
 * the context. Each trial runs the workload `repeats` times (>= 1, after one warm-up) and
 * scores the median, timed with HIP events on `stream`. The workload's output buffers are
 * overwritten. Knobs: stack kind, persistent blocks per CU (waves per CU / 4), refill
 * minimum, XCD chunk, and per workload: the render's ray order, the batch's sort flags, the
 * BLAS's width and leaf size. Results never depend on the knobs, only the time does. */
typedef struct vkhrt_tuning {
    int32_t stack_kind;
    int32_t blocks_per_cu;
    int32_t refill_min;
    int32_t ray_order;
    int32_t bvh_width;
    int32_t leaf_size;
    uint32_t batch_flags;  /* vkhrt_autotune_batch: the chosen VKHRT_BATCH_SORT_* flags */
    uint32_t trials;
    float baseline_ms;     /* the workload with the starting settings */
    float best_ms;         /* ... with the chosen ones (<= baseline_ms) */
    int32_t xcd_chunk;     /* VKHRT_OPTION_XCD_CHUNK */
    int32_t bvh_collapse;  /* VKHRT_OPTION_BVH_COLLAPSE (vkhrt_autotune_blas) */
    int32_t steps_per_iteration; /* VKHRT_OPTION_STEPS_PER_ITERATION */
} vkhrt_tuning;
/* vkhrt_render(tlas, camera, desc) as the workload; sets stack kind, blocks per CU, refill
 * minimum, ray order, XCD chunk and steps per iteration. */
int vkhrt_autotune_render(vkhrt_context* ctx, const vkhrt_tlas* tlas, const vkhrt_camera* camera,
                          const vkhrt_render_desc* desc, uint32_t repeats, vkhrt_tuning* out, void* stream);
/* vkhrt_render_batch as the workload; sets stack kind, blocks per CU, refill minimum, XCD chunk
 * and steps per iteration and reports the best sort flags in out->batch_flags (desc->flags' sort
 * bits are the start). */
int vkhrt_autotune_batch(vkhrt_context* ctx, const vkhrt_tlas* tlas, const vkhrt_view* views, uint32_t view_count,
                         const vkhrt_batch_desc* desc, uint32_t repeats, vkhrt_tuning* out, void* stream);
/* Builds `blas` (with its TLAS) for each BVH width in {4, 6, 8} and leaf size in {1, 2, 3},
 * renders `desc` from `camera` on each and sets the fastest width and leaf size as the
 * context's options for later builds (desc.leaf_size of `blas` is ignored here). */
int vkhrt_autotune_blas(vkhrt_context* ctx, const vkhrt_blas_desc* blas, const vkhrt_camera* camera,
                        const vkhrt_render_desc* desc, uint32_t repeats, vkhrt_tuning* out, void* stream);

/* ---- neural shadow visibility (extension; no reference counterpart)
 * A small MLP predicts the light visibility of a shadow ray from its origin and direction:
 * (p normalized by the scene box, d) -> 6 x 8 frequencies x {sin, cos} = 96 features ->
 * 4 x (64, ReLU) -> sigmoid. Trained with binary cross-entropy on shadow rays the exact
 * tracer labels (vkhrt_shadow_mlp_collect); BF16 MFMA forward/backward, FP32 master weights,
 * Adam. Samples are 8 floats (p.xyz, 0, d.xyz, 0); targets 1 = lit, 0 = occluded. */
#define VKHRT_SHADOW_MLP_PARAMS 18753 /* W1[64][96] b1 W2[64][64] b2 W3 b3 W4 b4 w5[64] b5 */
typedef struct vkhrt_shadow_mlp_desc {
    float scene_min[3]; /* box the positions are normalized by (e.g. the TLAS bounds) */
    float scene_max[3];
    uint64_t seed;      /* weight initialization */
} vkhrt_shadow_mlp_desc;
typedef struct vkhrt_shadow_mlp_train_desc {
    uint32_t batch;       /* samples per Adam step (e.g. 65536) */
    uint32_t epochs;
    float learning_rate;  /* e.g. 3e-3 */
    uint64_t seed;        /* epoch orders */
} vkhrt_shadow_mlp_train_desc;
int vkhrt_shadow_mlp_create(vkhrt_context* ctx, const vkhrt_shadow_mlp_desc* desc, vkhrt_shadow_mlp** out);
void vkhrt_shadow_mlp_destroy(vkhrt_shadow_mlp* mlp);
/* FP32 parameters in the flat layout (host buffers; synchronous). */
int vkhrt_shadow_mlp_get_params(vkhrt_context* ctx, const vkhrt_shadow_mlp* mlp, float* params);
int vkhrt_shadow_mlp_set_params(vkhrt_context* ctx, vkhrt_shadow_mlp* mlp, const float* params);
/* Training rays from the hair surface: `count` shadow rays from points drawn on the TLAS's
 * phantom curves (record uniform over all instances' records, curve parameter and angle about
 * the fibre uniform, radius as the intersector's) toward a point of `light` (one per ray,
 * fixed by (seed, i); light->samples is ignored) or, with `light` NULL, along (0, 1, 0).
 * Label them with vkhrt_shadow_mlp_collect. */
int vkhrt_generate_surface_shadow_rays(vkhrt_context* ctx, const vkhrt_tlas* tlas, const vkhrt_area_light* light,
                                       uint32_t count, uint64_t seed, vkhrt_ray* rays, void* stream);
/* Traces `count` shadow rays (any hit) and writes training samples and visibility targets. */
int vkhrt_shadow_mlp_collect(vkhrt_context* ctx, const vkhrt_tlas* tlas, const vkhrt_ray* shadow_rays, uint32_t count,
                             float* samples, float* targets, void* stream);
/* Adam on the mean binary cross-entropy; `mean_loss` (host, optional) receives the mean loss
 * of the last epoch (synchronizes the stream). */
int vkhrt_shadow_mlp_train(vkhrt_context* ctx, vkhrt_shadow_mlp* mlp, const float* samples, const float* targets,
                           uint32_t count, const vkhrt_shadow_mlp_train_desc* desc, float* mean_loss, void* stream);
/* Gradient of the mean binary cross-entropy over all `count` samples (device, PARAMS floats),
 * no update; `mean_loss` (host, optional). */
int vkhrt_shadow_mlp_gradient(vkhrt_context* ctx, vkhrt_shadow_mlp* mlp, const float* samples, const float* targets,
                              uint32_t count, float* gradient, float* mean_loss, void* stream);
/* Predicted visibility p per sample (device). */
int vkhrt_shadow_mlp_infer(vkhrt_context* ctx, const vkhrt_shadow_mlp* mlp, const float* samples, uint32_t count,
                           float* visibility, void* stream);

/* ---- polyline / triangle assets (glTF 2.0)
 * Replaces ModelLoader::LoadFromFile (source/resources/model/model_loader.cpp:274-291: assimp
 * 6.0.1 ReadFile with Triangulate | FlipUVs | GenNormals | GenBoundingBoxes), LoadModel
 * (:293-312), ProcessMesh (:139-206), ProcessNodes (:208-264), ProcessMaterial factors
 * (:60-137) and Node::GetWorldMatrix (source/resources/model/model.cpp:6-18).
 * Reads .gltf (external buffers or base64 data: URIs) and .glb. Each mesh primitive is one
 * mesh, in file order; LINES / LINE_LOOP / LINE_STRIP become line meshes (index pairs),
 * TRIANGLES / TRIANGLE_STRIP / TRIANGLE_FAN triangle meshes; POINTS and sparse accessors are
 * UNSUPPORTED. Indices are mesh-local (the reference adds a mesh's firstVertex both in
 * ProcessMesh :157 and in GenerateLines geometry_processor.cpp:61; it is applied once here).
 * Nodes are listed in pre-order from the root; several scene roots get an identity root
 * "ROOT" (as assimp does). Triangle meshes without normals get face normals written per vertex,
 * the last face of a vertex winning (aiProcess_GenNormals); texCoords are the file's (the
 * importer's V flip and FlipUVs cancel). Loading is host work: no context, no device. */
typedef struct vkhrt_asset vkhrt_asset;

#define VKHRT_PRIMITIVE_TRIANGLES 0 /* Mesh::PrimitiveType::eTriangles */
#define VKHRT_PRIMITIVE_LINES 1     /* Mesh::PrimitiveType::eLines */

typedef struct vkhrt_asset_info {
    uint32_t mesh_count;     /* mesh primitives */
    uint32_t node_count;
    uint32_t material_count;
    uint32_t instance_count; /* (node, mesh) pairs = BLASes made by vkhrt_asset_build */
    uint32_t line_mesh_count;
    uint32_t triangle_mesh_count;
} vkhrt_asset_info;

/* Mesh (model.hpp Mesh + its slice of the vertex and index buffers). Host pointers owned by
 * the asset. */
typedef struct vkhrt_asset_mesh {
    uint32_t primitive;      /* VKHRT_PRIMITIVE_* */
    int32_t material;        /* material index, -1 = none */
    uint32_t vertex_count;
    uint32_t index_count;    /* 2 per line, 3 per triangle */
    const float* vertices;   /* vertex_count x 8 floats: position, normal, texCoord (Vertex) */
    const uint32_t* indices; /* mesh-local, each < vertex_count */
    float bounds_min[3];     /* boundingBox of all vertices (aiProcess_GenBoundingBoxes) */
    float bounds_max[3];
    const char* name;
} vkhrt_asset_mesh;

/* Node (model.hpp Node): column-major 4x4 matrices as glm::mat4. */
typedef struct vkhrt_asset_node {
    int32_t parent;         /* node index, -1 for the root */
    uint32_t mesh_count;
    const uint32_t* meshes; /* host, owned by the asset */
    float local[16];        /* localMatrix */
    float world[16];        /* GetWorldMatrix(): parents multiplied in from the nearest */
    const char* name;
} vkhrt_asset_node;

/* Material factors (MaterialCreation; textures are not read). glTF defaults when absent. */
typedef struct vkhrt_asset_material {
    float base_color[4]; /* pbrMetallicRoughness.baseColorFactor */
    float metallic;
    float roughness;
    float emissive[3];
    float pad;
} vkhrt_asset_material;

/* One BLAS input: node's world matrix as a row-major 3x4 (BLASInput.transform). */
typedef struct vkhrt_asset_instance {
    uint32_t node;
    uint32_t mesh;
    float transform[12];
} vkhrt_asset_instance;

int vkhrt_asset_load(const char* path, vkhrt_asset** out);
/* `data` holds a .glb or the text of a .gltf; relative buffer URIs resolve against base_dir
 * (NULL: only data: URIs and the GLB binary chunk are allowed). */
int vkhrt_asset_load_memory(const void* data, size_t size, const char* base_dir, vkhrt_asset** out);
void vkhrt_asset_destroy(vkhrt_asset* asset);
/* Message of the last failed vkhrt_asset_* call on this thread. */
const char* vkhrt_asset_last_error(void);
int vkhrt_asset_info_get(const vkhrt_asset* asset, vkhrt_asset_info* out);
int vkhrt_asset_mesh_get(const vkhrt_asset* asset, uint32_t index, vkhrt_asset_mesh* out);
int vkhrt_asset_node_get(const vkhrt_asset* asset, uint32_t index, vkhrt_asset_node* out);
int vkhrt_asset_material_get(const vkhrt_asset* asset, uint32_t index, vkhrt_asset_material* out);
int vkhrt_asset_instance_get(const vkhrt_asset* asset, uint32_t index, vkhrt_asset_instance* out);

/* ProcessHairCurves / ProcessHairLSS / ProcessHairDOTS (geometry_processor.cpp:604-688,740-779)
 * and Renderer::InitializeBLAS (renderer.cpp:694-728) on the device. */
typedef struct vkhrt_asset_build_desc {
    uint32_t hair_mode;        /* line meshes: VKHRT_GEOMETRY_PHANTOM_CURVES, _LSS, _DOTS or _VOXELS */
    float radius;              /* fibre radius (0.02 in the reference) */
    float tension;             /* GenerateCurves tension (1) */
    uint32_t leaf_size;        /* 0 = the context's default */
    uint32_t default_material; /* material_index of meshes without a material */
    float voxel_size;          /* VOXELS: cell edge (ProcessHairVoxels uses 0.1, :712) */
    uint32_t voxel_lod;        /* VOXELS: level of detail */
    uint32_t voxel_flags;      /* VOXELS: VKHRT_BLAS_VOXEL_REFERENCE_SWEEP or 0 */
} vkhrt_asset_build_desc;
/* One BLAS per instance (vkhrt_asset_instance_get order), with the instance's transform and
 * its mesh's material. Line meshes: GenerateLines (mesh-local indices) -> GenerateCurves ->
 * phantom curves, or -> capsules (LSS) / lines (DOTS) / a voxel grid over the mesh bounds
 * (VOXELS, ProcessHairVoxels :690-737); triangle meshes as
 * VKHRT_GEOMETRY_TRIANGLES. Each mesh is generated once and shared by its instances; every
 * BLAS owns its records (the reference points every hair BLAS at the start of the curve
 * buffer, renderer.cpp:611). blases == NULL with capacity 0 only sets *count; otherwise
 * capacity must cover the instances. On failure no BLAS is left. */
int vkhrt_asset_build(vkhrt_context* ctx, const vkhrt_asset* asset, const vkhrt_asset_build_desc* desc, void* stream,
                      vkhrt_blas** blases, uint32_t capacity, uint32_t* count);
/* GenerateLines (geometry_processor.cpp:45-67) for an indexed line mesh on the device: line i
 * = positions of vertices[indices[2i]] and vertices[indices[2i+1]] (8-float vertices,
 * mesh-local indices), radius `radius` at both ends; `count` lines. INVALID_ARGUMENT when an
 * index is not below vertex_count (synchronizes the stream to check). */
int vkhrt_generate_mesh_lines(vkhrt_context* ctx, const float* vertices, uint32_t vertex_count,
                              const uint32_t* indices, uint32_t count, float radius, float* lines, void* stream);

/* ---- profiling: per-kernel HIP-event timing on the launch stream */
int vkhrt_profiling_enable(vkhrt_context* ctx, int enable);
/* name: "render_primary", "render_shadow", "render_batch", "trace_rays", "blas_build". */
int vkhrt_kernel_stats(vkhrt_context* ctx, const char* name, double* total_ms, uint32_t* launches);
int vkhrt_status_counters(vkhrt_context* ctx, uint32_t* stack_overflows);
/* Work counters of the counting build (the library compiled with VKHRT_WORK_COUNTERS, e.g.
 * vkhrt_amd/libvkhrt_count.so): per traversal kind, loop iterations, intersection rounds,
 * traversal steps, node and leaf tests, Prhi calls, busy lanes and shader-clock cycles per
 * phase, summed over waves since the last reset (layout: vk::WorkCounter in
 * vkhrt_amd/csrc/build.hpp). Waits for the device. The product build returns
 * VKHRT_ERROR_UNSUPPORTED. */
int vkhrt_work_counters(vkhrt_context* ctx, uint64_t* out, uint32_t count, int reset);

#ifdef __cplusplus
}
#endif

#endif /* VKHRT_H */
