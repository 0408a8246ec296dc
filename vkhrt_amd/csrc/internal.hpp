// Internal data layout shared by the C-ABI host code and the gfx950 kernels.
//
// HBM layout of one bottom-level structure (BLAS):
//   prims     : primitive records in leaf order, float4-aligned
//                 phantom curve : 4 x float4 (control point xyz, radius)      64 B
//                 LSS capsule   : 2 x float4 (a xyz, +-ra ; b xyz, rb)        32 B
//                 DOTS piece    : 4 x float4 (start,r0 ; end,r1 ; s ; t)      64 B
//                 triangle      : 3 x float4 (vertex positions)               48 B
//   voxels    : voxel BLAS only, one occupancy bit per cell (x fastest) in u32 words; no nodes
//   vbricks   : voxel BLAS only, one bit per 4x4x4-cell brick (x fastest), set when any of its
//               cells is filled, and vblocks the same per 16x16x16 cells; the walk crosses an
//               empty brick or block in one move
//   prim_ids  : original primitive index per leaf slot (u32)
//   nodes     : 8-wide BVH nodes with 8-bit quantized child boxes, 80 B each
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vk {

enum GeometryMode { kPhantom = 0, kLss = 1, kDots = 2, kTriangles = 3, kVoxels = 4 };

// 8-wide node. Child slot s has a box quantized to 8 bits per axis relative to the node
// origin with power-of-two scales; internal children are stored contiguously from
// child_base in slot order, leaf primitives contiguously from prim_base.
// meta[s]: 0 = empty slot, 0x80 = internal child, else (count << 5) | offset (count 1..3,
// offset 0..23 relative to prim_base). Slots are assigned so that visiting j = 0..7 with
// slot = j ^ octant(ray) approximates front-to-back order.
struct alignas(16) WideNode {
    float px, py, pz;
    uint32_t e_imask;  // ex | ey << 8 | ez << 16 | imask << 24 (biased float exponents)
    uint32_t child_base;
    uint32_t prim_base;
    uint32_t meta[2];
    uint32_t qlo_x[2], qlo_y[2], qlo_z[2];
    uint32_t qhi_x[2], qhi_y[2], qhi_z[2];
};
static_assert(sizeof(WideNode) == 80, "wide node is 80 bytes");

struct BlasDev {
    const float4* prims;
    const float* aux;  // phantom: precomputed early-out cylinder radius per leaf slot
    const uint32_t* prim_ids;
    const WideNode* nodes;
    const float4* normals;  // triangles: 3 vertex normals per original triangle (shading normal)
    const uint32_t* slot_of;  // phantom: leaf slot of each original primitive id
    uint32_t nprims;    // leaf primitives (curves / capsules / DOTS pieces)
    uint32_t nnodes;
    uint32_t mode;
    uint32_t radius_mode;  // phantom: 0 constant radius, 1 per control point
    float radius;          // phantom constant radius
    const uint32_t* voxels;  // voxels: occupancy bits of the traced level of detail
    const uint32_t* vbricks;  // voxels: one bit per 4x4x4 brick of cells (null: read every cell)
    const uint32_t* vblocks;  // voxels: one bit per 16x16x16 block of cells (null with vbricks)
    float vorigin[3];        // voxels: grid origin (object space)
    float vcell;             // voxels: cell edge at the traced level
    int32_t vres[3];         // voxels: cells per axis at the traced level
    uint32_t gbase;          // per instance: first global primitive id (merge keys); the ids of
                             // instance i are gbase(i) + [0, id range of its BLAS)
};

struct InstanceDev {
    BlasDev blas;
    float xf[12];      // world-from-object, row-major 3x4
    float inv[16];     // object-from-world, column-major 4x4
    float lo[3], hi[3];  // world bounds
    uint32_t identity;
    uint32_t material;  // BLAS material index
};

// Hit record, 32 bytes (matches vkhrt_hit in include/vkhrt.h).
struct HitRec {
    float t;
    uint32_t prim_id;
    uint32_t inst_id;
    float u;
    float n[3];
    float v;
};

enum TraceFlagsDev {
    kTraceAnyHit = 1u << 0,
    kTraceFp64Polish = 1u << 3,
};

}  // namespace vk
