// Host entry points of the geometry front-end kernels (geometry.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vk {

struct StrandParams {
    float sphere_radius, step, gravity, curl, r_root, r_tip;
};

hipError_t gen_strands(uint64_t seed, uint32_t nstrands, uint32_t nseg, const StrandParams& hp, float4* out,
                       hipStream_t st);
hipError_t gen_strand_lines(const float4* v, uint32_t nstrands, uint32_t nseg, float4* lines, hipStream_t st);
hipError_t gen_curves(const float4* lines, uint32_t n, float tension, float4* curves, hipStream_t st);
hipError_t gen_curve_lines(const float4* curves, uint32_t n, uint32_t pieces, float4* lines, hipStream_t st);
hipError_t gen_lss(const float4* lines, uint32_t n, float constant_radius, float4* caps, hipStream_t st);
hipError_t gen_dots(const float4* lines, uint32_t n, float constant_radius, float4* pieces, hipStream_t st);
hipError_t lss_chain(const float4* caps, uint32_t n, float4* out, hipStream_t st);
hipError_t gen_mesh_lines(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t n, float radius,
                          float4* lines, uint32_t* bad, hipStream_t st);
hipError_t gen_slot_of(const uint32_t* prim_ids, uint32_t n, uint32_t* slot_of, hipStream_t st);
hipError_t gen_tri_records(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t ntri, float4* recs,
                           float4* normals, uint32_t* bad, hipStream_t st);

}  // namespace vk
