// Voxel hair (BLASType::eVoxels): line capsules voxelized into a dense occupancy grid
// (GenerateVoxelMesh, source/resources/model/geometry_processor.cpp:515-602) and coarser
// levels of detail by 2x2x2 OR-reduction. The grid is one bit per voxel, x fastest
// (index x + y * rx + z * rx * ry, GetVoxelIndex1D :488-491), in u32 words.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

namespace vk {

struct VoxelGrid {
    float origin[3];  // the mesh bounds' minimum
    float size;       // voxel edge
    int32_t res[3];
    uint32_t count;   // res.x * res.y * res.z
};

// GenerateVoxelMesh's grid (:519-529): the bounds' extent rounded up to whole voxels with
// NextDivisible (ceil(e / s) * s, :472-476) and the resolution as the truncated quotient of
// that extent by s, all in float. False when the extent is not finite, a voxel is not
// positive, or the grid holds 2^31 voxels or more.
inline bool voxel_grid_params(const float lo[3], const float hi[3], float size, VoxelGrid& g) {
    if (!(size > 0.0f) || !std::isfinite(size)) return false;
    uint64_t count = 1;
    for (int k = 0; k < 3; ++k) {
        const float ext = hi[k] - lo[k];
        if (!std::isfinite(ext) || !std::isfinite(lo[k]) || ext < 0.0f) return false;
        const float whole = std::ceil(ext / size) * size;
        const float q = whole / size;
        if (!(q < 2147483648.0f)) return false;
        g.origin[k] = lo[k];
        g.res[k] = int32_t(q);
        count *= uint64_t(g.res[k]);
        if (count >= (1ull << 31)) return false;
    }
    g.size = size;
    g.count = uint32_t(count);
    return true;
}

// cells a line may visit before the build stops with an error (a radius or line length far
// above the voxel size)
constexpr uint32_t kVoxelMaxVisitsPerLine = 1u << 16;

hipError_t voxel_line_bounds(const float4* lines, uint32_t n, int32_t* keys /*device, 6*/, hipStream_t st);
void voxel_decode_bounds(const int32_t keys[6], float lo[3], float hi[3]);
hipError_t voxelize_lines(const float4* lines, uint32_t n, const VoxelGrid& g, float radius, bool reference_sweep,
                          uint32_t* bits, uint32_t* overflow, hipStream_t st);
hipError_t voxel_downsample(const uint32_t* fine, const int32_t fres[3], uint32_t* coarse, const int32_t cres[3],
                            hipStream_t st);
hipError_t voxel_count(const uint32_t* bits, uint32_t nwords, uint32_t* count, hipStream_t st);

}  // namespace vk
