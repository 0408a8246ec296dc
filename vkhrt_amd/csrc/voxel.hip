// Voxel hair kernels: capsule voxelization of line segments (GenerateVoxelMesh,
// source/resources/model/geometry_processor.cpp:515-602), the bounds of a line set, the
// 2x2x2 OR-reduction that makes each coarser level of detail, and the filled-voxel count.
// Compiled with -ffp-contract=off: the float steps round like the reference's host code.
#include <hip/hip_runtime.h>

#include <cstring>

#include "voxel.hpp"

namespace vk {
namespace {

constexpr int kBlock = 256;

// GetMajorAxes (:493-500): axes by decreasing |d|; std::sort on three entries is an insertion
// sort, so equal magnitudes keep the order x, y, z.
__device__ __forceinline__ void major_axes(const float d[3], int a[3]) {
    const float av[3] = {fabsf(d[0]), fabsf(d[1]), fabsf(d[2])};
    a[0] = 0;
    a[1] = 1;
    a[2] = 2;
    for (int i = 1; i < 3; ++i) {
        const int v = a[i];
        int j = i;
        while (j > 0 && av[v] > av[a[j - 1]]) {
            a[j] = a[j - 1];
            --j;
        }
        a[j] = v;
    }
}

// GetVoxelIndex3D (:483-486): floor((p - origin) / s), converted to int; kept within +-2^30 so
// that the loops below cannot overflow (the reference converts unclamped)
__device__ __forceinline__ int cell_of(float p, float origin, float size) {
    const float f = floorf((p - origin) / size);
    return f < -1073741824.0f ? -1073741824 : f > 1073741824.0f ? 1073741824 : int(f);
}

__device__ __forceinline__ void set_bit(uint32_t* bits, uint32_t idx) { atomicOr(&bits[idx >> 5], 1u << (idx & 31u)); }

// One thread per line. Reference sweep: FillVoxel (:502-512) at (minIndex.x, j, k) for the
// box's j, k (:585-593), the index formed in wrapping 32-bit arithmetic and kept when below
// the voxel count. Full sweep: every (i, j, k) of the box inside the grid.
__global__ void k_voxelize(const float4* __restrict__ L, uint32_t n, VoxelGrid g, float r, uint32_t ref_sweep,
                           uint32_t* __restrict__ bits, uint32_t* __restrict__ overflow) {
    const uint32_t li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= n) return;
    const float4 A = L[size_t(li) * 2], B = L[size_t(li) * 2 + 1];
    const float st[3] = {A.x, A.y, A.z}, en[3] = {B.x, B.y, B.z};
    const float d[3] = {en[0] - st[0], en[1] - st[1], en[2] - st[2]};
    int a[3];
    major_axes(d, a);
    const bool flip = d[a[0]] < 0.0f;  // the major axis runs positive (:544-546)
    float s[3], vr0[3], vr1[3];
    for (int k = 0; k < 3; ++k) {
        s[k] = d[k] / d[a[0]];
        const float sr = s[k] * r;  // capsule ends (:551-554)
        vr0[k] = (flip ? en[k] : st[k]) - sr;
        vr1[k] = (flip ? st[k] : en[k]) + sr;
    }
    // projected capsule radius on the minor axes (:556-559)
    const float dn = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const float q1 = d[a[1]] / dn, q2 = d[a[2]] / dn;
    const float r1 = r / sqrtf(1.0f - q1 * q1), r2 = r / sqrtf(1.0f - q2 * q2);
    const float tmin = vr0[a[0]], tmax = vr1[a[0]];
    float t0 = tmin;
    float p0[3] = {vr0[0], vr0[1], vr0[2]};
    const uint32_t rx = uint32_t(g.res[0]), ry = uint32_t(g.res[1]), rz = uint32_t(g.res[2]);
    uint32_t visits = 0;
    while (t0 < tmax) {
        const float tn = t0 + g.size;
        const float t1 = tn < tmax ? tn : tmax;
        if (!(t1 > t0)) break;  // no progress in float: the reference would not end
        float lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
            const float p1 = vr0[k] + s[k] * (t1 - tmin);
            lo[k] = p1 < p0[k] ? p1 : p0[k];
            hi[k] = p0[k] < p1 ? p1 : p0[k];
            p0[k] = p1;
        }
        lo[a[1]] -= r1;
        lo[a[2]] -= r2;
        hi[a[1]] += r1;
        hi[a[2]] += r2;
        int mi[3], ma[3];
        for (int k = 0; k < 3; ++k) {
            mi[k] = cell_of(lo[k], g.origin[k], g.size);
            ma[k] = cell_of(hi[k], g.origin[k], g.size);
        }
        if (ref_sweep) {
            for (int j = mi[1]; j <= ma[1]; ++j)
                for (int k = mi[2]; k <= ma[2]; ++k) {
                    if (++visits > kVoxelMaxVisitsPerLine) {
                        atomicAdd(overflow, 1u);
                        return;
                    }
                    const uint32_t idx = uint32_t(mi[0]) + uint32_t(j) * rx + uint32_t(k) * (rx * ry);
                    if (idx < g.count) set_bit(bits, idx);
                }
        } else {
            const int i0 = max(mi[0], 0), i1 = min(ma[0], int(rx) - 1);
            const int j0 = max(mi[1], 0), j1 = min(ma[1], int(ry) - 1);
            const int k0 = max(mi[2], 0), k1 = min(ma[2], int(rz) - 1);
            for (int k = k0; k <= k1; ++k)
                for (int j = j0; j <= j1; ++j)
                    for (int i = i0; i <= i1; ++i) {
                        if (++visits > kVoxelMaxVisitsPerLine) {
                            atomicAdd(overflow, 1u);
                            return;
                        }
                        set_bit(bits, uint32_t(i) + uint32_t(j) * rx + uint32_t(k) * (rx * ry));
                    }
        }
        t0 = t1;
    }
}

// float -> int32 key whose signed order is the float order
__device__ __forceinline__ int32_t order_key(float f) {
    const int32_t b = __float_as_int(f);
    return b >= 0 ? b : b ^ 0x7FFFFFFF;
}

// min / max of every line end point; NaN coordinates are skipped
__global__ void k_line_bounds(const float4* __restrict__ L, uint32_t n, int32_t* __restrict__ keys) {
    __shared__ int32_t sm[6][kBlock];
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < 2 * n; i += gridDim.x * blockDim.x) {
        const float4 p = L[i];
        const float c[3] = {p.x, p.y, p.z};
        for (int k = 0; k < 3; ++k) {
            if (c[k] < lo[k]) lo[k] = c[k];
            if (c[k] > hi[k]) hi[k] = c[k];
        }
    }
    for (int k = 0; k < 3; ++k) {
        sm[k][threadIdx.x] = order_key(lo[k]);
        sm[3 + k][threadIdx.x] = order_key(hi[k]);
    }
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w)
            for (int k = 0; k < 3; ++k) {
                sm[k][threadIdx.x] = min(sm[k][threadIdx.x], sm[k][threadIdx.x + w]);
                sm[3 + k][threadIdx.x] = max(sm[3 + k][threadIdx.x], sm[3 + k][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) {
            atomicMin(&keys[k], sm[k][0]);
            atomicMax(&keys[3 + k], sm[3 + k][0]);
        }
}

__global__ void k_bounds_init(int32_t* keys) {
    if (threadIdx.x < 3) keys[threadIdx.x] = order_key(INFINITY);
    else if (threadIdx.x < 6) keys[threadIdx.x] = order_key(-INFINITY);
}

// One thread per coarse word: cell (X, Y, Z) is filled when any fine cell (2X + dx, 2Y + dy,
// 2Z + dz) inside the fine grid is.
__global__ void k_downsample(const uint32_t* __restrict__ fine, int3 fr, uint32_t* __restrict__ coarse, int3 cr,
                             uint32_t nwords) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwords) return;
    const uint32_t ccount = uint32_t(cr.x) * uint32_t(cr.y) * uint32_t(cr.z);
    uint32_t word = 0;
    for (uint32_t b = 0; b < 32; ++b) {
        const uint32_t c = w * 32 + b;
        if (c >= ccount) break;
        const uint32_t X = c % uint32_t(cr.x), Y = (c / uint32_t(cr.x)) % uint32_t(cr.y),
                       Z = c / (uint32_t(cr.x) * uint32_t(cr.y));
        bool any = false;
        for (uint32_t dz = 0; dz < 2 && !any; ++dz)
            for (uint32_t dy = 0; dy < 2 && !any; ++dy)
                for (uint32_t dx = 0; dx < 2 && !any; ++dx) {
                    const uint32_t x = 2 * X + dx, y = 2 * Y + dy, z = 2 * Z + dz;
                    if (x >= uint32_t(fr.x) || y >= uint32_t(fr.y) || z >= uint32_t(fr.z)) continue;
                    const uint32_t f = x + y * uint32_t(fr.x) + z * (uint32_t(fr.x) * uint32_t(fr.y));
                    any = (fine[f >> 5] >> (f & 31u)) & 1u;
                }
        if (any) word |= 1u << b;
    }
    coarse[w] = word;
}

__global__ void k_popcount(const uint32_t* __restrict__ bits, uint32_t nwords, uint32_t* __restrict__ count) {
    __shared__ uint32_t sm[kBlock];
    uint32_t c = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += gridDim.x * blockDim.x)
        c += __popc(bits[i]);
    sm[threadIdx.x] = c;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(count, sm[0]);
}

inline dim3 grid_for(uint64_t n, uint32_t cap = 0xFFFFFFFFu) {
    uint64_t b = (n + kBlock - 1) / kBlock;
    return dim3(uint32_t(b < cap ? b : cap));
}

}  // namespace

hipError_t voxel_line_bounds(const float4* lines, uint32_t n, int32_t* keys, hipStream_t st) {
    hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(64), 0, st, keys);
    if (n) hipLaunchKernelGGL(k_line_bounds, grid_for(2ull * n, 1024), dim3(kBlock), 0, st, lines, n, keys);
    return hipGetLastError();
}

void voxel_decode_bounds(const int32_t keys[6], float lo[3], float hi[3]) {
    for (int k = 0; k < 6; ++k) {
        int32_t b = keys[k];
        b = b >= 0 ? b : b ^ 0x7FFFFFFF;
        float f;
        std::memcpy(&f, &b, 4);
        (k < 3 ? lo[k] : hi[k - 3]) = f;
    }
}

hipError_t voxelize_lines(const float4* lines, uint32_t n, const VoxelGrid& g, float radius, bool reference_sweep,
                          uint32_t* bits, uint32_t* overflow, hipStream_t st) {
    if (n == 0 || g.count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_voxelize, grid_for(n), dim3(kBlock), 0, st, lines, n, g, radius, reference_sweep ? 1u : 0u,
                       bits, overflow);
    return hipGetLastError();
}

hipError_t voxel_downsample(const uint32_t* fine, const int32_t fres[3], uint32_t* coarse, const int32_t cres[3],
                            hipStream_t st) {
    const uint64_t ccount = uint64_t(cres[0]) * uint64_t(cres[1]) * uint64_t(cres[2]);
    const uint32_t nwords = uint32_t((ccount + 31) / 32);
    if (nwords == 0) return hipSuccess;
    hipLaunchKernelGGL(k_downsample, grid_for(nwords), dim3(kBlock), 0, st, fine, make_int3(fres[0], fres[1], fres[2]),
                       coarse, make_int3(cres[0], cres[1], cres[2]), nwords);
    return hipGetLastError();
}

hipError_t voxel_count(const uint32_t* bits, uint32_t nwords, uint32_t* count, hipStream_t st) {
    if (nwords == 0) return hipSuccess;
    hipLaunchKernelGGL(k_popcount, grid_for(nwords, 1024), dim3(kBlock), 0, st, bits, nwords, count);
    return hipGetLastError();
}

}  // namespace vk
