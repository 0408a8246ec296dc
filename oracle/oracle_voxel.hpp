// Oracle restatement of the voxel hair technique (test infrastructure only).
//  - VoxelGridFor / Voxelize: GenerateVoxelMesh (source/resources/model/geometry_processor.cpp:
//    515-602) with NextDivisible (:472-476), GetVoxelIndex3D (:483-486), GetVoxelIndex1D
//    (:488-491), GetMajorAxes (:493-500) and FillVoxel (:502-512), in float like the reference.
//    The reference sweep fills (minIndex.x, j, k) only (:585-593) with the index formed in
//    wrapping 32-bit arithmetic; the full sweep fills the whole box inside the grid. Two guards
//    the reference lacks: a step that makes no progress in float ends the line, and a line
//    visiting more than 65536 cells fails the build.
//  - Downsample: 2x2x2 OR-reduction (a level of detail; the reference has none yet, README.md:18).
//  - VoxelWalk: the first filled cell a 3D DDA (Amanatides & Woo 1987) enters within
//    [tmin, tmax]; cell boundaries origin + i * size as the reference's voxel AABBs
//    (GetVoxelWorldPosition, :438-449). The reference reports every candidate box at t = 1.0
//    (voxel_intersection.rint:14), so its hit depends on the traversal order: parity unpinned.
//  - VoxelBrute: the filled cell with the smallest entry t (ties: lowest index) by testing every
//    cell's slab; it pins VoxelWalk.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace orc {

struct VGrid {
    float origin[3] = {0, 0, 0};
    float size = 0.0f;
    int32_t res[3] = {0, 0, 0};
    uint32_t count = 0;
};

inline bool VoxelGridFor(const float lo[3], const float hi[3], float size, VGrid& g) {
    if (!(size > 0.0f) || !std::isfinite(size)) return false;
    uint64_t count = 1;
    for (int k = 0; k < 3; ++k) {
        float e = hi[k] - lo[k];
        if (!std::isfinite(e) || !std::isfinite(lo[k]) || e < 0.0f) return false;
        e = std::ceil(e / size) * size;  // NextDivisible
        const float q = e / size;
        if (!(q < 2147483648.0f)) return false;
        g.res[k] = static_cast<int32_t>(q);
        g.origin[k] = lo[k];
        count *= static_cast<uint64_t>(g.res[k]);
        if (count >= (1ull << 31)) return false;
    }
    g.size = size;
    g.count = static_cast<uint32_t>(count);
    return true;
}

inline int VoxelCell(float p, float origin, float size) {
    const float f = std::floor((p - origin) / size);
    if (f < -1073741824.0f) return -1073741824;
    if (f > 1073741824.0f) return 1073741824;
    return static_cast<int>(f);
}

inline void SetVoxel(std::vector<uint32_t>& bits, uint32_t i) { bits[i >> 5] |= 1u << (i & 31u); }

// lines: n x 8 floats (start.xyz, -, end.xyz, -); bits: (count + 31) / 32 words, zeroed.
inline bool Voxelize(const float* lines, uint32_t n, const VGrid& g, float r, bool reference_sweep,
                     std::vector<uint32_t>& bits) {
    const uint32_t rx = static_cast<uint32_t>(g.res[0]), ry = static_cast<uint32_t>(g.res[1]),
                   rz = static_cast<uint32_t>(g.res[2]);
    if (g.count == 0) return true;
    for (uint32_t li = 0; li < n; ++li) {
        const float* L = lines + static_cast<size_t>(li) * 8;
        const float a[3] = {L[0], L[1], L[2]}, b[3] = {L[4], L[5], L[6]};
        float d[3];
        for (int k = 0; k < 3; ++k) d[k] = b[k] - a[k];
        // GetMajorAxes: a stable order by decreasing |d|
        int ax[3] = {0, 1, 2};
        const float m[3] = {std::fabs(d[0]), std::fabs(d[1]), std::fabs(d[2])};
        for (int i = 1; i < 3; ++i)
            for (int j = i; j > 0 && m[ax[j]] > m[ax[j - 1]]; --j) std::swap(ax[j], ax[j - 1]);
        const bool swap_ends = d[ax[0]] < 0.0f;
        const float* v0 = swap_ends ? b : a;
        const float* v1 = swap_ends ? a : b;
        float step[3], e0[3], e1[3];
        for (int k = 0; k < 3; ++k) {
            step[k] = d[k] / d[ax[0]];
            const float ext = step[k] * r;
            e0[k] = v0[k] - ext;
            e1[k] = v1[k] + ext;
        }
        const float len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        const float c1 = d[ax[1]] / len, c2 = d[ax[2]] / len;
        const float rad1 = r / std::sqrt(1.0f - c1 * c1), rad2 = r / std::sqrt(1.0f - c2 * c2);
        const float ta = e0[ax[0]], tb = e1[ax[0]];
        float t = ta;
        float p[3] = {e0[0], e0[1], e0[2]};
        uint32_t visits = 0;
        while (t < tb) {
            const float next = t + g.size;
            const float t1 = next < tb ? next : tb;
            if (!(t1 > t)) break;
            float lo[3], hi[3];
            for (int k = 0; k < 3; ++k) {
                const float q = e0[k] + step[k] * (t1 - ta);
                lo[k] = q < p[k] ? q : p[k];
                hi[k] = p[k] < q ? q : p[k];
                p[k] = q;
            }
            lo[ax[1]] -= rad1;
            lo[ax[2]] -= rad2;
            hi[ax[1]] += rad1;
            hi[ax[2]] += rad2;
            int c0[3], c9[3];
            for (int k = 0; k < 3; ++k) {
                c0[k] = VoxelCell(lo[k], g.origin[k], g.size);
                c9[k] = VoxelCell(hi[k], g.origin[k], g.size);
            }
            if (reference_sweep) {
                for (int j = c0[1]; j <= c9[1]; ++j)
                    for (int k = c0[2]; k <= c9[2]; ++k) {
                        if (++visits > 65536u) return false;
                        const uint32_t i = static_cast<uint32_t>(c0[0]) + static_cast<uint32_t>(j) * rx +
                                           static_cast<uint32_t>(k) * (rx * ry);
                        if (i < g.count) SetVoxel(bits, i);
                    }
            } else {
                for (int z = std::max(c0[2], 0); z <= std::min(c9[2], static_cast<int>(rz) - 1); ++z)
                    for (int y = std::max(c0[1], 0); y <= std::min(c9[1], static_cast<int>(ry) - 1); ++y)
                        for (int x = std::max(c0[0], 0); x <= std::min(c9[0], static_cast<int>(rx) - 1); ++x) {
                            if (++visits > 65536u) return false;
                            SetVoxel(bits, static_cast<uint32_t>(x) + static_cast<uint32_t>(y) * rx +
                                               static_cast<uint32_t>(z) * (rx * ry));
                        }
            }
            t = t1;
        }
    }
    return true;
}

inline bool VoxelAt(const std::vector<uint32_t>& bits, uint32_t i) { return (bits[i >> 5] >> (i & 31u)) & 1u; }

inline std::vector<uint32_t> Downsample(const std::vector<uint32_t>& fine, const int32_t fr[3], int32_t cr[3]) {
    for (int k = 0; k < 3; ++k) cr[k] = (fr[k] + 1) / 2;
    const uint64_t cc = uint64_t(cr[0]) * uint64_t(cr[1]) * uint64_t(cr[2]);
    std::vector<uint32_t> out((cc + 31) / 32, 0u);
    for (int32_t z = 0; z < fr[2]; ++z)
        for (int32_t y = 0; y < fr[1]; ++y)
            for (int32_t x = 0; x < fr[0]; ++x) {
                const uint32_t f = uint32_t(x) + uint32_t(y) * uint32_t(fr[0]) + uint32_t(z) * uint32_t(fr[0]) * uint32_t(fr[1]);
                if (!VoxelAt(fine, f)) continue;
                SetVoxel(out, uint32_t(x / 2) + uint32_t(y / 2) * uint32_t(cr[0]) +
                                  uint32_t(z / 2) * uint32_t(cr[0]) * uint32_t(cr[1]));
            }
    return out;
}

struct VoxelHit {
    float t;
    uint32_t cell;
    float n[3];
};

// Returns true with the first filled cell entered within [tmin, tmax].
inline bool VoxelWalk(const std::vector<uint32_t>& bits, const float O[3], float c, const int32_t R[3], const float o[3],
                      const float d[3], float tmin, float tmax, VoxelHit& h) {
    if (R[0] <= 0 || R[1] <= 0 || R[2] <= 0) return false;
    float inv[3];
    float t0 = tmin, t1 = tmax;
    int entry_axis = -1;
    for (int k = 0; k < 3; ++k) {
        const float top = O[k] + float(R[k]) * c;
        if (d[k] == 0.0f) {
            if (o[k] < O[k] || o[k] > top) return false;
            inv[k] = 0.0f;
            continue;
        }
        inv[k] = 1.0f / d[k];
        const float ta = (O[k] - o[k]) * inv[k], tb = (top - o[k]) * inv[k];
        const float near = ta < tb ? ta : tb, far = ta < tb ? tb : ta;
        if (near > t0) {
            t0 = near;
            entry_axis = k;
        }
        if (far < t1) t1 = far;
    }
    if (!(t0 <= t1)) return false;
    int cell[3], dir[3];
    float tnext[3];
    for (int k = 0; k < 3; ++k) {
        const float f = std::floor((o[k] + d[k] * t0 - O[k]) / c);
        cell[k] = f < 0.0f ? 0 : f > float(R[k] - 1) ? R[k] - 1 : static_cast<int>(f);
        dir[k] = d[k] > 0.0f ? 1 : d[k] < 0.0f ? -1 : 0;
        tnext[k] = dir[k] == 0 ? INFINITY : (O[k] + float(cell[k] + (dir[k] > 0 ? 1 : 0)) * c - o[k]) * inv[k];
    }
    float tenter = t0;
    for (int guard = R[0] + R[1] + R[2] + 3; guard > 0; --guard) {
        const uint32_t i = uint32_t(cell[0]) + uint32_t(cell[1]) * uint32_t(R[0]) +
                           uint32_t(cell[2]) * (uint32_t(R[0]) * uint32_t(R[1]));
        if (VoxelAt(bits, i)) {
            int a = entry_axis;
            if (a < 0) {
                a = 0;
                if (std::fabs(d[1]) > std::fabs(d[a])) a = 1;
                if (std::fabs(d[2]) > std::fabs(d[a])) a = 2;
            }
            h.t = tenter;
            h.cell = i;
            h.n[0] = h.n[1] = h.n[2] = 0.0f;
            h.n[a] = d[a] > 0.0f ? -1.0f : 1.0f;
            return true;
        }
        int a = 0;
        if (tnext[1] < tnext[a]) a = 1;
        if (tnext[2] < tnext[a]) a = 2;
        if (!(tnext[a] <= t1)) return false;
        tenter = tnext[a] > t0 ? tnext[a] : t0;
        entry_axis = a;
        cell[a] += dir[a];
        if (cell[a] < 0 || cell[a] >= R[a]) return false;
        tnext[a] = (O[a] + float(cell[a] + (dir[a] > 0 ? 1 : 0)) * c - o[a]) * inv[a];
    }
    return false;
}

// Every filled cell's slab [lo, hi] against [tmin, tmax] in double; the smallest
// max(entry, tmin), ties to the lowest index.
inline bool VoxelBrute(const std::vector<uint32_t>& bits, const float O[3], float c, const int32_t R[3], const float o[3],
                       const float d[3], float tmin, float tmax, VoxelHit& h) {
    bool found = false;
    double best = 0.0;
    const uint64_t n = uint64_t(R[0]) * uint64_t(R[1]) * uint64_t(R[2]);
    for (uint64_t i = 0; i < n; ++i) {
        if (!VoxelAt(bits, uint32_t(i))) continue;
        const int64_t ci[3] = {int64_t(i % uint64_t(R[0])), int64_t((i / uint64_t(R[0])) % uint64_t(R[1])),
                               int64_t(i / (uint64_t(R[0]) * uint64_t(R[1])))};
        double t0 = tmin, t1 = tmax;
        for (int k = 0; k < 3 && t0 <= t1; ++k) {
            const double lo = double(O[k] + float(ci[k]) * c), hi = double(O[k] + float(ci[k] + 1) * c);
            if (d[k] == 0.0f) {
                if (o[k] < lo || o[k] > hi) t0 = INFINITY;
                continue;
            }
            const double ta = (lo - o[k]) / d[k], tb = (hi - o[k]) / d[k];
            t0 = std::max(t0, std::min(ta, tb));
            t1 = std::min(t1, std::max(ta, tb));
        }
        if (!(t0 <= t1)) continue;
        if (!found || t0 < best) {
            found = true;
            best = t0;
            h.t = float(t0);
            h.cell = uint32_t(i);
        }
    }
    return found;
}

}  // namespace orc
