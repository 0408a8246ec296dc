// GPU builder for the bottom-level structure (replaces the driver-internal build behind
// vkCmdBuildAccelerationStructuresKHR, bottom_level_acceleration_structure.cpp:34-78).
//
// Pipeline (all on one stream, temporaries freed at the end):
//   1. primitive boxes (GenerateAABBs semantics, geometry_processor.cpp:422-436, plus a
//      small conservative pad) and centroid bounds (wave reduction + ordered-int atomics)
//   2. 63-bit Morton codes of box centroids
//   3. stable LSD radix sort of (key, index), 8 bits per pass: LDS histogram, one-block
//      scan, LDS-ranked stable scatter (wave ballots rank equal digits)
//   4. Karras (HPG 2012) binary radix tree, one thread per internal node
//   5. bottom-up box refit in level-synchronous rounds (one launch per tree level)
//   6. top-down collapse to 8-wide nodes, one level per launch, by one of two rules:
//      SAH: the refit rounds also compute, for every binary node, the least SAH cost
//      of covering its subtree with at most j child slots (j = 1..width), and each wide node is
//      opened along those choices (Ylitie, Karras, Laine, HPG 2017);
//      greedy (default, fewer node visits per ray on hair): each binary subtree is opened by
//      largest surface area until `width` (<= 8)
//      children, and subtrees with at most `leaf_max` primitives become leaf children.
//      Child slots are ordered per ray octant; child boxes are quantized to 8 bits against
//      power-of-two scales (outward rounding)
//   7. gather primitive records into leaf order
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "build.hpp"
#include "device_math.hpp"
#include "internal.hpp"

namespace vk {

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kLeafFlag = 0x80000000u;

__device__ __forceinline__ uint32_t f2ord(float f) {
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
    uint32_t b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    return __uint_as_float(b);
}

__global__ void k_init_bounds(uint32_t* cb) {
    if (threadIdx.x < 3) {
        cb[threadIdx.x] = 0xFFFFFFFFu;  // min (ordered)
        cb[3 + threadIdx.x] = 0u;       // max (ordered)
    }
}

// Conservative pad added to every primitive box: 1e-4 + 4e-7 * max|coordinate|.
__device__ __forceinline__ float box_pad(float lo[3], float hi[3]) {
    float m = 0.0f;
    for (int k = 0; k < 3; ++k) m = fmaxf(m, fmaxf(fabsf(lo[k]), fabsf(hi[k])));
    return 1e-4f + 4e-7f * m;
}

// Grid-stride over the records (a fixed grid), so the centroid bounds take one set of atomics
// per block: per-wave atomics on the same 6 words serialized large builds.
__global__ __launch_bounds__(kBlock) void k_prim_boxes(const float4* __restrict__ prims, uint32_t n, int mode,
                                                       int radius_mode, float radius, float4* __restrict__ blo,
                                                       float4* __restrict__ bhi, uint32_t* __restrict__ cbounds) {
    float cmn[3] = {INFINITY, INFINITY, INFINITY}, cmx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    float c[3] = {0, 0, 0};
    {
        if (mode == kPhantom) {
            float4 p[4];
            for (int k = 0; k < 4; ++k) p[k] = prims[size_t(i) * 4 + k];
            float r = radius;
            if (radius_mode == 1) r = gmax(gmax(p[0].w, p[1].w), gmax(p[2].w, p[3].w));
            const float px[4] = {p[0].x, p[1].x, p[2].x, p[3].x};
            const float py[4] = {p[0].y, p[1].y, p[2].y, p[3].y};
            const float pz[4] = {p[0].z, p[1].z, p[2].z, p[3].z};
            const float* pc[3] = {px, py, pz};
            for (int k = 0; k < 3; ++k) {
                // glm::min(glm::min(start, end), glm::min(cp1, cp2)) - r
                lo[k] = gmin(gmin(pc[k][0], pc[k][3]), gmin(pc[k][1], pc[k][2])) - r;
                hi[k] = gmax(gmax(pc[k][0], pc[k][3]), gmax(pc[k][1], pc[k][2])) + r;
            }
        } else if (mode == kLss) {
            float4 a = prims[size_t(i) * 2], b = prims[size_t(i) * 2 + 1];
            float ra = fabsf(a.w), rb = b.w;
            lo[0] = fminf(a.x - ra, b.x - rb);
            lo[1] = fminf(a.y - ra, b.y - rb);
            lo[2] = fminf(a.z - ra, b.z - rb);
            hi[0] = fmaxf(a.x + ra, b.x + rb);
            hi[1] = fmaxf(a.y + ra, b.y + rb);
            hi[2] = fmaxf(a.z + ra, b.z + rb);
        } else if (mode == kTriangles) {
            for (int v = 0; v < 3; ++v) {
                const float4 q = prims[size_t(i) * 3 + v];
                const float a[3] = {q.x, q.y, q.z};
                for (int k = 0; k < 3; ++k) {
                    lo[k] = fminf(lo[k], a[k]);
                    hi[k] = fmaxf(hi[k], a[k]);
                }
            }
        } else {
            float4 P[4];
            for (int k = 0; k < 4; ++k) P[k] = prims[size_t(i) * 4 + k];
            // the 6 distinct vertices of the two quads: start +- v r0, end +- v r1
            for (int f = 0; f < 2; ++f) {
                float4 v = P[2 + f];
                for (int sgn = -1; sgn <= 1; sgn += 2) {
                    float s0 = float(sgn) * P[0].w, s1 = float(sgn) * P[1].w;
                    float a[3] = {P[0].x + v.x * s0, P[0].y + v.y * s0, P[0].z + v.z * s0};
                    float b[3] = {P[1].x + v.x * s1, P[1].y + v.y * s1, P[1].z + v.z * s1};
                    for (int k = 0; k < 3; ++k) {
                        lo[k] = fminf(lo[k], fminf(a[k], b[k]));
                        hi[k] = fmaxf(hi[k], fmaxf(a[k], b[k]));
                    }
                }
            }
        }
        float pad = box_pad(lo, hi);
        for (int k = 0; k < 3; ++k) {
            lo[k] -= pad;
            hi[k] += pad;
            c[k] = 0.5f * lo[k] + 0.5f * hi[k];
        }
        blo[i] = make_float4(lo[0], lo[1], lo[2], 0.0f);
        bhi[i] = make_float4(hi[0], hi[1], hi[2], 0.0f);
    }
    for (int k = 0; k < 3; ++k) {
        cmn[k] = fminf(cmn[k], c[k]);
        cmx[k] = fmaxf(cmx[k], c[k]);
    }
    }
    // centroid bounds: wave reduction, block reduction in LDS, one atomic per block and word
    __shared__ float s_mn[kBlock / 64][3], s_mx[kBlock / 64][3];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = 0; k < 3; ++k) {
        float mn = cmn[k], mx = cmx[k];
        for (int off = 32; off > 0; off >>= 1) {
            mn = fminf(mn, __shfl_xor(mn, off));
            mx = fmaxf(mx, __shfl_xor(mx, off));
        }
        if (lane == 0) {
            s_mn[wave][k] = mn;
            s_mx[wave][k] = mx;
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int k = threadIdx.x;
        float mn = s_mn[0][k], mx = s_mx[0][k];
        for (int w = 1; w < kBlock / 64; ++w) {
            mn = fminf(mn, s_mn[w][k]);
            mx = fmaxf(mx, s_mx[w][k]);
        }
        if (mn <= mx) {
            atomicMin(&cbounds[k], f2ord(mn));
            atomicMax(&cbounds[3 + k], f2ord(mx));
        }
    }
}

__device__ __forceinline__ uint64_t spread21(uint64_t x) {
    x &= 0x1fffffull;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

__global__ void k_morton(const float4* __restrict__ blo, const float4* __restrict__ bhi, uint32_t n,
                         const uint32_t* __restrict__ cbounds, uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 a = blo[i], b = bhi[i];
    float c[3] = {0.5f * a.x + 0.5f * b.x, 0.5f * a.y + 0.5f * b.y, 0.5f * a.z + 0.5f * b.z};
    uint64_t q[3];
    for (int k = 0; k < 3; ++k) {
        float lo = ord2f(cbounds[k]), hi = ord2f(cbounds[3 + k]);
        float ext = hi - lo;
        float x = ext > 0.0f ? (c[k] - lo) / ext : 0.0f;
        x = fminf(fmaxf(x, 0.0f), 1.0f);
        q[k] = uint64_t(fminf(x * 2097152.0f, 2097151.0f));
    }
    keys[i] = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
    vals[i] = i;
}

// ------------------------------------------------------------------ radix sort
constexpr int kRsItems = 16;
constexpr int kRsTile = kBlock * kRsItems;

__global__ void k_rs_hist(const uint64_t* __restrict__ keys, uint32_t n, int shift, uint32_t* __restrict__ hist,
                          uint32_t nblocks) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint32_t base = blockIdx.x * kRsTile;
    for (int it = 0; it < kRsItems; ++it) {
        uint32_t idx = base + it * kBlock + threadIdx.x;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of `count` u32 in place with one block of 1024 threads.
__global__ void k_scan_single(uint32_t* __restrict__ data, uint32_t count) {
    __shared__ uint32_t part[1024];
    uint32_t per = (count + 1023) / 1024;
    uint32_t b = threadIdx.x * per, e = min(count, b + per);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += data[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint32_t v = threadIdx.x >= uint32_t(off) ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - s;
    for (uint32_t i = b; i < e; ++i) {
        uint32_t v = data[i];
        data[i] = run;
        run += v;
    }
}

// Exclusive prefix sum over `count` words in three launches: tile sums (4096 words per block,
// coalesced uint4 loads), one block scans the sums, then each block scans its tile from its
// offset.
constexpr uint32_t kScanTile = 4096;
__device__ __forceinline__ uint4 load4(const uint32_t* d, uint32_t i, uint32_t count) {
    if (i + 3 < count && (i & 3u) == 0u) return *reinterpret_cast<const uint4*>(d + i);
    return make_uint4(i < count ? d[i] : 0u, i + 1 < count ? d[i + 1] : 0u, i + 2 < count ? d[i + 2] : 0u,
                      i + 3 < count ? d[i + 3] : 0u);
}
__global__ __launch_bounds__(1024) void k_scan_tiles(const uint32_t* __restrict__ data, uint32_t count,
                                                     uint32_t* __restrict__ sums) {
    __shared__ uint32_t s_w[16];
    const uint4 v = load4(data, blockIdx.x * kScanTile + threadIdx.x * 4u, count);
    uint32_t t = v.x + v.y + v.z + v.w;
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t sum = 0;
        for (int w = 0; w < 16; ++w) sum += s_w[w];
        sums[blockIdx.x] = sum;
    }
}
__global__ __launch_bounds__(1024) void k_scan_apply(uint32_t* __restrict__ data, uint32_t count,
                                                     const uint32_t* __restrict__ sums) {
    __shared__ uint32_t s_part[1024];
    const uint32_t i = blockIdx.x * kScanTile + threadIdx.x * 4u;
    const uint4 v = load4(data, i, count);
    const uint32_t t = v.x + v.y + v.z + v.w;
    s_part[threadIdx.x] = t;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint32_t add = threadIdx.x >= off ? s_part[threadIdx.x - off] : 0u;
        __syncthreads();
        s_part[threadIdx.x] += add;
        __syncthreads();
    }
    uint32_t run = sums[blockIdx.x] + s_part[threadIdx.x] - t;
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
    for (uint32_t k = 0; k < 4; ++k)
        if (i + k < count) {
            data[i + k] = run;
            run += w[k];
        }
}

__global__ void k_rs_scatter(const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                             uint64_t* __restrict__ kout, uint32_t* __restrict__ vout, uint32_t n, int shift,
                             const uint32_t* __restrict__ offs, uint32_t nblocks) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[kBlock / 64][256];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    run[tid] = offs[tid * nblocks + blockIdx.x];
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t base = blockIdx.x * kRsTile;
    for (int it = 0; it < kRsItems; ++it) {
        for (int k = 0; k < kBlock / 64; ++k) wcnt[k][tid] = 0;
        __syncthreads();
        uint32_t idx = base + it * kBlock + tid;
        bool valid = idx < n;
        uint64_t key = valid ? kin[idx] : 0ull;
        uint32_t digit = uint32_t(key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < 8; ++b) {
            uint64_t m = __ballot((digit >> b) & 1u);
            peers &= ((digit >> b) & 1u) ? m : ~m;
        }
        uint32_t rank = __popcll(peers & lt);
        if (valid && rank == 0) wcnt[w][digit] = __popcll(peers);
        __syncthreads();
        {
            uint32_t acc = run[tid];
            for (int k = 0; k < kBlock / 64; ++k) {
                uint32_t c = wcnt[k][tid];
                wcnt[k][tid] = acc;
                acc += c;
            }
            run[tid] = acc;
        }
        __syncthreads();
        if (valid) {
            uint32_t pos = wcnt[w][digit] + rank;
            kout[pos] = key;
            vout[pos] = vin[idx];
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ Karras LBVH
__device__ __forceinline__ int delta(const uint64_t* k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint64_t a = k[i], b = k[j];
    if (a == b) return 64 + __clz(uint32_t(i ^ j));
    return __clzll(a ^ b);
}

__global__ void k_karras(const uint64_t* __restrict__ keys, int n, uint2* __restrict__ child,
                         uint2* __restrict__ range, uint32_t* __restrict__ parent_int, uint32_t* __restrict__ parent_leaf) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) > 0 ? 1 : -1;
    int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta(keys, n, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int gamma = i + s * d + min(d, 0);
    int lo = min(i, j), hi = max(i, j);
    uint32_t left = (lo == gamma) ? (kLeafFlag | uint32_t(gamma)) : uint32_t(gamma);
    uint32_t right = (hi == gamma + 1) ? (kLeafFlag | uint32_t(gamma + 1)) : uint32_t(gamma + 1);
    child[i] = make_uint2(left, right);
    range[i] = make_uint2(uint32_t(lo), uint32_t(hi));
    if (left & kLeafFlag) parent_leaf[left & ~kLeafFlag] = uint32_t(i); else parent_int[left] = uint32_t(i);
    if (right & kLeafFlag) parent_leaf[right & ~kLeafFlag] = uint32_t(i); else parent_int[right] = uint32_t(i);
    if (i == 0) parent_int[0] = 0xFFFFFFFFu;
}

__device__ __forceinline__ void child_box(uint32_t ref, const float4* blo, const float4* bhi, const uint32_t* vals,
                                          const float* nlo, const float* nhi, float lo[3], float hi[3]) {
    if (ref & kLeafFlag) {
        uint32_t p = vals[ref & ~kLeafFlag];
        float4 a = blo[p], b = bhi[p];
        lo[0] = a.x;
        lo[1] = a.y;
        lo[2] = a.z;
        hi[0] = b.x;
        hi[1] = b.y;
        hi[2] = b.z;
    } else {
        for (int k = 0; k < 3; ++k) {
            lo[k] = nlo[ref * 3 + k];
            hi[k] = nhi[ref * 3 + k];
        }
    }
}

__device__ __forceinline__ float box_area(const float lo[3], const float hi[3]) {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return dx * dy + dy * dz + dz * dx;
}

// SAH-optimal collapse (Ylitie, Karras, Laine, "Efficient Incoherent Ray Traversal on GPUs
// Through Compressed Wide BVHs", HPG 2017). Every binary node n gets the least SAH cost C(n, j)
// of covering its subtree with at most j child slots of a wide node (j = 1..width):
//   C(n, 1) = min(as a leaf slot: A(n) cnt(n) c_prim, if cnt(n) <= leaf_max;
//                 as a wide node: A(n) c_node + D(n, width))
//   C(n, j) = min(C(n, j - 1), D(n, j)),   D(n, j) = min over k of C(left, k) + C(right, j - k)
// with A the half surface area of the box; a binary leaf (one primitive) costs A c_prim for any
// j. The choices are recorded, and the top-down pass opens each wide node along them.
struct SahCost {
    float c[8];      // C(n, j), j = 1..width
    uint8_t dec[8];  // dec[0]: 1 leaf slot, 0 wide node; dec[j - 1] (j >= 2): 0 = C(n, j - 1), else the k of D(n, j)
    uint32_t kint;   // the k of D(n, width): slots of the wide node rooted at n from the left child
    uint32_t pad;
};
struct SahParams {
    SahCost* cost;  // null: greedy collapse
    const uint2* range;
    int leaf_max, width;
    float c_node, c_prim;
};

__device__ __forceinline__ float sah_child(const SahParams& sp, uint32_t ref, int j, const float lo[3],
                                           const float hi[3]) {
    return (ref & kLeafFlag) ? box_area(lo, hi) * sp.c_prim : sp.cost[ref].c[j - 1];
}

__device__ __forceinline__ void sah_node(const SahParams& sp, int i, uint2 c, const float lo[3], const float hi[3],
                                         const float alo[3], const float ahi[3], const float blo[3],
                                         const float bhi[3]) {
    const int W = sp.width;
    float cl[8], cr[8], dist[9];
    uint8_t dk[9];
#pragma unroll
    for (int j = 1; j <= 8; ++j) {
        cl[j - 1] = j <= W ? sah_child(sp, c.x, j, alo, ahi) : 0.0f;
        cr[j - 1] = j <= W ? sah_child(sp, c.y, j, blo, bhi) : 0.0f;
    }
    for (int j = 2; j <= W; ++j) {
        float best = INFINITY;
        uint8_t bk = 1;
        for (int k = 1; k < j; ++k) {
            const float v = cl[k - 1] + cr[j - k - 1];
            if (v < best) {
                best = v;
                bk = uint8_t(k);
            }
        }
        dist[j] = best;
        dk[j] = bk;
    }
    const float A = box_area(lo, hi);
    const uint2 rg = sp.range[i];
    const uint32_t cnt = rg.y - rg.x + 1;
    const float c_leaf = cnt <= uint32_t(sp.leaf_max) ? A * float(cnt) * sp.c_prim : INFINITY;
    const float c_int = A * sp.c_node + dist[W];
    SahCost out;
    out.c[0] = fminf(c_leaf, c_int);
    out.dec[0] = c_leaf <= c_int ? 1 : 0;
    for (int j = 2; j <= 8; ++j) {
        if (j <= W && dist[j] < out.c[j - 2]) {
            out.c[j - 1] = dist[j];
            out.dec[j - 1] = dk[j];
        } else {
            out.c[j - 1] = out.c[j - 2];
            out.dec[j - 1] = 0;
        }
    }
    out.kint = dk[W];
    out.pad = 0;
    sp.cost[i] = out;
}

// Level-synchronous refit, one round per launch: an internal node whose children were
// finished in earlier rounds (leaves always are) gets its box and records the round. Rounds =
// tree height. The kernel boundary makes the children's boxes visible, so the rounds need no
// atomics or fences (a per-node climb needs agent-scope release/acquire pairs, which write back
// and invalidate L2 across the XCDs at every step).
// With SAH costs requested, the round also computes the node's costs from its children's.
__global__ void k_refit_round(int nint, const uint2* __restrict__ child, const uint32_t* __restrict__ vals,
                              const float4* __restrict__ blo, const float4* __restrict__ bhi, float* __restrict__ nlo,
                              float* __restrict__ nhi, uint32_t* __restrict__ done, uint32_t round, SahParams sp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nint || done[i]) return;
    const uint2 c = child[i];
    const bool ra = (c.x & kLeafFlag) || (done[c.x] != 0u && done[c.x] < round);
    const bool rb = (c.y & kLeafFlag) || (done[c.y] != 0u && done[c.y] < round);
    if (!(ra && rb)) return;
    float alo[3], ahi[3], blo3[3], bhi3[3];
    child_box(c.x, blo, bhi, vals, nlo, nhi, alo, ahi);
    child_box(c.y, blo, bhi, vals, nlo, nhi, blo3, bhi3);
    float lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = fminf(alo[k], blo3[k]);
        hi[k] = fmaxf(ahi[k], bhi3[k]);
        nlo[i * 3 + k] = lo[k];
        nhi[i * 3 + k] = hi[k];
    }
    if (sp.cost) sah_node(sp, i, c, lo, hi, alo, ahi, blo3, bhi3);
    done[i] = round;
}

// ------------------------------------------------------------------ collapse to 8-wide
struct Task {
    uint32_t bref;   // binary node (internal index or kLeafFlag | sorted index)
    uint32_t wnode;  // wide node to fill
};

__device__ __forceinline__ float area(const float lo[3], const float hi[3]) {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return dx * dy + dy * dz + dz * dx;
}

__device__ __forceinline__ uint32_t fexp_for(float extent) {
    // smallest power of two s with 255 * s >= extent; returns the biased exponent byte
    float x = extent / 255.0f;
    if (!(x > 0.0f)) return 1u;
    int e;
    frexpf(x, &e);  // x = m * 2^e, m in [0.5, 1)  =>  2^e >= x
    int b = e + 127;
    return uint32_t(max(1, min(254, b)));
}

__global__ void k_collapse(const Task* __restrict__ tin, const uint32_t* __restrict__ ntin, Task* __restrict__ tout,
                           uint32_t* __restrict__ ntout, uint32_t* __restrict__ node_counter,
                           uint32_t* __restrict__ prim_counter, int n, int leaf_max, int width,
                           const uint2* __restrict__ child,
                           const uint2* __restrict__ range, const uint32_t* __restrict__ vals,
                           const float4* __restrict__ blo, const float4* __restrict__ bhi, const float* __restrict__ nlo,
                           const float* __restrict__ nhi, WideNode* __restrict__ nodes,
                           uint32_t* __restrict__ prim_slots, SahParams sp) {
    uint32_t ti = blockIdx.x * blockDim.x + threadIdx.x;
    if (ti >= *ntin) return;
    Task task = tin[ti];
    uint32_t ref[8], cnt[8], first[8];
    float lo[8][3], hi[8][3];
    int ne = 0;
    auto load = [&](int e, uint32_t r) {
        ref[e] = r;
        if (r & kLeafFlag) {
            cnt[e] = 1;
            first[e] = r & ~kLeafFlag;
            uint32_t p = vals[first[e]];
            float4 a = blo[p], b = bhi[p];
            lo[e][0] = a.x;
            lo[e][1] = a.y;
            lo[e][2] = a.z;
            hi[e][0] = b.x;
            hi[e][1] = b.y;
            hi[e][2] = b.z;
        } else {
            uint2 rg = range[r];
            cnt[e] = rg.y - rg.x + 1;
            first[e] = rg.x;
            for (int k = 0; k < 3; ++k) {
                lo[e][k] = nlo[r * 3 + k];
                hi[e][k] = nhi[r * 3 + k];
            }
        }
    };
    bool is_int[8];
    if (task.bref & kLeafFlag) {
        load(0, task.bref);
        ne = 1;
    } else if (sp.cost) {
        // the slots recorded by the cost pass: D(root, width) split, each side expanded along
        // its choices (a binary leaf, or one slot, ends the expansion)
        uint2 stk[16];
        int sn = 0;
        const uint2 c = child[task.bref];
        const int k = int(sp.cost[task.bref].kint);
        stk[sn++] = make_uint2(c.y, uint32_t(width - k));
        stk[sn++] = make_uint2(c.x, uint32_t(k));
        while (sn > 0 && ne < 8) {
            const uint2 t = stk[--sn];
            const uint32_t m = t.x;
            const int j = int(t.y);
            if ((m & kLeafFlag) || j <= 1) {
                load(ne, m);
                is_int[ne] = !(m & kLeafFlag) && sp.cost[m].dec[0] == 0;
                ++ne;
                continue;
            }
            const int d = sp.cost[m].dec[j - 1];
            if (d == 0) {
                stk[sn++] = make_uint2(m, uint32_t(j - 1));
            } else if (sn + 2 <= 16) {
                const uint2 cm = child[m];
                stk[sn++] = make_uint2(cm.y, uint32_t(j - d));
                stk[sn++] = make_uint2(cm.x, uint32_t(d));
            }
        }
    } else {
        uint2 c = child[task.bref];
        load(0, c.x);
        load(1, c.y);
        ne = 2;
    }
    while (!sp.cost && ne < width) {
        int best = -1;
        float ba = -1.0f;
        for (int e = 0; e < ne; ++e) {
            if ((ref[e] & kLeafFlag) || cnt[e] <= uint32_t(leaf_max)) continue;
            float a = area(lo[e], hi[e]);
            if (a > ba) {
                ba = a;
                best = e;
            }
        }
        if (best < 0) break;
        uint2 c = child[ref[best]];
        load(best, c.x);
        load(ne, c.y);
        ++ne;
    }
    // node box and centroid
    float nb_lo[3] = {INFINITY, INFINITY, INFINITY}, nb_hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int e = 0; e < ne; ++e)
        for (int k = 0; k < 3; ++k) {
            nb_lo[k] = fminf(nb_lo[k], lo[e][k]);
            nb_hi[k] = fmaxf(nb_hi[k], hi[e][k]);
        }
    // octant slot assignment: slot s should hold the child met first along direction
    // d_s = ((s&1)?-1:1, (s&2)?-1:1, (s&4)?-1:1), i.e. the lowest dot(centroid offset, d_s)
    float off[8][3];
    for (int e = 0; e < ne; ++e)
        for (int k = 0; k < 3; ++k)
            off[e][k] = (0.5f * lo[e][k] + 0.5f * hi[e][k]) - (0.5f * nb_lo[k] + 0.5f * nb_hi[k]);
    int slot_of[8];
    uint32_t used_slots = 0, used_e = 0;
    for (int round = 0; round < ne; ++round) {
        float bc = INFINITY;
        int be = -1, bs = -1;
        for (int e = 0; e < ne; ++e) {
            if (used_e & (1u << e)) continue;
            for (int s = 0; s < 8; ++s) {
                if (used_slots & (1u << s)) continue;
                float c = ((s & 1) ? -off[e][0] : off[e][0]) + ((s & 2) ? -off[e][1] : off[e][1]) +
                          ((s & 4) ? -off[e][2] : off[e][2]);
                if (c < bc) {
                    bc = c;
                    be = e;
                    bs = s;
                }
            }
        }
        if (be < 0) {  // NaN costs: take the first free pair
            for (int e = 0; e < ne && be < 0; ++e)
                if (!(used_e & (1u << e))) be = e;
            for (int s = 0; s < 8 && bs < 0; ++s)
                if (!(used_slots & (1u << s))) bs = s;
        }
        slot_of[be] = bs;
        used_e |= 1u << be;
        used_slots |= 1u << bs;
    }
    // quantization frame
    uint32_t ebyte[3];
    float scale[3];
    for (int k = 0; k < 3; ++k) {
        uint32_t eb = fexp_for(nb_hi[k] - nb_lo[k]);
        for (;;) {
            float s = __uint_as_float(eb << 23);
            if (nb_lo[k] + 255.0f * s >= nb_hi[k] || eb >= 254) break;
            ++eb;
        }
        ebyte[k] = eb;
        scale[k] = __uint_as_float(eb << 23);
    }
    // allocation of children
    int n_int = 0;
    uint32_t n_prims = 0;
    uint32_t slot_entry[8];
    for (int s = 0; s < 8; ++s) slot_entry[s] = 0xFFu;
    for (int e = 0; e < ne; ++e) slot_entry[slot_of[e]] = uint32_t(e);
    for (int e = 0; e < ne; ++e) {
        if (!sp.cost || (task.bref & kLeafFlag)) is_int[e] = !(ref[e] & kLeafFlag) && cnt[e] > uint32_t(leaf_max);
        if (is_int[e])
            ++n_int;
        else
            n_prims += cnt[e];
    }
    uint32_t cbase = n_int ? atomicAdd(node_counter, uint32_t(n_int)) : 0u;
    uint32_t pbase = n_prims ? atomicAdd(prim_counter, n_prims) : 0u;
    // capacity guards (a consistent tree never trips them; the host checks the counters)
    if (cbase + uint32_t(n_int) > uint32_t(n) || pbase + n_prims > uint32_t(n)) return;
    WideNode W;
    W.px = nb_lo[0];
    W.py = nb_lo[1];
    W.pz = nb_lo[2];
    uint32_t imask = 0;
    uint8_t meta[8], ql[3][8], qh[3][8];
    uint32_t rank = 0, poff = 0;
    for (int s = 0; s < 8; ++s) {
        meta[s] = 0;
        for (int k = 0; k < 3; ++k) {
            ql[k][s] = 0;
            qh[k][s] = 0;
        }
        uint32_t e = slot_entry[s];
        if (e == 0xFFu) continue;
        for (int k = 0; k < 3; ++k) {
            float org = nb_lo[k];
            int q0 = int(floorf((lo[e][k] - org) / scale[k]));
            q0 = max(0, min(255, q0));
            while (q0 > 0 && org + float(q0) * scale[k] > lo[e][k]) --q0;
            int q1 = int(ceilf((hi[e][k] - org) / scale[k]));
            q1 = max(0, min(255, q1));
            while (q1 < 255 && org + float(q1) * scale[k] < hi[e][k]) ++q1;
            ql[k][s] = uint8_t(q0);
            qh[k][s] = uint8_t(q1);
        }
        if (is_int[e]) {
            imask |= 1u << s;
            meta[s] = 0x80;
            uint32_t wi = cbase + rank;
            ++rank;
            uint32_t to = atomicAdd(ntout, 1u);
            if (to <= uint32_t(n)) tout[to] = Task{ref[e], wi};
        } else {
            meta[s] = uint8_t((cnt[e] << 5) | poff);
            for (uint32_t j = 0; j < cnt[e]; ++j) prim_slots[pbase + poff + j] = vals[first[e] + j];
            poff += cnt[e];
        }
    }
    W.e_imask = ebyte[0] | (ebyte[1] << 8) | (ebyte[2] << 16) | (imask << 24);
    W.child_base = cbase;
    W.prim_base = pbase;
    auto pack = [](const uint8_t* b, uint32_t* o) {
        o[0] = uint32_t(b[0]) | (uint32_t(b[1]) << 8) | (uint32_t(b[2]) << 16) | (uint32_t(b[3]) << 24);
        o[1] = uint32_t(b[4]) | (uint32_t(b[5]) << 8) | (uint32_t(b[6]) << 16) | (uint32_t(b[7]) << 24);
    };
    pack(meta, W.meta);
    pack(ql[0], W.qlo_x);
    pack(ql[1], W.qlo_y);
    pack(ql[2], W.qlo_z);
    pack(qh[0], W.qhi_x);
    pack(qh[1], W.qhi_y);
    pack(qh[2], W.qhi_z);
    nodes[task.wnode] = W;
}

__global__ void k_gather(const float4* __restrict__ in, const uint32_t* __restrict__ slots, uint32_t n, int rec,
                         float4* __restrict__ out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * uint32_t(rec)) return;
    uint32_t s = i / rec, k = i % rec;
    out[i] = in[size_t(slots[s]) * rec + k];
}

// Early-out cylinder radius of each curve in leaf order (hair_intersection.rint:21-22):
// CurveDistanceToCylinder(curve, SampleCurvePoint(curve, 0.5)) + radius, with the same IEEE
// sequence as the intersector (no contraction), so the per-ray test is unchanged bit for bit.
__global__ void k_phantom_rmax(const float4* __restrict__ prims, uint32_t n, int radius_mode, float radius,
                               float* __restrict__ rmax) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 a = prims[size_t(i) * 4], b = prims[size_t(i) * 4 + 1], c = prims[size_t(i) * 4 + 2],
                 d = prims[size_t(i) * 4 + 3];
    const f3 p0 = f3{a.x, a.y, a.z}, p1 = f3{b.x, b.y, b.z}, p2 = f3{c.x, c.y, c.z}, p3 = f3{d.x, d.y, d.z};
    float rpad = radius;
    if (radius_mode == 1) rpad = gmax(gmax(a.w, b.w), gmax(c.w, d.w));
    // SampleCurvePoint(curve, 0.5)
    const float t = 0.5f, u = 1.0f - t, tt = t * t, uu = u * u, uuu = uu * u, ttt = tt * t;
    const float w1 = 3.0f * uu * t, w2 = 3.0f * u * tt;
    const f3 mid = f3{uuu * p0.x + w1 * p1.x + w2 * p2.x + ttt * p3.x, uuu * p0.y + w1 * p1.y + w2 * p2.y + ttt * p3.y,
                      uuu * p0.z + w1 * p1.z + w2 * p2.z + ttt * p3.z};
    float r = len3(cross3(sub(mid, p0), sub(mid, p3))) / len3(sub(p3, p0));
    r += rpad;
    rmax[i] = r;
}

__global__ void k_set_u32(uint32_t* p, uint32_t v, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_root_task(Task* t, uint32_t* nt, uint32_t* node_counter, uint32_t bref) {
    t[0] = Task{bref, 0u};
    *nt = 1u;
    *node_counter = 1u;
}

template <class T>
hipError_t dalloc(T** p, size_t count) {
    return hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(1, count) * sizeof(T));
}

inline uint32_t blocks(uint64_t n, int b = kBlock) { return uint32_t((n + b - 1) / b); }

}  // namespace

#define VK_TRY(x)                                  \
    do {                                           \
        hipError_t e_ = (x);                       \
        if (e_ != hipSuccess) {                    \
            err = std::string(#x) + ": " + hipGetErrorString(e_); \
            goto fail;                             \
        }                                          \
    } while (0)

hipError_t build_bvh(int mode, const float4* prims_in, uint32_t n, int radius_mode, float radius, int leaf_max,
                     int width, int collapse, hipStream_t st, BuiltBvh* out, std::string& err) {
    const int rec = mode == kLss ? 2 : mode == kTriangles ? 3 : 4;
    float4 *blo = nullptr, *bhi = nullptr;
    uint32_t* cb = nullptr;
    uint64_t *k0 = nullptr, *k1 = nullptr;
    uint32_t *v0 = nullptr, *v1 = nullptr, *hist = nullptr, *scan_sums = nullptr;
    uint2 *child = nullptr, *range = nullptr;
    uint32_t *par_i = nullptr, *par_l = nullptr, *counters = nullptr;
    float *nlo = nullptr, *nhi = nullptr;
    Task *ta = nullptr, *tb = nullptr;
    SahParams sp{};
    uint32_t* cnt = nullptr;  // [0] ntasks a, [1] ntasks b, [2] nodes, [3] prims
    WideNode* nodes = nullptr;
    uint32_t* slots = nullptr;
    float4* prims_sorted = nullptr;
    float* aux = nullptr;
    hipError_t rc = hipSuccess;
    uint32_t nb = 0;
    uint32_t host_cnt[4] = {0, 0, 0, 0};
    uint32_t cur = 0;
    uint32_t cb_host[6];
    *out = BuiltBvh{};
    if (n == 0) return hipSuccess;
    if (leaf_max < 1) leaf_max = 1;
    if (leaf_max > 3) leaf_max = 3;
    if (width < 2) width = 2;
    if (width > 8) width = 8;

    VK_TRY(dalloc(&blo, n));
    VK_TRY(dalloc(&bhi, n));
    VK_TRY(dalloc(&cb, 6));
    VK_TRY(dalloc(&k0, n));
    VK_TRY(dalloc(&k1, n));
    VK_TRY(dalloc(&v0, n));
    VK_TRY(dalloc(&v1, n));
    nb = blocks(n, kRsTile);
    VK_TRY(dalloc(&hist, size_t(nb) * 256));
    VK_TRY(dalloc(&scan_sums, (size_t(nb) * 256 + kScanTile - 1) / kScanTile));
    VK_TRY(dalloc(&child, n));
    VK_TRY(dalloc(&range, n));
    VK_TRY(dalloc(&par_i, n));
    VK_TRY(dalloc(&par_l, n));
    VK_TRY(dalloc(&counters, n));
    VK_TRY(dalloc(&nlo, size_t(n) * 3));
    VK_TRY(dalloc(&nhi, size_t(n) * 3));
    VK_TRY(dalloc(&ta, n + 1));
    VK_TRY(dalloc(&tb, n + 1));
    VK_TRY(dalloc(&cnt, 4));
    VK_TRY(dalloc(&nodes, n));
    VK_TRY(dalloc(&slots, n));
    VK_TRY(dalloc(&prims_sorted, size_t(n) * rec));
    if (mode == kPhantom) VK_TRY(dalloc(&aux, n));
    if (collapse > 0 && n >= 2) {
        VK_TRY(dalloc(&sp.cost, n));
        sp.range = range;
        sp.leaf_max = leaf_max;
        sp.width = width;
        sp.c_node = 1.0f;
        sp.c_prim = 0.25f * float(collapse);
    }

    hipLaunchKernelGGL(k_init_bounds, dim3(1), dim3(64), 0, st, cb);
    hipLaunchKernelGGL(k_prim_boxes, dim3(std::min<uint32_t>(blocks(n), 2048u)), dim3(kBlock), 0, st, prims_in, n, mode,
                       radius_mode, radius, blo, bhi, cb);
    hipLaunchKernelGGL(k_morton, dim3(blocks(n)), dim3(kBlock), 0, st, blo, bhi, n, cb, k0, v0);
    VK_TRY(hipGetLastError());
    for (int pass = 0; pass < 8; ++pass) {
        int shift = pass * 8;
        hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(kBlock), 0, st, k0, n, shift, hist, nb);
        {
            const uint32_t count = nb * 256u, tiles = (count + kScanTile - 1) / kScanTile;
            hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(1024), 0, st, hist, count, scan_sums);
            hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(1024), 0, st, scan_sums, tiles);
            hipLaunchKernelGGL(k_scan_apply, dim3(tiles), dim3(1024), 0, st, hist, count, scan_sums);
        }
        hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(kBlock), 0, st, k0, v0, k1, v1, n, shift, hist, nb);
        std::swap(k0, k1);
        std::swap(v0, v1);
    }
    VK_TRY(hipGetLastError());
    if (n >= 2) {
        hipLaunchKernelGGL(k_karras, dim3(blocks(n - 1)), dim3(kBlock), 0, st, k0, int(n), child, range, par_i, par_l);
        hipLaunchKernelGGL(k_set_u32, dim3(blocks(n)), dim3(kBlock), 0, st, counters, 0u, n);
        // rounds in groups of 8; the root is the last node to finish
        uint32_t root_done = 0;
        for (uint32_t round = 1; !root_done; round += 8) {
            if (round > 8192) {
                err = "BVH refit did not reach the root";
                rc = hipErrorUnknown;
                goto fail;
            }
            for (uint32_t r = round; r < round + 8; ++r)
                hipLaunchKernelGGL(k_refit_round, dim3(blocks(n - 1)), dim3(kBlock), 0, st, int(n - 1), child, v0, blo,
                                   bhi, nlo, nhi, counters, r, sp);
            VK_TRY(hipGetLastError());
            VK_TRY(hipMemcpyAsync(&root_done, counters, 4, hipMemcpyDeviceToHost, st));
            VK_TRY(hipStreamSynchronize(st));
        }
    }
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, st, cnt, 0u, 4u);
    hipLaunchKernelGGL(k_root_task, dim3(1), dim3(1), 0, st, ta, cnt + 0, cnt + 2, n >= 2 ? 0u : kLeafFlag);
    VK_TRY(hipGetLastError());
    cur = 1;
    for (int level = 0; level < 4096 && cur > 0; ++level) {
        uint32_t* nin = cnt + (level & 1);
        uint32_t* nout = cnt + ((level & 1) ^ 1);
        hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, st, nout, 0u, 1u);
        hipLaunchKernelGGL(k_collapse, dim3(blocks(cur, 64)), dim3(64), 0, st, (level & 1) ? tb : ta, nin,
                           (level & 1) ? ta : tb, nout, cnt + 2, cnt + 3, int(n), leaf_max, width, child, range, v0, blo, bhi,
                           nlo, nhi, nodes, slots, sp);
        VK_TRY(hipGetLastError());
        VK_TRY(hipMemcpyAsync(&cur, nout, 4, hipMemcpyDeviceToHost, st));
        VK_TRY(hipStreamSynchronize(st));
        out->depth = uint32_t(level + 1);
    }
    VK_TRY(hipMemcpyAsync(host_cnt, cnt, 16, hipMemcpyDeviceToHost, st));
    VK_TRY(hipMemcpyAsync(cb_host, cb, 24, hipMemcpyDeviceToHost, st));
    hipLaunchKernelGGL(k_gather, dim3(blocks(uint64_t(n) * rec)), dim3(kBlock), 0, st, prims_in, slots, n, rec,
                       prims_sorted);
    if (mode == kPhantom)
        hipLaunchKernelGGL(k_phantom_rmax, dim3(blocks(n)), dim3(kBlock), 0, st, prims_sorted, n, radius_mode, radius,
                           aux);
    VK_TRY(hipGetLastError());
    VK_TRY(hipStreamSynchronize(st));
    if (host_cnt[3] != n || host_cnt[2] > n) {
        err = "BVH collapse placed " + std::to_string(host_cnt[3]) + " of " + std::to_string(n) + " primitives";
        rc = hipErrorUnknown;
        goto fail;
    }
    {
        uint32_t depth = out->depth;
        *out = BuiltBvh{};
        out->depth = depth;
    }
    out->nodes = nodes;
    out->nnodes = host_cnt[2];
    out->prim_ids = slots;
    out->prims = prims_sorted;
    out->aux = aux;
    aux = nullptr;
    nodes = nullptr;
    slots = nullptr;
    prims_sorted = nullptr;
    goto done;
fail:
    if (rc == hipSuccess) rc = hipErrorUnknown;
done:
    hipFree(blo);
    hipFree(bhi);
    hipFree(cb);
    hipFree(k0);
    hipFree(k1);
    hipFree(v0);
    hipFree(v1);
    hipFree(hist);
    hipFree(scan_sums);
    hipFree(child);
    hipFree(range);
    hipFree(par_i);
    hipFree(par_l);
    hipFree(counters);
    hipFree(nlo);
    hipFree(nhi);
    hipFree(ta);
    hipFree(tb);
    hipFree(cnt);
    hipFree(nodes);
    hipFree(slots);
    hipFree(prims_sorted);
    hipFree(aux);
    hipFree(sp.cost);
    return rc;
}

}  // namespace vk
