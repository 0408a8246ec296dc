// Oracle C API (test infrastructure only; see oracle_math.hpp). Loaded with ctypes by
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
//
// Semantics restated from the reference:
//  - traceRayEXT(tlas, gl_RayFlagsOpaqueEXT, 0xFF, ...) with tMin/tMax (ray_gen.rgen:26-45):
//    closest opaque hit with t in [tMin, tMax]; the phantom intersector reports t > 0 only
//    (hair_intersection.rint:146-149). Ties in t resolve to the lowest (instance, primitive),
//    a rule this build defines (the Vulkan spec leaves it open).
//  - Instance i of the TLAS has customIndex = i (top_level_acceleration_structure.cpp:30).
//  - Hair BLAS primitives are curve segments with GenerateAABBs bounds
//    (geometry_processor.cpp:422-436); the BVH is only an accelerator: brute force over all
//    primitives gives the same answer (pinned in tests).
// Additions with no reference counterpart: shadow rays toward -lightDirection
// (shading.glsl:4), the shadowed Shade, FP64 polish, the per-point radius.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <thread>
#include <vector>

#include "oracle_ff.hpp"
#include "oracle_count.hpp"
#include "oracle_geom.hpp"
#include "oracle_mlp.hpp"
#include "oracle_path.hpp"
#include "oracle_voxel.hpp"

using namespace orc;

namespace {

enum { MODE_PHANTOM = 0, MODE_LSS = 1, MODE_DOTS = 2, MODE_TRIS = 3, MODE_VOXELS = 4 };
enum {
    F_ANY_HIT = 1u << 0,
    F_BRUTE = 1u << 1,
    F_FP64 = 1u << 2,
    F_POLISH = 1u << 3,
    F_REF_NORMAL = 1u << 4,
    F_ACCURATE = 1u << 5,
};

struct Hit {
    float t;
    uint32_t prim;
    uint32_t inst;
    float u;
    float n[3];
    float v;
};
static_assert(sizeof(Hit) == 32, "hit record is 32 bytes");

struct Box {
    float mn[3], mx[3];
    void reset() {
        mn[0] = mn[1] = mn[2] = INFINITY;
        mx[0] = mx[1] = mx[2] = -INFINITY;
    }
    void grow(const Box& b) {
        for (int k = 0; k < 3; ++k) {
            mn[k] = std::min(mn[k], b.mn[k]);
            mx[k] = std::max(mx[k], b.mx[k]);
        }
    }
    float area() const {
        float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
        if (!(dx >= 0)) return 0.0f;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct Node {
    Box b;
    uint32_t left;   // internal: left child; leaf: first index into order[]
    uint32_t count;  // 0 = internal (right = left + 1 is NOT assumed: right stored separately)
    uint32_t right;
};

struct Blas {
    int mode = MODE_PHANTOM;
    bool per_point = false;
    float radius = 0.02f;
    std::vector<float> prims;        // phantom: 16/curve, LSS: 8/capsule, DOTS: 8/line
    std::vector<uint8_t> cap_a;      // LSS chained caps
    std::vector<DotsPiece> pieces;   // DOTS
    std::vector<float> tri, tnrm;    // triangle mesh: 3 vertex positions / normals per triangle
    std::vector<uint32_t> vbits;     // voxels: occupancy bits (x fastest)
    float vorigin[3] = {0, 0, 0};
    float vcell = 0.0f;
    int32_t vres[3] = {0, 0, 0};
    uint32_t nprims = 0;             // primitive count (DOTS: triangles)
    std::vector<Box> boxes;
    std::vector<uint32_t> order;
    std::vector<Node> nodes;
};

struct Instance {
    std::shared_ptr<Blas> blas;
    bool identity = true;
    uint32_t material = 0;  // batched path tracing: albedo tint index
    float xf[12];     // world-from-object, row-major 3x4
    M4<float> inv;    // object-from-world
};

struct Scene {
    std::vector<Instance> inst;
};

float pad_for(const Box& b) {
    float m = 0;
    for (int k = 0; k < 3; ++k) m = std::max(m, std::max(std::fabs(b.mn[k]), std::fabs(b.mx[k])));
    return 1e-4f + 4e-7f * m;
}

void prim_box(const Blas& B, uint32_t i, Box& bx, bool pad = true) {
    if (B.mode == MODE_PHANTOM) {
        CurveAABB(&B.prims[size_t(i) * 16], B.per_point, B.radius, bx.mn, bx.mx);
    } else if (B.mode == MODE_LSS) {
        const float* c = &B.prims[size_t(i) * 8];
        for (int k = 0; k < 3; ++k) {
            bx.mn[k] = std::min(c[k] - c[3], c[4 + k] - c[7]);
            bx.mx[k] = std::max(c[k] + c[3], c[4 + k] + c[7]);
        }
    } else if (B.mode == MODE_TRIS) {
        const float* v = &B.tri[size_t(i) * 9];
        for (int k = 0; k < 3; ++k) {
            bx.mn[k] = std::min(v[k], std::min(v[3 + k], v[6 + k]));
            bx.mx[k] = std::max(v[k], std::max(v[3 + k], v[6 + k]));
        }
    } else {
        const DotsPiece& p = B.pieces[i / 4];
        V3<float> v0, v1, v2;
        DotsTriangle(p, (i >> 1) & 1, i & 1, v0, v1, v2);
        for (int k = 0; k < 3; ++k) {
            float a = comp(v0, k), b = comp(v1, k), c = comp(v2, k);
            bx.mn[k] = std::min(a, std::min(b, c));
            bx.mx[k] = std::max(a, std::max(b, c));
        }
    }
    if (!pad) return;
    float p = pad_for(bx);
    for (int k = 0; k < 3; ++k) {
        bx.mn[k] -= p;
        bx.mx[k] += p;
    }
}

// ---------------------------------------------------------------- binned SAH BVH2
struct Builder {
    Blas& B;
    int leaf_size;
    std::vector<float> cen;  // 3 per prim
    std::atomic<uint32_t> next_node{0};

    uint32_t alloc() { return next_node.fetch_add(1); }

    void build(uint32_t ni, uint32_t first, uint32_t count, int depth) {
        Node& N = B.nodes[ni];
        Box bb;
        bb.reset();
        Box cb;
        cb.reset();
        for (uint32_t i = first; i < first + count; ++i) {
            uint32_t p = B.order[i];
            bb.grow(B.boxes[p]);
            for (int k = 0; k < 3; ++k) {
                cb.mn[k] = std::min(cb.mn[k], cen[p * 3 + k]);
                cb.mx[k] = std::max(cb.mx[k], cen[p * 3 + k]);
            }
        }
        N.b = bb;
        if (count <= uint32_t(leaf_size)) {
            N.left = first;
            N.count = count;
            return;
        }
        const int NB = 16;
        float best_cost = INFINITY;
        int best_axis = -1, best_split = 0;
        for (int ax = 0; ax < 3; ++ax) {
            float lo = cb.mn[ax], hi = cb.mx[ax];
            if (!(hi > lo)) continue;
            Box bins[NB];
            uint32_t cnt[NB] = {0};
            for (int b = 0; b < NB; ++b) bins[b].reset();
            float scale = NB / (hi - lo);
            for (uint32_t i = first; i < first + count; ++i) {
                uint32_t p = B.order[i];
                int b = std::min(NB - 1, int((cen[p * 3 + ax] - lo) * scale));
                cnt[b]++;
                bins[b].grow(B.boxes[p]);
            }
            float rarea[NB];
            uint32_t rcnt[NB];
            Box acc;
            acc.reset();
            uint32_t c = 0;
            for (int b = NB - 1; b > 0; --b) {
                acc.grow(bins[b]);
                c += cnt[b];
                rarea[b] = acc.area();
                rcnt[b] = c;
            }
            acc.reset();
            c = 0;
            for (int b = 0; b < NB - 1; ++b) {
                acc.grow(bins[b]);
                c += cnt[b];
                if (c == 0 || rcnt[b + 1] == 0) continue;
                float cost = acc.area() * c + rarea[b + 1] * rcnt[b + 1];
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = ax;
                    best_split = b;
                }
            }
        }
        uint32_t mid;
        if (best_axis < 0) {
            mid = first + count / 2;  // all centroids equal: split by index
        } else {
            float lo = cb.mn[best_axis], hi = cb.mx[best_axis];
            float scale = NB / (hi - lo);
            auto it = std::partition(B.order.begin() + first, B.order.begin() + first + count, [&](uint32_t p) {
                int b = std::min(NB - 1, int((cen[p * 3 + best_axis] - lo) * scale));
                return b <= best_split;
            });
            mid = uint32_t(it - B.order.begin());
            if (mid == first || mid == first + count) mid = first + count / 2;
        }
        uint32_t l = alloc(), r = alloc();
        N.left = l;
        N.right = r;
        N.count = 0;
        uint32_t lc = mid - first, rc = count - lc;
        if (depth < 4 && count > 50000) {
            std::thread th([&, l, first, lc, depth] { build(l, first, lc, depth + 1); });
            build(r, mid, rc, depth + 1);
            th.join();
        } else {
            build(l, first, lc, depth + 1);
            build(r, mid, rc, depth + 1);
        }
    }
};

void build_blas(Blas& B, int leaf_size) {
    uint32_t n = B.nprims;
    B.boxes.resize(n);
    B.order.resize(n);
    Builder bd{B, leaf_size};
    bd.cen.resize(size_t(n) * 3);
    for (uint32_t i = 0; i < n; ++i) {
        prim_box(B, i, B.boxes[i]);
        B.order[i] = i;
        for (int k = 0; k < 3; ++k) bd.cen[i * 3 + k] = 0.5f * (B.boxes[i].mn[k] + B.boxes[i].mx[k]);
    }
    B.nodes.assign(std::max<uint32_t>(1, 2 * n), Node{});
    if (n == 0) {
        B.nodes[0].b.reset();
        B.nodes[0].count = 0;
        B.nodes[0].left = 0;
        return;
    }
    bd.next_node = 1;
    bd.build(0, 0, n, 0);
    B.nodes.resize(bd.next_node.load());
}

// ---------------------------------------------------------------- per-primitive tests
template <class T>
bool test_prim(const Blas& B, uint32_t p, const Ray<float>& ray, float tmin, float tmax, uint32_t flags, Hit& h) {
    if (B.mode == MODE_PHANTOM) {
        const float* c = &B.prims[size_t(p) * 16];
        PrhiOpts o;
        o.radius_mode = B.per_point ? RADIUS_PER_POINT : RADIUS_CONSTANT;
        o.radius = B.radius;
        o.fp64_polish = (flags & F_POLISH) != 0;
        Curve<T> cv{mk(T(c[0]), T(c[1]), T(c[2])), mk(T(c[4]), T(c[5]), T(c[6])), mk(T(c[8]), T(c[9]), T(c[10])),
                    mk(T(c[12]), T(c[13]), T(c[14]))};
        T rr[4] = {T(c[3]), T(c[7]), T(c[11]), T(c[15])};
        Ray<T> r{cvt<T>(ray.origin), cvt<T>(ray.direction)};
        PrhiOut<T> out = Prhi(r, cv, rr, o);
        if (!(out.t > T(0))) return false;
        float t = float(out.t);
        if (!(t >= tmin && t <= tmax)) return false;
        h.t = t;
        h.u = float(out.u);
        V3<T> n = (flags & F_REF_NORMAL) ? out.n_ref : out.n_rcc;
        h.n[0] = float(n.x);
        h.n[1] = float(n.y);
        h.n[2] = float(n.z);
        h.v = out.polished ? 1.0f : 0.0f;
        return true;
    } else if (B.mode == MODE_LSS) {
        const float* c = &B.prims[size_t(p) * 8];
        LssOut<T> out;
        bool ok = IntersectLSS<T>(cvt<T>(ray.origin), cvt<T>(ray.direction), mk(T(c[0]), T(c[1]), T(c[2])), T(c[3]),
                                  mk(T(c[4]), T(c[5]), T(c[6])), T(c[7]), B.cap_a[p] != 0, T(tmin), T(tmax), out);
        if (!ok) return false;
        h.t = float(out.t);
        h.u = float(out.u);
        h.n[0] = float(out.n.x);
        h.n[1] = float(out.n.y);
        h.n[2] = float(out.n.z);
        h.v = 0.0f;
        return true;
    } else if (B.mode == MODE_TRIS) {
        // triangle_closest_hit.rchit:31-36: the interpolated vertex normal
        const float* v = &B.tri[size_t(p) * 9];
        const float* nv = &B.tnrm[size_t(p) * 9];
        const V3<float> v0 = mk(v[0], v[1], v[2]), v1 = mk(v[3], v[4], v[5]), v2 = mk(v[6], v[7], v[8]);
        TriOut<float> out;
        if (!IntersectTriangle(ray.origin, ray.direction, v0, v1, v2, tmin, tmax, out)) return false;
        h.t = out.t;
        h.u = out.b1;
        h.v = out.b2;
        const float b0 = 1.0f - out.b1 - out.b2;
        const V3<float> s = mk(nv[0], nv[1], nv[2]) * b0 + mk(nv[3], nv[4], nv[5]) * out.b1 +
                            mk(nv[6], nv[7], nv[8]) * out.b2;
        const V3<float> n = dot(s, s) > 0.0f ? normalize(s) : normalize(cross(v1 - v0, v2 - v0));
        h.n[0] = n.x;
        h.n[1] = n.y;
        h.n[2] = n.z;
        return true;
    } else {
        const DotsPiece& pc = B.pieces[p / 4];
        V3<float> v0, v1, v2;
        DotsTriangle(pc, (p >> 1) & 1, p & 1, v0, v1, v2);
        TriOut<float> out;
        if (!IntersectTriangle(ray.origin, ray.direction, v0, v1, v2, tmin, tmax, out)) return false;
        h.t = out.t;
        h.u = out.b1;
        h.v = out.b2;
        V3<float> n = normalize(cross(v1 - v0, v2 - v0));
        h.n[0] = n.x;
        h.n[1] = n.y;
        h.n[2] = n.z;
        return true;
    }
}

inline bool better(float t, uint32_t inst, uint32_t prim, const Hit& best) {
    if (t < best.t) return true;
    if (t == best.t) return inst < best.inst || (inst == best.inst && prim < best.prim);
    return false;
}

inline bool slab(const Box& b, const Ray<float>& r, const float inv[3], float tmin, float tmax) {
    float t0 = tmin, t1 = tmax;
    const float o[3] = {r.origin.x, r.origin.y, r.origin.z};
    for (int k = 0; k < 3; ++k) {
        float a = (b.mn[k] - o[k]) * inv[k];
        float c = (b.mx[k] - o[k]) * inv[k];
        if (std::isnan(a) || std::isnan(c)) {  // origin on a slab plane with a zero direction
            a = -INFINITY;
            c = INFINITY;
        }
        float lo = std::min(a, c), hi = std::max(a, c);
        t0 = std::max(t0, lo);
        t1 = std::min(t1, hi * (1.0f + 4e-7f));
    }
    return t0 <= t1;
}

// Traces one ray against one instance. Returns the number of primitive tests.
template <class T>
uint32_t trace_blas(const Blas& B, uint32_t inst_id, const Ray<float>& r, float tmin, float& tmax, uint32_t flags,
                    Hit& best, bool& found, std::vector<uint32_t>* cand) {
    uint32_t tests = 0;
    if (B.nprims == 0) return 0;
    const bool any = (flags & F_ANY_HIT) != 0;
    if (B.mode == MODE_VOXELS) {
        const float o[3] = {r.origin.x, r.origin.y, r.origin.z}, d[3] = {r.direction.x, r.direction.y, r.direction.z};
        VoxelHit vh{};
        const bool hit = (flags & F_BRUTE) ? VoxelBrute(B.vbits, B.vorigin, B.vcell, B.vres, o, d, tmin, tmax, vh)
                                           : VoxelWalk(B.vbits, B.vorigin, B.vcell, B.vres, o, d, tmin, tmax, vh);
        if (hit && (!found || better(vh.t, inst_id, vh.cell, best))) {
            best = Hit{};
            best.t = vh.t;
            best.prim = vh.cell;
            best.inst = inst_id;
            std::memcpy(best.n, vh.n, sizeof(best.n));
            found = true;
            if (!any) tmax = vh.t;
        }
        return 1;
    }
    if (flags & F_BRUTE) {
        for (uint32_t p = 0; p < B.nprims; ++p) {
            Hit h{};
            ++tests;
            if (test_prim<T>(B, p, r, tmin, tmax, flags, h)) {
                if (!found || better(h.t, inst_id, p, best)) {
                    best = h;
                    best.prim = p;
                    best.inst = inst_id;
                    found = true;
                    if (!any) tmax = h.t;
                }
                if (any) return tests;
            }
        }
        return tests;
    }
    float inv[3] = {1.0f / r.direction.x, 1.0f / r.direction.y, 1.0f / r.direction.z};
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const Node& N = B.nodes[stack[--sp]];
        if (!slab(N.b, r, inv, tmin, tmax)) continue;
        if (N.count) {
            for (uint32_t i = N.left; i < N.left + N.count; ++i) {
                uint32_t p = B.order[i];
                if (cand) {
                    Box exact;
                    prim_box(B, p, exact, false);  // GenerateAABBs box, no pad
                    if (slab(exact, r, inv, tmin, tmax)) cand->push_back(p);
                    continue;
                }
                Hit h{};
                ++tests;
                if (test_prim<T>(B, p, r, tmin, tmax, flags, h)) {
                    if (!found || better(h.t, inst_id, p, best)) {
                        best = h;
                        best.prim = p;
                        best.inst = inst_id;
                        found = true;
                        if (!any) tmax = h.t;
                    }
                    if (any) return tests;
                }
            }
        } else {
            stack[sp++] = N.left;
            stack[sp++] = N.right;
            if (sp > 126) {
                std::fprintf(stderr, "oracle: BVH2 stack overflow\n");
                std::abort();
            }
        }
    }
    return tests;
}

void obj_ray(const Instance& I, const Ray<float>& w, Ray<float>& o) {
    if (I.identity) {
        o = w;
        return;
    }
    o.origin = xform_point(I.inv, w.origin);
    float v[4] = {w.direction.x, w.direction.y, w.direction.z, 0.0f}, r[4];
    mul4(I.inv, v, r);
    o.direction = mk(r[0], r[1], r[2]);
}

void normal_to_world(const Instance& I, float n[3]) {
    if (I.identity) return;
    V3<float> a = mk(I.xf[0] * n[0] + I.xf[1] * n[1] + I.xf[2] * n[2], I.xf[4] * n[0] + I.xf[5] * n[1] + I.xf[6] * n[2],
                     I.xf[8] * n[0] + I.xf[9] * n[1] + I.xf[10] * n[2]);
    a = normalize(a);
    n[0] = a.x;
    n[1] = a.y;
    n[2] = a.z;
}

// Double-word re-run of Prhi for the final phantom hit (extension): the iteration of the FP64
// evaluation from the start, in FP32 pairs (oracle_ff.hpp), without the early-out test; t, u
// and the ray-centric normal are replaced when it converges with t > 0. Polished hits keep
// their FP64 values.
void accurate_phantom(const Instance& I, const Ray<float>& w, Hit& best) {
    const Blas& B = *I.blas;
    if (B.mode != MODE_PHANTOM || best.v != 0.0f) return;
    Ray<float> o;
    obj_ray(I, w, o);
    const float* c = &B.prims[size_t(best.prim) * 16];
    PrhiOpts opt;
    opt.radius_mode = B.per_point ? RADIUS_PER_POINT : RADIUS_CONSTANT;
    opt.radius = B.radius;
    opt.cull = false;
    Curve<FF> cv{mk(FF(c[0]), FF(c[1]), FF(c[2])), mk(FF(c[4]), FF(c[5]), FF(c[6])), mk(FF(c[8]), FF(c[9]), FF(c[10])),
                 mk(FF(c[12]), FF(c[13]), FF(c[14]))};
    FF rr[4] = {FF(c[3]), FF(c[7]), FF(c[11]), FF(c[15])};
    Ray<FF> r{cvt<FF>(o.origin), cvt<FF>(o.direction)};
    PrhiOut<FF> out = Prhi(r, cv, rr, opt);
    if (!(out.t > FF(0.0f))) return;
    best.t = float(out.t);
    best.u = float(out.u);
    best.n[0] = float(out.n_rcc.x);
    best.n[1] = float(out.n_rcc.y);
    best.n[2] = float(out.n_rcc.z);
    normal_to_world(I, best.n);
}

template <class T>
bool trace_scene(const Scene& S, const Ray<float>& r, float tmin, float tmax, uint32_t flags, Hit& best) {
    bool found = false;
    best = Hit{};
    best.t = tmax;
    float tcur = tmax;
    for (uint32_t i = 0; i < S.inst.size(); ++i) {
        const Instance& I = S.inst[i];
        Ray<float> o;
        obj_ray(I, r, o);
        bool was = found;
        uint32_t prev_inst = best.inst, prev_prim = best.prim;
        float prev_t = best.t;
        trace_blas<T>(*I.blas, i, o, tmin, tcur, flags, best, found, nullptr);
        bool updated = found && (!was || best.inst != prev_inst || best.prim != prev_prim || best.t != prev_t);
        if (updated) {
            normal_to_world(I, best.n);
            if (flags & F_ANY_HIT) return true;
        }
    }
    if (!found) {
        best.t = 0.0f;
        best.prim = best.inst = 0xFFFFFFFFu;
    } else if (flags & F_ACCURATE) {
        accurate_phantom(S.inst[best.inst], r, best);
    }
    return found;
}

template <class F>
void parallel_for(uint32_t n, int nthreads, F f) {
    if (nthreads <= 0) nthreads = int(std::max(1u, std::thread::hardware_concurrency()));
    std::atomic<uint32_t> next{0};
    const uint32_t chunk = 64;
    auto work = [&] {
        for (;;) {
            uint32_t b = next.fetch_add(chunk);
            if (b >= n) break;
            uint32_t e = std::min(n, b + chunk);
            for (uint32_t i = b; i < e; ++i) f(i);
        }
    };
    if (nthreads == 1) {
        work();
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) th.emplace_back(work);
    for (auto& t : th) t.join();
}

// Appends an instance of B with an optional row-major 3x4 world-from-object transform.
int add_instance(Scene& S, const std::shared_ptr<Blas>& B, const float* transform) {
    Instance I;
    I.blas = B;
    I.identity = true;
    if (transform) {
        std::memcpy(I.xf, transform, 48);
        const float id[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
        I.identity = std::memcmp(transform, id, 48) == 0;
        M4<float> M{};
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 4; ++c) M.m[c][r] = transform[r * 4 + c];
        M.m[0][3] = M.m[1][3] = M.m[2][3] = 0.0f;
        M.m[3][3] = 1.0f;
        I.inv = inverse(M);
    }
    S.inst.push_back(I);
    return int(S.inst.size() - 1);
}

inline Ray<float> load_ray(const float* r, float& tmin, float& tmax) {
    tmin = r[3];
    tmax = r[7];
    return Ray<float>{mk(r[0], r[1], r[2]), mk(r[4], r[5], r[6])};
}

}  // namespace

extern "C" {

typedef struct orc_scene orc_scene;

int orc_abi_version() { return 1; }
int orc_hardware_threads() { return int(std::max(1u, std::thread::hardware_concurrency())); }

// ---- generator / geometry processing
int orc_generate_strands(uint64_t seed, uint32_t nstrands, uint32_t nseg, float sphere_radius, float* out,
                         int nthreads) {
    HairParams hp;
    hp.sphere_radius = sphere_radius;
    parallel_for(nstrands, nthreads, [&](uint32_t s) {
        GenerateStrand(seed, s, int(nseg), hp, out + size_t(s) * (nseg + 1) * 4);
    });
    return 0;
}

// Strand polylines -> lines (GenerateLines over consecutive index pairs,
// geometry_processor.cpp:45-67): line k of strand s = (v[k], v[k+1]); 8 floats per line.
int orc_strands_to_lines(const float* verts, uint32_t nstrands, uint32_t nseg, float* lines) {
    for (uint32_t s = 0; s < nstrands; ++s)
        for (uint32_t k = 0; k < nseg; ++k) {
            const float* a = verts + (size_t(s) * (nseg + 1) + k) * 4;
            float* L = lines + (size_t(s) * nseg + k) * 8;
            L[0] = a[0];
            L[1] = a[1];
            L[2] = a[2];
            L[3] = a[3];
            L[4] = a[4];
            L[5] = a[5];
            L[6] = a[6];
            L[7] = a[7];
        }
    return 0;
}

static std::vector<LineR> lines_from(const float* lines, uint32_t n) {
    std::vector<LineR> v(n);
    for (uint32_t i = 0; i < n; ++i) {
        const float* L = lines + size_t(i) * 8;
        v[i].start = mk(L[0], L[1], L[2]);
        v[i].r0 = L[3];
        v[i].end = mk(L[4], L[5], L[6]);
        v[i].r1 = L[7];
    }
    return v;
}

int orc_generate_curves(const float* lines, uint32_t n, float* curves) {
    std::vector<float> out;
    GenerateCurves(lines_from(lines, n), out);
    std::memcpy(curves, out.data(), out.size() * sizeof(float));
    return 0;
}

// GenerateLines (geometry_processor.cpp:45-67) over an indexed line mesh whose indices are
// mesh-local: line i = (position[idx[2i]], radius, position[idx[2i+1]], radius), vertices of 8
// floats (Mesh::Vertex, model.hpp:44-49). The reference adds the mesh's firstVertex to indices
// that ProcessMesh (model_loader.cpp:157) already offset; here it is applied once. Returns -1
// (and writes nothing further) at the first index not below nv.
int orc_mesh_lines(const float* verts, uint32_t nv, const uint32_t* idx, uint32_t n, float radius, float* lines) {
    for (uint32_t i = 0; i < n; ++i)
        for (int k = 0; k < 2; ++k) {
            const uint32_t v = idx[size_t(i) * 2 + k];
            if (v >= nv) return -1;
            float* L = lines + size_t(i) * 8 + k * 4;
            for (int c = 0; c < 3; ++c) L[c] = verts[size_t(v) * 8 + c];
            L[3] = radius;
        }
    return 0;
}

// Node::GetWorldMatrix (source/resources/model/model.cpp:6-18): matrix = localMatrix, then
// matrix = parent->localMatrix * matrix for each ancestor from the nearest one, as glm::mat4
// products in float (column j of A * B = A * B[j], summed k = 0..3 in order). `locals` and
// `out` hold n column-major 4x4 matrices; parents[i] < 0 marks a root. -1 on a parent cycle.
int orc_world_matrix(const float* locals, const int32_t* parents, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        M4<float> m;
        std::memcpy(m.m, locals + size_t(i) * 16, 64);
        uint32_t steps = 0;
        for (int32_t p = parents[i]; p >= 0; p = parents[p]) {
            if (uint32_t(p) >= n || ++steps > n) return -1;
            M4<float> P, r;
            std::memcpy(P.m, locals + size_t(p) * 16, 64);
            for (int j = 0; j < 4; ++j) mul4(P, m.m[j], r.m[j]);
            m = r;
        }
        std::memcpy(out + size_t(i) * 16, m.m, 64);
    }
    return 0;
}

// Voxel grid of GenerateVoxelMesh over [lo, hi] (or the lines' end points when lo is null),
// filled from `lines` (n x 8 floats) and OR-reduced `lod` times. res_out: 3 cells per axis,
// geo_out: origin xyz, cell edge. bits: (res product + 31) / 32 words, or null to query the
// size. Returns the word count, -1 for an invalid grid, -2 when a line visits > 65536 cells.
long long orc_voxelize(const float* lines, uint32_t n, const float* lo, const float* hi, float radius, float size,
                       int reference_sweep, uint32_t lod, uint32_t* bits, int32_t* res_out, float* geo_out) {
    float blo[3] = {0, 0, 0}, bhi[3] = {0, 0, 0};
    if (lo && hi) {
        std::memcpy(blo, lo, sizeof(blo));
        std::memcpy(bhi, hi, sizeof(bhi));
    } else if (n) {
        for (int k = 0; k < 3; ++k) {
            blo[k] = INFINITY;
            bhi[k] = -INFINITY;
        }
        for (uint32_t i = 0; i < 2 * n; ++i)
            for (int k = 0; k < 3; ++k) {
                const float c = lines[size_t(i) * 4 + k];
                if (c < blo[k]) blo[k] = c;
                if (c > bhi[k]) bhi[k] = c;
            }
    }
    VGrid g;
    if (!VoxelGridFor(blo, bhi, size, g)) return -1;
    std::vector<uint32_t> v((size_t(g.count) + 31) / 32, 0u);
    if (!Voxelize(lines, n, g, radius, reference_sweep != 0, v)) return -2;
    int32_t res[3] = {g.res[0], g.res[1], g.res[2]};
    float cell = g.size;
    for (uint32_t l = 0; l < lod && g.count; ++l) {
        int32_t cr[3];
        v = Downsample(v, res, cr);
        std::memcpy(res, cr, sizeof(res));
        cell = cell * 2.0f;
    }
    std::memcpy(res_out, res, sizeof(res));
    for (int k = 0; k < 3; ++k) geo_out[k] = g.origin[k];
    geo_out[3] = cell;
    if (bits) std::memcpy(bits, v.data(), v.size() * sizeof(uint32_t));
    return (long long)v.size();
}

int orc_curves_to_lines(const float* curves, uint32_t n, uint32_t pieces, float* lines) {
    std::vector<float> c(curves, curves + size_t(n) * 16);
    std::vector<LineR> L;
    CurvesToLines(c, int(pieces), L);
    for (size_t i = 0; i < L.size(); ++i) {
        float* o = lines + i * 8;
        o[0] = L[i].start.x;
        o[1] = L[i].start.y;
        o[2] = L[i].start.z;
        o[3] = L[i].r0;
        o[4] = L[i].end.x;
        o[5] = L[i].end.y;
        o[6] = L[i].end.z;
        o[7] = L[i].r1;
    }
    return 0;
}

// GenerateLinearSweptSpheres (geometry_processor.cpp:273-298): one capsule per line,
// radius max(r, 0.001) (:285). constant_radius < 0 keeps the per-end radii.
int orc_generate_lss(const float* lines, uint32_t n, float constant_radius, float* caps) {
    for (uint32_t i = 0; i < n; ++i) {
        const float* L = lines + size_t(i) * 8;
        float* c = caps + size_t(i) * 8;
        for (int k = 0; k < 8; ++k) c[k] = L[k];
        float r0 = constant_radius >= 0 ? constant_radius : L[3];
        float r1 = constant_radius >= 0 ? constant_radius : L[7];
        c[3] = gmax(r0, 0.001f);
        c[7] = gmax(r1, 0.001f);
    }
    return 0;
}

// DOTS pieces: 16 floats (start, r0, end, r1, s, 0, t, 0).
int orc_generate_dots(const float* lines, uint32_t n, float constant_radius, float* pieces) {
    for (uint32_t i = 0; i < n; ++i) {
        const float* L = lines + size_t(i) * 8;
        float r0 = constant_radius >= 0 ? constant_radius : L[3];
        float r1 = constant_radius >= 0 ? constant_radius : L[7];
        DotsPiece p = MakeDotsPiece(mk(L[0], L[1], L[2]), mk(L[4], L[5], L[6]), r0, r1);
        float v[16] = {p.start.x, p.start.y, p.start.z, p.r0, p.end.x, p.end.y, p.end.z, p.r1,
                       p.s.x,     p.s.y,     p.s.z,     0.0f, p.t.x,   p.t.y,   p.t.z,   0.0f};
        std::memcpy(pieces + size_t(i) * 16, v, sizeof(v));
    }
    return 0;
}

// Expanded DOTS triangles (reference vertex layout, geometry_processor.cpp:240-266):
// 4 triangles x 3 vertices x xyz per piece.
int orc_dots_triangles(const float* pieces, uint32_t n, float* tris) {
    for (uint32_t i = 0; i < n; ++i) {
        const float* q = pieces + size_t(i) * 16;
        DotsPiece p;
        p.start = mk(q[0], q[1], q[2]);
        p.r0 = q[3];
        p.end = mk(q[4], q[5], q[6]);
        p.r1 = q[7];
        p.s = mk(q[8], q[9], q[10]);
        p.t = mk(q[12], q[13], q[14]);
        for (int f = 0; f < 2; ++f)
            for (int t = 0; t < 2; ++t) {
                V3<float> v[3];
                DotsTriangle(p, f, t, v[0], v[1], v[2]);
                float* o = tris + (size_t(i) * 4 + f * 2 + t) * 9;
                for (int j = 0; j < 3; ++j) {
                    o[j * 3 + 0] = v[j].x;
                    o[j * 3 + 1] = v[j].y;
                    o[j * 3 + 2] = v[j].z;
                }
            }
    }
    return 0;
}

int orc_curve_aabbs(const float* curves, uint32_t n, int per_point, float radius, float* aabbs) {
    for (uint32_t i = 0; i < n; ++i) CurveAABB(curves + size_t(i) * 16, per_point != 0, radius, aabbs + i * 6, aabbs + i * 6 + 3);
    return 0;
}

// ---- camera / ray generation
int orc_camera(const float eye[3], const float center[3], const float up[3], float tan_half_fovy, float aspect,
               float znear, float zfar, float view_inv[16], float proj_inv[16]) {
    M4<float> V = LookAtRH(mk(eye[0], eye[1], eye[2]), mk(center[0], center[1], center[2]), mk(up[0], up[1], up[2]));
    M4<float> P = PerspectiveRH_ZO_FlipY(tan_half_fovy, aspect, znear, zfar);
    M4<float> Vi = inverse(V), Pi = inverse(P);
    std::memcpy(view_inv, Vi.m, 64);
    std::memcpy(proj_inv, Pi.m, 64);
    return 0;
}

int orc_raygen(const float view_inv[16], const float proj_inv[16], uint32_t W, uint32_t H, float tmin, float tmax,
               float* rays) {
    M4<float> Vi, Pi;
    std::memcpy(Vi.m, view_inv, 64);
    std::memcpy(Pi.m, proj_inv, 64);
    for (uint32_t y = 0; y < H; ++y)
        for (uint32_t x = 0; x < W; ++x) {
            V3<float> o, d;
            RayGen(Vi, Pi, x, y, W, H, o, d);
            float* r = rays + (size_t(y) * W + x) * 8;
            r[0] = o.x;
            r[1] = o.y;
            r[2] = o.z;
            r[3] = tmin;
            r[4] = d.x;
            r[5] = d.y;
            r[6] = d.z;
            r[7] = tmax;
        }
    return 0;
}

// ---- single-function evaluators (used by the closed-form tests)
// fp64 != 0 evaluates the FP64 instantiation on the float inputs.
int orc_prhi(const float* ray /*8*/, const float* curve /*16*/, int per_point, float radius, int polish, int fp64,
             float* out /*t, u, n_ref[3], n_rcc[3], iters, polished*/) {
    PrhiOpts o;
    o.radius_mode = per_point ? RADIUS_PER_POINT : RADIUS_CONSTANT;
    o.radius = radius;
    o.fp64_polish = polish != 0;
    auto run = [&](auto zero) {
        using T = decltype(zero);
        Ray<T> r{mk(T(ray[0]), T(ray[1]), T(ray[2])), mk(T(ray[4]), T(ray[5]), T(ray[6]))};
        Curve<T> c{mk(T(curve[0]), T(curve[1]), T(curve[2])), mk(T(curve[4]), T(curve[5]), T(curve[6])),
                   mk(T(curve[8]), T(curve[9]), T(curve[10])), mk(T(curve[12]), T(curve[13]), T(curve[14]))};
        T rr[4] = {T(curve[3]), T(curve[7]), T(curve[11]), T(curve[15])};
        PrhiOut<T> p = Prhi(r, c, rr, o);
        double v[10] = {double(p.t),       double(p.u),       double(p.n_ref.x), double(p.n_ref.y), double(p.n_ref.z),
                        double(p.n_rcc.x), double(p.n_rcc.y), double(p.n_rcc.z), double(p.iters),   p.polished ? 1.0 : 0.0};
        for (int i = 0; i < 10; ++i) out[i] = float(v[i]);
    };
    if (fp64 == 2) {  // double-word FP32: the final-hit re-run (no early out, no polish)
        o.cull = false;
        o.fp64_polish = false;
        run(FF(0.0f));
    } else if (fp64) {
        run(0.0);
    } else {
        run(0.0f);
    }
    return 0;
}

// Double-word operations for pinning oracle_ff.hpp: a, b, out hold n (hi, lo) pairs;
// op 0 add, 1 sub, 2 mul, 3 div, 4 sqrt (of a).
int orc_ff_eval(const float* a, const float* b, uint32_t n, int op, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        const FF x(a[2 * i], a[2 * i + 1]), y(b[2 * i], b[2 * i + 1]);
        const FF z = op == 0 ? x + y : op == 1 ? x - y : op == 2 ? x * y : op == 3 ? x / y : sqrt(x);
        out[2 * i] = z.hi;
        out[2 * i + 1] = z.lo;
    }
    return 0;
}

// Same as orc_prhi but writes double results (for FP64 error references).
int orc_prhi_f64(const float* ray, const float* curve, int per_point, float radius, double* out) {
    PrhiOpts o;
    o.radius_mode = per_point ? RADIUS_PER_POINT : RADIUS_CONSTANT;
    o.radius = radius;
    Ray<double> r{mk(double(ray[0]), double(ray[1]), double(ray[2])), mk(double(ray[4]), double(ray[5]), double(ray[6]))};
    Curve<double> c{mk(double(curve[0]), double(curve[1]), double(curve[2])),
                    mk(double(curve[4]), double(curve[5]), double(curve[6])),
                    mk(double(curve[8]), double(curve[9]), double(curve[10])),
                    mk(double(curve[12]), double(curve[13]), double(curve[14]))};
    double rr[4] = {curve[3], curve[7], curve[11], curve[15]};
    PrhiOut<double> p = Prhi(r, c, rr, o);
    double v[10] = {p.t, p.u, p.n_ref.x, p.n_ref.y, p.n_ref.z, p.n_rcc.x, p.n_rcc.y, p.n_rcc.z, double(p.iters), 0.0};
    std::memcpy(out, v, sizeof(v));
    return 0;
}

int orc_rcc_matrix(const float* ray, int fp64, double* out /*16 column-major*/) {
    if (fp64) {
        Ray<double> r{mk(double(ray[0]), double(ray[1]), double(ray[2])), mk(double(ray[4]), double(ray[5]), double(ray[6]))};
        M4<double> M = CreateRCCMatrix(r);
        std::memcpy(out, M.m, 128);
    } else {
        Ray<float> r{mk(ray[0], ray[1], ray[2]), mk(ray[4], ray[5], ray[6])};
        M4<float> M = CreateRCCMatrix(r);
        for (int c = 0; c < 4; ++c)
            for (int k = 0; k < 4; ++k) out[c * 4 + k] = M.m[c][k];
    }
    return 0;
}

int orc_inverse4(const float* m, float* out) {
    M4<float> M;
    std::memcpy(M.m, m, 64);
    M4<float> R = inverse(M);
    std::memcpy(out, R.m, 64);
    return 0;
}

int orc_basis(const float* w, float* u, float* v) {
    V3<float> uu, vv;
    MakeOrthonormalBasis(uu, vv, mk(w[0], w[1], w[2]));
    u[0] = uu.x;
    u[1] = uu.y;
    u[2] = uu.z;
    v[0] = vv.x;
    v[1] = vv.y;
    v[2] = vv.z;
    return 0;
}

int orc_sample_curve(const float* curve, float t, int fp64, double* p /*3*/, double* axis /*3*/) {
    auto run = [&](auto zero) {
        using T = decltype(zero);
        Curve<T> c{mk(T(curve[0]), T(curve[1]), T(curve[2])), mk(T(curve[4]), T(curve[5]), T(curve[6])),
                   mk(T(curve[8]), T(curve[9]), T(curve[10])), mk(T(curve[12]), T(curve[13]), T(curve[14]))};
        V3<T> a = SampleCurvePoint(c, T(t)), b = SampleCurveAxis(c, T(t));
        p[0] = a.x;
        p[1] = a.y;
        p[2] = a.z;
        axis[0] = b.x;
        axis[1] = b.y;
        axis[2] = b.z;
    };
    if (fp64)
        run(0.0);
    else
        run(0.0f);
    return 0;
}

int orc_cylinder(const float* ray, const float* p0, const float* p1, float radius) {
    Ray<float> r{mk(ray[0], ray[1], ray[2]), mk(ray[4], ray[5], ray[6])};
    return RayCylinderIntersect(r, mk(p0[0], p0[1], p0[2]), mk(p1[0], p1[1], p1[2]), radius) ? 1 : 0;
}

int orc_cone(const float* center, float radius, const float* axis, float slant, float* out /*isect, s, dt, dp, dc, sp*/) {
    Cone<float> c{mk(center[0], center[1], center[2]), radius, mk(axis[0], axis[1], axis[2]), slant};
    RCI<float> r = RayConeIntersectRCC(c);
    out[0] = r.isIntersection ? 1.0f : 0.0f;
    out[1] = r.s;
    out[2] = r.dt;
    out[3] = r.dp;
    out[4] = r.dc;
    out[5] = r.sp;
    return 0;
}

int orc_lss(const float* ray, const float* cap, int cap_a, int fp64, double* out /*hit, t, u, n[3]*/) {
    auto run = [&](auto zero) {
        using T = decltype(zero);
        LssOut<T> o{};
        bool ok = IntersectLSS<T>(mk(T(ray[0]), T(ray[1]), T(ray[2])), mk(T(ray[4]), T(ray[5]), T(ray[6])),
                                  mk(T(cap[0]), T(cap[1]), T(cap[2])), T(cap[3]), mk(T(cap[4]), T(cap[5]), T(cap[6])),
                                  T(cap[7]), cap_a != 0, T(ray[3]), T(ray[7]), o);
        out[0] = ok ? 1 : 0;
        out[1] = o.t;
        out[2] = o.u;
        out[3] = o.n.x;
        out[4] = o.n.y;
        out[5] = o.n.z;
    };
    if (fp64)
        run(0.0);
    else
        run(0.0f);
    return 0;
}

float orc_det2(float a, float b, float c, float d) { return Det2Kahan(a, b, c, d); }

int orc_triangle(const float* ray, const float* v /*9*/, float* out /*hit, t, b1, b2*/) {
    TriOut<float> o{};
    bool ok = IntersectTriangle(mk(ray[0], ray[1], ray[2]), mk(ray[4], ray[5], ray[6]), mk(v[0], v[1], v[2]),
                                mk(v[3], v[4], v[5]), mk(v[6], v[7], v[8]), ray[3], ray[7], o);
    out[0] = ok ? 1.0f : 0.0f;
    out[1] = o.t;
    out[2] = o.b1;
    out[3] = o.b2;
    return 0;
}

int orc_perp_stark(const float* u, float* v) {
    V3<float> r = PerpStark(mk(u[0], u[1], u[2]));
    v[0] = r.x;
    v[1] = r.y;
    v[2] = r.z;
    return 0;
}

int orc_shade(const float* n, float vis, float* rgb) {
    V3<float> c = Shade(mk(n[0], n[1], n[2]), vis);
    rgb[0] = c.x;
    rgb[1] = c.y;
    rgb[2] = c.z;
    return 0;
}

int orc_debug_color(int32_t prim, float* rgb) {
    V3<float> c = GetDebugColor(prim);
    rgb[0] = c.x;
    rgb[1] = c.y;
    rgb[2] = c.z;
    return 0;
}

// ---- scenes
orc_scene* orc_scene_create() { return reinterpret_cast<orc_scene*>(new Scene()); }
void orc_scene_destroy(orc_scene* s) { delete reinterpret_cast<Scene*>(s); }

// mode 0: prims = curves (16 floats); 1: capsules (8 floats); 2: lines (8 floats) -> DOTS
// (4 triangles per line). transform: row-major 3x4 world-from-object or NULL.
int orc_scene_add_instance(orc_scene* sp, int mode, const float* prims, uint32_t n, int per_point, float radius,
                           const float* transform, int leaf_size) {
    Scene& S = *reinterpret_cast<Scene*>(sp);
    auto B = std::make_shared<Blas>();
    B->mode = mode;
    B->per_point = per_point != 0;
    B->radius = radius;
    if (mode == MODE_PHANTOM) {
        B->prims.assign(prims, prims + size_t(n) * 16);
        B->nprims = n;
    } else if (mode == MODE_LSS) {
        B->prims.assign(prims, prims + size_t(n) * 8);
        B->nprims = n;
        B->cap_a.resize(n);
        for (uint32_t i = 0; i < n; ++i) {
            const float* c = prims + size_t(i) * 8;
            bool chained = i > 0 && prims[(i - 1) * 8 + 4] == c[0] && prims[(i - 1) * 8 + 5] == c[1] &&
                           prims[(i - 1) * 8 + 6] == c[2];
            B->cap_a[i] = chained ? 0 : 1;
        }
    } else if (mode == MODE_DOTS) {
        B->pieces.resize(n);
        for (uint32_t i = 0; i < n; ++i) {
            const float* L = prims + size_t(i) * 8;
            B->pieces[i] = MakeDotsPiece(mk(L[0], L[1], L[2]), mk(L[4], L[5], L[6]), per_point ? L[3] : radius,
                                         per_point ? L[7] : radius);
        }
        B->nprims = n * 4;
    } else {
        return -1;
    }
    build_blas(*B, leaf_size > 0 ? leaf_size : 4);
    return add_instance(S, B, transform);
}

// Indexed triangle mesh: verts = nverts x 8 floats (position, normal, texCoord), indices =
// 3 x ntri. Returns the instance index, or -1 when an index is not below nverts.
int orc_scene_add_mesh(orc_scene* sp, const float* verts, uint32_t nverts, const uint32_t* indices, uint32_t ntri,
                       const float* transform, int leaf_size) {
    auto B = std::make_shared<Blas>();
    B->mode = MODE_TRIS;
    B->nprims = ntri;
    B->tri.resize(size_t(ntri) * 9);
    B->tnrm.resize(size_t(ntri) * 9);
    for (uint32_t t = 0; t < ntri; ++t)
        for (int k = 0; k < 3; ++k) {
            const uint32_t v = indices[size_t(t) * 3 + k];
            if (v >= nverts) return -1;
            for (int c = 0; c < 3; ++c) {
                B->tri[size_t(t) * 9 + k * 3 + c] = verts[size_t(v) * 8 + c];
                B->tnrm[size_t(t) * 9 + k * 3 + c] = verts[size_t(v) * 8 + 3 + c];
            }
        }
    build_blas(*B, leaf_size > 0 ? leaf_size : 4);
    return add_instance(*reinterpret_cast<Scene*>(sp), B, transform);
}

// A voxel instance over an occupancy grid (bits: res product bits, x fastest).
int orc_scene_add_voxels(orc_scene* sp, const uint32_t* bits, const int32_t* res, const float* origin, float cell,
                         const float* transform) {
    Scene& S = *reinterpret_cast<Scene*>(sp);
    auto B = std::make_shared<Blas>();
    B->mode = MODE_VOXELS;
    const uint64_t count = uint64_t(res[0]) * uint64_t(res[1]) * uint64_t(res[2]);
    B->vbits.assign(bits, bits + (count + 31) / 32);
    std::memcpy(B->vres, res, sizeof(B->vres));
    std::memcpy(B->vorigin, origin, sizeof(B->vorigin));
    B->vcell = cell;
    uint32_t filled = 0;
    for (uint32_t w : B->vbits) filled += uint32_t(__builtin_popcount(w));
    B->nprims = filled;
    return add_instance(S, B, transform);
}

int orc_scene_set_material(orc_scene* sp, uint32_t inst, uint32_t material) {
    Scene& S = *reinterpret_cast<Scene*>(sp);
    if (inst >= S.inst.size()) return -1;
    S.inst[inst].material = material;
    return 0;
}

int orc_scene_stats(orc_scene* sp, uint32_t inst, uint32_t* out /*nprims, nnodes*/) {
    Scene& S = *reinterpret_cast<Scene*>(sp);
    if (inst >= S.inst.size()) return -1;
    out[0] = S.inst[inst].blas->nprims;
    out[1] = uint32_t(S.inst[inst].blas->nodes.size());
    return 0;
}

// rays: 8 floats each (origin, tmin, dir, tmax); hits: 32-byte records.
int orc_trace(orc_scene* sp, const float* rays, uint32_t n, void* hits, uint32_t flags, int nthreads) {
    Scene& S = *reinterpret_cast<Scene*>(sp);
    Hit* H = reinterpret_cast<Hit*>(hits);
    parallel_for(n, nthreads, [&](uint32_t i) {
        float tmin, tmax;
        Ray<float> r = load_ray(rays + size_t(i) * 8, tmin, tmax);
        if (flags & F_FP64)
            trace_scene<double>(S, r, tmin, tmax, flags, H[i]);
        else
            trace_scene<float>(S, r, tmin, tmax, flags, H[i]);
    });
    return 0;
}

// N*_prim: number of primitives whose exact (unpadded) box overlaps [tmin, t_end] where
// t_end = hit t for hits, tmax for misses. Identity instances only.
int orc_count_candidates(orc_scene* sp, const float* rays, const void* hits, uint32_t n, uint32_t* counts,
                         int nthreads) {
    Scene& S = *reinterpret_cast<Scene*>(sp);
    const Hit* H = reinterpret_cast<const Hit*>(hits);
    parallel_for(n, nthreads, [&](uint32_t i) {
        float tmin, tmax;
        Ray<float> r = load_ray(rays + size_t(i) * 8, tmin, tmax);
        float tend = H[i].prim != 0xFFFFFFFFu ? H[i].t : tmax;
        uint32_t c = 0;
        for (uint32_t k = 0; k < S.inst.size(); ++k) {
            std::vector<uint32_t> cand;
            Hit dummy{};
            bool f = false;
            float te = tend;
            Ray<float> o;
            obj_ray(S.inst[k], r, o);
            trace_blas<float>(*S.inst[k].blas, k, o, tmin, te, 0, dummy, f, &cand);
            c += uint32_t(cand.size());
        }
        counts[i] = c;
    });
    return 0;
}

// FP32 operation census of the phantom intersector (test infrastructure; the oracle's own
// templates run with the counting float of oracle_count.hpp). For each ray against instance 0
// (a phantom BLAS), over the segments whose exact box overlaps [tmin, t_closest] (the
// candidate set of orc_count_candidates):
//   out[0] rays, out[1] FLOP of the ray frame (MakeOrthonormalBasis + the generic inverse of
//   CreateRCCMatrix, ray.glsl:13-33), summed over rays;
//   out[2] early-out tests, out[3] their FLOP (RayCylinderIntersect, cylinder.glsl:8-46, with
//   the precomputed radius rmax);
//   out[4] Prhi calls (candidates past the early-out), out[5] their FLOP after the early-out
//   (curve transform, start end, cone iterations, hit and normal; hair_intersection.rint:38-126,
//   the frame excluded), out[6] their cone iterations;
//   out[7..10] the compare/min/max/abs operations of the frame, early-out and Prhi classes
//   (not FLOP, reported apart).
int orc_flop_census(orc_scene* sp, const float* rays, uint32_t n, double* out) {
    Scene& S = *reinterpret_cast<Scene*>(sp);
    for (int i = 0; i < 11; ++i) out[i] = 0.0;
    if (S.inst.empty() || S.inst[0].blas->mode != MODE_PHANTOM) return -1;
    const Blas& B = *S.inst[0].blas;
    using orc::cnt::counts;
    for (uint32_t i = 0; i < n; ++i) {
        float tmin, tmax;
        Ray<float> rw = load_ray(rays + size_t(i) * 8, tmin, tmax);
        Ray<float> r;
        obj_ray(S.inst[0], rw, r);
        Hit best{};
        bool found = false;
        float tc = tmax;
        trace_blas<float>(B, 0, r, tmin, tc, 0, best, found, nullptr);
        std::vector<uint32_t> cand;
        Hit dummy{};
        bool f2 = false;
        float te = found ? best.t : tmax;
        trace_blas<float>(B, 0, r, tmin, te, 0, dummy, f2, &cand);
        const Ray<CountF> rc{cvt<CountF>(r.origin), cvt<CountF>(r.direction)};
        counts() = {};
        (void)CreateRCCMatrix(rc);
        const uint64_t frame = counts().flops();
        out[0] += 1;
        out[1] += double(frame);
        out[7] += double(counts().cmp);
        for (uint32_t p : cand) {
            const float* c = &B.prims[size_t(p) * 16];
            const Curve<float> cf{mk(c[0], c[1], c[2]), mk(c[4], c[5], c[6]), mk(c[8], c[9], c[10]), mk(c[12], c[13], c[14])};
            const float radf[4] = {c[3], c[7], c[11], c[15]};
            float rpad = B.radius;
            if (B.per_point) rpad = gmax(gmax(radf[0], radf[1]), gmax(radf[2], radf[3]));
            const float rmax = CurveDistanceToCylinder(cf, SampleCurvePoint(cf, 0.5f)) + rpad;
            const Curve<CountF> cc{cvt<CountF>(cf.start), cvt<CountF>(cf.cp1), cvt<CountF>(cf.cp2), cvt<CountF>(cf.end)};
            counts() = {};
            const bool pass = RayCylinderIntersect(rc, cc.start, cc.end, CountF(rmax));
            out[2] += 1;
            out[3] += double(counts().flops());
            out[8] += double(counts().cmp);
            if (!pass) continue;
            const CountF radc[4] = {CountF(radf[0]), CountF(radf[1]), CountF(radf[2]), CountF(radf[3])};
            PrhiOpts o;
            o.radius_mode = B.per_point ? RADIUS_PER_POINT : RADIUS_CONSTANT;
            o.radius = B.radius;
            o.cull = false;
            counts() = {};
            const PrhiOut<CountF> po = Prhi(rc, cc, radc, o);
            out[4] += 1;
            out[5] += double(counts().flops()) - double(frame);
            out[6] += double(po.iters);
            out[9] += double(counts().cmp);
        }
    }
    return 0;
}

// Primary + shadow render (ray_gen.rgen + closest hit + the shadow extension).
// flags: F_REF_NORMAL, F_POLISH. shadows: 0 = Shade(n) only, 1 = exact shadow ray.
// pixel_stride > 1 traces every stride-th pixel (row-major index) for bounded CPU timing.
// image: W*H*3 floats (untouched for skipped pixels); prim_ids: W*H (0xFFFFFFFF miss) or NULL.
// Returns the number of rays traced; *seconds gets the wall time of the tracing loop.
// miss.rmiss:17-38. DirectionToUV of the inverted direction: gamma = asin(v.y),
// theta = atan(v.x, -v.z), uv = (theta / 2pi, gamma / pi) + 0.5; the texture() lookup of the
// reference's default sampler (linear filter, repeat addressing, texel centres at
// (i + 0.5) / size, one mip level); then 1 - exp(-c exposure) and pow(c, 1 / 2.2).
static void EnvironmentColor(const float* env, uint32_t ew, uint32_t eh, float exposure, V3<float> dir, float rgb[3]) {
    const V3<float> v = normalize(mk(-dir.x, -dir.y, -dir.z));
    const float gamma = std::asin(v.y);
    const float theta = std::atan2(v.x, -v.z);
    const float u = theta * 0.3183098861837f * 0.5f + 0.5f;
    const float t = gamma * 0.3183098861837f + 0.5f;
    const float x = u * float(ew) - 0.5f, y = t * float(eh) - 0.5f;
    const float fx = std::floor(x), fy = std::floor(y);
    const float ax = x - fx, ay = y - fy;
    auto wrap = [](long i, long n) { long r = i % n; return r < 0 ? r + n : r; };
    const long x0 = wrap(long(fx), ew), x1 = wrap(long(fx) + 1, ew), y0 = wrap(long(fy), eh), y1 = wrap(long(fy) + 1, eh);
    auto tex = [&](long xx, long yy, int c) { return env[(size_t(yy) * ew + size_t(xx)) * 4 + c]; };
    auto lerp = [](float a, float b, float s) { return a + (b - a) * s; };
    for (int c = 0; c < 3; ++c) {
        const float val = lerp(lerp(tex(x0, y0, c), tex(x1, y0, c), ax), lerp(tex(x0, y1, c), tex(x1, y1, c), ax), ay);
        rgb[c] = std::pow(1.0f - std::exp(-val * exposure), 1.0f / 2.2f);
    }
}

// area_light (optional, 9 floats: center, edge u, edge v): `light_samples` shadow rays per hit
// toward center + a u + b v, a = 2 path_rand(seed, pixel, s, 0) - 1, b likewise with k = 1,
// tmax = distance; the visibility is their mean (sum, then / samples). NULL: one ray along
// (0, 1, 0).
// env (optional): an equirectangular RGBA32F map of ew x eh texels, the miss colour (miss.rmiss).
long long orc_render(orc_scene* sp, const float view_inv[16], const float proj_inv[16], uint32_t W, uint32_t H,
                     uint32_t flags, int shadows, float* image, uint32_t* prim_ids, int nthreads, uint32_t pixel_stride,
                     double* seconds, const float* area_light, uint32_t light_samples, uint64_t light_seed,
                     const float* env, uint32_t ew, uint32_t eh, float exposure) {
    Scene& S = *reinterpret_cast<Scene*>(sp);
    M4<float> Vi, Pi;
    std::memcpy(Vi.m, view_inv, 64);
    std::memcpy(Pi.m, proj_inv, 64);
    if (pixel_stride == 0) pixel_stride = 1;
    uint32_t npix = W * H, nsample = (npix + pixel_stride - 1) / pixel_stride;
    std::atomic<long long> nrays{0};
    auto t0 = std::chrono::steady_clock::now();
    parallel_for(nsample, nthreads, [&](uint32_t si) {
        uint32_t pix = si * pixel_stride;
        uint32_t x = pix % W, y = pix / W;
        V3<float> o, d;
        RayGen(Vi, Pi, x, y, W, H, o, d);
        Ray<float> r{o, d};
        Hit h;
        long long nr = 1;
        bool hit = trace_scene<float>(S, r, 0.001f, 10000.0f, flags & (F_REF_NORMAL | F_POLISH | F_ACCURATE), h);
        float rgb[3] = {0.0f, 0.0f, 0.0f};
        if (!hit && env) EnvironmentColor(env, ew, eh, exposure, d, rgb);
        if (hit) {
            float vis = 1.0f;
            if (shadows && !area_light) {
                V3<float> hp = o + h.t * d;
                Ray<float> sr{hp, mk(0.0f, 1.0f, 0.0f)};
                Hit sh;
                ++nr;
                if (trace_scene<float>(S, sr, 0.001f, 10000.0f, F_ANY_HIT | (flags & F_POLISH), sh)) vis = 0.0f;
            } else if (shadows) {
                const V3<float> hp = o + h.t * d;
                const float* L = area_light;
                float sum = 0.0f;
                for (uint32_t k = 0; k < light_samples; ++k) {
                    const float a = 2.0f * path_rand(light_seed, pix, k, 0) - 1.0f;
                    const float b = 2.0f * path_rand(light_seed, pix, k, 1) - 1.0f;
                    const V3<float> y = mk(L[0] + a * L[3] + b * L[6], L[1] + a * L[4] + b * L[7], L[2] + a * L[5] + b * L[8]);
                    const V3<float> w = y - hp;
                    Hit sh;
                    ++nr;
                    const bool occ = trace_scene<float>(S, Ray<float>{hp, normalize(w)}, 0.001f, std::sqrt(dot(w, w)),
                                                        F_ANY_HIT | (flags & F_POLISH), sh);
                    sum = k == 0 ? (occ ? 0.0f : 1.0f) : sum + (occ ? 0.0f : 1.0f);
                }
                vis = light_samples > 1 ? sum / float(light_samples) : sum;
            }
            V3<float> c = Shade(mk(h.n[0], h.n[1], h.n[2]), vis);
            rgb[0] = c.x;
            rgb[1] = c.y;
            rgb[2] = c.z;
        }
        nrays += nr;
        float* px = image + size_t(pix) * 3;
        px[0] = rgb[0];
        px[1] = rgb[1];
        px[2] = rgb[2];
        if (prim_ids) prim_ids[pix] = hit ? h.prim : 0xFFFFFFFFu;
    });
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    return nrays.load();
}

// ---- batched path tracing (oracle_path.hpp). Views: cams = n x 32 floats (view inverse,
// projection inverse, column-major), sizes = n x 2 (width, height); paths are the pixels of
// the views in order, row-major within a view. albedo: 16 x 3 tints. rgb: 3 floats per path,
// prim_ids: primary hit per path (0xFFFFFFFF = miss). Only every `path_stride`-th path is
// computed (the others are left untouched). counts[0] += bounce rays, counts[1] += shadow rays.
long long orc_render_batch(orc_scene* sp, const float* cams, const uint32_t* sizes, uint32_t nviews, uint32_t bounces,
                           uint64_t seed, const float env[3], const float* albedo, uint32_t flags, float* rgb,
                           uint32_t* prim_ids, unsigned long long* counts, int nthreads, uint32_t path_stride,
                           double* seconds) {
    Scene& S = *reinterpret_cast<Scene*>(sp);
    std::vector<uint64_t> first(nviews + 1, 0);
    for (uint32_t v = 0; v < nviews; ++v) first[v + 1] = first[v] + uint64_t(sizes[2 * v]) * sizes[2 * v + 1];
    const uint64_t total = first[nviews];
    if (path_stride == 0) path_stride = 1;
    const uint32_t nsample = uint32_t((total + path_stride - 1) / path_stride);
    std::atomic<long long> nrays{0};
    std::atomic<unsigned long long> nbounce{0}, nshadow{0};
    const uint32_t tflags = flags & (F_REF_NORMAL | F_POLISH);
    auto t0 = std::chrono::steady_clock::now();
    parallel_for(nsample, nthreads, [&](uint32_t si) {
        const uint64_t p = uint64_t(si) * path_stride;
        const uint32_t v = uint32_t(std::upper_bound(first.begin(), first.end(), p) - first.begin()) - 1;
        const uint32_t W = sizes[2 * v], H = sizes[2 * v + 1];
        const uint32_t local = uint32_t(p - first[v]);
        M4<float> Vi, Pi;
        std::memcpy(Vi.m, cams + size_t(v) * 32, 64);
        std::memcpy(Pi.m, cams + size_t(v) * 32 + 16, 64);
        V3<float> o, d;
        RayGen(Vi, Pi, local % W, local / W, W, H, o, d);
        V3<float> T = mk(1.0f, 1.0f, 1.0f), L = mk(0.0f, 0.0f, 0.0f);
        uint32_t prim0 = 0xFFFFFFFFu;
        long long nr = 0;
        unsigned long long nb = 0, ns = 0;
        for (uint32_t b = 0; b <= bounces; ++b) {
            Hit h;
            ++nr;
            const bool hit = trace_scene<float>(S, Ray<float>{o, d}, 0.001f, 10000.0f, tflags, h);
            if (b == 0 && hit) prim0 = h.prim;
            if (!hit) {
                L = mk(L.x + T.x * env[0], L.y + T.y * env[1], L.z + T.z * env[2]);
                break;
            }
            const V3<float> x = o + h.t * d;
            const V3<float> n = mk(h.n[0], h.n[1], h.n[2]);
            const V3<float> nf = dot(n, d) > 0.0f ? mk(-n.x, -n.y, -n.z) : n;
            const V3<float> s0 = Shade(n, 1.0f);
            const float* tint = albedo + 3 * std::min<uint32_t>(S.inst[h.inst].material, 15u);
            const V3<float> a = mk(tint[0] * s0.x, tint[1] * s0.y, tint[2] * s0.z);
            const float c = nf.y;
            if (c > 0.0f) {
                const V3<float> nee = mk(T.x * a.x * c, T.y * a.y * c, T.z * a.z * c);
                Hit sh;
                ++nr;
                ++ns;
                if (!trace_scene<float>(S, Ray<float>{x, mk(0.0f, 1.0f, 0.0f)}, 0.001f, 10000.0f,
                                        F_ANY_HIT | (flags & F_POLISH), sh))
                    L = mk(L.x + nee.x, L.y + nee.y, L.z + nee.z);
            }
            if (b == bounces) break;
            d = path_cosine_dir(nf, seed, uint32_t(p), b);
            T = mk(T.x * a.x, T.y * a.y, T.z * a.z);
            o = x;
            ++nb;
        }
        rgb[p * 3 + 0] = L.x;
        rgb[p * 3 + 1] = L.y;
        rgb[p * 3 + 2] = L.z;
        if (prim_ids) prim_ids[p] = prim0;
        nrays += nr;
        nbounce += nb;
        nshadow += ns;
    });
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    if (counts) {
        counts[0] += nbounce.load();
        counts[1] += nshadow.load();
    }
    return nrays.load();
}

float orc_path_rand(uint64_t seed, uint32_t p, uint32_t depth, uint32_t k) { return path_rand(seed, p, depth, k); }

int orc_path_basis(const float* n, float* t1, float* t2) {
    V3<float> a, b;
    path_basis(mk(n[0], n[1], n[2]), a, b);
    t1[0] = a.x, t1[1] = a.y, t1[2] = a.z;
    t2[0] = b.x, t2[1] = b.y, t2[2] = b.z;
    return 0;
}

// Directions for paths p = p0 .. p0 + count - 1 at one depth (count x 3 floats).
int orc_path_cosine_dirs(const float* n, uint64_t seed, uint32_t p0, uint32_t count, uint32_t depth, float* out) {
    for (uint32_t i = 0; i < count; ++i) {
        const V3<float> w = path_cosine_dir(mk(n[0], n[1], n[2]), seed, p0 + i, depth);
        out[3 * i] = w.x, out[3 * i + 1] = w.y, out[3 * i + 2] = w.z;
    }
    return 0;
}

// ---- neural shadow MLP (oracle_mlp.hpp). samples: n x 8 floats (p.xyz, 0, d.xyz, 0).
// fp64 = 1: exact double function without bf16 rounding (finite-difference checks);
// fp64 = 0: FP32 with the GPU kernels' bf16 roundings.
int orc_mlp_forward(const float* params, const float center[3], const float inv_half[3], const float* samples,
                    uint32_t n, int fp64, float* out_p) {
    if (fp64) {
        std::vector<double> P(params, params + kMlpParams);
        for (uint32_t s = 0; s < n; ++s) {
            MlpRef<double, false>::Acts a;
            MlpRef<double, false>::forward(P, samples + size_t(s) * 8, center, inv_half, a);
            out_p[s] = float(a.p);
        }
    } else {
        std::vector<float> P(params, params + kMlpParams);
        for (uint32_t s = 0; s < n; ++s) {
            MlpRef<float, true>::Acts a;
            MlpRef<float, true>::forward(P, samples + size_t(s) * 8, center, inv_half, a);
            out_p[s] = a.p;
        }
    }
    return 0;
}

// Gradient of the mean binary cross-entropy; params / grad in the flat layout. fp64 = 1
// takes and returns doubles.
int orc_mlp_grad(const void* params, const float center[3], const float inv_half[3], const float* samples,
                 const float* targets, uint32_t n, int fp64, void* grad, double* loss) {
    if (n == 0) return -1;
    if (fp64) {
        const double* pp = static_cast<const double*>(params);
        std::vector<double> P(pp, pp + kMlpParams), g;
        const double l = MlpRef<double, false>::grad(P, samples, targets, n, center, inv_half, g);
        std::memcpy(grad, g.data(), sizeof(double) * kMlpParams);
        if (loss) *loss = l;
    } else {
        const float* pp = static_cast<const float*>(params);
        std::vector<float> P(pp, pp + kMlpParams), g;
        const double l = MlpRef<float, true>::grad(P, samples, targets, n, center, inv_half, g);
        std::memcpy(grad, g.data(), sizeof(float) * kMlpParams);
        if (loss) *loss = l;
    }
    return 0;
}

int orc_mlp_param_count(void) { return kMlpParams; }

// The kernels' FP32 sin(pi v), cos(pi v) of the MLP features, elementwise.
int orc_sincospi_f32(const float* v, uint32_t n, float* s, float* c) {
    for (uint32_t i = 0; i < n; ++i) sincospi_f32(v[i], s[i], c[i]);
    return 0;
}

}  // extern "C"
