"""ctypes bindings to the oracle (oracle/build/liboracle.so).

Test infrastructure: used by tests/, __graft_entry__.smoke() and bench.py's CPU
baseline only. The product (vkhrt_amd) never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "build", "liboracle.so")

MODE_PHANTOM, MODE_LSS, MODE_DOTS, MODE_VOXELS = 0, 1, 2, 4
F_ANY_HIT, F_BRUTE, F_FP64, F_POLISH, F_REF_NORMAL, F_ACCURATE = 1, 2, 4, 8, 16, 32

HIT_DTYPE = np.dtype([("t", "<f4"), ("prim", "<u4"), ("inst", "<u4"), ("u", "<f4"),
                      ("n", "<f4", (3,)), ("v", "<f4")])
assert HIT_DTYPE.itemsize == 32

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    f32, u32, i32, u64 = ctypes.c_float, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint64
    sig = {
        "orc_abi_version": ([], i32),
        "orc_hardware_threads": ([], i32),
        "orc_generate_strands": ([u64, u32, u32, f32, P, i32], i32),
        "orc_strands_to_lines": ([P, u32, u32, P], i32),
        "orc_generate_curves": ([P, u32, P], i32),
        "orc_mesh_lines": ([P, u32, P, u32, f32, P], i32),
        "orc_world_matrix": ([P, P, u32, P], i32),
        "orc_voxelize": ([P, u32, P, P, f32, f32, i32, u32, P, P, P], ctypes.c_longlong),
        "orc_scene_add_voxels": ([P, P, P, P, f32, P], i32),
        "orc_curves_to_lines": ([P, u32, u32, P], i32),
        "orc_generate_lss": ([P, u32, f32, P], i32),
        "orc_generate_dots": ([P, u32, f32, P], i32),
        "orc_dots_triangles": ([P, u32, P], i32),
        "orc_curve_aabbs": ([P, u32, i32, f32, P], i32),
        "orc_camera": ([P, P, P, f32, f32, f32, f32, P, P], i32),
        "orc_raygen": ([P, P, u32, u32, f32, f32, P], i32),
        "orc_prhi": ([P, P, i32, f32, i32, i32, P], i32),
        "orc_prhi_f64": ([P, P, i32, f32, P], i32),
        "orc_rcc_matrix": ([P, i32, P], i32),
        "orc_inverse4": ([P, P], i32),
        "orc_basis": ([P, P, P], i32),
        "orc_sample_curve": ([P, f32, i32, P, P], i32),
        "orc_cylinder": ([P, P, P, f32], i32),
        "orc_cone": ([P, f32, P, f32, P], i32),
        "orc_lss": ([P, P, i32, i32, P], i32),
        "orc_triangle": ([P, P, P], i32),
        "orc_det2": ([f32, f32, f32, f32], f32),
        "orc_perp_stark": ([P, P], i32),
        "orc_shade": ([P, f32, P], i32),
        "orc_debug_color": ([ctypes.c_int32, P], i32),
        "orc_scene_create": ([], P),
        "orc_scene_destroy": ([P], None),
        "orc_scene_add_instance": ([P, i32, P, u32, i32, f32, P, i32], i32),
        "orc_scene_stats": ([P, u32, P], i32),
        "orc_trace": ([P, P, u32, P, u32, i32], i32),
        "orc_count_candidates": ([P, P, P, u32, P, i32], i32),
        "orc_flop_census": ([P, P, u32, P], i32),
        "orc_render": ([P, P, P, u32, u32, u32, i32, P, P, i32, u32, P, P, u32, u64, P, u32, u32, f32],
                       ctypes.c_longlong),
        "orc_scene_set_material": ([P, u32, u32], i32),
        "orc_scene_add_mesh": ([P, P, u32, P, u32, P, i32], i32),
        "orc_render_batch": ([P, P, P, u32, u32, u64, P, P, u32, P, P, P, i32, u32, P], ctypes.c_longlong),
        "orc_path_rand": ([u64, u32, u32, u32], f32),
        "orc_path_basis": ([P, P, P], i32),
        "orc_path_cosine_dirs": ([P, u64, u32, u32, u32, P], i32),
        "orc_mlp_forward": ([P, P, P, P, u32, i32, P], i32),
        "orc_mlp_grad": ([P, P, P, P, P, u32, i32, P, ctypes.POINTER(ctypes.c_double)], i32),
        "orc_mlp_param_count": ([], i32),
        "orc_sincospi_f32": ([P, u32, P, P], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


# ---------------------------------------------------------------- scene construction
def generate_strands(seed, nstrands, nseg, sphere_radius=100.0, nthreads=0):
    out = np.empty((nstrands, nseg + 1, 4), np.float32)
    lib().orc_generate_strands(seed, nstrands, nseg, sphere_radius, ptr(out), nthreads)
    return out


def strands_to_lines(verts):
    ns, nv, _ = verts.shape
    lines = np.empty((ns * (nv - 1), 8), np.float32)
    lib().orc_strands_to_lines(ptr(f32(verts)), ns, nv - 1, ptr(lines))
    return lines


def generate_curves(lines):
    lines = f32(lines)
    out = np.empty((lines.shape[0], 16), np.float32)
    lib().orc_generate_curves(ptr(lines), lines.shape[0], ptr(out))
    return out


def mesh_lines(verts, indices, radius=0.02):
    """GenerateLines of an indexed line mesh: (n, 8) lines from 8-float vertices and mesh-local
    index pairs; raises on an index past the vertices."""
    verts = np.ascontiguousarray(verts, np.float32).reshape(-1, 8)
    idx = np.ascontiguousarray(indices, np.uint32).reshape(-1)
    n = idx.size // 2
    out = np.zeros((n, 8), np.float32)
    if lib().orc_mesh_lines(ptr(verts), verts.shape[0], ptr(idx), n, radius, ptr(out)) != 0:
        raise ValueError("line index past the vertex count")
    return out


def world_matrices(locals_, parents):
    """Node::GetWorldMatrix for every node: (n, 16) column-major float32."""
    loc = np.ascontiguousarray(locals_, np.float32).reshape(-1, 16)
    par = np.ascontiguousarray(parents, np.int32)
    out = np.zeros_like(loc)
    if lib().orc_world_matrix(ptr(loc), ptr(par), loc.shape[0], ptr(out)) != 0:
        raise ValueError("parent cycle")
    return out


def voxelize(lines, voxel_size, radius=0.02, bounds=None, reference_sweep=False, lod=0):
    """GenerateVoxelMesh (+ lod OR-reductions): (bits uint32 words, res (3,), origin (3,), cell)."""
    lines = np.ascontiguousarray(lines, np.float32).reshape(-1, 8)
    lo = hi = None
    if bounds is not None:
        lo, hi = f32(bounds[0]), f32(bounds[1])
    res = np.zeros(3, np.int32)
    geo = np.zeros(4, np.float32)
    args = (ptr(lines), lines.shape[0], ptr(lo), ptr(hi), radius, voxel_size, int(reference_sweep), lod)
    nw = lib().orc_voxelize(*args, None, ptr(res), ptr(geo))
    if nw < 0:
        raise ValueError("invalid voxel grid" if nw == -1 else "a line covers more than 65536 cells")
    bits = np.zeros(max(nw, 1), np.uint32)
    lib().orc_voxelize(*args, ptr(bits), ptr(res), ptr(geo))
    return bits[:nw], res, geo[:3].copy(), float(geo[3])


def curves_to_lines(curves, pieces):
    curves = f32(curves)
    out = np.empty((curves.shape[0] * pieces, 8), np.float32)
    lib().orc_curves_to_lines(ptr(curves), curves.shape[0], pieces, ptr(out))
    return out


def generate_lss(lines, constant_radius=-1.0):
    lines = f32(lines)
    out = np.empty((lines.shape[0], 8), np.float32)
    lib().orc_generate_lss(ptr(lines), lines.shape[0], constant_radius, ptr(out))
    return out


def generate_dots(lines, constant_radius=-1.0):
    lines = f32(lines)
    out = np.empty((lines.shape[0], 16), np.float32)
    lib().orc_generate_dots(ptr(lines), lines.shape[0], constant_radius, ptr(out))
    return out


def dots_triangles(pieces):
    pieces = f32(pieces)
    out = np.empty((pieces.shape[0] * 4, 3, 3), np.float32)
    lib().orc_dots_triangles(ptr(pieces), pieces.shape[0], ptr(out))
    return out


def curve_aabbs(curves, per_point=False, radius=0.02):
    curves = f32(curves)
    out = np.empty((curves.shape[0], 6), np.float32)
    lib().orc_curve_aabbs(ptr(curves), curves.shape[0], int(per_point), radius, ptr(out))
    return out


def hair_curves(seed, nstrands, nseg):
    """Seeded strands -> lines -> GenerateCurves: (nstrands*nseg, 16) float32."""
    return generate_curves(strands_to_lines(generate_strands(seed, nstrands, nseg)))


def camera(eye, center, up=(0.0, 1.0, 0.0), tan_half_fovy=0.5773502691896257, aspect=1.0,
           znear=0.1, zfar=1000.0):
    vi = np.empty(16, np.float32)
    pi = np.empty(16, np.float32)
    lib().orc_camera(ptr(f32(eye)), ptr(f32(center)), ptr(f32(up)), tan_half_fovy, aspect,
                     znear, zfar, ptr(vi), ptr(pi))
    return vi, pi


def raygen(view_inv, proj_inv, W, H, tmin=0.001, tmax=10000.0):
    rays = np.empty((H * W, 8), np.float32)
    lib().orc_raygen(ptr(f32(view_inv)), ptr(f32(proj_inv)), W, H, tmin, tmax, ptr(rays))
    return rays


class Scene:
    """CPU scene: instances of BLASes, traced with a binned-SAH BVH2 or brute force."""

    def __init__(self):
        self.h = lib().orc_scene_create()
        self._keep = []

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_destroy(self.h)
            self.h = None

    def add(self, mode, prims, per_point=False, radius=0.02, transform=None, leaf_size=4):
        prims = f32(prims)
        xf = None if transform is None else f32(np.asarray(transform).reshape(12))
        r = lib().orc_scene_add_instance(self.h, mode, ptr(prims), prims.shape[0], int(per_point),
                                         radius, ptr(xf), leaf_size)
        assert r >= 0
        return r

    def trace(self, rays, flags=0, nthreads=0):
        rays = f32(rays)
        hits = np.zeros(rays.shape[0], HIT_DTYPE)
        lib().orc_trace(self.h, ptr(rays), rays.shape[0], ptr(hits), flags, nthreads)
        return hits

    def flop_census(self, rays):
        """FP32 operations of the phantom intersector on these rays against instance 0 (the
        counting build of the oracle's templates): dict of per-class calls and FLOP."""
        rays = f32(rays)
        out = np.zeros(11, np.float64)
        if lib().orc_flop_census(self.h, ptr(rays), rays.shape[0], out.ctypes.data_as(ctypes.c_void_p)) != 0:
            raise ValueError("instance 0 is not a phantom BLAS")
        return {"rays": out[0], "frame_flop": out[1], "early_out_tests": out[2], "early_out_flop": out[3],
                "prhi_calls": out[4], "prhi_flop": out[5], "prhi_iterations": out[6], "frame_cmp": out[7],
                "early_out_cmp": out[8], "prhi_cmp": out[9]}

    def count_candidates(self, rays, hits, nthreads=0):
        rays = f32(rays)
        counts = np.zeros(rays.shape[0], np.uint32)
        lib().orc_count_candidates(self.h, ptr(rays), ptr(hits), rays.shape[0], ptr(counts), nthreads)
        return counts

    def add_mesh(self, verts, indices, transform=None, leaf_size=4):
        """verts: (nv, 8) position, normal, texCoord; indices: (nt, 3) u32."""
        verts = f32(verts)
        idx = np.ascontiguousarray(indices, dtype=np.uint32)
        xf = None if transform is None else f32(np.asarray(transform).reshape(12))
        r = lib().orc_scene_add_mesh(self.h, ptr(verts), verts.shape[0], ptr(idx), idx.shape[0], ptr(xf), leaf_size)
        assert r >= 0, "mesh index out of range"
        return r

    def add_voxels(self, bits, res, origin, cell, transform=None):
        bits = np.ascontiguousarray(bits, np.uint32)
        if bits.size == 0:
            bits = np.zeros(1, np.uint32)
        xf = None if transform is None else f32(np.asarray(transform).reshape(12))
        r = lib().orc_scene_add_voxels(self.h, ptr(bits), ptr(np.ascontiguousarray(res, np.int32)), ptr(f32(origin)),
                                       cell, ptr(xf))
        assert r >= 0
        return r

    def set_material(self, inst, material):
        assert lib().orc_scene_set_material(self.h, inst, material) == 0

    def render_batch(self, views, bounces, seed, env=(0.0, 0.0, 0.0), albedo=None, flags=0, nthreads=0,
                     path_stride=1):
        """Batched path tracing: views = [(view_inv, proj_inv, W, H)]. Returns (per-view float
        images (H, W, 3), per-view primary ids, bounce rays, shadow rays, seconds)."""
        cams = np.concatenate([np.concatenate([f32(vi).reshape(16), f32(pi).reshape(16)]) for vi, pi, _, _ in views])
        sizes = np.array([[w, h] for _, _, w, h in views], np.uint32)
        total = int((sizes[:, 0].astype(np.int64) * sizes[:, 1]).sum())
        tints = np.ones((16, 3), np.float32)
        if albedo is not None:
            a = f32(albedo).reshape(-1, 3)
            tints[:a.shape[0]] = a
        rgb = np.zeros((total, 3), np.float32)
        ids = np.full(total, 0xFFFFFFFF, np.uint32)
        counts = np.zeros(2, np.uint64)
        secs = ctypes.c_double(0)
        lib().orc_render_batch(self.h, ptr(cams), ptr(sizes), len(views), bounces, seed, ptr(f32(env)), ptr(tints),
                               flags, ptr(rgb), ptr(ids), ptr(counts), nthreads, path_stride, ctypes.byref(secs))
        imgs, pids, o = [], [], 0
        for w, h in sizes:
            n = int(w) * int(h)
            imgs.append(rgb[o:o + n].reshape(int(h), int(w), 3))
            pids.append(ids[o:o + n].reshape(int(h), int(w)))
            o += n
        return imgs, pids, int(counts[0]), int(counts[1]), secs.value

    def render(self, view_inv, proj_inv, W, H, flags=0, shadows=1, nthreads=0, pixel_stride=1, area_light=None,
               light_samples=1, light_seed=0, env=None, exposure=1.0):
        """area_light: (center, edge_u, edge_v) of a parallelogram light, light_samples rays per hit.
        env: (h, w, 4) float32 equirectangular map coloring the misses."""
        img = np.zeros((H, W, 3), np.float32)
        ids = np.full((H, W), 0xFFFFFFFF, np.uint32)
        secs = ctypes.c_double(0)
        light = None if area_light is None else f32(np.asarray(area_light, np.float32).reshape(9))
        nrays = lib().orc_render(self.h, ptr(f32(view_inv)), ptr(f32(proj_inv)), W, H, flags,
                                 shadows, ptr(img), ptr(ids), nthreads, pixel_stride,
                                 ctypes.byref(secs), ptr(light), light_samples, light_seed,
                                 ptr(None if env is None else f32(env)), 0 if env is None else env.shape[1],
                                 0 if env is None else env.shape[0], exposure)
        return img, ids, int(nrays), secs.value


def prhi(ray, curve, per_point=False, radius=0.02, polish=False, fp64=False):
    out = np.zeros(10, np.float32)
    lib().orc_prhi(ptr(f32(ray)), ptr(f32(curve)), int(per_point), radius, int(polish),
                   int(fp64), ptr(out))
    return out


def prhi_f64(ray, curve, per_point=False, radius=0.02):
    out = np.zeros(10, np.float64)
    lib().orc_prhi_f64(ptr(f32(ray)), ptr(f32(curve)), int(per_point), radius, ptr(out))
    return out


# ---- neural shadow MLP
MLP_PARAMS = 18753
MLP_OFF = {"W1": 0, "b1": 6144, "W2": 6208, "b2": 10304, "W3": 10368, "b3": 14464, "W4": 14528,
           "b4": 18624, "w5": 18688, "b5": 18752}


def mlp_forward(params, center, inv_half, samples, fp64=False):
    samples = f32(samples)
    out = np.zeros(samples.shape[0], np.float32)
    lib().orc_mlp_forward(ptr(f32(params)), ptr(f32(center)), ptr(f32(inv_half)), ptr(samples),
                          samples.shape[0], int(fp64), ptr(out))
    return out


def mlp_grad(params, center, inv_half, samples, targets, fp64=False):
    samples = f32(samples)
    dt = np.float64 if fp64 else np.float32
    P = np.ascontiguousarray(params, dt)
    g = np.zeros(MLP_PARAMS, dt)
    loss = ctypes.c_double(0)
    lib().orc_mlp_grad(P.ctypes.data_as(ctypes.c_void_p), ptr(f32(center)), ptr(f32(inv_half)), ptr(samples),
                       ptr(f32(targets)), samples.shape[0], int(fp64), g.ctypes.data_as(ctypes.c_void_p),
                       ctypes.byref(loss))
    return g, loss.value


def mlp_init(seed):
    """He-uniform weights, zero biases (test parameters)."""
    rng = np.random.default_rng(seed)
    P = np.zeros(MLP_PARAMS, np.float32)
    for name, fan_in, size in [("W1", 96, 6144), ("W2", 64, 4096), ("W3", 64, 4096), ("W4", 64, 4096), ("w5", 64, 64)]:
        lim = np.sqrt(6.0 / fan_in)
        P[MLP_OFF[name]:MLP_OFF[name] + size] = rng.uniform(-lim, lim, size)
    return P


def path_rand(seed, p, depth, k):
    return lib().orc_path_rand(seed, p, depth, k)


def path_basis(n):
    t1 = np.zeros(3, np.float32)
    t2 = np.zeros(3, np.float32)
    lib().orc_path_basis(ptr(f32(n)), ptr(t1), ptr(t2))
    return t1, t2


def path_cosine_dirs(n, seed, p0, count, depth=0):
    out = np.zeros((count, 3), np.float32)
    lib().orc_path_cosine_dirs(ptr(f32(n)), seed, p0, count, depth, ptr(out))
    return out
