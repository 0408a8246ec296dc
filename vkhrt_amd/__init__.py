"""vkhrt_amd: MI355X-native hair ray-tracing core behind a C-ABI (include/vkhrt.h).

This module binds libvkhrt.so with ctypes and mirrors the reference's host interface:
``BottomLevelAccelerationStructure`` (include/bottom_level_acceleration_structure.hpp:31-50),
``TopLevelAccelerationStructure`` (include/top_level_acceleration_structure.hpp:9-23) and the
``traceRaysKHR`` dispatch (source/renderer.cpp:156-166) as ``Context.render``.

Device buffers are passed as integer device addresses (e.g. ``torch.Tensor.data_ptr()``)
and streams as ``hipStream_t`` integers (``torch.cuda.Stream.cuda_stream``). There is no
CPU fallback: every call goes to the HIP library, and loading fails loudly if it is absent.
"""
import ctypes
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# VKHRT_LIBRARY selects another build of the library, e.g. the counting build
# libvkhrt_count.so (work counters for measurement; same kernels and results)
LIB_PATH = os.environ.get("VKHRT_LIBRARY") or os.path.join(HERE, "libvkhrt.so")

GEOMETRY_PHANTOM_CURVES, GEOMETRY_LSS, GEOMETRY_DOTS, GEOMETRY_TRIANGLES, GEOMETRY_VOXELS = 0, 1, 2, 3, 4

BLAS_HOST_POINTERS = 1 << 0
BLAS_PER_POINT_RADIUS = 1 << 1
BLAS_HAS_TRANSFORM = 1 << 2
BLAS_HAS_BOUNDS = 1 << 3
BLAS_VOXEL_REFERENCE_SWEEP = 1 << 4

TRACE_TERMINATE_ON_FIRST_HIT = 1 << 0
TRACE_FP64_POLISH = 1 << 3
TRACE_ACCURATE_NORMALS = 1 << 5

RENDER_SHADOWS = 1 << 0
RENDER_DEBUG_PALETTE = 1 << 1
RENDER_RGBA8 = 1 << 2
RENDER_FP64_POLISH = 1 << 3
RENDER_NEURAL_SHADOWS = 1 << 4
RENDER_ACCURATE_NORMALS = 1 << 5

SHADOW_MLP_PARAMS = 18753

ABI_VERSION = 5  # VKHRT_ABI_VERSION of include/vkhrt.h: the struct layouts below

OPTION_STACK_KIND = 1
OPTION_LEAF_SIZE = 2
OPTION_PERSISTENT = 3
OPTION_REFILL_MIN = 4
OPTION_RAY_ORDER = 5
OPTION_BLOCKS_PER_CU = 6
OPTION_BVH_WIDTH = 7
OPTION_BVH_COLLAPSE = 9
OPTION_XCD_CHUNK = 8
OPTION_SHADOW_WAVE_LIMIT = 10
OPTION_SHADOW_GROUP_LANES = 11
OPTION_STEPS_PER_ITERATION = 12

BATCH_MAX_BOUNCES = 8
BATCH_MAX_MATERIALS = 16
BATCH_RGBA8 = 1 << 2
BATCH_FP64_POLISH = 1 << 3
BATCH_SORT_MATERIALS = 1 << 5
BATCH_GRAPH = 1 << 6
BATCH_SORT_RAYS = 1 << 7

PRIMITIVE_TRIANGLES, PRIMITIVE_LINES = 0, 1

STATUS = {0: "SUCCESS", -1: "INVALID_ARGUMENT", -2: "OUT_OF_MEMORY", -3: "DEVICE",
          -4: "UNSUPPORTED", -5: "STACK_OVERFLOW", -6: "NO_DEVICE"}

# Every symbol declared in include/vkhrt.h.
EXPORTS = [
    "vkhrt_abi_version", "vkhrt_context_create", "vkhrt_context_destroy", "vkhrt_last_error",
    "vkhrt_synchronize", "vkhrt_set_option", "vkhrt_blas_build", "vkhrt_blas_info_get",
    "vkhrt_blas_destroy", "vkhrt_tlas_build", "vkhrt_tlas_destroy", "vkhrt_trace_rays", "vkhrt_trace_rays_noalloc", "vkhrt_work_counters",
    "vkhrt_render", "vkhrt_camera_look_at", "vkhrt_generate_strands", "vkhrt_generate_lines",
    "vkhrt_generate_curves", "vkhrt_curves_to_lines", "vkhrt_generate_lss", "vkhrt_generate_dots",
    "vkhrt_profiling_enable", "vkhrt_kernel_stats", "vkhrt_status_counters",
    "vkhrt_generate_camera_rays", "vkhrt_generate_shadow_rays", "vkhrt_count_candidates",
    "vkhrt_render_ray_counts", "vkhrt_trace_stats", "vkhrt_render_stats_get",
    "vkhrt_shadow_mlp_create", "vkhrt_shadow_mlp_destroy", "vkhrt_shadow_mlp_get_params",
    "vkhrt_shadow_mlp_set_params", "vkhrt_shadow_mlp_collect", "vkhrt_shadow_mlp_train",
    "vkhrt_shadow_mlp_gradient", "vkhrt_shadow_mlp_infer", "vkhrt_render_batch", "vkhrt_batch_stats_get",
    "vkhrt_autotune_render", "vkhrt_autotune_batch", "vkhrt_autotune_blas", "vkhrt_generate_surface_shadow_rays",
    "vkhrt_asset_load", "vkhrt_asset_load_memory", "vkhrt_asset_destroy", "vkhrt_asset_last_error",
    "vkhrt_asset_info_get", "vkhrt_asset_mesh_get", "vkhrt_asset_node_get", "vkhrt_asset_material_get",
    "vkhrt_asset_instance_get", "vkhrt_asset_build", "vkhrt_generate_mesh_lines", "vkhrt_blas_voxel_grid_get",
    "vkhrt_blas_voxels_read",
]


class VkhrtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"vkhrt {STATUS.get(code, code)}: {msg}")
        self.code = code


class BlasDesc(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_uint32), ("flags", ctypes.c_uint32),
                ("primitives", ctypes.c_void_p), ("count", ctypes.c_uint32),
                ("radius", ctypes.c_float), ("material_index", ctypes.c_uint32),
                ("leaf_size", ctypes.c_uint32), ("transform", ctypes.c_float * 12),
                ("indices", ctypes.c_void_p), ("vertex_count", ctypes.c_uint32),
                ("voxel_size", ctypes.c_float), ("voxel_lod", ctypes.c_uint32),
                ("bounds_min", ctypes.c_float * 3), ("bounds_max", ctypes.c_float * 3)]


class BlasInfo(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_uint32), ("primitive_count", ctypes.c_uint32),
                ("leaf_records", ctypes.c_uint32), ("node_count", ctypes.c_uint32),
                ("depth", ctypes.c_uint32), ("device_bytes", ctypes.c_uint64),
                ("build_ms", ctypes.c_double)]


class Camera(ctypes.Structure):
    _fields_ = [("view_inverse", ctypes.c_float * 16), ("proj_inverse", ctypes.c_float * 16)]


class AreaLight(ctypes.Structure):
    _fields_ = [("center", ctypes.c_float * 3), ("edge_u", ctypes.c_float * 3), ("edge_v", ctypes.c_float * 3),
                ("samples", ctypes.c_uint32), ("seed", ctypes.c_uint64)]


def area_light(center, edge_u, edge_v, samples=2, seed=0):
    """Parallelogram light center + a edge_u + b edge_v, a, b in [-1, 1]."""
    L = AreaLight()
    for k in range(3):
        L.center[k], L.edge_u[k], L.edge_v[k] = float(center[k]), float(edge_u[k]), float(edge_v[k])
    L.samples, L.seed = samples, seed
    return L


class EnvironmentMap(ctypes.Structure):
    _fields_ = [("texels", ctypes.c_void_p), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("exposure", ctypes.c_float)]


class RenderDesc(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("flags", ctypes.c_uint32),
                ("image", ctypes.c_void_p), ("primitive_ids", ctypes.c_void_p),
                ("light_visibility", ctypes.c_void_p), ("shadow_mlp", ctypes.c_void_p),
                ("shadow_variance_threshold", ctypes.c_float), ("area_light", ctypes.POINTER(AreaLight)),
                ("environment", ctypes.POINTER(EnvironmentMap))]


class RenderStats(ctypes.Structure):
    _fields_ = [("primary_rays", ctypes.c_uint64), ("shadow_queries", ctypes.c_uint64),
                ("shadow_rays_traced", ctypes.c_uint64)]


class View(ctypes.Structure):
    _fields_ = [("camera", Camera), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("image", ctypes.c_void_p), ("primitive_ids", ctypes.c_void_p)]


class BatchDesc(ctypes.Structure):
    _fields_ = [("flags", ctypes.c_uint32), ("bounces", ctypes.c_uint32), ("seed", ctypes.c_uint64),
                ("environment", ctypes.c_float * 3), ("material_albedo", ctypes.c_void_p),
                ("material_count", ctypes.c_uint32)]


class BatchStats(ctypes.Structure):
    _fields_ = [("paths", ctypes.c_uint64), ("bounce_rays", ctypes.c_uint64), ("shadow_rays", ctypes.c_uint64),
                ("graph_captures", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class Tuning(ctypes.Structure):
    _fields_ = [("stack_kind", ctypes.c_int32), ("blocks_per_cu", ctypes.c_int32), ("refill_min", ctypes.c_int32),
                ("ray_order", ctypes.c_int32), ("bvh_width", ctypes.c_int32), ("leaf_size", ctypes.c_int32),
                ("batch_flags", ctypes.c_uint32), ("trials", ctypes.c_uint32), ("baseline_ms", ctypes.c_float),
                ("best_ms", ctypes.c_float), ("xcd_chunk", ctypes.c_int32), ("bvh_collapse", ctypes.c_int32),
                ("steps_per_iteration", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ShadowMlpDesc(ctypes.Structure):
    _fields_ = [("scene_min", ctypes.c_float * 3), ("scene_max", ctypes.c_float * 3),
                ("seed", ctypes.c_uint64)]


class ShadowMlpTrainDesc(ctypes.Structure):
    _fields_ = [("batch", ctypes.c_uint32), ("epochs", ctypes.c_uint32),
                ("learning_rate", ctypes.c_float), ("seed", ctypes.c_uint64)]


class StrandParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in
                ("sphere_radius", "step", "gravity", "curl", "r_root", "r_tip")]


class AssetInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in ("mesh_count", "node_count", "material_count", "instance_count",
                                               "line_mesh_count", "triangle_mesh_count")]


class AssetMesh(ctypes.Structure):
    _fields_ = [("primitive", ctypes.c_uint32), ("material", ctypes.c_int32), ("vertex_count", ctypes.c_uint32),
                ("index_count", ctypes.c_uint32), ("vertices", ctypes.POINTER(ctypes.c_float)),
                ("indices", ctypes.POINTER(ctypes.c_uint32)), ("bounds_min", ctypes.c_float * 3),
                ("bounds_max", ctypes.c_float * 3), ("name", ctypes.c_char_p)]


class AssetNode(ctypes.Structure):
    _fields_ = [("parent", ctypes.c_int32), ("mesh_count", ctypes.c_uint32), ("meshes", ctypes.POINTER(ctypes.c_uint32)),
                ("local", ctypes.c_float * 16), ("world", ctypes.c_float * 16), ("name", ctypes.c_char_p)]


class AssetMaterial(ctypes.Structure):
    _fields_ = [("base_color", ctypes.c_float * 4), ("metallic", ctypes.c_float), ("roughness", ctypes.c_float),
                ("emissive", ctypes.c_float * 3), ("pad", ctypes.c_float)]


class AssetInstance(ctypes.Structure):
    _fields_ = [("node", ctypes.c_uint32), ("mesh", ctypes.c_uint32), ("transform", ctypes.c_float * 12)]


class AssetBuildDesc(ctypes.Structure):
    _fields_ = [("hair_mode", ctypes.c_uint32), ("radius", ctypes.c_float), ("tension", ctypes.c_float),
                ("leaf_size", ctypes.c_uint32), ("default_material", ctypes.c_uint32),
                ("voxel_size", ctypes.c_float), ("voxel_lod", ctypes.c_uint32), ("voxel_flags", ctypes.c_uint32)]


class VoxelGrid(ctypes.Structure):
    _fields_ = [("res", ctypes.c_int32 * 3), ("origin", ctypes.c_float * 3), ("cell", ctypes.c_float),
                ("words", ctypes.c_uint64)]


_lib = None


def lib():
    """Loads libvkhrt.so; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() or "
                          "python vkhrt_amd/build.py")
    L = ctypes.CDLL(LIB_PATH)
    P, u32, i32, f32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_float
    sigs = {
        "vkhrt_abi_version": ([], i32),
        "vkhrt_context_create": ([i32, ctypes.POINTER(P)], i32),
        "vkhrt_context_destroy": ([P], None),
        "vkhrt_last_error": ([P], ctypes.c_char_p),
        "vkhrt_synchronize": ([P, P], i32),
        "vkhrt_set_option": ([P, i32, i32], i32),
        "vkhrt_blas_build": ([P, ctypes.POINTER(BlasDesc), P, ctypes.POINTER(P)], i32),
        "vkhrt_blas_info_get": ([P, ctypes.POINTER(BlasInfo)], i32),
        "vkhrt_blas_destroy": ([P], None),
        "vkhrt_tlas_build": ([P, ctypes.POINTER(P), u32, P, ctypes.POINTER(P)], i32),
        "vkhrt_tlas_destroy": ([P], None),
        "vkhrt_trace_rays": ([P, P, P, u32, P, u32, P], i32),
        "vkhrt_trace_rays_noalloc": ([P, P, P, u32, P, u32, P], i32),
        "vkhrt_work_counters": ([P, P, u32, i32], i32),
        "vkhrt_render": ([P, P, ctypes.POINTER(Camera), ctypes.POINTER(RenderDesc), P], i32),
        "vkhrt_camera_look_at": ([P, P, P, f32, f32, f32, f32, ctypes.POINTER(Camera)], i32),
        "vkhrt_generate_strands": ([P, ctypes.c_uint64, u32, u32, ctypes.POINTER(StrandParams), P, P], i32),
        "vkhrt_generate_lines": ([P, P, u32, u32, P, P], i32),
        "vkhrt_generate_curves": ([P, P, u32, f32, P, P], i32),
        "vkhrt_curves_to_lines": ([P, P, u32, u32, P, P], i32),
        "vkhrt_generate_lss": ([P, P, u32, f32, P, P], i32),
        "vkhrt_generate_dots": ([P, P, u32, f32, P, P], i32),
        "vkhrt_profiling_enable": ([P, i32], i32),
        "vkhrt_kernel_stats": ([P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(u32)], i32),
        "vkhrt_status_counters": ([P, ctypes.POINTER(u32)], i32),
        "vkhrt_generate_camera_rays": ([P, ctypes.POINTER(Camera), u32, u32, f32, f32, P, P], i32),
        "vkhrt_generate_shadow_rays": ([P, P, P, u32, P, f32, f32, P, P, P, P], i32),
        "vkhrt_count_candidates": ([P, P, P, P, u32, P, P], i32),
        "vkhrt_render_ray_counts": ([P, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)], i32),
        "vkhrt_trace_stats": ([P, P, P, u32, P, u32, P, P], i32),
        "vkhrt_render_stats_get": ([P, ctypes.POINTER(RenderStats)], i32),
        "vkhrt_shadow_mlp_create": ([P, ctypes.POINTER(ShadowMlpDesc), ctypes.POINTER(P)], i32),
        "vkhrt_shadow_mlp_destroy": ([P], None),
        "vkhrt_shadow_mlp_get_params": ([P, P, P], i32),
        "vkhrt_shadow_mlp_set_params": ([P, P, P], i32),
        "vkhrt_shadow_mlp_collect": ([P, P, P, u32, P, P, P], i32),
        "vkhrt_shadow_mlp_train": ([P, P, P, P, u32, ctypes.POINTER(ShadowMlpTrainDesc),
                                    ctypes.POINTER(f32), P], i32),
        "vkhrt_shadow_mlp_gradient": ([P, P, P, P, u32, P, ctypes.POINTER(f32), P], i32),
        "vkhrt_shadow_mlp_infer": ([P, P, P, u32, P, P], i32),
        "vkhrt_render_batch": ([P, P, ctypes.POINTER(View), u32, ctypes.POINTER(BatchDesc), P], i32),
        "vkhrt_generate_surface_shadow_rays": ([P, P, ctypes.POINTER(AreaLight), u32, ctypes.c_uint64, P, P], i32),
        "vkhrt_batch_stats_get": ([P, ctypes.POINTER(BatchStats)], i32),
        "vkhrt_autotune_render": ([P, P, ctypes.POINTER(Camera), ctypes.POINTER(RenderDesc), u32,
                                   ctypes.POINTER(Tuning), P], i32),
        "vkhrt_autotune_batch": ([P, P, ctypes.POINTER(View), u32, ctypes.POINTER(BatchDesc), u32,
                                  ctypes.POINTER(Tuning), P], i32),
        "vkhrt_autotune_blas": ([P, ctypes.POINTER(BlasDesc), ctypes.POINTER(Camera), ctypes.POINTER(RenderDesc), u32,
                                 ctypes.POINTER(Tuning), P], i32),
        "vkhrt_asset_load": ([ctypes.c_char_p, ctypes.POINTER(P)], i32),
        "vkhrt_asset_load_memory": ([P, ctypes.c_size_t, ctypes.c_char_p, ctypes.POINTER(P)], i32),
        "vkhrt_asset_destroy": ([P], None),
        "vkhrt_asset_last_error": ([], ctypes.c_char_p),
        "vkhrt_asset_info_get": ([P, ctypes.POINTER(AssetInfo)], i32),
        "vkhrt_asset_mesh_get": ([P, u32, ctypes.POINTER(AssetMesh)], i32),
        "vkhrt_asset_node_get": ([P, u32, ctypes.POINTER(AssetNode)], i32),
        "vkhrt_asset_material_get": ([P, u32, ctypes.POINTER(AssetMaterial)], i32),
        "vkhrt_asset_instance_get": ([P, u32, ctypes.POINTER(AssetInstance)], i32),
        "vkhrt_asset_build": ([P, P, ctypes.POINTER(AssetBuildDesc), P, ctypes.POINTER(P), u32,
                               ctypes.POINTER(u32)], i32),
        "vkhrt_generate_mesh_lines": ([P, P, u32, P, u32, f32, P, P], i32),
        "vkhrt_blas_voxel_grid_get": ([P, ctypes.POINTER(VoxelGrid)], i32),
        "vkhrt_blas_voxels_read": ([P, P, P, ctypes.c_uint64], i32),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    if L.vkhrt_abi_version() != ABI_VERSION:
        raise ImportError(f"{LIB_PATH} has ABI {L.vkhrt_abi_version()}, this module expects {ABI_VERSION}: "
                          "rebuild it with python vkhrt_amd/build.py")
    _lib = L
    return L


def _f3(v):
    return (ctypes.c_float * 3)(*[float(x) for x in v])


def camera_look_at(eye, center, up=(0.0, 1.0, 0.0), fov_deg=60.0, aspect=1.0, znear=0.1,
                   zfar=1000.0):
    """FlyCamera view/projection (fly_camera.cpp:25-35) inverted like renderer.cpp:189-195."""
    cam = Camera()
    rc = lib().vkhrt_camera_look_at(_f3(eye), _f3(center), _f3(up),
                                    math.tan(math.radians(fov_deg) * 0.5), aspect, znear, zfar,
                                    ctypes.byref(cam))
    if rc:
        raise VkhrtError(rc, "camera_look_at")
    return cam


class Context:
    """One device context (the analogue of VulkanContext for this path)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        rc = lib().vkhrt_context_create(device, ctypes.byref(h))
        if rc:
            raise VkhrtError(rc, "context_create")
        self.h = h
        # bound now: module globals may already be cleared when __del__ runs at exit
        self._destroy = lib().vkhrt_context_destroy

    def close(self):
        if getattr(self, "h", None):
            self._destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def check(self, rc, what=""):
        if rc:
            msg = lib().vkhrt_last_error(self.h)
            raise VkhrtError(rc, f"{what}: {msg.decode() if msg else ''}")

    def set_option(self, option, value):
        self.check(lib().vkhrt_set_option(self.h, option, value), "set_option")

    def synchronize(self, stream=None):
        self.check(lib().vkhrt_synchronize(self.h, stream), "synchronize")

    def trace_rays(self, tlas, rays_ptr, count, hits_ptr, flags=0, stream=None):
        """traceRayEXT over `count` vkhrt_ray records -> vkhrt_hit records (device)."""
        self.check(lib().vkhrt_trace_rays(self.h, tlas.h, rays_ptr, count, hits_ptr, flags, stream),
                   "trace_rays")

    def render(self, tlas, camera, width, height, image_ptr, flags=RENDER_SHADOWS,
               prim_ids_ptr=None, light_visibility_ptr=None, stream=None, shadow_mlp=None,
               variance_threshold=0.0, area_light=None, environment=None):
        """traceRaysKHR(W, H, 1) with ray_gen.rgen and the closest-hit shaders. With
        RENDER_NEURAL_SHADOWS, `shadow_mlp` answers shadow queries whose predicted variance
        p (1 - p) is at most `variance_threshold`; the rest trace exactly. `area_light`
        (an AreaLight) replaces the directional light with light samples per hit. `environment`
        (an EnvironmentMap over a device RGBA32F image) colours the primary misses."""
        d = RenderDesc(width, height, flags, image_ptr, prim_ids_ptr, light_visibility_ptr,
                       shadow_mlp.h if shadow_mlp is not None else None, variance_threshold,
                       ctypes.pointer(area_light) if area_light is not None else None,
                       ctypes.pointer(environment) if environment is not None else None)
        self.check(lib().vkhrt_render(self.h, tlas.h, ctypes.byref(camera), ctypes.byref(d), stream),
                   "render")

    @staticmethod
    def _batch_args(views, bounces, seed, environment, flags, material_albedo):
        arr = (View * len(views))()
        for i, v in enumerate(views):
            cam, w, h, img = v[:4]
            arr[i].camera = cam
            arr[i].width, arr[i].height, arr[i].image = w, h, img
            arr[i].primitive_ids = v[4] if len(v) > 4 else None
        d = BatchDesc()
        d.flags, d.bounces, d.seed = flags, bounces, seed
        for k in range(3):
            d.environment[k] = float(environment[k])
        tints = None
        if material_albedo:
            tints = (ctypes.c_float * (3 * len(material_albedo)))(*[float(x) for t in material_albedo for x in t])
            d.material_albedo = ctypes.cast(tints, ctypes.c_void_p)
            d.material_count = len(material_albedo)
        return arr, d, tints

    def render_batch(self, tlas, views, bounces=4, seed=0, environment=(0.0, 0.0, 0.0), flags=0,
                     material_albedo=None, stream=None):
        """Batched path tracing of many views of `tlas` in one pipeline. views: list of
        (camera, width, height, image_ptr[, primitive_ids_ptr]); material_albedo: list of
        (r, g, b) tints indexed by BLAS material index."""
        arr, d, _keep = self._batch_args(views, bounces, seed, environment, flags, material_albedo)
        self.check(lib().vkhrt_render_batch(self.h, tlas.h, arr, len(views), ctypes.byref(d), stream),
                   "render_batch")

    # autotuner: times the given workload under candidate settings, keeps the fastest
    def autotune_render(self, tlas, camera, width, height, image_ptr, flags=RENDER_SHADOWS, repeats=3, stream=None,
                        area_light=None, shadow_mlp=None, variance_threshold=0.0):
        """Times the render (the same arguments as render(), neural form included) under candidate
        traversal settings and leaves the context on the fastest."""
        d = RenderDesc(width, height, flags, image_ptr, None, None,
                       shadow_mlp.h if shadow_mlp is not None else None, variance_threshold,
                       ctypes.pointer(area_light) if area_light is not None else None)
        t = Tuning()
        self.check(lib().vkhrt_autotune_render(self.h, tlas.h, ctypes.byref(camera), ctypes.byref(d), repeats,
                                               ctypes.byref(t), stream), "autotune_render")
        return t.as_dict()

    def apply_tuning(self, t):
        """Sets the traversal options of a tuning (autotune_render's result) on the context."""
        for opt, key in ((OPTION_STACK_KIND, "stack_kind"), (OPTION_BLOCKS_PER_CU, "blocks_per_cu"),
                         (OPTION_REFILL_MIN, "refill_min"), (OPTION_RAY_ORDER, "ray_order"),
                         (OPTION_XCD_CHUNK, "xcd_chunk"), (OPTION_STEPS_PER_ITERATION, "steps_per_iteration")):
            self.set_option(opt, int(t[key]))

    def autotune_batch(self, tlas, views, bounces=4, seed=0, environment=(0.0, 0.0, 0.0), flags=0,
                       material_albedo=None, repeats=3, stream=None):
        arr, d, _keep = self._batch_args(views, bounces, seed, environment, flags, material_albedo)
        t = Tuning()
        self.check(lib().vkhrt_autotune_batch(self.h, tlas.h, arr, len(views), ctypes.byref(d), repeats,
                                              ctypes.byref(t), stream), "autotune_batch")
        return t.as_dict()

    def autotune_blas(self, mode, primitives_ptr, count, camera, width, height, image_ptr, blas_flags=0,
                      radius=0.02, render_flags=RENDER_SHADOWS, repeats=3, stream=None):
        b = BlasDesc()
        b.mode, b.flags, b.primitives, b.count, b.radius = mode, blas_flags, primitives_ptr, count, radius
        d = RenderDesc(width, height, render_flags, image_ptr, None, None, None, 0.0)
        t = Tuning()
        self.check(lib().vkhrt_autotune_blas(self.h, ctypes.byref(b), ctypes.byref(camera), ctypes.byref(d), repeats,
                                             ctypes.byref(t), stream), "autotune_blas")
        return t.as_dict()

    def batch_stats(self):
        """(paths, bounce rays, shadow rays, graph captures) of the last render_batch."""
        st = BatchStats()
        self.check(lib().vkhrt_batch_stats_get(self.h, ctypes.byref(st)), "batch_stats")
        return st.paths, st.bounce_rays, st.shadow_rays, st.graph_captures

    # geometry front-end
    def generate_strands(self, seed, nstrands, nseg, vertices_ptr, params=None, stream=None):
        p = None if params is None else ctypes.byref(StrandParams(*params))
        self.check(lib().vkhrt_generate_strands(self.h, seed, nstrands, nseg, p, vertices_ptr, stream),
                   "generate_strands")

    def generate_lines(self, vertices_ptr, nstrands, nseg, lines_ptr, stream=None):
        self.check(lib().vkhrt_generate_lines(self.h, vertices_ptr, nstrands, nseg, lines_ptr, stream),
                   "generate_lines")

    def generate_curves(self, lines_ptr, count, curves_ptr, tension=1.0, stream=None):
        self.check(lib().vkhrt_generate_curves(self.h, lines_ptr, count, tension, curves_ptr, stream),
                   "generate_curves")

    def curves_to_lines(self, curves_ptr, count, pieces, lines_ptr, stream=None):
        self.check(lib().vkhrt_curves_to_lines(self.h, curves_ptr, count, pieces, lines_ptr, stream),
                   "curves_to_lines")

    def generate_lss(self, lines_ptr, count, caps_ptr, constant_radius=-1.0, stream=None):
        self.check(lib().vkhrt_generate_lss(self.h, lines_ptr, count, constant_radius, caps_ptr, stream),
                   "generate_lss")

    def generate_dots(self, lines_ptr, count, pieces_ptr, constant_radius=-1.0, stream=None):
        self.check(lib().vkhrt_generate_dots(self.h, lines_ptr, count, constant_radius, pieces_ptr,
                                             stream), "generate_dots")

    def generate_mesh_lines(self, vertices_ptr, vertex_count, indices_ptr, count, lines_ptr, radius=0.02,
                            stream=None):
        """GenerateLines of an indexed line mesh (8-float vertices, mesh-local index pairs)."""
        self.check(lib().vkhrt_generate_mesh_lines(self.h, vertices_ptr, vertex_count, indices_ptr, count, radius,
                                                   lines_ptr, stream), "generate_mesh_lines")

    # auxiliary ray work
    def generate_camera_rays(self, camera, width, height, rays_ptr, tmin=0.001, tmax=10000.0,
                             stream=None):
        self.check(lib().vkhrt_generate_camera_rays(self.h, ctypes.byref(camera), width, height, tmin,
                                                    tmax, rays_ptr, stream), "generate_camera_rays")

    def generate_shadow_rays(self, rays_ptr, hits_ptr, count, out_ptr, count_out_ptr,
                             direction=(0.0, 1.0, 0.0), tmin=0.001, tmax=10000.0, source_ptr=None,
                             stream=None):
        self.check(lib().vkhrt_generate_shadow_rays(self.h, rays_ptr, hits_ptr, count, _f3(direction),
                                                    tmin, tmax, out_ptr, source_ptr, count_out_ptr,
                                                    stream), "generate_shadow_rays")

    def generate_surface_shadow_rays(self, tlas, count, seed, rays_ptr, light=None, stream=None):
        """Shadow rays from points drawn on the TLAS's hair surfaces toward `light` (AreaLight or
        None for the direction (0, 1, 0))."""
        self.check(lib().vkhrt_generate_surface_shadow_rays(self.h, tlas.h, ctypes.pointer(light) if light else None,
                                                            count, seed, rays_ptr, stream),
                   "generate_surface_shadow_rays")

    def count_candidates(self, tlas, rays_ptr, hits_ptr, count, counts_ptr, stream=None):
        self.check(lib().vkhrt_count_candidates(self.h, tlas.h, rays_ptr, hits_ptr, count, counts_ptr,
                                                stream), "count_candidates")

    def trace_stats(self, tlas, rays_ptr, count, hits_ptr, stats_ptr, flags=0, stream=None):
        """trace_rays with per-ray (nodes visited, primitives tested) into stats (2 x u32)."""
        self.check(lib().vkhrt_trace_stats(self.h, tlas.h, rays_ptr, count, hits_ptr, flags, stats_ptr,
                                           stream), "trace_stats")

    def render_stats(self):
        """(primary rays, shadow queries, shadow rays traced exactly) of the last render."""
        st = RenderStats()
        self.check(lib().vkhrt_render_stats_get(self.h, ctypes.byref(st)), "render_stats")
        return st.primary_rays, st.shadow_queries, st.shadow_rays_traced

    def render_ray_counts(self):
        p, s = ctypes.c_uint64(), ctypes.c_uint64()
        self.check(lib().vkhrt_render_ray_counts(self.h, ctypes.byref(p), ctypes.byref(s)),
                   "render_ray_counts")
        return p.value, s.value

    # profiling
    def profiling(self, enable):
        self.check(lib().vkhrt_profiling_enable(self.h, int(enable)), "profiling")

    def kernel_stats(self, name):
        ms = ctypes.c_double()
        n = ctypes.c_uint32()
        self.check(lib().vkhrt_kernel_stats(self.h, name.encode(), ctypes.byref(ms), ctypes.byref(n)),
                   "kernel_stats")
        return ms.value, n.value

    def stack_overflows(self):
        n = ctypes.c_uint32()
        self.check(lib().vkhrt_status_counters(self.h, ctypes.byref(n)), "status")
        return n.value


class BottomLevelAccelerationStructure:
    """BLAS over hair primitives (bottom_level_acceleration_structure.hpp:34).

    `primitives` is a device address (or a host address with BLAS_HOST_POINTERS), laid out per
    include/vkhrt.h. `transform` is a row-major 3x4 world-from-object matrix (BLASInput.transform).
    """

    def __init__(self, ctx, mode, primitives_ptr, count, radius=0.02, flags=0, transform=None,
                 leaf_size=0, material_index=0, stream=None, indices_ptr=None, vertex_count=0, voxel_size=0.1,
                 voxel_lod=0, bounds=None):
        """GEOMETRY_TRIANGLES: `primitives_ptr` holds `vertex_count` vertices (position, normal,
        texCoord = 8 floats), `indices_ptr` 3 u32 per triangle, `count` triangles.
        GEOMETRY_VOXELS: `count` lines (8 floats) voxelized as capsules of `radius` into cells of
        `voxel_size` over `bounds` ((min xyz, max xyz); default: the lines' box), traced at
        level of detail `voxel_lod`."""
        d = BlasDesc()
        d.mode, d.flags, d.primitives, d.count = mode, flags, primitives_ptr, count
        d.radius, d.material_index, d.leaf_size = radius, material_index, leaf_size
        d.indices, d.vertex_count = indices_ptr, vertex_count
        d.voxel_size, d.voxel_lod = voxel_size, voxel_lod
        if bounds is not None:
            d.flags |= BLAS_HAS_BOUNDS
            for k in range(3):
                d.bounds_min[k], d.bounds_max[k] = float(bounds[0][k]), float(bounds[1][k])
        if transform is not None:
            d.flags |= BLAS_HAS_TRANSFORM
            for i, v in enumerate(list(transform)[:12]):
                d.transform[i] = float(v)
        h = ctypes.c_void_p()
        ctx.check(lib().vkhrt_blas_build(ctx.h, ctypes.byref(d), stream, ctypes.byref(h)), "blas_build")
        self.ctx, self.h, self.mode = ctx, h, mode
        self._destroy = lib().vkhrt_blas_destroy

    @classmethod
    def _adopt(cls, ctx, handle, mode):
        """Wraps a vkhrt_blas* made by another entry point (it is destroyed with this object)."""
        obj = cls.__new__(cls)
        obj.ctx, obj.h, obj.mode = ctx, handle, mode
        obj._destroy = lib().vkhrt_blas_destroy
        return obj

    def voxel_grid(self):
        """Voxel BLAS: (occupancy as a bool array [z, y, x], origin (3,), cell edge)."""
        import numpy as np
        g = VoxelGrid()
        if lib().vkhrt_blas_voxel_grid_get(self.h, ctypes.byref(g)) != 0:
            raise VkhrtError(-1, "not a voxel BLAS")
        words = np.zeros(max(1, g.words), np.uint32)
        self.ctx.check(lib().vkhrt_blas_voxels_read(self.ctx.h, self.h, words.ctypes.data, g.words), "voxels_read")
        res = [g.res[0], g.res[1], g.res[2]]
        n = res[0] * res[1] * res[2]
        cells = np.unpackbits(words.view(np.uint8), bitorder="little")[:n].astype(bool).reshape(res[2], res[1], res[0])
        return cells, np.array(g.origin[:], np.float32), g.cell

    def info(self):
        i = BlasInfo()
        lib().vkhrt_blas_info_get(self.h, ctypes.byref(i))
        return {k: getattr(i, k) for k, _ in BlasInfo._fields_}

    def close(self):
        if getattr(self, "h", None):
            self._destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Asset:
    """A glTF 2.0 polyline / triangle asset: ModelLoader::LoadFromFile
    (source/resources/model/model_loader.cpp:274-312) and the SceneGraph it builds, read by the
    library's own glTF reader on the host. Meshes are mesh primitives in file order, nodes are
    in pre-order with ``world`` = Node::GetWorldMatrix (model.cpp:6-18), column-major.

    ``Asset(path)`` reads a .gltf / .glb file; ``Asset(data=bytes, base_dir=...)`` reads one
    from memory. Errors raise VkhrtError with the loader's message."""

    def __init__(self, path=None, data=None, base_dir=None):
        h = ctypes.c_void_p()
        if path is not None:
            rc = lib().vkhrt_asset_load(os.fsencode(path), ctypes.byref(h))
        else:
            buf = bytes(data)
            rc = lib().vkhrt_asset_load_memory(buf, len(buf), os.fsencode(base_dir) if base_dir else None,
                                               ctypes.byref(h))
        if rc:
            raise VkhrtError(rc, f"asset_load: {lib().vkhrt_asset_last_error().decode()}")
        self.h = h
        self._destroy = lib().vkhrt_asset_destroy

    def _get(self, fn, i, struct):
        out = struct()
        rc = fn(self.h, i, ctypes.byref(out))
        if rc:
            raise VkhrtError(rc, lib().vkhrt_asset_last_error().decode())
        return out

    def info(self):
        i = AssetInfo()
        lib().vkhrt_asset_info_get(self.h, ctypes.byref(i))
        return {k: getattr(i, k) for k, _ in AssetInfo._fields_}

    def mesh(self, i):
        """dict: primitive, material, vertices (n, 8) float32 (position, normal, texCoord),
        indices (uint32, mesh-local), bounds_min, bounds_max, name; arrays are copies."""
        import numpy as np
        m = self._get(lib().vkhrt_asset_mesh_get, i, AssetMesh)
        verts = (np.ctypeslib.as_array(m.vertices, (m.vertex_count * 8,)).reshape(-1, 8).copy()
                 if m.vertex_count else np.zeros((0, 8), np.float32))
        idx = (np.ctypeslib.as_array(m.indices, (m.index_count,)).copy() if m.index_count
               else np.zeros(0, np.uint32))
        return {"primitive": m.primitive, "material": m.material, "vertices": verts, "indices": idx,
                "bounds_min": np.array(m.bounds_min[:], np.float32), "bounds_max": np.array(m.bounds_max[:], np.float32),
                "name": (m.name or b"").decode()}

    def node(self, i):
        """dict: parent, meshes, local and world (16 floats, column-major), name."""
        import numpy as np
        n = self._get(lib().vkhrt_asset_node_get, i, AssetNode)
        return {"parent": n.parent, "meshes": [n.meshes[k] for k in range(n.mesh_count)],
                "local": np.array(n.local[:], np.float32), "world": np.array(n.world[:], np.float32),
                "name": (n.name or b"").decode()}

    def material(self, i):
        m = self._get(lib().vkhrt_asset_material_get, i, AssetMaterial)
        return {"base_color": list(m.base_color), "metallic": m.metallic, "roughness": m.roughness,
                "emissive": list(m.emissive)}

    def instance(self, i):
        """dict: node, mesh, transform (row-major 3x4 world-from-object, 12 floats)."""
        import numpy as np
        t = self._get(lib().vkhrt_asset_instance_get, i, AssetInstance)
        return {"node": t.node, "mesh": t.mesh, "transform": np.array(t.transform[:], np.float32)}

    @property
    def meshes(self):
        return [self.mesh(i) for i in range(self.info()["mesh_count"])]

    @property
    def nodes(self):
        return [self.node(i) for i in range(self.info()["node_count"])]

    @property
    def instances(self):
        return [self.instance(i) for i in range(self.info()["instance_count"])]

    def build(self, ctx, hair_mode=GEOMETRY_PHANTOM_CURVES, radius=0.02, tension=1.0, leaf_size=0,
              default_material=0, stream=None, voxel_size=0.1, voxel_lod=0, voxel_flags=0):
        """One BLAS per instance (Renderer::InitializeBLAS, renderer.cpp:694-728), built on the
        device; pass the list to TopLevelAccelerationStructure."""
        d = AssetBuildDesc(hair_mode, radius, tension, leaf_size, default_material, voxel_size, voxel_lod, voxel_flags)
        n = self.info()["instance_count"]
        arr = (ctypes.c_void_p * max(1, n))()
        cnt = ctypes.c_uint32()
        ctx.check(lib().vkhrt_asset_build(ctx.h, self.h, ctypes.byref(d), stream, arr, n, ctypes.byref(cnt)),
                  "asset_build")
        out = []
        for i in range(cnt.value):
            inst = self.instance(i)
            mode = hair_mode if self.mesh(inst["mesh"])["primitive"] == PRIMITIVE_LINES else GEOMETRY_TRIANGLES
            out.append(BottomLevelAccelerationStructure._adopt(ctx, ctypes.c_void_p(arr[i]), mode))
        return out

    def close(self):
        if getattr(self, "h", None):
            self._destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class TopLevelAccelerationStructure:
    """TLAS: instance i = blases[i], customIndex = i (top_level_acceleration_structure.cpp:23-43)."""

    def __init__(self, ctx, blases, stream=None):
        arr = (ctypes.c_void_p * max(1, len(blases)))(*[b.h.value for b in blases])
        h = ctypes.c_void_p()
        ctx.check(lib().vkhrt_tlas_build(ctx.h, arr, len(blases), stream, ctypes.byref(h)), "tlas_build")
        self.ctx, self.h, self.blases = ctx, h, list(blases)
        self._destroy = lib().vkhrt_tlas_destroy

    def close(self):
        if getattr(self, "h", None):
            self._destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class ShadowMLP:
    """Neural shadow-visibility MLP (BF16 MFMA training and inference on the device)."""

    def __init__(self, ctx, scene_min, scene_max, seed=1):
        self.ctx = ctx
        d = ShadowMlpDesc((ctypes.c_float * 3)(*scene_min), (ctypes.c_float * 3)(*scene_max), seed)
        h = ctypes.c_void_p()
        ctx.check(lib().vkhrt_shadow_mlp_create(ctx.h, ctypes.byref(d), ctypes.byref(h)), "shadow_mlp_create")
        self.h = h
        self._destroy = lib().vkhrt_shadow_mlp_destroy

    def close(self):
        if getattr(self, "h", None):
            self._destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def get_params(self):
        buf = (ctypes.c_float * SHADOW_MLP_PARAMS)()
        self.ctx.check(lib().vkhrt_shadow_mlp_get_params(self.ctx.h, self.h, buf), "shadow_mlp_get_params")
        return list(buf)

    def set_params(self, params):
        buf = (ctypes.c_float * SHADOW_MLP_PARAMS)(*[float(x) for x in params])
        self.ctx.check(lib().vkhrt_shadow_mlp_set_params(self.ctx.h, self.h, buf), "shadow_mlp_set_params")

    def collect(self, tlas, shadow_rays_ptr, count, samples_ptr, targets_ptr, stream=None):
        self.ctx.check(lib().vkhrt_shadow_mlp_collect(self.ctx.h, tlas.h, shadow_rays_ptr, count, samples_ptr,
                                                      targets_ptr, stream), "shadow_mlp_collect")

    def train(self, samples_ptr, targets_ptr, count, batch=65536, epochs=1, learning_rate=3e-3, seed=1,
              stream=None):
        d = ShadowMlpTrainDesc(batch, epochs, learning_rate, seed)
        loss = ctypes.c_float()
        self.ctx.check(lib().vkhrt_shadow_mlp_train(self.ctx.h, self.h, samples_ptr, targets_ptr, count,
                                                    ctypes.byref(d), ctypes.byref(loss), stream),
                       "shadow_mlp_train")
        return loss.value

    def gradient(self, samples_ptr, targets_ptr, count, grad_ptr, stream=None):
        loss = ctypes.c_float()
        self.ctx.check(lib().vkhrt_shadow_mlp_gradient(self.ctx.h, self.h, samples_ptr, targets_ptr, count,
                                                       grad_ptr, ctypes.byref(loss), stream),
                       "shadow_mlp_gradient")
        return loss.value

    def infer(self, samples_ptr, count, out_ptr, stream=None):
        self.ctx.check(lib().vkhrt_shadow_mlp_infer(self.ctx.h, self.h, samples_ptr, count, out_ptr, stream),
                       "shadow_mlp_infer")
