"""The allocation-free trace entry point (vkhrt_trace_rays_noalloc) against vkhrt_trace_rays,
through the C-ABI: bit-identical hit records on two calls in a row (the second with fewer
rays, so stale scratch contents from the first call would show), closest hit with and without
the FP64 polish, any hit, and a mixed-mode TLAS; a polished call above the scratch's ray limit
is refused with VKHRT_ERROR_INVALID_ARGUMENT."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

MAX_POLISH_RAYS = 1 << 18


def _call(ctx, fn, tlas, rays, flags):
    import gpu_util as G
    import vkhrt_amd as V
    r = G.dev(rays)
    h = torch.full((rays.shape[0] * 8,), -1.0, dtype=torch.float32, device="cuda")
    rc = getattr(V.lib(), fn)(ctx.h, tlas.h, ctypes.c_void_p(r.data_ptr()), rays.shape[0],
                              ctypes.c_void_p(h.data_ptr()), flags, ctypes.c_void_p(G.stream_ptr()))
    assert rc == 0, (fn, rc, V.lib().vkhrt_last_error(ctx.h))
    torch.cuda.synchronize()
    ctx.synchronize(G.stream_ptr())
    return h.cpu().numpy().view(G.HIT_DTYPE)


@pytest.fixture(scope="module")
def scene(gpu_ctx):
    import gpu_util as G
    import oracle_lib as O
    import scenes
    import vkhrt_amd as V
    curves = scenes.c0_curves()
    ph = G.build_blas(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, curves, per_point=True)
    lss = G.build_blas(gpu_ctx, V.GEOMETRY_LSS, O.generate_lss(scenes.lines_from_curves(curves[:3000], 4)))
    one = V.TopLevelAccelerationStructure(gpu_ctx, [ph])
    mixed = V.TopLevelAccelerationStructure(gpu_ctx, [ph, lss])
    rays = np.concatenate([scenes.grazing_rays(curves, 30000, seed=3), scenes.targeted_rays(curves, 90000, seed=4)])
    yield one, mixed, rays
    mixed.close()
    one.close()


@pytest.mark.parametrize("which,flags", [("one", 0), ("one", 8), ("one", 1), ("one", 9), ("mixed", 8),
                                         ("one", 8 | 32)])
def test_noalloc_matches_trace_rays(gpu_ctx, scene, which, flags):
    import gpu_util as G
    one, mixed, rays = scene
    tlas = one if which == "one" else mixed
    for n in (rays.shape[0], 7001):  # the second call is smaller than the first
        sub = np.ascontiguousarray(rays[:n])
        print(f"[noalloc] {which} flags {flags} rays {n}", flush=True)
        ref = _call(gpu_ctx, "vkhrt_trace_rays", tlas, sub, flags)
        got = _call(gpu_ctx, "vkhrt_trace_rays_noalloc", tlas, sub, flags)
        assert got.view(np.uint32).tobytes() == ref.view(np.uint32).tobytes(), f"{which} flags {flags} n {n}"
        assert (ref["prim"] != 0xFFFFFFFF).sum() > n // 10
    if flags & 8:
        print(f"polished roots {(ref['v'] != 0).sum()}")


def test_noalloc_rejects_oversized_polish(gpu_ctx, scene):
    import gpu_util as G
    import vkhrt_amd as V
    one, _, rays = scene
    n = MAX_POLISH_RAYS + 1
    big = np.ascontiguousarray(np.resize(rays, (n, 8)))
    r = G.dev(big)
    h = torch.zeros(n * 8, dtype=torch.float32, device="cuda")
    rc = V.lib().vkhrt_trace_rays_noalloc(gpu_ctx.h, one.h, ctypes.c_void_p(r.data_ptr()), n,
                                          ctypes.c_void_p(h.data_ptr()), V.TRACE_FP64_POLISH, None)
    assert rc == -1
    assert b"262144" in V.lib().vkhrt_last_error(gpu_ctx.h)
    # without the polish the same count is served
    rc = V.lib().vkhrt_trace_rays_noalloc(gpu_ctx.h, one.h, ctypes.c_void_p(r.data_ptr()), n,
                                          ctypes.c_void_p(h.data_ptr()), 0, None)
    assert rc == 0
    torch.cuda.synchronize()
    gpu_ctx.synchronize()
