"""GPU parity of the double-word final-hit re-run (VKHRT_TRACE_ACCURATE_NORMALS,
VKHRT_RENDER_ACCURATE_NORMALS) against the oracle's F_ACCURATE, bit for bit, and its accuracy
against the FP64 evaluation of the same intersector: the shading normal within 1e-6 rad and t
within 2e-7 relative on every sampled hit (the FP32 intersector alone reaches 1e-3 rad)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

XF = np.array([[0.0, -1.0, 0.0, 40.0], [1.0, 0.0, 0.0, -5.0], [0.0, 0.0, 1.0, 10.0]], np.float32)


@pytest.mark.parametrize("per_point", [True, False])
def test_accurate_trace_matches_oracle_and_fp64(gpu_ctx, oracle, per_point):
    import gpu_util as G
    import scenes
    import vkhrt_amd as V
    curves = scenes.c0_curves()[:20000]
    b0 = G.build_blas(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, curves, per_point=per_point)
    b1 = G.build_blas(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, curves, per_point=per_point, transform=XF.reshape(12))
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, [b0, b1])
    S = oracle.Scene()
    S.add(oracle.MODE_PHANTOM, curves, per_point=per_point)
    S.add(oracle.MODE_PHANTOM, curves, per_point=per_point, transform=XF)
    rays = scenes.targeted_rays(curves, 20000, seed=13, dist=(5.0, 300.0))
    print(f"[accurate trace] per-point {per_point}: 20k rays", flush=True)
    h = G.trace(gpu_ctx, tlas, rays, V.TRACE_ACCURATE_NORMALS)
    ref = S.trace(rays, flags=oracle.F_ACCURATE)
    nhit = G.compare_hits(h, ref, "accurate")
    assert nhit > 4000
    plain = G.trace(gpu_ctx, tlas, rays)
    hit = plain["prim"] != 0xFFFFFFFF
    assert np.array_equal(plain["prim"], h["prim"]) and np.array_equal(plain["inst"], h["inst"])
    assert (plain["n"][hit] != h["n"][hit]).any(axis=1).mean() > 0.5  # the normals were re-evaluated
    # against FP64 on identity-instance hits
    ang, rel = [], []
    for i in np.nonzero(hit & (h["inst"] == 0))[0][:800]:
        r64 = oracle.prhi_f64(rays[i], curves[h["prim"][i]], per_point=per_point)
        if r64[0] <= 0 or h["v"][i] != 0.0:
            continue
        n = h["n"][i].astype(np.float64)
        ang.append(math.acos(min(1.0, abs(n @ r64[5:8]) / np.linalg.norm(n) / np.linalg.norm(r64[5:8]))))
        rel.append(abs(float(h["t"][i]) - r64[0]) / r64[0])
    print(f"vs FP64: normal max {max(ang):.2g} rad, t rel max {max(rel):.2g} over {len(ang)} hits")
    assert len(ang) > 400 and max(ang) < 1e-6 and max(rel) < 2e-7
    # any-hit rays keep their plain records
    a0 = G.trace(gpu_ctx, tlas, rays, V.TRACE_TERMINATE_ON_FIRST_HIT)
    a1 = G.trace(gpu_ctx, tlas, rays, V.TRACE_TERMINATE_ON_FIRST_HIT | V.TRACE_ACCURATE_NORMALS)
    assert np.array_equal(a0.view(np.uint32), a1.view(np.uint32))


def test_accurate_render_matches_oracle(gpu_ctx, oracle):
    import gpu_util as G
    import scenes
    import vkhrt_amd as V
    curves = scenes.dense_ball_curves(nstrands=2500, nseg=12)
    blas = G.build_blas(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, curves, per_point=True)
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, [blas])
    S = oracle.Scene()
    S.add(oracle.MODE_PHANTOM, curves, per_point=True)
    W, H = 160, 120
    vi, pi = oracle.camera((30.0, 25.0, 30.0), (0.0, 8.0, 0.0), aspect=W / H)
    cam = V.Camera()
    for k in range(16):
        cam.view_inverse[k], cam.proj_inverse[k] = float(vi[k]), float(pi[k])
    for polish in (0, 1):
        flags = V.RENDER_SHADOWS | V.RENDER_ACCURATE_NORMALS | (V.RENDER_FP64_POLISH if polish else 0)
        print(f"[accurate render] polish {polish}", flush=True)
        img, ids = G.render(gpu_ctx, tlas, cam, W, H, flags=flags)
        ref, rids, _, _ = S.render(vi, pi, W, H, shadows=1,
                                   flags=oracle.F_ACCURATE | (oracle.F_POLISH if polish else 0))
        assert (rids != 0xFFFFFFFF).sum() > 2000
        assert np.array_equal(ids, rids) and np.array_equal(img[..., :3].view(np.uint32), ref.view(np.uint32))
    plain, _ = G.render(gpu_ctx, tlas, cam, W, H)
    assert not np.array_equal(plain, img)  # the shading follows the re-evaluated normals
    # a mixed TLAS takes the deferred pipeline with one thread per queued ray
    caps = oracle.generate_lss(scenes.lines_from_curves(curves[:100], 2))
    lss = G.build_blas(gpu_ctx, V.GEOMETRY_LSS, caps)
    mixed = V.TopLevelAccelerationStructure(gpu_ctx, [blas, lss])
    S.add(oracle.MODE_LSS, caps)
    img, ids = G.render(gpu_ctx, mixed, cam, W, H, flags=V.RENDER_SHADOWS | V.RENDER_ACCURATE_NORMALS)
    ref, rids, _, _ = S.render(vi, pi, W, H, shadows=1, flags=oracle.F_ACCURATE)
    assert np.array_equal(ids, rids) and np.array_equal(img[..., :3].view(np.uint32), ref.view(np.uint32))
    # without the deferred pipeline (persistent option off) the accurate re-run is refused
    gpu_ctx.set_option(V.OPTION_PERSISTENT, 0)
    try:
        with pytest.raises(V.VkhrtError) as e:
            G.render(gpu_ctx, mixed, cam, W, H, flags=V.RENDER_SHADOWS | V.RENDER_ACCURATE_NORMALS)
        assert e.value.code == -4
    finally:
        gpu_ctx.set_option(V.OPTION_PERSISTENT, 1)
