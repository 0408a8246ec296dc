"""GPU tests of the autotuner: it only changes time, never results. After tuning, the chosen
settings are valid, the best time is no worse than the starting one, and the same frame
renders bit for bit as before."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture
def fresh_ctx():
    import torch as T
    if not T.cuda.is_available():
        pytest.skip("no GPU")
    import vkhrt_amd as V
    ctx = V.Context(0)
    yield ctx
    ctx.close()


def _scene(ctx):
    import gpu_util as G
    import scenes
    import vkhrt_amd as V
    curves = scenes.dense_ball_curves(nstrands=4000, nseg=16)
    blas = G.build_blas(ctx, V.GEOMETRY_PHANTOM_CURVES, curves, per_point=True)
    return curves, blas, V.TopLevelAccelerationStructure(ctx, [blas])


def _check_ranges(t):
    assert 0 <= t["stack_kind"] <= 3 and 0 <= t["blocks_per_cu"] <= 32 and 1 <= t["refill_min"] <= 64
    assert t["ray_order"] in (0, 1, 2) and 4 <= t["bvh_width"] <= 8 and 1 <= t["leaf_size"] <= 3
    assert t["xcd_chunk"] == 0 or (64 <= t["xcd_chunk"] <= 1 << 20 and t["xcd_chunk"] & (t["xcd_chunk"] - 1) == 0)
    assert 0.0 < t["best_ms"] <= t["baseline_ms"]


def test_autotune_render_keeps_the_image(fresh_ctx):
    import gpu_util as G
    import vkhrt_amd as V
    ctx = fresh_ctx
    curves, blas, tlas = _scene(ctx)
    cam = V.camera_look_at((0.0, 20.0, 45.0), (0.0, 8.0, 0.0), aspect=1.0)
    flags = V.RENDER_SHADOWS | V.RENDER_FP64_POLISH
    before, ids0 = G.render(ctx, tlas, cam, 384, 384, flags)
    img = torch.zeros(384 * 384 * 4, dtype=torch.float32, device="cuda")
    print("[autotune render]", flush=True)
    t = ctx.autotune_render(tlas, cam, 384, 384, img.data_ptr(), flags=flags, repeats=2, stream=G.stream_ptr())
    print(t)
    _check_ranges(t)
    assert t["trials"] >= 10
    after, ids1 = G.render(ctx, tlas, cam, 384, 384, flags)
    assert np.array_equal(ids0, ids1) and np.array_equal(before.view(np.uint32), after.view(np.uint32))
    assert (ids0 != 0xFFFFFFFF).sum() > 1000


def test_autotune_blas_keeps_the_image(fresh_ctx):
    import gpu_util as G
    import scenes
    import vkhrt_amd as V
    ctx = fresh_ctx
    curves = scenes.dense_ball_curves(nstrands=3000, nseg=16)
    d = G.dev(curves)
    cam = V.camera_look_at((30.0, 25.0, 30.0), (0.0, 8.0, 0.0), aspect=1.0)
    blas0 = G.build_blas(ctx, V.GEOMETRY_PHANTOM_CURVES, curves, per_point=True)
    tlas0 = V.TopLevelAccelerationStructure(ctx, [blas0])
    before, ids0 = G.render(ctx, tlas0, cam, 256, 256)
    img = torch.zeros(256 * 256 * 4, dtype=torch.float32, device="cuda")
    print("[autotune blas]", flush=True)
    t = ctx.autotune_blas(V.GEOMETRY_PHANTOM_CURVES, d.data_ptr(), curves.shape[0], cam, 256, 256, img.data_ptr(),
                          blas_flags=V.BLAS_PER_POINT_RADIUS, repeats=2, stream=G.stream_ptr())
    print(t)
    _check_ranges(t)
    assert t["trials"] == 27  # collapse (greedy, SAH at primitive cost 0.5 and 1) x width x leaf size
    assert t["bvh_collapse"] in (0, 2, 4)
    blas1 = G.build_blas(ctx, V.GEOMETRY_PHANTOM_CURVES, curves, per_point=True)  # the tuned width and leaf size
    tlas1 = V.TopLevelAccelerationStructure(ctx, [blas1])
    assert blas1.info()["leaf_records"] == curves.shape[0]
    after, ids1 = G.render(ctx, tlas1, cam, 256, 256)
    assert np.array_equal(ids0, ids1) and np.array_equal(before.view(np.uint32), after.view(np.uint32))
    # every width builds a tree that renders the same image
    for width in (4, 6):
        ctx.set_option(V.OPTION_BVH_WIDTH, width)
        b = G.build_blas(ctx, V.GEOMETRY_PHANTOM_CURVES, curves, per_point=True)
        tl = V.TopLevelAccelerationStructure(ctx, [b])
        img_w, ids_w = G.render(ctx, tl, cam, 256, 256)
        assert np.array_equal(ids0, ids_w) and np.array_equal(before.view(np.uint32), img_w.view(np.uint32))
        assert b.info()["depth"] >= blas0.info()["depth"]


def test_autotune_batch_keeps_the_images(fresh_ctx):
    import gpu_util as G
    import scenes
    import vkhrt_amd as V
    ctx = fresh_ctx
    curves, blas, tlas = _scene(ctx)
    views, imgs = [], []
    for eye, tgt, th, (w, h) in scenes.DENSE_VIEWS:  # the test views at 4x their size
        w, h = 4 * w, 4 * h
        cam = V.camera_look_at(eye, tgt, fov_deg=60.0, aspect=w / h)
        imgs.append(torch.zeros(w * h * 4, dtype=torch.float32, device="cuda"))
        views.append((cam, w, h, imgs[-1].data_ptr()))

    def frame():
        ctx.render_batch(tlas, views, bounces=3, seed=5, environment=(0.3, 0.3, 0.3), stream=G.stream_ptr())
        torch.cuda.synchronize()
        ctx.synchronize(G.stream_ptr())
        return [im.cpu().numpy().copy() for im in imgs]

    before = frame()
    print("[autotune batch]", flush=True)
    t = ctx.autotune_batch(tlas, views, bounces=3, seed=5, environment=(0.3, 0.3, 0.3), repeats=2,
                           stream=G.stream_ptr())
    print(t)
    _check_ranges(t)
    assert t["batch_flags"] & ~(V.BATCH_SORT_MATERIALS | V.BATCH_SORT_RAYS) == 0
    ctx.render_batch(tlas, views, bounces=3, seed=5, environment=(0.3, 0.3, 0.3), flags=t["batch_flags"],
                     stream=G.stream_ptr())
    torch.cuda.synchronize()
    after = [im.cpu().numpy() for im in imgs]
    assert all(np.array_equal(a.view(np.uint32), b.view(np.uint32)) for a, b in zip(before, after))


def test_autotune_neural_render_keeps_the_image(fresh_ctx):
    """The neural form (area light, two samples, exact fallback above the threshold) tunes like
    the exact one, and switching between the two tunings with apply_tuning renders both forms
    bit for bit as before."""
    import gpu_util as G
    import vkhrt_amd as V
    ctx = fresh_ctx
    curves, blas, tlas = _scene(ctx)
    cam = V.camera_look_at((0.0, 20.0, 45.0), (0.0, 8.0, 0.0), aspect=1.0)
    light = V.area_light((0.0, 60.0, 0.0), (20.0, 0.0, 0.0), (0.0, 0.0, 20.0), samples=2, seed=3)
    mlp = V.ShadowMLP(ctx, [-30.0, -10.0, -30.0], [30.0, 40.0, 30.0], seed=5)  # untrained: p near 0.5
    W = H = 256
    flags = V.RENDER_SHADOWS | V.RENDER_FP64_POLISH

    def frame(neural):
        img = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
        kw = dict(shadow_mlp=mlp, variance_threshold=0.2499) if neural else {}
        ctx.render(tlas, cam, W, H, img.data_ptr(), flags=flags | (V.RENDER_NEURAL_SHADOWS if neural else 0),
                   stream=G.stream_ptr(), area_light=light, **kw)
        torch.cuda.synchronize()
        ctx.synchronize(G.stream_ptr())
        return img.cpu().numpy().view(np.uint32), ctx.render_stats()

    try:
        exact0, _ = frame(False)
        neural0, (_, q0, tr0) = frame(True)
        print(f"[autotune neural render] {tr0}/{q0} queries traced exactly", flush=True)
        assert tr0 > 0  # the exact fallback runs
        img = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
        print("[autotune neural render]", flush=True)
        te = ctx.autotune_render(tlas, cam, W, H, img.data_ptr(), flags=flags, repeats=2, stream=G.stream_ptr(),
                                 area_light=light)
        tn = ctx.autotune_render(tlas, cam, W, H, img.data_ptr(), flags=flags | V.RENDER_NEURAL_SHADOWS, repeats=2,
                                 stream=G.stream_ptr(), area_light=light, shadow_mlp=mlp, variance_threshold=0.2499)
        print(te, tn)
        _check_ranges(te)
        _check_ranges(tn)
        for _ in range(2):
            ctx.apply_tuning(tn)
            neural1, (_, q1, tr1) = frame(True)
            ctx.apply_tuning(te)
            exact1, _ = frame(False)
            assert (q1, tr1) == (q0, tr0)
            assert np.array_equal(neural0, neural1) and np.array_equal(exact0, exact1)
    finally:
        mlp.close()
