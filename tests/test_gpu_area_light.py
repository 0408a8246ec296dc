"""GPU tests of the area-light shadow form and of the surface sampler that feeds the shadow
MLP: renders against the oracle bit for bit, the neural form at threshold 0 against the exact
one, and the sampler's geometry (points on the fibre surface at the intersector's radius, rays
ending on the light) and distributions."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

LIGHT = np.array([[4.0, 45.0, -6.0], [25.0, 0.0, 5.0], [0.0, 3.0, 20.0]], np.float32)


def _light(V, samples, seed):
    return V.area_light(LIGHT[0], LIGHT[1], LIGHT[2], samples=samples, seed=seed)


@pytest.fixture(scope="module")
def ball(gpu_ctx, oracle):
    import gpu_util as G
    import scenes
    import vkhrt_amd as V
    curves = scenes.dense_ball_curves(nstrands=2500, nseg=12)
    S = oracle.Scene()
    S.add(oracle.MODE_PHANTOM, curves, per_point=True)
    blas = G.build_blas(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, curves, per_point=True)
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, [blas])
    yield S, tlas, curves
    tlas.close()
    blas.close()


@pytest.mark.parametrize("samples,polish", [(1, 0), (2, 0), (3, 1), (2, 1)])
def test_area_light_render_matches_oracle(gpu_ctx, oracle, ball, samples, polish):
    import gpu_util as G
    import vkhrt_amd as V
    S, tlas, _ = ball
    W, H = 160, 120
    vi, pi = oracle.camera((30.0, 25.0, 30.0), (0.0, 8.0, 0.0), aspect=W / H)
    cam = V.Camera()
    for k in range(16):
        cam.view_inverse[k], cam.proj_inverse[k] = float(vi[k]), float(pi[k])
    flags = V.RENDER_SHADOWS | (V.RENDER_FP64_POLISH if polish else 0)
    img = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
    ids = torch.zeros(W * H, dtype=torch.int32, device="cuda")
    print(f"[area light] samples {samples} polish {polish}", flush=True)
    gpu_ctx.render(tlas, cam, W, H, img.data_ptr(), flags=flags, prim_ids_ptr=ids.data_ptr(), stream=G.stream_ptr(),
                   area_light=_light(V, samples, 17))
    torch.cuda.synchronize()
    gpu_ctx.synchronize(G.stream_ptr())
    _, queries, traced = gpu_ctx.render_stats()
    got = img.cpu().numpy().reshape(H, W, 4)
    gids = ids.cpu().numpy().view(np.uint32).reshape(H, W)
    ref, rids, nrays, _ = S.render(vi, pi, W, H, flags=oracle.F_POLISH if polish else 0, shadows=1,
                                   area_light=LIGHT, light_samples=samples, light_seed=17)
    nhit = int((rids != 0xFFFFFFFF).sum())
    assert nhit > 2000 and queries == traced == samples * nhit and nrays == W * H + queries
    assert np.array_equal(gids, rids)
    assert np.array_equal(got[..., :3].view(np.uint32), ref.view(np.uint32))
    vis_levels = np.unique(np.round(((got[..., 0] - 0.3) / 0.4), 3))
    assert vis_levels.size > 3  # partly lit pixels occur


def test_area_light_neural_threshold_zero_is_exact(gpu_ctx, oracle, ball):
    import gpu_util as G
    import vkhrt_amd as V
    S, tlas, curves = ball
    W, H = 128, 96
    cam = V.camera_look_at((0.0, 20.0, 45.0), (0.0, 8.0, 0.0), aspect=W / H)
    pts = curves.reshape(-1, 4)[:, :3]
    mlp = V.ShadowMLP(gpu_ctx, pts.min(0) - 1.0, pts.max(0) + 1.0, seed=3)
    L = _light(V, 2, 5)
    exact, _ = G.render(gpu_ctx, tlas, cam, W, H)  # directional, for contrast
    out = {}
    for name, kw in (("exact", {}), ("t0", dict(shadow_mlp=mlp, variance_threshold=0.0)),
                     ("all", dict(shadow_mlp=mlp, variance_threshold=0.25))):
        img = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
        flags = V.RENDER_SHADOWS | (V.RENDER_NEURAL_SHADOWS if kw else 0)
        gpu_ctx.render(tlas, cam, W, H, img.data_ptr(), flags=flags, stream=G.stream_ptr(), area_light=L, **kw)
        torch.cuda.synchronize()
        out[name] = (img.cpu().numpy(), gpu_ctx.render_stats())
    assert np.array_equal(out["t0"][0], out["exact"][0])
    assert out["t0"][1][1] == out["t0"][1][2] == out["exact"][1][1] > 0
    assert out["all"][1][2] == 0 and out["all"][1][1] == out["exact"][1][1]
    assert not np.array_equal(out["exact"][0], exact.reshape(-1))


def test_area_light_gate_matches_infer(gpu_ctx, oracle, ball):
    """The gate of the neural form, at threshold 0.25 where the network answers every query: each
    hit pixel's visibility (read back through Shade) is the mean of the network's p over the
    pixel's two light samples, with p evaluated by the inference kernel on the queries formed
    from the traced primary hits and the light points (SplitMix64 through the oracle). The gate
    evaluates both samples of a hit in one wavefront column pair and shares their position
    features; a wrong pairing changes the mean."""
    import gpu_util as G
    import vkhrt_amd as V
    S, tlas, curves = ball
    W, H = 128, 96
    cam = V.camera_look_at((0.0, 20.0, 45.0), (0.0, 8.0, 0.0), aspect=W / H)
    pts = curves.reshape(-1, 4)[:, :3]
    lo, hi = pts.min(0) - 1.0, pts.max(0) + 1.0
    mlp = V.ShadowMLP(gpu_ctx, lo, hi, seed=7)
    seed = 11
    L = _light(V, 2, seed)
    img = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
    gpu_ctx.render(tlas, cam, W, H, img.data_ptr(), flags=V.RENDER_SHADOWS | V.RENDER_NEURAL_SHADOWS,
                   stream=G.stream_ptr(), area_light=L, shadow_mlp=mlp, variance_threshold=0.25)
    torch.cuda.synchronize()
    _, queries, traced = gpu_ctx.render_stats()
    assert traced == 0 and queries > 1000
    img = img.cpu().numpy().reshape(H * W, 4)
    rays_d = torch.empty(W * H * 8, dtype=torch.float32, device="cuda")
    gpu_ctx.generate_camera_rays(cam, W, H, rays_d.data_ptr(), stream=G.stream_ptr())
    torch.cuda.synchronize()
    rays = rays_d.cpu().numpy().reshape(-1, 8)
    hits = G.trace(gpu_ctx, tlas, rays)
    hit = np.nonzero(hits["prim"] != 0xFFFFFFFF)[0]
    assert 2 * len(hit) == queries
    # the queries as the render forms them, operation for operation in float32: origin o + t d,
    # light point c + a u + b v, direction w * (1 / sqrt(w . w))
    f32 = np.float32
    o = rays[hit, 0:3] + hits["t"][hit][:, None] * rays[hit, 4:7]
    smp = np.zeros((2 * len(hit), 8), np.float32)
    for s_ in range(2):
        a = np.array([f32(2.0) * f32(oracle.path_rand(seed, int(p), s_, 0)) - f32(1.0) for p in hit], np.float32)
        b = np.array([f32(2.0) * f32(oracle.path_rand(seed, int(p), s_, 1)) - f32(1.0) for p in hit], np.float32)
        y = (LIGHT[0][None, :] + a[:, None] * LIGHT[1][None, :]) + b[:, None] * LIGHT[2][None, :]
        w = y - o
        dd = (w[:, 0] * w[:, 0] + w[:, 1] * w[:, 1]) + w[:, 2] * w[:, 2]
        r = f32(1.0) / np.sqrt(dd)
        smp[s_::2, 0:3] = o
        smp[s_::2, 4:7] = w * r[:, None]
    sd = torch.from_numpy(smp).cuda()
    pd = torch.empty(len(smp), dtype=torch.float32, device="cuda")
    mlp.infer(sd.data_ptr(), len(smp), pd.data_ptr(), stream=G.stream_ptr())
    torch.cuda.synchronize()
    p = pd.cpu().numpy()
    v_want = (p[0::2] + p[1::2]) / 2.0
    ndl = np.abs(hits["n"][hit][:, 1])
    ok = ndl > 0.2
    v_got = (img[hit, 0] - 0.3) / (0.4 * ndl)
    err = np.abs(v_got - v_want)[ok]
    print(f"[gate vs infer] {ok.sum()} pixels, visibility error max {err.max():.3g}, mean {err.mean():.3g}; "
          f"p spread {p.min():.3f}..{p.max():.3f}")
    assert ok.sum() > 300 and err.max() < 1e-4
    mlp.close()


def _straight_fibres(n, rng):
    """n straight fibres (evenly spaced control points) with linearly tapered radii."""
    a = rng.uniform(-20, 20, (n, 3))
    b = a + rng.normal(size=(n, 3)) * 6.0
    r0, r1 = rng.uniform(0.2, 0.6, n), rng.uniform(0.05, 0.2, n)
    c = np.zeros((n, 4, 4), np.float32)
    for k in range(4):
        c[:, k, :3] = a + (b - a) * (k / 3)
        c[:, k, 3] = r0 + (r1 - r0) * (k / 3)
    return c.reshape(n, 16)


def test_surface_shadow_rays_geometry_and_distribution(gpu_ctx, oracle):
    import gpu_util as G
    import vkhrt_amd as V
    rng = np.random.default_rng(4)
    f0, f1 = _straight_fibres(300, rng), _straight_fibres(100, rng)
    xf = np.array([[0.0, -1.0, 0.0, 50.0], [1.0, 0.0, 0.0, -3.0], [0.0, 0.0, 1.0, 7.0]], np.float32)
    b0 = G.build_blas(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, f0, per_point=True)
    b1 = G.build_blas(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, f1, per_point=True, transform=xf.reshape(12))
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, [b0, b1])
    n = 200000
    rays = torch.zeros(n * 8, dtype=torch.float32, device="cuda")
    print("[surface sampler] 200k rays", flush=True)
    gpu_ctx.generate_surface_shadow_rays(tlas, n, 123, rays.data_ptr(), light=_light(V, 1, 0), stream=G.stream_ptr())
    torch.cuda.synchronize()
    r = rays.cpu().numpy().reshape(n, 8).astype(np.float64)
    o, d, tmax = r[:, 0:3], r[:, 4:7], r[:, 7]
    assert np.all(r[:, 3] == np.float32(0.001))
    assert np.abs(np.linalg.norm(d, axis=1) - 1).max() < 1e-5
    # the ray ends on the light parallelogram: o + tmax d = c + a u + b v with |a|, |b| <= 1
    end = o + tmax[:, None] * d - LIGHT[0]
    ab = np.linalg.lstsq(LIGHT[1:].T.astype(np.float64), end.T, rcond=None)[0].T
    resid = end - ab @ LIGHT[1:].astype(np.float64)
    assert np.abs(resid).max() < 2e-3 and np.abs(ab).max() <= 1.0 + 1e-4
    # each origin lies on one fibre's surface: distance to its axis = its radius at the
    # projection (straight fibres: the closest axis point is P(t), r(t) linear)
    world1 = np.concatenate([f1.reshape(-1, 4, 4)[:, :, :3] @ xf[:, :3].T + xf[:, 3],
                             f1.reshape(-1, 4, 4)[:, :, 3:]], axis=2).reshape(-1, 16)
    fib = np.concatenate([f0, world1]).reshape(-1, 4, 4).astype(np.float64)
    a, b = fib[:, 0, :3], fib[:, 3, :3]
    ra, rb = fib[:, 0, 3], fib[:, 3, 3]
    best = np.full(n, np.inf)
    which = np.zeros(n, np.int64)
    for i in range(fib.shape[0]):
        ab_ = b[i] - a[i]
        t = np.clip(((o - a[i]) @ ab_) / (ab_ @ ab_), 0.0, 1.0)
        dist = np.linalg.norm(o - (a[i] + t[:, None] * ab_[None, :]), axis=1)
        err = np.abs(dist - (ra[i] + (rb[i] - ra[i]) * t))
        upd = err < best
        best[upd] = err[upd]
        which[upd] = i
    assert np.quantile(best, 0.999) < 2e-4 and best.max() < 1e-3
    # records are drawn uniformly over both instances (300 : 100)
    counts = np.bincount(which, minlength=fib.shape[0])
    frac1 = counts[300:].sum() / n
    assert abs(frac1 - 0.25) < 4 * math.sqrt(0.25 * 0.75 / n)
    chi2 = ((counts - n / 400) ** 2 / (n / 400)).sum()
    assert chi2 < 400 + 6 * math.sqrt(800)
    tlas.close()
    b0.close()
    b1.close()


def test_surface_samples_train_the_mlp(gpu_ctx, oracle, ball):
    """Surface samples labelled by the exact tracer train the MLP: held-out samples (another
    seed) are classified better than by the majority label."""
    import gpu_util as G
    import vkhrt_amd as V
    _, tlas, curves = ball
    pts = curves.reshape(-1, 4)[:, :3]
    mlp = V.ShadowMLP(gpu_ctx, pts.min(0) - 1.0, pts.max(0) + 1.0, seed=2)
    L = _light(V, 1, 0)

    def data(n, seed):
        rays = torch.zeros(n * 8, dtype=torch.float32, device="cuda")
        gpu_ctx.generate_surface_shadow_rays(tlas, n, seed, rays.data_ptr(), light=L, stream=G.stream_ptr())
        smp = torch.zeros(n * 8, dtype=torch.float32, device="cuda")
        tgt = torch.zeros(n, dtype=torch.float32, device="cuda")
        mlp.collect(tlas, rays.data_ptr(), n, smp.data_ptr(), tgt.data_ptr(), stream=G.stream_ptr())
        torch.cuda.synchronize()
        return smp, tgt

    print("[surface training] 2^17 samples", flush=True)
    smp, tgt = data(1 << 17, 1)
    hs, ht = data(1 << 15, 2)
    mlp.train(smp.data_ptr(), tgt.data_ptr(), 1 << 17, batch=8192, epochs=12, learning_rate=3e-3)
    p = torch.zeros(1 << 15, dtype=torch.float32, device="cuda")
    mlp.infer(hs.data_ptr(), 1 << 15, p.data_ptr())
    torch.cuda.synchronize()
    y = ht.cpu().numpy()
    acc = float(((p.cpu().numpy() > 0.5) == (y > 0.5)).mean())
    base = max(y.mean(), 1 - y.mean())
    print(f"lit fraction {y.mean():.3f}; held-out accuracy {acc:.4f} vs majority {base:.4f}")
    assert 0.05 < y.mean() < 0.95
    assert acc > base + 0.03
