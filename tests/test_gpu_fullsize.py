"""Parity of the HIP path with the oracle at the workloads' full sizes.

- 100k strands x 32 tapered segments (3.2M phantom segments), 1920x1080 primary rays plus a
  shadow ray per hit, FP64 polish: the whole frame against the oracle, plain and with
  double-word accurate hits (then hit distance and normal against the FP64 intersector), and
  the frame shaded with the reference's literal normal formula.
- The same hair as 12.8M Linear Swept Spheres and as 12.8M DOTS pieces (51.2M triangles),
  same frame.
- The batch: 4 models (10k / 50k / 200k / 500k strands x 32 = 24.32M segments), 64 views
  512x512 .. 1920x1080, 4 path-traced bounces, FP64 polish; every 61st path against the
  oracle.
- 500k strands x 32 (16M segments) at 1920x1080 with 2 area-light samples per hit: the exact
  frame against the oracle, then the neural shadow form trained on 2^24 surface rays,
  against the exact frame with its exact fallback active.

Bars (the tolerances are written in each test): primitive ids agree on >= 99.99 % of the
pixels and the shaded float image reaches PSNR >= 50 dB (the kernels follow the oracle's
arithmetic, so both are met bit for bit today); the neural form reaches PSNR >= 40 dB.
"""
import math
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SEED = 20240601
W, H = 1920, 1080
EYE, TARGET = (0.0, 60.0, 300.0), (0.0, 40.0, 0.0)
ID_AGREEMENT = 0.9999
PSNR_DB = 50.0


def _log(*a):
    print(*a, flush=True)


def _gpu_hair(ctx, seed, nstrands, nseg):
    """Seeded strands -> lines -> GenerateCurves on the device: (count*16) float32 tensor."""
    import gpu_util as G
    n = nstrands * nseg
    verts = torch.empty(nstrands * (nseg + 1) * 4, dtype=torch.float32, device="cuda")
    lines = torch.empty(n * 8, dtype=torch.float32, device="cuda")
    curves = torch.empty(n * 16, dtype=torch.float32, device="cuda")
    ctx.generate_strands(seed, nstrands, nseg, verts.data_ptr(), stream=G.stream_ptr())
    ctx.generate_lines(verts.data_ptr(), nstrands, nseg, lines.data_ptr(), stream=G.stream_ptr())
    ctx.generate_curves(lines.data_ptr(), n, curves.data_ptr(), stream=G.stream_ptr())
    torch.cuda.synchronize()
    return curves


def _render(ctx, tlas, cam, w, h, flags, **kw):
    import gpu_util as G
    img = torch.zeros(w * h * 4, dtype=torch.float32, device="cuda")
    ids = torch.zeros(w * h, dtype=torch.int32, device="cuda")
    ctx.render(tlas, cam, w, h, img.data_ptr(), flags=flags, prim_ids_ptr=ids.data_ptr(), stream=G.stream_ptr(), **kw)
    torch.cuda.synchronize()
    ctx.synchronize(G.stream_ptr())
    return img.cpu().numpy().reshape(h, w, 4)[..., :3], ids.cpu().numpy().view(np.uint32).reshape(h, w)


def _agree(what, img, ids, ref_img, ref_ids, psnr_bar=PSNR_DB):
    import gpu_util as G
    same_id = float((ids == ref_ids).mean())
    same_px = float((img.view(np.uint32) == ref_img.view(np.uint32)).all(-1).mean())
    p = G.psnr(img, ref_img)
    hits = int((ref_ids != 0xFFFFFFFF).sum())
    _log(f"{what}: {ids.size} pixels, {hits} hits; primitive ids agree {same_id:.6f}, "
         f"bit-identical pixels {same_px:.6f}, PSNR {p:.2f} dB")
    assert hits > 0.05 * ids.size
    assert same_id >= ID_AGREEMENT, f"{what}: primitive-id agreement {same_id}"
    assert p >= psnr_bar, f"{what}: PSNR {p} dB"
    return same_id, same_px, p


@pytest.fixture(scope="module")
def c1(gpu_ctx, oracle):
    """The 100k x 32 hair on both sides, generated independently; the device curves equal the
    oracle's bit for bit."""
    import vkhrt_amd as V
    import gpu_util as G
    t0 = time.perf_counter()
    curves_d = _gpu_hair(gpu_ctx, SEED, 100000, 32)
    curves = oracle.hair_curves(SEED, 100000, 32)
    assert np.array_equal(curves_d.cpu().numpy().view(np.uint32).reshape(-1, 16), curves.view(np.uint32))
    blas = V.BottomLevelAccelerationStructure(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, curves_d.data_ptr(),
                                              curves.shape[0], flags=V.BLAS_PER_POINT_RADIUS, stream=G.stream_ptr())
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, [blas], stream=G.stream_ptr())
    S = oracle.Scene()
    S.add(oracle.MODE_PHANTOM, curves, per_point=True)
    cam = V.camera_look_at(EYE, TARGET, fov_deg=60.0, aspect=W / H)
    _log(f"[c1] 3.2M segments on both sides in {time.perf_counter() - t0:.1f} s")
    yield {"curves": curves, "curves_d": curves_d, "tlas": tlas, "S": S, "cam": cam}
    tlas.close()
    blas.close()


def test_c1_full_frame(gpu_ctx, oracle, c1):
    """The whole 1920x1080 primary + shadow frame with FP64 polish against the oracle."""
    import gpu_util as G
    import vkhrt_amd as V
    vi, pi = G.cam_arrays(c1["cam"])
    _log("[c1] frame: GPU")
    img, ids = _render(gpu_ctx, c1["tlas"], c1["cam"], W, H, V.RENDER_SHADOWS | V.RENDER_FP64_POLISH)
    n_primary, n_shadow = gpu_ctx.render_ray_counts()
    _log("[c1] frame: oracle")
    ref, rids, nrays, secs = c1["S"].render(vi, pi, W, H, flags=oracle.F_POLISH, shadows=1)
    assert n_primary + n_shadow == nrays
    _agree("c1 frame (FP64 polish)", img, ids, ref, rids)


def test_c1_accurate_hits_and_fp64_accuracy(gpu_ctx, oracle, c1):
    """The accurate frame (every primary phantom hit re-run in double-word FP32) against the
    oracle, and its hits against the FP64 intersector on rays spread over the frame: normal
    within 1e-4 rad and hit distance within 1e-5 relative error on every evaluated hit."""
    import accuracy
    import gpu_util as G
    import vkhrt_amd as V
    vi, pi = G.cam_arrays(c1["cam"])
    flags = V.RENDER_SHADOWS | V.RENDER_FP64_POLISH | V.RENDER_ACCURATE_NORMALS
    _log("[c1 accurate] frame")
    img, ids = _render(gpu_ctx, c1["tlas"], c1["cam"], W, H, flags)
    ref, rids, _, _ = c1["S"].render(vi, pi, W, H, flags=oracle.F_POLISH | oracle.F_ACCURATE, shadows=1)
    _agree("c1 accurate frame", img, ids, ref, rids)
    rays = oracle.raygen(vi, pi, W, H)
    sel = np.nonzero(rids.reshape(-1) != 0xFFFFFFFF)[0]
    sel = sel[np.linspace(0, sel.size - 1, 3000).astype(np.int64)]
    r = np.ascontiguousarray(rays[sel])
    out = {}
    for name, fl in (("fp32", V.TRACE_FP64_POLISH), ("accurate", V.TRACE_FP64_POLISH | V.TRACE_ACCURATE_NORMALS)):
        h = G.trace(gpu_ctx, c1["tlas"], r, flags=fl)
        out[name] = accuracy.phantom(r, h, c1["curves"], limit=3000)
        _log(f"[c1] vs FP64, {name}: {out[name]}")
    a = out["accurate"]
    assert a["hits"] > 2500
    assert a["normal_rad_max"] <= 1e-4 and a["t_rel_max"] <= 1e-5


def test_c1_reference_normal_formula(gpu_ctx, oracle, c1):
    """The product's frame against the oracle frame shaded with the reference's literal normal
    normalize(o + t d - SampleCurvePoint(t)) (hair_intersection.rint:74-76): same hits,
    PSNR >= 50 dB."""
    import gpu_util as G
    import vkhrt_amd as V
    vi, pi = G.cam_arrays(c1["cam"])
    img, ids = _render(gpu_ctx, c1["tlas"], c1["cam"], W, H, V.RENDER_SHADOWS | V.RENDER_FP64_POLISH)
    ref, rids, _, _ = c1["S"].render(vi, pi, W, H, flags=oracle.F_POLISH | oracle.F_REF_NORMAL, shadows=1)
    _agree("c1 frame vs the reference normal formula", img, ids, ref, rids)


@pytest.mark.parametrize("mode", ["lss", "dots"])
def test_c2_full_frame(gpu_ctx, oracle, c1, mode):
    """The hair as 4 linear pieces per cubic (12.8M), traced as LSS capsules or DOTS quads
    (51.2M triangles), same 1080p primary + shadow frame, against the oracle; the pieces the
    device generates equal the oracle's bit for bit. Hit accuracy against FP64 is printed."""
    import accuracy
    import gpu_util as G
    import vkhrt_amd as V
    ctx = gpu_ctx
    n = c1["curves"].shape[0] * 4
    pl = torch.empty(n * 8, dtype=torch.float32, device="cuda")
    ctx.curves_to_lines(c1["curves_d"].data_ptr(), c1["curves"].shape[0], 4, pl.data_ptr(), stream=G.stream_ptr())
    lines = oracle.curves_to_lines(c1["curves"], 4)
    S = oracle.Scene()
    if mode == "lss":
        prims_d = torch.empty(n * 8, dtype=torch.float32, device="cuda")
        ctx.generate_lss(pl.data_ptr(), n, prims_d.data_ptr(), stream=G.stream_ptr())
        prims = oracle.generate_lss(lines)
        torch.cuda.synchronize()
        assert np.array_equal(prims_d.cpu().numpy().view(np.uint32).reshape(-1, 8), prims.view(np.uint32))
        gmode, flags = V.GEOMETRY_LSS, 0
        S.add(oracle.MODE_LSS, prims)
    else:
        torch.cuda.synchronize()
        assert np.array_equal(pl.cpu().numpy().view(np.uint32).reshape(-1, 8), lines.view(np.uint32))
        prims_d, gmode, flags = pl, V.GEOMETRY_DOTS, V.BLAS_PER_POINT_RADIUS
        S.add(oracle.MODE_DOTS, lines, per_point=True)
    blas = V.BottomLevelAccelerationStructure(ctx, gmode, prims_d.data_ptr(), n, flags=flags, stream=G.stream_ptr())
    tlas = V.TopLevelAccelerationStructure(ctx, [blas], stream=G.stream_ptr())
    vi, pi = G.cam_arrays(c1["cam"])
    _log(f"[c2 {mode}] frame")
    img, ids = _render(ctx, tlas, c1["cam"], W, H, V.RENDER_SHADOWS)
    ref, rids, _, _ = S.render(vi, pi, W, H, shadows=1)
    _agree(f"c2 {mode} frame", img, ids, ref, rids)
    rays = oracle.raygen(vi, pi, W, H)
    sel = np.nonzero(rids.reshape(-1) != 0xFFFFFFFF)[0]
    r = np.ascontiguousarray(rays[sel[np.linspace(0, sel.size - 1, 2000).astype(np.int64)]])
    h = G.trace(ctx, tlas, r)
    acc = (accuracy.lss(r, h, prims, limit=2000) if mode == "lss"
           else accuracy.dots(r, h, oracle.generate_dots(lines), limit=2000))
    _log(f"[c2 {mode}] vs FP64: {acc}")
    assert acc["hits"] > 1500 and acc["t_rel_max"] <= 1e-5
    tlas.close()
    blas.close()


BATCH_MODELS = [(10000, 101), (50000, 102), (200000, 103), (500000, 104)]  # (strands, seed), 32 segments
BATCH_RES = [(512, 512), (1024, 576), (1280, 720), (1920, 1080)]
BATCH_TINTS = [(1.0, 1.0, 1.0), (0.9, 0.7, 0.5), (0.6, 0.8, 1.0), (1.0, 0.6, 0.6)]
BATCH_ENV = (0.25, 0.3, 0.35)


def test_c3_full_batch(gpu_ctx, oracle):
    """4 models of uneven size (24.32M segments) side by side, 64 views of mixed resolution
    (61.5M paths), 4 bounces, FP64 polish, as a captured graph with material- and ray-sorted
    queues; every 61st path (about 1M paths) against the oracle bit for bit (radiance and primary
    id), and the frame's bounce and shadow ray totals against the eager, unsorted frame."""
    import gpu_util as G
    import vkhrt_amd as V
    ctx = gpu_ctx
    S = oracle.Scene()
    blases, keep = [], []
    t0 = time.perf_counter()
    for k, (ns, seed) in enumerate(BATCH_MODELS):
        xf = np.array([1.0, 0.0, 0.0, 500.0 * k - 750.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0], np.float32)
        cd = _gpu_hair(ctx, seed, ns, 32)
        keep.append(cd)
        blases.append(V.BottomLevelAccelerationStructure(ctx, V.GEOMETRY_PHANTOM_CURVES, cd.data_ptr(), ns * 32,
                                                         flags=V.BLAS_PER_POINT_RADIUS, transform=xf,
                                                         material_index=k, stream=G.stream_ptr()))
        S.add(oracle.MODE_PHANTOM, oracle.hair_curves(seed, ns, 32), per_point=True, transform=xf.reshape(3, 4))
        S.set_material(k, k)
    tlas = V.TopLevelAccelerationStructure(ctx, blases, stream=G.stream_ptr())
    torch.cuda.synchronize()
    _log(f"[c3] 4 models on both sides in {time.perf_counter() - t0:.1f} s")
    views, gviews, imgs, pids = [], [], [], []
    for v in range(64):
        k, j = v % 4, v // 4
        w, h = BATCH_RES[(v + j) % 4]
        a = 2 * math.pi * (j + 0.5) / 16
        cx = 500.0 * k - 750.0
        eye = (cx + 280 * math.sin(a), 60.0 + 90 * math.sin(3 * a), 280 * math.cos(a))
        cam = V.camera_look_at(eye, (cx, 40.0, 0.0), fov_deg=60.0, aspect=w / h)
        vi, pi = G.cam_arrays(cam)
        views.append((vi, pi, w, h))
        img = torch.zeros(w * h * 4, dtype=torch.float32, device="cuda")
        pid = torch.zeros(w * h, dtype=torch.int32, device="cuda")
        imgs.append(img)
        pids.append(pid)
        gviews.append((cam, w, h, img.data_ptr(), pid.data_ptr()))
    paths = sum(w * h for _, _, w, h in views)
    totals = {}
    for name, fl in (("eager", V.BATCH_FP64_POLISH),
                     ("graph_sorted", V.BATCH_FP64_POLISH | V.BATCH_GRAPH | V.BATCH_SORT_MATERIALS |
                      V.BATCH_SORT_RAYS)):
        ctx.render_batch(tlas, gviews, bounces=4, seed=9, environment=BATCH_ENV, flags=fl,
                         material_albedo=BATCH_TINTS, stream=G.stream_ptr())
        torch.cuda.synchronize()
        ctx.synchronize(G.stream_ptr())
        totals[name] = ctx.batch_stats()
        _log(f"[c3] {name}: {paths} paths, stats {totals[name]}")
    assert totals["eager"][1:3] == totals["graph_sorted"][1:3]
    stride = 61
    t0 = time.perf_counter()
    ri, rp, nb, ns, _ = S.render_batch(views, bounces=4, seed=9, env=BATCH_ENV, albedo=BATCH_TINTS,
                                       flags=oracle.F_POLISH, path_stride=stride)
    _log(f"[c3] oracle on every {stride}th path in {time.perf_counter() - t0:.1f} s")
    g_img = np.concatenate([x.cpu().numpy().reshape(-1, 4)[:, :3] for x in imgs])[::stride]
    g_ids = np.concatenate([p.cpu().numpy().view(np.uint32) for p in pids])[::stride]
    r_img = np.concatenate([x.reshape(-1, 3) for x in ri])[::stride]
    r_ids = np.concatenate([p.reshape(-1) for p in rp])[::stride]
    same_id = float((g_ids == r_ids).mean())
    same_px = float((g_img.view(np.uint32) == r_img.view(np.uint32)).all(-1).mean())
    hits = int((r_ids != 0xFFFFFFFF).sum())
    _log(f"[c3] {g_ids.size} sampled paths, {hits} primary hits: ids agree {same_id:.6f}, "
         f"bit-identical radiance {same_px:.6f}")
    assert hits > 0.1 * g_ids.size
    assert same_id >= ID_AGREEMENT and same_px >= ID_AGREEMENT
    tlas.close()
    for b in blases:
        b.close()


AREA_LIGHT = ((0.0, 260.0, 0.0), (100.0, 0.0, 0.0), (0.0, 0.0, 100.0))


def test_c4_neural_shadows_full(gpu_ctx, oracle):
    """16M segments, 1080p, 2 area-light samples per hit. The exact frame against the oracle;
    the MLP trained on 2^24 exact surface shadow rays; the neural frame at the largest
    threshold below 0.25 of a fixed list whose frame keeps PSNR >= 40 dB against the exact one,
    with part of the queries traced by the exact fallback; threshold 0 reproduces the exact
    frame bit for bit."""
    import gpu_util as G
    import vkhrt_amd as V
    ctx = gpu_ctx
    ns, nseg, seed = 500000, 32, 104
    curves_d = _gpu_hair(ctx, seed, ns, nseg)
    blas = V.BottomLevelAccelerationStructure(ctx, V.GEOMETRY_PHANTOM_CURVES, curves_d.data_ptr(), ns * nseg,
                                              flags=V.BLAS_PER_POINT_RADIUS, stream=G.stream_ptr())
    tlas = V.TopLevelAccelerationStructure(ctx, [blas], stream=G.stream_ptr())
    cam = V.camera_look_at(EYE, TARGET, fov_deg=60.0, aspect=W / H)
    light = V.area_light(*AREA_LIGHT, samples=2, seed=7)
    flags = V.RENDER_SHADOWS | V.RENDER_FP64_POLISH
    _log("[c4] exact frame")
    exact, ids = _render(ctx, tlas, cam, W, H, flags, area_light=light)
    S = oracle.Scene()
    S.add(oracle.MODE_PHANTOM, oracle.hair_curves(seed, ns, nseg), per_point=True)
    vi, pi = G.cam_arrays(cam)
    ref, rids, _, _ = S.render(vi, pi, W, H, flags=oracle.F_POLISH, shadows=1,
                               area_light=np.array(AREA_LIGHT, np.float32), light_samples=2, light_seed=7)
    _agree("c4 exact frame (2 area-light samples)", exact, ids, ref, rids)
    del S
    cv = curves_d.view(-1, 4, 4)[:, :, :3].reshape(-1, 3)
    lo, hi = cv.min(0).values.tolist(), cv.max(0).values.tolist()
    pad = [0.02 * (b - a) + 1.0 for a, b in zip(lo, hi)]
    mlp = V.ShadowMLP(ctx, [a - p for a, p in zip(lo, pad)], [b + p for b, p in zip(hi, pad)], seed=1)
    n = 1 << 24
    rays = torch.empty(n * 8, dtype=torch.float32, device="cuda")
    ctx.generate_surface_shadow_rays(tlas, n, 1, rays.data_ptr(), light=light, stream=G.stream_ptr())
    smp = torch.empty(n * 8, dtype=torch.float32, device="cuda")
    tgt = torch.empty(n, dtype=torch.float32, device="cuda")
    mlp.collect(tlas, rays.data_ptr(), n, smp.data_ptr(), tgt.data_ptr(), stream=G.stream_ptr())
    del rays
    _log("[c4] training on 2^24 surface rays")
    loss = mlp.train(smp.data_ptr(), tgt.data_ptr(), n, batch=65536, epochs=24, learning_rate=3e-3,
                     stream=G.stream_ptr())
    del smp, tgt
    nflags = flags | V.RENDER_NEURAL_SHADOWS
    same, sids = _render(ctx, tlas, cam, W, H, nflags, area_light=light, shadow_mlp=mlp, variance_threshold=0.0)
    assert np.array_equal(same.view(np.uint32), exact.view(np.uint32)) and np.array_equal(sids, ids)
    chosen = None
    for thr in (0.2499, 0.249, 0.245, 0.24, 0.23, 0.2, 0.15, 0.1):
        img, _ = _render(ctx, tlas, cam, W, H, nflags, area_light=light, shadow_mlp=mlp, variance_threshold=thr)
        _, q, tr = ctx.render_stats()
        p = G.psnr(img, exact)
        _log(f"[c4] threshold {thr}: PSNR {p:.2f} dB vs exact, {tr}/{q} queries traced exactly")
        if p >= 40.0 and tr > 0:
            chosen = (thr, p, q, tr)
            break
    _log(f"[c4] training loss {loss:.4f}; chosen {chosen}")
    assert chosen is not None, "no threshold with exact fallback reaches 40 dB"
    assert 0 < chosen[3] < chosen[2]
    mlp.close()
    tlas.close()
    blas.close()
