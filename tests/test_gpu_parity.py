"""GPU parity: the HIP path (through the C-ABI) against the oracle on seeded inputs.

Integer / index results (hit/miss, primitive id, instance id) must agree exactly; the FP32
intersector arithmetic follows the same IEEE operation order on both sides, so t, u and
normals are compared bit-for-bit as well (tolerance 0).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def G():
    import gpu_util
    return gpu_util


@pytest.fixture(scope="module")
def curves():
    import scenes
    return scenes.c0_curves()


def _compare_hits(h, ref, what, exact=True):
    miss_g = h["prim"] == 0xFFFFFFFF
    miss_r = ref["prim"] == 0xFFFFFFFF
    agree = (h["prim"] == ref["prim"]) & (h["inst"] == ref["inst"])
    frac = agree.mean()
    print(f"{what}: {len(h)} rays, hits {(~miss_r).sum()}, primId agreement {frac:.6f}")
    assert (miss_g == miss_r).all(), f"{what}: hit/miss differs on {(miss_g != miss_r).sum()} rays"
    assert agree.all(), f"{what}: primitive id differs on {(~agree).sum()} rays"
    hit = ~miss_r
    if exact:
        assert np.array_equal(h["t"][hit].view(np.uint32), ref["t"][hit].view(np.uint32)), f"{what}: t differs"
        assert np.array_equal(h["u"][hit].view(np.uint32), ref["u"][hit].view(np.uint32)), f"{what}: u differs"
        assert np.array_equal(h["n"][hit].view(np.uint32), ref["n"][hit].view(np.uint32)), f"{what}: normal differs"


@pytest.mark.parametrize("per_point", [False, True])
def test_phantom_trace_targeted(gpu_ctx, G, oracle, curves, per_point):
    import scenes
    import vkhrt_amd as V
    rays = scenes.targeted_rays(curves, 20000, seed=7 + per_point)
    S = oracle.Scene()
    S.add(oracle.MODE_PHANTOM, curves, per_point=per_point)
    ref = S.trace(rays)
    blas = G.build_blas(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, curves, per_point=per_point)
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, [blas])
    h = G.trace(gpu_ctx, tlas, rays)
    _compare_hits(h, ref, f"phantom per_point={per_point}")
    # shadow (any-hit) form: visibility must agree exactly
    hs = G.trace(gpu_ctx, tlas, rays, flags=V.TRACE_TERMINATE_ON_FIRST_HIT)
    rs = S.trace(rays, flags=oracle.F_ANY_HIT)
    assert np.array_equal(hs["prim"] == 0xFFFFFFFF, rs["prim"] == 0xFFFFFFFF)


def test_phantom_render_c0(gpu_ctx, G, oracle, curves):
    import vkhrt_amd as V
    W = H = 256
    cam = V.camera_look_at((0.0, 50.0, 400.0), (0.0, 30.0, 0.0), fov_deg=60.0, aspect=1.0)
    vi, pi = G.cam_arrays(cam)
    S = oracle.Scene()
    S.add(oracle.MODE_PHANTOM, curves, per_point=True)
    ref_img, ref_ids, _, _ = S.render(vi, pi, W, H, shadows=1)
    blas = G.build_blas(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, curves, per_point=True)
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, [blas])
    img, ids = G.render(gpu_ctx, tlas, cam, W, H)
    assert np.array_equal(ids, ref_ids)
    p = G.psnr(img[..., :3], ref_img)
    print("C0 render PSNR", p)
    assert p >= 50.0
    assert np.array_equal(img[..., :3].view(np.uint32), ref_img.view(np.uint32))


def test_geometry_front_end(gpu_ctx, G, oracle):
    """Seeded strands, GenerateLines, GenerateCurves, cubic->lines, LSS and DOTS on the GPU
    are bit-identical to the oracle."""
    import vkhrt_amd as V
    ns, nseg, seed = 300, 16, 99
    verts = torch.zeros(ns * (nseg + 1) * 4, dtype=torch.float32, device="cuda")
    gpu_ctx.generate_strands(seed, ns, nseg, verts.data_ptr(), stream=G.stream_ptr())
    lines = torch.zeros(ns * nseg * 8, dtype=torch.float32, device="cuda")
    gpu_ctx.generate_lines(verts.data_ptr(), ns, nseg, lines.data_ptr(), stream=G.stream_ptr())
    curves = torch.zeros(ns * nseg * 16, dtype=torch.float32, device="cuda")
    gpu_ctx.generate_curves(lines.data_ptr(), ns * nseg, curves.data_ptr(), stream=G.stream_ptr())
    pl = torch.zeros(ns * nseg * 4 * 8, dtype=torch.float32, device="cuda")
    gpu_ctx.curves_to_lines(curves.data_ptr(), ns * nseg, 4, pl.data_ptr(), stream=G.stream_ptr())
    caps = torch.zeros(ns * nseg * 4 * 8, dtype=torch.float32, device="cuda")
    gpu_ctx.generate_lss(pl.data_ptr(), ns * nseg * 4, caps.data_ptr(), stream=G.stream_ptr())
    dots = torch.zeros(ns * nseg * 4 * 16, dtype=torch.float32, device="cuda")
    gpu_ctx.generate_dots(pl.data_ptr(), ns * nseg * 4, dots.data_ptr(), stream=G.stream_ptr())
    torch.cuda.synchronize()

    rv = oracle.generate_strands(seed, ns, nseg)
    rl = oracle.strands_to_lines(rv)
    rc = oracle.generate_curves(rl)
    rpl = oracle.curves_to_lines(rc, 4)
    rcaps = oracle.generate_lss(rpl)
    rdots = oracle.generate_dots(rpl)
    for name, g, r in [("strands", verts, rv), ("lines", lines, rl), ("curves", curves, rc),
                       ("pieces", pl, rpl), ("lss", caps, rcaps), ("dots", dots, rdots)]:
        gh = g.cpu().numpy().reshape(r.shape)
        assert np.array_equal(gh.view(np.uint32), r.view(np.uint32)), name


@pytest.mark.parametrize("steps", [1, 3, 6])
@pytest.mark.parametrize("mode", ["lss", "dots"])
def test_lss_dots_trace(gpu_ctx, G, oracle, curves, mode, steps):
    """Closest hits of capsules and triangle pieces bit for bit, with 1, 3 and 6 traversal steps
    per iteration of the persistent loop; any-hit decisions too."""
    import scenes
    import vkhrt_amd as V
    lines = scenes.lines_from_curves(curves[:4000], 4)
    rays = scenes.targeted_rays(curves[:4000], 20000, seed=11)
    S = oracle.Scene()
    if mode == "lss":
        caps = oracle.generate_lss(lines)
        S.add(oracle.MODE_LSS, caps)
        blas = G.build_blas(gpu_ctx, V.GEOMETRY_LSS, caps)
    else:
        S.add(oracle.MODE_DOTS, lines, per_point=True)
        blas = G.build_blas(gpu_ctx, V.GEOMETRY_DOTS, lines, per_point=True)
    ref = S.trace(rays)
    rs = S.trace(rays, flags=oracle.F_ANY_HIT)
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, [blas])
    gpu_ctx.set_option(V.OPTION_STEPS_PER_ITERATION, steps)
    try:
        h = G.trace(gpu_ctx, tlas, rays)
        hs = G.trace(gpu_ctx, tlas, rays, flags=V.TRACE_TERMINATE_ON_FIRST_HIT)
    finally:
        gpu_ctx.set_option(V.OPTION_STEPS_PER_ITERATION, 3)
    _compare_hits(h, ref, f"{mode}, {steps} steps per iteration")
    assert np.array_equal(hs["prim"] == 0xFFFFFFFF, rs["prim"] == 0xFFFFFFFF)


def test_instances_with_transforms(gpu_ctx, G, oracle, curves):
    """Several instances (identity and rigid transforms): closest hit across instances with
    the (instance, primitive) tie rule."""
    import scenes
    import vkhrt_amd as V
    xf1 = np.array([1, 0, 0, 40.0, 0, 1, 0, -10.0, 0, 0, 1, 5.0], np.float32)
    c, s = np.cos(0.3), np.sin(0.3)
    xf2 = np.array([c, 0, s, -30.0, 0, 1, 0, 0.0, -s, 0, c, 0.0], np.float32)
    sub = curves[:3000]
    rays = scenes.targeted_rays(sub, 8000, seed=5, dist=(50.0, 300.0))
    S = oracle.Scene()
    S.add(oracle.MODE_PHANTOM, sub, per_point=True)
    S.add(oracle.MODE_PHANTOM, sub, per_point=True, transform=xf1)
    S.add(oracle.MODE_PHANTOM, sub, per_point=True, transform=xf2)
    ref = S.trace(rays)
    b0 = G.build_blas(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, sub, per_point=True)
    b1 = G.build_blas(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, sub, per_point=True, transform=xf1)
    b2 = G.build_blas(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, sub, per_point=True, transform=xf2)
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, [b0, b1, b2])
    h = G.trace(gpu_ctx, tlas, rays)
    _compare_hits(h, ref, "instances")


@pytest.mark.parametrize("count", [3, 18, 64])
def test_render_many_instances(gpu_ctx, G, oracle, curves, count):
    """Renders of 3, 18 and 64 translated instances with shadows and the FP64 polish: the
    render kernels that stage the BLAS fields of up to 16 instances in LDS, and those that
    read them per lane beyond that, give the oracle's primitive ids (merge keys carry global
    primitive ids, so the instance count is not limited)."""
    import vkhrt_amd as V
    S = oracle.Scene()
    blases = []
    for i in range(count):
        # overlapping copies of the head, a few units apart
        xf = np.array([1, 0, 0, 3.0 * (i % 6) - 7.5, 0, 1, 0, 2.0 * (i // 6), 0, 0, 1, -1.5 * i], np.float32)
        S.add(oracle.MODE_PHANTOM, curves, per_point=True, transform=xf)
        blases.append(G.build_blas(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, curves, per_point=True, transform=xf))
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, blases)
    W, H = 320, 240
    cam = V.camera_look_at((0.0, 50.0, 400.0), (0.0, 30.0, 0.0), aspect=W / H)
    flags = V.RENDER_SHADOWS | V.RENDER_FP64_POLISH
    print(f"[render_many_instances] count={count}", flush=True)
    img, ids = G.render(gpu_ctx, tlas, cam, W, H, flags)
    vi, pi = G.cam_arrays(cam)
    rimg, rids, _, _ = S.render(vi, pi, W, H, oracle.F_POLISH, shadows=1, pixel_stride=3)
    mask = (np.arange(W * H).reshape(H, W) % 3) == 0
    assert (ids[mask] != 0xFFFFFFFF).sum() > 150
    assert np.array_equal(ids[mask], rids[mask])
    assert G.psnr(img[..., :3][mask], rimg[mask]) >= 50.0


def test_edge_cases(gpu_ctx, G, oracle, curves):
    """Empty BLAS, a single primitive, axis-aligned directions, tmin > tmax."""
    import vkhrt_amd as V
    empty = V.BottomLevelAccelerationStructure(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, None, 0)
    one = G.build_blas(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, curves[:1], per_point=True)
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, [empty, one])
    c = curves[0].reshape(4, 4)
    mid = (c[0, :3] + c[3, :3]) * 0.5
    rays = []
    for axis in range(3):
        for sgn in (1.0, -1.0):
            d = np.zeros(3, np.float32)
            d[axis] = sgn
            o = mid - d * 50.0
            rays.append(np.r_[o, 0.001, d, 10000.0])
    rays.append(np.r_[mid - np.array([0, 0, 50.0]), 20.0, np.array([0, 0, 1.0]), 10.0])  # empty interval
    rays = np.array(rays, np.float32)
    S = oracle.Scene()
    S.add(oracle.MODE_PHANTOM, np.zeros((0, 16), np.float32), per_point=True)
    S.add(oracle.MODE_PHANTOM, curves[:1], per_point=True)
    ref = S.trace(rays)
    h = G.trace(gpu_ctx, tlas, rays)
    _compare_hits(h, ref, "edge")


@pytest.mark.parametrize("per_point", [False, True])
def test_phantom_fp64_polish_trace(gpu_ctx, G, oracle, curves, per_point):
    """FP64 polish of grazing roots (deferred to a candidate pass on the GPU) equals the
    oracle's in-place polish bit for bit, grazing-heavy ray set included."""
    import scenes
    import vkhrt_amd as V
    rays = np.concatenate([scenes.targeted_rays(curves, 20000, seed=31 + per_point),
                           scenes.grazing_rays(curves, 20000, seed=41 + per_point, per_point=per_point)])
    S = oracle.Scene()
    S.add(oracle.MODE_PHANTOM, curves, per_point=per_point)
    ref = S.trace(rays, flags=oracle.F_POLISH)
    blas = G.build_blas(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, curves, per_point=per_point)
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, [blas])
    h = G.trace(gpu_ctx, tlas, rays, flags=V.TRACE_FP64_POLISH)
    _compare_hits(h, ref, f"polish per_point={per_point}")
    assert np.array_equal(h["v"].view(np.uint32), ref["v"].view(np.uint32))  # polished flag
    npol = int((ref["v"] == 1.0).sum())
    print("polished roots:", npol)
    assert npol > 50, "the grazing ray set must exercise the FP64 continuation"
    hs = G.trace(gpu_ctx, tlas, rays, flags=V.TRACE_FP64_POLISH | V.TRACE_TERMINATE_ON_FIRST_HIT)
    rs = S.trace(rays, flags=oracle.F_POLISH | oracle.F_ANY_HIT)
    assert np.array_equal(hs["prim"] == 0xFFFFFFFF, rs["prim"] == 0xFFFFFFFF)


def test_phantom_render_fp64_polish(gpu_ctx, G, oracle, curves):
    import vkhrt_amd as V
    W = H = 256
    cam = V.camera_look_at((0.0, 50.0, 400.0), (0.0, 30.0, 0.0), fov_deg=60.0, aspect=1.0)
    vi, pi = G.cam_arrays(cam)
    S = oracle.Scene()
    S.add(oracle.MODE_PHANTOM, curves, per_point=True)
    ref_img, ref_ids, _, _ = S.render(vi, pi, W, H, flags=oracle.F_POLISH, shadows=1)
    blas = G.build_blas(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, curves, per_point=True)
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, [blas])
    img, ids = G.render(gpu_ctx, tlas, cam, W, H, flags=V.RENDER_SHADOWS | V.RENDER_FP64_POLISH)
    assert np.array_equal(ids, ref_ids)
    assert np.array_equal(img[..., :3].view(np.uint32), ref_img.view(np.uint32))


@pytest.mark.parametrize("persistent,chunk", [(0, 0), (1, 0), (1, 64), (1, 4096)])
def test_persistent_and_plain_kernels_agree(gpu_ctx, G, oracle, curves, persistent, chunk):
    """One thread per ray, and persistent lanes fed by one work counter or by per-XCD counters
    in runs of 64 (313 runs, so waves steal from other XCDs' counters) or 4096 rays: every
    hit equals the oracle's."""
    import scenes
    import vkhrt_amd as V
    rays = scenes.targeted_rays(curves, 20000, seed=3)
    S = oracle.Scene()
    S.add(oracle.MODE_PHANTOM, curves, per_point=True)
    ref = S.trace(rays)
    blas = G.build_blas(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, curves, per_point=True)
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, [blas])
    gpu_ctx.set_option(V.OPTION_PERSISTENT, persistent)
    gpu_ctx.set_option(V.OPTION_REFILL_MIN, 8)
    gpu_ctx.set_option(V.OPTION_XCD_CHUNK, chunk)
    try:
        h = G.trace(gpu_ctx, tlas, rays)
    finally:
        gpu_ctx.set_option(V.OPTION_PERSISTENT, 1)
        gpu_ctx.set_option(V.OPTION_REFILL_MIN, 16)
        gpu_ctx.set_option(V.OPTION_XCD_CHUNK, 4096)
    _compare_hits(h, ref, f"persistent={persistent} chunk={chunk}")


@pytest.mark.parametrize("steps", [2, 3, 4, 8])
def test_steps_per_iteration_leave_results_unchanged(gpu_ctx, G, oracle, curves, steps):
    """Several traversal steps per pass over the persistent loop's bookkeeping: closest hits
    equal the oracle's bit for bit, any-hit decisions too, and the C0 render with shadows and
    FP64 polish equals the one-step render pixel for pixel."""
    import scenes
    import vkhrt_amd as V
    rays = scenes.targeted_rays(curves, 20000, seed=12)
    S = oracle.Scene()
    S.add(oracle.MODE_PHANTOM, curves, per_point=True)
    ref = S.trace(rays)
    rs = S.trace(rays, flags=oracle.F_ANY_HIT)
    blas = G.build_blas(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, curves, per_point=True)
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, [blas])
    cam = V.camera_look_at((0.0, 112.0, 140.0), (0.0, 105.0, 0.0), fov_deg=22.6, aspect=1.0)
    flags = V.RENDER_SHADOWS | V.RENDER_FP64_POLISH
    try:
        gpu_ctx.set_option(V.OPTION_STEPS_PER_ITERATION, 1)
        one, one_ids = G.render(gpu_ctx, tlas, cam, 256, 256, flags)
        gpu_ctx.set_option(V.OPTION_STEPS_PER_ITERATION, steps)
        h = G.trace(gpu_ctx, tlas, rays)
        hs = G.trace(gpu_ctx, tlas, rays, flags=V.TRACE_TERMINATE_ON_FIRST_HIT)
        img, ids = G.render(gpu_ctx, tlas, cam, 256, 256, flags)
    finally:
        gpu_ctx.set_option(V.OPTION_STEPS_PER_ITERATION, 3)
    _compare_hits(h, ref, f"steps per iteration {steps}")
    assert np.array_equal(hs["prim"] == 0xFFFFFFFF, rs["prim"] == 0xFFFFFFFF)
    assert np.array_equal(ids, one_ids) and np.array_equal(img.view(np.uint32), one.view(np.uint32))
    assert (ids != 0xFFFFFFFF).sum() > 1000
    for v in (0, 9):
        with pytest.raises(V.VkhrtError):
            gpu_ctx.set_option(V.OPTION_STEPS_PER_ITERATION, v)


@pytest.mark.parametrize("any_hit", [False, True])
def test_persistent_many_rays_per_lane(gpu_ctx, G, oracle, curves, any_hit):
    """1.5M rays (several per resident lane), led by 300k rays that miss everything so whole
    wavefronts draw all-miss batches: the persistent kernels (lane refill, compacted phantom
    intersection rounds) equal one thread per ray bit for bit, and the oracle on a sample."""
    import scenes
    import vkhrt_amd as V
    rng = np.random.default_rng(44)
    miss = np.zeros((300000, 8), np.float32)
    miss[:, 0:3] = rng.uniform(-50, 50, (300000, 3)) + np.array([0, 1000.0, 0], np.float32)
    miss[:, 3] = 0.001
    miss[:, 5] = 1.0  # pointing up, away from the hair
    miss[:, 7] = 10000.0
    rays = np.concatenate([miss, scenes.targeted_rays(curves, 1200000, seed=45)])
    blas = G.build_blas(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, curves, per_point=True)
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, [blas])
    flags = V.TRACE_TERMINATE_ON_FIRST_HIT if any_hit else 0
    print(f"[many_rays_per_lane] any_hit={any_hit} rays={rays.shape[0]}", flush=True)
    gpu_ctx.set_option(V.OPTION_PERSISTENT, 0)
    try:
        plain = G.trace(gpu_ctx, tlas, rays, flags)
    finally:
        gpu_ctx.set_option(V.OPTION_PERSISTENT, 1)
    pers = G.trace(gpu_ctx, tlas, rays, flags)
    hit = plain["prim"] != 0xFFFFFFFF
    assert hit[:300000].sum() == 0 and hit.mean() > 0.15
    if any_hit:
        # any accepted hit may be reported: compare the decision, and check the hit is real
        assert np.array_equal(pers["prim"] != 0xFFFFFFFF, hit)
    else:
        assert np.array_equal(pers.view(np.uint32), plain.view(np.uint32))
    sample = rng.choice(rays.shape[0], 20000, replace=False)
    S = oracle.Scene()
    S.add(oracle.MODE_PHANTOM, curves, per_point=True)
    ref = S.trace(rays[sample], oracle.F_ANY_HIT if any_hit else 0)
    if any_hit:
        assert np.array_equal(pers["prim"][sample] != 0xFFFFFFFF, ref["prim"] != 0xFFFFFFFF)
    else:
        _compare_hits(pers[sample], ref, "persistent many rays")


@pytest.mark.parametrize("polish,order,chunk", [(False, 0, 0), (True, 0, 0), (False, 1, 0), (True, 1, 0),
                                               (False, 2, 64), (True, 2, 4096), (True, 0, 1 << 20)])
def test_persistent_render_1024(gpu_ctx, G, oracle, curves, polish, order, chunk):
    """A 1000x1000 render (about 5 pixels per resident lane; 8x8 and 64x64 tiles cross the
    image edge) with shadows, row-major or tiled work order, rays from one counter or from
    per-XCD counters in runs of `chunk` (2^20: one run, so seven XCDs take all their rays from
    the eighth's counter): persistent and one-thread-per-ray kernels give the same image;
    primitive ids equal the oracle's."""
    import vkhrt_amd as V
    blas = G.build_blas(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, curves, per_point=True)
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, [blas])
    W = H = 1000
    cam = V.camera_look_at((0.0, 50.0, 400.0), (0.0, 30.0, 0.0), aspect=1.0)
    flags = V.RENDER_SHADOWS | (V.RENDER_FP64_POLISH if polish else 0)
    print(f"[persistent_render_1024] polish={polish} order={order} chunk={chunk}", flush=True)
    gpu_ctx.set_option(V.OPTION_RAY_ORDER, order)
    gpu_ctx.set_option(V.OPTION_XCD_CHUNK, chunk)
    try:
        # another view first: a pixel the frame under test failed to trace would keep its hit
        G.render(gpu_ctx, tlas, V.camera_look_at((60.0, 90.0, 300.0), (0.0, 30.0, 0.0), aspect=1.0), W, H, flags)
        img, ids = G.render(gpu_ctx, tlas, cam, W, H, flags)
    finally:
        gpu_ctx.set_option(V.OPTION_RAY_ORDER, 0)
        gpu_ctx.set_option(V.OPTION_XCD_CHUNK, 4096)
    if not polish:
        gpu_ctx.set_option(V.OPTION_PERSISTENT, 0)
        try:
            img0, ids0 = G.render(gpu_ctx, tlas, cam, W, H, flags)
        finally:
            gpu_ctx.set_option(V.OPTION_PERSISTENT, 1)
        assert np.array_equal(ids, ids0) and np.array_equal(img, img0)
    S = oracle.Scene()
    S.add(oracle.MODE_PHANTOM, curves, per_point=True)
    vi, pi = G.cam_arrays(cam)
    rimg, rids, _, _ = S.render(vi, pi, W, H, oracle.F_POLISH if polish else 0, shadows=1, pixel_stride=7)
    sel = rids != 0xFFFFFFFF
    assert (ids != 0xFFFFFFFF).sum() > 1500
    # the oracle traced every 7th pixel; the rest of its id image is the miss value
    stride_mask = (np.arange(W * H).reshape(H, W) % 7) == 0
    assert np.array_equal(ids[stride_mask], rids[stride_mask])
    assert G.psnr(img[..., :3][stride_mask], rimg[stride_mask]) >= 50.0
    assert sel.sum() > 0


def test_work_order_options_checked(gpu_ctx):
    """Ray order 0..2 and XCD chunks of 0 or a power of two in 64..2^20 are accepted; other
    values fail with INVALID_ARGUMENT."""
    import vkhrt_amd as V
    for v in (0, 1, 2):
        gpu_ctx.set_option(V.OPTION_RAY_ORDER, v)
    for v in (0, 64, 4096, 1 << 20):
        gpu_ctx.set_option(V.OPTION_XCD_CHUNK, v)
    for opt, v in ((V.OPTION_RAY_ORDER, 3), (V.OPTION_RAY_ORDER, -1), (V.OPTION_XCD_CHUNK, 32),
                   (V.OPTION_XCD_CHUNK, 100), (V.OPTION_XCD_CHUNK, 1 << 21), (V.OPTION_XCD_CHUNK, -64)):
        with pytest.raises(V.VkhrtError):
            gpu_ctx.set_option(opt, v)
    gpu_ctx.set_option(V.OPTION_RAY_ORDER, 0)
    gpu_ctx.set_option(V.OPTION_XCD_CHUNK, 4096)


@pytest.mark.gpu
def test_cpp_mirror_trace(tmp_path):
    """The C++ mirror (include/vkhrt.hpp) builds 3 translated instances from host pointers and
    traces through vkhrt_trace_rays; hits equal the oracle's bit for bit."""
    import subprocess
    import oracle_lib as O
    import scenes
    from test_host import cpp_test_binary
    curves = scenes.c0_curves(200, 16, seed=31)
    rays = scenes.targeted_rays(curves, 3000, seed=32)
    rays[1000:2000, 0] += 10.0  # aim a third of the rays at instance 1
    cpath, rpath, hpath = tmp_path / "c.bin", tmp_path / "r.bin", tmp_path / "h.bin"
    curves.astype(np.float32).tofile(cpath)
    rays.astype(np.float32).tofile(rpath)
    print("[cpp_mirror_trace] 3 instances x", curves.shape[0], "curves,", rays.shape[0], "rays", flush=True)
    r = subprocess.run([cpp_test_binary(), "trace", str(cpath), str(curves.shape[0]), str(rpath),
                        str(rays.shape[0]), str(hpath), "3"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(hpath, O.HIT_DTYPE)
    S = O.Scene()
    for i in range(3):
        xf = np.array([[1.0, 0, 0, 10.0 * i], [0, 1, 0, 0], [0, 0, 1, 0]], np.float32)
        S.add(O.MODE_PHANTOM, curves, per_point=True, transform=xf)
    ref = S.trace(rays)
    assert (ref["prim"] != 0xFFFFFFFF).mean() > 0.2
    assert np.array_equal(got["prim"], ref["prim"]) and np.array_equal(got["inst"], ref["inst"])
    assert np.array_equal(got["t"].view(np.uint32), ref["t"].view(np.uint32))


def _mesh_blas(ctx, G, verts, idx, transform=None, host=False):
    import vkhrt_amd as V
    if host:
        return V.BottomLevelAccelerationStructure(ctx, V.GEOMETRY_TRIANGLES, verts.ctypes.data, idx.shape[0],
                                                  flags=V.BLAS_HOST_POINTERS, transform=transform,
                                                  indices_ptr=idx.ctypes.data, vertex_count=verts.shape[0])
    dv, di = G.dev(verts), G.dev(idx)
    b = V.BottomLevelAccelerationStructure(ctx, V.GEOMETRY_TRIANGLES, dv.data_ptr(), idx.shape[0], transform=transform,
                                           indices_ptr=di.data_ptr(), vertex_count=verts.shape[0],
                                           stream=G.stream_ptr())
    torch.cuda.synchronize()
    return b


@pytest.mark.parametrize("host", [False, True])
def test_mesh_trace_matches_oracle(gpu_ctx, G, oracle, curves, host):
    """Triangle meshes (BLASType::eMesh): a mesh-only TLAS (the persistent mesh kernels) and a
    mesh + hair TLAS (mixed kernels) trace bit for bit like the oracle, interpolated normals
    included."""
    import scenes
    import vkhrt_amd as V
    verts, idx = scenes.uv_sphere(30.0, 32, 64, (0.0, 40.0, 0.0))
    xf = np.array([[0.0, -1, 0, 5], [1, 0, 0, 10], [0, 0, 1, -20]], np.float32)
    rng = np.random.default_rng(6)
    o = rng.uniform(-150, 150, (20000, 3))
    d = rng.uniform(-40, 40, (20000, 3)) + np.array([0, 40, 0]) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, np.full((20000, 1), 0.001), d, np.full((20000, 1), 1e4)], 1).astype(np.float32)
    print(f"[mesh trace] host pointers {host}", flush=True)
    m0 = _mesh_blas(gpu_ctx, G, verts, idx, host=host)
    m1 = _mesh_blas(gpu_ctx, G, verts, idx, transform=xf.reshape(12), host=host)
    tl = V.TopLevelAccelerationStructure(gpu_ctx, [m0, m1])
    S = oracle.Scene()
    S.add_mesh(verts, idx)
    S.add_mesh(verts, idx, transform=xf)
    _compare_hits(G.trace(gpu_ctx, tl, rays), S.trace(rays), "mesh")
    g_any = G.trace(gpu_ctx, tl, rays, V.TRACE_TERMINATE_ON_FIRST_HIT)
    r_any = S.trace(rays, flags=oracle.F_ANY_HIT)
    assert np.array_equal(g_any["prim"] != 0xFFFFFFFF, r_any["prim"] != 0xFFFFFFFF)
    hair = G.build_blas(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, curves, per_point=True)
    mixed = V.TopLevelAccelerationStructure(gpu_ctx, [hair, m1])
    S2 = oracle.Scene()
    S2.add(oracle.MODE_PHANTOM, curves, per_point=True)
    S2.add_mesh(verts, idx, transform=xf)
    _compare_hits(G.trace(gpu_ctx, mixed, rays), S2.trace(rays), "mesh + hair")


def test_mesh_render_and_batch_match_oracle(gpu_ctx, G, oracle):
    import scenes
    import vkhrt_amd as V
    verts, idx = scenes.uv_sphere(8.0, 24, 48, (0.0, 8.0, 0.0))
    blas = _mesh_blas(gpu_ctx, G, verts, idx)
    tl = V.TopLevelAccelerationStructure(gpu_ctx, [blas])
    S = oracle.Scene()
    S.add_mesh(verts, idx)
    W, H = 96, 64
    vi, pi = oracle.camera((0.0, 20.0, 45.0), (0.0, 8.0, 0.0), aspect=W / H)
    cam = V.Camera()
    for k in range(16):
        cam.view_inverse[k], cam.proj_inverse[k] = float(vi[k]), float(pi[k])
    img, ids = G.render(gpu_ctx, tl, cam, W, H)
    ref, rids, _, _ = S.render(vi, pi, W, H, shadows=1)
    assert (rids != 0xFFFFFFFF).sum() > 250
    assert np.array_equal(ids, rids) and np.array_equal(img[..., :3].view(np.uint32), ref.view(np.uint32))
    out = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
    gpu_ctx.render_batch(tl, [(cam, W, H, out.data_ptr())], bounces=2, seed=3, environment=(0.2, 0.3, 0.4),
                         stream=G.stream_ptr())
    torch.cuda.synchronize()
    ri, _, _, _, _ = S.render_batch([(vi, pi, W, H)], bounces=2, seed=3, env=(0.2, 0.3, 0.4))
    got = out.cpu().numpy().reshape(H, W, 4)[..., :3]
    assert np.array_equal(got.view(np.uint32), ri[0].view(np.uint32))


def test_mesh_rejects_out_of_range_indices(gpu_ctx, G):
    import vkhrt_amd as V
    verts = np.zeros((3, 8), np.float32)
    verts[1, 0] = verts[2, 1] = 1.0
    idx = np.array([[0, 1, 5]], np.uint32)
    with pytest.raises(V.VkhrtError) as e:
        _mesh_blas(gpu_ctx, G, verts, idx)
    assert e.value.code == -1


@pytest.mark.parametrize("mixed", [False, True])
def test_environment_map_render(gpu_ctx, G, oracle, mixed):
    """miss.rmiss with an environment map: hit pixels bit for bit, missed pixels within 1e-4 of the
    oracle, median 1e-6 (asin / atan2 / exp / pow come from different math libraries); the
    persistent render and the mixed-TLAS render."""
    import scenes
    import vkhrt_amd as V
    curves = scenes.dense_ball_curves(nstrands=1500)
    rng = np.random.default_rng(12)
    ew, eh = 96, 48
    xx, yy = np.meshgrid(np.linspace(0, 6.28, ew), np.linspace(0, 3.14, eh))
    env = np.stack([1 + np.sin(xx), 0.5 + 0.5 * np.cos(2 * yy), rng.uniform(0, 3, (eh, ew)), np.ones((eh, ew))],
                   -1).astype(np.float32)
    denv = G.dev(env)
    S = oracle.Scene()
    S.add(oracle.MODE_PHANTOM, curves, per_point=True)
    blases = [G.build_blas(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, curves, per_point=True)]
    if mixed:
        verts, idx = scenes.uv_sphere(4.0, 12, 24, (12.0, 4.0, 0.0))
        S.add_mesh(verts, idx)
        blases.append(_mesh_blas(gpu_ctx, G, verts, idx))
    tl = V.TopLevelAccelerationStructure(gpu_ctx, blases)
    W, H = 128, 96
    vi, pi = oracle.camera((0.0, 20.0, 45.0), (0.0, 8.0, 0.0), aspect=W / H)
    cam = V.Camera()
    for k in range(16):
        cam.view_inverse[k], cam.proj_inverse[k] = float(vi[k]), float(pi[k])
    em = V.EnvironmentMap(denv.data_ptr(), ew, eh, 1.0)
    img = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
    ids = torch.zeros(W * H, dtype=torch.int32, device="cuda")
    print(f"[environment map] mixed {mixed}", flush=True)
    gpu_ctx.render(tl, cam, W, H, img.data_ptr(), prim_ids_ptr=ids.data_ptr(), stream=G.stream_ptr(), environment=em)
    torch.cuda.synchronize()
    got = img.cpu().numpy().reshape(H, W, 4)[..., :3]
    gid = ids.cpu().numpy().view(np.uint32).reshape(H, W)
    ref, rid, _, _ = S.render(vi, pi, W, H, shadows=1, env=env)
    hit = rid != 0xFFFFFFFF
    assert np.array_equal(gid, rid) and 200 < hit.sum() < W * H - 200
    assert np.array_equal(got[hit].view(np.uint32), ref[hit].view(np.uint32))
    # a 1-ulp difference of asin / atan2 moves the lookup by ~1e-7 rad; the map's steepest channel
    # (uniform noise per texel) changes by up to ~3 per texel, i.e. ~5e-5 per 1e-7 rad at 96 texels
    # per turn, hence the bound on the maximum; the median difference is at rounding level
    diff = np.abs(got[~hit] - ref[~hit])
    assert diff.max() < 1e-4 and np.median(diff) < 1e-6
    assert np.ptp(ref[~hit][:, 2]) > 0.1  # the map varies across the missed pixels


@pytest.mark.parametrize("mode", ["phantom", "lss", "dots"])
def test_count_candidates_matches_oracle(gpu_ctx, G, oracle, curves, mode):
    """vkhrt_count_candidates (the N*_prim of the roofline's algorithmic bytes) equals the
    oracle's count ray for ray: camera rays of the C0 frame against their closest hits, and
    the shadow rays those hits spawn against their any-hit results."""
    import vkhrt_amd as V
    W = H = 256
    cam = V.camera_look_at((0.0, 112.0, 140.0), (0.0, 105.0, 0.0), fov_deg=22.62, aspect=1.0)
    vi, pi = G.cam_arrays(cam)
    S = oracle.Scene()
    if mode == "phantom":
        prims, gmode, kw = curves, V.GEOMETRY_PHANTOM_CURVES, {"per_point": True}
        S.add(oracle.MODE_PHANTOM, curves, per_point=True)
    else:
        lines = oracle.curves_to_lines(curves, 4)
        if mode == "lss":
            prims, gmode = oracle.generate_lss(lines), V.GEOMETRY_LSS
            S.add(oracle.MODE_LSS, prims)
        else:
            prims, gmode = lines, V.GEOMETRY_DOTS
            S.add(oracle.MODE_DOTS, lines, per_point=True)
        kw = {"per_point": mode == "dots"}
    blas = G.build_blas(gpu_ctx, gmode, prims, **kw)
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, [blas])
    rays = oracle.raygen(vi, pi, W, H)
    ref = S.trace(rays)
    h = G.trace(gpu_ctx, tlas, rays)
    _compare_hits(h, ref, f"{mode} C0 camera rays")
    hit = ref["prim"] != 0xFFFFFFFF
    srays = np.zeros((int(hit.sum()), 8), np.float32)  # shadow rays toward (0, 1, 0)
    srays[:, :3] = rays[hit, :3] + rays[hit, 4:7] * ref["t"][hit, None]
    srays[:, 3], srays[:, 5], srays[:, 7] = 0.001, 1.0, 10000.0
    sref = S.trace(srays, flags=oracle.F_ANY_HIT)
    sh = G.trace(gpu_ctx, tlas, srays, flags=V.TRACE_TERMINATE_ON_FIRST_HIT)
    assert np.array_equal(sh["prim"] == 0xFFFFFFFF, sref["prim"] == 0xFFFFFFFF)
    # the count runs over [tmin, t_closest]: an any-hit t depends on the traversal order, so the
    # shadow rays are counted against their closest hits (equal on both sides)
    sref = S.trace(srays)
    sh = G.trace(gpu_ctx, tlas, srays)
    _compare_hits(sh, sref, f"{mode} shadow rays, closest hit")
    for what, r, hh, rr in (("primary", rays, h, ref), ("shadow", srays, sh, sref)):
        d_r, d_h = G.dev(r), G.dev(hh.view(np.float32).reshape(-1))
        counts = torch.zeros(r.shape[0], dtype=torch.int32, device="cuda")
        gpu_ctx.count_candidates(tlas, d_r.data_ptr(), d_h.data_ptr(), r.shape[0], counts.data_ptr(),
                                 stream=G.stream_ptr())
        torch.cuda.synchronize()
        g = counts.cpu().numpy().view(np.uint32)
        o = S.count_candidates(r, rr)
        print(f"{mode} {what}: {r.shape[0]} rays, N* mean GPU {g.mean():.4f} oracle {o.mean():.4f}")
        assert g.sum() > 0
        assert np.array_equal(g, o), f"{mode} {what}: counts differ on {(g != o).sum()} rays"


def test_polish_trace_64_instances(gpu_ctx, G, oracle, curves):
    """The C0 hair split into 64 BLASes (one instance each) traced with grazing rays and the
    FP64 polish: the deferred candidates of every instance merge by global primitive id and
    equal the oracle's in-place polish bit for bit, polished roots included."""
    import scenes
    import vkhrt_amd as V
    parts = np.array_split(np.arange(curves.shape[0]), 64)
    S = oracle.Scene()
    blases = []
    for idx in parts:
        sub = np.ascontiguousarray(curves[idx])
        S.add(oracle.MODE_PHANTOM, sub, per_point=True)
        blases.append(G.build_blas(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, sub, per_point=True))
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, blases)
    rays = np.concatenate([scenes.grazing_rays(curves, 6000, seed=41), scenes.targeted_rays(curves, 4000, seed=42)])
    ref = S.trace(rays, flags=oracle.F_POLISH)
    h = G.trace(gpu_ctx, tlas, rays, flags=V.TRACE_FP64_POLISH)
    _compare_hits(h, ref, "64-instance polish")
    assert np.array_equal(h["v"].view(np.uint32), ref["v"].view(np.uint32))  # polished flag
    npol = int((ref["v"] != 0).sum())
    print(f"64 instances: {len(np.unique(ref['inst'][ref['prim'] != 0xFFFFFFFF]))} hit, {npol} polished roots")
    assert npol >= 50


@pytest.mark.parametrize("extra", ["polish", "polish+area+accurate"])
def test_mixed_tlas_render_polish(gpu_ctx, G, oracle, curves, extra):
    """A TLAS mixing phantom hair (per-point radius), LSS capsules, DOTS quads and a triangle
    mesh, rendered with the FP64 polish (the deferred pipeline with one thread per queued
    ray): bit for bit against the oracle, with an area light and accurate hits too."""
    import scenes
    import vkhrt_amd as V
    xf = np.array([1, 0, 0, 4.0, 0, 1, 0, -2.0, 0, 0, 1, 3.0], np.float32)
    lines = scenes.lines_from_curves(curves[4000:6000], 4)
    caps = oracle.generate_lss(lines)
    verts, tri = scenes.uv_sphere(12.0, 16, 32, (0.0, 96.0, 20.0))
    S = oracle.Scene()
    S.add(oracle.MODE_PHANTOM, curves[:4000], per_point=True)
    S.add(oracle.MODE_LSS, caps, transform=xf)
    S.add(oracle.MODE_DOTS, scenes.lines_from_curves(curves[6000:8000], 4), per_point=True)
    S.add_mesh(verts, tri)
    blases = [G.build_blas(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, curves[:4000], per_point=True),
              G.build_blas(gpu_ctx, V.GEOMETRY_LSS, caps, transform=xf),
              G.build_blas(gpu_ctx, V.GEOMETRY_DOTS, scenes.lines_from_curves(curves[6000:8000], 4), per_point=True)]
    dv, di = G.dev(verts), G.dev(tri)
    blases.append(V.BottomLevelAccelerationStructure(gpu_ctx, V.GEOMETRY_TRIANGLES, dv.data_ptr(), tri.shape[0],
                                                     indices_ptr=di.data_ptr(), vertex_count=verts.shape[0],
                                                     stream=G.stream_ptr()))
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, blases)
    W, H = 256, 256
    cam = V.camera_look_at((0.0, 112.0, 140.0), (0.0, 100.0, 0.0), fov_deg=30.0, aspect=1.0)
    vi, pi = G.cam_arrays(cam)
    flags, oflags, kw, okw = V.RENDER_SHADOWS | V.RENDER_FP64_POLISH, oracle.F_POLISH, {}, {}
    if extra != "polish":
        light = np.array([[0.0, 260.0, 0.0], [60.0, 0.0, 0.0], [0.0, 0.0, 60.0]], np.float32)
        kw["area_light"] = V.area_light(light[0], light[1], light[2], samples=2, seed=3)
        okw = {"area_light": light, "light_samples": 2, "light_seed": 3}
        flags |= V.RENDER_ACCURATE_NORMALS
        oflags |= oracle.F_ACCURATE
    img = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
    ids = torch.zeros(W * H, dtype=torch.int32, device="cuda")
    print(f"[mixed render] {extra}", flush=True)
    gpu_ctx.render(tlas, cam, W, H, img.data_ptr(), flags=flags, prim_ids_ptr=ids.data_ptr(), stream=G.stream_ptr(),
                   **kw)
    torch.cuda.synchronize()
    gpu_ctx.synchronize(G.stream_ptr())
    got = img.cpu().numpy().reshape(H, W, 4)[..., :3]
    gids = ids.cpu().numpy().view(np.uint32).reshape(H, W)
    ref, rids, _, _ = S.render(vi, pi, W, H, flags=oflags, shadows=1, **okw)
    assert (rids != 0xFFFFFFFF).sum() > 1000
    assert np.array_equal(gids, rids)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
