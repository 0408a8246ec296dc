"""GPU parity of batched path tracing (vkhrt_render_batch) against the oracle's restatement
(orc_render_batch): per-view float images and primary primitive ids bit for bit, and the
bounce / shadow ray counts. The path rule uses only correctly rounded operations on both
sides, so the tolerance is 0."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

ENV = (0.2, 0.3, 0.4)
TINTS = [(1.0, 1.0, 1.0), (0.5, 0.75, 1.25)]
XF = np.array([[1, 0, 0, 14.0], [0, 1, 0, 3.0], [0, 0, 1, -8.0]], np.float32)


def _vcam(vi, pi):
    import vkhrt_amd as V
    c = V.Camera()
    for k in range(16):
        c.view_inverse[k] = float(vi[k])
        c.proj_inverse[k] = float(pi[k])
    return c


def _views(oracle, specs):
    return [(*oracle.camera(eye, tgt, tan_half_fovy=th, aspect=w / h), w, h) for eye, tgt, th, (w, h) in specs]


@pytest.fixture(scope="module")
def scene(gpu_ctx, oracle):
    """Two phantom instances (the second transformed, material 1) on both sides."""
    import gpu_util as G
    import scenes
    import vkhrt_amd as V
    curves = scenes.dense_ball_curves()
    S = oracle.Scene()
    S.add(oracle.MODE_PHANTOM, curves, per_point=True)
    S.add(oracle.MODE_PHANTOM, curves[::3], per_point=True, transform=XF)
    S.set_material(1, 1)
    keep = [G.dev(curves), G.dev(np.ascontiguousarray(curves[::3]))]
    b0 = V.BottomLevelAccelerationStructure(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, keep[0].data_ptr(), curves.shape[0],
                                            flags=V.BLAS_PER_POINT_RADIUS, material_index=0, stream=G.stream_ptr())
    b1 = V.BottomLevelAccelerationStructure(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, keep[1].data_ptr(),
                                            curves[::3].shape[0], flags=V.BLAS_PER_POINT_RADIUS,
                                            transform=XF.reshape(12), material_index=1, stream=G.stream_ptr())
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, [b0, b1], stream=G.stream_ptr())
    torch.cuda.synchronize()
    yield S, tlas
    tlas.close()
    b0.close()
    b1.close()


def _gpu_batch(ctx, tlas, views, bounces, seed, flags=0, env=ENV, tints=TINTS, rgba8=False):
    import gpu_util as G
    import vkhrt_amd as V
    imgs, ids, vlist = [], [], []
    for vi, pi, w, h in views:
        img = torch.zeros(w * h * (1 if rgba8 else 4), dtype=torch.int32 if rgba8 else torch.float32, device="cuda")
        pid = torch.zeros(w * h, dtype=torch.int32, device="cuda")
        imgs.append(img)
        ids.append(pid)
        vlist.append((_vcam(vi, pi), w, h, img.data_ptr(), pid.data_ptr()))
    ctx.render_batch(tlas, vlist, bounces=bounces, seed=seed, environment=env,
                     flags=flags | (V.BATCH_RGBA8 if rgba8 else 0), material_albedo=tints, stream=G.stream_ptr())
    torch.cuda.synchronize()
    ctx.synchronize(G.stream_ptr())
    out_i, out_p = [], []
    for (vi, pi, w, h), img, pid in zip(views, imgs, ids):
        a = img.cpu().numpy()
        out_i.append(a.view(np.uint32).reshape(h, w) if rgba8 else a.reshape(h, w, 4))
        out_p.append(pid.cpu().numpy().view(np.uint32).reshape(h, w))
    return out_i, out_p, ctx.batch_stats()


def _assert_same(gi, gp, ri, rp, what):
    for v, (a, p, r, q) in enumerate(zip(gi, gp, ri, rp)):
        assert np.array_equal(p, q), f"{what}: view {v} primary ids differ on {(p != q).sum()} pixels"
        diff = (a[..., :3].view(np.uint32) != r.view(np.uint32)).any(-1)
        assert not diff.any(), f"{what}: view {v} radiance differs on {diff.sum()} pixels"
        assert np.all(a[..., 3] == 1.0)


@pytest.mark.parametrize("polish,sort,chunk", [(0, 0, 4096), (0, 1, 4096), (1, 0, 4096), (1, 1, 4096), (1, 2, 4096),
                                              (0, 3, 4096), (1, 0, 0), (0, 3, 64)])
def test_batch_matches_oracle(gpu_ctx, oracle, scene, polish, sort, chunk):
    """sort: bit 0 material-sorted shading, bit 1 ray-sorted bounce queues; chunk: the rays
    of each traversal launch from one work counter (0) or from per-XCD counters in runs of
    `chunk`."""
    import scenes
    import vkhrt_amd as V
    S, tlas = scene
    views = _views(oracle, scenes.DENSE_VIEWS)
    flags = ((V.BATCH_FP64_POLISH if polish else 0) | (V.BATCH_SORT_MATERIALS if sort & 1 else 0) |
             (V.BATCH_SORT_RAYS if sort & 2 else 0))
    print(f"[batch parity] polish {polish} sort {sort} chunk {chunk}: 3 views, 3 bounces", flush=True)
    gpu_ctx.set_option(V.OPTION_XCD_CHUNK, chunk)
    try:
        gi, gp, (paths, nb, ns, _) = _gpu_batch(gpu_ctx, tlas, views, 3, 21, flags)
    finally:
        gpu_ctx.set_option(V.OPTION_XCD_CHUNK, 4096)
    ri, rp, rnb, rns, _ = S.render_batch(views, bounces=3, seed=21, env=ENV, albedo=TINTS,
                                         flags=oracle.F_POLISH if polish else 0)
    print(f"paths {paths}: bounce rays {nb} (oracle {rnb}), shadow rays {ns} (oracle {rns})")
    assert paths == sum(w * h for _, _, w, h in views)
    assert (nb, ns) == (rnb, rns)
    assert nb > 500 and ns > 500
    _assert_same(gi, gp, ri, rp, f"polish {polish} sort {sort}")


def test_batch_graph_replays_match_eager(gpu_ctx, oracle, scene):
    """The frame captured as a graph and replayed with new seeds and cameras (and eager work
    between the replays) reproduces the eager frames bit for bit; a new bounce count recaptures."""
    import gpu_util as G
    import scenes
    import vkhrt_amd as V
    S, tlas = scene
    views = _views(oracle, scenes.DENSE_VIEWS)
    moved = _views(oracle, [((25.0, 5.0, 40.0), (0.0, 8.0, 0.0), 0.5, (48, 40))] + scenes.DENSE_VIEWS[1:])
    eager = {s: _gpu_batch(gpu_ctx, tlas, views, 3, s) for s in (1, 2)}
    eager_moved = _gpu_batch(gpu_ctx, tlas, moved, 3, 1)
    _, _, (_, _, _, cap0) = eager[1]
    print("[batch graph] capture + replays", flush=True)
    for step, (vs, seed, ref) in enumerate([(views, 1, eager[1]), (views, 2, eager[2]), (moved, 1, eager_moved),
                                            (views, 1, eager[1])]):
        gi, gp, (_, nb, ns, caps) = _gpu_batch(gpu_ctx, tlas, vs, 3, seed, V.BATCH_GRAPH)
        _assert_same(gi, gp, [x[..., :3] for x in ref[0]], ref[1], f"graph replay {step}")
        assert (nb, ns) == ref[2][1:3]
        assert caps == cap0 + 1, "one capture serves every replay"
        # eager work of the program between replays
        G.render(gpu_ctx, tlas, _vcam(*views[0][:2]), 32, 32)
        torch.ones(1 << 20, device="cuda").sum().item()
    gi, gp, (_, _, _, caps) = _gpu_batch(gpu_ctx, tlas, views, 2, 1, V.BATCH_GRAPH)
    assert caps == cap0 + 2
    ri, rp, _, _, _ = S.render_batch(views, bounces=2, seed=1, env=ENV, albedo=TINTS)
    _assert_same(gi, gp, ri, rp, "recaptured graph")


def test_batch_rgba8(gpu_ctx, oracle, scene):
    import scenes
    S, tlas = scene
    views = _views(oracle, scenes.DENSE_VIEWS)
    gi, _, _ = _gpu_batch(gpu_ctx, tlas, views, 2, 4, rgba8=True)
    fi, _, _ = _gpu_batch(gpu_ctx, tlas, views, 2, 4)
    for a, f in zip(gi, fi):
        q = (np.minimum(np.maximum(f[..., :3], 0.0), 1.0) * np.float32(255.0) + np.float32(0.5)).astype(np.uint32)
        assert np.array_equal(a & 0xFF, q[..., 0]) and np.array_equal((a >> 8) & 0xFF, q[..., 1])
        assert np.array_equal((a >> 16) & 0xFF, q[..., 2]) and np.all(a >> 24 == 255)


@pytest.mark.parametrize("mode", ["lss", "dots"])
def test_batch_lss_dots(gpu_ctx, oracle, mode):
    import gpu_util as G
    import scenes
    import vkhrt_amd as V
    curves = scenes.dense_ball_curves(nstrands=600)
    lines = scenes.lines_from_curves(curves, 4)
    S = oracle.Scene()
    if mode == "lss":
        prims = oracle.generate_lss(lines)
        S.add(oracle.MODE_LSS, prims)
        blas = G.build_blas(gpu_ctx, V.GEOMETRY_LSS, prims)
    else:
        S.add(oracle.MODE_DOTS, lines, per_point=True)
        blas = G.build_blas(gpu_ctx, V.GEOMETRY_DOTS, lines, per_point=True)
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, [blas])
    views = _views(oracle, scenes.DENSE_VIEWS)
    print(f"[batch {mode}] 3 views, 2 bounces", flush=True)
    gi, gp, (_, nb, ns, _) = _gpu_batch(gpu_ctx, tlas, views, 2, 8, tints=None)
    ri, rp, rnb, rns, _ = S.render_batch(views, bounces=2, seed=8, env=ENV)
    assert (nb, ns) == (rnb, rns) and nb > 100
    _assert_same(gi, gp, ri, rp, mode)


def test_batch_many_views_sampled(gpu_ctx, oracle):
    """A larger batch (16 views of uneven sizes over two hair models of uneven size, 4 bounces):
    unsorted, material-sorted, and ray-sorted + graph frames agree bitwise on every pixel; the
    oracle checks every 13th path."""
    import gpu_util as G
    import scenes
    import vkhrt_amd as V
    big = scenes.dense_ball_curves(nstrands=6000, nseg=16, seed=5, sphere_radius=20.0)
    small = scenes.dense_ball_curves(nstrands=800, nseg=16, seed=6)
    xf = np.array([[1, 0, 0, 60.0], [0, 1, 0, 0.0], [0, 0, 1, 0.0]], np.float32)
    S = oracle.Scene()
    S.add(oracle.MODE_PHANTOM, big, per_point=True)
    S.add(oracle.MODE_PHANTOM, small, per_point=True, transform=xf)
    S.set_material(1, 1)
    keep = [G.dev(big), G.dev(small)]
    b0 = V.BottomLevelAccelerationStructure(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, keep[0].data_ptr(), big.shape[0],
                                            flags=V.BLAS_PER_POINT_RADIUS, stream=G.stream_ptr())
    b1 = V.BottomLevelAccelerationStructure(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, keep[1].data_ptr(), small.shape[0],
                                            flags=V.BLAS_PER_POINT_RADIUS, transform=xf.reshape(12),
                                            material_index=1, stream=G.stream_ptr())
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, [b0, b1], stream=G.stream_ptr())
    sizes = [(128, 128), (160, 90), (96, 96), (200, 120)]
    specs = []
    for k in range(16):
        a = 2 * np.pi * k / 16
        tgt = (0.0, 10.0, 0.0) if k % 2 == 0 else (60.0, 8.0, 0.0)
        dist = 70.0 if k % 2 == 0 else 40.0
        eye = (tgt[0] + dist * np.sin(a), 25.0 + 10 * np.cos(3 * a), dist * np.cos(a))
        specs.append((eye, tgt, 0.5, sizes[k % 4]))
    views = _views(oracle, specs)
    print("[batch many views] 16 views, 4 bounces", flush=True)
    a, ap, (paths, nb, ns, _) = _gpu_batch(gpu_ctx, tlas, views, 4, 77, V.BATCH_FP64_POLISH)
    b, bp, stb = _gpu_batch(gpu_ctx, tlas, views, 4, 77, V.BATCH_FP64_POLISH | V.BATCH_SORT_MATERIALS)
    c, cp, stc = _gpu_batch(gpu_ctx, tlas, views, 4, 77,
                            V.BATCH_FP64_POLISH | V.BATCH_SORT_MATERIALS | V.BATCH_SORT_RAYS | V.BATCH_GRAPH)
    assert stb[1:3] == stc[1:3] == (nb, ns)
    for x, y, z, px, py, pz in zip(a, b, c, ap, bp, cp):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)) and np.array_equal(x.view(np.uint32), z.view(np.uint32))
        assert np.array_equal(px, py) and np.array_equal(px, pz)
    hits = sum(int((p != 0xFFFFFFFF).sum()) for p in ap)
    print(f"paths {paths}, primary hits {hits}, bounce rays {nb}, shadow rays {ns}")
    assert hits > paths // 10 and nb > hits
    ri, rp, _, _, _ = S.render_batch(views, bounces=4, seed=77, env=ENV, albedo=TINTS, flags=oracle.F_POLISH,
                                     path_stride=13)
    flat_g = np.concatenate([x[..., :3].reshape(-1, 3) for x in a])[::13]
    flat_r = np.concatenate([x.reshape(-1, 3) for x in ri])[::13]
    flat_gp = np.concatenate([p.reshape(-1) for p in ap])[::13]
    flat_rp = np.concatenate([p.reshape(-1) for p in rp])[::13]
    assert np.array_equal(flat_gp, flat_rp)
    assert np.array_equal(flat_g.view(np.uint32), flat_r.view(np.uint32))
    tlas.close()
    b0.close()
    b1.close()


def test_batch_errors(gpu_ctx, oracle, scene):
    import gpu_util as G
    import scenes
    import vkhrt_amd as V
    S, tlas = scene
    views = _views(oracle, scenes.DENSE_VIEWS[:1])
    with pytest.raises(V.VkhrtError) as e:
        _gpu_batch(gpu_ctx, tlas, views, V.BATCH_MAX_BOUNCES + 1, 1)
    assert e.value.code == -1
    with pytest.raises(V.VkhrtError) as e:
        gpu_ctx.render_batch(tlas, [], bounces=1)
    assert e.value.code == -1
    with pytest.raises(V.VkhrtError) as e:
        gpu_ctx.render_batch(tlas, [(_vcam(*views[0][:2]), 0, 4, 1)], bounces=1)
    assert e.value.code == -1
    curves = scenes.dense_ball_curves(nstrands=50)
    caps = oracle.generate_lss(scenes.lines_from_curves(curves, 2))
    b_ph = G.build_blas(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, curves, per_point=True)
    b_lss = G.build_blas(gpu_ctx, V.GEOMETRY_LSS, caps)
    mixed = V.TopLevelAccelerationStructure(gpu_ctx, [b_ph, b_lss])
    _gpu_batch(gpu_ctx, mixed, views, 1, 1, V.BATCH_FP64_POLISH)  # polish of a mixed TLAS is served


@pytest.mark.parametrize("flags", [0, 3, 4], ids=["plain", "sorted+graph", "polish"])
def test_mixed_tlas_batch_matches_oracle(gpu_ctx, oracle, flags):
    """Phantom hair, LSS capsules and a triangle mesh in one TLAS (the per-instance dispatch,
    one lane per path) against the oracle, with material tints per instance; with the FP64
    polish of the phantom instance's grazing roots too."""
    import gpu_util as G
    import scenes
    import vkhrt_amd as V
    curves = scenes.dense_ball_curves(nstrands=1500, nseg=8)
    caps = oracle.generate_lss(scenes.lines_from_curves(curves[::2], 2))
    verts, tri = scenes.uv_sphere(4.0, 16, 32, (0.0, 22.0, 0.0))
    S = oracle.Scene()
    S.add(oracle.MODE_PHANTOM, curves, per_point=True)
    S.add(oracle.MODE_LSS, caps, transform=XF)
    S.add_mesh(verts, tri)
    S.set_material(1, 1)
    S.set_material(2, 1)
    b_ph = G.build_blas(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, curves, per_point=True)
    d_caps = G.dev(caps)
    b_lss = V.BottomLevelAccelerationStructure(gpu_ctx, V.GEOMETRY_LSS, d_caps.data_ptr(), caps.shape[0],
                                               transform=XF.reshape(12), material_index=1, stream=G.stream_ptr())
    dv, di = G.dev(verts), G.dev(tri)
    b_tri = V.BottomLevelAccelerationStructure(gpu_ctx, V.GEOMETRY_TRIANGLES, dv.data_ptr(), tri.shape[0],
                                               indices_ptr=di.data_ptr(), vertex_count=verts.shape[0],
                                               material_index=1, stream=G.stream_ptr())
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, [b_ph, b_lss, b_tri], stream=G.stream_ptr())
    torch.cuda.synchronize()
    views = _views(oracle, scenes.DENSE_VIEWS)
    gflags = (V.BATCH_SORT_MATERIALS | V.BATCH_SORT_RAYS | V.BATCH_GRAPH) if flags & 3 else 0
    gflags |= V.BATCH_FP64_POLISH if flags & 4 else 0
    print(f"[mixed batch] flags {gflags}: 3 views, 3 bounces", flush=True)
    for seed in (4, 5):
        gi, gp, (_, nb, ns, _) = _gpu_batch(gpu_ctx, tlas, views, 3, seed, gflags)
        ri, rp, rnb, rns, _ = S.render_batch(views, bounces=3, seed=seed, env=ENV, albedo=TINTS,
                                             flags=oracle.F_POLISH if flags & 4 else 0)
        assert (nb, ns) == (rnb, rns) and nb > 500 and ns > 500
        _assert_same(gi, gp, ri, rp, f"mixed batch seed {seed}")
    tlas.close()


def test_cpp_mirror_batch(tmp_path, oracle):
    """The C++ mirror (TopLevelAccelerationStructure::RenderBatch, graph mode, FP64 polish) renders
    the same images as the oracle, bit for bit."""
    import subprocess
    import scenes
    from test_host import cpp_test_binary
    curves = scenes.dense_ball_curves(nstrands=900, nseg=8)
    views = _views(oracle, scenes.DENSE_VIEWS[:2])
    cams = np.concatenate([np.concatenate([vi, pi]) for vi, pi, _, _ in views]).astype(np.float32)
    cpath, vpath, opath = tmp_path / "c.bin", tmp_path / "v.bin", tmp_path / "o.bin"
    curves.astype(np.float32).tofile(cpath)
    cams.tofile(vpath)
    sizes = ",".join(f"{w}x{h}" for _, _, w, h in views)
    print("[cpp_mirror_batch] 2 views, 3 bounces", flush=True)
    r = subprocess.run([cpp_test_binary(), "batch", str(cpath), str(curves.shape[0]), str(vpath), "2", sizes, "3",
                        "41", str(opath)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(opath, np.float32).reshape(-1, 4)
    S = oracle.Scene()
    S.add(oracle.MODE_PHANTOM, curves, per_point=True)
    ri, _, nb, ns, _ = S.render_batch(views, bounces=3, seed=41, env=ENV, flags=oracle.F_POLISH)
    ref = np.concatenate([x.reshape(-1, 3) for x in ri])
    assert f"{nb} bounce rays, {ns} shadow rays" in r.stdout
    assert np.array_equal(got[:, :3].view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("bounces", [0, 8])
def test_batch_edge_views(gpu_ctx, oracle, scene, bounces):
    """1x1 and odd-sized views, a view that sees no hair (every path adds the environment), the
    bounce extremes, and a repeated call with the same seed: all bit for bit equal to the oracle
    and to the first call."""
    import scenes
    S, tlas = scene
    specs = [(scenes.DENSE_VIEWS[0][0], scenes.DENSE_VIEWS[0][1], 0.577, (1, 1)),
             (scenes.DENSE_VIEWS[1][0], scenes.DENSE_VIEWS[1][1], 0.4, (7, 3)),
             ((0.0, 20.0, 45.0), (0.0, 20.0, 500.0), 0.3, (16, 9)),  # looking away from the hair
             (scenes.DENSE_VIEWS[2][0], scenes.DENSE_VIEWS[2][1], 0.5, (33, 17))]
    views = _views(oracle, specs)
    print(f"[batch edge views] bounces {bounces}", flush=True)
    gi, gp, (paths, nb, ns, _) = _gpu_batch(gpu_ctx, tlas, views, bounces, 9)
    gi2, gp2, st2 = _gpu_batch(gpu_ctx, tlas, views, bounces, 9)
    ri, rp, rnb, rns, _ = S.render_batch(views, bounces=bounces, seed=9, env=ENV, albedo=TINTS)
    assert paths == 1 + 21 + 144 + 561 and (nb, ns) == (rnb, rns) == st2[1:3]
    _assert_same(gi, gp, ri, rp, f"edge views, bounces {bounces}")
    assert all(np.array_equal(a.view(np.uint32), b.view(np.uint32)) for a, b in zip(gi, gi2))
    away = gi[2][..., :3]
    assert np.all(away == np.float32(ENV)) and np.all(gp[2] == 0xFFFFFFFF)
    if bounces == 0:
        assert nb == 0
