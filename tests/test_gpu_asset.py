"""GPU parity of glTF assets built on the device (vkhrt_asset_build): ProcessHairCurves /
ProcessHairLSS / ProcessHairDOTS (geometry_processor.cpp:604-688,740-779) and
Renderer::InitializeBLAS (renderer.cpp:694-728) over a multi-mesh, multi-node asset.

The oracle scene is assembled from the loader's host arrays: GenerateLines over each mesh's own
(mesh-local) indices, GenerateCurves / LSS / DOTS, one instance per (node, mesh) with the node's
world matrix. Traces and renders must agree bit for bit, which also shows that every hair BLAS
reads its own mesh's curves in its own object space."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

RADIUS = 0.02
DEFAULT_MATERIAL = 2
ALBEDO = [(0.8, 0.6, 0.3), (0.3, 0.4, 0.9), (0.9, 0.9, 0.9)]


def _quat(axis, angle):
    axis = np.asarray(axis, np.float64)
    axis /= np.linalg.norm(axis)
    return list(np.r_[axis * np.sin(angle / 2), np.cos(angle / 2)])


def _asset(oracle, with_mesh=False):
    import vkhrt_amd as V
    return V.Asset(data=_gltf(oracle, with_mesh).glb())


def _gltf(oracle, with_mesh=False):
    """ROOT -> a (head: two line primitives) -> b (tail, rotated and scaled); ROOT -> c (tail
    again, translated); optionally ROOT -> d (a triangle ball)."""
    import gltf_writer as GW
    import scenes
    g = GW.Gltf()
    v0 = oracle.generate_strands(21, 900, 8, sphere_radius=12.0)
    v1 = oracle.generate_strands(22, 500, 6, sphere_radius=8.0)
    p0, i0 = GW.strand_lines_mesh(v0, 900, 8)
    p1, i1 = GW.strand_lines_mesh(v1, 500, 6)
    m0 = g.material(base_color=list(ALBEDO[0]) + [1.0])
    m1 = g.material(base_color=list(ALBEDO[1]) + [1.0])
    head = g.mesh([{"positions": p0, "indices": i0, "mode": 1, "material": m0},
                   {"positions": p1, "indices": i1, "mode": 1, "material": m1}])
    tail = g.mesh([{"positions": p1, "indices": i1, "mode": 1}])
    b = g.node(mesh=tail, translation=[0.0, -6.0, 25.0], rotation=_quat([1, 0, 0], 0.7), scale=[0.5, 0.5, 0.5])
    a = g.node(mesh=head, children=[b], translation=[3.0, 1.0, -2.0], rotation=_quat([0, 1, 0], 0.4))
    c = g.node(mesh=tail, translation=[-30.0, 4.0, 0.0])
    roots = [a, c]
    if with_mesh:
        sv, si = scenes.uv_sphere(5.0, 16, 32, (0.0, 0.0, 0.0))
        ball = g.mesh([{"positions": sv[:, :3], "normals": sv[:, 3:6], "indices": si.reshape(-1), "material": m1}])
        roots.append(g.node(mesh=ball, translation=[0.0, 25.0, 0.0]))
    g.scene(roots)
    return g


def _oracle_scene(oracle, asset, mode):
    import vkhrt_amd as V
    S = oracle.Scene()
    world_curves = []
    for k, inst in enumerate(asset.instances):
        m = asset.mesh(inst["mesh"])
        xf = inst["transform"].reshape(3, 4)
        if m["primitive"] == V.PRIMITIVE_LINES:
            lines = oracle.mesh_lines(m["vertices"], m["indices"], RADIUS)
            curves = oracle.generate_curves(lines)
            if mode == V.GEOMETRY_PHANTOM_CURVES:
                S.add(oracle.MODE_PHANTOM, curves, radius=RADIUS, transform=xf)
            elif mode == V.GEOMETRY_LSS:
                S.add(oracle.MODE_LSS, oracle.generate_lss(lines, RADIUS), transform=xf)
            else:
                S.add(oracle.MODE_DOTS, lines, radius=RADIUS, transform=xf)
            c = curves.reshape(-1, 4, 4).astype(np.float64)
            c[:, :, :3] = c[:, :, :3] @ xf[:, :3].T.astype(np.float64) + xf[:, 3]
            world_curves.append(c.reshape(-1, 16).astype(np.float32))
        else:
            S.add_mesh(m["vertices"], m["indices"].reshape(-1, 3), transform=xf)
        S.set_material(k, m["material"] if m["material"] >= 0 else DEFAULT_MATERIAL)
    return S, np.concatenate(world_curves)


@pytest.mark.parametrize("mode", [0, 1, 2], ids=["phantom", "lss", "dots"])
def test_asset_trace_matches_oracle(gpu_ctx, oracle, mode):
    import gpu_util as G
    import scenes
    import vkhrt_amd as V
    asset = _asset(oracle)
    assert [(i["node"], i["mesh"]) for i in asset.instances] == [(1, 0), (1, 1), (2, 2), (3, 2)]
    print(f"[asset trace] hair mode {mode}", flush=True)
    blases = asset.build(gpu_ctx, hair_mode=mode, radius=RADIUS, default_material=DEFAULT_MATERIAL,
                         stream=G.stream_ptr())
    torch.cuda.synchronize()
    assert len(blases) == 4 and all(b.mode == mode for b in blases)
    nlines = [asset.mesh(i["mesh"])["indices"].size // 2 for i in asset.instances]
    assert [b.info()["leaf_records"] for b in blases] == nlines
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, blases)
    S, wc = _oracle_scene(oracle, asset, mode)
    rays = scenes.targeted_rays(wc, 30000, seed=31, dist=(5.0, 120.0))
    h = G.trace(gpu_ctx, tlas, rays)
    nhit = G.compare_hits(h, S.trace(rays), f"asset mode {mode}")
    per_inst = np.bincount(h["inst"][h["prim"] != 0xFFFFFFFF], minlength=4)
    print(f"hits {nhit}, per instance {per_inst.tolist()}")
    assert nhit > 10000 and (per_inst > 500).all()
    any_g = G.trace(gpu_ctx, tlas, rays, V.TRACE_TERMINATE_ON_FIRST_HIT)
    any_r = S.trace(rays, flags=oracle.F_ANY_HIT)
    assert np.array_equal(any_g["prim"] != 0xFFFFFFFF, any_r["prim"] != 0xFFFFFFFF)
    tlas.close()
    for b in blases:
        b.close()


def test_asset_with_mesh_renders_like_the_oracle(gpu_ctx, oracle):
    """Hair and a triangle mesh from one asset: primary render with shadows and the batched path
    tracer (material tints from the asset's materials, the default for the material-less tail)
    agree with the oracle bit for bit."""
    import gpu_util as G
    import vkhrt_amd as V
    asset = _asset(oracle, with_mesh=True)
    blases = asset.build(gpu_ctx, hair_mode=V.GEOMETRY_PHANTOM_CURVES, radius=RADIUS,
                         default_material=DEFAULT_MATERIAL, stream=G.stream_ptr())
    assert [b.mode for b in blases] == [0, 0, 0, 0, V.GEOMETRY_TRIANGLES]
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, blases)
    S, _ = _oracle_scene(oracle, asset, V.GEOMETRY_PHANTOM_CURVES)
    W, H = 160, 120
    vi, pi = oracle.camera((5.0, 12.0, 50.0), (-6.0, 8.0, 5.0), aspect=W / H)  # every instance in view
    cam = V.Camera()
    for k in range(16):
        cam.view_inverse[k], cam.proj_inverse[k] = float(vi[k]), float(pi[k])
    print("[asset render] 160x120 with shadows", flush=True)
    img, ids = G.render(gpu_ctx, tlas, cam, W, H)
    ref, rids, _, _ = S.render(vi, pi, W, H, shadows=1)
    assert (rids != 0xFFFFFFFF).sum() > 2000
    assert np.array_equal(ids, rids) and np.array_equal(img[..., :3].view(np.uint32), ref.view(np.uint32))
    out = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
    print("[asset batch] 2 bounces", flush=True)
    gpu_ctx.render_batch(tlas, [(cam, W, H, out.data_ptr())], bounces=2, seed=5, environment=(0.2, 0.3, 0.4),
                         material_albedo=ALBEDO, stream=G.stream_ptr())
    torch.cuda.synchronize()
    ri, _, _, _, _ = S.render_batch([(vi, pi, W, H)], bounces=2, seed=5, env=(0.2, 0.3, 0.4), albedo=ALBEDO)
    got = out.cpu().numpy().reshape(H, W, 4)[..., :3]
    assert np.array_equal(got.view(np.uint32), ri[0].view(np.uint32))


def test_device_mesh_lines_match_the_oracle(gpu_ctx, oracle):
    import gpu_util as G
    rng = np.random.default_rng(12)
    verts = rng.normal(size=(5000, 8)).astype(np.float32)
    idx = rng.integers(0, 5000, 2 * 30000).astype(np.uint32)
    dv, di = G.dev(verts), G.dev(idx)
    out = torch.zeros(30000 * 8, dtype=torch.float32, device="cuda")
    gpu_ctx.generate_mesh_lines(dv.data_ptr(), 5000, di.data_ptr(), 30000, out.data_ptr(), radius=0.125,
                                stream=G.stream_ptr())
    torch.cuda.synchronize()
    ref = oracle.mesh_lines(verts, idx, 0.125)
    assert np.array_equal(out.cpu().numpy().reshape(-1, 8).view(np.uint32), ref.view(np.uint32))
    import vkhrt_amd as V
    bad = idx.copy()
    bad[7] = 5000
    db = G.dev(bad)
    with pytest.raises(V.VkhrtError, match="1 line indices are not below"):
        gpu_ctx.generate_mesh_lines(dv.data_ptr(), 5000, db.data_ptr(), 30000, out.data_ptr(), stream=G.stream_ptr())


def test_asset_build_arguments(gpu_ctx, oracle):
    import ctypes

    import vkhrt_amd as V
    asset = _asset(oracle)
    L = V.lib()
    d = V.AssetBuildDesc(V.GEOMETRY_PHANTOM_CURVES, RADIUS, 1.0, 0, 0)
    cnt = ctypes.c_uint32()
    assert L.vkhrt_asset_build(gpu_ctx.h, asset.h, ctypes.byref(d), None, None, 0, ctypes.byref(cnt)) == 0
    assert cnt.value == 4
    arr = (ctypes.c_void_p * 4)()
    assert L.vkhrt_asset_build(gpu_ctx.h, asset.h, ctypes.byref(d), None, arr, 3, ctypes.byref(cnt)) == -1
    assert b"capacity 3" in L.vkhrt_last_error(gpu_ctx.h)
    d.hair_mode = V.GEOMETRY_TRIANGLES
    assert L.vkhrt_asset_build(gpu_ctx.h, asset.h, ctypes.byref(d), None, arr, 4, ctypes.byref(cnt)) == -1
    d.hair_mode, d.radius = V.GEOMETRY_LSS, float("nan")
    assert L.vkhrt_asset_build(gpu_ctx.h, asset.h, ctypes.byref(d), None, arr, 4, ctypes.byref(cnt)) == -1
    assert all(p is None for p in arr)


def test_cpp_model_loader_trace(gpu_ctx, oracle, tmp_path):
    """The C++ mirror (ModelLoader::LoadFromFile -> Model::BuildBLASes -> TLAS -> TraceRays)
    traces the asset file like the oracle, bit for bit."""
    import subprocess

    import gpu_util as G
    import scenes
    from test_host import cpp_test_binary
    g = _gltf(oracle)
    path = tmp_path / "hair.gltf"
    g.write_gltf(str(path), bin_name="hair.bin")
    asset = _asset(oracle)
    S, wc = _oracle_scene(oracle, asset, 0)
    rays = scenes.targeted_rays(wc, 20000, seed=32, dist=(5.0, 120.0))
    rpath, hpath = tmp_path / "rays.bin", tmp_path / "hits.bin"
    rays.tofile(rpath)
    print("[asset] C++ mirror trace", flush=True)
    r = subprocess.run([cpp_test_binary(), "asset", str(path), "0", str(rpath), str(rays.shape[0]), str(hpath)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    h = np.fromfile(hpath, dtype=G.HIT_DTYPE)
    assert G.compare_hits(h, S.trace(rays), "C++ asset") > 5000
