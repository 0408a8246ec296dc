"""GPU parity of voxel hair (BLASType::eVoxels): the device voxelization (GenerateVoxelMesh,
geometry_processor.cpp:515-602, both sweeps, levels of detail) equals the oracle's cell for cell,
and traces, renders and batched path tracing through voxel instances (alone, transformed, and
mixed with phantom hair) equal the oracle's walk bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

XF = np.array([[0.0, -1.0, 0.0, 30.0], [1.0, 0.0, 0.0, 2.0], [0.0, 0.0, 1.0, -5.0]], np.float32)


def _lines(oracle, nstrands=1200, nseg=8, seed=5, radius=12.0):
    verts = oracle.generate_strands(seed, nstrands, nseg, sphere_radius=radius)
    return oracle.strands_to_lines(verts)


def _voxel_blas(ctx, G, lines, size, lod=0, ref=False, transform=None, bounds=None, radius=0.02):
    import vkhrt_amd as V
    d = G.dev(lines)
    b = V.BottomLevelAccelerationStructure(ctx, V.GEOMETRY_VOXELS, d.data_ptr(), lines.shape[0], radius=radius,
                                           flags=V.BLAS_VOXEL_REFERENCE_SWEEP if ref else 0, transform=transform,
                                           voxel_size=size, voxel_lod=lod, bounds=bounds, stream=G.stream_ptr())
    torch.cuda.synchronize()
    return b


def _cells(bits, res):
    n = int(res[0]) * int(res[1]) * int(res[2])
    return np.unpackbits(bits.view(np.uint8), bitorder="little")[:n].astype(bool).reshape(res[2], res[1], res[0])


@pytest.mark.parametrize("ref,lod,size", [(False, 0, 0.1), (True, 0, 0.1), (False, 2, 0.1), (True, 1, 0.25)])
def test_voxelization_matches_oracle(gpu_ctx, oracle, ref, lod, size):
    import gpu_util as G
    lines = _lines(oracle)
    print(f"[voxelize] {lines.shape[0]} lines, size {size}, reference sweep {ref}, lod {lod}", flush=True)
    b = _voxel_blas(gpu_ctx, G, lines, size, lod=lod, ref=ref)
    cells, origin, cell = b.voxel_grid()
    rbits, rres, rorg, rcell = oracle.voxelize(lines, size, 0.02, reference_sweep=ref, lod=lod)
    want = _cells(rbits, rres)
    assert cells.shape == want.shape and np.array_equal(origin, rorg) and np.float32(cell) == np.float32(rcell)
    assert np.array_equal(cells, want)
    assert b.info()["leaf_records"] == int(want.sum()) > 1000
    b.close()


def _rays(rng, n, lo, hi):
    """Rays from outside the box aimed at points inside it, some starting inside."""
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    c, ext = (lo + hi) / 2, (hi - lo)
    tgt = lo + rng.random((n, 3)) * ext
    o = c + rng.normal(size=(n, 3)) * ext.max()
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, np.full((n, 1), 0.001), d, np.full((n, 1), 1e4)], 1).astype(np.float32)
    inside = rng.random(n) < 0.15
    rays[inside, :3] = (tgt[inside] - d[inside] * 0.3).astype(np.float32)
    dd = rays[:, 4:7]
    rays[:, 4:7] = dd / np.sqrt((dd * dd).sum(1, keepdims=True))
    return rays


def test_voxel_trace_matches_oracle(gpu_ctx, oracle):
    """A voxel instance, the same grid transformed (level of detail 1), and phantom hair in one
    TLAS; closest and any hit."""
    import gpu_util as G
    import scenes
    import vkhrt_amd as V
    lines = _lines(oracle)
    v0 = _voxel_blas(gpu_ctx, G, lines, 0.1)
    v1 = _voxel_blas(gpu_ctx, G, lines, 0.1, lod=1, transform=XF.reshape(12))
    curves = scenes.dense_ball_curves(nstrands=600, nseg=8, seed=9)
    hair = G.build_blas(gpu_ctx, V.GEOMETRY_PHANTOM_CURVES, curves, per_point=True,
                        transform=np.array([1, 0, 0, -30.0, 0, 1, 0, 0, 0, 0, 1, 0], np.float32))
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, [v0, v1, hair])
    S = oracle.Scene()
    for lod, xf in ((0, None), (1, XF)):
        bits, res, org, cell = oracle.voxelize(lines, 0.1, 0.02, lod=lod)
        S.add_voxels(bits, res, org, cell, transform=xf)
    S.add(oracle.MODE_PHANTOM, curves, per_point=True, transform=np.array([[1, 0, 0, -30.0], [0, 1, 0, 0], [0, 0, 1, 0]]))
    rng = np.random.default_rng(3)
    rays = _rays(rng, 40000, (-45.0, -15.0, -20.0), (45.0, 30.0, 20.0))
    print("[voxel trace] 40k rays, 3 instances", flush=True)
    h = G.trace(gpu_ctx, tlas, rays)
    nhit = G.compare_hits(h, S.trace(rays), "voxels")
    per = np.bincount(h["inst"][h["prim"] != 0xFFFFFFFF], minlength=3)
    print(f"hits {nhit}, per instance {per.tolist()}")
    assert (per > 1000).all()
    g_any = G.trace(gpu_ctx, tlas, rays, V.TRACE_TERMINATE_ON_FIRST_HIT)
    r_any = S.trace(rays, flags=oracle.F_ANY_HIT)
    assert np.array_equal(g_any["prim"] != 0xFFFFFFFF, r_any["prim"] != 0xFFFFFFFF)
    tlas.close()


@pytest.mark.parametrize("palette", [False, True])
def test_voxel_render_and_batch_match_oracle(gpu_ctx, oracle, palette):
    """traceRaysKHR with voxel_closest_hit.rchit's debug colour of the cell index (palette) or
    the shaded face normal with shadow rays, and the batched path tracer; bit for bit."""
    import gpu_util as G
    import vkhrt_amd as V
    lines = _lines(oracle, nstrands=1500)
    blas = _voxel_blas(gpu_ctx, G, lines, 0.2, ref=True)
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, [blas])
    S = oracle.Scene()
    bits, res, org, cell = oracle.voxelize(lines, 0.2, 0.02, reference_sweep=True)
    S.add_voxels(bits, res, org, cell)
    W, H = 128, 96
    vi, pi = oracle.camera((10.0, 25.0, 45.0), (0.0, 6.0, 0.0), aspect=W / H)
    cam = V.Camera()
    for k in range(16):
        cam.view_inverse[k], cam.proj_inverse[k] = float(vi[k]), float(pi[k])
    flags = V.RENDER_DEBUG_PALETTE if palette else V.RENDER_SHADOWS
    print(f"[voxel render] palette {palette}", flush=True)
    img, ids = G.render(gpu_ctx, tlas, cam, W, H, flags=flags)
    ref, rids, _, _ = S.render(vi, pi, W, H, shadows=0 if palette else 1)
    assert (rids != 0xFFFFFFFF).sum() > 2000
    assert np.array_equal(ids, rids)
    if palette:  # GetDebugColor(gl_PrimitiveID) on hits, black misses
        L = oracle.lib()
        want = np.zeros((H, W, 3), np.float32)
        rgb = np.zeros(3, np.float32)
        for y, x in zip(*np.nonzero(rids != 0xFFFFFFFF)):
            L.orc_debug_color(int(rids[y, x]), oracle.ptr(rgb))  # cell indices are below 2^31
            want[y, x] = rgb
        assert np.array_equal(img[..., :3], want)
        return
    assert np.array_equal(img[..., :3].view(np.uint32), ref.view(np.uint32))
    out = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
    gpu_ctx.render_batch(tlas, [(cam, W, H, out.data_ptr())], bounces=2, seed=8, environment=(0.2, 0.3, 0.4),
                         stream=G.stream_ptr())
    torch.cuda.synchronize()
    ri, _, _, _, _ = S.render_batch([(vi, pi, W, H)], bounces=2, seed=8, env=(0.2, 0.3, 0.4))
    assert np.array_equal(out.cpu().numpy().reshape(H, W, 4)[..., :3].view(np.uint32), ri[0].view(np.uint32))


def test_asset_voxels_match_oracle(gpu_ctx, oracle):
    """ProcessHairVoxels on an asset: each line mesh voxelized over its own bounding box, one
    BLAS per (node, mesh) with the node's transform."""
    import gpu_util as G
    import scenes
    import vkhrt_amd as V
    from test_gpu_asset import _asset
    asset = _asset(oracle)
    blases = asset.build(gpu_ctx, hair_mode=V.GEOMETRY_VOXELS, voxel_size=0.1, stream=G.stream_ptr())
    assert all(b.mode == V.GEOMETRY_VOXELS for b in blases)
    tlas = V.TopLevelAccelerationStructure(gpu_ctx, blases)
    S = oracle.Scene()
    wc = []
    for k, inst in enumerate(asset.instances):
        m = asset.mesh(inst["mesh"])
        lines = oracle.mesh_lines(m["vertices"], m["indices"], 0.02)
        bits, res, org, cell = oracle.voxelize(lines, 0.1, 0.02, bounds=(m["bounds_min"], m["bounds_max"]))
        cells, gorg, gcell = blases[k].voxel_grid()
        assert np.array_equal(cells, _cells(bits, res)) and np.array_equal(gorg, org)
        xf = inst["transform"].reshape(3, 4)
        S.add_voxels(bits, res, org, cell, transform=xf)
        c = oracle.generate_curves(lines).reshape(-1, 4, 4).astype(np.float64)
        c[:, :, :3] = c[:, :, :3] @ xf[:, :3].T.astype(np.float64) + xf[:, 3]
        wc.append(c.reshape(-1, 16).astype(np.float32))
    rays = scenes.targeted_rays(np.concatenate(wc), 20000, seed=41, dist=(5.0, 100.0))
    print("[asset voxels] 20k rays", flush=True)
    assert G.compare_hits(G.trace(gpu_ctx, tlas, rays), S.trace(rays), "asset voxels") > 10000


def test_voxel_build_errors(gpu_ctx, oracle):
    import gpu_util as G
    import vkhrt_amd as V
    lines = _lines(oracle, nstrands=50)
    for kw, msg in ((dict(size=0.0), "voxel size"), (dict(size=0.1, lod=16), "level of detail"),
                    (dict(size=1e-4), r"2\^31 cells")):
        size = kw.pop("size")
        with pytest.raises(V.VkhrtError, match=msg):
            _voxel_blas(gpu_ctx, G, lines, size, **kw)
    # a radius far above the cell size: each line would visit over 65536 cells
    fat = np.array([[0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0]], np.float32)
    box = (np.full(3, -3.0, np.float32), np.full(3, 3.0, np.float32))
    with pytest.raises(ValueError):
        oracle.voxelize(fat, 0.05, 2.0, bounds=box)
    with pytest.raises(V.VkhrtError, match="65536 voxel cells"):
        _voxel_blas(gpu_ctx, G, fat, 0.05, radius=2.0, bounds=box)
    # an empty grid (no lines) traces as all misses
    b = V.BottomLevelAccelerationStructure(gpu_ctx, V.GEOMETRY_VOXELS, None, 0, voxel_size=0.1)
    assert b.info()["leaf_records"] == 0
    tl = V.TopLevelAccelerationStructure(gpu_ctx, [b])
    h = G.trace(gpu_ctx, tl, _rays(np.random.default_rng(0), 100, (-1, -1, -1), (1, 1, 1)))
    assert (h["prim"] == 0xFFFFFFFF).all()
