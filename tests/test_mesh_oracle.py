"""Pins of the oracle's triangle-mesh mode (the reference's BLASType::eMesh path,
triangle_closest_hit.rchit:19-41): watertight hits against closed forms, the interpolated
vertex normal, a sphere mesh against the analytic sphere, and the BVH2 against brute force."""
import numpy as np


def _tri_verts(normals):
    p = np.array([[0, 0, 0], [4, 0, 0], [0, 3, 0]], np.float32)
    return np.concatenate([p, np.asarray(normals, np.float32), np.zeros((3, 2), np.float32)], axis=1)


def test_single_triangle_barycentrics_and_normal(oracle):
    n = np.array([[0, 0, 1], [0.6, 0, 0.8], [0, -0.6, 0.8]], np.float64)
    S = oracle.Scene()
    S.add_mesh(_tri_verts(n), [[0, 1, 2]])
    rng = np.random.default_rng(3)
    b = rng.uniform(0.05, 0.9, (500, 2))
    b = b[b.sum(1) < 0.95]
    pts = np.stack([4 * b[:, 0], 3 * b[:, 1], np.zeros(len(b))], 1)
    rays = np.zeros((len(b), 8), np.float32)
    rays[:, 0:3] = pts + np.array([0.3, -0.2, 7.0])
    d = pts - rays[:, 0:3]
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 3], rays[:, 7] = 0.001, 100.0
    h = oracle.Scene.trace(S, rays)
    assert np.all(h["prim"] == 0)
    assert np.allclose(h["u"], b[:, 0], atol=2e-6) and np.allclose(h["v"], b[:, 1], atol=2e-6)
    assert np.allclose(h["t"], np.linalg.norm(d, axis=1), rtol=2e-6)
    s = (1 - b.sum(1))[:, None] * n[0] + b[:, :1] * n[1] + b[:, 1:] * n[2]
    s /= np.linalg.norm(s, axis=1, keepdims=True)
    assert np.abs(h["n"] - s).max() < 2e-6
    # a ray outside the triangle and one parallel to it miss
    miss = np.array([[5, 5, 5, 0.001, 0, 0, -1, 100], [-1, 1, 0.5, 0.001, 1, 0, 0, 100]], np.float32)
    assert np.all(oracle.Scene.trace(S, miss)["prim"] == 0xFFFFFFFF)


def test_zero_vertex_normals_fall_back_to_the_geometric_normal(oracle):
    S = oracle.Scene()
    S.add_mesh(_tri_verts(np.zeros((3, 3))), [[0, 1, 2]])
    h = S.trace(np.array([[1, 1, 5, 0.001, 0, 0, -1, 100]], np.float32))
    assert np.allclose(h["n"][0], [0, 0, 1], atol=1e-7)


def test_sphere_mesh_against_the_analytic_sphere(oracle):
    import scenes
    R, c = 10.0, np.array([1.0, 2.0, -3.0])
    verts, idx = scenes.uv_sphere(R, 48, 96, c)
    S = oracle.Scene()
    S.add_mesh(verts, idx)
    rng = np.random.default_rng(8)
    o = rng.normal(size=(4000, 3))
    o = c + 40 * o / np.linalg.norm(o, axis=1, keepdims=True)
    tgt = c + rng.uniform(-6, 6, (4000, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, np.full((4000, 1), 0.001), d, np.full((4000, 1), 1e4)], 1).astype(np.float32)
    h = S.trace(rays)
    oc = o - c
    bq = (oc * d).sum(1)
    disc = bq * bq - ((oc * oc).sum(1) - R * R)
    t_true = -bq - np.sqrt(np.maximum(disc, 0))
    hit = h["prim"] != 0xFFFFFFFF
    assert hit.mean() > 0.95
    # the mesh lies inside the sphere by at most the chord sag R (1 - cos(pi / 48))
    sag = R * (1 - np.cos(np.pi / 48))
    assert np.all(h["t"][hit] >= t_true[hit] - 1e-3) and np.all(h["t"][hit] <= t_true[hit] + 4 * sag)
    p = o[hit] + h["t"][hit, None] * d[hit]
    ang = np.arccos(np.clip((h["n"][hit] * (p - c) / np.linalg.norm(p - c, axis=1, keepdims=True)).sum(1), -1, 1))
    assert ang.max() < 0.05  # interpolated radial normals


def test_mesh_bvh_equals_brute_force(oracle):
    import scenes
    verts, idx = scenes.uv_sphere(10.0, 16, 32)
    xf = np.array([[0.0, -1, 0, 4], [1, 0, 0, 0], [0, 0, 1, 2]], np.float32)
    rng = np.random.default_rng(2)
    o = rng.uniform(-30, 30, (3000, 3))
    d = rng.uniform(-12, 12, (3000, 3)) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, np.full((3000, 1), 0.001), d, np.full((3000, 1), 1e4)], 1).astype(np.float32)
    S = oracle.Scene()
    S.add_mesh(verts, idx, transform=xf)
    a = S.trace(rays)
    b = S.trace(rays, flags=oracle.F_BRUTE)
    assert (a["prim"] != 0xFFFFFFFF).mean() > 0.3
    assert a.tobytes() == b.tobytes()
    any_hit = S.trace(rays, flags=oracle.F_ANY_HIT)
    assert np.array_equal(any_hit["prim"] != 0xFFFFFFFF, a["prim"] != 0xFFFFFFFF)


def test_mesh_rejects_out_of_range_indices(oracle):
    import pytest
    S = oracle.Scene()
    with pytest.raises(AssertionError):
        S.add_mesh(_tri_verts(np.zeros((3, 3))), [[0, 1, 3]])
