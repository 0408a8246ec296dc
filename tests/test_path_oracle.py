"""Pins of the oracle's batched path tracer (oracle/oracle_path.hpp, orc_render_batch).

The path rule has no reference counterpart, so each part is checked against what fixes it
independently: SplitMix64 re-implemented here in Python integers, the orthonormal-basis
invariants, the moments of the cosine lobe, and closed forms of whole paths built from the
oracle's already-pinned tracer (direct light only; one bounce off a lone convex fibre, whose
bounce rays must all escape).
"""
import math

import numpy as np
import pytest

M64 = (1 << 64) - 1


def _splitmix_out(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def _rand_py(seed, p, b, k):
    key = (p << 24) | (b << 16) | k
    bits = _splitmix_out(((seed ^ key) + 0x9E3779B97F4A7C15) & M64) >> 40
    return bits / float(1 << 24)


def test_rng_matches_splitmix64(oracle):
    rng = np.random.default_rng(5)
    for _ in range(2000):
        seed = int(rng.integers(0, 1 << 63)) * 2 + int(rng.integers(0, 2))
        p, b, k = int(rng.integers(0, 1 << 31)), int(rng.integers(0, 9)), int(rng.integers(0, 32))
        assert oracle.path_rand(seed, p, b, k) == _rand_py(seed, p, b, k)


def test_rng_uniform(oracle):
    u = np.array([oracle.path_rand(12345, p, 0, 0) for p in range(40000)])
    assert u.min() >= 0.0 and u.max() < 1.0
    assert np.all(u * (1 << 24) == np.floor(u * (1 << 24)))  # 24-bit grid
    # Kolmogorov-Smirnov against U[0, 1): D < 1.63 / sqrt(n) at the 1% level
    s = np.sort(u)
    n = len(s)
    d = max(np.max(np.arange(1, n + 1) / n - s), np.max(s - np.arange(n) / n))
    assert d < 1.63 / math.sqrt(n)
    # neighbouring keys are not correlated
    v = np.array([oracle.path_rand(12345, p, 0, 1) for p in range(40000)])
    assert abs(np.corrcoef(u, v)[0, 1]) < 0.02


def _normals(rng, n):
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    special = np.array([[0, 0, 1], [0, 0, -1], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0],
                        [0.6, 0.0, -0.8], [0, 0.6, 0.8]], np.float64)
    return np.concatenate([v, special]).astype(np.float32)


def test_basis_is_orthonormal_and_right_handed(oracle):
    for n in _normals(np.random.default_rng(7), 500):
        t1, t2 = oracle.path_basis(n)
        n64, a, b = n.astype(np.float64), t1.astype(np.float64), t2.astype(np.float64)
        assert abs(np.linalg.norm(a) - 1) < 1e-6 and abs(np.linalg.norm(b) - 1) < 1e-6
        assert abs(a @ b) < 1e-6 and abs(a @ n64) < 1e-6 and abs(b @ n64) < 1e-6
        assert np.allclose(np.cross(a, b), n64, atol=2e-6)
    # signed zero: n.z = -0 takes the negative branch without dividing by zero
    t1, t2 = oracle.path_basis(np.array([0.0, 0.0, -0.0], np.float32) + np.array([0.6, 0.8, 0.0], np.float32))
    assert np.all(np.isfinite(t1)) and np.all(np.isfinite(t2))


def test_cosine_lobe_moments(oracle):
    for n in _normals(np.random.default_rng(9), 6)[[0, 1, 2, 6, 7, 9]]:
        w = oracle.path_cosine_dirs(n, seed=99, p0=0, count=200000, depth=1).astype(np.float64)
        t1, t2 = (x.astype(np.float64) for x in oracle.path_basis(n))
        c = w @ n.astype(np.float64)
        assert np.abs(np.linalg.norm(w, axis=1) - 1).max() < 1e-6
        assert c.min() >= -1e-6  # the hemisphere about n
        # cos-weighted: E[cos] = 2/3, E[cos^2] = 1/2, tangential components mean zero
        assert abs(c.mean() - 2 / 3) < 0.004
        assert abs((c * c).mean() - 0.5) < 0.004
        assert abs((w @ t1).mean()) < 0.004 and abs((w @ t2).mean()) < 0.004
        # P(cos > 0.5) = 1 - 0.25 for the cosine lobe
        assert abs((c > 0.5).mean() - 0.75) < 0.005


@pytest.fixture(scope="module")
def small_scene(oracle):
    import scenes
    curves = scenes.dense_ball_curves()
    S = oracle.Scene()
    S.add(oracle.MODE_PHANTOM, curves, per_point=True)
    xf = np.array([[1, 0, 0, 14.0], [0, 1, 0, 3.0], [0, 0, 1, -8.0]], np.float32)
    S.add(oracle.MODE_PHANTOM, curves[::3], per_point=True, transform=xf)
    S.set_material(1, 1)
    return S, curves


def _views(oracle):
    import scenes
    out = []
    for eye, tgt, th, (w, h) in scenes.DENSE_VIEWS:
        vi, pi = oracle.camera(eye, tgt, tan_half_fovy=th, aspect=w / h)
        out.append((vi, pi, w, h))
    return out


def _direct_closed_form(oracle, S, views, env, tints):
    """Bounces = 0: miss -> env; hit -> tint * Shade(n) * nf.y when nf.y > 0 and the shadow ray
    toward (0, 1, 0) is unoccluded, from the oracle's traced primary and shadow hits."""
    imgs = []
    for vi, pi, w, h in views:
        rays = oracle.raygen(vi, pi, w, h)
        hits = S.trace(rays)
        hit = hits["prim"] != 0xFFFFFFFF
        o, d = rays[:, 0:3].astype(np.float64), rays[:, 4:7].astype(np.float64)
        n = hits["n"].astype(np.float64)
        x = o + hits["t"][:, None].astype(np.float64) * d
        nf = np.where(((n * d).sum(1) > 0)[:, None], -n, n)
        c = nf[:, 1]
        srays = np.zeros_like(rays)
        srays[:, 0:3] = x
        srays[:, 3] = 0.001
        srays[:, 4:7] = (0.0, 1.0, 0.0)
        srays[:, 7] = 10000.0
        vis = S.trace(srays, flags=oracle.F_ANY_HIT)["prim"] == 0xFFFFFFFF
        shade = np.abs(n[:, 1])[:, None] * np.array([0.4, 0.2, 0.1]) + 0.3
        tint = np.asarray(tints, np.float64)[np.where(hit, hits["inst"], 0)]
        val = tint * shade * (c * ((c > 0) & vis))[:, None]
        img = np.where(hit[:, None], val, np.asarray(env, np.float64)[None, :])
        imgs.append(img.reshape(h, w, 3))
    return imgs


def test_direct_light_closed_form(oracle, small_scene):
    S, _ = small_scene
    views = _views(oracle)
    env, tints = (0.1, 0.2, 0.3), [(1.0, 1.0, 1.0), (0.5, 0.75, 1.25)]
    imgs, ids, nb, ns, _ = S.render_batch(views, bounces=0, seed=3, env=env, albedo=tints)
    ref = _direct_closed_form(oracle, S, views, env, tints)
    nhit = 0
    for img, r, pid in zip(imgs, ref, ids):
        assert np.allclose(img, r, rtol=2e-6, atol=1e-7)
        nhit += int((pid != 0xFFFFFFFF).sum())
    assert nb == 0 and 0 < ns <= nhit
    assert nhit > 500


def test_one_bounce_off_a_convex_fibre(oracle):
    """A lone straight fibre (capped cylinder, convex): every bounce ray leaves the surface
    outward and escapes, so with bounces = 1 a hit pixel is tint * Shade(n) * (max(nf.y, 0)
    + env) (its shadow ray escapes too when nf.y > 0)."""
    p0, p1 = np.array([-30.0, 2.0, 0.0]), np.array([30.0, -2.0, 1.0])
    ctrl = [p0 + (p1 - p0) * k / 3 for k in range(4)]
    curve = np.array([[*c, 0.0] for c in ctrl], np.float32).reshape(1, 16)
    S = oracle.Scene()
    S.add(oracle.MODE_PHANTOM, curve, per_point=False, radius=1.5)
    vi, pi = oracle.camera((0.0, 8.0, 40.0), (0.0, 0.0, 0.0), aspect=1.5)
    view = (vi, pi, 96, 64)
    env = (0.25, 0.5, 1.0)
    imgs, ids, nb, ns, _ = S.render_batch([view], bounces=1, seed=11, env=env)
    rays = oracle.raygen(vi, pi, 96, 64)
    hits = S.trace(rays)
    hit = hits["prim"] != 0xFFFFFFFF
    assert hit.sum() > 300 and nb == hit.sum()
    n = hits["n"].astype(np.float64)
    d = rays[:, 4:7].astype(np.float64)
    nf = np.where(((n * d).sum(1) > 0)[:, None], -n, n)
    a = np.abs(n[:, 1])[:, None] * np.array([0.4, 0.2, 0.1]) + 0.3
    expect = a * (np.maximum(nf[:, 1], 0.0)[:, None] + np.asarray(env)[None, :])
    got = imgs[0].reshape(-1, 3)
    assert np.allclose(got[hit], expect[hit], rtol=3e-6, atol=1e-7)
    assert np.array_equal(got[~hit], np.broadcast_to(np.float32(env), got[~hit].shape))
    assert ns == int((hit & (nf[:, 1] > 0)).sum())


def test_paths_are_deterministic_and_stride_consistent(oracle, small_scene):
    S, _ = small_scene
    views = _views(oracle)
    a, ida, nb, ns, _ = S.render_batch(views, bounces=3, seed=5, env=(0.3, 0.3, 0.3))
    b, _, nb2, ns2, _ = S.render_batch(views, bounces=3, seed=5, env=(0.3, 0.3, 0.3))
    c, _, _, _, _ = S.render_batch(views, bounces=3, seed=6, env=(0.3, 0.3, 0.3))
    assert all(np.array_equal(x, y) for x, y in zip(a, b)) and (nb, ns) == (nb2, ns2)
    assert nb > 0
    assert any(not np.array_equal(x, y) for x, y in zip(a, c))  # bounce light depends on the seed
    # a path's value does not depend on which other paths are computed
    s, _, _, _, _ = S.render_batch(views, bounces=3, seed=5, env=(0.3, 0.3, 0.3), path_stride=7)
    flat_a = np.concatenate([x.reshape(-1, 3) for x in a])
    flat_s = np.concatenate([x.reshape(-1, 3) for x in s])
    assert np.array_equal(flat_a[::7], flat_s[::7])
    # more bounces only add light (every term is non-negative)
    d0, _, _, _, _ = S.render_batch(views, bounces=0, seed=5, env=(0.3, 0.3, 0.3))
    flat_0 = np.concatenate([x.reshape(-1, 3) for x in d0])
    assert np.all(flat_a >= flat_0 - 1e-6)


def _segment_line_distance(x, y, a, e):
    """Distance between segments x->y (rows) and the line a + s e (|e| = 1)."""
    d = y - x
    w0 = x - a
    aa = (d * d).sum(1)
    b = d @ e
    dd = (d * w0).sum(1)
    ee = w0 @ e
    den = np.maximum(aa - b * b, 1e-30)
    sc = np.clip((b * ee - dd) / den, 0.0, 1.0)
    tc = b * sc + ee
    return np.linalg.norm(w0 + sc[:, None] * d - tc[:, None] * e[None, :], axis=1)


def test_area_light_behind_an_occluder(oracle):
    """Area-light visibility of a lit fibre under a thick, long occluding fibre: light sample k
    of pixel p (the point of the light that SplitMix64, re-implemented here, fixes for
    (seed, p, k)) is blocked exactly when the segment from the hit point to it passes within the
    occluder's radius of its axis (the phantom intersector reports entering roots only, so a
    fibre does not shadow rays leaving its own surface, tangent rays aside). Each pixel's
    visibility is the mean over its samples and its colour Shade(n, visibility)."""
    lit = np.array([[-30.0, 2.0, 0.0], [30.0, -2.0, 1.0]])
    occ = np.array([[-200.0, 14.0, -4.0], [200.0, 14.0, -4.0]])
    curves = np.array([[[*(p0 + (p1 - p0) * k / 3), 0.0] for k in range(4)] for p0, p1 in (lit, occ)],
                      np.float32).reshape(2, 16)
    S = oracle.Scene()
    S.add(oracle.MODE_PHANTOM, curves[:1], per_point=False, radius=1.5)
    S.add(oracle.MODE_PHANTOM, curves[1:], per_point=False, radius=4.0)
    W, H = 96, 64
    vi, pi = oracle.camera((0.0, -6.0, 40.0), (0.0, 0.0, 0.0), aspect=W / H)
    light = np.array([[5.0, 40.0, -10.0], [30.0, 0.0, 0.0], [0.0, 0.0, 25.0]], np.float32)
    seed, ns = 99, 3
    img, ids, nrays, _ = S.render(vi, pi, W, H, shadows=1, area_light=light, light_samples=ns, light_seed=seed)
    rays = oracle.raygen(vi, pi, W, H)
    hits = S.trace(rays)
    hit = hits["prim"] != 0xFFFFFFFF
    on_lit = hit & (hits["inst"] == 0)
    assert on_lit.sum() > 300 and nrays == W * H + ns * int(hit.sum())
    x = rays[:, 0:3].astype(np.float64) + hits["t"][:, None] * rays[:, 4:7].astype(np.float64)
    n = hits["n"].astype(np.float64)
    e = np.array([1.0, 0.0, 0.0])
    vis = np.zeros(W * H)
    edge = np.zeros(W * H, bool)
    for k in range(ns):
        a = np.array([2 * _rand_py(seed, p, k, 0) - 1 for p in range(W * H)])
        b = np.array([2 * _rand_py(seed, p, k, 1) - 1 for p in range(W * H)])
        y = light[0][None, :] + a[:, None] * light[1][None, :] + b[:, None] * light[2][None, :]
        dist = _segment_line_distance(x, y, occ[0], e)
        vis += dist >= 4.0
        d = (y - x) / np.linalg.norm(y - x, axis=1, keepdims=True)
        # undecided by the closed form: the occluder's silhouette, and rays leaving almost
        # tangent to the own fibre (its surface may be re-entered within rounding)
        edge |= (np.abs(dist - 4.0) < 2e-3) | (np.abs((n * d).sum(1)) < 1e-2)
    vis /= ns
    expect = (np.abs(n[:, 1]) * vis)[:, None] * np.array([0.4, 0.2, 0.1]) + 0.3
    sel = on_lit & ~edge
    assert sel.sum() > 0.9 * on_lit.sum()
    assert np.any(vis[sel] == 0) and np.any(vis[sel] == 1) and np.any((vis[sel] > 0) & (vis[sel] < 1))
    got = img.reshape(-1, 3)
    assert np.allclose(got[sel], expect[sel], rtol=2e-6, atol=1e-6)
