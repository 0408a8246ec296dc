"""Pins of the oracle's miss shader (miss.rmiss:17-38): a constant map gives the tone-mapped
constant in every direction; gradient maps, inverted through the tone map, return the texel
coordinate that DirectionToUV fixes (computed here with numpy's arctan2 / arcsin), away from
the seam and the poles where the repeat addressing blends the opposite edge."""
import numpy as np


def _render_empty(oracle, env, W=64, H=48, eye=(0.0, 0.0, 0.0), tgt=(0.3, 0.2, -1.0), fov=1.2):
    S = oracle.Scene()
    vi, pi = oracle.camera(eye, tgt, tan_half_fovy=fov, aspect=W / H)
    img, ids, _, _ = S.render(vi, pi, W, H, shadows=1, env=env)
    rays = oracle.raygen(vi, pi, W, H)
    return img.reshape(-1, 3).astype(np.float64), rays[:, 4:7].astype(np.float64), ids


def test_constant_environment(oracle):
    env = np.zeros((8, 16, 4), np.float32)
    env[..., 0], env[..., 1], env[..., 2] = 0.5, 1.5, 3.0
    img, _, ids = _render_empty(oracle, env)
    assert np.all(ids == 0xFFFFFFFF)
    expect = (1 - np.exp(-np.array([0.5, 1.5, 3.0]))) ** (1 / 2.2)
    assert np.abs(img - expect).max() < 2e-6


def test_gradient_environment_fixes_direction_to_uv(oracle):
    ew, eh = 64, 32
    x, y = np.meshgrid(np.arange(ew), np.arange(eh))
    env = np.zeros((eh, ew, 4), np.float32)
    env[..., 0] = 2.0 * x / ew  # linear in the column, so a bilinear lookup returns 2 (u ew - 0.5) / ew
    env[..., 1] = 2.0 * y / eh
    for eye, tgt in [((0.0, 0.0, 0.0), (0.3, 0.2, -1.0)), ((0.0, 0.0, 0.0), (-1.0, -0.4, 0.2))]:
        img, d, _ = _render_empty(oracle, env, eye=eye, tgt=tgt)
        val = -np.log(1 - img[:, :2] ** 2.2)  # inverse tone map
        v = -d / np.linalg.norm(d, axis=1, keepdims=True)
        u = np.arctan2(v[:, 0], -v[:, 2]) / (2 * np.pi) + 0.5
        t = np.arcsin(v[:, 1]) / np.pi + 0.5
        ex, ey = 2 * (u * ew - 0.5) / ew, 2 * (t * eh - 0.5) / eh
        inside = (u * ew > 1.0) & (u * ew < ew - 1.0) & (t * eh > 1.0) & (t * eh < eh - 1.0)
        assert inside.mean() > 0.8
        assert np.abs(val[inside, 0] - ex[inside]).max() < 2e-4
        assert np.abs(val[inside, 1] - ey[inside]).max() < 2e-4


def test_hits_are_not_environment(oracle):
    import scenes
    curves = scenes.dense_ball_curves(nstrands=600)
    S = oracle.Scene()
    S.add(oracle.MODE_PHANTOM, curves, per_point=True)
    vi, pi = oracle.camera((0.0, 20.0, 45.0), (0.0, 8.0, 0.0), aspect=1.0)
    env = np.full((4, 8, 4), 2.0, np.float32)
    a, ida, _, _ = S.render(vi, pi, 48, 48, shadows=1, env=env)
    b, idb, _, _ = S.render(vi, pi, 48, 48, shadows=1)
    hit = ida != 0xFFFFFFFF
    assert hit.sum() > 100 and np.array_equal(ida, idb)
    assert np.array_equal(a[hit], b[hit]) and np.all(b[~hit] == 0)
    assert np.allclose(a[~hit], (1 - np.exp(-2.0)) ** (1 / 2.2), atol=2e-6)
