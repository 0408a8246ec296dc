"""Seeded test scenes and ray sets (small enough for the oracle to finish in seconds)."""
import numpy as np

import oracle_lib as O

C0_SEED = 20240601


def c0_curves(nstrands=1000, nseg=16, seed=C0_SEED):
    """Config 0 hair: 1,000 strands x 16 cubic segments (49 control points per strand)."""
    return O.hair_curves(seed, nstrands, nseg)


def dense_ball_curves(nstrands=1200, nseg=8, seed=77, sphere_radius=12.0):
    """The seeded generator on a small sphere: hair dense enough that small views (tens of
    pixels) see hundreds of hits and bounce rays hit other fibres."""
    v = O.generate_strands(seed, nstrands, nseg, sphere_radius=sphere_radius)
    return O.generate_curves(O.strands_to_lines(v))


DENSE_VIEWS = [((0.0, 20.0, 45.0), (0.0, 8.0, 0.0), 0.577, (48, 40)),
               ((30.0, 25.0, 30.0), (0.0, 8.0, 0.0), 0.4, (40, 32)),
               ((-35.0, 10.0, 20.0), (5.0, 6.0, 0.0), 0.5, (56, 24))]


def c0_camera(oracle_cam=True, W=256, H=256):
    return O.camera((0.0, 50.0, 400.0), (0.0, 30.0, 0.0), aspect=W / H)


def targeted_rays(curves, n, seed, spread=0.06, dist=(5.0, 400.0)):
    """Rays aimed at points near random curve samples from random directions, so most rays
    hit or graze hair. Returns (n, 8) float32 rays with tmin 0.001, tmax 10000."""
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, curves.shape[0], n)
    t = rng.random(n).astype(np.float32)
    c = curves[idx].reshape(n, 4, 4)[:, :, :3].astype(np.float64)
    u = 1.0 - t
    w = np.stack([u ** 3, 3 * u * u * t, 3 * u * t * t, t ** 3], axis=1)
    p = np.einsum("nk,nkj->nj", w, c)
    p += rng.normal(size=(n, 3)) * spread
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    L = rng.uniform(dist[0], dist[1], n)[:, None]
    o = p - d * L
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = 0.001
    rays[:, 4:7] = d
    rays[:, 7] = 10000.0
    # re-normalize in float32 (unit directions as the reference's rays)
    dd = rays[:, 4:7].astype(np.float32)
    dd = dd / np.sqrt((dd * dd).sum(1, keepdims=True))
    rays[:, 4:7] = dd
    return rays


def lines_from_curves(curves, pieces=4):
    return O.curves_to_lines(curves, pieces)


def grazing_rays(curves, n, seed, per_point=True, eps=2e-3, radius=0.02):
    """Rays tangent to the hair tube: through curve(t0) + r(t0)(1 + e) N with direction
    perpendicular to N, e uniform in [-eps, eps], so the cone discriminant is near zero
    (the FP64 polish trigger). Returns (n, 8) float32 rays with unit directions."""
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, curves.shape[0], n)
    t = rng.uniform(0.05, 0.95, n)
    c = curves[idx].reshape(n, 4, 4).astype(np.float64)
    u = 1.0 - t
    w = np.stack([u ** 3, 3 * u * u * t, 3 * u * t * t, t ** 3], axis=1)
    dw = np.stack([-3 * u * u, 3 * (3 * t * t - 4 * t + 1), 3 * (2 - 3 * t) * t, 3 * t * t], axis=1)
    p = np.einsum("nk,nkj->nj", w, c[:, :, :3])
    tan = np.einsum("nk,nkj->nj", dw, c[:, :, :3])
    tan /= np.linalg.norm(tan, axis=1, keepdims=True)
    r = np.einsum("nk,nk->n", w, c[:, :, 3]) if per_point else np.full(n, radius)
    a = rng.normal(size=(n, 3))
    nrm = a - (a * tan).sum(1, keepdims=True) * tan
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    b = np.cross(tan, nrm)
    ang = rng.uniform(-1.2, 1.2, n)[:, None]
    d = np.cos(ang) * b + np.sin(ang) * tan
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    e = rng.uniform(-eps, eps, n)[:, None]
    q = p + nrm * (r[:, None] * (1.0 + e))
    L = rng.uniform(20.0, 300.0, n)[:, None]
    o = q - d * L
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = 0.001
    dd = d.astype(np.float32)
    dd = dd / np.sqrt((dd * dd).sum(1, keepdims=True))
    rays[:, 4:7] = dd
    rays[:, 7] = 10000.0
    return rays


def uv_sphere(radius=10.0, nlat=24, nlon=48, center=(0.0, 0.0, 0.0)):
    """Latitude-longitude sphere mesh: vertices (position, radial normal, texCoord) as 8 floats,
    and u32 triangle indices (the poles are fans)."""
    c = np.asarray(center, np.float64)
    verts = []
    for i in range(nlat + 1):
        th = np.pi * i / nlat
        for j in range(nlon + 1):
            ph = 2 * np.pi * j / nlon
            n = np.array([np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph)])
            verts.append([*(c + radius * n), *n, j / nlon, i / nlat])
    idx = []
    for i in range(nlat):
        for j in range(nlon):
            a, b = i * (nlon + 1) + j, (i + 1) * (nlon + 1) + j
            if i != 0:
                idx.append([a, b, a + 1])
            if i != nlat - 1:
                idx.append([a + 1, b, b + 1])
    return np.asarray(verts, np.float32), np.asarray(idx, np.uint32)
