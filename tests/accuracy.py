"""Hit accuracy against an FP64 evaluation of the same intersector, per geometry mode.

Test infrastructure (also used by bench.py's CPU-baseline leg): for each hit record the FP64
reference is the oracle's intersector in double on the same ray and primitive
 - phantom: Prhi in double (hair_intersection.rint:15-130, oracle_math.hpp);
 - LSS: the round-cone capsule test in double (triangle_closest_hit.rchit:43-58);
 - DOTS: the ray/triangle plane intersection and geometric normal in float64 numpy, on the
   triangle the hit names (the FP32 vertices, widened exactly).
Reported: relative hit-distance error and shading-normal angle (radians, sign-free), as
median / p99 / max over the hits evaluated.
"""
import math

import numpy as np

import oracle_lib as O


def _stats(rel, ang):
    rel, ang = np.asarray(rel, np.float64), np.asarray(ang, np.float64)
    if rel.size == 0:
        return {"hits": 0}
    return {"hits": int(rel.size),
            "t_rel_median": float(np.median(rel)), "t_rel_p99": float(np.percentile(rel, 99)),
            "t_rel_max": float(rel.max()),
            "normal_rad_median": float(np.median(ang)), "normal_rad_p99": float(np.percentile(ang, 99)),
            "normal_rad_max": float(ang.max())}


def _angle(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    c = abs(float(a @ b)) / (np.linalg.norm(a) * np.linalg.norm(b))
    return math.acos(min(1.0, c))


def _hit_rows(hits, limit):
    idx = np.nonzero(hits["prim"] != 0xFFFFFFFF)[0]
    if limit and idx.size > limit:
        idx = idx[np.linspace(0, idx.size - 1, limit).astype(np.int64)]
    return idx


def phantom(rays, hits, curves, per_point=True, limit=1500):
    """hits: HIT_DTYPE records whose prim indexes `curves` (one instance)."""
    rel, ang = [], []
    for i in _hit_rows(hits, limit):
        r64 = O.prhi_f64(rays[i], curves[hits["prim"][i]], per_point=per_point)
        if r64[0] <= 0:
            continue
        rel.append(abs(float(hits["t"][i]) - r64[0]) / r64[0])
        ang.append(_angle(hits["n"][i], r64[5:8]))
    return _stats(rel, ang)


def lss(rays, hits, caps, limit=1500):
    """caps: (n, 8) capsule records; a capsule whose start equals the previous end is chained
    (its start cap belongs to the previous capsule), as the oracle and the BLAS build decide."""
    L = O.lib()
    rel, ang = [], []
    out = np.zeros(6, np.float64)
    for i in _hit_rows(hits, limit):
        p = int(hits["prim"][i])
        cap_a = not (p > 0 and np.array_equal(caps[p - 1, 4:7], caps[p, 0:3]))
        L.orc_lss(O.ptr(O.f32(rays[i])), O.ptr(O.f32(caps[p])), int(cap_a), 1, out.ctypes.data_as(O.ctypes.c_void_p))
        if out[0] == 0 or out[1] <= 0:
            continue
        rel.append(abs(float(hits["t"][i]) - out[1]) / out[1])
        ang.append(_angle(hits["n"][i], out[3:6]))
    return _stats(rel, ang)


def dots(rays, hits, pieces, limit=1500):
    """pieces: (n, 16) DOTS pieces; primitive id = 4 * piece + 2 * quad + triangle."""
    rel, ang = [], []
    for i in _hit_rows(hits, limit):
        p = int(hits["prim"][i])
        tri = O.dots_triangles(pieces[p // 4:p // 4 + 1])[p % 4].astype(np.float64)
        v0, v1, v2 = tri
        n = np.cross(v1 - v0, v2 - v0)
        o = rays[i, :3].astype(np.float64)
        d = rays[i, 4:7].astype(np.float64)
        den = float(n @ d)
        if den == 0.0:
            continue
        t = float(n @ (v0 - o)) / den
        if t <= 0:
            continue
        rel.append(abs(float(hits["t"][i]) - t) / t)
        ang.append(_angle(hits["n"][i], n))
    return _stats(rel, ang)
