"""Host-side checks that need no GPU: the C-ABI library (symbols, argument checks, camera),
the oracle's BVH2 traversal against brute force, and the candidate count that the roofline's
algorithmic bytes are built on."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle_lib as O
import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "vkhrt.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = re.sub(r"//[^\n]*", "", txt)
    return sorted(set(re.findall(r"\b(vkhrt_[a-z0-9_]+)\s*\(", txt)))


@pytest.fixture(scope="module")
def vk():
    import vkhrt_amd
    return vkhrt_amd


# ------------------------------------------------------------------ C-ABI library
def test_library_exports_every_declared_symbol(vk):
    L = ctypes.CDLL(vk.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(vk.EXPORTS) == syms
    assert vk.lib().vkhrt_abi_version() == vk.ABI_VERSION
    hdr = open(os.path.join(ROOT, "include", "vkhrt.h")).read()
    assert int(re.search(r"#define VKHRT_ABI_VERSION (\d+)", hdr).group(1)) == vk.ABI_VERSION
    assert ctypes.sizeof(vk.Tuning) == 13 * 4  # vkhrt_tuning: thirteen 4-byte fields


def test_null_handles_are_rejected(vk):
    L = vk.lib()
    bad = -1  # VKHRT_ERROR_INVALID_ARGUMENT
    assert L.vkhrt_synchronize(None, None) == bad
    assert L.vkhrt_set_option(None, 1, 1) == bad
    assert L.vkhrt_trace_rays(None, None, None, 0, None, 0, None) == bad
    assert L.vkhrt_render(None, None, None, None, None) == bad
    assert L.vkhrt_blas_build(None, None, None, None) == bad
    assert L.vkhrt_tlas_build(None, None, 0, None, None) == bad
    assert L.vkhrt_context_create(0, None) == bad
    L.vkhrt_context_destroy(None)  # no-op


def test_context_without_device(vk):
    h = ctypes.c_void_p()
    rc = vk.lib().vkhrt_context_create(0, ctypes.byref(h))
    if rc == 0:
        vk.lib().vkhrt_context_destroy(h)
        pytest.skip("a device is present")
    assert rc == -6  # VKHRT_ERROR_NO_DEVICE: no silent CPU path
    assert not h.value
    with pytest.raises(vk.VkhrtError):
        vk.Context(0)


def test_camera_look_at_matches_oracle(vk):
    for eye, center, aspect in [((0, 60, 300), (0, 40, 0), 1920 / 1080), ((0, 50, 400), (0, 30, 0), 1.0),
                                ((-120.5, 3, 17), (4, 5, -6), 0.75)]:
        cam = vk.camera_look_at(eye, center, aspect=aspect)
        vi, pi = O.camera(eye, center, aspect=aspect)
        assert np.array_equal(np.ctypeslib.as_array(cam.view_inverse), vi)
        assert np.array_equal(np.ctypeslib.as_array(cam.proj_inverse), pi)
    with pytest.raises(vk.VkhrtError):
        vk.camera_look_at((0, 0, 0), (0, 0, -1), fov_deg=0.0)


def _closed_form_dirs(eye, center, up, tan_half, aspect, W, H):
    """Pixel ray directions from glm's closed forms, in float64: lookAt gives the right-handed
    basis f = normalize(center - eye), s = normalize(f x up), u = s x f (fly_camera.cpp:27);
    perspectiveRH_ZO with [1][1] *= -1 (fly_camera.cpp:31-33) maps the NDC point (dx, dy, 1, 1)
    of ray_gen.rgen:18-24 back to the view-space point (dx*aspect*tan, -dy*tan, -1, 1/far).
    The world direction is therefore dx*aspect*tan*s - dy*tan*u + f, normalized."""
    eye, center, up = (np.asarray(v, np.float64) for v in (eye, center, up))
    f = center - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    x, y = np.meshgrid(np.arange(W), np.arange(H))
    dx = (x + 0.5) / W * 2.0 - 1.0
    dy = (y + 0.5) / H * 2.0 - 1.0
    d = (dx * aspect * tan_half)[..., None] * s - (dy * tan_half)[..., None] * u + f
    return d / np.linalg.norm(d, axis=-1, keepdims=True), s, u


def _raygen_f64(vi, pi, W, H):
    """ray_gen.rgen:18-24 evaluated in float64 on given inverse matrices (column-major)."""
    VI = np.asarray(vi, np.float64).reshape(4, 4).T
    PI = np.asarray(pi, np.float64).reshape(4, 4).T
    x, y = np.meshgrid(np.arange(W), np.arange(H))
    ndc = np.stack([(x + 0.5) / W * 2 - 1, (y + 0.5) / H * 2 - 1, np.ones_like(x, float), np.ones_like(x, float)], -1)
    tgt = ndc @ PI.T
    t3 = tgt[..., :3] / np.linalg.norm(tgt[..., :3], axis=-1, keepdims=True)
    d = np.concatenate([t3, np.zeros(t3.shape[:-1] + (1,))], -1) @ VI.T
    o = VI @ np.array([0, 0, 0, 1.0])
    return o[:3], d[..., :3]


@pytest.mark.parametrize("eye,center,aspect", [((0, 60, 300), (0, 40, 0), 1920 / 1080),
                                               ((-120.5, 3, 17), (4, 5, -6), 0.75),
                                               ((10, -40, 5), (0, 0, 0), 1.0)])
def test_camera_orientation_closed_form(vk, eye, center, aspect):
    """Which way the image points, from glm's closed forms rather than from the oracle: row 0
    looks toward +up (the Vulkan Y flip of fly_camera.cpp:33), column 0 toward -right (s = f x up),
    and every pixel's direction equals the closed form within FP32 rounding. Checked for the
    oracle's matrices and rays and for the library's vkhrt_camera_look_at. A mirrored image (flip
    dropped, or the cross product reversed) moves edge rays by about the field of view."""
    W, H, tan_half, up = 48, 30, 0.5773502691896257, (0.0, 1.0, 0.0)
    ref, s, u = _closed_form_dirs(eye, center, up, tan_half, aspect, W, H)
    vi_o, pi_o = O.camera(eye, center, up=up, tan_half_fovy=tan_half, aspect=aspect)
    cam = vk.camera_look_at(eye, center, up=up, fov_deg=60.0, aspect=aspect)
    vi_l, pi_l = np.ctypeslib.as_array(cam.view_inverse), np.ctypeslib.as_array(cam.proj_inverse)
    rays = O.raygen(vi_o, pi_o, W, H).reshape(H, W, 8)
    for name, o, d in [("oracle raygen", rays[0, 0, :3], rays[..., 4:7].astype(np.float64)),
                       ("oracle matrices", *_raygen_f64(vi_o, pi_o, W, H)),
                       ("library matrices", *_raygen_f64(vi_l, pi_l, W, H))]:
        d = d / np.linalg.norm(d, axis=-1, keepdims=True)
        assert np.allclose(o, eye, atol=1e-3), name
        err = np.degrees(np.arccos(np.clip((d * ref).sum(-1), -1, 1)))
        assert err.max() < 1e-3, (name, err.max())  # FP32 matrices: ~1e-5 degrees
        # image orientation, stated as signs
        assert (d[0] @ u > 0).all() and (d[-1] @ u < 0).all(), name          # row 0 is the top
        assert (d[:, 0] @ s < 0).all() and (d[:, -1] @ s > 0).all(), name    # column 0 is the left
    # the mirrored alternatives are far from every pixel's ray
    mirrored_y, _, _ = _closed_form_dirs(eye, center, up, tan_half, aspect, W, H)
    mirrored_y = mirrored_y[::-1]
    d = rays[..., 4:7].astype(np.float64)
    assert np.degrees(np.arccos(np.clip((d[0] * mirrored_y[0]).sum(-1), -1, 1))).min() > 30.0


# ------------------------------------------------------------------ oracle traversal
def _scene(kind):
    curves = scenes.c0_curves(300, 16, seed=11)
    S = O.Scene()
    if kind == "phantom_const":
        S.add(O.MODE_PHANTOM, curves)
    elif kind == "phantom_point":
        S.add(O.MODE_PHANTOM, curves, per_point=True)
    elif kind == "lss":
        S.add(O.MODE_LSS, O.generate_lss(scenes.lines_from_curves(curves)))
    elif kind == "dots":
        S.add(O.MODE_DOTS, O.generate_dots(scenes.lines_from_curves(curves)))
    else:  # several instances with transforms, mixed modes
        xf = np.array([[0.0, 0, 1, 5], [0, 1, 0, -3], [-1, 0, 0, 2]], np.float32)
        S.add(O.MODE_PHANTOM, curves, per_point=True)
        S.add(O.MODE_LSS, O.generate_lss(scenes.lines_from_curves(curves[:2000])), transform=xf)
        S.add(O.MODE_DOTS, O.generate_dots(scenes.lines_from_curves(curves[2000:3000])),
              transform=np.array([[1.0, 0, 0, 0.7], [0, 1, 0, 0.1], [0, 0, 1, -0.4]], np.float32))
    return curves, S


@pytest.mark.parametrize("kind", ["phantom_const", "phantom_point", "lss", "dots", "mixed"])
def test_bvh2_equals_brute_force(kind):
    curves, S = _scene(kind)
    rays = scenes.targeted_rays(curves, 2500, seed=3)
    a = S.trace(rays)
    b = S.trace(rays, O.F_BRUTE)
    hit = a["prim"] != 0xFFFFFFFF
    assert hit.mean() > 0.1
    assert np.array_equal(a["prim"], b["prim"]) and np.array_equal(a["inst"], b["inst"])
    assert np.array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32))
    assert np.array_equal(a["n"].view(np.uint32), b["n"].view(np.uint32))
    # any-hit: same hit/miss decision, and the reported primitive is a hit within range
    c = S.trace(rays, O.F_ANY_HIT)
    assert np.array_equal(c["prim"] != 0xFFFFFFFF, hit)
    assert (c["t"][hit] >= rays[hit, 3]).all() and (c["t"][hit] <= rays[hit, 7]).all()


def test_trace_ray_interval_and_empty():
    curves, S = _scene("phantom_const")
    rays = scenes.targeted_rays(curves, 500, seed=4)
    h = S.trace(rays)
    hit = h["prim"] != 0xFFFFFFFF
    # tmax just below the closest hit: that primitive can no longer be reported
    r2 = rays.copy()
    r2[hit, 7] = np.nextafter(h["t"][hit], np.float32(0))
    h2 = S.trace(r2)
    assert (h2["t"][hit][h2["prim"][hit] != 0xFFFFFFFF] < h["t"][hit][h2["prim"][hit] != 0xFFFFFFFF]).all()
    # empty interval: no hits
    r3 = rays.copy()
    r3[:, 7] = r3[:, 3] * 0.5
    assert (S.trace(r3)["prim"] == 0xFFFFFFFF).all()
    assert S.trace(np.zeros((0, 8), np.float32)).shape == (0,)
    # axis-aligned directions (zero components) trace like any other ray
    r4 = rays.copy()
    r4[:, 4:7] = 0
    r4[:, 5] = -1
    r4[:, 1] += 200
    a, b = S.trace(r4), S.trace(r4, O.F_BRUTE)
    assert np.array_equal(a["prim"], b["prim"])


def test_curve_aabbs_contain_the_tube():
    curves = scenes.c0_curves(50, 16, seed=8)
    for per_point in (False, True):
        boxes = O.curve_aabbs(curves, per_point=per_point)  # (n, 6): min xyz, max xyz
        ts = np.linspace(0, 1, 65)
        u = 1 - ts
        w = np.stack([u ** 3, 3 * u * u * ts, 3 * u * ts * ts, ts ** 3], 1)
        cp = curves.reshape(-1, 4, 4).astype(np.float64)
        P = np.einsum("sk,nkj->nsj", w, cp[:, :, :3])
        R = np.einsum("sk,nk->ns", w, cp[:, :, 3]) if per_point else np.full(P.shape[:2], 0.02)
        lo = (P - R[..., None]).min(1)
        hi = (P + R[..., None]).max(1)
        assert (boxes[:, :3] <= lo + 1e-5).all() and (boxes[:, 3:6] >= hi - 1e-5).all()


def test_candidate_count_matches_slab_brute_force():
    """N*_prim: segments whose exact (unpadded) box overlaps [tmin, t_closest]."""
    curves, S = _scene("phantom_point")
    rays = scenes.targeted_rays(curves, 400, seed=6)
    h = S.trace(rays)
    cnt = S.count_candidates(rays, h)
    cp = curves.reshape(-1, 4, 4)
    r = cp[:, :, 3].max(1)
    lo = cp[:, :, :3].min(1) - r[:, None]
    hi = cp[:, :, :3].max(1) + r[:, None]
    ref = []
    for i, ray in enumerate(rays.astype(np.float64)):
        o, d = ray[:3], ray[4:7]
        tend = h["t"][i] if h["prim"][i] != 0xFFFFFFFF else ray[7]
        with np.errstate(divide="ignore", invalid="ignore"):
            inv = 1.0 / d
            t0 = (lo - o) * inv
            t1 = (hi - o) * inv
        tn = np.nanmax(np.minimum(t0, t1), 1)
        tf = np.nanmin(np.maximum(t0, t1), 1)
        ref.append(int(((np.maximum(tn, ray[3]) <= np.minimum(tf, tend))).sum()))
    ref = np.array(ref)
    # boxes that touch the interval only at its end point can round either way
    assert np.abs(cnt.astype(np.int64) - ref).max() <= 1
    assert (cnt == ref).mean() > 0.98


# ------------------------------------------------------------------ C++ mirror of the reference classes
CPP_TEST = os.path.join(ROOT, "tests", "cpp", "host_api_test")


def cpp_test_binary():
    import subprocess
    if not os.path.exists(CPP_TEST):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")], check=True)
    return CPP_TEST


def test_cpp_mirror_without_device():
    """include/vkhrt.hpp: Context -> NO_DEVICE, BLASType/geometry pairing errors, camera."""
    import subprocess
    r = subprocess.run([cpp_test_binary(), "nodevice"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert lines[-1] == "ok"
    cam = np.array([float.fromhex(x) for x in lines[0].split()], np.float32)
    vi, pi = O.camera((0, 60, 300), (0, 40, 0), tan_half_fovy=np.float32(0.57735026), aspect=16.0 / 9.0)
    assert np.array_equal(cam[:16], vi) and np.array_equal(cam[16:], pi)
