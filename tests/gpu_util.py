"""Helpers for the GPU tests: device buffers through torch (plumbing only) and scene setup
through the vkhrt C-ABI."""
import numpy as np
import torch

import vkhrt_amd as V

HIT_DTYPE = np.dtype([("t", "<f4"), ("prim", "<u4"), ("inst", "<u4"), ("u", "<f4"),
                      ("n", "<f4", (3,)), ("v", "<f4")])


def dev(a):
    """numpy -> contiguous cuda tensor (float32/uint32 kept bitwise)."""
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        t = torch.from_numpy(a.view(np.int32).copy())
    else:
        t = torch.from_numpy(a.copy())
    return t.cuda()


def host(t, dtype=np.float32):
    return t.cpu().numpy().view(dtype)


def stream_ptr():
    return torch.cuda.current_stream().cuda_stream


def build_blas(ctx, mode, prims, per_point=False, radius=0.02, transform=None, leaf_size=0):
    prims = np.ascontiguousarray(prims, dtype=np.float32)
    d = dev(prims)
    flags = V.BLAS_PER_POINT_RADIUS if per_point else 0
    b = V.BottomLevelAccelerationStructure(ctx, mode, d.data_ptr(), prims.shape[0], radius=radius,
                                           flags=flags, transform=transform, leaf_size=leaf_size,
                                           stream=stream_ptr())
    torch.cuda.synchronize()
    return b


def trace(ctx, tlas, rays, flags=0):
    rays = np.ascontiguousarray(rays, dtype=np.float32)
    r = dev(rays)
    h = torch.zeros(rays.shape[0] * 8, dtype=torch.float32, device="cuda")
    ctx.trace_rays(tlas, r.data_ptr(), rays.shape[0], h.data_ptr(), flags, stream_ptr())
    torch.cuda.synchronize()
    ctx.synchronize(stream_ptr())
    return h.cpu().numpy().view(HIT_DTYPE)


def render(ctx, tlas, cam, W, H, flags=V.RENDER_SHADOWS):
    img = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
    ids = torch.zeros(W * H, dtype=torch.int32, device="cuda")
    ctx.render(tlas, cam, W, H, img.data_ptr(), flags=flags, prim_ids_ptr=ids.data_ptr(),
               stream=stream_ptr())
    torch.cuda.synchronize()
    ctx.synchronize(stream_ptr())
    return img.cpu().numpy().reshape(H, W, 4), ids.cpu().numpy().view(np.uint32).reshape(H, W)


def psnr(a, b):
    mse = float(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2))
    return float("inf") if mse == 0 else 10.0 * np.log10(1.0 / mse)


def cam_arrays(cam):
    return np.array(cam.view_inverse, np.float32), np.array(cam.proj_inverse, np.float32)


def compare_hits(h, ref, what):
    """Hit records agree bit for bit: hit/miss, (instance, primitive), t, u and normal."""
    miss_g = h["prim"] == 0xFFFFFFFF
    miss_r = ref["prim"] == 0xFFFFFFFF
    assert (miss_g == miss_r).all(), f"{what}: hit/miss differs on {(miss_g != miss_r).sum()} rays"
    agree = (h["prim"] == ref["prim"]) & (h["inst"] == ref["inst"])
    assert agree.all(), f"{what}: (instance, primitive) differs on {(~agree).sum()} rays"
    hit = ~miss_r
    for k in ("t", "u", "n"):
        assert np.array_equal(h[k][hit].view(np.uint32), ref[k][hit].view(np.uint32)), f"{what}: {k} differs"
    return int(hit.sum())
