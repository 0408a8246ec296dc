"""Neural-shadow frames of the 16M-segment area-light workload, for timing and counter runs.

    python -m vkhrt_amd.neural_frames [--threshold 0.2499] [--frames 3] [--save-params P.npy | --load-params P.npy]

Builds the 500,000-strand x 32-segment hair (seed 104), the 1920x1080 camera and the
parallelogram light with two samples per hit, trains the shadow MLP on 2^24 rays drawn on the
hair surface (24 epochs of 65,536 samples, as the bench does) or loads its parameters from a
file written by an earlier run, and renders `--frames` neural frames at each threshold. The
parameters of one training can be reused by later runs (a profiler run, for one), so that those
see the same network and the same exact-fallback rays. Prints one line per threshold with the
frame's query and fallback counts and the image's CRC-32 (the same network gives the same
image under every traversal setting).
"""
import argparse
import os
import sys
import zlib

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threshold", type=float, action="append", help="variance threshold(s); default 0.2499")
    ap.add_argument("--frames", type=int, default=3)
    ap.add_argument("--strands", type=int, default=500000)
    ap.add_argument("--save-params", default=None, help="write the trained parameters (numpy .npy)")
    ap.add_argument("--load-params", default=None, help="read the parameters instead of training")
    ap.add_argument("--wave-limit", type=int, default=None,
                    help="fallback queues up to this long are walked by lane groups (VKHRT_OPTION_SHADOW_WAVE_LIMIT)")
    ap.add_argument("--group-lanes", type=int, default=None, help="lanes per ray of that walk: 8, 16, 32 or 64")
    args = ap.parse_args()
    thresholds = args.threshold or [0.2499]

    import numpy as np
    import torch
    import vkhrt_amd as V
    assert torch.cuda.is_available(), "needs a GPU"
    print("[neural_frames] scene", flush=True)
    ctx = V.Context(0)
    sp = torch.cuda.current_stream().cuda_stream
    ns, nseg = args.strands, 32
    n = ns * nseg
    verts = torch.empty(ns * (nseg + 1) * 4, dtype=torch.float32, device="cuda")
    lines = torch.empty(n * 8, dtype=torch.float32, device="cuda")
    curves = torch.empty(n * 16, dtype=torch.float32, device="cuda")
    ctx.generate_strands(104, ns, nseg, verts.data_ptr(), stream=sp)
    ctx.generate_lines(verts.data_ptr(), ns, nseg, lines.data_ptr(), stream=sp)
    ctx.generate_curves(lines.data_ptr(), n, curves.data_ptr(), stream=sp)
    del verts, lines
    blas = V.BottomLevelAccelerationStructure(ctx, V.GEOMETRY_PHANTOM_CURVES, curves.data_ptr(), n,
                                              flags=V.BLAS_PER_POINT_RADIUS, stream=sp)
    tlas = V.TopLevelAccelerationStructure(ctx, [blas], stream=sp)
    torch.cuda.synchronize()
    W, H = 1920, 1080
    cam = V.camera_look_at((0.0, 60.0, 300.0), (0.0, 40.0, 0.0), fov_deg=60.0, aspect=W / H)
    image = torch.empty(W * H, dtype=torch.int32, device="cuda")
    light = V.area_light((0.0, 260.0, 0.0), (100.0, 0.0, 0.0), (0.0, 0.0, 100.0), samples=2, seed=7)
    cv = curves.view(-1, 4, 4)[:, :, :3].reshape(-1, 3)
    lo, hi = cv.min(0).values.tolist(), cv.max(0).values.tolist()
    del cv
    pad = [0.02 * (b - a) + 1.0 for a, b in zip(lo, hi)]
    mlp = V.ShadowMLP(ctx, [a - p for a, p in zip(lo, pad)], [b + p for b, p in zip(hi, pad)], seed=1)
    if args.load_params:
        mlp.set_params(np.load(args.load_params))
        print("[neural_frames] parameters loaded", flush=True)
    else:
        m = 1 << 24
        rays = torch.empty(m * 8, dtype=torch.float32, device="cuda")
        ctx.generate_surface_shadow_rays(tlas, m, 1, rays.data_ptr(), light=light, stream=sp)
        smp = torch.empty(m * 8, dtype=torch.float32, device="cuda")
        tgt = torch.empty(m, dtype=torch.float32, device="cuda")
        mlp.collect(tlas, rays.data_ptr(), m, smp.data_ptr(), tgt.data_ptr(), stream=sp)
        del rays
        mlp.train(smp.data_ptr(), tgt.data_ptr(), m, batch=65536, epochs=24, learning_rate=3e-3, stream=sp)
        del smp, tgt
        torch.cuda.synchronize()
        print("[neural_frames] trained", flush=True)
        if args.save_params:
            np.save(args.save_params, np.array(mlp.get_params(), np.float32))
    if args.wave_limit is not None:
        ctx.set_option(V.OPTION_SHADOW_WAVE_LIMIT, args.wave_limit)
    if args.group_lanes is not None:
        ctx.set_option(V.OPTION_SHADOW_GROUP_LANES, args.group_lanes)
    flags = V.RENDER_SHADOWS | V.RENDER_RGBA8 | V.RENDER_FP64_POLISH | V.RENDER_NEURAL_SHADOWS
    for thr in thresholds:
        def frame():
            ctx.render(tlas, cam, W, H, image.data_ptr(), flags=flags, shadow_mlp=mlp, variance_threshold=thr,
                       area_light=light, stream=sp)
        frame()  # warm-up
        torch.cuda.synchronize()
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record()
        for _ in range(args.frames):
            frame()
        t1.record()
        torch.cuda.synchronize()
        ms = t0.elapsed_time(t1) / max(args.frames, 1)
        prim, queries, traced = ctx.render_stats()
        crc = zlib.crc32(image.cpu().numpy().tobytes())
        print(f"[neural_frames] threshold {thr}: {queries} queries, {traced} traced exactly, {ms:.3f} ms/frame, "
              f"image crc32 {crc:08x}", flush=True)
    mlp.close()
    ctx.close()


if __name__ == "__main__":
    main()
