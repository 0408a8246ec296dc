"""Work counters of the traversal kernels on the 100k-strand hair frame, from the counting build
of the library (libvkhrt_count.so: the same kernels compiled with VKHRT_WORK_COUNTERS).

    python -m vkhrt_amd.workcount [--frames 5] [--accurate 1] [--area 0]

Renders the frame (100,000 seeded strands x 32 tapered segments, 1920x1080, a shadow ray per
hit, FP64 polish) and prints one JSON object: per frame and per traversal kind (closest hit,
any hit, and with --area 1 the lane-group walker of the 16M-segment area-light frame's neural
fallback) the loop iterations, intersection rounds, traversal steps, node and leaf tests, Prhi
calls, lane use and the shares of shader-clock cycles spent in rounds, steps and bookkeeping.
The counters add clock reads and atomics to the kernels, so their frame times are not the
product's; the counts are those of the product kernels (same code paths).
"""
import argparse
import json
import math
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

NAMES = ["iterations", "rounds", "steps", "node_tests", "leaf_tests", "prhi_calls", "round_lanes", "cycles_rounds",
         "cycles_steps", "cycles_other", "waves", "rays", "handovers", "step_lanes"]
WAVE_NAMES = ["rays", "steps", "node_tests", "leaf_tests", "prhi_calls", "record_steps", "cycles"]


def summarize(raw, base, frames):
    c = {k: raw[base + i] / frames for i, k in enumerate(NAMES)}
    cyc = c["cycles_rounds"] + c["cycles_steps"] + c["cycles_other"]
    rays = max(c["rays"], 1.0)
    return {
        "per_frame": {k: round(v, 1) for k, v in c.items()},
        "per_ray": {k: round(c[k] / rays, 3) for k in ("node_tests", "leaf_tests", "prhi_calls", "iterations")},
        "lane_use": {"steps": round(c["step_lanes"] / max(64 * c["steps"], 1), 4),
                     "rounds": round(c["round_lanes"] / max(64 * c["rounds"], 1), 4)},
        "cycle_share": {"rounds": round(c["cycles_rounds"] / max(cyc, 1), 4),
                        "steps": round(c["cycles_steps"] / max(cyc, 1), 4),
                        "other": round(c["cycles_other"] / max(cyc, 1), 4)},
        "iteration_share": {"rounds": round(c["rounds"] / max(c["iterations"], 1), 4),
                            "steps": round(c["steps"] / max(c["iterations"], 1), 4)},
    }


def _wave_loop(r, frames):
    """Loop timing of the lane-group walker's wavefronts (those that walked a ray)."""
    round_cyc, rounds, iters, loop_cyc, waves = (r[32 + 8 + i] / frames for i in range(5))
    waves = max(waves, 1.0)
    return {"waves": round(waves, 1), "iterations": round(iters / waves, 2), "rounds": round(rounds / waves, 2),
            "loop_cycles": round(loop_cyc / waves), "cycles_per_iteration": round(loop_cyc / max(iters, 1.0)),
            "cycles_per_round": round(round_cyc / max(rounds, 1.0)),
            "round_share": round(round_cyc / max(loop_cyc, 1.0), 4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--frames", type=int, default=5)
    ap.add_argument("--strands", type=int, default=100000)
    ap.add_argument("--segments", type=int, default=32)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--seed", type=int, default=20240601)
    ap.add_argument("--accurate", type=int, default=1)
    ap.add_argument("--area", type=int, default=0, help="also the 16M-segment area-light frame's neural fallback")
    ap.add_argument("--group-lanes", type=int, default=16, help="lanes per ray of the fallback's lane-group walker")
    args = ap.parse_args()

    import ctypes

    import torch
    import vkhrt_amd as V
    V.LIB_PATH = os.path.join(HERE, "libvkhrt_count.so")
    assert torch.cuda.is_available(), "needs a GPU"
    sp = torch.cuda.current_stream().cuda_stream
    ctx = V.Context(0)
    ctx.set_option(V.OPTION_SHADOW_GROUP_LANES, args.group_lanes)
    raw = (ctypes.c_uint64 * 48)()

    def read(reset=True):
        ctx.check(V.lib().vkhrt_work_counters(ctx.h, raw, 48, int(reset)), "work_counters")
        return list(raw)

    def hair(ns, nseg, seed):
        n = ns * nseg
        verts = torch.empty(ns * (nseg + 1) * 4, dtype=torch.float32, device="cuda")
        lines = torch.empty(n * 8, dtype=torch.float32, device="cuda")
        curves = torch.empty(n * 16, dtype=torch.float32, device="cuda")
        ctx.generate_strands(seed, ns, nseg, verts.data_ptr(), stream=sp)
        ctx.generate_lines(verts.data_ptr(), ns, nseg, lines.data_ptr(), stream=sp)
        ctx.generate_curves(lines.data_ptr(), n, curves.data_ptr(), stream=sp)
        blas = V.BottomLevelAccelerationStructure(ctx, V.GEOMETRY_PHANTOM_CURVES, curves.data_ptr(), n,
                                                  flags=V.BLAS_PER_POINT_RADIUS, stream=sp)
        tlas = V.TopLevelAccelerationStructure(ctx, [blas], stream=sp)
        torch.cuda.synchronize()
        return curves, blas, tlas

    W, H = args.width, args.height
    curves, blas, tlas = hair(args.strands, args.segments, args.seed)
    cam = V.camera_look_at((0.0, 60.0, 300.0), (0.0, 40.0, 0.0), fov_deg=60.0, aspect=W / H)
    image = torch.empty(W * H, dtype=torch.int32, device="cuda")
    flags = V.RENDER_SHADOWS | V.RENDER_RGBA8 | V.RENDER_FP64_POLISH | (V.RENDER_ACCURATE_NORMALS if args.accurate else 0)
    ctx.render(tlas, cam, W, H, image.data_ptr(), flags=flags, stream=sp)
    read()
    for _ in range(args.frames):
        ctx.render(tlas, cam, W, H, image.data_ptr(), flags=flags, stream=sp)
    r = read()
    out = {"workload": "phantom_100k_strands_x32_segments_1920x1080_primary_plus_shadow", "frames": args.frames,
           "primary": summarize(r, 0, args.frames), "shadow": summarize(r, 16, args.frames)}
    del curves, blas, tlas
    if args.area:
        curves, blas, tlas = hair(500000, 32, 104)
        light = V.area_light((0.0, 260.0, 0.0), (100.0, 0.0, 0.0), (0.0, 0.0, 100.0), samples=2, seed=7)
        cv = curves.view(-1, 4, 4)[:, :, :3].reshape(-1, 3)
        lo, hi = cv.min(0).values.tolist(), cv.max(0).values.tolist()
        del cv
        pad = [0.02 * (b - a) + 1.0 for a, b in zip(lo, hi)]
        mlp = V.ShadowMLP(ctx, [a - p for a, p in zip(lo, pad)], [b + p for b, p in zip(hi, pad)], seed=1)
        n = 1 << 24
        rays = torch.empty(n * 8, dtype=torch.float32, device="cuda")
        ctx.generate_surface_shadow_rays(tlas, n, 1, rays.data_ptr(), light=light, stream=sp)
        smp = torch.empty(n * 8, dtype=torch.float32, device="cuda")
        tgt = torch.empty(n, dtype=torch.float32, device="cuda")
        mlp.collect(tlas, rays.data_ptr(), n, smp.data_ptr(), tgt.data_ptr(), stream=sp)
        del rays
        mlp.train(smp.data_ptr(), tgt.data_ptr(), n, batch=65536, epochs=24, learning_rate=3e-3, stream=sp)
        del smp, tgt
        nflags = V.RENDER_SHADOWS | V.RENDER_RGBA8 | V.RENDER_FP64_POLISH | V.RENDER_NEURAL_SHADOWS
        thr = 0.2499
        ctx.render(tlas, cam, W, H, image.data_ptr(), flags=nflags, shadow_mlp=mlp, variance_threshold=thr,
                   area_light=light, stream=sp)
        read()
        for _ in range(args.frames):
            ctx.render(tlas, cam, W, H, image.data_ptr(), flags=nflags, shadow_mlp=mlp, variance_threshold=thr,
                       area_light=light, stream=sp)
        r = read()
        _, q, tr = ctx.render_stats()
        w = {k: r[32 + i] / args.frames for i, k in enumerate(WAVE_NAMES)}
        rays_w = max(w["rays"], 1.0)
        out["area_neural_fallback"] = {
            "threshold": thr, "queries": q, "traced": tr, "lanes_per_ray": args.group_lanes,
            "per_ray": {k: round(w[k] / rays_w, 2) for k in ("steps", "node_tests", "leaf_tests", "prhi_calls",
                                                            "record_steps", "cycles")},
            "slowest_ray_cycles": int(r[32 + len(WAVE_NAMES)]),
            "per_wave": _wave_loop(r, args.frames),
            "primary": summarize(r, 0, args.frames)}
        mlp.close()
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
