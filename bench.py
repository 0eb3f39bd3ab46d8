#!/usr/bin/env python3
"""Benchmark of the Lum voxel-lighting hot path on one MI355X.

Default workload (plain radiance update, FP32): a 48x48x16-block world over a
64-block random palette, every one of the 36,864 probes updated per step with
64 Fibonacci rays (randomly rotated per frame), at most 64 coarse steps,
blend alpha 0.1.  A step is one radiance update of the whole world; inputs
are resident in HBM before the timed region.

Prints one JSON line (rank 0).  `value` is probe-rays/s of the timed steps.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def radiance_bytes(keys, steps, n_probes, texel_bytes, request_bytes=0):
    """Gathered-bytes model of one radiance launch: 3 B (int16 block id +
    u8 voxel) per coarse sample and refinement fetch, 20 B of material plus
    8 trilinear texels per hit, and per probe the old texel read, the new
    texel written and its request."""
    samples = int(steps[:, 0].sum()) + int(steps[:, 1].sum())
    hits = int((keys != 0).sum())
    return 3 * samples + hits * (20 + 8 * texel_bytes) + n_probes * (2 * texel_bytes + request_bytes)


def cpu_baseline_plain(blocks, pal, mats, field, rays, alpha, max_steps, seconds=10.0):
    """Time the C++ oracle (all host threads) on repeated full updates."""
    import tests._oracle as orc
    threads = os.cpu_count() or 1
    w = orc.WorldData(blocks, pal, mats)
    n_rays = blocks.size * rays
    f = field
    done = 0
    t0 = time.perf_counter()
    frame = 0
    while True:
        f = orc.radiance_update(w, f, rays=rays, time_seed=frame, alpha=alpha, max_steps=max_steps,
                                threads=threads)
        done += n_rays
        frame += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "probe-rays/s", "cores": threads, "kind": "port",
            "sample": f"{frame} full updates of the same 48x48x16 world ({done} rays, {dt:.1f} s) by the "
                      f"C++ oracle on {threads} threads"}


def run_plain(args):
    import torch
    from lum_amd import synth
    from lum_amd.api import Context, World, radiance_params

    cfg = dict(synth.PLAIN)
    if args.size:
        cfg["size"] = tuple(int(v) for v in args.size.split(","))
    if args.rays:
        cfg["rays"] = args.rays
    if args.max_steps is not None:
        cfg["max_steps"] = args.max_steps
    rng = np.random.default_rng(cfg["seed"])
    size = cfg["size"]
    blocks = synth.make_world(size, rng, cfg["n_palette"])
    pal = synth.make_palette(rng, cfg["n_palette"])
    mats = synth.make_materials(rng)
    field = synth.make_field(size, rng)

    ctx = Context(0)
    w = World(ctx, size, static_blocks=cfg["n_palette"])
    w.set_blocks(blocks)
    w.set_palette(pal)
    w.set_materials(mats)
    w.write_radiance(field)
    w.upload_blocks()
    rays = cfg["rays"]
    n_probes = w.n_blocks

    def params(frame):
        return radiance_params(rays=rays, time_seed=frame, alpha=cfg["alpha"], max_dist=cfg["max_dist"],
                               max_steps=cfg["max_steps"])

    tuning, tune_ms = w.autotune(params(0), reps=5)
    # algorithmic bytes of one launch (traced, untimed)
    keys, steps = w.radiance_update(None, params(0), traced=True)
    torch.cuda.synchronize()
    bytes_per_launch = radiance_bytes(keys.cpu().numpy(), steps.cpu().numpy(), n_probes, 16)

    for i in range(args.warmup):
        w.radiance_update(None, params(1 + i))
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_start.record()
    for k in range(args.steps):
        ev[k][0].record()
        w.radiance_update(None, params(100 + k))
        ev[k][1].record()
    t_end.record()
    torch.cuda.synchronize()
    total_ms = t_start.elapsed_time(t_end)
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    rays_per_step = n_probes * rays
    value = rays_per_step * args.steps / (total_ms / 1e3)
    achieved = bytes_per_launch / (kern_ms / 1e3) / 1e9
    out = {
        "metric": "probe-rays/s (plain radiance update, %dx%dx%d world, %d rays/probe)" % (size + (rays,)),
        "value": value,
        "unit": "probe-rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Perlin-terrain world, random palette and materials)",
        "config": {"workload": "plain radiance update: %dx%dx%d blocks, 64-block palette, %d probes x %d rays, "
                               "max %d steps, alpha 0.1, FP32 field" % (size + (n_probes, rays, cfg["max_steps"])),
                   "probes": n_probes, "rays_per_probe": rays},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "radiance_tiled_kernel" if tuning["variant"] == 1 else "radiance_kernel",
                     "kernel_ms": kern_ms, "bytes_per_launch": bytes_per_launch},
        "tuning": tuning,
    }
    if args.copy_calibration_mb > 0:
        # A plain device copy of known size, for scaling PMC byte counters.
        n = int(args.copy_calibration_mb * (1 << 20) // 4)
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        out["copy_calibration_bytes"] = n * 4
    w.close()
    ctx.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_plain(blocks, pal, mats, field, rays, cfg["alpha"], cfg["max_steps"],
                                                 seconds=args.cpu_seconds)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="plain", choices=["plain"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--size", default="", help="override world size, e.g. 256,256,32")
    ap.add_argument("--rays", type=int, default=0)
    ap.add_argument("--max-steps", type=int, default=None)
    ap.add_argument("--copy-calibration-mb", type=float, default=0.0,
                    help="also run a plain device copy of this size (PMC byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the hot path has no multi-GPU form here); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    out = run_plain(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
