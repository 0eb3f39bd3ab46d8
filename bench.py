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
# VALU peak of MI355X in wave64 instructions/s: the published FP32 vector
# peak (157.3 TFLOP/s) over the 128 flops of one wave64 FMA instruction.
VALU_PEAK_GINST = 157.3e12 / 128 / 1e9


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed PMC summary
    (profiles/traffic.json: FETCH_SIZE and WRITE_SIZE runs of this bench,
    scaled by a 256 MiB copy), or None when no measurement is recorded."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None
    return None if rec is None else rec.get("traffic_bytes")


def sq_record(kernel):
    """SQ counters per launch of `kernel` from the committed profile
    (profiles/sq.json, written from a rocprofv3 --pmc run of this bench on the
    same launch), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "sq.json")) as f:
            return json.load(f).get(kernel)
    except (OSError, ValueError):
        return None


def roofline_entry(kernel, kern_ms, bytes_per_launch, launch):
    """Roofline of one launch of `kernel`, timed live (kern_ms).  The bound is
    VALU issue when the SQ profile of this launch is recorded (instructions
    per launch / time against the issue peak), else HBM.  The gathered-bytes
    rate (algorithmic bytes, mostly served by the caches) and the measured HBM
    rate (FETCH_SIZE + WRITE_SIZE per launch) are reported beside it."""
    t = kern_ms / 1e3
    gathered = bytes_per_launch / t / 1e9
    traffic = measured_traffic(kernel)
    hbm = None if traffic is None else traffic / t / 1e9
    sq = sq_record(kernel)
    if sq is not None:
        achieved = sq["valu_insts"] / t / 1e9
        entry = {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_GINST,
                 "unit": "G VALU wave-instructions/s", "frac": achieved / VALU_PEAK_GINST,
                 "valu_insts_per_launch": sq["valu_insts"], "valu_insts_per_wave": sq.get("valu_per_wave")}
    else:
        achieved = hbm if hbm is not None else gathered
        entry = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": achieved / HBM_PEAK_GBS}
    entry.update({"traffic": traffic, "kernel": kernel, "kernel_ms": kern_ms, "bytes_per_launch": bytes_per_launch,
                  "gathered_GBps": gathered, "gathered_frac_of_hbm_peak": gathered / HBM_PEAK_GBS,
                  "hbm_GBps": hbm, "hbm_frac": None if hbm is None else hbm / HBM_PEAK_GBS, "launch": launch})
    return entry


def radiance_bytes(keys, steps, n_probes, texel_bytes, request_bytes=0):
    """Gathered-bytes model of one radiance launch: 3 B (int16 block id +
    u8 voxel) per coarse sample and refinement fetch, 20 B of material plus
    8 trilinear texels per hit, and per probe the old texel read, the new
    texel written and its request.  keys and steps are numpy arrays or
    device tensors (reduced where they live)."""
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

    if args.sweep:
        sweep = []
        for variant, threads, P, eb, pb, k in ([(0, 0, 0, 0, 0, 1), (0, 0, 0, 0, 0, 2), (0, 0, 0, 0, 0, 4),
                                                 (2, 0, 0, 0, 0, 1)] +
                                                [(3, 0, P, 0, 0, 1) for P in (1, 2, 4) if rays == 64] +
                                                [(1, th, P, eb, pb, 1) for th in (256, 512, 1024) for P in (4, 16, 64)
                                                 for eb in (0, 1) for pb in (0, 1)]):
            if P * rays > 4096:
                continue
            try:
                w.set_tuning(variant=variant, threads=threads or 256, probes_per_batch=P or 16,
                             lds_block_bits=eb, lds_palette_bits=pb, wg_per_cu=0, rays_per_lane=k)
                w.radiance_update(None, params(0))
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                for frame in range(10):
                    w.radiance_update(None, params(frame))
                e1.record()
                torch.cuda.synchronize()
                sweep.append(dict(variant=variant, threads=threads, P=P, eb=eb, pb=pb, rays_per_lane=k,
                                  ms=e0.elapsed_time(e1) / 10))
            except Exception as ex:  # shape does not fit
                sweep.append(dict(variant=variant, threads=threads, P=P, eb=eb, pb=pb, rays_per_lane=k,
                                  error=str(ex)[:80]))
        print(json.dumps({"sweep": sweep}), file=sys.stderr, flush=True)
        w.write_radiance(field)
    tuning, tune_ms = w.autotune(params(0), reps=5)
    # algorithmic bytes of one launch (traced, untimed)
    keys, steps = w.radiance_update(None, params(0), traced=True)
    torch.cuda.synchronize()
    bytes_per_launch = radiance_bytes(keys, steps, n_probes, 16)

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
                     "variant": tuning["variant"],
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


def _time_frames(fn, n, torch):
    """Device time (ms) of n calls of fn(k) issued on torch's current stream."""
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record()
    for k in range(n):
        fn(k)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1), (time.perf_counter() - t0) * 1e3


def run_batched(args):
    """64 worlds (log-uniform 16x16x8..96x96x32 blocks, rays/probe from
    {32,64,128}, 4 models each, 640x360 glossy each): the map -> radiance ->
    glossy chain as one batched launch per pass captured in a hipGraph,
    against plain one-world-per-launch passes."""
    import torch
    from lum_amd import workloads
    from lum_amd.api import Context, Batch, Graph

    ctx = Context(0)
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    cases = workloads.batched_config(ctx, n_worlds=args.worlds)
    worlds = [c.world for c in cases]
    for c in cases:
        c.world.set_tuning(variant=0)
    batch = Batch(ctx, worlds)
    rays_per_frame = sum(c.probe_rays for c in cases)
    px_per_frame = sum(c.pixels for c in cases)
    gmax = 128

    def plain_frame(k):
        for c in cases:
            c.world.map_models_placed(c.models, c.placements)
            c.params.time_seed = k
            c.world.radiance_update(None, c.params)
            c.world.glossy(c.gbuf[0], c.gbuf[1], c.gbuf[2], 128.0, gmax, out=c.out)

    params = [c.params for c in cases]

    all_models = [c.models for c in cases]
    all_placements = [c.placements for c in cases]

    def batched_pass():
        batch.map_models(all_models, all_placements)
        batch.radiance(params)
        batch.glossy([c.gbuf[0] for c in cases], [c.gbuf[1] for c in cases], [c.gbuf[2] for c in cases],
                     [c.out for c in cases], 128.0, gmax)
        batch.advance_frame()

    def batched_frame(k):
        batched_pass()

    for c in cases:
        c.params.time_seed = 0
    batch_block, batch_rpl, _ = batch.autotune(params, reps=3)
    for k in range(args.warmup):
        plain_frame(k)
        batched_frame(k)
    frames = max(2, args.steps // 2 * 2)
    plain_ms, plain_wall = _time_frames(plain_frame, frames, torch)
    batched_ms, batched_wall = _time_frames(batched_frame, frames, torch)
    # the same batched chain captured once (two frames: even/odd fields)
    graph = Graph()
    torch.cuda.synchronize()
    graph.begin()
    batched_pass()
    batched_pass()
    graph.end()
    graph.launch()
    torch.cuda.synchronize()
    graph_ms, graph_wall = _time_frames(lambda k: graph.launch(), frames // 2, torch)
    plain_graph = Graph()
    plain_graph.begin()
    plain_frame(0)
    plain_frame(1)
    plain_graph.end()
    plain_graph.launch()
    torch.cuda.synchronize()
    pgraph_ms, pgraph_wall = _time_frames(lambda k: plain_graph.launch(), frames // 2, torch)
    # dominant kernel: the batched radiance launch alone
    def rad_only(k):
        batch.radiance(params)
    rad_ms, _ = _time_frames(rad_only, frames, torch)
    graph.close()
    plain_graph.close()
    per_frame = {
        "plain_eager_ms": plain_ms / frames,
        "plain_graph_ms": pgraph_ms / (frames // 2) / 2,
        "batched_eager_ms": batched_ms / frames,
        "batched_graph_ms": graph_ms / (frames // 2) / 2,
        "batched_radiance_ms": rad_ms / frames,
    }
    value = rays_per_frame / (per_frame["batched_graph_ms"] / 1e3)
    # Like for like: the per-world launches captured in a graph the same way.
    plain_value = rays_per_frame / (per_frame["plain_graph_ms"] / 1e3)
    out = {
        "metric": "probe-rays/s, batched map->radiance->glossy chain (hipGraph) vs plain per-world launches",
        "value": value,
        "unit": "probe-rays/s",
        "n_gpus": 1,
        "steps": frames,
        "warmup": args.warmup,
        "ms_per_step": per_frame["batched_graph_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded worlds, palettes, materials, models, G-buffers)",
        "config": {"workload": "batched: %d worlds log-uniform 16x16x8..96x96x32 blocks, rays/probe from "
                               "{32,64,128}, 4 models 16^3-32^3 per world, glossy 640x360 per world, FP32"
                               % len(cases),
                   "probe_rays_per_frame": rays_per_frame, "glossy_pixels_per_frame": px_per_frame},
        "plain_value": plain_value,
        "plain": "the same chain as per-world launches, captured in a hipGraph of two frames",
        "speedup_vs_plain": value / plain_value,
        "speedup_vs_plain_eager": value / (rays_per_frame / (per_frame["plain_eager_ms"] / 1e3)),
        "per_frame": per_frame,
        "batch_shape": {"block": batch_block, "rays_per_lane": batch_rpl},
    }
    batch.close()
    for c in cases:
        c.close()
    ctx.close()
    return out


REQUESTS_PER_STEP = 20


def run_serving(args):
    """200 resident worlds (log-uniform 16x16x8..128x128x32 blocks, FP16
    radiance) under an HBM budget of half their device state.  A stream of
    80% radiance updates (random 1-100% probe subset, 32/64/128 rays, no step
    cap) and 20% 640x360 reflection queries is served with continuous
    batching and, as the plain baseline, by the direct engine: one
    single-world call per request, queued on one stream with no wait between
    calls.  Reports saturated probe-rays/s of both and p50/p99 latency at a
    stated rate (70% of the plain engine's saturated request rate)."""
    import torch
    from lum_amd import synth
    from lum_amd.api import (Context, World, Batch, Server, radiance_params, camera_default, radiance_request,
                             glossy_request)
    from lum_amd._lib import RADIANCE_F16

    rng = np.random.default_rng(4242)
    ctx = Context(0)
    n_worlds = args.serving_worlds
    max_steps = args.max_steps if args.max_steps is not None else 0
    sizes = synth.log_uniform_sizes(rng, n_worlds, (16, 16, 8), (128, 128, 32))
    worlds, hosts, perms, perm_host, gbufs, outs = [], [], [], [], [], []
    W, H = 640, 360
    for size in sizes:
        blocks = synth.make_world(size, rng, 64)
        pal = synth.make_palette(rng, 64)
        mats = synth.make_materials(rng)
        field = synth.make_field(size, rng)
        w = World(ctx, size, static_blocks=64, palette_slots=64, radiance_format=RADIANCE_F16)
        w.set_blocks(blocks)
        w.set_palette(pal)
        w.set_materials(mats)
        w.write_radiance(field)
        w.upload_blocks()
        worlds.append(w)
        hosts.append((blocks, pal, mats, field))
        wx, wy, wz = size
        n = wx * wy * wz
        order = rng.permutation(n)
        q = np.stack([order % wx, (order // wx) % wy, order // (wx * wy), np.zeros(n, np.int64)], 1).astype(np.int32)
        q2 = np.concatenate([q, q], 0)
        perm_host.append(q2)
        perms.append(torch.from_numpy(q2).cuda())
        mn, dp = synth.make_gbuffer(W, H, mats, rng)
        gbufs.append((torch.from_numpy(mn).cuda(), torch.from_numpy(dp).cuda(), camera_default(W, H)))
        outs.append(torch.zeros((H, W, 4), dtype=torch.float32, device="cuda"))
    torch.cuda.synchronize()
    world_bytes = sum(w.resident()[1] for w in worlds)
    if args.hbm_budget_gb > 0:
        budget = int(args.hbm_budget_gb * (1 << 30))
    else:
        budget = int(args.hbm_budget_frac * world_bytes)
    budget = max(budget, max(w.resident()[1] for w in worlds)) if budget > 0 else 0

    # request trace
    # A serving step is REQUESTS_PER_STEP requests of the trace: --steps K
    # times exactly 20 K requests (4000 by default), --warmup W serves 20 W
    # untimed requests first.  --requests overrides the timed count.
    n_req = args.requests if args.requests else REQUESTS_PER_STEP * args.steps
    n_warm = REQUESTS_PER_STEP * args.warmup
    steps_out = n_req / REQUESTS_PER_STEP if args.requests else args.steps
    trace = []  # (kind, world, offset, count, rays)
    for i in range(n_req):
        wi = int(rng.integers(n_worlds))
        if rng.random() < 0.8:
            n = sizes[wi][0] * sizes[wi][1] * sizes[wi][2]
            cnt = max(1, int(round(rng.uniform(0.01, 1.0) * n)))
            trace.append(("r", wi, int(rng.integers(n)), cnt, int(rng.choice([32, 64, 128]))))
        else:
            trace.append(("g", wi, 0, 0, 0))

    def build(server):
        reqs = []
        for i, (kind, wi, off, cnt, rays) in enumerate(trace):
            if kind == "r":
                r = radiance_request(server.ids[wi], radiance_params(rays=rays, time_seed=i, alpha=0.1,
                                                                     max_steps=max_steps))
                r.requests_dev = perms[wi].data_ptr() + 16 * off
                r.n_requests = cnt
            else:
                g = gbufs[wi]
                r = glossy_request(server.ids[wi], g[0], g[1], g[2], outs[wi], 128.0, 128)
            reqs.append(r)
        return reqs

    probe_rays = sum(cnt * rays for kind, _, _, cnt, rays in trace if kind == "r")
    pixels = sum(W * H for kind, *_ in trace if kind == "g")

    # Launch shape of the served radiance launches (workgroup size x rays per
    # lane), chosen by the batch autotuner over max_batch random resident
    # worlds with rays from the request mix.  The trial launch is one batch of
    # the mix: a random probe subset of 1-100% per world (slices of the same
    # permutations the requests use).
    pick = rng.choice(n_worlds, size=min(args.max_batch, n_worlds), replace=False)
    tune_params, tune_reqs = [], []
    for i in pick:
        n = sizes[i][0] * sizes[i][1] * sizes[i][2]
        cnt = max(1, int(round(rng.uniform(0.01, 1.0) * n)))
        off = int(rng.integers(n))
        tune_reqs.append(perms[i][off:off + cnt])
        tune_params.append(radiance_params(rays=int(rng.choice([32, 64, 128])), time_seed=0, alpha=0.1,
                                           max_steps=max_steps))
    if args.served_shape:
        served_block, served_rpl = (int(v) for v in args.served_shape.split(","))
    else:
        tune_batch = Batch(ctx, [worlds[i] for i in pick])
        served_block, served_rpl, _ = tune_batch.autotune(tune_params, requests=tune_reqs, reps=5)
        tune_batch.close()
    torch.cuda.synchronize()

    def serve(engine, arrivals):
        if engine == "served":
            srv = Server(ctx, worlds, hbm_budget=budget, max_batch=args.max_batch, streams=args.streams,
                         block=served_block, rays_per_lane=served_rpl)
        elif engine == "plain":
            srv = Server(ctx, worlds, hbm_budget=budget, direct=True)
        else:  # one launch at a time, each awaited before the next
            srv = Server(ctx, worlds, hbm_budget=budget, max_batch=1, streams=1)
        lat, wall = srv.run_trace(build(srv)[:len(arrivals)], arrivals)
        st = srv.stats()
        srv.close()
        return lat, wall, st

    burst = [0.0] * n_req
    if n_warm > 0:
        for engine in ("plain", "served"):
            serve(engine, burst[: min(n_warm, n_req)])
    plain_lat_s, plain_wall_s, plain_st = serve("plain", burst)
    sync_lat_s, sync_wall_s, sync_st = serve("plain_sync", burst)
    served_lat_s, served_wall_s, served_st = serve("served", burst)
    rate = 0.7 * n_req / plain_wall_s
    gaps = rng.exponential(1.0 / rate, size=n_req)
    arrivals = list(np.cumsum(gaps) - gaps[0])
    plain_lat, plain_wall, _ = serve("plain", arrivals)
    served_lat, served_wall, served_rate_st = serve("served", arrivals)

    def pct(a, q):
        return float(np.percentile(np.asarray(a) * 1e3, q))

    value = probe_rays / served_wall_s
    plain_value = probe_rays / plain_wall_s

    # Dominant kernel: one batched radiance launch over every world (all probes,
    # 64 rays, FP16, no step cap), timed with events on this stream;
    # algorithmic bytes from a traced single-world update of each world with
    # the same parameters.
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    for w in worlds:
        w.restore()
    batch = Batch(ctx, worlds)
    batch.set_shape(served_block, served_rpl)
    kernel_name = "radiance_batch_kernel<1,false,false,%d>" % served_rpl
    rp = [radiance_params(rays=64, time_seed=0, alpha=0.1, max_steps=max_steps) for _ in worlds]
    bytes_per_launch = 0
    for w in worlds:
        keys, steps = w.radiance_update(None, rp[0], traced=True)
        bytes_per_launch += radiance_bytes(keys, steps, w.n_blocks, 8)
    for _ in range(2):
        batch.radiance(rp)
    torch.cuda.synchronize()
    reps = 10
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for a, b in ev:
        a.record()
        batch.radiance(rp)
        b.record()
    torch.cuda.synchronize()
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    batch.close()
    torch.cuda.set_stream(torch.cuda.default_stream())

    out = {
        "metric": "probe-rays/s and p50/p99 request latency, batched+served vs plain, 1xMI355X",
        "value": value,
        "unit": "probe-rays/s",
        "n_gpus": 1,
        "steps": steps_out,
        "warmup": min(n_warm, n_req) // REQUESTS_PER_STEP,
        "ms_per_step": served_wall_s * 1e3 / steps_out,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (fp16 radiance storage)",
        "data": "synthetic (seeded worlds, palettes, materials, G-buffers, request trace)",
        "config": {"workload": "serving: %d resident worlds log-uniform 16x16x8..128x128x32 blocks, FP16 radiance, "
                               "HBM budget %.0f%% of their device state; %d requests, 80%% radiance (1-100%% probe "
                               "subset, 32/64/128 rays, %s), 20%% glossy 640x360 (max 128 steps)"
                               % (n_worlds, 100.0 * budget / world_bytes, n_req,
                                  "max %d steps" % max_steps if max_steps else "no step cap"),
                   "requests": n_req, "requests_per_step": REQUESTS_PER_STEP,
                   "max_batch": args.max_batch, "streams": args.streams,
                   "served_shape": {"block": served_block, "rays_per_lane": served_rpl},
                   "hbm_budget_bytes": budget, "world_state_bytes": world_bytes,
                   "probe_rays": probe_rays, "glossy_pixels": pixels},
        "plain_value": plain_value,
        "speedup_vs_plain": value / plain_value,
        "plain_engine": "direct: one lumx_radiance_update / lumx_glossy call per request, queued on one stream "
                        "with no wait between calls",
        "swaps": {"served_evictions": served_st["evictions"], "served_restores": served_st["restores"],
                  "plain_evictions": plain_st["evictions"], "plain_restores": plain_st["restores"]},
        "latency_ms": {"rate_req_per_s": rate,
                       "plain_p50": pct(plain_lat, 50), "plain_p99": pct(plain_lat, 99),
                       "served_p50": pct(served_lat, 50), "served_p99": pct(served_lat, 99),
                       "plain_wall_s": plain_wall, "served_wall_s": served_wall},
        "saturated": {"plain_wall_s": plain_wall_s, "served_wall_s": served_wall_s,
                      "plain_launches": plain_st["launches"], "served_launches": served_st["launches"],
                      "plain_p50_ms": pct(plain_lat_s, 50), "served_p50_ms": pct(served_lat_s, 50),
                      "plain_sync_value": probe_rays / sync_wall_s, "plain_sync_wall_s": sync_wall_s},
        "roofline": roofline_entry(kernel_name, kern_ms, bytes_per_launch,
                                   "all %d worlds, every probe, 64 rays, FP16" % n_worlds),
    }
    if args.copy_calibration_mb > 0:
        out["copy_calibration_bytes"] = copy_calibration(args.copy_calibration_mb)
    if not args.no_cpu_baseline:
        out["cpu_baseline"], out["error_vs_fp64"] = cpu_baseline_serving(ctx, trace, hosts, perm_host, sizes,
                                                                         max_steps, seconds=args.cpu_seconds)
    for w in worlds:
        w.close()
    ctx.close()
    return out


def copy_calibration(mb):
    """A plain device copy of known size, for scaling PMC byte counters."""
    import torch
    n = int(mb * (1 << 20) // 4)
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    return n * 4


def cpu_baseline_serving(ctx, trace, hosts, perm_host, sizes, max_steps, seconds=10.0):
    """The C++ oracle (all host threads, FP32 field from the same initial
    values) on the trace's radiance requests in order, for about `seconds`.

    As the checker, the same leg states the served results' error: for the
    first radiance request of up to 8 worlds (at most 4M probe-rays each) the
    GPU's FP16 result of that request, computed on a fresh copy of its world
    (served results equal these alone results bit for bit), against the FP64
    oracle from the same initial field, over the requested probes' rgb."""
    import tests._oracle as orc
    threads = os.cpu_count() or 1
    fields = {}
    done = 0
    n = 0
    t0 = time.perf_counter()
    for i, (kind, wi, off, cnt, rays) in enumerate(trace):
        if kind != "r":
            continue
        blocks, pal, mats, field = hosts[wi]
        wd = orc.WorldData(blocks, pal, mats)
        f = fields.get(wi, field)
        req = perm_host[wi][off:off + cnt]
        fields[wi] = orc.radiance_update(wd, f, requests=req, rays=rays, time_seed=i, alpha=0.1,
                                         max_steps=max_steps, threads=threads)
        done += cnt * rays
        n += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    baseline = {"value": done / dt, "unit": "probe-rays/s", "cores": threads, "kind": "port",
                "sample": f"first {n} radiance requests of the serving trace ({done} probe-rays, {dt:.1f} s) by the "
                          f"C++ oracle on {threads} threads"}
    return baseline, served_error_vs_fp64(ctx, trace, hosts, perm_host, sizes, max_steps, threads)


def served_error_vs_fp64(ctx, trace, hosts, perm_host, sizes, max_steps, threads):
    import torch
    import tests._oracle as orc
    from lum_amd.api import World, radiance_params
    from lum_amd._lib import RADIANCE_F16
    seen, sample = set(), []
    for i, (kind, wi, off, cnt, rays) in enumerate(trace):
        if kind == "r" and wi not in seen:
            seen.add(wi)
            if cnt * rays <= 4_000_000:
                sample.append((i, wi, off, cnt, rays))
        if len(sample) == 8:
            break
    rels = []
    for i, wi, off, cnt, rays in sample:
        blocks, pal, mats, field = hosts[wi]
        req = perm_host[wi][off:off + cnt]
        w = World(ctx, sizes[wi], static_blocks=64, palette_slots=64, radiance_format=RADIANCE_F16)
        try:
            w.set_blocks(blocks)
            w.set_palette(pal)
            w.set_materials(mats)
            w.write_radiance(field)
            w.upload_blocks()
            w.radiance_update(torch.from_numpy(np.ascontiguousarray(req)).cuda(),
                              radiance_params(rays=rays, time_seed=i, alpha=0.1, max_steps=max_steps))
            torch.cuda.synchronize()
            got = w.read_radiance()
        finally:
            w.close()
        ref = orc.radiance_update(orc.WorldData(blocks, pal, mats), field.astype(np.float64), requests=req,
                                  rays=rays, time_seed=i, alpha=0.1, max_steps=max_steps, threads=threads, fp64=True)
        wx, wy = sizes[wi][0], sizes[wi][1]
        idx = req[:, 0] + wx * (req[:, 1] + wy * req[:, 2])
        g = got.reshape(-1, 4)[idx, :3].astype(np.float64)
        r = ref.reshape(-1, 4)[idx, :3]
        rels.append((np.abs(g - r) / np.maximum(np.abs(r), 1e-3)).ravel())
    rel = np.concatenate(rels) if rels else np.zeros(1)
    return {"storage": "fp16", "requests": len(sample), "texels": int(rel.size // 3),
            "mean": float(rel.mean()), "p99": float(np.percentile(rel, 99)), "max": float(rel.max()),
            "frac_within_1e-3": float((rel <= 1e-3).mean()),
            "sample": "first radiance request of %d worlds of the trace (<= 4M probe-rays each), FP16 result on the "
                      "GPU vs the FP64 oracle from the same initial field, rgb of the requested probes"
                      % len(sample)}


def run_chain(args):
    """The chain workload: the 48x48x16 world plus 32 models of 16^3-64^3
    (30% fill, uniform rotations) mapped, every probe updated (64 rays), and a
    1920x1080 glossy march (max 128 steps) from a synthetic G-buffer, FP32.
    A step is the whole map -> radiance -> glossy chain on one stream."""
    import torch
    from lum_amd import synth
    from lum_amd.api import Context, World, Model, placement, camera_default, radiance_params

    cfg = synth.PLAIN
    rng = np.random.default_rng(cfg["seed"])
    size = cfg["size"]
    blocks = synth.make_world(size, rng, cfg["n_palette"])
    pal = synth.make_palette(rng, cfg["n_palette"])
    mats = synth.make_materials(rng)
    field = synth.make_field(size, rng)
    vox, quats, shifts = synth.make_models(rng, n=32, lo=16, hi=64, fill=0.3, world_size=size)
    W, H = 1920, 1080
    mn, dp = synth.make_gbuffer(W, H, mats, rng)
    ctx = Context(0)
    w = World(ctx, size, static_blocks=cfg["n_palette"])
    w.set_blocks(blocks)
    w.set_palette(pal)
    w.set_materials(mats)
    w.write_radiance(field)
    ms = [Model(ctx, v) for v in vox]
    pls = [placement(q, s_, m.size) for q, s_, m in zip(quats, shifts, ms)]
    mn_d, dp_d = torch.from_numpy(mn).cuda(), torch.from_numpy(dp).cuda()
    cam = camera_default(W, H)
    out = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    map_threads = sum(p.groups[0] * p.groups[1] * p.groups[2] * 64 for p in pls)

    def step(k):
        w.map_models_placed(ms, pls)
        w.radiance_update(None, radiance_params(rays=64, time_seed=k, alpha=0.1, max_steps=64))
        w.glossy(mn_d, dp_d, cam, 128.0, 128, out=out)

    for k in range(args.warmup):
        step(k)
    torch.cuda.synchronize()
    pass_ms = {"map": [], "radiance": [], "glossy": []}
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for k in range(args.steps):
        step(100 + k)
        torch.cuda.synchronize()
        pass_ms["map"].append(w.last_pass_ms(1))
        pass_ms["radiance"].append(w.last_pass_ms(0))
        pass_ms["glossy"].append(w.last_pass_ms(2))
    e1.record()
    torch.cuda.synchronize()
    total = e0.elapsed_time(e1)
    per = {k: float(np.mean(v)) for k, v in pass_ms.items()}
    # The same chain with blockify and exec_copies on the device.
    dev_ms = {"map": [], "radiance": [], "glossy": []}
    e0.record()
    for k in range(args.steps):
        w.map_models_device(ms, pls)
        w.radiance_update(None, radiance_params(rays=64, time_seed=100 + k, alpha=0.1, max_steps=64))
        w.glossy(mn_d, dp_d, cam, 128.0, 128, out=out)
        torch.cuda.synchronize()
        dev_ms["map"].append(w.last_pass_ms(1))
        dev_ms["radiance"].append(w.last_pass_ms(0))
        dev_ms["glossy"].append(w.last_pass_ms(2))
    e1.record()
    torch.cuda.synchronize()
    dev_total = e0.elapsed_time(e1)
    rays = w.n_blocks * 64
    glossy_px = int((mats[mn[..., 0], 5] <= 0.5).sum())
    res = {
        "metric": "probe-rays/s (map -> radiance -> glossy chain, 48x48x16 world, 32 models, 1920x1080 glossy)",
        "value": rays / (per["radiance"] / 1e3),
        "unit": "probe-rays/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": total / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded world, palette, materials, 32 models, G-buffer)",
        "config": {"workload": "chain: 48x48x16 world + 32 models 16^3-64^3 at 30% fill (uniform rotations), "
                               "36864 probes x 64 rays, glossy 1920x1080 max 128 steps, FP32",
                   "palette_slots_used": w.palette_counter(), "map_threads": map_threads,
                   "glossy_pixels": glossy_px},
        "pass_ms": per,
        "device_map_chain": {"pass_ms": {k: float(np.mean(v)) for k, v in dev_ms.items()},
                             "ms_per_step": dev_total / args.steps},
        "map_voxels_per_s": map_threads / (per["map"] / 1e3),
        "glossy_rays_per_s": glossy_px / (per["glossy"] / 1e3),
    }
    # map.comp's byte model: per thread an int16 block id and a u8 model voxel,
    # per non-zero model voxel one u8 palette store; against the device map
    # pass (blockify, copies, bits and map kernels, HIP events of the library).
    map_bytes = 3 * map_threads + int(sum(int((v != 0).sum()) for v in vox))
    dev_map_ms = res["device_map_chain"]["pass_ms"]["map"]
    res["map"] = {"threads": map_threads, "bytes_model": map_bytes, "device_pass_ms": dev_map_ms,
                  "GBps_at_pass_time": map_bytes / (dev_map_ms / 1e3) / 1e9,
                  "frac_of_hbm_peak_at_pass_time": map_bytes / (dev_map_ms / 1e3) / 1e9 / HBM_PEAK_GBS}
    if args.copy_calibration_mb > 0:
        res["copy_calibration_bytes"] = copy_calibration(args.copy_calibration_mb)
    for m in ms:
        m.close()
    w.close()
    ctx.close()
    return res


def run_frames(args):
    """The drop-in frame loop: the chain world (48x48x16, 64-block palette,
    32 models of 16^3-64^3) and a 1920x1080 G-buffer driven through the C99
    API (include/clum.h) as a C host drives it: per frame start, one draw per
    model, prepare, submit.  Frames run map -> shift -> radiance (device probe
    selection, counted dispatch) -> glossy with two frames in flight; the
    timed region is K submissions followed by lum_wait_idle."""
    import ctypes as C
    import torch
    from lum_amd import synth, clum

    cfg = synth.PLAIN
    rng = np.random.default_rng(cfg["seed"])
    size = cfg["size"]
    static = cfg["n_palette"]
    blocks = synth.make_world(size, rng, static)
    pal = synth.make_palette(rng, static)
    mats = synth.make_materials(rng)
    synth.make_field(size, rng)
    vox, quats, shifts = synth.make_models(rng, n=32, lo=16, hi=64, fill=0.3, world_size=size)
    W, H = 1920, 1080
    mn, dp = synth.make_gbuffer(W, H, mats, rng)
    mn_d, dp_d = torch.from_numpy(mn).cuda(), torch.from_numpy(dp).cuda()
    out = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    L = clum.lib()
    s = clum.settings(size, static)
    lum = L.lum_create_instance(C.byref(s))

    def ptr(a):
        return a.ctypes.data_as(C.c_void_p)

    L.lum_init(lum, C.byref(s))
    b16 = np.ascontiguousarray(blocks, np.int16)
    L.lum_load_world(lum, ptr(b16))
    for b in range(1, static):
        L.lum_load_block(lum, b, ptr(np.ascontiguousarray(pal[b].transpose(2, 1, 0))))
    L.lum_upload_block_palette_to_gpu(lum)
    m32 = np.ascontiguousarray(mats, np.float32)
    L.lumx_lum_set_materials(lum, ptr(m32))
    L.lum_upload_material_palette_to_gpu(lum)
    keep = [np.ascontiguousarray(v) for v in vox]
    meshes = [L.lum_load_mesh_data(lum, ptr(v), v.shape[2], v.shape[1], v.shape[0]) for v in keep]
    trans = [clum.transform(q, sh) for q, sh in zip(quats, shifts)]
    L.lumx_lum_set_gbuffer(lum, C.c_void_p(mn_d.data_ptr()), C.c_void_p(dp_d.data_ptr()), W, H,
                           C.c_void_p(out.data_ptr()))

    def frame():
        L.lum_start_frame(lum)
        for m, t in zip(meshes, trans):
            L.lum_draw_model(lum, m, C.byref(t))
        L.lum_prepare_frame(lum)
        L.lum_submit_frame(lum)

    for _ in range(max(args.warmup, 2)):
        frame()
    L.lum_wait_idle(lum)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        frame()
    t_submit = time.perf_counter() - t0
    L.lum_wait_idle(lum)
    wall = time.perf_counter() - t0
    counter = C.c_int()
    L.lumx_world_palette_counter(L.lumx_lum_world(lum), C.byref(counter))
    passes = {"map_ms": L.lum_get_timestamp_difference(lum, 0, 1),
              "radiance_ms": L.lum_get_timestamp_difference(lum, 1, 2),
              "glossy_ms": L.lum_get_timestamp_difference(lum, 2, 3)}
    L.lum_cleanup(lum)
    return {
        "metric": "frames/s of the clum.h drop-in frame (map -> radiance -> glossy), 48x48x16 world, 32 models, "
                  "1920x1080 glossy",
        "value": args.steps / wall, "unit": "frames/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded world, palette, materials, 32 models, G-buffer)",
        "config": {"workload": "frames: lum_start_frame / 32 x lum_draw_model / lum_prepare_frame / "
                               "lum_submit_frame, two frames in flight", "palette_slots_used": counter.value},
        "host_submit_ms_per_frame": t_submit * 1e3 / args.steps,
        "pass_ms_of_a_finished_frame": passes,
    }


def run_precision(args):
    """The precision workload: a 256x256x32-block world (2,097,152 probes),
    128 rays per probe, max 96 steps, 8 accumulated frames, radiance stored in
    FP32, FP16 and BF16 (FP32 march and accumulation).  Each is compared with
    the FP64 CPU oracle run on the same inputs, on a seeded 4,096-probe
    subset."""
    import torch
    import tests._oracle as orc
    from lum_amd import synth
    from lum_amd.api import Context, World, radiance_params
    from lum_amd._lib import RADIANCE_F32, RADIANCE_F16, RADIANCE_BF16

    size = tuple(int(v) for v in args.size.split(",")) if args.size else (256, 256, 32)
    rays, steps, frames, alpha = 128, 96, args.frames, 0.1
    rng = np.random.default_rng(31337)
    blocks = synth.make_world(size, rng, 64)
    pal = synth.make_palette(rng, 64)
    mats = synth.make_materials(rng)
    field = synth.make_field(size, rng)
    n = size[0] * size[1] * size[2]
    subset = np.sort(rng.choice(n, size=min(4096, n), replace=False))
    ctx = Context(0)
    res_fmt = {}
    gpu_fields = {}
    for name, fmt in (("fp32", RADIANCE_F32), ("fp16", RADIANCE_F16), ("bf16", RADIANCE_BF16)):
        w = World(ctx, size, static_blocks=64, palette_slots=64, radiance_format=fmt)
        w.set_blocks(blocks)
        w.set_palette(pal)
        w.set_materials(mats)
        w.write_radiance(field)
        w.upload_blocks()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for f in range(frames):
            w.radiance_update(None, radiance_params(rays=rays, time_seed=f, alpha=alpha, max_steps=steps))
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / frames
        gpu_fields[name] = w.read_radiance().reshape(-1, 4)[subset, :3].astype(np.float64)
        res_fmt[name] = {"ms_per_frame": ms, "probe_rays_per_s": n * rays / (ms / 1e3)}
        w.close()
    ctx.close()
    # FP64 oracle: full field (bounce lookups read every probe), all host threads
    t0 = time.perf_counter()
    wd = orc.WorldData(blocks, pal, mats)
    f64 = field.astype(np.float64)
    for f in range(frames):
        f64 = orc.radiance_update(wd, f64, rays=rays, time_seed=f, alpha=alpha, max_steps=steps,
                                  threads=os.cpu_count() or 1, fp64=True)
    cpu_s = time.perf_counter() - t0
    ref = f64.reshape(-1, 4)[subset, :3]
    for name, got in gpu_fields.items():
        rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-3)
        res_fmt[name].update({"max_rel_err": float(rel.max()), "mean_rel_err": float(rel.mean()),
                              "p99_rel_err": float(np.percentile(rel, 99)),
                              "frac_within_1e-3": float((rel <= 1e-3).mean())})
    fp32 = gpu_fields["fp32"]
    for name in ("fp16", "bf16"):
        rel = np.abs(gpu_fields[name] - fp32) / np.maximum(np.abs(fp32), 1e-3)
        res_fmt[name]["max_rel_err_vs_fp32_storage"] = float(rel.max())
        res_fmt[name]["mean_rel_err_vs_fp32_storage"] = float(rel.mean())
    return {
        "metric": "radiance error of FP32/FP16/BF16 storage vs an FP64 CPU oracle (and probe-rays/s)",
        "value": res_fmt["fp16"]["probe_rays_per_s"],
        "unit": "probe-rays/s",
        "n_gpus": 1, "steps": frames, "warmup": 0,
        "ms_per_step": res_fmt["fp16"]["ms_per_frame"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 (fp16 storage)",
        "data": "synthetic (seeded world, palette, materials, field)",
        "config": {"workload": "precision: %dx%dx%d blocks, %d rays, max %d steps, %d frames, alpha %.1f; "
                               "error on a seeded %d-probe subset" % (size + (rays, steps, frames, alpha, len(subset)))},
        "formats": res_fmt,
        "fp64_oracle_cpu_s": cpu_s,
    }


def make_parser():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="serving",
                    choices=["serving", "plain", "batched", "chain", "precision", "frames"])
    ap.add_argument("--frames", type=int, default=8)
    ap.add_argument("--serving-worlds", type=int, default=200)
    ap.add_argument("--requests", type=int, default=0,
                    help="serving: timed requests (default 20 per step, 4000 at the default --steps)")
    ap.add_argument("--max-batch", type=int, default=32, help="serving: requests per launch")
    ap.add_argument("--streams", type=int, default=4, help="serving: launches in flight")
    ap.add_argument("--hbm-budget-gb", type=float, default=0.0,
                    help="serving: HBM budget in GiB (default: --hbm-budget-frac of the worlds' device state)")
    ap.add_argument("--hbm-budget-frac", type=float, default=0.5,
                    help="serving: HBM budget as a fraction of the worlds' device state (0: no budget)")
    ap.add_argument("--worlds", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--size", default="", help="override world size, e.g. 256,256,32")
    ap.add_argument("--rays", type=int, default=0)
    ap.add_argument("--max-steps", type=int, default=None)
    ap.add_argument("--sweep", action="store_true", help="time every radiance launch shape (stderr)")
    ap.add_argument("--served-shape", default="", help="serving: radiance launch shape 'block,rays_per_lane' "
                                                        "instead of the autotuner's")
    ap.add_argument("--copy-calibration-mb", type=float, default=0.0,
                    help="also run a plain device copy of this size (PMC byte-counter calibration)")
    return ap


def main():
    args = make_parser().parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the hot path has no multi-GPU form here); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        out = run_batched(args)
    elif args.workload == "chain":
        out = run_chain(args)
    elif args.workload == "precision":
        out = run_precision(args)
    elif args.workload == "frames":
        out = run_frames(args)
    elif args.workload == "plain":
        out = run_plain(args)
    else:
        out = run_serving(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
