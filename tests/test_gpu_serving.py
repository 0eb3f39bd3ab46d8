"""Serving engine: every request's result is the same alone and in the mix
(bit for bit), also when the HBM budget forces worlds out and back in."""
import numpy as np
import pytest

from lum_amd import synth
from lum_amd.api import (World, Server, radiance_params, camera_default, radiance_request, glossy_request)
from lum_amd._lib import RADIANCE_F16

pytestmark = pytest.mark.gpu

SIZES = [(16, 16, 8), (24, 16, 8), (12, 20, 10), (16, 16, 16), (20, 20, 6), (8, 8, 8)]


def make(ctx, size, seed, fmt=RADIANCE_F16):
    rng = np.random.default_rng(seed)
    blocks = synth.make_world(size, rng, 32)
    pal = synth.make_palette(rng, 32)
    mats = synth.make_materials(rng)
    field = synth.make_field(size, rng)
    w = World(ctx, size, static_blocks=32, palette_slots=32, radiance_format=fmt)
    w.set_blocks(blocks)
    w.set_palette(pal)
    w.set_materials(mats)
    w.write_radiance(field)
    w.upload_blocks()
    return w, mats


def request_plan(seed, n_req, worlds_mats):
    """(kind, world, args) tuples: 70% radiance (probe subsets, 32/64/128
    rays), 30% glossy, several per world, in submission order."""
    import torch
    rng = np.random.default_rng(seed)
    plan = []
    for i in range(n_req):
        wi = int(rng.integers(len(SIZES)))
        wx, wy, wz = SIZES[wi]
        if rng.random() < 0.7:
            n = wx * wy * wz
            k = int(rng.integers(1, n + 1))
            flat = rng.permutation(n)[:k]
            req = np.stack([flat % wx, (flat // wx) % wy, flat // (wx * wy), np.zeros(k, int)], 1).astype(np.int32)
            req_d = torch.from_numpy(req).cuda() if rng.random() < 0.8 else None
            p = radiance_params(rays=int(rng.choice([32, 64, 128])), time_seed=i, alpha=0.15,
                                max_steps=int(rng.choice([0, 64])))
            plan.append(("r", wi, (p, req_d)))
        else:
            W, H = 32, 24
            mn, dp = synth.make_gbuffer(W, H, worlds_mats[wi], rng, depth_hi=60.0)
            plan.append(("g", wi, (torch.from_numpy(mn).cuda(), torch.from_numpy(dp).cuda(),
                                   camera_default(W, H))))
    return plan


def run_alone(ctx, plan):
    import torch
    worlds = [make(ctx, s, 100 + i)[0] for i, s in enumerate(SIZES)]
    outs = []
    for kind, wi, a in plan:
        if kind == "r":
            worlds[wi].radiance_update(a[1], a[0])
            outs.append(None)
        else:
            o = torch.zeros((a[1].shape[0], a[1].shape[1], 4), dtype=torch.float32, device="cuda")
            worlds[wi].glossy(a[0], a[1], a[2], 128.0, 128, out=o)
            outs.append(o)
    torch.cuda.synchronize()
    fields = [w.read_radiance() for w in worlds]
    for w in worlds:
        w.close()
    return fields, outs


def run_served(ctx, plan, budget=0, max_batch=8, streams=2, block=0, rays_per_lane=0, direct=False):
    import torch
    worlds, mats = [], []
    for i, s in enumerate(SIZES):
        w, m = make(ctx, s, 100 + i)
        worlds.append(w)
        mats.append(m)
    torch.cuda.synchronize()
    srv = Server(ctx, worlds, hbm_budget=budget, max_batch=max_batch, streams=streams, block=block,
                 rays_per_lane=rays_per_lane, direct=direct)
    outs, tickets = [], []
    for kind, wi, a in plan:
        if kind == "r":
            tickets.append(srv.submit(radiance_request(srv.ids[wi], a[0], a[1])))
            outs.append(None)
        else:
            o = torch.zeros((a[1].shape[0], a[1].shape[1], 4), dtype=torch.float32, device="cuda")
            tickets.append(srv.submit(glossy_request(srv.ids[wi], a[0], a[1], a[2], o, 128.0, 128)))
            outs.append(o)
    lat = [srv.wait(t) for t in tickets]
    srv.drain()
    stats = srv.stats()
    srv.close()
    fields = []
    for w in worlds:
        w.restore()
        fields.append(w.read_radiance())
        w.close()
    return fields, outs, stats, lat


def _mats():
    return [synth.make_materials(_mat_rng(100 + i)) for i in range(len(SIZES))]


def _mat_rng(seed):
    rng = np.random.default_rng(seed)
    # advance exactly as make() does before drawing materials
    size = SIZES[seed - 100]
    synth.make_world(size, rng, 32)
    synth.make_palette(rng, 32)
    return rng


@pytest.mark.parametrize("block,rays_per_lane", [(0, 0), (512, 2), (1024, 4)])
def test_served_equals_alone(gpu_ctx, block, rays_per_lane):
    plan = request_plan(5, 40, _mats())
    ref_fields, ref_outs = run_alone(gpu_ctx, plan)
    fields, outs, stats, lat = run_served(gpu_ctx, plan, max_batch=8, streams=2, block=block,
                                          rays_per_lane=rays_per_lane)
    assert stats["completed"] == len(plan)
    assert stats["launches"] < len(plan)  # requests were packed together
    for a, b in zip(fields, ref_fields):
        assert np.array_equal(a, b)
    for a, b in zip(outs, ref_outs):
        if a is not None:
            assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
    assert min(lat) > 0


def test_served_under_budget_evicts_and_restores(gpu_ctx):
    plan = request_plan(6, 30, _mats())
    ref_fields, ref_outs = run_alone(gpu_ctx, plan)
    # room for about two of the worlds at a time
    w, _ = make(gpu_ctx, SIZES[3], 1)
    per_world = w.resident()[1]
    w.close()
    fields, outs, stats, _ = run_served(gpu_ctx, plan, budget=int(2.2 * per_world), max_batch=4, streams=2)
    assert stats["evictions"] > 0 and stats["restores"] > 0
    assert stats["resident_bytes"] <= int(2.2 * per_world)
    for a, b in zip(fields, ref_fields):
        assert np.array_equal(a, b)
    for a, b in zip(outs, ref_outs):
        if a is not None:
            assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())


def test_plain_mode_one_request_per_launch(gpu_ctx):
    plan = request_plan(7, 12, _mats())
    ref_fields, _ = run_alone(gpu_ctx, plan)
    fields, _, stats, _ = run_served(gpu_ctx, plan, max_batch=1, streams=1)
    assert stats["launches"] == len(plan)
    for a, b in zip(fields, ref_fields):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("budget_worlds", [0, 2.2])
def test_direct_engine_equals_alone(gpu_ctx, budget_worlds):
    """The plain engine (one single-world call per request, queued with no
    wait between calls), with and without a budget that forces swaps."""
    plan = request_plan(8, 30, _mats())
    ref_fields, ref_outs = run_alone(gpu_ctx, plan)
    w, _ = make(gpu_ctx, SIZES[3], 1)
    per_world = w.resident()[1]
    w.close()
    budget = int(budget_worlds * per_world)
    fields, outs, stats, lat = run_served(gpu_ctx, plan, budget=budget, direct=True)
    assert stats["launches"] == len(plan) and stats["completed"] == len(plan)
    if budget:
        assert stats["evictions"] > 0 and stats["restores"] > 0
        assert stats["resident_bytes"] <= budget
    for a, b in zip(fields, ref_fields):
        assert np.array_equal(a, b)
    for a, b in zip(outs, ref_outs):
        if a is not None:
            assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
    assert min(lat) > 0


def test_budget_rejects_a_world_larger_than_itself(gpu_ctx):
    w, _ = make(gpu_ctx, (16, 16, 8), 2)
    try:
        with pytest.raises(Exception, match="budget"):
            Server(gpu_ctx, [w], hbm_budget=w.resident()[1] // 2)
    finally:
        w.close()


def test_evict_restore_roundtrip(gpu_ctx):
    import torch
    w, _ = make(gpu_ctx, (16, 16, 8), 3)
    t, _ = make(gpu_ctx, (16, 16, 8), 3)
    try:
        p = radiance_params(rays=64, time_seed=1, alpha=0.1)
        w.radiance_update(None, p)
        t.radiance_update(None, p)
        torch.cuda.synchronize()
        before = w.read_radiance()
        w.evict()
        assert w.resident()[0] is False
        with pytest.raises(Exception):
            w.read_radiance()
        w.restore()
        assert np.array_equal(w.read_radiance(), before)
        p.time_seed = 2
        w.radiance_update(None, p)
        t.radiance_update(None, p)
        torch.cuda.synchronize()
        assert np.array_equal(w.read_radiance(), t.read_radiance())
    finally:
        w.close()
        t.close()


def test_submit_rejects_what_a_launch_would(gpu_ctx):
    """Requests a launch would reject are refused at submit, so they cannot
    fail the valid requests packed with them."""
    import torch
    w, mats = make(gpu_ctx, (16, 16, 8), 9)
    srv = Server(gpu_ctx, [w], max_batch=8, streams=1)
    try:
        wid = srv.ids[0]
        bad = [radiance_params(rays=64, size=0), radiance_params(rays=64, size=2, shift=2),
               radiance_params(rays=64, max_dist=0.0), radiance_params(rays=64, max_steps=-1),
               radiance_params(rays=48)]
        for p in bad:
            with pytest.raises(Exception):
                srv.submit(radiance_request(wid, p))
        mn, dp = synth.make_gbuffer(16, 8, mats, np.random.default_rng(1), depth_hi=60.0)
        mn_d, dp_d = torch.from_numpy(mn).cuda(), torch.from_numpy(dp).cuda()
        out = torch.zeros((8, 16, 4), dtype=torch.float32, device="cuda")
        with pytest.raises(Exception):
            srv.submit(glossy_request(wid, mn_d, dp_d, camera_default(16, 8), out, 128.0, -1))
        with pytest.raises(Exception):
            srv.submit(glossy_request(wid, mn_d, dp_d, camera_default(16, 8), out, -1.0, 64))
        tickets = [srv.submit(radiance_request(wid, radiance_params(rays=64, time_seed=k))) for k in range(3)]
        tickets.append(srv.submit(glossy_request(wid, mn_d, dp_d, camera_default(16, 8), out, 128.0, 64)))
        for t in tickets:
            assert srv.wait(t) > 0
        assert srv.stats()["completed"] == 4
    finally:
        srv.close()
        w.close()


def test_launch_restoring_more_than_sixteen_worlds(gpu_ctx):
    """40 evicted worlds, one request each, 32 per launch: the first launch
    restores 32 worlds at once (their blocks are set up in launches of 16);
    every world ends as its update alone leaves it."""
    import torch
    from lum_amd.api import World
    n, size = 40, (8, 8, 8)

    def build(i):
        rng = np.random.default_rng(500 + i)
        blocks = synth.make_world(size, rng, 32)
        pal = synth.make_palette(rng, 32)
        mats = synth.make_materials(rng)
        field = synth.make_field(size, rng)
        w = World(gpu_ctx, size, static_blocks=32, palette_slots=32, radiance_format=RADIANCE_F16)
        w.set_blocks(blocks)
        w.set_palette(pal)
        w.set_materials(mats)
        w.write_radiance(field)
        w.upload_blocks()
        return w

    rng = np.random.default_rng(77)
    reqs = []
    for i in range(n):
        k = int(rng.integers(1, 513))
        flat = rng.permutation(512)[:k]
        req = np.stack([flat % 8, (flat // 8) % 8, flat // 64, np.zeros(k, int)], 1).astype(np.int32)
        reqs.append((radiance_params(rays=64, time_seed=i, alpha=0.2, max_steps=64), torch.from_numpy(req).cuda()))
    alone = []
    for i in range(n):
        w = build(i)
        w.radiance_update(reqs[i][1], reqs[i][0])
        torch.cuda.synchronize()
        alone.append(w.read_radiance())
        w.close()
    worlds = [build(i) for i in range(n)]
    per_world = worlds[0].resident()[1]
    torch.cuda.synchronize()
    for w in worlds:
        w.evict()
    srv = Server(gpu_ctx, worlds, hbm_budget=n * per_world, max_batch=32, streams=2)
    tickets = [srv.submit(radiance_request(srv.ids[i], reqs[i][0], reqs[i][1])) for i in range(n)]
    for t in tickets:
        srv.wait(t)
    srv.drain()
    stats = srv.stats()
    srv.close()
    assert stats["completed"] == n and stats["restores"] == n and stats["evictions"] == 0
    assert stats["launches"] < n
    for i, w in enumerate(worlds):
        w.restore()
        assert np.array_equal(w.read_radiance(), alone[i]), i
        w.close()
