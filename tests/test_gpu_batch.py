"""Batched (many worlds per launch) passes and graph-captured chains.

Each world's result in a batch must equal, bit for bit, the same world
updated alone with lumx_radiance_update / lumx_glossy; a replayed graph must
equal the same passes issued eagerly, on every replay.
"""
import numpy as np
import pytest

from lum_amd import synth
from lum_amd.api import World, Model, Batch, Graph, placement, camera_default, radiance_params
import tests._oracle as orc

pytestmark = pytest.mark.gpu


def make_world(ctx, size, seed, n_palette=32):
    rng = np.random.default_rng(seed)
    blocks = synth.make_world(size, rng, n_palette)
    pal = synth.make_palette(rng, n_palette)
    mats = synth.make_materials(rng)
    field = synth.make_field(size, rng)
    w = World(ctx, size, static_blocks=n_palette)
    w.set_blocks(blocks)
    w.set_palette(pal)
    w.set_materials(mats)
    w.write_radiance(field)
    w.upload_blocks()
    return w, (blocks, pal, mats, field)


SPECS = [((16, 16, 8), 32, 1), ((24, 20, 10), 64, 2), ((20, 12, 6), 128, 3), ((33, 17, 9), 64, 4), ((8, 8, 8), 32, 5)]


def test_batch_radiance_equals_single_worlds(gpu_ctx):
    import torch
    batch_worlds, single_worlds, datas = [], [], []
    for size, rays, seed in SPECS:
        w, d = make_world(gpu_ctx, size, seed)
        batch_worlds.append(w)
        datas.append(d)
        s, _ = make_world(gpu_ctx, size, seed)
        s.set_tuning(variant=0)
        single_worlds.append(s)
    b = Batch(gpu_ctx, batch_worlds)
    try:
        params = [radiance_params(rays=r, time_seed=10 + i, alpha=0.2, max_steps=64 if i % 2 else 0)
                  for i, (_, r, _) in enumerate(SPECS)]
        # world 1 takes a request list with size/shift; the others every probe
        wx, wy, wz = SPECS[1][0]
        rng = np.random.default_rng(9)
        flat = rng.choice(wx * wy * wz, size=500, replace=False)
        req = np.stack([flat % wx, (flat // wx) % wy, flat // (wx * wy), np.zeros(500, int)], 1).astype(np.int32)
        params[1].size, params[1].shift = 2, 1
        req_d = torch.from_numpy(req).cuda()
        requests = [None, req_d, None, None, None]
        for frame in range(2):
            b.radiance(params, requests)
            for i, s in enumerate(single_worlds):
                s.radiance_update(requests[i], params[i])
            for p in params:
                p.time_seed += 1
        torch.cuda.synchronize()
        for i, (bw, sw) in enumerate(zip(batch_worlds, single_worlds)):
            assert np.array_equal(bw.read_radiance(), sw.read_radiance()), f"world {i}"
        # one world against the oracle too (two frames)
        blocks, pal, mats, field = datas[2]
        full = np.zeros((batch_worlds[2].palette_slots, 16, 16, 16), np.uint8)
        full[: pal.shape[0]] = pal
        wd = orc.WorldData(blocks, full, mats)
        ref = orc.radiance_update(wd, field, rays=128, time_seed=12, alpha=0.2)
        ref = orc.radiance_update(wd, ref, rays=128, time_seed=13, alpha=0.2)
        got = batch_worlds[2].read_radiance()
        assert (np.abs(got - ref) / np.maximum(np.abs(ref), 1e-3)).max() <= 1e-3
    finally:
        b.close()
        for w in batch_worlds + single_worlds:
            w.close()


def test_batch_compacted_subset_updates_equal_single(gpu_ctx):
    """Subset updates large enough for the batch to compact them (claimed
    probes in texel order) beside a small one it does not compact: each
    world equals its own single-world update run without compaction, bit for
    bit, with repeated and outside entries, over two frames."""
    import torch
    specs = [((16, 16, 16), 4096, 1500, 61), ((24, 20, 10), 3000, 2500, 62), ((12, 10, 6), 300, 40, 63)]
    bw, sw, reqs = [], [], []
    for size, n, distinct, seed in specs:
        b, _ = make_world(gpu_ctx, size, seed)
        s, _ = make_world(gpu_ctx, size, seed)
        s.set_tuning(variant=0, compact=-1)
        bw.append(b)
        sw.append(s)
        wx, wy, wz = size
        rng = np.random.default_rng(seed + 7)
        flat = rng.integers(0, distinct, size=n)
        q = np.stack([flat % wx, (flat // wx) % wy, flat // (wx * wy), np.zeros(n, int)], 1).astype(np.int32)
        q[3] = [wx, 0, 0, 0]
        reqs.append(torch.from_numpy(q).cuda())
    batch = Batch(gpu_ctx, bw)
    try:
        params = [radiance_params(rays=r, time_seed=40 + i, alpha=0.3, max_steps=64)
                  for i, r in enumerate((64, 128, 32))]
        for frame in range(2):
            batch.radiance(params, reqs)
            for i, s in enumerate(sw):
                s.radiance_update(reqs[i], params[i])
            for p in params:
                p.time_seed += 1
        torch.cuda.synchronize()
        for i, (x, y) in enumerate(zip(bw, sw)):
            assert np.array_equal(x.read_radiance(), y.read_radiance()), i
    finally:
        batch.close()
        for w in bw + sw:
            w.close()


def test_batch_compacted_updates_in_graph_replays(gpu_ctx):
    """Compacted subset updates recorded into a graph (claims, compaction,
    update, scatter) replay to the same fields as the same updates issued
    eagerly, on every replay, with eager work between replays."""
    import torch
    specs = [((16, 16, 16), 4096, 1500, 71), ((24, 20, 10), 3000, 2500, 72)]
    gw, ew, reqs = [], [], []
    for size, n, distinct, seed in specs:
        g, _ = make_world(gpu_ctx, size, seed)
        e, _ = make_world(gpu_ctx, size, seed)
        gw.append(g)
        ew.append(e)
        wx, wy, wz = size
        rng = np.random.default_rng(seed + 7)
        flat = rng.integers(0, distinct, size=n)
        q = np.stack([flat % wx, (flat // wx) % wy, flat // (wx * wy), np.zeros(n, int)], 1).astype(np.int32)
        reqs.append(torch.from_numpy(q).cuda())
    gb, eb = Batch(gpu_ctx, gw), Batch(gpu_ctx, ew)
    params = [radiance_params(rays=64, time_seed=50 + i, alpha=0.3, max_steps=64) for i in range(len(specs))]
    stream = torch.cuda.Stream()
    graph = Graph()
    try:
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            graph.begin()
            gb.radiance(params, reqs)
            graph.end()
        torch.cuda.synchronize()
        for rep in range(3):
            with torch.cuda.stream(stream):
                graph.launch()
            eb.radiance(params, reqs)
            torch.cuda.synchronize()
            junk = torch.ones(1 << 20, device="cuda") * rep
            torch.cuda.synchronize()
            del junk
            for i in range(len(specs)):
                assert np.array_equal(gw[i].read_radiance(), ew[i].read_radiance()), (rep, i)
    finally:
        graph.close()
        gb.close()
        eb.close()
        for w in gw + ew:
            w.close()


def test_batch_frame_counter_offsets_seed(gpu_ctx):
    import torch
    w, _ = make_world(gpu_ctx, (16, 16, 8), 21)
    s, _ = make_world(gpu_ctx, (16, 16, 8), 21)
    b = Batch(gpu_ctx, [w])
    try:
        b.set_frame(7)
        b.radiance([radiance_params(rays=64, time_seed=3, alpha=0.1)])
        s.radiance_update(None, radiance_params(rays=64, time_seed=10, alpha=0.1))
        torch.cuda.synchronize()
        assert np.array_equal(w.read_radiance(), s.read_radiance())
    finally:
        b.close()
        w.close()
        s.close()


SHAPES = [(256, 1), (512, 1), (1024, 1), (256, 2), (512, 2), (1024, 4), (256, 4), (0, 0)]


def test_batch_autotune_and_shapes(gpu_ctx):
    """The autotuner's trial launches leave every field as it was, and each
    launch shape it chooses from (workgroup size x rays per lane) gives the
    single-world result bit for bit (tail workgroups of odd probe counts with
    32, 64 and 128 rays included)."""
    import torch
    batch_worlds, single_worlds = [], []
    for size, rays, seed in SPECS:
        w, _ = make_world(gpu_ctx, size, 40 + seed)
        batch_worlds.append(w)
        s, _ = make_world(gpu_ctx, size, 40 + seed)
        s.set_tuning(variant=0)
        single_worlds.append(s)
    b = Batch(gpu_ctx, batch_worlds)
    try:
        params = [radiance_params(rays=r, time_seed=5 + i, alpha=0.3) for i, (_, r, _) in enumerate(SPECS)]
        before = [w.read_radiance() for w in batch_worlds]
        wx, wy, wz = SPECS[1][0]
        flat = np.random.default_rng(8).choice(wx * wy * wz, size=300, replace=False)
        req = np.stack([flat % wx, (flat // wx) % wy, flat // (wx * wy), np.zeros(300, int)], 1).astype(np.int32)
        req_d = torch.from_numpy(req).cuda()
        block, rpl, ms = b.autotune(params, requests=[None, req_d, None, None, None], reps=2)
        assert block in (256, 512, 1024) and rpl in (1, 2, 4) and ms > 0
        block, rpl, ms = b.autotune(params, reps=2)
        assert block in (256, 512, 1024) and rpl in (1, 2, 4) and ms > 0
        for w, f in zip(batch_worlds, before):
            assert np.array_equal(w.read_radiance(), f)
        for blk, k in SHAPES:
            b.set_shape(blk, k)
            b.radiance(params)
            for s, p in zip(single_worlds, params):
                s.radiance_update(None, p)
            torch.cuda.synchronize()
            for i, (bw, sw) in enumerate(zip(batch_worlds, single_worlds)):
                assert np.array_equal(bw.read_radiance(), sw.read_radiance()), f"shape {blk}/{k} world {i}"
            for p in params:
                p.time_seed += 1
        with pytest.raises(Exception):
            b.set_shape(384, 1)
        with pytest.raises(Exception):
            b.set_shape(256, 3)
    finally:
        b.close()
        for w in batch_worlds + single_worlds:
            w.close()


def test_batch_glossy_equals_single(gpu_ctx):
    import torch
    worlds, datas = [], []
    for size, _, seed in SPECS[:3]:
        w, d = make_world(gpu_ctx, size, seed)
        worlds.append(w)
        datas.append(d)
    b = Batch(gpu_ctx, worlds)
    try:
        rng = np.random.default_rng(3)
        dims = [(64, 36), (40, 30), (17, 9)]
        mns, dps, cams, outs, singles = [], [], [], [], []
        for (W, H), d in zip(dims, datas):
            mn, dp = synth.make_gbuffer(W, H, d[2], rng, depth_hi=60.0)
            mns.append(torch.from_numpy(mn).cuda())
            dps.append(torch.from_numpy(dp).cuda())
            cams.append(camera_default(W, H))
            outs.append(torch.zeros((H, W, 4), dtype=torch.float32, device="cuda"))
            singles.append(torch.zeros((H, W, 4), dtype=torch.float32, device="cuda"))
        b.glossy(mns, dps, cams, outs, max_dist=128.0, max_steps=128)
        for i, w in enumerate(worlds):
            w.glossy(mns[i], dps[i], cams[i], max_dist=128.0, max_steps=128, out=singles[i])
        torch.cuda.synchronize()
        for o, s in zip(outs, singles):
            assert torch.equal(o, s)
    finally:
        b.close()
        for w in worlds:
            w.close()


def _chain(worlds, batch, models, placements, gbufs, outs, params):
    """map -> radiance -> glossy -> advance, twice (one even/odd field pair)."""
    for w, ms, pl in zip(worlds, models, placements):
        w.map_models_placed(ms, pl)
    for _ in range(2):
        batch.radiance(params)
        batch.glossy([g[0] for g in gbufs], [g[1] for g in gbufs], [g[2] for g in gbufs], outs,
                     max_dist=128.0, max_steps=128)
        batch.advance_frame()


def test_graph_chain_replays_match_eager(gpu_ctx):
    import torch
    # 17 x 11 x 3 = 561 blocks: 1122 bytes of block ids, not a multiple of
    # 16, so the capture's staging of the copy list and counter is offset
    sizes = [((16, 16, 8), 11), ((20, 16, 8), 12), ((17, 11, 3), 13)]
    rng = np.random.default_rng(77)

    def build():
        worlds, models, placements, gbufs, outs = [], [], [], [], []
        for size, seed in sizes:
            w, d = make_world(gpu_ctx, size, seed)
            r = np.random.default_rng(seed + 100)
            vox, quats, shifts = synth.make_models(r, n=2, lo=8, hi=16, world_size=size)
            ms = [Model(gpu_ctx, v) for v in vox]
            worlds.append(w)
            models.append(ms)
            placements.append([placement(q, s, m.size) for q, s, m in zip(quats, shifts, ms)])
            mn, dp = synth.make_gbuffer(48, 32, d[2], np.random.default_rng(seed + 200), depth_hi=60.0)
            gbufs.append((torch.from_numpy(mn).cuda(), torch.from_numpy(dp).cuda(), camera_default(48, 32)))
            outs.append(torch.zeros((32, 48, 4), dtype=torch.float32, device="cuda"))
        return worlds, models, placements, gbufs, outs

    params = [radiance_params(rays=64, time_seed=5, alpha=0.1, max_steps=64) for _ in sizes]
    gw, gm, gp, gg, go = build()
    ew, em, ep, eg, eo = build()
    gb, eb = Batch(gpu_ctx, gw), Batch(gpu_ctx, ew)
    stream = torch.cuda.Stream()
    graph = Graph()
    try:
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            graph.begin()
            _chain(gw, gb, gm, gp, gg, go, params)
            graph.end()
        torch.cuda.synchronize()
        for rep in range(3):
            with torch.cuda.stream(stream):
                graph.launch()
            _chain(ew, eb, em, ep, eg, eo, params)
            torch.cuda.synchronize()
            # eager work of the program between replays
            junk = torch.ones(1 << 20, device="cuda") * rep
            torch.cuda.synchronize()
            del junk
            for i in range(len(sizes)):
                assert np.array_equal(gw[i].read_radiance(), ew[i].read_radiance()), (rep, i)
                assert torch.equal(go[i], eo[i]), (rep, i)
                c = ew[i].palette_counter()
                assert np.array_equal(gw[i].get_palette(0, c), ew[i].get_palette(0, c)), (rep, i)
    finally:
        graph.close()
        gb.close()
        eb.close()
        for ms in gm + em:
            for m in ms:
                m.close()
        for w in gw + ew:
            w.close()
