"""Batched map (lumx_batch_map_models) against single-world maps.

Bar: after the batched map every world holds, bit for bit, the palette slots
(including slots a previous, larger map used, which must read as zero) and
gives the radiance (plain and tiled kernels: block ids, empty-block bits and
palette occupancy bits) of the same world mapped alone with
lumx_map_models_placed; over successive maps, inside replayed graphs, and when
one world's palette overflows.
"""
import numpy as np
import pytest

from lum_amd import synth
from lum_amd.api import Batch, Graph, Model, placement, radiance_params
from lum_amd._lib import LumxError

from tests.test_gpu_parity import make_case, gpu_world

pytestmark = pytest.mark.gpu

# (size, seed, models, palette slots, dynamic origin ids, model partly below 0)
SPECS = [
    ((16, 16, 8), 31, 3, 0, False, False),
    ((17, 11, 3), 32, 0, 0, False, False),
    ((24, 20, 8), 33, 5, 0, False, True),
    ((16, 16, 8), 34, 2, 2048, True, False),
    ((20, 12, 6), 35, 4, 0, False, False),
]


def make_twins(ctx, specs):
    """Two identical worlds per spec and one set of models for both."""
    batch_w, single_w, models, pls = [], [], [], []
    for size, seed, n_models, slots, dynamic, negative in specs:
        blocks, pal, mats, field = make_case(size, seed, n_palette=32)
        if dynamic:
            blocks[0, 4:8, 4:8] = 40
        batch_w.append(gpu_world(ctx, blocks, pal, mats, field, slots=slots))
        single_w.append(gpu_world(ctx, blocks, pal, mats, field, slots=slots))
        rng = np.random.default_rng(seed + 1000)
        vox, quats, shifts = synth.make_models(rng, n=n_models, lo=8, hi=20, fill=0.3, world_size=size)
        ms = [Model(ctx, v) for v in vox]
        pl = [placement(q, s, m.size) for q, s, m in zip(quats, shifts, ms)]
        if negative and ms:
            pl[0] = placement(quats[0], [-6.0, 5.0, -3.0], ms[0].size)
        models.append(ms)
        pls.append(pl)
    return batch_w, single_w, models, pls


def close_all(batch, worlds, models):
    if batch is not None:
        batch.close()
    for ms in models:
        for m in ms:
            m.close()
    for w in worlds:
        w.close()


def same_world(a, b, upto, tag):
    c = b.palette_counter()
    assert a.palette_counter() == c, tag
    n = min(max(c, upto), b.palette_slots)
    assert np.array_equal(a.get_palette(0, n), b.get_palette(0, n)), tag
    return c


def same_radiance(a, b, variant, seed, tag):
    for w in (a, b):
        w.set_tuning(variant=variant)
        w.radiance_update(None, radiance_params(rays=64, time_seed=seed, alpha=0.3, max_steps=64))
    assert np.array_equal(a.read_radiance(), b.read_radiance()), tag


def test_batch_map_equals_single_maps(gpu_ctx):
    bw, sw, models, pls = make_twins(gpu_ctx, SPECS)
    b = Batch(gpu_ctx, bw)
    try:
        b.map_models(models, pls)
        counters = []
        for i, (x, y) in enumerate(zip(bw, sw)):
            y.map_models_placed(models[i], pls[i])
            counters.append(same_world(x, y, 0, ("first map", i)))
            same_radiance(x, y, 1, 3, ("first map, tiled", i))
        # a smaller second map: the slots the first one used read as zero again
        small = [ms[:1] for ms in models]
        small_pl = [p[:1] for p in pls]
        b.map_models(small, small_pl)
        for i, (x, y) in enumerate(zip(bw, sw)):
            y.map_models_placed(small[i], small_pl[i])
            same_world(x, y, counters[i], ("second map", i))
            same_radiance(x, y, 0, 4, ("second map, lane", i))
            same_radiance(x, y, 1, 5, ("second map, tiled", i))
    finally:
        close_all(b, bw + sw, models)


def test_batch_map_graph_replays_match_eager(gpu_ctx):
    import torch
    specs = SPECS[:3]
    bw, sw, models, pls = make_twins(gpu_ctx, specs)
    b = Batch(gpu_ctx, bw)
    stream = torch.cuda.Stream()
    graph = Graph()
    params = [radiance_params(rays=64, time_seed=7, alpha=0.2, max_steps=64) for _ in specs]
    try:
        for w in bw + sw:
            w.set_tuning(variant=0)
        # each replay follows an eager map of more models: the replay restores
        # the blocks and every slot they reference (slots the eager map used
        # beyond the recorded map's are never referenced by its blocks)
        big = [ms + ms for ms in models]
        big_pl = [p + p[::-1] for p in pls]
        torch.cuda.synchronize()
        # two radiance updates: a replay reads and writes the fields recorded
        # at capture, so the recorded chain covers one even/odd pair
        with torch.cuda.stream(stream):
            graph.begin()
            b.map_models(models, pls)
            b.radiance(params)
            b.radiance(params)
            graph.end()
        torch.cuda.synchronize()
        for rep in range(3):
            b.map_models(big, big_pl)
            torch.cuda.synchronize()
            with torch.cuda.stream(stream):
                graph.launch()
            torch.cuda.synchronize()
            junk = torch.ones(1 << 20, device="cuda") * rep
            torch.cuda.synchronize()
            del junk
            for i, y in enumerate(sw):
                y.map_models_placed(big[i], big_pl[i])
                y.map_models_placed(models[i], pls[i])
                y.radiance_update(None, params[i])
                y.radiance_update(None, params[i])
            torch.cuda.synchronize()
            for i, (x, y) in enumerate(zip(bw, sw)):
                n = y.palette_counter()
                assert np.array_equal(x.get_palette(0, n), y.get_palette(0, n)), (rep, i)
                assert np.array_equal(x.read_radiance(), y.read_radiance()), (rep, i)
    finally:
        graph.close()
        close_all(b, bw + sw, models)


def test_batch_map_overflow_commits_earlier_worlds(gpu_ctx):
    specs = [
        ((16, 16, 8), 41, 3, 0, False, False),
        ((16, 16, 8), 42, 4, 40, False, False),   # 8 dynamic slots: overflows
        ((20, 12, 6), 43, 3, 0, False, False),
    ]
    bw, sw, models, pls = make_twins(gpu_ctx, specs)
    # the overflowing world's models are large
    rng = np.random.default_rng(44)
    vox, quats, shifts = synth.make_models(rng, n=4, lo=20, hi=32, fill=0.3, world_size=specs[1][0])
    for m in models[1]:
        m.close()
    models[1] = [Model(gpu_ctx, v) for v in vox]
    pls[1] = [placement(q, s, m.size) for q, s, m in zip(quats, shifts, models[1])]
    b = Batch(gpu_ctx, bw)
    try:
        with pytest.raises(LumxError):
            b.map_models(models, pls)
        sw[0].map_models_placed(models[0], pls[0])
        with pytest.raises(LumxError):
            sw[1].map_models_placed(models[1], pls[1])
        for i, (x, y) in enumerate(zip(bw, sw)):
            same_world(x, y, 0, ("overflow", i))
            same_radiance(x, y, 0, 6, ("overflow", i))
        assert bw[1].palette_counter() == 32 and bw[2].palette_counter() == 32
        # the batch maps again once the overflowing world's models fit
        models_ok = [models[0], [], models[2]]
        pls_ok = [pls[0], [], pls[2]]
        b.map_models(models_ok, pls_ok)
        for i, y in enumerate(sw):
            y.map_models_placed(models_ok[i], pls_ok[i])
        for i, (x, y) in enumerate(zip(bw, sw)):
            same_world(x, y, 0, ("after overflow", i))
            same_radiance(x, y, 1, 7, ("after overflow", i))
    finally:
        close_all(b, bw + sw, models)


def test_batch_map_rejects_bad_arguments(gpu_ctx):
    bw, sw, models, pls = make_twins(gpu_ctx, SPECS[:2])
    b = Batch(gpu_ctx, bw)
    try:
        import ctypes as C
        from lum_amd._lib import Placement
        h0 = (C.c_void_p * 1)(models[0][0].handle.value)
        p0 = (Placement * 1)(pls[0][0])
        cnt = (C.c_int * 2)(1, 1)
        hs = (C.c_void_p * 2)(C.cast(h0, C.c_void_p), None)   # world 1: one model, no model list
        ps = (C.c_void_p * 2)(C.cast(p0, C.c_void_p), C.cast(p0, C.c_void_p))
        assert b.lib.lumx_batch_map_models(b.handle, cnt, hs, ps, None) != 0
        # nothing was mapped: the worlds still equal their twins
        for i, (x, y) in enumerate(zip(bw, sw)):
            same_world(x, y, 0, ("rejected", i))
    finally:
        close_all(b, bw + sw, models)
