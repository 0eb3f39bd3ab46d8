"""Device blockify (lumx_map_models_device) against the host path and the
oracle.

Bar: bit for bit the same blockified world, palette counter, palette slots
(including the slots a previous, larger map used, which must read as zero)
and radiance afterwards, over sequences of maps that mix both paths.
"""
import numpy as np
import pytest

from lum_amd import synth
from lum_amd.api import World, Model, Graph, placement, radiance_params
from lum_amd._lib import LumxError
import tests._oracle as orc

from tests.test_gpu_parity import make_case, gpu_world

pytestmark = pytest.mark.gpu


def twin(ctx, size, seed, static=32, slots=0):
    blocks, pal, mats, field = make_case(size, seed, n_palette=static)
    a = gpu_world(ctx, blocks, pal, mats, field, slots=slots)
    b = gpu_world(ctx, blocks, pal, mats, field, slots=slots)
    return a, b, (blocks, pal, mats, field)


def model_set(ctx, rng, n, size, lo=8, hi=24):
    vox, quats, shifts = synth.make_models(rng, n=n, lo=lo, hi=hi, fill=0.3, world_size=size)
    ms = [Model(ctx, v) for v in vox]
    return ms, [placement(q, s, m.size) for q, s, m in zip(quats, shifts, ms)]


def same_state(dev, host, upto):
    assert np.array_equal(dev.get_blocks(True), host.get_blocks(True))
    c = host.palette_counter()
    assert dev.palette_counter() == c
    n = max(c, upto)
    assert np.array_equal(dev.get_palette(0, n), host.get_palette(0, n))
    return c


def test_device_map_equals_host_map_and_oracle(gpu_ctx):
    size = (24, 20, 8)
    dev, host, (blocks, pal, mats, field) = twin(gpu_ctx, size, 5)
    rng = np.random.default_rng(6)
    ms, pls = model_set(gpu_ctx, rng, 6, size)
    try:
        dev.map_models_device(ms, pls)
        host.map_models_placed(ms, pls)
        c = same_state(dev, host, 0)
        # the oracle's blockify order
        cur = blocks.copy()
        copy_src = np.zeros(dev.palette_slots, np.int32)
        k = 32
        for p in pls:
            k = orc.blockify(size, cur, 32, k, list(p.aabb_min), list(p.aabb_max), copy_src)
        assert k == c and np.array_equal(dev.get_blocks(True), cur)
    finally:
        for m in ms:
            m.close()
        dev.close()
        host.close()


def test_map_sequences_mixing_paths_keep_slots_consistent(gpu_ctx):
    """Frames with more, then fewer, then no models; the device world maps
    on the device except every third frame; the twin always on the host."""
    import torch
    size = (32, 24, 8)
    dev, host, _ = twin(gpu_ctx, size, 9)
    rng = np.random.default_rng(10)
    sets = [model_set(gpu_ctx, rng, n, size) for n in (8, 3, 0, 5, 1, 6)]
    try:
        hi = 0
        for f, (ms, pls) in enumerate(sets):
            if f % 3 == 2:
                dev.map_models_placed(ms, pls)
            else:
                dev.map_models_device(ms, pls)
            host.map_models_placed(ms, pls)
            p = radiance_params(rays=32, time_seed=f, alpha=0.2, max_steps=64)
            dev.radiance_update(None, p)
            host.radiance_update(None, p)
            torch.cuda.synchronize()
            c = same_state(dev, host, hi)
            hi = max(hi, c)
            assert np.array_equal(dev.read_radiance(), host.read_radiance()), f
    finally:
        for ms, _ in sets:
            for m in ms:
                m.close()
        dev.close()
        host.close()


def test_device_map_with_dynamic_origin_ids_and_outside_models(gpu_ctx):
    """Origin ids at or above the static count are never reallocated; a model
    outside the world claims the clamped corner block, as the host path."""
    size = (16, 16, 8)
    blocks, pal, mats, field = make_case(size, 14, n_palette=32)
    blocks[0, 8:12, 8:12] = 40                   # dynamic ids in the origin world
    dev = gpu_world(gpu_ctx, blocks, pal, mats, field, slots=2048)
    host = gpu_world(gpu_ctx, blocks, pal, mats, field, slots=2048)
    rng = np.random.default_rng(15)
    ms, pls = model_set(gpu_ctx, rng, 3, size)
    outside = placement([1.0, 0.0, 0.0, 0.0], [-500.0, -500.0, -500.0], ms[0].size)
    try:
        dev.map_models_device(ms + [ms[0]], pls + [outside])
        host.map_models_placed(ms + [ms[0]], pls + [outside])
        same_state(dev, host, 0)
    finally:
        for m in ms:
            m.close()
        dev.close()
        host.close()


def test_device_map_overflow_is_reported(gpu_ctx):
    size = (16, 16, 8)
    blocks, pal, mats, field = make_case(size, 16, n_palette=32)
    w = gpu_world(gpu_ctx, blocks, pal, mats, field, slots=40)    # room for 8 dynamic slots
    rng = np.random.default_rng(17)
    ms, pls = model_set(gpu_ctx, rng, 4, size, lo=20, hi=32)
    try:
        w.map_models_device(ms, pls)
        with pytest.raises(LumxError):
            w.palette_counter()
        cur = w.get_blocks(True)
        assert cur.max() < 40
    finally:
        for m in ms:
            m.close()
        w.close()


def test_device_map_in_graph_replays(gpu_ctx):
    """A device map recorded into a graph (after one eager call sized its
    scratch) replays to the eager result, with eager work in between."""
    import torch
    size = (24, 24, 8)
    dev, host, _ = twin(gpu_ctx, size, 21)
    rng = np.random.default_rng(22)
    ms, pls = model_set(gpu_ctx, rng, 5, size)
    stream = torch.cuda.Stream()
    graph = Graph()
    try:
        dev.map_models_device(ms, pls)
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            graph.begin()
            dev.map_models_device(ms, pls)
            graph.end()
        torch.cuda.synchronize()
        host.map_models_placed(ms, pls)
        for rep in range(3):
            with torch.cuda.stream(stream):
                graph.launch()
            torch.cuda.synchronize()
            junk = torch.ones(1 << 20, device="cuda") * rep
            torch.cuda.synchronize()
            del junk
            same_state(dev, host, 0)
    finally:
        graph.close()
        for m in ms:
            m.close()
        dev.close()
        host.close()
