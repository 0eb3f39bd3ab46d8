"""The no-allocation device map (lumx_map_models_device_noalloc) against
lumx_map_models_device, through the C ABI.

Bar: bit for bit the same blockified world, palette counter and palette
slots (including slots a previous, larger map used, which must read as zero)
on two maps in a row, the second with fewer models and smaller boxes; and
the stated limits rejected with LUMX_EINVAL before any launch.
"""
import ctypes as C

import numpy as np
import pytest

from lum_amd import synth
from lum_amd.api import Model, placement, _stream
from lum_amd._lib import Placement, LUMX_EINVAL

from tests.test_gpu_parity import make_case, gpu_world

pytestmark = pytest.mark.gpu


def _map(world, models, placements, noalloc):
    n = len(models)
    hs = (C.c_void_p * max(n, 1))(*[m.handle.value for m in models])
    ps = (Placement * max(n, 1))(*placements)
    fn = world.lib.lumx_map_models_device_noalloc if noalloc else world.lib.lumx_map_models_device
    return fn(world.handle, n, hs, ps, _stream())


def _models(ctx, rng, n, size, lo, hi):
    vox, quats, shifts = synth.make_models(rng, n=n, lo=lo, hi=hi, fill=0.3, world_size=size)
    ms = [Model(ctx, v) for v in vox]
    return ms, [placement(q, s, m.size) for q, s, m in zip(quats, shifts, ms)]


def test_noalloc_map_equals_device_map(gpu_ctx):
    import torch
    size = (32, 24, 8)
    blocks, pal, mats, field = make_case(size, 21, n_palette=32)
    ref = gpu_world(gpu_ctx, blocks, pal, mats, field)
    got = gpu_world(gpu_ctx, blocks, pal, mats, field)
    rng = np.random.default_rng(22)
    sets = [_models(gpu_ctx, rng, 10, size, 12, 40), _models(gpu_ctx, rng, 3, size, 8, 16)]
    try:
        hi = 0
        for k, (ms, pls) in enumerate(sets):
            print(f"map {k}: {len(ms)} models", flush=True)
            assert _map(ref, ms, pls, False) == 0
            assert _map(got, ms, pls, True) == 0, got.lib.lumx_last_error()
            torch.cuda.synchronize()
            c = ref.palette_counter()
            assert got.palette_counter() == c, k
            assert c > 32, k
            assert np.array_equal(got.get_blocks(True), ref.get_blocks(True)), k
            upto = max(c, hi)
            assert np.array_equal(got.get_palette(0, upto), ref.get_palette(0, upto)), k
            hi = upto
    finally:
        for ms, _ in sets:
            for m in ms:
                m.close()
        ref.close()
        got.close()


def test_noalloc_map_rejects_beyond_its_limits(gpu_ctx):
    size = (16, 16, 8)
    blocks, pal, mats, field = make_case(size, 23, n_palette=16)
    w = gpu_world(gpu_ctx, blocks, pal, mats, field)
    rng = np.random.default_rng(24)
    ms, pls = _models(gpu_ctx, rng, 1, size, 8, 12)
    try:
        many = ms * 65
        assert _map(w, many, pls * 65, True) == LUMX_EINVAL
        assert b"at most 64" in w.lib.lumx_last_error()
        # 64 copies of one model are within the limit
        assert _map(w, ms * 64, pls * 64, True) == 0, w.lib.lumx_last_error()
    finally:
        for m in ms:
            m.close()
        w.close()
    # Model boxes over more blocks than the scratch holds: 64 boxes, each the
    # whole 128x128x32 world (33,554,432 elements; the ballot words alone
    # take 4 MiB).  Rejected before any launch.
    size = (128, 128, 32)
    big = gpu_world(gpu_ctx, np.zeros(size[::-1], np.int16), pal, mats, np.zeros(size[::-1] + (4,), np.float32),
                    slots=64)
    ms, pls = _models(gpu_ctx, rng, 1, size, 8, 12)
    try:
        whole = Placement.from_buffer_copy(pls[0])
        for a in range(3):
            whole.aabb_min[a] = 0.0
            whole.aabb_max[a] = float(size[a] * 16)
        assert _map(big, ms * 64, [whole] * 64, True) == LUMX_EINVAL
        assert b"scratch" in big.lib.lumx_last_error()
    finally:
        for m in ms:
            m.close()
        big.close()
