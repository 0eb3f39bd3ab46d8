"""Eviction to the pinned host snapshot and restore (lumx_world_evict /
lumx_world_restore).

Bar: a world that goes out and comes back continues bit for bit as its twin
that stayed resident, whichever field buffer held its current field when it
left (full updates write the other buffer and flip; subset updates land in
place) and whichever restore layout applies: one copy when every palette slot
is in use (the field then comes back as field 0), two copies otherwise.
Subset updates after a restore also check that the claim words came back idle.
"""
import numpy as np
import pytest

from lum_amd.api import radiance_params
from lum_amd._lib import RADIANCE_F16, RADIANCE_F32

from tests.test_gpu_parity import make_case, gpu_world

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("slots", [64, 0], ids=["one_copy", "two_copies"])
@pytest.mark.parametrize("fmt", [RADIANCE_F32, RADIANCE_F16], ids=["f32", "f16"])
def test_evict_restore_continues_as_the_twin(gpu_ctx, slots, fmt):
    import torch
    size = (20, 12, 6)
    blocks, pal, mats, field = make_case(size, 61)   # 64 static palette blocks
    a = gpu_world(gpu_ctx, blocks, pal, mats, field, fmt=fmt, slots=slots)
    b = gpu_world(gpu_ctx, blocks, pal, mats, field, fmt=fmt, slots=slots)
    n = size[0] * size[1] * size[2]
    flat = np.random.default_rng(62).permutation(n)[: n // 3]
    req = np.stack([flat % size[0], (flat // size[0]) % size[1], flat // (size[0] * size[1]),
                    np.zeros(len(flat), int)], 1).astype(np.int32)
    req_d = torch.from_numpy(req).cuda()
    try:
        # full, subset, full, full, subset: the current field sits in either
        # buffer at the evictions
        for k, r in enumerate([None, req_d, None, None, req_d]):
            p = radiance_params(rays=32, time_seed=k, alpha=0.25, max_steps=64)
            for w in (a, b):
                w.radiance_update(r, p)
            torch.cuda.synchronize()
            a.evict()
            assert not a.resident()[0]
            a.restore()
            assert a.resident()[0]
            assert np.array_equal(a.read_radiance(), b.read_radiance()), k
        assert not np.array_equal(a.read_radiance(), field)
    finally:
        a.close()
        b.close()
