"""Device probe selection, counted radiance updates and the field shift.

Bar: the device selection equals the host selection and the oracle's
(update_radiance's request build) entry for entry; a counted update equals
the plain update over the same list bit for bit; the shift equals the
oracle's restatement of shift_radiance bit for bit.
"""
import numpy as np
import pytest

from lum_amd import synth
from lum_amd.api import World, Model, placement, radiance_params
from lum_amd._lib import RADIANCE_F16, RADIANCE_BF16, RADIANCE_F32
import tests._oracle as orc

from tests.test_gpu_parity import make_case, gpu_world, TUNINGS

pytestmark = pytest.mark.gpu


def in_world(special, size):
    """The specials the device considers: inside the world, repeats kept
    (internal_render.cpp:1164-1168 appends every special whose probe the
    grid did not select)."""
    keep = [s for s in special
            if 0 <= int(s[0]) < size[0] and 0 <= int(s[1]) < size[1] and 0 <= int(s[2]) < size[2]]
    return np.array(keep, np.int32).reshape(-1, 4)


def device_select(w, special):
    import torch
    sp = None if special is None or len(special) == 0 else torch.from_numpy(np.ascontiguousarray(special)).cuda()
    req, cnt = w.select_probes_device(sp)
    torch.cuda.synchronize()
    n = int(cnt.cpu()[0])
    return req[:n].cpu().numpy(), req, cnt


SELECT_CASES = [
    ((20, 20, 10), 12, "half"),     # 4000 probes: two tiles, the second partial
    ((48, 48, 16), 3, "terrain"),   # 36864 probes: 18 full tiles
    ((7, 5, 3), 4, "tiny"),          # one partial tile
    ((64, 32, 2), 5, "empty"),      # no solid block: only specials
    ((16, 16, 8), 6, "full"),       # every block solid
    ((130, 9, 5), 7, "sparse"),
]


@pytest.mark.parametrize("size,seed,kind", SELECT_CASES, ids=[c[2] for c in SELECT_CASES])
def test_device_selection_matches_host_and_oracle(gpu_ctx, size, seed, kind):
    blocks, pal, mats, field = make_case(size, seed)
    if kind == "half":
        blocks[:5] = 0
    elif kind == "empty":
        blocks[:] = 0
    elif kind == "full":
        blocks[blocks == 0] = 1
    elif kind == "sparse":
        rng = np.random.default_rng(seed)
        blocks[:] = 0
        blocks.reshape(-1)[rng.choice(blocks.size, 12, replace=False)] = 3
    rng = np.random.default_rng(seed + 100)
    n_sp = 300
    special = np.stack([rng.integers(-1, size[0] + 1, n_sp), rng.integers(0, size[1], n_sp),
                        rng.integers(0, size[2], n_sp), rng.integers(0, 5, n_sp)], 1).astype(np.int32)
    special[10:20] = special[0:10]      # repeats
    w = gpu_world(gpu_ctx, blocks, pal, mats, field)
    try:
        got, _, _ = device_select(w, special)
        kept = in_world(special, size)
        ref = orc.select_probes(blocks, kept)
        assert np.array_equal(got, ref)
        assert len(got) <= blocks.size  # one entry per probe at most
        host = w.select_probes(kept)
        assert np.array_equal(got, host)
        got0, _, _ = device_select(w, None)
        assert np.array_equal(got0, orc.select_probes(blocks, None))
        # Nothing carries over between calls: the same call again gives the same list.
        again, _, _ = device_select(w, special)
        assert np.array_equal(again, got)
    finally:
        w.close()


def test_device_selection_follows_map(gpu_ctx):
    """After a map the selection sees the blockified world (stream-ordered)."""
    size = (32, 32, 8)
    blocks, pal, mats, field = make_case(size, 21)
    blocks[2:] = 0
    w = gpu_world(gpu_ctx, blocks, pal, mats, field)
    rng = np.random.default_rng(22)
    vox, quats, shifts = synth.make_models(rng, n=3, lo=16, hi=32, fill=0.3, world_size=size)
    ms = [Model(gpu_ctx, v) for v in vox]
    try:
        w.map_models_placed(ms, [placement(q, s, m.size) for q, s, m in zip(quats, shifts, ms)])
        got, _, _ = device_select(w, None)
        cur = w.get_blocks(True)
        assert not np.array_equal(cur, blocks)
        assert np.array_equal(got, orc.select_probes(cur, None))
    finally:
        for m in ms:
            m.close()
        w.close()


@pytest.mark.parametrize("tuning", ["lane", "lane_occ", "tiled", "lane_compacted"])
@pytest.mark.parametrize("size_shift", [(1, 0), (3, 2)])
def test_counted_update_equals_listed_update(gpu_ctx, tuning, size_shift):
    """lane_compacted: the counted update runs compacted (claimed probes in
    texel order, the count read on the device), the listed one never does."""
    import torch
    size = (24, 24, 12)
    blocks, pal, mats, field = make_case(size, 31)
    blocks[:4] = 0
    special = np.array([[1, 1, 11, 0], [2, 3, 10, 0]], np.int32)
    a = gpu_world(gpu_ctx, blocks, pal, mats, field)
    b = gpu_world(gpu_ctx, blocks, pal, mats, field)
    try:
        if tuning == "lane_compacted":
            a.set_tuning(**dict(TUNINGS["lane"], compact=1))
            b.set_tuning(**dict(TUNINGS["lane"], compact=-1))
        else:
            a.set_tuning(**TUNINGS[tuning])
            b.set_tuning(**TUNINGS[tuning])
        p = radiance_params(rays=64, time_seed=5, alpha=0.2, max_steps=64, size=size_shift[0],
                            shift=size_shift[1])
        listed, req, cnt = device_select(a, special)
        a.radiance_update_counted(req, cnt, p)
        b.radiance_update(torch.from_numpy(listed).cuda(), p)
        torch.cuda.synchronize()
        assert np.array_equal(a.read_radiance(), b.read_radiance())
        assert not np.array_equal(a.read_radiance(), field)
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("compact", [0, 1])
def test_counted_update_clamps_to_capacity(gpu_ctx, compact):
    """A device count above the list's capacity dispatches the capacity
    (compact=1: the counted update compacted, the listed one not)."""
    import torch
    size = (16, 16, 8)
    blocks, pal, mats, field = make_case(size, 41)
    a = gpu_world(gpu_ctx, blocks, pal, mats, field)
    b = gpu_world(gpu_ctx, blocks, pal, mats, field)
    try:
        if compact:
            a.set_tuning(**dict(TUNINGS["lane"], compact=1))
            b.set_tuning(**dict(TUNINGS["lane"], compact=-1))
        req = torch.from_numpy(orc.select_probes(blocks, None)[:100].copy()).cuda()
        cnt = torch.tensor([1 << 30], dtype=torch.int32, device="cuda")
        p = radiance_params(rays=32, time_seed=1, alpha=0.3, max_steps=64)
        a.radiance_update_counted(req, cnt, p)
        b.radiance_update(req, p)
        torch.cuda.synchronize()
        first = a.read_radiance()
        assert np.array_equal(first, b.read_radiance())
        assert not np.array_equal(first, field)
        cnt.fill_(0)
        a.radiance_update_counted(req, cnt, p)   # nothing dispatched: field copied through
        torch.cuda.synchronize()
        assert np.array_equal(a.read_radiance(), first)
    finally:
        a.close()
        b.close()


SHIFTS = [(1, 0, 0), (-1, 0, 0), (0, 2, -1), (-3, -2, 1), (5, 5, 5), (0, 0, 0), (24, 0, 0), (0, -10, 0),
          (0, 0, 9)]


@pytest.mark.parametrize("fmt", [RADIANCE_F32, RADIANCE_F16, RADIANCE_BF16])
def test_shift_matches_oracle(gpu_ctx, fmt):
    import torch
    size = (24, 10, 9)
    blocks, pal, mats, field = make_case(size, 51)
    w = gpu_world(gpu_ctx, blocks, pal, mats, field, fmt=fmt)
    try:
        cur = w.read_radiance()
        for s in SHIFTS:
            w.shift_radiance(s)
            torch.cuda.synchronize()
            got = w.read_radiance()
            ref = orc.shift_radiance(cur, s)
            assert np.array_equal(got, ref), s
            cur = got
    finally:
        w.close()


def test_frame_shift_select_counted_update_chain(gpu_ctx):
    """shift -> device selection -> counted update equals the host-driven
    chain (shift, host selection, listed update) over several frames."""
    import torch
    size = (32, 24, 8)
    blocks, pal, mats, field = make_case(size, 61)
    blocks[3:] = 0
    a = gpu_world(gpu_ctx, blocks, pal, mats, field, fmt=RADIANCE_F16)
    b = gpu_world(gpu_ctx, blocks, pal, mats, field, fmt=RADIANCE_F16)
    try:
        special = np.array([[5, 5, 7, 0]], np.int32)
        for frame in range(4):
            s = [(1, 0, 0), (0, -1, 0), (0, 0, 0), (-2, 1, 0)][frame]
            p = radiance_params(rays=64, time_seed=frame, alpha=0.1, max_steps=64, size=2, shift=frame % 2)
            a.shift_radiance(s)
            _, req, cnt = device_select(a, special)
            a.radiance_update_counted(req, cnt, p)
            b.shift_radiance(s)
            host = b.select_probes(special)
            b.radiance_update(torch.from_numpy(host).cuda(), p)
        torch.cuda.synchronize()
        assert np.array_equal(a.read_radiance(), b.read_radiance())
    finally:
        a.close()
        b.close()
