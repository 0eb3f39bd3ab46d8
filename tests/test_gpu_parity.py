"""GPU parity: liblumx HIP kernels vs the CPU oracle on identical seeded inputs.

Bar (per pass):
  radiance  hit keys agree on >= 99.9% of rays; radiance within 1e-3
            relative (|a-b| <= 1e-3 * max(|b|, 1e-3)) on every texel, FP32;
  map       bit-exact palette and blockified world;
  glossy    hit keys agree on >= 99.9% of pixels; output within 1e-3 relative.
"""
import numpy as np
import pytest

from lum_amd import synth
from lum_amd.api import World, Model, placement, camera_default, radiance_params
from lum_amd._lib import RADIANCE_F16, RADIANCE_BF16, RADIANCE_F32
import tests._oracle as orc

pytestmark = pytest.mark.gpu

REL = 1e-3


def rel_err(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-3)


def make_case(size, seed, n_palette=64):
    rng = np.random.default_rng(seed)
    blocks = synth.make_world(size, rng, n_palette=n_palette)
    pal = synth.make_palette(rng, n_palette)
    mats = synth.make_materials(rng)
    field = synth.make_field(size, rng)
    return blocks, pal, mats, field


def gpu_world(ctx, blocks, pal, mats, field, fmt=RADIANCE_F32, slots=0):
    wz, wy, wx = blocks.shape
    w = World(ctx, (wx, wy, wz), static_blocks=pal.shape[0], palette_slots=slots, radiance_format=fmt)
    w.set_blocks(blocks)
    w.set_palette(pal)
    w.set_materials(mats)
    w.write_radiance(field)
    w.upload_blocks()
    return w


def full_palette(world, pal):
    p = np.zeros((world.palette_slots, 16, 16, 16), np.uint8)
    p[: pal.shape[0]] = pal
    return p


TUNINGS = {
    "lane": dict(variant=0),
    "lane_occ": dict(variant=2),
    "lane_rpl4": dict(variant=0, rays_per_lane=4),
    "tiled": dict(variant=1, threads=256, probes_per_batch=16, lds_block_bits=1, lds_palette_bits=1),
    "tiled_global_bits": dict(variant=1, threads=1024, probes_per_batch=4, lds_block_bits=0, lds_palette_bits=0),
    "tiled_p32": dict(variant=1, threads=512, probes_per_batch=32, lds_block_bits=1, lds_palette_bits=0),
}


@pytest.mark.parametrize("tuning", sorted(TUNINGS))
@pytest.mark.parametrize("size,rays,seed,max_steps", [
    ((16, 16, 8), 64, 1, 0),
    ((48, 48, 16), 64, 1234, 64),
    ((24, 20, 12), 32, 7, 0),
    ((20, 24, 10), 128, 8, 96),
])
def test_radiance_all_probes(gpu_ctx, size, rays, seed, max_steps, tuning):
    import torch
    blocks, pal, mats, field = make_case(size, seed)
    w = gpu_world(gpu_ctx, blocks, pal, mats, field)
    w.set_tuning(**TUNINGS[tuning])
    try:
        keys, steps = w.radiance_update(None, radiance_params(rays=rays, time_seed=3, alpha=0.1,
                                                              max_steps=max_steps), traced=True)
        torch.cuda.synchronize()
        got = w.read_radiance()
        wd = orc.WorldData(blocks, full_palette(w, pal), mats)
        ref, rkeys, rsteps = orc.radiance_update(wd, field, rays=rays, time_seed=3, alpha=0.1,
                                                 max_steps=max_steps, traced=True)
        gk = keys.cpu().numpy()
        agree = (gk == rkeys).mean()
        print(f"size={size} rays={rays}: key agreement {agree:.6f}, "
              f"max rel {rel_err(got, ref).max():.3e}")
        assert agree >= 0.999
        assert rel_err(got, ref).max() <= REL
        assert np.array_equal(steps.cpu().numpy(), rsteps) or agree < 1.0
    finally:
        w.close()


@pytest.mark.parametrize("rays", [16, 32, 64, 128, 256])
def test_radiance_rays_per_lane_bit_exact(gpu_ctx, rays):
    """A thread tracing 2 or 4 rays adds them in the butterfly's order: the
    field equals the one-ray-per-thread field bit for bit (full updates and a
    request-list subset, 1,680 probes so the last workgroup is partial), and
    that equals the oracle's bit for bit (butterfly stages of every width,
    chunk sums beyond 64 rays)."""
    import torch
    size = (20, 12, 7)
    blocks, pal, mats, field = make_case(size, 31)
    wx, wy, wz = size
    rng = np.random.default_rng(rays)
    flat = rng.choice(wx * wy * wz, size=333, replace=False)
    req = np.stack([flat % wx, (flat // wx) % wy, flat // (wx * wy), np.zeros(333, int)], 1).astype(np.int32)
    req_d = torch.from_numpy(req).cuda()
    outs = {}
    for k in (1, 2, 4):
        w = gpu_world(gpu_ctx, blocks, pal, mats, field)
        w.set_tuning(variant=0, rays_per_lane=k)
        try:
            w.radiance_update(None, radiance_params(rays=rays, time_seed=9, alpha=0.2, max_steps=64))
            w.radiance_update(req_d, radiance_params(rays=rays, time_seed=10, alpha=0.2, max_steps=64))
            torch.cuda.synchronize()
            outs[k] = w.read_radiance()
            full = full_palette(w, pal)
        finally:
            w.close()
    assert np.array_equal(outs[2], outs[1])
    assert np.array_equal(outs[4], outs[1])
    wd = orc.WorldData(blocks, full, mats)
    ref = orc.radiance_update(wd, field, rays=rays, time_seed=9, alpha=0.2, max_steps=64)
    ref = orc.radiance_update(wd, ref, rays=rays, time_seed=10, alpha=0.2, max_steps=64, requests=req)
    # FP32 follows the oracle's arithmetic and reduction order exactly
    assert np.array_equal(outs[1], ref)


@pytest.mark.parametrize("tuning", sorted(TUNINGS))
def test_radiance_nan_stores_zero(gpu_ctx, tuning):
    """NaN reaches a probe mean through a NaN field (or a 0 * inf ray); every
    launch shape stores it as 0, like the reference's UNORM store, and
    clamps an old NaN to 0: bit-exact with the oracle."""
    import torch
    size = (16, 16, 8)
    blocks, pal, mats, field = make_case(size, 77)
    field[:2, ..., :3] = np.nan                        # NaN in the two bottom layers only
    field[0, ::2, 1::2, :3] = 0.25
    w = gpu_world(gpu_ctx, blocks, pal, mats, field)
    w.set_tuning(**TUNINGS[tuning])
    try:
        w.radiance_update(None, radiance_params(rays=64, time_seed=5, alpha=0.3, max_steps=64))
        torch.cuda.synchronize()
        got = w.read_radiance()
        wd = orc.WorldData(blocks, full_palette(w, pal), mats)
        ref = orc.radiance_update(wd, field, rays=64, time_seed=5, alpha=0.3, max_steps=64)
        assert not np.isnan(ref).any() and not np.isnan(got).any()
        assert np.array_equal(got, ref)
        assert (got[..., :3] == 0).any() and (got[..., :3] > 0).any()
    finally:
        w.close()


@pytest.mark.parametrize("tuning", ["lane", "tiled"])
def test_radiance_request_list_jacobi(gpu_ctx, tuning):
    """Subset of probes with size/shift: unrequested probes keep their value."""
    import torch
    size = (20, 18, 10)
    blocks, pal, mats, field = make_case(size, 42)
    rng = np.random.default_rng(5)
    wx, wy, wz = size
    n = 700
    flat = rng.choice(wx * wy * wz, size=n, replace=False)
    req = np.stack([flat % wx, (flat // wx) % wy, flat // (wx * wy), np.zeros(n, int)], 1).astype(np.int32)
    w = gpu_world(gpu_ctx, blocks, pal, mats, field)
    w.set_tuning(**TUNINGS[tuning])
    try:
        req_d = torch.from_numpy(req).cuda()
        w.radiance_update(req_d, radiance_params(rays=64, time_seed=11, size=3, shift=2, alpha=0.25))
        torch.cuda.synchronize()
        got = w.read_radiance()
        wd = orc.WorldData(blocks, full_palette(w, pal), mats)
        ref = orc.radiance_update(wd, field, requests=req, rays=64, time_seed=11, size=3, shift=2, alpha=0.25)
        assert rel_err(got, ref).max() <= REL
        touched = np.zeros(wx * wy * wz, bool)
        sel = req[2::3][: n // 3]
        touched[sel[:, 0] + wx * (sel[:, 1] + wy * sel[:, 2])] = True
        same = got.reshape(-1, 4)[~touched]
        assert np.array_equal(same, field.reshape(-1, 4)[~touched])
    finally:
        w.close()


@pytest.mark.parametrize("tuning", ["lane", "lane_occ", "lane_rpl4", "tiled"])
@pytest.mark.parametrize("size_shift", [(1, 0), (2, 1)])
def test_radiance_duplicate_requests_last_wins(gpu_ctx, tuning, size_shift):
    """A list naming some probes several times (and some outside the world):
    every entry is traced with its own RNG key, and of the entries naming one
    probe the last one's texel is stored, as in the oracle."""
    import torch
    size = (12, 10, 6)
    blocks, pal, mats, field = make_case(size, 77)
    rng = np.random.default_rng(78)
    wx, wy, wz = size
    n = 300
    flat = rng.integers(0, 40, size=n)  # 40 distinct probes, each named ~7 times
    req = np.stack([flat % wx, (flat // wx) % wy, flat // (wx * wy), np.zeros(n, int)], 1).astype(np.int32)
    req[5] = [wx, 0, 0, 0]       # outside: skipped
    req[17] = [-1, 2, 1, 0]
    w = gpu_world(gpu_ctx, blocks, pal, mats, field)
    w.set_tuning(**TUNINGS[tuning])
    sz, sh = size_shift
    try:
        for frame in range(2):  # the claims reset between updates
            w.radiance_update(torch.from_numpy(req).cuda(), radiance_params(rays=64, time_seed=3 + frame, size=sz,
                                                                            shift=sh, alpha=0.3))
        torch.cuda.synchronize()
        got = w.read_radiance()
        wd = orc.WorldData(blocks, full_palette(w, pal), mats)
        ref = field
        for frame in range(2):
            ref = orc.radiance_update(wd, ref, requests=req, rays=64, time_seed=3 + frame, size=sz, shift=sh,
                                      alpha=0.3)
        assert np.array_equal(got, ref)
        # the last entry of a repeated probe is the one stored: it differs
        # from what the first entry alone would give
        first = orc.radiance_update(wd, field, requests=req[:1], rays=64, time_seed=3, alpha=0.3)
        assert not np.array_equal(first.reshape(-1, 4)[flat[0]],
                                  orc.radiance_update(wd, field, requests=req, rays=64, time_seed=3,
                                                      alpha=0.3).reshape(-1, 4)[flat[0]])
    finally:
        w.close()


@pytest.mark.parametrize("tuning", ["lane", "lane_occ", "lane_rpl4"])
@pytest.mark.parametrize("rays,fmt", [(64, "f32"), (128, "f16"), (32, "f32"), (16, "f16")])
def test_radiance_compacted_update_bit_exact(gpu_ctx, tuning, rays, fmt):
    """A compacted subset update (claimed probes run in texel order, each
    entry's request index keying its rays) gives the oracle's field bit for
    bit, with repeats (the last wins), entries outside the world and
    size/shift, over two frames; and equals the same list through the tiled
    kernel, which never compacts."""
    import torch
    from lum_amd._lib import RADIANCE_F16, RADIANCE_F32
    size = (16, 16, 16)
    blocks, pal, mats, field = make_case(size, 301)
    wx, wy, wz = size
    rng = np.random.default_rng(302)
    n = 4096                                     # at most one dispatched probe per texel
    flat = rng.integers(0, 1400, size=n)         # 1400 distinct probes, each named ~3 times
    req = np.stack([flat % wx, (flat // wx) % wy, flat // (wx * wy), np.zeros(n, int)], 1).astype(np.int32)
    req[7] = [wx, 0, 0, 0]
    req[1001] = [0, -1, 3, 0]
    f = RADIANCE_F16 if fmt == "f16" else RADIANCE_F32
    results = []
    for tun in (dict(TUNINGS[tuning], compact=1), TUNINGS["tiled_global_bits"]):
        w = gpu_world(gpu_ctx, blocks, pal, mats, field, fmt=f)
        w.set_tuning(**tun)
        full = full_palette(w, pal)
        try:
            req_d = torch.from_numpy(req).cuda()
            w.radiance_update(req_d, radiance_params(rays=rays, time_seed=21, alpha=0.3))
            w.radiance_update(req_d, radiance_params(rays=rays, time_seed=22, size=2, shift=1, alpha=0.3))
            torch.cuda.synchronize()
            results.append(w.read_radiance())
        finally:
            w.close()
    assert np.array_equal(results[0], results[1])
    if fmt == "f32":
        wd = orc.WorldData(blocks, full, mats)
        ref = orc.radiance_update(wd, field, requests=req, rays=rays, time_seed=21, alpha=0.3)
        ref = orc.radiance_update(wd, ref, requests=req, rays=rays, time_seed=22, size=2, shift=1, alpha=0.3)
        assert np.array_equal(results[0], ref)


@pytest.mark.parametrize("tuning", ["lane", "lane_rpl4"])
def test_radiance_compacted_strided_update_bit_exact(gpu_ctx, tuning):
    """A compacted update without a request list (every fourth probe, from
    the third, in x-fastest order) gives the oracle's field bit for bit."""
    size = (16, 16, 32)
    blocks, pal, mats, field = make_case(size, 311)
    w = gpu_world(gpu_ctx, blocks, pal, mats, field)
    w.set_tuning(**dict(TUNINGS[tuning], compact=1))
    full = full_palette(w, pal)
    try:
        w.radiance_update(None, radiance_params(rays=64, time_seed=31, size=4, shift=2, alpha=0.3))
        got = w.read_radiance()
    finally:
        w.close()
    wd = orc.WorldData(blocks, full, mats)
    ref = orc.radiance_update(wd, field, rays=64, time_seed=31, size=4, shift=2, alpha=0.3)
    assert np.array_equal(got, ref)


def test_request_list_longer_than_world_rejected(gpu_ctx):
    """A list dispatching more probes than the world holds is refused before
    any launch: the staged texels live in one texel per probe."""
    import torch
    from lum_amd._lib import LumxError, LUMX_EINVAL
    size = (4, 4, 2)
    blocks, pal, mats, field = make_case(size, 5)
    w = gpu_world(gpu_ctx, blocks, pal, mats, field)
    try:
        n_blocks = 32
        req = torch.zeros((2 * (n_blocks + 1), 4), dtype=torch.int32, device="cuda")
        with pytest.raises(LumxError) as e:
            w.radiance_update(req[: n_blocks + 1], radiance_params(rays=32))
        assert e.value.code == LUMX_EINVAL
        with pytest.raises(LumxError):
            w.radiance_update(req, radiance_params(rays=32, size=2, shift=1))
        cnt = torch.tensor([1], dtype=torch.int32, device="cuda")
        with pytest.raises(LumxError):
            w.radiance_update_counted(req[: n_blocks + 1], cnt, radiance_params(rays=32))
        # one dispatched probe per block at stride 2 is accepted (32 entries
        # naming probe 0: the last one's texel is stored)
        w.radiance_update(req[: 2 * n_blocks + 1], radiance_params(rays=32, size=2, alpha=0.5))
        torch.cuda.synchronize()
        got = w.read_radiance()
        wd = orc.WorldData(blocks, full_palette(w, pal), mats)
        ref = orc.radiance_update(wd, field, requests=req[: 2 * n_blocks + 1].cpu().numpy(), rays=32, size=2,
                                  alpha=0.5)
        assert np.array_equal(got, ref)
    finally:
        w.close()


def test_radiance_frames_chain(gpu_ctx):
    """Three chained frames (time seed = frame) stay within tolerance."""
    import torch
    size = (16, 16, 8)
    blocks, pal, mats, field = make_case(size, 99)
    w = gpu_world(gpu_ctx, blocks, pal, mats, field)
    try:
        wd = orc.WorldData(blocks, full_palette(w, pal), mats)
        ref = field
        for frame in range(3):
            w.radiance_update(None, radiance_params(rays=64, time_seed=frame, alpha=0.1))
            ref = orc.radiance_update(wd, ref, rays=64, time_seed=frame, alpha=0.1)
        torch.cuda.synchronize()
        got = w.read_radiance()
        assert rel_err(got, ref).max() <= REL
    finally:
        w.close()


def _quantize(x, fmt):
    if fmt == RADIANCE_F16:
        return x.astype(np.float16).astype(np.float32)
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


@pytest.mark.parametrize("fmt", [RADIANCE_F16, RADIANCE_BF16])
def test_radiance_low_precision_storage(gpu_ctx, fmt):
    """FP16/BF16 storage, FP32 march: equals the FP32 oracle fed the stored
    (quantized) field, with its output rounded to the storage format."""
    import torch
    size = (16, 16, 8)
    blocks, pal, mats, field = make_case(size, 3)
    w = gpu_world(gpu_ctx, blocks, pal, mats, field, fmt=fmt)
    try:
        w.radiance_update(None, radiance_params(rays=64, time_seed=1, alpha=0.1))
        torch.cuda.synchronize()
        got = w.read_radiance()
        wd = orc.WorldData(blocks, full_palette(w, pal), mats)
        ref = _quantize(orc.radiance_update(wd, _quantize(field, fmt), rays=64, time_seed=1, alpha=0.1), fmt)
        ulp = 2.0 ** -10 if fmt == RADIANCE_F16 else 2.0 ** -7
        assert (np.abs(got - ref) <= ulp * np.maximum(np.abs(ref), 1e-3)).all()
        assert (got == ref).mean() >= 0.999
    finally:
        w.close()


def test_map_models_bit_exact(gpu_ctx):
    """blockify + exec_copies + map.comp for 8 seeded rotated models."""
    import torch
    size = (24, 24, 12)
    blocks, pal, mats, field = make_case(size, 17)
    rng = np.random.default_rng(21)
    models, quats, shifts = synth.make_models(rng, n=8, lo=8, hi=40, world_size=size)
    w = gpu_world(gpu_ctx, blocks, pal, mats, field)
    ms = [Model(gpu_ctx, m) for m in models]
    try:
        pls = [placement(q, s, m.size) for q, s, m in zip(quats, shifts, ms)]
        w.map_models_placed(ms, pls)
        torch.cuda.synchronize()
        counter = w.palette_counter()
        got_pal = w.get_palette(0, counter)
        got_world = w.get_blocks(blockified=True)
        # oracle: blockify -> exec_copies -> map, in order
        cur = blocks.copy()
        copy_src = np.zeros(w.palette_slots, np.int32)
        c = pal.shape[0]
        for p in pls:
            c = orc.blockify(size, cur, pal.shape[0], c, list(p.aabb_min), list(p.aabb_max), copy_src)
            assert c > 0
        assert c == counter
        ref_pal = full_palette(w, pal)
        orc.exec_copies(ref_pal, pal.shape[0], c, copy_src)
        for m, p in zip(models, pls):
            orc.map_model(size, cur, ref_pal, pal.shape[0], m, list(p.minv), list(p.shift)[:3], list(p.groups))
        assert np.array_equal(got_world, cur)
        assert np.array_equal(got_pal, ref_pal[:counter])
    finally:
        for m in ms:
            m.close()
        w.close()


@pytest.mark.parametrize("device_blockify", [False, True])
def test_map_models_crossing_origin_bit_exact(gpu_ctx, device_blockify):
    """Models whose dispatch box reaches below 0 on one, two or three axes:
    the threads at negative coordinates alias into block 0 as the shader's
    truncating division and OpSMod make them (map.comp:40-66), the highest
    dispatch index kept where threads collide; equal to the oracle bit for
    bit, together with models placed inside the world."""
    import torch
    size = (12, 10, 6)
    blocks, pal, mats, field = make_case(size, 23)
    rng = np.random.default_rng(24)
    models, quats, shifts = synth.make_models(rng, n=7, lo=6, hi=30, world_size=size)
    # Pull models 0-4 across the origin on various axes (half-voxel offsets
    # too), keep 5-6 inside.
    pulls = [(-7.0, 3.0, 2.0), (4.0, -5.5, 1.0), (-3.25, -2.0, 0.5), (-9.0, -9.0, -4.0), (2.0, 2.0, -12.5)]
    for i, d in enumerate(pulls):
        shifts[i] = tuple(float(v) for v in d)
    w = gpu_world(gpu_ctx, blocks, pal, mats, field)
    ms = [Model(gpu_ctx, m) for m in models]
    try:
        pls = [placement(q, s, m.size) for q, s, m in zip(quats, shifts, ms)]
        assert sum(min(p.shift[0], p.shift[1], p.shift[2]) < 0 for p in pls) == 5
        if device_blockify:
            w.map_models_device(ms, pls)
        else:
            w.map_models_placed(ms, pls)
        torch.cuda.synchronize()
        counter = w.palette_counter()
        got_pal = w.get_palette(0, counter)
        cur = blocks.copy()
        copy_src = np.zeros(w.palette_slots, np.int32)
        c = pal.shape[0]
        for p in pls:
            c = orc.blockify(size, cur, pal.shape[0], c, list(p.aabb_min), list(p.aabb_max), copy_src)
        assert c == counter
        ref_pal = full_palette(w, pal)
        orc.exec_copies(ref_pal, pal.shape[0], c, copy_src)
        before = ref_pal.copy()
        for m, p in zip(models, pls):
            orc.map_model(size, cur, ref_pal, pal.shape[0], m, list(p.minv), list(p.shift)[:3], list(p.groups))
        assert np.array_equal(w.get_blocks(blockified=True), cur)
        assert np.array_equal(got_pal, ref_pal[:counter])
        # block 0's slot was written through aliased coordinates
        b0 = cur[0, 0, 0]
        assert b0 >= pal.shape[0] and not np.array_equal(ref_pal[b0], before[b0])
    finally:
        for m in ms:
            m.close()
        w.close()


def test_glossy_parity(gpu_ctx):
    import torch
    size = (48, 48, 16)
    blocks, pal, mats, field = make_case(size, 55)
    rng = np.random.default_rng(56)
    W, H = 320, 180
    matnorm, depth = synth.make_gbuffer(W, H, mats, rng)
    w = gpu_world(gpu_ctx, blocks, pal, mats, field)
    try:
        cam = camera_default(W, H)
        out, keys = w.glossy(torch.from_numpy(matnorm).cuda(), torch.from_numpy(depth).cuda(), cam,
                             max_dist=128.0, max_steps=128, traced=True)
        torch.cuda.synchronize()
        ubo = np.array(list(cam.campos) + list(cam.camdir) + list(cam.horiz_scaled) + list(cam.verti_scaled)
                       + list(cam.light_dir) + list(cam.frame_size), np.float32)
        wd = orc.WorldData(blocks, full_palette(w, pal), mats)
        ref, rkeys = orc.glossy(wd, field, matnorm, depth, ubo, 128.0, 128)
        agree = (keys.cpu().numpy() == rkeys).mean()
        got = out.cpu().numpy()
        print(f"glossy key agreement {agree:.6f} max rel {rel_err(got, ref).max():.3e}")
        assert agree >= 0.999
        assert rel_err(got, ref).max() <= REL
    finally:
        w.close()


def test_select_probes_matches_oracle(gpu_ctx):
    size = (20, 20, 10)
    blocks, pal, mats, field = make_case(size, 12)
    blocks[:, :, :5] = 0
    w = gpu_world(gpu_ctx, blocks, pal, mats, field)
    try:
        special = np.array([[1, 1, 9, 0], [0, 0, 0, 0], [1, 1, 9, 0]], np.int32)
        got = w.select_probes(special)
        ref = orc.select_probes(blocks, special)
        assert np.array_equal(got, ref)
        # The repeated special is appended twice (internal_render.cpp:1164-1168).
        assert (got[:, :3] == [1, 1, 9]).all(1).sum() == 2
    finally:
        w.close()


def test_autotune_keeps_state_and_parity(gpu_ctx):
    """The autotuner leaves the field unchanged and its pick stays in parity."""
    import torch
    size = (24, 24, 12)
    blocks, pal, mats, field = make_case(size, 31)
    w = gpu_world(gpu_ctx, blocks, pal, mats, field)
    try:
        p = radiance_params(rays=64, time_seed=2, alpha=0.1, max_steps=64)
        best, ms = w.autotune(p, reps=2)
        assert ms > 0
        assert np.array_equal(w.read_radiance(), field)
        assert w.get_tuning() == best
        w.radiance_update(None, p)
        torch.cuda.synchronize()
        wd = orc.WorldData(blocks, full_palette(w, pal), mats)
        ref = orc.radiance_update(wd, field, rays=64, time_seed=2, alpha=0.1, max_steps=64)
        assert rel_err(w.read_radiance(), ref).max() <= REL
    finally:
        w.close()


def test_radiance_after_map_uses_dynamic_slots(gpu_ctx):
    """Radiance over a blockified world (models mapped) matches the oracle on
    the same palette, for the tiled kernel with palette bits in LDS and in
    global memory."""
    import torch
    size = (24, 24, 12)
    blocks, pal, mats, field = make_case(size, 41)
    rng = np.random.default_rng(42)
    models, quats, shifts = synth.make_models(rng, n=6, lo=8, hi=32, world_size=size)
    w = gpu_world(gpu_ctx, blocks, pal, mats, field)
    ms = [Model(gpu_ctx, m) for m in models]
    try:
        w.map_models(ms, quats, shifts)
        torch.cuda.synchronize()
        counter = w.palette_counter()
        full = np.zeros((w.palette_slots, 16, 16, 16), np.uint8)
        full[:counter] = w.get_palette(0, counter)
        wd = orc.WorldData(w.get_blocks(True), full, mats)
        ref = orc.radiance_update(wd, field, rays=64, time_seed=5, alpha=0.1)
        for pb in (1, 0):
            w.write_radiance(field)
            w.set_tuning(**dict(TUNINGS["tiled"], lds_palette_bits=pb))
            w.radiance_update(None, radiance_params(rays=64, time_seed=5, alpha=0.1))
            torch.cuda.synchronize()
            assert rel_err(w.read_radiance(), ref).max() <= REL
    finally:
        for m in ms:
            m.close()
        w.close()


@pytest.mark.parametrize("pool", [1, 2, 4])
@pytest.mark.parametrize("size,max_steps", [((16, 16, 8), 0), ((48, 48, 16), 64), ((13, 11, 7), 0)])
def test_radiance_refill_variant_bit_exact(gpu_ctx, pool, size, max_steps):
    """The lane-refill kernel (variant 3: a wave marches `pool` probes' rays as
    one pool, finished lanes taking the next ray) gives the oracle's field bit
    for bit: a full update (13x11x7 leaves a partial last pool), then a
    request list with size/shift, repeats and entries outside the world."""
    import torch
    blocks, pal, mats, field = make_case(size, 91)
    wx, wy, wz = size
    rng = np.random.default_rng(92)
    flat = rng.integers(0, wx * wy * wz, size=257)
    req = np.stack([flat % wx, (flat // wx) % wy, flat // (wx * wy), np.zeros(257, int)], 1).astype(np.int32)
    req[3] = [wx, 0, 0, 0]
    w = gpu_world(gpu_ctx, blocks, pal, mats, field)
    w.set_tuning(variant=3, probes_per_batch=pool)
    try:
        w.radiance_update(None, radiance_params(rays=64, time_seed=5, alpha=0.2, max_steps=max_steps))
        w.radiance_update(torch.from_numpy(req).cuda(),
                          radiance_params(rays=64, time_seed=6, size=2, shift=1, alpha=0.2, max_steps=max_steps))
        torch.cuda.synchronize()
        got = w.read_radiance()
        wd = orc.WorldData(blocks, full_palette(w, pal), mats)
    finally:
        w.close()
    ref = orc.radiance_update(wd, field, rays=64, time_seed=5, alpha=0.2, max_steps=max_steps)
    ref = orc.radiance_update(wd, ref, requests=req, rays=64, time_seed=6, size=2, shift=1, alpha=0.2,
                              max_steps=max_steps)
    assert np.array_equal(got, ref)


def test_radiance_refill_variant_fp16_and_rays(gpu_ctx):
    """FP16 storage through the refill kernel equals the lane kernel's; a
    ray count other than 64 is refused for it."""
    import torch
    from lum_amd._lib import RADIANCE_F16
    size = (20, 16, 8)
    blocks, pal, mats, field = make_case(size, 93)
    outs = []
    for variant in (0, 3):
        w = gpu_world(gpu_ctx, blocks, pal, mats, field, fmt=RADIANCE_F16)
        w.set_tuning(variant=variant, probes_per_batch=4)
        try:
            for f in range(2):
                w.radiance_update(None, radiance_params(rays=64, time_seed=f, alpha=0.1))
            torch.cuda.synchronize()
            outs.append(w.read_radiance())
            if variant == 3:
                with pytest.raises(Exception):
                    w.radiance_update(None, radiance_params(rays=32, time_seed=2, alpha=0.1))
        finally:
            w.close()
    assert np.array_equal(outs[0], outs[1])
