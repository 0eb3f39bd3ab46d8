"""Parity at the benchmark configurations' full sizes.

The chain workload: the 48x48x16 world with 32 models of 16^3-64^3 (30% fill,
uniform rotations) mapped bit-exactly, every probe updated (64 rays, max 64
steps) and a 1920x1080 glossy march (max 128 steps), each against the CPU
oracle on the same inputs.
"""
import numpy as np
import pytest

from lum_amd import synth
from lum_amd.api import World, Model, placement, camera_default, radiance_params
import tests._oracle as orc

pytestmark = pytest.mark.gpu


def test_chain_config_full_size(gpu_ctx):
    import torch
    cfg = synth.PLAIN
    rng = np.random.default_rng(cfg["seed"])
    size = cfg["size"]
    blocks = synth.make_world(size, rng, 64)
    pal = synth.make_palette(rng, 64)
    mats = synth.make_materials(rng)
    field = synth.make_field(size, rng)
    vox, quats, shifts = synth.make_models(rng, n=32, lo=16, hi=64, fill=0.3, world_size=size)
    W, H = 1920, 1080
    mn, dp = synth.make_gbuffer(W, H, mats, rng)
    w = World(gpu_ctx, size, static_blocks=64)
    w.set_blocks(blocks)
    w.set_palette(pal)
    w.set_materials(mats)
    w.write_radiance(field)
    ms = [Model(gpu_ctx, v) for v in vox]
    try:
        pls = [placement(q, s, m.size) for q, s, m in zip(quats, shifts, ms)]
        w.map_models_placed(ms, pls)
        w.radiance_update(None, radiance_params(rays=64, time_seed=9, alpha=0.1, max_steps=64))
        cam = camera_default(W, H)
        out, keys = w.glossy(torch.from_numpy(mn).cuda(), torch.from_numpy(dp).cuda(), cam, 128.0, 128,
                             traced=True)
        torch.cuda.synchronize()
        counter = w.palette_counter()
        print(f"palette slots in use: {counter}")
        # oracle chain
        cur = blocks.copy()
        copy_src = np.zeros(w.palette_slots, np.int32)
        c = 64
        for p in pls:
            c = orc.blockify(size, cur, 64, c, list(p.aabb_min), list(p.aabb_max), copy_src)
        assert c == counter
        ref_pal = np.zeros((w.palette_slots, 16, 16, 16), np.uint8)
        ref_pal[:64] = pal
        orc.exec_copies(ref_pal, 64, c, copy_src)
        for v, p in zip(vox, pls):
            orc.map_model(size, cur, ref_pal, 64, v, list(p.minv), list(p.shift)[:3], list(p.groups))
        assert np.array_equal(w.get_blocks(True), cur)
        assert np.array_equal(w.get_palette(0, counter), ref_pal[:counter]), "map.comp bit-exact"
        wd = orc.WorldData(cur, ref_pal, mats)
        ref_f = orc.radiance_update(wd, field, rays=64, time_seed=9, alpha=0.1, max_steps=64)
        got_f = w.read_radiance()
        rel = np.abs(got_f - ref_f) / np.maximum(np.abs(ref_f), 1e-3)
        assert rel.max() <= 1e-3
        ubo = np.array(list(cam.campos) + list(cam.camdir) + list(cam.horiz_scaled) + list(cam.verti_scaled)
                       + list(cam.light_dir) + list(cam.frame_size), np.float32)
        ref_g, ref_k = orc.glossy(wd, ref_f, mn, dp, ubo, 128.0, 128)
        agree = (keys.cpu().numpy() == ref_k).mean()
        g = out.cpu().numpy()
        grel = np.abs(g - ref_g) / np.maximum(np.abs(ref_g), 1e-3)
        print(f"glossy key agreement {agree:.6f}, max rel {grel.max():.3e}")
        assert agree >= 0.999
        assert (grel <= 1e-3).mean() >= 0.999
    finally:
        for m in ms:
            m.close()
        w.close()


def test_batched_config_worlds_against_oracle(gpu_ctx):
    """Four worlds drawn by the batched configuration's generator, updated in
    one batched launch, each against the oracle."""
    import torch
    from lum_amd import workloads
    from lum_amd.api import Batch
    cases = workloads.batched_config(gpu_ctx, n_worlds=4, n_models=2, gbuffer=None)
    b = Batch(gpu_ctx, [c.world for c in cases])
    try:
        for c in cases:
            c.world.map_models_placed(c.models, c.placements)
        params = [radiance_params(rays=c.rays, time_seed=2, alpha=0.1, max_steps=64) for c in cases]
        b.radiance(params)
        torch.cuda.synchronize()
        for c, p in zip(cases, params):
            counter = c.world.palette_counter()
            full = np.zeros((c.world.palette_slots, 16, 16, 16), np.uint8)
            full[:counter] = c.world.get_palette(0, counter)
            wd = orc.WorldData(c.world.get_blocks(True), full, c.mats)
            ref = orc.radiance_update(wd, c.field, rays=c.rays, time_seed=2, alpha=0.1, max_steps=64)
            got = c.world.read_radiance()
            assert (np.abs(got - ref) / np.maximum(np.abs(ref), 1e-3)).max() <= 1e-3
    finally:
        b.close()
        for c in cases:
            c.close()
