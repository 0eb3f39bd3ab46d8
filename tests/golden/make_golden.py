"""Writes tests/golden/lum_small.npz: small seeded inputs and the CPU
oracle's outputs for the radiance update (two frames, 64 and 128 rays),
blockify + map.comp (three rotated models) and the glossy pass.

The reference repository holds no golden vectors or fixtures for this path
and its shaders cannot run here, so these vectors freeze the oracle's
results (pinned separately by tests/test_oracle_pins.py) for regression and
for comparing the HIP kernels without rebuilding inputs.
Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from lum_amd import synth  # noqa: E402
from lum_amd.api import placement, camera_default  # noqa: E402
import tests._oracle as orc  # noqa: E402


def main():
    rng = np.random.default_rng(20261202)
    size = (8, 8, 4)
    static = 8
    blocks = synth.make_world(size, rng, static)
    pal = synth.make_palette(rng, static)
    mats = synth.make_materials(rng)
    field = synth.make_field(size, rng)
    slots = size[0] * size[1] * size[2] + static
    full = np.zeros((slots, 16, 16, 16), np.uint8)
    full[:static] = pal
    wd = orc.WorldData(blocks, full, mats)
    r1 = orc.radiance_update(wd, field, rays=64, time_seed=3, alpha=0.1, max_steps=64)
    r2 = orc.radiance_update(wd, r1, rays=128, time_seed=4, alpha=0.1)

    models, quats, shifts = synth.make_models(rng, n=3, lo=6, hi=14, world_size=size)
    pls = [placement(q, s, (m.shape[2], m.shape[1], m.shape[0])) for q, s, m in zip(quats, shifts, models)]
    cur = blocks.copy()
    copy_src = np.zeros(slots, np.int32)
    c = static
    for p in pls:
        c = orc.blockify(size, cur, static, c, list(p.aabb_min), list(p.aabb_max), copy_src)
    mapped = full.copy()
    orc.exec_copies(mapped, static, c, copy_src)
    for m, p in zip(models, pls):
        orc.map_model(size, cur, mapped, static, m, list(p.minv), list(p.shift)[:3], list(p.groups))

    W, H = 32, 16
    mn, dp = synth.make_gbuffer(W, H, mats, rng, depth_hi=60.0)
    cam = camera_default(W, H)
    ubo = np.array(list(cam.campos) + list(cam.camdir) + list(cam.horiz_scaled) + list(cam.verti_scaled)
                   + list(cam.light_dir) + list(cam.frame_size), np.float32)
    gl, gk = orc.glossy(orc.WorldData(cur, mapped, mats), r2, mn, dp, ubo, 128.0, 128)

    out = dict(blocks=blocks, palette=pal, mats=mats, field=field, radiance_f1=r1, radiance_f2=r2,
               blockified=cur, counter=np.int32(c), palette_mapped=mapped[:c], matnorm=mn, depth=dp, ubo=ubo,
               glossy=gl, glossy_keys=gk)
    for i, (m, p) in enumerate(zip(models, pls)):
        out[f"model{i}"] = m
        out[f"minv{i}"] = np.array(list(p.minv), np.float32)
        out[f"shift{i}"] = np.array(list(p.shift), np.int32)
        out[f"groups{i}"] = np.array(list(p.groups), np.int32)
        out[f"aabb{i}"] = np.array(list(p.aabb_min) + list(p.aabb_max), np.float32)
    np.savez_compressed(os.path.join(HERE, "lum_small.npz"), **out)
    print("wrote", os.path.join(HERE, "lum_small.npz"))


if __name__ == "__main__":
    main()
