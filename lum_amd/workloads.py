"""Seeded multi-world workloads of the benchmark configurations, built on the
device through lum_amd.api (synthetic data only)."""
import numpy as np

from . import synth
from .api import World, Model, placement, camera_default, radiance_params
from ._lib import RADIANCE_F32


class WorldCase:
    """One resident world with its models, G-buffer and parameters."""

    def __init__(self, ctx, size, rng, rays, n_palette=64, n_models=0, model_lo=16, model_hi=32,
                 gbuffer=None, radiance_format=RADIANCE_F32, max_steps=64, alpha=0.1):
        import torch
        self.size = tuple(size)
        self.rays = rays
        self.blocks = synth.make_world(size, rng, n_palette)
        self.palette = synth.make_palette(rng, n_palette)
        self.mats = synth.make_materials(rng)
        self.field = synth.make_field(size, rng)
        self.world = World(ctx, size, static_blocks=n_palette, radiance_format=radiance_format)
        self.world.set_blocks(self.blocks)
        self.world.set_palette(self.palette)
        self.world.set_materials(self.mats)
        self.world.write_radiance(self.field)
        self.world.upload_blocks()
        self.models, self.placements = [], []
        if n_models:
            vox, quats, shifts = synth.make_models(rng, n=n_models, lo=model_lo, hi=model_hi, world_size=size)
            self.models = [Model(ctx, v) for v in vox]
            self.placements = [placement(q, s, m.size) for q, s, m in zip(quats, shifts, self.models)]
        self.gbuf = None
        if gbuffer:
            W, H = gbuffer
            mn, dp = synth.make_gbuffer(W, H, self.mats, rng)
            self.gbuf = (torch.from_numpy(mn).cuda(), torch.from_numpy(dp).cuda(), camera_default(W, H))
            self.out = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
        self.params = radiance_params(rays=rays, time_seed=0, alpha=alpha, max_steps=max_steps)

    @property
    def probe_rays(self):
        n = self.size[0] * self.size[1] * self.size[2]
        return n * self.rays

    @property
    def pixels(self):
        return 0 if self.gbuf is None else int(self.gbuf[1].numel())

    def close(self):
        for m in self.models:
            m.close()
        self.world.close()


def batched_config(ctx, seed=2024, n_worlds=64, lo=(16, 16, 8), hi=(96, 96, 32), rays=(32, 64, 128),
                   n_models=4, gbuffer=(640, 360), radiance_format=RADIANCE_F32):
    """64 worlds of log-uniform size 16x16x8..96x96x32, rays per probe drawn
    from {32, 64, 128}, 4 models of 16^3-32^3 each, a 640x360 G-buffer each."""
    rng = np.random.default_rng(seed)
    sizes = synth.log_uniform_sizes(rng, n_worlds, lo, hi)
    cases = []
    for size in sizes:
        r = int(rng.choice(rays))
        cases.append(WorldCase(ctx, size, rng, r, n_models=n_models, gbuffer=gbuffer,
                               radiance_format=radiance_format))
    return cases
