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
        self.models, self.placements, self.model_voxels = [], [], []
        if n_models:
            vox, quats, shifts = synth.make_models(rng, n=n_models, lo=model_lo, hi=model_hi, world_size=size)
            self.model_voxels = vox
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


class ServingWorkload:
    """The serving configuration: n_worlds resident worlds of log-uniform size
    lo..hi blocks (one shared draw per world for every axis), each with a
    64-block random palette, random materials and field, FP16 radiance by
    default, a random probe permutation (request lists are windows of it)
    and a width x height G-buffer.  A trace is a seeded mix of 80% radiance
    updates (a random 1-100% probe subset, 32/64/128 rays) and 20% glossy
    queries."""

    def __init__(self, ctx, n_worlds=200, seed=4242, lo=(16, 16, 8), hi=(128, 128, 32), frame=(640, 360),
                 radiance_format=None, n_palette=64):
        import torch
        from ._lib import RADIANCE_F16
        fmt = RADIANCE_F16 if radiance_format is None else radiance_format
        self.rng = np.random.default_rng(seed)
        rng = self.rng
        self.sizes = synth.log_uniform_sizes(rng, n_worlds, lo, hi)
        self.frame = frame
        self.worlds, self.hosts, self.perms, self.perm_host, self.gbufs = [], [], [], [], []
        W, H = frame
        for size in self.sizes:
            blocks = synth.make_world(size, rng, n_palette)
            pal = synth.make_palette(rng, n_palette)
            mats = synth.make_materials(rng)
            field = synth.make_field(size, rng)
            w = World(ctx, size, static_blocks=n_palette, palette_slots=n_palette, radiance_format=fmt)
            w.set_blocks(blocks)
            w.set_palette(pal)
            w.set_materials(mats)
            w.write_radiance(field)
            w.upload_blocks()
            self.worlds.append(w)
            self.hosts.append((blocks, pal, mats, field))
            wx, wy, wz = size
            n = wx * wy * wz
            order = rng.permutation(n)
            q = np.stack([order % wx, (order // wx) % wy, order // (wx * wy), np.zeros(n, np.int64)],
                         1).astype(np.int32)
            q2 = np.concatenate([q, q], 0)  # a window starting anywhere wraps around
            self.perm_host.append(q2)
            self.perms.append(torch.from_numpy(q2).cuda())
            mn, dp = synth.make_gbuffer(W, H, mats, rng)
            self.gbufs.append((torch.from_numpy(mn).cuda(), torch.from_numpy(dp).cuda(), camera_default(W, H)))

    def n_probes(self, wi):
        s = self.sizes[wi]
        return s[0] * s[1] * s[2]

    def trace(self, n_req, rng=None):
        """n_req requests: ("r", world, window offset, count, rays) or
        ("g", world, 0, 0, 0)."""
        rng = self.rng if rng is None else rng
        out = []
        for _ in range(n_req):
            wi = int(rng.integers(len(self.worlds)))
            if rng.random() < 0.8:
                n = self.n_probes(wi)
                cnt = max(1, int(round(rng.uniform(0.01, 1.0) * n)))
                out.append(("r", wi, int(rng.integers(n)), cnt, int(rng.choice([32, 64, 128]))))
            else:
                out.append(("g", wi, 0, 0, 0))
        return out

    def request_list(self, wi, off, cnt):
        """Device view of the trace's probe list (int32x4, rows off..off+cnt)."""
        return self.perms[wi][off:off + cnt]

    def close(self):
        for w in self.worlds:
            w.close()
        self.worlds = []
