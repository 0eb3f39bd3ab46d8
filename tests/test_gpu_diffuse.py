"""diffuse.frag on the GPU against the oracle.

Bar: bit-exact with the oracle for every storage format (no march; the
oracle samples the field as stored, and the arithmetic contract is shared).
"""
import numpy as np
import pytest

from lum_amd import synth
from lum_amd.api import camera_default
from lum_amd._lib import RADIANCE_F16, RADIANCE_BF16, RADIANCE_F32
import tests._oracle as orc

from tests.test_gpu_parity import make_case, gpu_world, full_palette

pytestmark = pytest.mark.gpu


def ubo_of(cam):
    return np.array(list(cam.campos) + list(cam.camdir) + list(cam.horiz_scaled) + list(cam.verti_scaled)
                    + list(cam.light_dir) + list(cam.frame_size), np.float32)


def lightmap_case(rng, n=1024):
    """A D16 shadow map and an orthographic light projection over the
    world (column-major), as the lightmap pass would produce."""
    lm = rng.integers(0, 65536, (n, n), dtype=np.uint16)
    s = 1.0 / 400.0
    proj = np.array([[s, 0, 0, 0], [0, s, 0, 0], [0.2 * s, 0.1 * s, -0.002, 0], [-0.6, -0.6, -0.3, 1]],
                    np.float32)   # columns
    return lm, proj.reshape(16)


@pytest.mark.parametrize("fmt", [RADIANCE_F32, RADIANCE_F16, RADIANCE_BF16])
@pytest.mark.parametrize("with_lightmap", [False, True])
def test_diffuse_matches_oracle(gpu_ctx, fmt, with_lightmap):
    import torch
    size = (48, 48, 16)
    blocks, pal, mats, field = make_case(size, 81)
    w = gpu_world(gpu_ctx, blocks, pal, mats, field, fmt=fmt)
    try:
        rng = np.random.default_rng(82)
        W, H = 640, 360
        mn, dp = synth.make_gbuffer(W, H, mats, rng, rough_max=1.0)
        cam = camera_default(W, H)
        lm, proj = lightmap_case(rng) if with_lightmap else (None, None)
        out = w.diffuse(torch.from_numpy(mn).cuda(), torch.from_numpy(dp).cuda(), cam,
                        None if lm is None else torch.from_numpy(lm.view(np.int16)).cuda(), proj)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        # the oracle samples the field as stored
        stored = w.read_radiance()
        wd = orc.WorldData(blocks, full_palette(w, pal), mats)
        ref = orc.diffuse(wd, stored, mn, dp, ubo_of(cam), lm, proj)
        assert np.array_equal(got, ref)
        if with_lightmap:
            # the sun term is exercised: a good share of pixels see a lit tap
            no_sun = orc.diffuse(wd, stored, mn, dp, ubo_of(cam))
            lit = np.any(ref[..., :3] != no_sun[..., :3], axis=-1)
            assert lit.mean() > 0.05
    finally:
        w.close()
