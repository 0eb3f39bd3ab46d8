"""The kernels' own march and sampling source (lum_amd/csrc/march.hpp),
compiled for the host, against the CPU oracle ray by ray (no GPU).

This pins the product's arithmetic at the source level: cast_ray (byte
palette), cast_ray_occ (occupancy bits) and sample_radiance must give the
oracle's hit, voxel, sample counts, hit key, fraction and normal bit for bit
on seeded worlds and rays, including rays that start outside the world.
"""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from lum_amd import synth
import tests._oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def mh(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("march_host") / "libmarch_host.so")
    subprocess.run([hipcc, "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17", "-fPIC", "-shared",
                    "-ffp-contract=off", os.path.join(HERE, "march_host.cpp"), "-o", out],
                   check=True, capture_output=True)
    return C.CDLL(out)


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def occupancy_bits(blocks, palette):
    ids = blocks.reshape(-1)
    eb = np.zeros((ids.size + 31) // 32, np.uint32)
    empty = np.nonzero(ids == 0)[0]
    np.bitwise_or.at(eb, empty >> 5, (np.uint32(1) << (empty & 31).astype(np.uint32)))
    occ = (palette.reshape(palette.shape[0], 4096) != 0)
    pb = np.packbits(occ.reshape(-1, 32)[:, ::-1], axis=1).view(">u4").astype(np.uint32).reshape(-1)
    return eb, pb


def rays(rng, size, n):
    ext = np.array(size, np.float64) * 16
    o = rng.uniform(-0.2 * ext, 1.2 * ext, (n, 3))
    o[: n // 2] = rng.uniform(0.0, ext, (n // 2, 3))   # half start inside
    d = rng.normal(size=(n, 3))
    d[: n // 8, rng.integers(0, 3)] *= 1e-3             # near-axis directions
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o.astype(np.float32), d.astype(np.float32)


@pytest.mark.parametrize("occ", [0, 1])
@pytest.mark.parametrize("seed,size,max_steps", [(1, (12, 10, 6), 0), (2, (20, 16, 8), 64), (3, (9, 9, 9), 7)])
def test_host_built_cast_ray_matches_oracle(mh, occ, seed, size, max_steps):
    rng = np.random.default_rng(seed)
    blocks = synth.make_world(size, rng, 32)
    pal = synth.make_palette(rng, 32)
    pal[0] = 0
    pal[rng.integers(1, 32, 4)] = 0                      # some all-air slots
    mats = synth.make_materials(rng)
    wd = orc.WorldData(blocks, pal, mats)
    eb, pb = occupancy_bits(blocks, pal)
    n = 1500
    o, d = rays(rng, size, n)
    info = np.zeros((n, 5), np.int32)
    fn = np.zeros((n, 4), np.float32)
    mh.mh_cast_rays(_p(blocks), _p(pal), _p(mats), size[0], size[1], size[2], pal.shape[0], 1, _p(eb), _p(pb),
                    occ, n, _p(o), _p(d), C.c_float(128.0), max_steps, _p(info), _p(fn))
    field = np.zeros((size[2], size[1], size[0], 4), np.float32)
    hits = 0
    for i in range(n):
        r = orc.trace_ray(wd, field, o[i], d[i], 128.0, max_steps)
        ref = [int(r["hit"]), r["voxel"], r["coarse"], r["refine"], r["key"]]
        assert list(info[i]) == ref, (i, o[i], d[i])
        if r["hit"]:
            hits += 1
            assert fn[i, 0] == np.float32(r["fraction"]) and np.array_equal(fn[i, 1:], r["normal"]), i
    assert 0.2 * n < hits < n


def test_fast_division_is_exact_below_2_pow_28(mh):
    """Probe coordinates use n / d as a multiply-high and a shift: exact for
    every n < 2^28 and d in [1, 2^24] (world axes up to 4096, wx * wy up to
    2^24).  Checked on every power of two and its neighbours, the world sizes
    of the workloads and random divisors, with random and edge dividends."""
    rng = np.random.default_rng(3)
    ds = {1, 2, 3, 5, 7, 48, 48 * 48, 256, 256 * 256, 4096, 4096 * 4096, 130, 130 * 9, 1 << 24}
    for k in range(1, 25):
        ds.update({(1 << k) - 1, 1 << k, (1 << k) + 1})
    ds.update(int(v) for v in rng.integers(1, 1 << 24, 300))
    top = (1 << 28) - 1
    for d in sorted(x for x in ds if 1 <= x <= (1 << 24)):
        n = np.concatenate([rng.integers(0, 1 << 28, 2000), np.arange(0, 64), top - np.arange(64),
                            (np.arange(1, 200) * d) - 1, np.arange(1, 200) * d]).astype(np.uint64)
        n = n[n <= top].astype(np.uint32)
        out = np.zeros_like(n)
        mh.mh_fast_div(_p(n), len(n), C.c_uint32(d), _p(out))
        assert np.array_equal(out, n // np.uint32(d)), d


def test_host_built_sample_radiance_matches_oracle(mh):
    rng = np.random.default_rng(9)
    size = (7, 5, 4)
    field = rng.random((size[2], size[1], size[0], 4)).astype(np.float32)
    n = 4000
    # texel coordinates from below -n to beyond 2n on every axis: both the
    # two-subtract REPEAT wrap and its division fallback
    x = rng.uniform(-40, 150, (n, 3)).astype(np.float32)
    x[n // 2:] = rng.uniform(-260, 400, (n - n // 2, 3)).astype(np.float32)
    nr = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    out = np.zeros((n, 3), np.float32)
    mh.mh_sample_radiance(_p(field), size[0], size[1], size[2], n, _p(x), _p(nr), _p(out))
    for i in range(0, n, 7):
        assert np.array_equal(out[i], orc.sample_radiance(field, x[i], nr[i])), i
