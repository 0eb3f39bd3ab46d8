"""Seeded synthetic workloads of the benchmark configurations.

Worlds: int16 block ids, x-fastest, shape (wz, wy, wx); a block is solid
(a random static palette block 1..n_palette-1) below a seeded 2D Perlin
heightmap and air above.  Palette: n_palette random 16^3 uint8 blocks of
materials U{0..255}, block 0 all air.  Materials: 256 x {r, g, b, a,
emission, roughness} with rgb U[0,1], 2% emissive U[1,10], roughness U[0,1].
"""
import math

import numpy as np


def _perlin2(x, y, perm):
    """Classic 2D gradient noise in [-1, 1] (improved-noise fade curve)."""
    xi = np.floor(x).astype(np.int64)
    yi = np.floor(y).astype(np.int64)
    xf = x - xi
    yf = y - yi

    def fade(t):
        return t * t * t * (t * (t * 6 - 15) + 10)

    angles = (perm.astype(np.float64) / 256.0) * 2 * math.pi
    gx, gy = np.cos(angles), np.sin(angles)

    def grad(ix, iy, dx, dy):
        h = perm[(perm[ix & 255] + iy) & 255]
        return gx[h] * dx + gy[h] * dy

    n00 = grad(xi, yi, xf, yf)
    n10 = grad(xi + 1, yi, xf - 1, yf)
    n01 = grad(xi, yi + 1, xf, yf - 1)
    n11 = grad(xi + 1, yi + 1, xf - 1, yf - 1)
    u, v = fade(xf), fade(yf)
    nx0 = n00 + u * (n10 - n00)
    nx1 = n01 + u * (n11 - n01)
    return (nx0 + v * (nx1 - nx0)) * math.sqrt(2.0)


def heightmap(wx, wy, wz, rng, amplitude=None, wavelength=16.0):
    perm = rng.permutation(256)
    ys, xs = np.mgrid[0:wy, 0:wx].astype(np.float64)
    n = _perlin2(xs / wavelength + 0.37, ys / wavelength + 0.71, perm)
    amp = wz / 2.0 if amplitude is None else amplitude
    h = np.clip(np.round(wz / 2.0 + amp * n), 0, wz)
    return h.astype(np.int32)


def make_world(size, rng, n_palette=64, amplitude=None):
    wx, wy, wz = size
    h = heightmap(wx, wy, wz, rng, amplitude)
    z = np.arange(wz)[:, None, None]
    solid = z < h[None, :, :]
    ids = rng.integers(1, n_palette, size=(wz, wy, wx))
    return np.where(solid, ids, 0).astype(np.int16)


def make_palette(rng, n_palette=64):
    pal = rng.integers(0, 256, size=(n_palette, 16, 16, 16)).astype(np.uint8)
    pal[0] = 0
    return pal


def make_materials(rng):
    m = np.zeros((256, 6), np.float32)
    m[:, 0:3] = rng.random((256, 3))
    m[:, 3] = 1.0
    emissive = rng.random(256) < 0.02
    m[:, 4] = np.where(emissive, rng.uniform(1.0, 10.0, 256), 0.0)
    m[:, 5] = rng.random(256)
    m[0] = 0.0
    return m


def make_field(size, rng):
    wx, wy, wz = size
    f = rng.random((wz, wy, wx, 4)).astype(np.float32)
    f[..., 3] = 1.0
    return f


def uniform_quaternion(rng):
    """Shoemake's uniform random rotation, returned as (w, x, y, z)."""
    u1, u2, u3 = rng.random(3)
    a, b = math.sqrt(1 - u1), math.sqrt(u1)
    x = a * math.sin(2 * math.pi * u2)
    y = a * math.cos(2 * math.pi * u2)
    z = b * math.sin(2 * math.pi * u3)
    w = b * math.cos(2 * math.pi * u3)
    return (w, x, y, z)


def quat_matrix(q):
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ])


def make_models(rng, n=32, lo=16, hi=64, fill=0.3, world_size=(48, 48, 16)):
    """n models of edge U{lo..hi} per axis, `fill` of voxels set to U{1..255},
    each with a uniform rotation and a translation keeping its rotated box
    inside the world.  Returns (voxels[z][y][x] list, quats, shifts)."""
    ext = np.array(world_size, np.float64) * 16
    models, quats, shifts = [], [], []
    for _ in range(n):
        sx, sy, sz = (int(v) for v in rng.integers(lo, hi + 1, size=3))
        vox = np.where(rng.random((sz, sy, sx)) < fill, rng.integers(1, 256, size=(sz, sy, sx)), 0).astype(np.uint8)
        q = uniform_quaternion(rng)
        R = quat_matrix(q)
        corners = np.array([[cx, cy, cz] for cx in (0, sx - 1) for cy in (0, sy - 1) for cz in (0, sz - 1)], float)
        rc = corners @ R.T
        lo_c, hi_c = rc.min(0), rc.max(0)
        span = hi_c - lo_c
        shift = np.zeros(3)
        for a in range(3):
            room = ext[a] - 2.0 - span[a]
            t = rng.random() * max(room, 0.0)
            shift[a] = 1.0 + t - lo_c[a]
        models.append(vox)
        quats.append(q)
        shifts.append(tuple(float(v) for v in shift))
    return models, quats, shifts


def encode_normal(n):
    return np.clip(np.round((n * 0.5 + 0.5) * 255.0), 0, 255).astype(np.uint8)


def make_gbuffer(width, height, mats, rng, rough_max=0.3, depth_lo=1.0, depth_hi=200.0):
    """Synthetic G-buffer: material ids drawn among materials of roughness
    <= rough_max, normals uniform on the sphere (u8 encoded), view depth
    U[depth_lo, depth_hi] stored as depth / 1000."""
    glossy_ids = np.nonzero((mats[:, 5] <= rough_max) & (np.arange(256) > 0))[0]
    ids = rng.choice(glossy_ids, size=(height, width))
    v = rng.normal(size=(height, width, 3))
    v /= np.linalg.norm(v, axis=-1, keepdims=True)
    matnorm = np.zeros((height, width, 4), np.uint8)
    matnorm[..., 0] = ids
    matnorm[..., 1:] = encode_normal(v)
    depth = (rng.uniform(depth_lo, depth_hi, size=(height, width)) / 1000.0).astype(np.float32)
    return matnorm, depth


def log_uniform_sizes(rng, n, lo, hi):
    """n world sizes log-uniform between lo and hi, one shared draw per world
    applied to every axis (rounded to whole blocks)."""
    lo = np.array(lo, np.float64)
    hi = np.array(hi, np.float64)
    out = []
    for _ in range(n):
        t = rng.random()
        s = np.exp(np.log(lo) + t * (np.log(hi) - np.log(lo)))
        out.append(tuple(int(max(1, round(v))) for v in s))
    return out


PLAIN = dict(size=(48, 48, 16), n_palette=64, rays=64, alpha=0.1, max_dist=128.0, max_steps=64, seed=1234)
