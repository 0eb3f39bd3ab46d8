"""The `scene` world-file format (load_scene, src/renderer/src/load_stuff.cpp:
187-217) through lumx_scene_read, host only (no GPU).

Format: a 12-byte int32 header (stored size Sx, Sy, Sz) and int16 blocks at
x + Sx*y + Sx*Sy*z.  load_scene reads the region clamp(S, 0, world_size)
into origin_world, each block clamped to [0, static_block_palette_size];
blocks outside the region keep their values; fewer than 12 bytes clear the
world.  The expected worlds here are written with plain numpy slicing of
the file.  The reference's own assets/scene (a 48 x 48 x 16 world) is read
where the reference tree is present.
"""
import ctypes as C
import os
import struct

import numpy as np
import pytest

from lum_amd import _lib

SCENE = "/root/reference/assets/scene"


def read(data, world_size, static, prev=None):
    lib = _lib.load()
    wx, wy, wz = world_size
    blocks = np.zeros((wz, wy, wx), np.int16) if prev is None else prev.copy()
    buf = C.create_string_buffer(data, len(data)) if data else None
    ws = (C.c_int * 3)(*world_size)
    rc = lib.lumx_scene_read(buf, len(data), ws, static, blocks.ctypes.data_as(C.c_void_p))
    return rc, blocks


def scene_bytes(stored):
    """stored: int16 array [Sz][Sy][Sx] -> file image."""
    sz, sy, sx = stored.shape
    return struct.pack("<3i", sx, sy, sz) + np.ascontiguousarray(stored, "<i2").tobytes()


def expected(stored, world_size, static, prev):
    sz, sy, sx = stored.shape
    wx, wy, wz = world_size
    rx, ry, rz = min(max(sx, 0), wx), min(max(sy, 0), wy), min(max(sz, 0), wz)
    out = prev.copy()
    out[:rz, :ry, :rx] = np.clip(stored[:rz, :ry, :rx], 0, static)
    return out


@pytest.mark.skipif(not os.path.exists(SCENE), reason="reference tree not present")
def test_reference_scene_asset():
    data = open(SCENE, "rb").read()
    sx, sy, sz = struct.unpack("<3i", data[:12])
    assert (sx, sy, sz) == (48, 48, 16) and len(data) == 12 + 2 * sx * sy * sz
    stored = np.frombuffer(data[12:], "<i2").reshape(sz, sy, sx)
    for world_size, static in [((48, 48, 16), 64), ((48, 48, 16), 5), ((32, 50, 8), 64), ((64, 64, 20), 20)]:
        prev = np.full(world_size[::-1], 3, np.int16)
        rc, got = read(data, world_size, static, prev)
        assert rc == 0
        assert np.array_equal(got, expected(stored, world_size, static, prev)), (world_size, static)
    rc, got = read(data, (48, 48, 16), 64)
    assert (got != 0).any() and got.max() <= 64 and got.min() >= 0


def test_header_larger_than_world_and_out_of_range_ids():
    rng = np.random.default_rng(0)
    stored = rng.integers(-300, 300, size=(7, 6, 9)).astype(np.int16)   # Sz, Sy, Sx
    world = (5, 8, 4)
    prev = rng.integers(0, 50, size=(4, 8, 5)).astype(np.int16)
    rc, got = read(scene_bytes(stored), world, 17, prev)
    assert rc == 0
    exp = expected(stored, world, 17, prev)
    assert np.array_equal(got, exp)
    # x reads with the stored row length (Sx = 9), not the world's (5)
    assert got[1, 2, 3] == np.clip(stored[1, 2, 3], 0, 17)
    # the region beyond the stored y = 6 keeps its previous blocks
    assert np.array_equal(got[:, 6:, :], prev[:, 6:, :])


def test_short_files():
    prev = np.full((2, 3, 4), 9, np.int16)
    for data in (b"", b"\x01\x00\x00"):
        rc, got = read(data, (4, 3, 2), 10, prev)
        assert rc == 0 and not got.any()          # origin_world.set(0)
    stored = np.arange(24, dtype=np.int16).reshape(2, 3, 4)
    rc, _ = read(scene_bytes(stored)[:-2], (4, 3, 2), 30, prev)      # one block short
    assert rc == _lib.LUMX_EINVAL
    rc, got = read(struct.pack("<3i", 0, 3, 2), (4, 3, 2), 30, prev)  # empty region: nothing read
    assert rc == 0 and np.array_equal(got, prev)
