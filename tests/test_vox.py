""".vox import (lumx_vox_parse) against the reference's own reader.

The reference reads MagicaVoxel files with ogt_vox (common/ogt_vox.hpp);
oracle/_ref/libref_vox.so is that reader compiled from the reference tree
(oracle/Makefile, `make ref`) behind oracle/ref_vox_shim.cpp, which applies
Lum's extract_palette_mem (src/renderer/src/load_stuff.cpp:233-272).  Each
file must give the same first model (size and voxels), the same number of
distinct models and the same Material table, bit for bit.  Files: every
.vox asset the reference ships (where the reference tree is present) and
synthetic files written here that cover the default palette, the palette
rotation, MATL/MATT material types, IMAP reordering, empty and duplicate
models.  The synthetic files also carry closed-form checks that do not
depend on the reference build.
"""
import ctypes as C
import glob
import os
import struct

import numpy as np
import pytest

from lum_amd._lib import load

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SO = os.path.join(HERE, "..", "oracle", "_ref", "libref_vox.so")
ASSETS = "/root/reference/assets"


def parse(fn, data):
    buf = C.create_string_buffer(data, len(data))
    size = (C.c_int * 3)()
    n_models = C.c_int(-1)
    rc = fn(buf, len(data), size, None, 0, None, C.byref(n_models))
    if rc != 0:
        return rc, None, None, None, None
    vox = np.zeros(size[0] * size[1] * size[2], np.uint8)
    mats = np.zeros((256, 6), np.float32)
    rc = fn(buf, len(data), size, vox.ctypes.data_as(C.c_void_p), vox.size, mats.ctypes.data_as(C.c_void_p),
            C.byref(n_models))
    return rc, tuple(size), vox.reshape(size[2], size[1], size[0]), mats, n_models.value


def lumx_parse(data):
    return parse(load().lumx_vox_parse, data)


@pytest.fixture(scope="module")
def ref_parse():
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref/libref_vox.so not built (make -C oracle ref)")
    lib = C.CDLL(REF_SO)
    lib.ref_vox_parse.restype = C.c_int
    lib.ref_vox_parse.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p,
                                  C.c_void_p]
    return lambda data: parse(lib.ref_vox_parse, data)


# ---------------------------------------------------------------- a writer
def chunk(cid, content=b"", children=b""):
    return cid + struct.pack("<II", len(content), len(children)) + content + children


def vox_string(s):
    b = s.encode()
    return struct.pack("<I", len(b)) + b


def matl(mid, props):
    d = struct.pack("<I", len(props)) + b"".join(vox_string(k) + vox_string(v) for k, v in props.items())
    return chunk(b"MATL", struct.pack("<i", mid) + d)


def model_chunks(vox):
    """vox: uint8 (sz, sy, sx); returns SIZE + XYZI chunks."""
    sz, sy, sx = vox.shape
    z, y, x = np.nonzero(vox)
    xyzi = np.stack([x, y, z, vox[z, y, x]], 1).astype(np.uint8).tobytes()
    return chunk(b"SIZE", struct.pack("<III", sx, sy, sz)) + chunk(b"XYZI", struct.pack("<I", len(x)) + xyzi)


def vox_file(children, version=150):
    return b"VOX " + struct.pack("<I", version) + chunk(b"MAIN", b"", children)


def random_model(rng, shape, fill=0.3):
    return ((rng.random(shape) < fill) * rng.integers(1, 256, shape)).astype(np.uint8)


def synthetic_cases():
    rng = np.random.default_rng(11)
    cases = {}
    m = random_model(rng, (5, 6, 7))
    cases["default_palette"] = (vox_file(model_chunks(m)), m)
    pal = np.stack([np.arange(256), 255 - np.arange(256), np.full(256, 7), np.full(256, 200)], 1).astype(np.uint8)
    m2 = random_model(rng, (16, 16, 16))
    body = model_chunks(m2) + chunk(b"RGBA", pal.tobytes())
    body += matl(3, {"_type": "_emit", "_emit": "0.5", "_flux": "2"})
    body += matl(4, {"_type": "_metal", "_rough": "0.3", "_metal": "0.8"})
    body += matl(5, {"_type": "_glass", "_rough": "0.1"})
    body += matl(6, {"_type": "_blend"})
    body += matl(256, {"_type": "_emit", "_emit": "1.25"})
    cases["rgba_matl"] = (vox_file(body, 200), m2)
    m3 = random_model(rng, (3, 4, 5))
    empty = np.zeros((2, 2, 2), np.uint8)
    body = model_chunks(empty) + model_chunks(m3) + model_chunks(m3) + model_chunks(random_model(rng, (2, 3, 4)))
    cases["empty_and_duplicate_models"] = (vox_file(body), m3)
    m4 = random_model(rng, (4, 4, 4))
    imap = rng.permutation(256).astype(np.uint8)
    body = model_chunks(m4) + chunk(b"RGBA", pal.tobytes()) + chunk(b"IMAP", imap.tobytes())
    body += matl(9, {"_type": "_metal", "_rough": "0.5", "_metal": "0.25"})
    cases["imap"] = (vox_file(body), m4)
    matt = struct.pack("<iifI", 7, 3, 0.75, 0) + struct.pack("<f", 1.0)   # emissive, one extra property
    m5 = random_model(rng, (3, 3, 3))
    cases["matt_and_unknown_chunk"] = (vox_file(model_chunks(m5) + chunk(b"MATT", matt) +
                                                chunk(b"rOBJ", b"\x00" * 20)), m5)
    return cases


SYNTH = synthetic_cases()


@pytest.mark.parametrize("name", sorted(SYNTH))
def test_synthetic_files_match_the_reference_reader(ref_parse, name):
    data, _ = SYNTH[name]
    got, ref = lumx_parse(data), ref_parse(data)
    assert got[0] == 0 and ref[0] == 0
    assert got[1] == ref[1]
    assert np.array_equal(got[2], ref[2])
    assert np.array_equal(got[3], ref[3])
    assert got[4] == ref[4]


def reference_assets():
    return sorted(glob.glob(os.path.join(ASSETS, "*.vox")))


@pytest.mark.parametrize("path", reference_assets() or ["(no reference assets)"],
                         ids=[os.path.basename(p) for p in reference_assets()] or ["none"])
def test_reference_assets_match_the_reference_reader(ref_parse, path):
    if not os.path.exists(path):
        pytest.skip("reference assets not present")
    data = open(path, "rb").read()
    got, ref = lumx_parse(data), ref_parse(data)
    assert got[0] == ref[0] == 0
    assert got[1] == ref[1]
    assert np.array_equal(got[2], ref[2])
    assert np.array_equal(got[3], ref[3])
    assert got[4] == ref[4]


# ---------------------------------------------------------------- closed forms
def test_voxels_roundtrip_and_default_palette():
    data, m = SYNTH["default_palette"]
    rc, size, vox, mats, n = lumx_parse(data)
    assert rc == 0 and size == (7, 6, 5) and n == 1
    assert np.array_equal(vox, m)
    # colour index 1 = stored entry 0 = white; 255 = stored entry 254 = grey 0x11;
    # index 0 = the last stored entry (black) with alpha 0
    assert np.array_equal(mats[1, :4], np.float32([255, 255, 255, 255]) / 256)
    assert np.array_equal(mats[255, :4], np.float32([0x11, 0x11, 0x11, 255]) / 256)
    assert np.array_equal(mats[0, :4], np.float32([0, 0, 0, 0]))
    # cube: index 216 = stored 215 = first red ramp step 0xee; index 215 = (00, 00, 33)
    assert np.array_equal(mats[216, :4] * 256, [0xee, 0, 0, 255])
    assert np.array_equal(mats[215, :4] * 256, [0, 0, 0x33, 255])
    # no MATL: every material is diffuse, roughness 1, no emission
    assert np.all(mats[:, 4] == 0) and np.all(mats[:, 5] == 1)


def test_palette_rotation_and_material_types():
    data, m = SYNTH["rgba_matl"]
    rc, size, vox, mats, n = lumx_parse(data)
    assert rc == 0 and np.array_equal(vox, m)
    for i in [1, 2, 100, 255]:
        assert np.array_equal(mats[i, :4] * 256, [i - 1, 255 - (i - 1), 7, 200])
    assert np.array_equal(mats[0, :4] * 256, [255, 0, 7, 0])
    assert mats[3, 4] == np.float32(0.5 * (2 + 2 * 4)) and mats[3, 5] == np.float32(0.5)
    assert mats[4, 5] == np.float32((0.3 + (1 - 0.8)) / 2) and mats[4, 4] == 0
    assert mats[5, 5] == 0 and mats[6, 5] == 0            # glass, blend
    assert mats[0, 4] == np.float32(1.25 * 2) and mats[0, 5] == np.float32(0.5)   # MATL 256 is index 0
    assert mats[7, 5] == 1                                # untouched: diffuse


def test_empty_and_duplicate_models_are_dropped():
    data, m = SYNTH["empty_and_duplicate_models"]
    rc, size, vox, mats, n = lumx_parse(data)
    assert rc == 0 and n == 2 and size == (5, 4, 3) and np.array_equal(vox, m)


def test_not_a_vox_file_is_an_error():
    rc = lumx_parse(b"RIFF" + b"\x00" * 64)[0]
    assert rc != 0
    assert lumx_parse(vox_file(b""))[0] != 0               # no model


def test_oversized_models_and_short_buffers_are_errors():
    lib = load()
    # a side over 1024 voxels
    big = chunk(b"SIZE", struct.pack("<III", 2000, 1, 1)) + chunk(b"XYZI", struct.pack("<I", 1) + bytes([0, 0, 0, 1]))
    assert lumx_parse(vox_file(big))[0] != 0
    # a voxel buffer smaller than the model: an error, with the size reported
    data, m = SYNTH["default_palette"]
    buf = C.create_string_buffer(data, len(data))
    size = (C.c_int * 3)()
    small = np.zeros(m.size - 1, np.uint8)
    rc = lib.lumx_vox_parse(buf, len(data), size, small.ctypes.data_as(C.c_void_p), small.size, None, None)
    assert rc != 0 and tuple(size) == (m.shape[2], m.shape[1], m.shape[0])
