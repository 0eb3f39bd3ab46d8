"""The C-ABI library builds, loads without a GPU and exports every entry
point that include/lumx.h and include/clum.h declare."""
import ctypes as C
import os
import re

import pytest

from lum_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b((?:lumx|lum)_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_loads_and_exports_lumx():
    lib = _lib.load()
    names = declared("lumx.h")
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_lib.exported_symbols())


def test_library_exports_clum():
    lib = _lib.load()
    names = declared("clum.h")
    assert "lum_submit_frame" in names and "lum_load_mesh_data" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_error_without_gpu():
    lib = _lib.load()
    assert lib.lumx_version().startswith(b"lumx")
    # argument validation happens before any device call
    assert lib.lumx_world_create(None, None, None) == _lib.LUMX_EINVAL
    assert b"null" in lib.lumx_last_error()


def test_clum_null_safety_without_gpu():
    """The lum_* shim never aborts on bad handles (reference: assert/abort)."""
    lib = _lib.load()
    lib.lum_submit_frame.argtypes = [C.c_void_p]
    lib.lum_submit_frame(C.c_void_p(12345))
    lib.lum_get_world_block.restype = C.c_int16
    lib.lum_get_world_block.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
    assert lib.lum_get_world_block(C.c_void_p(0), 0, 0, 0) == 0
