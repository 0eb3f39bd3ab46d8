"""ctypes view of the C99 drop-in (include/clum.h) for tests and the bench.

The signatures are those of clum.h (the reference's api/crenderer.h plus the
lumx_lum_* extensions); the library is liblumx.so, loaded as in _lib.
"""
import ctypes as C

from ._lib import load


class LumSettings(C.Structure):
    _fields_ = [("world_size", C.c_int * 3), ("static_block_palette_size", C.c_int),
                ("maxParticleCount", C.c_int), ("timestampCount", C.c_int), ("fif", C.c_int),
                ("vsync", C.c_char), ("fullscreen", C.c_char), ("debug", C.c_char), ("profile", C.c_char)]


class LumMeshTransform(C.Structure):
    _fields_ = [("rot", C.c_float * 4), ("shift", C.c_float * 3)]


_P = C.c_void_p
_SIGS = {
    "lum_create_instance": (_P, [C.POINTER(LumSettings)]),
    "lum_init": (None, [_P, C.POINTER(LumSettings)]),
    "lum_load_world": (None, [_P, _P]),
    "lum_load_block": (None, [_P, C.c_int16, _P]),
    "lum_upload_block_palette_to_gpu": (None, [_P]),
    "lum_upload_material_palette_to_gpu": (None, [_P]),
    "lum_load_mesh_data": (_P, [_P, _P, C.c_int, C.c_int, C.c_int]),
    "lum_start_frame": (None, [_P]),
    "lum_draw_model": (None, [_P, _P, C.POINTER(LumMeshTransform)]),
    "lum_prepare_frame": (None, [_P]),
    "lum_submit_frame": (None, [_P]),
    "lum_wait_idle": (None, [_P]),
    "lum_cleanup": (None, [_P]),
    "lum_get_timestamp": (C.c_double, [_P, C.c_int]),
    "lum_get_timestamp_difference": (C.c_double, [_P, C.c_int, C.c_int]),
    "lumx_lum_set_materials": (None, [_P, _P]),
    "lumx_lum_request_radiance_update": (None, [_P, C.c_int, C.c_int, C.c_int]),
    "lumx_lum_set_gbuffer": (None, [_P, _P, _P, C.c_int, C.c_int, _P]),
    "lumx_lum_set_radiance_shift": (None, [_P, C.c_int, C.c_int, C.c_int]),
    "lumx_lum_world": (_P, [_P]),
    "lumx_world_read_radiance": (C.c_int, [_P, _P]),
    "lumx_world_palette_counter": (C.c_int, [_P, C.POINTER(C.c_int)]),
}


def lib():
    """liblumx.so with the clum.h entry points typed."""
    L = load()
    for name, (res, args) in _SIGS.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def settings(world_size, static_blocks):
    return LumSettings((C.c_int * 3)(*world_size), static_blocks, 8128, 1024, 2, b"\0", b"\0", b"\0", b"\0")


def transform(rot_wxyz, shift):
    return LumMeshTransform((C.c_float * 4)(*rot_wxyz), (C.c_float * 3)(*shift))
