"""ctypes binding of liblumx.so (include/lumx.h).

Loading fails loudly when the library has not been built: there is no
fallback path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblumx.so")

LUMX_OK = 0
LUMX_EINVAL = -1
LUMX_ENOMEM = -2
LUMX_EDEVICE = -3
LUMX_EOVERFLOW = -4
LUMX_EBUSY = -5

RADIANCE_F32 = 0
RADIANCE_F16 = 1
RADIANCE_BF16 = 2
PALETTE_XYZ_XFAST = 0
PALETTE_REF_ZFAST = 1


class LumxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"lumx error {code}: {msg}")
        self.code = code


class WorldDesc(C.Structure):
    _fields_ = [("size", C.c_int * 3), ("static_blocks", C.c_int), ("palette_slots", C.c_int),
                ("radiance_format", C.c_int)]


class RadianceParams(C.Structure):
    _fields_ = [("rays", C.c_int), ("time_seed", C.c_int), ("size", C.c_int), ("shift", C.c_int),
                ("alpha", C.c_float), ("max_dist", C.c_float), ("max_steps", C.c_int), ("reserved", C.c_int)]


class Transform(C.Structure):
    _fields_ = [("rot", C.c_float * 4), ("shift", C.c_float * 3)]


class Placement(C.Structure):
    _fields_ = [("minv", C.c_float * 16), ("shift", C.c_int * 4), ("groups", C.c_int * 3),
                ("aabb_min", C.c_float * 3), ("aabb_max", C.c_float * 3)]


class Tuning(C.Structure):
    _fields_ = [("variant", C.c_int), ("threads", C.c_int), ("probes_per_batch", C.c_int),
                ("lds_block_bits", C.c_int), ("lds_palette_bits", C.c_int), ("wg_per_cu", C.c_int),
                ("rays_per_lane", C.c_int), ("compact", C.c_int)]


class Camera(C.Structure):
    _fields_ = [("campos", C.c_float * 3), ("camdir", C.c_float * 3), ("horiz_scaled", C.c_float * 3),
                ("verti_scaled", C.c_float * 3), ("light_dir", C.c_float * 3), ("frame_size", C.c_float * 2)]


class Lightmap(C.Structure):
    _fields_ = [("depth_dev", C.c_void_p), ("width", C.c_int), ("height", C.c_int), ("proj", C.c_float * 16)]


class ServerDesc(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_batch", C.c_int), ("streams", C.c_int), ("block", C.c_int),
                ("rays_per_lane", C.c_int), ("direct", C.c_int)]


class Request(C.Structure):
    _fields_ = [("kind", C.c_int), ("world", C.c_int), ("requests_dev", C.c_void_p), ("n_requests", C.c_int),
                ("params", RadianceParams), ("matnorm_dev", C.c_void_p), ("depth_dev", C.c_void_p),
                ("width", C.c_int), ("height", C.c_int), ("cam", Camera), ("max_dist", C.c_float),
                ("max_steps", C.c_int), ("out_dev", C.c_void_p)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_longlong), ("completed", C.c_longlong), ("launches", C.c_longlong),
                ("probe_rays", C.c_longlong), ("pixels", C.c_longlong), ("evictions", C.c_longlong),
                ("restores", C.c_longlong), ("resident_bytes", C.c_size_t)]


REQ_RADIANCE = 0
REQ_GLOSSY = 1

# name -> (restype, argtypes)
_P = C.c_void_p
_I = C.c_int
_SIGS = {
    "lumx_context_create": (_I, [_I, C.POINTER(_P)]),
    "lumx_context_destroy": (_I, [_P]),
    "lumx_context_sync": (_I, [_P]),
    "lumx_context_stream": (_P, [_P]),
    "lumx_last_error": (C.c_char_p, []),
    "lumx_version": (C.c_char_p, []),
    "lumx_world_create": (_I, [_P, C.POINTER(WorldDesc), C.POINTER(_P)]),
    "lumx_world_destroy": (_I, [_P]),
    "lumx_world_info": (_I, [_P, C.POINTER(WorldDesc)]),
    "lumx_world_set_blocks": (_I, [_P, _P]),
    "lumx_world_get_blocks": (_I, [_P, _P, _I]),
    "lumx_world_set_block": (_I, [_P, _I, _I, _I, C.c_int16]),
    "lumx_world_get_block": (_I, [_P, _I, _I, _I, C.POINTER(C.c_int16)]),
    "lumx_world_set_palette": (_I, [_P, _I, _I, _P, _I]),
    "lumx_world_get_palette": (_I, [_P, _I, _I, _P]),
    "lumx_world_set_materials": (_I, [_P, _P]),
    "lumx_world_write_radiance": (_I, [_P, _P]),
    "lumx_world_read_radiance": (_I, [_P, _P]),
    "lumx_world_radiance_dev": (_I, [_P, C.POINTER(_P), C.POINTER(C.c_size_t)]),
    "lumx_world_upload_blocks": (_I, [_P, _P]),
    "lumx_model_create": (_I, [_P, _P, _I, _I, _I, C.POINTER(_P)]),
    "lumx_model_destroy": (_I, [_P]),
    "lumx_model_size": (_I, [_P, C.POINTER(C.c_int)]),
    "lumx_model_placement": (_I, [C.POINTER(Transform), C.POINTER(C.c_int), C.POINTER(Placement)]),
    "lumx_map_models": (_I, [_P, _I, C.POINTER(_P), C.POINTER(Transform), _P]),
    "lumx_map_models_placed": (_I, [_P, _I, C.POINTER(_P), C.POINTER(Placement), _P]),
    "lumx_world_palette_counter": (_I, [_P, C.POINTER(C.c_int)]),
    "lumx_radiance_update": (_I, [_P, _P, _I, C.POINTER(RadianceParams), _P]),
    "lumx_radiance_update_traced": (_I, [_P, _P, _I, C.POINTER(RadianceParams), _P, _P, _P]),
    "lumx_select_probes": (_I, [_P, _P, _I, _P, C.POINTER(C.c_int)]),
    "lumx_select_probes_device": (_I, [_P, _P, _I, _P, _P, _P]),
    "lumx_radiance_update_counted": (_I, [_P, _P, _P, _I, C.POINTER(RadianceParams), _P]),
    "lumx_world_shift_radiance": (_I, [_P, _I, _I, _I, _P]),
    "lumx_glossy": (_I, [_P, _P, _P, _I, _I, C.POINTER(Camera), C.c_float, _I, _P, _P, _P]),
    "lumx_batch_autotune": (_I, [_P, _P, _P, C.POINTER(RadianceParams), _I, C.POINTER(C.c_int),
                                 C.POINTER(C.c_int), C.POINTER(C.c_float)]),
    "lumx_batch_set_shape": (_I, [_P, _I, _I]),
    "lumx_map_models_device": (_I, [_P, _I, _P, _P, _P]),
    "lumx_map_models_device_noalloc": (_I, [_P, _I, _P, _P, _P]),
    "lumx_vox_parse": (_I, [_P, C.c_size_t, _P, _P, C.c_size_t, _P, _P]),
    "lumx_diffuse": (_I, [_P, _P, _P, _I, _I, C.POINTER(Camera), C.POINTER(Lightmap), _P, _P]),
    "lumx_camera_default": (_I, [_I, _I, C.POINTER(Camera)]),
    "lumx_scene_read": (_I, [_P, C.c_size_t, C.POINTER(C.c_int), _I, _P]),
    "lumx_world_last_pass_ms": (_I, [_P, _I, C.POINTER(C.c_float)]),
    "lumx_world_set_tuning": (_I, [_P, C.POINTER(Tuning)]),
    "lumx_world_get_tuning": (_I, [_P, C.POINTER(Tuning)]),
    "lumx_autotune_radiance": (_I, [_P, C.POINTER(RadianceParams), _I, C.POINTER(Tuning), C.POINTER(C.c_float)]),
    "lumx_batch_create": (_I, [_P, _I, C.POINTER(_P), C.POINTER(_P)]),
    "lumx_batch_destroy": (_I, [_P]),
    "lumx_batch_radiance": (_I, [_P, C.POINTER(_P), C.POINTER(C.c_int), C.POINTER(RadianceParams), _P]),
    "lumx_batch_glossy": (_I, [_P, C.POINTER(_P), C.POINTER(_P), C.POINTER(C.c_int), C.POINTER(C.c_int),
                               C.POINTER(Camera), C.c_float, _I, C.POINTER(_P), _P]),
    "lumx_batch_map_models": (_I, [_P, C.POINTER(C.c_int), C.POINTER(_P), C.POINTER(_P), _P]),
    "lumx_batch_set_frame": (_I, [_P, _I, _P]),
    "lumx_batch_advance_frame": (_I, [_P, _P]),
    "lumx_world_evict": (_I, [_P]),
    "lumx_world_restore": (_I, [_P]),
    "lumx_world_resident": (_I, [_P, C.POINTER(C.c_int), C.POINTER(C.c_size_t)]),
    "lumx_server_create": (_I, [_P, C.POINTER(ServerDesc), C.POINTER(_P)]),
    "lumx_server_destroy": (_I, [_P]),
    "lumx_server_add_world": (_I, [_P, _P, C.POINTER(C.c_int)]),
    "lumx_server_submit": (_I, [_P, C.POINTER(Request), C.POINTER(C.c_longlong)]),
    "lumx_server_wait": (_I, [_P, C.c_longlong, C.POINTER(C.c_double)]),
    "lumx_server_drain": (_I, [_P]),
    "lumx_server_get_stats": (_I, [_P, C.POINTER(ServerStats)]),
    "lumx_server_run_trace": (_I, [_P, _I, C.POINTER(Request), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double)]),
    "lumx_graph_begin": (_I, [_P]),
    "lumx_graph_end": (_I, [_P, C.POINTER(_P)]),
    "lumx_graph_launch": (_I, [_P, _P]),
    "lumx_graph_destroy": (_I, [_P]),
}

_lib = None


def exported_symbols():
    """Names declared in include/lumx.h that the binding expects."""
    return sorted(_SIGS)


def load():
    """Load liblumx.so (raises if it is missing: no fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m lum_amd.build`")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != LUMX_OK:
        msg = load().lumx_last_error().decode(errors="replace")
        raise LumxError(rc, msg)
    return rc
