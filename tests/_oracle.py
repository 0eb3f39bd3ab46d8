"""ctypes view of the CPU oracle (oracle/liblum_oracle.so), for tests, the
smoke check and the benchmark's CPU-baseline leg only."""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.environ.get("LUMX_ORACLE_SO") or os.path.join(ORACLE_DIR, "liblum_oracle.so")


class OrcWorld(C.Structure):
    _fields_ = [("wx", C.c_int), ("wy", C.c_int), ("wz", C.c_int), ("blocks", C.c_void_p),
                ("palette", C.c_void_p), ("palette_slots", C.c_int), ("materials", C.c_void_p)]


class OrcRadParams(C.Structure):
    _fields_ = [("rays", C.c_int), ("time_seed", C.c_int), ("size", C.c_int), ("shift", C.c_int),
                ("alpha", C.c_float), ("max_dist", C.c_float), ("max_steps", C.c_int), ("n_req", C.c_int)]


_lib = None
_lib_fma = None
ORACLE_FMA_SO = os.path.join(ORACLE_DIR, "liblum_oracle_fma.so")


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        _lib = C.CDLL(ORACLE_SO)
    return _lib


def load_contracted():
    """The oracle built with FMA contraction on (a fusing shader compiler's
    stand-in); only for measuring how much contraction moves results."""
    global _lib_fma
    if _lib_fma is None:
        if not os.path.exists(ORACLE_FMA_SO):
            build()
        _lib_fma = C.CDLL(ORACLE_FMA_SO)
    return _lib_fma


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class WorldData:
    """Host copy of a world: blocks (wz,wy,wx) int16, palette (slots,16,16,16)
    uint8 x-fastest, materials (256,6) float32."""

    def __init__(self, blocks, palette, mats):
        self.blocks = np.ascontiguousarray(blocks, np.int16)
        self.palette = np.ascontiguousarray(palette, np.uint8)
        self.mats = np.ascontiguousarray(mats, np.float32)
        wz, wy, wx = self.blocks.shape
        self.size = (wx, wy, wz)
        self.c = OrcWorld(wx, wy, wz, _p(self.blocks).value, _p(self.palette).value, self.palette.shape[0],
                          _p(self.mats).value)


def hash33(seed):
    out = np.zeros(3, np.float32)
    s = np.array(seed, np.uint32)
    load().orc_hash33(_p(s), _p(out))
    return out


def hash33_u(seed):
    out = np.zeros(3, np.uint32)
    s = np.array(seed, np.uint32)
    load().orc_hash33_u(_p(s), _p(out))
    return out


def fibonacci(i, n):
    out = np.zeros(3, np.float32)
    load().orc_fibonacci(C.c_int(i), C.c_int(n), _p(out))
    return out


def trace_ray(world, field, o, d, max_dist=128.0, max_steps=0):
    light = np.zeros(3, np.float32)
    info = np.zeros(5, np.int32)
    fn = np.zeros(4, np.float32)
    f = np.ascontiguousarray(field, np.float32)
    o = np.asarray(o, np.float32)
    d = np.asarray(d, np.float32)
    load().orc_trace_ray_f32(C.byref(world.c), _p(f), _p(o), _p(d), C.c_float(max_dist), C.c_int(max_steps),
                             _p(light), _p(info), _p(fn))
    return dict(light=light, hit=bool(info[0]), voxel=int(info[1]), coarse=int(info[2]), refine=int(info[3]),
                key=int(info[4]), fraction=float(fn[0]), normal=fn[1:4].copy())


def sample_radiance(field, x, n):
    f = np.ascontiguousarray(field, np.float32)
    wz, wy, wx = f.shape[:3]
    out = np.zeros(3, np.float32)
    load().orc_sample_radiance_f32(wx, wy, wz, _p(f), _p(np.asarray(x, np.float32)),
                                   _p(np.asarray(n, np.float32)), _p(out))
    return out


def radiance_update(world, field, requests=None, rays=64, time_seed=0, size=1, shift=0, alpha=0.1,
                    max_dist=128.0, max_steps=0, threads=8, traced=False, fp64=False, contracted=False):
    """Returns the new field (and per-ray keys / steps when traced).
    requests None = every probe in x-fastest order."""
    wx, wy, wz = world.size
    if requests is None:
        zz, yy, xx = np.meshgrid(np.arange(wz), np.arange(wy), np.arange(wx), indexing="ij")
        requests = np.stack([xx.ravel(), yy.ravel(), zz.ravel(), np.zeros(xx.size, int)], 1)
    req = np.ascontiguousarray(requests, np.int32)
    n_req = len(req)
    p = OrcRadParams(rays, time_seed, size, shift, alpha, max_dist, max_steps, n_req)
    n_rays = (n_req // size) * rays
    keys = np.zeros(max(n_rays, 1), np.int32) if traced else None
    steps = np.zeros(max(2 * n_rays, 2), np.int32) if traced else None
    kp = _p(keys) if traced else None
    sp = _p(steps) if traced else None
    if fp64:
        fin = np.ascontiguousarray(field, np.float64)
        fout = np.zeros_like(fin)
        load().orc_radiance_update_f64(C.byref(world.c), _p(fin), _p(fout), _p(req), C.byref(p), threads, kp, sp)
    else:
        fin = np.ascontiguousarray(field, np.float32)
        fout = np.zeros_like(fin)
        lib = load_contracted() if contracted else load()
        lib.orc_radiance_update_f32(C.byref(world.c), _p(fin), _p(fout), _p(req), C.byref(p), threads, kp, sp)
    if traced:
        return fout, keys[:n_rays], steps[: 2 * n_rays].reshape(-1, 2)
    return fout


def map_model(world_size, blocks, palette, static_count, model, minv, shift, groups, fused=False):
    """map.comp for one model, in place on `palette` (slots,16,16,16)."""
    wx, wy, wz = world_size
    b = np.ascontiguousarray(blocks, np.int16)
    m = np.ascontiguousarray(model, np.uint8)
    mz, my, mx = m.shape
    mv = np.ascontiguousarray(minv, np.float32)
    sh = np.ascontiguousarray(shift, np.int32)
    gr = np.ascontiguousarray(groups, np.int32)
    assert palette.flags["C_CONTIGUOUS"] and palette.dtype == np.uint8
    load().orc_map_model(wx, wy, wz, _p(b), _p(palette), palette.shape[0], static_count, _p(m), mx, my, mz,
                         _p(mv), _p(sh), _p(gr), 1 if fused else 0)


def blockify(world_size, world, static_count, counter, aabb_min, aabb_max, copy_src):
    wx, wy, wz = world_size
    assert world.flags["C_CONTIGUOUS"] and world.dtype == np.int16
    assert copy_src.dtype == np.int32
    return load().orc_blockify(wx, wy, wz, _p(world), static_count, counter,
                               _p(np.asarray(aabb_min, np.float32)), _p(np.asarray(aabb_max, np.float32)),
                               _p(copy_src), len(copy_src))


def exec_copies(palette, static_count, counter, copy_src):
    load().orc_exec_copies(_p(palette), palette.shape[0], static_count, counter, _p(copy_src))


def select_probes(world_blocks, special=None):
    b = np.ascontiguousarray(world_blocks, np.int16)
    wz, wy, wx = b.shape
    sp = np.zeros((0, 4), np.int32) if special is None else np.ascontiguousarray(special, np.int32)
    out = np.zeros((b.size + len(sp), 4), np.int32)
    n = load().orc_select_probes(wx, wy, wz, _p(b), _p(sp), len(sp), _p(out))
    return out[:n]


def glossy(world, field, matnorm, depth, ubo17, max_dist=128.0, max_steps=0, threads=8):
    f = np.ascontiguousarray(field, np.float32)
    mn = np.ascontiguousarray(matnorm, np.uint8)
    dp = np.ascontiguousarray(depth, np.float32)
    h, w = dp.shape
    u = np.ascontiguousarray(ubo17, np.float32)
    out = np.zeros((h, w, 4), np.float32)
    keys = np.zeros((h, w), np.int32)
    load().orc_glossy_f32(C.byref(world.c), _p(f), _p(mn), _p(dp), w, h, _p(u), C.c_float(max_dist),
                          C.c_int(max_steps), _p(out), _p(keys), threads)
    return out, keys


def shift_radiance(field, shift):
    """field: float32 (wz, wy, wx, c); returns the shifted copy."""
    f = np.ascontiguousarray(field, np.float32).copy()
    wz, wy, wx, c = f.shape
    load().orc_shift_radiance(wx, wy, wz, c, _p(f), int(shift[0]), int(shift[1]), int(shift[2]))
    return f


def diffuse(world, field, matnorm, depth, ubo17, lightmap=None, lm_proj=None):
    """diffuse.frag main; lightmap: uint16 (h, w) or None; lm_proj: 16
    floats column-major."""
    f = np.ascontiguousarray(field, np.float32)
    mn = np.ascontiguousarray(matnorm, np.uint8)
    dp = np.ascontiguousarray(depth, np.float32)
    h, w = dp.shape
    u = np.ascontiguousarray(ubo17, np.float32)
    out = np.zeros((h, w, 4), np.float32)
    if lightmap is None:
        lm, lw, lh = None, 0, 0
        pr = np.zeros(16, np.float32)
    else:
        lm = np.ascontiguousarray(lightmap, np.uint16)
        lh, lw = lm.shape
        pr = np.ascontiguousarray(lm_proj, np.float32).reshape(16)
    load().orc_diffuse_f32(C.byref(world.c), _p(f), _p(mn), _p(dp), w, h, _p(u),
                           None if lm is None else _p(lm), lw, lh, _p(pr), _p(out))
    return out
