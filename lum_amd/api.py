"""Python view of the lumx C ABI (include/lumx.h).

Device buffers are torch tensors on the HIP device (torch is used for HBM
allocation and streams only); every pass runs in liblumx's HIP kernels on
torch's current stream.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import (Camera, Placement, RadianceParams, Transform, Tuning, WorldDesc, check, load, RADIANCE_F32,
                   PALETTE_XYZ_XFAST, ServerDesc, Request, ServerStats, REQ_RADIANCE, REQ_GLOSSY)

TUNING_FIELDS = ("variant", "threads", "probes_per_batch", "lds_block_bits", "lds_palette_bits", "wg_per_cu",
                 "rays_per_lane")


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return C.c_void_p(0 if t is None else t.data_ptr())


def _np_ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Context:
    def __init__(self, device=0):
        self.lib = load()
        self.handle = C.c_void_p()
        check(self.lib.lumx_context_create(device, C.byref(self.handle)))
        self.device = device

    def sync(self):
        check(self.lib.lumx_context_sync(self.handle))

    def close(self):
        if self.handle:
            check(self.lib.lumx_context_destroy(self.handle))
            self.handle = C.c_void_p()


class Model:
    def __init__(self, ctx, voxels):
        """voxels: uint8 array indexed [z][y][x] (x fastest in memory)."""
        v = np.ascontiguousarray(voxels, dtype=np.uint8)
        self.lib = ctx.lib
        self.size = (v.shape[2], v.shape[1], v.shape[0])
        self.handle = C.c_void_p()
        check(self.lib.lumx_model_create(ctx.handle, _np_ptr(v), self.size[0], self.size[1], self.size[2],
                                         C.byref(self.handle)))

    def close(self):
        if self.handle:
            check(self.lib.lumx_model_destroy(self.handle))
            self.handle = C.c_void_p()


def placement(rot_wxyz, shift, size):
    """Host math of map_mesh for one model: returns a Placement."""
    lib = load()
    t = Transform((C.c_float * 4)(*rot_wxyz), (C.c_float * 3)(*shift))
    sz = (C.c_int * 3)(*size)
    p = Placement()
    check(lib.lumx_model_placement(C.byref(t), sz, C.byref(p)))
    return p


def camera_default(width, height):
    cam = Camera()
    check(load().lumx_camera_default(width, height, C.byref(cam)))
    return cam


def radiance_params(rays=64, time_seed=0, size=1, shift=0, alpha=0.1, max_dist=128.0, max_steps=0):
    return RadianceParams(rays, time_seed, size, shift, alpha, max_dist, max_steps, 0)


class World:
    def __init__(self, ctx, size, static_blocks, palette_slots=0, radiance_format=RADIANCE_F32):
        self.ctx = ctx
        self.lib = ctx.lib
        self.size = tuple(int(s) for s in size)
        d = WorldDesc((C.c_int * 3)(*self.size), static_blocks, palette_slots, radiance_format)
        self.handle = C.c_void_p()
        check(self.lib.lumx_world_create(ctx.handle, C.byref(d), C.byref(self.handle)))
        info = WorldDesc()
        check(self.lib.lumx_world_info(self.handle, C.byref(info)))
        self.static_blocks = info.static_blocks
        self.palette_slots = info.palette_slots
        self.radiance_format = radiance_format

    @property
    def n_blocks(self):
        return self.size[0] * self.size[1] * self.size[2]

    def close(self):
        if self.handle:
            check(self.lib.lumx_world_destroy(self.handle))
            self.handle = C.c_void_p()

    # -- state --
    def set_blocks(self, blocks):
        b = np.ascontiguousarray(blocks, dtype=np.int16)
        assert b.size == self.n_blocks
        check(self.lib.lumx_world_set_blocks(self.handle, _np_ptr(b)))

    def get_blocks(self, blockified=True):
        out = np.zeros((self.size[2], self.size[1], self.size[0]), np.int16)
        check(self.lib.lumx_world_get_blocks(self.handle, _np_ptr(out), 1 if blockified else 0))
        return out

    def set_palette(self, voxels, first=0, layout=PALETTE_XYZ_XFAST):
        v = np.ascontiguousarray(voxels, dtype=np.uint8)
        count = v.size // 4096
        check(self.lib.lumx_world_set_palette(self.handle, first, count, _np_ptr(v), layout))

    def get_palette(self, first, count):
        out = np.zeros((count, 16, 16, 16), np.uint8)
        check(self.lib.lumx_world_get_palette(self.handle, first, count, _np_ptr(out)))
        return out

    def set_materials(self, mats):
        m = np.ascontiguousarray(mats, dtype=np.float32)
        assert m.size == 256 * 6
        check(self.lib.lumx_world_set_materials(self.handle, _np_ptr(m)))

    def write_radiance(self, rgba):
        r = np.ascontiguousarray(rgba, dtype=np.float32)
        assert r.size == self.n_blocks * 4
        check(self.lib.lumx_world_write_radiance(self.handle, _np_ptr(r)))

    def read_radiance(self):
        out = np.zeros((self.size[2], self.size[1], self.size[0], 4), np.float32)
        check(self.lib.lumx_world_read_radiance(self.handle, _np_ptr(out)))
        return out

    def upload_blocks(self):
        check(self.lib.lumx_world_upload_blocks(self.handle, _stream()))

    def palette_counter(self):
        n = C.c_int()
        check(self.lib.lumx_world_palette_counter(self.handle, C.byref(n)))
        return n.value

    def last_pass_ms(self, kind):
        ms = C.c_float()
        check(self.lib.lumx_world_last_pass_ms(self.handle, kind, C.byref(ms)))
        return ms.value

    def set_tuning(self, **kw):
        t = Tuning()
        check(self.lib.lumx_world_get_tuning(self.handle, C.byref(t)))
        for k, v in kw.items():
            setattr(t, k, int(v))
        check(self.lib.lumx_world_set_tuning(self.handle, C.byref(t)))

    def get_tuning(self):
        t = Tuning()
        check(self.lib.lumx_world_get_tuning(self.handle, C.byref(t)))
        return {k: getattr(t, k) for k in TUNING_FIELDS}

    def autotune(self, params, reps=5):
        t = Tuning()
        ms = C.c_float()
        check(self.lib.lumx_autotune_radiance(self.handle, C.byref(params), reps, C.byref(t), C.byref(ms)))
        return {k: getattr(t, k) for k in TUNING_FIELDS}, ms.value

    def evict(self):
        check(self.lib.lumx_world_evict(self.handle))

    def restore(self):
        check(self.lib.lumx_world_restore(self.handle))

    def resident(self):
        r = C.c_int()
        b = C.c_size_t()
        check(self.lib.lumx_world_resident(self.handle, C.byref(r), C.byref(b)))
        return bool(r.value), b.value

    # -- passes --
    def map_models(self, models, rots, shifts):
        n = len(models)
        hs = (C.c_void_p * max(n, 1))(*[m.handle.value for m in models])
        ts = (Transform * max(n, 1))(*[Transform((C.c_float * 4)(*r), (C.c_float * 3)(*s))
                                       for r, s in zip(rots, shifts)])
        check(self.lib.lumx_map_models(self.handle, n, hs, ts, _stream()))

    def map_models_placed(self, models, placements):
        n = len(models)
        hs = (C.c_void_p * max(n, 1))(*[m.handle.value for m in models])
        ps = (Placement * max(n, 1))(*placements)
        check(self.lib.lumx_map_models_placed(self.handle, n, hs, ps, _stream()))

    def map_models_device(self, models, placements):
        """Map with blockify and exec_copies on the device."""
        n = len(models)
        hs = (C.c_void_p * max(n, 1))(*[m.handle.value for m in models])
        ps = (Placement * max(n, 1))(*placements)
        check(self.lib.lumx_map_models_device(self.handle, n, hs, ps, _stream()))

    def map_models_device_noalloc(self, models, placements):
        """map_models_device with the library's fixed scratch: allocates
        nothing (at most 64 models)."""
        n = len(models)
        hs = (C.c_void_p * max(n, 1))(*[m.handle.value for m in models])
        ps = (Placement * max(n, 1))(*placements)
        check(self.lib.lumx_map_models_device_noalloc(self.handle, n, hs, ps, _stream()))

    def radiance_update(self, requests=None, params=None, traced=False, **kw):
        """requests: torch int32 (n, 4) on the device, or None for every probe."""
        import torch
        p = params if params is not None else radiance_params(**kw)
        n = 0 if requests is None else int(requests.shape[0])
        if requests is not None and n == 0:
            # An empty list is not "every probe": torch reports a null data
            # pointer for empty tensors, so pass a real one with n = 0.
            requests = torch.zeros((1, 4), dtype=torch.int32, device=requests.device)
        if not traced:
            check(self.lib.lumx_radiance_update(self.handle, _ptr(requests), n, C.byref(p), _stream()))
            return None
        n_total = n if requests is not None else self.n_blocks
        n_rays = (n_total // p.size) * p.rays
        keys = torch.zeros(max(n_rays, 1), dtype=torch.int32, device="cuda")
        steps = torch.zeros(max(2 * n_rays, 2), dtype=torch.int32, device="cuda")
        check(self.lib.lumx_radiance_update_traced(self.handle, _ptr(requests), n, C.byref(p), _ptr(keys),
                                                   _ptr(steps), _stream()))
        return keys[:n_rays], steps[:2 * n_rays].view(-1, 2)

    def select_probes(self, special=None):
        sp = np.zeros((0, 4), np.int32) if special is None else np.ascontiguousarray(special, np.int32)
        out = np.zeros((self.n_blocks + len(sp), 4), np.int32)
        n = C.c_int()
        check(self.lib.lumx_select_probes(self.handle, _np_ptr(sp), len(sp), _np_ptr(out), C.byref(n)))
        return out[: n.value]

    def select_probes_device(self, special=None):
        """Device selection: returns (requests int32 (n_blocks, 4), count int32
        (1,)), both torch tensors on the device; rows past count[0] are
        unspecified."""
        import torch
        n_sp = 0 if special is None else int(special.shape[0])
        out = torch.empty((self.n_blocks, 4), dtype=torch.int32, device="cuda")
        count = torch.empty(1, dtype=torch.int32, device="cuda")
        check(self.lib.lumx_select_probes_device(self.handle, _ptr(special), n_sp, _ptr(out), _ptr(count),
                                                 _stream()))
        return out, count

    def radiance_update_counted(self, requests, count, params=None, **kw):
        """Radiance update over requests[:count[0]], the count read on the device."""
        p = params if params is not None else radiance_params(**kw)
        check(self.lib.lumx_radiance_update_counted(self.handle, _ptr(requests), _ptr(count),
                                                    int(requests.shape[0]), C.byref(p), _stream()))

    def shift_radiance(self, shift):
        check(self.lib.lumx_world_shift_radiance(self.handle, int(shift[0]), int(shift[1]), int(shift[2]),
                                                 _stream()))

    def glossy(self, matnorm, depth, cam, max_dist=128.0, max_steps=0, out=None, traced=False):
        """matnorm: torch uint8 (h, w, 4); depth: torch float32 (h, w); returns float32 (h, w, 4)."""
        import torch
        h, w = int(depth.shape[0]), int(depth.shape[1])
        if out is None:
            out = torch.zeros((h, w, 4), dtype=torch.float32, device="cuda")
        keys = torch.zeros((h, w), dtype=torch.int32, device="cuda") if traced else None
        check(self.lib.lumx_glossy(self.handle, _ptr(matnorm), _ptr(depth), w, h, C.byref(cam),
                                   float(max_dist), int(max_steps), _ptr(out), _ptr(keys), _stream()))
        return (out, keys) if traced else out


    def diffuse(self, matnorm, depth, cam, lightmap=None, lightmap_proj=None, out=None):
        """diffuse.frag: matnorm torch uint8 (h, w, 4), depth float32 (h, w);
        lightmap: torch int16/uint16 (lh, lw) D16 values or None; lightmap_proj:
        16 floats, column-major.  Returns float32 (h, w, 4)."""
        import torch
        h, w = int(depth.shape[0]), int(depth.shape[1])
        if out is None:
            out = torch.zeros((h, w, 4), dtype=torch.float32, device="cuda")
        lm = None
        if lightmap is not None:
            proj = np.ascontiguousarray(lightmap_proj, np.float32).reshape(16)
            lm = _lib.Lightmap(lightmap.data_ptr(), int(lightmap.shape[1]), int(lightmap.shape[0]),
                               (C.c_float * 16)(*proj))
        check(self.lib.lumx_diffuse(self.handle, _ptr(matnorm), _ptr(depth), w, h, C.byref(cam),
                                    None if lm is None else C.byref(lm), _ptr(out), _stream()))
        return out


class Batch:
    """Worlds updated together in one launch per pass (lumx_batch_*)."""

    def __init__(self, ctx, worlds):
        self.lib = ctx.lib
        self.worlds = list(worlds)
        hs = (C.c_void_p * len(self.worlds))(*[w.handle.value for w in self.worlds])
        self.handle = C.c_void_p()
        check(self.lib.lumx_batch_create(ctx.handle, len(self.worlds), hs, C.byref(self.handle)))

    def close(self):
        if self.handle:
            check(self.lib.lumx_batch_destroy(self.handle))
            self.handle = C.c_void_p()

    def radiance(self, params, requests=None):
        """params: one RadianceParams per world; requests: None or a list of
        torch int32 (n, 4) device tensors (None entries = every probe)."""
        n = len(self.worlds)
        ps = (RadianceParams * n)(*params)
        if requests is None:
            check(self.lib.lumx_batch_radiance(self.handle, None, None, ps, _stream()))
            return
        rq = (C.c_void_p * n)(*[0 if r is None else r.data_ptr() for r in requests])
        nr = (C.c_int * n)(*[0 if r is None else int(r.shape[0]) for r in requests])
        check(self.lib.lumx_batch_radiance(self.handle, rq, nr, ps, _stream()))

    def autotune(self, params, requests=None, reps=3):
        """Choose the batched radiance launch shape on the launch radiance()
        would make with these arguments; returns (block, rays_per_lane, ms)."""
        n = len(self.worlds)
        ps = (RadianceParams * n)(*params)
        blk = C.c_int()
        rpl = C.c_int()
        ms = C.c_float()
        rq = nr = None
        if requests is not None:
            rq = (C.c_void_p * n)(*[0 if r is None else r.data_ptr() for r in requests])
            nr = (C.c_int * n)(*[0 if r is None else int(r.shape[0]) for r in requests])
        check(self.lib.lumx_batch_autotune(self.handle, rq, nr, ps, reps, C.byref(blk), C.byref(rpl),
                                           C.byref(ms)))
        return blk.value, rpl.value, ms.value

    def set_shape(self, block=0, rays_per_lane=0):
        check(self.lib.lumx_batch_set_shape(self.handle, int(block), int(rays_per_lane)))

    def map_models(self, models, placements):
        """models[i], placements[i]: world i's models and placements (lists;
        empty = the world's origin).  Equal to World.map_models_placed on
        each world in order."""
        n = len(self.worlds)
        keep = []
        hs = (C.c_void_p * n)()
        ps = (C.c_void_p * n)()
        cnt = (C.c_int * n)()
        for i in range(n):
            k = len(models[i])
            cnt[i] = k
            if k:
                h = (C.c_void_p * k)(*[m.handle.value for m in models[i]])
                p = (Placement * k)(*placements[i])
                keep += [h, p]
                hs[i] = C.cast(h, C.c_void_p)
                ps[i] = C.cast(p, C.c_void_p)
        check(self.lib.lumx_batch_map_models(self.handle, cnt, hs, ps, _stream()))

    def glossy(self, matnorms, depths, cams, outs, max_dist=128.0, max_steps=0):
        n = len(self.worlds)
        mn = (C.c_void_p * n)(*[m.data_ptr() for m in matnorms])
        dp = (C.c_void_p * n)(*[d.data_ptr() for d in depths])
        ws = (C.c_int * n)(*[int(d.shape[1]) for d in depths])
        hs = (C.c_int * n)(*[int(d.shape[0]) for d in depths])
        cs = (Camera * n)(*cams)
        os_ = (C.c_void_p * n)(*[o.data_ptr() for o in outs])
        check(self.lib.lumx_batch_glossy(self.handle, mn, dp, ws, hs, cs, float(max_dist), int(max_steps), os_,
                                         _stream()))

    def set_frame(self, frame):
        check(self.lib.lumx_batch_set_frame(self.handle, int(frame), _stream()))

    def advance_frame(self):
        check(self.lib.lumx_batch_advance_frame(self.handle, _stream()))


class Graph:
    """hipGraph of the lumx passes issued on torch's current (non-null) stream
    between begin() and end()."""

    def __init__(self):
        self.lib = load()
        self.handle = C.c_void_p()

    def begin(self):
        check(self.lib.lumx_graph_begin(_stream()))

    def end(self):
        check(self.lib.lumx_graph_end(_stream(), C.byref(self.handle)))

    def launch(self):
        check(self.lib.lumx_graph_launch(self.handle, _stream()))

    def close(self):
        if self.handle:
            check(self.lib.lumx_graph_destroy(self.handle))
            self.handle = C.c_void_p()


def radiance_request(world_id, params, requests=None):
    """requests must stay alive until the request completes."""
    if requests is not None and int(requests.shape[0]) == 0:
        raise ValueError("an empty request list updates nothing; pass None for every probe")
    r = Request()
    r.kind = REQ_RADIANCE
    r.world = world_id
    r.requests_dev = 0 if requests is None else requests.data_ptr()
    r.n_requests = 0 if requests is None else int(requests.shape[0])
    r.params = params
    return r


def glossy_request(world_id, matnorm, depth, cam, out, max_dist=128.0, max_steps=0):
    r = Request()
    r.kind = REQ_GLOSSY
    r.world = world_id
    r.matnorm_dev = matnorm.data_ptr()
    r.depth_dev = depth.data_ptr()
    r.height, r.width = int(depth.shape[0]), int(depth.shape[1])
    r.cam = cam
    r.max_dist = float(max_dist)
    r.max_steps = int(max_steps)
    r.out_dev = out.data_ptr()
    return r


class Server:
    """Continuous-batching request engine (lumx_server_*)."""

    def __init__(self, ctx, worlds, hbm_budget=0, max_batch=32, streams=2, block=0, rays_per_lane=0, direct=False):
        """direct=True: the plain engine, one single-world call per request
        queued on one stream with no wait between calls."""
        self.lib = ctx.lib
        d = ServerDesc(int(hbm_budget), int(max_batch), int(streams), int(block), int(rays_per_lane),
                       1 if direct else 0)
        self.handle = C.c_void_p()
        check(self.lib.lumx_server_create(ctx.handle, C.byref(d), C.byref(self.handle)))
        self.ids = []
        try:
            for w in worlds:
                i = C.c_int()
                check(self.lib.lumx_server_add_world(self.handle, w.handle, C.byref(i)))
                self.ids.append(i.value)
        except Exception:
            self.close()
            raise

    def submit(self, req):
        t = C.c_longlong()
        check(self.lib.lumx_server_submit(self.handle, C.byref(req), C.byref(t)))
        return t.value

    def wait(self, ticket):
        lat = C.c_double()
        check(self.lib.lumx_server_wait(self.handle, int(ticket), C.byref(lat)))
        return lat.value

    def drain(self):
        check(self.lib.lumx_server_drain(self.handle))

    def stats(self):
        st = ServerStats()
        check(self.lib.lumx_server_get_stats(self.handle, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def run_trace(self, reqs, arrivals):
        n = len(reqs)
        if len(arrivals) != n:
            raise ValueError("one arrival time per request")
        arr = (Request * max(n, 1))(*reqs)
        at = (C.c_double * max(n, 1))(*arrivals)
        lat = (C.c_double * max(n, 1))()
        wall = C.c_double()
        check(self.lib.lumx_server_run_trace(self.handle, n, arr, at, lat, C.byref(wall)))
        return np.array(lat[:n]), wall.value

    def close(self):
        if self.handle:
            check(self.lib.lumx_server_destroy(self.handle))
            self.handle = C.c_void_p()
