"""The C99 boundary (include/clum.h) driven as a C host would drive it.

The sequence follows the reference's examples/cdemo.c: create, init, load the
world, upload block and material palettes, load a mesh, then per frame
start / draw / prepare / submit.  Each frame's radiance field and glossy
output must equal, bit for bit, the same frame driven through the pass-level
API (include/lumx.h): map, shift_radiance with the stored shift, probe
selection with the special requests, radiance update (64 rays, alpha 0.01,
size 1 while the measured pass stays under 1.8 ms), glossy.
"""
import ctypes as C

import numpy as np
import pytest

from lum_amd import synth
from lum_amd.api import World, Model, radiance_params, camera_default
from lum_amd.clum import lib as clum_lib, LumSettings, LumMeshTransform

pytestmark = pytest.mark.gpu


def clum():
    return clum_lib()


def np_ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def test_lum_frames_equal_pass_level_chain(gpu_ctx):
    import torch
    lib = clum()
    size, static = (32, 32, 8), 16
    rng = np.random.default_rng(77)
    blocks = synth.make_world(size, rng, n_palette=static)
    pal = synth.make_palette(rng, static)
    mats = synth.make_materials(rng)
    vox = (rng.random((20, 24, 28)) < 0.3).astype(np.uint8) * rng.integers(1, 256, (20, 24, 28)).astype(np.uint8)
    rot = [0.9238795, 0.0, 0.0, 0.3826834]            # 45 degrees about z
    shift = [250.0, 240.0, 40.0]                       # in view of Lum's camera
    W, H = 96, 64
    mn, dp = synth.make_gbuffer(W, H, mats, rng, depth_hi=60.0)
    mn_d, dp_d = torch.from_numpy(mn).cuda(), torch.from_numpy(dp).cuda()
    out_lum = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    frames = [((0, 0, 0), []), ((1, 0, 0), [(3, 4, 7)]), ((0, -1, 0), [(3, 4, 7), (31, 0, 0)]),
              ((0, 0, 0), [])]

    s = LumSettings((C.c_int * 3)(*size), static, 8128, 1024, 2, b"\0", b"\0", b"\0", b"\0")
    lum = lib.lum_create_instance(C.byref(s))
    assert lum
    ref = None
    mesh_ref = None
    try:
        lib.lum_init(lum, C.byref(s))
        b16 = np.ascontiguousarray(blocks, np.int16)
        lib.lum_load_world(lum, np_ptr(b16))
        for b in range(1, static):
            bx = np.ascontiguousarray(pal[b].transpose(2, 1, 0))   # BlockVoxels[x][y][z]
            lib.lum_load_block(lum, b, np_ptr(bx))
        lib.lum_upload_block_palette_to_gpu(lum)
        m32 = np.ascontiguousarray(mats, np.float32)
        lib.lumx_lum_set_materials(lum, np_ptr(m32))
        lib.lum_upload_material_palette_to_gpu(lum)
        v = np.ascontiguousarray(vox)
        mesh = lib.lum_load_mesh_data(lum, np_ptr(v), v.shape[2], v.shape[1], v.shape[0])
        assert mesh
        lib.lumx_lum_set_gbuffer(lum, C.c_void_p(mn_d.data_ptr()), C.c_void_p(dp_d.data_ptr()), W, H,
                                 C.c_void_p(out_lum.data_ptr()))
        lw = lib.lumx_lum_world(lum)

        ref = World(gpu_ctx, size, static_blocks=static)
        ref.set_blocks(blocks)
        pal0 = pal.copy()
        pal0[0] = 0                       # the C API never loads block 0 (air)
        ref.set_palette(pal0)
        ref.set_materials(mats)
        mesh_ref = Model(gpu_ctx, vox)
        t = LumMeshTransform((C.c_float * 4)(*rot), (C.c_float * 3)(*shift))
        for f, (rshift, specials) in enumerate(frames):
            lib.lumx_lum_set_radiance_shift(lum, *rshift)
            for sp in specials:
                lib.lumx_lum_request_radiance_update(lum, *sp)
            lib.lum_start_frame(lum)
            lib.lum_draw_model(lum, mesh, C.byref(t))
            lib.lum_prepare_frame(lum)
            lib.lum_submit_frame(lum)
            counter = C.c_int()
            assert lib.lumx_world_palette_counter(lw, C.byref(counter)) == 0
            # Lum's camera is set at the first submit: frame 0 culls the model.
            drawn = counter.value > static
            assert drawn == (f > 0)
            assert lib.lum_get_timestamp_difference(lum, 1, 2) < 1.8   # controller keeps size 1

            ref.map_models([mesh_ref] if drawn else [], [rot] if drawn else [], [shift] if drawn else [])
            ref.shift_radiance(rshift)
            sp = np.array([[x, y, z, 0] for x, y, z in specials], np.int32).reshape(-1, 4)
            req = ref.select_probes(sp)
            ref.radiance_update(torch.from_numpy(req).cuda(),
                                radiance_params(rays=64, time_seed=f, size=1, shift=0, alpha=0.01, max_dist=128.0))
            out_ref = ref.glossy(mn_d, dp_d, camera_default(W, H), 128.0, 0)
            torch.cuda.synchronize()
            got = np.zeros((size[2], size[1], size[0], 4), np.float32)
            assert lib.lumx_world_read_radiance(lw, np_ptr(got)) == 0
            assert np.array_equal(got, ref.read_radiance()), f"frame {f}"
            assert torch.equal(out_lum, out_ref), f"frame {f}"
        assert np.abs(got[..., :3]).max() > 0
    finally:
        lib.lum_cleanup(lum)
        if mesh_ref is not None:
            mesh_ref.close()
        if ref is not None:
            ref.close()


def test_lum_frames_in_flight_equal_pass_level_chain(gpu_ctx):
    """Six frames submitted back to back with no host read in between (two in
    flight, specials staged per frame slot), then compared once: the field and
    the last glossy frame equal the pass-level chain of the same six frames."""
    import torch
    lib = clum()
    size, static = (32, 32, 8), 16
    rng = np.random.default_rng(78)
    blocks = synth.make_world(size, rng, n_palette=static)
    pal = synth.make_palette(rng, static)
    mats = synth.make_materials(rng)
    vox = (rng.random((18, 20, 22)) < 0.3).astype(np.uint8) * rng.integers(1, 256, (18, 20, 22)).astype(np.uint8)
    rot = [0.9238795, 0.0, 0.0, 0.3826834]
    # one model per frame, moved every frame, in view of Lum's camera
    shifts = [[250.0 + 3 * f, 240.0 - 2 * f, 40.0 + (f % 2)] for f in range(6)]
    W, H = 64, 48
    mn, dp = synth.make_gbuffer(W, H, mats, rng, depth_hi=60.0)
    mn_d, dp_d = torch.from_numpy(mn).cuda(), torch.from_numpy(dp).cuda()
    out_lum = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    frames = [((0, 0, 0), [(2, 3, 4)]), ((0, 0, 0), [(5, 5, 5), (2, 3, 4), (2, 3, 4)]), ((1, 0, 0), []),
              ((0, 0, 0), [(31, 31, 7)] * 3), ((0, 1, 0), [(0, 0, 0)]), ((0, 0, 0), [])]
    s = LumSettings((C.c_int * 3)(*size), static, 8128, 1024, 2, b"\0", b"\0", b"\0", b"\0")
    lum = lib.lum_create_instance(C.byref(s))
    assert lum
    ref = mesh_ref = None
    try:
        lib.lum_init(lum, C.byref(s))
        b16 = np.ascontiguousarray(blocks, np.int16)
        lib.lum_load_world(lum, np_ptr(b16))
        for b in range(1, static):
            lib.lum_load_block(lum, b, np_ptr(np.ascontiguousarray(pal[b].transpose(2, 1, 0))))
        lib.lum_upload_block_palette_to_gpu(lum)
        m32 = np.ascontiguousarray(mats, np.float32)
        lib.lumx_lum_set_materials(lum, np_ptr(m32))
        lib.lum_upload_material_palette_to_gpu(lum)
        v = np.ascontiguousarray(vox)
        mesh = lib.lum_load_mesh_data(lum, np_ptr(v), v.shape[2], v.shape[1], v.shape[0])
        lib.lumx_lum_set_gbuffer(lum, C.c_void_p(mn_d.data_ptr()), C.c_void_p(dp_d.data_ptr()), W, H,
                                 C.c_void_p(out_lum.data_ptr()))
        ts = [LumMeshTransform((C.c_float * 4)(*rot), (C.c_float * 3)(*sh)) for sh in shifts]
        for f, (rshift, specials) in enumerate(frames):
            lib.lumx_lum_set_radiance_shift(lum, *rshift)
            for sp in specials:
                lib.lumx_lum_request_radiance_update(lum, *sp)
            lib.lum_start_frame(lum)
            lib.lum_draw_model(lum, mesh, C.byref(ts[f]))
            lib.lum_prepare_frame(lum)
            lib.lum_submit_frame(lum)
        lib.lum_wait_idle(lum)
        lw = lib.lumx_lum_world(lum)
        got = np.zeros((size[2], size[1], size[0], 4), np.float32)
        assert lib.lumx_world_read_radiance(lw, np_ptr(got)) == 0

        ref = World(gpu_ctx, size, static_blocks=static)
        ref.set_blocks(blocks)
        pal0 = pal.copy()
        pal0[0] = 0
        ref.set_palette(pal0)
        ref.set_materials(mats)
        mesh_ref = Model(gpu_ctx, vox)
        for f, (rshift, specials) in enumerate(frames):
            # Lum's camera is set at the first submit: frame 0 culls the model.
            n = 0 if f == 0 else 1
            ref.map_models([mesh_ref] * n, [rot] * n, [shifts[f]] * n)
            ref.shift_radiance(rshift)
            sp = np.array([[x, y, z, 0] for x, y, z in specials], np.int32).reshape(-1, 4)
            req = ref.select_probes(sp)
            ref.radiance_update(torch.from_numpy(req).cuda(),
                                radiance_params(rays=64, time_seed=f, size=1, shift=0, alpha=0.01, max_dist=128.0))
            out_ref = ref.glossy(mn_d, dp_d, camera_default(W, H), 128.0, 0)
        torch.cuda.synchronize()
        assert np.array_equal(got, ref.read_radiance())
        assert torch.equal(out_lum, out_ref)
        assert lib.lum_get_timestamp_difference(lum, 0, 3) > 0   # a finished frame's pass times
    finally:
        lib.lum_cleanup(lum)
        if mesh_ref is not None:
            mesh_ref.close()
        if ref is not None:
            ref.close()
