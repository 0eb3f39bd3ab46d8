"""Committed fixtures (tests/golden/lum_small.npz, written by
tests/golden/make_golden.py): the oracle reproduces them exactly (CPU), and
the HIP kernels match them (GPU)."""
import os

import numpy as np
import pytest

import tests._oracle as orc

G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lum_small.npz"))
SIZE = (8, 8, 4)
STATIC = 8


def full_palette(p, slots):
    out = np.zeros((slots, 16, 16, 16), np.uint8)
    out[: p.shape[0]] = p
    return out


def test_oracle_reproduces_golden():
    slots = SIZE[0] * SIZE[1] * SIZE[2] + STATIC
    wd = orc.WorldData(G["blocks"], full_palette(G["palette"], slots), G["mats"])
    r1 = orc.radiance_update(wd, G["field"], rays=64, time_seed=3, alpha=0.1, max_steps=64)
    r2 = orc.radiance_update(wd, r1, rays=128, time_seed=4, alpha=0.1)
    assert np.array_equal(r1, G["radiance_f1"]) and np.array_equal(r2, G["radiance_f2"])
    cur = G["blocks"].copy()
    copy_src = np.zeros(slots, np.int32)
    c = STATIC
    for i in range(3):
        a = G[f"aabb{i}"]
        c = orc.blockify(SIZE, cur, STATIC, c, a[:3], a[3:], copy_src)
    assert c == int(G["counter"]) and np.array_equal(cur, G["blockified"])
    mapped = full_palette(G["palette"], slots)
    orc.exec_copies(mapped, STATIC, c, copy_src)
    for i in range(3):
        orc.map_model(SIZE, cur, mapped, STATIC, G[f"model{i}"], G[f"minv{i}"], G[f"shift{i}"][:3], G[f"groups{i}"])
    assert np.array_equal(mapped[:c], G["palette_mapped"])
    gl, gk = orc.glossy(orc.WorldData(cur, mapped, G["mats"]), r2, G["matnorm"], G["depth"], G["ubo"], 128.0, 128)
    assert np.array_equal(gl, G["glossy"]) and np.array_equal(gk, G["glossy_keys"])


@pytest.mark.gpu
def test_kernels_match_golden(gpu_ctx):
    import ctypes as C
    import torch
    from lum_amd.api import World, Model
    from lum_amd._lib import Placement, Camera
    w = World(gpu_ctx, SIZE, static_blocks=STATIC)
    w.set_blocks(G["blocks"])
    w.set_palette(G["palette"])
    w.set_materials(G["mats"])
    w.write_radiance(G["field"])
    ms = []
    try:
        ms = [Model(gpu_ctx, G[f"model{i}"]) for i in range(3)]
        pls = []
        for i in range(3):
            p = Placement()
            p.minv = (C.c_float * 16)(*G[f"minv{i}"].tolist())
            p.shift = (C.c_int * 4)(*G[f"shift{i}"].tolist())
            p.groups = (C.c_int * 3)(*G[f"groups{i}"].tolist())
            a = G[f"aabb{i}"].tolist()
            p.aabb_min = (C.c_float * 3)(*a[:3])
            p.aabb_max = (C.c_float * 3)(*a[3:])
            pls.append(p)
        # radiance on the unmapped world first (as the fixture), then map
        w.upload_blocks()
        from lum_amd.api import radiance_params
        w.radiance_update(None, radiance_params(rays=64, time_seed=3, alpha=0.1, max_steps=64))
        w.radiance_update(None, radiance_params(rays=128, time_seed=4, alpha=0.1))
        torch.cuda.synchronize()
        got = w.read_radiance()
        assert np.array_equal(got, G["radiance_f2"])
        w.map_models_placed(ms, pls)
        torch.cuda.synchronize()
        c = w.palette_counter()
        assert c == int(G["counter"])
        assert np.array_equal(w.get_blocks(True), G["blockified"])
        assert np.array_equal(w.get_palette(0, c), G["palette_mapped"])
        u = G["ubo"].tolist()
        cam = Camera((C.c_float * 3)(*u[0:3]), (C.c_float * 3)(*u[3:6]), (C.c_float * 3)(*u[6:9]),
                     (C.c_float * 3)(*u[9:12]), (C.c_float * 3)(*u[12:15]), (C.c_float * 2)(*u[15:17]))
        out, keys = w.glossy(torch.from_numpy(G["matnorm"]).cuda(), torch.from_numpy(G["depth"]).cuda(), cam,
                             128.0, 128, traced=True)
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), G["glossy"])
        assert np.array_equal(keys.cpu().numpy(), G["glossy_keys"])
    finally:
        for m in ms:
            m.close()
        w.close()
