"""Mutation check of the oracle pins (CPU only).

Each mutant is the oracle source with one plausible restatement mistake
planted (a swapped swizzle, a dropped term, a flipped comparison, a wrong
order), compiled into a temporary directory and loaded by the pin suite
(tests/test_oracle_pins.py) in a child process through LUMX_ORACLE_SO.  Every
mutant must make at least one pin fail; the unmutated source must pass them
all.  The contraction-sensitivity measurement is left out of these runs: it
compares with the unmutated build and would catch every mutant without
pinning anything.
"""
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "lum_oracle.cpp")

# name -> (text in the oracle source, replacement)
MUTANTS = {
    "hash_swizzle": ("uint32_t nx = ((x >> 8) ^ y) * kHashMul;", "uint32_t nx = ((x >> 8) ^ z) * kHashMul;"),
    "dropped_emission": ("light = light + (prop * m.color + m.color * m.emit);", "light = light + (prop * m.color);"),
    "sample_radiance_no_x16": ("return sample_texel<T>(f, (x + n * (T)16) / (T)16);",
                               "return sample_texel<T>(f, (x + n) / (T)16);"),
    "blend_weights_swapped": ("out4[0] = unorm_nan(o.x * ((T)1 - a) + mean.x * a);",
                              "out4[0] = unorm_nan(o.x * a + mean.x * ((T)1 - a));"),
    "dda_tie_strict": ("st.y = (tMax.x >= tMax.y && tMax.y <= tMax.z) ? (T)1 : (T)0;",
                       "st.y = (tMax.x > tMax.y && tMax.y <= tMax.z) ? (T)1 : (T)0;"),
    "map_no_half_voxel": ("float wvx = ((float)gx + 0.5f) + (float)shift[0];",
                          "float wvx = ((float)gx) + (float)shift[0];"),
    "blockify_no_margin": ("mn[a] = ((int)aabb_min[a] - 1) / 16;", "mn[a] = ((int)aabb_min[a]) / 16;"),
    # radiance.comp:372-381, the ray setup
    "ray_parity_swapped": ("if (lane % 2 != 0) dir = -dir;", "if (lane % 2 == 0) dir = -dir;"),
    "ray_offset_after_flip": ("V3<T> pos = center + dir * (T)6.0;",
                              "V3<T> pos = center + (lane % 2 != 0 ? -dir : dir) * (T)6.0;"),
    "ray_no_direction_jitter": ("V3<T> dir = normalize(fd + (T)0.1 * ((rr0 - half) * (T)2.0));",
                                "V3<T> dir = normalize(fd);"),
    "ray_no_origin_jitter": ("pos = pos + (rr1 - half) * (T)2.0;", "pos = pos + (rr1 - rr1);"),
    "ray_key_skips_second_draw": ("seed[0]++;", "(void)seed;"),
    # glossy.frag:312-326,341-345, the hit branch
    "glossy_hit_R_replaced": ("R = R * hm.color;", "R = hm.color;"),
    "glossy_hit_no_emission": ("L = L + R * (V3<float>{hm.emit, hm.emit, hm.emit} + d2);", "L = L + R * d2;"),
    # map.comp negative coordinates, repeated requests and specials
    "map_skips_negative": ("float wvx = ((float)gx + 0.5f) + (float)shift[0];",
                           "if (gx + shift[0] < 0 || gy + shift[1] < 0 || gz + shift[2] < 0) continue;\n"
                           "        float wvx = ((float)gx + 0.5f) + (float)shift[0];"),
    "map_first_thread_wins": ("for (int gx = 0; gx < nx; ++gx) {", "for (int gx = nx - 1; gx >= 0; --gx) {"),
    "duplicate_first_wins": ("for (int wg = 0; wg < n_wg; ++wg) {\n        if (!stored[wg]) continue;",
                             "for (int wg = n_wg - 1; wg >= 0; --wg) {\n        if (!stored[wg]) continue;"),
    "specials_deduplicated": ("out[4 * n + 3] = special[4 * i + 3];\n            ++n;",
                              "out[4 * n + 3] = special[4 * i + 3];\n            set[idx] = 1;\n            ++n;"),
}

CXXFLAGS = ["-O1", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math", "-std=c++17", "-fPIC", "-shared",
            "-pthread"]


def _build(src_text, out_dir, name):
    path = os.path.join(out_dir, name + ".cpp")
    with open(path, "w") as f:
        f.write(src_text)
    so = os.path.join(out_dir, name + ".so")
    subprocess.run(["g++"] + CXXFLAGS + [path, "-o", so], check=True, capture_output=True)
    return so


def _pins_pass(so):
    env = dict(os.environ, LUMX_ORACLE_SO=so)
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests", "test_oracle_pins.py"), "-x",
                        "-q", "-p", "no:cacheprovider", "-k", "not fma_contraction"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    return r.returncode == 0, r.stdout[-2000:]


def test_every_mutant_fails_a_pin():
    with open(SRC) as f:
        src = f.read()
    for name, (old, new) in MUTANTS.items():
        assert src.count(old) == 1, f"{name}: the mutated text must occur exactly once"
    with tempfile.TemporaryDirectory() as tmp:
        jobs = {"unmutated": src}
        jobs.update({name: src.replace(old, new) for name, (old, new) in MUTANTS.items()})
        workers = min(8, os.cpu_count() or 1)
        with ThreadPoolExecutor(workers) as ex:
            libs = dict(zip(jobs, ex.map(lambda kv: _build(kv[1], tmp, kv[0]), jobs.items())))
            results = dict(zip(libs, ex.map(_pins_pass, libs.values())))
    ok, log = results.pop("unmutated")
    assert ok, "the unmutated oracle must pass every pin:\n" + log
    survivors = [name for name, (passed, _) in results.items() if passed]
    assert not survivors, f"mutants no pin catches: {survivors}"
