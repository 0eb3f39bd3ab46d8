// march_host.cpp - the kernels' march and sampling helpers compiled for the
// host, so tests can compare the product's own source with the CPU oracle
// ray by ray without a GPU.  Built by tests/test_kernel_source_cpu.py with
// `hipcc -x hip --cuda-host-only`; lum_amd/csrc/march.hpp is the code under
// test (its helpers are __host__ __device__).
#include "../lum_amd/csrc/march.hpp"

using namespace lumx;

extern "C" {

// One cast per ray: info = {hit, voxel, coarse, refine, key}, fn = {fraction, normal.xyz}.
// occ != 0 marches the occupancy bits (cast_ray_occ) instead of the u8 palette.
void mh_cast_rays(const int16_t* blocks, const uint8_t* palette, const float* mats, int wx, int wy, int wz,
                  int palette_slots, int block0_empty, const uint32_t* empty_bits, const uint32_t* pal_bits,
                  int occ, int n, const float* o, const float* d, float max_dist, int max_steps, int* info,
                  float* fn) {
    WorldView w;
    w.blocks = blocks;
    w.palette = palette;
    w.mats = mats;
    w.wx = wx;
    w.wy = wy;
    w.wz = wz;
    w.palette_slots = palette_slots;
    w.block0_empty = block0_empty;
    OccView ov;
    ov.ids = blocks;
    ov.palette = palette;
    ov.eb = empty_bits;
    ov.pb = pal_bits;
    ov.wx = wx;
    ov.wy = wy;
    ov.wz = wz;
    ov.palette_slots = palette_slots;
    ov.block0_empty = block0_empty;
    for (int i = 0; i < n; ++i) {
        const f3 oo = {o[3 * i], o[3 * i + 1], o[3 * i + 2]};
        const f3 dd = {d[3 * i], d[3 * i + 1], d[3 * i + 2]};
        RayHit h = occ ? cast_ray_occ(ov, oo, dd, max_dist, max_steps) : cast_ray(w, oo, dd, max_dist, max_steps);
        info[5 * i + 0] = h.hit;
        info[5 * i + 1] = h.voxel;
        info[5 * i + 2] = h.coarse;
        info[5 * i + 3] = h.refine;
        info[5 * i + 4] = h.hit ? h.key : 0;
        fn[4 * i + 0] = h.fraction;
        fn[4 * i + 1] = h.normal.x;
        fn[4 * i + 2] = h.normal.y;
        fn[4 * i + 3] = h.normal.z;
    }
}

// sample_radiance over an FP32 field (float4 texels).
void mh_sample_radiance(const float* field, int wx, int wy, int wz, int n, const float* x, const float* nrm,
                        float* out) {
    for (int i = 0; i < n; ++i) {
        f3 r = sample_radiance<STORE_F32>(field, wx, wy, wz, f3{x[3 * i], x[3 * i + 1], x[3 * i + 2]},
                                          f3{nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]});
        out[3 * i] = r.x;
        out[3 * i + 1] = r.y;
        out[3 * i + 2] = r.z;
    }
}

// n[i] / d through the kernels' multiply-high division.
void mh_fast_div(const uint32_t* n, int count, uint32_t d, uint32_t* out) {
    const Divisor v = make_divisor(d);
    for (int i = 0; i < count; ++i) out[i] = fast_div(n[i], v);
}

}  // extern "C"
