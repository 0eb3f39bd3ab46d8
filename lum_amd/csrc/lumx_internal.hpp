// lumx_internal.hpp - kernel argument blocks and launchers shared between the
// HIP kernels and the host runtime of liblumx.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "march.hpp"

namespace lumx {

struct RadianceArgs {
    WorldView world;
    const void* field_in;   // texels of the previous field
    void* field_out;        // texels of the new field
    const int4* requests;   // probe list, or nullptr = every probe x-fastest
    const float4* fib;      // per-lane Fibonacci directions (xyz)
    int rays;
    int n_probes;           // dispatched probes (= n_requests / size)
    int time_seed;
    int size;
    int shift;
    float alpha;
    float max_dist;
    int max_steps;
    int32_t* ray_keys;      // traced launches only
    int32_t* ray_steps;
};

struct MapArgs {
    const int16_t* blocks;  // blockified world
    uint8_t* palette;
    const uint8_t* model;   // x-fastest u8 voxels
    int wx, wy, wz;
    int palette_slots;
    int static_count;
    int mx, my, mz;
    int nx, ny, nz;         // threads per axis (4 * groups)
    int shift_x, shift_y, shift_z;
    float m[12];            // rows 0..2 of the inverse transform: m[r*4 + c]
};

struct CopyArgs {
    uint8_t* palette;
    const int32_t* copy_src;  // per slot: static source block (0 = zero fill)
    int first_slot;
    int n_slots;
};

struct GlossyArgs {
    WorldView world;
    const void* field;
    const uchar4* matnorm;
    const float* depth;
    float4* out;
    int32_t* keys;
    int width, height;
    float campos[3], camdir[3], horiz[3], verti[3], light[3];
    float frame_w, frame_h;
    float max_dist;
    int max_steps;
};

hipError_t launch_radiance(const RadianceArgs& a, int store, bool traced, hipStream_t stream);
hipError_t launch_map(const MapArgs& a, hipStream_t stream);
hipError_t launch_exec_copies(const CopyArgs& a, hipStream_t stream);
hipError_t launch_glossy(const GlossyArgs& a, int store, bool traced, hipStream_t stream);

}  // namespace lumx
