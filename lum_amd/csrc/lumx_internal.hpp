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
    int log2_rays;          // log2(rays) when rays is a power of two, else -1
    int n_probes;           // dispatched probes (= n_requests / size)
    const int* n_requests_dev;  // device request count (nullptr = n_probes is exact)
    Divisor div_x, div_xy;  // wx and wx * wy, for probe coordinates of x-fastest indices
    int staged;             // 1: probe i's new texel goes to field_out[i] (a subset update,
                            // applied in place by the scatter kernel), 0: to its own texel
    int* claim;             // staged updates followed by a scatter: per texel, the highest
                            // request index naming it (set by the claim pass before the
                            // update, -1 between updates), so the last of duplicate
                            // requests is the one applied; nullptr = none
    int time_seed;
    int size;
    int shift;
    float alpha;
    float max_dist;
    int max_steps;
    int32_t* ray_keys;      // traced launches only
    int32_t* ray_steps;
    // tiled kernel only
    const uint32_t* empty_bits;  // per block: 1 = block id 0
    const uint32_t* pal_bits;    // per slot: 128 words of voxel-occupancy bits
    int eb_words;
    int pb_words;
    int probes_per_batch;        // P: probes per batch (power of two)
    int n_batches;
    unsigned int* batch_counter; // 2 words: next batch, finished workgroups (self-resetting)
    // Compacted subset update (lane kernels): probe j of the launch is the
    // j-th claimed texel in x-fastest order.  Its entry int2 {texel, request
    // index} sits at byte j * list_stride of field_out, the slot the probe's
    // own new texel later overwrites; the request index keys its rays.
    int compact;
    int list_stride;             // texel bytes
    int* tiles;                  // per tile of 2048 texels: count, then offset; [n_tiles] = entries
    int n_tiles;
};

// Launch shape of the tiled radiance kernel (chosen by the autotuner).
struct RadianceTune {
    int threads;          // workgroup size (multiple of 64, <= 1024)
    int probes_per_batch; // P
    int eb_lds;           // stage empty-block bits in LDS
    int pb_lds;           // stage palette occupancy bits in LDS
    int wg_per_cu;        // persistent workgroups per CU (0 = occupancy limit)
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
    int cls;                // sign class of the threads this launch maps (7: all coordinates >= 0)
    int tiles_x, tiles_y;   // workgroup tiles of 16 x 4 x kMapTileZ voxels: (nx + 15) / 16, ny / 4
    Divisor div_tiles_x, div_tiles_y;
};

// A map workgroup of 256 threads covers 16 x 4 x kMapTileZ voxels of a
// model's dispatch box, each thread kMapTileZ / 4 of them along z.
constexpr int kMapTileZ = 16;
inline long long map_tiles(const MapArgs& a) {
    return (long long)a.tiles_x * a.tiles_y * ((a.nz + kMapTileZ - 1) / kMapTileZ);
}

// Models of one map launch (kernel argument, at most kMapGroup models).
constexpr int kMapGroup = 16;
struct MapGroup {
    MapArgs m[kMapGroup];
    int begin[kMapGroup + 1];  // workgroup ranges of 256 threads, begin[n] = grid size
    int n;
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

struct DiffuseArgs {
    WorldView world;
    const void* field;
    const uchar4* matnorm;
    const float* depth;
    float4* out;
    int width, height;
    float campos[3], camdir[3], horiz[3], verti[3], light[3];
    float frame_w, frame_h;
    const uint16_t* lightmap;  // D16 shadow map (nullptr = no sun term)
    int lm_w, lm_h;
    float lm_proj[16];         // column-major mat4
};

// Device probe selection (probes.hip).
struct SelectArgs {
    const int16_t* blocks;      // current world, x-fastest
    int wx, wy, wz, n;          // n = wx * wy * wz
    int n_tiles;                // select_tiles(n)
    unsigned long long* bits;   // one ballot word per 64 probes
    int* tile_counts;           // per tile: count, then offset
    const int4* special;        // special requests (x, y, z, w)
    int n_special;
    int4* out;                  // n entries
    int* count;                 // entries written (at most n)
};

// Field shift (probes.hip): texel p + shift lands on p inside the destination box.
struct ShiftArgs {
    const void* in;
    void* out;
    int wx, wy, wz;
    long long n;
    int shift[3];
    int dst[3];
    int extent[3];
};

// Device blockify (blockify.hip).
struct BlockifyArgs {
    const int16_t* origin;      // origin world (device)
    int16_t* blocks;            // current world, a copy of origin on entry
    int32_t* copy_src;          // per slot: static source block
    uint8_t* palette;
    int* claim;                 // per block: first model covering it, idle value otherwise
    int claim_idle;
    int wx, wy, wz;
    int static_count, palette_slots;
    int n_models, n_elements, n_tiles;
    const int* first;           // per model: index of its first element (prefix of box volumes)
    const int4* boxes;          // per model: {first block of its box}, {box extent in blocks}
    unsigned long long* bits;   // one ballot word per 64 elements
    int* tile_counts;           // per tile: count, then offset
    int* state;                 // [0] counter of the last map, [1] copy bound, [2] overflow
};

size_t blockify_tiles(int n_elements);
hipError_t launch_blockify(const BlockifyArgs& a, int copy_slots, hipStream_t stream);

// No-allocation device map (blockify.hip): the model boxes, ballot words and
// tile counts live in one fixed device array per GPU instead of the world's
// scratch buffer, and the boxes arrive as a kernel argument instead of a
// staged copy.
constexpr int kNoAllocModels = 64;
constexpr size_t kNoAllocScratchBytes = size_t(4) << 20;
struct BlockifyDesc {
    int4 boxes[2 * kNoAllocModels];  // per model: first block, extent
    int first[kNoAllocModels];       // per model: index of its first element
    int n;
};
// Device address of the fixed array on `device`, resolved once per device
// (the first call makes `device` current).
hipError_t noalloc_scratch(int device, void** out);
// Writes d's boxes and prefix to `boxes` and `first` on the device.
hipError_t launch_blockify_desc(const BlockifyDesc& d, int4* boxes, int* first, hipStream_t stream);

// World attach (residency.hip): the parts of a world's device block that
// are not kept while it is evicted.
struct AttachArgs {
    uint2* field_other;         // the field that is not current
    long long n_field_words;    // its size in 8-byte words
    uint8_t* palette;
    int used, slots;            // palette slots in use (kept) and in total
    const int16_t* blocks;
    int n_blocks;
    uint32_t* empty_bits;       // per block: 1 = id 0
    uint32_t* pal_bits;         // per slot: 128 words of voxel occupancy
    int* sel_claim;
    int claim_idle;
    int* upd_claim;
    unsigned int* batch_counter;
    int* blockify_state;
    int counter;                // blockify_state[0]
};
hipError_t launch_world_attach(const AttachArgs& a, hipStream_t stream);
// Worlds a[0..n) set up by launches of up to kAttachBatch worlds each.
constexpr int kAttachBatch = 16;
hipError_t launch_world_attach_batch(const AttachArgs* a, int n, hipStream_t stream);

size_t select_tiles(int n_probes);
size_t select_bits_bytes(int n_probes);
int select_claim_idle();
hipError_t launch_select(const SelectArgs& a, hipStream_t stream);
hipError_t launch_shift(const ShiftArgs& a, int texel_bytes, hipStream_t stream);

// rays_per_lane 2 or 4 (plain march, not traced) when rays is a power of
// two >= it, else 1; rays_per_lane_for gives the value a launch uses.
int rays_per_lane_for(int rays, int want);
hipError_t launch_radiance(const RadianceArgs& a, int store, bool traced, bool occ, int rays_per_lane,
                           hipStream_t stream);
// Claims of a staged subset update (a.claim set), launched before it.
hipError_t launch_claim(const RadianceArgs& a, hipStream_t stream);
hipError_t launch_claim_batch(const RadianceArgs* args_dev, const int* pbegin_dev, int n_worlds, int total_probes,
                              hipStream_t stream);
// Compaction of a claimed subset update into x-fastest texel order, and its
// scatter (a.compact set): tile counts, per-world scan, entries; after the
// update, field_in[texel of j] = field_out[j] and the claims reset.  The
// batched forms take per-world tile ranges tbegin[w] .. tbegin[w + 1] - 1.
hipError_t launch_compact(const RadianceArgs& a, hipStream_t stream);
hipError_t launch_compact_scatter(const RadianceArgs& a, int texel_bytes, hipStream_t stream);
hipError_t launch_compact_batch(const RadianceArgs* args_dev, const int* tbegin_dev, int n_worlds, int total_tiles,
                                hipStream_t stream);
hipError_t launch_compact_scatter_batch(const RadianceArgs* args_dev, const int* tbegin_dev, int n_worlds,
                                        int total_tiles, int texel_bytes, hipStream_t stream);
// Whether a staged update of n_probes probes on a world of n_blocks is
// compacted (results are the same either way).
bool compact_pays(long long n_probes, long long n_blocks);
// Applies a staged subset update: field_in[probe of i] = field_out[i].
hipError_t launch_scatter(const RadianceArgs& a, int texel_bytes, hipStream_t stream);
hipError_t launch_scatter_batch(const RadianceArgs* args_dev, const int* pbegin_dev, int n_worlds, int total_probes,
                                int texel_bytes, hipStream_t stream);
hipError_t launch_radiance_tiled(RadianceArgs a, int store, bool traced, const RadianceTune& t, int n_cu,
                                 hipStream_t stream);
// Lane-refill variant: 64 rays per probe, P probes per wave (1..4).
hipError_t launch_radiance_refill(const RadianceArgs& a, int store, int P, hipStream_t stream);
size_t radiance_tiled_lds_bytes(const RadianceArgs& a, const RadianceTune& t);
hipError_t launch_block_bits(const int16_t* blocks, int n_blocks, uint32_t* bits, hipStream_t stream);
hipError_t launch_palette_bits(const uint8_t* palette, int first_slot, int n_slots, uint32_t* bits,
                               hipStream_t stream);
hipError_t launch_map_group(const MapGroup& g, hipStream_t stream);

// Batched map over many worlds (map.hip): per world the staged result of its
// host blockify and where it goes.
// The blockified world is its origin world with block slot_block[k] set to
// k for each slot k in st .. counter - 1 (what blockify allocated), so only
// that list crosses PCIe; the origin comes from the device copy, or from the
// staging (and is written to the device copy) when it changed.
struct MapWorldDesc {
    int16_t* blocks;          // device: blockified world
    int16_t* origin;          // device: copy of the origin world
    int32_t* copy_src;        // device: copy list
    int* state;               // device: blockify state ([0] = counter)
    uint32_t* empty_bits;
    uint8_t* palette;
    uint32_t* pal_bits;
    long long stage_origin;   // byte offset of the staged origin world, -1: the device copy is current
    unsigned long long stage_copy;    // byte offset of the staged copy list (slots st .. hi - 1)
    unsigned long long stage_slots;   // byte offset of the staged slot blocks (slots st .. counter - 1)
    int group_begin;          // first workgroup of this world in the commit launch
    int n_blocks;
    int st, hi;               // slots st .. hi - 1 are copied and get their occupancy bits
    int counter;
    int slot_begin;           // first slot of this world in the concatenated st .. hi - 1 ranges
};
// origin blocks + their empty-block bits + copy list + counter of every
// world: world w owns workgroups group_begin .. (next world's group_begin) - 1
// of 256 threads, one element each.
int map_commit_elements(const MapWorldDesc& d);
hipError_t launch_map_commit_batch(const MapWorldDesc* d_dev, int n, int total_groups, const char* stage_dev,
                                   hipStream_t stream);
// exec_copies of every world's slots st .. hi - 1, and each allocated slot's
// block (its id, its empty bit cleared).
hipError_t launch_map_copies_batch(const MapWorldDesc* d_dev, int n, int total_slots, const char* stage_dev,
                                   hipStream_t stream);
// Models of many worlds in one launch: model i owns workgroups begin[i] .. begin[i + 1] - 1.
hipError_t launch_map_batch(const MapArgs* args_dev, const int* begin_dev, int n_models, int total_groups,
                            hipStream_t stream);
// Occupancy bits of every world's slots st .. hi - 1.
hipError_t launch_palette_bits_batch(const MapWorldDesc* d_dev, int n, int total_slots, hipStream_t stream);
hipError_t launch_exec_copies(const CopyArgs& a, hipStream_t stream);
hipError_t launch_glossy(const GlossyArgs& a, int store, bool traced, hipStream_t stream);
hipError_t launch_diffuse(const DiffuseArgs& a, int store, hipStream_t stream);
hipError_t launch_radiance_batch(const RadianceArgs* args_dev, const int* begin_dev, int n_worlds,
                                 int total_threads, int block, int rays_per_lane, int store, bool served,
                                 const int* frame_dev, hipStream_t stream);
// Frame w owns 16 x 16 pixel tiles begin[w] .. begin[w + 1] - 1 (one workgroup each).
hipError_t launch_glossy_batch(const GlossyArgs* args_dev, const int* begin_dev, int n, int total_tiles, int store,
                               hipStream_t stream);

}  // namespace lumx
