// host_internal.hpp - host-side objects of liblumx shared by host.cpp and
// batch.cpp: contexts, worlds, models, error reporting.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lumx.h"
#include "lumx_internal.hpp"

namespace lumx {

// Records the message of the calling thread's last failure; returns code.
int fail(int code, const char* fmt, ...);
void clear_error();

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return ::lumx::fail(LUMX_EDEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                __FILE__, __LINE__);                                               \
    } while (0)

// n elements of T: on the stack up to N, else on the heap.  The
// no-allocation paths stay below N.
template <class T, int N>
struct SmallBuf {
    T local[N];
    std::vector<T> heap;
    T* p;
    explicit SmallBuf(int n, const T& v = T()) {
        if (n <= N) {
            for (int i = 0; i < N; ++i) local[i] = v;
            p = local;
        } else {
            heap.assign((size_t)n, v);
            p = heap.data();
        }
    }
    SmallBuf(const SmallBuf&) = delete;
    SmallBuf& operator=(const SmallBuf&) = delete;
    T& operator[](int i) { return p[i]; }
    const T& operator[](int i) const { return p[i]; }
};

// Two timing events destroyed on every return path.
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    EventPair() = default;
    EventPair(const EventPair&) = delete;
    EventPair& operator=(const EventPair&) = delete;
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

}  // namespace lumx

struct lumx_context {
    int device = 0;
    int n_cu = 256;
    hipStream_t stream = nullptr;
    std::map<int, float4*> fib_tables;  // rays -> device table of per-lane directions
    void* noalloc_scratch = nullptr;     // the device's fixed scratch array of the no-allocation map
    std::mutex mu;
};

struct lumx_model {
    lumx_context* ctx = nullptr;
    int size[3] = {0, 0, 0};
    uint8_t* voxels = nullptr;  // device, x-fastest
};

struct lumx_world {
    lumx_context* ctx = nullptr;
    lumx_world_desc desc{};
    int n_blocks = 0;
    std::vector<int16_t> origin;   // CPU origin world (Lum's origin_world)
    std::vector<int16_t> current;  // blockified world of the last map
    std::vector<int32_t> copy_src; // per palette slot: static source block
    std::vector<int32_t> slot_block; // per palette slot the last host blockify allocated: its block
    int counter = 0;               // palette slots in use (paletteCounter)
    bool block0_empty = true;
    std::vector<uint8_t> static_palette;  // host copy of the static slots
    int16_t* blocks = nullptr;     // device blockified world
    uint8_t* palette = nullptr;    // device palette slots
    float* mats = nullptr;         // device materials
    int32_t* copy_src_dev = nullptr;
    void* field[2] = {nullptr, nullptr};
    int cur = 0;
    size_t texel_bytes = 16;
    hipEvent_t ev[3][2] = {};
    bool ev_valid[3] = {false, false, false};
    int prev_counter = 0;               // palette counter of the previous map
    uint32_t* empty_bits = nullptr;     // device: per block, 1 = id 0
    uint32_t* pal_bits = nullptr;       // device: per slot, 128 words of occupancy bits
    unsigned int* batch_counter = nullptr;  // device: tiled kernel work counter (2 words)
    unsigned long long* sel_bits = nullptr; // device probe selection: flags, one word per 64 probes
    int* sel_tiles = nullptr;               // device probe selection: per-tile counts / offsets
    int* sel_claim = nullptr;               // per block claim word shared by probe selection and
                                            // device blockify (idle between calls)
    int* upd_claim = nullptr;               // per probe: last request of a subset update that
                                            // named it (-1 between updates)
    int* upd_tiles = nullptr;               // compacted subset update: per tile of 2048 probes
                                            // the count, then the offset; then the total
    // Device blockify.
    int16_t* origin_dev = nullptr;          // device copy of the origin world
    bool origin_dirty = true;               // origin changed since its last upload
    bool origin_ids_dirty = true;           // origin changed since origin_has_dynamic was computed
    bool origin_has_dynamic = false;        // an origin id >= static_blocks (a slot two blocks may share)
    int* blockify_state = nullptr;          // device: counter, copy bound, overflow
    void* blockify_scratch = nullptr;       // device: model boxes, ballot words, tile counts
    size_t blockify_scratch_bytes = 0;
    bool host_stale = false;                // current/counter/copy_src live on the device
    int dev_counter_bound = 0;              // upper bound of the device counter
    bool dev_overflow = false;              // the last device map ran out of palette slots
    void* stage_desc = nullptr;             // pinned staging of the device map's model boxes
    size_t stage_desc_bytes = 0;
    int16_t* stage_blocks = nullptr;    // pinned host staging of the blockified world
    hipEvent_t stage_done = nullptr;    // recorded after the staging uploads of the last map
    bool stage_pending = false;
    std::vector<void*> graph_stage;     // pinned staging owned by maps recorded into graphs
    // Device block holding every device part of the world (world_carve), and
    // whether the world allocated it (else a server arena lent it).
    void* dev_mem = nullptr;
    bool dev_own = false;
    // Pinned host snapshot of the device state while evicted.
    void* snap_host = nullptr;
    size_t snap_bytes = 0;
    int snap_used = 0;                  // palette slots in the snapshot
    bool snap_static_valid = false;     // the snapshot holds the static state (blocks .. palette)
    int32_t* stage_copy_src = nullptr;  // pinned host staging of the copy list
    lumx_radiance_tuning tune{};
    size_t device_bytes = 0;            // HBM held by this world (its device block)
};

namespace lumx {

// Bytes of a world's device block.
size_t world_bytes(const lumx_world* w);
// Residency in steps on a stream (the server's copy stream, or the null
// stream for lumx_world_evict / lumx_world_restore):
//   world_snapshot_reserve  pinned snapshot space for `used` palette slots;
//   world_snapshot_async    device state -> snapshot on s (only the field
//                           when `field_only`: the caller knows the static
//                           state has not changed since the last snapshot),
//                           then the world forgets its block (freed after a
//                           sync of s when it is its own; a lent block is the
//                           lender's again);
//   world_attach_async      the world takes `block` (its own when `own`) and
//                           fills it from the snapshot on s: the copies
//                           (world_attach_copy_async), then one kernel that
//                           sets up the rest of the block (world_attach_init:
//                           the other field and unused slots zeroed, claims
//                           idle, counters, occupancy bits), which may run
//                           later on another stream ordered after the copies
//                           and before any other use of the world.
int world_snapshot_reserve(lumx_world* w, int used, hipStream_t s);
int world_snapshot_async(lumx_world* w, hipStream_t s, bool field_only);
// The static state into the snapshot now (synchronous; the world stays
// resident), so later evictions copy only the field.
int world_snapshot_static(lumx_world* w);
int world_attach_async(lumx_world* w, void* block, bool own, hipStream_t s);
int world_attach_copy_async(lumx_world* w, void* block, bool own, hipStream_t s);
int world_attach_init(lumx_world* w, hipStream_t s);
// world_attach_init of n worlds, in launches of up to 16 worlds.
int world_attach_init_many(lumx_world* const* ws, int n, hipStream_t s);
// The world's device parts now live in `block` (a copy of its old block).
void world_rebase(lumx_world* w, void* block, bool own);

// Host-blockified maps (lumx_map_models_placed and its batched form):
//   map_check    arguments of one world's map;
//   map_prepare  start_blockify and blockify of each model on the host world
//                (back at its origin on failure);
//   map_plan     the world's map launches in order: group k is
//                args[ends[k - 1] .. ends[k] - 1] (appended).
int map_check(const lumx_world* w, int n, const lumx_model_t* models, const lumx_placement* pl);
int map_prepare(lumx_world* w, int n, const lumx_placement* pl);
int map_plan(lumx_world* w, int n, const lumx_model_t* models, const lumx_placement* pl, std::vector<MapArgs>* args,
             std::vector<int>* ends);

// Validation shared by single-world and batched radiance updates.
int radiance_check(const lumx_world* w, const int32_t* req, int n_req, const lumx_radiance_params* p);
// The checks of radiance_check that do not need the world.
int radiance_params_check(const lumx_radiance_params* p);
// Kernel arguments of one world's radiance update (field[cur] -> field[cur ^ 1]).
int radiance_args(lumx_world* w, const int32_t* req, int n_req, const lumx_radiance_params* p, RadianceArgs* a);
// Marks a staged update's arguments as compacted (claims of the world's
// subset update, tiles in the world's update scratch).
void compact_args(const lumx_world* w, RadianceArgs* a);
// Kernel arguments of one glossy pass over a G-buffer.
void glossy_args(lumx_world* w, const uint8_t* matnorm, const float* depth, int width, int height,
                 const lumx_camera* cam, float max_dist, int max_steps, float* out, int32_t* keys, GlossyArgs* a);
WorldView view_of(const lumx_world* w);
bool stream_capturing(hipStream_t s);

// Batched launches over any set of distinct worlds (batch.cpp).
struct Launcher;
int launcher_create(lumx_context* ctx, int max_worlds, int ring_slots, Launcher** out);
void launcher_destroy(Launcher* L);
// Radiance launch shape: workgroup size 0 (= 256), 256, 512, 1024; rays per lane 0 (= 1), 1, 2, 4.
int launcher_set_shape(Launcher* L, int block, int rays_per_lane);
int launcher_set_frame(Launcher* L, int value, int add, hipStream_t s);
int launcher_radiance(Launcher* L, lumx_world* const* worlds, int n, const int32_t* const* requests,
                      const int* n_requests, const lumx_radiance_params* params, bool use_frame, hipStream_t s,
                      bool served, bool commit = true);
// Host-blockified maps of many worlds in one upload and a few launches; the
// same result as lumx_map_models_placed on each world in order.
int launcher_map(Launcher* L, lumx_world* const* worlds, int n, const int* n_models,
                 const lumx_model_t* const* models, const lumx_placement* const* pl, hipStream_t s);
int launcher_glossy(Launcher* L, lumx_world* const* worlds, int n, const uint8_t* const* matnorm,
                    const float* const* depth, const int* width, const int* height, const lumx_camera* cams,
                    const float* max_dist, const int* max_steps, float* const* out, hipStream_t s);

}  // namespace lumx
