// batch.cpp - many worlds per launch (batched radiance and glossy passes),
// the device frame counter of graph-captured chains, and the graph helpers.
//
// A batched pass writes one kernel-argument record per world plus the first
// thread of each world into a descriptor slot (pinned host staging copied to
// device memory in stream order) and launches one kernel over all worlds.
// Eager passes rotate through a ring of slots and reuse a slot only after the
// kernel that read it has finished; a pass recorded into a graph gets a slot
// of its own that is never rewritten, so every replay reads the same
// descriptors.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "host_internal.hpp"

using namespace lumx;

namespace {

__global__ void frame_kernel(int* frame, int value, int add) {
    if (threadIdx.x == 0) frame[0] = add ? frame[0] + value : value;
}

size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

size_t record_bytes() { return std::max(sizeof(RadianceArgs), sizeof(GlossyArgs)); }

}  // namespace

namespace lumx {

struct DescSlot {
    void* host = nullptr;  // pinned
    void* dev = nullptr;
    size_t bytes = 0;
    hipEvent_t done = nullptr;
    bool pending = false;
};

struct Launcher {
    lumx_context* ctx = nullptr;
    int max_worlds = 0;
    size_t slot_bytes = 0;
    std::vector<DescSlot> ring;
    std::vector<DescSlot> capture_pool;  // pre-allocated, handed to recorded passes
    std::vector<DescSlot> captured;      // owned by recorded graphs until destroy
    int ring_next = 0;
    // Map staging: blocks, copy lists and launch descriptors of a batched
    // map, sized on demand.  Eager maps alternate between two; a map recorded
    // into a graph gets its own, kept until destroy.
    DescSlot map_ring[2];
    int map_next = 0;
    std::vector<DescSlot> map_captured;
    int* frame_dev = nullptr;
    int block = 0;          // radiance workgroup size (0 = 256)
    int rays_per_lane = 0;  // rays per thread of the radiance launches (0 = 1)
};

// Fraction by which an autotuned batch shape must beat the default.
static const float kShapeMargin = 0.03f;

int launcher_set_shape(Launcher* L, int block, int rays_per_lane) {
    if (block != 0 && block != 256 && block != 512 && block != 1024)
        return fail(LUMX_EINVAL, "batched workgroup size %d is not 0, 256, 512 or 1024", block);
    if (rays_per_lane < 0 || rays_per_lane > 4 || rays_per_lane == 3)
        return fail(LUMX_EINVAL, "rays per lane %d is not 0, 1, 2 or 4", rays_per_lane);
    L->block = block;
    L->rays_per_lane = rays_per_lane;
    return LUMX_OK;
}

static void free_slot(DescSlot& s) {
    if (s.done) hipEventDestroy(s.done);
    if (s.host) hipHostFree(s.host);
    if (s.dev) hipFree(s.dev);
    s = DescSlot{};
}

static int alloc_slot(DescSlot& s, size_t bytes) {
    HIP_TRY(hipHostMalloc(&s.host, bytes, hipHostMallocDefault));
    HIP_TRY(hipMalloc(&s.dev, bytes));
    HIP_TRY(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
    s.bytes = bytes;
    return LUMX_OK;
}

void launcher_destroy(Launcher* L) {
    if (!L) return;
    hipSetDevice(L->ctx->device);
    hipDeviceSynchronize();
    for (DescSlot& s : L->ring) free_slot(s);
    for (DescSlot& s : L->capture_pool) free_slot(s);
    for (DescSlot& s : L->captured) free_slot(s);
    for (DescSlot& s : L->map_ring) free_slot(s);
    for (DescSlot& s : L->map_captured) free_slot(s);
    hipFree(L->frame_dev);
    delete L;
}

int launcher_create(lumx_context* ctx, int max_worlds, int ring_slots, Launcher** out) {
    HIP_TRY(hipSetDevice(ctx->device));
    Launcher* L = new Launcher();
    L->ctx = ctx;
    L->max_worlds = max_worlds;
    L->slot_bytes = round_up((size_t)max_worlds * record_bytes(), 256) +
                    3 * round_up(((size_t)max_worlds + 1) * sizeof(int), 256);
    L->ring.resize(ring_slots);
    L->capture_pool.resize(8);
    L->captured.reserve(64);
    for (std::vector<DescSlot>* v : {&L->ring, &L->capture_pool})
        for (DescSlot& s : *v) {
            int rc = alloc_slot(s, L->slot_bytes);
            if (rc) {
                launcher_destroy(L);
                return rc;
            }
        }
    if (hipMalloc(&L->frame_dev, sizeof(int)) != hipSuccess || hipMemset(L->frame_dev, 0, sizeof(int)) != hipSuccess) {
        launcher_destroy(L);
        return fail(LUMX_ENOMEM, "frame counter allocation failed");
    }
    *out = L;
    return LUMX_OK;
}

// A slot whose staging may be rewritten now.
static int acquire_slot(Launcher* L, hipStream_t s, DescSlot** out) {
    if (stream_capturing(s)) {
        DescSlot slot;
        if (!L->capture_pool.empty()) {
            slot = L->capture_pool.back();
            L->capture_pool.pop_back();
        } else {
            int rc = alloc_slot(slot, L->slot_bytes);
            if (rc) return rc;
        }
        L->captured.push_back(slot);
        *out = &L->captured.back();
        return LUMX_OK;
    }
    DescSlot& slot = L->ring[L->ring_next];
    L->ring_next = (L->ring_next + 1) % (int)L->ring.size();
    if (slot.pending) HIP_TRY(hipEventSynchronize(slot.done));
    slot.pending = false;
    *out = &slot;
    return LUMX_OK;
}

// After the kernel that reads the slot: the slot may be reused once this
// event completes (its staging and its device copy are then both free).
static int release_slot(DescSlot* slot, hipStream_t s) {
    if (stream_capturing(s)) return LUMX_OK;
    HIP_TRY(hipEventRecord(slot->done, s));
    slot->pending = true;
    return LUMX_OK;
}

int launcher_set_frame(Launcher* L, int value, int add, hipStream_t s) {
    HIP_TRY(hipSetDevice(L->ctx->device));
    hipLaunchKernelGGL(frame_kernel, dim3(1), dim3(64), 0, s, L->frame_dev, value, add);
    HIP_TRY(hipGetLastError());
    return LUMX_OK;
}

int launcher_radiance(Launcher* L, lumx_world* const* worlds, int n, const int32_t* const* requests,
                      const int* n_requests, const lumx_radiance_params* params, bool use_frame, hipStream_t s,
                      bool served, bool commit) {
    if (n <= 0) return LUMX_OK;
    if (n > L->max_worlds) return fail(LUMX_EINVAL, "%d worlds exceed the launcher's %d", n, L->max_worlds);
    const int format = worlds[0]->desc.radiance_format;
    int max_r = 1, min_r = 1 << 30;
    for (int i = 0; i < n; ++i) {
        const int32_t* req = requests ? requests[i] : nullptr;
        if (req && !n_requests) return fail(LUMX_EINVAL, "request lists need their counts");
        int nr = req ? n_requests[i] : 0;
        int rc = radiance_check(worlds[i], req, nr, &params[i]);
        if (rc) return rc;
        if (worlds[i]->desc.radiance_format != format)
            return fail(LUMX_EINVAL, "worlds of one batched launch share one radiance format");
        const int r = params[i].rays;
        if (r & (r - 1)) return fail(LUMX_EINVAL, "batched rays per probe must be a power of two (world %d: %d)", i, r);
        for (int j = 0; j < i; ++j)
            if (worlds[j] == worlds[i]) return fail(LUMX_EINVAL, "world %d appears twice in one launch", i);
        max_r = std::max(max_r, r);
        min_r = std::min(min_r, r);
    }
    HIP_TRY(hipSetDevice(L->ctx->device));
    // K rays per thread: a probe of R rays owns R / K threads of its world's range.
    const int K = rays_per_lane_for(min_r, L->rays_per_lane);
    const int block = std::max(L->block > 0 ? L->block : 256, max_r / K);
    DescSlot* slot = nullptr;
    int rc = acquire_slot(L, s, &slot);
    if (rc) return rc;
    const size_t args_bytes = round_up((size_t)n * record_bytes(), 256);
    const size_t begin_bytes = round_up(((size_t)n + 1) * sizeof(int), 256);
    RadianceArgs* args = (RadianceArgs*)slot->host;
    int* begin = (int*)((char*)slot->host + args_bytes);
    int* pbegin = (int*)((char*)slot->host + args_bytes + begin_bytes);  // probe ranges of staged worlds
    int* tbegin = (int*)((char*)slot->host + args_bytes + 2 * begin_bytes);  // tile ranges of compacted ones
    long long total = 0, total_staged = 0, total_tiles = 0;
    for (int i = 0; i < n; ++i) {
        lumx_world* w = worlds[i];
        const int32_t* req = requests ? requests[i] : nullptr;
        int nr = req ? n_requests[i] : 0;
        rc = radiance_args(w, req, nr, &params[i], &args[i]);
        if (rc) return rc;
        // A subset update claims its probes before the launch; the scatter
        // that applies it consumes (and resets) the claims.  A large one runs
        // over its claimed probes in texel order (compacted).
        if (args[i].staged && commit) {
            args[i].claim = w->upd_claim;
            if (compact_pays(args[i].n_probes, w->n_blocks)) compact_args(w, &args[i]);
        }
        tbegin[i] = (int)total_tiles;
        total_tiles += args[i].compact ? args[i].n_tiles : 0;
        if (total_tiles > (1ll << 31) - 1) return fail(LUMX_EINVAL, "batch too large");
        begin[i] = (int)total;
        total += round_up((size_t)args[i].n_probes * (args[i].rays / K), (size_t)block);
        if (total > (1ll << 31) - 1) return fail(LUMX_EINVAL, "batch too large");
        pbegin[i] = (int)total_staged;
        if (args[i].staged) total_staged += args[i].n_probes;
    }
    pbegin[n] = (int)total_staged;
    tbegin[n] = (int)total_tiles;
    HIP_TRY(hipMemcpyAsync(slot->dev, slot->host, args_bytes + 3 * begin_bytes, hipMemcpyHostToDevice, s));
    const RadianceArgs* args_dev = (const RadianceArgs*)slot->dev;
    const int* tbegin_dev = (const int*)((char*)slot->dev + args_bytes + 2 * begin_bytes);
    if (commit) {
        HIP_TRY(launch_claim_batch(args_dev, (const int*)((char*)slot->dev + args_bytes + begin_bytes), n,
                                   (int)total_staged, s));
        HIP_TRY(launch_compact_batch(args_dev, tbegin_dev, n, (int)total_tiles, s));
    }
    HIP_TRY(launch_radiance_batch(args_dev, (const int*)((char*)slot->dev + args_bytes), n, (int)total, block, K,
                                  format, served, use_frame ? L->frame_dev : nullptr, s));
    // Subset updates land in their current fields in place.  A trial launch
    // (commit = false) writes the free field buffers only.
    if (commit) {
        bool listed = false;  // a staged world that is not compacted
        for (int i = 0; i < n && !listed; ++i) listed = args[i].staged && !args[i].compact;
        if (listed)
            HIP_TRY(launch_scatter_batch(args_dev, (const int*)((char*)slot->dev + args_bytes + begin_bytes), n,
                                         (int)total_staged, (int)worlds[0]->texel_bytes, s));
        HIP_TRY(launch_compact_scatter_batch(args_dev, tbegin_dev, n, (int)total_tiles,
                                             (int)worlds[0]->texel_bytes, s));
        for (int i = 0; i < n; ++i)
            if (!args[i].staged) worlds[i]->cur ^= 1;  // full updates wrote the other field
    }
    return release_slot(slot, s);
}

// Batched map: per world the host half of lumx_map_models_placed
// (map_prepare), then for all worlds together one upload of the allocated
// slots' blocks, copy lists and launch descriptors (and of an origin world
// that changed), one commit launch (the origin into the blocks with its
// empty-block bits, copy lists, counters), one exec_copies launch that also
// gives each allocated slot's block its id, the map launches (group k of
// every world in launch k: groups of one world stay in order, different
// worlds write disjoint memory) and one palette-bits launch.
int launcher_map(Launcher* L, lumx_world* const* worlds, int n, const int* n_models,
                 const lumx_model_t* const* models, const lumx_placement* const* pl, hipStream_t s) {
    if (n <= 0) return LUMX_OK;
    if (!worlds || !n_models) return fail(LUMX_EINVAL, "null argument");
    for (int i = 0; i < n; ++i) {
        if (!worlds[i] || worlds[i]->ctx != L->ctx) return fail(LUMX_EINVAL, "world %d invalid", i);
        const bool some = n_models[i] > 0;
        int rc = map_check(worlds[i], n_models[i], some && models ? models[i] : nullptr,
                           some && pl ? pl[i] : nullptr);
        if (rc) return rc;
        for (int j = 0; j < i; ++j)
            if (worlds[j] == worlds[i]) return fail(LUMX_EINVAL, "world %d appears twice in one map", i);
    }
    HIP_TRY(hipSetDevice(L->ctx->device));
    const bool capturing = stream_capturing(s);
    // Launch plans (they do not depend on blockify); launch k takes group k
    // of every world that has one.
    std::vector<std::vector<MapArgs>> args(n);
    std::vector<std::vector<int>> ends(n);
    for (int i = 0; i < n; ++i) {
        int rc = map_plan(worlds[i], n_models[i], n_models[i] > 0 ? models[i] : nullptr,
                          n_models[i] > 0 ? pl[i] : nullptr, &args[i], &ends[i]);
        if (rc) return rc;
    }
    // Each map launch's workgroups, checked before blockify changes any world.
    {
        size_t n_launches = 0;
        for (int i = 0; i < n; ++i) n_launches = std::max(n_launches, ends[i].size());
        for (size_t k = 0; k < n_launches; ++k) {
            long long g = 0;
            for (int i = 0; i < n; ++i) {
                if (k >= ends[i].size()) continue;
                for (int j = k ? ends[i][k - 1] : 0; j < ends[i][k]; ++j) g += map_tiles(args[i][j]);
            }
            if (g > (1ll << 31) - 1) return fail(LUMX_EINVAL, "batched map launch too large");
        }
    }
    // Host blockify, world by world.  When world k fails the worlds before it
    // are still committed, as a sequence of single-world maps would leave them.
    int failed_rc = LUMX_OK;
    int m = n;
    for (int i = 0; i < n; ++i) {
        int rc = map_prepare(worlds[i], n_models[i], n_models[i] > 0 ? pl[i] : nullptr);
        if (rc) {
            failed_rc = rc;
            m = i;
            break;
        }
    }
    std::string saved_error;
    if (failed_rc) saved_error = lumx_last_error();
    if (m == 0) return failed_rc;
    size_t n_launches = 0;
    for (int i = 0; i < m; ++i) n_launches = std::max(n_launches, ends[i].size());
    // Staging layout.
    std::vector<MapWorldDesc> desc(m);
    size_t off = round_up((size_t)m * sizeof(MapWorldDesc), 256);
    std::vector<size_t> launch_off(n_launches), launch_begin_off(n_launches);
    std::vector<int> launch_count(n_launches, 0);
    for (size_t k = 0; k < n_launches; ++k) {
        for (int i = 0; i < m; ++i)
            if (k < ends[i].size()) launch_count[k] += ends[i][k] - (k ? ends[i][k - 1] : 0);
        launch_off[k] = off;
        off += round_up((size_t)launch_count[k] * sizeof(MapArgs), 256);
        launch_begin_off[k] = off;
        off += round_up(((size_t)launch_count[k] + 1) * sizeof(int), 256);
    }
    long long groups = 0, slots = 0;
    for (int i = 0; i < m; ++i) {
        lumx_world* w = worlds[i];
        MapWorldDesc& d = desc[i];
        const int st = w->desc.static_blocks;
        const int hi = std::max(w->counter, w->prev_counter);
        for (int k = w->counter; k < hi; ++k) w->copy_src[k] = 0;
        d.blocks = w->blocks;
        d.origin = w->origin_dev;
        d.copy_src = w->copy_src_dev;
        d.state = w->blockify_state;
        d.empty_bits = w->empty_bits;
        d.palette = w->palette;
        d.pal_bits = w->pal_bits;
        d.n_blocks = w->n_blocks;
        d.st = st;
        d.hi = std::max(hi, st);
        d.counter = w->counter;
        // The origin world crosses PCIe only when it changed since its last
        // upload; the blockified world is rebuilt from it on the device.
        d.stage_origin = -1;
        if (w->origin_dirty) {
            d.stage_origin = (long long)off;
            off += round_up((size_t)w->n_blocks * 2, 256);
        }
        d.stage_copy = off;
        off += round_up((size_t)(d.hi - st) * 4, 256);
        d.stage_slots = off;
        off += round_up((size_t)std::max(w->counter - st, 0) * 4, 256);
        d.group_begin = (int)groups;
        groups += (map_commit_elements(d) + 255) / 256;
        d.slot_begin = (int)slots;
        slots += d.hi - st;
    }
    if (groups > (1ll << 31) - 1) return fail(LUMX_EINVAL, "batched map too large");
    const size_t need = off;
    // The staging this map writes.
    DescSlot* slot = nullptr;
    if (capturing) {
        DescSlot fresh;
        int rc = alloc_slot(fresh, need);
        if (rc) return rc;
        L->map_captured.push_back(fresh);
        slot = &L->map_captured.back();
    } else {
        slot = &L->map_ring[L->map_next];
        L->map_next ^= 1;
        if (slot->pending) HIP_TRY(hipEventSynchronize(slot->done));
        slot->pending = false;
        if (slot->bytes < need) {
            free_slot(*slot);
            int rc = alloc_slot(*slot, round_up(need + need / 4, 1 << 16));
            if (rc) return rc;
        }
    }
    char* h = (char*)slot->host;
    char* dv = (char*)slot->dev;
    std::memcpy(h, desc.data(), (size_t)m * sizeof(MapWorldDesc));
    std::vector<int> total_groups(n_launches, 0);
    for (size_t k = 0; k < n_launches; ++k) {
        MapArgs* a = (MapArgs*)(h + launch_off[k]);
        int* begin = (int*)(h + launch_begin_off[k]);
        int c = 0;
        long long g = 0;
        for (int i = 0; i < m; ++i) {
            if (k >= ends[i].size()) continue;
            for (int j = k ? ends[i][k - 1] : 0; j < ends[i][k]; ++j) {
                const MapArgs& x = args[i][j];
                a[c] = x;
                begin[c] = (int)g;
                g += map_tiles(x);
                ++c;
            }
        }
        begin[c] = (int)g;  // below 2^31: checked before blockify
        total_groups[k] = (int)g;
    }
    for (int i = 0; i < m; ++i) {
        lumx_world* w = worlds[i];
        const MapWorldDesc& d = desc[i];
        if (d.stage_origin >= 0) {
            std::memcpy(h + d.stage_origin, w->origin.data(), (size_t)w->n_blocks * 2);
            // A recorded upload runs only when the graph does: eager maps keep uploading.
            if (!capturing) w->origin_dirty = false;
        }
        if (d.hi > d.st) std::memcpy(h + d.stage_copy, w->copy_src.data() + d.st, (size_t)(d.hi - d.st) * 4);
        if (d.counter > d.st)
            std::memcpy(h + d.stage_slots, w->slot_block.data() + d.st, (size_t)(d.counter - d.st) * 4);
    }
    HIP_TRY(hipMemcpyAsync(dv, h, need, hipMemcpyHostToDevice, s));
    const MapWorldDesc* d_dev = (const MapWorldDesc*)dv;
    HIP_TRY(launch_map_commit_batch(d_dev, m, (int)groups, dv, s));
    HIP_TRY(launch_map_copies_batch(d_dev, m, (int)slots, dv, s));
    for (size_t k = 0; k < n_launches; ++k)
        HIP_TRY(launch_map_batch((const MapArgs*)(dv + launch_off[k]), (const int*)(dv + launch_begin_off[k]),
                                 launch_count[k], total_groups[k], s));
    HIP_TRY(launch_palette_bits_batch(d_dev, m, (int)slots, s));
    for (int i = 0; i < m; ++i) {
        worlds[i]->prev_counter = worlds[i]->counter;
        worlds[i]->ev_valid[1] = false;  // batched maps are not timed per world
    }
    if (!capturing) {
        HIP_TRY(hipEventRecord(slot->done, s));
        slot->pending = true;
    }
    if (failed_rc) return fail(failed_rc, "%s", saved_error.c_str());
    return LUMX_OK;
}

int launcher_glossy(Launcher* L, lumx_world* const* worlds, int n, const uint8_t* const* matnorm,
                    const float* const* depth, const int* width, const int* height, const lumx_camera* cams,
                    const float* max_dist, const int* max_steps, float* const* out, hipStream_t s) {
    if (n <= 0) return LUMX_OK;
    if (n > L->max_worlds) return fail(LUMX_EINVAL, "%d frames exceed the launcher's %d", n, L->max_worlds);
    const int format = worlds[0]->desc.radiance_format;
    for (int i = 0; i < n; ++i) {
        if (!worlds[i] || !matnorm[i] || !depth[i] || !out[i]) return fail(LUMX_EINVAL, "frame %d buffers missing", i);
        if (width[i] <= 0 || height[i] <= 0 || width[i] > 16384 || height[i] > 16384)
            return fail(LUMX_EINVAL, "frame %d size %dx%d invalid", i, width[i], height[i]);
        if (!(max_dist[i] > 0.0f && max_dist[i] <= 65536.0f) || max_steps[i] < 0)
            return fail(LUMX_EINVAL, "max_dist must be in (0, 65536] and max_steps >= 0");
        if (worlds[i]->desc.radiance_format != format)
            return fail(LUMX_EINVAL, "worlds of one batched launch share one radiance format");
        if (!worlds[i]->blocks) return fail(LUMX_EINVAL, "world %d is evicted", i);
    }
    HIP_TRY(hipSetDevice(L->ctx->device));
    DescSlot* slot = nullptr;
    int rc = acquire_slot(L, s, &slot);
    if (rc) return rc;
    const size_t args_bytes = round_up((size_t)n * record_bytes(), 256);
    GlossyArgs* args = (GlossyArgs*)slot->host;
    int* begin = (int*)((char*)slot->host + args_bytes);
    long long total = 0;  // 16 x 16 pixel tiles
    for (int i = 0; i < n; ++i) {
        glossy_args(worlds[i], matnorm[i], depth[i], width[i], height[i], &cams[i], max_dist[i], max_steps[i], out[i],
                    nullptr, &args[i]);
        begin[i] = (int)total;
        total += (long long)((width[i] + 15) / 16) * ((height[i] + 15) / 16);
        if (total * 256 > (1ll << 31) - 1) return fail(LUMX_EINVAL, "batch too large");
    }
    begin[n] = (int)total;
    HIP_TRY(hipMemcpyAsync(slot->dev, slot->host, args_bytes + ((size_t)n + 1) * sizeof(int), hipMemcpyHostToDevice,
                           s));
    HIP_TRY(launch_glossy_batch((const GlossyArgs*)slot->dev, (const int*)((char*)slot->dev + args_bytes), n,
                                (int)total, format, s));
    return release_slot(slot, s);
}

}  // namespace lumx

struct lumx_batch {
    Launcher* L = nullptr;
    std::vector<lumx_world*> worlds;
};

struct lumx_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" {

int lumx_batch_create(lumx_context_t ctx, int n, const lumx_world_t* worlds, lumx_batch_t* out) {
    if (!ctx || n <= 0 || !worlds || !out) return fail(LUMX_EINVAL, "bad argument");
    for (int i = 0; i < n; ++i) {
        if (!worlds[i] || worlds[i]->ctx != ctx) return fail(LUMX_EINVAL, "world %d invalid", i);
        if (worlds[i]->desc.radiance_format != worlds[0]->desc.radiance_format)
            return fail(LUMX_EINVAL, "worlds of one batch share one radiance format");
        for (int j = 0; j < i; ++j)
            if (worlds[j] == worlds[i]) return fail(LUMX_EINVAL, "world %d appears twice", i);
    }
    lumx_batch* b = new lumx_batch();
    int rc = launcher_create(ctx, n, 8, &b->L);
    if (rc) {
        delete b;
        return rc;
    }
    b->worlds.assign(worlds, worlds + n);
    *out = b;
    return LUMX_OK;
}

int lumx_batch_destroy(lumx_batch_t b) {
    if (!b) return fail(LUMX_EINVAL, "null batch");
    launcher_destroy(b->L);
    delete b;
    return LUMX_OK;
}

int lumx_batch_set_frame(lumx_batch_t b, int frame, void* stream) {
    if (!b) return fail(LUMX_EINVAL, "null batch");
    return launcher_set_frame(b->L, frame, 0, (hipStream_t)stream);
}

int lumx_batch_advance_frame(lumx_batch_t b, void* stream) {
    if (!b) return fail(LUMX_EINVAL, "null batch");
    return launcher_set_frame(b->L, 1, 1, (hipStream_t)stream);
}

int lumx_batch_radiance(lumx_batch_t b, const int32_t* const* requests, const int* n_requests,
                        const lumx_radiance_params* params, void* stream) {
    if (!b || !params) return fail(LUMX_EINVAL, "null argument");
    return launcher_radiance(b->L, b->worlds.data(), (int)b->worlds.size(), requests, n_requests, params, true,
                             (hipStream_t)stream, false, true);
}

int lumx_batch_autotune(lumx_batch_t b, const int32_t* const* requests, const int* n_requests,
                        const lumx_radiance_params* params, int reps, int* best_block, int* best_rays_per_lane,
                        float* best_ms) {
    if (!b || !params) return fail(LUMX_EINVAL, "null argument");
    if (reps < 1) reps = 1;
    Launcher* L = b->L;
    HIP_TRY(hipSetDevice(L->ctx->device));
    hipStream_t s = L->ctx->stream;
    const int n = (int)b->worlds.size();
    int min_r = 1 << 30;
    for (int i = 0; i < n; ++i) min_r = std::min(min_r, params[i].rays);
    HIP_TRY(hipDeviceSynchronize());  // work queued on other streams must not overlap the timings
    // Candidates: the default (256 threads, 1 ray per lane) first.
    struct Cand {
        int block, k;
        float ms;
    };
    std::vector<Cand> cands;
    for (int k : {1, 2, 4}) {
        if (k > 1 && rays_per_lane_for(min_r, k) != k) continue;
        for (int blk : {256, 512, 1024}) cands.push_back({blk, k, 1e30f});
    }
    EventPair ev;
    const int saved_block = L->block, saved_k = L->rays_per_lane;
    int rc = LUMX_OK;
    hipError_t e = hipEventCreate(&ev.a);
    if (e == hipSuccess) e = hipEventCreate(&ev.b);
    hipEvent_t e0 = ev.a, e1 = ev.b;
    // Two passes over the candidates, keeping each one's faster pass, so the
    // clock ramp-up at the start does not count against the first ones.
    for (int pass = 0; pass < 2; ++pass)
        for (Cand& c : cands) {
            if (e != hipSuccess || rc) break;
            L->block = c.block;
            L->rays_per_lane = c.k;
            // The launch lumx_batch_radiance would make, written to the free
            // field buffers only (no scatter, no flip): the fields the worlds
            // read from are left as they were.
            rc = launcher_radiance(L, b->worlds.data(), n, requests, n_requests, params, true, s, false, false);
            if (rc) break;
            e = hipEventRecord(e0, s);
            for (int r = 0; r < reps && rc == 0 && e == hipSuccess; ++r)
                rc = launcher_radiance(L, b->worlds.data(), n, requests, n_requests, params, true, s, false, false);
            if (e == hipSuccess) e = hipEventRecord(e1, s);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.0f;
            if (e == hipSuccess && !rc) e = hipEventElapsedTime(&ms, e0, e1);
            if (e == hipSuccess && !rc) c.ms = std::min(c.ms, ms / (float)reps);
        }
    L->block = saved_block;
    L->rays_per_lane = saved_k;
    if (rc) return rc;
    if (e != hipSuccess) return fail(LUMX_EDEVICE, "autotune: %s", hipGetErrorString(e));
    // The trial launch stands for launches of other sizes and mixes: another
    // shape replaces the default only when it is clearly faster on it.
    Cand best = cands[0];
    for (const Cand& c : cands)
        if (c.ms < best.ms) best = c;
    if (best.ms > (1.0f - kShapeMargin) * cands[0].ms) best = cands[0];
    L->block = best.block;
    L->rays_per_lane = best.k;
    if (best_block) *best_block = best.block;
    if (best_rays_per_lane) *best_rays_per_lane = best.k;
    if (best_ms) *best_ms = best.ms;
    clear_error();
    return LUMX_OK;
}

int lumx_batch_set_shape(lumx_batch_t b, int block, int rays_per_lane) {
    if (!b) return fail(LUMX_EINVAL, "null batch");
    int rc = launcher_set_shape(b->L, block, rays_per_lane);
    if (rc) return rc;
    clear_error();
    return LUMX_OK;
}

int lumx_batch_map_models(lumx_batch_t b, const int* n_models, const lumx_model_t* const* models,
                          const lumx_placement* const* placements, void* stream) {
    if (!b || !n_models) return fail(LUMX_EINVAL, "null argument");
    return launcher_map(b->L, b->worlds.data(), (int)b->worlds.size(), n_models, models, placements,
                        (hipStream_t)stream);
}

int lumx_batch_glossy(lumx_batch_t b, const uint8_t* const* matnorm, const float* const* depth, const int* width,
                      const int* height, const lumx_camera* cams, float max_dist, int max_steps, float* const* out,
                      void* stream) {
    if (!b || !matnorm || !depth || !width || !height || !cams || !out) return fail(LUMX_EINVAL, "null argument");
    const int n = (int)b->worlds.size();
    std::vector<float> md(n, max_dist);
    std::vector<int> ms(n, max_steps);
    return launcher_glossy(b->L, b->worlds.data(), n, matnorm, depth, width, height, cams, md.data(), ms.data(), out,
                           (hipStream_t)stream);
}

// ---- graphs ----
int lumx_graph_begin(void* stream) {
    if (!stream) return fail(LUMX_EINVAL, "graph capture needs a created stream (not the null stream)");
    HIP_TRY(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeRelaxed));
    return LUMX_OK;
}

int lumx_graph_end(void* stream, lumx_graph_t* out) {
    if (!stream || !out) return fail(LUMX_EINVAL, "bad argument");
    lumx_graph* g = new lumx_graph();
    hipError_t e = hipStreamEndCapture((hipStream_t)stream, &g->graph);
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        if (g->graph) hipGraphDestroy(g->graph);
        delete g;
        return fail(LUMX_EDEVICE, "graph capture failed: %s", hipGetErrorString(e));
    }
    *out = g;
    return LUMX_OK;
}

int lumx_graph_launch(lumx_graph_t g, void* stream) {
    if (!g) return fail(LUMX_EINVAL, "null graph");
    HIP_TRY(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return LUMX_OK;
}

int lumx_graph_destroy(lumx_graph_t g) {
    if (!g) return fail(LUMX_EINVAL, "null graph");
    hipGraphExecDestroy(g->exec);
    hipGraphDestroy(g->graph);
    delete g;
    return LUMX_OK;
}

}  // extern "C"
