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
    hipFree(L->frame_dev);
    delete L;
}

int launcher_create(lumx_context* ctx, int max_worlds, int ring_slots, Launcher** out) {
    HIP_TRY(hipSetDevice(ctx->device));
    Launcher* L = new Launcher();
    L->ctx = ctx;
    L->max_worlds = max_worlds;
    L->slot_bytes = round_up((size_t)max_worlds * record_bytes(), 256) +
                    2 * round_up(((size_t)max_worlds + 1) * sizeof(int), 256);
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
    long long total = 0, total_staged = 0;
    for (int i = 0; i < n; ++i) {
        lumx_world* w = worlds[i];
        const int32_t* req = requests ? requests[i] : nullptr;
        int nr = req ? n_requests[i] : 0;
        rc = radiance_args(w, req, nr, &params[i], &args[i]);
        if (rc) return rc;
        // The scatter that applies a subset update consumes (and resets) the claims.
        if (args[i].staged && commit) args[i].claim = w->upd_claim;
        begin[i] = (int)total;
        total += round_up((size_t)args[i].n_probes * (args[i].rays / K), (size_t)block);
        if (total > (1ll << 31) - 1) return fail(LUMX_EINVAL, "batch too large");
        pbegin[i] = (int)total_staged;
        if (args[i].staged) total_staged += args[i].n_probes;
    }
    pbegin[n] = (int)total_staged;
    HIP_TRY(hipMemcpyAsync(slot->dev, slot->host, args_bytes + 2 * begin_bytes, hipMemcpyHostToDevice, s));
    const RadianceArgs* args_dev = (const RadianceArgs*)slot->dev;
    HIP_TRY(launch_radiance_batch(args_dev, (const int*)((char*)slot->dev + args_bytes), n, (int)total, block, K,
                                  format, served, use_frame ? L->frame_dev : nullptr, s));
    // Subset updates land in their current fields in place.  A trial launch
    // (commit = false) writes the free field buffers only.
    if (commit) {
        HIP_TRY(launch_scatter_batch(args_dev, (const int*)((char*)slot->dev + args_bytes + begin_bytes), n,
                                     (int)total_staged, (int)worlds[0]->texel_bytes, s));
        for (int i = 0; i < n; ++i)
            if (!args[i].staged) worlds[i]->cur ^= 1;  // full updates wrote the other field
    }
    return release_slot(slot, s);
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
    long long total = 0;
    for (int i = 0; i < n; ++i) {
        glossy_args(worlds[i], matnorm[i], depth[i], width[i], height[i], &cams[i], max_dist[i], max_steps[i], out[i],
                    nullptr, &args[i]);
        begin[i] = (int)total;
        total += round_up((size_t)width[i] * height[i], 256);
        if (total > (1ll << 31) - 1) return fail(LUMX_EINVAL, "batch too large");
    }
    HIP_TRY(hipMemcpyAsync(slot->dev, slot->host, args_bytes + (size_t)n * sizeof(int), hipMemcpyHostToDevice, s));
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
