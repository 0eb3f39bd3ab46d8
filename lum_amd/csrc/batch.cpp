// batch.cpp - many worlds per launch (batched radiance and glossy passes),
// the device frame counter of graph-captured chains, and the graph helpers.
//
// A batch pass writes one kernel-argument record per world plus the first
// thread of each world into a descriptor slot (pinned host staging copied to
// device memory in stream order) and launches one kernel over all worlds.
// Eager passes rotate through a ring of slots, waiting for a slot's previous
// copy before rewriting it; a pass recorded into a graph gets a slot of its
// own that is never rewritten, so every replay reads the same descriptors.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "host_internal.hpp"

using namespace lumx;

namespace {

struct DescSlot {
    void* host = nullptr;  // pinned
    void* dev = nullptr;
    size_t bytes = 0;
    hipEvent_t copied = nullptr;
    bool pending = false;
};

__global__ void frame_kernel(int* frame, int value, int add) {
    if (threadIdx.x == 0) frame[0] = add ? frame[0] + value : value;
}

}  // namespace

struct lumx_batch {
    lumx_context* ctx = nullptr;
    std::vector<lumx_world*> worlds;
    int format = 0;
    size_t slot_bytes = 0;
    std::vector<DescSlot> ring;
    std::vector<DescSlot> capture_pool;  // pre-allocated, handed to captured passes
    std::vector<DescSlot> captured;      // owned by recorded graphs until destroy
    int ring_next = 0;
    int* frame_dev = nullptr;
};

struct lumx_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static void free_slot(DescSlot& s) {
    if (s.copied) hipEventDestroy(s.copied);
    if (s.host) hipHostFree(s.host);
    if (s.dev) hipFree(s.dev);
    s = DescSlot{};
}

static int alloc_slot(DescSlot& s, size_t bytes) {
    HIP_TRY(hipHostMalloc(&s.host, bytes, hipHostMallocDefault));
    HIP_TRY(hipMalloc(&s.dev, bytes));
    HIP_TRY(hipEventCreateWithFlags(&s.copied, hipEventDisableTiming));
    s.bytes = bytes;
    return LUMX_OK;
}

// A slot whose staging may be rewritten now.
static int acquire_slot(lumx_batch* b, hipStream_t s, DescSlot** out) {
    if (stream_capturing(s)) {
        // A recorded pass keeps its slot for the batch's lifetime.
        DescSlot slot;
        if (!b->capture_pool.empty()) {
            slot = b->capture_pool.back();
            b->capture_pool.pop_back();
        } else {
            int rc = alloc_slot(slot, b->slot_bytes);
            if (rc) return rc;
        }
        b->captured.push_back(slot);
        *out = &b->captured.back();
        return LUMX_OK;
    }
    DescSlot& slot = b->ring[b->ring_next];
    b->ring_next = (b->ring_next + 1) % (int)b->ring.size();
    if (slot.pending) HIP_TRY(hipEventSynchronize(slot.copied));
    slot.pending = false;
    *out = &slot;
    return LUMX_OK;
}

static int upload_slot(DescSlot* slot, size_t bytes, hipStream_t s) {
    HIP_TRY(hipMemcpyAsync(slot->dev, slot->host, bytes, hipMemcpyHostToDevice, s));
    return LUMX_OK;
}

// After the kernel that reads the slot: the slot may be reused once this
// event completes (both its staging and its device copy are then free).
static int release_slot(DescSlot* slot, hipStream_t s) {
    if (stream_capturing(s)) return LUMX_OK;
    HIP_TRY(hipEventRecord(slot->copied, s));
    slot->pending = true;
    return LUMX_OK;
}

static size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

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
    HIP_TRY(hipSetDevice(ctx->device));
    lumx_batch* b = new lumx_batch();
    b->ctx = ctx;
    b->worlds.assign(worlds, worlds + n);
    b->format = worlds[0]->desc.radiance_format;
    const size_t rec = std::max(sizeof(RadianceArgs), sizeof(GlossyArgs));
    b->slot_bytes = round_up((size_t)n * rec, 256) + round_up((size_t)n * sizeof(int), 256);
    b->ring.resize(8);
    b->capture_pool.resize(8);
    b->captured.reserve(64);
    for (std::vector<DescSlot>* v : {&b->ring, &b->capture_pool})
        for (DescSlot& s : *v) {
            int rc = alloc_slot(s, b->slot_bytes);
            if (rc) {
                for (DescSlot& t : b->ring) free_slot(t);
                for (DescSlot& t : b->capture_pool) free_slot(t);
                delete b;
                return rc;
            }
        }
    if (hipMalloc(&b->frame_dev, sizeof(int)) != hipSuccess || hipMemset(b->frame_dev, 0, sizeof(int)) != hipSuccess) {
        for (DescSlot& t : b->ring) free_slot(t);
        for (DescSlot& t : b->capture_pool) free_slot(t);
        delete b;
        return fail(LUMX_ENOMEM, "batch frame counter allocation failed");
    }
    *out = b;
    return LUMX_OK;
}

int lumx_batch_destroy(lumx_batch_t b) {
    if (!b) return fail(LUMX_EINVAL, "null batch");
    hipSetDevice(b->ctx->device);
    hipDeviceSynchronize();
    for (DescSlot& s : b->ring) free_slot(s);
    for (DescSlot& s : b->capture_pool) free_slot(s);
    for (DescSlot& s : b->captured) free_slot(s);
    hipFree(b->frame_dev);
    delete b;
    return LUMX_OK;
}

int lumx_batch_set_frame(lumx_batch_t b, int frame, void* stream) {
    if (!b) return fail(LUMX_EINVAL, "null batch");
    HIP_TRY(hipSetDevice(b->ctx->device));
    hipLaunchKernelGGL(frame_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, b->frame_dev, frame, 0);
    HIP_TRY(hipGetLastError());
    return LUMX_OK;
}

int lumx_batch_advance_frame(lumx_batch_t b, void* stream) {
    if (!b) return fail(LUMX_EINVAL, "null batch");
    HIP_TRY(hipSetDevice(b->ctx->device));
    hipLaunchKernelGGL(frame_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, b->frame_dev, 1, 1);
    HIP_TRY(hipGetLastError());
    return LUMX_OK;
}

int lumx_batch_radiance(lumx_batch_t b, const int32_t* const* requests, const int* n_requests,
                        const lumx_radiance_params* params, void* stream) {
    if (!b || !params) return fail(LUMX_EINVAL, "null argument");
    const int n = (int)b->worlds.size();
    HIP_TRY(hipSetDevice(b->ctx->device));
    hipStream_t s = (hipStream_t)stream;
    int max_r = 1;
    for (int i = 0; i < n; ++i) {
        const int32_t* req = requests ? requests[i] : nullptr;
        int nr = (req && n_requests) ? n_requests[i] : 0;
        if (req && !n_requests) return fail(LUMX_EINVAL, "request lists need their counts");
        int rc = radiance_check(b->worlds[i], req, nr, &params[i]);
        if (rc) return rc;
        const int r = params[i].rays;
        if (r & (r - 1)) return fail(LUMX_EINVAL, "batched rays per probe must be a power of two (world %d: %d)", i, r);
        max_r = std::max(max_r, r);
    }
    const int block = std::max(256, max_r);
    DescSlot* slot = nullptr;
    int rc = acquire_slot(b, s, &slot);
    if (rc) return rc;
    RadianceArgs* args = (RadianceArgs*)slot->host;
    const size_t args_bytes = round_up((size_t)n * std::max(sizeof(RadianceArgs), sizeof(GlossyArgs)), 256);
    int* begin = (int*)((char*)slot->host + args_bytes);
    long long total = 0;
    for (int i = 0; i < n; ++i) {
        lumx_world* w = b->worlds[i];
        const int32_t* req = requests ? requests[i] : nullptr;
        int nr = req ? n_requests[i] : 0;
        rc = radiance_args(w, req, nr, &params[i], &args[i]);
        if (rc) return rc;
        begin[i] = (int)total;
        total += round_up((size_t)args[i].n_probes * args[i].rays, (size_t)block);
        if (total > (1ll << 31) - 1) return fail(LUMX_EINVAL, "batch too large");
        const bool full = (req == nullptr && params[i].size == 1);
        if (!full)
            HIP_TRY(hipMemcpyAsync(args[i].field_out, args[i].field_in, (size_t)w->n_blocks * w->texel_bytes,
                                   hipMemcpyDeviceToDevice, s));
    }
    rc = upload_slot(slot, args_bytes + (size_t)n * sizeof(int), s);
    if (rc) return rc;
    HIP_TRY(launch_radiance_batch((const RadianceArgs*)slot->dev, (const int*)((char*)slot->dev + args_bytes), n,
                                  (int)total, block, b->format, false, b->frame_dev, s));
    rc = release_slot(slot, s);
    if (rc) return rc;
    for (lumx_world* w : b->worlds) w->cur ^= 1;
    return LUMX_OK;
}

int lumx_batch_glossy(lumx_batch_t b, const uint8_t* const* matnorm, const float* const* depth, const int* width,
                      const int* height, const lumx_camera* cams, float max_dist, int max_steps, float* const* out,
                      void* stream) {
    if (!b || !matnorm || !depth || !width || !height || !cams || !out) return fail(LUMX_EINVAL, "null argument");
    if (!(max_dist > 0.0f && max_dist <= 65536.0f) || max_steps < 0)
        return fail(LUMX_EINVAL, "max_dist must be in (0, 65536] and max_steps >= 0");
    const int n = (int)b->worlds.size();
    HIP_TRY(hipSetDevice(b->ctx->device));
    hipStream_t s = (hipStream_t)stream;
    DescSlot* slot = nullptr;
    int rc = acquire_slot(b, s, &slot);
    if (rc) return rc;
    GlossyArgs* args = (GlossyArgs*)slot->host;
    const size_t args_bytes = round_up((size_t)n * std::max(sizeof(RadianceArgs), sizeof(GlossyArgs)), 256);
    int* begin = (int*)((char*)slot->host + args_bytes);
    long long total = 0;
    for (int i = 0; i < n; ++i) {
        if (!matnorm[i] || !depth[i] || !out[i]) return fail(LUMX_EINVAL, "frame %d buffers missing", i);
        if (width[i] <= 0 || height[i] <= 0 || width[i] > 16384 || height[i] > 16384)
            return fail(LUMX_EINVAL, "frame %d size %dx%d invalid", i, width[i], height[i]);
        glossy_args(b->worlds[i], matnorm[i], depth[i], width[i], height[i], &cams[i], max_dist, max_steps, out[i],
                    nullptr, &args[i]);
        begin[i] = (int)total;
        total += round_up((size_t)width[i] * height[i], 256);
        if (total > (1ll << 31) - 1) return fail(LUMX_EINVAL, "batch too large");
    }
    rc = upload_slot(slot, args_bytes + (size_t)n * sizeof(int), s);
    if (rc) return rc;
    HIP_TRY(launch_glossy_batch((const GlossyArgs*)slot->dev, (const int*)((char*)slot->dev + args_bytes), n,
                                (int)total, b->format, s));
    return release_slot(slot, s);
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
