// host.cpp - host runtime of liblumx: contexts, world/palette/material/model
// state in HBM, the host logic of the hot passes and the C ABI of lumx.h.
//
// Host logic restated from the reference (paths relative to it):
//   blockify_mesh / start_blockify / end_blockify   src/renderer/src/internal_render.cpp:977-1097
//   get_shift                                       src/renderer/src/internal_render.cpp:989-1013
//   map_mesh push constants and dispatch            src/renderer/src/internal_render.cpp:1485-1518
//   exec_copies                                     src/renderer/src/internal_render.cpp:1375-1476
//   update_radiance probe selection                 src/renderer/src/internal_render.cpp:1142-1168
//   Camera::updateCamera + UBO fill                 src/renderer/src/internal_render.cpp:923-934,1580-1592
//   updateBlockPalette layout transpose             src/renderer/src/load_stuff.cpp:36-86
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "host_internal.hpp"

using namespace lumx;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_last_error;

int lumx::fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

void lumx::clear_error() { g_last_error.clear(); }

bool lumx::stream_capturing(hipStream_t s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) != hipSuccess) return false;
    return st == hipStreamCaptureStatusActive;
}

static const size_t kLdsBudget = 60 * 1024;  // dynamic LDS per tiled workgroup
// A single-world subset update compacts automatically above this field size
// (one XCD's L2 holds a smaller field whatever the probe order).
static const size_t kCompactFieldBytes = size_t(4) << 20;

static lumx_radiance_tuning default_tuning() {
    lumx_radiance_tuning t{};
    t.variant = 0;
    t.threads = 256;
    t.probes_per_batch = 16;
    t.lds_block_bits = 1;
    t.lds_palette_bits = 1;
    t.wg_per_cu = 0;
    return t;
}

// NULL is the HIP null stream (as for every HIP API); the context's own
// non-blocking stream is available through lumx_context_stream().
static hipStream_t pick_stream(lumx_context*, void* s) { return (hipStream_t)s; }

static int sync_host_state(lumx_world* w);

#define REQUIRE_RESIDENT(w) \
    do { \
        if (!(w)->blocks) return fail(LUMX_EINVAL, "world is evicted (lumx_world_restore first)"); \
    } while (0)

// ---------------------------------------------------------------------------
// host math: quaternion -> matrix, AABB of get_shift, inverse, camera
// ---------------------------------------------------------------------------
namespace {

// Rotation matrix of a unit quaternion (w, x, y, z), FP32 as Lum's
// toMat4(quat); m[c][r] column-major.
void quat_to_mat(const float q[4], float m[4][4]) {
    const float w = q[0], x = q[1], y = q[2], z = q[3];
    const float xx = x * x, yy = y * y, zz = z * z;
    const float xz = x * z, xy = x * y, yz = y * z;
    const float wx = w * x, wy = w * y, wz = w * z;
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) m[c][r] = 0.0f;
    m[0][0] = 1.0f - 2.0f * (yy + zz);
    m[0][1] = 2.0f * (xy + wz);
    m[0][2] = 2.0f * (xz - wy);
    m[1][0] = 2.0f * (xy - wz);
    m[1][1] = 1.0f - 2.0f * (xx + zz);
    m[1][2] = 2.0f * (yz + wx);
    m[2][0] = 2.0f * (xz + wy);
    m[2][1] = 2.0f * (yz - wx);
    m[2][2] = 1.0f - 2.0f * (xx + yy);
    m[3][3] = 1.0f;
}

// General 4x4 inverse (Gauss-Jordan, partial pivoting) in double; m[c][r].
bool invert4(const double a[4][4], double out[4][4]) {
    // Work row-major: A[r][c] = a[c][r].
    double A[4][4], inv[4][4];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) A[r][c] = a[c][r];
    // Gauss-Jordan with partial pivoting.
    double M[4][8];
    for (int r = 0; r < 4; ++r) {
        for (int c = 0; c < 4; ++c) M[r][c] = A[r][c];
        for (int c = 0; c < 4; ++c) M[r][4 + c] = (r == c) ? 1.0 : 0.0;
    }
    for (int col = 0; col < 4; ++col) {
        int piv = col;
        for (int r = col + 1; r < 4; ++r)
            if (std::fabs(M[r][col]) > std::fabs(M[piv][col])) piv = r;
        if (M[piv][col] == 0.0) return false;
        if (piv != col)
            for (int c = 0; c < 8; ++c) std::swap(M[piv][c], M[col][c]);
        double d = M[col][col];
        for (int c = 0; c < 8; ++c) M[col][c] /= d;
        for (int r = 0; r < 4; ++r) {
            if (r == col) continue;
            double f = M[r][col];
            if (f == 0.0) continue;
            for (int c = 0; c < 8; ++c) M[r][c] -= f * M[col][c];
        }
    }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) inv[r][c] = M[r][4 + c];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) out[c][r] = inv[r][c];
    return true;
}

// M * (x, y, z, 1) in double, summed (c0*x + c1*y) + (c2*z + c3*1), then
// rounded to float (get_shift's vec3(trans * vec4(corner, 1.0))).
void xform_point(const double m[4][4], const float p[3], float out[3]) {
    for (int r = 0; r < 3; ++r) {
        double a = m[0][r] * (double)p[0] + m[1][r] * (double)p[1];
        double b = m[2][r] * (double)p[2] + m[3][r] * 1.0;
        out[r] = (float)(a + b);
    }
}

void dnormalize(double v[3]) {
    double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] /= l; v[1] /= l; v[2] /= l;
}
void fnormalize(float v[3]) {
    float l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] /= l; v[1] /= l; v[2] /= l;
}
void dcross(const double a[3], const double b[3], double o[3]) {
    o[0] = a[1] * b[2] - b[1] * a[2];
    o[1] = a[2] * b[0] - b[2] * a[0];
    o[2] = a[0] * b[1] - b[0] * a[1];
}

// Per-lane ray directions of sphericalFibonacci(i, n) (radiance.comp:293-306):
// FP32 operands, cos/sin of the FP32 angle correctly rounded.
void fib_table(int n, std::vector<float4>& t) {
    t.resize(n);
    const float PI = 3.1415926535897932384626433832795f;
    const float PHI = std::sqrt(5.0f) * 0.5f + 0.5f;
    for (int i = 0; i < n; ++i) {
        float fi = (float)i;
        float a = fi * (PHI - 1.0f);
        float frac = a - std::floor(a);
        float phi = 2.0f * PI * frac;
        float ct = 1.0f - (2.0f * fi + 1.0f) * (1.0f / (float)n);
        float s2 = 1.0f - ct * ct;
        s2 = std::min(std::max(s2, 0.0f), 1.0f);
        float st = std::sqrt(s2);
        float c = (float)std::cos((double)phi), s = (float)std::sin((double)phi);
        t[i] = make_float4(c * st, s * st, ct, 0.0f);
    }
}

bool valid_rays(int r) {
    if (r <= 0 || r > 1024) return false;
    if (r <= 64) return (r & (r - 1)) == 0;
    return r % 64 == 0;
}

}  // namespace

static int get_fib(lumx_context* c, int rays, const float4** out) {
    std::lock_guard<std::mutex> lk(c->mu);
    auto it = c->fib_tables.find(rays);
    if (it != c->fib_tables.end()) {
        *out = it->second;
        return LUMX_OK;
    }
    std::vector<float4> t;
    fib_table(rays, t);
    float4* d = nullptr;
    HIP_TRY(hipMalloc(&d, sizeof(float4) * rays));
    HIP_TRY(hipMemcpy(d, t.data(), sizeof(float4) * rays, hipMemcpyHostToDevice));
    c->fib_tables[rays] = d;
    *out = d;
    return LUMX_OK;
}

WorldView lumx::view_of(const lumx_world* w) {
    WorldView v;
    v.blocks = w->blocks;
    v.palette = w->palette;
    v.mats = w->mats;
    v.wx = w->desc.size[0];
    v.wy = w->desc.size[1];
    v.wz = w->desc.size[2];
    v.palette_slots = w->desc.palette_slots;
    v.block0_empty = w->block0_empty ? 1 : 0;
    return v;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* lumx_last_error(void) { return g_last_error.c_str(); }
const char* lumx_version(void) { return "lumx 0.1 (gfx950)"; }

int lumx_context_create(int device, lumx_context_t* out) {
    if (!out) return fail(LUMX_EINVAL, "null output");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(LUMX_EINVAL, "device %d out of range (%d devices)", device, n);
    HIP_TRY(hipSetDevice(device));
    lumx_context* c = new lumx_context();
    c->device = device;
    hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device);
    if (c->n_cu <= 0) c->n_cu = 1;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(LUMX_EDEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    e = noalloc_scratch(device, &c->noalloc_scratch);
    if (e != hipSuccess) {
        hipStreamDestroy(c->stream);
        delete c;
        return fail(LUMX_EDEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    }
    // Direction tables of every power-of-two ray count up front, so that no
    // pass allocates while a graph is being captured.
    for (int r = 1; r <= 1024; r <<= 1) {
        const float4* t = nullptr;
        int rc = get_fib(c, r, &t);
        if (rc) {
            lumx_context_destroy(c);
            return rc;
        }
    }
    *out = c;
    return LUMX_OK;
}

int lumx_context_destroy(lumx_context_t c) {
    if (!c) return fail(LUMX_EINVAL, "null context");
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    for (auto& kv : c->fib_tables) hipFree(kv.second);
    hipStreamDestroy(c->stream);
    delete c;
    return LUMX_OK;
}

int lumx_context_sync(lumx_context_t c) {
    if (!c) return fail(LUMX_EINVAL, "null context");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipDeviceSynchronize());
    return LUMX_OK;
}

void* lumx_context_stream(lumx_context_t c) { return c ? (void*)c->stream : nullptr; }

// ---- world ----
// Device state of a world, carved from one block, each part 256-byte
// aligned, in this order:
//   blocks | copy list | materials | palette slots | field[0] | field[1] |
//   occupancy bits, claims, counters, origin world
// What a world keeps while evicted is its current field and the range from
// the blocks to the last palette slot in use (its "static" state: serving
// only changes fields, so a server copies that range out once).
struct WorldLayout {
    size_t blocks, copy_src, mats, palette, field0, field1, empty_bits, pal_bits, batch_counter, sel_bits,
        sel_tiles, sel_claim, upd_claim, upd_tiles, origin, blockify_state, total;
};

static WorldLayout world_layout(const lumx_world* w) {
    const size_t nb = (size_t)w->n_blocks, slots = (size_t)w->desc.palette_slots;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        off = (off + 255) & ~(size_t)255;
        const size_t at = off;
        off += bytes;
        return at;
    };
    WorldLayout L;
    L.blocks = take(nb * sizeof(int16_t));
    L.copy_src = take(slots * sizeof(int32_t));
    L.mats = take(256 * 6 * sizeof(float));
    L.palette = take(slots * 4096);
    L.field0 = take(nb * w->texel_bytes);
    L.field1 = take(nb * w->texel_bytes);
    L.empty_bits = take(((nb + 31) / 32) * sizeof(uint32_t));
    L.pal_bits = take(slots * 128 * sizeof(uint32_t));
    L.batch_counter = take(2 * sizeof(unsigned int));
    L.sel_bits = take(lumx::select_bits_bytes(w->n_blocks));
    L.sel_tiles = take(lumx::select_tiles(w->n_blocks) * sizeof(int));
    L.sel_claim = take(nb * sizeof(int));
    L.upd_claim = take(nb * sizeof(int));
    L.upd_tiles = take((lumx::select_tiles(w->n_blocks) + 1) * sizeof(int));
    L.origin = take(nb * sizeof(int16_t));
    L.blockify_state = take(4 * sizeof(int));
    L.total = (off + 4095) & ~(size_t)4095;
    return L;
}

static void world_carve(lumx_world* w, char* base) {
    const WorldLayout L = world_layout(w);
    w->blocks = (int16_t*)(base + L.blocks);
    w->copy_src_dev = (int32_t*)(base + L.copy_src);
    w->mats = (float*)(base + L.mats);
    w->palette = (uint8_t*)(base + L.palette);
    w->field[0] = base + L.field0;
    w->field[1] = base + L.field1;
    w->empty_bits = (uint32_t*)(base + L.empty_bits);
    w->pal_bits = (uint32_t*)(base + L.pal_bits);
    w->batch_counter = (unsigned int*)(base + L.batch_counter);
    w->sel_bits = (unsigned long long*)(base + L.sel_bits);
    w->sel_tiles = (int*)(base + L.sel_tiles);
    w->sel_claim = (int*)(base + L.sel_claim);
    w->upd_claim = (int*)(base + L.upd_claim);
    w->upd_tiles = (int*)(base + L.upd_tiles);
    w->origin_dev = (int16_t*)(base + L.origin);
    w->blockify_state = (int*)(base + L.blockify_state);
}

// The host snapshot: the static range (blocks .. palette slot `used`) at
// offset 0, the current field after it.
static size_t snap_static_bytes(const lumx_world* w, int used) {
    return world_layout(w).palette + (size_t)used * 4096;
}
static size_t snap_field_offset(const lumx_world* w, int used) {
    return (snap_static_bytes(w, used) + 255) & ~(size_t)255;
}

extern "C++" {
size_t lumx::world_bytes(const lumx_world* w) { return world_layout(w).total; }

void lumx::world_rebase(lumx_world* w, void* block, bool own) {
    w->dev_mem = block;
    w->dev_own = own;
    world_carve(w, (char*)block);
}
}  // extern "C++"

// What world_attach_kernel needs to set up the rest of w's block, palette
// slots below `used` kept (the next map re-uploads the origin world).
static lumx::AttachArgs attach_args(lumx_world* w, int used) {
    lumx::AttachArgs a{};
    a.field_other = (uint2*)w->field[w->cur ^ 1];
    a.n_field_words = (long long)w->n_blocks * (long long)(w->texel_bytes / 8);
    a.palette = w->palette;
    a.used = used;
    a.slots = w->desc.palette_slots;
    a.blocks = w->blocks;
    a.n_blocks = w->n_blocks;
    a.empty_bits = w->empty_bits;
    a.pal_bits = w->pal_bits;
    a.sel_claim = w->sel_claim;
    a.claim_idle = lumx::select_claim_idle();
    a.upd_claim = w->upd_claim;
    a.batch_counter = w->batch_counter;
    a.blockify_state = w->blockify_state;
    a.counter = w->counter;
    w->origin_dirty = true;
    return a;
}

// Idle values of everything a world does not keep while evicted: the
// other field and the palette slots above `used` zeroed, claims idle,
// counters zero (the map counter `counter`), occupancy bits rebuilt from the
// blocks and the palette (one kernel on s).
static hipError_t world_init_rest(lumx_world* w, int used, hipStream_t s) {
    return lumx::launch_world_attach(attach_args(w, used), s);
}

// A world's own device block, every part zeroed (claims and counters idle).
static hipError_t world_alloc_device(lumx_world* w) {
    const size_t nb = (size_t)w->n_blocks, slots = (size_t)w->desc.palette_slots;
    void* mem = nullptr;
    hipError_t e = hipMalloc(&mem, w->device_bytes);
    if (e != hipSuccess) return e;
    w->dev_mem = mem;
    w->dev_own = true;
    world_carve(w, (char*)mem);
    auto chk = [&](hipError_t r) { if (e == hipSuccess) e = r; };
    chk(hipMemsetAsync(w->blocks, 0, nb * sizeof(int16_t), nullptr));
    chk(hipMemsetAsync(w->mats, 0, 256 * 6 * sizeof(float), nullptr));
    chk(hipMemsetAsync(w->copy_src_dev, 0, slots * sizeof(int32_t), nullptr));
    chk(hipMemsetAsync(w->field[w->cur], 0, nb * w->texel_bytes, nullptr));
    chk(world_init_rest(w, 0, nullptr));
    chk(hipMemsetD32Async((hipDeviceptr_t)w->blockify_state, 0, 1, nullptr));
    chk(hipStreamSynchronize(nullptr));
    return e;
}

// Forgets the device block (freeing it when it is the world's own).
static void world_free_device(lumx_world* w) {
    if (w->dev_own && w->dev_mem) hipFree(w->dev_mem);
    w->dev_mem = nullptr;
    w->dev_own = false;
    // hipFree waits for the whole device even for a null pointer: an
    // eviction from the server's arena must not.
    if (w->blockify_scratch) hipFree(w->blockify_scratch);
    w->blockify_scratch = nullptr;
    w->blockify_scratch_bytes = 0;
    w->blocks = nullptr;
    w->palette = nullptr;
    w->mats = nullptr;
    w->copy_src_dev = nullptr;
    w->field[0] = w->field[1] = nullptr;
    w->empty_bits = nullptr;
    w->pal_bits = nullptr;
    w->batch_counter = nullptr;
    w->sel_bits = nullptr;
    w->sel_tiles = nullptr;
    w->sel_claim = nullptr;
    w->upd_claim = nullptr;
    w->origin_dev = nullptr;
    w->blockify_state = nullptr;
}

extern "C++" {
int lumx::world_snapshot_reserve(lumx_world* w, int used, hipStream_t s) {
    const size_t need = snap_field_offset(w, used) + (size_t)w->n_blocks * w->texel_bytes;
    if (need <= w->snap_bytes) return LUMX_OK;
    if (w->snap_host) {
        // The previous snapshot may still be read or written by copies on s.
        HIP_TRY(hipStreamSynchronize(s));
        hipHostFree(w->snap_host);
        w->snap_host = nullptr;
        w->snap_bytes = 0;
        w->snap_static_valid = false;
    }
    HIP_TRY(hipHostMalloc(&w->snap_host, need, hipHostMallocDefault));
    w->snap_bytes = need;
    return LUMX_OK;
}

int lumx::world_snapshot_async(lumx_world* w, hipStream_t s, bool field_only) {
    int rc = sync_host_state(w);
    if (rc) return rc;
    if (w->stage_pending) HIP_TRY(hipEventSynchronize(w->stage_done));
    w->stage_pending = false;
    const int used = std::max(w->counter, w->prev_counter);
    if (field_only && !(w->snap_static_valid && used == w->snap_used))
        return fail(LUMX_EINVAL, "no snapshot of the world's static state to keep");
    rc = world_snapshot_reserve(w, used, s);
    if (rc) return rc;
    // Copies from the device on s, so the snapshot is written in stream
    // order after whatever copy last read it.
    char* h = (char*)w->snap_host;
    if (!field_only || !w->snap_static_valid)
        HIP_TRY(hipMemcpyAsync(h, w->dev_mem, snap_static_bytes(w, used), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(h + snap_field_offset(w, used), w->field[w->cur], (size_t)w->n_blocks * w->texel_bytes,
                           hipMemcpyDeviceToHost, s));
    w->snap_used = used;
    w->snap_static_valid = true;
    if (w->dev_own) HIP_TRY(hipStreamSynchronize(s));  // the block is freed below
    world_free_device(w);
    return LUMX_OK;
}

int lumx::world_snapshot_static(lumx_world* w) {
    const int used = std::max(w->counter, w->prev_counter);
    int rc = world_snapshot_reserve(w, used, nullptr);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(w->snap_host, w->dev_mem, snap_static_bytes(w, used), hipMemcpyDeviceToHost));
    w->snap_used = used;
    w->snap_static_valid = true;
    return LUMX_OK;
}

int lumx::world_attach_copy_async(lumx_world* w, void* block, bool own, hipStream_t s) {
    if (!w->snap_host) return fail(LUMX_EINVAL, "world has no snapshot to restore");
    w->dev_mem = block;
    w->dev_own = own;
    world_carve(w, (char*)block);
    const int used = w->snap_used;
    const char* h = (const char*)w->snap_host;
    const size_t field_bytes = (size_t)w->n_blocks * w->texel_bytes;
    if ((char*)w->field[0] == (char*)block + snap_field_offset(w, used)) {
        // Every palette slot in use: the snapshot is the block's own layout
        // up to field 0's end, so one copy, and the field comes back as
        // field 0 (which of the two buffers holds it is the world's own
        // bookkeeping: the other is zeroed by world_attach_init).
        HIP_TRY(hipMemcpyAsync(block, h, snap_field_offset(w, used) + field_bytes, hipMemcpyHostToDevice, s));
        w->cur = 0;
    } else {
        HIP_TRY(hipMemcpyAsync(block, h, snap_static_bytes(w, used), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(w->field[w->cur], h + snap_field_offset(w, used), field_bytes, hipMemcpyHostToDevice,
                               s));
    }
    return LUMX_OK;
}

int lumx::world_attach_init(lumx_world* w, hipStream_t s) {
    HIP_TRY(world_init_rest(w, w->snap_used, s));
    return LUMX_OK;
}

int lumx::world_attach_init_many(lumx_world* const* ws, int n, hipStream_t s) {
    std::vector<lumx::AttachArgs> a((size_t)n);
    for (int i = 0; i < n; ++i) a[i] = attach_args(ws[i], ws[i]->snap_used);
    HIP_TRY(lumx::launch_world_attach_batch(a.data(), n, s));
    return LUMX_OK;
}

int lumx::world_attach_async(lumx_world* w, void* block, bool own, hipStream_t s) {
    int rc = world_attach_copy_async(w, block, own, s);
    if (rc) return rc;
    return world_attach_init(w, s);
}
}  // extern "C++"

static void world_free(lumx_world* w) {
    world_free_device(w);
    if (w->snap_host) hipHostFree(w->snap_host);
    if (w->stage_blocks) hipHostFree(w->stage_blocks);
    if (w->stage_copy_src) hipHostFree(w->stage_copy_src);
    if (w->stage_desc) hipHostFree(w->stage_desc);
    if (w->stage_done) hipEventDestroy(w->stage_done);
    for (void* p : w->graph_stage) hipHostFree(p);
    for (int p = 0; p < 3; ++p)
        for (int k = 0; k < 2; ++k)
            if (w->ev[p][k]) hipEventDestroy(w->ev[p][k]);
}

int lumx_world_create(lumx_context_t c, const lumx_world_desc* d, lumx_world_t* out) {
    if (!c || !d || !out) return fail(LUMX_EINVAL, "null argument");
    for (int a = 0; a < 3; ++a)
        if (d->size[a] <= 0 || d->size[a] > 4096) return fail(LUMX_EINVAL, "world size[%d]=%d invalid", a, d->size[a]);
    long long nb = (long long)d->size[0] * d->size[1] * d->size[2];
    if (nb > (1ll << 28)) return fail(LUMX_EINVAL, "world too large (%lld blocks)", nb);
    if (d->static_blocks < 1 || d->static_blocks > 32767)
        return fail(LUMX_EINVAL, "static_blocks=%d invalid", d->static_blocks);
    if (d->radiance_format < 0 || d->radiance_format > 2)
        return fail(LUMX_EINVAL, "radiance_format=%d invalid", d->radiance_format);
    HIP_TRY(hipSetDevice(c->device));
    lumx_world* w = new lumx_world();
    w->ctx = c;
    w->desc = *d;
    w->n_blocks = (int)nb;
    long long slots = d->palette_slots;
    if (slots <= 0) slots = std::min<long long>(nb + d->static_blocks, 32768);
    if (slots < d->static_blocks || slots > 32768) {
        delete w;
        return fail(LUMX_EINVAL, "palette_slots=%lld invalid", slots);
    }
    w->desc.palette_slots = (int)slots;
    w->texel_bytes = d->radiance_format == LUMX_RADIANCE_F32 ? 16 : 8;
    w->origin.assign(nb, 0);
    w->current.assign(nb, 0);
    w->copy_src.assign(slots, 0);
    w->slot_block.assign(slots, 0);
    w->static_palette.assign((size_t)d->static_blocks * 4096, 0);
    w->counter = d->static_blocks;
    w->prev_counter = d->static_blocks;
    w->tune = default_tuning();
    w->device_bytes = lumx::world_bytes(w);
    hipError_t e = world_alloc_device(w);
    auto chk = [&](hipError_t r) { if (e == hipSuccess) e = r; };
    chk(hipHostMalloc((void**)&w->stage_blocks, nb * sizeof(int16_t), hipHostMallocDefault));
    chk(hipEventCreateWithFlags(&w->stage_done, hipEventDisableTiming));
    chk(hipHostMalloc((void**)&w->stage_copy_src, ((size_t)slots + 4) * sizeof(int32_t), hipHostMallocDefault));
    for (int p = 0; p < 3; ++p)
        for (int k = 0; k < 2; ++k) chk(hipEventCreate(&w->ev[p][k]));
    if (e != hipSuccess) {
        world_free(w);
        delete w;
        if (e == hipErrorOutOfMemory) return fail(LUMX_ENOMEM, "device allocation failed");
        return fail(LUMX_EDEVICE, "world allocation: %s", hipGetErrorString(e));
    }
    *out = w;
    return LUMX_OK;
}

int lumx_world_destroy(lumx_world_t w) {
    if (!w) return fail(LUMX_EINVAL, "null world");
    hipSetDevice(w->ctx->device);
    hipDeviceSynchronize();
    world_free(w);
    delete w;
    return LUMX_OK;
}

int lumx_world_info(lumx_world_t w, lumx_world_desc* out) {
    if (!w || !out) return fail(LUMX_EINVAL, "null argument");
    *out = w->desc;
    return LUMX_OK;
}

static int check_block_ids(const lumx_world* w, const int16_t* b, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (b[i] < 0 || b[i] >= w->desc.palette_slots)
            return fail(LUMX_EINVAL, "block id %d at %zu outside the palette [0, %d)", (int)b[i], i,
                        w->desc.palette_slots);
    return LUMX_OK;
}

int lumx_world_set_blocks(lumx_world_t w, const int16_t* blocks) {
    if (!w || !blocks) return fail(LUMX_EINVAL, "null argument");
    int rc = check_block_ids(w, blocks, w->n_blocks);
    if (rc) return rc;
    std::memcpy(w->origin.data(), blocks, (size_t)w->n_blocks * 2);
    w->origin_dirty = true;
    w->origin_ids_dirty = true;
    return LUMX_OK;
}

int lumx_world_get_blocks(lumx_world_t w, int16_t* blocks, int blockified) {
    if (!w || !blocks) return fail(LUMX_EINVAL, "null argument");
    if (blockified) {
        int rc = sync_host_state(w);
        if (rc) return rc;
    }
    std::memcpy(blocks, blockified ? w->current.data() : w->origin.data(), (size_t)w->n_blocks * 2);
    return LUMX_OK;
}

static bool in_world(const lumx_world* w, int x, int y, int z) {
    return x >= 0 && y >= 0 && z >= 0 && x < w->desc.size[0] && y < w->desc.size[1] && z < w->desc.size[2];
}

int lumx_world_set_block(lumx_world_t w, int x, int y, int z, int16_t b) {
    if (!w) return fail(LUMX_EINVAL, "null world");
    if (!in_world(w, x, y, z)) return fail(LUMX_EINVAL, "block (%d,%d,%d) outside the world", x, y, z);
    if (b < 0 || b >= w->desc.palette_slots) return fail(LUMX_EINVAL, "block id %d invalid", (int)b);
    w->origin[x + w->desc.size[0] * (y + w->desc.size[1] * z)] = b;
    w->origin_dirty = true;
    w->origin_ids_dirty = true;
    return LUMX_OK;
}

int lumx_world_get_block(lumx_world_t w, int x, int y, int z, int16_t* out) {
    if (!w || !out) return fail(LUMX_EINVAL, "null argument");
    if (!in_world(w, x, y, z)) return fail(LUMX_EINVAL, "block (%d,%d,%d) outside the world", x, y, z);
    *out = w->origin[x + w->desc.size[0] * (y + w->desc.size[1] * z)];
    return LUMX_OK;
}

int lumx_world_set_palette(lumx_world_t w, int first, int count, const uint8_t* vox, int layout) {
    if (!w || (!vox && count > 0)) return fail(LUMX_EINVAL, "null argument");
    REQUIRE_RESIDENT(w);
    if (first < 0 || count < 0 || first + count > w->desc.static_blocks)
        return fail(LUMX_EINVAL, "palette range [%d, %d) outside the static palette [0, %d)", first, first + count,
                    w->desc.static_blocks);
    if (layout != LUMX_PALETTE_XYZ_XFAST && layout != LUMX_PALETTE_REF_ZFAST)
        return fail(LUMX_EINVAL, "palette layout %d invalid", layout);
    std::vector<uint8_t> tmp((size_t)count * 4096);
    for (int b = 0; b < count; ++b) {
        const uint8_t* src = vox + (size_t)b * 4096;
        uint8_t* dst = tmp.data() + (size_t)b * 4096;
        if (layout == LUMX_PALETTE_XYZ_XFAST) {
            std::memcpy(dst, src, 4096);
        } else {
            // BlockVoxels[x][y][z]: z fastest in memory.
            for (int x = 0; x < 16; ++x)
                for (int y = 0; y < 16; ++y)
                    for (int z = 0; z < 16; ++z) dst[x + 16 * y + 256 * z] = src[(x * 16 + y) * 16 + z];
        }
    }
    std::memcpy(w->static_palette.data() + (size_t)first * 4096, tmp.data(), tmp.size());
    HIP_TRY(hipSetDevice(w->ctx->device));
    HIP_TRY(hipDeviceSynchronize());
    bool e0 = true;
    for (int i = 0; i < 4096; ++i)
        if (w->static_palette[i]) { e0 = false; break; }
    w->block0_empty = e0;
    HIP_TRY(hipMemcpy(w->palette + (size_t)first * 4096, tmp.data(), tmp.size(), hipMemcpyHostToDevice));
    HIP_TRY(launch_palette_bits(w->palette, first, count, w->pal_bits, nullptr));
    HIP_TRY(hipDeviceSynchronize());
    return LUMX_OK;
}

int lumx_world_get_palette(lumx_world_t w, int first, int count, uint8_t* out) {
    if (!w || (!out && count > 0)) return fail(LUMX_EINVAL, "null argument");
    REQUIRE_RESIDENT(w);
    if (first < 0 || count < 0 || first + count > w->desc.palette_slots)
        return fail(LUMX_EINVAL, "palette range outside [0, %d)", w->desc.palette_slots);
    HIP_TRY(hipSetDevice(w->ctx->device));
    HIP_TRY(hipStreamSynchronize(w->ctx->stream));
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(out, w->palette + (size_t)first * 4096, (size_t)count * 4096, hipMemcpyDeviceToHost));
    return LUMX_OK;
}

int lumx_world_set_materials(lumx_world_t w, const float* m) {
    if (!w || !m) return fail(LUMX_EINVAL, "null argument");
    REQUIRE_RESIDENT(w);
    HIP_TRY(hipSetDevice(w->ctx->device));
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(w->mats, m, 256 * 6 * sizeof(float), hipMemcpyHostToDevice));
    return LUMX_OK;
}

static uint16_t host_f32_to_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
static float host_bf16_to_f32(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

int lumx_world_write_radiance(lumx_world_t w, const float* rgba) {
    if (!w || !rgba) return fail(LUMX_EINVAL, "null argument");
    REQUIRE_RESIDENT(w);
    size_t n = (size_t)w->n_blocks;
    HIP_TRY(hipSetDevice(w->ctx->device));
    HIP_TRY(hipDeviceSynchronize());
    if (w->desc.radiance_format == LUMX_RADIANCE_F32) {
        HIP_TRY(hipMemcpy(w->field[w->cur], rgba, n * 16, hipMemcpyHostToDevice));
        return LUMX_OK;
    }
    std::vector<uint16_t> h(n * 4);
    for (size_t i = 0; i < n * 4; ++i) {
        if (w->desc.radiance_format == LUMX_RADIANCE_F16) {
            _Float16 v = (_Float16)rgba[i];
            std::memcpy(&h[i], &v, 2);
        } else {
            h[i] = host_f32_to_bf16(rgba[i]);
        }
    }
    HIP_TRY(hipMemcpy(w->field[w->cur], h.data(), n * 8, hipMemcpyHostToDevice));
    return LUMX_OK;
}

int lumx_world_read_radiance(lumx_world_t w, float* rgba) {
    if (!w || !rgba) return fail(LUMX_EINVAL, "null argument");
    REQUIRE_RESIDENT(w);
    size_t n = (size_t)w->n_blocks;
    HIP_TRY(hipSetDevice(w->ctx->device));
    HIP_TRY(hipDeviceSynchronize());
    if (w->desc.radiance_format == LUMX_RADIANCE_F32) {
        HIP_TRY(hipMemcpy(rgba, w->field[w->cur], n * 16, hipMemcpyDeviceToHost));
        return LUMX_OK;
    }
    std::vector<uint16_t> h(n * 4);
    HIP_TRY(hipMemcpy(h.data(), w->field[w->cur], n * 8, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n * 4; ++i) {
        if (w->desc.radiance_format == LUMX_RADIANCE_F16) {
            _Float16 v;
            std::memcpy(&v, &h[i], 2);
            rgba[i] = (float)v;
        } else {
            rgba[i] = host_bf16_to_f32(h[i]);
        }
    }
    return LUMX_OK;
}

int lumx_world_radiance_dev(lumx_world_t w, void** ptr, size_t* bytes) {
    if (!w || !ptr || !bytes) return fail(LUMX_EINVAL, "null argument");
    REQUIRE_RESIDENT(w);
    *ptr = w->field[w->cur];
    *bytes = (size_t)w->n_blocks * w->texel_bytes;
    return LUMX_OK;
}

// end_blockify + exec_copies (internal_render.cpp:1092-1097,1375-1476): upload
// the blockified world, rebuild its empty-block bits, fill every slot
// allocated by blockify with its static source (or zeros) and zero the slots
// the previous frame allocated beyond the new counter, so every slot at or
// above the counter reads as air like the reference's cleared palette.
static int commit_blocks(lumx_world* w, hipStream_t s) {
    // Pinned staging: the copies read it when they execute.  Eager maps share
    // one staging and wait for the previous map's copies before rewriting it;
    // a map recorded into a graph gets staging of its own, read on every replay.
    int16_t* stage_blocks = w->stage_blocks;
    int32_t* stage_copy = w->stage_copy_src;
    if (stream_capturing(s)) {
        // Blocks, then (16-byte aligned) the copy list and the counter word:
        // the same layout as the eager staging's palette_slots + 4 words.
        void* p = nullptr;
        const size_t copy_off = (((size_t)w->n_blocks * 2 + 15) / 16) * 16;
        const size_t bytes = copy_off + ((size_t)w->desc.palette_slots + 4) * 4;
        HIP_TRY(hipHostMalloc(&p, bytes, hipHostMallocDefault));
        w->graph_stage.push_back(p);
        stage_blocks = (int16_t*)p;
        stage_copy = (int32_t*)((char*)p + copy_off);
    } else {
        if (w->stage_pending) HIP_TRY(hipEventSynchronize(w->stage_done));
        w->stage_pending = false;
    }
    std::memcpy(stage_blocks, w->current.data(), (size_t)w->n_blocks * 2);
    HIP_TRY(hipMemcpyAsync(w->blocks, stage_blocks, (size_t)w->n_blocks * 2, hipMemcpyHostToDevice, s));
    HIP_TRY(launch_block_bits(w->blocks, w->n_blocks, w->empty_bits, s));
    const int st = w->desc.static_blocks;
    const int hi = std::max(w->counter, w->prev_counter);
    for (int k = w->counter; k < hi; ++k) w->copy_src[k] = 0;
    if (hi > st) {
        std::memcpy(stage_copy + st, w->copy_src.data() + st, (size_t)(hi - st) * 4);
        HIP_TRY(hipMemcpyAsync(w->copy_src_dev + st, stage_copy + st, (size_t)(hi - st) * 4,
                               hipMemcpyHostToDevice, s));
        CopyArgs ca{w->palette, w->copy_src_dev, st, hi - st};
        HIP_TRY(launch_exec_copies(ca, s));
    }
    // The device copy of the counter: the next device map's previous counter.
    stage_copy[w->desc.palette_slots] = w->counter;
    HIP_TRY(hipMemcpyAsync(w->blockify_state, stage_copy + w->desc.palette_slots, sizeof(int), hipMemcpyHostToDevice,
                           s));
    return LUMX_OK;
}

// After a device map the blockified world, the counter and the copy list live
// on the device; host code that reads them brings them back first.
static int sync_host_state(lumx_world* w) {
    if (!w->host_stale) return LUMX_OK;
    HIP_TRY(hipSetDevice(w->ctx->device));
    HIP_TRY(hipDeviceSynchronize());
    int state[4];
    HIP_TRY(hipMemcpy(state, w->blockify_state, sizeof(state), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(w->current.data(), w->blocks, (size_t)w->n_blocks * 2, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(w->copy_src.data(), w->copy_src_dev, (size_t)w->desc.palette_slots * 4, hipMemcpyDeviceToHost));
    w->counter = state[0];
    w->prev_counter = state[0];
    w->dev_overflow = state[2] != 0;
    w->host_stale = false;
    return LUMX_OK;
}

// Occupancy bits of the dynamic slots after the palette changed on device.
static int commit_palette_bits(lumx_world* w, hipStream_t s) {
    const int st = w->desc.static_blocks;
    const int hi = std::max(w->counter, w->prev_counter);
    if (hi > st) HIP_TRY(launch_palette_bits(w->palette, st, hi - st, w->pal_bits, s));
    w->prev_counter = w->counter;
    return LUMX_OK;
}

int lumx_world_upload_blocks(lumx_world_t w, void* stream) {
    if (!w) return fail(LUMX_EINVAL, "null world");
    REQUIRE_RESIDENT(w);
    HIP_TRY(hipSetDevice(w->ctx->device));
    int rc0 = sync_host_state(w);
    if (rc0) return rc0;
    w->current = w->origin;
    w->counter = w->desc.static_blocks;
    hipStream_t s = pick_stream(w->ctx, stream);
    int rc = commit_blocks(w, s);
    if (rc) return rc;
    rc = commit_palette_bits(w, s);
    if (rc) return rc;
    if (stream_capturing(s)) return LUMX_OK;
    HIP_TRY(hipEventRecord(w->stage_done, s));
    w->stage_pending = true;
    return LUMX_OK;
}

int lumx_world_evict(lumx_world_t w) {
    if (!w) return fail(LUMX_EINVAL, "null world");
    if (!w->blocks) return LUMX_OK;
    HIP_TRY(hipSetDevice(w->ctx->device));
    // Work on the world may be queued on any stream.
    HIP_TRY(hipDeviceSynchronize());
    return world_snapshot_async(w, nullptr, false);
}

int lumx_world_restore(lumx_world_t w) {
    if (!w) return fail(LUMX_EINVAL, "null world");
    if (w->blocks) return LUMX_OK;
    HIP_TRY(hipSetDevice(w->ctx->device));
    void* mem = nullptr;
    hipError_t e = hipMalloc(&mem, w->device_bytes);
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? LUMX_ENOMEM : LUMX_EDEVICE, "restore: %s", hipGetErrorString(e));
    int rc = world_attach_async(w, mem, true, nullptr);
    if (rc == LUMX_OK) {
        hipError_t se = hipStreamSynchronize(nullptr);
        if (se != hipSuccess) rc = fail(LUMX_EDEVICE, "restore: %s", hipGetErrorString(se));
    }
    if (rc) world_free_device(w);
    return rc;
}

int lumx_world_resident(lumx_world_t w, int* resident, size_t* device_bytes) {
    if (!w) return fail(LUMX_EINVAL, "null world");
    if (resident) *resident = w->blocks != nullptr;
    if (device_bytes) *device_bytes = w->device_bytes;
    return LUMX_OK;
}

int lumx_world_palette_counter(lumx_world_t w, int* out) {
    if (!w || !out) return fail(LUMX_EINVAL, "null argument");
    int rc = sync_host_state(w);
    if (rc) return rc;
    *out = w->counter;
    if (w->dev_overflow) return fail(LUMX_EOVERFLOW, "the last device map ran out of palette slots (%d)",
                                     w->desc.palette_slots);
    return LUMX_OK;
}

int lumx_world_last_pass_ms(lumx_world_t w, int pass, float* ms) {
    if (!w || !ms || pass < 0 || pass > 2) return fail(LUMX_EINVAL, "bad argument");
    if (!w->ev_valid[pass]) return fail(LUMX_EINVAL, "no timed pass of kind %d yet", pass);
    HIP_TRY(hipEventSynchronize(w->ev[pass][1]));
    HIP_TRY(hipEventElapsedTime(ms, w->ev[pass][0], w->ev[pass][1]));
    return LUMX_OK;
}

// ---- models ----
int lumx_model_create(lumx_context_t c, const uint8_t* vox, int sx, int sy, int sz, lumx_model_t* out) {
    if (!c || !vox || !out) return fail(LUMX_EINVAL, "null argument");
    if (sx <= 0 || sy <= 0 || sz <= 0 || sx > 1024 || sy > 1024 || sz > 1024)
        return fail(LUMX_EINVAL, "model size %dx%dx%d invalid", sx, sy, sz);
    HIP_TRY(hipSetDevice(c->device));
    lumx_model* m = new lumx_model();
    m->ctx = c;
    m->size[0] = sx; m->size[1] = sy; m->size[2] = sz;
    size_t n = (size_t)sx * sy * sz;
    hipError_t e = hipMalloc(&m->voxels, n);
    if (e == hipSuccess) e = hipMemcpy(m->voxels, vox, n, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hipFree(m->voxels);
        delete m;
        return fail(e == hipErrorOutOfMemory ? LUMX_ENOMEM : LUMX_EDEVICE, "model upload: %s", hipGetErrorString(e));
    }
    *out = m;
    return LUMX_OK;
}

int lumx_model_destroy(lumx_model_t m) {
    if (!m) return fail(LUMX_EINVAL, "null model");
    hipSetDevice(m->ctx->device);
    hipDeviceSynchronize();
    hipFree(m->voxels);
    delete m;
    return LUMX_OK;
}

int lumx_model_size(lumx_model_t m, int size[3]) {
    if (!m || !size) return fail(LUMX_EINVAL, "null argument");
    size[0] = m->size[0]; size[1] = m->size[1]; size[2] = m->size[2];
    return LUMX_OK;
}

int lumx_model_placement(const lumx_transform* t, const int size[3], lumx_placement* out) {
    if (!t || !size || !out) return fail(LUMX_EINVAL, "null argument");
    float r[4][4];
    quat_to_mat(t->rot, r);
    // translate(I, shift) * rotate == [R | shift] exactly; held in double.
    double m[4][4];
    for (int c = 0; c < 4; ++c)
        for (int k = 0; k < 4; ++k) m[c][k] = (double)r[c][k];
    m[3][0] = (double)t->shift[0];
    m[3][1] = (double)t->shift[1];
    m[3][2] = (double)t->shift[2];
    m[3][3] = 1.0;
    // get_shift: AABB of M over the corners of [0, size-1].
    float box[3] = {(float)(size[0] - 1), (float)(size[1] - 1), (float)(size[2] - 1)};
    float mn[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
    float mx[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
    for (int k = 0; k < 8; ++k) {
        float p[3] = {(k & 1) ? box[0] : 0.0f, (k & 2) ? box[1] : 0.0f, (k & 4) ? box[2] : 0.0f};
        float q[3];
        xform_point(m, p, q);
        for (int a = 0; a < 3; ++a) {
            mn[a] = std::min(mn[a], q[a]);
            mx[a] = std::max(mx[a], q[a]);
        }
    }
    double inv[4][4];
    if (!invert4(m, inv)) return fail(LUMX_EINVAL, "transform is singular");
    for (int c = 0; c < 4; ++c)
        for (int k = 0; k < 4; ++k) out->minv[c * 4 + k] = (float)inv[c][k];
    for (int a = 0; a < 3; ++a) {
        int lo = (int)std::floor(mn[a]);
        int hi = (int)std::ceil(mx[a]);
        out->shift[a] = lo;
        out->groups[a] = std::max((hi - lo) + 3, 0) / 4;
        out->aabb_min[a] = mn[a];
        out->aabb_max[a] = mx[a];
    }
    out->shift[3] = 0;
    return LUMX_OK;
}

// blockify_mesh for one placed model (internal_render.cpp:1029-1076).
static int blockify(lumx_world* w, const lumx_placement& p) {
    int mn[3], mx[3];
    for (int a = 0; a < 3; ++a) {
        mn[a] = ((int)p.aabb_min[a] - 1) / 16;
        mx[a] = ((int)p.aabb_max[a] + 1) / 16;
        mn[a] = std::min(std::max(mn[a], 0), w->desc.size[a] - 1);
        mx[a] = std::min(std::max(mx[a], 0), w->desc.size[a] - 1);
    }
    const int wx = w->desc.size[0], wy = w->desc.size[1];
    for (int z = mn[2]; z <= mx[2]; ++z)
        for (int y = mn[1]; y <= mx[1]; ++y)
            for (int x = mn[0]; x <= mx[0]; ++x) {
                const int idx = x + wx * (y + wy * z);
                int16_t& b = w->current[idx];
                if (b < w->desc.static_blocks) {
                    if (w->counter >= w->desc.palette_slots)
                        return fail(LUMX_EOVERFLOW, "block palette full (%d slots)", w->desc.palette_slots);
                    w->copy_src[w->counter] = b;
                    w->slot_block[w->counter] = idx;
                    b = (int16_t)w->counter;
                    w->counter++;
                }
            }
    return LUMX_OK;
}

// map.comp for each placed model, in order, over the current device world.
// Models go in layers: a model's layer is one more than that of the last
// earlier model whose write box (the voxels its dispatch covers) intersects
// its own.  Models of one layer write disjoint palette bytes, so a layer is
// one launch (in groups of kMapGroup) and the result equals submission-order
// mapping.  An origin world holding dynamic ids could alias two blocks to one
// slot; it maps one model per launch.  A model whose dispatch box reaches
// below 0 on some axis writes through aliased coordinates into block 0
// (map.hip), outside its box: it gets a layer of its own, mapped as one launch
// per sign class in increasing class order.
extern "C++" {
template <class Emit>
static int plan_map_launches(lumx_world* w, int n, const lumx_model_t* models, const lumx_placement* pl,
                             Emit&& emit) {
    const int st = w->desc.static_blocks;
    SmallBuf<MapArgs, kNoAllocModels> args(std::max(n, 1));
    SmallBuf<int, kNoAllocModels> layer(std::max(n, 1), 0);
    SmallBuf<char, kNoAllocModels> negative(std::max(n, 1), 0);
    if (w->origin_ids_dirty) {
        w->origin_has_dynamic = false;
        for (int16_t b : w->origin)
            if (b >= st) {
                w->origin_has_dynamic = true;
                break;
            }
        w->origin_ids_dirty = false;
    }
    const bool aliasing = w->origin_has_dynamic;
    for (int i = 0; i < n; ++i) {
        const lumx_placement& p = pl[i];
        MapArgs& a = args[i];
        a.blocks = w->blocks;
        a.palette = w->palette;
        a.model = models[i]->voxels;
        a.wx = w->desc.size[0]; a.wy = w->desc.size[1]; a.wz = w->desc.size[2];
        a.palette_slots = w->desc.palette_slots;
        a.static_count = st;
        a.mx = models[i]->size[0]; a.my = models[i]->size[1]; a.mz = models[i]->size[2];
        a.nx = p.groups[0] * 4; a.ny = p.groups[1] * 4; a.nz = p.groups[2] * 4;
        a.shift_x = p.shift[0]; a.shift_y = p.shift[1]; a.shift_z = p.shift[2];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 4; ++c) a.m[r * 4 + c] = p.minv[c * 4 + r];
        a.cls = 7;
        a.tiles_x = (a.nx + 15) / 16;
        a.tiles_y = a.ny / 4;
        if (map_tiles(a) >= (1ll << 28))
            return fail(LUMX_EINVAL, "model %d dispatch too large", i);
        a.div_tiles_x = make_divisor((uint32_t)std::max(a.tiles_x, 1));
        a.div_tiles_y = make_divisor((uint32_t)std::max(a.tiles_y, 1));
        negative[i] = a.shift_x < 0 || a.shift_y < 0 || a.shift_z < 0;
        layer[i] = 0;
        for (int j = 0; j < i; ++j) {
            const MapArgs& b = args[j];
            const bool overlap = aliasing || negative[i] || negative[j] ||
                (a.shift_x < b.shift_x + b.nx && b.shift_x < a.shift_x + a.nx && a.shift_y < b.shift_y + b.ny &&
                 b.shift_y < a.shift_y + a.ny && a.shift_z < b.shift_z + b.nz && b.shift_z < a.shift_z + a.nz);
            if (overlap) layer[i] = std::max(layer[i], layer[j] + 1);
        }
    }
    const int n_layers = n ? *std::max_element(&layer[0], &layer[0] + n) + 1 : 0;
    SmallBuf<MapArgs, kNoAllocModels> group(std::max(n, 1));
    for (int l = 0; l < n_layers; ++l) {
        int g = 0;
        for (int i = 0; i < n; ++i) {
            if (layer[i] != l) continue;
            const MapArgs& a = args[i];
            const long long groups = map_tiles(a);
            if (groups <= 0) continue;
            if (negative[i]) {
                // Alone in its layer: one launch per sign class present.
                const int lo[3] = {a.shift_x, a.shift_y, a.shift_z};
                const int hi[3] = {a.shift_x + a.nx, a.shift_y + a.ny, a.shift_z + a.nz};
                for (int cls = 0; cls < 8; ++cls) {
                    bool present = true;
                    for (int k = 0; k < 3; ++k) present &= (cls >> k & 1) ? hi[k] > 0 : lo[k] < 0;
                    if (!present) continue;
                    MapArgs one = a;
                    one.cls = cls;
                    int rc = emit(&one, 1);
                    if (rc) return rc;
                }
                continue;
            }
            group[g++] = a;
        }
        if (g > 0) {
            int rc = emit(&group[0], g);
            if (rc) return rc;
        }
    }
    return LUMX_OK;
}
}  // extern "C++"

// The launches of plan_map_launches in order, up to kMapGroup models per
// launch (descriptors as the kernel argument).  Allocates nothing.
static int launch_models(lumx_world* w, int n, const lumx_model_t* models, const lumx_placement* pl,
                         hipStream_t s) {
    return plan_map_launches(w, n, models, pl, [&](const MapArgs* group, int count) -> int {
        MapGroup g;
        g.n = 0;
        g.begin[0] = 0;
        for (int k = 0; k <= count; ++k) {
            if (k < count) {
                const MapArgs& a = group[k];
                const long long groups = map_tiles(a);
                if ((long long)g.begin[g.n] + groups > (1ll << 31) - 1)
                    return fail(LUMX_EINVAL, "map launch too large");
                g.m[g.n] = a;
                g.begin[g.n + 1] = g.begin[g.n] + (int)groups;
                g.n++;
            }
            if (g.n == kMapGroup || (k == count && g.n > 0)) {
                HIP_TRY(launch_map_group(g, s));
                g.n = 0;
                g.begin[0] = 0;
            }
        }
        return LUMX_OK;
    });
}

// Block box of blockify_mesh for a placed model (internal_render.cpp:1032-1037).
static void block_box(const lumx_world* w, const lumx_placement& p, int mn[3], int mx[3]) {
    for (int a = 0; a < 3; ++a) {
        mn[a] = ((int)p.aabb_min[a] - 1) / 16;
        mx[a] = ((int)p.aabb_max[a] + 1) / 16;
        mn[a] = std::min(std::max(mn[a], 0), w->desc.size[a] - 1);
        mx[a] = std::min(std::max(mx[a], 0), w->desc.size[a] - 1);
    }
}

// Device map: blockify, exec_copies and map.comp on the device.  With
// `fixed` the scratch is the device's fixed array and the model boxes reach it
// as a kernel argument, so nothing is allocated (host or device).
static int map_device(lumx_world_t w, int n, const lumx_model_t* models, const lumx_placement* pl, void* stream,
                      bool fixed) {
    if (!w || n < 0 || (n > 0 && (!models || !pl))) return fail(LUMX_EINVAL, "bad argument");
    if (fixed && n > kNoAllocModels)
        return fail(LUMX_EINVAL, "no-allocation map: %d models (at most %d)", n, kNoAllocModels);
    for (int i = 0; i < n; ++i)
        if (!models[i] || models[i]->ctx != w->ctx) return fail(LUMX_EINVAL, "model %d invalid", i);
    REQUIRE_RESIDENT(w);
    HIP_TRY(hipSetDevice(w->ctx->device));
    hipStream_t s = pick_stream(w->ctx, stream);
    const bool capturing = stream_capturing(s);
    const int st = w->desc.static_blocks, slots = w->desc.palette_slots;
    // Model boxes and their prefix over the concatenated element sequence.
    SmallBuf<int, kNoAllocModels> first(std::max(n, 1), 0);
    SmallBuf<int, 8 * kNoAllocModels> box(8 * std::max(n, 1), 0);  // per model: int4 min, int4 extent
    long long total = 0;
    for (int i = 0; i < n; ++i) {
        int mn[3], mx[3];
        block_box(w, pl[i], mn, mx);
        first[i] = (int)total;
        for (int a = 0; a < 3; ++a) {
            box[8 * i + a] = mn[a];
            box[8 * i + 4 + a] = mx[a] - mn[a] + 1;
        }
        box[8 * i + 3] = box[8 * i + 7] = 0;
        total += (long long)box[8 * i + 4] * box[8 * i + 5] * box[8 * i + 6];
    }
    if (total > (1ll << 30)) return fail(LUMX_EINVAL, "model boxes cover too many blocks (%lld)", total);
    const int n_el = (int)total;
    const size_t tiles = blockify_tiles(n_el);
    auto align = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t off_box = align((size_t)std::max(n, 1) * 4);
    const size_t off_bits = off_box + align((size_t)std::max(n, 1) * 32);
    const size_t off_tiles = off_bits + align(tiles * (2048 / 8));
    const size_t need = off_tiles + align(std::max<size_t>(tiles, 1) * 4);
    if (fixed && need > kNoAllocScratchBytes)
        return fail(LUMX_EINVAL, "no-allocation map: model boxes cover %d blocks, scratch of %zu bytes holds "
                                 "at most %zu", n_el, need, kNoAllocScratchBytes);
    if (fixed && capturing && w->origin_dirty)
        return fail(LUMX_EINVAL, "no-allocation map inside a capture: the origin world changed since its last "
                                 "upload; map once outside the capture first");
    if (fixed && !w->ctx->noalloc_scratch) return fail(LUMX_EDEVICE, "no-allocation scratch not resolved");
    if (!fixed && need > w->blockify_scratch_bytes) {
        if (capturing)
            return fail(LUMX_EINVAL, "device map scratch too small inside a capture: run the same map once "
                                     "outside the capture first");
        HIP_TRY(hipStreamSynchronize(s));
        if (w->blockify_scratch) hipFree(w->blockify_scratch);
        w->blockify_scratch = nullptr;
        w->blockify_scratch_bytes = 0;
        HIP_TRY(hipMalloc(&w->blockify_scratch, need));
        w->blockify_scratch_bytes = need;
    }
    // Staging of the model boxes (and of the origin world when it changed):
    // pinned, shared by eager maps once the previous map's copies are done; a
    // map recorded into a graph gets its own, read on every replay.
    const size_t desc_bytes = off_bits;
    const size_t origin_bytes = w->origin_dirty ? (size_t)w->n_blocks * 2 : 0;
    char* base = fixed ? (char*)w->ctx->noalloc_scratch : (char*)w->blockify_scratch;
    char* stage = nullptr;
    if (fixed) {
        // Boxes as a kernel argument; the origin world (when it changed)
        // through the world's own pinned staging.
        BlockifyDesc d{};
        d.n = n;
        for (int i = 0; i < n; ++i) {
            d.first[i] = first[i];
            d.boxes[2 * i] = make_int4(box[8 * i], box[8 * i + 1], box[8 * i + 2], 0);
            d.boxes[2 * i + 1] = make_int4(box[8 * i + 4], box[8 * i + 5], box[8 * i + 6], 0);
        }
        if (!capturing) HIP_TRY(hipEventRecord(w->ev[1][0], s));
        HIP_TRY(launch_blockify_desc(d, (int4*)(base + off_box), (int*)base, s));
        if (origin_bytes) {  // not capturing (checked above)
            if (w->stage_pending) HIP_TRY(hipEventSynchronize(w->stage_done));
            w->stage_pending = false;
            std::memcpy(w->stage_blocks, w->origin.data(), origin_bytes);
            HIP_TRY(hipMemcpyAsync(w->origin_dev, w->stage_blocks, origin_bytes, hipMemcpyHostToDevice, s));
            w->origin_dirty = false;
        }
    } else if (capturing) {
        void* p = nullptr;
        HIP_TRY(hipHostMalloc(&p, desc_bytes + origin_bytes + 16, hipHostMallocDefault));
        w->graph_stage.push_back(p);
        stage = (char*)p;
    } else {
        if (w->stage_pending) HIP_TRY(hipEventSynchronize(w->stage_done));
        w->stage_pending = false;
        if (desc_bytes > w->stage_desc_bytes) {
            if (w->stage_desc) hipHostFree(w->stage_desc);
            w->stage_desc = nullptr;
            w->stage_desc_bytes = 0;
            HIP_TRY(hipHostMalloc(&w->stage_desc, desc_bytes, hipHostMallocDefault));
            w->stage_desc_bytes = desc_bytes;
        }
        stage = (char*)w->stage_desc;
    }
    if (!fixed) {
        std::memcpy(stage, &first[0], (size_t)std::max(n, 1) * 4);
        std::memcpy(stage + off_box, &box[0], (size_t)std::max(n, 1) * 32);
        if (!capturing) HIP_TRY(hipEventRecord(w->ev[1][0], s));
        HIP_TRY(hipMemcpyAsync(w->blockify_scratch, stage, desc_bytes, hipMemcpyHostToDevice, s));
    }
    if (!fixed && origin_bytes) {
        int16_t* so = capturing ? (int16_t*)(stage + desc_bytes) : w->stage_blocks;
        std::memcpy(so, w->origin.data(), origin_bytes);
        HIP_TRY(hipMemcpyAsync(w->origin_dev, so, origin_bytes, hipMemcpyHostToDevice, s));
        // A recorded upload runs only when the graph does: eager maps keep uploading.
        if (!capturing) w->origin_dirty = false;
    }
    // start_blockify: the current world is the origin world.
    HIP_TRY(hipMemcpyAsync(w->blocks, w->origin_dev, (size_t)w->n_blocks * 2, hipMemcpyDeviceToDevice, s));
    BlockifyArgs a{};
    a.origin = w->origin_dev;
    a.blocks = w->blocks;
    a.copy_src = w->copy_src_dev;
    a.palette = w->palette;
    a.claim = w->sel_claim;
    a.claim_idle = select_claim_idle();
    a.wx = w->desc.size[0];
    a.wy = w->desc.size[1];
    a.wz = w->desc.size[2];
    a.static_count = st;
    a.palette_slots = slots;
    a.n_models = n;
    a.n_elements = n_el;
    a.n_tiles = (int)tiles;
    a.first = (const int*)base;
    a.boxes = (const int4*)(base + off_box);
    a.bits = (unsigned long long*)(base + off_bits);
    a.tile_counts = (int*)(base + off_tiles);
    a.state = w->blockify_state;
    // Slots that may need copying: up to the larger of this map's and the
    // previous map's counters.
    const int prev_bound = w->host_stale ? w->dev_counter_bound : std::max(w->counter, w->prev_counter);
    const int bound = (int)std::min<long long>(slots, std::max<long long>(st + total, prev_bound));
    HIP_TRY(launch_blockify(a, bound - st, s));
    HIP_TRY(launch_block_bits(w->blocks, w->n_blocks, w->empty_bits, s));
    int rc = launch_models(w, n, models, pl, s);
    if (rc) return rc;
    if (bound > st) HIP_TRY(launch_palette_bits(w->palette, st, bound - st, w->pal_bits, s));
    w->host_stale = true;
    w->dev_counter_bound = (int)std::min<long long>(slots, st + total);
    if (capturing) return LUMX_OK;
    HIP_TRY(hipEventRecord(w->ev[1][1], s));
    w->ev_valid[1] = true;
    HIP_TRY(hipEventRecord(w->stage_done, s));
    w->stage_pending = true;
    return LUMX_OK;
}

int lumx_map_models_device(lumx_world_t w, int n, const lumx_model_t* models, const lumx_placement* pl,
                           void* stream) {
    return map_device(w, n, models, pl, stream, false);
}

int lumx_map_models_device_noalloc(lumx_world_t w, int n, const lumx_model_t* models, const lumx_placement* pl,
                                   void* stream) {
    return map_device(w, n, models, pl, stream, true);
}

extern "C++" {
int lumx::map_check(const lumx_world* w, int n, const lumx_model_t* models, const lumx_placement* pl) {
    if (!w || n < 0 || (n > 0 && (!models || !pl))) return fail(LUMX_EINVAL, "bad argument");
    for (int i = 0; i < n; ++i)
        if (!models[i] || models[i]->ctx != w->ctx) return fail(LUMX_EINVAL, "model %d invalid", i);
    REQUIRE_RESIDENT(w);
    return LUMX_OK;
}

// start_blockify + blockify_mesh of each model (internal_render.cpp:1006-1092)
// on the host world; on failure the host world is back at its origin.
int lumx::map_prepare(lumx_world* w, int n, const lumx_placement* pl) {
    HIP_TRY(hipSetDevice(w->ctx->device));
    int rc0 = sync_host_state(w);
    if (rc0) return rc0;
    w->dev_overflow = false;
    w->current = w->origin;
    w->counter = w->desc.static_blocks;
    for (int i = 0; i < n; ++i) {
        int rc = blockify(w, pl[i]);
        if (rc) {
            w->current = w->origin;
            w->counter = w->desc.static_blocks;
            return rc;
        }
    }
    return LUMX_OK;
}

int lumx::map_plan(lumx_world* w, int n, const lumx_model_t* models, const lumx_placement* pl,
                   std::vector<MapArgs>* args, std::vector<int>* ends) {
    return plan_map_launches(w, n, models, pl, [&](const MapArgs* group, int count) -> int {
        args->insert(args->end(), group, group + count);
        ends->push_back((int)args->size());
        return LUMX_OK;
    });
}
}  // extern "C++"

int lumx_map_models_placed(lumx_world_t w, int n, const lumx_model_t* models, const lumx_placement* pl,
                           void* stream) {
    int rc0 = map_check(w, n, models, pl);
    if (rc0) return rc0;
    rc0 = map_prepare(w, n, pl);
    if (rc0) return rc0;
    hipStream_t s = pick_stream(w->ctx, stream);
    // end_blockify + exec_copies + map.
    const bool capturing = stream_capturing(s);
    if (!capturing) HIP_TRY(hipEventRecord(w->ev[1][0], s));
    int rc = commit_blocks(w, s);
    if (rc) return rc;
    rc = launch_models(w, n, models, pl, s);
    if (rc) return rc;
    rc = commit_palette_bits(w, s);
    if (rc) return rc;
    if (capturing) return LUMX_OK;  // a graph re-reads the pinned staging on replay
    HIP_TRY(hipEventRecord(w->ev[1][1], s));
    w->ev_valid[1] = true;
    HIP_TRY(hipEventRecord(w->stage_done, s));
    w->stage_pending = true;
    return LUMX_OK;
}

int lumx_map_models(lumx_world_t w, int n, const lumx_model_t* models, const lumx_transform* t, void* stream) {
    if (!w || n < 0 || (n > 0 && (!models || !t))) return fail(LUMX_EINVAL, "bad argument");
    std::vector<lumx_placement> pl(n);
    for (int i = 0; i < n; ++i) {
        if (!models[i]) return fail(LUMX_EINVAL, "model %d is null", i);
        int rc = lumx_model_placement(&t[i], models[i]->size, &pl[i]);
        if (rc) return rc;
    }
    return lumx_map_models_placed(w, n, models, pl.data(), stream);
}

// ---- radiance ----
extern "C++" {
int lumx::radiance_check(const lumx_world* w, const int32_t* req, int n_req, const lumx_radiance_params* p) {
    if (!w || !p) return fail(LUMX_EINVAL, "null argument");
    REQUIRE_RESIDENT(w);
    int rc = radiance_params_check(p);
    if (rc) return rc;
    if (req && n_req < 0) return fail(LUMX_EINVAL, "n_requests=%d invalid", n_req);
    // A subset update stages probe i's texel in slot i of the free field, which
    // holds one texel per probe: the reference's request buffer likewise holds
    // one entry per probe (internal_render.cpp:198-200).
    if (req && n_req / p->size > w->n_blocks)
        return fail(LUMX_EINVAL, "%d requests at stride %d dispatch more probes than the world's %d", n_req,
                    p->size, w->n_blocks);
    return LUMX_OK;
}

int lumx::radiance_params_check(const lumx_radiance_params* p) {
    if (!valid_rays(p->rays))
        return fail(LUMX_EINVAL, "rays=%d: must be a power of two <= 64 or a multiple of 64 <= 1024", p->rays);
    if (p->size < 1 || p->shift < 0 || p->shift >= p->size)
        return fail(LUMX_EINVAL, "size=%d shift=%d invalid", p->size, p->shift);
    if (!(p->max_dist > 0.0f && p->max_dist <= 65536.0f) || p->max_steps < 0)
        return fail(LUMX_EINVAL, "max_dist must be in (0, 65536] and max_steps >= 0");
    return LUMX_OK;
}
}  // extern "C++"

static int check_tuning(const lumx_world* w, const lumx_radiance_tuning& t, int rays) {
    if (t.variant < 0 || t.variant > 3) return fail(LUMX_EINVAL, "tuning variant %d invalid", t.variant);
    if (t.rays_per_lane != 0 && t.rays_per_lane != 1 && t.rays_per_lane != 2 && t.rays_per_lane != 4)
        return fail(LUMX_EINVAL, "tuning rays_per_lane %d is not 1, 2 or 4", t.rays_per_lane);
    if (t.compact < -1 || t.compact > 1) return fail(LUMX_EINVAL, "tuning compact %d is not -1, 0 or 1", t.compact);
    if (t.variant == 3 && (t.probes_per_batch < 1 || t.probes_per_batch > 4))
        return fail(LUMX_EINVAL, "tuning probes_per_batch %d invalid for the refill variant (1..4)",
                    t.probes_per_batch);
    if (t.variant != 1) return LUMX_OK;
    if (t.threads < 64 || t.threads > 1024 || t.threads % 64)
        return fail(LUMX_EINVAL, "tuning threads %d invalid", t.threads);
    int P = t.probes_per_batch;
    if (P < 1 || P > 64 || (P & (P - 1))) return fail(LUMX_EINVAL, "tuning probes_per_batch %d invalid", P);
    if (t.wg_per_cu < 0) return fail(LUMX_EINVAL, "tuning wg_per_cu invalid");
    (void)w;
    (void)rays;
    return LUMX_OK;
}

extern "C++" {
int lumx::radiance_args(lumx_world* w, const int32_t* req, int n_req, const lumx_radiance_params* p,
                        RadianceArgs* out) {
    int n_total = req ? n_req : w->n_blocks;
    int n_probes = n_total / p->size;
    const float4* fib = nullptr;
    int rc = get_fib(w->ctx, p->rays, &fib);
    if (rc) return rc;
    RadianceArgs& a = *out;
    a = RadianceArgs{};
    a.world = view_of(w);
    a.field_in = w->field[w->cur];
    a.field_out = w->field[w->cur ^ 1];
    a.requests = (const int4*)req;
    a.fib = fib;
    a.rays = p->rays;
    a.log2_rays = (p->rays & (p->rays - 1)) == 0 ? __builtin_ctz((unsigned)p->rays) : -1;
    a.n_probes = n_probes;
    a.time_seed = p->time_seed;
    a.size = p->size;
    a.shift = p->shift;
    a.alpha = p->alpha;
    a.max_dist = p->max_dist;
    a.max_steps = p->max_steps;
    a.empty_bits = w->empty_bits;
    a.pal_bits = w->pal_bits;
    a.eb_words = (w->n_blocks + 31) / 32;
    a.pb_words = w->desc.palette_slots * 128;
    a.batch_counter = w->batch_counter;
    a.staged = (req != nullptr || p->size != 1) ? 1 : 0;
    a.div_x = make_divisor((uint32_t)w->desc.size[0]);
    a.div_xy = make_divisor((uint32_t)w->desc.size[0] * (uint32_t)w->desc.size[1]);
    return LUMX_OK;
}

void lumx::compact_args(const lumx_world* w, RadianceArgs* a) {
    a->compact = 1;
    a->list_stride = (int)w->texel_bytes;
    a->tiles = w->upd_tiles;
    a->n_tiles = (int)select_tiles(w->n_blocks);
}
}  // extern "C++"

// Launch one radiance update from field[cur] into field[cur ^ 1] with tuning t.
// A full update (every probe) writes field[cur ^ 1], which then becomes
// current.  A subset update writes probe i's new texel to field[cur ^ 1][i]
// (free scratch between updates) and the scatter kernel then stores it in
// field[cur] in place, so only the requested probes move and the current
// field stays current.  apply = false (the autotuner) leaves out the scatter.
static int radiance_launch(lumx_world* w, const int32_t* req, int n_req, const lumx_radiance_params* p,
                           int32_t* keys, int32_t* steps, bool traced, hipStream_t s,
                           const lumx_radiance_tuning& tune, const int32_t* count_dev = nullptr,
                           bool apply = true) {
    RadianceArgs a;
    int rc = radiance_args(w, req, n_req, p, &a);
    if (rc) return rc;
    a.n_requests_dev = count_dev;
    a.ray_keys = keys;
    a.ray_steps = steps;
    // A subset update claims its probes right before it runs (after every
    // check that can fail); the scatter that applies it consumes (and resets)
    // the claims.
    if (a.staged && apply) a.claim = w->upd_claim;
    if (tune.variant == 3 && !traced) {
        if (p->rays != 64) return fail(LUMX_EINVAL, "the refill variant takes 64 rays per probe");
        HIP_TRY(launch_claim(a, s));
        HIP_TRY(launch_radiance_refill(a, w->desc.radiance_format, tune.probes_per_batch, s));
        if (a.staged && apply) HIP_TRY(launch_scatter(a, (int)w->texel_bytes, s));
        return LUMX_OK;
    }
    if (tune.variant == 0 || tune.variant == 2 || tune.variant == 3) {
        // (a traced launch of the refill variant traces with the lane kernel: the same results)
        // A large subset update runs over its claimed probes in texel order
        // (compacted): the same results, better locality.
        // One world whose field fits a cache gains nothing from texel order,
        // so automatic compaction waits for larger fields.
        const bool compact = a.staged && apply && !traced && tune.compact >= 0 &&
                             (tune.compact > 0 || (compact_pays(a.n_probes, w->n_blocks) &&
                                                   (size_t)w->n_blocks * w->texel_bytes > kCompactFieldBytes));
        if (compact) compact_args(w, &a);
        HIP_TRY(launch_claim(a, s));
        if (compact) HIP_TRY(launch_compact(a, s));
        HIP_TRY(launch_radiance(a, w->desc.radiance_format, traced, tune.variant == 2, tune.rays_per_lane, s));
        if (a.staged && apply)
            HIP_TRY(compact ? launch_compact_scatter(a, (int)w->texel_bytes, s)
                            : launch_scatter(a, (int)w->texel_bytes, s));
        return LUMX_OK;
    }
    RadianceTune t;
    t.threads = tune.threads;
    t.probes_per_batch = tune.probes_per_batch;
    t.wg_per_cu = tune.wg_per_cu;
    t.eb_lds = 0;
    t.pb_lds = 0;
    // Stage what fits the LDS budget: block bits first, then palette bits.
    size_t base = radiance_tiled_lds_bytes(a, t);
    if (tune.lds_block_bits && base + 4 * (size_t)a.eb_words <= kLdsBudget) {
        t.eb_lds = 1;
        base += 4 * (size_t)a.eb_words;
    }
    const int used_slots = w->host_stale ? w->dev_counter_bound : std::max(w->counter, w->prev_counter);
    if (tune.lds_palette_bits && base + 4 * (size_t)used_slots * 128 <= kLdsBudget) {
        t.pb_lds = 1;
        a.pb_words = used_slots * 128;
    }
    if (base > kLdsBudget) return fail(LUMX_EINVAL, "tiled radiance batch does not fit LDS (%zu B)", base);
    HIP_TRY(launch_claim(a, s));
    HIP_TRY(launch_radiance_tiled(a, w->desc.radiance_format, traced, t, w->ctx->n_cu, s));
    if (a.staged && apply) HIP_TRY(launch_scatter(a, (int)w->texel_bytes, s));
    return LUMX_OK;
}

static int radiance_impl(lumx_world* w, const int32_t* req, int n_req, const lumx_radiance_params* p,
                         int32_t* keys, int32_t* steps, bool traced, void* stream,
                         const int32_t* count_dev = nullptr) {
    int rc = radiance_check(w, req, n_req, p);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(w->ctx->device));
    hipStream_t s = pick_stream(w->ctx, stream);
    const bool timed = !stream_capturing(s);
    if (timed) HIP_TRY(hipEventRecord(w->ev[0][0], s));
    rc = radiance_launch(w, req, n_req, p, keys, steps, traced, s, w->tune, count_dev);
    if (rc) return rc;
    if (timed) {
        HIP_TRY(hipEventRecord(w->ev[0][1], s));
        w->ev_valid[0] = true;
    }
    if (req == nullptr && p->size == 1) w->cur ^= 1;  // a full update wrote the other field
    return LUMX_OK;
}

int lumx_world_set_tuning(lumx_world_t w, const lumx_radiance_tuning* t) {
    if (!w || !t) return fail(LUMX_EINVAL, "null argument");
    int rc = check_tuning(w, *t, 64);
    if (rc) return rc;
    w->tune = *t;
    return LUMX_OK;
}

int lumx_world_get_tuning(lumx_world_t w, lumx_radiance_tuning* out) {
    if (!w || !out) return fail(LUMX_EINVAL, "null argument");
    *out = w->tune;
    return LUMX_OK;
}

int lumx_autotune_radiance(lumx_world_t w, const lumx_radiance_params* p, int reps, lumx_radiance_tuning* best,
                           float* best_ms) {
    int rc = radiance_check(w, nullptr, 0, p);
    if (rc) return rc;
    if (reps < 1) reps = 1;
    HIP_TRY(hipSetDevice(w->ctx->device));
    hipStream_t s = w->ctx->stream;
    std::vector<lumx_radiance_tuning> cands;
    lumx_radiance_tuning c0{};
    c0.variant = 0;
    cands.push_back(c0);
    for (int k : {2, 4})
        if (rays_per_lane_for(p->rays, k) == k) {
            c0.rays_per_lane = k;
            cands.push_back(c0);
        }
    c0.rays_per_lane = 0;
    c0.variant = 2;
    cands.push_back(c0);
    if (p->rays == 64)
        for (int P : {1, 2, 4}) {
            lumx_radiance_tuning c{};
            c.variant = 3;
            c.probes_per_batch = P;
            cands.push_back(c);
        }
    const int threads[] = {256, 512, 1024};
    const int probes[] = {4, 8, 16, 32};
    for (int th : threads)
        for (int P : probes)
            for (int pb = 0; pb <= 1; ++pb) {
                if ((long long)P * p->rays > 4096) continue;
                lumx_radiance_tuning c{};
                c.variant = 1;
                c.threads = th;
                c.probes_per_batch = P;
                c.lds_block_bits = 1;
                c.lds_palette_bits = pb;
                cands.push_back(c);
            }
    EventPair ev;
    HIP_TRY(hipEventCreate(&ev.a));
    HIP_TRY(hipEventCreate(&ev.b));
    hipEvent_t e0 = ev.a, e1 = ev.b;
    float best_t = 1e30f;
    lumx_radiance_tuning best_c = w->tune;
    std::vector<float> cand_ms(cands.size(), 1e30f);
    // Two passes over the candidates, keeping each one's faster pass, so the
    // clock ramp-up at the start does not count against the first ones.
    for (int pass = 0; pass < 2; ++pass)
        for (size_t ci = 0; ci < cands.size(); ++ci) {
            const lumx_radiance_tuning& c = cands[ci];
            // Every trial writes field[cur ^ 1] only; the current field is unchanged.
            rc = radiance_launch(w, nullptr, 0, p, nullptr, nullptr, false, s, c, nullptr, false);
            if (rc) continue;  // shape does not fit this world
            HIP_TRY(hipEventRecord(e0, s));
            for (int r = 0; r < reps; ++r) {
                rc = radiance_launch(w, nullptr, 0, p, nullptr, nullptr, false, s, c, nullptr, false);
                if (rc) break;
            }
            HIP_TRY(hipEventRecord(e1, s));
            HIP_TRY(hipEventSynchronize(e1));
            if (rc) continue;
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            cand_ms[ci] = std::min(cand_ms[ci], ms / (float)reps);
        }
    for (size_t ci = 0; ci < cands.size(); ++ci)
        if (cand_ms[ci] < best_t) {
            best_t = cand_ms[ci];
            best_c = cands[ci];
        }
    w->tune = best_c;
    if (best) *best = best_c;
    if (best_ms) *best_ms = best_t;
    clear_error();
    return LUMX_OK;
}

int lumx_radiance_update(lumx_world_t w, const int32_t* req, int n, const lumx_radiance_params* p, void* stream) {
    return radiance_impl(w, req, n, p, nullptr, nullptr, false, stream);
}

int lumx_radiance_update_traced(lumx_world_t w, const int32_t* req, int n, const lumx_radiance_params* p,
                                int32_t* keys, int32_t* steps, void* stream) {
    return radiance_impl(w, req, n, p, keys, steps, true, stream);
}

int lumx_radiance_update_counted(lumx_world_t w, const int32_t* req, const int32_t* count_dev, int max_requests,
                                 const lumx_radiance_params* p, void* stream) {
    if (!req || !count_dev) return fail(LUMX_EINVAL, "counted update needs a request list and its device count");
    return radiance_impl(w, req, max_requests, p, nullptr, nullptr, false, stream, count_dev);
}

int lumx_select_probes_device(lumx_world_t w, const int32_t* special_dev, int n_special, int32_t* out_dev,
                              int32_t* count_dev, void* stream) {
    if (!w || !out_dev || !count_dev || n_special < 0 || (n_special > 0 && !special_dev))
        return fail(LUMX_EINVAL, "bad argument");
    if (n_special > (1 << 30)) return fail(LUMX_EINVAL, "too many special requests");
    REQUIRE_RESIDENT(w);
    HIP_TRY(hipSetDevice(w->ctx->device));
    hipStream_t s = pick_stream(w->ctx, stream);
    lumx::SelectArgs a{};
    a.blocks = w->blocks;
    a.wx = w->desc.size[0];
    a.wy = w->desc.size[1];
    a.wz = w->desc.size[2];
    a.n = w->n_blocks;
    a.n_tiles = (int)lumx::select_tiles(w->n_blocks);
    a.bits = w->sel_bits;
    a.tile_counts = w->sel_tiles;
    a.special = (const int4*)special_dev;
    a.n_special = n_special;
    a.out = (int4*)out_dev;
    a.count = count_dev;
    HIP_TRY(lumx::launch_select(a, s));
    return LUMX_OK;
}

int lumx_world_shift_radiance(lumx_world_t w, int dx, int dy, int dz, void* stream) {
    if (!w) return fail(LUMX_EINVAL, "null world");
    REQUIRE_RESIDENT(w);
    const int d[3] = {dx, dy, dz};
    // A shift of a whole world size or more leaves no intersection: nothing moves.
    for (int k = 0; k < 3; ++k)
        if (std::llabs((long long)d[k]) >= w->desc.size[k]) return LUMX_OK;
    if (dx == 0 && dy == 0 && dz == 0) return LUMX_OK;
    HIP_TRY(hipSetDevice(w->ctx->device));
    hipStream_t s = pick_stream(w->ctx, stream);
    lumx::ShiftArgs a{};
    a.in = w->field[w->cur];
    a.out = w->field[w->cur ^ 1];
    a.wx = w->desc.size[0];
    a.wy = w->desc.size[1];
    a.wz = w->desc.size[2];
    a.n = w->n_blocks;
    for (int k = 0; k < 3; ++k) {
        a.shift[k] = d[k];
        a.dst[k] = d[k] >= 0 ? 0 : -d[k];
        a.extent[k] = w->desc.size[k] - (int)std::llabs((long long)d[k]);
    }
    HIP_TRY(lumx::launch_shift(a, (int)w->texel_bytes, s));
    w->cur ^= 1;
    return LUMX_OK;
}

int lumx_select_probes(lumx_world_t w, const int32_t* special, int n_special, int32_t* out, int* n_out) {
    if (!w || !out || !n_out || n_special < 0 || (n_special > 0 && !special)) return fail(LUMX_EINVAL, "bad argument");
    int rc0 = sync_host_state(w);
    if (rc0) return rc0;
    const int wx = w->desc.size[0], wy = w->desc.size[1], wz = w->desc.size[2];
    const std::vector<int16_t>& cw = w->current;
    std::vector<uint8_t> set((size_t)w->n_blocks, 0);
    int n = 0;
    for (int z = 0; z < wz; ++z)
        for (int y = 0; y < wy; ++y)
            for (int x = 0; x < wx; ++x) {
                int sum = 0;
                for (int dz = -1; dz <= 1; ++dz) {
                    int tz = std::min(std::max(z + dz, 0), wz - 1);
                    for (int dy = -1; dy <= 1; ++dy) {
                        int ty = std::min(std::max(y + dy, 0), wy - 1);
                        for (int dx = -1; dx <= 1; ++dx) {
                            int tx = std::min(std::max(x + dx, 0), wx - 1);
                            sum += cw[tx + wx * (ty + wy * tz)];
                        }
                    }
                }
                if (sum > 0) {
                    int32_t* o = out + 4 * n;
                    o[0] = x; o[1] = y; o[2] = z; o[3] = 0;
                    set[x + (size_t)wx * (y + (size_t)wy * z)] = 1;
                    ++n;
                }
            }
    // Every special whose probe the grid did not select is appended, repeats
    // included (the reference marks only grid probes in `set`); the list ends
    // at one entry per probe, the capacity of the reference's request buffer.
    for (int i = 0; i < n_special; ++i) {
        int x = special[4 * i], y = special[4 * i + 1], z = special[4 * i + 2];
        if (!in_world(w, x, y, z)) return fail(LUMX_EINVAL, "special request %d outside the world", i);
        size_t idx = x + (size_t)wx * (y + (size_t)wy * z);
        if (!set[idx] && n < w->n_blocks) {
            int32_t* o = out + 4 * n;
            o[0] = x; o[1] = y; o[2] = z; o[3] = special[4 * i + 3];
            ++n;
        }
    }
    *n_out = n;
    return LUMX_OK;
}

// ---- glossy ----
extern "C++" {
void lumx::glossy_args(lumx_world* w, const uint8_t* matnorm, const float* depth, int width, int height,
                       const lumx_camera* cam, float max_dist, int max_steps, float* out, int32_t* keys,
                       GlossyArgs* ap) {
    GlossyArgs& a = *ap;
    a = GlossyArgs{};
    a.world = view_of(w);
    a.field = w->field[w->cur];
    a.matnorm = (const uchar4*)matnorm;
    a.depth = depth;
    a.out = (float4*)out;
    a.keys = keys;
    a.width = width;
    a.height = height;
    for (int k = 0; k < 3; ++k) {
        a.campos[k] = cam->campos[k];
        a.camdir[k] = cam->camdir[k];
        a.horiz[k] = cam->horiz_scaled[k];
        a.verti[k] = cam->verti_scaled[k];
        a.light[k] = cam->light_dir[k];
    }
    a.frame_w = cam->frame_size[0];
    a.frame_h = cam->frame_size[1];
    a.max_dist = max_dist;
    a.max_steps = max_steps;
}
}  // extern "C++"

int lumx_glossy(lumx_world_t w, const uint8_t* matnorm, const float* depth, int width, int height,
                const lumx_camera* cam, float max_dist, int max_steps, float* out, int32_t* keys, void* stream) {
    if (!w || !matnorm || !depth || !cam || !out) return fail(LUMX_EINVAL, "null argument");
    REQUIRE_RESIDENT(w);
    if (width <= 0 || height <= 0 || width > 16384 || height > 16384)
        return fail(LUMX_EINVAL, "frame %dx%d invalid", width, height);
    if (!(max_dist > 0.0f && max_dist <= 65536.0f) || max_steps < 0)
        return fail(LUMX_EINVAL, "max_dist must be in (0, 65536] and max_steps >= 0");
    HIP_TRY(hipSetDevice(w->ctx->device));
    hipStream_t s = pick_stream(w->ctx, stream);
    GlossyArgs a;
    glossy_args(w, matnorm, depth, width, height, cam, max_dist, max_steps, out, keys, &a);
    const bool timed = !stream_capturing(s);
    if (timed) HIP_TRY(hipEventRecord(w->ev[2][0], s));
    HIP_TRY(launch_glossy(a, w->desc.radiance_format, keys != nullptr, s));
    if (timed) {
        HIP_TRY(hipEventRecord(w->ev[2][1], s));
        w->ev_valid[2] = true;
    }
    return LUMX_OK;
}

int lumx_diffuse(lumx_world_t w, const uint8_t* matnorm, const float* depth, int width, int height,
                 const lumx_camera* cam, const lumx_lightmap* lm, float* out, void* stream) {
    if (!w || !matnorm || !depth || !cam || !out) return fail(LUMX_EINVAL, "null argument");
    REQUIRE_RESIDENT(w);
    if (width <= 0 || height <= 0 || width > 16384 || height > 16384)
        return fail(LUMX_EINVAL, "frame %dx%d invalid", width, height);
    if (lm && lm->depth_dev && (lm->width <= 0 || lm->height <= 0 || lm->width > 16384 || lm->height > 16384))
        return fail(LUMX_EINVAL, "lightmap %dx%d invalid", lm->width, lm->height);
    HIP_TRY(hipSetDevice(w->ctx->device));
    hipStream_t s = pick_stream(w->ctx, stream);
    DiffuseArgs a{};
    a.world = view_of(w);
    a.field = w->field[w->cur];
    a.matnorm = (const uchar4*)matnorm;
    a.depth = depth;
    a.out = (float4*)out;
    a.width = width;
    a.height = height;
    for (int k = 0; k < 3; ++k) {
        a.campos[k] = cam->campos[k];
        a.camdir[k] = cam->camdir[k];
        a.horiz[k] = cam->horiz_scaled[k];
        a.verti[k] = cam->verti_scaled[k];
        a.light[k] = cam->light_dir[k];
    }
    a.frame_w = cam->frame_size[0];
    a.frame_h = cam->frame_size[1];
    if (lm && lm->depth_dev) {
        a.lightmap = lm->depth_dev;
        a.lm_w = lm->width;
        a.lm_h = lm->height;
        for (int k = 0; k < 16; ++k) a.lm_proj[k] = lm->proj[k];
    }
    HIP_TRY(launch_diffuse(a, w->desc.radiance_format, s));
    return LUMX_OK;
}

int lumx_camera_default(int frame_w, int frame_h, lumx_camera* out) {
    if (!out || frame_w <= 0 || frame_h <= 0) return fail(LUMX_EINVAL, "bad argument");
    // Camera defaults (internal_types.hpp:54-66): position (60, 0, 194),
    // direction normalize(vec3(0.61, 1.0, -0.8)) in FP32, 5 px per voxel over
    // a 1920x1080 view.
    float fdir[3] = {0.61f, 1.0f, -0.8f};
    fnormalize(fdir);
    double pos[3] = {60.0, 0.0, 194.0};
    double dir[3] = {fdir[0], fdir[1], fdir[2]};
    double view[2] = {1920.0 / 5.0, 1080.0 / 5.0};
    double plane[3] = {dir[0], dir[1], 0.0};
    dnormalize(plane);
    double up[3] = {0.0, 0.0, 1.0};
    double horiz[3], verti[3];
    dcross(plane, up, horiz);
    dnormalize(horiz);
    dcross(dir, horiz, verti);
    dnormalize(verti);
    float fl[3] = {0.5f, 0.5f, -0.9f};
    fnormalize(fl);
    for (int k = 0; k < 3; ++k) {
        out->campos[k] = (float)pos[k];
        out->camdir[k] = (float)dir[k];
        out->horiz_scaled[k] = (float)(horiz[k] * view[0] / 2.0);
        out->verti_scaled[k] = (float)(verti[k] * view[1] / 2.0);
        out->light_dir[k] = fl[k];
    }
    out->frame_size[0] = (float)frame_w;
    out->frame_size[1] = (float)frame_h;
    return LUMX_OK;
}

}  // extern "C"
