// clum.cpp - the headless lum_* C API of include/clum.h on top of lumx.h.
//
// Mirrors the hot-path subset of the reference's C99 API
// (src/renderer/api/crenderer.cpp:12-209) and the frame sequencing of
// Lum::Renderer (src/renderer/api/renderer.cpp:52-549): host-side origin
// world and block palette, per-frame model queue with screen-space culling
// (is_model_visible, renderer.cpp:220-257) and camera-distance ordering
// (prepareFrame, renderer.cpp:401-420), then on submit the GPU passes
// map -> radiance -> glossy.  The radiance work-size controller of
// update_radiance (internal_render.cpp:1190-1201) is kept, driven by the
// measured radiance pass time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "../../include/clum.h"
#include "../../include/lumx.h"
#include "host_internal.hpp"

namespace {

struct DrawRequest {
    lumx_model_t model;
    LumMeshTransform trans;
    double cam_dist;
};

struct Renderer {
    LumSettings settings{};
    lumx_context_t ctx = nullptr;
    lumx_world_t world = nullptr;
    std::vector<std::array<uint8_t, 4096>> blocks;  // BlockVoxels[x][y][z] per static block
    std::vector<float> mats = std::vector<float>(256 * 6, 0.0f);
    std::vector<DrawRequest> mesh_que;
    std::vector<int32_t> special;  // specialRadianceUpdates, int32x4
    std::vector<lumx_model_t> models;
    double cam[4][4];              // cameraTransform (identity until the first frame)
    int frame = 0;
    int magic_size = 2;
    float radiance_ms = 0.0f;
    int32_t* req_dev = nullptr;      // probe list of update_radiance (n_blocks entries)
    int32_t* count_dev = nullptr;    // its length, written by the device selection
    int32_t* special_dev = nullptr;  // specials of the frame being submitted (stream-ordered)
    size_t special_cap = 0;          // entries special_dev holds
    // Frames in flight (MAX_FRAMES_IN_FLIGHT = 2, internal_types.hpp:20): per
    // slot the frame's timing events and the pinned staging of its specials,
    // reused once the frame two submissions back has finished.
    struct Slot {
        hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};  // start, map, radiance, glossy (= done)
        bool pending = false;
        int32_t* special_pin = nullptr;
        size_t special_pin_cap = 0;
    } slot[2];
    int radiance_shift[3] = {0, 0, 0};  // Renderer::stored_radiance_shift
    bool has_palette = false;           // a Material table was taken from a .vox file
    const uint8_t* gb_matnorm = nullptr;
    const float* gb_depth = nullptr;
    float* gb_out = nullptr;
    int gb_w = 0, gb_h = 0;
    double ts[4] = {0, 0, 0, 0};
    bool should_close = false;
};

std::mutex g_live_mu;
std::set<Renderer*> g_live;

Renderer* get(LumRenderer p) {
    std::lock_guard<std::mutex> lk(g_live_mu);
    Renderer* r = (Renderer*)p;
    return g_live.count(r) ? r : nullptr;
}

void ident(double m[4][4]) {
    for (int c = 0; c < 4; ++c)
        for (int k = 0; k < 4; ++k) m[c][k] = (c == k) ? 1.0 : 0.0;
}

void matmul(const double a[4][4], const double b[4][4], double o[4][4]) {
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k) s += a[k][r] * b[c][k];
            o[c][r] = s;
        }
}

double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
void cross3(const double a[3], const double b[3], double o[3]) {
    o[0] = a[1] * b[2] - b[1] * a[2];
    o[1] = a[2] * b[0] - b[2] * a[0];
    o[2] = a[0] * b[1] - b[0] * a[1];
}
void norm3(double v[3]) {
    double l = std::sqrt(dot3(v, v));
    v[0] /= l; v[1] /= l; v[2] /= l;
}

// Camera::updateCamera (internal_render.cpp:923-934) with the default camera:
// worldToScreen = ortho(-w/2, w/2, h/2, -h/2, 0, 2000) * lookAt(pos, pos+dir, +z),
// right-handed, clip depth in [-1, 1].
void update_camera(double out[4][4]) {
    float fd[3] = {0.61f, 1.0f, -0.8f};
    float l = std::sqrt(fd[0] * fd[0] + fd[1] * fd[1] + fd[2] * fd[2]);
    double dir[3] = {(double)(fd[0] / l), (double)(fd[1] / l), (double)(fd[2] / l)};
    double eye[3] = {60.0, 0.0, 194.0};
    double up[3] = {0.0, 0.0, 1.0};
    double f[3] = {dir[0], dir[1], dir[2]};
    norm3(f);
    double s[3];
    cross3(f, up, s);
    norm3(s);
    double u[3];
    cross3(s, f, u);
    double view[4][4];
    ident(view);
    view[0][0] = s[0]; view[1][0] = s[1]; view[2][0] = s[2];
    view[0][1] = u[0]; view[1][1] = u[1]; view[2][1] = u[2];
    view[0][2] = -f[0]; view[1][2] = -f[1]; view[2][2] = -f[2];
    view[3][0] = -dot3(s, eye);
    view[3][1] = -dot3(u, eye);
    view[3][2] = dot3(f, eye);
    const double vw = 1920.0 / 5.0, vh = 1080.0 / 5.0;
    const double L = -vw / 2.0, R = vw / 2.0, B = vh / 2.0, T = -vh / 2.0, N = -0.0, F = 2000.0;
    double proj[4][4];
    ident(proj);
    proj[0][0] = 2.0 / (R - L);
    proj[1][1] = 2.0 / (T - B);
    proj[2][2] = -2.0 / (F - N);
    proj[3][0] = -(R + L) / (R - L);
    proj[3][1] = -(T + B) / (T - B);
    proj[3][2] = -(F + N) / (F - N);
    matmul(proj, view, out);
}

// Rotate v by unit quaternion q = (w, x, y, z) (FP32).
void quat_rotate(const float q[4], const float v[3], float o[3]) {
    const float qv[3] = {q[1], q[2], q[3]};
    float uv[3] = {qv[1] * v[2] - v[1] * qv[2], qv[2] * v[0] - v[2] * qv[0], qv[0] * v[1] - v[0] * qv[1]};
    float uuv[3] = {qv[1] * uv[2] - uv[1] * qv[2], qv[2] * uv[0] - uv[2] * qv[0], qv[0] * uv[1] - uv[0] * qv[1]};
    for (int k = 0; k < 3; ++k) o[k] = v[k] + ((uv[k] * q[0]) + uuv[k]) * 2.0f;
}

// is_model_visible (renderer.cpp:220-257): any transformed corner of the
// model box inside the clip cube of the FP32 view-projection.
bool model_visible(const int size[3], const LumMeshTransform& t, const double cam[4][4]) {
    float m[4][4];
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) m[c][r] = (float)cam[c][r];
    for (int x = 0; x <= 1; ++x)
        for (int y = 0; y <= 1; ++y)
            for (int z = 0; z <= 1; ++z) {
                float corner[3] = {(float)x * (float)size[0], (float)y * (float)size[1], (float)z * (float)size[2]};
                float p[3];
                quat_rotate(t.rot, corner, p);
                for (int k = 0; k < 3; ++k) p[k] += t.shift[k];
                float clip[4];
                for (int r = 0; r < 4; ++r) clip[r] = m[0][r] * p[0] + m[1][r] * p[1] + m[2][r] * p[2] + m[3][r];
                if (clip[3] != 0.0f)
                    for (int r = 0; r < 4; ++r) clip[r] /= clip[3];
                if (clip[0] >= -1.0f && clip[1] >= -1.0f && clip[2] >= -1.0f && clip[0] <= 1.0f &&
                    clip[1] <= 1.0f && clip[2] <= 1.0f)
                    return true;
            }
    return false;
}

bool read_file(const char* path, std::vector<uint8_t>* out) {
    FILE* f = path ? std::fopen(path, "rb") : nullptr;
    if (!f) return false;
    uint8_t tmp[65536];
    size_t n;
    out->clear();
    while ((n = std::fread(tmp, 1, sizeof(tmp), f)) > 0) out->insert(out->end(), tmp, tmp + n);
    std::fclose(f);
    return true;
}

// A .vox file's first model and Material table (load_stuff.cpp:233-342).
bool parse_vox_file(const char* path, int size[3], std::vector<uint8_t>* voxels, float* mats) {
    std::vector<uint8_t> buf;
    if (!read_file(path, &buf)) {
        lumx::fail(LUMX_EINVAL, "cannot read %s", path ? path : "(null)");
        return false;
    }
    if (lumx_vox_parse(buf.data(), buf.size(), size, nullptr, 0, nullptr, nullptr) != LUMX_OK) return false;
    voxels->assign((size_t)size[0] * size[1] * size[2], 0);
    return lumx_vox_parse(buf.data(), buf.size(), size, voxels->data(), voxels->size(), mats, nullptr) == LUMX_OK;
}

// load_scene (load_stuff.cpp:187-217) into the instance's origin world; an
// unreadable file reads as empty (readFileBuffer returns no bytes).
void load_scene_file(Renderer* r, const char* path) {
    if (!r->world || !path) return;
    std::vector<uint8_t> buf;
    if (!read_file(path, &buf)) buf.clear();
    const int* ws = r->settings.world_size;
    std::vector<int16_t> world((size_t)ws[0] * ws[1] * ws[2], 0);
    if (lumx_world_get_blocks(r->world, world.data(), 0) != LUMX_OK) return;
    if (lumx_scene_read(buf.data(), buf.size(), ws, r->settings.static_block_palette_size, world.data()) != LUMX_OK)
        return;
    lumx_world_set_blocks(r->world, world.data());
}

}  // namespace

extern "C" {

int lumx_scene_read(const void* data, size_t bytes, const int world_size[3], int static_blocks,
                    int16_t* blocks) {
    if ((!data && bytes) || !world_size || !blocks) return lumx::fail(LUMX_EINVAL, "null argument");
    const int wx = world_size[0], wy = world_size[1], wz = world_size[2];
    if (wx <= 0 || wy <= 0 || wz <= 0) return lumx::fail(LUMX_EINVAL, "world size invalid");
    const size_t n_world = (size_t)wx * wy * wz;
    // Fewer bytes than the header: the world is cleared (origin_world.set(0)).
    if (bytes < 12) {
        std::fill(blocks, blocks + n_world, (int16_t)0);
        return LUMX_OK;
    }
    int hs[3];
    std::memcpy(hs, data, 12);
    const uint8_t* stored = (const uint8_t*)data + 12;
    const size_t n_stored = (bytes - 12) / 2;
    // size2read = clamp(stored size, 0, world size); the stored blocks are
    // indexed with the stored size, x + Sx * y + Sx * Sy * z.
    const int rx = std::min(std::max(hs[0], 0), wx), ry = std::min(std::max(hs[1], 0), wy),
              rz = std::min(std::max(hs[2], 0), wz);
    if (rx > 0 && ry > 0 && rz > 0) {
        const long long last = (long long)(rx - 1) + (long long)hs[0] * (ry - 1) + (long long)hs[0] * hs[1] * (rz - 1);
        if (last >= (long long)n_stored)
            return lumx::fail(LUMX_EINVAL, "scene holds %zu blocks, its %dx%dx%d header needs more", n_stored, hs[0],
                              hs[1], hs[2]);
    }
    const int16_t hi = (int16_t)std::min(std::max(static_blocks, 0), 32767);
    for (int z = 0; z < rz; ++z)
        for (int y = 0; y < ry; ++y)
            for (int x = 0; x < rx; ++x) {
                const size_t idx = (size_t)x + (size_t)hs[0] * y + (size_t)hs[0] * hs[1] * z;
                int16_t b;
                std::memcpy(&b, stored + 2 * idx, 2);
                blocks[x + (size_t)wx * (y + (size_t)wy * z)] = std::min(std::max(b, (int16_t)0), hi);
            }
    return LUMX_OK;
}

LumRenderer lum_create_instance(const LumSettings* settings) {
    Renderer* r = new Renderer();
    if (settings) r->settings = *settings;
    ident(r->cam);
    std::lock_guard<std::mutex> lk(g_live_mu);
    g_live.insert(r);
    return r;
}

static void release(Renderer* r) {
    if (r->ctx) lumx_context_sync(r->ctx);
    for (lumx_model_t m : r->models) lumx_model_destroy(m);
    r->models.clear();
    if (r->world) lumx_world_destroy(r->world);
    r->world = nullptr;
    if (r->req_dev) hipFree(r->req_dev);
    r->req_dev = nullptr;
    if (r->count_dev) hipFree(r->count_dev);
    r->count_dev = nullptr;
    if (r->special_dev) hipFree(r->special_dev);
    r->special_dev = nullptr;
    r->special_cap = 0;
    for (Renderer::Slot& fs : r->slot) {
        for (hipEvent_t& e : fs.ev) {
            if (e) (void)hipEventDestroy(e);
            e = nullptr;
        }
        if (fs.special_pin) hipHostFree(fs.special_pin);
        fs.special_pin = nullptr;
        fs.special_pin_cap = 0;
        fs.pending = false;
    }
    if (r->ctx) lumx_context_destroy(r->ctx);
    r->ctx = nullptr;
}

void lum_destroy_instance(LumRenderer instance) {
    Renderer* r = get(instance);
    if (!r) return;
    release(r);
    {
        std::lock_guard<std::mutex> lk(g_live_mu);
        g_live.erase(r);
    }
    delete r;
}

void lum_demo_load_scene(LumRenderer instance, const char* path) {
    Renderer* r = get(instance);
    if (r) load_scene_file(r, path);
}

void lum_init(LumRenderer instance, const LumSettings* settings) {
    Renderer* r = get(instance);
    if (!r || !settings) return;
    r->settings = *settings;
    if (lumx_context_create(0, &r->ctx) != LUMX_OK) return;
    lumx_world_desc d{};
    for (int a = 0; a < 3; ++a) d.size[a] = settings->world_size[a];
    d.static_blocks = settings->static_block_palette_size;
    d.palette_slots = 0;
    d.radiance_format = LUMX_RADIANCE_F32;
    if (lumx_world_create(r->ctx, &d, &r->world) != LUMX_OK) return;
    r->blocks.assign(settings->static_block_palette_size, std::array<uint8_t, 4096>{});
    size_t nb = (size_t)d.size[0] * d.size[1] * d.size[2];
    // Every probe appears at most once in the selection: n_blocks entries suffice.
    if (hipMalloc(&r->req_dev, nb * 16) != hipSuccess) r->req_dev = nullptr;
    if (hipMalloc(&r->count_dev, sizeof(int32_t)) != hipSuccess) r->count_dev = nullptr;
    for (Renderer::Slot& fs : r->slot)
        for (hipEvent_t& e : fs.ev)
            if (hipEventCreate(&e) != hipSuccess) e = nullptr;
}

void lum_load_world(LumRenderer instance, const LumMeshBlock* world_data) {
    Renderer* r = get(instance);
    if (r && r->world && world_data) lumx_world_set_blocks(r->world, world_data);
}

void lum_load_world_file(LumRenderer instance, const char* file) {
    Renderer* r = get(instance);
    if (r) load_scene_file(r, file);
}

void lum_set_world_block(LumRenderer instance, int x, int y, int z, LumMeshBlock block) {
    Renderer* r = get(instance);
    if (r && r->world) lumx_world_set_block(r->world, x, y, z, block);
}

LumMeshBlock lum_get_world_block(const LumRenderer instance, int x, int y, int z) {
    Renderer* r = get(instance);
    int16_t b = 0;
    if (r && r->world) lumx_world_get_block(r->world, x, y, z, &b);
    return b;
}

void lum_load_block(LumRenderer instance, LumMeshBlock block, void* block_data) {
    Renderer* r = get(instance);
    if (!r || !block_data || block < 0 || block >= (int)r->blocks.size()) return;
    std::memcpy(r->blocks[block].data(), block_data, 4096);
}

void lum_load_block_file(LumRenderer instance, LumMeshBlock block, const char* file) {
    // load_block (load_stuff.cpp:316-342): BlockVoxels[x][y][z] from the
    // file's first model; a model larger than 16 per axis is cut to 16.
    Renderer* r = get(instance);
    if (!r || block < 0 || block >= (int)r->blocks.size()) return;
    int size[3];
    std::vector<uint8_t> v;
    if (!parse_vox_file(file, size, &v, nullptr)) return;
    std::array<uint8_t, 4096>& dst = r->blocks[block];
    dst.fill(0);
    for (int x = 0; x < std::min(size[0], 16); ++x)
        for (int y = 0; y < std::min(size[1], 16); ++y)
            for (int z = 0; z < std::min(size[2], 16); ++z)
                dst[x * 256 + y * 16 + z] = v[x + (size_t)size[0] * (y + (size_t)size[1] * z)];
}

void lum_upload_block_palette_to_gpu(LumRenderer instance) {
    Renderer* r = get(instance);
    if (!r || !r->world || r->blocks.empty()) return;
    std::vector<uint8_t> flat(r->blocks.size() * 4096);
    for (size_t b = 0; b < r->blocks.size(); ++b) std::memcpy(flat.data() + b * 4096, r->blocks[b].data(), 4096);
    lumx_world_set_palette(r->world, 0, (int)r->blocks.size(), flat.data(), LUMX_PALETTE_REF_ZFAST);
}

void lum_upload_material_palette_to_gpu(LumRenderer instance) {
    Renderer* r = get(instance);
    if (r && r->world) lumx_world_set_materials(r->world, r->mats.data());
}

LumMeshModel lum_load_mesh_file(LumRenderer instance, const char* file, bool try_extract_palette) {
    // loadModel(file) -> load_mesh (load_stuff.cpp:288-306): the first model;
    // the file's Material table replaces the palette when asked and no
    // palette was taken from a file before.
    Renderer* r = get(instance);
    if (!r || !r->ctx) return nullptr;
    int size[3];
    std::vector<uint8_t> v;
    std::vector<float> mats(256 * 6);
    if (!parse_vox_file(file, size, &v, mats.data())) return nullptr;
    lumx_model_t m = nullptr;
    if (lumx_model_create(r->ctx, v.data(), size[0], size[1], size[2], &m) != LUMX_OK) return nullptr;
    r->models.push_back(m);
    if (try_extract_palette && !r->has_palette) {
        r->mats = mats;
        r->has_palette = true;
    }
    return m;
}

LumMeshModel lum_load_mesh_data(LumRenderer instance, void* mesh_data, int x_size, int y_size, int z_size) {
    Renderer* r = get(instance);
    if (!r || !r->ctx) return nullptr;
    lumx_model_t m = nullptr;
    if (lumx_model_create(r->ctx, (const uint8_t*)mesh_data, x_size, y_size, z_size, &m) != LUMX_OK) return nullptr;
    r->models.push_back(m);
    return m;
}

void lum_free_mesh(LumRenderer instance, LumMeshModel model) {
    Renderer* r = get(instance);
    if (!r || !model) return;
    auto it = std::find(r->models.begin(), r->models.end(), (lumx_model_t)model);
    if (it == r->models.end()) return;
    lumx_model_destroy(*it);
    r->models.erase(it);
}

LumMeshFoliage lum_load_foliage(LumRenderer, const char*, int, int) { return nullptr; }
LumMeshVolumetric lum_load_volumetric(LumRenderer, float, float, const uint8_t*) { return nullptr; }
LumMeshLiquid lum_load_liquid(LumRenderer, int, int) { return nullptr; }

void lum_start_frame(LumRenderer instance) {
    Renderer* r = get(instance);
    if (r) r->mesh_que.clear();
}

void lum_draw_world(LumRenderer) {}
void lum_draw_particles(LumRenderer) {}

void lum_draw_model(LumRenderer instance, const LumMeshModel mesh, const LumMeshTransform* trans) {
    Renderer* r = get(instance);
    if (!r || !mesh || !trans) return;
    int size[3];
    if (lumx_model_size((lumx_model_t)mesh, size) != LUMX_OK) return;
    if (!model_visible(size, *trans, r->cam)) return;
    r->mesh_que.push_back(DrawRequest{(lumx_model_t)mesh, *trans, 0.0});
}

void lum_draw_foliage_block(LumRenderer, const LumMeshFoliage, const int*) {}
void lum_draw_liquid_block(LumRenderer, const LumMeshLiquid, const int*) {}
void lum_draw_volumetric_block(LumRenderer, const LumMeshVolumetric, const int*) {}

void lum_prepare_frame(LumRenderer instance) {
    Renderer* r = get(instance);
    if (!r) return;
    for (DrawRequest& d : r->mesh_que) {
        double z = 0.0;
        for (int k = 0; k < 3; ++k) z += r->cam[k][2] * (double)d.trans.shift[k];
        z += r->cam[3][2];
        d.cam_dist = -z;
    }
    std::stable_sort(r->mesh_que.begin(), r->mesh_que.end(),
                     [](const DrawRequest& a, const DrawRequest& b) { return a.cam_dist > b.cam_dist; });
}

// One frame: map.comp with blockify on the device, shift_radiance,
// update_radiance (device probe selection, work-size controller, dispatch
// over the device-side count) and glossy, queued on the context's stream with
// no wait for the GPU.  Two frames are in flight (renderer.cpp:71-73): a
// submission first waits for the frame two submissions back, the reference's
// fence wait, and takes its pass times for the timestamps and the controller.
void lum_submit_frame(LumRenderer instance) {
    Renderer* r = get(instance);
    if (!r || !r->world || !r->req_dev || !r->count_dev) return;
    Renderer::Slot& fs = r->slot[r->frame & 1];
    for (hipEvent_t e : fs.ev)
        if (!e) return;
    if (fs.pending) {
        if (hipEventSynchronize(fs.ev[3]) != hipSuccess) return;
        float t[3] = {0.0f, 0.0f, 0.0f};
        for (int k = 0; k < 3; ++k) (void)hipEventElapsedTime(&t[k], fs.ev[k], fs.ev[k + 1]);
        r->radiance_ms = t[1];
        r->ts[0] = 0.0;
        r->ts[1] = t[0];
        r->ts[2] = r->ts[1] + t[1];
        r->ts[3] = r->ts[2] + t[2];
        fs.pending = false;
    }
    update_camera(r->cam);
    hipStream_t stream = (hipStream_t)lumx_context_stream(r->ctx);
    if (hipEventRecord(fs.ev[0], stream) != hipSuccess) return;
    // blockify + exec_copies + map.comp, blockify on the device: only the
    // model boxes cross PCIe (the origin world too, after it changed).
    std::vector<lumx_model_t> models;
    std::vector<lumx_placement> placed;
    for (const DrawRequest& d : r->mesh_que) {
        lumx_transform t;
        std::memcpy(t.rot, d.trans.rot, sizeof(t.rot));
        std::memcpy(t.shift, d.trans.shift, sizeof(t.shift));
        int size[3];
        lumx_placement p;
        if (lumx_model_size(d.model, size) != LUMX_OK || lumx_model_placement(&t, size, &p) != LUMX_OK) return;
        models.push_back(d.model);
        placed.push_back(p);
    }
    if (lumx_map_models_device(r->world, (int)models.size(), models.data(), placed.data(), stream) != LUMX_OK) return;
    if (hipEventRecord(fs.ev[1], stream) != hipSuccess) return;
    // shift_radiance with the stored shift (renderer.cpp:436), then
    // update_radiance.
    if (lumx_world_shift_radiance(r->world, r->radiance_shift[0], r->radiance_shift[1], r->radiance_shift[2],
                                  stream) != LUMX_OK)
        return;
    const size_t n_special = r->special.size() / 4;
    if (n_special > fs.special_pin_cap) {
        // The slot's previous frame has finished (waited for above).
        if (fs.special_pin) hipHostFree(fs.special_pin);
        fs.special_pin = nullptr;
        fs.special_pin_cap = 0;
        if (hipHostMalloc((void**)&fs.special_pin, n_special * 16, hipHostMallocDefault) != hipSuccess) return;
        fs.special_pin_cap = n_special;
    }
    if (n_special > r->special_cap) {
        if (r->special_dev) hipFree(r->special_dev);
        r->special_cap = 0;
        if (hipMalloc(&r->special_dev, n_special * 16) != hipSuccess) {
            r->special_dev = nullptr;
            return;
        }
        r->special_cap = n_special;
    }
    if (n_special) {
        std::memcpy(fs.special_pin, r->special.data(), n_special * 16);
        if (hipMemcpyAsync(r->special_dev, fs.special_pin, n_special * 16, hipMemcpyHostToDevice, stream) !=
            hipSuccess)
            return;
    }
    r->special.clear();
    if (lumx_select_probes_device(r->world, r->special_dev, (int)n_special, r->req_dev, r->count_dev, stream) !=
        LUMX_OK)
        return;
    // Work-size controller on the radiance time of the last finished frame
    // (internal_render.cpp:1192-1196).
    if (r->radiance_ms < 1.8f) r->magic_size--;
    else if (r->radiance_ms > 2.2f) r->magic_size++;
    r->magic_size = std::min(std::max(r->magic_size, 1), 10);
    lumx_radiance_params p{};
    p.rays = 64;
    p.time_seed = r->frame;
    p.size = r->magic_size;
    p.shift = r->frame % r->magic_size;
    p.alpha = 0.01f;  // REACTIVNESS default (radiance.comp:28)
    p.max_dist = 128.0f;
    p.max_steps = 0;
    lumx_world_desc wd;
    lumx_world_info(r->world, &wd);
    const int n_blocks = wd.size[0] * wd.size[1] * wd.size[2];
    if (lumx_radiance_update_counted(r->world, r->req_dev, r->count_dev, n_blocks, &p, stream) != LUMX_OK) return;
    if (hipEventRecord(fs.ev[2], stream) != hipSuccess) return;
    if (r->gb_matnorm && r->gb_depth && r->gb_out) {
        lumx_camera cam;
        lumx_camera_default(r->gb_w, r->gb_h, &cam);
        if (lumx_glossy(r->world, r->gb_matnorm, r->gb_depth, r->gb_w, r->gb_h, &cam, 128.0f, 0, r->gb_out, nullptr,
                        stream) != LUMX_OK)
            return;
    }
    if (hipEventRecord(fs.ev[3], stream) != hipSuccess) return;
    fs.pending = true;
    r->frame++;
}

void lum_cleanup(LumRenderer instance) {
    // The reference's lum_cleanup also deletes the instance (crenderer.cpp:168-171).
    lum_destroy_instance(instance);
}

void* lum_get_glfw_ptr(const LumRenderer) { return nullptr; }

bool lum_get_should_close(const LumRenderer instance) {
    Renderer* r = get(instance);
    return r ? r->should_close : true;
}

bool lum_set_should_close(const LumRenderer instance, bool should_close) {
    Renderer* r = get(instance);
    if (!r) return should_close;
    r->should_close = should_close;
    return r->should_close;
}

void lum_wait_idle(LumRenderer instance) {
    Renderer* r = get(instance);
    if (r && r->ctx) lumx_context_sync(r->ctx);
}

bool lum_is_profiling_enabled(const LumRenderer instance) { return get(instance) != nullptr; }

int lum_get_current_timestamp_count(const LumRenderer instance) { return get(instance) ? 4 : 0; }

void lum_get_timestamp_name(const LumRenderer instance, int index, char* out_name, size_t name_length) {
    static const char* names[4] = {"start_frame", "map", "update_radiance", "glossy"};
    if (!out_name || name_length == 0) return;
    const char* n = (get(instance) && index >= 0 && index < 4) ? names[index] : "";
    std::strncpy(out_name, n, name_length - 1);
    out_name[name_length - 1] = '\0';
}

double lum_get_timestamp(const LumRenderer instance, int index) {
    Renderer* r = get(instance);
    return (r && index >= 0 && index < 4) ? r->ts[index] : 0.0;
}

double lum_get_timestamp_difference(const LumRenderer instance, int index_start, int index_end) {
    return lum_get_timestamp(instance, index_end) - lum_get_timestamp(instance, index_start);
}

void lumx_lum_set_materials(LumRenderer instance, const float* materials) {
    Renderer* r = get(instance);
    if (r && materials) std::memcpy(r->mats.data(), materials, sizeof(float) * 256 * 6);
}

void lumx_lum_request_radiance_update(LumRenderer instance, int x, int y, int z) {
    Renderer* r = get(instance);
    if (!r) return;
    const int* ws = r->settings.world_size;
    if (x < 0 || y < 0 || z < 0 || x >= ws[0] || y >= ws[1] || z >= ws[2]) return;
    r->special.insert(r->special.end(), {x, y, z, 0});
}

void lumx_lum_set_gbuffer(LumRenderer instance, const uint8_t* matnorm_dev, const float* depth_dev, int width,
                          int height, float* out_dev) {
    Renderer* r = get(instance);
    if (!r) return;
    r->gb_matnorm = matnorm_dev;
    r->gb_depth = depth_dev;
    r->gb_out = out_dev;
    r->gb_w = width;
    r->gb_h = height;
}

void lumx_lum_load_palette_file(LumRenderer instance, const char* file) {
    // Renderer::loadPalette -> extract_palette (load_stuff.cpp:274-286).
    Renderer* r = get(instance);
    if (!r) return;
    int size[3];
    std::vector<uint8_t> v;
    std::vector<float> mats(256 * 6);
    if (!parse_vox_file(file, size, &v, mats.data())) return;
    r->mats = mats;
    r->has_palette = true;
}

void lumx_lum_set_radiance_shift(LumRenderer instance, int x, int y, int z) {
    Renderer* r = get(instance);
    if (!r) return;
    r->radiance_shift[0] = x;
    r->radiance_shift[1] = y;
    r->radiance_shift[2] = z;
}

void lumx_lum_get_radiance_shift(LumRenderer instance, int* out3) {
    Renderer* r = get(instance);
    if (!r || !out3) return;
    for (int k = 0; k < 3; ++k) out3[k] = r->radiance_shift[k];
}

void* lumx_lum_world(LumRenderer instance) {
    Renderer* r = get(instance);
    return r ? (void*)r->world : nullptr;
}

}  // extern "C"
