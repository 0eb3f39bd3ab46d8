/*
 * clum.h - headless, MI355X-native drop-in for the hot-path subset of Lum's
 * C99 API (reference: src/renderer/api/crenderer.h:19-114, implemented in
 * src/renderer/api/crenderer.cpp:12-209 and src/renderer/api/renderer.cpp:52-549).
 *
 * Same type names, layouts and signatures as the reference.  What differs:
 *  - no window: lum_get_glfw_ptr returns NULL, foliage / liquid / volumetric
 *    handles are NULL and their draw calls are no-ops (raster passes are not
 *    part of this library);
 *  - lum_submit_frame runs, on the GPU: blockify -> world upload ->
 *    exec_copies -> map.comp per drawn model -> radiance update over the
 *    selected probes -> glossy over the attached G-buffer (if any);
 *  - lum_load_world uses the data it is given (the reference loads
 *    "assets/scene" and ignores it);
 *  - failures never abort: they are recorded and returned by lumx_last_error().
 * Extensions (the reference has no C entry point for these) are prefixed
 * lumx_lum_.
 */
#ifndef LUM_C_BINDINGS_H
#define LUM_C_BINDINGS_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* LumRenderer;
typedef void* LumMeshModel;
typedef int16_t LumMeshBlock;
typedef void* LumMeshFoliage;
typedef void* LumMeshLiquid;
typedef void* LumMeshVolumetric;
typedef void* LumMeshUi;

typedef struct {
    int world_size[3];
    int static_block_palette_size;
    int maxParticleCount;

    int timestampCount;
    int fif;
    char vsync;
    char fullscreen;
    char debug;
    char profile;
} LumSettings;

#define LUM_DEFAULT_SETTINGS ((LumSettings){{48, 48, 16}, 15, 8128, 1024, 2, false, false, false, false})

typedef struct LumMeshTransform {
    float rot[4];   /* rotation quaternion (w, x, y, z) */
    float shift[3];
} LumMeshTransform;
#define LUM_DEFAULT_MESH_TRANSFORM ((LumMeshTransform){{1, 0, 0, 0}, {0, 0, 0}})

LumRenderer lum_create_instance(const LumSettings* settings);
void lum_destroy_instance(LumRenderer instance);

void lum_demo_load_scene(LumRenderer instance, const char* path);

void lum_init(LumRenderer instance, const LumSettings* settings);
void lum_load_world(LumRenderer instance, const LumMeshBlock* world_data);
void lum_load_world_file(LumRenderer instance, const char* file);

void lum_set_world_block(LumRenderer instance, int x, int y, int z, LumMeshBlock block);
LumMeshBlock lum_get_world_block(const LumRenderer instance, int x, int y, int z);

void lum_load_block(LumRenderer instance, LumMeshBlock block, void* block_data);
void lum_load_block_file(LumRenderer instance, LumMeshBlock block, const char* file);

void lum_upload_block_palette_to_gpu(LumRenderer instance);
void lum_upload_material_palette_to_gpu(LumRenderer instance);

LumMeshModel lum_load_mesh_file(LumRenderer instance, const char* file, bool try_extract_palette);
LumMeshModel lum_load_mesh_data(LumRenderer instance, void* mesh_data, int x_size, int y_size, int z_size);
void lum_free_mesh(LumRenderer instance, LumMeshModel model);

LumMeshFoliage lum_load_foliage(LumRenderer instance, const char* vertex_shader_file, int vertices_per_foliage,
                                int density);
LumMeshVolumetric lum_load_volumetric(LumRenderer instance, float max_density, float density_deviation,
                                      const uint8_t color[3]);
LumMeshLiquid lum_load_liquid(LumRenderer instance, int main_mat, int foam_mat);

void lum_start_frame(LumRenderer instance);
void lum_draw_world(LumRenderer instance);
void lum_draw_particles(LumRenderer instance);
void lum_draw_model(LumRenderer instance, const LumMeshModel mesh, const LumMeshTransform* trans);
void lum_draw_foliage_block(LumRenderer instance, const LumMeshFoliage foliage, const int pos[3]);
void lum_draw_liquid_block(LumRenderer instance, const LumMeshLiquid liquid, const int pos[3]);
void lum_draw_volumetric_block(LumRenderer instance, const LumMeshVolumetric volumetric, const int pos[3]);

void lum_prepare_frame(LumRenderer instance);
void lum_submit_frame(LumRenderer instance);

void lum_cleanup(LumRenderer instance);
void* lum_get_glfw_ptr(const LumRenderer instance);
bool lum_get_should_close(const LumRenderer instance);
bool lum_set_should_close(const LumRenderer instance, bool should_close);
void lum_wait_idle(LumRenderer instance);

bool lum_is_profiling_enabled(const LumRenderer instance);
int lum_get_current_timestamp_count(const LumRenderer instance);
void lum_get_timestamp_name(const LumRenderer instance, int index, char* out_name, size_t name_length);
double lum_get_timestamp(const LumRenderer instance, int index);
double lum_get_timestamp_difference(const LumRenderer instance, int index_start, int index_end);

/* ---- extensions (no reference counterpart in the C API) ---- */
/* Material palette: 256 x {r, g, b, a, emission, roughness} (the C++ API's
 * public mat_palette member, opaque_renderer_members.hpp:50). */
void lumx_lum_set_materials(LumRenderer instance, const float* materials);
/* Renderer::specialRadianceUpdates (internal_render.hpp:411). */
void lumx_lum_request_radiance_update(LumRenderer instance, int x, int y, int z);
/* G-buffer for the glossy pass of the next frames (device pointers:
 * matnorm u8x4, depth = view depth / 1000, out float4); NULL detaches. */
void lumx_lum_set_gbuffer(LumRenderer instance, const uint8_t* matnorm_dev, const float* depth_dev, int width,
                          int height, float* out_dev);
/* Renderer::loadPalette (renderer.cpp:150-152): the Material table of a
 * .vox file (upload it with lum_upload_material_palette_to_gpu). */
void lumx_lum_load_palette_file(LumRenderer instance, const char* file);
/* Renderer::setStoredRadianceShift / getStoredRadianceShift
 * (renderer.hpp:125-127, renderer.cpp:604-609): the shift applied to the
 * radiance field at every lum_submit_frame, before its radiance update. */
void lumx_lum_set_radiance_shift(LumRenderer instance, int x, int y, int z);
void lumx_lum_get_radiance_shift(LumRenderer instance, int* out3);
/* The pass-level world handle (lumx_world_t of lumx.h). */
void* lumx_lum_world(LumRenderer instance);

#ifdef __cplusplus
}
#endif

#endif /* LUM_C_BINDINGS_H */
