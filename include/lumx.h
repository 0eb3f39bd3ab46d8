/*
 * lumx.h - C ABI of the MI355X-native Lum voxel-lighting hot path.
 *
 * Pass-level entry points for the three hot passes of Lum
 * (radiance probe update, model->world voxelization, glossy reflection march)
 * plus the world / block-palette / material / model / radiance-buffer state
 * they run on.  Plain C types only: host pointers are marked _host, device
 * (HBM) pointers _dev, and `stream` is a hipStream_t passed as void* (NULL =
 * the HIP null stream, as everywhere in HIP; lumx_context_stream() returns the
 * context's own non-blocking stream).
 *
 * Every function returns LUMX_OK (0) or a negative LUMX_E* status and never
 * aborts; lumx_last_error() returns the message of the calling thread's last
 * failure.  One world may not be used from two threads at once.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   lumx_world_*           LumInternalRenderer::createImages world/radiance/palette/material
 *                          resources (src/renderer/src/internal_render.cpp:107-209)
 *                          and Lum::Renderer::setWorldBlock/getWorldBlock
 *                          (src/renderer/api/renderer.cpp:116-133)
 *   lumx_world_set_palette updateBlockPalette (src/renderer/src/load_stuff.cpp:36-86)
 *   lumx_world_set_materials updateMaterialPalette (src/renderer/src/load_stuff.cpp:88-111)
 *   lumx_model_create      load_mesh / create_RayTrace_VoxelImages
 *                          (src/renderer/src/load_stuff.cpp:376-385, src/renderer/src/setup.cpp:4-39)
 *   lumx_model_placement   get_shift + map_mesh push constants + blockify_mesh AABB
 *                          (src/renderer/src/internal_render.cpp:989-1013,1029-1035,1494-1511)
 *   lumx_map_models*       start_blockify/blockify_mesh/end_blockify + exec_copies +
 *                          start_map/map_mesh/end_map -> shaders/map.comp
 *                          (src/renderer/src/internal_render.cpp:977-1097,1375-1527)
 *   lumx_radiance_update*  update_radiance -> shaders/radiance.comp
 *                          (src/renderer/src/internal_render.cpp:1129-1214)
 *   lumx_select_probes*    update_radiance probe selection
 *                          (src/renderer/src/internal_render.cpp:1142-1168)
 *   lumx_world_shift_radiance shift_radiance, applied by Lum::Renderer with
 *                          setStoredRadianceShift's value
 *                          (src/renderer/src/internal_render.cpp:1218-1335,
 *                          src/renderer/api/renderer.cpp:436,604-609)
 *   lumx_glossy            glossy_raygen stencil + glossy -> shaders/glossy.frag
 *                          (src/renderer/src/internal_render.cpp:1990-1998,2027-2037)
 *   lumx_diffuse           diffuse -> shaders/diffuse.frag
 *                          (src/renderer/src/internal_render.cpp:1968-1977)
 *   lumx_camera_default    Camera::updateCamera + start_raygen UBO fill
 *                          (src/renderer/src/internal_render.cpp:923-934,1580-1592)
 */
#ifndef LUMX_H
#define LUMX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LUMX_OK 0
#define LUMX_EINVAL (-1)     /* bad argument / handle / size */
#define LUMX_ENOMEM (-2)     /* device allocation failed or HBM budget exceeded */
#define LUMX_EDEVICE (-3)    /* HIP runtime error */
#define LUMX_EOVERFLOW (-4)  /* block palette slots exhausted */
#define LUMX_EBUSY (-5)      /* serving queue full / budget refused */

#define LUMX_RADIANCE_F32 0  /* RGBA float32 texels, 16 B */
#define LUMX_RADIANCE_F16 1  /* RGBA float16 texels, 8 B, round to nearest even */
#define LUMX_RADIANCE_BF16 2 /* RGBA bfloat16 texels, 8 B, round to nearest even */

#define LUMX_PALETTE_XYZ_XFAST 0  /* block voxels indexed x + 16y + 256z */
#define LUMX_PALETTE_REF_ZFAST 1  /* Lum BlockVoxels[x][y][z] (z fastest) */

typedef struct lumx_context* lumx_context_t;
typedef struct lumx_world* lumx_world_t;
typedef struct lumx_model* lumx_model_t;
typedef struct lumx_batch* lumx_batch_t;
typedef struct lumx_graph* lumx_graph_t;

typedef struct {
    int size[3];          /* world size in 16^3 blocks (x, y, z) */
    int static_blocks;    /* static block palette size (slots 0..static-1) */
    int palette_slots;    /* total palette slots; 0 = world blocks + static (capped at 32768) */
    int radiance_format;  /* LUMX_RADIANCE_* */
} lumx_world_desc;

/* Radiance update parameters (radiance.comp push constants + constants). */
typedef struct {
    int rays;          /* rays per probe: power of two <= 64 or multiple of 64 <= 1024 */
    int time_seed;     /* pco.timeSeed (frame index) */
    int size;          /* pco.size: request stride (>= 1) */
    int shift;         /* pco.sihft: request offset (0 <= shift < size) */
    float alpha;       /* REACTIVNESS temporal blend weight */
    float max_dist;    /* march distance cap in voxels (reference: 128) */
    int max_steps;     /* coarse-step cap, 0 = none (reference) */
    int reserved;
} lumx_radiance_params;

/* Mesh transform, identical layout to LumMeshTransform: rot = (w, x, y, z). */
typedef struct {
    float rot[4];
    float shift[3];
} lumx_transform;

/* Everything the map pass and blockify consume for one placed model. */
typedef struct {
    float minv[16];     /* inverse model transform, column-major (GLSL mat4) */
    int shift[4];       /* border.min (dispatch origin), w unused */
    int groups[3];      /* dispatch size in 4x4x4 workgroups */
    float aabb_min[3];  /* float AABB of get_shift() (used by blockify) */
    float aabb_max[3];
} lumx_placement;

/* The glossy pass UBO fields (common/ubo.glsl) it reads. */
typedef struct {
    float campos[3];
    float camdir[3];
    float horiz_scaled[3];  /* horizline * viewSize.x / 2 */
    float verti_scaled[3];  /* vertiline * viewSize.y / 2 */
    float light_dir[3];     /* globalLightDir */
    float frame_size[2];    /* framebuffer size in pixels */
} lumx_camera;

/* Shadow map of the diffuse pass's sun term: the reference's D16_UNORM
 * lightmap (1024 x 1024, internal_render.cpp:71,119-126; internal_render.hpp:450)
 * and ubo.lightmap_proj. */
typedef struct {
    const uint16_t* depth_dev; /* row-major width x height D16 values; NULL = no sun term */
    int width, height;
    float proj[16];            /* column-major mat4 */
} lumx_lightmap;

/* Launch shape of the radiance update.  No reference counterpart: it
 * replaces the fixed 64-lane workgroup of radiance.comp:37,360. */
typedef struct {
    int variant;            /* 0 = one thread per ray, a probe per wave; 1 = tiled persistent;
                               2 = as 0, marching the occupancy bits instead of the u8 palette;
                               3 = lane refill: a wave marches the rays of probes_per_batch (1..4)
                               probes as one pool, finished lanes taking the next ray (64 rays) */
    int threads;            /* workgroup size of the tiled kernel (multiple of 64, <= 1024) */
    int probes_per_batch;   /* probes per batch of the tiled kernel (power of two, 1..64); probes per
                               wave of the refill variant */
    int lds_block_bits;     /* stage the empty-block bits in LDS when they fit */
    int lds_palette_bits;   /* stage the palette occupancy bits in LDS when they fit */
    int wg_per_cu;          /* persistent workgroups per CU, 0 = occupancy limit */
    int rays_per_lane;      /* variant 0: rays traced one after another by one thread (0 or 1, 2, 4);
                               used when rays per probe is a power of two >= it, else 1.  The
                               results do not depend on it. */
    int compact;            /* subset updates of variants 0 and 2: run the requested probes once
                               each in texel order (claims, then an ordered compaction of the
                               claimed texels) instead of in list order.  0 = when it pays (at
                               least 2048 probes and a sixteenth of the world, and a field larger
                               than 4 MiB), 1 = always, -1 = never.  The results do not depend
                               on it.  Batched and served launches compact when it pays. */
} lumx_radiance_tuning;

/* ---- context / errors ---- */
int lumx_context_create(int device, lumx_context_t* out);
int lumx_context_destroy(lumx_context_t ctx);
int lumx_context_sync(lumx_context_t ctx);
void* lumx_context_stream(lumx_context_t ctx);
const char* lumx_last_error(void);
const char* lumx_version(void);

/* ---- world state ---- */
int lumx_world_create(lumx_context_t ctx, const lumx_world_desc* desc, lumx_world_t* out);
int lumx_world_destroy(lumx_world_t w);
int lumx_world_info(lumx_world_t w, lumx_world_desc* out);
/* Origin world (x-fastest int16, 0 = air), copied at call time. */
int lumx_world_set_blocks(lumx_world_t w, const int16_t* blocks_host);
int lumx_world_get_blocks(lumx_world_t w, int16_t* blocks_host, int blockified);
int lumx_world_set_block(lumx_world_t w, int x, int y, int z, int16_t block);
int lumx_world_get_block(lumx_world_t w, int x, int y, int z, int16_t* out);
/* Static palette blocks [first, first+count): count*4096 voxels. */
int lumx_world_set_palette(lumx_world_t w, int first, int count, const uint8_t* voxels_host, int layout);
/* Device palette readback: slots [first, first+count) x-fastest. */
int lumx_world_get_palette(lumx_world_t w, int first, int count, uint8_t* voxels_host);
/* 256 materials x {r, g, b, a, emission, roughness}. */
int lumx_world_set_materials(lumx_world_t w, const float* materials_host);
/* Radiance field, RGBA float per probe (x-fastest); converted to the storage format. */
int lumx_world_write_radiance(lumx_world_t w, const float* rgba_host);
int lumx_world_read_radiance(lumx_world_t w, float* rgba_host);
/* Device pointer and byte size of the current radiance field. */
int lumx_world_radiance_dev(lumx_world_t w, void** ptr, size_t* bytes);
/* Upload the origin world as the current world (no models mapped). */
int lumx_world_upload_blocks(lumx_world_t w, void* stream);
/* Residency: evict copies the world's device state (blockified world,
 * palette slots in use, materials, current radiance field) to host memory and
 * frees its HBM; restore re-creates it bit for bit (in a new device block,
 * the current field possibly in the other of its two buffers), so graphs
 * recorded before an eviction must be recorded again.  Passes on an evicted
 * world fail with LUMX_EINVAL.  No reference counterpart (the reference keeps
 * one world resident); used by the serving engine's HBM budget. */
int lumx_world_evict(lumx_world_t w);
int lumx_world_restore(lumx_world_t w);
int lumx_world_resident(lumx_world_t w, int* resident, size_t* device_bytes);

/* ---- models ---- */
int lumx_model_create(lumx_context_t ctx, const uint8_t* voxels_host, int sx, int sy, int sz,
                      lumx_model_t* out);
int lumx_model_destroy(lumx_model_t m);
int lumx_model_size(lumx_model_t m, int size[3]);
/* Host math of map_mesh/get_shift for one transform and model size. */
int lumx_model_placement(const lumx_transform* t, const int size[3], lumx_placement* out);

/* ---- passes ---- */
/* blockify (host) + world upload + exec_copies + map.comp for n models in
 * order (later models win on overlap). */
int lumx_map_models(lumx_world_t w, int n, const lumx_model_t* models, const lumx_transform* transforms,
                    void* stream);
int lumx_map_models_placed(lumx_world_t w, int n, const lumx_model_t* models,
                           const lumx_placement* placements, void* stream);
/* The same map with blockify and exec_copies on the device: the blockified
 * world is built from a device copy of the origin world, so nothing
 * proportional to the world crosses PCIe (only the model boxes).  Results
 * equal lumx_map_models_placed bit for bit.  Host queries of the blockified
 * state (lumx_world_get_blocks, lumx_world_palette_counter,
 * lumx_select_probes) synchronise the device and read it back.  Slots beyond
 * the palette are not allocated; lumx_world_palette_counter then returns
 * LUMX_EOVERFLOW. */
int lumx_map_models_device(lumx_world_t w, int n, const lumx_model_t* models,
                           const lumx_placement* placements, void* stream);
/* lumx_map_models_device without any allocation, host or device.  No
 * counterpart in the reference: it is for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Its scratch (model boxes, ballot words, tile counts) is one
 * fixed 4 MiB device array per GPU, held until the process ends, and the
 * model boxes reach it as a kernel argument.  Limits: at most 64 models, and
 * model boxes covering at most about 16,000 x 2048 blocks in total (the
 * scratch); beyond either it returns LUMX_EINVAL.  Calls that overlap on one
 * GPU (any worlds, any streams) race on the array: order them on one stream
 * or with events.  Inside a capture the world's origin must already be
 * uploaded (any earlier map did it), else LUMX_EINVAL.  Results equal
 * lumx_map_models_device bit for bit. */
int lumx_map_models_device_noalloc(lumx_world_t w, int n, const lumx_model_t* models,
                                   const lumx_placement* placements, void* stream);
/* Palette slots allocated by the last map (the reference's paletteCounter). */
int lumx_world_palette_counter(lumx_world_t w, int* out);

/* radiance.comp: requests_dev = int32x4 (x, y, z, _) probe list in HBM, or
 * NULL for every probe in x-fastest order (n_requests ignored).  Reads the
 * current field, writes a new one (Jacobi), then makes it current.  A list
 * dispatches n_requests / size probes, at most the world's block count (the
 * capacity of the reference's request buffer; more is LUMX_EINVAL).  Entries
 * outside the world are skipped; where several dispatched entries name one
 * probe, the last of them is the one stored. */
int lumx_radiance_update(lumx_world_t w, const int32_t* requests_dev, int n_requests,
                         const lumx_radiance_params* p, void* stream);
/* Same, also writing per-ray hit keys (0 = miss) and {coarse, refine} sample
 * counts for every dispatched ray (either may be NULL). */
int lumx_radiance_update_traced(lumx_world_t w, const int32_t* requests_dev, int n_requests,
                                const lumx_radiance_params* p, int32_t* ray_keys_dev,
                                int32_t* ray_steps_dev, void* stream);
/* Launch shape used by this world's radiance updates. */
int lumx_world_set_tuning(lumx_world_t w, const lumx_radiance_tuning* t);
int lumx_world_get_tuning(lumx_world_t w, lumx_radiance_tuning* out);
/* Autotuner: times candidate launch shapes of lumx_radiance_update with
 * these parameters on this world (its state is left unchanged), keeps the
 * fastest and returns it (either output may be NULL).  No reference
 * counterpart. */
int lumx_autotune_radiance(lumx_world_t w, const lumx_radiance_params* p, int reps,
                           lumx_radiance_tuning* best, float* best_ms);
/* Probe selection of update_radiance on the current (blockified) world: the
 * probes whose 27 clamped neighbours hold a non-zero block, x-fastest, then
 * every special whose probe the grid did not select, repeats included, in
 * order.  The list ends at the world's block count (the capacity of the
 * reference's request buffer), so requests_host needs room for n_blocks
 * int32x4 entries.  Specials outside the world are an error. */
int lumx_select_probes(lumx_world_t w, const int32_t* special_host, int n_special,
                       int32_t* requests_host, int* n_out);
/* The same selection on the device, stream-ordered after the last map:
 * requests_dev (room for n_blocks int32x4 entries) receives the list,
 * *count_dev its length.  Specials outside the world are skipped. */
int lumx_select_probes_device(lumx_world_t w, const int32_t* special_dev, int n_special,
                              int32_t* requests_dev, int32_t* count_dev, void* stream);
/* lumx_radiance_update over a list whose length is read on the device
 * (*count_dev, at most max_requests), e.g. the output of
 * lumx_select_probes_device: no host round trip between the two. */
int lumx_radiance_update_counted(lumx_world_t w, const int32_t* requests_dev, const int32_t* count_dev,
                                 int max_requests, const lumx_radiance_params* p, void* stream);
/* shift_radiance (src/renderer/src/internal_render.cpp:1218-1335): the
 * texel at p + (dx, dy, dz) moves to p wherever both lie in the world;
 * other texels keep their value.  A shift of a whole world size or more on
 * any axis changes nothing.  The shifted field becomes current. */
int lumx_world_shift_radiance(lumx_world_t w, int dx, int dy, int dz, void* stream);

/* glossy.frag over a width x height G-buffer (matnorm u8x4, depth = view
 * depth / 1000).  Pixels whose material roughness > 0.5 are not written. */
int lumx_glossy(lumx_world_t w, const uint8_t* matnorm_dev, const float* depth_dev, int width,
                int height, const lumx_camera* cam, float max_dist, int max_steps, float* out_dev,
                int32_t* hit_keys_dev, void* stream);
/* diffuse.frag (the radiance field's consumer) over a width x height
 * G-buffer: out_dev float4 per pixel = ((2 * field sample at origin + 6 *
 * normal + emission + sun) * colour / 8, 1).  lightmap may be NULL. */
int lumx_diffuse(lumx_world_t w, const uint8_t* matnorm_dev, const float* depth_dev, int width, int height,
                 const lumx_camera* cam, const lumx_lightmap* lightmap, float* out_dev, void* stream);
/* ---- MagicaVoxel .vox import ----
 * The reference reads .vox through ogt_vox (common/ogt_vox.hpp,
 * vox_read_scene) in load_mesh / load_block / extract_palette and builds its
 * Material table in extract_palette_mem (src/renderer/src/load_stuff.cpp:
 * 233-342).  Parses a .vox file image: size_out[3] gets the first distinct
 * non-empty model's size; voxels (x-fastest colour indices, 0 = empty), when
 * non-NULL, needs voxels_capacity >= size_x * size_y * size_z (else
 * LUMX_EINVAL, with size_out set: call once with NULL to size it); models
 * are at most 1024 voxels a side and 2 GiB dense in total; materials (256 x
 * {r, g, b, a, emission, roughness}) is written when non-NULL;
 * n_models_out (optional) gets the number of distinct non-empty models. */
int lumx_vox_parse(const void* data, size_t bytes, int* size_out, uint8_t* voxels, size_t voxels_capacity,
                   float* materials, int* n_models_out);
/* ---- world file ("scene") ----
 * load_scene (src/renderer/src/load_stuff.cpp:187-217): a file image of a
 * 12-byte int32 header (stored size Sx, Sy, Sz) and int16 blocks indexed
 * x + Sx * y + Sx * Sy * z.  Into blocks (x-fastest, world_size[0..2]) it
 * writes the stored blocks of the region clamp(S, 0, world_size), each
 * clamped to [0, static_blocks]; blocks outside that region keep their
 * values.  Fewer than 12 bytes clear the whole world.  A file holding fewer
 * blocks than its header's region needs is LUMX_EINVAL (the reference
 * asserts).  Host only; lum_load_world_file reads files with it. */
int lumx_scene_read(const void* data, size_t bytes, const int world_size[3], int static_blocks, int16_t* blocks);
/* Lum's default camera (internal_types.hpp:54-66) for a frame size. */
int lumx_camera_default(int frame_w, int frame_h, lumx_camera* out);

/* ---- batches: many worlds per launch (no reference counterpart: the
 * reference renders one world per frame) ---- */
/* Worlds (distinct, one radiance format) updated together. */
int lumx_batch_create(lumx_context_t ctx, int n, const lumx_world_t* worlds, lumx_batch_t* out);
int lumx_batch_destroy(lumx_batch_t b);
/* One radiance update of every world in one launch: requests_dev[i] /
 * n_requests[i] as for lumx_radiance_update (requests_dev NULL = every probe
 * of every world), params[i] per world (rays a power of two).  Each world's
 * result equals its own lumx_radiance_update with the same arguments. */
int lumx_batch_radiance(lumx_batch_t b, const int32_t* const* requests_dev, const int* n_requests,
                        const lumx_radiance_params* params, void* stream);
/* Autotuner of the batched radiance launch: times every launch shape
 * (workgroups of 256, 512, 1024 threads x 1, 2, 4 rays per lane) on the
 * launch lumx_batch_radiance would make with these arguments (requests_dev
 * NULL = every probe of every world), keeps the fastest for this batch's
 * later launches and returns it.  A shape other than the default (256
 * threads, 1 ray per lane) is kept only when it is at least 3% faster, as
 * the trial stands for launches of other sizes.  The trial launches write the free field
 * buffers only: every world's field is left unchanged.  No reference
 * counterpart. */
int lumx_batch_autotune(lumx_batch_t b, const int32_t* const* requests_dev, const int* n_requests,
                        const lumx_radiance_params* params, int reps, int* best_block, int* best_rays_per_lane,
                        float* best_ms);
/* Pins the batched radiance launch shape: workgroup size 256, 512 or 1024
 * (0 = 256) and rays per lane 1, 2 or 4 (0 = 1; a launch whose smallest
 * rays per probe is below it uses 1).  Results do not depend on either. */
int lumx_batch_set_shape(lumx_batch_t b, int block, int rays_per_lane);
/* Map of every world of the batch: world i gets n_models[i] models
 * models[i][0 ..] at placements[i][0 ..] (models[i] / placements[i] may be
 * NULL when n_models[i] is 0).  Blockify runs on the host as for
 * lumx_map_models_placed, then all worlds go up in one copy and are
 * committed, copied and mapped by a handful of launches shared by every
 * world.  Each world's result equals its own lumx_map_models_placed with the
 * same arguments, bit for bit.  When world k's palette overflows, worlds
 * 0 .. k-1 are mapped, world k is back at its origin, later worlds are
 * untouched, and the overflow is returned (as k + 1 single-world calls would
 * leave them).  Inside a graph capture the blockified worlds are those of
 * capture time, uploaded again on every replay.  Batched maps set no
 * per-world map timing (lumx_world_last_pass_ms). */
int lumx_batch_map_models(lumx_batch_t b, const int* n_models, const lumx_model_t* const* models,
                          const lumx_placement* const* placements, void* stream);
/* Glossy pass of every world in one launch (one G-buffer and camera per world). */
int lumx_batch_glossy(lumx_batch_t b, const uint8_t* const* matnorm_dev, const float* const* depth_dev,
                      const int* width, const int* height, const lumx_camera* cams, float max_dist, int max_steps,
                      float* const* out_dev, void* stream);
/* Device frame counter added to every time_seed of this batch's radiance
 * updates (0 at creation), so a replayed graph advances the ray rotation. */
int lumx_batch_set_frame(lumx_batch_t b, int frame, void* stream);
int lumx_batch_advance_frame(lumx_batch_t b, void* stream);

/* ---- serving: continuous batching over resident worlds (no reference
 * counterpart: the reference renders one world per frame) ----
 * A worker thread admits submitted requests in arrival order, at most one
 * per world per launch and in submission order per world, restoring evicted
 * worlds (and evicting the idle ones the pending requests need last) to stay
 * within the HBM budget.  With a budget, max_batch >= 4 and more than two
 * launches' worth queued, a launch first takes up to three quarters of its requests from
 * worlds already resident among the next sixteen launches' worth of the queue,
 * then the rest in arrival order (each world's requests still in submission
 * order).  It packs up to max_batch requests into one batched radiance launch
 * plus one batched glossy launch, and keeps up to `streams` launches in
 * flight while it admits the next.  Each request computes exactly what its
 * own lumx_radiance_update / lumx_glossy call would.  Worlds added to a
 * server belong to it until it is destroyed.
 * With an HBM budget the server allocates one device arena of the budget's
 * size at creation and every world lives there or in a pinned host snapshot:
 * lumx_server_add_world moves a resident world into the arena while it has
 * room (copying its static state to its snapshot, so evictions while serving
 * copy only its field), else to the host; a world larger than the budget is
 * LUMX_EINVAL.  Evictions and restores are copies on the server's own stream,
 * never a host wait.  At destroy, worlds in the arena
 * move into device blocks of their own; evicted ones stay on the host
 * (lumx_world_restore brings them back). */
#define LUMX_REQ_RADIANCE 0
#define LUMX_REQ_GLOSSY 1

typedef struct lumx_server* lumx_server_t;

typedef struct {
    size_t hbm_budget;  /* bytes of resident world state, 0 = no limit */
    int max_batch;      /* requests per launch; 1 = one request per launch */
    int streams;        /* launches in flight at once (>= 1) */
    int block;          /* radiance workgroup size: 0 (= 256), 256, 512 or 1024 */
    int rays_per_lane;  /* rays per thread of the radiance launches: 0 (= 1), 1, 2 or 4 */
    int direct;         /* 1: the plain engine - every request is its own
                         * lumx_radiance_update / lumx_glossy call, queued in
                         * arrival order on one stream as soon as its world is
                         * resident, up to 256 calls deep, with no wait between
                         * them (max_batch, streams, block and rays_per_lane are
                         * ignored) */
} lumx_server_desc;

typedef struct {
    int kind;                    /* LUMX_REQ_RADIANCE or LUMX_REQ_GLOSSY */
    int world;                   /* id from lumx_server_add_world */
    const int32_t* requests_dev; /* radiance: probe list (int32x4) or NULL = every probe */
    int n_requests;
    lumx_radiance_params params; /* radiance (rays a power of two) */
    const uint8_t* matnorm_dev;  /* glossy: G-buffer, camera, output */
    const float* depth_dev;
    int width, height;
    lumx_camera cam;
    float max_dist;
    int max_steps;
    float* out_dev;
} lumx_request;

typedef struct {
    long long submitted, completed, launches, probe_rays, pixels, evictions, restores;
    size_t resident_bytes;
} lumx_server_stats;

int lumx_server_create(lumx_context_t ctx, const lumx_server_desc* d, lumx_server_t* out);
int lumx_server_destroy(lumx_server_t s);
int lumx_server_add_world(lumx_server_t s, lumx_world_t w, int* id);
/* Thread-safe.  The ticket identifies the request in lumx_server_wait. */
int lumx_server_submit(lumx_server_t s, const lumx_request* r, long long* ticket);
/* Blocks until the request finished; latency = completion - submission. */
int lumx_server_wait(lumx_server_t s, long long ticket, double* latency_s);
int lumx_server_drain(lumx_server_t s);
int lumx_server_get_stats(lumx_server_t s, lumx_server_stats* out);
/* Open-loop load: submits reqs[i] at arrival_s[i] seconds after the call
 * (from a submitter thread), waits for all of them and returns each latency
 * and the wall time from the first arrival to the last completion. */
int lumx_server_run_trace(lumx_server_t s, int n, const lumx_request* reqs, const double* arrival_s,
                          double* latency_s, double* wall_s);

/* ---- graphs (hipGraph capture of any sequence of passes on a stream) ---- */
int lumx_graph_begin(void* stream);
int lumx_graph_end(void* stream, lumx_graph_t* out);
int lumx_graph_launch(lumx_graph_t g, void* stream);
int lumx_graph_destroy(lumx_graph_t g);

/* ---- timing ---- */
/* Time (ms) of the last pass of each kind on this world (0 radiance, 1 map, 2 glossy),
 * measured with HIP events on the pass's stream; valid after a sync. */
int lumx_world_last_pass_ms(lumx_world_t w, int pass, float* ms);

#ifdef __cplusplus
}
#endif
#endif /* LUMX_H */
