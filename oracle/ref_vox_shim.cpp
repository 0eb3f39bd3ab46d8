// ref_vox_shim.cpp - TEST INFRASTRUCTURE ONLY.  Wraps the reference's own
// .vox reader (common/ogt_vox.hpp, compiled from the reference tree by
// oracle/Makefile into oracle/_ref/) behind a C entry point with the same
// outputs as lumx_vox_parse, so tests can compare the product's reader with
// the reference's on the same files.  The Material table follows Lum's
// extract_palette_mem (src/renderer/src/load_stuff.cpp:233-272); the model is
// scene->models[0] as in load_mesh / load_block (load_stuff.cpp:288-342).
#include <cstdint>
#include <cstring>

#include "ogt_vox.hpp"

extern "C" int ref_vox_parse(const void* data, size_t bytes, int* size_out, uint8_t* voxels,
                             size_t voxels_capacity, float* materials, int* n_models_out) {
    const ogt::vox_scene* scene = ogt::vox_read_scene((const uint8_t*)data, (uint32_t)bytes);
    if (!scene) return -1;
    if (n_models_out) *n_models_out = (int)scene->num_models;
    if (materials) {
        for (int i = 0; i < 256; ++i) {
            float* o = materials + 6 * i;
            o[0] = (float)(scene->palette.color[i].r / 256.0);
            o[1] = (float)(scene->palette.color[i].g / 256.0);
            o[2] = (float)(scene->palette.color[i].b / 256.0);
            o[3] = (float)(scene->palette.color[i].a / 256.0);
            float emit = 0.0f, rough;
            const ogt::vox_matl& m = scene->materials.matl[i];
            switch (m.type) {
                case ogt::matl_type_diffuse: rough = 1.0f; break;
                case ogt::matl_type_emit:
                    emit = (float)(m.emit * (2.0 + m.flux * 4.0));
                    rough = 0.5f;
                    break;
                case ogt::matl_type_metal: rough = (float)((m.rough + (1.0 - m.metal)) / 2.0); break;
                default: rough = 0.0f; break;
            }
            o[4] = emit;
            o[5] = rough;
        }
    }
    int rc = 0;
    if (scene->num_models == 0) {
        size_out[0] = size_out[1] = size_out[2] = 0;
        rc = -1;
    } else {
        const ogt::vox_model* m = scene->models[0];
        size_out[0] = (int)m->size_x;
        size_out[1] = (int)m->size_y;
        size_out[2] = (int)m->size_z;
        const size_t n = (size_t)m->size_x * m->size_y * m->size_z;
        if (voxels && voxels_capacity >= n) std::memcpy(voxels, m->voxel_data, n);
    }
    ogt::vox_destroy_scene(scene);
    return rc;
}
