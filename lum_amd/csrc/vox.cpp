// vox.cpp - MagicaVoxel .vox import for meshes, blocks and material palettes.
//
// The reference reads .vox files through ogt_vox (common/ogt_vox.hpp,
// vox_read_scene) in load_mesh, load_block and extract_palette
// (src/renderer/src/load_stuff.cpp:233-342) and turns the palette and the
// MATL/MATT materials into its Material table in extract_palette_mem
// (load_stuff.cpp:233-272).  This reader keeps the parts of that behaviour
// the lighting path consumes:
//   * the file is a "VOX " header (version 150 or 200) followed by a flat
//     sequence of chunks {id, size, children size}; MAIN's children follow
//     it directly, chunks this reader does not use are skipped by size;
//   * each SIZE + XYZI pair is a model (x-fastest colour indices, 0 = empty);
//     a model with no voxels is dropped, a model identical to an earlier one
//     is dropped, and the first remaining model is the one a mesh or block
//     gets;
//   * the palette is MagicaVoxel's default one unless an RGBA chunk replaces
//     it; an IMAP chunk reorders palette, materials and voxel indices into
//     display order; the palette is then rotated by one so colour index i
//     reads entry i (entry 0, the empty colour, gets alpha 0);
//   * MATL (dictionary: _type, _metal, _rough, _emit, _flux) and legacy MATT
//     chunks set per-index material type and weights;
//   * the Material table: colour = rgba / 256, and by type: diffuse ->
//     roughness 1; emit -> emission = emit * (2 + 4 * flux), roughness 0.5;
//     metal -> roughness (rough + 1 - metal) / 2; other types -> 0.
// Host code only; no GPU work.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "host_internal.hpp"

namespace {

inline uint32_t chunk_id(char a, char b, char c, char d) {
    return (uint32_t)(uint8_t)a | ((uint32_t)(uint8_t)b << 8) | ((uint32_t)(uint8_t)c << 16) |
           ((uint32_t)(uint8_t)d << 24);
}

struct Reader {
    const uint8_t* p;
    size_t n, off = 0;
    size_t left() const { return n - off; }
    bool u32(uint32_t* v) {
        if (left() < 4) {
            off = n;
            return false;
        }
        std::memcpy(v, p + off, 4);
        off += 4;
        return true;
    }
    bool bytes(void* dst, size_t k) {
        if (left() < k) {
            off = n;
            return false;
        }
        std::memcpy(dst, p + off, k);
        off += k;
        return true;
    }
    void skip(size_t k) { off += k < left() ? k : left(); }
    bool str(std::string* s) {
        uint32_t len = 0;
        if (!u32(&len) || left() < len) {
            off = n;
            return false;
        }
        s->assign((const char*)p + off, len);
        off += len;
        return true;
    }
};

enum MatType { kDiffuse = 0, kMetal = 1, kGlass = 2, kEmit = 3, kBlend = 4, kMedia = 5 };

struct Matl {
    int type = kDiffuse;
    float metal = 0.0f, rough = 0.0f, emit = 0.0f, flux = 0.0f, trans = 0.0f;
};

struct Model {
    uint32_t sx, sy, sz;
    std::vector<uint8_t> v;
};

// MagicaVoxel's default palette as stored in a file (entry k is colour
// index k + 1): the 6x6x6 cube of {ff, cc, 99, 66, 33, 00} per channel with
// red slowest and black left out, then ten-step ramps of red, green, blue and
// grey, then black.
void default_palette(uint8_t pal[256][4]) {
    static const uint8_t cube[6] = {0xff, 0xcc, 0x99, 0x66, 0x33, 0x00};
    static const uint8_t ramp[10] = {0xee, 0xdd, 0xbb, 0xaa, 0x88, 0x77, 0x55, 0x44, 0x22, 0x11};
    int k = 0;
    for (int r = 0; r < 6; ++r)
        for (int g = 0; g < 6; ++g)
            for (int b = 0; b < 6; ++b) {
                if (r == 5 && g == 5 && b == 5) continue;
                pal[k][0] = cube[r];
                pal[k][1] = cube[g];
                pal[k][2] = cube[b];
                pal[k][3] = 0xff;
                ++k;
            }
    for (int c = 0; c < 4; ++c)
        for (int i = 0; i < 10; ++i) {
            for (int ch = 0; ch < 3; ++ch) pal[k][ch] = (c == 3 || c == ch) ? ramp[i] : 0;
            pal[k][3] = 0xff;
            ++k;
        }
    pal[255][0] = pal[255][1] = pal[255][2] = 0;
    pal[255][3] = 0xff;
}

bool read_dict(Reader& r, std::vector<std::pair<std::string, std::string>>* d) {
    uint32_t n = 0;
    if (!r.u32(&n)) return false;
    d->clear();
    for (uint32_t i = 0; i < n; ++i) {
        std::string k, v;
        if (!r.str(&k) || !r.str(&v)) return false;
        d->emplace_back(std::move(k), std::move(v));
    }
    return true;
}

const std::string* dict_get(const std::vector<std::pair<std::string, std::string>>& d, const char* key) {
    for (const auto& kv : d)
        if (kv.first == key) return &kv.second;
    return nullptr;
}

float to_float(const std::string& s) { return (float)std::strtod(s.c_str(), nullptr); }

}  // namespace

extern "C" int lumx_vox_parse(const void* data, size_t bytes, int* size_out, uint8_t* voxels, size_t voxels_capacity,
                              float* materials, int* n_models_out) {
    if (!data || !size_out) return lumx::fail(LUMX_EINVAL, "null argument");
    Reader r{(const uint8_t*)data, bytes};
    uint32_t magic = 0, version = 0;
    r.u32(&magic);
    r.u32(&version);
    if (magic != chunk_id('V', 'O', 'X', ' ') || (version != 150 && version != 200))
        return lumx::fail(LUMX_EINVAL, "not a .vox file (header %08x, version %u)", magic, version);

    uint8_t pal[256][4];
    default_palette(pal);
    Matl matl[256];
    uint8_t imap[256];
    bool have_imap = false;
    std::vector<Model> models;   // empty voxel vector = a model with no voxels
    uint32_t sx = 0, sy = 0, sz = 0;
    size_t model_bytes = 0;      // all dense models held so far
    std::vector<std::pair<std::string, std::string>> dict;

    while (r.left() >= 12) {
        uint32_t id = 0, size = 0, child = 0;
        r.u32(&id);
        r.u32(&size);
        r.u32(&child);
        if (id == chunk_id('M', 'A', 'I', 'N')) {
            continue;
        } else if (id == chunk_id('S', 'I', 'Z', 'E')) {
            r.u32(&sx);
            r.u32(&sy);
            r.u32(&sz);
        } else if (id == chunk_id('X', 'Y', 'Z', 'I')) {
            if (!(sx && sy && sz)) return lumx::fail(LUMX_EINVAL, ".vox: XYZI chunk before a SIZE chunk");
            // Models are at most 1024 voxels a side (the largest a lumx model
            // takes), and all of a file's models together at most 2 GiB dense.
            if (sx > 1024 || sy > 1024 || sz > 1024)
                return lumx::fail(LUMX_EINVAL, ".vox: model %ux%ux%u larger than 1024 a side", sx, sy, sz);
            uint32_t n = 0;
            r.u32(&n);
            Model m{sx, sy, sz, {}};
            if (n != 0) {
                model_bytes += (size_t)sx * sy * sz;
                if (model_bytes > (1ull << 31)) return lumx::fail(LUMX_EINVAL, ".vox: models too large in total");
                m.v.assign((size_t)sx * sy * sz, 0);
                const size_t avail = r.left() / 4 < n ? r.left() / 4 : n;
                const uint8_t* q = r.p + r.off;
                for (size_t i = 0; i < avail; ++i) {
                    const uint32_t x = q[4 * i], y = q[4 * i + 1], z = q[4 * i + 2];
                    if (x < sx && y < sy && z < sz) m.v[x + (size_t)sx * (y + (size_t)sy * z)] = q[4 * i + 3];
                }
                r.skip((size_t)n * 4);
            }
            models.push_back(std::move(m));
        } else if (id == chunk_id('R', 'G', 'B', 'A')) {
            r.bytes(pal, sizeof(pal));
        } else if (id == chunk_id('I', 'M', 'A', 'P')) {
            if (r.bytes(imap, 256)) have_imap = true;
        } else if (id == chunk_id('M', 'A', 'T', 'L')) {
            const size_t end = r.off + size;
            uint32_t mid = 0;
            r.u32(&mid);
            Matl& m = matl[mid & 0xff];
            if (read_dict(r, &dict)) {
                if (const std::string* t = dict_get(dict, "_type")) {
                    if (*t == "_diffuse") m.type = kDiffuse;
                    else if (*t == "_metal") m.type = kMetal;
                    else if (*t == "_glass") m.type = kGlass;
                    else if (*t == "_emit") m.type = kEmit;
                    else if (*t == "_blend") m.type = kBlend;
                    else if (*t == "_media") m.type = kMedia;
                }
                if (const std::string* s = dict_get(dict, "_metal")) m.metal = to_float(*s);
                if (const std::string* s = dict_get(dict, "_rough")) m.rough = to_float(*s);
                if (const std::string* s = dict_get(dict, "_flux")) m.flux = to_float(*s);
                if (const std::string* s = dict_get(dict, "_emit")) m.emit = to_float(*s);
                if (const std::string* s = dict_get(dict, "_trans")) m.trans = to_float(*s);
            }
            if (r.off < end) r.skip(end - r.off);
        } else if (id == chunk_id('M', 'A', 'T', 'T')) {
            uint32_t mid = 0, type = 0, bits = 0;
            float weight = 0.0f;
            r.u32(&mid);
            r.u32(&type);
            r.bytes(&weight, 4);
            r.u32(&bits);
            Matl& m = matl[mid & 0xff];
            m.type = (int)type;
            if (type == kMetal) m.metal = weight;
            else if (type == kGlass) m.trans = weight;
            else if (type == kEmit) m.emit = weight;
            r.skip(size >= 16 ? size - 16 : 0);
        } else {
            r.skip(size);
        }
    }

    if (have_imap) {
        uint8_t inverse[256];
        for (int i = 0; i < 256; ++i) inverse[imap[i]] = (uint8_t)i;
        uint8_t old_pal[256][4];
        std::memcpy(old_pal, pal, sizeof(pal));
        for (int i = 0; i < 256; ++i) std::memcpy(pal[i], old_pal[(imap[i] + 255) & 0xff], 4);
        Matl old_matl[256];
        for (int i = 0; i < 256; ++i) old_matl[i] = matl[i];
        for (int i = 0; i < 256; ++i) matl[i] = old_matl[imap[(i + 255) & 0xff]];
        for (Model& m : models)
            for (uint8_t& v : m.v) v = (uint8_t)(1 + inverse[v]);
    }
    {
        uint8_t last[4];
        std::memcpy(last, pal[255], 4);
        for (int i = 255; i > 0; --i) std::memcpy(pal[i], pal[i - 1], 4);
        std::memcpy(pal[0], last, 4);
        pal[0][3] = 0;
    }

    // Distinct non-empty models, in file order.
    std::vector<const Model*> kept;
    for (const Model& m : models) {
        if (m.v.empty()) continue;
        bool dup = false;
        for (const Model* k : kept)
            if (k->sx == m.sx && k->sy == m.sy && k->sz == m.sz && k->v == m.v) {
                dup = true;
                break;
            }
        if (!dup) kept.push_back(&m);
    }
    if (n_models_out) *n_models_out = (int)kept.size();
    if (materials) {
        for (int i = 0; i < 256; ++i) {
            float* o = materials + 6 * i;
            for (int c = 0; c < 4; ++c) o[c] = (float)(pal[i][c] / 256.0);
            float emit = 0.0f, rough;
            const Matl& m = matl[i];
            switch (m.type) {
                case kDiffuse: rough = 1.0f; break;
                case kEmit:
                    emit = (float)(m.emit * (2.0 + m.flux * 4.0));
                    rough = 0.5f;
                    break;
                case kMetal: rough = (float)((m.rough + (1.0 - m.metal)) / 2.0); break;
                default: rough = 0.0f; break;
            }
            o[4] = emit;
            o[5] = rough;
        }
    }
    if (kept.empty()) {
        size_out[0] = size_out[1] = size_out[2] = 0;
        return lumx::fail(LUMX_EINVAL, ".vox: no model with voxels");
    }
    const Model& m0 = *kept[0];
    size_out[0] = (int)m0.sx;
    size_out[1] = (int)m0.sy;
    size_out[2] = (int)m0.sz;
    if (voxels) {
        if (voxels_capacity < m0.v.size())
            return lumx::fail(LUMX_EINVAL, ".vox: voxel buffer holds %zu bytes, the model needs %zu", voxels_capacity,
                              m0.v.size());
        std::memcpy(voxels, m0.v.data(), m0.v.size());
    }
    return LUMX_OK;
}
