// march.hpp - device-side voxel march shared by the radiance and glossy
// kernels (gfx950).
//
// Semantics follow the reference shaders exactly (paths relative to the
// reference repository):
//   GetBlock / voxel_in_palette / GetVoxel  shaders/radiance.comp:97-143
//   GetMat                                  shaders/radiance.comp:145-157
//   sample_radiance                         shaders/radiance.comp:161-166
//   (sample_texel also serves diffuse.frag:92-97)
//   CastRay_fast                            shaders/radiance.comp:183-291
// Arithmetic rules: no FMA contraction (the pragma below and
// -ffp-contract=off), IEEE division and sqrt, fminf/fmaxf, float->int
// truncation toward zero, x % 16 == x & 15 (sign of the divisor), block
// fetches outside the world return 0 and palette slots outside
// [0, palette_slots) read as voxel 0.
//
// HBM layout consumed here:
//   blocks   int16[wz][wy][wx]                       (2 B per block)
//   palette  uint8[slot][z][y][x], 4096 B per slot    (slot-contiguous)
//   mats     float[256][6] = {r, g, b, a, emission, roughness}
//   field    texel[wz][wy][wx], texel = float4 / half4 / bf16x4
#pragma once
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <stdint.h>

// Every helper here compiles for the host too, so the kernels' arithmetic can
// be exercised on the CPU from the same source.
#define LUMX_HD __host__ __device__ __forceinline__

// Every buffer the kernels read or write lives in global memory.  Pointers
// taken from a descriptor in memory (the batched kernels) are otherwise
// generic and compile to flat accesses; these helpers make them global.
#if defined(__HIP_DEVICE_COMPILE__)
#define LUMX_GLOBAL __attribute__((address_space(1)))
#else
#define LUMX_GLOBAL
#endif
template <class T>
LUMX_HD T gld(const T* p, size_t i) {
    return ((const LUMX_GLOBAL T*)p)[i];
}
template <class T>
LUMX_HD void gst(T* p, size_t i, T v) {
    ((LUMX_GLOBAL T*)p)[i] = v;
}

namespace lumx {

struct f3 {
    float x, y, z;
};

LUMX_HD f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
LUMX_HD f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
LUMX_HD f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
LUMX_HD f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
LUMX_HD f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
LUMX_HD f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
LUMX_HD f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
LUMX_HD float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
LUMX_HD f3 normalize3(f3 v) {
    float len = sqrtf(dot3(v, v));
    return {v.x / len, v.y / len, v.z / len};
}

struct WorldView {
    const int16_t* blocks;
    const uint8_t* palette;
    const float* mats;
    int wx, wy, wz;
    int palette_slots;
    int block0_empty;  // palette slot 0 is all zero: skip its voxel fetch
};

// store_probe's image is A2B10G10R10 UNORM in the reference, whose store
// turns a NaN component into 0; the float stores here keep that rule.
LUMX_HD f3 unorm_nan_to_zero(f3 v) {
    return {v.x != v.x ? 0.0f : v.x, v.y != v.y ? 0.0f : v.y, v.z != v.z ? 0.0f : v.z};
}

// Radiance texel storage formats.
enum { STORE_F32 = 0, STORE_F16 = 1, STORE_BF16 = 2 };

LUMX_HD float bf16_to_f32(uint16_t h) { return __builtin_bit_cast(float, ((uint32_t)h) << 16); }
LUMX_HD uint16_t f32_to_bf16(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

template <int STORE> struct Texel;
template <> struct Texel<STORE_F32> {
    typedef float4 type;
    LUMX_HD static f3 load(const void* base, unsigned i) {
        // 32-bit byte offsets: a field holds at most 2^28 texels of 16 B.
        float4 v = gld((const float4*)((const char*)base + (i << 4)), 0);
        return {v.x, v.y, v.z};
    }
    LUMX_HD static void store(void* base, unsigned i, f3 v) {
        gst((float4*)((char*)base + (i << 4)), 0, make_float4(v.x, v.y, v.z, 1.0f));
    }
};
template <> struct Texel<STORE_F16> {
    LUMX_HD static f3 load(const void* base, unsigned i) {
        uint2 raw = gld((const uint2*)((const char*)base + (i << 3)), 0);
        _Float16 h[4];
        __builtin_memcpy(h, &raw, 8);
        return {(float)h[0], (float)h[1], (float)h[2]};
    }
    LUMX_HD static void store(void* base, unsigned i, f3 v) {
        _Float16 h[4] = {(_Float16)v.x, (_Float16)v.y, (_Float16)v.z, (_Float16)1.0f};
        uint2 raw;
        __builtin_memcpy(&raw, h, 8);
        gst((uint2*)((char*)base + (i << 3)), 0, raw);
    }
};
template <> struct Texel<STORE_BF16> {
    LUMX_HD static f3 load(const void* base, unsigned i) {
        uint2 raw = gld((const uint2*)((const char*)base + (i << 3)), 0);
        return {bf16_to_f32((uint16_t)(raw.x & 0xffffu)), bf16_to_f32((uint16_t)(raw.x >> 16)),
                bf16_to_f32((uint16_t)(raw.y & 0xffffu))};
    }
    LUMX_HD static void store(void* base, unsigned i, f3 v) {
        uint2 raw;
        raw.x = (uint32_t)f32_to_bf16(v.x) | ((uint32_t)f32_to_bf16(v.y) << 16);
        raw.y = (uint32_t)f32_to_bf16(v.z) | ((uint32_t)0x3f80u << 16);
        gst((uint2*)((char*)base + (i << 3)), 0, raw);
    }
};

// a * b for operands below 2^24 (a full-rate 24-bit multiply on the GPU).
LUMX_HD unsigned umin_(unsigned a, unsigned b) { return a < b ? a : b; }

LUMX_HD unsigned mul24(unsigned a, unsigned b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul24(a, b);
#else
    return a * b;
#endif
}

// Division of n < 2^28 by an invariant d in [1, 2^24] as a multiply-high
// and a shift (division by invariant integers, Granlund & Montgomery 1994):
// k = max(32, 28 + ceil(log2 d)), m = ceil(2^k / d), n / d = (n * m) >> k.
// With e = m * d - 2^k < d, n * e < 2^28 * d <= 2^k keeps every quotient
// exact.  d = 1 is m = 0 (n itself).
struct Divisor {
    uint32_t m;
    uint32_t shift;  // k - 32
};
inline Divisor make_divisor(uint32_t d) {
    if (d <= 1) return {0u, 0u};
    uint32_t L = 0;
    while ((1ull << L) < d) ++L;
    const uint32_t k = 28 + L > 32 ? 28 + L : 32;
    const unsigned long long m = ((1ull << k) + d - 1) / d;
    return {(uint32_t)m, k - 32};
}
LUMX_HD uint32_t fast_div(uint32_t n, Divisor v) {
    if (v.m == 0) return n;
#if defined(__HIP_DEVICE_COMPILE__)
    return __umulhi(n, v.m) >> v.shift;
#else
    return (uint32_t)(((unsigned long long)n * v.m) >> 32) >> v.shift;
#endif
}

// Block id at in-bounds block coordinates.  Every operand is below 2^24
// (axes <= 4096, wy * wz <= 2^24) and the byte offset is formed in 32 bits
// (worlds hold at most 2^28 blocks), so the load takes a scalar base plus a
// 32-bit vector offset.
LUMX_HD int block_id(const WorldView& w, int bx, int by, int bz) {
    const unsigned lin = (unsigned)bx + mul24((unsigned)w.wx, (unsigned)by + mul24((unsigned)w.wy, (unsigned)bz));
    return gld((const int16_t*)((const char*)w.blocks + (lin << 1)), 0);
}

// voxel_in_palette: voxel (ix, iy, iz) mod 16 of block b.
LUMX_HD int voxel_in_block(const WorldView& w, int b, int ix, int iy, int iz) {
    if (b == 0 && w.block0_empty) return 0;
    if ((unsigned)b >= (unsigned)w.palette_slots) return 0;
    uint32_t off = ((uint32_t)b << 12) | (uint32_t)((ix & 15) | ((iy & 15) << 4) | ((iz & 15) << 8));
    return gld(w.palette, off);
}

// GetBlock + palette fetch for an integer voxel position.
// The block coordinate i / 16 (truncating) lies in [0, W) exactly when i lies
// in [-15, 16 W), and there it equals max(i, 0) >> 4: one add and compare for
// the bounds, a max and a shift for the fetch, instead of a signed division.
LUMX_HD bool block_axis_in(int i, int w) { return (unsigned)i + 15u < (unsigned)(16 * w + 15); }
LUMX_HD int block_axis(int i) { return (i > 0 ? i : 0) >> 4; }

LUMX_HD int get_voxel(const WorldView& w, int ix, int iy, int iz, int* blk_out) {
    int b = 0;
    if (block_axis_in(ix, w.wx) && block_axis_in(iy, w.wy) && block_axis_in(iz, w.wz))
        b = block_id(w, block_axis(ix), block_axis(iy), block_axis(iz));
    *blk_out = b;
    return voxel_in_block(w, b, ix, iy, iz);
}

struct MatF {
    f3 color;
    float emit, rough;
};
LUMX_HD MatF get_mat(const WorldView& w, int v) {
    const unsigned o = (unsigned)v * 6u;
    MatF r;
    r.color = {gld(w.mats, o), gld(w.mats, o + 1), gld(w.mats, o + 2)};
    r.emit = gld(w.mats, o + 4);
    r.rough = gld(w.mats, o + 5);
    return r;
}

// i mod n in [0, n) for n in [1, 2^30).  For i in [-n, 2n), u = i + n lies
// in [0, 3n) and two conditional subtracts finish it; min(u, u - n) over
// unsigned values is u - n exactly when u >= n.  Other i divide.
LUMX_HD int wrap_i(int i, int n) {
    unsigned u = (unsigned)i + (unsigned)n;
    if (u >= 3u * (unsigned)n) {
        const int r = i % n;
        return r < 0 ? r + n : r;
    }
    u = umin_(u, u - (unsigned)n);
    return (int)umin_(u, u - (unsigned)n);
}

// (x + 1) mod n for x in [0, n).
LUMX_HD int wrap_next(int x, int n) {
    const unsigned u = (unsigned)x + 1u;
    return (int)umin_(u, u - (unsigned)n);
}

// Trilinear LINEAR filter with REPEAT wrap at texel-space coordinate bp
// (texel i's centre at bp = i, i.e. normalized coordinate (bp + 0.5) / size).
template <int STORE>
LUMX_HD f3 sample_texel(const void* field, int wx, int wy, int wz, f3 bp) {
    float fx0 = floorf(bp.x), fy0 = floorf(bp.y), fz0 = floorf(bp.z);
    float tx = bp.x - fx0, ty = bp.y - fy0, tz = bp.z - fz0;
    int ix = (int)fx0, iy = (int)fy0, iz = (int)fz0;
    // REPEAT: the second tap of each axis is the first one's successor mod n.
    int x0 = wrap_i(ix, wx), x1 = wrap_next(x0, wx);
    int y0 = wrap_i(iy, wy), y1 = wrap_next(y0, wy);
    int z0 = wrap_i(iz, wz), z1 = wrap_next(z0, wz);
    // Operands below 2^24 (axes <= 4096, wy * wz <= 2^24): 24-bit multiplies.
    const unsigned r00 = mul24((unsigned)wx, (unsigned)y0 + mul24((unsigned)wy, (unsigned)z0));
    const unsigned r10 = mul24((unsigned)wx, (unsigned)y1 + mul24((unsigned)wy, (unsigned)z0));
    const unsigned r01 = mul24((unsigned)wx, (unsigned)y0 + mul24((unsigned)wy, (unsigned)z1));
    const unsigned r11 = mul24((unsigned)wx, (unsigned)y1 + mul24((unsigned)wy, (unsigned)z1));
    f3 t000 = Texel<STORE>::load(field, r00 + (unsigned)x0);
    f3 t100 = Texel<STORE>::load(field, r00 + (unsigned)x1);
    f3 t010 = Texel<STORE>::load(field, r10 + (unsigned)x0);
    f3 t110 = Texel<STORE>::load(field, r10 + (unsigned)x1);
    f3 t001 = Texel<STORE>::load(field, r01 + (unsigned)x0);
    f3 t101 = Texel<STORE>::load(field, r01 + (unsigned)x1);
    f3 t011 = Texel<STORE>::load(field, r11 + (unsigned)x0);
    f3 t111 = Texel<STORE>::load(field, r11 + (unsigned)x1);
    float ux = 1.0f - tx, uy = 1.0f - ty, uz = 1.0f - tz;
    f3 c00 = t000 * ux + t100 * tx;
    f3 c10 = t010 * ux + t110 * tx;
    f3 c01 = t001 * ux + t101 * tx;
    f3 c11 = t011 * ux + t111 * tx;
    f3 c0 = c00 * uy + c10 * ty;
    f3 c1 = c01 * uy + c11 * ty;
    return c0 * uz + c1 * tz;
}

// sample_radiance: trilinear at texel coordinate (x + n*16)/16, REPEAT wrap.
template <int STORE>
LUMX_HD f3 sample_radiance(const void* field, int wx, int wy, int wz, f3 x, f3 n) {
    return sample_texel<STORE>(field, wx, wy, wz, (x + n * 16.0f) / 16.0f);
}

struct RayHit {
    int hit;
    int voxel;
    float fraction;
    f3 normal;
    int coarse;
    int refine;
    int key;
};

// A ray of CastRay_fast with the values its march reuses every sample.
struct MarchRay {
    f3 o, d, inv, corner;
};
LUMX_HD MarchRay march_ray(f3 o, f3 d) {
    MarchRay m;
    m.o = o;
    m.d = d;
    m.inv = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    m.corner = {d.x > 0.0f ? 16.0f : 0.0f, d.y > 0.0f ? 16.0f : 0.0f, d.z > 0.0f ? 16.0f : 0.0f};
    return m;
}

// One coarse sample of CastRay_fast (radiance.comp:205-244): fraction += 0.5,
// the voxel there, the empty-block skip.  Returns 1 on a hit, -1 when the ray
// leaves (outside the world, past max_dist, or sample `it` reached
// max_steps), 0 to go on.
LUMX_HD int coarse_sample(const WorldView& w, const MarchRay& m, float max_dist, int max_steps, int it,
                          float* fraction_io) {
    float fraction = *fraction_io + 0.5f;
    f3 pos = m.o + m.d * fraction;
    const int ix = (int)pos.x, iy = (int)pos.y, iz = (int)pos.z;
    // Block coordinate ivec3(p) / 16 (truncating) as trunc(p / 16): the
    // same integer, p / 16 being exact, without the signed division.
    const f3 ps = pos * 0.0625f;
    const int bx = (int)ps.x, by = (int)ps.y, bz = (int)ps.z;
    const bool inb = (unsigned)bx < (unsigned)w.wx && (unsigned)by < (unsigned)w.wy && (unsigned)bz < (unsigned)w.wz;
    const int blk = inb ? block_id(w, bx, by, bz) : 0;
    const int cur = voxel_in_block(w, blk, ix, iy, iz);
    if (blk == 0) {
        // (float)(b * 16) + corner as trunc(p / 16) * 16 + corner: the same
        // integers, exact in float, in one FMA per axis.  (Lanes whose b
        // would overflow are outside the world and leave below without a
        // hit, so their fraction is never read.)
        f3 box = {fmaf(truncf(ps.x), 16.0f, m.corner.x), fmaf(truncf(ps.y), 16.0f, m.corner.y),
                  fmaf(truncf(ps.z), 16.0f, m.corner.z)};
        f3 tmp = (-pos) * m.inv;
        f3 bc = box * m.inv + tmp;
        float f = fminf(bc.x, fminf(bc.y, bc.z));
        f = fmaxf(f, 0.01f);
        fraction += f;
    }
    *fraction_io = fraction;
    if (cur != 0) return 1;
    // p outside [0, 16 W): a negative coordinate, or (p >= 0) a block
    // coordinate past the far edge, since trunc(p / 16) < W <=> p < 16 W.
    if (!inb || fminf(pos.x, fminf(pos.y, pos.z)) < 0.0f) return -1;
    if (fraction >= max_dist) return -1;
    if (max_steps > 0 && it >= max_steps) return -1;
    return 0;
}

// The refinement after a coarse hit at `fraction` (radiance.comp:246-288):
// up to five voxel-DDA steps from origin + dir * (fraction - 1.5).
LUMX_HD RayHit refine_hit(const WorldView& w, const MarchRay& m, float fraction) {
    RayHit r;
    r.hit = 1;
    r.coarse = 0;
    const f3 o = m.o, d = m.d, inv = m.inv;
    f3 bh = o + d * (fraction - 1.5f);
    int sx = d.x > 0.0f ? 1 : -1, sy = d.y > 0.0f ? 1 : -1, sz = d.z > 0.0f ? 1 : -1;
    f3 fl = {floorf(bh.x), floorf(bh.y), floorf(bh.z)};
    f3 c1 = {(fl.x - bh.x) / d.x, (fl.y - bh.y) / d.y, (fl.z - bh.z) / d.z};
    f3 c2 = {(fl.x - bh.x) / d.x + inv.x, (fl.y - bh.y) / d.y + inv.y, (fl.z - bh.z) / d.z + inv.z};
    f3 tMax = {fmaxf(c1.x, c2.x), fmaxf(c1.y, c2.y), fmaxf(c1.z, c2.z)};
    // 1/|d| == |1/d| exactly (round-to-nearest is sign-symmetric).
    f3 tDelta = {fabsf(inv.x), fabsf(inv.y), fabsf(inv.z)};
    int vx = (int)bh.x, vy = (int)bh.y, vz = (int)bh.z;
    f3 st = {0.0f, 0.0f, 0.0f};
    int blk;
    int cv = get_voxel(w, vx, vy, vz, &blk);
    r.refine = 1;
    int iters = 0;
    while ((iters++ <= 4) && (cv == 0)) {
        st.x = (tMax.x <= tMax.y && tMax.x <= tMax.z) ? 1.0f : 0.0f;
        st.y = (tMax.x >= tMax.y && tMax.y <= tMax.z) ? 1.0f : 0.0f;
        st.z = (tMax.z <= tMax.x && tMax.z <= tMax.y) ? 1.0f : 0.0f;
        vx += st.x != 0.0f ? sx : 0;  // sx * int(mask), without an integer multiply
        vy += st.y != 0.0f ? sy : 0;
        vz += st.z != 0.0f ? sz : 0;
        tMax = {tMax.x + tDelta.x * st.x, tMax.y + tDelta.y * st.y, tMax.z + tDelta.z * st.z};
        cv = get_voxel(w, vx, vy, vz, &blk);
        r.refine++;
    }
    r.normal = -(f3{(float)sx, (float)sy, (float)sz} * st);
    f3 tf = tMax - tDelta;
    r.fraction = dot3(tf, st);
    r.voxel = cv;
    r.key = 1 + ((vx + 512) & 1023) + (((vy + 512) & 1023) << 10) + (((vz + 512) & 1023) << 20);
    return r;
}

// CastRay_fast: coarse 0.5-voxel march with empty-block skipping, then up to
// five voxel-DDA refinement steps from origin + dir*(fraction - 1.5).
LUMX_HD RayHit cast_ray(const WorldView& w, f3 o, f3 d, float max_dist, int max_steps) {
    const MarchRay m = march_ray(o, d);
    float fraction = 0.0f;
    int it = 0;
    int coarse = 0;
    int st;
    while (true) {
        // Samples begun, counted before any lane can leave, so the count
        // stays wave-uniform (scalar) on the GPU.
        ++it;
        ++coarse;
        st = coarse_sample(w, m, max_dist, max_steps, it, &fraction);
        if (st != 0) break;
    }
    RayHit r;
    if (st < 0) {
        r.hit = 0;
        r.voxel = 0;
        r.fraction = 0.0f;
        r.normal = {0.0f, 0.0f, 0.0f};
        r.refine = 0;
        r.key = 0;
    } else {
        r = refine_hit(w, m, fraction);
    }
    r.coarse = coarse;
    return r;
}

// globalLightDir of radiance.comp:42, normalize(vec3(0.5, 0.5, -0.9)) in FP32.
LUMX_HD f3 radiance_light_dir() { return normalize3(f3{0.5f, 0.5f, -0.9f}); }

}  // namespace lumx

namespace lumx {

// Occupancy view used by the tiled kernels: per-block "id == 0" bits and
// per-slot voxel-occupancy bits (1 = voxel != 0), either of which may live
// in LDS.  A sample needs the block id (global int16) only for blocks whose
// id is non-zero, and the voxel value only at the refined hit voxel.
//   empty_bits: bit (x + wx*(y + wy*z)) of word >> 5
//   pal_bits:   slot * 128 words; voxel index i = x + 16y + 256z -> bit i
struct OccView {
    const int16_t* ids;
    const uint8_t* palette;
    const uint32_t* eb;
    const uint32_t* pb;
    int wx, wy, wz;
    int palette_slots;
    int block0_empty;

    LUMX_HD int block_at(int ix, int iy, int iz) const {
        int bx = ix / 16, by = iy / 16, bz = iz / 16;
        if ((unsigned)bx >= (unsigned)wx || (unsigned)by >= (unsigned)wy || (unsigned)bz >= (unsigned)wz) return 0;
        unsigned lin = (unsigned)(bx + wx * (by + wy * bz));
        if ((eb[lin >> 5] >> (lin & 31)) & 1u) return 0;  // eb and pb may point into LDS
        return gld(ids, lin);
    }
    LUMX_HD bool occupied(int b, int ix, int iy, int iz) const {
        if (b == 0 && block0_empty) return false;
        if ((unsigned)b >= (unsigned)palette_slots) return false;
        int idx = (ix & 15) | ((iy & 15) << 4) | ((iz & 15) << 8);
        return (pb[((uint32_t)b << 7) | (uint32_t)(idx >> 5)] >> (idx & 31)) & 1u;
    }
    LUMX_HD int voxel_value(int b, int ix, int iy, int iz) const {
        int idx = (ix & 15) | ((iy & 15) << 4) | ((iz & 15) << 8);
        return gld(palette, ((uint32_t)b << 12) | (uint32_t)idx);
    }
};

// CastRay_fast over an OccView; same arithmetic and results as cast_ray().
LUMX_HD RayHit cast_ray_occ(const OccView& w, f3 o, f3 d, float max_dist, int max_steps) {
    RayHit r;
    r.hit = 0;
    r.voxel = 0;
    r.fraction = 0.0f;
    r.normal = {0.0f, 0.0f, 0.0f};
    r.coarse = 0;
    r.refine = 0;
    r.key = 0;
    const f3 inv = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    const f3 corner = {d.x > 0.0f ? 16.0f : 0.0f, d.y > 0.0f ? 16.0f : 0.0f, d.z > 0.0f ? 16.0f : 0.0f};
    const float ext_x = (float)(w.wx * 16), ext_y = (float)(w.wy * 16), ext_z = (float)(w.wz * 16);
    float fraction = 0.0f;
    int it = 0;
    while (true) {
        // Iterations begun, counted before any lane can leave, so the count
        // stays wave-uniform (scalar) on the GPU.
        ++it;
        fraction += 0.5f;
        f3 pos = o + d * fraction;
        int ix = (int)pos.x, iy = (int)pos.y, iz = (int)pos.z;
        int blk = w.block_at(ix, iy, iz);
        bool occ = w.occupied(blk, ix, iy, iz);
        r.coarse++;
        if (blk == 0) {
            f3 box = {(float)((ix / 16) * 16) + corner.x, (float)((iy / 16) * 16) + corner.y,
                      (float)((iz / 16) * 16) + corner.z};
            f3 tmp = (-pos) * inv;
            f3 bc = box * inv + tmp;
            float f = fminf(bc.x, fminf(bc.y, bc.z));
            f = fmaxf(f, 0.01f);
            fraction += f;
        }
        if (occ) {
            r.hit = 1;
            break;
        }
        if (pos.x < 0.0f || pos.y < 0.0f || pos.z < 0.0f || pos.x >= ext_x || pos.y >= ext_y || pos.z >= ext_z)
            break;
        if (fraction >= max_dist) break;
        if (max_steps > 0 && it >= max_steps) break;
    }
    if (!r.hit) return r;

    f3 bh = o + d * (fraction - 1.5f);
    int sx = d.x > 0.0f ? 1 : -1, sy = d.y > 0.0f ? 1 : -1, sz = d.z > 0.0f ? 1 : -1;
    f3 fl = {floorf(bh.x), floorf(bh.y), floorf(bh.z)};
    f3 c1 = {(fl.x - bh.x) / d.x, (fl.y - bh.y) / d.y, (fl.z - bh.z) / d.z};
    f3 c2 = {(fl.x - bh.x) / d.x + inv.x, (fl.y - bh.y) / d.y + inv.y, (fl.z - bh.z) / d.z + inv.z};
    f3 tMax = {fmaxf(c1.x, c2.x), fmaxf(c1.y, c2.y), fmaxf(c1.z, c2.z)};
    // 1/|d| == |1/d| exactly (round-to-nearest is sign-symmetric).
    f3 tDelta = {fabsf(inv.x), fabsf(inv.y), fabsf(inv.z)};
    int vx = (int)bh.x, vy = (int)bh.y, vz = (int)bh.z;
    f3 st = {0.0f, 0.0f, 0.0f};
    int vb = w.block_at(vx, vy, vz);
    bool cv = w.occupied(vb, vx, vy, vz);
    r.refine = 1;
    int iters = 0;
    while ((iters++ <= 4) && !cv) {
        st.x = (tMax.x <= tMax.y && tMax.x <= tMax.z) ? 1.0f : 0.0f;
        st.y = (tMax.x >= tMax.y && tMax.y <= tMax.z) ? 1.0f : 0.0f;
        st.z = (tMax.z <= tMax.x && tMax.z <= tMax.y) ? 1.0f : 0.0f;
        vx += st.x != 0.0f ? sx : 0;  // sx * int(mask), without an integer multiply
        vy += st.y != 0.0f ? sy : 0;
        vz += st.z != 0.0f ? sz : 0;
        tMax = {tMax.x + tDelta.x * st.x, tMax.y + tDelta.y * st.y, tMax.z + tDelta.z * st.z};
        vb = w.block_at(vx, vy, vz);
        cv = w.occupied(vb, vx, vy, vz);
        r.refine++;
    }
    r.normal = -(f3{(float)sx, (float)sy, (float)sz} * st);
    f3 tf = tMax - tDelta;
    r.fraction = dot3(tf, st);
    r.voxel = cv ? w.voxel_value(vb, vx, vy, vz) : 0;
    r.key = 1 + ((vx + 512) & 1023) + (((vy + 512) & 1023) << 10) + (((vz + 512) & 1023) << 20);
    return r;
}

}  // namespace lumx
