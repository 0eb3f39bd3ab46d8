// lum_oracle.cpp - CPU oracle for the Lum voxel-lighting hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, on the CPU, the algorithms of
// the reference shaders and host functions listed below.  Only the test suite,
// the smoke check and the benchmark's CPU-baseline leg load it, and only as the
// checker; the product library (lum_amd/csrc) never links, calls or falls back
// to it.
//
// Restated reference code (paths relative to the reference repository):
//   shaders/radiance.comp:51-66     hash33 / Random3D                -> hash33()
//   shaders/radiance.comp:293-314   sphericalFibonacci               -> fib_dir()
//   shaders/radiance.comp:97-143    GetBlock / voxel_in_palette / GetVoxel -> World::voxel()
//   shaders/radiance.comp:145-157   GetMat                            -> World::mat()
//   shaders/radiance.comp:161-166   sample_radiance (LINEAR + REPEAT) -> sample_radiance()
//   shaders/radiance.comp:183-291   CastRay_fast                      -> cast_ray()
//   shaders/radiance.comp:316-347   trace_ray                         -> trace_ray()
//   shaders/radiance.comp:349-354   store_probe                       -> radiance_update()
//   shaders/radiance.comp:361-413   main                              -> radiance_update()
//   shaders/map.comp:33-68          main                              -> map_model()
//   shaders/glossy.frag:43-49,312-360,367-382,444-493                  -> glossy()
//   src/renderer/src/internal_render.cpp:977-1097,2182-2186 blockify  -> blockify()
//   src/renderer/src/internal_render.cpp:1142-1168 probe selection    -> select_probes()
//   src/renderer/src/internal_render.cpp:1375-1476 exec_copies        -> exec_copies()
//   src/renderer/src/internal_render.cpp:1218-1335 shift_radiance     -> shift_radiance()
//   shaders/diffuse.frag:92-97,105-137,161-245 diffuse main            -> diffuse()
//
// Arithmetic conventions (shared, by definition, with the HIP kernels):
//   * every float expression is evaluated unfused, left to right, in the order
//     the GLSL writes it (build with -ffp-contract=off);
//   * division and sqrt are IEEE correctly rounded; min/max follow fminf/fmaxf;
//   * ivec3(float) truncates toward zero, signed '/' truncates (OpSDiv) and
//     signed '%' by 16 takes the sign of the divisor (OpSMod), i.e. x & 15;
//   * out-of-bounds world fetches return block 0, out-of-range palette slots
//     and out-of-bounds model fetches return voxel 0 (robust access);
//   * the 64-per-chunk butterfly reduction order of the kernels is reproduced
//     in reduce_mean().
// Where the reference leaves a choice to the driver (FMA contraction, the
// sampler's sub-texel precision, the transcendental ulps, the subgroup sum
// order) this restatement fixes one and documents it; those choices are
// "parity unpinned" against a real driver run.
//
// Instantiated for T=float (the FP32 oracle the kernels must match) and
// T=double (the FP64 oracle used to state storage-precision error).

#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>
#include <algorithm>

namespace {

template <class T> struct V3 { T x, y, z; };
template <class T> inline V3<T> mk(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator*(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> normalize(V3<T> v) {
    T len = std::sqrt(dot(v, v));
    return V3<T>{v.x / len, v.y / len, v.z / len};
}
template <class T> inline T tmin(T a, T b) { return std::fmin(a, b); }
template <class T> inline T tmax(T a, T b) { return std::fmax(a, b); }

inline int trunc_i(double v) { return (int)v; }  // ivec3(float): toward zero

// --------------------------------------------------------------------------
// hash33 / Random3D (radiance.comp:51-66).  Integer-exact; the float
// conversion is round-to-nearest and 1/float(0xffffffff) == 2^-32.
// --------------------------------------------------------------------------
const uint32_t kHashMul = 1103515245u;

inline void hash33_u(const uint32_t in[3], uint32_t out[3]) {
    uint32_t x = in[0], y = in[1], z = in[2];
    for (int r = 0; r < 3; ++r) {
        uint32_t nx = ((x >> 8) ^ y) * kHashMul;
        uint32_t ny = ((y >> 8) ^ z) * kHashMul;
        uint32_t nz = ((z >> 8) ^ x) * kHashMul;
        x = nx; y = ny; z = nz;
    }
    out[0] = x; out[1] = y; out[2] = z;
}

// The GLSL result is a float vec3; both oracle precisions consume the float.
inline V3<float> hash33(const uint32_t s[3]) {
    uint32_t h[3];
    hash33_u(s, h);
    const float inv = 1.0f / 4294967296.0f;  // 1.0/float(0xffffffffU)
    return V3<float>{(float)h[0] * inv, (float)h[1] * inv, (float)h[2] * inv};
}

// --------------------------------------------------------------------------
// sphericalFibonacci(i, n) (radiance.comp:293-306), FP32 operands; cos/sin
// of the FP32 angle are taken in double and rounded once (the correctly
// rounded value).  Returned as float: the kernels receive the same table.
// --------------------------------------------------------------------------
V3<float> fib_dir(int i, int n) {
    const float PI = 3.1415926535897932384626433832795f;
    const float PHI = std::sqrt(5.0f) * 0.5f + 0.5f;
    float fi = (float)i;
    float a = fi * (PHI - 1.0f);
    float frac = a - std::floor(a);
    float phi = 2.0f * PI * frac;
    float cos_theta = 1.0f - (2.0f * fi + 1.0f) * (1.0f / (float)n);
    float s2 = 1.0f - cos_theta * cos_theta;
    s2 = std::min(std::max(s2, 0.0f), 1.0f);
    float sin_theta = std::sqrt(s2);
    float c = (float)std::cos((double)phi);
    float s = (float)std::sin((double)phi);
    return V3<float>{c * sin_theta, s * sin_theta, cos_theta};
}

// --------------------------------------------------------------------------
// World view: blocks (int16, x-fastest), palette (slot-linear 16^3 u8 blocks,
// voxel index x + 16y + 256z), materials (256 x 6 floats).
// --------------------------------------------------------------------------
struct World {
    int wx, wy, wz;
    const int16_t* blocks;
    const uint8_t* palette;
    int palette_slots;
    const float* mats;

    // GetBlock: texelFetch with out-of-bounds -> 0.
    inline int block(int bx, int by, int bz) const {
        if (bx < 0 || by < 0 || bz < 0 || bx >= wx || by >= wy || bz >= wz) return 0;
        return blocks[bx + wx * (by + wy * bz)];
    }
    inline int pal(int slot, int rx, int ry, int rz) const {
        if (slot < 0 || slot >= palette_slots) return 0;
        return palette[((size_t)slot << 12) | (size_t)(rx | (ry << 4) | (rz << 8))];
    }
    // GetVoxel(ivec3 ip): returns voxel, block id.
    inline int voxel_i(int ix, int iy, int iz, int* blk) const {
        int b = block(ix / 16, iy / 16, iz / 16);
        *blk = b;
        return pal(b, ix & 15, iy & 15, iz & 15);
    }
};

template <class T> struct Mat { V3<T> color; T emit, rough; };

template <class T> inline Mat<T> get_mat(const World& w, int v) {
    const float* m = w.mats + (size_t)v * 6;
    return Mat<T>{V3<T>{(T)m[0], (T)m[1], (T)m[2]}, (T)m[4], (T)m[5]};
}

// Radiance field: RGBA texels, x-fastest.  T-typed so the FP64 oracle can
// accumulate frames in double.
template <class T> struct Field {
    int wx, wy, wz;
    const T* rgba;
};

inline int wrap(int i, int n) { int r = i % n; return r < 0 ? r + n : r; }

// sample_radiance (radiance.comp:161-166): trilinear at texel coordinate
// bp = (x + n*16)/16 with REPEAT wrap (normalized LINEAR sampler, the
// texel-centre mapping (bp+0.5)/W * W - 0.5 == bp is taken exactly).
// Trilinear LINEAR filter with REPEAT wrap at texel-space coordinate bp
// (texel i's centre at bp = i), i.e. at normalized (bp + 0.5) / size.
template <class T> V3<T> sample_texel(const Field<T>& f, V3<T> bp) {
    T fx0 = std::floor(bp.x), fy0 = std::floor(bp.y), fz0 = std::floor(bp.z);
    T tx = bp.x - fx0, ty = bp.y - fy0, tz = bp.z - fz0;
    int ix = (int)fx0, iy = (int)fy0, iz = (int)fz0;
    int x0 = wrap(ix, f.wx), x1 = wrap(ix + 1, f.wx);
    int y0 = wrap(iy, f.wy), y1 = wrap(iy + 1, f.wy);
    int z0 = wrap(iz, f.wz), z1 = wrap(iz + 1, f.wz);
    auto tex = [&](int a, int b, int c) -> V3<T> {
        const T* p = f.rgba + 4 * ((size_t)a + (size_t)f.wx * ((size_t)b + (size_t)f.wy * c));
        return V3<T>{p[0], p[1], p[2]};
    };
    auto lerp = [](V3<T> a, V3<T> b, T t) -> V3<T> {
        T u = (T)1 - t;
        return V3<T>{a.x * u + b.x * t, a.y * u + b.y * t, a.z * u + b.z * t};
    };
    V3<T> c00 = lerp(tex(x0, y0, z0), tex(x1, y0, z0), tx);
    V3<T> c10 = lerp(tex(x0, y1, z0), tex(x1, y1, z0), tx);
    V3<T> c01 = lerp(tex(x0, y0, z1), tex(x1, y0, z1), tx);
    V3<T> c11 = lerp(tex(x0, y1, z1), tex(x1, y1, z1), tx);
    V3<T> c0 = lerp(c00, c10, ty);
    V3<T> c1 = lerp(c01, c11, ty);
    return lerp(c0, c1, tz);
}

template <class T> V3<T> sample_radiance(const Field<T>& f, V3<T> x, V3<T> n) {
    return sample_texel<T>(f, (x + n * (T)16) / (T)16);
}

template <class T> struct RayResult {
    bool hit;
    T fraction;          // t_local after refinement (the reference quirk)
    V3<T> normal;
    int voxel;           // material id after refinement (0 if refinement found none)
    int coarse_steps;    // coarse-loop samples taken
    int refine_steps;    // voxel-DDA refinement fetches (incl. the initial one)
    int hit_key;         // packed hit voxel position + 1 (0 = miss)
};

// CastRay_fast (radiance.comp:183-291 == glossy.frag:119-226).
template <class T>
RayResult<T> cast_ray(const World& w, V3<T> o, V3<T> d, T max_dist, int max_steps) {
    RayResult<T> r{};
    V3<T> inv = {(T)1 / d.x, (T)1 / d.y, (T)1 / d.z};
    V3<T> corner = {d.x > 0 ? (T)16 : (T)0, d.y > 0 ? (T)16 : (T)0, d.z > 0 ? (T)16 : (T)0};
    const T ext_x = (T)(w.wx * 16), ext_y = (T)(w.wy * 16), ext_z = (T)(w.wz * 16);
    T fraction = 0;
    int cur_voxel = 0;
    int it = 0;
    while (true) {
        fraction += (T)0.5;
        V3<T> pos = o + d * fraction;
        int ix = trunc_i(pos.x), iy = trunc_i(pos.y), iz = trunc_i(pos.z);
        int blk;
        cur_voxel = w.voxel_i(ix, iy, iz, &blk);
        r.coarse_steps++;
        if (blk == 0) {
            V3<T> box = {(T)((ix / 16) * 16) + corner.x, (T)((iy / 16) * 16) + corner.y,
                         (T)((iz / 16) * 16) + corner.z};
            V3<T> tmp = (-pos) * inv;
            V3<T> bc = box * inv + tmp;
            T f = tmin(bc.x, tmin(bc.y, bc.z));
            f = tmax(f, (T)0.01);
            fraction += f;
        }
        if (cur_voxel != 0) { r.hit = true; break; }
        if (pos.x < 0 || pos.y < 0 || pos.z < 0 || pos.x >= ext_x || pos.y >= ext_y || pos.z >= ext_z) {
            r.hit = false; break;
        }
        if (fraction >= max_dist) { r.hit = false; break; }
        if (max_steps > 0 && ++it >= max_steps) { r.hit = false; break; }
    }
    if (!r.hit) return r;

    V3<T> bh = o + d * (fraction - (T)1.5);
    int sx = d.x > 0 ? 1 : -1, sy = d.y > 0 ? 1 : -1, sz = d.z > 0 ? 1 : -1;
    V3<T> fl = {std::floor(bh.x), std::floor(bh.y), std::floor(bh.z)};
    V3<T> c1 = V3<T>{(fl.x - bh.x) / d.x, (fl.y - bh.y) / d.y, (fl.z - bh.z) / d.z};
    V3<T> c2 = V3<T>{(fl.x - bh.x) / d.x + (T)1 / d.x, (fl.y - bh.y) / d.y + (T)1 / d.y,
                     (fl.z - bh.z) / d.z + (T)1 / d.z};
    V3<T> tMax = {tmax(c1.x, c2.x), tmax(c1.y, c2.y), tmax(c1.z, c2.z)};
    V3<T> tDelta = {(T)1 / std::fabs(d.x), (T)1 / std::fabs(d.y), (T)1 / std::fabs(d.z)};
    int vx = trunc_i(bh.x), vy = trunc_i(bh.y), vz = trunc_i(bh.z);
    V3<T> st = {0, 0, 0};
    int blk;
    int cv = w.voxel_i(vx, vy, vz, &blk);
    r.refine_steps = 1;
    int iters = 0;
    while ((iters++ <= 4) && (cv == 0)) {
        st.x = (tMax.x <= tMax.y && tMax.x <= tMax.z) ? (T)1 : (T)0;
        st.y = (tMax.x >= tMax.y && tMax.y <= tMax.z) ? (T)1 : (T)0;
        st.z = (tMax.z <= tMax.x && tMax.z <= tMax.y) ? (T)1 : (T)0;
        vx += sx * (int)st.x; vy += sy * (int)st.y; vz += sz * (int)st.z;
        tMax = V3<T>{tMax.x + tDelta.x * st.x, tMax.y + tDelta.y * st.y, tMax.z + tDelta.z * st.z};
        cv = w.voxel_i(vx, vy, vz, &blk);
        r.refine_steps++;
    }
    r.normal = -(V3<T>{(T)sx, (T)sy, (T)sz} * st);
    V3<T> tf = tMax - tDelta;
    r.fraction = dot(tf, st);
    r.voxel = cv;
    // Hit key: refined voxel position (offset so negatives pack) + 1.
    r.hit_key = 1 + ((vx + 512) & 1023) + (((vy + 512) & 1023) << 10) + (((vz + 512) & 1023) << 20);
    return r;
}

// globalLightDir = normalize(vec3(0.5, 0.5, -0.9)) (radiance.comp:42).
template <class T> inline V3<T> light_dir() {
    return normalize(V3<T>{(T)0.5, (T)0.5, (T)-0.9});
}

// trace_ray (radiance.comp:316-347).
template <class T>
V3<T> trace_ray(const World& w, const Field<T>& f, V3<T> o, V3<T> d, T max_dist, int max_steps,
                RayResult<T>* out) {
    RayResult<T> r = cast_ray<T>(w, o, d, max_dist, max_steps);
    if (out) *out = r;
    V3<T> light = {0, 0, 0};
    if (r.hit) {
        Mat<T> m = get_mat<T>(w, r.voxel);
        V3<T> prop = sample_radiance<T>(f, o + r.fraction * d, r.normal);
        light = light + (prop * m.color + m.color * m.emit);
    } else {
        T gl = -dot(d, light_dir<T>());
        if (gl > (T)0.95) {
            V3<T> sun = V3<T>{(T)0.9, (T)0.9, (T)0.6} * (T)3.0;
            light = light + gl * sun;
        }
        V3<T> sky = V3<T>{(T)0.53, (T)0.81, (T)0.92} * (T)0.1;
        light = light + sky;
    }
    return light;
}

// Fixed-order mean over R ray lights (R a power of two <= 64, or a multiple
// of 64): chunks of C = min(R, 64) lanes reduced in butterfly order
// (s = C/2..1, v[i] += v[i+s]), chunk sums added in chunk order, then one
// IEEE division by R.
template <class T> V3<T> reduce_mean(std::vector<V3<T>>& v, int R) {
    const int C = R < 64 ? R : 64;
    V3<T> total = {0, 0, 0};
    for (int c = 0; c * C < R; ++c) {
        V3<T> buf[64];
        for (int i = 0; i < C; ++i) buf[i] = v[c * C + i];
        for (int s = C / 2; s >= 1; s >>= 1)
            for (int i = 0; i < s; ++i) buf[i] = buf[i] + buf[i + s];
        total = c == 0 ? buf[0] : total + buf[0];
    }
    return total / (T)R;
}

struct RadParams {
    int rays;        // rays per probe (RAYS_PER_PROBE)
    int time_seed;   // pco.timeSeed
    int size;        // pco.size   (request stride)
    int shift;       // pco.sihft  (request offset)
    float alpha;     // REACTIVNESS
    float max_dist;  // 16*8 in the reference
    int max_steps;   // 0 = no cap (reference)
    int n_req;       // requests in the list
};

// One probe of radiance.comp:main.  'wg' is the workgroup (dispatch) index.
// The new texel goes to out4[0..3]; returns false (nothing to store) for a
// request outside the world.
template <class T>
bool radiance_probe(const World& w, const Field<T>& fin, T* out4, const int32_t* reqs,
                    const RadParams& p, const V3<float>* fib, int wg, int* ray_keys, int* ray_steps) {
    int req = wg * p.size + p.shift;
    const int32_t* q = reqs + 4 * (size_t)req;
    int px = q[0], py = q[1], pz = q[2];
    if (px < 0 || py < 0 || pz < 0 || px >= w.wx || py >= w.wy || pz >= w.wz) return false;
    std::vector<V3<T>> lights(p.rays);
    for (int lane = 0; lane < p.rays; ++lane) {
        uint32_t gidx = (uint32_t)wg * (uint32_t)p.rays + (uint32_t)lane;
        uint32_t seed[3] = {gidx + (uint32_t)p.time_seed, (uint32_t)p.time_seed, (uint32_t)p.time_seed};
        V3<float> r0 = hash33(seed);
        seed[0]++;
        V3<float> r1 = hash33(seed);
        V3<T> corner = V3<T>{(T)px, (T)py, (T)pz} * (T)16;
        V3<T> center = corner + V3<T>{(T)8, (T)8, (T)8};
        V3<T> fd = {(T)fib[lane].x, (T)fib[lane].y, (T)fib[lane].z};
        V3<T> rr0 = {(T)r0.x, (T)r0.y, (T)r0.z};
        V3<T> rr1 = {(T)r1.x, (T)r1.y, (T)r1.z};
        V3<T> half = {(T)0.5, (T)0.5, (T)0.5};
        V3<T> dir = normalize(fd + (T)0.1 * ((rr0 - half) * (T)2.0));
        V3<T> pos = center + dir * (T)6.0;
        pos = pos + (rr1 - half) * (T)2.0;
        if (lane % 2 != 0) dir = -dir;
        RayResult<T> rr;
        lights[lane] = trace_ray<T>(w, fin, pos, dir, (T)p.max_dist, p.max_steps, &rr);
        if (ray_keys) ray_keys[(size_t)wg * p.rays + lane] = rr.hit ? rr.hit_key : 0;
        if (ray_steps) {
            ray_steps[2 * ((size_t)wg * p.rays + lane) + 0] = rr.coarse_steps;
            ray_steps[2 * ((size_t)wg * p.rays + lane) + 1] = rr.refine_steps;
        }
    }
    V3<T> mean = reduce_mean<T>(lights, p.rays);
    const T* old = fin.rgba + 4 * ((size_t)px + (size_t)fin.wx * ((size_t)py + (size_t)fin.wy * pz));
    V3<T> o = {std::fmin(std::fmax(old[0], (T)0), (T)1), std::fmin(std::fmax(old[1], (T)0), (T)1),
               std::fmin(std::fmax(old[2], (T)0), (T)1)};
    T a = (T)p.alpha;
    // A NaN component stores as 0, as the reference's UNORM image stores it.
    auto unorm_nan = [](T v) { return v != v ? (T)0 : v; };
    out4[0] = unorm_nan(o.x * ((T)1 - a) + mean.x * a);
    out4[1] = unorm_nan(o.y * ((T)1 - a) + mean.y * a);
    out4[2] = unorm_nan(o.z * ((T)1 - a) + mean.z * a);
    out4[3] = (T)1;
    return true;
}

// Every dispatched probe reads field_in (Jacobi); the new texels are then
// stored in dispatch order, so of several requests naming one probe the last
// one's value remains (the reference's write order there is a race).
template <class T>
void radiance_update(const World& w, const T* field_in, T* field_out, const int32_t* reqs,
                     const RadParams& p, int threads, int* ray_keys, int* ray_steps) {
    Field<T> fin{w.wx, w.wy, w.wz, field_in};
    size_t n_tex = (size_t)w.wx * w.wy * w.wz * 4;
    if (field_out != field_in) std::memcpy(field_out, field_in, n_tex * sizeof(T));
    std::vector<V3<float>> fib(p.rays);
    for (int i = 0; i < p.rays; ++i) fib[i] = fib_dir(i, p.rays);
    int n_wg = p.size > 0 ? p.n_req / p.size : 0;
    if (threads < 1) threads = 1;
    std::vector<T> texels((size_t)n_wg * 4);
    std::vector<uint8_t> stored((size_t)n_wg, 0);
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) {
        pool.emplace_back([&, t]() {
            for (int wg = t; wg < n_wg; wg += threads)
                stored[wg] = radiance_probe<T>(w, fin, texels.data() + 4 * (size_t)wg, reqs, p, fib.data(), wg,
                                               ray_keys, ray_steps);
        });
    }
    for (auto& th : pool) th.join();
    for (int wg = 0; wg < n_wg; ++wg) {
        if (!stored[wg]) continue;
        const int32_t* q = reqs + 4 * ((size_t)wg * p.size + p.shift);
        T* dst = field_out + 4 * ((size_t)q[0] + (size_t)w.wx * ((size_t)q[1] + (size_t)w.wy * q[2]));
        for (int c = 0; c < 4; ++c) dst[c] = texels[4 * (size_t)wg + c];
    }
}

}  // namespace

// ==========================================================================
// C interface (ctypes)
// ==========================================================================
extern "C" {

struct orc_world {
    int wx, wy, wz;
    const int16_t* blocks;
    const uint8_t* palette;
    int palette_slots;
    const float* materials;
};

struct orc_radiance_params {
    int rays, time_seed, size, shift;
    float alpha, max_dist;
    int max_steps, n_req;
};

static World to_world(const orc_world* w) {
    return World{w->wx, w->wy, w->wz, w->blocks, w->palette, w->palette_slots, w->materials};
}
static RadParams to_params(const orc_radiance_params* p) {
    return RadParams{p->rays, p->time_seed, p->size, p->shift, p->alpha, p->max_dist, p->max_steps, p->n_req};
}

void orc_hash33(const uint32_t* seed, float* out) {
    V3<float> r = hash33(seed);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

void orc_hash33_u(const uint32_t* seed, uint32_t* out) { hash33_u(seed, out); }

void orc_fibonacci(int i, int n, float* out) {
    V3<float> r = fib_dir(i, n);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

// Single ray through CastRay_fast + trace_ray shading, FP32.
// out_light[3], out_info: {hit, voxel, coarse_steps, refine_steps, hit_key}, out_frac_normal[4].
void orc_trace_ray_f32(const orc_world* w, const float* field, const float* o, const float* d,
                       float max_dist, int max_steps, float* out_light, int* out_info,
                       float* out_frac_normal) {
    World ww = to_world(w);
    Field<float> f{w->wx, w->wy, w->wz, field};
    RayResult<float> r;
    V3<float> L = trace_ray<float>(ww, f, V3<float>{o[0], o[1], o[2]}, V3<float>{d[0], d[1], d[2]},
                                   max_dist, max_steps, &r);
    out_light[0] = L.x; out_light[1] = L.y; out_light[2] = L.z;
    out_info[0] = r.hit; out_info[1] = r.voxel; out_info[2] = r.coarse_steps;
    out_info[3] = r.refine_steps; out_info[4] = r.hit ? r.hit_key : 0;
    out_frac_normal[0] = r.fraction; out_frac_normal[1] = r.normal.x;
    out_frac_normal[2] = r.normal.y; out_frac_normal[3] = r.normal.z;
}

void orc_sample_radiance_f32(int wx, int wy, int wz, const float* field, const float* x,
                             const float* n, float* out) {
    Field<float> f{wx, wy, wz, field};
    V3<float> r = sample_radiance<float>(f, V3<float>{x[0], x[1], x[2]}, V3<float>{n[0], n[1], n[2]});
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

// Radiance update, FP32.  field_in/field_out: W texels x float4.  ray_keys
// (optional) receives one hit key per dispatched ray (0 = miss), ray_steps
// (optional) two counts per ray {coarse, refine}.
void orc_radiance_update_f32(const orc_world* w, const float* field_in, float* field_out,
                             const int32_t* reqs, const orc_radiance_params* p, int threads,
                             int* ray_keys, int* ray_steps) {
    radiance_update<float>(to_world(w), field_in, field_out, reqs, to_params(p), threads, ray_keys,
                           ray_steps);
}

// Same, in double (the FP64 oracle); fields are double RGBA.
void orc_radiance_update_f64(const orc_world* w, const double* field_in, double* field_out,
                             const int32_t* reqs, const orc_radiance_params* p, int threads,
                             int* ray_keys, int* ray_steps) {
    radiance_update<double>(to_world(w), field_in, field_out, reqs, to_params(p), threads, ray_keys,
                            ray_steps);
}

// --------------------------------------------------------------------------
// map.comp:33-68 for one model.  minv: column-major float mat4 (GLSL
// layout, minv[c*4+r]); shift: border.min; groups: dispatch size in 4^3
// workgroups.  fused != 0 evaluates the mat4 x vec4 rows as
// fma(m2,z,fma(m1,y,m0*x)) + m3 (a contracted driver), else unfused
// ((m0*x + m1*y) + m2*z) + m3.  Threads whose world coordinate is negative
// alias as the shader's arithmetic makes them: ivec3() truncates toward zero,
// iw / 16 truncates (OpSDiv) and iw % 16 takes the divisor's sign (OpSMod,
// iw & 15), so -16 < iw < 0 lands in block 0.  Threads run in dispatch order
// (x fastest, then y, then z), so where several threads write one voxel the
// highest dispatch index is kept (the reference's write order there is a
// race).
// --------------------------------------------------------------------------
void orc_map_model(int wx, int wy, int wz, const int16_t* blocks, uint8_t* palette,
                   int palette_slots, int static_count, const uint8_t* model, int mx, int my,
                   int mz, const float* minv, const int* shift, const int* groups, int fused) {
    int nx = groups[0] * 4, ny = groups[1] * 4, nz = groups[2] * 4;
    for (int gz = 0; gz < nz; ++gz)
    for (int gy = 0; gy < ny; ++gy)
    for (int gx = 0; gx < nx; ++gx) {
        float wvx = ((float)gx + 0.5f) + (float)shift[0];
        float wvy = ((float)gy + 0.5f) + (float)shift[1];
        float wvz = ((float)gz + 0.5f) + (float)shift[2];
        int iwx = (int)wvx, iwy = (int)wvy, iwz = (int)wvz;
        float mv[3];
        for (int r = 0; r < 3; ++r) {
            float m0 = minv[0 * 4 + r], m1 = minv[1 * 4 + r], m2 = minv[2 * 4 + r], m3 = minv[3 * 4 + r];
            if (fused) {
                mv[r] = std::fma(m2, wvz, std::fma(m1, wvy, m0 * wvx)) + m3;
            } else {
                float a = m0 * wvx;
                float b = m1 * wvy;
                float c = m2 * wvz;
                mv[r] = ((a + b) + c) + m3;
            }
        }
        int imx = (int)mv[0], imy = (int)mv[1], imz = (int)mv[2];
        int tbx = iwx / 16, tby = iwy / 16, tbz = iwz / 16;
        int target = 0;
        if (tbx >= 0 && tby >= 0 && tbz >= 0 && tbx < wx && tby < wy && tbz < wz)
            target = blocks[tbx + wx * (tby + wy * tbz)];
        if (target >= static_count) {
            int v = 0;
            if (imx >= 0 && imy >= 0 && imz >= 0 && imx < mx && imy < my && imz < mz)
                v = model[imx + (size_t)mx * (imy + (size_t)my * imz)];
            if (v != 0 && target < palette_slots) {
                palette[((size_t)target << 12) | (size_t)((iwx & 15) | ((iwy & 15) << 4) | ((iwz & 15) << 8))] =
                    (uint8_t)v;
            }
        }
    }
}

// --------------------------------------------------------------------------
// blockify_mesh (internal_render.cpp:1029-1076) for one model, in queue
// order.  aabb_min/max: the float AABB of get_shift() for this model.
// world: the current (blockified) world, updated in place.  Returns the new
// palette counter; copy_src[slot] receives the static source block of each
// newly allocated slot (the exec_copies queue; 0 = zero-filled slot).
// --------------------------------------------------------------------------
int orc_blockify(int wx, int wy, int wz, int16_t* world, int static_count, int counter,
                 const float* aabb_min, const float* aabb_max, int32_t* copy_src, int capacity) {
    int mn[3], mx[3];
    int ws[3] = {wx, wy, wz};
    for (int a = 0; a < 3; ++a) {
        mn[a] = ((int)aabb_min[a] - 1) / 16;
        mx[a] = ((int)aabb_max[a] + 1) / 16;
        mn[a] = std::min(std::max(mn[a], 0), ws[a] - 1);
        mx[a] = std::min(std::max(mx[a], 0), ws[a] - 1);
    }
    for (int z = mn[2]; z <= mx[2]; ++z)
    for (int y = mn[1]; y <= mx[1]; ++y)
    for (int x = mn[0]; x <= mx[0]; ++x) {
        int16_t& b = world[x + wx * (y + wy * z)];
        if (b < static_count) {
            if (counter >= capacity) return -1;
            copy_src[counter] = b;
            b = (int16_t)counter;
            counter++;
        }
    }
    return counter;
}

// exec_copies (internal_render.cpp:1375-1476): palette = static blocks,
// allocated slots = copy of their static source (or zeros), rest zero.
void orc_exec_copies(uint8_t* palette, int palette_slots, int static_count, int counter,
                     const int32_t* copy_src) {
    for (int s = static_count; s < palette_slots; ++s) {
        uint8_t* dst = palette + ((size_t)s << 12);
        if (s < counter && copy_src[s] != 0)
            std::memcpy(dst, palette + ((size_t)copy_src[s] << 12), 4096);
        else
            std::memset(dst, 0, 4096);
    }
}

// update_radiance probe selection (internal_render.cpp:1142-1168): probes
// whose clamped 3x3x3 neighbourhood sums to > 0, x-fastest, then every
// special request whose probe the grid did not select, repeats included (the
// reference sets `set` for grid probes only).  The list ends at one entry per
// probe, the size of the reference's request buffer
// (internal_render.cpp:198-200).  out: int32x4 per request.  Returns the count.
int orc_select_probes(int wx, int wy, int wz, const int16_t* world, const int32_t* special,
                      int n_special, int32_t* out) {
    std::vector<uint8_t> set((size_t)wx * wy * wz, 0);
    int n = 0;
    for (int z = 0; z < wz; ++z)
    for (int y = 0; y < wy; ++y)
    for (int x = 0; x < wx; ++x) {
        int sum = 0;
        for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            int tx = std::min(std::max(x + dx, 0), wx - 1);
            int ty = std::min(std::max(y + dy, 0), wy - 1);
            int tz = std::min(std::max(z + dz, 0), wz - 1);
            sum += world[tx + wx * (ty + wy * tz)];
        }
        if (sum > 0) {
            out[4 * n + 0] = x; out[4 * n + 1] = y; out[4 * n + 2] = z; out[4 * n + 3] = 0;
            set[x + (size_t)wx * (y + (size_t)wy * z)] = 1;
            ++n;
        }
    }
    for (int i = 0; i < n_special; ++i) {
        int x = special[4 * i], y = special[4 * i + 1], z = special[4 * i + 2];
        size_t idx = x + (size_t)wx * (y + (size_t)wy * z);
        if (!set[idx] && n < wx * wy * wz) {
            out[4 * n + 0] = x; out[4 * n + 1] = y; out[4 * n + 2] = z; out[4 * n + 3] = special[4 * i + 3];
            ++n;
        }
    }
    return n;
}

// --------------------------------------------------------------------------
// glossy.frag:444-493 over a W x H G-buffer.  matnorm: u8x4 per pixel
// (x = material id, gba = encoded normal), depth: the stored depth value
// (view depth / 1000).  ubo: {campos.xyz, camdir.xyz, horiz.xyz, verti.xyz,
// light.xyz, frame_w, frame_h} as 17 floats.  Pixels whose material
// roughness > 0.5 are masked (fillStencilGlossy.frag:29-35) and not
// written.  out: float4 per pixel (pre-blend (L/8, 1 - rough)).
// out_keys (optional): hit key per pixel (0 = miss or masked).
// --------------------------------------------------------------------------
void orc_glossy_f32(const orc_world* w, const float* field, const uint8_t* matnorm,
                    const float* depth, int width, int height, const float* ubo, float max_dist,
                    int max_steps, float* out, int* out_keys, int threads) {
    World ww = to_world(w);
    Field<float> f{w->wx, w->wy, w->wz, field};
    V3<float> campos = {ubo[0], ubo[1], ubo[2]};
    V3<float> camdir = {ubo[3], ubo[4], ubo[5]};
    V3<float> horiz = {ubo[6], ubo[7], ubo[8]};
    V3<float> verti = {ubo[9], ubo[10], ubo[11]};
    V3<float> gl = {ubo[12], ubo[13], ubo[14]};
    float fw = ubo[15], fh = ubo[16];
    if (threads < 1) threads = 1;
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) {
        pool.emplace_back([&, t]() {
            for (int py = t; py < height; py += threads)
            for (int px = 0; px < width; ++px) {
                size_t pi = (size_t)py * width + px;
                const uint8_t* mn = matnorm + 4 * pi;
                Mat<float> mat = get_mat<float>(ww, mn[0]);
                if (out_keys) out_keys[pi] = 0;
                if (mat.rough > 0.5f) continue;
                float cx = ((float)px + 0.5f) / fw * 2.0f - 1.0f;
                float cy = ((float)py + 0.5f) / fh * 2.0f - 1.0f;
                float dep = depth[pi] * 1000.0f;
                V3<float> origin = campos + horiz * cx + verti * cy + camdir * dep;
                V3<float> normal = {((float)mn[1] / 255.0f) * 2.0f - 1.0f,
                                    ((float)mn[2] / 255.0f) * 2.0f - 1.0f,
                                    ((float)mn[3] / 255.0f) * 2.0f - 1.0f};
                V3<float> L = {0, 0, 0}, R = {1, 1, 1};
                V3<float> dir = camdir;
                // ProcessHit(origin, direction, 0, normal, mat, L, R)
                origin = origin + 0.0f * dir;
                V3<float> diff = sample_radiance<float>(f, origin, normal);
                R = R * mat.color;
                L = L + R * (V3<float>{mat.emit, mat.emit, mat.emit} + diff);
                float dn = dot(normal, dir);
                dir = dir - (2.0f * dn) * normal;
                // trace_glossy_ray
                RayResult<float> r = cast_ray<float>(ww, origin, dir, max_dist, max_steps);
                if (r.hit) {
                    Mat<float> hm = get_mat<float>(ww, r.voxel);
                    V3<float> o2 = origin + r.fraction * dir;
                    V3<float> d2 = sample_radiance<float>(f, o2, r.normal);
                    R = R * hm.color;
                    L = L + R * (V3<float>{hm.emit, hm.emit, hm.emit} + d2);
                    if (out_keys) out_keys[pi] = r.hit_key;
                } else {
                    float s = -dot(dir, gl);
                    if (s > 0.9f) {
                        V3<float> sun = V3<float>{0.9f, 0.9f, 0.6f} * 0.5f;
                        L = L + ((sun * R) * (s - 0.9f)) * 10.0f;
                    }
                    V3<float> sky = V3<float>{0.53f, 0.81f, 0.92f} * 0.1f;
                    L = L + sky * R;
                }
                float* o = out + 4 * pi;
                o[0] = L.x / 8.0f; o[1] = L.y / 8.0f; o[2] = L.z / 8.0f; o[3] = 1.0f - mat.rough;
            }
        });
    }
    for (auto& th : pool) th.join();
}

// --------------------------------------------------------------------------
// diffuse.frag:228-245 (main) over a W x H G-buffer, with
// sample_radiance(position) (:92-97: textureLod at (position/16)/world_size,
// LINEAR + REPEAT, i.e. texel-space coordinate position/16 - 0.5),
// get_origin_from_depth (:131-137), load_norm/load_mat/load_depth (:105-118)
// and sample_lightmap (:154-218: five NEAREST shadow taps 1/1024 apart, depth
// compare LESS against the D16 lightmap, MIRRORED_REPEAT).  lightmap may be
// null: the sun term is then 0.  The reference value of the compare is
// clamped to [0, 1] (a UNORM depth format); the stored depth is v / 65535.
// ubo: as orc_glossy_f32; lm_proj: column-major mat4 (ubo.lightmap_proj).
// out: float4 per pixel = (final / 8, 1).
// --------------------------------------------------------------------------
static int mirror_repeat(int i, int size) {
    int m = i % (2 * size);
    if (m < 0) m += 2 * size;
    int a = m - size;
    int mir = a >= 0 ? a : -(1 + a);
    return (size - 1) - mir;
}

void orc_diffuse_f32(const orc_world* w, const float* field, const uint8_t* matnorm, const float* depth,
                     int width, int height, const float* ubo, const uint16_t* lightmap, int lm_w, int lm_h,
                     const float* lm_proj, float* out) {
    World ww = to_world(w);
    Field<float> f{w->wx, w->wy, w->wz, field};
    V3<float> campos = {ubo[0], ubo[1], ubo[2]};
    V3<float> camdir = {ubo[3], ubo[4], ubo[5]};
    V3<float> horiz = {ubo[6], ubo[7], ubo[8]};
    V3<float> verti = {ubo[9], ubo[10], ubo[11]};
    V3<float> gl = {ubo[12], ubo[13], ubo[14]};
    float fw = ubo[15], fh = ubo[16];
    for (int py = 0; py < height; ++py)
        for (int px = 0; px < width; ++px) {
            size_t pi = (size_t)py * width + px;
            const uint8_t* mn = matnorm + 4 * pi;
            Mat<float> mat = get_mat<float>(ww, mn[0]);
            float cx = ((float)px + 0.5f) / fw * 2.0f - 1.0f;
            float cy = ((float)py + 0.5f) / fh * 2.0f - 1.0f;
            float dep = depth[pi] * 1000.0f;
            V3<float> origin = campos + horiz * cx + verti * cy + camdir * dep;
            V3<float> normal = {((float)mn[1] / 255.0f) * 2.0f - 1.0f, ((float)mn[2] / 255.0f) * 2.0f - 1.0f,
                                ((float)mn[3] / 255.0f) * 2.0f - 1.0f};
            V3<float> pos = origin + normal * 6.0f;
            V3<float> bp = pos / 16.0f;
            V3<float> incoming = sample_texel<float>(f, V3<float>{bp.x - 0.5f, bp.y - 0.5f, bp.z - 0.5f});
            float sun = 0.0f;
            if (lightmap) {
                V3<float> biased = origin;
                if (dot(normal, gl) > 0.0f) biased = biased - normal * 0.9f;
                else biased = biased + normal * 0.9f;
                float c[3];
                for (int r = 0; r < 3; ++r)
                    c[r] = lm_proj[0 * 4 + r] * biased.x + lm_proj[1 * 4 + r] * biased.y +
                           lm_proj[2 * 4 + r] * biased.z + lm_proj[3 * 4 + r];
                float world_depth = 1.0f + c[2];
                float u = (c[0] + 1.0f) / 2.0f, v = (c[1] + 1.0f) / 2.0f;
                float dref = std::fmin(std::fmax(world_depth, 0.0f), 1.0f);
                const int taps[5][2] = {{-1, 0}, {0, 0}, {1, 0}, {0, -1}, {0, 1}};
                float total = 0.0f;
                for (int k = 0; k < 5; ++k) {
                    float su = u + (float)taps[k][0] * (1.0f / 1024.0f);
                    float sv = v + (float)taps[k][1] * (1.0f / 1024.0f);
                    int i = (int)std::floor(su * (float)lm_w), j = (int)std::floor(sv * (float)lm_h);
                    i = mirror_repeat(i, lm_w);
                    j = mirror_repeat(j, lm_h);
                    float dtex = (float)lightmap[(size_t)j * lm_w + i] / 65535.0f;
                    total = total + (dref < dtex ? 1.0f : 0.0f);
                }
                sun = (total / 5.0f) * 0.15f;
            }
            V3<float> fin = {(2.0f * incoming.x + mat.emit + sun) * mat.color.x,
                             (2.0f * incoming.y + mat.emit + sun) * mat.color.y,
                             (2.0f * incoming.z + mat.emit + sun) * mat.color.z};
            float* o = out + 4 * pi;
            o[0] = fin.x / 8.0f; o[1] = fin.y / 8.0f; o[2] = fin.z / 8.0f; o[3] = 1.0f;
        }
}

// shift_radiance (internal_render.cpp:1218-1335): per axis, source offset =
// max(shift, 0), destination offset = max(-shift, 0), extent = size - |shift|;
// the intersection is copied to a temporary image at the origin and from there
// to the destination offset of the field.  texel_floats = floats per texel.
// The reference tests only |shift.x| >= size.x (three times); this
// restatement leaves the field unchanged when any axis has |shift| >= size.
void orc_shift_radiance(int wx, int wy, int wz, int texel_floats, float* field, int sx, int sy, int sz) {
    const int size[3] = {wx, wy, wz}, shift[3] = {sx, sy, sz};
    int src[3], dst[3], ext[3];
    for (int k = 0; k < 3; ++k) {
        if (std::abs(shift[k]) >= size[k]) return;
        src[k] = shift[k] >= 0 ? shift[k] : 0;
        dst[k] = shift[k] >= 0 ? 0 : -shift[k];
        ext[k] = size[k] - std::abs(shift[k]);
    }
    const int c = texel_floats;
    std::vector<float> temp((size_t)ext[0] * ext[1] * ext[2] * c);
    for (int z = 0; z < ext[2]; ++z)
        for (int y = 0; y < ext[1]; ++y)
            for (int x = 0; x < ext[0]; ++x) {
                const size_t t = ((size_t)x + (size_t)ext[0] * (y + (size_t)ext[1] * z)) * c;
                const size_t f = ((size_t)(x + src[0]) + (size_t)wx * ((y + src[1]) + (size_t)wy * (z + src[2]))) * c;
                for (int i = 0; i < c; ++i) temp[t + i] = field[f + i];
            }
    for (int z = 0; z < ext[2]; ++z)
        for (int y = 0; y < ext[1]; ++y)
            for (int x = 0; x < ext[0]; ++x) {
                const size_t t = ((size_t)x + (size_t)ext[0] * (y + (size_t)ext[1] * z)) * c;
                const size_t f = ((size_t)(x + dst[0]) + (size_t)wx * ((y + dst[1]) + (size_t)wy * (z + dst[2]))) * c;
                for (int i = 0; i < c; ++i) field[f + i] = temp[t + i];
            }
}

}  // extern "C"
