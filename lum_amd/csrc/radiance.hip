// radiance.hip - per-probe radiance update (gfx950).
//
// Restates shaders/radiance.comp:361-413 (main), :51-66 (hash33/Random3D),
// :316-347 (trace_ray) and :349-354 (store_probe) of the reference.
//
// Launch shape: one thread per ray (or K = 2, 4 rays per thread, traced one
// after another).  A probe's R / K threads are consecutive lanes; with
// R / K <= 64 a wave64 holds whole probes, otherwise a probe spans whole
// waves of one workgroup.  The probe mean is reduced in a fixed order
// (butterfly over min(R,64) rays, then chunk sums in chunk order, then one
// division by R) for every K, so results depend neither on timing nor on
// the launch shape.  Jacobi semantics: rays sample field_in, the probe is
// written to field_out.
#include <type_traits>

#include "march.hpp"
#include "lumx_internal.hpp"

namespace lumx {

__device__ __forceinline__ void hash33(uint32_t x, uint32_t y, uint32_t z, f3* out) {
    const uint32_t k = 1103515245u;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        uint32_t nx = ((x >> 8) ^ y) * k;
        uint32_t ny = ((y >> 8) ^ z) * k;
        uint32_t nz = ((z >> 8) ^ x) * k;
        x = nx;
        y = ny;
        z = nz;
    }
    const float inv = 2.3283064365386963e-10f;  // 1.0 / float(0xffffffffu) == 2^-32
    *out = {(float)x * inv, (float)y * inv, (float)z * inv};
}

// total / R.  For R a power of two the quotient is total scaled by 2^-k,
// which the product with the exact 1/R rounds identically; other R divide.
__device__ __forceinline__ f3 probe_mean(f3 total, int R) {
    if ((R & (R - 1)) == 0) {
        const float inv = 1.0f / (float)R;
        return total * inv;
    }
    return total / (float)R;
}

// Coordinates of probe p of the x-fastest order (wx * wy * wz < 2^28).
__device__ __forceinline__ void probe_xyz(const RadianceArgs& a, uint32_t p, int* x, int* y, int* z) {
    const uint32_t q = fast_div(p, a.div_x);    // p / wx
    const uint32_t r = fast_div(p, a.div_xy);   // p / (wx * wy)
    *x = (int)(p - mul24(q, (uint32_t)a.world.wx));
    *y = (int)(q - mul24(r, (uint32_t)a.world.wy));
    *z = (int)r;
}

// Probes of this launch: n_probes, or fewer when the request count lives on
// the device (the reference's wg_count = n_requests / size).
__device__ __forceinline__ int dispatched_probes(const RadianceArgs& a) {
    if (!a.n_requests_dev) return a.n_probes;
    return min(a.n_probes, gld(a.n_requests_dev, 0) / a.size);
}

// Probe wg's block coordinates (from the request list, the x-fastest index
// or the compacted list) and the request index that keys its rays (wg, or
// the entry's for a compacted update); false when outside the world or not
// dispatched.
__device__ __forceinline__ bool probe_coords(const RadianceArgs& a, int wg, int* px, int* py, int* pz, int* key) {
    *key = wg;
    if (a.compact) {
        if (wg >= min(a.n_probes, gld((const int*)a.tiles, (size_t)a.n_tiles))) return false;
        const int2 e = gld((const int2*)((const char*)a.field_out + (size_t)wg * (size_t)a.list_stride), 0);
        probe_xyz(a, (uint32_t)e.x, px, py, pz);
        *key = e.y;
        return true;
    }
    if (wg >= dispatched_probes(a)) return false;
    if (a.requests) {
        const int4 q = gld(a.requests, (size_t)(wg * a.size + a.shift));
        *px = q.x;
        *py = q.y;
        *pz = q.z;
    } else {
        probe_xyz(a, (uint32_t)wg, px, py, pz);
    }
    const WorldView& w = a.world;
    return (unsigned)*px < (unsigned)w.wx && (unsigned)*py < (unsigned)w.wy && (unsigned)*pz < (unsigned)w.wz;
}

// Ray r of the probe of request index `key` at block (px, py, pz)
// (radiance.comp:370-381): origin and direction.
__device__ __forceinline__ void ray_setup(const RadianceArgs& a, int key, int r, int px, int py, int pz,
                                          f3* pos_out, f3* dir_out) {
    const int R = a.rays;
    const uint32_t t = (uint32_t)a.time_seed;
    const uint32_t gidx = (uint32_t)key * (uint32_t)R + (uint32_t)r;
    f3 r0, r1;
    hash33(gidx + t, t, t, &r0);
    hash33(gidx + t + 1u, t, t, &r1);
    const float4 fb = gld(a.fib, (size_t)r);
    f3 corner = f3{(float)px, (float)py, (float)pz} * 16.0f;
    f3 center = corner + f3{8.0f, 8.0f, 8.0f};
    const f3 half = {0.5f, 0.5f, 0.5f};
    f3 dir = normalize3(f3{fb.x, fb.y, fb.z} + 0.1f * ((r0 - half) * 2.0f));
    f3 pos = center + dir * 6.0f;
    pos = pos + (r1 - half) * 2.0f;
    if (r & 1) dir = -dir;
    *pos_out = pos;
    *dir_out = dir;
}

// trace_ray's shading of a cast ray from pos along dir (radiance.comp:316-347).
template <int STORE>
__device__ __forceinline__ f3 shade_ray(const RadianceArgs& a, const RayHit& h, f3 pos, f3 dir) {
    const WorldView& w = a.world;
    f3 light = {0.0f, 0.0f, 0.0f};
    if (h.hit) {
        MatF m = get_mat(w, h.voxel);
        f3 prop = sample_radiance<STORE>(a.field_in, w.wx, w.wy, w.wz, pos + h.fraction * dir, h.normal);
        light = light + (prop * m.color + m.color * m.emit);
    } else {
        float gl = -dot3(dir, radiance_light_dir());
        if (gl > 0.95f) {
            f3 sun = f3{0.9f, 0.9f, 0.6f} * 3.0f;
            light = light + gl * sun;
        }
        f3 sky = f3{0.53f, 0.81f, 0.92f} * 0.1f;
        light = light + sky;
    }
    return light;
}

// Light gathered by ray r of the probe of request index `key` at block
// (px, py, pz).  Traced launches (never compacted: key is the probe's index
// in the launch) record the ray at key * R + r.
template <int STORE, bool TRACED, bool OCC>
__device__ __forceinline__ f3 probe_ray_light(const RadianceArgs& a, int key, int r, int px, int py, int pz) {
    const WorldView& w = a.world;
    const int R = a.rays;
    f3 pos, dir;
    ray_setup(a, key, r, px, py, pz, &pos, &dir);

    RayHit h;
    if (OCC) {
        OccView occ;
        occ.ids = w.blocks;
        occ.palette = w.palette;
        occ.eb = a.empty_bits;
        occ.pb = a.pal_bits;
        occ.wx = w.wx;
        occ.wy = w.wy;
        occ.wz = w.wz;
        occ.palette_slots = w.palette_slots;
        occ.block0_empty = w.block0_empty;
        h = cast_ray_occ(occ, pos, dir, a.max_dist, a.max_steps);
    } else {
        h = cast_ray(w, pos, dir, a.max_dist, a.max_steps);
    }
    const f3 light = shade_ray<STORE>(a, h, pos, dir);
    if (TRACED) {
        size_t ri = (size_t)key * R + r;
        if (a.ray_keys) gst(a.ray_keys, ri, h.hit ? h.key : 0);
        if (a.ray_steps) {
            gst(a.ray_steps, 2 * ri + 0, h.coarse);
            gst(a.ray_steps, 2 * ri + 1, h.refine);
        }
    }
    return light;
}

// One butterfly stage, v[i] + v[i ^ S], with every lane of the wave active.
// S = 32: a swap of the two 32-lane halves hands each lane both its own and
// its partner's value (in some order), and their sum is that stage's sum;
// S = 16..4: swizzle in XOR mode; S = 2, 1: DPP quad permutes, which the
// compiler folds into the add.
template <int S>
__device__ __forceinline__ float bfly_stage(float v) {
    if constexpr (S == 32) {
        const auto p = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
        return __uint_as_float(p[0]) + __uint_as_float(p[1]);
    } else if constexpr (S >= 4) {
        return v + __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x1F | (S << 10)));
    } else {
        constexpr int quad_perm = S == 2 ? 0x4E : 0xB1;  // [2,3,0,1] : [1,0,3,2]
        return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), quad_perm, 0xF, 0xF, true));
    }
}

template <int S>
__device__ __forceinline__ f3 bfly_stage3(f3 v) {
    return {bfly_stage<S>(v.x), bfly_stage<S>(v.y), bfly_stage<S>(v.z)};
}

// v[i] += v[i ^ s] for s = width/2 .. 1 (width a power of two <= 64; the
// whole wave executes it).
__device__ __forceinline__ f3 butterfly(f3 v, int width) {
    if (width > 32) v = bfly_stage3<32>(v);
    if (width > 16) v = bfly_stage3<16>(v);
    if (width > 8) v = bfly_stage3<8>(v);
    if (width > 4) v = bfly_stage3<4>(v);
    if (width > 2) v = bfly_stage3<2>(v);
    if (width > 1) v = bfly_stage3<1>(v);
    return v;
}

// Blend of the probe mean into the old texel (radiance.comp:349-354, 405-412).
template <int STORE>
__device__ __forceinline__ void store_probe(const RadianceArgs& a, int wg, int px, int py, int pz, f3 total) {
    const WorldView& w = a.world;
    f3 mean = probe_mean(total, a.rays);
    const unsigned ti = (unsigned)px + (unsigned)w.wx * ((unsigned)py + (unsigned)w.wy * (unsigned)pz);
    f3 old = Texel<STORE>::load(a.field_in, ti);
    old = {fminf(fmaxf(old.x, 0.0f), 1.0f), fminf(fmaxf(old.y, 0.0f), 1.0f), fminf(fmaxf(old.z, 0.0f), 1.0f)};
    const float al = a.alpha, bl = 1.0f - a.alpha;
    f3 nv = {old.x * bl + mean.x * al, old.y * bl + mean.y * al, old.z * bl + mean.z * al};
    Texel<STORE>::store(a.field_out, a.staged ? (unsigned)wg : ti, unorm_nan_to_zero(nv));
}

// K rays per lane.  tid is the thread's index within this world's launch
// range, which starts on a workgroup boundary; a probe owns L = R / K
// consecutive threads.
//
// The probe total is reduced in one fixed order for every K: per chunk of
// C = min(R, 64) rays a butterfly (v[i] += v[i ^ s], s = C/2 .. 1), then the
// chunk sums in chunk order.  With K > 1 (R a power of two, R >= K) a lane
// traces several rays one after another and adds them as the butterfly's
// first stages would, so the total is bitwise the K = 1 total:
//   Kc = min(K, R / C) chunks per lane, Kw = K / Kc rays per lane and chunk,
//   Lc = C / Kw lanes per chunk.  Lane m = g * Lc + p traces, for each of its
//   chunks c = g * Kc + jc, the rays c * C + p + o * Lc, o = 0, 1 (Kw = 2)
//   or o = 0, 2, 1, 3 (Kw = 4): stage C/2 pairs rays p and p + C/2 (o = 0, 2
//   for Kw = 4), stage C/4 pairs their sums.
template <int STORE, bool TRACED, bool OCC, int K>
__device__ __forceinline__ void radiance_lane_body(const RadianceArgs& a, const int tid, f3* chunk_sum) {
    const int R = a.rays;
    constexpr int LOG2K = K == 1 ? 0 : (K == 2 ? 1 : 2);
    const int L = R >> LOG2K;  // K > 1: R is a power of two
    int wg = K == 1 ? (a.log2_rays >= 0 ? (tid >> a.log2_rays) : tid / R) : (tid >> (a.log2_rays - LOG2K));
    // With L >= 64 a wave holds lanes of one probe: keep the probe on the scalar unit.
    if (L >= 64) wg = __builtin_amdgcn_readfirstlane(wg);
    const int m = tid - wg * L;
    int px = 0, py = 0, pz = 0, key = wg;
    const bool valid = probe_coords(a, wg, &px, &py, &pz, &key);
    const f3 zero = {0.0f, 0.0f, 0.0f};
    const int C = R < 64 ? R : 64;
    f3 total;
    if (K == 1) {
        f3 light = valid ? probe_ray_light<STORE, TRACED, OCC>(a, key, m, px, py, pz) : zero;
        total = butterfly(light, C);
        if (R > 64) {
            const int wave = threadIdx.x >> 6;
            if ((threadIdx.x & 63) == 0) chunk_sum[wave] = total;
            __syncthreads();
            if (m == 0) {
                const int n_chunks = R >> 6;
                total = chunk_sum[wave];
                for (int c = 1; c < n_chunks; ++c) total = total + chunk_sum[wave + c];
            }
        }
    } else {
        const int n_chunks = R / C;
        const int Kc = K < n_chunks ? K : n_chunks;
        const int Kw = K / Kc;
        const int Lc = C / Kw;
        const int g = m / Lc;
        const int p = m - g * Lc;
        f3 acc[K];
#pragma unroll
        for (int q = 0; q < K; ++q) acc[q] = zero;
        f3 t0 = zero, t1 = zero, cur = zero;
#pragma unroll 1
        for (int j = 0; j < K; ++j) {
            const int jc = j / Kw, jj = j - jc * Kw;
            const int o = Kw == 4 ? (((jj & 1) << 1) | (jj >> 1)) : jj;
            const int r = (g * Kc + jc) * C + p + o * Lc;
            const f3 v = valid ? probe_ray_light<STORE, TRACED, OCC>(a, key, r, px, py, pz) : zero;
            if (Kw == 1) {
                cur = v;
            } else if (Kw == 2) {
                if (jj == 0) t0 = v;
                else cur = t0 + v;
            } else {
                if (jj == 0) t0 = v;
                else if (jj == 1) t0 = t0 + v;
                else if (jj == 2) t1 = v;
                else cur = t0 + (t1 + v);
            }
            if (jj == Kw - 1) {
#pragma unroll
                for (int q = 0; q < K; ++q)
                    if (q == jc) acc[q] = cur;
            }
        }
#pragma unroll
        for (int q = 0; q < K; ++q)
            if (q < Kc) acc[q] = butterfly(acc[q], Lc);
        if (n_chunks == Kc) {
            total = acc[0];
#pragma unroll
            for (int q = 1; q < K; ++q)
                if (q < Kc) total = total + acc[q];
        } else {
            // Kc = K, Lc = 64: wave g of the probe holds chunks g*K .. g*K+K-1.
            const int wave = threadIdx.x >> 6;
            if ((threadIdx.x & 63) == 0) {
#pragma unroll
                for (int q = 0; q < K; ++q) chunk_sum[wave * K + q] = acc[q];
            }
            __syncthreads();
            total = acc[0];
            if (m == 0) {
                for (int c = 1; c < n_chunks; ++c) total = total + chunk_sum[wave * K + c];
            }
        }
    }
    if (valid && m == 0) store_probe<STORE>(a, wg, px, py, pz, total);
}

template <int STORE, bool TRACED, bool OCC, int K>
__global__ void __launch_bounds__(1024) radiance_kernel(RadianceArgs a) {
    extern __shared__ f3 chunk_sum[];
    radiance_lane_body<STORE, TRACED, OCC, K>(a, blockIdx.x * blockDim.x + threadIdx.x, chunk_sum);
}

// Lane refill (tuning variant 3; 64 rays per probe).  A wave owns probes
// p0 .. p0 + P - 1 as one pool of 64 P rays.  Every lane sets up its home
// rays (ray `lane` of each probe) into LDS; then lanes march pool rays one
// coarse sample per iteration, and a lane whose ray ends takes the next
// unmarched ray of the pool (ballot of the finished lanes, rank by mbcnt),
// so no lane idles while rays remain.  The coarse result of every ray (its
// fraction at a hit, -1 for a miss) goes to LDS; each home lane then refines
// and shades its rays and the probes are reduced and stored exactly as in
// the one-ray-per-lane kernel, so results are bitwise the same.
constexpr int kRefillFloats = 10;  // per pool ray: origin, direction, 1/direction, coarse fraction

size_t radiance_refill_lds_bytes(int P) { return (size_t)4 * P * 64 * kRefillFloats * sizeof(float); }

template <int STORE>
__global__ void __launch_bounds__(256) radiance_refill_kernel(RadianceArgs a, int P) {
    extern __shared__ float refill_pool[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int NR = P * 64;
    float* rs = refill_pool + (size_t)wave * NR * kRefillFloats;
    const int wg0 = (blockIdx.x * 4 + wave) * P;
    const WorldView& w = a.world;
    for (int j = 0; j < P; ++j) {
        int px = 0, py = 0, pz = 0, key = 0;
        f3 pos = {-64.0f, -64.0f, -64.0f}, dir = {0.57735027f, 0.57735027f, 0.57735027f};
        // A probe not dispatched or outside the world marches from outside it
        // and leaves at its first sample; its light is not used.
        if (probe_coords(a, wg0 + j, &px, &py, &pz, &key)) ray_setup(a, key, lane, px, py, pz, &pos, &dir);
        const MarchRay m = march_ray(pos, dir);
        float* r = rs + (j * 64 + lane) * kRefillFloats;
        r[0] = m.o.x; r[1] = m.o.y; r[2] = m.o.z;
        r[3] = m.d.x; r[4] = m.d.y; r[5] = m.d.z;
        r[6] = m.inv.x; r[7] = m.inv.y; r[8] = m.inv.z;
    }
    __syncthreads();
    auto load_ray = [&](int q) {
        const float* r = rs + q * kRefillFloats;
        MarchRay m;
        m.o = {r[0], r[1], r[2]};
        m.d = {r[3], r[4], r[5]};
        m.inv = {r[6], r[7], r[8]};
        m.corner = {m.d.x > 0.0f ? 16.0f : 0.0f, m.d.y > 0.0f ? 16.0f : 0.0f, m.d.z > 0.0f ? 16.0f : 0.0f};
        return m;
    };
    int ray = lane;
    MarchRay m = load_ray(ray);
    float fraction = 0.0f;
    int it = 0;
    bool active = true;
    int next = 64;  // wave-uniform: the next pool ray to hand out
    while (true) {
        int st = 0;
        if (active) {
            ++it;
            st = coarse_sample(w, m, a.max_dist, a.max_steps, it, &fraction);
        }
        const bool done = active && st != 0;
        if (done) rs[ray * kRefillFloats + 9] = st > 0 ? fraction : -1.0f;
        const unsigned long long fin = __ballot(done);
        if (fin) {
            const int rank = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(fin >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((unsigned)fin, 0u));
            if (done) {
                const int q = next + rank;
                if (q < NR) {
                    ray = q;
                    m = load_ray(q);
                    fraction = 0.0f;
                    it = 0;
                } else {
                    active = false;
                }
            }
            next += __popcll(fin);
        }
        if (__ballot(active) == 0ull) break;
    }
    __syncthreads();
    const f3 zero = {0.0f, 0.0f, 0.0f};
    for (int j = 0; j < P; ++j) {
        int px = 0, py = 0, pz = 0, key = 0;
        const bool valid = probe_coords(a, wg0 + j, &px, &py, &pz, &key);
        const int q = j * 64 + lane;
        const MarchRay hm = load_ray(q);
        const float fr = rs[q * kRefillFloats + 9];
        RayHit h;
        if (fr >= 0.0f) {
            h = refine_hit(w, hm, fr);
        } else {
            h.hit = 0;
            h.voxel = 0;
            h.fraction = 0.0f;
            h.normal = zero;
        }
        const f3 light = valid ? shade_ray<STORE>(a, h, hm.o, hm.d) : zero;
        const f3 total = butterfly(light, 64);
        if (valid && lane == 0) store_probe<STORE>(a, wg0 + j, px, py, pz, total);
    }
}

hipError_t launch_radiance_refill(const RadianceArgs& a, int store, int P, hipStream_t stream) {
    if (a.rays != 64 || P < 1 || P > 4) return hipErrorInvalidValue;
    if (a.n_probes <= 0) return hipSuccess;
    const int grid = (a.n_probes + 4 * P - 1) / (4 * P);
    const size_t lds = radiance_refill_lds_bytes(P);
    switch (store) {
        case STORE_F16:
            hipLaunchKernelGGL(radiance_refill_kernel<STORE_F16>, dim3(grid), dim3(256), lds, stream, a, P);
            break;
        case STORE_BF16:
            hipLaunchKernelGGL(radiance_refill_kernel<STORE_BF16>, dim3(grid), dim3(256), lds, stream, a, P);
            break;
        default:
            hipLaunchKernelGGL(radiance_refill_kernel<STORE_F32>, dim3(grid), dim3(256), lds, stream, a, P);
            break;
    }
    return hipGetLastError();
}

// Many worlds in one launch.  World w owns threads [begin[w], begin[w+1]),
// each range starting on a workgroup boundary; a workgroup finds its world by
// a scalar binary search over begin[].  The per-probe arithmetic, RNG keying
// (request index within the world's own update) and reduction order are those
// of a single-world launch, so every world's result is the same alone and in
// the batch.  time_seed of every world is offset by *frame (0 when null).
// SERVED only separates the serving engine's launches from the batch API's
// in profiles; the code is the same.
template <int STORE, bool OCC, bool SERVED, int K>
__global__ void __launch_bounds__(1024)
radiance_batch_kernel(const RadianceArgs* args, const int* begin, int n_worlds, const int* frame) {
    extern __shared__ f3 chunk_sum[];
    const int b0 = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * blockDim.x));
    int lo = 0, hi = n_worlds - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (begin[mid] <= b0) lo = mid;
        else hi = mid - 1;
    }
    lo = __builtin_amdgcn_readfirstlane(lo);
    RadianceArgs a = args[lo];
    if (frame) a.time_seed += *frame;
    radiance_lane_body<STORE, false, OCC, K>(a, (int)(blockIdx.x * blockDim.x + threadIdx.x) - begin[lo], chunk_sum);
}

template <int STORE, bool SERVED>
static void launch_batch_k(int K, dim3 grid, dim3 block, size_t lds, hipStream_t stream, const RadianceArgs* args,
                           const int* begin, int n, const int* frame) {
    if (K == 4)
        hipLaunchKernelGGL((radiance_batch_kernel<STORE, false, SERVED, 4>), grid, block, lds, stream, args, begin,
                           n, frame);
    else if (K == 2)
        hipLaunchKernelGGL((radiance_batch_kernel<STORE, false, SERVED, 2>), grid, block, lds, stream, args, begin,
                           n, frame);
    else
        hipLaunchKernelGGL((radiance_batch_kernel<STORE, false, SERVED, 1>), grid, block, lds, stream, args, begin,
                           n, frame);
}

hipError_t launch_radiance_batch(const RadianceArgs* args_dev, const int* begin_dev, int n_worlds,
                                 int total_threads, int block, int rays_per_lane, int store, bool served,
                                 const int* frame_dev, hipStream_t stream) {
    if (total_threads <= 0 || n_worlds <= 0) return hipSuccess;
    const int grid = (total_threads + block - 1) / block;
    const size_t lds = sizeof(f3) * (block / 64) * rays_per_lane;
#define LUMX_BATCH_LAUNCH(S, V) \
    launch_batch_k<S, V>(rays_per_lane, dim3(grid), dim3(block), lds, stream, args_dev, begin_dev, n_worlds, frame_dev)
    switch (store) {
        case STORE_F16: if (served) LUMX_BATCH_LAUNCH(STORE_F16, true); else LUMX_BATCH_LAUNCH(STORE_F16, false); break;
        case STORE_BF16: if (served) LUMX_BATCH_LAUNCH(STORE_BF16, true); else LUMX_BATCH_LAUNCH(STORE_BF16, false); break;
        default: if (served) LUMX_BATCH_LAUNCH(STORE_F32, true); else LUMX_BATCH_LAUNCH(STORE_F32, false); break;
    }
#undef LUMX_BATCH_LAUNCH
    return hipGetLastError();
}

// K rays per lane (K > 1 needs R a power of two, R >= K): a probe owns
// L = R / K threads.
template <int STORE, bool TRACED, bool OCC, int K>
static hipError_t launch_store(const RadianceArgs& a, hipStream_t stream) {
    const int R = a.rays, L = R / K;
    int bs = L >= 256 ? L : L * (256 / L);
    long long threads = (long long)a.n_probes * L;
    int grid = (int)((threads + bs - 1) / bs);
    size_t lds = sizeof(f3) * (bs / 64) * K;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL((radiance_kernel<STORE, TRACED, OCC, K>), dim3(grid), dim3(bs), lds, stream, a);
    return hipGetLastError();
}

template <bool OCC>
static hipError_t launch_lane(const RadianceArgs& a, int store, bool traced, int K, hipStream_t stream) {
    if (traced) {
        switch (store) {
            case STORE_F16: return launch_store<STORE_F16, true, OCC, 1>(a, stream);
            case STORE_BF16: return launch_store<STORE_BF16, true, OCC, 1>(a, stream);
            default: return launch_store<STORE_F32, true, OCC, 1>(a, stream);
        }
    }
    // Several rays per lane: the plain (u8 palette) march only.
    if (!OCC && K == 2) {
        switch (store) {
            case STORE_F16: return launch_store<STORE_F16, false, false, 2>(a, stream);
            case STORE_BF16: return launch_store<STORE_BF16, false, false, 2>(a, stream);
            default: return launch_store<STORE_F32, false, false, 2>(a, stream);
        }
    }
    if (!OCC && K == 4) {
        switch (store) {
            case STORE_F16: return launch_store<STORE_F16, false, false, 4>(a, stream);
            case STORE_BF16: return launch_store<STORE_BF16, false, false, 4>(a, stream);
            default: return launch_store<STORE_F32, false, false, 4>(a, stream);
        }
    }
    switch (store) {
        case STORE_F16: return launch_store<STORE_F16, false, OCC, 1>(a, stream);
        case STORE_BF16: return launch_store<STORE_BF16, false, OCC, 1>(a, stream);
        default: return launch_store<STORE_F32, false, OCC, 1>(a, stream);
    }
}

// Texel of probe i of a staged subset update (request i * size + shift);
// false when i is not dispatched or the probe lies outside the world.
__device__ __forceinline__ bool staged_texel(const RadianceArgs& a, int i, unsigned* ti) {
    if (i >= dispatched_probes(a)) return false;
    const WorldView& w = a.world;
    int px, py, pz;
    if (a.requests) {
        const int4 q = gld(a.requests, (size_t)(i * a.size + a.shift));
        px = q.x;
        py = q.y;
        pz = q.z;
    } else {
        const int p = i * a.size + a.shift;
        probe_xyz(a, (uint32_t)p, &px, &py, &pz);
    }
    if ((unsigned)px >= (unsigned)w.wx || (unsigned)py >= (unsigned)w.wy || (unsigned)pz >= (unsigned)w.wz) return false;
    *ti = (unsigned)px + (unsigned)w.wx * ((unsigned)py + (unsigned)w.wy * (unsigned)pz);
    return true;
}

// Claims of a staged subset update, before it runs: claim[texel] = the
// highest i among the probes naming it (duplicate requests of one probe:
// the last one is applied).  A pass of its own keeps the atomics out of the
// march kernels, whose waves would otherwise hold their slots until the
// atomic returns.
__device__ __forceinline__ void claim_probe(const RadianceArgs& a, int i) {
    unsigned ti;
    if (staged_texel(a, i, &ti)) atomicMax(a.claim + ti, i);
}

// Scatter of a staged subset update: probe i of the launch (request
// i * size + shift) gets field_out[i] in field_in, in place.  Probes outside
// the world were not updated and are skipped, as in the update itself.
// Where several requests name one probe, the last of them (the highest i,
// recorded in claim[] before the update) is applied and the claim word is
// reset for the next update; the others leave the texel alone.
template <int BYTES>
__device__ __forceinline__ void scatter_probe(const RadianceArgs& a, int i) {
    typedef typename std::conditional<BYTES == 16, uint4, uint2>::type T;
    unsigned ti;
    if (!staged_texel(a, i, &ti)) return;
    if (a.claim) {
        if (a.claim[ti] != i) return;
        a.claim[ti] = -1;
    }
    gst((T*)a.field_in, ti, gld((const T*)a.field_out, (size_t)i));
}

template <int BYTES>
__global__ void __launch_bounds__(256) scatter_kernel(RadianceArgs a) {
    scatter_probe<BYTES>(a, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}

// Many worlds: world w owns probes [pbegin[w], pbegin[w + 1]).
template <int BYTES>
__global__ void __launch_bounds__(256) scatter_batch_kernel(const RadianceArgs* args, const int* pbegin, int n) {
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pbegin[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    const RadianceArgs& a = args[lo];
    if (!a.staged || a.compact) return;
    scatter_probe<BYTES>(a, t - pbegin[lo]);
}

__global__ void __launch_bounds__(256) claim_kernel(RadianceArgs a) {
    claim_probe(a, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}

// Many worlds: world w owns probes [pbegin[w], pbegin[w + 1]).
__global__ void __launch_bounds__(256) claim_batch_kernel(const RadianceArgs* args, const int* pbegin, int n) {
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pbegin[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    const RadianceArgs& a = args[lo];
    if (!a.staged || !a.claim) return;
    claim_probe(a, t - pbegin[lo]);
}

hipError_t launch_claim(const RadianceArgs& a, hipStream_t stream) {
    if (a.n_probes <= 0 || !a.claim) return hipSuccess;
    hipLaunchKernelGGL(claim_kernel, dim3((a.n_probes + 255) / 256), dim3(256), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_claim_batch(const RadianceArgs* args_dev, const int* pbegin_dev, int n_worlds, int total_probes,
                              hipStream_t stream) {
    if (n_worlds <= 0 || total_probes <= 0) return hipSuccess;
    hipLaunchKernelGGL(claim_batch_kernel, dim3((total_probes + 255) / 256), dim3(256), 0, stream, args_dev,
                       pbegin_dev, n_worlds);
    return hipGetLastError();
}

// ---- compacted subset updates ----
// The claimed texels of a world (claim >= 0), in x-fastest order, are the
// probes of the update: tile t holds texels 2048 t .. 2048 t + 2047, 8 rounds
// of 256 threads; the tile's claimed texels are ranked by round, wave and
// lane, which is texel order.
constexpr int kCompactTile = 2048;
constexpr int kCompactThreads = 256;

bool compact_pays(long long n_probes, long long n_blocks) {
    // Three passes over the world's claims (about 12 B per texel) against
    // the update's own work: worth it from a sixteenth of the world and a
    // few thousand probes up.
    return n_probes >= 2048 && n_probes * 16 >= n_blocks;
}

__device__ __forceinline__ int world_texels(const RadianceArgs& a) {
    return a.world.wx * a.world.wy * a.world.wz;
}

// Ballot words of tile t's claimed texels (8 rounds x 4 waves, word
// 4 * round + wave) in LDS; returns the tile's count to thread 0.
__device__ __forceinline__ int compact_ballots(const RadianceArgs& a, int t, unsigned long long* s_words) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = world_texels(a);
    for (int k = 0; k < kCompactTile / kCompactThreads; ++k) {
        const int p = t * kCompactTile + k * kCompactThreads + (int)threadIdx.x;
        const bool flag = p < n && gld((const int*)a.claim, (size_t)p) >= 0;
        const unsigned long long m = __ballot(flag);
        if (lane == 0) s_words[k * (kCompactThreads / 64) + wave] = m;
    }
    __syncthreads();
    int total = 0;
    if (threadIdx.x == 0)
        for (int c = 0; c < kCompactTile / 64; ++c) total += __popcll(s_words[c]);
    return total;
}

// Rank of thread's texel of round k within tile (claimed texels before it).
__device__ __forceinline__ int compact_rank(const unsigned long long* s_words, const int* s_pref, int k) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c = k * (kCompactThreads / 64) + wave;
    return s_pref[c] + __popcll(s_words[c] & ((1ull << lane) - 1ull));
}

__device__ __forceinline__ void compact_prefix(const unsigned long long* s_words, int* s_pref) {
    if (threadIdx.x < kCompactTile / 64) {
        int s = 0;
        for (int c = 0; c < (int)threadIdx.x; ++c) s += __popcll(s_words[c]);
        s_pref[threadIdx.x] = s;
    }
    __syncthreads();
}

__device__ __forceinline__ void compact_count_tile(const RadianceArgs& a, int t) {
    __shared__ unsigned long long s_words[kCompactTile / 64];
    const int total = compact_ballots(a, t, s_words);
    if (threadIdx.x == 0) a.tiles[t] = total;
}

// Exclusive scan of one world's tile counts (one workgroup of 256, small
// enough to start between the waves of a busy GPU): the tile offsets, and
// the total after them.
__device__ __forceinline__ void compact_scan_world(const RadianceArgs& a) {
    __shared__ int s_wave[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, n_waves = blockDim.x >> 6;
    int running = 0;
    for (int base = 0; base < a.n_tiles; base += blockDim.x) {
        const int i = base + (int)threadIdx.x;
        const int v = i < a.n_tiles ? a.tiles[i] : 0;
        int x = v;
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(x, d);
            if (lane >= d) x += y;
        }
        if (lane == 63) s_wave[wave] = x;
        __syncthreads();
        if (wave == 0) {
            int w = lane < n_waves ? s_wave[lane] : 0;
            for (int d = 1; d < 16; d <<= 1) {
                const int y = __shfl_up(w, d);
                if (lane >= d) w += y;
            }
            if (lane < n_waves) s_wave[lane] = w;
        }
        __syncthreads();
        const int before = wave > 0 ? s_wave[wave - 1] : 0;
        if (i < a.n_tiles) a.tiles[i] = running + before + x - v;
        running += s_wave[n_waves - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.tiles[a.n_tiles] = running;
}

// Entries {texel, request index} of tile t's claimed texels.
__device__ __forceinline__ void compact_list_tile(const RadianceArgs& a, int t) {
    __shared__ unsigned long long s_words[kCompactTile / 64];
    __shared__ int s_pref[kCompactTile / 64];
    compact_ballots(a, t, s_words);
    compact_prefix(s_words, s_pref);
    const int offset = a.tiles[t];
    const int n = world_texels(a);
    for (int k = 0; k < kCompactTile / kCompactThreads; ++k) {
        const int p = t * kCompactTile + k * kCompactThreads + (int)threadIdx.x;
        if (p >= n) break;
        const int c = gld((const int*)a.claim, (size_t)p);
        if (c < 0) continue;
        const int j = offset + compact_rank(s_words, s_pref, k);
        gst((int2*)((char*)a.field_out + (size_t)j * (size_t)a.list_stride), 0, make_int2(p, c));
    }
}

// After the update: field_in[texel of entry j] = field_out[j], claims reset.
template <int BYTES>
__device__ __forceinline__ void compact_scatter_tile(const RadianceArgs& a, int t) {
    typedef typename std::conditional<BYTES == 16, uint4, uint2>::type T;
    __shared__ unsigned long long s_words[kCompactTile / 64];
    __shared__ int s_pref[kCompactTile / 64];
    compact_ballots(a, t, s_words);
    compact_prefix(s_words, s_pref);
    const int offset = a.tiles[t];
    const int n = world_texels(a);
    for (int k = 0; k < kCompactTile / kCompactThreads; ++k) {
        const int p = t * kCompactTile + k * kCompactThreads + (int)threadIdx.x;
        if (p >= n) break;
        if (gld((const int*)a.claim, (size_t)p) < 0) continue;
        const int j = offset + compact_rank(s_words, s_pref, k);
        gst((T*)a.field_in, (size_t)p, gld((const T*)a.field_out, (size_t)j));
        gst(a.claim, (size_t)p, -1);
    }
}

__global__ void __launch_bounds__(kCompactThreads) compact_count_kernel(RadianceArgs a) {
    compact_count_tile(a, (int)blockIdx.x);
}
__global__ void __launch_bounds__(kCompactThreads) compact_scan_kernel(RadianceArgs a) { compact_scan_world(a); }
__global__ void __launch_bounds__(kCompactThreads) compact_list_kernel(RadianceArgs a) {
    compact_list_tile(a, (int)blockIdx.x);
}
template <int BYTES>
__global__ void __launch_bounds__(kCompactThreads) compact_scatter_kernel(RadianceArgs a) {
    compact_scatter_tile<BYTES>(a, (int)blockIdx.x);
}

// Many worlds: world w owns tiles tbegin[w] .. tbegin[w + 1] - 1 (none when
// not compacted); a workgroup finds its world by a scalar search.
__device__ __forceinline__ int tile_world(const int* tbegin, int n, int b) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tbegin[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

__global__ void __launch_bounds__(kCompactThreads) compact_count_batch_kernel(const RadianceArgs* args,
                                                                              const int* tbegin, int n) {
    const int w = tile_world(tbegin, n, (int)blockIdx.x);
    compact_count_tile(args[w], (int)blockIdx.x - tbegin[w]);
}
__global__ void __launch_bounds__(kCompactThreads) compact_scan_batch_kernel(const RadianceArgs* args) {
    const RadianceArgs& a = args[blockIdx.x];
    if (a.compact) compact_scan_world(a);
}
__global__ void __launch_bounds__(kCompactThreads) compact_list_batch_kernel(const RadianceArgs* args,
                                                                             const int* tbegin, int n) {
    const int w = tile_world(tbegin, n, (int)blockIdx.x);
    compact_list_tile(args[w], (int)blockIdx.x - tbegin[w]);
}
template <int BYTES>
__global__ void __launch_bounds__(kCompactThreads) compact_scatter_batch_kernel(const RadianceArgs* args,
                                                                                const int* tbegin, int n) {
    const int w = tile_world(tbegin, n, (int)blockIdx.x);
    compact_scatter_tile<BYTES>(args[w], (int)blockIdx.x - tbegin[w]);
}

hipError_t launch_compact(const RadianceArgs& a, hipStream_t stream) {
    if (!a.compact || a.n_tiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(compact_count_kernel, dim3(a.n_tiles), dim3(kCompactThreads), 0, stream, a);
    hipLaunchKernelGGL(compact_scan_kernel, dim3(1), dim3(kCompactThreads), 0, stream, a);
    hipLaunchKernelGGL(compact_list_kernel, dim3(a.n_tiles), dim3(kCompactThreads), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_compact_scatter(const RadianceArgs& a, int texel_bytes, hipStream_t stream) {
    if (!a.compact || a.n_tiles <= 0) return hipSuccess;
    if (texel_bytes == 16)
        hipLaunchKernelGGL(compact_scatter_kernel<16>, dim3(a.n_tiles), dim3(kCompactThreads), 0, stream, a);
    else
        hipLaunchKernelGGL(compact_scatter_kernel<8>, dim3(a.n_tiles), dim3(kCompactThreads), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_compact_batch(const RadianceArgs* args_dev, const int* tbegin_dev, int n_worlds, int total_tiles,
                                hipStream_t stream) {
    if (n_worlds <= 0 || total_tiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(compact_count_batch_kernel, dim3(total_tiles), dim3(kCompactThreads), 0, stream, args_dev,
                       tbegin_dev, n_worlds);
    hipLaunchKernelGGL(compact_scan_batch_kernel, dim3(n_worlds), dim3(kCompactThreads), 0, stream, args_dev);
    hipLaunchKernelGGL(compact_list_batch_kernel, dim3(total_tiles), dim3(kCompactThreads), 0, stream, args_dev,
                       tbegin_dev, n_worlds);
    return hipGetLastError();
}

hipError_t launch_compact_scatter_batch(const RadianceArgs* args_dev, const int* tbegin_dev, int n_worlds,
                                        int total_tiles, int texel_bytes, hipStream_t stream) {
    if (n_worlds <= 0 || total_tiles <= 0) return hipSuccess;
    if (texel_bytes == 16)
        hipLaunchKernelGGL(compact_scatter_batch_kernel<16>, dim3(total_tiles), dim3(kCompactThreads), 0, stream,
                           args_dev, tbegin_dev, n_worlds);
    else
        hipLaunchKernelGGL(compact_scatter_batch_kernel<8>, dim3(total_tiles), dim3(kCompactThreads), 0, stream,
                           args_dev, tbegin_dev, n_worlds);
    return hipGetLastError();
}

hipError_t launch_scatter(const RadianceArgs& a, int texel_bytes, hipStream_t stream) {
    if (a.n_probes <= 0) return hipSuccess;
    const int grid = (a.n_probes + 255) / 256;
    if (texel_bytes == 16) hipLaunchKernelGGL(scatter_kernel<16>, dim3(grid), dim3(256), 0, stream, a);
    else hipLaunchKernelGGL(scatter_kernel<8>, dim3(grid), dim3(256), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_scatter_batch(const RadianceArgs* args_dev, const int* pbegin_dev, int n_worlds, int total_probes,
                                int texel_bytes, hipStream_t stream) {
    if (n_worlds <= 0 || total_probes <= 0) return hipSuccess;
    const int grid = (total_probes + 255) / 256;
    if (texel_bytes == 16)
        hipLaunchKernelGGL(scatter_batch_kernel<16>, dim3(grid), dim3(256), 0, stream, args_dev, pbegin_dev, n_worlds);
    else
        hipLaunchKernelGGL(scatter_batch_kernel<8>, dim3(grid), dim3(256), 0, stream, args_dev, pbegin_dev, n_worlds);
    return hipGetLastError();
}

// occ: march over the occupancy bits (global memory) instead of the u8 palette.
int rays_per_lane_for(int rays, int want) {
    if (want != 2 && want != 4) return 1;
    if ((rays & (rays - 1)) || rays < want) return 1;
    return want;
}

hipError_t launch_radiance(const RadianceArgs& a, int store, bool traced, bool occ, int rays_per_lane,
                           hipStream_t stream) {
    const int K = traced || occ ? 1 : rays_per_lane_for(a.rays, rays_per_lane);
    return occ ? launch_lane<true>(a, store, traced, K, stream) : launch_lane<false>(a, store, traced, K, stream);
}

}  // namespace lumx

// ---------------------------------------------------------------------------
// Tiled persistent radiance kernel.
//
// A batch is P consecutive probes x R rays.  Ray q of a batch is direction
// l = q / P of probe p = q % P, so the 64 lanes of a wave trace 64/P
// directions from P neighbouring probes: similar lengths and neighbouring
// addresses.  Per-ray lights go to LDS and each probe's mean is reduced there
// in the fixed order (butterfly-equivalent tree over min(R,64) rays, chunk
// sums in chunk order, one division by R).  Workgroups are persistent and
// take batches from a global counter that the last workgroup resets to 0.
// LDS holds the material table and, when they fit, the empty-block bits and
// the palette occupancy bits.
// ---------------------------------------------------------------------------
namespace lumx {

size_t radiance_tiled_lds_bytes(const RadianceArgs& a, const RadianceTune& t) {
    size_t b = 1536 * sizeof(float);                                   // materials
    b += 3 * sizeof(float) * (size_t)t.probes_per_batch * a.rays;      // per-ray light
    if (t.eb_lds) b += sizeof(uint32_t) * (size_t)a.eb_words;
    if (t.pb_lds) b += sizeof(uint32_t) * (size_t)a.pb_words;
    return b + 16;
}

template <int STORE, bool TRACED, bool EB_LDS, bool PB_LDS>
__global__ void __launch_bounds__(1024) radiance_tiled_kernel(RadianceArgs a) {
    extern __shared__ float smem[];
    __shared__ int s_batch;
    const int T = blockDim.x, tid = threadIdx.x;
    const int R = a.rays, P = a.probes_per_batch, NR = P * R;
    float* s_mats = smem;
    float* s_lx = s_mats + 1536;
    float* s_ly = s_lx + NR;
    float* s_lz = s_ly + NR;
    uint32_t* s_eb = (uint32_t*)(s_lz + NR);
    uint32_t* s_pb = s_eb + (EB_LDS ? a.eb_words : 0);

    for (int i = tid; i < 1536; i += T) s_mats[i] = a.world.mats[i];
    if (EB_LDS)
        for (int i = tid; i < a.eb_words; i += T) s_eb[i] = a.empty_bits[i];
    if (PB_LDS)
        for (int i = tid; i < a.pb_words; i += T) s_pb[i] = a.pal_bits[i];

    const WorldView& wv = a.world;
    OccView occ;
    occ.ids = wv.blocks;
    occ.palette = wv.palette;
    occ.eb = EB_LDS ? s_eb : a.empty_bits;
    occ.pb = PB_LDS ? s_pb : a.pal_bits;
    occ.wx = wv.wx;
    occ.wy = wv.wy;
    occ.wz = wv.wz;
    // Staged palette bits cover the slots in use; every slot above them is air.
    occ.palette_slots = PB_LDS ? a.pb_words / 128 : wv.palette_slots;
    occ.block0_empty = wv.block0_empty;
    const int C = R < 64 ? R : 64;
    const uint32_t t = (uint32_t)a.time_seed;

    while (true) {
        __syncthreads();  // staging done / previous batch fully consumed
        if (tid == 0) s_batch = (int)atomicAdd(&a.batch_counter[0], 1u);
        __syncthreads();
        const int batch = s_batch;
        if (batch >= a.n_batches) break;

        for (int q = tid; q < NR; q += T) {
            const int l = q / P, p = q - l * P;
            const int wg = batch * P + p;
            f3 light = {0.0f, 0.0f, 0.0f};
            int px = 0, py = 0, pz = 0;
            bool valid = false;
            if (wg < dispatched_probes(a)) {
                if (a.requests) {
                    int4 rq = a.requests[wg * a.size + a.shift];
                    px = rq.x;
                    py = rq.y;
                    pz = rq.z;
                } else {
                    probe_xyz(a, (uint32_t)wg, &px, &py, &pz);
                }
                valid = (unsigned)px < (unsigned)wv.wx && (unsigned)py < (unsigned)wv.wy &&
                        (unsigned)pz < (unsigned)wv.wz;
            }
            if (valid) {
                const uint32_t gidx = (uint32_t)wg * (uint32_t)R + (uint32_t)l;
                f3 r0, r1;
                hash33(gidx + t, t, t, &r0);
                hash33(gidx + t + 1u, t, t, &r1);
                const float4 fb = a.fib[l];
                f3 corner = f3{(float)px, (float)py, (float)pz} * 16.0f;
                f3 center = corner + f3{8.0f, 8.0f, 8.0f};
                const f3 half = {0.5f, 0.5f, 0.5f};
                f3 dir = normalize3(f3{fb.x, fb.y, fb.z} + 0.1f * ((r0 - half) * 2.0f));
                f3 pos = center + dir * 6.0f;
                pos = pos + (r1 - half) * 2.0f;
                if (l & 1) dir = -dir;
                RayHit h = cast_ray_occ(occ, pos, dir, a.max_dist, a.max_steps);
                if (h.hit) {
                    const float* m = s_mats + h.voxel * 6;
                    const f3 color = {m[0], m[1], m[2]};
                    const float emit = m[4];
                    f3 prop = sample_radiance<STORE>(a.field_in, wv.wx, wv.wy, wv.wz, pos + h.fraction * dir,
                                                     h.normal);
                    light = light + (prop * color + color * emit);
                } else {
                    float gl = -dot3(dir, radiance_light_dir());
                    if (gl > 0.95f) {
                        f3 sun = f3{0.9f, 0.9f, 0.6f} * 3.0f;
                        light = light + gl * sun;
                    }
                    f3 sky = f3{0.53f, 0.81f, 0.92f} * 0.1f;
                    light = light + sky;
                }
                if (TRACED) {
                    size_t ri = (size_t)wg * R + l;
                    if (a.ray_keys) a.ray_keys[ri] = h.hit ? h.key : 0;
                    if (a.ray_steps) {
                        a.ray_steps[2 * ri + 0] = h.coarse;
                        a.ray_steps[2 * ri + 1] = h.refine;
                    }
                }
            }
            s_lx[q] = light.x;
            s_ly[q] = light.y;
            s_lz[q] = light.z;
        }
        __syncthreads();
        // Tree over rays within chunks of C: light[l] += light[l + s], s = C/2 .. 1.
        for (int s = C >> 1; s >= 1; s >>= 1) {
            const int n_pairs = (R / C) * s * P;
            for (int i = tid; i < n_pairs; i += T) {
                const int p = i % P;
                const int rest = i / P;
                const int l = rest % s;
                const int c = rest / s;
                const int q0 = (c * C + l) * P + p, q1 = q0 + s * P;
                s_lx[q0] = s_lx[q0] + s_lx[q1];
                s_ly[q0] = s_ly[q0] + s_ly[q1];
                s_lz[q0] = s_lz[q0] + s_lz[q1];
            }
            __syncthreads();
        }
        for (int p = tid; p < P; p += T) {
            const int wg = batch * P + p;
            if (wg >= dispatched_probes(a)) continue;
            int px, py, pz;
            if (a.requests) {
                int4 rq = a.requests[wg * a.size + a.shift];
                px = rq.x;
                py = rq.y;
                pz = rq.z;
            } else {
                probe_xyz(a, (uint32_t)wg, &px, &py, &pz);
            }
            if ((unsigned)px >= (unsigned)wv.wx || (unsigned)py >= (unsigned)wv.wy || (unsigned)pz >= (unsigned)wv.wz)
                continue;
            f3 total = {s_lx[p], s_ly[p], s_lz[p]};
            for (int c = 1; c < R / C; ++c) {
                const int q = c * C * P + p;
                total = total + f3{s_lx[q], s_ly[q], s_lz[q]};
            }
            f3 mean = probe_mean(total, R);
            const unsigned ti = (unsigned)px + (unsigned)wv.wx * ((unsigned)py + (unsigned)wv.wy * (unsigned)pz);
            f3 old = Texel<STORE>::load(a.field_in, ti);
            old = {fminf(fmaxf(old.x, 0.0f), 1.0f), fminf(fmaxf(old.y, 0.0f), 1.0f), fminf(fmaxf(old.z, 0.0f), 1.0f)};
            const float al = a.alpha, bl = 1.0f - a.alpha;
            f3 nv = {old.x * bl + mean.x * al, old.y * bl + mean.y * al, old.z * bl + mean.z * al};
            Texel<STORE>::store(a.field_out, a.staged ? (unsigned)wg : ti, unorm_nan_to_zero(nv));
        }
    }
    // The last workgroup to finish resets the batch counter for the next launch.
    if (tid == 0) {
        __threadfence();
        unsigned int done = atomicAdd(&a.batch_counter[1], 1u);
        if (done == gridDim.x - 1) {
            atomicExch(&a.batch_counter[0], 0u);
            atomicExch(&a.batch_counter[1], 0u);
        }
    }
}

template <int STORE, bool TRACED, bool EB, bool PB>
static hipError_t launch_tiled_inst(const RadianceArgs& a, const RadianceTune& t, int n_cu, hipStream_t stream) {
    const size_t lds = radiance_tiled_lds_bytes(a, t);
    auto kern = radiance_tiled_kernel<STORE, TRACED, EB, PB>;
    // Occupancy per (threads, LDS size) for this instantiation, cached.
    static thread_local int cached_threads = -1;
    static thread_local size_t cached_lds = 0;
    static thread_local int cached_per_cu = 0;
    int per_cu = 0;
    if (cached_threads == t.threads && cached_lds == lds) {
        per_cu = cached_per_cu;
    } else {
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, t.threads, lds);
        if (e != hipSuccess) return e;
        cached_threads = t.threads;
        cached_lds = lds;
        cached_per_cu = per_cu;
    }
    if (per_cu < 1) return hipErrorInvalidConfiguration;
    if (t.wg_per_cu > 0 && t.wg_per_cu < per_cu) per_cu = t.wg_per_cu;
    long long grid = (long long)per_cu * n_cu;
    if (grid > a.n_batches) grid = a.n_batches;
    if (grid < 1) return hipSuccess;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(t.threads), lds, stream, a);
    return hipGetLastError();
}

template <int STORE, bool TRACED>
static hipError_t launch_tiled_store(const RadianceArgs& a, const RadianceTune& t, int n_cu, hipStream_t s) {
    if (t.eb_lds && t.pb_lds) return launch_tiled_inst<STORE, TRACED, true, true>(a, t, n_cu, s);
    if (t.eb_lds) return launch_tiled_inst<STORE, TRACED, true, false>(a, t, n_cu, s);
    if (t.pb_lds) return launch_tiled_inst<STORE, TRACED, false, true>(a, t, n_cu, s);
    return launch_tiled_inst<STORE, TRACED, false, false>(a, t, n_cu, s);
}

hipError_t launch_radiance_tiled(RadianceArgs a, int store, bool traced, const RadianceTune& t, int n_cu,
                                 hipStream_t stream) {
    if (a.n_probes <= 0) return hipSuccess;
    a.probes_per_batch = t.probes_per_batch;
    a.n_batches = (a.n_probes + t.probes_per_batch - 1) / t.probes_per_batch;
    if (traced) {
        switch (store) {
            case STORE_F16: return launch_tiled_store<STORE_F16, true>(a, t, n_cu, stream);
            case STORE_BF16: return launch_tiled_store<STORE_BF16, true>(a, t, n_cu, stream);
            default: return launch_tiled_store<STORE_F32, true>(a, t, n_cu, stream);
        }
    }
    switch (store) {
        case STORE_F16: return launch_tiled_store<STORE_F16, false>(a, t, n_cu, stream);
        case STORE_BF16: return launch_tiled_store<STORE_BF16, false>(a, t, n_cu, stream);
        default: return launch_tiled_store<STORE_F32, false>(a, t, n_cu, stream);
    }
}

// Per-block "id == 0" bits, one thread per 32 blocks.
__global__ void __launch_bounds__(256) block_bits_kernel(const int16_t* blocks, int n, uint32_t* bits) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    const int base = w * 32;
    if (base >= n) return;
    uint32_t v = 0;
    for (int k = 0; k < 32 && base + k < n; ++k) v |= (blocks[base + k] == 0 ? 1u : 0u) << k;
    bits[w] = v;
}

// Per-slot voxel-occupancy bits: one workgroup of 128 threads per slot, one
// output word (32 voxels) per thread.
__global__ void __launch_bounds__(128) palette_bits_kernel(const uint8_t* palette, int first, uint32_t* bits) {
    const int slot = first + blockIdx.x;
    const uint8_t* src = palette + ((size_t)slot << 12) + threadIdx.x * 32;
    const uint4 a0 = ((const uint4*)src)[0];
    const uint4 a1 = ((const uint4*)src)[1];
    const uint32_t w[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
    uint32_t v = 0;
#pragma unroll
    for (int k = 0; k < 32; ++k) v |= (((w[k >> 2] >> ((k & 3) * 8)) & 0xffu) != 0 ? 1u : 0u) << k;
    bits[((size_t)slot << 7) + threadIdx.x] = v;
}

hipError_t launch_block_bits(const int16_t* blocks, int n_blocks, uint32_t* bits, hipStream_t stream) {
    int words = (n_blocks + 31) / 32;
    if (words == 0) return hipSuccess;
    hipLaunchKernelGGL(block_bits_kernel, dim3((words + 255) / 256), dim3(256), 0, stream, blocks, n_blocks, bits);
    return hipGetLastError();
}

hipError_t launch_palette_bits(const uint8_t* palette, int first_slot, int n_slots, uint32_t* bits,
                               hipStream_t stream) {
    if (n_slots <= 0) return hipSuccess;
    hipLaunchKernelGGL(palette_bits_kernel, dim3(n_slots), dim3(128), 0, stream, palette, first_slot, bits);
    return hipGetLastError();
}

}  // namespace lumx
