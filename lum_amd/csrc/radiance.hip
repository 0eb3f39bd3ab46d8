// radiance.hip - per-probe radiance update (gfx950).
//
// Restates shaders/radiance.comp:361-413 (main), :51-66 (hash33/Random3D),
// :316-347 (trace_ray) and :349-354 (store_probe) of the reference.
//
// Launch shape: one thread per ray.  A probe's R rays are R consecutive
// lanes; with R <= 64 a wave64 holds 64/R whole probes, with R a multiple
// of 64 a probe spans R/64 whole waves of one workgroup.  The probe mean is
// reduced in a fixed order (butterfly over min(R,64) lanes, then chunk sums
// in chunk order, then one division by R) so results do not depend on
// timing.  Jacobi semantics: rays sample field_in, the probe is written to
// field_out.
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

template <int STORE, bool TRACED>
__global__ void __launch_bounds__(1024) radiance_kernel(RadianceArgs a) {
    extern __shared__ f3 chunk_sum[];
    const int R = a.rays;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int wg = tid / R;
    const int lane = tid - wg * R;
    const bool active = wg < a.n_probes;

    const WorldView& w = a.world;
    int px = 0, py = 0, pz = 0;
    bool valid = false;
    f3 light = {0.0f, 0.0f, 0.0f};
    if (active) {
        if (a.requests) {
            int4 q = a.requests[wg * a.size + a.shift];
            px = q.x;
            py = q.y;
            pz = q.z;
        } else {
            px = wg % w.wx;
            py = (wg / w.wx) % w.wy;
            pz = wg / (w.wx * w.wy);
        }
        valid = (unsigned)px < (unsigned)w.wx && (unsigned)py < (unsigned)w.wy && (unsigned)pz < (unsigned)w.wz;
    }
    if (valid) {
        const uint32_t t = (uint32_t)a.time_seed;
        const uint32_t gidx = (uint32_t)wg * (uint32_t)R + (uint32_t)lane;
        f3 r0, r1;
        hash33(gidx + t, t, t, &r0);
        hash33(gidx + t + 1u, t, t, &r1);
        const float4 fb = a.fib[lane];
        f3 corner = f3{(float)px, (float)py, (float)pz} * 16.0f;
        f3 center = corner + f3{8.0f, 8.0f, 8.0f};
        const f3 half = {0.5f, 0.5f, 0.5f};
        f3 dir = normalize3(f3{fb.x, fb.y, fb.z} + 0.1f * ((r0 - half) * 2.0f));
        f3 pos = center + dir * 6.0f;
        pos = pos + (r1 - half) * 2.0f;
        if (lane & 1) dir = -dir;

        RayHit h = cast_ray(w, pos, dir, a.max_dist, a.max_steps);
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
        if (TRACED) {
            size_t ri = (size_t)wg * R + lane;
            if (a.ray_keys) a.ray_keys[ri] = h.hit ? h.key : 0;
            if (a.ray_steps) {
                a.ray_steps[2 * ri + 0] = h.coarse;
                a.ray_steps[2 * ri + 1] = h.refine;
            }
        }
    }

    // Butterfly over C = min(R, 64) lanes: v[i] += v[i ^ s], s = C/2 .. 1.
    const int C = R < 64 ? R : 64;
    for (int s = C >> 1; s >= 1; s >>= 1) {
        float ox = __shfl_xor(light.x, s);
        float oy = __shfl_xor(light.y, s);
        float oz = __shfl_xor(light.z, s);
        light = light + f3{ox, oy, oz};
    }
    f3 total = light;
    if (R > 64) {
        const int wave = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) chunk_sum[wave] = light;
        __syncthreads();
        if (lane == 0) {
            const int n_chunks = R >> 6;
            total = chunk_sum[wave];
            for (int c = 1; c < n_chunks; ++c) total = total + chunk_sum[wave + c];
        }
    }
    if (valid && lane == 0) {
        f3 mean = total / (float)R;
        size_t ti = (size_t)px + (size_t)w.wx * ((size_t)py + (size_t)w.wy * pz);
        f3 old = Texel<STORE>::load(a.field_in, ti);
        old = {fminf(fmaxf(old.x, 0.0f), 1.0f), fminf(fmaxf(old.y, 0.0f), 1.0f), fminf(fmaxf(old.z, 0.0f), 1.0f)};
        const float al = a.alpha, bl = 1.0f - a.alpha;
        f3 nv = {old.x * bl + mean.x * al, old.y * bl + mean.y * al, old.z * bl + mean.z * al};
        Texel<STORE>::store(a.field_out, ti, nv);
    }
}

template <int STORE, bool TRACED>
static hipError_t launch_store(const RadianceArgs& a, hipStream_t stream) {
    const int R = a.rays;
    int bs = R >= 256 ? R : R * (256 / R);
    long long threads = (long long)a.n_probes * R;
    int grid = (int)((threads + bs - 1) / bs);
    size_t lds = R > 64 ? sizeof(f3) * (bs / 64) : 0;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL((radiance_kernel<STORE, TRACED>), dim3(grid), dim3(bs), lds, stream, a);
    return hipGetLastError();
}

hipError_t launch_radiance(const RadianceArgs& a, int store, bool traced, hipStream_t stream) {
    if (traced) {
        switch (store) {
            case STORE_F16: return launch_store<STORE_F16, true>(a, stream);
            case STORE_BF16: return launch_store<STORE_BF16, true>(a, stream);
            default: return launch_store<STORE_F32, true>(a, stream);
        }
    }
    switch (store) {
        case STORE_F16: return launch_store<STORE_F16, false>(a, stream);
        case STORE_BF16: return launch_store<STORE_BF16, false>(a, stream);
        default: return launch_store<STORE_F32, false>(a, stream);
    }
}

}  // namespace lumx
