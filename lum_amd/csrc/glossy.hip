// glossy.hip - per-pixel glossy reflection march (gfx950).
//
// Restates shaders/glossy.frag:444-493 (main) with get_origin_from_depth
// (:43-49), ProcessHit (:312-326), trace_glossy_ray (:328-360), load_norm /
// load_mat / load_depth (:367-382) and the stencil mask of
// shaders/fillStencilGlossy.frag:29-35 (roughness > 0.5 -> pixel untouched).
// The output is the fragment's pre-blend value vec4(L / 8, 1 - roughness).
// Launch: 16 x 16 pixel tiles, one thread per pixel.
#include "lumx_internal.hpp"

namespace lumx {

template <int STORE, bool TRACED>
__device__ __forceinline__ void glossy_pixel(const GlossyArgs& a, const int px, const int py) {
    const unsigned pi = (unsigned)py * (unsigned)a.width + (unsigned)px;
    const WorldView& w = a.world;
    const uchar4 mn = a.matnorm[pi];
    const MatF mat = get_mat(w, mn.x);
    if (TRACED) a.keys[pi] = 0;
    if (mat.rough > 0.5f) return;

    const float cx = ((float)px + 0.5f) / a.frame_w * 2.0f - 1.0f;
    const float cy = ((float)py + 0.5f) / a.frame_h * 2.0f - 1.0f;
    const float dep = a.depth[pi] * 1000.0f;
    const f3 campos = {a.campos[0], a.campos[1], a.campos[2]};
    const f3 camdir = {a.camdir[0], a.camdir[1], a.camdir[2]};
    const f3 horiz = {a.horiz[0], a.horiz[1], a.horiz[2]};
    const f3 verti = {a.verti[0], a.verti[1], a.verti[2]};
    f3 origin = campos + horiz * cx + verti * cy + camdir * dep;
    const f3 normal = {((float)mn.y / 255.0f) * 2.0f - 1.0f, ((float)mn.z / 255.0f) * 2.0f - 1.0f,
                       ((float)mn.w / 255.0f) * 2.0f - 1.0f};
    f3 L = {0.0f, 0.0f, 0.0f}, R = {1.0f, 1.0f, 1.0f};
    f3 dir = camdir;
    // ProcessHit at the surface with fraction 0.
    origin = origin + 0.0f * dir;
    f3 diff = sample_radiance<STORE>(a.field, w.wx, w.wy, w.wz, origin, normal);
    R = R * mat.color;
    L = L + R * (f3{mat.emit, mat.emit, mat.emit} + diff);
    const float dn = dot3(normal, dir);
    dir = dir - (2.0f * dn) * normal;

    RayHit h = cast_ray(w, origin, dir, a.max_dist, a.max_steps);
    if (h.hit) {
        MatF hm = get_mat(w, h.voxel);
        f3 o2 = origin + h.fraction * dir;
        f3 d2 = sample_radiance<STORE>(a.field, w.wx, w.wy, w.wz, o2, h.normal);
        R = R * hm.color;
        L = L + R * (f3{hm.emit, hm.emit, hm.emit} + d2);
        if (TRACED) a.keys[pi] = h.key;
    } else {
        const f3 gl = {a.light[0], a.light[1], a.light[2]};
        float s = -dot3(dir, gl);
        if (s > 0.9f) {
            f3 sun = f3{0.9f, 0.9f, 0.6f} * 0.5f;
            L = L + ((sun * R) * (s - 0.9f)) * 10.0f;
        }
        f3 sky = f3{0.53f, 0.81f, 0.92f} * 0.1f;
        L = L + sky * R;
    }
    a.out[pi] = make_float4(L.x / 8.0f, L.y / 8.0f, L.z / 8.0f, 1.0f - mat.rough);
}

template <int STORE, bool TRACED>
__global__ void __launch_bounds__(256) glossy_kernel(GlossyArgs a) {
    const int px = blockIdx.x * 16 + (threadIdx.x & 15);
    const int py = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (px >= a.width || py >= a.height) return;
    glossy_pixel<STORE, TRACED>(a, px, py);
}

// Many frames in one launch: frame w owns workgroups [begin[w], begin[w + 1]),
// one 16 x 16 pixel tile each, tiles row-major (the single-frame kernel's
// tiling, so a wave's reflection rays start from a 16 x 4 patch).
template <int STORE>
__global__ void __launch_bounds__(256) glossy_batch_kernel(const GlossyArgs* args, const int* begin, int n) {
    const int b = (int)blockIdx.x;
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (begin[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    lo = __builtin_amdgcn_readfirstlane(lo);
    const GlossyArgs a = args[lo];
    const int t = b - begin[lo];
    const int tiles_x = (a.width + 15) >> 4;
    const int ty = t / tiles_x;
    const int px = (t - ty * tiles_x) * 16 + (int)(threadIdx.x & 15);
    const int py = ty * 16 + (int)(threadIdx.x >> 4);
    if (px >= a.width || py >= a.height) return;
    glossy_pixel<STORE, false>(a, px, py);
}

hipError_t launch_glossy_batch(const GlossyArgs* args_dev, const int* begin_dev, int n, int total_tiles, int store,
                               hipStream_t stream) {
    if (n <= 0 || total_tiles <= 0) return hipSuccess;
    const int grid = total_tiles;
    switch (store) {
        case STORE_F16:
            hipLaunchKernelGGL(glossy_batch_kernel<STORE_F16>, dim3(grid), dim3(256), 0, stream, args_dev, begin_dev, n);
            break;
        case STORE_BF16:
            hipLaunchKernelGGL(glossy_batch_kernel<STORE_BF16>, dim3(grid), dim3(256), 0, stream, args_dev, begin_dev, n);
            break;
        default:
            hipLaunchKernelGGL(glossy_batch_kernel<STORE_F32>, dim3(grid), dim3(256), 0, stream, args_dev, begin_dev, n);
            break;
    }
    return hipGetLastError();
}

template <int STORE, bool TRACED>
static hipError_t launch_g(const GlossyArgs& a, hipStream_t stream) {
    dim3 grid((a.width + 15) / 16, (a.height + 15) / 16);
    if (grid.x == 0 || grid.y == 0) return hipSuccess;
    hipLaunchKernelGGL((glossy_kernel<STORE, TRACED>), grid, dim3(256), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_glossy(const GlossyArgs& a, int store, bool traced, hipStream_t stream) {
    if (traced) {
        switch (store) {
            case STORE_F16: return launch_g<STORE_F16, true>(a, stream);
            case STORE_BF16: return launch_g<STORE_BF16, true>(a, stream);
            default: return launch_g<STORE_F32, true>(a, stream);
        }
    }
    switch (store) {
        case STORE_F16: return launch_g<STORE_F16, false>(a, stream);
        case STORE_BF16: return launch_g<STORE_BF16, false>(a, stream);
        default: return launch_g<STORE_F32, false>(a, stream);
    }
}

}  // namespace lumx
