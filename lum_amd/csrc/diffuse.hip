// diffuse.hip - per-pixel diffuse shading from the radiance field (gfx950).
//
// Restates shaders/diffuse.frag:228-245 (main) with its one-argument
// sample_radiance (:92-97: textureLod of the field at (position/16) /
// world_size, LINEAR + REPEAT), get_origin_from_depth (:131-137),
// load_norm / load_mat / load_depth (:105-118), GetMat (:119-129) and
// sample_lightmap (:161-218): five NEAREST taps of the D16 shadow map, 1/1024
// apart, each a depth compare LESS with MIRRORED_REPEAT addressing.  The
// radiance sample is at position = origin + normal * 6, i.e. texel-space
// coordinate position/16 - 0.5.  Output per pixel: vec4(final / 8, 1) with
// final = (2 * incoming + emission + sunlight) * colour.
// Launch: 16 x 16 pixel tiles, one thread per pixel.  Per pixel the kernel
// reads 4 B of G-buffer, 4 B of depth, 8 field texels (L2-resident), up to 5
// lightmap texels and one material row, and writes 16 B.
#include "lumx_internal.hpp"

namespace lumx {

__device__ __forceinline__ int mirror_repeat(int i, int size) {
    int m = i % (2 * size);
    if (m < 0) m += 2 * size;
    const int a = m - size;
    const int mir = a >= 0 ? a : -(1 + a);
    return (size - 1) - mir;
}

template <int STORE>
__global__ void __launch_bounds__(256) diffuse_kernel(DiffuseArgs a) {
    const int px = blockIdx.x * 16 + (threadIdx.x & 15);
    const int py = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (px >= a.width || py >= a.height) return;
    const unsigned pi = (unsigned)py * (unsigned)a.width + (unsigned)px;
    const WorldView& w = a.world;
    const uchar4 mn = a.matnorm[pi];
    const MatF mat = get_mat(w, mn.x);
    const float cx = ((float)px + 0.5f) / a.frame_w * 2.0f - 1.0f;
    const float cy = ((float)py + 0.5f) / a.frame_h * 2.0f - 1.0f;
    const float dep = a.depth[pi] * 1000.0f;
    const f3 campos = {a.campos[0], a.campos[1], a.campos[2]};
    const f3 camdir = {a.camdir[0], a.camdir[1], a.camdir[2]};
    const f3 horiz = {a.horiz[0], a.horiz[1], a.horiz[2]};
    const f3 verti = {a.verti[0], a.verti[1], a.verti[2]};
    const f3 origin = campos + horiz * cx + verti * cy + camdir * dep;
    const f3 normal = {((float)mn.y / 255.0f) * 2.0f - 1.0f, ((float)mn.z / 255.0f) * 2.0f - 1.0f,
                       ((float)mn.w / 255.0f) * 2.0f - 1.0f};
    const f3 bp = (origin + normal * 6.0f) / 16.0f;
    const f3 incoming = sample_texel<STORE>(a.field, w.wx, w.wy, w.wz, f3{bp.x - 0.5f, bp.y - 0.5f, bp.z - 0.5f});

    float sun = 0.0f;
    if (a.lightmap) {
        const f3 gl = {a.light[0], a.light[1], a.light[2]};
        const f3 biased = dot3(normal, gl) > 0.0f ? origin - normal * 0.9f : origin + normal * 0.9f;
        float c[3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
            c[r] = a.lm_proj[0 * 4 + r] * biased.x + a.lm_proj[1 * 4 + r] * biased.y +
                   a.lm_proj[2 * 4 + r] * biased.z + a.lm_proj[3 * 4 + r];
        const float world_depth = 1.0f + c[2];
        const float u = (c[0] + 1.0f) / 2.0f, v = (c[1] + 1.0f) / 2.0f;
        const float dref = fminf(fmaxf(world_depth, 0.0f), 1.0f);
        const int tap_x[5] = {-1, 0, 1, 0, 0}, tap_y[5] = {0, 0, 0, -1, 1};
        float total = 0.0f;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const float su = u + (float)tap_x[k] * (1.0f / 1024.0f);
            const float sv = v + (float)tap_y[k] * (1.0f / 1024.0f);
            const int i = mirror_repeat((int)floorf(su * (float)a.lm_w), a.lm_w);
            const int j = mirror_repeat((int)floorf(sv * (float)a.lm_h), a.lm_h);
            const float dtex = (float)a.lightmap[(unsigned)j * (unsigned)a.lm_w + (unsigned)i] / 65535.0f;
            total = total + (dref < dtex ? 1.0f : 0.0f);
        }
        sun = (total / 5.0f) * 0.15f;
    }
    const float fx = (2.0f * incoming.x + mat.emit + sun) * mat.color.x;
    const float fy = (2.0f * incoming.y + mat.emit + sun) * mat.color.y;
    const float fz = (2.0f * incoming.z + mat.emit + sun) * mat.color.z;
    a.out[pi] = make_float4(fx / 8.0f, fy / 8.0f, fz / 8.0f, 1.0f);
}

hipError_t launch_diffuse(const DiffuseArgs& a, int store, hipStream_t stream) {
    dim3 grid((a.width + 15) / 16, (a.height + 15) / 16);
    if (grid.x == 0 || grid.y == 0) return hipSuccess;
    switch (store) {
        case STORE_F16: hipLaunchKernelGGL(diffuse_kernel<STORE_F16>, grid, dim3(256), 0, stream, a); break;
        case STORE_BF16: hipLaunchKernelGGL(diffuse_kernel<STORE_BF16>, grid, dim3(256), 0, stream, a); break;
        default: hipLaunchKernelGGL(diffuse_kernel<STORE_F32>, grid, dim3(256), 0, stream, a); break;
    }
    return hipGetLastError();
}

}  // namespace lumx
