// map.hip - model->world voxelization and palette slot copies (gfx950).
//
// map_kernel restates shaders/map.comp:33-68 of the reference: one thread per
// candidate world voxel of the model's dispatch box, FP32 inverse transform
// (unfused, rows evaluated ((m0*x + m1*y) + m2*z) + m3), truncation to model
// voxel, predicated byte store into the palette slot the blockified world
// assigns to that block.  Threads whose world coordinate is negative write
// nothing (there the reference's writes alias and race).
//
// copy_kernel restates the transfer part of exec_copies
// (src/renderer/src/internal_render.cpp:1375-1476): every slot allocated by
// blockify receives a copy of its static source block, or zeros.
#include "lumx_internal.hpp"

namespace lumx {

__device__ __forceinline__ void map_thread(const MapArgs& a, long long t) {
    if (t >= (long long)a.nx * a.ny * a.nz) return;
    const int gx = (int)(t % a.nx);
    const int gy = (int)((t / a.nx) % a.ny);
    const int gz = (int)(t / ((long long)a.nx * a.ny));
    if (gx + a.shift_x < 0 || gy + a.shift_y < 0 || gz + a.shift_z < 0) return;

    const float wvx = ((float)gx + 0.5f) + (float)a.shift_x;
    const float wvy = ((float)gy + 0.5f) + (float)a.shift_y;
    const float wvz = ((float)gz + 0.5f) + (float)a.shift_z;
    const int iwx = (int)wvx, iwy = (int)wvy, iwz = (int)wvz;

    const int tbx = iwx / 16, tby = iwy / 16, tbz = iwz / 16;
    int target = 0;
    if ((unsigned)tbx < (unsigned)a.wx && (unsigned)tby < (unsigned)a.wy && (unsigned)tbz < (unsigned)a.wz)
        target = a.blocks[tbx + a.wx * (tby + a.wy * tbz)];
    if (target < a.static_count || target >= a.palette_slots) return;

    float mv[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const float p0 = a.m[r * 4 + 0] * wvx;
        const float p1 = a.m[r * 4 + 1] * wvy;
        const float p2 = a.m[r * 4 + 2] * wvz;
        mv[r] = ((p0 + p1) + p2) + a.m[r * 4 + 3];
    }
    const int imx = (int)mv[0], imy = (int)mv[1], imz = (int)mv[2];
    if ((unsigned)imx >= (unsigned)a.mx || (unsigned)imy >= (unsigned)a.my || (unsigned)imz >= (unsigned)a.mz)
        return;
    const uint8_t v = a.model[imx + (size_t)a.mx * (imy + (size_t)a.my * imz)];
    if (v == 0) return;
    const uint32_t off = ((uint32_t)target << 12) | (uint32_t)((iwx & 15) | ((iwy & 15) << 4) | ((iwz & 15) << 8));
    a.palette[off] = v;
}

__global__ void __launch_bounds__(256) map_kernel(MapArgs a) {
    map_thread(a, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// Several models whose write boxes are disjoint in one launch: model i owns
// workgroups [begin[i], begin[i + 1]).  Disjoint boxes write disjoint palette
// bytes, so the result equals mapping them one after another.
__global__ void __launch_bounds__(256) map_group_kernel(MapGroup g) {
    const int b = blockIdx.x;
    int i = 0;
    while (i + 1 < g.n && g.begin[i + 1] <= b) ++i;
    map_thread(g.m[i], (long long)(b - g.begin[i]) * 256 + threadIdx.x);
}

// One workgroup of 256 threads per slot, 16 B per thread.
__global__ void __launch_bounds__(256) copy_kernel(CopyArgs a) {
    const int slot = a.first_slot + blockIdx.x;
    const int src = a.copy_src[slot];
    uint4* dst = (uint4*)(a.palette + ((size_t)slot << 12));
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (src != 0) v = ((const uint4*)(a.palette + ((size_t)src << 12)))[threadIdx.x];
    dst[threadIdx.x] = v;
}

hipError_t launch_map(const MapArgs& a, hipStream_t stream) {
    long long total = (long long)a.nx * a.ny * a.nz;
    if (total <= 0) return hipSuccess;
    int grid = (int)((total + 255) / 256);
    hipLaunchKernelGGL(map_kernel, dim3(grid), dim3(256), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_map_group(const MapGroup& g, hipStream_t stream) {
    if (g.n <= 0 || g.begin[g.n] <= 0) return hipSuccess;
    hipLaunchKernelGGL(map_group_kernel, dim3(g.begin[g.n]), dim3(256), 0, stream, g);
    return hipGetLastError();
}

hipError_t launch_exec_copies(const CopyArgs& a, hipStream_t stream) {
    if (a.n_slots <= 0) return hipSuccess;
    hipLaunchKernelGGL(copy_kernel, dim3(a.n_slots), dim3(256), 0, stream, a);
    return hipGetLastError();
}

}  // namespace lumx
