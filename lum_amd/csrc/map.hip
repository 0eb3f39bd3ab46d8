// map.hip - model->world voxelization and palette slot copies (gfx950).
//
// map_group_kernel restates shaders/map.comp:33-68 of the reference: per
// candidate world voxel of the model's dispatch box (one shader invocation
// each; here a thread takes four, see map_tile), FP32 inverse transform
// (unfused, rows evaluated ((m0*x + m1*y) + m2*z) + m3), truncation to model
// voxel, predicated byte store into the palette slot the blockified world
// assigns to that block.
//
// Threads whose world coordinate g = gid + shift is negative on an axis
// alias, as in the shader: the coordinate truncates toward zero (g = -1 gives
// world voxel 0), its block is the truncating iw / 16 (block 0 for
// -16 < iw < 0) and its voxel within the block is iw % 16 with the sign of
// the divisor, iw & 15 (OpSMod).  So they write voxels of block 0 that
// non-negative threads write too.  Where several threads of one model write
// one voxel, the one with the highest dispatch index (z, then y, then x) is
// the value kept.  Per axis the map from g to the voxel is one-to-one within
// g < 0 and within g >= 0, and a thread with g >= 0 has the higher index, so
// a model's threads split into sign classes (bit 2: z >= 0, bit 1: y >= 0,
// bit 0: x >= 0) that write disjoint voxels, and mapping the classes in
// increasing order keeps the highest index.  MapArgs::cls selects one class.
//
// copy_kernel restates the transfer part of exec_copies
// (src/renderer/src/internal_render.cpp:1375-1476): every slot allocated by
// blockify receives a copy of its static source block, or zeros.
#include <algorithm>

#include "lumx_internal.hpp"

namespace lumx {

// Voxel (gx, gy, gz) of the model's dispatch box: loads its block id and
// model voxel; true (with the palette byte offset and value) when it stores.
__device__ __forceinline__ bool map_voxel(const MapArgs& a, int gx, int gy, int gz, uint32_t* off, uint8_t* val) {
    const int cls = (gx + a.shift_x >= 0 ? 1 : 0) | (gy + a.shift_y >= 0 ? 2 : 0) | (gz + a.shift_z >= 0 ? 4 : 0);
    if (cls != a.cls) return false;

    const float wvx = ((float)gx + 0.5f) + (float)a.shift_x;
    const float wvy = ((float)gy + 0.5f) + (float)a.shift_y;
    const float wvz = ((float)gz + 0.5f) + (float)a.shift_z;
    const int iwx = (int)wvx, iwy = (int)wvy, iwz = (int)wvz;

    const int tbx = iwx / 16, tby = iwy / 16, tbz = iwz / 16;
    float mv[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const float p0 = a.m[r * 4 + 0] * wvx;
        const float p1 = a.m[r * 4 + 1] * wvy;
        const float p2 = a.m[r * 4 + 2] * wvz;
        mv[r] = ((p0 + p1) + p2) + a.m[r * 4 + 3];
    }
    const int imx = (int)mv[0], imy = (int)mv[1], imz = (int)mv[2];
    // The block id and the model voxel are independent loads: both are issued
    // before either is tested.
    int target = 0;
    if ((unsigned)tbx < (unsigned)a.wx && (unsigned)tby < (unsigned)a.wy && (unsigned)tbz < (unsigned)a.wz)
        target = a.blocks[tbx + a.wx * (tby + a.wy * tbz)];
    uint8_t v = 0;
    if ((unsigned)imx < (unsigned)a.mx && (unsigned)imy < (unsigned)a.my && (unsigned)imz < (unsigned)a.mz)
        v = a.model[imx + (size_t)a.mx * (imy + (size_t)a.my * imz)];
    if (target < a.static_count || target >= a.palette_slots || v == 0) return false;
    *off = ((uint32_t)target << 12) | (uint32_t)((iwx & 15) | ((iwy & 15) << 4) | ((iwz & 15) << 8));
    *val = v;
    return true;
}

// Several models whose write boxes are disjoint in one launch: model i owns
// workgroups [begin[i], begin[i + 1]).  Disjoint boxes write disjoint palette
// bytes, so the result equals mapping them one after another.  A workgroup of
// 256 threads covers a 16 x 4 x kMapTileZ tile of its model's dispatch box
// (x fastest, so a row of 16 threads stores into one 16-byte palette row);
// each thread takes kMapTileZ / 4 voxels 4 apart in z, issuing all their
// loads before any store.  Within one sign class no two voxels of a model
// write one palette byte, so which thread takes which voxel does not matter.
// The tile is found by 32-bit divisions of the workgroup index, once per
// workgroup on the scalar unit.
__device__ __forceinline__ void map_tile(const MapArgs& a, uint32_t t) {
    constexpr int kPer = kMapTileZ / 4;
    const uint32_t r = fast_div(t, a.div_tiles_x);
    const uint32_t bx = t - r * (uint32_t)a.tiles_x;
    const uint32_t bz = fast_div(r, a.div_tiles_y);
    const uint32_t by = r - bz * (uint32_t)a.tiles_y;
    const int gx = (int)(bx * 16u + (threadIdx.x & 15u));
    const int gy = (int)(by * 4u + ((threadIdx.x >> 4) & 3u));
    const int gz0 = (int)(bz * (uint32_t)kMapTileZ + (threadIdx.x >> 6));
    if (gx >= a.nx) return;
    uint32_t off[kPer];
    uint8_t val[kPer];
    bool st[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int gz = gz0 + 4 * k;
        st[k] = gz < a.nz && map_voxel(a, gx, gy, gz, &off[k], &val[k]);
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k)
        if (st[k]) a.palette[off[k]] = val[k];
}

__global__ void __launch_bounds__(256) map_group_kernel(MapGroup g) {
    const int b = blockIdx.x;
    int i = 0;
    while (i + 1 < g.n && g.begin[i + 1] <= b) ++i;
    map_tile(g.m[i], (uint32_t)(b - g.begin[i]));
}

// Last index i of begin[0 .. n - 1] (non-decreasing) with begin[i] <= x, on
// the scalar unit when x is wave-uniform.
template <class T>
__device__ __forceinline__ int find_range(const T* begin, int n, T x) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (begin[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Models of many worlds, descriptors in memory; the same per-thread work as
// map_group_kernel.
__global__ void __launch_bounds__(256) map_batch_kernel(const MapArgs* args, const int* begin, int n) {
    const int b = (int)blockIdx.x;
    const int i = __builtin_amdgcn_readfirstlane(find_range(begin, n, b));
    const MapArgs a = args[i];
    map_tile(a, (uint32_t)(b - begin[i]));
}

// Elements of a world in the commit launch: its blocks, its empty-block bit
// words, its copy list (slots st .. hi - 1), its counter.
__host__ __device__ inline int commit_elements(int n_blocks, int st, int hi) {
    return n_blocks + (n_blocks + 31) / 32 + (hi - st) + 1;
}

int map_commit_elements(const MapWorldDesc& d) { return commit_elements(d.n_blocks, d.st, d.hi); }

// Last index i of d[0 .. n - 1] whose range starts at or before workgroup b
// (ranges of empty worlds start where the next one does).
template <int MapWorldDesc::*Begin>
__device__ __forceinline__ int world_of(const MapWorldDesc* d, int n, int b) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (d[mid].*Begin <= b) lo = mid;
        else hi = mid - 1;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

__global__ void __launch_bounds__(256) map_commit_batch_kernel(const MapWorldDesc* d, int n, const char* stage) {
    const MapWorldDesc& w = d[world_of<&MapWorldDesc::group_begin>(d, n, (int)blockIdx.x)];
    int j = ((int)blockIdx.x - w.group_begin) * 256 + (int)threadIdx.x;
    const bool staged = w.stage_origin >= 0;
    const int16_t* sb = staged ? (const int16_t*)(stage + w.stage_origin) : w.origin;
    if (j < w.n_blocks) {
        const int16_t b = sb[j];
        w.blocks[j] = b;
        if (staged) w.origin[j] = b;
        return;
    }
    j -= w.n_blocks;
    const int ebw = (w.n_blocks + 31) / 32;
    if (j < ebw) {
        const int base = j * 32;
        uint32_t v = 0;
        for (int k = 0; k < 32 && base + k < w.n_blocks; ++k) v |= (sb[base + k] == 0 ? 1u : 0u) << k;
        w.empty_bits[j] = v;
        return;
    }
    j -= ebw;
    if (j < w.hi - w.st) {
        w.copy_src[w.st + j] = ((const int32_t*)(stage + w.stage_copy))[j];
        return;
    }
    if (j == w.hi - w.st) w.state[0] = w.counter;
}

// One workgroup of 256 threads per slot, 16 B per thread: exec_copies; the
// first thread also gives an allocated slot's block its id (after the commit
// launch wrote the origin there).
__global__ void __launch_bounds__(256) map_copies_batch_kernel(const MapWorldDesc* d, int n, const char* stage) {
    const int b = (int)blockIdx.x;
    const MapWorldDesc& w = d[world_of<&MapWorldDesc::slot_begin>(d, n, b)];
    const int slot = w.st + (b - w.slot_begin);
    if (threadIdx.x == 0 && slot < w.counter) {
        const int blk = ((const int32_t*)(stage + w.stage_slots))[slot - w.st];
        w.blocks[blk] = (int16_t)slot;
        atomicAnd(w.empty_bits + (blk >> 5), ~(1u << (blk & 31)));
    }
    const int src = w.copy_src[slot];
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (src != 0) v = ((const uint4*)(w.palette + ((size_t)src << 12)))[threadIdx.x];
    ((uint4*)(w.palette + ((size_t)slot << 12)))[threadIdx.x] = v;
}

// One workgroup of 128 threads per slot: the slot's voxel-occupancy words.
__global__ void __launch_bounds__(128) palette_bits_batch_kernel(const MapWorldDesc* d, int n) {
    const int b = (int)blockIdx.x;
    const MapWorldDesc& w = d[world_of<&MapWorldDesc::slot_begin>(d, n, b)];
    const int slot = w.st + (b - w.slot_begin);
    const uint8_t* src = w.palette + ((size_t)slot << 12) + threadIdx.x * 32;
    const uint4 a0 = ((const uint4*)src)[0];
    const uint4 a1 = ((const uint4*)src)[1];
    const uint32_t words[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
    uint32_t v = 0;
#pragma unroll
    for (int k = 0; k < 32; ++k) v |= (((words[k >> 2] >> ((k & 3) * 8)) & 0xffu) != 0 ? 1u : 0u) << k;
    w.pal_bits[((size_t)slot << 7) + threadIdx.x] = v;
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

hipError_t launch_map_group(const MapGroup& g, hipStream_t stream) {
    if (g.n <= 0 || g.begin[g.n] <= 0) return hipSuccess;
    hipLaunchKernelGGL(map_group_kernel, dim3(g.begin[g.n]), dim3(256), 0, stream, g);
    return hipGetLastError();
}

hipError_t launch_map_batch(const MapArgs* args_dev, const int* begin_dev, int n_models, int total_groups,
                            hipStream_t stream) {
    if (n_models <= 0 || total_groups <= 0) return hipSuccess;
    hipLaunchKernelGGL(map_batch_kernel, dim3(total_groups), dim3(256), 0, stream, args_dev, begin_dev, n_models);
    return hipGetLastError();
}

hipError_t launch_map_commit_batch(const MapWorldDesc* d_dev, int n, int total_groups, const char* stage_dev,
                                   hipStream_t stream) {
    if (n <= 0 || total_groups <= 0) return hipSuccess;
    hipLaunchKernelGGL(map_commit_batch_kernel, dim3(total_groups), dim3(256), 0, stream, d_dev, n, stage_dev);
    return hipGetLastError();
}

hipError_t launch_map_copies_batch(const MapWorldDesc* d_dev, int n, int total_slots, const char* stage_dev,
                                   hipStream_t stream) {
    if (n <= 0 || total_slots <= 0) return hipSuccess;
    hipLaunchKernelGGL(map_copies_batch_kernel, dim3(total_slots), dim3(256), 0, stream, d_dev, n, stage_dev);
    return hipGetLastError();
}

hipError_t launch_palette_bits_batch(const MapWorldDesc* d_dev, int n, int total_slots, hipStream_t stream) {
    if (n <= 0 || total_slots <= 0) return hipSuccess;
    hipLaunchKernelGGL(palette_bits_batch_kernel, dim3(total_slots), dim3(128), 0, stream, d_dev, n);
    return hipGetLastError();
}

hipError_t launch_exec_copies(const CopyArgs& a, hipStream_t stream) {
    if (a.n_slots <= 0) return hipSuccess;
    hipLaunchKernelGGL(copy_kernel, dim3(a.n_slots), dim3(256), 0, stream, a);
    return hipGetLastError();
}

}  // namespace lumx
