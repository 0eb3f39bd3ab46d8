// residency.hip - the device half of bringing a world into a device block
// (restore from its host snapshot, or creation): after the kept state (current
// field, blocks, copy list, materials, palette slots in use) has been copied
// in, one kernel (for one world, or for up to kAttachBatch worlds at once)
// sets everything else: the other field and the palette slots
// above the ones in use to zero, the claim words idle, the counters, and the
// occupancy bits of the blocks and of the palette slots rebuilt from the data.
#include <algorithm>

#include "lumx_internal.hpp"

namespace lumx {

namespace {

// One world's pass over the concatenation of five index ranges: elements
// first, first + stride, ...; `lead` (one thread of the world's threads)
// also resets the counters.
__device__ __forceinline__ void attach_world(const AttachArgs& a, long long first, long long stride, bool lead) {
    const long long n_field = a.n_field_words;
    const long long n_pal = (long long)(a.slots - a.used) * 512;  // uint2 words of the unused slots
    const long long n_claim = a.n_blocks;
    const long long n_eb = (a.n_blocks + 31) / 32;
    const long long n_pb = (long long)a.slots * 128;
    const long long total = n_field + n_pal + n_claim + n_eb + n_pb;
    if (lead) {
        a.batch_counter[0] = 0;
        a.batch_counter[1] = 0;
        a.blockify_state[0] = a.counter;
        a.blockify_state[1] = 0;
        a.blockify_state[2] = 0;
        a.blockify_state[3] = 0;
    }
    for (long long i = first; i < total; i += stride) {
        long long j = i;
        if (j < n_field) {
            a.field_other[j] = make_uint2(0u, 0u);
            continue;
        }
        j -= n_field;
        if (j < n_pal) {
            ((uint2*)(a.palette + ((size_t)a.used << 12)))[j] = make_uint2(0u, 0u);
            continue;
        }
        j -= n_pal;
        if (j < n_claim) {
            a.sel_claim[j] = a.claim_idle;
            a.upd_claim[j] = -1;
            continue;
        }
        j -= n_claim;
        if (j < n_eb) {
            const long long base = j * 32;
            uint32_t v = 0;
            for (int k = 0; k < 32 && base + k < a.n_blocks; ++k) v |= (a.blocks[base + k] == 0 ? 1u : 0u) << k;
            a.empty_bits[j] = v;
            continue;
        }
        j -= n_eb;
        // pal_bits word j: slot j / 128, voxels 32 (j % 128) .. + 31
        const int slot = (int)(j >> 7);
        uint32_t v = 0;
        if (slot < a.used) {
            const uint4* src = (const uint4*)(a.palette + ((size_t)j << 5));
            const uint4 a0 = src[0], a1 = src[1];
            const uint32_t w[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
#pragma unroll
            for (int k = 0; k < 32; ++k) v |= (((w[k >> 2] >> ((k & 3) * 8)) & 0xffu) != 0 ? 1u : 0u) << k;
        }
        a.pal_bits[j] = v;
    }
}

__global__ __launch_bounds__(256) void world_attach_kernel(AttachArgs a) {
    attach_world(a, (long long)blockIdx.x * blockDim.x + threadIdx.x, (long long)gridDim.x * blockDim.x,
                 blockIdx.x == 0 && threadIdx.x == 0);
}

// Up to kAttachBatch worlds in one launch, their descriptors passed by value:
// world k owns workgroups k * per .. k * per + per - 1.
struct AttachBatch {
    AttachArgs w[kAttachBatch];
    int n, per;
};

static_assert(sizeof(AttachBatch) <= 4096, "kernel arguments are limited to 4 KiB");

__global__ __launch_bounds__(256) void world_attach_batch_kernel(AttachBatch b) {
    const int k = (int)blockIdx.x / b.per;
    if (k >= b.n) return;
    const int g = (int)blockIdx.x - k * b.per;
    attach_world(b.w[k], (long long)g * blockDim.x + threadIdx.x, (long long)b.per * blockDim.x,
                 g == 0 && threadIdx.x == 0);
}

// Workgroups of one world's pass (each thread then takes a grid stride).
long long attach_groups(const AttachArgs& a) {
    const long long total = a.n_field_words + (long long)(a.slots - a.used) * 512 + a.n_blocks +
                            (a.n_blocks + 31) / 32 + (long long)a.slots * 128;
    return std::max<long long>(std::min<long long>((total + 255) / 256, 2048), 1);
}

}  // namespace

hipError_t launch_world_attach(const AttachArgs& a, hipStream_t stream) {
    hipLaunchKernelGGL(world_attach_kernel, dim3((unsigned)attach_groups(a)), dim3(256), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_world_attach_batch(const AttachArgs* a, int n, hipStream_t stream) {
    for (int first = 0; first < n; first += kAttachBatch) {
        AttachBatch b{};
        b.n = std::min(kAttachBatch, n - first);
        long long per = 1;
        for (int k = 0; k < b.n; ++k) {
            b.w[k] = a[first + k];
            per = std::max(per, attach_groups(b.w[k]));
        }
        b.per = (int)per;  // <= 2048, so the grid stays below 2^15 workgroups
        hipLaunchKernelGGL(world_attach_batch_kernel, dim3((unsigned)(b.n * b.per)), dim3(256), 0, stream, b);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace lumx
