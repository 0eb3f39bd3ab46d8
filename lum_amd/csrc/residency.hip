// residency.hip - the device half of bringing a world into a device block
// (restore from its host snapshot, or creation): after the kept state (current
// field, blocks, copy list, materials, palette slots in use) has been copied
// in, one kernel sets everything else: the other field and the palette slots
// above the ones in use to zero, the claim words idle, the counters, and the
// occupancy bits of the blocks and of the palette slots rebuilt from the data.
#include <algorithm>

#include "lumx_internal.hpp"

namespace lumx {

namespace {

// One grid-stride pass over the concatenation of five index ranges.
__global__ __launch_bounds__(256) void world_attach_kernel(AttachArgs a) {
    const long long n_field = a.n_field_words;
    const long long n_pal = (long long)(a.slots - a.used) * 512;  // uint2 words of the unused slots
    const long long n_claim = a.n_blocks;
    const long long n_eb = (a.n_blocks + 31) / 32;
    const long long n_pb = (long long)a.slots * 128;
    const long long total = n_field + n_pal + n_claim + n_eb + n_pb;
    const long long stride = (long long)gridDim.x * blockDim.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.batch_counter[0] = 0;
        a.batch_counter[1] = 0;
        a.blockify_state[0] = a.counter;
        a.blockify_state[1] = 0;
        a.blockify_state[2] = 0;
        a.blockify_state[3] = 0;
    }
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
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

}  // namespace

hipError_t launch_world_attach(const AttachArgs& a, hipStream_t stream) {
    const long long total = a.n_field_words + (long long)(a.slots - a.used) * 512 + a.n_blocks +
                            (a.n_blocks + 31) / 32 + (long long)a.slots * 128;
    const long long groups = std::min<long long>((total + 255) / 256, 2048);
    hipLaunchKernelGGL(world_attach_kernel, dim3((unsigned)std::max<long long>(groups, 1)), dim3(256), 0, stream, a);
    return hipGetLastError();
}

}  // namespace lumx
