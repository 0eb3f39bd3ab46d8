// blockify.hip - blockify + exec_copies on the device (gfx950).
//
// Restates start_blockify / blockify_mesh / end_blockify and the copy queue of
// exec_copies (src/renderer/src/internal_render.cpp:977-987,1029-1097,
// 1375-1476).  The reference walks each model's block box in z, y, x order,
// model after model, and gives every block whose current id is static the next
// palette slot (copy source: its origin id).  A block is therefore allocated
// iff some model's box covers it and its origin id is static, by the first
// model (lowest index) that covers it, and slots are handed out in the order of
// the sequence (model, z, y, x) restricted to those first touches.  On the
// device that is an ordered stream compaction over the concatenated boxes:
//
//   blockify_claim_kernel    every (model, block) element whose block has a
//                            static origin id claims the block with atomicMin
//                            of its model index;
//   blockify_flags_kernel    one workgroup per 2048 elements: an element is an
//                            allocation iff its model holds the claim; flags
//                            as wave ballots, per-tile counts;
//   blockify_scan_kernel     one workgroup: tile offsets, the new counter, the
//                            copy bound max(counter, previous counter);
//   blockify_scatter_kernel  slot = static + offset + rank: the block's new id,
//                            the slot's copy source, the claim released;
//   copy_bounded_kernel      exec_copies over the slots below the copy bound:
//                            copy from the source block (0 = zero fill), zero
//                            fill above the new counter.
// The current world starts as a device copy of the origin world, so no world
// crosses PCIe.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "lumx_internal.hpp"

// Scratch of the no-allocation device map (lumx_map_models_device_noalloc):
// model boxes and prefix, ballot words, tile counts.  Every GPU holds its own
// copy until the process ends; host code reaches it only through the address
// noalloc_scratch resolves, never through this name.
static __device__ __align__(256) unsigned char scratch[lumx::kNoAllocScratchBytes];

namespace lumx {

namespace {

constexpr int kThreads = 256;
constexpr int kTile = 2048;

__global__ __launch_bounds__(64) void blockify_desc_kernel(BlockifyDesc d, int4* boxes, int* first) {
    const int i = threadIdx.x;
    if (i >= d.n) return;
    first[i] = d.first[i];
    boxes[2 * i] = d.boxes[2 * i];
    boxes[2 * i + 1] = d.boxes[2 * i + 1];
}

// Element t of the concatenated boxes: its model and block index.
__device__ __forceinline__ void element(const BlockifyArgs& a, int t, int* m_out, unsigned* b_out) {
    int lo = 0, hi = a.n_models - 1;
    while (lo < hi) {  // last model whose first element is <= t
        const int mid = (lo + hi + 1) >> 1;
        if (a.first[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    const int4 bmin = a.boxes[2 * lo], ext = a.boxes[2 * lo + 1];
    const int l = t - a.first[lo];
    const int lx = l % ext.x, ly = (l / ext.x) % ext.y, lz = l / (ext.x * ext.y);
    *m_out = lo;
    *b_out = (unsigned)(bmin.x + lx) + (unsigned)a.wx * ((unsigned)(bmin.y + ly) + (unsigned)a.wy * (unsigned)(bmin.z + lz));
}

__global__ __launch_bounds__(kThreads) void blockify_claim_kernel(BlockifyArgs a) {
    const int t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= a.n_elements) return;
    int m;
    unsigned b;
    element(a, t, &m, &b);
    if (a.origin[b] < a.static_count) atomicMin(&a.claim[b], m);
}

__global__ __launch_bounds__(kThreads) void blockify_flags_kernel(BlockifyArgs a) {
    __shared__ int s_count[kThreads / 64];
    const int tile = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int count = 0;
    for (int k = 0; k < kTile / kThreads; ++k) {
        const int t = tile * kTile + k * kThreads + threadIdx.x;
        bool flag = false;
        if (t < a.n_elements) {
            int m;
            unsigned b;
            element(a, t, &m, &b);
            flag = a.origin[b] < a.static_count && a.claim[b] == m;
        }
        const unsigned long long bits = __ballot(flag);
        if (lane == 0) a.bits[(size_t)tile * (kTile / 64) + k * (kThreads / 64) + wave] = bits;
        count += __popcll(bits);
    }
    if (lane == 0) s_count[wave] = count;
    __syncthreads();
    if (threadIdx.x == 0) {
        int c = 0;
        for (int i = 0; i < kThreads / 64; ++i) c += s_count[i];
        a.tile_counts[tile] = c;
    }
}

__global__ __launch_bounds__(1024) void blockify_scan_kernel(BlockifyArgs a) {
    __shared__ int s_wave[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, n_waves = blockDim.x >> 6;
    int running = 0;
    for (int base = 0; base < a.n_tiles; base += blockDim.x) {
        const int i = base + threadIdx.x;
        const int v = i < a.n_tiles ? a.tile_counts[i] : 0;
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
        const int total = s_wave[n_waves - 1];
        if (i < a.n_tiles) a.tile_counts[i] = running + before + x - v;
        running += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        // Slots past the palette are not handed out (the host path reports
        // the same case as an overflow); the flag records it.
        const int capacity = a.palette_slots - a.static_count;
        const int allocated = running < capacity ? running : capacity;
        const int counter = a.static_count + allocated;
        const int prev = a.state[0] > a.static_count ? a.state[0] : a.static_count;
        a.state[0] = counter;
        a.state[1] = counter > prev ? counter : prev;  // copy bound of this map
        a.state[2] = running > capacity ? 1 : 0;      // overflow
    }
}

__global__ __launch_bounds__(kThreads) void blockify_scatter_kernel(BlockifyArgs a) {
    constexpr int kWords = kTile / 64;
    __shared__ int s_pref[kWords];
    const int tile = blockIdx.x;
    const unsigned long long* words = a.bits + (size_t)tile * kWords;
    if (threadIdx.x < kWords) {
        int s = 0;
        for (int c = 0; c < (int)threadIdx.x; ++c) s += __popcll(words[c]);
        s_pref[threadIdx.x] = s;
    }
    __syncthreads();
    const int offset = a.tile_counts[tile];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = 0; k < kTile / kThreads; ++k) {
        const int c = k * (kThreads / 64) + wave;
        const unsigned long long bits = words[c];
        if (!((bits >> lane) & 1ull)) continue;
        const int t = tile * kTile + c * 64 + lane;
        int m;
        unsigned b;
        element(a, t, &m, &b);
        const int slot = a.static_count + offset + s_pref[c] + __popcll(bits & ((1ull << lane) - 1ull));
        if (slot < a.palette_slots) {
            a.copy_src[slot] = a.origin[b];
            a.blocks[b] = (int16_t)slot;
        }
        a.claim[b] = a.claim_idle;
    }
}

// exec_copies with the bounds on the device: slot s in [first, first + n)
// below state[1]: a copy of block copy_src[s] below state[0] (0 = zero
// fill), zero fill above it.  256 threads x 16 B per slot.
__global__ __launch_bounds__(256) void copy_bounded_kernel(uint8_t* palette, const int32_t* copy_src,
                                                           const int* state, int first) {
    const int slot = first + blockIdx.x;
    if (slot >= state[1]) return;
    const int src = slot < state[0] ? copy_src[slot] : 0;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (src != 0) v = ((const uint4*)(palette + (size_t)src * 4096))[threadIdx.x];
    ((uint4*)(palette + (size_t)slot * 4096))[threadIdx.x] = v;
}

}  // namespace

size_t blockify_tiles(int n_elements) { return (size_t)(n_elements + kTile - 1) / kTile; }

hipError_t launch_blockify(const BlockifyArgs& a, int copy_slots, hipStream_t stream) {
    if (a.n_elements > 0) {
        const int groups = (a.n_elements + kThreads - 1) / kThreads;
        hipLaunchKernelGGL(blockify_claim_kernel, dim3(groups), dim3(kThreads), 0, stream, a);
        hipLaunchKernelGGL(blockify_flags_kernel, dim3(a.n_tiles), dim3(kThreads), 0, stream, a);
    }
    hipLaunchKernelGGL(blockify_scan_kernel, dim3(1), dim3(1024), 0, stream, a);
    if (a.n_elements > 0)
        hipLaunchKernelGGL(blockify_scatter_kernel, dim3(a.n_tiles), dim3(kThreads), 0, stream, a);
    if (copy_slots > 0)
        hipLaunchKernelGGL(copy_bounded_kernel, dim3(copy_slots), dim3(256), 0, stream, a.palette, a.copy_src,
                           a.state, a.static_count);
    return hipGetLastError();
}

hipError_t noalloc_scratch(int device, void** out) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void* ptr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(mu);
    if (!ptr[device]) {
        hipError_t e = hipSetDevice(device);
        if (e != hipSuccess) return e;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        ptr[device] = p;
    }
    *out = ptr[device];
    return hipSuccess;
}

hipError_t launch_blockify_desc(const BlockifyDesc& d, int4* boxes, int* first, hipStream_t stream) {
    if (d.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(blockify_desc_kernel, dim3(1), dim3(64), 0, stream, d, boxes, first);
    return hipGetLastError();
}

}  // namespace lumx
