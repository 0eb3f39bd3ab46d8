// probes.hip - probe selection and radiance field shift on the device (gfx950).
//
// Probe selection restates update_radiance's request build
// (src/renderer/src/internal_render.cpp:1142-1168): every probe, in
// x-fastest order, whose 27 neighbours (clamped to the world) sum to a
// positive block id, followed by the special requests whose probe was not
// selected.  The order of the output is the order of the reference's loops,
// so the selection is an ordered stream compaction:
//
//   select_flags_kernel   one workgroup per tile of 2048 probes: the flags as
//                         64-bit wave ballots (one word per 64 probes) and
//                         the tile's count;
//   select_scan_kernel    one workgroup: exclusive scan of the tile counts,
//                         then the specials in order (kept = inside the world
//                         and its probe not flagged; repeats are kept, as the
//                         reference marks only grid probes as taken), written
//                         after the grid probes up to one entry per probe
//                         (the capacity of the reference's request buffer),
//                         count stored;
//   select_scatter_kernel one workgroup per tile: each flagged probe goes to
//                         tile offset + rank within the tile.
//
// Field shift restates shift_radiance (src/renderer/src/internal_render.cpp:
// 1218-1335): the intersection of the field and the field moved by `shift`
// is copied so that the texel at p + shift lands on p; texels outside the
// destination box keep their value.  One pass writes the whole field into
// the other buffer of the pair.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lumx_internal.hpp"

namespace lumx {

namespace {

constexpr int kSelThreads = 256;
constexpr int kSelTile = 2048;  // probes per tile: 8 per thread, 32 ballot words
constexpr int kClaimIdle = 0x7f7f7f7f;

__device__ __forceinline__ int neighbour_sum(const int16_t* blocks, int wx, int wy, int wz, int x, int y, int z) {
    int s = 0;
    for (int dz = -1; dz <= 1; ++dz) {
        const int tz = min(max(z + dz, 0), wz - 1);
        for (int dy = -1; dy <= 1; ++dy) {
            const int ty = min(max(y + dy, 0), wy - 1);
            const int row = wx * (ty + wy * tz);
            for (int dx = -1; dx <= 1; ++dx) {
                const int tx = min(max(x + dx, 0), wx - 1);
                s += blocks[row + tx];
            }
        }
    }
    return s;
}

__global__ __launch_bounds__(kSelThreads) void select_flags_kernel(SelectArgs a) {
    const int tile = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __shared__ int s_count[kSelThreads / 64];
    int count = 0;
    for (int k = 0; k < kSelTile / kSelThreads; ++k) {
        const int p = tile * kSelTile + k * kSelThreads + threadIdx.x;
        bool flag = false;
        if (p < a.n) {
            const int x = p % a.wx, y = (p / a.wx) % a.wy, z = p / (a.wx * a.wy);
            flag = neighbour_sum(a.blocks, a.wx, a.wy, a.wz, x, y, z) > 0;
        }
        const unsigned long long m = __ballot(flag);
        // Word c of the tile holds probes tile*2048 + 64c .. +63, c = 4k + wave,
        // so probe p's flag is bit p % 64 of word p / 64.
        if (lane == 0) a.bits[(size_t)tile * (kSelTile / 64) + k * (kSelThreads / 64) + wave] = m;
        count += __popcll(m);
    }
    if (lane == 0) s_count[wave] = count;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int i = 0; i < kSelThreads / 64; ++i) t += s_count[i];
        a.tile_counts[tile] = t;
    }
}

// Inclusive scan of v over the workgroup (blockDim.x a multiple of 64, at
// most 1024).  Returns the inclusive prefix; *total gets the group's sum.
__device__ __forceinline__ int group_scan(int v, int* s_wave, int* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, n_waves = blockDim.x >> 6;
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
    *total = s_wave[n_waves - 1];
    __syncthreads();  // s_wave is reused by the next call
    return x + before;
}

__global__ __launch_bounds__(1024) void select_scan_kernel(SelectArgs a) {
    __shared__ int s_wave[16];
    int running = 0;
    for (int base = 0; base < a.n_tiles; base += blockDim.x) {
        const int i = base + threadIdx.x;
        const int v = i < a.n_tiles ? a.tile_counts[i] : 0;
        int total;
        const int incl = group_scan(v, s_wave, &total);
        if (i < a.n_tiles) a.tile_counts[i] = running + incl - v;  // now the tile's offset
        running += total;
    }
    for (int base = 0; base < a.n_special; base += blockDim.x) {
        const int i = base + threadIdx.x;
        int4 q = make_int4(0, 0, 0, 0);
        int keep = 0;
        if (i < a.n_special) {
            q = a.special[i];
            if ((unsigned)q.x < (unsigned)a.wx && (unsigned)q.y < (unsigned)a.wy && (unsigned)q.z < (unsigned)a.wz) {
                const int p = q.x + a.wx * (q.y + a.wy * q.z);
                keep = !((a.bits[p >> 6] >> (p & 63)) & 1ull);  // not selected by the grid
            }
        }
        int total;
        const int incl = group_scan(keep, s_wave, &total);
        const int pos = running + incl - 1;
        if (keep && pos < a.n) a.out[pos] = q;
        running += total;
    }
    if (threadIdx.x == 0) a.count[0] = min(running, a.n);
}

__global__ __launch_bounds__(kSelThreads) void select_scatter_kernel(SelectArgs a) {
    constexpr int kWords = kSelTile / 64;
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
    for (int k = 0; k < kSelTile / kSelThreads; ++k) {
        const int c = k * (kSelThreads / 64) + wave;
        const unsigned long long m = words[c];
        if ((m >> lane) & 1ull) {
            const int p = tile * kSelTile + c * 64 + lane;
            const int rank = __popcll(m & ((1ull << lane) - 1ull));
            const int x = p % a.wx, y = (p / a.wx) % a.wy, z = p / (a.wx * a.wy);
            a.out[offset + s_pref[c] + rank] = make_int4(x, y, z, 0);
        }
    }
}

template <int BYTES>
struct TexelWords;
template <>
struct TexelWords<16> { typedef uint4 T; };
template <>
struct TexelWords<8> { typedef uint2 T; };

template <int BYTES>
__global__ __launch_bounds__(256) void shift_kernel(ShiftArgs a) {
    typedef typename TexelWords<BYTES>::T T;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    const int x = (int)(p % a.wx), y = (int)((p / a.wx) % a.wy), z = (int)(p / ((long long)a.wx * a.wy));
    long long src = p;
    const int rx = x - a.dst[0], ry = y - a.dst[1], rz = z - a.dst[2];
    if ((unsigned)rx < (unsigned)a.extent[0] && (unsigned)ry < (unsigned)a.extent[1] &&
        (unsigned)rz < (unsigned)a.extent[2])
        src = (x + a.shift[0]) + (long long)a.wx * ((y + a.shift[1]) + (long long)a.wy * (z + a.shift[2]));
    ((T*)a.out)[p] = ((const T*)a.in)[src];
}

}  // namespace

size_t select_tiles(int n_probes) { return (size_t)(n_probes + kSelTile - 1) / kSelTile; }
size_t select_bits_bytes(int n_probes) { return select_tiles(n_probes) * (kSelTile / 8); }
int select_claim_idle() { return kClaimIdle; }

hipError_t launch_select(const SelectArgs& a, hipStream_t stream) {
    if (a.n_tiles > 0)
        hipLaunchKernelGGL(select_flags_kernel, dim3(a.n_tiles), dim3(kSelThreads), 0, stream, a);
    hipLaunchKernelGGL(select_scan_kernel, dim3(1), dim3(1024), 0, stream, a);
    if (a.n_tiles > 0)
        hipLaunchKernelGGL(select_scatter_kernel, dim3(a.n_tiles), dim3(kSelThreads), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_shift(const ShiftArgs& a, int texel_bytes, hipStream_t stream) {
    if (a.n <= 0) return hipSuccess;
    const int grid = (int)((a.n + 255) / 256);
    if (texel_bytes == 16)
        hipLaunchKernelGGL(shift_kernel<16>, dim3(grid), dim3(256), 0, stream, a);
    else
        hipLaunchKernelGGL(shift_kernel<8>, dim3(grid), dim3(256), 0, stream, a);
    return hipGetLastError();
}

}  // namespace lumx
