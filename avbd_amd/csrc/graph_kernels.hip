// graph_kernels.hip -- DSATUR colouring of every scene on the device.
//
// ColorList::greedyColor2 (common/ColorList.hpp:149-241): repeatedly take the
// uncoloured vertex with the largest saturation (distinct colours among its
// coloured neighbours), ties to the largest degree, ties to the lowest index,
// and give it the lowest colour none of its neighbours has.  The order is
// inherently sequential, so one workgroup colours one scene (a batch colours
// all its scenes at once), and the parallelism is inside each round:
//   * every vertex has a packed key, kept current in memory,
//     key = saturation << 48 | degree << 32 | (0xFFFFFFFF - local index)
//     (0 once coloured), so the largest key is exactly the reference's
//     choice; every thread owns the vertices t, t + 1024, ... and keeps the
//     largest key among them;
//   * a block-wide max picks the round's vertex v;
//   * v's colour is the lowest zero bit of its neighbour-colour set (a 256-bit
//     set per vertex, the reference's ColorBitset<4> limit), kept up to date
//     incrementally, so no neighbour scan is needed to colour it;
//   * the lanes of the first wavefront walk v's neighbours: an uncoloured
//     neighbour that did not have this colour yet gains it and one unit of
//     saturation, and the thread that owns it is marked to refresh its key;
//   * only the marked threads rescan their vertices.
// Every workgroup runs exactly as many rounds as its scene has vertices.
#include "internal.hpp"

namespace avbd {
namespace {

constexpr uint32_t kDsaturBlock = 1024;
constexpr uint32_t kUncoloured = 0xFFFFFFFFu;

__device__ __forceinline__ uint64_t dsatur_key(uint32_t sat, uint32_t deg, uint32_t local) {
    return ((uint64_t)sat << 48) | ((uint64_t)deg << 32) | (uint64_t)(0xFFFFFFFFu - local);
}

__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint64_t o = __shfl_xor(v, off);
        v = o > v ? o : v;
    }
    return v;
}

// scene_off: n_scenes + 1 particle offsets; off / nb: the global static
// vertex adjacency (neighbours of a vertex lie in its scene); key, seen:
// per-vertex scratch (the packed key of every vertex, 0 once coloured; the
// neighbour-colour set, 4 words per vertex), seen zeroed by the caller;
// color: output; scene_colors: colours used per scene; status bit 0: a scene
// needed more than 256 colours; bit 1: a degree beyond 65535.
__global__ __launch_bounds__(kDsaturBlock) void dsatur_kernel(const uint32_t* __restrict__ scene_off,
                                                              const uint32_t* __restrict__ off,
                                                              const uint32_t* __restrict__ nb, uint64_t* key,
                                                              uint64_t* seen, uint32_t* color,
                                                              uint32_t* __restrict__ scene_colors, uint32_t* status) {
    __shared__ uint64_t wave_best[kDsaturBlock / 64];
    __shared__ uint8_t refresh[kDsaturBlock];
    __shared__ uint32_t used_colors;
    const uint32_t s = blockIdx.x, lo = scene_off[s], hi = scene_off[s + 1], n = hi - lo;
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    for (uint32_t v = lo + t; v < hi; v += kDsaturBlock) {
        const uint32_t deg = off[v + 1] - off[v];
        if (deg > 0xFFFFu) atomicOr(status, 2u);
        color[v] = kUncoloured;
        key[v] = dsatur_key(0, min(deg, 0xFFFFu), v - lo);
    }
    if (t == 0) used_colors = 0;
    refresh[t] = 1;
    __syncthreads();
    uint64_t mine = 0;
    for (uint32_t round = 0; round < n; round++) {
        if (refresh[t]) {
            mine = 0;
            for (uint32_t v = lo + t; v < hi; v += kDsaturBlock) {
                const uint64_t k = key[v];
                mine = k > mine ? k : mine;
            }
            refresh[t] = 0;
        }
        const uint64_t wm = wave_max_u64(mine);
        if (lane == 0) wave_best[wave] = wm;
        __syncthreads();
        if (wave == 0) {
            const uint64_t b = wave_max_u64(lane < kDsaturBlock / 64 ? wave_best[lane] : 0ull);
            const uint32_t v = lo + (0xFFFFFFFFu - (uint32_t)(b & 0xFFFFFFFFull));
            // lowest colour absent from v's neighbours
            const uint64_t* sv = seen + 4 * (size_t)v;
            uint32_t c = 256;
#pragma unroll
            for (int w = 3; w >= 0; w--) {
                const uint64_t free_bits = ~sv[w];
                if (free_bits) c = 64u * (uint32_t)w + (uint32_t)__builtin_ctzll(free_bits);
            }
            if (c >= 256) {
                if (lane == 0) atomicOr(status, 1u);
                c = 255;   // keeps the round well defined; the host reports the failure
            }
            if (lane == 0) {
                color[v] = c;
                key[v] = 0;
                if (c + 1 > used_colors) used_colors = c + 1;
                refresh[(v - lo) % kDsaturBlock] = 1;
            }
            // uncoloured neighbours gain the colour (the neighbour list has no
            // repeats, so every lane updates its own vertices)
            const uint32_t e0 = off[v], e1 = off[v + 1];
            const uint64_t bit = 1ull << (c & 63u);
            for (uint32_t e = e0 + lane; e < e1; e += 64) {
                const uint32_t w = nb[e];
                if (w == v || color[w] != kUncoloured) continue;
                uint64_t* sw = seen + 4 * (size_t)w + (c >> 6);
                if (*sw & bit) continue;
                *sw |= bit;
                key[w] += 1ull << 48;
                refresh[(w - lo) % kDsaturBlock] = 1;
            }
        }
        __syncthreads();
    }
    if (t == 0) scene_colors[s] = used_colors;
}

}  // namespace

void launch_dsatur(hipStream_t s, uint32_t n_scenes, const uint32_t* scene_off, const uint32_t* off,
                   const uint32_t* nb, uint64_t* key, uint64_t* seen, uint32_t* color, uint32_t* scene_colors,
                   uint32_t* status) {
    if (n_scenes == 0) return;
    hipLaunchKernelGGL(dsatur_kernel, dim3(n_scenes), dim3(kDsaturBlock), 0, s, scene_off, off, nb, key, seen, color,
                       scene_colors, status);
}

}  // namespace avbd
