// lbvh_kernels.hip -- device LBVH broad phase over triangle primitives.
//
// Bit-exact with the reference's CPU builder on identical inputs:
//   * bounds, centroids and Morton codes are computed in FP64 from the FP32
//     state, without contraction (this file is compiled with -ffp-contract=off):
//     CollisionPrimitivePool::speculativeGlobalBounds (collision/CollisionPrimitivePool.hpp:160-174),
//     LBVHBuilder::computeAABB_MortonCode (collision/LBVHBuilder.cpp:98-127),
//     morton3D_64 / expandBits21 (collision/LBVHBuilder.hpp:65-86);
//   * the sort orders by (code, original index) -- what a stable LSD radix sort
//     of the codes produces (common/Algorithm.hpp:22-56);
//   * the hierarchy is Karras (2012) with the reference's delta, including the
//     index tie-break of commonPrefixLen (collision/LBVHBuilder.hpp:33-41,
//     collision/LBVHBuilder.cpp:181-286);
//   * refit is bottom-up min/max with an arrival counter per internal node
//     (collision/LBVHBuilder.cpp:410-453) -- min/max are exact, so the order of
//     arrival does not change a bit;
//   * traversal reports every leaf pair (a < b) whose closed boxes overlap,
//     the set produced by LBVHTraversal::traverseSelfIterative
//     (collision/LBVHTraversal.hpp:30-88); one thread walks the tree per leaf.
#include "lbvh.hpp"

namespace avbd {

namespace {

constexpr uint32_t INV = 0xFFFFFFFFu;
constexpr int SORT_THREADS = 1024;
constexpr int SORT_CHUNK = 12;   // LDS sort handles n <= 12288

__device__ __forceinline__ uint64_t expand_bits21(uint64_t v) {
    v &= 0x1FFFFFull;
    v = (v | (v << 32)) & 0x1F00000000FFFFull;
    v = (v | (v << 16)) & 0x1F0000FF0000FFull;
    v = (v | (v << 8)) & 0x100F00F00F00F00Full;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

__device__ __forceinline__ uint64_t morton_axis(double p) {
    double s = p * 2097152.0;
    s = s < 0.0 ? 0.0 : (s > 2097151.0 ? 2097151.0 : s);   // std::clamp(s, 0, 2097151)
    return expand_bits21(static_cast<uint64_t>(s));
}

__device__ __forceinline__ int clz64(uint64_t x) { return x == 0 ? 64 : __builtin_clzll(x); }

__device__ __forceinline__ int common_prefix_len(uint64_t ci, uint32_t i, uint64_t cj, uint32_t j) {
    if (ci == cj) return 64 + clz64(static_cast<uint64_t>(i) ^ static_cast<uint64_t>(j));
    return clz64(ci ^ cj);
}

__global__ void tri_aabb_kernel(const uint32_t* __restrict__ tris, uint32_t n, const float4* __restrict__ x,
                                const float4* __restrict__ v, double dt, double* __restrict__ aabb) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double b[6] = {__DBL_MAX__, __DBL_MAX__, __DBL_MAX__, -__DBL_MAX__, -__DBL_MAX__, -__DBL_MAX__};
    for (int c = 0; c < 3; c++) {
        const uint32_t p = tris[3 * t + c];
        const float4 xp = x[p], vp = v[p];
        const double a[3] = {xp.x, xp.y, xp.z};
        const double w[3] = {vp.x, vp.y, vp.z};
        for (int k = 0; k < 3; k++) {
            const double e = a[k] + dt * w[k];
            b[k] = fmin(b[k], a[k]);
            b[3 + k] = fmax(b[3 + k], a[k]);
            b[k] = fmin(b[k], e);
            b[3 + k] = fmax(b[3 + k], e);
        }
    }
    for (int k = 0; k < 6; k++) aabb[6 * t + k] = b[k];
}

// One block: scene box = union of the primitive boxes; also resets the pair
// counter and status word of this build.
__global__ __launch_bounds__(1024) void scene_box_kernel(const double* __restrict__ aabb, uint32_t n,
                                                         double* __restrict__ sbox, uint32_t* __restrict__ counters) {
    __shared__ double red[6][1024];
    double b[6] = {__DBL_MAX__, __DBL_MAX__, __DBL_MAX__, -__DBL_MAX__, -__DBL_MAX__, -__DBL_MAX__};
    for (uint32_t t = threadIdx.x; t < n; t += blockDim.x)
        for (int k = 0; k < 3; k++) {
            b[k] = fmin(b[k], aabb[6 * t + k]);
            b[3 + k] = fmax(b[3 + k], aabb[6 * t + 3 + k]);
        }
    for (int k = 0; k < 6; k++) red[k][threadIdx.x] = b[k];
    __syncthreads();
    for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < 3; k++) {
                red[k][threadIdx.x] = fmin(red[k][threadIdx.x], red[k][threadIdx.x + s]);
                red[3 + k][threadIdx.x] = fmax(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) sbox[threadIdx.x] = red[threadIdx.x][0];
    if (threadIdx.x == 0) {
        counters[0] = 0;   // pair count
        counters[1] = 0;   // status bits
    }
}

__global__ void morton_kernel(const double* __restrict__ aabb, uint32_t n, const double* __restrict__ sbox,
                              uint64_t* __restrict__ keys) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    uint64_t code = 0;
    for (int k = 0; k < 3; k++) {
        const double ext = fmax(sbox[3 + k] - sbox[k], 1e-6);
        const double c = (aabb[6 * t + k] + aabb[6 * t + 3 + k]) * 0.5;
        code |= morton_axis((c - sbox[k]) / ext) << k;
    }
    keys[t] = code;
}

// Stable LSD radix sort of (key, original index), 4-bit digits, in LDS, one
// block of 1024 threads; thread t owns elements [12t, 12t+12) so the pass is
// stable.  Per-thread digit counts are packed 16 bits per digit into four
// 64-bit words and scanned across the block.
__device__ __forceinline__ uint32_t field16(uint64_t w0, uint64_t w1, uint64_t w2, uint64_t w3, uint32_t d) {
    const uint64_t w = (d < 4) ? w0 : (d < 8) ? w1 : (d < 12) ? w2 : w3;
    return static_cast<uint32_t>((w >> (16 * (d & 3))) & 0xFFFFull);
}

__device__ __forceinline__ uint64_t wave_incl_scan64(uint64_t v, int lane) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t o = __shfl_up(v, off);
        if (lane >= off) v += o;
    }
    return v;
}

__global__ __launch_bounds__(SORT_THREADS) void lds_radix_sort_kernel(const uint64_t* __restrict__ keys_in, uint32_t n,
                                                                      uint64_t* __restrict__ keys_out,
                                                                      uint32_t* __restrict__ vals_out) {
    __shared__ uint64_t skey[SORT_THREADS * SORT_CHUNK];
    __shared__ uint32_t sval[SORT_THREADS * SORT_CHUNK];
    __shared__ uint64_t wsum[16][4];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    uint64_t k_[SORT_CHUNK];
    uint32_t v_[SORT_CHUNK];
#pragma unroll
    for (int k = 0; k < SORT_CHUNK; k++) {
        const uint32_t e = t * SORT_CHUNK + k;
        k_[k] = e < n ? keys_in[e] : ~0ull;   // sentinels sort last, in index order
        v_[k] = e;
    }
    for (int shift = 0; shift < 64; shift += 4) {
        uint64_t c[4] = {0, 0, 0, 0};
#pragma unroll
        for (int k = 0; k < SORT_CHUNK; k++) {
            const uint32_t d = (k_[k] >> shift) & 15u;
            const uint64_t inc = 1ull << (16 * (d & 3));
            c[0] += (d >> 2) == 0 ? inc : 0;
            c[1] += (d >> 2) == 1 ? inc : 0;
            c[2] += (d >> 2) == 2 ? inc : 0;
            c[3] += (d >> 2) == 3 ? inc : 0;
        }
        uint64_t incl[4];
#pragma unroll
        for (int w = 0; w < 4; w++) incl[w] = wave_incl_scan64(c[w], lane);
        if (lane == 63)
            for (int w = 0; w < 4; w++) wsum[wave][w] = incl[w];
        __syncthreads();
        uint64_t pre[4] = {0, 0, 0, 0}, tot[4] = {0, 0, 0, 0};
        for (int q = 0; q < 16; q++)
            for (int w = 0; w < 4; w++) {
                if (q < wave) pre[w] += wsum[q][w];
                tot[w] += wsum[q][w];
            }
        uint64_t excl[4];
        for (int w = 0; w < 4; w++) excl[w] = pre[w] + incl[w] - c[w];
        uint32_t base[16];
        uint32_t run = 0;
        for (uint32_t d = 0; d < 16; d++) {
            base[d] = run;
            run += field16(tot[0], tot[1], tot[2], tot[3], d);
        }
        uint64_t seen[4] = {0, 0, 0, 0};
#pragma unroll
        for (int k = 0; k < SORT_CHUNK; k++) {
            const uint32_t d = (k_[k] >> shift) & 15u;
            uint32_t b = 0;
            for (uint32_t q = 0; q < 16; q++) b = (q == d) ? base[q] : b;
            const uint32_t pos = b + field16(excl[0], excl[1], excl[2], excl[3], d) +
                                 field16(seen[0], seen[1], seen[2], seen[3], d);
            const uint64_t inc = 1ull << (16 * (d & 3));
            seen[0] += (d >> 2) == 0 ? inc : 0;
            seen[1] += (d >> 2) == 1 ? inc : 0;
            seen[2] += (d >> 2) == 2 ? inc : 0;
            seen[3] += (d >> 2) == 3 ? inc : 0;
            skey[pos] = k_[k];
            sval[pos] = v_[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < SORT_CHUNK; k++) {
            k_[k] = skey[t * SORT_CHUNK + k];
            v_[k] = sval[t * SORT_CHUNK + k];
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < SORT_CHUNK; k++) {
        const uint32_t e = t * SORT_CHUNK + k;
        if (e < n) {
            keys_out[e] = k_[k];
            vals_out[e] = v_[k];
        }
    }
}

// Any n: rank of (key, index) by counting smaller keys (tiles staged in LDS).
__global__ __launch_bounds__(256) void rank_sort_kernel(const uint64_t* __restrict__ keys, uint32_t n,
                                                        uint64_t* __restrict__ keys_out,
                                                        uint32_t* __restrict__ vals_out) {
    __shared__ uint64_t tile[256];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t ki = i < n ? keys[i] : 0;
    uint32_t rank = 0;
    for (uint32_t base = 0; base < n; base += 256) {
        const uint32_t j = base + threadIdx.x;
        tile[threadIdx.x] = j < n ? keys[j] : ~0ull;
        __syncthreads();
        const uint32_t m = n - base < 256 ? n - base : 256;
        for (uint32_t q = 0; q < m; q++) {
            const uint64_t kj = tile[q];
            rank += (kj < ki || (kj == ki && base + q < i)) ? 1u : 0u;
        }
        __syncthreads();
    }
    if (i < n) {
        keys_out[rank] = ki;
        vals_out[rank] = i;
    }
}

__global__ void karras_kernel(const uint64_t* __restrict__ code, const uint32_t* __restrict__ sorted, uint32_t n,
                              uint32_t* __restrict__ left, uint32_t* __restrict__ right,
                              uint32_t* __restrict__ parent, uint32_t* __restrict__ range_hi,
                              uint32_t* __restrict__ visited) {
    const int N = (int)n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) parent[N > 1 ? N : 0] = INV;   // the root has no parent
    if (i < N) range_hi[i] = (uint32_t)i;      // a leaf covers itself
    if (i >= N - 1) return;
    visited[i] = 0;
    auto delta = [&](int a, int b) -> int {
        if (b < 0 || b >= N) return -1;
        return common_prefix_len(code[a], sorted[a], code[b], sorted[b]);
    };
    const int d = (delta(i, i + 1) - delta(i, i - 1) >= 0) ? 1 : -1;
    const int dmin = delta(i, i - d);
    int lmax = 128;
    while (delta(i, i + lmax * d) > dmin) lmax *= 4;
    int l = 0;
    for (int t = lmax; t >= 1; t /= 2)
        if (delta(i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(i, j);
    int s = 0;
    for (int t = (l + 1) / 2; t >= 1; t = (t == 1 ? 0 : (t + 1) / 2)) {
        const int ns = s + t;
        if (ns < l && delta(i, i + ns * d) > dnode) s = ns;
    }
    const int gamma = i + s * d + (d < 0 ? d : 0);
    const uint32_t g = (uint32_t)i + n;
    const uint32_t lc = (min(i, j) == gamma) ? (uint32_t)gamma : n + (uint32_t)gamma;
    const uint32_t rc = (max(i, j) == gamma + 1) ? (uint32_t)gamma + 1 : n + (uint32_t)gamma + 1;
    left[g] = lc;
    right[g] = rc;
    parent[lc] = g;
    parent[rc] = g;
    range_hi[g] = (uint32_t)max(i, j);
}

__device__ __forceinline__ double aload(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void refit_kernel(const double* __restrict__ aabb, const uint32_t* __restrict__ sorted, uint32_t n,
                             const uint32_t* __restrict__ left, const uint32_t* __restrict__ right,
                             const uint32_t* __restrict__ parent, uint32_t* __restrict__ visited,
                             double* node_box) {
    const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    const uint32_t prim = sorted[l];
    for (int k = 0; k < 6; k++) node_box[6 * l + k] = aabb[6 * prim + k];
    __threadfence();
    uint32_t node = parent[l];
    while (node != INV) {
        if (atomicAdd(&visited[node - n], 1u) == 0) return;   // the sibling finishes the union
        __threadfence();
        const uint32_t a = left[node], b = right[node];
        double box[6];
        for (int k = 0; k < 3; k++) {
            box[k] = fmin(aload(node_box + 6 * a + k), aload(node_box + 6 * b + k));
            box[3 + k] = fmax(aload(node_box + 6 * a + 3 + k), aload(node_box + 6 * b + 3 + k));
        }
        for (int k = 0; k < 6; k++) node_box[6 * node + k] = box[k];
        __threadfence();
        node = parent[node];
    }
}

__device__ __forceinline__ bool overlap(const double* a, const double* b) {
    return a[0] <= b[3] && a[3] >= b[0] && a[1] <= b[4] && a[4] >= b[1] && a[2] <= b[5] && a[5] >= b[2];
}

constexpr int TRAVERSE_STACK = 128;

__global__ void traverse_kernel(uint32_t n, uint32_t root, const uint32_t* __restrict__ left,
                                const uint32_t* __restrict__ right, const uint32_t* __restrict__ range_hi,
                                const double* __restrict__ node_box, uint32_t* __restrict__ pairs, uint32_t cap,
                                uint32_t* __restrict__ counters) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double bi[6];
    for (int k = 0; k < 6; k++) bi[k] = node_box[6 * i + k];
    uint32_t stack[TRAVERSE_STACK];
    int sp = 0;
    stack[sp++] = root;
    while (sp > 0) {
        const uint32_t node = stack[--sp];
        if (range_hi[node] <= i) continue;             // every leaf below is <= i
        if (!overlap(bi, node_box + 6 * node)) continue;
        if (node < n) {
            const uint32_t slot = atomicAdd(&counters[0], 1u);
            if (slot < cap) {
                pairs[2 * slot] = i;
                pairs[2 * slot + 1] = node;
            }
            continue;
        }
        if (sp + 2 > TRAVERSE_STACK) {
            atomicOr(&counters[1], 1u);   // stack overflow: reported, traversal truncated
            break;
        }
        stack[sp++] = right[node];
        stack[sp++] = left[node];
    }
}

inline uint32_t nblk(uint32_t n, uint32_t b) { return (n + b - 1) / b; }

}  // namespace

bool lbvh_lds_sort_fits(uint32_t n) { return n <= (uint32_t)(SORT_THREADS * SORT_CHUNK); }

void launch_lbvh_build(hipStream_t s, const LbvhDev& b, const float4* x, const float4* v, double dt) {
    const uint32_t n = b.n;
    if (n == 0) return;
    hipLaunchKernelGGL(tri_aabb_kernel, dim3(nblk(n, 256)), dim3(256), 0, s, b.tris, n, x, v, dt, b.aabb);
    hipLaunchKernelGGL(scene_box_kernel, dim3(1), dim3(1024), 0, s, b.aabb, n, b.sbox, b.counters);
    hipLaunchKernelGGL(morton_kernel, dim3(nblk(n, 256)), dim3(256), 0, s, b.aabb, n, b.sbox, b.keys);
    if (lbvh_lds_sort_fits(n))
        hipLaunchKernelGGL(lds_radix_sort_kernel, dim3(1), dim3(SORT_THREADS), 0, s, b.keys, n, b.keys_sorted,
                           b.sorted);
    else
        hipLaunchKernelGGL(rank_sort_kernel, dim3(nblk(n, 256)), dim3(256), 0, s, b.keys, n, b.keys_sorted,
                           b.sorted);
    hipLaunchKernelGGL(karras_kernel, dim3(nblk(n, 256)), dim3(256), 0, s, b.keys_sorted, b.sorted, n, b.left,
                       b.right, b.parent, b.range_hi, b.visited);
    hipLaunchKernelGGL(refit_kernel, dim3(nblk(n, 256)), dim3(256), 0, s, b.aabb, b.sorted, n, b.left, b.right,
                       b.parent, b.visited, b.node_box);
    hipLaunchKernelGGL(traverse_kernel, dim3(nblk(n, 128)), dim3(128), 0, s, n, n > 1 ? n : 0u, b.left, b.right,
                       b.range_hi, b.node_box, b.pairs, b.pair_cap, b.counters);
}

}  // namespace avbd
