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
//   * node boxes are the min/max over the leaves a node covers -- what the
//     reference's bottom-up refit computes (collision/LBVHBuilder.cpp:410-453);
//     min/max are exact, so any evaluation order gives the same bits;
//   * traversal reports every leaf pair (a < b) whose closed boxes overlap,
//     the set produced by LBVHTraversal::traverseSelfIterative
//     (collision/LBVHTraversal.hpp:30-88); one thread walks the tree per leaf,
//     reading a packed record of both children per visit.
#include "lbvh.hpp"

namespace avbd {

namespace {

constexpr uint32_t INV = 0xFFFFFFFFu;

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

constexpr uint32_t BUCKET_SHIFT = 50;   // codes have 63 bits: the top 13 pick the bucket
constexpr uint32_t N_BUCKETS = 1u << 13;
constexpr uint32_t BUCKETS_PER_THREAD = N_BUCKETS / 1024;

// One block: scene box = union of the primitive boxes; also resets the pair
// counter, the status word and the bucket counts of this build.
__global__ __launch_bounds__(1024) void scene_box_kernel(const double* __restrict__ aabb, uint32_t n,
                                                         double* __restrict__ sbox, uint32_t* __restrict__ counters,
                                                         uint32_t* __restrict__ bucket_count) {
    __shared__ double red[6][1024];
    for (uint32_t k = threadIdx.x; k < N_BUCKETS; k += blockDim.x) bucket_count[k] = 0;
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
                              uint64_t* __restrict__ keys, uint32_t* __restrict__ bucket_count) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    uint64_t code = 0;
    for (int k = 0; k < 3; k++) {
        const double ext = fmax(sbox[3 + k] - sbox[k], 1e-6);
        const double c = (aabb[6 * t + k] + aabb[6 * t + 3 + k]) * 0.5;
        code |= morton_axis((c - sbox[k]) / ext) << k;
    }
    keys[t] = code;
    atomicAdd(&bucket_count[code >> BUCKET_SHIFT], 1u);
}

// Sort by (code, original index) -- the order a stable LSD radix sort of the
// codes produces -- in short fully parallel passes: morton_kernel counts codes
// per bucket of their top 13 bits; one block scans the 8,192 counts; every
// primitive takes a slot in its bucket; every primitive counts the keys of
// its bucket that precede it and lands at bucket offset + count.
__global__ __launch_bounds__(1024) void bucket_scan_kernel(uint32_t* __restrict__ count, uint32_t* __restrict__ offset,
                                                           uint32_t n) {
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x;   // BUCKETS_PER_THREAD consecutive buckets per thread
    uint32_t local = 0;
    for (uint32_t k = 0; k < BUCKETS_PER_THREAD; k++) local += count[BUCKETS_PER_THREAD * t + k];
    part[t] = local;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint32_t v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - local;
    for (uint32_t k = 0; k < BUCKETS_PER_THREAD; k++) {
        const uint32_t b = BUCKETS_PER_THREAD * t + k;
        const uint32_t c = count[b];
        offset[b] = run;
        run += c;
        count[b] = 0;   // reused as the fill cursor by bucket_place_kernel
    }
    if (t == 0) offset[N_BUCKETS] = n;
}

__global__ void bucket_place_kernel(const uint64_t* __restrict__ keys, uint32_t n, const uint32_t* __restrict__ offset,
                                    uint32_t* __restrict__ cursor, uint32_t* __restrict__ slots) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t b = (uint32_t)(keys[i] >> BUCKET_SHIFT);
    slots[offset[b] + atomicAdd(&cursor[b], 1u)] = i;
}

__global__ void bucket_rank_kernel(const uint64_t* __restrict__ keys, uint32_t n, const uint32_t* __restrict__ offset,
                                   const uint32_t* __restrict__ slots, uint64_t* __restrict__ keys_out,
                                   uint32_t* __restrict__ vals_out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t ki = keys[i];
    const uint32_t b = (uint32_t)(ki >> BUCKET_SHIFT);
    const uint32_t lo = offset[b], hi = offset[b + 1];
    uint32_t rank = lo;
    for (uint32_t s = lo; s < hi; s++) {
        const uint32_t j = slots[s];
        const uint64_t kj = keys[j];
        rank += (kj < ki || (kj == ki && j < i)) ? 1u : 0u;
    }
    keys_out[rank] = ki;
    vals_out[rank] = i;
}

// Aligned range-union table over the sorted leaves: entry (k, j) of level k is
// the union of leaves [j 2^k, (j+1) 2^k).  An internal node covers a
// contiguous leaf range, so its box is the union of O(2 log n) entries.
__device__ __forceinline__ uint32_t level_offset(uint32_t n, uint32_t k) {
    uint32_t off = 0;
    for (uint32_t q = 0; q < k; q++) off += (n + (1u << q) - 1) >> q;
    return off;
}

__global__ __launch_bounds__(256) void leaf_table_kernel(const double* __restrict__ aabb,
                                                         const uint32_t* __restrict__ sorted, uint32_t n,
                                                         double* __restrict__ node_box, double* __restrict__ table,
                                                         uint32_t levels_local) {
    __shared__ double sb[6][256];
    const uint32_t l = blockIdx.x * 256 + threadIdx.x;
    double b[6];
    if (l < n) {
        const uint32_t prim = sorted[l];
        for (int k = 0; k < 6; k++) b[k] = aabb[6 * prim + k];
        for (int k = 0; k < 6; k++) {
            node_box[6 * l + k] = b[k];
            table[6 * l + k] = b[k];
        }
    } else {
        for (int k = 0; k < 3; k++) {
            b[k] = __DBL_MAX__;
            b[3 + k] = -__DBL_MAX__;
        }
    }
    for (int k = 0; k < 6; k++) sb[k][threadIdx.x] = b[k];
    __syncthreads();
    for (uint32_t lev = 1; lev <= levels_local; lev++) {
        const uint32_t w = 1u << lev;
        if ((threadIdx.x & (w - 1)) == 0) {
            const uint32_t o = threadIdx.x + (w >> 1);
            for (int k = 0; k < 3; k++) {
                sb[k][threadIdx.x] = fmin(sb[k][threadIdx.x], sb[k][o]);
                sb[3 + k][threadIdx.x] = fmax(sb[3 + k][threadIdx.x], sb[3 + k][o]);
            }
            const uint32_t j = l >> lev;
            if (j < ((n + w - 1) >> lev)) {
                const uint32_t base = level_offset(n, lev) + j;
                for (int k = 0; k < 6; k++) table[6 * base + k] = sb[k][threadIdx.x];
            }
        }
        __syncthreads();
    }
}

// Levels above one block: a single block folds pairs level by level.
__global__ __launch_bounds__(1024) void upper_table_kernel(uint32_t n, double* table, uint32_t first_level,
                                                           uint32_t last_level) {
    for (uint32_t lev = first_level; lev <= last_level; lev++) {
        const uint32_t cnt = (n + (1u << lev) - 1) >> lev;
        const uint32_t prev_cnt = (n + (1u << (lev - 1)) - 1) >> (lev - 1);
        const uint32_t po = level_offset(n, lev - 1), o = level_offset(n, lev);
        for (uint32_t j = threadIdx.x; j < cnt; j += blockDim.x) {
            const double* a = table + 6 * (po + 2 * j);
            double b[6];
            for (int k = 0; k < 6; k++) b[k] = a[k];
            if (2 * j + 1 < prev_cnt) {
                const double* c = a + 6;
                for (int k = 0; k < 3; k++) {
                    b[k] = fmin(b[k], c[k]);
                    b[3 + k] = fmax(b[3 + k], c[3 + k]);
                }
            }
            for (int k = 0; k < 6; k++) table[6 * (o + j) + k] = b[k];
        }
        __syncthreads();
    }
}

__device__ __forceinline__ void range_union(const double* __restrict__ table, uint32_t n, uint32_t lo, uint32_t hi,
                                            double* box) {
    for (int k = 0; k < 3; k++) {
        box[k] = __DBL_MAX__;
        box[3 + k] = -__DBL_MAX__;
    }
    uint32_t l = lo;
    const uint32_t r = hi + 1;
    while (l < r) {
        uint32_t k = 31u - (uint32_t)__builtin_clz(r - l);   // floor(log2(r - l))
        if (l) k = min(k, (uint32_t)__builtin_ctz(l));       // alignment of l
        const double* e = table + 6 * (level_offset(n, k) + (l >> k));
        for (int c = 0; c < 3; c++) {
            box[c] = fmin(box[c], e[c]);
            box[3 + c] = fmax(box[3 + c], e[3 + c]);
        }
        l += 1u << k;
    }
}

__global__ void karras_kernel(const uint64_t* __restrict__ code, const uint32_t* __restrict__ sorted, uint32_t n,
                              uint32_t* __restrict__ left, uint32_t* __restrict__ right,
                              uint32_t* __restrict__ parent, uint32_t* __restrict__ range_hi,
                              const double* __restrict__ table, double* __restrict__ node_box,
                              TravNode* __restrict__ trav) {
    const int N = (int)n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) parent[N > 1 ? N : 0] = INV;   // the root has no parent
    if (i < N) range_hi[i] = (uint32_t)i;      // a leaf covers itself
    if (i >= N - 1) return;
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
    // children cover [lo, gamma] and [gamma + 1, hi]; the node's box is the
    // union of theirs (min/max are exact: the same bits as one range union)
    TravNode t;
    t.child[0] = lc;
    t.child[1] = rc;
    t.lo[0] = (uint32_t)min(i, j);
    t.hi[0] = (uint32_t)gamma;
    t.lo[1] = (uint32_t)gamma + 1;
    t.hi[1] = (uint32_t)max(i, j);
    t.pad[0] = t.pad[1] = 0;
    range_union(table, n, t.lo[0], t.hi[0], t.box[0]);
    range_union(table, n, t.lo[1], t.hi[1], t.box[1]);
    trav[i] = t;
    for (int k = 0; k < 3; k++) {
        node_box[6 * g + k] = fmin(t.box[0][k], t.box[1][k]);
        node_box[6 * g + 3 + k] = fmax(t.box[0][3 + k], t.box[1][3 + k]);
    }
}

// Closed-box overlap; every operand is already in registers, so the six
// tests are evaluated together (no short-circuit chain of dependent reads).
__device__ __forceinline__ bool overlap(const double* a, const double* b) {
    return (a[0] <= b[3]) & (a[3] >= b[0]) & (a[1] <= b[4]) & (a[4] >= b[1]) & (a[2] <= b[5]) & (a[5] >= b[2]);
}

constexpr int TRAVERSE_STACK = 64;
constexpr int TRAVERSE_BLOCK = 64;
constexpr int TRAVERSE_LOCAL = 32;   // pairs buffered per thread before the block flush

// One thread per leaf i; descends from the root.  A visit reads the node's
// 128-byte TravNode (both children's boxes and ranges) in one go and tests
// both children; the deferred child goes to a per-thread stack in LDS.
// Children whose last leaf is <= i are skipped, leaf children that overlap
// are recorded as pairs (i, j), j > i.  Pairs are buffered in LDS and written
// with one global atomic reservation per block (a pair counter hit by every
// pair serialises at L2).
__global__ __launch_bounds__(TRAVERSE_BLOCK) void traverse_kernel(uint32_t n, const TravNode* __restrict__ trav,
                                                                  const double* __restrict__ node_box,
                                                                  uint32_t* __restrict__ pairs, uint32_t cap,
                                                                  uint32_t* __restrict__ counters) {
    __shared__ uint32_t stack_mem[TRAVERSE_STACK * TRAVERSE_BLOCK];
    __shared__ uint32_t out_j[TRAVERSE_LOCAL * TRAVERSE_BLOCK];
    __shared__ uint32_t block_base;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t cnt = 0;
    auto emit = [&](uint32_t j) {
        if (cnt < TRAVERSE_LOCAL) {
            out_j[cnt * TRAVERSE_BLOCK + threadIdx.x] = j;
        } else {   // rare overflow of the local buffer: reserve directly
            const uint32_t slot = atomicAdd(&counters[0], 1u);
            if (slot < cap) {
                pairs[2 * slot] = i;
                pairs[2 * slot + 1] = j;
            }
        }
        cnt++;
    };
    if (i + 1 < n) {   // the last leaf has no partner j > i; n > 1 here, so the root n is internal
        double bi[6];
        for (int k = 0; k < 6; k++) bi[k] = node_box[6 * i + k];
        int sp = 0;
        uint32_t node = n;
        while (true) {
            const TravNode r = trav[node - n];
            bool h0 = (r.hi[0] > i) & overlap(bi, r.box[0]);
            bool h1 = (r.hi[1] > i) & overlap(bi, r.box[1]);
            if (h0 && r.child[0] < n) {
                emit(r.child[0]);
                h0 = false;
            }
            if (h1 && r.child[1] < n) {
                emit(r.child[1]);
                h1 = false;
            }
            if (h0 && h1) {
                if (sp >= TRAVERSE_STACK) {
                    atomicOr(&counters[1], 1u);   // stack overflow: reported, traversal truncated
                    break;
                }
                stack_mem[(sp++) * TRAVERSE_BLOCK + threadIdx.x] = r.child[1];
                node = r.child[0];
            } else if (h0) {
                node = r.child[0];
            } else if (h1) {
                node = r.child[1];
            } else if (sp > 0) {
                node = stack_mem[(--sp) * TRAVERSE_BLOCK + threadIdx.x];
            } else {
                break;
            }
        }
    }
    // block flush: exclusive prefix of the buffered counts (one wave per block)
    const uint32_t mine = cnt < TRAVERSE_LOCAL ? cnt : TRAVERSE_LOCAL;
    uint32_t incl = mine;
    for (int off = 1; off < TRAVERSE_BLOCK; off <<= 1) {
        const uint32_t o = __shfl_up(incl, off);
        if ((int)threadIdx.x >= off) incl += o;
    }
    if (threadIdx.x == TRAVERSE_BLOCK - 1) block_base = incl ? atomicAdd(&counters[0], incl) : 0u;
    __syncthreads();
    const uint32_t base = block_base + incl - mine;
    for (uint32_t k = 0; k < mine; k++) {
        const uint32_t slot = base + k;
        if (slot < cap) {
            pairs[2 * slot] = i;
            pairs[2 * slot + 1] = out_j[k * TRAVERSE_BLOCK + threadIdx.x];
        }
    }
}

inline uint32_t nblk(uint32_t n, uint32_t b) { return (n + b - 1) / b; }

}  // namespace

uint32_t lbvh_table_entries(uint32_t n) {
    uint32_t total = 0;
    for (uint32_t k = 0; (1u << k) < 2 * n; k++) total += (n + (1u << k) - 1) >> k;
    return total;
}

void launch_lbvh_bounds(hipStream_t s, const LbvhDev& b, const float4* x, const float4* v, double dt) {
    if (b.n == 0) return;
    hipLaunchKernelGGL(tri_aabb_kernel, dim3(nblk(b.n, 256)), dim3(256), 0, s, b.tris, b.n, x, v, dt, b.aabb);
}

void launch_lbvh_tree(hipStream_t s, const LbvhDev& b) {
    const uint32_t n = b.n;
    if (n == 0) return;
    uint32_t top = 0;   // highest table level: 2^top >= n
    while ((1u << top) < n) top++;
    hipLaunchKernelGGL(scene_box_kernel, dim3(1), dim3(1024), 0, s, b.aabb, n, b.sbox, b.counters, b.bucket_count);
    hipLaunchKernelGGL(morton_kernel, dim3(nblk(n, 256)), dim3(256), 0, s, b.aabb, n, b.sbox, b.keys,
                       b.bucket_count);
    hipLaunchKernelGGL(bucket_scan_kernel, dim3(1), dim3(1024), 0, s, b.bucket_count, b.bucket_offset, n);
    hipLaunchKernelGGL(bucket_place_kernel, dim3(nblk(n, 256)), dim3(256), 0, s, b.keys, n, b.bucket_offset,
                       b.bucket_count, b.slots);
    hipLaunchKernelGGL(bucket_rank_kernel, dim3(nblk(n, 256)), dim3(256), 0, s, b.keys, n, b.bucket_offset, b.slots,
                       b.keys_sorted, b.sorted);
    hipLaunchKernelGGL(leaf_table_kernel, dim3(nblk(n, 256)), dim3(256), 0, s, b.aabb, b.sorted, n, b.node_box,
                       b.table, top < 8 ? top : 8u);
    if (top > 8) hipLaunchKernelGGL(upper_table_kernel, dim3(1), dim3(1024), 0, s, n, b.table, 9u, top);
    hipLaunchKernelGGL(karras_kernel, dim3(nblk(n, 256)), dim3(256), 0, s, b.keys_sorted, b.sorted, n, b.left,
                       b.right, b.parent, b.range_hi, b.table, b.node_box, b.trav);
    hipLaunchKernelGGL(traverse_kernel, dim3(nblk(n, TRAVERSE_BLOCK)), dim3(TRAVERSE_BLOCK), 0, s, n, b.trav,
                       b.node_box, b.pairs, b.pair_cap, b.counters);
}

void launch_lbvh_build(hipStream_t s, const LbvhDev& b, const float4* x, const float4* v, double dt) {
    launch_lbvh_bounds(s, b, x, v, dt);
    launch_lbvh_tree(s, b);
}

}  // namespace avbd
