// lbvh.hpp -- device buffers and launch wrapper of the LBVH broad phase.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace avbd {

// Traversal record of one internal node: both children with their boxes and
// leaf ranges, so a visit is one contiguous 128-byte read.
struct alignas(16) TravNode {
    double box[2][6];      // child boxes (min xyz, max xyz), FP64
    uint32_t child[2];     // child node ids (leaves are < n)
    uint32_t lo[2], hi[2]; // sorted-leaf range each child covers
    uint32_t pad[2];
};
static_assert(sizeof(TravNode) == 128, "TravNode is two 64-byte halves");

// Node numbering follows the reference (collision/LBVH.hpp:10-16): leaves
// [0, n) in sorted order, internal nodes [n, 2n-1), root n (n > 1).
struct LbvhDev {
    uint32_t n = 0;             // triangles
    uint32_t* tris = nullptr;   // 3 particle indices per triangle
    double* aabb = nullptr;     // 6 per triangle (min xyz, max xyz), FP64
    double* sbox = nullptr;     // scene box (6)
    uint64_t* keys = nullptr;   // Morton code per triangle
    uint64_t* keys_sorted = nullptr;
    uint32_t* sorted = nullptr;     // sorted_order: original triangle per sorted slot
    uint32_t* left = nullptr;       // 2n-1
    uint32_t* right = nullptr;
    uint32_t* parent = nullptr;
    uint32_t* range_hi = nullptr;   // last leaf covered by each node
    uint32_t* bucket_count = nullptr;    // 8,192 (zeroed by the build)
    uint32_t* bucket_offset = nullptr;   // 8,193
    uint32_t* slots = nullptr;           // n: primitives grouped by bucket
    double* table = nullptr;        // aligned range unions, 6 per entry (lbvh_table_entries)
    double* node_box = nullptr;     // 6 per node
    TravNode* trav = nullptr;       // n - 1 internal nodes (record g - n of node g)
    uint32_t* pairs = nullptr;      // candidate leaf pairs (a < b)
    uint32_t pair_cap = 0;
    uint32_t* counters = nullptr;   // [0] pair count, [1] status (bit0 stack overflow)
};

uint32_t lbvh_table_entries(uint32_t n);
// AABBs + Morton (FP64) -> sort -> Karras -> refit -> self traversal.
void launch_lbvh_build(hipStream_t s, const LbvhDev& b, const float4* x, const float4* v, double dt);
// The same in two parts: the speculative bounds read the particle state; the
// rest (codes, sort, hierarchy, boxes, traversal) reads only the bounds, so
// it may run concurrently with work that moves the particles.
void launch_lbvh_bounds(hipStream_t s, const LbvhDev& b, const float4* x, const float4* v, double dt);
void launch_lbvh_tree(hipStream_t s, const LbvhDev& b);

}  // namespace avbd
