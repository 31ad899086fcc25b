// lbvh_static.hip -- allocation-free variant of the triangle broad phase.
//
// avbd_detect_collisions_noalloc runs the same LBVH build and self traversal
// as avbd_detect_collisions (speculative bounds, FP64 Morton codes, sort,
// Karras hierarchy, range-union boxes, traversal), but over caller-owned
// device buffers and without a context: every intermediate array lives in one
// file-scope device array, `scratch`, so a call performs no device allocation
// and no host synchronisation (callers under stream capture, serving loops).
//
// Each GPU holds its own copy of `scratch` (kStaticBytes) from the first call
// on that device until the process ends.  Its device address is looked up
// once per device with hipGetSymbolAddress; host code only ever passes that
// address, never the symbol itself.  Calls on one device share the array, so
// calls that overlap in time on one GPU (two streams, two threads) race.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "../../include/avbd.h"
#include "internal.hpp"
#include "lbvh.hpp"

namespace avbd {
namespace {

constexpr uint32_t kStaticMaxTris = AVBD_NOALLOC_MAX_TRIANGLES;
constexpr uint32_t kBuckets = 1u << 13;   // bucket_count / bucket_offset sizes of the sort

constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }

constexpr uint32_t table_entries(uint32_t n) {
    uint32_t total = 0;
    for (uint32_t k = 0; (1u << k) < 2 * n; k++) total += (n + (1u << k) - 1) >> k;
    return total;
}

// Byte offsets of the LBVH arrays for n triangles inside the scratch array;
// the counters and the pair list are the caller's output buffers.
struct Carve {
    size_t aabb, sbox, keys, keys_sorted, sorted, left, right, parent, range_hi, bcount, boffset, slots, table,
        node_box, trav, total;
};

constexpr Carve carve(uint32_t n) {
    const size_t nn = 2 * (size_t)n - 1;
    Carve c{};
    size_t o = 0;
    c.aabb = o;        o += align256(6 * sizeof(double) * (size_t)n);
    c.sbox = o;        o += align256(6 * sizeof(double));
    c.keys = o;        o += align256(sizeof(uint64_t) * (size_t)n);
    c.keys_sorted = o; o += align256(sizeof(uint64_t) * (size_t)n);
    c.sorted = o;      o += align256(sizeof(uint32_t) * (size_t)n);
    c.left = o;        o += align256(sizeof(uint32_t) * nn);
    c.right = o;       o += align256(sizeof(uint32_t) * nn);
    c.parent = o;      o += align256(sizeof(uint32_t) * nn);
    c.range_hi = o;    o += align256(sizeof(uint32_t) * nn);
    c.bcount = o;      o += align256(sizeof(uint32_t) * kBuckets);
    c.boffset = o;     o += align256(sizeof(uint32_t) * (kBuckets + 1));
    c.slots = o;       o += align256(sizeof(uint32_t) * (size_t)n);
    c.table = o;       o += align256(6 * sizeof(double) * (size_t)table_entries(n));
    c.node_box = o;    o += align256(6 * sizeof(double) * nn);
    c.trav = o;        o += align256(sizeof(TravNode) * (size_t)(n > 1 ? n - 1 : 1));
    c.total = o;
    return c;
}

// Every carve of n <= kStaticMaxTris fits: each array grows with n.
constexpr size_t kStaticBytes = carve(kStaticMaxTris).total;
static_assert(kStaticBytes <= (size_t)8 << 20, "the static scratch array is meant to stay below 8 MiB");

__device__ __align__(256) unsigned char scratch[kStaticBytes];

// Device address of `scratch` per device ordinal (looked up on the first call
// made with that device current).
constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_scratch[kMaxDevices] = {};

hipError_t scratch_address(int dev, unsigned char** out) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

// counters[0] = pairs found, counters[1] = status bits (an empty input never
// reaches the build kernels, which clear them otherwise).
__global__ void clear_counters_kernel(uint32_t* counters) {
    if (threadIdx.x < 2) counters[threadIdx.x] = 0u;
}

}  // namespace
}  // namespace avbd

using namespace avbd;

extern "C" int avbd_detect_collisions_noalloc(const float* x4, const float* v4, const uint32_t* tris,
                                              uint32_t n_tris, double dt, uint32_t* pairs, uint32_t pair_cap,
                                              uint32_t* counters, void* stream) {
    if (n_tris > kStaticMaxTris) {
        char msg[160];
        std::snprintf(msg, sizeof msg, "avbd_detect_collisions_noalloc: %u triangles exceed the static maximum of %u",
                      n_tris, kStaticMaxTris);
        set_context_free_error(msg);
        return AVBD_E_CAPACITY;
    }
    if (!counters || (n_tris && (!x4 || !v4 || !tris)) || (pair_cap && !pairs)) {
        set_context_free_error("avbd_detect_collisions_noalloc: null device pointer");
        return AVBD_E_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess && (dev < 0 || dev >= kMaxDevices)) e = hipErrorInvalidDevice;
    unsigned char* base = nullptr;
    if (e == hipSuccess) e = scratch_address(dev, &base);
    if (e != hipSuccess) {
        set_context_free_error(std::string("avbd_detect_collisions_noalloc: ") + hipGetErrorString(e));
        return AVBD_E_HIP;
    }
    if (n_tris == 0) {
        hipLaunchKernelGGL(clear_counters_kernel, dim3(1), dim3(64), 0, s, counters);
    } else {
        const Carve c = carve(n_tris);
        LbvhDev b;
        b.n = n_tris;
        b.tris = const_cast<uint32_t*>(tris);
        b.aabb = reinterpret_cast<double*>(base + c.aabb);
        b.sbox = reinterpret_cast<double*>(base + c.sbox);
        b.keys = reinterpret_cast<uint64_t*>(base + c.keys);
        b.keys_sorted = reinterpret_cast<uint64_t*>(base + c.keys_sorted);
        b.sorted = reinterpret_cast<uint32_t*>(base + c.sorted);
        b.left = reinterpret_cast<uint32_t*>(base + c.left);
        b.right = reinterpret_cast<uint32_t*>(base + c.right);
        b.parent = reinterpret_cast<uint32_t*>(base + c.parent);
        b.range_hi = reinterpret_cast<uint32_t*>(base + c.range_hi);
        b.bucket_count = reinterpret_cast<uint32_t*>(base + c.bcount);
        b.bucket_offset = reinterpret_cast<uint32_t*>(base + c.boffset);
        b.slots = reinterpret_cast<uint32_t*>(base + c.slots);
        b.table = reinterpret_cast<double*>(base + c.table);
        b.node_box = reinterpret_cast<double*>(base + c.node_box);
        b.trav = reinterpret_cast<TravNode*>(base + c.trav);
        b.pairs = pairs;
        b.pair_cap = pair_cap;
        b.counters = counters;
        launch_lbvh_build(s, b, reinterpret_cast<const float4*>(x4), reinterpret_cast<const float4*>(v4), dt);
    }
    e = hipGetLastError();
    if (e != hipSuccess) {
        set_context_free_error(std::string("avbd_detect_collisions_noalloc: kernel launch: ") + hipGetErrorString(e));
        return AVBD_E_HIP;
    }
    return AVBD_OK;
}
