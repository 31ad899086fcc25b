// lbvh.cpp -- FP64 oracle of the LBVH broad phase over triangle primitives
// (test infrastructure only; see oracle.hpp).
//
// Restates:
//   CollisionPrimitivePool::speculativeGlobalBounds (Triangle)  collision/CollisionPrimitivePool.hpp:160-174
//   LBVHBuilder::computeAABB_MortonCode                          collision/LBVHBuilder.cpp:98-127
//   expandBits21 / morton3D_64                                    collision/LBVHBuilder.hpp:65-86
//   Algorithm::radixSort (stable LSD by code)                     common/Algorithm.hpp:22-56
//   commonPrefixLen + LBVHBuilder::constructTree (Karras 2012)    collision/LBVHBuilder.hpp:33-41,
//                                                                 collision/LBVHBuilder.cpp:181-286
//   LBVHBuilder::assembleBVH (bottom-up refit)                    collision/LBVHBuilder.cpp:410-453
//   LBVHTraversal::traverseSelfIterative                          collision/LBVHTraversal.hpp:30-88
//   AABB::overlaps (closed boxes)                                 collision/AABB.hpp:40-44
#include <algorithm>
#include <bit>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <utility>
#include <vector>

namespace {

uint64_t expand_bits21(uint64_t v) {
    v &= 0x1FFFFFull;
    v = (v | (v << 32)) & 0x1F00000000FFFFull;
    v = (v | (v << 16)) & 0x1F0000FF0000FFull;
    v = (v | (v << 8)) & 0x100F00F00F00F00Full;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

uint64_t morton3d_64(const double p[3]) {
    uint64_t c[3];
    for (int k = 0; k < 3; k++) {
        double s = std::clamp(p[k] * 2097152.0, 0.0, 2097151.0);
        c[k] = expand_bits21(static_cast<uint64_t>(s));
    }
    return c[0] | (c[1] << 1) | (c[2] << 2);
}

int common_prefix_len(uint64_t ci, unsigned i, uint64_t cj, unsigned j) {
    if (ci == cj) return 64 + std::countl_zero(static_cast<uint64_t>(i) ^ static_cast<uint64_t>(j));
    return std::countl_zero(ci ^ cj);
}

bool overlaps(const double* a, const double* b) {
    for (int k = 0; k < 3; k++)
        if (!(a[k] <= b[3 + k] && a[3 + k] >= b[k])) return false;
    return true;
}

}  // namespace

extern "C" {

void oracle_lbvh_build_boxes(uint32_t n, const double* aabb, uint64_t* morton, uint32_t* sorted, uint32_t* left,
                             uint32_t* right, uint32_t* parent, double* node_box, uint32_t* root);

// Bit-interleave of one 21-bit value (for unit tests of the closed form).
uint64_t oracle_expand_bits21(uint64_t v) { return expand_bits21(v); }

// Build the LBVH over n triangles.  x, v: particle positions and velocities
// (3 doubles each); tris: 3 particle indices per triangle.
// Outputs (caller-allocated):
//   aabb[6n] (min xyz, max xyz), morton[n], sorted[n],
//   left/right/parent[2n-1] (leaves [0,n), internal [n,2n-1); 0xFFFFFFFF = none),
//   node_box[6(2n-1)], *root.
void oracle_lbvh_build(uint32_t n, const uint32_t* tris, const double* x, const double* v, double dt,
                       double* aabb, uint64_t* morton, uint32_t* sorted, uint32_t* left, uint32_t* right,
                       uint32_t* parent, double* node_box, uint32_t* root) {
    // speculative bounds: box over p and p + dt v for the 3 vertices
    for (uint32_t t = 0; t < n; t++) {
        double* b = aabb + 6 * t;
        for (int k = 0; k < 3; k++) {
            b[k] = std::numeric_limits<double>::max();
            b[3 + k] = std::numeric_limits<double>::lowest();
        }
        for (int c = 0; c < 3; c++) {
            const uint32_t p = tris[3 * t + c];
            for (int k = 0; k < 3; k++) {
                const double a = x[3 * p + k];
                const double e = x[3 * p + k] + dt * v[3 * p + k];
                b[k] = std::min(b[k], a);
                b[3 + k] = std::max(b[3 + k], a);
                b[k] = std::min(b[k], e);
                b[3 + k] = std::max(b[3 + k], e);
            }
        }
    }
    oracle_lbvh_build_boxes(n, aabb, morton, sorted, left, right, parent, node_box, root);
}

// LBVHBuilder::computeAABB_MortonCode .. assembleBVH over given boxes
// (collision/LBVHBuilder.cpp:98-127,181-286,410-453).
void oracle_lbvh_build_boxes(uint32_t n, const double* aabb, uint64_t* morton, uint32_t* sorted, uint32_t* left,
                             uint32_t* right, uint32_t* parent, double* node_box, uint32_t* root) {
    const uint32_t INV = 0xFFFFFFFFu;
    double smin[3] = {std::numeric_limits<double>::max(), std::numeric_limits<double>::max(),
                      std::numeric_limits<double>::max()};
    double smax[3] = {std::numeric_limits<double>::lowest(), std::numeric_limits<double>::lowest(),
                      std::numeric_limits<double>::lowest()};
    for (uint32_t t = 0; t < n; t++) {
        const double* b = aabb + 6 * t;
        for (int k = 0; k < 3; k++) {
            smin[k] = std::min(smin[k], b[k]);
            smax[k] = std::max(smax[k], b[3 + k]);
        }
    }
    double ext[3];
    for (int k = 0; k < 3; k++) ext[k] = std::max(smax[k] - smin[k], 1e-6);
    for (uint32_t t = 0; t < n; t++) {
        const double* b = aabb + 6 * t;
        double q[3];
        for (int k = 0; k < 3; k++) {
            const double c = (b[k] + b[3 + k]) * 0.5;
            q[k] = (c - smin[k]) / ext[k];
        }
        morton[t] = morton3d_64(q);
    }
    // stable sort by code == sort by (code, original index)
    std::iota(sorted, sorted + n, 0u);
    std::stable_sort(sorted, sorted + n, [&](uint32_t a, uint32_t b) { return morton[a] < morton[b]; });

    const uint32_t nn = 2 * n - 1;
    for (uint32_t i = 0; i < nn; i++) left[i] = right[i] = parent[i] = INV;
    const int N = (int)n;
    auto delta = [&](int i, int j) {
        if (j < 0 || j >= N) return -1;
        return common_prefix_len(morton[sorted[i]], sorted[i], morton[sorted[j]], sorted[j]);
    };
    for (int i = 0; i < N - 1; i++) {
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
        const int gamma = i + s * d + std::min(d, 0);
        const uint32_t g = (uint32_t)i + n;
        if (std::min(i, j) == gamma) {
            left[g] = (uint32_t)gamma;
            parent[gamma] = g;
        } else {
            left[g] = n + (uint32_t)gamma;
            parent[n + gamma] = g;
        }
        if (std::max(i, j) == gamma + 1) {
            right[g] = (uint32_t)gamma + 1;
            parent[gamma + 1] = g;
        } else {
            right[g] = n + (uint32_t)gamma + 1;
            parent[n + gamma + 1] = g;
        }
    }
    uint32_t node = 0;
    while (parent[node] != INV) node = parent[node];
    *root = node;

    // refit: leaves take their primitive's box; internal = union of children,
    // computed children-first (any topological order gives the same min/max)
    for (uint32_t l = 0; l < n; l++) std::memcpy(node_box + 6 * l, aabb + 6 * sorted[l], 6 * sizeof(double));
    std::vector<uint8_t> done(nn, 0);
    std::fill(done.begin(), done.begin() + n, 1);
    std::vector<uint32_t> order;
    order.reserve(n);
    // post-order via explicit stack from the root
    if (n > 1) {
        std::vector<std::pair<uint32_t, bool>> st{{*root, false}};
        while (!st.empty()) {
            auto [k, expanded] = st.back();
            st.pop_back();
            if (k < n) continue;
            if (expanded) {
                double* b = node_box + 6 * k;
                const double* L = node_box + 6 * left[k];
                const double* R = node_box + 6 * right[k];
                for (int c = 0; c < 3; c++) {
                    b[c] = std::min(L[c], R[c]);
                    b[3 + c] = std::max(L[3 + c], R[3 + c]);
                }
            } else {
                st.push_back({k, true});
                st.push_back({right[k], false});
                st.push_back({left[k], false});
            }
        }
    }
}

// Self traversal from (root, root): writes leaf pairs (a < b, sorted-order
// leaf indices) whose closed boxes overlap; returns the pair count (pairs are
// written only while count < cap).
uint32_t oracle_lbvh_self_pairs(uint32_t n, const uint32_t* left, const uint32_t* right, const double* node_box,
                                uint32_t root, uint32_t* pairs, uint32_t cap) {
    // subtree sizes for the descent heuristic
    const uint32_t nn = 2 * n - 1;
    std::vector<uint32_t> size(nn, 1);
    {
        std::vector<std::pair<uint32_t, bool>> st{{root, false}};
        while (!st.empty() && n > 1) {
            auto [k, ex] = st.back();
            st.pop_back();
            if (k < n) continue;
            if (ex) size[k] = size[left[k]] + size[right[k]];
            else {
                st.push_back({k, true});
                st.push_back({right[k], false});
                st.push_back({left[k], false});
            }
        }
    }
    uint32_t count = 0;
    std::vector<std::pair<uint32_t, uint32_t>> work{{root, root}};
    while (!work.empty()) {
        auto [a, b] = work.back();
        work.pop_back();
        if (!overlaps(node_box + 6 * a, node_box + 6 * b)) continue;
        const bool la = a < n, lb = b < n;
        if (la && lb) {
            if (a < b) {
                if (count < cap) {
                    pairs[2 * count] = a;
                    pairs[2 * count + 1] = b;
                }
                count++;
            }
            continue;
        }
        if (a == b) {
            work.push_back({left[a], left[a]});
            work.push_back({right[a], right[a]});
            work.push_back({left[a], right[a]});
            continue;
        }
        if (la || (!lb && size[b] > size[a])) {
            work.push_back({a, left[b]});
            work.push_back({a, right[b]});
        } else {
            work.push_back({left[a], b});
            work.push_back({right[a], b});
        }
    }
    return count;
}

}  // extern "C"

extern "C" {

// Karras hierarchy from given codes (stable order by code, index tie-break):
// the constructTree part of oracle_lbvh_build for unit tests.
void oracle_lbvh_from_codes(uint32_t n, const uint64_t* morton, uint32_t* sorted, uint32_t* left, uint32_t* right,
                            uint32_t* parent, uint32_t* root) {
    // reuse the builder with degenerate geometry by sorting here directly
    const uint32_t INV = 0xFFFFFFFFu;
    std::iota(sorted, sorted + n, 0u);
    std::stable_sort(sorted, sorted + n, [&](uint32_t a, uint32_t b) { return morton[a] < morton[b]; });
    const uint32_t nn = 2 * n - 1;
    for (uint32_t i = 0; i < nn; i++) left[i] = right[i] = parent[i] = INV;
    const int N = (int)n;
    auto delta = [&](int i, int j) {
        if (j < 0 || j >= N) return -1;
        return common_prefix_len(morton[sorted[i]], sorted[i], morton[sorted[j]], sorted[j]);
    };
    for (int i = 0; i < N - 1; i++) {
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
        const int gamma = i + s * d + std::min(d, 0);
        const uint32_t g = (uint32_t)i + n;
        const uint32_t lc = (std::min(i, j) == gamma) ? (uint32_t)gamma : n + (uint32_t)gamma;
        const uint32_t rc = (std::max(i, j) == gamma + 1) ? (uint32_t)gamma + 1 : n + (uint32_t)gamma + 1;
        left[g] = lc;
        right[g] = rc;
        parent[lc] = g;
        parent[rc] = g;
    }
    uint32_t node = 0;
    while (parent[node] != INV) node = parent[node];
    *root = node;
}

}  // extern "C"

// Broad phase used by the oracle's time step: build (serial) + self traversal
// split over worker ranges of leaves (per-leaf descent; the union over leaves
// is exactly the pair set of traverseSelfIterative).  Results are discarded by
// the step -- the triangle-triangle narrow phase of the reference is a no-op
// (collision/CollisionDetector.cpp:589-599) -- but the work is done, as the
// reference does every step.
#include "oracle.hpp"

namespace oracle {

void BroadPhase::build(const std::vector<uint32_t>& tris, const std::vector<V3>& x, const std::vector<V3>& v,
                       double dt) {
    n = (uint32_t)(tris.size() / 3);
    if (!n) return;
    xs.resize(3 * x.size());
    vs.resize(3 * x.size());
    for (size_t i = 0; i < x.size(); i++)
        for (int k = 0; k < 3; k++) {
            xs[3 * i + k] = x[i][k];
            vs[3 * i + k] = v[i][k];
        }
    const uint32_t nn = 2 * n - 1;
    aabb.resize(6 * n);
    morton.resize(n);
    sorted.resize(n);
    left.resize(nn);
    right.resize(nn);
    parent.resize(nn);
    node_box.resize(6 * nn);
    hi.assign(nn, 0);
    oracle_lbvh_build(n, tris.data(), xs.data(), vs.data(), dt, aabb.data(), morton.data(), sorted.data(),
                      left.data(), right.data(), parent.data(), node_box.data(), &root);
    // last leaf under each node (leaves are in sorted order)
    for (uint32_t l = 0; l < n; l++) hi[l] = l;
    std::vector<std::pair<uint32_t, bool>> st{{root, false}};
    while (!st.empty() && n > 1) {
        auto [k, ex] = st.back();
        st.pop_back();
        if (k < n) continue;
        if (ex) hi[k] = hi[right[k]];
        else {
            st.push_back({k, true});
            st.push_back({right[k], false});
            st.push_back({left[k], false});
        }
    }
}

void BroadPhase::build_boxes(const std::vector<double>& boxes) {
    n = (uint32_t)(boxes.size() / 6);
    if (!n) return;
    const uint32_t nn = 2 * n - 1;
    aabb = boxes;
    morton.resize(n);
    sorted.resize(n);
    left.resize(nn);
    right.resize(nn);
    parent.resize(nn);
    node_box.resize(6 * nn);
    hi.assign(nn, 0);
    oracle_lbvh_build_boxes(n, aabb.data(), morton.data(), sorted.data(), left.data(), right.data(), parent.data(),
                            node_box.data(), &root);
    for (uint32_t l = 0; l < n; l++) hi[l] = l;
    std::vector<std::pair<uint32_t, bool>> st{{root, false}};
    while (!st.empty() && n > 1) {
        auto [k, ex] = st.back();
        st.pop_back();
        if (k < n) continue;
        if (ex) hi[k] = hi[right[k]];
        else {
            st.push_back({k, true});
            st.push_back({right[k], false});
            st.push_back({left[k], false});
        }
    }
}

void BroadPhase::traverse_pairs(std::vector<std::pair<uint32_t, uint32_t>>& out) const {
    out.clear();
    std::vector<uint32_t> stack;
    for (uint32_t i = 0; i < n; i++) {
        const double* bi = node_box.data() + 6 * i;
        stack.assign(1, root);
        while (!stack.empty()) {
            const uint32_t k = stack.back();
            stack.pop_back();
            if (hi[k] <= i || !overlaps(bi, node_box.data() + 6 * k)) continue;
            if (k < n) out.emplace_back(i, k);
            else {
                stack.push_back(right[k]);
                stack.push_back(left[k]);
            }
        }
    }
}

uint64_t BroadPhase::traverse_leaves(uint32_t lo, uint32_t hi_leaf) const {
    uint64_t count = 0;
    std::vector<uint32_t> stack;
    for (uint32_t i = lo; i < hi_leaf; i++) {
        const double* bi = node_box.data() + 6 * i;
        stack.assign(1, root);
        while (!stack.empty()) {
            const uint32_t k = stack.back();
            stack.pop_back();
            if (hi[k] <= i || !overlaps(bi, node_box.data() + 6 * k)) continue;
            if (k < n) count++;
            else {
                stack.push_back(right[k]);
                stack.push_back(left[k]);
            }
        }
    }
    return count;
}

}  // namespace oracle
