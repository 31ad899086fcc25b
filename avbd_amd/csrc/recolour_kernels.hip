// recolour_kernels.hip -- incremental recolouring and work-list layout of the
// dynamic-contact path on the device.
//
// ColorList::incrementalRecoloring + finalizeConflicts
// (common/ColorList.hpp:366-522), run after every detection round:
//   * the particles of newly added contacts are marked dirty (a particle that
//     is already dirty, including the ones a previous call left unresolved,
//     is not queued twice);
//   * at most six rounds: every dirty particle that shares its colour with a
//     lower-index neighbour (static or dynamic) takes the lowest colour none
//     of its neighbours has; the candidates are computed from the colours at
//     the start of the round and then applied, and every neighbour of a
//     particle that changed becomes dirty for the next round;
//   * the particles still dirty after the last round, and their neighbours
//     of the same colour, are flagged conflicted (their solves are buffered
//     and committed after the colour phase).
// Because every round reads only the colours from its start, the result does
// not depend on the order of the dirty list, so one workgroup can process a
// round with all of its threads and atomics for the list appends.
//
// The work list is then laid out again as the reference's solve walks it:
// per colour, positional items, oriented items and conflicted particles, each
// in ascending particle order.
#include "contacts.hpp"

namespace avbd {
namespace {

constexpr uint32_t kRecolourBlock = 1024;
constexpr uint32_t kNotDirty = 0xFFFFFFFFu;   // last_dirty sentinel
constexpr uint32_t kMaxRounds = 6;

template <class F>
__device__ __forceinline__ void each_neighbour(const RecolourDev& r, uint32_t p, F&& f) {
    for (uint32_t e = r.s_off[p]; e < r.s_off[p + 1]; e++) f(r.s_nb[e]);
    for (uint32_t e = r.d_off[p]; e < r.d_off[p + 1]; e++) f(r.d_nb[e]);
}

__global__ __launch_bounds__(kRecolourBlock) void recolour_kernel(RecolourDev r, const uint32_t* __restrict__ seeds,
                                                                   uint32_t n_seeds) {
    __shared__ uint32_t s_nd, s_add, s_next, s_touched, s_ntc, s_nc, s_status, s_changed;
    const uint32_t t = threadIdx.x;
    if (t == 0) {
        s_nd = r.counters[0];
        s_ntc = r.counters[1];
        s_nc = r.counters[2];
        s_status = 0;
        s_add = 0;
        s_touched = 0;
        s_changed = s_ntc != 0;   // clearing a conflict set changes the layout
    }
    __syncthreads();
    // the previous call's conflicts are cleared (they are found again below)
    for (uint32_t i = t; i < s_ntc; i += kRecolourBlock) r.conflicted[r.touched_conf[i]] = 0;
    __syncthreads();
    if (t == 0) s_ntc = 0;
    // ColorList::markParticleDirty for the particles of this round's new contacts
    const uint32_t nd0 = s_nd;
    for (uint32_t i = t; i < n_seeds; i += kRecolourBlock) {
        const uint32_t p = seeds[i];
        if (atomicExch(&r.last_dirty[p], 0u) != 0u) {
            r.dirty[nd0 + atomicAdd(&s_add, 1u)] = p;
            r.touched[atomicAdd(&s_touched, 1u)] = p;
        }
    }
    __syncthreads();
    if (t == 0) s_nd = nd0 + s_add;
    __syncthreads();
    uint32_t* cur = r.dirty;
    uint32_t* nxt = r.next;
    uint32_t nd = s_nd;
    for (uint32_t round = 0; nd > 0 && round < kMaxRounds; round++) {
        // candidates from the colours at the start of the round
        for (uint32_t i = t; i < nd; i += kRecolourBlock) {
            const uint32_t p = cur[i];
            const uint32_t cp = r.color[p];
            uint64_t bits[4] = {0, 0, 0, 0};
            bool must = false;
            each_neighbour(r, p, [&](uint32_t q) {
                const uint32_t c = r.color[q];
                bits[c >> 6] |= 1ull << (c & 63u);
                if (c == cp && q < p) must = true;
            });
            uint32_t cand = cp;
            if (must) {
                uint32_t c = 256;
                for (uint32_t w = 0; w < 4 && c == 256; w++)
                    if (~bits[w]) c = w * 64 + (uint32_t)__builtin_ctzll(~bits[w]);
                if (c == 256) atomicOr(&s_status, 1u);   // out of colours: the host fails the step
                else {
                    cand = c;
                    atomicMax(&s_nc, c + 1);
                }
            }
            r.candidate[p] = cand;
        }
        if (t == 0) s_next = 0;
        __syncthreads();
        // apply, and queue the neighbours of every particle that changed
        for (uint32_t i = t; i < nd; i += kRecolourBlock) {
            const uint32_t p = cur[i];
            const uint32_t cand = r.candidate[p];
            if (cand == r.color[p]) continue;
            r.color[p] = cand;
            s_changed = 1;
            each_neighbour(r, p, [&](uint32_t q) {
                if (atomicExch(&r.last_dirty[q], round + 1) != round + 1) {
                    nxt[atomicAdd(&s_next, 1u)] = q;
                    r.touched[atomicAdd(&s_touched, 1u)] = q;
                }
            });
        }
        __syncthreads();
        nd = s_next;
        uint32_t* tmp = cur;
        cur = nxt;
        nxt = tmp;
        __syncthreads();
    }
    // finalizeConflicts: the particles still dirty and their same-colour neighbours
    for (uint32_t i = t; i < nd; i += kRecolourBlock) {
        const uint32_t p = cur[i];
        const uint32_t cp = r.color[p];
        each_neighbour(r, p, [&](uint32_t q) {
            if (r.color[q] != cp) return;
            if (atomicExch(&r.conflicted[p], 1u) == 0u) r.touched_conf[atomicAdd(&s_ntc, 1u)] = p;
            if (atomicExch(&r.conflicted[q], 1u) == 0u) r.touched_conf[atomicAdd(&s_ntc, 1u)] = q;
        });
    }
    __syncthreads();
    // every particle touched this call is clean again, except the ones still dirty
    for (uint32_t i = t; i < s_touched; i += kRecolourBlock) r.last_dirty[r.touched[i]] = kNotDirty;
    __syncthreads();
    for (uint32_t i = t; i < nd; i += kRecolourBlock) {
        const uint32_t p = cur[i];
        r.last_dirty[p] = 0;
        if (cur != r.dirty) r.dirty[i] = p;   // the unresolved list stays in `dirty` for the next call
    }
    __syncthreads();
    if (t == 0) {
        r.counters[0] = nd;
        r.counters[1] = s_ntc;
        r.counters[4] = (s_changed || s_ntc != 0 || s_nc != r.counters[2]) ? 1u : 0u;
        r.counters[2] = s_nc;
        r.counters[3] |= s_status;
    }
}

// Work-list layout (the positional items, the oriented items and the
// conflicted particles, each grouped by colour in ascending particle order).
// Every wavefront owns one contiguous sixteenth of the particles: it counts
// its particles per (list, colour) in LDS, the counts are scanned over the
// wavefronts and the colours, and each wavefront then places its own
// particles from its cursors, 64 at a time, ranking the lanes of one colour
// with a ballot.
__global__ __launch_bounds__(kRecolourBlock) void relayout_kernel(RecolourDev r, VItem* __restrict__ items,
                                                                   OItem* __restrict__ oitems,
                                                                   uint32_t* __restrict__ conflicts, bool force) {
    constexpr uint32_t W = kRecolourBlock / 64;
    if (!force && !r.counters[4]) return;   // colours and conflicts as before: the layout stands
    __shared__ uint32_t cnt[3][W][256];
    __shared__ uint32_t tot[3][256];
    const uint32_t t = threadIdx.x, w = t >> 6, lane = t & 63u;
    const uint32_t nc = r.counters[2];
    const uint32_t seg = (r.n + W - 1) / W;
    const uint32_t v0 = min(r.n, w * seg), v1 = min(r.n, v0 + seg);
    for (uint32_t i = t; i < 3 * W * 256; i += kRecolourBlock) (&cnt[0][0][0])[i] = 0;
    __syncthreads();
    for (uint32_t v = v0 + lane; v < v1; v += 64) {
        const uint32_t c = r.color[v], k = r.kind[v];
        if (k == 1) atomicAdd(&cnt[0][w][c], 1u);
        else if (k == 2) atomicAdd(&cnt[1][w][c], 1u);
        if (r.conflicted[v]) atomicAdd(&cnt[2][w][c], 1u);
    }
    __syncthreads();
    // per (list, colour): exclusive scan over the wavefronts
    if (t < 3 * 256) {
        const uint32_t L = t >> 8, c = t & 255u;
        uint32_t run = 0;
        for (uint32_t q = 0; q < W; q++) {
            const uint32_t x = cnt[L][q][c];
            cnt[L][q][c] = run;
            run += x;
        }
        tot[L][c] = run;
    }
    __syncthreads();
    // per list: exclusive scan over the colours
    if (t < 3) {
        uint32_t run = 0;
        for (uint32_t c = 0; c < nc; c++) {
            const uint32_t k = tot[t][c];
            tot[t][c] = run;
            r.offsets[t * 257 + c] = run;
            run += k;
        }
        r.offsets[t * 257 + nc] = run;
    }
    __syncthreads();
    if (t < 3 * 256) {
        const uint32_t L = t >> 8, c = t & 255u;
        for (uint32_t q = 0; q < W; q++) cnt[L][q][c] += tot[L][c];
    }
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
    for (uint32_t base = v0; base < v1; base += 64) {
        const uint32_t v = base + lane;
        const bool in = v < v1;
        const uint32_t c = in ? r.color[v] : 0u, k = in ? r.kind[v] : 0u;
        const bool cf = in && r.conflicted[v] != 0;
        for (uint32_t L = 0; L < 3; L++) {
            bool pending = L == 0 ? k == 1 : (L == 1 ? k == 2 : cf);
            for (;;) {
                const uint64_t m = __ballot(pending);
                if (m == 0) break;
                const uint32_t c0 = __shfl(c, (int)(__ffsll((long long)m) - 1));
                const bool mine = pending && c == c0;
                const uint64_t mm = __ballot(mine);
                const uint32_t at = cnt[L][w][c0];
                if (mine) {
                    const uint32_t pos = at + (uint32_t)__popcll(mm & below);
                    if (L == 0) {
                        VItem it = r.item_of[v];
                        it.inc_count = (it.inc_count & ~VITEM_CONFLICT) | (cf ? VITEM_CONFLICT : 0u);
                        items[pos] = it;
                    } else if (L == 1) {
                        OItem it = r.oitem_of[v];
                        it.n_cos = (it.n_cos & ~VITEM_CONFLICT) | (cf ? VITEM_CONFLICT : 0u);
                        oitems[pos] = it;
                    } else {
                        conflicts[pos] = v;
                    }
                    pending = false;
                }
                if (lane == 0) cnt[L][w][c0] = at + (uint32_t)__popcll(mm);
            }
        }
    }
}

}  // namespace

void launch_recolour(hipStream_t s, const RecolourDev& r, const uint32_t* seeds, uint32_t n_seeds) {
    hipLaunchKernelGGL(recolour_kernel, dim3(1), dim3(kRecolourBlock), 0, s, r, seeds, n_seeds);
}

void launch_relayout(hipStream_t s, const RecolourDev& r, VItem* items, OItem* oitems, uint32_t* conflicts,
                     bool force) {
    hipLaunchKernelGGL(relayout_kernel, dim3(1), dim3(kRecolourBlock), 0, s, r, items, oitems, conflicts, force);
}

}  // namespace avbd
