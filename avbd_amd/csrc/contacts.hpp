// contacts.hpp -- dynamic contacts of the avbd context: collision primitives
// beyond triangles, the device narrow phase, and the host-side contact
// lifecycle, dynamic adjacency and incremental recolouring that the
// reference runs serially on its first worker every step
// (CollisionDetector::detectCollisionsAndRecolor_Parallel,
// collision/CollisionDetector.cpp:57-129).
#pragma once

#include <array>
#include <cstdint>
#include <vector>

#include "internal.hpp"
#include "lbvh.hpp"

namespace avbd {

// CollisionGeometryType values (collision/CollisionGeometryType.hpp)
constexpr uint32_t PRIM_TRIANGLE = 0u, PRIM_ROD_SEGMENT = 1u, PRIM_SPHERE = 2u;
// DetectedCollisionType values (collision/DetectedCollision.hpp:6-15); the
// rigid-rigid (sphere-sphere) pair is detected here, where the reference
// throws (collision/CollisionDetector.cpp:678-686), and so is the
// triangle-triangle vertex-face pair, whose narrow phase the reference leaves
// empty (collision/CollisionDetector.cpp:589-599)
constexpr uint32_t DET_KIND_VERTEX_FACE = 0u, DET_KIND_TRI_RIGID = 2u, DET_KIND_TRI_ROD = 3u,
                   DET_KIND_RIGID_RIGID = 6u;
// narrow-phase status word bits
constexpr uint32_t NARROW_UNSUPPORTED = 1u;

// Device primitive table (one entry per CollisionPrimitivePool slot).
struct DevPrims {
    uint32_t n = 0;
    uint32_t* type = nullptr;
    uint32_t* idx = nullptr;      // 3 per primitive: triangle vertices / rod nodes / sphere centre
    uint32_t* object = nullptr;
    double* radius = nullptr;
    // vertex-face contacts (self collision): on/off, the owner of every
    // particle (the lowest-numbered triangle holding it, ~0 for none), and
    // the widening of the triangle bounds (the contact margin) that lets the
    // broad phase report every face within the margin of a vertex
    uint32_t vertex_face = 0;
    uint32_t* owner = nullptr;
    double tri_pad = 0;
};

struct HostPrim {
    uint32_t type, idx[3], object;
    double radius;
};

void launch_prim_bounds(hipStream_t s, const DevPrims& p, const LbvhDev& b, const DevState& d, double dt);
// status[0]: NARROW_UNSUPPORTED; status[1]: detected count (may exceed det_cap)
void launch_narrow(hipStream_t s, const DevPrims& p, const LbvhDev& b, const DevState& d, double dt, double margin,
                   DetRec* det, uint32_t det_cap, uint32_t* status);

// Device state of the incremental recolouring (ColorList::incrementalRecoloring,
// common/ColorList.hpp:366-522) and of the work-list layout that follows it.
struct RecolourDev {
    uint32_t n = 0;
    uint32_t* color = nullptr;        // per particle
    uint32_t* last_dirty = nullptr;   // per particle: round that queued it, 0 = dirty, ~0 = clean
    uint32_t* candidate = nullptr;
    uint32_t* dirty = nullptr;        // [0, counters[0]): dirty particles left unresolved by the last call
    uint32_t* next = nullptr;
    uint32_t* touched = nullptr;      // particles queued during a call (capacity 7n)
    uint32_t* touched_conf = nullptr; // [0, counters[1]): the conflicted particles
    uint32_t* conflicted = nullptr;   // per particle flag
    uint32_t* counters = nullptr;     // [0] dirty, [1] conflicted, [2] colours, [3] status (1: out of colours),
                                      // [4] layout changed by the last recolouring
    const uint32_t* s_off = nullptr;  // static vertex adjacency (CSR)
    const uint32_t* s_nb = nullptr;
    const uint32_t* d_off = nullptr;  // dynamic vertex adjacency of this round (CSR)
    const uint32_t* d_nb = nullptr;
    const uint8_t* kind = nullptr;    // per particle: 0 no item, 1 positional, 2 oriented
    const VItem* item_of = nullptr;   // per particle: its work item
    const OItem* oitem_of = nullptr;
    uint32_t* offsets = nullptr;      // 3 x 257 per-colour offsets: items, oriented items, conflicted
};

// One workgroup: mark the particles of this round's new contacts dirty, then
// recolour and flag the conflicts; the new colour count is in counters[2].
void launch_recolour(hipStream_t s, const RecolourDev& r, const uint32_t* seeds, uint32_t n_seeds);
// One workgroup: lay the work list out again from the colours and conflicts
// (only when the last recolouring changed them, unless forced).
void launch_relayout(hipStream_t s, const RecolourDev& r, VItem* items, OItem* oitems, uint32_t* conflicts,
                     bool force);

// Host side of one detection round.
class ContactSystem {
public:
    double margin = 1e-3;
    uint32_t cap = 0;   // records per kind

    void init(uint32_t n_particles, uint32_t capacity, double collision_margin);

    // Merge this round's detected contacts (any order) with the previous
    // round's and rebuild the dynamic adjacency.  Fills ops / active /
    // dyn_off / dyn_ent / nb_off / nb and the particles of the added contacts
    // (seeds), which the device recolouring marks dirty.
    void process(std::vector<DetRec>& det, const std::vector<HostPrim>& prims, const std::vector<double>& mass,
                 double dt);

    std::vector<DynOp> ops;
    std::vector<uint32_t> active;            // record indices, pool 0 then pool 1, ascending slot
    std::vector<uint32_t> dyn_off, dyn_ent;  // per-particle CSR of (record << 3) | local
    std::vector<uint32_t> nb_off, nb;        // dynamic vertex adjacency (with duplicates)
    std::vector<uint32_t> seeds;             // particles of the contacts added this round, in add order

    struct Slots {   // TombstonePool (common/TombstonePool.hpp): LIFO stack of free slots
        uint32_t top = 0, count = 0;
        std::vector<uint32_t> stack;   // stack[i] for i < size, else i
        std::vector<uint8_t> active;
        std::vector<uint32_t> gen;
        std::vector<std::array<uint32_t, 5>> idx;
        uint32_t alloc(uint32_t cap);
        void free(uint32_t slot);
    };
    Slots pool[DYN_KINDS];   // triangle-rigid, triangle-rod, rigid-rigid, vertex-triangle

    struct Entry {   // one detected contact of the previous round, key order
        uint64_t key;
        uint32_t kind, slot, gen1, gen2;
    };
    std::vector<Entry> prev;

    // checkpoint: append the host state to out / restore it from [p, end)
    void save(std::vector<uint8_t>& out) const;
    void load(const uint8_t*& p, const uint8_t* end);

private:
    uint32_t n = 0;
};

}  // namespace avbd
