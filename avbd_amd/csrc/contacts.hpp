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
// DetectedCollisionType values (collision/DetectedCollision.hpp:6-15)
constexpr uint32_t DET_KIND_TRI_RIGID = 2u, DET_KIND_TRI_ROD = 3u;
// narrow-phase status word bits
constexpr uint32_t NARROW_UNSUPPORTED = 1u;

// Device primitive table (one entry per CollisionPrimitivePool slot).
struct DevPrims {
    uint32_t n = 0;
    uint32_t* type = nullptr;
    uint32_t* idx = nullptr;      // 3 per primitive: triangle vertices / rod nodes / sphere centre
    uint32_t* object = nullptr;
    double* radius = nullptr;
};

struct HostPrim {
    uint32_t type, idx[3], object;
    double radius;
};

void launch_prim_bounds(hipStream_t s, const DevPrims& p, const LbvhDev& b, const DevState& d, double dt);
// status[0]: NARROW_UNSUPPORTED; status[1]: detected count (may exceed det_cap)
void launch_narrow(hipStream_t s, const DevPrims& p, const LbvhDev& b, const DevState& d, double dt, double margin,
                   DetRec* det, uint32_t det_cap, uint32_t* status);

// Host side of one detection round.
class ContactSystem {
public:
    double margin = 1e-3;
    uint32_t cap = 0;   // records per kind

    void init(uint32_t n_particles, uint32_t capacity, double collision_margin);

    // Merge this round's detected contacts (any order) with the previous
    // round's, then rebuild the dynamic adjacency and recolour incrementally.
    // Fills ops / active / dyn_off / dyn_ent; returns true when the colouring
    // or the conflict set changed (plan relaid out).
    bool process(std::vector<DetRec>& det, const std::vector<HostPrim>& prims, const std::vector<double>& mass,
                 double dt, ColorPlan& plan);

    std::vector<DynOp> ops;
    std::vector<uint32_t> active;            // record indices, pool 0 then pool 1, ascending slot
    std::vector<uint32_t> dyn_off, dyn_ent;  // per-particle CSR of (record << 3) | local
    std::vector<uint8_t> conflicted;

    struct Slots {   // TombstonePool (common/TombstonePool.hpp): LIFO stack of free slots
        uint32_t top = 0, count = 0;
        std::vector<uint32_t> stack;   // stack[i] for i < size, else i
        std::vector<uint8_t> active;
        std::vector<uint32_t> gen;
        std::vector<std::array<uint32_t, 5>> idx;
        uint32_t alloc(uint32_t cap);
        void free(uint32_t slot);
    };
    Slots pool[2];

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
    std::vector<uint32_t> last_dirty_round, candidate, dirty, next_dirty, touched_this_frame, touched_conflicted;
    std::vector<uint32_t> nb_off, nb;   // dynamic vertex adjacency (with duplicates)
    void mark_dirty(uint32_t p);
    bool recolour(ColorPlan& plan);
};

}  // namespace avbd
