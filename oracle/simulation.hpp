// simulation.hpp -- the oracle's restatement of one reference Simulation
// (SimulationContext + VBDSolver), positional and oriented particles.
// Test infrastructure only; see oracle.hpp.
#pragma once

#include "contact.hpp"
#include "oracle.hpp"
#include "rigid.hpp"

namespace oracle {

class Simulation {
public:
    Params params;
    // ParticlePool, common/ParticlePool.hpp:78-90
    std::vector<V3> x, x_buf, y, x_prev, x_last, x_lastlast, v, v_prev;
    // per-particle external acceleration added to gravity (empty = none; an
    // extension of the reference, whose only external acceleration is gravity)
    std::vector<V3> f_ext;
    std::vector<Real> mass;
    std::vector<uint8_t> fixed;
    RotState rs;   // RotationPool through rotation_idx

    // EnergyRegistry (energy/EnergyRegistry.hpp:17-27)
    std::vector<NHTet> nh;
    std::vector<CosseratRod> cos;
    std::vector<GroundContact> ground;
    std::vector<RigidGround> rground;
    std::vector<Joint> joints;
    std::vector<Joint2> joints2;   // two-sided fixed joints (new)

    std::vector<uint32_t> tris;     // collision triangles (3 particle indices each)
    BroadPhase bp;
    uint64_t last_candidates = 0;

    // Collision primitives (CollisionPrimitivePool, one object per add call)
    // and the dynamic contact pipeline; off until enable_dynamic_contacts().
    std::vector<Prim> prims;
    uint32_t next_object = 0;
    bool dynamic = false;
    Real collision_margin = 1e-3;
    bool self_collision = false;   // vertex-face contacts between collision triangles (new)
    DynPool dyn[DYN_COUNT];   // TRIANGLE_RIGID_COLLISION, TRIANGLE_ROD_COLLISION, rigid-rigid, vertex-triangle (new)
    std::vector<uint32_t> vertex_owner;   // per particle: lowest-numbered triangle primitive holding it
    std::vector<Detected> prev_det, cur_det;
    std::vector<uint32_t> prev_order, cur_order;
    DynAdjacency dadj;
    Recoloring rc;

    StaticAdjacency adj;
    Coloring coloring;
    std::vector<uint8_t> is_conflicted;

    // Setup (simobject/TetMeshObject.cpp:46-96 and simulation/Simulation.hpp:117-126).
    uint32_t add_tet_mesh(uint32_t nv, const Real* rest_x, uint32_t nt, const uint32_t* tets,
                          Real E, Real nu, Real density, Real kd, bool with_ground,
                          Real mu_s, Real mu_k);
    // ParticlePool::addParticle (common/ParticlePool.hpp:213-238) for n free particles.
    uint32_t add_particles(uint32_t n, const Real* x, const Real* m);
    // ParticlePool::addOrientedParticle (common/ParticlePool.hpp:248-263).
    uint32_t add_oriented(V3 p, Quat q, Real m, V3 inertia);
    // Rod ctor (simobject/Rod.cpp:10-78) + per-node rigid ground contacts
    // (simulation/Simulation.hpp:128-138).  Returns the first node.
    uint32_t add_rod(uint32_t elements, Real length, V3 curvature, Real E, Real nu, Real density, Real diameter,
                     V3 pos, Quat rot, V3 vel, bool base_fixed, bool tip_fixed, bool with_ground);
    // RigidSphere ctor (simobject/rigid/RigidSphere.cpp:8-22, RigidObjectBase.cpp:10-20)
    // + ground contact (simulation/Simulation.hpp:108-115).
    uint32_t add_sphere(Real radius, Real density, V3 pos, Quat rot, bool fixed_, bool with_ground);
    // OneSidedFixedJointEnergyPool::addEnergy (energy/joint/OneSidedFixedJointEnergyPool.hpp:33-50).
    uint32_t add_joint(uint32_t p, V3 pos_offset, Quat rot_offset, V3 ref_position, Quat ref_rotation);
    // Two-sided fixed joint between oriented particles p0 and p1 (new), k
    // starting at the one-sided joint's k_start.
    uint32_t add_joint2(uint32_t p0, uint32_t p1, V3 pos_offset0, Quat rot_offset0, V3 pos_offset1, Quat rot_offset1);

    void build_graph();     // adjacency + DSATUR + work list (simulation/Simulation.cpp:87-89)

    // CollisionPrimitivePool::addObject (collision/CollisionPrimitivePool.cpp:36-87)
    uint32_t add_collision_triangles(uint32_t n, const uint32_t* t);
    uint32_t add_collision_segments(uint32_t n_nodes, const uint32_t* nodes, Real radius);
    uint32_t add_collision_sphere(uint32_t p, Real radius);
    // Run the reference's per-step detectCollisionsAndRecolor over every
    // primitive (capacity: slots per dynamic pool).  After build_graph().
    void enable_dynamic_contacts(Real margin, uint32_t capacity);
    void detect_and_recolor();

    // One time step of VBDSolver::solve_Parallel, single threaded.
    void step();
    // Same step on n_threads SPMD workers with a spin barrier between phases.
    void step_mt(uint32_t n_threads, uint32_t n_steps);

    Stats stats() const;
    // sum over free particles in [lo, hi) of |m/dt^2 (x - y) + G|^2 (positional block)
    Real residual2(uint32_t lo, uint32_t hi) const;
    Real mg2() const;
    uint32_t last_iters = 0;
    uint64_t total_iters = 0;

    void accumulate_vertex(uint32_t p, V3& G, M3& H) const;
    void accumulate_vertex6(uint32_t p, V6& G, M6& H) const;
    bool oriented(uint32_t p) const { return !rs.rot_of.empty() && rs.rot_of[p] >= 0; }

private:
    void process_pair(uint32_t a, uint32_t b, std::vector<Detected>& out);
    void mark_dirty(uint32_t p);
    void add_contact(Detected& d);
    void remove_contact(const Detected& d);
    void update_contact(const Detected& d);
    void handle_detected();
    void build_dynamic_adjacency();
    void incremental_recoloring();
    void phase_after_timestep_dyn();
    void phase_dual_dyn();
    void phase_commit(uint32_t c);
    void phase_after_timestep(uint32_t lo_nh, uint32_t hi_nh, uint32_t lo_g, uint32_t hi_g);
    void phase_after_timestep_rigid(uint32_t lo, uint32_t hi);
    void phase_inertial(uint32_t lo, uint32_t hi);
    void phase_color(uint32_t c, uint32_t lo, uint32_t hi);
    void phase_chebyshev(Real omega, uint32_t lo, uint32_t hi);
    void phase_dual(uint32_t lo, uint32_t hi);
    void phase_dual_rigid(uint32_t lo, uint32_t hi);
    void phase_velocity(uint32_t lo, uint32_t hi);
    void solve_particle(uint32_t p);
    void solve_particle6(uint32_t p);
};

}  // namespace oracle
