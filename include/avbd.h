/*
 * avbd.h -- C ABI of the MI355X-native AVBD per-step solver (libavbd_amd.so).
 *
 * Plain pointers and sizes only; no C++ or torch types cross this boundary.
 * Every entry point returns 0 on success or a negative AVBD_E_* code; the
 * message for the last failure on a context is returned by avbd_last_error().
 * Host arrays are borrowed for the duration of the call.  One caller thread
 * per context (contexts are independent; each owns one HIP stream).
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   VBDSolver(SimulationContext*, unsigned solver_iters, Real iter_acceleration)
 *       + solve_Parallel(WorkerThreadContext&)     simulation/VBDSolver.hpp:37-103
 *   SimulationExecutor::timeStep                   simulation/SimulationExecutor.hpp:107-156
 *   SimulationContext(capacities...) + params      simulation/SimulationContext.hpp:18-60,
 *                                                  simulation/SimulationParams.hpp:8-15
 *   ParticlePool::addParticle / addOrientedParticle common/ParticlePool.hpp:213-263
 *   EnergyRegistry pools' addEnergy(...)           energy/EnergyRegistry.hpp:17-173,
 *                                                  energy/elastic/NeoHookeanEnergyPool.hpp:40-67,
 *                                                  energy/collision/GroundCollisionEnergyPool.hpp:26-38
 *   TetMeshObject::setup (setup formulas)          simobject/TetMeshObject.cpp:21-113
 *   Rod / RigidSphere constructors                 simobject/Rod.cpp:10-78,
 *                                                  simobject/rigid/RigidSphere.cpp:8-22,
 *                                                  simobject/rigid/RigidObjectBase.cpp:10-20
 *   Simulation::_addGroundCollisionConstraintsForObject  simulation/Simulation.hpp:117-126
 *   StaticParticleAdjacency::buildAdjacency + ColorList::buildInitialColorList
 *                                                  simulation/Simulation.cpp:87-89
 */
#ifndef AVBD_H
#define AVBD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (reference: exceptions at common/TombstonePool.hpp:64,
 *      common/ColorList.hpp:36, collision/CollisionDetector.cpp:211,686) ---- */
#define AVBD_OK 0
#define AVBD_E_ARG (-1)                 /* invalid argument / state                  */
#define AVBD_E_CAPACITY (-2)            /* "Pool has exceeded capacity"              */
#define AVBD_E_COLORS (-3)              /* "Ran out of colors!" (more than 256)      */
#define AVBD_E_UNSUPPORTED_CONTACT (-4) /* "Unsupported collision!"                  */
#define AVBD_E_NONFINITE (-5)           /* NaN/Inf detected in the state             */
#define AVBD_E_HIP (-6)                 /* HIP runtime error                         */
#define AVBD_E_NOT_BUILT (-7)           /* avbd_build_graph() has not been called    */

typedef struct avbd_ctx avbd_ctx;

/* ---- energy types, in the reference's EnergyType order (common/common.hpp:156-163) ---- */
typedef enum {
    AVBD_NEO_HOOKEAN = 0,
    AVBD_COSSERAT_ROD = 1,
    AVBD_GROUND_COLLISION = 2,
    AVBD_RIGID_BODY_GROUND_COLLISION = 3,
    AVBD_TRIANGLE_RIGID_COLLISION = 4,
    AVBD_TRIANGLE_ROD_COLLISION = 5,
    AVBD_ONE_SIDED_FIXED_JOINT = 6,
    AVBD_FIXED_JOINT = 7 /* new: two-sided fixed joint between two oriented particles */
} avbd_energy_type;

/* ---- context creation flags ---- */
#define AVBD_FLAG_NO_GRAPH 1u    /* launch kernels eagerly instead of replaying a hipGraph */
#define AVBD_FLAG_NO_PERSISTENT 4u  /* one kernel launch per solver phase (the default) */
#define AVBD_FLAG_PERSISTENT 8u     /* rejected by avbd_create (AVBD_E_ARG): the persistent step
                                       kernel was removed.  Every grid barrier needed agent-scope
                                       release / acquire fences, which write back and invalidate
                                       the L2 of every XCD, so on MI355X one launch per phase
                                       replayed from a hipGraph was 25x faster */
#define AVBD_FLAG_SERIAL_BROADPHASE 2u  /* run the triangle broad phase on the solve stream
                                           instead of concurrently on a second stream */
#define AVBD_FLAG_FP64_STATE 32u /* keep positions, the inertial target and the Chebyshev history in
                                    FP64 on the device and evaluate the vertex-block solve, the
                                    ground terms and the force residual in FP64 (material and contact
                                    records stay FP32).  An FP32 position resolves ~6e-8 m at 0.5 m,
                                    which bounds the relative force residual an FP32 state can reach
                                    near 1e-3..1e-2 on fine stiff meshes; this mode reaches 1e-4 and
                                    below.  Positional particles only (no oriented particles);
                                    dynamic contacts between collision triangles only (the
                                    vertex-face contacts of avbd_set_self_collision), AVBD_E_ARG
                                    for rod segments and spheres */
#define AVBD_FLAG_HOST_COLORING 64u /* run the DSATUR colouring of avbd_build_graph on the host
                                       instead of the device (same colours; without a device the
                                       host colours anyway) */
#define AVBD_FLAG_SEPARATE_CHEBYSHEV 16u /* run the Chebyshev / dual update of every iteration as
                                            its own launch.  By default, in scenes without oriented
                                            particles, it is fused into the iteration's last colour
                                            phase over two position buffers (identical results) */

/* SimulationContext capacities + SimulationParams + VBDSolver arguments.
 * Reference: simulation/SimulationContext.hpp:46-59, simulation/SimulationParams.hpp:8-15,
 * simulation/VBDSolver.hpp:37-42.  Gravity is (0, -g_accel, 0) (y-up,
 * simulation/VBDSolver.hpp:51). */
typedef struct {
    uint32_t particle_capacity;
    uint32_t energy_capacity;      /* per energy type */
    double dt;
    double g_accel;
    uint32_t solver_iters;
    double iter_acceleration;      /* Chebyshev rho */
    uint32_t flags;
    int32_t device;                /* HIP device ordinal */
    uint32_t lanes_per_vertex;     /* 0 = autotuned; else 4, 8, 16 or 32 */
    uint32_t block_size;           /* 0 = autotuned; else 64..1024, multiple of 64 */
    uint32_t gather_batch;         /* tet incidences a lane gathers at once: 0 = autotuned
                                      (1 when lanes and block size are both given); else 1, 2 or 4 */
} avbd_scene_desc;

/* NeoHookeanEnergyInfo, energy/elastic/NeoHookeanEnergyPool.hpp:8-17 (FP64, row-major 3x3). */
typedef struct {
    uint32_t particle_indices[4];
    double lambda, mu, kd;
    double Q[9];
    double rest_volume;
    double E_prev[9];
} avbd_neo_hookean_record;

/* GroundCollisionEnergyInfo, energy/collision/CollisionConstraintEnergyPool.hpp:9-30
 * + energy/collision/GroundCollisionEnergyPool.hpp:8-13. */
typedef struct {
    uint32_t particle_index;
    uint32_t use_static;
    uint32_t unclamped_lambda_tb;
    uint32_t pad_;
    double k_start, k_n, k_t, k_b;
    double lambda_n, lambda_t, lambda_b;
    double C_n_prev;
    double mu_s, mu_k;
    double cp_x, cp_z;
} avbd_ground_record;

/* CosseratRodEnergyInfo, energy/elastic/CosseratRodEnergyPool.hpp:8-14:
 * stiffness = diag(GA, GA, EA, EI, EI, GJ), rest length, rest curvature. */
typedef struct {
    uint32_t particle_indices[2];
    double stiffness[6];
    double rest_length;
    double rest_curvature[3];
} avbd_cosserat_record;

/* Collision geometry of a rigid-ground contact: the reference's
 * CollisionGeometryType values (collision/CollisionGeometryType.hpp). */
#define AVBD_SHAPE_ROD_SEGMENT 1
#define AVBD_SHAPE_SPHERE 2
#define AVBD_SHAPE_BOX 3
#define AVBD_SHAPE_CAPSULE 4

/* RigidBodyGroundCollisionEnergyInfo (energy/collision/RigidBodyGroundCollisionEnergyPool.hpp:8-15)
 * + the AL state of CollisionConstraintEnergyInfo; normal y, tangents x and z. */
typedef struct {
    uint32_t particle_index;
    uint32_t shape;                /* AVBD_SHAPE_* */
    uint32_t use_static;
    uint32_t unclamped_lambda_tb;
    double radius;                 /* sphere / rod radius (lowest-point contact) */
    double cp_rb_local[3];         /* contact point in the body frame */
    double cp_ground[3];
    double k_start, k_n, k_t, k_b;
    double lambda_n, lambda_t, lambda_b;
    double C_n_prev;
    double mu_s, mu_k;
} avbd_rigid_ground_record;

/* OneSidedFixedJointEnergyInfo (energy/joint/OneSidedFixedJointEnergyPool.hpp:8-16)
 * + the vector AL state of HardConstraintEnergyPool.  Quaternions are
 * (w, x, y, z). */
typedef struct {
    uint32_t particle_index;
    uint32_t pad_;
    double position_offset[3];
    double rotation_offset[4];
    double ref_position[3];
    double ref_rotation[4];
    double k[6];
    double lambda[6];
    double C_prev[6];
} avbd_joint_record;

/* Two-sided fixed joint (new; the reference's joint holds one particle to a
 * fixed world pose, energy/joint/OneSidedFixedJointEnergySolver.hpp:16-34,
 * so a rigid body attached to a rod tip has to be welded into the node).
 * Joint frame i = (x_i + R_i position_offset[i], q_i rotation_offset[i]) on
 * oriented particle particle_index[i];
 * C = [frame_0 position - frame_1 position ; Minus_S3(frame_0 rotation,
 * frame_1 rotation)] under the one-sided joint's vector AL (k_start 1e4,
 * unbounded multipliers).  With particle 1 fixed it is the one-sided joint
 * with ref = frame 1.  The two particles are adjacent in the colouring.
 * Quaternions are (w, x, y, z). */
typedef struct {
    uint32_t particle_index[2];
    double position_offset[2][3];
    double rotation_offset[2][4];
    double k[6];
    double lambda[6];
    double C_prev[6];
} avbd_fixed_joint_record;

/* Per-step statistics (new: the reference has no residual; see the
 * definitions in the oracle and the solver sources). */
typedef struct {
    double residual_abs;    /* ||r||_2 over free particles                     */
    double residual_rel;    /* ||r||_2 / ||M g||_2                              */
    double incremental_potential;
    uint32_t num_colors;
    uint32_t num_particles;
    uint32_t num_tets;
    uint32_t num_ground;
    uint64_t steps_taken;
    uint64_t vertex_updates; /* cumulative vertex-block solves                 */
    double last_step_ms;     /* device time of the last avbd_step() call        */
    uint64_t iterations_total;      /* Gauss-Seidel iterations over all steps   */
    uint32_t last_step_iterations;  /* iterations of the last step              */
    uint32_t pad_;
    uint64_t warm_start_predictions; /* scene-steps started from the network    */
    uint64_t warm_start_fallbacks;   /* of those, kept the inertial start       */
    uint64_t steps_at_budget;        /* residual mode: steps that used max_iters
                                        without reaching the tolerance           */
} avbd_stats;

/* Context lifetime -------------------------------------------------------- */
int avbd_create(const avbd_scene_desc* desc, avbd_ctx** out_ctx);
void avbd_destroy(avbd_ctx* ctx);
/* Last error message for ctx (or the last avbd_create failure if ctx is NULL). */
const char* avbd_last_error(avbd_ctx* ctx);

/* Pools (host-side staging until avbd_build_graph) ------------------------- */

/* ParticlePool::addParticle (common/ParticlePool.hpp:213-238) for n particles:
 * x, v, v_prev, x_prev are 3n doubles (NULL = zeros / x), mass n doubles,
 * fixed n bytes (NULL = not fixed).  *first receives the first new index. */
int avbd_add_particles(avbd_ctx* ctx, uint32_t n, const double* x, const double* v, const double* v_prev,
                       const double* x_prev, const double* mass, const uint8_t* fixed, uint32_t* first);

/* ParticlePool::addOrientedParticle (common/ParticlePool.hpp:248-263) for n
 * particles: x, v, w 3n doubles (NULL v/w = zero), q 4n doubles (w, x, y, z;
 * NULL = identity), mass n, inertia 3n (principal moments), fixed n bytes
 * (NULL = not fixed).  *first receives the first new index. */
int avbd_add_oriented_particles(avbd_ctx* ctx, uint32_t n, const double* x, const double* q, const double* v,
                                const double* w, const double* mass, const double* inertia, const uint8_t* fixed,
                                uint32_t* first);

/* Rod ctor (simobject/Rod.cpp:10-78): elements+1 oriented nodes from pos along
 * the body z axis of rot (quaternion w, x, y, z), bent by the rest curvature;
 * lumped masses and inertias; one Cosserat energy per element; base/tip
 * fixed flags; with_ground adds one rigid-ground contact per node
 * (simulation/Simulation.hpp:128-138, k_start = m/dt^2, mu 0.5/0.2).
 * *first receives the first node. */
int avbd_add_rod(avbd_ctx* ctx, uint32_t elements, double length, const double* curvature, double E, double nu,
                 double density, double diameter, const double* pos, const double* rot, const double* vel,
                 int base_fixed, int tip_fixed, int with_ground, uint32_t* first);

/* RigidSphere ctor (simobject/rigid/RigidSphere.cpp:8-22, RigidObjectBase.cpp:10-20):
 * one oriented particle of mass 4/3 pi r^3 density and inertia 2/5 m r^2;
 * with_ground adds its rigid-ground contact (simulation/Simulation.hpp:108-115). */
int avbd_add_sphere(avbd_ctx* ctx, double radius, double density, const double* pos, const double* rot, int fixed,
                    int with_ground, uint32_t* first);

/* EnergyRegistry pools: append count records of the given type
 * (avbd_neo_hookean_record / avbd_ground_record / avbd_cosserat_record /
 * avbd_rigid_ground_record / avbd_joint_record / avbd_fixed_joint_record). */
int avbd_add_energies(avbd_ctx* ctx, int type, const void* records, uint32_t count);

/* TetMeshObject::setup formulas (simobject/TetMeshObject.cpp:46-96) for a mesh
 * given in its rest state: Lame parameters from (E, nu), Q = X^-1,
 * V = |det X|/6, lumped masses 0.25 V density; plus, if with_ground, one
 * ground energy per vertex with k_start = m/dt^2 (simulation/Simulation.hpp:117-126).
 * Particles are appended; *first receives the first new particle index. */
int avbd_add_tet_mesh(avbd_ctx* ctx, uint32_t nv, const double* rest_x, uint32_t nt, const uint32_t* tets,
                      double E, double nu, double density, double kd, int with_ground, double mu_s,
                      double mu_k, uint32_t* first);

/* Overwrite the state of particles [first, first+n) (3n doubles each; NULL =
 * leave unchanged).  Valid before and after avbd_build_graph. */
int avbd_set_particle_state(avbd_ctx* ctx, uint32_t first, uint32_t n, const double* x, const double* x_prev,
                            const double* v, const double* v_prev);
int avbd_set_fixed(avbd_ctx* ctx, uint32_t first, uint32_t n, const uint8_t* fixed);
/* Per-particle external acceleration field (N/kg, 3n doubles) added to gravity
 * for particles [first, first+n) (new; the reference's only external
 * acceleration is gravity, simulation/VBDSolver.hpp:51).  It enters the
 * inertial target and the adaptive initial guess of _particleInertialUpdate
 * (simulation/VBDSolver.hpp:208-248) as gravity + a.  A field set before
 * avbd_build_graph can be changed afterwards; a scene built without one
 * keeps none (AVBD_E_ARG). */
int avbd_set_external_accel(avbd_ctx* ctx, uint32_t first, uint32_t n, const double* a);
/* Add mass dm and principal inertia dI (3, NULL = none; oriented particles
 * only) to particle p: a rigid body welded to it, e.g. a sphere at a rod tip
 * (new; the reference builds such bodies as separate objects). */
int avbd_add_mass(avbd_ctx* ctx, uint32_t p, double dm, const double* dI);
/* Orientation q (4n: w, x, y, z) and angular velocity w (3n) of the oriented
 * particles among [first, first+n), also taken as their previous-step values
 * (rows of positional particles are ignored; NULL = unchanged). */
int avbd_set_rotation_state(avbd_ctx* ctx, uint32_t first, uint32_t n, const double* q, const double* w);

/* Residual-targeted iteration (new; the reference runs a fixed solver_iters):
 * every step iterates until, in every scene s (avbd_set_scenes; one scene by
 * default), the relative force residual ||r_s||_2 / ||M_s g||_2 (free
 * particles of the scene; r = m/dt^2 (x - y) + dE/dx at the current iterate)
 * is <= rel_tol, checked every check_every iterations, or until max_iters
 * iterations (rounded up to a multiple of check_every).  rel_tol = 0
 * restores the fixed count.  Call before avbd_build_graph. */
int avbd_set_residual_target(avbd_ctx* ctx, double rel_tol, uint32_t max_iters, uint32_t check_every);
/* Relative residual ||r||_2 / ||M g||_2 over all scenes at each check of the
 * last step (residual-targeted mode): rel[k] after iteration
 * (k+1)*check_every; *count = checks that ran (the first 1024 are kept; at
 * most cap are written). */
int avbd_get_residual_history(avbd_ctx* ctx, double* rel, uint32_t cap, uint32_t* count);
/* The same checks' largest per-scene relative residual (the quantity the
 * stopping rule compares with rel_tol). */
int avbd_get_scene_residual_history(avbd_ctx* ctx, double* worst, uint32_t cap, uint32_t* count);
/* Iterations each scene ran in the last step (residual-targeted mode; new).
 * In a batch a scene stops iterating at the first check where it meets
 * rel_tol, while the others go on, so every scene ends its step where it
 * would on its own; the step's count is the largest.  *count = scenes (at
 * most cap are written). */
int avbd_get_scene_iterations(avbd_ctx* ctx, uint32_t* iters, uint32_t cap, uint32_t* count);

/* Scene partition for batching: scene s owns particles
 * [particle_offsets[s], particle_offsets[s+1]); each scene is coloured on its
 * own.  Default: one scene holding every particle. */
int avbd_set_scenes(avbd_ctx* ctx, uint32_t n_scenes, const uint32_t* particle_offsets);

/* StaticParticleAdjacency::buildAdjacency + ColorList::buildInitialColorList
 * (DSATUR) per scene, then upload every pool to the device as FP32 SoA. */
int avbd_build_graph(avbd_ctx* ctx);

/* Time stepping ------------------------------------------------------------ */

/* n_steps calls of VBDSolver::solve_Parallel (asynchronous on the ctx stream;
 * in residual-targeted mode the call returns once the last step's final
 * check intervals are queued: it waits for the check two intervals back to
 * decide whether to queue more, so it runs at most two check intervals ahead
 * of the device).  Fails with AVBD_E_NONFINITE once a synchronising call has
 * reported a non-finite state, until the state is set again. */
int avbd_step(avbd_ctx* ctx, uint32_t n_steps);
/* Waits for the ctx stream.  Every step checks its new positions and
 * velocities on the device; AVBD_E_NONFINITE when one was NaN or Inf (the
 * downloads and avbd_get_stats report it too, after filling their outputs).
 * avbd_set_particle_state and avbd_load_checkpoint clear it. */
int avbd_synchronize(avbd_ctx* ctx);

/* Readback (synchronising) -------------------------------------------------- */
int avbd_download_state(avbd_ctx* ctx, double* x, double* v);
int avbd_download_inertial(avbd_ctx* ctx, double* y);
/* Per particle: orientation (4: w, x, y, z; identity for positional
 * particles) and angular velocity (3; zero for positional particles). */
int avbd_download_rotations(avbd_ctx* ctx, double* q, double* w);
int avbd_download_energies(avbd_ctx* ctx, int type, void* records, uint32_t count);
/* color: one entry per particle; *num_colors = max over scenes. */
int avbd_download_coloring(avbd_ctx* ctx, uint32_t* color, uint32_t* num_colors);
/* Force residual / incremental potential at the current state (device pass). */
int avbd_get_stats(avbd_ctx* ctx, avbd_stats* out);

/* Introspection for tests and the bench -------------------------------------- */
uint32_t avbd_num_particles(avbd_ctx* ctx);
uint32_t avbd_num_energies(avbd_ctx* ctx, int type);
/* Kernel launch geometry actually in use (after autotuning); *gather_batch =
 * tet incidences one lane gathers at once (1, 2 or 4). Any pointer may be null. */
int avbd_get_launch_config(avbd_ctx* ctx, uint32_t* lanes_per_vertex, uint32_t* block_size, uint32_t* gather_batch);
/* *persistent_grid = blocks of the persistent step kernel, or 0 when the
 * step runs as one launch per phase; *fused_chebyshev = 1 when the Chebyshev /
 * dual update runs inside each iteration's last colour phase
 * (AVBD_FLAG_SEPARATE_CHEBYSHEV). Either pointer may be null. */
int avbd_get_step_mode(avbd_ctx* ctx, uint32_t* persistent_grid, uint32_t* fused_chebyshev);
/* Kernel timing: while enabled, avbd_step() launches eagerly and brackets
 * every per-colour vertex-solve launch with HIP events on the ctx stream.
 * avbd_get_solve_timing() returns the average duration (ms) of those
 * launches over the last avbd_step() call and how many it covers. */
int avbd_set_timing(avbd_ctx* ctx, int enable);
int avbd_get_solve_timing(avbd_ctx* ctx, double* avg_ms, uint32_t* launches);
/* Average per-step duration (ms) of the broad phase over the last timed
 * avbd_step() call (0 without collision triangles); with dynamic contacts,
 * of the whole detection round over the last avbd_step() call (device
 * bounds, LBVH and narrow phase, host lifecycle and recolouring, uploads;
 * host wall clock). */
int avbd_get_broadphase_timing(avbd_ctx* ctx, double* avg_ms);

/* Collision primitives and broad phase ------------------------------------
 * CollisionPrimitivePool::addObject(TetMeshObject) (collision/CollisionPrimitivePool.cpp:36-50):
 * n triangles (3 particle indices each) join the LBVH.  When any exist, every
 * avbd_step() first rebuilds the LBVH and self-traverses it, as
 * CollisionDetector::detectCollisionsAndRecolor_Parallel does before
 * VBDSolver::solve_Parallel (simulation/SimulationExecutor.hpp:142-156). */
int avbd_add_collision_triangles(avbd_ctx* ctx, uint32_t n, const uint32_t* tris);
/* Run the broad phase now on the current state (LBVHBuilder::buildBVH +
 * LBVHTraversal self traversal, collision/LBVHBuilder.cpp:13-25,
 * collision/LBVHTraversal.hpp:30-88). */
int avbd_detect_collisions(avbd_ctx* ctx);
/* Last LBVH: morton[n] and sorted[n] per primitive, left/right/parent[2n-1],
 * node_box[6(2n-1)] (min xyz, max xyz), *root (any pointer may be NULL). */
int avbd_download_bvh(avbd_ctx* ctx, uint64_t* morton, uint32_t* sorted, uint32_t* left, uint32_t* right,
                      uint32_t* parent, double* node_box, uint32_t* root);
/* Candidate leaf pairs (a < b, sorted-order leaf indices) of the last
 * traversal: up to cap pairs into pairs[2*cap]; *count = total found
 * (> cap means truncated).  Order is unspecified. */
int avbd_download_candidates(avbd_ctx* ctx, uint32_t* pairs, uint32_t cap, uint32_t* count);

/* Dynamic contacts --------------------------------------------------------
 * CollisionPrimitivePool::addObject(Rod) (collision/CollisionPrimitivePool.cpp:52-72):
 * the n_nodes - 1 segments between consecutive nodes (oriented particles),
 * radius `radius`, as one object.  Primitives are numbered in call order
 * across avbd_add_collision_triangles / _segments / _sphere. */
int avbd_add_collision_segments(avbd_ctx* ctx, uint32_t n_nodes, const uint32_t* nodes, double radius);
/* CollisionPrimitivePool::addObject(RigidSphere) (collision/CollisionPrimitivePool.cpp:74-87):
 * a sphere of `radius` centred on oriented particle p, as one object. */
int avbd_add_collision_sphere(avbd_ctx* ctx, uint32_t p, double radius);
/* Turn on the reference's per-step contact pipeline (after avbd_build_graph):
 * every avbd_step() then runs CollisionDetector::detectCollisionsAndRecolor_Parallel
 * (collision/CollisionDetector.cpp:57-129) before the solve -- LBVH over
 * every primitive and the narrow phase on the device; contact lifecycle
 * (_handleDetectedCollisions, :272-501) and dynamic adjacency on the host;
 * incremental recolouring (common/ColorList.hpp:381-522) and the work-list
 * layout on the device -- and the solve includes the triangle-rigid /
 * triangle-rod AL contacts.
 * collision_margin: SimulationParams::collision_margin; capacity: contacts
 * per kind (0 = 65,536).  Steps then synchronise once per step (the host
 * needs the narrow-phase output); fixed-count steps launch eagerly, and a
 * residual target replays head / interval / tail graphs captured for the
 * round's layout (captured again when the launch sizes, the active contacts
 * or the adjacency buffers change).  A step with no active contact on the
 * set-up work list runs the static kernels.  Recording warm-start training
 * data is not available in this mode (AVBD_E_ARG).  With
 * AVBD_FLAG_FP64_STATE only collision triangles are allowed (their
 * vertex-face contacts are evaluated from the FP64 positions).  Sphere-sphere pairs become
 * rigid-rigid AL contacts (new; the reference throws at :686-687): normal
 * and friction on both bodies, C = dir . (x_2 + R_2 cp_2 - x_1 - R_1 cp_1).
 * A rod-sphere pair that shares no particle fails the step with
 * AVBD_E_UNSUPPORTED_CONTACT, as the reference throws "Unsupported
 * collision!" (:210-213); a sphere welded to a rod's node does not collide
 * with that rod. */
int avbd_set_dynamic_contacts(avbd_ctx* ctx, double collision_margin, uint32_t capacity);
/* New: vertex-face contacts between collision triangles, of one mesh (self
 * collision) or of two.  The reference declares TriangleTriangle_VertexFace
 * (collision/DetectedCollision.hpp:7,29) but its triangle-triangle narrow
 * phase is empty (collision/CollisionDetector.cpp:589-599), which is the
 * behaviour with this off (the default).  On: the triangle bounds are
 * widened by the collision margin, and for every candidate triangle pair
 * each vertex of either triangle that is not a vertex of the other, tested
 * from its owner (the lowest-numbered triangle holding it, so that each
 * vertex-face pair is found once), is in contact with the other triangle
 * when its distance to it is below the margin.  The normal is the face
 * normal turned towards the face as seen from the vertex (a persisting
 * contact keeps its side), and the contact is the triangle-rigid AL contact
 * with the vertex as the contact point: C = dir . (x_v - sum_i barys_i x_i)
 * for dir = n, t, b, normal and friction on all four particles (kind 3 of
 * avbd_download_contacts; DetectedCollisionType 0).  Call before
 * avbd_set_dynamic_contacts. */
int avbd_set_self_collision(avbd_ctx* ctx, int enable);

/* One dynamic contact (CollisionConstraintEnergyInfo + TriangleRigid /
 * TriangleRod info, energy/collision/CollisionConstraintEnergyPool.hpp:9-30,
 * energy/collision/TriangleRigidCollisionEnergyPool.hpp:8-14,
 * energy/collision/TriangleRodCollisionEnergyPool.hpp:8-15). */
typedef struct avbd_contact {
    uint32_t slot;          /* pool slot (TombstonePool index) */
    uint32_t idx[5];        /* triangle vertices, then the body or both rod nodes;
                               rigid-rigid: the two bodies; vertex-triangle:
                               the face's vertices, then the vertex */
    double n[3], t[3], b[3];
    double barys[3];        /* rigid-rigid: the second body's contact point (body frame) */
    double s;               /* triangle-rod: contact point along the segment */
    double cp_local[3];
    double k_start, k_n, k_t, k_b;
    double lambda_n, lambda_t, lambda_b, C_n_prev;
    uint8_t use_static, unclamped, pad_[6];
} avbd_contact;
/* Active contacts of one kind (0 triangle-rigid, 1 triangle-rod, 2
 * rigid-rigid, 3 vertex-triangle) in slot order: up to cap into out;
 * *count = active total. */
int avbd_download_contacts(avbd_ctx* ctx, uint32_t kind, avbd_contact* out, uint32_t cap, uint32_t* count);
/* Contacts detected by the last step's detection in key order
 * (DetectedCollision::generateKey, collision/DetectedCollision.hpp:43-66):
 * key, pool slot, DetectedCollisionType (0 vertex-face, 2 triangle-rigid,
 * 3 triangle-rod, 6 rigid-rigid). */
int avbd_download_detected(avbd_ctx* ctx, uint64_t* keys, uint32_t* slots, uint8_t* kinds, uint32_t cap,
                           uint32_t* count);
/* is_conflicted per particle after the last recolouring (n bytes). */
int avbd_download_conflicts(avbd_ctx* ctx, uint8_t* flags);

/* Checkpoint / resume (new; the reference keeps no checkpoints) ---------------
 * Every mutable part of a built context -- particle state and histories,
 * E_prev, contact / joint multipliers and stiffnesses, rotation state, the
 * residual-control word, the dynamic contact records and, with dynamic
 * contacts, the colouring and the host contact state -- into a caller
 * buffer of avbd_checkpoint_size() bytes.  Loading it into a context built
 * from the same scene with the same modes resumes the run: the following
 * steps reproduce the original run bit for bit.  Synchronises. */
int avbd_checkpoint_size(avbd_ctx* ctx, uint64_t* bytes);
int avbd_save_checkpoint(avbd_ctx* ctx, void* buf, uint64_t cap, uint64_t* written);
int avbd_load_checkpoint(avbd_ctx* ctx, const void* buf, uint64_t size);

/* Learned warm start (new; replaces the inertial / adaptive initial guess of
 * _particleInertialUpdate, simulation/VBDSolver.hpp:208-248, whenever it
 * lowers the incremental potential of the scene).  A per-vertex MLP
 * (24 features -> 128 -> 128 -> 128 -> 3, SiLU, BF16 weights and activations,
 * FP32 accumulation on the matrix cores) predicts the step displacement from
 * the initial guess.  Parameters: for each layer W (out x in, row-major, in
 * padded 24 -> 32 and out 3 -> 16 with zeros) then b (out); then feature
 * mean (24), feature std (24), target mean (3), target std (3).  Features and
 * targets are FP32 rows of 24 and 3 (targets in units of dt^2 g). */
uint64_t avbd_mlp_param_count(void);
int avbd_mlp_init(avbd_ctx* ctx, uint64_t seed);
int avbd_mlp_set_params(avbd_ctx* ctx, const float* params, uint64_t count);
int avbd_mlp_get_params(avbd_ctx* ctx, float* params, uint64_t count);
/* feature mean (24), feature std (24), target mean (3), target std (3) */
int avbd_mlp_set_normalization(avbd_ctx* ctx, const float* norm);
/* One forward + backward pass (MSE on the normalised targets); grads has the
 * parameter layout without the normalisation block; *loss = mean loss. */
int avbd_mlp_forward_backward(avbd_ctx* ctx, const float* feats, const float* tgts, uint32_t rows, float* grads,
                              uint64_t count, float* loss);
int avbd_mlp_predict(avbd_ctx* ctx, const float* feats, uint32_t rows, float* out);
/* Adam over the data set for `epochs` passes of minibatches (batch % 16 == 0);
 * *loss = mean loss of the last epoch. */
int avbd_mlp_train(avbd_ctx* ctx, const float* feats, const float* tgts, uint32_t rows, uint32_t epochs,
                   uint32_t batch, float lr, uint64_t seed, float* loss);
/* n_steps time steps that also record, per positional work item, the
 * features at the start of the step and the target (x_end - x_start)/(dt^2 g);
 * rows = n_steps * items, in work-item order per step. */
int avbd_record_steps(avbd_ctx* ctx, uint32_t n_steps, float* feats, float* tgts, uint32_t cap_rows,
                      uint32_t* rows);
/* Start every step from the network's prediction (with the per-scene
 * incremental-potential fallback) or, with enable = 0, from the inertial guess. */
int avbd_set_warm_start(avbd_ctx* ctx, int enable);
/* The warm start's choice in the last step, per scene (new): ip_delta[3 s + j]
 * = IP(x0 + a_j p) - IP(x0) for the step lengths a = 1, 1/2, 1/4 along the
 * prediction p from the inertial start x0, and step[s] = the a whose change
 * is the most negative, or 0 when none is negative (the inertial start is
 * kept).  *n_scenes = scenes (at most cap are written). */
int avbd_get_warm_start_decision(avbd_ctx* ctx, double* ip_delta, float* step, uint32_t cap, uint32_t* n_scenes);
/* Test hook: C (M x N) = A (M x K) B (K x N) through the BF16 MFMA GEMM
 * kernel, with A (ta) or B (tb) stored transposed on the device; FP32 row-major
 * host arrays, sizes multiples of 16. */
int avbd_mlp_gemm_test(avbd_ctx* ctx, uint32_t M, uint32_t N, uint32_t K, int ta, int tb, const float* A,
                       const float* B, float* C);

/* Measurement helper with no counterpart in the reference: `reps` rounds of
 * the per-colour vertex-solve launches back to back on the ctx stream (into a
 * scratch output, state untouched), bracketed by HIP events; *avg_ms = average
 * duration per launch. */
int avbd_time_solve_launches(avbd_ctx* ctx, uint32_t reps, double* avg_ms, uint32_t* launches);

/* Measurement helper with no counterpart in the reference: runs a plain
 * device-to-device streaming copy kernel (calib_copy_kernel) of `bytes` bytes
 * `iters` times on the ctx stream and returns the average duration, so that
 * HBM byte counters can be scaled against a kernel of known traffic. */
int avbd_calibrate_copy(avbd_ctx* ctx, uint64_t bytes, uint32_t iters, double* avg_ms);

/* Allocation-free broad phase.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under stream capture,
 * serving loops, harnesses that forbid hipMalloc).  The same LBVH build and
 * self traversal as avbd_detect_collisions + avbd_download_candidates, over
 * caller-owned device buffers and without a context:
 *   x4, v4    device arrays of float4 per particle (xyz + any w), the
 *             positions and velocities the speculative bounds are built from;
 *   tris      device array of 3 particle indices per triangle (n_tris);
 *   pairs     device output of up to pair_cap candidate leaf pairs (a < b,
 *             sorted-order leaf indices, 2 uint32 each, order unspecified);
 *   counters  device output of 2 uint32: [0] pairs found (> pair_cap means
 *             truncated), [1] status (bit 0: traversal stack overflow);
 *   stream    hipStream_t to launch on (NULL = the null stream).
 * Asynchronous: outputs are valid once the stream has reached this point.
 * Every intermediate array lives in one static device array of the library
 * (about 6.7 MiB per GPU, held from the first call on a device until the
 * process ends), so a call allocates nothing and never synchronises.  At most
 * AVBD_NOALLOC_MAX_TRIANGLES triangles (AVBD_E_CAPACITY beyond).  Calls on one
 * GPU share that array: calls that overlap in time on one GPU race.  Errors
 * are reported through avbd_last_error(NULL). */
#define AVBD_NOALLOC_MAX_TRIANGLES 16384u
int avbd_detect_collisions_noalloc(const float* x4, const float* v4, const uint32_t* tris, uint32_t n_tris,
                                   double dt, uint32_t* pairs, uint32_t pair_cap, uint32_t* counters, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* AVBD_H */
