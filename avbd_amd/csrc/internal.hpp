// internal.hpp -- device record layouts and the solver context behind the
// avbd C ABI.  Host staging pools are FP64 (the reference's Real = double,
// common/common.hpp:77); device pools are FP32 SoA/AoS records.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "../../include/avbd.h"
#include "mlp.hpp"

namespace avbd {

// AL constants CONSTRAINT_ALPHA, STIFFNESS_BETA, STIFFNESS_GAMMA (common/common.hpp:13-15)
constexpr float kAlpha = 0.95f;
constexpr float kBeta = 10.0f;
constexpr float kGamma = 0.99f;

// ---------------------------------------------------------------------------
// Device records
// ---------------------------------------------------------------------------

// Neo-Hookean tet, static part: one 64-byte record (half a 128-byte line),
// fetched with four 16-byte loads.  VL = V*lambda, VM = V*mu and
// alpha = 1 + mu/lambda are folded at upload (energy/elastic/NeoHookeanEnergySolver.hpp:195-196).
struct alignas(16) TetRec {
    uint32_t idx[4];
    float q[9];   // Q, row-major
    float VL;
    float VM;
    float alpha;
};
static_assert(sizeof(TetRec) == 64, "TetRec must be 64 bytes");

// Strain-rate damping part (only read when some tet has kd != 0).
// E_prev is symmetric (0.5 (F^T F - I)), so 6 entries are stored.
struct alignas(16) TetDamp {
    float e00, e11, e22, e01;
    float e02, e12, damp;   // damp = 2/dt * V * kd
    float pad;
};
static_assert(sizeof(TetDamp) == 32, "TetDamp must be 32 bytes");

// Ground AL contact (CollisionConstraintEnergyInfo + cp_x/cp_z), 64 bytes.
struct alignas(16) GroundRec {
    float k_start, k_n, k_t, k_b;
    float lambda_n, lambda_t, lambda_b, C_n_prev;
    float mu_s, mu_k, cp_x, cp_z;
    uint32_t p;
    uint32_t flags;   // bit0 use_static, bit1 unclamped_lambda_tb
    uint32_t pad0, pad1;
};
static_assert(sizeof(GroundRec) == 64, "GroundRec must be 64 bytes");

constexpr uint32_t GF_USE_STATIC = 1u;
constexpr uint32_t GF_UNCLAMPED = 2u;

// Work item of one colour phase: vertex, first incidence, NH incidence count
// (bit 31 = conflicted: buffer the update), ground record (or -1).
struct alignas(16) VItem {
    uint32_t v;
    uint32_t inc_begin;
    uint32_t inc_count;
    int32_t ground;
};
constexpr uint32_t VITEM_CONFLICT = 0x80000000u;

// Cosserat rod element (CosseratRodEnergyInfo, energy/elastic/CosseratRodEnergyPool.hpp:8-14):
// nodes, rest length, stiffness diag(GA, GA, EA, EI, EI, GJ), rest curvature.
struct alignas(16) CosRec {
    uint32_t i0, i1;
    float L, k0;
    float k1, k2, k3, k4;
    float k5, kx, ky, kz;
};
static_assert(sizeof(CosRec) == 48, "CosRec must be 48 bytes");

// Rigid-body ground contact (RigidBodyGroundCollisionEnergyInfo,
// energy/collision/RigidBodyGroundCollisionEnergyPool.hpp:8-15).  Normal and
// tangents are the world axes y, x, z.  shape: the reference's
// CollisionGeometryType (collision/CollisionGeometryType.hpp).
struct alignas(16) RGroundRec {
    float k_start, k_n, k_t, k_b;
    float lambda_n, lambda_t, lambda_b, C_n_prev;
    float mu_s, mu_k, radius, cpg_x;
    float cpg_y, cpg_z, cpl_x, cpl_y;
    float cpl_z;
    uint32_t p, shape, flags;   // flags: GF_USE_STATIC | GF_UNCLAMPED
};
static_assert(sizeof(RGroundRec) == 80, "RGroundRec must be 80 bytes");
constexpr uint32_t SHAPE_ROD_SEGMENT = 1u;
constexpr uint32_t SHAPE_SPHERE = 2u;

// One-sided fixed joint (OneSidedFixedJointEnergyInfo,
// energy/joint/OneSidedFixedJointEnergyPool.hpp:8-16) with its vector AL state.
// A two-sided fixed joint (new) uses the same record: p2 is its second body,
// and ref_pos / ref_rot hold that body's position and rotation offsets
// instead of a fixed world pose.
struct alignas(16) JointRec {
    float4 pos_off;   // xyz
    float4 rot_off;   // quaternion (x, y, z, w)
    float4 ref_pos;   // xyz (two-sided: body p2's position offset)
    float4 ref_rot;   // quaternion (x, y, z, w) (two-sided: body p2's rotation offset)
    float k[6], lambda[6], C_prev[6];
    uint32_t p, p2;   // p2 = JOINT_ONE_SIDED for the one-sided joint
};
static_assert(sizeof(JointRec) == 144, "JointRec must be 144 bytes");
constexpr float kJointKStart = 1e4f;   // OneSidedFixedJointEnergyPool k_start
constexpr uint32_t JOINT_ONE_SIDED = 0xFFFFFFFFu;
// oriented incidence entry of a joint: the joint record, with this bit set
// when the particle is the second body of a two-sided joint
constexpr uint32_t OINC_JOINT_SIDE1 = 0x80000000u;

// Work item of an oriented particle: vertex, rotation slot, first entry in
// the oriented incidence list, counts of Cosserat incidences ((e << 1) | local,
// bit 31 = conflicted), rigid-ground records and joints that follow it, and
// the positional ground record (or -1).
struct alignas(16) OItem {
    uint32_t v, rot, inc_begin, n_cos;
    uint32_t n_rg, n_joint;
    int32_t ground;
    uint32_t pad;
};
constexpr uint32_t PREC_JOINT = 0x80000000u;   // per-rotation record list: joint flag

// Dynamic triangle-rigid / triangle-rod AL contact (CollisionConstraintEnergyInfo
// + TriangleRigidCollisionEnergyInfo / TriangleRodCollisionEnergyInfo,
// energy/collision/CollisionConstraintEnergyPool.hpp:9-30,
// energy/collision/TriangleRigidCollisionEnergyPool.hpp:8-14,
// energy/collision/TriangleRodCollisionEnergyPool.hpp:8-15).  idx: the
// triangle's vertices, then the rigid body (kind 0) or the two rod nodes
// (kind 1); kind 2 (rigid-rigid, new) holds the two bodies in idx[0..1] and
// their body-frame contact points in cpl and bary; kind 3 (vertex-triangle,
// new) holds the triangle's vertices and then the vertex, all positional.
// Record index = kind * capacity + pool slot.
struct alignas(16) DynRec {
    float4 n;      // normal, w = k_start
    float4 t;      // tangent, w = mu_s
    float4 b;      // binormal, w = mu_k
    float4 bary;   // barycentrics on the triangle, w = s (rod parameter); rigid-rigid: the second body's contact point
    float4 cpl;    // contact point in the body / interpolated rod frame, w = C_n_prev
    float4 k;      // k_n, k_t, k_b, unused
    float4 lam;    // lambda_n, lambda_t, lambda_b, unused
    uint32_t idx[5];
    uint32_t kind_flags;   // kind in bits 0-7, GF_USE_STATIC / GF_UNCLAMPED in bits 8-9
    uint32_t pad[2];
};
static_assert(sizeof(DynRec) == 144, "DynRec must be 144 bytes");
constexpr uint32_t DYN_TRI_RIGID = 0u, DYN_TRI_ROD = 1u, DYN_RIGID_RIGID = 2u, DYN_VERTEX_TRI = 3u, DYN_KINDS = 4u;
// particles per record of each kind: triangle + body, triangle + two rod
// nodes, two bodies, triangle + vertex
__host__ __device__ constexpr uint32_t dyn_particles(uint32_t kind) {
    return kind == 1u ? 5u : (kind == 2u ? 2u : 4u);
}
constexpr uint32_t DYN_FLAG_SHIFT = 8u;
constexpr float kContactMuS = 0.4f, kContactMuK = 0.2f;   // _addCollision (collision/CollisionDetector.cpp:365-366)

// Host -> device contact update of one detection round: a new record (ADD)
// or new normal / contact point of a persisting one (UPDATE, which rotates
// the friction frame on the device, CollisionConstraintEnergySolver::updateCollisionFrame).
struct alignas(16) DynOp {
    uint32_t rec, op, kind, pad;
    uint32_t idx[5];
    float k_start;
    float n[3], t[3], b[3], bary[3], s, cpl[3];
};
constexpr uint32_t DYN_OP_ADD = 0u, DYN_OP_UPDATE = 1u;

// Narrow-phase output (device -> host): DetectedCollision fields of one pair.
struct alignas(8) DetRec {
    uint64_t key;
    uint32_t kind;       // DetectedCollisionType: 0 vertex-face, 2 triangle-rigid, 3 triangle-rod, 6 rigid-rigid
    uint32_t prim_a, prim_b, pad;   // vertex-face: face, the vertex's owner, the vertex's corner in it
    double n[3], bary[3], cpl[3], s;
};
static_assert(sizeof(DetRec) == 104, "DetRec must be 104 bytes");

// ---------------------------------------------------------------------------
// Kernel launch wrappers (defined in vbd_kernels.hip)
// ---------------------------------------------------------------------------
struct DevState {
    float4* x = nullptr;        // xyz + mass in w
    float4* x_buf = nullptr;
    float4* y = nullptr;
    float4* x_prev = nullptr;
    float4* x_last = nullptr;
    float4* x_lastlast = nullptr;
    float4* v = nullptr;
    float4* v_prev = nullptr;
    float4* x_scratch = nullptr;   // autotuner output sink
    uint8_t* fixed = nullptr;
    int32_t* ground_of = nullptr;  // ground record per particle or -1
    TetRec* tets = nullptr;
    TetDamp* damp = nullptr;
    GroundRec* ground = nullptr;
    uint32_t* inc = nullptr;       // (tet << 2) | local, grouped per work item
    TetRec* irec = nullptr;        // copy of the tet record per incidence, in inc order,
                                   // local vertex index in the top two bits of idx[3]
    VItem* items = nullptr;        // all colours back to back
    double* reduce = nullptr;      // residual / IP partial sums
    uint32_t n = 0, n_tets = 0, n_ground = 0, n_inc = 0, n_items = 0;

    // oriented particles (RotationPool through rotation_idx, common/ParticlePool.hpp:7-69)
    int32_t* rot_of = nullptr;     // rotation slot per particle or -1
    float4* q = nullptr;           // orientation (x, y, z, w)
    float4* q_buf = nullptr;
    float4* q_in = nullptr;        // inertial target orientation
    float4* q_prev = nullptr;
    float4* q_last = nullptr;
    float4* q_lastlast = nullptr;
    float4* w = nullptr;           // angular velocity (body frame)
    float4* w_prev = nullptr;
    float4* inertia = nullptr;     // principal rotational inertia
    CosRec* cos = nullptr;
    RGroundRec* rground = nullptr;
    JointRec* joints = nullptr;
    OItem* oitems = nullptr;
    uint32_t* oinc = nullptr;
    uint32_t* prec_off = nullptr;  // per rotation slot: rigid-ground / joint records
    uint32_t* prec = nullptr;
    uint32_t n_rot = 0, n_cos = 0, n_rground = 0, n_joints = 0, n_oitems = 0;
    uint32_t n_joints1 = 0;   // joints [0, n_joints1) are one-sided, the rest two-sided

    // dynamic contacts (null unless enabled): per-particle CSR of incident
    // records, entries (record << 3) | local, in the reference's dynamic
    // adjacency order (triangle-rigid pool, then triangle-rod, ascending slot)
    const uint32_t* dyn_off = nullptr;
    const uint32_t* dyn_ent = nullptr;
    DynRec* dyn = nullptr;

    // bit 0: a non-finite position or velocity was produced (set by the
    // velocity update, read and cleared by the host)
    uint32_t* status = nullptr;

    // residual mode over several scenes: per-scene convergence flags (null:
    // no freezing) and the scene of every particle
    const uint32_t* frozen = nullptr;
    const uint32_t* scene_of = nullptr;

    // per-particle external acceleration added to gravity (xyz), or null
    // when the scene has no such field (avbd_set_external_accel)
    const float4* a_ext = nullptr;
};

// The particle state in FP64 (AVBD_FLAG_FP64_STATE, vbd_fp64.hip): double4
// streams (xyz + mass in w) beside the FP32 ones of DevState, which then
// hold a rounded mirror of the positions at step boundaries.
struct F64State {
    double4* x = nullptr;
    double4* y = nullptr;
    double4* x_prev = nullptr;
    double4* x_last = nullptr;
    double4* x_lastlast = nullptr;
    double4* x_alt = nullptr;      // second position buffer of the fused Chebyshev phase
    double4* x0 = nullptr;         // warm start: the initial guess of the step
    double4* x_scratch = nullptr;  // autotuner output sink
    float4* x32 = nullptr;         // the FP32 mirror (= DevState::x)
    float4* y32 = nullptr;         // FP32 mirrors of y and x_prev (= DevState::y, ::x_prev), read by
    float4* x_prev32 = nullptr;    // the warm-start features
    float4* v = nullptr;
    float4* v_prev = nullptr;
    const uint8_t* fixed = nullptr;
    const int32_t* ground_of = nullptr;
    const TetRec* tets = nullptr;
    TetDamp* damp = nullptr;
    GroundRec* ground = nullptr;
    const uint32_t* inc = nullptr;
    const TetRec* irec = nullptr;
    const VItem* items = nullptr;
    const float4* a_ext = nullptr;  // = DevState::a_ext
    const uint32_t* frozen = nullptr;    // = DevState::frozen
    const uint32_t* scene_of = nullptr;  // = DevState::scene_of
    // dynamic contacts (null unless enabled; = DevState::dyn_off / dyn_ent /
    // dyn): with the FP64 state every particle is positional, so the only
    // kind is the vertex-triangle contact; x_buf: the buffered positions of
    // conflicted particles
    const uint32_t* dyn_off = nullptr;
    const uint32_t* dyn_ent = nullptr;
    DynRec* dyn = nullptr;
    double4* x_buf = nullptr;
    uint32_t n = 0, n_tets = 0, n_ground = 0, n_items = 0;
};

struct LaunchCfg {
    uint32_t lanes = 8;
    uint32_t block = 256;
    uint32_t batch = 1;   // incidences a lane gathers at once (1, 2 or 4)
};

// Residual-targeted iteration: a step iterates in check intervals of
// `every` iterations; after each interval the force residual is reduced per
// scene and the step ends once every scene has ||r_s|| <= tol ||M_s g||.
// Per check k of the current step: rel[k] = ||r|| / ||M g|| over the whole
// scene set, worst[k] = the largest per-scene ratio (the first
// kMaxResidualChecks checks are kept).
constexpr uint32_t kMaxResidualChecks = 1024;
struct ResCtl {
    double rel[kMaxResidualChecks];
    double worst[kMaxResidualChecks];
    uint32_t done;         // set when the current step has converged
    uint32_t last_iters;   // iterations used by the last step
    uint32_t checks;       // checks run in the current step
    uint32_t iter;         // iterations run in the current step (Chebyshev weight index)
    unsigned long long total_iters;
    unsigned long long budget_steps;   // steps that ended at max_iters unconverged
    unsigned long long vertex_solves;  // vertex-block solves run (frozen scenes skip theirs)
};

// Per-scene freezing of a batch in residual mode: a scene that has reached
// the target stops iterating (its solves, Chebyshev steps and dual updates
// return at once) while the others go on, as independent runs would.
struct SceneFreeze {
    uint32_t* frozen = nullptr;          // per scene: 1 once converged in the current step
    uint32_t* iters = nullptr;           // per scene: iterations it ran in the current step
    const uint32_t* items = nullptr;     // per scene: free work items (vertex-block solves per iteration)
    const uint32_t* scene_of = nullptr;  // per particle
};

void launch_eprev(hipStream_t s, const DevState& d, float dt);
void launch_solve_ocolor(hipStream_t s, const DevState& d, uint32_t item_begin, uint32_t item_count, float dt,
                         const uint32_t* done = nullptr);
void launch_step_begin(hipStream_t s, const DevState& d, float dt, float g);
void launch_solve_color(hipStream_t s, const DevState& d, const LaunchCfg& cfg, uint32_t item_begin,
                        uint32_t item_count, bool damping, float dt, float4* x_out, const uint32_t* done = nullptr);
void launch_commit(hipStream_t s, const DevState& d, const uint32_t* list, uint32_t count,
                   const uint32_t* done = nullptr);
// rc / omega_tab / j: in residual mode the launch takes the Chebyshev weight
// omega_tab[rc->iter + j] and returns at once once the step has converged;
// with rc null it uses `omega`.
void launch_cheb_dual(hipStream_t s, const DevState& d, float omega, const ResCtl* rc = nullptr,
                      const float* omega_tab = nullptr, uint32_t j = 0);
// residual-targeted iteration (per-phase path): reset at step start, one
// last colour phase of an iteration fused with the Chebyshev / dual update
// (no oriented particles, no conflicts): reads d.x, writes the extrapolated
// positions of every particle to x_next (cheb_list: the particles that are
// not work items of this colour)
void launch_solve_cheb(hipStream_t s, const DevState& d, const LaunchCfg& cfg, uint32_t begin, uint32_t count,
                       bool damping, float dt, float omega, const ResCtl* rc, const float* omega_tab, uint32_t j,
                       const uint32_t* cheb_list, uint32_t n_cheb, float4* x_next);
// end of a double-buffered step: x = x_alt when the iteration count is odd
void launch_settle(hipStream_t s, float4* x, const float4* x_alt, uint32_t n, const uint32_t* iters_word,
                   uint32_t iters);
// residual checks: reset at step start; the pass writes |r_v|^2 per particle
// into rr (FP32 state; launch_residual_f64 for the FP64 state); the decision
// sums rr per scene and sets rc->done once every scene is within tol2
// (relative, squared); the bookkeeping at step end
void launch_residual_pass(hipStream_t s, const DevState& d, const LaunchCfg& cfg, bool damping, float dt,
                          const ResCtl* rc, double* rr);
void launch_residual_decide(hipStream_t s, ResCtl* rc, const double* rr, const uint32_t* scene_off,
                            uint32_t n_scenes, double* acc_scene, const double* mg2_scene, double tol2,
                            uint32_t every, const SceneFreeze& fz);
void launch_residual_reset(hipStream_t s, ResCtl* rc, const SceneFreeze& fz, uint32_t n_scenes);
void launch_residual_finish(hipStream_t s, ResCtl* rc, const SceneFreeze& fz, uint32_t n_scenes,
                            uint32_t n_items);
// learned warm start
void launch_warm_save(hipStream_t s, const DevState& d, const WarmIO& io);
void launch_warm_features(hipStream_t s, const DevState& d, const WarmIO& io, uint32_t lanes, float dt, float g,
                          uint32_t rows_pad);
void launch_warm_apply(hipStream_t s, const DevState& d, const WarmIO& io, float dt2g);
void launch_warm_ip(hipStream_t s, const DevState& d, const WarmIO& io, bool damping, float dt);
void launch_warm_choose(hipStream_t s, const DevState& d, const WarmIO& io);
void launch_warm_target(hipStream_t s, const DevState& d, const WarmIO& io, float inv_dt2g);
void launch_velocity(hipStream_t s, const DevState& d, float dt);
// dynamic contacts: apply a detection round's ADD / UPDATE ops; the
// start-of-step and per-iteration AL updates of the active records
void launch_dyn_apply(hipStream_t s, const DevState& d, const DynOp* ops, uint32_t n_ops);
void launch_dyn_step_begin(hipStream_t s, const DevState& d, const uint32_t* active, uint32_t n_active);
void launch_dyn_dual(hipStream_t s, const DevState& d, const uint32_t* active, uint32_t n_active,
                     const uint32_t* done = nullptr);
// two-sided fixed joints: start-of-step update (before launch_step_begin) and
// dual update (after the Chebyshev step); no-ops without such joints
void launch_joint2_step_begin(hipStream_t s, const DevState& d);
void launch_joint2_dual(hipStream_t s, const DevState& d, const uint32_t* done = nullptr);
// plain streaming copy of n_float4 16-byte words (HBM counter calibration)
void launch_calib_copy(hipStream_t s, const float4* src, float4* dst, uint64_t n_float4);
// residual/IP pass: writes 3 partial sums per block into d.reduce; returns block count
uint32_t launch_stats(hipStream_t s, const DevState& d, const LaunchCfg& cfg, bool damping, float dt, float g);

// FP64-state kernels (vbd_fp64.hip)
// FP64 state with dynamic contacts: the vertex-triangle records' start-of-step
// and per-iteration AL updates, and the commit of buffered positions
void launch_dyn_step_begin_f64(hipStream_t s, const F64State& f, const uint32_t* active, uint32_t n_active);
void launch_dyn_dual_f64(hipStream_t s, const F64State& f, const uint32_t* active, uint32_t n_active,
                         const uint32_t* done);
void launch_commit_f64(hipStream_t s, const F64State& f, const uint32_t* list, uint32_t count, const uint32_t* done);
void launch_solve_color_f64(hipStream_t s, const F64State& f, const LaunchCfg& cfg, uint32_t begin, uint32_t count,
                            bool damping, double dt, double4* x_out, const uint32_t* done);
void launch_solve_cheb_f64(hipStream_t s, const F64State& f, const LaunchCfg& cfg, uint32_t begin, uint32_t count,
                           bool damping, double dt, float omega, const ResCtl* rc, const float* omega_tab,
                           uint32_t j, const uint32_t* cheb_list, uint32_t n_cheb, const double4* x_cur,
                           double4* x_next);
void launch_cheb_dual_f64(hipStream_t s, const F64State& f, uint32_t n, float omega, const ResCtl* rc,
                          const float* omega_tab, uint32_t j);
void launch_step_begin_f64(hipStream_t s, const F64State& f, uint32_t n, double dt, double g);
void launch_eprev_f64(hipStream_t s, const F64State& f, uint32_t n_tets);
void launch_velocity_f64(hipStream_t s, const F64State& f, uint32_t n, double dt, uint32_t* status);
void launch_settle_f64(hipStream_t s, double4* x, const double4* x_alt, uint32_t n, const uint32_t* iters_word,
                       uint32_t iters);
void launch_residual_f64(hipStream_t s, const F64State& f, const LaunchCfg& cfg, bool damping, double dt,
                         const double4* x, double* rr, const ResCtl* rc);
// avbd_get_stats with the FP64 state: 4 sums per block into out; returns the block count
uint32_t launch_stats_f64(hipStream_t s, const F64State& f, const LaunchCfg& cfg, bool damping, double dt, double g,
                          double* out);
void launch_warm_save_f64(hipStream_t s, const F64State& f, uint32_t n);
void launch_warm_ip_f64(hipStream_t s, const F64State& f, const WarmIO& io, bool damping, double dt);
void launch_warm_place_f64(hipStream_t s, const F64State& f, const float4* pred, const uint32_t* scene_of,
                           const float* step_of_scene);
void launch_warm_target_f64(hipStream_t s, const F64State& f, float* tgt, double inv_s);

// ---------------------------------------------------------------------------
// Host-side graph build (host_graph.cpp)
// ---------------------------------------------------------------------------
struct HostNH {
    uint32_t idx[4];
    double lambda, mu, kd;
    double Q[9];
    double V;
    double E_prev[9];
};
struct HostGround {
    uint32_t p;
    double k_start, k_n, k_t, k_b, lambda_n, lambda_t, lambda_b, C_n_prev, mu_s, mu_k, cp_x, cp_z;
    bool use_static, unclamped;
};

struct HostCos {
    uint32_t idx[2];
    double stiffness[6];
    double L;
    double kappa0[3];
};
struct HostRGround {
    uint32_t p, shape;
    double radius;
    double cp_rb_local[3], cp_ground[3];
    double k_start, k_n, k_t, k_b, lambda_n, lambda_t, lambda_b, C_n_prev, mu_s, mu_k;
    bool use_static, unclamped;
};
struct HostJoint {
    uint32_t p;
    double pos_offset[3], rot_offset[4];   // quaternions (w, x, y, z)
    double ref_position[3], ref_rotation[4];
    double k[6], lambda[6], C_prev[6];
};
// Two-sided fixed joint between oriented particles p[0] and p[1] (new).
struct HostJoint2 {
    uint32_t p[2];
    double pos_offset[2][3], rot_offset[2][4];   // quaternions (w, x, y, z)
    double k[6], lambda[6], C_prev[6];
};

// Everything the colour plan is built from.
struct HostModel {
    uint32_t n = 0;
    const std::vector<HostNH>* nh = nullptr;
    const std::vector<HostGround>* ground = nullptr;
    const std::vector<HostCos>* cos = nullptr;
    const std::vector<HostRGround>* rground = nullptr;
    const std::vector<HostJoint>* joints = nullptr;
    const std::vector<HostJoint2>* joints2 = nullptr;
    const std::vector<uint8_t>* fixed = nullptr;
    const std::vector<int32_t>* rot_of = nullptr;
    const std::vector<uint32_t>* scene_offsets = nullptr;
};

struct ColorPlan {
    uint32_t num_colors = 0;
    std::vector<uint32_t> color;              // per particle
    std::vector<uint32_t> item_offsets;       // per colour: [begin, end) in items
    std::vector<VItem> items;
    std::vector<uint32_t> inc;
    std::vector<uint32_t> oitem_offsets;      // per colour: [begin, end) in oitems
    std::vector<OItem> oitems;
    std::vector<uint32_t> oinc;
    std::vector<uint32_t> prec_off;           // per rotation slot (n_rot + 1)
    std::vector<uint32_t> prec;
    std::vector<uint32_t> conflict_offsets;   // per colour
    std::vector<uint32_t> conflict_list;
    // per particle: its work item as laid out at build time (kind 0 none /
    // fixed, 1 positional, 2 oriented), so that a recolouring can regroup the
    // items without moving the incidence lists; and the static vertex
    // adjacency over all particles (deduplicated)
    std::vector<uint8_t> item_kind;
    std::vector<VItem> item_of;
    std::vector<OItem> oitem_of;
    std::vector<uint32_t> static_off, static_nb;
};

// Colours every scene of a global static adjacency CSR (scene offsets,
// vertex offsets, neighbours) into color, with the colour count per scene.
using SceneColourer = std::function<void(const std::vector<uint32_t>&, const std::vector<uint32_t>&,
                                         const std::vector<uint32_t>&, std::vector<uint32_t>&,
                                         std::vector<uint32_t>&)>;

// Build adjacency and colour every scene with DSATUR (with `colour` when it
// is set -- the device colourer -- else on the host); then lay out the
// per-colour work items.  Throws std::runtime_error on failure.
void build_color_plan(const HostModel& m, ColorPlan& plan, const SceneColourer& colour = nullptr);

// DSATUR of every scene on the device (graph_kernels.hip): one workgroup
// per scene; key and seen are per-vertex scratch (seen zeroed, 4 words per
// vertex); status bit 0 = more than 256 colours, bit 1 = a degree > 65535.
void launch_dsatur(hipStream_t s, uint32_t n_scenes, const uint32_t* scene_off, const uint32_t* off,
                   const uint32_t* nb, uint64_t* key, uint64_t* seen, uint32_t* color, uint32_t* scene_colors,
                   uint32_t* status);


// Message of the last failure of an entry point that takes no context
// (returned by avbd_last_error(NULL)).
void set_context_free_error(const std::string& msg);

// DSATUR on one CSR graph (ColorList::greedyColor2, common/ColorList.hpp:149-241).
uint32_t dsatur_csr(uint32_t n, const uint32_t* p_offsets, const uint32_t* p_neighbors, uint32_t* color);

}  // namespace avbd
