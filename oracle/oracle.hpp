// oracle.hpp -- FP64 CPU restatement of the avbd per-step AVBD solve.
//
// TEST INFRASTRUCTURE ONLY.  This code is the checker for the HIP path: only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
// The product library (avbd_amd) never links, calls or falls back to it.
//
// It restates, in plain C++ (no Eigen), the behaviour of these reference files
// (paths relative to the reference repository root):
//   simulation/VBDSolver.hpp:48-566            step orchestration, vertex solve
//   simulation/SimulationExecutor.hpp:142-156  phase order
//   energy/elastic/NeoHookeanEnergySolver.hpp:20-227,330-573
//   energy/collision/CollisionConstraintEnergySolver.hpp:60-285
//   energy/collision/GroundCollisionEnergySolver.hpp:14-91
//   energy/collision/CollisionConstraintEnergyPool.hpp:48-69 (initial values)
//   common/StaticParticleAdjacency.hpp:42-133   CSR adjacency
//   common/ColorList.hpp:149-241,345-363        DSATUR + work list
//   common/WorkerThreadContext.hpp:48-59         SPMD range split
//   common/SpinBarrier.hpp:22-36                 phase barrier
//   simobject/TetMeshObject.cpp:46-96            setup formulas
//   simulation/Simulation.hpp:117-126            per-vertex ground energies
// Each function below cites the lines it follows.
#pragma once

#include <cstdint>
#include <cmath>
#include <vector>
#include <array>
#include <atomic>
#include <string>
#include <utility>

namespace oracle {

using Real = double;

// AL constants, common/common.hpp:13-15
constexpr Real STIFFNESS_BETA = 10;
constexpr Real STIFFNESS_GAMMA = 0.99;
constexpr Real CONSTRAINT_ALPHA = 0.95;

struct V3 {
    Real x = 0, y = 0, z = 0;
    V3() = default;
    V3(Real a, Real b, Real c) : x(a), y(b), z(c) {}
    Real& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    Real operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(Real s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(V3 a, Real s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3& operator+=(V3& a, V3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
inline Real dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline Real norm(V3 a) { return std::sqrt(dot(a, a)); }

// Row-major 3x3: m[r][c].
struct M3 {
    Real m[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    static M3 identity() { M3 r; r.m[0][0] = r.m[1][1] = r.m[2][2] = 1; return r; }
    V3 col(int c) const { return {m[0][c], m[1][c], m[2][c]}; }
    V3 row(int r) const { return {m[r][0], m[r][1], m[r][2]}; }
};
M3 matmul(const M3& a, const M3& b);
M3 transpose(const M3& a);
V3 matvec(const M3& a, V3 v);
Real det(const M3& a);
M3 inverse(const M3& a);
// Dense solve with partial pivoting (Eigen partialPivLu semantics,
// simulation/VBDSolver.hpp:472,498).  n <= 6, A row-major n*n.
void lu_solve(int n, const Real* A, const Real* b, Real* x);

// --------------------------------------------------------------------------
// Pools
// --------------------------------------------------------------------------

// energy/elastic/NeoHookeanEnergyPool.hpp:8-17
struct NHTet {
    uint32_t idx[4];
    Real lambda, mu, kd;
    M3 Q;
    Real V;
    M3 E_prev;
};

// energy/collision/CollisionConstraintEnergyPool.hpp:9-30 +
// energy/collision/GroundCollisionEnergyPool.hpp:8-13
struct GroundContact {
    uint32_t p;
    Real k_start, k_n, k_t, k_b;
    Real lambda_n, lambda_t, lambda_b;
    Real C_n_prev;
    bool use_static, unclamped_lambda_tb;
    Real mu_s, mu_k;
    Real cp_x, cp_z;
};

// Static energy types in reference order (common/common.hpp:167-172).
enum StaticType : uint8_t { ST_NEO_HOOKEAN = 0, ST_COSSERAT = 1, ST_GROUND = 2, ST_RIGID_GROUND = 3, ST_JOINT = 4, ST_COUNT = 5 };

struct AdjEntry {
    uint8_t type;
    uint32_t e_idx;
    uint16_t local;
};

// common/StaticParticleAdjacency.hpp:17-34
struct StaticAdjacency {
    std::vector<std::array<uint32_t, ST_COUNT>> e_offsets;   // N+1 (last row: [0] = total)
    std::vector<AdjEntry> e_entries;
    std::vector<uint32_t> p_offsets;                        // N+1
    std::vector<uint32_t> p_neighbors;
};

// common/ColorList.hpp:62-78
struct Coloring {
    uint32_t num_colors = 0;
    std::vector<uint32_t> color;
    std::vector<uint32_t> color_counts;
    std::vector<uint32_t> color_offsets;
    std::vector<uint32_t> work_list;
};

struct Params {
    Real dt = 1.0 / 60.0;
    Real g = 9.81;
    uint32_t iters = 20;
    Real rho = 0.95;   // iter-acceleration (Chebyshev)
    // residual-targeted iteration (new; 0 = fixed `iters`): iterate until
    // ||r|| <= tol ||M g||, checked every check_every iterations, at most max_iters
    Real tol = 0;
    uint32_t max_iters = 0, check_every = 0;
};

struct Stats {
    Real residual_abs = 0;   // ||r||_2 over free vertices
    Real residual_rel = 0;   // ||r||_2 / ||M g||_2
    Real ip = 0;             // incremental potential
};

extern "C" void oracle_lbvh_build(uint32_t n, const uint32_t* tris, const double* x, const double* v, double dt,
                                  double* aabb, uint64_t* morton, uint32_t* sorted, uint32_t* left,
                                  uint32_t* right, uint32_t* parent, double* node_box, uint32_t* root);
// LBVH over n given boxes (aabb[6n] is input here).
extern "C" void oracle_lbvh_build_boxes(uint32_t n, const double* aabb, uint64_t* morton, uint32_t* sorted,
                                        uint32_t* left, uint32_t* right, uint32_t* parent, double* node_box,
                                        uint32_t* root);

// LBVH over the collision triangles, rebuilt every step (lbvh.cpp).
struct BroadPhase {
    uint32_t n = 0, root = 0;
    std::vector<double> xs, vs, aabb, node_box;
    std::vector<uint64_t> morton;
    std::vector<uint32_t> sorted, left, right, parent, hi;
    void build(const std::vector<uint32_t>& tris, const std::vector<V3>& x, const std::vector<V3>& v, double dt);
    // the same over given primitive boxes (6 per primitive: min xyz, max xyz)
    void build_boxes(const std::vector<double>& boxes);
    uint64_t traverse_leaves(uint32_t lo, uint32_t hi_leaf) const;   // candidate count
    // every overlapping leaf pair (i < j, sorted-order leaf indices)
    void traverse_pairs(std::vector<std::pair<uint32_t, uint32_t>>& out) const;
};


// Energy kernels exposed for unit tests.
Real nh_energy(const NHTet& t, const std::vector<V3>& x, Real dt);
void nh_accumulate(const NHTet& t, const std::vector<V3>& x, uint32_t local, V3& G, M3& H, Real dt);
void nh_accumulate4(const NHTet* const t[4], const std::vector<V3>& x, const uint32_t local[4],
                    V3& G, M3& H, Real dt);
void nh_update_after_timestep(NHTet& t, const std::vector<V3>& x_prev);

void ground_eval(const GroundContact& c, const std::vector<V3>& x, Real& Cn, Real& Ct, Real& Cb);
void ground_accumulate(const GroundContact& c, const std::vector<V3>& x, V3& G, M3& H);
void ground_update_contact_points(GroundContact& c, const std::vector<V3>& x);
void ground_update_after_timestep(GroundContact& c, const std::vector<V3>& x);
void ground_update_after_iteration(GroundContact& c, const std::vector<V3>& x);

// Graph pieces exposed for unit tests.
struct CosseratRod;
struct RigidGround;
struct Joint;
struct Joint2;
// joint entries: the one-sided joints (e_idx = joint index, local 0), then
// both sides of every two-sided joint (e_idx = |joints| + index, local = side)
void build_static_adjacency(uint32_t n, const std::vector<NHTet>& nh, const std::vector<CosseratRod>& cos,
                            const std::vector<GroundContact>& ground, const std::vector<RigidGround>& rground,
                            const std::vector<Joint>& joints, const std::vector<Joint2>& joints2,
                            StaticAdjacency& adj);
void dsatur(const StaticAdjacency& adj, uint32_t n, Coloring& col);
void rebuild_work_list(uint32_t n, Coloring& col);

}  // namespace oracle
