// contact.hpp -- FP64 oracle: collision primitives, the triangle narrow
// phase, the contact lifecycle, the dynamic adjacency, incremental
// recolouring and the dynamic triangle-rigid / triangle-rod AL contacts.
// Test infrastructure only; see oracle.hpp.
//
// Restates (paths relative to the reference repository root):
//   collision/CollisionPrimitivePool.hpp:85-236, collision/CollisionPrimitivePool.cpp:36-87
//   collision/SDF.hpp:16-107                       sphere SDF
//   collision/CollisionDetector.cpp:132-215        _shouldSkip, _processPotentialCollision
//   collision/CollisionDetector.cpp:272-501        _handleDetectedCollisions, add / remove / update
//   collision/CollisionDetector.cpp:504-666        _triangleSphere, _triangleRod
//   collision/CollisionDetector.cpp:689-876        _triangleSDF_CCD
//   collision/DetectedCollision.hpp:43-66          generateKey
//   common/Algorithm.hpp:11-15,253-299             lerp, goldenSectionSearch
//   common/Math.hpp:323-327,421-681                basis, barycentric, closest points
//   common/TombstonePool.hpp:13-180                slot stack
//   common/DynamicParticleAdjacency.hpp:33-118
//   common/ColorList.hpp:366-522                   markParticleDirty, incrementalRecoloring, finalizeConflicts
//   energy/collision/CollisionConstraintEnergySolver.hpp:17-285
//   energy/collision/CollisionConstraintEnergyPool.hpp:48-69
//   energy/collision/TriangleRigidCollisionEnergySolver.hpp:14-117
//   energy/collision/TriangleRodCollisionEnergySolver.hpp:14-150
#pragma once

#include <array>
#include <utility>

#include "rigid.hpp"

namespace oracle {

// CollisionGeometryType (collision/CollisionGeometryType.hpp)
enum PrimType : uint8_t { PRIM_TRIANGLE = 0, PRIM_ROD_SEGMENT = 1, PRIM_SPHERE = 2 };

// One CollisionPrimitivePool slot.
struct Prim {
    uint8_t type = PRIM_TRIANGLE;
    uint8_t np = 3;             // particles: 3 triangle, 2 rod segment, 1 sphere (its centre)
    uint32_t idx[3] = {0, 0, 0};
    uint32_t object = 0;
    uint32_t gen = 1;           // generation of the slot (allocated once)
    Real radius = 0;            // rod segment / sphere
};

// Geometry (common/Math.hpp)
V3 unit_orthogonal(V3 n);                             // Eigen MatrixBase::unitOrthogonal, 3-vectors
void complete_basis(V3 n, V3& t, V3& b);              // Math::completeOrthonormalBasisGivenNormal
V3 barycentric(V3 p, V3 a, V3 b, V3 c);               // Math::barycentricCoordinates
V3 closest_point_triangle(V3 p, V3 a, V3 b, V3 c);    // Math::closestPoint_PointTriangle
void closest_segment_segment(V3 p1, V3 q1, V3 p2, V3 q2, Real& s, Real& t);
bool point_in_triangle(V3 p, V3 a, V3 b, V3 c);
void closest_segment_triangle(V3 p0, V3 p1, V3 a, V3 b, V3 c, Real& s, V3& barys, Real& dist);
// Algorithm::goldenSectionSearch (tol 1e-3) of (t - c)^2 over [a, b]: a
// known-answer hook for the tests.
Real golden_section_quadratic(Real a, Real b, Real c);

// DetectedCollision::generateKey: 5 type bits | 26 | 26 | 7 code bits.
uint64_t contact_key(uint32_t type, uint32_t a, uint32_t b, uint32_t code);

// CollisionPrimitivePool::speculativeGlobalBounds: box[6] = min xyz, max xyz;
// triangle boxes widened by tri_pad on every side (self collision, new).
void prim_bounds(const Prim& pr, const std::vector<V3>& x, const std::vector<V3>& v, const RotState& rs, Real dt,
                 double box[6], Real tri_pad = 0);

// TombstonePool: a stack of free slots (initially 0, 1, 2, ...), LIFO reuse.
struct SlotPool {
    uint32_t capacity = 0, count = 0, highest = 0, top = 0;
    std::vector<uint8_t> active;
    std::vector<uint32_t> gen;
    std::vector<uint32_t> slots;   // slots[i] for i < slots.size(); i beyond it
    uint32_t alloc();
    void free(uint32_t slot);
    uint32_t slot_at(uint32_t i) const { return i < slots.size() ? slots[i] : i; }
};

// DynamicEnergyType order (common/common.hpp) and DetectedCollisionType
// values; the rigid-rigid (sphere-sphere) contact is new: the reference
// declares its detection record (collision/DetectedCollision.hpp:33) but
// throws on sphere-sphere pairs (collision/CollisionDetector.cpp:678-686).
// The vertex-face contact between surface triangles is new as well: the
// reference declares TriangleTriangle_VertexFace (collision/DetectedCollision.hpp:7,29)
// but its triangle-triangle narrow phase is empty
// (collision/CollisionDetector.cpp:589-599) and adding one throws (:333-336).
enum DynType : uint8_t { DYN_TRI_RIGID = 0, DYN_TRI_ROD = 1, DYN_RIGID_RIGID = 2, DYN_VERTEX_TRI = 3, DYN_COUNT = 4 };
enum DetType : uint8_t { DET_VERTEX_FACE = 0, DET_TRI_RIGID = 2, DET_TRI_ROD = 3, DET_RIGID_RIGID = 6 };

// CollisionConstraintEnergyInfo + TriangleRigid / TriangleRod info.
// idx: triangle vertices, then the rigid body (tri-rigid) or both rod nodes;
// rigid-rigid: the two bodies (idx[0], idx[1]) with their body-frame contact
// points cpl and cpl2; vertex-triangle: the triangle's vertices, then the
// vertex (idx[3]).
struct DynContact {
    uint32_t idx[5] = {0, 0, 0, 0, 0};
    Real k_start = 0, k_n = 0, k_t = 0, k_b = 0;
    Real lambda_n = 0, lambda_t = 0, lambda_b = 0, C_n_prev = 0;
    V3 n, t, b;
    bool use_static = false, unclamped = false;
    Real mu_s = 0, mu_k = 0;
    V3 barys;
    Real s = 0;   // tri-rod: contact point along the segment
    V3 cpl;       // contact point in the body (tri-rigid) or interpolated rod frame
    V3 cpl2;      // rigid-rigid: contact point in the second body's frame
};

struct DynPool {
    SlotPool sp;
    std::vector<DynContact> data;
    uint32_t add(const DynContact& c);
};

// DetectedCollision (collision/DetectedCollision.hpp:20-35), the two kinds
// the reference creates.
struct Detected {
    uint8_t type = DET_TRI_RIGID;
    uint32_t e_idx = 0;
    uint64_t key = 0;
    uint32_t gen1 = 0, gen2 = 0;
    V3 normal;
    uint32_t tri[3] = {0, 0, 0};
    V3 barys;
    uint32_t rb = 0;                 // tri-rigid; rigid-rigid: the first body
    uint32_t rod[2] = {0, 0};        // tri-rod
    Real s = 0;
    V3 cpl;
    uint32_t rb2 = 0;                // rigid-rigid: the second body
    V3 cpl2;
    uint32_t vtx = 0;                // vertex-face: the vertex (tri: the face)
};

struct DynAdjEntry {
    uint8_t type;
    uint32_t e_idx;
    uint16_t local;
};
// DynamicParticleAdjacency (common/DynamicParticleAdjacency.hpp:16-31)
struct DynAdjacency {
    std::vector<std::array<uint32_t, DYN_COUNT>> e_offsets;   // N+1 (last row: [0] = total)
    std::vector<DynAdjEntry> e_entries;
    std::vector<uint32_t> p_offsets, p_neighbors;             // with duplicates
};

// Incremental-recolouring state of ColorList (common/ColorList.hpp:80-92).
struct Recoloring {
    static constexpr uint32_t SENTINEL = 0xFFFFFFFFu;
    std::vector<uint32_t> dirty, next_dirty, last_dirty_round, candidate, touched_this_frame, touched_conflicted;
    std::vector<uint32_t> conflict_offsets, conflict_entries;   // per colour CSR of conflicted particles
};

// Constraint value and gradients of one dynamic contact for one of its
// particles (constraintGradientHessian<DOF>).  g*: 3 (local < 3) or 6 entries;
// h*: the rotational 3x3 block (6-DOF, tri-rigid and rigid-rigid; zero
// otherwise).  Rigid-rigid: C = dir . (cp_2 - cp_1) for dir = n, t, b with
// cp_k = x_k + R_k cpl_k; local 0 is the first body (sign -1), local 1 the
// second (sign +1); both bodies carry the normal and the friction terms.
// Vertex-triangle: C = dir . (x_v - sum_i barys_i x_i); the triangle's
// vertices take -barys_i dir, the vertex (local 3) +dir, friction included.
void dyn_constraint(uint8_t type, const DynContact& c, const std::vector<V3>& x, const RotState& rs, uint32_t local,
                    int dof, Real& Cn, Real& Ct, Real& Cb, Real gn[6], Real gt[6], Real gb[6], M3& hn, M3& ht, M3& hb);
void dyn_eval(uint8_t type, const DynContact& c, const std::vector<V3>& x, const RotState& rs, Real& Cn, Real& Ct,
              Real& Cb);
void dyn_accumulate3(uint8_t type, const DynContact& c, const std::vector<V3>& x, const RotState& rs, uint32_t local,
                     V3& G, M3& H);
void dyn_accumulate6(uint8_t type, const DynContact& c, const std::vector<V3>& x, const RotState& rs, uint32_t local,
                     V6& G, M6& H);
void dyn_update_after_iteration(uint8_t type, DynContact& c, const std::vector<V3>& x, const RotState& rs);
void dyn_update_after_timestep(uint8_t type, DynContact& c, const std::vector<V3>& x, const RotState& rs);
void update_collision_frame(DynContact& c, V3 new_normal);

// CollisionDetector::_triangleSDF_CCD for a sphere SDF.  tri/tv: triangle
// vertices and velocities; rb_*: body position, orientation, velocity and
// body-frame angular velocity.
// Vertex-face detection (new): vertex p is in contact with triangle (a, b, c)
// when its distance to the triangle is below the margin.  The normal is the
// face normal f = unit((b - a) x (c - a)) turned towards the face as seen
// from the vertex (-f when f . (p - cp) >= 0, else f), so that
// C_n = n . (p - cp_tri) is negative while the vertex is apart and positive
// once it is through; barys: the closest point.  An existing contact keeps
// the side it was created with: its update turns a new normal that points
// against the current one around (a vertex that went through the face is
// pushed back, not further).
bool vertex_face(V3 p, V3 a, V3 b, V3 c, Real margin, V3& normal, V3& barys);
// Sphere-sphere detection (new): bodies at xa, xb with radii ra, rb are in
// contact when |xa - xb| < ra + rb + margin; n points from b to a, and the
// contact points are the surface points facing each other, in body frames.
bool sphere_sphere(V3 xa, const Quat& qa, Real ra, V3 xb, const Quat& qb, Real rb, Real margin, V3& normal, V3& cpl_a,
                   V3& cpl_b);

bool tri_sphere_ccd(const V3 tri[3], const V3 tv[3], V3 rb_pos, const Quat& rb_rot, V3 rb_v, V3 rb_w, Real radius,
                    Real dt, Real margin, V3& normal, V3& barys, V3& cpl);

}  // namespace oracle
