// rigid.hpp -- FP64 oracle: quaternions, S3/SO3 maps and the 6-DOF energies
// (Cosserat rod, rigid-body ground contact, one-sided and two-sided fixed joints).
// Test infrastructure only; see oracle.hpp.
//
// Restates:
//   common/Math.hpp:13-47,134-246        Skew3, J_r, J_r^-1, Exp_so3, Exp_s3, Log_S3, Plus_S3, Minus_S3
//   energy/elastic/CosseratRodEnergySolver.hpp:20-159
//   energy/collision/RigidBodyGroundCollisionEnergySolver.hpp:14-183
//   energy/joint/OneSidedFixedJointEnergySolver.hpp:16-90
//   energy/HardConstraintEnergySolver.hpp:17-155 (vector AL)
#pragma once

#include <cmath>

#include "oracle.hpp"

namespace oracle {

// Hamilton quaternion (Eigen's convention: q = w + xi + yj + zk).
struct Quat {
    Real w = 1, x = 0, y = 0, z = 0;
    Quat() = default;
    Quat(Real w_, Real x_, Real y_, Real z_) : w(w_), x(x_), y(y_), z(z_) {}
};
inline Quat qmul(const Quat& a, const Quat& b) {
    return {a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
            a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x, a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
}
inline Quat qconj(const Quat& q) { return {q.w, -q.x, -q.y, -q.z}; }
M3 qmat(const Quat& q);                       // rotation matrix of a unit quaternion
V3 qrot(const Quat& q, V3 v);                 // q * v (rotation)

M3 skew3(V3 v);
M3 exp_so3(V3 v);                             // Math::Exp_so3
M3 jr(V3 theta);                              // Math::ExpMap_RightJacobian
M3 jr_inv(V3 theta);                          // Math::ExpMap_InvRightJacobian
Quat exp_s3(V3 v);                            // Math::Exp_s3
V3 log_s3(const Quat& q);                     // Math::Log_S3
inline Quat plus_s3(const Quat& q, V3 v) { return qmul(q, exp_s3(v)); }
inline V3 minus_s3(const Quat& q1, const Quat& q2) { return log_s3(qmul(qconj(q2), q1)); }

// 6-vector / symmetric-agnostic dense 6x6 (row-major) helpers
struct V6 {
    Real v[6] = {0, 0, 0, 0, 0, 0};
};
struct M6 {
    Real m[6][6] = {};
};

// CosseratRodEnergyInfo, energy/elastic/CosseratRodEnergyPool.hpp:8-14
struct CosseratRod {
    uint32_t idx[2];
    Real stiffness[6];
    Real L;
    V3 kappa0;
};

// Shapes of the rigid-ground contact (collision/CollisionGeometryType.hpp)
enum ShapeType : uint32_t { SHAPE_SPHERE = 0, SHAPE_ROD = 1, SHAPE_OTHER = 2 };

// RigidBodyGroundCollisionEnergyInfo, energy/collision/RigidBodyGroundCollisionEnergyPool.hpp:8-15
struct RigidGround {
    uint32_t p;
    uint32_t shape;
    Real radius;
    V3 cp_rb_local, cp_ground;
    V3 n{0, 1, 0}, t{1, 0, 0}, b{0, 0, 1};
    Real k_start, k_n, k_t, k_b;
    Real lambda_n, lambda_t, lambda_b;
    Real C_n_prev;
    bool use_static, unclamped_lambda_tb;
    Real mu_s, mu_k;
};

// OneSidedFixedJointEnergyInfo, energy/joint/OneSidedFixedJointEnergyPool.hpp:8-16
struct Joint {
    uint32_t p;
    V3 pos_offset;
    Quat rot_offset;
    V3 ref_position;
    Quat ref_rotation;
    Real k[6], lambda[6], C_prev[6];
};
constexpr Real JOINT_K_START = 1e4;   // OneSidedFixedJointEnergyPool ctor

// Two-sided fixed joint (new: the reference's joint holds one particle to a
// fixed world pose, energy/joint/OneSidedFixedJointEnergySolver.hpp:16-34).
// Joint frame i = (x_i + R_i pos_offset[i], q_i rot_offset[i]) on body
// p[i]; C = [frame_0 position - frame_1 position ; Minus_S3(frame_0
// rotation, frame_1 rotation)] under the same vector AL, so that a fixed body
// 1 reproduces the one-sided joint with ref = frame 1.
struct Joint2 {
    uint32_t p[2];
    V3 pos_offset[2];
    Quat rot_offset[2];
    Real k[6], lambda[6], C_prev[6];
};

// Kernels (positions x, rotations by particle through rot_of)
struct RotState;
void cosserat_accumulate(const CosseratRod& e, const std::vector<V3>& x, const RotState& r, uint32_t local, V6& G,
                         M6& H);
Real cosserat_energy(const CosseratRod& e, const std::vector<V3>& x, const RotState& r);

void rground_eval(const RigidGround& c, const std::vector<V3>& x, const RotState& r, Real& Cn, Real& Ct, Real& Cb);
void rground_accumulate(const RigidGround& c, const std::vector<V3>& x, const RotState& r, V6& G, M6& H);
void rground_update_contact_points(RigidGround& c, const std::vector<V3>& x, const RotState& r);
void rground_update_after_timestep(RigidGround& c, const std::vector<V3>& x, const RotState& r);
void rground_update_after_iteration(RigidGround& c, const std::vector<V3>& x, const RotState& r);

void joint_eval(const Joint& j, const std::vector<V3>& x, const RotState& r, Real C[6]);
void joint_accumulate(const Joint& j, const std::vector<V3>& x, const RotState& r, V6& G, M6& H);
void joint_update_after_timestep(Joint& j, const std::vector<V3>& x, const RotState& r);
void joint_update_after_iteration(Joint& j, const std::vector<V3>& x, const RotState& r);

void joint2_eval(const Joint2& j, const std::vector<V3>& x, const RotState& r, Real C[6]);
// gradient rows of C with respect to body `side`'s (x, theta), theta the
// body-frame increment of q [+] theta
void joint2_grad(const Joint2& j, uint32_t side, const std::vector<V3>& x, const RotState& r, Real grad[6][6]);
void joint2_accumulate(const Joint2& j, uint32_t side, const std::vector<V3>& x, const RotState& r, V6& G, M6& H);
void joint2_update_after_timestep(Joint2& j, const std::vector<V3>& x, const RotState& r);
void joint2_update_after_iteration(Joint2& j, const std::vector<V3>& x, const RotState& r);

// RotationPool (common/ParticlePool.hpp:7-69), indexed by particle via rot_of
struct RotState {
    std::vector<int32_t> rot_of;   // per particle, -1 = not oriented
    std::vector<Quat> q, q_buf, q_in, q_prev, q_last, q_lastlast;
    std::vector<V3> w, w_prev, inertia;
    const Quat& rot(uint32_t p) const { return q[rot_of[p]]; }
};

}  // namespace oracle
