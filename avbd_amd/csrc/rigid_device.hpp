// rigid_device.hpp -- device side of the oriented (6-DOF) particles: the
// Cosserat rod, rigid-body ground contact and one-sided fixed joint energies,
// their start-of-step and per-iteration AL updates, the rotation parts of the
// inertial / Chebyshev / velocity phases, and the 6x6 vertex block solve.
// Included by vbd_kernels.hip (one translation unit, so the persistent step
// kernel can inline every phase).
#pragma once

#include "dyn_device.hpp"
#include "internal.hpp"
#include "rot_math.hpp"

namespace avbd {
namespace rigid {

using namespace rot;

__device__ __forceinline__ float3 xyz(float4 v) { return make_float3(v.x, v.y, v.z); }

// ---------------------------------------------------------------------------
// Cosserat rod (CosseratRodEnergySolver, energy/elastic/CosseratRodEnergySolver.hpp:20-159)
// ---------------------------------------------------------------------------
struct CosStrain {
    float3 q_diff;   // q2 [-] q1
    float4 q_mid;    // q1 [+] q_diff / 2
    float e[6];      // shear/stretch, bend/twist strain
};

__device__ __forceinline__ CosStrain cos_strain(const CosRec& c, float3 p1, float3 p2, float4 q1, float4 q2) {
    CosStrain s;
    const float inv_L = 1.f / c.L;
    s.q_diff = minus_s3(q2, q1);
    s.q_mid = plus_s3(q1, scl(0.5f, s.q_diff));
    const float3 d = qrot(qconj(s.q_mid), sub(p2, p1));
    s.e[0] = inv_L * d.x;
    s.e[1] = inv_L * d.y;
    s.e[2] = inv_L * d.z - 1.f;
    s.e[3] = inv_L * s.q_diff.x - c.kx;
    s.e[4] = inv_L * s.q_diff.y - c.ky;
    s.e[5] = inv_L * s.q_diff.z - c.kz;
    return s;
}

__device__ __forceinline__ float cos_energy(const CosRec& c, float3 p1, float3 p2, float4 q1, float4 q2) {
    const CosStrain s = cos_strain(c, p1, p2, q1, q2);
    const float k[6] = {c.k0, c.k1, c.k2, c.k3, c.k4, c.k5};
    float acc = 0.f;
#pragma unroll
    for (int i = 0; i < 6; i++) acc += s.e[i] * k[i] * s.e[i];
    return 0.5f * c.L * acc;
}

// Gradient L J^T K e and Gauss-Newton Hessian L J^T K J for node `local`,
// with the strain Jacobian J = [[A, B], [0, C]] (3x3 blocks).
__device__ __forceinline__ void cos_accumulate(const CosRec& c, float3 p1, float3 p2, float4 q1, float4 q2,
                                               uint32_t local, Vec6& G, Sym6& H) {
    const CosStrain s = cos_strain(c, p1, p2, q1, q2);
    const float inv_L = 1.f / c.L;
    const M3f gam_inv = jr_inv(s.q_diff);
    const float3 theta = scl(0.5f, s.q_diff);
    const M3f R = qmat(s.q_mid);
    const M3f Rt = mtrans(R);
    const M3f S = skew(mvec(Rt, scl(inv_L, sub(p2, p1))));
    const M3f gam_theta = jr(theta);
    M3f A, B, C;
    if (local == 0) {
        const M3f git = mtrans(gam_inv);
        A = mscale(-inv_L, Rt);
        M3f inner = mtrans(exp_so3(theta));
        const M3f gd = mmul(gam_theta, mscale(-0.5f, git));
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) inner.m[i][j] += gd.m[i][j];
        B = mmul(S, inner);
        C = mscale(-inv_L, git);
    } else {
        A = mscale(inv_L, Rt);
        B = mmul(mmul(S, gam_theta), mscale(0.5f, gam_inv));
        C = mscale(inv_L, gam_inv);
    }
    const float k[6] = {c.k0, c.k1, c.k2, c.k3, c.k4, c.k5};
    float ke[6];
#pragma unroll
    for (int i = 0; i < 6; i++) ke[i] = k[i] * s.e[i];
    // G = L [A^T ke1 ; B^T ke1 + C^T ke2]
#pragma unroll
    for (int j = 0; j < 3; j++) {
        G.v[j] += c.L * (A.m[0][j] * ke[0] + A.m[1][j] * ke[1] + A.m[2][j] * ke[2]);
        G.v[3 + j] += c.L * (B.m[0][j] * ke[0] + B.m[1][j] * ke[1] + B.m[2][j] * ke[2] + C.m[0][j] * ke[3] +
                             C.m[1][j] * ke[4] + C.m[2][j] * ke[5]);
    }
    // H = L [[A^T K1 A, A^T K1 B], [., B^T K1 B + C^T K2 C]]
#pragma unroll
    for (int a = 0; a < 3; a++) {
#pragma unroll
        for (int b = a; b < 3; b++)
            H.a[s6(a, b)] += c.L * (A.m[0][a] * k[0] * A.m[0][b] + A.m[1][a] * k[1] * A.m[1][b] +
                                    A.m[2][a] * k[2] * A.m[2][b]);
#pragma unroll
        for (int b = 0; b < 3; b++)
            H.a[s6(a, 3 + b)] += c.L * (A.m[0][a] * k[0] * B.m[0][b] + A.m[1][a] * k[1] * B.m[1][b] +
                                        A.m[2][a] * k[2] * B.m[2][b]);
    }
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = a; b < 3; b++)
            H.a[s6(3 + a, 3 + b)] +=
                c.L * (B.m[0][a] * k[0] * B.m[0][b] + B.m[1][a] * k[1] * B.m[1][b] + B.m[2][a] * k[2] * B.m[2][b] +
                       C.m[0][a] * k[3] * C.m[0][b] + C.m[1][a] * k[4] * C.m[1][b] + C.m[2][a] * k[5] * C.m[2][b]);
}

// ---------------------------------------------------------------------------
// Rigid-body ground contact (RigidBodyGroundCollisionConstraintSolver,
// energy/collision/RigidBodyGroundCollisionEnergySolver.hpp:14-183).
// n = y, t = x, b = z.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool rg_lowest_point(const RGroundRec& c) {
    return c.shape == SHAPE_SPHERE || c.shape == SHAPE_ROD_SEGMENT;
}

__device__ __forceinline__ float3 rg_eval(const RGroundRec& c, float3 p, float4 q) {
    const float3 cpl = v3(c.cpl_x, c.cpl_y, c.cpl_z);
    const float3 cpg = v3(c.cpg_x, c.cpg_y, c.cpg_z);
    const float3 diff = sub(cpg, add(p, qrot(q, cpl)));
    const float Cn = rg_lowest_point(c) ? c.cpg_y - (p.y - c.radius) : diff.y;
    return v3(Cn, diff.x, diff.z);   // (C_n, C_t, C_b)
}

// sym(-[a]x [cp]x) = (a.cp) I - (cp a^T + a cp^T)/2, a = R^T v: the rotational
// Hessian of v . (-R(q Exp(theta)) cp) at theta = 0.
__device__ __forceinline__ void rg_hess_block(Sym6& H, float s, float3 a, float3 cp) {
    const float ac = dotv(a, cp);
    const float av[3] = {a.x, a.y, a.z}, cv[3] = {cp.x, cp.y, cp.z};
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = i; j < 3; j++)
            H.a[s6(3 + i, 3 + j)] += s * ((i == j ? ac : 0.f) - 0.5f * (cv[i] * av[j] + av[i] * cv[j]));
}

__device__ __forceinline__ void rg_accumulate(const RGroundRec& c, float3 p, float4 q, Vec6& G, Sym6& H) {
    const float3 C = rg_eval(c, p, q);
    const M3f R = qmat(q);
    const float3 cpl = v3(c.cpl_x, c.cpl_y, c.cpl_z);
    const M3f Rr = mmul(R, skew(cpl));
    // rotational gradient rows: (R [cp]x)^T v = row of R [cp]x picked by the axis v
    const float3 rt = v3(Rr.m[0][0], Rr.m[0][1], Rr.m[0][2]);
    const float3 rn = v3(Rr.m[1][0], Rr.m[1][1], Rr.m[1][2]);
    const float3 rb = v3(Rr.m[2][0], Rr.m[2][1], Rr.m[2][2]);
    const bool lp = rg_lowest_point(c);
    const float gn[6] = {0.f, -1.f, 0.f, lp ? 0.f : rn.x, lp ? 0.f : rn.y, lp ? 0.f : rn.z};
    const float gt[6] = {-1.f, 0.f, 0.f, rt.x, rt.y, rt.z};
    const float gb[6] = {0.f, 0.f, -1.f, rb.x, rb.y, rb.z};

    const float Ccorr = C.x - kAlpha * c.C_n_prev;
    const float an = c.k_n * Ccorr + c.lambda_n;
    const float lnp = fmaxf(0.f, an);
    const float at = c.k_t * C.y + c.lambda_t, ab = c.k_b * C.z + c.lambda_b;
    float lt = at, lb = ab;
    const float mag = sqrtf(lt * lt + lb * lb);
    const float lmax = ((c.flags & GF_USE_STATIC) ? c.mu_s : c.mu_k) * lnp;
    if (mag > lmax) {
        const float nrm = mag + 1e-8f;
        lt = lt / nrm * lmax;
        lb = lb / nrm * lmax;
    }
#pragma unroll
    for (int i = 0; i < 6; i++) G.v[i] += lnp * gn[i] + lt * gt[i] + lb * gb[i];
    // a = R^T v for v = x, z (and y for shapes without the lowest-point rule)
    const float3 a_t = v3(R.m[0][0], R.m[0][1], R.m[0][2]);
    const float3 a_b = v3(R.m[2][0], R.m[2][1], R.m[2][2]);
    rg_hess_block(H, at, a_t, cpl);
    rg_hess_block(H, ab, a_b, cpl);
    if (!lp) rg_hess_block(H, an, v3(R.m[1][0], R.m[1][1], R.m[1][2]), cpl);
    sym6_outer(H, gn, c.k_n);
    sym6_outer(H, gt, c.k_t);
    sym6_outer(H, gb, c.k_b);
}

// updateContactPoints + CollisionConstraintEnergySolver::updateAfterTimeStep
// (RigidBodyGroundCollisionEnergySolver.hpp:161-183, CollisionConstraintEnergySolver.hpp:141-167).
__device__ __forceinline__ void rg_step_begin(RGroundRec& c, float3 p, float4 q) {
    c.cpg_x = p.x;
    c.cpg_z = p.z;
    if (rg_lowest_point(c)) {
        const float3 l = qrot(qconj(q), v3(0.f, -c.radius, 0.f));
        c.cpl_x = l.x;
        c.cpl_y = l.y;
        c.cpl_z = l.z;
    }
    if (c.lambda_n > 0.f) c.C_n_prev = rg_eval(c, p, q).x;
    if ((c.flags & GF_UNCLAMPED) && !(c.flags & GF_USE_STATIC)) c.flags |= GF_USE_STATIC;
    c.lambda_n *= kAlpha * kGamma;
    c.lambda_t *= kGamma;
    c.lambda_b *= kGamma;
    c.k_n = fmaxf(kGamma * c.k_n, c.k_start);
    c.k_t = fmaxf(kGamma * c.k_t, c.k_start);
    c.k_b = fmaxf(kGamma * c.k_b, c.k_start);
}

// CollisionConstraintEnergySolver::updateAfterIteration (CollisionConstraintEnergySolver.hpp:60-134).
__device__ __forceinline__ void rg_dual(RGroundRec& c, float3 p, float4 q) {
    const float3 C = rg_eval(c, p, q);
    const float Ccorr = C.x - kAlpha * c.C_n_prev;
    const float lnp = c.k_n * Ccorr + c.lambda_n;
    const float ltp = c.k_t * C.y + c.lambda_t;
    const float lbp = c.k_b * C.z + c.lambda_b;
    if (lnp > 0.f) c.k_n += kBeta * fabsf(Ccorr);
    c.lambda_n = fmaxf(0.f, lnp);
    const float mag = sqrtf(ltp * ltp + lbp * lbp);
    float lmax = ((c.flags & GF_USE_STATIC) ? c.mu_s : c.mu_k) * c.lambda_n;
    if (mag < lmax) {
        c.k_t += kBeta * fabsf(C.y);
        c.k_b += kBeta * fabsf(C.z);
        c.lambda_t = ltp;
        c.lambda_b = lbp;
        c.flags |= GF_UNCLAMPED;
    } else {
        if (c.flags & GF_USE_STATIC) {
            c.flags &= ~GF_USE_STATIC;
            lmax = c.mu_k * lnp;
        }
        c.lambda_t = ltp / (mag + 1e-8f) * lmax;
        c.lambda_b = lbp / (mag + 1e-8f) * lmax;
        c.flags &= ~GF_UNCLAMPED;
    }
}

// ---------------------------------------------------------------------------
// One-sided fixed joint (OneSidedFixedJointConstraintSolver,
// energy/joint/OneSidedFixedJointEnergySolver.hpp:16-90) under the vector AL of
// HardConstraintEnergySolver (energy/HardConstraintEnergySolver.hpp:17-155);
// multiplier bounds are (-inf, inf), so no clamping and no stiffness rescale.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void joint_eval(const JointRec& j, float3 p, float4 q, float C[6]) {
    const float3 jp = add(p, qrot(q, xyz(j.pos_off)));
    const float3 th = minus_s3(qmul(q, j.rot_off), j.ref_rot);
    C[0] = jp.x - j.ref_pos.x;
    C[1] = jp.y - j.ref_pos.y;
    C[2] = jp.z - j.ref_pos.z;
    C[3] = th.x;
    C[4] = th.y;
    C[5] = th.z;
}

__device__ __forceinline__ void joint_accumulate(const JointRec& j, float3 p, float4 q, Vec6& G, Sym6& H) {
    float C[6];
    joint_eval(j, p, q, C);
    const M3f dCp = mscale(-1.f, mmul(qmat(q), skew(xyz(j.pos_off))));
    const M3f dCR = mmul(jr_inv(v3(C[3], C[4], C[5])), mtrans(qmat(j.rot_off)));
#pragma unroll
    for (int ci = 0; ci < 6; ci++) {
        float g[6];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            g[a] = (ci < 3 && a == ci) ? 1.f : 0.f;
            g[3 + a] = ci < 3 ? dCp.m[ci][a] : dCR.m[ci - 3][a];
        }
        const float lam = j.k[ci] * (C[ci] - kAlpha * j.C_prev[ci]) + j.lambda[ci];
#pragma unroll
        for (int a = 0; a < 6; a++) G.v[a] += g[a] * lam;
        sym6_outer(H, g, j.k[ci]);
    }
}

// HardConstraintEnergySolver::updateAfterTimeStep (HardConstraintEnergySolver.hpp:79-88)
__device__ __forceinline__ void joint_step_begin(JointRec& j, float3 p, float4 q) {
#pragma unroll
    for (int i = 0; i < 6; i++) {
        j.lambda[i] = kAlpha * kGamma * j.lambda[i];
        j.k[i] = fmaxf(kJointKStart, kGamma * j.k[i]);
    }
    joint_eval(j, p, q, j.C_prev);
}

// HardConstraintEnergySolver::updateAfterIteration (HardConstraintEnergySolver.hpp:42-72)
__device__ __forceinline__ void joint_dual(JointRec& j, float3 p, float4 q) {
    float C[6];
    joint_eval(j, p, q, C);
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const float Cc = C[i] - kAlpha * j.C_prev[i];
        const float lp = j.k[i] * Cc + j.lambda[i];
        if (lp > -3.402823466e38f && lp < 3.402823466e38f) j.k[i] += kBeta * Cc;
        j.lambda[i] = lp;
    }
}

// ---------------------------------------------------------------------------
// Two-sided fixed joint (new: the reference's joint holds one particle to a
// fixed world pose).  Frame i = (x_i + R_i o_i, q_i r_i) with body 0's
// offsets in pos_off / rot_off and body 1's in ref_pos / ref_rot;
// C = [frame_0 position - frame_1 position ; Minus_S3(frame_0 rotation,
// frame_1 rotation)], so a fixed body 1 is the one-sided joint.  Rows of the
// gradient: body 0 [I, -R_0 [o_0]x ; 0, J_r^-1(theta) R(r_0)^T], body 1
// [-I, R_1 [o_1]x ; 0, -J_r^-1(-theta) R(r_1)^T].  Same vector AL.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void joint2_eval(const JointRec& j, float3 p0, float4 q0, float3 p1, float4 q1,
                                            float C[6]) {
    const float3 a = add(p0, qrot(q0, xyz(j.pos_off)));
    const float3 b = add(p1, qrot(q1, xyz(j.ref_pos)));
    const float3 th = minus_s3(qmul(q0, j.rot_off), qmul(q1, j.ref_rot));
    C[0] = a.x - b.x;
    C[1] = a.y - b.y;
    C[2] = a.z - b.z;
    C[3] = th.x;
    C[4] = th.y;
    C[5] = th.z;
}

// G, H of body `side` (pos, q: that body; po, qo: the other one)
__device__ __forceinline__ void joint2_accumulate(const JointRec& j, uint32_t side, float3 pos, float4 q, float3 po,
                                                  float4 qo, Vec6& G, Sym6& H) {
    float C[6];
    if (side == 0) joint2_eval(j, pos, q, po, qo, C);
    else joint2_eval(j, po, qo, pos, q, C);
    const float sg = side == 0 ? 1.f : -1.f;
    const float3 off = side == 0 ? xyz(j.pos_off) : xyz(j.ref_pos);
    const float4 roff = side == 0 ? j.rot_off : j.ref_rot;
    const M3f dCp = mscale(-sg, mmul(qmat(q), skew(off)));
    const M3f dCR = mscale(sg, mmul(jr_inv(v3(sg * C[3], sg * C[4], sg * C[5])), mtrans(qmat(roff))));
#pragma unroll
    for (int ci = 0; ci < 6; ci++) {
        float g[6];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            g[a] = (ci < 3 && a == ci) ? sg : 0.f;
            g[3 + a] = ci < 3 ? dCp.m[ci][a] : dCR.m[ci - 3][a];
        }
        const float lam = j.k[ci] * (C[ci] - kAlpha * j.C_prev[ci]) + j.lambda[ci];
#pragma unroll
        for (int a = 0; a < 6; a++) G.v[a] += g[a] * lam;
        sym6_outer(H, g, j.k[ci]);
    }
}

// Body frames (x, q) of both bodies of a two-sided joint from the state.
__device__ __forceinline__ void joint2_bodies(const DevState& d, const JointRec& j, float3& p0, float4& q0,
                                              float3& p1, float4& q1) {
    p0 = xyz(d.x[j.p]);
    q0 = d.q[d.rot_of[j.p]];
    p1 = xyz(d.x[j.p2]);
    q1 = d.q[d.rot_of[j.p2]];
}

// HardConstraintEnergySolver::updateAfterTimeStep / updateAfterIteration
// (HardConstraintEnergySolver.hpp:42-88) with the two-sided constraint.
__device__ __forceinline__ void joint2_step_begin(const DevState& d, JointRec& j) {
#pragma unroll
    for (int i = 0; i < 6; i++) {
        j.lambda[i] = kAlpha * kGamma * j.lambda[i];
        j.k[i] = fmaxf(kJointKStart, kGamma * j.k[i]);
    }
    float3 p0, p1;
    float4 q0, q1;
    joint2_bodies(d, j, p0, q0, p1, q1);
    joint2_eval(j, p0, q0, p1, q1, j.C_prev);
}

__device__ __forceinline__ void joint2_dual(const DevState& d, JointRec& j) {
    float3 p0, p1;
    float4 q0, q1;
    joint2_bodies(d, j, p0, q0, p1, q1);
    float C[6];
    joint2_eval(j, p0, q0, p1, q1, C);
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const float Cc = C[i] - kAlpha * j.C_prev[i];
        const float lp = j.k[i] * Cc + j.lambda[i];
        if (lp > -3.402823466e38f && lp < 3.402823466e38f) j.k[i] += kBeta * Cc;
        j.lambda[i] = lp;
    }
}

// ---------------------------------------------------------------------------
// Per-particle phases of an oriented particle (rotation slot r)
// ---------------------------------------------------------------------------

// Start-of-step updates of the particle's rigid-ground and one-sided joint
// records (all read the start-of-step state, before the inertial update).
// Two-sided joints depend on two particles and have kernels of their own.
__device__ __forceinline__ void records_step_begin(const DevState& d, uint32_t p, uint32_t r) {
    const float3 pos = xyz(d.x[p]);
    const float4 q = d.q[r];
    for (uint32_t k = d.prec_off[r]; k < d.prec_off[r + 1]; k++) {
        const uint32_t e = d.prec[k];
        if (e & PREC_JOINT) {
            JointRec j = d.joints[e & ~PREC_JOINT];
            joint_step_begin(j, pos, q);
            d.joints[e & ~PREC_JOINT] = j;
        } else {
            RGroundRec c = d.rground[e];
            rg_step_begin(c, pos, q);
            d.rground[e] = c;
        }
    }
}

// Dual / stiffness update of the particle's rigid-ground and joint records.
__device__ __forceinline__ void records_dual(const DevState& d, float3 pos, float4 q, uint32_t r) {
    for (uint32_t k = d.prec_off[r]; k < d.prec_off[r + 1]; k++) {
        const uint32_t e = d.prec[k];
        if (e & PREC_JOINT) {
            JointRec j = d.joints[e & ~PREC_JOINT];
            joint_dual(j, pos, q);
            d.joints[e & ~PREC_JOINT] = j;
        } else {
            RGroundRec c = d.rground[e];
            rg_dual(c, pos, q);
            d.rground[e] = c;
        }
    }
}

// _particleRotationInertialUpdate (simulation/VBDSolver.hpp:253-283):
// q_in = q [+] (dt w - dt^2 I^-1 (w x I w)), q = q [+] dt w.
__device__ __forceinline__ void rotation_inertial(const DevState& d, uint32_t r, float dt) {
    const float3 w = xyz(d.w[r]);
    const float3 I = xyz(d.inertia[r]);
    const float3 c = crossv(w, v3(I.x * w.x, I.y * w.y, I.z * w.z));
    const float dt2 = dt * dt;
    const float3 dq_in = v3(dt * w.x - dt2 * c.x / I.x, dt * w.y - dt2 * c.y / I.y, dt * w.z - dt2 * c.z / I.z);
    const float4 q = d.q[r];
    d.q_in[r] = plus_s3(q, dq_in);
    const float4 qn = plus_s3(q, scl(dt, w));
    d.q[r] = qn;
    d.q_last[r] = qn;
    d.q_lastlast[r] = qn;
}

// Rotation part of _particleChebyshevAcceleration (VBDSolver.hpp:513-536).
__device__ __forceinline__ float4 rotation_chebyshev(const DevState& d, uint32_t r, float omega) {
    const float4 qll = d.q_lastlast[r];
    const float4 q = plus_s3(qll, scl(omega, minus_s3(d.q[r], qll)));
    d.q[r] = q;
    d.q_lastlast[r] = d.q_last[r];
    d.q_last[r] = q;
    return q;
}

// Rotation part of _particleVelocityUpdate (VBDSolver.hpp:538-566).
__device__ __forceinline__ void rotation_velocity(const DevState& d, uint32_t r, float inv_dt) {
    const float4 q = d.q[r];
    d.w_prev[r] = d.w[r];
    const float3 w = scl(inv_dt, minus_s3(q, d.q_prev[r]));
    d.w[r] = make_float4(w.x, w.y, w.z, 0.f);
    d.q_prev[r] = q;
}

// Gradient and Hessian of every static energy on oriented particle it.v, in
// the reference's StaticEnergyType order: Cosserat, ground (positional block),
// rigid-ground, joint (simulation/VBDSolver.hpp:297-397).
template <bool DYN = false>
__device__ __forceinline__ void accumulate6(const DevState& d, const OItem& it, float3 pos, float4 q, Vec6& G,
                                            Sym6& H) {
    const uint32_t ncos = it.n_cos & ~VITEM_CONFLICT;
    uint32_t k = it.inc_begin;
    for (uint32_t e = 0; e < ncos; e++, k++) {
        const uint32_t code = d.oinc[k];
        const CosRec c = d.cos[code >> 1];
        const uint32_t local = code & 1u;
        const uint32_t other = local ? c.i0 : c.i1;
        const float3 po = xyz(d.x[other]);
        const float4 qo = d.q[d.rot_of[other]];
        if (local == 0) cos_accumulate(c, pos, po, q, qo, 0, G, H);
        else cos_accumulate(c, po, pos, qo, q, 1, G, H);
    }
    if (it.ground >= 0) {
        // CollisionConstraintEnergySolver::accumulate<6> with the ground constraint: positional block only
        const GroundRec c = d.ground[it.ground];
        const float Cn = -pos.y, Ct = pos.x - c.cp_x, Cb = pos.z - c.cp_z;
        const float Ccorr = Cn - kAlpha * c.C_n_prev;
        const float lnp = fmaxf(0.0f, c.k_n * Ccorr + c.lambda_n);
        float lt = c.k_t * Ct + c.lambda_t, lb = c.k_b * Cb + c.lambda_b;
        const float mag = sqrtf(lt * lt + lb * lb);
        const float lmax = ((c.flags & GF_USE_STATIC) ? c.mu_s : c.mu_k) * lnp;
        if (mag > lmax) {
            const float nrm = mag + 1e-8f;
            lt = lt / nrm * lmax;
            lb = lb / nrm * lmax;
        }
        G.v[0] += lt;
        G.v[1] += -lnp;
        G.v[2] += lb;
        H.a[s6(0, 0)] += c.k_t;
        H.a[s6(1, 1)] += c.k_n;
        H.a[s6(2, 2)] += c.k_b;
    }
    for (uint32_t e = 0; e < it.n_rg; e++, k++) rg_accumulate(d.rground[d.oinc[k]], pos, q, G, H);
    for (uint32_t e = 0; e < it.n_joint; e++, k++) {
        const uint32_t code = d.oinc[k];
        const JointRec& j = d.joints[code & ~OINC_JOINT_SIDE1];
        if (j.p2 == JOINT_ONE_SIDED) {
            joint_accumulate(j, pos, q, G, H);
        } else {
            const uint32_t side = code >> 31;
            const uint32_t other = side ? j.p : j.p2;
            joint2_accumulate(j, side, pos, q, xyz(d.x[other]), d.q[d.rot_of[other]], G, H);
        }
    }
    // dynamic contacts after the static energies (simulation/VBDSolver.hpp:400-438)
    if constexpr (DYN) {
        const dyn::View view = dyn::view_of(d);
        const uint32_t e1 = d.dyn_off[it.v + 1];
        for (uint32_t e = d.dyn_off[it.v]; e < e1; e++) {
            const uint32_t code = d.dyn_ent[e];
            const DynRec c = d.dyn[code >> 3];
            dyn::accumulate6(view, c, code & 7u, G, H);
        }
    }
}

// _solveParticle<6> (simulation/VBDSolver.hpp:450-486): one thread per oriented vertex.
template <bool DYN = false>
__device__ __forceinline__ void solve_oitem(const DevState& d, const OItem& it, float inv_dt2) {
    const float4 xv = d.x[it.v];
    const float3 pos = xyz(xv);
    const float4 q = d.q[it.rot];
    Vec6 G;
    Sym6 H;
    vec6_zero(G);
    sym6_zero(H);
    accumulate6<DYN>(d, it, pos, q, G, H);
    const float md = xv.w * inv_dt2;
    const float3 I = xyz(d.inertia[it.rot]);
    const float3 yv = xyz(d.y[it.v]);
    const float3 th = minus_s3(q, d.q_in[it.rot]);
    float b[6] = {-md * (pos.x - yv.x) - G.v[0], -md * (pos.y - yv.y) - G.v[1], -md * (pos.z - yv.z) - G.v[2],
                  -(I.x * inv_dt2) * th.x - G.v[3], -(I.y * inv_dt2) * th.y - G.v[4],
                  -(I.z * inv_dt2) * th.z - G.v[5]};
    H.a[s6(0, 0)] += md;
    H.a[s6(1, 1)] += md;
    H.a[s6(2, 2)] += md;
    H.a[s6(3, 3)] += I.x * inv_dt2;
    H.a[s6(4, 4)] += I.y * inv_dt2;
    H.a[s6(5, 5)] += I.z * inv_dt2;
    float dx[6];
    lu6(H, b, dx);
    const float4 nx = make_float4(pos.x + dx[0], pos.y + dx[1], pos.z + dx[2], xv.w);
    const float4 nq = plus_s3(q, v3(dx[3], dx[4], dx[5]));
    if (it.n_cos & VITEM_CONFLICT) {
        d.x_buf[it.v] = nx;
        d.q_buf[it.rot] = nq;
    } else {
        d.x[it.v] = nx;
        d.q[it.rot] = nq;
    }
}

}  // namespace rigid
}  // namespace avbd
