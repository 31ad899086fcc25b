// rigid.cpp -- FP64 oracle of the 6-DOF energies and S3 maps (test
// infrastructure only; see oracle.hpp and rigid.hpp for the cited sources).
#include "rigid.hpp"

#include <algorithm>

namespace oracle {

M3 qmat(const Quat& q) {
    const Real tx = 2 * q.x, ty = 2 * q.y, tz = 2 * q.z;
    const Real twx = tx * q.w, twy = ty * q.w, twz = tz * q.w;
    const Real txx = tx * q.x, txy = ty * q.x, txz = tz * q.x;
    const Real tyy = ty * q.y, tyz = tz * q.y, tzz = tz * q.z;
    M3 r;
    r.m[0][0] = 1 - (tyy + tzz);
    r.m[0][1] = txy - twz;
    r.m[0][2] = txz + twy;
    r.m[1][0] = txy + twz;
    r.m[1][1] = 1 - (txx + tzz);
    r.m[1][2] = tyz - twx;
    r.m[2][0] = txz - twy;
    r.m[2][1] = tyz + twx;
    r.m[2][2] = 1 - (txx + tyy);
    return r;
}

V3 qrot(const Quat& q, V3 v) {
    // v + 2w (u x v) + 2 u x (u x v), u = vector part
    const V3 u(q.x, q.y, q.z);
    V3 uv = cross(u, v);
    uv = uv + uv;
    return v + q.w * uv + cross(u, uv);
}

M3 skew3(V3 v) {
    M3 s;
    s.m[0][1] = -v.z;
    s.m[0][2] = v.y;
    s.m[1][0] = v.z;
    s.m[1][2] = -v.x;
    s.m[2][0] = -v.y;
    s.m[2][1] = v.x;
    return s;
}

static M3 madd(const M3& a, Real sb, const M3& b) {
    M3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.m[i][j] = a.m[i][j] + sb * b.m[i][j];
    return r;
}

// common/Math.hpp:147-156
M3 exp_so3(V3 v) {
    const M3 S = skew3(v);
    const Real mag = norm(v);
    if (mag < 1e-8) return madd(M3::identity(), 1.0, S);
    return madd(madd(M3::identity(), std::sin(mag) / mag, S), (1 - std::cos(mag)) / (mag * mag), matmul(S, S));
}

// common/Math.hpp:28-38
M3 jr(V3 th) {
    const Real n = norm(th);
    const M3 S = skew3(th);
    if (n < 1e-8) return madd(M3::identity(), -0.5, S);
    return madd(madd(M3::identity(), -(1 - std::cos(n)) / (n * n), S), (n - std::sin(n)) / (n * n * n), matmul(S, S));
}

// common/Math.hpp:134-145
M3 jr_inv(V3 th) {
    const Real n = norm(th);
    const M3 S = skew3(th);
    if (n < 1e-8) return madd(M3::identity(), 0.5, S);
    return madd(madd(M3::identity(), 0.5, S), 1 / (n * n) - (1 + std::cos(n)) / (2 * n * std::sin(n)), matmul(S, S));
}

// common/Math.hpp:197-221
Quat exp_s3(V3 v) {
    const Real th = norm(v);
    if (th < 1e-6) {
        Quat q(1.0, 0.5 * v.x, 0.5 * v.y, 0.5 * v.z);
        const Real nq = std::sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
        return {q.w / nq, q.x / nq, q.y / nq, q.z / nq};
    }
    const Real h = 0.5 * th;
    const Real s = std::sin(h) / th;
    return {std::cos(h), s * v.x, s * v.y, s * v.z};
}

// common/Math.hpp:223-235
V3 log_s3(const Quat& q) {
    const V3 v(q.x, q.y, q.z);
    const Real vm = norm(v);
    if (vm < 1e-6) return 2.0 * v;
    const Real th = 2 * std::atan2(vm, q.w);
    return (th / vm) * v;
}

// ---------------------------------------------------------------------------
// Cosserat rod (energy/elastic/CosseratRodEnergySolver.hpp:20-159)
// ---------------------------------------------------------------------------
static void cosserat_strain(const CosseratRod& e, const std::vector<V3>& x, const RotState& r, V3& q_diff,
                            Quat& q_mid, Real eps[6]) {
    const Quat& q1 = r.rot(e.idx[0]);
    const Quat& q2 = r.rot(e.idx[1]);
    const V3 p1 = x[e.idx[0]], p2 = x[e.idx[1]];
    const Real inv_L = 1.0 / e.L;
    q_diff = minus_s3(q2, q1);
    q_mid = plus_s3(q1, 0.5 * q_diff);
    const V3 s = qrot(qconj(q_mid), p2 - p1);
    eps[0] = inv_L * s.x;
    eps[1] = inv_L * s.y;
    eps[2] = inv_L * s.z - 1;
    eps[3] = inv_L * q_diff.x - e.kappa0.x;
    eps[4] = inv_L * q_diff.y - e.kappa0.y;
    eps[5] = inv_L * q_diff.z - e.kappa0.z;
}

Real cosserat_energy(const CosseratRod& e, const std::vector<V3>& x, const RotState& r) {
    V3 qd;
    Quat qm;
    Real eps[6];
    cosserat_strain(e, x, r, qd, qm, eps);
    Real s = 0;
    for (int i = 0; i < 6; i++) s += eps[i] * e.stiffness[i] * eps[i];
    return 0.5 * e.L * s;
}

void cosserat_accumulate(const CosseratRod& e, const std::vector<V3>& x, const RotState& r, uint32_t local, V6& G,
                         M6& H) {
    V3 q_diff;
    Quat q_mid;
    Real eps[6];
    cosserat_strain(e, x, r, q_diff, q_mid, eps);
    const Real inv_L = 1.0 / e.L;
    const V3 p1 = x[e.idx[0]], p2 = x[e.idx[1]];
    const M3 gam_inv = jr_inv(q_diff);
    const V3 theta = 0.5 * q_diff;
    const M3 exp_theta = exp_so3(theta);
    const M3 gam_theta = jr(theta);
    const M3 R = qmat(q_mid);
    const V3 dp_ds = inv_L * (p2 - p1);
    const M3 Rt = transpose(R);
    const M3 S = skew3(matvec(Rt, dp_ds));
    Real J[6][6] = {};
    M3 A, B, Cm;
    if (local == 0) {
        M3 dth = transpose(gam_inv);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) dth.m[i][j] *= -0.5;
        M3 dths = transpose(gam_inv);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) dths.m[i][j] *= -inv_L;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) A.m[i][j] = -inv_L * Rt.m[i][j];
        M3 inner = transpose(exp_theta);
        const M3 gd = matmul(gam_theta, dth);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) inner.m[i][j] += gd.m[i][j];
        B = matmul(S, inner);
        Cm = dths;
    } else {
        M3 dth = gam_inv;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) dth.m[i][j] *= 0.5;
        M3 dths = gam_inv;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) dths.m[i][j] *= inv_L;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) A.m[i][j] = inv_L * Rt.m[i][j];
        B = matmul(matmul(S, gam_theta), dth);
        Cm = dths;
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            J[i][j] = A.m[i][j];
            J[i][3 + j] = B.m[i][j];
            J[3 + i][3 + j] = Cm.m[i][j];
        }
    // G += L J^T K eps ; H += L J^T K J (Gauss-Newton)
    for (int c = 0; c < 6; c++) {
        Real s = 0;
        for (int k = 0; k < 6; k++) s += eps[k] * e.stiffness[k] * J[k][c];
        G.v[c] += e.L * s;
    }
    for (int a = 0; a < 6; a++)
        for (int b = 0; b < 6; b++) {
            Real s = 0;
            for (int k = 0; k < 6; k++) s += J[k][a] * e.stiffness[k] * J[k][b];
            H.m[a][b] += e.L * s;
        }
}

// ---------------------------------------------------------------------------
// Rigid-body ground contact (energy/collision/RigidBodyGroundCollisionEnergySolver.hpp)
// ---------------------------------------------------------------------------
void rground_eval(const RigidGround& c, const std::vector<V3>& x, const RotState& r, Real& Cn, Real& Ct, Real& Cb) {
    const V3 p = x[c.p];
    const V3 cp_rb = p + qrot(r.rot(c.p), c.cp_rb_local);
    const V3 diff = c.cp_ground - cp_rb;
    Ct = dot(c.t, diff);
    Cb = dot(c.b, diff);
    if (c.shape == SHAPE_SPHERE || c.shape == SHAPE_ROD) {
        const V3 cp_global = p - V3(0, c.radius, 0);
        Cn = dot(c.n, c.cp_ground - cp_global);
    } else {
        Cn = dot(c.n, diff);
    }
}

static void outer_add(M6& H, const Real a[6], Real s) {
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) H.m[i][j] += s * a[i] * a[j];
}

void rground_accumulate(const RigidGround& c, const std::vector<V3>& x, const RotState& r, V6& G, M6& H) {
    Real Cn, Ct, Cb;
    rground_eval(c, x, r, Cn, Ct, Cb);
    const M3 R = qmat(r.rot(c.p));
    const M3 Sc = skew3(c.cp_rb_local);
    const M3 Rr = matmul(R, Sc);
    const M3 Rrt = transpose(Rr);
    // gradients (position part, rotation part)
    Real gn[6], gt[6], gb[6];
    const V3 rt = matvec(Rrt, c.t), rb = matvec(Rrt, c.b), rn = matvec(Rrt, c.n);
    for (int i = 0; i < 3; i++) {
        gn[i] = -c.n[i];
        gt[i] = -c.t[i];
        gb[i] = -c.b[i];
        gt[3 + i] = rt[i];
        gb[3 + i] = rb[i];
        gn[3 + i] = (c.shape == SHAPE_SPHERE || c.shape == SHAPE_ROD) ? 0.0 : rn[i];
    }
    // rotational Hessian blocks: sym(-R^T [v]x R [cp]x)
    auto hblock = [&](V3 v) {
        M3 h = matmul(matmul(transpose(R), skew3(v)), Rr);
        M3 s;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) s.m[i][j] = 0.5 * (-h.m[i][j] + -h.m[j][i]);
        return s;
    };
    const M3 Ht = hblock(c.t), Hb = hblock(c.b);
    const bool n_hess = !(c.shape == SHAPE_SPHERE || c.shape == SHAPE_ROD);
    const M3 Hn = n_hess ? hblock(c.n) : M3();

    const Real C_corr_n = Cn - CONSTRAINT_ALPHA * c.C_n_prev;
    const Real ln_plus = std::max(Real(0), c.k_n * C_corr_n + c.lambda_n);
    Real lt = c.k_t * Ct + c.lambda_t, lb = c.k_b * Cb + c.lambda_b;
    const Real mag = std::sqrt(lt * lt + lb * lb);
    const Real mu = c.use_static ? c.mu_s : c.mu_k;
    const Real lmax = mu * ln_plus;
    if (mag > lmax) {
        const Real nrm = std::sqrt(lt * lt + lb * lb) + 1e-8;
        lt = lt / nrm * lmax;
        lb = lb / nrm * lmax;
    }
    for (int i = 0; i < 6; i++) G.v[i] += ln_plus * gn[i] + lt * gt[i] + lb * gb[i];
    const Real an = c.k_n * C_corr_n + c.lambda_n, at = c.k_t * Ct + c.lambda_t, ab = c.k_b * Cb + c.lambda_b;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            H.m[3 + i][3 + j] += an * Hn.m[i][j] + at * Ht.m[i][j] + ab * Hb.m[i][j];
    outer_add(H, gn, c.k_n);
    outer_add(H, gt, c.k_t);
    outer_add(H, gb, c.k_b);
}

void rground_update_contact_points(RigidGround& c, const std::vector<V3>& x, const RotState& r) {
    c.cp_ground.x = x[c.p].x;
    c.cp_ground.z = x[c.p].z;
    if (c.shape == SHAPE_SPHERE || c.shape == SHAPE_ROD) c.cp_rb_local = qrot(qconj(r.rot(c.p)), V3(0, -c.radius, 0));
}

void rground_update_after_timestep(RigidGround& c, const std::vector<V3>& x, const RotState& r) {
    if (c.lambda_n > 0) {
        Real Cn, Ct, Cb;
        rground_eval(c, x, r, Cn, Ct, Cb);
        c.C_n_prev = Cn;
    }
    if (c.unclamped_lambda_tb && !c.use_static) c.use_static = true;
    c.lambda_n *= CONSTRAINT_ALPHA * STIFFNESS_GAMMA;
    c.lambda_t *= STIFFNESS_GAMMA;
    c.lambda_b *= STIFFNESS_GAMMA;
    c.k_n = std::max(STIFFNESS_GAMMA * c.k_n, c.k_start);
    c.k_t = std::max(STIFFNESS_GAMMA * c.k_t, c.k_start);
    c.k_b = std::max(STIFFNESS_GAMMA * c.k_b, c.k_start);
}

void rground_update_after_iteration(RigidGround& c, const std::vector<V3>& x, const RotState& r) {
    Real Cn, Ct, Cb;
    rground_eval(c, x, r, Cn, Ct, Cb);
    const Real C_corr_n = Cn - CONSTRAINT_ALPHA * c.C_n_prev;
    const Real ln_plus = c.k_n * C_corr_n + c.lambda_n;
    const Real lt_plus = c.k_t * Ct + c.lambda_t;
    const Real lb_plus = c.k_b * Cb + c.lambda_b;
    if (ln_plus > 0) c.k_n += STIFFNESS_BETA * std::fabs(C_corr_n);
    c.lambda_n = std::max(Real(0), ln_plus);
    const Real mag = std::sqrt(lt_plus * lt_plus + lb_plus * lb_plus);
    Real lmax = (c.use_static ? c.mu_s : c.mu_k) * c.lambda_n;
    if (mag < lmax) {
        c.k_t += STIFFNESS_BETA * std::fabs(Ct);
        c.k_b += STIFFNESS_BETA * std::fabs(Cb);
        c.lambda_t = lt_plus;
        c.lambda_b = lb_plus;
        c.unclamped_lambda_tb = true;
    } else {
        if (c.use_static) {
            c.use_static = false;
            lmax = c.mu_k * ln_plus;
        }
        c.lambda_t = lt_plus / (mag + 1e-8) * lmax;
        c.lambda_b = lb_plus / (mag + 1e-8) * lmax;
        c.unclamped_lambda_tb = false;
    }
}

// ---------------------------------------------------------------------------
// One-sided fixed joint (energy/joint/OneSidedFixedJointEnergySolver.hpp,
// energy/HardConstraintEnergySolver.hpp); lambda bounds are (-inf, +inf).
// ---------------------------------------------------------------------------
void joint_eval(const Joint& j, const std::vector<V3>& x, const RotState& r, Real C[6]) {
    const Quat& q = r.rot(j.p);
    const V3 jp = x[j.p] + qrot(q, j.pos_offset);
    const Quat jq = qmul(q, j.rot_offset);
    const V3 th = minus_s3(jq, j.ref_rotation);
    for (int i = 0; i < 3; i++) {
        C[i] = jp[i] - j.ref_position[i];
        C[3 + i] = th[i];
    }
}

void joint_accumulate(const Joint& j, const std::vector<V3>& x, const RotState& r, V6& G, M6& H) {
    Real Craw[6];
    joint_eval(j, x, r, Craw);
    const Quat& q = r.rot(j.p);
    const M3 R = qmat(q);
    const M3 S = skew3(j.pos_offset);
    M3 dCp = matmul(R, S);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) dCp.m[a][b] = -dCp.m[a][b];
    const V3 th(Craw[3], Craw[4], Craw[5]);
    const M3 dCR = matmul(jr_inv(th), transpose(qmat(j.rot_offset)));
    Real grad[6][6] = {};
    for (int i = 0; i < 3; i++) {
        grad[i][i] = 1;
        for (int k = 0; k < 3; k++) {
            grad[i][3 + k] = dCp.m[i][k];
            grad[3 + i][3 + k] = dCR.m[i][k];
        }
    }
    for (int ci = 0; ci < 6; ci++) {
        const Real C = Craw[ci] - CONSTRAINT_ALPHA * j.C_prev[ci];
        const Real lam = j.k[ci] * C + j.lambda[ci];   // unbounded: never clamped
        for (int a = 0; a < 6; a++) G.v[a] += grad[ci][a] * lam;
        outer_add(H, grad[ci], j.k[ci]);
    }
}

void joint_update_after_timestep(Joint& j, const std::vector<V3>& x, const RotState& r) {
    for (int i = 0; i < 6; i++) {
        j.lambda[i] = CONSTRAINT_ALPHA * STIFFNESS_GAMMA * j.lambda[i];
        j.k[i] = std::max(JOINT_K_START, STIFFNESS_GAMMA * j.k[i]);
    }
    joint_eval(j, x, r, j.C_prev);
}

void joint_update_after_iteration(Joint& j, const std::vector<V3>& x, const RotState& r) {
    Real C[6];
    joint_eval(j, x, r, C);
    for (int i = 0; i < 6; i++) {
        const Real Cc = C[i] - CONSTRAINT_ALPHA * j.C_prev[i];
        const Real lp = j.k[i] * Cc + j.lambda[i];
        if (lp > std::numeric_limits<Real>::lowest() && lp < std::numeric_limits<Real>::max())
            j.k[i] += STIFFNESS_BETA * Cc;
        j.lambda[i] = lp;
    }
}

// ---------------------------------------------------------------------------
// Two-sided fixed joint (new; same vector AL as the one-sided joint).
// Perturbing body 0 by q0 [+] d: frame rotation f0 exp(R(r0)^T d), so
// dC_rot/dtheta_0 = J_r^-1(theta) R(r0)^T, the one-sided joint's form.
// Perturbing body 1: f1 exp(R(r1)^T d), and
// (f1 exp(u))^-1 f0 = exp(-u) f1^-1 f0, whose Log is theta - J_l^-1(theta) u
// to first order with J_l^-1(theta) = J_r^-1(-theta); so
// dC_rot/dtheta_1 = -J_r^-1(-theta) R(r1)^T.  Positions: dC/dx_0 = I,
// dC/dx_1 = -I, dC/dtheta_0 = -R0 [p0]x, dC/dtheta_1 = +R1 [p1]x.
// ---------------------------------------------------------------------------
void joint2_eval(const Joint2& j, const std::vector<V3>& x, const RotState& r, Real C[6]) {
    const Quat& q0 = r.rot(j.p[0]);
    const Quat& q1 = r.rot(j.p[1]);
    const V3 d = (x[j.p[0]] + qrot(q0, j.pos_offset[0])) - (x[j.p[1]] + qrot(q1, j.pos_offset[1]));
    const V3 th = minus_s3(qmul(q0, j.rot_offset[0]), qmul(q1, j.rot_offset[1]));
    for (int i = 0; i < 3; i++) {
        C[i] = d[i];
        C[3 + i] = th[i];
    }
}

void joint2_grad(const Joint2& j, uint32_t side, const std::vector<V3>& x, const RotState& r, Real grad[6][6]) {
    Real C[6];
    joint2_eval(j, x, r, C);
    const Real sgn = side == 0 ? 1.0 : -1.0;
    const M3 dCp = matmul(qmat(r.rot(j.p[side])), skew3(j.pos_offset[side]));   // R [p]x
    const V3 th = side == 0 ? V3(C[3], C[4], C[5]) : V3(-C[3], -C[4], -C[5]);
    const M3 dCR = matmul(jr_inv(th), transpose(qmat(j.rot_offset[side])));
    for (int a = 0; a < 6; a++)
        for (int b = 0; b < 6; b++) grad[a][b] = 0;
    for (int i = 0; i < 3; i++) {
        grad[i][i] = sgn;
        for (int k = 0; k < 3; k++) {
            grad[i][3 + k] = -sgn * dCp.m[i][k];
            grad[3 + i][3 + k] = sgn * dCR.m[i][k];
        }
    }
}

void joint2_accumulate(const Joint2& j, uint32_t side, const std::vector<V3>& x, const RotState& r, V6& G, M6& H) {
    Real Craw[6], grad[6][6];
    joint2_eval(j, x, r, Craw);
    joint2_grad(j, side, x, r, grad);
    for (int ci = 0; ci < 6; ci++) {
        const Real C = Craw[ci] - CONSTRAINT_ALPHA * j.C_prev[ci];
        const Real lam = j.k[ci] * C + j.lambda[ci];
        for (int a = 0; a < 6; a++) G.v[a] += grad[ci][a] * lam;
        outer_add(H, grad[ci], j.k[ci]);
    }
}

void joint2_update_after_timestep(Joint2& j, const std::vector<V3>& x, const RotState& r) {
    for (int i = 0; i < 6; i++) {
        j.lambda[i] = CONSTRAINT_ALPHA * STIFFNESS_GAMMA * j.lambda[i];
        j.k[i] = std::max(JOINT_K_START, STIFFNESS_GAMMA * j.k[i]);
    }
    joint2_eval(j, x, r, j.C_prev);
}

void joint2_update_after_iteration(Joint2& j, const std::vector<V3>& x, const RotState& r) {
    Real C[6];
    joint2_eval(j, x, r, C);
    for (int i = 0; i < 6; i++) {
        const Real Cc = C[i] - CONSTRAINT_ALPHA * j.C_prev[i];
        const Real lp = j.k[i] * Cc + j.lambda[i];
        if (lp > std::numeric_limits<Real>::lowest() && lp < std::numeric_limits<Real>::max())
            j.k[i] += STIFFNESS_BETA * Cc;
        j.lambda[i] = lp;
    }
}

}  // namespace oracle
