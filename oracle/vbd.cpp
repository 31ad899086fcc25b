// vbd.cpp -- FP64 restatement of the reference AVBD step (test infrastructure
// only; see oracle.hpp for the rules on who may load this).
#include "simulation.hpp"

#include <algorithm>
#include <stdexcept>
#include <thread>
#include <immintrin.h>

namespace oracle {

// ---------------------------------------------------------------------------
// small dense math
// ---------------------------------------------------------------------------
M3 matmul(const M3& a, const M3& b) {
    M3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            Real s = 0;
            for (int k = 0; k < 3; k++) s += a.m[i][k] * b.m[k][j];
            r.m[i][j] = s;
        }
    return r;
}
M3 transpose(const M3& a) {
    M3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.m[i][j] = a.m[j][i];
    return r;
}
V3 matvec(const M3& a, V3 v) {
    return {a.m[0][0] * v.x + a.m[0][1] * v.y + a.m[0][2] * v.z,
            a.m[1][0] * v.x + a.m[1][1] * v.y + a.m[1][2] * v.z,
            a.m[2][0] * v.x + a.m[2][1] * v.y + a.m[2][2] * v.z};
}
Real det(const M3& a) {
    return a.m[0][0] * (a.m[1][1] * a.m[2][2] - a.m[1][2] * a.m[2][1]) -
           a.m[0][1] * (a.m[1][0] * a.m[2][2] - a.m[1][2] * a.m[2][0]) +
           a.m[0][2] * (a.m[1][0] * a.m[2][1] - a.m[1][1] * a.m[2][0]);
}
M3 inverse(const M3& a) {
    // adjugate / determinant (simobject/TetMeshObject.cpp:80 uses X.inverse())
    Real d = det(a);
    M3 r;
    r.m[0][0] = (a.m[1][1] * a.m[2][2] - a.m[1][2] * a.m[2][1]) / d;
    r.m[0][1] = (a.m[0][2] * a.m[2][1] - a.m[0][1] * a.m[2][2]) / d;
    r.m[0][2] = (a.m[0][1] * a.m[1][2] - a.m[0][2] * a.m[1][1]) / d;
    r.m[1][0] = (a.m[1][2] * a.m[2][0] - a.m[1][0] * a.m[2][2]) / d;
    r.m[1][1] = (a.m[0][0] * a.m[2][2] - a.m[0][2] * a.m[2][0]) / d;
    r.m[1][2] = (a.m[0][2] * a.m[1][0] - a.m[0][0] * a.m[1][2]) / d;
    r.m[2][0] = (a.m[1][0] * a.m[2][1] - a.m[1][1] * a.m[2][0]) / d;
    r.m[2][1] = (a.m[0][1] * a.m[2][0] - a.m[0][0] * a.m[2][1]) / d;
    r.m[2][2] = (a.m[0][0] * a.m[1][1] - a.m[0][1] * a.m[1][0]) / d;
    return r;
}

void lu_solve(int n, const Real* Ain, const Real* b, Real* x) {
    // Gaussian elimination with partial (row) pivoting: the algorithm behind
    // Eigen's partialPivLu used at simulation/VBDSolver.hpp:472,498.
    Real A[36], r[6];
    for (int i = 0; i < n * n; i++) A[i] = Ain[i];
    for (int i = 0; i < n; i++) r[i] = b[i];
    for (int k = 0; k < n; k++) {
        int piv = k;
        Real best = std::fabs(A[k * n + k]);
        for (int i = k + 1; i < n; i++) {
            Real v = std::fabs(A[i * n + k]);
            if (v > best) { best = v; piv = i; }
        }
        if (piv != k) {
            for (int j = 0; j < n; j++) std::swap(A[k * n + j], A[piv * n + j]);
            std::swap(r[k], r[piv]);
        }
        Real inv = 1.0 / A[k * n + k];
        for (int i = k + 1; i < n; i++) {
            Real f = A[i * n + k] * inv;
            if (f == 0) continue;
            for (int j = k + 1; j < n; j++) A[i * n + j] -= f * A[k * n + j];
            r[i] -= f * r[k];
        }
    }
    for (int i = n - 1; i >= 0; i--) {
        Real s = r[i];
        for (int j = i + 1; j < n; j++) s -= A[i * n + j] * x[j];
        x[i] = s / A[i * n + i];
    }
}

// ---------------------------------------------------------------------------
// Neo-Hookean tet energy
// ---------------------------------------------------------------------------

// F = [x0-x3, x1-x3, x2-x3] * Q  (NeoHookeanEnergySolver.hpp:96-110)
static M3 compute_F(V3 a, V3 b, V3 c, V3 d, const M3& Q) {
    M3 D;
    V3 c0 = a - d, c1 = b - d, c2 = c - d;
    for (int r = 0; r < 3; r++) { D.m[r][0] = c0[r]; D.m[r][1] = c1[r]; D.m[r][2] = c2[r]; }
    return matmul(D, Q);
}

// NeoHookeanEnergySolver::energy, NeoHookeanEnergySolver.hpp:20-53
Real nh_energy(const NHTet& t, const std::vector<V3>& x, Real dt) {
    M3 F = compute_F(x[t.idx[0]], x[t.idx[1]], x[t.idx[2]], x[t.idx[3]], t.Q);
    M3 FtF = matmul(transpose(F), F);
    Real damp = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            Real E = 0.5 * (FtF.m[i][j] - (i == j ? 1.0 : 0.0));
            Real Ed = (E - t.E_prev.m[i][j]) / dt;
            damp += Ed * Ed;
        }
    damp *= t.V * t.kd;
    Real fac = det(F) - t.mu / t.lambda - 1;
    Real hyd = 0.5 * t.lambda * t.V * fac * fac;
    Real dev = 0.5 * t.mu * t.V * (FtF.m[0][0] + FtF.m[1][1] + FtF.m[2][2] - 3);
    return hyd + dev + damp;
}

// NeoHookeanEnergySolver::updateAfterTimeStep, NeoHookeanEnergySolver.hpp:70-90
void nh_update_after_timestep(NHTet& t, const std::vector<V3>& xp) {
    M3 F = compute_F(xp[t.idx[0]], xp[t.idx[1]], xp[t.idx[2]], xp[t.idx[3]], t.Q);
    M3 FtF = matmul(transpose(F), F);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) t.E_prev.m[i][j] = 0.5 * (FtF.m[i][j] - (i == j ? 1.0 : 0.0));
}

// Per-incidence contribution shared by the scalar and the 4-wide paths.
// Follows NeoHookeanEnergySolver::accumulate, NeoHookeanEnergySolver.hpp:121-227:
//   gradient  = sign*hyd*(cof q) + sign*dev*(F q) + sign*(2V kd/dt) F Edot q
//   Hessian   = V lambda (cof q)(cof q)^T + V mu |q|^2 I
//             + (2V kd/dt) (q^T Edot q) I + (V kd/dt^2)(F F^T |q|^2 + Fq (Fq)^T)
static void nh_incidence(const NHTet& t, const std::vector<V3>& x, uint32_t local, Real dt,
                         Real g[3], Real h[3][3]) {
    M3 F = compute_F(x[t.idx[0]], x[t.idx[1]], x[t.idx[2]], x[t.idx[3]], t.Q);
    V3 f0 = F.col(0), f1 = F.col(1), f2 = F.col(2);
    V3 c0 = cross(f1, f2), c1 = cross(f2, f0), c2 = cross(f0, f1);
    Real detF = dot(f0, c0);
    V3 qi; Real sign;
    if (local < 3) { qi = t.Q.row(local); sign = 1.0; }
    else { qi = t.Q.row(0) + t.Q.row(1) + t.Q.row(2); sign = -1.0; }
    V3 Fqi = matvec(F, qi);
    V3 Fc = qi.x * c0 + qi.y * c1 + qi.z * c2;
    Real hyd = t.V * t.lambda * (detF - t.mu / t.lambda - 1);
    Real dev = t.V * t.mu;
    M3 FtF = matmul(transpose(F), F);
    Real q2 = dot(qi, qi);
    Real damp = 2 / dt * t.V * t.kd;
    M3 Ed;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            Ed.m[i][j] = ((0.5 * FtF.m[i][j] - (i == j ? 0.5 : 0.0)) - t.E_prev.m[i][j]) / dt;
    V3 Edq = matvec(Ed, qi);
    V3 FEdq = matvec(F, Edq);
    for (int i = 0; i < 3; i++)
        g[i] = sign * hyd * Fc[i] + sign * dev * Fqi[i] + sign * damp * FEdq[i];
    M3 FFt = matmul(F, transpose(F));
    Real qEq = dot(qi, Edq);
    Real VL = t.V * t.lambda;
    Real alpha = 0.5 * damp / dt;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            Real v = VL * Fc[i] * Fc[j];
            if (i == j) v += t.V * t.mu * q2 + damp * qEq;
            v += alpha * (q2 * FFt.m[i][j] + Fqi[i] * Fqi[j]);
            h[i][j] = v;
        }
}

void nh_accumulate(const NHTet& t, const std::vector<V3>& x, uint32_t local, V3& G, M3& H, Real dt) {
    Real g[3], h[3][3];
    nh_incidence(t, x, local, dt, g, h);
    G.x += g[0]; G.y += g[1]; G.z += g[2];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) H.m[i][j] += h[i][j];
}

// accumulate4 (NeoHookeanEnergySolver.hpp:330-573): four incidences are
// evaluated lane-parallel and folded with hsum_pd (common/AVX.hpp:96-105),
// i.e. (l0 + l2) + (l1 + l3), before being added to the vertex totals.
void nh_accumulate4(const NHTet* const t[4], const std::vector<V3>& x, const uint32_t local[4],
                    V3& G, M3& H, Real dt) {
    Real g[4][3], h[4][3][3];
    for (int l = 0; l < 4; l++) nh_incidence(*t[l], x, local[l], dt, g[l], h[l]);
    auto hsum = [](Real a0, Real a1, Real a2, Real a3) { return (a0 + a2) + (a1 + a3); };
    G.x += hsum(g[0][0], g[1][0], g[2][0], g[3][0]);
    G.y += hsum(g[0][1], g[1][1], g[2][1], g[3][1]);
    G.z += hsum(g[0][2], g[1][2], g[2][2], g[3][2]);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) H.m[i][j] += hsum(h[0][i][j], h[1][i][j], h[2][i][j], h[3][i][j]);
}

// ---------------------------------------------------------------------------
// Ground contact (augmented Lagrangian, friction cone)
// ---------------------------------------------------------------------------

// GroundCollisionConstraintSolver::evaluateConstraint, GroundCollisionEnergySolver.hpp:14-25
void ground_eval(const GroundContact& c, const std::vector<V3>& x, Real& Cn, Real& Ct, Real& Cb) {
    const V3& p = x[c.p];
    Cn = -p.y;
    Ct = p.x - c.cp_x;
    Cb = p.z - c.cp_z;
}

// CollisionConstraintEnergySolver::accumulate<3>, CollisionConstraintEnergySolver.hpp:180-285,
// with the ground constraint gradients (0,-1,0), (1,0,0), (0,0,1) and zero constraint Hessians
// (GroundCollisionEnergySolver.hpp:27-60).
void ground_accumulate(const GroundContact& c, const std::vector<V3>& x, V3& G, M3& H) {
    Real Cn, Ct, Cb;
    ground_eval(c, x, Cn, Ct, Cb);
    Real C_corr_n = Cn - CONSTRAINT_ALPHA * c.C_n_prev;
    Real ln_plus = std::max(Real(0), c.k_n * C_corr_n + c.lambda_n);
    Real lt = c.k_t * Ct + c.lambda_t, lb = c.k_b * Cb + c.lambda_b;
    Real mag = std::sqrt(lt * lt + lb * lb);
    Real mu = c.use_static ? c.mu_s : c.mu_k;
    Real lmax = mu * ln_plus;
    if (mag > lmax) {
        // lambda_tb_plus / (norm + 1e-8) * lambda_tb_max, component-wise
        Real nrm = std::sqrt(lt * lt + lb * lb) + 1e-8;
        lt = lt / nrm * lmax;
        lb = lb / nrm * lmax;
    }
    // gradients: n = (0,-1,0), t = (1,0,0), b = (0,0,1)
    V3 gn(0, -ln_plus, 0), gt(lt, 0, 0), gb(0, 0, lb);
    // Hessians: k * grad grad^T (constraint Hessians are zero)
    H.m[1][1] += c.k_n;    // (0,-1,0)(0,-1,0)^T
    H.m[0][0] += c.k_t;
    H.m[2][2] += c.k_b;
    G += gn; G += gt; G += gb;
}

// GroundCollisionEnergySolver::updateContactPoints, GroundCollisionEnergySolver.hpp:81-91
void ground_update_contact_points(GroundContact& c, const std::vector<V3>& x) {
    c.cp_x = x[c.p].x;
    c.cp_z = x[c.p].z;
}

// CollisionConstraintEnergySolver::updateAfterTimeStep, CollisionConstraintEnergySolver.hpp:141-167
void ground_update_after_timestep(GroundContact& c, const std::vector<V3>& x) {
    if (c.lambda_n > 0) {
        Real Cn, Ct, Cb;
        ground_eval(c, x, Cn, Ct, Cb);
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

// CollisionConstraintEnergySolver::updateAfterIteration, CollisionConstraintEnergySolver.hpp:60-134
void ground_update_after_iteration(GroundContact& c, const std::vector<V3>& x) {
    Real Cn, Ct, Cb;
    ground_eval(c, x, Cn, Ct, Cb);
    Real C_corr_n = Cn - CONSTRAINT_ALPHA * c.C_n_prev;
    Real ln_plus = c.k_n * C_corr_n + c.lambda_n;
    Real lt_plus = c.k_t * Ct + c.lambda_t;
    Real lb_plus = c.k_b * Cb + c.lambda_b;
    if (ln_plus > 0) c.k_n += STIFFNESS_BETA * std::fabs(C_corr_n);
    c.lambda_n = std::max(Real(0), ln_plus);
    Real mag = std::sqrt(lt_plus * lt_plus + lb_plus * lb_plus);
    Real mu = c.use_static ? c.mu_s : c.mu_k;
    Real lmax = mu * c.lambda_n;
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
// ---------------------------------------------------------------------------
// Setup
// ---------------------------------------------------------------------------

static void push_particle(Simulation& s, V3 p, Real m) {
    s.x.push_back(p); s.x_buf.push_back(p); s.y.push_back(p); s.x_prev.push_back(p);
    s.x_last.push_back(p); s.x_lastlast.push_back(p);
    s.v.push_back(V3()); s.v_prev.push_back(V3());
    s.mass.push_back(m); s.fixed.push_back(0);
    s.rs.rot_of.push_back(-1);
}

// TetMeshObject::setup, simobject/TetMeshObject.cpp:46-96, and the per-vertex
// ground energies of simulation/Simulation.hpp:117-126.
uint32_t Simulation::add_tet_mesh(uint32_t nv, const Real* rest_x, uint32_t nt, const uint32_t* tets,
                                  Real E, Real nu, Real density, Real kd, bool with_ground,
                                  Real mu_s, Real mu_k) {
    uint32_t base = (uint32_t)x.size();
    for (uint32_t i = 0; i < nv; i++) push_particle(*this, V3(rest_x[3 * i], rest_x[3 * i + 1], rest_x[3 * i + 2]), 0);
    Real mu = E / (2 * (1 + nu));
    Real lambda = (E * nu) / ((1 + nu) * (1 - 2 * nu));
    for (uint32_t e = 0; e < nt; e++) {
        NHTet t;
        for (int k = 0; k < 4; k++) t.idx[k] = base + tets[4 * e + k];
        V3 v1 = x[t.idx[0]], v2 = x[t.idx[1]], v3 = x[t.idx[2]], v4 = x[t.idx[3]];
        M3 X;
        V3 a = v1 - v4, b = v2 - v4, c = v3 - v4;
        for (int r = 0; r < 3; r++) { X.m[r][0] = a[r]; X.m[r][1] = b[r]; X.m[r][2] = c[r]; }
        t.Q = inverse(X);
        t.V = std::fabs(det(X) / 6.0);
        t.lambda = lambda; t.mu = mu; t.kd = kd;
        t.E_prev = M3();
        nh.push_back(t);
        for (int k = 0; k < 4; k++) mass[t.idx[k]] += 0.25 * t.V * density;
    }
    if (with_ground) {
        for (uint32_t i = 0; i < nv; i++) {
            uint32_t p = base + i;
            Real k_start = mass[p] / (params.dt * params.dt);
            GroundContact c{};
            // CollisionConstraintEnergyPool::addEnergy, CollisionConstraintEnergyPool.hpp:48-69
            c.p = p;
            c.lambda_n = c.lambda_t = c.lambda_b = 0;
            c.k_start = c.k_n = c.k_t = c.k_b = k_start;
            c.C_n_prev = 0;
            c.mu_s = mu_s; c.mu_k = mu_k;
            c.use_static = false; c.unclamped_lambda_tb = false;
            c.cp_x = 0; c.cp_z = 0;
            ground.push_back(c);
        }
    }
    return base;
}

uint32_t Simulation::add_particles(uint32_t n, const Real* xs, const Real* m) {
    uint32_t base = (uint32_t)x.size();
    for (uint32_t i = 0; i < n; i++) push_particle(*this, V3(xs[3 * i], xs[3 * i + 1], xs[3 * i + 2]), m[i]);
    return base;
}

uint32_t Simulation::add_oriented(V3 p, Quat q, Real m, V3 inertia) {
    const uint32_t idx = (uint32_t)x.size();
    push_particle(*this, p, m);
    rs.rot_of[idx] = (int32_t)rs.q.size();
    // RotationPool::addParticle, common/ParticlePool.hpp:37-51 (angular velocity starts at zero)
    rs.q.push_back(q); rs.q_buf.push_back(q); rs.q_in.push_back(q); rs.q_prev.push_back(q);
    rs.q_last.push_back(q); rs.q_lastlast.push_back(q);
    rs.w.push_back(V3()); rs.w_prev.push_back(V3()); rs.inertia.push_back(inertia);
    return idx;
}

static RigidGround make_rground(uint32_t p, uint32_t shape, Real radius, V3 cp_local, Real k_start, Real mu_s,
                                Real mu_k) {
    // RigidBodyGroundCollisionEnergyPool::addEnergy (+ CollisionConstraintEnergyPool::addEnergy)
    RigidGround c{};
    c.p = p; c.shape = shape; c.radius = radius; c.cp_rb_local = cp_local; c.cp_ground = V3();
    c.n = V3(0, 1, 0); c.t = V3(1, 0, 0); c.b = V3(0, 0, 1);
    c.k_start = c.k_n = c.k_t = c.k_b = k_start;
    c.lambda_n = c.lambda_t = c.lambda_b = 0; c.C_n_prev = 0;
    c.use_static = false; c.unclamped_lambda_tb = false;
    c.mu_s = mu_s; c.mu_k = mu_k;
    return c;
}

uint32_t Simulation::add_rod(uint32_t elements, Real length, V3 curvature, Real E, Real nu, Real density,
                             Real diameter, V3 pos, Quat rot, V3 vel, bool base_fixed, bool tip_fixed,
                             bool with_ground) {
    const Real radius = diameter / 2;
    const Real G = E / (2 * (1 + nu));
    const uint32_t nn = elements + 1;
    const Real ds = length / (nn - 1);
    const V3 dR = ds * curvature;
    std::vector<uint32_t> nodes(nn);
    V3 cp = pos;
    Quat cq = rot;
    for (uint32_t i = 0; i < nn; i++) {
        nodes[i] = add_oriented(cp, cq, 0, V3());
        v[nodes[i]] = vel;   // previous_velocities stay zero (simobject/Rod.cpp:33)
        cp = cp + qrot(cq, V3(0, 0, ds));
        cq = qmul(cq, exp_s3(dR));
    }
    const Real pi = 3.14159265358979323846;
    const Real area = pi * radius * radius;
    const Real Ix = pi * radius * radius * radius * radius / 4.0, Iy = Ix, Iz = 2 * Ix;
    const Real elem_len = length / (nn - 1);
    const Real elem_mass = elem_len * area * density;
    const V3 elem_inertia = (density * elem_len) * V3(Ix, Iy, Iz);
    for (uint32_t e = 0; e < elements; e++) {
        mass[nodes[e]] += 0.5 * elem_mass;
        mass[nodes[e + 1]] += 0.5 * elem_mass;
        rs.inertia[rs.rot_of[nodes[e]]] += 0.5 * elem_inertia;
        rs.inertia[rs.rot_of[nodes[e + 1]]] += 0.5 * elem_inertia;
    }
    for (uint32_t e = 0; e < elements; e++) {
        CosseratRod r{};
        r.idx[0] = nodes[e]; r.idx[1] = nodes[e + 1];
        const Real st[6] = {G * area, G * area, E * area, E * Ix, E * Iy, G * Iz};
        for (int k = 0; k < 6; k++) r.stiffness[k] = st[k];
        r.L = ds;
        r.kappa0 = curvature;
        cos.push_back(r);
    }
    if (base_fixed) fixed[nodes.front()] = 1;
    if (tip_fixed) fixed[nodes.back()] = 1;
    if (with_ground)
        for (uint32_t n_idx : nodes)
            rground.push_back(make_rground(n_idx, SHAPE_ROD, radius, V3(0, -radius, 0),
                                           mass[n_idx] / (params.dt * params.dt), 0.5, 0.2));
    return nodes.front();
}

uint32_t Simulation::add_sphere(Real radius, Real density, V3 pos, Quat rot, bool fixed_, bool with_ground) {
    const Real pi = 3.14159265358979323846;
    const uint32_t p = add_oriented(pos, rot, 0, V3());
    fixed[p] = fixed_ ? 1 : 0;
    const Real m = 4 * pi / 3 * radius * radius * radius * density;
    mass[p] = m;
    rs.inertia[rs.rot_of[p]] = (0.4 * m * radius * radius) * V3(1, 1, 1);
    if (with_ground)
        rground.push_back(make_rground(p, SHAPE_SPHERE, radius, V3(0, -radius, 0), m / (params.dt * params.dt),
                                       0.5, 0.2));
    return p;
}

uint32_t Simulation::add_joint(uint32_t p, V3 pos_offset, Quat rot_offset, V3 ref_position, Quat ref_rotation) {
    Joint j{};
    j.p = p; j.pos_offset = pos_offset; j.rot_offset = rot_offset;
    j.ref_position = ref_position; j.ref_rotation = ref_rotation;
    for (int i = 0; i < 6; i++) { j.k[i] = JOINT_K_START; j.lambda[i] = 0; j.C_prev[i] = 0; }
    joints.push_back(j);
    return (uint32_t)joints.size() - 1;
}

uint32_t Simulation::add_joint2(uint32_t p0, uint32_t p1, V3 pos_offset0, Quat rot_offset0, V3 pos_offset1,
                                Quat rot_offset1) {
    if (p0 == p1) throw std::runtime_error("two-sided joint with one particle twice");
    if (p0 >= x.size() || p1 >= x.size() || rs.rot_of[p0] < 0 || rs.rot_of[p1] < 0)
        throw std::runtime_error("joint needs oriented particles");
    Joint2 j{};
    j.p[0] = p0; j.p[1] = p1;
    j.pos_offset[0] = pos_offset0; j.pos_offset[1] = pos_offset1;
    j.rot_offset[0] = rot_offset0; j.rot_offset[1] = rot_offset1;
    for (int i = 0; i < 6; i++) { j.k[i] = JOINT_K_START; j.lambda[i] = 0; j.C_prev[i] = 0; }
    joints2.push_back(j);
    return (uint32_t)joints2.size() - 1;
}

void Simulation::build_graph() {
    uint32_t n = (uint32_t)x.size();
    build_static_adjacency(n, nh, cos, ground, rground, joints, joints2, adj);
    dsatur(adj, n, coloring);
    rebuild_work_list(n, coloring);
    is_conflicted.assign(n, 0);
}

// ---------------------------------------------------------------------------
// Step phases (VBDSolver.hpp)
// ---------------------------------------------------------------------------

void Simulation::phase_after_timestep(uint32_t lo_nh, uint32_t hi_nh, uint32_t lo_g, uint32_t hi_g) {
    for (uint32_t e = lo_nh; e < hi_nh; e++) nh_update_after_timestep(nh[e], x_prev);
    for (uint32_t e = lo_g; e < hi_g; e++) {
        ground_update_contact_points(ground[e], x);
        ground_update_after_timestep(ground[e], x);
    }
}

void Simulation::phase_after_timestep_rigid(uint32_t lo, uint32_t hi) {
    // records [0, |rground|), then the joints, then the two-sided joints
    const size_t nj = rground.size() + joints.size();
    for (uint32_t e = lo; e < hi; e++) {
        if (e < rground.size()) {
            rground_update_contact_points(rground[e], x, rs);
            rground_update_after_timestep(rground[e], x, rs);
        } else if (e < nj) {
            joint_update_after_timestep(joints[e - rground.size()], x, rs);
        } else {
            joint2_update_after_timestep(joints2[e - nj], x, rs);
        }
    }
}

// _particleInertialUpdate + _particleRotationInertialUpdate, VBDSolver.hpp:208-283
// _particleInertialUpdate (simulation/VBDSolver.hpp:208-248).  With a
// per-particle field f_ext (not in the reference) the external acceleration
// of particle p is gravity + f_ext[p], in the inertial target and in the
// adaptive initial guess alike.
void Simulation::phase_inertial(uint32_t lo, uint32_t hi) {
    const V3 gravity(0, -params.g, 0);
    Real dt = params.dt;
    for (uint32_t p = lo; p < hi; p++) {
        if (fixed[p]) continue;
        const V3 a_ext = f_ext.empty() ? gravity : gravity + f_ext[p];
        V3 a = (v[p] - v_prev[p]) / dt;
        V3 a_tilde_vec;
        Real an = norm(a_ext);
        if (an > 1e-12) {
            Real along = dot(a, a_ext) / an;
            Real at;
            if (along > an) at = 1;
            else if (along < 0) at = 0;
            else at = along / an;
            a_tilde_vec = at * a_ext;
        }
        y[p] = x[p] + dt * v[p] + (dt * dt) * a_ext;
        x[p] = x[p] + (dt * v[p] + (dt * dt) * a_tilde_vec);
        x_last[p] = x[p];
        x_lastlast[p] = x[p];
        if (oriented(p)) {
            const int32_t r = rs.rot_of[p];
            const V3 w = rs.w[r], I = rs.inertia[r];
            const V3 Iinv(1 / I.x, 1 / I.y, 1 / I.z);
            const V3 Iw(I.x * w.x, I.y * w.y, I.z * w.z);
            const V3 c = cross(w, Iw);
            const V3 corr(-Iinv.x * c.x, -Iinv.y * c.y, -Iinv.z * c.z);   // alpha_ext = 0
            const V3 dq_in = dt * w + (dt * dt) * corr;
            rs.q_in[r] = plus_s3(rs.q[r], dq_in);
            rs.q[r] = plus_s3(rs.q[r], dt * w);
            rs.q_last[r] = rs.q[r];
            rs.q_lastlast[r] = rs.q[r];
        }
    }
}

void Simulation::accumulate_vertex(uint32_t p, V3& G, M3& H) const {
    Real dt = params.dt;
    const auto& off = adj.e_offsets[p];
    uint32_t end = adj.e_offsets[p + 1][0];
    for (int ty = 0; ty < ST_COUNT; ty++) {
        uint32_t e = off[ty];
        uint32_t fe = (ty + 1 == ST_COUNT) ? end : off[ty + 1];
        if (ty == ST_NEO_HOOKEAN) {
            // chunks of 4 then the scalar remainder (VBDSolver.hpp:316-358)
            for (; e + 3 < fe; e += 4) {
                const NHTet* t4[4];
                uint32_t l4[4];
                for (int l = 0; l < 4; l++) {
                    t4[l] = &nh[adj.e_entries[e + l].e_idx];
                    l4[l] = adj.e_entries[e + l].local;
                }
                nh_accumulate4(t4, x, l4, G, H, dt);
            }
            for (; e < fe; e++) nh_accumulate(nh[adj.e_entries[e].e_idx], x, adj.e_entries[e].local, G, H, dt);
        } else if (ty == ST_GROUND) {
            for (; e < fe; e++) ground_accumulate(ground[adj.e_entries[e].e_idx], x, G, H);
        }
        // Cosserat, rigid ground and joints support oriented particles only
    }
    // dynamic energies after the static ones (VBDSolver.hpp:400-438)
    if (dynamic && p + 1 < dadj.e_offsets.size()) {
        for (uint32_t e = dadj.e_offsets[p][0]; e < dadj.e_offsets[p + 1][0]; e++) {
            const DynAdjEntry& en = dadj.e_entries[e];
            dyn_accumulate3(en.type, dyn[en.type].data[en.e_idx], x, rs, en.local, G, H);
        }
    }
}

// Static energies of an oriented particle in StaticEnergyType order
// (VBDSolver.hpp:297-397): NH (positional only: skipped), Cosserat, ground
// (accumulate<6>: positional block), rigid ground, joint.
void Simulation::accumulate_vertex6(uint32_t p, V6& G, M6& H) const {
    const auto& off = adj.e_offsets[p];
    uint32_t end = adj.e_offsets[p + 1][0];
    for (int ty = 0; ty < ST_COUNT; ty++) {
        uint32_t e = off[ty];
        uint32_t fe = (ty + 1 == ST_COUNT) ? end : off[ty + 1];
        for (; e < fe; e++) {
            const AdjEntry& en = adj.e_entries[e];
            if (ty == ST_COSSERAT) {
                cosserat_accumulate(cos[en.e_idx], x, rs, en.local, G, H);
            } else if (ty == ST_GROUND) {
                V3 g3; M3 h3;
                ground_accumulate(ground[en.e_idx], x, g3, h3);
                for (int i = 0; i < 3; i++) {
                    G.v[i] += g3[i];
                    for (int j = 0; j < 3; j++) H.m[i][j] += h3.m[i][j];
                }
            } else if (ty == ST_RIGID_GROUND) {
                rground_accumulate(rground[en.e_idx], x, rs, G, H);
            } else if (ty == ST_JOINT) {
                if (en.e_idx < joints.size()) joint_accumulate(joints[en.e_idx], x, rs, G, H);
                else joint2_accumulate(joints2[en.e_idx - joints.size()], en.local, x, rs, G, H);
            }
        }
    }
    if (dynamic && p + 1 < dadj.e_offsets.size()) {
        for (uint32_t e = dadj.e_offsets[p][0]; e < dadj.e_offsets[p + 1][0]; e++) {
            const DynAdjEntry& en = dadj.e_entries[e];
            dyn_accumulate6(en.type, dyn[en.type].data[en.e_idx], x, rs, en.local, G, H);
        }
    }
}

// _solveParticle<3>, VBDSolver.hpp:285-449,487-510
void Simulation::solve_particle(uint32_t p) {
    if (fixed[p]) return;
    if (oriented(p)) {
        solve_particle6(p);
        return;
    }
    V3 G; M3 H;
    accumulate_vertex(p, G, H);
    Real dt = params.dt;
    Real m = mass[p];
    V3 rhs = (-m / (dt * dt)) * (x[p] - y[p]);
    rhs = rhs - G;
    Real A[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) A[3 * i + j] = (i == j ? m / (dt * dt) : 0.0) + H.m[i][j];
    Real b[3] = {rhs.x, rhs.y, rhs.z}, dx[3];
    lu_solve(3, A, b, dx);
    V3 d(dx[0], dx[1], dx[2]);
    if (is_conflicted[p]) x_buf[p] = x[p] + d;
    else x[p] += d;
}

// _solveParticle<6>, VBDSolver.hpp:450-486
void Simulation::solve_particle6(uint32_t p) {
    V6 G; M6 H;
    accumulate_vertex6(p, G, H);
    const Real dt = params.dt;
    const Real m = mass[p];
    const int32_t r = rs.rot_of[p];
    const V3 I = rs.inertia[r];
    const V3 dpos = (-1 / (dt * dt) * m) * (x[p] - y[p]);
    const V3 th = minus_s3(rs.q[r], rs.q_in[r]);
    const Real s = -1 / (dt * dt);
    Real rhs[6] = {dpos.x, dpos.y, dpos.z, s * (I.x * th.x), s * (I.y * th.y), s * (I.z * th.z)};
    for (int i = 0; i < 6; i++) rhs[i] -= G.v[i];
    Real A[36];
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) A[6 * i + j] = H.m[i][j];
    for (int i = 0; i < 3; i++) {
        A[6 * i + i] += m / (dt * dt);
        A[6 * (3 + i) + 3 + i] += I[i] / (dt * dt);
    }
    Real dx[6];
    lu_solve(6, A, rhs, dx);
    const V3 dp(dx[0], dx[1], dx[2]), dth(dx[3], dx[4], dx[5]);
    if (is_conflicted[p]) {
        x_buf[p] = x[p] + dp;
        rs.q_buf[r] = plus_s3(rs.q[r], dth);
    } else {
        x[p] += dp;
        rs.q[r] = plus_s3(rs.q[r], dth);
    }
}

void Simulation::phase_color(uint32_t c, uint32_t lo, uint32_t hi) {
    uint32_t start = coloring.color_offsets[c];
    for (uint32_t k = start + lo; k < start + hi; k++) solve_particle(coloring.work_list[k]);
}

// _particleChebyshevAcceleration, VBDSolver.hpp:513-536
void Simulation::phase_chebyshev(Real omega, uint32_t lo, uint32_t hi) {
    for (uint32_t p = lo; p < hi; p++) {
        x[p] = omega * (x[p] - x_lastlast[p]) + x_lastlast[p];
        if (oriented(p)) {
            const int32_t r = rs.rot_of[p];
            rs.q[r] = plus_s3(rs.q_lastlast[r], omega * minus_s3(rs.q[r], rs.q_lastlast[r]));
        }
        x_lastlast[p] = x_last[p];
        x_last[p] = x[p];
        if (oriented(p)) {
            const int32_t r = rs.rot_of[p];
            rs.q_lastlast[r] = rs.q_last[r];
            rs.q_last[r] = rs.q[r];
        }
    }
}

// _commitConflictedInColor (VBDSolver.hpp:108-119): positions only -- the
// buffered rotation of a conflicted oriented particle is not committed.
void Simulation::phase_commit(uint32_t c) {
    if (!dynamic || c + 1 >= rc.conflict_offsets.size()) return;
    for (uint32_t k = rc.conflict_offsets[c]; k < rc.conflict_offsets[c + 1]; k++) {
        const uint32_t p = rc.conflict_entries[k];
        x[p] = x_buf[p];
    }
}

void Simulation::phase_after_timestep_dyn() {
    for (int ty = 0; ty < DYN_COUNT; ty++)
        for (uint32_t e = 0; e < dyn[ty].sp.active.size(); e++)
            if (dyn[ty].sp.active[e]) dyn_update_after_timestep((uint8_t)ty, dyn[ty].data[e], x, rs);
}

void Simulation::phase_dual_dyn() {
    for (int ty = 0; ty < DYN_COUNT; ty++)
        for (uint32_t e = 0; e < dyn[ty].sp.active.size(); e++)
            if (dyn[ty].sp.active[e]) dyn_update_after_iteration((uint8_t)ty, dyn[ty].data[e], x, rs);
}

void Simulation::phase_dual(uint32_t lo, uint32_t hi) {
    for (uint32_t e = lo; e < hi; e++) ground_update_after_iteration(ground[e], x);
}

void Simulation::phase_dual_rigid(uint32_t lo, uint32_t hi) {
    for (uint32_t e = lo; e < hi; e++) {
        if (e < rground.size()) rground_update_after_iteration(rground[e], x, rs);
        else if (e < rground.size() + joints.size()) joint_update_after_iteration(joints[e - rground.size()], x, rs);
        else joint2_update_after_iteration(joints2[e - rground.size() - joints.size()], x, rs);
    }
}

// _particleVelocityUpdate, VBDSolver.hpp:538-566
void Simulation::phase_velocity(uint32_t lo, uint32_t hi) {
    Real dt = params.dt;
    for (uint32_t p = lo; p < hi; p++) {
        v_prev[p] = v[p];
        v[p] = (x[p] - x_prev[p]) / dt;
        x_prev[p] = x[p];
        if (oriented(p)) {
            const int32_t r = rs.rot_of[p];
            rs.w_prev[r] = rs.w[r];
            rs.w[r] = minus_s3(rs.q[r], rs.q_prev[r]) / dt;
            rs.q_prev[r] = rs.q[r];
        }
    }
}

static inline Real omega_at(uint32_t i, Real rho, Real omega) {
    // VBDSolver.hpp:65-74
    if (i > 2) return 4 / (4 - rho * rho * omega);
    if (i == 2) return 2 / (2 - rho * rho);
    return 1;
}

void Simulation::step() {
    uint32_t n = (uint32_t)x.size();
    if (dynamic) {
        // CollisionDetector::detectCollisionsAndRecolor_Parallel over every
        // primitive: broad + narrow phase, contact lifecycle, recolouring
        detect_and_recolor();
    } else if (!tris.empty()) {
        // triangles only: the triangle-triangle narrow phase is a no-op, so
        // the broad phase is the whole detection
        bp.build(tris, x, v, params.dt);
        last_candidates = bp.traverse_leaves(0, bp.n);
    }
    const uint32_t n_rigid = (uint32_t)(rground.size() + joints.size() + joints2.size());
    phase_after_timestep(0, (uint32_t)nh.size(), 0, (uint32_t)ground.size());
    phase_after_timestep_rigid(0, n_rigid);
    if (dynamic) phase_after_timestep_dyn();
    phase_inertial(0, n);
    Real omega = 1;
    const bool target = params.tol > 0;
    const uint32_t iters = target ? params.max_iters : params.iters;
    const Real tol2 = target ? params.tol * params.tol * mg2() : 0;
    uint32_t used = iters;
    for (uint32_t i = 0; i < iters; i++) {
        omega = omega_at(i, params.rho, omega);
        for (uint32_t c = 0; c < coloring.num_colors; c++) {
            phase_color(c, 0, coloring.color_counts[c]);
            phase_commit(c);
        }
        phase_chebyshev(omega, 0, n);
        phase_dual(0, (uint32_t)ground.size());
        phase_dual_rigid(0, n_rigid);
        if (dynamic) phase_dual_dyn();
        if (target && (i + 1) % params.check_every == 0 && i + 1 < iters && residual2(0, n) <= tol2) {
            used = i + 1;
            break;
        }
    }
    last_iters = used;
    total_iters += used;
    phase_velocity(0, n);
}

Real Simulation::residual2(uint32_t lo, uint32_t hi) const {
    const Real dt = params.dt;
    Real r2 = 0;
    for (uint32_t p = lo; p < hi; p++) {
        if (fixed[p]) continue;
        V3 G;
        if (oriented(p)) {
            V6 G6;
            M6 H6;
            accumulate_vertex6(p, G6, H6);
            G = V3(G6.v[0], G6.v[1], G6.v[2]);
        } else {
            M3 H;
            accumulate_vertex(p, G, H);
        }
        const V3 r = (mass[p] / (dt * dt)) * (x[p] - y[p]) + G;
        r2 += dot(r, r);
    }
    return r2;
}

Real Simulation::mg2() const {
    Real s = 0;
    for (size_t p = 0; p < x.size(); p++)
        if (!fixed[p]) s += (mass[p] * params.g) * (mass[p] * params.g);
    return s;
}

// ---------------------------------------------------------------------------
// SPMD multithreaded step: same decomposition as the reference executor
// (WorkerThreadContext::computeStartEnd + SpinBarrier between phases).
// ---------------------------------------------------------------------------
namespace {
struct SpinBarrier {
    alignas(64) std::atomic<unsigned> arrived{0};
    alignas(64) std::atomic<unsigned> generation{0};
    unsigned count;
    explicit SpinBarrier(unsigned c) : count(c) {}
    void arrive_and_wait() {
        unsigned gen = generation.load(std::memory_order_acquire);
        if (arrived.fetch_add(1) == count - 1) {
            arrived.store(0, std::memory_order_relaxed);
            generation.fetch_add(1, std::memory_order_release);
            return;
        }
        while (generation.load(std::memory_order_acquire) == gen) _mm_pause();
    }
};
inline void range(uint32_t total, uint32_t idx, uint32_t nt, uint32_t& lo, uint32_t& hi) {
    lo = (uint32_t)((uint64_t)total * idx / nt);
    hi = (uint32_t)((uint64_t)total * (idx + 1) / nt);
}
}  // namespace

void Simulation::step_mt(uint32_t nthreads, uint32_t n_steps) {
    if (nthreads <= 1 || dynamic) {
        for (uint32_t s = 0; s < n_steps; s++) step();
        return;
    }
    SpinBarrier bar(nthreads);
    std::vector<uint64_t> cand(nthreads, 0);
    std::vector<Real> part(nthreads, 0);
    const bool target = params.tol > 0;
    const uint32_t iters = target ? params.max_iters : params.iters;
    const Real tol2 = target ? params.tol * params.tol * mg2() : 0;
    const uint32_t n_rigid = (uint32_t)(rground.size() + joints.size() + joints2.size());
    auto worker = [&](uint32_t w) {
        uint32_t n = (uint32_t)x.size();
        uint32_t lo, hi, lo2, hi2;
        for (uint32_t s = 0; s < n_steps; s++) {
            if (!tris.empty()) {
                if (w == 0) bp.build(tris, x, v, params.dt);
                bar.arrive_and_wait();
                range(bp.n, w, nthreads, lo, hi);
                cand[w] = bp.traverse_leaves(lo, hi);
                bar.arrive_and_wait();
                if (w == 0) {
                    last_candidates = 0;
                    for (auto cc : cand) last_candidates += cc;
                }
            }
            range((uint32_t)nh.size(), w, nthreads, lo, hi);
            phase_after_timestep(lo, hi, 0, 0);
            bar.arrive_and_wait();
            range((uint32_t)ground.size(), w, nthreads, lo2, hi2);
            phase_after_timestep(0, 0, lo2, hi2);
            bar.arrive_and_wait();
            if (n_rigid) {
                range(n_rigid, w, nthreads, lo, hi);
                phase_after_timestep_rigid(lo, hi);
                bar.arrive_and_wait();
            }
            range(n, w, nthreads, lo, hi);
            phase_inertial(lo, hi);
            bar.arrive_and_wait();
            Real omega = 1;
            uint32_t used = iters;
            for (uint32_t i = 0; i < iters; i++) {
                omega = omega_at(i, params.rho, omega);
                for (uint32_t c = 0; c < coloring.num_colors; c++) {
                    range(coloring.color_counts[c], w, nthreads, lo, hi);
                    phase_color(c, lo, hi);
                    bar.arrive_and_wait();
                }
                range(n, w, nthreads, lo, hi);
                phase_chebyshev(omega, lo, hi);
                bar.arrive_and_wait();
                if (!ground.empty()) {
                    range((uint32_t)ground.size(), w, nthreads, lo, hi);
                    phase_dual(lo, hi);
                    bar.arrive_and_wait();
                }
                if (n_rigid) {
                    range(n_rigid, w, nthreads, lo, hi);
                    phase_dual_rigid(lo, hi);
                    bar.arrive_and_wait();
                }
                if (target && (i + 1) % params.check_every == 0 && i + 1 < iters) {
                    range(n, w, nthreads, lo, hi);
                    part[w] = residual2(lo, hi);
                    bar.arrive_and_wait();
                    Real tot = 0;
                    for (Real v : part) tot += v;   // same order on every worker
                    bar.arrive_and_wait();          // partials are read before reuse
                    if (tot <= tol2) {
                        used = i + 1;
                        break;
                    }
                }
            }
            if (w == 0) {
                last_iters = used;
                total_iters += used;
            }
            range(n, w, nthreads, lo, hi);
            phase_velocity(lo, hi);
            bar.arrive_and_wait();
        }
    };
    std::vector<std::thread> th;
    for (uint32_t w = 1; w < nthreads; w++) th.emplace_back(worker, w);
    worker(0);
    for (auto& t : th) t.join();
}

// ---------------------------------------------------------------------------
// Force residual and incremental potential (new definitions)
//   r_v = m_v/dt^2 (x_v - y_v) + sum_e dE_e/dx_v  (AL terms with lambda^+)
//       = -RHS of VBDSolver.hpp:493 at the current iterate, free vertices only
//       (positional block for oriented particles)
//   rel = ||r||_2 / ||M g||_2 over free vertices
//   IP  = sum m/(2dt^2)|x-y|^2 + sum_e E_e + sum_ground (1/2 k C^2 + lambda C)
// ---------------------------------------------------------------------------
Stats Simulation::stats() const {
    Stats s;
    Real dt = params.dt;
    Real r2 = 0, mg2 = 0;
    uint32_t n = (uint32_t)x.size();
    for (uint32_t p = 0; p < n; p++) {
        if (fixed[p]) continue;
        V3 G;
        if (oriented(p)) {
            V6 G6; M6 H6;
            accumulate_vertex6(p, G6, H6);
            G = V3(G6.v[0], G6.v[1], G6.v[2]);
        } else {
            M3 H;
            accumulate_vertex(p, G, H);
        }
        V3 r = (mass[p] / (dt * dt)) * (x[p] - y[p]) + G;
        r2 += dot(r, r);
        Real mg = mass[p] * params.g;
        mg2 += mg * mg;
        V3 d = x[p] - y[p];
        s.ip += 0.5 * mass[p] / (dt * dt) * dot(d, d);
    }
    for (const auto& t : nh) s.ip += nh_energy(t, x, dt);
    for (const auto& e : cos) s.ip += cosserat_energy(e, x, rs);
    for (const auto& c : ground) {
        Real Cn, Ct, Cb;
        ground_eval(c, x, Cn, Ct, Cb);
        Real Cc = Cn - CONSTRAINT_ALPHA * c.C_n_prev;
        if (c.k_n * Cc + c.lambda_n > 0) s.ip += 0.5 * c.k_n * Cc * Cc + c.lambda_n * Cc;
        s.ip += 0.5 * c.k_t * Ct * Ct + c.lambda_t * Ct + 0.5 * c.k_b * Cb * Cb + c.lambda_b * Cb;
    }
    if (dynamic) {
        for (int ty = 0; ty < DYN_COUNT; ty++)
            for (uint32_t e = 0; e < dyn[ty].sp.active.size(); e++) {
                if (!dyn[ty].sp.active[e]) continue;
                const DynContact& c = dyn[ty].data[e];
                Real Cn, Ct, Cb;
                dyn_eval((uint8_t)ty, c, x, rs, Cn, Ct, Cb);
                Real Cc = Cn - CONSTRAINT_ALPHA * c.C_n_prev;
                if (c.k_n * Cc + c.lambda_n > 0) s.ip += 0.5 * c.k_n * Cc * Cc + c.lambda_n * Cc;
                s.ip += 0.5 * c.k_t * Ct * Ct + c.lambda_t * Ct + 0.5 * c.k_b * Cb * Cb + c.lambda_b * Cb;
            }
    }
    s.residual_abs = std::sqrt(r2);
    s.residual_rel = mg2 > 0 ? std::sqrt(r2) / std::sqrt(mg2) : 0;
    return s;
}

}  // namespace oracle
