// contact.cpp -- FP64 oracle of the dynamic contact pipeline (see contact.hpp).
// Test infrastructure only; see oracle.hpp.
#include "contact.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>

#include "simulation.hpp"

namespace oracle {

// ---------------------------------------------------------------------------
// Geometry helpers
// ---------------------------------------------------------------------------

// Eigen's unitOrthogonal for 3-vectors (OrthoMethods.h, unitOrthogonal_selector<.,3>):
// unless x and y are both negligible against z (|.| <= 1e-12 |z|, Eigen's
// dummy precision for double), return (-y, x, 0) / |(x, y)|; else
// (0, -z, y) / |(y, z)|.
V3 unit_orthogonal(V3 n) {
    const Real prec = 1e-12;
    const bool x_small = std::fabs(n.x) <= std::fabs(n.z) * prec;
    const bool y_small = std::fabs(n.y) <= std::fabs(n.z) * prec;
    if (!x_small || !y_small) {
        const Real inv = 1.0 / std::sqrt(n.x * n.x + n.y * n.y);
        return {-n.y * inv, n.x * inv, 0.0};
    }
    const Real inv = 1.0 / std::sqrt(n.y * n.y + n.z * n.z);
    return {0.0, -n.z * inv, n.y * inv};
}

// common/Math.hpp:323-327
void complete_basis(V3 n, V3& t, V3& b) {
    t = unit_orthogonal(n);
    b = cross(n, t);
}

// common/Math.hpp:421-438
V3 barycentric(V3 p, V3 a, V3 b, V3 c) {
    const V3 v0 = b - a, v1 = c - a, v2 = p - a;
    const Real d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1);
    const Real d20 = dot(v2, v0), d21 = dot(v2, v1);
    const Real denom = d00 * d11 - d01 * d01;
    const Real v = (d11 * d20 - d01 * d21) / denom;
    const Real w = (d00 * d21 - d01 * d20) / denom;
    const Real u = 1 - v - w;
    return {u, v, w};
}

// common/Math.hpp:446-499 (Embree's region test)
V3 closest_point_triangle(V3 p, V3 a, V3 b, V3 c) {
    const V3 ab = b - a, ac = c - a, ap = p - a;
    const Real d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) return a;
    const V3 bp = p - b;
    const Real d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) return b;
    const V3 cp = p - c;
    const Real d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) return c;
    const Real vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        const Real v = d1 / (d1 - d3);
        return a + v * ab;
    }
    const Real vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        const Real v = d2 / (d2 - d6);
        return a + v * ac;
    }
    const Real va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        const Real v = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        return b + v * (c - b);
    }
    const Real denom = 1 / (va + vb + vc);
    const Real v = vb * denom, w = vc * denom;
    return a + v * ab + w * ac;
}

static Real clamp01(Real v) { return std::min(std::max(v, Real(0)), Real(1)); }

// common/Math.hpp:508-559 (Ericson's ClosestPtSegmentSegment)
void closest_segment_segment(V3 p1, V3 q1, V3 p2, V3 q2, Real& s, Real& t) {
    const V3 d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
    const Real a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    const Real EPS = 1e-8;
    if (a <= EPS && e <= EPS) {
        s = t = 0;
        return;
    }
    if (a <= EPS) {
        s = 0;
        t = clamp01(f / e);
        return;
    }
    const Real c = dot(d1, r);
    if (e <= EPS) {
        t = 0;
        s = clamp01(-c / a);
        return;
    }
    const Real b = dot(d1, d2);
    const Real denom = a * e - b * b;
    s = denom != 0 ? clamp01((b * f - c * e) / denom) : 0;
    t = (b * s + f) / e;
    if (t < 0) {
        t = 0;
        s = clamp01(-c / a);
    } else if (t > 1) {
        t = 1;
        s = clamp01((b - c) / a);
    }
}

// common/Math.hpp:561-591
bool point_in_triangle(V3 p, V3 a, V3 b, V3 c) {
    const V3 v0 = b - a, v1 = c - a, v2 = p - a;
    const Real d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1);
    const Real d20 = dot(v2, v0), d21 = dot(v2, v1);
    const Real denom = d00 * d11 - d01 * d01;
    if (std::fabs(denom) < 1e-12) return false;
    const Real v = (d11 * d20 - d01 * d21) / denom;
    const Real w = (d00 * d21 - d01 * d20) / denom;
    const Real u = 1.0 - v - w;
    const Real EPS = 1e-8;
    return u >= -EPS && v >= -EPS && w >= -EPS;
}

// common/Math.hpp:593-681.  As in the reference, the edge test measures the
// distance from the segment point at s to the edge point at the *segment's*
// parameter s ((1-s) q0 + s q1), not at the edge parameter t; t only sets
// the barycentric weights.
void closest_segment_triangle(V3 p0, V3 p1, V3 a, V3 b, V3 c, Real& s_out, V3& barys_out, Real& dist_out) {
    const Real EPS = 1e-8;
    const V3 dir = p1 - p0, ab = b - a, ac = c - a;
    const V3 nrm = cross(ab, ac);
    const Real nl2 = dot(nrm, nrm);
    s_out = 0;
    barys_out = V3(1, 0, 0);
    dist_out = std::numeric_limits<Real>::max();
    if (nl2 > EPS) {
        const Real denom = dot(nrm, dir);
        if (std::fabs(denom) > EPS) {
            const Real s = dot(nrm, a - p0) / denom;
            if (s >= 0.0 && s <= 1.0) {
                const V3 p = p0 + s * dir;
                if (point_in_triangle(p, a, b, c)) {
                    s_out = s;
                    barys_out = barycentric(p, a, b, c);
                    dist_out = 0;
                    return;
                }
            }
        }
    }
    auto test_edge = [&](V3 q0, V3 q1, int i0, int i1) {
        Real s, t;
        closest_segment_segment(p0, p1, q0, q1, s, t);
        const V3 cp = (1 - s) * p0 + s * p1;
        const V3 cq = (1 - s) * q0 + s * q1;
        const V3 d = cp - cq;
        const Real d2 = dot(d, d);
        if (d2 < dist_out * dist_out) {
            s_out = s;
            barys_out = V3(0, 0, 0);
            barys_out[i0] = 1 - t;
            barys_out[i1] = t;
            dist_out = std::sqrt(d2);
        }
    };
    test_edge(a, b, 0, 1);
    test_edge(a, c, 0, 2);
    test_edge(b, c, 1, 2);
    const V3 c0 = closest_point_triangle(p0, a, b, c);
    const V3 e0 = p0 - c0;
    const Real sq0 = dot(e0, e0);
    if (sq0 < dist_out * dist_out) {
        s_out = 0;
        barys_out = barycentric(c0, a, b, c);
        dist_out = std::sqrt(sq0);
    }
    const V3 c1 = closest_point_triangle(p1, a, b, c);
    const V3 e1 = p1 - c1;
    const Real sq1 = dot(e1, e1);
    if (sq1 < dist_out * dist_out) {
        s_out = 1;
        barys_out = barycentric(c1, a, b, c);
        dist_out = std::sqrt(sq1);
    }
}

// Algorithm::goldenSectionSearch (common/Algorithm.hpp:253-299) with
// lerp(a, b, u) = (1 - u) a + u b (:11-15); T = Real or V3.
namespace {
inline Real gss_width(Real v) { return std::fabs(v); }
inline Real gss_width(V3 v) { return norm(v); }
template <class T>
inline T lerp(const T& a, const T& b, Real u) {
    return (1 - u) * a + u * b;
}
inline Real lerp(Real a, Real b, Real u) { return (1 - u) * a + u * b; }

template <class T, class F>
T golden(const T& ts, const T& te, F&& f, Real tol = 1e-3) {
    const Real phi = 1.61803398875;
    const Real r = 1 / phi;
    const Real s = 1 - r;
    Real a0 = 0, a1 = s, a2 = r, a3 = 1;
    T t0 = ts, t1 = lerp(ts, te, a1), t2 = lerp(ts, te, a2), t3 = te;
    Real f0 = f(t0), f1 = f(t1), f2 = f(t2), f3 = f(t3);
    while (gss_width(t3 - t0) > tol) {
        if (std::min(f0, f1) < std::min(f2, f3)) {
            a3 = a2; t3 = t2; f3 = f2;
            a2 = a1; t2 = t1; f2 = f1;
            a1 = r * a2 + s * a0;
            t1 = lerp(ts, te, a1);
            f1 = f(t1);
        } else {
            a0 = a1; t0 = t1; f0 = f1;
            a1 = a2; t1 = t2; f1 = f2;
            a2 = r * a1 + s * a3;
            t2 = lerp(ts, te, a2);
            f2 = f(t2);
        }
    }
    const T tm = lerp(ts, te, 0.5 * (a0 + a3));
    const Real fm = f(tm);
    if (f0 < fm && f0 < f3) return t0;
    if (fm < f3) return tm;
    return t3;
}

// Eigen Quaternion::inverse: conjugate / squared norm
inline Quat qinverse(const Quat& q) {
    const Real n2 = q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z;
    return {q.w / n2, -q.x / n2, -q.y / n2, -q.z / n2};
}
}  // namespace

Real golden_section_quadratic(Real a, Real b, Real c) {
    return golden(a, b, [&](Real t) { return (t - c) * (t - c); });
}

// collision/DetectedCollision.hpp:43-66
uint64_t contact_key(uint32_t type, uint32_t a, uint32_t b, uint32_t code) {
    if (a > b) std::swap(a, b);
    const uint32_t IDX = (1u << 26) - 1, CODE = (1u << 7) - 1;
    uint64_t key = (uint64_t)type & ((1u << 5) - 1);
    key = (key << 26) | (a & IDX);
    key = (key << 26) | (b & IDX);
    key = (key << 7) | (code & CODE);
    return key;
}

// collision/CollisionPrimitivePool.hpp:160-236 (sphere: SDF::localBounds, collision/SDF.hpp:16-23)
void prim_bounds(const Prim& pr, const std::vector<V3>& x, const std::vector<V3>& v, const RotState& rs, Real dt,
                 double box[6], Real tri_pad) {
    for (int k = 0; k < 3; k++) {
        box[k] = std::numeric_limits<double>::max();
        box[3 + k] = std::numeric_limits<double>::lowest();
    }
    auto expand_pt = [&](V3 p) {
        for (int k = 0; k < 3; k++) {
            box[k] = std::min(box[k], p[k]);
            box[3 + k] = std::max(box[3 + k], p[k]);
        }
    };
    auto expand_box = [&](const double o[6]) {
        for (int k = 0; k < 3; k++) {
            box[k] = std::min(box[k], o[k]);
            box[3 + k] = std::max(box[3 + k], o[3 + k]);
        }
    };
    if (pr.type == PRIM_TRIANGLE) {
        for (int c = 0; c < 3; c++) {
            expand_pt(x[pr.idx[c]]);
            expand_pt(x[pr.idx[c]] + dt * v[pr.idx[c]]);
        }
        for (int k = 0; k < 3; k++) {   // self collision (new): widened by the contact margin
            box[k] -= tri_pad;
            box[3 + k] += tri_pad;
        }
    } else if (pr.type == PRIM_ROD_SEGMENT) {
        const V3 p1 = x[pr.idx[0]], p2 = x[pr.idx[1]];
        const Real r = pr.radius;
        double cur[6], nxt[6];
        for (int k = 0; k < 3; k++) {
            cur[k] = std::min(p1[k], p2[k]) - r;
            cur[3 + k] = std::max(p1[k], p2[k]) + r;
        }
        const V3 n1 = p1 + dt * v[pr.idx[0]], n2 = p2 + dt * v[pr.idx[1]];
        for (int k = 0; k < 3; k++) {
            nxt[k] = std::min(n1[k], n2[k]) - r;
            nxt[3 + k] = std::max(n1[k], n2[k]) + r;
        }
        expand_box(cur);
        expand_box(nxt);
    } else {
        // local box [-r, r]^3: centre 0, half extent r
        const uint32_t p = pr.idx[0];
        const Real r = pr.radius;
        const Quat& q = rs.rot(p);
        auto rotated = [&](const Quat& qq, V3 centre, double out[6]) {
            const M3 R = qmat(qq);
            for (int i = 0; i < 3; i++) {
                const Real e = std::fabs(R.m[i][0]) * r + std::fabs(R.m[i][1]) * r + std::fabs(R.m[i][2]) * r;
                out[i] = centre[i] - e;
                out[3 + i] = centre[i] + e;
            }
        };
        double cur[6], nxt[6];
        rotated(q, x[p], cur);
        const Quat qn = plus_s3(q, dt * rs.w[rs.rot_of[p]]);
        rotated(qn, x[p] + dt * v[p], nxt);
        expand_box(cur);
        expand_box(nxt);
    }
}

// ---------------------------------------------------------------------------
// TombstonePool slot stack (common/TombstonePool.hpp:64-110)
// ---------------------------------------------------------------------------
uint32_t SlotPool::alloc() {
    if (top >= capacity) throw std::runtime_error("Pool has exceeded capacity");
    const uint32_t idx = top++;
    const uint32_t slot = slot_at(idx);
    if (active.size() <= slot) {
        active.resize(slot + 1, 0);
        gen.resize(slot + 1, 0);
    }
    active[slot] = 1;
    gen[slot]++;
    count++;
    highest = std::max(highest, slot);
    return slot;
}

void SlotPool::free(uint32_t slot) {
    if (slot >= active.size() || !active[slot]) throw std::runtime_error("freeSlot called on inactive slot");
    // the reference lowers highest_index only past inactive slots, and tests
    // the slot being freed before clearing it: it never moves down
    active[slot] = 0;
    count--;
    const uint32_t idx = --top;
    while (slots.size() <= idx) slots.push_back((uint32_t)slots.size());
    slots[idx] = slot;
}

uint32_t DynPool::add(const DynContact& c) {
    const uint32_t slot = sp.alloc();
    if (data.size() <= slot) data.resize(slot + 1);
    data[slot] = c;
    return slot;
}

// ---------------------------------------------------------------------------
// Dynamic contact energies
// ---------------------------------------------------------------------------
namespace {
struct RodMid {
    V3 q_diff;
    Quat q_mid;
    V3 p_mid;
};
RodMid rod_mid(const DynContact& c, const std::vector<V3>& x, const RotState& rs) {
    RodMid m;
    const Quat& q1 = rs.rot(c.idx[3]);
    const Quat& q2 = rs.rot(c.idx[4]);
    m.q_diff = minus_s3(q2, q1);
    m.q_mid = plus_s3(q1, c.s * m.q_diff);
    m.p_mid = (1 - c.s) * x[c.idx[3]] + c.s * x[c.idx[4]];
    return m;
}
M3 sym(const M3& a) {
    M3 s;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) s.m[i][j] = 0.5 * (a.m[i][j] + a.m[j][i]);
    return s;
}
}  // namespace

// TriangleRigidCollisionConstraintSolver::evaluateConstraint (:14-42) and
// TriangleRodCollisionConstraintSolver::evaluateConstraint (:14-49)
void dyn_eval(uint8_t type, const DynContact& c, const std::vector<V3>& x, const RotState& rs, Real& Cn, Real& Ct,
              Real& Cb) {
    if (type == DYN_RIGID_RIGID) {
        const V3 cp1 = x[c.idx[0]] + qrot(rs.rot(c.idx[0]), c.cpl);
        const V3 cp2 = x[c.idx[1]] + qrot(rs.rot(c.idx[1]), c.cpl2);
        const V3 d = cp2 - cp1;
        Cn = dot(c.n, d);
        Ct = dot(c.t, d);
        Cb = dot(c.b, d);
        return;
    }
    const V3 cp_tri = c.barys.x * x[c.idx[0]] + c.barys.y * x[c.idx[1]] + c.barys.z * x[c.idx[2]];
    V3 cp;
    if (type == DYN_VERTEX_TRI) {
        cp = x[c.idx[3]];
    } else if (type == DYN_TRI_RIGID) {
        cp = x[c.idx[3]] + qrot(rs.rot(c.idx[3]), c.cpl);
    } else {
        const RodMid m = rod_mid(c, x, rs);
        cp = m.p_mid + qrot(m.q_mid, c.cpl);
    }
    const V3 d = cp - cp_tri;
    Cn = dot(c.n, d);
    Ct = dot(c.t, d);
    Cb = dot(c.b, d);
}

// constraintGradientHessian<DOF> of both solvers.
void dyn_constraint(uint8_t type, const DynContact& c, const std::vector<V3>& x, const RotState& rs, uint32_t local,
                    int dof, Real& Cn, Real& Ct, Real& Cb, Real gn[6], Real gt[6], Real gb[6], M3& hn, M3& ht,
                    M3& hb) {
    dyn_eval(type, c, x, rs, Cn, Ct, Cb);
    for (int i = 0; i < 6; i++) gn[i] = gt[i] = gb[i] = 0;
    hn = ht = hb = M3();
    if (type == DYN_RIGID_RIGID) {
        // C = dir . (x_2 + R_2 cpl_2 - x_1 - R_1 cpl_1): per body sign * [dir ; -(R [cpl]x)^T dir],
        // rotational Hessians sign * sym(R^T [dir]x R [cpl]x), as the tri-rigid body side
        if (dof != 6 || local > 1) throw std::runtime_error("dynamic contact: local index out of bounds");
        const Real sg = local == 0 ? -1.0 : 1.0;
        const uint32_t body = c.idx[local];
        const V3 cpl = local == 0 ? c.cpl : c.cpl2;
        const M3 R = qmat(rs.rot(body));
        const M3 Sc = skew3(cpl);
        const M3 Rr = matmul(R, Sc);
        auto grad = [&](V3 dir, Real g[6]) {
            const V3 r = matvec(transpose(Rr), dir);
            for (int i = 0; i < 3; i++) {
                g[i] = sg * dir[i];
                g[3 + i] = -sg * r[i];
            }
        };
        auto hblock = [&](V3 v) {
            M3 h = sym(matmul(matmul(matmul(transpose(R), skew3(v)), R), Sc));
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) h.m[i][j] *= sg;
            return h;
        };
        grad(c.n, gn);
        grad(c.t, gt);
        grad(c.b, gb);
        hn = hblock(c.n);
        ht = hblock(c.t);
        hb = hblock(c.b);
        return;
    }
    if (type == DYN_VERTEX_TRI) {
        if (dof != 3 || local > 3) throw std::runtime_error("dynamic contact: local index out of bounds");
        const Real w = local == 3 ? 1.0 : -c.barys[(int)local];
        for (int i = 0; i < 3; i++) {
            gn[i] = w * c.n[i];
            gt[i] = w * c.t[i];
            gb[i] = w * c.b[i];
        }
        return;
    }
    if (dof == 3) {
        if (local >= 3) throw std::runtime_error("dynamic contact: local index out of bounds for a positional particle");
        const Real bary = c.barys[(int)local];
        for (int i = 0; i < 3; i++) {
            gn[i] = -bary * c.n[i];
            gt[i] = -bary * c.t[i];
            gb[i] = -bary * c.b[i];
        }
        return;
    }
    if (type == DYN_TRI_RIGID) {
        // TriangleRigidCollisionEnergySolver.hpp:77-102: tangent gradients are
        // zeroed; the rotational Hessians are sym(R^T [v]x R [cp]x)
        const M3 R = qmat(rs.rot(c.idx[3]));
        const M3 Sc = skew3(c.cpl);
        const M3 Rr = matmul(R, Sc);
        const V3 rn = matvec(transpose(Rr), c.n);
        for (int i = 0; i < 3; i++) {
            gn[i] = c.n[i];
            gn[3 + i] = -rn[i];
        }
        auto hblock = [&](V3 v) { return sym(matmul(matmul(matmul(transpose(R), skew3(v)), R), Sc)); };
        hn = hblock(c.n);
        ht = hblock(c.t);
        hb = hblock(c.b);
    } else {
        // TriangleRodCollisionEnergySolver.hpp:92-136 (Hessians zero)
        if (local != 3 && local != 4) throw std::runtime_error("dynamic contact: local index out of bounds");
        const RodMid m = rod_mid(c, x, rs);
        const M3 R = qmat(m.q_mid);
        const M3 Sc = skew3(c.cpl);
        M3 Rr = matmul(R, Sc);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) Rr.m[i][j] = -Rr.m[i][j];
        const M3 gam_inv = jr_inv(m.q_diff);
        const M3 gam = jr(c.s * m.q_diff);
        M3 dR;
        Real scaling;
        if (local == 3) {
            const M3 e = transpose(exp_so3(c.s * m.q_diff));
            const M3 g2 = matmul(gam, transpose(gam_inv));
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) dR.m[i][j] = e.m[i][j] - c.s * g2.m[i][j];
            scaling = 1 - c.s;
        } else {
            const M3 g2 = matmul(gam, gam_inv);
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) dR.m[i][j] = c.s * g2.m[i][j];
            scaling = c.s;
        }
        const M3 M = matmul(Rr, dR);
        const V3 mn = matvec(transpose(M), c.n);
        for (int i = 0; i < 3; i++) {
            gn[i] = scaling * c.n[i];
            gn[3 + i] = -mn[i];
        }
    }
}

namespace {
// CollisionConstraintEnergySolver::accumulate<DOF> (:179-285) on the
// constraint data of one particle; dof 3 or 6.
void al_accumulate(const DynContact& c, int dof, Real Cn, Real Ct, Real Cb, const Real gn[6], const Real gt[6],
                   const Real gb[6], const M3& hn, const M3& ht, const M3& hb, Real* G, Real* H /* dof x dof */) {
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
    for (int i = 0; i < dof; i++) G[i] += ln_plus * gn[i] + lt * gt[i] + lb * gb[i];
    const Real an = c.k_n * C_corr_n + c.lambda_n, at = c.k_t * Ct + c.lambda_t, ab = c.k_b * Cb + c.lambda_b;
    for (int i = 0; i < dof; i++)
        for (int j = 0; j < dof; j++) {
            Real h = c.k_n * gn[i] * gn[j] + c.k_t * gt[i] * gt[j] + c.k_b * gb[i] * gb[j];
            if (dof == 6 && i >= 3 && j >= 3)
                h += an * hn.m[i - 3][j - 3] + at * ht.m[i - 3][j - 3] + ab * hb.m[i - 3][j - 3];
            H[i * dof + j] += h;
        }
}
}  // namespace

void dyn_accumulate3(uint8_t type, const DynContact& c, const std::vector<V3>& x, const RotState& rs, uint32_t local,
                     V3& G, M3& H) {
    Real Cn, Ct, Cb, gn[6], gt[6], gb[6];
    M3 hn, ht, hb;
    dyn_constraint(type, c, x, rs, local, 3, Cn, Ct, Cb, gn, gt, gb, hn, ht, hb);
    Real g[3] = {G.x, G.y, G.z}, h[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) h[3 * i + j] = H.m[i][j];
    al_accumulate(c, 3, Cn, Ct, Cb, gn, gt, gb, hn, ht, hb, g, h);
    G = V3(g[0], g[1], g[2]);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) H.m[i][j] = h[3 * i + j];
}

void dyn_accumulate6(uint8_t type, const DynContact& c, const std::vector<V3>& x, const RotState& rs, uint32_t local,
                     V6& G, M6& H) {
    Real Cn, Ct, Cb, gn[6], gt[6], gb[6];
    M3 hn, ht, hb;
    dyn_constraint(type, c, x, rs, local, 6, Cn, Ct, Cb, gn, gt, gb, hn, ht, hb);
    Real h[36];
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) h[6 * i + j] = H.m[i][j];
    al_accumulate(c, 6, Cn, Ct, Cb, gn, gt, gb, hn, ht, hb, G.v, h);
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) H.m[i][j] = h[6 * i + j];
}

// CollisionConstraintEnergySolver::updateAfterIteration (:60-134)
void dyn_update_after_iteration(uint8_t type, DynContact& c, const std::vector<V3>& x, const RotState& rs) {
    Real Cn, Ct, Cb;
    dyn_eval(type, c, x, rs, Cn, Ct, Cb);
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
        c.unclamped = true;
    } else {
        if (c.use_static) {
            c.use_static = false;
            lmax = c.mu_k * ln_plus;
        }
        c.lambda_t = lt_plus / (mag + 1e-8) * lmax;
        c.lambda_b = lb_plus / (mag + 1e-8) * lmax;
        c.unclamped = false;
    }
}

// CollisionConstraintEnergySolver::updateAfterTimeStep (:141-167)
void dyn_update_after_timestep(uint8_t type, DynContact& c, const std::vector<V3>& x, const RotState& rs) {
    if (c.lambda_n > 0) {
        Real Cn, Ct, Cb;
        dyn_eval(type, c, x, rs, Cn, Ct, Cb);
        c.C_n_prev = Cn;
    }
    if (c.unclamped && !c.use_static) c.use_static = true;
    c.lambda_n *= CONSTRAINT_ALPHA * STIFFNESS_GAMMA;
    c.lambda_t *= STIFFNESS_GAMMA;
    c.lambda_b *= STIFFNESS_GAMMA;
    c.k_n = std::max(STIFFNESS_GAMMA * c.k_n, c.k_start);
    c.k_t = std::max(STIFFNESS_GAMMA * c.k_t, c.k_start);
    c.k_b = std::max(STIFFNESS_GAMMA * c.k_b, c.k_start);
}

// CollisionConstraintEnergySolver::updateCollisionFrame (:17-53)
void update_collision_frame(DynContact& c, V3 nn) {
    const V3 f = c.lambda_t * c.t + c.lambda_b * c.b;
    const V3 tp = c.t - dot(c.t, nn) * nn;
    const Real m = norm(tp);
    if (m > 1e-8) {
        c.t = tp / m;
        c.b = cross(nn, c.t);
    } else {
        complete_basis(nn, c.t, c.b);
    }
    c.n = nn;
    c.lambda_t = dot(f, c.t);
    c.lambda_b = dot(f, c.b);
}

// ---------------------------------------------------------------------------
// Narrow phase: CollisionDetector::_triangleSDF_CCD (:689-876), sphere SDF
// ---------------------------------------------------------------------------
bool sphere_sphere(V3 xa, const Quat& qa, Real ra, V3 xb, const Quat& qb, Real rb, Real margin, V3& normal, V3& cpl_a,
                   V3& cpl_b) {
    const V3 d = xa - xb;
    const Real dist = norm(d);
    if (!(dist < ra + rb + margin)) return false;
    normal = dist > 1e-12 ? d / dist : V3(1, 0, 0);
    cpl_a = qrot(qinverse(qa), -ra * normal);
    cpl_b = qrot(qinverse(qb), rb * normal);
    return true;
}

bool vertex_face(V3 p, V3 a, V3 b, V3 c, Real margin, V3& normal, V3& barys) {
    const V3 cp = closest_point_triangle(p, a, b, c);
    if (!(norm(p - cp) < margin)) return false;
    const V3 f = cross(b - a, c - a);
    const Real fm = norm(f);
    const V3 fu = fm > 0 ? (1.0 / fm) * f : V3(0, 1, 0);
    normal = dot(fu, p - cp) >= 0 ? -1.0 * fu : fu;
    barys = barycentric(cp, a, b, c);
    return true;
}

bool tri_sphere_ccd(const V3 tri[3], const V3 tv[3], V3 rb_pos, const Quat& rb_rot, V3 rb_v, V3 rb_w, Real radius,
                    Real dt, Real margin, V3& normal, V3& cp_barys, V3& cpl) {
    auto pos_t = [&](Real t) { return rb_pos + t * rb_v; };
    auto rot_t = [&](Real t) { return qmul(rb_rot, exp_s3(t * rb_w)); };
    auto lin_vel_on_rb = [&](V3 xt, Real t) {
        const V3 wg = qrot(rot_t(t), rb_w);
        return rb_v + cross(wg, xt - pos_t(t));
    };
    auto rel_vel = [&](V3 xt, V3 vt, Real t) { return qrot(qinverse(rot_t(t)), vt - lin_vel_on_rb(xt, t)); };
    auto local_pos = [&](V3 xt, Real t) { return qrot(qinverse(rot_t(t)), xt - pos_t(t)); };
    auto sdf = [&](V3 p) { return norm(p) - radius; };                 // SDF::evaluate, sphere
    auto grad = [&](V3 p) {                                            // SDF::gradient, sphere
        const Real d = norm(p);
        return d > 1e-8 ? p / d : V3(1, 0, 0);
    };
    auto eval0 = [&](V3 si, V3 vi) { return dot(rel_vel(si, vi, 0), grad(local_pos(si, 0))); };
    const Real e1 = eval0(tri[0], tv[0]), e2 = eval0(tri[1], tv[1]), e3 = eval0(tri[2], tv[2]);
    Real t_start = 0, t_end = dt, t = 0, u, v, w;
    if (e1 < e2 && e1 < e3) {
        u = 1; v = 0; w = 0;
    } else if (e2 < e1 && e2 < e3) {
        u = 0; v = 1; w = 0;
    } else {
        u = 0; v = 0; w = 1;
    }
    auto interp = [&](Real tt) { return u * (tri[0] + tt * tv[0]) + v * (tri[1] + tt * tv[1]) + w * (tri[2] + tt * tv[2]); };
    auto interp_vel = [&]() { return u * tv[0] + v * tv[1] + w * tv[2]; };
    auto sd_time = [&](Real tt) { return sdf(local_pos(interp(tt), tt)); };
    auto ud_time = [&](Real tt) { return std::fabs(sd_time(tt)); };
    auto sd_point = [&](V3 p) { return sdf(p); };
    Real sdf_x = 0;
    V3 grad_x, x_ti;
    for (int iter = 0; iter < 16; iter++) {
        const Real t_old = t;
        V3 xg = interp(t), vg = interp_vel();
        x_ti = local_pos(xg, t);
        V3 v_ti = rel_vel(xg, vg, t);
        sdf_x = sdf(x_ti);
        if (sdf_x <= 0) {
            t_end = std::min(t, t_end);
            t = golden(t_start, t, ud_time);
        } else {
            const Real vdg = dot(v_ti, grad(x_ti));
            const Real d = vdg > 0 ? -1 : 1;
            if (d < 0) t = golden(t_start, t, sd_time);
            else t = golden(t, t_end, sd_time);
        }
        xg = interp(t);
        x_ti = local_pos(xg, t);
        sdf_x = sdf(x_ti);
        grad_x = grad(x_ti);
        if (sdf_x <= 0) t_end = std::min(t, t_end);
        const V3 l1 = local_pos(tri[0] + t * tv[0], t);
        const V3 l2 = local_pos(tri[1] + t * tv[1], t);
        const V3 l3 = local_pos(tri[2] + t * tv[2], t);
        const Real s1 = dot(l1, grad_x), s2 = dot(l2, grad_x), s3 = dot(l3, grad_x);
        V3 sp;
        if (s1 < s2 && s1 < s3) sp = l1;
        else if (s2 < s1 && s2 < s3) sp = l2;
        else sp = l3;
        const V3 x_new = golden(x_ti, sp, sd_point);
        const V3 bc = barycentric(x_new, l1, l2, l3);
        u = bc.x;
        v = bc.y;
        w = bc.z;
        if (std::fabs(t_old - t) < 1e-6 && norm(x_new - x_ti) < 1e-6) break;
    }
    cp_barys = V3(u, v, w);
    cpl = x_ti - sdf_x * grad_x;
    normal = qrot(rot_t(t), grad_x);
    return sdf_x < margin;
}

// ---------------------------------------------------------------------------
// Simulation: detection, lifecycle, dynamic adjacency, recolouring
// ---------------------------------------------------------------------------

uint32_t Simulation::add_collision_triangles(uint32_t n, const uint32_t* t) {
    const uint32_t first = (uint32_t)prims.size();
    const uint32_t obj = next_object++;
    for (uint32_t i = 0; i < n; i++) {
        Prim p;
        p.type = PRIM_TRIANGLE;
        p.np = 3;
        for (int k = 0; k < 3; k++) p.idx[k] = t[3 * i + k];
        p.object = obj;
        prims.push_back(p);
    }
    tris.insert(tris.end(), t, t + 3 * (size_t)n);
    return first;
}

uint32_t Simulation::add_collision_segments(uint32_t n_nodes, const uint32_t* nodes, Real radius) {
    const uint32_t first = (uint32_t)prims.size();
    const uint32_t obj = next_object++;
    for (uint32_t i = 0; i + 1 < n_nodes; i++) {
        Prim p;
        p.type = PRIM_ROD_SEGMENT;
        p.np = 2;
        p.idx[0] = nodes[i];
        p.idx[1] = nodes[i + 1];
        p.object = obj;
        p.radius = radius;
        prims.push_back(p);
    }
    return first;
}

uint32_t Simulation::add_collision_sphere(uint32_t p_idx, Real radius) {
    Prim p;
    p.type = PRIM_SPHERE;
    p.np = 1;
    p.idx[0] = p_idx;
    p.object = next_object++;
    p.radius = radius;
    prims.push_back(p);
    return (uint32_t)prims.size() - 1;
}

void Simulation::enable_dynamic_contacts(Real margin, uint32_t capacity) {
    dynamic = true;
    collision_margin = margin;
    for (int k = 0; k < DYN_COUNT; k++) {
        dyn[k] = DynPool();
        dyn[k].sp.capacity = capacity;
    }
    const uint32_t n = (uint32_t)x.size();
    rc = Recoloring();
    rc.last_dirty_round.assign(n, Recoloring::SENTINEL);
    rc.candidate.assign(n, 0);
    rc.conflict_offsets.assign(coloring.num_colors + 1, 0);
}

// _shouldSkip (collision/CollisionDetector.cpp:132-164)
static bool should_skip(const Prim& a, const Prim& b) {
    if (a.object != b.object) return false;
    for (int i = 0; i < a.np; i++)
        for (int j = 0; j < b.np; j++)
            if (a.idx[i] == b.idx[j]) return true;
    return false;
}

// _processPotentialCollision (:166-215) with _triangleSphere (:504-536) and
// _triangleRod (:601-666).
void Simulation::process_pair(uint32_t a, uint32_t b, std::vector<Detected>& out) {
    if (should_skip(prims[a], prims[b])) return;
    uint8_t ta = prims[a].type, tb = prims[b].type;
    if (ta > tb) {
        std::swap(ta, tb);
        std::swap(a, b);
    }
    const Real dt = params.dt;
    if (ta == PRIM_TRIANGLE && tb == PRIM_TRIANGLE) {
        if (!self_collision) return;   // the reference's _triangleTriangle is empty
        // _triangleTriangle, vertex-face only (new; the reference's is empty):
        // each vertex of either triangle against the other triangle.  A
        // vertex is tested from its owner, the lowest-numbered triangle that
        // holds it, so that a vertex-face pair is found once; the key is the
        // reference's generateKey(TriangleTriangle_VertexFace, owner, face,
        // code) with code = the vertex's corner in the owner, + 3 when the
        // owner is the higher-numbered primitive.
        for (int side = 0; side < 2; side++) {
            const uint32_t own = side == 0 ? a : b, face = side == 0 ? b : a;
            const Prim& O = prims[own];
            const Prim& F = prims[face];
            for (uint32_t k = 0; k < 3; k++) {
                const uint32_t pv = O.idx[k];
                if (vertex_owner[pv] != own) continue;
                if (pv == F.idx[0] || pv == F.idx[1] || pv == F.idx[2]) continue;
                V3 nrm, bc;
                if (!vertex_face(x[pv], x[F.idx[0]], x[F.idx[1]], x[F.idx[2]], collision_margin, nrm, bc)) continue;
                Detected d;
                d.type = DET_VERTEX_FACE;
                d.key = contact_key(DET_VERTEX_FACE, own, face, own < face ? k : 3 + k);
                d.gen1 = F.gen;
                d.gen2 = O.gen;
                d.normal = nrm;
                for (int q = 0; q < 3; q++) d.tri[q] = F.idx[q];
                d.barys = bc;
                d.vtx = pv;
                out.push_back(d);
            }
        }
        return;
    }
    if (ta == PRIM_ROD_SEGMENT && tb == PRIM_ROD_SEGMENT) return;  // _rodRod: no-op
    if (ta == PRIM_SPHERE && tb == PRIM_SPHERE) {
        // new: the reference throws here (collision/CollisionDetector.cpp:678-686)
        if (a > b) std::swap(a, b);
        const Prim& A = prims[a];
        const Prim& B = prims[b];
        const uint32_t pa = A.idx[0], pb = B.idx[0];
        V3 nrm, ca, cb;
        if (sphere_sphere(x[pa], rs.rot(pa), A.radius, x[pb], rs.rot(pb), B.radius, collision_margin, nrm, ca, cb)) {
            Detected d;
            d.type = DET_RIGID_RIGID;
            d.key = contact_key(DET_RIGID_RIGID, a, b, 0);
            d.gen1 = A.gen;
            d.gen2 = B.gen;
            d.normal = nrm;
            d.rb = pa;
            d.rb2 = pb;
            d.cpl = ca;
            d.cpl2 = cb;
            out.push_back(d);
        }
        return;
    }
    if (ta == PRIM_TRIANGLE && tb == PRIM_SPHERE) {
        const Prim& tr = prims[a];
        const Prim& sp = prims[b];
        const uint32_t rb = sp.idx[0];
        const V3 T[3] = {x[tr.idx[0]], x[tr.idx[1]], x[tr.idx[2]]};
        const V3 TV[3] = {v[tr.idx[0]], v[tr.idx[1]], v[tr.idx[2]]};
        V3 nrm, bc, cpl;
        if (tri_sphere_ccd(T, TV, x[rb], rs.rot(rb), v[rb], rs.w[rs.rot_of[rb]], sp.radius, dt, collision_margin, nrm,
                           bc, cpl)) {
            Detected d;
            d.type = DET_TRI_RIGID;
            d.key = contact_key(DET_TRI_RIGID, a, b, 0);
            d.gen1 = tr.gen;
            d.gen2 = sp.gen;
            d.normal = nrm;
            for (int k = 0; k < 3; k++) d.tri[k] = tr.idx[k];
            d.barys = bc;
            d.rb = rb;
            d.cpl = cpl;
            out.push_back(d);
        }
        return;
    }
    if (ta == PRIM_TRIANGLE && tb == PRIM_ROD_SEGMENT) {
        const Prim& tr = prims[a];
        const Prim& sg = prims[b];
        const V3 v1 = x[tr.idx[0]], v2 = x[tr.idx[1]], v3 = x[tr.idx[2]];
        const V3 s1 = x[sg.idx[0]], s2 = x[sg.idx[1]];
        Real cs, dist;
        V3 bc;
        closest_segment_triangle(s1, s2, v1, v2, v3, cs, bc, dist);
        if (dist < sg.radius + collision_margin) {
            const V3 cp_line = (1 - cs) * s1 + cs * s2;
            const V3 cp_tri = bc.x * v1 + bc.y * v2 + bc.z * v3;
            const V3 diff = cp_tri - cp_line;
            const Real dm = norm(diff);
            const V3 nrm = dm > 1e-8 ? diff / dm : V3(1, 0, 0);
            const Quat& q1 = rs.rot(sg.idx[0]);
            const Quat& q2 = rs.rot(sg.idx[1]);
            const Quat qm = plus_s3(q1, cs * minus_s3(q2, q1));
            Detected d;
            d.type = DET_TRI_ROD;
            d.key = contact_key(DET_TRI_ROD, a, b, 0);
            d.gen1 = tr.gen;
            d.gen2 = sg.gen;
            d.normal = nrm;
            for (int k = 0; k < 3; k++) d.tri[k] = tr.idx[k];
            d.rod[0] = sg.idx[0];
            d.rod[1] = sg.idx[1];
            d.barys = bc;
            d.s = cs;
            d.cpl = qrot(qinverse(qm), sg.radius * nrm);
            out.push_back(d);
        }
        return;
    }
    // rod segment vs sphere: a sphere welded to a rod's node does not collide
    // with that rod (the shared-particle rule of _shouldSkip); any other pair
    // is unsupported, as in the reference
    for (int i = 0; i < prims[a].np; i++)
        for (int j = 0; j < prims[b].np; j++)
            if (prims[a].idx[i] == prims[b].idx[j]) return;
    throw std::runtime_error("Unsupported collision!");
}

static uint8_t dyn_type_of(uint8_t det_type) {
    if (det_type == DET_VERTEX_FACE) return DYN_VERTEX_TRI;
    return det_type == DET_TRI_RIGID ? DYN_TRI_RIGID : (det_type == DET_TRI_ROD ? DYN_TRI_ROD : DYN_RIGID_RIGID);
}

// ColorList::markParticleDirty (common/ColorList.hpp:366-377)
void Simulation::mark_dirty(uint32_t p) {
    if (rc.last_dirty_round[p] != 0) {
        rc.last_dirty_round[p] = 0;
        rc.touched_this_frame.push_back(p);
        rc.dirty.push_back(p);
    }
}

// _addCollision (collision/CollisionDetector.cpp:329-430)
void Simulation::add_contact(Detected& d) {
    const Real dt = params.dt;
    DynContact c;
    uint8_t ty;
    Real min_mass;
    if (d.type == DET_RIGID_RIGID) {
        ty = DYN_RIGID_RIGID;
        min_mass = std::min(mass[d.rb], mass[d.rb2]);
        c.idx[0] = d.rb;
        c.idx[1] = d.rb2;
        c.cpl2 = d.cpl2;
    } else if (d.type == DET_VERTEX_FACE) {
        ty = DYN_VERTEX_TRI;
        min_mass = std::min({mass[d.tri[0]], mass[d.tri[1]], mass[d.tri[2]], mass[d.vtx]});
        c.idx[3] = d.vtx;
    } else if (d.type == DET_TRI_RIGID) {
        ty = DYN_TRI_RIGID;
        min_mass = std::min({mass[d.tri[0]], mass[d.tri[1]], mass[d.tri[2]], mass[d.rb]});
        c.idx[3] = d.rb;
    } else {
        ty = DYN_TRI_ROD;
        min_mass = std::min({mass[d.tri[0]], mass[d.tri[1]], mass[d.tri[2]], mass[d.rod[0]], mass[d.rod[1]]});
        c.idx[3] = d.rod[0];
        c.idx[4] = d.rod[1];
        c.s = d.s;
    }
    if (ty != DYN_RIGID_RIGID)
        for (int k = 0; k < 3; k++) c.idx[k] = d.tri[k];
    const Real k_start = min_mass / (dt * dt);
    // CollisionConstraintEnergyPool::addEnergy (energy/collision/CollisionConstraintEnergyPool.hpp:48-69)
    c.k_start = c.k_n = c.k_t = c.k_b = k_start;
    c.lambda_n = c.lambda_t = c.lambda_b = 0;
    c.C_n_prev = 0;
    c.mu_s = 0.4;
    c.mu_k = 0.2;
    c.use_static = false;
    c.unclamped = false;
    c.n = d.normal;
    complete_basis(d.normal, c.t, c.b);
    c.barys = d.barys;
    c.cpl = d.cpl;
    d.e_idx = dyn[ty].add(c);
    if (ty == DYN_RIGID_RIGID) {
        mark_dirty(d.rb);
        mark_dirty(d.rb2);
        return;
    }
    for (int k = 0; k < 3; k++) mark_dirty(d.tri[k]);
    if (ty == DYN_VERTEX_TRI) {
        mark_dirty(d.vtx);
    } else if (ty == DYN_TRI_RIGID) {
        mark_dirty(d.rb);
    } else {
        mark_dirty(d.rod[0]);
        mark_dirty(d.rod[1]);
    }
}

void Simulation::remove_contact(const Detected& d) { dyn[dyn_type_of(d.type)].sp.free(d.e_idx); }

// _updateCollision (:454-501)
void Simulation::update_contact(const Detected& d) {
    DynContact& c = dyn[dyn_type_of(d.type)].data[d.e_idx];
    if (d.type == DET_TRI_ROD) c.s = d.s;
    c.cpl = d.cpl;
    c.cpl2 = d.cpl2;
    c.barys = d.barys;
    // a vertex-face contact keeps the side it was created with
    const bool flip = d.type == DET_VERTEX_FACE && dot(d.normal, c.n) < 0;
    update_collision_frame(c, flip ? -1.0 * d.normal : d.normal);
}

// _handleDetectedCollisions (:272-327): merge the key-sorted lists.
void Simulation::handle_detected() {
    cur_order.resize(cur_det.size());
    for (uint32_t i = 0; i < cur_order.size(); i++) cur_order[i] = i;
    std::stable_sort(cur_order.begin(), cur_order.end(),
                     [&](uint32_t a, uint32_t b) { return cur_det[a].key < cur_det[b].key; });
    size_t ci = 0, pi = 0;
    while (ci < cur_det.size() && pi < prev_det.size()) {
        Detected& cur = cur_det[cur_order[ci]];
        Detected& prev = prev_det[prev_order[pi]];
        if (cur.key < prev.key) {
            add_contact(cur);
            ci++;
        } else if (cur.key > prev.key) {
            remove_contact(prev);
            pi++;
        } else {
            if (cur.gen1 == prev.gen1 && cur.gen2 == prev.gen2) {
                cur.e_idx = prev.e_idx;
                update_contact(cur);
            } else {
                remove_contact(prev);
                add_contact(cur);
            }
            ci++;
            pi++;
        }
    }
    for (; ci < cur_det.size(); ci++) add_contact(cur_det[cur_order[ci]]);
    for (; pi < prev_det.size(); pi++) remove_contact(prev_det[prev_order[pi]]);
}

// DynamicParticleAdjacency::buildAdjacency (common/DynamicParticleAdjacency.hpp:33-118)
void Simulation::build_dynamic_adjacency() {
    const uint32_t n = (uint32_t)x.size();
    std::vector<std::array<uint32_t, DYN_COUNT>> val(n, {0, 0, 0, 0});
    std::vector<uint32_t> deg(n, 0);
    const uint32_t npe[DYN_COUNT] = {4, 5, 2, 4};
    for (int ty = 0; ty < DYN_COUNT; ty++) {
        const DynPool& P = dyn[ty];
        for (uint32_t e = 0; e < P.sp.active.size(); e++) {
            if (!P.sp.active[e]) continue;
            for (uint32_t k = 0; k < npe[ty]; k++) {
                val[P.data[e].idx[k]][ty]++;
                deg[P.data[e].idx[k]] += npe[ty] - 1;
            }
        }
    }
    dadj.e_offsets.assign(n + 1, {0, 0, 0, 0});
    dadj.p_offsets.assign(n + 1, 0);
    for (uint32_t i = 0; i < n; i++) {
        for (int ty = 1; ty < DYN_COUNT; ty++) dadj.e_offsets[i][ty] = dadj.e_offsets[i][ty - 1] + val[i][ty - 1];
        dadj.e_offsets[i + 1][0] = dadj.e_offsets[i][DYN_COUNT - 1] + val[i][DYN_COUNT - 1];
        dadj.p_offsets[i + 1] = dadj.p_offsets[i] + deg[i];
    }
    dadj.e_entries.resize(dadj.e_offsets[n][0]);
    dadj.p_neighbors.resize(dadj.p_offsets[n]);
    auto ecur = dadj.e_offsets;
    auto pcur = dadj.p_offsets;
    for (int ty = 0; ty < DYN_COUNT; ty++) {
        const DynPool& P = dyn[ty];
        for (uint32_t e = 0; e < P.sp.active.size(); e++) {
            if (!P.sp.active[e]) continue;
            for (uint32_t k = 0; k < npe[ty]; k++) {
                const uint32_t p = P.data[e].idx[k];
                dadj.e_entries[ecur[p][ty]++] = {(uint8_t)ty, e, (uint16_t)k};
                for (uint32_t k2 = 0; k2 < npe[ty]; k2++)
                    if (k2 != k) dadj.p_neighbors[pcur[p]++] = P.data[e].idx[k2];
            }
        }
    }
}

// ColorList::incrementalRecoloring + finalizeConflicts (common/ColorList.hpp:381-522)
void Simulation::incremental_recoloring() {
    const uint32_t n = (uint32_t)x.size();
    auto for_each_neighbor = [&](uint32_t p, auto&& f) {
        for (uint32_t k = adj.p_offsets[p]; k < adj.p_offsets[p + 1]; k++) f(adj.p_neighbors[k]);
        for (uint32_t k = dadj.p_offsets[p]; k < dadj.p_offsets[p + 1]; k++) f(dadj.p_neighbors[k]);
    };
    for (uint32_t p : rc.touched_conflicted) is_conflicted[p] = 0;
    rc.touched_conflicted.clear();
    uint32_t iter = 0;
    while (!rc.dirty.empty() && iter < 6) {
        for (uint32_t p : rc.dirty) {
            uint64_t bits[4] = {0, 0, 0, 0};   // ColorBitset<4>: 256 colours
            bool must = false;
            for_each_neighbor(p, [&](uint32_t nb) {
                const uint32_t cn = coloring.color[nb];
                bits[cn / 64] |= 1ull << (cn % 64);
                if (cn == coloring.color[p] && nb < p) must = true;
            });
            if (must) {
                uint32_t c = 256;
                for (uint32_t w = 0; w < 4 && c == 256; w++)
                    if (~bits[w]) c = w * 64 + (uint32_t)__builtin_ctzll(~bits[w]);
                if (c == 256) throw std::runtime_error("Ran out of colors!");
                rc.candidate[p] = c;
                if (c >= coloring.num_colors) coloring.num_colors++;
            } else {
                rc.candidate[p] = coloring.color[p];
            }
        }
        rc.next_dirty.clear();
        for (uint32_t p : rc.dirty) {
            if (rc.candidate[p] != coloring.color[p]) {
                coloring.color[p] = rc.candidate[p];
                for_each_neighbor(p, [&](uint32_t nb) {
                    if (rc.last_dirty_round[nb] != iter + 1) {
                        rc.last_dirty_round[nb] = iter + 1;
                        rc.touched_this_frame.push_back(nb);
                        rc.next_dirty.push_back(nb);
                    }
                });
            }
        }
        rc.dirty.swap(rc.next_dirty);
        iter++;
    }
    if (!rc.dirty.empty()) {
        for (uint32_t p : rc.dirty) {
            for_each_neighbor(p, [&](uint32_t nb) {
                if (coloring.color[nb] == coloring.color[p]) {
                    if (!is_conflicted[p]) {
                        is_conflicted[p] = 1;
                        rc.touched_conflicted.push_back(p);
                    }
                    if (!is_conflicted[nb]) {
                        is_conflicted[nb] = 1;
                        rc.touched_conflicted.push_back(nb);
                    }
                }
            });
        }
        rc.conflict_offsets.assign(coloring.num_colors + 1, 0);
        for (uint32_t p : rc.touched_conflicted) rc.conflict_offsets[coloring.color[p] + 1]++;
        for (uint32_t c = 0; c < coloring.num_colors; c++) rc.conflict_offsets[c + 1] += rc.conflict_offsets[c];
        rc.conflict_entries.resize(rc.conflict_offsets[coloring.num_colors]);
        auto cur = rc.conflict_offsets;
        for (uint32_t p : rc.touched_conflicted) rc.conflict_entries[cur[coloring.color[p]]++] = p;
    } else {
        rc.conflict_offsets.assign(coloring.num_colors + 1, 0);
        rc.conflict_entries.clear();
    }
    rebuild_work_list(n, coloring);
    for (uint32_t p : rc.touched_this_frame) rc.last_dirty_round[p] = Recoloring::SENTINEL;
    rc.touched_this_frame.clear();
    for (uint32_t p : rc.dirty) rc.last_dirty_round[p] = 0;
}

// CollisionDetector::detectCollisionsAndRecolor_Parallel (:57-129)
void Simulation::detect_and_recolor() {
    prev_det.swap(cur_det);
    prev_order.swap(cur_order);
    cur_det.clear();
    cur_order.clear();
    last_candidates = 0;
    if (!prims.empty()) {
        std::vector<double> boxes(6 * prims.size());
        const Real pad = self_collision ? collision_margin : 0.0;
        for (size_t i = 0; i < prims.size(); i++)
            prim_bounds(prims[i], x, v, rs, params.dt, boxes.data() + 6 * i, pad);
        // owner of every triangle vertex: the lowest-numbered triangle holding it
        vertex_owner.assign(x.size(), UINT32_MAX);
        for (uint32_t i = (uint32_t)prims.size(); i-- > 0;)
            if (prims[i].type == PRIM_TRIANGLE)
                for (int k = 0; k < 3; k++) vertex_owner[prims[i].idx[k]] = i;
        bp.build_boxes(boxes);
        std::vector<std::pair<uint32_t, uint32_t>> pairs;
        bp.traverse_pairs(pairs);
        last_candidates = pairs.size();
        for (const auto& pr : pairs) process_pair(bp.sorted[pr.first], bp.sorted[pr.second], cur_det);
    }
    handle_detected();
    build_dynamic_adjacency();
    incremental_recoloring();
}

}  // namespace oracle
