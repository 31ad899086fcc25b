// dyn_device.hpp -- device side of the dynamic triangle-rigid / triangle-rod
// AL contacts: constraint values, the 3-DOF and 6-DOF vertex-block
// contributions, and the per-iteration / start-of-step multiplier and
// stiffness updates.  FP32.  Included by vbd_kernels.hip.
//
// Follows:
//   energy/collision/TriangleRigidCollisionEnergySolver.hpp:14-117
//   energy/collision/TriangleRodCollisionEnergySolver.hpp:14-150
//   energy/collision/CollisionConstraintEnergySolver.hpp:17-285
#pragma once

#include "internal.hpp"
#include "rot_math.hpp"

namespace avbd {
namespace dyn {

using namespace rot;

__device__ __forceinline__ float3 p3(float4 v) { return make_float3(v.x, v.y, v.z); }

// The particle state a contact reads: positions and, through rot_of, orientations.
struct View {
    const float4* x;
    const float4* q;
    const int32_t* rot_of;
};
__device__ __forceinline__ View view_of(const DevState& d) { return View{d.x, d.q, d.rot_of}; }
__device__ __forceinline__ uint32_t kind_of(const DynRec& c) { return c.kind_flags & 0xffu; }
__device__ __forceinline__ uint32_t flags_of(const DynRec& c) { return c.kind_flags >> DYN_FLAG_SHIFT; }

// Contact-point geometry of one record at the current state: the point on
// the triangle, the point on the body / rod, and (rod) the interpolated frame.
struct Geo {
    float3 diff;     // cp_body - cp_tri
    float4 q;        // body orientation (kind 0) or q_mid (kind 1)
    float3 q_diff;   // kind 1: q2 [-] q1
};

__device__ __forceinline__ Geo geometry(const View& d, const DynRec& c) {
    Geo g;
    if (kind_of(c) == DYN_RIGID_RIGID) {
        // cp_2 - cp_1 with cp_k = x_k + R_k cpl_k (cpl_1 in cpl, cpl_2 in bary)
        const float4 q1 = d.q[d.rot_of[c.idx[0]]], q2 = d.q[d.rot_of[c.idx[1]]];
        const float3 cp1 = add(p3(d.x[c.idx[0]]), qrot(q1, make_float3(c.cpl.x, c.cpl.y, c.cpl.z)));
        const float3 cp2 = add(p3(d.x[c.idx[1]]), qrot(q2, make_float3(c.bary.x, c.bary.y, c.bary.z)));
        g.q = q1;
        g.q_diff = make_float3(0.f, 0.f, 0.f);
        g.diff = sub(cp2, cp1);
        return g;
    }
    const float3 x0 = p3(d.x[c.idx[0]]), x1 = p3(d.x[c.idx[1]]), x2 = p3(d.x[c.idx[2]]);
    const float3 cp_tri = add(add(scl(c.bary.x, x0), scl(c.bary.y, x1)), scl(c.bary.z, x2));
    const float3 cpl = make_float3(c.cpl.x, c.cpl.y, c.cpl.z);
    float3 cp;
    if (kind_of(c) == DYN_VERTEX_TRI) {
        // vertex-triangle (new): the vertex itself, no orientation
        g.q = make_float4(0.f, 0.f, 0.f, 1.f);
        g.q_diff = make_float3(0.f, 0.f, 0.f);
        cp = p3(d.x[c.idx[3]]);
    } else if (kind_of(c) == DYN_TRI_RIGID) {
        g.q = d.q[d.rot_of[c.idx[3]]];
        g.q_diff = make_float3(0.f, 0.f, 0.f);
        cp = add(p3(d.x[c.idx[3]]), qrot(g.q, cpl));
    } else {
        const float s = c.bary.w;
        const float4 q1 = d.q[d.rot_of[c.idx[3]]], q2 = d.q[d.rot_of[c.idx[4]]];
        g.q_diff = minus_s3(q2, q1);
        g.q = plus_s3(q1, scl(s, g.q_diff));
        const float3 pm = add(scl(1.f - s, p3(d.x[c.idx[3]])), scl(s, p3(d.x[c.idx[4]])));
        cp = add(pm, qrot(g.q, cpl));
    }
    g.diff = sub(cp, cp_tri);
    return g;
}

__device__ __forceinline__ void eval(const View& d, const DynRec& c, float& Cn, float& Ct, float& Cb) {
    const Geo g = geometry(d, c);
    Cn = dotv(p3(c.n), g.diff);
    Ct = dotv(p3(c.t), g.diff);
    Cb = dotv(p3(c.b), g.diff);
}

// lambda+ of the three directions with the friction cone (accumulate, :219-240)
struct Mult {
    float Ccorr, lnp, lt, lb;
};
__device__ __forceinline__ Mult multipliers(const DynRec& c, float Cn, float Ct, float Cb) {
    Mult m;
    m.Ccorr = Cn - kAlpha * c.cpl.w;
    m.lnp = fmaxf(0.f, c.k.x * m.Ccorr + c.lam.x);
    m.lt = c.k.y * Ct + c.lam.y;
    m.lb = c.k.z * Cb + c.lam.z;
    const float mag = sqrtf(m.lt * m.lt + m.lb * m.lb);
    const float mu = (flags_of(c) & GF_USE_STATIC) ? c.t.w : c.b.w;
    const float lmax = mu * m.lnp;
    if (mag > lmax) {
        const float nrm = mag + 1e-8f;
        m.lt = m.lt / nrm * lmax;
        m.lb = m.lb / nrm * lmax;
    }
    return m;
}

// 3-DOF contribution of a triangle vertex (local 0..2): gradients -bary n, t, b;
// and of the vertex of a vertex-triangle contact (local 3): +n, t, b.
template <class Sym>
__device__ __forceinline__ void accumulate3(const View& d, const DynRec& c, uint32_t local, float3& G, Sym& H) {
    float Cn, Ct, Cb;
    eval(d, c, Cn, Ct, Cb);
    const Mult m = multipliers(c, Cn, Ct, Cb);
    const float beta = local == 0 ? c.bary.x : (local == 1 ? c.bary.y : (local == 2 ? c.bary.z : -1.f));
    const float3 n = p3(c.n), t = p3(c.t), b = p3(c.b);
    G = add(G, scl(-beta, add(add(scl(m.lnp, n), scl(m.lt, t)), scl(m.lb, b))));
    const float b2 = beta * beta;
    const float kn = c.k.x * b2, kt = c.k.y * b2, kb = c.k.z * b2;
    H.a00 += kn * n.x * n.x + kt * t.x * t.x + kb * b.x * b.x;
    H.a11 += kn * n.y * n.y + kt * t.y * t.y + kb * b.y * b.y;
    H.a22 += kn * n.z * n.z + kt * t.z * t.z + kb * b.z * b.z;
    H.a01 += kn * n.x * n.y + kt * t.x * t.y + kb * b.x * b.y;
    H.a02 += kn * n.x * n.z + kt * t.x * t.z + kb * b.x * b.z;
    H.a12 += kn * n.y * n.z + kt * t.y * t.z + kb * b.y * b.z;
}

// sym([v_b]x [c]x) = (c v_b^T + v_b c^T) / 2 - (v_b . c) I, added with weight w
__device__ __forceinline__ void add_rot_hess(Sym6& H, float w, float3 vb, float3 c) {
    const float dd = dotv(vb, c);
    const float a[3] = {vb.x, vb.y, vb.z}, e[3] = {c.x, c.y, c.z};
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = i; j < 3; j++) {
            float h = 0.5f * (e[i] * a[j] + a[i] * e[j]);
            if (i == j) h -= dd;
            H.a[s6(3 + i, 3 + j)] += w * h;
        }
}

// 6-DOF contribution of the body (kind 0, local 3) or a rod node (kind 1,
// local 3 / 4).  The tangent gradients of the oriented side are zero in the
// reference (TriangleRigidCollisionEnergySolver.hpp:90,
// TriangleRodCollisionEnergySolver.hpp:130).
// Rigid-rigid contact, body `local` (0 first, sign -1; 1 second, sign +1):
// per direction dir = n, t, b the gradient sign * [dir ; cpl x R^T dir], the
// Gauss-Newton terms k g g^T, and the rotational blocks
// sign * sym(R^T [dir]x R [cpl]x) weighted by the unclamped k C + lambda,
// with the friction multipliers clamped to the cone (new; the same terms as
// the body side of the triangle-rigid contact, plus friction).
__device__ __forceinline__ void accumulate6_rigid_rigid(const View& d, const DynRec& c, uint32_t local, Vec6& G,
                                                        Sym6& H) {
    const Geo g = geometry(d, c);
    const float3 n = p3(c.n), t = p3(c.t), b = p3(c.b);
    const float Cn = dotv(n, g.diff), Ct = dotv(t, g.diff), Cb = dotv(b, g.diff);
    const Mult m = multipliers(c, Cn, Ct, Cb);
    const float sg = local == 0 ? -1.f : 1.f;
    const uint32_t body = c.idx[local];
    const float3 cpl = local == 0 ? make_float3(c.cpl.x, c.cpl.y, c.cpl.z) : make_float3(c.bary.x, c.bary.y, c.bary.z);
    const M3f R = qmat(d.q[d.rot_of[body]]);
    float gn[6], gt[6], gb[6];
    auto grad = [&](float3 dir, float* gv) {
        const float3 r = crossv(cpl, mtvec(R, dir));
        gv[0] = sg * dir.x; gv[1] = sg * dir.y; gv[2] = sg * dir.z;
        gv[3] = sg * r.x; gv[4] = sg * r.y; gv[5] = sg * r.z;
    };
    grad(n, gn);
    grad(t, gt);
    grad(b, gb);
#pragma unroll
    for (int i = 0; i < 6; i++) G.v[i] += m.lnp * gn[i] + m.lt * gt[i] + m.lb * gb[i];
    sym6_outer(H, gn, c.k.x);
    sym6_outer(H, gt, c.k.y);
    sym6_outer(H, gb, c.k.z);
    const float an = c.k.x * m.Ccorr + c.lam.x, at = c.k.y * Ct + c.lam.y, ab = c.k.z * Cb + c.lam.z;
    add_rot_hess(H, sg * an, mtvec(R, n), cpl);
    add_rot_hess(H, sg * at, mtvec(R, t), cpl);
    add_rot_hess(H, sg * ab, mtvec(R, b), cpl);
}

__device__ __forceinline__ void accumulate6(const View& d, const DynRec& c, uint32_t local, Vec6& G, Sym6& H) {
    if (kind_of(c) == DYN_RIGID_RIGID) {
        accumulate6_rigid_rigid(d, c, local, G, H);
        return;
    }
    const Geo g = geometry(d, c);
    const float3 n = p3(c.n), t = p3(c.t), b = p3(c.b);
    const float Cn = dotv(n, g.diff), Ct = dotv(t, g.diff), Cb = dotv(b, g.diff);
    const Mult m = multipliers(c, Cn, Ct, Cb);
    const float3 cpl = make_float3(c.cpl.x, c.cpl.y, c.cpl.z);
    const M3f R = qmat(g.q);
    const float3 nb = mtvec(R, n);          // R^T n
    float gn[6];
    if (kind_of(c) == DYN_TRI_RIGID) {
        // [n ; -(R [cp]x)^T n] = [n ; cp x R^T n]
        const float3 r = crossv(cpl, nb);
        gn[0] = n.x; gn[1] = n.y; gn[2] = n.z;
        gn[3] = r.x; gn[4] = r.y; gn[5] = r.z;
        // rotational Hessian blocks sym(R^T [v]x R [cp]x), weights k C + lambda (unclamped)
        const float an = c.k.x * m.Ccorr + c.lam.x, at = c.k.y * Ct + c.lam.y, ab = c.k.z * Cb + c.lam.z;
        add_rot_hess(H, an, nb, cpl);
        add_rot_hess(H, at, mtvec(R, t), cpl);
        add_rot_hess(H, ab, mtvec(R, b), cpl);
    } else {
        // R_rloc = -R [cp]x; gradient -(R_rloc dRs/dR)^T n = dRs/dR^T (R [cp]x)^T n
        const float s = c.bary.w;
        const M3f gam_inv = jr_inv(g.q_diff);
        const M3f gam = jr(scl(s, g.q_diff));
        M3f dR;
        float scaling;
        if (local == 3) {
            const M3f e = mtrans(exp_so3(scl(s, g.q_diff)));
            const M3f g2 = mmul(gam, mtrans(gam_inv));
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) dR.m[i][j] = e.m[i][j] - s * g2.m[i][j];
            scaling = 1.f - s;
        } else {
            const M3f g2 = mmul(gam, gam_inv);
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) dR.m[i][j] = s * g2.m[i][j];
            scaling = s;
        }
        const float3 v1 = scl(-1.f, crossv(cpl, nb));   // (R [cp]x)^T n
        const float3 r = mtvec(dR, v1);
        gn[0] = scaling * n.x; gn[1] = scaling * n.y; gn[2] = scaling * n.z;
        gn[3] = r.x; gn[4] = r.y; gn[5] = r.z;
    }
#pragma unroll
    for (int i = 0; i < 6; i++) G.v[i] += m.lnp * gn[i];
    sym6_outer(H, gn, c.k.x);
}

// CollisionConstraintEnergySolver::updateAfterIteration (:60-134) at the
// constraint values Cn, Ct, Cb of the current state.
__device__ __forceinline__ void al_update(DynRec& c, float Cn, float Ct, float Cb) {
    const float Ccorr = Cn - kAlpha * c.cpl.w;
    const float lnp = c.k.x * Ccorr + c.lam.x;
    const float ltp = c.k.y * Ct + c.lam.y;
    const float lbp = c.k.z * Cb + c.lam.z;
    uint32_t fl = flags_of(c);
    if (lnp > 0.f) c.k.x += kBeta * fabsf(Ccorr);
    c.lam.x = fmaxf(0.f, lnp);
    const float mag = sqrtf(ltp * ltp + lbp * lbp);
    float lmax = ((fl & GF_USE_STATIC) ? c.t.w : c.b.w) * c.lam.x;
    if (mag < lmax) {
        c.k.y += kBeta * fabsf(Ct);
        c.k.z += kBeta * fabsf(Cb);
        c.lam.y = ltp;
        c.lam.z = lbp;
        fl |= GF_UNCLAMPED;
    } else {
        if (fl & GF_USE_STATIC) {
            fl &= ~GF_USE_STATIC;
            lmax = c.b.w * lnp;
        }
        c.lam.y = ltp / (mag + 1e-8f) * lmax;
        c.lam.z = lbp / (mag + 1e-8f) * lmax;
        fl &= ~GF_UNCLAMPED;
    }
    c.kind_flags = (c.kind_flags & 0xffu) | (fl << DYN_FLAG_SHIFT);
}

__device__ __forceinline__ void update_after_iteration(const View& d, DynRec& c) {
    float Cn, Ct, Cb;
    eval(d, c, Cn, Ct, Cb);
    al_update(c, Cn, Ct, Cb);
}

// CollisionConstraintEnergySolver::updateAfterTimeStep (:141-167) with the
// normal constraint value Cn of the start-of-step state (read only when the
// normal multiplier is active).
__device__ __forceinline__ void al_step_begin(DynRec& c, float Cn) {
    if (c.lam.x > 0.f) c.cpl.w = Cn;
    uint32_t fl = flags_of(c);
    if ((fl & GF_UNCLAMPED) && !(fl & GF_USE_STATIC)) fl |= GF_USE_STATIC;
    c.kind_flags = (c.kind_flags & 0xffu) | (fl << DYN_FLAG_SHIFT);
    c.lam.x *= kAlpha * kGamma;
    c.lam.y *= kGamma;
    c.lam.z *= kGamma;
    const float ks = c.n.w;
    c.k.x = fmaxf(kGamma * c.k.x, ks);
    c.k.y = fmaxf(kGamma * c.k.y, ks);
    c.k.z = fmaxf(kGamma * c.k.z, ks);
}

__device__ __forceinline__ void update_after_timestep(const View& d, DynRec& c) {
    float Cn = 0.f, Ct, Cb;
    if (c.lam.x > 0.f) eval(d, c, Cn, Ct, Cb);
    al_step_begin(c, Cn);
}

// Eigen's unitOrthogonal for 3-vectors with the double-precision threshold
// of the reference (Real = double): (-y, x, 0)/|(x, y)| unless x and y are
// both below 1e-12 |z|.
__device__ __forceinline__ float3 unit_orthogonal(float3 v) {
    const bool xs = fabsf(v.x) <= fabsf(v.z) * 1e-12f, ys = fabsf(v.y) <= fabsf(v.z) * 1e-12f;
    if (!xs || !ys) {
        const float inv = 1.f / sqrtf(v.x * v.x + v.y * v.y);
        return make_float3(-v.y * inv, v.x * inv, 0.f);
    }
    const float inv = 1.f / sqrtf(v.y * v.y + v.z * v.z);
    return make_float3(0.f, -v.z * inv, v.y * inv);
}

// CollisionConstraintEnergySolver::updateCollisionFrame (:17-53)
__device__ __forceinline__ void update_frame(DynRec& c, float3 nn) {
    float3 t = p3(c.t), b = p3(c.b);
    const float3 f = add(scl(c.lam.y, t), scl(c.lam.z, b));
    const float3 tp = sub(t, scl(dotv(t, nn), nn));
    const float m = normv(tp);
    if (m > 1e-8f) {
        t = scl(1.f / m, tp);
        b = crossv(nn, t);
    } else {
        t = unit_orthogonal(nn);
        b = crossv(nn, t);
    }
    c.n = make_float4(nn.x, nn.y, nn.z, c.n.w);
    c.t = make_float4(t.x, t.y, t.z, c.t.w);
    c.b = make_float4(b.x, b.y, b.z, c.b.w);
    c.lam.y = dotv(f, t);
    c.lam.z = dotv(f, b);
}

}  // namespace dyn
}  // namespace avbd
