// vbd_fp64.hip -- the static solve path (Neo-Hookean tets + ground contact)
// with the particle state in FP64 (AVBD_FLAG_FP64_STATE).
//
// Why: an FP32 position resolves 6e-8 m at |x| ~ 0.5 m; times a vertex
// stiffness of E h (3e4 N/m on the 32^3 cube at E = 1e6) that is a force of
// 1e-3 .. 1e-2 of a vertex's weight, so an FP32 state cannot reach a relative
// force residual of 1e-4 however many iterations run.  Here positions, the
// inertial target and the Chebyshev history are FP64 (double4: xyz + mass),
// and everything a vertex block solve computes from them -- deformation
// gradient, stress, gradient and Hessian sums, the 3x3 solve, the ground
// constraint, the Chebyshev step and the force residual -- is FP64.  The
// material records (Q, V lambda, V mu, alpha) and the AL records stay FP32:
// they are constants of the model within a step, so they change the system
// being solved by a rounding, not the precision the iteration can reach.
// Velocities stay FP32 (they only form the next step's inertial target).
// The FP32 position stream d.x is kept as a mirror of the state at step
// boundaries (xyz rounded, mass in w) for the broad phase and the
// warm-start features, which read FP32 positions.
//
// Same phase structure and formulas as the FP32 kernels of vbd_kernels.hip
// (VBDSolver::solve_Parallel, simulation/VBDSolver.hpp:48-103; the Neo-Hookean
// accumulate, energy/elastic/NeoHookeanEnergySolver.hpp:121-227; the ground AL
// terms, energy/collision/CollisionConstraintEnergySolver.hpp:60-285).
#include "dyn_device.hpp"
#include "internal.hpp"
#include "kernarg.hpp"
#include "warm_device.hpp"

namespace avbd {
namespace {

struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 dd3(double4 v) { return D3{v.x, v.y, v.z}; }
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 operator*(double s, D3 a) { return D3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ double dot3(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 cross3(D3 a, D3 b) {
    return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

struct SymD {
    double a00, a11, a22, a01, a02, a12;
};

__device__ __forceinline__ float4 mirror(double4 p) {
    return make_float4((float)p.x, (float)p.y, (float)p.z, (float)p.w);
}

__device__ __forceinline__ void load_rec(const TetRec* __restrict__ recs, uint32_t k, uint4& idx, float4& r1,
                                         float4& r2, float4& r3) {
    const uint4* p = reinterpret_cast<const uint4*>(recs + k);
    idx = p[0];
    const float4* f = reinterpret_cast<const float4*>(p);
    r1 = f[1];
    r2 = f[2];
    r3 = f[3];
}

// Deformation gradient columns f_c = sum_r d_r Q[r][c] of a record
// (r1 = q00 q01 q02 q10, r2 = q11 q12 q20 q21, r3 = q22 VL VM alpha).
__device__ __forceinline__ void deform(D3 p0, D3 p1, D3 p2, D3 p3, float4 r1, float4 r2, float4 r3, D3& f0, D3& f1,
                                       D3& f2) {
    const D3 d0 = p0 - p3, d1 = p1 - p3, d2 = p2 - p3;
    f0 = (double)r1.x * d0 + (double)r1.w * d1 + (double)r2.z * d2;
    f1 = (double)r1.y * d0 + (double)r2.x * d1 + (double)r2.w * d2;
    f2 = (double)r1.z * d0 + (double)r2.y * d1 + (double)r3.x * d2;
}

// One tet incidence of the stable Neo-Hookean energy in FP64
// (NeoHookeanEnergySolver::accumulate, energy/elastic/NeoHookeanEnergySolver.hpp:121-227):
// Gauss-Newton hydrostatic term, exact deviatoric term, strain-rate damping.
template <bool DAMP>
__device__ __forceinline__ void nh_incidence(D3 p0, D3 p1, D3 p2, D3 p3, float4 r1, float4 r2, float4 r3,
                                             uint32_t local, const TetDamp* __restrict__ damp, uint32_t t,
                                             double inv_dt, D3& g, SymD& h) {
    D3 f0, f1, f2;
    deform(p0, p1, p2, p3, r1, r2, r3, f0, f1, f2);
    const double q00 = r1.x, q01 = r1.y, q02 = r1.z, q10 = r1.w, q11 = r2.x, q12 = r2.y, q20 = r2.z, q21 = r2.w,
                 q22 = r3.x;
    const double VL = r3.y, VM = r3.z, alpha = r3.w;
    const D3 c0 = cross3(f1, f2), c1 = cross3(f2, f0), c2 = cross3(f0, f1);
    const double J = dot3(f0, c0);
    D3 q;
    if (local == 0) q = D3{q00, q01, q02};
    else if (local == 1) q = D3{q10, q11, q12};
    else if (local == 2) q = D3{q20, q21, q22};
    else q = D3{-(q00 + q10 + q20), -(q01 + q11 + q21), -(q02 + q12 + q22)};
    const D3 Fq = q.x * f0 + q.y * f1 + q.z * f2;
    const D3 Fc = q.x * c0 + q.y * c1 + q.z * c2;
    const double hyd = VL * (J - alpha);
    const double q2 = dot3(q, q);
    g = g + (hyd * Fc + VM * Fq);
    const double dev = VM * q2;
    h.a00 += VL * Fc.x * Fc.x + dev;
    h.a11 += VL * Fc.y * Fc.y + dev;
    h.a22 += VL * Fc.z * Fc.z + dev;
    h.a01 += VL * Fc.x * Fc.y;
    h.a02 += VL * Fc.x * Fc.z;
    h.a12 += VL * Fc.y * Fc.z;
    if constexpr (DAMP) {
        const TetDamp dr = damp[t];
        const double dm = dr.damp;
        const double e00 = (0.5 * dot3(f0, f0) - 0.5 - dr.e00) * inv_dt;
        const double e11 = (0.5 * dot3(f1, f1) - 0.5 - dr.e11) * inv_dt;
        const double e22 = (0.5 * dot3(f2, f2) - 0.5 - dr.e22) * inv_dt;
        const double e01 = (0.5 * dot3(f0, f1) - dr.e01) * inv_dt;
        const double e02 = (0.5 * dot3(f0, f2) - dr.e02) * inv_dt;
        const double e12 = (0.5 * dot3(f1, f2) - dr.e12) * inv_dt;
        const D3 Eq = D3{e00 * q.x + e01 * q.y + e02 * q.z, e01 * q.x + e11 * q.y + e12 * q.z,
                         e02 * q.x + e12 * q.y + e22 * q.z};
        const D3 FEq = Eq.x * f0 + Eq.y * f1 + Eq.z * f2;
        g = g + dm * FEq;
        const double diag = dm * dot3(q, Eq);
        const double a = 0.5 * dm * inv_dt;
        const double ff00 = f0.x * f0.x + f1.x * f1.x + f2.x * f2.x;
        const double ff11 = f0.y * f0.y + f1.y * f1.y + f2.y * f2.y;
        const double ff22 = f0.z * f0.z + f1.z * f1.z + f2.z * f2.z;
        const double ff01 = f0.x * f0.y + f1.x * f1.y + f2.x * f2.y;
        const double ff02 = f0.x * f0.z + f1.x * f1.z + f2.x * f2.z;
        const double ff12 = f0.y * f0.z + f1.y * f1.z + f2.y * f2.z;
        h.a00 += diag + a * (q2 * ff00 + Fq.x * Fq.x);
        h.a11 += diag + a * (q2 * ff11 + Fq.y * Fq.y);
        h.a22 += diag + a * (q2 * ff22 + Fq.z * Fq.z);
        h.a01 += a * (q2 * ff01 + Fq.x * Fq.y);
        h.a02 += a * (q2 * ff02 + Fq.x * Fq.z);
        h.a12 += a * (q2 * ff12 + Fq.y * Fq.z);
    }
}

// Ground AL contact (CollisionConstraintEnergySolver::accumulate<3> with the
// ground constraint, energy/collision/GroundCollisionEnergySolver.hpp:27-60).
__device__ __forceinline__ void ground_accumulate(const GroundRec& c, D3 p, D3& g, SymD& h) {
    const double Cn = -p.y, Ct = p.x - (double)c.cp_x, Cb = p.z - (double)c.cp_z;
    const double Ccorr = Cn - (double)kAlpha * c.C_n_prev;
    const double lnp = fmax(0.0, (double)c.k_n * Ccorr + c.lambda_n);
    double lt = (double)c.k_t * Ct + c.lambda_t;
    double lb = (double)c.k_b * Cb + c.lambda_b;
    const double mag = sqrt(lt * lt + lb * lb);
    const double mu = (c.flags & GF_USE_STATIC) ? c.mu_s : c.mu_k;
    const double lmax = mu * lnp;
    if (mag > lmax) {
        const double nrm = mag + 1e-8;
        lt = lt / nrm * lmax;
        lb = lb / nrm * lmax;
    }
    g.x += lt;
    g.y += -lnp;
    g.z += lb;
    h.a00 += c.k_t;
    h.a11 += c.k_n;
    h.a22 += c.k_b;
}

// Dual + stiffness-ramp update of one ground record at the extrapolated
// position (CollisionConstraintEnergySolver::updateAfterIteration,
// energy/collision/CollisionConstraintEnergySolver.hpp:60-134).
__device__ __forceinline__ void ground_dual_update(GroundRec& c, D3 xp) {
    const double Cn = -xp.y, Ct = xp.x - (double)c.cp_x, Cb = xp.z - (double)c.cp_z;
    const double Ccorr = Cn - (double)kAlpha * c.C_n_prev;
    const double lnp = (double)c.k_n * Ccorr + c.lambda_n;
    const double ltp = (double)c.k_t * Ct + c.lambda_t;
    const double lbp = (double)c.k_b * Cb + c.lambda_b;
    if (lnp > 0.0) c.k_n = (float)(c.k_n + (double)kBeta * fabs(Ccorr));
    const double ln = fmax(0.0, lnp);
    c.lambda_n = (float)ln;
    const double mag = sqrt(ltp * ltp + lbp * lbp);
    double lmax = ((c.flags & GF_USE_STATIC) ? c.mu_s : c.mu_k) * ln;
    if (mag < lmax) {
        c.k_t = (float)(c.k_t + (double)kBeta * fabs(Ct));
        c.k_b = (float)(c.k_b + (double)kBeta * fabs(Cb));
        c.lambda_t = (float)ltp;
        c.lambda_b = (float)lbp;
        c.flags |= GF_UNCLAMPED;
    } else {
        if (c.flags & GF_USE_STATIC) {
            c.flags &= ~GF_USE_STATIC;
            lmax = c.mu_k * lnp;
        }
        c.lambda_t = (float)(ltp / (mag + 1e-8) * lmax);
        c.lambda_b = (float)(lbp / (mag + 1e-8) * lmax);
        c.flags &= ~GF_UNCLAMPED;
    }
}

// Vertex-triangle contact (the only dynamic kind with the FP64 state, where
// every particle is positional) for particle `local` of record c: the
// triangle-rigid AL contact with the vertex as the contact point,
// C = dir . (x_v - sum_i barys_i x_i) for dir = n, t, b, evaluated in FP64
// from the FP64 positions (CollisionConstraintEnergySolver::accumulate<3>,
// energy/collision/CollisionConstraintEnergySolver.hpp:179-285).
__device__ __forceinline__ void vt_constraint(const double4* x, const DynRec& c, double& Cn, double& Ct, double& Cb) {
    const D3 x0 = dd3(x[c.idx[0]]), x1 = dd3(x[c.idx[1]]), x2 = dd3(x[c.idx[2]]), xv = dd3(x[c.idx[3]]);
    const D3 cp = (double)c.bary.x * x0 + (double)c.bary.y * x1 + (double)c.bary.z * x2;
    const D3 d = xv - cp;
    Cn = (double)c.n.x * d.x + (double)c.n.y * d.y + (double)c.n.z * d.z;
    Ct = (double)c.t.x * d.x + (double)c.t.y * d.y + (double)c.t.z * d.z;
    Cb = (double)c.b.x * d.x + (double)c.b.y * d.y + (double)c.b.z * d.z;
}

__device__ __forceinline__ void vt_accumulate(const double4* x, const DynRec& c, uint32_t local, D3& g, SymD& h) {
    if (dyn::kind_of(c) != DYN_VERTEX_TRI) return;
    double Cn, Ct, Cb;
    vt_constraint(x, c, Cn, Ct, Cb);
    const double Ccorr = Cn - (double)kAlpha * c.cpl.w;
    const double lnp = fmax(0.0, (double)c.k.x * Ccorr + c.lam.x);
    double lt = (double)c.k.y * Ct + c.lam.y, lb = (double)c.k.z * Cb + c.lam.z;
    const double mag = sqrt(lt * lt + lb * lb);
    const double lmax = (double)((dyn::flags_of(c) & GF_USE_STATIC) ? c.t.w : c.b.w) * lnp;
    if (mag > lmax) {
        const double nrm = mag + 1e-8;
        lt = lt / nrm * lmax;
        lb = lb / nrm * lmax;
    }
    const double w = local == 3 ? 1.0 : -(double)(local == 0 ? c.bary.x : (local == 1 ? c.bary.y : c.bary.z));
    const D3 n = D3{c.n.x, c.n.y, c.n.z}, t = D3{c.t.x, c.t.y, c.t.z}, b = D3{c.b.x, c.b.y, c.b.z};
    g = g + w * (lnp * n + lt * t + lb * b);
    const double w2 = w * w, kn = w2 * c.k.x, kt = w2 * c.k.y, kb = w2 * c.k.z;
    h.a00 += kn * n.x * n.x + kt * t.x * t.x + kb * b.x * b.x;
    h.a11 += kn * n.y * n.y + kt * t.y * t.y + kb * b.y * b.y;
    h.a22 += kn * n.z * n.z + kt * t.z * t.z + kb * b.z * b.z;
    h.a01 += kn * n.x * n.y + kt * t.x * t.y + kb * b.x * b.y;
    h.a02 += kn * n.x * n.z + kt * t.x * t.z + kb * b.x * b.z;
    h.a12 += kn * n.y * n.z + kt * t.y * t.z + kb * b.y * b.z;
}

// The dynamic contacts of particle v, after its static energies (simulation/VBDSolver.hpp:400-438).
__device__ __forceinline__ void dyn_terms(const uint32_t* __restrict__ off, const uint32_t* __restrict__ ent,
                                          const DynRec* __restrict__ rec, const double4* x, uint32_t v, D3& g,
                                          SymD& h) {
    const uint32_t e1 = off[v + 1];
    for (uint32_t e = off[v]; e < e1; e++) {
        const uint32_t code = ent[e];
        vt_accumulate(x, rec[code >> 3], code & 7u, g, h);
    }
}

__device__ __forceinline__ D3 ldlt3(const SymD& A, D3 b) {
    const double d0 = A.a00;
    const double l10 = A.a01 / d0, l20 = A.a02 / d0;
    const double d1 = A.a11 - l10 * A.a01;
    const double l21 = (A.a12 - l20 * A.a01) / d1;
    const double d2 = A.a22 - l20 * A.a02 - l21 * l21 * d1;
    const double z0 = b.x;
    const double z1 = b.y - l10 * z0;
    const double z2 = b.z - l20 * z0 - l21 * z1;
    const double x2 = z2 / d2;
    const double x1 = z1 / d1 - l21 * x2;
    const double x0 = z0 / d0 - l10 * x1 - l20 * x2;
    return D3{x0, x1, x2};
}

template <int L>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ __forceinline__ uint32_t xcd_logical_block() {
    const uint32_t b = blockIdx.x, G = gridDim.x;
    const uint32_t x = b % 8u, q = b / 8u;
    return x * (G / 8u) + min(x, G % 8u) + q;
}

// Sum of the tet incidences of one work item over its L lanes (every lane
// gets the sums); U incidences per lane are gathered at once (records, then
// positions with the next records in flight), summed in incidence order.
template <int L, int U, bool DAMP, bool HESS>
__device__ __forceinline__ void item_gradient(const VItem& it, uint32_t lane, const uint32_t* __restrict__ inc,
                                              const TetRec* __restrict__ irec, const TetDamp* __restrict__ damp,
                                              const double4* x, double inv_dt, D3& g, SymD& h) {
    const uint32_t cnt = it.inc_count & ~VITEM_CONFLICT;
    g = D3{0.0, 0.0, 0.0};
    h = SymD{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    uint4 idx[U];
    float4 r1[U], r2[U], r3[U];
#pragma unroll
    for (int u = 0; u < U; u++)
        if (lane + u * L < cnt) load_rec(irec, it.inc_begin + lane + u * L, idx[u], r1[u], r2[u], r3[u]);
    for (uint32_t j0 = lane; j0 < cnt; j0 += U * L) {
        D3 P[U][4];
#pragma unroll
        for (int u = 0; u < U; u++)
            if (j0 + u * L < cnt) {
                P[u][0] = dd3(x[idx[u].x]);
                P[u][1] = dd3(x[idx[u].y]);
                P[u][2] = dd3(x[idx[u].z]);
                P[u][3] = dd3(x[idx[u].w & 0x3fffffffu]);
            }
        uint4 nidx[U];
        float4 n1[U], n2[U], n3[U];
#pragma unroll
        for (int u = 0; u < U; u++)
            if (j0 + (U + u) * L < cnt) load_rec(irec, it.inc_begin + j0 + (U + u) * L, nidx[u], n1[u], n2[u], n3[u]);
#pragma unroll
        for (int u = 0; u < U; u++)
            if (j0 + u * L < cnt)
                nh_incidence<DAMP>(P[u][0], P[u][1], P[u][2], P[u][3], r1[u], r2[u], r3[u], idx[u].w >> 30, damp,
                                   DAMP ? (inc[it.inc_begin + j0 + u * L] >> 2) : 0u, inv_dt, g, h);
#pragma unroll
        for (int u = 0; u < U; u++) {
            idx[u] = nidx[u];
            r1[u] = n1[u];
            r2[u] = n2[u];
            r3[u] = n3[u];
        }
    }
    g.x = group_sum<L>(g.x);
    g.y = group_sum<L>(g.y);
    g.z = group_sum<L>(g.z);
    if constexpr (HESS) {
        h.a00 = group_sum<L>(h.a00);
        h.a11 = group_sum<L>(h.a11);
        h.a22 = group_sum<L>(h.a22);
        h.a01 = group_sum<L>(h.a01);
        h.a02 = group_sum<L>(h.a02);
        h.a12 = group_sum<L>(h.a12);
    }
}

// The group leader's solve: ground terms, inertia, 3x3 LDL^T; new position
// (mass kept in w).
__device__ __forceinline__ double4 item_solve(const VItem& it, double4 xv, double4 yv, const GroundRec& gr, D3 g,
                                              SymD h, double inv_dt2) {
    const D3 p = dd3(xv);
    if (it.ground >= 0) ground_accumulate(gr, p, g, h);
    const double md = xv.w * inv_dt2;
    h.a00 += md;
    h.a11 += md;
    h.a22 += md;
    const D3 b = D3{-md * (p.x - yv.x) - g.x, -md * (p.y - yv.y) - g.y, -md * (p.z - yv.z) - g.z};
    const D3 dx = ldlt3(h, b);
    return make_double4(p.x + dx.x, p.y + dx.y, p.z + dx.z, xv.w);
}

// One colour phase (_solveParticle<3>, simulation/VBDSolver.hpp:285-449,487-510).
// DYN: the vertex-triangle contacts are added after the static energies and
// a conflicted particle's result is buffered in dyn.x_buf
// (_commitConflictedInColor, VBDSolver.hpp:108-119).
struct DynF64 {
    const uint32_t* off;
    const uint32_t* ent;
    const DynRec* rec;
    double4* x_buf;
};

template <int L, bool DAMP, int U, bool DYN = false>
__global__ __launch_bounds__(256) void solve_color_f64_kernel(const VItem* __restrict__ items, uint32_t n_items,
                                                              const uint32_t* __restrict__ inc,
                                                              const TetRec* __restrict__ irec,
                                                              const TetDamp* __restrict__ damp, double4* x,
                                                              const double4* __restrict__ y,
                                                              const GroundRec* __restrict__ ground, double4* x_out,
                                                              double inv_dt, double inv_dt2,
                                                              const uint32_t* __restrict__ done,
                                                              const uint32_t* __restrict__ frozen,
                                                              const uint32_t* __restrict__ scene_of,
                                                              DynF64 dy = DynF64{}) {
    kernarg_prefetch(items, n_items, inc, irec, x, y, ground, x_out, done, frozen, scene_of, inv_dt, inv_dt2,
                     blockDim.x, gridDim.x);
    const uint32_t tid = xcd_logical_block() * blockDim.x + threadIdx.x;
    const uint32_t item = tid / L, lane = tid % L;
    if (item >= n_items) return;   // whole L-lane groups exit together
    const VItem it = items[item];   // in flight with the converged flag's load
    if (done && *done) return;      // residual target reached earlier in this step
    if (frozen && frozen[scene_of[it.v]]) return;   // the item's scene has converged in this step
    double4 xv = make_double4(0.0, 0.0, 0.0, 0.0), yv = xv;
    GroundRec gr{};
    const bool lead = lane == 0;
    if (lead) {
        xv = x[it.v];
        yv = y[it.v];
        if (it.ground >= 0) gr = ground[it.ground];
    }
    D3 g;
    SymD h;
    item_gradient<L, U, DAMP, true>(it, lane, inc, irec, damp, x, inv_dt, g, h);
    if (!lead) return;
    if constexpr (DYN) {
        // static energies (tets, ground) first, then the dynamic contacts
        VItem st = it;
        if (it.ground >= 0) ground_accumulate(gr, dd3(xv), g, h);
        st.ground = -1;
        dyn_terms(dy.off, dy.ent, dy.rec, x, it.v, g, h);
        const double4 nx = item_solve(st, xv, yv, gr, g, h, inv_dt2);
        if (it.inc_count & VITEM_CONFLICT) dy.x_buf[it.v] = nx;
        else x_out[it.v] = nx;
    } else {
        x_out[it.v] = item_solve(it, xv, yv, gr, g, h, inv_dt2);
    }
}

// Chebyshev weight of local iteration j of the running interval: from the
// per-iteration table at the step's iteration counter (residual mode), else
// the value baked into the launch.
__device__ __forceinline__ double cheb_omega(float omega, const ResCtl* rc, const float* omega_tab, uint32_t j) {
    return rc ? (double)omega_tab[rc->iter + j] : (double)omega;
}

// The last colour phase fused with the Chebyshev / dual update (layout and
// buffer discipline of solve_cheb_kernel in vbd_kernels.hip).
template <int L, bool DAMP, int U>
__global__ __launch_bounds__(256) void solve_cheb_f64_kernel(F64State f, const VItem* __restrict__ items,
                                                             uint32_t n_items, uint32_t sx, uint32_t cx,
                                                             const uint32_t* __restrict__ cheb_list, uint32_t n_cheb,
                                                             const double4* __restrict__ x_cur, double4* x_next,
                                                             float omega, const ResCtl* rc, const float* omega_tab,
                                                             uint32_t j, double inv_dt, double inv_dt2) {
    kernarg_prefetch(f.inc, f.irec, f.damp, f.y, f.ground, f.x_last, f.x_lastlast, f.ground_of, f.frozen, f.scene_of,
                     items, n_items, sx, cx, cheb_list, n_cheb, x_cur, x_next, rc, inv_dt, inv_dt2, blockDim.x);
    const uint32_t xcd = blockIdx.x % 8u, q = blockIdx.x / 8u;
    if (q < sx) {
        const uint32_t tid = (xcd * sx + q) * blockDim.x + threadIdx.x;
        const uint32_t item = tid / L, lane = tid % L;
        if (item >= n_items) return;
        const VItem it = items[item];   // in flight with the converged flag's load
        if (rc && rc->done) return;
        const double w = cheb_omega(omega, rc, omega_tab, j);
        const uint32_t v = it.v;
        if (f.frozen && f.frozen[f.scene_of[v]]) {   // converged scene: keep its position in both buffers
            if (lane == 0) x_next[v] = x_cur[v];
            return;
        }
        double4 xv = make_double4(0.0, 0.0, 0.0, 0.0), yv = xv, xll = xv;
        GroundRec gr{};
        const bool lead = lane == 0;
        if (lead) {
            xv = x_cur[v];
            yv = f.y[v];
            xll = f.x_lastlast[v];
            if (it.ground >= 0) gr = f.ground[it.ground];
        }
        D3 g;
        SymD h;
        item_gradient<L, U, DAMP, true>(it, lane, f.inc, f.irec, f.damp, x_cur, inv_dt, g, h);
        if (!lead) return;
        const double4 xs = item_solve(it, xv, yv, gr, g, h, inv_dt2);
        const double4 xp = make_double4(w * (xs.x - xll.x) + xll.x, w * (xs.y - xll.y) + xll.y,
                                        w * (xs.z - xll.z) + xll.z, xs.w);
        x_next[v] = xp;
        f.x_lastlast[v] = f.x_last[v];
        f.x_last[v] = xp;
        if (it.ground >= 0) {
            ground_dual_update(gr, dd3(xp));
            f.ground[it.ground] = gr;
        }
    } else {
        const uint32_t k = (xcd * cx + (q - sx)) * blockDim.x + threadIdx.x;
        if (k >= n_cheb) return;
        const uint32_t p = cheb_list[k];
        if (rc && rc->done) return;
        const double w = cheb_omega(omega, rc, omega_tab, j);
        if (f.frozen && f.frozen[f.scene_of[p]]) {
            x_next[p] = x_cur[p];
            return;
        }
        const double4 xl = f.x_last[p], xll = f.x_lastlast[p], xc = x_cur[p];
        const double4 xp = make_double4(w * (xc.x - xll.x) + xll.x, w * (xc.y - xll.y) + xll.y,
                                        w * (xc.z - xll.z) + xll.z, xc.w);
        x_next[p] = xp;
        f.x_lastlast[p] = xl;
        f.x_last[p] = xp;
        const int32_t gi = f.ground_of[p];
        if (gi >= 0) {
            GroundRec c = f.ground[gi];
            ground_dual_update(c, dd3(xp));
            f.ground[gi] = c;
        }
    }
}

// _particleChebyshevAcceleration (VBDSolver.hpp:513-536) + the ground dual update.
__global__ void cheb_dual_f64_kernel(F64State f, uint32_t n, float omega, const ResCtl* rc, const float* omega_tab,
                                     uint32_t j) {
    if (rc && rc->done) return;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || (f.frozen && f.frozen[f.scene_of[p]])) return;
    const double w = cheb_omega(omega, rc, omega_tab, j);
    const double4 xl = f.x_last[p], xll = f.x_lastlast[p], xc = f.x[p];
    const double4 xp = make_double4(w * (xc.x - xll.x) + xll.x, w * (xc.y - xll.y) + xll.y,
                                    w * (xc.z - xll.z) + xll.z, xc.w);
    f.x[p] = xp;
    f.x_lastlast[p] = xl;
    f.x_last[p] = xp;
    const int32_t gi = f.ground_of[p];
    if (gi >= 0) {
        GroundRec c = f.ground[gi];
        ground_dual_update(c, dd3(xp));
        f.ground[gi] = c;
    }
}

// Start of step per particle: ground contact point + AL warm start, inertial
// target and adaptive initial guess (GroundCollisionEnergySolver::updateContactPoints,
// CollisionConstraintEnergySolver::updateAfterTimeStep, _particleInertialUpdate,
// VBDSolver.hpp:208-248), and the FP32 mirror of the start position.
__global__ void step_begin_f64_kernel(F64State f, uint32_t n, double dt, double g) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double4 xp = f.x[p];
    const int32_t gi = f.ground_of[p];
    if (gi >= 0) {
        GroundRec c = f.ground[gi];
        c.cp_x = (float)xp.x;
        c.cp_z = (float)xp.z;
        if (c.lambda_n > 0.f) c.C_n_prev = (float)(-xp.y);
        if ((c.flags & GF_UNCLAMPED) && !(c.flags & GF_USE_STATIC)) c.flags |= GF_USE_STATIC;
        c.lambda_n *= kAlpha * kGamma;
        c.lambda_t *= kGamma;
        c.lambda_b *= kGamma;
        c.k_n = fmaxf(kGamma * c.k_n, c.k_start);
        c.k_t = fmaxf(kGamma * c.k_t, c.k_start);
        c.k_b = fmaxf(kGamma * c.k_b, c.k_start);
        f.ground[gi] = c;
    }
    if (!f.fixed[p]) {
        // external acceleration: gravity, plus the per-particle field when set
        // (not in the reference), in the inertial target and the adaptive guess
        const float4 vv = f.v[p], vp = f.v_prev[p];
        double ex = 0.0, ey = -g, ez = 0.0;
        if (f.a_ext) {
            const float4 fa = f.a_ext[p];
            ex += fa.x;
            ey += fa.y;
            ez += fa.z;
        }
        const double an = sqrt(ex * ex + ey * ey + ez * ez);
        double a_tilde = 0.0;
        if (an > 1e-12) {
            const double along =
                (((double)vv.x - vp.x) * ex + ((double)vv.y - vp.y) * ey + ((double)vv.z - vp.z) * ez) / dt / an;
            if (along > an) a_tilde = 1.0;
            else if (along < 0.0) a_tilde = 0.0;
            else a_tilde = along / an;
        }
        const double dt2 = dt * dt;
        const double4 yv = make_double4(xp.x + dt * vv.x + dt2 * ex, xp.y + dt * vv.y + dt2 * ey,
                                        xp.z + dt * vv.z + dt2 * ez, 0.0);
        f.y[p] = yv;
        f.y32[p] = mirror(yv);
        xp.x = xp.x + (dt * vv.x + dt2 * (a_tilde * ex));
        xp.y = xp.y + (dt * vv.y + dt2 * (a_tilde * ey));
        xp.z = xp.z + (dt * vv.z + dt2 * (a_tilde * ez));
        f.x[p] = xp;
        f.x_last[p] = xp;
        f.x_lastlast[p] = xp;
    }
    f.x32[p] = mirror(xp);
}

// E_prev = 0.5 (F_prev^T F_prev - I) (NeoHookeanEnergySolver::updateAfterTimeStep,
// energy/elastic/NeoHookeanEnergySolver.hpp:70-90) from the FP64 previous positions.
__global__ void eprev_f64_kernel(const TetRec* __restrict__ tets, TetDamp* __restrict__ damp,
                                 const double4* __restrict__ x_prev, uint32_t n_tets) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tets) return;
    uint4 idx;
    float4 r1, r2, r3;
    load_rec(tets, t, idx, r1, r2, r3);
    D3 f0, f1, f2;
    deform(dd3(x_prev[idx.x]), dd3(x_prev[idx.y]), dd3(x_prev[idx.z]), dd3(x_prev[idx.w]), r1, r2, r3, f0, f1, f2);
    TetDamp dr = damp[t];
    dr.e00 = (float)(0.5 * (dot3(f0, f0) - 1.0));
    dr.e11 = (float)(0.5 * (dot3(f1, f1) - 1.0));
    dr.e22 = (float)(0.5 * (dot3(f2, f2) - 1.0));
    dr.e01 = (float)(0.5 * dot3(f0, f1));
    dr.e02 = (float)(0.5 * dot3(f0, f2));
    dr.e12 = (float)(0.5 * dot3(f1, f2));
    damp[t] = dr;
}

// _particleVelocityUpdate (VBDSolver.hpp:538-566) + FP32 mirror; a
// non-finite position or velocity raises bit 0 of the status word.
__global__ void velocity_f64_kernel(F64State f, uint32_t n, double inv_dt, uint32_t* status) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double4 xp = f.x[p], xo = f.x_prev[p];
    const float4 vv = f.v[p];
    f.v_prev[p] = vv;
    const float4 nv = make_float4((float)((xp.x - xo.x) * inv_dt), (float)((xp.y - xo.y) * inv_dt),
                                  (float)((xp.z - xo.z) * inv_dt), 0.f);
    f.v[p] = nv;
    f.x_prev[p] = xp;
    f.x32[p] = mirror(xp);
    f.x_prev32[p] = mirror(xp);
    const bool bad = !(isfinite(xp.x) && isfinite(xp.y) && isfinite(xp.z) && isfinite(nv.x) && isfinite(nv.y) &&
                       isfinite(nv.z));
    if (bad) atomicOr(status, 1u);
}

// Odd number of iterations over two position buffers: the result is in x_alt.
__global__ void settle_f64_kernel(double4* __restrict__ x, const double4* __restrict__ x_alt, uint32_t n,
                                  const uint32_t* __restrict__ iters_word, uint32_t iters) {
    const uint32_t it = iters_word ? *iters_word : iters;
    if ((it & 1u) == 0) return;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) x[p] = x_alt[p];
}

// Force residual r_v = m/dt^2 (x_v - y_v) + dE/dx_v per work item, |r_v|^2
// into rr[v] (lane 0 of each group).
template <int L, bool DAMP>
__global__ __launch_bounds__(256) void residual_f64_kernel(F64State f, const VItem* __restrict__ items,
                                                           uint32_t n_items, const double4* __restrict__ x,
                                                           double* __restrict__ rr, const ResCtl* rc, double inv_dt,
                                                           double inv_dt2) {
    kernarg_prefetch(f.inc, f.irec, f.damp, f.y, f.ground, f.frozen, f.scene_of, f.dyn_off, items, n_items, x, rr, rc,
                     inv_dt, inv_dt2, blockDim.x);
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t item = tid / L, lane = tid % L;
    if (item >= n_items) return;
    const VItem it = items[item];   // in flight with the converged flag's load
    if (rc && rc->done) return;
    if (f.frozen && f.frozen[f.scene_of[it.v]]) return;   // keeps the residual it converged with
    D3 g;
    SymD h;
    item_gradient<L, 1, DAMP, false>(it, lane, f.inc, f.irec, f.damp, x, inv_dt, g, h);
    if (lane != 0) return;
    const double4 xv = x[it.v], yv = f.y[it.v];
    if (it.ground >= 0) ground_accumulate(f.ground[it.ground], dd3(xv), g, h);
    if (f.dyn_off) dyn_terms(f.dyn_off, f.dyn_ent, f.dyn, x, it.v, g, h);
    const double md = xv.w * inv_dt2;
    const double rx = md * (xv.x - yv.x) + g.x, ry = md * (xv.y - yv.y) + g.y, rz = md * (xv.z - yv.z) + g.z;
    rr[it.v] = rx * rx + ry * ry + rz * rz;
}

// Stable Neo-Hookean energy of tet t in FP64 (NeoHookeanEnergySolver::energy,
// energy/elastic/NeoHookeanEnergySolver.hpp:20-53), with damping.
template <bool DAMP>
__device__ __forceinline__ double nh_energy(const double4* __restrict__ x, const TetRec* __restrict__ tets,
                                            const TetDamp* __restrict__ damp, uint32_t t, double inv_dt) {
    uint4 idx;
    float4 r1, r2, r3;
    load_rec(tets, t, idx, r1, r2, r3);
    D3 f0, f1, f2;
    deform(dd3(x[idx.x]), dd3(x[idx.y]), dd3(x[idx.z]), dd3(x[idx.w]), r1, r2, r3, f0, f1, f2);
    const double J = dot3(f0, cross3(f1, f2));
    const double VL = r3.y, VM = r3.z, al = r3.w;
    const double trc = dot3(f0, f0) + dot3(f1, f1) + dot3(f2, f2);
    double e = 0.5 * VL * (J - al) * (J - al) + 0.5 * VM * (trc - 3.0);
    if constexpr (DAMP) {
        const TetDamp dr = damp[t];
        const double E[6] = {0.5 * (dot3(f0, f0) - 1.0), 0.5 * (dot3(f1, f1) - 1.0), 0.5 * (dot3(f2, f2) - 1.0),
                             0.5 * dot3(f0, f1), 0.5 * dot3(f0, f2), 0.5 * dot3(f1, f2)};
        const double Pv[6] = {dr.e00, dr.e11, dr.e22, dr.e01, dr.e02, dr.e12};
        double sq = 0;
        for (int i = 0; i < 6; i++) {
            const double dd = (E[i] - Pv[i]) * inv_dt;
            sq += (i < 3 ? 1.0 : 2.0) * dd * dd;
        }
        e += 0.5 * dr.damp / inv_dt * sq;
    }
    return e;
}

__device__ __forceinline__ double ground_energy(const GroundRec& c, double4 xp) {
    const double Cn = -xp.y, Ct = xp.x - (double)c.cp_x, Cb = xp.z - (double)c.cp_z;
    const double Cc = Cn - (double)kAlpha * c.C_n_prev;
    double e = 0;
    if ((double)c.k_n * Cc + c.lambda_n > 0) e += 0.5 * c.k_n * Cc * Cc + (double)c.lambda_n * Cc;
    return e + 0.5 * c.k_t * Ct * Ct + (double)c.lambda_t * Ct + 0.5 * c.k_b * Cb * Cb + (double)c.lambda_b * Cb;
}

// avbd_get_stats: per block, slots 0..2 = sums of |r|^2, |m g|^2 and the
// inertia term over the work items of the block, slot 3 = tet + ground energy.
template <int L, bool DAMP>
__global__ __launch_bounds__(256) void stats_f64_kernel(F64State f, const VItem* __restrict__ items, uint32_t n_items,
                                                        uint32_t n_tets, uint32_t n_ground, double* __restrict__ out,
                                                        double inv_dt, double inv_dt2, double g_acc) {
    __shared__ double red[4][256];
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t item = tid / L, lane = tid % L;
    double r2 = 0, mg2 = 0, ip = 0, e = 0;
    if (item < n_items) {
        const VItem it = items[item];
        D3 g;
        SymD h;
        item_gradient<L, 1, DAMP, false>(it, lane, f.inc, f.irec, f.damp, f.x, inv_dt, g, h);
        if (lane == 0) {
            const double4 xv = f.x[it.v], yv = f.y[it.v];
            if (it.ground >= 0) ground_accumulate(f.ground[it.ground], dd3(xv), g, h);
            if (f.dyn_off) dyn_terms(f.dyn_off, f.dyn_ent, f.dyn, f.x, it.v, g, h);
            const double md = xv.w * inv_dt2;
            const double dx = xv.x - yv.x, dy = xv.y - yv.y, dz = xv.z - yv.z;
            const double rx = md * dx + g.x, ry = md * dy + g.y, rz = md * dz + g.z;
            r2 = rx * rx + ry * ry + rz * rz;
            mg2 = (xv.w * g_acc) * (xv.w * g_acc);
            ip = 0.5 * md * (dx * dx + dy * dy + dz * dz);
        }
    }
    if (tid < n_tets) e = nh_energy<DAMP>(f.x, f.tets, f.damp, tid, inv_dt);
    else if (tid < n_tets + n_ground) {
        const GroundRec c = f.ground[tid - n_tets];
        e = ground_energy(c, f.x[c.p]);
    }
    red[0][threadIdx.x] = r2;
    red[1][threadIdx.x] = mg2;
    red[2][threadIdx.x] = ip;
    red[3][threadIdx.x] = e;
    __syncthreads();
    for (uint32_t st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            for (int i = 0; i < 4; i++) red[i][threadIdx.x] += red[i][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x < 4) out[blockIdx.x * 4 + threadIdx.x] = red[threadIdx.x][0];
}

// Dynamic vertex-triangle records with the FP64 state: the AL updates of
// dyn_device.hpp at constraint values evaluated from the FP64 positions.
__global__ void dyn_step_begin_f64_kernel(const double4* __restrict__ x, DynRec* __restrict__ dyn,
                                          const uint32_t* __restrict__ active, uint32_t n_active) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_active) return;
    DynRec c = dyn[active[k]];
    double Cn = 0, Ct, Cb;
    if (c.lam.x > 0.f) vt_constraint(x, c, Cn, Ct, Cb);
    dyn::al_step_begin(c, (float)Cn);
    dyn[active[k]] = c;
}

__global__ void dyn_dual_f64_kernel(const double4* __restrict__ x, DynRec* __restrict__ dyn,
                                    const uint32_t* __restrict__ active, uint32_t n_active,
                                    const uint32_t* __restrict__ done) {
    if (done && *done) return;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_active) return;
    DynRec c = dyn[active[k]];
    double Cn, Ct, Cb;
    vt_constraint(x, c, Cn, Ct, Cb);
    dyn::al_update(c, (float)Cn, (float)Ct, (float)Cb);
    dyn[active[k]] = c;
}

// _commitConflictedInColor (VBDSolver.hpp:108-119): buffered positions of one colour.
__global__ void commit_f64_kernel(double4* __restrict__ x, const double4* __restrict__ x_buf,
                                  const uint32_t* __restrict__ list, uint32_t count, const uint32_t* __restrict__ done) {
    if (done && *done) return;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) x[list[k]] = x_buf[list[k]];
}

// Learned warm start with the FP64 state: keep the start (x0 = x), apply the
// chosen step along the prediction per scene, training targets.
__global__ void warm_save_f64_kernel(F64State f, uint32_t n) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) f.x0[p] = f.x[p];
}

__global__ void warm_place_f64_kernel(F64State f, const VItem* __restrict__ items, uint32_t n_items,
                                      const float4* __restrict__ pred, const uint32_t* __restrict__ scene_of,
                                      const float* __restrict__ step_of_scene) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_items) return;
    const uint32_t v = items[i].v;
    const double a = step_of_scene[scene_of[v]];
    const double4 b = f.x0[v];
    const float4 d = pred[v];
    const double4 x = make_double4(b.x + a * d.x, b.y + a * d.y, b.z + a * d.z, b.w);
    f.x[v] = x;
    f.x_last[v] = x;
    f.x_lastlast[v] = x;
    f.x32[v] = mirror(x);
}

__global__ void warm_target_f64_kernel(F64State f, const VItem* __restrict__ items, uint32_t n_items,
                                       float* __restrict__ tgt, double inv_s) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_items) return;
    const uint32_t v = items[i].v;
    const double4 x = f.x[v], x0 = f.x0[v];
    tgt[(size_t)i * 3 + 0] = (float)((x.x - x0.x) * inv_s);
    tgt[(size_t)i * 3 + 1] = (float)((x.y - x0.y) * inv_s);
    tgt[(size_t)i * 3 + 2] = (float)((x.z - x0.z) * inv_s);
}

// Per scene and step length a: incremental potential at x0 + a pred minus at
// x0, in FP64 from the FP64 start (the terms that change along the
// prediction: Neo-Hookean with damping, ground AL, inertia).  An FP32
// evaluation cannot resolve these differences once the start is already
// close to the solution (the energies are ~1e4 times larger than their
// change there), and would then pick predictions that raise the potential.
__device__ __forceinline__ D3 along(const double4* x0, const float4* pred, uint32_t p, double a) {
    const double4 b = x0[p];
    const float4 d = pred[p];
    return D3{b.x + a * d.x, b.y + a * d.y, b.z + a * d.z};
}

template <bool DAMP>
__device__ __forceinline__ double nh_energy_at(D3 p0, D3 p1, D3 p2, D3 p3, float4 r1, float4 r2, float4 r3,
                                               const TetDamp* __restrict__ damp, uint32_t t, double inv_dt) {
    D3 f0, f1, f2;
    deform(p0, p1, p2, p3, r1, r2, r3, f0, f1, f2);
    const double J = dot3(f0, cross3(f1, f2));
    const double VL = r3.y, VM = r3.z, al = r3.w;
    const double trc = dot3(f0, f0) + dot3(f1, f1) + dot3(f2, f2);
    double e = 0.5 * VL * (J - al) * (J - al) + 0.5 * VM * (trc - 3.0);
    if constexpr (DAMP) {
        const TetDamp dr = damp[t];
        const double E[6] = {0.5 * (dot3(f0, f0) - 1.0), 0.5 * (dot3(f1, f1) - 1.0), 0.5 * (dot3(f2, f2) - 1.0),
                             0.5 * dot3(f0, f1), 0.5 * dot3(f0, f2), 0.5 * dot3(f1, f2)};
        const double Pv[6] = {dr.e00, dr.e11, dr.e22, dr.e01, dr.e02, dr.e12};
        double sq = 0;
        for (int i = 0; i < 6; i++) {
            const double dd = (E[i] - Pv[i]) * inv_dt;
            sq += (i < 3 ? 1.0 : 2.0) * dd * dd;
        }
        e += 0.5 * dr.damp / inv_dt * sq;
    }
    return e;
}

__device__ __forceinline__ double ground_energy_at(const GroundRec& c, D3 p) {
    return ground_energy(c, make_double4(p.x, p.y, p.z, 0.0));
}

template <bool DAMP>
__global__ __launch_bounds__(256) void warm_ip_f64_kernel(F64State f, WarmIO io, double inv_dt) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    double e[kWarmAlphas] = {0.0, 0.0, 0.0};
    uint32_t scene = 0;
    bool active = false;
    if (k < f.n_tets) {
        uint4 idx;
        float4 r1, r2, r3;
        load_rec(f.tets, k, idx, r1, r2, r3);
        const uint32_t v[4] = {idx.x, idx.y, idx.z, idx.w};
        D3 P[4];
        for (int i = 0; i < 4; i++) P[i] = along(f.x0, io.pred, v[i], 0.0);
        const double e0 = nh_energy_at<DAMP>(P[0], P[1], P[2], P[3], r1, r2, r3, f.damp, k, inv_dt);
        for (int j = 0; j < kWarmAlphas; j++) {
            for (int i = 0; i < 4; i++) P[i] = along(f.x0, io.pred, v[i], kWarmSteps[j]);
            e[j] = nh_energy_at<DAMP>(P[0], P[1], P[2], P[3], r1, r2, r3, f.damp, k, inv_dt) - e0;
        }
        scene = io.scene_of[idx.x];
        active = true;
    } else if (k < f.n_tets + f.n_ground) {
        const GroundRec c = f.ground[k - f.n_tets];
        const double e0 = ground_energy_at(c, along(f.x0, io.pred, c.p, 0.0));
        for (int j = 0; j < kWarmAlphas; j++) e[j] = ground_energy_at(c, along(f.x0, io.pred, c.p, kWarmSteps[j])) - e0;
        scene = io.scene_of[c.p];
        active = true;
    } else if (k < f.n_tets + f.n_ground + f.n_items) {
        const uint32_t v = f.items[k - f.n_tets - f.n_ground].v;
        const double4 x0 = f.x0[v], y = f.y[v];
        const double md = 0.5 * x0.w * inv_dt * inv_dt;
        const D3 d0 = D3{x0.x - y.x, x0.y - y.y, x0.z - y.z};
        const double b = dot3(d0, d0);
        for (int j = 0; j < kWarmAlphas; j++) {
            const D3 x = along(f.x0, io.pred, v, kWarmSteps[j]);
            const D3 d = D3{x.x - y.x, x.y - y.y, x.z - y.z};
            e[j] = md * (dot3(d, d) - b);
        }
        scene = io.scene_of[v];
        active = true;
    }
    add_per_scene_block<kWarmAlphas>(io.ipd, io.n_scenes, scene, e, active);
}

__global__ void warm_ip_reset_f64_kernel(WarmIO io) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < io.n_scenes * kWarmAlphas) io.ipd[s] = 0.0;
}

inline uint32_t blocks_for(uint32_t n, uint32_t b) { return (n + b - 1) / b; }

}  // namespace

void launch_warm_ip_f64(hipStream_t s, const F64State& f, const WarmIO& io, bool damping, double dt) {
    hipLaunchKernelGGL(warm_ip_reset_f64_kernel, dim3(blocks_for(io.n_scenes * kWarmAlphas, 256)), dim3(256), 0, s,
                       io);
    const uint32_t n = f.n_tets + f.n_ground + f.n_items;
    if (n == 0) return;
    if (damping)
        hipLaunchKernelGGL(warm_ip_f64_kernel<true>, dim3(blocks_for(n, 256)), dim3(256), 0, s, f, io, 1.0 / dt);
    else
        hipLaunchKernelGGL(warm_ip_f64_kernel<false>, dim3(blocks_for(n, 256)), dim3(256), 0, s, f, io, 1.0 / dt);
}

// ---- launch wrappers ---------------------------------------------------------

void launch_solve_color_f64(hipStream_t s, const F64State& f, const LaunchCfg& cfg, uint32_t begin, uint32_t count,
                            bool damping, double dt, double4* x_out, const uint32_t* done) {
    if (count == 0) return;
    const uint32_t block = cfg.block > 256 ? 256u : cfg.block;
    const uint32_t L = cfg.lanes;
    const dim3 grid(blocks_for(count * L, block)), blk(block);
    const double inv_dt = 1.0 / dt, inv_dt2 = inv_dt * inv_dt;
#define AVBD_F64_LAUNCH(LL, DD, UU)                                                                                  \
    hipLaunchKernelGGL((solve_color_f64_kernel<LL, DD, UU>), grid, blk, 0, s, f.items + begin, count, f.inc, f.irec, \
                       f.damp, f.x, f.y, f.ground, x_out, inv_dt, inv_dt2, done, f.frozen, f.scene_of, DynF64{})
    // with dynamic contacts: one incidence gathered per lane at a time
#define AVBD_F64_U(LL, DD)                                                                                         \
    if (f.dyn_off)                                                                                                 \
        hipLaunchKernelGGL((solve_color_f64_kernel<LL, DD, 1, true>), grid, blk, 0, s, f.items + begin, count,    \
                           f.inc, f.irec, f.damp, f.x, f.y, f.ground, x_out, inv_dt, inv_dt2, done, f.frozen,      \
                           f.scene_of, DynF64{f.dyn_off, f.dyn_ent, f.dyn, f.x_buf});                               \
    else if (cfg.batch >= 4) AVBD_F64_LAUNCH(LL, DD, 4);                                                           \
    else if (cfg.batch >= 2) AVBD_F64_LAUNCH(LL, DD, 2);                                                           \
    else AVBD_F64_LAUNCH(LL, DD, 1);
#define AVBD_F64_CASE(LL)                      \
    case LL:                                    \
        if (damping) { AVBD_F64_U(LL, true) }   \
        else { AVBD_F64_U(LL, false) }          \
        break;
    switch (L) {
        AVBD_F64_CASE(4)
        AVBD_F64_CASE(8)
        AVBD_F64_CASE(16)
        AVBD_F64_CASE(32)
        default: break;
    }
#undef AVBD_F64_CASE
#undef AVBD_F64_U
#undef AVBD_F64_LAUNCH
}

void launch_dyn_step_begin_f64(hipStream_t s, const F64State& f, const uint32_t* active, uint32_t n_active) {
    if (n_active)
        hipLaunchKernelGGL(dyn_step_begin_f64_kernel, dim3(blocks_for(n_active, 256)), dim3(256), 0, s, f.x, f.dyn,
                           active, n_active);
}

void launch_dyn_dual_f64(hipStream_t s, const F64State& f, const uint32_t* active, uint32_t n_active,
                         const uint32_t* done) {
    if (n_active)
        hipLaunchKernelGGL(dyn_dual_f64_kernel, dim3(blocks_for(n_active, 256)), dim3(256), 0, s, f.x, f.dyn, active,
                           n_active, done);
}

void launch_commit_f64(hipStream_t s, const F64State& f, const uint32_t* list, uint32_t count, const uint32_t* done) {
    if (count)
        hipLaunchKernelGGL(commit_f64_kernel, dim3(blocks_for(count, 256)), dim3(256), 0, s, f.x, f.x_buf, list, count,
                           done);
}

void launch_solve_cheb_f64(hipStream_t s, const F64State& f, const LaunchCfg& cfg, uint32_t begin, uint32_t count,
                           bool damping, double dt, float omega, const ResCtl* rc, const float* omega_tab,
                           uint32_t j, const uint32_t* cheb_list, uint32_t n_cheb, const double4* x_cur,
                           double4* x_next) {
    const uint32_t block = cfg.block > 256 ? 256u : cfg.block;
    const uint32_t L = cfg.lanes;
    const uint32_t sx = (blocks_for(count * L, block) + 7u) / 8u, cx = (blocks_for(n_cheb, block) + 7u) / 8u;
    if (sx + cx == 0) return;
    const dim3 grid(8u * (sx + cx)), blk(block);
    const double inv_dt = 1.0 / dt, inv_dt2 = inv_dt * inv_dt;
#define AVBD_F64C_LAUNCH(LL, DD, UU)                                                                              \
    hipLaunchKernelGGL((solve_cheb_f64_kernel<LL, DD, UU>), grid, blk, 0, s, f, f.items + begin, count, sx, cx,   \
                       cheb_list, n_cheb, x_cur, x_next, omega, rc, omega_tab, j, inv_dt, inv_dt2)
#define AVBD_F64C_U(LL, DD)                           \
    if (cfg.batch >= 4) AVBD_F64C_LAUNCH(LL, DD, 4);  \
    else if (cfg.batch >= 2) AVBD_F64C_LAUNCH(LL, DD, 2); \
    else AVBD_F64C_LAUNCH(LL, DD, 1);
#define AVBD_F64C_CASE(LL)                     \
    case LL:                                    \
        if (damping) { AVBD_F64C_U(LL, true) }  \
        else { AVBD_F64C_U(LL, false) }         \
        break;
    switch (L) {
        AVBD_F64C_CASE(4)
        AVBD_F64C_CASE(8)
        AVBD_F64C_CASE(16)
        AVBD_F64C_CASE(32)
        default: break;
    }
#undef AVBD_F64C_CASE
#undef AVBD_F64C_U
#undef AVBD_F64C_LAUNCH
}

void launch_cheb_dual_f64(hipStream_t s, const F64State& f, uint32_t n, float omega, const ResCtl* rc,
                          const float* omega_tab, uint32_t j) {
    if (n) hipLaunchKernelGGL(cheb_dual_f64_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, f, n, omega, rc,
                              omega_tab, j);
}

void launch_step_begin_f64(hipStream_t s, const F64State& f, uint32_t n, double dt, double g) {
    if (n) hipLaunchKernelGGL(step_begin_f64_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, f, n, dt, g);
}

void launch_eprev_f64(hipStream_t s, const F64State& f, uint32_t n_tets) {
    if (n_tets)
        hipLaunchKernelGGL(eprev_f64_kernel, dim3(blocks_for(n_tets, 256)), dim3(256), 0, s, f.tets, f.damp, f.x_prev,
                           n_tets);
}

void launch_velocity_f64(hipStream_t s, const F64State& f, uint32_t n, double dt, uint32_t* status) {
    if (n) hipLaunchKernelGGL(velocity_f64_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, f, n, 1.0 / dt, status);
}

void launch_settle_f64(hipStream_t s, double4* x, const double4* x_alt, uint32_t n, const uint32_t* iters_word,
                       uint32_t iters) {
    if (n) hipLaunchKernelGGL(settle_f64_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, x, x_alt, n, iters_word,
                              iters);
}

void launch_residual_f64(hipStream_t s, const F64State& f, const LaunchCfg& cfg, bool damping, double dt,
                         const double4* x, double* rr, const ResCtl* rc) {
    const uint32_t L = cfg.lanes, n = f.n_items;
    if (n == 0) return;
    const dim3 grid(blocks_for(n * L, 256)), blk(256);
    const double inv_dt = 1.0 / dt, inv_dt2 = inv_dt * inv_dt;
#define AVBD_F64R_CASE(LL)                                                                                        \
    case LL:                                                                                                      \
        if (damping)                                                                                              \
            hipLaunchKernelGGL((residual_f64_kernel<LL, true>), grid, blk, 0, s, f, f.items, n, x, rr, rc, inv_dt, \
                               inv_dt2);                                                                          \
        else                                                                                                      \
            hipLaunchKernelGGL((residual_f64_kernel<LL, false>), grid, blk, 0, s, f, f.items, n, x, rr, rc, inv_dt, \
                               inv_dt2);                                                                          \
        break;
    switch (L) {
        AVBD_F64R_CASE(4)
        AVBD_F64R_CASE(8)
        AVBD_F64R_CASE(16)
        AVBD_F64R_CASE(32)
        default: break;
    }
#undef AVBD_F64R_CASE
}

uint32_t launch_stats_f64(hipStream_t s, const F64State& f, const LaunchCfg& cfg, bool damping, double dt, double g,
                          double* out) {
    const uint32_t L = cfg.lanes;
    const uint32_t threads = std::max(f.n_items * L, f.n_tets + f.n_ground);
    const uint32_t nb = blocks_for(threads, 256);
    if (nb == 0) return 0;
    const double inv_dt = 1.0 / dt, inv_dt2 = inv_dt * inv_dt;
#define AVBD_F64S_CASE(LL)                                                                                         \
    case LL:                                                                                                       \
        if (damping)                                                                                               \
            hipLaunchKernelGGL((stats_f64_kernel<LL, true>), dim3(nb), dim3(256), 0, s, f, f.items, f.n_items,      \
                               f.n_tets, f.n_ground, out, inv_dt, inv_dt2, g);                                      \
        else                                                                                                       \
            hipLaunchKernelGGL((stats_f64_kernel<LL, false>), dim3(nb), dim3(256), 0, s, f, f.items, f.n_items,     \
                               f.n_tets, f.n_ground, out, inv_dt, inv_dt2, g);                                      \
        break;
    switch (L) {
        AVBD_F64S_CASE(4)
        AVBD_F64S_CASE(8)
        AVBD_F64S_CASE(16)
        AVBD_F64S_CASE(32)
        default: break;
    }
#undef AVBD_F64S_CASE
    return nb;
}

void launch_warm_save_f64(hipStream_t s, const F64State& f, uint32_t n) {
    if (n) hipLaunchKernelGGL(warm_save_f64_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, f, n);
}

void launch_warm_place_f64(hipStream_t s, const F64State& f, const float4* pred, const uint32_t* scene_of,
                           const float* step_of_scene) {
    if (f.n_items)
        hipLaunchKernelGGL(warm_place_f64_kernel, dim3(blocks_for(f.n_items, 256)), dim3(256), 0, s, f, f.items,
                           f.n_items, pred, scene_of, step_of_scene);
}

void launch_warm_target_f64(hipStream_t s, const F64State& f, float* tgt, double inv_s) {
    if (f.n_items)
        hipLaunchKernelGGL(warm_target_f64_kernel, dim3(blocks_for(f.n_items, 256)), dim3(256), 0, s, f, f.items,
                           f.n_items, tgt, inv_s);
}

}  // namespace avbd
