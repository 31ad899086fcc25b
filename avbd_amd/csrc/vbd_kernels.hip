// vbd_kernels.hip -- gfx950 kernels of one AVBD time step.
//
// Phase structure of VBDSolver::solve_Parallel (simulation/VBDSolver.hpp:48-103),
// one kernel per phase, every barrier of the reference becomes a kernel boundary:
//   eprev        NeoHookeanEnergySolver::updateAfterTimeStep      (only with damping)
//   step_begin   GroundCollisionEnergySolver::updateContactPoints +
//                CollisionConstraintEnergySolver::updateAfterTimeStep +
//                _particleInertialUpdate (adaptive initialisation), fused per particle
//   solve_color  _solveParticle<3> for every vertex of one colour
//   commit       _commitConflictedInColor
//   cheb_dual    _particleChebyshevAcceleration + CollisionConstraintEnergySolver::updateAfterIteration
//   velocity     _particleVelocityUpdate
//
// Vertex block solve: a group of L lanes (L | 64, groups never straddle a
// wavefront) owns one vertex; lane j handles incidences j, j+L, ...; the 3
// gradient and 6 (symmetric) Hessian partial sums are folded with
// cross-lane xor shuffles; the group leader adds the ground AL terms,
// solves the 3x3 system by LDL^T in registers and writes the vertex.
// Oriented (6-DOF) vertices of a colour are solved in the same phase, one
// thread each (rigid_device.hpp).
#include "internal.hpp"
#include "kernarg.hpp"
#include "dyn_device.hpp"
#include "rigid_device.hpp"
#include "warm_device.hpp"

namespace avbd {

namespace {

struct Sym3 {
    float a00, a11, a22, a01, a02, a12;
};

__device__ __forceinline__ float3 f3(float4 v) { return make_float3(v.x, v.y, v.z); }
__device__ __forceinline__ float3 operator-(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 operator+(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 operator*(float s, float3 a) { return make_float3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 cross3(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// Load a 64-byte tet record with four 16-byte loads.
__device__ __forceinline__ void load_tet(const TetRec* __restrict__ tets, uint32_t t, uint4& idx, float4& r1,
                                         float4& r2, float4& r3) {
    const uint4* p = reinterpret_cast<const uint4*>(tets + t);
    idx = p[0];
    const float4* f = reinterpret_cast<const float4*>(p);
    r1 = f[1];
    r2 = f[2];
    r3 = f[3];
}

// One tet incidence of the stable Neo-Hookean energy
// (NeoHookeanEnergySolver::accumulate, energy/elastic/NeoHookeanEnergySolver.hpp:121-227).
// q = Q.row(local) for local < 3, -(Q.row0 + Q.row1 + Q.row2) for local 3 (the
// reference's sign = -1 folded into q; every term is linear in q or quadratic).
// The record (idx, r1..r3) is already loaded; t indexes the damping record.
// nh_incidence_pos takes the four vertex positions already gathered.
template <bool DAMP>
__device__ __forceinline__ void nh_incidence_pos(float3 p0, float3 p1, float3 p2, float3 p3, float4 r1, float4 r2,
                                                 float4 r3, uint32_t local, const TetDamp* __restrict__ damp,
                                                 uint32_t t, float inv_dt, float3& g, Sym3& h) {
    // r1 = q00 q01 q02 q10 ; r2 = q11 q12 q20 q21 ; r3 = q22 VL VM alpha
    const float3 d0 = p0 - p3, d1 = p1 - p3, d2 = p2 - p3;
    const float q00 = r1.x, q01 = r1.y, q02 = r1.z, q10 = r1.w, q11 = r2.x, q12 = r2.y, q20 = r2.z, q21 = r2.w,
                q22 = r3.x;
    const float VL = r3.y, VM = r3.z, alpha = r3.w;
    // F = D Q: column c of F = d0 Q[0][c] + d1 Q[1][c] + d2 Q[2][c]
    const float3 f0 = q00 * d0 + q10 * d1 + q20 * d2;
    const float3 f1 = q01 * d0 + q11 * d1 + q21 * d2;
    const float3 f2 = q02 * d0 + q12 * d1 + q22 * d2;
    const float3 c0 = cross3(f1, f2), c1 = cross3(f2, f0), c2 = cross3(f0, f1);
    const float J = dot3(f0, c0);
    float3 q;
    if (local == 0) q = make_float3(q00, q01, q02);
    else if (local == 1) q = make_float3(q10, q11, q12);
    else if (local == 2) q = make_float3(q20, q21, q22);
    else q = make_float3(-(q00 + q10 + q20), -(q01 + q11 + q21), -(q02 + q12 + q22));
    const float3 Fq = q.x * f0 + q.y * f1 + q.z * f2;
    const float3 Fc = q.x * c0 + q.y * c1 + q.z * c2;
    const float hyd = VL * (J - alpha);
    const float q2 = dot3(q, q);
    g = g + (hyd * Fc + VM * Fq);
    const float dev = VM * q2;
    h.a00 += VL * Fc.x * Fc.x + dev;
    h.a11 += VL * Fc.y * Fc.y + dev;
    h.a22 += VL * Fc.z * Fc.z + dev;
    h.a01 += VL * Fc.x * Fc.y;
    h.a02 += VL * Fc.x * Fc.z;
    h.a12 += VL * Fc.y * Fc.z;
    if constexpr (DAMP) {
        const TetDamp dr = damp[t];
        const float dm = dr.damp;
        // Edot = (0.5 F^T F - 0.5 I - E_prev) / dt, symmetric
        const float e00 = (0.5f * dot3(f0, f0) - 0.5f - dr.e00) * inv_dt;
        const float e11 = (0.5f * dot3(f1, f1) - 0.5f - dr.e11) * inv_dt;
        const float e22 = (0.5f * dot3(f2, f2) - 0.5f - dr.e22) * inv_dt;
        const float e01 = (0.5f * dot3(f0, f1) - dr.e01) * inv_dt;
        const float e02 = (0.5f * dot3(f0, f2) - dr.e02) * inv_dt;
        const float e12 = (0.5f * dot3(f1, f2) - dr.e12) * inv_dt;
        const float3 Eq = make_float3(e00 * q.x + e01 * q.y + e02 * q.z, e01 * q.x + e11 * q.y + e12 * q.z,
                                      e02 * q.x + e12 * q.y + e22 * q.z);
        const float3 FEq = Eq.x * f0 + Eq.y * f1 + Eq.z * f2;
        g = g + dm * FEq;
        const float diag = dm * dot3(q, Eq);
        const float a = 0.5f * dm * inv_dt;
        // F F^T rows: (F F^T)_ij = sum_c F_ic F_jc
        const float ff00 = f0.x * f0.x + f1.x * f1.x + f2.x * f2.x;
        const float ff11 = f0.y * f0.y + f1.y * f1.y + f2.y * f2.y;
        const float ff22 = f0.z * f0.z + f1.z * f1.z + f2.z * f2.z;
        const float ff01 = f0.x * f0.y + f1.x * f1.y + f2.x * f2.y;
        const float ff02 = f0.x * f0.z + f1.x * f1.z + f2.x * f2.z;
        const float ff12 = f0.y * f0.z + f1.y * f1.z + f2.y * f2.z;
        h.a00 += diag + a * (q2 * ff00 + Fq.x * Fq.x);
        h.a11 += diag + a * (q2 * ff11 + Fq.y * Fq.y);
        h.a22 += diag + a * (q2 * ff22 + Fq.z * Fq.z);
        h.a01 += a * (q2 * ff01 + Fq.x * Fq.y);
        h.a02 += a * (q2 * ff02 + Fq.x * Fq.z);
        h.a12 += a * (q2 * ff12 + Fq.y * Fq.z);
    }
}

template <bool DAMP>
__device__ __forceinline__ void nh_incidence_rec(const float4* x, uint4 idx, float4 r1, float4 r2, float4 r3,
                                                 uint32_t local, const TetDamp* __restrict__ damp, uint32_t t,
                                                 float inv_dt, float3& g, Sym3& h) {
    nh_incidence_pos<DAMP>(f3(x[idx.x]), f3(x[idx.y]), f3(x[idx.z]), f3(x[idx.w]), r1, r2, r3, local, damp, t, inv_dt,
                           g, h);
}

template <bool DAMP>
__device__ __forceinline__ void nh_incidence(const float4* x, const TetRec* __restrict__ tets,
                                             const TetDamp* __restrict__ damp, uint32_t e, float inv_dt,
                                             float3& g, Sym3& h) {
    uint4 idx;
    float4 r1, r2, r3;
    load_tet(tets, e >> 2, idx, r1, r2, r3);
    nh_incidence_rec<DAMP>(x, idx, r1, r2, r3, e & 3u, damp, e >> 2, inv_dt, g, h);
}

// Incidence k of the work-item order, read from the per-incidence copy of
// the tet records (contiguous for an item's lanes, no index indirection);
// the local vertex index sits in the top two bits of idx.w.
template <bool DAMP>
__device__ __forceinline__ void nh_incidence_k(const float4* x, const TetRec* __restrict__ irec,
                                               const uint32_t* __restrict__ inc, const TetDamp* __restrict__ damp,
                                               uint32_t k, float inv_dt, float3& g, Sym3& h) {
    uint4 idx;
    float4 r1, r2, r3;
    load_tet(irec, k, idx, r1, r2, r3);
    const uint32_t local = idx.w >> 30;
    idx.w &= 0x3fffffffu;
    nh_incidence_rec<DAMP>(x, idx, r1, r2, r3, local, damp, DAMP ? (inc[k] >> 2) : 0u, inv_dt, g, h);
}

// Ground AL contact contribution (CollisionConstraintEnergySolver::accumulate<3>,
// energy/collision/CollisionConstraintEnergySolver.hpp:180-285 with the ground
// constraint of energy/collision/GroundCollisionEnergySolver.hpp:27-60).
__device__ __forceinline__ void ground_accumulate(const GroundRec& c, float3 p, float3& g, Sym3& h) {
    const float Cn = -p.y, Ct = p.x - c.cp_x, Cb = p.z - c.cp_z;
    const float Ccorr = Cn - kAlpha * c.C_n_prev;
    const float lnp = fmaxf(0.0f, c.k_n * Ccorr + c.lambda_n);
    float lt = c.k_t * Ct + c.lambda_t;
    float lb = c.k_b * Cb + c.lambda_b;
    const float mag = sqrtf(lt * lt + lb * lb);
    const float mu = (c.flags & GF_USE_STATIC) ? c.mu_s : c.mu_k;
    const float lmax = mu * lnp;
    if (mag > lmax) {
        const float nrm = mag + 1e-8f;
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

// Dynamic contacts incident to one positional particle (the reference
// accumulates them after the static energies, simulation/VBDSolver.hpp:400-438).
struct DynArgs {
    const uint32_t* off;   // null: no dynamic contacts
    const uint32_t* ent;
    const DynRec* rec;
    const float4* q;
    const int32_t* rot_of;
};
__host__ __device__ __forceinline__ DynArgs dyn_args(const DevState& d) {
    return DynArgs{d.dyn_off, d.dyn_ent, d.dyn, d.q, d.rot_of};
}
__device__ __forceinline__ void dyn_add3(const DynArgs& a, const float4* x, uint32_t v, float3& g, Sym3& h) {
    if (!a.off) return;
    const dyn::View view{x, a.q, a.rot_of};
    const uint32_t e1 = a.off[v + 1];
    for (uint32_t e = a.off[v]; e < e1; e++) {
        const uint32_t code = a.ent[e];
        const DynRec c = a.rec[code >> 3];
        dyn::accumulate3(view, c, code & 7u, g, h);
    }
}

// Solve (m/dt^2 I + H) dx = b for symmetric positive definite H by LDL^T.
__device__ __forceinline__ float3 ldlt3(const Sym3& A, float3 b) {
    const float d0 = A.a00;
    const float l10 = A.a01 / d0, l20 = A.a02 / d0;
    const float d1 = A.a11 - l10 * A.a01;
    const float l21 = (A.a12 - l20 * A.a01) / d1;
    const float d2 = A.a22 - l20 * A.a02 - l21 * l21 * d1;
    const float z0 = b.x;
    const float z1 = b.y - l10 * z0;
    const float z2 = b.z - l20 * z0 - l21 * z1;
    const float x2 = z2 / d2;
    const float x1 = z1 / d1 - l21 * x2;
    const float x0 = z0 / d0 - l10 * x1 - l20 * x2;
    return make_float3(x0, x1, x2);
}

template <int L>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Workgroups are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8),
// each with its own L2.  Remap so that XCD x runs a contiguous range of
// logical blocks: work items are ordered by vertex index, i.e. spatially, so
// each XCD's L2 keeps the tets and positions of one slab of the mesh across
// colour phases and iterations.
__device__ __forceinline__ uint32_t xcd_logical_block() {
    const uint32_t b = blockIdx.x, G = gridDim.x;
    const uint32_t x = b % 8u, q = b / 8u;
    return x * (G / 8u) + min(x, G % 8u) + q;
}

// A lane's incidences in batches of U (j0, j0 + L, ..., j0 + (U-1) L, those
// < cnt): a batch's records are loaded together, then its vertex positions
// together with the next batch's records, then its terms are accumulated in
// incidence order -- the same order and arithmetic as one incidence at a
// time, with fewer dependent memory round trips and the next records in
// flight during the arithmetic.
template <int U>
struct IncBatch {
    uint4 idx[U];
    float4 r1[U], r2[U], r3[U];
};

template <int L, int U>
__device__ __forceinline__ void load_inc_batch(const TetRec* __restrict__ irec, uint32_t base, uint32_t j0,
                                               uint32_t cnt, IncBatch<U>& b) {
#pragma unroll
    for (int u = 0; u < U; u++)
        if (j0 + u * L < cnt) load_tet(irec, base + j0 + u * L, b.idx[u], b.r1[u], b.r2[u], b.r3[u]);
}

template <int L, int U, bool DAMP>
__device__ __forceinline__ void nh_incidences_batch(const float4* x, const TetRec* __restrict__ irec,
                                                    const uint32_t* __restrict__ inc,
                                                    const TetDamp* __restrict__ damp, uint32_t base, uint32_t j0,
                                                    uint32_t cnt, float inv_dt, float3& g, Sym3& h,
                                                    IncBatch<U>& cur) {
    const uint4* idx = cur.idx;
    const float4 *r1 = cur.r1, *r2 = cur.r2, *r3 = cur.r3;
    float3 P[U][4];
#pragma unroll
    for (int u = 0; u < U; u++)
        if (j0 + u * L < cnt) {
            P[u][0] = f3(x[idx[u].x]);
            P[u][1] = f3(x[idx[u].y]);
            P[u][2] = f3(x[idx[u].z]);
            P[u][3] = f3(x[idx[u].w & 0x3fffffffu]);
        }
    IncBatch<U> next;
    if (j0 + U * L < cnt) load_inc_batch<L, U>(irec, base, j0 + U * L, cnt, next);
#pragma unroll
    for (int u = 0; u < U; u++)
        if (j0 + u * L < cnt)
            nh_incidence_pos<DAMP>(P[u][0], P[u][1], P[u][2], P[u][3], r1[u], r2[u], r3[u], idx[u].w >> 30, damp,
                                   DAMP ? (inc[base + j0 + u * L] >> 2) : 0u, inv_dt, g, h);
    cur = next;
}

// The tet incidences of one work item summed over the L lanes of its group
// (every lane gets the sums; every lane of the group must call it).  U > 1
// gathers U incidences of a lane at once (nh_incidences_batch).
template <int L, bool DAMP, int U>
__device__ __forceinline__ void item_gradient(const VItem& it, uint32_t lane, const uint32_t* __restrict__ inc,
                                              const TetRec* __restrict__ irec, const TetDamp* __restrict__ damp,
                                              const float4* x, float inv_dt, float3& g, Sym3& h) {
    const uint32_t cnt = it.inc_count & ~VITEM_CONFLICT;
    g = make_float3(0.f, 0.f, 0.f);
    h = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if constexpr (U > 1) {
        IncBatch<U> cur;
        if (lane < cnt) load_inc_batch<L, U>(irec, it.inc_begin, lane, cnt, cur);
        for (uint32_t j0 = lane; j0 < cnt; j0 += U * L)
            nh_incidences_batch<L, U, DAMP>(x, irec, inc, damp, it.inc_begin, j0, cnt, inv_dt, g, h, cur);
    } else {
        for (uint32_t j = lane; j < cnt; j += L)
            nh_incidence_k<DAMP>(x, irec, inc, damp, it.inc_begin + j, inv_dt, g, h);
    }
    g.x = group_sum<L>(g.x);
    g.y = group_sum<L>(g.y);
    g.z = group_sum<L>(g.z);
    h.a00 = group_sum<L>(h.a00);
    h.a11 = group_sum<L>(h.a11);
    h.a22 = group_sum<L>(h.a22);
    h.a01 = group_sum<L>(h.a01);
    h.a02 = group_sum<L>(h.a02);
    h.a12 = group_sum<L>(h.a12);
}

// The group leader's part of the vertex block solve: ground AL terms,
// dynamic contacts, inertia, and the 3x3 solve; returns the new position.
template <bool DYN>
__device__ __forceinline__ float4 item_solve(const VItem& it, float4 xv, float4 yv, const GroundRec& gr, float3 g,
                                             Sym3 h, const float4* x, const DynArgs& da, float inv_dt2) {
    const float3 p = f3(xv);
    if (it.ground >= 0) ground_accumulate(gr, p, g, h);
    if constexpr (DYN) dyn_add3(da, x, it.v, g, h);
    const float md = xv.w * inv_dt2;   // m / dt^2
    h.a00 += md;
    h.a11 += md;
    h.a22 += md;
    const float3 b = make_float3(-md * (p.x - yv.x) - g.x, -md * (p.y - yv.y) - g.y, -md * (p.z - yv.z) - g.z);
    const float3 dx = ldlt3(h, b);
    return make_float4(p.x + dx.x, p.y + dx.y, p.z + dx.z, xv.w);
}

// One vertex block solve (_solveParticle<3>, simulation/VBDSolver.hpp:285-449,487-510)
// by the L lanes of a group; every lane of the group must call it.
template <int L, bool DAMP, bool DYN = false, int U = 1>
__device__ __forceinline__ void solve_item(const VItem& it, uint32_t lane, const uint32_t* __restrict__ inc,
                                           const TetRec* __restrict__ irec, const TetDamp* __restrict__ damp,
                                           float4* x, const float4* __restrict__ y,
                                           const GroundRec* __restrict__ ground, float4* x_out, float4* x_buf,
                                           float inv_dt, float inv_dt2, const DynArgs& da) {
    // the group leader's own state is loaded first, so that its latency
    // overlaps the incidence gathers instead of following the reduction
    float4 xv = make_float4(0.f, 0.f, 0.f, 0.f), yv = xv;
    GroundRec gr{};
    const bool lead = lane == 0;
    if (lead) {
        xv = x[it.v];
        yv = y[it.v];
        if (it.ground >= 0) gr = ground[it.ground];
    }
    float3 g;
    Sym3 h;
    item_gradient<L, DAMP, U>(it, lane, inc, irec, damp, x, inv_dt, g, h);
    if (!lead) return;
    const float4 nx = item_solve<DYN>(it, xv, yv, gr, g, h, x, da, inv_dt2);
    if (it.inc_count & VITEM_CONFLICT) x_buf[it.v] = nx;
    else x_out[it.v] = nx;
}

template <int L, bool DAMP, bool DYN, int U>
__global__ __launch_bounds__(U > 1 ? 256 : 1024) void solve_color_kernel(const VItem* __restrict__ items, uint32_t n_items,
                                                           const uint32_t* __restrict__ inc,
                                                           const TetRec* __restrict__ irec,
                                                           const TetDamp* __restrict__ damp, float4* x,
                                                           const float4* __restrict__ y,
                                                           const GroundRec* __restrict__ ground, float4* x_out,
                                                           float4* x_buf, float inv_dt, float inv_dt2,
                                                           const uint32_t* __restrict__ done, DynArgs da,
                                                           const uint32_t* __restrict__ frozen,
                                                           const uint32_t* __restrict__ scene_of) {
    kernarg_prefetch(items, n_items, inc, irec, x, y, ground, x_out, done, frozen, scene_of, inv_dt, inv_dt2,
                     blockDim.x, gridDim.x);
    const uint32_t tid = xcd_logical_block() * blockDim.x + threadIdx.x;
    const uint32_t item = tid / L;
    const uint32_t lane = tid % L;
    if (item >= n_items) return;   // whole L-lane groups exit together (L divides the block)
    const VItem it = items[item];   // in flight with the converged flag's load
    if (done && *done) return;      // residual target reached earlier in this step
    if (frozen && frozen[scene_of[it.v]]) return;   // the item's scene has converged in this step
    solve_item<L, DAMP, DYN, U>(it, lane, inc, irec, damp, x, y, ground, x_out, x_buf, inv_dt, inv_dt2, da);
}

// E_prev = 0.5 (F_prev^T F_prev - I) from previous positions
// (NeoHookeanEnergySolver::updateAfterTimeStep, energy/elastic/NeoHookeanEnergySolver.hpp:70-90).
__device__ __forceinline__ void eprev_one(uint32_t t, const TetRec* __restrict__ tets, TetDamp* __restrict__ damp,
                                          const float4* __restrict__ x_prev) {
    uint4 idx;
    float4 r1, r2, r3;
    load_tet(tets, t, idx, r1, r2, r3);
    const float3 p0 = f3(x_prev[idx.x]), p1 = f3(x_prev[idx.y]), p2 = f3(x_prev[idx.z]), p3 = f3(x_prev[idx.w]);
    const float3 d0 = p0 - p3, d1 = p1 - p3, d2 = p2 - p3;
    const float3 f0 = r1.x * d0 + r1.w * d1 + r2.z * d2;
    const float3 f1 = r1.y * d0 + r2.x * d1 + r2.w * d2;
    const float3 f2 = r1.z * d0 + r2.y * d1 + r3.x * d2;
    TetDamp dr = damp[t];
    dr.e00 = 0.5f * (dot3(f0, f0) - 1.0f);
    dr.e11 = 0.5f * (dot3(f1, f1) - 1.0f);
    dr.e22 = 0.5f * (dot3(f2, f2) - 1.0f);
    dr.e01 = 0.5f * dot3(f0, f1);
    dr.e02 = 0.5f * dot3(f0, f2);
    dr.e12 = 0.5f * dot3(f1, f2);
    damp[t] = dr;
}

__global__ void eprev_kernel(const TetRec* __restrict__ tets, TetDamp* __restrict__ damp,
                             const float4* __restrict__ x_prev, uint32_t n_tets) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n_tets) eprev_one(t, tets, damp, x_prev);
}

// Stable Neo-Hookean energy of tet t at positions P (NeoHookeanEnergySolver::energy,
// energy/elastic/NeoHookeanEnergySolver.hpp:20-53), with the strain-rate
// damping term when DAMP, in FP64 accumulation.
template <bool DAMP>
__device__ __forceinline__ double nh_energy_dev(const float4* P, const TetRec* __restrict__ tets,
                                                const TetDamp* __restrict__ damp, uint32_t t, float inv_dt,
                                                uint32_t damp_index) {
    uint4 idx;
    float4 r1, r2, r3;
    load_tet(tets, t, idx, r1, r2, r3);
    const float3 p0 = f3(P[idx.x]), p1 = f3(P[idx.y]), p2 = f3(P[idx.z]), p3 = f3(P[idx.w]);
    const float3 d0 = p0 - p3, d1 = p1 - p3, d2 = p2 - p3;
    const float3 f0 = r1.x * d0 + r1.w * d1 + r2.z * d2;
    const float3 f1 = r1.y * d0 + r2.x * d1 + r2.w * d2;
    const float3 f2 = r1.z * d0 + r2.y * d1 + r3.x * d2;
    const double J = dot3(f0, cross3(f1, f2));
    const double VL = r3.y, VM = r3.z, al = r3.w;
    const double trc = (double)dot3(f0, f0) + dot3(f1, f1) + dot3(f2, f2);
    double e = 0.5 * VL * (J - al) * (J - al) + 0.5 * VM * (trc - 3.0);
    if constexpr (DAMP) {
        const TetDamp dr = damp[damp_index];
        const double E[6] = {0.5 * (dot3(f0, f0) - 1.0), 0.5 * (dot3(f1, f1) - 1.0), 0.5 * (dot3(f2, f2) - 1.0),
                             0.5 * dot3(f0, f1), 0.5 * dot3(f0, f2), 0.5 * dot3(f1, f2)};
        const double Pv[6] = {dr.e00, dr.e11, dr.e22, dr.e01, dr.e02, dr.e12};
        double sq = 0;
        for (int i = 0; i < 6; i++) {
            const double dd = (E[i] - Pv[i]) * inv_dt;
            sq += (i < 3 ? 1.0 : 2.0) * dd * dd;
        }
        e += 0.5 * dr.damp / inv_dt * sq;   // V kd |Edot|^2 with damp = 2 V kd / dt
    }
    return e;
}

// Start of step for one particle: ground contact point + AL warm start, the
// rigid-ground / joint records of an oriented particle, then the inertial
// target and adaptive initial guess (and the rotation inertial update).
__device__ __forceinline__ void step_begin_one(const DevState& d, uint32_t p, float dt, float g) {
    float4 xp = d.x[p];
    const int32_t gi = d.ground_of[p];
    if (gi >= 0) {
        GroundRec c = d.ground[gi];
        // updateContactPoints (GroundCollisionEnergySolver.hpp:81-91)
        c.cp_x = xp.x;
        c.cp_z = xp.z;
        // updateAfterTimeStep (CollisionConstraintEnergySolver.hpp:141-167)
        if (c.lambda_n > 0.f) c.C_n_prev = -xp.y;
        if ((c.flags & GF_UNCLAMPED) && !(c.flags & GF_USE_STATIC)) c.flags |= GF_USE_STATIC;
        c.lambda_n *= kAlpha * kGamma;
        c.lambda_t *= kGamma;
        c.lambda_b *= kGamma;
        c.k_n = fmaxf(kGamma * c.k_n, c.k_start);
        c.k_t = fmaxf(kGamma * c.k_t, c.k_start);
        c.k_b = fmaxf(kGamma * c.k_b, c.k_start);
        d.ground[gi] = c;
    }
    const int32_t r = d.n_rot ? d.rot_of[p] : -1;
    if (r >= 0) rigid::records_step_begin(d, p, (uint32_t)r);
    if (d.fixed[p]) return;
    const float4 vv = d.v[p], vp = d.v_prev[p];
    if (d.a_ext) {
        // with a per-particle field f (not in the reference): the external
        // acceleration gravity + f, in the inertial target and the adaptive
        // initial guess alike
        const float4 f = d.a_ext[p];
        const float ex = f.x, ey = f.y - g, ez = f.z;
        const float an = sqrtf(ex * ex + ey * ey + ez * ez);
        float at = 0.f;
        if (an > 1e-12f) {
            const float along = ((vv.x - vp.x) * ex + (vv.y - vp.y) * ey + (vv.z - vp.z) * ez) / dt / an;
            if (along > an) at = 1.f;
            else if (along < 0.f) at = 0.f;
            else at = along / an;
        }
        const float dt2 = dt * dt;
        d.y[p] = make_float4(xp.x + dt * vv.x + dt2 * ex, xp.y + dt * vv.y + dt2 * ey, xp.z + dt * vv.z + dt2 * ez, 0.f);
        xp.x = xp.x + (dt * vv.x + dt2 * (at * ex));
        xp.y = xp.y + (dt * vv.y + dt2 * (at * ey));
        xp.z = xp.z + (dt * vv.z + dt2 * (at * ez));
        d.x[p] = xp;
        d.x_last[p] = xp;
        d.x_lastlast[p] = xp;
        if (r >= 0) rigid::rotation_inertial(d, (uint32_t)r, dt);
        return;
    }
    // _particleInertialUpdate (VBDSolver.hpp:208-248): gravity (0,-g,0)
    const float ay = (vv.y - vp.y) / dt;
    float a_tilde = 0.f;
    if (g > 1e-12f) {
        const float along = (ay * -g) / g;   // a . a_ext / |a_ext|
        if (along > g) a_tilde = 1.f;
        else if (along < 0.f) a_tilde = 0.f;
        else a_tilde = along / g;
    }
    const float dt2 = dt * dt;
    d.y[p] = make_float4(xp.x + dt * vv.x, xp.y + dt * vv.y + dt2 * -g, xp.z + dt * vv.z, 0.f);
    xp.x = xp.x + dt * vv.x;
    xp.y = xp.y + (dt * vv.y + dt2 * (a_tilde * -g));
    xp.z = xp.z + dt * vv.z;
    d.x[p] = xp;
    d.x_last[p] = xp;
    d.x_lastlast[p] = xp;
    if (r >= 0) rigid::rotation_inertial(d, (uint32_t)r, dt);
}

__global__ void step_begin_kernel(DevState d, float dt, float g) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < d.n) step_begin_one(d, p, dt, g);
}

// _commitConflictedInColor (VBDSolver.hpp:108-119) for one buffered particle:
// positions only, as in the reference -- the buffered rotation of a
// conflicted oriented particle is not committed.
__device__ __forceinline__ void commit_one(const DevState& d, uint32_t p) { d.x[p] = d.x_buf[p]; }

__global__ void commit_kernel(DevState d, const uint32_t* __restrict__ list, uint32_t count,
                              const uint32_t* __restrict__ done) {
    if (done && *done) return;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) commit_one(d, list[k]);
}

// _particleChebyshevAcceleration (VBDSolver.hpp:513-536) of one position.
__device__ __forceinline__ float4 cheb_apply(float4 xp, float4 xll, float omega) {
    xp.x = omega * (xp.x - xll.x) + xll.x;
    xp.y = omega * (xp.y - xll.y) + xll.y;
    xp.z = omega * (xp.z - xll.z) + xll.z;
    return xp;
}

// CollisionConstraintEnergySolver::updateAfterIteration (CollisionConstraintEnergySolver.hpp:60-134)
// of one ground record at the extrapolated position xp.
__device__ __forceinline__ void ground_dual_update(GroundRec& c, float4 xp) {
    const float Cn = -xp.y, Ct = xp.x - c.cp_x, Cb = xp.z - c.cp_z;
    const float Ccorr = Cn - kAlpha * c.C_n_prev;
    const float lnp = c.k_n * Ccorr + c.lambda_n;
    const float ltp = c.k_t * Ct + c.lambda_t;
    const float lbp = c.k_b * Cb + c.lambda_b;
    if (lnp > 0.f) c.k_n += kBeta * fabsf(Ccorr);
    c.lambda_n = fmaxf(0.f, lnp);
    const float mag = sqrtf(ltp * ltp + lbp * lbp);
    float lmax = ((c.flags & GF_USE_STATIC) ? c.mu_s : c.mu_k) * c.lambda_n;
    if (mag < lmax) {
        c.k_t += kBeta * fabsf(Ct);
        c.k_b += kBeta * fabsf(Cb);
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

// Chebyshev step (positions and orientations) and the dual / stiffness update
// of the particle's contact and joint records; the position is read from
// x_in and written to x_out (the same buffer unless the step is fused into
// the last colour phase, solve_cheb_kernel).
__device__ __forceinline__ void cheb_dual_io(const DevState& d, uint32_t p, float omega, const float4* x_in,
                                             float4* x_out) {
    const float4 xl = d.x_last[p], xll = d.x_lastlast[p];
    const float4 xp = cheb_apply(x_in[p], xll, omega);
    x_out[p] = xp;
    d.x_lastlast[p] = xl;
    d.x_last[p] = xp;
    const int32_t r = d.n_rot ? d.rot_of[p] : -1;
    float4 q = make_float4(0.f, 0.f, 0.f, 1.f);
    if (r >= 0) q = rigid::rotation_chebyshev(d, (uint32_t)r, omega);
    const int32_t gi = d.ground_of[p];
    if (gi >= 0) {
        GroundRec c = d.ground[gi];
        ground_dual_update(c, xp);
        d.ground[gi] = c;
    }
    if (r >= 0) rigid::records_dual(d, make_float3(xp.x, xp.y, xp.z), q, (uint32_t)r);
}

__device__ __forceinline__ void cheb_dual_one(const DevState& d, uint32_t p, float omega) {
    cheb_dual_io(d, p, omega, d.x, d.x);
}

// Chebyshev weight of local iteration j of the running check interval: from
// the per-iteration table at the step's iteration counter in residual mode
// (one interval graph serves every interval of a step), else the weight
// baked into the launch.
__device__ __forceinline__ float cheb_weight(float omega, const ResCtl* rc, const float* omega_tab, uint32_t j) {
    return rc ? omega_tab[rc->iter + j] : omega;
}

__global__ void cheb_dual_kernel(DevState d, float omega, const ResCtl* rc, const float* __restrict__ omega_tab,
                                 uint32_t j) {
    if (rc && rc->done) return;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.n || (d.frozen && d.frozen[d.scene_of[p]])) return;
    cheb_dual_one(d, p, cheb_weight(omega, rc, omega_tab, j));
}

// The last colour phase of an iteration fused with the Chebyshev / dual
// update that follows it (scenes without oriented particles or conflicts).
// Positions are double-buffered across iterations: this iteration's solves
// read and write d.x; the extrapolated positions go to x_next, which the
// next iteration uses as d.x.  Per XCD the grid holds `sx` solve blocks and
// then `cx` Chebyshev blocks (hardware block b runs on XCD b mod 8):
//   * a solve group computes its vertex of this colour and its leader
//     extrapolates that result straight into x_next and updates its ground
//     record, exactly as cheb_dual_io would after the phase;
//   * a Chebyshev thread extrapolates one vertex of the other colours (and
//     the unsolved ones) from d.x into x_next.
// d.x is not written, so every solve still reads its neighbours' positions
// before extrapolation, as with a separate Chebyshev launch.
template <int L, bool DAMP, int U>
__global__ __launch_bounds__(U > 1 ? 256 : 1024) void solve_cheb_kernel(
    DevState d, const VItem* __restrict__ items, uint32_t n_items, uint32_t sx, uint32_t cx,
    const uint32_t* __restrict__ cheb_list, uint32_t n_cheb, float4* __restrict__ x_next, float omega_fixed,
    const ResCtl* rc, const float* __restrict__ omega_tab, uint32_t j, float inv_dt, float inv_dt2) {
    kernarg_prefetch(d.x, d.y, d.x_last, d.x_lastlast, d.ground, d.ground_of, d.inc, d.irec, d.damp, d.frozen,
                     d.scene_of, items, n_items, sx, cx, cheb_list, n_cheb, x_next, rc, inv_dt, inv_dt2, blockDim.x);
    const uint32_t xcd = blockIdx.x % 8u, q = blockIdx.x / 8u;
    if (q < sx) {
        const uint32_t tid = (xcd * sx + q) * blockDim.x + threadIdx.x;
        const uint32_t item = tid / L, lane = tid % L;
        if (item >= n_items) return;   // whole L-lane groups exit together
        const VItem it = items[item];   // in flight with the converged flag's load
        if (rc && rc->done) return;
        const float omega = cheb_weight(omega_fixed, rc, omega_tab, j);
        const uint32_t v = it.v;
        if (d.frozen && d.frozen[d.scene_of[v]]) {   // converged scene: keep its position in both buffers
            if (lane == 0) x_next[v] = d.x[v];
            return;
        }
        float4 xv = make_float4(0.f, 0.f, 0.f, 0.f), yv = xv, xl = xv, xll = xv;
        GroundRec gr{};
        const bool lead = lane == 0;
        if (lead) {
            xv = d.x[v];
            yv = d.y[v];
            xl = d.x_last[v];
            xll = d.x_lastlast[v];
            if (it.ground >= 0) gr = d.ground[it.ground];
        }
        float3 g;
        Sym3 h;
        item_gradient<L, DAMP, U>(it, lane, d.inc, d.irec, d.damp, d.x, inv_dt, g, h);
        if (!lead) return;
        const float4 xp = cheb_apply(item_solve<false>(it, xv, yv, gr, g, h, d.x, DynArgs{}, inv_dt2), xll, omega);
        x_next[v] = xp;
        d.x_lastlast[v] = xl;
        d.x_last[v] = xp;
        if (it.ground >= 0) {
            ground_dual_update(gr, xp);
            d.ground[it.ground] = gr;
        }
    } else {
        const uint32_t k = (xcd * cx + (q - sx)) * blockDim.x + threadIdx.x;
        if (k < n_cheb) {
            const uint32_t p = cheb_list[k];
            if (rc && rc->done) return;
            const float omega = cheb_weight(omega_fixed, rc, omega_tab, j);
            if (d.frozen && d.frozen[d.scene_of[p]]) x_next[p] = d.x[p];
            else cheb_dual_io(d, p, omega, d.x, x_next);
        }
    }
}

// End of a step whose iterations ran on two position buffers: when an odd
// number of iterations ran (iters_word, or `iters` when it is null), the
// final positions are in x_alt.
__global__ void settle_kernel(float4* __restrict__ x, const float4* __restrict__ x_alt, uint32_t n,
                              const uint32_t* __restrict__ iters_word, uint32_t iters) {
    const uint32_t it = iters_word ? *iters_word : iters;
    if ((it & 1u) == 0) return;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) x[p] = x_alt[p];
}

// _particleVelocityUpdate (VBDSolver.hpp:538-566) for one particle.
// A non-finite position or velocity (or orientation / angular velocity)
// raises bit 0 of d.status; the host reports it at the next synchronising call.
__device__ __forceinline__ void velocity_one(const DevState& d, uint32_t p, float inv_dt) {
    const float4 xp = d.x[p], xo = d.x_prev[p], vv = d.v[p];
    d.v_prev[p] = vv;
    const float4 nv = make_float4((xp.x - xo.x) * inv_dt, (xp.y - xo.y) * inv_dt, (xp.z - xo.z) * inv_dt, 0.f);
    d.v[p] = nv;
    d.x_prev[p] = xp;
    bool ok = isfinite(xp.x) && isfinite(xp.y) && isfinite(xp.z) && isfinite(nv.x) && isfinite(nv.y) && isfinite(nv.z);
    const int32_t r = d.n_rot ? d.rot_of[p] : -1;
    if (r >= 0) {
        rigid::rotation_velocity(d, (uint32_t)r, inv_dt);
        const float4 q = d.q[r], w = d.w[r];
        ok = ok && isfinite(q.x) && isfinite(q.y) && isfinite(q.z) && isfinite(q.w) && isfinite(w.x) &&
             isfinite(w.y) && isfinite(w.z);
    }
    if (!ok && d.status) atomicOr(d.status, 1u);
}

__global__ void velocity_kernel(DevState d, float inv_dt) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < d.n) velocity_one(d, p, inv_dt);
}

// ---- dynamic contacts -------------------------------------------------------
// One detection round's ADD / UPDATE ops (CollisionConstraintEnergyPool::addEnergy,
// energy/collision/CollisionConstraintEnergyPool.hpp:48-69, and _updateCollision,
// collision/CollisionDetector.cpp:454-501).
__global__ void dyn_apply_kernel(DynRec* __restrict__ recs, const DynOp* __restrict__ ops, uint32_t n_ops) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_ops) return;
    const DynOp o = ops[k];
    DynRec& c = recs[o.rec];
    if (o.op == DYN_OP_ADD) {
        DynRec r;
        r.n = make_float4(o.n[0], o.n[1], o.n[2], o.k_start);
        r.t = make_float4(o.t[0], o.t[1], o.t[2], kContactMuS);
        r.b = make_float4(o.b[0], o.b[1], o.b[2], kContactMuK);
        r.bary = make_float4(o.bary[0], o.bary[1], o.bary[2], o.s);
        r.cpl = make_float4(o.cpl[0], o.cpl[1], o.cpl[2], 0.f);
        r.k = make_float4(o.k_start, o.k_start, o.k_start, 0.f);
        r.lam = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int i = 0; i < 5; i++) r.idx[i] = o.idx[i];
        r.kind_flags = o.kind;
        r.pad[0] = r.pad[1] = 0u;
        c = r;
    } else {
        DynRec r = c;
        if (dyn::kind_of(r) == DYN_TRI_ROD) r.bary.w = o.s;
        r.cpl = make_float4(o.cpl[0], o.cpl[1], o.cpl[2], r.cpl.w);
        r.bary = make_float4(o.bary[0], o.bary[1], o.bary[2], r.bary.w);
        float3 nn = make_float3(o.n[0], o.n[1], o.n[2]);
        // a vertex-triangle contact keeps the side it was created with
        if (dyn::kind_of(r) == DYN_VERTEX_TRI && nn.x * r.n.x + nn.y * r.n.y + nn.z * r.n.z < 0.f)
            nn = make_float3(-nn.x, -nn.y, -nn.z);
        dyn::update_frame(r, nn);
        c = r;
    }
}

__global__ void dyn_step_begin_kernel(DevState d, const uint32_t* __restrict__ active, uint32_t n_active) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_active) return;
    DynRec c = d.dyn[active[k]];
    dyn::update_after_timestep(dyn::view_of(d), c);
    d.dyn[active[k]] = c;
}

// Two-sided fixed joints (records [n_joints1, n_joints)): their start-of-step
// update, before the inertial update moves either body, and their dual /
// stiffness update after the Chebyshev step of both bodies.  Kept out of the
// per-particle kernels because each reads two particles.
__global__ void joint2_step_begin_kernel(DevState d) {
    const uint32_t e = d.n_joints1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= d.n_joints) return;
    JointRec j = d.joints[e];
    rigid::joint2_step_begin(d, j);
    d.joints[e] = j;
}

__global__ void joint2_dual_kernel(DevState d, const uint32_t* __restrict__ done) {
    if (done && *done) return;
    const uint32_t e = d.n_joints1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= d.n_joints) return;
    JointRec j = d.joints[e];
    if (d.frozen && d.frozen[d.scene_of[j.p]]) return;
    rigid::joint2_dual(d, j);
    d.joints[e] = j;
}

__global__ void dyn_dual_kernel(DevState d, const uint32_t* __restrict__ active, uint32_t n_active,
                                const uint32_t* __restrict__ done) {
    if (done && *done) return;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_active) return;
    DynRec c = d.dyn[active[k]];
    dyn::update_after_iteration(dyn::view_of(d), c);
    d.dyn[active[k]] = c;
}

// Oriented vertices of one colour, one thread each.
template <bool DYN>
__global__ __launch_bounds__(256) void solve_ocolor_kernel(DevState d, uint32_t begin, uint32_t count,
                                                           float inv_dt2, const uint32_t* __restrict__ done) {
    if (done && *done) return;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const OItem it = d.oitems[begin + k];
    if (d.frozen && d.frozen[d.scene_of[it.v]]) return;
    rigid::solve_oitem<DYN>(d, it, inv_dt2);
}

// ---- force residual (residual-targeted iteration) --------------------------
// r_v = m/dt^2 (x_v - y_v) + dE/dx_v for one positional work item: the
// gradient of the vertex solve at the current iterate; |r_v|^2 is returned by
// lane 0 of the group (0 elsewhere).  Every lane of the group must call it.
template <int L, bool DAMP>
__device__ __forceinline__ float residual_item(const VItem& it, uint32_t lane, const DevState& d, float inv_dt,
                                               float inv_dt2) {
    const uint32_t cnt = it.inc_count & ~VITEM_CONFLICT;
    float3 g = make_float3(0.f, 0.f, 0.f);
    Sym3 h = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (uint32_t j = lane; j < cnt; j += L) nh_incidence_k<DAMP>(d.x, d.irec, d.inc, d.damp, it.inc_begin + j, inv_dt, g, h);
    g.x = group_sum<L>(g.x);
    g.y = group_sum<L>(g.y);
    g.z = group_sum<L>(g.z);
    if (lane != 0) return 0.f;
    const float4 xv = d.x[it.v], yv = d.y[it.v];
    if (it.ground >= 0) ground_accumulate(d.ground[it.ground], f3(xv), g, h);
    const float md = xv.w * inv_dt2;
    const float rx = md * (xv.x - yv.x) + g.x, ry = md * (xv.y - yv.y) + g.y, rz = md * (xv.z - yv.z) + g.z;
    return rx * rx + ry * ry + rz * rz;
}

// Positional block of the residual of an oriented work item.
__device__ __forceinline__ float residual_oitem(const DevState& d, const OItem& it, float inv_dt2) {
    const float4 xv = d.x[it.v];
    rot::Vec6 G;
    rot::Sym6 H;
    rot::vec6_zero(G);
    rot::sym6_zero(H);
    rigid::accumulate6(d, it, rigid::xyz(xv), d.q[it.rot], G, H);
    const float4 yv = d.y[it.v];
    const float md = xv.w * inv_dt2;
    const float rx = md * (xv.x - yv.x) + G.v[0], ry = md * (xv.y - yv.y) + G.v[1], rz = md * (xv.z - yv.z) + G.v[2];
    return rx * rx + ry * ry + rz * rz;
}

// |r_v|^2 of every work item into rr[v] (positional and oriented work items
// are disjoint particles; fixed particles keep the 0 the buffer starts with).
template <int L, bool DAMP>
__global__ __launch_bounds__(256) void residual_pass_kernel(DevState d, const ResCtl* rc, double* __restrict__ rr,
                                                            float inv_dt, float inv_dt2) {
    if (rc->done) return;
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t item = tid / L, lane = tid % L;
    // a converged scene keeps the residual it converged with
    if (item < d.n_items) {
        const VItem it = d.items[item];
        if (!(d.frozen && d.frozen[d.scene_of[it.v]])) {
            const float r = residual_item<L, DAMP>(it, lane, d, inv_dt, inv_dt2);
            if (lane == 0) rr[it.v] = r;
        }
    }
    if (tid < d.n_oitems) {
        const OItem it = d.oitems[tid];
        if (!(d.frozen && d.frozen[d.scene_of[it.v]])) rr[it.v] = residual_oitem(d, it, inv_dt2);
    }
}

// One block per scene: sum of rr over the scene's particles in a fixed order
// (reproducible, no atomics).
// A scene that is within the target now (or has no free particle) freezes
// for the rest of the step; its iteration count is the step's count so far.
__global__ __launch_bounds__(1024) void residual_scene_kernel(const ResCtl* rc, const double* __restrict__ rr,
                                                              const uint32_t* __restrict__ scene_off,
                                                              double* __restrict__ acc_scene,
                                                              const double* __restrict__ mg2_scene, double tol2,
                                                              uint32_t every, SceneFreeze fz) {
    kernarg_prefetch(rc, rr, scene_off, acc_scene, mg2_scene, tol2, every, fz.frozen, fz.iters);
    __shared__ double red[16];
    const uint32_t s = blockIdx.x, b = scene_off[s], e = scene_off[s + 1];   // in flight with the flag's load
    if (rc->done) return;
    if (fz.frozen && fz.frozen[s]) return;   // its sum is the one it converged with
    // four independent loads in flight per thread; the grouping depends only
    // on the offset from the scene's first particle, so a scene sums the
    // same way wherever it sits in a batch
    double v0 = 0, v1 = 0, v2 = 0, v3 = 0;
    uint32_t p = b + threadIdx.x;
    for (; p + 3 * 1024 < e; p += 4 * 1024) {
        const double a0 = rr[p], a1 = rr[p + 1024], a2 = rr[p + 2048], a3 = rr[p + 3072];
        v0 += a0;
        v1 += a1;
        v2 += a2;
        v3 += a3;
    }
    for (; p < e; p += 1024) v0 += rr[p];
    double v = (v0 + v1) + (v2 + v3);
    // wavefront butterfly, then the 16 wavefront sums in order
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63u) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int w = 0; w < 16; w++) t += red[w];
        red[0] = t;
        acc_scene[s] = red[0];
        if (fz.frozen && red[0] <= tol2 * mg2_scene[s]) {
            fz.frozen[s] = 1;
            fz.iters[s] = rc->iter + every;
        }
    }
}

// The decision after a check interval of `every` iterations: every scene
// must reach its own target, ||r_s||^2 <= tol^2 ||M_s g||^2 (scenes without
// a free particle are skipped).  Records the check (the whole batch's
// relative residual and the worst scene's) and advances the step's
// iteration counter, which the Chebyshev launches read their weight by.
__global__ __launch_bounds__(256) void residual_decide_kernel(ResCtl* rc, const double* __restrict__ acc_scene,
                                                              const double* __restrict__ mg2_scene,
                                                              uint32_t n_scenes, double tol2, uint32_t every) {
    kernarg_prefetch(rc, acc_scene, mg2_scene, n_scenes, tol2, every);
    __shared__ double r_s[256], m_s[256], w_s[256];
    double sum = 0, mg = 0, w = 0;
    for (uint32_t s = threadIdx.x; s < n_scenes; s += 256) {
        const double a = acc_scene[s], m = mg2_scene[s];
        sum += a;
        mg += m;
        if (m > 0) w = fmax(w, a / m);
    }
    if (rc->done) return;   // the same for the whole block, tested after the sums' loads are in flight
    r_s[threadIdx.x] = sum;
    m_s[threadIdx.x] = mg;
    w_s[threadIdx.x] = w;
    __syncthreads();
    for (uint32_t st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            r_s[threadIdx.x] += r_s[threadIdx.x + st];
            m_s[threadIdx.x] += m_s[threadIdx.x + st];
            w_s[threadIdx.x] = fmax(w_s[threadIdx.x], w_s[threadIdx.x + st]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const uint32_t k = rc->checks;
        rc->checks = k + 1;
        rc->iter += every;
        if (k < kMaxResidualChecks) {
            rc->rel[k] = m_s[0] > 0 ? sqrt(r_s[0] / m_s[0]) : 0.0;
            rc->worst[k] = sqrt(w_s[0]);
        }
        if (w_s[0] <= tol2) {
            rc->done = 1;
            rc->last_iters = rc->iter;
        }
    }
}

// Start of a residual-targeted step: nothing converged, no iteration run.
__global__ void residual_reset_kernel(ResCtl* rc, SceneFreeze fz, uint32_t n_scenes) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) {
        rc->done = 0;
        rc->checks = 0;
        rc->iter = 0;
    }
    if (fz.frozen && t < n_scenes) fz.frozen[t] = 0;
}

// End of a residual-targeted step: a step that never converged used the
// whole budget (the iterations it ran); a scene still running at the end ran
// all of them; the vertex-block solves of the step are every scene's
// iterations times its work items.
__global__ __launch_bounds__(256) void residual_finish_kernel(ResCtl* rc, SceneFreeze fz, uint32_t n_scenes,
                                                              uint32_t n_items) {
    __shared__ unsigned long long red[256];
    const uint32_t iters_run = rc->done ? rc->last_iters : rc->iter;
    unsigned long long v = 0;
    if (fz.frozen) {
        for (uint32_t s = threadIdx.x; s < n_scenes; s += 256) {
            if (!fz.frozen[s]) fz.iters[s] = iters_run;
            v += (unsigned long long)fz.iters[s] * fz.items[s];
        }
    } else if (threadIdx.x == 0) {
        v = (unsigned long long)iters_run * n_items;
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (!rc->done) rc->last_iters = rc->iter;
        rc->total_iters += rc->last_iters;
        if (!rc->done) rc->budget_steps += 1;
        rc->vertex_solves += red[0];
    }
}

// ---- learned warm start: features, prediction apply, IP comparison ---------
// Green-strain norm ||E||_F, E = (F^T F - I)/2, of tet t at positions P;
// nfixed += 1 when one of its vertices is fixed.
__device__ __forceinline__ float tet_strain(const float4* P, const TetRec* __restrict__ tets,
                                            const uint8_t* __restrict__ fixed, uint32_t t, float& nfixed) {
    uint4 idx;
    float4 r1, r2, r3;
    load_tet(tets, t, idx, r1, r2, r3);
    const float3 p0 = f3(P[idx.x]), p1 = f3(P[idx.y]), p2 = f3(P[idx.z]), p3 = f3(P[idx.w]);
    const float3 d0 = p0 - p3, d1 = p1 - p3, d2 = p2 - p3;
    const float3 f0 = r1.x * d0 + r1.w * d1 + r2.z * d2;
    const float3 f1 = r1.y * d0 + r2.x * d1 + r2.w * d2;
    const float3 f2 = r1.z * d0 + r2.y * d1 + r3.x * d2;
    const float e00 = 0.5f * (dot3(f0, f0) - 1.f), e11 = 0.5f * (dot3(f1, f1) - 1.f), e22 = 0.5f * (dot3(f2, f2) - 1.f);
    const float e01 = 0.5f * dot3(f0, f1), e02 = 0.5f * dot3(f0, f2), e12 = 0.5f * dot3(f1, f2);
    nfixed += (fixed[idx.x] | fixed[idx.y] | fixed[idx.z] | fixed[idx.w]) ? 1.f : 0.f;
    return sqrtf(e00 * e00 + e11 * e11 + e22 * e22 + 2.f * (e01 * e01 + e02 * e02 + e12 * e12));
}

template <int L>
__device__ __forceinline__ float group_max(float v) {
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
    return v;
}

__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Per positional work item at the start of a step (after the inertial /
// adaptive initialisation): the 24 features of the warm-start network.
//   0-2   (y - x_prev) / (dt^2 g)          inertial displacement
//   3-5   dt v / (dt^2 g)                   previous-step displacement
//   6-8   dt (v - v_prev) / (dt^2 g)        change of displacement
//   9     log10 m
//   10,11 mean / max Green-strain norm of the incident tets at x_prev
//   12    contact flag (ground multiplier > 0 or within 1 mm of the ground)
//   13    fraction of incident tets that touch a fixed vertex
//   14,15 log10 E, nu of the incident material
//   16    height above the ground (0 without a ground record)
//   17    ground normal multiplier / (m g)
//   18-20 elastic force at x_prev / (m g)
//   21-23 force residual at the initial guess / (m g)
// Normalised features go to x16 (bf16, rows of 32, zero padding); raw ones
// to raw (rows of 24) when it is given.
template <int L>
__global__ __launch_bounds__(256) void warm_features_kernel(DevState d, WarmIO io, float dt, float g_acc,
                                                            uint32_t rows_pad) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t item = tid / L, lane = tid % L;
    if (item >= rows_pad) return;   // whole groups
    if (item >= d.n_items) {
        if (lane == 0)
            for (int f = 0; f < kMlpInPad; f++) io.x16[(size_t)item * kMlpInPad + f] = 0;
        return;
    }
    const VItem it = d.items[item];
    const uint32_t cnt = it.inc_count & ~VITEM_CONFLICT;
    const float inv_dt = 1.f / dt;
    float3 gp = make_float3(0.f, 0.f, 0.f), gi = make_float3(0.f, 0.f, 0.f);
    Sym3 h = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float ssum = 0.f, smax = 0.f, nfix = 0.f;
    for (uint32_t j = lane; j < cnt; j += L) {
        const uint32_t e = d.inc[it.inc_begin + j];
        nh_incidence<false>(d.x_prev, d.tets, d.damp, e, inv_dt, gp, h);
        nh_incidence<false>(d.x, d.tets, d.damp, e, inv_dt, gi, h);
        const float sn = tet_strain(d.x_prev, d.tets, d.fixed, e >> 2, nfix);
        ssum += sn;
        smax = fmaxf(smax, sn);
    }
    gp.x = group_sum<L>(gp.x);
    gp.y = group_sum<L>(gp.y);
    gp.z = group_sum<L>(gp.z);
    gi.x = group_sum<L>(gi.x);
    gi.y = group_sum<L>(gi.y);
    gi.z = group_sum<L>(gi.z);
    ssum = group_sum<L>(ssum);
    nfix = group_sum<L>(nfix);
    smax = group_max<L>(smax);
    if (lane != 0) return;
    const float4 x0 = d.x[it.v], xp = d.x_prev[it.v], yv = d.y[it.v], vv = d.v[it.v], vpv = d.v_prev[it.v];
    const float m = x0.w, mg = m * g_acc, sc = 1.f / (dt * dt * g_acc);
    float f[kMlpIn];
    f[0] = (yv.x - xp.x) * sc;
    f[1] = (yv.y - xp.y) * sc;
    f[2] = (yv.z - xp.z) * sc;
    f[3] = dt * vv.x * sc;
    f[4] = dt * vv.y * sc;
    f[5] = dt * vv.z * sc;
    f[6] = dt * (vv.x - vpv.x) * sc;
    f[7] = dt * (vv.y - vpv.y) * sc;
    f[8] = dt * (vv.z - vpv.z) * sc;
    f[9] = log10f(fmaxf(m, 1e-30f));
    f[10] = cnt ? ssum / (float)cnt : 0.f;
    f[11] = smax;
    f[12] = 0.f;
    f[16] = 0.f;
    f[17] = 0.f;
    if (it.ground >= 0) {
        const GroundRec c = d.ground[it.ground];
        f[12] = (c.lambda_n > 0.f || xp.y < 1e-3f) ? 1.f : 0.f;
        f[16] = xp.y;
        f[17] = c.lambda_n / mg;
        ground_accumulate(c, f3(x0), gi, h);
    }
    f[13] = cnt ? nfix / (float)cnt : 0.f;
    const float2 mat = io.mat[it.v];
    f[14] = mat.x;
    f[15] = mat.y;
    f[18] = gp.x / mg;
    f[19] = gp.y / mg;
    f[20] = gp.z / mg;
    const float md = m / (dt * dt);
    f[21] = (md * (x0.x - yv.x) + gi.x) / mg;
    f[22] = (md * (x0.y - yv.y) + gi.y) / mg;
    f[23] = (md * (x0.z - yv.z) + gi.z) / mg;
    uint16_t* row = io.x16 + (size_t)item * kMlpInPad;
    for (int k = 0; k < kMlpIn; k++) {
        row[k] = to_bf16((f[k] - io.norm[k]) / io.norm[kMlpIn + k]);
        if (io.raw) io.raw[(size_t)item * kMlpIn + k] = f[k];
    }
    for (int k = kMlpIn; k < kMlpInPad; k++) row[k] = 0;
}

// Keep the initial guess of every particle (x -> x_scratch) and clear the
// predicted step (pred = 0 for particles the network does not move).
__global__ void warm_save_kernel(DevState d, WarmIO io) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.n) return;
    d.x_scratch[p] = d.x[p];
    io.pred[p] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// pred = dt^2 g * (network output, de-normalised) for every positional work item.
__global__ void warm_apply_kernel(DevState d, WarmIO io, float s) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.n_items) return;
    const uint32_t v = d.items[i].v;
    const float* o = io.out + (size_t)i * kMlpOutPad;
    const float* tn = io.norm + 2 * kMlpIn;   // target mean (3), std (3)
    io.pred[v] = make_float4(s * (o[0] * tn[3] + tn[0]), s * (o[1] * tn[4] + tn[1]), s * (o[2] * tn[5] + tn[2]), 0.f);
}

__device__ __forceinline__ double ground_al_energy(const GroundRec& c, float4 xp) {
    const double Cn = -xp.y, Ct = (double)xp.x - c.cp_x, Cb = (double)xp.z - c.cp_z;
    const double Cc = Cn - (double)kAlpha * c.C_n_prev;
    double e = 0;
    if ((double)c.k_n * Cc + c.lambda_n > 0) e += 0.5 * c.k_n * Cc * Cc + (double)c.lambda_n * Cc;
    return e + 0.5 * c.k_t * Ct * Ct + (double)c.lambda_t * Ct + 0.5 * c.k_b * Cb * Cb + (double)c.lambda_b * Cb;
}


__device__ __forceinline__ float4 along(const float4* x0, const float4* pred, uint32_t p, float a) {
    const float4 b = x0[p], d = pred[p];
    return make_float4(b.x + a * d.x, b.y + a * d.y, b.z + a * d.z, b.w);
}

// Stable Neo-Hookean energy (+ damping) of tet t at x0 + a * pred.
template <bool DAMP>
__device__ __forceinline__ double nh_energy_along(const float4* x0, const float4* pred, float a,
                                                  const TetRec* __restrict__ tets, const TetDamp* __restrict__ damp,
                                                  uint32_t t, float inv_dt) {
    const uint4 idx = reinterpret_cast<const uint4*>(tets + t)[0];
    float4 P[4];
    P[0] = along(x0, pred, idx.x, a);
    P[1] = along(x0, pred, idx.y, a);
    P[2] = along(x0, pred, idx.z, a);
    P[3] = along(x0, pred, idx.w, a);
    // local positions table: energy of the tet whose vertices are 0..3 of P
    TetRec r = tets[t];
    r.idx[0] = 0;
    r.idx[1] = 1;
    r.idx[2] = 2;
    r.idx[3] = 3;
    return nh_energy_dev<DAMP>(P, &r, damp, 0, inv_dt, t);
}

// Per scene and step length a: incremental potential at x0 + a * pred minus
// at x0 -- Neo-Hookean (+ damping), ground AL and inertia terms (the terms
// that change along the prediction).
template <bool DAMP>
__global__ __launch_bounds__(256) void warm_ip_kernel(DevState d, WarmIO io, float dt) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const float inv_dt = 1.f / dt;
    double e[kWarmAlphas] = {0.0, 0.0, 0.0};
    uint32_t scene = 0;
    bool active = false;
    if (k < d.n_tets) {
        const double e0 = nh_energy_along<DAMP>(d.x_scratch, io.pred, 0.f, d.tets, d.damp, k, inv_dt);
        for (int j = 0; j < kWarmAlphas; j++)
            e[j] = nh_energy_along<DAMP>(d.x_scratch, io.pred, kWarmSteps[j], d.tets, d.damp, k, inv_dt) - e0;
        scene = io.scene_of[d.tets[k].idx[0]];
        active = true;
    } else if (k < d.n_tets + d.n_ground) {
        const GroundRec c = d.ground[k - d.n_tets];
        const double e0 = ground_al_energy(c, d.x_scratch[c.p]);
        for (int j = 0; j < kWarmAlphas; j++)
            e[j] = ground_al_energy(c, along(d.x_scratch, io.pred, c.p, kWarmSteps[j])) - e0;
        scene = io.scene_of[c.p];
        active = true;
    } else if (k < d.n_tets + d.n_ground + d.n_items) {
        const uint32_t v = d.items[k - d.n_tets - d.n_ground].v;
        const float4 x0 = d.x_scratch[v], y = d.y[v];
        const double md = 0.5 * (double)x0.w * inv_dt * inv_dt;
        const double b = ((double)x0.x - y.x) * (x0.x - y.x) + ((double)x0.y - y.y) * (x0.y - y.y) +
                         ((double)x0.z - y.z) * (x0.z - y.z);
        for (int j = 0; j < kWarmAlphas; j++) {
            const float4 x = along(d.x_scratch, io.pred, v, kWarmSteps[j]);
            const double a = ((double)x.x - y.x) * (x.x - y.x) + ((double)x.y - y.y) * (x.y - y.y) +
                             ((double)x.z - y.z) * (x.z - y.z);
            e[j] = md * (a - b);
        }
        scene = io.scene_of[v];
        active = true;
    }
    add_per_scene_block<kWarmAlphas>(io.ipd, io.n_scenes, scene, e, active);
}

__global__ void warm_ip_reset_kernel(WarmIO io) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < io.n_scenes * kWarmAlphas) io.ipd[s] = 0.0;
}

// Per scene: the step length along the prediction with the lowest
// incremental potential, if it is below the initial guess's, else the initial
// guess; restart the Chebyshev history at the chosen start.
__device__ __forceinline__ float warm_choice(const WarmIO& io, uint32_t s) {
    double best = 0.0;
    float a = 0.f;
    for (int j = 0; j < kWarmAlphas; j++) {
        const double v = io.ipd[j * io.n_scenes + s];
        if (v < best) {
            best = v;
            a = kWarmSteps[j];
        }
    }
    return a;
}

__global__ void warm_choose_kernel(DevState d, WarmIO io) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < io.n_scenes) {
        const float a = warm_choice(io, i);
        atomicAdd(&io.counters[0], 1ull);
        if (a == 0.f) atomicAdd(&io.counters[1], 1ull);
        if (io.step_of) io.step_of[i] = a;
    }
    if (i >= d.n_items) return;
    const uint32_t v = d.items[i].v;
    const float4 x = along(d.x_scratch, io.pred, v, warm_choice(io, io.scene_of[v]));
    d.x[v] = x;
    d.x_last[v] = x;
    d.x_lastlast[v] = x;
}

// Training target of a recorded step: (x_end - x0) / (dt^2 g) per work item.
__global__ void warm_target_kernel(DevState d, WarmIO io, float inv_s) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.n_items) return;
    const uint32_t v = d.items[i].v;
    const float4 x = d.x[v], x0 = d.x_scratch[v];
    io.tgt[(size_t)i * 3 + 0] = (x.x - x0.x) * inv_s;
    io.tgt[(size_t)i * 3 + 1] = (x.y - x0.y) * inv_s;
    io.tgt[(size_t)i * 3 + 2] = (x.z - x0.z) * inv_s;
}

// ---- residual / incremental potential (new; see avbd_get_stats) ----------
// Per free vertex: r = m/dt^2 (x - y) + G(x); partial sums of |r|^2, |m g|^2,
// and the inertia term m/(2 dt^2)|x-y|^2 per block.
template <int L, bool DAMP>
__global__ __launch_bounds__(1024) void stats_items_kernel(const VItem* __restrict__ items, uint32_t n_items,
                                                           const uint32_t* __restrict__ inc,
                                                           const TetRec* __restrict__ tets,
                                                           const TetDamp* __restrict__ damp,
                                                           const float4* __restrict__ x,
                                                           const float4* __restrict__ y,
                                                           const GroundRec* __restrict__ ground,
                                                           double* __restrict__ out, float inv_dt, float inv_dt2,
                                                           float g_acc, DynArgs da) {
    __shared__ double red[3][1024 / L];
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t item = tid / L;
    const uint32_t lane = tid % L;
    const uint32_t grp = threadIdx.x / L;
    double r2 = 0, mg2 = 0, ip = 0;
    if (item < n_items) {
        const VItem it = items[item];
        const uint32_t cnt = it.inc_count & ~VITEM_CONFLICT;
        float3 g = make_float3(0.f, 0.f, 0.f);
        Sym3 h = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (uint32_t j = lane; j < cnt; j += L)
            nh_incidence<DAMP>(x, tets, damp, inc[it.inc_begin + j], inv_dt, g, h);
        g.x = group_sum<L>(g.x);
        g.y = group_sum<L>(g.y);
        g.z = group_sum<L>(g.z);
        if (lane == 0) {
            const float4 xv = x[it.v], yv = y[it.v];
            if (it.ground >= 0) ground_accumulate(ground[it.ground], f3(xv), g, h);
            dyn_add3(da, x, it.v, g, h);
            const double md = (double)xv.w * (double)inv_dt2;
            const double dx = (double)xv.x - yv.x, dy = (double)xv.y - yv.y, dz = (double)xv.z - yv.z;
            const double rx = md * dx + g.x, ry = md * dy + g.y, rz = md * dz + g.z;
            r2 = rx * rx + ry * ry + rz * rz;
            const double mg = (double)xv.w * g_acc;
            mg2 = mg * mg;
            ip = 0.5 * md * (dx * dx + dy * dy + dz * dz);
        }
    } else {
        (void)lane;
    }
    if (lane == 0) {
        red[0][grp] = r2;
        red[1][grp] = mg2;
        red[2][grp] = ip;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double s = 0;
        for (uint32_t k = 0; k < blockDim.x / L; k++) s += red[threadIdx.x][k];
        out[blockIdx.x * 4 + threadIdx.x] = s;
    }
}

// Energy of every tet (NeoHookeanEnergySolver::energy, NeoHookeanEnergySolver.hpp:20-53)
// and ground record (AL form 1/2 k C^2 + lambda C) summed per block into out[b*4+3].
template <bool DAMP>
__global__ void stats_energy_kernel(const TetRec* __restrict__ tets, const TetDamp* __restrict__ damp,
                                    uint32_t n_tets, const GroundRec* __restrict__ ground, uint32_t n_ground,
                                    const float4* __restrict__ x, double* __restrict__ out, float inv_dt,
                                    DevState d) {
    __shared__ double red[256];
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0;
    if (k < n_tets) {
        e = nh_energy_dev<DAMP>(x, tets, damp, k, inv_dt, k);
    } else if (k < n_tets + n_ground) {
        const GroundRec c = ground[k - n_tets];
        e = ground_al_energy(c, x[c.p]);
    } else if (k < n_tets + n_ground + d.n_cos) {
        // CosseratRodEnergySolver::energy (CosseratRodEnergySolver.hpp:20-49)
        const CosRec c = d.cos[k - n_tets - n_ground];
        e = rigid::cos_energy(c, rigid::xyz(x[c.i0]), rigid::xyz(x[c.i1]), d.q[d.rot_of[c.i0]], d.q[d.rot_of[c.i1]]);
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x * 4 + 3] = red[0];
}

// Residual / inertia term of the oriented vertices (positional block of the
// 6-DOF gradient), written to slots 0..2 of blocks [base_block, ...).
__global__ __launch_bounds__(256) void stats_oitems_kernel(DevState d, double* __restrict__ out,
                                                           uint32_t base_block, float inv_dt2, float g_acc) {
    __shared__ double red[3][256];
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    double r2 = 0, mg2 = 0, ip = 0;
    if (k < d.n_oitems) {
        const OItem it = d.oitems[k];
        const float4 xv = d.x[it.v];
        rot::Vec6 G;
        rot::Sym6 H;
        rot::vec6_zero(G);
        rot::sym6_zero(H);
        if (d.dyn_off) rigid::accumulate6<true>(d, it, rigid::xyz(xv), d.q[it.rot], G, H);
        else rigid::accumulate6<false>(d, it, rigid::xyz(xv), d.q[it.rot], G, H);
        const float4 yv = d.y[it.v];
        const double md = (double)xv.w * (double)inv_dt2;
        const double dx = (double)xv.x - yv.x, dy = (double)xv.y - yv.y, dz = (double)xv.z - yv.z;
        const double rx = md * dx + G.v[0], ry = md * dy + G.v[1], rz = md * dz + G.v[2];
        r2 = rx * rx + ry * ry + rz * rz;
        const double mg = (double)xv.w * g_acc;
        mg2 = mg * mg;
        ip = 0.5 * md * (dx * dx + dy * dy + dz * dz);
    }
    red[0][threadIdx.x] = r2;
    red[1][threadIdx.x] = mg2;
    red[2][threadIdx.x] = ip;
    __syncthreads();
    for (uint32_t st = blockDim.x / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            for (int i = 0; i < 3; i++) red[i][threadIdx.x] += red[i][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x < 3) out[(base_block + blockIdx.x) * 4 + threadIdx.x] = red[threadIdx.x][0];
}

inline uint32_t blocks_for(uint32_t n, uint32_t b) { return (n + b - 1) / b; }

template <int L, bool DAMP>
void solve_dispatch(hipStream_t s, const DevState& d, uint32_t block, uint32_t batch, uint32_t begin, uint32_t count,
                    float dt, float4* x_out, const uint32_t* done) {
    const uint32_t threads = count * L;
    const dim3 grid(blocks_for(threads, block)), blk(block);
#define AVBD_SOLVE_ARGS \
    d.items + begin, count, d.inc, d.irec, d.damp, d.x, d.y, d.ground, x_out, d.x_buf, 1.0f / dt, 1.0f / (dt * dt), \
        done, dyn_args(d), d.frozen, d.scene_of
    if (block > 256) batch = 1;   // the batched variants are built for blocks of <= 256 threads
    if (d.dyn_off) hipLaunchKernelGGL((solve_color_kernel<L, DAMP, true, 1>), grid, blk, 0, s, AVBD_SOLVE_ARGS);
    else if (batch >= 4) hipLaunchKernelGGL((solve_color_kernel<L, DAMP, false, 4>), grid, blk, 0, s, AVBD_SOLVE_ARGS);
    else if (batch >= 2) hipLaunchKernelGGL((solve_color_kernel<L, DAMP, false, 2>), grid, blk, 0, s, AVBD_SOLVE_ARGS);
    else hipLaunchKernelGGL((solve_color_kernel<L, DAMP, false, 1>), grid, blk, 0, s, AVBD_SOLVE_ARGS);
#undef AVBD_SOLVE_ARGS
}

template <int L, bool DAMP>
void residual_dispatch(hipStream_t s, const DevState& d, float dt, const ResCtl* rc, double* rr) {
    const uint32_t threads = max(d.n_items * L, d.n_oitems);
    if (threads == 0) return;
    hipLaunchKernelGGL((residual_pass_kernel<L, DAMP>), dim3(blocks_for(threads, 256)), dim3(256), 0, s, d, rc, rr,
                       1.0f / dt, 1.0f / (dt * dt));
}

template <int L, bool DAMP>
uint32_t stats_dispatch(hipStream_t s, const DevState& d, uint32_t block, float dt, float g) {
    const uint32_t threads = d.n_items * L;
    const uint32_t nb = blocks_for(threads, block);
    if (nb == 0) return 0;
    hipLaunchKernelGGL((stats_items_kernel<L, DAMP>), dim3(nb), dim3(block), 0, s, d.items, d.n_items, d.inc,
                       d.tets, d.damp, d.x, d.y, d.ground, d.reduce, 1.0f / dt, 1.0f / (dt * dt), g, dyn_args(d));
    return nb;
}

}  // namespace

void launch_eprev(hipStream_t s, const DevState& d, float) {
    if (d.n_tets == 0) return;
    hipLaunchKernelGGL(eprev_kernel, dim3(blocks_for(d.n_tets, 256)), dim3(256), 0, s, d.tets, d.damp, d.x_prev,
                       d.n_tets);
}

void launch_step_begin(hipStream_t s, const DevState& d, float dt, float g) {
    if (d.n == 0) return;
    hipLaunchKernelGGL(step_begin_kernel, dim3(blocks_for(d.n, 256)), dim3(256), 0, s, d, dt, g);
}

void launch_solve_ocolor(hipStream_t s, const DevState& d, uint32_t begin, uint32_t count, float dt,
                         const uint32_t* done) {
    if (count == 0) return;
    if (d.dyn_off)
        hipLaunchKernelGGL(solve_ocolor_kernel<true>, dim3(blocks_for(count, 256)), dim3(256), 0, s, d, begin, count,
                           1.0f / (dt * dt), done);
    else
        hipLaunchKernelGGL(solve_ocolor_kernel<false>, dim3(blocks_for(count, 256)), dim3(256), 0, s, d, begin,
                           count, 1.0f / (dt * dt), done);
}

void launch_residual_reset(hipStream_t s, ResCtl* rc, const SceneFreeze& fz, uint32_t n_scenes) {
    const uint32_t n = fz.frozen ? (n_scenes > 0 ? n_scenes : 1u) : 1u;
    hipLaunchKernelGGL(residual_reset_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, rc, fz, n_scenes);
}

void launch_residual_pass(hipStream_t s, const DevState& d, const LaunchCfg& cfg, bool damping, float dt,
                          const ResCtl* rc, double* rr) {
#define AVBD_RES_CASE(LL)                                                                                    \
    case LL:                                                                                                  \
        if (damping) residual_dispatch<LL, true>(s, d, dt, rc, rr);                                          \
        else residual_dispatch<LL, false>(s, d, dt, rc, rr);                                                 \
        break;
    switch (cfg.lanes) {
        AVBD_RES_CASE(4)
        AVBD_RES_CASE(8)
        AVBD_RES_CASE(16)
        AVBD_RES_CASE(32)
        default: break;
    }
#undef AVBD_RES_CASE
}

void launch_residual_decide(hipStream_t s, ResCtl* rc, const double* rr, const uint32_t* scene_off,
                            uint32_t n_scenes, double* acc_scene, const double* mg2_scene, double tol2,
                            uint32_t every, const SceneFreeze& fz) {
    hipLaunchKernelGGL(residual_scene_kernel, dim3(n_scenes), dim3(1024), 0, s, rc, rr, scene_off, acc_scene,
                       mg2_scene, tol2, every, fz);
    hipLaunchKernelGGL(residual_decide_kernel, dim3(1), dim3(256), 0, s, rc, acc_scene, mg2_scene, n_scenes, tol2,
                       every);
}

void launch_residual_finish(hipStream_t s, ResCtl* rc, const SceneFreeze& fz, uint32_t n_scenes, uint32_t n_items) {
    hipLaunchKernelGGL(residual_finish_kernel, dim3(1), dim3(256), 0, s, rc, fz, n_scenes, n_items);
}

void launch_dyn_apply(hipStream_t s, const DevState& d, const DynOp* ops, uint32_t n_ops) {
    if (n_ops == 0) return;
    hipLaunchKernelGGL(dyn_apply_kernel, dim3(blocks_for(n_ops, 256)), dim3(256), 0, s, d.dyn, ops, n_ops);
}

void launch_dyn_step_begin(hipStream_t s, const DevState& d, const uint32_t* active, uint32_t n_active) {
    if (n_active == 0) return;
    hipLaunchKernelGGL(dyn_step_begin_kernel, dim3(blocks_for(n_active, 256)), dim3(256), 0, s, d, active, n_active);
}

void launch_dyn_dual(hipStream_t s, const DevState& d, const uint32_t* active, uint32_t n_active,
                     const uint32_t* done) {
    if (n_active == 0) return;
    hipLaunchKernelGGL(dyn_dual_kernel, dim3(blocks_for(n_active, 256)), dim3(256), 0, s, d, active, n_active,
                       done);
}

void launch_joint2_step_begin(hipStream_t s, const DevState& d) {
    const uint32_t n2 = d.n_joints - d.n_joints1;
    if (n2) hipLaunchKernelGGL(joint2_step_begin_kernel, dim3(blocks_for(n2, 256)), dim3(256), 0, s, d);
}

void launch_joint2_dual(hipStream_t s, const DevState& d, const uint32_t* done) {
    const uint32_t n2 = d.n_joints - d.n_joints1;
    if (n2) hipLaunchKernelGGL(joint2_dual_kernel, dim3(blocks_for(n2, 256)), dim3(256), 0, s, d, done);
}

void launch_solve_color(hipStream_t s, const DevState& d, const LaunchCfg& cfg, uint32_t begin, uint32_t count,
                        bool damping, float dt, float4* x_out, const uint32_t* done) {
    if (count == 0) return;
#define AVBD_SOLVE_CASE(LL)                                                                \
    case LL:                                                                                \
        if (damping) solve_dispatch<LL, true>(s, d, cfg.block, cfg.batch, begin, count, dt, x_out, done); \
        else solve_dispatch<LL, false>(s, d, cfg.block, cfg.batch, begin, count, dt, x_out, done);      \
        break;
    switch (cfg.lanes) {
        AVBD_SOLVE_CASE(4)
        AVBD_SOLVE_CASE(8)
        AVBD_SOLVE_CASE(16)
        AVBD_SOLVE_CASE(32)
        default: break;
    }
#undef AVBD_SOLVE_CASE
}

void launch_solve_cheb(hipStream_t s, const DevState& d, const LaunchCfg& cfg, uint32_t begin, uint32_t count,
                       bool damping, float dt, float omega, const ResCtl* rc, const float* omega_tab, uint32_t j,
                       const uint32_t* cheb_list, uint32_t n_cheb, float4* x_next) {
    uint32_t block = cfg.block, batch = cfg.batch;
    if (block > 256) batch = 1;   // the batched variants are built for blocks of <= 256 threads
    const uint32_t L = cfg.lanes;
    const uint32_t sx = (blocks_for(count * L, block) + 7u) / 8u, cx = (blocks_for(n_cheb, block) + 7u) / 8u;
    if (sx + cx == 0) return;
    const dim3 grid(8u * (sx + cx)), blk(block);
    const float inv_dt = 1.0f / dt, inv_dt2 = 1.0f / (dt * dt);
#define AVBD_CHEB_LAUNCH(LL, DD, UU)                                                                           \
    hipLaunchKernelGGL((solve_cheb_kernel<LL, DD, UU>), grid, blk, 0, s, d, d.items + begin, count, sx, cx,     \
                       cheb_list, n_cheb, x_next, omega, rc, omega_tab, j, inv_dt, inv_dt2)
#define AVBD_CHEB_U(LL, DD)                          \
    if (batch >= 4) AVBD_CHEB_LAUNCH(LL, DD, 4);     \
    else if (batch >= 2) AVBD_CHEB_LAUNCH(LL, DD, 2); \
    else AVBD_CHEB_LAUNCH(LL, DD, 1);
#define AVBD_CHEB_CASE(LL)                     \
    case LL:                                    \
        if (damping) { AVBD_CHEB_U(LL, true) }  \
        else { AVBD_CHEB_U(LL, false) }         \
        break;
    switch (L) {
        AVBD_CHEB_CASE(4)
        AVBD_CHEB_CASE(8)
        AVBD_CHEB_CASE(16)
        AVBD_CHEB_CASE(32)
        default: break;
    }
#undef AVBD_CHEB_CASE
#undef AVBD_CHEB_U
#undef AVBD_CHEB_LAUNCH
}

void launch_settle(hipStream_t s, float4* x, const float4* x_alt, uint32_t n, const uint32_t* iters_word,
                   uint32_t iters) {
    if (n == 0) return;
    hipLaunchKernelGGL(settle_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, x, x_alt, n, iters_word, iters);
}

void launch_commit(hipStream_t s, const DevState& d, const uint32_t* list, uint32_t count, const uint32_t* done) {
    if (count == 0) return;
    hipLaunchKernelGGL(commit_kernel, dim3(blocks_for(count, 256)), dim3(256), 0, s, d, list, count, done);
}

void launch_cheb_dual(hipStream_t s, const DevState& d, float omega, const ResCtl* rc, const float* omega_tab,
                      uint32_t j) {
    if (d.n == 0) return;
    hipLaunchKernelGGL(cheb_dual_kernel, dim3(blocks_for(d.n, 256)), dim3(256), 0, s, d, omega, rc, omega_tab, j);
}

void launch_velocity(hipStream_t s, const DevState& d, float dt) {
    if (d.n == 0) return;
    hipLaunchKernelGGL(velocity_kernel, dim3(blocks_for(d.n, 256)), dim3(256), 0, s, d, 1.0f / dt);
}

void launch_warm_save(hipStream_t s, const DevState& d, const WarmIO& io) {
    if (d.n) hipLaunchKernelGGL(warm_save_kernel, dim3(blocks_for(d.n, 256)), dim3(256), 0, s, d, io);
}

void launch_warm_features(hipStream_t s, const DevState& d, const WarmIO& io, uint32_t lanes, float dt, float g,
                          uint32_t rows_pad) {
    if (rows_pad == 0) return;
#define AVBD_FEAT_CASE(LL)                                                                                  \
    case LL:                                                                                                \
        hipLaunchKernelGGL((warm_features_kernel<LL>), dim3(blocks_for(rows_pad * LL, 256)), dim3(256), 0, s, d, \
                           io, dt, g, rows_pad);                                                            \
        break;
    switch (lanes) {
        AVBD_FEAT_CASE(4)
        AVBD_FEAT_CASE(8)
        AVBD_FEAT_CASE(16)
        AVBD_FEAT_CASE(32)
        default: break;
    }
#undef AVBD_FEAT_CASE
}

void launch_warm_apply(hipStream_t s, const DevState& d, const WarmIO& io, float dt2g) {
    if (d.n_items) hipLaunchKernelGGL(warm_apply_kernel, dim3(blocks_for(d.n_items, 256)), dim3(256), 0, s, d, io, dt2g);
}

void launch_warm_ip(hipStream_t s, const DevState& d, const WarmIO& io, bool damping, float dt) {
    hipLaunchKernelGGL(warm_ip_reset_kernel, dim3(blocks_for(io.n_scenes * kWarmAlphas, 256)), dim3(256), 0, s, io);
    const uint32_t n = d.n_tets + d.n_ground + d.n_items;
    if (n == 0) return;
    if (damping)
        hipLaunchKernelGGL(warm_ip_kernel<true>, dim3(blocks_for(n, 256)), dim3(256), 0, s, d, io, dt);
    else
        hipLaunchKernelGGL(warm_ip_kernel<false>, dim3(blocks_for(n, 256)), dim3(256), 0, s, d, io, dt);
}

void launch_warm_choose(hipStream_t s, const DevState& d, const WarmIO& io) {
    const uint32_t n = d.n_items > io.n_scenes ? d.n_items : io.n_scenes;
    if (n) hipLaunchKernelGGL(warm_choose_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, d, io);
}

void launch_warm_target(hipStream_t s, const DevState& d, const WarmIO& io, float inv_dt2g) {
    if (d.n_items)
        hipLaunchKernelGGL(warm_target_kernel, dim3(blocks_for(d.n_items, 256)), dim3(256), 0, s, d, io, inv_dt2g);
}

uint32_t launch_stats(hipStream_t s, const DevState& d, const LaunchCfg& cfg, bool damping, float dt, float g) {
    // the items pass writes slots 0..2 of its blocks; the energy pass slot 3 of its
    // blocks; both grids are sized so the larger one is returned and the host
    // zero-fills the rest before launching (see the caller).
    uint32_t nb_items = 0;
#define AVBD_STATS_CASE(LL)                                                                   \
    case LL:                                                                                   \
        nb_items = damping ? stats_dispatch<LL, true>(s, d, cfg.block, dt, g)                  \
                           : stats_dispatch<LL, false>(s, d, cfg.block, dt, g);                \
        break;
    switch (cfg.lanes) {
        AVBD_STATS_CASE(4)
        AVBD_STATS_CASE(8)
        AVBD_STATS_CASE(16)
        AVBD_STATS_CASE(32)
        default: break;
    }
#undef AVBD_STATS_CASE
    const uint32_t ne = d.n_tets + d.n_ground + d.n_cos;
    const uint32_t nb_e = blocks_for(ne, 256);
    if (ne) {
        if (damping)
            hipLaunchKernelGGL(stats_energy_kernel<true>, dim3(nb_e), dim3(256), 0, s, d.tets, d.damp, d.n_tets,
                               d.ground, d.n_ground, d.x, d.reduce, 1.0f / dt, d);
        else
            hipLaunchKernelGGL(stats_energy_kernel<false>, dim3(nb_e), dim3(256), 0, s, d.tets, d.damp, d.n_tets,
                               d.ground, d.n_ground, d.x, d.reduce, 1.0f / dt, d);
    }
    const uint32_t nb_o = blocks_for(d.n_oitems, 256);
    if (nb_o)
        hipLaunchKernelGGL(stats_oitems_kernel, dim3(nb_o), dim3(256), 0, s, d, d.reduce, nb_items, 1.0f / (dt * dt),
                           g);
    return nb_items + nb_o > nb_e ? nb_items + nb_o : nb_e;
}

}  // namespace avbd

namespace avbd {

namespace {
// Plain device copy with four independent 16-byte loads in flight per thread
// per trip (all issued before the stores), then a one-load tail.
__global__ void calib_copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const float4 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
        dst[i] = a;
        dst[i + stride] = b;
        dst[i + 2 * stride] = c;
        dst[i + 3 * stride] = d;
    }
    for (; i < n; i += stride) dst[i] = src[i];
}
}  // namespace

void launch_calib_copy(hipStream_t s, const float4* src, float4* dst, uint64_t n_float4) {
    hipLaunchKernelGGL(calib_copy_kernel, dim3(2048), dim3(256), 0, s, src, dst, n_float4);
}

}  // namespace avbd

