// contact_kernels.hip -- collision primitives beyond triangles and the
// narrow phase of the dynamic contacts, in FP64 from the FP32 state
// (compiled with -ffp-contract=off, like the LBVH).
//
//   prim_bounds_kernel   CollisionPrimitivePool::speculativeGlobalBounds
//                        (collision/CollisionPrimitivePool.hpp:160-236) for
//                        triangles, rod segments and spheres
//   narrow_kernel        CollisionDetector::_processPotentialCollision with
//                        _shouldSkip, _triangleSphere / _triangleSDF_CCD and
//                        _triangleRod (collision/CollisionDetector.cpp:132-215,
//                        504-666, 689-876): one thread per candidate pair,
//                        detected contacts appended to a list the host sorts
//                        by key
#include "contacts.hpp"

namespace avbd {

namespace {

struct d3 {
    double x, y, z;
};
__device__ __forceinline__ d3 mk(double x, double y, double z) { return d3{x, y, z}; }
__device__ __forceinline__ d3 operator+(d3 a, d3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ d3 operator-(d3 a, d3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ d3 operator*(double s, d3 a) { return mk(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ d3 cross(d3 a, d3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double norm(d3 a) { return sqrt(dot(a, a)); }
__device__ __forceinline__ double comp(d3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
__device__ __forceinline__ d3 ld(float4 v) { return mk((double)v.x, (double)v.y, (double)v.z); }

// Hamilton quaternion, scalar first in this file
struct dq {
    double w, x, y, z;
};
__device__ __forceinline__ dq ldq(float4 v) { return dq{(double)v.w, (double)v.x, (double)v.y, (double)v.z}; }
__device__ __forceinline__ dq qmul(dq a, dq b) {
    return dq{a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
              a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x, a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
}
__device__ __forceinline__ dq qconj(dq q) { return dq{q.w, -q.x, -q.y, -q.z}; }
// Eigen Quaternion::inverse: conjugate / squared norm
__device__ __forceinline__ dq qinv(dq q) {
    const double n2 = q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z;
    return dq{q.w / n2, -q.x / n2, -q.y / n2, -q.z / n2};
}
// q * v (Eigen's _transformVector)
__device__ __forceinline__ d3 qrot(dq q, d3 v) {
    const d3 u = mk(q.x, q.y, q.z);
    d3 uv = cross(u, v);
    uv = uv + uv;
    return v + q.w * uv + cross(u, uv);
}
// Math::Exp_s3 / Log_S3 (common/Math.hpp:197-235)
__device__ __forceinline__ dq exp_s3(d3 v) {
    const double th = norm(v);
    if (th < 1e-6) {
        const dq q{1.0, 0.5 * v.x, 0.5 * v.y, 0.5 * v.z};
        const double nq = sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
        return dq{q.w / nq, q.x / nq, q.y / nq, q.z / nq};
    }
    const double h = 0.5 * th;
    const double s = sin(h) / th;
    return dq{cos(h), s * v.x, s * v.y, s * v.z};
}
__device__ __forceinline__ d3 log_s3(dq q) {
    const d3 v = mk(q.x, q.y, q.z);
    const double vm = norm(v);
    if (vm < 1e-6) return 2.0 * v;
    const double th = 2 * atan2(vm, q.w);
    return (th / vm) * v;
}
__device__ __forceinline__ dq plus_s3(dq q, d3 v) { return qmul(q, exp_s3(v)); }
__device__ __forceinline__ d3 minus_s3(dq q1, dq q2) { return log_s3(qmul(qconj(q2), q1)); }

// Quaternion::toRotationMatrix, row i dotted with |.| and a half extent r
__device__ __forceinline__ void abs_rot_extent(dq q, double r, double e[3]) {
    const double tx = 2 * q.x, ty = 2 * q.y, tz = 2 * q.z;
    const double twx = tx * q.w, twy = ty * q.w, twz = tz * q.w;
    const double txx = tx * q.x, txy = ty * q.x, txz = tz * q.x;
    const double tyy = ty * q.y, tyz = tz * q.y, tzz = tz * q.z;
    const double R[3][3] = {{1 - (tyy + tzz), txy - twz, txz + twy},
                            {txy + twz, 1 - (txx + tzz), tyz - twx},
                            {txz - twy, tyz + twx, 1 - (txx + tyy)}};
    for (int i = 0; i < 3; i++) e[i] = fabs(R[i][0]) * r + fabs(R[i][1]) * r + fabs(R[i][2]) * r;
}

__global__ void prim_bounds_kernel(DevPrims p, const float4* __restrict__ x, const float4* __restrict__ v,
                                   const float4* __restrict__ q, const float4* __restrict__ w,
                                   const int32_t* __restrict__ rot_of, double dt, double* __restrict__ aabb) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= p.n) return;
    double b[6] = {__DBL_MAX__, __DBL_MAX__, __DBL_MAX__, -__DBL_MAX__, -__DBL_MAX__, -__DBL_MAX__};
    auto expand = [&](const double lo[3], const double hi[3]) {
        for (int i = 0; i < 3; i++) {
            b[i] = fmin(b[i], lo[i]);
            b[3 + i] = fmax(b[3 + i], hi[i]);
        }
    };
    const uint32_t ty = p.type[k];
    const uint32_t* id = p.idx + 3 * k;
    if (ty == PRIM_TRIANGLE) {
        for (int c = 0; c < 3; c++) {
            const d3 a = ld(x[id[c]]);
            const d3 e = a + dt * ld(v[id[c]]);
            const double pa[3] = {a.x, a.y, a.z}, pe[3] = {e.x, e.y, e.z};
            expand(pa, pa);
            expand(pe, pe);
        }
        // self collision (new): widened by the contact margin
        for (int i = 0; i < 3; i++) {
            b[i] -= p.tri_pad;
            b[3 + i] += p.tri_pad;
        }
    } else if (ty == PRIM_ROD_SEGMENT) {
        const double r = p.radius[k];
        const d3 p1 = ld(x[id[0]]), p2 = ld(x[id[1]]);
        const d3 n1 = p1 + dt * ld(v[id[0]]), n2 = p2 + dt * ld(v[id[1]]);
        double lo[3], hi[3];
        for (int i = 0; i < 3; i++) {
            lo[i] = fmin(comp(p1, i), comp(p2, i)) - r;
            hi[i] = fmax(comp(p1, i), comp(p2, i)) + r;
        }
        expand(lo, hi);
        for (int i = 0; i < 3; i++) {
            lo[i] = fmin(comp(n1, i), comp(n2, i)) - r;
            hi[i] = fmax(comp(n1, i), comp(n2, i)) + r;
        }
        expand(lo, hi);
    } else {
        // sphere: local box [-r, r]^3 rotated, at the current and the predicted pose
        const double r = p.radius[k];
        const uint32_t c = id[0];
        const dq qc = ldq(q[rot_of[c]]);
        const d3 pc = ld(x[c]);
        double e[3], lo[3], hi[3];
        abs_rot_extent(qc, r, e);
        for (int i = 0; i < 3; i++) {
            lo[i] = comp(pc, i) - e[i];
            hi[i] = comp(pc, i) + e[i];
        }
        expand(lo, hi);
        const dq qn = plus_s3(qc, dt * ld(w[rot_of[c]]));
        const d3 pn = pc + dt * ld(v[c]);
        abs_rot_extent(qn, r, e);
        for (int i = 0; i < 3; i++) {
            lo[i] = comp(pn, i) - e[i];
            hi[i] = comp(pn, i) + e[i];
        }
        expand(lo, hi);
    }
    for (int i = 0; i < 6; i++) aabb[6 * k + i] = b[i];
}

// ---- geometry (common/Math.hpp:421-681) ----------------------------------------
__device__ __forceinline__ d3 barycentric(d3 p, d3 a, d3 b, d3 c) {
    const d3 v0 = b - a, v1 = c - a, v2 = p - a;
    const double d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1), d20 = dot(v2, v0), d21 = dot(v2, v1);
    const double den = d00 * d11 - d01 * d01;
    const double vv = (d11 * d20 - d01 * d21) / den;
    const double ww = (d00 * d21 - d01 * d20) / den;
    return mk(1 - vv - ww, vv, ww);
}

__device__ __forceinline__ d3 closest_point_triangle(d3 p, d3 a, d3 b, d3 c) {
    const d3 ab = b - a, ac = c - a, ap = p - a;
    const double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) return a;
    const d3 bp = p - b;
    const double d3_ = dot(ab, bp), d4 = dot(ac, bp);
    if (d3_ >= 0 && d4 <= d3_) return b;
    const d3 cp = p - c;
    const double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) return c;
    const double vc = d1 * d4 - d3_ * d2;
    if (vc <= 0 && d1 >= 0 && d3_ <= 0) return a + (d1 / (d1 - d3_)) * ab;
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) return a + (d2 / (d2 - d6)) * ac;
    const double va = d3_ * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3_) >= 0 && (d5 - d6) >= 0) return b + ((d4 - d3_) / ((d4 - d3_) + (d5 - d6))) * (c - b);
    const double den = 1 / (va + vb + vc);
    return a + (vb * den) * ab + (vc * den) * ac;
}

__device__ __forceinline__ double clamp01(double v) { return fmin(fmax(v, 0.0), 1.0); }

__device__ __forceinline__ void closest_segment_segment(d3 p1, d3 q1, d3 p2, d3 q2, double& s, double& t) {
    const d3 d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
    const double a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    const double EPS = 1e-8;
    if (a <= EPS && e <= EPS) {
        s = t = 0;
        return;
    }
    if (a <= EPS) {
        s = 0;
        t = clamp01(f / e);
        return;
    }
    const double c = dot(d1, r);
    if (e <= EPS) {
        t = 0;
        s = clamp01(-c / a);
        return;
    }
    const double b = dot(d1, d2);
    const double den = a * e - b * b;
    s = den != 0 ? clamp01((b * f - c * e) / den) : 0;
    t = (b * s + f) / e;
    if (t < 0) {
        t = 0;
        s = clamp01(-c / a);
    } else if (t > 1) {
        t = 1;
        s = clamp01((b - c) / a);
    }
}

__device__ __forceinline__ bool point_in_triangle(d3 p, d3 a, d3 b, d3 c) {
    const d3 v0 = b - a, v1 = c - a, v2 = p - a;
    const double d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1), d20 = dot(v2, v0), d21 = dot(v2, v1);
    const double den = d00 * d11 - d01 * d01;
    if (fabs(den) < 1e-12) return false;
    const double vv = (d11 * d20 - d01 * d21) / den;
    const double ww = (d00 * d21 - d01 * d20) / den;
    const double uu = 1.0 - vv - ww;
    const double EPS = 1e-8;
    return uu >= -EPS && vv >= -EPS && ww >= -EPS;
}

// Math::closestPoint_SegmentTriangle; the edge test measures to the edge
// point at the segment's parameter s, as the reference does.
__device__ void closest_segment_triangle(d3 p0, d3 p1, d3 a, d3 b, d3 c, double& s_out, double bc_out[3],
                                         double& dist) {
    const double EPS = 1e-8;
    const d3 dir = p1 - p0, ab = b - a, ac = c - a;
    const d3 nrm = cross(ab, ac);
    s_out = 0;
    bc_out[0] = 1;
    bc_out[1] = bc_out[2] = 0;
    dist = __DBL_MAX__;
    if (dot(nrm, nrm) > EPS) {
        const double den = dot(nrm, dir);
        if (fabs(den) > EPS) {
            const double s = dot(nrm, a - p0) / den;
            if (s >= 0.0 && s <= 1.0) {
                const d3 p = p0 + s * dir;
                if (point_in_triangle(p, a, b, c)) {
                    s_out = s;
                    const d3 w = barycentric(p, a, b, c);
                    bc_out[0] = w.x;
                    bc_out[1] = w.y;
                    bc_out[2] = w.z;
                    dist = 0;
                    return;
                }
            }
        }
    }
    const d3 E[3][2] = {{a, b}, {a, c}, {b, c}};
    const int I[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int k = 0; k < 3; k++) {
        double s, t;
        closest_segment_segment(p0, p1, E[k][0], E[k][1], s, t);
        const d3 cp = (1 - s) * p0 + s * p1;
        const d3 cq = (1 - s) * E[k][0] + s * E[k][1];
        const d3 dd = cp - cq;
        const double d2 = dot(dd, dd);
        if (d2 < dist * dist) {
            s_out = s;
            bc_out[0] = bc_out[1] = bc_out[2] = 0;
            bc_out[I[k][0]] = 1 - t;
            bc_out[I[k][1]] = t;
            dist = sqrt(d2);
        }
    }
    for (int k = 0; k < 2; k++) {
        const d3 pe = k == 0 ? p0 : p1;
        const d3 cpe = closest_point_triangle(pe, a, b, c);
        const d3 e = pe - cpe;
        const double sq = dot(e, e);
        if (sq < dist * dist) {
            s_out = k == 0 ? 0.0 : 1.0;
            const d3 w = barycentric(cpe, a, b, c);
            bc_out[0] = w.x;
            bc_out[1] = w.y;
            bc_out[2] = w.z;
            dist = sqrt(sq);
        }
    }
}

// Algorithm::goldenSectionSearch (common/Algorithm.hpp:253-299), tol 1e-3
__device__ __forceinline__ double gwidth(double v) { return fabs(v); }
__device__ __forceinline__ double gwidth(d3 v) { return norm(v); }
__device__ __forceinline__ double glerp(double a, double b, double u) { return (1 - u) * a + u * b; }
__device__ __forceinline__ d3 glerp(d3 a, d3 b, double u) { return (1 - u) * a + u * b; }

template <class T, class F>
__device__ T golden(T ts, T te, F&& f) {
    const double tol = 1e-3;
    const double phi = 1.61803398875;
    const double r = 1 / phi;
    const double s = 1 - r;
    double a0 = 0, a1 = s, a2 = r, a3 = 1;
    T t0 = ts, t1 = glerp(ts, te, a1), t2 = glerp(ts, te, a2), t3 = te;
    double f0 = f(t0), f1 = f(t1), f2 = f(t2), f3 = f(t3);
    while (gwidth(t3 - t0) > tol) {
        if (fmin(f0, f1) < fmin(f2, f3)) {
            a3 = a2; t3 = t2; f3 = f2;
            a2 = a1; t2 = t1; f2 = f1;
            a1 = r * a2 + s * a0;
            t1 = glerp(ts, te, a1);
            f1 = f(t1);
        } else {
            a0 = a1; t0 = t1; f0 = f1;
            a1 = a2; t1 = t2; f1 = f2;
            a2 = r * a1 + s * a3;
            t2 = glerp(ts, te, a2);
            f2 = f(t2);
        }
    }
    const T tm = glerp(ts, te, 0.5 * (a0 + a3));
    const double fm = f(tm);
    if (f0 < fm && f0 < f3) return t0;
    if (fm < f3) return tm;
    return t3;
}

// CollisionDetector::_triangleSDF_CCD for a sphere SDF (|p| - r).
// The reference evaluates every trial point in the body frame at the trial
// time, R(t)^T (x - c(t)) with R(t) = R0 exp(t w).  For a sphere the SDF, its
// gradient, the normal velocity <v_rel, grad>, the barycentric coordinates
// and every comparison of the search are invariant under that rotation, so
// the search here runs on world-frame offsets x - c(t): the same search up
// to rounding, without a quaternion exponential per trial point.  The world
// angular velocity R(t) w equals R0 w (exp(t w) fixes w).  R(t) itself is
// formed only for the body-frame contact point and for the gradient at the
// centre (the reference's body-frame (1, 0, 0)).
__device__ bool tri_sphere_ccd(const d3 tri[3], const d3 tv[3], d3 rb_pos, dq rb_rot, d3 rb_v, d3 rb_w, double radius,
                               double dt, double margin, d3& normal, double bc_out[3], d3& cpl) {
    auto pos_t = [&](double t) { return rb_pos + t * rb_v; };
    auto rot_t = [&](double t) { return qmul(rb_rot, exp_s3(t * rb_w)); };
    const d3 wg = qrot(rb_rot, rb_w);
    auto rel_vel = [&](d3 xt, d3 vt, double t) { return vt - (rb_v + cross(wg, xt - pos_t(t))); };
    auto local_pos = [&](d3 xt, double t) { return xt - pos_t(t); };
    auto sdf = [&](d3 p) { return norm(p) - radius; };
    auto grad = [&](d3 p, double t) {
        const double dd = norm(p);
        return dd > 1e-8 ? (1.0 / dd) * p : qrot(rot_t(t), mk(1, 0, 0));
    };
    auto eval0 = [&](d3 si, d3 vi) { return dot(rel_vel(si, vi, 0), grad(local_pos(si, 0), 0)); };
    const double e1 = eval0(tri[0], tv[0]), e2 = eval0(tri[1], tv[1]), e3 = eval0(tri[2], tv[2]);
    double t_start = 0, t_end = dt, t = 0, u, v, w;
    if (e1 < e2 && e1 < e3) {
        u = 1; v = 0; w = 0;
    } else if (e2 < e1 && e2 < e3) {
        u = 0; v = 1; w = 0;
    } else {
        u = 0; v = 0; w = 1;
    }
    auto interp = [&](double tt) { return u * (tri[0] + tt * tv[0]) + v * (tri[1] + tt * tv[1]) + w * (tri[2] + tt * tv[2]); };
    auto interp_vel = [&]() { return u * tv[0] + v * tv[1] + w * tv[2]; };
    auto sd_time = [&](double tt) { return sdf(local_pos(interp(tt), tt)); };
    auto ud_time = [&](double tt) { return fabs(sd_time(tt)); };
    auto sd_point = [&](d3 p) { return sdf(p); };
    double sdf_x = 0;
    d3 grad_x = mk(1, 0, 0), x_ti = mk(0, 0, 0);
    for (int iter = 0; iter < 16; iter++) {
        const double t_old = t;
        d3 xg = interp(t);
        const d3 vg = interp_vel();
        x_ti = local_pos(xg, t);
        const d3 v_ti = rel_vel(xg, vg, t);
        sdf_x = sdf(x_ti);
        if (sdf_x <= 0) {
            t_end = fmin(t, t_end);
            t = golden(t_start, t, ud_time);
        } else {
            const double vdg = dot(v_ti, grad(x_ti, t));
            if (vdg > 0) t = golden(t_start, t, sd_time);
            else t = golden(t, t_end, sd_time);
        }
        xg = interp(t);
        x_ti = local_pos(xg, t);
        sdf_x = sdf(x_ti);
        grad_x = grad(x_ti, t);
        if (sdf_x <= 0) t_end = fmin(t, t_end);
        const d3 l1 = local_pos(tri[0] + t * tv[0], t);
        const d3 l2 = local_pos(tri[1] + t * tv[1], t);
        const d3 l3 = local_pos(tri[2] + t * tv[2], t);
        const double s1 = dot(l1, grad_x), s2 = dot(l2, grad_x), s3 = dot(l3, grad_x);
        d3 sp;
        if (s1 < s2 && s1 < s3) sp = l1;
        else if (s2 < s1 && s2 < s3) sp = l2;
        else sp = l3;
        const d3 x_new = golden(x_ti, sp, sd_point);
        const d3 bc = barycentric(x_new, l1, l2, l3);
        u = bc.x;
        v = bc.y;
        w = bc.z;
        if (fabs(t_old - t) < 1e-6 && norm(x_new - x_ti) < 1e-6) break;
    }
    bc_out[0] = u;
    bc_out[1] = v;
    bc_out[2] = w;
    // body-frame contact point and world normal, as the reference reports them
    cpl = qrot(qinv(rot_t(t)), x_ti - sdf_x * grad_x);
    normal = grad_x;
    return sdf_x < margin;
}

// collision/DetectedCollision.hpp:43-66
__device__ __forceinline__ uint64_t contact_key(uint32_t type, uint32_t a, uint32_t b, uint32_t code = 0) {
    if (a > b) {
        const uint32_t t = a;
        a = b;
        b = t;
    }
    const uint32_t M = (1u << 26) - 1;
    return (((((uint64_t)type & 31u) << 26 | (a & M)) << 26 | (b & M)) << 7) | (code & 127u);
}


__global__ void narrow_kernel(DevPrims p, const uint32_t* __restrict__ pairs, const uint32_t* __restrict__ counters,
                              uint32_t pair_cap, const uint32_t* __restrict__ sorted, const float4* __restrict__ x,
                              const float4* __restrict__ v, const float4* __restrict__ q,
                              const float4* __restrict__ w, const int32_t* __restrict__ rot_of, double dt,
                              double margin, DetRec* __restrict__ det, uint32_t det_cap, uint32_t* __restrict__ status) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n_pairs = min(counters[0], pair_cap);
    if (k >= n_pairs) return;
    uint32_t a = sorted[pairs[2 * k]], b = sorted[pairs[2 * k + 1]];
    // _shouldSkip: same object and a shared particle
    if (p.object[a] == p.object[b]) {
        const uint32_t na = p.type[a] == PRIM_TRIANGLE ? 3u : (p.type[a] == PRIM_ROD_SEGMENT ? 2u : 1u);
        const uint32_t nb = p.type[b] == PRIM_TRIANGLE ? 3u : (p.type[b] == PRIM_ROD_SEGMENT ? 2u : 1u);
        for (uint32_t i = 0; i < na; i++)
            for (uint32_t j = 0; j < nb; j++)
                if (p.idx[3 * a + i] == p.idx[3 * b + j]) return;
    }
    uint32_t ta = p.type[a], tb = p.type[b];
    if (ta > tb) {
        const uint32_t t = ta;
        ta = tb;
        tb = t;
        const uint32_t s = a;
        a = b;
        b = s;
    }
    if (ta == PRIM_TRIANGLE && tb == PRIM_TRIANGLE) {
        // _triangleTriangle: empty in the reference; with self collision on
        // (new), each vertex of either triangle against the other one.  A
        // vertex is tested from its owner (the lowest-numbered triangle
        // holding it) only, so a vertex-face pair is reported once; key =
        // generateKey(TriangleTriangle_VertexFace, owner, face, code), code =
        // the vertex's corner in the owner, + 3 when the owner is the
        // higher-numbered primitive.  Normal: the face normal turned towards
        // the face as seen from the vertex.
        if (!p.vertex_face) return;
        for (uint32_t side = 0; side < 2; side++) {
            const uint32_t own = side == 0 ? a : b, face = side == 0 ? b : a;
            const uint32_t* fi = p.idx + 3 * face;
            const d3 fa = ld(x[fi[0]]), fb = ld(x[fi[1]]), fc = ld(x[fi[2]]);
            for (uint32_t k = 0; k < 3; k++) {
                const uint32_t pv = p.idx[3 * own + k];
                if (p.owner[pv] != own || pv == fi[0] || pv == fi[1] || pv == fi[2]) continue;
                const d3 pp = ld(x[pv]);
                const d3 cp = closest_point_triangle(pp, fa, fb, fc);
                if (!(norm(pp - cp) < margin)) continue;
                const d3 f = cross(fb - fa, fc - fa);
                const double fm = norm(f);
                const d3 fu = fm > 0 ? (1.0 / fm) * f : mk(0, 1, 0);
                const d3 nrm = dot(fu, pp - cp) >= 0 ? (-1.0) * fu : fu;
                const d3 bc = barycentric(cp, fa, fb, fc);
                DetRec r;
                r.prim_a = face;
                r.prim_b = own;
                r.pad = k;
                r.kind = DET_KIND_VERTEX_FACE;
                r.n[0] = nrm.x; r.n[1] = nrm.y; r.n[2] = nrm.z;
                r.bary[0] = bc.x; r.bary[1] = bc.y; r.bary[2] = bc.z;
                r.cpl[0] = 0; r.cpl[1] = 0; r.cpl[2] = 0;
                r.s = 0;
                r.key = contact_key(r.kind, own, face, own < face ? k : 3 + k);
                const uint32_t slot = atomicAdd(status + 1, 1u);
                if (slot < det_cap) det[slot] = r;
            }
        }
        return;
    }
    if (ta == PRIM_ROD_SEGMENT && tb == PRIM_ROD_SEGMENT) return;        // _rodRod: no-op
    if (ta == PRIM_SPHERE) {
        // sphere-sphere (new; the reference throws): in contact when the
        // centres are closer than r_a + r_b + margin; n from b to a, contact
        // points on the facing surfaces in body frames; the lower primitive
        // index is the first body
        if (a > b) {
            const uint32_t s = a;
            a = b;
            b = s;
        }
        const uint32_t ca = p.idx[3 * a], cb = p.idx[3 * b];
        const double ra = p.radius[a], rb = p.radius[b];
        const d3 d = ld(x[ca]) - ld(x[cb]);
        const double dist = norm(d);
        if (!(dist < ra + rb + margin)) return;
        const d3 nrm = dist > 1e-12 ? (1.0 / dist) * d : mk(1, 0, 0);
        const d3 cpa = qrot(qinv(ldq(q[rot_of[ca]])), (-ra) * nrm);
        const d3 cpb = qrot(qinv(ldq(q[rot_of[cb]])), rb * nrm);
        DetRec r;
        r.prim_a = a;
        r.prim_b = b;
        r.pad = 0;
        r.kind = DET_KIND_RIGID_RIGID;
        r.n[0] = nrm.x; r.n[1] = nrm.y; r.n[2] = nrm.z;
        r.bary[0] = cpb.x; r.bary[1] = cpb.y; r.bary[2] = cpb.z;
        r.cpl[0] = cpa.x; r.cpl[1] = cpa.y; r.cpl[2] = cpa.z;
        r.s = 0;
        r.key = contact_key(r.kind, a, b);
        const uint32_t slot = atomicAdd(status + 1, 1u);
        if (slot < det_cap) det[slot] = r;
        return;
    }
    if (ta != PRIM_TRIANGLE) {
        // rod segment - sphere: a sphere welded to one of the segment's nodes
        // does not collide with it; any other pair is unsupported (the
        // reference throws)
        if (p.idx[3 * b] == p.idx[3 * a] || p.idx[3 * b] == p.idx[3 * a + 1]) return;
        atomicOr(status, NARROW_UNSUPPORTED);
        return;
    }
    const uint32_t* ti = p.idx + 3 * a;
    const d3 v1 = ld(x[ti[0]]), v2 = ld(x[ti[1]]), v3 = ld(x[ti[2]]);
    DetRec r;
    r.prim_a = a;
    r.prim_b = b;
    r.pad = 0;
    if (tb == PRIM_SPHERE) {
        const uint32_t c = p.idx[3 * b];
        const d3 T[3] = {v1, v2, v3};
        const d3 TV[3] = {ld(v[ti[0]]), ld(v[ti[1]]), ld(v[ti[2]])};
        // Every point the CCD evaluates is a point of the triangle at some
        // t in [0, dt] minus the sphere centre at t: with linear motion it
        // lies in the box of the vertices' offsets at t = 0 and t = dt.  When
        // that box is farther than r + margin from the centre (with 1e-6 m
        // to spare for rounding) the CCD cannot report a contact: skip it.
        {
            const d3 cc = ld(x[c]), cv = ld(v[c]);
            double lo[3] = {__DBL_MAX__, __DBL_MAX__, __DBL_MAX__}, hi[3] = {-__DBL_MAX__, -__DBL_MAX__, -__DBL_MAX__};
            for (int k = 0; k < 3; k++) {
                const d3 q0 = T[k] - cc, q1 = (T[k] + dt * TV[k]) - (cc + dt * cv);
                const double a0[3] = {q0.x, q0.y, q0.z}, a1[3] = {q1.x, q1.y, q1.z};
                for (int e = 0; e < 3; e++) {
                    lo[e] = fmin(lo[e], fmin(a0[e], a1[e]));
                    hi[e] = fmax(hi[e], fmax(a0[e], a1[e]));
                }
            }
            double d2 = 0;
            for (int e = 0; e < 3; e++) {
                const double g = fmax(0.0, fmax(lo[e], -hi[e]));
                d2 += g * g;
            }
            const double reach = p.radius[b] + margin + 1e-6;
            if (d2 > reach * reach) return;
        }
        d3 nrm, cpl;
        double bc[3];
        if (!tri_sphere_ccd(T, TV, ld(x[c]), ldq(q[rot_of[c]]), ld(v[c]), ld(w[rot_of[c]]), p.radius[b], dt, margin,
                            nrm, bc, cpl))
            return;
        r.kind = DET_KIND_TRI_RIGID;
        r.n[0] = nrm.x; r.n[1] = nrm.y; r.n[2] = nrm.z;
        r.bary[0] = bc[0]; r.bary[1] = bc[1]; r.bary[2] = bc[2];
        r.cpl[0] = cpl.x; r.cpl[1] = cpl.y; r.cpl[2] = cpl.z;
        r.s = 0;
    } else {
        const uint32_t n0 = p.idx[3 * b], n1 = p.idx[3 * b + 1];
        const d3 s1 = ld(x[n0]), s2 = ld(x[n1]);
        const double radius = p.radius[b];
        double cs, dist, bc[3];
        closest_segment_triangle(s1, s2, v1, v2, v3, cs, bc, dist);
        if (!(dist < radius + margin)) return;
        const d3 cp_line = (1 - cs) * s1 + cs * s2;
        const d3 cp_tri = bc[0] * v1 + bc[1] * v2 + bc[2] * v3;
        const d3 diff = cp_tri - cp_line;
        const double dm = norm(diff);
        const d3 nrm = dm > 1e-8 ? (1.0 / dm) * diff : mk(1, 0, 0);
        const dq q1 = ldq(q[rot_of[n0]]), q2 = ldq(q[rot_of[n1]]);
        const dq qm = plus_s3(q1, cs * minus_s3(q2, q1));
        const d3 cpl = qrot(qinv(qm), radius * nrm);
        r.kind = DET_KIND_TRI_ROD;
        r.n[0] = nrm.x; r.n[1] = nrm.y; r.n[2] = nrm.z;
        r.bary[0] = bc[0]; r.bary[1] = bc[1]; r.bary[2] = bc[2];
        r.cpl[0] = cpl.x; r.cpl[1] = cpl.y; r.cpl[2] = cpl.z;
        r.s = cs;
    }
    r.key = contact_key(r.kind, a, b);
    const uint32_t slot = atomicAdd(status + 1, 1u);
    if (slot < det_cap) det[slot] = r;
}

inline uint32_t nblk(uint32_t n, uint32_t b) { return (n + b - 1) / b; }

}  // namespace

void launch_prim_bounds(hipStream_t s, const DevPrims& p, const LbvhDev& b, const DevState& d, double dt) {
    if (p.n == 0) return;
    hipLaunchKernelGGL(prim_bounds_kernel, dim3(nblk(p.n, 256)), dim3(256), 0, s, p, d.x, d.v, d.q, d.w, d.rot_of,
                       dt, b.aabb);
}

void launch_narrow(hipStream_t s, const DevPrims& p, const LbvhDev& b, const DevState& d, double dt, double margin,
                   DetRec* det, uint32_t det_cap, uint32_t* status) {
    // a thread per reserved pair slot; the kernel reads the actual count
    hipLaunchKernelGGL(narrow_kernel, dim3(nblk(b.pair_cap, 256)), dim3(256), 0, s, p, b.pairs, b.counters,
                       b.pair_cap, b.sorted, d.x, d.v, d.q, d.w, d.rot_of, dt, margin, det, det_cap, status);
}

}  // namespace avbd
