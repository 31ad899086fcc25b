// rot_math.hpp -- FP32 device math of the oriented (6-DOF) particles:
// quaternions, the S3 exponential/logarithm, SO(3) exponential and its right
// Jacobians, small dense 3x3 helpers and an in-register 6x6 LDL^T solve.
//
// Quaternions are float4 (x, y, z, w) with w the scalar part (Hamilton
// product, the convention of Eigen::Quaternion used by the reference,
// common/Math.hpp:197-246).  Every map is written in a form that stays
// accurate in FP32 near zero angle: 1 - cos(n) is formed as 2 sin^2(n/2), and
// the coefficients that cancel catastrophically, (n - sin n)/n^3 and
// 1/n^2 - (1 + cos n)/(2 n sin n), switch to their Taylor series below
// n = 0.5 (next omitted term < 2e-7 relative there).
#pragma once

#include <hip/hip_runtime.h>

namespace avbd {
namespace rot {

struct M3f {
    float m[3][3];
};

__device__ __forceinline__ float3 v3(float x, float y, float z) { return make_float3(x, y, z); }
__device__ __forceinline__ float3 add(float3 a, float3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 sub(float3 a, float3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 scl(float s, float3 a) { return v3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ float dotv(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 crossv(float3 a, float3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float normv(float3 a) { return sqrtf(dotv(a, a)); }
__device__ __forceinline__ float comp(float3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// Hamilton product a * b.
__device__ __forceinline__ float4 qmul(float4 a, float4 b) {
    return make_float4(a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
                       a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z);
}
__device__ __forceinline__ float4 qconj(float4 q) { return make_float4(-q.x, -q.y, -q.z, q.w); }
__device__ __forceinline__ float4 qnormalize(float4 q) {
    const float s = rsqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    return make_float4(q.x * s, q.y * s, q.z * s, q.w * s);
}

// Rotate v by unit q: v + w t + u x t with t = 2 u x v.
__device__ __forceinline__ float3 qrot(float4 q, float3 v) {
    const float3 u = v3(q.x, q.y, q.z);
    const float3 t = scl(2.f, crossv(u, v));
    return add(add(v, scl(q.w, t)), crossv(u, t));
}

// Rotation matrix of a unit quaternion.
__device__ __forceinline__ M3f qmat(float4 q) {
    const float tx = 2.f * q.x, ty = 2.f * q.y, tz = 2.f * q.z;
    const float twx = tx * q.w, twy = ty * q.w, twz = tz * q.w;
    const float txx = tx * q.x, txy = ty * q.x, txz = tz * q.x;
    const float tyy = ty * q.y, tyz = tz * q.y, tzz = tz * q.z;
    M3f r;
    r.m[0][0] = 1.f - (tyy + tzz);
    r.m[0][1] = txy - twz;
    r.m[0][2] = txz + twy;
    r.m[1][0] = txy + twz;
    r.m[1][1] = 1.f - (txx + tzz);
    r.m[1][2] = tyz - twx;
    r.m[2][0] = txz - twy;
    r.m[2][1] = tyz + twx;
    r.m[2][2] = 1.f - (txx + tyy);
    return r;
}

// sin(h)/h, accurate for small h.
__device__ __forceinline__ float sinc(float h) {
    if (fabsf(h) < 1e-3f) return 1.f - h * h * (1.f / 6.f);
    return sinf(h) / h;
}

// Exp_s3(v) = (cos(|v|/2), sin(|v|/2) v/|v|).
__device__ __forceinline__ float4 exp_s3(float3 v) {
    const float th = normv(v);
    const float h = 0.5f * th;
    const float s = 0.5f * sinc(h);   // sin(h) / th
    return make_float4(s * v.x, s * v.y, s * v.z, cosf(h));
}

// Log_S3(q) = 2 atan2(|u|, w) u / |u|, u the vector part; 2u for |u| -> 0.
__device__ __forceinline__ float3 log_s3(float4 q) {
    const float3 u = v3(q.x, q.y, q.z);
    const float vm = normv(u);
    if (vm < 1e-6f) return scl(2.f, u);
    return scl(2.f * atan2f(vm, q.w) / vm, u);
}

__device__ __forceinline__ float4 plus_s3(float4 q, float3 v) { return qnormalize(qmul(q, exp_s3(v))); }
// q1 [-] q2 = Log(q2^-1 q1)
__device__ __forceinline__ float3 minus_s3(float4 q1, float4 q2) { return log_s3(qmul(qconj(q2), q1)); }

// A = a I + b [v]x + c [v]x^2, with [v]x^2 = v v^T - |v|^2 I.
__device__ __forceinline__ M3f so3_poly(float a, float b, float c, float3 v) {
    const float n2 = dotv(v, v);
    M3f r;
    const float vv[3] = {v.x, v.y, v.z};
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) r.m[i][j] = c * vv[i] * vv[j] + (i == j ? a - c * n2 : 0.f);
    r.m[0][1] += -b * v.z;
    r.m[0][2] += b * v.y;
    r.m[1][0] += b * v.z;
    r.m[1][2] += -b * v.x;
    r.m[2][0] += -b * v.y;
    r.m[2][1] += b * v.x;
    return r;
}

// Exp_so3(v) = I + sin(n)/n [v] + (1-cos n)/n^2 [v]^2.
__device__ __forceinline__ M3f exp_so3(float3 v) {
    const float n = normv(v);
    const float sh = sinc(0.5f * n);
    return so3_poly(1.f, sinc(n), 0.5f * sh * sh, v);
}

// Right Jacobian J_r(t) = I - (1-cos n)/n^2 [t] + (n - sin n)/n^3 [t]^2.
__device__ __forceinline__ M3f jr(float3 t) {
    const float n = normv(t), n2 = n * n;
    const float sh = sinc(0.5f * n);
    const float a = 0.5f * sh * sh;
    const float c = n < 0.5f ? (1.f / 6.f) - n2 * ((1.f / 120.f) - n2 * (1.f / 5040.f)) : (n - sinf(n)) / (n2 * n);
    return so3_poly(1.f, -a, c, t);
}

// Inverse right Jacobian J_r^-1(t) = I + 1/2 [t] + (1/n^2 - (1+cos n)/(2 n sin n)) [t]^2.
__device__ __forceinline__ M3f jr_inv(float3 t) {
    const float n = normv(t), n2 = n * n;
    const float c = n < 0.5f ? (1.f / 12.f) + n2 * ((1.f / 720.f) + n2 * (1.f / 30240.f))
                             : 1.f / n2 - (1.f + cosf(n)) / (2.f * n * sinf(n));
    return so3_poly(1.f, 0.5f, c, t);
}

__device__ __forceinline__ M3f mmul(const M3f& a, const M3f& b) {
    M3f r;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j];
    return r;
}
__device__ __forceinline__ M3f mtrans(const M3f& a) {
    M3f r;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) r.m[i][j] = a.m[j][i];
    return r;
}
__device__ __forceinline__ float3 mvec(const M3f& a, float3 v) {
    return v3(a.m[0][0] * v.x + a.m[0][1] * v.y + a.m[0][2] * v.z, a.m[1][0] * v.x + a.m[1][1] * v.y + a.m[1][2] * v.z,
              a.m[2][0] * v.x + a.m[2][1] * v.y + a.m[2][2] * v.z);
}
// a^T v
__device__ __forceinline__ float3 mtvec(const M3f& a, float3 v) {
    return v3(a.m[0][0] * v.x + a.m[1][0] * v.y + a.m[2][0] * v.z, a.m[0][1] * v.x + a.m[1][1] * v.y + a.m[2][1] * v.z,
              a.m[0][2] * v.x + a.m[1][2] * v.y + a.m[2][2] * v.z);
}
__device__ __forceinline__ M3f skew(float3 v) { return so3_poly(0.f, 1.f, 0.f, v); }
__device__ __forceinline__ M3f mscale(float s, const M3f& a) {
    M3f r;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) r.m[i][j] = s * a.m[i][j];
    return r;
}

// Symmetric 6x6 (upper triangle, row-major packing: 21 entries) and 6-vector.
struct Sym6 {
    float a[21];
};
struct Vec6 {
    float v[6];
};
__device__ __forceinline__ constexpr int s6(int i, int j) {
    return i <= j ? i * 6 - (i * (i - 1)) / 2 + (j - i) : j * 6 - (j * (j - 1)) / 2 + (i - j);
}
__device__ __forceinline__ void sym6_zero(Sym6& h) {
#pragma unroll
    for (int k = 0; k < 21; k++) h.a[k] = 0.f;
}
__device__ __forceinline__ void vec6_zero(Vec6& g) {
#pragma unroll
    for (int k = 0; k < 6; k++) g.v[k] = 0.f;
}
// h += s * a a^T
__device__ __forceinline__ void sym6_outer(Sym6& h, const float a[6], float s) {
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
        for (int j = i; j < 6; j++) h.a[s6(i, j)] += s * a[i] * a[j];
}

// Solve A x = b for the symmetric 6x6 block system by Gaussian elimination
// with partial (row) pivoting -- the algorithm of Eigen's partialPivLu that the
// reference uses (simulation/VBDSolver.hpp:472,498).  The block is not
// always positive definite: the rotational Hessian of a rigid-ground contact
// at a point off the centre (box, capsule) and of the triangle-rigid
// contacts is indefinite, and a large multiplier can outweigh I/dt^2.  The
// pivot is the first row of largest magnitude in the column; the row swaps
// are register selects, so every index stays compile-time.
__device__ __forceinline__ void lu6(const Sym6& S, const float b[6], float x[6]) {
    float A[6][6], r[6];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        r[i] = b[i];
#pragma unroll
        for (int j = 0; j < 6; j++) A[i][j] = S.a[i <= j ? s6(i, j) : s6(j, i)];
    }
#pragma unroll
    for (int k = 0; k < 6; k++) {
        int piv = k;
        float best = fabsf(A[k][k]);
#pragma unroll
        for (int i = k + 1; i < 6; i++) {
            const float v = fabsf(A[i][k]);
            if (v > best) {
                best = v;
                piv = i;
            }
        }
#pragma unroll
        for (int i = k + 1; i < 6; i++) {
            const bool sw = piv == i;
#pragma unroll
            for (int j = 0; j < 6; j++) {
                const float t = A[k][j];
                A[k][j] = sw ? A[i][j] : t;
                A[i][j] = sw ? t : A[i][j];
            }
            const float t = r[k];
            r[k] = sw ? r[i] : t;
            r[i] = sw ? t : r[i];
        }
        const float inv = 1.f / A[k][k];
#pragma unroll
        for (int i = k + 1; i < 6; i++) {
            const float f = A[i][k] * inv;
#pragma unroll
            for (int j = k + 1; j < 6; j++) A[i][j] -= f * A[k][j];
            r[i] -= f * r[k];
        }
    }
#pragma unroll
    for (int i = 5; i >= 0; i--) {
        float t = r[i];
#pragma unroll
        for (int j = i + 1; j < 6; j++) t -= A[i][j] * x[j];
        x[i] = t / A[i][i];
    }
}

}  // namespace rot
}  // namespace avbd
