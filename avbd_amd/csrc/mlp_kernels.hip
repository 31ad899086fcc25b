// mlp_kernels.hip -- the learned warm start: a per-vertex MLP
// (24 -> 128 -> 128 -> 128 -> 3, SiLU) that predicts the converged step
// displacement from 1-ring features and replaces the inertial initial guess.
//
// Every matrix product, in inference and in training (forward and backward),
// runs on gfx950's BF16 MFMA with FP32 accumulation
// (v_mfma_f32_16x16x32_bf16, 32-deep products): bf16 operands in HBM,
// operand transposition chosen at compile time, and fused epilogues (bias,
// SiLU, SiLU derivative, FP32 / BF16 stores).  Weights keep an FP32 master
// copy updated by Adam; the bf16 copy is refreshed after each update.
//
// MFMA operand layout (16x16x32, 64 lanes, lane l, r = l % 16, g = l / 16):
//   A (16 x 32, rows x k): lane holds A[r][8g .. 8g+7]
//   B (32 x 16, k x cols): lane holds B[8g .. 8g+7][r]
//   D (16 x 16):           lane holds D[4g .. 4g+3][r]
#include <hip/hip_runtime.h>

#include <cstdint>

#include "internal.hpp"
#include "mlp.hpp"

namespace avbd {

namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __bf16 as_bf16(uint16_t u) { return __builtin_bit_cast(__bf16, u); }

__device__ __forceinline__ uint16_t f2bf(float f) {
    // round to nearest even (NaN kept quiet)
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ float sigmoidf_(float z) { return 1.f / (1.f + __expf(-z)); }

// C[M x N] = A(M x K) B(K x N) with A(m, k) = TA ? a[k * lda + m] : a[m * lda + k]
// and B(k, n) = TB ? b[n * ldb + k] : b[k * ldb + n].  M, N, K multiples of
// 16; a K that is not a multiple of 32 is zero-filled to the next one.
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(const uint16_t* __restrict__ a, int lda,
                                                        const uint16_t* __restrict__ b, int ldb,
                                                        float* __restrict__ c32, uint16_t* __restrict__ c16, int ldc,
                                                        const float* __restrict__ bias, const float* __restrict__ aux,
                                                        int ldaux, int M, int N, int K) {
    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int tiles_n = N >> 4;
    if (tile >= (M >> 4) * tiles_n) return;   // whole wavefronts leave together
    const int tm = (tile / tiles_n) << 4, tn = (tile % tiles_n) << 4;
    const int r = lane & 15, g = lane >> 4;
    float4_t acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 32) {
        bf16x8_t av, bv;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int k = k0 + 8 * g + j;
            const int m = tm + r, n = tn + r;
            const bool in = k < K;
            av[j] = as_bf16(in ? (TA ? a[(size_t)k * lda + m] : a[(size_t)m * lda + k]) : (uint16_t)0);
            bv[j] = as_bf16(in ? (TB ? b[(size_t)n * ldb + k] : b[(size_t)k * ldb + n]) : (uint16_t)0);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
    }
    const int n = tn + r;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int m = tm + 4 * g + j;
        float v = acc[j];
        const size_t o = (size_t)m * ldc + n;
        if constexpr (EPI == MLP_EPI_F32) {
            c32[o] = v;
        } else if constexpr (EPI == MLP_EPI_BIAS_F32) {
            c32[o] = v + bias[n];
        } else if constexpr (EPI == MLP_EPI_BIAS_SILU) {
            const float z = v + bias[n];
            if (c32) c32[o] = z;                       // pre-activation, kept for the backward pass
            c16[o] = f2bf(z * sigmoidf_(z));
        } else if constexpr (EPI == MLP_EPI_SILU_BWD) {
            const float z = aux[(size_t)m * ldaux + n];
            const float s = sigmoidf_(z);
            c16[o] = f2bf(v * s * (1.f + z * (1.f - s)));
        }
    }
}

// Eight consecutive bf16 (16-byte aligned) as an MFMA operand fragment.
__device__ __forceinline__ bf16x8_t ld_frag(const uint16_t* p) {
    return __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const uint4*>(p));
}

// The forward products: C[M x N] = A(M x K) B(K x N) with A row-major and
// B(k, n) = b[n * ldb + k] (weights stored out x in), N = 16 NT, K a
// multiple of 32, lda and ldb multiples of 8 (16-byte fragments).  One
// wavefront owns a 16-row strip and all NT output tiles of it, so each A
// fragment is read once (the one-tile-per-wavefront kernel above reads it NT
// times) and the weight fragments, the same for every strip, come from the
// caches.  The k loop and the MFMA sequence per tile are those of
// gemm_bf16_kernel, so the results are identical.
template <int EPI, int NT>
__global__ __launch_bounds__(256) void gemm_rows_kernel(const uint16_t* __restrict__ a, int lda,
                                                        const uint16_t* __restrict__ b, int ldb,
                                                        float* __restrict__ c32, uint16_t* __restrict__ c16, int ldc,
                                                        const float* __restrict__ bias, int M, int K) {
    const int lane = threadIdx.x & 63;
    const int strip = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (strip >= (M >> 4)) return;   // whole wavefronts leave together
    const int tm = strip << 4;
    const int r = lane & 15, g = lane >> 4;
    float4_t acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = float4_t{0.f, 0.f, 0.f, 0.f};
    const uint16_t* arow = a + (size_t)(tm + r) * lda + 8 * g;
    for (int k0 = 0; k0 < K; k0 += 32) {
        const bf16x8_t av = ld_frag(arow + k0);
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const bf16x8_t bv = ld_frag(b + (size_t)(t * 16 + r) * ldb + k0 + 8 * g);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int n = t * 16 + r;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int m = tm + 4 * g + j;
            const float v = acc[t][j];
            const size_t o = (size_t)m * ldc + n;
            if constexpr (EPI == MLP_EPI_F32) {
                c32[o] = v;
            } else if constexpr (EPI == MLP_EPI_BIAS_F32) {
                c32[o] = v + bias[n];
            } else if constexpr (EPI == MLP_EPI_BIAS_SILU) {
                const float z = v + bias[n];
                if (c32) c32[o] = z;   // pre-activation, kept for the backward pass
                c16[o] = f2bf(z * sigmoidf_(z));
            }
        }
    }
}

// Output-layer loss gradient: dY = 2 (Y - T) / rows for the first `outs`
// columns of the first `rows` rows (zero in the padding), as bf16; adds the
// squared-error sum to *loss.
__global__ void mse_grad_kernel(const float* __restrict__ y, const float* __restrict__ t, uint16_t* __restrict__ dy,
                                int rows, int rows_pad, int ld, int outs, float* __restrict__ loss) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows_pad * ld) return;
    const int col = i % ld, row = i / ld;
    float d = 0.f;
    if (col < outs && row < rows) {
        d = y[i] - t[(size_t)row * outs + col];
        atomicAdd(loss, d * d);
    }
    dy[i] = f2bf(2.f * d / (float)rows);
}

// Column sums of a bf16 matrix (bias gradient): out[n] = sum_m d[m][n].
__global__ void colsum_kernel(const uint16_t* __restrict__ d, int rows, int ld, int cols, float* __restrict__ out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= cols) return;
    float s = 0.f;
    for (int m = 0; m < rows; m++) s += bf2f(d[(size_t)m * ld + n]);
    out[n] = s;
}

// Adam on FP32 master weights; refreshes the bf16 copy.
__global__ void adam_kernel(float* __restrict__ w, uint16_t* __restrict__ w16, const float* __restrict__ gw,
                            float* __restrict__ m1, float* __restrict__ m2, int n, float lr, float b1, float b2,
                            float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = gw[i];
    const float a = b1 * m1[i] + (1.f - b1) * g;
    const float b = b2 * m2[i] + (1.f - b2) * g * g;
    m1[i] = a;
    m2[i] = b;
    const float v = w[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
    w[i] = v;
    if (w16) w16[i] = f2bf(v);
}

__global__ void f32_to_bf16_kernel(const float* __restrict__ x, uint16_t* __restrict__ y, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

inline int blocks_for(int n, int b) { return (n + b - 1) / b; }

}  // namespace

void launch_gemm_bf16(hipStream_t s, bool ta, bool tb, int epi, const uint16_t* a, int lda, const uint16_t* b,
                      int ldb, float* c32, uint16_t* c16, int ldc, const float* bias, const float* aux, int ldaux,
                      int M, int N, int K) {
    if (!ta && tb && epi != MLP_EPI_SILU_BWD && (N == 128 || N == 16) && K % 32 == 0 && lda % 8 == 0 &&
        ldb % 8 == 0) {
        const dim3 grid(blocks_for(M / 16, 4)), block(256);
#define AVBD_ROWS(E_, NT_)                                                                                      \
    if (epi == E_ && N == 16 * NT_) {                                                                           \
        hipLaunchKernelGGL((gemm_rows_kernel<E_, NT_>), grid, block, 0, s, a, lda, b, ldb, c32, c16, ldc, bias, \
                           M, K);                                                                               \
        return;                                                                                                 \
    }
        AVBD_ROWS(MLP_EPI_BIAS_SILU, 8)
        AVBD_ROWS(MLP_EPI_BIAS_F32, 8)
        AVBD_ROWS(MLP_EPI_F32, 8)
        AVBD_ROWS(MLP_EPI_BIAS_SILU, 1)
        AVBD_ROWS(MLP_EPI_BIAS_F32, 1)
        AVBD_ROWS(MLP_EPI_F32, 1)
#undef AVBD_ROWS
    }
    const int tiles = (M / 16) * (N / 16);
    const dim3 grid(blocks_for(tiles, 4)), block(256);
#define AVBD_GEMM(TA_, TB_, E_)                                                                                 \
    if (ta == TA_ && tb == TB_ && epi == E_) {                                                                  \
        hipLaunchKernelGGL((gemm_bf16_kernel<TA_, TB_, E_>), grid, block, 0, s, a, lda, b, ldb, c32, c16, ldc,  \
                           bias, aux, ldaux, M, N, K);                                                          \
        return;                                                                                                 \
    }
    // forward: activations (rows x k) times W stored transposed (out x in)
    AVBD_GEMM(false, true, MLP_EPI_BIAS_SILU)
    AVBD_GEMM(false, true, MLP_EPI_BIAS_F32)
    AVBD_GEMM(false, true, MLP_EPI_F32)
    // backward: dH = dZ W (W stored out x in, i.e. B(k=out, n=in) = w[k * in + n])
    AVBD_GEMM(false, false, MLP_EPI_SILU_BWD)
    AVBD_GEMM(false, false, MLP_EPI_F32)
    // weight gradient: dW (out x in) = dZ^T (out x rows) H (rows x in)
    AVBD_GEMM(true, false, MLP_EPI_F32)
#undef AVBD_GEMM
}

void launch_mse_grad(hipStream_t s, const float* y, const float* t, uint16_t* dy, int rows, int rows_pad, int ld,
                     int outs, float* loss) {
    hipLaunchKernelGGL(mse_grad_kernel, dim3(blocks_for(rows_pad * ld, 256)), dim3(256), 0, s, y, t, dy, rows,
                       rows_pad, ld, outs, loss);
}

void launch_colsum(hipStream_t s, const uint16_t* d, int rows, int ld, int cols, float* out) {
    hipLaunchKernelGGL(colsum_kernel, dim3(blocks_for(cols, 128)), dim3(128), 0, s, d, rows, ld, cols, out);
}

void launch_adam(hipStream_t s, float* w, uint16_t* w16, const float* gw, float* m1, float* m2, int n, float lr,
                 float b1, float b2, float eps, int t) {
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, w, w16, gw, m1, m2, n, lr, b1, b2,
                       eps, c1, c2);
}

void launch_f32_to_bf16(hipStream_t s, const float* x, uint16_t* y, int n) {
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, x, y, n);
}

}  // namespace avbd
