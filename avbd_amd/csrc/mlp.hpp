// mlp.hpp -- the learned warm start: network shape, device state and the
// launch wrappers of mlp_kernels.hip / the feature and apply kernels of
// vbd_kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace avbd {

// Network: 24 input features (padded to 32) -> 128 -> 128 -> 128 -> 3
// (padded to 16), SiLU after every hidden layer.
constexpr int kMlpIn = 24;
constexpr int kMlpInPad = 32;
constexpr int kMlpHidden = 128;
constexpr int kMlpOut = 3;
constexpr int kMlpOutPad = 16;
constexpr int kMlpLayers = 4;
constexpr int kWarmAlphas = 3;   // step lengths tried along the prediction: 1, 1/2, 1/4

// GEMM epilogues
constexpr int MLP_EPI_F32 = 0;        // C32 = acc
constexpr int MLP_EPI_BIAS_F32 = 1;   // C32 = acc + bias
constexpr int MLP_EPI_BIAS_SILU = 2;  // Z = acc + bias -> C32 (optional), C16 = bf16(silu(Z))
constexpr int MLP_EPI_SILU_BWD = 3;   // C16 = bf16(acc * silu'(aux))

inline int mlp_layer_in(int l) { return l == 0 ? kMlpInPad : kMlpHidden; }
inline int mlp_layer_out(int l) { return l == kMlpLayers - 1 ? kMlpOutPad : kMlpHidden; }

// Device buffers of the in-step warm start (features -> network -> apply ->
// incremental-potential comparison -> choice).
struct WarmIO {
    uint16_t* x16 = nullptr;             // rows_pad x 32 normalised bf16 features
    float* raw = nullptr;                // n_items x 24 raw features (recording) or null
    const float* out = nullptr;          // rows_pad x 16 network output (FP32)
    const float* norm = nullptr;         // feature mean/std (24 + 24), target mean/std (3 + 3)
    float* tgt = nullptr;                // n_items x 3 training targets (recording)
    const float2* mat = nullptr;         // per particle: log10 E, nu of the incident material
    const uint32_t* scene_of = nullptr;  // per particle
    float4* pred = nullptr;              // per particle: predicted step from the initial guess
    double* ipd = nullptr;               // per step length and scene: IP(x0 + a pred) - IP(x0)
    unsigned long long* counters = nullptr;   // [0] scene-steps predicted, [1] fallbacks to the inertial start
    float* step_of = nullptr;            // per scene: the chosen step along the prediction (0 = fallback), or null
    uint32_t n_scenes = 0;
};

void launch_gemm_bf16(hipStream_t s, bool ta, bool tb, int epi, const uint16_t* a, int lda, const uint16_t* b,
                      int ldb, float* c32, uint16_t* c16, int ldc, const float* bias, const float* aux, int ldaux,
                      int M, int N, int K);
void launch_mse_grad(hipStream_t s, const float* y, const float* t, uint16_t* dy, int rows, int rows_pad, int ld,
                     int outs, float* loss);
void launch_colsum(hipStream_t s, const uint16_t* d, int rows, int ld, int cols, float* out);
void launch_adam(hipStream_t s, float* w, uint16_t* w16, const float* gw, float* m1, float* m2, int n, float lr,
                 float b1, float b2, float eps, int t);
void launch_f32_to_bf16(hipStream_t s, const float* x, uint16_t* y, int n);

}  // namespace avbd
