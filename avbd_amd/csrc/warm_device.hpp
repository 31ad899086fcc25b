// warm_device.hpp -- device helpers shared by the FP32 and FP64 forms of the
// learned warm start: the step lengths tried along the prediction and the
// per-scene sums of the incremental-potential differences.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "mlp.hpp"

namespace avbd {

// Step lengths tried along the prediction: x0 + a * pred.
__device__ constexpr float kWarmSteps[kWarmAlphas] = {1.0f, 0.5f, 0.25f};

// Add v[j] to acc[j * stride + scene] for NV values over a block of 256
// threads.  When every active lane of the block has one scene (a one-scene
// run, or the interior of a scene in a batch) the block adds one sum per
// value; otherwise each wavefront whose active lanes share a scene adds its
// sums, and the lanes of a mixed wavefront add their own values.  Every
// thread of the block must call it.
template <int NV>
__device__ __forceinline__ void add_per_scene_block(double* acc, uint32_t stride, uint32_t scene, const double* v,
                                                    bool active) {
    __shared__ double part[4][NV];
    __shared__ uint32_t wscene[4], wstate[4];   // state: 0 no active lane, 1 one scene, 2 mixed
    const unsigned long long act = __ballot(active);
    const uint32_t wave = threadIdx.x / 64u, lane = __lane_id();
    uint32_t s0 = 0, state = 0;
    if (act) {
        const int first = __ffsll((long long)act) - 1;
        s0 = __shfl(scene, first);
        state = __all(!active || scene == s0) ? 1u : 2u;
    }
    double t[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) {
        t[j] = active ? v[j] : 0.0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) t[j] += __shfl_xor(t[j], off);
    }
    if (lane == 0) {
        wscene[wave] = s0;
        wstate[wave] = state;
#pragma unroll
        for (int j = 0; j < NV; j++) part[wave][j] = t[j];
    }
    __syncthreads();
    bool uniform = true;
    uint32_t sb = 0;
    bool any = false;
    for (int w = 0; w < 4; w++) {
        if (wstate[w] == 2u) uniform = false;
        if (wstate[w] == 1u) {
            if (any && wscene[w] != sb) uniform = false;
            sb = wscene[w];
            any = true;
        }
    }
    if (uniform) {
        if (any && threadIdx.x == 0) {
#pragma unroll
            for (int j = 0; j < NV; j++)
                atomicAdd(&acc[j * stride + sb], part[0][j] + part[1][j] + part[2][j] + part[3][j]);
        }
        return;
    }
    if (state == 1u) {
        if (lane == 0) {
#pragma unroll
            for (int j = 0; j < NV; j++) atomicAdd(&acc[j * stride + s0], t[j]);
        }
    } else if (state == 2u && active) {
#pragma unroll
        for (int j = 0; j < NV; j++) atomicAdd(&acc[j * stride + scene], v[j]);
    }
}

}  // namespace avbd
