// mlp_model.hpp -- host-side model of the learned warm start (see mlp_host.cpp).
#pragma once

#include <vector>

#include "mlp.hpp"

namespace avbd {

struct MlpModel {
    struct Layer {
        int in = 0, out = 0;
        float* w = nullptr;        // FP32 master, out x in (row-major)
        uint16_t* w16 = nullptr;   // bf16 copy
        float* b = nullptr;
        float *gw = nullptr, *gb = nullptr;
        float *mw = nullptr, *vw = nullptr, *mb = nullptr, *vb = nullptr;   // Adam moments
    };

    explicit MlpModel(hipStream_t s);
    ~MlpModel();
    MlpModel(const MlpModel&) = delete;
    MlpModel& operator=(const MlpModel&) = delete;

    // Flat parameter layout: for each layer W (out x in, padded) then b (out),
    // then feature mean (24), feature std (24), target mean (3), target std (3).
    static size_t param_count();
    void init(uint64_t seed);
    void set_params(const float* p);
    void get_params(float* p) const;
    void get_grads(float* p) const;   // W and b gradients in the same layout (no normalisation block)
    void set_norm(const float* norm); // 2*24 + 2*3 floats

    // One forward + backward pass over `rows` raw samples (features rows x 24,
    // targets rows x 3, both FP32 host); gradients stay on the device.
    float forward_backward(const float* feat, const float* tgt, int rows);
    void predict(const float* feat, int rows, float* out);
    float train(const float* feat, const float* tgt, int rows, int epochs, int batch, float lr, uint64_t seed);

    // device-side pieces used by the in-step inference
    void ensure_batch(int rows_pad);
    void forward(int rows_pad, const uint16_t* x, bool keep_pre = true);

    hipStream_t stream;
    Layer layers[kMlpLayers];
    float* d_norm = nullptr;
    float* d_loss = nullptr;
    std::vector<float> host_norm;
    bool loaded = false;
    int adam_t = 0;

    // batch buffers
    int cap_rows = 0;
    uint16_t* x16 = nullptr;
    float* z32[kMlpLayers] = {};
    uint16_t* h16[kMlpLayers] = {};
    uint16_t* d16[kMlpLayers] = {};
    float* t32 = nullptr;

private:
    void upload_layer(int l, const float* w, const float* b);
    void zero_adam();
    void stage_batch(const float* feat, const float* tgt, int rows, int rows_pad);
    void backward(int rows, int rows_pad, const uint16_t* x);
    void adam_step(float lr);
    std::vector<void*> owned, batch_owned;
};

}  // namespace avbd
