// mlp_host.cpp -- host side of the learned warm start: parameter storage
// (FP32 master + bf16 copies), Adam state, batch buffers, and the forward /
// backward / training drivers over the BF16 MFMA GEMM kernel.
#include "mlp_model.hpp"

#include <cmath>
#include <cstring>
#include <algorithm>
#include <random>
#include <stdexcept>
#include <string>

namespace avbd {

namespace {
void hchk(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
template <class T>
T* dalloc(std::vector<void*>& owned, size_t n) {
    void* p = nullptr;
    hchk(hipMalloc(&p, (n ? n : 1) * sizeof(T)), "hipMalloc");
    owned.push_back(p);
    return static_cast<T*>(p);
}
uint16_t host_f2bf(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
}  // namespace

size_t MlpModel::param_count() {
    size_t n = 0;
    for (int l = 0; l < kMlpLayers; l++) n += (size_t)mlp_layer_out(l) * mlp_layer_in(l) + mlp_layer_out(l);
    return n + 2 * kMlpIn + 2 * kMlpOut;
}

MlpModel::MlpModel(hipStream_t s) : stream(s) {
    for (int l = 0; l < kMlpLayers; l++) {
        const int in = mlp_layer_in(l), out = mlp_layer_out(l);
        Layer& L = layers[l];
        L.in = in;
        L.out = out;
        L.w = dalloc<float>(owned, (size_t)out * in);
        L.w16 = dalloc<uint16_t>(owned, (size_t)out * in);
        L.b = dalloc<float>(owned, out);
        L.gw = dalloc<float>(owned, (size_t)out * in);
        L.gb = dalloc<float>(owned, out);
        L.mw = dalloc<float>(owned, (size_t)out * in);
        L.vw = dalloc<float>(owned, (size_t)out * in);
        L.mb = dalloc<float>(owned, out);
        L.vb = dalloc<float>(owned, out);
    }
    d_norm = dalloc<float>(owned, 2 * kMlpIn + 2 * kMlpOut);
    d_loss = dalloc<float>(owned, 1);
    host_norm.assign(2 * kMlpIn + 2 * kMlpOut, 0.f);
    for (int i = 0; i < kMlpIn; i++) host_norm[kMlpIn + i] = 1.f;            // feature std
    for (int i = 0; i < kMlpOut; i++) host_norm[2 * kMlpIn + kMlpOut + i] = 1.f;   // target std
}

MlpModel::~MlpModel() {
    for (void* p : owned) (void)hipFree(p);
}

void MlpModel::init(uint64_t seed) {
    std::mt19937_64 rng(seed);
    for (int l = 0; l < kMlpLayers; l++) {
        Layer& L = layers[l];
        // uniform(-a, a) with a = sqrt(6 / fan_in) (He); padded inputs/outputs zero
        const int real_in = l == 0 ? kMlpIn : kMlpHidden;
        const int real_out = l == kMlpLayers - 1 ? kMlpOut : kMlpHidden;
        const float a = std::sqrt(6.f / (float)real_in) * (l == kMlpLayers - 1 ? 0.1f : 1.f);
        std::uniform_real_distribution<float> U(-a, a);
        std::vector<float> w((size_t)L.out * L.in, 0.f), b(L.out, 0.f);
        for (int o = 0; o < real_out; o++)
            for (int i = 0; i < real_in; i++) w[(size_t)o * L.in + i] = U(rng);
        upload_layer(l, w.data(), b.data());
    }
    adam_t = 0;
    zero_adam();
}

void MlpModel::zero_adam() {
    for (auto& L : layers) {
        const size_t nw = (size_t)L.out * L.in;
        hchk(hipMemsetAsync(L.mw, 0, nw * 4, stream), "memset");
        hchk(hipMemsetAsync(L.vw, 0, nw * 4, stream), "memset");
        hchk(hipMemsetAsync(L.mb, 0, L.out * 4, stream), "memset");
        hchk(hipMemsetAsync(L.vb, 0, L.out * 4, stream), "memset");
    }
}

void MlpModel::upload_layer(int l, const float* w, const float* b) {
    Layer& L = layers[l];
    const size_t nw = (size_t)L.out * L.in;
    std::vector<uint16_t> w16(nw);
    for (size_t i = 0; i < nw; i++) w16[i] = host_f2bf(w[i]);
    hchk(hipMemcpyAsync(L.w, w, nw * 4, hipMemcpyHostToDevice, stream), "H2D");
    hchk(hipMemcpyAsync(L.w16, w16.data(), nw * 2, hipMemcpyHostToDevice, stream), "H2D");
    hchk(hipMemcpyAsync(L.b, b, L.out * 4, hipMemcpyHostToDevice, stream), "H2D");
    hchk(hipStreamSynchronize(stream), "sync");
}

void MlpModel::set_params(const float* p) {
    size_t off = 0;
    for (int l = 0; l < kMlpLayers; l++) {
        const Layer& L = layers[l];
        const size_t nw = (size_t)L.out * L.in;
        upload_layer(l, p + off, p + off + nw);
        off += nw + L.out;
    }
    host_norm.assign(p + off, p + off + 2 * kMlpIn + 2 * kMlpOut);
    hchk(hipMemcpyAsync(d_norm, host_norm.data(), host_norm.size() * 4, hipMemcpyHostToDevice, stream), "H2D");
    hchk(hipStreamSynchronize(stream), "sync");
    loaded = true;
}

void MlpModel::get_params(float* p) const {
    size_t off = 0;
    for (int l = 0; l < kMlpLayers; l++) {
        const Layer& L = layers[l];
        const size_t nw = (size_t)L.out * L.in;
        hchk(hipMemcpyAsync(p + off, L.w, nw * 4, hipMemcpyDeviceToHost, stream), "D2H");
        hchk(hipMemcpyAsync(p + off + nw, L.b, L.out * 4, hipMemcpyDeviceToHost, stream), "D2H");
        off += nw + L.out;
    }
    hchk(hipStreamSynchronize(stream), "sync");
    std::memcpy(p + off, host_norm.data(), host_norm.size() * 4);
}

void MlpModel::get_grads(float* p) const {
    size_t off = 0;
    for (int l = 0; l < kMlpLayers; l++) {
        const Layer& L = layers[l];
        const size_t nw = (size_t)L.out * L.in;
        hchk(hipMemcpyAsync(p + off, L.gw, nw * 4, hipMemcpyDeviceToHost, stream), "D2H");
        hchk(hipMemcpyAsync(p + off + nw, L.gb, L.out * 4, hipMemcpyDeviceToHost, stream), "D2H");
        off += nw + L.out;
    }
    hchk(hipStreamSynchronize(stream), "sync");
}

void MlpModel::set_norm(const float* norm) {
    host_norm.assign(norm, norm + 2 * kMlpIn + 2 * kMlpOut);
    hchk(hipMemcpyAsync(d_norm, host_norm.data(), host_norm.size() * 4, hipMemcpyHostToDevice, stream), "H2D");
    hchk(hipStreamSynchronize(stream), "sync");
}

void MlpModel::ensure_batch(int rows) {
    const int pad = (rows + 15) / 16 * 16;
    if (pad <= cap_rows) return;
    for (void* p : batch_owned) (void)hipFree(p);
    batch_owned.clear();
    cap_rows = pad;
    x16 = dalloc<uint16_t>(batch_owned, (size_t)pad * kMlpInPad);
    for (int l = 0; l < kMlpLayers; l++) {
        z32[l] = dalloc<float>(batch_owned, (size_t)pad * mlp_layer_out(l));
        h16[l] = dalloc<uint16_t>(batch_owned, (size_t)pad * mlp_layer_out(l));
        d16[l] = dalloc<uint16_t>(batch_owned, (size_t)pad * mlp_layer_out(l));
    }
    t32 = dalloc<float>(batch_owned, (size_t)pad * kMlpOut);
}

// Forward pass of rows_pad rows of x16 (bf16, normalised, padded): with
// keep_pre the hidden layers keep their pre-activations (z32) for the
// backward pass (inference skips those stores); the output layer leaves FP32
// in z32[3].
void MlpModel::forward(int rows_pad, const uint16_t* x, bool keep_pre) {
    const uint16_t* in = x;
    for (int l = 0; l < kMlpLayers; l++) {
        const Layer& L = layers[l];
        const bool last = l == kMlpLayers - 1;
        launch_gemm_bf16(stream, false, true, last ? MLP_EPI_BIAS_F32 : MLP_EPI_BIAS_SILU, in, L.in, L.w16, L.in,
                         (last || keep_pre) ? z32[l] : nullptr, last ? nullptr : h16[l], L.out, L.b, nullptr, 0,
                         rows_pad, L.out, L.in);
        in = h16[l];
    }
}

// Backward pass for MSE on the first kMlpOut outputs against t32 (rows rows,
// padded to rows_pad with zero rows); writes gw / gb of every layer and the
// loss sum into d_loss.
void MlpModel::backward(int rows, int rows_pad, const uint16_t* x) {
    hchk(hipMemsetAsync(d_loss, 0, 4, stream), "memset");
    launch_mse_grad(stream, z32[kMlpLayers - 1], t32, d16[kMlpLayers - 1], rows, rows_pad, kMlpOutPad, kMlpOut,
                    d_loss);
    for (int l = kMlpLayers - 1; l >= 0; l--) {
        const Layer& L = layers[l];
        const uint16_t* act_in = l == 0 ? x : h16[l - 1];
        // dW (out x in) = dZ^T H_in ; db = colsum dZ
        launch_gemm_bf16(stream, true, false, MLP_EPI_F32, d16[l], L.out, act_in, L.in, L.gw, nullptr, L.in, nullptr,
                         nullptr, 0, L.out, L.in, rows_pad);
        launch_colsum(stream, d16[l], rows_pad, L.out, L.out, L.gb);
        if (l > 0) {
            // dZ_{l-1} = (dZ_l W_l) * silu'(Z_{l-1})
            launch_gemm_bf16(stream, false, false, MLP_EPI_SILU_BWD, d16[l], L.out, L.w16, L.in, nullptr, d16[l - 1],
                             L.in, nullptr, z32[l - 1], L.in, rows_pad, L.in, L.out);
        }
    }
}

void MlpModel::adam_step(float lr) {
    adam_t++;
    for (auto& L : layers) {
        launch_adam(stream, L.w, L.w16, L.gw, L.mw, L.vw, L.out * L.in, lr, 0.9f, 0.999f, 1e-8f, adam_t);
        launch_adam(stream, L.b, nullptr, L.gb, L.mb, L.vb, L.out, lr, 0.9f, 0.999f, 1e-8f, adam_t);
    }
}

// Normalise raw features (rows x 24, FP32 host) into x16 and targets into t32.
void MlpModel::stage_batch(const float* feat, const float* tgt, int rows, int rows_pad) {
    std::vector<uint16_t> xb((size_t)rows_pad * kMlpInPad, 0);
    for (int r = 0; r < rows; r++)
        for (int i = 0; i < kMlpIn; i++) {
            const float mu = host_norm[i], sd = host_norm[kMlpIn + i];
            xb[(size_t)r * kMlpInPad + i] = host_f2bf((feat[(size_t)r * kMlpIn + i] - mu) / sd);
        }
    hchk(hipMemcpyAsync(x16, xb.data(), xb.size() * 2, hipMemcpyHostToDevice, stream), "H2D");
    if (tgt) {
        std::vector<float> tb((size_t)rows_pad * kMlpOut, 0.f);
        for (int r = 0; r < rows; r++)
            for (int j = 0; j < kMlpOut; j++) {
                const float mu = host_norm[2 * kMlpIn + j], sd = host_norm[2 * kMlpIn + kMlpOut + j];
                tb[(size_t)r * kMlpOut + j] = (tgt[(size_t)r * kMlpOut + j] - mu) / sd;
            }
        hchk(hipMemcpyAsync(t32, tb.data(), tb.size() * 4, hipMemcpyHostToDevice, stream), "H2D");
    }
}

float MlpModel::forward_backward(const float* feat, const float* tgt, int rows) {
    const int pad = (rows + 15) / 16 * 16;
    ensure_batch(pad);
    stage_batch(feat, tgt, rows, pad);
    forward(pad, x16);
    backward(rows, pad, x16);
    float loss = 0;
    hchk(hipMemcpyAsync(&loss, d_loss, 4, hipMemcpyDeviceToHost, stream), "D2H");
    hchk(hipStreamSynchronize(stream), "sync");
    return loss / (float)rows;
}

void MlpModel::predict(const float* feat, int rows, float* out) {
    const int pad = (rows + 15) / 16 * 16;
    ensure_batch(pad);
    stage_batch(feat, nullptr, rows, pad);
    forward(pad, x16);
    std::vector<float> y((size_t)pad * kMlpOutPad);
    hchk(hipMemcpyAsync(y.data(), z32[kMlpLayers - 1], y.size() * 4, hipMemcpyDeviceToHost, stream), "D2H");
    hchk(hipStreamSynchronize(stream), "sync");
    for (int r = 0; r < rows; r++)
        for (int j = 0; j < kMlpOut; j++)
            out[(size_t)r * kMlpOut + j] =
                y[(size_t)r * kMlpOutPad + j] * host_norm[2 * kMlpIn + kMlpOut + j] + host_norm[2 * kMlpIn + j];
}

// Minibatch Adam over the dataset (rows x 24 features, rows x 3 targets) in
// a fixed shuffled order per epoch; returns the mean loss of the last epoch.
float MlpModel::train(const float* feat, const float* tgt, int rows, int epochs, int batch, float lr,
                      uint64_t seed) {
    std::mt19937_64 rng(seed);
    std::vector<int> order(rows);
    for (int i = 0; i < rows; i++) order[i] = i;
    std::vector<float> bf((size_t)batch * kMlpIn), bt((size_t)batch * kMlpOut);
    float last = 0;
    for (int e = 0; e < epochs; e++) {
        std::shuffle(order.begin(), order.end(), rng);
        double sum = 0;
        int nb = 0;
        for (int s = 0; s + batch <= rows; s += batch) {
            for (int k = 0; k < batch; k++) {
                std::memcpy(&bf[(size_t)k * kMlpIn], feat + (size_t)order[s + k] * kMlpIn, kMlpIn * 4);
                std::memcpy(&bt[(size_t)k * kMlpOut], tgt + (size_t)order[s + k] * kMlpOut, kMlpOut * 4);
            }
            sum += forward_backward(bf.data(), bt.data(), batch);
            adam_step(lr);
            nb++;
        }
        last = nb ? (float)(sum / nb) : 0.f;
    }
    hchk(hipStreamSynchronize(stream), "sync");
    loaded = true;
    return last;
}

}  // namespace avbd
