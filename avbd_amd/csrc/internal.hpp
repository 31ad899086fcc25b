// internal.hpp -- device record layouts and the solver context behind the
// avbd C ABI.  Host staging pools are FP64 (the reference's Real = double,
// common/common.hpp:77); device pools are FP32 SoA/AoS records.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/avbd.h"

namespace avbd {

// ---------------------------------------------------------------------------
// Device records
// ---------------------------------------------------------------------------

// Neo-Hookean tet, static part: one 64-byte record (half a 128-byte line),
// fetched with four 16-byte loads.  VL = V*lambda, VM = V*mu and
// alpha = 1 + mu/lambda are folded at upload (energy/elastic/NeoHookeanEnergySolver.hpp:195-196).
struct alignas(16) TetRec {
    uint32_t idx[4];
    float q[9];   // Q, row-major
    float VL;
    float VM;
    float alpha;
};
static_assert(sizeof(TetRec) == 64, "TetRec must be 64 bytes");

// Strain-rate damping part (only read when some tet has kd != 0).
// E_prev is symmetric (0.5 (F^T F - I)), so 6 entries are stored.
struct alignas(16) TetDamp {
    float e00, e11, e22, e01;
    float e02, e12, damp;   // damp = 2/dt * V * kd
    float pad;
};
static_assert(sizeof(TetDamp) == 32, "TetDamp must be 32 bytes");

// Ground AL contact (CollisionConstraintEnergyInfo + cp_x/cp_z), 64 bytes.
struct alignas(16) GroundRec {
    float k_start, k_n, k_t, k_b;
    float lambda_n, lambda_t, lambda_b, C_n_prev;
    float mu_s, mu_k, cp_x, cp_z;
    uint32_t p;
    uint32_t flags;   // bit0 use_static, bit1 unclamped_lambda_tb
    uint32_t pad0, pad1;
};
static_assert(sizeof(GroundRec) == 64, "GroundRec must be 64 bytes");

constexpr uint32_t GF_USE_STATIC = 1u;
constexpr uint32_t GF_UNCLAMPED = 2u;

// Work item of one colour phase: vertex, first incidence, NH incidence count
// (bit 31 = conflicted: buffer the update), ground record (or -1).
struct alignas(16) VItem {
    uint32_t v;
    uint32_t inc_begin;
    uint32_t inc_count;
    int32_t ground;
};
constexpr uint32_t VITEM_CONFLICT = 0x80000000u;

// ---------------------------------------------------------------------------
// Kernel launch wrappers (defined in vbd_kernels.hip)
// ---------------------------------------------------------------------------
struct DevState {
    float4* x = nullptr;        // xyz + mass in w
    float4* x_buf = nullptr;
    float4* y = nullptr;
    float4* x_prev = nullptr;
    float4* x_last = nullptr;
    float4* x_lastlast = nullptr;
    float4* v = nullptr;
    float4* v_prev = nullptr;
    float4* x_scratch = nullptr;   // autotuner output sink
    uint8_t* fixed = nullptr;
    int32_t* ground_of = nullptr;  // ground record per particle or -1
    TetRec* tets = nullptr;
    TetDamp* damp = nullptr;
    GroundRec* ground = nullptr;
    uint32_t* inc = nullptr;       // (tet << 2) | local, grouped per work item
    VItem* items = nullptr;        // all colours back to back
    double* reduce = nullptr;      // residual / IP partial sums
    uint32_t n = 0, n_tets = 0, n_ground = 0, n_inc = 0, n_items = 0;
};

struct LaunchCfg {
    uint32_t lanes = 8;
    uint32_t block = 256;
};

// Grid barrier of the persistent step kernel (zeroed once at build time).
struct GridBarrier {
    uint32_t count;   // blocks arrived at the current barrier
    uint32_t gen;     // barrier generation
    uint32_t abort;   // set when a wait timed out; every block then leaves
    uint32_t pad;
};

struct StepArgs {
    DevState d;
    const uint32_t* color_offsets = nullptr;      // num_colors + 1
    const uint32_t* conflict_offsets = nullptr;   // num_colors + 1
    const uint32_t* conflicts = nullptr;
    const float* omega = nullptr;                 // Chebyshev omega per iteration
    GridBarrier* bar = nullptr;
    uint32_t iters = 0, num_colors = 0;
    float dt = 0.f, g = 0.f;
};

// Blocks of 256 threads that fit on the device at once for the persistent
// kernel with `lanes` lanes per vertex.
uint32_t persistent_grid(int device, uint32_t lanes, bool damping);
void launch_step_persistent(hipStream_t s, const StepArgs& a, uint32_t lanes, bool damping, uint32_t grid);

void launch_eprev(hipStream_t s, const DevState& d, float dt);
void launch_step_begin(hipStream_t s, const DevState& d, float dt, float g);
void launch_solve_color(hipStream_t s, const DevState& d, const LaunchCfg& cfg, uint32_t item_begin,
                        uint32_t item_count, bool damping, float dt, float4* x_out);
void launch_commit(hipStream_t s, const DevState& d, const uint32_t* list, uint32_t count);
void launch_cheb_dual(hipStream_t s, const DevState& d, float omega);
void launch_velocity(hipStream_t s, const DevState& d, float dt);
// plain streaming copy of n_float4 16-byte words (HBM counter calibration)
void launch_calib_copy(hipStream_t s, const float4* src, float4* dst, uint64_t n_float4);
// residual/IP pass: writes 3 partial sums per block into d.reduce; returns block count
uint32_t launch_stats(hipStream_t s, const DevState& d, const LaunchCfg& cfg, bool damping, float dt, float g);

// ---------------------------------------------------------------------------
// Host-side graph build (host_graph.cpp)
// ---------------------------------------------------------------------------
struct HostNH {
    uint32_t idx[4];
    double lambda, mu, kd;
    double Q[9];
    double V;
    double E_prev[9];
};
struct HostGround {
    uint32_t p;
    double k_start, k_n, k_t, k_b, lambda_n, lambda_t, lambda_b, C_n_prev, mu_s, mu_k, cp_x, cp_z;
    bool use_static, unclamped;
};

struct ColorPlan {
    uint32_t num_colors = 0;
    std::vector<uint32_t> color;              // per particle
    std::vector<uint32_t> item_offsets;       // per colour: [begin, end) in items
    std::vector<VItem> items;
    std::vector<uint32_t> inc;
    std::vector<uint32_t> conflict_offsets;   // per colour
    std::vector<uint32_t> conflict_list;
};

// Build adjacency and colour every scene with DSATUR; then lay out the
// per-colour work items.  Throws std::runtime_error on failure.
void build_color_plan(uint32_t n, const std::vector<HostNH>& nh, const std::vector<HostGround>& ground,
                      const std::vector<uint8_t>& fixed, const std::vector<uint32_t>& scene_offsets,
                      ColorPlan& plan);

// DSATUR on one CSR graph (ColorList::greedyColor2, common/ColorList.hpp:149-241).
uint32_t dsatur_csr(uint32_t n, const uint32_t* p_offsets, const uint32_t* p_neighbors, uint32_t* color);

}  // namespace avbd
