// api.cpp -- the avbd C ABI (include/avbd.h): context lifetime, FP64 host
// staging pools, upload to FP32 device records, hipGraph capture of one time
// step, readback and statistics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "contacts.hpp"
#include "internal.hpp"
#include "mlp_model.hpp"
#include "lbvh.hpp"

using namespace avbd;

struct avbd_ctx {
    avbd_scene_desc desc{};
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side_stream = nullptr;   // broad phase, concurrent with the solve
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;

    // FP64 staging (ParticlePool / EnergyRegistry layout, reference Real = double)
    std::vector<double> x, v, v_prev, x_prev, mass;
    std::vector<uint8_t> fixed;
    std::vector<double> f_ext;   // per-particle external acceleration (3n; empty = none)
    std::vector<HostNH> nh;
    std::vector<HostGround> ground;
    std::vector<uint32_t> scene_offsets;   // empty = one scene
    std::vector<uint32_t> tris;            // collision triangles (3 particles each)
    // RotationPool (common/ParticlePool.hpp:7-69): slot per oriented particle
    std::vector<int32_t> rot_of;           // per particle, -1 = positional
    std::vector<double> q, w, inertia;     // per slot: (w, x, y, z), body-frame w, principal inertia
    std::vector<HostCos> cos;
    std::vector<HostRGround> rground;
    std::vector<HostJoint> joints;
    std::vector<HostJoint2> joints2;       // two-sided fixed joints (new)

    bool planned = false;   // host colour plan done
    bool built = false;     // device buffers uploaded
    bool damping = false;
    ColorPlan plan;
    DevState d;
    uint32_t* d_conflicts = nullptr;
    std::vector<void*> allocations;
    uint32_t reduce_blocks = 0;
    LbvhDev lb;

    // FP64 particle state (AVBD_FLAG_FP64_STATE)
    bool fp64 = false;
    F64State f;
    // a non-finite state was reported by the device (sticky until the state is reset)
    bool nonfinite = false;

    LaunchCfg cfg;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // residual mode: the step's graphs (head + first check interval, one
    // further interval replayed as often as needed, tail; see capture_graph),
    // an event after each interval, and per interval its converged flag,
    // copied back to pinned host memory after the interval
    hipGraphExec_t head_exec = nullptr, interval_exec = nullptr, tail_exec = nullptr;
    std::vector<hipEvent_t> part_ev;
    uint32_t* h_conv = nullptr;
    std::vector<float> omega_seq;   // Chebyshev weight of every iteration
    // Chebyshev / dual update fused into each iteration's last colour phase:
    // positions alternate between d.x (even iterations) and x_alt
    bool fuse_cheb = false;
    float4* x_alt = nullptr;
    uint32_t* cheb_list = nullptr;  // particles that are not work items of the last colour
    uint32_t n_cheb = 0;

    bool timing = false;
    std::vector<hipEvent_t> events;
    double solve_avg_ms = 0;
    double lbvh_avg_ms = 0;
    hipEvent_t ev_lb0 = nullptr, ev_lb1 = nullptr;
    uint32_t solve_launches = 0;
    hipEvent_t ev_a = nullptr, ev_b = nullptr;

    uint64_t steps_taken = 0, vertex_updates = 0;
    double last_step_ms = 0;

    // residual-targeted iteration (avbd_set_residual_target)
    double res_tol = 0;            // 0: fixed iteration count
    uint32_t res_max_iters = 0, res_check_every = 0;
    ResCtl* rc = nullptr;
    uint32_t n_res_scenes = 0;
    double res_mg2 = 0;            // ||M g||^2 over every free particle
    double* res_rr = nullptr;      // |r_v|^2 per particle
    uint32_t* res_scene_off = nullptr;
    double* res_mg2_scene = nullptr;
    double* res_acc_scene = nullptr;
    float* omega_tab = nullptr;    // Chebyshev weight per iteration of a step
    // a batch freezes each scene once it meets the target (per scene: flag,
    // iterations it ran, its work items; per particle: its scene)
    uint32_t* res_frozen = nullptr;
    uint32_t* res_iters = nullptr;
    uint32_t* res_items = nullptr;
    uint32_t* res_scene_of = nullptr;

    // learned warm start (mlp_host.cpp): the network, its in-step buffers,
    // and the recording of training data
    std::unique_ptr<MlpModel> mlp;
    bool warm = false;             // prediction replaces the inertial start
    bool recording = false;        // record features / targets of the steps
    WarmIO wio;
    uint32_t rows_pad = 0;
    float* d_raw = nullptr;        // n_items x 24
    float* d_tgt = nullptr;        // n_items x 3

    // dynamic contacts (avbd_set_dynamic_contacts)
    std::vector<HostPrim> prims;   // CollisionPrimitivePool, call order
    uint32_t next_object = 0;
    bool dynamic = false;
    bool self_collision = false;   // vertex-face contacts (avbd_set_self_collision)
    ContactSystem cs;
    DevPrims dp;
    LbvhDev lbd;                   // LBVH over every primitive
    DetRec* d_det = nullptr;
    uint32_t det_cap = 0;
    uint32_t* d_status = nullptr;  // [0] narrow-phase flags, [1] detected count
    RecolourDev rdev;              // device recolouring and work-list layout
    uint32_t* h_layout = nullptr;  // pinned: the 3 x 257 per-colour offsets, then the 5 recolouring counters
    // dynamic contacts: a detection round has changed the colours or the
    // conflict set since set-up (the work list is no longer the set-up one)
    bool layout_changed = false;
    std::vector<uint32_t> setup_color;   // the set-up colouring (dynamic contacts)
    struct Grow {                  // device buffer that grows on demand
        void* p = nullptr;
        size_t bytes = 0;
    };
    Grow g_ops, g_active, g_off, g_ent, g_nboff, g_nb, g_seeds;
    // residual mode with dynamic contacts: what the captured step graphs
    // were recorded with (launch sizes and device buffers); they are
    // captured again when a detection round changes any of it
    std::vector<uint64_t> dyn_graph_key;
    // pinned staging for a detection round's uploads (written only after the
    // round's stream synchronisation, so one buffer serves every round)
    uint8_t* stage = nullptr;
    size_t stage_cap = 0, stage_used = 0;

    // iterations per step: the budget in residual mode, else solver_iters
    uint32_t iters() const { return res_tol > 0 ? res_max_iters : desc.solver_iters; }
};

namespace {

std::string g_create_error;

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct CodeError : std::runtime_error {
    int code;
    CodeError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

const char* nonfinite_message() {
    return "a time step produced a non-finite position or velocity (NaN/Inf); the state stays unusable until it "
           "is set again (avbd_set_particle_state, avbd_load_checkpoint)";
}

// After a synchronisation: the device's non-finite flag (sticky on the host
// until the state is set again).
void check_status(avbd_ctx* c) {
    if (!c->built || !c->d.status) return;
    uint32_t st = 0;
    hip_check(hipMemcpy(&st, c->d.status, sizeof st, hipMemcpyDeviceToHost), "D2H");
    if (st & 1u) c->nonfinite = true;
    if (c->nonfinite) throw CodeError(AVBD_E_NONFINITE, nonfinite_message());
}

// The state was overwritten from the host: clear the non-finite flag.
void clear_status(avbd_ctx* c) {
    c->nonfinite = false;
    if (c->d.status) hip_check(hipMemsetAsync(c->d.status, 0, sizeof(uint32_t), c->stream), "hipMemsetAsync");
}

template <class F>
int guarded(avbd_ctx* c, F&& f) {
    if (!c) return AVBD_E_ARG;
    try {
        f();
        return AVBD_OK;
    } catch (const CodeError& e) {
        c->err = e.what();
        return e.code;
    } catch (const HipError& e) {
        c->err = e.what();
        return AVBD_E_HIP;
    } catch (const std::exception& e) {
        c->err = e.what();
        const std::string m = e.what();
        if (m.find("Ran out of colors") != std::string::npos) return AVBD_E_COLORS;
        if (m.find("capacity") != std::string::npos) return AVBD_E_CAPACITY;
        if (m.find("Unsupported collision") != std::string::npos) return AVBD_E_UNSUPPORTED_CONTACT;
        return AVBD_E_ARG;
    }
}

template <class T>
T* dev_alloc(avbd_ctx* c, size_t count) {
    if (count == 0) count = 1;
    void* p = nullptr;
    hip_check(hipMalloc(&p, count * sizeof(T)), "hipMalloc");
    c->allocations.push_back(p);
    return static_cast<T*>(p);
}

// Make g hold at least `bytes`; the old contents are not kept.
void* grow(avbd_ctx* c, avbd_ctx::Grow& g, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (g.bytes >= bytes) return g.p;
    if (g.p) {
        c->allocations.erase(std::remove(c->allocations.begin(), c->allocations.end(), g.p), c->allocations.end());
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        (void)hipFree(g.p);
        g.p = nullptr;
        g.bytes = 0;
    }
    const size_t nb = std::max(bytes, 2 * g.bytes);
    hip_check(hipMalloc(&g.p, nb), "hipMalloc");
    c->allocations.push_back(g.p);
    g.bytes = nb;
    return g.p;
}

template <class T>
void upload(avbd_ctx* c, T* dst, const std::vector<T>& src) {
    if (src.empty()) return;
    hip_check(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, c->stream),
              "hipMemcpyAsync H2D");
}

float4 f4(const double* p, double w) { return make_float4((float)p[0], (float)p[1], (float)p[2], (float)w); }

GroundRec to_dev(const HostGround& g) {
    GroundRec r{};
    r.k_start = (float)g.k_start;
    r.k_n = (float)g.k_n;
    r.k_t = (float)g.k_t;
    r.k_b = (float)g.k_b;
    r.lambda_n = (float)g.lambda_n;
    r.lambda_t = (float)g.lambda_t;
    r.lambda_b = (float)g.lambda_b;
    r.C_n_prev = (float)g.C_n_prev;
    r.mu_s = (float)g.mu_s;
    r.mu_k = (float)g.mu_k;
    r.cp_x = (float)g.cp_x;
    r.cp_z = (float)g.cp_z;
    r.p = g.p;
    r.flags = (g.use_static ? GF_USE_STATIC : 0u) | (g.unclamped ? GF_UNCLAMPED : 0u);
    return r;
}

// quaternion (w, x, y, z) FP64 -> device float4 (x, y, z, w)
float4 qf4(const double* q) { return make_float4((float)q[1], (float)q[2], (float)q[3], (float)q[0]); }

CosRec to_dev(const HostCos& h) {
    CosRec r{};
    r.i0 = h.idx[0];
    r.i1 = h.idx[1];
    r.L = (float)h.L;
    r.k0 = (float)h.stiffness[0];
    r.k1 = (float)h.stiffness[1];
    r.k2 = (float)h.stiffness[2];
    r.k3 = (float)h.stiffness[3];
    r.k4 = (float)h.stiffness[4];
    r.k5 = (float)h.stiffness[5];
    r.kx = (float)h.kappa0[0];
    r.ky = (float)h.kappa0[1];
    r.kz = (float)h.kappa0[2];
    return r;
}

RGroundRec to_dev(const HostRGround& h) {
    RGroundRec r{};
    r.k_start = (float)h.k_start;
    r.k_n = (float)h.k_n;
    r.k_t = (float)h.k_t;
    r.k_b = (float)h.k_b;
    r.lambda_n = (float)h.lambda_n;
    r.lambda_t = (float)h.lambda_t;
    r.lambda_b = (float)h.lambda_b;
    r.C_n_prev = (float)h.C_n_prev;
    r.mu_s = (float)h.mu_s;
    r.mu_k = (float)h.mu_k;
    r.radius = (float)h.radius;
    r.cpg_x = (float)h.cp_ground[0];
    r.cpg_y = (float)h.cp_ground[1];
    r.cpg_z = (float)h.cp_ground[2];
    r.cpl_x = (float)h.cp_rb_local[0];
    r.cpl_y = (float)h.cp_rb_local[1];
    r.cpl_z = (float)h.cp_rb_local[2];
    r.p = h.p;
    r.shape = h.shape;
    r.flags = (h.use_static ? GF_USE_STATIC : 0u) | (h.unclamped ? GF_UNCLAMPED : 0u);
    return r;
}

JointRec to_dev(const HostJoint& h) {
    JointRec r{};
    r.pos_off = make_float4((float)h.pos_offset[0], (float)h.pos_offset[1], (float)h.pos_offset[2], 0.f);
    r.rot_off = qf4(h.rot_offset);
    r.ref_pos = make_float4((float)h.ref_position[0], (float)h.ref_position[1], (float)h.ref_position[2], 0.f);
    r.ref_rot = qf4(h.ref_rotation);
    for (int i = 0; i < 6; i++) {
        r.k[i] = (float)h.k[i];
        r.lambda[i] = (float)h.lambda[i];
        r.C_prev[i] = (float)h.C_prev[i];
    }
    r.p = h.p;
    r.p2 = JOINT_ONE_SIDED;
    return r;
}

JointRec to_dev(const HostJoint2& h) {
    JointRec r{};
    r.pos_off = make_float4((float)h.pos_offset[0][0], (float)h.pos_offset[0][1], (float)h.pos_offset[0][2], 0.f);
    r.rot_off = qf4(h.rot_offset[0]);
    r.ref_pos = make_float4((float)h.pos_offset[1][0], (float)h.pos_offset[1][1], (float)h.pos_offset[1][2], 0.f);
    r.ref_rot = qf4(h.rot_offset[1]);
    for (int i = 0; i < 6; i++) {
        r.k[i] = (float)h.k[i];
        r.lambda[i] = (float)h.lambda[i];
        r.C_prev[i] = (float)h.C_prev[i];
    }
    r.p = h.p[0];
    r.p2 = h.p[1];
    return r;
}

// Append one particle to the FP64 staging pools (ParticlePool::addParticle).
void push_particle(avbd_ctx* c, const double* x, const double* v, const double* v_prev, const double* x_prev,
                   double mass, uint8_t fixed) {
    for (int k = 0; k < 3; k++) {
        c->x.push_back(x[k]);
        c->v.push_back(v ? v[k] : 0.0);
        c->v_prev.push_back(v_prev ? v_prev[k] : 0.0);
        c->x_prev.push_back(x_prev ? x_prev[k] : x[k]);
    }
    c->mass.push_back(mass);
    c->fixed.push_back(fixed);
    c->rot_of.push_back(-1);
    if (!c->f_ext.empty()) c->f_ext.insert(c->f_ext.end(), 3, 0.0);
}

// Append one oriented particle (ParticlePool::addOrientedParticle, common/ParticlePool.hpp:248-263).
uint32_t push_oriented(avbd_ctx* c, const double* x, const double* q, const double* v, const double* w, double mass,
                       const double* inertia, uint8_t fixed) {
    const uint32_t idx = (uint32_t)c->mass.size();
    push_particle(c, x, v, nullptr, nullptr, mass, fixed);
    c->rot_of[idx] = (int32_t)(c->q.size() / 4);
    const double qi[4] = {1, 0, 0, 0};
    for (int k = 0; k < 4; k++) c->q.push_back(q ? q[k] : qi[k]);
    for (int k = 0; k < 3; k++) {
        c->w.push_back(w ? w[k] : 0.0);
        c->inertia.push_back(inertia ? inertia[k] : 0.0);
    }
    return idx;
}

void require_built(avbd_ctx* c) {
    if (!c->built) throw CodeError(AVBD_E_NOT_BUILT, "avbd_build_graph() has not been called");
}

void require_not_built(avbd_ctx* c) {
    if (c->built) throw CodeError(AVBD_E_ARG, "pools are frozen after avbd_build_graph()");
}

void require_not_planned(avbd_ctx* c) {
    if (c->planned) throw CodeError(AVBD_E_ARG, "pools are frozen after avbd_build_graph()");
}

void check_capacity(avbd_ctx* c, size_t needed, uint32_t cap) {
    if (cap && needed > cap) throw CodeError(AVBD_E_CAPACITY, "Pool has exceeded capacity");
    (void)c;
}

// Upload particle state [first, first+n) from the FP64 staging arrays.
void upload_particle_range(avbd_ctx* c, uint32_t first, uint32_t n) {
    std::vector<float4> bx(n), bxp(n), bv(n), bvp(n);
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t p = first + i;
        bx[i] = f4(&c->x[3 * p], c->mass[p]);
        bxp[i] = f4(&c->x_prev[3 * p], 0.0);
        bv[i] = f4(&c->v[3 * p], 0.0);
        bvp[i] = f4(&c->v_prev[3 * p], 0.0);
    }
    auto cp = [&](float4* dst, const std::vector<float4>& src) {
        hip_check(hipMemcpyAsync(dst + first, src.data(), n * sizeof(float4), hipMemcpyHostToDevice, c->stream),
                  "hipMemcpyAsync H2D");
    };
    cp(c->d.x, bx);
    cp(c->d.x_prev, bxp);
    cp(c->d.v, bv);
    cp(c->d.v_prev, bvp);
    // ParticlePool::addParticle initialises y, x_last, x_lastlast to the position
    cp(c->d.y, bx);
    cp(c->d.x_last, bx);
    cp(c->d.x_lastlast, bx);
    if (c->fp64) {
        // the FP64 streams take the staged values exactly
        std::vector<double4> dx(n), dxp(n);
        for (uint32_t i = 0; i < n; i++) {
            const uint32_t p = first + i;
            dx[i] = make_double4(c->x[3 * p], c->x[3 * p + 1], c->x[3 * p + 2], c->mass[p]);
            dxp[i] = make_double4(c->x_prev[3 * p], c->x_prev[3 * p + 1], c->x_prev[3 * p + 2], 0.0);
        }
        auto cpd = [&](double4* dst, const std::vector<double4>& src) {
            hip_check(hipMemcpyAsync(dst + first, src.data(), n * sizeof(double4), hipMemcpyHostToDevice, c->stream),
                      "hipMemcpyAsync H2D");
        };
        cpd(c->f.x, dx);
        cpd(c->f.x_prev, dxp);
        cpd(c->f.y, dx);
        cpd(c->f.x_last, dx);
        cpd(c->f.x_lastlast, dx);
    }
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

// Upload orientation state of rotation slots [first, first+n): q and its
// history copies take the staged orientation, w and w_prev the staged
// angular velocity.
void upload_rotation_range(avbd_ctx* c, uint32_t first, uint32_t n) {
    if (n == 0) return;
    std::vector<float4> bq(n), bw(n), bi(n);
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t r = first + i;
        bq[i] = qf4(&c->q[4 * r]);
        bw[i] = f4(&c->w[3 * r], 0.0);
        bi[i] = f4(&c->inertia[3 * r], 0.0);
    }
    auto cp = [&](float4* dst, const std::vector<float4>& src) {
        hip_check(hipMemcpyAsync(dst + first, src.data(), n * sizeof(float4), hipMemcpyHostToDevice, c->stream),
                  "hipMemcpyAsync H2D");
    };
    for (float4* dst : {c->d.q, c->d.q_buf, c->d.q_in, c->d.q_prev, c->d.q_last, c->d.q_lastlast}) cp(dst, bq);
    cp(c->d.w, bw);
    cp(c->d.w_prev, bw);
    cp(c->d.inertia, bi);
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

// Device buffers of an LBVH over n primitives (bounds are written by the caller's kernel).
void alloc_lbvh(avbd_ctx* c, LbvhDev& b, uint32_t n) {
    const uint32_t nn = 2 * n - 1;
    b.n = n;
    b.aabb = dev_alloc<double>(c, 6 * (size_t)n);
    b.sbox = dev_alloc<double>(c, 6);
    b.keys = dev_alloc<uint64_t>(c, n);
    b.keys_sorted = dev_alloc<uint64_t>(c, n);
    b.sorted = dev_alloc<uint32_t>(c, n);
    b.left = dev_alloc<uint32_t>(c, nn);
    b.right = dev_alloc<uint32_t>(c, nn);
    b.parent = dev_alloc<uint32_t>(c, nn);
    b.range_hi = dev_alloc<uint32_t>(c, nn);
    b.bucket_count = dev_alloc<uint32_t>(c, 1u << 13);
    b.bucket_offset = dev_alloc<uint32_t>(c, (1u << 13) + 1);
    b.slots = dev_alloc<uint32_t>(c, n);
    b.table = dev_alloc<double>(c, 6 * (size_t)lbvh_table_entries(n));
    b.node_box = dev_alloc<double>(c, 6 * (size_t)nn);
    b.trav = dev_alloc<TravNode>(c, n > 1 ? n - 1 : 1);
    b.pair_cap = std::max<uint32_t>(1024u, 64u * n);
    b.pairs = dev_alloc<uint32_t>(c, 2 * (size_t)b.pair_cap);
    b.counters = dev_alloc<uint32_t>(c, 2);
    // left/right of leaves stay "none" (0xFFFFFFFF) for the whole run
    std::vector<uint32_t> none(nn, 0xFFFFFFFFu);
    upload(c, b.left, none);
    upload(c, b.right, none);
    upload(c, b.parent, none);
}

double omega_at(uint32_t i, double rho, double omega) {
    // Chebyshev schedule, simulation/VBDSolver.hpp:65-74
    if (i > 2) return 4 / (4 - rho * rho * omega);
    if (i == 2) return 2 / (2 - rho * rho);
    return 1;
}

// One time step (VBDSolver::solve_Parallel, simulation/VBDSolver.hpp:48-103) is
// enqueued as: the head (start-of-step work), iterations, the tail
// (end-of-step work).  In residual-targeted mode the iterations come in check
// intervals of res_check_every iterations, each followed by a residual check;
// the Chebyshev launches then read their weight from the device table at the
// step's iteration counter, so every interval is the same launch sequence.

// The broad phase that runs before the solve (CollisionDetector::
// detectCollisionsAndRecolor_Parallel).  Triangle-triangle pairs produce no
// contact energy (the reference's narrow phase for them is a no-op,
// collision/CollisionDetector.cpp:589-599), so nothing in this step's solve
// consumes the broad-phase output: after the bounds kernel has read the
// start-of-step state, the rest of the build and the traversal run on a
// second stream, concurrently with the solve, and join later in the step.
// (With dynamic contacts the detection round has already run, see detect_round.)
bool forks_broadphase(const avbd_ctx* c) { return c->lb.n && !c->dynamic && c->side_stream; }

void join_broadphase(avbd_ctx* c) {
    if (forks_broadphase(c)) hip_check(hipStreamWaitEvent(c->stream, c->ev_join, 0), "hipStreamWaitEvent");
}

// Converged-scene flags of a residual-targeted batch (none for one scene):
// a scene that meets the target stops iterating while the others go on, so
// each scene of a batch ends its step as it would on its own.
SceneFreeze scene_freeze(const avbd_ctx* c) {
    SceneFreeze fz;
    if (c->res_frozen) {
        fz.frozen = c->res_frozen;
        fz.iters = c->res_iters;
        fz.items = c->res_items;
        fz.scene_of = c->res_scene_of;
    }
    return fz;
}

// Dynamic contacts, but none active and the work list still the set-up one
// (no colour change, no conflict): the step is the static one.
bool dyn_quiet(const avbd_ctx* c) {
    return c->dynamic && c->cs.active.empty() && !c->layout_changed && !c->plan.conflict_offsets.empty() &&
           c->plan.conflict_offsets.back() == 0;
}

// The FP64 state with this round's dynamic-contact arrays (the adjacency
// buffers can move when a detection round grows them).
F64State f64_state(const avbd_ctx* c) {
    F64State f = c->f;
    if (c->dynamic) {
        f.dyn_off = c->d.dyn_off;
        f.dyn_ent = c->d.dyn_ent;
        f.dyn = c->d.dyn;
    }
    return f;
}

void enqueue_head(avbd_ctx* c, bool with_events) {
    const float dt = (float)c->desc.dt, g = (float)c->desc.g_accel;
    hipStream_t s = c->stream;
    const bool tri_bp = c->lb.n && !c->dynamic;
    const bool fork = forks_broadphase(c);
    if (tri_bp) {
        hipStream_t sb = fork ? c->side_stream : s;
        if (with_events) hip_check(hipEventRecord(c->ev_lb0, s), "hipEventRecord");
        launch_lbvh_bounds(s, c->lb, c->d.x, c->d.v, c->desc.dt);
        if (fork) {
            hip_check(hipEventRecord(c->ev_fork, s), "hipEventRecord");
            hip_check(hipStreamWaitEvent(sb, c->ev_fork, 0), "hipStreamWaitEvent");
        }
        launch_lbvh_tree(sb, c->lb);
        if (with_events) hip_check(hipEventRecord(c->ev_lb1, sb), "hipEventRecord");
        if (fork) hip_check(hipEventRecord(c->ev_join, sb), "hipEventRecord");
    }
    if (c->damping) {
        if (c->fp64) launch_eprev_f64(s, c->f, c->d.n_tets);
        else launch_eprev(s, c->d, dt);
    }
    // dynamic contacts' updateAfterTimeStep reads the start-of-step positions
    const uint32_t n_active = c->dynamic ? (uint32_t)c->cs.active.size() : 0u;
    if (n_active) {
        const uint32_t* act = static_cast<const uint32_t*>(c->g_active.p);
        if (c->fp64) launch_dyn_step_begin_f64(s, f64_state(c), act, n_active);
        else launch_dyn_step_begin(s, c->d, act, n_active);
    }
    if (c->fp64) {
        launch_step_begin_f64(s, c->f, c->d.n, c->desc.dt, c->desc.g_accel);
    } else {
        launch_joint2_step_begin(s, c->d);   // reads both bodies before the inertial update
        launch_step_begin(s, c->d, dt, g);
    }
    // learned warm start / recording: features, the network, the choice of the start
    if (c->warm || c->recording) {
        WarmIO io = c->wio;
        io.raw = c->recording ? c->d_raw : nullptr;
        io.tgt = c->d_tgt;
        io.norm = c->mlp->d_norm;
        if (c->warm) io.out = c->mlp->z32[kMlpLayers - 1];
        launch_warm_save(s, c->d, io);
        if (c->fp64) launch_warm_save_f64(s, c->f, c->d.n);
        launch_warm_features(s, c->d, io, c->cfg.lanes, dt, g, c->rows_pad);
        if (c->warm) {
            c->mlp->forward((int)c->rows_pad, io.x16, false);
            launch_warm_apply(s, c->d, io, dt * dt * g);
            if (c->fp64) launch_warm_ip_f64(s, c->f, io, c->damping, c->desc.dt);
            else launch_warm_ip(s, c->d, io, c->damping, dt);
            launch_warm_choose(s, c->d, io);
            if (c->fp64) launch_warm_place_f64(s, c->f, io.pred, io.scene_of, io.step_of);
        }
    }
    if (c->rc) launch_residual_reset(s, c->rc, scene_freeze(c), c->n_res_scenes);
    hip_check(hipGetLastError(), "kernel launch");
}

// Iterations [i0, i1) of a step.  With `check` (residual mode: i0 = 0,
// i1 = res_check_every) the Chebyshev weights come from the device table and
// a residual check follows the last iteration.
void enqueue_iters(avbd_ctx* c, bool with_events, uint32_t i0, uint32_t i1, bool check) {
    const float dt = (float)c->desc.dt;
    hipStream_t s = c->stream;
    const uint32_t* done = c->rc ? &c->rc->done : nullptr;
    const ResCtl* rc = c->rc;
    const uint32_t* d_active = static_cast<const uint32_t*>(c->g_active.p);
    const uint32_t n_active = c->dynamic ? (uint32_t)c->cs.active.size() : 0u;
    // a dynamic-contact step with no active contact on the set-up work list
    // is the static step: the static kernels, fused as at set-up
    const bool quiet = dyn_quiet(c);
    const bool fuse = c->fuse_cheb && (!c->dynamic || quiet);
    size_t ev = (size_t)2 * i0 * c->plan.num_colors;
    const uint32_t last = c->plan.num_colors - 1;
    // the step's state with the converged-scene flags of a batch
    const SceneFreeze fz = scene_freeze(c);
    DevState dz = c->d;
    F64State f0 = f64_state(c);
    dz.frozen = f0.frozen = fz.frozen;
    dz.scene_of = f0.scene_of = fz.scene_of;
    if (quiet) {
        dz.dyn_off = f0.dyn_off = nullptr;
        dz.dyn_ent = f0.dyn_ent = nullptr;
    }
    for (uint32_t i = i0; i < i1; i++) {
        const float omega = c->omega_seq[i];
        const uint32_t j = i - i0;
        const bool odd = (i & 1u) != 0;
        if (c->fp64) {
            F64State fi = f0;
            double4* x_next = c->f.x;
            if (fuse) {
                fi.x = odd ? c->f.x_alt : c->f.x;
                x_next = odd ? c->f.x : c->f.x_alt;
            }
            for (uint32_t col = 0; col < c->plan.num_colors; col++) {
                const uint32_t b = c->plan.item_offsets[col], e = c->plan.item_offsets[col + 1];
                if (with_events) hip_check(hipEventRecord(c->events[ev++], s), "hipEventRecord");
                if (fuse && col == last)
                    launch_solve_cheb_f64(s, fi, c->cfg, b, e - b, c->damping, c->desc.dt, omega, rc, c->omega_tab, j,
                                          c->cheb_list, c->n_cheb, fi.x, x_next);
                else
                    launch_solve_color_f64(s, fi, c->cfg, b, e - b, c->damping, c->desc.dt, fi.x, done);
                if (with_events) hip_check(hipEventRecord(c->events[ev++], s), "hipEventRecord");
                const uint32_t cb = c->plan.conflict_offsets[col], ce = c->plan.conflict_offsets[col + 1];
                if (ce > cb) launch_commit_f64(s, fi, c->d_conflicts + cb, ce - cb, done);
            }
            if (!fuse) launch_cheb_dual_f64(s, f0, c->d.n, omega, rc, c->omega_tab, j);
            if (n_active) launch_dyn_dual_f64(s, f0, d_active, n_active, done);
            continue;
        }
        // fused: iteration i works on x_cur and extrapolates into x_next
        DevState di = dz;
        float4* x_next = c->d.x;
        if (fuse) {
            di.x = odd ? c->x_alt : c->d.x;
            x_next = odd ? c->d.x : c->x_alt;
        }
        for (uint32_t col = 0; col < c->plan.num_colors; col++) {
            const uint32_t b = c->plan.item_offsets[col], e = c->plan.item_offsets[col + 1];
            if (with_events) hip_check(hipEventRecord(c->events[ev++], s), "hipEventRecord");
            if (fuse && col == last)
                launch_solve_cheb(s, di, c->cfg, b, e - b, c->damping, dt, omega, rc, c->omega_tab, j, c->cheb_list,
                                  c->n_cheb, x_next);
            else
                launch_solve_color(s, di, c->cfg, b, e - b, c->damping, dt, di.x, done);
            const uint32_t ob = c->plan.oitem_offsets[col], oe = c->plan.oitem_offsets[col + 1];
            launch_solve_ocolor(s, di, ob, oe - ob, dt, done);
            if (with_events) hip_check(hipEventRecord(c->events[ev++], s), "hipEventRecord");
            const uint32_t cb = c->plan.conflict_offsets[col], ce = c->plan.conflict_offsets[col + 1];
            if (ce > cb) launch_commit(s, di, c->d_conflicts + cb, ce - cb, done);
        }
        if (!fuse) launch_cheb_dual(s, dz, omega, rc, c->omega_tab, j);
        launch_joint2_dual(s, dz, done);   // oriented particles: never fused
        if (n_active) launch_dyn_dual(s, c->d, d_active, n_active, done);
    }
    if (check) {
        // an interval has an even number of iterations whenever the step is
        // fused (see avbd_build_graph), so the iterate is back in x
        if (c->fp64) launch_residual_f64(s, f0, c->cfg, c->damping, c->desc.dt, c->f.x, c->res_rr, c->rc);
        else launch_residual_pass(s, dz, c->cfg, c->damping, dt, c->rc, c->res_rr);
        launch_residual_decide(s, c->rc, c->res_rr, c->res_scene_off, c->n_res_scenes, c->res_acc_scene,
                               c->res_mg2_scene, c->res_tol * c->res_tol, c->res_check_every, fz);
    }
    hip_check(hipGetLastError(), "kernel launch");
}

void enqueue_tail(avbd_ctx* c) {
    const float dt = (float)c->desc.dt, g = (float)c->desc.g_accel;
    hipStream_t s = c->stream;
    const uint32_t iters = c->iters();
    if (c->rc)
        launch_residual_finish(s, c->rc, scene_freeze(c), c->n_res_scenes, c->d.n_items + c->d.n_oitems);
    // fused: the final positions are in x_alt after an odd number of
    // iterations (never in residual mode, whose intervals are even)
    if (c->fuse_cheb && (!c->dynamic || dyn_quiet(c)) && !c->rc && (iters & 1u)) {
        if (c->fp64) launch_settle_f64(s, c->f.x, c->f.x_alt, c->d.n, nullptr, iters);
        else launch_settle(s, c->d.x, c->x_alt, c->d.n, nullptr, iters);
    }
    if (c->recording) {
        if (c->fp64) launch_warm_target_f64(s, c->f, c->d_tgt, 1.0 / (c->desc.dt * c->desc.dt * c->desc.g_accel));
        else {
            WarmIO io = c->wio;
            io.tgt = c->d_tgt;
            launch_warm_target(s, c->d, io, 1.0f / (dt * dt * g));
        }
    }
    if (c->fp64) launch_velocity_f64(s, c->f, c->d.n, c->desc.dt, c->d.status);
    else launch_velocity(s, c->d, dt);
    hip_check(hipGetLastError(), "kernel launch");
}

// One whole step at the fixed iteration count (or, in residual mode, the
// eager form: every interval with its check, the done flag turning the rest
// into no-ops).
void enqueue_step(avbd_ctx* c, bool with_events) {
    enqueue_head(c, with_events);
    if (c->rc) {
        for (uint32_t k = 0; k < c->iters() / c->res_check_every; k++)
            enqueue_iters(c, false, 0, c->res_check_every, true);
    } else {
        enqueue_iters(c, with_events, 0, c->iters(), false);
    }
    enqueue_tail(c);
    join_broadphase(c);
}

void destroy_graph(avbd_ctx* c) {
    for (hipGraphExec_t* e : {&c->graph_exec, &c->head_exec, &c->interval_exec, &c->tail_exec}) {
        if (*e) (void)hipGraphExecDestroy(*e);
        *e = nullptr;
    }
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graph = nullptr;
}

template <class F>
hipGraphExec_t capture(avbd_ctx* c, F&& body) {
    hip_check(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
        body();
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(c->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    hipGraph_t g = nullptr;
    hip_check(hipStreamEndCapture(c->stream, &g), "hipStreamEndCapture");
    hipGraphExec_t e = nullptr;
    const hipError_t r = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    hip_check(r, "hipGraphInstantiate");
    return e;
}

// Residual mode replays three graphs per step: the head with the first check
// interval (the concurrent broad phase joins there), one interval replayed as
// often as the host decides (avbd_step), and the tail.
void capture_graph(avbd_ctx* c) {
    destroy_graph(c);
    c->dyn_graph_key.clear();
    // dynamic contacts: the fixed-count step launches eagerly after its
    // detection round; residual-mode graphs are captured per detection round
    // when its layout differs from the captured one (dyn_residual_graphs)
    if (c->dynamic) return;
    if (c->rc) {
        const uint32_t every = c->res_check_every;
        c->head_exec = capture(c, [&] {
            enqueue_head(c, false);
            enqueue_iters(c, false, 0, every, true);
            join_broadphase(c);
        });
        c->interval_exec = capture(c, [&] { enqueue_iters(c, false, 0, every, true); });
        c->tail_exec = capture(c, [&] { enqueue_tail(c); });
        return;
    }
    c->graph_exec = capture(c, [&] { enqueue_step(c, false); });
}

// Residual mode with dynamic contacts: the head / interval / tail graphs bake
// in the colour-phase and conflict launch sizes, the number of active
// contacts and the dynamic-adjacency buffers; capture them again when a
// detection round changed any of these (the cube of the headline keeps one
// layout for every step while it has no self contact).
void dyn_residual_graphs(avbd_ctx* c) {
    const ColorPlan& p = c->plan;
    std::vector<uint64_t> key;
    key.push_back(p.num_colors);
    key.insert(key.end(), p.item_offsets.begin(), p.item_offsets.end());
    key.insert(key.end(), p.oitem_offsets.begin(), p.oitem_offsets.end());
    key.insert(key.end(), p.conflict_offsets.begin(), p.conflict_offsets.end());
    key.push_back(c->cs.active.size());
    key.push_back(dyn_quiet(c) ? 1u : 0u);
    for (const void* ptr : {(const void*)c->g_active.p, (const void*)c->d.dyn_off, (const void*)c->d.dyn_ent})
        key.push_back((uint64_t)(uintptr_t)ptr);
    if (c->head_exec && key == c->dyn_graph_key) return;
    destroy_graph(c);
    const uint32_t every = c->res_check_every;
    c->head_exec = capture(c, [&] {
        enqueue_head(c, false);
        enqueue_iters(c, false, 0, every, true);
    });
    c->interval_exec = capture(c, [&] { enqueue_iters(c, false, 0, every, true); });
    c->tail_exec = capture(c, [&] { enqueue_tail(c); });
    c->dyn_graph_key = std::move(key);
}

// Residual mode, one step: interval k is queued once the check of interval
// k - kAhead has been seen unconverged, so at most kAhead - 1 intervals of
// early-exiting launches follow the converged check instead of the rest of
// the iteration budget.  Each interval's converged flag is copied to pinned
// host memory behind the interval (a stream-ordered copy), then an event is
// recorded.  `graphs` replays the captured graphs; otherwise the same
// sequence is launched eagerly.
void run_residual_step(avbd_ctx* c, bool graphs) {
    constexpr uint32_t kAhead = 2;
    const uint32_t intervals = c->iters() / c->res_check_every;
    hipStream_t s = c->stream;
    for (uint32_t k = 0; k < intervals; k++) {
        if (k >= kAhead) {
            hip_check(hipEventSynchronize(c->part_ev[k - kAhead]), "hipEventSynchronize");
            if (c->h_conv[k - kAhead]) break;
        }
        if (k == 0) {
            if (graphs) hip_check(hipGraphLaunch(c->head_exec, s), "hipGraphLaunch");
            else {
                enqueue_head(c, false);
                enqueue_iters(c, false, 0, c->res_check_every, true);
                join_broadphase(c);
            }
        } else {
            if (graphs) hip_check(hipGraphLaunch(c->interval_exec, s), "hipGraphLaunch");
            else enqueue_iters(c, false, 0, c->res_check_every, true);
        }
        hip_check(hipMemcpyAsync(c->h_conv + k, &c->rc->done, sizeof(uint32_t), hipMemcpyDeviceToHost, s),
                  "hipMemcpyAsync D2H");
        hip_check(hipEventRecord(c->part_ev[k], s), "hipEventRecord");
    }
    if (graphs) hip_check(hipGraphLaunch(c->tail_exec, s), "hipGraphLaunch");
    else enqueue_tail(c);
}

// Pinned staging of a detection round's host arrays: reserve the round's
// total first (may reallocate: the stream is idle then), then each array is
// copied in and sent with a truly asynchronous H2D copy.
void stage_reserve(avbd_ctx* c, size_t bytes) {
    c->stage_used = 0;
    if (bytes <= c->stage_cap) return;
    if (c->stage) hip_check(hipHostFree(c->stage), "hipHostFree");
    c->stage = nullptr;
    c->stage_cap = 0;
    void* p = nullptr;
    const size_t cap = bytes + bytes / 2 + 4096;
    hip_check(hipHostMalloc(&p, cap, hipHostMallocDefault), "hipHostMalloc");
    c->stage = static_cast<uint8_t*>(p);
    c->stage_cap = cap;
}

constexpr size_t kStageAlign = 256;
inline size_t stage_bytes(size_t bytes) { return (bytes + kStageAlign - 1) / kStageAlign * kStageAlign; }

template <class T>
void stage_upload(avbd_ctx* c, T* dst, const std::vector<T>& src) {
    if (src.empty()) return;
    const size_t bytes = src.size() * sizeof(T);
    if (c->stage_used + stage_bytes(bytes) > c->stage_cap) throw std::logic_error("staging buffer overrun");
    uint8_t* p = c->stage + c->stage_used;
    std::memcpy(p, src.data(), bytes);
    c->stage_used += stage_bytes(bytes);
    hip_check(hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, c->stream), "hipMemcpyAsync H2D");
}

// One round of CollisionDetector::detectCollisionsAndRecolor_Parallel
// (collision/CollisionDetector.cpp:57-129) at the start of a step: device
// bounds, LBVH and narrow phase; the detected list comes back to the host
// for the lifecycle, the dynamic adjacency and the recolouring; the
// resulting record ops, adjacency and (when the colouring changed) work
// items go back before the solve.
// The per-colour offsets of the work list the device laid out, and the
// colour count (synchronising): the host sizes the colour-phase launches by
// them.
void read_layout(avbd_ctx* c) {
    hipStream_t s = c->stream;
    hip_check(hipMemcpyAsync(c->h_layout, c->rdev.offsets, 3 * 257 * sizeof(uint32_t), hipMemcpyDeviceToHost, s),
              "D2H");
    hip_check(hipMemcpyAsync(c->h_layout + 3 * 257, c->rdev.counters, 5 * sizeof(uint32_t), hipMemcpyDeviceToHost, s),
              "D2H");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    const uint32_t* cnt = c->h_layout + 3 * 257;
    if (cnt[3] & 1u) throw CodeError(AVBD_E_COLORS, "Ran out of colors!");
    if (cnt[4]) c->layout_changed = true;
    const uint32_t K = cnt[2];
    ColorPlan& plan = c->plan;
    plan.num_colors = K;
    plan.item_offsets.assign(c->h_layout, c->h_layout + K + 1);
    plan.oitem_offsets.assign(c->h_layout + 257, c->h_layout + 257 + K + 1);
    plan.conflict_offsets.assign(c->h_layout + 514, c->h_layout + 514 + K + 1);
    // per-launch timing events (fixed iteration counts only; residual mode
    // runs without them)
    const size_t need = c->rc ? 0 : 2 * (size_t)c->iters() * K;
    while (c->events.size() < need) {
        hipEvent_t ev;
        hip_check(hipEventCreate(&ev), "hipEventCreate");
        c->events.push_back(ev);
    }
}

void detect_round(avbd_ctx* c) {
    hipStream_t s = c->stream;
    const double dt = c->desc.dt;
    hip_check(hipMemsetAsync(c->d_status, 0, 2 * sizeof(uint32_t), s), "hipMemsetAsync");
    launch_prim_bounds(s, c->dp, c->lbd, c->d, dt);
    launch_lbvh_tree(s, c->lbd);
    launch_narrow(s, c->dp, c->lbd, c->d, dt, c->cs.margin, c->d_det, c->det_cap, c->d_status);
    hip_check(hipGetLastError(), "kernel launch");
    uint32_t st[4] = {0, 0, 0, 0};
    hip_check(hipMemcpyAsync(st, c->d_status, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s), "D2H");
    hip_check(hipMemcpyAsync(st + 2, c->lbd.counters, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s), "D2H");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (st[2] > c->lbd.pair_cap || (st[3] & 1u))
        throw CodeError(AVBD_E_CAPACITY, "broad phase: candidate pairs exceeded capacity");
    if (st[0] & NARROW_UNSUPPORTED) throw CodeError(AVBD_E_UNSUPPORTED_CONTACT, "Unsupported collision!");
    if (st[1] > c->det_cap) throw CodeError(AVBD_E_CAPACITY, "narrow phase: detected contacts exceeded capacity");
    std::vector<DetRec> det(st[1]);
    if (!det.empty()) {
        hip_check(hipMemcpyAsync(det.data(), c->d_det, det.size() * sizeof(DetRec), hipMemcpyDeviceToHost, s), "D2H");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    }
    c->cs.process(det, c->prims, c->mass, dt);
    ContactSystem& cs = c->cs;
    // the stream is idle here (synchronised above): reserve the staging for
    // every upload of this round, then queue them; the recolouring and the
    // solve follow on the same stream
    const size_t total = stage_bytes(cs.ops.size() * sizeof(DynOp)) + stage_bytes(cs.active.size() * sizeof(uint32_t)) +
                         stage_bytes(cs.dyn_off.size() * sizeof(uint32_t)) +
                         stage_bytes(cs.dyn_ent.size() * sizeof(uint32_t)) +
                         stage_bytes(cs.nb_off.size() * sizeof(uint32_t)) + stage_bytes(cs.nb.size() * sizeof(uint32_t)) +
                         stage_bytes(cs.seeds.size() * sizeof(uint32_t));
    stage_reserve(c, total);
    if (!cs.ops.empty()) {
        auto* d_ops = static_cast<DynOp*>(grow(c, c->g_ops, cs.ops.size() * sizeof(DynOp)));
        stage_upload(c, d_ops, cs.ops);
        launch_dyn_apply(s, c->d, d_ops, (uint32_t)cs.ops.size());
    }
    auto* d_act = static_cast<uint32_t*>(grow(c, c->g_active, cs.active.size() * sizeof(uint32_t)));
    stage_upload(c, d_act, cs.active);
    auto* d_off = static_cast<uint32_t*>(grow(c, c->g_off, cs.dyn_off.size() * sizeof(uint32_t)));
    stage_upload(c, d_off, cs.dyn_off);
    auto* d_ent = static_cast<uint32_t*>(grow(c, c->g_ent, cs.dyn_ent.size() * sizeof(uint32_t)));
    stage_upload(c, d_ent, cs.dyn_ent);
    c->d.dyn_off = d_off;
    c->d.dyn_ent = d_ent;
    // incremental recolouring and the new work list, on the device
    auto* d_nboff = static_cast<uint32_t*>(grow(c, c->g_nboff, cs.nb_off.size() * sizeof(uint32_t)));
    stage_upload(c, d_nboff, cs.nb_off);
    auto* d_nb = static_cast<uint32_t*>(grow(c, c->g_nb, std::max<size_t>(cs.nb.size(), 1) * sizeof(uint32_t)));
    stage_upload(c, d_nb, cs.nb);
    auto* d_seeds = static_cast<uint32_t*>(grow(c, c->g_seeds, std::max<size_t>(cs.seeds.size(), 1) * sizeof(uint32_t)));
    stage_upload(c, d_seeds, cs.seeds);
    c->rdev.d_off = d_nboff;
    c->rdev.d_nb = d_nb;
    // no new contact, nothing left dirty or conflicted by the last round: the
    // recolouring would change nothing (ColorList only recolours from dirty
    // particles), so the colours, the work list and its offsets stand
    const uint32_t* last = c->h_layout + 3 * 257;
    if (cs.seeds.empty() && last[0] == 0 && last[1] == 0) {
        hip_check(hipGetLastError(), "kernel launch");
        return;
    }
    launch_recolour(s, c->rdev, d_seeds, (uint32_t)cs.seeds.size());
    launch_relayout(s, c->rdev, c->d.items, c->d.oitems, c->d_conflicts, false);
    hip_check(hipGetLastError(), "kernel launch");
    read_layout(c);
}

// Autotuner: time the colour-phase solve of one iteration for each candidate
// (lanes per vertex, block size), writing into a scratch sink so that the
// state is untouched, and keep the fastest.
// Temporary device buffers, freed when the holder goes out of scope.
struct DevTemp {
    std::vector<void*> ptrs;
    ~DevTemp() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    T* get(size_t count) {
        void* p = nullptr;
        hip_check(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)), "hipMalloc");
        ptrs.push_back(p);
        return static_cast<T*>(p);
    }
};

// DSATUR of every scene on the device (ColorList::greedyColor2,
// common/ColorList.hpp:149-241): the static adjacency goes up, the colours
// come back for the work-list layout.
void device_dsatur(avbd_ctx* c, const std::vector<uint32_t>& scene_off, const std::vector<uint32_t>& off,
                   const std::vector<uint32_t>& nb, std::vector<uint32_t>& color, std::vector<uint32_t>& scene_colors) {
    const uint32_t n = (uint32_t)color.size(), ns = (uint32_t)scene_off.size() - 1;
    hipStream_t s = c->stream;
    DevTemp t;
    uint32_t* d_so = t.get<uint32_t>(scene_off.size());
    uint32_t* d_off = t.get<uint32_t>(off.size());
    uint32_t* d_nb = t.get<uint32_t>(nb.size());
    uint64_t* d_key = t.get<uint64_t>(n);
    uint64_t* d_seen = t.get<uint64_t>(4 * (size_t)n);
    uint32_t* d_color = t.get<uint32_t>(n);
    uint32_t* d_sc = t.get<uint32_t>(ns);
    uint32_t* d_status = t.get<uint32_t>(1);
    auto h2d = [&](void* dst, const void* src, size_t bytes) {
        if (bytes) hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
    };
    h2d(d_so, scene_off.data(), scene_off.size() * sizeof(uint32_t));
    h2d(d_off, off.data(), off.size() * sizeof(uint32_t));
    h2d(d_nb, nb.data(), nb.size() * sizeof(uint32_t));
    hip_check(hipMemsetAsync(d_seen, 0, 4 * (size_t)n * sizeof(uint64_t), s), "hipMemsetAsync");
    hip_check(hipMemsetAsync(d_status, 0, sizeof(uint32_t), s), "hipMemsetAsync");
    launch_dsatur(s, ns, d_so, d_off, d_nb, d_key, d_seen, d_color, d_sc, d_status);
    hip_check(hipGetLastError(), "dsatur launch");
    uint32_t status = 0;
    hip_check(hipMemcpyAsync(color.data(), d_color, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost, s), "D2H");
    hip_check(hipMemcpyAsync(scene_colors.data(), d_sc, (size_t)ns * sizeof(uint32_t), hipMemcpyDeviceToHost, s),
              "D2H");
    hip_check(hipMemcpyAsync(&status, d_status, sizeof(uint32_t), hipMemcpyDeviceToHost, s), "D2H");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (status & 1u) throw std::runtime_error("Ran out of colors!");
    if (status & 2u) throw std::runtime_error("a vertex has more than 65535 neighbours");
}

// One colour phase of vertex solves into the scratch output (state untouched).
void solve_into_scratch(avbd_ctx* c, const LaunchCfg& cfg, uint32_t b, uint32_t count) {
    if (c->fp64)
        // the variant the next step launches: contact-aware unless the step is quiet
        launch_solve_color_f64(c->stream, dyn_quiet(c) ? c->f : f64_state(c), cfg, b, count, c->damping, c->desc.dt,
                               c->f.x_scratch, nullptr);
    else {
        DevState d = c->d;
        if (dyn_quiet(c)) d.dyn_off = d.dyn_ent = nullptr;
        launch_solve_color(c->stream, d, cfg, b, count, c->damping, (float)c->desc.dt, c->d.x_scratch);
    }
}

// Lanes per vertex, block size and gather batch: time every candidate on
// the scene's colour phases and keep the fastest.
void autotune(avbd_ctx* c) {
    const uint32_t lanes_c[] = {4, 8, 16, 32};
    const uint32_t block_c[] = {64, 128, 256};
    const uint32_t batch_c[] = {1, 2, 4};
    float best = 1e30f;
    LaunchCfg best_cfg = c->cfg;
    for (uint32_t L : lanes_c) {
        if (c->desc.lanes_per_vertex && L != c->desc.lanes_per_vertex) continue;
        for (uint32_t B : block_c)
        for (uint32_t U : batch_c) {
            if (c->desc.block_size && B != c->desc.block_size) continue;
            if (c->desc.gather_batch && U != c->desc.gather_batch) continue;
            LaunchCfg cfg{L, B, U};
            auto run = [&] {
                for (uint32_t col = 0; col < c->plan.num_colors; col++) {
                    const uint32_t b = c->plan.item_offsets[col], e = c->plan.item_offsets[col + 1];
                    solve_into_scratch(c, cfg, b, e - b);
                }
            };
            for (int w = 0; w < 3; w++) run();
            hip_check(hipEventRecord(c->ev_a, c->stream), "hipEventRecord");
            for (int r = 0; r < 10; r++) run();
            hip_check(hipEventRecord(c->ev_b, c->stream), "hipEventRecord");
            hip_check(hipEventSynchronize(c->ev_b), "hipEventSynchronize");
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, c->ev_a, c->ev_b), "hipEventElapsedTime");
            if (ms < best) {
                best = ms;
                best_cfg = cfg;
            }
        }
    }
    c->cfg = best_cfg;
}

// Fused or separate Chebyshev / dual update: time the last colour phase plus
// the update both ways on the current state and keep the faster.  Both
// forms write x_last, x_lastlast and the ground records (and the separate
// one x), so those are saved first and restored afterwards.
void tune_fusion(avbd_ctx* c) {
    if (!c->fuse_cheb) return;
    const uint32_t last = c->plan.num_colors - 1;
    const uint32_t b = c->plan.item_offsets[last], e = c->plan.item_offsets[last + 1];
    const float dt = (float)c->desc.dt;
    const size_t nx = (size_t)c->d.n * (c->fp64 ? sizeof(double4) : sizeof(float4));
    const size_t ng = (size_t)c->d.n_ground * sizeof(GroundRec);
    void* px = c->fp64 ? (void*)c->f.x : (void*)c->d.x;
    void* pl = c->fp64 ? (void*)c->f.x_last : (void*)c->d.x_last;
    void* pll = c->fp64 ? (void*)c->f.x_lastlast : (void*)c->d.x_lastlast;
    std::vector<uint8_t> save(3 * nx + ng);
    auto d2h = [&](void* dst, const void* src, size_t bytes) {
        if (bytes) hip_check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "D2H");
    };
    auto h2d = [&](void* dst, const void* src, size_t bytes) {
        if (bytes) hip_check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "H2D");
    };
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    d2h(save.data(), px, nx);
    d2h(save.data() + nx, pl, nx);
    d2h(save.data() + 2 * nx, pll, nx);
    d2h(save.data() + 3 * nx, c->d.ground, ng);
    auto timed = [&](auto&& body) {
        for (int w = 0; w < 3; w++) body();
        hip_check(hipEventRecord(c->ev_a, c->stream), "hipEventRecord");
        for (int r = 0; r < 10; r++) body();
        hip_check(hipEventRecord(c->ev_b, c->stream), "hipEventRecord");
        hip_check(hipEventSynchronize(c->ev_b), "hipEventSynchronize");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, c->ev_a, c->ev_b), "hipEventElapsedTime");
        return ms;
    };
    float t_sep = 0, t_fused = 0;
    if (c->fp64) {
        t_sep = timed([&] {
            solve_into_scratch(c, c->cfg, b, e - b);
            launch_cheb_dual_f64(c->stream, c->f, c->d.n, 1.0f, nullptr, nullptr, 0);
        });
        t_fused = timed([&] {
            launch_solve_cheb_f64(c->stream, c->f, c->cfg, b, e - b, c->damping, c->desc.dt, 1.0f, nullptr, nullptr, 0,
                                  c->cheb_list, c->n_cheb, c->f.x, c->f.x_alt);
        });
    } else {
        t_sep = timed([&] {
            solve_into_scratch(c, c->cfg, b, e - b);
            launch_cheb_dual(c->stream, c->d, 1.0f);
        });
        t_fused = timed([&] {
            launch_solve_cheb(c->stream, c->d, c->cfg, b, e - b, c->damping, dt, 1.0f, nullptr, nullptr, 0,
                              c->cheb_list, c->n_cheb, c->x_alt);
        });
    }
    h2d(px, save.data(), nx);
    h2d(pl, save.data() + nx, nx);
    h2d(pll, save.data() + 2 * nx, nx);
    h2d(c->d.ground, save.data() + 3 * nx, ng);
    c->fuse_cheb = t_fused < t_sep;
}

}  // namespace

namespace avbd {
void set_context_free_error(const std::string& msg) { g_create_error = msg; }
}  // namespace avbd

extern "C" {

int avbd_create(const avbd_scene_desc* desc, avbd_ctx** out) {
    if (!desc || !out) {
        g_create_error = "null argument";
        return AVBD_E_ARG;
    }
    *out = nullptr;
    if (!(desc->dt > 0) || desc->solver_iters == 0) {
        g_create_error = "dt must be > 0 and solver_iters >= 1";
        return AVBD_E_ARG;
    }
    const uint32_t L = desc->lanes_per_vertex, B = desc->block_size;
    if (L && L != 4 && L != 8 && L != 16 && L != 32) {
        g_create_error = "lanes_per_vertex must be 0, 4, 8, 16 or 32";
        return AVBD_E_ARG;
    }
    if (B && (B < 64 || B > 1024 || B % 64)) {
        g_create_error = "block_size must be 0 or a multiple of 64 in [64, 1024]";
        return AVBD_E_ARG;
    }
    const uint32_t U = desc->gather_batch;
    if (U && U != 1 && U != 2 && U != 4) {
        g_create_error = "gather_batch must be 0, 1, 2 or 4";
        return AVBD_E_ARG;
    }
    if (desc->flags & AVBD_FLAG_PERSISTENT) {
        g_create_error = "AVBD_FLAG_PERSISTENT is no longer supported: one launch per solver phase, replayed from a "
                         "hipGraph, is faster on MI355X than one persistent launch with grid barriers";
        return AVBD_E_ARG;
    }
    // HIP is initialised lazily by avbd_build_graph(): staging and the host
    // graph build work without a device.
    auto* c = new avbd_ctx();
    c->desc = *desc;
    c->fp64 = (desc->flags & AVBD_FLAG_FP64_STATE) != 0;
    c->device = desc->device;
    c->cfg.lanes = L ? L : 8;
    c->cfg.block = B ? B : 256;
    c->cfg.batch = U ? U : 1;
    *out = c;
    return AVBD_OK;
}

void avbd_destroy(avbd_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    destroy_graph(c);
    c->mlp.reset();
    for (void* p : c->allocations) (void)hipFree(p);
    for (auto ev : c->events) (void)hipEventDestroy(ev);
    for (auto ev : c->part_ev) (void)hipEventDestroy(ev);
    if (c->h_conv) (void)hipHostFree(c->h_conv);
    if (c->h_layout) (void)hipHostFree(c->h_layout);
    if (c->stage) (void)hipHostFree(c->stage);
    if (c->ev_lb0) (void)hipEventDestroy(c->ev_lb0);
    if (c->ev_lb1) (void)hipEventDestroy(c->ev_lb1);
    if (c->ev_a) (void)hipEventDestroy(c->ev_a);
    if (c->ev_b) (void)hipEventDestroy(c->ev_b);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->side_stream) (void)hipStreamDestroy(c->side_stream);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* avbd_last_error(avbd_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int avbd_add_particles(avbd_ctx* c, uint32_t n, const double* x, const double* v, const double* v_prev,
                       const double* x_prev, const double* mass, const uint8_t* fixed, uint32_t* first) {
    return guarded(c, [&] {
        require_not_planned(c);
        if (n && !x) throw std::runtime_error("x is required");
        const size_t base = c->mass.size();
        check_capacity(c, base + n, c->desc.particle_capacity);
        for (uint32_t i = 0; i < n; i++)
            push_particle(c, x + 3 * i, v ? v + 3 * i : nullptr, v_prev ? v_prev + 3 * i : nullptr,
                          x_prev ? x_prev + 3 * i : nullptr, mass ? mass[i] : 0.0, fixed ? fixed[i] : 0);
        if (first) *first = (uint32_t)base;
    });
}

int avbd_add_energies(avbd_ctx* c, int type, const void* records, uint32_t count) {
    return guarded(c, [&] {
        require_not_planned(c);
        if (count && !records) throw std::runtime_error("records is null");
        if (type == AVBD_NEO_HOOKEAN) {
            check_capacity(c, c->nh.size() + count, c->desc.energy_capacity);
            const auto* r = static_cast<const avbd_neo_hookean_record*>(records);
            for (uint32_t i = 0; i < count; i++) {
                HostNH h;
                std::memcpy(h.idx, r[i].particle_indices, sizeof(h.idx));
                h.lambda = r[i].lambda;
                h.mu = r[i].mu;
                h.kd = r[i].kd;
                std::memcpy(h.Q, r[i].Q, sizeof(h.Q));
                h.V = r[i].rest_volume;
                std::memcpy(h.E_prev, r[i].E_prev, sizeof(h.E_prev));
                c->nh.push_back(h);
            }
        } else if (type == AVBD_GROUND_COLLISION) {
            check_capacity(c, c->ground.size() + count, c->desc.energy_capacity);
            const auto* r = static_cast<const avbd_ground_record*>(records);
            for (uint32_t i = 0; i < count; i++) {
                HostGround g;
                g.p = r[i].particle_index;
                g.k_start = r[i].k_start;
                g.k_n = r[i].k_n;
                g.k_t = r[i].k_t;
                g.k_b = r[i].k_b;
                g.lambda_n = r[i].lambda_n;
                g.lambda_t = r[i].lambda_t;
                g.lambda_b = r[i].lambda_b;
                g.C_n_prev = r[i].C_n_prev;
                g.mu_s = r[i].mu_s;
                g.mu_k = r[i].mu_k;
                g.cp_x = r[i].cp_x;
                g.cp_z = r[i].cp_z;
                g.use_static = r[i].use_static != 0;
                g.unclamped = r[i].unclamped_lambda_tb != 0;
                c->ground.push_back(g);
            }
        } else if (type == AVBD_COSSERAT_ROD) {
            check_capacity(c, c->cos.size() + count, c->desc.energy_capacity);
            const auto* r = static_cast<const avbd_cosserat_record*>(records);
            for (uint32_t i = 0; i < count; i++) {
                HostCos h{};
                h.idx[0] = r[i].particle_indices[0];
                h.idx[1] = r[i].particle_indices[1];
                std::memcpy(h.stiffness, r[i].stiffness, sizeof(h.stiffness));
                h.L = r[i].rest_length;
                std::memcpy(h.kappa0, r[i].rest_curvature, sizeof(h.kappa0));
                if (!(h.L > 0)) throw std::runtime_error("Cosserat rest length must be > 0");
                c->cos.push_back(h);
            }
        } else if (type == AVBD_RIGID_BODY_GROUND_COLLISION) {
            check_capacity(c, c->rground.size() + count, c->desc.energy_capacity);
            const auto* r = static_cast<const avbd_rigid_ground_record*>(records);
            for (uint32_t i = 0; i < count; i++) {
                HostRGround h{};
                h.p = r[i].particle_index;
                h.shape = r[i].shape;
                h.radius = r[i].radius;
                std::memcpy(h.cp_rb_local, r[i].cp_rb_local, sizeof(h.cp_rb_local));
                std::memcpy(h.cp_ground, r[i].cp_ground, sizeof(h.cp_ground));
                h.k_start = r[i].k_start;
                h.k_n = r[i].k_n;
                h.k_t = r[i].k_t;
                h.k_b = r[i].k_b;
                h.lambda_n = r[i].lambda_n;
                h.lambda_t = r[i].lambda_t;
                h.lambda_b = r[i].lambda_b;
                h.C_n_prev = r[i].C_n_prev;
                h.mu_s = r[i].mu_s;
                h.mu_k = r[i].mu_k;
                h.use_static = r[i].use_static != 0;
                h.unclamped = r[i].unclamped_lambda_tb != 0;
                if (h.shape < AVBD_SHAPE_ROD_SEGMENT || h.shape > AVBD_SHAPE_CAPSULE)
                    throw std::runtime_error("rigid-ground shape must be a rod segment, sphere, box or capsule");
                c->rground.push_back(h);
            }
        } else if (type == AVBD_ONE_SIDED_FIXED_JOINT) {
            check_capacity(c, c->joints.size() + count, c->desc.energy_capacity);
            const auto* r = static_cast<const avbd_joint_record*>(records);
            for (uint32_t i = 0; i < count; i++) {
                HostJoint h{};
                h.p = r[i].particle_index;
                std::memcpy(h.pos_offset, r[i].position_offset, sizeof(h.pos_offset));
                std::memcpy(h.rot_offset, r[i].rotation_offset, sizeof(h.rot_offset));
                std::memcpy(h.ref_position, r[i].ref_position, sizeof(h.ref_position));
                std::memcpy(h.ref_rotation, r[i].ref_rotation, sizeof(h.ref_rotation));
                std::memcpy(h.k, r[i].k, sizeof(h.k));
                std::memcpy(h.lambda, r[i].lambda, sizeof(h.lambda));
                std::memcpy(h.C_prev, r[i].C_prev, sizeof(h.C_prev));
                c->joints.push_back(h);
            }
        } else if (type == AVBD_FIXED_JOINT) {
            check_capacity(c, c->joints2.size() + count, c->desc.energy_capacity);
            const auto* r = static_cast<const avbd_fixed_joint_record*>(records);
            for (uint32_t i = 0; i < count; i++) {
                HostJoint2 h{};
                std::memcpy(h.p, r[i].particle_index, sizeof(h.p));
                std::memcpy(h.pos_offset, r[i].position_offset, sizeof(h.pos_offset));
                std::memcpy(h.rot_offset, r[i].rotation_offset, sizeof(h.rot_offset));
                std::memcpy(h.k, r[i].k, sizeof(h.k));
                std::memcpy(h.lambda, r[i].lambda, sizeof(h.lambda));
                std::memcpy(h.C_prev, r[i].C_prev, sizeof(h.C_prev));
                c->joints2.push_back(h);
            }
        } else {
            throw CodeError(AVBD_E_ARG, "energy type not supported by this build");
        }
    });
}

int avbd_add_oriented_particles(avbd_ctx* c, uint32_t n, const double* x, const double* q, const double* v,
                                const double* w, const double* mass, const double* inertia, const uint8_t* fixed,
                                uint32_t* first) {
    return guarded(c, [&] {
        require_not_planned(c);
        if (n && (!x || !inertia)) throw std::runtime_error("x and inertia are required");
        const size_t base = c->mass.size();
        check_capacity(c, base + n, c->desc.particle_capacity);
        for (uint32_t i = 0; i < n; i++)
            push_oriented(c, x + 3 * i, q ? q + 4 * i : nullptr, v ? v + 3 * i : nullptr, w ? w + 3 * i : nullptr,
                          mass ? mass[i] : 0.0, inertia + 3 * i, fixed ? fixed[i] : 0);
        if (first) *first = (uint32_t)base;
    });
}

namespace {
// Hamilton product of (w, x, y, z) quaternions
void qmul_d(const double* a, const double* b, double* o) {
    const double w = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
    const double x = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
    const double y = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
    const double z = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
    o[0] = w;
    o[1] = x;
    o[2] = y;
    o[3] = z;
}
// rotate v by unit q (w, x, y, z)
void qrot_d(const double* q, const double* v, double* o) {
    const double u[3] = {q[1], q[2], q[3]};
    const double t[3] = {2 * (u[1] * v[2] - u[2] * v[1]), 2 * (u[2] * v[0] - u[0] * v[2]),
                         2 * (u[0] * v[1] - u[1] * v[0])};
    o[0] = v[0] + q[0] * t[0] + (u[1] * t[2] - u[2] * t[1]);
    o[1] = v[1] + q[0] * t[1] + (u[2] * t[0] - u[0] * t[2]);
    o[2] = v[2] + q[0] * t[2] + (u[0] * t[1] - u[1] * t[0]);
}
// Exp_s3 (common/Math.hpp:197-221)
void exp_s3_d(const double* v, double* q) {
    const double th = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (th < 1e-6) {
        double a[4] = {1.0, 0.5 * v[0], 0.5 * v[1], 0.5 * v[2]};
        const double nq = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + a[3] * a[3]);
        for (int k = 0; k < 4; k++) q[k] = a[k] / nq;
        return;
    }
    const double s = std::sin(0.5 * th) / th;
    q[0] = std::cos(0.5 * th);
    q[1] = s * v[0];
    q[2] = s * v[1];
    q[3] = s * v[2];
}
HostRGround make_rground(uint32_t p, uint32_t shape, double radius, double k_start) {
    // RigidBodyGroundCollisionEnergyPool::addEnergy with mu_s 0.5, mu_k 0.2
    // (simulation/Simulation.hpp:108-138)
    HostRGround h{};
    h.p = p;
    h.shape = shape;
    h.radius = radius;
    h.cp_rb_local[1] = -radius;
    h.k_start = h.k_n = h.k_t = h.k_b = k_start;
    h.mu_s = 0.5;
    h.mu_k = 0.2;
    return h;
}
}  // namespace

int avbd_add_rod(avbd_ctx* c, uint32_t elements, double length, const double* curvature, double E, double nu,
                 double density, double diameter, const double* pos, const double* rot, const double* vel,
                 int base_fixed, int tip_fixed, int with_ground, uint32_t* first) {
    return guarded(c, [&] {
        require_not_planned(c);
        if (elements == 0 || !(length > 0) || !(diameter > 0) || !pos)
            throw std::runtime_error("rod needs elements >= 1, length > 0, diameter > 0 and a position");
        const uint32_t nn = elements + 1;
        check_capacity(c, c->mass.size() + nn, c->desc.particle_capacity);
        check_capacity(c, c->cos.size() + elements, c->desc.energy_capacity);
        if (with_ground) check_capacity(c, c->rground.size() + nn, c->desc.energy_capacity);
        const double radius = diameter / 2, G = E / (2 * (1 + nu));
        const double ds = length / (nn - 1);
        const double kap[3] = {curvature ? curvature[0] : 0.0, curvature ? curvature[1] : 0.0,
                               curvature ? curvature[2] : 0.0};
        const double dR[3] = {kap[0] * ds, kap[1] * ds, kap[2] * ds};
        double cp[3] = {pos[0], pos[1], pos[2]};
        double cq[4] = {1, 0, 0, 0};
        if (rot) std::memcpy(cq, rot, sizeof(cq));
        const double zero3[3] = {0, 0, 0};
        std::vector<uint32_t> nodes(nn);
        for (uint32_t i = 0; i < nn; i++) {
            nodes[i] = push_oriented(c, cp, cq, vel, nullptr, 0.0, zero3, 0);
            double step[3];
            const double ez[3] = {0, 0, ds};
            qrot_d(cq, ez, step);
            for (int k = 0; k < 3; k++) cp[k] += step[k];
            double e[4], nq[4];
            exp_s3_d(dR, e);
            qmul_d(cq, e, nq);
            std::memcpy(cq, nq, sizeof(cq));
        }
        const double pi = 3.14159265358979323846;
        const double area = pi * radius * radius;
        const double Ix = pi * radius * radius * radius * radius / 4.0, Iy = Ix, Iz = 2 * Ix;
        const double elem_mass = ds * area * density;
        const double elem_I[3] = {density * ds * Ix, density * ds * Iy, density * ds * Iz};
        for (uint32_t e = 0; e < elements; e++)
            for (uint32_t j = e; j <= e + 1; j++) {
                c->mass[nodes[j]] += 0.5 * elem_mass;
                for (int k = 0; k < 3; k++) c->inertia[3 * c->rot_of[nodes[j]] + k] += 0.5 * elem_I[k];
            }
        for (uint32_t e = 0; e < elements; e++) {
            HostCos h{};
            h.idx[0] = nodes[e];
            h.idx[1] = nodes[e + 1];
            const double st[6] = {G * area, G * area, E * area, E * Ix, E * Iy, G * Iz};
            std::memcpy(h.stiffness, st, sizeof(st));
            h.L = ds;
            std::memcpy(h.kappa0, kap, sizeof(kap));
            c->cos.push_back(h);
        }
        if (base_fixed) c->fixed[nodes.front()] = 1;
        if (tip_fixed) c->fixed[nodes.back()] = 1;
        if (with_ground) {
            const double dt2 = c->desc.dt * c->desc.dt;
            for (uint32_t nd : nodes)
                c->rground.push_back(make_rground(nd, AVBD_SHAPE_ROD_SEGMENT, radius, c->mass[nd] / dt2));
        }
        if (first) *first = nodes.front();
    });
}

int avbd_add_sphere(avbd_ctx* c, double radius, double density, const double* pos, const double* rot, int fixed,
                    int with_ground, uint32_t* first) {
    return guarded(c, [&] {
        require_not_planned(c);
        if (!(radius > 0) || !pos) throw std::runtime_error("sphere needs radius > 0 and a position");
        check_capacity(c, c->mass.size() + 1, c->desc.particle_capacity);
        if (with_ground) check_capacity(c, c->rground.size() + 1, c->desc.energy_capacity);
        const double pi = 3.14159265358979323846;
        const double m = 4 * pi / 3 * radius * radius * radius * density;
        const double I = 0.4 * m * radius * radius;
        const double inertia[3] = {I, I, I};
        const uint32_t p = push_oriented(c, pos, rot, nullptr, nullptr, m, inertia, fixed ? 1 : 0);
        if (with_ground)
            c->rground.push_back(make_rground(p, AVBD_SHAPE_SPHERE, radius, m / (c->desc.dt * c->desc.dt)));
        if (first) *first = p;
    });
}

// TetMeshObject::setup (simobject/TetMeshObject.cpp:46-96) and the per-vertex
// ground energies of simulation/Simulation.hpp:117-126, in FP64.
int avbd_add_tet_mesh(avbd_ctx* c, uint32_t nv, const double* rest_x, uint32_t nt, const uint32_t* tets, double E,
                      double nu, double density, double kd, int with_ground, double mu_s, double mu_k,
                      uint32_t* first) {
    return guarded(c, [&] {
        require_not_planned(c);
        if (!rest_x || (nt && !tets)) throw std::runtime_error("null mesh arrays");
        const uint32_t base = (uint32_t)c->mass.size();
        check_capacity(c, (size_t)base + nv, c->desc.particle_capacity);
        check_capacity(c, c->nh.size() + nt, c->desc.energy_capacity);
        if (with_ground) check_capacity(c, c->ground.size() + nv, c->desc.energy_capacity);
        for (uint32_t t = 0; t < nt; t++)
            for (int k = 0; k < 4; k++)
                if (tets[4 * t + k] >= nv) throw std::runtime_error("tet index out of range");
        for (uint32_t i = 0; i < nv; i++) push_particle(c, rest_x + 3 * i, nullptr, nullptr, nullptr, 0.0, 0);
        const double mu = E / (2 * (1 + nu));
        const double lambda = (E * nu) / ((1 + nu) * (1 - 2 * nu));
        for (uint32_t t = 0; t < nt; t++) {
            HostNH h{};
            for (int k = 0; k < 4; k++) h.idx[k] = base + tets[4 * t + k];
            const double* p[4];
            for (int k = 0; k < 4; k++) p[k] = &c->x[3 * h.idx[k]];
            double X[3][3];
            for (int r = 0; r < 3; r++)
                for (int col = 0; col < 3; col++) X[r][col] = p[col][r] - p[3][r];
            const double det = X[0][0] * (X[1][1] * X[2][2] - X[1][2] * X[2][1]) -
                               X[0][1] * (X[1][0] * X[2][2] - X[1][2] * X[2][0]) +
                               X[0][2] * (X[1][0] * X[2][1] - X[1][1] * X[2][0]);
            if (!(std::fabs(det) > 0)) throw std::runtime_error("degenerate tet in rest state");
            const double inv = 1.0 / det;
            h.Q[0] = (X[1][1] * X[2][2] - X[1][2] * X[2][1]) * inv;
            h.Q[1] = (X[0][2] * X[2][1] - X[0][1] * X[2][2]) * inv;
            h.Q[2] = (X[0][1] * X[1][2] - X[0][2] * X[1][1]) * inv;
            h.Q[3] = (X[1][2] * X[2][0] - X[1][0] * X[2][2]) * inv;
            h.Q[4] = (X[0][0] * X[2][2] - X[0][2] * X[2][0]) * inv;
            h.Q[5] = (X[0][2] * X[1][0] - X[0][0] * X[1][2]) * inv;
            h.Q[6] = (X[1][0] * X[2][1] - X[1][1] * X[2][0]) * inv;
            h.Q[7] = (X[0][1] * X[2][0] - X[0][0] * X[2][1]) * inv;
            h.Q[8] = (X[0][0] * X[1][1] - X[0][1] * X[1][0]) * inv;
            h.V = std::fabs(det / 6.0);
            h.lambda = lambda;
            h.mu = mu;
            h.kd = kd;
            c->nh.push_back(h);
            for (int k = 0; k < 4; k++) c->mass[h.idx[k]] += 0.25 * h.V * density;
        }
        if (with_ground) {
            const double dt2 = c->desc.dt * c->desc.dt;
            for (uint32_t i = 0; i < nv; i++) {
                HostGround g{};
                g.p = base + i;
                g.k_start = g.k_n = g.k_t = g.k_b = c->mass[base + i] / dt2;
                g.mu_s = mu_s;
                g.mu_k = mu_k;
                c->ground.push_back(g);
            }
        }
        if (first) *first = base;
    });
}

int avbd_set_particle_state(avbd_ctx* c, uint32_t first, uint32_t n, const double* x, const double* x_prev,
                            const double* v, const double* v_prev) {
    return guarded(c, [&] {
        if ((size_t)first + n > c->mass.size()) throw std::runtime_error("particle range out of bounds");
        if (c->built) {
            // pull the current device state so that NULL fields stay unchanged
            // (a non-finite state is pulled too: this call is how it is replaced)
            std::vector<double> cx(3 * c->mass.size()), cv(3 * c->mass.size());
            const int rc = avbd_download_state(c, cx.data(), cv.data());
            if (rc != AVBD_OK && rc != AVBD_E_NONFINITE) throw std::runtime_error(c->err);
        }
        for (uint32_t i = 0; i < 3 * n; i++) {
            const size_t k = 3 * (size_t)first + i;
            if (x) c->x[k] = x[i];
            if (x_prev) c->x_prev[k] = x_prev[i];
            if (v) c->v[k] = v[i];
            if (v_prev) c->v_prev[k] = v_prev[i];
        }
        if (c->built) {
            upload_particle_range(c, first, n);
            clear_status(c);
        }
    });
}

int avbd_set_external_accel(avbd_ctx* c, uint32_t first, uint32_t n, const double* a) {
    return guarded(c, [&] {
        if ((size_t)first + n > c->mass.size()) throw std::runtime_error("particle range out of bounds");
        if (!a) throw CodeError(AVBD_E_ARG, "a is null");
        if (c->built) {
            if (!c->d.a_ext)
                throw CodeError(AVBD_E_ARG, "an external acceleration field must first be set before avbd_build_graph");
            std::vector<float4> fa(n);
            for (uint32_t i = 0; i < n; i++)
                fa[i] = make_float4((float)a[3 * i], (float)a[3 * i + 1], (float)a[3 * i + 2], 0.f);
            hip_check(hipMemcpyAsync(const_cast<float4*>(c->d.a_ext) + first, fa.data(), n * sizeof(float4),
                                     hipMemcpyHostToDevice, c->stream),
                      "hipMemcpyAsync H2D");
            hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        }
        if (c->f_ext.empty()) c->f_ext.assign(3 * c->mass.size(), 0.0);
        std::copy(a, a + 3 * (size_t)n, c->f_ext.begin() + 3 * (size_t)first);
    });
}

int avbd_set_fixed(avbd_ctx* c, uint32_t first, uint32_t n, const uint8_t* fixed) {
    return guarded(c, [&] {
        require_not_planned(c);
        if ((size_t)first + n > c->mass.size() || !fixed) throw std::runtime_error("bad fixed range");
        for (uint32_t i = 0; i < n; i++) c->fixed[first + i] = fixed[i];
    });
}

int avbd_add_mass(avbd_ctx* c, uint32_t p, double dm, const double* dI) {
    return guarded(c, [&] {
        require_not_planned(c);
        if (p >= c->mass.size()) throw std::runtime_error("particle out of range");
        c->mass[p] += dm;
        const int32_t r = c->rot_of[p];
        if (dI && r >= 0)
            for (int k = 0; k < 3; k++) c->inertia[3 * r + k] += dI[k];
    });
}

int avbd_set_rotation_state(avbd_ctx* c, uint32_t first, uint32_t n, const double* q, const double* w) {
    return guarded(c, [&] {
        if ((size_t)first + n > c->mass.size()) throw std::runtime_error("particle range out of bounds");
        if (c->built) {
            // pull the current device orientations so that NULL fields stay unchanged
            const size_t np = c->mass.size();
            std::vector<double> cq(4 * np), cw(3 * np);
            if (avbd_download_rotations(c, cq.data(), cw.data()) != AVBD_OK) throw std::runtime_error(c->err);
            for (size_t p = 0; p < np; p++) {
                const int32_t r = c->rot_of[p];
                if (r < 0) continue;
                for (int k = 0; k < 4; k++) c->q[4 * r + k] = cq[4 * p + k];
                for (int k = 0; k < 3; k++) c->w[3 * r + k] = cw[3 * p + k];
            }
        }
        for (uint32_t i = 0; i < n; i++) {
            const int32_t r = c->rot_of[first + i];
            if (r < 0) continue;
            if (q)
                for (int k = 0; k < 4; k++) c->q[4 * r + k] = q[4 * i + k];
            if (w)
                for (int k = 0; k < 3; k++) c->w[3 * r + k] = w[3 * i + k];
        }
        if (c->built) upload_rotation_range(c, 0, c->d.n_rot);
    });
}

int avbd_set_residual_target(avbd_ctx* c, double rel_tol, uint32_t max_iters, uint32_t check_every) {
    return guarded(c, [&] {
        require_not_built(c);
        if (rel_tol < 0) throw std::runtime_error("rel_tol must be >= 0");
        if (rel_tol > 0) {
            if (max_iters == 0 || check_every == 0) throw std::runtime_error("max_iters and check_every must be >= 1");
            // whole check intervals: the budget is rounded up to a multiple of check_every
            max_iters = (max_iters + check_every - 1) / check_every * check_every;
        }
        c->res_tol = rel_tol;
        c->res_max_iters = max_iters;
        c->res_check_every = check_every;
    });
}

namespace {
// Per-check history of the last step (which = 0: whole scene set, 1: worst scene).
void residual_history(avbd_ctx* c, int which, double* out, uint32_t cap, uint32_t* count) {
    require_built(c);
    if (!c->rc) throw std::runtime_error("residual-targeted iteration is off");
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    ResCtl rc{};
    hip_check(hipMemcpy(&rc, c->rc, sizeof(ResCtl), hipMemcpyDeviceToHost), "D2H");
    const uint32_t n = std::min(rc.checks, kMaxResidualChecks);
    const double* src = which ? rc.worst : rc.rel;
    for (uint32_t k = 0; k < n && k < cap; k++) out[k] = src[k];
    if (count) *count = n;
}
}  // namespace

int avbd_get_residual_history(avbd_ctx* c, double* rel, uint32_t cap, uint32_t* count) {
    return guarded(c, [&] { residual_history(c, 0, rel, cap, count); });
}

int avbd_get_scene_residual_history(avbd_ctx* c, double* worst, uint32_t cap, uint32_t* count) {
    return guarded(c, [&] { residual_history(c, 1, worst, cap, count); });
}

int avbd_get_scene_iterations(avbd_ctx* c, uint32_t* iters, uint32_t cap, uint32_t* count) {
    return guarded(c, [&] {
        require_built(c);
        if (!c->rc) throw std::runtime_error("residual-targeted iteration is off");
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        const uint32_t ns = c->n_res_scenes;
        std::vector<uint32_t> h(ns, 0);
        if (c->res_iters) {
            hip_check(hipMemcpy(h.data(), c->res_iters, ns * sizeof(uint32_t), hipMemcpyDeviceToHost), "D2H");
        } else {
            ResCtl rc{};
            hip_check(hipMemcpy(&rc, c->rc, sizeof(ResCtl), hipMemcpyDeviceToHost), "D2H");
            h.assign(ns, rc.last_iters);
        }
        for (uint32_t s = 0; s < ns && s < cap; s++) iters[s] = h[s];
        if (count) *count = ns;
    });
}

int avbd_set_scenes(avbd_ctx* c, uint32_t n_scenes, const uint32_t* offsets) {
    return guarded(c, [&] {
        require_not_planned(c);
        if (n_scenes == 0 || !offsets) throw std::runtime_error("need at least one scene");
        if (offsets[0] != 0 || offsets[n_scenes] != c->mass.size())
            throw std::runtime_error("scene offsets must start at 0 and end at the particle count");
        for (uint32_t s = 0; s < n_scenes; s++)
            if (offsets[s + 1] < offsets[s]) throw std::runtime_error("scene offsets must be non-decreasing");
        c->scene_offsets.assign(offsets, offsets + n_scenes + 1);
    });
}

int avbd_build_graph(avbd_ctx* c) {
    return guarded(c, [&] {
        require_not_built(c);
        const uint32_t n = (uint32_t)c->mass.size();
        // the device first: the colouring runs on it when there is one
        const hipError_t dev_err = hipSetDevice(c->device);
        if (dev_err == hipSuccess && !c->stream)
            hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
        if (!c->planned) {
            std::vector<uint32_t> so = c->scene_offsets;
            if (so.empty()) so = {0u, n};
            if (so.back() != n) throw std::runtime_error("scene offsets do not cover the particle pool");
            HostModel m;
            m.n = n;
            m.nh = &c->nh;
            m.ground = &c->ground;
            m.cos = &c->cos;
            m.rground = &c->rground;
            m.joints = &c->joints;
            m.joints2 = &c->joints2;
            m.fixed = &c->fixed;
            m.rot_of = &c->rot_of;
            m.scene_offsets = &so;
            SceneColourer colour;
            if (dev_err == hipSuccess && !(c->desc.flags & AVBD_FLAG_HOST_COLORING))
                colour = [c](const std::vector<uint32_t>& scene_off, const std::vector<uint32_t>& off,
                             const std::vector<uint32_t>& nb, std::vector<uint32_t>& color,
                             std::vector<uint32_t>& scene_colors) {
                    device_dsatur(c, scene_off, off, nb, color, scene_colors);
                };
            build_color_plan(m, c->plan, colour);
            c->planned = true;
        }
        // device part
        hip_check(dev_err, "hipSetDevice");
        if (!c->tris.empty() && !(c->desc.flags & AVBD_FLAG_SERIAL_BROADPHASE)) {
            hip_check(hipStreamCreateWithFlags(&c->side_stream, hipStreamNonBlocking), "hipStreamCreate");
            hip_check(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming), "hipEventCreate");
            hip_check(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming), "hipEventCreate");
        }
        if (!c->ev_a) hip_check(hipEventCreate(&c->ev_a), "hipEventCreate");
        if (!c->ev_b) hip_check(hipEventCreate(&c->ev_b), "hipEventCreate");
        c->damping = false;
        for (const auto& t : c->nh)
            if (t.kd != 0.0) c->damping = true;

        DevState& d = c->d;
        d.n = n;
        d.n_tets = (uint32_t)c->nh.size();
        d.n_ground = (uint32_t)c->ground.size();
        d.n_items = (uint32_t)c->plan.items.size();
        d.n_inc = (uint32_t)c->plan.inc.size();
        d.x = dev_alloc<float4>(c, n);
        d.x_buf = dev_alloc<float4>(c, n);
        d.y = dev_alloc<float4>(c, n);
        d.x_prev = dev_alloc<float4>(c, n);
        d.x_last = dev_alloc<float4>(c, n);
        d.x_lastlast = dev_alloc<float4>(c, n);
        d.v = dev_alloc<float4>(c, n);
        d.v_prev = dev_alloc<float4>(c, n);
        d.x_scratch = dev_alloc<float4>(c, n);
        d.fixed = dev_alloc<uint8_t>(c, n);
        if (c->fp64) {
            if (!c->q.empty())
                throw CodeError(AVBD_E_ARG, "the FP64 state covers positional particles only (no oriented particles)");
            F64State& f = c->f;
            f.x = dev_alloc<double4>(c, n);
            f.y = dev_alloc<double4>(c, n);
            f.x_prev = dev_alloc<double4>(c, n);
            f.x_last = dev_alloc<double4>(c, n);
            f.x_lastlast = dev_alloc<double4>(c, n);
            f.x0 = dev_alloc<double4>(c, n);
            f.x_scratch = dev_alloc<double4>(c, n);
        }
        d.ground_of = dev_alloc<int32_t>(c, n);
        d.tets = dev_alloc<TetRec>(c, d.n_tets);
        d.damp = dev_alloc<TetDamp>(c, d.n_tets);
        d.ground = dev_alloc<GroundRec>(c, d.n_ground);
        d.inc = dev_alloc<uint32_t>(c, d.n_inc);
        d.items = dev_alloc<VItem>(c, d.n_items);
        c->d_conflicts = dev_alloc<uint32_t>(c, c->plan.conflict_list.size());

        // oriented particles and their energies
        d.n_rot = (uint32_t)(c->q.size() / 4);
        d.n_cos = (uint32_t)c->cos.size();
        d.n_rground = (uint32_t)c->rground.size();
        d.n_joints1 = (uint32_t)c->joints.size();
        d.n_joints = d.n_joints1 + (uint32_t)c->joints2.size();
        d.n_oitems = (uint32_t)c->plan.oitems.size();
        d.rot_of = dev_alloc<int32_t>(c, n);
        d.q = dev_alloc<float4>(c, d.n_rot);
        d.q_buf = dev_alloc<float4>(c, d.n_rot);
        d.q_in = dev_alloc<float4>(c, d.n_rot);
        d.q_prev = dev_alloc<float4>(c, d.n_rot);
        d.q_last = dev_alloc<float4>(c, d.n_rot);
        d.q_lastlast = dev_alloc<float4>(c, d.n_rot);
        d.w = dev_alloc<float4>(c, d.n_rot);
        d.w_prev = dev_alloc<float4>(c, d.n_rot);
        d.inertia = dev_alloc<float4>(c, d.n_rot);
        d.cos = dev_alloc<CosRec>(c, d.n_cos);
        d.rground = dev_alloc<RGroundRec>(c, d.n_rground);
        d.joints = dev_alloc<JointRec>(c, d.n_joints);
        d.oitems = dev_alloc<OItem>(c, d.n_oitems);
        d.oinc = dev_alloc<uint32_t>(c, c->plan.oinc.size());
        d.prec_off = dev_alloc<uint32_t>(c, c->plan.prec_off.size());
        d.prec = dev_alloc<uint32_t>(c, c->plan.prec.size());
        {
            std::vector<CosRec> cr(d.n_cos);
            for (uint32_t i = 0; i < d.n_cos; i++) cr[i] = to_dev(c->cos[i]);
            std::vector<RGroundRec> rr(d.n_rground);
            for (uint32_t i = 0; i < d.n_rground; i++) rr[i] = to_dev(c->rground[i]);
            std::vector<JointRec> jr(d.n_joints);
            for (uint32_t i = 0; i < d.n_joints1; i++) jr[i] = to_dev(c->joints[i]);
            for (uint32_t i = d.n_joints1; i < d.n_joints; i++) jr[i] = to_dev(c->joints2[i - d.n_joints1]);
            upload(c, d.cos, cr);
            upload(c, d.rground, rr);
            upload(c, d.joints, jr);
            upload(c, d.rot_of, c->rot_of);
            upload(c, d.oitems, c->plan.oitems);
            upload(c, d.oinc, c->plan.oinc);
            upload(c, d.prec_off, c->plan.prec_off);
            upload(c, d.prec, c->plan.prec);
            upload_rotation_range(c, 0, d.n_rot);
        }
        // reduction slots: enough blocks for the stats passes at any launch config
        c->reduce_blocks = std::max((d.n_items * 32 + 63) / 64 + (d.n_oitems + 255) / 256,
                                    (d.n_tets + d.n_ground + d.n_cos + 255) / 256) +
                           1;
        d.reduce = dev_alloc<double>(c, (size_t)c->reduce_blocks * 4);

        // learned warm start buffers: features, per-particle material and
        // scene, per-scene IP difference, counters, recording buffers
        {
            c->rows_pad = std::max<uint32_t>(16u, (d.n_items + 15u) / 16u * 16u);
            c->wio.x16 = dev_alloc<uint16_t>(c, (size_t)c->rows_pad * kMlpInPad);
            c->d_raw = dev_alloc<float>(c, (size_t)std::max<uint32_t>(1u, d.n_items) * kMlpIn);
            c->d_tgt = dev_alloc<float>(c, (size_t)std::max<uint32_t>(1u, d.n_items) * 3);
            std::vector<double> le(n, 0.0), lnu(n, 0.0), cnt(n, 0.0);
            for (const auto& t : c->nh) {
                const double E = t.mu * (3 * t.lambda + 2 * t.mu) / (t.lambda + t.mu);
                const double nu = t.lambda / (2 * (t.lambda + t.mu));
                for (int k = 0; k < 4; k++) {
                    le[t.idx[k]] += std::log10(E);
                    lnu[t.idx[k]] += nu;
                    cnt[t.idx[k]] += 1;
                }
            }
            std::vector<float2> mat(n);
            for (uint32_t p = 0; p < n; p++)
                mat[p] = cnt[p] > 0 ? make_float2((float)(le[p] / cnt[p]), (float)(lnu[p] / cnt[p])) : make_float2(0.f, 0.f);
            std::vector<uint32_t> so = c->scene_offsets;
            if (so.empty()) so = {0u, n};
            std::vector<uint32_t> scene_of(n, 0);
            for (uint32_t sc = 0; sc + 1 < so.size(); sc++)
                for (uint32_t p = so[sc]; p < so[sc + 1]; p++) scene_of[p] = sc;
            c->wio.n_scenes = (uint32_t)(so.size() - 1);
            float2* d_mat = dev_alloc<float2>(c, n);
            uint32_t* d_scene = dev_alloc<uint32_t>(c, n);
            c->wio.ipd = dev_alloc<double>(c, (size_t)c->wio.n_scenes * kWarmAlphas);
            c->wio.pred = dev_alloc<float4>(c, n);
            c->wio.counters = dev_alloc<unsigned long long>(c, 2);
            c->wio.step_of = dev_alloc<float>(c, c->wio.n_scenes);
            upload(c, d_mat, mat);
            upload(c, d_scene, scene_of);
            hip_check(hipMemsetAsync(c->wio.counters, 0, 2 * sizeof(unsigned long long), c->stream), "hipMemsetAsync");
            c->wio.mat = d_mat;
            c->wio.scene_of = d_scene;
        }

        std::vector<TetRec> tr(d.n_tets);
        std::vector<TetDamp> td(d.n_tets);
        for (uint32_t t = 0; t < d.n_tets; t++) {
            const HostNH& h = c->nh[t];
            std::memcpy(tr[t].idx, h.idx, sizeof(h.idx));
            for (int k = 0; k < 9; k++) tr[t].q[k] = (float)h.Q[k];
            tr[t].VL = (float)(h.V * h.lambda);
            tr[t].VM = (float)(h.V * h.mu);
            tr[t].alpha = (float)(h.mu / h.lambda + 1.0);
            td[t].e00 = (float)h.E_prev[0];
            td[t].e11 = (float)h.E_prev[4];
            td[t].e22 = (float)h.E_prev[8];
            td[t].e01 = (float)h.E_prev[1];
            td[t].e02 = (float)h.E_prev[2];
            td[t].e12 = (float)h.E_prev[5];
            td[t].damp = (float)(2.0 / c->desc.dt * h.V * h.kd);
            td[t].pad = 0.f;
        }
        std::vector<GroundRec> gr(d.n_ground);
        for (uint32_t i = 0; i < d.n_ground; i++) gr[i] = to_dev(c->ground[i]);
        std::vector<int32_t> gof(n, -1);
        for (uint32_t i = 0; i < d.n_ground; i++) gof[c->ground[i].p] = (int32_t)i;
        upload(c, d.tets, tr);
        upload(c, d.damp, td);
        upload(c, d.ground, gr);
        upload(c, d.ground_of, gof);
        upload(c, d.fixed, c->fixed);
        upload(c, d.inc, c->plan.inc);
        {
            // per-incidence record copies (work-item order) for the vertex solve
            if (n >= (1u << 30)) throw std::runtime_error("more than 2^30 particles");
            std::vector<TetRec> ir(c->plan.inc.size());
            for (size_t k = 0; k < ir.size(); k++) {
                ir[k] = tr[c->plan.inc[k] >> 2];
                ir[k].idx[3] |= (c->plan.inc[k] & 3u) << 30;
            }
            d.irec = dev_alloc<TetRec>(c, ir.size());
            upload(c, d.irec, ir);
        }
        upload(c, d.items, c->plan.items);
        upload(c, c->d_conflicts, c->plan.conflict_list);
        if (!c->f_ext.empty()) {
            // the field, if any, is fixed at build time (its buffer is part of the captured steps)
            std::vector<float4> fa(n);
            for (uint32_t p = 0; p < n; p++)
                fa[p] = make_float4((float)c->f_ext[3 * p], (float)c->f_ext[3 * p + 1], (float)c->f_ext[3 * p + 2], 0.f);
            float4* d_fa = dev_alloc<float4>(c, n);
            upload(c, d_fa, fa);
            d.a_ext = d_fa;
        }
        if (c->fp64) {
            F64State& f = c->f;
            f.a_ext = d.a_ext;
            f.x32 = d.x;
            f.y32 = d.y;
            f.x_prev32 = d.x_prev;
            f.v = d.v;
            f.v_prev = d.v_prev;
            f.fixed = d.fixed;
            f.ground_of = d.ground_of;
            f.tets = d.tets;
            f.damp = d.damp;
            f.ground = d.ground;
            f.inc = d.inc;
            f.irec = d.irec;
            f.items = d.items;
            f.n = n;
            f.n_tets = d.n_tets;
            f.n_ground = d.n_ground;
            f.n_items = d.n_items;
        }
        upload_particle_range(c, 0, n);

        // LBVH broad phase buffers
        const uint32_t nt = (uint32_t)(c->tris.size() / 3);
        if (nt) {
            for (uint32_t p : c->tris)
                if (p >= n) throw std::runtime_error("collision triangle references a particle out of range");
            alloc_lbvh(c, c->lb, nt);
            c->lb.tris = dev_alloc<uint32_t>(c, 3 * (size_t)nt);
            upload(c, c->lb.tris, c->tris);
        }
        // Chebyshev schedule (per iteration; residual mode reads it on the device)
        {
            std::vector<float> om(c->iters());
            double w = 1;
            for (uint32_t i = 0; i < c->iters(); i++) om[i] = (float)(w = omega_at(i, c->desc.iter_acceleration, w));
            c->omega_seq = om;
        }
        d.status = dev_alloc<uint32_t>(c, 1);
        hip_check(hipMemsetAsync(d.status, 0, sizeof(uint32_t), c->stream), "hipMemsetAsync");
        if (c->res_tol > 0) {
            // per scene: particle range and ||M_s g||^2 over its free particles,
            // as in the residual definition; |r_v|^2 per particle
            std::vector<uint32_t> so = c->scene_offsets;
            if (so.empty()) so = {0u, n};
            c->n_res_scenes = (uint32_t)(so.size() - 1);
            std::vector<double> mg2(c->n_res_scenes, 0.0);
            for (uint32_t sc = 0; sc < c->n_res_scenes; sc++)
                for (uint32_t p = so[sc]; p < so[sc + 1]; p++)
                    if (!c->fixed[p]) {
                        const double mg = c->mass[p] * c->desc.g_accel;
                        mg2[sc] += mg * mg;
                    }
            double mg2_all = 0;
            for (double v : mg2) mg2_all += v;
            c->res_mg2 = mg2_all;
            c->rc = dev_alloc<ResCtl>(c, 1);
            c->res_rr = dev_alloc<double>(c, n);
            c->res_scene_off = dev_alloc<uint32_t>(c, so.size());
            c->res_mg2_scene = dev_alloc<double>(c, c->n_res_scenes);
            c->res_acc_scene = dev_alloc<double>(c, c->n_res_scenes);
            c->omega_tab = dev_alloc<float>(c, c->omega_seq.size());
            upload(c, c->res_scene_off, so);
            upload(c, c->res_mg2_scene, mg2);
            upload(c, c->omega_tab, c->omega_seq);
            if (c->n_res_scenes > 1) {
                std::vector<uint32_t> scene_of(n, 0), items(c->n_res_scenes, 0);
                for (uint32_t sc = 0; sc < c->n_res_scenes; sc++)
                    for (uint32_t p = so[sc]; p < so[sc + 1]; p++) scene_of[p] = sc;
                for (const VItem& it : c->plan.items) items[scene_of[it.v]]++;
                for (const OItem& it : c->plan.oitems) items[scene_of[it.v]]++;
                c->res_frozen = dev_alloc<uint32_t>(c, c->n_res_scenes);
                c->res_iters = dev_alloc<uint32_t>(c, c->n_res_scenes);
                c->res_items = dev_alloc<uint32_t>(c, c->n_res_scenes);
                c->res_scene_of = dev_alloc<uint32_t>(c, n);
                upload(c, c->res_items, items);
                upload(c, c->res_scene_of, scene_of);
                hip_check(hipMemsetAsync(c->res_frozen, 0, c->n_res_scenes * sizeof(uint32_t), c->stream),
                          "hipMemsetAsync");
                hip_check(hipMemsetAsync(c->res_iters, 0, c->n_res_scenes * sizeof(uint32_t), c->stream),
                          "hipMemsetAsync");
            }
            hip_check(hipMemsetAsync(c->rc, 0, sizeof(ResCtl), c->stream), "hipMemsetAsync");
            hip_check(hipMemsetAsync(c->res_rr, 0, (size_t)n * sizeof(double), c->stream), "hipMemsetAsync");
            // converged flag of every interval of a step, copied back after the interval
            const uint32_t intervals = c->res_max_iters / c->res_check_every;
            void* hc = nullptr;
            hip_check(hipHostMalloc(&hc, (size_t)intervals * sizeof(uint32_t), hipHostMallocDefault), "hipHostMalloc");
            c->h_conv = static_cast<uint32_t*>(hc);
            std::memset(c->h_conv, 0, (size_t)intervals * sizeof(uint32_t));
            while (c->part_ev.size() < intervals) {
                hipEvent_t e = nullptr;
                hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
                c->part_ev.push_back(e);
            }
        }
        // Chebyshev / dual update fused into the last colour phase (scenes
        // without oriented particles or conflicted vertices)
        // (residual mode: with an even check interval, so that every interval
        // starts on the same buffer and one interval graph serves them all)
        if (!(c->desc.flags & AVBD_FLAG_SEPARATE_CHEBYSHEV) && c->d.n_rot == 0 && c->d.n_oitems == 0 &&
            c->plan.num_colors > 0 && c->plan.conflict_list.empty() && (!c->rc || c->res_check_every % 2 == 0)) {
            const uint32_t last = c->plan.num_colors - 1;
            std::vector<uint8_t> in_last(n, 0);
            for (uint32_t k = c->plan.item_offsets[last]; k < c->plan.item_offsets[last + 1]; k++)
                in_last[c->plan.items[k].v] = 1;
            std::vector<uint32_t> rest;
            for (uint32_t p = 0; p < n; p++)
                if (!in_last[p]) rest.push_back(p);
            if (c->fp64) c->f.x_alt = dev_alloc<double4>(c, n);
            else c->x_alt = dev_alloc<float4>(c, n);
            c->n_cheb = (uint32_t)rest.size();
            c->cheb_list = dev_alloc<uint32_t>(c, std::max<size_t>(1, rest.size()));
            if (!rest.empty()) upload(c, c->cheb_list, rest);
            c->fuse_cheb = true;
        }
        c->built = true;

        // timing events bracket the solve launches of a fixed-count step
        const uint32_t n_solves = c->rc ? 0u : c->iters() * c->plan.num_colors;
        c->events.resize(2 * n_solves);
        hip_check(hipEventCreate(&c->ev_lb0), "hipEventCreate");
        hip_check(hipEventCreate(&c->ev_lb1), "hipEventCreate");
        for (auto& ev : c->events) hip_check(hipEventCreate(&ev), "hipEventCreate");

        if (c->desc.lanes_per_vertex == 0 || c->desc.block_size == 0) {
            autotune(c);
            tune_fusion(c);
        }
        if (!(c->desc.flags & AVBD_FLAG_NO_GRAPH)) capture_graph(c);
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    });
}

int avbd_step(avbd_ctx* c, uint32_t n_steps) {
    return guarded(c, [&] {
        require_built(c);
        hip_check(hipEventRecord(c->ev_a, c->stream), "hipEventRecord");
        if (c->nonfinite) throw CodeError(AVBD_E_NONFINITE, nonfinite_message());
        if (c->dynamic) {
            // per step: detection round (synchronising), then the solve
            double total = 0, detect_ms = 0;
            uint32_t launches = 0;
            for (uint32_t st = 0; st < n_steps; st++) {
                const auto t0 = std::chrono::steady_clock::now();
                detect_round(c);
                detect_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                if (c->rc) {
                    // residual-targeted: the interval sequence, replayed from
                    // graphs captured for this round's layout (eager with
                    // AVBD_FLAG_NO_GRAPH or timing)
                    const bool graphs = !(c->desc.flags & AVBD_FLAG_NO_GRAPH) && !c->timing;
                    if (graphs) dyn_residual_graphs(c);
                    run_residual_step(c, graphs);
                    continue;
                }
                enqueue_step(c, c->timing);
                c->vertex_updates += (uint64_t)c->desc.solver_iters * (c->d.n_items + c->d.n_oitems);
                if (c->timing) {
                    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
                    for (size_t k = 0; k + 1 < 2 * (size_t)c->iters() * c->plan.num_colors; k += 2) {
                        float ms = 0;
                        hip_check(hipEventElapsedTime(&ms, c->events[k], c->events[k + 1]), "hipEventElapsedTime");
                        total += ms;
                        launches++;
                    }
                }
            }
            if (c->timing) {
                c->solve_avg_ms = launches ? total / launches : 0;
                c->solve_launches = launches;
            }
            c->lbvh_avg_ms = n_steps ? detect_ms / n_steps : 0;
            hip_check(hipEventRecord(c->ev_b, c->stream), "hipEventRecord");
            c->steps_taken += n_steps;
            return;
        }
        if (c->rc) {
            for (uint32_t st = 0; st < n_steps; st++) run_residual_step(c, c->head_exec != nullptr && !c->timing);
        } else if (c->timing) {
            double total = 0, lbvh_total = 0;
            uint32_t launches = 0;
            for (uint32_t s = 0; s < n_steps; s++) {
                enqueue_step(c, true);
                hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
                for (size_t k = 0; k + 1 < c->events.size(); k += 2) {
                    float ms = 0;
                    hip_check(hipEventElapsedTime(&ms, c->events[k], c->events[k + 1]), "hipEventElapsedTime");
                    total += ms;
                    launches++;
                }
                if (c->lb.n) {
                    float ms = 0;
                    hip_check(hipEventElapsedTime(&ms, c->ev_lb0, c->ev_lb1), "hipEventElapsedTime");
                    lbvh_total += ms;
                }
            }
            c->solve_avg_ms = launches ? total / launches : 0;
            c->lbvh_avg_ms = lbvh_total / std::max(1u, n_steps);
            c->solve_launches = launches;
        } else if (c->graph_exec) {
            for (uint32_t s = 0; s < n_steps; s++) hip_check(hipGraphLaunch(c->graph_exec, c->stream), "hipGraphLaunch");
        } else {
            for (uint32_t s = 0; s < n_steps; s++) enqueue_step(c, false);
        }
        hip_check(hipEventRecord(c->ev_b, c->stream), "hipEventRecord");
        c->steps_taken += n_steps;
        if (!c->rc) c->vertex_updates += (uint64_t)n_steps * c->desc.solver_iters * (c->d.n_items + c->d.n_oitems);
    });
}

int avbd_synchronize(avbd_ctx* c) {
    return guarded(c, [&] {
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        if (c->built) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, c->ev_a, c->ev_b) == hipSuccess) c->last_step_ms = ms;
            check_status(c);
        }
    });
}

int avbd_download_state(avbd_ctx* c, double* x, double* v) {
    return guarded(c, [&] {
        require_built(c);
        const uint32_t n = c->d.n;
        // positions: the FP64 stream when the state is FP64, else the FP32 one
        std::vector<double> px(3 * (size_t)n), pxp(3 * (size_t)n);
        std::vector<float4> bv(n), bvp(n);
        auto positions = [&](const float4* src32, const double4* src64, std::vector<double>& out) {
            if (c->fp64) {
                std::vector<double4> b(n);
                hip_check(hipMemcpyAsync(b.data(), src64, n * sizeof(double4), hipMemcpyDeviceToHost, c->stream),
                          "D2H");
                hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
                for (uint32_t i = 0; i < n; i++) {
                    out[3 * i] = b[i].x;
                    out[3 * i + 1] = b[i].y;
                    out[3 * i + 2] = b[i].z;
                }
            } else {
                std::vector<float4> b(n);
                hip_check(hipMemcpyAsync(b.data(), src32, n * sizeof(float4), hipMemcpyDeviceToHost, c->stream),
                          "D2H");
                hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
                for (uint32_t i = 0; i < n; i++) {
                    out[3 * i] = b[i].x;
                    out[3 * i + 1] = b[i].y;
                    out[3 * i + 2] = b[i].z;
                }
            }
        };
        positions(c->d.x, c->f.x, px);
        positions(c->d.x_prev, c->f.x_prev, pxp);
        hip_check(hipMemcpyAsync(bv.data(), c->d.v, n * sizeof(float4), hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipMemcpyAsync(bvp.data(), c->d.v_prev, n * sizeof(float4), hipMemcpyDeviceToHost, c->stream),
                  "D2H");
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        for (uint32_t i = 0; i < n; i++) {
            const float* fv = &bv[i].x;
            const float* fvp = &bvp[i].x;
            for (int k = 0; k < 3; k++) {
                if (x) x[3 * i + k] = px[3 * i + k];
                if (v) v[3 * i + k] = fv[k];
                // keep the staging copy current (used by avbd_set_particle_state)
                c->x[3 * i + k] = px[3 * i + k];
                c->v[3 * i + k] = fv[k];
                c->x_prev[3 * i + k] = pxp[3 * i + k];
                c->v_prev[3 * i + k] = fvp[k];
            }
        }
        check_status(c);   // the arrays are filled either way
    });
}

int avbd_download_inertial(avbd_ctx* c, double* y) {
    return guarded(c, [&] {
        require_built(c);
        const uint32_t n = c->d.n;
        if (c->fp64) {
            std::vector<double4> by(n);
            hip_check(hipMemcpyAsync(by.data(), c->f.y, n * sizeof(double4), hipMemcpyDeviceToHost, c->stream), "D2H");
            hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
            for (uint32_t i = 0; i < n; i++) {
                y[3 * i] = by[i].x;
                y[3 * i + 1] = by[i].y;
                y[3 * i + 2] = by[i].z;
            }
            return;
        }
        std::vector<float4> by(n);
        hip_check(hipMemcpyAsync(by.data(), c->d.y, n * sizeof(float4), hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        for (uint32_t i = 0; i < n; i++) {
            y[3 * i] = by[i].x;
            y[3 * i + 1] = by[i].y;
            y[3 * i + 2] = by[i].z;
        }
    });
}

int avbd_download_rotations(avbd_ctx* c, double* q, double* w) {
    return guarded(c, [&] {
        require_built(c);
        const uint32_t n = c->d.n, nr = c->d.n_rot;
        std::vector<float4> bq(nr), bw(nr);
        if (nr) {
            hip_check(hipMemcpyAsync(bq.data(), c->d.q, nr * sizeof(float4), hipMemcpyDeviceToHost, c->stream), "D2H");
            hip_check(hipMemcpyAsync(bw.data(), c->d.w, nr * sizeof(float4), hipMemcpyDeviceToHost, c->stream), "D2H");
        }
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        for (uint32_t p = 0; p < n; p++) {
            const int32_t r = c->rot_of[p];
            if (q) {
                if (r < 0) {
                    q[4 * p] = 1;
                    q[4 * p + 1] = q[4 * p + 2] = q[4 * p + 3] = 0;
                } else {
                    q[4 * p] = bq[r].w;
                    q[4 * p + 1] = bq[r].x;
                    q[4 * p + 2] = bq[r].y;
                    q[4 * p + 3] = bq[r].z;
                }
            }
            if (w) {
                w[3 * p] = r < 0 ? 0.0 : bw[r].x;
                w[3 * p + 1] = r < 0 ? 0.0 : bw[r].y;
                w[3 * p + 2] = r < 0 ? 0.0 : bw[r].z;
            }
        }
        check_status(c);   // the arrays are filled either way
    });
}

int avbd_download_energies(avbd_ctx* c, int type, void* records, uint32_t count) {
    return guarded(c, [&] {
        require_built(c);
        if (type == AVBD_NEO_HOOKEAN) {
            if (count > c->d.n_tets) throw std::runtime_error("count exceeds the pool");
            std::vector<TetDamp> td(c->d.n_tets);
            hip_check(hipMemcpyAsync(td.data(), c->d.damp, td.size() * sizeof(TetDamp), hipMemcpyDeviceToHost,
                                     c->stream),
                      "D2H");
            hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
            auto* r = static_cast<avbd_neo_hookean_record*>(records);
            for (uint32_t i = 0; i < count; i++) {
                const HostNH& h = c->nh[i];
                std::memcpy(r[i].particle_indices, h.idx, sizeof(h.idx));
                r[i].lambda = h.lambda;
                r[i].mu = h.mu;
                r[i].kd = h.kd;
                std::memcpy(r[i].Q, h.Q, sizeof(h.Q));
                r[i].rest_volume = h.V;
                const double E[9] = {td[i].e00, td[i].e01, td[i].e02, td[i].e01, td[i].e11,
                                     td[i].e12, td[i].e02, td[i].e12, td[i].e22};
                std::memcpy(r[i].E_prev, E, sizeof(E));
            }
        } else if (type == AVBD_GROUND_COLLISION) {
            if (count > c->d.n_ground) throw std::runtime_error("count exceeds the pool");
            std::vector<GroundRec> gr(c->d.n_ground);
            hip_check(hipMemcpyAsync(gr.data(), c->d.ground, gr.size() * sizeof(GroundRec), hipMemcpyDeviceToHost,
                                     c->stream),
                      "D2H");
            hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
            auto* r = static_cast<avbd_ground_record*>(records);
            for (uint32_t i = 0; i < count; i++) {
                const GroundRec& g = gr[i];
                r[i].particle_index = g.p;
                r[i].use_static = (g.flags & GF_USE_STATIC) ? 1 : 0;
                r[i].unclamped_lambda_tb = (g.flags & GF_UNCLAMPED) ? 1 : 0;
                r[i].pad_ = 0;
                r[i].k_start = g.k_start;
                r[i].k_n = g.k_n;
                r[i].k_t = g.k_t;
                r[i].k_b = g.k_b;
                r[i].lambda_n = g.lambda_n;
                r[i].lambda_t = g.lambda_t;
                r[i].lambda_b = g.lambda_b;
                r[i].C_n_prev = g.C_n_prev;
                r[i].mu_s = g.mu_s;
                r[i].mu_k = g.mu_k;
                r[i].cp_x = g.cp_x;
                r[i].cp_z = g.cp_z;
            }
        } else if (type == AVBD_COSSERAT_ROD) {
            if (count > c->d.n_cos) throw std::runtime_error("count exceeds the pool");
            auto* r = static_cast<avbd_cosserat_record*>(records);
            for (uint32_t i = 0; i < count; i++) {
                const HostCos& h = c->cos[i];
                r[i].particle_indices[0] = h.idx[0];
                r[i].particle_indices[1] = h.idx[1];
                std::memcpy(r[i].stiffness, h.stiffness, sizeof(h.stiffness));
                r[i].rest_length = h.L;
                std::memcpy(r[i].rest_curvature, h.kappa0, sizeof(h.kappa0));
            }
        } else if (type == AVBD_RIGID_BODY_GROUND_COLLISION) {
            if (count > c->d.n_rground) throw std::runtime_error("count exceeds the pool");
            std::vector<RGroundRec> gr(c->d.n_rground);
            if (!gr.empty())
                hip_check(hipMemcpyAsync(gr.data(), c->d.rground, gr.size() * sizeof(RGroundRec),
                                         hipMemcpyDeviceToHost, c->stream),
                          "D2H");
            hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
            auto* r = static_cast<avbd_rigid_ground_record*>(records);
            for (uint32_t i = 0; i < count; i++) {
                const RGroundRec& g = gr[i];
                r[i].particle_index = g.p;
                r[i].shape = g.shape;
                r[i].use_static = (g.flags & GF_USE_STATIC) ? 1 : 0;
                r[i].unclamped_lambda_tb = (g.flags & GF_UNCLAMPED) ? 1 : 0;
                r[i].radius = g.radius;
                r[i].cp_rb_local[0] = g.cpl_x;
                r[i].cp_rb_local[1] = g.cpl_y;
                r[i].cp_rb_local[2] = g.cpl_z;
                r[i].cp_ground[0] = g.cpg_x;
                r[i].cp_ground[1] = g.cpg_y;
                r[i].cp_ground[2] = g.cpg_z;
                r[i].k_start = g.k_start;
                r[i].k_n = g.k_n;
                r[i].k_t = g.k_t;
                r[i].k_b = g.k_b;
                r[i].lambda_n = g.lambda_n;
                r[i].lambda_t = g.lambda_t;
                r[i].lambda_b = g.lambda_b;
                r[i].C_n_prev = g.C_n_prev;
                r[i].mu_s = g.mu_s;
                r[i].mu_k = g.mu_k;
            }
        } else if (type == AVBD_ONE_SIDED_FIXED_JOINT) {
            if (count > c->d.n_joints1) throw std::runtime_error("count exceeds the pool");
            std::vector<JointRec> jr(c->d.n_joints1);
            if (!jr.empty())
                hip_check(hipMemcpyAsync(jr.data(), c->d.joints, jr.size() * sizeof(JointRec), hipMemcpyDeviceToHost,
                                         c->stream),
                          "D2H");
            hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
            auto* r = static_cast<avbd_joint_record*>(records);
            for (uint32_t i = 0; i < count; i++) {
                const HostJoint& h = c->joints[i];
                r[i].particle_index = h.p;
                r[i].pad_ = 0;
                std::memcpy(r[i].position_offset, h.pos_offset, sizeof(h.pos_offset));
                std::memcpy(r[i].rotation_offset, h.rot_offset, sizeof(h.rot_offset));
                std::memcpy(r[i].ref_position, h.ref_position, sizeof(h.ref_position));
                std::memcpy(r[i].ref_rotation, h.ref_rotation, sizeof(h.ref_rotation));
                for (int k = 0; k < 6; k++) {
                    r[i].k[k] = jr[i].k[k];
                    r[i].lambda[k] = jr[i].lambda[k];
                    r[i].C_prev[k] = jr[i].C_prev[k];
                }
            }
        } else if (type == AVBD_FIXED_JOINT) {
            const uint32_t n1 = c->d.n_joints1, n2 = c->d.n_joints - c->d.n_joints1;
            if (count > n2) throw std::runtime_error("count exceeds the pool");
            std::vector<JointRec> jr(n2);
            if (n2)
                hip_check(hipMemcpyAsync(jr.data(), c->d.joints + n1, n2 * sizeof(JointRec), hipMemcpyDeviceToHost,
                                         c->stream),
                          "D2H");
            hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
            auto* r = static_cast<avbd_fixed_joint_record*>(records);
            for (uint32_t i = 0; i < count; i++) {
                const HostJoint2& h = c->joints2[i];
                std::memcpy(r[i].particle_index, h.p, sizeof(h.p));
                std::memcpy(r[i].position_offset, h.pos_offset, sizeof(h.pos_offset));
                std::memcpy(r[i].rotation_offset, h.rot_offset, sizeof(h.rot_offset));
                for (int k = 0; k < 6; k++) {
                    r[i].k[k] = jr[i].k[k];
                    r[i].lambda[k] = jr[i].lambda[k];
                    r[i].C_prev[k] = jr[i].C_prev[k];
                }
            }
        } else {
            throw CodeError(AVBD_E_ARG, "energy type not supported by this build");
        }
    });
}

int avbd_download_coloring(avbd_ctx* c, uint32_t* color, uint32_t* num_colors) {
    return guarded(c, [&] {
        // the set-up colouring is host-side: available once avbd_build_graph()
        // has planned it, even if the device part of that call failed; with
        // dynamic contacts the device recolours every step
        if (!c->planned) throw CodeError(AVBD_E_NOT_BUILT, "avbd_build_graph() has not been called");
        if (c->dynamic && color) {
            hip_check(hipMemcpyAsync(c->plan.color.data(), c->rdev.color, c->plan.color.size() * sizeof(uint32_t),
                                     hipMemcpyDeviceToHost, c->stream),
                      "D2H");
            hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        }
        if (color) std::copy(c->plan.color.begin(), c->plan.color.end(), color);
        if (num_colors) *num_colors = c->plan.num_colors;
    });
}

int avbd_get_stats(avbd_ctx* c, avbd_stats* out) {
    return guarded(c, [&] {
        require_built(c);
        if (!out) throw std::runtime_error("null output");
        avbd_stats st{};
        hip_check(hipMemsetAsync(c->d.reduce, 0, (size_t)c->reduce_blocks * 4 * sizeof(double), c->stream),
                  "hipMemsetAsync");
        if (c->fp64)
            launch_stats_f64(c->stream, f64_state(c), c->cfg, c->damping, c->desc.dt, c->desc.g_accel, c->d.reduce);
        else
            launch_stats(c->stream, c->d, c->cfg, c->damping, (float)c->desc.dt, (float)c->desc.g_accel);
        hip_check(hipGetLastError(), "stats launch");
        std::vector<double> red((size_t)c->reduce_blocks * 4);
        hip_check(hipMemcpyAsync(red.data(), c->d.reduce, red.size() * sizeof(double), hipMemcpyDeviceToHost,
                                 c->stream),
                  "D2H");
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        double r2 = 0, mg2 = 0, ip = 0, en = 0;
        for (uint32_t b = 0; b < c->reduce_blocks; b++) {
            r2 += red[4 * b];
            mg2 += red[4 * b + 1];
            ip += red[4 * b + 2];
            en += red[4 * b + 3];
        }
        st.residual_abs = std::sqrt(r2);
        st.residual_rel = mg2 > 0 ? std::sqrt(r2) / std::sqrt(mg2) : 0;
        st.incremental_potential = ip + en;
        st.num_colors = c->plan.num_colors;
        st.num_particles = c->d.n;
        st.num_tets = c->d.n_tets;
        st.num_ground = c->d.n_ground;
        st.steps_taken = c->steps_taken;
        {
            unsigned long long wc[2] = {0, 0};
            hip_check(hipMemcpy(wc, c->wio.counters, sizeof(wc), hipMemcpyDeviceToHost), "D2H");
            st.warm_start_predictions = wc[0];
            st.warm_start_fallbacks = wc[1];
        }
        if (c->rc) {
            ResCtl rc{};
            hip_check(hipMemcpy(&rc, c->rc, sizeof(ResCtl), hipMemcpyDeviceToHost), "D2H");
            st.iterations_total = rc.total_iters;
            st.last_step_iterations = rc.last_iters;
            st.vertex_updates = rc.vertex_solves;
            st.steps_at_budget = rc.budget_steps;
        } else {
            st.iterations_total = c->steps_taken * c->desc.solver_iters;
            st.last_step_iterations = c->steps_taken ? c->desc.solver_iters : 0;
            st.vertex_updates = c->vertex_updates;
        }
        st.last_step_ms = c->last_step_ms;
        *out = st;
        check_status(c);   // the statistics are filled either way
    });
}

uint32_t avbd_num_particles(avbd_ctx* c) { return c ? (uint32_t)c->mass.size() : 0; }

uint32_t avbd_num_energies(avbd_ctx* c, int type) {
    if (!c) return 0;
    if (type == AVBD_NEO_HOOKEAN) return (uint32_t)c->nh.size();
    if (type == AVBD_GROUND_COLLISION) return (uint32_t)c->ground.size();
    if (type == AVBD_COSSERAT_ROD) return (uint32_t)c->cos.size();
    if (type == AVBD_RIGID_BODY_GROUND_COLLISION) return (uint32_t)c->rground.size();
    if (type == AVBD_ONE_SIDED_FIXED_JOINT) return (uint32_t)c->joints.size();
    if (type == AVBD_FIXED_JOINT) return (uint32_t)c->joints2.size();
    return 0;
}

int avbd_get_launch_config(avbd_ctx* c, uint32_t* lanes, uint32_t* block, uint32_t* batch) {
    return guarded(c, [&] {
        if (lanes) *lanes = c->cfg.lanes;
        if (block) *block = c->cfg.block;
        if (batch) *batch = c->cfg.batch;
    });
}

int avbd_get_step_mode(avbd_ctx* c, uint32_t* persistent_grid, uint32_t* fused_chebyshev) {
    if (!c) return AVBD_E_ARG;
    if (persistent_grid) *persistent_grid = 0u;   // one launch per phase (the persistent step kernel is gone)
    if (fused_chebyshev) *fused_chebyshev = (c->fuse_cheb && (!c->dynamic || dyn_quiet(c))) ? 1u : 0u;
    return AVBD_OK;
}

int avbd_set_timing(avbd_ctx* c, int enable) {
    return guarded(c, [&] { c->timing = enable != 0; });
}

int avbd_get_broadphase_timing(avbd_ctx* c, double* avg_ms) {
    return guarded(c, [&] {
        if (avg_ms) *avg_ms = c->lbvh_avg_ms;
    });
}

int avbd_get_solve_timing(avbd_ctx* c, double* avg_ms, uint32_t* launches) {
    return guarded(c, [&] {
        if (avg_ms) *avg_ms = c->solve_avg_ms;
        if (launches) *launches = c->solve_launches;
    });
}

int avbd_time_solve_launches(avbd_ctx* c, uint32_t reps, double* avg_ms, uint32_t* launches) {
    return guarded(c, [&] {
        require_built(c);
        if (reps == 0) throw std::runtime_error("reps must be >= 1");
        uint32_t n = 0;
        auto round = [&] {
            for (uint32_t col = 0; col < c->plan.num_colors; col++) {
                const uint32_t b = c->plan.item_offsets[col], e = c->plan.item_offsets[col + 1];
                if (e > b) {
                    solve_into_scratch(c, c->cfg, b, e - b);
                    n++;
                }
            }
        };
        round();   // warm-up
        n = 0;
        hip_check(hipEventRecord(c->ev_a, c->stream), "hipEventRecord");
        for (uint32_t r = 0; r < reps; r++) round();
        hip_check(hipEventRecord(c->ev_b, c->stream), "hipEventRecord");
        hip_check(hipEventSynchronize(c->ev_b), "hipEventSynchronize");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, c->ev_a, c->ev_b), "hipEventElapsedTime");
        if (avg_ms) *avg_ms = n ? ms / n : 0.0;
        if (launches) *launches = n;
    });
}

int avbd_calibrate_copy(avbd_ctx* c, uint64_t bytes, uint32_t iters, double* avg_ms) {
    return guarded(c, [&] {
        require_built(c);
        const uint64_t n = bytes / sizeof(float4);
        if (n == 0 || iters == 0) throw std::runtime_error("bytes must be >= 16 and iters >= 1");
        float4 *src = nullptr, *dst = nullptr;
        hip_check(hipMalloc(&src, n * sizeof(float4)), "hipMalloc");
        if (hipMalloc(&dst, n * sizeof(float4)) != hipSuccess) {
            (void)hipFree(src);
            throw HipError("hipMalloc failed");
        }
        hipError_t e = hipMemsetAsync(src, 0, n * sizeof(float4), c->stream);
        if (e == hipSuccess) launch_calib_copy(c->stream, src, dst, n);   // warm-up
        if (e == hipSuccess) e = hipEventRecord(c->ev_a, c->stream);
        for (uint32_t i = 0; i < iters && e == hipSuccess; i++) launch_calib_copy(c->stream, src, dst, n);
        if (e == hipSuccess) e = hipEventRecord(c->ev_b, c->stream);
        if (e == hipSuccess) e = hipEventSynchronize(c->ev_b);
        float ms = 0;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev_a, c->ev_b);
        (void)hipFree(src);
        (void)hipFree(dst);
        hip_check(e, "calibration copy");
        if (avg_ms) *avg_ms = ms / iters;
    });
}

}  // extern "C"

namespace {
void ensure_model(avbd_ctx* c) {
    require_built(c);
    if (!c->mlp) c->mlp = std::make_unique<MlpModel>(c->stream);
}
void require_not_warm(avbd_ctx* c) {
    if (c->warm) throw CodeError(AVBD_E_ARG, "disable the warm start before training or evaluating the network");
}
}  // namespace

extern "C" {

uint64_t avbd_mlp_param_count(void) { return MlpModel::param_count(); }

int avbd_mlp_init(avbd_ctx* c, uint64_t seed) {
    return guarded(c, [&] {
        ensure_model(c);
        require_not_warm(c);
        c->mlp->init(seed);
    });
}

int avbd_mlp_set_params(avbd_ctx* c, const float* p, uint64_t n) {
    return guarded(c, [&] {
        ensure_model(c);
        require_not_warm(c);
        if (!p || n != MlpModel::param_count()) throw std::runtime_error("parameter vector has the wrong length");
        c->mlp->set_params(p);
    });
}

int avbd_mlp_get_params(avbd_ctx* c, float* p, uint64_t n) {
    return guarded(c, [&] {
        ensure_model(c);
        if (!p || n != MlpModel::param_count()) throw std::runtime_error("parameter vector has the wrong length");
        c->mlp->get_params(p);
    });
}

int avbd_mlp_set_normalization(avbd_ctx* c, const float* norm) {
    return guarded(c, [&] {
        ensure_model(c);
        require_not_warm(c);
        if (!norm) throw std::runtime_error("null normalisation");
        for (int i = 0; i < kMlpIn; i++)
            if (!(norm[kMlpIn + i] > 0)) throw std::runtime_error("feature std must be > 0");
        c->mlp->set_norm(norm);
    });
}

int avbd_mlp_forward_backward(avbd_ctx* c, const float* feats, const float* tgts, uint32_t rows, float* grads,
                              uint64_t n, float* loss) {
    return guarded(c, [&] {
        ensure_model(c);
        require_not_warm(c);
        if (!feats || !tgts || rows == 0) throw std::runtime_error("empty batch");
        const float l = c->mlp->forward_backward(feats, tgts, (int)rows);
        if (grads) {
            if (n != MlpModel::param_count() - 2 * kMlpIn - 2 * kMlpOut)
                throw std::runtime_error("gradient vector has the wrong length");
            c->mlp->get_grads(grads);
        }
        if (loss) *loss = l;
    });
}

int avbd_mlp_predict(avbd_ctx* c, const float* feats, uint32_t rows, float* out) {
    return guarded(c, [&] {
        ensure_model(c);
        require_not_warm(c);
        if (!feats || !out || rows == 0) throw std::runtime_error("empty batch");
        c->mlp->predict(feats, (int)rows, out);
    });
}

int avbd_mlp_train(avbd_ctx* c, const float* feats, const float* tgts, uint32_t rows, uint32_t epochs,
                   uint32_t batch, float lr, uint64_t seed, float* loss) {
    return guarded(c, [&] {
        ensure_model(c);
        require_not_warm(c);
        if (!feats || !tgts || rows < batch || batch == 0 || batch % 16) throw std::runtime_error("bad training set / batch");
        const float l = c->mlp->train(feats, tgts, (int)rows, (int)epochs, (int)batch, lr, seed);
        if (loss) *loss = l;
    });
}

int avbd_mlp_gemm_test(avbd_ctx* c, uint32_t M, uint32_t N, uint32_t K, int ta, int tb, const float* A,
                       const float* B, float* C) {
    return guarded(c, [&] {
        require_built(c);
        if (!A || !B || !C || M % 16 || N % 16 || K % 16 || M == 0 || N == 0 || K == 0 || (ta && tb))
            throw std::runtime_error("gemm test: sizes must be multiples of 16 and at most one operand transposed");
        // A is M x K and B is K x N (row-major FP32); stored as the kernel reads them
        std::vector<uint16_t> a16((size_t)M * K), b16((size_t)K * N);
        auto bf = [](float f) {
            uint32_t u;
            std::memcpy(&u, &f, 4);
            u += 0x7fffu + ((u >> 16) & 1u);
            return (uint16_t)(u >> 16);
        };
        for (uint32_t m = 0; m < M; m++)
            for (uint32_t k = 0; k < K; k++) a16[ta ? (size_t)k * M + m : (size_t)m * K + k] = bf(A[(size_t)m * K + k]);
        for (uint32_t k = 0; k < K; k++)
            for (uint32_t n = 0; n < N; n++) b16[tb ? (size_t)n * K + k : (size_t)k * N + n] = bf(B[(size_t)k * N + n]);
        uint16_t *da = nullptr, *db = nullptr;
        float* dc = nullptr;
        hip_check(hipMalloc(&da, a16.size() * 2), "hipMalloc");
        hipError_t e = hipMalloc(&db, b16.size() * 2);
        if (e == hipSuccess) e = hipMalloc(&dc, (size_t)M * N * 4);
        if (e == hipSuccess) e = hipMemcpy(da, a16.data(), a16.size() * 2, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(db, b16.data(), b16.size() * 2, hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            launch_gemm_bf16(c->stream, ta != 0, tb != 0, MLP_EPI_F32, da, ta ? (int)M : (int)K, db,
                             tb ? (int)K : (int)N, dc, nullptr, (int)N, nullptr, nullptr, 0, (int)M, (int)N, (int)K);
            e = hipStreamSynchronize(c->stream);
        }
        if (e == hipSuccess) e = hipMemcpy(C, dc, (size_t)M * N * 4, hipMemcpyDeviceToHost);
        (void)hipFree(da);
        (void)hipFree(db);
        (void)hipFree(dc);
        hip_check(e, "gemm test");
    });
}

int avbd_record_steps(avbd_ctx* c, uint32_t n_steps, float* feats, float* tgts, uint32_t cap_rows, uint32_t* rows) {
    return guarded(c, [&] {
        ensure_model(c);
        require_not_warm(c);
        if (c->dynamic) throw CodeError(AVBD_E_ARG, "recording is not available with dynamic contacts");
        const uint32_t ni = c->d.n_items;
        if ((uint64_t)n_steps * ni > cap_rows) throw std::runtime_error("record buffer too small");
        c->recording = true;
        std::vector<float> raw((size_t)ni * kMlpIn), tg((size_t)ni * 3);
        try {
            for (uint32_t st = 0; st < n_steps; st++) {
                // residual mode: intervals stop being queued once converged
                if (c->rc) run_residual_step(c, false);
                else enqueue_step(c, false);
                hip_check(hipMemcpyAsync(raw.data(), c->d_raw, raw.size() * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
                hip_check(hipMemcpyAsync(tg.data(), c->d_tgt, tg.size() * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
                hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
                check_status(c);
                std::memcpy(feats + (size_t)st * ni * kMlpIn, raw.data(), raw.size() * 4);
                std::memcpy(tgts + (size_t)st * ni * 3, tg.data(), tg.size() * 4);
                c->steps_taken++;
            }
        } catch (...) {
            c->recording = false;
            throw;
        }
        c->recording = false;
        if (rows) *rows = n_steps * ni;
    });
}

int avbd_get_warm_start_decision(avbd_ctx* c, double* ip_delta, float* step, uint32_t cap, uint32_t* n_scenes) {
    return guarded(c, [&] {
        require_built(c);
        if (!c->warm) throw CodeError(AVBD_E_ARG, "the warm start is off");
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        const uint32_t ns = c->wio.n_scenes;
        std::vector<double> ipd((size_t)kWarmAlphas * ns);
        std::vector<float> st(ns);
        hip_check(hipMemcpy(ipd.data(), c->wio.ipd, ipd.size() * sizeof(double), hipMemcpyDeviceToHost), "D2H");
        hip_check(hipMemcpy(st.data(), c->wio.step_of, ns * sizeof(float), hipMemcpyDeviceToHost), "D2H");
        for (uint32_t s = 0; s < ns && s < cap; s++) {
            if (ip_delta)
                for (int j = 0; j < kWarmAlphas; j++) ip_delta[(size_t)kWarmAlphas * s + j] = ipd[(size_t)j * ns + s];
            if (step) step[s] = st[s];
        }
        if (n_scenes) *n_scenes = ns;
    });
}

int avbd_set_warm_start(avbd_ctx* c, int enable) {
    return guarded(c, [&] {
        ensure_model(c);
        if (enable && !c->mlp->loaded) throw CodeError(AVBD_E_ARG, "the warm-start network has no trained parameters");
        if (enable) c->mlp->ensure_batch((int)c->rows_pad);
        c->warm = enable != 0;
        if (!(c->desc.flags & AVBD_FLAG_NO_GRAPH)) {
            capture_graph(c);
            hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        }
    });
}

}  // extern "C"


extern "C" {

int avbd_add_collision_triangles(avbd_ctx* c, uint32_t n, const uint32_t* tris) {
    return guarded(c, [&] {
        require_not_planned(c);
        if (n && !tris) throw std::runtime_error("tris is null");
        c->tris.insert(c->tris.end(), tris, tris + 3 * (size_t)n);
        const uint32_t obj = c->next_object++;
        for (uint32_t i = 0; i < n; i++)
            c->prims.push_back(HostPrim{PRIM_TRIANGLE, {tris[3 * i], tris[3 * i + 1], tris[3 * i + 2]}, obj, 0.0});
    });
}

int avbd_add_collision_segments(avbd_ctx* c, uint32_t n_nodes, const uint32_t* nodes, double radius) {
    return guarded(c, [&] {
        require_not_planned(c);
        if (n_nodes < 2 || !nodes) throw std::runtime_error("a rod needs at least two nodes");
        if (!(radius >= 0)) throw std::runtime_error("radius must be >= 0");
        const uint32_t obj = c->next_object++;
        for (uint32_t i = 0; i + 1 < n_nodes; i++)
            c->prims.push_back(HostPrim{PRIM_ROD_SEGMENT, {nodes[i], nodes[i + 1], 0u}, obj, radius});
    });
}

int avbd_set_self_collision(avbd_ctx* c, int enable) {
    return guarded(c, [&] {
        if (c->dynamic) throw CodeError(AVBD_E_ARG, "set self collision before avbd_set_dynamic_contacts");
        c->self_collision = enable != 0;
    });
}

int avbd_add_collision_sphere(avbd_ctx* c, uint32_t p, double radius) {
    return guarded(c, [&] {
        require_not_planned(c);
        if (!(radius > 0)) throw std::runtime_error("radius must be > 0");
        c->prims.push_back(HostPrim{PRIM_SPHERE, {p, 0u, 0u}, c->next_object++, radius});
    });
}

int avbd_set_dynamic_contacts(avbd_ctx* c, double margin, uint32_t capacity) {
    return guarded(c, [&] {
        require_built(c);
        if (c->dynamic) throw CodeError(AVBD_E_ARG, "dynamic contacts are already on");
        if (c->prims.empty()) throw CodeError(AVBD_E_ARG, "no collision primitives");
        if (!(margin >= 0)) throw CodeError(AVBD_E_ARG, "collision margin must be >= 0");
        if (c->recording) throw CodeError(AVBD_E_ARG, "recording is not available with dynamic contacts");
        // the scenes of a batch are independent: their primitives must not meet
        if (c->scene_offsets.size() > 2)
            throw CodeError(AVBD_E_ARG, "dynamic contacts take one scene, not a batch (avbd_set_scenes)");
        if (c->fp64)
            for (const HostPrim& pr : c->prims)
                if (pr.type != PRIM_TRIANGLE)
                    throw CodeError(AVBD_E_ARG, "the FP64 state takes collision triangles only (vertex-face contacts)");
        const uint32_t n = c->d.n;
        auto oriented = [&](uint32_t p) { return p < n && c->rot_of[p] >= 0; };
        for (const HostPrim& pr : c->prims) {
            if (pr.type == PRIM_TRIANGLE) {
                for (int k = 0; k < 3; k++)
                    if (pr.idx[k] >= n || oriented(pr.idx[k]))
                        throw CodeError(AVBD_E_ARG, "collision triangle needs three positional particles");
            } else if (pr.type == PRIM_ROD_SEGMENT) {
                if (!oriented(pr.idx[0]) || !oriented(pr.idx[1]))
                    throw CodeError(AVBD_E_ARG, "rod segment needs two oriented particles");
            } else if (!oriented(pr.idx[0])) {
                throw CodeError(AVBD_E_ARG, "collision sphere needs an oriented particle");
            }
        }
        const uint32_t cap = capacity ? capacity : 65536u;
        if ((uint64_t)DYN_KINDS * cap > (1ull << 29)) throw CodeError(AVBD_E_ARG, "capacity too large");
        const uint32_t np = (uint32_t)c->prims.size();
        destroy_graph(c);
        alloc_lbvh(c, c->lbd, np);
        std::vector<uint32_t> ty(np), id(3 * (size_t)np), ob(np);
        std::vector<double> rad(np);
        for (uint32_t i = 0; i < np; i++) {
            ty[i] = c->prims[i].type;
            for (int k = 0; k < 3; k++) id[3 * i + k] = c->prims[i].idx[k];
            ob[i] = c->prims[i].object;
            rad[i] = c->prims[i].radius;
        }
        c->dp.n = np;
        if (c->self_collision) {
            // owner of every particle: the lowest-numbered triangle holding it
            std::vector<uint32_t> owner(n, 0xFFFFFFFFu);
            for (uint32_t i = np; i-- > 0;)
                if (c->prims[i].type == PRIM_TRIANGLE)
                    for (int k = 0; k < 3; k++) owner[c->prims[i].idx[k]] = i;
            c->dp.owner = dev_alloc<uint32_t>(c, n);
            upload(c, c->dp.owner, owner);
            c->dp.vertex_face = 1;
            c->dp.tri_pad = margin;
        }
        c->dp.type = dev_alloc<uint32_t>(c, np);
        c->dp.idx = dev_alloc<uint32_t>(c, 3 * (size_t)np);
        c->dp.object = dev_alloc<uint32_t>(c, np);
        c->dp.radius = dev_alloc<double>(c, np);
        upload(c, c->dp.type, ty);
        upload(c, c->dp.idx, id);
        upload(c, c->dp.object, ob);
        upload(c, c->dp.radius, rad);
        c->det_cap = c->lbd.pair_cap;
        c->d_det = dev_alloc<DetRec>(c, c->det_cap);
        c->d_status = dev_alloc<uint32_t>(c, 2);
        c->d.dyn = dev_alloc<DynRec>(c, DYN_KINDS * (size_t)cap);
        if (c->fp64) c->f.x_buf = dev_alloc<double4>(c, n);   // conflicted particles' buffered positions
        c->cs.init(n, cap, margin);
        auto* d_off = static_cast<uint32_t*>(grow(c, c->g_off, (n + 1) * sizeof(uint32_t)));
        upload(c, d_off, c->cs.dyn_off);
        c->d.dyn_off = d_off;
        c->d.dyn_ent = static_cast<uint32_t*>(grow(c, c->g_ent, 16));
        // device recolouring state: the set-up colouring, the static
        // adjacency, each particle's work item; nothing dirty or conflicted
        RecolourDev& r = c->rdev;
        const ColorPlan& plan = c->plan;
        r.n = n;
        r.color = dev_alloc<uint32_t>(c, n);
        r.last_dirty = dev_alloc<uint32_t>(c, n);
        r.candidate = dev_alloc<uint32_t>(c, n);
        r.dirty = dev_alloc<uint32_t>(c, n);
        r.next = dev_alloc<uint32_t>(c, n);
        r.touched = dev_alloc<uint32_t>(c, 7 * (size_t)n);
        r.touched_conf = dev_alloc<uint32_t>(c, n);
        r.conflicted = dev_alloc<uint32_t>(c, n);
        r.counters = dev_alloc<uint32_t>(c, 8);
        r.offsets = dev_alloc<uint32_t>(c, 3 * 257);
        upload(c, r.color, plan.color);
        c->setup_color = plan.color;
        c->layout_changed = false;
        hip_check(hipMemsetAsync(r.last_dirty, 0xFF, (size_t)n * sizeof(uint32_t), c->stream), "hipMemsetAsync");
        hip_check(hipMemsetAsync(r.conflicted, 0, (size_t)n * sizeof(uint32_t), c->stream), "hipMemsetAsync");
        const std::vector<uint32_t> cnt0{0u, 0u, plan.num_colors, 0u, 0u, 0u, 0u, 0u};
        upload(c, r.counters, cnt0);
        uint32_t* s_off = dev_alloc<uint32_t>(c, plan.static_off.size());
        uint32_t* s_nb = dev_alloc<uint32_t>(c, std::max<size_t>(plan.static_nb.size(), 1));
        upload(c, s_off, plan.static_off);
        upload(c, s_nb, plan.static_nb);
        r.s_off = s_off;
        r.s_nb = s_nb;
        uint8_t* kind = dev_alloc<uint8_t>(c, n);
        VItem* item_of = dev_alloc<VItem>(c, n);
        OItem* oitem_of = dev_alloc<OItem>(c, n);
        upload(c, kind, plan.item_kind);
        upload(c, item_of, plan.item_of);
        upload(c, oitem_of, plan.oitem_of);
        r.kind = kind;
        r.item_of = item_of;
        r.oitem_of = oitem_of;
        c->d_conflicts = dev_alloc<uint32_t>(c, n);   // at most every particle is conflicted
        if (!c->h_layout) {
            void* hl = nullptr;
            hip_check(hipHostMalloc(&hl, (3 * 257 + 8) * sizeof(uint32_t), hipHostMallocDefault), "hipHostMalloc");
            c->h_layout = static_cast<uint32_t*>(hl);
        }
        // the per-colour offsets of the set-up layout (later rounds keep them
        // when nothing changes)
        launch_relayout(c->stream, r, c->d.items, c->d.oitems, c->d_conflicts, true);
        hip_check(hipGetLastError(), "kernel launch");
        c->dynamic = true;
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    });
}

int avbd_download_contacts(avbd_ctx* c, uint32_t kind, avbd_contact* out, uint32_t cap, uint32_t* count) {
    return guarded(c, [&] {
        require_built(c);
        if (!c->dynamic) throw CodeError(AVBD_E_ARG, "dynamic contacts are off");
        if (kind >= DYN_KINDS) throw CodeError(AVBD_E_ARG, "kind must be 0, 1, 2 or 3");
        const auto& P = c->cs.pool[kind];
        std::vector<DynRec> recs(c->cs.cap);
        hip_check(hipMemcpyAsync(recs.data(), c->d.dyn + (size_t)kind * c->cs.cap, recs.size() * sizeof(DynRec),
                                 hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        uint32_t k = 0;
        for (uint32_t sl = 0; sl < P.active.size(); sl++) {
            if (!P.active[sl]) continue;
            if (out && k < cap) {
                const DynRec& r = recs[sl];
                avbd_contact& o = out[k];
                std::memset(&o, 0, sizeof(o));
                o.slot = sl;
                for (int i = 0; i < 5; i++) o.idx[i] = r.idx[i];
                const float4* v4[6] = {&r.n, &r.t, &r.b, &r.bary, &r.cpl, nullptr};
                double* dst[5] = {o.n, o.t, o.b, o.barys, o.cp_local};
                for (int a = 0; a < 5; a++) {
                    dst[a][0] = v4[a]->x;
                    dst[a][1] = v4[a]->y;
                    dst[a][2] = v4[a]->z;
                }
                o.s = r.bary.w;
                o.k_start = r.n.w;
                o.k_n = r.k.x;
                o.k_t = r.k.y;
                o.k_b = r.k.z;
                o.lambda_n = r.lam.x;
                o.lambda_t = r.lam.y;
                o.lambda_b = r.lam.z;
                o.C_n_prev = r.cpl.w;
                const uint32_t fl = r.kind_flags >> DYN_FLAG_SHIFT;
                o.use_static = (fl & GF_USE_STATIC) ? 1 : 0;
                o.unclamped = (fl & GF_UNCLAMPED) ? 1 : 0;
            }
            k++;
        }
        if (count) *count = k;
    });
}

int avbd_download_detected(avbd_ctx* c, uint64_t* keys, uint32_t* slots, uint8_t* kinds, uint32_t cap,
                           uint32_t* count) {
    return guarded(c, [&] {
        require_built(c);
        if (!c->dynamic) throw CodeError(AVBD_E_ARG, "dynamic contacts are off");
        const auto& E = c->cs.prev;
        for (uint32_t i = 0; i < E.size() && i < cap; i++) {
            if (keys) keys[i] = E[i].key;
            if (slots) slots[i] = E[i].slot;
            if (kinds)
                kinds[i] = (uint8_t)(E[i].kind == DYN_TRI_RIGID    ? DET_KIND_TRI_RIGID
                                     : E[i].kind == DYN_TRI_ROD    ? DET_KIND_TRI_ROD
                                     : E[i].kind == DYN_VERTEX_TRI ? DET_KIND_VERTEX_FACE
                                                                   : DET_KIND_RIGID_RIGID);
        }
        if (count) *count = (uint32_t)E.size();
    });
}

}  // extern "C"

namespace {
// Checkpoint layout: header, then the device arrays in a fixed order, then
// (dynamic contacts) the contact records, the colouring and the host
// contact state.
struct CkptHeader {
    char magic[8];
    uint32_t version, n, n_tets, n_ground, n_rot, n_rground, n_joints;
    uint32_t dynamic, has_damp, has_rc, num_colors;
    uint32_t fp64, dyn_cap;   // FP64 state; contact records per kind
    uint64_t steps_taken, vertex_updates;
    uint64_t array_bytes;     // device arrays that follow the header
};
constexpr uint32_t kCkptVersion = 4;
constexpr char kCkptMagic[8] = {'A', 'V', 'B', 'D', 'C', 'K', 'P', 'T'};

// (device pointer, bytes) of every mutable device array, in checkpoint order
std::vector<std::pair<void*, size_t>> ckpt_arrays(avbd_ctx* c) {
    const DevState& d = c->d;
    std::vector<std::pair<void*, size_t>> a;
    const size_t p4 = (size_t)d.n * sizeof(float4), r4 = (size_t)d.n_rot * sizeof(float4);
    for (float4* v : {d.x, d.x_buf, d.y, d.x_prev, d.x_last, d.x_lastlast, d.v, d.v_prev}) a.emplace_back(v, p4);
    if (c->damping) a.emplace_back(d.damp, (size_t)d.n_tets * sizeof(TetDamp));
    a.emplace_back(d.ground, (size_t)d.n_ground * sizeof(GroundRec));
    if (d.n_rot)
        for (float4* v : {d.q, d.q_buf, d.q_in, d.q_prev, d.q_last, d.q_lastlast, d.w, d.w_prev}) a.emplace_back(v, r4);
    if (d.n_rground) a.emplace_back(d.rground, (size_t)d.n_rground * sizeof(RGroundRec));
    if (d.n_joints) a.emplace_back(d.joints, (size_t)d.n_joints * sizeof(JointRec));
    if (c->rc) a.emplace_back(c->rc, sizeof(ResCtl));
    if (c->dynamic) {
        a.emplace_back(d.dyn, DYN_KINDS * (size_t)c->cs.cap * sizeof(DynRec));
        // the device recolouring state (colours, dirty marks and list, conflicts)
        const RecolourDev& r = c->rdev;
        const size_t pu = (size_t)d.n * sizeof(uint32_t);
        for (uint32_t* v : {r.color, r.last_dirty, r.dirty, r.touched_conf, r.conflicted}) a.emplace_back(v, pu);
        a.emplace_back(r.counters, 8 * sizeof(uint32_t));
    }
    if (c->fp64) {
        const size_t p8 = (size_t)d.n * sizeof(double4);
        for (double4* v : {c->f.x, c->f.y, c->f.x_prev, c->f.x_last, c->f.x_lastlast}) a.emplace_back(v, p8);
    }
    return a;
}

CkptHeader ckpt_header(avbd_ctx* c) {
    CkptHeader h{};
    std::memcpy(h.magic, kCkptMagic, 8);
    h.version = kCkptVersion;
    h.n = c->d.n;
    h.n_tets = c->d.n_tets;
    h.n_ground = c->d.n_ground;
    h.n_rot = c->d.n_rot;
    h.n_rground = c->d.n_rground;
    h.n_joints = c->d.n_joints;
    h.dynamic = c->dynamic ? 1u : 0u;
    h.has_damp = c->damping ? 1u : 0u;
    h.has_rc = c->rc ? 1u : 0u;
    h.num_colors = c->plan.num_colors;
    h.fp64 = c->fp64 ? 1u : 0u;
    h.dyn_cap = c->dynamic ? c->cs.cap : 0u;
    h.steps_taken = c->steps_taken;
    h.vertex_updates = c->vertex_updates;
    for (const auto& a : ckpt_arrays(c)) h.array_bytes += a.second;
    return h;
}

std::vector<uint8_t> ckpt_host_tail(avbd_ctx* c) {
    std::vector<uint8_t> t;
    if (c->dynamic) c->cs.save(t);
    return t;
}
}  // namespace

extern "C" {

int avbd_checkpoint_size(avbd_ctx* c, uint64_t* bytes) {
    return guarded(c, [&] {
        require_built(c);
        uint64_t total = sizeof(CkptHeader);
        for (const auto& a : ckpt_arrays(c)) total += a.second;
        total += ckpt_host_tail(c).size();
        if (bytes) *bytes = total;
    });
}

int avbd_save_checkpoint(avbd_ctx* c, void* buf, uint64_t cap, uint64_t* written) {
    return guarded(c, [&] {
        require_built(c);
        if (!buf) throw CodeError(AVBD_E_ARG, "buf is null");
        const auto arrays = ckpt_arrays(c);
        const std::vector<uint8_t> tail = ckpt_host_tail(c);
        uint64_t total = sizeof(CkptHeader) + tail.size();
        for (const auto& a : arrays) total += a.second;
        if (cap < total) throw CodeError(AVBD_E_CAPACITY, "checkpoint buffer too small (see avbd_checkpoint_size)");
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        uint8_t* p = static_cast<uint8_t*>(buf);
        const CkptHeader h = ckpt_header(c);
        std::memcpy(p, &h, sizeof(h));
        p += sizeof(h);
        for (const auto& a : arrays) {
            hip_check(hipMemcpyAsync(p, a.first, a.second, hipMemcpyDeviceToHost, c->stream), "D2H");
            p += a.second;
        }
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        if (!tail.empty()) std::memcpy(p, tail.data(), tail.size());
        if (written) *written = total;
    });
}

int avbd_load_checkpoint(avbd_ctx* c, const void* buf, uint64_t size) {
    return guarded(c, [&] {
        require_built(c);
        if (!buf || size < sizeof(CkptHeader)) throw CodeError(AVBD_E_ARG, "not a checkpoint");
        const uint8_t* p = static_cast<const uint8_t*>(buf);
        const uint8_t* end = p + size;
        CkptHeader h;
        std::memcpy(&h, p, sizeof(h));
        p += sizeof(h);
        const CkptHeader mine = ckpt_header(c);
        if (std::memcmp(h.magic, kCkptMagic, 8) != 0 || h.version != kCkptVersion)
            throw CodeError(AVBD_E_ARG, "not a checkpoint");
        // everything is validated on the host before any part of the context changes
        if (h.n != mine.n || h.n_tets != mine.n_tets || h.n_ground != mine.n_ground || h.n_rot != mine.n_rot ||
            h.n_rground != mine.n_rground || h.n_joints != mine.n_joints || h.dynamic != mine.dynamic ||
            h.has_damp != mine.has_damp || h.has_rc != mine.has_rc || h.fp64 != mine.fp64 ||
            h.dyn_cap != mine.dyn_cap || h.array_bytes != mine.array_bytes ||
            (!c->dynamic && h.num_colors != mine.num_colors))
            throw CodeError(AVBD_E_ARG, "checkpoint does not match this context's scene and modes");
        if ((uint64_t)(end - p) < h.array_bytes) throw CodeError(AVBD_E_ARG, "checkpoint is truncated");
        const uint8_t* arrays = p;
        p += h.array_bytes;
        ContactSystem cs;
        if (c->dynamic) {
            cs = c->cs;
            cs.load(p, end);   // throws on a malformed tail, before anything is applied
        }
        // apply
        for (const auto& a : ckpt_arrays(c)) {
            hip_check(hipMemcpyAsync(a.first, arrays, a.second, hipMemcpyHostToDevice, c->stream), "H2D");
            arrays += a.second;
        }
        if (c->dynamic) {
            c->cs = std::move(cs);
            // the work list of the restored colouring, laid out on the device
            launch_relayout(c->stream, c->rdev, c->d.items, c->d.oitems, c->d_conflicts, true);
            hip_check(hipGetLastError(), "kernel launch");
            read_layout(c);
            // the restored work list is the set-up one only if the colours are
            // and nothing is conflicted
            std::vector<uint32_t> col(c->d.n);
            hip_check(hipMemcpy(col.data(), c->rdev.color, col.size() * sizeof(uint32_t), hipMemcpyDeviceToHost), "D2H");
            c->layout_changed = col != c->setup_color || c->plan.conflict_offsets.back() != 0;
            auto* d_act = static_cast<uint32_t*>(grow(c, c->g_active, c->cs.active.size() * sizeof(uint32_t)));
            upload(c, d_act, c->cs.active);
            auto* d_off = static_cast<uint32_t*>(grow(c, c->g_off, c->cs.dyn_off.size() * sizeof(uint32_t)));
            upload(c, d_off, c->cs.dyn_off);
            auto* d_ent = static_cast<uint32_t*>(grow(c, c->g_ent, c->cs.dyn_ent.size() * sizeof(uint32_t)));
            upload(c, d_ent, c->cs.dyn_ent);
            c->d.dyn_off = d_off;
            c->d.dyn_ent = d_ent;
        }
        c->steps_taken = h.steps_taken;
        c->vertex_updates = h.vertex_updates;
        clear_status(c);
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    });
}

int avbd_download_conflicts(avbd_ctx* c, uint8_t* flags) {
    return guarded(c, [&] {
        require_built(c);
        if (!flags) throw CodeError(AVBD_E_ARG, "flags is null");
        const uint32_t n = c->d.n;
        if (c->dynamic) {
            std::vector<uint32_t> f(n);
            hip_check(hipMemcpyAsync(f.data(), c->rdev.conflicted, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                     c->stream),
                      "D2H");
            hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
            for (uint32_t p = 0; p < n; p++) flags[p] = f[p] ? 1u : 0u;
        } else {
            std::memset(flags, 0, n);
        }
    });
}

int avbd_detect_collisions(avbd_ctx* c) {
    return guarded(c, [&] {
        require_built(c);
        if (!c->lb.n) throw CodeError(AVBD_E_ARG, "no collision primitives");
        launch_lbvh_build(c->stream, c->lb, c->d.x, c->d.v, c->desc.dt);
        hip_check(hipGetLastError(), "lbvh launch");
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    });
}

int avbd_download_bvh(avbd_ctx* c, uint64_t* morton, uint32_t* sorted, uint32_t* left, uint32_t* right,
                      uint32_t* parent, double* node_box, uint32_t* root) {
    return guarded(c, [&] {
        require_built(c);
        const LbvhDev& b = c->lb;
        if (!b.n) throw CodeError(AVBD_E_ARG, "no collision primitives");
        const size_t nn = 2 * (size_t)b.n - 1;
        auto d2h = [&](void* dst, const void* src, size_t bytes) {
            if (dst) hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream), "D2H");
        };
        d2h(morton, b.keys, b.n * sizeof(uint64_t));
        d2h(sorted, b.sorted, b.n * sizeof(uint32_t));
        d2h(left, b.left, nn * sizeof(uint32_t));
        d2h(right, b.right, nn * sizeof(uint32_t));
        d2h(parent, b.parent, nn * sizeof(uint32_t));
        d2h(node_box, b.node_box, 6 * nn * sizeof(double));
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        if (root) *root = b.n > 1 ? b.n : 0;
    });
}

int avbd_download_candidates(avbd_ctx* c, uint32_t* pairs, uint32_t cap, uint32_t* count) {
    return guarded(c, [&] {
        require_built(c);
        const LbvhDev& b = c->lb;
        if (!b.n) throw CodeError(AVBD_E_ARG, "no collision primitives");
        uint32_t ctr[2] = {0, 0};
        hip_check(hipMemcpyAsync(ctr, b.counters, sizeof(ctr), hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        if (ctr[1] & 1u) throw CodeError(AVBD_E_CAPACITY, "LBVH traversal stack overflow");
        const uint32_t stored = std::min(ctr[0], b.pair_cap);
        const uint32_t m = std::min(stored, cap);
        if (pairs && m)
            hip_check(hipMemcpyAsync(pairs, b.pairs, 2 * (size_t)m * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                     c->stream),
                      "D2H");
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        if (count) *count = ctr[0];
        if (ctr[0] > b.pair_cap) throw CodeError(AVBD_E_CAPACITY, "candidate pair buffer exceeded");
    });
}

}  // extern "C"
