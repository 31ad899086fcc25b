// capi.cpp -- C entry points of the FP64 oracle for ctypes (test
// infrastructure only: tests/, smoke() and bench.py's cpu_baseline leg).
#include "simulation.hpp"

#include <cstring>
#include <exception>
#include <string>

using namespace oracle;

namespace {
thread_local std::string g_err;
void m3_out(const M3& m, double* o) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) o[3 * i + j] = m.m[i][j];
}
template <class F>
int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}
}  // namespace

extern "C" {

const char* oracle_last_error() { return g_err.c_str(); }

void* oracle_new(double dt, double g, uint32_t iters, double rho) {
    auto* s = new Simulation();
    s->params.dt = dt;
    s->params.g = g;
    s->params.iters = iters;
    s->params.rho = rho;
    return s;
}

void oracle_free(void* h) { delete static_cast<Simulation*>(h); }

// Residual-targeted iteration (tol = 0: fixed iteration count).
void oracle_set_residual_target(void* h, double tol, uint32_t max_iters, uint32_t check_every) {
    auto* s = static_cast<Simulation*>(h);
    s->params.tol = tol;
    s->params.max_iters = max_iters;
    s->params.check_every = check_every;
}
uint32_t oracle_last_iters(void* h) { return static_cast<Simulation*>(h)->last_iters; }
uint64_t oracle_total_iters(void* h) { return static_cast<Simulation*>(h)->total_iters; }

int64_t oracle_add_tet_mesh(void* h, uint32_t nv, const double* rest_x, uint32_t nt, const uint32_t* tets,
                            double E, double nu, double density, double kd, int with_ground,
                            double mu_s, double mu_k) {
    int64_t base = -1;
    int rc = guard([&] {
        base = static_cast<Simulation*>(h)->add_tet_mesh(nv, rest_x, nt, tets, E, nu, density, kd,
                                                          with_ground != 0, mu_s, mu_k);
    });
    return rc == 0 ? base : -1;
}

uint32_t oracle_add_particles(void* h, uint32_t n, const double* x, const double* m) {
    return static_cast<Simulation*>(h)->add_particles(n, x, m);
}

// Replace the collision triangles by one mesh object of n triangles.
void oracle_set_collision_triangles(void* h, uint32_t n, const uint32_t* tris) {
    auto* s = static_cast<Simulation*>(h);
    s->tris.clear();
    s->prims.clear();
    s->next_object = 0;
    s->add_collision_triangles(n, tris);
}
// CollisionPrimitivePool::addObject, one object per call; primitives are
// numbered in call order.
uint32_t oracle_add_collision_triangles(void* h, uint32_t n, const uint32_t* tris) {
    return static_cast<Simulation*>(h)->add_collision_triangles(n, tris);
}
uint32_t oracle_add_collision_segments(void* h, uint32_t n_nodes, const uint32_t* nodes, double radius) {
    return static_cast<Simulation*>(h)->add_collision_segments(n_nodes, nodes, radius);
}
uint32_t oracle_add_collision_sphere(void* h, uint32_t p, double radius) {
    return static_cast<Simulation*>(h)->add_collision_sphere(p, radius);
}
uint32_t oracle_num_prims(void* h) { return (uint32_t)static_cast<Simulation*>(h)->prims.size(); }
void oracle_set_self_collision(void* h, int enable) { static_cast<Simulation*>(h)->self_collision = enable != 0; }
void oracle_enable_dynamic_contacts(void* h, double margin, uint32_t capacity) {
    static_cast<Simulation*>(h)->enable_dynamic_contacts(margin, capacity);
}
// Run the per-step detection + lifecycle + recolouring alone (tests).
int oracle_detect_and_recolor(void* h) {
    return guard([&] { static_cast<Simulation*>(h)->detect_and_recolor(); });
}
uint32_t oracle_num_contacts(void* h, uint32_t type) {
    auto* s = static_cast<Simulation*>(h);
    return type < DYN_COUNT ? s->dyn[type].sp.count : 0;
}
// Active contacts of one dynamic pool in slot order, 32 doubles each:
// slot, idx[5], n, t, b, barys (rigid-rigid: cpl2), s, cpl, k_start, k_n,
// k_t, k_b, lambda_n, lambda_t, lambda_b, C_n_prev, use_static, unclamped.
uint32_t oracle_get_contacts(void* h, uint32_t type, double* rec32, uint32_t cap) {
    auto* s = static_cast<Simulation*>(h);
    if (type >= DYN_COUNT) return 0;
    const DynPool& P = s->dyn[type];
    uint32_t k = 0;
    for (uint32_t e = 0; e < P.sp.active.size(); e++) {
        if (!P.sp.active[e]) continue;
        if (k < cap) {
            double* r = rec32 + 32 * (size_t)k;
            const DynContact& c = P.data[e];
            r[0] = e;
            for (int i = 0; i < 5; i++) r[1 + i] = c.idx[i];
            for (int i = 0; i < 3; i++) {
                r[6 + i] = c.n[i];
                r[9 + i] = c.t[i];
                r[12 + i] = c.b[i];
                r[15 + i] = type == DYN_RIGID_RIGID ? c.cpl2[i] : c.barys[i];
                r[19 + i] = c.cpl[i];
            }
            r[18] = c.s;
            r[22] = c.k_start; r[23] = c.k_n; r[24] = c.k_t; r[25] = c.k_b;
            r[26] = c.lambda_n; r[27] = c.lambda_t; r[28] = c.lambda_b; r[29] = c.C_n_prev;
            r[30] = c.use_static; r[31] = c.unclamped;
        }
        k++;
    }
    return k;
}
// Contacts detected by the last detection, in key order: key, pool slot, type
// (DetectedCollisionType).
uint32_t oracle_get_detected(void* h, uint64_t* keys, uint32_t* slots, uint8_t* types, uint32_t cap) {
    auto* s = static_cast<Simulation*>(h);
    const uint32_t n = (uint32_t)s->cur_det.size();
    for (uint32_t i = 0; i < n && i < cap; i++) {
        const Detected& d = s->cur_det[s->cur_order[i]];
        keys[i] = d.key;
        slots[i] = d.e_idx;
        types[i] = d.type;
    }
    return n;
}
// Conflicted particles (is_conflicted) after the last recolouring.
uint32_t oracle_get_conflicted(void* h, uint8_t* flags) {
    auto* s = static_cast<Simulation*>(h);
    uint32_t k = 0;
    for (size_t p = 0; p < s->is_conflicted.size(); p++) {
        flags[p] = s->is_conflicted[p];
        k += s->is_conflicted[p];
    }
    return k;
}
uint64_t oracle_last_candidates(void* h) { return static_cast<Simulation*>(h)->last_candidates; }

uint32_t oracle_num_particles(void* h) { return (uint32_t)static_cast<Simulation*>(h)->x.size(); }
uint32_t oracle_num_tets(void* h) { return (uint32_t)static_cast<Simulation*>(h)->nh.size(); }
uint32_t oracle_num_ground(void* h) { return (uint32_t)static_cast<Simulation*>(h)->ground.size(); }

// Overwrite positions (and previous positions) of particles [0, n).
void oracle_set_positions(void* h, const double* x, int also_prev) {
    auto* s = static_cast<Simulation*>(h);
    for (size_t i = 0; i < s->x.size(); i++) {
        V3 p(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
        s->x[i] = p;
        if (also_prev) s->x_prev[i] = p;
    }
}
void oracle_set_velocities(void* h, const double* v, const double* v_prev) {
    auto* s = static_cast<Simulation*>(h);
    for (size_t i = 0; i < s->x.size(); i++) {
        s->v[i] = V3(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
        s->v_prev[i] = V3(v_prev[3 * i], v_prev[3 * i + 1], v_prev[3 * i + 2]);
    }
}
void oracle_set_fixed(void* h, const uint8_t* fixed) {
    auto* s = static_cast<Simulation*>(h);
    for (size_t i = 0; i < s->x.size(); i++) s->fixed[i] = fixed[i];
}

int oracle_build_graph(void* h) {
    return guard([&] { static_cast<Simulation*>(h)->build_graph(); });
}

int oracle_step(void* h, uint32_t n_steps, uint32_t n_threads) {
    return guard([&] { static_cast<Simulation*>(h)->step_mt(n_threads, n_steps); });
}

void oracle_get_positions(void* h, double* out) {
    auto* s = static_cast<Simulation*>(h);
    for (size_t i = 0; i < s->x.size(); i++) {
        out[3 * i] = s->x[i].x; out[3 * i + 1] = s->x[i].y; out[3 * i + 2] = s->x[i].z;
    }
}
void oracle_get_velocities(void* h, double* out) {
    auto* s = static_cast<Simulation*>(h);
    for (size_t i = 0; i < s->x.size(); i++) {
        out[3 * i] = s->v[i].x; out[3 * i + 1] = s->v[i].y; out[3 * i + 2] = s->v[i].z;
    }
}
void oracle_get_inertial(void* h, double* out) {
    auto* s = static_cast<Simulation*>(h);
    for (size_t i = 0; i < s->x.size(); i++) {
        out[3 * i] = s->y[i].x; out[3 * i + 1] = s->y[i].y; out[3 * i + 2] = s->y[i].z;
    }
}
void oracle_get_masses(void* h, double* out) {
    auto* s = static_cast<Simulation*>(h);
    for (size_t i = 0; i < s->x.size(); i++) out[i] = s->mass[i];
}

// Tet records: per tet 4 indices, then Q (9, row-major), V, lambda, mu, kd, E_prev (9).
void oracle_get_tets(void* h, uint32_t* idx, double* Q, double* V, double* lam, double* mu, double* kd,
                     double* Eprev) {
    auto* s = static_cast<Simulation*>(h);
    for (size_t e = 0; e < s->nh.size(); e++) {
        const auto& t = s->nh[e];
        for (int k = 0; k < 4; k++) idx[4 * e + k] = t.idx[k];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                Q[9 * e + 3 * i + j] = t.Q.m[i][j];
                Eprev[9 * e + 3 * i + j] = t.E_prev.m[i][j];
            }
        V[e] = t.V; lam[e] = t.lambda; mu[e] = t.mu; kd[e] = t.kd;
    }
}
void oracle_set_tet_eprev(void* h, const double* Eprev) {
    auto* s = static_cast<Simulation*>(h);
    for (size_t e = 0; e < s->nh.size(); e++)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) s->nh[e].E_prev.m[i][j] = Eprev[9 * e + 3 * i + j];
}

// Ground records: k_start,k_n,k_t,k_b,lambda_n,lambda_t,lambda_b,C_n_prev,mu_s,mu_k,cp_x,cp_z,use_static,unclamped
void oracle_get_ground(void* h, uint32_t* p, double* rec14) {
    auto* s = static_cast<Simulation*>(h);
    for (size_t e = 0; e < s->ground.size(); e++) {
        const auto& c = s->ground[e];
        p[e] = c.p;
        double* r = rec14 + 14 * e;
        r[0] = c.k_start; r[1] = c.k_n; r[2] = c.k_t; r[3] = c.k_b;
        r[4] = c.lambda_n; r[5] = c.lambda_t; r[6] = c.lambda_b; r[7] = c.C_n_prev;
        r[8] = c.mu_s; r[9] = c.mu_k; r[10] = c.cp_x; r[11] = c.cp_z;
        r[12] = c.use_static ? 1 : 0; r[13] = c.unclamped_lambda_tb ? 1 : 0;
    }
}
void oracle_set_ground(void* h, const double* rec14) {
    auto* s = static_cast<Simulation*>(h);
    for (size_t e = 0; e < s->ground.size(); e++) {
        auto& c = s->ground[e];
        const double* r = rec14 + 14 * e;
        c.k_start = r[0]; c.k_n = r[1]; c.k_t = r[2]; c.k_b = r[3];
        c.lambda_n = r[4]; c.lambda_t = r[5]; c.lambda_b = r[6]; c.C_n_prev = r[7];
        c.mu_s = r[8]; c.mu_k = r[9]; c.cp_x = r[10]; c.cp_z = r[11];
        c.use_static = r[12] != 0; c.unclamped_lambda_tb = r[13] != 0;
    }
}

// One ground contact on particle p with the initial values of
// CollisionConstraintEnergyPool::addEnergy (CollisionConstraintEnergyPool.hpp:48-69).
void oracle_add_ground(void* h, uint32_t p, double k_start, double mu_s, double mu_k) {
    auto* s = static_cast<Simulation*>(h);
    GroundContact c{};
    c.p = p;
    c.k_start = c.k_n = c.k_t = c.k_b = k_start;
    c.mu_s = mu_s;
    c.mu_k = mu_k;
    c.use_static = false;
    c.unclamped_lambda_tb = false;
    s->ground.push_back(c);
}

uint32_t oracle_num_colors(void* h) { return static_cast<Simulation*>(h)->coloring.num_colors; }
void oracle_get_colors(void* h, uint32_t* color) {
    auto* s = static_cast<Simulation*>(h);
    std::memcpy(color, s->coloring.color.data(), s->coloring.color.size() * sizeof(uint32_t));
}
uint32_t oracle_num_adj_entries(void* h) { return (uint32_t)static_cast<Simulation*>(h)->adj.e_entries.size(); }
// e_offsets: (n+1)*5, entries: type/e_idx/local triplets
void oracle_get_adjacency(void* h, uint32_t* e_offsets, uint32_t* entries, uint32_t* p_offsets,
                          uint32_t* p_neighbors) {
    auto* s = static_cast<Simulation*>(h);
    const auto& a = s->adj;
    for (size_t i = 0; i < a.e_offsets.size(); i++)
        for (int t = 0; t < ST_COUNT; t++) e_offsets[ST_COUNT * i + t] = a.e_offsets[i][t];
    for (size_t k = 0; k < a.e_entries.size(); k++) {
        entries[3 * k] = a.e_entries[k].type;
        entries[3 * k + 1] = a.e_entries[k].e_idx;
        entries[3 * k + 2] = a.e_entries[k].local;
    }
    if (p_offsets) std::memcpy(p_offsets, a.p_offsets.data(), a.p_offsets.size() * sizeof(uint32_t));
    if (p_neighbors) std::memcpy(p_neighbors, a.p_neighbors.data(), a.p_neighbors.size() * sizeof(uint32_t));
}
uint32_t oracle_num_p_neighbors(void* h) { return (uint32_t)static_cast<Simulation*>(h)->adj.p_neighbors.size(); }

// Standalone DSATUR on a caller-provided CSR graph (for hand-made cases).
uint32_t oracle_dsatur_csr(uint32_t n, const uint32_t* p_offsets, const uint32_t* p_neighbors, uint32_t* color_out) {
    StaticAdjacency adj;
    adj.p_offsets.assign(p_offsets, p_offsets + n + 1);
    adj.p_neighbors.assign(p_neighbors, p_neighbors + p_offsets[n]);
    Coloring c;
    dsatur(adj, n, c);
    std::memcpy(color_out, c.color.data(), n * sizeof(uint32_t));
    return c.num_colors;
}

void oracle_stats(void* h, double* out3) {
    Stats st = static_cast<Simulation*>(h)->stats();
    out3[0] = st.residual_abs;
    out3[1] = st.residual_rel;
    out3[2] = st.ip;
}

// Single-tet kernels for unit tests.  x: 4 positions (12), Q row-major (9),
// Eprev (9).  Returns the energy; writes G (3) and H (9, row-major) for `local`.
double oracle_nh_single(const double* x12, const double* Q9, const double* Eprev9, double V, double lambda,
                        double mu, double kd, double dt, uint32_t local, double* G3, double* H9, int use4) {
    std::vector<V3> x(4);
    for (int i = 0; i < 4; i++) x[i] = V3(x12[3 * i], x12[3 * i + 1], x12[3 * i + 2]);
    NHTet t;
    for (int k = 0; k < 4; k++) t.idx[k] = k;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            t.Q.m[i][j] = Q9[3 * i + j];
            t.E_prev.m[i][j] = Eprev9[3 * i + j];
        }
    t.V = V; t.lambda = lambda; t.mu = mu; t.kd = kd;
    V3 G; M3 H;
    if (use4) {
        // the same incidence four times through the 4-wide path, divided by 4
        // (the reference's own check, test/test_neohookean_accumulate.cpp:51-84)
        const NHTet* t4[4] = {&t, &t, &t, &t};
        uint32_t l4[4] = {local, local, local, local};
        nh_accumulate4(t4, x, l4, G, H, dt);
        G = 0.25 * G;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) H.m[i][j] *= 0.25;
    } else {
        nh_accumulate(t, x, local, G, H, dt);
    }
    G3[0] = G.x; G3[1] = G.y; G3[2] = G.z;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) H9[3 * i + j] = H.m[i][j];
    return nh_energy(t, x, dt);
}

void oracle_lu_solve(int n, const double* A, const double* b, double* x) { lu_solve(n, A, b, x); }

// ---------------------------------------------------------------------------
// Oriented (6-DOF) particles.  Quaternions are passed as (w, x, y, z).
// ---------------------------------------------------------------------------
static inline Quat q_in(const double* q) { return Quat(q[0], q[1], q[2], q[3]); }
static inline V3 v_in(const double* v) { return V3(v[0], v[1], v[2]); }
static inline void q_out(const Quat& q, double* o) { o[0] = q.w; o[1] = q.x; o[2] = q.y; o[3] = q.z; }

int64_t oracle_add_rod(void* h, uint32_t elements, double length, const double* curvature3, double E, double nu,
                       double density, double diameter, const double* pos3, const double* rot4, const double* vel3,
                       int base_fixed, int tip_fixed, int with_ground) {
    int64_t first = -1;
    int rc = guard([&] {
        first = static_cast<Simulation*>(h)->add_rod(elements, length, v_in(curvature3), E, nu, density, diameter,
                                                      v_in(pos3), q_in(rot4), v_in(vel3), base_fixed != 0,
                                                      tip_fixed != 0, with_ground != 0);
    });
    return rc == 0 ? first : -1;
}

int64_t oracle_add_sphere(void* h, double radius, double density, const double* pos3, const double* rot4, int fixed,
                          int with_ground) {
    return static_cast<Simulation*>(h)->add_sphere(radius, density, v_in(pos3), q_in(rot4), fixed != 0,
                                                    with_ground != 0);
}

int64_t oracle_add_oriented(void* h, const double* pos3, const double* rot4, double mass, const double* inertia3) {
    return static_cast<Simulation*>(h)->add_oriented(v_in(pos3), q_in(rot4), mass, v_in(inertia3));
}

// Cosserat element between two existing oriented particles.
void oracle_add_cosserat(void* h, uint32_t a, uint32_t b, const double* stiffness6, double L, const double* kappa3) {
    CosseratRod r{};
    r.idx[0] = a; r.idx[1] = b;
    for (int k = 0; k < 6; k++) r.stiffness[k] = stiffness6[k];
    r.L = L; r.kappa0 = v_in(kappa3);
    static_cast<Simulation*>(h)->cos.push_back(r);
}

int64_t oracle_add_joint(void* h, uint32_t p, const double* pos_offset3, const double* rot_offset4,
                         const double* ref_pos3, const double* ref_rot4) {
    return static_cast<Simulation*>(h)->add_joint(p, v_in(pos_offset3), q_in(rot_offset4), v_in(ref_pos3),
                                                   q_in(ref_rot4));
}

// Two-sided fixed joint (new): offsets3 = body 0's then body 1's position
// offset, rot_offsets4 = body 0's then body 1's rotation offset.
int64_t oracle_add_joint2(void* h, uint32_t p0, uint32_t p1, const double* offsets6, const double* rot_offsets8) {
    return static_cast<Simulation*>(h)->add_joint2(p0, p1, v_in(offsets6), q_in(rot_offsets8), v_in(offsets6 + 3),
                                                    q_in(rot_offsets8 + 4));
}

// Add mass (and, for an oriented particle, principal inertia) to particle p:
// a rigid body welded to it (a sphere at a rod tip).
void oracle_add_mass(void* h, uint32_t p, double dm, const double* dI3) {
    auto* s = static_cast<Simulation*>(h);
    s->mass[p] += dm;
    if (dI3 && s->oriented(p)) s->rs.inertia[s->rs.rot_of[p]] += v_in(dI3);
}

// RigidBodyGroundCollisionEnergyPool::addEnergy with a lowest-point shape
// (shape 0 sphere, 1 rod) of the given radius.
void oracle_add_rground(void* h, uint32_t p, uint32_t shape, double radius, double k_start, double mu_s, double mu_k) {
    auto* s = static_cast<Simulation*>(h);
    RigidGround c{};
    c.p = p;
    c.shape = shape;
    c.radius = radius;
    c.cp_rb_local = V3(0, -radius, 0);
    c.cp_ground = V3();
    c.n = V3(0, 1, 0);
    c.t = V3(1, 0, 0);
    c.b = V3(0, 0, 1);
    c.k_start = c.k_n = c.k_t = c.k_b = k_start;
    c.mu_s = mu_s;
    c.mu_k = mu_k;
    s->rground.push_back(c);
}

// Per-particle external acceleration added to gravity (3n doubles; NULL clears it).
void oracle_set_external_accel(void* h, const double* a) {
    auto* s = static_cast<Simulation*>(h);
    s->f_ext.clear();
    if (!a) return;
    s->f_ext.resize(s->x.size());
    for (size_t p = 0; p < s->x.size(); p++) s->f_ext[p] = v_in(a + 3 * p);
}

// RigidBodyGroundCollisionEnergyPool::addEnergy for a shape without the
// lowest-point rule (box, capsule): a fixed contact point in the body frame
// (energy/collision/RigidBodyGroundCollisionEnergySolver.hpp:111-120,161-183).
void oracle_add_rground_point(void* h, uint32_t p, const double* cp_local3, double k_start, double mu_s,
                              double mu_k) {
    auto* s = static_cast<Simulation*>(h);
    RigidGround c{};
    c.p = p;
    c.shape = SHAPE_OTHER;
    c.radius = 0;
    c.cp_rb_local = v_in(cp_local3);
    c.cp_ground = V3();
    c.n = V3(0, 1, 0);
    c.t = V3(1, 0, 0);
    c.b = V3(0, 0, 1);
    c.k_start = c.k_n = c.k_t = c.k_b = k_start;
    c.mu_s = mu_s;
    c.mu_k = mu_k;
    s->rground.push_back(c);
}

uint32_t oracle_num_cosserat(void* h) { return (uint32_t)static_cast<Simulation*>(h)->cos.size(); }
uint32_t oracle_num_rground(void* h) { return (uint32_t)static_cast<Simulation*>(h)->rground.size(); }
uint32_t oracle_num_joints(void* h) { return (uint32_t)static_cast<Simulation*>(h)->joints.size(); }
uint32_t oracle_num_joints2(void* h) { return (uint32_t)static_cast<Simulation*>(h)->joints2.size(); }

// Per particle: orientation (4, identity for positional particles), angular
// velocity (3), inertia (3), oriented flag.
void oracle_get_rotations(void* h, double* q4, double* w3, double* inertia3, uint8_t* oriented) {
    auto* s = static_cast<Simulation*>(h);
    for (size_t p = 0; p < s->x.size(); p++) {
        const bool o = s->oriented((uint32_t)p);
        if (oriented) oriented[p] = o ? 1 : 0;
        const int32_t r = o ? s->rs.rot_of[p] : -1;
        if (q4) q_out(o ? s->rs.q[r] : Quat(), q4 + 4 * p);
        for (int k = 0; k < 3; k++) {
            if (w3) w3[3 * p + k] = o ? s->rs.w[r][k] : 0.0;
            if (inertia3) inertia3[3 * p + k] = o ? s->rs.inertia[r][k] : 0.0;
        }
    }
}

// Sets orientation and angular velocity (and their previous-step copies) of
// every oriented particle; rows of positional particles are ignored.
void oracle_set_rotation_state(void* h, const double* q4, const double* w3) {
    auto* s = static_cast<Simulation*>(h);
    for (size_t p = 0; p < s->x.size(); p++) {
        if (!s->oriented((uint32_t)p)) continue;
        const int32_t r = s->rs.rot_of[p];
        if (q4) {
            const Quat q = q_in(q4 + 4 * p);
            s->rs.q[r] = s->rs.q_prev[r] = s->rs.q_last[r] = s->rs.q_lastlast[r] = s->rs.q_buf[r] = q;
        }
        if (w3) s->rs.w[r] = s->rs.w_prev[r] = v_in(w3 + 3 * p);
    }
}

void oracle_get_mass(void* h, double* m, uint8_t* fixed) {
    auto* s = static_cast<Simulation*>(h);
    for (size_t p = 0; p < s->x.size(); p++) {
        if (m) m[p] = s->mass[p];
        if (fixed) fixed[p] = s->fixed[p];
    }
}

// Rigid-ground records: k_start,k_n,k_t,k_b,lambda_n,lambda_t,lambda_b,C_n_prev,
// cp_rb_local(3),cp_ground(3),use_static,unclamped (16 per record)
void oracle_get_rground(void* h, double* rec16) {
    auto* s = static_cast<Simulation*>(h);
    for (size_t e = 0; e < s->rground.size(); e++) {
        const auto& c = s->rground[e];
        double* r = rec16 + 16 * e;
        r[0] = c.k_start; r[1] = c.k_n; r[2] = c.k_t; r[3] = c.k_b;
        r[4] = c.lambda_n; r[5] = c.lambda_t; r[6] = c.lambda_b; r[7] = c.C_n_prev;
        for (int k = 0; k < 3; k++) { r[8 + k] = c.cp_rb_local[k]; r[11 + k] = c.cp_ground[k]; }
        r[14] = c.use_static; r[15] = c.unclamped_lambda_tb;
    }
}

// Joint records: k(6), lambda(6), C_prev(6) (18 per record)
void oracle_get_joints(void* h, double* rec18) {
    auto* s = static_cast<Simulation*>(h);
    for (size_t e = 0; e < s->joints.size(); e++) {
        const auto& j = s->joints[e];
        for (int k = 0; k < 6; k++) {
            rec18[18 * e + k] = j.k[k];
            rec18[18 * e + 6 + k] = j.lambda[k];
            rec18[18 * e + 12 + k] = j.C_prev[k];
        }
    }
}

// Two-sided joint records: k(6), lambda(6), C_prev(6) (18 per record)
void oracle_get_joints2(void* h, double* rec18) {
    auto* s = static_cast<Simulation*>(h);
    for (size_t e = 0; e < s->joints2.size(); e++) {
        const auto& j = s->joints2[e];
        for (int k = 0; k < 6; k++) {
            rec18[18 * e + k] = j.k[k];
            rec18[18 * e + 6 + k] = j.lambda[k];
            rec18[18 * e + 12 + k] = j.C_prev[k];
        }
    }
}

// Single-energy evaluations for unit tests (G: 6, H: 36 row-major)
static void out6(const V6& G, const M6& H, double* G6, double* H36) {
    for (int i = 0; i < 6; i++) {
        G6[i] = G.v[i];
        for (int j = 0; j < 6; j++) H36[6 * i + j] = H.m[i][j];
    }
}

double oracle_cosserat_single(const double* x6, const double* q8, const double* stiffness6, double L,
                              const double* kappa3, uint32_t local, double* G6, double* H36) {
    std::vector<V3> x = {v_in(x6), v_in(x6 + 3)};
    RotState rs;
    rs.rot_of = {0, 1};
    rs.q = {q_in(q8), q_in(q8 + 4)};
    CosseratRod r{};
    r.idx[0] = 0; r.idx[1] = 1;
    for (int k = 0; k < 6; k++) r.stiffness[k] = stiffness6[k];
    r.L = L; r.kappa0 = v_in(kappa3);
    V6 G; M6 H;
    cosserat_accumulate(r, x, rs, local, G, H);
    out6(G, H, G6, H36);
    return cosserat_energy(r, x, rs);
}

// Joint at particle 0; writes the raw constraint vector C (6) as well.
void oracle_joint_single(const double* x3, const double* q4, const double* pos_offset3, const double* rot_offset4,
                         const double* ref_pos3, const double* ref_rot4, const double* k6, const double* lambda6,
                         const double* C_prev6, double* C6, double* G6, double* H36) {
    std::vector<V3> x = {v_in(x3)};
    RotState rs;
    rs.rot_of = {0};
    rs.q = {q_in(q4)};
    Joint j{};
    j.p = 0; j.pos_offset = v_in(pos_offset3); j.rot_offset = q_in(rot_offset4);
    j.ref_position = v_in(ref_pos3); j.ref_rotation = q_in(ref_rot4);
    for (int k = 0; k < 6; k++) { j.k[k] = k6[k]; j.lambda[k] = lambda6[k]; j.C_prev[k] = C_prev6[k]; }
    joint_eval(j, x, rs, C6);
    V6 G; M6 H;
    joint_accumulate(j, x, rs, G, H);
    out6(G, H, G6, H36);
}

// Two-sided joint between particles 0 and 1 (x6: both positions, q8: both
// orientations), accumulated for body `side`; writes C (6) and the gradient
// rows dC/d(x, theta) of that side (36, row-major) as well.
void oracle_joint2_single(const double* x6, const double* q8, const double* offsets6, const double* rot_offsets8,
                          const double* k6, const double* lambda6, const double* C_prev6, uint32_t side, double* C6,
                          double* grad36, double* G6, double* H36) {
    std::vector<V3> x = {v_in(x6), v_in(x6 + 3)};
    RotState rs;
    rs.rot_of = {0, 1};
    rs.q = {q_in(q8), q_in(q8 + 4)};
    Joint2 j{};
    j.p[0] = 0; j.p[1] = 1;
    j.pos_offset[0] = v_in(offsets6); j.pos_offset[1] = v_in(offsets6 + 3);
    j.rot_offset[0] = q_in(rot_offsets8); j.rot_offset[1] = q_in(rot_offsets8 + 4);
    for (int k = 0; k < 6; k++) { j.k[k] = k6[k]; j.lambda[k] = lambda6[k]; j.C_prev[k] = C_prev6[k]; }
    joint2_eval(j, x, rs, C6);
    Real grad[6][6];
    joint2_grad(j, side, x, rs, grad);
    for (int a = 0; a < 6; a++)
        for (int b = 0; b < 6; b++) grad36[6 * a + b] = grad[a][b];
    V6 G; M6 H;
    joint2_accumulate(j, side, x, rs, G, H);
    out6(G, H, G6, H36);
}

// Rigid ground at particle 0.  rec: k_n,k_t,k_b,lambda_n,lambda_t,lambda_b,C_n_prev,
// mu_s,mu_k,use_static,radius,shape; cp_local3 / cp_ground3.  Writes C (3).
void oracle_rground_single(const double* x3, const double* q4, const double* rec12, const double* cp_local3,
                           const double* cp_ground3, double* C3, double* G6, double* H36) {
    std::vector<V3> x = {v_in(x3)};
    RotState rs;
    rs.rot_of = {0};
    rs.q = {q_in(q4)};
    RigidGround c{};
    c.p = 0;
    c.k_n = rec12[0]; c.k_t = rec12[1]; c.k_b = rec12[2];
    c.lambda_n = rec12[3]; c.lambda_t = rec12[4]; c.lambda_b = rec12[5]; c.C_n_prev = rec12[6];
    c.mu_s = rec12[7]; c.mu_k = rec12[8]; c.use_static = rec12[9] != 0; c.radius = rec12[10];
    c.shape = (uint32_t)rec12[11];
    c.cp_rb_local = v_in(cp_local3); c.cp_ground = v_in(cp_ground3);
    rground_eval(c, x, rs, C3[0], C3[1], C3[2]);
    V6 G; M6 H;
    rground_accumulate(c, x, rs, G, H);
    out6(G, H, G6, H36);
}

// Geometry and narrow-phase hooks for the tests
void oracle_closest_segment_triangle(const double* p0, const double* p1, const double* a, const double* b,
                                     const double* c, double* s, double* barys3, double* dist) {
    V3 bc;
    closest_segment_triangle(v_in(p0), v_in(p1), v_in(a), v_in(b), v_in(c), *s, bc, *dist);
    for (int k = 0; k < 3; k++) barys3[k] = bc[k];
}
void oracle_closest_point_triangle(const double* p, const double* a, const double* b, const double* c, double* out3) {
    const V3 r = closest_point_triangle(v_in(p), v_in(a), v_in(b), v_in(c));
    for (int k = 0; k < 3; k++) out3[k] = r[k];
}
double oracle_golden_section_quadratic(double a, double b, double c) { return golden_section_quadratic(a, b, c); }
void oracle_unit_orthogonal(const double* n3, double* t3, double* b3) {
    V3 t, b;
    complete_basis(v_in(n3), t, b);
    for (int k = 0; k < 3; k++) {
        t3[k] = t[k];
        b3[k] = b[k];
    }
}
uint64_t oracle_contact_key(uint32_t type, uint32_t a, uint32_t b, uint32_t code) { return contact_key(type, a, b, code); }
int oracle_vertex_face(const double* p3, const double* tri9, double margin, double* normal3, double* barys3) {
    V3 n, bc;
    const bool hit = vertex_face(v_in(p3), v_in(tri9), v_in(tri9 + 3), v_in(tri9 + 6), margin, n, bc);
    for (int k = 0; k < 3; k++) {
        normal3[k] = n[k];
        barys3[k] = bc[k];
    }
    return hit ? 1 : 0;
}

int oracle_tri_sphere_ccd(const double* tri9, const double* tv9, const double* pos3, const double* q4, const double* v3,
                          const double* w3, double radius, double dt, double margin, double* normal3, double* barys3,
                          double* cpl3) {
    const V3 T[3] = {v_in(tri9), v_in(tri9 + 3), v_in(tri9 + 6)};
    const V3 TV[3] = {v_in(tv9), v_in(tv9 + 3), v_in(tv9 + 6)};
    V3 n, bc, cp;
    const bool hit = tri_sphere_ccd(T, TV, v_in(pos3), q_in(q4), v_in(v3), v_in(w3), radius, dt, margin, n, bc, cp);
    for (int k = 0; k < 3; k++) {
        normal3[k] = n[k];
        barys3[k] = bc[k];
        cpl3[k] = cp[k];
    }
    return hit ? 1 : 0;
}
// One dynamic contact over particles 0..4 (0-2 triangle, 3 body or 3-4 rod
// nodes; rigid-rigid: bodies 3 and 4, cpl2 in the barys slot; q20:
// quaternions (w, x, y, z) of particles 0..4, only 3-4 used).
// rec: n, t, b, barys, s, cpl (19), then k_n, k_t, k_b, lambda_n, lambda_t,
// lambda_b, C_n_prev, mu_s, mu_k, use_static (10).  Writes C (3), the raw
// gradients g (3 x 6: n, t, b), the rotational Hessian blocks h (3 x 9) and
// the accumulated G (dof) / H (dof x dof).
void oracle_dyn_single(uint32_t type, const double* x15, const double* q20, const double* rec26, uint32_t local,
                       uint32_t dof, double* C3, double* g18, double* h27, double* G, double* H) {
    std::vector<V3> x;
    RotState rs;
    for (int i = 0; i < 5; i++) {
        x.push_back(v_in(x15 + 3 * i));
        rs.rot_of.push_back(i >= 3 ? i - 3 : -1);
    }
    rs.q = {q_in(q20 + 12), q_in(q20 + 16)};
    DynContact c;
    for (int i = 0; i < 5; i++) c.idx[i] = i;
    if (type == DYN_RIGID_RIGID) {   // the two bodies are the oriented particles 3 and 4
        c.idx[0] = 3;
        c.idx[1] = 4;
    }
    c.n = v_in(rec26);
    c.t = v_in(rec26 + 3);
    c.b = v_in(rec26 + 6);
    c.barys = v_in(rec26 + 9);
    if (type == DYN_RIGID_RIGID) c.cpl2 = c.barys;   // the rigid-rigid record carries cpl2 there
    c.s = rec26[12];
    c.cpl = v_in(rec26 + 13);
    c.k_n = rec26[16]; c.k_t = rec26[17]; c.k_b = rec26[18];
    c.lambda_n = rec26[19]; c.lambda_t = rec26[20]; c.lambda_b = rec26[21]; c.C_n_prev = rec26[22];
    c.mu_s = rec26[23]; c.mu_k = rec26[24]; c.use_static = rec26[25] != 0;
    M3 hn, ht, hb;
    dyn_constraint((uint8_t)type, c, x, rs, local, (int)dof, C3[0], C3[1], C3[2], g18, g18 + 6, g18 + 12, hn, ht, hb);
    m3_out(hn, h27);
    m3_out(ht, h27 + 9);
    m3_out(hb, h27 + 18);
    if (dof == 3) {
        V3 g3;
        M3 H3;
        dyn_accumulate3((uint8_t)type, c, x, rs, local, g3, H3);
        for (int i = 0; i < 3; i++) {
            G[i] = g3[i];
            for (int j = 0; j < 3; j++) H[3 * i + j] = H3.m[i][j];
        }
    } else {
        V6 g6;
        M6 H6;
        dyn_accumulate6((uint8_t)type, c, x, rs, local, g6, H6);
        out6(g6, H6, G, H);
    }
}

// TombstonePool slot stack: op >= 0 frees that slot, op < 0 allocates (the
// slot goes to out[i]; out[i] = 0xFFFFFFFF for frees).  Returns 0 or -1.
int oracle_slot_sequence(uint32_t n_ops, const int32_t* ops, uint32_t capacity, uint32_t* out) {
    return guard([&] {
        SlotPool sp;
        sp.capacity = capacity;
        for (uint32_t i = 0; i < n_ops; i++) {
            if (ops[i] < 0) out[i] = sp.alloc();
            else {
                sp.free((uint32_t)ops[i]);
                out[i] = 0xFFFFFFFFu;
            }
        }
    });
}
// updateCollisionFrame on (n, t, b, lambda_t, lambda_b) in frame12 =
// n, t, b, lambda_t, lambda_b, unused; writes the updated values back.
void oracle_update_collision_frame(double* frame12, const double* new_n3) {
    DynContact c;
    c.n = v_in(frame12);
    c.t = v_in(frame12 + 3);
    c.b = v_in(frame12 + 6);
    c.lambda_t = frame12[9];
    c.lambda_b = frame12[10];
    update_collision_frame(c, v_in(new_n3));
    for (int k = 0; k < 3; k++) {
        frame12[k] = c.n[k];
        frame12[3 + k] = c.t[k];
        frame12[6 + k] = c.b[k];
    }
    frame12[9] = c.lambda_t;
    frame12[10] = c.lambda_b;
}

// S3 / SO3 maps
void oracle_exp_s3(const double* v3, double* q4) { q_out(exp_s3(v_in(v3)), q4); }
void oracle_log_s3(const double* q4, double* v3) {
    const V3 v = log_s3(q_in(q4));
    v3[0] = v.x; v3[1] = v.y; v3[2] = v.z;
}
void oracle_qmat(const double* q4, double* M9) { m3_out(qmat(q_in(q4)), M9); }
void oracle_exp_so3(const double* v3, double* M9) { m3_out(exp_so3(v_in(v3)), M9); }
void oracle_jr(const double* v3, double* M9) { m3_out(jr(v_in(v3)), M9); }
void oracle_jr_inv(const double* v3, double* M9) { m3_out(jr_inv(v_in(v3)), M9); }

}  // extern "C"
