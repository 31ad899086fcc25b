// contacts_host.cpp -- host side of the dynamic contacts: contact lifecycle
// (add / update / remove against the previous round, with the reference's
// slot-stack pools) and the dynamic adjacency.  The reference runs these on
// its first worker thread between the narrow phase and the solve; here they
// run on the host between the device narrow phase and the device
// recolouring (recolour_kernels.hip) of the same step.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "contacts.hpp"

namespace avbd {

namespace {

// Eigen's unitOrthogonal for 3-vectors (double precision, as the reference).
void unit_orthogonal(const double n[3], double t[3]) {
    const double prec = 1e-12;
    const bool xs = std::fabs(n[0]) <= std::fabs(n[2]) * prec, ys = std::fabs(n[1]) <= std::fabs(n[2]) * prec;
    if (!xs || !ys) {
        const double inv = 1.0 / std::sqrt(n[0] * n[0] + n[1] * n[1]);
        t[0] = -n[1] * inv;
        t[1] = n[0] * inv;
        t[2] = 0.0;
    } else {
        const double inv = 1.0 / std::sqrt(n[1] * n[1] + n[2] * n[2]);
        t[0] = 0.0;
        t[1] = -n[2] * inv;
        t[2] = n[1] * inv;
    }
}
}  // namespace

// TombstonePool::allocSlot / freeSlot (common/TombstonePool.hpp:64-110)
uint32_t ContactSystem::Slots::alloc(uint32_t capacity) {
    if (top >= capacity) throw std::runtime_error("dynamic contact pool has exceeded its capacity");
    const uint32_t i = top++;
    const uint32_t slot = i < stack.size() ? stack[i] : i;
    if (active.size() <= slot) {
        active.resize(slot + 1, 0);
        gen.resize(slot + 1, 0);
        idx.resize(slot + 1);
    }
    active[slot] = 1;
    gen[slot]++;
    count++;
    return slot;
}

void ContactSystem::Slots::free(uint32_t slot) {
    if (slot >= active.size() || !active[slot]) throw std::runtime_error("freeSlot called on inactive slot");
    active[slot] = 0;
    count--;
    const uint32_t i = --top;
    while (stack.size() <= i) stack.push_back((uint32_t)stack.size());
    stack[i] = slot;
}

void ContactSystem::init(uint32_t n_particles, uint32_t capacity, double collision_margin) {
    n = n_particles;
    cap = capacity;
    margin = collision_margin;
    for (Slots& P : pool) P = Slots();
    prev.clear();
    seeds.clear();
    dyn_off.assign(n + 1, 0);
    dyn_ent.clear();
    nb_off.assign(n + 1, 0);
    nb.clear();
    ops.clear();
    active.clear();
}

void ContactSystem::process(std::vector<DetRec>& det, const std::vector<HostPrim>& prims,
                            const std::vector<double>& mass, double dt) {
    ops.clear();
    seeds.clear();
    // _handleDetectedCollisions (collision/CollisionDetector.cpp:272-327): key order
    std::sort(det.begin(), det.end(), [](const DetRec& a, const DetRec& b) { return a.key < b.key; });
    std::vector<Entry> cur(det.size());
    auto kind_of = [](uint32_t det_kind) {
        if (det_kind == DET_KIND_VERTEX_FACE) return DYN_VERTEX_TRI;
        return det_kind == DET_KIND_TRI_RIGID ? DYN_TRI_RIGID : (det_kind == DET_KIND_TRI_ROD ? DYN_TRI_ROD : DYN_RIGID_RIGID);
    };
    auto add = [&](size_t i) {
        const DetRec& r = det[i];
        const uint32_t kind = kind_of(r.kind);
        const HostPrim& pa = prims[r.prim_a];
        const HostPrim& pb = prims[r.prim_b];
        // triangle-rigid: triangle, body; triangle-rod: triangle, two nodes;
        // rigid-rigid: the two bodies; vertex-triangle: the face (prim_a),
        // then corner `pad` of the vertex's owner (prim_b)
        std::array<uint32_t, 5> id{};
        if (kind == DYN_RIGID_RIGID) id = {pa.idx[0], pb.idx[0], 0u, 0u, 0u};
        else if (kind == DYN_VERTEX_TRI) id = {pa.idx[0], pa.idx[1], pa.idx[2], pb.idx[r.pad], 0u};
        else id = {pa.idx[0], pa.idx[1], pa.idx[2], pb.idx[0], kind == DYN_TRI_ROD ? pb.idx[1] : 0u};
        const uint32_t np = dyn_particles(kind);
        double mmin = mass[id[0]];
        for (uint32_t k = 1; k < np; k++) mmin = std::min(mmin, mass[id[k]]);
        const uint32_t slot = pool[kind].alloc(cap);
        pool[kind].idx[slot] = id;
        DynOp o{};
        o.rec = kind * cap + slot;
        o.op = DYN_OP_ADD;
        o.kind = kind;
        for (int k = 0; k < 5; k++) o.idx[k] = id[k];
        o.k_start = (float)(mmin / (dt * dt));
        double t[3], b[3];
        unit_orthogonal(r.n, t);
        b[0] = r.n[1] * t[2] - r.n[2] * t[1];
        b[1] = r.n[2] * t[0] - r.n[0] * t[2];
        b[2] = r.n[0] * t[1] - r.n[1] * t[0];
        for (int k = 0; k < 3; k++) {
            o.n[k] = (float)r.n[k];
            o.t[k] = (float)t[k];
            o.b[k] = (float)b[k];
            o.bary[k] = (float)r.bary[k];
            o.cpl[k] = (float)r.cpl[k];
        }
        o.s = (float)r.s;
        ops.push_back(o);
        cur[i] = Entry{r.key, kind, slot, 1u, 1u};
        for (uint32_t k = 0; k < np; k++) seeds.push_back(id[k]);   // ColorList::markParticleDirty
    };
    auto update = [&](size_t i, uint32_t slot) {
        const DetRec& r = det[i];
        const uint32_t kind = kind_of(r.kind);
        DynOp o{};
        o.rec = kind * cap + slot;
        o.op = DYN_OP_UPDATE;
        o.kind = kind;
        for (int k = 0; k < 3; k++) {
            o.n[k] = (float)r.n[k];
            o.bary[k] = (float)r.bary[k];
            o.cpl[k] = (float)r.cpl[k];
        }
        o.s = (float)r.s;
        ops.push_back(o);
        cur[i] = Entry{r.key, kind, slot, 1u, 1u};
    };
    auto remove = [&](const Entry& e) { pool[e.kind].free(e.slot); };
    size_t ci = 0, pi = 0;
    while (ci < det.size() && pi < prev.size()) {
        if (det[ci].key < prev[pi].key) {
            add(ci++);
        } else if (det[ci].key > prev[pi].key) {
            remove(prev[pi++]);
        } else {
            // primitives are never removed, so both generations always match
            update(ci++, prev[pi++].slot);
        }
    }
    for (; ci < det.size(); ci++) add(ci);
    for (; pi < prev.size(); pi++) remove(prev[pi]);
    prev.swap(cur);

    // active records and the dynamic adjacency (DynamicParticleAdjacency::buildAdjacency,
    // common/DynamicParticleAdjacency.hpp:33-118): pool 0 then pool 1, ascending slot
    active.clear();
    std::vector<uint32_t> val(n + 1, 0), deg(n + 1, 0);
    for (uint32_t kind = 0; kind < DYN_KINDS; kind++) {
        const uint32_t np = dyn_particles(kind);
        const Slots& P = pool[kind];
        for (uint32_t sl = 0; sl < P.active.size(); sl++) {
            if (!P.active[sl]) continue;
            active.push_back(kind * cap + sl);
            for (uint32_t k = 0; k < np; k++) {
                val[P.idx[sl][k] + 1]++;
                deg[P.idx[sl][k] + 1] += np - 1;
            }
        }
    }
    for (uint32_t v = 0; v < n; v++) {
        val[v + 1] += val[v];
        deg[v + 1] += deg[v];
    }
    dyn_off = val;
    nb_off = deg;
    dyn_ent.resize(val[n]);
    nb.resize(deg[n]);
    for (uint32_t kind = 0; kind < DYN_KINDS; kind++) {
        const uint32_t np = dyn_particles(kind);
        const Slots& P = pool[kind];
        for (uint32_t sl = 0; sl < P.active.size(); sl++) {
            if (!P.active[sl]) continue;
            const uint32_t rec = kind * cap + sl;
            for (uint32_t k = 0; k < np; k++) {
                const uint32_t p = P.idx[sl][k];
                dyn_ent[val[p]++] = (rec << 3) | k;
                for (uint32_t k2 = 0; k2 < np; k2++)
                    if (k2 != k) nb[deg[p]++] = P.idx[sl][k2];
            }
        }
    }
}

namespace {
template <class T>
void put(std::vector<uint8_t>& o, const T& v) {
    const auto* b = reinterpret_cast<const uint8_t*>(&v);
    o.insert(o.end(), b, b + sizeof(T));
}
template <class T>
void put_vec(std::vector<uint8_t>& o, const std::vector<T>& v) {
    put<uint64_t>(o, v.size());
    const auto* b = reinterpret_cast<const uint8_t*>(v.data());
    o.insert(o.end(), b, b + v.size() * sizeof(T));
}
template <class T>
T get(const uint8_t*& p, const uint8_t* end) {
    if ((size_t)(end - p) < sizeof(T)) throw std::runtime_error("checkpoint is truncated");
    T v;
    std::memcpy(&v, p, sizeof(T));
    p += sizeof(T);
    return v;
}
template <class T>
void get_vec(const uint8_t*& p, const uint8_t* end, std::vector<T>& v) {
    const uint64_t n = get<uint64_t>(p, end);
    if ((uint64_t)(end - p) / sizeof(T) < n) throw std::runtime_error("checkpoint is truncated");
    v.resize(n);
    std::memcpy(v.data(), p, n * sizeof(T));
    p += n * sizeof(T);
}
}  // namespace

void ContactSystem::save(std::vector<uint8_t>& o) const {
    put(o, n);
    put(o, cap);
    put(o, margin);
    for (const Slots& s : pool) {
        put(o, s.top);
        put(o, s.count);
        put_vec(o, s.stack);
        put_vec(o, s.active);
        put_vec(o, s.gen);
        put_vec(o, s.idx);
    }
    put_vec(o, prev);
    for (const auto* v : {&active, &dyn_off, &dyn_ent, &nb_off, &nb}) put_vec(o, *v);
}

void ContactSystem::load(const uint8_t*& p, const uint8_t* end) {
    if (get<uint32_t>(p, end) != n || get<uint32_t>(p, end) != cap)
        throw std::runtime_error("checkpoint does not match this context's contact pools");
    margin = get<double>(p, end);
    for (Slots& s : pool) {
        s.top = get<uint32_t>(p, end);
        s.count = get<uint32_t>(p, end);
        get_vec(p, end, s.stack);
        get_vec(p, end, s.active);
        get_vec(p, end, s.gen);
        get_vec(p, end, s.idx);
    }
    get_vec(p, end, prev);
    for (auto* v : {&active, &dyn_off, &dyn_ent, &nb_off, &nb}) get_vec(p, end, *v);
    ops.clear();
    seeds.clear();
}

}  // namespace avbd
