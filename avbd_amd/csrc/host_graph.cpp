// host_graph.cpp -- setup-time graph work of the avbd context: vertex
// adjacency, DSATUR colouring per scene, and the per-colour work-item layout
// that the vertex-solve kernel consumes.
#include "internal.hpp"

#include <algorithm>
#include <set>
#include <stdexcept>
#include <thread>
#include <tuple>

namespace avbd {

// ColorList::greedyColor2 (common/ColorList.hpp:149-241).  The reference scans
// all uncoloured vertices each round and keeps the first vertex with the
// lexicographically largest (saturation, degree); that is the minimum of the
// key (-saturation, -degree, index), kept here in an ordered set so a round
// costs O(log n) instead of O(n).  Saturation = number of distinct colours
// among coloured neighbours, maintained with a 256-bit set per vertex (the
// reference's ColorBitset<4> limit, common/ColorList.hpp:11-38).
uint32_t dsatur_csr(uint32_t n, const uint32_t* off, const uint32_t* nb, uint32_t* color) {
    constexpr uint32_t UNC = 0xFFFFFFFFu;
    std::vector<std::array<uint64_t, 4>> seen(n, std::array<uint64_t, 4>{0, 0, 0, 0});
    std::vector<uint32_t> sat(n, 0);
    using Key = std::tuple<int64_t, int64_t, uint32_t>;
    std::set<Key> heap;
    for (uint32_t v = 0; v < n; v++) {
        color[v] = UNC;
        heap.emplace(0, -(int64_t)(off[v + 1] - off[v]), v);
    }
    uint32_t num_colors = 0;
    while (!heap.empty()) {
        const uint32_t best = std::get<2>(*heap.begin());
        heap.erase(heap.begin());
        std::array<uint64_t, 4> used = {0, 0, 0, 0};
        for (uint32_t e = off[best]; e < off[best + 1]; e++) {
            const uint32_t c = color[nb[e]];
            if (c != UNC) used[c >> 6] |= 1ull << (c & 63);
        }
        uint32_t c = 0;
        while (c < 256 && (used[c >> 6] >> (c & 63) & 1ull)) c++;
        if (c >= 256) throw std::runtime_error("Ran out of colors!");
        color[best] = c;
        num_colors = std::max(num_colors, c + 1);
        for (uint32_t e = off[best]; e < off[best + 1]; e++) {
            const uint32_t w = nb[e];
            if (color[w] != UNC) continue;
            if (seen[w][c >> 6] >> (c & 63) & 1ull) continue;
            heap.erase(Key(-(int64_t)sat[w], -(int64_t)(off[w + 1] - off[w]), w));
            seen[w][c >> 6] |= 1ull << (c & 63);
            sat[w]++;
            heap.emplace(-(int64_t)sat[w], -(int64_t)(off[w + 1] - off[w]), w);
        }
    }
    return num_colors;
}

namespace {

struct SceneGraph {
    std::vector<uint32_t> off, nb;
};

// Vertex-vertex adjacency of one scene (StaticParticleAdjacency::buildAdjacency,
// common/StaticParticleAdjacency.hpp:42-133): deduplicated neighbour sets of the
// static multi-particle energies (Neo-Hookean tets and Cosserat elements, and
// the two-sided fixed joints, new).
void scene_adjacency(uint32_t lo, uint32_t hi, const std::vector<HostNH>& nh, const std::vector<uint32_t>& tets,
                     const std::vector<HostCos>& cos, const std::vector<uint32_t>& rods,
                     const std::vector<HostJoint2>& joints2, const std::vector<uint32_t>& links, SceneGraph& g) {
    const uint32_t n = hi - lo;
    std::vector<std::vector<uint32_t>> sets(n);
    for (uint32_t t : tets) {
        const HostNH& r = nh[t];
        for (int k = 0; k < 4; k++)
            for (int k2 = 0; k2 < 4; k2++)
                if (k != k2) sets[r.idx[k] - lo].push_back(r.idx[k2] - lo);
    }
    for (uint32_t e : rods) {
        sets[cos[e].idx[0] - lo].push_back(cos[e].idx[1] - lo);
        sets[cos[e].idx[1] - lo].push_back(cos[e].idx[0] - lo);
    }
    for (uint32_t e : links) {
        sets[joints2[e].p[0] - lo].push_back(joints2[e].p[1] - lo);
        sets[joints2[e].p[1] - lo].push_back(joints2[e].p[0] - lo);
    }
    g.off.assign(n + 1, 0);
    for (uint32_t v = 0; v < n; v++) {
        auto& s = sets[v];
        std::sort(s.begin(), s.end());
        s.erase(std::unique(s.begin(), s.end()), s.end());
        g.off[v + 1] = g.off[v] + (uint32_t)s.size();
    }
    g.nb.resize(g.off[n]);
    for (uint32_t v = 0; v < n; v++) std::copy(sets[v].begin(), sets[v].end(), g.nb.begin() + g.off[v]);
}

// CSR of (particle -> entries) from (particle, entry) pairs in pool order.
void csr_by_particle(uint32_t n, const std::vector<std::pair<uint32_t, uint32_t>>& pairs, std::vector<uint32_t>& off,
                     std::vector<uint32_t>& ent) {
    off.assign(n + 1, 0);
    for (const auto& pr : pairs) off[pr.first + 1]++;
    for (uint32_t v = 0; v < n; v++) off[v + 1] += off[v];
    ent.resize(pairs.size());
    std::vector<uint32_t> cur(off.begin(), off.end() - 1);
    for (const auto& pr : pairs) ent[cur[pr.first]++] = pr.second;
}

}  // namespace

void build_color_plan(const HostModel& m, ColorPlan& plan, const SceneColourer& colour) {
    const uint32_t n = m.n;
    const auto& nh = *m.nh;
    const auto& ground = *m.ground;
    const auto& cos = *m.cos;
    const auto& rground = *m.rground;
    const auto& joints = *m.joints;
    static const std::vector<HostJoint2> no_joints2;
    const auto& joints2 = m.joints2 ? *m.joints2 : no_joints2;
    const auto& fixed = *m.fixed;
    const auto& rot_of = *m.rot_of;
    const auto& scene_offsets = *m.scene_offsets;
    const uint32_t n_scenes = (uint32_t)scene_offsets.size() - 1;
    std::vector<uint32_t> scene_of(n, 0);
    for (uint32_t s = 0; s < n_scenes; s++)
        for (uint32_t v = scene_offsets[s]; v < scene_offsets[s + 1]; v++) scene_of[v] = s;
    auto oriented = [&](uint32_t p) { return rot_of[p] >= 0; };

    // energies per scene; every multi-particle energy must lie inside one scene.
    // Neo-Hookean supports positional particles only and the 6-DOF energies
    // oriented ones only (SupportsPositional / SupportsOriented of the solvers).
    std::vector<std::vector<uint32_t>> scene_tets(n_scenes), scene_rods(n_scenes);
    for (uint32_t t = 0; t < nh.size(); t++) {
        for (int k = 0; k < 4; k++) {
            if (nh[t].idx[k] >= n) throw std::runtime_error("tet references a particle out of range");
            if (oriented(nh[t].idx[k])) throw std::runtime_error("Neo-Hookean energy on an oriented particle");
        }
        const uint32_t s = scene_of[nh[t].idx[0]];
        for (int k = 0; k < 4; k++)
            if (scene_of[nh[t].idx[k]] != s) throw std::runtime_error("tet spans two scenes");
        scene_tets[s].push_back(t);
    }
    for (uint32_t e = 0; e < cos.size(); e++) {
        for (int k = 0; k < 2; k++) {
            if (cos[e].idx[k] >= n) throw std::runtime_error("Cosserat element references a particle out of range");
            if (!oriented(cos[e].idx[k])) throw std::runtime_error("Cosserat element on a positional particle");
        }
        if (cos[e].idx[0] == cos[e].idx[1]) throw std::runtime_error("Cosserat element with one particle twice");
        if (scene_of[cos[e].idx[0]] != scene_of[cos[e].idx[1]]) throw std::runtime_error("rod spans two scenes");
        scene_rods[scene_of[cos[e].idx[0]]].push_back(e);
    }
    for (const auto& r : rground)
        if (r.p >= n || !oriented(r.p)) throw std::runtime_error("rigid-ground contact needs an oriented particle");
    for (const auto& j : joints)
        if (j.p >= n || !oriented(j.p)) throw std::runtime_error("joint needs an oriented particle");
    std::vector<std::vector<uint32_t>> scene_links(n_scenes);
    for (uint32_t e = 0; e < joints2.size(); e++) {
        const auto& j = joints2[e];
        for (int k = 0; k < 2; k++)
            if (j.p[k] >= n || !oriented(j.p[k])) throw std::runtime_error("joint needs an oriented particle");
        if (j.p[0] == j.p[1]) throw std::runtime_error("two-sided joint with one particle twice");
        if (scene_of[j.p[0]] != scene_of[j.p[1]]) throw std::runtime_error("joint spans two scenes");
        scene_links[scene_of[j.p[0]]].push_back(e);
    }

    // vertex adjacency of every scene (threads over scenes), kept as one
    // global CSR (also used by the incremental recolouring)
    std::vector<SceneGraph> graphs(n_scenes);
    const uint32_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    {
        std::vector<std::thread> pool;
        for (uint32_t w = 0; w < hw; w++)
            pool.emplace_back([&, w] {
                for (uint32_t s = w; s < n_scenes; s += hw)
                    scene_adjacency(scene_offsets[s], scene_offsets[s + 1], nh, scene_tets[s], cos, scene_rods[s],
                                    joints2, scene_links[s], graphs[s]);
            });
        for (auto& t : pool) t.join();
    }
    plan.static_off.assign(n + 1, 0);
    plan.static_nb.clear();
    for (uint32_t s = 0; s < n_scenes; s++) {
        const uint32_t lo = scene_offsets[s], hi = scene_offsets[s + 1];
        const SceneGraph& g = graphs[s];
        for (uint32_t v = lo; v < hi; v++) {
            for (uint32_t e = g.off[v - lo]; e < g.off[v - lo + 1]; e++) plan.static_nb.push_back(g.nb[e] + lo);
            plan.static_off[v + 1] = (uint32_t)plan.static_nb.size();
        }
    }

    // colour every scene on its own: DSATUR on the device when a colourer is
    // given, else on the host (threads over scenes)
    plan.color.assign(n, 0);
    std::vector<uint32_t> scene_colors(n_scenes, 0);
    if (colour) {
        colour(scene_offsets, plan.static_off, plan.static_nb, plan.color, scene_colors);
    } else {
        std::vector<std::string> errors(n_scenes);
        std::vector<std::thread> pool;
        for (uint32_t w = 0; w < hw; w++) {
            pool.emplace_back([&, w] {
                for (uint32_t s = w; s < n_scenes; s += hw) {
                    try {
                        const SceneGraph& g = graphs[s];
                        const uint32_t lo = scene_offsets[s], hi = scene_offsets[s + 1];
                        scene_colors[s] = dsatur_csr(hi - lo, g.off.data(), g.nb.data(), plan.color.data() + lo);
                    } catch (const std::exception& e) {
                        errors[s] = e.what();
                    }
                }
            });
        }
        for (auto& t : pool) t.join();
        for (uint32_t s = 0; s < n_scenes; s++)
            if (!errors[s].empty()) throw std::runtime_error(errors[s]);
    }
    plan.num_colors = 0;
    for (uint32_t s = 0; s < n_scenes; s++) plan.num_colors = std::max(plan.num_colors, scene_colors[s]);
    graphs.clear();

    // per-vertex incidences in pool order (ascending energy index, then local)
    std::vector<std::pair<uint32_t, uint32_t>> pr;
    for (uint32_t t = 0; t < nh.size(); t++)
        for (uint32_t k = 0; k < 4; k++) pr.emplace_back(nh[t].idx[k], (t << 2) | k);
    std::vector<uint32_t> inc_off, inc_all;
    csr_by_particle(n, pr, inc_off, inc_all);
    pr.clear();
    for (uint32_t e = 0; e < cos.size(); e++)
        for (uint32_t k = 0; k < 2; k++) pr.emplace_back(cos[e].idx[k], (e << 1) | k);
    std::vector<uint32_t> cos_off, cos_all;
    csr_by_particle(n, pr, cos_off, cos_all);
    pr.clear();
    for (uint32_t e = 0; e < rground.size(); e++) pr.emplace_back(rground[e].p, e);
    std::vector<uint32_t> rg_off, rg_all;
    csr_by_particle(n, pr, rg_off, rg_all);
    pr.clear();
    for (uint32_t e = 0; e < joints.size(); e++) pr.emplace_back(joints[e].p, e);
    std::vector<uint32_t> jt_off, jt_all;
    csr_by_particle(n, pr, jt_off, jt_all);
    // the solve's joint terms: the one-sided joints, then both bodies of every
    // two-sided one (records follow the one-sided ones on the device)
    const uint32_t n_j1 = (uint32_t)joints.size();
    for (uint32_t e = 0; e < joints2.size(); e++) {
        pr.emplace_back(joints2[e].p[0], n_j1 + e);
        pr.emplace_back(joints2[e].p[1], (n_j1 + e) | OINC_JOINT_SIDE1);
    }
    std::vector<uint32_t> js_off, js_all;
    csr_by_particle(n, pr, js_off, js_all);

    std::vector<int32_t> ground_of(n, -1);
    for (uint32_t gi = 0; gi < ground.size(); gi++) {
        const uint32_t p = ground[gi].p;
        if (p >= n) throw std::runtime_error("ground energy references a particle out of range");
        if (ground_of[p] >= 0) throw std::runtime_error("more than one ground energy on one particle");
        ground_of[p] = (int32_t)gi;
    }

    // per rotation slot: its rigid-ground then joint records (start-of-step
    // and dual updates run per particle, fixed ones included)
    uint32_t n_rot = 0;
    for (uint32_t v = 0; v < n; v++)
        if (oriented(v)) n_rot = std::max(n_rot, (uint32_t)rot_of[v] + 1);
    std::vector<int32_t> particle_of_rot(n_rot, -1);
    for (uint32_t v = 0; v < n; v++)
        if (oriented(v)) particle_of_rot[rot_of[v]] = (int32_t)v;
    plan.prec_off.assign(n_rot + 1, 0);
    plan.prec.clear();
    for (uint32_t r = 0; r < n_rot; r++) {
        plan.prec_off[r] = (uint32_t)plan.prec.size();
        const int32_t v = particle_of_rot[r];
        if (v < 0) continue;
        for (uint32_t k = rg_off[v]; k < rg_off[v + 1]; k++) plan.prec.push_back(rg_all[k]);
        for (uint32_t k = jt_off[v]; k < jt_off[v + 1]; k++) plan.prec.push_back(jt_all[k] | PREC_JOINT);
    }
    plan.prec_off[n_rot] = (uint32_t)plan.prec.size();

    // work items: colour-major, then scene order, then ascending vertex index
    // (ColorList::rebuildWorkList order, common/ColorList.hpp:345-363); fixed
    // particles are dropped (they return immediately, VBDSolver.hpp:288).
    plan.items.clear();
    plan.inc.clear();
    plan.oitems.clear();
    plan.oinc.clear();
    plan.item_kind.assign(n, 0);
    plan.item_of.assign(n, VItem{});
    plan.oitem_of.assign(n, OItem{});
    plan.item_offsets.assign(plan.num_colors + 1, 0);
    plan.oitem_offsets.assign(plan.num_colors + 1, 0);
    for (uint32_t c = 0; c < plan.num_colors; c++) {
        plan.item_offsets[c] = (uint32_t)plan.items.size();
        plan.oitem_offsets[c] = (uint32_t)plan.oitems.size();
        for (uint32_t v = 0; v < n; v++) {
            if (plan.color[v] != c || fixed[v]) continue;
            if (!oriented(v)) {
                VItem it;
                it.v = v;
                it.inc_begin = (uint32_t)plan.inc.size();
                it.inc_count = inc_off[v + 1] - inc_off[v];
                it.ground = ground_of[v];
                plan.inc.insert(plan.inc.end(), inc_all.begin() + inc_off[v], inc_all.begin() + inc_off[v + 1]);
                plan.items.push_back(it);
                plan.item_kind[v] = 1;
                plan.item_of[v] = it;
            } else {
                OItem it{};
                it.v = v;
                it.rot = (uint32_t)rot_of[v];
                it.inc_begin = (uint32_t)plan.oinc.size();
                it.n_cos = cos_off[v + 1] - cos_off[v];
                it.n_rg = rg_off[v + 1] - rg_off[v];
                it.n_joint = js_off[v + 1] - js_off[v];
                it.ground = ground_of[v];
                plan.oinc.insert(plan.oinc.end(), cos_all.begin() + cos_off[v], cos_all.begin() + cos_off[v + 1]);
                plan.oinc.insert(plan.oinc.end(), rg_all.begin() + rg_off[v], rg_all.begin() + rg_off[v + 1]);
                plan.oinc.insert(plan.oinc.end(), js_all.begin() + js_off[v], js_all.begin() + js_off[v + 1]);
                plan.oitems.push_back(it);
                plan.item_kind[v] = 2;
                plan.oitem_of[v] = it;
            }
        }
    }
    plan.item_offsets[plan.num_colors] = (uint32_t)plan.items.size();
    plan.oitem_offsets[plan.num_colors] = (uint32_t)plan.oitems.size();
    plan.conflict_offsets.assign(plan.num_colors + 1, 0);
    plan.conflict_list.clear();
}

}  // namespace avbd
