// graph.cpp -- FP64 oracle: static adjacency CSR and DSATUR colouring
// (test infrastructure only; see oracle.hpp).
#include "oracle.hpp"
#include "rigid.hpp"

#include <algorithm>
#include <stdexcept>

namespace oracle {

// StaticParticleAdjacency::buildAdjacency, common/StaticParticleAdjacency.hpp:42-133.
// Entries for (vertex, type) appear in pool order (ascending energy index, then
// local index k).  The vertex-vertex neighbour set is deduplicated; the
// reference stores it in unordered_set iteration order, which no consumer
// depends on (DSATUR only uses it as a set), so it is kept ascending here.
void build_static_adjacency(uint32_t n, const std::vector<NHTet>& nh, const std::vector<CosseratRod>& cos,
                            const std::vector<GroundContact>& ground, const std::vector<RigidGround>& rground,
                            const std::vector<Joint>& joints, const std::vector<Joint2>& joints2,
                            StaticAdjacency& adj) {
    std::vector<std::array<uint32_t, ST_COUNT>> val(n);
    for (auto& a : val) a.fill(0);
    std::vector<std::vector<uint32_t>> nb(n);
    for (const auto& t : nh)
        for (int k = 0; k < 4; k++) {
            val[t.idx[k]][ST_NEO_HOOKEAN]++;
            for (int k2 = 0; k2 < 4; k2++)
                if (k2 != k) nb[t.idx[k]].push_back(t.idx[k2]);
        }
    for (const auto& r : cos)
        for (int k = 0; k < 2; k++) {
            val[r.idx[k]][ST_COSSERAT]++;
            nb[r.idx[k]].push_back(r.idx[1 - k]);
        }
    for (const auto& g : ground) val[g.p][ST_GROUND]++;
    for (const auto& g : rground) val[g.p][ST_RIGID_GROUND]++;
    for (const auto& j : joints) val[j.p][ST_JOINT]++;
    // a two-sided joint (new) couples its two bodies like a rod element
    for (const auto& j : joints2)
        for (int k = 0; k < 2; k++) {
            val[j.p[k]][ST_JOINT]++;
            nb[j.p[k]].push_back(j.p[1 - k]);
        }

    adj.e_offsets.assign(n + 1, {});
    adj.e_offsets[0][0] = 0;
    adj.p_offsets.assign(n + 1, 0);
    for (uint32_t i = 0; i < n; i++) {
        for (int e = 0; e < ST_COUNT; e++) {
            if (e == ST_COUNT - 1) adj.e_offsets[i + 1][0] = adj.e_offsets[i][e] + val[i][e];
            else adj.e_offsets[i][e + 1] = adj.e_offsets[i][e] + val[i][e];
        }
        auto& s = nb[i];
        std::sort(s.begin(), s.end());
        s.erase(std::unique(s.begin(), s.end()), s.end());
        adj.p_offsets[i + 1] = adj.p_offsets[i] + (uint32_t)s.size();
    }
    // the last row only carries the total in slot 0; fill the rest with the
    // total so that e_offsets[n][t] is a valid end for every type
    for (int e = 1; e < ST_COUNT; e++) adj.e_offsets[n][e] = adj.e_offsets[n][0];

    adj.e_entries.assign(adj.e_offsets[n][0], AdjEntry{});
    auto cursor = adj.e_offsets;
    for (uint32_t e = 0; e < nh.size(); e++)
        for (uint16_t k = 0; k < 4; k++) {
            uint32_t p = nh[e].idx[k];
            adj.e_entries[cursor[p][ST_NEO_HOOKEAN]++] = AdjEntry{ST_NEO_HOOKEAN, e, k};
        }
    for (uint32_t e = 0; e < cos.size(); e++)
        for (uint16_t k = 0; k < 2; k++) {
            uint32_t p = cos[e].idx[k];
            adj.e_entries[cursor[p][ST_COSSERAT]++] = AdjEntry{ST_COSSERAT, e, k};
        }
    for (uint32_t e = 0; e < ground.size(); e++) {
        uint32_t p = ground[e].p;
        adj.e_entries[cursor[p][ST_GROUND]++] = AdjEntry{ST_GROUND, e, 0};
    }
    for (uint32_t e = 0; e < rground.size(); e++) {
        uint32_t p = rground[e].p;
        adj.e_entries[cursor[p][ST_RIGID_GROUND]++] = AdjEntry{ST_RIGID_GROUND, e, 0};
    }
    for (uint32_t e = 0; e < joints.size(); e++) {
        uint32_t p = joints[e].p;
        adj.e_entries[cursor[p][ST_JOINT]++] = AdjEntry{ST_JOINT, e, 0};
    }
    for (uint32_t e = 0; e < joints2.size(); e++)
        for (uint16_t k = 0; k < 2; k++) {
            uint32_t p = joints2[e].p[k];
            adj.e_entries[cursor[p][ST_JOINT]++] = AdjEntry{ST_JOINT, (uint32_t)joints.size() + e, k};
        }
    adj.p_neighbors.resize(adj.p_offsets[n]);
    for (uint32_t i = 0; i < n; i++)
        std::copy(nb[i].begin(), nb[i].end(), adj.p_neighbors.begin() + adj.p_offsets[i]);
}

// ColorList::greedyColor2 (DSATUR), common/ColorList.hpp:149-241.
// Pick: max saturation; tie -> max degree (p_offsets difference); tie -> lowest
// index (strict comparisons, scan in index order).  Colour: lowest colour not
// used by a coloured neighbour.  num_colors = number of colours ever used.
void dsatur(const StaticAdjacency& adj, uint32_t n, Coloring& col) {
    const uint32_t UNC = 0xFFFFFFFFu;
    col.color.assign(n, UNC);
    std::vector<uint32_t> sat(n, 0);
    std::vector<uint8_t> used;
    auto deg = [&](uint32_t v) { return adj.p_offsets[v + 1] - adj.p_offsets[v]; };
    for (uint32_t i = 0; i < n; i++) {
        uint32_t best = UNC;
        for (uint32_t j = 0; j < n; j++) {
            if (col.color[j] != UNC) continue;
            if (best == UNC) best = j;
            if (sat[j] > sat[best]) best = j;
            else if (sat[j] == sat[best] && deg(j) > deg(best)) best = j;
        }
        if (best == UNC) throw std::runtime_error("best color is UNCOLORED!");
        std::fill(used.begin(), used.end(), 0);
        for (uint32_t e = adj.p_offsets[best]; e < adj.p_offsets[best + 1]; e++) {
            uint32_t nbv = adj.p_neighbors[e];
            if (col.color[nbv] != UNC) {
                if (col.color[nbv] >= used.size()) used.resize(col.color[nbv] + 1, 0);
                used[col.color[nbv]] = 1;
            }
        }
        uint32_t c = 0;
        while (c < used.size() && used[c]) c++;
        col.color[best] = c;
        if (used.size() < c + 1) used.resize(c + 1, 0);
        for (uint32_t e = adj.p_offsets[best]; e < adj.p_offsets[best + 1]; e++) {
            uint32_t nbv = adj.p_neighbors[e];
            if (col.color[nbv] != UNC) continue;
            std::fill(used.begin(), used.end(), 0);
            for (uint32_t k = adj.p_offsets[nbv]; k < adj.p_offsets[nbv + 1]; k++) {
                uint32_t nn = adj.p_neighbors[k];
                if (col.color[nn] != UNC) used[col.color[nn]] = 1;
            }
            uint32_t cnt = 0;
            for (auto u : used) cnt += u ? 1 : 0;
            sat[nbv] = cnt;
        }
    }
    col.num_colors = (uint32_t)used.size();
}

// ColorList::rebuildWorkList + countColors, common/ColorList.hpp:243-253,345-363
void rebuild_work_list(uint32_t n, Coloring& col) {
    col.color_counts.assign(col.num_colors, 0);
    for (uint32_t v = 0; v < n; v++) col.color_counts[col.color[v]]++;
    col.color_offsets.assign(col.num_colors + 1, 0);
    for (uint32_t c = 0; c < col.num_colors; c++) col.color_offsets[c + 1] = col.color_offsets[c] + col.color_counts[c];
    col.work_list.assign(n, 0);
    auto cursor = col.color_offsets;
    for (uint32_t v = 0; v < n; v++) col.work_list[cursor[col.color[v]]++] = v;
}

}  // namespace oracle
