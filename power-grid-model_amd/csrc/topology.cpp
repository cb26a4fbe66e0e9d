// topology.cpp -- bus ordering (bit-exact with the reference) and the device factorisation plan.
#include "topology.hpp"

#include <algorithm>
#include <limits>
#include <numeric>
#include <string>
#include <set>
#include <stdexcept>

namespace pgm_amd {

// ------------------------------------------------------------------------------------------------
// minimum degree ordering -- same steps as sparse_ordering.hpp:20-219 (vertex with the smallest
// degree, ties to the smallest vertex id; indistinguishable neighbours eliminated together; the
// remaining neighbours made a clique, every new edge recorded as a fill-in in clique-map order).
// The adjacency map is held in vectors indexed by vertex id with a presence flag.
// ------------------------------------------------------------------------------------------------
namespace {
struct AdjGraph {
    std::vector<std::vector<Idx>> adj;
    std::vector<char> present;
    Idx n_present = 0;
    explicit AdjGraph(Idx n) : adj(n), present(n, 0) {}
    std::vector<Idx>& at(Idx v) {
        if (!present[v]) throw std::out_of_range("minimum degree: vertex not in graph");
        return adj[v];
    }
    std::vector<Idx>& get_or_create(Idx v) {
        if (!present[v]) {
            present[v] = 1;
            ++n_present;
            adj[v].clear();
        }
        return adj[v];
    }
    void erase(Idx v) {
        if (present[v]) {
            present[v] = 0;
            --n_present;
            adj[v].clear();
        }
    }
    bool has_edge(Idx a, Idx b) const {
        return present[a] && std::find(adj[a].begin(), adj[a].end(), b) != adj[a].end();
    }
    Idx first_present() const {
        for (Idx v = 0; v < Idx(present.size()); ++v)
            if (present[v]) return v;
        return -1;
    }
};
struct DegreeSet {
    std::set<std::pair<Idx, Idx>> by_degree;  // (degree, vertex)
    std::vector<Idx> deg;
    explicit DegreeSet(Idx n) : deg(n, -1) {}
    void set(Idx v, Idx d) {
        if (deg[v] >= 0) by_degree.erase({deg[v], v});
        deg[v] = d;
        by_degree.insert({d, v});
    }
    void erase(Idx v) {
        if (deg[v] >= 0) {
            by_degree.erase({deg[v], v});
            deg[v] = -1;
        }
    }
    Idx min_vertex() const { return by_degree.begin()->second; }
};
}  // namespace

std::pair<std::vector<Idx>, std::vector<std::pair<Idx, Idx>>>
minimum_degree(std::vector<Idx> const& keys, std::vector<std::vector<Idx>> const& adjacency) {
    Idx nv = 0;
    for (size_t i = 0; i < keys.size(); ++i) {
        nv = std::max(nv, keys[i] + 1);
        for (Idx e : adjacency[i]) nv = std::max(nv, e + 1);
    }
    AdjGraph g(nv);
    for (size_t i = 0; i < keys.size(); ++i) {
        auto& a = g.get_or_create(keys[i]);
        a.insert(a.end(), adjacency[i].begin(), adjacency[i].end());
    }
    // symmetric closure with sorted unique neighbour lists
    for (size_t i = 0; i < keys.size(); ++i)
        for (Idx e : adjacency[i]) g.get_or_create(e).push_back(keys[i]);
    for (Idx v = 0; v < nv; ++v)
        if (g.present[v]) {
            auto& a = g.adj[v];
            std::sort(a.begin(), a.end());
            a.erase(std::unique(a.begin(), a.end()), a.end());
        }
    DegreeSet dgd(nv);
    for (Idx v = 0; v < nv; ++v)
        if (g.present[v]) dgd.set(v, Idx(g.adj[v].size()));
    Idx const n = g.n_present;
    std::vector<Idx> alpha;
    std::vector<std::pair<Idx, Idx>> fills;
    for (Idx k = 0; k < n; ++k) {
        Idx const u = dgd.min_vertex();
        alpha.push_back(u);
        if (g.n_present == 2) {
            Idx const from = g.first_present();
            Idx const to = g.adj[from][0];
            alpha.push_back(alpha.back() == from ? to : from);
            return {alpha, fills};
        }
        // neighbours of u, and those indistinguishable from u
        std::vector<Idx> nbs = g.at(u);
        std::vector<Idx> lu = nbs;
        lu.push_back(u);
        std::sort(lu.begin(), lu.end());
        std::vector<Idx> rl;
        for (Idx v : nbs) {
            std::vector<Idx> lv = g.at(v);
            lv.push_back(v);
            std::sort(lv.begin(), lv.end());
            if (lu == lv) rl.push_back(v);
        }
        std::vector<Idx> const eliminated_with_u = rl;
        rl.push_back(u);
        for (Idx uu : rl) {
            if (uu != u) nbs.erase(std::remove(nbs.begin(), nbs.end(), uu), nbs.end());
            dgd.erase(uu);
            std::vector<Idx> empty_after;
            std::vector<Idx> const around = g.get_or_create(uu);
            for (Idx e : around) {
                auto& a = g.get_or_create(e);
                a.erase(std::remove(a.begin(), a.end(), uu), a.end());
                if (a.empty()) empty_after.push_back(e);
            }
            empty_after.push_back(uu);
            for (Idx e : empty_after) g.erase(e);
        }
        // clique on the remaining neighbours, visited in increasing vertex order
        std::vector<Idx> order = nbs;
        std::sort(order.begin(), order.end());
        for (Idx kk : order) {
            for (Idx e : nbs) {
                if (e == kk) continue;
                if (!g.has_edge(kk, e)) {
                    g.get_or_create(kk).push_back(e);
                    g.get_or_create(e).push_back(kk);
                    fills.emplace_back(kk, e);
                }
            }
        }
        for (Idx e : nbs) dgd.set(e, g.present[e] ? Idx(g.adj[e].size()) : 0);
        alpha.insert(alpha.end(), eliminated_with_u.begin(), eliminated_with_u.end());
        if (g.n_present == 0) return {alpha, fills};
    }
    return {alpha, fills};
}

// ------------------------------------------------------------------------------------------------
// topology (topology.hpp:196-368 for the bus order, :370-560 for the coupling of components)
// ------------------------------------------------------------------------------------------------
Topology build_topology(Model const& m, ComponentState const& st, std::vector<char> const* couple_open) {
    Idx const nn = m.n_node();
    Idx const nb = m.n_branch();
    Idx const nl = m.n_line();
    auto bfrom = [&](Idx b) { return b < nl ? m.line_f[b] : m.tr_f[b - nl]; };
    auto bto = [&](Idx b) { return b < nl ? m.line_t[b] : m.tr_t[b - nl]; };
    auto bfs = [&](Idx b) { return b < nl ? st.line_fs[b] != 0 : st.tr_fs[b - nl] != 0; };
    auto bts = [&](Idx b) { return b < nl ? st.line_ts[b] != 0 : st.tr_ts[b - nl] != 0; };

    // out-edges grouped by vertex, each group in insertion order: for branch (i, j) the edge i->j
    // precedes j->i, branches in component order
    std::vector<Idx> deg(nn + 1, 0);
    for (Idx b = 0; b < nb; ++b)
        if (bfs(b) && bts(b) && bfrom(b) != bto(b)) {
            ++deg[bfrom(b) + 1];
            ++deg[bto(b) + 1];
        }
    std::partial_sum(deg.begin(), deg.end(), deg.begin());
    std::vector<Idx> target(deg[nn]);
    // phase shift along an edge: theta_to = theta_from - shift for the branch's from->to direction
    std::vector<double> edge_shift(deg[nn]);
    // line slots carry the generic branches' angle theta (0 for lines and links), transformer slots their clock
    auto shift = [&](Idx b) { return branch_phase_shift(m, b); };
    {
        std::vector<Idx> fillp(deg.begin(), deg.end() - 1);
        for (Idx b = 0; b < nb; ++b)
            if (bfs(b) && bts(b) && bfrom(b) != bto(b)) {
                edge_shift[fillp[bfrom(b)]] = -shift(b);
                target[fillp[bfrom(b)]++] = bto(b);
                edge_shift[fillp[bto(b)]] = shift(b);
                target[fillp[bto(b)]++] = bfrom(b);
            }
    }
    Topology t;
    t.node_group.assign(nn, -1);
    t.node_pos.assign(nn, -1);
    t.node_phase.assign(nn, 0.0);
    std::vector<Idx> parent(nn);
    std::iota(parent.begin(), parent.end(), Idx{0});
    std::vector<char> state(nn, 0);      // 0 new, 1 on the DFS stack, 2 finished
    std::vector<char> in_cycle(nn, 0);   // on a cycle or between the source and a cycle
    std::vector<Idx> cursor(nn);
    for (Idx s = 0; s < Idx(m.source.size()); ++s) {
        if (!st.src_status[s]) continue;
        Idx const root = m.src_node[s];
        if (t.node_group[root] != -1) continue;
        Idx const grp = Idx(t.models.size());
        std::vector<Idx> order;
        std::vector<std::pair<Idx, Idx>> back;
        std::vector<Idx> stack{root};
        state[root] = 1;
        t.node_group[root] = grp;
        order.push_back(root);
        cursor[root] = deg[root];
        while (!stack.empty()) {
            Idx const u = stack.back();
            if (cursor[u] == deg[u + 1]) {
                state[u] = 2;
                stack.pop_back();
                continue;
            }
            Idx const e = cursor[u]++;
            Idx const v = target[e];
            if (state[v] == 0) {
                t.node_phase[v] = t.node_phase[u] + edge_shift[e];
                parent[v] = u;
                state[v] = 1;
                t.node_group[v] = grp;
                order.push_back(v);
                cursor[v] = deg[v];
                stack.push_back(v);
            } else if (state[v] == 1 && parent[u] != v) {
                back.emplace_back(u, v);
            }
        }
        MathModel mm;
        mm.radial = back.empty();
        if (back.empty()) {
            mm.bus_node.assign(order.rbegin(), order.rend());
        } else {
            for (auto const& e : back)
                for (Idx x = e.first; !in_cycle[x]; x = parent[x]) in_cycle[x] = 1;
            for (auto it = order.rbegin(); it != order.rend(); ++it)
                if (!in_cycle[*it]) mm.bus_node.push_back(*it);
            std::vector<Idx> cyc;
            for (Idx x : order)
                if (in_cycle[x]) cyc.push_back(x);
            if (cyc.size() < 4) {
                mm.bus_node.insert(mm.bus_node.end(), cyc.rbegin(), cyc.rend());
            } else {
                // adjacency map: each cyclic node to its DFS parent, then the back edges not yet present
                std::vector<Idx> keys;
                std::vector<std::vector<Idx>> adj;
                std::vector<Idx> slot(nn, -1);
                std::vector<Idx> cyc_sorted = cyc;
                std::sort(cyc_sorted.begin(), cyc_sorted.end());
                auto key_slot = [&](Idx x) -> Idx {
                    if (slot[x] < 0) {
                        slot[x] = Idx(keys.size());
                        keys.push_back(x);
                        adj.emplace_back();
                    }
                    return slot[x];
                };
                for (Idx x : cyc_sorted)
                    if (parent[x] != x) adj[key_slot(x)].push_back(parent[x]);
                for (auto const& [a, b] : back) {
                    bool present = slot[a] >= 0 &&
                                   std::find(adj[slot[a]].begin(), adj[slot[a]].end(), b) != adj[slot[a]].end();
                    if (!present) adj[key_slot(a)].push_back(b);
                }
                // keys must be visited in increasing order
                std::vector<Idx> perm(keys.size());
                std::iota(perm.begin(), perm.end(), Idx{0});
                std::sort(perm.begin(), perm.end(), [&](Idx a, Idx b) { return keys[a] < keys[b]; });
                std::vector<Idx> skeys;
                std::vector<std::vector<Idx>> sadj;
                for (Idx p : perm) {
                    skeys.push_back(keys[p]);
                    sadj.push_back(adj[p]);
                }
                auto [reordered, fl] = minimum_degree(skeys, sadj);
                Idx const base = Idx(mm.bus_node.size());
                std::vector<Idx> newpos(nn, -1);
                for (Idx i = 0; i < Idx(reordered.size()); ++i) newpos[reordered[i]] = base + i;
                mm.bus_node.insert(mm.bus_node.end(), reordered.begin(), reordered.end());
                for (auto const& [a, b] : fl) mm.fill_in.push_back({newpos[a], newpos[b]});
            }
        }
        mm.n_bus = Idx(mm.bus_node.size());
        for (Idx i = 0; i < mm.n_bus; ++i) t.node_pos[mm.bus_node[i]] = i;
        mm.slack = t.node_pos[root];
        t.models.push_back(std::move(mm));
    }
    Idx const nmm = Idx(t.models.size());
    // branches: coupled to the model of the first connected, energised side
    t.branch_group.assign(nb, -1);
    t.branch_pos.assign(nb, -1);
    for (Idx b = 0; b < nb; ++b) {
        Idx const f = bfrom(b), to = bto(b);
        Idx g = -1;
        if (bfs(b) && t.node_group[f] != -1)
            g = t.node_group[f];
        else if (bts(b) && t.node_group[to] != -1)
            g = t.node_group[to];
        else if (couple_open && (*couple_open)[b] && t.node_group[f] != -1 && t.node_group[f] == t.node_group[to])
            g = t.node_group[f];  // open on both sides here, closed in some lane of a union group
        if (g < 0) continue;
        auto& mm = t.models[g];
        t.branch_group[b] = g;
        t.branch_pos[b] = Idx(mm.branch.size());
        mm.branch.push_back(b);
        mm.branch_bus.push_back({bfs(b) ? t.node_pos[f] : -1, bts(b) ? t.node_pos[to] : -1});
    }
    // appliances: grouped by bus, component order kept within a bus
    auto couple = [&](std::vector<Idx> const& node_of, std::vector<char> const& include,
                      std::vector<Idx> MathModel::*list, std::vector<Idx>& grp, std::vector<Idx>& pos) {
        Idx const n = Idx(node_of.size());
        grp.assign(n, -1);
        pos.assign(n, -1);
        std::vector<std::vector<Idx>> count(nmm);
        for (Idx g = 0; g < nmm; ++g) count[g].assign(t.models[g].n_bus + 1, 0);
        for (Idx i = 0; i < n; ++i)
            if (include[i] && t.node_group[node_of[i]] >= 0) ++count[t.node_group[node_of[i]]][t.node_pos[node_of[i]] + 1];
        for (Idx g = 0; g < nmm; ++g) {
            std::partial_sum(count[g].begin(), count[g].end(), count[g].begin());
            (t.models[g].*list).assign(count[g].back(), -1);
        }
        for (Idx i = 0; i < n; ++i) {
            if (!include[i]) continue;
            Idx const g = t.node_group[node_of[i]];
            if (g < 0) continue;
            Idx const p = count[g][t.node_pos[node_of[i]]]++;
            (t.models[g].*list)[p] = i;
            grp[i] = g;
            pos[i] = p;
        }
    };
    std::vector<char> all_sh(m.shunt.size(), 1), all_lg(m.lg_id.size(), 1), src_on(m.source.size());
    for (size_t i = 0; i < m.source.size(); ++i) src_on[i] = st.src_status[i] != 0;
    couple(m.sh_node, all_sh, &MathModel::shunt, t.sh_group, t.sh_pos);
    couple(m.lg_node, all_lg, &MathModel::load, t.lg_group, t.lg_pos);
    couple(m.src_node, src_on, &MathModel::source, t.src_group, t.src_pos);
    return t;
}

// ------------------------------------------------------------------------------------------------
// device plan
// ------------------------------------------------------------------------------------------------

void build_schedule(Plan& p, int target_parallel_tasks) {
    Idx const n = p.n;
    auto find = [&](int r, int c) {
        auto const b = p.col.begin() + p.row_ptr[r], e = p.col.begin() + p.row_ptr[r + 1];
        auto const it = std::lower_bound(b, e, c);
        if (it == e || *it != c) throw std::logic_error("LU pattern entry missing");
        return int(it - p.col.begin());
    };
    // Schur updates in pull form: for pivot k, every L entry (i,k) times every U entry (k,j) updates
    // (i,j); the update belongs to pivot min(i,j) and is listed in increasing k
    std::vector<std::vector<std::array<int, 3>>> own(n);
    for (Idx k = 0; k < n; ++k) {
        for (int lu = p.diag[k] + 1; lu < p.row_ptr[k + 1]; ++lu) {
            int const i = p.col[lu];
            int const l_entry = p.transpose[lu];
            for (int uu = p.diag[k] + 1; uu < p.row_ptr[k + 1]; ++uu) {
                int const j = p.col[uu];
                int const tgt = find(i, j);
                own[std::min(i, j)].push_back({tgt, l_entry, uu});
            }
        }
    }
    p.schur_ptr.assign(n + 1, 0);
    for (Idx r = 0; r < n; ++r) p.schur_ptr[r + 1] = p.schur_ptr[r] + int(own[r].size());
    p.schur.reserve(p.schur_ptr[n]);
    for (Idx r = 0; r < n; ++r) p.schur.insert(p.schur.end(), own[r].begin(), own[r].end());
    p.keep.assign(n, 0);
    for (Idx r = 0; r < n; ++r) {
        bool k = p.row_ptr[r + 1] - p.diag[r] - 1 == 1;
        for (auto const& u : own[r]) k = k && u[0] == p.diag[r];
        p.keep[r] = k ? 1 : 0;
    }

    // tasks over the elimination tree: parent(p) = first column right of the diagonal in row p
    std::vector<int> parent(n, -1);
    std::vector<std::vector<int>> children(n);
    for (Idx r = 0; r < n; ++r)
        if (p.diag[r] + 1 < p.row_ptr[r + 1]) {
            parent[r] = p.col[p.diag[r] + 1];
            children[parent[r]].push_back(int(r));
        }
    std::vector<double> w(n);
    for (Idx r = 0; r < n; ++r) w[r] = 2.0 + (p.row_ptr[r + 1] - p.row_ptr[r]) + (p.schur_ptr[r + 1] - p.schur_ptr[r]);
    std::vector<int> level_of(n, -1);
    std::vector<char> remaining(n, 1);
    Idx n_rem = n;
    // one task size bound for every level: maximal subtrees of the remaining forest up to T form the
    // tasks of the next level, so a task never waits on a task of its own level
    int const max_levels = 256;
    p.level_ptr.push_back(0);
    p.task_ptr.push_back(0);
    std::vector<double> W(n);
    double total0 = 0;
    for (Idx r = 0; r < n; ++r) total0 += w[r];
    double const T = std::max(48.0, total0 / std::max(1, target_parallel_tasks));
    for (int lvl = 0; n_rem > 0; ++lvl) {
        // subtree weights within the remaining forest (children have smaller indices)
        double total = 0;
        for (Idx r = 0; r < n; ++r) {
            if (!remaining[r]) continue;
            W[r] = w[r];
            for (int c : children[r])
                if (remaining[c]) W[r] += W[c];
            total += w[r];
        }
        bool const last = lvl == max_levels - 1 || total <= T;
        std::vector<int> roots;
        for (Idx r = 0; r < n; ++r) {
            if (!remaining[r]) continue;
            bool const parent_rem = parent[r] >= 0 && remaining[parent[r]];
            bool has_child = false;
            for (int c : children[r]) has_child = has_child || remaining[c];
            // a remaining leaf heavier than T still forms a task of its own
            bool const small = W[r] <= T || !has_child;
            if (last ? !parent_rem : (small && (!parent_rem || W[parent[r]] > T))) roots.push_back(int(r));
        }
        std::vector<std::pair<double, int>> level_tasks;  // (weight, task)
        for (int root : roots) {
            // collect the subtree of root among remaining pivots, in increasing order
            std::vector<int> members{root};
            for (size_t q = 0; q < members.size(); ++q)
                for (int c : children[members[q]])
                    if (remaining[c]) members.push_back(c);
            std::sort(members.begin(), members.end());
            double tw = 0;
            for (int v : members) {
                p.task_pivot.push_back(v);
                remaining[v] = 0;
                level_of[v] = lvl;
                tw += w[v];
            }
            n_rem -= Idx(members.size());
            p.task_ptr.push_back(int(p.task_pivot.size()));
            p.task_mm.push_back(p.bus_mm[root]);
            level_tasks.emplace_back(tw, int(p.task_mm.size()) - 1);
        }
        // heaviest tasks first: waves start longest-first, and packed waves (several tasks of a level
        // per wave) hold tasks of similar length
        std::stable_sort(level_tasks.begin(), level_tasks.end(),
                         [](auto const& a, auto const& b) { return a.first > b.first; });
        for (auto const& lt : level_tasks) p.level_task.push_back(lt.second);
        p.level_ptr.push_back(int(p.level_task.size()));
        if (roots.empty()) throw std::logic_error("task decomposition made no progress");
    }
}

// ------------------------------------------------------------------------------------------------
Plan build_plan(Model const& m, ComponentState const& st, Topology const& topo, std::vector<Idx> const& models,
                bool symmetric, int target_parallel_tasks) {
    Plan p;
    p.P = symmetric ? 1 : 3;
    int const PP = p.P * p.P;
    p.n_mm = Idx(models.size());
    p.mm_bus_begin.push_back(0);
    for (Idx g : models) p.mm_bus_begin.push_back(p.mm_bus_begin.back() + int(topo.models[g].n_bus));
    p.n = p.mm_bus_begin.back();
    Idx const n = p.n;
    p.bus_mm.resize(n);
    p.phase.resize(n);
    for (Idx k = 0; k < p.n_mm; ++k) {
        auto const& mm = topo.models[models[k]];
        for (Idx i = 0; i < mm.n_bus; ++i) p.phase[p.mm_bus_begin[k] + i] = topo.node_phase[mm.bus_node[i]];
    }
    for (Idx k = 0; k < p.n_mm; ++k)
        for (int b = p.mm_bus_begin[k]; b < p.mm_bus_begin[k + 1]; ++b) p.bus_mm[b] = int(k);

    // pattern: diagonal of every bus, both directions of every closed branch, fill-ins
    std::vector<std::vector<int>> rows(n);
    for (Idx b = 0; b < n; ++b) rows[b].push_back(int(b));
    std::vector<std::vector<int>> fill_rows(n);
    for (Idx k = 0; k < p.n_mm; ++k) {
        auto const& mm = topo.models[models[k]];
        int const off = p.mm_bus_begin[k];
        for (auto const& bb : mm.branch_bus)
            if (bb[0] >= 0 && bb[1] >= 0) {
                rows[off + bb[0]].push_back(int(off + bb[1]));
                rows[off + bb[1]].push_back(int(off + bb[0]));
            }
        for (auto const& f : mm.fill_in) {
            rows[off + f[0]].push_back(int(off + f[1]));
            rows[off + f[1]].push_back(int(off + f[0]));
            fill_rows[off + f[0]].push_back(int(off + f[1]));
            fill_rows[off + f[1]].push_back(int(off + f[0]));
        }
    }
    p.row_ptr.assign(n + 1, 0);
    for (Idx r = 0; r < n; ++r) {
        auto& v = rows[r];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        p.row_ptr[r + 1] = p.row_ptr[r] + int(v.size());
    }
    p.E = p.row_ptr[n];
    p.col.resize(p.E);
    p.diag.resize(n);
    p.fill.assign(p.E, 0);
    for (Idx r = 0; r < n; ++r) {
        std::copy(rows[r].begin(), rows[r].end(), p.col.begin() + p.row_ptr[r]);
        for (int k = p.row_ptr[r]; k < p.row_ptr[r + 1]; ++k) {
            if (p.col[k] == r) p.diag[r] = k;
            if (std::find(fill_rows[r].begin(), fill_rows[r].end(), p.col[k]) != fill_rows[r].end()) p.fill[k] = 1;
        }
    }
    auto find = [&](int r, int c) {
        auto const b = p.col.begin() + p.row_ptr[r], e = p.col.begin() + p.row_ptr[r + 1];
        auto const it = std::lower_bound(b, e, c);
        if (it == e || *it != c) throw std::logic_error("LU pattern entry missing");
        return int(it - p.col.begin());
    };
    p.transpose.resize(p.E);
    for (Idx r = 0; r < n; ++r)
        for (int k = p.row_ptr[r]; k < p.row_ptr[r + 1]; ++k) p.transpose[k] = find(p.col[k], int(r));

    // admittance per entry: branches in math order (ff, ft, tf, tt), then shunts
    p.ydata.assign(size_t(p.E) * PP, cd{});
    auto add_y = [&](int e, Tensor3 const& y) {
        for (int i = 0; i < PP; ++i) p.ydata[size_t(e) * PP + i] += y.m[i];
    };
    for (Idx k = 0; k < p.n_mm; ++k) {
        auto const& mm = topo.models[models[k]];
        int const off = p.mm_bus_begin[k];
        for (size_t b = 0; b < mm.branch.size(); ++b) {
            auto const y = branch_param(m, st, mm.branch[b], symmetric);
            int const f = mm.branch_bus[b][0], t = mm.branch_bus[b][1];
            p.branch_from.push_back(f >= 0 ? off + f : -1);
            p.branch_to.push_back(t >= 0 ? off + t : -1);
            p.branch_comp.push_back(mm.branch[b]);
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < PP; ++j) p.branch_y.push_back(y.v[i].m[j]);
            if (f >= 0) add_y(find(off + f, off + f), y.v[0]);
            if (f >= 0 && t >= 0) {
                add_y(find(off + f, off + t), y.v[1]);
                add_y(find(off + t, off + f), y.v[2]);
            }
            if (t >= 0) add_y(find(off + t, off + t), y.v[3]);
        }
    }
    p.shunt_ptr.assign(n + 1, 0);
    p.load_ptr.assign(n + 1, 0);
    p.src_ptr.assign(n + 1, 0);
    for (Idx k = 0; k < p.n_mm; ++k) {
        auto const& mm = topo.models[models[k]];
        int const off = p.mm_bus_begin[k];
        for (Idx s : mm.shunt) {
            Idx const bus = off + topo.node_pos[m.sh_node[s]];
            auto const y = shunt_param(m, st, s, symmetric);
            add_y(p.diag[bus], y);
            ++p.shunt_ptr[bus + 1];
            p.shunt_comp.push_back(s);
            for (int j = 0; j < PP; ++j) p.shunt_y.push_back(y.m[j]);
        }
        for (Idx l : mm.load) {
            ++p.load_ptr[off + topo.node_pos[m.lg_node[l]] + 1];
            p.load_comp.push_back(l);
            p.load_type.push_back(m.lg_type[l]);
        }
        for (Idx s : mm.source) {
            ++p.src_ptr[off + topo.node_pos[m.src_node[s]] + 1];
            p.src_comp.push_back(s);
        }
    }
    std::partial_sum(p.shunt_ptr.begin(), p.shunt_ptr.end(), p.shunt_ptr.begin());
    std::partial_sum(p.load_ptr.begin(), p.load_ptr.end(), p.load_ptr.begin());
    std::partial_sum(p.src_ptr.begin(), p.src_ptr.end(), p.src_ptr.begin());

    p.row_end.assign(p.row_ptr.begin() + 1, p.row_ptr.end());
    p.bus_label.resize(n);
    std::iota(p.bus_label.begin(), p.bus_label.end(), 0);
    p.mm_src_ptr.assign(p.n_mm + 1, 0);
    for (Idx b = 0; b < n; ++b) p.mm_src_ptr[p.bus_mm[b] + 1] += p.src_ptr[b + 1] - p.src_ptr[b];
    std::partial_sum(p.mm_src_ptr.begin(), p.mm_src_ptr.end(), p.mm_src_ptr.begin());
    p.mm_src.resize(p.src_comp.size());
    {
        std::vector<int> fillp(p.mm_src_ptr.begin(), p.mm_src_ptr.end() - 1);
        for (Idx b = 0; b < n; ++b)
            for (int q = p.src_ptr[b]; q < p.src_ptr[b + 1]; ++q) p.mm_src[fillp[p.bus_mm[b]]++] = q;
    }

    build_schedule(p, target_parallel_tasks);
    return p;
}

void interleave_plan(Plan& p, int tpw) {
    tpw = std::max(tpw, 1);
    Idx const n = p.n;
    int const PP = p.P * p.P;
    auto row_stride = [&](int b) { return p.row_stride.empty() ? 1 : p.row_stride[b]; };
    auto row_len = [&](int b) { return (p.row_end[b] - p.row_ptr[b]) / row_stride(b); };
    std::vector<int> newbus(n, -1), stride_of(n, 1);
    std::vector<long long> newent(p.E, -1);
    long long B = 0, Ent = 0;
    for (size_t l = 0; l + 1 < p.level_ptr.size(); ++l)
        for (int q0 = p.level_ptr[l]; q0 < p.level_ptr[l + 1]; q0 += tpw) {
            int const q1 = std::min(q0 + tpw, p.level_ptr[l + 1]);
            int K = 0, R = 0;
            for (int q = q0; q < q1; ++q) {
                int const tk = p.level_task[q];
                K = std::max(K, p.task_ptr[tk + 1] - p.task_ptr[tk]);
                for (int i = p.task_ptr[tk]; i < p.task_ptr[tk + 1]; ++i) R = std::max(R, row_len(p.task_pivot[i]));
            }
            // the chunk's stride is its task count: full chunks interleave 64 / S lanes, a one-task chunk
            // stays contiguous
            int const nt = q1 - q0;
            for (int q = q0; q < q1; ++q) {
                int const t = q - q0, tk = p.level_task[q];
                for (int k = 0; k < p.task_ptr[tk + 1] - p.task_ptr[tk]; ++k) {
                    int const piv = p.task_pivot[p.task_ptr[tk] + k];
                    newbus[piv] = int(B + (long long)(k) * nt + t);
                    stride_of[piv] = nt;
                    for (int j = 0; j < row_len(piv); ++j)
                        newent[p.row_ptr[piv] + j * row_stride(piv)] = Ent + ((long long)(k) * R + j) * nt + t;
                }
            }
            B += (long long)(K) * nt;
            Ent += (long long)(K) * R * nt;
        }
    if (B > std::numeric_limits<int>::max() / 2 || Ent > std::numeric_limits<int>::max() / 2)
        throw std::logic_error("interleaved plan too large");
    for (Idx b = 0; b < n; ++b)
        if (newbus[b] < 0) throw std::logic_error("interleave: bus without a task");
    int const nL = int(B), EL = int(Ent);
    std::vector<int> oldbus(nL, -1);
    for (Idx b = 0; b < n; ++b) oldbus[newbus[b]] = int(b);
    // buses
    std::vector<int> row_ptr(nL + 1, 0), row_end(nL, 0), row_str(nL, 1), diag(nL, 0), bus_mm(nL, -1);
    std::vector<int8_t> keep(nL, 0);
    std::vector<double> phase(nL, 0.0);
    for (Idx b = 0; b < n; ++b) {
        int const L = newbus[b];
        row_ptr[L] = int(newent[p.row_ptr[b]]);
        row_end[L] = row_ptr[L] + row_len(int(b)) * stride_of[b];
        row_str[L] = stride_of[b];
        diag[L] = int(newent[p.diag[b]]);
        keep[L] = p.keep[b];
        bus_mm[L] = p.bus_mm[b];
        phase[L] = p.phase[b];
    }
    row_ptr[nL] = EL;
    // entries (fill-flagged zero entries for the padding positions)
    std::vector<int> col(EL, 0), transpose(EL, 0);
    std::vector<int8_t> fill(EL, 1);
    std::vector<cd> ydata(size_t(EL) * PP, cd{});
    for (int e = 0; e < p.E; ++e) {
        if (newent[e] < 0) continue;  // not an entry of any row (a padding position of the old plan)
        int const E2 = int(newent[e]);
        col[E2] = newbus[p.col[e]];
        transpose[E2] = int(newent[p.transpose[e]]);
        fill[E2] = p.fill[e];
        for (int q = 0; q < PP; ++q) ydata[size_t(E2) * PP + q] = p.ydata[size_t(e) * PP + q];
    }
    // Schur lists in new bus order
    std::vector<int> schur_ptr(nL + 1, 0);
    std::vector<std::array<int, 3>> schur;
    schur.reserve(p.schur.size());
    for (int L = 0; L < nL; ++L) {
        int const b = oldbus[L];
        if (b >= 0)
            for (int q = p.schur_ptr[b]; q < p.schur_ptr[b + 1]; ++q)
                schur.push_back({int(newent[p.schur[q][0]]), int(newent[p.schur[q][1]]), int(newent[p.schur[q][2]])});
        schur_ptr[L + 1] = int(schur.size());
    }
    for (auto& v : p.task_pivot) v = newbus[v];
    // appliances regrouped in new bus order
    auto regroup = [&](std::vector<int>& ptr, auto&& move_item) {
        std::vector<int> nptr(nL + 1, 0);
        int cnt = 0;
        for (int L = 0; L < nL; ++L) {
            int const b = oldbus[L];
            if (b >= 0)
                for (int q = ptr[b]; q < ptr[b + 1]; ++q) move_item(q, cnt++);
            nptr[L + 1] = cnt;
        }
        ptr = std::move(nptr);
    };
    {
        auto lt = p.load_type;
        auto lc = p.load_comp;
        regroup(p.load_ptr, [&](int q, int i) {
            p.load_type[i] = lt[q];
            p.load_comp[i] = lc[q];
        });
    }
    {
        auto sc = p.src_comp;
        regroup(p.src_ptr, [&](int q, int i) { p.src_comp[i] = sc[q]; });
    }
    {
        auto hc = p.shunt_comp;
        auto hy = p.shunt_y;
        regroup(p.shunt_ptr, [&](int q, int i) {
            p.shunt_comp[i] = hc[q];
            for (int k = 0; k < PP; ++k) p.shunt_y[size_t(i) * PP + k] = hy[size_t(q) * PP + k];
        });
    }
    for (auto& v : p.branch_from) v = v >= 0 ? newbus[v] : -1;
    for (auto& v : p.branch_to) v = v >= 0 ? newbus[v] : -1;
    for (auto& v : p.bus_label) v = newbus[v];
    p.row_ptr = std::move(row_ptr);
    p.row_end = std::move(row_end);
    p.row_stride = std::move(row_str);
    p.diag = std::move(diag);
    p.keep = std::move(keep);
    p.bus_mm = std::move(bus_mm);
    p.phase = std::move(phase);
    p.col = std::move(col);
    p.transpose = std::move(transpose);
    p.fill = std::move(fill);
    p.ydata = std::move(ydata);
    p.schur_ptr = std::move(schur_ptr);
    p.schur = std::move(schur);
    p.n = nL;
    p.E = EL;
    // sources of every math model, in the new source order
    p.mm_src_ptr.assign(p.n_mm + 1, 0);
    for (int L = 0; L < nL; ++L)
        if (p.bus_mm[L] >= 0) p.mm_src_ptr[p.bus_mm[L] + 1] += p.src_ptr[L + 1] - p.src_ptr[L];
    std::partial_sum(p.mm_src_ptr.begin(), p.mm_src_ptr.end(), p.mm_src_ptr.begin());
    p.mm_src.assign(p.src_comp.size(), 0);
    std::vector<int> fillp(p.mm_src_ptr.begin(), p.mm_src_ptr.end() - 1);
    for (int L = 0; L < nL; ++L)
        for (int q = p.src_ptr[L]; q < p.src_ptr[L + 1]; ++q) p.mm_src[fillp[p.bus_mm[L]]++] = q;
}

void check_plan(Plan const& p) {
    auto fail = [](char const* what) { throw std::logic_error(std::string("plan check: ") + what); };
    auto stride = [&](int r) { return p.row_stride.empty() ? 1 : p.row_stride[r]; };
    auto in_row = [&](int r, int e) {
        return e >= p.row_ptr[r] && e < p.row_end[r] && (e - p.row_ptr[r]) % stride(r) == 0;
    };
    std::vector<char> pivot(p.n, 0);
    for (int tk = 0; tk + 1 < int(p.task_ptr.size()); ++tk)
        for (int i = p.task_ptr[tk]; i < p.task_ptr[tk + 1]; ++i) {
            int const b = p.task_pivot[i];
            if (b < 0 || b >= p.n || pivot[b]) fail("task pivots");
            pivot[b] = 1;
        }
    for (int r = 0; r < p.n; ++r) {
        if (!pivot[r]) {
            if (p.bus_mm[r] >= 0) fail("bus not in a task");
            continue;
        }
        // L entries (columns eliminated before r) precede the diagonal, U entries follow it in
        // increasing column order
        int prev = r, ndiag = 0;
        for (int k = p.row_ptr[r]; k < p.row_end[r]; k += stride(r)) {
            int const c = p.col[k];
            if (k < p.diag[r] && c >= r) fail("L entry after the diagonal column");
            if (k > p.diag[r]) {
                if (c <= prev) fail("U part not sorted");
                prev = c;
            }
            if (c == r) {
                ++ndiag;
                if (k != p.diag[r]) fail("diagonal");
            }
            int const t = p.transpose[k];
            if (p.transpose[t] != k || p.col[t] != r || !in_row(c, t)) fail("transpose");
            if (p.fill[t] != p.fill[k]) fail("fill symmetry");
        }
        if (ndiag != 1) fail("one diagonal per row");
        for (int q = p.schur_ptr[r]; q < p.schur_ptr[r + 1]; ++q) {
            int const tgt = p.schur[q][0], le = p.schur[q][1], ue = p.schur[q][2];
            int const kk = p.col[le];
            if (!in_row(kk, ue)) fail("Schur U entry outside its pivot row");
            if (!(in_row(r, tgt) || p.col[tgt] == r)) fail("Schur target");
        }
    }
}

}  // namespace pgm_amd
