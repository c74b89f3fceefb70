// plan.cpp — query preparation on the host (see plan.h).

#include "plan.h"

#include <algorithm>
#include <limits>
#include <map>
#include <queue>

#include "psm_abi.h"

namespace psm {

uint32_t QGraph::max_degree() const {
    uint32_t m = 0;
    for (uint32_t u = 0; u < n; ++u) m = std::max(m, degree(u));
    return m;
}

bool QGraph::has_edge(uint32_t a, uint32_t b) const {
    return std::binary_search(adj.begin() + off[a], adj.begin() + off[a + 1], b);
}

static void index_qgraph(QGraph& q) {
    uint32_t maxl = 0;
    for (uint32_t u = 0; u < q.n; ++u) maxl = std::max(maxl, q.lab[u]);
    q.L = q.n ? maxl + 1 : 0;
    q.nlf.assign(q.n, {});
    for (uint32_t u = 0; u < q.n; ++u) {
        std::map<uint32_t, uint32_t> h;
        for (uint32_t i = q.off[u]; i < q.off[u + 1]; ++i) h[q.lab[q.adj[i]]]++;
        for (auto& kv : h) q.nlf[u].push_back(kv);
    }
    // 2-core: repeatedly drop vertices of degree < 2 (only membership of the
    // core table, graphoperations.cpp:5-76, is ever consulted)
    std::vector<uint32_t> d(q.n), stack;
    q.core2.assign(q.n, 1);
    for (uint32_t u = 0; u < q.n; ++u) {
        d[u] = q.degree(u);
        if (d[u] < 2) { q.core2[u] = 0; stack.push_back(u); }
    }
    while (!stack.empty()) {
        const uint32_t u = stack.back();
        stack.pop_back();
        for (uint32_t i = q.off[u]; i < q.off[u + 1]; ++i) {
            const uint32_t w = q.adj[i];
            if (q.core2[w] && --d[w] < 2) { q.core2[w] = 0; stack.push_back(w); }
        }
    }
    q.core2_size = 0;
    for (uint8_t c : q.core2) q.core2_size += c;
}

int make_qgraph(uint32_t n, const uint32_t* off, const uint32_t* adj, const uint32_t* lab,
                QGraph& q) {
    if (n == 0 || n > PSM_MAX_QUERY_VERTICES) return PSM_EINVAL;
    if (off[0] != 0) return PSM_EINVAL;
    q.n = n;
    q.off.assign(off, off + n + 1);
    for (uint32_t u = 0; u < n; ++u)
        if (q.off[u + 1] < q.off[u]) return PSM_EINVAL;
    q.adj.assign(adj, adj + q.off[n]);
    q.lab.assign(lab, lab + n);
    for (uint32_t u = 0; u < n; ++u) {
        std::sort(q.adj.begin() + q.off[u], q.adj.begin() + q.off[u + 1]);
        for (uint32_t i = q.off[u]; i < q.off[u + 1]; ++i) {
            const uint32_t w = q.adj[i];
            if (w >= n || w == u) return PSM_EINVAL;
            if (i > q.off[u] && q.adj[i - 1] == w) return PSM_EINVAL;
        }
    }
    for (uint32_t u = 0; u < n; ++u)     // undirected: every edge listed both ways
        for (uint32_t i = q.off[u]; i < q.off[u + 1]; ++i)
            if (!std::binary_search(q.adj.begin() + q.off[q.adj[i]], q.adj.begin() + q.off[q.adj[i] + 1], u))
                return PSM_EINVAL;
    index_qgraph(q);
    return PSM_OK;
}

QGraph unit_graph(const QGraph& q) {
    QGraph g;
    g.n = q.n;
    g.off = q.off;
    g.adj = q.adj;
    g.lab.resize(q.n);
    std::vector<uint32_t> remap(q.L, ~0u);
    uint32_t next = 0;
    for (uint32_t u = 0; u < q.n; ++u) {
        uint32_t& r = remap[q.lab[u]];
        if (r == ~0u) r = next++;
        g.lab[u] = r;
    }
    index_qgraph(g);
    return g;
}

// Start = fewest candidates, ties to the larger degree; then repeatedly the
// unvisited neighbour of the visited set with the fewest candidates (ties to
// larger degree, then smaller id). Pivot = the first earlier neighbour.
void gql_order(const QGraph& q, const std::vector<uint32_t>& cc, uint32_t data_V, Plan& p) {
    const uint32_t n = q.n;
    p.order.assign(n, 0);
    p.pivot.assign(n, 0);
    std::vector<uint8_t> in(n, 0), frontier(n, 0);
    auto add = [&](uint32_t u) {
        in[u] = 1;
        for (uint32_t i = q.off[u]; i < q.off[u + 1]; ++i) frontier[q.adj[i]] = 1;
    };
    uint32_t s = 0;
    for (uint32_t u = 1; u < n; ++u)
        if (cc[u] < cc[s] || (cc[u] == cc[s] && q.degree(u) > q.degree(s))) s = u;
    p.order[0] = s;
    add(s);
    for (uint32_t i = 1; i < n; ++i) {
        uint64_t best_c = (uint64_t)data_V + 1;
        uint32_t best = 0;
        for (uint32_t u = 0; u < n; ++u) {
            if (in[u] || !frontier[u]) continue;
            if (cc[u] < best_c) { best_c = cc[u]; best = u; }
            else if (cc[u] == best_c && q.degree(u) > q.degree(best)) best = u;
        }
        p.order[i] = best;
        add(best);
    }
    for (uint32_t i = 1; i < n; ++i)
        for (uint32_t j = 0; j < i; ++j)
            if (q.has_edge(p.order[i], p.order[j])) { p.pivot[i] = p.order[j]; break; }
}

void backward_neighbours(const QGraph& q, Plan& p, bool drop_pivot) {
    const uint32_t n = q.n;
    p.bn.assign(n, {});
    std::vector<uint8_t> seen(n, 0);
    seen[p.order[0]] = 1;
    for (uint32_t i = 1; i < n; ++i) {
        const uint32_t u = p.order[i];
        for (uint32_t k = q.off[u]; k < q.off[u + 1]; ++k) {
            const uint32_t w = q.adj[k];
            if (seen[w] && !(drop_pivot && w == p.pivot[i])) p.bn[i].push_back(w);
        }
        seen[u] = 1;
    }
}

// Children are appended in adjacency order (graphoperations.cpp:156-195).
static void bfs(const QGraph& q, uint32_t root, CflTree& t) {
    const uint32_t n = q.n;
    t.parent.assign(n, 0);
    t.level.assign(n, 0);
    t.children.assign(n, {});
    t.order.clear();
    std::vector<uint8_t> seen(n, 0);
    std::queue<uint32_t> fifo;
    fifo.push(root);
    seen[root] = 1;
    while (!fifo.empty()) {
        const uint32_t u = fifo.front();
        fifo.pop();
        t.order.push_back(u);
        for (uint32_t i = q.off[u]; i < q.off[u + 1]; ++i) {
            const uint32_t w = q.adj[i];
            if (seen[w]) continue;
            seen[w] = 1;
            fifo.push(w);
            t.parent[w] = u;
            t.level[w] = t.level[u] + 1;
            t.children[u].push_back(w);
        }
    }
}

int cfl_filter_plan(const QGraph& q, uint32_t data_L, const std::vector<uint32_t>& ri_off,
                    uint32_t data_max_freq, const NlfCountFn& nlf_count, CflTree& t) {
    // Rank core vertices (all vertices when there is no 2-core) by
    // freq(label)/deg, keep the three smallest through the same max-heap
    // discipline as the reference, then pick the strict minimum of
    // |NLF candidates|/deg in heap pop order.
    using Item = std::pair<uint32_t, double>;
    auto cmp = [](const Item& a, const Item& b) { return a.second < b.second; };
    std::priority_queue<Item, std::vector<Item>, decltype(cmp)> heap(cmp);
    for (uint32_t u = 0; u < q.n; ++u) {
        if (q.core2_size != 0 && !q.core2[u]) continue;
        const uint32_t l = q.lab[u];
        const uint32_t freq = l < data_L ? ri_off[l + 1] - ri_off[l] : 0;
        heap.push({u, freq / (double)q.degree(u)});
    }
    while (heap.size() > 3) heap.pop();
    uint32_t root = 0;
    double best = data_max_freq + 1;
    while (!heap.empty()) {
        const uint32_t u = heap.top().first;
        heap.pop();
        uint32_t c = 0;
        int r = nlf_count(u, c);
        if (r != PSM_OK) return r;
        const double score = c / (double)q.degree(u);
        if (score < best) { best = score; root = u; }
    }
    bfs(q, root, t);
    const uint32_t n = q.n;
    std::vector<uint32_t> rank(n);
    for (uint32_t i = 0; i < n; ++i) rank[t.order[i]] = i;
    t.bn.assign(n, {});
    t.fn.assign(n, {});
    t.under.assign(n, {});
    t.level_off.clear();
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t u = t.order[i];
        if (i == 0 || t.level[u] != t.level[t.order[i - 1]]) t.level_off.push_back(i);
        for (uint32_t k = q.off[u]; k < q.off[u + 1]; ++k) {
            const uint32_t w = q.adj[k];
            if (t.level[w] < t.level[u] || (t.level[w] == t.level[u] && rank[w] < rank[u]))
                t.bn[u].push_back(w);
            else if (t.level[w] == t.level[u])
                t.fn[u].push_back(w);
            else
                t.under[u].push_back(w);
        }
    }
    t.level_off.push_back(n);
    return PSM_OK;
}

int dpiso_filter_plan(const QGraph& q, uint32_t data_V, const NlfCountFn& ldf_count, CflTree& t) {
    // start: strict minimum of |LDF candidates|/deg over vertices of degree
    // 2 or more, from |V(data)|; vertex 0 when there is none
    uint32_t root = 0;
    double best = data_V;
    for (uint32_t u = 0; u < q.n; ++u) {
        if (q.degree(u) <= 1) continue;
        uint32_t c = 0;
        int r = ldf_count(u, c);
        if (r != PSM_OK) return r;
        const double score = c / (double)q.degree(u);
        if (score < best) { best = score; root = u; }
    }
    bfs(q, root, t);
    const uint32_t n = q.n;
    std::vector<uint32_t> rank(n);
    for (uint32_t i = 0; i < n; ++i) rank[t.order[i]] = i;
    t.bn.assign(n, {});
    t.fn.assign(n, {});
    t.under.assign(n, {});
    t.level_off.clear();
    for (uint32_t u = 0; u < n; ++u)
        for (uint32_t k = q.off[u]; k < q.off[u + 1]; ++k) {
            const uint32_t w = q.adj[k];
            (rank[w] < rank[u] ? t.bn[u] : t.fn[u]).push_back(w);
        }
    return PSM_OK;
}

void dpiso_order(const CflTree& t, Plan& p) {
    const uint32_t n = (uint32_t)t.order.size();
    p.order = t.order;
    p.pivot.assign(n, 0);
    for (uint32_t i = 1; i < n; ++i) p.pivot[i] = t.parent[p.order[i]];
}

namespace {

// Estimated embeddings of every suffix of a tree path, from the candidate
// edge tables along it (GenerateQueryPlan.cpp:680-721).
int path_estimates(const std::vector<uint32_t>& path, const TableFn& table,
                   std::vector<uint64_t>& est) {
    const size_t m = path.size();
    est.assign(m - 1, 0);
    std::vector<uint32_t> off, idx;
    int r = table(path[m - 2], path[m - 1], off, idx);
    if (r) return r;
    std::vector<uint64_t> below(off.size() - 1), above;
    uint64_t sum = 0;
    for (size_t j = 0; j + 1 < off.size(); ++j) sum += below[j] = off[j + 1] - off[j];
    est[m - 2] = sum;
    for (size_t i = m - 2; i >= 1; --i) {
        r = table(path[i - 1], path[i], off, idx);
        if (r) return r;
        above.assign(off.size() - 1, 0);
        sum = 0;
        for (size_t j = 0; j + 1 < off.size(); ++j) {
            uint64_t s = 0;
            for (uint32_t k = off[j]; k < off[j + 1]; ++k) s += below[idx[k]];
            above[j] = s;
            sum += s;
        }
        est[i - 1] = sum;
        below.swap(above);
    }
    return PSM_OK;
}

void collect_paths(const QGraph& q, const CflTree& t, uint32_t u, bool core, std::vector<uint32_t>& cur,
                   std::vector<std::vector<uint32_t>>& out) {
    cur.push_back(u);
    bool leaf = true;
    for (uint32_t c : t.children[u]) {
        const bool follow = core ? q.core2[c] != 0 : q.degree(c) > 1;
        if (!follow) continue;
        collect_paths(q, t, c, core, cur, out);
        leaf = false;
    }
    if (leaf && (core || cur.size() > 1)) out.push_back(cur);
    cur.pop_back();
}

}  // namespace

int cfl_order(const QGraph& q, const CflTree& t, const std::vector<uint32_t>& cc,
              const TableFn& table, Plan& p) {
    const uint32_t n = q.n;
    const uint32_t root = t.order[0];
    p.order.assign(n, 0);
    p.pivot.assign(n, 0);
    std::vector<uint8_t> placed(n, 0);
    uint32_t cnt = 0;
    auto place = [&](uint32_t u, uint32_t piv) {
        p.order[cnt] = u;
        p.pivot[cnt] = piv;
        ++cnt;
        placed[u] = 1;
    };
    std::vector<std::vector<uint32_t>> core_paths;
    std::vector<std::vector<std::vector<uint32_t>>> forests;
    std::vector<uint32_t> cur;
    if (q.core2[root]) {
        collect_paths(q, t, root, true, cur, core_paths);
        for (uint32_t u = 0; u < n; ++u) {
            if (!q.core2[u]) continue;
            std::vector<std::vector<uint32_t>> tp;
            collect_paths(q, t, u, false, cur, tp);
            if (!tp.empty()) forests.push_back(std::move(tp));
        }
    } else {
        std::vector<std::vector<uint32_t>> tp;
        collect_paths(q, t, root, false, cur, tp);
        if (!tp.empty()) forests.push_back(std::move(tp));
    }
    place(root, 0);
    constexpr double kMax = std::numeric_limits<double>::max();
    // first index of a path not yet placed, minus one (the path's attachment)
    auto attach = [&](const std::vector<uint32_t>& path, bool tree) -> uint32_t {
        for (uint32_t j = 0; j < path.size(); ++j)
            if (!placed[path[j]]) return tree && j == 0 ? 0 : j - 1;
        return 0;
    };
    if (!core_paths.empty()) {
        std::vector<std::vector<uint64_t>> est(core_paths.size());
        std::vector<uint32_t> weight;
        for (size_t i = 0; i < core_paths.size(); ++i) {
            const auto& path = core_paths[i];
            uint32_t nte = q.degree(path[0]) - (uint32_t)t.children[path[0]].size();
            for (size_t j = 1; j < path.size(); ++j)
                nte += q.degree(path[j]) - (uint32_t)t.children[path[j]].size() - 1;
            weight.push_back(nte + 1);
            int r = path_estimates(path, table, est[i]);
            if (r) return r;
        }
        double bestv = kMax;
        size_t sel = 0;
        for (size_t i = 0; i < core_paths.size(); ++i) {
            const double v = est[i][0] / (double)weight[i];
            if (v < bestv) { bestv = v; sel = i; }
        }
        for (size_t j = 1; j < core_paths[sel].size(); ++j) place(core_paths[sel][j], core_paths[sel][j - 1]);
        core_paths.erase(core_paths.begin() + sel);
        est.erase(est.begin() + sel);
        while (!core_paths.empty()) {
            bestv = kMax;
            sel = 0;
            for (size_t i = 0; i < core_paths.size(); ++i) {
                const uint32_t a = attach(core_paths[i], false);
                const double v = est[i][a] / (double)cc[core_paths[i][a]];
                if (v < bestv) { bestv = v; sel = i; }
            }
            for (size_t j = 1; j < core_paths[sel].size(); ++j)
                if (!placed[core_paths[sel][j]]) place(core_paths[sel][j], core_paths[sel][j - 1]);
            core_paths.erase(core_paths.begin() + sel);
            est.erase(est.begin() + sel);
        }
    }
    for (auto& paths : forests) {
        std::vector<std::vector<uint64_t>> est(paths.size());
        for (size_t i = 0; i < paths.size(); ++i) {
            int r = path_estimates(paths[i], table, est[i]);
            if (r) return r;
        }
        while (!paths.empty()) {
            double bestv = kMax;
            size_t sel = 0;
            for (size_t i = 0; i < paths.size(); ++i) {
                const uint32_t a = attach(paths[i], true);
                const double v = est[i][a] / (double)cc[paths[i][a]];
                if (v < bestv) { bestv = v; sel = i; }
            }
            for (size_t j = 0; j < paths[sel].size(); ++j)
                if (!placed[paths[sel][j]]) place(paths[sel][j], j ? paths[sel][j - 1] : 0);
            paths.erase(paths.begin() + sel);
            est.erase(est.begin() + sel);
        }
    }
    std::vector<uint32_t> leaves;
    for (uint32_t u = 0; u < n; ++u)
        if (q.degree(u) == 1) leaves.push_back(u);
    while (!leaves.empty()) {
        double bestv = kMax;
        size_t sel = 0;
        for (size_t i = 0; i < leaves.size(); ++i) {
            const double v = (double)cc[leaves[i]];
            if (v < bestv) { bestv = v; sel = i; }
        }
        if (!placed[leaves[sel]]) place(leaves[sel], t.parent[leaves[sel]]);
        leaves.erase(leaves.begin() + sel);
    }
    return cnt == n ? PSM_OK : PSM_EINVAL;
}

}  // namespace psm
