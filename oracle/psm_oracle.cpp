// psm_oracle.cpp — CPU restatement of the reference matcher's hot path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker that the GPU path is
// compared against. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it; the product (psm_benchmark_amd) never does.
//
// It restates, in code written for this project, the algorithms of
// scorpiontornado/PSM_Benchmark that sit on the accelerated path. Each
// function names the reference lines it follows. The restatement is pinned
// (tests/test_oracle.py) by
//   * the reference's recorded logs test_anony/C_output/HPRD/L15/Q10/*.txt
//     (embedding counts, call counts and the BuildTable memory checksum for
//     GQL_GQL_GQL, CFL_CFL_EXPLORE and DPiso_DPiso_LFTJ at one thread),
//   * brute-force enumeration on tiny graphs and closed forms (k-cliques),
//   * the reference binary built from its own sources (oracle/_ref) on
//     seeded synthetic graphs, when that binary is present.
//
// Exported C ABI (ctypes): orc_graph_new/free, orc_run/free and accessors.

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <iterator>
#include <limits>
#include <numeric>
#include <queue>
#include <string>
#include <thread>
#include <vector>

namespace {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;  // marks a pruned candidate

// ---------------------------------------------------------------------------
// CSR graph with the derived indexes the reference keeps.
// graph/graph.cpp:15-52 (reverse index, NLF), :92-172 (load), :210-219 (core).
// ---------------------------------------------------------------------------
struct OGraph {
    uint32_t V = 0;
    uint64_t E = 0;                       // undirected edge count
    std::vector<uint32_t> off, adj, lab;
    uint32_t L = 0;                       // labels_count_ = max label + 1
    std::vector<uint32_t> freq;           // per label
    std::vector<uint32_t> ri_off, ri;     // vertices grouped by label, ascending
    std::vector<uint32_t> nlf_off;        // per vertex: sorted (label,count) runs
    std::vector<uint32_t> nlf_lab, nlf_cnt;
    uint32_t max_deg = 0, max_freq = 0;
    std::vector<uint8_t> core2;           // 1 if vertex is in the 2-core
    uint32_t core2_size = 0;

    uint32_t deg(uint32_t v) const { return off[v + 1] - off[v]; }
    const uint32_t* nbrs(uint32_t v) const { return adj.data() + off[v]; }
    uint32_t label_freq(uint32_t l) const { return l < L ? freq[l] : 0; }

    // graph.h:203-234: binary search in the adjacency of the lower-degree end.
    bool has_edge(uint32_t u, uint32_t v) const {
        if (deg(u) < deg(v)) std::swap(u, v);
        const uint32_t* a = nbrs(v);
        return std::binary_search(a, a + deg(v), u);
    }

    void finish() {
        max_deg = 0;
        uint32_t maxl = 0;
        for (uint32_t v = 0; v < V; ++v) {
            std::sort(adj.begin() + off[v], adj.begin() + off[v + 1]);
            max_deg = std::max(max_deg, deg(v));
            maxl = std::max(maxl, lab[v]);
        }
        L = V ? maxl + 1 : 0;
        freq.assign(L, 0);
        for (uint32_t v = 0; v < V; ++v) freq[lab[v]]++;
        max_freq = 0;
        for (uint32_t f : freq) max_freq = std::max(max_freq, f);
        ri_off.assign(L + 1, 0);
        for (uint32_t l = 0; l < L; ++l) ri_off[l + 1] = ri_off[l] + freq[l];
        ri.assign(V, 0);
        std::vector<uint32_t> fill(ri_off.begin(), ri_off.end() - 1);
        for (uint32_t v = 0; v < V; ++v) ri[fill[lab[v]]++] = v;
        // neighbour-label frequency as sorted runs
        nlf_off.assign(V + 1, 0);
        nlf_lab.clear();
        nlf_cnt.clear();
        std::vector<uint32_t> tmp;
        for (uint32_t v = 0; v < V; ++v) {
            tmp.clear();
            for (uint32_t i = off[v]; i < off[v + 1]; ++i) tmp.push_back(lab[adj[i]]);
            std::sort(tmp.begin(), tmp.end());
            for (size_t i = 0; i < tmp.size();) {
                size_t j = i;
                while (j < tmp.size() && tmp[j] == tmp[i]) ++j;
                nlf_lab.push_back(tmp[i]);
                nlf_cnt.push_back((uint32_t)(j - i));
                i = j;
            }
            nlf_off[v + 1] = (uint32_t)nlf_lab.size();
        }
        // 2-core membership: peel vertices of degree < 2 (core value > 1 test
        // of graphoperations.cpp:5-76 only ever asks for membership).
        std::vector<uint32_t> d(V);
        std::vector<uint32_t> stack;
        core2.assign(V, 1);
        for (uint32_t v = 0; v < V; ++v) {
            d[v] = deg(v);
            if (d[v] < 2) { core2[v] = 0; stack.push_back(v); }
        }
        while (!stack.empty()) {
            uint32_t v = stack.back();
            stack.pop_back();
            for (uint32_t i = off[v]; i < off[v + 1]; ++i) {
                uint32_t w = adj[i];
                if (core2[w] && --d[w] < 2) { core2[w] = 0; stack.push_back(w); }
            }
        }
        core2_size = 0;
        for (uint32_t v = 0; v < V; ++v) core2_size += core2[v];
    }
};

// FilterVertices.cpp:786-803: every (label,count) of the query vertex is
// dominated by the data vertex.
bool nlf_dominates(const OGraph& d, uint32_t v, const OGraph& q, uint32_t u) {
    uint32_t i = d.nlf_off[v], ie = d.nlf_off[v + 1];
    for (uint32_t j = q.nlf_off[u]; j < q.nlf_off[u + 1]; ++j) {
        uint32_t l = q.nlf_lab[j];
        while (i < ie && d.nlf_lab[i] < l) ++i;
        if (i == ie || d.nlf_lab[i] != l || d.nlf_cnt[i] < q.nlf_cnt[j]) return false;
    }
    return true;
}

using Cands = std::vector<std::vector<uint32_t>>;

// FilterVertices.cpp:767-813 (computeCandidateWithNLF), :19-46 (LDF).
// `qlabel` is passed separately because the per-unit CFL start-vertex choice
// runs this with the unit graph's remapped labels.
void nlf_candidates(const OGraph& d, const OGraph& q, uint32_t u, bool use_nlf,
                    std::vector<uint32_t>& out) {
    out.clear();
    uint32_t l = q.lab[u];
    if (l >= d.L) return;
    for (uint32_t i = d.ri_off[l]; i < d.ri_off[l + 1]; ++i) {
        uint32_t v = d.ri[i];
        if (d.deg(v) < q.deg(u)) continue;
        if (use_nlf && !nlf_dominates(d, v, q, u)) continue;
        out.push_back(v);
    }
}

// FilterVertices.cpp:19-65: stops at the first empty set (later sets stay
// empty), as the reference returns early.
bool ldf_nlf_filter(const OGraph& d, const OGraph& q, bool use_nlf, Cands& C) {
    C.assign(q.V, {});
    for (uint32_t u = 0; u < q.V; ++u) {
        nlf_candidates(d, q, u, use_nlf, C[u]);
        if (C[u].empty()) return false;
    }
    return true;
}

// Left-saturating bipartite matching test (Kuhn's augmenting paths). The
// reference uses match_bfs (graphoperations.cpp:95-149); whether a matching
// saturating the left side exists does not depend on the algorithm.
bool left_saturating(uint32_t nl, const std::vector<std::vector<uint32_t>>& adjl,
                     uint32_t nr) {
    std::vector<int> match_r(nr, -1);
    std::vector<uint32_t> seen(nr, 0);
    uint32_t stamp = 0;
    // iterative DFS
    for (uint32_t s = 0; s < nl; ++s) {
        ++stamp;
        struct Fr { uint32_t l; uint32_t i; };
        std::vector<Fr> st;
        std::vector<uint32_t> via;  // right vertex chosen at each frame
        st.push_back({s, 0});
        bool found = false;
        while (!st.empty() && !found) {
            Fr& f = st.back();
            if (f.i >= adjl[f.l].size()) {
                st.pop_back();
                if (!via.empty()) via.pop_back();
                continue;
            }
            uint32_t r = adjl[f.l][f.i++];
            if (seen[r] == stamp) continue;
            seen[r] = stamp;
            if (match_r[r] < 0) {
                via.push_back(r);
                found = true;
                break;
            }
            via.push_back(r);
            st.push_back({(uint32_t)match_r[r], 0});
        }
        if (!found) return false;
        // augment: st[k].l is matched to via[k]
        for (size_t k = 0; k < st.size(); ++k) match_r[via[k]] = (int)st[k].l;
    }
    return true;
}

// FilterVertices.cpp:67-143 (GQLFilter) + :704-741 (verifyExactTwigIso):
// NLF, then two in-place (Gauss-Seidel) rounds over u = 0..n-1.
bool gql_filter(const OGraph& d, const OGraph& q, Cands& C) {
    if (!ldf_nlf_filter(d, q, true, C)) return false;
    uint32_t n = q.V;
    std::vector<std::vector<uint8_t>> valid(n, std::vector<uint8_t>(d.V, 0));
    for (uint32_t u = 0; u < n; ++u)
        for (uint32_t v : C[u]) valid[u][v] = 1;
    std::vector<std::vector<uint32_t>> adjl;
    for (int round = 0; round < 2; ++round) {
        for (uint32_t u = 0; u < n; ++u) {
            uint32_t k = q.deg(u);
            for (uint32_t& v : C[u]) {
                if (v == kInvalid) continue;
                uint32_t dv = d.deg(v);
                adjl.assign(k, {});
                for (uint32_t i = 0; i < k; ++i) {
                    uint32_t up = q.nbrs(u)[i];
                    for (uint32_t j = 0; j < dv; ++j)
                        if (valid[up][d.nbrs(v)[j]]) adjl[i].push_back(j);
                }
                if (!left_saturating(k, adjl, dv)) {
                    valid[u][v] = 0;
                    v = kInvalid;
                }
            }
        }
    }
    bool ok = true;
    for (uint32_t u = 0; u < n; ++u) {  // compactCandidates, :743-757
        C[u].erase(std::remove(C[u].begin(), C[u].end(), kInvalid), C[u].end());
        if (C[u].empty()) ok = false;
    }
    return ok;
}

// ---------------------------------------------------------------------------
// BFS tree (graphoperations.cpp:156-195) and CFL filtering plan
// (GenerateFilteringPlan.cpp:15-140, start vertex :343-385).
// ---------------------------------------------------------------------------
struct Tree {
    std::vector<uint32_t> parent, level;
    std::vector<std::vector<uint32_t>> children, bn, fn, under;
    std::vector<uint32_t> order;          // BFS order
    std::vector<uint32_t> level_off;      // level boundaries in `order`
};

void bfs_tree(const OGraph& q, uint32_t root, Tree& t) {
    uint32_t n = q.V;
    t.parent.assign(n, 0);
    t.level.assign(n, 0);
    t.children.assign(n, {});
    t.order.clear();
    std::vector<uint8_t> vis(n, 0);
    std::queue<uint32_t> bq;
    bq.push(root);
    vis[root] = 1;
    while (!bq.empty()) {
        uint32_t u = bq.front();
        bq.pop();
        t.order.push_back(u);
        for (uint32_t i = 0; i < q.deg(u); ++i) {
            uint32_t w = q.nbrs(u)[i];
            if (vis[w]) continue;
            vis[w] = 1;
            bq.push(w);
            t.parent[w] = u;
            t.level[w] = t.level[u] + 1;
            t.children[u].push_back(w);
        }
    }
}

// Start vertex: among 2-core vertices (all, if the 2-core is empty) keep the
// three with the smallest freq(label)/deg via the same max-heap the reference
// uses, then take the strict minimum of |NLF candidates|/deg in pop order.
uint32_t cfl_start_vertex(const OGraph& d, const OGraph& q) {
    auto cmp = [](std::pair<uint32_t, double> a, std::pair<uint32_t, double> b) {
        return a.second < b.second;
    };
    std::priority_queue<std::pair<uint32_t, double>, std::vector<std::pair<uint32_t, double>>,
                        decltype(cmp)> pq(cmp);
    for (uint32_t u = 0; u < q.V; ++u) {
        if (q.core2_size == 0 || q.core2[u]) {
            double rank = d.label_freq(q.lab[u]) / (double)q.deg(u);
            pq.push({u, rank});
        }
    }
    while (pq.size() > 3) pq.pop();
    uint32_t start = 0;
    double best = d.max_freq + 1;
    std::vector<uint32_t> tmp;
    while (!pq.empty()) {
        uint32_t u = pq.top().first;
        nlf_candidates(d, q, u, true, tmp);
        double score = tmp.size() / (double)q.deg(u);
        if (score < best) { start = u; best = score; }
        pq.pop();
    }
    return start;
}

void cfl_plan(const OGraph& d, const OGraph& q, Tree& t) {
    uint32_t root = cfl_start_vertex(d, q);
    bfs_tree(q, root, t);
    uint32_t n = q.V;
    std::vector<uint32_t> pos(n);
    for (uint32_t i = 0; i < n; ++i) pos[t.order[i]] = i;
    t.bn.assign(n, {});
    t.fn.assign(n, {});
    t.under.assign(n, {});
    t.level_off.clear();
    uint32_t cur = kInvalid;
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t u = t.order[i];
        if (t.level[u] != cur) { cur = t.level[u]; t.level_off.push_back(i); }
        for (uint32_t j = 0; j < q.deg(u); ++j) {
            uint32_t w = q.nbrs(u)[j];
            if (t.level[w] == t.level[u]) {
                (pos[w] < pos[u] ? t.bn[u] : t.fn[u]).push_back(w);
            } else if (t.level[w] < t.level[u]) {
                t.bn[u].push_back(w);
            } else {
                t.under[u].push_back(w);
            }
        }
    }
    t.level_off.push_back(n);
}

// FilterVertices.cpp:841-912 (generateCandidates) and :914-966
// (pruneCandidates), restated with set semantics: a data vertex qualifies when
// it has the query vertex's label, enough degree, and a neighbour among the
// live candidates of every pivot.
void pivot_support(const OGraph& d, const OGraph& q, uint32_t u,
                   const std::vector<uint32_t>& pivots, const Cands& C,
                   std::vector<uint32_t>& cnt, std::vector<uint32_t>& touched) {
    touched.clear();
    uint32_t lu = q.lab[u], du = q.deg(u);
    for (uint32_t i = 0; i < pivots.size(); ++i) {
        for (uint32_t v : C[pivots[i]]) {
            if (v == kInvalid) continue;
            for (uint32_t k = d.off[v]; k < d.off[v + 1]; ++k) {
                uint32_t w = d.adj[k];
                if (cnt[w] == i && d.lab[w] == lu && d.deg(w) >= du) {
                    if (i == 0) touched.push_back(w);
                    cnt[w] = i + 1;
                }
            }
        }
    }
}

void cfl_generate(const OGraph& d, const OGraph& q, uint32_t u,
                  const std::vector<uint32_t>& pivots, Cands& C,
                  std::vector<uint32_t>& cnt) {
    std::vector<uint32_t> touched;
    pivot_support(d, q, u, pivots, C, cnt, touched);
    for (uint32_t w : touched)
        if (cnt[w] == pivots.size() && nlf_dominates(d, w, q, u)) C[u].push_back(w);
    for (uint32_t w : touched) cnt[w] = 0;
}

void cfl_prune(const OGraph& d, const OGraph& q, uint32_t u,
               const std::vector<uint32_t>& pivots, Cands& C, std::vector<uint32_t>& cnt) {
    std::vector<uint32_t> touched;
    pivot_support(d, q, u, pivots, C, cnt, touched);
    for (uint32_t& v : C[u])
        if (v != kInvalid && cnt[v] != pivots.size()) v = kInvalid;
    for (uint32_t w : touched) cnt[w] = 0;
}

// FilterVertices.cpp:188-250.
bool cfl_filter(const OGraph& d, const OGraph& q, Cands& C) {
    Tree t;
    cfl_plan(d, q, t);
    uint32_t n = q.V;
    C.assign(n, {});
    nlf_candidates(d, q, t.order[0], true, C[t.order[0]]);
    std::vector<uint32_t> cnt(d.V, 0);
    uint32_t nlev = (uint32_t)t.level_off.size() - 1;
    for (uint32_t i = 1; i < nlev; ++i) {
        for (uint32_t j = t.level_off[i]; j < t.level_off[i + 1]; ++j) {
            uint32_t u = t.order[j];
            cfl_generate(d, q, u, t.bn[u], C, cnt);
        }
        for (int j = (int)t.level_off[i + 1] - 1; j >= (int)t.level_off[i]; --j) {
            uint32_t u = t.order[j];
            if (!t.fn[u].empty()) cfl_prune(d, q, u, t.fn[u], C, cnt);
        }
    }
    for (int i = (int)nlev - 2; i >= 0; --i) {
        for (uint32_t j = t.level_off[i]; j < t.level_off[i + 1]; ++j) {
            uint32_t u = t.order[j];
            if (!t.under[u].empty()) cfl_prune(d, q, u, t.under[u], C, cnt);
        }
    }
    bool ok = true;
    for (uint32_t u = 0; u < n; ++u) {
        C[u].erase(std::remove(C[u].begin(), C[u].end(), kInvalid), C[u].end());
        if (C[u].empty()) ok = false;
    }
    return ok;
}

// ---------------------------------------------------------------------------
// DPiso filter (FilterVertices.cpp:252-294) and its plan
// (GenerateFilteringPlan.cpp:183-214, start vertex :387-406).
// ---------------------------------------------------------------------------
// The query vertex of degree >= 2 with the fewest LDF candidates per unit of
// degree: strict minimum starting from |V(data)|, vertex 0 when no vertex has
// degree 2 or more.
uint32_t dpiso_start_vertex(const OGraph& d, const OGraph& q) {
    double best = d.V;
    uint32_t start = 0;
    std::vector<uint32_t> tmp;
    for (uint32_t u = 0; u < q.V; ++u) {
        if (q.deg(u) <= 1) continue;
        nlf_candidates(d, q, u, false, tmp);
        double score = tmp.size() / (double)q.deg(u);
        if (score < best) { best = score; start = u; }
    }
    return start;
}

// BFS order from the start vertex; bn = neighbours earlier in that order,
// fn = neighbours later, both in adjacency order.
void dpiso_plan(const OGraph& d, const OGraph& q, Tree& t) {
    bfs_tree(q, dpiso_start_vertex(d, q), t);
    uint32_t n = q.V;
    std::vector<uint32_t> pos(n);
    for (uint32_t i = 0; i < n; ++i) pos[t.order[i]] = i;
    t.bn.assign(n, {});
    t.fn.assign(n, {});
    for (uint32_t u = 0; u < n; ++u)
        for (uint32_t j = 0; j < q.deg(u); ++j) {
            uint32_t w = q.nbrs(u)[j];
            (pos[w] < pos[u] ? t.bn[u] : t.fn[u]).push_back(w);
        }
}

// LDF candidates, then three pruning sweeps over the plan order (k = 0, 1, 2):
// even sweeps run forward from the second vertex against its backward
// neighbours, the odd sweep runs backward from the second-to-last vertex
// against its forward neighbours. A vertex without pivots keeps its set.
bool dpiso_filter(const OGraph& d, const OGraph& q, Cands& C) {
    if (!ldf_nlf_filter(d, q, false, C)) return false;
    Tree t;
    dpiso_plan(d, q, t);
    uint32_t n = q.V;
    std::vector<uint32_t> cnt(d.V, 0);
    for (uint32_t k = 0; k < 3; ++k) {
        if (k % 2 == 0) {
            for (uint32_t i = 1; i < n; ++i) {
                uint32_t u = t.order[i];
                cfl_prune(d, q, u, t.bn[u], C, cnt);
            }
        } else {
            for (int i = (int)n - 2; i >= 0; --i) {
                uint32_t u = t.order[i];
                cfl_prune(d, q, u, t.fn[u], C, cnt);
            }
        }
    }
    bool ok = true;
    for (uint32_t u = 0; u < n; ++u) {
        C[u].erase(std::remove(C[u].begin(), C[u].end(), kInvalid), C[u].end());
        if (C[u].empty()) ok = false;
    }
    return ok;
}

// ---------------------------------------------------------------------------
// Candidate-edge tables, BuildTable.cpp:9-151, by their definition:
// T[a][b] row j = ascending positions k with C(b)[k] adjacent to C(a)[j].
// ---------------------------------------------------------------------------
struct Table {
    bool present = false;
    std::vector<uint32_t> off, idx;
};
using Tables = std::vector<std::vector<Table>>;

void build_tables(const OGraph& d, const OGraph& q, const Cands& C, Tables& T) {
    uint32_t n = q.V;
    T.assign(n, std::vector<Table>(n));
    std::vector<uint32_t> pos(d.V, kInvalid);
    for (uint32_t a = 0; a < n; ++a) {
        for (uint32_t i = 0; i < q.deg(a); ++i) {
            uint32_t b = q.nbrs(a)[i];
            for (uint32_t k = 0; k < C[b].size(); ++k) pos[C[b][k]] = k;
            Table& t = T[a][b];
            t.present = true;
            t.off.assign(C[a].size() + 1, 0);
            t.idx.clear();
            for (uint32_t j = 0; j < C[a].size(); ++j) {
                uint32_t v = C[a][j];
                for (uint32_t e = d.off[v]; e < d.off[v + 1]; ++e) {
                    uint32_t p = pos[d.adj[e]];
                    if (p != kInvalid) t.idx.push_back(p);
                }
                t.off[j + 1] = (uint32_t)t.idx.size();
            }
            for (uint32_t v : C[b]) pos[v] = kInvalid;
        }
    }
}

// BuildTable.cpp:300-326, in 4-byte words.
uint64_t memory_words(const OGraph& q, const Cands& C, const Tables& T) {
    uint64_t w = 0;
    for (uint32_t u = 0; u < q.V; ++u) w += C[u].size();
    for (uint32_t a = 0; a < q.V; ++a)
        for (uint32_t b = a + 1; b < q.V; ++b)
            if (q.has_edge(a, b))
                w += T[a][b].idx.size() + C[a].size() + T[b][a].idx.size() + C[b].size();
    return w;
}

// ---------------------------------------------------------------------------
// Query unit for the undecomposed query (SplitGraph.h:622-636 ->
// graph.cpp:585-734): same vertex ids and edges, labels renumbered in order of
// first appearance over vertices 0..n-1.
// ---------------------------------------------------------------------------
OGraph unit_graph(const OGraph& q) {
    OGraph g;
    g.V = q.V;
    g.E = q.E;
    g.off = q.off;
    g.adj = q.adj;
    g.lab.resize(q.V);
    std::vector<uint32_t> remap(q.L, kInvalid);
    uint32_t next = 0;
    for (uint32_t v = 0; v < q.V; ++v) {
        uint32_t l = q.lab[v];
        if (remap[l] == kInvalid) remap[l] = next++;
        g.lab[v] = remap[l];
    }
    g.finish();
    return g;
}

// ---------------------------------------------------------------------------
// Matching orders.
// ---------------------------------------------------------------------------
struct Plan {
    std::vector<uint32_t> order, pivot;
    std::vector<std::vector<uint32_t>> bn;  // per depth
};

// GenerateQueryPlan.cpp:9-58 + :149-184.
void gql_order(const OGraph& q, const Cands& C, uint32_t dataV, Plan& p) {
    uint32_t n = q.V;
    p.order.assign(n, 0);
    p.pivot.assign(n, 0);
    std::vector<uint8_t> vis(n, 0), adjc(n, 0);
    auto take = [&](uint32_t u) {
        vis[u] = 1;
        for (uint32_t i = 0; i < q.deg(u); ++i) adjc[q.nbrs(u)[i]] = 1;
    };
    uint32_t s = 0;
    for (uint32_t u = 1; u < n; ++u) {
        if (C[u].size() < C[s].size() ||
            (C[u].size() == C[s].size() && q.deg(u) > q.deg(s)))
            s = u;
    }
    p.order[0] = s;
    take(s);
    for (uint32_t i = 1; i < n; ++i) {
        uint32_t best = 0;
        uint64_t minv = (uint64_t)dataV + 1;
        for (uint32_t u = 0; u < n; ++u) {
            if (vis[u] || !adjc[u]) continue;
            if (C[u].size() < minv) { minv = C[u].size(); best = u; }
            else if (C[u].size() == minv && q.deg(u) > q.deg(best)) best = u;
        }
        p.order[i] = best;
        take(best);
    }
    for (uint32_t i = 1; i < n; ++i)
        for (uint32_t j = 0; j < i; ++j)
            if (q.has_edge(p.order[i], p.order[j])) { p.pivot[i] = p.order[j]; break; }
}

// GenerateQueryPlan.cpp:680-721: estimated embeddings of each path suffix.
std::vector<uint64_t> path_estimates(const std::vector<uint32_t>& path, const Tables& T) {
    size_t m = path.size();
    std::vector<uint64_t> est(m - 1), child, parent;
    const Table& last = T[path[m - 2]][path[m - 1]];
    uint32_t rows = (uint32_t)last.off.size() - 1;
    child.resize(rows);
    uint64_t sum = 0;
    for (uint32_t i = 0; i < rows; ++i) { child[i] = last.off[i + 1] - last.off[i]; sum += child[i]; }
    est[m - 2] = sum;
    for (int i = (int)m - 2; i >= 1; --i) {
        const Table& e = T[path[i - 1]][path[i]];
        uint32_t r = (uint32_t)e.off.size() - 1;
        parent.assign(r, 0);
        sum = 0;
        for (uint32_t j = 0; j < r; ++j) {
            uint64_t s = 0;
            for (uint32_t k = e.off[j]; k < e.off[j + 1]; ++k) s += child[e.idx[k]];
            parent[j] = s;
            sum += s;
        }
        est[i - 1] = sum;
        parent.swap(child);
    }
    return est;
}

void core_paths(const OGraph& q, const Tree& t, uint32_t u, std::vector<uint32_t>& cur,
                std::vector<std::vector<uint32_t>>& out) {
    cur.push_back(u);
    bool leaf = true;
    for (uint32_t c : t.children[u])
        if (q.core2[c]) { core_paths(q, t, c, cur, out); leaf = false; }
    if (leaf) out.push_back(cur);
    cur.pop_back();
}

void tree_paths(const OGraph& q, const Tree& t, uint32_t u, std::vector<uint32_t>& cur,
                std::vector<std::vector<uint32_t>>& out) {
    cur.push_back(u);
    bool leaf = true;
    for (uint32_t c : t.children[u])
        if (q.deg(c) > 1) { tree_paths(q, t, c, cur, out); leaf = false; }
    if (leaf && cur.size() > 1) out.push_back(cur);
    cur.pop_back();
}

// GenerateQueryPlan.cpp:270-471 on the unit graph with the unit's BFS tree
// (GenerateFilteringPlan.cpp:15-79).
void cfl_order(const OGraph& d, const OGraph& q, const Cands& C, const Tables& T, Plan& p) {
    OGraph ug = unit_graph(q);
    Tree t;
    uint32_t root = cfl_start_vertex(d, ug);
    bfs_tree(ug, root, t);
    uint32_t n = q.V;
    p.order.assign(n, 0);
    p.pivot.assign(n, 0);
    std::vector<uint8_t> vis(n, 0);
    std::vector<std::vector<uint32_t>> cps;
    std::vector<std::vector<std::vector<uint32_t>>> forests;
    std::vector<uint32_t> leaves, cur;
    for (uint32_t u = 0; u < n; ++u)
        if (ug.deg(u) == 1) leaves.push_back(u);
    if (ug.core2[root]) {
        core_paths(ug, t, root, cur, cps);
        for (uint32_t u = 0; u < n; ++u) {
            if (!ug.core2[u]) continue;
            std::vector<std::vector<uint32_t>> tp;
            tree_paths(ug, t, u, cur, tp);
            if (!tp.empty()) forests.push_back(tp);
        }
    } else {
        std::vector<std::vector<uint32_t>> tp;
        tree_paths(ug, t, root, cur, tp);
        if (!tp.empty()) forests.push_back(tp);
    }
    uint32_t cnt = 0;
    p.order[cnt++] = root;
    vis[root] = 1;
    auto div = [](uint64_t a, uint64_t b) { return a / (double)b; };
    if (!cps.empty()) {
        std::vector<std::vector<uint64_t>> est;
        std::vector<uint32_t> nte;
        for (auto& path : cps) {
            uint32_t c = ug.deg(path[0]) - (uint32_t)t.children[path[0]].size();
            for (size_t i = 1; i < path.size(); ++i)
                c += ug.deg(path[i]) - (uint32_t)t.children[path[i]].size() - 1;
            nte.push_back(c + 1);
            est.push_back(path_estimates(path, T));
        }
        double minv = std::numeric_limits<double>::max();
        uint32_t sel = 0;
        for (uint32_t i = 0; i < cps.size(); ++i) {
            double v = div(est[i][0], nte[i]);
            if (v < minv) { minv = v; sel = i; }
        }
        for (size_t i = 1; i < cps[sel].size(); ++i) {
            p.order[cnt] = cps[sel][i];
            p.pivot[cnt] = cps[sel][i - 1];
            ++cnt;
            vis[cps[sel][i]] = 1;
        }
        cps.erase(cps.begin() + sel);
        est.erase(est.begin() + sel);
        nte.erase(nte.begin() + sel);
        while (!cps.empty()) {
            minv = std::numeric_limits<double>::max();
            sel = 0;
            for (uint32_t i = 0; i < cps.size(); ++i) {
                uint32_t ri = 0;
                for (uint32_t j = 0; j < cps[i].size(); ++j) {
                    if (vis[cps[i][j]]) continue;
                    ri = j - 1;
                    break;
                }
                double v = div(est[i][ri], C[cps[i][ri]].size());
                if (v < minv) { minv = v; sel = i; }
            }
            for (size_t i = 1; i < cps[sel].size(); ++i) {
                if (vis[cps[sel][i]]) continue;
                p.order[cnt] = cps[sel][i];
                p.pivot[cnt] = cps[sel][i - 1];
                ++cnt;
                vis[cps[sel][i]] = 1;
            }
            cps.erase(cps.begin() + sel);
            est.erase(est.begin() + sel);
        }
    }
    for (auto& tps : forests) {
        std::vector<std::vector<uint64_t>> est;
        for (auto& path : tps) est.push_back(path_estimates(path, T));
        while (!tps.empty()) {
            double minv = std::numeric_limits<double>::max();
            uint32_t sel = 0;
            for (uint32_t i = 0; i < tps.size(); ++i) {
                uint32_t ri = 0;
                for (uint32_t j = 0; j < tps[i].size(); ++j) {
                    if (vis[tps[i][j]]) continue;
                    ri = j == 0 ? j : j - 1;
                    break;
                }
                double v = div(est[i][ri], C[tps[i][ri]].size());
                if (v < minv) { minv = v; sel = i; }
            }
            for (size_t i = 0; i < tps[sel].size(); ++i) {
                if (vis[tps[sel][i]]) continue;
                p.order[cnt] = tps[sel][i];
                p.pivot[cnt] = i ? tps[sel][i - 1] : 0;  // i == 0 never happens: path roots are placed
                ++cnt;
                vis[tps[sel][i]] = 1;
            }
            tps.erase(tps.begin() + sel);
            est.erase(est.begin() + sel);
        }
    }
    while (!leaves.empty()) {
        double minv = std::numeric_limits<double>::max();
        uint32_t sel = 0;
        for (uint32_t i = 0; i < leaves.size(); ++i) {
            double v = (double)C[leaves[i]].size();
            if (v < minv) { minv = v; sel = i; }
        }
        uint32_t u = leaves[sel];
        if (!vis[u]) {
            p.order[cnt] = u;
            p.pivot[cnt] = t.parent[u];
            ++cnt;
            vis[u] = 1;
        }
        leaves.erase(leaves.begin() + sel);
    }
}

// EvaluateQuery.cpp:1594-1666: bn[i] = earlier neighbours of order[i], in
// neighbour-id order, optionally without the pivot.
// GenerateQueryPlan.cpp:1133-1160 after GenerateFilteringPlan.cpp:145-181,
// both on the query unit (labels renumbered, so the start-vertex rule counts
// LDF candidates of the renumbered labels): the matching order is the unit's
// DPiso BFS order and each later vertex's pivot is its BFS parent.
void dpiso_order(const OGraph& d, const OGraph& q, Plan& p) {
    OGraph ug = unit_graph(q);
    Tree t;
    bfs_tree(ug, dpiso_start_vertex(d, ug), t);
    uint32_t n = q.V;
    p.order = t.order;
    p.pivot.assign(n, 0);
    for (uint32_t i = 1; i < n; ++i) p.pivot[i] = t.parent[p.order[i]];
}

void backward_neighbours(const OGraph& q, Plan& p, bool drop_pivot) {
    uint32_t n = q.V;
    p.bn.assign(n, {});
    std::vector<uint8_t> vis(n, 0);
    vis[p.order[0]] = 1;
    for (uint32_t i = 1; i < n; ++i) {
        uint32_t u = p.order[i];
        for (uint32_t j = 0; j < q.deg(u); ++j) {
            uint32_t w = q.nbrs(u)[j];
            if (vis[w] && !(drop_pivot && w == p.pivot[i])) p.bn[i].push_back(w);
        }
        vis[u] = 1;
    }
}

// ---------------------------------------------------------------------------
// Enumeration engines (at one thread, one task holding every root candidate;
// enumerate() below also runs them with several threads):
// GQL   EvaluateQuery.cpp:980-1050 + :2156-2194
// EXPLORE :1052-1126 + :2369-2426
// LFTJ  :231-372 + :2024-2082
// `calls` counts assignments exactly where the reference increments
// l_call_count; `cap` stops after the leaf that reaches it.
// ---------------------------------------------------------------------------
enum Engine { kGQL = 0, kExplore = 1, kLFTJ = 2 };

struct EnumResult {
    uint64_t count = 0, calls = 0;
    std::vector<uint64_t> depth_nodes;    // assignments per depth
    uint64_t algo_bytes = 0;              // search bytes as defined below
};

// Algorithmic bytes of the search (the measure the device kernels count):
// every search-tree node below the root reads, for each earlier query
// neighbour b of the vertex it extends, the candidate-edge table row
// T[b][u] at b's assigned candidate (4 bytes per entry), and writes its valid
// extensions (4 bytes each) unless they are leaves. Independent of the engine:
// the rows and the valid sets are fixed by the candidates, tables and order.
uint64_t node_row_entries(const Tables& T, const std::vector<uint32_t>& earlier, uint32_t u,
                          const std::vector<uint32_t>& eidx) {
    uint64_t s = 0;
    for (uint32_t b : earlier) {
        const Table& t = T[b][u];
        s += t.off[eidx[b] + 1] - t.off[eidx[b]];
    }
    return s;
}

// Shared state of one enumeration run by `threads` workers. The root
// candidates are claimed one at a time from `next` (the reference chunks them
// into 16 * threads tasks and donates halves of running tasks to idle
// threads, EvaluateQuery.cpp:58-99 / TaskPool.cpp; a root claimed per task is
// the simplest schedule with the same counts). With one worker the claims
// follow the root order, so calls and the cap stop equal the sequential
// engines' exactly; with several the count is min(cap, total) and calls are
// exact when the cap is not reached.
struct SearchShared {
    uint64_t cap = UINT64_MAX;
    uint32_t threads = 1;
    double time_limit_s = 0;   // 0: none
    std::chrono::steady_clock::time_point t0;
    std::atomic<uint64_t> next{0};
    std::atomic<uint64_t> count{0};
    std::atomic<bool> stop{false};
    std::atomic<bool> timed_out{false};
};

void enumerate_worker(const OGraph& d, const OGraph& q, const Cands& C, const Tables& T,
                      const Plan& p, Engine eng, SearchShared& sh, EnumResult& r) {
    uint32_t n = q.V;
    r = EnumResult();
    r.depth_nodes.assign(n, 0);
    if (n == 0) return;
    std::vector<std::vector<uint32_t>> cand(n);   // vertex ids (GQL) or indices
    std::vector<uint32_t> pos(n, 0), emb(n, 0), eidx(n, 0), tmp;
    std::vector<uint8_t> visited(d.V, 0);
    // every earlier neighbour of order[i] (the device's operand rows)
    std::vector<std::vector<uint32_t>> earlier(n);
    {
        std::vector<uint8_t> seen(n, 0);
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t u = p.order[i];
            for (uint32_t j = 0; j < q.deg(u); ++j)
                if (seen[q.nbrs(u)[j]]) earlier[i].push_back(q.nbrs(u)[j]);
            seen[u] = 1;
        }
    }
    uint32_t root = p.order[0];
    const uint64_t n_roots = C[root].size();

    auto gen = [&](uint32_t dep) {
        uint32_t u = p.order[dep];
        std::vector<uint32_t>& out = cand[dep];
        out.clear();
        if (eng == kGQL) {
            for (uint32_t v : C[u]) {
                if (visited[v]) continue;
                bool ok = true;
                for (uint32_t b : p.bn[dep])
                    if (!d.has_edge(v, emb[b])) { ok = false; break; }
                if (ok) out.push_back(v);
            }
        } else if (eng == kExplore) {
            uint32_t pv = p.pivot[dep];
            const Table& t = T[pv][u];
            for (uint32_t k = t.off[eidx[pv]]; k < t.off[eidx[pv] + 1]; ++k) {
                uint32_t ci = t.idx[k];
                uint32_t v = C[u][ci];
                if (visited[v]) continue;
                bool ok = true;
                for (uint32_t b : p.bn[dep])
                    if (!d.has_edge(v, emb[b])) { ok = false; break; }
                if (ok) out.push_back(ci);
            }
        } else {
            const auto& bn = p.bn[dep];
            const Table& t0 = T[bn[0]][u];
            out.assign(t0.idx.begin() + t0.off[eidx[bn[0]]], t0.idx.begin() + t0.off[eidx[bn[0]] + 1]);
            for (size_t i = 1; i < bn.size(); ++i) {
                const Table& t = T[bn[i]][u];
                tmp.clear();
                std::set_intersection(out.begin(), out.end(), t.idx.begin() + t.off[eidx[bn[i]]],
                                      t.idx.begin() + t.off[eidx[bn[i]] + 1], std::back_inserter(tmp));
                out.swap(tmp);
            }
        }
    };

    const bool single = sh.threads <= 1;
    uint64_t unflushed = 0;   // leaves not yet added to sh.count (several workers)
    uint64_t tick = 0;
    while (!sh.stop.load(std::memory_order_relaxed)) {
        // claim the next root candidate: this worker's task is its subtree
        const uint64_t k = sh.next.fetch_add(1, std::memory_order_relaxed);
        if (k >= n_roots) break;
        cand[0].assign(1, eng == kGQL ? C[root][k] : (uint32_t)k);
        int dep = 0;
        pos[0] = 0;
        while (true) {
            while (pos[dep] < cand[dep].size()) {
                uint32_t u = p.order[dep];
                uint32_t x = cand[dep][pos[dep]++];
                uint32_t v = eng == kGQL ? x : C[u][x];
                r.calls++;
                if ((++tick & 0xFFFF) == 0) {
                    if (sh.stop.load(std::memory_order_relaxed)) goto out;
                    if (sh.time_limit_s > 0 &&
                        std::chrono::duration<double>(std::chrono::steady_clock::now() - sh.t0).count() >
                            sh.time_limit_s) {
                        sh.timed_out = true;
                        sh.stop = true;
                        goto out;
                    }
                }
                if (eng == kLFTJ && visited[v]) continue;
                r.depth_nodes[dep]++;
                emb[u] = v;
                // candidate index of the assignment (the GQL engine walks
                // vertex ids; C(u) is sorted)
                eidx[u] = eng == kGQL ? (uint32_t)(std::lower_bound(C[u].begin(), C[u].end(), v) - C[u].begin()) : x;
                visited[v] = 1;
                if (dep == (int)n - 1) {
                    r.count++;
                    visited[v] = 0;
                    if (single) {
                        if (r.count >= sh.cap) { sh.stop = true; goto out; }
                    } else if (++unflushed == 256) {
                        if (sh.count.fetch_add(unflushed) + unflushed >= sh.cap) sh.stop = true;
                        unflushed = 0;
                        if (sh.stop.load(std::memory_order_relaxed)) goto out;
                    }
                } else {
                    if (dep >= 1) r.algo_bytes += 4;   // a survivor written below the root level
                    ++dep;
                    pos[dep] = 0;
                    gen(dep);
                    r.algo_bytes += 4 * node_row_entries(T, earlier[dep], p.order[dep], eidx);
                }
            }
            --dep;
            if (dep < 0) break;
            visited[emb[p.order[dep]]] = 0;
        }
    }
out:
    if (unflushed) sh.count.fetch_add(unflushed);
}

// Runs the engine with `threads` workers (0: every hardware thread) under an
// optional time limit; r.count = min(cap, total).
void enumerate(const OGraph& d, const OGraph& q, const Cands& C, const Tables& T,
               const Plan& p, Engine eng, uint64_t cap, EnumResult& r, uint32_t threads = 1,
               double time_limit_s = 0, bool* timed_out = nullptr) {
    SearchShared sh;
    sh.cap = cap;
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    sh.threads = threads;
    sh.time_limit_s = time_limit_s;
    sh.t0 = std::chrono::steady_clock::now();
    if (threads == 1) {
        enumerate_worker(d, q, C, T, p, eng, sh, r);
    } else {
        std::vector<EnumResult> part(threads);
        std::vector<std::thread> pool;
        for (uint32_t t = 0; t < threads; ++t)
            pool.emplace_back([&, t] { enumerate_worker(d, q, C, T, p, eng, sh, part[t]); });
        for (auto& th : pool) th.join();
        r = EnumResult();
        r.depth_nodes.assign(q.V, 0);
        for (auto& x : part) {
            r.calls += x.calls;
            r.algo_bytes += x.algo_bytes;
            for (uint32_t i = 0; i < x.depth_nodes.size(); ++i) r.depth_nodes[i] += x.depth_nodes[i];
        }
        r.count = std::min<uint64_t>(cap, sh.count.load());
    }
    if (timed_out) *timed_out = sh.timed_out.load();
}

// ---------------------------------------------------------------------------
// Whole plain path at one thread (StudyPerformance.cpp:1164-1781).
// ---------------------------------------------------------------------------
struct Run {
    Cands C;
    Tables T;
    Plan plan;
    EnumResult er;
    uint64_t mem_words = 0;
    bool filter_ok = true;
};

}  // namespace

extern "C" {

struct orc_result {
    uint64_t count;
    uint64_t calls;
    uint64_t mem_words;
    uint32_t n;
    uint32_t filter_ok;
    uint32_t cand_count[64];
    uint32_t order[64];
    uint32_t pivot[64];
    uint64_t depth_nodes[64];
    uint64_t algo_bytes;
};

void* orc_graph_new(uint32_t V, const uint32_t* off, const uint32_t* adj, const uint32_t* lab) {
    OGraph* g = new OGraph();
    g->V = V;
    g->off.assign(off, off + V + 1);
    g->adj.assign(adj, adj + off[V]);
    g->lab.assign(lab, lab + V);
    g->E = off[V] / 2;
    g->finish();
    return g;
}

void orc_graph_free(void* g) { delete static_cast<OGraph*>(g); }

uint32_t orc_graph_labels(void* g) { return static_cast<OGraph*>(g)->L; }

// Phase times of one run (steady_clock, seconds), the phases the reference
// logs (Args.cpp:517-535): filter (with sortCandidates), build table,
// generate query plan, enumerate.
struct orc_times {
    double filter_s, build_s, plan_s, enum_s, total_s;
    uint32_t threads;
    uint32_t timed_out;
};

// filter: "LDF" "NLF" "GQL" "CFL" "DPiso"; order: "GQL" "CFL" "DPiso";
// engine: "GQL" "EXPLORE" "LFTJ" (not EXPLORE with the DPiso order).
// threads: enumeration workers (0: every hardware thread); time_limit_s > 0
// stops the enumeration once that many seconds have passed since the start
// of filtering (the reference's -time_limit), setting times->timed_out.
// Returns a run handle (NULL on a bad argument).
void* orc_run_mt(void* dg, void* qg, const char* filter, const char* order, const char* engine,
                 uint64_t cap, uint32_t threads, double time_limit_s, orc_result* out, orc_times* times) {
    using clk = std::chrono::steady_clock;
    auto secs = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); };
    const OGraph& d = *static_cast<OGraph*>(dg);
    const OGraph& q = *static_cast<OGraph*>(qg);
    if (q.V == 0 || q.V > 64) return nullptr;
    std::string f(filter), o(order), e(engine);
    Engine eng;
    if (e == "GQL") eng = kGQL;
    else if (e == "EXPLORE") eng = kExplore;
    else if (e == "LFTJ") eng = kLFTJ;
    else return nullptr;
    if (!(o == "GQL" || o == "CFL" || (o == "DPiso" && eng != kExplore))) return nullptr;
    Run* r = new Run();
    const clk::time_point t0 = clk::now();
    if (f == "LDF") r->filter_ok = ldf_nlf_filter(d, q, false, r->C);
    else if (f == "NLF") r->filter_ok = ldf_nlf_filter(d, q, true, r->C);
    else if (f == "GQL") r->filter_ok = gql_filter(d, q, r->C);
    else if (f == "CFL") r->filter_ok = cfl_filter(d, q, r->C);
    else if (f == "DPiso") r->filter_ok = dpiso_filter(d, q, r->C);
    else { delete r; return nullptr; }
    for (auto& c : r->C) std::sort(c.begin(), c.end());   // sortCandidates
    const clk::time_point t1 = clk::now();
    build_tables(d, q, r->C, r->T);
    r->mem_words = memory_words(q, r->C, r->T);
    const clk::time_point t2 = clk::now();
    if (o == "GQL") {
        gql_order(q, r->C, d.V, r->plan);
        backward_neighbours(q, r->plan, false);
    } else if (o == "CFL") {
        cfl_order(d, q, r->C, r->T, r->plan);
        backward_neighbours(q, r->plan, true);
    } else {
        // the reference plans DPiso orders with plain backward neighbours
        // (StudyPerformance.cpp:1397-1414), which the pivot-based EXPLORE
        // engine cannot use
        dpiso_order(d, q, r->plan);
        backward_neighbours(q, r->plan, false);
    }
    const clk::time_point t3 = clk::now();
    double left = 0;
    if (time_limit_s > 0) left = std::max(1e-6, time_limit_s - secs(t0, t3));
    bool to = false;
    enumerate(d, q, r->C, r->T, r->plan, eng, cap ? cap : UINT64_MAX, r->er, threads, left, &to);
    const clk::time_point t4 = clk::now();
    if (times) {
        times->filter_s = secs(t0, t1);
        times->build_s = secs(t1, t2);
        times->plan_s = secs(t2, t3);
        times->enum_s = secs(t3, t4);
        times->total_s = secs(t0, t4);
        times->threads = threads ? threads : std::max(1u, std::thread::hardware_concurrency());
        times->timed_out = to;
    }
    if (out) {
        memset(out, 0, sizeof(*out));
        out->count = r->er.count;
        out->calls = r->er.calls;
        out->mem_words = r->mem_words;
        out->n = q.V;
        out->filter_ok = r->filter_ok;
        out->algo_bytes = r->er.algo_bytes;
        for (uint32_t u = 0; u < q.V; ++u) {
            out->cand_count[u] = (uint32_t)r->C[u].size();
            out->order[u] = r->plan.order[u];
            out->pivot[u] = r->plan.pivot[u];
            out->depth_nodes[u] = r->er.depth_nodes[u];
        }
    }
    return r;
}

// One thread, no time limit.
void* orc_run(void* dg, void* qg, const char* filter, const char* order, const char* engine,
              uint64_t cap, orc_result* out) {
    return orc_run_mt(dg, qg, filter, order, engine, cap, 1, 0.0, out, nullptr);
}

void orc_run_free(void* r) { delete static_cast<Run*>(r); }

// Copies C(u) into out (which holds at least cand_count[u] entries).
void orc_run_cands(void* rr, uint32_t u, uint32_t* out) {
    Run* r = static_cast<Run*>(rr);
    std::copy(r->C[u].begin(), r->C[u].end(), out);
}

// Table T[a][b]: returns its entry count, or -1 when a-b is not a query edge.
// off receives |C(a)|+1 entries, idx receives the returned count.
int64_t orc_run_table(void* rr, uint32_t a, uint32_t b, uint32_t* off, uint32_t* idx) {
    Run* r = static_cast<Run*>(rr);
    const Table& t = r->T[a][b];
    if (!t.present) return -1;
    if (off) std::copy(t.off.begin(), t.off.end(), off);
    if (idx) std::copy(t.idx.begin(), t.idx.end(), idx);
    return (int64_t)t.idx.size();
}

// NLF candidate count of one query vertex (used to pin the start-vertex rule).
uint32_t orc_nlf_count(void* dg, void* qg, uint32_t u) {
    std::vector<uint32_t> tmp;
    nlf_candidates(*static_cast<OGraph*>(dg), *static_cast<OGraph*>(qg), u, true, tmp);
    return (uint32_t)tmp.size();
}

}  // extern "C"
