// psm_dropin.cpp — implementation of the reference's C++ call surfaces over
// the C ABI (see psm_dropin.h). Host-side bookkeeping only; every candidate
// set, table and enumeration is computed by libpsm.so on the GPU.

#include "psm_dropin.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <unordered_map>

namespace {

std::mutex g_mu;
psm_ctx* g_ctx = nullptr;
std::string g_log_path;
std::mutex g_log_mu;

// One device query per query Graph (GLOBALARGS/UnitArgs device state).
std::unordered_map<const Graph*, psm_query*> g_sessions;
// the last filter per query stopped at the time limit
std::unordered_map<const Graph*, bool> g_filter_overtime;
const Graph* g_last_filtered = nullptr;

void check(int r, const char* what) {
    if (r != PSM_OK && r != PSM_EEMPTY && r != PSM_ETIMEOUT) {
        std::string msg = std::string(what) + " failed with status " + std::to_string(r) + ": " +
                          psm_last_error(psm_dropin::context());
        throw std::runtime_error(msg);
    }
}

psm_query* session(const Graph* data, const Graph* query, bool fresh) {
    // context() and device() take g_mu themselves: resolve them first
    psm_ctx* ctx = psm_dropin::context();
    psm_graph* dg = data->device();
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_sessions.find(query);
    if (it != g_sessions.end() && !fresh) return it->second;
    if (it != g_sessions.end()) {
        psm_query_destroy(it->second);
        g_sessions.erase(it);
    }
    psm_query* q = nullptr;
    const auto& o = query->offsets();
    const auto& a = query->neighbors();
    const auto& l = query->labels();
    ui dummy = 0;
    const int r = psm_query_create(ctx, dg, query->getVerticesCount(), o.data(),
                                   a.empty() ? &dummy : a.data(), l.data(), &q);
    if (r != PSM_OK)
        throw std::runtime_error("psm_query_create failed with status " + std::to_string(r) + ": " +
                                 psm_last_error(ctx));
    g_sessions[query] = q;
    return q;
}

psm_query* existing(const Graph* query) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_sessions.find(query);
    if (it == g_sessions.end()) throw std::runtime_error("query graph has not been filtered");
    return it->second;
}

// The reference allocates n rows of max-label-frequency entries
// (FilterVertices.cpp:689-702); the caller owns them. The time limit counts
// from here, as the reference sets its deadline before filtering.
bool run_filter(const Graph* data, const Graph* query, int method, ui**& candidates, ui*& counts) {
    psm_query* q = session(data, query, true);
    psm_set_time_limit(psm_dropin::context(), (double)FilterVertices::time_limit);
    const int r = psm_filter(q, method);
    check(r, "psm_filter");
    {
        std::lock_guard<std::mutex> lk(g_mu);
        g_filter_overtime[query] = r == PSM_ETIMEOUT;
        g_last_filtered = query;
    }
    const ui n = query->getVerticesCount();
    counts = new ui[n];
    check(psm_query_candidate_counts(q, counts), "psm_query_candidate_counts");
    candidates = new ui*[n];
    const ui row = std::max<ui>(1, data->getGraphMaxLabelFrequency());
    for (ui u = 0; u < n; ++u) {
        candidates[u] = new ui[row];
        check(psm_query_candidates(q, u, candidates[u]), "psm_query_candidates");
    }
    return r == PSM_OK;
}

// Fills tree nodes from a BFS order / parent / level (graphoperations.cpp:
// 156-195) and each node's bn / fn / under-level neighbours by the CFL
// filtering plan's rule (GenerateFilteringPlan.cpp:81-140): a neighbour on
// the same level is bn when it comes earlier in the BFS order, fn otherwise;
// one on an upper level is bn, one on a lower level under-level.
void fill_tree(const Graph* g, const std::vector<ui>& order, const std::vector<ui>& parent,
               const std::vector<ui>& level, TreeNode* tree, ui* tree_order) {
    const ui n = g->getVerticesCount();
    std::vector<ui> index(n);
    for (ui i = 0; i < n; ++i) index[order[i]] = i;
    for (ui u = 0; u < n; ++u) {
        tree[u].initialize(n);
        tree[u].id_ = u;
        tree[u].parent_ = parent[u];
        tree[u].level_ = level[u];
        tree[u].children_count_ = tree[u].bn_count_ = tree[u].fn_count_ = tree[u].under_level_count_ = 0;
    }
    for (ui i = 0; i < n; ++i) {
        const ui u = order[i];
        if (tree_order) tree_order[i] = u;
        if (i > 0) tree[parent[u]].children_[tree[parent[u]].children_count_++] = u;
        ui cnt;
        const ui* nb = g->getVertexNeighbors(u, cnt);
        for (ui j = 0; j < cnt; ++j) {
            const ui w = nb[j];
            if (level[w] == level[u]) {
                if (index[w] < index[u]) tree[u].bn_[tree[u].bn_count_++] = w;
                else tree[u].fn_[tree[u].fn_count_++] = w;
            } else if (level[u] > level[w]) {
                tree[u].bn_[tree[u].bn_count_++] = w;
            } else {
                tree[u].under_level_[tree[u].under_level_count_++] = w;
            }
        }
    }
}

// BFS tree the library used for `query`'s last CFL / DPiso step.
void device_tree(const Graph* query, int which, TreeNode* tree, ui* tree_order) {
    const ui n = query->getVerticesCount();
    std::vector<ui> order(n), parent(n), level(n);
    check(psm_query_bfs_tree(existing(query), which, order.data(), parent.data(), level.data()),
          "psm_query_bfs_tree");
    fill_tree(query, order, parent, level, tree, tree_order);
}

bool filter_with_tree(const Graph* data, const Graph* query, int method, ui**& candidates, ui*& counts,
                      ui*& order, TreeNode*& tree) {
    const bool ok = run_filter(data, query, method, candidates, counts);
    const ui n = query->getVerticesCount();
    if (!order) order = new ui[n];
    if (!tree) tree = new TreeNode[n];
    device_tree(query, PSM_TREE_FILTER, tree, order);
    return ok;
}

// The query whose device session a unit's plan and enumeration use.
const Graph* unit_query(const UnitArgs* u) {
    if (!u || !u->g_args) throw std::runtime_error("map the unit's candidates (mapUidCandidates) first");
    return u->g_args->query_graph;
}

// A plan generated on the query (original ids) moved to the unit's ids.
void plan_to_unit(UnitArgs* u, int method, bool tree_too) {
    const Graph* query = unit_query(u);
    psm_query* q = existing(query);
    const ui nq = query->getVerticesCount();
    const ui n = u->vertices_count;
    std::vector<ui> order(nq), pivot(nq);
    check(psm_generate_plan(q, method, order.data(), pivot.data()), "psm_generate_plan");
    if (n != nq) throw std::runtime_error("a unit smaller than its query has no device plan");
    const ui* o2n = u->query_graph->old2news;
    for (ui i = 0; i < n; ++i) {
        u->order[i] = o2n[order[i]];
        u->pivots[i] = o2n[pivot[i]];
    }
    if (tree_too) {
        if (!u->tree) u->tree = new TreeNode[n];
        std::vector<ui> to(nq), par(nq), lev(nq), uo(n), upar(n), ulev(n);
        check(psm_query_bfs_tree(q, PSM_TREE_PLAN, to.data(), par.data(), lev.data()), "psm_query_bfs_tree");
        for (ui i = 0; i < n; ++i) {
            uo[i] = o2n[to[i]];
            upar[o2n[i]] = o2n[par[i]];
            ulev[o2n[i]] = lev[i];
        }
        fill_tree(u->query_graph->subgraph, uo, upar, ulev, u->tree, u->tree_order);
    }
}

// One unit enumerated on the device in the unit's order: its counts and the
// engine's Call Count go to the global arguments.
void run_unit(TaskSlot& task, bool lftj) {
    GLOBALARGS* g = TaskSlot::g_args;
    UnitArgs* u = task.u_args;
    if (!g || !u) throw std::runtime_error("TaskSlot::g_args and the task's unit must be set");
    const Graph* query = unit_query(u);
    psm_query* q = existing(query);
    const ui n = u->vertices_count;
    const ui* n2o = u->query_graph->new2olds;
    std::vector<ui> order(n), pivot(n, 0);
    for (ui i = 0; i < n; ++i) {
        order[i] = n2o[u->order[i]];
        if (u->pivots && i > 0) pivot[i] = n2o[u->pivots[i]];
    }
    // a pivot that is not an earlier neighbour (plans without pivots) is dropped
    const int rp = psm_set_plan(q, order.data(), pivot.data());
    if (rp == PSM_EINVAL) check(psm_set_plan(q, order.data(), nullptr), "psm_set_plan");
    else check(rp, "psm_set_plan");
    const uint64_t cap = g->output_limit_num <= 0 || g->output_limit_num >= INT64_MAX ? 0
                                                                                      : (uint64_t)g->output_limit_num;
    psm_enum_stats st;
    const auto t0 = std::chrono::steady_clock::now();
    const int r = psm_evaluate(q, cap, &st);
    const int64_t ns =
        std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count();
    check(r, "psm_evaluate");
    const ul calls = (ul)(st.calls + (lftj ? st.visited_rejects : 0));
    g->embedding_count += (int64_t)st.count;
    g->call_count += calls;
    if (r == PSM_ETIMEOUT) g->overtime = true;
    g->total_task_num += (int)st.tasks;
    const auto tid = std::this_thread::get_id();
    g->addTimeEnum(tid, ns);
    g->addTimeExec(tid, (int64_t)(st.kernel_ms * 1e6));
    g->addTimeSplit(tid, (int64_t)(st.expand_ms * 1e6));
    g->addCallCount(tid, (int64_t)calls);
    g->addEmbeddingCount(tid, (int64_t)st.count);
}

}  // namespace

namespace psm_dropin {
psm_ctx* context() {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ctx) {
        const char* d = getenv("PSM_DEVICE");
        int r = psm_ctx_create(d ? atoi(d) : 0, &g_ctx);
        if (r != PSM_OK) throw std::runtime_error("psm_ctx_create failed: " + std::to_string(r));
    }
    return g_ctx;
}

void set_log_file(const std::string& path) {
    std::lock_guard<std::mutex> lk(g_log_mu);
    g_log_path = path;
}

LogLine::LogLine() { oss_ << "[Thread 1] "; }

LogLine::~LogLine() {
    std::lock_guard<std::mutex> lk(g_log_mu);
    if (g_log_path.empty()) {
        std::cout << oss_.str();
    } else {
        std::ofstream f(g_log_path, std::ios::app);
        f << oss_.str();
    }
}
}  // namespace psm_dropin

// ---------------------------------------------------------------------------
// Graph
// ---------------------------------------------------------------------------
Graph::Graph(bool) {}

Graph::~Graph() {
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_sessions.find(this);
        if (it != g_sessions.end()) {
            psm_query_destroy(it->second);
            g_sessions.erase(it);
        }
    }
    if (dev_) psm_graph_destroy(dev_);
}

// Text format of graph.cpp:92-172: "t V E", "v id label degree", "e u v".
void Graph::loadGraphFromFile(const std::string& path) {
    // parsed by libpsm's multi-threaded reader (every hardware thread)
    psm_text_graph* tg = nullptr;
    const int r = psm_text_graph_read(path.c_str(), 0, &tg);
    if (r == PSM_ENOTFOUND) {
        std::cout << "Can not open the graph file " << path << " ." << std::endl;
        exit(-1);
    }
    if (r != PSM_OK) {
        std::cerr << "Bad graph file " << path << ": " << psm_text_graph_last_error() << std::endl;
        exit(-1);
    }
    ui V = 0;
    uint64_t E = 0;
    psm_text_graph_info(tg, &V, &E);
    loadFromCSR(V, psm_text_graph_offsets(tg), psm_text_graph_neighbors(tg), psm_text_graph_labels(tg));
    psm_text_graph_free(tg);
}

// Binary format of graph.cpp:221-403: [i32 4][u32 V][u32 E][u32 deg[V]],
// neighbours vertex by vertex, u32 labels. A file whose degrees do not sum to
// 2E, or that ends early, is rejected.
void Graph::loadGraphFromFileCompressed(const std::string& dp, const std::string& ep,
                                        const std::string& lp) {
    std::ifstream df(dp, std::ios::binary), ef(ep, std::ios::binary), lf(lp, std::ios::binary);
    if (!df || !ef || !lf) throw std::runtime_error("cannot open the binary graph files");
    auto read = [](std::ifstream& f, void* dst, size_t bytes, const char* what) {
        f.read(static_cast<char*>(dst), (std::streamsize)bytes);
        if (!f || (size_t)f.gcount() != bytes)
            throw std::runtime_error(std::string("binary graph file ends early: ") + what);
    };
    int32_t isz = 0;
    ui V = 0, E = 0;
    read(df, &isz, 4, "header");
    read(df, &V, 4, "header");
    read(df, &E, 4, "header");
    if (isz != 4) throw std::runtime_error("binary graph: element size is not 4");
    std::vector<ui> deg(V), off(V + 1, 0), lab(V);
    read(df, deg.data(), 4ull * V, "degrees");
    uint64_t sum = 0;
    for (ui v = 0; v < V; ++v) {
        sum += deg[v];
        if (sum > 2ull * E) break;
        off[v + 1] = (ui)sum;
    }
    if (sum != 2ull * E)
        throw std::runtime_error("binary graph: degrees sum to " + std::to_string(sum) + ", not 2E = " +
                                 std::to_string(2ull * E));
    std::vector<ui> adj((size_t)E * 2);
    read(ef, adj.data(), 4ull * adj.size(), "neighbours");
    read(lf, lab.data(), 4ull * V, "labels");
    for (ui x : adj)
        if (x >= V) throw std::runtime_error("binary graph: neighbour id out of range");
    loadFromCSR(V, off.data(), adj.data(), lab.data());
}

void Graph::storeComparessedGraph(const std::string& dp, const std::string& ep, const std::string& lp) {
    std::ofstream df(dp, std::ios::binary), ef(ep, std::ios::binary), lf(lp, std::ios::binary);
    int32_t isz = 4;
    df.write(reinterpret_cast<const char*>(&isz), 4);
    df.write(reinterpret_cast<const char*>(&vertices_count_), 4);
    df.write(reinterpret_cast<const char*>(&edges_count_), 4);
    for (ui v = 0; v < vertices_count_; ++v) {
        ui d = getVertexDegree(v);
        df.write(reinterpret_cast<const char*>(&d), 4);
    }
    ef.write(reinterpret_cast<const char*>(neighbors_.data()), 4ull * neighbors_.size());
    lf.write(reinterpret_cast<const char*>(labels_.data()), 4ull * labels_.size());
}

void Graph::loadFromCSR(ui V, const ui* off, const ui* adj, const ui* lab) {
    vertices_count_ = V;
    offsets_.assign(off, off + V + 1);
    neighbors_.assign(adj, adj + off[V]);
    labels_.assign(lab, lab + V);
    edges_count_ = off[V] / 2;
    for (ui v = 0; v < V; ++v) std::sort(neighbors_.begin() + off[v], neighbors_.begin() + off[v + 1]);
    finish();
    if (dev_) {
        psm_graph_destroy(dev_);
        dev_ = nullptr;
    }
}

void Graph::finish() {
    max_degree_ = 0;
    ui maxl = 0;
    for (ui v = 0; v < vertices_count_; ++v) {
        max_degree_ = std::max(max_degree_, getVertexDegree(v));
        maxl = std::max(maxl, labels_[v]);
    }
    labels_count_ = vertices_count_ ? maxl + 1 : 0;
    ri_off_.assign(labels_count_ + 1, 0);
    for (ui v = 0; v < vertices_count_; ++v) ri_off_[labels_[v] + 1]++;
    max_label_frequency_ = 0;
    for (ui l = 0; l < labels_count_; ++l) {
        max_label_frequency_ = std::max(max_label_frequency_, ri_off_[l + 1]);
        ri_off_[l + 1] += ri_off_[l];
    }
    ri_.assign(vertices_count_, 0);
    std::vector<ui> fill(ri_off_.begin(), ri_off_.end() - 1);
    for (ui v = 0; v < vertices_count_; ++v) ri_[fill[labels_[v]]++] = v;
}

// Core values via degree peeling (graphoperations.cpp:5-76 computes the
// full core decomposition; only "core > 1" is consulted by the matcher).
void Graph::buildCoreTable() {
    const ui V = vertices_count_;
    std::vector<ui> d(V), order(V), pos(V), bin(max_degree_ + 2, 0);
    for (ui v = 0; v < V; ++v) { d[v] = getVertexDegree(v); bin[d[v]]++; }
    ui start = 0;
    for (ui k = 0; k <= max_degree_; ++k) { ui c = bin[k]; bin[k] = start; start += c; }
    for (ui v = 0; v < V; ++v) { pos[v] = bin[d[v]]++; order[pos[v]] = v; }
    for (ui k = max_degree_; k > 0; --k) bin[k] = bin[k - 1];
    bin[0] = 0;
    for (ui i = 0; i < V; ++i) {
        const ui v = order[i];
        for (ui p = offsets_[v]; p < offsets_[v + 1]; ++p) {
            const ui u = neighbors_[p];
            if (d[u] > d[v]) {
                const ui du = d[u], pu = pos[u], pw = bin[du], w = order[pw];
                if (u != w) { pos[u] = pw; order[pu] = w; pos[w] = pu; order[pw] = u; }
                bin[du]++;
                d[u]--;
            }
        }
    }
    core_ = d;
    core_length_ = 0;
    for (ui v = 0; v < V; ++v) core_length_ += core_[v] > 1;
}

ui Graph::getLabelsFrequency(LabelID label) const {
    return label < labels_count_ ? ri_off_[label + 1] - ri_off_[label] : 0;
}

const ui* Graph::getVerticesByLabel(LabelID id, ui& count) const {
    if (id >= labels_count_) { count = 0; return ri_.data(); }
    count = ri_off_[id + 1] - ri_off_[id];
    return ri_.data() + ri_off_[id];
}

bool Graph::checkEdgeExistence(VertexID u, VertexID v) const {
    if (getVertexDegree(u) < getVertexDegree(v)) std::swap(u, v);
    return std::binary_search(neighbors_.begin() + offsets_[v], neighbors_.begin() + offsets_[v + 1], u);
}

psm_graph* Graph::device() const {
    if (!dev_) {
        ui dummy = 0;
        int r = psm_graph_upload(psm_dropin::context(), vertices_count_, offsets_.data(),
                                 neighbors_.empty() ? &dummy : neighbors_.data(), labels_.data(), &dev_);
        check(r, "psm_graph_upload");
    }
    return dev_;
}

// ---------------------------------------------------------------------------
// FilterVertices / BuildTable
// ---------------------------------------------------------------------------
int64_t FilterVertices::time_limit = 0;

bool FilterVertices::LDFFilter(const Graph* d, const Graph* q, ui**& c, ui*& n) {
    return run_filter(d, q, PSM_FILTER_LDF, c, n);
}
bool FilterVertices::NLFFilter(const Graph* d, const Graph* q, ui**& c, ui*& n) {
    return run_filter(d, q, PSM_FILTER_NLF, c, n);
}
bool FilterVertices::GQLFilter(const Graph* d, const Graph* q, ui**& c, ui*& n) {
    return run_filter(d, q, PSM_FILTER_GQL, c, n);
}
bool FilterVertices::CFLFilter(const Graph* d, const Graph* q, ui**& c, ui*& n, ui*& order, TreeNode*& tree) {
    return filter_with_tree(d, q, PSM_FILTER_CFL, c, n, order, tree);
}
bool FilterVertices::DPisoFilter(const Graph* d, const Graph* q, ui**& c, ui*& n, ui*& order, TreeNode*& tree) {
    return filter_with_tree(d, q, PSM_FILTER_DPISO, c, n, order, tree);
}

void FilterVertices::sortCandidates(ui** candidates, ui* counts, ui num) {
    for (ui i = 0; i < num; ++i) std::sort(candidates[i], candidates[i] + counts[i]);
}

bool FilterVertices::checkOverTime() {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_filter_overtime.find(g_last_filtered);
    return it != g_filter_overtime.end() && it->second;
}

// edge_matrix: caller-allocated n x n array of Edges* (StudyPerformance.cpp:
// 1251-1254); entries of query edges receive new Edges, the rest nullptr.
void BuildTable::buildTables(const Graph* data, const Graph* query, ui**, ui* counts, Edges*** edge_matrix) {
    (void)data;
    psm_query* q = existing(query);
    check(psm_build_tables(q), "psm_build_tables");
    const ui n = query->getVerticesCount();
    for (ui a = 0; a < n; ++a)
        for (ui b = 0; b < n; ++b) edge_matrix[a][b] = nullptr;
    for (ui a = 0; a < n; ++a) {
        ui cnt;
        const ui* nb = query->getVertexNeighbors(a, cnt);
        for (ui i = 0; i < cnt; ++i) {
            const ui b = nb[i];
            uint64_t nnz = 0;
            check(psm_query_table(q, a, b, nullptr, nullptr, &nnz), "psm_query_table");
            Edges* e = new Edges();
            e->vertex_count_ = counts[a];
            e->edge_count_ = (ui)nnz;
            e->offset_ = new ui[counts[a] + 1];
            e->edge_ = new ui[std::max<uint64_t>(nnz, 1)];
            check(psm_query_table(q, a, b, e->offset_, e->edge_, &nnz), "psm_query_table");
            for (ui j = 0; j < counts[a]; ++j)
                e->max_degree_ = std::max(e->max_degree_, e->offset_[j + 1] - e->offset_[j]);
            edge_matrix[a][b] = e;
        }
    }
}

size_t BuildTable::computeMemoryCostInBytes(const Graph* query, ui* counts, Edges*** edge_matrix) {
    size_t bytes = 0;
    const ui n = query->getVerticesCount();
    for (ui u = 0; u < n; ++u) bytes += counts[u] * sizeof(ui);
    for (ui a = 0; a < n; ++a)
        for (ui b = a + 1; b < n; ++b)
            if (query->checkEdgeExistence(a, b)) {
                bytes += (size_t)(edge_matrix[a][b]->edge_count_ + edge_matrix[a][b]->vertex_count_) * sizeof(ui);
                bytes += (size_t)(edge_matrix[b][a]->edge_count_ + edge_matrix[b][a]->vertex_count_) * sizeof(ui);
            }
    return bytes;
}

// ---------------------------------------------------------------------------
// plans
// ---------------------------------------------------------------------------
// The unit's BFS tree and the matching order come out of one device plan
// (the CFL order is built on the unit's tree, GenerateQueryPlan.cpp:270-471),
// so the filtering-plan call already places the order; the query-plan call
// generates it again, which gives the same result.
void GenerateFilteringPlan::generateCFLFilterPlan(const Graph*, UnitArgs* u_args) {
    plan_to_unit(u_args, PSM_ORDER_CFL, true);
}
void GenerateFilteringPlan::generateDPisoFilterPlan(const Graph*, UnitArgs* u_args) {
    plan_to_unit(u_args, PSM_ORDER_DPISO, true);
}
void GenerateQueryPlan::generateGQLQueryPlan(const Graph*, const Graph*, UnitArgs* u_args) {
    plan_to_unit(u_args, PSM_ORDER_GQL, false);
}
void GenerateQueryPlan::generateCFLQueryPlan(const Graph*, UnitArgs* u_args) {
    plan_to_unit(u_args, PSM_ORDER_CFL, false);
}
void GenerateQueryPlan::generateDPisoQueryPlan(const Graph*, UnitArgs* u_args) {
    plan_to_unit(u_args, PSM_ORDER_DPISO, false);
}

// ---------------------------------------------------------------------------
// EvaluateQuery
// ---------------------------------------------------------------------------
std::function<void(TaskSlot&, int)> EvaluateQuery::enum_method;
GLOBALARGS* TaskSlot::g_args = nullptr;

size_t EvaluateQuery::ParallelExecute(TaskPool*, SplitGraph* splitGraph, UnitArgs** unitArgsVec) {
    GLOBALARGS* g = TaskSlot::g_args;
    if (!g) throw std::runtime_error("TaskSlot::g_args is not set");
    if (!enum_method) throw std::runtime_error("EvaluateQuery::enum_method is not set");
    // SPLITMODE::Q: the device decomposition runs in JoinCollect
    if (g->splitMode == SPLITMODE::Q) return 0;
    const int units = splitGraph->getSubgraphCount();
    for (int i = 0; i < units; ++i) {
        TaskSlot task;
        task.id_unit = i;
        task.u_args = unitArgsVec[i];
        enum_method(task, i);
    }
    if (g->output_limit_num > 0 && g->embedding_count > g->output_limit_num)
        g->embedding_count = g->output_limit_num;
    return (size_t)g->embedding_count.load();
}

void EvaluateQuery::LFTJ(TaskPool*, TaskSlot& task, int) { run_unit(task, true); }
void EvaluateQuery::exploreGraphQLStyle(TaskPool*, TaskSlot& task, int) { run_unit(task, false); }
void EvaluateQuery::exploreGraph(TaskPool*, TaskSlot& task, int) { run_unit(task, false); }

void EvaluateQuery::generateBN(const Graph* query_graph, ui* order, ui* pivot, ui** bn, ui* bn_count) {
    const ui n = query_graph->getVerticesCount();
    std::fill(bn_count, bn_count + n, 0);
    std::vector<bool> seen(n, false);
    seen[order[0]] = true;
    for (ui i = 1; i < n; ++i) {
        const ui u = order[i];
        ui cnt;
        const ui* nb = query_graph->getVertexNeighbors(u, cnt);
        for (ui j = 0; j < cnt; ++j)
            if (seen[nb[j]] && (!pivot || nb[j] != pivot[i])) bn[i][bn_count[i]++] = nb[j];
        seen[u] = true;
    }
}

void EvaluateQuery::generateBN(const Graph* query_graph, ui* order, ui** bn, ui* bn_count) {
    generateBN(query_graph, order, nullptr, bn, bn_count);
}

void EvaluateQuery::generateBN(UnitArgs* u) {
    generateBN(u->query_graph->subgraph, u->order, nullptr, u->bn, u->bn_count);
}

void EvaluateQuery::generateBNWithPivot(UnitArgs* u) {
    generateBN(u->query_graph->subgraph, u->order, u->pivots, u->bn, u->bn_count);
}

// ---------------------------------------------------------------------------
// units, arguments, scheduler
// ---------------------------------------------------------------------------
InducedSubgraph::InducedSubgraph(ui vertices_count, ui label_count) : total_vertices_count(vertices_count) {
    path = new ui[vertices_count];
    subgraph = new Graph(true);
    old2news = new ui[vertices_count];
    new2olds = new ui[vertices_count];
    l_old2news = new ui[std::max<ui>(1, label_count)];
    l_new2olds = new ui[std::max<ui>(1, label_count)];
}

InducedSubgraph::~InducedSubgraph() {
    delete[] path;
    delete subgraph;
    delete[] old2news;
    delete[] new2olds;
    delete[] l_old2news;
    delete[] l_new2olds;
}

SplitGraph::SplitGraph(Graph* d, Graph* g, JoinParadigm jp, SplitPattern sp, SPLITMODE sm, int64_t tl)
    : d_graph(d), graph(g), join_paradigm(jp), split_pattern(sp), split_mode(sm), time_limit(tl) {
    // the whole query as one unit: ids kept, labels renumbered by first
    // appearance (graph.cpp:585-734)
    const ui n = g->getVerticesCount();
    InducedSubgraph* u = new InducedSubgraph(n, g->getLabelsCount());
    std::vector<ui> lab(n);
    const ui L = g->getLabelsCount();
    std::fill(u->l_old2news, u->l_old2news + std::max<ui>(1, L), (ui)-1);
    ui next = 0;
    for (ui v = 0; v < n; ++v) {
        u->path[v] = v;
        u->old2news[v] = u->new2olds[v] = v;
        const ui l = g->getVertexLabel(v);
        if (u->l_old2news[l] == (ui)-1) {
            u->l_old2news[l] = next;
            u->l_new2olds[next] = l;
            ++next;
        }
        lab[v] = u->l_old2news[l];
    }
    u->subgraph->loadFromCSR(n, g->offsets().data(), g->neighbors().data(), lab.data());
    u->subgraph->buildCoreTable();
    subgraphs_.push_back(u);
    if (sm == SPLITMODE::Q) jseq = new JoinSequence(sp);
}

SplitGraph::~SplitGraph() {
    for (InducedSubgraph* u : subgraphs_) delete u;
    delete jseq;
}

PartialMatch::PartialMatch(ui len, ui total, ui cap) : length(len), init_capacity(cap), total_vcnt(total) {
    data.reserve((size_t)cap * len);
    new2olds = new ui[std::max<ui>(1, total)];
    old2news = new ui[std::max<ui>(1, total)];
}

PartialMatch::~PartialMatch() {
    delete[] new2olds;
    delete[] old2news;
}

void PartialMatch::add(const ui* match) {
    data.insert(data.end(), match, match + length);
    ++size;
}

UnitArgs::UnitArgs(InducedSubgraph* q, size_t limit)
    : query_graph(q), output_limit((ui)std::min<size_t>(limit, UINT_MAX)), vertices_count(q->getVerticesCount()) {
    const ui n = vertices_count;
    tree_order = new ui[n];
    order = new ui[n];
    pivots = new ui[n];
    bn = new ui*[n];
    for (ui i = 0; i < n; ++i) bn[i] = new ui[n];
    bn_count = new ui[n];
    std::fill(order, order + n, 0);
    std::fill(pivots, pivots + n, 0);
    std::fill(bn_count, bn_count + n, 0);
    unit_partial_match = new PartialMatch(n, q->total_vertices_count, 0);
    std::copy(q->new2olds, q->new2olds + q->total_vertices_count, unit_partial_match->new2olds);
    std::copy(q->old2news, q->old2news + q->total_vertices_count, unit_partial_match->old2news);
}

UnitArgs::~UnitArgs() {
    delete[] tree_order;
    delete[] order;
    delete[] pivots;
    if (bn)
        for (ui i = 0; i < vertices_count; ++i) delete[] bn[i];
    delete[] bn;
    delete[] bn_count;
    delete[] tree;
    // candidate rows and tables belong to GLOBALARGS (Args.cpp:104-160)
    delete[] candidates;
    delete[] candidates_count;
    if (edge_matrix)
        for (ui i = 0; i < vertices_count; ++i) delete[] edge_matrix[i];
    delete[] edge_matrix;
    delete unit_partial_match;
}

void UnitArgs::mapUidCandidates(GLOBALARGS* g) {
    g_args = g;
    needCandidates = true;
    candidates = new ui*[vertices_count];
    candidates_count = new ui[vertices_count];
    for (ui i = 0; i < vertices_count; ++i) {
        const ui old = query_graph->new2olds[i];
        candidates[i] = g->candidates[old];
        candidates_count[i] = g->candidates_count[old];
    }
}

void UnitArgs::mapUidEdgeMatrix(GLOBALARGS* g) {
    g_args = g;
    needEdgeMatrix = true;
    edge_matrix = new Edges**[vertices_count];
    for (ui i = 0; i < vertices_count; ++i) {
        edge_matrix[i] = new Edges*[vertices_count];
        for (ui j = 0; j < vertices_count; ++j)
            edge_matrix[i][j] = g->edge_matrix[query_graph->new2olds[i]][query_graph->new2olds[j]];
    }
}

void UnitArgs::addPartialMatch(ui* embedding) {
    std::lock_guard<std::mutex> lk(pm_mtx_);
    unit_partial_match->add(embedding);
}

GLOBALARGS::GLOBALARGS(const Graph* d, const Graph* q, Edges*** em, ui** c, ui* cc, size_t limit, int64_t tl)
    : data_graph(d), query_graph(q), edge_matrix(em), candidates(c), candidates_count(cc),
      output_limit_num((int64_t)std::min<size_t>(limit, (size_t)INT64_MAX)), time_limit(tl) {}

GLOBALARGS::~GLOBALARGS() {
    const ui n = query_graph->getVerticesCount();
    if (edge_matrix) {
        for (ui i = 0; i < n; ++i) {
            for (ui j = 0; j < n; ++j) delete edge_matrix[i][j];
            delete[] edge_matrix[i];
        }
        delete[] edge_matrix;
    }
    if (candidates)
        for (ui i = 0; i < n; ++i) delete[] candidates[i];
    delete[] candidates;
    delete[] candidates_count;
}

void GLOBALARGS::addTimeExec(std::thread::id t, int64_t v) {
    std::lock_guard<std::mutex> lk(global_mutex);
    thread_times_exec[t] += v;
}
void GLOBALARGS::addTimeEnum(std::thread::id t, int64_t v) {
    std::lock_guard<std::mutex> lk(global_mutex);
    thread_times_enum[t] += v;
}
void GLOBALARGS::addTimeSplit(std::thread::id t, int64_t v) {
    std::lock_guard<std::mutex> lk(global_mutex);
    thread_times_split[t] += v;
}
void GLOBALARGS::addCallCount(std::thread::id t, int64_t v) {
    std::lock_guard<std::mutex> lk(global_mutex);
    thread_call_count[t] += v;
}
void GLOBALARGS::addEmbeddingCount(std::thread::id t, int64_t v) {
    std::lock_guard<std::mutex> lk(global_mutex);
    thread_embedding_count[t] += v;
}
void GLOBALARGS::clearStatistics() {
    thread_times_exec.clear();
    thread_times_split.clear();
    thread_call_count.clear();
    thread_embedding_count.clear();
}

void GLOBALARGS::log4Metrics() {
    // threads in a fixed order (first seen in the enumeration times)
    std::vector<std::thread::id> tids;
    for (const auto& kv : thread_times_enum) tids.push_back(kv.first);
    auto line = [&](const std::string& title, const std::function<int64_t(std::thread::id)>& v, bool secs) {
        std::ostringstream oss;
        oss << title;
        for (size_t i = 0; i < tids.size(); ++i) {
            oss << (i ? ", " : " ");
            const int64_t x = v(tids[i]);
            if (secs) oss << NANOSECTOSEC(x);
            else oss << x;
        }
        LOG() << oss.str() << std::endl;
    };
    auto get = [](std::unordered_map<std::thread::id, int64_t>& m) {
        return [&m](std::thread::id t) { auto it = m.find(t); return it == m.end() ? (int64_t)0 : it->second; };
    };
    line("Thread Enum Time:", get(thread_times_enum), true);
    line("Thread Execution Time:", get(thread_times_exec), true);
    line("Thread Split Time:", get(thread_times_split), true);
    line("Thread Idle Time:", [&](std::thread::id t) { return get(thread_times_enum)(t) - get(thread_times_exec)(t); },
         true);
    line("Thread Call Count:", get(thread_call_count), false);
    line("Thread Embedding Count:", get(thread_embedding_count), false);
    LOG() << "TaskNum: " << total_task_num << std::endl;
}

void GLOBALARGS::logGlobalStatistics() {
    LOG() << "Load graphs time (seconds): " << NANOSECTOSEC(load_graphs_time_in_ns) << std::endl;
    LOG() << "Build table time (seconds): " << NANOSECTOSEC(build_table_time_in_ns) << std::endl;
    LOG() << "Filter vertices time (seconds): " << NANOSECTOSEC(filter_vertices_time_in_ns) << std::endl;
    LOG() << "Split Qeury Pattern Time: " << NANOSECTOSEC(split_query_time_in_ns) << std::endl;
    LOG() << "Generate query plan time (seconds): " << NANOSECTOSEC(generate_query_plan_time_in_ns) << std::endl;
    LOG() << "Enumerate time (seconds): " << NANOSECTOSEC(enumeration_time_in_ns) << std::endl;
    LOG() << "Preprocessing time (seconds): " << NANOSECTOSEC(preprocessing_time_in_ns) << std::endl;
    LOG() << "Join time (seconds): " << NANOSECTOSEC(join_time_in_ns) << std::endl;
    LOG() << "Total time (seconds): " << NANOSECTOSEC(total_time_in_ns) << std::endl;
    LOG() << "Memory cost (MB): " << BYTESTOMB(memory_cost_in_bytes) << std::endl;
    LOG() << "#Embeddings: " << embedding_count << std::endl;
    LOG() << "Call Count: " << call_count.load() << std::endl;
    log4Metrics();
    LOG() << "Per Call Time (nanoseconds): "
          << enumeration_time_in_ns / (call_count.load() == 0 ? 1 : (int64_t)call_count.load()) << std::endl;
    LOG() << "Overtime: " << overtime.load() << std::endl;
    LOG() << "End the query" << std::endl;
}

TaskPool::TaskPool(uint64_t threads, uint64_t capacity) : num_threads(threads), QUEUE_CAPACITY((int64_t)capacity) {}

// ---------------------------------------------------------------------------
// JoinCollect
// ---------------------------------------------------------------------------
JoinCollect::JoinCollect(Graph* d, Graph* g, JoinSequence* js, ui cnt, int64_t tl, TaskPool* pool)
    : data_graph(d), graph(g), jseq(js), unit_cnt(cnt), time_limit(tl), taskPool(pool) {
    partial_matches = new PartialMatch*[std::max<ui>(1, cnt)];
    for (ui i = 0; i < std::max<ui>(1, cnt); ++i) partial_matches[i] = nullptr;
}

JoinCollect::~JoinCollect() { delete[] partial_matches; }

PartialMatch* JoinCollect::sequence_collect() {
    const ui n = graph->getVerticesCount();
    PartialMatch* res = new PartialMatch(n, n, 0);
    if (!jseq) {
        // one unit: its rows, as the reference's single-unit sequence gives
        if (unit_cnt >= 1 && partial_matches[0]) {
            res->data = partial_matches[0]->data;
            res->size = partial_matches[0]->size;
        }
        return res;
    }
    int p;
    switch (jseq->pattern) {
        case SplitPattern::STAR: p = PSM_SPLIT_STAR; break;
        case SplitPattern::TWINTWIG: p = PSM_SPLIT_TWINTWIG; break;
        case SplitPattern::SEED: p = PSM_SPLIT_SEED; break;
        default: delete res; throw std::runtime_error("split pattern not supported (STAR, TWINTWIG, SEED)");
    }
    psm_set_time_limit(psm_dropin::context(), (double)time_limit);
    psm_split_join_stats st;
    const int r = psm_split_join_count(existing(graph), p, 0, &st);
    overtime = r == PSM_ETIMEOUT;
    if (!overtime) check(r, "psm_split_join_count");
    res->size = overtime ? 0 : (ui)std::min<uint64_t>(st.count, UINT_MAX);
    call_count = overtime ? 0 : st.calls;
    enumerate_seconds = overtime ? 0 : st.enum_ms / 1e3;
    join_seconds = overtime ? 0 : st.join_ms / 1e3;
    return res;
}

// ---------------------------------------------------------------------------
// ComputeSetIntersection (host utility)
// ---------------------------------------------------------------------------
namespace {
ui gallop_intersect(const VertexID* small, ui ns, const VertexID* big, ui nb, VertexID* out) {
    ui c = 0, lo = 0;
    for (ui i = 0; i < ns && lo < nb; ++i) {
        const VertexID x = small[i];
        ui step = 1, hi = lo;
        while (hi < nb && big[hi] < x) { lo = hi + 1; hi += step; step <<= 1; }
        hi = std::min(hi, nb);
        lo = (ui)(std::lower_bound(big + lo, big + hi, x) - big);
        if (lo < nb && big[lo] == x) {
            if (out) out[c] = x;
            ++c;
            ++lo;
        }
    }
    return c;
}
}  // namespace

void ComputeSetIntersection::ComputeCandidates(const VertexID* l, ui lc, const VertexID* r, ui rc, VertexID* cn,
                                               ui& cnc) {
    if (lc == 0 || rc == 0) { cnc = 0; return; }
    if ((uint64_t)lc * 50 < rc) { cnc = gallop_intersect(l, lc, r, rc, cn); return; }
    if ((uint64_t)rc * 50 < lc) { cnc = gallop_intersect(r, rc, l, lc, cn); return; }
    ui i = 0, j = 0, c = 0;
    while (i < lc && j < rc) {
        if (l[i] < r[j]) ++i;
        else if (l[i] > r[j]) ++j;
        else { if (cn) cn[c] = l[i]; ++c; ++i; ++j; }
    }
    cnc = c;
}

void ComputeSetIntersection::ComputeCandidates(const VertexID* l, ui lc, const VertexID* r, ui rc, ui& cnc) {
    ComputeCandidates(l, lc, r, rc, nullptr, cnc);
}

// ---------------------------------------------------------------------------
// MatchingCommand: "-key value" pairs (utility/commandparser.cpp:3-19)
// ---------------------------------------------------------------------------
MatchingCommand::MatchingCommand(int argc, char** argv) {
    for (int i = 1; i + 1 < argc; ++i)
        if (argv[i][0] == '-') opts_[argv[i]] = argv[i + 1];
}

std::string MatchingCommand::get(const std::string& key, const std::string& def) const {
    auto it = opts_.find(key);
    return it == opts_.end() || it->second.empty() ? def : it->second;
}
