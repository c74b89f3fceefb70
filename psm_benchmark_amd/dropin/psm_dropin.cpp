// psm_dropin.cpp — implementation of the reference's C++ call surfaces over
// the C ABI (see psm_dropin.h). Host-side bookkeeping only; every candidate
// set, table and enumeration is computed by libpsm.so on the GPU.

#include "psm_dropin.h"

#include <algorithm>
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

// One device query per query Graph (GLOBALARGS/UnitArgs device state).
std::unordered_map<const Graph*, psm_query*> g_sessions;

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
// (FilterVertices.cpp:689-702); the caller owns them.
bool run_filter(const Graph* data, const Graph* query, int method, ui**& candidates, ui*& counts) {
    psm_query* q = session(data, query, true);
    const int r = psm_filter(q, method);
    check(r, "psm_filter");
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
// neighbours vertex by vertex, u32 labels.
void Graph::loadGraphFromFileCompressed(const std::string& dp, const std::string& ep,
                                        const std::string& lp) {
    std::ifstream df(dp, std::ios::binary), ef(ep, std::ios::binary), lf(lp, std::ios::binary);
    if (!df || !ef || !lf) {
        std::cerr << "Cannot open the binary graph files." << std::endl;
        exit(-1);
    }
    int32_t isz;
    ui V, E;
    df.read(reinterpret_cast<char*>(&isz), 4);
    df.read(reinterpret_cast<char*>(&V), 4);
    df.read(reinterpret_cast<char*>(&E), 4);
    std::vector<ui> deg(V), off(V + 1, 0), adj((size_t)E * 2), lab(V);
    df.read(reinterpret_cast<char*>(deg.data()), 4ull * V);
    for (ui v = 0; v < V; ++v) off[v + 1] = off[v] + deg[v];
    ef.read(reinterpret_cast<char*>(adj.data()), 4ull * off[V]);
    lf.read(reinterpret_cast<char*>(lab.data()), 4ull * V);
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
// FilterVertices / BuildTable / GenerateQueryPlan / EvaluateQuery
// ---------------------------------------------------------------------------
bool FilterVertices::LDFFilter(const Graph* d, const Graph* q, ui**& c, ui*& n) {
    return run_filter(d, q, PSM_FILTER_LDF, c, n);
}
bool FilterVertices::NLFFilter(const Graph* d, const Graph* q, ui**& c, ui*& n) {
    return run_filter(d, q, PSM_FILTER_NLF, c, n);
}
bool FilterVertices::GQLFilter(const Graph* d, const Graph* q, ui**& c, ui*& n) {
    return run_filter(d, q, PSM_FILTER_GQL, c, n);
}
bool FilterVertices::CFLFilter(const Graph* d, const Graph* q, ui**& c, ui*& n) {
    return run_filter(d, q, PSM_FILTER_CFL, c, n);
}

bool FilterVertices::DPisoFilter(const Graph* d, const Graph* q, ui**& c, ui*& n) {
    return run_filter(d, q, PSM_FILTER_DPISO, c, n);
}

void FilterVertices::sortCandidates(ui** candidates, ui* counts, ui num) {
    for (ui i = 0; i < num; ++i) std::sort(candidates[i], candidates[i] + counts[i]);
}

// edge_matrix: caller-allocated n x n array of Edges* (StudyPerformance.cpp:
// 1251-1254); entries of query edges receive new Edges, the rest nullptr.
void BuildTable::buildTables(const Graph* data, const Graph* query, ui**, ui* counts,
                             Edges*** edge_matrix) {
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

void GenerateQueryPlan::generateGQLQueryPlan(const Graph*, const Graph* query, ui*, ui* order, ui* pivot) {
    check(psm_generate_plan(existing(query), PSM_ORDER_GQL, order, pivot), "psm_generate_plan");
}

void GenerateQueryPlan::generateCFLQueryPlan(const Graph*, const Graph* query, ui*, ui* order, ui* pivot) {
    check(psm_generate_plan(existing(query), PSM_ORDER_CFL, order, pivot), "psm_generate_plan");
}

void GenerateQueryPlan::generateDPisoQueryPlan(const Graph*, const Graph* query, ui*, ui* order, ui* pivot) {
    check(psm_generate_plan(existing(query), PSM_ORDER_DPISO, order, pivot), "psm_generate_plan");
}

std::string EvaluateQuery::engine_type = "GQL";

size_t EvaluateQuery::evaluate(const Graph*, const Graph* query, size_t limit, size_t& calls,
                               bool& overtime, double time_limit_s) {
    psm_set_time_limit(psm_dropin::context(), time_limit_s);
    psm_enum_stats st;
    const uint64_t cap = limit >= (size_t)INT64_MAX ? 0 : (uint64_t)limit;
    const int r = psm_evaluate(existing(query), cap, &st);
    check(r, "psm_evaluate");
    overtime = r == PSM_ETIMEOUT;
    calls = st.calls + (engine_type == "LFTJ" ? st.visited_rejects : 0);
    return st.count;
}

size_t JoinCollect::sequence_collect(const Graph* query, const std::string& pattern, size_t& calls,
                                     bool& overtime, double& enum_s, double& join_s, double time_limit_s) {
    int p;
    if (pattern == "STAR") p = PSM_SPLIT_STAR;
    else if (pattern == "TWINTWIG") p = PSM_SPLIT_TWINTWIG;
    else if (pattern == "SEED") p = PSM_SPLIT_SEED;
    else throw std::runtime_error("split pattern " + pattern + " is not supported");
    psm_set_time_limit(psm_dropin::context(), time_limit_s);
    psm_split_join_stats st;
    const int r = psm_split_join_count(existing(query), p, 0, &st);
    overtime = r == PSM_ETIMEOUT;
    if (overtime) {
        calls = 0;
        enum_s = join_s = 0;
        return 0;
    }
    check(r, "psm_split_join_count");
    calls = st.calls;
    enum_s = st.enum_ms / 1e3;
    join_s = st.join_ms / 1e3;
    return st.count;
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
