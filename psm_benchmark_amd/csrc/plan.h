// plan.h — host-side query preparation: the query graph with its derived
// indexes, the CFL and DPiso BFS filtering plans and the GQL / CFL / DPiso
// matching orders.
// These run on the host in the reference too (n <= 64, microseconds).
#pragma once

#include <cstdint>
#include <functional>
#include <utility>
#include <vector>

namespace psm {

struct QGraph {
    uint32_t n = 0;
    uint32_t L = 0;                                               // max label + 1
    std::vector<uint32_t> off, adj, lab;                          // CSR, adjacency ascending
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> nlf;  // label-sorted (label, count)
    std::vector<uint8_t> core2;                                   // 2-core membership
    uint32_t core2_size = 0;

    uint32_t degree(uint32_t u) const { return off[u + 1] - off[u]; }
    uint32_t max_degree() const;
    bool has_edge(uint32_t a, uint32_t b) const;
};

// Validates and indexes a query CSR (graph.cpp:92-172 + buildCoreTable
// :210-219). Returns a PSM status.
int make_qgraph(uint32_t n, const uint32_t* off, const uint32_t* adj, const uint32_t* lab,
                QGraph& out);

// The undecomposed query unit (SplitGraph.h:622-636 -> graph.cpp:585-734):
// same vertices and edges, labels renumbered by first appearance.
QGraph unit_graph(const QGraph& q);

struct Plan {
    std::vector<uint32_t> order, pivot;
    std::vector<std::vector<uint32_t>> bn;   // per depth: backward neighbours
};

// GenerateQueryPlan.cpp:9-58,149-184.
void gql_order(const QGraph& q, const std::vector<uint32_t>& cand_count, uint32_t data_V, Plan& p);
// EvaluateQuery.cpp:1594-1666.
void backward_neighbours(const QGraph& q, Plan& p, bool drop_pivot);

struct CflTree {
    std::vector<uint32_t> parent, level, order, level_off;
    std::vector<std::vector<uint32_t>> children, bn, fn, under;
};

using NlfCountFn = std::function<int(uint32_t u, uint32_t& count)>;
using TableFn = std::function<int(uint32_t a, uint32_t b, std::vector<uint32_t>& off,
                                  std::vector<uint32_t>& idx)>;

// GenerateFilteringPlan.cpp:81-140 (+ start vertex :343-385): BFS tree from
// the start vertex and each vertex's same/upper-level (bn), same-level-later
// (fn) and lower-level (under) neighbours. nlf_count(u) gives |NLF
// candidates| of u for the start-vertex score.
int cfl_filter_plan(const QGraph& q, uint32_t data_L, const std::vector<uint32_t>& data_ri_off,
                    uint32_t data_max_freq, const NlfCountFn& nlf_count, CflTree& t);

// GenerateFilteringPlan.cpp:183-214 (+ start vertex :387-406): BFS tree from
// the vertex of degree >= 2 with the fewest LDF candidates per unit of degree
// (ldf_count(u) gives |LDF candidates| of u); bn = neighbours earlier in the
// BFS order, fn = neighbours later (under and level_off stay empty).
int dpiso_filter_plan(const QGraph& q, uint32_t data_V, const NlfCountFn& ldf_count, CflTree& t);

// GenerateQueryPlan.cpp:1133-1160: the matching order is the DPiso BFS order
// of the unit graph and each later vertex's pivot is its BFS parent.
void dpiso_order(const CflTree& t, Plan& p);

// GenerateQueryPlan.cpp:270-471 on the unit graph and its BFS tree; table(a,b)
// supplies candidate-edge tables for the path embedding estimates
// (GenerateQueryPlan.cpp:680-721).
int cfl_order(const QGraph& ug, const CflTree& t, const std::vector<uint32_t>& cand_count,
              const TableFn& table, Plan& p);

}  // namespace psm
