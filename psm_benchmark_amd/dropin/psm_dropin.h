// psm_dropin.h — the reference matcher's C++ call surfaces over the C ABI
// of libpsm.so (include/psm_abi.h). Class and method names, argument meaning,
// ownership and return conventions follow the reference so that its driver
// code can call these in place of its CPU implementations:
//
//   Graph                      graph/graph.h:21-237
//   FilterVertices             matching/FilterVertices.h:7-60
//   BuildTable                 matching/BuildTable.h:7-26
//   GenerateQueryPlan          matching/GenerateQueryPlan.h:7-85
//   EvaluateQuery              matching/EvaluateQuery.h:28-134
//   MatchingCommand            matching/matchingcommand.h:32-160
//
// Device state (candidates, tables) of one (data graph, query graph) pair is
// held in a session keyed by the query Graph and released with it.
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "psm_abi.h"

typedef unsigned int ui;
typedef uint32_t VertexID;
typedef ui LabelID;

// Candidate-edge table (configuration/types.h:63-83).
class Edges {
public:
    ui* offset_ = nullptr;
    ui* edge_ = nullptr;
    ui vertex_count_ = 0;
    ui edge_count_ = 0;
    ui max_degree_ = 0;
    ~Edges() {
        delete[] offset_;
        delete[] edge_;
    }
};

namespace psm_dropin {
psm_ctx* context();   // process-wide context on GPU 0 (PSM_DEVICE overrides)
}

class Graph {
public:
    explicit Graph(bool enable_label_offset = true);
    ~Graph();
    Graph(const Graph&) = delete;
    Graph& operator=(const Graph&) = delete;

    void loadGraphFromFile(const std::string& file_path);
    void loadGraphFromFileCompressed(const std::string& degree_path, const std::string& edge_path,
                                     const std::string& label_path);
    void storeComparessedGraph(const std::string& degree_path, const std::string& edge_path,
                               const std::string& label_path);
    void loadFromCSR(ui vertices, const ui* offsets, const ui* neighbors, const ui* labels);
    void buildCoreTable();

    ui getLabelsCount() const { return labels_count_; }
    ui getVerticesCount() const { return vertices_count_; }
    ui getEdgesCount() const { return edges_count_; }
    ui getGraphMaxDegree() const { return max_degree_; }
    ui getGraphMaxLabelFrequency() const { return max_label_frequency_; }
    ui getVertexDegree(VertexID id) const { return offsets_[id + 1] - offsets_[id]; }
    ui getLabelsFrequency(LabelID label) const;
    ui getCoreValue(VertexID id) const { return core_[id]; }
    ui get2CoreSize() const { return core_length_; }
    LabelID getVertexLabel(VertexID id) const { return labels_[id]; }
    const ui* getVertexNeighbors(VertexID id, ui& count) const {
        count = offsets_[id + 1] - offsets_[id];
        return neighbors_.data() + offsets_[id];
    }
    const ui* getVerticesByLabel(LabelID id, ui& count) const;
    bool checkEdgeExistence(VertexID u, VertexID v) const;

    // device twin (uploaded on first use)
    psm_graph* device() const;
    const std::vector<ui>& offsets() const { return offsets_; }
    const std::vector<ui>& neighbors() const { return neighbors_; }
    const std::vector<ui>& labels() const { return labels_; }

private:
    void finish();
    ui vertices_count_ = 0, edges_count_ = 0, labels_count_ = 0, max_degree_ = 0;
    ui max_label_frequency_ = 0, core_length_ = 0;
    std::vector<ui> offsets_, neighbors_, labels_, ri_off_, ri_, core_;
    mutable psm_graph* dev_ = nullptr;
};

class FilterVertices {
public:
    static bool LDFFilter(const Graph* data_graph, const Graph* query_graph, ui**& candidates,
                          ui*& candidates_count);
    static bool NLFFilter(const Graph* data_graph, const Graph* query_graph, ui**& candidates,
                          ui*& candidates_count);
    static bool GQLFilter(const Graph* data_graph, const Graph* query_graph, ui**& candidates,
                          ui*& candidates_count);
    static bool CFLFilter(const Graph* data_graph, const Graph* query_graph, ui**& candidates,
                          ui*& candidates_count);
    // FilterVertices.h:17; the DPiso tree and order stay inside the library
    // (generateDPisoQueryPlan rebuilds them on the query unit, as the
    // reference's driver does).
    static bool DPisoFilter(const Graph* data_graph, const Graph* query_graph, ui**& candidates,
                            ui*& candidates_count);
    // Candidates come out sorted from the device; kept for call-site parity.
    static void sortCandidates(ui** candidates, ui* candidates_count, ui num);
};

class BuildTable {
public:
    static void buildTables(const Graph* data_graph, const Graph* query_graph, ui** candidates,
                            ui* candidates_count, Edges*** edge_matrix);
    static size_t computeMemoryCostInBytes(const Graph* query_graph, ui* candidates_count,
                                           Edges*** edge_matrix);
};

class GenerateQueryPlan {
public:
    // order / pivot: caller-allocated arrays of |V(query)| entries.
    static void generateGQLQueryPlan(const Graph* data_graph, const Graph* query_graph,
                                     ui* candidates_count, ui* order, ui* pivot);
    static void generateCFLQueryPlan(const Graph* data_graph, const Graph* query_graph,
                                     ui* candidates_count, ui* order, ui* pivot);
    // GenerateQueryPlan.h:55 (order = the unit's DPiso BFS order, pivot = BFS
    // parent; operands are all earlier neighbours).
    static void generateDPisoQueryPlan(const Graph* data_graph, const Graph* query_graph,
                                       ui* candidates_count, ui* order, ui* pivot);
};

class EvaluateQuery {
public:
    // The engine whose Call Count evaluate() reports ("GQL", "EXPLORE" or
    // "LFTJ"); the reference picks EvaluateQuery::enum_method by engine
    // (EvaluateQuery.h:133, StudyPerformance.cpp:1692-1706). Counts do not
    // depend on it.
    static std::string engine_type;
    // Counts embeddings of the query prepared by the calls above (at most
    // output_limit_num); call_count receives the reference's Call Count.
    // Returns the embedding count. overtime is set when time_limit_s passes.
    static size_t evaluate(const Graph* data_graph, const Graph* query_graph, size_t output_limit_num,
                           size_t& call_count, bool& overtime, double time_limit_s = 60.0);
};

// Command line options and defaults of the reference (matchingcommand.h:70-159).
// Q mode (-QorCandi Q, StudyPerformance.cpp:429-437): the query prepared by
// the filter and buildTables calls above is split by -Qpattern (STAR,
// TWINTWIG or SEED), its units are enumerated on the GPU and joined one by
// one (JoinCollect.cpp:60-153). Returns the embedding count; call_count
// receives the assignments made enumerating the units (the units are this
// library's own cover, so it differs from the reference's Q-mode count).
class JoinCollect {
public:
    static size_t sequence_collect(const Graph* query_graph, const std::string& pattern, size_t& call_count,
                                   bool& overtime, double& enumerate_seconds, double& join_seconds,
                                   double time_limit_s = 60.0);
};

class MatchingCommand {
public:
    MatchingCommand(int argc, char** argv);
    std::string get(const std::string& key, const std::string& def = "") const;
    std::string getDataGraphFilePath() const { return get("-d"); }
    std::string getQueryGraphFilePath() const { return get("-q"); }
    std::string getMaximumEmbeddingNum() const { return get("-num", "MAX"); }
    std::string getTimeLimit() const { return get("-time_limit", "60"); }
    std::string getOutputFile() const { return get("-OutputFile", ""); }
    std::string getThreadNumbers() const { return get("-threadnums", "1"); }
    std::string getSplitType() const { return get("-split", "shixuan"); }
    std::string getScheduleType() const { return get("-schedule", "steal"); }
    std::string getQorCandiType() const { return get("-QorCandi", "NOSPLIT"); }
    std::string getQpatternType() const { return get("-Qpattern", "STAR"); }
    std::string getJoinMethod() const { return get("-JoinMethod", "LEFTDEEP"); }
    std::string getBackMethodTypes() const { return get("-BackMethods", "DPiso"); }
    std::string getFilterType() const { return get("-filter", "CFL"); }
    std::string getOrderType() const { return get("-order", "GQL"); }
    std::string getEngineType() const { return get("-engine", "LFTJ"); }
    std::string getAlgorithm() const { return get("-a"); }
    std::string getIndexType() const { return get("-i", "VertexCentric"); }
    std::string getDepthThreshold() const { return get("-d0", "0"); }
    std::string getWidthThreshold() const { return get("-w0", "1"); }
    std::string getOrderNum() const { return get("-order_num", "100"); }
    std::string getDistributionFilePath() const { return get("-dis_file", "temp.distribution"); }
    std::string getCSRFilePath() const { return get("-csr"); }

private:
    std::map<std::string, std::string> opts_;
};
