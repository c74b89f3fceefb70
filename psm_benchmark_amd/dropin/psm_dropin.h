// psm_dropin.h — the reference matcher's C++ call surfaces over the C ABI
// of libpsm.so (include/psm_abi.h). Class and method names, signatures,
// argument meaning, ownership and return conventions follow the reference so
// that its driver code (split_C_test, matching/StudyPerformance.cpp:
// 1164-1716) calls these in place of its CPU implementations:
//
//   Graph, TreeNode, Edges          graph/graph.h:21-237, configuration/types.h:18-83
//   FilterVertices                  matching/FilterVertices.h:7-60
//   BuildTable                      matching/BuildTable.h:7-26
//   GenerateFilteringPlan           matching/GenerateFilteringPlan.h:10-30
//   GenerateQueryPlan               matching/GenerateQueryPlan.h:28-60
//   EvaluateQuery                   matching/EvaluateQuery.h:28-134
//   UnitArgs, GLOBALARGS            matching/Args.h:30-236
//   TaskSlot, TaskPool              matching/TaskPool.h:31-120
//   SplitGraph, InducedSubgraph     utility/splitGraph/SplitGraph.h:10-30,82-147,602-702
//   PartialMatch, JoinCollect       utility/relations/partialMatch.h:11-40, matching/JoinCollect.h:9-49
//   ComputeSetIntersection          utility/computesetintersection.h:13-23
//   MatchingCommand                 matching/matchingcommand.h:32-160
//
// Candidates, tables and enumeration are computed by libpsm.so on the GPU.
// Device state of one (data graph, query graph) pair lives in a session
// keyed by the query Graph and is released with it. The query-split path
// (SPLITMODE::Q) keeps the library's own decomposition: SplitGraph then holds
// the whole query as one unit and JoinCollect::sequence_collect splits,
// enumerates and joins on the device.
#pragma once

#include <atomic>
#include <cstdint>
#include <functional>
#include <map>
#include <mutex>
#include <ostream>
#include <sstream>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "psm_abi.h"

typedef unsigned int ui;
typedef unsigned long ul;
typedef uint32_t VertexID;
typedef ui LabelID;

#define NANOSECTOSEC(elapsed_time) ((elapsed_time) / (double)1000000000)
#define BYTESTOMB(memory_cost) ((memory_cost) / (double)(1024 * 1024))

namespace psm_dropin {
psm_ctx* context();   // process-wide context on GPU 0 (PSM_DEVICE overrides)
// Where LOG() lines go: the file set here (appended), or stdout when empty.
void set_log_file(const std::string& path);
// One log line, "[Thread 1] " + text, written when it goes out of scope.
class LogLine {
public:
    LogLine();
    ~LogLine();
    template <class T>
    LogLine& operator<<(const T& v) {
        oss_ << v;
        return *this;
    }
    LogLine& operator<<(std::ostream& (*m)(std::ostream&)) {
        oss_ << m;
        return *this;
    }

private:
    std::ostringstream oss_;
};
}  // namespace psm_dropin

#define LOG() psm_dropin::LogLine()

// BFS tree node (configuration/types.h:18-61).
class TreeNode {
public:
    VertexID id_ = 0;
    VertexID parent_ = 0;
    ui level_ = 0;
    ui under_level_count_ = 0;
    ui children_count_ = 0;
    ui bn_count_ = 0;
    ui fn_count_ = 0;
    VertexID* under_level_ = nullptr;
    VertexID* children_ = nullptr;
    VertexID* bn_ = nullptr;
    VertexID* fn_ = nullptr;
    size_t estimated_embeddings_num_ = 0;

    TreeNode() = default;
    TreeNode(const TreeNode&) = delete;
    TreeNode& operator=(const TreeNode&) = delete;
    ~TreeNode() {
        delete[] under_level_;
        delete[] bn_;
        delete[] fn_;
        delete[] children_;
    }
    void initialize(const ui size) {
        delete[] under_level_;
        delete[] bn_;
        delete[] fn_;
        delete[] children_;
        under_level_ = new VertexID[size];
        bn_ = new VertexID[size];
        fn_ = new VertexID[size];
        children_ = new VertexID[size];
    }
};

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

class Graph {
public:
    explicit Graph(bool enable_label_offset = true);
    ~Graph();
    Graph(const Graph&) = delete;
    Graph& operator=(const Graph&) = delete;

    void loadGraphFromFile(const std::string& file_path);
    // The binary triple (graph.cpp:221-403). Throws std::runtime_error on a
    // missing file, a short read or degrees that do not sum to 2E.
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

// ---- query split (utility/splitGraph/SplitGraph.h) ------------------------
enum SPLITMODE { Q = 0, C = 1, NOSPLIT = 2 };
enum SplitPattern { HYBRIDPATH_RANDOM = 0, TWINTWIG = 1, STAR = 2, SEED = 3 };
enum JoinParadigm { LEFTDEEP = 0, BUSHY = 1 };

// A query unit: its own graph (ids 0..n'-1, labels renumbered by first
// appearance, graph.cpp:585-734) and the maps to and from the query's ids.
class InducedSubgraph {
public:
    InducedSubgraph(ui vertices_count, ui label_count);
    ~InducedSubgraph();
    InducedSubgraph(const InducedSubgraph&) = delete;
    InducedSubgraph& operator=(const InducedSubgraph&) = delete;
    ui getVerticesCount() const { return subgraph->getVerticesCount(); }

    ui total_vertices_count;
    ui* path = nullptr;
    Graph* subgraph = nullptr;
    ui* old2news = nullptr;
    ui* new2olds = nullptr;
    ui* l_old2news = nullptr;
    ui* l_new2olds = nullptr;
};

// The join order of the units (SplitGraph.h:602-613); here the split pattern
// the device decomposition applies.
class JoinSequence {
public:
    explicit JoinSequence(SplitPattern p) : pattern(p) {}
    SplitPattern pattern;
};

class SplitGraph {
public:
    // NOSPLIT / C: one unit, the whole query (SplitGraph.h:622-636). Q: one
    // unit as well, plus the join sequence for JoinCollect, whose
    // sequence_collect decomposes the query on the device.
    SplitGraph(Graph* d_graph, Graph* graph, JoinParadigm join_paradigm, SplitPattern split_pattern,
               SPLITMODE split_mode, int64_t time_limit);
    ~SplitGraph();
    SplitGraph(const SplitGraph&) = delete;
    SplitGraph& operator=(const SplitGraph&) = delete;
    std::vector<InducedSubgraph*> getSubgraphs() { return subgraphs_; }
    int getSubgraphCount() { return (int)subgraphs_.size(); }

    Graph* d_graph;
    Graph* graph;
    JoinParadigm join_paradigm;
    SplitPattern split_pattern;
    SPLITMODE split_mode;
    int64_t time_limit;
    JoinSequence* jseq = nullptr;

private:
    std::vector<InducedSubgraph*> subgraphs_;
};

// Rows of embeddings (utility/relations/partialMatch.h:11-40).
class PartialMatch {
public:
    PartialMatch() = default;
    PartialMatch(ui length, ui total_vcnt, ui init_capacity = 0);
    ~PartialMatch();
    void add(const ui* match);
    ui getSize() { return size; }
    ui* getTuple(ui row_idx) { return data.data() + (size_t)row_idx * length; }

    std::vector<ui> data;
    ui size = 0;
    ui length = 0;
    ui init_capacity = 0;
    ui total_vcnt = 0;
    ui* new2olds = nullptr;
    ui* old2news = nullptr;
};

// ---- arguments (matching/Args.h:30-236) -----------------------------------
class GLOBALARGS;

class UnitArgs {
public:
    UnitArgs() = default;
    UnitArgs(InducedSubgraph* query_graph, size_t output_limit_num);
    ~UnitArgs();
    UnitArgs(const UnitArgs&) = delete;
    UnitArgs& operator=(const UnitArgs&) = delete;

    // candidates / tables of the unit's vertices point into the query's
    // (Args.cpp:104-160); they also tie the unit to the query's device session
    void mapUidCandidates(GLOBALARGS* g_args);
    void mapUidEdgeMatrix(GLOBALARGS* g_args);
    void addPartialMatch(ui* embedding);

    const InducedSubgraph* query_graph = nullptr;
    Edges*** edge_matrix = nullptr;
    ui** candidates = nullptr;
    ui* candidates_count = nullptr;
    TreeNode* tree = nullptr;
    ui* tree_order = nullptr;   // BFS order of the unit's filtering plan
    ui* order = nullptr;        // matching order (unit ids)
    ui* pivots = nullptr;
    ui** bn = nullptr;
    ui* bn_count = nullptr;
    ui output_limit = 0;
    ui vertices_count = 0;
    PartialMatch* unit_partial_match = nullptr;
    bool needCandidates = false;
    bool needEdgeMatrix = false;
    GLOBALARGS* g_args = nullptr;   // set by mapUidCandidates / mapUidEdgeMatrix

private:
    std::mutex pm_mtx_;
};

class GLOBALARGS {
public:
    // Takes ownership of edge_matrix, candidates and candidates_count
    // (Args.cpp:366-400).
    GLOBALARGS(const Graph* data_graph, const Graph* query_graph, Edges*** edge_matrix, ui** candidates,
               ui* candidates_count, size_t output_limit_num, int64_t time_limit);
    ~GLOBALARGS();
    GLOBALARGS(const GLOBALARGS&) = delete;
    GLOBALARGS& operator=(const GLOBALARGS&) = delete;

    void addTimeExec(std::thread::id thread_id, int64_t elapsed_time);
    void addTimeEnum(std::thread::id thread_id, int64_t elapsed_time);
    void addTimeSplit(std::thread::id thread_id, int64_t elapsed_time);
    void addCallCount(std::thread::id thread_id, int64_t new_call_count);
    void addEmbeddingCount(std::thread::id thread_id, int64_t new_embedding_count);
    void clearStatistics();
    // Per-thread lines (Args.cpp:456-515): the host threads that drove the
    // device; "Execution" is device enumeration time, "Split" the task
    // pre-expansion time, "Idle" the rest of the thread's enumeration time.
    void log4Metrics();
    void logGlobalStatistics();   // Args.cpp:517-535

    SPLITMODE splitMode = SPLITMODE::NOSPLIT;
    int count_unit = 0;
    UnitArgs** unitArgsVec = nullptr;

    const Graph* data_graph = nullptr;
    const Graph* query_graph = nullptr;
    Edges*** edge_matrix = nullptr;
    ui** candidates = nullptr;
    ui* candidates_count = nullptr;
    TreeNode* tree = nullptr;
    ui* tree_order = nullptr;
    ui* order = nullptr;
    ui* pivots = nullptr;
    std::string engine_type;

    int64_t output_limit_num = 0;
    std::atomic<int64_t> embedding_count{0};
    std::atomic<ul> call_count{0};
    int64_t time_limit = 0;
    std::atomic<bool> overtime{false};
    int total_task_num = 0;
    bool needCandidates = false;
    bool needEdgeMatrix = false;

    std::mutex global_mutex;
    std::unordered_map<std::thread::id, int64_t> thread_times_enum, thread_times_exec, thread_times_split,
        thread_call_count, thread_embedding_count;

    int64_t load_graphs_time_in_ns = 0;
    int64_t filter_vertices_time_in_ns = 0;
    int64_t split_query_time_in_ns = 0;
    int64_t build_table_time_in_ns = 0;
    int64_t generate_query_plan_time_in_ns = 0;
    int64_t enumeration_time_in_ns = 0;
    int64_t preprocessing_time_in_ns = 0;
    int64_t join_time_in_ns = 0;
    int64_t total_time_in_ns = 0;
    int64_t memory_cost_in_bytes = 0;
};

// ---- scheduler (matching/TaskPool.h) --------------------------------------
enum ScheduleType { STATIC = 0, TIMEOUT = 1, BUSY2IDLENOSTOP = 2, BUSY2IDLEDEPTHSTOP = 3, STATICWORKLOAD = 4 };
enum SplitType { LINEAR = 1, WORKLOAD = 2, UPPERONE = 3, LAYER = 4 };

// One unit of enumeration work: a query unit (its device session). The
// device's own work queue splits it further (psm_evaluate).
class TaskSlot {
public:
    TaskSlot() = default;
    int id_unit = -1;
    UnitArgs* u_args = nullptr;
    static GLOBALARGS* g_args;
};

// The reference's host thread pool; here its settings are recorded and the
// device work queue does the scheduling (schedule and split types do not
// change results).
class TaskPool {
public:
    TaskPool(uint64_t num_threads, uint64_t queue_capacity);
    bool Empty() { return true; }
    std::size_t Size() { return 0; }
    uint64_t num_threads;
    int64_t QUEUE_CAPACITY;
    ScheduleType schedule_type = ScheduleType::BUSY2IDLENOSTOP;
    SplitType split_type = SplitType::LINEAR;
};

// ---- the matching path ----------------------------------------------------
class FilterVertices {
public:
    static bool LDFFilter(const Graph* data_graph, const Graph* query_graph, ui**& candidates,
                          ui*& candidates_count);
    static bool NLFFilter(const Graph* data_graph, const Graph* query_graph, ui**& candidates,
                          ui*& candidates_count);
    static bool GQLFilter(const Graph* data_graph, const Graph* query_graph, ui**& candidates,
                          ui*& candidates_count);
    // order / tree: caller-allocated |V(query)| entries (StudyPerformance.cpp:
    // 1175-1188); they receive the BFS order and tree the filter used.
    static bool CFLFilter(const Graph* data_graph, const Graph* query_graph, ui**& candidates,
                          ui*& candidates_count, ui*& order, TreeNode*& tree);
    static bool DPisoFilter(const Graph* data_graph, const Graph* query_graph, ui**& candidates,
                            ui*& candidates_count, ui*& order, TreeNode*& tree);
    // Candidates come out sorted from the device; kept for call-site parity.
    static void sortCandidates(ui** candidates, ui* candidates_count, ui num);
    // true when the last filter stopped at the time limit set through
    // time_limit (seconds, 0: none; FilterVertices.h:49-50)
    static bool checkOverTime();
    static int64_t time_limit;
};

class BuildTable {
public:
    static void buildTables(const Graph* data_graph, const Graph* query_graph, ui** candidates,
                            ui* candidates_count, Edges*** edge_matrix);
    static size_t computeMemoryCostInBytes(const Graph* query_graph, ui* candidates_count,
                                           Edges*** edge_matrix);
};

class GenerateFilteringPlan {
public:
    // The unit's CFL / DPiso BFS tree (GenerateFilteringPlan.cpp:15-79):
    // u_args->tree (|V| nodes, allocated here when null) and tree_order.
    static void generateCFLFilterPlan(const Graph* data_graph, UnitArgs* u_args);
    static void generateDPisoFilterPlan(const Graph* data_graph, UnitArgs* u_args);
};

class GenerateQueryPlan {
public:
    // u_args->order / pivots receive the matching order (GenerateQueryPlan.h:32,49,55)
    static void generateGQLQueryPlan(const Graph* data_graph, const Graph* query_graph, UnitArgs* u_args);
    static void generateCFLQueryPlan(const Graph* query_graph, UnitArgs* u_args);
    static void generateDPisoQueryPlan(const Graph* query_graph, UnitArgs* u_args);
};

class EvaluateQuery {
public:
    // Runs enum_method on one task per unit and returns the embedding count
    // (EvaluateQuery.cpp:26-112); g_args receives counts and statistics.
    static size_t ParallelExecute(TaskPool* taskPool, SplitGraph* splitGraph, UnitArgs** unitArgsVec);
    // The engines: the task's unit is enumerated on the device in the unit's
    // order (psm_set_plan + psm_evaluate); counts are the same for all three
    // and each adds its own Call Count (LFTJ: + visited rejects).
    static void LFTJ(TaskPool* taskPool, TaskSlot& task, int task_id);
    static void exploreGraphQLStyle(TaskPool* taskPool, TaskSlot& task, int task_id);
    static void exploreGraph(TaskPool* taskPool, TaskSlot& task, int task_id);
    static std::function<void(TaskSlot&, int)> enum_method;

    // EvaluateQuery.cpp:1594-1666
    static void generateBN(const Graph* query_graph, ui* order, ui* pivot, ui** bn, ui* bn_count);
    static void generateBN(const Graph* query_graph, ui* order, ui** bn, ui* bn_count);
    static void generateBN(UnitArgs* u_args);
    static void generateBNWithPivot(UnitArgs* u_args);
};

class JoinCollect {
public:
    JoinCollect(Graph* data_graph, Graph* g, JoinSequence* jseq, ui cnt, int64_t time_limit, TaskPool* taskPool);
    ~JoinCollect();
    JoinCollect(const JoinCollect&) = delete;
    JoinCollect& operator=(const JoinCollect&) = delete;
    // With a join sequence (SPLITMODE::Q): the query is split by the
    // sequence's pattern, its units enumerated and joined on the device; the
    // result's size is the embedding count (rows stay on the device).
    // Without one: the single unit's rows (none unless materialised).
    PartialMatch* sequence_collect();

    PartialMatch** partial_matches = nullptr;
    Graph* data_graph = nullptr;
    Graph* graph = nullptr;
    JoinSequence* jseq = nullptr;
    ui unit_cnt = 0;
    int64_t time_limit = 0;
    bool overtime = false;
    TaskPool* taskPool = nullptr;
    // filled by sequence_collect on the device path
    ul call_count = 0;
    double enumerate_seconds = 0, join_seconds = 0;
};

class ComputeSetIntersection {
public:
    // sorted l ∩ sorted r (galloping when one side is more than 50 times the
    // other, merge otherwise; computesetintersection.cpp:5-40)
    static void ComputeCandidates(const VertexID* larray, ui l_count, const VertexID* rarray, ui r_count,
                                  VertexID* cn, ui& cn_count);
    static void ComputeCandidates(const VertexID* larray, ui l_count, const VertexID* rarray, ui r_count,
                                  ui& cn_count);
};

// Command line options and defaults of the reference (matchingcommand.h:70-159).
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
