// subgraph_matching.cpp — command-line driver with the reference's options
// and result log, running on the GPU through the drop-in classes. It makes
// the calls of the reference's split_C_test, in its order and with its
// signatures (matching/StudyPerformance.cpp:1164-1716): filter (CFL / DPiso
// with their order and tree), sortCandidates, buildTables, GLOBALARGS,
// SplitGraph, UnitArgs with mapUidCandidates / mapUidEdgeMatrix, the
// filtering and query plans with generateBN / generateBNWithPivot, TaskPool,
// EvaluateQuery::enum_method + ParallelExecute, JoinCollect::sequence_collect
// and GLOBALARGS::logGlobalStatistics (Args.cpp:517-535, with the per-thread
// lines of log4Metrics). -QorCandi C / NOSPLIT enumerates the query as one
// unit; Q with -Qpattern STAR, TWINTWIG or SEED splits and joins it on the
// device. Its log can be read by the reference's harness and parsers.
//
//   SubgraphMatching.out -d data.txt -q query.txt -num MAX -time_limit 60
//       -BackMethods GQL_GQL_GQL,CFL_CFL_EXPLORE -QorCandi C -OutputFile log.txt

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "psm_dropin.h"

namespace {

std::vector<std::string> split(const std::string& s, char c) {
    std::vector<std::string> out;
    std::stringstream ss(s);
    std::string t;
    while (std::getline(ss, t, c)) out.push_back(t);
    return out;
}

int64_t ns_since(std::chrono::steady_clock::time_point a) {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - a).count();
}

}  // namespace

int main(int argc, char** argv) {
    MatchingCommand command(argc, argv);
    const std::string qorc = command.getQorCandiType();
    if (qorc != "C" && qorc != "NOSPLIT" && qorc != "Q") {
        std::cerr << "Error: -QorCandi must be C, NOSPLIT or Q." << std::endl;
        return 1;
    }
    const std::string qpattern = command.getQpatternType();
    if (qorc == "Q" && qpattern != "STAR" && qpattern != "TWINTWIG" && qpattern != "SEED") {
        std::cerr << "Error: -Qpattern " << qpattern << " is not supported (STAR, TWINTWIG or SEED)." << std::endl;
        return 1;
    }
    psm_dropin::set_log_file(command.getOutputFile());
    auto start = std::chrono::steady_clock::now();
    Graph* data_graph = new Graph(true);
    Graph* query_graph = new Graph(true);
    data_graph->loadGraphFromFile(command.getDataGraphFilePath());
    query_graph->loadGraphFromFile(command.getQueryGraphFilePath());
    query_graph->buildCoreTable();
    const int64_t load_graphs_time_in_ns = ns_since(start);

    const int64_t time_limit = std::stoll(command.getTimeLimit());
    size_t output_limit = std::numeric_limits<int64_t>::max();
    if (command.getMaximumEmbeddingNum() != "MAX") output_limit = std::stoull(command.getMaximumEmbeddingNum());
    const ui n = query_graph->getVerticesCount();

    for (const std::string& threads : split(command.getThreadNumbers(), ',')) {
        const int num_threads = std::stoi(threads);
        for (const std::string& method : split(command.getBackMethodTypes(), ',')) {
            auto parts = split(method, '_');
            if (parts.size() != 3) {
                std::cerr << "bad method " << method << std::endl;
                return 1;
            }
            const std::string& filter_type = parts[0];
            const std::string& order_type = parts[1];
            const std::string& engine_type = parts[2];
            LOG() << "Start the query" << std::endl;
            LOG() << "Input Query Graph File: " << command.getQueryGraphFilePath() << std::endl;
            LOG() << "Input Data Graph File: " << command.getDataGraphFilePath() << std::endl;
            LOG() << "Maximum Embedding Number: " << command.getMaximumEmbeddingNum() << std::endl;
            LOG() << "Time Limit: " << command.getTimeLimit() << std::endl;
            LOG() << "Num Threads: " << threads << std::endl;
            LOG() << "Split Type: " << command.getSplitType() << std::endl;
            if (qorc == "Q") LOG() << "QuerySplit Pattern: " << qpattern << std::endl;
            LOG() << "Schedule Type: " << command.getScheduleType() << std::endl;
            LOG() << "Filter Type: " << filter_type << std::endl;
            LOG() << "Order Type: " << order_type << std::endl;
            LOG() << "Engine Type: " << engine_type << std::endl;
            LOG() << "Output Limit: " << (int64_t)std::min<size_t>(output_limit, std::numeric_limits<int64_t>::max())
                  << std::endl;
            if (engine_type != "GQL" && engine_type != "EXPLORE" && engine_type != "LFTJ") {
                LOG() << "Not supported engine type: " << engine_type << std::endl;
                return 1;
            }

            // ---- filter (StudyPerformance.cpp:1164-1200)
            FilterVertices::time_limit = time_limit;
            start = std::chrono::steady_clock::now();
            ui** candidates = nullptr;
            ui* candidates_count = nullptr;
            ui* cfl_order = nullptr;
            TreeNode* cfl_tree = nullptr;
            ui* dpiso_order = nullptr;
            TreeNode* dpiso_tree = nullptr;
            if (filter_type == "LDF") {
                FilterVertices::LDFFilter(data_graph, query_graph, candidates, candidates_count);
            } else if (filter_type == "NLF") {
                FilterVertices::NLFFilter(data_graph, query_graph, candidates, candidates_count);
            } else if (filter_type == "GQL") {
                FilterVertices::GQLFilter(data_graph, query_graph, candidates, candidates_count);
            } else if (filter_type == "CFL") {
                cfl_order = new ui[n];
                cfl_tree = new TreeNode[n];
                FilterVertices::CFLFilter(data_graph, query_graph, candidates, candidates_count, cfl_order, cfl_tree);
            } else if (filter_type == "DPiso") {
                dpiso_order = new ui[n];
                dpiso_tree = new TreeNode[n];
                FilterVertices::DPisoFilter(data_graph, query_graph, candidates, candidates_count, dpiso_order,
                                            dpiso_tree);
            } else {
                LOG() << "The specified filter type '" << filter_type << "' is not supported." << std::endl;
                return 1;
            }
            FilterVertices::sortCandidates(candidates, candidates_count, n);
            const int64_t filter_vertices_time_in_ns = ns_since(start);
            if (FilterVertices::checkOverTime()) {
                LOG() << "Timeout 1." << std::endl;
                LOG() << "Load graphs time (seconds): " << NANOSECTOSEC(load_graphs_time_in_ns) << std::endl;
                LOG() << "Build table time (seconds): " << 0 << std::endl;
                LOG() << "Filter vertices time (seconds): " << NANOSECTOSEC(filter_vertices_time_in_ns) << std::endl;
                LOG() << "Generate query plan time (seconds): " << 0 << std::endl;
                LOG() << "Enumerate time (seconds): " << 0 << std::endl;
                LOG() << "Preprocessing time (seconds): " << 0 << std::endl;
                LOG() << "Total time (seconds): " << NANOSECTOSEC(filter_vertices_time_in_ns) << std::endl;
                LOG() << "Memory cost (MB): " << 0 << std::endl;
                LOG() << "#Embeddings: " << 0 << std::endl;
                LOG() << "Call Count: " << 0 << std::endl;
                LOG() << "Per Call Time (nanoseconds): " << 0 << std::endl;
                LOG() << "Overtime: " << 1 << std::endl;
                LOG() << "End the query" << std::endl;
                for (ui i = 0; i < n; ++i) delete[] candidates[i];
                delete[] candidates;
                delete[] candidates_count;
                delete[] cfl_order;
                delete[] cfl_tree;
                delete[] dpiso_order;
                delete[] dpiso_tree;
                continue;
            }

            // ---- tables (:1247-1271)
            LOG() << "-----" << std::endl;
            LOG() << "Build indices..." << std::endl;
            start = std::chrono::steady_clock::now();
            Edges*** edge_matrix = new Edges**[n];
            for (ui i = 0; i < n; ++i) edge_matrix[i] = new Edges*[n];
            BuildTable::buildTables(data_graph, query_graph, candidates, candidates_count, edge_matrix);
            const int64_t build_table_time_in_ns = ns_since(start);
            const size_t memory_cost_in_bytes =
                BuildTable::computeMemoryCostInBytes(query_graph, candidates_count, edge_matrix);

            // ---- arguments, units, plans (:1272-1589)
            LOG() << "-----" << std::endl;
            LOG() << "Generate a matching order..." << std::endl;
            start = std::chrono::steady_clock::now();
            TaskSlot::g_args = new GLOBALARGS(data_graph, query_graph, edge_matrix, candidates, candidates_count,
                                              output_limit, time_limit);
            if (num_threads == 1 && qorc != "Q") TaskSlot::g_args->splitMode = SPLITMODE::NOSPLIT;
            else if (qorc == "Q") TaskSlot::g_args->splitMode = SPLITMODE::Q;
            else if (qorc == "C") TaskSlot::g_args->splitMode = SPLITMODE::C;
            else TaskSlot::g_args->splitMode = SPLITMODE::NOSPLIT;
            if (order_type == "DPiso") {
                TaskSlot::g_args->tree = dpiso_tree;
                TaskSlot::g_args->tree_order = dpiso_order;
            } else if (order_type == "CFL") {
                TaskSlot::g_args->tree = cfl_tree;
                TaskSlot::g_args->tree_order = cfl_order;
            } else if (order_type != "GQL") {
                LOG() << "The specified order type '" << order_type << "' is not supported." << std::endl;
                return 1;
            }
            TaskSlot::g_args->needCandidates = true;
            TaskSlot::g_args->needEdgeMatrix = true;
            TaskSlot::g_args->engine_type = engine_type;

            const auto split_start = std::chrono::steady_clock::now();
            const SplitPattern pattern = qpattern == "TWINTWIG" ? SplitPattern::TWINTWIG
                                         : qpattern == "SEED"   ? SplitPattern::SEED
                                                                : SplitPattern::STAR;
            SplitGraph* multigraphs = new SplitGraph(data_graph, query_graph, JoinParadigm::LEFTDEEP,
                                                     qorc == "Q" ? pattern : SplitPattern::SEED,
                                                     TaskSlot::g_args->splitMode, time_limit);
            const int64_t split_query_time_in_ns = ns_since(split_start);
            const int units = multigraphs->getSubgraphCount();
            UnitArgs** unitArgsVec = new UnitArgs*[units];
            for (int id_unit = 0; id_unit < units; ++id_unit) {
                unitArgsVec[id_unit] = new UnitArgs(multigraphs->getSubgraphs()[id_unit], output_limit);
                unitArgsVec[id_unit]->mapUidCandidates(TaskSlot::g_args);
                unitArgsVec[id_unit]->mapUidEdgeMatrix(TaskSlot::g_args);
            }
            TaskSlot::g_args->unitArgsVec = unitArgsVec;
            TaskSlot::g_args->count_unit = units;
            for (int id_unit = 0; id_unit < units; ++id_unit) {
                UnitArgs* u_args = unitArgsVec[id_unit];
                Graph* q_graph = u_args->query_graph->subgraph;
                if (order_type == "DPiso") {
                    GenerateFilteringPlan::generateDPisoFilterPlan(data_graph, u_args);
                    GenerateQueryPlan::generateDPisoQueryPlan(q_graph, u_args);
                    EvaluateQuery::generateBN(u_args);
                } else if (order_type == "GQL") {
                    GenerateQueryPlan::generateGQLQueryPlan(data_graph, q_graph, u_args);
                    EvaluateQuery::generateBN(u_args);
                } else {
                    GenerateFilteringPlan::generateCFLFilterPlan(data_graph, u_args);
                    GenerateQueryPlan::generateCFLQueryPlan(q_graph, u_args);
                    EvaluateQuery::generateBNWithPivot(u_args);
                }
            }
            const int64_t generate_query_plan_time_in_ns = ns_since(start);

            // ---- enumeration (:1590-1716)
            LOG() << "-----" << std::endl;
            LOG() << "Enumerate..." << std::endl;
            start = std::chrono::steady_clock::now();
            TaskPool* taskPool = new TaskPool(num_threads, (uint64_t)num_threads * 16);
            if (engine_type == "LFTJ") {
                EvaluateQuery::enum_method = [taskPool](TaskSlot& task, int task_id) {
                    EvaluateQuery::LFTJ(taskPool, task, task_id);
                };
            } else if (engine_type == "GQL") {
                EvaluateQuery::enum_method = [taskPool](TaskSlot& task, int task_id) {
                    EvaluateQuery::exploreGraphQLStyle(taskPool, task, task_id);
                };
            } else {
                EvaluateQuery::enum_method = [taskPool](TaskSlot& task, int task_id) {
                    EvaluateQuery::exploreGraph(taskPool, task, task_id);
                };
            }
            EvaluateQuery::ParallelExecute(taskPool, multigraphs, unitArgsVec);
            int64_t enumeration_time_in_ns = ns_since(start);

            JoinCollect* jc = new JoinCollect(data_graph, query_graph, multigraphs->jseq, units, time_limit, taskPool);
            for (int id_unit = 0; id_unit < units; ++id_unit)
                jc->partial_matches[id_unit] = unitArgsVec[id_unit]->unit_partial_match;
            PartialMatch* res = jc->sequence_collect();
            int64_t join_time_in_ns = 0;
            if (res && SPLITMODE::Q == TaskSlot::g_args->splitMode) {
                // the count is reported as min(cap, total)
                TaskSlot::g_args->embedding_count = (int64_t)std::min<size_t>(output_limit, res->size);
                TaskSlot::g_args->call_count = jc->call_count;
                enumeration_time_in_ns = (int64_t)(jc->enumerate_seconds * 1e9);
                join_time_in_ns = (int64_t)(jc->join_seconds * 1e9);
            }
            if (jc->overtime) TaskSlot::g_args->overtime = true;
            delete jc;
            delete res;

            GLOBALARGS* g = TaskSlot::g_args;
            g->load_graphs_time_in_ns = load_graphs_time_in_ns;
            g->filter_vertices_time_in_ns = filter_vertices_time_in_ns;
            g->split_query_time_in_ns = split_query_time_in_ns;
            g->build_table_time_in_ns = build_table_time_in_ns;
            g->generate_query_plan_time_in_ns = generate_query_plan_time_in_ns;
            g->enumeration_time_in_ns = enumeration_time_in_ns;
            g->join_time_in_ns = join_time_in_ns;
            g->preprocessing_time_in_ns =
                filter_vertices_time_in_ns + build_table_time_in_ns + generate_query_plan_time_in_ns;
            g->total_time_in_ns = g->preprocessing_time_in_ns + enumeration_time_in_ns + join_time_in_ns;
            g->memory_cost_in_bytes = (int64_t)memory_cost_in_bytes;
            g->logGlobalStatistics();

            for (int id_unit = 0; id_unit < units; ++id_unit) delete unitArgsVec[id_unit];
            delete[] unitArgsVec;
            delete multigraphs;
            delete taskPool;
            delete TaskSlot::g_args;   // frees candidates and tables (Args.cpp:366-400)
            TaskSlot::g_args = nullptr;
            delete[] cfl_order;
            delete[] cfl_tree;
            delete[] dpiso_order;
            delete[] dpiso_tree;
        }
    }
    delete query_graph;
    delete data_graph;
    return 0;
}
