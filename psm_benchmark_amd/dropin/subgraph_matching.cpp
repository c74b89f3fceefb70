// subgraph_matching.cpp — command-line driver with the reference's options
// and result log (matching/StudyPerformance.cpp:951-1831; logGlobalStatistics
// Args.cpp:517-535), running on the GPU through the drop-in classes: the
// undecomposed path (-QorCandi C / NOSPLIT) and the query-split path
// (-QorCandi Q with -Qpattern STAR, TWINTWIG or SEED). Its log can be read by
// the reference's harness and parsers.
//
//   SubgraphMatching.out -d data.txt -q query.txt -num MAX -time_limit 60
//       -BackMethods GQL_GQL_GQL,CFL_CFL_EXPLORE -QorCandi C -OutputFile log.txt

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "psm_dropin.h"

namespace {

std::string g_log_path;

struct Line {
    std::ostringstream oss;
    Line() { oss << "[Thread 1] "; }
    ~Line() {
        if (g_log_path.empty()) {
            std::cout << oss.str();
        } else {
            std::ofstream f(g_log_path, std::ios::app);
            f << oss.str();
        }
    }
    template <class T>
    Line& operator<<(const T& v) { oss << v; return *this; }
    Line& operator<<(std::ostream& (*m)(std::ostream&)) { oss << m; return *this; }
};

double sec(int64_t ns) { return ns / 1e9; }

std::vector<std::string> split(const std::string& s, char c) {
    std::vector<std::string> out;
    std::stringstream ss(s);
    std::string t;
    while (std::getline(ss, t, c)) out.push_back(t);
    return out;
}

}  // namespace

int main(int argc, char** argv) {
    MatchingCommand cmd(argc, argv);
    const bool qmode = cmd.getQorCandiType() == "Q";
    if (cmd.getQorCandiType() != "C" && cmd.getQorCandiType() != "NOSPLIT" && !qmode) {
        std::cerr << "Error: -QorCandi must be C, NOSPLIT or Q." << std::endl;
        return 1;
    }
    const std::string qpattern = cmd.getQpatternType();
    if (qmode && qpattern != "STAR" && qpattern != "TWINTWIG" && qpattern != "SEED") {
        std::cerr << "Error: -Qpattern " << qpattern << " is not supported (STAR, TWINTWIG or SEED)." << std::endl;
        return 1;
    }
    g_log_path = cmd.getOutputFile();
    auto now = [] { return std::chrono::steady_clock::now(); };
    auto ns = [](auto a, auto b) {
        return (int64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(b - a).count();
    };
    auto t0 = now();
    Graph data(true), query(true);
    data.loadGraphFromFile(cmd.getDataGraphFilePath());
    query.loadGraphFromFile(cmd.getQueryGraphFilePath());
    query.buildCoreTable();
    const int64_t load_ns = ns(t0, now());

    double time_limit = std::stod(cmd.getTimeLimit());
    size_t limit = std::numeric_limits<int64_t>::max();
    if (cmd.getMaximumEmbeddingNum() != "MAX") limit = std::stoull(cmd.getMaximumEmbeddingNum());
    const ui n = query.getVerticesCount();

    for (const std::string& threads : split(cmd.getThreadNumbers(), ',')) {
        for (const std::string& method : split(cmd.getBackMethodTypes(), ',')) {
            auto parts = split(method, '_');
            if (parts.size() != 3) { std::cerr << "bad method " << method << std::endl; return 1; }
            const std::string& f = parts[0];
            const std::string& o = parts[1];
            const std::string& e = parts[2];
            Line() << "Start the query" << std::endl;
            Line() << "Input Query Graph File: " << cmd.getQueryGraphFilePath() << std::endl;
            Line() << "Input Data Graph File: " << cmd.getDataGraphFilePath() << std::endl;
            Line() << "Maximum Embedding Number: " << cmd.getMaximumEmbeddingNum() << std::endl;
            Line() << "Time Limit: " << cmd.getTimeLimit() << std::endl;
            Line() << "Num Threads: " << threads << std::endl;
            Line() << "Split Type: " << cmd.getSplitType() << std::endl;
            if (qmode) Line() << "QuerySplit Pattern: " << qpattern << std::endl;
            Line() << "Schedule Type: " << cmd.getScheduleType() << std::endl;
            Line() << "Filter Type: " << f << std::endl;
            Line() << "Order Type: " << o << std::endl;
            Line() << "Engine Type: " << e << std::endl;
            Line() << "Output Limit: " << (int64_t)std::min<size_t>(limit, std::numeric_limits<int64_t>::max())
                   << std::endl;

            auto t = now();
            ui** cand = nullptr;
            ui* cnt = nullptr;
            if (f == "LDF") FilterVertices::LDFFilter(&data, &query, cand, cnt);
            else if (f == "NLF") FilterVertices::NLFFilter(&data, &query, cand, cnt);
            else if (f == "GQL") FilterVertices::GQLFilter(&data, &query, cand, cnt);
            else if (f == "CFL") FilterVertices::CFLFilter(&data, &query, cand, cnt);
            else if (f == "DPiso") FilterVertices::DPisoFilter(&data, &query, cand, cnt);
            else { Line() << "The specified filter type '" << f << "' is not supported." << std::endl; return 1; }
            FilterVertices::sortCandidates(cand, cnt, n);
            const int64_t filter_ns = ns(t, now());
            Line() << "-----" << std::endl;
            Line() << "Build indices..." << std::endl;
            t = now();
            std::vector<Edges**> em(n);
            for (ui i = 0; i < n; ++i) em[i] = new Edges*[n];
            BuildTable::buildTables(&data, &query, cand, cnt, em.data());
            const int64_t table_ns = ns(t, now());
            const size_t mem = BuildTable::computeMemoryCostInBytes(&query, cnt, em.data());
            Line() << "-----" << std::endl;
            Line() << "Generate a matching order..." << std::endl;
            t = now();
            std::vector<ui> order(n), pivot(n);
            if (o == "GQL") GenerateQueryPlan::generateGQLQueryPlan(&data, &query, cnt, order.data(), pivot.data());
            else if (o == "CFL") GenerateQueryPlan::generateCFLQueryPlan(&data, &query, cnt, order.data(), pivot.data());
            else if (o == "DPiso") GenerateQueryPlan::generateDPisoQueryPlan(&data, &query, cnt, order.data(), pivot.data());
            else { Line() << "The specified order type '" << o << "' is not supported." << std::endl; return 1; }
            if (e != "GQL" && e != "EXPLORE" && e != "LFTJ") {
                Line() << "The specified engine type '" << e << "' is not supported." << std::endl;
                return 1;
            }
            const int64_t plan_ns = ns(t, now());
            Line() << "-----" << std::endl;
            Line() << "Enumerate..." << std::endl;
            t = now();
            size_t calls = 0, emb = 0;
            bool overtime = false;
            int64_t enum_ns = 0, join_ns = 0;
            if (qmode) {
                // units enumerated and joined on the GPU; the count is reported
                // as min(cap, total)
                double es = 0, js = 0;
                emb = std::min(limit, JoinCollect::sequence_collect(&query, qpattern, calls, overtime, es, js,
                                                                     time_limit));
                enum_ns = (int64_t)(es * 1e9);
                join_ns = (int64_t)(js * 1e9);
            } else {
                EvaluateQuery::engine_type = e;
                emb = EvaluateQuery::evaluate(&data, &query, limit, calls, overtime, time_limit);
                enum_ns = ns(t, now());
            }
            const int64_t pre_ns = filter_ns + table_ns + plan_ns;
            Line() << "Load graphs time (seconds): " << sec(load_ns) << std::endl;
            Line() << "Build table time (seconds): " << sec(table_ns) << std::endl;
            Line() << "Filter vertices time (seconds): " << sec(filter_ns) << std::endl;
            Line() << "Split Qeury Pattern Time: " << 0 << std::endl;
            Line() << "Generate query plan time (seconds): " << sec(plan_ns) << std::endl;
            Line() << "Enumerate time (seconds): " << sec(enum_ns) << std::endl;
            Line() << "Preprocessing time (seconds): " << sec(pre_ns) << std::endl;
            Line() << "Join time (seconds): " << sec(join_ns) << std::endl;
            Line() << "Total time (seconds): " << sec(pre_ns + enum_ns + join_ns) << std::endl;
            Line() << "Memory cost (MB): " << mem / (double)(1024 * 1024) << std::endl;
            Line() << "#Embeddings: " << emb << std::endl;
            Line() << "Call Count: " << calls << std::endl;
            Line() << "Per Call Time (nanoseconds): " << enum_ns / (calls == 0 ? 1 : (int64_t)calls) << std::endl;
            Line() << "Overtime: " << (overtime ? 1 : 0) << std::endl;
            Line() << "End the query" << std::endl;
            for (ui i = 0; i < n; ++i) {
                for (ui j = 0; j < n; ++j) delete em[i][j];
                delete[] em[i];
                delete[] cand[i];
            }
            delete[] cand;
            delete[] cnt;
        }
    }
    return 0;
}
