/*
 * psm_abi.h — C ABI of the MI355X subgraph matcher (libpsm.so).
 *
 * Every entry point replaces a call surface of the reference matcher
 * (scorpiontornado/PSM_Benchmark); the comment on each names it. The ABI is
 * plain C: opaque handles, pointers and sizes, no C++ or torch types.
 *
 * Conventions
 *   - Every function returns int: 0 = PSM_OK, > 0 a PSM status below,
 *     < 0 a negated hipError_t. Nothing throws or aborts across the ABI.
 *   - Host inputs are borrowed for the duration of the call. Device objects
 *     are owned by their handle until the matching *_destroy.
 *   - One psm_ctx owns one HIP stream; calls on one context are serialised
 *     by the caller, calls on different contexts may run concurrently.
 *   - The text of the last error of a context is in psm_last_error().
 *   - A context must outlive the graphs and queries created with it.
 */
#ifndef PSM_ABI_H
#define PSM_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define PSM_OK 0
#define PSM_EINVAL 1      /* bad argument (size, method name, query too big) */
#define PSM_EEMPTY 2      /* a candidate set is empty: the reference's filters return false */
#define PSM_ENOMEM 3      /* device memory exhausted or an admission budget exceeded */
#define PSM_ETIMEOUT 4    /* time limit hit (reference: Overtime: 1) */
#define PSM_EOVERFLOW 5   /* a 64-bit counter would overflow */
#define PSM_ENOTFOUND 6   /* generator could not produce the requested object */
#define PSM_ESTATE 7      /* call out of order (e.g. enumerate before build tables) */

/* filter methods: FilterVertices::{LDF,NLF,GQL,CFL,DPiso}Filter
 * (matching/FilterVertices.h:9-17) */
#define PSM_FILTER_LDF 0
#define PSM_FILTER_NLF 1
#define PSM_FILTER_GQL 2
#define PSM_FILTER_CFL 3
#define PSM_FILTER_DPISO 4

/* order methods: GenerateQueryPlan::generate{GQL,CFL,DPiso}QueryPlan
 * (matching/GenerateQueryPlan.h:32,49,55). The DPiso order keeps every
 * earlier neighbour as an operand (EvaluateQuery::generateBN), the plan the
 * reference's LFTJ and GQL engines run. */
#define PSM_ORDER_GQL 0
#define PSM_ORDER_CFL 1
#define PSM_ORDER_DPISO 2

/* the largest query the device path accepts (config.h:9 MAXIMUM_QUERY_GRAPH_SIZE) */
#define PSM_MAX_QUERY_VERTICES 64

typedef struct psm_ctx psm_ctx;
typedef struct psm_graph psm_graph;
typedef struct psm_query psm_query;

/* Enumeration statistics. `count` is min(cap, total) when capped. `calls` is
 * the reference's "Call Count" of the GQL and EXPLORE engines (one per
 * assignment, EvaluateQuery.cpp:1009) and is exact only for uncapped runs. Times are device-side (HIP events). */
typedef struct psm_enum_stats {
    uint64_t count;
    uint64_t calls;
    uint32_t capped;          /* 1 when the cap was reached */
    uint32_t tasks;           /* device tasks executed */
    double kernel_ms;         /* enumeration kernels */
    double expand_ms;         /* task pre-expansion kernels */
    uint64_t probe_bytes;     /* bytes the enumeration kernels touched (base elements,
                                 search probes, rows located) */
    uint64_t algo_bytes;      /* algorithmic bytes of the search: for every search-tree
                                 node below the root, 4 per entry of each operand row
                                 (table row of an earlier query neighbour) plus 4 per
                                 survivor written unless the survivors are leaves;
                                 schedule-independent when the count is uncapped */
    uint64_t visited_rejects; /* extensions present in every operand row whose data
                                 vertex is already in the partial embedding: the LFTJ
                                 engine's Call Count is calls + visited_rejects
                                 (EvaluateQuery.cpp:278-295); exact when uncapped */
} psm_enum_stats;

/* ---- context -------------------------------------------------------- */
int psm_ctx_create(int device, psm_ctx** out);
int psm_ctx_destroy(psm_ctx* ctx);
const char* psm_last_error(const psm_ctx* ctx);

/* Per-kernel-class statistics of a context, accumulated since the last
 * reset: out3 = {launches, total milliseconds (HIP events on the context's
 * stream), algorithmic bytes counted by the kernels}. Classes: 0 GQL
 * refinement, 1 table count, 2 table fill, 3 enumeration DFS, 4 task
 * pre-expansion, 5 compaction, 6 bitmaps. Instrumentation of this library;
 * the reference logs phase times only (Args.cpp:517-535). */
int psm_kernel_stats(psm_ctx* ctx, int kind, double* out3);
int psm_kernel_stats_reset(psm_ctx* ctx);

/* The reference's -time_limit (StudyPerformance.cpp:1134-1148): counted from
 * the start of filtering, as the reference sets its deadline before
 * FilterVertices, so filter, tables, plan and enumeration share it;
 * enumeration stops when it passes and returns PSM_ETIMEOUT with partial
 * counts (reference: "Overtime: 1"). Default 60 s; 0 disables the limit. */
int psm_set_time_limit(psm_ctx* ctx, double seconds);

/* Enumeration launch shape, set by the autotuner (psm_benchmark_amd/autotune)
 * or by hand; results never depend on it. Keys:
 *   PSM_TUNE_DFS_WAVES_PER_CU       wavefronts per CU of the persistent DFS
 *                                   grid, 1..32 (default 1)
 *   PSM_TUNE_EXPAND_TASKS_PER_LANE  pre-expansion target: tasks per DFS
 *                                   worker (lane or wavefront) before the
 *                                   persistent kernel starts (2)
 *   PSM_TUNE_DFS_STRATEGY           PSM_DFS_LANE: one search per lane;
 *                                   PSM_DFS_WAVE: one search per wavefront
 *                                   with chunked cooperative intersection
 *   PSM_TUNE_QUEUE_SLOTS            free slots of the task ring beyond the
 *                                   pre-expanded tasks (0: automatic, the
 *                                   larger of 2^18 and 4 per DFS lane); the
 *                                   ring reuses slots, so this bounds only
 *                                   the tasks queued at one time
 *   PSM_TUNE_CLAIM_DIVISOR          pre-expanded tasks per DFS lane above
 *                                   which lanes whose previous tasks were
 *                                   short take several (up to 8) per claim,
 *                                   1..1024 (default 1): fewer claims for
 *                                   millions of small tasks, one task per
 *                                   claim for heavy ones
 *   PSM_TUNE_DFS_BLOCK_LANES        lanes per workgroup of the lane DFS: 64,
 *                                   128 or 256, or 0 for 64 (default 0);
 *                                   halved while the block's LDS state
 *                                   would not fit. Blocks of one wavefront
 *                                   spread a small grid over every CU
 *   PSM_TUNE_EXPAND_DEPTH           deepest level the BFS pre-expansion
 *                                   assigns before the DFS takes over (the
 *                                   DFS stack's starting depth), 1..31
 *                                   (default 31; the leaf level is never
 *                                   assigned there)
 *   PSM_TUNE_INTERSECT              set intersection of the DFS kernels:
 *                                   PSM_INTERSECT_ADAPTIVE (default: the
 *                                   lane kernel advances galloping cursors
 *                                   over up to four operand rows, the
 *                                   wavefront kernel tests rows of 1024+
 *                                   entries cooperatively through an
 *                                   LDS-staged window) or
 *                                   PSM_INTERSECT_BSEARCH (every operand row
 *                                   by binary search, per lane)
 *   PSM_TUNE_SERVE_RESIDENT         servers created on this context admit
 *                                   prepared batches into one resident
 *                                   enumeration launch (1, default) or run
 *                                   one launch per batch (0)
 * No counterpart in the reference (its only knob is -threadnums). */
#define PSM_TUNE_DFS_WAVES_PER_CU 0
#define PSM_TUNE_EXPAND_TASKS_PER_LANE 1
#define PSM_TUNE_DFS_STRATEGY 2
#define PSM_TUNE_QUEUE_SLOTS 3
#define PSM_TUNE_CLAIM_DIVISOR 4
#define PSM_TUNE_DFS_BLOCK_LANES 5
#define PSM_TUNE_EXPAND_DEPTH 6
#define PSM_TUNE_INTERSECT 7
#define PSM_TUNE_SERVE_RESIDENT 8
#define PSM_DFS_LANE 0
#define PSM_DFS_WAVE 1
#define PSM_INTERSECT_ADAPTIVE 0
#define PSM_INTERSECT_BSEARCH 1
int psm_set_tuning(psm_ctx* ctx, int key, double value);
int psm_get_tuning(psm_ctx* ctx, int key, double* value);

/* ---- data graph: class Graph (graph/graph.h:21-237) ------------------
 * Builds on the host, from a CSR data graph (adjacency sorted ascending per
 * vertex), the label reverse index (graph.cpp:15-33), the
 * neighbour-label-frequency runs (graph.cpp:36-52) and the label-grouped
 * adjacency, and uploads the graph with them to HBM, where it stays resident
 * until psm_graph_destroy. */
int psm_graph_upload(psm_ctx* ctx, uint32_t num_vertices, const uint32_t* offsets,
                     const uint32_t* neighbors, const uint32_t* labels, psm_graph** out);
int psm_graph_destroy(psm_graph* g);
/* |V|, |E|, labels count (max label + 1), max degree, max label frequency */
int psm_graph_info(const psm_graph* g, uint64_t* info5);

/* ---- one query run (plain path) ---------------------------------------
 * A psm_query holds the device state the reference keeps in GLOBALARGS and
 * UnitArgs (matching/Args.h:30-181): candidates, candidate-edge tables and
 * the matching plan of one query graph against one data graph. */
int psm_query_create(psm_ctx* ctx, const psm_graph* g, uint32_t n, const uint32_t* q_offsets,
                     const uint32_t* q_neighbors, const uint32_t* q_labels, psm_query** out);
int psm_query_destroy(psm_query* q);

/* FilterVertices::{LDF,NLF,GQL,CFL}Filter + sortCandidates
 * (FilterVertices.cpp:19-250,1012-1016). Candidate sets come out sorted.
 * Returns PSM_EEMPTY where the reference's filter returns false. */
int psm_filter(psm_query* q, int method);
/* candidates_count (n entries) and a host copy of C(u) (candidate_count[u]
 * entries): the ui** / ui* outputs of the reference's filters. */
int psm_query_candidate_counts(const psm_query* q, uint32_t* counts);
int psm_query_candidates(const psm_query* q, uint32_t u, uint32_t* out);

/* BuildTable::buildTables (BuildTable.cpp:9-151): one table per ordered query
 * edge (a,b); row j lists ascending positions k in C(b) of the neighbours of
 * C(a)[j]. */
int psm_build_tables(psm_query* q);
/* Host copy of table (a,b): off has candidate_count[a]+1 entries, idx has
 * *nnz entries (pass idx = NULL to query nnz). PSM_EINVAL if a-b is no edge. */
int psm_query_table(const psm_query* q, uint32_t a, uint32_t b, uint32_t* off, uint32_t* idx,
                    uint64_t* nnz);
/* BuildTable::computeMemoryCostInBytes (BuildTable.cpp:300-326), in bytes. */
int psm_memory_cost_bytes(const psm_query* q, uint64_t* bytes);

/* GenerateQueryPlan::generateGQLQueryPlan / generateCFLQueryPlan
 * (GenerateQueryPlan.cpp:9-83,270-471) followed by EvaluateQuery::generateBN
 * / generateBNWithPivot (EvaluateQuery.cpp:1594-1666). order/pivot receive n
 * entries each (either may be NULL). */
int psm_generate_plan(psm_query* q, int order_method, uint32_t* order, uint32_t* pivot);

/* A caller-supplied matching order (and pivots: pivot[i] an earlier query
 * neighbour of order[i], i >= 1; NULL for none), as the reference's engines
 * take UnitArgs::order / pivots from whichever plan generator ran
 * (EvaluateQuery.cpp:1594-1666 derives the backward neighbours from them).
 * Extensions at each depth are the intersection of the table rows of every
 * earlier neighbour, so the pivot never changes counts. Needs
 * psm_build_tables; PSM_EINVAL unless order is a permutation of 0..n-1. */
int psm_set_plan(psm_query* q, const uint32_t* order, const uint32_t* pivot);

/* The BFS tree a CFL or DPiso step used (GraphOperations::bfsTraversal,
 * graphoperations.cpp:156-195): PSM_TREE_FILTER the one the last CFL / DPiso
 * filter built on the query graph (FilterVertices::CFLFilter's order and
 * tree, FilterVertices.h:15-18), PSM_TREE_PLAN the one the last CFL / DPiso
 * order built on the query unit (GenerateFilteringPlan::generateCFLFilterPlan
 * (data, UnitArgs*), GenerateFilteringPlan.h:15). n entries each: BFS order,
 * parent (the root's is 0, as in the reference), level. PSM_ESTATE when
 * there is none. */
#define PSM_TREE_FILTER 0
#define PSM_TREE_PLAN 1
int psm_query_bfs_tree(const psm_query* q, int which, uint32_t* order, uint32_t* parent, uint32_t* level);

/* EvaluateQuery::ParallelExecute with the GQL / EXPLORE engines
 * (EvaluateQuery.cpp:26-112,980-1126): counts embeddings, at most `cap`
 * (0 = no cap). Returns PSM_ETIMEOUT with the partial count when the time
 * limit (psm_set_time_limit) stops it. */
int psm_evaluate(psm_query* q, uint64_t cap, psm_enum_stats* stats);

/* psm_evaluate without device memory allocation: no counterpart in the
 * reference. For callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * counts, call counts, visited rejects and search bytes as psm_evaluate
 * (calls exact when uncapped); calls no hipMalloc / hipFree. Its plan,
 * operands, counters and task ring live in a fixed device scratch of about
 * 11 MB per GPU (allocated with the code object, held until the process
 * ends): at most 2048 intersection operands (every query within
 * PSM_MAX_QUERY_VERTICES fits; PSM_EINVAL beyond), 4096 root tasks and 65536
 * ring slots. Every call on one GPU uses that GPU's one scratch: calls that
 * overlap on one GPU (from any context) race and must be serialised by the
 * caller. Needs psm_filter, psm_build_tables and psm_generate_plan first. */
int psm_evaluate_noalloc(psm_query* q, uint64_t cap, psm_enum_stats* stats);

/* Query decomposition + join: JoinCollect::sequence_collect over two units
 * (JoinCollect.cpp:60-153,237-315; NoPartitionHashJoin hashjoin.cpp:337-781).
 * Unit A is the first ceil(n/2) vertices of the generated plan's order, unit
 * B every query edge leaving A; both are enumerated on the GPU into rows and
 * joined by a device hash join on the shared vertices with the injectivity
 * check across units. count equals the plain count (no cap). Needs
 * psm_generate_plan. max_rows bounds each unit (0: 64M). */
typedef struct psm_join_stats {
    uint64_t count;
    uint64_t rows_a, rows_b;
    uint32_t keys;
    double enum_ms, join_ms;
} psm_join_stats;

int psm_join_count(psm_query* q, uint64_t max_rows, psm_join_stats* stats);

/* Query split + a join sequence: the STAR / TWINTWIG / SEED splits
 * (SplitGraph.cpp:874-980) as greedy covers (vertices in id order take their
 * query edges not yet taken: one star each, or twin twigs of at most two
 * edges, or twin twigs closed into triangles where the query has the edge)
 * and JoinCollect::sequence_collect (JoinCollect.cpp:60-153) over the units'
 * device-enumerated rows: smallest unit first, then the connected unit with
 * the fewest rows, hash-joined on the shared vertices with the injectivity
 * check across sides. count equals the plain count (no cap); the units are
 * this library's own covers, not the reference's decomposition search.
 * Needs psm_build_tables. max_rows bounds every unit and intermediate result
 * (0: 64M); PSM_ENOMEM beyond it. */
#define PSM_SPLIT_STAR 0
#define PSM_SPLIT_TWINTWIG 1
#define PSM_SPLIT_SEED 2

typedef struct psm_split_join_stats {
    uint64_t count;
    uint32_t units;
    uint32_t pad;
    uint64_t unit_rows;       /* rows of all units */
    uint64_t max_inter_rows;  /* the largest intermediate result */
    double enum_ms, join_ms;
    uint64_t calls;           /* assignments made enumerating the units */
} psm_split_join_stats;

int psm_split_join_count(psm_query* q, int pattern, uint64_t max_rows, psm_split_join_stats* st);

/* Convenience: filter + tables + plan + evaluate in one call, the sequence of
 * split_C_test (StudyPerformance.cpp:1164-1716). */
int psm_match(psm_query* q, int filter, int order, uint64_t cap, psm_enum_stats* stats);

/* ---- batched form -------------------------------------------------------
 * Many queries against one resident data graph, each stage launched once for
 * the whole batch and enumeration in one persistent kernel. Queries are given
 * as concatenated CSR arrays: query i has n[i] vertices, its offsets are
 * q_offsets[off_base[i] .. off_base[i]+n[i]] (relative to its own neighbours,
 * which start at nbr_base[i]), labels at lab_base[i]. counts/calls receive
 * one entry per query (calls exact only for queries that did not hit cap);
 * status[i] is PSM_OK, PSM_EEMPTY (a filter emptied a candidate set) or
 * PSM_ETIMEOUT (partial count: the time limit stopped it).
 * No counterpart in the reference, which matches one query per process. */
typedef struct psm_batch_stats {
    double filter_ms, table_ms, plan_ms, enum_ms, total_ms;
    uint64_t tasks;
    uint64_t probe_bytes;
    /* DFS lane utilisation: loop iterations in which a lane held a task, and
     * all loop iterations of all lanes */
    uint64_t dfs_lane_steps, dfs_lane_iters;
    uint64_t algo_bytes;      /* as psm_enum_stats.algo_bytes, summed over the batch */
} psm_batch_stats;

int psm_batch_match(psm_ctx* ctx, const psm_graph* g, uint32_t num_queries, const uint32_t* n,
                    const uint64_t* off_base, const uint64_t* nbr_base, const uint64_t* lab_base,
                    const uint32_t* q_offsets, const uint32_t* q_neighbors, const uint32_t* q_labels,
                    int filter, int order, uint64_t cap, uint64_t* counts, uint64_t* calls,
                    uint32_t* status, psm_batch_stats* stats);

/* The batched form stage by stage, so that the enumeration can run again
 * (different cap, limit or launch shape) without filtering again:
 * psm_batch_create takes the queries as psm_batch_match does;
 * psm_batch_prepare runs filter, tables and plans; psm_batch_evaluate
 * enumerates. enum_time_limit_s < 0 applies the context's time limit counted
 * from the start of psm_batch_prepare (the reference's deadline), >= 0 a
 * limit of that many seconds counted from this call (0: none). psm_batch_match
 * is create + prepare + evaluate(-1) + destroy. */
typedef struct psm_batch psm_batch;
int psm_batch_create(psm_ctx* ctx, const psm_graph* g, uint32_t num_queries, const uint32_t* n,
                     const uint64_t* off_base, const uint64_t* nbr_base, const uint64_t* lab_base,
                     const uint32_t* q_offsets, const uint32_t* q_neighbors, const uint32_t* q_labels,
                     psm_batch** out);
int psm_batch_prepare(psm_batch* b, int filter, int order);
int psm_batch_evaluate(psm_batch* b, uint64_t cap, double enum_time_limit_s, uint64_t* counts,
                       uint64_t* calls, uint32_t* status, psm_batch_stats* stats);
/* Per-query latency of the last psm_batch_evaluate: milliseconds from
 * psm_batch_create until the query's count was final (its cap reached or its
 * last enumeration task ended, read from the device clock), one entry per
 * query; a query without enumeration tasks reports the end of the evaluate
 * call. No counterpart in the reference. */
int psm_batch_latencies(const psm_batch* b, double* ms);
int psm_batch_destroy(psm_batch* b);

/* ---- serving form ----------------------------------------------------------
 * Continuous batching over a resident data graph: `workers` threads, each with
 * its own HIP stream, admit the oldest pending requests as one batch while the
 * batch holds fewer than max_batch requests and the device footprint of
 * running batches stays within mem_budget (0: half the free device memory;
 * a request counts a bound of its bitmaps, membership words and candidate
 * sets until its tables are built, then their real size), and prepare it
 * (filter, tables, plan, task pre-expansion).
 * A prepared batch is published into the resident enumeration launch: one
 * persistent kernel that keeps running while batches join it, so each
 * batch's search starts when its preparation ends, beside earlier batches'
 * searches. Each request's result is returned once its own count is final
 * (its cap reached, its last task ended, or its time limit passed: then
 * PSM_ETIMEOUT with the partial count). The time limit counts from the start
 * of the batch's preparation. The launch ends when it has had nothing to run
 * for a short while or its batches hold half the memory budget; the next
 * batch starts a new one. With PSM_TUNE_SERVE_RESIDENT 0 on the creating
 * context (and for queries above 32 vertices) every batch runs its own
 * enumeration launch instead, batches of different workers overlapping on
 * the device. `cap` and the creating context's launch shape
 * (psm_set_tuning) apply to both. Times in psm_result are milliseconds since
 * server creation. No counterpart in the reference, which runs one query per
 * process. */
typedef struct psm_server psm_server;
typedef struct psm_result {
    uint64_t request_id;
    uint64_t count;
    uint64_t calls;
    int32_t status;           /* PSM_OK, PSM_EEMPTY, PSM_ETIMEOUT or an error */
    uint32_t batch_size;
    double submit_ms, start_ms, done_ms;
} psm_result;

int psm_server_create(psm_ctx* ctx, const psm_graph* g, uint32_t workers, uint32_t max_batch,
                      uint64_t mem_budget, int filter, int order, uint64_t cap, double time_limit_s,
                      psm_server** out);
int psm_server_submit(psm_server* s, uint64_t request_id, uint32_t n, const uint32_t* q_offsets,
                      const uint32_t* q_neighbors, const uint32_t* q_labels);
/* Moves up to max_results finished results to out; waits up to wait_ms for
 * the first one when none is ready. */
int psm_server_poll(psm_server* s, psm_result* out, uint32_t max_results, uint32_t* n_out,
                    double wait_ms);
/* submitted, completed, batches formed, peak admitted footprint, pending */
int psm_server_stats(psm_server* s, uint64_t* out5);
int psm_server_destroy(psm_server* s);

/* ---- graph files (host) -------------------------------------------------
 * Graph::loadGraphFromFile (graph/graph.cpp:92-172): the "t V E" /
 * "v id label degree" / "e u v" text format, memory-mapped and parsed by
 * `threads` host threads (0: every hardware thread, at most 64) into a CSR
 * with ascending adjacency. Adjacency comes from the e lines and every v
 * line's degree field must agree with them (the reference places neighbours
 * by the degree fields); PSM_EINVAL on a malformed line, an id out of range,
 * a vertex listed twice or counts other than the header's, PSM_ENOTFOUND
 * when the file cannot be opened; psm_text_graph_last_error() says which
 * (per calling thread). The arrays belong to the handle. */
typedef struct psm_text_graph psm_text_graph;
int psm_text_graph_read(const char* path, uint32_t threads, psm_text_graph** out);
int psm_text_graph_info(const psm_text_graph* g, uint32_t* num_vertices, uint64_t* num_edges);
const uint32_t* psm_text_graph_offsets(const psm_text_graph* g);    /* V+1 entries */
const uint32_t* psm_text_graph_neighbors(const psm_text_graph* g);  /* 2E entries */
const uint32_t* psm_text_graph_labels(const psm_text_graph* g);     /* V entries */
void psm_text_graph_free(psm_text_graph* g);
const char* psm_text_graph_last_error(void);

/* ---- seeded generators (host) -------------------------------------------
 * Data graphs for the benchmark configurations. offsets has V+1 entries,
 * neighbors 2E. */
int psm_gen_chung_lu(uint32_t V, uint64_t E, double gamma, uint64_t seed, uint32_t* offsets,
                     uint32_t* neighbors);
int psm_gen_rmat(uint32_t scale, uint64_t E, double a, double b, double c, uint64_t seed,
                 uint32_t* offsets, uint32_t* neighbors);
/* zipf_s <= 0 draws labels uniformly */
int psm_gen_labels(uint32_t V, uint32_t L, double zipf_s, uint64_t seed, uint32_t* labels);
/* A connected n-vertex query drawn from the data graph: dense = 0 the tree of
 * a random walk, 1 the induced subgraph of a greedy densest growth, 2 the
 * induced subgraph of a random walk's vertices (cycles), 3 the induced
 * subgraph of a degree-biased random walk's vertices (each step picks a
 * neighbour with probability proportional to its degree). Returns the number
 * of query edges (pairs written to out_edges) or a negative status. */
int64_t psm_gen_query(uint32_t V, const uint32_t* offsets, const uint32_t* neighbors, uint32_t n,
                      int dense, uint64_t seed, uint32_t* out_vertices, uint32_t* out_edges,
                      uint64_t max_edges);

#ifdef __cplusplus
}
#endif

#endif /* PSM_ABI_H */
