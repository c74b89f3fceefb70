// psm_abi.cpp — host side of the C ABI declared in include/psm_abi.h.
//
// A Batch runs the matching pipeline for a set of queries against one
// resident data graph, each stage as a handful of launches over all queries:
//   filter   LDF / NLF bucket scans, GQL refinement rounds, CFL passes
//   tables   candidate bitmaps + ranks, BuildTable count / scan / fill
//   plan     GQL or CFL matching order on the host (n <= 64 per query)
//   evaluate task pre-expansion, then one persistent DFS kernel
// The plain path (psm_query) is a batch of one query driven stage by stage.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <atomic>
#include <functional>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "internal.h"
#include "plan.h"
#include "psm_abi.h"

using namespace psm;

// ---------------------------------------------------------------------------
// context, errors, device buffers
// ---------------------------------------------------------------------------
// The library's default enumeration launch shape (what the plain path uses):
// one lane-per-search wavefront per CU. Since lanes claim queue slots in
// batches, a single search runs fastest on few waves (fewer wavefronts
// contending for L2 and the queue): the 9,460-vertex C1 query enumerates in
// 15 ms at 1 wave per CU, 19 ms at 2 and 48 ms at 8.
constexpr uint32_t kDefaultWavesPerCu = 1;
constexpr uint32_t kDefaultBlockLanes = 0;   // lanes per workgroup of the lane DFS (0: 64)
constexpr uint32_t kDefaultClaimDiv = 1;     // array tasks per lane above which a claim takes several
constexpr double kDefaultTasksPerLane = 2.0;

struct psm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cu = 256;
    std::string err;
    // per-kernel-class timing: HIP events around launches on `stream`
    struct KStat { double launches = 0, ms = 0, bytes = 0; };
    KStat kstat[kKNumKinds];
    struct Pending { hipEvent_t a, b; int kind; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> ev_pool;
    unsigned long long* d_kbytes = nullptr;   // kernel-counted algorithmic bytes per class
    double time_limit_s = 60.0;               // reference default -time_limit 60
    // enumeration launch shape (psm_set_tuning)
    uint32_t dfs_waves_per_cu = kDefaultWavesPerCu;        // persistent DFS grid
    double expand_tasks_per_lane = kDefaultTasksPerLane;   // pre-expansion target
    uint32_t dfs_strategy = PSM_DFS_LANE;     // lane-per-task or wavefront-per-task DFS
    uint32_t queue_slots = 0;                 // ring slots for split-off work (0: automatic)
    uint32_t claim_div = kDefaultClaimDiv;    // array tasks per lane above which claims take several
    uint32_t block_lanes = kDefaultBlockLanes;  // lanes per workgroup of the lane DFS (0: 64)
    uint32_t expand_depth = kExpandMax;         // deepest level pre-expansion assigns
    uint32_t intersect = PSM_INTERSECT_ADAPTIVE;  // set intersection of the DFS kernels
    uint32_t serve_resident = 1;                // servers admit batches into one resident launch
    double wall_khz = 100000.0;               // device wall clock rate
    bool verbose = false;
};

namespace {

// Stage helpers may run on several host threads of one context (the CFL
// filter of a batch): the error text is written under a lock.
std::mutex& err_mu() {
    static std::mutex m;
    return m;
}

int set_err(psm_ctx* ctx, int code, const std::string& msg) {
    if (ctx) {
        std::lock_guard<std::mutex> lk(err_mu());
        ctx->err = msg;
    }
    return code;
}

int hip_fail(psm_ctx* ctx, hipError_t e, const char* what, int line) {
    char buf[512];
    snprintf(buf, sizeof(buf), "%s failed (%s) at psm_abi.cpp:%d", what, hipGetErrorString(e), line);
    if (ctx) {
        std::lock_guard<std::mutex> lk(err_mu());
        ctx->err = buf;
    }
    return e == hipErrorOutOfMemory ? PSM_ENOMEM : -(int)e;
}

#define HIP_TRY(ctx, expr)                                                      \
    do {                                                                        \
        hipError_t _e = (expr);                                                 \
        if (_e != hipSuccess) return hip_fail((ctx), _e, #expr, __LINE__);      \
    } while (0)

#define PSM_TRY(expr)                     \
    do {                                  \
        int _r = (expr);                  \
        if (_r != PSM_OK) return _r;      \
    } while (0)

// Device memory comes from a process-wide cache of size-bucketed blocks:
// hipFree synchronises the whole device, which would serialise concurrent
// contexts (serving workers), so released blocks are kept for reuse. A block
// is returned to the cache only after the stream that last used it has
// drained, so a recycled block is never still read by an earlier launch.
struct DevCache {
    std::mutex mu;
    std::multimap<size_t, void*> free_blocks;
    size_t cached = 0;
    static constexpr size_t kMaxCached = 48ull << 30;
};
DevCache& dev_cache() {
    static DevCache c;
    return c;
}

size_t bucket_size(size_t b) {
    if (b <= 4096) return 4096;
    size_t s = 4096;
    while (s < b) s <<= 1;
    // above 64 MiB, round to 1/8 steps of the power of two instead
    if (s > (64ull << 20)) {
        const size_t step = s / 16;
        return (b + step - 1) / step * step;
    }
    return s;
}

thread_local hipStream_t tls_stream = nullptr;   // stream of the running batch stage

struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;          // capacity of the block
    hipStream_t st = nullptr;  // stream the block was allocated for
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (!p) return;
        if (st) (void)hipStreamSynchronize(st);
        else (void)hipDeviceSynchronize();
        DevCache& c = dev_cache();
        bool keep;
        {
            std::lock_guard<std::mutex> lk(c.mu);
            keep = c.cached + bytes <= DevCache::kMaxCached;
            if (keep) {
                c.free_blocks.emplace(bytes, p);
                c.cached += bytes;
            }
        }
        if (!keep) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t alloc(size_t b) {
        release();
        const size_t want = bucket_size(b ? b : 16);
        st = tls_stream;
        DevCache& c = dev_cache();
        {
            std::lock_guard<std::mutex> lk(c.mu);
            auto it = c.free_blocks.find(want);
            if (it != c.free_blocks.end()) {
                p = it->second;
                c.cached -= want;
                c.free_blocks.erase(it);
                bytes = want;
                return hipSuccess;
            }
        }
        hipError_t e = hipMalloc(&p, want);
        if (e == hipErrorOutOfMemory) {
            // give cached blocks back to the runtime and try once more
            (void)hipGetLastError();
            std::lock_guard<std::mutex> lk(c.mu);
            for (auto& kv : c.free_blocks) (void)hipFree(kv.second);
            c.free_blocks.clear();
            c.cached = 0;
            e = hipMalloc(&p, want);
        }
        if (e != hipSuccess) { p = nullptr; return e; }
        bytes = want;
        return hipSuccess;
    }
    void swap(DBuf& o) {
        std::swap(p, o.p);
        std::swap(bytes, o.bytes);
        std::swap(st, o.st);
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Host vectors handed to upload() are often temporaries: the copy completes
// before returning so the source may be freed at once.
template <class T>
hipError_t upload(DBuf& buf, const std::vector<T>& v, hipStream_t s) {
    hipError_t e = buf.alloc(v.size() * sizeof(T));
    if (e != hipSuccess || v.empty()) return e;
    e = hipMemcpyAsync(buf.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

template <class T>
hipError_t download(std::vector<T>& v, const void* src, size_t n, hipStream_t s) {
    v.resize(n);
    if (n == 0) return hipSuccess;
    hipError_t e = hipMemcpyAsync(v.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

hipEvent_t take_event(psm_ctx* ctx) {
    if (!ctx->ev_pool.empty()) {
        hipEvent_t e = ctx->ev_pool.back();
        ctx->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

// Marks the start of a timed launch of class `kind`; returns its index.
size_t timer_begin(psm_ctx* ctx, int kind) {
    psm_ctx::Pending p{take_event(ctx), take_event(ctx), kind};
    (void)hipEventRecord(p.a, ctx->stream);
    ctx->pending.push_back(p);
    return ctx->pending.size() - 1;
}

void timer_end(psm_ctx* ctx, size_t i) { (void)hipEventRecord(ctx->pending[i].b, ctx->stream); }

// Waits for the stream and folds finished timers and byte counters into the
// context's statistics.
hipError_t timers_flush(psm_ctx* ctx) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return e;
    for (auto& p : ctx->pending) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            ctx->kstat[p.kind].ms += ms;
            ctx->kstat[p.kind].launches += 1;
        }
        ctx->ev_pool.push_back(p.a);
        ctx->ev_pool.push_back(p.b);
    }
    ctx->pending.clear();
    if (ctx->d_kbytes) {
        unsigned long long h[kKNumKinds];
        e = hipMemcpy(h, ctx->d_kbytes, sizeof(h), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return e;
        for (int k = 0; k < kKNumKinds; ++k) ctx->kstat[k].bytes += (double)h[k];
        e = hipMemset(ctx->d_kbytes, 0, sizeof(h));
    }
    return e;
}

#define TIMED(ctx, kind, expr)                              \
    do {                                                    \
        const size_t _ti = timer_begin((ctx), (kind));      \
        HIP_TRY((ctx), (expr));                             \
        timer_end((ctx), _ti);                              \
    } while (0)

double now_ms();

// Stage progress on stderr when PSM_VERBOSE is set.
#define verbose(...)                                                  \
    do {                                                              \
        if (ctx->verbose) {                                           \
            fprintf(stderr, "[psm] " __VA_ARGS__);                    \
            fputc('\n', stderr);                                     \
        }                                                             \
    } while (0)

double now_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// ---------------------------------------------------------------------------
// data graph
// ---------------------------------------------------------------------------
struct psm_graph {
    psm_ctx* ctx = nullptr;
    uint32_t V = 0, L = 0, max_deg = 0, max_freq = 0;
    uint64_t nnz = 0;
    std::vector<uint32_t> h_ri_off;   // host copy: bucket sizes drive work items
    DBuf off, adj, lab, ri_off, ri, nlf_off, nlf, adj_bl, nlf_beg, deg_ri, nlf_ri_off, nlf_ri;
    DevGraph dg{};
};

// Reverse index (graph.cpp:15-33) and neighbour-label runs (graph.cpp:36-52)
// built on the host with all cores, then uploaded. Also the label-grouped
// adjacency (the reference's OPTIMIZED_LABELED_GRAPH layout, graph.cpp:35-90):
// adj_bl holds each vertex's neighbours ordered by (label, id), and
// nlf_beg[i] is where run i (label nlf[i].x, nlf[i].y neighbours) starts in it.
static void build_graph_indexes(uint32_t V, const uint32_t* off, const uint32_t* adj,
                                const uint32_t* lab, uint32_t L, std::vector<uint32_t>& ri_off,
                                std::vector<uint32_t>& ri, std::vector<uint32_t>& nlf_off,
                                std::vector<uint2>& nlf, std::vector<uint32_t>& adj_bl,
                                std::vector<uint32_t>& nlf_beg) {
    ri_off.assign(L + 1, 0);
    for (uint32_t v = 0; v < V; ++v) ri_off[lab[v] + 1]++;
    for (uint32_t l = 0; l < L; ++l) ri_off[l + 1] += ri_off[l];
    ri.resize(V);
    {
        std::vector<uint32_t> fill(ri_off.begin(), ri_off.end() - 1);
        for (uint32_t v = 0; v < V; ++v) ri[fill[lab[v]]++] = v;
    }
    unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    std::vector<uint32_t> runs(V, 0);
    auto pass = [&](int write, unsigned t) {
        std::vector<uint32_t> cnt(L, 0), touched;
        for (uint32_t v = t; v < V; v += nt) {
            touched.clear();
            for (uint32_t p = off[v]; p < off[v + 1]; ++p) {
                uint32_t l = lab[adj[p]];
                if (cnt[l]++ == 0) touched.push_back(l);
            }
            if (write) {
                std::sort(touched.begin(), touched.end());
                uint32_t o = nlf_off[v], at = off[v];
                for (uint32_t l : touched) {
                    nlf[o] = make_uint2(l, cnt[l]);
                    nlf_beg[o++] = at;
                    const uint32_t c = cnt[l];
                    cnt[l] = at;   // next free position of label l
                    at += c;
                }
                // stable counting sort: ids stay ascending inside a run
                for (uint32_t p = off[v]; p < off[v + 1]; ++p) adj_bl[cnt[lab[adj[p]]]++] = adj[p];
            } else {
                runs[v] = (uint32_t)touched.size();
            }
            for (uint32_t l : touched) cnt[l] = 0;
        }
    };
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) th.emplace_back(pass, 0, t);
    for (auto& x : th) x.join();
    th.clear();
    nlf_off.assign(V + 1, 0);
    for (uint32_t v = 0; v < V; ++v) nlf_off[v + 1] = nlf_off[v] + runs[v];
    nlf.resize(nlf_off[V]);
    nlf_beg.resize(nlf_off[V]);
    adj_bl.resize(off[V]);
    for (unsigned t = 0; t < nt; ++t) th.emplace_back(pass, 1, t);
    for (auto& x : th) x.join();
}

// The degrees and neighbour-label runs in reverse-index order: position j
// describes vertex ri[j] (what a label-bucket scan reads, as streams).
static void build_ri_order(uint32_t V, const uint32_t* off, const std::vector<uint32_t>& ri,
                           const std::vector<uint32_t>& nlf_off, const std::vector<uint2>& nlf,
                           std::vector<uint32_t>& deg_ri, std::vector<uint32_t>& nlf_ri_off,
                           std::vector<uint2>& nlf_ri) {
    deg_ri.resize(V);
    nlf_ri_off.assign(V + 1, 0);
    for (uint32_t j = 0; j < V; ++j) {
        const uint32_t v = ri[j];
        deg_ri[j] = off[v + 1] - off[v];
        nlf_ri_off[j + 1] = nlf_ri_off[j] + (nlf_off[v + 1] - nlf_off[v]);
    }
    nlf_ri.resize(nlf.size());
    const unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            for (uint32_t j = t; j < V; j += nt) {
                const uint32_t v = ri[j];
                std::copy(nlf.begin() + nlf_off[v], nlf.begin() + nlf_off[v + 1], nlf_ri.begin() + nlf_ri_off[j]);
            }
        });
    for (auto& x : th) x.join();
}

// ---------------------------------------------------------------------------
// query host state
// ---------------------------------------------------------------------------
namespace {

struct QueryHost {
    QGraph qg;
    int status = PSM_OK;              // filter result (PSM_EEMPTY like the reference's false)
    std::vector<uint32_t> cand_count;
    std::vector<uint64_t> cand_begin;
    uint32_t slot_base = 0;           // bitmap slot of (this query, u) = slot_base + u
    std::vector<uint64_t> tab_row;    // n*n: first row of table (a,b), ~0 when no edge
    std::vector<uint64_t> tab_nnz;    // n*n
    Plan plan;
    bool planned = false;
    uint64_t count = 0, calls = 0, rejects = 0, probe_bytes = 0, algo_bytes = 0;
    bool capped = false;
    bool timed_out = false;           // partial count: the reference's "Overtime: 1"
    double done_ms = -1;              // host clock (now_ms) when the count became final
    CflTree filter_tree;              // BFS tree of the last CFL / DPiso filter (query graph)
    CflTree plan_tree;                // BFS tree of the last CFL / DPiso order (query unit)
    bool has_filter_tree = false, has_plan_tree = false;
};

constexpr uint64_t kNoTable = ~0ull;

struct JoinResult {
    uint64_t rows_a = 0, rows_b = 0, count = 0;
    uint32_t keys = 0;
    double enum_ms = 0, join_ms = 0;
};

struct SplitJoinResult {
    uint64_t count = 0, unit_rows = 0, max_inter_rows = 0, calls = 0;
    uint32_t units = 0;
    double enum_ms = 0, join_ms = 0;
};

struct Batch {
    psm_ctx* ctx = nullptr;
    const psm_graph* g = nullptr;
    std::vector<QueryHost*> qs;
    // candidates
    DBuf cand_pool;
    uint64_t cand_total = 0;
    bool bitmaps_current = false;     // bitmaps == final candidate sets
    // bitmaps
    DBuf bitmaps, wprefix;
    uint64_t wps = 0;
    uint32_t n_slots = 0;
    // tables
    DBuf row_off, idx;
    uint64_t n_rows = 0, n_idx = 0;
    bool tables_built = false;
    // plan + enumeration
    DBuf plans, ops, qcnt;
    double t_filter = 0, t_table = 0, t_plan = 0, t_enum = 0, t_expand = 0;
    uint64_t tasks_run = 0;
    uint64_t lane_steps = 0, lane_iters = 0;   // DFS lane utilisation
    double t_begin = -1;   // start of filtering: the time limit counts from here
    double enum_limit_s = -1;   // >= 0: an enumeration time limit counted from t_eval instead
    double t_eval = 0;
    // the last DFS launch: host clock just before it and the device wall
    // clock when its first block started (maps done_tick to host time)
    double dfs_launch_ms = 0;
    unsigned long long dfs_start_tick = 0;
    // serving: host-visible completion words of the queries (4 per query,
    // see DQueue::notify) and a callback run while the DFS launch is in flight
    unsigned long long* notify = nullptr;
    std::function<void()> poll;

    hipStream_t s() const { return ctx->stream; }

    int filter(int method);
    int bucket_filter(bool use_nlf, std::vector<QueryHost*>& active);
    int ensure_bitmaps(bool from_candidates);
    int compact_pool();
    int gql_refine(std::vector<QueryHost*>& active);
    int for_each_query(size_t count, const std::function<int(size_t, uint32_t, hipStream_t)>& fn,
                       uint32_t& workers, const std::function<void(uint32_t)>& worker_done = nullptr);
    int cfl_filter(std::vector<QueryHost*>& active, bool dpiso);
    int cfl_filter_query(QueryHost* q, hipStream_t st, DBuf& cnt, uint32_t& epoch, DBuf& defer, bool dpiso,
                         std::vector<std::vector<uint32_t>>& sets);
    int build_tables();
    int make_plans(int order_method);
    int run_plans(std::vector<DPlan> hp, std::vector<DOp> hops, EmitSink sink,
                  std::vector<QCounters>& qc, bool& timed_out);
    int expand_tasks(const std::vector<DPlan>& hp, uint64_t target_total, DBuf& d_out, uint64_t& n_out,
                     std::vector<uint64_t>& ntask);
    // points a plan at this batch's candidate pool and tables
    void bind_tables(DPlan& P) const {
        P.row_off = row_off.as<uint64_t>();
        P.idx = idx.as<uint32_t>();
        P.cand_pool = cand_pool.as<uint32_t>();
    }
    int evaluate(uint64_t cap, double enum_limit = -1);
    int join_count(QueryHost& q, uint64_t max_rows, JoinResult& jr);
    int split_join_count(QueryHost& q, int pattern, uint64_t max_rows, SplitJoinResult& jr);
    int host_table(const QueryHost& q, uint32_t a, uint32_t b, std::vector<uint32_t>& off,
                   std::vector<uint32_t>& idxv, hipStream_t st);
    int nlf_count(const QGraph& q, uint32_t u, uint32_t label_override, uint32_t& out, hipStream_t st,
                  uint32_t mode = kModeNLF);
};

// ---- candidate filters ------------------------------------------------------

// NLF / LDF over label buckets for every (query, u) with a live query: one
// label-shared scan over the buckets of the labels in use (CGroup).
// Candidate pool = concatenation of all sets, in (query, u) order.
int Batch::bucket_filter(bool use_nlf, std::vector<QueryHost*>& active) {
    std::vector<QVert> qv;
    std::vector<uint2> qnlf;
    std::vector<CItem> items;
    uint64_t chunks64 = 0;
    std::map<uint32_t, std::vector<uint32_t>> by_label;   // label -> its items
    for (QueryHost* q : qs) {
        q->cand_count.assign(q->qg.n, 0);
        q->cand_begin.assign(q->qg.n, 0);
        for (uint32_t u = 0; u < q->qg.n; ++u) {
            QVert v;
            v.label = q->qg.lab[u];
            v.degree = q->qg.degree(u);
            v.nlf_begin = (uint32_t)qnlf.size();
            v.nlf_count = (uint32_t)q->qg.nlf[u].size();
            for (auto& r : q->qg.nlf[u]) qnlf.push_back(make_uint2(r.first, r.second));
            CItem it{};
            const uint32_t l = v.label;
            it.src_begin = l < g->L ? g->h_ri_off[l] : 0;
            it.len = l < g->L ? g->h_ri_off[l + 1] - g->h_ri_off[l] : 0;
            it.block_start = (uint32_t)chunks64;   // first chunk of 64 elements (checked below)
            it.mode = use_nlf ? kModeNLF : kModeLDF;
            it.qv = (uint32_t)qv.size();
            if (it.len) by_label[l].push_back((uint32_t)items.size());
            qv.push_back(v);
            items.push_back(it);
            chunks64 += (it.len + kBucketChunk - 1) / kBucketChunk;
            if (chunks64 > 0xFFFFFFFFull) return set_err(ctx, PSM_ENOMEM, "batch too large for one bucket scan");
        }
    }
    const uint32_t chunks = (uint32_t)chunks64;
    std::vector<CGroup> groups;
    std::vector<CBucketItem> bitems;
    uint64_t waves64 = 0;
    for (auto& [l, list] : by_label) {
        CGroup gr{};
        gr.src_begin = g->h_ri_off[l];
        gr.len = g->h_ri_off[l + 1] - g->h_ri_off[l];
        gr.wave_start = (uint32_t)waves64;
        gr.item_begin = (uint32_t)bitems.size();
        gr.n_items = (uint32_t)list.size();
        gr.n_sub = (gr.n_items + kBucketChunk - 1) / kBucketChunk;
        for (uint32_t i : list) {
            const QVert& v = qv[items[i].qv];
            bitems.push_back(CBucketItem{items[i].block_start, v.degree, v.nlf_begin,
                                         v.nlf_count | (use_nlf ? 0u : 1u << 31)});
        }
        groups.push_back(gr);
        waves64 += (uint64_t)((gr.len + kBucketChunk - 1) / kBucketChunk) * gr.n_sub;
    }
    if (waves64 > 0xFFFFFFFFull) return set_err(ctx, PSM_ENOMEM, "batch too large for one bucket scan");
    const uint32_t waves = (uint32_t)waves64;
    DBuf d_qnlf, d_groups, d_bitems, d_masks, d_cc, d_cp, d_tmp, d_starts, d_ends, d_cnt, d_beg;
    HIP_TRY(ctx, upload(d_qnlf, qnlf, s()));
    HIP_TRY(ctx, upload(d_groups, groups, s()));
    HIP_TRY(ctx, upload(d_bitems, bitems, s()));
    HIP_TRY(ctx, d_masks.alloc((size_t)chunks * 8));
    HIP_TRY(ctx, d_cc.alloc((size_t)chunks * 4));
    HIP_TRY(ctx, d_cp.alloc((size_t)(chunks + 1) * 8));
    HIP_TRY(ctx, d_tmp.alloc(scan_tmp_entries(chunks) * 8));
    HIP_TRY(ctx, launch_bucket_flag(g->dg, d_groups.as<CGroup>(), (uint32_t)groups.size(), waves,
                                    d_bitems.as<CBucketItem>(), d_qnlf.as<uint2>(),
                                    d_masks.as<unsigned long long>(), d_cc.as<uint32_t>(), s()));
    HIP_TRY(ctx, launch_scan_u32(d_cc.as<uint32_t>(), chunks, d_cp.as<uint64_t>(),
                                 d_tmp.as<unsigned long long>(), s()));
    // per item: begin = prefix[first chunk], count = prefix[next item's first chunk] - begin
    const uint32_t ni = (uint32_t)items.size();
    std::vector<uint64_t> starts(ni), ends(ni), zeros(ni, 0);
    for (uint32_t i = 0; i < ni; ++i) {
        starts[i] = items[i].block_start;
        ends[i] = i + 1 < ni ? items[i + 1].block_start : chunks;
    }
    HIP_TRY(ctx, upload(d_starts, starts, s()));
    HIP_TRY(ctx, upload(d_ends, ends, s()));
    HIP_TRY(ctx, d_cnt.alloc((size_t)ni * 8));
    HIP_TRY(ctx, d_beg.alloc((size_t)ni * 8));
    HIP_TRY(ctx, launch_span(d_cp.as<uint64_t>(), d_starts.as<uint64_t>(), d_ends.as<uint64_t>(), ni,
                             d_cnt.as<uint64_t>(), s()));
    {
        DBuf d_zero;
        HIP_TRY(ctx, upload(d_zero, zeros, s()));
        HIP_TRY(ctx, launch_span(d_cp.as<uint64_t>(), d_zero.as<uint64_t>(), d_starts.as<uint64_t>(),
                                 ni, d_beg.as<uint64_t>(), s()));
        std::vector<uint64_t> cnt, beg;
        HIP_TRY(ctx, download(cnt, d_cnt.p, ni, s()));
        HIP_TRY(ctx, download(beg, d_beg.p, ni, s()));
        uint32_t i = 0;
        cand_total = 0;
        for (QueryHost* q : qs) {
            for (uint32_t u = 0; u < q->qg.n; ++u, ++i) {
                q->cand_count[u] = (uint32_t)cnt[i];
                q->cand_begin[u] = beg[i];
                cand_total = std::max(cand_total, beg[i] + cnt[i]);
            }
        }
    }
    HIP_TRY(ctx, cand_pool.alloc(cand_total * 4));
    HIP_TRY(ctx, launch_bucket_scatter(g->dg, d_groups.as<CGroup>(), (uint32_t)groups.size(), waves,
                                       d_bitems.as<CBucketItem>(), d_masks.as<unsigned long long>(),
                                       d_cp.as<uint64_t>(), cand_pool.as<uint32_t>(), s()));
    HIP_TRY(ctx, hipStreamSynchronize(s()));
    // FilterVertices.cpp:19-65: the reference stops at the first empty set
    active.clear();
    for (QueryHost* q : qs) {
        q->status = PSM_OK;
        for (uint32_t u = 0; u < q->qg.n; ++u) {
            if (q->status != PSM_OK) q->cand_count[u] = 0;
            else if (q->cand_count[u] == 0) q->status = PSM_EEMPTY;
        }
        if (q->status == PSM_OK) active.push_back(q);
    }
    bitmaps_current = false;
    return PSM_OK;
}

// Bitmap slots for every (query, u); set from the current candidates.
int Batch::ensure_bitmaps(bool from_candidates) {
    wps = ((uint64_t)g->V + 31) / 32;
    uint32_t slots = 0;
    for (QueryHost* q : qs) {
        q->slot_base = slots;
        slots += q->qg.n;
    }
    n_slots = slots;
    if (bitmaps.bytes < (size_t)slots * wps * 4) HIP_TRY(ctx, bitmaps.alloc((size_t)slots * wps * 4));
    HIP_TRY(ctx, hipMemsetAsync(bitmaps.p, 0, (size_t)slots * wps * 4, s()));
    if (!from_candidates) return PSM_OK;
    std::vector<uint64_t> sb;
    std::vector<uint32_t> sl, ss;
    uint32_t maxlen = 0;
    for (QueryHost* q : qs)
        for (uint32_t u = 0; u < q->qg.n; ++u) {
            if (!q->cand_count[u]) continue;
            sb.push_back(q->cand_begin[u]);
            sl.push_back(q->cand_count[u]);
            ss.push_back(q->slot_base + u);
            maxlen = std::max(maxlen, q->cand_count[u]);
        }
    DBuf d_sb, d_sl, d_ss;
    HIP_TRY(ctx, upload(d_sb, sb, s()));
    HIP_TRY(ctx, upload(d_sl, sl, s()));
    HIP_TRY(ctx, upload(d_ss, ss, s()));
    for (size_t i = 0; i < sb.size(); i += 65535) {
        uint32_t ns = (uint32_t)std::min<size_t>(65535, sb.size() - i);
        HIP_TRY(ctx, launch_set_bits(cand_pool.as<uint32_t>(), d_sb.as<uint64_t>() + i,
                                     d_sl.as<uint32_t>() + i, d_ss.as<uint32_t>() + i, ns, maxlen,
                                     bitmaps.as<uint32_t>(), wps, s()));
    }
    HIP_TRY(ctx, hipStreamSynchronize(s()));
    bitmaps_current = true;
    return PSM_OK;
}

// Removes kInvalid entries from every candidate list, preserving order.
int Batch::compact_pool() {
    std::vector<CItem> items;
    uint32_t blocks = 0;
    for (QueryHost* q : qs)
        for (uint32_t u = 0; u < q->qg.n; ++u) {
            CItem it{};
            it.src_begin = q->cand_begin[u];
            it.len = q->cand_count[u];
            it.block_start = blocks;
            it.mode = kModeValid;
            items.push_back(it);
            blocks += (it.len + 255) / 256;
        }
    const uint32_t ni = (uint32_t)items.size();
    DBuf d_items, d_masks, d_bc, d_bp, d_tmp, d_starts, d_ends, d_zero, d_cnt, d_beg, out;
    HIP_TRY(ctx, upload(d_items, items, s()));
    HIP_TRY(ctx, d_masks.alloc((size_t)blocks * 4 * 8));
    HIP_TRY(ctx, d_bc.alloc((size_t)blocks * 4));
    HIP_TRY(ctx, d_bp.alloc((size_t)(blocks + 1) * 8));
    HIP_TRY(ctx, d_tmp.alloc(scan_tmp_entries(blocks) * 8));
    HIP_TRY(ctx, launch_compact_flag(g->dg, d_items.as<CItem>(), ni, blocks, nullptr, nullptr,
                                     cand_pool.as<uint32_t>(), nullptr,
                                     d_masks.as<unsigned long long>(), d_bc.as<uint32_t>(), s()));
    HIP_TRY(ctx, launch_scan_u32(d_bc.as<uint32_t>(), blocks, d_bp.as<uint64_t>(),
                                 d_tmp.as<unsigned long long>(), s()));
    std::vector<uint64_t> starts(ni), ends(ni), zeros(ni, 0), cnt, beg;
    for (uint32_t i = 0; i < ni; ++i) {
        starts[i] = items[i].block_start;
        ends[i] = i + 1 < ni ? items[i + 1].block_start : blocks;
    }
    HIP_TRY(ctx, upload(d_starts, starts, s()));
    HIP_TRY(ctx, upload(d_ends, ends, s()));
    HIP_TRY(ctx, upload(d_zero, zeros, s()));
    HIP_TRY(ctx, d_cnt.alloc((size_t)ni * 8));
    HIP_TRY(ctx, d_beg.alloc((size_t)ni * 8));
    HIP_TRY(ctx, launch_span(d_bp.as<uint64_t>(), d_starts.as<uint64_t>(), d_ends.as<uint64_t>(), ni,
                             d_cnt.as<uint64_t>(), s()));
    HIP_TRY(ctx, launch_span(d_bp.as<uint64_t>(), d_zero.as<uint64_t>(), d_starts.as<uint64_t>(), ni,
                             d_beg.as<uint64_t>(), s()));
    HIP_TRY(ctx, download(cnt, d_cnt.p, ni, s()));
    HIP_TRY(ctx, download(beg, d_beg.p, ni, s()));
    uint64_t total = 0;
    for (uint32_t i = 0; i < ni; ++i) total = std::max(total, beg[i] + cnt[i]);
    HIP_TRY(ctx, out.alloc(total * 4));
    HIP_TRY(ctx, launch_compact_scatter(d_items.as<CItem>(), ni, blocks, nullptr, cand_pool.as<uint32_t>(),
                                        d_masks.as<unsigned long long>(), d_bp.as<uint64_t>(),
                                        out.as<uint32_t>(), s()));
    HIP_TRY(ctx, hipStreamSynchronize(s()));
    cand_pool.swap(out);
    cand_total = total;
    uint32_t i = 0;
    for (QueryHost* q : qs)
        for (uint32_t u = 0; u < q->qg.n; ++u, ++i) {
            q->cand_count[u] = (uint32_t)cnt[i];
            q->cand_begin[u] = beg[i];
        }
    return PSM_OK;
}

// FilterVertices.cpp:67-143: two in-place rounds over u = 0..n-1; queries
// advance in lock step, one launch per (round, u).
int Batch::gql_refine(std::vector<QueryHost*>& active) {
    PSM_TRY(ensure_bitmaps(true));
    uint32_t max_n = 0;
    for (QueryHost* q : active) max_n = std::max(max_n, q->qg.n);
    // membership words: member[a * V + w] bit u <=> w in C(u) of active query a
    DBuf member;
    {
        const size_t words = active.size() * (size_t)g->V;
        HIP_TRY(ctx, member.alloc(words * 8));
        HIP_TRY(ctx, hipMemsetAsync(member.p, 0, words * 8, s()));
        std::vector<uint64_t> sb, sm;
        std::vector<uint32_t> sl, sbit;
        uint32_t mx = 0;
        for (size_t a = 0; a < active.size(); ++a) {
            QueryHost* q = active[a];
            for (uint32_t u = 0; u < q->qg.n; ++u) {
                if (!q->cand_count[u]) continue;
                sb.push_back(q->cand_begin[u]);
                sl.push_back(q->cand_count[u]);
                sm.push_back(a * (uint64_t)g->V);
                sbit.push_back(u);
                mx = std::max(mx, q->cand_count[u]);
            }
        }
        DBuf d_sb, d_sl, d_sm, d_sbit;
        HIP_TRY(ctx, upload(d_sb, sb, s()));
        HIP_TRY(ctx, upload(d_sl, sl, s()));
        HIP_TRY(ctx, upload(d_sm, sm, s()));
        HIP_TRY(ctx, upload(d_sbit, sbit, s()));
        TIMED(ctx, kKBitmap,
              launch_set_member(cand_pool.as<uint32_t>(), d_sb.as<uint64_t>(), d_sl.as<uint32_t>(),
                                d_sm.as<uint64_t>(), d_sbit.as<uint32_t>(), (uint32_t)sb.size(), mx,
                                member.as<unsigned long long>(), s()));
        HIP_TRY(ctx, hipStreamSynchronize(s()));
    }
    // work items of every query vertex position p (the candidates of vertex p
    // of each query); built again for the second round after the pool is
    // compacted (the first round's drops leave most slots kInvalid)
    std::vector<GItem> items;
    std::vector<uint32_t> left, left_lab, labs, item_begin(max_n + 1, 0), cands(max_n, 0);
    uint32_t max_cands = 0;
    DBuf d_items, d_left, d_left_lab, d_labs, d_lists, d_nlists;
    auto build_items = [&]() -> int {
        items.clear();
        left.clear();
        left_lab.clear();
        labs.clear();
        max_cands = 0;
        for (uint32_t p = 0; p < max_n; ++p) {
            item_begin[p] = (uint32_t)items.size();
            uint32_t waves = 0;
            for (size_t a = 0; a < active.size(); ++a) {
                QueryHost* q = active[a];
                if (p >= q->qg.n || q->cand_count[p] == 0) continue;
                GItem it{};
                it.cand_begin = q->cand_begin[p];
                it.len = q->cand_count[p];
                it.wave_start = waves;
                it.bm_self = q->slot_base + p;
                it.u = p;
                it.mem_base = a * (uint64_t)g->V;
                it.k = q->qg.degree(p);
                it.left_begin = (uint32_t)left.size();
                it.lab_begin = (uint32_t)labs.size();
                const size_t l0 = labs.size();
                for (uint32_t i = q->qg.off[p]; i < q->qg.off[p + 1]; ++i) {
                    left.push_back(q->slot_base + q->qg.adj[i]);
                    left_lab.push_back(q->qg.lab[q->qg.adj[i]]);
                    labs.push_back(q->qg.lab[q->qg.adj[i]]);
                }
                std::sort(labs.begin() + l0, labs.end());
                labs.erase(std::unique(labs.begin() + l0, labs.end()), labs.end());
                it.nlab = (uint32_t)(labs.size() - l0);
                items.push_back(it);
                waves += it.len;
            }
            cands[p] = waves;
            max_cands = std::max(max_cands, waves);
        }
        item_begin[max_n] = (uint32_t)items.size();
        HIP_TRY(ctx, upload(d_items, items, s()));
        HIP_TRY(ctx, upload(d_left, left, s()));
        HIP_TRY(ctx, upload(d_left_lab, left_lab, s()));
        HIP_TRY(ctx, upload(d_labs, labs, s()));
        return PSM_OK;
    };
    PSM_TRY(build_items());
    // deferred pairs of one (round, p): class c at c * max_cands; reused by
    // every launch group (stream order)
    HIP_TRY(ctx, d_lists.alloc(4ull * max_cands * sizeof(uint2)));
    // per (round, p): the four deferral counters and the wavefront kernel's
    // cursor (eight slots), cleared by a fill kernel
    constexpr uint64_t kCounterSlots = 8;
    const uint64_t n_counters = 2ull * max_n * kCounterSlots;
    HIP_TRY(ctx, d_nlists.alloc(n_counters * 4));
    HIP_TRY(ctx, launch_fill_u32(d_nlists.as<uint32_t>(), 0u, n_counters, s()));
    // removal flags per round and bitmap slot, and the second round's skip
    // flags per item (GqlRound)
    uint32_t n_slots = 0;
    for (QueryHost* q : active) n_slots = std::max(n_slots, q->slot_base + q->qg.n);
    DBuf d_removed, d_skip;
    HIP_TRY(ctx, d_removed.alloc(2ull * n_slots * 4));
    HIP_TRY(ctx, launch_fill_u32(d_removed.as<uint32_t>(), 0u, 2ull * n_slots, s()));
    HIP_TRY(ctx, d_skip.alloc(std::max<size_t>(1, items.size()) * 4));
    // Launch groups follow one another on the stream without host syncs: the
    // deferred-class kernels read their counts from d_nlists.
    const uint32_t lists_cap = max_cands;   // d_lists holds 4 classes of this many pairs
    for (int round = 0; round < 2; ++round) {
        if (round == 1) {
            // the second round tests only the first round's survivors: the
            // pool is compacted (order kept) and the items rebuilt on it
            HIP_TRY(ctx, timers_flush(ctx));
            PSM_TRY(compact_pool());
            PSM_TRY(build_items());
            if (max_cands > lists_cap) return set_err(ctx, PSM_EINVAL, "candidate sets grew between rounds");
        }
        uint32_t* removed = d_removed.as<uint32_t>() + (uint64_t)round * n_slots;
        for (uint32_t p = 0; p < max_n; ++p) {
            const uint32_t ni = item_begin[p + 1] - item_begin[p];
            if (!ni) continue;
            const GItem* its = d_items.as<GItem>() + item_begin[p];
            uint32_t* cnt = d_nlists.as<uint32_t>() + ((uint64_t)round * max_n + p) * kCounterSlots;
            uint2* lists = d_lists.as<uint2>();
            GqlRound rd{removed, nullptr};
            if (round == 1) {
                uint32_t* skip = d_skip.as<uint32_t>() + item_begin[p];
                TIMED(ctx, kKGqlRefine,
                      launch_gql_skip(its, ni, d_left.as<uint32_t>(), d_removed.as<uint32_t>(), removed, skip, s()));
                rd.skip = skip;
            }
            TIMED(ctx, kKGqlRefine,
                  launch_gql_refine_small(g->dg, its, ni, cands[p], d_left.as<uint32_t>(), d_labs.as<uint32_t>(),
                                          cand_pool.as<uint32_t>(),
                                          bitmaps.as<uint32_t>(), wps, member.as<unsigned long long>(), lists,
                                          max_cands, cnt, rd, ctx->d_kbytes + kKGqlRefine, s()));
            const uint32_t lanes[3] = {16, 32, 64};
            for (int c = 0; c < 3; ++c)
                TIMED(ctx, kKGqlRefine,
                      launch_gql_hall(lanes[c], g->dg, its, lists + (uint64_t)c * max_cands, cnt + c, cands[p],
                                      d_left.as<uint32_t>(), d_labs.as<uint32_t>(), cand_pool.as<uint32_t>(), bitmaps.as<uint32_t>(), wps,
                                      member.as<unsigned long long>(), rd, ctx->d_kbytes + kKGqlRefine, s()));
            TIMED(ctx, kKGqlRefine,
                  launch_gql_refine(g->dg, its, lists + 3ull * max_cands, cnt + 3, cnt + 4, cands[p], d_left.as<uint32_t>(),
                                    d_left_lab.as<uint32_t>(), d_labs.as<uint32_t>(), cand_pool.as<uint32_t>(), bitmaps.as<uint32_t>(), wps,
                                    member.as<unsigned long long>(), rd, ctx->d_kbytes + kKGqlRefine, s()));
        }
    }
    if (ctx->verbose) {
        std::vector<uint32_t> nd, sk;
        HIP_TRY(ctx, download(nd, d_nlists.p, n_counters, s()));
        HIP_TRY(ctx, download(sk, d_skip.p, items.size(), s()));
        for (int round = 0; round < 2; ++round)
            for (uint32_t p = 0; p < max_n; ++p) {
                const uint32_t* c = nd.data() + ((uint64_t)round * max_n + p) * kCounterSlots;
                uint64_t skipped = 0;
                if (round == 1)
                    for (uint32_t i = item_begin[p]; i < item_begin[p + 1]; ++i) skipped += sk[i] ? items[i].len : 0;
                verbose("gql round %d u %u: %u candidates, %llu skipped, deferred %u / %u / %u / %u", round, p,
                        cands[p], (unsigned long long)skipped, c[0], c[1], c[2], c[3]);
            }
    }
    HIP_TRY(ctx, timers_flush(ctx));
    PSM_TRY(compact_pool());
    bitmaps_current = true;   // refinement kept bitmaps == surviving candidates
    for (QueryHost* q : active)
        for (uint32_t u = 0; u < q->qg.n; ++u)
            if (q->cand_count[u] == 0) q->status = PSM_EEMPTY;
    return PSM_OK;
}

int Batch::filter(int method) {
    tls_stream = s();
    const double t0 = now_ms();
    // the reference sets its deadline before FilterVertices
    // (StudyPerformance.cpp:1134-1148): filter, tables, plan and enumeration
    // share one time limit
    t_begin = t0;
    std::vector<QueryHost*> active;
    for (QueryHost* q : qs) {
        if (q->qg.n == 0 || q->qg.n > kMaxQ)
            return set_err(ctx, PSM_EINVAL, "query size must be in [1, 64]");
        if (method == PSM_FILTER_GQL && q->qg.max_degree() > kMaxGqlDegree)
            return set_err(ctx, PSM_EINVAL, "GQL filter: query vertex degree above 32");
    }
    tables_built = false;
    for (QueryHost* q : qs) q->planned = false;
    switch (method) {
        case PSM_FILTER_LDF: PSM_TRY(bucket_filter(false, active)); break;
        case PSM_FILTER_NLF: PSM_TRY(bucket_filter(true, active)); break;
        case PSM_FILTER_GQL:
            PSM_TRY(bucket_filter(true, active));
            if (!active.empty()) PSM_TRY(gql_refine(active));
            break;
        case PSM_FILTER_CFL:
        case PSM_FILTER_DPISO:
            for (QueryHost* q : qs) active.push_back(q);
            PSM_TRY(cfl_filter(active, method == PSM_FILTER_DPISO));
            break;
        default: return set_err(ctx, PSM_EINVAL, "unknown filter method");
    }
    t_filter += now_ms() - t0;
    verbose("filter %d: %zu queries, %llu candidates, %.1f ms", method, qs.size(),
            (unsigned long long)cand_total, now_ms() - t0);
    return PSM_OK;
}

// ---- BuildTable --------------------------------------------------------------
int Batch::build_tables() {
    tls_stream = s();
    const double t0 = now_ms();
    if (!bitmaps_current) PSM_TRY(ensure_bitmaps(true));
    if (wprefix.bytes < (size_t)n_slots * wps * 4) HIP_TRY(ctx, wprefix.alloc((size_t)n_slots * wps * 4));
    HIP_TRY(ctx, launch_slot_prefix(bitmaps.as<uint32_t>(), wprefix.as<uint32_t>(), n_slots, wps, s()));
    std::vector<TItem> items;
    std::vector<uint64_t> row_start;
    std::vector<std::pair<QueryHost*, uint32_t>> item_of;  // (query, a*n+b)
    n_rows = 0;
    for (QueryHost* q : qs) {
        const uint32_t n = q->qg.n;
        q->tab_row.assign((size_t)n * n, kNoTable);
        q->tab_nnz.assign((size_t)n * n, 0);
        for (uint32_t a = 0; a < n; ++a)
            for (uint32_t i = q->qg.off[a]; i < q->qg.off[a + 1]; ++i) {
                const uint32_t b = q->qg.adj[i];
                TItem it{};
                it.cand_a = q->cand_begin[a];
                it.len_a = q->cand_count[a];
                it.bm_b = q->slot_base + b;
                it.lab_b = q->qg.lab[b];
                it.row_base = n_rows;
                q->tab_row[(size_t)a * n + b] = n_rows;
                items.push_back(it);
                row_start.push_back(n_rows);
                item_of.push_back({q, a * n + b});
                n_rows += it.len_a;
            }
    }
    const uint32_t ni = (uint32_t)items.size();
    DBuf d_items, d_rs, d_deg, d_tmp, d_ends, d_nnz, d_long, d_nlong;
    HIP_TRY(ctx, upload(d_items, items, s()));
    HIP_TRY(ctx, upload(d_rs, row_start, s()));
    HIP_TRY(ctx, d_deg.alloc(n_rows * 8));
    HIP_TRY(ctx, row_off.alloc((n_rows + 1) * 8));
    HIP_TRY(ctx, d_tmp.alloc(scan_tmp_entries(n_rows) * 8));
    HIP_TRY(ctx, d_long.alloc(std::max<uint64_t>(n_rows, 1) * 8));
    HIP_TRY(ctx, d_nlong.alloc(8));
    HIP_TRY(ctx, hipMemsetAsync(d_nlong.p, 0, 8, s()));
    TIMED(ctx, kKTableCount,
          launch_table_count(g->dg, d_items.as<TItem>(), ni, d_rs.as<uint64_t>(), n_rows,
                             cand_pool.as<uint32_t>(), bitmaps.as<uint32_t>(), wps,
                             d_deg.as<uint64_t>(), d_long.as<uint64_t>(), d_nlong.as<unsigned long long>(),
                             ctx->d_kbytes + kKTableCount, s()));
    uint64_t n_long = 0;
    HIP_TRY(ctx, hipMemcpyAsync(&n_long, d_nlong.p, 8, hipMemcpyDeviceToHost, s()));
    HIP_TRY(ctx, hipStreamSynchronize(s()));
    TIMED(ctx, kKTableCount,
          launch_table_long(0, g->dg, d_items.as<TItem>(), ni, d_rs.as<uint64_t>(), d_long.as<uint64_t>(), n_long,
                            cand_pool.as<uint32_t>(), bitmaps.as<uint32_t>(), nullptr, wps, d_deg.as<uint64_t>(),
                            nullptr, nullptr, ctx->d_kbytes + kKTableCount, s()));
    HIP_TRY(ctx, launch_scan_u64(d_deg.as<uint64_t>(), n_rows, row_off.as<uint64_t>(),
                                 d_tmp.as<unsigned long long>(), s()));
    HIP_TRY(ctx, hipMemcpyAsync(&n_idx, row_off.as<uint64_t>() + n_rows, 8, hipMemcpyDeviceToHost, s()));
    HIP_TRY(ctx, hipStreamSynchronize(s()));
    HIP_TRY(ctx, idx.alloc(n_idx * 4));
    TIMED(ctx, kKTableFill,
          launch_table_fill(g->dg, d_items.as<TItem>(), ni, d_rs.as<uint64_t>(), n_rows,
                            cand_pool.as<uint32_t>(), bitmaps.as<uint32_t>(),
                            wprefix.as<uint32_t>(), wps, row_off.as<uint64_t>(),
                            idx.as<uint32_t>(), ctx->d_kbytes + kKTableFill, s()));
    TIMED(ctx, kKTableFill,
          launch_table_long(1, g->dg, d_items.as<TItem>(), ni, d_rs.as<uint64_t>(), d_long.as<uint64_t>(), n_long,
                            cand_pool.as<uint32_t>(), bitmaps.as<uint32_t>(), wprefix.as<uint32_t>(), wps,
                            nullptr, row_off.as<uint64_t>(), idx.as<uint32_t>(), ctx->d_kbytes + kKTableFill,
                            s()));
    std::vector<uint64_t> ends(ni), nnz;
    for (uint32_t i = 0; i < ni; ++i) ends[i] = row_start[i] + items[i].len_a;
    HIP_TRY(ctx, upload(d_ends, ends, s()));
    HIP_TRY(ctx, d_nnz.alloc((size_t)ni * 8));
    HIP_TRY(ctx, launch_span(row_off.as<uint64_t>(), d_rs.as<uint64_t>(), d_ends.as<uint64_t>(), ni,
                             d_nnz.as<uint64_t>(), s()));
    HIP_TRY(ctx, download(nnz, d_nnz.p, ni, s()));
    for (uint32_t i = 0; i < ni; ++i) item_of[i].first->tab_nnz[item_of[i].second] = nnz[i];
    HIP_TRY(ctx, timers_flush(ctx));
    tables_built = true;
    t_table += now_ms() - t0;
    verbose("tables: %u tables, %llu rows, %llu entries, %.1f ms", ni, (unsigned long long)n_rows,
            (unsigned long long)n_idx, now_ms() - t0);
    return PSM_OK;
}

// Host copy of one table (CFL path-estimate input; psm_query_table).
int Batch::host_table(const QueryHost& q, uint32_t a, uint32_t b, std::vector<uint32_t>& off,
                      std::vector<uint32_t>& idxv, hipStream_t st) {
    const uint32_t n = q.qg.n;
    const uint64_t r0 = q.tab_row[(size_t)a * n + b];
    if (r0 == kNoTable) return PSM_EINVAL;
    const uint32_t rows = q.cand_count[a];
    std::vector<uint64_t> ro;
    HIP_TRY(ctx, download(ro, row_off.as<uint64_t>() + r0, rows + 1, st));
    off.resize(rows + 1);
    for (uint32_t i = 0; i <= rows; ++i) off[i] = (uint32_t)(ro[i] - ro[0]);
    HIP_TRY(ctx, download(idxv, idx.as<uint32_t>() + ro[0], ro[rows] - ro[0], st));
    return PSM_OK;
}

// |C(u)| under the NLF filter for one query vertex whose label is replaced by
// `label_override` (the CFL unit start-vertex rule runs on relabelled units).
int Batch::nlf_count(const QGraph& q, uint32_t u, uint32_t label, uint32_t& out, hipStream_t st,
                     uint32_t mode) {
    std::vector<QVert> qv(1);
    std::vector<uint2> qn;
    qv[0].label = label;
    qv[0].degree = q.degree(u);
    qv[0].nlf_begin = 0;
    qv[0].nlf_count = (uint32_t)q.nlf[u].size();
    for (auto& r : q.nlf[u]) qn.push_back(make_uint2(r.first, r.second));
    out = 0;
    if (label >= g->L) return PSM_OK;
    std::vector<CItem> items(1);
    items[0] = CItem{};
    items[0].src_begin = g->h_ri_off[label];
    items[0].len = g->h_ri_off[label + 1] - g->h_ri_off[label];
    items[0].mode = mode;
    const uint32_t blocks = (items[0].len + 255) / 256;
    if (!blocks) return PSM_OK;
    DBuf d_items, d_qv, d_qn, d_masks, d_bc;
    HIP_TRY(ctx, upload(d_items, items, st));
    HIP_TRY(ctx, upload(d_qv, qv, st));
    if (qn.empty()) qn.push_back(make_uint2(0, 0));
    HIP_TRY(ctx, upload(d_qn, qn, st));
    HIP_TRY(ctx, d_masks.alloc((size_t)blocks * 32));
    HIP_TRY(ctx, d_bc.alloc((size_t)blocks * 4));
    HIP_TRY(ctx, launch_compact_flag(g->dg, d_items.as<CItem>(), 1, blocks, d_qv.as<QVert>(),
                                     d_qn.as<uint2>(), nullptr, nullptr,
                                     d_masks.as<unsigned long long>(), d_bc.as<uint32_t>(), st));
    std::vector<uint32_t> bc;
    HIP_TRY(ctx, download(bc, d_bc.p, blocks, st));
    for (uint32_t c : bc) out += c;
    return PSM_OK;
}

// ---- CFL filter (FilterVertices.cpp:188-250) ------------------------------
// Support counters use an epoch per pass so they never need clearing: a
// pass with P pivots moves cnt[w] from base to base+P, base = epoch*64.
// The CFL filter of one query on stream `st` with its own support counters;
// its candidate sets (one per query vertex) go to `sets`. Its steps depend on
// the query's own BFS levels. With `dpiso`, the DPiso filter
// (FilterVertices.cpp:252-294) instead: LDF sets, then three pruning sweeps
// along the DPiso BFS order with the same support and prune kernels.
int Batch::cfl_filter_query(QueryHost* q, hipStream_t st, DBuf& cnt, uint32_t& epoch, DBuf& defer, bool dpiso,
                            std::vector<std::vector<uint32_t>>& sets) {
    tls_stream = st;
    const QGraph& qg = q->qg;
    const uint32_t n = qg.n;
    std::vector<uint32_t> nlf_counts(n, kInvalid);
    auto nlf_of = [&](uint32_t u, uint32_t& c) -> int {
        if (nlf_counts[u] == kInvalid) PSM_TRY(nlf_count(qg, u, qg.lab[u], nlf_counts[u], st));
        c = nlf_counts[u];
        return PSM_OK;
    };
    CflTree t;
    if (dpiso) {
        PSM_TRY(dpiso_filter_plan(qg, g->V, [&](uint32_t u, uint32_t& c) {
            return nlf_count(qg, u, qg.lab[u], c, st, kModeLDF);
        }, t));
    } else {
        PSM_TRY(cfl_filter_plan(qg, g->L, g->h_ri_off, g->max_freq,
                                [&](uint32_t u, uint32_t& c) { return nlf_of(u, c); }, t));
    }
    q->filter_tree = t;
    q->has_filter_tree = true;
    // per-vertex candidate lists live in their own device buffers while
    // the passes run
    std::vector<std::unique_ptr<DBuf>> cb(n);
    std::vector<uint32_t> len(n, 0);
    auto run_compact = [&](uint32_t u, uint32_t mode, uint32_t target) -> int {
        std::vector<QVert> qv(1);
        std::vector<uint2> qn;
        qv[0].label = qg.lab[u];
        qv[0].degree = qg.degree(u);
        qv[0].nlf_begin = 0;
        qv[0].nlf_count = (uint32_t)qg.nlf[u].size();
        for (auto& r : qg.nlf[u]) qn.push_back(make_uint2(r.first, r.second));
        if (qn.empty()) qn.push_back(make_uint2(0, 0));
        std::vector<CItem> items(1);
        items[0] = CItem{};
        const uint32_t l = qg.lab[u];
        items[0].src_begin = l < g->L ? g->h_ri_off[l] : 0;
        items[0].len = l < g->L ? g->h_ri_off[l + 1] - g->h_ri_off[l] : 0;
        items[0].mode = mode;
        items[0].target = target;
        const uint32_t blocks = (items[0].len + 255) / 256;
        cb[u].reset(new DBuf());
        len[u] = 0;
        if (!blocks) { HIP_TRY(ctx, cb[u]->alloc(4)); return PSM_OK; }
        DBuf d_items, d_qv, d_qn, d_masks, d_bc, d_bp, d_tmp;
        HIP_TRY(ctx, upload(d_items, items, st));
        HIP_TRY(ctx, upload(d_qv, qv, st));
        HIP_TRY(ctx, upload(d_qn, qn, st));
        HIP_TRY(ctx, d_masks.alloc((size_t)blocks * 32));
        HIP_TRY(ctx, d_bc.alloc((size_t)blocks * 4));
        HIP_TRY(ctx, d_bp.alloc((size_t)(blocks + 1) * 8));
        HIP_TRY(ctx, d_tmp.alloc(scan_tmp_entries(blocks) * 8));
        HIP_TRY(ctx, launch_compact_flag(g->dg, d_items.as<CItem>(), 1, blocks, d_qv.as<QVert>(),
                                         d_qn.as<uint2>(), nullptr, cnt.as<uint32_t>(),
                                         d_masks.as<unsigned long long>(), d_bc.as<uint32_t>(), st));
        HIP_TRY(ctx, launch_scan_u32(d_bc.as<uint32_t>(), blocks, d_bp.as<uint64_t>(),
                                     d_tmp.as<unsigned long long>(), st));
        uint64_t total = 0;
        HIP_TRY(ctx, hipMemcpyAsync(&total, d_bp.as<uint64_t>() + blocks, 8, hipMemcpyDeviceToHost, st));
        HIP_TRY(ctx, hipStreamSynchronize(st));
        HIP_TRY(ctx, cb[u]->alloc(total * 4));
        HIP_TRY(ctx, launch_compact_scatter(d_items.as<CItem>(), 1, blocks, g->ri.as<uint32_t>(), nullptr,
                                            d_masks.as<unsigned long long>(), d_bp.as<uint64_t>(),
                                            cb[u]->as<uint32_t>(), st));
        len[u] = (uint32_t)total;
        return PSM_OK;
    };
    auto support = [&](uint32_t u, const std::vector<uint32_t>& pivots, uint32_t& target) -> int {
        ++epoch;
        const uint32_t base = epoch * 64;
        for (uint32_t i = 0; i < pivots.size(); ++i) {
            const uint32_t p = pivots[i];
            HIP_TRY(ctx, launch_cfl_support(g->dg, cb[p]->as<uint32_t>(), 0, len[p], qg.lab[u],
                                            qg.degree(u), cnt.as<uint32_t>(),
                                            i == 0 ? 0u : base + i, base + i + 1, defer.as<uint32_t>(),
                                            defer.as<uint32_t>() + g->V, st));
        }
        target = base + (uint32_t)pivots.size();
        return PSM_OK;
    };
    auto prune = [&](uint32_t u, const std::vector<uint32_t>& piv) -> int {
        uint32_t target;
        PSM_TRY(support(u, piv, target));
        HIP_TRY(ctx, launch_cfl_prune(cnt.as<uint32_t>(), cb[u]->as<uint32_t>(), 0, len[u], target, st));
        return PSM_OK;
    };
    auto finish = [&]() -> int {
        HIP_TRY(ctx, hipStreamSynchronize(st));
        sets.assign(n, {});
        for (uint32_t u = 0; u < n; ++u) HIP_TRY(ctx, download(sets[u], cb[u]->p, len[u], st));
        return PSM_OK;
    };
    if (dpiso) {
        // LDFFilter stops at the first empty set (later sets stay empty and
        // nothing is pruned); a vertex without pivots keeps its set
        bool empty = false;
        for (uint32_t u = 0; u < n; ++u) {
            if (empty) { cb[u].reset(new DBuf()); HIP_TRY(ctx, cb[u]->alloc(4)); len[u] = 0; continue; }
            PSM_TRY(run_compact(u, kModeLDF, 0));
            empty = len[u] == 0;
        }
        if (!empty)
            for (uint32_t k = 0; k < 3; ++k) {
                if (k % 2 == 0) {
                    for (uint32_t i = 1; i < n; ++i) {
                        const uint32_t u = t.order[i];
                        if (!t.bn[u].empty()) PSM_TRY(prune(u, t.bn[u]));
                    }
                } else {
                    for (int i = (int)n - 2; i >= 0; --i) {
                        const uint32_t u = t.order[i];
                        if (!t.fn[u].empty()) PSM_TRY(prune(u, t.fn[u]));
                    }
                }
            }
        return finish();
    }
    const uint32_t root = t.order[0];
    PSM_TRY(run_compact(root, kModeNLF, 0));
    for (uint32_t i = 0; i < n; ++i)
        if (!cb[i]) { cb[i].reset(new DBuf()); HIP_TRY(ctx, cb[i]->alloc(4)); }
    const uint32_t nlev = (uint32_t)t.level_off.size() - 1;
    for (uint32_t lv = 1; lv < nlev; ++lv) {
        for (uint32_t j = t.level_off[lv]; j < t.level_off[lv + 1]; ++j) {
            const uint32_t u = t.order[j];
            uint32_t target;
            PSM_TRY(support(u, t.bn[u], target));
            PSM_TRY(run_compact(u, kModeCflGen, target));
        }
        for (int j = (int)t.level_off[lv + 1] - 1; j >= (int)t.level_off[lv]; --j) {
            const uint32_t u = t.order[j];
            if (!t.fn[u].empty()) PSM_TRY(prune(u, t.fn[u]));
        }
    }
    for (int lv = (int)nlev - 2; lv >= 0; --lv)
        for (uint32_t j = t.level_off[lv]; j < t.level_off[lv + 1]; ++j) {
            const uint32_t u = t.order[j];
            if (!t.under[u].empty()) PSM_TRY(prune(u, t.under[u]));
        }
    return finish();
}

// Runs fn(i, worker, stream) for i in [0, count) on up to 8 host threads of
// this context, each with its own HIP stream (queries whose steps depend on
// their own structure run concurrently this way). Returns the first error.
// worker_done(w) runs on the worker's thread before its stream is destroyed
// (buffers tied to that stream are released there).
int Batch::for_each_query(size_t count, const std::function<int(size_t, uint32_t, hipStream_t)>& fn,
                          uint32_t& workers, const std::function<void(uint32_t)>& worker_done) {
    const uint32_t nw = (uint32_t)std::min<size_t>(count, 8);
    workers = nw;
    std::vector<int> rc(nw, PSM_OK);
    std::atomic<size_t> next{0};
    auto worker = [&](uint32_t w) {
        (void)hipSetDevice(ctx->device);
        hipStream_t st = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            rc[w] = set_err(ctx, PSM_ENOMEM, "hipStreamCreate failed");
            return;
        }
        tls_stream = st;
        while (rc[w] == PSM_OK) {
            const size_t i = next.fetch_add(1);
            if (i >= count) break;
            rc[w] = fn(i, w, st);
        }
        (void)hipStreamSynchronize(st);
        if (worker_done) worker_done(w);
        tls_stream = nullptr;
        (void)hipStreamDestroy(st);
    };
    if (nw <= 1) {
        if (nw == 1) worker(0);
    } else {
        std::vector<std::thread> th;
        for (uint32_t w = 0; w < nw; ++w) th.emplace_back(worker, w);
        for (auto& t : th) t.join();
    }
    tls_stream = s();
    for (int r : rc) PSM_TRY(r);
    return PSM_OK;
}

// Queries are independent: they are filtered concurrently, each worker with
// its own support counters.
int Batch::cfl_filter(std::vector<QueryHost*>& active, bool dpiso) {
    std::vector<std::vector<std::vector<uint32_t>>> sets(active.size());
    std::vector<std::unique_ptr<DBuf>> cnts(8), defers(8);
    std::vector<uint32_t> epochs(8, 0);
    uint32_t nw = 0;
    PSM_TRY(for_each_query(active.size(), [&](size_t i, uint32_t w, hipStream_t st) -> int {
        if (!cnts[w]) {
            cnts[w].reset(new DBuf());
            HIP_TRY(ctx, cnts[w]->alloc((size_t)g->V * 4));
            HIP_TRY(ctx, hipMemsetAsync(cnts[w]->p, 0, (size_t)g->V * 4, st));
            defers[w].reset(new DBuf());   // hub pivot candidates + their count
            HIP_TRY(ctx, defers[w]->alloc(((size_t)g->V + 1) * 4));
        }
        const int r = cfl_filter_query(active[i], st, *cnts[w], epochs[w], *defers[w], dpiso, sets[i]);
        tls_stream = st;
        return r;
    }, nw, [&](uint32_t w) {
        cnts[w].reset();
        defers[w].reset();
    }));
    std::vector<uint32_t> all;   // concatenated final candidates
    for (size_t i = 0; i < active.size(); ++i) {
        QueryHost* q = active[i];
        const uint32_t n = q->qg.n;
        q->cand_count.assign(n, 0);
        q->cand_begin.assign(n, 0);
        for (uint32_t u = 0; u < n; ++u) {
            q->cand_begin[u] = all.size();
            q->cand_count[u] = (uint32_t)sets[i][u].size();
            all.insert(all.end(), sets[i][u].begin(), sets[i][u].end());
        }
    }
    for (QueryHost* q : qs)
        if (q->cand_count.size() != q->qg.n) {
            q->cand_count.assign(q->qg.n, 0);
            q->cand_begin.assign(q->qg.n, 0);
        }
    HIP_TRY(ctx, upload(cand_pool, all, s()));
    cand_total = all.size();
    PSM_TRY(compact_pool());
    bitmaps_current = false;
    for (QueryHost* q : qs) {
        q->status = PSM_OK;
        for (uint32_t u = 0; u < q->qg.n; ++u)
            if (q->cand_count[u] == 0) q->status = PSM_EEMPTY;
    }
    return PSM_OK;
}

// ---- plans ----------------------------------------------------------------
int Batch::make_plans(int order_method) {
    tls_stream = s();
    const double t0 = now_ms();
    if (!tables_built) return set_err(ctx, PSM_ESTATE, "build tables before generating a plan");
    if (order_method == PSM_ORDER_GQL) {
        for (QueryHost* q : qs) {
            gql_order(q->qg, q->cand_count, g->V, q->plan);
            backward_neighbours(q->qg, q->plan, false);
            q->has_plan_tree = false;
            q->planned = true;
        }
    } else if (order_method == PSM_ORDER_CFL) {
        // the CFL order reads unit NLF counts and tables from the device:
        // queries are planned concurrently
        uint32_t nw = 0;
        PSM_TRY(for_each_query(qs.size(), [&](size_t i, uint32_t, hipStream_t st) -> int {
            QueryHost* q = qs[i];
            const QGraph ug = unit_graph(q->qg);
            std::vector<uint32_t> cache(ug.n, kInvalid);
            auto nlf_of = [&](uint32_t u, uint32_t& c) -> int {
                if (cache[u] == kInvalid) PSM_TRY(nlf_count(ug, u, ug.lab[u], cache[u], st));
                c = cache[u];
                return PSM_OK;
            };
            CflTree t;
            PSM_TRY(cfl_filter_plan(ug, g->L, g->h_ri_off, g->max_freq, nlf_of, t));
            auto table = [&](uint32_t a, uint32_t b, std::vector<uint32_t>& off,
                             std::vector<uint32_t>& ix) { return host_table(*q, a, b, off, ix, st); };
            PSM_TRY(cfl_order(ug, t, q->cand_count, table, q->plan));
            backward_neighbours(q->qg, q->plan, true);
            q->plan_tree = t;
            q->has_plan_tree = true;
            q->planned = true;
            return PSM_OK;
        }, nw));
    } else if (order_method == PSM_ORDER_DPISO) {
        // StudyPerformance.cpp:1397-1414: the unit's DPiso filtering plan
        // (its start-vertex rule counts LDF candidates of the unit's
        // renumbered labels), its BFS order, plain backward neighbours
        uint32_t nw = 0;
        PSM_TRY(for_each_query(qs.size(), [&](size_t i, uint32_t, hipStream_t st) -> int {
            QueryHost* q = qs[i];
            const QGraph ug = unit_graph(q->qg);
            CflTree t;
            PSM_TRY(dpiso_filter_plan(ug, g->V, [&](uint32_t u, uint32_t& c) {
                return nlf_count(ug, u, ug.lab[u], c, st, kModeLDF);
            }, t));
            dpiso_order(t, q->plan);
            backward_neighbours(q->qg, q->plan, false);
            q->plan_tree = t;
            q->has_plan_tree = true;
            q->planned = true;
            return PSM_OK;
        }, nw));
    } else {
        return set_err(ctx, PSM_EINVAL, "unknown order method");
    }
    t_plan += now_ms() - t0;
    return PSM_OK;
}

// Device plan of one enumeration unit: `order` lists query vertices by depth;
// the operands of a depth are its earlier neighbours over the unit's edges
// (all query edges for a whole query).
void make_plan(const QueryHost& q, const std::vector<uint32_t>& order,
               const std::vector<uint8_t>* unit_edge, uint64_t cap, DPlan& P, std::vector<DOp>& ops) {
    memset(&P, 0, sizeof(P));
    const uint32_t n = q.qg.n;
    const uint32_t m = (uint32_t)order.size();
    P.n = m;
    P.cap = cap ? cap : ~0ull;
    std::vector<uint32_t> pos(n, kInvalid);
    for (uint32_t d = 0; d < m; ++d) pos[order[d]] = d;
    for (uint32_t d = 0; d < m; ++d) {
        const uint32_t u = order[d];
        P.cand_base[d] = q.cand_begin[u];
        P.cand_len[d] = q.cand_count[u];
        P.op_begin[d] = (uint32_t)ops.size();
        for (uint32_t k = q.qg.off[u]; k < q.qg.off[u + 1]; ++k) {
            const uint32_t b = q.qg.adj[k];
            if (pos[b] == kInvalid || pos[b] >= d) continue;
            if (unit_edge && !(*unit_edge)[(size_t)b * n + u]) continue;
            DOp op{};
            op.src_depth = pos[b];
            op.row_base = q.tab_row[(size_t)b * n + u];
            ops.push_back(op);
        }
        P.op_count[d] = (uint32_t)ops.size() - P.op_begin[d];
        uint64_t msk = 0;
        for (uint32_t e = 0; e < d; ++e)
            if (q.qg.lab[order[e]] == q.qg.lab[u]) msk |= 1ull << e;
        P.same_label[d] = msk;
        P.m_needed |= msk;
    }
}

// ---- enumeration ------------------------------------------------------------
// Pre-expansion of the root tasks of the uploaded plans (plans, ops and qcnt
// hold hp.size() queries): root tasks are expanded level by level until
// each query has its share of target_total tasks or reaches the depth limit.
// The final task array goes to d_out (n_out tasks, ntask per query).
int Batch::expand_tasks(const std::vector<DPlan>& hp, uint64_t target_total, DBuf& d_out, uint64_t& n_out,
                        std::vector<uint64_t>& ntask) {
    const uint32_t nq = (uint32_t)hp.size();
    std::vector<DTask> tasks;
    ntask.assign(nq, 0);
    for (uint32_t i = 0; i < nq; ++i) {
        if (hp[i].n == 0 || hp[i].cand_len[0] == 0) continue;
        DTask t{};
        t.q = i;
        t.hi = kAutoHi;
        tasks.push_back(t);
        ntask[i] = 1;
    }
    const uint64_t budget_total = std::max<uint64_t>(target_total * 8, 1u << 22);
    uint32_t active_q = 0;
    for (uint32_t i = 0; i < nq; ++i) active_q += ntask[i] ? 1 : 0;
    const uint64_t per_q_target = active_q ? std::max<uint64_t>(64, target_total / active_q) : 0;

    // The array holds the tasks of the queries still being expanded, sorted
    // by query (children follow their parents). A query whose expansion
    // ends (enough tasks, or the depth limit) has its segment moved once to
    // the final array, so each round reads only the tasks it expands.
    DBuf d_tasks, d_tasks2, d_final, d_cnt, d_pos, d_tmp, d_expand, d_qchild, d_seg;
    HIP_TRY(ctx, upload(d_tasks, tasks, s()));
    uint64_t n_tasks = tasks.size();   // tasks in d_tasks
    uint64_t n_final = 0;              // tasks in d_final
    std::vector<uint64_t> final_cnt(nq, 0);   // tasks per query in d_final
    std::vector<uint64_t> final_at(nq, 0);    // where query i's segment starts in d_final
    HIP_TRY(ctx, d_final.alloc(std::max<uint64_t>(n_tasks, 1) * sizeof(DTask)));
    // moves the segments of the queries with retire[i] from d_tasks to the
    // end of d_final (segments in query order, as in d_tasks)
    auto retire_segments = [&](const std::vector<uint8_t>& retire) -> int {
        std::vector<uint64_t> src, dst;
        uint64_t off = 0, add = 0;
        for (uint32_t i = 0; i < nq; ++i) {
            if (!ntask[i]) continue;
            if (retire[i]) {
                src.push_back(off);
                dst.push_back(add);
                final_at[i] = n_final + add;
                add += ntask[i];
                final_cnt[i] += ntask[i];
            }
            off += ntask[i];
        }
        if (!add) return PSM_OK;
        dst.push_back(add);
        const uint32_t nseg = (uint32_t)src.size();
        std::vector<uint64_t> seg(src);
        seg.insert(seg.end(), dst.begin(), dst.end());
        HIP_TRY(ctx, upload(d_seg, seg, s()));
        HIP_TRY(ctx, d_tasks2.alloc((n_final + add) * sizeof(DTask)));
        if (n_final)
            HIP_TRY(ctx, hipMemcpyAsync(d_tasks2.p, d_final.p, n_final * sizeof(DTask), hipMemcpyDeviceToDevice,
                                        s()));
        HIP_TRY(ctx, launch_gather_tasks(d_tasks.as<DTask>(), d_seg.as<uint64_t>(), d_seg.as<uint64_t>() + nseg,
                                         nseg, add, d_tasks2.as<DTask>() + n_final, s()));
        d_final.swap(d_tasks2);
        n_final += add;
        return PSM_OK;
    };
    const double te0 = now_ms();
    // a query is expanded while it has fewer tasks than its share and some
    // of its tasks are above the depth limit (the kernel leaves the others
    // as they are); each round takes every task at least one level deeper
    std::vector<uint8_t> progress(nq, 1);
    const uint32_t max_rounds = ctx->expand_depth;   // each round assigns at least one level
    for (uint32_t round = 0; round <= max_rounds && n_tasks; ++round) {
        std::vector<uint8_t> ex(nq, 0), retire(nq, 0);
        bool any = false;
        for (uint32_t i = 0; i < nq; ++i) {
            if (!ntask[i]) continue;
            if (round < max_rounds && ntask[i] < per_q_target && progress[i]) {
                // a query with one task may cross a chain of single
                // candidates in this round (k_expand's mode 2)
                ex[i] = ntask[i] == 1 ? 2 : 1;
                any = true;
            } else {
                retire[i] = 1;
            }
        }
        PSM_TRY(retire_segments(retire));
        if (!any) {
            n_tasks = 0;
            break;
        }
        HIP_TRY(ctx, upload(d_expand, ex, s()));
        HIP_TRY(ctx, d_qchild.alloc((size_t)nq * 16));
        HIP_TRY(ctx, hipMemsetAsync(d_qchild.p, 0, (size_t)nq * 16, s()));
        HIP_TRY(ctx, d_cnt.alloc(n_tasks * 4));
        HIP_TRY(ctx, d_pos.alloc((n_tasks + 1) * 8));
        HIP_TRY(ctx, d_tmp.alloc(scan_tmp_entries(n_tasks) * 8));
        TIMED(ctx, kKExpand,
              launch_expand(plans.as<DPlan>(), ops.as<DOp>(), d_tasks.as<DTask>(), (uint32_t)n_tasks,
                            d_expand.as<uint8_t>(), d_cnt.as<uint32_t>(), nullptr, nullptr,
                            qcnt.as<QCounters>(), d_qchild.as<unsigned long long>(), nq, ctx->expand_depth, 0,
                            s()));
        HIP_TRY(ctx, launch_scan_u32(d_cnt.as<uint32_t>(), n_tasks, d_pos.as<uint64_t>(),
                                     d_tmp.as<unsigned long long>(), s()));
        // the tasks per query that replace this round's (one read-back per
        // round; their sum is the size of the next array) and how many of
        // its tasks were expanded
        std::vector<unsigned long long> qchild;
        HIP_TRY(ctx, download(qchild, d_qchild.p, 2 * (size_t)nq, s()));
        uint64_t total = 0;
        for (uint32_t i = 0; i < nq; ++i)
            if (ex[i]) total += qchild[i];
        if (total > budget_total || total > 0x7FFFFFFFull) {
            // too many: the expanding queries keep the tasks they have
            PSM_TRY(retire_segments(ex));
            n_tasks = 0;
            break;
        }
        HIP_TRY(ctx, d_tasks2.alloc(std::max<uint64_t>(total, 1) * sizeof(DTask)));
        TIMED(ctx, kKExpand,
              launch_expand(plans.as<DPlan>(), ops.as<DOp>(), d_tasks.as<DTask>(), (uint32_t)n_tasks,
                            d_expand.as<uint8_t>(), nullptr, d_pos.as<uint64_t>(), d_tasks2.as<DTask>(),
                            qcnt.as<QCounters>(), nullptr, nq, ctx->expand_depth, 1, s()));
        d_tasks.swap(d_tasks2);
        verbose("expand round %u: %llu tasks -> %llu children, %.2f ms so far", round,
                (unsigned long long)n_tasks, (unsigned long long)total, now_ms() - te0);
        for (uint32_t i = 0; i < nq; ++i) {
            if (ex[i]) {
                ntask[i] = qchild[i];
                progress[i] = qchild[nq + i] != 0;
            } else {
                ntask[i] = 0;
            }
        }
        n_tasks = total;
    }
    // The final task array interleaves the queries in chunks of
    // kInterleave tasks, round after round in query order (segments were
    // appended in the order the queries' expansions ended). Lanes claim the
    // array from the front, so every query's search starts at once: in a
    // query-ordered array the last queries of a batch started only after
    // every earlier query's tasks had been claimed, and a capped query that
    // needs a few milliseconds alone finished last.
    {
        constexpr uint64_t kInterleave = 256;
        std::vector<uint64_t> src, dst;
        uint64_t at = 0;
        for (uint64_t r = 0;; ++r) {
            bool any = false;
            for (uint32_t i = 0; i < nq; ++i) {
                const uint64_t from = r * kInterleave;
                if (final_cnt[i] <= from) continue;
                any = true;
                src.push_back(final_at[i] + from);
                dst.push_back(at);
                at += std::min(kInterleave, final_cnt[i] - from);
            }
            if (!any) break;
        }
        dst.push_back(at);
        const uint32_t nseg = (uint32_t)src.size();
        if (nseg > 1) {
            std::vector<uint64_t> seg(src);
            seg.insert(seg.end(), dst.begin(), dst.end());
            HIP_TRY(ctx, upload(d_seg, seg, s()));
            HIP_TRY(ctx, d_tasks.alloc(std::max<uint64_t>(at, 1) * sizeof(DTask)));
            HIP_TRY(ctx, launch_gather_tasks(d_final.as<DTask>(), d_seg.as<uint64_t>(),
                                             d_seg.as<uint64_t>() + nseg, nseg, at, d_tasks.as<DTask>(), s()));
        } else {
            d_tasks.swap(d_final);
        }
    }
    n_out = n_final;
    ntask = final_cnt;
    d_out.swap(d_tasks);
    HIP_TRY(ctx, timers_flush(ctx));
    t_expand += now_ms() - te0;
    verbose("expand: %llu tasks in %.1f ms", (unsigned long long)n_final, now_ms() - te0);
    return PSM_OK;
}


// Root tasks (one per plan) are pre-expanded level by level until each plan
// has enough tasks to spread over the chip, then one persistent DFS kernel
// drains a shared work queue that running lanes refill by splitting their
// ranges. `sink` optionally materialises the embeddings (one plan only).
int Batch::run_plans(std::vector<DPlan> hp, std::vector<DOp> hops, EmitSink sink,
                     std::vector<QCounters>& qc, bool& timed_out) {
    const uint32_t nq = (uint32_t)hp.size();
    if (hops.empty()) hops.push_back(DOp{});
    for (DPlan& P : hp) bind_tables(P);
    HIP_TRY(ctx, upload(plans, hp, s()));
    HIP_TRY(ctx, upload(ops, hops, s()));
    HIP_TRY(ctx, qcnt.alloc(sizeof(QCounters) * nq));
    HIP_TRY(ctx, hipMemsetAsync(qcnt.p, 0, sizeof(QCounters) * nq, s()));

    uint32_t max_n = 1;
    for (uint32_t i = 0; i < nq; ++i) max_n = std::max(max_n, hp[i].n);
    // persistent grid: dfs_waves_per_cu wavefronts per CU; the wavefront
    // kernel takes one task per wavefront, the lane kernel one per lane
    const bool wave_dfs = ctx->dfs_strategy == PSM_DFS_WAVE && sink.rows == nullptr;
    const uint32_t grid_waves_full = (uint32_t)ctx->num_cu * ctx->dfs_waves_per_cu;
    const uint64_t target_total = std::max<uint64_t>(
        64, (uint64_t)((double)grid_waves_full * (wave_dfs ? 1 : 64) * ctx->expand_tasks_per_lane));
    DBuf d_tasks;
    uint64_t n_tasks = 0;
    std::vector<uint64_t> ntask;
    PSM_TRY(expand_tasks(hp, target_total, d_tasks, n_tasks, ntask));

    unsigned int donated = 0;
    timed_out = false;
    if (n_tasks) {
        // queue: the expanded tasks as a fixed array, and a ring for
        // split-off work; ring slots are reused, so its size bounds only the
        // split-off tasks queued at once (splitting runs while the ring is
        // shorter than the idle lanes, or than an eighth of all lanes for a
        // query with few tasks left)
        const uint64_t lanes_all = (uint64_t)grid_waves_full * 64;
        const uint64_t room = ctx->queue_slots ? ctx->queue_slots : std::max<uint64_t>(1u << 18, 4 * lanes_all);
        uint64_t ring = 64;
        while (ring < room) ring <<= 1;
        if (ring > (1ull << 31)) return set_err(ctx, PSM_ENOMEM, "task ring beyond 2^31 slots");
        DBuf d_queue, d_seq, d_q, d_state, d_pending;
        HIP_TRY(ctx, d_queue.alloc(ring * sizeof(DTask)));
        HIP_TRY(ctx, d_seq.alloc(ring * 4));
        HIP_TRY(ctx, launch_ring_init(d_seq.as<uint32_t>(), (uint32_t)ring, 0u, s()));
        DQueue hq{};
        hq.n0 = (unsigned int)n_tasks;
        hq.mask = (unsigned int)(ring - 1);
        hq.claim_div = ctx->claim_div;
        hq.intersect = ctx->intersect;
        hq.notify = sink.rows ? nullptr : notify;
        if (enum_limit_s > 0) {
            const double left_s = enum_limit_s - (now_ms() - t_eval) / 1000.0;
            hq.limit_ticks = (unsigned long long)std::max(1.0, left_s * ctx->wall_khz * 1000.0);
        } else if (enum_limit_s < 0 && ctx->time_limit_s > 0) {
            // what is left of the time limit, counted from the start of filtering
            double left_s = ctx->time_limit_s;
            if (t_begin >= 0) left_s -= (now_ms() - t_begin) / 1000.0;
            hq.limit_ticks = (unsigned long long)std::max(1.0, left_s * ctx->wall_khz * 1000.0);
        }
        HIP_TRY(ctx, d_q.alloc(sizeof(DQueue)));
        HIP_TRY(ctx, hipMemcpyAsync(d_q.p, &hq, sizeof(hq), hipMemcpyHostToDevice, s()));
        const uint32_t gw = grid_waves_full / 4 * 4;
        std::vector<uint32_t> pend(nq);
        for (uint32_t i = 0; i < nq; ++i) pend[i] = (uint32_t)ntask[i];
        HIP_TRY(ctx, upload(d_pending, pend, s()));
        const double td0 = now_ms();
        dfs_launch_ms = td0;
        if (wave_dfs) {
            TIMED(ctx, kKDfs,
                  launch_dfs_wave(plans.as<DPlan>(), ops.as<DOp>(), d_tasks.as<DTask>(), d_queue.as<DTask>(),
                                  d_seq.as<uint32_t>(),
                                  d_q.as<DQueue>(), qcnt.as<QCounters>(), d_pending.as<uint32_t>(), max_n, gw,
                                  s()));
        } else {
            TIMED(ctx, kKDfs,
                  launch_dfs(plans.as<DPlan>(), ops.as<DOp>(), d_tasks.as<DTask>(), d_queue.as<DTask>(),
                             d_seq.as<uint32_t>(),
                             d_q.as<DQueue>(), qcnt.as<QCounters>(), d_pending.as<uint32_t>(), max_n, gw,
                             ctx->block_lanes, sink, s()));
        }
        if (poll && hq.notify) {
            // results of queries that finish early are taken while the launch runs
            hipEvent_t ev = nullptr;
            HIP_TRY(ctx, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
            HIP_TRY(ctx, hipEventRecord(ev, s()));
            hipError_t qe;
            while ((qe = hipEventQuery(ev)) == hipErrorNotReady) {
                poll();
                std::this_thread::sleep_for(std::chrono::microseconds(50));
            }
            (void)hipEventDestroy(ev);
            if (qe != hipSuccess) return hip_fail(ctx, qe, "hipEventQuery", __LINE__);
        }
        HIP_TRY(ctx, hipMemcpyAsync(&hq, d_q.p, sizeof(hq), hipMemcpyDeviceToHost, s()));
        HIP_TRY(ctx, timers_flush(ctx));
        t_enum += now_ms() - td0;
        timed_out = hq.timeout != 0;
        donated = hq.donated;
        dfs_start_tick = hq.start_tick;
        lane_steps += hq.lane_steps;
        lane_iters += hq.lane_iters;
        verbose("dfs: %.1f ms, %u tasks split off (ring %llu slots, %u skipped: slot held), timeout %u, "
                "lanes busy %.3f", now_ms() - td0, donated, (unsigned long long)ring, hq.full, hq.timeout,
                hq.lane_iters ? (double)hq.lane_steps / (double)hq.lane_iters : 0.0);
        // a query with queued or unfinished tasks and no cap reached is incomplete
        std::vector<uint32_t> left;
        HIP_TRY(ctx, download(left, d_pending.p, nq, s()));
        HIP_TRY(ctx, download(qc, qcnt.p, nq, s()));
        for (uint32_t i = 0; i < nq; ++i)
            if (left[i] && !qc[i].done) qc[i].timed_out = 1;
        HIP_TRY(ctx, hipMemcpyAsync(qcnt.p, qc.data(), sizeof(QCounters) * nq, hipMemcpyHostToDevice, s()));
        HIP_TRY(ctx, hipStreamSynchronize(s()));
    }
    tasks_run += n_tasks + donated;
    HIP_TRY(ctx, download(qc, qcnt.p, nq, s()));
    // per-kernel statistics carry the algorithmic bytes of the nodes each
    // kernel entered (what the roofline divides by launch time)
    for (uint32_t i = 0; i < nq; ++i) {
        ctx->kstat[kKDfs].bytes += (double)qc[i].algo_bytes;
        ctx->kstat[kKExpand].bytes += (double)qc[i].expand_algo_bytes;
    }
    return PSM_OK;
}

// enum_limit >= 0 replaces the context's time limit (counted from the start
// of filtering) by one counted from this call (0: none).
int Batch::evaluate(uint64_t cap, double enum_limit) {
    tls_stream = s();
    enum_limit_s = enum_limit;
    t_eval = now_ms();
    for (QueryHost* q : qs)
        if (!q->planned) return set_err(ctx, PSM_ESTATE, "generate a plan before evaluating");
    std::vector<DPlan> hp(qs.size());
    std::vector<DOp> hops;
    for (size_t i = 0; i < qs.size(); ++i) make_plan(*qs[i], qs[i]->plan.order, nullptr, cap, hp[i], hops);
    std::vector<QCounters> qc;
    bool timed_out = false;
    dfs_start_tick = 0;
    PSM_TRY(run_plans(hp, hops, EmitSink{nullptr, nullptr, 0, nullptr}, qc, timed_out));
    const double t_end = now_ms();
    for (size_t i = 0; i < qs.size(); ++i) {
        QueryHost* q = qs[i];
        // the device clock runs at wall_khz ticks per millisecond
        const unsigned long long tick = qc[i].cap_tick ? qc[i].cap_tick : qc[i].end_tick;
        q->done_ms = tick && dfs_start_tick && tick >= dfs_start_tick
                         ? std::min(t_end, dfs_launch_ms + (double)(tick - dfs_start_tick) / ctx->wall_khz)
                         : t_end;
        q->count = cap ? std::min<uint64_t>(cap, qc[i].count) : qc[i].count;
        q->capped = cap && qc[i].count >= cap;
        q->calls = qc[i].calls;
        q->rejects = qc[i].rejects;
        q->probe_bytes = qc[i].probe_bytes;
        q->algo_bytes = qc[i].algo_bytes + qc[i].expand_algo_bytes;
        q->timed_out = qc[i].timed_out != 0;
    }
    if (timed_out) set_err(ctx, PSM_ETIMEOUT, "enumeration passed the time limit");
    return PSM_OK;
}

// ---- two-unit decomposition + hash join --------------------------------------
// Unit A: the first ceil(n/2) vertices of the matching order (a connected
// prefix) with the query edges among them. Unit B: the remaining vertices and
// every query edge with an endpoint outside A, plus the A vertices those edges
// reach (the keys). Every query edge is in exactly one unit, so joining the
// units' embeddings on the keys, with data vertices distinct across the units,
// gives exactly the query's embeddings.
int Batch::join_count(QueryHost& q, uint64_t max_rows, JoinResult& jr) {
    tls_stream = s();
    const uint32_t n = q.qg.n;
    jr = JoinResult();
    if (!q.planned) return set_err(ctx, PSM_ESTATE, "generate a plan before joining");
    const uint32_t k = (n + 1) / 2;
    std::vector<uint8_t> inA(n, 0), inB(n, 0), eA((size_t)n * n, 0), eB((size_t)n * n, 0);
    for (uint32_t d = 0; d < k; ++d) inA[q.plan.order[d]] = 1;
    for (uint32_t u = 0; u < n; ++u)
        for (uint32_t i = q.qg.off[u]; i < q.qg.off[u + 1]; ++i) {
            const uint32_t w = q.qg.adj[i];
            if (inA[u] && inA[w]) eA[(size_t)u * n + w] = 1;
            else { eB[(size_t)u * n + w] = 1; inB[u] = inB[w] = 1; }
        }
    std::vector<uint32_t> orderA(q.plan.order.begin(), q.plan.order.begin() + k), orderB;
    // unit B order: greedy by candidate count, preferring vertices adjacent
    // (over unit-B edges) to those already placed
    {
        std::vector<uint8_t> placed(n, 0);
        uint32_t nb = 0;
        for (uint32_t u = 0; u < n; ++u) nb += inB[u];
        while (orderB.size() < nb) {
            uint32_t best = kInvalid;
            bool best_adj = false;
            for (uint32_t u = 0; u < n; ++u) {
                if (!inB[u] || placed[u]) continue;
                bool adj = false;
                for (uint32_t w : orderB) adj |= eB[(size_t)w * n + u] != 0;
                if (best == kInvalid || (adj && !best_adj) ||
                    (adj == best_adj && q.cand_count[u] < q.cand_count[best])) {
                    best = u;
                    best_adj = adj;
                }
            }
            placed[best] = 1;
            orderB.push_back(best);
        }
    }
    jr.rows_a = jr.rows_b = 0;
    jr.keys = 0;
    if (orderB.empty()) return set_err(ctx, PSM_EINVAL, "query too small to split");
    // key columns (shared vertices, ascending query id) and the rest
    std::vector<uint32_t> posA(n, kInvalid), posB(n, kInvalid);
    for (uint32_t d = 0; d < orderA.size(); ++d) posA[orderA[d]] = d;
    for (uint32_t d = 0; d < orderB.size(); ++d) posB[orderB[d]] = d;
    JoinKeys ka{}, kb{};
    for (uint32_t u = 0; u < n; ++u) {
        if (inA[u] && inB[u]) {
            ka.key[ka.nkeys++] = posA[u];
            kb.key[kb.nkeys++] = posB[u];
        } else if (inA[u]) {
            ka.rest[ka.nrest++] = posA[u];
        } else if (inB[u]) {
            kb.rest[kb.nrest++] = posB[u];
        }
    }
    jr.keys = ka.nkeys;
    // materialise both units, each through one staging buffer of max_rows
    // rows and then kept at its exact size
    DBuf rows[2], cnt, ovf, stage;
    HIP_TRY(ctx, cnt.alloc(8));
    HIP_TRY(ctx, ovf.alloc(4));
    const std::vector<uint32_t>* orders[2] = {&orderA, &orderB};
    const std::vector<uint8_t>* edges[2] = {&eA, &eB};
    uint64_t nrows[2] = {0, 0};
    HIP_TRY(ctx, stage.alloc(std::max<uint64_t>(max_rows, 1) * std::max(orderA.size(), orderB.size()) * 4));
    const double t0 = now_ms();
    for (int side = 0; side < 2; ++side) {
        const uint32_t w = (uint32_t)orders[side]->size();
        HIP_TRY(ctx, hipMemsetAsync(cnt.p, 0, 8, s()));
        HIP_TRY(ctx, hipMemsetAsync(ovf.p, 0, 4, s()));
        std::vector<DPlan> hp(1);
        std::vector<DOp> hops;
        make_plan(q, *orders[side], edges[side], 0, hp[0], hops);
        std::vector<QCounters> qc;
        bool to = false;
        EmitSink sink{stage.as<uint32_t>(), cnt.as<unsigned long long>(), max_rows, ovf.as<uint32_t>()};
        PSM_TRY(run_plans(hp, hops, sink, qc, to));
        uint32_t overflow = 0;
        HIP_TRY(ctx, hipMemcpyAsync(&nrows[side], cnt.p, 8, hipMemcpyDeviceToHost, s()));
        HIP_TRY(ctx, hipMemcpyAsync(&overflow, ovf.p, 4, hipMemcpyDeviceToHost, s()));
        HIP_TRY(ctx, hipStreamSynchronize(s()));
        if (to) return set_err(ctx, PSM_ETIMEOUT, "join unit enumeration passed the time limit");
        if (overflow) return set_err(ctx, PSM_ENOMEM, "join unit has more rows than max_rows");
        HIP_TRY(ctx, rows[side].alloc(std::max<uint64_t>(nrows[side], 1) * w * 4));
        if (nrows[side])
            HIP_TRY(ctx, hipMemcpyAsync(rows[side].p, stage.p, nrows[side] * w * 4, hipMemcpyDeviceToDevice, s()));
    }
    stage.release();
    jr.rows_a = nrows[0];
    jr.rows_b = nrows[1];
    jr.enum_ms = now_ms() - t0;
    // build on the smaller side (hashjoin.cpp:60-69), probe with the other
    const double t1 = now_ms();
    const int bs = nrows[0] <= nrows[1] ? 0 : 1, ps = 1 - bs;
    const JoinKeys& bk = bs == 0 ? ka : kb;
    const JoinKeys& pk = bs == 0 ? kb : ka;
    uint64_t slots = 16;
    while (slots < 2 * nrows[bs]) slots <<= 1;
    DBuf fp, head, next, out;
    HIP_TRY(ctx, fp.alloc(slots * 8));
    HIP_TRY(ctx, hipMemsetAsync(fp.p, 0, slots * 8, s()));
    HIP_TRY(ctx, head.alloc(slots * 4));
    HIP_TRY(ctx, launch_fill_u32(head.as<uint32_t>(), kInvalid, slots, s()));
    HIP_TRY(ctx, next.alloc(std::max<uint64_t>(nrows[bs], 1) * 4));
    HIP_TRY(ctx, out.alloc(8));
    HIP_TRY(ctx, hipMemsetAsync(out.p, 0, 8, s()));
    TIMED(ctx, kKJoin,
          launch_join_build(rows[bs].as<uint32_t>(), nrows[bs], (uint32_t)orders[bs]->size(), bk,
                            fp.as<unsigned long long>(), head.as<uint32_t>(), slots - 1, next.as<uint32_t>(), s()));
    TIMED(ctx, kKJoin,
          launch_join_probe(rows[bs].as<uint32_t>(), (uint32_t)orders[bs]->size(), bk, rows[ps].as<uint32_t>(),
                            nrows[ps], (uint32_t)orders[ps]->size(), pk, fp.as<unsigned long long>(),
                            head.as<uint32_t>(), slots - 1, next.as<uint32_t>(), out.as<unsigned long long>(),
                            ctx->d_kbytes + kKJoin, s()));
    HIP_TRY(ctx, hipMemcpyAsync(&jr.count, out.p, 8, hipMemcpyDeviceToHost, s()));
    HIP_TRY(ctx, timers_flush(ctx));
    jr.join_ms = now_ms() - t1;
    verbose("join: %llu x %llu rows on %u keys -> %llu in %.1f ms", (unsigned long long)nrows[0],
            (unsigned long long)nrows[1], jr.keys, (unsigned long long)jr.count, jr.join_ms);
    return PSM_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// public ABI
// ---------------------------------------------------------------------------
struct psm_query {
    psm_ctx* ctx = nullptr;
    const psm_graph* g = nullptr;
    QueryHost qh;
    Batch batch;
    bool filtered = false;
};

extern "C" {

int psm_ctx_create(int device, psm_ctx** out) {
    if (!out) return PSM_EINVAL;
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) return -(int)e;
    if (device < 0 || device >= n) return PSM_EINVAL;
    e = hipSetDevice(device);
    if (e != hipSuccess) return -(int)e;
    // the allocation-free entry point's scratch: its address on this device
    e = scratch_setup(device);
    if (e != hipSuccess) return -(int)e;
    psm_ctx* c = new psm_ctx();
    c->device = device;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; return -(int)e; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cu = prop.multiProcessorCount;
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
        c->wall_khz = khz;
    const char* vb = getenv("PSM_VERBOSE");
    c->verbose = vb && *vb && *vb != '0';
    e = hipMalloc(&c->d_kbytes, sizeof(unsigned long long) * kKNumKinds);
    if (e == hipSuccess) e = hipMemset(c->d_kbytes, 0, sizeof(unsigned long long) * kKNumKinds);
    if (e != hipSuccess) { (void)hipStreamDestroy(c->stream); delete c; return -(int)e; }
    *out = c;
    return PSM_OK;
}

int psm_ctx_destroy(psm_ctx* ctx) {
    if (!ctx) return PSM_OK;
    (void)hipSetDevice(ctx->device);
    (void)timers_flush(ctx);
    for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
    if (ctx->d_kbytes) (void)hipFree(ctx->d_kbytes);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PSM_OK;
}

const char* psm_last_error(const psm_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int psm_kernel_stats(psm_ctx* ctx, int kind, double* out3) {
    if (!ctx || !out3 || kind < 0 || kind >= kKNumKinds) return PSM_EINVAL;
    HIP_TRY(ctx, timers_flush(ctx));
    out3[0] = ctx->kstat[kind].launches;
    out3[1] = ctx->kstat[kind].ms;
    out3[2] = ctx->kstat[kind].bytes;
    return PSM_OK;
}

int psm_set_time_limit(psm_ctx* ctx, double seconds) {
    if (!ctx || seconds < 0) return PSM_EINVAL;
    ctx->time_limit_s = seconds;
    return PSM_OK;
}

int psm_set_tuning(psm_ctx* ctx, int key, double value) {
    if (!ctx) return PSM_EINVAL;
    switch (key) {
        case PSM_TUNE_DFS_WAVES_PER_CU:
            if (value < 1 || value > 32) return set_err(ctx, PSM_EINVAL, "dfs waves per CU outside [1, 32]");
            ctx->dfs_waves_per_cu = (uint32_t)value;
            return PSM_OK;
        case PSM_TUNE_EXPAND_TASKS_PER_LANE:
            if (!(value > 0 && value <= 64)) return set_err(ctx, PSM_EINVAL, "expand tasks per lane outside (0, 64]");
            ctx->expand_tasks_per_lane = value;
            return PSM_OK;
        case PSM_TUNE_DFS_STRATEGY:
            if (value != PSM_DFS_LANE && value != PSM_DFS_WAVE)
                return set_err(ctx, PSM_EINVAL, "dfs strategy must be PSM_DFS_LANE or PSM_DFS_WAVE");
            ctx->dfs_strategy = (uint32_t)value;
            return PSM_OK;
        case PSM_TUNE_CLAIM_DIVISOR:
            if (!(value >= 1 && value <= 1024)) return set_err(ctx, PSM_EINVAL, "claim divisor outside [1, 1024]");
            ctx->claim_div = (uint32_t)value;
            return PSM_OK;
        case PSM_TUNE_DFS_BLOCK_LANES:
            if (value != 0 && value != 64 && value != 128 && value != 256)
                return set_err(ctx, PSM_EINVAL, "dfs block lanes must be 0, 64, 128 or 256");
            ctx->block_lanes = (uint32_t)value;
            return PSM_OK;
        case PSM_TUNE_EXPAND_DEPTH:
            if (!(value >= 1 && value <= kExpandMax)) return set_err(ctx, PSM_EINVAL, "expand depth outside [1, 31]");
            ctx->expand_depth = (uint32_t)value;
            return PSM_OK;
        case PSM_TUNE_SERVE_RESIDENT:
            if (value != 0 && value != 1) return set_err(ctx, PSM_EINVAL, "serve resident must be 0 or 1");
            ctx->serve_resident = (uint32_t)value;
            return PSM_OK;
        case PSM_TUNE_INTERSECT:
            if (value != PSM_INTERSECT_ADAPTIVE && value != PSM_INTERSECT_BSEARCH)
                return set_err(ctx, PSM_EINVAL, "intersect must be PSM_INTERSECT_ADAPTIVE or PSM_INTERSECT_BSEARCH");
            ctx->intersect = (uint32_t)value;
            return PSM_OK;
        case PSM_TUNE_QUEUE_SLOTS:
            if (!(value >= 0 && value <= (double)(1u << 28)))
                return set_err(ctx, PSM_EINVAL, "queue slots outside [0, 2^28]");
            ctx->queue_slots = (uint32_t)value;
            return PSM_OK;
        default:
            return set_err(ctx, PSM_EINVAL, "unknown tuning key");
    }
}

int psm_get_tuning(psm_ctx* ctx, int key, double* value) {
    if (!ctx || !value) return PSM_EINVAL;
    switch (key) {
        case PSM_TUNE_DFS_WAVES_PER_CU: *value = ctx->dfs_waves_per_cu; return PSM_OK;
        case PSM_TUNE_EXPAND_TASKS_PER_LANE: *value = ctx->expand_tasks_per_lane; return PSM_OK;
        case PSM_TUNE_DFS_STRATEGY: *value = ctx->dfs_strategy; return PSM_OK;
        case PSM_TUNE_QUEUE_SLOTS: *value = ctx->queue_slots; return PSM_OK;
        case PSM_TUNE_CLAIM_DIVISOR: *value = ctx->claim_div; return PSM_OK;
        case PSM_TUNE_DFS_BLOCK_LANES: *value = ctx->block_lanes; return PSM_OK;
        case PSM_TUNE_EXPAND_DEPTH: *value = ctx->expand_depth; return PSM_OK;
        case PSM_TUNE_INTERSECT: *value = ctx->intersect; return PSM_OK;
        case PSM_TUNE_SERVE_RESIDENT: *value = ctx->serve_resident; return PSM_OK;
        default: return set_err(ctx, PSM_EINVAL, "unknown tuning key");
    }
}

int psm_kernel_stats_reset(psm_ctx* ctx) {
    if (!ctx) return PSM_EINVAL;
    HIP_TRY(ctx, timers_flush(ctx));
    for (auto& k : ctx->kstat) k = psm_ctx::KStat();
    return PSM_OK;
}

int psm_graph_upload(psm_ctx* ctx, uint32_t V, const uint32_t* off, const uint32_t* adj,
                     const uint32_t* lab, psm_graph** out) {
    if (!ctx || !out || !off || !lab || (off[V] && !adj)) return PSM_EINVAL;
    *out = nullptr;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    tls_stream = nullptr;   // graph buffers outlive batches: release with a device sync
    std::unique_ptr<psm_graph> g(new psm_graph());
    g->ctx = ctx;
    g->V = V;
    g->nnz = off[V];
    uint32_t maxl = 0;
    for (uint32_t v = 0; v < V; ++v) {
        maxl = std::max(maxl, lab[v]);
        g->max_deg = std::max(g->max_deg, off[v + 1] - off[v]);
        if (off[v + 1] < off[v]) return set_err(ctx, PSM_EINVAL, "offsets must be non-decreasing");
    }
    g->L = V ? maxl + 1 : 0;
    std::vector<uint32_t> ri, nlf_off, adj_bl, nlf_beg;
    std::vector<uint2> nlf;
    build_graph_indexes(V, off, adj, lab, g->L, g->h_ri_off, ri, nlf_off, nlf, adj_bl, nlf_beg);
    std::vector<uint32_t> deg_ri, nlf_ri_off;
    std::vector<uint2> nlf_ri;
    build_ri_order(V, off, ri, nlf_off, nlf, deg_ri, nlf_ri_off, nlf_ri);
    for (uint32_t l = 0; l < g->L; ++l)
        g->max_freq = std::max(g->max_freq, g->h_ri_off[l + 1] - g->h_ri_off[l]);
    hipStream_t s = ctx->stream;
    auto up = [&](DBuf& b, const void* src, size_t bytes) -> hipError_t {
        hipError_t e = b.alloc(bytes);
        if (e != hipSuccess || !bytes) return e;
        return hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, s);
    };
    HIP_TRY(ctx, up(g->off, off, (size_t)(V + 1) * 4));
    HIP_TRY(ctx, up(g->adj, adj, (size_t)g->nnz * 4));
    HIP_TRY(ctx, up(g->lab, lab, (size_t)V * 4));
    HIP_TRY(ctx, up(g->ri_off, g->h_ri_off.data(), g->h_ri_off.size() * 4));
    HIP_TRY(ctx, up(g->ri, ri.data(), ri.size() * 4));
    HIP_TRY(ctx, up(g->nlf_off, nlf_off.data(), nlf_off.size() * 4));
    HIP_TRY(ctx, up(g->nlf, nlf.data(), nlf.size() * sizeof(uint2)));
    HIP_TRY(ctx, up(g->adj_bl, adj_bl.data(), adj_bl.size() * 4));
    HIP_TRY(ctx, up(g->nlf_beg, nlf_beg.data(), nlf_beg.size() * 4));
    HIP_TRY(ctx, up(g->deg_ri, deg_ri.data(), deg_ri.size() * 4));
    HIP_TRY(ctx, up(g->nlf_ri_off, nlf_ri_off.data(), nlf_ri_off.size() * 4));
    HIP_TRY(ctx, up(g->nlf_ri, nlf_ri.data(), nlf_ri.size() * sizeof(uint2)));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    DevGraph& d = g->dg;
    d.V = V;
    d.L = g->L;
    d.nnz = g->nnz;
    d.off = g->off.as<uint32_t>();
    d.adj = g->adj.as<uint32_t>();
    d.lab = g->lab.as<uint32_t>();
    d.ri_off = g->ri_off.as<uint32_t>();
    d.ri = g->ri.as<uint32_t>();
    d.nlf_off = g->nlf_off.as<uint32_t>();
    d.nlf = g->nlf.as<uint2>();
    d.adj_bl = g->adj_bl.as<uint32_t>();
    d.nlf_beg = g->nlf_beg.as<uint32_t>();
    d.deg_ri = g->deg_ri.as<uint32_t>();
    d.nlf_ri_off = g->nlf_ri_off.as<uint32_t>();
    d.nlf_ri = g->nlf_ri.as<uint2>();
    *out = g.release();
    return PSM_OK;
}

int psm_graph_destroy(psm_graph* g) {
    delete g;
    return PSM_OK;
}

int psm_graph_info(const psm_graph* g, uint64_t* info) {
    if (!g || !info) return PSM_EINVAL;
    info[0] = g->V;
    info[1] = g->nnz / 2;
    info[2] = g->L;
    info[3] = g->max_deg;
    info[4] = g->max_freq;
    return PSM_OK;
}

int psm_query_create(psm_ctx* ctx, const psm_graph* g, uint32_t n, const uint32_t* q_off,
                     const uint32_t* q_adj, const uint32_t* q_lab, psm_query** out) {
    if (!ctx || !g || !out || !q_off || !q_lab) return PSM_EINVAL;
    if (n == 0 || n > kMaxQ) return set_err(ctx, PSM_EINVAL, "query size must be in [1, 64]");
    std::unique_ptr<psm_query> q(new psm_query());
    q->ctx = ctx;
    q->g = g;
    int r = make_qgraph(n, q_off, q_adj, q_lab, q->qh.qg);
    if (r != PSM_OK) return set_err(ctx, r, "malformed query graph");
    q->batch.ctx = ctx;
    q->batch.g = g;
    q->batch.qs.push_back(&q->qh);
    *out = q.release();
    return PSM_OK;
}

int psm_query_destroy(psm_query* q) {
    delete q;
    return PSM_OK;
}

int psm_filter(psm_query* q, int method) {
    if (!q) return PSM_EINVAL;
    HIP_TRY(q->ctx, hipSetDevice(q->ctx->device));
    PSM_TRY(q->batch.filter(method));
    q->filtered = true;
    return q->qh.status;
}

int psm_query_candidate_counts(const psm_query* q, uint32_t* counts) {
    if (!q || !counts) return PSM_EINVAL;
    if (!q->filtered) return PSM_ESTATE;
    std::copy(q->qh.cand_count.begin(), q->qh.cand_count.end(), counts);
    return PSM_OK;
}

int psm_query_candidates(const psm_query* q, uint32_t u, uint32_t* out) {
    if (!q || !out || u >= q->qh.qg.n) return PSM_EINVAL;
    if (!q->filtered) return PSM_ESTATE;
    const uint32_t c = q->qh.cand_count[u];
    if (!c) return PSM_OK;
    psm_ctx* ctx = q->ctx;
    HIP_TRY(ctx, hipMemcpyAsync(out, q->batch.cand_pool.as<uint32_t>() + q->qh.cand_begin[u], (size_t)c * 4,
                                hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return PSM_OK;
}

int psm_build_tables(psm_query* q) {
    if (!q) return PSM_EINVAL;
    if (!q->filtered) return set_err(q->ctx, PSM_ESTATE, "filter before building tables");
    HIP_TRY(q->ctx, hipSetDevice(q->ctx->device));
    return q->batch.build_tables();
}

int psm_query_table(const psm_query* q, uint32_t a, uint32_t b, uint32_t* off, uint32_t* idx,
                    uint64_t* nnz) {
    if (!q || a >= q->qh.qg.n || b >= q->qh.qg.n) return PSM_EINVAL;
    if (!q->batch.tables_built) return PSM_ESTATE;
    const uint32_t n = q->qh.qg.n;
    if (q->qh.tab_row[(size_t)a * n + b] == kNoTable) return PSM_EINVAL;
    if (nnz) *nnz = q->qh.tab_nnz[(size_t)a * n + b];
    if (!off && !idx) return PSM_OK;
    std::vector<uint32_t> o, ix;
    PSM_TRY(const_cast<Batch&>(q->batch).host_table(q->qh, a, b, o, ix, q->ctx->stream));
    if (off) std::copy(o.begin(), o.end(), off);
    if (idx) std::copy(ix.begin(), ix.end(), idx);
    return PSM_OK;
}

int psm_memory_cost_bytes(const psm_query* q, uint64_t* bytes) {
    if (!q || !bytes) return PSM_EINVAL;
    if (!q->batch.tables_built) return PSM_ESTATE;
    const QGraph& qg = q->qh.qg;
    const uint32_t n = qg.n;
    uint64_t w = 0;
    for (uint32_t u = 0; u < n; ++u) w += q->qh.cand_count[u];
    for (uint32_t a = 0; a < n; ++a)
        for (uint32_t k = qg.off[a]; k < qg.off[a + 1]; ++k) {
            const uint32_t b = qg.adj[k];
            if (a < b)
                w += q->qh.tab_nnz[(size_t)a * n + b] + q->qh.cand_count[a] +
                     q->qh.tab_nnz[(size_t)b * n + a] + q->qh.cand_count[b];
        }
    *bytes = w * 4;
    return PSM_OK;
}

int psm_set_plan(psm_query* q, const uint32_t* order, const uint32_t* pivot) {
    if (!q || !order) return PSM_EINVAL;
    QueryHost& h = q->qh;
    const uint32_t n = h.qg.n;
    if (!q->batch.tables_built) return set_err(q->ctx, PSM_ESTATE, "build tables before setting a plan");
    std::vector<uint8_t> seen(n, 0);
    for (uint32_t i = 0; i < n; ++i) {
        if (order[i] >= n || seen[order[i]]) return set_err(q->ctx, PSM_EINVAL, "order is not a permutation");
        seen[order[i]] = 1;
    }
    h.plan.order.assign(order, order + n);
    h.plan.pivot.assign(n, 0);
    if (pivot) {
        for (uint32_t i = 1; i < n; ++i) {
            const uint32_t p = pivot[i];
            bool earlier = false;
            for (uint32_t e = 0; e < i; ++e) earlier |= order[e] == p;
            if (p >= n || !earlier || !h.qg.has_edge(p, order[i]))
                return set_err(q->ctx, PSM_EINVAL, "pivot " + std::to_string(i) + " is not an earlier neighbour");
            h.plan.pivot[i] = p;
        }
    }
    backward_neighbours(h.qg, h.plan, false);
    h.has_plan_tree = false;
    h.planned = true;
    return PSM_OK;
}

int psm_query_bfs_tree(const psm_query* q, int which, uint32_t* order, uint32_t* parent, uint32_t* level) {
    if (!q || (which != PSM_TREE_FILTER && which != PSM_TREE_PLAN)) return PSM_EINVAL;
    const QueryHost& h = q->qh;
    const bool has = which == PSM_TREE_FILTER ? h.has_filter_tree : h.has_plan_tree;
    if (!has) return set_err(q->ctx, PSM_ESTATE, "no BFS tree: run a CFL or DPiso filter / order first");
    const CflTree& t = which == PSM_TREE_FILTER ? h.filter_tree : h.plan_tree;
    const uint32_t n = h.qg.n;
    for (uint32_t i = 0; i < n; ++i) {
        if (order) order[i] = t.order[i];
        if (parent) parent[i] = t.parent[i];
        if (level) level[i] = t.level[i];
    }
    return PSM_OK;
}

int psm_generate_plan(psm_query* q, int order_method, uint32_t* order, uint32_t* pivot) {
    if (!q) return PSM_EINVAL;
    HIP_TRY(q->ctx, hipSetDevice(q->ctx->device));
    PSM_TRY(q->batch.make_plans(order_method));
    if (order) std::copy(q->qh.plan.order.begin(), q->qh.plan.order.end(), order);
    if (pivot) std::copy(q->qh.plan.pivot.begin(), q->qh.plan.pivot.end(), pivot);
    return PSM_OK;
}

// ---- query split + join sequence ---------------------------------------------
// Splits (SplitGraph.cpp:874-980 restated as greedy covers): vertices in id
// order take each of their query edges not yet taken, so every query edge
// lies in exactly one unit. STAR: a vertex's taken edges form one star
// (centre + leaves). TWINTWIG: its leaves, by ascending candidate count, are
// cut into twin twigs of at most two edges. SEED: a twin twig whose two
// leaves share a query edge not yet taken takes that edge too (a triangle).
// Each unit is enumerated on the device into rows (centre first); the units
// are then joined one at a time (JoinCollect::sequence_collect,
// JoinCollect.cpp:60-153): start from the unit with the fewest rows, and
// join next the unit sharing a vertex with the joined ones that has the
// fewest rows (more shared vertices on ties), on the shared vertices
// (ascending query id, keyTraits.h) with data vertices distinct across the
// two sides. Each join counts its matches first; an intermediate result is
// then materialised at exactly that size, the last join only counts. The
// count equals the plain count.
int Batch::split_join_count(QueryHost& q, int pattern, uint64_t max_rows, SplitJoinResult& jr) {
    tls_stream = s();
    const uint32_t n = q.qg.n;
    jr = SplitJoinResult();
    if (!tables_built) return set_err(ctx, PSM_ESTATE, "build tables before joining");
    struct Unit {
        std::vector<uint32_t> cols;   // query vertex of each row column (= enumeration order)
        std::vector<uint8_t> edge;    // n*n: the unit's query edges
        std::unique_ptr<DBuf> rows;
        uint64_t nrows = 0;
    };
    if (pattern != PSM_SPLIT_STAR && pattern != PSM_SPLIT_TWINTWIG && pattern != PSM_SPLIT_SEED)
        return set_err(ctx, PSM_EINVAL, "unknown split pattern");
    std::vector<Unit> units;
    std::vector<uint8_t> taken((size_t)n * n, 0);
    auto take = [&](Unit& un, uint32_t a, uint32_t b) {
        taken[(size_t)a * n + b] = taken[(size_t)b * n + a] = 1;
        un.edge[(size_t)a * n + b] = un.edge[(size_t)b * n + a] = 1;
    };
    for (uint32_t u = 0; u < n; ++u) {
        std::vector<uint32_t> leaves;
        for (uint32_t i = q.qg.off[u]; i < q.qg.off[u + 1]; ++i) {
            const uint32_t w = q.qg.adj[i];
            if (!taken[(size_t)u * n + w]) leaves.push_back(w);
        }
        if (leaves.empty()) continue;
        std::stable_sort(leaves.begin(), leaves.end(),
                         [&](uint32_t a, uint32_t b) { return q.cand_count[a] < q.cand_count[b]; });
        const size_t per = pattern == PSM_SPLIT_STAR ? leaves.size() : 2;
        for (size_t i0 = 0; i0 < leaves.size(); i0 += per) {
            Unit un;
            un.rows.reset(new DBuf());
            un.edge.assign((size_t)n * n, 0);
            un.cols.push_back(u);
            for (size_t i = i0; i < std::min(leaves.size(), i0 + per); ++i) {
                un.cols.push_back(leaves[i]);
                take(un, u, leaves[i]);
            }
            if (pattern == PSM_SPLIT_SEED && un.cols.size() == 3) {
                const uint32_t a = un.cols[1], b = un.cols[2];
                if (q.qg.has_edge(a, b) && !taken[(size_t)a * n + b]) take(un, a, b);
            }
            units.push_back(std::move(un));
        }
    }
    if (units.empty()) return set_err(ctx, PSM_EINVAL, "query has no edge to split");
    jr.units = (uint32_t)units.size();
    DBuf cnt, ovf, stage;
    HIP_TRY(ctx, cnt.alloc(8));
    HIP_TRY(ctx, ovf.alloc(4));
    // units are enumerated into one staging buffer of max_rows rows and kept
    // at their exact size
    size_t max_w = 0;
    for (const Unit& un : units) max_w = std::max(max_w, un.cols.size());
    HIP_TRY(ctx, stage.alloc(std::max<uint64_t>(max_rows, 1) * max_w * 4));
    const double t0 = now_ms();
    for (Unit& un : units) {
        const uint32_t w = (uint32_t)un.cols.size();
        HIP_TRY(ctx, hipMemsetAsync(cnt.p, 0, 8, s()));
        HIP_TRY(ctx, hipMemsetAsync(ovf.p, 0, 4, s()));
        std::vector<DPlan> hp(1);
        std::vector<DOp> hops;
        make_plan(q, un.cols, &un.edge, 0, hp[0], hops);
        std::vector<QCounters> qc;
        bool to = false;
        EmitSink sink{stage.as<uint32_t>(), cnt.as<unsigned long long>(), max_rows, ovf.as<uint32_t>()};
        PSM_TRY(run_plans(hp, hops, sink, qc, to));
        uint32_t overflow = 0;
        HIP_TRY(ctx, hipMemcpyAsync(&un.nrows, cnt.p, 8, hipMemcpyDeviceToHost, s()));
        HIP_TRY(ctx, hipMemcpyAsync(&overflow, ovf.p, 4, hipMemcpyDeviceToHost, s()));
        HIP_TRY(ctx, hipStreamSynchronize(s()));
        if (to) return set_err(ctx, PSM_ETIMEOUT, "unit enumeration passed the time limit");
        if (overflow) return set_err(ctx, PSM_ENOMEM, "a unit has more rows than max_rows");
        HIP_TRY(ctx, un.rows->alloc(std::max<uint64_t>(un.nrows, 1) * w * 4));
        if (un.nrows)
            HIP_TRY(ctx, hipMemcpyAsync(un.rows->p, stage.p, un.nrows * w * 4, hipMemcpyDeviceToDevice, s()));
        jr.unit_rows += un.nrows;
        jr.calls += qc[0].calls;
    }
    stage.release();
    jr.enum_ms = now_ms() - t0;
    const double t1 = now_ms();
    // the joined result so far
    std::vector<uint8_t> used(units.size(), 0);
    size_t first = 0;
    for (size_t i = 1; i < units.size(); ++i)
        if (units[i].nrows < units[first].nrows) first = i;
    used[first] = 1;
    std::vector<uint32_t> cols = units[first].cols;
    std::unique_ptr<DBuf> cur;      // owned intermediate rows (none while the first unit is the result)
    const uint32_t* cur_rows = units[first].rows->as<uint32_t>();
    uint64_t cur_n = units[first].nrows;
    if (units.size() == 1) jr.count = cur_n;
    for (size_t step = 1; step < units.size(); ++step) {
        std::vector<uint8_t> in_cur(n, 0);
        for (uint32_t c : cols) in_cur[c] = 1;
        size_t nxt = units.size();
        uint32_t nxt_shared = 0;
        for (size_t i = 0; i < units.size(); ++i) {
            if (used[i]) continue;
            uint32_t sh = 0;
            for (uint32_t c : units[i].cols) sh += in_cur[c];
            if (!sh) continue;
            if (nxt == units.size() || units[i].nrows < units[nxt].nrows ||
                (units[i].nrows == units[nxt].nrows && sh > nxt_shared)) {
                nxt = i;
                nxt_shared = sh;
            }
        }
        if (nxt == units.size()) return set_err(ctx, PSM_EINVAL, "query is not connected");
        used[nxt] = 1;
        Unit& un = units[nxt];
        const bool last = step + 1 == units.size();
        if (cur_n == 0 || un.nrows == 0) {   // nothing can match
            jr.count = 0;
            break;
        }
        // key columns on both sides, ascending query vertex id
        JoinKeys kc{}, ku{};
        std::vector<uint32_t> pos_c(n, kInvalid), pos_u(n, kInvalid);
        for (uint32_t i = 0; i < cols.size(); ++i) pos_c[cols[i]] = i;
        for (uint32_t i = 0; i < un.cols.size(); ++i) pos_u[un.cols[i]] = i;
        for (uint32_t v = 0; v < n; ++v) {
            if (pos_c[v] != kInvalid && pos_u[v] != kInvalid) {
                kc.key[kc.nkeys++] = pos_c[v];
                ku.key[ku.nkeys++] = pos_u[v];
            } else if (pos_c[v] != kInvalid) {
                kc.rest[kc.nrest++] = pos_c[v];
            } else if (pos_u[v] != kInvalid) {
                ku.rest[ku.nrest++] = pos_u[v];
            }
        }
        // build on the smaller side (hashjoin.cpp:60-69)
        const bool build_cur = cur_n <= un.nrows;
        const uint32_t* b_rows = build_cur ? cur_rows : un.rows->as<uint32_t>();
        const uint32_t* p_rows = build_cur ? un.rows->as<uint32_t>() : cur_rows;
        const uint64_t b_n = build_cur ? cur_n : un.nrows, p_n = build_cur ? un.nrows : cur_n;
        const uint32_t b_w = build_cur ? (uint32_t)cols.size() : (uint32_t)un.cols.size();
        const uint32_t p_w = build_cur ? (uint32_t)un.cols.size() : (uint32_t)cols.size();
        const JoinKeys& bk = build_cur ? kc : ku;
        const JoinKeys& pk = build_cur ? ku : kc;
        const std::vector<uint32_t>& b_cols = build_cur ? cols : un.cols;
        const std::vector<uint32_t>& p_cols = build_cur ? un.cols : cols;
        if (b_n >= kInvalid) return set_err(ctx, PSM_ENOMEM, "join side above 2^32 rows");
        uint64_t slots = 16;
        while (slots < 2 * b_n) slots <<= 1;
        DBuf fp, head, next;
        HIP_TRY(ctx, fp.alloc(slots * 8));
        HIP_TRY(ctx, hipMemsetAsync(fp.p, 0, slots * 8, s()));
        HIP_TRY(ctx, head.alloc(slots * 4));
        HIP_TRY(ctx, launch_fill_u32(head.as<uint32_t>(), kInvalid, slots, s()));
        HIP_TRY(ctx, next.alloc(b_n * 4));
        TIMED(ctx, kKJoin,
              launch_join_build(b_rows, b_n, b_w, bk, fp.as<unsigned long long>(), head.as<uint32_t>(), slots - 1,
                                next.as<uint32_t>(), s()));
        // count the matches (the last join stops here), then materialise
        // exactly that many rows
        HIP_TRY(ctx, hipMemsetAsync(cnt.p, 0, 8, s()));
        TIMED(ctx, kKJoin,
              launch_join_probe(b_rows, b_w, bk, p_rows, p_n, p_w, pk, fp.as<unsigned long long>(),
                                head.as<uint32_t>(), slots - 1, next.as<uint32_t>(), cnt.as<unsigned long long>(),
                                ctx->d_kbytes + kKJoin, s()));
        uint64_t matches = 0;
        HIP_TRY(ctx, hipMemcpyAsync(&matches, cnt.p, 8, hipMemcpyDeviceToHost, s()));
        HIP_TRY(ctx, hipStreamSynchronize(s()));
        if (last) {
            jr.count = matches;
            break;
        }
        if (matches > max_rows) return set_err(ctx, PSM_ENOMEM, "an intermediate join result has more rows than max_rows");
        if (matches == 0) {
            jr.count = 0;
            break;
        }
        // output columns: the probe row, then the build row's non-key columns
        std::vector<uint32_t> out_cols = p_cols;
        for (uint32_t i = 0; i < bk.nrest; ++i) out_cols.push_back(b_cols[bk.rest[i]]);
        std::unique_ptr<DBuf> out(new DBuf());
        HIP_TRY(ctx, out->alloc(matches * out_cols.size() * 4));
        HIP_TRY(ctx, hipMemsetAsync(cnt.p, 0, 8, s()));
        HIP_TRY(ctx, hipMemsetAsync(ovf.p, 0, 4, s()));
        TIMED(ctx, kKJoin,
              launch_join_probe_emit(b_rows, b_w, bk, p_rows, p_n, p_w, pk, fp.as<unsigned long long>(),
                                     head.as<uint32_t>(), slots - 1, next.as<uint32_t>(), out->as<uint32_t>(),
                                     matches, cnt.as<unsigned long long>(), ovf.as<uint32_t>(), s()));
        uint64_t out_n = 0;
        uint32_t overflow = 0;
        HIP_TRY(ctx, hipMemcpyAsync(&out_n, cnt.p, 8, hipMemcpyDeviceToHost, s()));
        HIP_TRY(ctx, hipMemcpyAsync(&overflow, ovf.p, 4, hipMemcpyDeviceToHost, s()));
        HIP_TRY(ctx, hipStreamSynchronize(s()));
        if (overflow || out_n != matches) return set_err(ctx, PSM_EINVAL, "join probe passes disagree");
        jr.max_inter_rows = std::max(jr.max_inter_rows, out_n);
        cur = std::move(out);
        cur_rows = cur->as<uint32_t>();
        cur_n = out_n;
        cols = out_cols;
    }
    HIP_TRY(ctx, timers_flush(ctx));
    jr.join_ms = now_ms() - t1;
    verbose("split join: %u units, %llu unit rows, largest intermediate %llu -> %llu in %.1f ms", jr.units,
            (unsigned long long)jr.unit_rows, (unsigned long long)jr.max_inter_rows,
            (unsigned long long)jr.count, jr.join_ms);
    return PSM_OK;
}

int psm_evaluate(psm_query* q, uint64_t cap, psm_enum_stats* st) {
    if (!q) return PSM_EINVAL;
    HIP_TRY(q->ctx, hipSetDevice(q->ctx->device));
    q->batch.t_enum = q->batch.t_expand = 0;
    PSM_TRY(q->batch.evaluate(cap));
    const int r = q->qh.timed_out ? PSM_ETIMEOUT : PSM_OK;
    if (st) {
        memset(st, 0, sizeof(*st));
        st->count = q->qh.count;
        st->calls = q->qh.calls;
        st->capped = q->qh.capped;
        st->tasks = (uint32_t)q->batch.tasks_run;
        st->kernel_ms = q->batch.t_enum;
        st->expand_ms = q->batch.t_expand;
        st->probe_bytes = q->qh.probe_bytes;
        st->algo_bytes = q->qh.algo_bytes;
        st->visited_rejects = q->qh.rejects;
    }
    return r;
}

// psm_evaluate through the fixed device scratch (device.hip): no device
// memory is allocated or freed. The plan, operands, counters, task ring and
// root tasks live in the scratch of the context's GPU; the query's
// candidates and tables (its inputs, built by psm_filter / psm_build_tables)
// are read where they are. Root range tasks replace the pre-expansion
// kernels (which size their buffers per round); the persistent DFS splits
// work off through the ring as psm_evaluate's does.
int psm_evaluate_noalloc(psm_query* q, uint64_t cap, psm_enum_stats* st) {
    if (!q) return PSM_EINVAL;
    psm_ctx* ctx = q->ctx;
    Batch& b = q->batch;
    QueryHost& h = q->qh;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (!h.planned) return set_err(ctx, PSM_ESTATE, "generate a plan before evaluating");
    unsigned char* scr = static_cast<unsigned char*>(scratch_address(ctx->device));
    if (!scr) return set_err(ctx, PSM_ESTATE, "device scratch not set up on this device");
    const double t0 = now_ms();
    DPlan P;
    std::vector<DOp> hops;
    hops.reserve(kScratchOps);
    make_plan(h, h.plan.order, nullptr, cap, P, hops);
    b.bind_tables(P);
    if (hops.size() > kScratchOps)
        return set_err(ctx, PSM_EINVAL, "query needs " + std::to_string(hops.size()) +
                                            " intersection operands; the device scratch holds " +
                                            std::to_string(kScratchOps));
    if (P.n == 0 || P.n > kMaxQ) return set_err(ctx, PSM_EINVAL, "query size outside the device scratch");
    if (hops.empty()) hops.push_back(DOp{});
    const uint64_t roots = P.cand_len[0];
    h.count = h.calls = h.rejects = h.probe_bytes = h.algo_bytes = 0;
    h.capped = h.timed_out = false;
    if (roots == 0) {
        if (st) memset(st, 0, sizeof(*st));
        return PSM_OK;
    }
    const uint32_t n_tasks = (uint32_t)std::min<uint64_t>(roots, kScratchRootTasks);
    DQueue* dq = reinterpret_cast<DQueue*>(scr + kScrQueue);
    QCounters* dqc = reinterpret_cast<QCounters*>(scr + kScrCounters);
    uint32_t* dpend = reinterpret_cast<uint32_t*>(scr + kScrPending);
    DPlan* dplan = reinterpret_cast<DPlan*>(scr + kScrPlan);
    DOp* dops = reinterpret_cast<DOp*>(scr + kScrOps);
    uint32_t* dseq = reinterpret_cast<uint32_t*>(scr + kScrSeq);
    DTask* dring = reinterpret_cast<DTask*>(scr + kScrRing);
    DTask* dinit = reinterpret_cast<DTask*>(scr + kScrInit);
    // the header, counters and pending count start at zero (a fill kernel,
    // no memset node); the scratch keeps the last call's contents otherwise
    HIP_TRY(ctx, launch_fill_u32(reinterpret_cast<uint32_t*>(scr), 0u, kScrZeroEnd / 4, ctx->stream));
    DQueue hq{};
    hq.n0 = n_tasks;
    hq.mask = kScratchSlots - 1;
    hq.claim_div = ctx->claim_div;
    hq.intersect = ctx->intersect;
    if (ctx->time_limit_s > 0) {
        double left_s = ctx->time_limit_s;
        if (b.t_begin >= 0) left_s -= (now_ms() - b.t_begin) / 1000.0;
        hq.limit_ticks = (unsigned long long)std::max(1.0, left_s * ctx->wall_khz * 1000.0);
    }
    const uint32_t pend = n_tasks;
    HIP_TRY(ctx, hipMemcpyAsync(dq, &hq, sizeof(hq), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(dpend, &pend, 4, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(dplan, &P, sizeof(P), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(dops, hops.data(), hops.size() * sizeof(DOp), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, launch_root_tasks(dinit, n_tasks, 0, roots, ctx->stream));
    HIP_TRY(ctx, launch_ring_init(dseq, kScratchSlots, 0u, ctx->stream));
    const uint32_t gw = (uint32_t)ctx->num_cu * ctx->dfs_waves_per_cu / 4 * 4;
    const double tk = now_ms();
    if (ctx->dfs_strategy == PSM_DFS_WAVE) {
        HIP_TRY(ctx, launch_dfs_wave(dplan, dops, dinit, dring, dseq, dq, dqc, dpend, P.n, gw, ctx->stream));
    } else {
        HIP_TRY(ctx, launch_dfs(dplan, dops, dinit, dring, dseq, dq, dqc, dpend, P.n, gw, ctx->block_lanes,
                                EmitSink{nullptr, nullptr, 0, nullptr}, ctx->stream));
    }
    QCounters qc{};
    uint32_t left = 0;
    HIP_TRY(ctx, hipMemcpyAsync(&hq, dq, sizeof(hq), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(&qc, dqc, sizeof(qc), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(&left, dpend, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    const double t1 = now_ms();
    h.count = cap ? std::min<uint64_t>(cap, qc.count) : qc.count;
    h.capped = cap && qc.count >= cap;
    h.calls = qc.calls;
    h.rejects = qc.rejects;
    h.probe_bytes = qc.probe_bytes;
    h.algo_bytes = qc.algo_bytes;
    h.timed_out = left && !qc.done;
    if (st) {
        memset(st, 0, sizeof(*st));
        st->count = h.count;
        st->calls = h.calls;
        st->capped = h.capped;
        st->tasks = n_tasks + hq.donated;
        st->kernel_ms = t1 - tk;
        st->expand_ms = tk - t0;
        st->probe_bytes = h.probe_bytes;
        st->algo_bytes = h.algo_bytes;
        st->visited_rejects = h.rejects;
    }
    if (h.timed_out) {
        set_err(ctx, PSM_ETIMEOUT, "enumeration passed the time limit");
        return PSM_ETIMEOUT;
    }
    return PSM_OK;
}

int psm_split_join_count(psm_query* q, int pattern, uint64_t max_rows, psm_split_join_stats* st) {
    if (!q || !st) return PSM_EINVAL;
    HIP_TRY(q->ctx, hipSetDevice(q->ctx->device));
    SplitJoinResult jr;
    PSM_TRY(q->batch.split_join_count(q->qh, pattern, max_rows ? max_rows : (64ull << 20), jr));
    memset(st, 0, sizeof(*st));
    st->count = jr.count;
    st->units = jr.units;
    st->unit_rows = jr.unit_rows;
    st->max_inter_rows = jr.max_inter_rows;
    st->calls = jr.calls;
    st->enum_ms = jr.enum_ms;
    st->join_ms = jr.join_ms;
    return PSM_OK;
}

int psm_join_count(psm_query* q, uint64_t max_rows, psm_join_stats* st) {
    if (!q || !st) return PSM_EINVAL;
    HIP_TRY(q->ctx, hipSetDevice(q->ctx->device));
    JoinResult jr;
    PSM_TRY(q->batch.join_count(q->qh, max_rows ? max_rows : (64ull << 20), jr));
    st->count = jr.count;
    st->rows_a = jr.rows_a;
    st->rows_b = jr.rows_b;
    st->keys = jr.keys;
    st->enum_ms = jr.enum_ms;
    st->join_ms = jr.join_ms;
    return PSM_OK;
}

int psm_match(psm_query* q, int filter, int order, uint64_t cap, psm_enum_stats* st) {
    int r = psm_filter(q, filter);
    if (r != PSM_OK && r != PSM_EEMPTY) return r;
    PSM_TRY(psm_build_tables(q));
    PSM_TRY(psm_generate_plan(q, order, nullptr, nullptr));
    return psm_evaluate(q, cap, st);
}

}  // extern "C"

struct psm_batch {
    psm_ctx* ctx = nullptr;
    std::vector<std::unique_ptr<QueryHost>> hosts;
    Batch b;
    bool prepared = false;
    double t0 = 0;
};

extern "C" {

int psm_batch_create(psm_ctx* ctx, const psm_graph* g, uint32_t nq, const uint32_t* n, const uint64_t* off_base,
                     const uint64_t* nbr_base, const uint64_t* lab_base, const uint32_t* q_off,
                     const uint32_t* q_adj, const uint32_t* q_lab, psm_batch** out) {
    if (!ctx || !g || !n || !off_base || !nbr_base || !lab_base || !q_off || !q_lab || !out) return PSM_EINVAL;
    *out = nullptr;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    std::unique_ptr<psm_batch> B(new psm_batch());
    B->ctx = ctx;
    B->t0 = now_ms();
    B->hosts.resize(nq);
    B->b.ctx = ctx;
    B->b.g = g;
    for (uint32_t i = 0; i < nq; ++i) {
        B->hosts[i].reset(new QueryHost());
        const int r = make_qgraph(n[i], q_off + off_base[i], q_adj + nbr_base[i], q_lab + lab_base[i],
                                  B->hosts[i]->qg);
        if (r != PSM_OK) return set_err(ctx, r, "malformed query graph " + std::to_string(i));
        B->b.qs.push_back(B->hosts[i].get());
    }
    *out = B.release();
    return PSM_OK;
}

int psm_batch_prepare(psm_batch* B, int filter, int order) {
    if (!B) return PSM_EINVAL;
    HIP_TRY(B->ctx, hipSetDevice(B->ctx->device));
    B->prepared = false;
    PSM_TRY(B->b.filter(filter));
    PSM_TRY(B->b.build_tables());
    PSM_TRY(B->b.make_plans(order));
    B->prepared = true;
    return PSM_OK;
}

int psm_batch_evaluate(psm_batch* B, uint64_t cap, double enum_time_limit_s, uint64_t* counts, uint64_t* calls,
                       uint32_t* status, psm_batch_stats* st) {
    if (!B) return PSM_EINVAL;
    if (!B->prepared) return set_err(B->ctx, PSM_ESTATE, "prepare the batch before evaluating");
    HIP_TRY(B->ctx, hipSetDevice(B->ctx->device));
    Batch& b = B->b;
    const double e0 = b.t_enum + b.t_expand;
    const uint64_t steps0 = b.lane_steps, iters0 = b.lane_iters, tasks0 = b.tasks_run;
    PSM_TRY(b.evaluate(cap, enum_time_limit_s));
    for (size_t i = 0; i < B->hosts.size(); ++i) {
        const QueryHost* h = B->hosts[i].get();
        if (counts) counts[i] = h->count;
        if (calls) calls[i] = h->calls;
        if (status) status[i] = h->timed_out ? PSM_ETIMEOUT : h->status;
    }
    if (st) {
        memset(st, 0, sizeof(*st));
        st->filter_ms = b.t_filter;
        st->table_ms = b.t_table;
        st->plan_ms = b.t_plan;
        st->enum_ms = b.t_enum + b.t_expand - e0;
        st->total_ms = now_ms() - B->t0;
        st->tasks = b.tasks_run - tasks0;
        st->dfs_lane_steps = b.lane_steps - steps0;
        st->dfs_lane_iters = b.lane_iters - iters0;
        for (auto& h : B->hosts) {
            st->probe_bytes += h->probe_bytes;
            st->algo_bytes += h->algo_bytes;
        }
    }
    return PSM_OK;
}

int psm_batch_latencies(const psm_batch* B, double* ms) {
    if (!B || !ms) return PSM_EINVAL;
    for (size_t i = 0; i < B->hosts.size(); ++i) {
        const QueryHost* h = B->hosts[i].get();
        ms[i] = h->done_ms >= 0 ? h->done_ms - B->t0 : -1.0;
    }
    return PSM_OK;
}

int psm_batch_destroy(psm_batch* B) {
    delete B;
    return PSM_OK;
}

int psm_batch_match(psm_ctx* ctx, const psm_graph* g, uint32_t nq, const uint32_t* n,
                    const uint64_t* off_base, const uint64_t* nbr_base, const uint64_t* lab_base,
                    const uint32_t* q_off, const uint32_t* q_adj, const uint32_t* q_lab, int filter,
                    int order, uint64_t cap, uint64_t* counts, uint64_t* calls, uint32_t* status,
                    psm_batch_stats* st) {
    psm_batch* B = nullptr;
    PSM_TRY(psm_batch_create(ctx, g, nq, n, off_base, nbr_base, lab_base, q_off, q_adj, q_lab, &B));
    std::unique_ptr<psm_batch> own(B);
    PSM_TRY(psm_batch_prepare(B, filter, order));
    return psm_batch_evaluate(B, cap, -1.0, counts, calls, status, st);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Serving: continuous batching over a resident data graph. Worker threads,
// each with its own context and HIP stream, take the oldest pending requests
// as one batch while slots and the memory budget allow, and prepare it:
// filter, tables, plans and the pre-expansion of its tasks. A prepared batch
// is admitted into the resident enumeration launch, one persistent lane-DFS
// launch that keeps running while batches are published into it (a segment
// each): its search starts as soon as its preparation ends, beside the
// searches of the batches before it, and each query's result is returned
// once its own count is final. Device memory that the launch reads (the
// batches' candidates, tables and tasks) is recycled when the launch ends:
// the host closes it when it has nothing left to run or memory runs short,
// and the next batch starts a new one. A batch whose queries exceed the
// resident launch's size runs its own enumeration launch instead, as it does
// with the resident launch switched off (PSM_TUNE_SERVE_RESIDENT 0).
// No counterpart in the reference, which runs one query per process.
// ---------------------------------------------------------------------------
#include <condition_variable>
#include <deque>

namespace {

struct Request {
    uint64_t id = 0;
    std::unique_ptr<QueryHost> q;
    uint64_t est_bytes = 0;
    double submit_ms = 0;
};

constexpr uint32_t kEngSlots = 8192;       // query slots of one resident launch
constexpr uint32_t kEngSegs = 1024;        // segments (admitted batches) of one launch
constexpr uint32_t kEngOps = 1u << 20;     // intersection operands of one launch
constexpr uint32_t kEngMaxN = 32;          // query vertices the resident launch takes
// A batch's slots start on a 64-slot boundary and its operands on a
// 16-operand one: its plans, counters, pending counts and operands then fill
// 256-byte lines of their own, which the launch has not read before the
// batch is published (the host writes them while the launch runs).
constexpr uint32_t kSlotAlign = 64;
constexpr uint32_t kOpAlign = 16;
static_assert(sizeof(DPlan) * kSlotAlign % 256 == 0 && sizeof(QCounters) * kSlotAlign % 256 == 0 &&
                  sizeof(uint32_t) * kSlotAlign % 256 == 0 && sizeof(DOp) * kOpAlign % 256 == 0,
              "a batch's plans, counters, pending counts and operands start on 256-byte lines");
static_assert(sizeof(DSeg) == 128, "one segment descriptor per 128-byte line");
constexpr double kEngIdleExitMs = 500.0;   // a drained launch the host has not closed leaves after this
constexpr double kEngIdleCloseMs = 2.0;    // the host closes a launch that has had nothing to run this long
constexpr double kResidentShare = 8.0;         // a batch is pre-expanded for 1/8 of the resident grid
// The wavefront-per-search kernel has a resident mode (k_dfs_wave with
// DQueue::segs) that the server does not use: a served test of it once left
// the GPU faulted and the cause has not been found, so resident launches
// always run the lane kernel.
constexpr bool kResidentWave = false;

// A batch prepared for the resident launch.
struct Prepared {
    std::vector<Request> reqs;
    double start_ms = 0;      // server clock when its preparation began
    double t_begin = 0;       // now_ms() then: the time limit counts from here
    uint64_t est = 0;         // admission estimate charged while it was prepared
    uint64_t held = 0;        // device bytes it holds until its launch ends
    int rc = PSM_OK;          // preparation failed: every request gets rc
    std::vector<DPlan> hp;    // bound to its tables, operand indices local
    std::vector<DOp> hops;
    std::vector<QCounters> qc0;   // counters of the pre-expansion (calls, bytes)
    std::vector<uint64_t> ntask;
    DBuf tasks, cand_pool, row_off, idx;
    uint64_t task_off = 0;          // tasks before this position ran in an earlier launch
    uint64_t n_tasks = 0;           // tasks from task_off on
    uint32_t q_base = 0, seg = 0;   // placement in the running launch
    std::vector<uint8_t> delivered;
};

// The resident launch and its host side.
struct Engine {
    psm_ctx* ctx = nullptr;         // launch stream and launch shape
    hipStream_t us = nullptr;       // fills and uploads beside the running launch
    uint32_t grid_waves = 0, ring = 0;
    bool wave = false;              // wavefront-per-search kernel (else lane-per-search)
    DBuf d_plans, d_ops, d_qc, d_pend, d_segs, d_heads, d_queue, d_seq, d_q;
    unsigned int* hword = nullptr;  // coherent pinned: segments published | closed << 31
    unsigned int* d_hword = nullptr;
    unsigned long long* notify = nullptr;   // coherent pinned, 4 words per slot
    unsigned long long* d_notify = nullptr;
    hipEvent_t ev = nullptr;        // end of the running launch
    bool running = false, closed = false;
    uint32_t next_slot = 0, next_op = 0, n_segs = 0;
    std::vector<std::unique_ptr<Prepared>> batches;   // published into the running launch
    std::vector<std::pair<Prepared*, uint32_t>> owner;   // slot -> (batch, query)
    std::vector<uint32_t> open;     // slots whose result is not delivered
    double idle_since = -1;
    uint64_t held = 0;              // device bytes the running launch's batches hold
    uint64_t launches = 0;
    double t_start = 0;             // host clock at the running launch's start
    const char* close_reason = "";
    ~Engine() {
        // a launch still running reads the pinned words and the arrays below:
        // close it and let it drain first
        if (running && hword) {
            __atomic_store_n(hword, n_segs | 0x80000000u, __ATOMIC_RELEASE);
            if (ctx) (void)hipStreamSynchronize(ctx->stream);
        }
        d_plans.release(); d_ops.release(); d_qc.release(); d_pend.release(); d_segs.release();
        d_heads.release(); d_queue.release(); d_seq.release(); d_q.release();
        if (hword) (void)hipHostFree(hword);
        if (notify) (void)hipHostFree(notify);
        if (ev) (void)hipEventDestroy(ev);
        if (us) (void)hipStreamDestroy(us);
        if (ctx) psm_ctx_destroy(ctx);
    }
};

}  // namespace

struct psm_server {
    const psm_graph* g = nullptr;
    int device = 0;
    int filter = PSM_FILTER_GQL, order = PSM_ORDER_GQL;
    uint64_t cap = 0;
    double time_limit_s = 60.0;
    uint32_t max_batch = 64;
    uint64_t mem_budget = 0;
    std::vector<psm_ctx*> ctxs;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_eng;
    std::deque<Request> pending;
    std::deque<psm_result> done;
    uint64_t mem_in_use = 0, peak_mem = 0, submitted = 0, completed = 0, batches = 0;
    bool stop = false;
    double t0 = 0;
    // resident launch
    std::unique_ptr<Engine> eng;          // null: every batch runs its own launch
    std::thread eng_thread;
    std::deque<std::unique_ptr<Prepared>> ready;   // prepared, not yet published
    uint32_t preparing = 0;               // batches workers are preparing
};

namespace {

// A request's device footprint up to its tables, bounded before it runs:
// bitmaps + ranks of its query vertices, its GQL membership words, and its
// candidate sets (label buckets, with their compaction scratch). Its tables
// are charged at their real size once built: bounded by the buckets and
// the degrees they are 10-25x larger than what C4's batches build.
uint64_t footprint(const psm_graph* g, const QGraph& q) {
    const uint64_t wps = ((uint64_t)g->V + 31) / 32;
    uint64_t b = (uint64_t)q.n * wps * 8 + (uint64_t)g->V * 8;
    for (uint32_t u = 0; u < q.n; ++u) {
        const uint32_t l = q.lab[u];
        const uint64_t bucket = l < g->L ? g->h_ri_off[l + 1] - g->h_ri_off[l] : 0;
        b += bucket * 4 * 2;
    }
    return b + (64ull << 20);
}

constexpr size_t kSmallBatch = 4;   // batches below this size use the default launch shape

// Returns request i of a prepared batch (caller holds S->mu).
void deliver_prepared(psm_server* S, Prepared& b, uint32_t i, uint64_t count, uint64_t calls, int status) {
    if (b.delivered[i]) return;
    b.delivered[i] = 1;
    psm_result res{};
    res.request_id = b.reqs[i].id;
    res.count = S->cap ? std::min<uint64_t>(S->cap, count) : count;
    res.calls = calls;
    res.status = status;
    res.batch_size = (uint32_t)b.reqs.size();
    res.submit_ms = b.reqs[i].submit_ms;
    res.start_ms = b.start_ms;
    res.done_ms = now_ms() - S->t0;
    S->done.push_back(res);
    ++S->completed;
}

// Worker side of the resident launch: the batch's plans (bound to its
// tables) and its pre-expanded tasks, with the device memory the launch
// will read moved into p.
int prepare_resident(psm_server* S, Batch& b, Prepared& p) {
    psm_ctx* ctx = b.ctx;
    tls_stream = b.s();
    const uint32_t nq = (uint32_t)b.qs.size();
    p.hp.assign(nq, DPlan{});
    p.hops.clear();
    for (uint32_t i = 0; i < nq; ++i) {
        make_plan(*b.qs[i], b.qs[i]->plan.order, nullptr, S->cap, p.hp[i], p.hops);
        b.bind_tables(p.hp[i]);
    }
    if (p.hops.empty()) p.hops.push_back(DOp{});
    HIP_TRY(ctx, upload(b.plans, p.hp, b.s()));
    HIP_TRY(ctx, upload(b.ops, p.hops, b.s()));
    HIP_TRY(ctx, b.qcnt.alloc(sizeof(QCounters) * nq));
    HIP_TRY(ctx, hipMemsetAsync(b.qcnt.p, 0, sizeof(QCounters) * nq, b.s()));
    // a batch shares the resident grid with the batches around it: the
    // lane kernel's batches are expanded for a share of its lanes, the
    // wavefront kernel's for its wavefronts
    const bool wave = S->eng->wave;
    const double workers = (double)S->eng->grid_waves * (wave ? 1 : 64);
    const uint64_t target_total = std::max<uint64_t>(
        64, (uint64_t)(workers * ctx->expand_tasks_per_lane / (wave ? 1.0 : kResidentShare)));
    PSM_TRY(b.expand_tasks(p.hp, target_total, p.tasks, p.n_tasks, p.ntask));
    HIP_TRY(ctx, download(p.qc0, b.qcnt.p, nq, b.s()));
    p.cand_pool.swap(b.cand_pool);
    p.row_off.swap(b.row_off);
    p.idx.swap(b.idx);
    // released by the engine once its launch has ended: nothing else is
    // left to wait for then
    for (DBuf* d : {&p.tasks, &p.cand_pool, &p.row_off, &p.idx}) {
        d->st = S->eng->us;
        p.held += d->bytes;
    }
    return PSM_OK;
}

void serve_worker(psm_server* S, psm_ctx* ctx) {
    (void)hipSetDevice(S->device);
    // coherent pinned words the DFS kernel posts finished queries to (4 per
    // request of a batch), so each result leaves as soon as it is final
    unsigned long long* words = nullptr;
    unsigned long long* dwords = nullptr;
    if (hipHostMalloc((void**)&words, (size_t)S->max_batch * 32, hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer((void**)&dwords, words, 0) != hipSuccess) {
        if (words) (void)hipHostFree(words);
        words = dwords = nullptr;   // results then leave when their batch ends
    }
    while (true) {
        std::vector<Request> batch;
        uint64_t est = 0;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [&] { return S->stop || !S->pending.empty(); });
            if (S->pending.empty()) {   // stopping
                if (words) (void)hipHostFree(words);
                return;
            }
            // admission: oldest first, while slots and memory allow (a request
            // that fits nowhere else runs alone once the device is idle)
            while (!S->pending.empty() && batch.size() < S->max_batch) {
                Request& r = S->pending.front();
                const bool fits = S->mem_in_use + est + r.est_bytes <= S->mem_budget;
                if (!fits && !(batch.empty() && S->mem_in_use == 0)) break;
                est += r.est_bytes;
                batch.push_back(std::move(r));
                S->pending.pop_front();
            }
            if (batch.empty()) {
                S->cv_done.wait(lk);   // memory frees up when another batch completes
                continue;
            }
            S->mem_in_use += est;
            S->peak_mem = std::max(S->peak_mem, S->mem_in_use);
            ++S->batches;
            ++S->preparing;
        }
        const double start = now_ms() - S->t0;
        Batch b;
        b.ctx = ctx;
        b.g = S->g;
        uint32_t max_n = 0;
        for (Request& r : batch) {
            b.qs.push_back(r.q.get());
            max_n = std::max(max_n, r.q->qg.n);
        }
        if (S->eng && max_n <= kEngMaxN) {
            // resident launch: prepare, then hand the batch to the engine
            std::unique_ptr<Prepared> p(new Prepared());
            p->start_ms = start;
            p->t_begin = now_ms();
            p->est = est;
            int rc = b.filter(S->filter);
            if (rc == PSM_OK) rc = b.build_tables();
            if (rc == PSM_OK) rc = b.make_plans(S->order);
            if (rc == PSM_OK) rc = prepare_resident(S, b, *p);
            p->rc = rc;
            p->reqs = std::move(batch);
            p->delivered.assign(p->reqs.size(), 0);
            {
                std::lock_guard<std::mutex> lk(S->mu);
                // the estimate gives way to what the launch will hold
                S->mem_in_use = S->mem_in_use - est + p->held;
                S->peak_mem = std::max(S->peak_mem, S->mem_in_use);
                --S->preparing;
                S->ready.push_back(std::move(p));
            }
            S->cv_eng.notify_one();
            S->cv_done.notify_all();
            continue;
        }
        // A batch of a few requests runs with the library's default launch
        // shape (the plain path's): the served shape is tuned for full
        // batches, and its chip-wide pre-expansion of one query would cost
        // more than the query itself.
        const uint32_t tuned_waves = ctx->dfs_waves_per_cu, tuned_strategy = ctx->dfs_strategy;
        const double tuned_expand = ctx->expand_tasks_per_lane;
        if (batch.size() < kSmallBatch) {
            ctx->dfs_waves_per_cu = kDefaultWavesPerCu;
            ctx->expand_tasks_per_lane = kDefaultTasksPerLane;
            ctx->dfs_strategy = PSM_DFS_LANE;
        }
        std::vector<uint8_t> delivered(batch.size(), 0);
        auto deliver = [&](size_t i, uint64_t count, uint64_t calls, int status) {
            psm_result res{};
            res.request_id = batch[i].id;
            res.count = count;
            res.calls = calls;
            res.status = status;
            res.batch_size = (uint32_t)batch.size();
            res.submit_ms = batch[i].submit_ms;
            res.start_ms = start;
            res.done_ms = now_ms() - S->t0;
            S->done.push_back(res);
            ++S->completed;
            delivered[i] = 1;
        };
        if (words) {
            memset(words, 0, batch.size() * 32);
            b.notify = dwords;
            b.poll = [&] {
                bool any = false;
                for (size_t i = 0; i < batch.size(); ++i) {
                    if (delivered[i] || !__atomic_load_n(words + 4 * i, __ATOMIC_ACQUIRE)) continue;
                    const uint64_t c = __atomic_load_n(words + 4 * i + 1, __ATOMIC_RELAXED);
                    const uint64_t k = __atomic_load_n(words + 4 * i + 2, __ATOMIC_RELAXED);
                    if (!any) S->mu.lock();
                    any = true;
                    deliver(i, S->cap ? std::min<uint64_t>(S->cap, c) : c, k, PSM_OK);
                }
                if (any) {
                    S->mu.unlock();
                    S->cv_done.notify_all();
                }
            };
        }
        int rc = b.filter(S->filter);
        if (rc == PSM_OK) rc = b.build_tables();
        const uint64_t tables = b.row_off.bytes + b.idx.bytes;   // charged at their real size
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->mem_in_use += tables;
            S->peak_mem = std::max(S->peak_mem, S->mem_in_use);
        }
        if (rc == PSM_OK) rc = b.make_plans(S->order);
        if (rc == PSM_OK) rc = b.evaluate(S->cap);
        b.poll = nullptr;
        ctx->dfs_waves_per_cu = tuned_waves;
        ctx->expand_tasks_per_lane = tuned_expand;
        ctx->dfs_strategy = tuned_strategy;
        std::lock_guard<std::mutex> lk(S->mu);
        for (size_t i = 0; i < batch.size(); ++i) {
            if (delivered[i]) continue;
            const QueryHost* q = batch[i].q.get();
            deliver(i, q->count, q->calls, rc != PSM_OK ? rc : (q->timed_out ? PSM_ETIMEOUT : q->status));
        }
        // the batch's device buffers (candidates, tables, queue) are freed now
        S->mem_in_use -= est + tables;
        --S->preparing;
        S->cv_done.notify_all();
        S->cv_eng.notify_one();
    }
}

// ---- resident launch (engine thread) ----------------------------------------

int engine_setup(psm_server* S, Engine& E, const psm_ctx* proto) {
    PSM_TRY(psm_ctx_create(S->device, &E.ctx));
    psm_ctx* c = E.ctx;
    c->claim_div = proto->claim_div;
    // the lane kernel's own intersection (galloping cursors): a tuned
    // PSM_INTERSECT_BSEARCH is chosen for the wavefront kernel, and served C4
    // through the resident lane kernel at two thirds of the rate
    c->intersect = PSM_INTERSECT_ADAPTIVE;
    c->queue_slots = proto->queue_slots;
    c->block_lanes = 0;
    c->verbose = proto->verbose;
    HIP_TRY(c, hipSetDevice(S->device));
    HIP_TRY(c, hipStreamCreateWithFlags(&E.us, hipStreamNonBlocking));
    HIP_TRY(c, hipEventCreateWithFlags(&E.ev, hipEventDisableTiming));
    HIP_TRY(c, hipHostMalloc((void**)&E.hword, 256, hipHostMallocCoherent));
    HIP_TRY(c, hipHostGetDevicePointer((void**)&E.d_hword, E.hword, 0));
    HIP_TRY(c, hipHostMalloc((void**)&E.notify, (size_t)kEngSlots * 32, hipHostMallocCoherent));
    HIP_TRY(c, hipHostGetDevicePointer((void**)&E.d_notify, E.notify, 0));
    // The lane kernel, one wavefront per CU. The resident launch shares
    // every CU with the kernels that prepare the next batches, and they are
    // most of a served request's device time: a further resident wavefront
    // per CU holds LDS (35 KB per wavefront of 32-vertex searches) and
    // registers they need for as long as the launch runs, and with two per
    // CU C4 was served at half the rate of one per CU. (The wavefront
    // kernel's resident mode is not used: see kResidentWave.)
    E.wave = kResidentWave && proto->dfs_strategy == PSM_DFS_WAVE;
    E.grid_waves = std::max<uint32_t>(4, (uint32_t)c->num_cu / 4 * 4);
    const uint64_t lanes_all = (uint64_t)E.grid_waves * 64;
    const uint64_t room = c->queue_slots ? c->queue_slots : std::max<uint64_t>(1u << 18, 4 * lanes_all);
    uint64_t ring = 64;
    while (ring < room) ring <<= 1;
    E.ring = (uint32_t)ring;
    hipStream_t prev = tls_stream;
    tls_stream = E.us;
    hipError_t e = hipSuccess;
    for (auto pr : {std::make_pair(&E.d_plans, sizeof(DPlan) * kEngSlots),
                    std::make_pair(&E.d_ops, sizeof(DOp) * kEngOps),
                    std::make_pair(&E.d_qc, sizeof(QCounters) * kEngSlots),
                    std::make_pair(&E.d_pend, sizeof(uint32_t) * kEngSlots),
                    std::make_pair(&E.d_segs, sizeof(DSeg) * kEngSegs),
                    std::make_pair(&E.d_heads, sizeof(uint32_t) * kEngSegs),
                    std::make_pair(&E.d_queue, sizeof(DTask) * ring),
                    std::make_pair(&E.d_seq, sizeof(uint32_t) * ring),
                    std::make_pair(&E.d_q, sizeof(DQueue))})
        if (e == hipSuccess) e = pr.first->alloc(pr.second);
    tls_stream = prev;
    HIP_TRY(c, e);
    E.owner.assign(kEngSlots, {nullptr, 0});
    return PSM_OK;
}

void engine_close(Engine& E, const char* reason) {
    if (!E.running || E.closed) return;
    __atomic_store_n(E.hword, E.n_segs | 0x80000000u, __ATOMIC_RELEASE);
    E.closed = true;
    E.close_reason = reason;
}

// A new launch: counters, pending counts and segment heads cleared, the
// ring reset, nothing published; then the persistent kernel.
int engine_start(Engine& E) {
    psm_ctx* c = E.ctx;
    HIP_TRY(c, launch_fill_u32(E.d_qc.as<uint32_t>(), 0u, sizeof(QCounters) * kEngSlots / 4, E.us));
    HIP_TRY(c, launch_fill_u32(E.d_pend.as<uint32_t>(), 0u, kEngSlots, E.us));
    HIP_TRY(c, launch_fill_u32(E.d_heads.as<uint32_t>(), 0u, kEngSegs, E.us));
    HIP_TRY(c, launch_ring_init(E.d_seq.as<uint32_t>(), E.ring, 0u, E.us));
    DQueue hq{};
    hq.mask = E.ring - 1;
    hq.claim_div = c->claim_div;
    hq.intersect = c->intersect;
    hq.notify = E.d_notify;
    hq.segs = E.d_segs.as<DSeg>();
    hq.seg_head = E.d_heads.as<unsigned int>();
    hq.host_word = E.d_hword;
    hq.idle_ticks = (unsigned long long)(kEngIdleExitMs * c->wall_khz);
    HIP_TRY(c, hipMemcpyAsync(E.d_q.p, &hq, sizeof(hq), hipMemcpyHostToDevice, E.us));
    __atomic_store_n(E.hword, 0u, __ATOMIC_SEQ_CST);
    memset(E.notify, 0, (size_t)kEngSlots * 32);
    HIP_TRY(c, hipStreamSynchronize(E.us));
    // Whatever the launch call returns, the kernel may have started: it is
    // tracked (its end event recorded) from here on, and a failed start is
    // closed at once, so nothing it reads is released before it has left.
    (void)hipGetLastError();   // the launch wrappers report hipGetLastError()
    hipError_t le;
    if (E.wave)
        le = launch_dfs_wave(E.d_plans.as<DPlan>(), E.d_ops.as<DOp>(), nullptr, E.d_queue.as<DTask>(),
                             E.d_seq.as<uint32_t>(), E.d_q.as<DQueue>(), E.d_qc.as<QCounters>(),
                             E.d_pend.as<uint32_t>(), kEngMaxN, E.grid_waves, c->stream);
    else
        le = launch_dfs(E.d_plans.as<DPlan>(), E.d_ops.as<DOp>(), nullptr, E.d_queue.as<DTask>(),
                        E.d_seq.as<uint32_t>(), E.d_q.as<DQueue>(), E.d_qc.as<QCounters>(),
                        E.d_pend.as<uint32_t>(), kEngMaxN, E.grid_waves, 0, EmitSink{nullptr, nullptr, 0, nullptr},
                        c->stream);
    E.running = true;
    E.closed = false;
    E.next_slot = E.next_op = E.n_segs = 0;
    E.idle_since = -1;
    ++E.launches;
    const hipError_t re = hipEventRecord(E.ev, c->stream);
    E.t_start = now_ms();
    if (le != hipSuccess || re != hipSuccess) {
        engine_close(E, "start failed");
        if (re != hipSuccess) {   // no end event to wait for: wait for the stream itself
            (void)hipStreamSynchronize(c->stream);
            E.running = false;
        }
        return hip_fail(c, le != hipSuccess ? le : re, "resident launch start", __LINE__);
    }
    return PSM_OK;
}

bool engine_fits(const Engine& E, const Prepared& b) {
    const uint64_t q_base = (E.next_slot + kSlotAlign - 1) / kSlotAlign * kSlotAlign;
    const uint64_t op_base = (E.next_op + kOpAlign - 1) / kOpAlign * kOpAlign;
    return E.n_segs < kEngSegs && q_base + b.reqs.size() <= kEngSlots && op_base + b.hops.size() <= kEngOps;
}

// Writes a prepared batch's plans, operands, counters, pending counts and
// segment descriptor into the running launch (memory it has not read), then
// publishes the segment; the launch then owns the batch (bp is left empty).
int engine_publish(psm_server* S, Engine& E, std::unique_ptr<Prepared>& bp) {
    psm_ctx* c = E.ctx;
    Prepared& b = *bp;
    const uint32_t nq = (uint32_t)b.reqs.size();
    const uint32_t q_base = (E.next_slot + kSlotAlign - 1) / kSlotAlign * kSlotAlign;
    const uint32_t op_base = (E.next_op + kOpAlign - 1) / kOpAlign * kOpAlign;
    std::vector<DPlan> hp = b.hp;
    uint64_t limit_ticks = 0;
    if (S->time_limit_s > 0) {
        const double left_s = S->time_limit_s - (now_ms() - b.t_begin) / 1000.0;
        limit_ticks = (uint64_t)std::max(1.0, left_s * c->wall_khz * 1000.0);
    }
    for (DPlan& P : hp) {
        for (uint32_t d = 0; d < P.n; ++d) P.op_begin[d] += op_base;
        P.limit_ticks = limit_ticks;
    }
    std::vector<uint32_t> pend(nq);
    for (uint32_t i = 0; i < nq; ++i) pend[i] = (uint32_t)b.ntask[i];
    DSeg seg{};
    seg.tasks = b.tasks.as<DTask>() + b.task_off;
    seg.n = (uint32_t)b.n_tasks;
    seg.q_base = q_base;
    HIP_TRY(c, hipMemcpyAsync(E.d_plans.as<DPlan>() + q_base, hp.data(), sizeof(DPlan) * nq,
                              hipMemcpyHostToDevice, E.us));
    HIP_TRY(c, hipMemcpyAsync(E.d_ops.as<DOp>() + op_base, b.hops.data(), sizeof(DOp) * b.hops.size(),
                              hipMemcpyHostToDevice, E.us));
    HIP_TRY(c, hipMemcpyAsync(E.d_qc.as<QCounters>() + q_base, b.qc0.data(), sizeof(QCounters) * nq,
                              hipMemcpyHostToDevice, E.us));
    HIP_TRY(c, hipMemcpyAsync(E.d_pend.as<uint32_t>() + q_base, pend.data(), sizeof(uint32_t) * nq,
                              hipMemcpyHostToDevice, E.us));
    HIP_TRY(c, hipMemcpyAsync(E.d_segs.as<DSeg>() + E.n_segs, &seg, sizeof(seg), hipMemcpyHostToDevice, E.us));
    HIP_TRY(c, hipStreamSynchronize(E.us));
    E.next_slot = q_base + nq;
    E.next_op = op_base + (uint32_t)b.hops.size();
    b.q_base = q_base;
    b.seg = E.n_segs;
    {
        // a query without tasks left (or without candidates) is final now
        std::lock_guard<std::mutex> lk(S->mu);
        for (uint32_t i = 0; i < nq; ++i) {
            E.owner[q_base + i] = {&b, i};
            const QueryHost* q = b.reqs[i].q.get();
            if (b.delivered[i]) continue;
            if (b.ntask[i] == 0 || q->status != PSM_OK)
                deliver_prepared(S, b, i, b.qc0[i].count, b.qc0[i].calls,
                                 q->status != PSM_OK ? q->status : (b.qc0[i].timed_out ? PSM_ETIMEOUT : PSM_OK));
            else
                E.open.push_back(q_base + i);
        }
    }
    S->cv_done.notify_all();
    ++E.n_segs;
    __atomic_store_n(E.hword, E.n_segs, __ATOMIC_RELEASE);
    E.held += b.held;
    E.batches.push_back(std::move(bp));
    return PSM_OK;
}

// Results the launch has posted.
void engine_poll(psm_server* S, Engine& E) {
    bool any = false;
    for (size_t k = 0; k < E.open.size();) {
        const uint32_t slot = E.open[k];
        const unsigned long long* w = E.notify + 4ull * slot;
        const unsigned long long flag = __atomic_load_n(w, __ATOMIC_ACQUIRE);
        if (!flag) {
            ++k;
            continue;
        }
        const uint64_t count = __atomic_load_n(w + 1, __ATOMIC_RELAXED);
        const uint64_t calls = __atomic_load_n(w + 2, __ATOMIC_RELAXED);
        if (!any) S->mu.lock();
        any = true;
        deliver_prepared(S, *E.owner[slot].first, E.owner[slot].second, count, calls,
                         flag == 2 ? PSM_ETIMEOUT : PSM_OK);
        E.open[k] = E.open.back();
        E.open.pop_back();
    }
    if (any) {
        S->mu.unlock();
        S->cv_done.notify_all();
    }
}

// The launch has ended: results it did not post come from its counters. A
// batch whose tasks it did not all claim (published after its wavefronts
// began to leave) goes back to the queue with what is left, its queries'
// counters and pending counts carried over, for the next launch. The memory
// of the other batches is recycled.
int engine_finish(psm_server* S, Engine& E, std::deque<std::unique_ptr<Prepared>>& again, int failed) {
    psm_ctx* c = E.ctx;
    engine_poll(S, E);   // results posted before the end (their words are cleared at the next start)
    std::vector<uint32_t> heads, pend;
    std::vector<QCounters> qc;
    int rc = failed;
    if (rc == PSM_OK) {
        hipError_t e = download(heads, E.d_heads.p, E.n_segs, E.us);
        if (e == hipSuccess) e = download(pend, E.d_pend.p, E.next_slot, E.us);
        if (e == hipSuccess) e = download(qc, E.d_qc.p, E.next_slot, E.us);
        if (e != hipSuccess) rc = hip_fail(c, e, "resident launch read-back", __LINE__);
    }
    uint64_t freed = 0;
    if (c->verbose)
        fprintf(stderr, "[psm] resident launch %llu: %.1f ms, %u segments, %u slots, %.1f MB held, closed: %s\n",
                (unsigned long long)E.launches, now_ms() - E.t_start, E.n_segs, E.next_slot, E.held / 1e6,
                E.closed ? E.close_reason : "no (left idle)");
    {
        std::lock_guard<std::mutex> lk(S->mu);
        for (auto& bp : E.batches) {
            Prepared& b = *bp;
            if (rc == PSM_OK && b.n_tasks && heads[b.seg] < b.n_tasks) {
                const uint32_t nq = (uint32_t)b.reqs.size();
                b.task_off += heads[b.seg];
                b.n_tasks -= heads[b.seg];
                for (uint32_t i = 0; i < nq; ++i) {
                    b.qc0[i] = qc[b.q_base + i];
                    b.ntask[i] = pend[b.q_base + i];
                }
                again.push_back(std::move(bp));
                continue;
            }
            for (uint32_t i = 0; i < (uint32_t)b.reqs.size(); ++i) {
                if (b.delivered[i]) continue;
                if (rc != PSM_OK) {
                    deliver_prepared(S, b, i, 0, 0, rc);
                    continue;
                }
                const uint32_t slot = b.q_base + i;
                const QCounters& k = qc[slot];
                const bool partial = k.timed_out || (pend[slot] && !k.done);
                deliver_prepared(S, b, i, k.count, k.calls, partial ? PSM_ETIMEOUT : PSM_OK);
            }
            freed += b.held;
        }
        S->mem_in_use -= freed;
    }
    // the blocks go back to the cache only now: the launch that read them has ended
    E.batches.clear();
    E.open.clear();
    E.held = 0;
    E.running = false;
    S->cv_done.notify_all();
    return rc;
}

void engine_loop(psm_server* S) {
    Engine& E = *S->eng;
    (void)hipSetDevice(S->device);
    tls_stream = E.us;
    std::deque<std::unique_ptr<Prepared>> again;   // batches for the next launch
    while (true) {
        std::deque<std::unique_ptr<Prepared>> got;
        bool stopping = false;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            if (!E.running && again.empty())
                S->cv_eng.wait(lk, [&] {
                    return !S->ready.empty() || (S->stop && S->pending.empty() && S->preparing == 0);
                });
            got.swap(S->ready);
            stopping = S->stop && S->pending.empty() && S->preparing == 0;
        }
        while (!got.empty()) {
            again.push_back(std::move(got.front()));
            got.pop_front();
        }
        if (again.empty() && !E.running && stopping) break;
        // publish what fits into the running launch (a new one if none runs)
        while (!again.empty()) {
            std::unique_ptr<Prepared>& b = again.front();
            if (b->rc != PSM_OK) {
                {
                    std::lock_guard<std::mutex> lk(S->mu);
                    for (uint32_t i = 0; i < (uint32_t)b->reqs.size(); ++i) deliver_prepared(S, *b, i, 0, 0, b->rc);
                    S->mem_in_use -= b->held;
                }
                S->cv_done.notify_all();
                again.pop_front();
                continue;
            }
            if (E.running && (E.closed || !engine_fits(E, *b))) {
                engine_close(E, "slots");   // the next launch takes it
                break;
            }
            int rc = E.running ? PSM_OK : engine_start(E);
            if (rc == PSM_OK) rc = engine_publish(S, E, b);
            if (rc != PSM_OK) {
                if (b) {
                    std::lock_guard<std::mutex> lk(S->mu);
                    for (uint32_t i = 0; i < (uint32_t)b->reqs.size(); ++i) deliver_prepared(S, *b, i, 0, 0, rc);
                    S->mem_in_use -= b->held;
                }
                S->cv_done.notify_all();
            }
            again.pop_front();
        }
        if (!E.running) continue;
        engine_poll(S, E);
        // close the launch when it has had nothing to run for 2 ms, even
        // while workers prepare further batches (a long preparation would
        // otherwise run beside an idle launch that holds LDS and registers
        // its kernels need; the next batch starts a new launch, and the
        // memory of this one is recycled), at once when stopping, or when
        // the memory its batches hold reaches 60 % of the budget
        const double now = now_ms();
        if (!E.closed) {
            const bool idle = E.open.empty() && again.empty();
            if (!idle) {
                E.idle_since = -1;
            } else if (stopping) {
                engine_close(E, "stopping");
            } else if (E.idle_since < 0) {
                E.idle_since = now;
            } else if (now - E.idle_since > kEngIdleCloseMs) {
                engine_close(E, "idle");
            }
            if ((double)E.held > 0.6 * (double)S->mem_budget) engine_close(E, "memory");
        }
        const hipError_t qe = hipEventQuery(E.ev);
        if (qe == hipSuccess) {
            engine_finish(S, E, again, PSM_OK);
        } else if (qe != hipErrorNotReady) {
            // the launch is over only once its kernel has left: close it and
            // wait, so nothing it reads is released under it
            const int rc = hip_fail(E.ctx, qe, "resident launch", __LINE__);
            engine_close(E, "launch error");
            (void)hipStreamSynchronize(E.ctx->stream);
            engine_finish(S, E, again, rc);
        } else {
            std::this_thread::sleep_for(std::chrono::microseconds(50));
        }
    }
}

}  // namespace

extern "C" {

int psm_server_create(psm_ctx* ctx, const psm_graph* g, uint32_t workers, uint32_t max_batch,
                      uint64_t mem_budget, int filter, int order, uint64_t cap, double time_limit_s,
                      psm_server** out) {
    if (!ctx || !g || !out || workers == 0 || workers > 16 || max_batch == 0) return PSM_EINVAL;
    if (filter < PSM_FILTER_LDF || filter > PSM_FILTER_DPISO || order < PSM_ORDER_GQL || order > PSM_ORDER_DPISO)
        return PSM_EINVAL;
    std::unique_ptr<psm_server> S(new psm_server());
    S->g = g;
    S->device = ctx->device;
    S->filter = filter;
    S->order = order;
    S->cap = cap;
    S->time_limit_s = time_limit_s;
    S->max_batch = max_batch;
    if (!mem_budget) {
        size_t free_b = 0, total_b = 0;
        (void)hipSetDevice(ctx->device);
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) mem_budget = free_b / 2;
        else mem_budget = 16ull << 30;
    }
    S->mem_budget = mem_budget;
    S->t0 = now_ms();
    if (ctx->serve_resident) {
        S->eng.reset(new Engine());
        const int r = engine_setup(S.get(), *S->eng, ctx);
        if (r != PSM_OK) {
            set_err(ctx, r, "resident launch: " + (S->eng->ctx ? S->eng->ctx->err : std::string("no context")));
            return r;
        }
    }
    for (uint32_t i = 0; i < workers; ++i) {
        psm_ctx* c = nullptr;
        const int r = psm_ctx_create(ctx->device, &c);
        if (r != PSM_OK) {
            for (psm_ctx* x : S->ctxs) psm_ctx_destroy(x);
            return r;
        }
        c->time_limit_s = time_limit_s;
        c->verbose = ctx->verbose;
        // workers run with the creating context's launch shape, the
        // persistent DFS grid split between them: batches of different
        // workers overlap on the device, and grids sized for the whole chip
        // would queue behind one another instead of sharing the CUs
        c->dfs_waves_per_cu = std::max<uint32_t>(1, ctx->dfs_waves_per_cu / workers);
        c->expand_tasks_per_lane = ctx->expand_tasks_per_lane;
        c->dfs_strategy = ctx->dfs_strategy;
        c->expand_depth = ctx->expand_depth;
        c->intersect = ctx->intersect;
        S->ctxs.push_back(c);
    }
    for (psm_ctx* c : S->ctxs) S->workers.emplace_back(serve_worker, S.get(), c);
    if (S->eng) S->eng_thread = std::thread(engine_loop, S.get());
    *out = S.release();
    return PSM_OK;
}

int psm_server_submit(psm_server* S, uint64_t request_id, uint32_t n, const uint32_t* q_off,
                      const uint32_t* q_adj, const uint32_t* q_lab) {
    if (!S || !q_off || !q_lab) return PSM_EINVAL;
    Request r;
    r.id = request_id;
    r.q.reset(new QueryHost());
    const int rc = make_qgraph(n, q_off, q_adj, q_lab, r.q->qg);
    if (rc != PSM_OK) return rc;
    if (S->filter == PSM_FILTER_GQL && r.q->qg.max_degree() > kMaxGqlDegree) return PSM_EINVAL;
    r.est_bytes = footprint(S->g, r.q->qg);
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop) return PSM_ESTATE;
        r.submit_ms = now_ms() - S->t0;
        S->pending.push_back(std::move(r));
        ++S->submitted;
    }
    S->cv_work.notify_one();
    return PSM_OK;
}

int psm_server_poll(psm_server* S, psm_result* out, uint32_t max_results, uint32_t* n_out, double wait_ms) {
    if (!S || !n_out || (!out && max_results)) return PSM_EINVAL;
    std::unique_lock<std::mutex> lk(S->mu);
    if (S->done.empty() && wait_ms > 0)
        S->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(wait_ms),
                            [&] { return !S->done.empty(); });
    uint32_t k = 0;
    while (k < max_results && !S->done.empty()) {
        out[k++] = S->done.front();
        S->done.pop_front();
    }
    *n_out = k;
    return PSM_OK;
}

int psm_server_stats(psm_server* S, uint64_t* out5) {
    if (!S || !out5) return PSM_EINVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    out5[0] = S->submitted;
    out5[1] = S->completed;
    out5[2] = S->batches;
    out5[3] = S->peak_mem;
    out5[4] = S->pending.size();
    return PSM_OK;
}

int psm_server_destroy(psm_server* S) {
    if (!S) return PSM_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    S->cv_done.notify_all();
    S->cv_eng.notify_all();
    for (auto& t : S->workers) t.join();
    // the workers have handed over every batch: the engine publishes them,
    // closes its launch once everything has run, and returns
    S->cv_eng.notify_all();
    if (S->eng_thread.joinable()) S->eng_thread.join();
    for (psm_ctx* c : S->ctxs) psm_ctx_destroy(c);
    delete S;
    return PSM_OK;
}

}  // extern "C"
