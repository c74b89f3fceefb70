// internal.h — structures shared by the device kernels (device.hip) and the
// host orchestration (psm_abi.cpp). Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "psm_abi.h"

namespace psm {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;   // pruned candidate
constexpr uint32_t kMaxQ = 64;               // query vertices (config.h:9)
constexpr uint32_t kMaxGqlDegree = 32;       // query degree limit of the GQL refinement kernel
constexpr uint32_t kPrefixMax = 32;          // prefix entries a task record carries
constexpr uint32_t kExpandMax = kPrefixMax - 1;   // depth limit of task pre-expansion (a task at depth d carries d prefix entries)
constexpr uint64_t kAutoHi = ~0ull;          // task range = the whole base list
constexpr uint32_t kIdentity = 0xFFFFFFFFu;  // "base list is C(u) itself"

// Device-resident data graph (uint32 CSR + derived indexes).
struct DevGraph {
    uint32_t V;
    uint32_t L;            // labels count
    uint64_t nnz;          // 2E
    const uint32_t* off;   // V+1
    const uint32_t* adj;   // nnz
    const uint32_t* lab;   // V
    const uint32_t* ri_off;  // L+1: vertices grouped by label
    const uint32_t* ri;      // V, ascending within a label
    const uint32_t* nlf_off; // V+1: neighbour-label runs per vertex
    const uint2* nlf;        // (label, count), sorted by label
    const uint32_t* adj_bl;  // nnz: neighbours of each vertex ordered by (label, id)
    const uint32_t* nlf_beg; // per run: its first position in adj_bl
    // the degree and neighbour-label runs again, in reverse-index order
    // (position j describes vertex ri[j]): a scan of a label bucket reads
    // them as streams instead of one random line per vertex and field
    const uint32_t* deg_ri;     // V
    const uint32_t* nlf_ri_off; // V+1
    const uint2* nlf_ri;        // the runs of ri[0], ri[1], ...
};

// One query vertex as the candidate filters see it.
struct QVert {
    uint32_t label;
    uint32_t degree;
    uint32_t nlf_begin;    // into the batch's query NLF run array
    uint32_t nlf_count;
};

// A segment of elements to filter and compact in order. Blocks of 256
// elements are numbered consecutively over all items.
enum CompactMode : uint32_t {
    kModeLDF = 0,      // src = reverse index bucket, keep deg >= d(u)
    kModeNLF = 1,      // + NLF dominance
    kModeValid = 2,    // src = candidate pool, keep != kInvalid
    kModeCflGen = 3,   // src = reverse index bucket, keep cnt == target && NLF
};

struct CItem {
    uint64_t src_begin;    // first element in the source array
    uint32_t len;
    uint32_t block_start;  // first block of this item
    uint32_t mode;
    uint32_t qv;           // QVert index
    uint64_t cnt_base;     // kModeCflGen: per-query support counter array offset
    uint32_t target;
    uint32_t pad;
};

// Label-shared bucket scan (LDF / NLF over every (query, u) of a batch): the
// items of one label all scan that label's bucket, so a wavefront takes one
// 64-element chunk of the bucket and tests up to 64 items of the label
// against it (a subset; the chunk's degrees and neighbour-label runs are
// read once per subset instead of once per item). Wavefront w of a group
// takes chunk w / n_sub and subset w % n_sub (the subsets of one chunk are
// adjacent, so they share its lines in the caches). An item's keep mask and
// count for chunk r of its bucket sit at chunk_start + r.
constexpr uint32_t kBucketChunk = 64;   // bucket elements per wavefront (one per lane)
struct CGroup {
    uint64_t src_begin;    // first element of the bucket (reverse-index order)
    uint32_t len;          // bucket length
    uint32_t wave_start;   // first wavefront of this group
    uint32_t item_begin;   // into the CBucketItem array
    uint32_t n_items;
    uint32_t n_sub;        // item subsets of 64: ceil(n_items / 64)
    uint32_t pad;
};
// One (query, u) of a label-shared scan, as the scan needs it.
struct CBucketItem {
    uint32_t chunk_start;
    uint32_t degree;       // d(u)
    uint32_t nlf_begin;    // into the query NLF run array
    uint32_t nlf_count;    // runs; bit 31 set: LDF (no NLF test)
};

// GQL refinement work item: the candidates of one query vertex u.
// The candidate sets of one query are also held as membership words:
// member[mem_base + w] has bit u set when data vertex w is in C(u).
struct GItem {
    uint64_t cand_begin;
    uint32_t len;
    uint32_t wave_start;   // first wavefront of this item
    uint32_t bm_self;      // bitmap slot of (query, u)
    uint32_t k;            // query degree of u
    uint32_t left_begin;   // into the left-slot array: bitmap slots of N_q(u)
    uint32_t u;            // the query vertex (bm_self - u = the query's first slot)
    uint64_t mem_base;     // the query's membership words
    uint32_t lab_begin;    // into the label array: the distinct labels of N_q(u), ascending
    uint32_t nlab;
};

// BuildTable work item: ordered query edge (a, b) of one query.
struct TItem {
    uint64_t cand_a;       // candidate pool offset of C(a)
    uint32_t len_a;
    uint32_t bm_b;         // bitmap slot of (query, b)
    uint64_t row_base;     // first row of this table in the global row arrays
    uint32_t lab_b;        // label of b: only that run of N(v) can hold C(b) vertices
    uint32_t pad;
};

// Enumeration plan of one query (depth d holds query vertex order[d]).
struct DPlan {
    uint32_t n;
    uint32_t pad;
    uint64_t cap;
    uint64_t m_needed;            // depths whose data vertex a later depth compares
    // the query's candidate pool and candidate-edge tables (device memory of
    // the batch that prepared it): the kernels read every query's data
    // through its own plan, so one launch can enumerate queries of several
    // batches
    const uint64_t* row_off;
    const uint32_t* idx;
    const uint32_t* cand_pool;
    uint64_t cand_base[kMaxQ];
    uint32_t cand_len[kMaxQ];
    uint32_t op_begin[kMaxQ];
    uint32_t op_count[kMaxQ];
    uint32_t pad2[kMaxQ];
    uint64_t same_label[kMaxQ];   // earlier depths whose vertex has the same label
    // resident launch: wall-clock ticks the query may run from the start of
    // its first task (0: no limit)
    uint64_t limit_ticks;
};

// One intersection operand of a depth: the table row E[b -> u_d] selected by
// the candidate index assigned at depth src_depth.
struct DOp {
    uint32_t src_depth;
    uint32_t pad;
    uint64_t row_base;
};

// A unit of enumeration work: query q with the prefix I[0..depth-1] fixed;
// the lane enumerates positions [lo, hi) of the depth's base list (operand
// `sel`, or C(u) itself when sel == kIdentity). hi == kAutoHi: the whole
// list, located by the lane from the prefix.
struct DTask {
    uint32_t q;
    uint32_t depth;
    uint32_t sel;
    uint32_t pad;
    uint64_t lo;
    uint64_t hi;
    uint32_t prefix[kPrefixMax];
};

// Work queue of the persistent DFS, in two parts. The pre-expanded tasks are
// a fixed array init[0, n0), claimed in order through head0. Work that
// running lanes split off goes to a ring of `mask + 1` task records (a power
// of two) addressed by 32-bit positions; position p lives in slot p & mask,
// positions [head, tail) are queued, and idle lanes claim from the ring
// before the array: split-off work belongs to searches already under way,
// whose ends bound the launch. seq[slot] is the slot's sequence word: p + 1
// once the record of position p is written (claimable), p + mask + 1 once
// its claimant has read it (free for position p + mask + 1). A producer
// takes position tail only when its slot is free, so slots are reused for as
// long as the launch runs and splitting never stops at a fixed count.
// Positions wrap modulo 2^32; only differences of positions are compared.
struct DQueue {
    unsigned int head;
    unsigned int tail;
    unsigned int busy;         // lanes holding a task
    unsigned int timeout;      // the time limit passed
    unsigned int mask;         // ring slots - 1
    unsigned int donated;      // tasks split off by running lanes
    unsigned long long limit_ticks;   // wall-clock ticks allowed (0: no limit)
    unsigned long long deadline;      // set by the first lane to start
    unsigned long long lane_steps;    // loop iterations in which a lane held a task
    unsigned long long lane_iters;    // loop iterations of all lanes
    unsigned long long start_tick;    // wall clock when the first block started
    unsigned int full;         // donations skipped because the next slot was still held
    unsigned int head0;        // next task of the pre-expanded array
    unsigned int n0;           // pre-expanded tasks
    unsigned int claim_div;    // array tasks per lane above which a lane takes several per claim
    unsigned int intersect;    // PSM_INTERSECT_ADAPTIVE or PSM_INTERSECT_BSEARCH (psm_abi.h)
    // Host-visible (coherent pinned) words, 4 per query, or null: once a
    // query's count is final the kernel writes its count, calls and visited
    // rejects to words 1..3 and then 1 to word 0 (2: the query passed its
    // time limit and the count is partial), so a caller can return it while
    // the launch still runs.
    unsigned long long* notify;
    // Resident launch (segs != null, lane kernel only). The pre-expanded
    // array is a list of segments that the host publishes while the launch
    // runs: host_word (coherent pinned memory, written by the host) holds the
    // number of segments published, and bit 31 once the host has closed the
    // launch (no further segments). A segment's descriptor, plans, operands,
    // counters and pending counts are written before it is published, into
    // memory the launch has not read. The launch ends once it is closed and
    // drained, or after idle_ticks drained with nothing new published (then
    // as a whole: a segment published meanwhile is left untouched, for the
    // host to publish again in its next launch).
    const struct DSeg* segs;
    unsigned int* seg_head;        // next task of each segment (zero at the launch)
    const unsigned int* host_word;
    unsigned int pub;              // the latest host_word a wavefront read
    unsigned int seg_cur;          // segments before this one are known to be exhausted
    unsigned long long poll_tick;  // wall clock of the latest read of host_word
    unsigned long long idle_ticks; // 0: leave only once closed
    unsigned int exiting;          // set by the first wavefront to leave after idle_ticks: no
                                   // segment is claimed after it, every wavefront leaves
};

// One published segment of a resident launch: a prepared batch's
// pre-expanded tasks (query indices local to the batch) and the launch slot
// of the batch's first query. One descriptor per 128-byte line.
struct alignas(128) DSeg {
    const DTask* tasks;
    uint32_t n;
    uint32_t q_base;
};

// Where the DFS writes complete embeddings (rows of n data vertex ids in
// depth order) when materialising a join unit; rows == nullptr: count only.
struct EmitSink {
    uint32_t* rows;
    unsigned long long* count;
    unsigned long long cap;
    uint32_t* overflow;
};

// Column roles of one join side: key columns (shared query vertices, in the
// same key order on both sides) and the remaining columns.
struct JoinKeys {
    uint32_t nkeys;
    uint32_t nrest;
    uint32_t key[kMaxQ];
    uint32_t rest[kMaxQ];
};

// Per-query enumeration counters.
struct QCounters {
    unsigned long long count;
    unsigned long long calls;
    // extensions in every operand row whose data vertex an earlier depth
    // holds (the LFTJ engine's Call Count adds them)
    unsigned long long rejects;
    unsigned long long probe_bytes;    // bytes the DFS kernel touched (probes, base elements, rows located)
    unsigned long long expand_bytes;   // the same for task pre-expansion
    // Algorithmic enumeration bytes: for every search-tree node below the
    // root, 4 per entry of each of its operand rows, plus 4 per survivor it
    // writes unless its survivors are leaves; split by the kernel that
    // entered the node.
    unsigned long long algo_bytes;         // nodes of the DFS kernels
    unsigned long long expand_algo_bytes;  // nodes of task pre-expansion
    uint32_t done;        // cap reached
    uint32_t timed_out;   // work was dropped at the time limit
    // wall clock (wall_clock64 ticks) when the query's count reached its cap
    // (0: not reached), and the latest end of one of its tasks: the count is
    // final at the first, or else at the second
    unsigned long long cap_tick;
    unsigned long long end_tick;
    uint32_t notified;    // the final count has been posted to DQueue::notify
    uint32_t pad;
    // resident launch: the query's deadline in wall-clock ticks, set by the
    // first lane to start one of its tasks (DPlan::limit_ticks)
    unsigned long long deadline;
};

// Kernel classes timed by the library (psm_kernel_stats).
enum KernelKind : int {
    kKGqlRefine = 0,
    kKTableCount = 1,
    kKTableFill = 2,
    kKDfs = 3,
    kKExpand = 4,
    kKCompact = 5,
    kKBitmap = 6,
    kKJoin = 7,
    kKNumKinds = 8,
};

// ---- kernel launchers (device.hip) --------------------------------------
hipError_t launch_compact_flag(const DevGraph& g, const CItem* items, uint32_t n_items,
                               uint32_t n_blocks, const QVert* qv, const uint2* qnlf,
                               const uint32_t* src_pool, const uint32_t* cnt_pool,
                               unsigned long long* masks, uint32_t* block_counts, hipStream_t s);
hipError_t launch_compact_scatter(const CItem* items, uint32_t n_items, uint32_t n_blocks,
                                  const uint32_t* src_ri, const uint32_t* src_pool,
                                  const unsigned long long* masks, const uint64_t* block_prefix,
                                  uint32_t* out, hipStream_t s);
// Label-shared bucket scan (CGroup): masks / chunk_counts per item chunk,
// then the survivors' vertex ids written at chunk_prefix (exclusive scan of
// chunk_counts). n_waves: the wavefronts of all groups.
hipError_t launch_bucket_flag(const DevGraph& g, const CGroup* groups, uint32_t n_groups, uint32_t n_waves,
                              const CBucketItem* bitems, const uint2* qnlf, unsigned long long* masks,
                              uint32_t* chunk_counts, hipStream_t s);
hipError_t launch_bucket_scatter(const DevGraph& g, const CGroup* groups, uint32_t n_groups, uint32_t n_waves,
                                 const CBucketItem* bitems, const unsigned long long* masks,
                                 const uint64_t* chunk_prefix, uint32_t* out, hipStream_t s);
hipError_t launch_set_bits(const uint32_t* pool, const uint64_t* seg_begin, const uint32_t* seg_len,
                           const uint32_t* seg_slot, uint32_t n_seg, uint32_t total_max,
                           uint32_t* bitmaps, uint64_t words_per_slot, hipStream_t s);
hipError_t launch_set_member(const uint32_t* pool, const uint64_t* seg_begin, const uint32_t* seg_len,
                             const uint64_t* seg_mem, const uint32_t* seg_bit, uint32_t n_seg,
                             uint32_t total_max, unsigned long long* member, hipStream_t s);
hipError_t launch_popc_words(const uint32_t* bitmaps, uint32_t* out, uint64_t n_words, hipStream_t s);
// The 8-lane kernel tests candidates with k <= 8 and deg(v) <= 8 and defers
// the rest by class (16, 32, 64 lanes, wavefront kernel) to
// lists + c * list_stride (list_stride >= n_cands), counts added to
// n_lists[0..3]. The deferred-class kernels read their count from n_list;
// max_n bounds it (sizes the grid).
// One (round, vertex position) step of the GQL refinement: removed[slot] is
// set when a candidate of bitmap slot `slot` is dropped in this round; skip
// (second round only, else null) holds one flag per item of the step: set
// when none of the item's left vertices lost a candidate since the item's
// first-round test, so every surviving candidate passes again untested.
struct GqlRound {
    uint32_t* removed;
    const uint32_t* skip;
};
hipError_t launch_gql_refine_small(const DevGraph& g, const GItem* items, uint32_t n_items,
                                   uint32_t n_cands, const uint32_t* left_slots, const uint32_t* left_labs,
                                   uint32_t* cand_pool,
                                   uint32_t* bitmaps, uint64_t words_per_slot, unsigned long long* member,
                                   uint2* lists, uint64_t list_stride, uint32_t* n_lists, GqlRound rd,
                                   unsigned long long* bytes, hipStream_t s);
hipError_t launch_gql_hall(uint32_t lanes, const DevGraph& g, const GItem* items, const uint2* list,
                           const uint32_t* n_list, uint32_t max_n, const uint32_t* left_slots,
                           const uint32_t* left_labs, uint32_t* cand_pool, uint32_t* bitmaps, uint64_t words_per_slot,
                           unsigned long long* member, GqlRound rd, unsigned long long* bytes, hipStream_t s);
// cursor: zero before the launch (the wavefronts' shared position in list);
// left_lab: the label of each left vertex (parallel to left_slots)
hipError_t launch_gql_refine(const DevGraph& g, const GItem* items, const uint2* list, const uint32_t* n_list,
                             uint32_t* cursor, uint32_t max_n, const uint32_t* left_slots, const uint32_t* left_lab,
                             const uint32_t* left_labs, uint32_t* cand_pool, uint32_t* bitmaps,
                             uint64_t words_per_slot, unsigned long long* member, GqlRound rd,
                             unsigned long long* bytes, hipStream_t s);
// Second-round skip flags of the n_items items of one vertex position (see
// GqlRound): removed1 / removed2 are the first and second rounds' flags.
hipError_t launch_gql_skip(const GItem* items, uint32_t n_items, const uint32_t* left_slots,
                           const uint32_t* removed1, const uint32_t* removed2, uint32_t* skip, hipStream_t s);
// thread per row; rows of candidates with more than 64 neighbours are listed
// in long_rows (count in *n_long) for launch_table_long
hipError_t launch_table_count(const DevGraph& g, const TItem* items, uint32_t n_items,
                              const uint64_t* item_row_start, uint64_t n_rows,
                              const uint32_t* cand_pool, const uint32_t* bitmaps,
                              uint64_t words_per_slot, uint64_t* row_deg, uint64_t* long_rows,
                              unsigned long long* n_long, unsigned long long* bytes, hipStream_t s);
// wavefront per listed row: the count pass (fill = 0) or the fill pass
hipError_t launch_table_long(int fill, const DevGraph& g, const TItem* items, uint32_t n_items,
                             const uint64_t* item_row_start, const uint64_t* long_rows, uint64_t n_long,
                             const uint32_t* cand_pool, const uint32_t* bitmaps, const uint32_t* word_prefix,
                             uint64_t words_per_slot, uint64_t* row_deg, const uint64_t* row_off, uint32_t* idx,
                             unsigned long long* bytes, hipStream_t s);
hipError_t launch_table_fill(const DevGraph& g, const TItem* items, uint32_t n_items,
                             const uint64_t* item_row_start, uint64_t n_rows,
                             const uint32_t* cand_pool, const uint32_t* bitmaps,
                             const uint32_t* word_prefix, uint64_t words_per_slot,
                             const uint64_t* row_off, uint32_t* idx, unsigned long long* bytes,
                             hipStream_t s);
// defer: room for len pivot candidates (those with more than 64 neighbours
// are handled by wavefronts), n_defer: one counter
hipError_t launch_cfl_support(const DevGraph& g, const uint32_t* cand_pool, uint64_t cand_begin,
                              uint32_t len, uint32_t label, uint32_t degree, uint32_t* cnt,
                              uint32_t expect, uint32_t next, uint32_t* defer, uint32_t* n_defer,
                              hipStream_t s);
hipError_t launch_cfl_prune(const uint32_t* cnt, uint32_t* cand_pool, uint64_t cand_begin,
                            uint32_t len, uint32_t target, hipStream_t s);
hipError_t launch_expand(const DPlan* plans, const DOp* ops, const DTask* tasks,
                         uint32_t n_tasks, const uint8_t* expand_q, uint32_t* out_count,
                         const uint64_t* out_pos, DTask* out_tasks, QCounters* qc,
                         unsigned long long* qchild, uint32_t nq, uint32_t max_depth, int write,
                         hipStream_t s);
// Copies segments of task records (dst_off: exclusive prefix of the segment
// lengths, n_seg + 1 entries; total = dst_off[n_seg]).
hipError_t launch_gather_tasks(const DTask* src, const uint64_t* src_off, const uint64_t* dst_off,
                               uint32_t n_seg, uint64_t total, DTask* dst, hipStream_t s);
hipError_t launch_dfs(const DPlan* plans, const DOp* ops, const DTask* init, DTask* queue, uint32_t* seq,
                      DQueue* dq, QCounters* qc, uint32_t* pending, uint32_t max_n, uint32_t grid_waves,
                      uint32_t block_lanes, EmitSink sink, hipStream_t s);
hipError_t launch_join_build(const uint32_t* rows, uint64_t n_rows, uint32_t width, JoinKeys keys,
                             unsigned long long* fp, uint32_t* head, uint64_t mask, uint32_t* next,
                             hipStream_t s);
hipError_t launch_join_probe_emit(const uint32_t* build_rows, uint32_t build_width, JoinKeys build_keys,
                                  const uint32_t* probe_rows, uint64_t n_probe, uint32_t probe_width,
                                  JoinKeys probe_keys, const unsigned long long* fp, const uint32_t* head,
                                  uint64_t mask, const uint32_t* next, uint32_t* out, unsigned long long cap,
                                  unsigned long long* count, uint32_t* overflow, hipStream_t s);
hipError_t launch_join_probe(const uint32_t* build_rows, uint32_t build_width, JoinKeys build_keys,
                             const uint32_t* probe_rows, uint64_t n_probe, uint32_t probe_width,
                             JoinKeys probe_keys, const unsigned long long* fp, const uint32_t* head,
                             uint64_t mask, const uint32_t* next, unsigned long long* count,
                             unsigned long long* bytes, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s);
// n root range tasks of query q splitting C(order[0]) (len entries) evenly,
// written to tasks[0, n)
hipError_t launch_root_tasks(DTask* tasks, uint32_t n, uint32_t q, uint64_t len, hipStream_t s);
// ring sequence words: seq[i] = i + 1 for the n_tasks queued positions, i for
// the free slots
hipError_t launch_ring_init(uint32_t* seq, uint32_t slots, uint32_t n_tasks, hipStream_t s);
size_t dfs_lane_lds_bytes(uint32_t max_n);
size_t dfs_wave_lds_bytes(uint32_t max_n);
hipError_t launch_dfs_wave(const DPlan* plans, const DOp* ops, const DTask* init, DTask* queue, uint32_t* seq,
                           DQueue* dq, QCounters* qc, uint32_t* pending, uint32_t max_n, uint32_t grid_waves,
                           hipStream_t s);
uint64_t scan_tmp_entries(uint64_t n);

// ---- fixed device scratch of the allocation-free enumeration -------------
// Layout (byte offsets, 256-aligned): queue header, query counters, pending
// count, one plan, the operands, the ring's sequence words and records, the
// root tasks.
constexpr uint32_t kScratchOps = 2048;       // >= the edges of a 64-vertex query
constexpr uint32_t kScratchSlots = 1u << 16;  // task ring slots
constexpr uint32_t kScratchRootTasks = 4096;  // root range tasks at most
constexpr size_t kScratchAlign(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t kScrQueue = 0;
constexpr size_t kScrCounters = kScrQueue + kScratchAlign(sizeof(DQueue));
constexpr size_t kScrPending = kScrCounters + kScratchAlign(sizeof(QCounters));
constexpr size_t kScrZeroEnd = kScrPending + 256;   // [0, kScrZeroEnd) is cleared per call
constexpr size_t kScrPlan = kScrZeroEnd;
constexpr size_t kScrOps = kScrPlan + kScratchAlign(sizeof(DPlan));
constexpr size_t kScrSeq = kScrOps + kScratchAlign(sizeof(DOp) * kScratchOps);
constexpr size_t kScrRing = kScrSeq + kScratchAlign(4ull * kScratchSlots);
constexpr size_t kScrInit = kScrRing + kScratchAlign(sizeof(DTask) * kScratchSlots);
constexpr size_t kScratchBytes = kScrInit + kScratchAlign(sizeof(DTask) * kScratchRootTasks);
constexpr uint32_t kScratchMaxDevices = 64;
static_assert(kScratchBytes <= (12u << 20), "the scratch stays modest");
// resolves the scratch's address on `device` (which must be current); once
// per device, thread-safe
hipError_t scratch_setup(int device);
// the address scratch_setup resolved on `device`, nullptr before
void* scratch_address(int device);
hipError_t launch_scan_u32(const uint32_t* in, uint64_t n, uint64_t* out, unsigned long long* tmp,
                           hipStream_t s);
hipError_t launch_scan_u64(const uint64_t* in, uint64_t n, uint64_t* out, unsigned long long* tmp,
                           hipStream_t s);
hipError_t launch_slot_prefix(const uint32_t* bitmaps, uint32_t* prefix, uint32_t n_slots,
                              uint64_t words_per_slot, hipStream_t s);
hipError_t launch_span(const uint64_t* row_off, const uint64_t* starts, const uint64_t* ends,
                       uint32_t n, uint64_t* out, hipStream_t s);

}  // namespace psm
