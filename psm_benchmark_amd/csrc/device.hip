// device.hip — HIP kernels of the subgraph matcher, written for gfx950
// (CDNA4, 64-wide wavefronts). All arithmetic is integer; nothing here is a
// dense contraction, so the kernels are bound by memory traffic and latency.
//
//   compaction   order-preserving filter of label buckets / candidate lists
//                (LDF, NLF, CFL generation, removal of pruned candidates)
//   bitmaps      per (query, vertex) candidate bitmaps + per-word ranks
//   gql_refine   GQL global refinement: one wavefront per candidate tests
//                for a matching that saturates N_q(u) inside N(v)
//   tables       BuildTable rows (count pass, scan, fill pass)
//   cfl          CFL support counting and pruning
//   expand/dfs   enumeration: task pre-expansion and the persistent
//                lane-per-task DFS fed by a global atomic work queue

#include <algorithm>
#include <mutex>

#include "internal.h"

namespace psm {

namespace {

constexpr int kBlock = 256;
constexpr int kWave = 64;

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & (kWave - 1); }

__device__ __forceinline__ unsigned long long lanemask_lt() {
    return (1ull << lane_id()) - 1ull;
}

// Sum of v over the wavefront, added once to *p (all lanes must call).
__device__ __forceinline__ void wave_add(unsigned long long* p, unsigned long long v) {
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)v, o), hi = __shfl_xor((uint32_t)(v >> 32), o);
        v += ((unsigned long long)hi << 32) | lo;
    }
    if (lane_id() == 0 && v) atomicAdd(p, v);
}

// Sum of v over the wavefront, added once to *p (all lanes must call).
__device__ __forceinline__ void wave_add_u32(unsigned int* p, uint32_t v) {
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (lane_id() == 0 && v) atomicAdd(p, v);
}

// Sum of v over the block, added once to *p. Every thread of the block must
// call it (it synchronises the block).
__device__ __forceinline__ void block_add(unsigned long long* p, unsigned long long v) {
    __shared__ unsigned long long acc;
    if (threadIdx.x == 0) acc = 0;
    __syncthreads();
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)v, o), hi = __shfl_xor((uint32_t)(v >> 32), o);
        v += ((unsigned long long)hi << 32) | lo;
    }
    if (lane_id() == 0 && v) atomicAdd(&acc, v);
    __syncthreads();
    if (threadIdx.x == 0 && acc) atomicAdd(p, acc);
}

// Largest i in [0, n) with starts[i] <= key (starts ascending, starts[0] == 0).
template <typename T, typename K>
__device__ __forceinline__ uint32_t find_item(const T* items, uint32_t n, K key, K T::*field) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (items[mid].*field <= key) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t find_u64(const uint64_t* starts, uint32_t n, uint64_t key) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (starts[mid] <= key) lo = mid;
        else hi = mid;
    }
    return lo;
}

// FilterVertices.cpp:786-803: the data vertex's neighbour-label counts
// (runs[i, ie)) dominate the query vertex's, both held as label-sorted runs.
__device__ bool nlf_dominates(const uint2* __restrict__ runs, uint32_t i, uint32_t ie, const uint2* qr,
                              uint32_t qn) {
    for (uint32_t j = 0; j < qn; ++j) {
        const uint2 r = qr[j];
        bool found = false;
        while (i < ie) {
            const uint2 d = runs[i];
            if (d.x < r.x) { ++i; continue; }
            if (d.x == r.x && d.y >= r.y) { found = true; ++i; }
            break;
        }
        if (!found) return false;
    }
    return true;
}

// The run of label l among v's neighbours in the label-grouped adjacency:
// (first position in adj_bl, length); length 0 when v has no such neighbour.
__device__ __forceinline__ uint2 label_run(const DevGraph& g, uint32_t v, uint32_t l) {
    const uint32_t end = g.nlf_off[v + 1];
    uint32_t lo = g.nlf_off[v], hi = end;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (g.nlf[mid].x < l) lo = mid + 1;
        else hi = mid;
    }
    if (lo < end) {
        const uint2 r = g.nlf[lo];
        if (r.x == l) return make_uint2(g.nlf_beg[lo], r.y);
    }
    return make_uint2(0u, 0u);
}

// ---------------------------------------------------------------------------
// Order-preserving compaction, pass 1: predicate per element, one 64-bit
// mask per wavefront and a survivor count per block.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_compact_flag(
    DevGraph g, const CItem* __restrict__ items, uint32_t n_items, const QVert* __restrict__ qv,
    const uint2* __restrict__ qnlf, const uint32_t* __restrict__ src_pool,
    const uint32_t* __restrict__ cnt_pool, unsigned long long* __restrict__ masks,
    uint32_t* __restrict__ block_counts) {
    __shared__ uint32_t s_item;
    __shared__ uint32_t s_cnt[kBlock / kWave];
    const uint32_t b = blockIdx.x;
    if (threadIdx.x == 0) s_item = find_item(items, n_items, b, &CItem::block_start);
    __syncthreads();
    const CItem it = items[s_item];
    const uint32_t e = (b - it.block_start) * kBlock + threadIdx.x;
    bool keep = false;
    if (e < it.len) {
        if (it.mode == kModeValid) {
            keep = src_pool[it.src_begin + e] != kInvalid;
        } else {
            // the bucket's degrees and runs are read in reverse-index order
            const uint32_t j = it.src_begin + e;
            const QVert q = qv[it.qv];
            keep = g.deg_ri[j] >= q.degree;
            if (keep && it.mode == kModeCflGen) keep = cnt_pool[it.cnt_base + g.ri[j]] == it.target;
            if (keep && it.mode != kModeLDF)
                keep = nlf_dominates(g.nlf_ri, g.nlf_ri_off[j], g.nlf_ri_off[j + 1], qnlf + q.nlf_begin, q.nlf_count);
        }
    }
    const unsigned long long m = __ballot(keep);
    const uint32_t w = threadIdx.x / kWave;
    if (lane_id() == 0) {
        masks[(size_t)b * (kBlock / kWave) + w] = m;
        s_cnt[w] = __popcll(m);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t c = 0;
        for (int i = 0; i < kBlock / kWave; ++i) c += s_cnt[i];
        block_counts[b] = c;
    }
}

// Pass 3: survivors written at their block's exclusive offset.
__global__ __launch_bounds__(kBlock) void k_compact_scatter(
    const CItem* __restrict__ items, uint32_t n_items, const uint32_t* __restrict__ src_ri,
    const uint32_t* __restrict__ src_pool, const unsigned long long* __restrict__ masks,
    const uint64_t* __restrict__ block_prefix, uint32_t* __restrict__ out) {
    __shared__ uint32_t s_item;
    const uint32_t b = blockIdx.x;
    if (threadIdx.x == 0) s_item = find_item(items, n_items, b, &CItem::block_start);
    __syncthreads();
    const CItem it = items[s_item];
    const uint32_t w = threadIdx.x / kWave;
    const unsigned long long m = masks[(size_t)b * (kBlock / kWave) + w];
    if (!((m >> lane_id()) & 1ull)) return;
    uint64_t pos = block_prefix[b];
    for (uint32_t i = 0; i < w; ++i) pos += __popcll(masks[(size_t)b * (kBlock / kWave) + i]);
    pos += __popcll(m & lanemask_lt());
    const uint32_t e = (b - it.block_start) * kBlock + threadIdx.x;
    out[pos] = it.mode == kModeValid ? src_pool[it.src_begin + e] : src_ri[it.src_begin + e];
}

// Label-shared bucket scan, pass 1 (FilterVertices.cpp:19-65 LDF / NLF for
// every (query, u) of a batch): a wavefront takes chunk r of its group's
// bucket (lane l: element r * 64 + l) and subset s of the group's items
// (lane k holds item s * 64 + k), and tests the chunk against each item of
// the subset in turn: the element's degree and run bounds stay in
// registers, its runs come from the L1 after the first item, the item's
// fields come from lane k by readlane. Lane k stores item k's mask and
// count at the end.
static_assert(kBucketChunk == kWave, "a bucket chunk is one element per lane");
__global__ __launch_bounds__(kBlock) void k_bucket_flag(
    DevGraph g, const CGroup* __restrict__ groups, uint32_t n_groups, uint32_t n_waves,
    const CBucketItem* __restrict__ bitems, const uint2* __restrict__ qnlf,
    unsigned long long* __restrict__ masks, uint32_t* __restrict__ chunk_counts) {
    const uint32_t lane = lane_id();
    const uint32_t wv = (uint32_t)__builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave));
    if (wv >= n_waves) return;   // whole wavefronts: no barrier below
    const CGroup gr = groups[find_item(groups, n_groups, wv, &CGroup::wave_start)];
    const uint32_t local = wv - gr.wave_start;
    const uint32_t r = local / gr.n_sub, sub = local % gr.n_sub;
    const uint32_t first = sub * kWave;
    const uint32_t n = min(gr.n_items - first, (uint32_t)kWave);
    CBucketItem mine{};
    if (lane < n) mine = bitems[gr.item_begin + first + lane];
    const uint32_t e = r * kWave + lane;
    const bool in = e < gr.len;
    const uint64_t j = gr.src_begin + e;
    uint32_t deg = 0, o0 = 0, o1 = 0;
    if (in) {
        deg = g.deg_ri[j];
        o0 = g.nlf_ri_off[j];
        o1 = g.nlf_ri_off[j + 1];
    }
    unsigned long long my_m = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t qd = (uint32_t)__builtin_amdgcn_readlane((int)mine.degree, (int)k);
        const uint32_t qb = (uint32_t)__builtin_amdgcn_readlane((int)mine.nlf_begin, (int)k);
        const uint32_t qc = (uint32_t)__builtin_amdgcn_readlane((int)mine.nlf_count, (int)k);
        bool keep = in && deg >= qd;
        if (keep && !(qc >> 31)) keep = nlf_dominates(g.nlf_ri, o0, o1, qnlf + qb, qc);
        const unsigned long long m = __ballot(keep);
        if (lane == k) my_m = m;
    }
    if (lane < n) {
        masks[mine.chunk_start + r] = my_m;
        chunk_counts[mine.chunk_start + r] = (uint32_t)__popcll(my_m);
    }
}

// Pass 2: each item's survivors of chunk r at the chunk's exclusive offset,
// as vertex ids (the bucket's reverse index); lane k loads item k's mask and
// offset, and the loop reads them by readlane.
__global__ __launch_bounds__(kBlock) void k_bucket_scatter(
    DevGraph g, const CGroup* __restrict__ groups, uint32_t n_groups, uint32_t n_waves,
    const CBucketItem* __restrict__ bitems, const unsigned long long* __restrict__ masks,
    const uint64_t* __restrict__ chunk_prefix, uint32_t* __restrict__ out) {
    const uint32_t lane = lane_id();
    const uint32_t wv = (uint32_t)__builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave));
    if (wv >= n_waves) return;
    const CGroup gr = groups[find_item(groups, n_groups, wv, &CGroup::wave_start)];
    const uint32_t local = wv - gr.wave_start;
    const uint32_t r = local / gr.n_sub, sub = local % gr.n_sub;
    const uint32_t first = sub * kWave;
    const uint32_t n = min(gr.n_items - first, (uint32_t)kWave);
    unsigned long long my_m = 0;
    uint64_t my_p = 0;
    if (lane < n) {
        const uint32_t c = bitems[gr.item_begin + first + lane].chunk_start + r;
        my_m = masks[c];
        my_p = chunk_prefix[c];
    }
    const uint32_t e = r * kWave + lane;
    const uint32_t vid = e < gr.len ? g.ri[gr.src_begin + e] : 0u;
    const unsigned long long below = lanemask_lt();
    for (uint32_t k = 0; k < n; ++k) {
        const unsigned long long m = ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(my_m >> 32), (int)k) << 32) |
                                     (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)my_m, (int)k);
        if (!m) continue;
        const uint64_t p = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(my_p >> 32), (int)k) << 32) |
                           (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)my_p, (int)k);
        if ((m >> lane) & 1ull) out[p + __popcll(m & below)] = vid;
    }
}

// ---------------------------------------------------------------------------
// Candidate bitmaps: one bit per data vertex per (query, query vertex) slot.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_set_bits(
    const uint32_t* __restrict__ pool, const uint64_t* __restrict__ seg_begin,
    const uint32_t* __restrict__ seg_len, const uint32_t* __restrict__ seg_slot, uint32_t n_seg,
    uint32_t* __restrict__ bitmaps, uint64_t wps) {
    // grid.y = segment, grid.x strides its elements
    const uint32_t s = blockIdx.y;
    if (s >= n_seg) return;
    const uint32_t len = seg_len[s];
    uint32_t* bm = bitmaps + (size_t)seg_slot[s] * wps;
    const uint32_t* src = pool + seg_begin[s];
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < len; i += gridDim.x * kBlock) {
        const uint32_t v = src[i];
        if (v != kInvalid) atomicOr(bm + (v >> 5), 1u << (v & 31));
    }
}

// member[mem_base + v] |= 1 << u for every candidate v of segment (q, u).
__global__ __launch_bounds__(kBlock) void k_set_member(
    const uint32_t* __restrict__ pool, const uint64_t* __restrict__ seg_begin,
    const uint32_t* __restrict__ seg_len, const uint64_t* __restrict__ seg_mem,
    const uint32_t* __restrict__ seg_bit, uint32_t n_seg, unsigned long long* __restrict__ member) {
    const uint32_t s = blockIdx.y;
    if (s >= n_seg) return;
    const uint32_t len = seg_len[s];
    unsigned long long* mw = member + seg_mem[s];
    const unsigned long long bit = 1ull << seg_bit[s];
    const uint32_t* src = pool + seg_begin[s];
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < len; i += gridDim.x * kBlock) {
        const uint32_t v = src[i];
        if (v != kInvalid) atomicOr(mw + v, bit);
    }
}

__global__ __launch_bounds__(kBlock) void k_popc_words(const uint32_t* __restrict__ bm,
                                                        uint32_t* __restrict__ out, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * kBlock)
        out[i] = __popc(bm[i]);
}

// ---------------------------------------------------------------------------
// GQL global refinement (FilterVertices.cpp:104-124, verifyExactTwigIso
// :704-741). One wavefront per candidate v of query vertex u. Left side:
// the k query neighbours u'_i of u; right side: N(v); edge (i, w) iff w is a
// live candidate of u'_i. v survives iff a matching saturates the left side.
//
// A left vertex with at least k neighbours can always be matched after the
// others (at most k-1 right vertices are taken by then), so only the set T
// of left vertices with fewer than k neighbours needs an explicit matching;
// their neighbours number < k*|T| <= 31*31. Those are compacted into LDS and
// matched by Kuhn's augmenting paths, each path search run by the whole
// wavefront (64 right vertices per step).
// ---------------------------------------------------------------------------
constexpr int kRightCap = 1024;
constexpr int kGqlWavesPerBlock = kBlock / kWave;

struct GqlShared {
    uint32_t rmask[kGqlWavesPerBlock][kRightCap];
    uint8_t match_r[kGqlWavesPerBlock][kRightCap];
    uint8_t seen[kGqlWavesPerBlock][kRightCap];
    uint32_t st_left[kGqlWavesPerBlock][kMaxGqlDegree + 1];
    uint32_t st_pos[kGqlWavesPerBlock][kMaxGqlDegree + 1];
    uint32_t st_via[kGqlWavesPerBlock][kMaxGqlDegree + 1];
};

// Membership words: bit u of member[mem_base + w] is set when data vertex w
// is a live candidate of query vertex u, so the left vertices a neighbour w
// can serve come from one load; left vertex i of an item (the i-th query
// neighbour of u, bitmap slot left_slots[left_begin + i]) owns bit
// left_slots[left_begin + i] - (bm_self - u). The kernels keep those bit
// positions in registers, one per lane.

constexpr uint32_t kScanUnroll = 4;   // 64-neighbour chunks a wavefront loads before testing

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t x) {
    for (int o = kWave / 2; o > 0; o >>= 1) x = min(x, (uint32_t)__shfl_xor(x, o));
    return x;
}
__device__ __forceinline__ uint32_t wave_or_u32(uint32_t x) {
    for (int o = kWave / 2; o > 0; o >>= 1) x |= (uint32_t)__shfl_xor(x, o);
    return x;
}

// The run index of concatenated position p: the largest li < nlab with
// start[li] <= p, where lane li holds start[li] (ascending). The loop count
// depends on nlab only, so every lane of the wavefront runs it (it calls
// __shfl with a per-lane source).
__device__ __forceinline__ uint32_t run_of(uint32_t start, uint32_t nlab, uint32_t p) {
    uint32_t lo = 0;
    for (uint32_t s = 32; s > 0; s >>= 1)
        if (s < nlab) {
            const uint32_t st = __shfl(start, min(lo + s, nlab - 1));
            if (lo + s < nlab && st <= p) lo += s;
        }
    return lo;
}

// The first concatenated position at or after `from` inside a run whose bit
// is set in `labels` (lane li holds its run's start and length), or
// kInvalid. Every lane calls it.
__device__ __forceinline__ uint32_t next_pos(uint32_t start, uint32_t len, uint32_t labels, uint32_t from) {
    const uint32_t lane = lane_id();
    const bool in = lane < 32 && ((labels >> lane) & 1u) && start + len > from;
    return wave_min_u32(in ? max(start, from) : kInvalid);
}

// One candidate (item it, index j) per wavefront; returns its algorithmic
// bytes (on every lane). wl: the wavefront's index in its block.
__device__ __forceinline__ unsigned long long gql_refine_one(
    const DevGraph& g, const GItem& it, uint32_t j, const uint32_t* __restrict__ left_slots,
    const uint32_t* __restrict__ left_lab, const uint32_t* __restrict__ left_labs, uint32_t* __restrict__ cand_pool, uint32_t* __restrict__ bitmaps, uint64_t wps,
    unsigned long long* __restrict__ member, uint32_t* __restrict__ removed, GqlShared& sh, uint32_t wl) {
    const uint32_t lane = lane_id();
    if (j >= it.len) return 0;
    const uint32_t v = cand_pool[it.cand_begin + j];
    if (v == kInvalid) return 0;
    const uint32_t k = it.k;
    const uint32_t nlab = it.nlab;   // distinct left labels, at most k <= 32
    // Lane i < k holds left vertex i's bit in a membership word and lane
    // li < nlab the li-th distinct left label (one load each per candidate;
    // the loops below read them with readlane instead of reloading them)
    const uint32_t my_shift = lane < k ? left_slots[it.left_begin + lane] - (it.bm_self - it.u) : 0u;
    const uint32_t my_want = lane < nlab ? left_labs[it.lab_begin + lane] : kInvalid;
    // Right vertices can only be neighbours carrying a left vertex's label,
    // and the neighbours of one label can only serve the left vertices of
    // that label (candidate sets are label-pure): the scans below walk the
    // runs of the left labels in the label-grouped adjacency, taken one
    // after the other as one sequence.
    // Lane li < nlab gets the run (start in adj_bl, length) of its label my_want
    // among v's neighbours: v's runs are read 64 at a time (coalesced) and
    // matched against the wanted labels by ballots.
    uint32_t run_x = 0, run_y = 0;
    {
        const uint32_t r0 = g.nlf_off[v], r1 = g.nlf_off[v + 1];
        for (uint32_t base = r0; base < r1; base += kWave) {
            const uint32_t jj = base + lane;
            uint32_t lab_j = kInvalid, beg_j = 0, cnt_j = 0;
            if (jj < r1) {
                const uint2 r = g.nlf[jj];
                lab_j = r.x;
                cnt_j = r.y;
                beg_j = g.nlf_beg[jj];
            }
            for (uint32_t li = 0; li < nlab; ++li) {
                const unsigned long long bal =
                    __ballot(lab_j == (uint32_t)__builtin_amdgcn_readlane((int)my_want, (int)li));
                if (bal) {
                    const uint32_t src = __ffsll((long long)bal) - 1;
                    const uint32_t x = __shfl(beg_j, src), y = __shfl(cnt_j, src);
                    if (lane == li) {
                        run_x = x;
                        run_y = y;
                    }
                }
            }
        }
    }
    // the runs' positions in the concatenation (exclusive prefix)
    uint32_t incl = lane < nlab ? run_y : 0u;
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o);
        if ((int)lane >= o) incl += t;
    }
    const uint32_t start = incl - (lane < nlab ? run_y : 0u);
    const uint32_t total = __shfl(incl, kWave - 1);
    // lane i < k: left vertex i, its label's run index and that run's end,
    // and how many left vertices share its label
    const uint32_t my_lab = lane < k ? left_lab[it.left_begin + lane] : kInvalid;
    uint32_t my_li = 0, my_sat = 0;
    for (uint32_t li = 0; li < nlab; ++li) {
        const uint32_t want = (uint32_t)__builtin_amdgcn_readlane((int)my_want, (int)li);
        const uint32_t same = (uint32_t)__popcll(__ballot(my_lab == want));
        if (my_lab == want) {
            my_li = li;
            my_sat = same;
        }
    }
    const uint32_t my_end = __shfl(incl, my_li);
    const uint32_t my_len = __shfl(run_y, my_li);
    // pass 1: lane i (< k) counts the right neighbours of left vertex i,
    // saturating at my_sat, the number of left vertices with its label. A
    // right vertex serves only left vertices of its own label, so the
    // bipartite graph is a disjoint union of one part per label, and in its
    // part a left vertex with my_sat right neighbours can always be matched
    // last (each in turn still finds an unused one). The rest of a run is
    // skipped once every left vertex of its label is saturated; a left
    // vertex whose run holds no right neighbour fails the candidate. A step
    // loads kScanUnroll chunks of 64 positions together (their neighbour and
    // membership loads are independent and in flight at once).
    bool ok = __ballot(lane < k && my_len == 0) == 0;
    uint32_t my_cnt = 0;
    uint32_t scanned = 0, rescanned = 0;
    const uint32_t all_labels = nlab >= 32 ? 0xFFFFFFFFu : (1u << nlab) - 1u;
    uint32_t o = ok ? next_pos(start, run_y, all_labels, 0u) : kInvalid;
    while (o < total) {
        uint32_t w[kScanUnroll];
        unsigned long long mw[kScanUnroll];
#pragma unroll
        for (uint32_t c = 0; c < kScanUnroll; ++c) {
            const uint32_t p = o + c * kWave + lane;
            const uint32_t li = run_of(start, nlab, min(p, total - 1));
            const uint32_t rx = __shfl(run_x, li), rs = __shfl(start, li);
            w[c] = p < total ? g.adj_bl[rx + (p - rs)] : kInvalid;
        }
#pragma unroll
        for (uint32_t c = 0; c < kScanUnroll; ++c) mw[c] = w[c] != kInvalid ? member[it.mem_base + w[c]] : 0ull;
        for (uint32_t i = 0; i < k; ++i) {
            const uint32_t sh_i = (uint32_t)__builtin_amdgcn_readlane((int)my_shift, (int)i);
            uint32_t cnt = 0;
#pragma unroll
            for (uint32_t c = 0; c < kScanUnroll; ++c) cnt += __popcll(__ballot((mw[c] >> sh_i) & 1ull));
            if (lane == i) my_cnt = min(my_cnt + cnt, my_sat);
        }
        const uint32_t nxt = o + kScanUnroll * kWave;
        scanned += min(kScanUnroll * kWave, total - o);
        // a left vertex whose run has been passed without a right neighbour
        if (__ballot(lane < k && my_cnt == 0 && my_end <= nxt)) {
            ok = false;
            break;
        }
        // labels that still have an unsaturated left vertex
        const uint32_t unsat = wave_or_u32(lane < k && my_cnt < my_sat ? 1u << my_li : 0u);
        if (!unsat) break;
        o = next_pos(start, run_y, unsat, nxt);
    }
    // every left vertex needs a right neighbour (the scan above leaves no
    // run of an unsaturated left vertex unread)
    if (ok && __ballot(lane < k && my_cnt == 0)) ok = false;
    const uint32_t T = ok ? (uint32_t)__ballot(lane < k && my_cnt < my_sat) : 0u;
    const bool pass2 = ok && T;
    if (pass2) {
        // pass 2: right vertices adjacent to T (only the runs of T's labels),
        // compacted into LDS
        const uint32_t tlabels = wave_or_u32(((T >> lane) & 1u) ? 1u << my_li : 0u);
        uint32_t nr = 0;
        uint32_t o2 = next_pos(start, run_y, tlabels, 0u);
        while (o2 < total) {
            const uint32_t p = o2 + lane;
            const uint32_t li = run_of(start, nlab, min(p, total - 1));
            const uint32_t rx = __shfl(run_x, li), rs = __shfl(start, li);
            const bool in = p < total && ((tlabels >> li) & 1u);
            const unsigned long long mw = in ? member[it.mem_base + g.adj_bl[rx + (p - rs)]] : 0ull;
            uint32_t m = 0;
            for (uint32_t tt = T; tt; tt &= tt - 1) {
                const uint32_t i = __ffs(tt) - 1;
                m |= (uint32_t)((mw >> (uint32_t)__builtin_amdgcn_readlane((int)my_shift, (int)i)) & 1ull) << i;
            }
            rescanned += (uint32_t)__popcll(__ballot(in));
            const unsigned long long bal = __ballot(m != 0);
            if (m) {
                const uint32_t r = nr + __popcll(bal & lanemask_lt());
                if (r < kRightCap) sh.rmask[wl][r] = m;
            }
            nr += __popcll(bal);
            o2 = next_pos(start, run_y, tlabels, o2 + kWave);
        }
        nr = min(nr, (uint32_t)kRightCap);  // each of T's vertices has < 32 right neighbours: < 32*32
        for (uint32_t r = lane; r < nr; r += kWave) sh.match_r[wl][r] = 0xFF;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        // Kuhn: augment from each left vertex of T in turn
        for (uint32_t i = 0; i < k && ok; ++i) {
            if (!((T >> i) & 1u)) continue;
            for (uint32_t r = lane; r < nr; r += kWave) sh.seen[wl][r] = 0;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            int sp = 0;
            if (lane == 0) { sh.st_left[wl][0] = i; sh.st_pos[wl][0] = 0; }
            bool found = false;
            while (sp >= 0) {
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                const uint32_t l = sh.st_left[wl][sp];
                const uint32_t st = sh.st_pos[wl][sp];
                uint32_t hit = kInvalid;
                for (uint32_t c0 = st & ~(uint32_t)(kWave - 1); c0 < nr; c0 += kWave) {
                    const uint32_t r = c0 + lane;
                    const bool h = r >= st && r < nr && ((sh.rmask[wl][r] >> l) & 1u) &&
                                   !sh.seen[wl][r];
                    const unsigned long long bal = __ballot(h);
                    if (bal) { hit = c0 + __ffsll((long long)bal) - 1; break; }
                }
                if (hit == kInvalid) { --sp; continue; }
                const uint32_t owner = sh.match_r[wl][hit];
                if (lane == 0) {
                    sh.seen[wl][hit] = 1;
                    sh.st_pos[wl][sp] = hit + 1;
                    sh.st_via[wl][sp] = hit;
                }
                if (owner == 0xFF) { found = true; break; }
                ++sp;
                if (lane == 0) { sh.st_left[wl][sp] = owner; sh.st_pos[wl][sp] = 0; }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            if (!found) { ok = false; break; }
            if (lane == 0)
                for (int s = 0; s <= sp; ++s) sh.match_r[wl][sh.st_via[wl][s]] = (uint8_t)sh.st_left[wl][s];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        }
    }
    if (!ok && lane == 0) {
        cand_pool[it.cand_begin + j] = kInvalid;
        atomicAnd(bitmaps + (size_t)it.bm_self * wps + (v >> 5), ~(1u << (v & 31)));
        atomicAnd(member + it.mem_base + v, ~(1ull << it.u));
        removed[it.bm_self] = 1u;
    }
    // algorithmic bytes: candidate, one run lookup per left label, the
    // scanned part of the label runs, one membership word per (neighbour,
    // left vertex) probe; pass 2 rescans the runs when it runs
    return 4ull + 16ull * nlab * (pass2 ? 2 : 1) + 4ull * (scanned + rescanned) * (1 + k);
}

// The wavefront test over the deferred (item, index) pairs of class 3; the
// pair count is read from device memory (written by the 8-lane kernel).
// Wavefronts take pairs one at a time from a shared cursor (zero before the
// launch): a candidate's cost follows the length of its label runs, which is
// heavy-tailed on power-law graphs, so a fixed striding would leave the
// wavefronts that drew hubs finishing long after the rest.
__global__ __launch_bounds__(kBlock) void k_gql_refine(
    DevGraph g, const GItem* __restrict__ items, const uint2* __restrict__ list,
    const uint32_t* __restrict__ n_list, uint32_t* __restrict__ cursor, const uint32_t* __restrict__ left_slots,
    const uint32_t* __restrict__ left_lab, const uint32_t* __restrict__ left_labs,
    uint32_t* __restrict__ cand_pool, uint32_t* __restrict__ bitmaps,
    uint64_t wps, unsigned long long* __restrict__ member, GqlRound rd, unsigned long long* __restrict__ bytes) {
    __shared__ GqlShared sh;
    const uint32_t n = *n_list;
    const uint32_t wl = threadIdx.x / kWave;
    unsigned long long by = 0;
    while (true) {
        uint32_t w = 0;
        if (lane_id() == 0) w = atomicAdd(cursor, 1u);
        w = __shfl(w, 0);
        if (w >= n) break;
        const uint2 e = list[w];
        const GItem it = items[e.x];   // in registers for the whole test
        by += gql_refine_one(g, it, e.y, left_slots, left_lab, left_labs, cand_pool, bitmaps, wps, member,
                             rd.removed, sh, wl);
    }
    block_add(bytes, lane_id() == 0 ? by : 0ull);
}


// GQL refinement for small candidates: 8 lanes per candidate v when
// deg(v) <= 8 and k = deg_q(u) <= 8 (the common case on power-law graphs).
// Lane j of the group holds neighbour w_j and its k-bit mask; the group's
// right-neighbour set R_i of every left vertex i comes out of one ballot per
// i. By Hall's theorem a matching saturating the left side exists iff every
// subset S of left vertices has |union of R_i, i in S| >= |S|; the 2^k - 1
// subsets are split over the group's 8 lanes. Candidates outside the limits
// are deferred by degree class: k <= 8 and deg(v) <= 16 / 32 / 64 to the
// same test with 16 / 32 / 64 lanes per candidate (k_gql_hall), the rest to
// the wavefront kernel above. Deferred lists: class c at lists + c * cap.
constexpr uint32_t kGroup = 8;
constexpr uint32_t kDeferClasses = 4;   // 16, 32, 64 lanes, wavefront kernel
constexpr uint32_t kItemsInLds = 512;   // item starts the 8-lane kernel keeps in LDS

// The last index i with starts[i] <= key (starts ascending, starts[0] <= key).
__device__ __forceinline__ uint32_t find_start(const uint32_t* starts, uint32_t n, uint32_t key) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (starts[mid] <= key) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t defer_class(uint32_t deg, uint32_t k) {
    if (k > kGroup || deg > 64) return 3;
    return deg <= 16 ? 0 : deg <= 32 ? 1 : 2;
}

template <uint32_t G>
__device__ __forceinline__ unsigned long long group_bits(unsigned long long bal, uint32_t gbase) {
    if constexpr (G == 64) return bal;
    else return (bal >> gbase) & ((1ull << G) - 1ull);
}

// The k-bit left mask of a membership word for a G-lane group: lane i < k of
// the group holds left vertex i's bit position (my_shift), read by shuffles
// within the group. Every lane of the wavefront calls it.
template <uint32_t G>
__device__ __forceinline__ uint32_t group_mask(unsigned long long mw, uint32_t my_shift, uint32_t k) {
    uint32_t m = 0;
#pragma unroll
    for (uint32_t i = 0; i < kGroup; ++i) {
        const uint32_t sh = __shfl(my_shift, i, G);
        if (i < k) m |= (uint32_t)((mw >> sh) & 1ull) << i;
    }
    return m;
}

// The neighbours of v that can be right vertices of item it are the label
// runs of its left labels in adj_bl, taken one after the other (at most
// kGroup labels here: callers use the positions only for k <= kGroup, and
// only the length otherwise). Every lane of a G-lane group gets their total
// length; lane lig also gets the adj_bl position of element lig (when lig
// is below the total). Every lane of the wavefront calls it.
template <uint32_t G>
__device__ __forceinline__ uint32_t group_relevant(const DevGraph& g, const GItem& it,
                                                   const uint32_t* __restrict__ left_labs, uint32_t v,
                                                   bool live, uint32_t lig, uint32_t& pos) {
    uint2 run = make_uint2(0u, 0u);
    uint32_t r0 = 0, R = 0;
    if (live) {
        r0 = g.nlf_off[v];
        R = g.nlf_off[v + 1] - r0;
    }
    const uint32_t want = live && lig < it.nlab && lig < kGroup ? left_labs[it.lab_begin + lig] : kInvalid;
    if (R <= G) {
        // v has at most G label runs: the group's lanes load one each in a
        // single round, and every lane finds its label by a binary search
        // over the lanes (R is the same on all lanes of a group)
        const uint2 e = lig < R ? g.nlf[r0 + lig] : make_uint2(kInvalid, 0u);
        const uint32_t eb = lig < R ? g.nlf_beg[r0 + lig] : 0u;
        uint32_t lo = 0, len = R;
        for (uint32_t round = G; round; round >>= 1) {   // len <= G halves every round
            const uint32_t half = len >> 1;
            const uint32_t lab = __shfl(e.x, min(lo + half, G - 1), G);
            if (len) {
                if (lab < want) {
                    lo += half + 1;
                    len -= half + 1;
                } else {
                    len = half;
                }
            }
        }
        const uint32_t at = min(lo, G - 1);
        const uint32_t lab = __shfl(e.x, at, G), cnt = __shfl(e.y, at, G), beg = __shfl(eb, at, G);
        if (want != kInvalid && lo < R && lab == want) run = make_uint2(beg, cnt);
    } else if (want != kInvalid) {
        run = label_run(g, v, want);
    }
    uint32_t incl = run.y;   // inclusive prefix of run lengths over the group's lanes
#pragma unroll
    for (uint32_t o = 1; o < kGroup; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o, G);
        if (lig >= o) incl += t;
    }
    const uint32_t total = __shfl(incl, kGroup - 1, G);
    pos = 0;
#pragma unroll
    for (uint32_t r = 0; r < kGroup; ++r) {
        const uint32_t e = __shfl(incl, r, G), b = __shfl(run.x, r, G), l = __shfl(run.y, r, G);
        if (lig + l >= e && lig < e) pos = b + (lig + l - e);
    }
    return total;
}

// Hall's condition for one candidate tested by a group of G lanes: lane lig
// holds right vertex lig's k-bit left mask m (0 past deg); R[i] = the group's
// lanes adjacent to left vertex i; the 2^k - 1 left subsets are split over
// the group. True when some subset S has fewer than |S| right neighbours
// (no matching saturates the left side). Every lane of the wavefront calls it.
template <uint32_t G>
__device__ __forceinline__ bool hall_fails(uint32_t m, bool live, uint32_t k, uint32_t lig, uint32_t gbase) {
    unsigned long long R[kGroup];
#pragma unroll
    for (uint32_t i = 0; i < kGroup; ++i) R[i] = group_bits<G>(__ballot((m >> i) & 1u), gbase);
    bool fail = false;
    if (live) {
        const uint32_t full = (1u << k) - 1u;
        for (uint32_t sub = lig + 1; sub <= full && !fail; sub += G) {
            unsigned long long uni = 0;
#pragma unroll
            for (uint32_t i = 0; i < kGroup; ++i)
                if ((sub >> i) & 1u) uni |= R[i];
            fail = __popcll(uni) < __popc(sub);
        }
    }
    return group_bits<G>(__ballot(fail), gbase) != 0;
}

// Removes candidate j of item it (value v): pool slot, bitmap bit, membership bit.
__device__ __forceinline__ void gql_drop(const GItem& it, uint32_t j, uint32_t v, uint32_t* __restrict__ cand_pool,
                                         uint32_t* __restrict__ bitmaps, uint64_t wps,
                                         unsigned long long* __restrict__ member, uint32_t* __restrict__ removed) {
    cand_pool[it.cand_begin + j] = kInvalid;
    atomicAnd(bitmaps + (size_t)it.bm_self * wps + (v >> 5), ~(1u << (v & 31)));
    atomicAnd(member + it.mem_base + v, ~(1ull << it.u));
    removed[it.bm_self] = 1u;
}

// Second-round skip flag of each item (GqlRound): the item's candidates of
// u were tested in the first round against the sets of u's query neighbours
// as they stood then; since, those of later neighbours (u' > u) changed only
// in the first round, those of earlier ones (u' < u) only in the second. If
// none changed, the second test would see the same bipartite graphs.
__global__ __launch_bounds__(kBlock) void k_gql_skip(const GItem* __restrict__ items, uint32_t n_items,
                                                     const uint32_t* __restrict__ left_slots,
                                                     const uint32_t* __restrict__ removed1,
                                                     const uint32_t* __restrict__ removed2,
                                                     uint32_t* __restrict__ skip) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_items) return;
    const GItem it = items[i];
    const uint32_t first = it.bm_self - it.u;
    uint32_t changed = 0;
    for (uint32_t e = 0; e < it.k; ++e) {
        const uint32_t slot = left_slots[it.left_begin + e];
        changed |= slot - first > it.u ? removed1[slot] : removed2[slot];
    }
    skip[i] = changed ? 0u : 1u;
}

// The Hall test with G lanes per candidate over the (item, index) pairs the
// 8-lane kernel deferred to this class (k <= 8, deg(v) <= G). The pair count
// is read from device memory; a fixed grid strides over the list, so the
// launch needs no count from the host.
template <uint32_t G>
__global__ __launch_bounds__(kBlock) void k_gql_hall(
    DevGraph g, const GItem* __restrict__ items, const uint2* __restrict__ list,
    const uint32_t* __restrict__ n_list, const uint32_t* __restrict__ left_slots,
    const uint32_t* __restrict__ left_labs, uint32_t* __restrict__ cand_pool, uint32_t* __restrict__ bitmaps,
    uint64_t wps, unsigned long long* __restrict__ member, GqlRound rd, unsigned long long* __restrict__ bytes) {
    constexpr uint32_t kPerIter = kBlock / G;
    const uint32_t n = *n_list;
    const uint32_t lane = lane_id();
    const uint32_t lig = lane & (G - 1);
    const uint32_t gbase = lane & ~(G - 1);
    unsigned long long by = 0;
    // block-uniform trip count: every lane takes part in the ballots
    for (uint64_t first = (uint64_t)blockIdx.x * kPerIter; first < n; first += (uint64_t)gridDim.x * kPerIter) {
        const uint64_t gid = first + threadIdx.x / G;
        bool live = gid < n;
        uint32_t j = 0, v = kInvalid, k = 0, deg = 0, my_shift = 0;
        GItem it{};
        if (live) {
            const uint2 e = list[gid];
            it = items[e.x];
            j = e.y;
            v = cand_pool[it.cand_begin + j];
            // lane lig < k: left vertex lig's bit in a membership word
            if (lig < min(it.k, kGroup)) my_shift = left_slots[it.left_begin + lig] - (it.bm_self - it.u);
            live = v != kInvalid;
        }
        uint32_t pos = 0;
        // relevant degree <= G by construction of the list
        deg = min(group_relevant<G>(g, it, left_labs, v, live, lig, pos), G);
        if (live) k = min(it.k, kGroup);
        const bool has = live && lig < deg;
        const unsigned long long mw = has ? member[it.mem_base + g.adj_bl[pos]] : 0ull;
        const uint32_t m0 = group_mask<G>(mw, my_shift, k);   // every lane shuffles
        const uint32_t m = has ? m0 : 0u;
        const bool fail = hall_fails<G>(m, live, k, lig, gbase);
        if (live && lig == 0) {
            if (fail) gql_drop(it, j, v, cand_pool, bitmaps, wps, member, rd.removed);
            // algorithmic bytes: candidate, one run lookup per left label,
            // the runs, one membership word per neighbour and left vertex
            by += 4ull + 16ull * it.nlab + 4ull * deg * (1 + k);
        }
    }
    block_add(bytes, by);
}

// Deferred pairs are staged per block in LDS and appended to the global class
// lists with one atomic per class per flush (a per-wavefront atomic on the
// four list counters serialises the whole launch on four addresses).
constexpr uint32_t kDeferBuf = 512;

struct DeferStage {
    uint2 buf[kDeferClasses][kDeferBuf];
    uint32_t count[kDeferClasses];
    uint32_t base[kDeferClasses];
};

// Appends the staged pairs of every class holding more than `threshold` to
// lists + c * stride. Every thread of the block calls it after a barrier.
__device__ __forceinline__ void defer_flush(DeferStage& st, uint2* __restrict__ lists, uint64_t stride,
                                            uint32_t* __restrict__ n_lists, uint32_t threshold) {
    if (threadIdx.x < kDeferClasses) {
        const uint32_t c = threadIdx.x, m = st.count[c];
        st.base[c] = (m > threshold && m > 0) ? atomicAdd(n_lists + c, m) : kInvalid;
    }
    __syncthreads();
    for (uint32_t c = 0; c < kDeferClasses; ++c) {
        const uint32_t b = st.base[c];
        if (b == kInvalid) continue;
        const uint32_t m = st.count[c];
        for (uint32_t i = threadIdx.x; i < m; i += kBlock) lists[(uint64_t)c * stride + b + i] = st.buf[c][i];
    }
    __syncthreads();
    if (threadIdx.x < kDeferClasses && st.base[threadIdx.x] != kInvalid) st.count[threadIdx.x] = 0;
    __syncthreads();
}

__global__ __launch_bounds__(kBlock) void k_gql_refine_small(
    DevGraph g, const GItem* __restrict__ items, uint32_t n_items, uint32_t n_cands,
    const uint32_t* __restrict__ left_slots, const uint32_t* __restrict__ left_labs,
    uint32_t* __restrict__ cand_pool, uint32_t* __restrict__ bitmaps, uint64_t wps,
    unsigned long long* __restrict__ member, uint2* __restrict__ lists, uint64_t list_stride,
    uint32_t* __restrict__ n_lists, GqlRound rd,
    unsigned long long* __restrict__ bytes) {
    constexpr uint32_t kPerIter = kBlock / kGroup;   // candidates per block iteration
    __shared__ DeferStage st;
    __shared__ uint32_t s_start[kItemsInLds];   // the items' first positions, searched in LDS
    const uint32_t lane = lane_id();
    const uint32_t lig = lane & (kGroup - 1);
    const uint32_t gbase = lane & ~(kGroup - 1);
    const bool starts_in_lds = n_items <= kItemsInLds;
    if (threadIdx.x < kDeferClasses) st.count[threadIdx.x] = 0;
    if (starts_in_lds)
        for (uint32_t i = threadIdx.x; i < n_items; i += kBlock) s_start[i] = items[i].wave_start;
    __syncthreads();
    unsigned long long by = 0;
    // block-uniform trip count: barriers and ballots see every lane
    for (uint64_t first = (uint64_t)blockIdx.x * kPerIter; first < n_cands;
         first += (uint64_t)gridDim.x * kPerIter) {
        const uint64_t gid = first + threadIdx.x / kGroup;
        bool live = gid < n_cands;
        uint32_t ii = 0, j = 0, v = kInvalid, k = 0;
        GItem it{};
        if (live) {
            ii = starts_in_lds ? find_start(s_start, n_items, (uint32_t)gid)
                               : find_item(items, n_items, (uint32_t)gid, &GItem::wave_start);
            it = items[ii];
            j = (uint32_t)gid - it.wave_start;
            live = j < it.len && !(rd.skip && rd.skip[ii]);
        }
        uint32_t my_shift = 0;
        if (live) {
            v = cand_pool[it.cand_begin + j];
            // lane lig < k: left vertex lig's bit in a membership word
            if (lig < min(it.k, kGroup)) my_shift = left_slots[it.left_begin + lig] - (it.bm_self - it.u);
            live = v != kInvalid;
        }
        uint32_t pos = 0;
        const uint32_t deg = group_relevant<kGroup>(g, it, left_labs, v, live, lig, pos);
        if (live) k = it.k;
        // defer the candidates outside the limits into the block's stage
        const bool defer = live && lig == 0 && (deg > kGroup || k > kGroup);
        const uint32_t dc = defer ? defer_class(deg, k) : kDeferClasses;
#pragma unroll
        for (uint32_t c = 0; c < kDeferClasses; ++c) {
            const unsigned long long bal = __ballot(dc == c);
            if (!bal) continue;
            const uint32_t leader = __ffsll((long long)bal) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&st.count[c], (uint32_t)__popcll(bal));
            base = __shfl(base, leader);
            if (dc == c) st.buf[c][base + __popcll(bal & lanemask_lt())] = make_uint2(ii, j);
        }
        if (live && (deg > kGroup || k > kGroup)) live = false;
        const bool has = live && lig < deg;
        const unsigned long long mw = has ? member[it.mem_base + g.adj_bl[pos]] : 0ull;
        const uint32_t m0 = group_mask<kGroup>(mw, my_shift, min(k, kGroup));
        const uint32_t m = has ? m0 : 0u;
        const bool fail = hall_fails<kGroup>(m, live, k, lig, gbase);
        if (live && lig == 0) {
            if (fail) gql_drop(it, j, v, cand_pool, bitmaps, wps, member, rd.removed);
            by += 4ull + 16ull * it.nlab + 4ull * deg * (1 + k);
        }
        __syncthreads();
        // at most kPerIter pairs per class join the stage per iteration
        defer_flush(st, lists, list_stride, n_lists, kDeferBuf - kPerIter);
    }
    defer_flush(st, lists, list_stride, n_lists, 0);
    block_add(bytes, by);
}

// ---------------------------------------------------------------------------
// BuildTable (BuildTable.cpp:9-151) by definition: row j of table (a, b) lists
// the ranks in C(b) of the neighbours of C(a)[j] that are in C(b). The rank
// of w is the count of set bits before it in the (query, b) bitmap. Only
// neighbours with b's label can be in C(b): a row scans that label's run of
// the label-grouped adjacency (ids ascending inside it, so ranks come out
// ascending), not all of N(v).
// ---------------------------------------------------------------------------
constexpr uint32_t kRowCoop = 64;   // rows whose label run is longer go to a wavefront

__device__ __forceinline__ uint32_t bit_rank(const uint32_t* bm, const uint32_t* wp, uint32_t w) {
    const uint32_t word = w >> 5;
    return (wp[word] - wp[0]) + __popc(bm[word] & ((1u << (w & 31)) - 1u));
}

__global__ __launch_bounds__(kBlock) void k_table_count(
    DevGraph g, const TItem* __restrict__ items, uint32_t n_items,
    const uint64_t* __restrict__ item_row_start, uint64_t n_rows,
    const uint32_t* __restrict__ cand_pool, const uint32_t* __restrict__ bitmaps, uint64_t wps,
    uint64_t* __restrict__ row_deg, uint64_t* __restrict__ long_rows, unsigned long long* __restrict__ n_long,
    unsigned long long* __restrict__ bytes) {
    unsigned long long by = 0;
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    const uint64_t rows_rounded = (n_rows + stride - 1) / stride * stride;
    for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < rows_rounded; r += stride) {
        bool defer = false;
        if (r < n_rows) {
            const uint32_t ii = find_u64(item_row_start, n_items, r);
            const TItem it = items[ii];
            const uint32_t v = cand_pool[it.cand_a + (r - item_row_start[ii])];
            const uint2 run = label_run(g, v, it.lab_b);
            if (run.y > kRowCoop) {
                defer = true;   // a wavefront scans it (k_table_count_long)
            } else {
                const uint32_t* bm = bitmaps + (size_t)it.bm_b * wps;
                uint32_t c = 0;
                for (uint32_t p = run.x; p < run.x + run.y; ++p) {
                    const uint32_t w = g.adj_bl[p];
                    c += (bm[w >> 5] >> (w & 31)) & 1u;
                }
                row_deg[r] = c;
                // candidate + run lookup + the run + one bitmap word per
                // neighbour in it + the count
                by += 4 + 16 + 8ull * run.y + 8;
            }
        }
        const unsigned long long bal = __ballot(defer);
        if (bal) {
            const uint32_t leader = __ffsll((long long)bal) - 1;
            unsigned long long base = 0;
            if (lane_id() == leader) base = atomicAdd(n_long, (unsigned long long)__popcll(bal));
            base = __shfl(base, leader);
            if (defer) long_rows[base + __popcll(bal & lanemask_lt())] = r;
        }
    }
    wave_add(bytes, by);
}

// Rows of candidates with more than kRowCoop neighbours, one wavefront each:
// lanes read 64 consecutive neighbours at a time.
__global__ __launch_bounds__(kBlock) void k_table_count_long(
    DevGraph g, const TItem* __restrict__ items, uint32_t n_items,
    const uint64_t* __restrict__ item_row_start, const uint64_t* __restrict__ long_rows, uint64_t n_long,
    const uint32_t* __restrict__ cand_pool, const uint32_t* __restrict__ bitmaps, uint64_t wps,
    uint64_t* __restrict__ row_deg, unsigned long long* __restrict__ bytes) {
    const uint32_t lane = lane_id();
    const uint64_t waves = (uint64_t)gridDim.x * (kBlock / kWave);
    unsigned long long by = 0;
    for (uint64_t i = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) / kWave; i < n_long; i += waves) {
        const uint64_t r = long_rows[i];
        const uint32_t ii = find_u64(item_row_start, n_items, r);
        const TItem it = items[ii];
        const uint32_t v = cand_pool[it.cand_a + (r - item_row_start[ii])];
        const uint2 run = label_run(g, v, it.lab_b);
        const uint32_t vb = run.x, ve = run.x + run.y;
        const uint32_t* bm = bitmaps + (size_t)it.bm_b * wps;
        uint64_t c = 0;
        for (uint32_t base = vb; base < ve; base += kWave) {
            const uint32_t p = base + lane;
            bool hit = false;
            if (p < ve) {
                const uint32_t w = g.adj_bl[p];
                hit = (bm[w >> 5] >> (w & 31)) & 1u;
            }
            c += __popcll(__ballot(hit));
        }
        if (lane == 0) {
            row_deg[r] = c;
            by += 4 + 16 + 8ull * run.y + 8;
        }
    }
    wave_add(bytes, by);
}


__global__ __launch_bounds__(kBlock) void k_table_fill(
    DevGraph g, const TItem* __restrict__ items, uint32_t n_items,
    const uint64_t* __restrict__ item_row_start, uint64_t n_rows,
    const uint32_t* __restrict__ cand_pool, const uint32_t* __restrict__ bitmaps,
    const uint32_t* __restrict__ word_prefix, uint64_t wps, const uint64_t* __restrict__ row_off,
    uint32_t* __restrict__ idx, unsigned long long* __restrict__ bytes) {
    unsigned long long by = 0;
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    const uint64_t rows_rounded = (n_rows + stride - 1) / stride * stride;
    for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < rows_rounded; r += stride) {
        if (r >= n_rows) continue;
        const uint32_t ii = find_u64(item_row_start, n_items, r);
        const TItem it = items[ii];
        const uint32_t v = cand_pool[it.cand_a + (r - item_row_start[ii])];
        const uint2 run = label_run(g, v, it.lab_b);
        if (run.y > kRowCoop) continue;   // k_table_fill_long
        const uint32_t* bm = bitmaps + (size_t)it.bm_b * wps;
        const uint32_t* wp = word_prefix + (size_t)it.bm_b * wps;
        const uint64_t o0 = row_off[r];
        uint64_t o = o0;
        for (uint32_t p = run.x; p < run.x + run.y; ++p) {
            const uint32_t w = g.adj_bl[p];
            if ((bm[w >> 5] >> (w & 31)) & 1u) idx[o++] = bit_rank(bm, wp, w);
        }
        // candidate + run lookup + row offset + the run + bitmap word per
        // neighbour in it + per entry: rank words (2) and the entry written
        by += 4 + 16 + 8 + 8ull * run.y + 12ull * (o - o0);
    }
    wave_add(bytes, by);
}

// Long rows of the fill pass, one wavefront each: every 64-neighbour chunk is
// written in order (ballot prefix), so the row comes out ascending.
__global__ __launch_bounds__(kBlock) void k_table_fill_long(
    DevGraph g, const TItem* __restrict__ items, uint32_t n_items,
    const uint64_t* __restrict__ item_row_start, const uint64_t* __restrict__ long_rows, uint64_t n_long,
    const uint32_t* __restrict__ cand_pool, const uint32_t* __restrict__ bitmaps,
    const uint32_t* __restrict__ word_prefix, uint64_t wps, const uint64_t* __restrict__ row_off,
    uint32_t* __restrict__ idx, unsigned long long* __restrict__ bytes) {
    const uint32_t lane = lane_id();
    const uint64_t waves = (uint64_t)gridDim.x * (kBlock / kWave);
    unsigned long long by = 0;
    for (uint64_t i = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) / kWave; i < n_long; i += waves) {
        const uint64_t r = long_rows[i];
        const uint32_t ii = find_u64(item_row_start, n_items, r);
        const TItem it = items[ii];
        const uint32_t v = cand_pool[it.cand_a + (r - item_row_start[ii])];
        const uint2 run = label_run(g, v, it.lab_b);
        const uint32_t vb = run.x, ve = run.x + run.y;
        const uint32_t* bm = bitmaps + (size_t)it.bm_b * wps;
        const uint32_t* wp = word_prefix + (size_t)it.bm_b * wps;
        const uint64_t o0 = row_off[r];
        uint64_t o = o0;
        for (uint32_t base = vb; base < ve; base += kWave) {
            const uint32_t p = base + lane;
            bool hit = false;
            uint32_t w = 0;
            if (p < ve) {
                w = g.adj_bl[p];
                hit = (bm[w >> 5] >> (w & 31)) & 1u;
            }
            const unsigned long long bal = __ballot(hit);
            if (hit) idx[o + __popcll(bal & lanemask_lt())] = bit_rank(bm, wp, w);
            o += __popcll(bal);
        }
        if (lane == 0) by += 4 + 16 + 8 + 8ull * run.y + 12ull * (o - o0);
    }
    wave_add(bytes, by);
}

// ---------------------------------------------------------------------------
// CFL support counting (FilterVertices.cpp:841-966 restated with set
// semantics): cnt[w] advances from `expect` to `next` once per pivot for a
// data vertex w with the query vertex's label and degree that neighbours a
// live candidate of that pivot. A pivot candidate's scan walks only the run
// of that label in the label-grouped adjacency.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void cfl_mark(const DevGraph& g, uint32_t w, uint32_t label, uint32_t degree,
                                         uint32_t* __restrict__ cnt, uint32_t expect, uint32_t next) {
    if (g.lab[w] != label || g.off[w + 1] - g.off[w] < degree) return;
    if (expect == 0) {
        // first pivot: any stale value below `next` becomes `next`
        atomicMax(cnt + w, next);
    } else {
        atomicCAS(cnt + w, expect, next);
    }
}

// Thread per pivot candidate; candidates with more than kRowCoop neighbours
// of the label are listed for k_cfl_support_long (one atomic per wavefront).
__global__ __launch_bounds__(kBlock) void k_cfl_support(
    DevGraph g, const uint32_t* __restrict__ cand_pool, uint64_t cand_begin, uint32_t len,
    uint32_t label, uint32_t degree, uint32_t* __restrict__ cnt, uint32_t expect, uint32_t next,
    uint32_t* __restrict__ defer, uint32_t* __restrict__ n_defer) {
    const uint32_t stride = gridDim.x * kBlock;
    const uint32_t rounded = (len + stride - 1) / stride * stride;
    for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < rounded; j += stride) {
        uint32_t v = kInvalid;
        bool hub = false;
        if (j < len) {
            v = cand_pool[cand_begin + j];
            if (v != kInvalid) {
                const uint2 run = label_run(g, v, label);
                hub = run.y > kRowCoop;
                if (!hub)
                    for (uint32_t p = run.x; p < run.x + run.y; ++p)
                        cfl_mark(g, g.adj_bl[p], label, degree, cnt, expect, next);
            }
        }
        const unsigned long long bal = __ballot(hub);
        if (bal) {
            const uint32_t leader = __ffsll((long long)bal) - 1;
            uint32_t base = 0;
            if (lane_id() == leader) base = atomicAdd(n_defer, (uint32_t)__popcll(bal));
            base = __shfl(base, leader);
            if (hub) defer[base + __popcll(bal & lanemask_lt())] = v;
        }
    }
}

// The listed pivot candidates, one wavefront each over 64 neighbours at a time.
__global__ __launch_bounds__(kBlock) void k_cfl_support_long(
    DevGraph g, const uint32_t* __restrict__ defer, const uint32_t* __restrict__ n_defer, uint32_t label,
    uint32_t degree, uint32_t* __restrict__ cnt, uint32_t expect, uint32_t next) {
    const uint32_t lane = lane_id();
    const uint32_t n = *n_defer;
    const uint32_t waves = gridDim.x * (kBlock / kWave);
    for (uint32_t i = (blockIdx.x * kBlock + threadIdx.x) / kWave; i < n; i += waves) {
        const uint32_t v = defer[i];
        const uint2 run = label_run(g, v, label);
        for (uint32_t p = run.x + lane; p < run.x + run.y; p += kWave)
            cfl_mark(g, g.adj_bl[p], label, degree, cnt, expect, next);
    }
}

__global__ __launch_bounds__(kBlock) void k_cfl_prune(const uint32_t* __restrict__ cnt,
                                                       uint32_t* __restrict__ cand_pool,
                                                       uint64_t cand_begin, uint32_t len,
                                                       uint32_t target) {
    for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < len; j += gridDim.x * kBlock) {
        const uint32_t v = cand_pool[cand_begin + j];
        if (v != kInvalid && cnt[v] != target) cand_pool[cand_begin + j] = kInvalid;
    }
}

// ---------------------------------------------------------------------------
// Enumeration. A partial embedding at depth d assigns candidate index I[e]
// (and data vertex M[e]) to query vertex order[e] for e < d. The valid
// extensions at depth d are the candidate indices present in every operand
// row E[b -> order[d]][I[src(b)]] whose data vertex is not already used.
// They are produced lazily: the shortest operand row is walked, every
// element is binary-searched in the other rows and checked against the
// earlier same-label assignments. Call counting follows the GQL / EXPLORE
// engines (one call per assignment, EvaluateQuery.cpp:1009,1081); the
// extensions present in every operand row whose data vertex an earlier depth
// holds are counted apart (QCounters::rejects): the LFTJ engine also counts
// those (EvaluateQuery.cpp:278-295 takes the visited test after the call).
// ---------------------------------------------------------------------------
struct LevelBase {
    uint64_t lo, hi;
    uint32_t sel;
};

template <typename State>
__device__ __forceinline__ LevelBase level_base(const DPlan& P, const DOp* __restrict__ ops,
                                                const uint64_t* __restrict__ row_off, uint32_t d,
                                                const State& st, unsigned long long& bytes,
                                                unsigned long long& rows) {
    LevelBase b;
    const uint32_t nops = P.op_count[d];
    if (nops == 0) {
        b.lo = 0;
        b.hi = P.cand_len[d];
        b.sel = kIdentity;
        return b;
    }
    b.sel = 0;
    b.lo = 0;
    b.hi = ~0ull;
    uint64_t best = ~0ull;
    for (uint32_t j = 0; j < nops; ++j) {
        const DOp op = ops[P.op_begin[d] + j];
        const uint64_t r = op.row_base + st.I(op.src_depth);
        const uint64_t lo = row_off[r], hi = row_off[r + 1];
        rows += hi - lo;
        if (hi - lo < best) { best = hi - lo; b.lo = lo; b.hi = hi; b.sel = j; }
    }
    bytes += 16ull * nops;
    return b;
}

// 0: x is not in every operand row; 1: x is a valid extension (v_out set);
// 3: x is in every operand row but its data vertex is already used.
template <typename State>
__device__ __forceinline__ int test_element(const DPlan& P, const DOp* __restrict__ ops,
                                             const uint64_t* __restrict__ row_off,
                                             const uint32_t* __restrict__ idx,
                                             const uint32_t* __restrict__ cand_pool, uint32_t d,
                                             uint32_t x, uint32_t sel, const State& st,
                                             uint32_t& v_out, unsigned long long& bytes) {
    const uint32_t nops = P.op_count[d];
    for (uint32_t j = 0; j < nops; ++j) {
        if (j == sel) continue;
        const DOp op = ops[P.op_begin[d] + j];
        const uint64_t r = op.row_base + st.I(op.src_depth);
        const uint64_t hi0 = row_off[r + 1];
        uint64_t lo = row_off[r], hi = hi0;
        bytes += 16;
        // lower_bound of x in idx[lo, hi)
        while (lo < hi) {
            const uint64_t mid = (lo + hi) >> 1;
            bytes += 4;
            if (idx[mid] < x) lo = mid + 1;
            else hi = mid;
        }
        if (lo == hi0 || idx[lo] != x) return 0;
    }
    unsigned long long mask = P.same_label[d];
    if (!mask) {
        v_out = kInvalid;   // no earlier vertex shares the label: injective by construction
        return 1;
    }
    const uint32_t v = cand_pool[P.cand_base[d] + x];
    bytes += 4;
    while (mask) {
        const uint32_t e = __ffsll((long long)mask) - 1;
        mask &= mask - 1;
        if (st.M(e) == v) return 3;
    }
    v_out = v;
    return 1;
}

// Per-lane state of the persistent DFS, in LDS: per level the assigned
// candidate index I and data vertex M, and the level's cursor and end as
// offsets from the start of its base row (restored by re-reading the row
// start on backtrack), plus the base operand. Laid out [field][level][lane
// of the block], so the lanes of a wavefront hit distinct banks whatever
// level each is at. 17 bytes per level per lane.
constexpr uint32_t kLaneStateBytes = 17;
constexpr uint8_t kSelIdentity = 0xFF;

struct LaneState {
    uint32_t* I_;
    uint32_t* M_;
    uint32_t* cur_;
    uint32_t* end_;
    uint8_t* sel_;
    uint32_t stride;   // lanes per block
    uint32_t lane;     // this lane's index in the block
    __device__ __forceinline__ uint32_t at(uint32_t lvl) const { return lvl * stride + lane; }
    __device__ __forceinline__ uint32_t I(uint32_t e) const { return I_[at(e)]; }
    __device__ __forceinline__ uint32_t M(uint32_t e) const { return M_[at(e)]; }
};

// Task pre-expansion state: the prefix held in the task record.
struct PrefixState {
    const uint32_t* I_;
    const uint32_t* M_;
    __device__ __forceinline__ uint32_t I(uint32_t e) const { return I_[e]; }
    __device__ __forceinline__ uint32_t M(uint32_t e) const { return M_[e]; }
};

constexpr uint64_t kExpandChunk = 256;   // longest base range one expansion thread walks

// Adds v to counter[q] for every lane of the wavefront (all lanes call; lanes
// with q == kInvalid add nothing). Task arrays are sorted by query, so the
// lanes of a wavefront usually share q: their sum then goes in one atomic.
__device__ __forceinline__ void add_by_query(unsigned long long* counter, size_t stride, uint32_t q,
                                             unsigned long long v) {
    const unsigned long long valid = __ballot(q != kInvalid);
    if (!valid) return;
    const uint32_t lead = __ffsll((long long)valid) - 1;
    const uint32_t ql = __shfl(q, lead);
    if (__ballot(q != kInvalid && q != ql) == 0) {
        unsigned long long sum = q != kInvalid ? v : 0ull;
        for (int o = kWave / 2; o > 0; o >>= 1) {
            const uint32_t lo = __shfl_xor((uint32_t)sum, o), hi = __shfl_xor((uint32_t)(sum >> 32), o);
            sum += ((unsigned long long)hi << 32) | lo;
        }
        if (lane_id() == lead && sum) atomicAdd(counter + (size_t)ql * stride, sum);
    } else if (q != kInvalid && v) {
        atomicAdd(counter + (size_t)q * stride, v);
    }
}

// One thread per task of the queries being expanded (the caller keeps the
// tasks of the other queries apart): enumerate the valid extensions at the
// task's depth. For a query with expand_q[q] == 2 (the caller sets it while
// the query has a single task), a level with exactly one valid extension is
// assigned at once and the thread goes on to the next level, so a chain of
// single candidates at the top of a search is crossed in one round; it stops
// at a level with none or several, or at the depth limit min(n - 1,
// kExpandMax) (the leaf level is never assigned here). Otherwise every task
// goes one level deeper, so a query's tasks stay at one depth and of
// comparable size. Count pass (write = 0): out_count[t] = the tasks that replace task
// t (its children, or t itself when it is at the limit), qchild[q] += them,
// qchild[nq + q] += 1 if t was expanded. Write pass: those tasks at
// out_pos[t]; calls += one per assignment made, once the expansion is
// committed. A task of a query that is not expanded (expand_q[q] == 0) is
// dropped.
__global__ __launch_bounds__(kBlock) void k_expand(
    const DPlan* __restrict__ plans, const DOp* __restrict__ ops, const DTask* __restrict__ tasks, uint32_t n_tasks,
    const uint8_t* __restrict__ expand_q, uint32_t* __restrict__ out_count,
    const uint64_t* __restrict__ out_pos, DTask* __restrict__ out_tasks, QCounters* qc,
    unsigned long long* __restrict__ qchild, uint32_t nq, uint32_t max_depth, int write) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    // every lane reaches the per-query sums at the end
    uint32_t q = kInvalid, out = 0, rej = 0, calls = 0, expanded = 0;
    unsigned long long bytes = 0, algo = 0;
    const uint32_t mode = t < n_tasks ? expand_q[tasks[t].q] : 0u;
    if (mode) {
        const DTask& task = tasks[t];
        q = task.q;
        const DPlan& P = plans[q];
        const uint64_t* __restrict__ row_off = P.row_off;
        const uint32_t* __restrict__ idx = P.idx;
        const uint32_t* __restrict__ cand_pool = P.cand_pool;
        const uint32_t lim = min(P.n - 1, min(max_depth, kExpandMax));
        uint32_t I[kExpandMax + 1], M[kExpandMax + 1];
        uint32_t d = task.depth;
        for (uint32_t e = 0; e < d; ++e) {
            I[e] = task.prefix[e];
            M[e] = cand_pool[P.cand_base[e] + I[e]];
        }
        PrefixState st{I, M};
        uint64_t pos = write ? out_pos[t] : 0;
        LevelBase b;
        bool fresh = task.hi == kAutoHi;
        if (!fresh) {
            b.lo = task.lo;
            b.hi = task.hi;
            b.sel = task.sel;
        }
        if (d >= lim) {
            // at the limit: the task stays as it is
            out = 1;
            if (write) out_tasks[pos] = task;
        } else {
            expanded = 1;
            while (true) {
                unsigned long long rows = 0;
                if (fresh) b = level_base(P, ops, row_off, d, st, bytes, rows);   // a node entered here
                if (b.hi - b.lo > kExpandChunk) {
                    // a long base range (a hub's row) is not walked by one
                    // thread: it is cut into range tasks of at most
                    // kExpandChunk positions at the same depth (no assignment
                    // is made), expanded next round
                    const uint64_t pieces = (b.hi - b.lo + kExpandChunk - 1) / kExpandChunk;
                    if (write) {
                        for (uint64_t i = 0; i < pieces; ++i) {
                            DTask ch = task;
                            ch.depth = d;
                            for (uint32_t e = 0; e < d; ++e) ch.prefix[e] = I[e];
                            ch.sel = b.sel;
                            ch.lo = b.lo + i * kExpandChunk;
                            ch.hi = min(b.hi, ch.lo + kExpandChunk);
                            out_tasks[pos++] = ch;
                        }
                    }
                    out = (uint32_t)pieces;
                    algo += 4ull * rows;
                    break;
                }
                // the valid extensions of this level: counted, and written
                // unless the level turns out to have exactly one
                uint32_t c = 0, first = 0;
                for (uint64_t cur = b.lo; cur < b.hi; ++cur) {
                    const uint32_t x = b.sel == kIdentity ? (uint32_t)cur : idx[cur];
                    uint32_t v;
                    const int r = test_element(P, ops, row_off, idx, cand_pool, d, x, b.sel, st, v, bytes);
                    if (r != 1) {
                        rej += r == 3;
                        continue;
                    }
                    if (c == 0) first = x;
                    ++c;
                }
                // each valid extension is one assignment at this depth (a
                // call); those below the root level are written survivors
                calls += c;
                algo += 4ull * rows + (d >= 1 ? 4ull * c : 0ull);
                if (c == 1 && d + 1 < lim && mode == 2) {
                    // a single extension: assign it and go on
                    I[d] = first;
                    M[d] = cand_pool[P.cand_base[d] + first];
                    ++d;
                    fresh = true;
                    continue;
                }
                if (write && c) {
                    for (uint64_t cur = b.lo; cur < b.hi; ++cur) {
                        const uint32_t x = b.sel == kIdentity ? (uint32_t)cur : idx[cur];
                        uint32_t v;
                        unsigned long long scratch_bytes = 0;
                        if (test_element(P, ops, row_off, idx, cand_pool, d, x, b.sel, st, v, scratch_bytes) != 1)
                            continue;
                        DTask ch;
                        ch.q = q;
                        ch.depth = d + 1;
                        ch.sel = 0;
                        ch.pad = 0;
                        ch.lo = 0;
                        ch.hi = kAutoHi;
                        for (uint32_t e = 0; e < kPrefixMax; ++e) ch.prefix[e] = e < d ? I[e] : 0;
                        ch.prefix[d] = x;
                        out_tasks[pos++] = ch;
                    }
                }
                out = c;
                break;
            }
        }
    }
    if (!write) {
        if (t < n_tasks) out_count[t] = out;
        add_by_query(qchild, 1, q, out);
        add_by_query(qchild + nq, 1, q, expanded);
        return;
    }
    static_assert(sizeof(QCounters) % sizeof(unsigned long long) == 0, "QCounters is an array of u64 strides");
    const size_t stride = sizeof(QCounters) / sizeof(unsigned long long);
    add_by_query(&qc->calls, stride, q, calls);
    add_by_query(&qc->rejects, stride, q, rej);
    add_by_query(&qc->expand_bytes, stride, q, bytes);
    add_by_query(&qc->expand_algo_bytes, stride, q, algo);
}

// Copies segments of task records: segment s moves src[src_off[s] ..
// + len) to dst[dst_off[s] ..], where dst_off is the exclusive prefix of
// the lengths (n_seg + 1 entries, the last one the total).
__global__ __launch_bounds__(kBlock) void k_gather_tasks(const DTask* __restrict__ src,
                                                          const uint64_t* __restrict__ src_off,
                                                          const uint64_t* __restrict__ dst_off, uint32_t n_seg,
                                                          DTask* __restrict__ dst) {
    const uint64_t total = dst_off[n_seg];
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * kBlock) {
        const uint32_t sg = find_u64(dst_off, n_seg, i);
        dst[i] = src[src_off[sg] + (i - dst_off[sg])];
    }
}

// Current level of a lane's search: the base list [cur, end) (operand `sel`
// or C(u) itself) and merge cursors over up to four other operand rows.
// Elements of the base list come in increasing order, so every cursor only
// moves forward; an exhausted operand ends the level. Operands beyond the
// fourth are binary-searched.
constexpr uint32_t kCursors = 4;

struct Level {
    uint64_t cur, end;
    uint64_t base0;         // start of the base row (0 for C(u) itself)
    uint32_t sel, nb, nx;   // base operand, cached cursors, all other operands
    uint64_t c0, c1, c2, c3, e0, e1, e2, e3;
};

// Adds a cursor in front, shifting the others back: only constant field
// positions are used, so the level stays in registers (an indexed store into
// the cursors would place the whole level in scratch memory). The order in
// which cursors are tested does not matter: an element must be in all rows.
__device__ __forceinline__ void push_cursor(Level& L, uint64_t lo, uint64_t hi) {
    L.c3 = L.c2; L.e3 = L.e2;
    L.c2 = L.c1; L.e2 = L.e1;
    L.c1 = L.c0; L.e1 = L.e0;
    L.c0 = lo;   L.e0 = hi;
}

// Sets up level d. forced_sel != kAuto keeps the given base operand (a
// restored or donated range); otherwise the shortest row is the base.
constexpr uint32_t kAutoSel = 0xFFFFFFFEu;

__device__ __forceinline__ void enter_level(const DPlan& P, const DOp* __restrict__ ops,
                                            const uint64_t* __restrict__ row_off, uint32_t d,
                                            const LaneState& st, uint32_t forced_sel, Level& L,
                                            unsigned long long& bytes, unsigned long long& algo,
                                            uint32_t max_cursors) {
    const uint32_t nops = P.op_count[d];
    L.nb = L.nx = 0;
    if (nops == 0) {
        L.sel = kIdentity;
        L.cur = 0;
        L.base0 = 0;
        L.end = P.cand_len[d];
        return;
    }
    uint32_t sel = forced_sel;
    if (sel == kAutoSel) {
        uint64_t best = ~0ull;
        for (uint32_t j = 0; j < nops; ++j) {
            const DOp op = ops[P.op_begin[d] + j];
            const uint64_t r = op.row_base + st.I(op.src_depth);
            const uint64_t len = row_off[r + 1] - row_off[r];
            algo += 4ull * len;   // a fresh node: its operand rows
            if (len < best) { best = len; sel = j; }
        }
        bytes += 16ull * nops;
    }
    L.sel = sel;
    {
        const DOp op = ops[P.op_begin[d] + sel];
        const uint64_t r = op.row_base + st.I(op.src_depth);
        L.cur = row_off[r];
        L.base0 = L.cur;
        L.end = row_off[r + 1];
    }
    for (uint32_t j = 0; j < nops; ++j) {
        if (j == sel) continue;
        if (L.nb < max_cursors) {
            const DOp op = ops[P.op_begin[d] + j];
            const uint64_t r = op.row_base + st.I(op.src_depth);
            push_cursor(L, row_off[r], row_off[r + 1]);
            ++L.nb;
        }
        ++L.nx;
    }
    bytes += 16ull * (1 + L.nb);
}

// First position p in [c, e) with idx[p] >= x, galloping from c.
__device__ __forceinline__ uint64_t gallop(const uint32_t* __restrict__ idx, uint64_t c, uint64_t e,
                                           uint32_t x, unsigned long long& bytes) {
    if (c >= e) return e;
    bytes += 4;
    if (idx[c] >= x) return c;
    uint64_t lo = c, step = 1;   // idx[lo] < x
    while (lo + step < e) {
        bytes += 4;
        if (idx[lo + step] >= x) break;
        lo += step;
        step <<= 1;
    }
    uint64_t hi = min(lo + step, e);   // idx[hi] >= x or hi == e
    ++lo;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        bytes += 4;
        if (idx[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// 0: x fails; 1: x is a valid extension (v_out set); 2: an operand is
// exhausted, no later element of this level can match; 3 (test_merge): x is
// in every operand row but its data vertex is already used.
__device__ __forceinline__ int step_cursor(const uint32_t* __restrict__ idx, uint64_t& c, uint64_t e,
                                           uint32_t x, unsigned long long& bytes) {
    c = gallop(idx, c, e, x, bytes);
    if (c >= e) return 2;
    return idx[c] == x ? 1 : 0;
}

// keep_v: the data vertex is needed beyond the injectivity test (a later
// depth compares it, or embeddings are materialised); otherwise it is only
// loaded when an earlier depth shares the label.
__device__ __forceinline__ int test_merge(const DPlan& P, const DOp* __restrict__ ops,
                                          const uint64_t* __restrict__ row_off,
                                          const uint32_t* __restrict__ idx,
                                          const uint32_t* __restrict__ cand_pool, uint32_t d,
                                          uint32_t x, Level& L, const LaneState& st, uint32_t& v_out,
                                          unsigned long long& bytes, bool keep_v) {
    int r;
    if (L.nb > 0 && (r = step_cursor(idx, L.c0, L.e0, x, bytes)) != 1) return r;
    if (L.nb > 1 && (r = step_cursor(idx, L.c1, L.e1, x, bytes)) != 1) return r;
    if (L.nb > 2 && (r = step_cursor(idx, L.c2, L.e2, x, bytes)) != 1) return r;
    if (L.nb > 3 && (r = step_cursor(idx, L.c3, L.e3, x, bytes)) != 1) return r;
    if (L.nx > L.nb) {
        // operands past the cached ones: binary search
        const uint32_t nops = P.op_count[d];
        uint32_t pos = 0;
        for (uint32_t j = 0; j < nops; ++j) {
            if (j == L.sel) continue;
            if (pos++ < L.nb) continue;
            const DOp op = ops[P.op_begin[d] + j];
            const uint64_t rr = op.row_base + st.I(op.src_depth);
            const uint64_t hi0 = row_off[rr + 1];
            uint64_t lo = row_off[rr], hi = hi0;
            bytes += 16;
            while (lo < hi) {
                const uint64_t mid = (lo + hi) >> 1;
                bytes += 4;
                if (idx[mid] < x) lo = mid + 1;
                else hi = mid;
            }
            if (lo == hi0 || idx[lo] != x) return 0;
        }
    }
    unsigned long long mask = P.same_label[d];
    if (!mask && !keep_v) {
        v_out = kInvalid;
        return 1;
    }
    const uint32_t v = cand_pool[P.cand_base[d] + x];
    bytes += 4;
    while (mask) {
        const uint32_t e = __ffsll((long long)mask) - 1;
        mask &= mask - 1;
        if (st.M(e) == v) return 3;
    }
    v_out = v;
    return 1;
}

// Device-coherent loads of a queued task record. A lane that splits a task
// off writes the record, makes it visible at agent scope (release fence) and
// then stores the slot's sequence word. The claiming
// lane loads the sequence word (relaxed), waits for that load to return
// (record_wait: s_waitcnt on the vector memory counter, so no record load can
// issue before the word was observed), and then reads the record with
// agent-scope coherent loads, which are served by the L2 the release wrote
// to. This is an acquire without its L1 invalidate: an acquire fence per
// claim would invalidate the claiming CU's vector L1 millions of times per
// launch, and the coherent loads do not use the L1 anyway.
__device__ __forceinline__ uint32_t coherent_u32(const uint32_t* p) {
    return __hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t coherent_u64(const uint64_t* p) {
    return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void record_wait() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }
// Makes a split-off record visible at agent scope before its sequence word:
// a release fence (L2 writeback and store wait, no L1 invalidate).
__device__ __forceinline__ void record_publish_fence() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }

// Sequence word of ring position p (see DQueue).
__device__ __forceinline__ uint32_t seq_load(uint32_t* seq, uint32_t mask, uint32_t p) {
    return __hip_atomic_load(seq + (p & mask), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// The record of position p is written: claimable. The caller has made the
// record visible at agent scope first (record_publish_fence).
__device__ __forceinline__ void seq_publish(uint32_t* seq, uint32_t mask, uint32_t p) {
    __hip_atomic_store(seq + (p & mask), p + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// The record of position p has been read: its slot is free for p + mask + 1.
// The record loads have returned (record_wait) before the word is stored, so
// a producer that sees it can only overwrite a record nobody reads any more.
// Nothing the claimant wrote needs to be visible, so no release (on gfx950 an
// agent-scope release writes back the L2, at every claim).
__device__ __forceinline__ void seq_release(uint32_t* seq, uint32_t mask, uint32_t p) {
    record_wait();
    __hip_atomic_store(seq + (p & mask), p + mask + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Posts query q's final count to the host-visible notify words (one writer
// per query): the values first, then the flag, with system-scope stores.
__device__ __forceinline__ void post_final(unsigned long long* notify, QCounters* qc, uint32_t q,
                                           unsigned long long count, unsigned long long calls,
                                           unsigned long long rejects, unsigned long long flag = 1ull) {
    if (atomicCAS(&qc[q].notified, 0u, 1u) != 0u) return;
    unsigned long long* w = notify + 4ull * q;
    __hip_atomic_store(w + 1, count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(w + 2, calls, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(w + 3, rejects, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(w, flag, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Query q passed its time limit (resident launch): its remaining tasks are
// dropped, and the count so far is posted as partial (flag 2).
__device__ __forceinline__ void post_timeout(unsigned long long* notify, QCounters* qc, uint32_t q) {
    atomicExch(&qc[q].timed_out, 1u);
    if (notify) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        post_final(notify, qc, q, __hip_atomic_load(&qc[q].count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                   __hip_atomic_load(&qc[q].calls, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                   __hip_atomic_load(&qc[q].rejects, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), 2ull);
    }
}

// Query q's deadline (resident launch): set by the first lane that starts
// one of its tasks, lim ticks from then.
__device__ __forceinline__ unsigned long long query_deadline(QCounters* qc, uint32_t q, unsigned long long lim) {
    const unsigned long long d = __hip_atomic_load(&qc[q].deadline, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (d) return d;
    const unsigned long long want = (unsigned long long)wall_clock64() + lim;
    const unsigned long long old = atomicCAS(&qc[q].deadline, 0ull, want);
    return old ? old : want;
}

// Resident launch: the segments published so far (bit 31: closed). The
// host's word is read over the bus at most once per kPollTicks for the whole
// launch, by whichever wavefront asks first; the others read the copy.
constexpr unsigned long long kPollTicks = 200;   // 2 us of the 100 MHz wall clock
__device__ __forceinline__ uint32_t published(DQueue* Q) {
    uint32_t v = __hip_atomic_load(&Q->pub, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long now = (unsigned long long)wall_clock64();
    const unsigned long long last = __hip_atomic_load(&Q->poll_tick, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (now - last >= kPollTicks && atomicCAS(&Q->poll_tick, last, now) == last) {
        const uint32_t h =
            __hip_atomic_load(const_cast<unsigned int*>(Q->host_word), __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
        if (h > v) {
            atomicMax(&Q->pub, h);
            v = h;
        }
    }
    return v;
}

// Resident launch: tasks left in the current segment, plus one if a later
// segment is published (a hint for the donation rules).
__device__ __forceinline__ uint32_t segment_left(DQueue* Q, const DSeg* segs, unsigned int* seg_head) {
    const uint32_t np = __hip_atomic_load(&Q->pub, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 0x7FFFFFFFu;
    const uint32_t sc = __hip_atomic_load(&Q->seg_cur, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (sc >= np) return 0;
    const uint32_t sn = segs[sc].n;
    const uint32_t sh = __hip_atomic_load(&seg_head[sc], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return (sh < sn ? sn - sh : 0u) + (sc + 1 < np ? 1u : 0u);
}

// Resident launch, one lane of an idle wavefront: is the launch drained
// (every published segment claimed, the ring empty, no task held), and
// should the wavefront leave? It leaves once the launch is drained and
// closed by the host, or drained for idle_ticks since the wavefront last had
// work, with nothing new published. The first to leave that way marks the
// launch exiting: from then on no segment is claimed, and every wavefront
// leaves once the ring is empty and no task is held (a claim that raced
// with the mark runs to its end). What a segment published meanwhile has
// left unclaimed is published again by the host in its next launch, so a
// new segment never runs on the few wavefronts a leaving launch still has.
__device__ __forceinline__ bool resident_idle(DQueue* Q, const DSeg* segs, unsigned int* seg_head,
                                              unsigned long long& idle_since, bool& leave) {
    if (__hip_atomic_load(&Q->exiting, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        leave = __hip_atomic_load(&Q->head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                    __hip_atomic_load(&Q->tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &&
                __hip_atomic_load(&Q->busy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
        return true;
    }
    const uint32_t pv = published(Q);
    const uint32_t np = pv & 0x7FFFFFFFu;
    const uint32_t sc = __hip_atomic_load(&Q->seg_cur, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const bool arrays =
        sc >= np ||
        (sc + 1 == np && __hip_atomic_load(&seg_head[sc], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= segs[sc].n);
    const bool drained = arrays &&
                         __hip_atomic_load(&Q->head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                             __hip_atomic_load(&Q->tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &&
                         __hip_atomic_load(&Q->busy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
    const unsigned long long now = (unsigned long long)wall_clock64();
    leave = false;
    if (!drained) {
        idle_since = 0;
    } else if (pv >> 31) {
        leave = true;
    } else if (!idle_since) {
        idle_since = now;
    } else if (Q->idle_ticks && now - idle_since > Q->idle_ticks) {
        atomicExch(&Q->exiting, 1u);
        leave = true;
    }
    return drained;
}

// Resident launch: an idle wavefront's nap, growing with every poll that
// found nothing (to 8 x 3.4 us once the launch is drained, 4 x 3.4 us while
// other searches run and may split off work). A resident launch spends most
// of its time with most wavefronts idle, beside the kernels that prepare the
// next batches; each poll is a handful of device-scope loads and atomics on
// the queue's two lines, the lines the running wavefronts claim and split
// on. Polling every microsecond from every idle wavefront of a grid of 4 per
// CU served C4 at a seventh of the rate of one per CU.
__device__ __forceinline__ void resident_nap(bool drained, uint32_t& nap) {
    nap = min(nap + 1u, drained ? 8u : 4u);
    for (uint32_t i = 0; i < nap; ++i) __builtin_amdgcn_s_sleep(127);
}

// Query q's count reached its cap now (the first caller's time is kept).
__device__ __forceinline__ void mark_final(QCounters* qc, uint32_t q, unsigned long long cap,
                                           unsigned long long* notify) {
    if (__hip_atomic_load(&qc[q].cap_tick, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
        atomicCAS(&qc[q].cap_tick, 0ull, (unsigned long long)wall_clock64());
    if (notify) post_final(notify, qc, q, cap, 0ull, 0ull);
}
// Subtracts k ended tasks from pending[q] and records the time of the ends:
// the query's last task end is the largest. Without notify words no atomic
// here returns a value, so the lane does not wait for these shared counters.
// With them, the counts the lane flushed before are released ahead of the
// subtraction, and the lane that ends the query's last task reads the final
// counts (acquire) and posts them.
__device__ __forceinline__ void pending_sub(uint32_t* pending, QCounters* qc, uint32_t q, uint32_t k,
                                            unsigned long long* notify) {
    if (notify) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if (atomicSub(pending + q, k) == k && !__atomic_load_n(&qc[q].done, __ATOMIC_RELAXED)) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            post_final(notify, qc, q,
                       __hip_atomic_load(&qc[q].count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                       __hip_atomic_load(&qc[q].calls, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                       __hip_atomic_load(&qc[q].rejects, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        }
    } else {
        atomicSub(pending + q, k);
    }
    atomicMax(&qc[q].end_tick, (unsigned long long)wall_clock64());
}

__device__ __forceinline__ uint32_t load_relaxed(unsigned int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool past_deadline(DQueue* Q) {
    return Q->limit_ticks &&
           (unsigned long long)wall_clock64() >
               __hip_atomic_load(&Q->deadline, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The launch's deadline in wall-clock ticks, 0 without a time limit: the
// first wavefront to arrive sets it, every wavefront keeps the value in a
// register (called by one lane).
__device__ __forceinline__ unsigned long long launch_deadline(DQueue* Q, unsigned long long now) {
    const unsigned long long lim = Q->limit_ticks;
    if (!lim) return 0ull;
    const unsigned long long old = atomicCAS(&Q->deadline, 0ull, now + lim);
    return old ? old : now + lim;
}
__device__ __forceinline__ bool passed(unsigned long long deadline) {
    return deadline && (unsigned long long)wall_clock64() > deadline;
}

// base elements one step of the lane DFS tests at most: the cap and queue
// checks count steps, so a shorter step lets capped queries stop sooner
// (C4 plain path 37 -> 44 queries/s from 32 to 16) at about 1 % of an
// uncapped search's rate
constexpr uint32_t kLeafBurst = 16;
constexpr uint32_t kClaimMax = 8;     // queue slots a lane takes per claim at most
constexpr uint32_t kFewTasks = 64;
constexpr uint32_t kShortTask = 1024; // steps below which a finished task counts as short for claims    // a query with fewer queued and running tasks splits while the ring is short

// Adds a lane's running count to query q's count (setting the query's done
// flag at its cap) and clears it.
__device__ __forceinline__ void lane_flush_count(QCounters* qc, uint32_t q, uint64_t cap,
                                                 unsigned long long& lcount, bool& near_cap,
                                                 unsigned long long* notify) {
    if (lcount) {
        const unsigned long long prev = atomicAdd(&qc[q].count, lcount);
        if (prev + lcount >= cap) {
            atomicExch(&qc[q].done, 1u);
            mark_final(qc, q, cap, notify);
        }
        near_cap = prev + lcount >= cap / 2;
        lcount = 0;
    }
}

// Adds a lane's running count and statistics to query q's counters and
// clears them.
__device__ __forceinline__ void lane_flush(QCounters* qc, uint32_t q, uint64_t cap, unsigned long long& lcount,
                                           unsigned long long& lcalls, unsigned long long& lrej,
                                           unsigned long long& lbytes, unsigned long long& lalgo,
                                           bool& near_cap, unsigned long long* notify) {
    lane_flush_count(qc, q, cap, lcount, near_cap, notify);
    if (lcalls) atomicAdd(&qc[q].calls, lcalls);
    if (lrej) atomicAdd(&qc[q].rejects, lrej);
    if (lbytes) atomicAdd(&qc[q].probe_bytes, lbytes);
    if (lalgo) atomicAdd(&qc[q].algo_bytes, lalgo);
    lcalls = lrej = lbytes = lalgo = 0;
}

// Persistent DFS. Every lane runs its own depth-first search over one task at
// a time. Idle lanes claim the next queued tasks (one claim per wavefront).
// A running lane splits off the second half of the remaining range at its
// shallowest level with at least two entries left, as a new task, while the
// queue is short (the reference's TaskPool donates half the siblings at the
// shallowest splittable depth to idle threads, TaskPool.cpp:960-1362).
// pending[q] counts the queued and running tasks of query q; a task that
// runs to its end decrements it, so a query left with pending work when the
// deadline stops the search is known to be incomplete. A query whose count
// reaches its cap is marked done and its tasks are dropped (count =
// min(cap, total)).
__global__ __launch_bounds__(kBlock) void k_dfs(
    const DPlan* __restrict__ plans, const DOp* __restrict__ ops, const DTask* __restrict__ init, DTask* queue,
    uint32_t* seq, DQueue* Q, QCounters* qc, uint32_t* pending, uint32_t max_n, EmitSink sink) {
    extern __shared__ __align__(16) unsigned char lane_smem[];
    const uint32_t lane = lane_id();
    const uint32_t per_field = blockDim.x * max_n;
    LaneState st;
    st.I_ = reinterpret_cast<uint32_t*>(lane_smem);
    st.M_ = st.I_ + per_field;
    st.cur_ = st.M_ + per_field;
    st.end_ = st.cur_ + per_field;
    st.sel_ = reinterpret_cast<uint8_t*>(st.end_ + per_field);
    st.stride = blockDim.x;
    st.lane = threadIdx.x;

    unsigned long long deadline = 0;
    if (lane == 0) {
        const unsigned long long now = (unsigned long long)wall_clock64();
        deadline = launch_deadline(Q, now);
        if (threadIdx.x == 0 &&
            __hip_atomic_load(&Q->start_tick, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
            atomicCAS(&Q->start_tick, 0ull, now);
    }
    deadline = __shfl(deadline, 0);
    const uint32_t qmask = Q->mask;
    const uint32_t n0 = Q->n0;
    const uint32_t claim_div = max(1u, Q->claim_div);
    // intersection strategy: merge cursors over up to kCursors operand rows
    // (galloping), or binary search in every row
    const uint32_t max_cursors = Q->intersect == PSM_INTERSECT_BSEARCH ? 0u : kCursors;
    unsigned long long* const notify = Q->notify;
    // resident launch: published segments instead of one array
    const DSeg* const segs = Q->segs;
    unsigned int* const seg_head = Q->seg_head;
    uint32_t wseg = 0;                  // segment of the lane's claimed array positions
    unsigned long long qdl = 0;         // deadline of the lane's query (0: none)
    unsigned long long idle_since = 0;  // lane 0: since when the wavefront has seen the launch drained
    uint32_t nap = 0;                   // resident launch: naps of an idle wavefront, growing

    bool active = false, waiting = false, wring = false;   // wring: the claimed positions are ring positions
    uint32_t wslot = 0, wend = 0, backoff = 0;
    uint32_t busy_done = 0, pend_done = 0;   // task ends not yet subtracted from busy / pending[q]
    uint32_t q = 0, tdepth = 0, d = 0, n = 0;
    Level L{};
    Level Lp{};   // the level above the leaf depth, kept while the lane is at the leaf
    unsigned long long lcount = 0, lcalls = 0, lrej = 0, lbytes = 0, lalgo = 0;
    unsigned long long busy_iters = 0, all_iters = 0;
    uint32_t ldonated = 0, lfull = 0;   // this lane's donations, and those skipped for a held slot
    unsigned long long flush_at = 4096;   // leaf counts a lane holds before adding them to its query
    bool near_cap = false;                // the query's count was at half its cap or more at the last flush
    bool hungry = true;                   // wavefront view: the queue holds fewer tasks than idle lanes
    bool ring_short = true;               // wavefront view: the ring holds fewer than lanes / 8 tasks
    bool few_left = false;                // the lane's query had fewer than kFewTasks tasks left
    bool stop = false;                    // wavefront view: the time limit has passed
    bool array_left = true;               // wavefront view: the pre-expanded array still has tasks
    bool check = false;                   // this iteration checks the lanes' caps and the limit
    uint32_t witer = 0;
    uint32_t steps = 0;
    uint32_t last_steps = kShortTask;   // steps of the lane's previous task (none yet: counts as long)
    const DPlan* P = plans;
    // the held query's candidate pool and tables (set with its first task: in
    // a resident launch no plan is read before it is published)
    const uint64_t* __restrict__ row_off = nullptr;
    const uint32_t* __restrict__ idx = nullptr;
    const uint32_t* __restrict__ cand_pool = nullptr;
    const uint32_t lanes_total = gridDim.x * blockDim.x;

    while (true) {
        // Shared state is read on the wavefront's iteration count, with the
        // wavefront converged, never on a lane's own step count: lanes reach
        // their step counts at different iterations, so per-lane checks issue
        // one agent-scope load per lane, and with one query every one of them
        // goes to the same L2 line. Every 16 iterations one lane reads the
        // queue's state (hungry: the ring and the array together hold fewer
        // tasks than there are idle lanes; ring_short: the ring holds fewer
        // than an eighth of the lanes) and the time limit, and, while the
        // ring is short, the running lanes read their queries' task counts
        // for the donation rule (one load per distinct query). The lanes'
        // caps are checked every 16 iterations while a lane's query is near
        // its cap, otherwise every 256.
        check = false;
        if ((witer++ & 15) == 0) {
            uint32_t h = 0;
            if (lane == 0) {
                const uint32_t len = load_relaxed(&Q->tail) - load_relaxed(&Q->head);
                uint32_t left0;
                if (segs) {
                    left0 = segment_left(Q, segs, seg_head);
                } else {
                    const uint32_t h0 = load_relaxed(&Q->head0);
                    left0 = h0 < n0 ? n0 - h0 : 0u;
                }
                bool late = load_relaxed(&Q->timeout) != 0;
                if (!late && passed(deadline)) {
                    atomicExch(&Q->timeout, 1u);
                    late = true;
                }
                h = (len + left0 < lanes_total - min(load_relaxed(&Q->busy), lanes_total) ? 1u : 0u) |
                    (len < (lanes_total >> 3) ? 2u : 0u) | (late ? 4u : 0u) | (left0 ? 8u : 0u);
            }
            h = __shfl(h, 0);
            hungry = (h & 1u) != 0;
            ring_short = (h & 2u) != 0;
            stop = (h & 4u) != 0;
            array_left = (h & 8u) != 0;
            few_left = false;
            if (!hungry && ring_short && active) few_left = load_relaxed(pending + q) < kFewTasks;
            check = stop || (witer & 255u) == 1u || __ballot(active && near_cap) != 0;
        }
        // claim queued tasks for idle lanes, one atomic per wavefront (none
        // once the time limit has passed): split-off work from the ring
        // first, one task per lane, then the pre-expanded array. While the
        // array holds many tasks per lane (more than claim_div) and most of
        // the claiming lanes' previous tasks were short, each lane gets
        // several consecutive positions per claim (up to kClaimMax) and takes
        // them one after the other: claims go through one compare-and-swap
        // on the shared head, whose rate (not the search) bounded launches
        // with millions of small tasks. Otherwise a lane takes one task per
        // claim: a lane holding several heavy subtrees would run them in
        // turn while other lanes wait.
        const unsigned long long want = __ballot(!active && !waiting && !stop);
        // lanes whose previous task was short (small tasks, many claims)
        const unsigned long long quick = __ballot(!active && !waiting && last_steps < kShortTask);
        // While the pre-expanded array holds tasks, claims are fetch-adds
        // that never fail and idle lanes claim at once. Once only split-off
        // work remains, a wavefront with running lanes claims for its idle
        // ones on every 16th iteration: each claim is a compare-and-swap
        // round trip that the wavefront's running lanes wait for, and
        // thousands of wavefronts racing for each donated task slow the
        // donors down; a wavefront with no running lane claims at once.
        const bool claim_now = array_left || (witer & 15) == 1 || __ballot(active) == 0;
        if (want && backoff == 0 && claim_now) {
            const uint32_t leader = __ffsll((long long)want) - 1;
            uint32_t base = 0, got = 0, per = 1, from_ring = 0, empty = 0, cseg = 0;
            if (lane == leader) {
                const uint32_t nw = (uint32_t)__popcll(want);
                const uint32_t h = load_relaxed(&Q->head), t = load_relaxed(&Q->tail);
                // busy counts claimed, unfinished tasks; it rises before a
                // claim so that empty queues with busy == 0 are only seen
                // once no lane holds or is taking a task
                if (t != h) {
                    // ring positions are claimed only up to the tail (a
                    // position past it may never be published), so a claim
                    // that loses the race to another wavefront takes nothing
                    // and is tried again on the next iteration
                    const uint32_t k = min(nw, t - h);
                    atomicAdd(&Q->busy, k);
                    if (atomicCAS(&Q->head, h, h + k) == h) {
                        base = h;
                        got = k;
                        from_ring = 1;
                    } else {
                        atomicSub(&Q->busy, k);
                    }
                } else if (segs) {
                    // resident launch: the oldest published segment with
                    // tasks left, claimed as the array is (one fetch-add);
                    // an exhausted one moves the shared cursor on
                    const uint32_t np = load_relaxed(&Q->exiting) ? 0u : published(Q) & 0x7FFFFFFFu;
                    const uint32_t sc = load_relaxed(&Q->seg_cur);
                    if (sc < np) {
                        const uint32_t sn = segs[sc].n;
                        const uint32_t sh = load_relaxed(&seg_head[sc]);
                        if (sh < sn) {
                            if (2u * (uint32_t)__popcll(quick & want) >= nw)
                                per = min(kClaimMax, max(1u, (sn - sh) / (claim_div * lanes_total)));
                            const uint32_t k = min(nw * per, sn - sh);
                            atomicAdd(&Q->busy, k);
                            const uint32_t old = atomicAdd(&seg_head[sc], k);
                            got = old < sn ? min(k, sn - old) : 0u;
                            base = old;
                            cseg = sc;
                            if (got < k) atomicSub(&Q->busy, k - got);
                        } else {
                            atomicCAS(&Q->seg_cur, sc, sc + 1);
                        }
                    } else {
                        empty = 1;
                    }
                } else {
                    const uint32_t h0 = load_relaxed(&Q->head0);
                    if (h0 < n0) {
                        // several per claim while most idle lanes ran short
                        // tasks (the queue then turns over fast); one per
                        // claim after long ones
                        if (2u * (uint32_t)__popcll(quick & want) >= nw)
                            per = min(kClaimMax, max(1u, (n0 - h0) / (claim_div * lanes_total)));
                        const uint32_t k = min(nw * per, n0 - h0);
                        // the array is complete before the launch, so a claim
                        // is one fetch-add that cannot fail: concurrent
                        // claims get consecutive ranges, and one that starts
                        // past n0 gets nothing (head0 overshoots n0 by at
                        // most one claim per wavefront)
                        atomicAdd(&Q->busy, k);
                        const uint32_t old = atomicAdd(&Q->head0, k);
                        got = old < n0 ? min(k, n0 - old) : 0u;
                        base = old;
                        if (got < k) atomicSub(&Q->busy, k - got);
                    } else {
                        empty = 1;
                    }
                }
            }
            base = __shfl(base, leader);
            got = __shfl(got, leader);
            per = __shfl(per, leader);
            from_ring = __shfl(from_ring, leader);
            empty = __shfl(empty, leader);
            cseg = __shfl(cseg, leader);
            if (!active && !waiting) {
                const uint32_t r = __popcll(want & lanemask_lt());
                if (r * per < got) {
                    waiting = true;
                    wring = from_ring != 0;
                    wseg = cseg;
                    wslot = base + r * per;
                    wend = min(wslot + per, base + got);
                }
            }
            // back off only when both queues were empty: a lost race is
            // retried at once (the idle lanes have nothing else to do)
            if (!got && empty) backoff = 16;
        } else if (backoff) {
            --backoff;
        }
        if (waiting && (!wring || seq_load(seq, qmask, wslot) == wslot + 1u)) {
            waiting = false;
            if (wring) record_wait();
            // the record is read field by field with coherent loads (a local
            // copy of its prefix array would live in scratch memory)
            const DTask* task = wring ? queue + (wslot & qmask) : (segs ? segs[wseg].tasks : init) + wslot;
            // array tasks of a segment hold the query's index in its batch
            const uint32_t tq = coherent_u32(&task->q) + (wring || !segs ? 0u : segs[wseg].q_base);
            if (tq != q) {
                // the lane's counters so far belong to the previous query
                lane_flush(qc, q, P->cap, lcount, lcalls, lrej, lbytes, lalgo, near_cap, notify);
                if (pend_done) pending_sub(pending, qc, q, pend_done, notify);
                pend_done = 0;
            }
            q = tq;
            P = plans + q;
            row_off = P->row_off;
            idx = P->idx;
            cand_pool = P->cand_pool;
            qdl = P->limit_ticks ? query_deadline(qc, q, P->limit_ticks) : 0ull;
            const bool late = passed(qdl);
            if (late) post_timeout(notify, qc, q);
            if (stop || late || __atomic_load_n(&qc[q].done, __ATOMIC_RELAXED)) {
                if (wring) seq_release(seq, qmask, wslot);
                ++busy_done;
                if (wslot + 1 < wend) {   // the lane's next claimed slot
                    ++wslot;
                    waiting = true;
                } else {
                    atomicSub(&Q->busy, busy_done);
                    busy_done = 0;
                    if (pend_done) pending_sub(pending, qc, q, pend_done, notify);
                    pend_done = 0;
                }
            } else {
                n = P->n;
                tdepth = coherent_u32(&task->depth);
                d = tdepth;
                for (uint32_t e = 0; e < tdepth; ++e) st.I_[st.at(e)] = coherent_u32(&task->prefix[e]);
                const uint64_t thi = coherent_u64(&task->hi);
                const uint64_t tlo = coherent_u64(&task->lo);
                const uint32_t tsel = coherent_u32(&task->sel);
                if (wring) seq_release(seq, qmask, wslot);
                for (uint32_t e = 0; e < tdepth; ++e) st.M_[st.at(e)] = cand_pool[P->cand_base[e] + st.I_[st.at(e)]];
                if (thi == kAutoHi) {
                    enter_level(*P, ops, row_off, d, st, kAutoSel, L, lbytes, lalgo, max_cursors);
                } else {
                    enter_level(*P, ops, row_off, d, st, tsel, L, lbytes, lalgo, max_cursors);
                    L.cur = tlo;
                    L.end = thi;
                }
                active = true;
                steps = 0;
                // a capped query learns its running count sooner (smaller
                // flushes, more frequent checks): every lane's unflushed
                // embeddings and unchecked steps are searched past the cap
                near_cap = false;
                flush_at = P->cap != ~0ull ? min(4096ull, max(16ull, P->cap >> 12)) : 4096ull;
            }
        }
        if (__ballot(active || waiting) == 0) {
            // the whole wavefront is idle: leave once no work can appear
            uint32_t leave = 0, drained_w = 0;
            if (lane == 0 && segs) {
                bool lv = false;
                drained_w = resident_idle(Q, segs, seg_head, idle_since, lv) ? 1u : 0u;
                leave = lv ? 1u : 0u;
            } else if (lane == 0) {
                leave = load_relaxed(&Q->timeout) != 0 ||
                        (load_relaxed(&Q->head) == load_relaxed(&Q->tail) && load_relaxed(&Q->head0) >= n0 &&
                         load_relaxed(&Q->busy) == 0);
            }
            if (__shfl(leave, 0)) break;
            if (segs) {
                resident_nap(__shfl(drained_w, 0) != 0, nap);
                continue;
            }
            // a short nap: idle wavefronts must take donated work at once
            // (longer waits between polls made lone searches slower)
            __builtin_amdgcn_s_sleep(4);
            continue;
        }
        nap = 0;
        idle_since = 0;
        ++all_iters;
        if (!active) continue;
        ++busy_iters;

        // one step of this lane's search: a burst of up to kLeafBurst tests of
        // base elements of the current level. A lane at the leaf depth counts
        // the valid ones; a lane above it stops at its first valid one and
        // descends after the burst. Lanes at any depth thus skip failing
        // elements at the same pace, and descents happen together.
        bool finished = false;
        if (L.cur < L.end) {
            const bool count_leaf = d + 1 == n && !sink.rows;
            const bool keep_v = sink.rows != nullptr || ((P->m_needed >> d) & 1ull);
            bool found = false;
            uint32_t fx = 0, fv = 0;
            uint32_t burst = kLeafBurst;
            while (L.cur < L.end && burst--) {
                const uint32_t x = L.sel == kIdentity ? (uint32_t)L.cur : idx[L.cur];
                lbytes += L.sel == kIdentity ? 0 : 4;
                ++L.cur;
                uint32_t v;
                const int r = test_merge(*P, ops, row_off, idx, cand_pool, d, x, L, st, v, lbytes, keep_v);
                if (r == 2) {
                    L.cur = L.end;   // an operand ran out: nothing further at this level
                } else if (r == 3) {
                    ++lrej;
                } else if (r == 1) {
                    ++lcalls;
                    if (count_leaf) {
                        ++lcount;
                    } else {
                        found = true;
                        fx = x;
                        fv = v;
                        break;
                    }
                }
            }
            if (found && d + 1 == n) {
                ++lcount;
                // materialise the embedding (data vertex ids in depth order)
                const unsigned long long row = atomicAdd(sink.count, 1ull);
                if (row < sink.cap) {
                    uint32_t* out = sink.rows + row * n;
                    for (uint32_t e = 0; e < d; ++e) out[e] = st.M_[st.at(e)];
                    out[d] = fv;
                } else {
                    // no room: the unit is rejected, so stop searching it
                    atomicExch(sink.overflow, 1u);
                    atomicExch(&qc[q].done, 1u);
                }
            } else if (found) {
                const uint32_t a = st.at(d);
                st.cur_[a] = (uint32_t)(L.cur - L.base0);
                st.end_[a] = (uint32_t)(L.end - L.base0);
                st.sel_[a] = L.sel == kIdentity ? kSelIdentity : (uint8_t)L.sel;
                st.I_[a] = fx;
                st.M_[a] = fv;
                if (d >= 1) lalgo += 4;   // a survivor written below the root level
                ++d;
                if (d + 1 == n) Lp = L;   // restored on the way back up
                enter_level(*P, ops, row_off, d, st, kAutoSel, L, lbytes, lalgo, max_cursors);
            }
        } else if (d == tdepth) {
            finished = true;
            active = false;
        } else {
            --d;
            const uint32_t a = st.at(d);
            if (d + 2 == n) {
                // back from the leaf depth: the level's cursors are still in
                // registers; only its end may have moved (donation)
                L = Lp;
                L.end = L.base0 + st.end_[a];
            } else {
                const uint32_t sel = st.sel_[a] == kSelIdentity ? kIdentity : st.sel_[a];
                enter_level(*P, ops, row_off, d, st, sel, L, lbytes, lalgo, max_cursors);
                L.cur = L.base0 + st.cur_[a];
                L.end = L.base0 + st.end_[a];
            }
        }
        ++steps;
        // the lanes' caps and the time limit (see the top of the loop)
        if (check && active) {
            if (stop || __atomic_load_n(&qc[q].done, __ATOMIC_RELAXED)) {
                active = false;
            } else if (passed(qdl)) {
                // the query's own time limit (resident launch): post what
                // it has counted, the lane's share included
                lane_flush(qc, q, P->cap, lcount, lcalls, lrej, lbytes, lalgo, near_cap, notify);
                post_timeout(notify, qc, q);
                active = false;
            }
        }
        // donation: split the shallowest level that still has >= 2 entries
        // into the ring when it holds fewer tasks than there are lanes
        // without one (work goes to lanes that wait for it), or, while it is
        // short, when the lane's query has few tasks left (a search running
        // on a handful of lanes is what the launch waits for; its split-off
        // work is claimed before fresh array tasks). While every lane is busy
        // with well-spread queries, splitting would only add claims and
        // records.
        if (active && (steps & 15) == 0 && (hungry || few_left)) {
            uint32_t e = tdepth;
            uint64_t ce = 0, ee = 0;
            for (; e <= d; ++e) {
                ce = e == d ? L.cur : st.cur_[st.at(e)];
                ee = e == d ? L.end : st.end_[st.at(e)];
                if (ee - ce >= 2) break;
            }
            if (e <= d && e <= kPrefixMax) {
                const uint32_t t = load_relaxed(&Q->tail);
                const bool free_slot = seq_load(seq, qmask, t) == t;
                if (!free_slot) {
                    ++lfull;
                } else if (atomicCAS(&Q->tail, t, t + 1) == t) {
                    // shallower levels hold offsets from their base row: locate it
                    uint64_t base = 0;
                    uint32_t esel = L.sel;
                    if (e < d) {
                        const uint8_t sb = st.sel_[st.at(e)];
                        esel = sb == kSelIdentity ? kIdentity : sb;
                        if (esel != kIdentity) {
                            const DOp op = ops[P->op_begin[e] + esel];
                            base = row_off[op.row_base + st.I(op.src_depth)];
                        }
                    }
                    const uint64_t mid = ce + (ee - ce) / 2;
                    atomicAdd(pending + q, 1u);
                    DTask& nt = queue[t & qmask];
                    nt.q = q;
                    nt.depth = e;
                    nt.sel = esel;
                    nt.pad = 0;
                    nt.lo = base + mid;
                    nt.hi = base + ee;
                    for (uint32_t k = 0; k < e; ++k) nt.prefix[k] = st.I_[st.at(k)];
                    record_publish_fence();
                    seq_publish(seq, qmask, t);
                    if (e == d) L.end = mid;
                    else st.end_[st.at(e)] = (uint32_t)mid;
                    ++ldonated;
                }
            }
        }
        // A lane's counters go to its query and to the queue when it has no
        // claimed slot left (or moves to another query), not at every task
        // end: with millions of small tasks, device-scope atomics on the few
        // shared counters queue behind one another and every later load of
        // the issuing wavefront waits for them. The running count alone goes
        // to the query as often as its cap needs (every flush_at embeddings);
        // the statistics wait for the task's end or its next 1024 steps.
        const bool more = !active && wslot + 1 < wend;   // another claimed slot follows
        const bool idle_next = !active && !more;
        if (idle_next || (steps & 1023) == 0)
            lane_flush(qc, q, P->cap, lcount, lcalls, lrej, lbytes, lalgo, near_cap, notify);
        else if (lcount >= flush_at)
            lane_flush_count(qc, q, P->cap, lcount, near_cap, notify);
        if (finished) ++pend_done;
        if (!active) {
            last_steps = steps;
            ++busy_done;
            if (more) {
                ++wslot;
                waiting = true;
            } else {
                atomicSub(&Q->busy, busy_done);
                busy_done = 0;
                if (pend_done) pending_sub(pending, qc, q, pend_done, notify);
                pend_done = 0;
            }
        }
    }
    lane_flush(qc, q, P->cap, lcount, lcalls, lrej, lbytes, lalgo, near_cap, notify);
    wave_add(&Q->lane_steps, busy_iters);
    wave_add(&Q->lane_iters, all_iters);
    // statistics of the queue, once per wavefront (on the queue's line every
    // update would contend with claims and donations)
    wave_add_u32(&Q->donated, ldonated);
    wave_add_u32(&Q->full, lfull);
}

// ---------------------------------------------------------------------------
// Wavefront-cooperative DFS. One wavefront (a one-wave workgroup) walks the
// search tree of one task at a time, with its stack in LDS. The unit of work
// is a generation: producing the candidates of depth g for a set of members,
// the siblings of one run at depth g-1, which share every assignment above
// g-1 (committed in the LDS path I/M). The members' base lists (each
// member's shortest operand row) are concatenated and cut into chunks of 64
// positions; every lane tests one position (coalesced load of the base
// element, binary search in the member's other operand rows, injectivity
// against the committed path and the member), and the survivors of a chunk
// are compacted into the depth's LDS frame tagged with their member.
// Survivors with the same tag form a run: the wavefront commits their parent
// and generates depth g+1 from that run. The leaf depth only counts
// survivors. Short base lists of many siblings thus share one chunk, a hub's
// long list fills whole chunks, and every load of a chunk is independent
// across lanes. Work is split off to other wavefronts as in the lane-per-task
// kernel: half of the remaining base range of the current member at the
// shallowest depth that has at least two positions left.
// ---------------------------------------------------------------------------
constexpr uint32_t kWCache = 8;          // operand rows per member kept in LDS besides the base
constexpr uint32_t kStage = 1024;        // entries of an operand row window staged in LDS
constexpr uint32_t kCoopRow = 1024;      // operand rows at least this long are tested cooperatively
constexpr uint64_t kFresh = ~0ull;       // generation cursor: start at the member's base

// One wavefront's LDS, for queries of up to N vertices.
template <uint32_t N>
struct WaveFrame {
    uint32_t I[N];          // committed candidate indices
    uint32_t M[N];          // committed data vertices
    uint32_t run_lo[N];     // current run of the depth's frame
    uint32_t run_hi[N];
    uint32_t ccount[N];     // entries in the depth's frame
    uint32_t gen_m[N];      // generation cursor: member,
    uint32_t gen_cnt[N];    //   number of members,
    uint32_t gen_sel[N];    //   base operand of the cursor member,
    uint32_t counted[N];    //   members whose operand rows are counted
    uint64_t gen_pos[N];    //   next base position (kFresh: not started),
    uint64_t gen_end[N];    //   end of the cursor member's base range
    uint32_t cx[N][kWave];  // frame: candidate index
    uint8_t ctag[N][kWave]; // frame: member tag
    // member cache of the active generation
    uint64_t m_lo[kWave];   // first base position
    uint64_t m_pre[kWave + 1];   // exclusive prefix of base lengths
    uint32_t m_len[kWave];
    uint32_t m_sel[kWave];
    uint32_t m_x[kWave];    // member candidate index (depth g-1)
    uint32_t m_v[kWave];    // member data vertex
    uint64_t m_olo[kWCache][kWave];   // other operand rows: start
    uint32_t m_olen[kWCache][kWave];  //                     length
    uint32_t stage[kStage];           // window of one operand row (cooperative test)
};

// First position p in idx[lo, lo + len) with idx[p] >= v, found by the whole
// wavefront (every lane calls it with the same lo, len, v): each round the
// 64 lanes probe 64 evenly spaced positions and a ballot keeps the one
// interval that holds the answer, so a row of m entries takes
// log64(m) rounds of independent loads instead of log2(m) dependent ones.
__device__ __forceinline__ uint64_t wave_lower_bound(const uint32_t* __restrict__ idx, uint64_t lo, uint64_t len,
                                                     uint32_t v, unsigned long long& bytes) {
    const uint32_t lane = threadIdx.x;
    while (len > kWave) {
        const uint64_t stride = (len + kWave - 1) / kWave;
        const uint64_t p = lo + (uint64_t)lane * stride;
        const bool valid = p < lo + len;
        const bool less = valid && idx[p] < v;
        if (valid) bytes += 4;
        const uint32_t c = (uint32_t)__popcll(__ballot(less));
        if (c == 0) return lo;
        const uint64_t left = lo + (uint64_t)(c - 1) * stride + 1;
        const uint64_t right = min(lo + (uint64_t)c * stride, lo + len);
        lo = left;
        len = right - left;
    }
    const bool valid = lane < len;
    const bool less = valid && idx[lo + lane] < v;
    if (valid) bytes += 4;
    return lo + (uint32_t)__popcll(__ballot(less));
}

// Is x in the sorted stage[0, n)?
__device__ __forceinline__ bool stage_has(const uint32_t* st, uint32_t n, uint32_t x) {
    const uint32_t end = n;
    uint32_t lo = 0;
    while (n > 0) {
        const uint32_t half = n >> 1;
        if (st[lo + half] < x) {
            lo += half + 1;
            n -= half + 1;
        } else {
            n = half;
        }
    }
    return lo < end && st[lo] == x;
}

// Is x in the sorted row idx[lo, lo + len)? (lower bound, then compare)
__device__ __forceinline__ bool row_has(const uint32_t* __restrict__ idx, uint64_t lo, uint64_t len,
                                        uint32_t x, unsigned long long& bytes) {
    const uint64_t end = lo + len;
    while (len > 0) {
        const uint64_t half = len >> 1;
        bytes += 4;
        if (idx[lo + half] < x) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    if (lo >= end) return false;
    bytes += 4;
    return idx[lo] == x;
}

// Candidate index of the parent of depth-g members' operand rows: operand
// rows of depth g whose source is depth g-1 take the member's own index,
// the others the committed path (the root generation g == t has no members:
// the whole prefix is committed).
template <typename F>
__device__ __forceinline__ uint32_t op_source(const F& S, const DOp& op, uint32_t g, uint32_t t,
                                              uint32_t member_x) {
    return (g > t && op.src_depth == g - 1) ? member_x : S.I[op.src_depth];
}

// Fills the member cache of generation g (all lanes; lane k = member k) and
// returns the total base length of the remaining members.
template <typename F>
__device__ __forceinline__ uint64_t wave_build_cache(const DPlan& P, const DOp* __restrict__ ops,
                                     const uint64_t* __restrict__ row_off,
                                     const uint32_t* __restrict__ cand_pool, F& S, uint32_t g,
                                     uint32_t t, uint32_t& cnt_out, unsigned long long& bytes,
                                     unsigned long long& algo) {
    const uint32_t lane = threadIdx.x;
    const uint32_t cnt = g == t ? 1u : S.run_hi[g - 1] - S.run_lo[g - 1];
    const uint32_t gm = S.gen_m[g];
    const uint64_t gpos = S.gen_pos[g];
    uint64_t len = 0;
    if (lane < cnt) {
        uint32_t x = 0, v = 0;
        if (g > t) {
            x = S.cx[g - 1][S.run_lo[g - 1] + lane];
            if ((P.m_needed >> (g - 1)) & 1ull) {
                v = cand_pool[P.cand_base[g - 1] + x];
                bytes += 4;
            }
        }
        const uint32_t nops = P.op_count[g];
        uint64_t lo = 0;
        uint32_t sel = kIdentity;
        if (lane < gm) {
            len = 0;
            sel = 0;
        } else if (lane == gm && gpos != kFresh) {
            lo = gpos;
            len = S.gen_end[g] - gpos;
            sel = S.gen_sel[g];
        } else if (nops == 0) {
            lo = 0;
            len = P.cand_len[g];
            sel = kIdentity;
        } else {
            uint64_t best = ~0ull, rows = 0;
            for (uint32_t j = 0; j < nops; ++j) {
                const DOp op = ops[P.op_begin[g] + j];
                const uint64_t r = op.row_base + op_source(S, op, g, t, x);
                const uint64_t a = row_off[r], b = row_off[r + 1];
                rows += b - a;
                if (b - a < best) { best = b - a; lo = a; sel = j; }
            }
            bytes += 16ull * nops;
            // a fresh member is a node entered here; a cache rebuilt for the
            // same run counts it only once
            if (lane >= S.counted[g]) algo += 4ull * rows;
            len = best;
        }
        if (sel != kIdentity && len > 0) {
            uint32_t slot = 0;
            for (uint32_t j = 0; j < nops && slot < kWCache; ++j) {
                if (j == sel) continue;
                const DOp op = ops[P.op_begin[g] + j];
                const uint64_t r = op.row_base + op_source(S, op, g, t, x);
                const uint64_t a = row_off[r], b = row_off[r + 1];
                S.m_olo[slot][lane] = a;
                S.m_olen[slot][lane] = (uint32_t)(b - a);
                ++slot;
            }
            bytes += 16ull * slot;
        }
        S.m_lo[lane] = lo;
        S.m_len[lane] = (uint32_t)len;
        S.m_sel[lane] = sel;
        S.m_x[lane] = x;
        S.m_v[lane] = v;
    }
    // exclusive prefix of the lengths over the members
    uint64_t incl = len;
    for (uint32_t o = 1; o < kWave; o <<= 1) {
        const uint32_t lo32 = __shfl_up((uint32_t)incl, o), hi32 = __shfl_up((uint32_t)(incl >> 32), o);
        if (lane >= o) incl += ((uint64_t)hi32 << 32) | lo32;
    }
    S.m_pre[lane + 1] = incl;
    const uint32_t counted = S.counted[g];
    if (lane == 0) {
        S.m_pre[0] = 0;
        S.gen_cnt[g] = cnt;
    }
    __syncthreads();
    if (lane == 0) S.counted[g] = max(counted, cnt);
    const uint32_t tl = __shfl((uint32_t)incl, kWave - 1), th = __shfl((uint32_t)(incl >> 32), kWave - 1);
    __syncthreads();
    cnt_out = cnt;
    return ((uint64_t)th << 32) | tl;
}

// Size of the contiguous run of equal tags starting at frame position lo.
template <typename F>
__device__ __forceinline__ uint32_t run_length(const F& S, uint32_t depth, uint32_t lo) {
    const uint32_t lane = threadIdx.x;
    const uint32_t n = S.ccount[depth];
    const uint8_t tag = S.ctag[depth][lo];
    const bool in = lane >= lo && lane < n && S.ctag[depth][lane] == tag;
    return (uint32_t)__popcll(__ballot(in));
}

template <uint32_t N>
__global__ __launch_bounds__(kWave) void k_dfs_wave(
    const DPlan* __restrict__ plans, const DOp* __restrict__ ops, const DTask* __restrict__ init, DTask* queue,
    uint32_t* seq, DQueue* Q, QCounters* qc, uint32_t* pending) {
    const uint32_t lane = threadIdx.x;
    __shared__ WaveFrame<N> S;
    if (lane == 0) {
        const unsigned long long now = (unsigned long long)wall_clock64();
        if (Q->limit_ticks) atomicCAS(&Q->deadline, 0ull, now + Q->limit_ticks);
        if (__hip_atomic_load(&Q->start_tick, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
            atomicCAS(&Q->start_tick, 0ull, now);
    }
    const uint32_t qmask = Q->mask;
    const uint32_t n0 = Q->n0;
    unsigned long long* const notify = Q->notify;
    const uint32_t hungry_below = gridDim.x / 2;   // queue shorter than half the wavefronts
    // intersection strategy: rows of kCoopRow+ entries tested cooperatively
    // through an LDS-staged window, or every row by per-lane binary search
    const uint64_t coop_row = Q->intersect == PSM_INTERSECT_BSEARCH ? ~0ull : (uint64_t)kCoopRow;
    unsigned long long bytes = 0, algo = 0, busy_pos = 0, all_pos = 0;
    // resident launch: published segments instead of one array
    const DSeg* const segs = Q->segs;
    unsigned int* const seg_head = Q->seg_head;
    unsigned long long idle_since = 0;   // lane 0: since when the launch has looked drained
    uint32_t nap = 0;

    while (true) {
        // ---- claim one task (none once the time limit has passed: queued
        // tasks stay counted in pending[] and their queries report a timeout)
        // split-off work from the ring first, then the pre-expanded array
        uint32_t slot = kInvalid, late = 0, from_ring = 0, sseg = 0;
        if (lane == 0) late = load_relaxed(&Q->timeout);
        if (__shfl(late, 0)) break;
        if (lane == 0) {
            const uint32_t h = load_relaxed(&Q->head), tl = load_relaxed(&Q->tail);
            if (tl != h) {
                atomicAdd(&Q->busy, 1u);   // before the claim, as in k_dfs
                if (atomicCAS(&Q->head, h, h + 1) == h) {
                    slot = h;
                    from_ring = 1;
                } else {
                    atomicSub(&Q->busy, 1u);
                }
            } else if (segs) {
                // resident launch: the oldest published segment with tasks
                // left (as in k_dfs)
                const uint32_t np = load_relaxed(&Q->exiting) ? 0u : published(Q) & 0x7FFFFFFFu;
                const uint32_t sc = load_relaxed(&Q->seg_cur);
                if (sc < np) {
                    const uint32_t sn = segs[sc].n;
                    if (load_relaxed(&seg_head[sc]) < sn) {
                        atomicAdd(&Q->busy, 1u);
                        const uint32_t old = atomicAdd(&seg_head[sc], 1u);
                        if (old < sn) {
                            slot = old;
                            sseg = sc;
                        } else {
                            atomicSub(&Q->busy, 1u);
                        }
                    } else {
                        atomicCAS(&Q->seg_cur, sc, sc + 1);
                    }
                }
            } else {
                const uint32_t h0 = load_relaxed(&Q->head0);
                if (h0 < n0) {
                    // one fetch-add on the complete array (as in k_dfs)
                    atomicAdd(&Q->busy, 1u);
                    const uint32_t old = atomicAdd(&Q->head0, 1u);
                    if (old < n0) slot = old;
                    else atomicSub(&Q->busy, 1u);
                }
            }
        }
        slot = __shfl(slot, 0);
        from_ring = __shfl(from_ring, 0);
        sseg = __shfl(sseg, 0);
        if (slot == kInvalid) {
            all_pos += 1;
            uint32_t stop = 0, drained_w = 0;
            if (lane == 0 && segs) {
                bool lv = false;
                drained_w = resident_idle(Q, segs, seg_head, idle_since, lv) ? 1u : 0u;
                stop = lv ? 1u : 0u;
            } else if (lane == 0) {
                stop = load_relaxed(&Q->timeout) ||
                       (load_relaxed(&Q->head) == load_relaxed(&Q->tail) && load_relaxed(&Q->head0) >= n0 &&
                        load_relaxed(&Q->busy) == 0);
            }
            if (__shfl(stop, 0)) break;
            if (segs) {
                resident_nap(__shfl(drained_w, 0) != 0, nap);
                continue;
            }
            __builtin_amdgcn_s_sleep(8);
            continue;
        }
        nap = 0;
        idle_since = 0;
        if (from_ring) {
            while (seq_load(seq, qmask, slot) != slot + 1u) __builtin_amdgcn_s_sleep(1);
            record_wait();
        }
        // read with coherent loads (see coherent_u32)
        const DTask* task = from_ring ? queue + (slot & qmask) : (segs ? segs[sseg].tasks : init) + slot;
        // every lane reads the same record: q (and the plan's pointers) are
        // wavefront-uniform; array tasks of a segment hold the query's index
        // in its batch
        const uint32_t q = __builtin_amdgcn_readfirstlane(
            coherent_u32(&task->q) + (from_ring || !segs ? 0u : segs[sseg].q_base));
        const DPlan& P = plans[q];
        const uint64_t* __restrict__ row_off = P.row_off;
        const uint32_t* __restrict__ idx = P.idx;
        const uint32_t* __restrict__ cand_pool = P.cand_pool;
        uint32_t drop = 0;
        unsigned long long qdl = 0;   // lane 0: the query's own deadline (resident launch)
        if (lane == 0) {
            drop = load_relaxed(&Q->timeout) || __atomic_load_n(&qc[q].done, __ATOMIC_RELAXED);
            if (!drop && past_deadline(Q)) {
                atomicExch(&Q->timeout, 1u);
                drop = 1;
            }
            if (P.limit_ticks) qdl = query_deadline(qc, q, P.limit_ticks);
            if (!drop && passed(qdl)) {
                post_timeout(notify, qc, q);
                drop = 1;
            }
            if (drop) {
                if (from_ring) seq_release(seq, qmask, slot);
                atomicSub(&Q->busy, 1u);
            }
        }
        if (__shfl(drop, 0)) continue;
        const uint32_t n = P.n, t = coherent_u32(&task->depth);
        if (lane < t) {
            const uint32_t x = coherent_u32(&task->prefix[lane]);
            S.I[lane] = x;
            S.M[lane] = cand_pool[P.cand_base[lane] + x];
        }
        if (lane == 0) {
            S.gen_m[t] = 0;
            S.gen_cnt[t] = 1;
            S.counted[t] = 0;
            const uint64_t thi = coherent_u64(&task->hi);
            if (thi == kAutoHi) {
                S.gen_pos[t] = kFresh;
            } else {
                S.gen_pos[t] = coherent_u64(&task->lo);
                S.gen_end[t] = thi;
                S.gen_sel[t] = coherent_u32(&task->sel);
            }
        }
        __syncthreads();
        // one wavefront: lane 0's release waits for every lane's record loads
        if (lane == 0 && from_ring) seq_release(seq, qmask, slot);

        uint32_t g = t, cnt = 0, it = 0;
        int cache_g = -1;
        uint64_t total = 0, flat = 0;
        unsigned long long lcount = 0, lcalls = 0, lrej = 0;
        bool finished = false;
        while (true) {
            if (cache_g != (int)g) {
                total = wave_build_cache(P, ops, row_off, cand_pool, S, g, t, cnt, bytes, algo);
                flat = 0;
                cache_g = (int)g;
            }
            if (flat >= total) {
                // generation g exhausted: its members' run in frame g-1 is done
                if (g == t) {
                    finished = true;
                    break;
                }
                const uint32_t f = g - 1;
                const uint32_t lo = S.run_hi[f];
                if (lo < S.ccount[f]) {
                    const uint32_t len = run_length(S, f, lo);
                    const uint32_t tag = S.ctag[f][lo];
                    __syncthreads();
                    if (lane == 0) {
                        S.run_lo[f] = lo;
                        S.run_hi[f] = lo + len;
                        if (f > t) {
                            // parent of the new run: member `tag` of generation f
                            const uint32_t x = S.cx[f - 1][S.run_lo[f - 1] + tag];
                            S.I[f - 1] = x;
                            S.M[f - 1] = cand_pool[P.cand_base[f - 1] + x];
                        }
                        S.gen_m[g] = 0;
                        S.gen_pos[g] = kFresh;
                        S.counted[g] = 0;
                    }
                    __syncthreads();
                } else {
                    g = f;   // resume generation f at its cursor
                }
                cache_g = -1;
                continue;
            }

            // ---- one chunk of 64 positions of generation g
            const uint64_t fpos = flat + lane;
            bool ok = false, rej = false;
            uint32_t xx = 0, k = 0, sel = kIdentity;
            const bool act = fpos < total;
            all_pos += 1;
            if (act) {
                ++busy_pos;
                uint32_t lo = 0, hi = cnt;
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (S.m_pre[mid] <= fpos) lo = mid;
                    else hi = mid;
                }
                k = lo;
                const uint64_t pos = S.m_lo[k] + (fpos - S.m_pre[k]);
                sel = S.m_sel[k];
                if (sel == kIdentity) {
                    xx = (uint32_t)pos;
                } else {
                    xx = idx[pos];
                    bytes += 4;
                }
                ok = true;
            }
            // When every active lane holds a position of one member (a long
            // base list), its cached operand rows are tested cooperatively:
            // the elements ascend with the lane, so only the window of each
            // row between the chunk's smallest and largest element can hold
            // them; the wavefront locates it (wave_lower_bound), stages it in
            // LDS when it fits and every lane searches LDS. Chunks spread
            // over several members test lane by lane.
            const unsigned long long am = __ballot(act);
            const uint32_t k0 = __shfl(k, 0);   // lane 0 is active whenever any lane is
            const bool uni = am != 0 && __ballot(act && k != k0) == 0;
            const uint32_t nops = P.op_count[g];
            const uint32_t nother = sel == kIdentity ? 0u : nops - 1;
            const uint32_t ncache = min(nother, kWCache);
            const uint32_t sel0 = __shfl(sel, 0);
            const uint32_t ncache0 = sel0 == kIdentity ? 0u : min(nops - 1, kWCache);   // wave-uniform
            if (uni && ncache0 > 0) {
                const uint32_t last = 63u - (uint32_t)__clzll(am);
                const uint32_t xmin = __shfl(xx, 0), xmax = __shfl(xx, last);
                for (uint32_t j = 0; j < ncache0; ++j) {
                    if (__ballot(ok) == 0) break;
                    const uint64_t rlo = S.m_olo[j][k0];
                    const uint64_t rlen = S.m_olen[j][k0];
                    if (rlen < coop_row) {
                        // a short row: each lane's binary search is a few
                        // cached probes, cheaper than locating a window
                        if (ok) ok = row_has(idx, rlo, rlen, xx, bytes);
                        continue;
                    }
                    const uint64_t wlo = wave_lower_bound(idx, rlo, rlen, xmin, bytes);
                    const uint64_t whi = xmax == 0xFFFFFFFFu ? rlo + rlen
                                                              : wave_lower_bound(idx, wlo, rlo + rlen - wlo, xmax + 1u, bytes);
                    const uint64_t wl = whi - wlo;
                    if (wl == 0) {
                        ok = false;
                    } else if (wl <= kStage) {
                        for (uint32_t i = lane; i < (uint32_t)wl; i += kWave) S.stage[i] = idx[wlo + i];
                        bytes += 4ull * (lane < wl ? (wl - lane + kWave - 1) / kWave : 0);
                        __syncthreads();
                        if (ok) ok = stage_has(S.stage, (uint32_t)wl, xx);
                        __syncthreads();
                    } else if (ok) {
                        ok = row_has(idx, wlo, wl, xx, bytes);
                    }
                }
            } else if (act) {
                for (uint32_t j = 0; j < ncache && ok; ++j)
                    ok = row_has(idx, S.m_olo[j][k], S.m_olen[j][k], xx, bytes);
            }
            if (act) {
                if (ok && nother > kWCache) {
                    uint32_t slot2 = 0;
                    for (uint32_t j = 0; j < nops && ok; ++j) {
                        if (j == sel) continue;
                        if (slot2++ < kWCache) continue;
                        const DOp op = ops[P.op_begin[g] + j];
                        const uint64_t r = op.row_base + op_source(S, op, g, t, S.m_x[k]);
                        const uint64_t a = row_off[r], b = row_off[r + 1];
                        bytes += 16;
                        ok = row_has(idx, a, b - a, xx, bytes);
                    }
                }
                if (ok && P.same_label[g]) {
                    const uint32_t v = cand_pool[P.cand_base[g] + xx];
                    bytes += 4;
                    unsigned long long mask = P.same_label[g];
                    while (mask) {
                        const uint32_t e = __ffsll((long long)mask) - 1;
                        mask &= mask - 1;
                        const uint32_t me = (g > t && e == g - 1) ? S.m_v[k] : S.M[e];
                        if (me == v) {
                            ok = false;
                            rej = true;
                            break;
                        }
                    }
                }
            }
            const unsigned long long bal = __ballot(ok);
            const uint32_t ns = (uint32_t)__popcll(bal);
            const bool leaf = g + 1 == n;
            if (!leaf && ok) {
                const uint32_t p = (uint32_t)__popcll(bal & lanemask_lt());
                S.cx[g][p] = xx;
                S.ctag[g][p] = (uint8_t)k;
            }
            // advance the cursor past this chunk
            flat += min<uint64_t>(kWave, total - flat);
            {
                const bool here = lane < cnt && S.m_pre[lane] <= flat && flat < S.m_pre[lane] + S.m_len[lane];
                const unsigned long long hb = __ballot(here);
                if (here) {
                    S.gen_m[g] = lane;
                    S.gen_pos[g] = S.m_lo[lane] + (flat - S.m_pre[lane]);
                    S.gen_end[g] = S.m_lo[lane] + S.m_len[lane];
                    S.gen_sel[g] = S.m_sel[lane];
                }
                if (!hb && lane == 0) S.gen_m[g] = cnt;
            }
            lcalls += ns;
            lrej += (uint32_t)__popcll(__ballot(rej));
            if (!leaf && g >= 1 && lane == 0) algo += 4ull * ns;   // survivors written below the root level
            if (leaf) {
                lcount += ns;
            } else if (ns) {
                if (lane == 0) {
                    S.ccount[g] = ns;
                    S.run_lo[g] = 0;
                }
                __syncthreads();
                const uint32_t len = run_length(S, g, 0);
                const uint32_t tag = S.ctag[g][0];
                if (lane == 0) {
                    S.run_hi[g] = len;
                    if (g > t) {
                        S.I[g - 1] = S.m_x[tag];
                        S.M[g - 1] = S.m_v[tag];
                    }
                    S.gen_m[g + 1] = 0;
                    S.gen_pos[g + 1] = kFresh;
                    S.counted[g + 1] = 0;
                }
                ++g;
                cache_g = -1;
            }
            __syncthreads();

            ++it;
            if ((it & 15) == 0 || lcount >= (P.cap != ~0ull ? min(65536ull, max(64ull, P.cap >> 10)) : 65536ull)) {
                uint32_t stop = 0;
                if (lane == 0) {
                    if (lcount | lcalls) {
                        const unsigned long long prev = atomicAdd(&qc[q].count, lcount);
                        atomicAdd(&qc[q].calls, lcalls);
                        if (lcount && prev + lcount >= P.cap) {
                            atomicExch(&qc[q].done, 1u);
                            mark_final(qc, q, P.cap, notify);
                        }
                    }
                    if (lrej) atomicAdd(&qc[q].rejects, lrej);
                    stop = __atomic_load_n(&qc[q].done, __ATOMIC_RELAXED) || load_relaxed(&Q->timeout);
                    if (!stop && past_deadline(Q)) {
                        atomicExch(&Q->timeout, 1u);
                        stop = 1;
                    }
                    if (!stop && passed(qdl)) {
                        // the query's own time limit: post what it has counted
                        post_timeout(notify, qc, q);
                        stop = 1;
                    }
                }
                lcount = lcalls = lrej = 0;
                if (__shfl(stop, 0)) break;
            }
            // split off work while the queue is short
            if ((it & 7) == 0) {
                uint32_t hungry = 0;
                if (lane == 0) {
                    const uint32_t h0 = load_relaxed(&Q->head0);
                    const uint32_t left0 = segs ? segment_left(Q, segs, seg_head) : (h0 < n0 ? n0 - h0 : 0u);
                    hungry = load_relaxed(&Q->tail) - load_relaxed(&Q->head) + left0 < hungry_below;
                }
                if (__shfl(hungry, 0)) {
                    const uint32_t e = t + lane;
                    bool can = false;
                    if (e <= g && e <= kPrefixMax && (e < g || cache_g == (int)g)) {
                        can = S.gen_m[e] < S.gen_cnt[e] && S.gen_pos[e] != kFresh &&
                              S.gen_end[e] - S.gen_pos[e] >= 2;
                    }
                    const unsigned long long cb = __ballot(can);
                    if (cb) {
                        const uint32_t ed = t + (uint32_t)(__ffsll((long long)cb) - 1);
                        uint32_t ns2 = kInvalid;
                        if (lane == 0) {
                            const uint32_t tl = load_relaxed(&Q->tail);
                            if (seq_load(seq, qmask, tl) != tl) atomicAdd(&Q->full, 1u);
                            else if (atomicCAS(&Q->tail, tl, tl + 1) == tl) ns2 = tl;
                        }
                        ns2 = __shfl(ns2, 0);
                        if (ns2 != kInvalid) {
                            const uint64_t gp = S.gen_pos[ed], ge = S.gen_end[ed];
                            const uint64_t mid = gp + (ge - gp) / 2;
                            DTask& nt = queue[ns2 & qmask];
                            if (lane < ed) {
                                const bool member = ed > t && lane == ed - 1;
                                nt.prefix[lane] =
                                    member ? S.cx[ed - 1][S.run_lo[ed - 1] + S.gen_m[ed]] : S.I[lane];
                            }
                            if (lane == 0) {
                                nt.q = q;
                                nt.depth = ed;
                                nt.sel = S.gen_sel[ed];
                                nt.pad = 0;
                                nt.lo = mid;
                                nt.hi = ge;
                                atomicAdd(pending + q, 1u);
                                atomicAdd(&Q->donated, 1u);
                            }
                            record_publish_fence();
                            __syncthreads();
                            if (lane == 0) {
                                seq_publish(seq, qmask, ns2);
                                S.gen_end[ed] = mid;
                            }
                            __syncthreads();
                            if ((int)ed == cache_g) cache_g = -1;
                        }
                    }
                }
            }
        }
        // task ended (finished, or stopped by the cap / time limit)
        if (lane == 0) {
            if (lcount | lcalls) {
                const unsigned long long prev = atomicAdd(&qc[q].count, lcount);
                atomicAdd(&qc[q].calls, lcalls);
                if (lcount && prev + lcount >= P.cap) {
                    atomicExch(&qc[q].done, 1u);
                    mark_final(qc, q, P.cap, notify);
                }
            }
            if (lrej) atomicAdd(&qc[q].rejects, lrej);
            if (finished) pending_sub(pending, qc, q, 1u, notify);
            atomicSub(&Q->busy, 1u);
        }
        wave_add(&qc[q].probe_bytes, bytes);
        wave_add(&qc[q].algo_bytes, algo);
        bytes = algo = 0;
    }
    wave_add(&Q->lane_steps, busy_pos);
    wave_add(&Q->lane_iters, all_pos);
}

// ---------------------------------------------------------------------------
// Two-unit hash join (JoinCollect.cpp:60-153 + hashjoin.cpp:337-781 restated):
// rows of the build unit are chained under a 64-bit fingerprint of their key
// columns in an open-addressing table; each probe row walks the chain of its
// fingerprint, checks the key columns exactly and counts the build rows whose
// non-key data vertices are disjoint from its own (injectivity across units).
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long key_fingerprint(const uint32_t* row, const JoinKeys& k) {
    unsigned long long h = 0x9E3779B97F4A7C15ull ^ k.nkeys;
    for (uint32_t i = 0; i < k.nkeys; ++i) {
        h ^= row[k.key[i]] + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
        h *= 0xBF58476D1CE4E5B9ull;
        h ^= h >> 31;
    }
    return h | 1ull;   // 0 marks an empty slot
}

// A probe row matches build row `br` when the key columns agree and no
// non-key data vertex of one equals a non-key data vertex of the other.
__device__ __forceinline__ bool join_match(const uint32_t* br, const JoinKeys& bk, const uint32_t* pr,
                                           const JoinKeys& pk) {
    for (uint32_t i = 0; i < pk.nkeys; ++i)
        if (br[bk.key[i]] != pr[pk.key[i]]) return false;
    for (uint32_t i = 0; i < bk.nrest; ++i)
        for (uint32_t j = 0; j < pk.nrest; ++j)
            if (br[bk.rest[i]] == pr[pk.rest[j]]) return false;
    return true;
}

__global__ __launch_bounds__(kBlock) void k_join_build(const uint32_t* __restrict__ rows, uint64_t n_rows,
                                                        uint32_t width, JoinKeys keys,
                                                        unsigned long long* fp, uint32_t* head,
                                                        uint64_t mask, uint32_t* next) {
    for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < n_rows;
         r += (uint64_t)gridDim.x * kBlock) {
        const unsigned long long h = key_fingerprint(rows + r * width, keys);
        uint64_t slot = h & mask;
        while (true) {
            const unsigned long long prev = atomicCAS(fp + slot, 0ull, h);
            if (prev == 0ull || prev == h) break;
            slot = (slot + 1) & mask;
        }
        next[r] = atomicExch(head + slot, (uint32_t)r);
    }
}

__global__ __launch_bounds__(kBlock) void k_join_probe(
    const uint32_t* __restrict__ build_rows, uint32_t build_width, JoinKeys bk,
    const uint32_t* __restrict__ probe_rows, uint64_t n_probe, uint32_t probe_width, JoinKeys pk,
    const unsigned long long* __restrict__ fp, const uint32_t* __restrict__ head, uint64_t mask,
    const uint32_t* __restrict__ next, unsigned long long* count, unsigned long long* bytes) {
    unsigned long long c = 0, by = 0;
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    const uint64_t rounded = (n_probe + stride - 1) / stride * stride;
    for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < rounded; r += stride) {
        if (r >= n_probe) continue;
        const uint32_t* pr = probe_rows + r * probe_width;
        const unsigned long long h = key_fingerprint(pr, pk);
        by += 4ull * probe_width;
        uint64_t slot = h & mask;
        uint32_t a = kInvalid;
        while (true) {
            const unsigned long long f = fp[slot];
            by += 8;
            if (f == 0ull) break;
            if (f == h) { a = head[slot]; break; }
            slot = (slot + 1) & mask;
        }
        for (; a != kInvalid; a = next[a]) {
            const uint32_t* br = build_rows + (uint64_t)a * build_width;
            by += 4ull * build_width + 4;
            c += join_match(br, bk, pr, pk);
        }
    }
    wave_add(count, c);
    wave_add(bytes, by);
}

// Materialising probe of a multi-unit join (JoinCollect::concateTwoTabels,
// JoinCollect.cpp:237-315, restated): each output row is the probe row
// followed by the build row's non-key columns. A thread counts its probe
// row's matches, the wavefront reserves room for all of its matches with one
// atomic, and the thread walks the chain again to write them; rows past `cap`
// are dropped and flag `overflow`.
__global__ __launch_bounds__(kBlock) void k_join_probe_emit(
    const uint32_t* __restrict__ build_rows, uint32_t build_width, JoinKeys bk,
    const uint32_t* __restrict__ probe_rows, uint64_t n_probe, uint32_t probe_width, JoinKeys pk,
    const unsigned long long* __restrict__ fp, const uint32_t* __restrict__ head, uint64_t mask,
    const uint32_t* __restrict__ next, uint32_t* __restrict__ out, unsigned long long cap,
    unsigned long long* count, uint32_t* overflow) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    const uint64_t rounded = (n_probe + stride - 1) / stride * stride;
    const uint32_t out_width = probe_width + bk.nrest;
    const uint32_t lane = lane_id();
    for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < rounded; r += stride) {
        const bool live = r < n_probe;
        const uint32_t* pr = probe_rows + (live ? r : 0) * probe_width;
        uint32_t a0 = kInvalid;
        unsigned long long m = 0;
        if (live) {
            const unsigned long long h = key_fingerprint(pr, pk);
            uint64_t slot = h & mask;
            while (true) {
                const unsigned long long f = fp[slot];
                if (f == 0ull) break;
                if (f == h) { a0 = head[slot]; break; }
                slot = (slot + 1) & mask;
            }
            for (uint32_t a = a0; a != kInvalid; a = next[a])
                m += join_match(build_rows + (uint64_t)a * build_width, bk, pr, pk);
        }
        // inclusive scan of m over the wavefront; lane 63 holds the total
        unsigned long long inc = m;
        for (uint32_t o = 1; o < kWave; o <<= 1) {
            const uint32_t lo = __shfl_up((uint32_t)inc, o), hi = __shfl_up((uint32_t)(inc >> 32), o);
            if (lane >= o) inc += ((unsigned long long)hi << 32) | lo;
        }
        const unsigned long long total =
            ((unsigned long long)__shfl((uint32_t)(inc >> 32), kWave - 1) << 32) | __shfl((uint32_t)inc, kWave - 1);
        unsigned long long base = 0;
        if (lane == 0 && total) base = atomicAdd(count, total);
        base = ((unsigned long long)__shfl((uint32_t)(base >> 32), 0) << 32) | __shfl((uint32_t)base, 0);
        unsigned long long pos = base + inc - m;
        if (m) {
            for (uint32_t a = a0; a != kInvalid; a = next[a]) {
                const uint32_t* br = build_rows + (uint64_t)a * build_width;
                if (!join_match(br, bk, pr, pk)) continue;
                if (pos < cap) {
                    uint32_t* o = out + pos * out_width;
                    for (uint32_t c = 0; c < probe_width; ++c) o[c] = pr[c];
                    for (uint32_t c = 0; c < bk.nrest; ++c) o[probe_width + c] = br[bk.rest[c]];
                } else {
                    *overflow = 1u;
                }
                ++pos;
            }
        }
    }
}

// Root range tasks of query q: task i enumerates positions
// [i * len / n, (i + 1) * len / n) of C(order[0]) at depth 0.
__global__ void k_root_tasks(DTask* tasks, uint32_t n, uint32_t q, uint64_t len) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    DTask t;
    t.q = q;
    t.depth = 0;
    t.sel = kIdentity;
    t.pad = 0;
    t.lo = (uint64_t)i * len / n;
    t.hi = (uint64_t)(i + 1) * len / n;
    for (uint32_t e = 0; e < kPrefixMax; ++e) t.prefix[e] = 0;
    tasks[i] = t;
}

__global__ void k_ring_init(uint32_t* seq, uint32_t slots, uint32_t n_tasks) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < slots; i += (uint64_t)gridDim.x * blockDim.x)
        seq[i] = (uint32_t)i + (i < n_tasks ? 1u : 0u);
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock)
        p[i] = v;
}

// ---------------------------------------------------------------------------
// Exclusive scans (64-bit totals) and per-slot bitmap word ranks.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long shfl_up64(unsigned long long x, uint32_t o) {
    const uint32_t lo = __shfl_up((uint32_t)x, o), hi = __shfl_up((uint32_t)(x >> 32), o);
    return ((unsigned long long)hi << 32) | lo;
}

// Block-wide exclusive scan of one value per thread; `total` gets the sum.
__device__ __forceinline__ unsigned long long block_scan(unsigned long long x,
                                                         unsigned long long& total) {
    __shared__ unsigned long long wsum[kBlock / kWave];
    const uint32_t lane = lane_id(), w = threadIdx.x / kWave;
    unsigned long long inc = x;
    for (uint32_t o = 1; o < kWave; o <<= 1) {
        const unsigned long long y = shfl_up64(inc, o);
        if (lane >= o) inc += y;
    }
    if (lane == kWave - 1) wsum[w] = inc;
    __syncthreads();
    unsigned long long base = 0;
    total = 0;
    for (uint32_t i = 0; i < kBlock / kWave; ++i) {
        if (i < w) base += wsum[i];
        total += wsum[i];
    }
    __syncthreads();
    return base + inc - x;
}

constexpr uint32_t kScanPerThread = 8;
constexpr uint32_t kScanChunk = kBlock * kScanPerThread;

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scan_reduce(const T* __restrict__ in, uint64_t n,
                                                        unsigned long long* __restrict__ sums) {
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
    unsigned long long x = 0;
    for (uint32_t i = 0; i < kScanPerThread; ++i) {
        const uint64_t p = base + (uint64_t)i * kBlock + threadIdx.x;
        if (p < n) x += in[p];
    }
    unsigned long long total;
    block_scan(x, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// One block: exclusive scan of the chunk sums in place.
__global__ __launch_bounds__(kBlock) void k_scan_sums(unsigned long long* __restrict__ sums,
                                                      uint32_t n_chunks) {
    unsigned long long run = 0;
    for (uint32_t b = 0; b < n_chunks; b += kBlock) {
        const uint32_t i = b + threadIdx.x;
        const unsigned long long x = i < n_chunks ? sums[i] : 0;
        unsigned long long total;
        const unsigned long long ex = block_scan(x, total);
        if (i < n_chunks) sums[i] = run + ex;
        run += total;
    }
}

// out[i] = sum of in[0..i), for i in [0, n]; out has n+1 entries.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_scan_down(const T* __restrict__ in, uint64_t n,
                                                      const unsigned long long* __restrict__ sums,
                                                      uint64_t* __restrict__ out) {
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk + (uint64_t)threadIdx.x * kScanPerThread;
    unsigned long long v[kScanPerThread];
    unsigned long long x = 0;
    for (uint32_t i = 0; i < kScanPerThread; ++i) {
        const uint64_t p = base + i;
        v[i] = p < n ? (unsigned long long)in[p] : 0ull;
        x += v[i];
    }
    unsigned long long total;
    unsigned long long run = sums[blockIdx.x] + block_scan(x, total);
    for (uint32_t i = 0; i < kScanPerThread; ++i) {
        const uint64_t p = base + i;
        if (p < n) out[p] = run;
        if (p == n - 1) out[n] = run + v[i];
        run += v[i];
    }
}

// Per slot: prefix[slot][w] = set bits of words [0, w) of that slot's bitmap.
__global__ __launch_bounds__(kBlock) void k_slot_prefix(const uint32_t* __restrict__ bitmaps,
                                                        uint32_t* __restrict__ prefix, uint64_t wps) {
    const uint32_t* bm = bitmaps + (size_t)blockIdx.x * wps;
    uint32_t* out = prefix + (size_t)blockIdx.x * wps;
    unsigned long long run = 0;
    for (uint64_t b = 0; b < wps; b += kBlock) {
        const uint64_t i = b + threadIdx.x;
        const unsigned long long x = i < wps ? __popc(bm[i]) : 0u;
        unsigned long long total;
        const unsigned long long ex = block_scan(x, total);
        if (i < wps) out[i] = (uint32_t)(run + ex);
        run += total;
    }
}

// out[i] = row_off[end_i] - row_off[start_i] for item i.
__global__ __launch_bounds__(kBlock) void k_span(const uint64_t* __restrict__ row_off,
                                                 const uint64_t* __restrict__ starts,
                                                 const uint64_t* __restrict__ ends, uint32_t n,
                                                 uint64_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) out[i] = row_off[ends[i]] - row_off[starts[i]];
}

}  // namespace

template <typename T>
static hipError_t scan_impl(const T* in, uint64_t n, uint64_t* out, unsigned long long* tmp,
                            hipStream_t s) {
    if (n == 0) return hipMemsetAsync(out, 0, sizeof(uint64_t), s);
    const uint64_t chunks = (n + kScanChunk - 1) / kScanChunk;
    hipLaunchKernelGGL(k_scan_reduce<T>, dim3((uint32_t)chunks), dim3(kBlock), 0, s, in, n, tmp);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, s, tmp, (uint32_t)chunks);
    hipLaunchKernelGGL(k_scan_down<T>, dim3((uint32_t)chunks), dim3(kBlock), 0, s, in, n, tmp, out);
    return hipGetLastError();
}

uint64_t scan_tmp_entries(uint64_t n) { return (n + kScanChunk - 1) / kScanChunk + 1; }

hipError_t launch_scan_u32(const uint32_t* in, uint64_t n, uint64_t* out,
                           unsigned long long* tmp, hipStream_t s) {
    return scan_impl(in, n, out, tmp, s);
}

hipError_t launch_scan_u64(const uint64_t* in, uint64_t n, uint64_t* out,
                           unsigned long long* tmp, hipStream_t s) {
    return scan_impl(in, n, out, tmp, s);
}

hipError_t launch_slot_prefix(const uint32_t* bitmaps, uint32_t* prefix, uint32_t n_slots,
                              uint64_t words_per_slot, hipStream_t s) {
    if (n_slots == 0) return hipSuccess;
    hipLaunchKernelGGL(k_slot_prefix, dim3(n_slots), dim3(kBlock), 0, s, bitmaps, prefix,
                       words_per_slot);
    return hipGetLastError();
}

hipError_t launch_span(const uint64_t* row_off, const uint64_t* starts, const uint64_t* ends,
                       uint32_t n, uint64_t* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_span, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, row_off, starts,
                       ends, n, out);
    return hipGetLastError();
}

namespace {
}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_compact_flag(const DevGraph& g, const CItem* items, uint32_t n_items,
                               uint32_t n_blocks, const QVert* qv, const uint2* qnlf,
                               const uint32_t* src_pool, const uint32_t* cnt_pool,
                               unsigned long long* masks, uint32_t* block_counts, hipStream_t s) {
    if (n_blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_compact_flag, dim3(n_blocks), dim3(kBlock), 0, s, g, items, n_items, qv,
                       qnlf, src_pool, cnt_pool, masks, block_counts);
    return hipGetLastError();
}

hipError_t launch_compact_scatter(const CItem* items, uint32_t n_items, uint32_t n_blocks,
                                  const uint32_t* src_ri, const uint32_t* src_pool,
                                  const unsigned long long* masks, const uint64_t* block_prefix,
                                  uint32_t* out, hipStream_t s) {
    if (n_blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_compact_scatter, dim3(n_blocks), dim3(kBlock), 0, s, items, n_items,
                       src_ri, src_pool, masks, block_prefix, out);
    return hipGetLastError();
}

hipError_t launch_bucket_flag(const DevGraph& g, const CGroup* groups, uint32_t n_groups, uint32_t n_waves,
                              const CBucketItem* bitems, const uint2* qnlf, unsigned long long* masks,
                              uint32_t* chunk_counts, hipStream_t s) {
    if (n_waves == 0 || n_groups == 0) return hipSuccess;
    const uint32_t waves_per_block = kBlock / kWave;
    hipLaunchKernelGGL(k_bucket_flag, dim3((n_waves + waves_per_block - 1) / waves_per_block), dim3(kBlock), 0, s,
                       g, groups, n_groups, n_waves, bitems, qnlf, masks, chunk_counts);
    return hipGetLastError();
}

hipError_t launch_bucket_scatter(const DevGraph& g, const CGroup* groups, uint32_t n_groups, uint32_t n_waves,
                                 const CBucketItem* bitems, const unsigned long long* masks,
                                 const uint64_t* chunk_prefix, uint32_t* out, hipStream_t s) {
    if (n_waves == 0 || n_groups == 0) return hipSuccess;
    const uint32_t waves_per_block = kBlock / kWave;
    hipLaunchKernelGGL(k_bucket_scatter, dim3((n_waves + waves_per_block - 1) / waves_per_block), dim3(kBlock), 0,
                       s, g, groups, n_groups, n_waves, bitems, masks, chunk_prefix, out);
    return hipGetLastError();
}

hipError_t launch_set_bits(const uint32_t* pool, const uint64_t* seg_begin, const uint32_t* seg_len,
                           const uint32_t* seg_slot, uint32_t n_seg, uint32_t total_max,
                           uint32_t* bitmaps, uint64_t words_per_slot, hipStream_t s) {
    if (n_seg == 0) return hipSuccess;
    uint32_t gx = (total_max + kBlock - 1) / kBlock;
    gx = gx < 1 ? 1 : (gx > 64 ? 64 : gx);
    hipLaunchKernelGGL(k_set_bits, dim3(gx, n_seg), dim3(kBlock), 0, s, pool, seg_begin, seg_len,
                       seg_slot, n_seg, bitmaps, words_per_slot);
    return hipGetLastError();
}

hipError_t launch_set_member(const uint32_t* pool, const uint64_t* seg_begin, const uint32_t* seg_len,
                             const uint64_t* seg_mem, const uint32_t* seg_bit, uint32_t n_seg,
                             uint32_t total_max, unsigned long long* member, hipStream_t s) {
    if (n_seg == 0 || total_max == 0) return hipSuccess;
    const uint32_t gx = min((total_max + kBlock - 1) / kBlock, 64u);
    hipLaunchKernelGGL(k_set_member, dim3(gx, n_seg), dim3(kBlock), 0, s, pool, seg_begin, seg_len, seg_mem,
                       seg_bit, n_seg, member);
    return hipGetLastError();
}

hipError_t launch_popc_words(const uint32_t* bitmaps, uint32_t* out, uint64_t n_words, hipStream_t s) {
    if (n_words == 0) return hipSuccess;
    uint64_t blocks = (n_words + kBlock - 1) / kBlock;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_popc_words, dim3((uint32_t)blocks), dim3(kBlock), 0, s, bitmaps, out, n_words);
    return hipGetLastError();
}

// Grids of the refinement kernels: fixed, strided over their work, so that
// launches that read their counts from device memory need no host sync.
constexpr uint32_t kGqlGrid = 2048;

hipError_t launch_gql_refine(const DevGraph& g, const GItem* items, const uint2* list, const uint32_t* n_list,
                             uint32_t* cursor, uint32_t max_n, const uint32_t* left_slots, const uint32_t* left_lab,
                             const uint32_t* left_labs, uint32_t* cand_pool, uint32_t* bitmaps,
                             uint64_t words_per_slot, unsigned long long* member, GqlRound rd,
                             unsigned long long* bytes, hipStream_t s) {
    if (max_n == 0) return hipSuccess;
    // six blocks per CU fit the LDS of gfx950's 256 CUs (26 KB per block)
    const uint32_t blocks = std::min<uint32_t>(1536, (max_n + kGqlWavesPerBlock - 1) / kGqlWavesPerBlock);
    hipLaunchKernelGGL(k_gql_refine, dim3(blocks), dim3(kBlock), 0, s, g, items, list, n_list, cursor, left_slots,
                       left_lab, left_labs, cand_pool, bitmaps, words_per_slot, member, rd, bytes);
    return hipGetLastError();
}

hipError_t launch_gql_skip(const GItem* items, uint32_t n_items, const uint32_t* left_slots,
                           const uint32_t* removed1, const uint32_t* removed2, uint32_t* skip, hipStream_t s) {
    if (n_items == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gql_skip, dim3((n_items + kBlock - 1) / kBlock), dim3(kBlock), 0, s, items, n_items,
                       left_slots, removed1, removed2, skip);
    return hipGetLastError();
}

hipError_t launch_gql_refine_small(const DevGraph& g, const GItem* items, uint32_t n_items,
                                   uint32_t n_cands, const uint32_t* left_slots, const uint32_t* left_labs,
                                   uint32_t* cand_pool,
                                   uint32_t* bitmaps, uint64_t words_per_slot, unsigned long long* member,
                                   uint2* lists, uint64_t list_stride, uint32_t* n_lists, GqlRound rd,
                                   unsigned long long* bytes, hipStream_t s) {
    if (n_cands == 0) return hipSuccess;
    if (list_stride < n_cands) return hipErrorInvalidValue;
    const uint64_t want = ((uint64_t)n_cands * kGroup + kBlock - 1) / kBlock;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(want, kGqlGrid);
    hipLaunchKernelGGL(k_gql_refine_small, dim3(blocks), dim3(kBlock), 0, s, g, items, n_items, n_cands,
                       left_slots, left_labs, cand_pool, bitmaps, words_per_slot, member, lists, list_stride, n_lists,
                       rd, bytes);
    return hipGetLastError();
}

hipError_t launch_gql_hall(uint32_t lanes, const DevGraph& g, const GItem* items, const uint2* list,
                           const uint32_t* n_list, uint32_t max_n, const uint32_t* left_slots,
                           const uint32_t* left_labs, uint32_t* cand_pool, uint32_t* bitmaps, uint64_t words_per_slot,
                           unsigned long long* member, GqlRound rd, unsigned long long* bytes, hipStream_t s) {
    if (max_n == 0) return hipSuccess;
    const uint64_t want = ((uint64_t)max_n * lanes + kBlock - 1) / kBlock;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(want, kGqlGrid);
    if (lanes == 16)
        hipLaunchKernelGGL(k_gql_hall<16>, dim3(blocks), dim3(kBlock), 0, s, g, items, list, n_list, left_slots,
                           left_labs, cand_pool, bitmaps, words_per_slot, member, rd, bytes);
    else if (lanes == 32)
        hipLaunchKernelGGL(k_gql_hall<32>, dim3(blocks), dim3(kBlock), 0, s, g, items, list, n_list, left_slots,
                           left_labs, cand_pool, bitmaps, words_per_slot, member, rd, bytes);
    else if (lanes == 64)
        hipLaunchKernelGGL(k_gql_hall<64>, dim3(blocks), dim3(kBlock), 0, s, g, items, list, n_list, left_slots,
                           left_labs, cand_pool, bitmaps, words_per_slot, member, rd, bytes);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

static uint32_t grid_for(uint64_t n) {
    uint64_t b = (n + kBlock - 1) / kBlock;
    if (b > 131072) b = 131072;
    return b ? (uint32_t)b : 1u;
}

hipError_t launch_table_count(const DevGraph& g, const TItem* items, uint32_t n_items,
                              const uint64_t* item_row_start, uint64_t n_rows,
                              const uint32_t* cand_pool, const uint32_t* bitmaps,
                              uint64_t words_per_slot, uint64_t* row_deg, uint64_t* long_rows,
                              unsigned long long* n_long, unsigned long long* bytes, hipStream_t s) {
    if (n_rows == 0) return hipSuccess;
    hipLaunchKernelGGL(k_table_count, dim3(grid_for(n_rows)), dim3(kBlock), 0, s, g, items, n_items,
                       item_row_start, n_rows, cand_pool, bitmaps, words_per_slot, row_deg, long_rows, n_long,
                       bytes);
    return hipGetLastError();
}

hipError_t launch_table_long(int fill, const DevGraph& g, const TItem* items, uint32_t n_items,
                             const uint64_t* item_row_start, const uint64_t* long_rows, uint64_t n_long,
                             const uint32_t* cand_pool, const uint32_t* bitmaps, const uint32_t* word_prefix,
                             uint64_t words_per_slot, uint64_t* row_deg, const uint64_t* row_off, uint32_t* idx,
                             unsigned long long* bytes, hipStream_t s) {
    if (n_long == 0) return hipSuccess;
    const uint64_t want = (n_long + (kBlock / kWave) - 1) / (kBlock / kWave);
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(want, 65536);
    if (fill)
        hipLaunchKernelGGL(k_table_fill_long, dim3(blocks), dim3(kBlock), 0, s, g, items, n_items, item_row_start,
                           long_rows, n_long, cand_pool, bitmaps, word_prefix, words_per_slot, row_off, idx, bytes);
    else
        hipLaunchKernelGGL(k_table_count_long, dim3(blocks), dim3(kBlock), 0, s, g, items, n_items, item_row_start,
                           long_rows, n_long, cand_pool, bitmaps, words_per_slot, row_deg, bytes);
    return hipGetLastError();
}

hipError_t launch_table_fill(const DevGraph& g, const TItem* items, uint32_t n_items,
                             const uint64_t* item_row_start, uint64_t n_rows,
                             const uint32_t* cand_pool, const uint32_t* bitmaps,
                             const uint32_t* word_prefix, uint64_t words_per_slot,
                             const uint64_t* row_off, uint32_t* idx, unsigned long long* bytes,
                             hipStream_t s) {
    if (n_rows == 0) return hipSuccess;
    hipLaunchKernelGGL(k_table_fill, dim3(grid_for(n_rows)), dim3(kBlock), 0, s, g, items, n_items,
                       item_row_start, n_rows, cand_pool, bitmaps, word_prefix, words_per_slot,
                       row_off, idx, bytes);
    return hipGetLastError();
}

hipError_t launch_cfl_support(const DevGraph& g, const uint32_t* cand_pool, uint64_t cand_begin,
                              uint32_t len, uint32_t label, uint32_t degree, uint32_t* cnt,
                              uint32_t expect, uint32_t next, uint32_t* defer, uint32_t* n_defer,
                              hipStream_t s) {
    if (len == 0) return hipSuccess;
    hipError_t e = hipMemsetAsync(n_defer, 0, 4, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_cfl_support, dim3(grid_for(len)), dim3(kBlock), 0, s, g, cand_pool,
                       cand_begin, len, label, degree, cnt, expect, next, defer, n_defer);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint32_t blocks = std::min<uint32_t>((len + (kBlock / kWave) - 1) / (kBlock / kWave), 1024);
    hipLaunchKernelGGL(k_cfl_support_long, dim3(blocks), dim3(kBlock), 0, s, g, defer, n_defer, label, degree,
                       cnt, expect, next);
    return hipGetLastError();
}

hipError_t launch_cfl_prune(const uint32_t* cnt, uint32_t* cand_pool, uint64_t cand_begin,
                            uint32_t len, uint32_t target, hipStream_t s) {
    if (len == 0) return hipSuccess;
    hipLaunchKernelGGL(k_cfl_prune, dim3(grid_for(len)), dim3(kBlock), 0, s, cnt, cand_pool,
                       cand_begin, len, target);
    return hipGetLastError();
}

hipError_t launch_expand(const DPlan* plans, const DOp* ops, const DTask* tasks, uint32_t n_tasks, const uint8_t* expand_q, uint32_t* out_count,
                         const uint64_t* out_pos, DTask* out_tasks, QCounters* qc,
                         unsigned long long* qchild, uint32_t nq, uint32_t max_depth, int write,
                         hipStream_t s) {
    if (n_tasks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_expand, dim3((n_tasks + kBlock - 1) / kBlock), dim3(kBlock), 0, s, plans,
                       ops, tasks, n_tasks, expand_q, out_count, out_pos,
                       out_tasks, qc, qchild, nq, max_depth, write);
    return hipGetLastError();
}

hipError_t launch_gather_tasks(const DTask* src, const uint64_t* src_off, const uint64_t* dst_off,
                               uint32_t n_seg, uint64_t total, DTask* dst, hipStream_t s) {
    if (n_seg == 0 || total == 0) return hipSuccess;
    const uint64_t blocks = std::min<uint64_t>((total + kBlock - 1) / kBlock, 4096);
    hipLaunchKernelGGL(k_gather_tasks, dim3((uint32_t)blocks), dim3(kBlock), 0, s, src, src_off, dst_off,
                       n_seg, dst);
    return hipGetLastError();
}

size_t lane_lds_bytes(uint32_t lanes, uint32_t max_n) {
    return ((size_t)lanes * max_n * kLaneStateBytes + 15) & ~(size_t)15;
}

// Lanes per block of the lane DFS: want (64 / 128 / 256; 0: 64) or less,
// halved until the block's LDS state fits the device's per-block limit
// (gfx950: 160 KB). One-wavefront blocks let the dispatcher spread a grid of
// one wavefront per CU over every CU (four-wavefront blocks of the same grid
// fill a quarter of them); a CU then holds as many of them as its LDS allows.
uint32_t dfs_lanes_per_block(uint32_t max_n, uint32_t want) {
    int dev = 0, lim = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || lim <= 0)
        lim = 64 * 1024;
    uint32_t lanes = want ? std::min<uint32_t>(want, kBlock) : kWave;
    while (lanes > kWave && lane_lds_bytes(lanes, max_n) > (size_t)lim) lanes /= 2;
    return lanes;
}

size_t dfs_lane_lds_bytes(uint32_t max_n) { return lane_lds_bytes(dfs_lanes_per_block(max_n, 0), max_n); }

hipError_t launch_dfs(const DPlan* plans, const DOp* ops, const DTask* init, DTask* queue, uint32_t* seq,
                      DQueue* dq, QCounters* qc, uint32_t* pending, uint32_t max_n, uint32_t grid_waves,
                      uint32_t block_lanes, EmitSink sink, hipStream_t s) {
    if (grid_waves == 0) return hipSuccess;
    if (max_n == 0 || max_n > kMaxQ) return hipErrorInvalidValue;
    const uint32_t lanes = dfs_lanes_per_block(max_n, block_lanes);
    const size_t lds = lane_lds_bytes(lanes, max_n);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_dfs),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    uint32_t blocks = std::max<uint32_t>(1, grid_waves / (lanes / kWave));
    // a persistent grid larger than what the CUs hold at once only adds
    // blocks that start after the work has drained: keep it resident
    int per_cu = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&k_dfs), (int)lanes,
                                                     lds) == hipSuccess &&
        hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && per_cu > 0 &&
        cus > 0)
        blocks = std::min<uint32_t>(blocks, (uint32_t)(per_cu * cus));
    hipLaunchKernelGGL(k_dfs, dim3(blocks), dim3(lanes), lds, s, plans, ops, init, queue, seq, dq, qc, pending,
                       max_n, sink);
    return hipGetLastError();
}

size_t dfs_wave_lds_bytes(uint32_t max_n) {
    return max_n <= 16 ? sizeof(WaveFrame<16>) : max_n <= 32 ? sizeof(WaveFrame<32>) : sizeof(WaveFrame<64>);
}

hipError_t launch_dfs_wave(const DPlan* plans, const DOp* ops, const DTask* init, DTask* queue, uint32_t* seq,
                           DQueue* dq, QCounters* qc, uint32_t* pending, uint32_t max_n, uint32_t grid_waves,
                           hipStream_t s) {
    if (grid_waves == 0) return hipSuccess;
    if (max_n == 0 || max_n > kMaxQ) return hipErrorInvalidValue;
    if (max_n <= 16)
        hipLaunchKernelGGL(k_dfs_wave<16>, dim3(grid_waves), dim3(kWave), 0, s, plans, ops, init, queue, seq,
                           dq, qc, pending);
    else if (max_n <= 32)
        hipLaunchKernelGGL(k_dfs_wave<32>, dim3(grid_waves), dim3(kWave), 0, s, plans, ops, init, queue, seq,
                           dq, qc, pending);
    else
        hipLaunchKernelGGL(k_dfs_wave<64>, dim3(grid_waves), dim3(kWave), 0, s, plans, ops, init, queue, seq,
                           dq, qc, pending);
    return hipGetLastError();
}

hipError_t launch_join_build(const uint32_t* rows, uint64_t n_rows, uint32_t width, JoinKeys keys,
                             unsigned long long* fp, uint32_t* head, uint64_t mask, uint32_t* next,
                             hipStream_t s) {
    if (n_rows == 0) return hipSuccess;
    hipLaunchKernelGGL(k_join_build, dim3(grid_for(n_rows)), dim3(kBlock), 0, s, rows, n_rows, width, keys,
                       fp, head, mask, next);
    return hipGetLastError();
}

hipError_t launch_join_probe(const uint32_t* build_rows, uint32_t build_width, JoinKeys build_keys,
                             const uint32_t* probe_rows, uint64_t n_probe, uint32_t probe_width,
                             JoinKeys probe_keys, const unsigned long long* fp, const uint32_t* head,
                             uint64_t mask, const uint32_t* next, unsigned long long* count,
                             unsigned long long* bytes, hipStream_t s) {
    if (n_probe == 0) return hipSuccess;
    hipLaunchKernelGGL(k_join_probe, dim3(grid_for(n_probe)), dim3(kBlock), 0, s, build_rows, build_width,
                       build_keys, probe_rows, n_probe, probe_width, probe_keys, fp, head, mask, next, count,
                       bytes);
    return hipGetLastError();
}

hipError_t launch_join_probe_emit(const uint32_t* build_rows, uint32_t build_width, JoinKeys build_keys,
                                  const uint32_t* probe_rows, uint64_t n_probe, uint32_t probe_width,
                                  JoinKeys probe_keys, const unsigned long long* fp, const uint32_t* head,
                                  uint64_t mask, const uint32_t* next, uint32_t* out, unsigned long long cap,
                                  unsigned long long* count, uint32_t* overflow, hipStream_t s) {
    if (n_probe == 0) return hipSuccess;
    hipLaunchKernelGGL(k_join_probe_emit, dim3(grid_for(n_probe)), dim3(kBlock), 0, s, build_rows, build_width,
                       build_keys, probe_rows, n_probe, probe_width, probe_keys, fp, head, mask, next, out, cap,
                       count, overflow);
    return hipGetLastError();
}

hipError_t launch_root_tasks(DTask* tasks, uint32_t n, uint32_t q, uint64_t len, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_root_tasks, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, tasks, n, q, len);
    return hipGetLastError();
}

hipError_t launch_ring_init(uint32_t* seq, uint32_t slots, uint32_t n_tasks, hipStream_t s) {
    if (slots == 0) return hipSuccess;
    hipLaunchKernelGGL(k_ring_init, dim3(grid_for(slots)), dim3(kBlock), 0, s, seq, slots, n_tasks);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n)), dim3(kBlock), 0, s, p, v, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Fixed device scratch of the allocation-free entry point
// (psm_evaluate_noalloc): one array per GPU the library runs on, held until
// the process ends. Host code never passes `scratch` itself to a kernel (on
// the host its name is a stand-in, not a device address): psm_ctx_create
// resolves the array's address on each device once, with that device
// current, and every later use goes through that address.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_addr[kScratchMaxDevices] = {};
}  // namespace

hipError_t scratch_setup(int device) {
    if (device < 0 || device >= (int)kScratchMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_addr[device]) return hipSuccess;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    g_scratch_addr[device] = ptr;
    return hipSuccess;
}

void* scratch_address(int device) {
    if (device < 0 || device >= (int)kScratchMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_addr[device];
}

}  // namespace psm
