// gen.cpp — seeded synthetic data graphs and query graphs.
//
// Data graphs: Chung-Lu power law and R-MAT, with Zipf or uniform labels.
// Queries: connected vertex sets grown from a random start by random walk
// ("sparse": the walk's tree edges), by a degree-biased random walk or greedy
// densest growth ("dense": the induced subgraph), keeping the data vertices' labels, so every query has at
// least one embedding. Everything is deterministic for a given seed
// (xoshiro256** seeded through splitmix64).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "psm_abi.h"

namespace {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        for (int i = 0; i < 4; ++i) {
            seed += 0x9E3779B97F4A7C15ull;
            uint64_t z = seed;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            s[i] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    uint64_t below(uint64_t n) { return n ? next() % n : 0; }
};

// Open-addressing set of undirected edges packed as (min << 32 | max).
struct EdgeSet {
    std::vector<uint64_t> slots;
    uint64_t mask;
    explicit EdgeSet(uint64_t expected) {
        uint64_t cap = 16;
        while (cap < expected * 2) cap <<= 1;
        slots.assign(cap, ~0ull);
        mask = cap - 1;
    }
    bool insert(uint64_t key) {
        uint64_t h = key * 0x9E3779B97F4A7C15ull;
        uint64_t i = (h ^ (h >> 29)) & mask;
        while (true) {
            if (slots[i] == ~0ull) { slots[i] = key; return true; }
            if (slots[i] == key) return false;
            i = (i + 1) & mask;
        }
    }
};

void sample_discrete(const std::vector<double>& cdf, Rng& r, uint32_t* out, uint64_t n) {
    double total = cdf.back();
    for (uint64_t i = 0; i < n; ++i) {
        double x = r.uniform() * total;
        out[i] = (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), x) - cdf.begin());
        if (out[i] >= cdf.size()) out[i] = (uint32_t)cdf.size() - 1;
    }
}

// Builds sorted CSR from an edge list (u < v pairs, each once).
void edges_to_csr(uint32_t V, const std::vector<uint64_t>& edges, uint32_t* offsets, uint32_t* nbrs) {
    std::vector<uint32_t> deg(V, 0);
    for (uint64_t e : edges) { deg[e >> 32]++; deg[(uint32_t)e]++; }
    offsets[0] = 0;
    for (uint32_t v = 0; v < V; ++v) offsets[v + 1] = offsets[v] + deg[v];
    std::vector<uint32_t> fill(offsets, offsets + V);
    for (uint64_t e : edges) {
        uint32_t a = (uint32_t)(e >> 32), b = (uint32_t)e;
        nbrs[fill[a]++] = b;
        nbrs[fill[b]++] = a;
    }
    for (uint32_t v = 0; v < V; ++v) std::sort(nbrs + offsets[v], nbrs + offsets[v + 1]);
}

}  // namespace

extern "C" {

int psm_gen_chung_lu(uint32_t V, uint64_t E, double gamma, uint64_t seed,
                     uint32_t* offsets, uint32_t* nbrs) {
    if (V < 2 || gamma <= 1.0 || E > (uint64_t)V * (V - 1) / 2) return PSM_EINVAL;
    Rng r(seed);
    std::vector<double> cdf(V);
    double acc = 0, ex = -1.0 / (gamma - 1.0);
    for (uint32_t i = 0; i < V; ++i) { acc += std::pow((double)(i + 1), ex); cdf[i] = acc; }
    EdgeSet set(E);
    std::vector<uint64_t> edges;
    edges.reserve(E);
    std::vector<uint32_t> buf(2 * 65536);
    uint64_t attempts = 0;
    while (edges.size() < E) {
        sample_discrete(cdf, r, buf.data(), buf.size());
        for (size_t i = 0; i + 1 < buf.size() && edges.size() < E; i += 2) {
            uint32_t a = buf[i], b = buf[i + 1];
            if (a == b) continue;
            if (a > b) std::swap(a, b);
            uint64_t key = ((uint64_t)a << 32) | b;
            if (set.insert(key)) edges.push_back(key);
        }
        if (++attempts > 1000000) return PSM_EINVAL;  // unreachable density
    }
    edges_to_csr(V, edges, offsets, nbrs);
    return PSM_OK;
}

int psm_gen_rmat(uint32_t scale, uint64_t E, double a, double b, double c, uint64_t seed,
                 uint32_t* offsets, uint32_t* nbrs) {
    if (scale == 0 || scale > 31) return PSM_EINVAL;
    uint32_t V = 1u << scale;
    if (E > (uint64_t)V * (V - 1) / 2 || a + b + c >= 1.0) return PSM_EINVAL;
    Rng r(seed);
    EdgeSet set(E);
    std::vector<uint64_t> edges;
    edges.reserve(E);
    double ab = a + b, abc = a + b + c;
    uint64_t tries = 0, limit = E * 64 + 1000000;
    while (edges.size() < E && tries++ < limit) {
        uint32_t x = 0, y = 0;
        for (uint32_t l = 0; l < scale; ++l) {
            double p = r.uniform();
            uint32_t bx = 0, by = 0;
            if (p < a) { }
            else if (p < ab) { by = 1; }
            else if (p < abc) { bx = 1; }
            else { bx = 1; by = 1; }
            x = (x << 1) | bx;
            y = (y << 1) | by;
        }
        if (x == y) continue;
        if (x > y) std::swap(x, y);
        uint64_t key = ((uint64_t)x << 32) | y;
        if (set.insert(key)) edges.push_back(key);
    }
    if (edges.size() < E) return PSM_EINVAL;
    edges_to_csr(V, edges, offsets, nbrs);
    return PSM_OK;
}

int psm_gen_labels(uint32_t V, uint32_t L, double zipf_s, uint64_t seed, uint32_t* labels) {
    if (L == 0) return PSM_EINVAL;
    Rng r(seed);
    if (zipf_s <= 0.0) {
        for (uint32_t v = 0; v < V; ++v) labels[v] = (uint32_t)r.below(L);
        return PSM_OK;
    }
    std::vector<double> cdf(L);
    double acc = 0;
    for (uint32_t k = 0; k < L; ++k) { acc += std::pow((double)(k + 1), -zipf_s); cdf[k] = acc; }
    sample_discrete(cdf, r, labels, V);
    return PSM_OK;
}

// One connected query of n vertices drawn from the data graph.
// dense = 0: random walk, edges = first-visit (tree) edges.
// dense = 1: greedy growth that adds the frontier vertex with most edges into
//            the current set (random tie-break), edges = induced subgraph.
// dense = 2: random walk, edges = induced subgraph of the walked vertices
//            (the walk tree plus every data edge among them: cycles).
// dense = 3: degree-biased random walk (each step moves to a neighbour w with
//            probability deg(w) / sum of the neighbours' degrees, from the
//            endpoint of a random edge), edges = induced subgraph. The walk
//            stays among high-degree vertices, whose induced subgraphs are
//            dense on power-law graphs where a uniform walk's are trees.
// Output: vertex ids (data ids, query vertex i = out_vertices[i]) and the
// query's undirected edges as pairs of query-local ids. Returns the number of
// edges, or a negative PSM status.
int64_t psm_gen_query(uint32_t V, const uint32_t* offsets, const uint32_t* nbrs,
                      uint32_t n, int dense, uint64_t seed,
                      uint32_t* out_vertices, uint32_t* out_edges, uint64_t max_edges) {
    if (n == 0 || n > 64 || V == 0 || dense < 0 || dense > 3) return -PSM_EINVAL;
    const bool walk = dense != 1, induced = dense != 0, biased = dense == 3;
    Rng r(seed);
    for (int attempt = 0; attempt < 1000; ++attempt) {
        uint32_t start = (uint32_t)r.below(V);
        if (offsets[start + 1] == offsets[start]) continue;
        std::vector<uint32_t> verts{start};
        std::vector<std::pair<uint32_t, uint32_t>> tree;
        if (biased) {
            const uint64_t nnz = offsets[V];
            const uint64_t e = r.below(nnz);
            uint32_t cur = (uint32_t)(std::upper_bound(offsets, offsets + V + 1, (uint32_t)e) - offsets - 1);
            verts[0] = cur;
            std::vector<uint64_t> cum;
            uint64_t steps = 0;
            while (verts.size() < n && steps++ < 100000) {
                const uint32_t d = offsets[cur + 1] - offsets[cur];
                cum.resize(d);
                uint64_t acc = 0;
                for (uint32_t k = 0; k < d; ++k) {
                    const uint32_t w = nbrs[offsets[cur] + k];
                    acc += offsets[w + 1] - offsets[w];
                    cum[k] = acc;
                }
                const uint64_t x = r.below(acc);
                const uint32_t k = (uint32_t)(std::upper_bound(cum.begin(), cum.end(), x) - cum.begin());
                const uint32_t w = nbrs[offsets[cur] + k];
                if (std::find(verts.begin(), verts.end(), w) == verts.end()) verts.push_back(w);
                cur = w;
            }
        } else if (walk) {
            uint32_t cur = 0;  // index into verts
            uint64_t steps = 0;
            while (verts.size() < n && steps++ < 100000) {
                uint32_t v = verts[cur];
                uint32_t d = offsets[v + 1] - offsets[v];
                uint32_t w = nbrs[offsets[v] + r.below(d)];
                auto it = std::find(verts.begin(), verts.end(), w);
                if (it == verts.end()) {
                    tree.push_back({cur, (uint32_t)verts.size()});
                    verts.push_back(w);
                    cur = (uint32_t)verts.size() - 1;
                } else {
                    cur = (uint32_t)(it - verts.begin());
                }
                // restart from a random visited vertex now and then so walks
                // do not get trapped on a degree-1 tail
                if (r.below(8) == 0) cur = (uint32_t)r.below(verts.size());
            }
        } else {
            // start at a degree-weighted vertex (endpoint of a random edge):
            // dense neighbourhoods sit around the hubs
            const uint64_t nnz = offsets[V];
            const uint64_t e = r.below(nnz);
            verts[0] = (uint32_t)(std::upper_bound(offsets, offsets + V + 1, (uint32_t)e) - offsets - 1);
            std::unordered_map<uint32_t, uint32_t> score;   // frontier vertex -> edges into the set
            auto absorb = [&](uint32_t v) {
                score.erase(v);
                const uint32_t d = offsets[v + 1] - offsets[v];
                const uint32_t lim = std::min<uint32_t>(d, 4096);
                for (uint32_t k = 0; k < lim; ++k) {
                    const uint32_t w = nbrs[offsets[v] + (d > 4096 ? r.below(d) : k)];
                    if (std::find(verts.begin(), verts.end(), w) != verts.end()) continue;
                    score[w]++;
                }
            };
            absorb(verts[0]);
            while (verts.size() < n && !score.empty()) {
                uint32_t best = 0, pick = 0, ties = 0;
                for (auto& kv : score) {
                    if (kv.second > best) { best = kv.second; pick = kv.first; ties = 1; }
                    else if (kv.second == best && r.below(++ties) == 0) pick = kv.first;
                }
                verts.push_back(pick);
                absorb(pick);
            }
        }
        if (verts.size() < n) continue;
        std::vector<std::pair<uint32_t, uint32_t>> edges;
        if (!induced) {
            edges = tree;
        } else {
            for (uint32_t i = 0; i < n; ++i) {
                uint32_t v = verts[i];
                for (uint32_t j = i + 1; j < n; ++j) {
                    uint32_t w = verts[j];
                    if (std::binary_search(nbrs + offsets[v], nbrs + offsets[v + 1], w))
                        edges.push_back({i, j});
                }
            }
        }
        if (edges.size() > max_edges) return -PSM_EINVAL;
        for (uint32_t i = 0; i < n; ++i) out_vertices[i] = verts[i];
        for (size_t i = 0; i < edges.size(); ++i) {
            out_edges[2 * i] = edges[i].first;
            out_edges[2 * i + 1] = edges[i].second;
        }
        return (int64_t)edges.size();
    }
    return -PSM_ENOTFOUND;
}

}  // extern "C"
