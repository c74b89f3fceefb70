// textio.cpp — the reference's text graph format read by several host threads.
//
// Format (Graph::loadGraphFromFile, graph/graph.cpp:92-172):
//     t <V> <E>
//     v <id> <label> <degree>      one per vertex
//     e <u> <v>                    one per undirected edge
// The file is memory-mapped and cut into one byte range per thread at line
// boundaries. Each thread parses its range: v lines go straight into the
// label and degree-field arrays (ids are disjoint), e lines into a
// thread-local edge list. Degrees are then counted from the edges with
// atomic increments, offsets are an exclusive scan, every thread scatters its
// edges through per-vertex atomic cursors, and adjacency lists are sorted in
// parallel over vertex ranges of equal edge counts. The reference places
// neighbours by the v lines' degree fields; here adjacency is built from the
// e lines and each degree field must agree with them, so a malformed file is
// an error instead of a corrupt graph.

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "psm_abi.h"

struct psm_text_graph {
    uint32_t V = 0;
    uint64_t E = 0;
    std::vector<uint32_t> off, adj, lab;
};

namespace {

thread_local std::string g_err;

int fail(const std::string& m) {
    g_err = m;
    return PSM_EINVAL;
}

struct Mapped {
    const char* p = nullptr;
    size_t n = 0;
    int fd = -1;
    ~Mapped() {
        if (p && n) munmap(const_cast<char*>(p), n);
        if (fd >= 0) close(fd);
    }
};

inline void skip_blank(const char*& p, const char* e) {
    while (p < e && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
}

// Unsigned decimal after optional blanks; false when absent or above `max`.
inline bool read_uint(const char*& p, const char* e, uint64_t max, uint64_t& x) {
    skip_blank(p, e);
    if (p >= e || *p < '0' || *p > '9') return false;
    x = 0;
    while (p < e && *p >= '0' && *p <= '9') {
        x = x * 10 + (uint64_t)(*p - '0');
        if (x > max) return false;
        ++p;
    }
    return true;
}

inline const char* line_end(const char* p, const char* e) {
    const void* nl = memchr(p, '\n', (size_t)(e - p));
    return nl ? static_cast<const char*>(nl) : e;
}

struct Part {
    std::vector<std::pair<uint32_t, uint32_t>> edges;
    uint64_t vlines = 0;
    std::string err;
};

// Parses the lines of [b, e) (b at a line start, e at a line end or the end
// of the file).
void parse_range(const char* b, const char* e, uint32_t V, uint32_t* lab, uint32_t* degf,
                 uint8_t* seen, Part& out) {
    const char* p = b;
    while (p < e && out.err.empty()) {
        const char* le = line_end(p, e);
        const char* q = p;
        skip_blank(q, le);
        if (q < le) {
            const char t = *q++;
            uint64_t a = 0, c = 0, d = 0;
            if (t == 'v') {
                if (!read_uint(q, le, V ? V - 1 : 0, a) || !read_uint(q, le, UINT32_MAX, c) ||
                    !read_uint(q, le, UINT32_MAX, d) || V == 0) {
                    out.err = "malformed or out-of-range v line: " + std::string(p, le);
                } else if (__atomic_exchange_n(seen + a, (uint8_t)1, __ATOMIC_RELAXED)) {
                    out.err = "vertex " + std::to_string(a) + " listed twice";
                } else {
                    lab[a] = (uint32_t)c;
                    degf[a] = (uint32_t)d;
                    ++out.vlines;
                }
            } else if (t == 'e') {
                if (!read_uint(q, le, V ? V - 1 : 0, a) || !read_uint(q, le, V ? V - 1 : 0, c) || V == 0)
                    out.err = "malformed or out-of-range e line: " + std::string(p, le);
                else
                    out.edges.emplace_back((uint32_t)a, (uint32_t)c);
            } else {
                out.err = "unexpected line: " + std::string(p, le);
            }
        }
        p = le + 1;
    }
}

}  // namespace

extern "C" {

const char* psm_text_graph_last_error(void) { return g_err.c_str(); }

int psm_text_graph_read(const char* path, uint32_t threads, psm_text_graph** out) {
    if (!path || !out) return PSM_EINVAL;
    *out = nullptr;
    g_err.clear();
    Mapped m;
    m.fd = open(path, O_RDONLY);
    if (m.fd < 0) {
        g_err = std::string("cannot open ") + path;
        return PSM_ENOTFOUND;
    }
    struct stat sb;
    if (fstat(m.fd, &sb) != 0) return fail(std::string("cannot stat ") + path);
    m.n = (size_t)sb.st_size;
    if (m.n == 0) return fail("empty graph file");
    void* mp = mmap(nullptr, m.n, PROT_READ, MAP_PRIVATE, m.fd, 0);
    if (mp == MAP_FAILED) {
        m.n = 0;
        return fail(std::string("cannot map ") + path);
    }
    m.p = static_cast<const char*>(mp);
    (void)madvise(mp, m.n, MADV_SEQUENTIAL);
    const char* const end = m.p + m.n;

    // header: the first non-blank line, "t V E"
    const char* p = m.p;
    uint64_t V = 0, E = 0;
    for (;;) {
        if (p >= end) return fail("missing 't' header line");
        const char* le = line_end(p, end);
        const char* q = p;
        skip_blank(q, le);
        if (q < le) {
            if (*q != 't') return fail("first line is not the 't V E' header");
            ++q;
            // (read_uint's bound keeps x * 10 within 64 bits)
            if (!read_uint(q, le, UINT32_MAX, V) || !read_uint(q, le, UINT32_MAX, E))
                return fail("malformed or out-of-range header: " + std::string(p, le));
            p = le + 1;
            break;
        }
        p = le + 1;
    }
    if (2 * E > UINT32_MAX) return fail("more than 2^31 edges: offsets are 32-bit");

    std::unique_ptr<psm_text_graph> g(new psm_text_graph());
    g->V = (uint32_t)V;
    g->E = E;
    g->lab.assign(V, 0);
    std::vector<uint32_t> degf(V, 0);
    std::vector<uint8_t> seen(V, 0);

    uint32_t T = threads ? threads : std::max(1u, std::thread::hardware_concurrency());
    T = std::min<uint32_t>(T, 64);
    const char* body = std::min(p, end);
    const size_t body_n = (size_t)(end - body);
    if (body_n < (size_t)T * 4096) T = std::max<uint32_t>(1, (uint32_t)(body_n / 4096));
    std::vector<const char*> cut(T + 1);
    cut[0] = body;
    cut[T] = end;
    for (uint32_t i = 1; i < T; ++i) {
        const char* c = body + body_n / T * i;
        c = std::max(c, cut[i - 1]);
        const char* le = line_end(c, end);
        cut[i] = le < end ? le + 1 : end;
    }
    std::vector<Part> parts(T);
    auto run = [&](const std::function<void(uint32_t)>& fn) {
        std::vector<std::thread> th;
        for (uint32_t i = 1; i < T; ++i) th.emplace_back(fn, i);
        fn(0);
        for (auto& t : th) t.join();
    };
    run([&](uint32_t i) {
        parse_range(cut[i], cut[i + 1], (uint32_t)V, g->lab.data(), degf.data(), seen.data(), parts[i]);
    });
    uint64_t vlines = 0, elines = 0;
    for (const Part& pt : parts) {
        if (!pt.err.empty()) return fail(pt.err);
        vlines += pt.vlines;
        elines += pt.edges.size();
    }
    if (vlines != V)
        return fail("header lists " + std::to_string(V) + " vertices, file has " + std::to_string(vlines) +
                    " v lines");
    if (elines != E)
        return fail("header lists " + std::to_string(E) + " edges, file has " + std::to_string(elines) +
                    " e lines");

    // degrees from the edges, offsets, scatter through per-vertex cursors
    // (one thread needs no atomic read-modify-writes)
    std::vector<uint32_t> cnt(V, 0);
    run([&](uint32_t i) {
        if (T == 1) {
            for (const auto& e : parts[i].edges) {
                ++cnt[e.first];
                ++cnt[e.second];
            }
            return;
        }
        for (const auto& e : parts[i].edges) {
            __atomic_fetch_add(&cnt[e.first], 1u, __ATOMIC_RELAXED);
            __atomic_fetch_add(&cnt[e.second], 1u, __ATOMIC_RELAXED);
        }
    });
    g->off.assign(V + 1, 0);
    for (uint64_t v = 0; v < V; ++v) {
        if (cnt[v] != degf[v])
            return fail("vertex " + std::to_string(v) + ": degree field " + std::to_string(degf[v]) +
                        ", e lines give " + std::to_string(cnt[v]));
        g->off[v + 1] = g->off[v] + cnt[v];
    }
    g->adj.assign(2 * E, 0);
    std::vector<uint32_t> cur(g->off.begin(), g->off.begin() + V);
    run([&](uint32_t i) {
        if (T == 1) {
            for (const auto& e : parts[i].edges) {
                g->adj[cur[e.first]++] = e.second;
                g->adj[cur[e.second]++] = e.first;
            }
        } else {
            for (const auto& e : parts[i].edges) {
                g->adj[__atomic_fetch_add(&cur[e.first], 1u, __ATOMIC_RELAXED)] = e.second;
                g->adj[__atomic_fetch_add(&cur[e.second], 1u, __ATOMIC_RELAXED)] = e.first;
            }
        }
        std::vector<std::pair<uint32_t, uint32_t>>().swap(parts[i].edges);
    });
    // ascending adjacency; thread i sorts the vertices whose lists start in
    // the i-th of T equal slices of the neighbour array
    auto first_vertex = [&](uint32_t i) -> uint64_t {
        if (i == 0) return 0;
        if (i == T) return V;
        const uint32_t target = (uint32_t)(2 * E * i / T);
        return (uint64_t)(std::lower_bound(g->off.begin(), g->off.begin() + V, target) - g->off.begin());
    };
    run([&](uint32_t i) {
        const uint64_t a = first_vertex(i), b = first_vertex(i + 1);
        for (uint64_t v = a; v < b; ++v) std::sort(g->adj.begin() + g->off[v], g->adj.begin() + g->off[v + 1]);
    });
    *out = g.release();
    return PSM_OK;
}

int psm_text_graph_info(const psm_text_graph* g, uint32_t* num_vertices, uint64_t* num_edges) {
    if (!g) return PSM_EINVAL;
    if (num_vertices) *num_vertices = g->V;
    if (num_edges) *num_edges = g->E;
    return PSM_OK;
}

const uint32_t* psm_text_graph_offsets(const psm_text_graph* g) { return g ? g->off.data() : nullptr; }
const uint32_t* psm_text_graph_neighbors(const psm_text_graph* g) { return g ? g->adj.data() : nullptr; }
const uint32_t* psm_text_graph_labels(const psm_text_graph* g) { return g ? g->lab.data() : nullptr; }

void psm_text_graph_free(psm_text_graph* g) { delete g; }

}  // extern "C"
