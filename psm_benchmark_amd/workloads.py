"""The benchmark workloads, seeded and regenerated on demand.

C1  plain:    Chung-Lu 9,460 V / 34,998 E, 307 Zipf(1.1) labels; one 8-vertex
              random-walk query; GQL filter + GQL order; uncapped.
C2  batched:  Chung-Lu 1.1M V / 3.0M E, 100 Zipf(1.1) labels; 512 queries of
              4/8/16/24/32 vertices, half dense (induced subgraph of a
              degree-biased walk, average degree > 3; K4 at 4 vertices) and
              half sparse (walk tree); GQL filter + GQL order; cap 1e5.
C3  join:     Chung-Lu 4M V / 35M E, 32 Zipf(1.0) labels; 64 cyclic queries of
              12-16 vertices; CFL filter + CFL order; uncapped.
C4  serving:  the C3 graph; Poisson arrivals; cap 1e5.
C5  stress:   R-MAT (0.57, 0.19, 0.19) 2^24 V / 100M E, 8 uniform labels;
              32 dense queries of 6-10 vertices; uncapped.
"""
from __future__ import annotations

import numpy as np

from . import gen
from .graph_io import CSRGraph

GRAPHS = {
    "C1": dict(kind="chung_lu", V=9460, E=34998, labels=307, zipf=1.1, seed=1),
    "C2": dict(kind="chung_lu", V=1_100_000, E=3_000_000, labels=100, zipf=1.1, seed=2),
    "C3": dict(kind="chung_lu", V=4_000_000, E=35_000_000, labels=32, zipf=1.0, seed=3),
    "C5": dict(kind="rmat", scale=24, E=100_000_000, labels=8, seed=5),
}


def data_graph(name: str) -> CSRGraph:
    p = GRAPHS["C3" if name == "C4" else name]
    if p["kind"] == "chung_lu":
        return gen.chung_lu_graph(p["V"], p["E"], p["labels"], p["zipf"], seed=p["seed"])
    return gen.rmat_graph(p["scale"], p["E"], p["labels"], seed=p["seed"])


def _avg_degree(q: CSRGraph) -> float:
    return 2.0 * q.num_edges / q.num_vertices


def is_dense(q: CSRGraph) -> bool:
    """The dense half's rule: average degree above 3; a 4-vertex query can
    reach 3 at most (K4), so there dense means K4."""
    n = q.num_vertices
    return q.num_edges == 6 if n == 4 else 2.0 * q.num_edges > 3.0 * n


def c2_queries(g: CSRGraph, count: int = 512, seed: int = 22):
    """Random-walk queries; sizes cycle through 4/8/16/24/32, even index
    dense, odd sparse. Sparse: the tree of a uniform random walk (average
    degree 2(n-1)/n <= 3). Dense: the induced subgraph of a degree-biased
    random walk (gen.query mode 3), redrawn until it meets is_dense. On a
    power-law graph a uniform walk's induced subgraph is nearly a tree, so
    the dense half walks towards high-degree vertices instead."""
    sizes = [4, 8, 16, 24, 32]
    qs, meta = [], []
    rng = np.random.default_rng(seed)
    for i in range(count):
        n = sizes[i % len(sizes)]
        dense = i % 2 == 0
        base = int(rng.integers(0, 2**62))
        if dense:
            for t in range(1 << 16):
                q = gen.query(g, n, 3, base + t)
                if is_dense(q):
                    break
            else:
                raise RuntimeError(f"no dense {n}-vertex walk in 65536 draws")
        else:
            q = gen.query(g, n, False, base)
        qs.append(q)
        meta.append(dict(n=n, dense=dense, edges=q.num_edges))
    return qs, meta


def c1_query(g: CSRGraph, seed: int = 11) -> CSRGraph:
    return gen.query(g, 8, True, seed)


def _adjsets(q: CSRGraph):
    off = q.offsets.astype(np.int64)
    return [set(int(x) for x in q.neighbors[off[v]:off[v + 1]]) for v in range(q.num_vertices)]


def has_cycle(q: CSRGraph) -> bool:
    """A connected query has a cycle iff it has at least as many edges as vertices."""
    return q.num_edges >= q.num_vertices


def clique_size(q: CSRGraph, k: int) -> bool:
    """True when the query contains a k-clique (k <= 4, small queries)."""
    import itertools
    adj = _adjsets(q)
    for c in itertools.combinations(range(q.num_vertices), k):
        if all(b in adj[a] for a, b in itertools.combinations(c, 2)):
            return True
    return False


def c3_queries(g: CSRGraph, count: int = 64, seed: int = 33):
    """64 queries of 12-16 vertices whose induced subgraph has a cycle: the
    vertices of a random walk with every data edge among them."""
    rng = np.random.default_rng(seed)
    qs, meta = [], []
    while len(qs) < count:
        n = int(rng.integers(12, 17))
        q = gen.query(g, n, 2, int(rng.integers(0, 2**62)))
        if not has_cycle(q):
            continue
        qs.append(q)
        meta.append(dict(n=n, edges=q.num_edges))
    return qs, meta


def c5_queries(g: CSRGraph, count: int = 32, seed: int = 55):
    """32 dense queries of 6-10 vertices (greedy densest growth); every query
    has a triangle and every other one a 4-clique."""
    rng = np.random.default_rng(seed)
    qs, meta = [], []
    while len(qs) < count:
        n = int(rng.integers(6, 11))
        want4 = len(qs) % 2 == 0
        for _ in range(64):
            q = gen.query(g, n, 1, int(rng.integers(0, 2**62)))
            if clique_size(q, 3) and (not want4 or clique_size(q, 4)):
                break
        else:
            continue
        qs.append(q)
        meta.append(dict(n=n, edges=q.num_edges, k4=clique_size(q, 4)))
    return qs, meta


def c3_join_queries(g: CSRGraph, count: int = 16, seed: int = 35, rare_labels: int = 16,
                    sizes=(6, 9)):
    """Join-checkable queries on the C3 graph: `sizes` (inclusive) vertices
    with a cycle, the vertices of a random walk over data vertices of the
    `rare_labels` least frequent labels, with every data edge among them.
    Chosen so that their counts can be enumerated in full and the two-unit
    join compared with the plain count at C3 scale (12-16-vertex queries of
    this graph have more than 1e10 embeddings each, even on the rare
    labels)."""
    freq = np.bincount(g.labels, minlength=int(g.labels.max()) + 1)
    allowed = np.zeros(freq.shape[0], bool)
    allowed[np.argsort(freq, kind="stable")[:rare_labels]] = True
    ok = allowed[g.labels]
    off = g.offsets.astype(np.int64)
    rng = np.random.default_rng(seed)
    starts = np.flatnonzero(ok)
    qs, meta = [], []
    tries = 0
    while len(qs) < count and tries < 100 * count:
        tries += 1
        n = int(rng.integers(sizes[0], sizes[1] + 1))
        verts = [int(rng.choice(starts))]
        cur = verts[0]
        for _ in range(50 * n):
            if len(verts) >= n:
                break
            nb = g.neighbors[off[cur]:off[cur + 1]]
            nb = nb[ok[nb]]
            if nb.size == 0:
                cur = verts[int(rng.integers(0, len(verts)))]
                continue
            w = int(nb[int(rng.integers(0, nb.size))])
            if w not in verts:
                verts.append(w)
            cur = w if rng.random() < 0.85 else verts[int(rng.integers(0, len(verts)))]
        if len(verts) < n:
            continue
        pos = {v: i for i, v in enumerate(verts)}
        edges = []
        for i, v in enumerate(verts):
            for w in g.neighbors[off[v]:off[v + 1]]:
                j = pos.get(int(w))
                if j is not None and i < j:
                    edges.append((i, j))
        from .graph_io import csr_from_edges
        q = csr_from_edges(n, np.array(edges, np.int64).reshape(-1, 2), g.labels[np.array(verts)])
        if not has_cycle(q):
            continue
        qs.append(q)
        meta.append(dict(n=n, edges=q.num_edges))
    return qs, meta
