"""Brute-force subgraph isomorphism counting for tiny graphs (test checker).

Counts injective maps f: V(q) -> V(G) with label(f(u)) == label(u) and every
query edge mapped onto a data edge (non-induced matching, the semantics of the
reference's engines). Pure Python; meant for graphs of a dozen vertices.
"""
from __future__ import annotations

import itertools

import numpy as np


def adjacency_sets(g):
    off = g.offsets.astype(np.int64)
    return [set(int(x) for x in g.neighbors[off[v]:off[v + 1]]) for v in range(g.num_vertices)]


def count_embeddings(data, query) -> int:
    da, qa = adjacency_sets(data), adjacency_sets(query)
    n = query.num_vertices
    by_label = {}
    for v in range(data.num_vertices):
        by_label.setdefault(int(data.labels[v]), []).append(v)
    order = list(range(n))
    count = 0
    assign = [-1] * n

    def rec(i, used):
        nonlocal count
        if i == n:
            count += 1
            return
        u = order[i]
        for v in by_label.get(int(query.labels[u]), []):
            if v in used:
                continue
            if all(assign[w] in da[v] for w in qa[u] if w < u):
                assign[u] = v
                used.add(v)
                rec(i + 1, used)
                used.discard(v)
                assign[u] = -1

    rec(0, set())
    return count


def random_graph(rng, V, p, num_labels):
    from psm_benchmark_amd.graph_io import csr_from_edges
    edges = [(a, b) for a, b in itertools.combinations(range(V), 2) if rng.random() < p]
    return csr_from_edges(V, np.array(edges, np.int64).reshape(-1, 2), rng.integers(0, num_labels, V))


def random_connected_query(rng, data, n):
    """Connected n-vertex query taken from the data graph (walk tree + some
    induced edges), labels kept."""
    from psm_benchmark_amd.graph_io import csr_from_edges
    da = adjacency_sets(data)
    for _ in range(200):
        start = int(rng.integers(0, data.num_vertices))
        verts = [start]
        edges = []
        cur = start
        for _ in range(50 * n):
            if len(verts) == n:
                break
            nb = sorted(da[cur])
            if not nb:
                break
            w = int(nb[rng.integers(0, len(nb))])
            if w not in verts:
                edges.append((verts.index(cur), len(verts)))
                verts.append(w)
            cur = w
        if len(verts) < n:
            continue
        for i in range(n):
            for j in range(i + 1, n):
                if verts[j] in da[verts[i]] and (i, j) not in edges and rng.random() < 0.5:
                    edges.append((i, j))
        return csr_from_edges(n, np.array(edges, np.int64).reshape(-1, 2), data.labels[verts])
    return None
