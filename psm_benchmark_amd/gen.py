"""Seeded synthetic workloads (C++ generators in libpsm.so, csrc/gen.cpp).

Data graphs: Chung-Lu power law (weights ~ (i+1)^(-1/(gamma-1))) and R-MAT;
labels Zipf(s) or uniform. Queries: connected vertex sets drawn from the data
graph, so each has at least one embedding (the walk itself).
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from .graph_io import CSRGraph, csr_from_edges


def chung_lu(V: int, E: int, gamma: float, seed: int) -> tuple[np.ndarray, np.ndarray]:
    off = np.zeros(V + 1, np.uint32)
    adj = np.zeros(2 * E, np.uint32)
    r = L.lib().psm_gen_chung_lu(V, E, gamma, seed, L.u32p(off), L.u32p(adj))
    if r != L.PSM_OK:
        raise L.PSMError(r, "psm_gen_chung_lu")
    return off, adj


def rmat(scale: int, E: int, a: float, b: float, c: float, seed: int):
    V = 1 << scale
    off = np.zeros(V + 1, np.uint32)
    adj = np.zeros(2 * E, np.uint32)
    r = L.lib().psm_gen_rmat(scale, E, a, b, c, seed, L.u32p(off), L.u32p(adj))
    if r != L.PSM_OK:
        raise L.PSMError(r, "psm_gen_rmat")
    return off, adj


def labels(V: int, num_labels: int, zipf_s: float, seed: int) -> np.ndarray:
    lab = np.zeros(max(1, V), np.uint32)
    r = L.lib().psm_gen_labels(V, num_labels, zipf_s, seed, L.u32p(lab))
    if r != L.PSM_OK:
        raise L.PSMError(r, "psm_gen_labels")
    return lab[:V]


def chung_lu_graph(V, E, num_labels, zipf_s, seed, gamma=2.5) -> CSRGraph:
    off, adj = chung_lu(V, E, gamma, seed)
    return CSRGraph(off, adj, labels(V, num_labels, zipf_s, seed ^ 0x5A5A5A5A))


def rmat_graph(scale, E, num_labels, seed, a=0.57, b=0.19, c=0.19) -> CSRGraph:
    off, adj = rmat(scale, E, a, b, c, seed)
    return CSRGraph(off, adj, labels(1 << scale, num_labels, 0.0, seed ^ 0x5A5A5A5A))


def query(data: CSRGraph, n: int, dense, seed: int) -> CSRGraph:
    """An n-vertex connected query drawn from `data` (labels kept). dense:
    False/0 random-walk tree, True/1 greedy densest induced subgraph, 2 the
    induced subgraph of a random walk (with cycles)."""
    verts = np.zeros(n, np.uint32)
    max_e = n * (n - 1) // 2
    edges = np.zeros(2 * max(1, max_e), np.uint32)
    off = np.ascontiguousarray(data.offsets, np.uint32)
    adj = np.ascontiguousarray(data.neighbors, np.uint32)
    m = L.lib().psm_gen_query(data.num_vertices, L.u32p(off), L.u32p(adj), n, int(dense), seed,
                              L.u32p(verts), L.u32p(edges), max_e)
    if m < 0:
        raise L.PSMError(int(-m), "psm_gen_query")
    e = edges[: 2 * m].reshape(-1, 2).astype(np.int64)
    return csr_from_edges(n, e, data.labels[verts])
