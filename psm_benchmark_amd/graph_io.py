"""Graph files in the reference's formats, held as uint32 CSR numpy arrays.

Text format (graph/graph.cpp:92-172 of the reference)::

    t <V> <E>
    v <id> <label> <degree>
    e <u> <v>

Binary format (graph/graph.cpp:221-403, matching/EdgeListToCSR.cpp:77-121):
``<name>_deg.bin`` = [i32 4][u32 V][u32 E][u32 deg[V]], ``<name>_adj.bin`` =
u32 neighbours vertex by vertex, ``<name>_label.bin`` = u32 label[V].
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class CSRGraph:
    """Undirected labelled graph: adjacency sorted ascending per vertex."""

    offsets: np.ndarray   # uint32[V+1]
    neighbors: np.ndarray  # uint32[2E]
    labels: np.ndarray     # uint32[V]

    @property
    def num_vertices(self) -> int:
        return int(self.labels.shape[0])

    @property
    def num_edges(self) -> int:
        return int(self.neighbors.shape[0]) // 2

    @property
    def degrees(self) -> np.ndarray:
        return np.diff(self.offsets.astype(np.int64)).astype(np.uint32)

    def edges(self) -> np.ndarray:
        """Undirected edges (u < v) as an int64[E, 2] array."""
        deg = self.degrees.astype(np.int64)
        src = np.repeat(np.arange(self.num_vertices, dtype=np.int64), deg)
        dst = self.neighbors.astype(np.int64)
        m = src < dst
        return np.stack([src[m], dst[m]], axis=1)


def csr_from_edges(num_vertices: int, edges: np.ndarray, labels: np.ndarray) -> CSRGraph:
    """Builds a CSR graph from undirected edges (each listed once)."""
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    src = np.concatenate([e[:, 0], e[:, 1]])
    dst = np.concatenate([e[:, 1], e[:, 0]])
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    counts = np.bincount(src, minlength=num_vertices)
    offsets = np.zeros(num_vertices + 1, dtype=np.uint64)
    np.cumsum(counts, out=offsets[1:])
    return CSRGraph(offsets.astype(np.uint32), dst.astype(np.uint32),
                    np.asarray(labels, dtype=np.uint32).copy())


def load_text(path: str) -> CSRGraph:
    """Reads the reference's ``t/v/e`` text format."""
    labels = None
    edges = []
    with open(path) as f:
        for line in f:
            p = line.split()
            if not p:
                continue
            if p[0] == "t":
                labels = np.zeros(int(p[1]), dtype=np.uint32)
            elif p[0] == "v":
                labels[int(p[1])] = int(p[2])
            elif p[0] == "e":
                edges.append((int(p[1]), int(p[2])))
    if labels is None:
        raise ValueError(f"{path}: missing 't' header line")
    return csr_from_edges(labels.shape[0], np.array(edges, dtype=np.int64).reshape(-1, 2), labels)


def load_text_parallel(path: str, threads: int = 0) -> CSRGraph:
    """The graph load_text returns, read by libpsm's multi-threaded parser
    (psm_text_graph_read; threads 0 = every hardware thread). Raises
    PSMError on a malformed file, including v-line degree fields that
    disagree with the e lines."""
    import ctypes
    from . import _lib as L
    h = ctypes.c_void_p()
    r = L.lib().psm_text_graph_read(path.encode(), threads, ctypes.byref(h))
    if r != L.PSM_OK:
        raise L.PSMError(r, L.lib().psm_text_graph_last_error().decode())
    try:
        v, e = ctypes.c_uint32(), ctypes.c_uint64()
        L.lib().psm_text_graph_info(h, ctypes.byref(v), ctypes.byref(e))
        n, m = v.value, 2 * e.value

        def take(ptr, count):
            if count == 0:
                return np.zeros(0, np.uint32)
            return np.ctypeslib.as_array(ptr, shape=(count,)).copy()
        return CSRGraph(take(L.lib().psm_text_graph_offsets(h), n + 1),
                        take(L.lib().psm_text_graph_neighbors(h), m),
                        take(L.lib().psm_text_graph_labels(h), n))
    finally:
        L.lib().psm_text_graph_free(h)


def save_text(g: CSRGraph, path: str) -> None:
    """Writes the ``t/v/e`` text format (degree field included: the
    reference's loader sizes adjacency from it)."""
    deg = g.degrees
    e = g.edges()
    with open(path, "w") as f:
        f.write(f"t {g.num_vertices} {e.shape[0]}\n")
        vlines = np.stack([np.arange(g.num_vertices), g.labels, deg], axis=1)
        np.savetxt(f, vlines, fmt="v %d %d %d")
        if e.shape[0]:
            np.savetxt(f, e, fmt="e %d %d")


def save_binary(g: CSRGraph, prefix: str) -> None:
    """Writes ``<prefix>_deg.bin``, ``_adj.bin`` and ``_label.bin``."""
    with open(prefix + "_deg.bin", "wb") as f:
        np.array([4, g.num_vertices, g.num_edges], dtype=np.uint32).tofile(f)
        g.degrees.astype(np.uint32).tofile(f)
    g.neighbors.astype(np.uint32).tofile(prefix + "_adj.bin")
    g.labels.astype(np.uint32).tofile(prefix + "_label.bin")


def load_binary(prefix: str) -> CSRGraph:
    """Reads the binary deg/adj/label triple (adjacency re-sorted per vertex)."""
    head = np.fromfile(prefix + "_deg.bin", dtype=np.uint32, count=3)
    v = int(head[1])
    deg = np.fromfile(prefix + "_deg.bin", dtype=np.uint32, offset=12, count=v)
    offsets = np.zeros(v + 1, dtype=np.uint64)
    np.cumsum(deg, out=offsets[1:])
    adj = np.fromfile(prefix + "_adj.bin", dtype=np.uint32)
    src = np.repeat(np.arange(v, dtype=np.int64), deg.astype(np.int64))
    order = np.lexsort((adj, src))
    labels = np.fromfile(prefix + "_label.bin", dtype=np.uint32, count=v)
    return CSRGraph(offsets.astype(np.uint32), adj[order].astype(np.uint32), labels)
