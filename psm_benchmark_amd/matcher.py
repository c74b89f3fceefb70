"""Python driver over the C ABI: device context, resident data graph, the
plain per-query pipeline and the batched form.

Method strings follow the reference's ``-BackMethods`` (``FILTER_ORDER_ENGINE``,
e.g. ``GQL_GQL_GQL`` or ``CFL_CFL_EXPLORE``, matching/StudyPerformance.cpp:
1018-1030). Engines GQL, EXPLORE and LFTJ count the same embeddings; the
device path computes every depth's extensions as the intersection of the
candidate-edge table rows of all earlier neighbours. The Call Count follows
the engine: one per assignment for GQL and EXPLORE, plus the extensions
rejected because their data vertex is already used for LFTJ.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .graph_io import CSRGraph


def _check(ctx_handle, code: int, what: str, ok=(L.PSM_OK,)):
    if code not in ok:
        msg = L.lib().psm_last_error(ctx_handle).decode() if ctx_handle else ""
        raise L.PSMError(code, f"{what}: {msg}")
    return code


def parse_method(method: str):
    parts = method.split("_")
    if len(parts) != 3:
        raise ValueError(f"method must be FILTER_ORDER_ENGINE, got {method!r}")
    f, o, e = parts
    if f not in L.FILTERS or o not in L.ORDERS or e not in ("GQL", "EXPLORE", "LFTJ"):
        raise ValueError(f"unsupported method {method!r}")
    return L.FILTERS[f], L.ORDERS[o], e


class Device:
    """One HIP stream on one GPU (psm_ctx)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        _check(None, L.lib().psm_ctx_create(device, ctypes.byref(h)), "psm_ctx_create")
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            L.lib().psm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class DeviceGraph:
    """A data graph resident in HBM (psm_graph)."""

    def __init__(self, dev: Device, g: CSRGraph):
        self.dev = dev
        self.host = g
        off = np.ascontiguousarray(g.offsets, np.uint32)
        adj = np.ascontiguousarray(g.neighbors, np.uint32)
        lab = np.ascontiguousarray(g.labels, np.uint32)
        if adj.size == 0:
            adj = np.zeros(1, np.uint32)
        h = ctypes.c_void_p()
        _check(dev.h, L.lib().psm_graph_upload(dev.h, lab.shape[0], L.u32p(off), L.u32p(adj),
                                               L.u32p(lab), ctypes.byref(h)), "psm_graph_upload")
        self.h = h

    def info(self):
        a = np.zeros(5, np.uint64)
        L.lib().psm_graph_info(self.h, L.u64p(a))
        return dict(V=int(a[0]), E=int(a[1]), labels=int(a[2]), max_degree=int(a[3]),
                    max_label_frequency=int(a[4]))

    def close(self):
        if getattr(self, "h", None):
            L.lib().psm_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Query:
    """One query against one resident data graph, stage by stage (psm_query)."""

    def __init__(self, dg: DeviceGraph, q: CSRGraph):
        self.dg = dg
        self.n = q.num_vertices
        off = np.ascontiguousarray(q.offsets, np.uint32)
        adj = np.ascontiguousarray(q.neighbors, np.uint32)
        lab = np.ascontiguousarray(q.labels, np.uint32)
        if adj.size == 0:
            adj = np.zeros(1, np.uint32)
        h = ctypes.c_void_p()
        _check(dg.dev.h, L.lib().psm_query_create(dg.dev.h, dg.h, self.n, L.u32p(off), L.u32p(adj),
                                                  L.u32p(lab), ctypes.byref(h)), "psm_query_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            L.lib().psm_query_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def filter(self, method: str) -> bool:
        """FilterVertices::<method>Filter; False where the reference returns false."""
        r = _check(self.dg.dev.h, L.lib().psm_filter(self.h, L.FILTERS[method]), "psm_filter",
                   ok=(L.PSM_OK, L.PSM_EEMPTY))
        return r == L.PSM_OK

    def candidate_counts(self) -> np.ndarray:
        c = np.zeros(self.n, np.uint32)
        _check(self.dg.dev.h, L.lib().psm_query_candidate_counts(self.h, L.u32p(c)), "counts")
        return c

    def candidates(self, u: int) -> np.ndarray:
        c = self.candidate_counts()
        out = np.zeros(max(1, int(c[u])), np.uint32)
        _check(self.dg.dev.h, L.lib().psm_query_candidates(self.h, u, L.u32p(out)), "candidates")
        return out[: int(c[u])]

    def build_tables(self):
        _check(self.dg.dev.h, L.lib().psm_build_tables(self.h), "psm_build_tables")

    def table(self, a: int, b: int):
        nnz = ctypes.c_uint64()
        r = L.lib().psm_query_table(self.h, a, b, None, None, ctypes.byref(nnz))
        if r == L.PSM_EINVAL:
            return None
        _check(self.dg.dev.h, r, "psm_query_table")
        c = self.candidate_counts()
        off = np.zeros(int(c[a]) + 1, np.uint32)
        idx = np.zeros(max(1, nnz.value), np.uint32)
        _check(self.dg.dev.h, L.lib().psm_query_table(self.h, a, b, L.u32p(off), L.u32p(idx),
                                                      ctypes.byref(nnz)), "psm_query_table")
        return off, idx[: nnz.value]

    def memory_cost_bytes(self) -> int:
        b = ctypes.c_uint64()
        _check(self.dg.dev.h, L.lib().psm_memory_cost_bytes(self.h, ctypes.byref(b)), "memcost")
        return b.value

    def generate_plan(self, order: str):
        o = np.zeros(self.n, np.uint32)
        p = np.zeros(self.n, np.uint32)
        _check(self.dg.dev.h, L.lib().psm_generate_plan(self.h, L.ORDERS[order], L.u32p(o), L.u32p(p)),
               "psm_generate_plan")
        return o, p

    def join_count(self, max_rows: int = 0) -> L.JoinStats:
        """Two-unit decomposition of the planned query + device hash join."""
        st = L.JoinStats()
        _check(self.dg.dev.h, L.lib().psm_join_count(self.h, max_rows, ctypes.byref(st)), "psm_join_count")
        return st

    def split_join_count(self, pattern: str = "STAR", max_rows: int = 0) -> L.SplitJoinStats:
        """STAR / TWINTWIG / SEED split of the query + a sequence of device
        hash joins."""
        st = L.SplitJoinStats()
        _check(self.dg.dev.h, L.lib().psm_split_join_count(self.h, L.SPLITS[pattern], max_rows, ctypes.byref(st)),
               "psm_split_join_count")
        return st

    def evaluate(self, cap: int = 0) -> L.EnumStats:
        """Counts embeddings; st.timed_out is set when the time limit cut it short."""
        st = L.EnumStats()
        r = _check(self.dg.dev.h, L.lib().psm_evaluate(self.h, cap, ctypes.byref(st)), "psm_evaluate",
                   ok=(L.PSM_OK, L.PSM_ETIMEOUT))
        st.timed_out = r == L.PSM_ETIMEOUT
        return st


@dataclass
class MatchResult:
    count: int
    calls: int
    mem_words: int
    filter_ok: bool
    order: list
    capped: bool
    kernel_ms: float
    probe_bytes: int
    cand_counts: list = field(default_factory=list)
    timed_out: bool = False
    algo_bytes: int = 0


def match(dg: DeviceGraph, q: CSRGraph, method: str = "GQL_GQL_GQL", cap: int = 0) -> MatchResult:
    """The plain path of split_C_test at one query: filter, sort, BuildTable,
    plan, enumerate (StudyPerformance.cpp:1164-1716)."""
    f, o, engine = parse_method(method)
    fname = method.split("_")[0]
    oname = method.split("_")[1]
    run = Query(dg, q)
    try:
        ok = run.filter(fname)
        run.build_tables()
        mem = run.memory_cost_bytes() // 4
        order, _ = run.generate_plan(oname)
        st = run.evaluate(cap)
        calls = int(st.calls) + (int(st.visited_rejects) if engine == "LFTJ" else 0)
        return MatchResult(int(st.count), calls, int(mem), ok, [int(x) for x in order],
                           bool(st.capped), float(st.kernel_ms + st.expand_ms), int(st.probe_bytes),
                           [int(x) for x in run.candidate_counts()], bool(st.timed_out), int(st.algo_bytes))
    finally:
        run.close()


def join_match(dg: DeviceGraph, q: CSRGraph, method: str = "CFL_CFL_EXPLORE", max_rows: int = 0):
    """Filter, tables and plan as the plain path, then the decomposed count."""
    parse_method(method)
    f, o, _ = method.split("_")
    run = Query(dg, q)
    try:
        run.filter(f)
        run.build_tables()
        run.generate_plan(o)
        return run.join_count(max_rows)
    finally:
        run.close()


def split_join_match(dg: DeviceGraph, q: CSRGraph, pattern: str = "STAR", filter: str = "GQL",
                     max_rows: int = 0):
    """Filter and tables as the plain path, then the count over a STAR,
    TWINTWIG or SEED split joined unit by unit."""
    if filter not in L.FILTERS or pattern not in L.SPLITS:
        raise ValueError(f"unsupported filter {filter!r} or split {pattern!r}")
    run = Query(dg, q)
    try:
        run.filter(filter)
        run.build_tables()
        return run.split_join_count(pattern, max_rows)
    finally:
        run.close()


def pack_queries(queries):
    """Concatenated CSR arrays for psm_batch_match."""
    n = np.array([q.num_vertices for q in queries], np.uint32)
    off_base = np.zeros(len(queries), np.uint64)
    nbr_base = np.zeros(len(queries), np.uint64)
    lab_base = np.zeros(len(queries), np.uint64)
    offs, nbrs, labs = [], [], []
    o = nb = lb = 0
    for i, q in enumerate(queries):
        off_base[i], nbr_base[i], lab_base[i] = o, nb, lb
        offs.append(np.asarray(q.offsets, np.uint32))
        nbrs.append(np.asarray(q.neighbors, np.uint32))
        labs.append(np.asarray(q.labels, np.uint32))
        o += q.num_vertices + 1
        nb += q.neighbors.shape[0]
        lb += q.num_vertices
    cat = lambda xs: np.ascontiguousarray(np.concatenate(xs) if xs else np.zeros(1, np.uint32), np.uint32)
    adj = cat(nbrs)
    if adj.size == 0:
        adj = np.zeros(1, np.uint32)
    return n, off_base, nbr_base, lab_base, cat(offs), adj, cat(labs)


def batch_match(dg: DeviceGraph, queries, filter: str = "GQL", order: str = "GQL", cap: int = 0):
    """All queries through the batched stages and one persistent enumeration
    kernel. Returns (counts, calls, status, stats)."""
    n, ob, nb, lb, off, adj, lab = pack_queries(queries)
    m = len(queries)
    counts = np.zeros(max(1, m), np.uint64)
    calls = np.zeros(max(1, m), np.uint64)
    status = np.zeros(max(1, m), np.uint32)
    st = L.BatchStats()
    _check(dg.dev.h, L.lib().psm_batch_match(dg.dev.h, dg.h, m, L.u32p(n), L.u64p(ob), L.u64p(nb),
                                             L.u64p(lb), L.u32p(off), L.u32p(adj), L.u32p(lab),
                                             L.FILTERS[filter], L.ORDERS[order], cap, L.u64p(counts),
                                             L.u64p(calls), L.u32p(status), ctypes.byref(st)),
           "psm_batch_match")
    return counts[:m], calls[:m], status[:m], st


class Batch:
    """The batched form stage by stage (psm_batch): prepare once (filter,
    tables, plans), evaluate as often as needed. `dev` is the context (HIP
    stream) the batch runs on, by default the graph's; any context on the
    graph's GPU may run batches over it."""

    def __init__(self, dg: DeviceGraph, queries, dev: Device | None = None):
        self.dg = dg
        self.dev = dev or dg.dev
        self.m = len(queries)
        self._arrays = pack_queries(queries)
        n, ob, nb, lb, off, adj, lab = self._arrays
        h = ctypes.c_void_p()
        _check(self.dev.h, L.lib().psm_batch_create(self.dev.h, dg.h, self.m, L.u32p(n), L.u64p(ob),
                                                    L.u64p(nb), L.u64p(lb), L.u32p(off), L.u32p(adj),
                                                    L.u32p(lab), ctypes.byref(h)), "psm_batch_create")
        self.h = h

    def prepare(self, filter: str = "GQL", order: str = "GQL"):
        _check(self.dev.h, L.lib().psm_batch_prepare(self.h, L.FILTERS[filter], L.ORDERS[order]),
               "psm_batch_prepare")

    def evaluate(self, cap: int = 0, enum_time_limit: float = -1.0):
        counts = np.zeros(max(1, self.m), np.uint64)
        calls = np.zeros(max(1, self.m), np.uint64)
        status = np.zeros(max(1, self.m), np.uint32)
        st = L.BatchStats()
        _check(self.dev.h, L.lib().psm_batch_evaluate(self.h, cap, enum_time_limit, L.u64p(counts),
                                                         L.u64p(calls), L.u32p(status), ctypes.byref(st)),
               "psm_batch_evaluate")
        return counts[: self.m], calls[: self.m], status[: self.m], st

    def latencies(self) -> np.ndarray:
        """Per query: ms from the batch's creation until its count was final
        in the last evaluate."""
        ms = np.zeros(max(1, self.m), np.float64)
        _check(self.dev.h, L.lib().psm_batch_latencies(self.h, ms.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
               "psm_batch_latencies")
        return ms[: self.m]

    def close(self):
        if getattr(self, "h", None):
            L.lib().psm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


@dataclass
class BatchRun:
    counts: np.ndarray
    calls: np.ndarray
    status: np.ndarray
    stats: object
    latencies_ms: np.ndarray   # per query, from the batch's creation
    start_s: float             # time.perf_counter() at the batch's creation
    end_s: float               # ... and when its evaluate returned


class Pipeline:
    """A batched job of many batches over one resident graph: `depth`
    contexts on the graph's GPU, each with its own HIP stream and host
    thread, take the batches in turn (batch i on context i % depth), so the
    filter, table and plan launches of one batch run while the enumeration of
    another drains on the few lanes its longest queries still occupy. Every
    batch's results are those it has alone; only the launches overlap. The
    contexts take the graph context's launch shape and the given time limit.
    """

    def __init__(self, dg: DeviceGraph, depth: int = 2, time_limit: float = 60.0):
        from . import autotune
        if depth < 1:
            raise ValueError("depth must be at least 1")
        self.dg = dg
        self.devs = [Device(dg.dev.device) for _ in range(depth)]
        cfg = autotune.get_config(dg.dev)
        for d in self.devs:
            autotune.set_config(d, cfg)
            _check(d.h, L.lib().psm_set_time_limit(d.h, time_limit), "psm_set_time_limit")

    def run(self, batches, filter: str = "GQL", order: str = "GQL", cap: int = 0) -> list:
        """Runs every batch (a list of query lists); returns a BatchRun per
        batch, in order. The first error of any context is raised after all
        of them have stopped."""
        import threading
        import time
        out = [None] * len(batches)
        errors = []

        def work(w):
            try:
                for i in range(w, len(batches), len(self.devs)):
                    if errors:
                        return
                    t = time.perf_counter()
                    b = Batch(self.dg, batches[i], dev=self.devs[w])
                    try:
                        b.prepare(filter, order)
                        counts, calls, status, st = b.evaluate(cap)
                        end = time.perf_counter()
                        out[i] = BatchRun(counts, calls, status, st, b.latencies(), t, end)
                    finally:
                        b.close()
            except BaseException as e:   # re-raised on the caller's thread
                errors.append(e)

        threads = [threading.Thread(target=work, args=(w,)) for w in range(len(self.devs))]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            raise errors[0]
        return out

    def kernel_stats_reset(self):
        for d in self.devs:
            L.lib().psm_kernel_stats_reset(d.h)

    def close(self):
        for d in self.devs:
            d.close()
        self.devs = []

    def __del__(self):
        self.close()
