"""Serving form: continuous batching over a resident data graph (psm_server_*
in the C ABI), and a seeded Poisson load generator that reports p50/p99
latency and sustained queries/s."""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _lib as L
from .graph_io import CSRGraph
from .matcher import DeviceGraph, _check


class Server:
    """resident: admit prepared batches into one resident enumeration launch
    (True, the library's default) or run one launch per batch (False); None
    keeps the device context's setting (PSM_TUNE_SERVE_RESIDENT)."""

    def __init__(self, dg: DeviceGraph, workers: int = 2, max_batch: int = 64, mem_budget: int = 0,
                 filter: str = "GQL", order: str = "GQL", cap: int = 0, time_limit: float = 60.0,
                 resident: bool | None = None):
        self.dg = dg
        h = ctypes.c_void_p()
        lib = L.lib()
        prev = ctypes.c_double()
        _check(dg.dev.h, lib.psm_get_tuning(dg.dev.h, L.PSM_TUNE_SERVE_RESIDENT, ctypes.byref(prev)),
               "psm_get_tuning")
        if resident is not None:
            _check(dg.dev.h, lib.psm_set_tuning(dg.dev.h, L.PSM_TUNE_SERVE_RESIDENT, 1.0 if resident else 0.0),
                   "psm_set_tuning")
        try:
            _check(dg.dev.h, lib.psm_server_create(dg.dev.h, dg.h, workers, max_batch, mem_budget,
                                                   L.FILTERS[filter], L.ORDERS[order], cap, time_limit,
                                                   ctypes.byref(h)), "psm_server_create")
        finally:
            lib.psm_set_tuning(dg.dev.h, L.PSM_TUNE_SERVE_RESIDENT, prev.value)
        self.h = h
        self._keep = {}

    def submit(self, request_id: int, q: CSRGraph):
        off = np.ascontiguousarray(q.offsets, np.uint32)
        adj = np.ascontiguousarray(q.neighbors, np.uint32)
        lab = np.ascontiguousarray(q.labels, np.uint32)
        if adj.size == 0:
            adj = np.zeros(1, np.uint32)
        _check(self.dg.dev.h, L.lib().psm_server_submit(self.h, request_id, q.num_vertices, L.u32p(off),
                                                        L.u32p(adj), L.u32p(lab)), "psm_server_submit")

    def poll(self, max_results: int = 1024, wait_ms: float = 0.0):
        buf = (L.Result * max_results)()
        n = ctypes.c_uint32()
        _check(self.dg.dev.h, L.lib().psm_server_poll(self.h, buf, max_results, ctypes.byref(n), wait_ms),
               "psm_server_poll")
        return [buf[i] for i in range(n.value)]

    def stats(self):
        a = np.zeros(5, np.uint64)
        L.lib().psm_server_stats(self.h, L.u64p(a))
        return dict(submitted=int(a[0]), completed=int(a[1]), batches=int(a[2]), peak_mem=int(a[3]),
                    pending=int(a[4]))

    def close(self):
        if getattr(self, "h", None):
            L.lib().psm_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def run_poisson(server: Server, queries, rate_qps: float, n_requests: int, seed: int = 0):
    """Submits n_requests drawn cyclically from `queries` at Poisson arrival
    times of the given rate (seeded), collects every result, and returns the
    per-request results plus the latency summary. Latency = completion time -
    arrival time, on the server's clock."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_qps, n_requests)
    arrivals = np.cumsum(gaps)
    results = {}
    t0 = time.perf_counter()
    i = 0
    while len(results) < n_requests:
        now = time.perf_counter() - t0
        while i < n_requests and arrivals[i] <= now:
            server.submit(i, queries[i % len(queries)])
            i += 1
        wait = 0.0 if i < n_requests else 50.0
        if i < n_requests:
            wait = max(0.0, (arrivals[i] - (time.perf_counter() - t0)) * 1e3)
            wait = min(wait, 5.0)
        for r in server.poll(4096, wait):
            results[int(r.request_id)] = r
    wall = time.perf_counter() - t0
    lat = np.array([results[k].done_ms - results[k].submit_ms for k in range(n_requests)])
    fifth = max(1, n_requests // 5)
    # offered_qps: the Poisson stream's own rate (requests / last arrival);
    # a server that keeps up has the same latency in the last fifth of the
    # stream as in the first, one that falls behind a growing one
    return results, dict(rate_qps=rate_qps, requests=n_requests, wall_s=round(wall, 3),
                         offered_qps=round(n_requests / float(arrivals[-1]), 2),
                         sustained_qps=round(n_requests / wall, 2),
                         p50_ms=round(float(np.percentile(lat, 50)), 3),
                         p99_ms=round(float(np.percentile(lat, 99)), 3),
                         p50_first_fifth_ms=round(float(np.percentile(lat[:fifth], 50)), 3),
                         p50_last_fifth_ms=round(float(np.percentile(lat[-fifth:], 50)), 3),
                         mean_batch=round(float(np.mean([results[k].batch_size for k in range(n_requests)])), 2),
                         timeouts=int(sum(1 for k in range(n_requests) if results[k].status == L.PSM_ETIMEOUT)))
