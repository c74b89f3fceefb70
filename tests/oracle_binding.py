"""ctypes binding to the CPU oracle (oracle/liboracle.so).

Test infrastructure: the oracle is the checker the GPU path is compared
against; nothing in psm_benchmark_amd imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
SRC_PATH = os.path.join(ROOT, "oracle", "psm_oracle.cpp")


class OrcResult(ctypes.Structure):
    _fields_ = [
        ("count", ctypes.c_uint64),
        ("calls", ctypes.c_uint64),
        ("mem_words", ctypes.c_uint64),
        ("n", ctypes.c_uint32),
        ("filter_ok", ctypes.c_uint32),
        ("cand_count", ctypes.c_uint32 * 64),
        ("order", ctypes.c_uint32 * 64),
        ("pivot", ctypes.c_uint32 * 64),
        ("depth_nodes", ctypes.c_uint64 * 64),
        ("algo_bytes", ctypes.c_uint64),
    ]


class OrcTimes(ctypes.Structure):
    _fields_ = [
        ("filter_s", ctypes.c_double),
        ("build_s", ctypes.c_double),
        ("plan_s", ctypes.c_double),
        ("enum_s", ctypes.c_double),
        ("total_s", ctypes.c_double),
        ("threads", ctypes.c_uint32),
        ("timed_out", ctypes.c_uint32),
    ]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(SRC_PATH):
        subprocess.check_call(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-o", LIB_PATH, SRC_PATH,
                               "-lpthread"])
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        l = ctypes.CDLL(build())
        P = ctypes.c_void_p
        U32P = ctypes.POINTER(ctypes.c_uint32)
        l.orc_graph_new.restype = P
        l.orc_graph_new.argtypes = [ctypes.c_uint32, U32P, U32P, U32P]
        l.orc_graph_free.argtypes = [P]
        l.orc_run.restype = P
        l.orc_run.argtypes = [P, P, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                              ctypes.c_uint64, ctypes.POINTER(OrcResult)]
        l.orc_run_mt.restype = P
        l.orc_run_mt.argtypes = [P, P, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                 ctypes.c_uint64, ctypes.c_uint32, ctypes.c_double,
                                 ctypes.POINTER(OrcResult), ctypes.POINTER(OrcTimes)]
        l.orc_run_free.argtypes = [P]
        l.orc_run_cands.argtypes = [P, ctypes.c_uint32, U32P]
        l.orc_run_table.restype = ctypes.c_int64
        l.orc_run_table.argtypes = [P, ctypes.c_uint32, ctypes.c_uint32, U32P, U32P]
        l.orc_nlf_count.restype = ctypes.c_uint32
        l.orc_nlf_count.argtypes = [P, P, ctypes.c_uint32]
        _lib = l
    return _lib


def _u32p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


class OGraph:
    def __init__(self, g):
        self._arrs = (np.ascontiguousarray(g.offsets, np.uint32),
                      np.ascontiguousarray(g.neighbors, np.uint32),
                      np.ascontiguousarray(g.labels, np.uint32))
        o, a, l_ = self._arrs
        if a.size == 0:
            a = np.zeros(1, np.uint32)
        self.h = lib().orc_graph_new(l_.shape[0], _u32p(o), _u32p(a), _u32p(l_))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_graph_free(self.h)
            self.h = None


class ORun:
    """One plain-path run: filter -> sort -> BuildTable -> order -> enumerate."""

    def __init__(self, data: OGraph, query: OGraph, method: str = "GQL_GQL_GQL", cap: int = 0,
                 threads: int = 1, time_limit: float = 0.0):
        """threads: enumeration workers (0: every hardware thread);
        time_limit > 0: seconds from the start of filtering (timed_out)."""
        f, o, e = method.split("_")
        self.res = OrcResult()
        self.times = OrcTimes()
        self.h = lib().orc_run_mt(data.h, query.h, f.encode(), o.encode(), e.encode(), cap, threads,
                                  time_limit, ctypes.byref(self.res), ctypes.byref(self.times))
        if not self.h:
            raise ValueError(f"oracle rejected method {method}")
        self.n = self.res.n

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_run_free(self.h)
            self.h = None

    count = property(lambda s: int(s.res.count))
    calls = property(lambda s: int(s.res.calls))
    mem_words = property(lambda s: int(s.res.mem_words))
    order = property(lambda s: [int(s.res.order[i]) for i in range(s.n)])
    pivot = property(lambda s: [int(s.res.pivot[i]) for i in range(s.n)])
    cand_counts = property(lambda s: [int(s.res.cand_count[i]) for i in range(s.n)])
    depth_nodes = property(lambda s: [int(s.res.depth_nodes[i]) for i in range(s.n)])
    algo_bytes = property(lambda s: int(s.res.algo_bytes))
    timed_out = property(lambda s: bool(s.times.timed_out))

    def candidates(self, u: int) -> np.ndarray:
        out = np.zeros(max(1, self.res.cand_count[u]), np.uint32)
        lib().orc_run_cands(self.h, u, _u32p(out))
        return out[: self.res.cand_count[u]]

    def table(self, a: int, b: int):
        m = lib().orc_run_table(self.h, a, b, None, None)
        if m < 0:
            return None
        off = np.zeros(self.res.cand_count[a] + 1, np.uint32)
        idx = np.zeros(max(1, m), np.uint32)
        lib().orc_run_table(self.h, a, b, _u32p(off), _u32p(idx))
        return off, idx[:m]
