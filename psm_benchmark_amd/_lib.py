"""ctypes binding to libpsm.so, the C ABI declared in include/psm_abi.h.

The library is built in-tree (psm_benchmark_amd/_build.py). Loading fails
loudly when it is missing: there is no CPU fallback for the device path.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpsm.so")

PSM_OK = 0
PSM_EINVAL = 1
PSM_EEMPTY = 2
PSM_ENOMEM = 3
PSM_ETIMEOUT = 4
PSM_EOVERFLOW = 5
PSM_ENOTFOUND = 6
PSM_ESTATE = 7

FILTERS = {"LDF": 0, "NLF": 1, "GQL": 2, "CFL": 3, "DPiso": 4}
ORDERS = {"GQL": 0, "CFL": 1, "DPiso": 2}

U32P = ctypes.POINTER(ctypes.c_uint32)
U64P = ctypes.POINTER(ctypes.c_uint64)
VP = ctypes.c_void_p


class EnumStats(ctypes.Structure):
    _fields_ = [
        ("count", ctypes.c_uint64),
        ("calls", ctypes.c_uint64),
        ("capped", ctypes.c_uint32),
        ("tasks", ctypes.c_uint32),
        ("kernel_ms", ctypes.c_double),
        ("expand_ms", ctypes.c_double),
        ("probe_bytes", ctypes.c_uint64),
        ("algo_bytes", ctypes.c_uint64),
        ("visited_rejects", ctypes.c_uint64),
    ]


class BatchStats(ctypes.Structure):
    _fields_ = [
        ("filter_ms", ctypes.c_double),
        ("table_ms", ctypes.c_double),
        ("plan_ms", ctypes.c_double),
        ("enum_ms", ctypes.c_double),
        ("total_ms", ctypes.c_double),
        ("tasks", ctypes.c_uint64),
        ("probe_bytes", ctypes.c_uint64),
        ("dfs_lane_steps", ctypes.c_uint64),
        ("dfs_lane_iters", ctypes.c_uint64),
        ("algo_bytes", ctypes.c_uint64),
    ]


class JoinStats(ctypes.Structure):
    _fields_ = [
        ("count", ctypes.c_uint64),
        ("rows_a", ctypes.c_uint64),
        ("rows_b", ctypes.c_uint64),
        ("keys", ctypes.c_uint32),
        ("enum_ms", ctypes.c_double),
        ("join_ms", ctypes.c_double),
    ]


class SplitJoinStats(ctypes.Structure):
    _fields_ = [
        ("count", ctypes.c_uint64),
        ("units", ctypes.c_uint32),
        ("pad", ctypes.c_uint32),
        ("unit_rows", ctypes.c_uint64),
        ("max_inter_rows", ctypes.c_uint64),
        ("enum_ms", ctypes.c_double),
        ("join_ms", ctypes.c_double),
        ("calls", ctypes.c_uint64),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("request_id", ctypes.c_uint64),
        ("count", ctypes.c_uint64),
        ("calls", ctypes.c_uint64),
        ("status", ctypes.c_int32),
        ("batch_size", ctypes.c_uint32),
        ("submit_ms", ctypes.c_double),
        ("start_ms", ctypes.c_double),
        ("done_ms", ctypes.c_double),
    ]


PSM_TUNE_DFS_WAVES_PER_CU = 0
PSM_TUNE_EXPAND_TASKS_PER_LANE = 1
PSM_TUNE_DFS_STRATEGY = 2
PSM_TUNE_QUEUE_SLOTS = 3
PSM_TUNE_CLAIM_DIVISOR = 4
PSM_TUNE_DFS_BLOCK_LANES = 5
PSM_TUNE_EXPAND_DEPTH = 6
PSM_TUNE_INTERSECT = 7
PSM_TUNE_SERVE_RESIDENT = 8
PSM_INTERSECT_ADAPTIVE = 0
PSM_INTERSECT_BSEARCH = 1
PSM_TREE_FILTER = 0
PSM_TREE_PLAN = 1
PSM_DFS_LANE = 0
PSM_DFS_WAVE = 1
SPLITS = {"STAR": 0, "TWINTWIG": 1, "SEED": 2}

# name -> (restype, argtypes); every symbol include/psm_abi.h declares
SIGNATURES = {
    "psm_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(VP)]),
    "psm_ctx_destroy": (ctypes.c_int, [VP]),
    "psm_last_error": (ctypes.c_char_p, [VP]),
    "psm_kernel_stats": (ctypes.c_int, [VP, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    "psm_kernel_stats_reset": (ctypes.c_int, [VP]),
    "psm_set_time_limit": (ctypes.c_int, [VP, ctypes.c_double]),
    "psm_set_tuning": (ctypes.c_int, [VP, ctypes.c_int, ctypes.c_double]),
    "psm_get_tuning": (ctypes.c_int, [VP, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    "psm_graph_upload": (ctypes.c_int, [VP, ctypes.c_uint32, U32P, U32P, U32P, ctypes.POINTER(VP)]),
    "psm_graph_destroy": (ctypes.c_int, [VP]),
    "psm_graph_info": (ctypes.c_int, [VP, U64P]),
    "psm_query_create": (ctypes.c_int, [VP, VP, ctypes.c_uint32, U32P, U32P, U32P, ctypes.POINTER(VP)]),
    "psm_query_destroy": (ctypes.c_int, [VP]),
    "psm_filter": (ctypes.c_int, [VP, ctypes.c_int]),
    "psm_query_candidate_counts": (ctypes.c_int, [VP, U32P]),
    "psm_query_candidates": (ctypes.c_int, [VP, ctypes.c_uint32, U32P]),
    "psm_build_tables": (ctypes.c_int, [VP]),
    "psm_query_table": (ctypes.c_int, [VP, ctypes.c_uint32, ctypes.c_uint32, U32P, U32P, U64P]),
    "psm_memory_cost_bytes": (ctypes.c_int, [VP, U64P]),
    "psm_generate_plan": (ctypes.c_int, [VP, ctypes.c_int, U32P, U32P]),
    "psm_set_plan": (ctypes.c_int, [VP, U32P, U32P]),
    "psm_query_bfs_tree": (ctypes.c_int, [VP, ctypes.c_int, U32P, U32P, U32P]),
    "psm_evaluate": (ctypes.c_int, [VP, ctypes.c_uint64, ctypes.POINTER(EnumStats)]),
    "psm_evaluate_noalloc": (ctypes.c_int, [VP, ctypes.c_uint64, ctypes.POINTER(EnumStats)]),
    "psm_join_count": (ctypes.c_int, [VP, ctypes.c_uint64, ctypes.POINTER(JoinStats)]),
    "psm_split_join_count": (ctypes.c_int, [VP, ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(SplitJoinStats)]),
    "psm_match": (ctypes.c_int, [VP, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(EnumStats)]),
    "psm_batch_match": (ctypes.c_int, [VP, VP, ctypes.c_uint32, U32P, U64P, U64P, U64P, U32P, U32P, U32P,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_uint64, U64P, U64P, U32P,
                                       ctypes.POINTER(BatchStats)]),
    "psm_batch_create": (ctypes.c_int, [VP, VP, ctypes.c_uint32, U32P, U64P, U64P, U64P, U32P, U32P, U32P,
                                        ctypes.POINTER(VP)]),
    "psm_batch_prepare": (ctypes.c_int, [VP, ctypes.c_int, ctypes.c_int]),
    "psm_batch_evaluate": (ctypes.c_int, [VP, ctypes.c_uint64, ctypes.c_double, U64P, U64P, U32P,
                                          ctypes.POINTER(BatchStats)]),
    "psm_batch_latencies": (ctypes.c_int, [VP, ctypes.POINTER(ctypes.c_double)]),
    "psm_batch_destroy": (ctypes.c_int, [VP]),
    "psm_server_create": (ctypes.c_int, [VP, VP, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_uint64, ctypes.c_double, ctypes.POINTER(VP)]),
    "psm_server_submit": (ctypes.c_int, [VP, ctypes.c_uint64, ctypes.c_uint32, U32P, U32P, U32P]),
    "psm_server_poll": (ctypes.c_int, [VP, ctypes.POINTER(Result), ctypes.c_uint32,
                                       ctypes.POINTER(ctypes.c_uint32), ctypes.c_double]),
    "psm_server_stats": (ctypes.c_int, [VP, U64P]),
    "psm_server_destroy": (ctypes.c_int, [VP]),
    "psm_text_graph_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint32, ctypes.POINTER(VP)]),
    "psm_text_graph_info": (ctypes.c_int, [VP, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint64)]),
    "psm_text_graph_offsets": (ctypes.POINTER(ctypes.c_uint32), [VP]),
    "psm_text_graph_neighbors": (ctypes.POINTER(ctypes.c_uint32), [VP]),
    "psm_text_graph_labels": (ctypes.POINTER(ctypes.c_uint32), [VP]),
    "psm_text_graph_free": (None, [VP]),
    "psm_text_graph_last_error": (ctypes.c_char_p, []),
    "psm_gen_chung_lu": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double, ctypes.c_uint64,
                                        U32P, U32P]),
    "psm_gen_rmat": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_uint64, U32P, U32P]),
    "psm_gen_labels": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double, ctypes.c_uint64, U32P]),
    "psm_gen_query": (ctypes.c_int64, [ctypes.c_uint32, U32P, U32P, ctypes.c_uint32, ctypes.c_int,
                                       ctypes.c_uint64, U32P, U32P, ctypes.c_uint64]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """Loads libpsm.so; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m psm_benchmark_amd._build` "
                "(the matcher has no CPU fallback)")
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


class PSMError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"psm status {code}: {msg}")
        self.code = code


def u32p(a):
    return a.ctypes.data_as(U32P)


def u64p(a):
    return a.ctypes.data_as(U64P)
