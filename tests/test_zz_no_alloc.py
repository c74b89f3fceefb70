"""psm_evaluate_noalloc, the entry point that allocates no device memory
(its plan, counters and task ring live in the library's fixed device
scratch), against psm_evaluate on the same prepared queries: equal counts,
call counts, visited rejects and search bytes, uncapped and capped, with
both enumeration kernels; two calls in a row, the second with a smaller
input, and repeated calls over the scratch the previous call left behind."""
import ctypes
import os

import pytest

from psm_benchmark_amd.graph_io import load_text

pytestmark = pytest.mark.gpu

HPRD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hprd")


def _prepared(dg, q, method):
    from psm_benchmark_amd.matcher import Query
    f, o, _ = method.split("_")
    run = Query(dg, q)
    run.filter(f)
    run.build_tables()
    run.generate_plan(o)
    return run


def _both(run, cap):
    from psm_benchmark_amd import _lib as L
    a, b = L.EnumStats(), L.EnumStats()
    ra = L.lib().psm_evaluate(run.h, cap, ctypes.byref(a))
    rb = L.lib().psm_evaluate_noalloc(run.h, cap, ctypes.byref(b))
    assert ra == L.PSM_OK and rb == L.PSM_OK, (ra, rb)
    return a, b


def _fields(s):
    return (int(s.count), int(s.calls), int(s.visited_rejects), int(s.algo_bytes), int(s.capped))


@pytest.mark.parametrize("strategy", ["lane", "wave"])
def test_noalloc_equals_evaluate(strategy):
    from psm_benchmark_amd import autotune
    from psm_benchmark_amd import workloads as W
    from psm_benchmark_amd import _lib as L
    from psm_benchmark_amd.matcher import Device, DeviceGraph
    dev = Device(0)
    L.lib().psm_set_time_limit(dev.h, 60.0)
    autotune.set_config(dev, {"dfs_strategy": strategy})
    c1 = W.data_graph("C1")
    hprd = load_text(os.path.join(HPRD, "HPRD-15.txt"))
    dg1, dgh = DeviceGraph(dev, c1), DeviceGraph(dev, hprd)
    cases = [
        (dg1, W.c1_query(c1), "GQL_GQL_GQL", 157_066_360),                  # the larger input first
        (dgh, load_text(os.path.join(HPRD, "Q10-2.txt")), "GQL_GQL_GQL", 1272),   # then a smaller one
        (dgh, load_text(os.path.join(HPRD, "Q10-10.txt")), "CFL_CFL_EXPLORE", 1_015_922),
        (dgh, load_text(os.path.join(HPRD, "Q10-5.txt")), "DPiso_DPiso_LFTJ", 497_052),
    ]
    for dg, q, method, total in cases:
        run = _prepared(dg, q, method)
        for cap in (0, 1000, total, total + 1):
            a, b = _both(run, cap)
            print(strategy, method, q.num_vertices, cap, _fields(a), _fields(b), flush=True)
            assert a.count == b.count == (min(cap, total) if cap else total)
            assert bool(b.capped) == (cap != 0 and cap <= total)
            if not a.capped:
                assert _fields(a) == _fields(b)
        # again over what the previous call left in the scratch
        b2 = L.EnumStats()
        assert L.lib().psm_evaluate_noalloc(run.h, 0, ctypes.byref(b2)) == L.PSM_OK
        assert _fields(b2) == _fields(a)
        run.close()
    dg1.close()
    dgh.close()
    dev.close()


def test_noalloc_needs_a_plan():
    from psm_benchmark_amd import _lib as L
    from psm_benchmark_amd.matcher import Device, DeviceGraph, Query
    dev = Device(0)
    dg = DeviceGraph(dev, load_text(os.path.join(HPRD, "HPRD-15.txt")))
    run = Query(dg, load_text(os.path.join(HPRD, "Q10-2.txt")))
    run.filter("GQL")
    run.build_tables()
    st = L.EnumStats()
    assert L.lib().psm_evaluate_noalloc(run.h, 0, ctypes.byref(st)) == L.PSM_ESTATE
    assert L.lib().psm_evaluate_noalloc(None, 0, ctypes.byref(st)) == L.PSM_EINVAL
    run.close()
    dg.close()
    dev.close()
