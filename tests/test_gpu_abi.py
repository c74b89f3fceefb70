"""Error behaviour and state of the C ABI entry points that steer the device
path (launch shape, staged batches), through libpsm.so on the GPU."""
import ctypes
import os

import pytest

from psm_benchmark_amd.graph_io import load_text

pytestmark = pytest.mark.gpu

HPRD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hprd")


@pytest.fixture(scope="module")
def dg():
    from psm_benchmark_amd.matcher import Device, DeviceGraph
    dev = Device(0)
    g = DeviceGraph(dev, load_text(os.path.join(HPRD, "HPRD-15.txt")))
    yield g
    g.close()
    dev.close()


def test_tuning_round_trip_and_rejection(dg):
    from psm_benchmark_amd import _lib as L
    from psm_benchmark_amd import autotune
    h = dg.dev.h
    before = autotune.get_config(dg.dev)
    assert before["dfs_strategy"] in ("lane", "wave")
    cfg = {"dfs_strategy": "wave", "dfs_waves_per_cu": 6, "expand_tasks_per_lane": 3.0}
    autotune.set_config(dg.dev, cfg)
    got = autotune.get_config(dg.dev)
    assert {k: got[k] for k in cfg} == {**cfg, "dfs_waves_per_cu": 6.0}
    autotune.set_config(dg.dev, {"claim_divisor": 3, "dfs_block_lanes": 256})
    got = autotune.get_config(dg.dev)
    assert got["claim_divisor"] == 3.0 and got["dfs_block_lanes"] == 256.0
    autotune.set_config(dg.dev, {"intersect": "bsearch", "expand_depth": 5})
    got = autotune.get_config(dg.dev)
    assert got["intersect"] == "bsearch" and got["expand_depth"] == 5.0
    for key, bad in ((L.PSM_TUNE_DFS_WAVES_PER_CU, 0), (L.PSM_TUNE_DFS_WAVES_PER_CU, 33),
                     (L.PSM_TUNE_EXPAND_TASKS_PER_LANE, 0.0), (L.PSM_TUNE_DFS_STRATEGY, 7),
                     (L.PSM_TUNE_CLAIM_DIVISOR, 0), (L.PSM_TUNE_QUEUE_SLOTS, -1),
                     (L.PSM_TUNE_DFS_BLOCK_LANES, 32), (L.PSM_TUNE_DFS_BLOCK_LANES, 512),
                     (L.PSM_TUNE_EXPAND_DEPTH, 0), (L.PSM_TUNE_EXPAND_DEPTH, 32),
                     (L.PSM_TUNE_INTERSECT, 2), (99, 1)):
        assert L.lib().psm_set_tuning(h, key, float(bad)) == L.PSM_EINVAL
    v = ctypes.c_double()
    assert L.lib().psm_get_tuning(h, 99, ctypes.byref(v)) == L.PSM_EINVAL
    autotune.set_config(dg.dev, before)


def test_staged_batch_state_and_repeat(dg):
    """evaluate before prepare is a state error; a prepared batch evaluates
    twice with the same results; the counts equal the reference's logs."""
    import json
    from psm_benchmark_amd import _lib as L
    from psm_benchmark_amd.matcher import Batch
    recs = [r for r in json.load(open(os.path.join(HPRD, "golden.json")))
            if r["threads"] == 1 and r["filter"] == "GQL" and r["engine"] == "GQL"][:4]
    qs = [load_text(os.path.join(HPRD, r["query"] + ".txt")) for r in recs]
    b = Batch(dg, qs)
    st = L.BatchStats()
    import numpy as np
    c = np.zeros(len(qs), np.uint64)
    r = L.lib().psm_batch_evaluate(b.h, 0, -1.0, L.u64p(c), None, None, ctypes.byref(st))
    assert r == L.PSM_ESTATE
    b.prepare("GQL", "GQL")
    c1, k1, s1, _ = b.evaluate(0, 0.0)
    c2, k2, s2, _ = b.evaluate(0, 0.0)
    b.close()
    assert c1.tolist() == c2.tolist() == [r["embeddings"] for r in recs]
    assert k1.tolist() == k2.tolist() == [r["calls"] for r in recs]


def test_caller_supplied_plan_and_bfs_trees(dg):
    """psm_set_plan: any order gives the golden count (extensions intersect
    every earlier neighbour's row), bad orders and pivots are rejected;
    psm_query_bfs_tree returns the CFL filter's and the CFL order's BFS
    trees (a BFS order with parents earlier and adjacent, levels one more)."""
    import json
    import numpy as np
    from psm_benchmark_amd import _lib as L
    from psm_benchmark_amd.matcher import Query
    rec = next(r for r in json.load(open(os.path.join(HPRD, "golden.json")))
               if r["query"] == "Q10-3" and r["threads"] == 1 and r["filter"] == "CFL")
    q = load_text(os.path.join(HPRD, "Q10-3.txt"))
    run = Query(dg, q)
    try:
        run.filter("CFL")
        run.build_tables()
        n = q.num_vertices
        lib = L.lib()
        order, pivot = run.generate_plan("CFL")
        off = q.offsets.astype(np.int64)
        for which in (L.PSM_TREE_FILTER, L.PSM_TREE_PLAN):
            bo, par, lev = (np.zeros(n, np.uint32) for _ in range(3))
            assert lib.psm_query_bfs_tree(run.h, which, L.u32p(bo), L.u32p(par), L.u32p(lev)) == L.PSM_OK
            assert sorted(bo.tolist()) == list(range(n))
            pos = {int(v): i for i, v in enumerate(bo)}
            assert lev[bo[0]] == 0
            for v in bo[1:]:
                p = int(par[v])
                assert pos[p] < pos[int(v)] and int(v) in q.neighbors[off[p]:off[p + 1]]
                assert lev[v] == lev[p] + 1
        assert run.evaluate().calls == rec["calls"]
        assert lib.psm_set_plan(run.h, L.u32p(np.arange(n, dtype=np.uint32)), None) == L.PSM_OK
        assert run.evaluate().count == rec["embeddings"]
        # a plan set by the caller has no BFS tree
        z = np.zeros(n, np.uint32)
        assert lib.psm_query_bfs_tree(run.h, L.PSM_TREE_PLAN, L.u32p(z), None, None) == L.PSM_ESTATE
        rev = np.ascontiguousarray(order[::-1])
        assert lib.psm_set_plan(run.h, L.u32p(rev), None) == L.PSM_OK
        assert run.evaluate().count == rec["embeddings"]
        o, pv = np.ascontiguousarray(order), np.ascontiguousarray(pivot)
        assert lib.psm_set_plan(run.h, L.u32p(o), L.u32p(pv)) == L.PSM_OK
        st = run.evaluate()
        assert (st.count, st.calls) == (rec["embeddings"], rec["calls"])
        assert lib.psm_set_plan(run.h, L.u32p(np.zeros(n, np.uint32)), None) == L.PSM_EINVAL
        badp = pv.copy()
        badp[1] = order[-1]
        assert lib.psm_set_plan(run.h, L.u32p(o), L.u32p(np.ascontiguousarray(badp))) == L.PSM_EINVAL
    finally:
        run.close()
