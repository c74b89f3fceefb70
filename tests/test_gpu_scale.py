"""Parity at the benchmark configurations' full sizes, through the C ABI.

* C2: Chung-Lu 1.1M V / 3.0M E, all 512 random-walk queries, cap 1e5, no
  time limit: the batched lane kernel, the batched wavefront kernel, three
  batches through a two-context pipeline and the plain path report the same
  min(cap, total) for every query and none stops early; the oracle agrees
  on a sample; a prepared batch evaluated again at a smaller cap reports
  min(cap, count).
* C3: Chung-Lu 4M V / 35M E, 32 Zipf(1.0) labels, CFL_CFL_EXPLORE: capped
  counts of the 64 cyclic 12-16-vertex queries agree across the lane
  kernel, the wavefront kernel and the plain path; on the join-checkable
  queries of the same graph (finite counts) the uncapped plain count, both
  kernels, the two-unit device join and the STAR split join agree.
* C4: the C3 graph resident, a seeded Poisson stream of random-walk requests
  (4-32 vertices, cap 1e5) through the server: every served count equals the
  count of the same query alone.
* C5: R-MAT 2^24 V / 100M E, 8 uniform labels, 32 dense queries with
  triangles and 4-cliques: capped counts agree across both kernels and the
  plain path, and the oracle agrees on two of them.

Every enumeration runs under a 120 s safety limit and the tests assert that
none reached it.
"""
import numpy as np
import pytest

from oracle_binding import OGraph, ORun

pytestmark = pytest.mark.gpu

SAFETY_S = 120.0


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def dev():
    from psm_benchmark_amd import _lib as L
    from psm_benchmark_amd.matcher import Device
    d = Device(0)
    L.lib().psm_set_time_limit(d.h, SAFETY_S)
    yield d
    d.close()


def _set(dev, strategy):
    from psm_benchmark_amd import autotune
    autotune.set_config(dev, {"dfs_strategy": strategy})


def _batched(dg, qs, filt, order, cap):
    from psm_benchmark_amd import _lib as L
    from psm_benchmark_amd.matcher import Batch
    b = Batch(dg, qs)
    try:
        b.prepare(filt, order)
        counts, calls, status, st = b.evaluate(cap)
    finally:
        b.close()
    assert not (status == L.PSM_ETIMEOUT).any(), np.flatnonzero(status == L.PSM_ETIMEOUT)
    return counts, st


# ---- C2 ----------------------------------------------------------------------
@pytest.fixture(scope="module")
def c2(dev):
    from psm_benchmark_amd import workloads as W
    from psm_benchmark_amd.matcher import DeviceGraph
    g = W.data_graph("C2")
    qs, meta = W.c2_queries(g)
    dg = DeviceGraph(dev, g)
    yield g, qs, dg
    dg.close()


def test_c2_all_queries_lane_wave_plain_agree(dev, c2):
    from psm_benchmark_amd import _lib as L
    from psm_benchmark_amd.matcher import Batch, match
    g, qs, dg = c2
    cap = 100_000
    _set(dev, "lane")
    lane, st = _batched(dg, qs, "GQL", "GQL", cap)
    _say("C2 lane", int(lane.sum()), f"{st.enum_ms:.0f} ms")
    _set(dev, "wave")
    wave, st = _batched(dg, qs, "GQL", "GQL", cap)
    _say("C2 wave", int(wave.sum()), f"{st.enum_ms:.0f} ms")
    assert np.array_equal(lane, wave), np.flatnonzero(lane != wave)[:10]
    _set(dev, "lane")
    for i, q in enumerate(qs):
        r = match(dg, q, "GQL_GQL_GQL", cap=cap)
        assert not r.timed_out and r.count == lane[i], i
    # three batches through a two-context pipeline (overlapping launches)
    from psm_benchmark_amd.matcher import Pipeline
    p = Pipeline(dg, 2, time_limit=SAFETY_S)
    try:
        for r in p.run([qs] * 3, "GQL", "GQL", cap):
            assert not (r.status == L.PSM_ETIMEOUT).any()
            assert np.array_equal(r.counts, lane)
    finally:
        p.close()
    # a prepared batch evaluated again at a smaller cap
    b = Batch(dg, qs[:64])
    b.prepare("GQL", "GQL")
    small, _, status, _ = b.evaluate(1000)
    b.close()
    assert not (status == L.PSM_ETIMEOUT).any()
    assert np.array_equal(small, np.minimum(lane[:64], 1000))


def test_c2_sample_matches_oracle(dev, c2):
    from psm_benchmark_amd.matcher import match
    g, qs, dg = c2
    og = OGraph(g)
    _set(dev, "lane")
    for i in range(0, 512, 97):
        ref = ORun(og, OGraph(qs[i]), "GQL_GQL_GQL", cap=100_000, threads=0)
        r = match(dg, qs[i], "GQL_GQL_GQL", cap=100_000)
        _say("C2 query", i, "n", qs[i].num_vertices, "count", r.count, "oracle", ref.count)
        assert r.count == ref.count and r.cand_counts == ref.cand_counts and r.order == ref.order, i


# ---- C3 / C4 (one resident graph) --------------------------------------------
@pytest.fixture(scope="module")
def c3(dev):
    from psm_benchmark_amd import workloads as W
    from psm_benchmark_amd.matcher import DeviceGraph
    g = W.data_graph("C3")
    _say("C3 graph", g.num_vertices, g.num_edges)
    dg = DeviceGraph(dev, g)
    yield g, dg
    dg.close()


def test_c3_capped_lane_wave_plain_agree(dev, c3):
    from psm_benchmark_amd import workloads as W
    from psm_benchmark_amd.matcher import match
    g, dg = c3
    qs, _ = W.c3_queries(g)
    cap = 100_000
    got = {}
    for strategy in ("lane", "wave"):
        _set(dev, strategy)
        got[strategy], st = _batched(dg, qs, "CFL", "CFL", cap)
        _say("C3", strategy, int(got[strategy].sum()), f"filter {st.filter_ms:.0f} ms enum {st.enum_ms:.0f} ms")
    assert np.array_equal(got["lane"], got["wave"])
    _set(dev, "lane")
    for i in range(0, len(qs), 4):
        r = match(dg, qs[i], "CFL_CFL_EXPLORE", cap=cap)
        assert not r.timed_out and r.count == got["lane"][i], i


def test_c3_join_queries_uncapped_all_forms_agree(dev, c3):
    from psm_benchmark_amd import workloads as W
    from psm_benchmark_amd import _lib as L
    from psm_benchmark_amd.matcher import join_match, match, split_join_match
    g, dg = c3
    jq, _ = W.c3_join_queries(g)
    counts = {}
    for strategy in ("lane", "wave"):
        _set(dev, strategy)
        counts[strategy], _ = _batched(dg, jq, "CFL", "CFL", 0)
    assert np.array_equal(counts["lane"], counts["wave"])
    _set(dev, "lane")
    checked = 0
    for i, q in enumerate(jq):
        r = match(dg, q, "CFL_CFL_EXPLORE")
        assert not r.timed_out and r.count == counts["lane"][i], i
        for name, fn in (("join", lambda: join_match(dg, q, "CFL_CFL_EXPLORE")),
                         ("star", lambda: split_join_match(dg, q, "STAR", "CFL"))):
            try:
                c = int(fn().count)
            except L.PSMError as e:
                # a unit with more rows than the row bound (64M) is refused
                # with PSM_ENOMEM and compared no further
                assert e.code == L.PSM_ENOMEM, (i, name, e)
                _say("C3 join query", i, name, "refused:", e)
                continue
            _say("C3 join query", i, "n", q.num_vertices, "count", r.count, name, c)
            assert c == r.count, (i, name)
            checked += 1
    assert checked >= len(jq), checked


def test_c4_served_equals_alone(dev, c3):
    from psm_benchmark_amd import workloads as W
    from psm_benchmark_amd import _lib as L
    from psm_benchmark_amd.matcher import match
    from psm_benchmark_amd.serving import Server, run_poisson
    g, dg = c3
    qs, _ = W.c2_queries(g, count=96, seed=44)
    cap = 100_000
    _set(dev, "lane")
    alone = []
    for q in qs:
        r = match(dg, q, "GQL_GQL_GQL", cap=cap)
        assert not r.timed_out
        alone.append(r.count)
    # served through the resident launch (lane kernel) and through one
    # launch per batch
    for strategy, resident in (("lane", True), ("wave", False)):
        _set(dev, strategy)
        srv = Server(dg, workers=4, max_batch=32, filter="GQL", order="GQL", cap=cap, time_limit=SAFETY_S,
                     resident=resident)
        try:
            res, summary = run_poisson(srv, qs, rate_qps=1000.0, n_requests=192, seed=4)
        finally:
            srv.close()
        _say("C4", strategy, "resident" if resident else "per batch", summary)
        for k, r in res.items():
            assert r.status == L.PSM_OK, (strategy, resident, k)
            assert r.count == alone[k % len(qs)], (strategy, resident, k)
    _set(dev, "lane")


# ---- C5 ----------------------------------------------------------------------
@pytest.fixture(scope="module")
def c5(dev):
    from psm_benchmark_amd import workloads as W
    from psm_benchmark_amd.matcher import DeviceGraph
    g = W.data_graph("C5")
    _say("C5 graph", g.num_vertices, g.num_edges)
    dg = DeviceGraph(dev, g)
    yield g, dg
    dg.close()


def test_c5_capped_lane_wave_plain_oracle_agree(dev, c5):
    from psm_benchmark_amd import workloads as W
    from psm_benchmark_amd.matcher import match
    g, dg = c5
    qs, _ = W.c5_queries(g)
    cap = 10_000
    got = {}
    for strategy in ("lane", "wave"):
        _set(dev, strategy)
        got[strategy], st = _batched(dg, qs, "GQL", "GQL", cap)
        _say("C5", strategy, int(got[strategy].sum()), f"filter {st.filter_ms:.0f} ms enum {st.enum_ms:.0f} ms")
    assert np.array_equal(got["lane"], got["wave"])
    _set(dev, "lane")
    for i in range(0, len(qs), 4):
        r = match(dg, qs[i], "GQL_GQL_GQL", cap=cap)
        assert not r.timed_out and r.count == got["lane"][i], i
    og = OGraph(g)
    for i in (0, 1):
        ref = ORun(og, OGraph(qs[i]), "GQL_GQL_GQL", cap=1000, threads=0)
        r = match(dg, qs[i], "GQL_GQL_GQL", cap=1000)
        _say("C5 query", i, "count", r.count, "oracle", ref.count, "cands", ref.cand_counts)
        assert r.count == ref.count and r.cand_counts == ref.cand_counts, i
