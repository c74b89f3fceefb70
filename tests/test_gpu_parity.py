"""Parity of the HIP path (libpsm.so, through its C ABI) with the reference's
golden vectors and the CPU oracle. Bit-exact: counts, call counts, memory
checksums, candidate sets and candidate-edge tables."""
import ctypes
import itertools
import json
import math
import os
import time

import numpy as np
import pytest

from brute import count_embeddings, random_graph, random_connected_query
from oracle_binding import OGraph, ORun
from psm_benchmark_amd.graph_io import csr_from_edges, load_text

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
HPRD = os.path.join(HERE, "golden", "hprd")


@pytest.fixture(scope="module", params=["lane", "wave"])
def dev(request):
    """Every device test runs with both enumeration kernels: one search per
    lane, and one search per wavefront with chunked cooperative tests."""
    from psm_benchmark_amd import autotune
    from psm_benchmark_amd.matcher import Device
    d = Device(0)
    autotune.set_config(d, {"dfs_strategy": request.param})
    print("dfs strategy", request.param, flush=True)
    yield d
    d.close()


@pytest.fixture(scope="module")
def hprd(dev):
    from psm_benchmark_amd.matcher import DeviceGraph
    g = load_text(os.path.join(HPRD, "HPRD-15.txt"))
    return g, DeviceGraph(dev, g), OGraph(g)


def golden(methods):
    recs = json.load(open(os.path.join(HPRD, "golden.json")))
    return [r for r in recs if r["threads"] == 1 and f'{r["filter"]}_{r["order"]}_{r["engine"]}' in methods]


@pytest.mark.parametrize("rec", golden(("GQL_GQL_GQL", "CFL_CFL_EXPLORE")),
                         ids=lambda r: f'{r["query"]}-{r["filter"]}')
def test_hprd_golden_on_device(hprd, rec):
    from psm_benchmark_amd.matcher import match
    _, dg, _ = hprd
    q = load_text(os.path.join(HPRD, rec["query"] + ".txt"))
    r = match(dg, q, f'{rec["filter"]}_{rec["order"]}_{rec["engine"]}')
    print(rec["query"], rec["filter"], r.count, r.calls, r.mem_words, flush=True)
    assert (r.count, r.calls, r.mem_words) == (rec["embeddings"], rec["calls"], rec["mem_words"])


@pytest.mark.parametrize("rec", golden(("DPiso_DPiso_LFTJ",)), ids=lambda r: r["query"])
def test_hprd_dpiso_golden_on_device(hprd, rec):
    # the LFTJ Call Count (assignments + extensions whose data vertex is
    # already used) against the reference's log, and the assignments alone
    # against the oracle's GQL engine on the same DPiso plan
    from psm_benchmark_amd.matcher import match
    _, dg, og = hprd
    q = load_text(os.path.join(HPRD, rec["query"] + ".txt"))
    r = match(dg, q, "DPiso_DPiso_LFTJ")
    g = match(dg, q, "DPiso_DPiso_GQL")
    ref = ORun(og, OGraph(q), "DPiso_DPiso_GQL")
    print(rec["query"], r.count, r.calls, r.mem_words, r.order, flush=True)
    assert (r.count, r.calls, r.mem_words) == (rec["embeddings"], rec["calls"], rec["mem_words"])
    assert r.order == ref.order and g.calls == ref.calls


@pytest.mark.parametrize("qname", ["Q10-2", "Q10-5", "Q10-10"])
def test_lftj_calls_match_oracle(hprd, qname):
    # GQL order + LFTJ engine (the oracle's LFTJ calls are pinned by the
    # reference binary's GQL_GQL_LFTJ vectors)
    from psm_benchmark_amd.matcher import match
    _, dg, og = hprd
    q = load_text(os.path.join(HPRD, qname + ".txt"))
    r = match(dg, q, "GQL_GQL_LFTJ")
    ref = ORun(og, OGraph(q), "GQL_GQL_LFTJ")
    assert (r.count, r.calls) == (ref.count, ref.calls)
    assert r.calls > match(dg, q, "GQL_GQL_GQL").calls


@pytest.mark.parametrize("qname", ["Q10-1", "Q10-4", "Q10-9"])
def test_dpiso_plan_matches_oracle(hprd, qname):
    from psm_benchmark_amd.matcher import Query
    _, dg, og = hprd
    q = load_text(os.path.join(HPRD, qname + ".txt"))
    ref = ORun(og, OGraph(q), "DPiso_DPiso_LFTJ")
    run = Query(dg, q)
    run.filter("DPiso")
    run.build_tables()
    o, p = run.generate_plan("DPiso")
    assert o.tolist() == ref.order
    assert p.tolist()[1:] == ref.pivot[1:]


@pytest.mark.parametrize("method", ["LDF", "NLF", "GQL", "CFL", "DPiso"])
@pytest.mark.parametrize("qname", ["Q10-1", "Q10-3", "Q10-8"])
def test_candidates_and_tables_match_oracle(hprd, method, qname):
    from psm_benchmark_amd.matcher import Query
    _, dg, og = hprd
    q = load_text(os.path.join(HPRD, qname + ".txt"))
    ref = ORun(og, OGraph(q), f"{method}_GQL_GQL")
    run = Query(dg, q)
    run.filter(method)
    assert run.candidate_counts().tolist() == ref.cand_counts
    for u in range(q.num_vertices):
        assert np.array_equal(run.candidates(u), ref.candidates(u)), u
    run.build_tables()
    for a, b in itertools.product(range(q.num_vertices), repeat=2):
        t, rt = run.table(a, b), ref.table(a, b)
        assert (t is None) == (rt is None)
        if t is not None:
            assert np.array_equal(t[0], rt[0]) and np.array_equal(t[1], rt[1]), (a, b)
    assert run.memory_cost_bytes() // 4 == ref.mem_words
    o, _ = run.generate_plan("GQL")
    assert o.tolist() == ref.order


@pytest.mark.parametrize("n,labels,dense", [(12, 1, 1), (24, 1, 1), (36, 1, 1), (40, 1, 2), (40, 2, 2)])
def test_gql_candidates_with_large_label_groups(dev, n, labels, dense):
    """GQL refinement where many query vertices share a label: the
    membership words are 2, 4 and 8 bytes wide (label groups above 8, 16
    and 32 vertices); dense queries on a power-law graph make the
    refinement drop about half of the candidates. Candidates are compared
    with the oracle's."""
    from psm_benchmark_amd import gen
    from psm_benchmark_amd.matcher import DeviceGraph, Query
    data = gen.chung_lu_graph(2000, 16000, labels, 1.0, seed=n + labels)
    q = gen.query(data, n, dense, 3 * n + labels)
    dg = DeviceGraph(dev, data)
    ref = ORun(OGraph(data), OGraph(q), "GQL_GQL_GQL", cap=1, time_limit=0.5)
    run = Query(dg, q)
    run.filter("GQL")
    assert run.candidate_counts().tolist() == ref.cand_counts
    for u in range(n):
        assert np.array_equal(run.candidates(u), ref.candidates(u)), u


def test_gql_batch_with_wide_membership_words_keeps_small_query_counts(dev):
    """Small queries batched with a 40-vertex query whose largest label
    group holds 33 vertices (8-byte membership words for the whole batch)
    count and call as the oracle does. The cap (1e5) stops the large
    query's search early and lies above every small query's count."""
    from psm_benchmark_amd import gen
    from psm_benchmark_amd.matcher import Batch, DeviceGraph
    data = gen.chung_lu_graph(2000, 16000, 2, 1.0, seed=42)
    big = gen.query(data, 40, 2, 122)
    small = [gen.query(data, 3 + i % 4, 1, 300 + i) for i in (1, 3, 4, 6)]
    dg = DeviceGraph(dev, data)
    b = Batch(dg, small + [big])
    b.prepare("GQL", "GQL")
    counts, calls, status, _ = b.evaluate(100000, 30.0)
    b.close()
    assert (int(status[4]), int(counts[4])) == (0, 100000)
    od = OGraph(data)
    for i, q in enumerate(small):
        ref = ORun(od, OGraph(q), "GQL_GQL_GQL")
        assert (int(status[i]), int(counts[i]), int(calls[i])) == (0, ref.count, ref.calls), i


def test_c1_reference_vectors_on_device(dev):
    from psm_benchmark_amd import gen
    from psm_benchmark_amd.matcher import DeviceGraph, match
    data = json.load(open(os.path.join(HERE, "golden", "c1_ref.json")))
    gp = data["graph"]
    g = gen.chung_lu_graph(gp["V"], gp["E"], gp["labels"], gp["zipf"], seed=gp["seed"])
    dg = DeviceGraph(dev, g)
    for rec in data["records"]:
        q = gen.query(g, rec["n"], bool(rec["dense"]), rec["qseed"])
        r = match(dg, q, rec["method"])
        assert r.count == rec["embeddings"], rec
        assert r.mem_words == rec["mem_words"], rec
        assert r.calls == rec["calls"], rec


@pytest.mark.parametrize("seed", range(8))
def test_bruteforce_tiny_on_device(dev, seed):
    from psm_benchmark_amd.matcher import DeviceGraph, match
    rng = np.random.default_rng(100 + seed)
    data = random_graph(rng, 11, 0.45, 2)
    q = random_connected_query(rng, data, int(rng.integers(2, 5)))
    dg = DeviceGraph(dev, data)
    expect = count_embeddings(data, q)
    for m in ("LDF_GQL_GQL", "NLF_GQL_GQL", "GQL_GQL_GQL", "CFL_CFL_EXPLORE", "DPiso_DPiso_LFTJ"):
        assert match(dg, q, m).count == expect, m


def test_clique_closed_form_on_device(dev):
    from psm_benchmark_amd.matcher import DeviceGraph, match
    n, k = 9, 5
    kn = csr_from_edges(n, np.array(list(itertools.combinations(range(n), 2))), np.zeros(n, np.uint32))
    kk = csr_from_edges(k, np.array(list(itertools.combinations(range(k), 2))), np.zeros(k, np.uint32))
    dg = DeviceGraph(dev, kn)
    r = match(dg, kk, "GQL_GQL_GQL")
    assert r.count == math.factorial(n) // math.factorial(n - k)
    assert r.calls == sum(math.factorial(n) // math.factorial(n - d - 1) for d in range(k))


def test_search_bytes_clique_closed_form_on_device(dev):
    from psm_benchmark_amd.matcher import DeviceGraph, match
    from test_oracle import clique_search_bytes
    n, k = 9, 5
    kn = csr_from_edges(n, np.array(list(itertools.combinations(range(n), 2))), np.zeros(n, np.uint32))
    kk = csr_from_edges(k, np.array(list(itertools.combinations(range(k), 2))), np.zeros(k, np.uint32))
    dg = DeviceGraph(dev, kn)
    assert match(dg, kk, "GQL_GQL_GQL").algo_bytes == clique_search_bytes(n, k)


def test_tiny_task_ring_keeps_golden_counts(dev, hprd):
    """A ring of 64 slots for split-off work (and 64 pre-expanded tasks per
    query): the ring wraps many times over, slots are reused as they are
    consumed, and splitting waits when the next slot is still held; counts
    and call counts still equal the golden records, batched and plain."""
    from psm_benchmark_amd import autotune
    from psm_benchmark_amd import _lib as L
    from psm_benchmark_amd.matcher import batch_match, match
    _, dg, _ = hprd
    before = autotune.get_config(dev)
    v = ctypes.c_double()
    try:
        assert L.lib().psm_set_tuning(dev.h, L.PSM_TUNE_QUEUE_SLOTS, 64.0) == L.PSM_OK
        autotune.set_config(dev, {"expand_tasks_per_lane": 0.001})
        assert L.lib().psm_get_tuning(dev.h, L.PSM_TUNE_QUEUE_SLOTS, ctypes.byref(v)) == L.PSM_OK and v.value == 64
        for method in ("GQL_GQL_GQL", "CFL_CFL_EXPLORE"):
            recs = golden((method,))
            qs = [load_text(os.path.join(HPRD, r["query"] + ".txt")) for r in recs]
            f, o, _ = method.split("_")
            counts, calls, status, _ = batch_match(dg, qs, f, o, 0)
            assert counts.tolist() == [r["embeddings"] for r in recs], method
            assert calls.tolist() == [r["calls"] for r in recs], method
            for r, q in zip(recs[-3:], qs[-3:]):
                m = match(dg, q, method)
                assert (m.count, m.calls) == (r["embeddings"], r["calls"]), (method, r["query"])
    finally:
        L.lib().psm_set_tuning(dev.h, L.PSM_TUNE_QUEUE_SLOTS, 0.0)
        autotune.set_config(dev, before)


def test_block_widths_give_golden_counts(dev, hprd):
    """The lane DFS in workgroups of 128 and 256 lanes (the default is one
    wavefront) and at four wavefronts per CU: batched counts and call counts
    still equal the golden records."""
    from psm_benchmark_amd import autotune
    from psm_benchmark_amd.matcher import batch_match
    _, dg, _ = hprd
    before = autotune.get_config(dev)
    recs = golden(("GQL_GQL_GQL",))
    qs = [load_text(os.path.join(HPRD, r["query"] + ".txt")) for r in recs]
    try:
        for lanes, waves in ((128, 1), (256, 1), (0, 4)):
            autotune.set_config(dev, {"dfs_block_lanes": lanes, "dfs_waves_per_cu": waves})
            counts, calls, status, _ = batch_match(dg, qs, "GQL", "GQL", 0)
            assert counts.tolist() == [r["embeddings"] for r in recs], lanes
            assert calls.tolist() == [r["calls"] for r in recs], lanes
    finally:
        autotune.set_config(dev, before)


def test_intersection_and_depth_knobs_give_golden_counts(dev, hprd):
    """Binary search in every operand row (no galloping cursors, no
    LDS-staged windows) and shallow pre-expansion depths: batched counts and
    call counts still equal the golden records, on both DFS kernels."""
    from psm_benchmark_amd import autotune
    from psm_benchmark_amd.matcher import batch_match
    _, dg, _ = hprd
    before = autotune.get_config(dev)
    recs = golden(("GQL_GQL_GQL",))
    qs = [load_text(os.path.join(HPRD, r["query"] + ".txt")) for r in recs]
    try:
        for inter, depth in (("bsearch", 31), ("bsearch", 1), ("adaptive", 2)):
            autotune.set_config(dev, {"intersect": inter, "expand_depth": depth})
            counts, calls, status, _ = batch_match(dg, qs, "GQL", "GQL", 0)
            assert counts.tolist() == [r["embeddings"] for r in recs], (inter, depth)
            assert calls.tolist() == [r["calls"] for r in recs], (inter, depth)
    finally:
        autotune.set_config(dev, before)


def test_pipeline_batches_equal_golden(dev, hprd):
    """Batches through a two-context pipeline (two HIP streams, their
    launches overlapping) each give the golden counts and call counts."""
    from psm_benchmark_amd import _lib as L
    from psm_benchmark_amd.matcher import Pipeline
    _, dg, _ = hprd
    recs = golden(("GQL_GQL_GQL",))
    qs = [load_text(os.path.join(HPRD, r["query"] + ".txt")) for r in recs]
    halves = [qs[::2], qs[1::2], qs, qs[::-1]]
    want = [[r["embeddings"] for r in recs[::2]], [r["embeddings"] for r in recs[1::2]],
            [r["embeddings"] for r in recs], [r["embeddings"] for r in recs[::-1]]]
    p = Pipeline(dg, 2, time_limit=60.0)
    try:
        runs = p.run(halves, "GQL", "GQL", 0)
    finally:
        p.close()
    for r, w in zip(runs, want):
        assert r.counts.tolist() == w
        assert not (r.status == L.PSM_ETIMEOUT).any()
        assert r.end_s >= r.start_s and (r.latencies_ms >= 0).all()
    assert runs[2].calls.tolist() == [r["calls"] for r in recs]


def test_batch_latencies_are_within_the_step(hprd):
    """Per-query completion times of a batch lie between the start of the
    batch and the end of its evaluate call, and capped queries finish."""
    import time
    from psm_benchmark_amd.matcher import Batch
    _, dg, _ = hprd
    qs = [load_text(os.path.join(HPRD, f"Q10-{i}.txt")) for i in range(1, 11)]
    t = time.perf_counter()
    b = Batch(dg, qs)
    b.prepare("GQL", "GQL")
    counts, _, _, _ = b.evaluate(10000)
    wall_ms = (time.perf_counter() - t) * 1e3
    lat = b.latencies()
    b.close()
    assert counts.tolist() == [min(10000, r["embeddings"]) for r in golden(("GQL_GQL_GQL",))]
    assert (lat > 0).all() and (lat <= wall_ms + 1.0).all(), (lat, wall_ms)


@pytest.mark.parametrize("method", ["GQL_GQL_GQL", "CFL_CFL_EXPLORE"])
def test_search_bytes_match_oracle(hprd, method):
    """The algorithmic bytes the enumeration kernels count (the roofline's
    numerator) equal the oracle's sum over the same search tree, whatever
    the schedule (expansion, lanes or wavefronts, split-off ranges)."""
    from psm_benchmark_amd.matcher import match
    _, dg, og = hprd
    for qname in ("Q10-1", "Q10-2", "Q10-3", "Q10-4", "Q10-9"):
        q = load_text(os.path.join(HPRD, qname + ".txt"))
        r, ref = match(dg, q, method), ORun(og, OGraph(q), method)
        print(qname, method, r.algo_bytes, ref.algo_bytes, flush=True)
        assert r.count == ref.count and r.algo_bytes == ref.algo_bytes, qname


def test_cap_reports_min_cap_total(hprd):
    from psm_benchmark_amd.matcher import match
    _, dg, _ = hprd
    q = load_text(os.path.join(HPRD, "Q10-8.txt"))
    total = match(dg, q, "GQL_GQL_GQL").count
    assert total == 994203
    for cap in (1, 777, 100000, total, total + 1):
        r = match(dg, q, "GQL_GQL_GQL", cap=cap)
        assert r.count == min(cap, total)
        assert r.capped == (cap <= total)


def test_batch_equals_golden(hprd):
    from psm_benchmark_amd.matcher import batch_match
    _, dg, _ = hprd
    recs = golden(("GQL_GQL_GQL",))
    qs = [load_text(os.path.join(HPRD, r["query"] + ".txt")) for r in recs]
    counts, calls, status, st = batch_match(dg, qs, "GQL", "GQL", 0)
    assert counts.tolist() == [r["embeddings"] for r in recs]
    assert calls.tolist() == [r["calls"] for r in recs]
    recs = golden(("CFL_CFL_EXPLORE",))
    counts, calls, status, st = batch_match(dg, qs, "CFL", "CFL", 0)
    assert counts.tolist() == [r["embeddings"] for r in recs]
    assert calls.tolist() == [r["calls"] for r in recs]
    recs = golden(("DPiso_DPiso_LFTJ",))
    counts, calls, status, st = batch_match(dg, qs, "DPiso", "DPiso", 0)
    assert counts.tolist() == [r["embeddings"] for r in recs]


@pytest.mark.parametrize("method", ["GQL_GQL_GQL", "CFL_CFL_EXPLORE"])
def test_join_equals_plain_on_hprd(hprd, method):
    """Two-unit decomposition + device hash join counts every HPRD query's
    embeddings exactly (the reference's Q-split result, JoinCollect.cpp)."""
    from psm_benchmark_amd.matcher import join_match
    _, dg, _ = hprd
    for rec in golden((method,)):
        q = load_text(os.path.join(HPRD, rec["query"] + ".txt"))
        st = join_match(dg, q, method)
        print(rec["query"], st.rows_a, st.rows_b, st.keys, st.count, flush=True)
        assert st.count == rec["embeddings"], rec["query"]


@pytest.mark.parametrize("seed", range(6))
def test_join_bruteforce_tiny(dev, seed):
    from psm_benchmark_amd.matcher import DeviceGraph, join_match
    rng = np.random.default_rng(200 + seed)
    data = random_graph(rng, 12, 0.45, 2)
    q = random_connected_query(rng, data, int(rng.integers(3, 6)))
    dg = DeviceGraph(dev, data)
    assert join_match(dg, q, "GQL_GQL_GQL").count == count_embeddings(data, q)


@pytest.mark.parametrize("pattern,filt", [("STAR", "GQL"), ("STAR", "LDF"), ("TWINTWIG", "GQL"),
                                          ("SEED", "GQL"), ("SEED", "LDF")])
def test_split_join_equals_plain_on_hprd(hprd, pattern, filt):
    """A STAR / TWINTWIG / SEED split joined unit by unit on the device counts
    every HPRD query's embeddings exactly (the reference's Q-split result)."""
    from psm_benchmark_amd.matcher import split_join_match
    _, dg, _ = hprd
    for rec in golden(("GQL_GQL_GQL",)):
        q = load_text(os.path.join(HPRD, rec["query"] + ".txt"))
        st = split_join_match(dg, q, pattern, filt, max_rows=1 << 24)
        print(rec["query"], pattern, filt, st.units, st.unit_rows, st.max_inter_rows, st.count, flush=True)
        assert st.count == rec["embeddings"], rec["query"]
        assert st.units >= 2


@pytest.mark.parametrize("seed", range(8))
def test_split_join_bruteforce_tiny(dev, seed):
    from psm_benchmark_amd.matcher import DeviceGraph, split_join_match
    rng = np.random.default_rng(300 + seed)
    data = random_graph(rng, 12, 0.5, 2)
    q = random_connected_query(rng, data, int(rng.integers(2, 7)))
    dg = DeviceGraph(dev, data)
    expect = count_embeddings(data, q)
    for pattern in ("STAR", "TWINTWIG", "SEED"):
        for filt in ("LDF", "GQL"):
            assert split_join_match(dg, q, pattern, filt).count == expect, (pattern, filt)


def test_split_join_equals_reference_vectors_on_c1(dev):
    """The split joins on the seeded C1-shaped graph (307 labels) against the
    reference binary's counts (tests/golden/c1_ref.json), and batched DPiso
    against its DPiso_DPiso_LFTJ counts."""
    from psm_benchmark_amd import gen
    from psm_benchmark_amd.matcher import DeviceGraph, batch_match, split_join_match
    data = json.load(open(os.path.join(HERE, "golden", "c1_ref.json")))
    gp = data["graph"]
    g = gen.chung_lu_graph(gp["V"], gp["E"], gp["labels"], gp["zipf"], seed=gp["seed"])
    dg = DeviceGraph(dev, g)
    recs = [r for r in data["records"] if r["method"] == "GQL_GQL_GQL" and r["n"] >= 4][:24]
    qs = [gen.query(g, r["n"], bool(r["dense"]), r["qseed"]) for r in recs]
    for pattern in ("STAR", "SEED"):
        for q, r in zip(qs, recs):
            st = split_join_match(dg, q, pattern, "GQL", max_rows=1 << 24)
            assert st.count == r["embeddings"], (pattern, r)
    drecs = {(r["qseed"], r["dense"], r["n"]): r for r in data["records"] if r["method"] == "DPiso_DPiso_LFTJ"}
    keys = [(r["qseed"], r["dense"], r["n"]) for r in recs if (r["qseed"], r["dense"], r["n"]) in drecs]
    dqs = [gen.query(g, n, bool(dense), qseed) for qseed, dense, n in keys]
    counts, calls, status, st = batch_match(dg, dqs, "DPiso", "DPiso", 0)
    assert counts.tolist() == [drecs[k]["embeddings"] for k in keys]


def test_split_join_row_bound_is_enforced(hprd):
    from psm_benchmark_amd import _lib as L
    from psm_benchmark_amd.matcher import split_join_match
    _, dg, _ = hprd
    q = load_text(os.path.join(HPRD, "Q10-10.txt"))
    with pytest.raises(L.PSMError) as e:
        split_join_match(dg, q, "STAR", "GQL", max_rows=8)
    assert e.value.code == L.PSM_ENOMEM


def test_time_limit_reports_partial_count(hprd):
    """A zero-length time budget stops enumeration and reports PSM_ETIMEOUT
    with a partial count, as the reference logs Overtime: 1."""
    from psm_benchmark_amd import _lib as L
    from psm_benchmark_amd.matcher import match
    _, dg, _ = hprd
    q = load_text(os.path.join(HPRD, "Q10-10.txt"))
    L.lib().psm_set_time_limit(dg.dev.h, 1e-9)
    try:
        r = match(dg, q, "GQL_GQL_GQL")
    finally:
        L.lib().psm_set_time_limit(dg.dev.h, 60.0)
    assert r.timed_out and r.count < 1015922
    assert match(dg, q, "GQL_GQL_GQL").count == 1015922


def _resident_lane_only(dev):
    """Resident launches always run the lane kernel (the wavefront kernel's
    resident mode is switched off: a served test of it once left the GPU
    faulted, cause not found), so the wavefront parametrization would only
    repeat the lane one."""
    from psm_benchmark_amd import _lib as L
    v = ctypes.c_double()
    L.lib().psm_get_tuning(dev.h, L.PSM_TUNE_DFS_STRATEGY, ctypes.byref(v))
    if int(v.value) == L.PSM_DFS_WAVE:
        pytest.skip("resident launches run the lane kernel; covered by the lane parametrization")


@pytest.mark.parametrize("resident", [True, False], ids=["resident", "per_batch"])
def test_served_results_equal_alone_and_in_the_mix(dev, hprd, resident):
    """Requests served through continuous batching (several workers, mixed
    arrivals; batches admitted into one resident launch, or one launch per
    batch) give each query exactly its count and Call Count when matched
    alone."""
    from psm_benchmark_amd.serving import Server, run_poisson
    if resident:
        _resident_lane_only(dev)
    g, dg, _ = hprd
    recs = golden(("GQL_GQL_GQL",))
    qs = [load_text(os.path.join(HPRD, r["query"] + ".txt")) for r in recs]
    srv = Server(dg, workers=3, max_batch=4, filter="GQL", order="GQL", resident=resident)
    try:
        results, summary = run_poisson(srv, qs, rate_qps=500.0, n_requests=60, seed=7)
        print(summary, srv.stats(), flush=True)
    finally:
        srv.close()
    for k, r in results.items():
        assert r.status == 0
        assert r.count == recs[k % len(recs)]["embeddings"], k
        assert r.calls == recs[k % len(recs)]["calls"], k
    assert max(r.batch_size for r in results.values()) >= 1


def test_resident_launch_reopens_and_takes_single_request_batches(dev, hprd):
    """One request per batch, in two waves with a pause between them: the
    first launch closes once it has nothing to run, the second wave starts a
    new one; every count and Call Count is the golden one."""
    from psm_benchmark_amd.serving import Server
    _resident_lane_only(dev)
    g, dg, _ = hprd
    recs = golden(("GQL_GQL_GQL",))
    qs = [load_text(os.path.join(HPRD, r["query"] + ".txt")) for r in recs]
    srv = Server(dg, workers=4, max_batch=1, filter="GQL", order="GQL", resident=True)
    results = {}
    try:
        for wave in range(2):
            for k in range(len(qs)):
                srv.submit(wave * len(qs) + k, qs[k])
            t = time.time()
            while len(results) < (wave + 1) * len(qs) and time.time() - t < 60:
                for r in srv.poll(64, 20.0):
                    results[int(r.request_id)] = r
            time.sleep(0.05)
    finally:
        srv.close()
    assert len(results) == 2 * len(qs)
    for k, r in results.items():
        assert r.status == 0 and r.batch_size == 1, k
        assert r.count == recs[k % len(recs)]["embeddings"], k
        assert r.calls == recs[k % len(recs)]["calls"], k


def test_resident_launch_time_limit_gives_partial_counts(dev, hprd):
    """A time limit the heavy query cannot meet: the resident launch drops
    each query's work at its own deadline and returns its partial count with
    PSM_ETIMEOUT (the reference's Overtime: 1); with a generous limit the
    same requests are exact."""
    from psm_benchmark_amd import _lib as L
    from psm_benchmark_amd.serving import Server
    _resident_lane_only(dev)
    g, dg, _ = hprd
    recs = golden(("GQL_GQL_GQL",))
    qs = [load_text(os.path.join(HPRD, r["query"] + ".txt")) for r in recs]
    full = [r["embeddings"] for r in recs]
    qs.append(load_text(os.path.join(HPRD, "Q10-10.txt")))
    full.append(1015922)
    for limit in (1e-9, 60.0):
        srv = Server(dg, workers=2, max_batch=4, filter="GQL", order="GQL", time_limit=limit, resident=True)
        results = {}
        try:
            for k in range(len(qs)):
                srv.submit(k, qs[k])
            t = time.time()
            while len(results) < len(qs) and time.time() - t < 60:
                for r in srv.poll(64, 20.0):
                    results[int(r.request_id)] = r
        finally:
            srv.close()
        assert len(results) == len(qs)
        for k, r in results.items():
            if limit >= 1:
                assert r.status == 0 and r.count == full[k], k
            elif r.status == L.PSM_ETIMEOUT:
                assert r.count < full[k], k
            else:   # a light query can end within its first task
                assert r.status == 0 and r.count == full[k], k
        if limit < 1:   # the heavy one cannot
            assert results[len(qs) - 1].status == L.PSM_ETIMEOUT


def test_resident_server_runs_large_queries_beside_the_resident_launch(dev, hprd):
    """A batch holding a query above the resident launch's 32 vertices runs
    its own enumeration launch beside the resident one: with small and
    40-vertex queries mixed in one stream, every served count equals the
    query's count matched alone (cap 1,000)."""
    from psm_benchmark_amd import gen
    from psm_benchmark_amd.matcher import match
    from psm_benchmark_amd.serving import Server, run_poisson
    _resident_lane_only(dev)
    g, dg, _ = hprd
    recs = golden(("GQL_GQL_GQL",))
    qs = [load_text(os.path.join(HPRD, r["query"] + ".txt")) for r in recs[:4]]
    qs += [gen.query(g, 40, False, seed) for seed in (3, 5)]
    cap = 1000
    alone = [match(dg, q, "GQL_GQL_GQL", cap=cap).count for q in qs]
    srv = Server(dg, workers=3, max_batch=3, filter="GQL", order="GQL", cap=cap, resident=True)
    try:
        results, summary = run_poisson(srv, qs, rate_qps=300.0, n_requests=36, seed=5)
    finally:
        srv.close()
    assert max(q.num_vertices for q in qs) > 32
    for k, r in results.items():
        assert r.status == 0, k
        assert r.count == alone[k % len(qs)], k


def test_served_dpiso_results_equal_golden(hprd):
    """The serving loop with the DPiso filter and order: every request's count
    equals the reference's DPiso_DPiso_LFTJ log, and its assignment count the
    oracle's GQL engine on the same plan."""
    from psm_benchmark_amd.serving import Server, run_poisson
    g, dg, og = hprd
    recs = golden(("DPiso_DPiso_LFTJ",))
    qs = [load_text(os.path.join(HPRD, r["query"] + ".txt")) for r in recs]
    calls = [ORun(og, OGraph(q), "DPiso_DPiso_GQL").calls for q in qs]
    srv = Server(dg, workers=2, max_batch=4, filter="DPiso", order="DPiso")
    try:
        results, summary = run_poisson(srv, qs, rate_qps=500.0, n_requests=30, seed=11)
    finally:
        srv.close()
    for k, r in results.items():
        assert r.status == 0
        assert r.count == recs[k % len(recs)]["embeddings"], k
        assert r.calls == calls[k % len(recs)], k


def hub_star_query(g, hub: int, leaves: int, seed: int):
    """A query around a data hub: the hub, `leaves` of its neighbours, and the
    data edges among those neighbours; labels are the data labels (so at
    least one embedding exists). The hub's query degree exceeds 8, which
    sends its candidates to the wavefront refinement test."""
    rng = np.random.default_rng(seed)
    off = g.offsets.astype(np.int64)
    nb = g.neighbors[off[hub]:off[hub + 1]]
    pick = [int(x) for x in rng.choice(nb, size=leaves, replace=False)]
    verts = [hub] + pick
    idx = {v: i for i, v in enumerate(verts)}
    edges = [(0, i) for i in range(1, len(verts))]
    for a in pick:
        for b in g.neighbors[off[a]:off[a + 1]]:
            b = int(b)
            if b in idx and idx[a] < idx[b]:
                edges.append((idx[a], idx[b]))
    return csr_from_edges(len(verts), np.array(edges), g.labels[np.array(verts)])


@pytest.mark.parametrize("labels,seed", [(3, 7), (12, 8)])
def test_wavefront_refinement_on_hubs_matches_oracle(dev, labels, seed):
    """Candidate sets and tables of GQL filtering on a power-law graph with
    few labels (long label runs on the hubs) for queries whose centre has
    query degree 10-12 (the wavefront test with its per-label early exits
    and saturation), against the oracle element for element."""
    from psm_benchmark_amd import gen
    from psm_benchmark_amd.matcher import DeviceGraph, Query
    g = gen.chung_lu_graph(20000, 200000, labels, 1.0, seed=seed)
    og = OGraph(g)
    dg = DeviceGraph(dev, g)
    deg = np.diff(g.offsets.astype(np.int64))
    hubs = np.argsort(-deg)[:40:8]
    for i, hub in enumerate(hubs):
        q = hub_star_query(g, int(hub), 10 + (i % 3), seed + i)
        ref = ORun(og, OGraph(q), "GQL_GQL_GQL", cap=1000)
        run = Query(dg, q)
        run.filter("GQL")
        print("hub", int(hub), "deg", int(deg[hub]), "query n", q.num_vertices, "cands", ref.cand_counts, flush=True)
        assert run.candidate_counts().tolist() == ref.cand_counts
        for u in range(q.num_vertices):
            assert np.array_equal(run.candidates(u), ref.candidates(u)), u
        run.build_tables()
        for a, b in itertools.product(range(q.num_vertices), repeat=2):
            t, rt = run.table(a, b), ref.table(a, b)
            assert (t is None) == (rt is None)
            if t is not None:
                assert np.array_equal(t[0], rt[0]) and np.array_equal(t[1], rt[1]), (a, b)
        run.close()
