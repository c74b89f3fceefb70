"""CPU checks of host-side logic around the device path: the benchmark
workload generators (C3 cyclic queries, C5 triangle / 4-clique queries, the
C3 join-checkable set) and the rocprofv3 PMC summary that turns FETCH_SIZE /
WRITE_SIZE counters into HBM bytes."""
import csv
import json
import os
import sys

import numpy as np

from psm_benchmark_amd import gen
from psm_benchmark_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _simple_connected(q):
    off = q.offsets.astype(np.int64)
    n = q.num_vertices
    seen, stack = {0}, [0]
    while stack:
        v = stack.pop()
        for w in q.neighbors[off[v]:off[v + 1]]:
            w = int(w)
            assert w != v
            if w not in seen:
                seen.add(w)
                stack.append(w)
    return len(seen) == n


def test_c3_queries_are_cyclic_and_sized():
    g = gen.chung_lu_graph(50_000, 300_000, 32, 1.0, seed=3)
    qs, meta = W.c3_queries(g, count=12)
    assert len(qs) == 12
    for q, m in zip(qs, meta):
        assert 12 <= q.num_vertices <= 16 and m["n"] == q.num_vertices
        assert W.has_cycle(q) and _simple_connected(q)
    # deterministic for a seed
    qs2, _ = W.c3_queries(g, count=12)
    assert all(np.array_equal(a.neighbors, b.neighbors) and np.array_equal(a.labels, b.labels)
               for a, b in zip(qs, qs2))


def test_c5_queries_have_triangles_and_four_cliques():
    g = gen.rmat_graph(14, 120_000, 8, seed=5)
    qs, meta = W.c5_queries(g, count=8)
    assert len(qs) == 8
    for i, (q, m) in enumerate(zip(qs, meta)):
        assert 6 <= q.num_vertices <= 10 and _simple_connected(q)
        assert W.clique_size(q, 3)
        if i % 2 == 0:
            assert W.clique_size(q, 4) and m["k4"]


def test_clique_detection_closed_cases():
    from psm_benchmark_amd.graph_io import csr_from_edges
    tri = csr_from_edges(3, np.array([[0, 1], [1, 2], [0, 2]]), np.zeros(3, np.uint32))
    path = csr_from_edges(4, np.array([[0, 1], [1, 2], [2, 3]]), np.zeros(4, np.uint32))
    k4 = csr_from_edges(4, np.array([[a, b] for a in range(4) for b in range(a + 1, 4)]), np.zeros(4, np.uint32))
    assert W.clique_size(tri, 3) and not W.clique_size(tri, 4)
    assert not W.clique_size(path, 3) and not W.has_cycle(path)
    assert W.clique_size(k4, 4) and W.has_cycle(k4)


def test_c3_join_queries_use_rare_labels_only():
    g = gen.chung_lu_graph(200_000, 1_500_000, 32, 1.0, seed=3)
    qs, _ = W.c3_join_queries(g, count=6)
    freq = np.bincount(g.labels, minlength=32)
    rare = set(np.argsort(freq, kind="stable")[:16].tolist())
    for q in qs:
        assert 6 <= q.num_vertices <= 9 and W.has_cycle(q) and _simple_connected(q)
        assert set(int(x) for x in q.labels) <= rare


def _write_counters(path, rows):
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "x_counter_collection.csv"), "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for r in rows:
            w.writerow(r)


def test_pmc_summary_scales_counters_by_the_copy_calibration(tmp_path):
    sys.path.insert(0, ROOT)
    import bench
    d = {k: str(tmp_path / k) for k in ("kf", "kw", "cf", "cw")}
    # the kernel: two dispatches, FETCH_SIZE split over two counter instances
    _write_counters(d["kf"], [
        dict(Dispatch_Id=1, Kernel_Name="k_dfs(...)", Counter_Name="FETCH_SIZE", Counter_Value=300),
        dict(Dispatch_Id=1, Kernel_Name="k_dfs(...)", Counter_Name="FETCH_SIZE", Counter_Value=100),
        dict(Dispatch_Id=2, Kernel_Name="k_dfs(...)", Counter_Name="FETCH_SIZE", Counter_Value=600),
        dict(Dispatch_Id=3, Kernel_Name="k_other(...)", Counter_Name="FETCH_SIZE", Counter_Value=9999),
    ])
    _write_counters(d["kw"], [dict(Dispatch_Id=1, Kernel_Name="k_dfs(...)", Counter_Name="WRITE_SIZE",
                                   Counter_Value=50)])
    # copy of 1 MiB: the fetch counter reads half of it, the write counter all of it
    _write_counters(d["cf"], [dict(Dispatch_Id=1, Kernel_Name="elementwise_copy", Counter_Name="FETCH_SIZE",
                                   Counter_Value=512)])
    _write_counters(d["cw"], [dict(Dispatch_Id=1, Kernel_Name="elementwise_copy", Counter_Name="WRITE_SIZE",
                                   Counter_Value=1024)])
    for k in ("cf", "cw"):
        with open(os.path.join(d[k], "copy.json"), "w") as fh:
            fh.write(json.dumps({"copy_bytes": 1 << 20, "reps": 1}) + "\n")
    out = str(tmp_path / "t.json")
    bench.pmc_summary(d["kf"], d["kw"], d["cf"], d["cw"], "dfs", out)
    r = json.load(open(out))
    assert r["kernel"] == "dfs"
    assert r["dispatches"] == [2, 1]
    assert r["fetch_kib"] == 500.0 and r["write_kib"] == 50.0
    assert r["scale_fetch"] == 2.0 and r["scale_write"] == 1.0
    assert r["traffic_bytes"] == 500 * 1024 * 2 + 50 * 1024


def test_pmc_summary_averages_a_kernel_class_over_its_kernels(tmp_path):
    # the GQL refinement class spans several kernels: all of their dispatches
    # count, other kernels do not
    sys.path.insert(0, ROOT)
    import bench
    d = {k: str(tmp_path / k) for k in ("kf", "kw", "cf", "cw")}
    _write_counters(d["kf"], [
        dict(Dispatch_Id=1, Kernel_Name="psm::k_gql_refine_small(...)", Counter_Name="FETCH_SIZE", Counter_Value=100),
        dict(Dispatch_Id=2, Kernel_Name="void psm::k_gql_hall<16u>(...)", Counter_Name="FETCH_SIZE", Counter_Value=300),
        dict(Dispatch_Id=3, Kernel_Name="psm::k_table_fill(...)", Counter_Name="FETCH_SIZE", Counter_Value=9999),
    ])
    _write_counters(d["kw"], [
        dict(Dispatch_Id=1, Kernel_Name="psm::k_gql_refine(...)", Counter_Name="WRITE_SIZE", Counter_Value=8),
        dict(Dispatch_Id=2, Kernel_Name="psm::k_gql_skip(...)", Counter_Name="WRITE_SIZE", Counter_Value=2),
    ])
    _write_counters(d["cf"], [dict(Dispatch_Id=1, Kernel_Name="elementwise_copy", Counter_Name="FETCH_SIZE",
                                   Counter_Value=1024)])
    _write_counters(d["cw"], [dict(Dispatch_Id=1, Kernel_Name="elementwise_copy", Counter_Name="WRITE_SIZE",
                                   Counter_Value=1024)])
    for k in ("cf", "cw"):
        with open(os.path.join(d[k], "copy.json"), "w") as fh:
            fh.write(json.dumps({"copy_bytes": 1 << 20, "reps": 1}) + "\n")
    out = str(tmp_path / "g.json")
    bench.pmc_summary(d["kf"], d["kw"], d["cf"], d["cw"], "gql_refine", out)
    r = json.load(open(out))
    assert r["kernel"] == "gql_refine" and r["dispatches"] == [2, 2]
    assert r["fetch_kib"] == 200.0 and r["write_kib"] == 5.0
    assert r["traffic_bytes"] == 205 * 1024


def test_c2_queries_follow_the_dense_sparse_rule():
    # even-indexed queries are dense (average degree above 3; K4 at 4
    # vertices, the densest 4-vertex graph), odd-indexed ones are the trees
    # of random walks (average degree 2(n-1)/n <= 3); sizes cycle through
    # 4/8/16/24/32; every query is a connected subgraph of the data graph
    g = W.data_graph("C2")
    qs, meta = W.c2_queries(g, count=40)
    for i, (q, m) in enumerate(zip(qs, meta)):
        n = q.num_vertices
        assert n == [4, 8, 16, 24, 32][i % 5] == m["n"]
        assert _simple_connected(q)
        avg = 2.0 * q.num_edges / n
        if i % 2 == 0:
            assert m["dense"] and W.is_dense(q)
            assert (q.num_edges == 6) if n == 4 else avg > 3.0
        else:
            assert not m["dense"] and q.num_edges == n - 1 and avg <= 3.0


def test_autotune_trial_order():
    # completed trials rank by enumeration time (a wider search overshoots
    # per-query caps), time-limited ones by search throughput, after them
    from psm_benchmark_amd.autotune import trial_key
    trials = [dict(name="fast_rate", timeouts=0, enum_ms=80.0, nodes_per_ms=9e6),
              dict(name="short", timeouts=0, enum_ms=50.0, nodes_per_ms=5e6),
              dict(name="limited_hi", timeouts=2, enum_ms=800.0, nodes_per_ms=9e6),
              dict(name="limited_lo", timeouts=2, enum_ms=800.0, nodes_per_ms=1e6)]
    assert [t["name"] for t in sorted(trials, key=trial_key)] == ["short", "fast_rate", "limited_hi", "limited_lo"]


def test_autotune_pick_prefers_throughput_among_near_ties():
    from psm_benchmark_amd.autotune import pick
    trials = [dict(name="short", timeouts=0, enum_ms=100.0, nodes_per_ms=5e5),
              dict(name="tie_fast", timeouts=0, enum_ms=108.0, nodes_per_ms=7e5),
              dict(name="slow_fast", timeouts=0, enum_ms=150.0, nodes_per_ms=9e5)]
    assert pick(trials)["name"] == "tie_fast"
    limited = [dict(name="a", timeouts=3, enum_ms=800.0, nodes_per_ms=2e5),
               dict(name="b", timeouts=3, enum_ms=800.0, nodes_per_ms=4e5)]
    assert pick(limited)["name"] == "b"
