"""Pins the CPU oracle (oracle/psm_oracle.cpp) before it is trusted as the
checker of the device path.

* HPRD-15 golden logs of the reference (test_anony/C_output/HPRD/L15/Q10):
  embeddings, call counts and BuildTable memory checksum, GQL_GQL_GQL,
  CFL_CFL_EXPLORE and DPiso_DPiso_LFTJ at one thread.
* Vectors from the reference binary built from its own sources on a seeded
  C1-shaped workload (tests/golden/c1_ref.json, made by make_ref_golden.py),
  including GQL_GQL_LFTJ (call counts that include visited rejects) and
  DPiso_DPiso_LFTJ.
* Brute force on tiny random graphs; k-cliques on a one-label complete graph
  (n!/(n-k)! embeddings); NLF/LDF candidate sets against their definition;
  the transpose symmetry of candidate-edge tables.
"""
import itertools
import json
import math
import os

import numpy as np
import pytest

from oracle_binding import OGraph, ORun
from brute import count_embeddings, random_graph, random_connected_query
from psm_benchmark_amd.graph_io import csr_from_edges, load_text

HERE = os.path.dirname(os.path.abspath(__file__))
HPRD = os.path.join(HERE, "golden", "hprd")


@pytest.fixture(scope="module")
def hprd():
    return OGraph(load_text(os.path.join(HPRD, "HPRD-15.txt")))


def golden_records(methods=("GQL_GQL_GQL", "CFL_CFL_EXPLORE", "DPiso_DPiso_LFTJ")):
    recs = json.load(open(os.path.join(HPRD, "golden.json")))
    return [r for r in recs if r["threads"] == 1 and f'{r["filter"]}_{r["order"]}_{r["engine"]}' in methods]


@pytest.mark.parametrize("rec", golden_records(), ids=lambda r: f'{r["query"]}-{r["filter"]}')
def test_hprd_golden(hprd, rec):
    q = OGraph(load_text(os.path.join(HPRD, rec["query"] + ".txt")))
    r = ORun(hprd, q, f'{rec["filter"]}_{rec["order"]}_{rec["engine"]}')
    assert (r.count, r.calls, r.mem_words) == (rec["embeddings"], rec["calls"], rec["mem_words"])


def test_hprd_counts_agree_across_all_reference_methods():
    # every method and thread count of the reference logs the same #Embeddings
    recs = json.load(open(os.path.join(HPRD, "golden.json")))
    by_q = {}
    for r in recs:
        by_q.setdefault(r["query"], set()).add(r["embeddings"])
    assert all(len(v) == 1 for v in by_q.values())


def test_c1_reference_vectors():
    from psm_benchmark_amd import gen
    data = json.load(open(os.path.join(HERE, "golden", "c1_ref.json")))
    gp = data["graph"]
    g = gen.chung_lu_graph(gp["V"], gp["E"], gp["labels"], gp["zipf"], seed=gp["seed"])
    d = OGraph(g)
    cache = {}
    for rec in data["records"]:
        key = (rec["qseed"], rec["dense"], rec["n"])
        if key not in cache:
            cache[key] = OGraph(gen.query(g, rec["n"], bool(rec["dense"]), rec["qseed"]))
        r = ORun(d, cache[key], rec["method"])
        assert (r.count, r.calls, r.mem_words) == (rec["embeddings"], rec["calls"], rec["mem_words"]), rec


@pytest.mark.parametrize("seed", range(12))
def test_bruteforce_tiny(seed):
    rng = np.random.default_rng(seed)
    data = random_graph(rng, 11, 0.45, 2)
    q = random_connected_query(rng, data, int(rng.integers(2, 5)))
    assert q is not None
    expect = count_embeddings(data, q)
    d, qq = OGraph(data), OGraph(q)
    for m in ("LDF_GQL_GQL", "NLF_GQL_EXPLORE", "GQL_GQL_GQL", "CFL_CFL_EXPLORE", "GQL_GQL_LFTJ",
              "DPiso_DPiso_LFTJ", "DPiso_DPiso_GQL"):
        assert ORun(d, qq, m).count == expect, m


@pytest.mark.parametrize("n,k", [(6, 3), (7, 4), (8, 5)])
def test_clique_closed_form(n, k):
    kn = csr_from_edges(n, np.array(list(itertools.combinations(range(n), 2))), np.zeros(n, np.uint32))
    kk = csr_from_edges(k, np.array(list(itertools.combinations(range(k), 2))), np.zeros(k, np.uint32))
    expect = math.factorial(n) // math.factorial(n - k)
    for m in ("GQL_GQL_GQL", "CFL_CFL_EXPLORE", "GQL_GQL_LFTJ", "DPiso_DPiso_LFTJ"):
        r = ORun(OGraph(kn), OGraph(kk), m)
        assert r.count == expect
        # every partial injective map extends: calls = sum_d n!/(n-d-1)!
        assert r.calls == sum(math.factorial(n) // math.factorial(n - d - 1) for d in range(k))


def clique_search_bytes(n: int, k: int) -> int:
    """Algorithmic search bytes of a k-clique query on a one-label K_n: a
    node at depth d (1 <= d < k) reads d table rows of n-1 entries and there
    are n!/(n-d)! of them; the survivors at depths 1..k-2 are written."""
    P = lambda d: math.factorial(n) // math.factorial(n - d)
    rows = sum(P(d) * d * (n - 1) for d in range(1, k))
    survivors = sum(P(d + 1) for d in range(1, k - 1))
    return 4 * (rows + survivors)


@pytest.mark.parametrize("n,k", [(6, 3), (7, 4), (8, 5)])
def test_search_bytes_clique_closed_form(n, k):
    kn = csr_from_edges(n, np.array(list(itertools.combinations(range(n), 2))), np.zeros(n, np.uint32))
    kk = csr_from_edges(k, np.array(list(itertools.combinations(range(k), 2))), np.zeros(k, np.uint32))
    for m in ("GQL_GQL_GQL", "CFL_CFL_EXPLORE", "GQL_GQL_LFTJ"):
        assert ORun(OGraph(kn), OGraph(kk), m).algo_bytes == clique_search_bytes(n, k), m


@pytest.mark.parametrize("qname", ["Q10-1", "Q10-2", "Q10-3", "Q10-9"])
def test_search_bytes_independent_of_engine(hprd, qname):
    # the measure is fixed by candidates, tables and order: the three engines
    # walk different mechanics over the same search tree (the GQL order keeps
    # the pivot among the backward neighbours, which all three engines need)
    q = OGraph(load_text(os.path.join(HPRD, qname + ".txt")))
    for f in ("GQL", "NLF"):
        vals = {e: ORun(hprd, q, f"{f}_GQL_{e}").algo_bytes for e in ("GQL", "EXPLORE", "LFTJ")}
        assert len(set(vals.values())) == 1 and vals["GQL"] > 0, vals


def test_cap_is_min_cap_total(hprd):
    q = OGraph(load_text(os.path.join(HPRD, "Q10-1.txt")))
    total = ORun(hprd, q, "GQL_GQL_GQL").count
    for cap in (1, 1000, total, total + 5):
        assert ORun(hprd, q, "GQL_GQL_GQL", cap=cap).count == min(cap, total)


@pytest.mark.parametrize("qname", ["Q10-1", "Q10-5", "Q10-10"])
@pytest.mark.parametrize("method", ["GQL_GQL_GQL", "CFL_CFL_EXPLORE", "DPiso_DPiso_LFTJ"])
def test_threaded_enumeration_matches_golden(hprd, qname, method):
    # the threaded leg (root candidates claimed one at a time by workers)
    # walks the same search tree: uncapped counts, call counts and search
    # bytes equal the one-thread golden values at every thread count
    f, o, e = method.split("_")
    rec = next(r for r in golden_records((method,)) if r["query"] == qname)
    q = OGraph(load_text(os.path.join(HPRD, qname + ".txt")))
    one = ORun(hprd, q, method)
    for t in (2, 5, 8):
        r = ORun(hprd, q, method, threads=t)
        assert (r.count, r.calls) == (rec["embeddings"], rec["calls"]), t
        assert r.algo_bytes == one.algo_bytes and r.depth_nodes == one.depth_nodes, t
        assert r.times.threads == t and not r.timed_out


def test_threaded_cap_is_min_cap_total(hprd):
    q = OGraph(load_text(os.path.join(HPRD, "Q10-8.txt")))
    total = ORun(hprd, q, "GQL_GQL_GQL").count
    for cap in (1, 300, 1000, 100000, total, total + 5):
        for t in (1, 4, 8):
            assert ORun(hprd, q, "GQL_GQL_GQL", cap=cap, threads=t).count == min(cap, total), (cap, t)


def test_time_limit_reports_timeout(hprd):
    q = OGraph(load_text(os.path.join(HPRD, "Q10-10.txt")))
    for t in (1, 4):
        r = ORun(hprd, q, "GQL_GQL_GQL", threads=t, time_limit=1e-4)
        assert r.timed_out and r.count < 1015922, t
        full = ORun(hprd, q, "GQL_GQL_GQL", threads=t, time_limit=600.0)
        assert not full.timed_out and full.count == 1015922, t


def _nlf(g, v):
    off = g.offsets.astype(np.int64)
    labs = g.labels[g.neighbors[off[v]:off[v + 1]]]
    u, c = np.unique(labs, return_counts=True)
    return dict(zip(u.tolist(), c.tolist()))


@pytest.mark.parametrize("method", ["LDF", "NLF"])
def test_local_filters_by_definition(method):
    rng = np.random.default_rng(3)
    data = random_graph(rng, 40, 0.15, 3)
    q = random_connected_query(rng, data, 5)
    r = ORun(OGraph(data), OGraph(q), f"{method}_GQL_GQL")
    for u in range(q.num_vertices):
        expect = []
        for v in range(data.num_vertices):
            if data.labels[v] != q.labels[u] or data.degrees[v] < q.degrees[u]:
                continue
            if method == "NLF":
                dv, du = _nlf(data, v), _nlf(q, u)
                if any(dv.get(l, 0) < c for l, c in du.items()):
                    continue
            expect.append(v)
        assert r.candidates(u).tolist() == expect


def test_tables_are_transposes(hprd):
    q = load_text(os.path.join(HPRD, "Q10-3.txt"))
    r = ORun(hprd, OGraph(q), "GQL_GQL_GQL")
    for a in range(q.num_vertices):
        for b in range(q.num_vertices):
            t = r.table(a, b)
            if t is None:
                continue
            off, idx = t
            toff, tidx = r.table(b, a)
            fwd = {(j, int(k)) for j in range(len(off) - 1) for k in idx[off[j]:off[j + 1]]}
            bwd = {(int(k), j) for j in range(len(toff) - 1) for k in tidx[toff[j]:toff[j + 1]]}
            assert fwd == bwd
            for j in range(len(off) - 1):
                row = idx[off[j]:off[j + 1]]
                assert np.all(np.diff(row.astype(np.int64)) > 0)


@pytest.mark.parametrize("qname", ["Q10-1", "Q10-4", "Q10-9"])
def test_dpiso_sets_are_pruned_ldf_sets(hprd, qname):
    # DPiso prunes the LDF sets (it never adds a vertex) and keeps every
    # vertex that takes part in an embedding: each vertex of an embedding
    # found under the LDF sets is in the DPiso set of its query vertex
    q = OGraph(load_text(os.path.join(HPRD, qname + ".txt")))
    ldf = ORun(hprd, q, "LDF_GQL_GQL")
    dp = ORun(hprd, q, "DPiso_DPiso_LFTJ")
    assert dp.count == ldf.count
    total_ldf = total_dp = 0
    for u in range(dp.n):
        a, b = set(ldf.candidates(u).tolist()), set(dp.candidates(u).tolist())
        assert b <= a
        total_ldf, total_dp = total_ldf + len(a), total_dp + len(b)
    assert total_dp < total_ldf


def test_dpiso_order_is_a_bfs_order_with_parent_pivots(hprd):
    for k in range(1, 11):
        qg = load_text(os.path.join(HPRD, f"Q10-{k}.txt"))
        r = ORun(hprd, OGraph(qg), "DPiso_DPiso_LFTJ")
        order, pivot = r.order, r.pivot
        off = qg.offsets.astype(np.int64)
        nbrs = lambda u: set(qg.neighbors[off[u]:off[u + 1]].tolist())
        assert sorted(order) == list(range(qg.num_vertices))
        pos = {u: i for i, u in enumerate(order)}
        for i in range(1, len(order)):
            p = pivot[i]
            # the pivot is a neighbour placed earlier, and the first one in
            # BFS order (the vertex that discovered it)
            assert p in nbrs(order[i]) and pos[p] < i
            assert p == min(nbrs(order[i]), key=lambda w: pos[w])
