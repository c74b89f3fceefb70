"""The C++ drop-in layer (psm_benchmark_amd/dropin): the reference's classes
and signatures over the C ABI, and a SubgraphMatching.out driver that makes
split_C_test's calls (StudyPerformance.cpp:1164-1716) and writes the
reference's log (parsed with the same parser as the reference's logs).
Host-side pieces (binary graph files against the reference's own writer,
set intersection, units and backward neighbours) are checked on the CPU by
tests/model/dropin_host_test.cpp."""
import json
import os
import subprocess

import numpy as np
import pytest

import ref_runner as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "psm_benchmark_amd", "SubgraphMatching.out")
HPRD = os.path.join(ROOT, "tests", "golden", "hprd")
LIBDIR = os.path.join(ROOT, "psm_benchmark_amd")
REF_WRITER = os.path.join(ROOT, "oracle", "_ref", "EdgeListToCSR.out")


@pytest.fixture(scope="module")
def host_test(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("dropin") / "dropin_host_test")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"),
                           "-I" + os.path.join(LIBDIR, "dropin"), os.path.join(ROOT, "tests", "model",
                                                                                "dropin_host_test.cpp"),
                           "-L" + LIBDIR, "-lpsm_dropin", "-lpsm", "-Wl,-rpath," + LIBDIR, "-o", out])
    return out


def _load_bin(host_test, prefix):
    out = subprocess.run([host_test, "load_bin", prefix], capture_output=True, text=True, check=True).stdout
    lines = out.strip().splitlines()
    V, E = map(int, lines[0].split())
    rows = [list(map(int, l.split())) for l in lines[1:]]
    return V, E, {r[0]: (r[1], r[2:]) for r in rows}


@pytest.mark.skipif(not os.access(REF_WRITER, os.X_OK), reason="the reference's writer is built here only")
def test_binary_reader_reads_the_reference_writer(host_test, tmp_path):
    """matching/EdgeListToCSR.cpp:77-121 writes the deg / adj / label triple
    that graph.cpp:221-403 reads; the drop-in's reader loads it into the
    same graph as the edge list (labels are random in the writer: range
    checked)."""
    rng = np.random.default_rng(3)
    V = 300
    edges = sorted({(int(a), int(b)) for a, b in rng.integers(0, V, (1500, 2)) if a < b})
    V = max(b for _, b in edges) + 1
    el = tmp_path / "g.txt"
    el.write_text("".join(f"{a} {b}\n" for a, b in edges))
    prefix = str(tmp_path / "g")
    subprocess.run([REF_WRITER, str(el), prefix, "5"], check=True, capture_output=True)
    nV, nE, rows = _load_bin(host_test, prefix)
    assert (nV, nE) == (V, len(edges))
    adj = {v: [] for v in range(V)}
    for a, b in edges:
        adj[a].append(b)
        adj[b].append(a)
    for v in range(V):
        lab, nb = rows[v]
        assert nb == sorted(adj[v]) and 0 <= lab < 5, v


def test_binary_round_trip_and_rejections(host_test, tmp_path):
    prefix = str(tmp_path / "h")
    subprocess.run([host_test, "store_bin", os.path.join(HPRD, "HPRD-15.txt"), prefix], check=True)
    from psm_benchmark_amd.graph_io import load_text
    g = load_text(os.path.join(HPRD, "HPRD-15.txt"))
    V, E, rows = _load_bin(host_test, prefix)
    assert (V, E) == (g.num_vertices, g.num_edges)
    off = g.offsets.astype(np.int64)
    for v in (0, 1, 17, 4000, V - 1):
        assert rows[v] == (int(g.labels[v]), [int(x) for x in g.neighbors[off[v]:off[v + 1]]])
    # truncated adjacency, then degrees that do not sum to 2E
    adj = open(prefix + "_adj.bin", "rb").read()
    open(prefix + "_adj.bin", "wb").write(adj[:-8])
    r = subprocess.run([host_test, "bad_bin", prefix], capture_output=True, text=True)
    assert r.returncode == 0 and "ends early" in r.stdout
    open(prefix + "_adj.bin", "wb").write(adj)
    deg = bytearray(open(prefix + "_deg.bin", "rb").read())
    deg[12] = (deg[12] + 1) % 256
    open(prefix + "_deg.bin", "wb").write(bytes(deg))
    r = subprocess.run([host_test, "bad_bin", prefix], capture_output=True, text=True)
    assert r.returncode == 0 and "2E" in r.stdout


def test_set_intersection_and_units(host_test):
    for seed in (1, 2):
        r = subprocess.run([host_test, "intersect", str(seed)], capture_output=True, text=True)
        assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout
    r = subprocess.run([host_test, "plan"], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "ok", (r.returncode, r.stderr)


def test_driver_built_and_rejects_unsupported_modes():
    assert os.access(DRIVER, os.X_OK)
    r = subprocess.run([DRIVER, "-d", "x", "-q", "y", "-QorCandi", "Z"], capture_output=True, text=True)
    assert r.returncode == 1 and "QorCandi" in r.stderr
    r = subprocess.run([DRIVER, "-d", "x", "-q", "y", "-QorCandi", "Q", "-Qpattern", "SKETCHTREE"],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "Qpattern" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("qname", ["Q10-1", "Q10-4", "Q10-9"])
def test_driver_log_matches_golden(tmp_path, qname):
    log = str(tmp_path / "log.txt")
    subprocess.run([DRIVER, "-d", os.path.join(HPRD, "HPRD-15.txt"), "-q", os.path.join(HPRD, qname + ".txt"),
                    "-num", "MAX", "-time_limit", "60", "-threadnums", "1", "-QorCandi", "C",
                    "-BackMethods", "GQL_GQL_GQL,CFL_CFL_EXPLORE,DPiso_DPiso_LFTJ", "-OutputFile", log],
                   check=True, timeout=120)
    blocks = R.parse_log(log)
    gold = {(r["filter"], r["order"], r["engine"]): r for r in json.load(open(os.path.join(HPRD, "golden.json")))
            if r["query"] == qname and r["threads"] == 1}
    assert len(blocks) == 3
    for b in blocks:
        g = gold[(b["Filter Type"], b["Order Type"], b["Engine Type"])]
        assert int(b["#Embeddings"]) == g["embeddings"]
        assert int(b["Call Count"]) == g["calls"]
        assert round(float(b["Memory cost (MB)"]) * 2**20 / 4) == g["mem_words"]
        assert b["Overtime"] == "0"
        # log4Metrics (Args.cpp:456-515): one host thread drove the device
        assert int(b["Thread Call Count"]) == g["calls"]
        assert int(b["Thread Embedding Count"]) == g["embeddings"]
        assert float(b["Thread Enum Time"]) >= float(b["Thread Execution Time"]) > 0
        assert int(b["TaskNum"]) >= 1


@pytest.mark.gpu
@pytest.mark.parametrize("pattern", ["STAR", "TWINTWIG", "SEED"])
def test_driver_query_split_mode_counts_golden(tmp_path, pattern):
    log = str(tmp_path / "log.txt")
    gold = {r["query"]: r["embeddings"] for r in json.load(open(os.path.join(HPRD, "golden.json")))
            if r["threads"] == 1 and r["filter"] == "GQL" and r["engine"] == "GQL"}
    for qname in ("Q10-1", "Q10-4", "Q10-9"):
        subprocess.run([DRIVER, "-d", os.path.join(HPRD, "HPRD-15.txt"), "-q", os.path.join(HPRD, qname + ".txt"),
                        "-num", "MAX", "-time_limit", "60", "-threadnums", "1", "-QorCandi", "Q",
                        "-Qpattern", pattern, "-BackMethods", "GQL_GQL_GQL", "-OutputFile", log],
                       check=True, timeout=120)
        b = R.parse_log(log)[-1]
        assert int(b["#Embeddings"]) == gold[qname], (qname, pattern)
        assert b["Overtime"] == "0" and int(b["Call Count"]) > 0
