"""The multi-threaded text graph reader (psm_text_graph_read, host code of
libpsm.so): the same CSR as the pure-Python loader on the reference's HPRD
data graph and on seeded synthetic graphs at several thread counts, and an
error, not a corrupt graph, on malformed files."""
import os

import numpy as np
import pytest

from psm_benchmark_amd import _lib as L
from psm_benchmark_amd import gen
from psm_benchmark_amd.graph_io import load_text, load_text_parallel, save_text

HERE = os.path.dirname(os.path.abspath(__file__))
HPRD = os.path.join(HERE, "golden", "hprd")


def same(a, b):
    return (np.array_equal(a.offsets, b.offsets) and np.array_equal(a.neighbors, b.neighbors)
            and np.array_equal(a.labels, b.labels))


@pytest.mark.parametrize("threads", [1, 2, 7, 0])
def test_hprd_matches_python_loader(threads):
    path = os.path.join(HPRD, "HPRD-15.txt")
    assert same(load_text_parallel(path, threads), load_text(path))


@pytest.mark.parametrize("threads", [1, 3, 16])
def test_synthetic_graph_round_trip(tmp_path, threads):
    g = gen.chung_lu_graph(20000, 90000, 40, 1.1, seed=5)
    p = str(tmp_path / "g.txt")
    save_text(g, p)
    r = load_text_parallel(p, threads)
    assert same(r, g)
    assert same(r, load_text(p))


def test_query_files_and_edgeless_graph(tmp_path):
    for k in (1, 5, 10):
        p = os.path.join(HPRD, f"Q10-{k}.txt")
        assert same(load_text_parallel(p), load_text(p))
    p = tmp_path / "iso.txt"
    p.write_text("t 3 0\nv 0 1 0\nv 1 2 0\nv 2 1 0\n")
    g = load_text_parallel(str(p))
    assert g.offsets.tolist() == [0, 0, 0, 0] and g.labels.tolist() == [1, 2, 1] and g.num_edges == 0


@pytest.mark.parametrize("text,what", [
    ("t 2 1\nv 0 0 1\nv 1 0 1\ne 0 2\n", "out-of-range"),
    ("t 2 1\nv 0 0 1\nv 1 0 2\ne 0 1\n", "degree field"),
    ("t 2 1\nv 0 0 1\nv 0 0 1\ne 0 1\n", "twice"),
    ("t 3 1\nv 0 0 1\nv 1 0 1\ne 0 1\n", "v lines"),
    ("t 2 2\nv 0 0 1\nv 1 0 1\ne 0 1\n", "e lines"),
    ("t 2 1\nv 0 0 1\nv 1 0 1\nx 0 1\n", "unexpected"),
    ("v 0 0 1\n", "header"),
])
def test_malformed_files_are_errors(tmp_path, text, what):
    p = tmp_path / "bad.txt"
    p.write_text(text)
    with pytest.raises(L.PSMError) as e:
        load_text_parallel(str(p))
    assert e.value.code == L.PSM_EINVAL and what in str(e.value)


def test_missing_file(tmp_path):
    with pytest.raises(L.PSMError) as e:
        load_text_parallel(str(tmp_path / "nope.txt"))
    assert e.value.code == L.PSM_ENOTFOUND


def test_crlf_blank_lines_and_trailing_fields(tmp_path):
    # Windows line ends, blank lines and a final line without a newline parse
    # to the same graph as the plain file
    plain = "t 3 2\nv 0 1 1\nv 1 2 2\nv 2 1 1\ne 0 1\ne 1 2\n"
    crlf = "t 3 2\r\n\r\nv 0 1 1\r\nv 1 2 2\r\n  v 2 1 1\r\n\r\ne 0 1\r\ne 1 2"
    a, b = tmp_path / "a.txt", tmp_path / "b.txt"
    a.write_text(plain)
    b.write_bytes(crlf.encode())
    ga, gb = load_text_parallel(str(a)), load_text_parallel(str(b))
    assert same(ga, gb) and same(ga, load_text(str(a)))
    assert ga.offsets.tolist() == [0, 1, 3, 4] and ga.neighbors.tolist() == [1, 0, 2, 1]
