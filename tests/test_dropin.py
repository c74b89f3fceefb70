"""The C++ drop-in layer (psm_benchmark_amd/dropin): the reference's class
names over the C ABI and a SubgraphMatching.out driver whose log has the
reference's keys (parsed with the same parser as the reference's logs)."""
import json
import os
import subprocess

import pytest

import ref_runner as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "psm_benchmark_amd", "SubgraphMatching.out")
HPRD = os.path.join(ROOT, "tests", "golden", "hprd")


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
