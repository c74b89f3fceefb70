"""Runs the reference binary built from its own sources (oracle/_ref,
recipe oracle/Makefile.ref) on graph files and parses its log.

Test infrastructure only: a checker for the oracle on synthetic inputs and the
one-time calibration of the oracle restatement. Absent binary -> available() is False.
"""
from __future__ import annotations

import os
import re
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "SubgraphMatching.out")


def available() -> bool:
    return os.path.exists(BIN) and os.access(BIN, os.X_OK)


def parse_log(path):
    blocks, cur = [], None
    with open(path) as f:
        for line in f:
            line = re.sub(r"^\[Thread \d+\] ", "", line.rstrip("\n"))
            if line == "Start the query":
                cur = {}
            elif line == "End the query":
                blocks.append(cur)
                cur = None
            elif cur is not None and ": " in line:
                k, v = line.split(": ", 1)
                cur[k] = v
    return blocks


def run(data_path: str, query_path: str, methods: str = "GQL_GQL_GQL", threads: str = "1",
        cap: str = "MAX", time_limit: int = 600, timeout: float | None = None):
    """Returns one dict per (threads, method) block of the reference's log."""
    with tempfile.TemporaryDirectory() as td:
        log = os.path.join(td, "log.txt")
        cmd = [BIN, "-d", data_path, "-q", query_path, "-num", cap, "-time_limit", str(time_limit),
               "-threadnums", threads, "-split", "linear", "-schedule", "busy2idlenostop",
               "-QorCandi", "C", "-BackMethods", methods, "-OutputFile", log]
        subprocess.run(cmd, check=True, timeout=timeout, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL)
        out = []
        for b in parse_log(log):
            out.append({
                "threads": int(b["Num Threads"]),
                "method": f'{b["Filter Type"]}_{b["Order Type"]}_{b["Engine Type"]}',
                "embeddings": int(b["#Embeddings"]),
                "calls": int(b["Call Count"]),
                "mem_words": round(float(b["Memory cost (MB)"]) * 2**20 / 4),
                "overtime": int(b["Overtime"]),
                "filter_s": float(b.get("Filter vertices time (seconds)", 0)),
                "build_s": float(b.get("Build table time (seconds)", 0)),
                "plan_s": float(b.get("Generate query plan time (seconds)", 0)),
                "enum_s": float(b.get("Enumerate time (seconds)", 0)),
                "total_s": float(b.get("Total time (seconds)", 0)),
            })
        return out


def deep_dir(base: str) -> str:
    """The reference's driver needs >= 4 '/' in query paths (StudyPerformance.cpp:39-56)."""
    d = os.path.join(base, "a", "b", "c", "d")
    os.makedirs(d, exist_ok=True)
    return d
