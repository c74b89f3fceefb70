"""Writes tests/golden/hprd/golden.json from the reference's recorded logs.

Source: test_anony/C_output/HPRD/L15/Q10/Q10-{1..10}.txt of the reference
(runs of SubgraphMatching.out on datasets/HPRD/L15 with -time_limit 2,
uncapped). Each log holds blocks "Start the query" .. "End the query"; this
keeps, per (query, threads, filter, order, engine), the embedding count, the
call count and the memory cost converted to 4-byte words
(round(MB * 2**20 / 4), the BuildTable checksum).

Run once in a container that has the reference mounted:
    python tests/golden/extract_hprd_golden.py /root/reference
"""
import json
import os
import re
import sys


def parse(path):
    blocks, cur = [], None
    for line in open(path):
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


def main(ref):
    out = []
    base = os.path.join(ref, "test_anony/C_output/HPRD/L15/Q10")
    for i in range(1, 11):
        for b in parse(os.path.join(base, f"Q10-{i}.txt")):
            out.append({
                "query": f"Q10-{i}",
                "threads": int(b["Num Threads"]),
                "filter": b["Filter Type"],
                "order": b["Order Type"],
                "engine": b["Engine Type"],
                "embeddings": int(b["#Embeddings"]),
                "calls": int(b["Call Count"]),
                "mem_words": round(float(b["Memory cost (MB)"]) * 2**20 / 4),
                "overtime": int(b["Overtime"]),
            })
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "hprd", "golden.json"), "w") as f:
        json.dump(out, f, indent=0)
    print(len(out), "records")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
