"""Benchmark of the batched matching path on one MI355X.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--workload C1|C2|C3|C4|C5]

Workload C2 (seeded, synthetic): Chung-Lu 1.1M vertices / 3.0M edges,
100 Zipf(1.1) labels; 512 random-walk queries of 4/8/16/24/32 vertices, half
dense and half sparse; GQL filter + GQL order; count per query capped at 1e5
and reported as min(cap, total), no time limit (every query completes). A
step is one batched pass over the 512 queries (filter, candidate-edge
tables, plans, enumeration) with the data graph resident in HBM. The JSON
line carries queries/s, embeddings/s, per-query p50/p99 latencies of the
batched form and of the plain one-query-at-a-time GPU path, the roofline of
the dominant kernel and the CPU baseline: the oracle restatement of the
reference path on the host cores at 1 and nproc threads, on a fixed sample
of the queries on which the GPU is timed as well.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

# The serving form runs one HIP stream per worker plus the resident launch's
# two; with HIP's default of 4 hardware queues per process, streams share
# queues and a stream's kernels can wait behind another's. Read by the HIP
# runtime when it starts, so set before anything touches the GPU.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E
PLAIN_SHAPE = "lane,1,2"   # the library's default launch shape (psm_ctx)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_sample(qs, stride: int = 33):
    """The CPU baseline's bounded sample: every stride-th query (stride odd and
    prime to the five sizes, so sizes and the dense/sparse halves all occur)."""
    return list(range(0, len(qs), stride))


def cpu_baseline(g, qs, method: str, cap: int, threads: int, time_limit: float):
    """The oracle restatement (oracle/psm_oracle.cpp, the project's own C++
    port of the reference path) on the host cores, `threads` enumeration
    workers (0: every hardware thread), each query under `time_limit` seconds
    counted from the start of its filtering. Queries that hit it are timeouts:
    their time is reported, they are not counted as completed."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_binding import OGraph, ORun
    d = OGraph(g)
    spent, done, emb, timeouts, used = 0.0, 0, 0, 0, 1
    phases = dict(filter_s=0.0, build_s=0.0, plan_s=0.0, enum_s=0.0)
    counts = []
    for q in qs:
        t = time.perf_counter()
        r = ORun(d, OGraph(q), method, cap=cap, threads=threads, time_limit=time_limit)
        spent += time.perf_counter() - t
        used = int(r.times.threads)
        for k in phases:
            phases[k] += getattr(r.times, k)
        if r.timed_out:
            timeouts += 1
            counts.append(None)
        else:
            done += 1
            emb += r.count
            counts.append(r.count)
    return dict(value=done / spent if spent else float("nan"), unit="queries/s", cores=used, kind="port",
                completed=done, timeouts=timeouts, seconds=round(spent, 3),
                embeddings_per_s=emb / spent if spent else float("nan"),
                phase_s={k: round(v, 3) for k, v in phases.items()}, counts=counts)


def copy_calibration(mib: int, reps: int) -> None:
    """A plain device copy of known size (mib MiB read + mib MiB written per
    launch), the reference point that scales the FETCH_SIZE / WRITE_SIZE
    counters into HBM bytes."""
    import torch
    n = mib << 20
    src = torch.empty(n, dtype=torch.uint8, device="cuda:0").random_(0, 255)
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n, "reps": reps}))


def _pmc_rows(path: str, counter: str):
    """(kernel name, value) per dispatch from a rocprofv3 counter_collection.csv
    tree (values summed over the counter's instances)."""
    import csv
    import glob
    per = {}
    for f in glob.glob(os.path.join(path, "**", "*counter_collection.csv"), recursive=True):
        with open(f, newline="") as fh:
            for row in csv.DictReader(fh):
                if row.get("Counter_Name") != counter:
                    continue
                key = (f, row.get("Dispatch_Id") or row.get("Correlation_Id"))
                name, val = row["Kernel_Name"], float(row["Counter_Value"])
                old = per.get(key)
                per[key] = (name, val + (old[1] if old else 0.0))
    return list(per.values())


def pmc_summary(fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, kernel, out_json):
    """HBM traffic per launch of the kernel class `kernel` (a name of KERNELS:
    its dispatches are those whose names match KERNEL_MATCH[kernel]):
    FETCH_SIZE and WRITE_SIZE (KiB per dispatch) scaled by what the same
    counters read for the copy calibration."""
    import re
    pattern = re.compile(KERNEL_MATCH.get(kernel, re.escape(kernel)))
    def avg(d, c, match):
        v = [x for n, x in _pmc_rows(d, c) if match(n)]
        return (sum(v) / len(v), len(v)) if v else (float("nan"), 0)
    is_copy = lambda n: "elementwise" in n.lower() or "copy" in n.lower()
    copy_bytes = None
    for d in (copy_fetch_dir, copy_write_dir):
        f = os.path.join(d, "copy.json")
        if os.path.exists(f):
            for line in open(f):
                if '"copy_bytes"' in line:
                    copy_bytes = json.loads(line)["copy_bytes"]
    cf, ncf = avg(copy_fetch_dir, "FETCH_SIZE", is_copy)
    cw, ncw = avg(copy_write_dir, "WRITE_SIZE", is_copy)
    kf, nkf = avg(fetch_dir, "FETCH_SIZE", lambda n: pattern.search(n) is not None)
    kw, nkw = avg(write_dir, "WRITE_SIZE", lambda n: pattern.search(n) is not None)
    sf = copy_bytes / (cf * 1024) if copy_bytes else float("nan")
    sw = copy_bytes / (cw * 1024) if copy_bytes else float("nan")
    res = dict(kernel=kernel, match=pattern.pattern, fetch_kib=kf, write_kib=kw, dispatches=[nkf, nkw],
               copy_bytes=copy_bytes, copy_fetch_kib=cf, copy_write_kib=cw, copy_dispatches=[ncf, ncw],
               scale_fetch=sf, scale_write=sw,
               traffic_bytes=kf * 1024 * sf + kw * 1024 * sw)
    with open(out_json, "w") as fh:
        json.dump(res, fh, indent=1)
    print(json.dumps(res))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="C2")
    ap.add_argument("--cap", type=int, default=100_000)
    ap.add_argument("--plain", type=int, default=512, help="queries timed through the plain path")
    ap.add_argument("--cpu-stride", type=int, default=17,
                    help="C2: the CPU baseline runs every N-th query (and the GPU is timed on the same ones)")
    ap.add_argument("--cpu-limit", type=float, default=60.0,
                    help="per-query time limit of the CPU baseline (s); a query that hits it is a timeout")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--queries", type=int, default=512, help="use the first N workload queries")
    ap.add_argument("--pipeline", type=int, default=2,
                    help="C2: batches in flight at once, each on its own context and HIP stream")
    ap.add_argument("--time-limit", type=float, default=None,
                    help="time limit in seconds (default none for C1 and C2, whose queries all "
                         "finish; 20 s for the uncapped C3 / C5 workloads)")
    ap.add_argument("--plain-limit", type=float, default=5.0,
                    help="C3 / C5: time limit of one plain-path query (s)")
    ap.add_argument("--plain-budget", type=float, default=60.0,
                    help="C3 / C5: total time spent on plain-path and join queries (s)")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--shape", default=None, metavar="STRATEGY,WAVES_PER_CU,TASKS_PER_WORKER",
                    help="set this launch shape instead of autotuning (e.g. lane,4,4)")
    ap.add_argument("--requests", type=int, default=1000, help="C4: requests per arrival rate")
    ap.add_argument("--no-serving", action="store_true", help="C2: leave out the C4 serving leg")
    ap.add_argument("--serving-plain", type=int, default=128, help="C2 line's C4 leg: plain-path queries")
    ap.add_argument("--serving-shape", default="lane,1,2", help="C2 line's C4 leg: launch shape of the server")
    ap.add_argument("--workers", type=int, default=8, help="C4: server workers (HIP streams)")
    ap.add_argument("--max-batch", type=int, default=128, help="C4: requests per batch at most")
    ap.add_argument("--low-rate", type=float, default=10.0, help="C4: extra rate below capacity (QPS)")
    ap.add_argument("--low-requests", type=int, default=100, help="C4: requests at the extra rate")
    ap.add_argument("--serve-mode", default="both", choices=("resident", "per-batch", "both"),
                    help="C4: admit prepared batches into one resident enumeration launch, run one launch per "
                         "batch, or both (resident first; the extra low rate only for it)")
    ap.add_argument("--tune-sample", type=int, default=128, help="queries per autotuner trial")
    ap.add_argument("--tune-limit", type=float, default=0.8, help="time limit of an autotuner trial (s)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain MIB-MiB device copy (PMC counter calibration)")
    ap.add_argument("--pmc-summary", nargs=6,
                    metavar=("FETCH_DIR", "WRITE_DIR", "COPY_FETCH_DIR", "COPY_WRITE_DIR", "KERNEL", "OUT"),
                    help="summarise rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs into OUT (json)")
    ap.add_argument("--traffic-dir", default=os.path.join(ROOT, "profiles"),
                    help="where the HBM traffic summaries (<class>_c2_batched_traffic.json, from "
                         "--pmc-summary) of the dominant kernel class are read")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration, 8)
        return
    if args.pmc_summary:
        pmc_summary(*args.pmc_summary)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (batched and served "
              "queries shard independently across GPUs, one copy of the data graph per GPU)", file=sys.stderr)
        sys.exit(2)
    if args.time_limit is None:
        args.time_limit = 0.0 if args.workload in ("C1", "C2") else 20.0
    if args.workload == "C1":
        plain_workload(args)
        return
    if args.workload in ("C3", "C5"):
        large_workload(args)
        return
    if args.workload == "C4":
        serving_workload(args)
        return
    if args.workload == "C4-batch":
        c4_batch_workload(args)
        return

    from psm_benchmark_amd import autotune
    from psm_benchmark_amd import workloads as W
    from psm_benchmark_amd.matcher import Batch, Device, DeviceGraph, Pipeline, match
    from psm_benchmark_amd import _lib as L

    t0 = time.time()
    g = W.data_graph(args.workload)
    qs, meta = W.c2_queries(g)
    qs, meta = qs[: args.queries], meta[: args.queries]
    log(f"workload {args.workload}: V={g.num_vertices} E={g.num_edges} queries={len(qs)} "
        f"({time.time() - t0:.1f}s to generate)")
    dev = Device(0)
    L.lib().psm_set_time_limit(dev.h, args.time_limit)
    t0 = time.time()
    dg = DeviceGraph(dev, g)
    log(f"graph resident in HBM ({time.time() - t0:.1f}s)")

    tuned, trials = None, []
    if args.shape:
        tuned = apply_shape(dev, args.shape)
    elif not args.no_tune:
        t0 = time.time()
        tuned, trials = autotune.tune(dg, qs, "GQL", "GQL", args.cap, sample=args.tune_sample,
                                      trial_limit_s=args.tune_limit, restore_limit_s=args.time_limit,
                                      log=log)
        log(f"autotune picked {tuned} ({time.time() - t0:.1f}s)")

    def step(queries):
        """One batched pass on the graph's own context: create, filter +
        tables + plans, enumeration."""
        t = time.perf_counter()
        b = Batch(dg, queries)
        try:
            b.prepare("GQL", "GQL")
            counts, calls, status, st = b.evaluate(args.cap)
            dt = time.perf_counter() - t
            lat = b.latencies()
        finally:
            b.close()
        return dt, counts, calls, status, st, lat

    # The timed steps: one batch of all the queries each, with --pipeline
    # batches in flight on their own contexts (HIP streams), so one batch's
    # filter and table launches run beside another's enumeration tail.
    pipe = Pipeline(dg, args.pipeline, time_limit=args.time_limit)
    for r in pipe.run([qs] * args.warmup, "GQL", "GQL", args.cap):
        st = r.stats
        log(f"warmup: {(r.end_s - r.start_s) * 1e3:.1f} ms (filter {st.filter_ms:.1f}, table {st.table_ms:.1f}, "
            f"plan {st.plan_ms:.1f}, enum {st.enum_ms:.1f})")
    pipe.kernel_stats_reset()
    t = time.perf_counter()
    runs = pipe.run([qs] * args.steps, "GQL", "GQL", args.cap)
    total = time.perf_counter() - t
    stats, lats, nodes = [], [], 0
    counts = runs[0].counts
    for r in runs:
        st = r.stats
        stats.append(st)
        lats.append(r.latencies_ms)
        nodes += int(r.calls.sum())
        if not np.array_equal(r.counts, counts):
            raise SystemExit("two batches of the same queries gave different counts")
        log(f"step: {(r.end_s - r.start_s) * 1e3:.1f} ms in flight (filter {st.filter_ms:.1f}, "
            f"table {st.table_ms:.1f}, plan {st.plan_ms:.1f}, enum {st.enum_ms:.1f}), "
            f"timeouts {int((r.status == L.PSM_ETIMEOUT).sum())}")
    status = runs[-1].status
    ks = kernel_stats(*pipe.devs)
    pipe.close()
    timeouts = int(max((r.status == L.PSM_ETIMEOUT).sum() for r in runs))
    nq = len(qs)
    emb = int(counts.sum())
    step_ms = total / args.steps * 1e3
    lat_all = np.concatenate(lats)

    # plain path: one query at a time through the same library, with the
    # library's default launch shape (what a caller gets without tuning; the
    # batched shape expands every query to a chip-wide task count)
    apply_shape(dev, PLAIN_SHAPE)
    log(f"plain path over {min(args.plain, nq)} queries")
    lat = []
    plain_counts, plain_timeouts = [], []
    for q in qs[: args.plain]:
        t = time.perf_counter()
        r = match(dg, q, "GQL_GQL_GQL", cap=args.cap)
        lat.append(time.perf_counter() - t)
        plain_counts.append(r.count)
        plain_timeouts.append(r.timed_out)
    mismatch = [i for i, c in enumerate(plain_counts)
                if c != counts[i] and status[i] != L.PSM_ETIMEOUT and not plain_timeouts[i]]
    if mismatch:
        raise SystemExit(f"batched and plain counts differ for queries {mismatch[:10]}")
    lat_ms = np.array(lat if lat else [float("nan")]) * 1e3
    plain_qps = len(lat) / sum(lat) if lat else float("nan")
    if tuned:
        autotune.set_config(dev, tuned)

    # dominant kernel roofline (bytes counted by the kernels themselves,
    # durations from HIP events on the library's stream)
    dom = max(ks.values(), key=lambda k: k["ms"])
    achieved = dom["bytes"] / dom["launches"] / (dom["ms"] / dom["launches"] / 1e3) / 1e9
    traffic = None
    tfile = os.path.join(args.traffic_dir, f"{dom['name']}_c2_batched_traffic.json")
    if os.path.exists(tfile):
        tr = json.load(open(tfile))
        if tr.get("kernel") == dom["name"] and tr.get("traffic_bytes") == tr.get("traffic_bytes"):
            traffic = int(tr["traffic_bytes"])
    roofline = dict(kernel=dom["name"], bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS,
                    unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                    avg_launch_ms=round(dom["ms"] / dom["launches"], 4),
                    bytes_per_launch=int(dom["bytes"] / dom["launches"]))
    if dom["name"] == "dfs":
        roofline["bytes_definition"] = ("search-tree nodes entered by the kernel: 4 B per entry of each "
                                        "operand row + 4 B per survivor written below the root unless "
                                        "a leaf (kernel-counted; equals the oracle's sum when uncapped)")
        roofline["touched_bytes_per_launch"] = int(sum(s.probe_bytes for s in stats) / args.steps)
    elif dom["name"] == "gql_refine":
        roofline["bytes_definition"] = ("GQL refinement (all its kernels; launch = one kernel of a (round, query "
                                        "vertex) group): per candidate tested 4 B + 16 B per left-label run "
                                        "located + 4 B x (1 + k) per label-run entry scanned (the neighbour and "
                                        "its membership word per left vertex), kernel-counted")

    cpu = cpu_n = gpu_sample = None
    if not args.no_cpu:
        idx = cpu_sample(qs, args.cpu_stride)
        sample = [qs[i] for i in idx]
        # the batched GPU form on exactly the CPU's queries (the like-for-like ratio)
        reps = []
        for _ in range(max(3, args.steps // 4)):
            dt, scounts, _, sstatus, _, _ = step(sample)
            reps.append(dt)
        gpu_sample = dict(queries=len(sample), ms_per_pass=round(float(np.mean(reps)) * 1e3, 3),
                          queries_per_s=round(len(sample) / float(np.mean(reps)), 2))
        sample_desc = (f"queries {idx[0]}, {idx[1]}, ..., {idx[-1]} (every {args.cpu_stride}th of the {nq}: "
                       f"{len(sample)} queries, all sizes, dense and sparse), GQL_GQL_GQL, cap {args.cap}, "
                       f"per-query limit {args.cpu_limit:g} s; the GPU batched form is timed on the same queries")
        log(f"cpu baseline, 1 thread, {len(sample)} queries")
        cpu = cpu_baseline(g, sample, "GQL_GQL_GQL", args.cap, 1, args.cpu_limit)
        nproc = min(os.cpu_count() or 1, 16)   # the GPU host's CPU share
        log(f"cpu baseline, {nproc} threads")
        cpu_n = cpu_baseline(g, sample, "GQL_GQL_GQL", args.cap, nproc, args.cpu_limit)
        for c in (cpu, cpu_n):
            bad = [i for i, x in zip(idx, c["counts"]) if x is not None and x != int(counts[i])]
            if bad:
                raise SystemExit(f"oracle and device counts differ for queries {bad[:10]}")
            c["sample"] = sample_desc
            del c["counts"]

    out = {
        "metric": "queries/s + embeddings/s on 1 MI355X; p50/p99 latency batched vs plain",
        "value": round(nq * args.steps / total, 2),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_ms, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded Chung-Lu graph and random-walk queries)",
        "config": {"workload": "C2 batched: Chung-Lu 1.1M V / 3.0M E, 100 Zipf(1.1) labels, "
                               "512 random-walk queries of 4-32 vertices (half dense), GQL_GQL_GQL, cap 1e5",
                   "queries": nq, "cap": args.cap, "time_limit_s": args.time_limit or None,
                   "batches_in_flight": args.pipeline},
        "timeouts": timeouts,
        "completed_queries_per_s": round((nq - timeouts) * args.steps / total, 2),
        "embeddings_per_s": round(emb * args.steps / total, 1),
        "embeddings_per_step": emb,
        "search_nodes_per_s": round(nodes / total, 1),
        "dfs_lanes_busy": round(sum(s.dfs_lane_steps for s in stats) / max(1, sum(s.dfs_lane_iters for s in stats)), 4),
        "launch_shape": tuned,
        "latency_ms": {"batched_p50": round(float(np.percentile(lat_all, 50)), 3),
                       "batched_p99": round(float(np.percentile(lat_all, 99)), 3),
                       "batched_definition": "per query: from the start of its batched step (the batch's "
                                             "creation on its context) until its count was final (cap "
                                             "reached or last task ended, device clock)",
                       "plain_p50": round(float(np.percentile(lat_ms, 50)), 3),
                       "plain_p99": round(float(np.percentile(lat_ms, 99)), 3)},
        "plain_gpu_queries_per_s": round(plain_qps, 2),
        "plain_launch_shape": PLAIN_SHAPE,
        "plain_timeouts": int(sum(plain_timeouts)),
        "batched_over_plain": round(nq * args.steps / total / plain_qps, 2),
        "stage_ms": {k: round(float(np.mean([getattr(s, k) for s in stats])), 2)
                     for k in ("filter_ms", "table_ms", "plan_ms", "enum_ms")},
        "kernels": {k: dict(launches=round(v["launches"] / args.steps, 2), ms=round(v["ms"] / args.steps, 3),
                            bytes=int(v["bytes"] / args.steps)) for k, v in ks.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_on_cpu_sample"] = gpu_sample
        out["batched_over_cpu"] = round(gpu_sample["queries_per_s"] / cpu["value"], 1)
        out["cpu_baseline_nproc"] = cpu_n
        out["batched_over_cpu_nproc"] = round(gpu_sample["queries_per_s"] / cpu_n["value"], 1)
    if not args.no_serving:
        # the serving form on its own workload (C4), reported beside the line
        dg.close()
        out["serving"] = run_serving(args, dev=dev, plain_queries=args.serving_plain, shape=args.serving_shape,
                                     tune=False)
    print(json.dumps(out))


def plain_workload(args):
    """C1, the reference's CPU-runnable case: Chung-Lu 9,460 V / 34,998 E, 307
    Zipf(1.1) labels, one 8-vertex random-walk query (seeded), GQL_GQL_GQL,
    uncapped. A step is one plain-path match of the query on the GPU (filter,
    tables, plan, enumeration; the graph resident in HBM), at the launch shape
    the autotuner picks for it (--shape overrides; the library's default
    shape is timed beside it). The reference binary runs the same query on the
    host at one thread as the CPU baseline, and its embedding count must
    equal the GPU's."""
    from psm_benchmark_amd import autotune
    from psm_benchmark_amd import workloads as W
    from psm_benchmark_amd import _lib as L
    from psm_benchmark_amd.matcher import Device, DeviceGraph, match

    g = W.data_graph("C1")
    q = W.c1_query(g)
    log(f"workload C1: V={g.num_vertices} E={g.num_edges} query n={q.num_vertices} e={q.num_edges}")
    dev = Device(0)
    L.lib().psm_set_time_limit(dev.h, args.time_limit)
    dg = DeviceGraph(dev, g)
    default_rate = None
    if args.shape:
        tuned = apply_shape(dev, args.shape)
    elif not args.no_tune:
        # the library default first, for reference
        apply_shape(dev, PLAIN_SHAPE)
        match(dg, q, "GQL_GQL_GQL")
        t = time.perf_counter()
        n_default = max(1, min(10, args.steps))
        for _ in range(n_default):
            match(dg, q, "GQL_GQL_GQL")
        default_rate = round(n_default / (time.perf_counter() - t), 2)
        tuned, _ = autotune.tune(dg, [q], "GQL", "GQL", 0, sample=1, trial_limit_s=max(1.0, args.tune_limit),
                                 restore_limit_s=args.time_limit, log=log)
        log(f"autotune picked {tuned}")
    else:
        tuned = apply_shape(dev, PLAIN_SHAPE)
    L.lib().psm_set_time_limit(dev.h, args.time_limit)
    for _ in range(args.warmup):
        match(dg, q, "GQL_GQL_GQL")
    L.lib().psm_kernel_stats_reset(dev.h)
    lat, r = [], None
    for _ in range(args.steps):
        t = time.perf_counter()
        r = match(dg, q, "GQL_GQL_GQL")
        lat.append(time.perf_counter() - t)
    ks = kernel_stats(dev)
    total = sum(lat)
    lat_ms = np.array(lat) * 1e3
    cpu = cpu_n = None
    if not args.no_cpu:
        nproc = min(os.cpu_count() or 1, 16)   # the GPU host's CPU share
        log("cpu baseline, 1 thread")
        cpu = cpu_baseline(g, [q], "GQL_GQL_GQL", 0, 1, args.cpu_limit)
        log(f"cpu baseline, {nproc} threads")
        cpu_n = cpu_baseline(g, [q], "GQL_GQL_GQL", 0, nproc, args.cpu_limit)
        for c in (cpu, cpu_n):
            if c["counts"][0] is not None and c["counts"][0] != r.count:
                raise SystemExit(f"oracle count {c['counts'][0]} != device count {r.count}")
            c["sample"] = f"the same C1 query once, GQL_GQL_GQL uncapped, limit {args.cpu_limit:g} s"
            del c["counts"]
    out = {
        "metric": "queries/s + embeddings/s on 1 MI355X; p50/p99 latency batched vs plain",
        "value": round(args.steps / total, 2), "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(total / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic (seeded graph and query)",
        "config": {"workload": "C1 plain: Chung-Lu 9,460 V / 34,998 E, 307 Zipf(1.1) labels, one 8-vertex "
                               "random-walk query, GQL_GQL_GQL, uncapped", "time_limit_s": args.time_limit},
        "embeddings": r.count, "calls": r.calls, "timed_out": r.timed_out,
        "launch_shape": tuned, "default_shape_queries_per_s": default_rate, "default_shape": PLAIN_SHAPE,
        "embeddings_per_s": round(r.count * args.steps / total, 1),
        "latency_ms": {"p50": round(float(np.percentile(lat_ms, 50)), 3),
                       "p99": round(float(np.percentile(lat_ms, 99)), 3)},
        "kernels": {k: dict(launches=round(v["launches"] / args.steps, 2), ms=round(v["ms"] / args.steps, 4),
                            bytes=int(v["bytes"] / args.steps)) for k, v in ks.items()},
        "roofline": roofline_of(ks, args.steps),
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu"] = round(out["value"] / cpu["value"], 2)
        out["cpu_baseline_nproc"] = cpu_n
        out["gpu_over_cpu_nproc"] = round(out["value"] / cpu_n["value"], 2)
    print(json.dumps(out))


def apply_shape(dev, text):
    """STRATEGY,WAVES_PER_CU,TASKS_PER_WORKER[,BLOCK_LANES] (block lanes of the
    lane DFS; 0 or absent: the library's choice)."""
    from psm_benchmark_amd import autotune
    f = text.split(",")
    st, w, e = f[:3]
    cfg = {"dfs_strategy": st, "dfs_waves_per_cu": int(w), "expand_tasks_per_lane": float(e),
           "dfs_block_lanes": int(f[3]) if len(f) > 3 else 0}
    autotune.set_config(dev, cfg)
    return cfg


def roofline_of(ks, steps, traffic_file=None):
    dom = max(ks.values(), key=lambda k: k["ms"])
    achieved = dom["bytes"] / dom["launches"] / (dom["ms"] / dom["launches"] / 1e3) / 1e9
    traffic = None
    if traffic_file and os.path.exists(traffic_file):
        tr = json.load(open(traffic_file))
        if tr.get("kernel") == dom["name"] and tr.get("traffic_bytes") == tr.get("traffic_bytes"):
            traffic = int(tr["traffic_bytes"])
    return dict(kernel=dom["name"], bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS,
                unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                avg_launch_ms=round(dom["ms"] / dom["launches"], 4),
                bytes_per_launch=int(dom["bytes"] / dom["launches"]))


def large_workload(args):
    """C3 (Chung-Lu 4M V / 35M E, 32 Zipf(1.0) labels, 64 cyclic queries of
    12-16 vertices, CFL_CFL_EXPLORE, plus the two-unit join) and C5 (R-MAT
    2^24 V / 100M E, 8 uniform labels, 32 dense queries of 6-10 vertices with
    triangles and 4-cliques, GQL_GQL_GQL): uncapped counts under a time limit.
    The batched form runs all queries in one pipeline; the plain path runs
    query by query within a time budget and must agree with it on every
    query that neither stopped early. C3 also compares the uncapped plain
    count with the two-unit device join on 16 join-checkable queries of the
    same graph (cyclic walks over the rarest labels, finite counts)."""
    from psm_benchmark_amd import autotune
    from psm_benchmark_amd import workloads as W
    from psm_benchmark_amd import _lib as L
    from psm_benchmark_amd.matcher import Device, DeviceGraph, batch_match, join_match, match

    name = args.workload
    method = "CFL_CFL_EXPLORE" if name == "C3" else "GQL_GQL_GQL"
    f, o, _ = method.split("_")
    t0 = time.time()
    g = W.data_graph(name)
    qs, meta = (W.c3_queries if name == "C3" else W.c5_queries)(g)
    qs, meta = qs[: args.queries], meta[: args.queries]
    log(f"workload {name}: V={g.num_vertices} E={g.num_edges} queries={len(qs)} ({time.time() - t0:.1f}s)")
    dev = Device(0)
    dg = DeviceGraph(dev, g)
    tuned = None
    if args.shape:
        tuned = apply_shape(dev, args.shape)
    elif not args.no_tune:
        tuned, _ = autotune.tune(dg, qs, f, o, 0, sample=len(qs), trial_limit_s=max(1.0, args.tune_limit),
                                 restore_limit_s=args.time_limit, log=log)
        log(f"autotune picked {tuned}")
    L.lib().psm_set_time_limit(dev.h, args.time_limit)
    L.lib().psm_kernel_stats_reset(dev.h)
    times, nodes = [], 0
    for _ in range(args.steps):
        t = time.perf_counter()
        counts, calls, status, st = batch_match(dg, qs, f, o, 0)
        times.append(time.perf_counter() - t)
        nodes += int(calls.sum())
        log(f"batched step: {times[-1] * 1e3:.1f} ms (filter {st.filter_ms:.1f}, table {st.table_ms:.1f}, "
            f"plan {st.plan_ms:.1f}, enum {st.enum_ms:.1f}), timeouts {int((status == L.PSM_ETIMEOUT).sum())}")
    ks = kernel_stats(dev)
    total = sum(times)
    done = status != L.PSM_ETIMEOUT
    # plain path (and the join for C3), query by query within the budget,
    # with the library's default launch shape
    apply_shape(dev, PLAIN_SHAPE)
    L.lib().psm_set_time_limit(dev.h, args.plain_limit)
    plain, joins, lat = [], [], []
    spent = 0.0
    for i, q in enumerate(qs):
        if spent > args.plain_budget:
            break
        t = time.perf_counter()
        r = match(dg, q, method)
        lat.append(time.perf_counter() - t)
        spent += lat[-1]
        plain.append((i, r.count, r.timed_out))
        if name == "C3" and not r.timed_out:
            t = time.perf_counter()
            try:
                js = join_match(dg, q, method)
                joins.append((i, int(js.count), js.count == r.count))
            except L.PSMError as e:
                joins.append((i, None, str(e)))
            spent += time.perf_counter() - t
    if name == "C3":
        # join-checkable queries (rare labels): uncapped plain count vs the
        # two-unit device join, on the same resident C3 graph
        jq, _ = W.c3_join_queries(g)
        for i, q in enumerate(jq):
            r = match(dg, q, method)
            if r.timed_out:
                joins.append((f"j{i}", None, "plain timed out"))
                continue
            t = time.perf_counter()
            try:
                js = join_match(dg, q, method)
                joins.append((f"j{i}", int(js.count), js.count == r.count))
                log(f"join query j{i}: n={q.num_vertices} plain {r.count} join {js.count} "
                    f"rows {js.rows_a}/{js.rows_b} ({(time.perf_counter() - t) * 1e3:.1f} ms)")
            except L.PSMError as e:
                joins.append((f"j{i}", None, str(e)))
    mismatch = [i for i, c, to in plain if not to and done[i] and c != counts[i]]
    if mismatch:
        raise SystemExit(f"batched and plain counts differ for queries {mismatch[:10]}")
    bad_join = [i for i, c, okj in joins if okj is False]
    if bad_join:
        raise SystemExit(f"join and plain counts differ for queries {bad_join[:10]}")
    lat_ms = np.array(lat if lat else [float("nan")]) * 1e3
    nq = len(qs)
    out = {
        "metric": "queries/s + embeddings/s on 1 MI355X; p50/p99 latency batched vs plain",
        "value": round(nq * args.steps / total, 3),
        "unit": "queries/s",
        "n_gpus": 1, "steps": args.steps, "warmup": 0,
        "ms_per_step": round(total / args.steps * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (seeded graph and queries)",
        "config": {"workload": ("C3: Chung-Lu 4M V / 35M E, 32 Zipf(1.0) labels, 64 cyclic queries of "
                                "12-16 vertices, CFL_CFL_EXPLORE, uncapped" if name == "C3" else
                                "C5: R-MAT(0.57,0.19,0.19) 2^24 V / 100M E, 8 uniform labels, 32 dense "
                                "queries of 6-10 vertices with triangles and 4-cliques, GQL_GQL_GQL, uncapped"),
                   "queries": nq, "time_limit_s": args.time_limit},
        "timeouts": int((~done).sum()),
        "completed_queries_per_s": round(int(done.sum()) * args.steps / total, 3),
        "embeddings_per_s": round(float(counts.sum()) * args.steps / total, 1),
        "embeddings_completed": int(counts[done].sum()),
        "search_nodes_per_s": round(nodes / total, 1),
        "launch_shape": tuned,
        "plain": {"queries": len(plain), "timeouts": int(sum(to for _, _, to in plain)),
                  "time_limit_s": args.plain_limit,
                  "p50_ms": round(float(np.percentile(lat_ms, 50)), 2),
                  "p99_ms": round(float(np.percentile(lat_ms, 99)), 2),
                  "queries_per_s": round(len(lat) / max(1e-9, sum(lat)), 3)},
        "kernels": {k: dict(launches=round(v["launches"] / args.steps, 2), ms=round(v["ms"] / args.steps, 3),
                            bytes=int(v["bytes"] / args.steps)) for k, v in ks.items()},
        "roofline": roofline_of(ks, args.steps),
        "cpu_baseline": None,
    }
    if name == "C3":
        out["join"] = {"queries": len(joins), "equal_to_plain": int(sum(1 for j in joins if j[2] is True)),
                       "embeddings_checked": int(sum(j[1] for j in joins if j[2] is True)),
                       "not_compared": {str(j[0]): j[2] for j in joins if isinstance(j[2], str)}}
    print(json.dumps(out))


def run_serving(args, g=None, dev=None, plain_queries=None, shape=None, tune=True):
    """C4: the C3 graph resident in HBM; random-walk queries of 4-32 vertices
    (dense / sparse 50/50) arrive as seeded Poisson streams at 200, 1,000 and
    5,000 QPS into the continuous-batching server (memory-based admission,
    cap 1e5, GQL_GQL_GQL); each result is returned as soon as its count is
    final on the device. Reports p50/p99 latency (completion - arrival) and
    sustained queries/s per rate beside the plain one-query-at-a-time path,
    and checks every served count against the plain count of the same query.
    Returns the summary dict."""
    from psm_benchmark_amd import autotune
    from psm_benchmark_amd import workloads as W
    from psm_benchmark_amd import _lib as L
    from psm_benchmark_amd.matcher import Device, DeviceGraph, match
    from psm_benchmark_amd.serving import Server, run_poisson

    t0 = time.time()
    if g is None:
        g = W.data_graph("C4")
    qs, meta = W.c2_queries(g, count=args.queries, seed=44)
    log(f"workload C4: V={g.num_vertices} E={g.num_edges} queries={len(qs)} ({time.time() - t0:.1f}s)")
    dev = dev or Device(0)
    dg = DeviceGraph(dev, g)
    tuned = None
    if shape:
        tuned = apply_shape(dev, shape)
    elif tune:
        tuned, _ = autotune.tune(dg, qs, "GQL", "GQL", args.cap, sample=min(len(qs), args.tune_sample),
                                 trial_limit_s=args.tune_limit, restore_limit_s=args.time_limit, log=log)
        log(f"autotune picked {tuned}")
    L.lib().psm_set_time_limit(dev.h, args.time_limit)
    # plain path: one query at a time, with the library's default launch shape
    served_shape = autotune.get_config(dev)
    apply_shape(dev, PLAIN_SHAPE)
    lat, plain_counts, plain_to = [], [], []
    for q in qs[: plain_queries if plain_queries is not None else args.plain]:
        t = time.perf_counter()
        r = match(dg, q, "GQL_GQL_GQL", cap=args.cap)
        lat.append(time.perf_counter() - t)
        plain_counts.append(r.count)
        plain_to.append(r.timed_out)
    lat_ms = np.array(lat) * 1e3
    plain = {"queries": len(lat), "queries_per_s": round(len(lat) / sum(lat), 2), "launch_shape": PLAIN_SHAPE,
             "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
             "p99_ms": round(float(np.percentile(lat_ms, 99)), 3), "timeouts": int(sum(plain_to))}
    log(f"plain: {plain}")
    autotune.set_config(dev, served_shape)   # server workers inherit the served shape
    modes = ["resident", "per-batch"] if args.serve_mode == "both" else [args.serve_mode]
    by_mode = {}
    for mode in modes:
        rates = []
        # the three offered rates of the workload, and (first mode only) one
        # below the server's capacity so that unloaded latency shows beside them
        for rate, nreq in ((200.0, args.requests), (1000.0, args.requests), (5000.0, args.requests),
                           (args.low_rate, args.low_requests if mode == modes[0] else 0)):
            if nreq <= 0:
                continue
            srv = Server(dg, workers=args.workers, max_batch=args.max_batch, filter="GQL", order="GQL",
                         cap=args.cap, time_limit=args.time_limit, resident=mode == "resident")
            try:
                res, summary = run_poisson(srv, qs, rate_qps=rate, n_requests=nreq, seed=int(rate))
                summary["server"] = srv.stats()
            finally:
                srv.close()
            bad = [k for k, r in res.items() if k % len(qs) < len(plain_counts) and r.status == 0
                   and not plain_to[k % len(qs)] and r.count != plain_counts[k % len(qs)]]
            if bad:
                raise SystemExit(f"served counts differ from the plain path for requests {bad[:10]}")
            summary["checked_against_plain"] = sum(1 for k in res if k % len(qs) < len(plain_counts))
            log(f"{mode} rate {rate}: {summary}")
            rates.append(summary)
        by_mode[mode] = rates
    dg.close()
    best = {m: max(r["sustained_qps"] for r in by_mode[m][:3]) for m in modes}
    out = {"workload": "C4 serving: Chung-Lu 4M V / 35M E resident, 32 Zipf(1.0) labels, random-walk queries "
                       "of 4-32 vertices (dense/sparse 50/50), Poisson 200/1000/5000 QPS, GQL_GQL_GQL, cap 1e5",
           "requests_per_rate": args.requests, "workers": args.workers, "max_batch": args.max_batch,
           "serve_mode": modes[0], "rates": by_mode[modes[0]], "plain": plain, "best_sustained_qps": best[modes[0]],
           "served_over_plain": round(best[modes[0]] / plain["queries_per_s"], 2), "launch_shape": tuned}
    for m in modes[1:]:
        key = m.replace("-", "_")
        out[f"rates_{key}"] = by_mode[m]
        out[f"best_sustained_qps_{key}"] = best[m]
    return out


def serving_workload(args):
    """The C4 serving workload alone (--workload C4)."""
    sv = run_serving(args, shape=args.shape, tune=not args.no_tune)
    out = {
        "metric": "queries/s + embeddings/s on 1 MI355X; p50/p99 latency batched vs plain",
        "value": sv["best_sustained_qps"], "unit": "queries/s", "n_gpus": 1, "steps": len(sv["rates"]),
        "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32", "data": "synthetic (seeded graph, queries and Poisson arrivals)",
        "config": {"workload": sv["workload"], "requests_per_rate": args.requests, "workers": args.workers,
                   "max_batch": args.max_batch, "cap": args.cap, "time_limit_s": args.time_limit,
                   "serve_mode": args.serve_mode},
        "rates": sv["rates"], "plain": sv["plain"], "served_over_plain": sv["served_over_plain"],
        "launch_shape": sv["launch_shape"], "cpu_baseline": None,
    }
    for k, v in sv.items():
        if k.startswith("rates_") or k.startswith("best_sustained_qps_"):
            out[k] = v
    print(json.dumps(out))


def c4_batch_workload(args):
    """The C4 requests as plain batches (--workload C4-batch): the C3 graph
    resident, the first --queries (default 256 here) C4 random-walk queries
    as one batch per step, GQL_GQL_GQL, cap 1e5, launch shape --shape or the
    library default. A diagnostic line for the filter and table kernels that
    bound the serving form (stage and per-kernel times), not a headline."""
    from psm_benchmark_amd import workloads as W
    from psm_benchmark_amd import _lib as L
    from psm_benchmark_amd.matcher import Batch, Device, DeviceGraph

    nq = args.queries if args.queries != 512 else 256
    g = W.data_graph("C4")
    qs, _ = W.c2_queries(g, count=max(nq, 1), seed=44)
    qs = qs[:nq]
    dev = Device(0)
    L.lib().psm_set_time_limit(dev.h, args.time_limit)
    dg = DeviceGraph(dev, g)
    if args.shape:
        apply_shape(dev, args.shape)
    stats, times, timeouts = [], [], 0
    for i in range(args.warmup + args.steps):
        if i == args.warmup:
            L.lib().psm_kernel_stats_reset(dev.h)
        t = time.perf_counter()
        b = Batch(dg, qs)
        try:
            b.prepare("GQL", "GQL")
            counts, calls, status, st = b.evaluate(args.cap)
        finally:
            b.close()
        if i >= args.warmup:
            times.append(time.perf_counter() - t)
            stats.append(st)
            timeouts += int((status == L.PSM_ETIMEOUT).sum())
        log(f"batch {i}: {(time.perf_counter() - t) * 1e3:.1f} ms (filter {st.filter_ms:.1f}, "
            f"table {st.table_ms:.1f}, enum {st.enum_ms:.1f})")
    ks = kernel_stats(dev)
    dg.close()
    total = sum(times)
    print(json.dumps({
        "metric": "queries/s of C4 requests as plain batches (diagnostic)", "value": round(nq * len(times) / total, 2),
        "unit": "queries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(total / len(times) * 1e3, 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic (seeded graph and random-walk queries)",
        "config": {"workload": f"C4 requests as batches: Chung-Lu 4M V / 35M E, 32 Zipf(1.0) labels, {nq} "
                               "random-walk queries of 4-32 vertices per batch, GQL_GQL_GQL, cap 1e5",
                   "queries": nq, "cap": args.cap},
        "timeouts": timeouts,
        "stage_ms": {k: round(float(np.mean([getattr(s, k) for s in stats])), 2)
                     for k in ("filter_ms", "table_ms", "plan_ms", "enum_ms")},
        "kernels": {k: dict(launches=round(v["launches"] / len(times), 2), ms=round(v["ms"] / len(times), 3),
                            bytes=int(v["bytes"] / len(times))) for k, v in ks.items()},
        "roofline": None, "cpu_baseline": None}))


KERNELS = ["gql_refine", "table_count", "table_fill", "dfs", "expand", "compact", "bitmap"]
# the device kernels of each class, by name (rocprofv3 counter summaries)
KERNEL_MATCH = {"gql_refine": r"k_gql_", "table_count": r"k_table_count", "table_fill": r"k_table_fill",
                "dfs": r"k_dfs", "expand": r"k_expand", "compact": r"k_compact_",
                "bitmap": r"k_set_(bits|member)|k_slot_prefix"}


def kernel_stats(*devs):
    """Launches, HIP-event milliseconds and kernel-counted bytes per kernel
    class, summed over the contexts."""
    import ctypes

    from psm_benchmark_amd import _lib as L
    out = {}
    for i, name in enumerate(KERNELS):
        tot = [0.0, 0.0, 0.0]
        for dev in devs:
            a = (ctypes.c_double * 3)()
            L.lib().psm_kernel_stats(dev.h, i, a)
            tot = [x + y for x, y in zip(tot, a)]
        if tot[0] > 0:
            out[name] = dict(name=name, launches=int(tot[0]), ms=float(tot[1]), bytes=float(tot[2]))
    return out


if __name__ == "__main__":
    main()
