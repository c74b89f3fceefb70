"""Autotuner of the enumeration launch shape.

Five knobs of the enumeration: the search strategy (one search per lane,
or one per wavefront with chunked cooperative tests), the persistent DFS
grid (wavefronts per CU), the task pre-expansion target (tasks per DFS
worker before the persistent kernel starts, i.e. the frontier chunk the BFS
phase hands to the DFS phase), the pre-expansion depth limit (the level at
which the DFS stacks start) and the set-intersection strategy (galloping
cursors and LDS-staged cooperative windows, or binary search alone). The
first three are searched on a grid; the last two are then tried around the
best shape. They trade occupancy against memory-system
contention and load balance, and the best point depends on the data graph
and the query mix. The tuner prepares (filter, tables, plans) a sample of
the workload once, runs its enumeration for each candidate setting (twice,
averaged: the sample's enumeration time is set by a few tails and varies)
under a short time limit and keeps the best by `pick`: among trials that
finish every query of the sample, those within 10 % of the shortest
enumeration are near ties, and the one with the highest search throughput
among them wins (with per-query caps, a wider search overshoots the caps and
does more work for the same counts, so throughput alone would pick a slower
setting; among near ties it favours the shape that keeps a long search
busiest); trials that run into the limit rank below those and by search
throughput (assignments per millisecond, the reference's "Call Count" rate).
Results never depend on the setting: only the launch shape changes.
"""
from __future__ import annotations

import ctypes
import json
import os

from . import _lib as L
from .matcher import Batch

# candidate launch shapes: (strategy, wavefronts per CU, tasks per worker)
GRID = [("lane", w, e) for w in (1, 2, 4, 8) for e in (2.0, 4.0, 8.0)] + \
       [("wave", w, e) for w in (1, 2, 4) for e in (1.0, 4.0)]
NAMES = ("dfs_strategy", "dfs_waves_per_cu", "expand_tasks_per_lane")
# tried around the best grid shape: (intersection strategy, pre-expansion depth)
REFINE = [("adaptive", 31), ("bsearch", 31), ("adaptive", 8), ("bsearch", 8), ("adaptive", 4)]
REPEATS = 2   # evaluations per trial, averaged
KEYS = {
    "dfs_waves_per_cu": L.PSM_TUNE_DFS_WAVES_PER_CU,
    "expand_tasks_per_lane": L.PSM_TUNE_EXPAND_TASKS_PER_LANE,
    "dfs_strategy": L.PSM_TUNE_DFS_STRATEGY,
    "claim_divisor": L.PSM_TUNE_CLAIM_DIVISOR,
    "dfs_block_lanes": L.PSM_TUNE_DFS_BLOCK_LANES,
    "expand_depth": L.PSM_TUNE_EXPAND_DEPTH,
    "intersect": L.PSM_TUNE_INTERSECT,
}
STRATEGIES = {"lane": L.PSM_DFS_LANE, "wave": L.PSM_DFS_WAVE}
INTERSECTS = {"adaptive": L.PSM_INTERSECT_ADAPTIVE, "bsearch": L.PSM_INTERSECT_BSEARCH}
CHOICES = {"dfs_strategy": STRATEGIES, "intersect": INTERSECTS}


def get_config(dev) -> dict:
    out = {}
    for name, key in KEYS.items():
        v = ctypes.c_double()
        r = L.lib().psm_get_tuning(dev.h, key, ctypes.byref(v))
        if r != L.PSM_OK:
            raise L.PSMError(r, f"psm_get_tuning {name}")
        out[name] = v.value
    for name, table in CHOICES.items():
        out[name] = {c: n for n, c in table.items()}[int(out[name])]
    return out


def set_config(dev, cfg: dict) -> None:
    for name, value in cfg.items():
        if name in CHOICES:
            value = CHOICES[name][value]
        r = L.lib().psm_set_tuning(dev.h, KEYS[name], float(value))
        if r != L.PSM_OK:
            raise L.PSMError(r, f"psm_set_tuning {name}={value}: {L.lib().psm_last_error(dev.h).decode()}")


def trial_key(t: dict):
    """Sort key of a trial (smaller is better): fewer timed-out queries
    first; then, when none timed out, the shorter enumeration; otherwise the
    higher search throughput."""
    return (t["timeouts"], t["enum_ms"] if t["timeouts"] == 0 else -t["nodes_per_ms"])


NEAR_TIE = 1.10   # completed trials within 10 % of the shortest are near ties


def pick(trials: list) -> dict:
    """The chosen trial: the best by trial_key, or, when it finished every
    query, the highest-throughput trial among its near ties."""
    best = min(trials, key=trial_key)
    if best["timeouts"]:
        return best
    ties = [t for t in trials if t["timeouts"] == 0 and t["enum_ms"] <= NEAR_TIE * best["enum_ms"]]
    return max(ties, key=lambda t: t["nodes_per_ms"])


def tune(dg, queries, filter: str = "GQL", order: str = "GQL", cap: int = 0, sample: int = 128,
         trial_limit_s: float = 0.8, restore_limit_s: float = 60.0, grid=None, refine=None,
         cache: str | None = None, log=None):
    """Picks and sets the launch shape for `queries` on `dg`.

    sample: queries used per trial (every k-th of the workload);
    trial_limit_s: enumeration time limit of a trial; restore_limit_s: the
    context's limit afterwards; grid: (strategy, waves per CU, tasks per worker)
    tuples, timed with the default intersection and depth; refine:
    (intersection strategy, pre-expansion depth) pairs timed with the best
    grid shape (REFINE by default, [] to skip). Returns (best config, list of
    trials)."""
    grid = grid or GRID
    refine = REFINE if refine is None else refine
    dev = dg.dev
    step = max(1, len(queries) // max(1, sample))
    sub = list(queries[::step])[:sample]
    info = dg.info()
    sig = dict(V=info["V"], E=info["E"], labels=info["labels"], queries=len(sub),
               sizes=sorted({q.num_vertices for q in sub}), filter=filter, order=order, cap=cap)
    if cache and os.path.exists(cache):
        try:
            saved = json.load(open(cache))
            if saved.get("signature") == sig:
                set_config(dev, saved["best"])
                return saved["best"], saved["trials"]
        except (ValueError, KeyError):
            pass
    trials = []
    # filter, tables and plans once; each trial times the enumeration alone
    # under its own limit
    b = Batch(dg, sub)

    def trial(cfg):
        set_config(dev, cfg)
        runs = []
        for _ in range(REPEATS):
            counts, calls, status, st = b.evaluate(cap, trial_limit_s)
            runs.append((int(calls.sum()), st.enum_ms, int((status == L.PSM_ETIMEOUT).sum())))
        nodes = sum(r[0] for r in runs) // len(runs)
        ms = sum(r[1] for r in runs) / len(runs)
        rate = nodes / ms if ms > 0 else 0.0
        trials.append(dict(config=cfg, nodes=nodes, enum_ms=round(ms, 3), nodes_per_ms=round(rate, 1),
                           timeouts=max(r[2] for r in runs)))
        if log:
            log(f"autotune {cfg}: {nodes} nodes in {ms:.1f} ms ({rate:.0f}/ms)")

    try:
        b.prepare(filter, order)
        base = dict(intersect="adaptive", expand_depth=31)
        for values in grid:
            trial({**dict(zip(NAMES, values)), **base})
        shape = {k: pick(trials)["config"][k] for k in NAMES}
        for inter, depth in refine:
            if (inter, depth) != (base["intersect"], base["expand_depth"]):
                trial({**shape, "intersect": inter, "expand_depth": depth})
    finally:
        b.close()
        L.lib().psm_set_time_limit(dev.h, restore_limit_s)
    best = pick(trials)["config"]
    set_config(dev, best)
    if cache:
        with open(cache, "w") as fh:
            json.dump(dict(signature=sig, best=best, trials=trials), fh, indent=1)
    return best, trials
