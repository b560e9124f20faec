"""Tile-shape autotuner for the grouped block GEMM.

For each contraction artifact of a workload the tuner times the workload
step with every GEMM tile configuration of that artifact (the other
artifacts held fixed) and keeps the fastest -- coordinate descent over the
artifacts, in the order they run.  A workload that can record its step as a
hipGraph is timed on graph replays: eager steps of a few tens of
microseconds are bound by host issue, and their timings would rank the
configurations by noise.  Results are keyed by the artifact's
job-shape signature (the sorted (m, k, n) list), kept on the Runtime and,
when TENET_AMD_TILE_CACHE names a file, persisted as JSON so later runs
start tuned.
"""
import hashlib
import json
import os
import time

from . import _native as N

CONFIGS = (0, 1, 2, 3, 5, 6)  # 128x128, 64x64, 128x64, 64x128, 128x128 on 8 waves, 32x32 per wave
                              # (narrow jobs keep the vector kernel)
MARGIN = 0.03  # relative step-time gain a configuration needs to replace the current best


def signature(contraction):
    d = contraction.desc.contents
    shapes = sorted((d.jobs[i].rows, d.jobs[i].contracted, d.jobs[i].cols) for i in range(d.n_jobs))
    return hashlib.sha1(repr(shapes).encode()).hexdigest()[:16]


def _cache_path():
    return os.environ.get("TENET_AMD_TILE_CACHE")


def load_cache():
    p = _cache_path()
    if p and os.path.exists(p):
        with open(p) as fh:
            return json.load(fh)
    return {}


def save_cache(table):
    p = _cache_path()
    if p:
        with open(p, "w") as fh:
            json.dump(table, fh, indent=1, sort_keys=True)


def apply(rt, contraction, cfg):
    rt._check(N.hip().tnt_contract_set_tile_config(rt._ctx, contraction.art, int(cfg)))
    contraction.tile_cfg = int(cfg)


def _time(rt, step, reps):
    step()
    rt.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        step()
    rt.synchronize()
    return (time.perf_counter() - t) / reps


def _time_graph(rt, graph_step, reps):
    """Seconds per hipGraph replay of graph_step's step (recorded afresh:
    the tile tables of the configuration under test)."""
    graph_step.capture()
    try:
        for _ in range(2):
            graph_step.replay()
        rt.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            graph_step.replay()
        rt.synchronize()
        return (time.perf_counter() - t) / reps
    finally:
        rt._check(N.hip().tnt_graph_destroy(rt._ctx, graph_step.graph))
        graph_step.graph = None


def autotune(rt, contractions, step, reps=5, log=None, graph_step=None):
    """Pick a tile configuration per contraction artifact; returns {signature: cfg}.

    graph_step: an object with capture() / replay() / graph (a GraphStep)
    whose step is `step`; when given, configurations are timed on replays."""
    table = getattr(rt, "tile_table", None)
    if table is None:
        table = rt.tile_table = load_cache()
    for c in contractions:
        sig = signature(c)
        if sig in table:
            apply(rt, c, table[sig])
            continue
        # two passes over the configurations, the faster time of each kept: the
        # first configuration timed after a change of artifact reads slow
        times = {}
        for _ in range(2):
            for cfg in (-1,) + CONFIGS:
                apply(rt, c, cfg)
                t = _time_graph(rt, graph_step, max(reps, 10)) if graph_step is not None else _time(rt, step, reps)
                times[cfg] = min(t, times.get(cfg, t))
        if log:
            log(f"[autotune] {sig} " + ", ".join(f"cfg {k}: {v * 1e3:.3f} ms" for k, v in times.items()))
        # leave the per-job default only for a clear win (timing noise is ~1-2%)
        best, best_t = -1, times[-1]
        for cfg in CONFIGS:
            if times[cfg] < times[-1] * (1.0 - MARGIN) and times[cfg] < best_t:
                best, best_t = cfg, times[cfg]
        apply(rt, c, best)
        table[sig] = best
    save_cache(table)
    return table
