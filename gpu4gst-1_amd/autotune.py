"""Autotuner of the grow phase: degree split (thread vs wavefront per vertex), edge chunk
per wavefront item, and near-far step, chosen per (graph, k) by timing a sample of
queries on the device.

The tuning changes only the schedule of the same exact DP, so every candidate returns
the same optimum; the tuner checks that on its sample and refuses a candidate that
disagrees (it would be a bug).
"""
from __future__ import annotations

import itertools
import time

import numpy as np

DEFAULT_GRID = {
    "degree_split": (16, 32, 64, 128, 256),
    "chunk_edges": (512, 1024, 4096),
    "near_far_delta": (0, 64),
}


def tune(eng, queries, D=-1, grid=None, repeats=1, budget_s=60.0, log=None, batched=False):
    """Time every grid point on `queries` and apply the fastest: one query at a time on
    the plain path, or, with batched=True, all of them as one batch (the batched form
    favours other settings: with many jobs in flight, lanes idle behind a long thread
    item are covered by other waves, while per-item overheads are not).

    Returns (best_config dict, table of (config, seconds per query))."""
    from . import engine

    grid = grid or DEFAULT_GRID
    keys = list(grid)
    table = []
    ref = None
    t_start = time.perf_counter()
    best = None
    for vals in itertools.product(*(grid[k] for k in keys)):
        cfg = dict(zip(keys, vals))
        eng.set_tuning(cfg["degree_split"], cfg["chunk_edges"], cfg["near_far_delta"])
        if batched:
            batch = [engine.Query(np.asarray(q, np.int32), D) for q in queries]
            eng.solve_batch(batch[:2])  # warm (graph capture, slot pool)
            t = time.perf_counter()
            for _ in range(repeats):
                costs = [r["cost"] for r in eng.solve_batch(batch)]
        else:
            eng.solve(queries[0], D=D)  # warm (allocations, code paths)
            t = time.perf_counter()
            costs = []
            for _ in range(repeats):
                costs = [eng.solve(q, D=D)["cost"] for q in queries]
        dt = (time.perf_counter() - t) / (repeats * len(queries))
        if ref is None:
            ref = costs
        elif costs != ref:
            raise RuntimeError(f"tuning {cfg} changed the optimum: {costs} vs {ref}")
        table.append((cfg, dt))
        if log:
            log(f"autotune {cfg} {dt * 1e3:.2f} ms/query")
        if best is None or dt < best[1]:
            best = (cfg, dt)
        if time.perf_counter() - t_start > budget_s:
            break
    cfg = best[0]
    eng.set_tuning(cfg["degree_split"], cfg["chunk_edges"], cfg["near_far_delta"])
    return cfg, table


def degree_profile(graph):
    """Summary used to key tuned configurations: vertices, arcs, degree quantiles."""
    d = graph.degrees()
    return {"n": int(graph.n), "m": int(graph.m), "p50": float(np.median(d)), "p99": float(np.percentile(d, 99)),
            "max": int(d.max())}
