"""Parity at the synthetic scales of the benchmark configurations:
  C2  Chung-Lu 168k vertices / 6.8M edges, 16 queries at each of k = 3, 5, 7 (plain and
      one batch of all 48);
  C3  Chung-Lu 1.1M / 3M, diameter bounds D = 3, 4, 5 at k = 3, 5, 7;
  C4  Chung-Lu beta 2.0, 3.07M / 117M: 8 plain k = 5 queries, and one batch mixing
      k = 3, 5 and 7;
  C5  Chung-Lu 4.8M / 69M: a served mix with Poisson-like arrivals, k in {3, 5, 7} and
      D in {none, 3, 4, 5}; each result equals the same query solved alone.
Expected weights come from the CPU oracle: live for the cheap C3 queries, from committed
fixtures (tests/golden/scale_*.json, written by tests/golden/make_scale_golden.py) for
queries that take the oracle minutes. Every tree is validated (tree, coverage, weight,
diameter)."""
import os
import time

import numpy as np
import pytest

import pyoracle
from conftest import load_scale
from gpu4gst_1_amd import datasets, engine

pytestmark = pytest.mark.gpu
THREADS = min(os.cpu_count() or 1, 16)


def _check(wl, q, D, r, expect_status, expect_cost, what):
    if expect_status != 0:
        assert r["status"] == engine.GST_NO_SOLUTION, (what, q, D, r["status"])
        return
    assert r["status"] == engine.GST_OK and r["cost"] == expect_cost, (what, q, D, r["status"], r["cost"], expect_cost)
    assert r["tree_cost"] == r["cost"], (what, q, D)
    ok, why = engine.validate_tree(wl.graph, wl.groups, q, r["edges"], r["cost"], D)
    assert ok, (what, q, D, why)


def test_cl168k_plain_and_batched():
    wl, fx = load_scale("cl168k")
    rows = fx["samples"]["plain"]
    with engine.Engine(wl.graph, wl.groups) as eng:
        qs = [engine.Query(np.array(x["gids"], np.int32)) for x in rows]
        print("cl168k batch of", len(qs), flush=True)
        batched = eng.solve_batch(qs)
        for x, b in zip(rows, batched):
            q = np.array(x["gids"], np.int32)
            print("cl168k k", len(q), q, flush=True)
            r = eng.solve(q)
            _check(wl, q, -1, r, x["status"], x["cost"], "plain")
            _check(wl, q, -1, b, x["status"], x["cost"], "batched")


@pytest.fixture(scope="module")
def cl1m():
    wl = datasets.make_workload("cl1m")
    return wl, pyoracle.Instance(wl.graph, wl.groups)


@pytest.mark.parametrize("D", [3, 4, 5])
def test_cl1m_hop(cl1m, D):
    wl, inst = cl1m
    outs = []
    for sel in (wl.queries[3][:3], wl.queries[5][:1]):
        outs += list(zip(sel, pyoracle.batch(inst, sel, algo=1, D=D, n_threads=THREADS)))
    with engine.Engine(wl.graph, wl.groups) as eng:
        for q, o in outs:
            print("cl1m D", D, q, flush=True)
            r = eng.solve(q, D=D)
            _check(wl, q, D, r, o["status"], o["tree_cost"], "plain")


def test_cl1m_hop_k5_k7_plain_and_batched():
    """k = 5 and 7 with D = 3, 4, 5 (the k = 7 oracle queries hold gigabytes of states:
    their answers are fixtures)."""
    wl, fx = load_scale("cl1m")
    rows = fx["samples"]["hop"]
    with engine.Engine(wl.graph, wl.groups) as eng:
        qs = [engine.Query(np.array(x["gids"], np.int32), x["D"]) for x in rows]
        print("cl1m hop batch of", len(qs), flush=True)
        batched = eng.solve_batch(qs)
        for x, b in zip(rows, batched):
            q = np.array(x["gids"], np.int32)
            print("cl1m hop k", len(q), "D", x["D"], q, flush=True)
            r = eng.solve(q, D=x["D"])
            _check(wl, q, x["D"], r, x["status"], x["cost"], "plain")
            _check(wl, q, x["D"], b, x["status"], x["cost"], "batched")


@pytest.fixture(scope="module")
def cl3m():
    return load_scale("cl3m")


def test_cl3m_plain_k5(cl3m):
    wl, fx = cl3m
    with engine.Engine(wl.graph, wl.groups) as eng:
        for x in fx["samples"]["plain"]:
            q = np.array(x["gids"], np.int32)
            print("cl3m k5", q, flush=True)
            _check(wl, q, -1, eng.solve(q), x["status"], x["cost"], "plain")


def test_cl3m_batch_mixing_k(cl3m):
    """One batch of 16 queries mixing k = 3, 5 and 7 (DP tables of 98 MB to 1.57 GB)."""
    wl, fx = cl3m
    rows = fx["samples"]["mixed"]
    assert sorted({len(x["gids"]) for x in rows}) == [3, 5, 7]
    with engine.Engine(wl.graph, wl.groups) as eng:
        qs = [engine.Query(np.array(x["gids"], np.int32)) for x in rows]
        print("cl3m mixed batch of", len(qs), flush=True)
        batched = eng.solve_batch(qs)
        for x, b in zip(rows, batched):
            q = np.array(x["gids"], np.int32)
            print("cl3m mixed k", len(q), q, flush=True)
            _check(wl, q, -1, b, x["status"], x["cost"], "batched")
            assert eng.solve(q, flags=engine.F_NO_TREE)["cost"] == b["cost"]


def test_cl4m_served_mix():
    """Queries arrive while others run (exponential gaps) and are admitted as slots and
    budget free up; every result equals the same query solved alone, and the oracle's
    weight where the fixture has one (hop queries at k = 7 are too large for the oracle)."""
    wl, fx = load_scale("cl4m")
    rows = fx["samples"]["served"]
    rng = np.random.default_rng(4)
    with engine.Engine(wl.graph, wl.groups, max_batch=16, dp_budget_bytes=200 * 10 ** 9) as eng:
        qs = [engine.Query(np.array(x["gids"], np.int32), x["D"]) for x in rows]
        tickets = []
        for q in qs:
            tickets += eng.submit([q])
            time.sleep(float(rng.exponential(0.01)))
        print("cl4m served", len(tickets), "queries submitted", flush=True)
        served = [eng.wait(t) for t in tickets]
        n_oracle = 0
        for x, q, s in zip(rows, qs, served):
            print("cl4m k", len(q.gids), "D", q.D, q.gids, flush=True)
            alone = eng.solve(q.gids, D=q.D)
            assert s["status"] == alone["status"] and s["cost"] == alone["cost"], (q.gids, q.D, s["cost"], alone["cost"])
            if "cost" in x:
                n_oracle += 1
                _check(wl, q.gids, q.D, s, x["status"], x["cost"], "served")
            elif s["status"] == engine.GST_OK:
                ok, why = engine.validate_tree(wl.graph, wl.groups, q.gids, s["edges"], s["cost"], q.D)
                assert ok, (q.gids, q.D, why)
        assert n_oracle >= 16
