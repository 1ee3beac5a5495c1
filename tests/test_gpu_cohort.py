"""Cohorts: queued non-hop queries of one k grown together over a member-interleaved level
block (the batched and served forms), then handed to the per-query jobs. Every member's
result equals the same query solved alone and the CPU oracle; trees are valid; the
finite-state count pins that every state of every member was written. The engines here
form cohorts from two queries on (the default waits for 8); `ceng_all` keeps a cohort
together through level k-1 (the merges of its members then write into the cohort's level
block, and every level but the last comes back from it)."""
import os

import numpy as np
import pytest

import pyoracle
from gpu4gst_1_amd import engine

pytestmark = pytest.mark.gpu
THREADS = min(os.cpu_count() or 1, 16)


@pytest.fixture(scope="module")
def ceng(rmat):
    eng = engine.Engine(rmat.graph, rmat.groups, device=0, cohort_min=2)
    yield eng
    eng.close()


@pytest.fixture(scope="module")
def ceng_all(rmat):
    eng = engine.Engine(rmat.graph, rmat.groups, device=0, cohort_min=2, cohort_levels=7)
    yield eng
    eng.close()


def _queries(rmat, ks, per_k, seed, D=-1):
    rng = np.random.default_rng(seed)
    return [engine.Query(rng.choice(rmat.groups.n_groups, k, replace=False).astype(np.int32), D)
            for k in ks for _ in range(per_k)]


def _check(rmat, inst, q, r, alone):
    o = pyoracle.prunedpp(inst, q.gids) if q.D < 0 else pyoracle.dprunedpp(inst, q.gids, q.D)
    if o["status"] != 0:
        assert r["status"] == alone["status"] == engine.GST_NO_SOLUTION, (q.gids, r["status"])
        return
    assert r["status"] == engine.GST_OK, (q.gids, q.D, r["status"])
    assert r["cost"] == alone["cost"] == o["tree_cost"], (q.gids, q.D, r["cost"], alone["cost"], o["tree_cost"])
    assert r["tree_cost"] == r["cost"]
    ok, why = engine.validate_tree(rmat.graph, rmat.groups, q.gids, r["edges"], r["cost"], q.D)
    assert ok, (q.gids, why)


@pytest.mark.parametrize("levels", ["1", "all"])
def test_cohort_k5(ceng, ceng_all, rmat, rmat_inst, levels):
    eng = ceng if levels == "1" else ceng_all
    qs = _queries(rmat, [5], 24, 2024)
    print("cohort of", len(qs), "k=5 queries, levels", levels, flush=True)
    res = eng.solve_batch(qs)
    for q, r in zip(qs, res):
        _check(rmat, rmat_inst, q, r, eng.solve(q.gids))


@pytest.mark.parametrize("levels", ["1", "all"])
def test_cohorts_mixed_k_and_hop(ceng, ceng_all, rmat, rmat_inst, levels):
    """One batch: cohorts of k = 2..7 beside diameter-bounded queries, which run on their
    own. The 40 k = 7 queries form one cohort when it grows level 1 only (a row of 7 labels
    per member), and two when it keeps together through level 6 (rows of up to 35 labels
    per member: 29 members at most)."""
    eng = ceng if levels == "1" else ceng_all
    qs = _queries(rmat, [2, 3, 4, 6], 5, 7) + _queries(rmat, [7], 40, 8) + _queries(rmat, [3, 5], 4, 9, D=3)
    rng = np.random.default_rng(10)
    order = rng.permutation(len(qs))
    qs = [qs[i] for i in order]
    print("mixed batch of", len(qs), "levels", levels, flush=True)
    res = eng.solve_batch(qs)
    for q, r in zip(qs, res):
        _check(rmat, rmat_inst, q, r, eng.solve(q.gids, D=q.D))


def test_cohort_k7_full_width(ceng, rmat, rmat_inst):
    """64 queries at k = 7: one cohort of 64 members whose level-1 rows take 448 labels, the
    widest level-1 row (the 8-chunk grow instantiation); every member equals the same
    query alone and the oracle."""
    qs = _queries(rmat, [7], 64, 4242)
    ceng.cohort_stats(reset=True)
    print("cohort of", len(qs), "k=7 queries", flush=True)
    res = ceng.solve_batch(qs)
    st = ceng.cohort_stats(reset=True)
    assert st["cohorts"] == 1 and st["members"] == len(qs), st
    exp = pyoracle.batch(rmat_inst, [q.gids for q in qs], algo=0, n_threads=THREADS)
    for q, r, o in zip(qs, res, exp):
        alone = ceng.solve(q.gids)
        assert r["status"] == alone["status"] == engine.GST_OK and o["status"] == 0, (q.gids, r["status"])
        assert r["cost"] == alone["cost"] == o["tree_cost"], (q.gids, r["cost"], alone["cost"], o["tree_cost"])
        assert r["tree_cost"] == r["cost"]
        ok, why = engine.validate_tree(rmat.graph, rmat.groups, q.gids, r["edges"], r["cost"], -1)
        assert ok, (q.gids, why)


def test_cohort_equals_single_jobs(rmat):
    """cohort_max = 1 turns cohorts off: the same batch through per-query jobs gives the
    same costs; without pruning, the same trees (the arenas then hold the exact DP values
    and the rebuild's choices are canonical)."""
    qs = _queries(rmat, [3, 4, 5], 6, 31)
    with engine.Engine(rmat.graph, rmat.groups, device=0, cohort_min=2) as a, \
            engine.Engine(rmat.graph, rmat.groups, device=0, cohort_max=1) as b:
        for flags in (0, engine.F_NO_PRUNE):
            qf = [engine.Query(q.gids, q.D, flags) for q in qs]
            print("cohort vs single, flags", flags, flush=True)
            ra, rb = a.solve_batch(qf), b.solve_batch(qf)
            for q, x, y in zip(qf, ra, rb):
                assert x["status"] == y["status"] == engine.GST_OK and x["cost"] == y["cost"], (q.gids, x, y)
                if flags:
                    assert np.array_equal(x["edges"], y["edges"]), q.gids


@pytest.mark.parametrize("levels", ["1", "all"])
@pytest.mark.parametrize("k", [2, 4, 6])
def test_cohort_stats_count_unpruned(ceng, ceng_all, rmat, k, levels):
    """Unpruned, on the connected R-MAT graph, every state of every member is finite:
    n (2^k - 1) per member, whichever levels come back from the cohort block."""
    eng = ceng if levels == "1" else ceng_all
    qs = [engine.Query(q.gids, -1, engine.F_STATS | engine.F_NO_PRUNE) for q in _queries(rmat, [k], 5, 50 + k)]
    res = eng.solve_batch(qs)
    for q, r in zip(qs, res):
        assert r["states_finite"] == rmat.graph.n * (2 ** k - 1), (q.gids, r["states_finite"])
        assert r["cost"] == eng.solve(q.gids)["cost"]


def test_cohort_instrumentation(rmat):
    """gst_get_cohort_stats counts the cohorts of a batch, their members and their level-1
    grow launches, device time, arcs and relaxations; a reset clears the totals. The
    members' level-1 counters hold an even share of its arc and relaxation totals."""
    qs = _queries(rmat, [4], 12, 77)
    with engine.Engine(rmat.graph, rmat.groups, device=0, cohort_min=2) as eng:
        eng.cohort_stats(reset=True)
        res = eng.solve_batch(qs)
        st = eng.cohort_stats(reset=True)
        print("cohort stats", st, flush=True)
        assert st["cohorts"] == 1 and st["members"] == len(qs)
        assert st["grow_launches"] >= 1 and st["grow_device_s"] > 0
        assert st["grow_arcs"] > 0 and st["grow_relax"] > 0
        # each member's level-1 counters include an even share of the cohort's level 1
        assert all(r["level_edges"][1] >= st["grow_arcs"] // len(qs) for r in res)
        assert all(r["level_relax"][1] >= st["grow_relax"] // len(qs) for r in res)
        assert eng.cohort_stats()["cohorts"] == 0
        eng.solve(qs[0].gids)  # the plain path runs no cohort
        assert eng.cohort_stats()["grow_launches"] == 0
