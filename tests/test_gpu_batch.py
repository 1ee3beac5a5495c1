"""Batched persistent form and serving: every result equals the single-query result
(exact integers) and the CPU oracle; trees are valid."""
import time

import numpy as np
import pytest

import pyoracle
from gpu4gst_1_amd import engine

pytestmark = pytest.mark.gpu


def _mixed_queries(rmat, n, seed):
    rng = np.random.default_rng(seed)
    qs = []
    for i in range(n):
        k = int(rng.choice([3, 5, 7]))
        D = int(rng.choice([-1, -1, 2, 3, 4]))
        qs.append(engine.Query(rng.choice(rmat.groups.n_groups, k, replace=False).astype(np.int32), D))
    return qs


def _oracle_cost(inst, q):
    o = pyoracle.prunedpp(inst, q.gids) if q.D < 0 else pyoracle.dprunedpp(inst, q.gids, q.D)
    return o["tree_cost"] if o["status"] == 0 else None


def _assert_same(rmat, inst, q, r, alone):
    oc = _oracle_cost(inst, q)
    if oc is None:
        assert r["status"] == engine.GST_NO_SOLUTION and alone["status"] == engine.GST_NO_SOLUTION
        return
    assert r["status"] == engine.GST_OK, r
    assert r["cost"] == alone["cost"] == oc, (q, r["cost"], alone["cost"], oc)
    assert r["tree_cost"] == r["cost"]
    ok, why = engine.validate_tree(rmat.graph, rmat.groups, q.gids, r["edges"], r["cost"], q.D)
    assert ok, why


def test_batch_matches_plain_and_oracle(gpu_engine, rmat, rmat_inst):
    qs = _mixed_queries(rmat, 24, 1)
    print("batch of", len(qs), flush=True)
    res = gpu_engine.solve_batch(qs)
    for q, r in zip(qs, res):
        alone = gpu_engine.solve(q.gids, D=q.D)
        _assert_same(rmat, rmat_inst, q, r, alone)


def test_batch_after_tuning_change(rmat, rmat_inst):
    """The batched rounds are captured with the degree split and chunk size of their
    time; a tuning change between two batched calls must not run stale ones."""
    eng = engine.Engine(rmat.graph, rmat.groups, device=0)
    try:
        qs = _mixed_queries(rmat, 12, 5)
        for split, chunk in [(16, 64), (48, 128), (256, 4096)]:
            eng.set_tuning(split, chunk)
            print("tuning", split, chunk, flush=True)
            res = eng.solve_batch(qs)
            for q, r in zip(qs, res):
                alone = eng.solve(q.gids, D=q.D)
                _assert_same(rmat, rmat_inst, q, r, alone)
    finally:
        eng.close()


def test_batch_small_budget_serialises(rmat, rmat_inst):
    # a budget that fits the largest arena of the set, but never all of them at once,
    # forces queuing; results unchanged
    qs = _mixed_queries(rmat, 10, 2)
    with engine.Engine(rmat.graph, rmat.groups, max_batch=8) as probe:
        charged = [probe.job_bytes(len(q.gids), q.D) for q in qs]
    assert all(c > engine.arena_bytes(rmat.graph.n, len(q.gids), q.D) for c, q in zip(charged, qs))
    budget = max(charged) + 1
    assert sum(charged) > budget
    with engine.Engine(rmat.graph, rmat.groups, dp_budget_bytes=budget, max_batch=8) as eng:
        res = eng.solve_batch(qs)
        for q, r in zip(qs, res):
            alone = eng.solve(q.gids, D=q.D)
            _assert_same(rmat, rmat_inst, q, r, alone)
        info = eng.info()
        assert info["completed"] >= 10 and info["in_use_bytes"] == 0
        assert info["budget_bytes"] == budget


def test_serving_continuous_admission(gpu_engine, rmat, rmat_inst):
    """Requests submitted while others are on the device; each result equals the
    result of the same query alone."""
    qs = _mixed_queries(rmat, 30, 3)
    tickets = []
    for i in range(0, len(qs), 5):
        tickets += gpu_engine.submit(qs[i:i + 5])
        time.sleep(0.002)
    got = {}
    deadline = time.time() + 60
    while len(got) < len(tickets) and time.time() < deadline:
        for t in tickets:
            if t not in got:
                r = gpu_engine.poll(t)
                if r is not None:
                    got[t] = r
        time.sleep(0.001)
    assert len(got) == len(tickets)
    for q, t in zip(qs, tickets):
        alone = gpu_engine.solve(q.gids, D=q.D)
        _assert_same(rmat, rmat_inst, q, got[t], alone)


@pytest.mark.parametrize("delta", [8, 64])
def test_near_far_plain_and_mixed_batch(rmat, rmat_inst, delta):
    """Near-far frontier (states above the threshold wait in a far pile) on the plain path
    and in batches that mix jobs, whose wavefronts then push for several jobs and lists
    in one step: every result equals the oracle and the label-correcting default."""
    qs = _mixed_queries(rmat, 24, 40 + delta)
    with engine.Engine(rmat.graph, rmat.groups, device=0, near_far_delta=delta) as eng:
        for split, chunk in [(32, 1024), (4, 64)]:  # (4, 64): hubs as many wavefront items
            eng.set_tuning(split, chunk, delta)
            print("near-far delta", delta, "tuning", split, chunk, flush=True)
            res = eng.solve_batch(qs)
            for q, r in zip(qs, res):
                alone = eng.solve(q.gids, D=q.D)
                _assert_same(rmat, rmat_inst, q, r, alone)


def test_progressive_bounds(rmat, rmat_inst):
    """gst_poll on queries in flight reports the best tree weight found so far (ub) and a
    lower bound (lb): ub never increases, lb never decreases, lb <= optimum <= ub at every
    poll, and the final result has ub == lb == cost (the anytime view of PrunedDP++,
    CPUNONHOP.h:684-726)."""
    rng = np.random.default_rng(77)
    qs = [engine.Query(rng.choice(rmat.groups.n_groups, 7, replace=False).astype(np.int32), D)
          for D in (-1, -1, -1, 4)]
    with engine.Engine(rmat.graph, rmat.groups, device=0, graph_rounds=1) as eng:
        tickets = eng.submit(qs)
        seen = {t: [] for t in tickets}
        final = {}
        deadline = time.time() + 60
        while len(final) < len(tickets) and time.time() < deadline:
            for t in tickets:
                if t in final:
                    continue
                r = eng.poll(t, progress=True)
                if r["status"] == engine.GST_PENDING:
                    seen[t].append((r["ub"], r["lb"], r["level"]))
                else:
                    final[t] = r
        assert len(final) == len(tickets)
        n_interim = 0
        for q, t in zip(qs, tickets):
            r = final[t]
            oc = _oracle_cost(rmat_inst, q)
            assert r["status"] == engine.GST_OK and r["cost"] == oc, (q, r["cost"], oc)
            assert r["ub"] == r["lb"] == r["cost"]
            ubs = [u for u, _, _ in seen[t] if u >= 0]
            lbs = [lo for _, lo, _ in seen[t]]
            assert all(a >= b for a, b in zip(ubs, ubs[1:])), ubs
            assert all(a <= b for a, b in zip(lbs, lbs[1:])), lbs
            assert all(u >= oc for u in ubs) and all(lo <= oc for lo in lbs), (oc, ubs, lbs)
            n_interim += sum(1 for u, lo, _ in seen[t] if u >= 0 and lo > 0)
        assert n_interim > 0  # some poll saw both bounds before the end


def test_serving_rejects_bad_and_oversized(rmat):
    budget = 1 << 16
    assert engine.arena_bytes(rmat.graph.n, 3) > budget
    with engine.Engine(rmat.graph, rmat.groups, dp_budget_bytes=budget) as eng:
        t = eng.submit([engine.Query(np.array([0, 1, 2], np.int32)),
                        engine.Query(np.array([rmat.groups.n_groups], np.int32))])
        r0 = eng.wait(t[0])
        r1 = eng.wait(t[1])
        assert r0["status"] == engine.GST_OVER_BUDGET
        assert r1["status"] == engine.GST_BAD_INPUT


def test_per_k_tuning_in_mixed_batch(rmat, rmat_inst):
    """Each k on its own grow tuning and merge tile (gst_set_tuning_k): a batch mixing k,
    where every job runs with its k's settings, equals the oracle and the plain path."""
    qs = _mixed_queries(rmat, 30, 21)
    with engine.Engine(rmat.graph, rmat.groups, device=0) as eng:
        eng.set_tuning_k(3, 4, 64, 8)        # hubs as many small wavefront items, tiny tiles
        eng.set_tuning_k(5, 256, 4096, 0)    # nearly every vertex a thread item
        eng.set_tuning_k(7, 48, 128, 3)
        print("per-k tuning batch of", len(qs), flush=True)
        res = eng.solve_batch(qs)
        for q, r in zip(qs, res):
            alone = eng.solve(q.gids, D=q.D)
            _assert_same(rmat, rmat_inst, q, r, alone)
        with pytest.raises(engine.GstError):
            eng.set_tuning_k(8, 32, 1024, 0)
