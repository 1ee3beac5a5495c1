"""GPU parity: the HIP engine against the CPU oracles, bit-exact (integer costs)."""
import numpy as np
import pytest

import pyoracle
from gpu4gst_1_amd import engine

pytestmark = pytest.mark.gpu


def _check(eng, graph, groups, inst, q, D=-1, flags=0):
    r = eng.solve(q, D=D, flags=flags)
    o = pyoracle.prunedpp(inst, q) if D < 0 else pyoracle.dprunedpp(inst, q, D)
    assert r["status"] == engine.GST_OK, r
    assert r["cost"] == o["tree_cost"] == o["dp_cost"], (q, D, r["cost"], o["tree_cost"], o["dp_cost"])
    assert r["tree_cost"] == r["cost"], (q, D, r["tree_cost"], r["cost"])
    ok, why = engine.validate_tree(graph, groups, q, r["edges"], r["cost"], D)
    assert ok, (q, D, why)
    return r


def test_rmat_k3_all_queries(gpu_engine, rmat, rmat_inst):
    for i, q in enumerate(rmat.queries[3]):
        print("k3 query", i, q, flush=True)
        _check(gpu_engine, rmat.graph, rmat.groups, rmat_inst, q)


@pytest.mark.parametrize("k", [1, 2, 4, 5, 6, 7])
def test_rmat_random_k(gpu_engine, rmat, rmat_inst, k):
    rng = np.random.default_rng(100 + k)
    for i in range(6 if k < 7 else 3):
        q = rng.choice(rmat.groups.n_groups, k, replace=False).astype(np.int32)
        print("k", k, "query", i, q, flush=True)
        _check(gpu_engine, rmat.graph, rmat.groups, rmat_inst, q)


def test_rmat_unpruned_matches_pruned(gpu_engine, rmat, rmat_inst):
    for q in rmat.queries[3][:5]:
        a = gpu_engine.solve(q)
        b = gpu_engine.solve(q, flags=engine.F_NO_PRUNE)
        assert a["cost"] == b["cost"], (q, a["cost"], b["cost"])


@pytest.mark.parametrize("D", [0, 1, 2, 3, 4, 5])
def test_rmat_hop(gpu_engine, rmat, rmat_inst, D):
    for i, q in enumerate(rmat.queries[3][:8]):
        print("hop D", D, "query", i, q, flush=True)
        r = gpu_engine.solve(q, D=D)
        o = pyoracle.dprunedpp(rmat_inst, q, D)
        if o["status"] != 0:
            assert r["status"] == engine.GST_NO_SOLUTION, (q, D, r)
            continue
        assert r["status"] == engine.GST_OK and r["cost"] == o["tree_cost"] == o["dp_cost"], (q, D, r["cost"], o)
        ok, why = engine.validate_tree(rmat.graph, rmat.groups, q, r["edges"], r["cost"], D)
        assert ok, (q, D, why)


def test_rmat_hop_k5(gpu_engine, rmat, rmat_inst):
    rng = np.random.default_rng(7)
    for i in range(4):
        q = rng.choice(rmat.groups.n_groups, 5, replace=False).astype(np.int32)
        for D in (3, 4):
            print("hop k5 D", D, q, flush=True)
            r = gpu_engine.solve(q, D=D)
            o = pyoracle.dprunedpp(rmat_inst, q, D)
            if o["status"] != 0:
                assert r["status"] == engine.GST_NO_SOLUTION and o["tree_cost"] == pyoracle.INF, (q, D, r, o)
                assert gpu_engine.DP_gpu(q, D)[0] == 100000
                continue
            assert r["cost"] == o["tree_cost"], (q, D, r["cost"], o["tree_cost"])
            ok, why = engine.validate_tree(rmat.graph, rmat.groups, q, r["edges"], r["cost"], D)
            assert ok, (q, D, why)


def test_caller_order_matches_renumbered(gpu_engine, rmat, rmat_inst):
    """gst_open renumbers vertices by degree unless GST_O_KEEP_ORDER is set; both engines
    return the same optimum, and every tree, root included, is in the caller's ids."""
    keep = engine.Engine(rmat.graph, rmat.groups, device=0, flags=engine.O_KEEP_ORDER)
    try:
        rng = np.random.default_rng(11)
        for i in range(6):
            k = int(rng.integers(2, 6))
            D = int(rng.choice([-1, 2, 4]))
            q = rng.choice(rmat.groups.n_groups, k, replace=False).astype(np.int32)
            print("order query", i, k, D, q, flush=True)
            a = gpu_engine.solve(q, D=D)
            b = keep.solve(q, D=D)
            assert a["status"] == b["status"] and a["cost"] == b["cost"], (q, D, a["cost"], b["cost"])
            if a["status"] != engine.GST_OK:
                continue
            for r in (a, b):
                ok, why = engine.validate_tree(rmat.graph, rmat.groups, q, r["edges"], r["cost"], D)
                assert ok, (q, D, why)
                if len(r["edges"]):
                    assert r["root"] in set(r["edges"][:, :2].ravel().tolist()), (q, D, r["root"])
    finally:
        keep.close()


def test_hub_index_matches_no_index(gpu_engine, rmat, rmat_inst):
    """Queries start from the best hub star as their upper bound (the default) or from
    INF (GST_O_NO_HUB_INDEX): the same optimum either way, equal to PrunedDP++, and the
    hub star really bounds it (the result's ub never exceeds the first one polled)."""
    plain = engine.Engine(rmat.graph, rmat.groups, device=0, flags=engine.O_NO_HUB_INDEX)
    try:
        rng = np.random.default_rng(23)
        qs = [rng.choice(rmat.groups.n_groups, int(rng.integers(2, 7)), replace=False).astype(np.int32)
              for _ in range(10)]
        for q in qs:
            print("hub query", q, flush=True)
            o = pyoracle.prunedpp(rmat_inst, q)
            a = gpu_engine.solve(q)
            b = plain.solve(q)
            assert a["status"] == b["status"] == o["status"], (q, a["status"], b["status"], o["status"])
            if a["status"] != engine.GST_OK:
                continue
            assert a["cost"] == b["cost"] == o["tree_cost"], (q, a["cost"], b["cost"], o["tree_cost"])
            ok, why = engine.validate_tree(rmat.graph, rmat.groups, q, a["edges"], a["cost"])
            assert ok, (q, why)
    finally:
        plain.close()


@pytest.mark.parametrize("D", [1, 2, 3, 5])
def test_rmat_hop_pruned_matches_unpruned(gpu_engine, rmat, D):
    """The diameter-bounded search prunes by a lower bound against an upper bound; with
    GST_F_NO_PRUNE it evaluates every state. Both give the same optimum."""
    rng = np.random.default_rng(300 + D)
    for i in range(5):
        k = int(rng.integers(2, 6))
        q = rng.choice(rmat.groups.n_groups, k, replace=False).astype(np.int32)
        print("hop prune D", D, "query", i, q, flush=True)
        a = gpu_engine.solve(q, D=D)
        b = gpu_engine.solve(q, D=D, flags=engine.F_NO_PRUNE)
        assert a["status"] == b["status"] and a["cost"] == b["cost"], (q, D, a["cost"], b["cost"])
        if a["status"] == engine.GST_OK:
            ok, why = engine.validate_tree(rmat.graph, rmat.groups, q, a["edges"], a["cost"], D)
            assert ok, (q, D, why)


@pytest.mark.parametrize("D", [8, 15])
def test_rmat_hop_large_D(gpu_engine, rmat, rmat_inst, D):
    """Diameter bounds up to GST_MAX_D (16 heights per state: the largest hop tiles)."""
    rng = np.random.default_rng(900 + D)
    for i in range(3):
        k = int(rng.integers(3, 5))
        q = rng.choice(rmat.groups.n_groups, k, replace=False).astype(np.int32)
        print("large D", D, "query", i, q, flush=True)
        r = gpu_engine.solve(q, D=D)
        o = pyoracle.dprunedpp(rmat_inst, q, D)
        assert r["status"] == engine.GST_OK and r["cost"] == o["tree_cost"] == o["dp_cost"], (q, D, r["cost"], o)
        ok, why = engine.validate_tree(rmat.graph, rmat.groups, q, r["edges"], r["cost"], D)
        assert ok, (q, D, why)

@pytest.mark.parametrize("k", [4, 5, 6, 7])
def test_rmat_hop_full_tiles(gpu_engine, rmat, rmat_inst, k):
    """D = 0..2 give the hop merge its widest tiles (244 vertices at k = 4, D = 0, where one
    staged level segment is 1464 words; 81 at k = 4, D = 2): every level segment must reach
    LDS. The queries take groups of distinct neighbours of one vertex, so a tree of
    diameter 2 exists and D = 2 has a finite optimum."""
    g, gr = rmat.graph, rmat.groups
    groups_of = {}
    for gi in range(gr.n_groups):
        for v in gr.members(gi):
            groups_of.setdefault(int(v), []).append(gi)
    rng = np.random.default_rng(1300 + k)
    n_finite = 0
    for i in range(2):
        while True:
            v = int(rng.integers(g.n))
            nb = g.adj[g.beg_pos[v]:g.beg_pos[v + 1]]
            cand = sorted({x for u in nb for x in groups_of.get(int(u), [])})
            if len(cand) >= k:
                q = rng.choice(cand, k, replace=False).astype(np.int32)
                break
        for D in (0, 1, 2):
            print("full tiles k", k, "D", D, q, flush=True)
            r = gpu_engine.solve(q, D=D)
            o = pyoracle.dprunedpp(rmat_inst, q, D)
            if o["status"] != 0:
                assert r["status"] == engine.GST_NO_SOLUTION, (q, D, r)
                continue
            n_finite += 1
            assert r["status"] == engine.GST_OK and r["cost"] == o["tree_cost"] == o["dp_cost"], (q, D, r["cost"], o)
            ok, why = engine.validate_tree(g, gr, q, r["edges"], r["cost"], D)
            assert ok, (q, D, why)
    assert n_finite >= 2

@pytest.mark.parametrize("k", [2, 3, 4])
def test_stats_count_unpruned(gpu_engine, rmat, k):
    """GST_F_STATS without pruning: on a connected graph every state (v, p) is finite, so
    the count is n (2^k - 1) exactly, for the plain and the batched path. Every state of
    levels >= 2 must have been written by its level's merge (only level 1 is filled at
    query start)."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import connected_components
    g = rmat.graph
    a = csr_matrix((np.ones(g.m), g.adj, g.beg_pos), shape=(g.n, g.n))
    assert connected_components(a, directed=False)[0] == 1
    rng = np.random.default_rng(1500 + k)
    q = rng.choice(rmat.groups.n_groups, k, replace=False).astype(np.int32)
    flags = engine.F_STATS | engine.F_NO_PRUNE
    r = gpu_engine.solve(q, flags=flags)
    assert r["status"] == engine.GST_OK and r["states_finite"] == g.n * (2 ** k - 1), (q, r["states_finite"])
    b = gpu_engine.solve_batch([engine.Query(q, -1, flags)])[0]
    assert b["cost"] == r["cost"] and b["states_finite"] == g.n * (2 ** k - 1), (q, b["states_finite"])
