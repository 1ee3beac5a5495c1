"""Seeded fuzz over small random graphs: parallel edges, isolated vertices, empty and
overlapping groups, k = 1..5, D in {none, 0..4}. Every plain and batched result equals
the oracle (PrunedDP++ / D-PrunedDP++) and every tree is valid. On even seeds the batched
engine groups any two queued non-hop queries of one k into a cohort, so cohorts meet the
same edge cases."""
import numpy as np
import pytest

import pyoracle
from gpu4gst_1_amd import datasets, engine

pytestmark = pytest.mark.gpu


def _case(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 40))
    m = int(rng.integers(0, 3 * n))
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    keep = u != v
    u, v = u[keep], v[keep]
    w = rng.integers(1, 12, len(u))
    g = datasets.from_edges(n, u.astype(np.int32), v.astype(np.int32), w.astype(np.int32))
    ng = int(rng.integers(2, 9))
    members = []
    for _ in range(ng):
        sz = int(rng.integers(0, 4))  # empty groups included
        members.append(sorted(set(rng.integers(0, n, sz).tolist())))
    offs = np.zeros(ng + 1, np.int64)
    offs[1:] = np.cumsum([len(x) for x in members])
    gr = datasets.Groups(offs, np.array([x for mm in members for x in mm], np.int32))
    queries = []
    for _ in range(4):
        k = int(rng.integers(1, min(5, ng) + 1))
        q = rng.choice(ng, k, replace=False).astype(np.int32)
        D = int(rng.choice([-1, -1, 0, 1, 2, 3, 4]))
        queries.append((q, D))
    return g, gr, queries


@pytest.mark.parametrize("seed", list(range(150)))
def test_fuzz_small_graphs(seed):
    g, gr, queries = _case(seed)
    inst = pyoracle.Instance(g, gr)
    eng = engine.Engine(g, gr, device=0, cohort_min=2 if seed % 2 == 0 else 0)
    try:
        batch = eng.solve_batch([engine.Query(q, D) for q, D in queries])
        for (q, D), rb in zip(queries, batch):
            print("seed", seed, "query", q.tolist(), "D", D, flush=True)
            r = eng.solve(q, D=D)
            o = pyoracle.prunedpp(inst, q) if D < 0 else pyoracle.dprunedpp(inst, q, D)
            if o["status"] != 0:
                assert r["status"] == engine.GST_NO_SOLUTION and rb["status"] == engine.GST_NO_SOLUTION, (r, rb, o)
                continue
            assert r["status"] == engine.GST_OK and r["cost"] == o["tree_cost"] == rb["cost"], (r["cost"], rb["cost"], o)
            assert r["tree_cost"] == r["cost"]
            ok, why = engine.validate_tree(g, gr, q, r["edges"], r["cost"], D)
            assert ok, why
    finally:
        eng.close()
