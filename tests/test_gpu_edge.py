"""Edge cases at the query boundary, on a hand-built graph whose answers are stated below:
argument checks, duplicate group ids, groups in different components, an empty group,
groups that share a vertex, and diameter bounds that cut the only tree. Plain and batched
forms must agree."""
import numpy as np
import pytest

from gpu4gst_1_amd import datasets, engine

pytestmark = pytest.mark.gpu

# path 0-1-2-3 (weights 1, 2, 3), triangle 4-5-6 (5, 5, 1), isolated vertex 7
_U, _V, _W = [0, 1, 2, 4, 5, 6], [1, 2, 3, 5, 6, 4], [1, 2, 3, 5, 5, 1]
# groups: 0 {0}, 1 {3}, 2 {1, 5}, 3 {4}, 4 {} (empty), 5 {2}, 6 {2}, 7 {7}
_MEMBERS = [[0], [3], [1, 5], [4], [], [2], [2], [7]]

# (gids, D, expected cost or None for no solution, expected edges)
CASES = [
    ([0, 1], -1, 6, {(0, 1, 1), (1, 2, 2), (2, 3, 3)}),
    ([0, 1], 3, 6, {(0, 1, 1), (1, 2, 2), (2, 3, 3)}),
    ([0, 1], 2, None, None),                 # the only tree has diameter 3
    ([0, 3], -1, None, None),                # different components
    ([0, 3], 4, None, None),
    ([0, 4], -1, None, None),                # empty group
    ([0, 7], -1, None, None),                # isolated vertex
    ([5, 6], -1, 0, set()),                  # both groups are {2}
    ([5, 6], 0, 0, set()),
    ([2], -1, 0, set()),                     # k = 1
    ([0, 0, 1, 1], -1, 6, {(0, 1, 1), (1, 2, 2), (2, 3, 3)}),  # duplicates collapse to k = 2
    ([0, 2, 5], -1, 3, {(0, 1, 1), (1, 2, 2)}),
    ([2, 3], -1, 5, {(4, 5, 5)}),            # group 2 via its member 5, in the triangle
    ([2, 3], 0, None, None),                 # D = 0: no single vertex in both
]


@pytest.fixture(scope="module")
def tiny():
    g = datasets.from_edges(8, _U, _V, _W)
    offs = np.zeros(len(_MEMBERS) + 1, np.int64)
    offs[1:] = np.cumsum([len(m) for m in _MEMBERS])
    verts = np.array([x for m in _MEMBERS for x in m], np.int32)
    gr = datasets.Groups(offs, verts)
    eng = engine.Engine(g, gr, device=0)
    yield g, gr, eng
    eng.close()


@pytest.mark.parametrize("gids, D, cost, edges", CASES)
def test_tiny_graph_answers(tiny, gids, D, cost, edges):
    g, gr, eng = tiny
    r = eng.solve(np.array(gids, np.int32), D=D)
    if cost is None:
        assert r["status"] == engine.GST_NO_SOLUTION and r["cost"] == -1, r
        return
    assert r["status"] == engine.GST_OK and r["cost"] == cost and r["tree_cost"] == cost, r
    assert r["k"] == len(set(gids))
    assert {tuple(int(x) for x in e) for e in r["edges"]} == edges
    ok, why = engine.validate_tree(g, gr, np.array(sorted(set(gids)), np.int32), r["edges"], cost, D)
    assert ok, why
    if not edges:
        assert r["root"] in set(np.concatenate([gr.members(x) for x in gids]).tolist())


def test_tiny_graph_batched_equals_plain(tiny):
    g, gr, eng = tiny
    qs = [engine.Query(np.array(gids, np.int32), D) for gids, D, _, _ in CASES]
    res = eng.solve_batch(qs)
    for (gids, D, cost, _), r in zip(CASES, res):
        if cost is None:
            assert r["status"] == engine.GST_NO_SOLUTION, (gids, D, r)
        else:
            assert r["status"] == engine.GST_OK and r["cost"] == cost, (gids, D, r)


def test_tiny_graph_cohort_equals_plain(tiny):
    """The same cases through an engine that groups queued non-hop queries of one k into a
    cohort from two on: the k = 2 cases (groups in different components, an empty group,
    shared vertices) then grow their level 1 together."""
    g, gr, _ = tiny
    qs = [engine.Query(np.array(gids, np.int32), D) for gids, D, _, _ in CASES]
    with engine.Engine(g, gr, device=0, cohort_min=2) as ceng:
        ceng.cohort_stats(reset=True)
        res = ceng.solve_batch(qs)
        assert ceng.cohort_stats()["cohorts"] >= 1
    for (gids, D, cost, _), r in zip(CASES, res):
        if cost is None:
            assert r["status"] == engine.GST_NO_SOLUTION, (gids, D, r)
        else:
            assert r["status"] == engine.GST_OK and r["cost"] == cost and r["tree_cost"] == cost, (gids, D, r)
            ok, why = engine.validate_tree(g, gr, np.array(sorted(set(gids)), np.int32), r["edges"], cost, D)
            assert ok, why


@pytest.mark.parametrize("gids, D", [([-1, 0], -1), ([0, 8], -1), ([], -1), (list(range(8)), -1),
                                     ([0, 1], 16)])
def test_tiny_graph_rejects_bad_queries(tiny, gids, D):
    g, gr, eng = tiny
    with pytest.raises(engine.GstError) as e:
        eng.solve(np.array(gids, np.int32), D=D)
    assert e.value.code == engine.GST_BAD_INPUT


@pytest.mark.parametrize("split, chunk", [(0, 1024), (1 << 25, 1024), (32, 32), (32, 1 << 25)])
def test_tuning_rejects_out_of_range(tiny, split, chunk):
    g, gr, eng = tiny
    with pytest.raises(engine.GstError) as e:
        eng.set_tuning(split, chunk)
    assert e.value.code == engine.GST_BAD_INPUT
    eng.set_tuning(32, 1024)  # still usable
    assert eng.solve(np.array([0, 1], np.int32))["cost"] == 6
