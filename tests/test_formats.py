"""Generators and the reference's on-disk formats (CPU only)."""
import os

import numpy as np
import pytest

from gpu4gst_1_amd import datasets as ds


def _check_csr(g):
    assert g.beg_pos[0] == 0 and g.beg_pos[-1] == g.m
    assert np.all(np.diff(g.beg_pos) >= 0)
    src = np.repeat(np.arange(g.n), np.diff(g.beg_pos))
    assert np.all(g.adj != src), "self loop"
    # symmetric: sorted (u,v,w) equals sorted (v,u,w)
    a = np.stack([src, g.adj, g.w], 1)
    b = np.stack([g.adj, src, g.w], 1)
    assert np.array_equal(a[np.lexsort(a.T[::-1])], b[np.lexsort(b.T[::-1])])
    for v in range(0, g.n, max(1, g.n // 200)):
        nb = g.adj[g.beg_pos[v]:g.beg_pos[v + 1]]
        assert np.all(np.diff(nb) > 0), "adjacency not strictly ascending (duplicate edge)"


def _connected(g):
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import connected_components
    m = csr_matrix((np.ones(g.m), g.adj, g.beg_pos), shape=(g.n, g.n))
    return connected_components(m, directed=False)[0] == 1


def test_rmat_shape_and_determinism(rmat):
    g = rmat.graph
    _check_csr(g)
    assert _connected(g)
    assert 1 <= g.w.min() and g.w.max() <= 100
    assert 15000 < g.n <= 32768
    g2 = ds.gen_rmat(15, 8, seed=ds.SEED_BASE)
    assert np.array_equal(g.adj, g2.adj) and np.array_equal(g.w, g2.w)


def test_chunglu_small():
    g = ds.gen_chunglu(20000, 100000, 2.1, seed=7)
    _check_csr(g)
    assert _connected(g)
    assert g.n_edges <= 100000
    d = g.degrees()
    assert d.max() > 20 * np.median(d), "power-law skew expected"
    g2 = ds.gen_chunglu(20000, 100000, 2.1, seed=7)
    assert np.array_equal(g.beg_pos, g2.beg_pos) and np.array_equal(g.adj, g2.adj)


def test_groups_and_queries():
    gr = ds.gen_groups_zipf(5000, 300, 1.5, 1, 50, seed=3)
    sizes = np.diff(gr.offs)
    assert sizes.min() >= 1 and sizes.max() <= 50
    for gi in range(gr.n_groups):
        m = gr.members(gi)
        assert np.all(np.diff(m) > 0) and m.min() >= 0 and m.max() < 5000
    q = ds.gen_queries(300, 40, 5, seed=1)
    assert q.shape == (40, 5)
    for row in q:
        assert len(set(row.tolist())) == 5


def test_round_trip(tmp_path, rmat):
    ds.write_workload(rmat, str(tmp_path))
    g = ds.read_graph_in(str(tmp_path / "rmat32k.in"))
    assert np.array_equal(g.beg_pos, rmat.graph.beg_pos) and np.array_equal(g.adj, rmat.graph.adj)
    assert np.array_equal(g.w, rmat.graph.w)
    gb = ds.read_graph_bin(str(tmp_path / "rmat32k"))
    assert np.array_equal(gb.adj, rmat.graph.adj) and np.array_equal(gb.beg_pos, rmat.graph.beg_pos)
    # int64 binary files with V+1 beg_pos entries (graph.hpp:25)
    assert os.path.getsize(tmp_path / "rmat32k_beg_pos.bin") == 8 * (rmat.graph.n + 1)
    gr = ds.read_groups(str(tmp_path / "rmat32k.g"))
    assert np.array_equal(gr.offs, rmat.groups.offs) and np.array_equal(gr.verts, rmat.groups.verts)
    q = ds.read_queries(str(tmp_path / "rmat32k3.csv"))
    assert np.array_equal(q, rmat.queries[3])


def test_reader_semantics(tmp_path):
    # .in: duplicates are kept (no dedup), both arcs inserted, adjacency sorted
    (tmp_path / "a.in").write_text("3 3\n0 1 5\n1 0 7\n1 2 1\nnot an edge\n")
    g = ds.read_graph_in(str(tmp_path / "a.in"))
    assert g.n == 3 and g.m == 6
    assert list(g.adj[g.beg_pos[1]:g.beg_pos[2]]) == [0, 0, 2]
    # .g: 1-based "g<n>:" lines stored at n-1, trailing space tolerated, members sorted
    (tmp_path / "a.g").write_text("g2:2 0 \ng1:1\n")
    gr = ds.read_groups(str(tmp_path / "a.g"))
    assert gr.n_groups == 2 and list(gr.members(0)) == [1] and list(gr.members(1)) == [0, 2]
    (tmp_path / "bad.g").write_text("x1:1 2\n")
    with pytest.raises(RuntimeError):
        ds.read_groups(str(tmp_path / "bad.g"))
    (tmp_path / "q3.csv").write_text("1 2 3 \n4 5 6\n")
    assert ds.read_queries(str(tmp_path / "q3.csv")).tolist() == [[1, 2, 3], [4, 5, 6]]
    (tmp_path / "bad.csv").write_text("1 2 3\n4 5\n")
    with pytest.raises(RuntimeError):
        ds.read_queries(str(tmp_path / "bad.csv"))
