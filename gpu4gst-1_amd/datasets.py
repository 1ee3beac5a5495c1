"""Seeded synthetic GST workloads and the reference's on-disk formats.

The graph topology comes from libgstdata.so (C++: R-MAT / Chung-Lu, dedup, largest
connected component, CSR); groups and queries are drawn here with numpy's PCG64.

Workloads (seed = 20261203 + index):
  rmat32k   R-MAT scale 15, edge factor 8, a=.57 b=c=.19, 1,000 groups of 2-10 vertices,
            50 queries of k=3
  cl168k    Chung-Lu beta=2.1, 168k vertices / 6.8M edges, 5,000 Zipf(1.5) groups in [1,50],
            300 queries at each of k=3,5,7
  cl1m      Chung-Lu beta=2.1, 1.1M / 3M, 10,000 groups, 300 queries at k=3,5,7 (hop D=3,4,5)
  cl3m      Chung-Lu beta=2.0, 3.07M / 117M, 20,000 groups, batches mixing k=3,5,7
  cl4m      Chung-Lu beta=2.1, 4.8M / 69M, 50,000 groups, Poisson arrivals
Every graph is simple, undirected, restricted to its largest connected component and
relabelled in original-id order; edge weights are int32 uniform in [1,100].

File formats follow the reference readers
(code/PrunedDP++/include/graph_hash_of_mixed_weighted_read_for_GSTP.h:15-218,
code/D-TrimCDP-WB/include/graph.hpp:7-118): .in, _beg_pos/_csr/_weight.bin (int64),
.g ("g<n>:" 1-based), <name><k>.csv (0-based group ids).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

SEED_BASE = 20261203


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libgstdata.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(path)
        P = ctypes.c_void_p
        i64, i32, u64, d = ctypes.c_int64, ctypes.c_int32, ctypes.c_uint64, ctypes.c_double
        lib.gstd_gen_rmat.restype = P
        lib.gstd_gen_rmat.argtypes = [ctypes.c_int, ctypes.c_int, d, d, d, ctypes.c_int, ctypes.c_int, u64]
        lib.gstd_gen_chunglu.restype = P
        lib.gstd_gen_chunglu.argtypes = [i64, i64, d, ctypes.c_int, ctypes.c_int, u64]
        lib.gstd_from_edges.restype = P
        lib.gstd_from_edges.argtypes = [i64, i64, P, P, P]
        for name in ("gstd_num_vertices", "gstd_num_arcs"):
            getattr(lib, name).restype = i64
            getattr(lib, name).argtypes = [P]
        for name in ("gstd_beg_pos", "gstd_adj", "gstd_weight"):
            getattr(lib, name).restype = P
            getattr(lib, name).argtypes = [P]
        lib.gstd_free.argtypes = [P]
        lib.gstd_free_buf.argtypes = [P]
        lib.gstd_write_in.argtypes = [P, ctypes.c_char_p]
        lib.gstd_write_bin.argtypes = [P, ctypes.c_char_p]
        lib.gstd_read_in.restype = P
        lib.gstd_read_in.argtypes = [ctypes.c_char_p]
        lib.gstd_read_bin.restype = P
        lib.gstd_read_bin.argtypes = [ctypes.c_char_p]
        lib.gstd_write_groups.argtypes = [ctypes.c_char_p, i32, P, P]
        lib.gstd_read_groups.restype = i32
        lib.gstd_read_groups.argtypes = [ctypes.c_char_p, ctypes.POINTER(P), ctypes.POINTER(P)]
        lib.gstd_write_queries.argtypes = [ctypes.c_char_p, i32, i32, P]
        lib.gstd_read_queries.restype = i32
        lib.gstd_read_queries.argtypes = [ctypes.c_char_p, ctypes.POINTER(i32), ctypes.POINTER(P)]
        lib.gstd_last_error.restype = ctypes.c_char_p
        _LIB = lib
    return _LIB


def _err(msg):
    raise RuntimeError(f"{msg}: {_lib().gstd_last_error().decode()}")


@dataclass
class Graph:
    """Symmetric CSR: beg_pos[n+1] (int64), adj[m], w[m] (int32), m = 2 x undirected edges."""

    beg_pos: np.ndarray
    adj: np.ndarray
    w: np.ndarray

    @property
    def n(self) -> int:
        return len(self.beg_pos) - 1

    @property
    def m(self) -> int:
        return len(self.adj)

    @property
    def n_edges(self) -> int:
        return self.m // 2

    def degrees(self) -> np.ndarray:
        return np.diff(self.beg_pos)


def _take(h) -> Graph:
    lib = _lib()
    if not h:
        _err("graph construction failed")
    n, m = lib.gstd_num_vertices(h), lib.gstd_num_arcs(h)
    beg = np.ctypeslib.as_array(ctypes.cast(lib.gstd_beg_pos(h), ctypes.POINTER(ctypes.c_int64)), (n + 1,)).copy()
    if m:
        adj = np.ctypeslib.as_array(ctypes.cast(lib.gstd_adj(h), ctypes.POINTER(ctypes.c_int32)), (m,)).copy()
        w = np.ctypeslib.as_array(ctypes.cast(lib.gstd_weight(h), ctypes.POINTER(ctypes.c_int32)), (m,)).copy()
    else:
        adj = np.zeros(0, np.int32)
        w = np.zeros(0, np.int32)
    lib.gstd_free(h)
    return Graph(beg, adj, w)


def gen_rmat(scale, edge_factor, a=0.57, b=0.19, c=0.19, seed=0, wlo=1, whi=100) -> Graph:
    return _take(_lib().gstd_gen_rmat(scale, edge_factor, a, b, c, wlo, whi, seed))


def gen_chunglu(n, m, beta, seed=0, wlo=1, whi=100) -> Graph:
    return _take(_lib().gstd_gen_chunglu(n, m, beta, wlo, whi, seed))


def from_edges(n, u, v, w) -> Graph:
    """Reference .in reader semantics: both arcs inserted, adjacency sorted, no dedup."""
    u = np.ascontiguousarray(u, np.int32)
    v = np.ascontiguousarray(v, np.int32)
    w = np.ascontiguousarray(w, np.int32)
    return _take(_lib().gstd_from_edges(n, len(u), u.ctypes.data, v.ctypes.data, w.ctypes.data))


@dataclass
class Groups:
    """Group table: group g (0-based) = verts[offs[g]:offs[g+1]], members ascending."""

    offs: np.ndarray
    verts: np.ndarray

    @property
    def n_groups(self) -> int:
        return len(self.offs) - 1

    def members(self, g: int) -> np.ndarray:
        return self.verts[self.offs[g]:self.offs[g + 1]]


def gen_groups_uniform(n_vertices, n_groups, lo, hi, seed) -> Groups:
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = rng.integers(lo, hi + 1, size=n_groups)
    return _fill_groups(rng, n_vertices, sizes)


def gen_groups_zipf(n_vertices, n_groups, s, lo, hi, seed) -> Groups:
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = np.clip(rng.zipf(s, size=n_groups), lo, hi)
    return _fill_groups(rng, n_vertices, sizes)


def _fill_groups(rng, n_vertices, sizes) -> Groups:
    sizes = np.minimum(np.asarray(sizes, np.int64), n_vertices)
    offs = np.zeros(len(sizes) + 1, np.int64)
    np.cumsum(sizes, out=offs[1:])
    verts = np.empty(int(offs[-1]), np.int32)
    for g, sz in enumerate(sizes):
        verts[offs[g]:offs[g + 1]] = np.sort(rng.choice(n_vertices, size=int(sz), replace=False))
    return Groups(offs, verts)


def gen_queries(n_groups, n_queries, k, seed) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.stack([rng.choice(n_groups, size=k, replace=False) for _ in range(n_queries)]).astype(np.int32) \
        if n_queries else np.zeros((0, k), np.int32)


# ---------------------------------------------------------------- workloads
@dataclass
class Workload:
    name: str
    graph: Graph
    groups: Groups
    queries: dict = field(default_factory=dict)  # k -> int32[n_queries, k]


WORKLOADS = {
    # name: (generator, graph args, groups spec, queries {k: count}, seed index)
    "rmat32k": ("rmat", dict(scale=15, edge_factor=8), ("uniform", 1000, 2, 10), {3: 50}, 0),
    "cl168k": ("chunglu", dict(n=168_000, m=6_800_000, beta=2.1), ("zipf", 5000, 1, 50), {3: 300, 5: 300, 7: 300}, 1),
    "cl1m": ("chunglu", dict(n=1_100_000, m=3_000_000, beta=2.1), ("zipf", 10000, 1, 50), {3: 300, 5: 300, 7: 300}, 2),
    "cl3m": ("chunglu", dict(n=3_070_000, m=117_000_000, beta=2.0), ("zipf", 20000, 1, 50), {3: 300, 5: 300, 7: 300}, 3),
    "cl4m": ("chunglu", dict(n=4_800_000, m=69_000_000, beta=2.1), ("zipf", 50000, 1, 50), {3: 300, 5: 300, 7: 300}, 4),
}


def make_workload(name: str, scale_down: int = 1) -> Workload:
    """Build a named workload. scale_down > 1 shrinks vertex and edge counts (tests only)."""
    gen, gargs, gspec, qspec, idx = WORKLOADS[name]
    seed = SEED_BASE + idx
    if gen == "rmat":
        sc = gargs["scale"] - int(np.log2(scale_down)) if scale_down > 1 else gargs["scale"]
        graph = gen_rmat(sc, gargs["edge_factor"], seed=seed)
    else:
        graph = gen_chunglu(gargs["n"] // scale_down, gargs["m"] // scale_down, gargs["beta"], seed=seed)
    kind, ng, lo, hi = gspec
    ng = max(8, ng // scale_down)
    if kind == "uniform":
        groups = gen_groups_uniform(graph.n, ng, lo, hi, seed + 1000)
    else:
        groups = gen_groups_zipf(graph.n, ng, 1.5, lo, hi, seed + 1000)
    queries = {k: gen_queries(ng, cnt, k, seed + 2000 + k) for k, cnt in qspec.items()}
    return Workload(name, graph, groups, queries)


# ---------------------------------------------------------------- file formats
def write_workload(wl: Workload, directory: str, binary: bool = True) -> None:
    os.makedirs(directory, exist_ok=True)
    write_graph_in(wl.graph, os.path.join(directory, wl.name + ".in"))
    if binary:
        write_graph_bin(wl.graph, os.path.join(directory, wl.name))
    write_groups(wl.groups, os.path.join(directory, wl.name + ".g"))
    for k, q in wl.queries.items():
        write_queries(q, os.path.join(directory, f"{wl.name}{k}.csv"))


def _graph_handle(g: Graph):
    """Wrap a Graph as a temporary native handle (edge list u<v)."""
    src = np.repeat(np.arange(g.n, dtype=np.int32), np.diff(g.beg_pos))
    keep = src < g.adj
    return _lib().gstd_from_edges(g.n, int(keep.sum()), np.ascontiguousarray(src[keep]).ctypes.data,
                                  np.ascontiguousarray(g.adj[keep]).ctypes.data,
                                  np.ascontiguousarray(g.w[keep]).ctypes.data)


def write_graph_in(g: Graph, path: str) -> None:
    src = np.repeat(np.arange(g.n, dtype=np.int64), np.diff(g.beg_pos))
    keep = src < g.adj
    with open(path, "w") as f:
        f.write(f"{g.n} {int(keep.sum())}\n")
        np.savetxt(f, np.stack([src[keep], g.adj[keep].astype(np.int64), g.w[keep].astype(np.int64)], 1), fmt="%d")


def write_graph_bin(g: Graph, prefix: str) -> None:
    g.beg_pos.astype(np.int64).tofile(prefix + "_beg_pos.bin")
    g.adj.astype(np.int64).tofile(prefix + "_csr.bin")
    g.w.astype(np.int64).tofile(prefix + "_weight.bin")


def read_graph_in(path: str) -> Graph:
    return _take(_lib().gstd_read_in(path.encode()))


def read_graph_bin(prefix: str) -> Graph:
    return _take(_lib().gstd_read_bin(prefix.encode()))


def write_groups(gr: Groups, path: str) -> None:
    offs = np.ascontiguousarray(gr.offs, np.int64)
    verts = np.ascontiguousarray(gr.verts, np.int32)
    if _lib().gstd_write_groups(path.encode(), gr.n_groups, offs.ctypes.data, verts.ctypes.data):
        _err("write_groups")


def read_groups(path: str) -> Groups:
    lib = _lib()
    po, pv = ctypes.c_void_p(), ctypes.c_void_p()
    n = lib.gstd_read_groups(path.encode(), ctypes.byref(po), ctypes.byref(pv))
    if n < 0:
        _err("read_groups")
    offs = np.ctypeslib.as_array(ctypes.cast(po, ctypes.POINTER(ctypes.c_int64)), (n + 1,)).copy()
    total = int(offs[-1])
    verts = np.ctypeslib.as_array(ctypes.cast(pv, ctypes.POINTER(ctypes.c_int32)), (max(total, 1),))[:total].copy()
    lib.gstd_free_buf(po)
    lib.gstd_free_buf(pv)
    return Groups(offs, verts)


def write_queries(q: np.ndarray, path: str) -> None:
    q = np.ascontiguousarray(q, np.int32)
    if _lib().gstd_write_queries(path.encode(), q.shape[0], q.shape[1] if q.ndim == 2 else 0, q.ctypes.data):
        _err("write_queries")


def read_queries(path: str) -> np.ndarray:
    lib = _lib()
    k = ctypes.c_int32()
    pg = ctypes.c_void_p()
    n = lib.gstd_read_queries(path.encode(), ctypes.byref(k), ctypes.byref(pg))
    if n < 0:
        _err("read_queries")
    arr = np.ctypeslib.as_array(ctypes.cast(pg, ctypes.POINTER(ctypes.c_int32)), (max(n * k.value, 1),))[:n * k.value].copy()
    lib.gstd_free_buf(pg)
    return arr.reshape(n, k.value)
