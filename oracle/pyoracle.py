"""ctypes front end of liboracle.so and pure-Python brute force. TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module,
as the parity checker or the reported CPU baseline. The product never does.

  prunedpp / dprunedpp   restatements of PrunedDP++ / D-PrunedDP++ (oracle/prunedpp.cpp)
  dpbf / dpbf_hop        independent exact DPs (oracle/dpbf.c)
  group_sssp             multi-source Dijkstra per query group
  brute_force            min over vertex subsets S covering all groups with G[S] connected
                         of MST(G[S]) (tiny graphs)
  brute_force_hop        min over edge subsets forming a tree that covers all groups with
                         unweighted diameter <= D (tiny graphs)
  batch                  one query per worker thread (CPU baseline)
"""
from __future__ import annotations

import ctypes
import itertools
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
INF = 100000


class OrcGraph(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("beg", ctypes.c_void_p), ("adj", ctypes.c_void_p),
                ("w", ctypes.c_void_p), ("n_groups", ctypes.c_int32), ("goffs", ctypes.c_void_p),
                ("gverts", ctypes.c_void_p)]


class OrcResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("dp_cost", ctypes.c_int32), ("tree_cost", ctypes.c_int32),
                ("n_edges", ctypes.c_int32), ("root", ctypes.c_int32), ("root_d", ctypes.c_int32),
                ("counts", ctypes.c_int64), ("process", ctypes.c_int64),
                ("t_search_s", ctypes.c_double), ("t_total_s", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (make -C oracle)")
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        i32 = ctypes.c_int32
        R = ctypes.POINTER(OrcResult)
        G = ctypes.POINTER(OrcGraph)
        L.orc_prunedpp.argtypes = [G, P, i32, P, i32, R]
        L.orc_dprunedpp.argtypes = [G, P, i32, i32, P, i32, R]
        L.orc_dpbf.argtypes = [G, P, i32, P, i32, R]
        L.orc_dpbf_hop.argtypes = [G, P, i32, i32, R]
        L.orc_group_sssp.argtypes = [G, P, i32, P]
        L.orc_batch.argtypes = [G, P, i32, i32, i32, i32, i32, ctypes.c_double, R]
        _LIB = L
    return _LIB


class Instance:
    """Keeps the numpy arrays alive behind an OrcGraph."""

    def __init__(self, graph, groups):
        self.beg = np.ascontiguousarray(graph.beg_pos, np.int64)
        self.adj = np.ascontiguousarray(graph.adj, np.int32)
        self.w = np.ascontiguousarray(graph.w, np.int32)
        self.goffs = np.ascontiguousarray(groups.offs, np.int64)
        self.gverts = np.ascontiguousarray(groups.verts, np.int32)
        self.n = len(self.beg) - 1
        self.n_groups = len(self.goffs) - 1
        self.g = OrcGraph(self.n, self.beg.ctypes.data, self.adj.ctypes.data, self.w.ctypes.data,
                          self.n_groups, self.goffs.ctypes.data, self.gverts.ctypes.data)


def _run(fn, inst, gids, *extra, edges=True):
    gids = np.ascontiguousarray(gids, np.int32)
    res = OrcResult()
    cap = 1 << 16
    buf = np.zeros(3 * cap, np.int32) if edges else None
    args = [ctypes.byref(inst.g), gids.ctypes.data, len(gids), *extra]
    if edges:
        args += [buf.ctypes.data, cap]
    rc = fn(*args, ctypes.byref(res))
    out = res.as_dict()
    out["rc"] = rc
    if edges:
        out["edges"] = buf[:3 * min(res.n_edges, cap)].reshape(-1, 3).copy()
    return out


def prunedpp(inst, gids):
    return _run(lib().orc_prunedpp, inst, gids)


def dprunedpp(inst, gids, D):
    return _run(lib().orc_dprunedpp, inst, gids, D)


def dpbf(inst, gids):
    return _run(lib().orc_dpbf, inst, gids)


def dpbf_hop(inst, gids, D):
    return _run(lib().orc_dpbf_hop, inst, gids, D, edges=False)


def group_sssp(inst, gids):
    gids = np.ascontiguousarray(gids, np.int32)
    out = np.zeros((len(gids), inst.n), np.int32)
    rc = lib().orc_group_sssp(ctypes.byref(inst.g), gids.ctypes.data, len(gids), out.ctypes.data)
    if rc:
        raise ValueError("bad group id")
    return out


def batch(inst, queries, algo=0, D=0, n_threads=None, time_budget_s=0.0):
    """algo 0 = PrunedDP++, 1 = D-PrunedDP++, 2 = DPBF. Returns list of result dicts."""
    q = np.ascontiguousarray(queries, np.int32)
    n_q, k = q.shape
    res = (OrcResult * n_q)()
    n_threads = n_threads or os.cpu_count() or 1
    lib().orc_batch(ctypes.byref(inst.g), q.ctypes.data, n_q, k, D, algo, n_threads, time_budget_s, res)
    return [r.as_dict() for r in res]


# ------------------------------------------------------------------ brute force
def _adjacency(graph):
    n = graph.n
    adj = [dict() for _ in range(n)]
    for v in range(n):
        for e in range(graph.beg_pos[v], graph.beg_pos[v + 1]):
            u = int(graph.adj[e])
            w = int(graph.w[e])
            if u != v and (u not in adj[v] or w < adj[v][u]):
                adj[v][u] = w
    return adj


def brute_force(graph, groups, gids):
    """Exact GST weight by enumeration of vertex subsets (n <= ~14)."""
    n = graph.n
    adj = _adjacency(graph)
    gsets = [set(int(x) for x in groups.members(g)) for g in dict.fromkeys(int(x) for x in gids)]
    best = INF
    for r in range(1, n + 1):
        for S in itertools.combinations(range(n), r):
            Ss = set(S)
            if not all(gs & Ss for gs in gsets):
                continue
            # Prim on G[S]
            start = S[0]
            seen = {start}
            total = 0
            frontier = [(w, u) for u, w in adj[start].items() if u in Ss]
            import heapq
            heapq.heapify(frontier)
            while frontier and len(seen) < r:
                w, u = heapq.heappop(frontier)
                if u in seen:
                    continue
                seen.add(u)
                total += w
                for x, wx in adj[u].items():
                    if x in Ss and x not in seen:
                        heapq.heappush(frontier, (wx, x))
            if len(seen) == r and total < best:
                best = total
    return best


def brute_force_hop(graph, groups, gids, D):
    """Exact diameter-bounded GST weight by enumeration of edge subsets (m <= ~20)."""
    adj = _adjacency(graph)
    edges = sorted({(min(v, u), max(v, u), w) for v in range(graph.n) for u, w in adj[v].items()})
    gsets = [set(int(x) for x in groups.members(g)) for g in dict.fromkeys(int(x) for x in gids)]
    best = INF
    # single vertex trees
    for v in range(graph.n):
        if all(v in gs for gs in gsets):
            return 0
    for r in range(1, len(edges) + 1):
        for sub in itertools.combinations(edges, r):
            verts = set()
            for a, b, _ in sub:
                verts.add(a)
                verts.add(b)
            if len(verts) != r + 1:
                continue
            if not all(gs & verts for gs in gsets):
                continue
            nb = {v: [] for v in verts}
            for a, b, _ in sub:
                nb[a].append(b)
                nb[b].append(a)

            def bfs(s):
                dist = {s: 0}
                order = [s]
                for x in order:
                    for y in nb[x]:
                        if y not in dist:
                            dist[y] = dist[x] + 1
                            order.append(y)
                return dist

            d0 = bfs(next(iter(verts)))
            if len(d0) != len(verts):
                continue
            far = max(d0, key=d0.get)
            d1 = bfs(far)
            if max(d1.values()) > D:
                continue
            wsum = sum(w for _, _, w in sub)
            best = min(best, wsum)
    return best
