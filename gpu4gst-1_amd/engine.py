"""ctypes mirror of include/gst.h (libgst.so).

Host-side names mirror the reference's per-query solve interface:
  Engine.DP_GPU(gids)        non-hop solve (code/TrimCDP/include/DPQ.cuh:63)
  Engine.DP_gpu(gids, D)     hop solve     (code/D-TrimCDP/include/DPQ.cuh:63)
Both return (real_cost, result dict) where real_cost follows the reference drivers'
no-solution convention (TrimCDP: 0, DPQ.cu:571-574; D-TrimCDP: 100000,
code/D-TrimCDP/src/DPQ.cu:229-242), while the result dict carries the status code.
The engine refuses to run without the HIP library: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

GST_OK, GST_NO_SOLUTION, GST_PENDING, GST_TREE_OMITTED = 0, 1, 2, 3
GST_BAD_INPUT, GST_OVER_BUDGET, GST_DEVICE_ERROR, GST_INTERNAL, GST_NO_DEVICE = -1, -2, -3, -4, -5
GST_COST_RANGE = -6
COST_CEILING = 0x3fffffff
F_STATS, F_NO_TREE, F_NO_PRUNE = 1, 2, 4
O_KEEP_ORDER = 0x100  # gst_opts.flags: keep the caller's vertex numbering on the device
O_NO_HUB_INDEX = 0x200  # gst_opts.flags: no hub distance index (queries start at UB = INF)
MAX_K, MAX_D = 7, 15


class GstCsr(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("m", ctypes.c_int64), ("beg_pos", ctypes.c_void_p),
                ("adj", ctypes.c_void_p), ("w", ctypes.c_void_p)]


class GstGroups(ctypes.Structure):
    _fields_ = [("n_groups", ctypes.c_int32), ("offs", ctypes.c_void_p), ("verts", ctypes.c_void_p)]


class GstOpts(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("degree_split", ctypes.c_int32), ("chunk_edges", ctypes.c_int32),
                ("max_batch", ctypes.c_int32), ("dp_budget_bytes", ctypes.c_int64), ("flags", ctypes.c_int32),
                ("graph_rounds", ctypes.c_int32), ("near_far_delta", ctypes.c_int32), ("cohort_max", ctypes.c_int32),
                ("cohort_min", ctypes.c_int32), ("cohort_levels", ctypes.c_int32)]


class GstQuery(ctypes.Structure):
    _fields_ = [("gids", ctypes.c_void_p), ("k", ctypes.c_int32), ("D", ctypes.c_int32), ("flags", ctypes.c_int32)]


class GstResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("cost", ctypes.c_int32), ("tree_cost", ctypes.c_int32),
                ("n_edges", ctypes.c_int32), ("edges", ctypes.POINTER(ctypes.c_int32)),
                ("root", ctypes.c_int32), ("root_height", ctypes.c_int32), ("k", ctypes.c_int32),
                ("D", ctypes.c_int32), ("rounds", ctypes.c_int32), ("t_total_s", ctypes.c_double),
                ("t_dp_s", ctypes.c_double), ("t_queue_s", ctypes.c_double), ("t_grow_s", ctypes.c_double),
                ("t_merge_s", ctypes.c_double), ("states_finite", ctypes.c_int64),
                ("states_processed", ctypes.c_int64), ("grow_vertices", ctypes.c_int64),
                ("grow_edges", ctypes.c_int64), ("grow_relax", ctypes.c_int64), ("grow_pushes", ctypes.c_int64),
                ("arena_bytes", ctypes.c_int64), ("level_edges", ctypes.c_int64 * 8),
                ("level_relax", ctypes.c_int64 * 8), ("grow_rounds", ctypes.c_int32),
                ("ub", ctypes.c_int32), ("lb", ctypes.c_int32), ("level", ctypes.c_int32), ("pad_", ctypes.c_int32)]


class GstInfo(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("m", ctypes.c_int64), ("graph_bytes", ctypes.c_int64),
                ("budget_bytes", ctypes.c_int64), ("in_use_bytes", ctypes.c_int64),
                ("active_jobs", ctypes.c_int32), ("queued_jobs", ctypes.c_int32), ("completed", ctypes.c_int64)]


class GstCohortStats(ctypes.Structure):
    _fields_ = [("cohorts", ctypes.c_int64), ("members", ctypes.c_int64), ("grow_launches", ctypes.c_int64),
                ("grow_device_s", ctypes.c_double), ("grow_arcs", ctypes.c_int64), ("grow_relax", ctypes.c_int64)]


EXPORTS = ["gst_open", "gst_solve", "gst_solve_batch", "gst_submit", "gst_poll", "gst_wait", "gst_info_get",
           "gst_set_tuning", "gst_result_free", "gst_close", "gst_strerror", "gst_last_error", "gst_arena_bytes",
           "gst_job_bytes", "gst_solve_noalloc", "gst_set_tuning_k", "gst_get_cohort_stats"]


def lib_path() -> str:
    return os.path.join(_HERE, "libgst.so")


def lib():
    global _LIB
    if _LIB is None:
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build() (no CPU fallback exists)")
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        i32 = ctypes.c_int32
        L.gst_open.argtypes = [ctypes.POINTER(GstCsr), ctypes.POINTER(GstGroups), ctypes.POINTER(GstOpts),
                               ctypes.POINTER(P)]
        L.gst_solve.argtypes = [P, P, i32, i32, i32, ctypes.POINTER(GstResult)]
        L.gst_solve_batch.argtypes = [P, ctypes.POINTER(GstQuery), i32, ctypes.POINTER(GstResult)]
        L.gst_submit.argtypes = [P, ctypes.POINTER(GstQuery), i32, ctypes.POINTER(ctypes.c_uint64)]
        L.gst_poll.argtypes = [P, ctypes.c_uint64, ctypes.POINTER(GstResult)]
        L.gst_wait.argtypes = [P, ctypes.c_uint64, ctypes.POINTER(GstResult)]
        L.gst_info_get.argtypes = [P, ctypes.POINTER(GstInfo)]
        L.gst_set_tuning.argtypes = [P, i32, i32, i32]
        L.gst_set_tuning_k.argtypes = [P, i32, i32, i32, i32]
        L.gst_result_free.argtypes = [ctypes.POINTER(GstResult)]
        L.gst_close.argtypes = [P]
        L.gst_strerror.restype = ctypes.c_char_p
        L.gst_strerror.argtypes = [ctypes.c_int]
        L.gst_last_error.restype = ctypes.c_char_p
        L.gst_arena_bytes.restype = ctypes.c_int64
        L.gst_arena_bytes.argtypes = [i32, i32, i32]
        L.gst_solve_noalloc.argtypes = [P, P, i32, i32, i32, P, i32, ctypes.POINTER(GstResult)]
        L.gst_job_bytes.restype = ctypes.c_int64
        L.gst_job_bytes.argtypes = [P, i32, i32]
        L.gst_get_cohort_stats.argtypes = [P, ctypes.POINTER(GstCohortStats), i32]
        _LIB = L
    return _LIB


class GstError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"{lib().gst_strerror(code).decode()} ({code}): {msg}")
        self.code = code


def _result_dict(r: GstResult) -> dict:
    d = {f: getattr(r, f) for f, _ in GstResult._fields_ if f not in ("edges", "level_edges", "level_relax", "pad_")}
    d["level_edges"] = list(r.level_edges)
    d["level_relax"] = list(r.level_relax)
    n = max(r.n_edges, 0)
    d["edges"] = np.ctypeslib.as_array(r.edges, (3 * n,)).reshape(n, 3).copy() if n and r.edges else \
        np.zeros((0, 3), np.int32)
    return d


@dataclass
class Query:
    gids: np.ndarray
    D: int = -1
    flags: int = 0


class Engine:
    """One resident graph + group table on one GPU."""

    def __init__(self, graph, groups, device=0, degree_split=0, chunk_edges=0, max_batch=0,
                 dp_budget_bytes=0, flags=0, graph_rounds=0, near_far_delta=0, cohort_max=0, cohort_min=0,
                 cohort_levels=0):
        L = lib()
        self._beg = np.ascontiguousarray(graph.beg_pos, np.int64)
        self._adj = np.ascontiguousarray(graph.adj, np.int32)
        self._w = np.ascontiguousarray(graph.w, np.int32)
        self._goffs = np.ascontiguousarray(groups.offs, np.int64)
        self._gverts = np.ascontiguousarray(groups.verts, np.int32)
        csr = GstCsr(len(self._beg) - 1, len(self._adj), self._beg.ctypes.data, self._adj.ctypes.data,
                     self._w.ctypes.data)
        grp = GstGroups(len(self._goffs) - 1, self._goffs.ctypes.data, self._gverts.ctypes.data)
        opts = GstOpts(device, degree_split, chunk_edges, max_batch, dp_budget_bytes, flags, graph_rounds,
                       near_far_delta, cohort_max, cohort_min, cohort_levels)
        h = ctypes.c_void_p()
        rc = L.gst_open(ctypes.byref(csr), ctypes.byref(grp), ctypes.byref(opts), ctypes.byref(h))
        if rc != GST_OK:
            raise GstError(rc, L.gst_last_error().decode())
        self._h = h
        self.n = csr.n
        self.m = csr.m

    def close(self):
        if getattr(self, "_h", None):
            lib().gst_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -------------------------------------------------------------- plain path
    def solve(self, gids, D=-1, flags=0) -> dict:
        L = lib()
        g = np.ascontiguousarray(gids, np.int32)
        r = GstResult()
        rc = L.gst_solve(self._h, g.ctypes.data, len(g), int(D), int(flags), ctypes.byref(r))
        try:
            if rc < 0 and rc not in (GST_INTERNAL, GST_COST_RANGE):
                raise GstError(rc, L.gst_last_error().decode())
            d = _result_dict(r)
            if rc in (GST_INTERNAL, GST_COST_RANGE):
                d["error"] = L.gst_last_error().decode()
            return d
        finally:
            L.gst_result_free(ctypes.byref(r))

    def solve_noalloc(self, gids, D=-1, flags=0, edge_cap=None) -> dict:
        """gst_solve_noalloc: no device or host allocation in the call; the tree goes to a
        caller buffer of edge_cap triples (default n - 1, which always suffices)."""
        L = lib()
        g = np.ascontiguousarray(gids, np.int32)
        cap = max(self.n - 1, 1) if edge_cap is None else int(edge_cap)
        buf = np.zeros(3 * max(cap, 1), np.int32)
        r = GstResult()
        rc = L.gst_solve_noalloc(self._h, g.ctypes.data, len(g), int(D), int(flags), buf.ctypes.data, cap,
                                 ctypes.byref(r))
        if rc < 0 and rc not in (GST_INTERNAL, GST_COST_RANGE):
            raise GstError(rc, L.gst_last_error().decode())
        d = _result_dict(r)
        n = max(r.n_edges, 0) if r.status in (GST_OK,) else 0
        d["edges"] = buf[:3 * n].reshape(n, 3).copy()
        return d

    def DP_GPU(self, gids):
        """Non-hop solve with the reference driver's cost convention (0 = no solution)."""
        d = self.solve(gids, -1)
        return (d["cost"] if d["status"] == GST_OK else 0), d

    def DP_gpu(self, gids, D):
        """Hop solve with the reference driver's cost convention (100000 = no solution)."""
        d = self.solve(gids, D)
        return (d["cost"] if d["status"] == GST_OK else 100000), d

    def set_tuning(self, degree_split, chunk_edges, near_far_delta=0):
        rc = lib().gst_set_tuning(self._h, int(degree_split), int(chunk_edges), int(near_far_delta))
        if rc:
            raise GstError(rc, lib().gst_last_error().decode())

    def set_tuning_k(self, k, degree_split, chunk_edges, merge_tile=0):
        """Grow tuning and merge tile for queries of k groups only."""
        rc = lib().gst_set_tuning_k(self._h, int(k), int(degree_split), int(chunk_edges), int(merge_tile))
        if rc:
            raise GstError(rc, lib().gst_last_error().decode())

    # -------------------------------------------------------------- batched / served
    @staticmethod
    def _queries(qs):
        keep, arr = [], (GstQuery * len(qs))()
        for i, q in enumerate(qs):
            if not isinstance(q, Query):
                q = Query(np.asarray(q[0]), *q[1:]) if isinstance(q, tuple) else Query(np.asarray(q))
            g = np.ascontiguousarray(q.gids, np.int32)
            keep.append(g)
            arr[i] = GstQuery(g.ctypes.data, len(g), int(q.D), int(q.flags))
        return arr, keep

    def solve_batch(self, qs) -> list:
        L = lib()
        arr, keep = self._queries(qs)
        res = (GstResult * len(qs))()
        rc = L.gst_solve_batch(self._h, arr, len(qs), res)
        if rc < 0:
            raise GstError(rc, L.gst_last_error().decode())
        out = []
        for r in res:
            out.append(_result_dict(r))
            L.gst_result_free(ctypes.byref(r))
        del keep
        return out

    def submit(self, qs) -> list:
        L = lib()
        arr, keep = self._queries(qs)
        tickets = (ctypes.c_uint64 * len(qs))()
        rc = L.gst_submit(self._h, arr, len(qs), tickets)
        if rc < 0:
            raise GstError(rc, L.gst_last_error().decode())
        del keep
        return list(tickets)

    def poll(self, ticket, progress=False):
        """Result of a finished query; for one in flight None, or with progress=True a dict
        with status GST_PENDING and the bounds so far (ub: best tree weight found, -1 if
        none yet; lb: lower bound on the optimum; level: DP level reached)."""
        L = lib()
        r = GstResult()
        rc = L.gst_poll(self._h, ctypes.c_uint64(ticket), ctypes.byref(r))
        if rc == GST_PENDING:
            return _result_dict(r) if progress else None
        if rc < 0 and r.status == 0:
            raise GstError(rc, L.gst_last_error().decode())
        try:
            return _result_dict(r)
        finally:
            L.gst_result_free(ctypes.byref(r))

    def wait(self, ticket):
        L = lib()
        r = GstResult()
        rc = L.gst_wait(self._h, ctypes.c_uint64(ticket), ctypes.byref(r))
        if rc < 0 and r.status == 0:
            raise GstError(rc, L.gst_last_error().decode())
        try:
            return _result_dict(r)
        finally:
            L.gst_result_free(ctypes.byref(r))

    def job_bytes(self, k, D=-1) -> int:
        """Device bytes one admitted query of this shape is charged against the budget."""
        b = lib().gst_job_bytes(self._h, int(k), int(D))
        if b < 0:
            raise GstError(int(b), "bad query shape")
        return b

    def cohort_stats(self, reset=False) -> dict:
        """Totals over the cohorts run since open or the last reset: cohorts, members, and
        their level-1 grow launches, device seconds, arcs scanned and label relaxations."""
        st = GstCohortStats()
        rc = lib().gst_get_cohort_stats(self._h, ctypes.byref(st), 1 if reset else 0)
        if rc != GST_OK:
            raise GstError(rc, "gst_get_cohort_stats")
        return {f: getattr(st, f) for f, _ in GstCohortStats._fields_}

    def info(self) -> dict:
        i = GstInfo()
        lib().gst_info_get(self._h, ctypes.byref(i))
        return {f: getattr(i, f) for f, _ in GstInfo._fields_}


def arena_bytes(n, k, D=-1):
    return lib().gst_arena_bytes(n, k, D)


def validate_tree(graph, groups, gids, edges, cost, D=-1):
    """Check that edges form a tree of weight `cost` covering every query group and, when
    D >= 0, of unweighted diameter <= D. Returns (ok, reason)."""
    gids = list(dict.fromkeys(int(x) for x in gids))
    edges = np.asarray(edges).reshape(-1, 3)
    verts = set()
    for a, b, _ in edges:
        verts.add(int(a))
        verts.add(int(b))
    if len(edges) == 0:
        # single vertex tree: some vertex must be in every group
        common = None
        for g in gids:
            mem = set(int(x) for x in groups.members(g))
            common = mem if common is None else common & mem
        return (cost == 0 and bool(common)), "single-vertex tree"
    if len(verts) != len(edges) + 1:
        return False, "not a tree (|V| != |E| + 1)"
    # edge weights must match the graph
    for a, b, w in edges:
        a, b = int(a), int(b)
        nb = graph.adj[graph.beg_pos[a]:graph.beg_pos[a + 1]]
        ws = graph.w[graph.beg_pos[a]:graph.beg_pos[a + 1]]
        hit = np.nonzero(nb == b)[0]
        if len(hit) == 0 or int(ws[hit].min()) != int(w):
            return False, f"edge ({a},{b},{w}) not in graph"
    if int(edges[:, 2].sum()) != int(cost):
        return False, f"weight {int(edges[:, 2].sum())} != cost {cost}"
    nbr = {v: [] for v in verts}
    for a, b, _ in edges:
        nbr[int(a)].append(int(b))
        nbr[int(b)].append(int(a))

    def bfs(s):
        dist = {s: 0}
        order = [s]
        for x in order:
            for y in nbr[x]:
                if y not in dist:
                    dist[y] = dist[x] + 1
                    order.append(y)
        return dist

    d0 = bfs(next(iter(verts)))
    if len(d0) != len(verts):
        return False, "disconnected"
    for g in gids:
        if not (set(int(x) for x in groups.members(g)) & verts):
            return False, f"group {g} not covered"
    if D >= 0:
        far = max(d0, key=d0.get)
        if max(bfs(far).values()) > D:
            return False, "diameter exceeds D"
    return True, "ok"
