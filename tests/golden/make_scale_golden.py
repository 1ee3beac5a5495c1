"""Writes tests/golden/scale_<workload>.json: optimal GST weights of fixed query samples on
the large synthetic workloads, computed by the CPU oracle (oracle/prunedpp.cpp, the
restatement of PrunedDP++ / D-PrunedDP++), so that the GPU tests can check the engine
at full size without re-running oracle queries that take minutes each.

Every workload is regenerated bit-identically from its seed (gpu4gst-1_amd/datasets.py);
the fixture stores a fingerprint of the graph and group table, and the tests refuse a
fixture whose fingerprint differs from the graph they generated.

Run: python tests/golden/make_scale_golden.py [workload ...] [--threads N]
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import __graft_entry__  # noqa: E402

__graft_entry__.load_package()
from gpu4gst_1_amd import datasets  # noqa: E402


def fingerprint(wl):
    h = hashlib.sha256()
    for a in (wl.graph.beg_pos, wl.graph.adj, wl.graph.w, wl.groups.offs, wl.groups.verts):
        h.update(np.ascontiguousarray(a).view(np.uint8))
    return h.hexdigest()[:24]


def served_mix(wl, n, seed):
    """The served-form sample: k uniform in {3,5,7}, D uniform in {none,3,4,5}."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        k = int(rng.choice([3, 5, 7]))
        D = int(rng.choice([-1, 3, 4, 5]))
        out.append((rng.choice(wl.groups.n_groups, k, replace=False).astype(np.int32), D))
    return out


# Query samples. (kind, queries, D): plain queries come from the workload's query lists.
def samples(wl):
    q = wl.queries
    if wl.name == "cl168k":
        return {"plain": [(x, -1) for k in (3, 5, 7) for x in q[k][:16]]}
    if wl.name == "cl1m":
        return {"hop": [(x, D) for D in (3, 4, 5) for x in list(q[5][:3]) + list(q[7][:2])]}
    if wl.name == "cl3m":
        return {"plain": [(x, -1) for x in q[5][:8]],
                "mixed": [(x, -1) for x in list(q[3][:6]) + list(q[5][8:13]) + list(q[7][:5])]}
    if wl.name == "cl4m":
        # hop queries at k = 7 need more than ORACLE_BYTES in the oracle: they are checked
        # for "alone == in the mix" only, and carry no expected cost here
        mix = served_mix(wl, 24, datasets.SEED_BASE + 4 + 555)
        return {"served": [(x, D) for x, D in mix]}
    raise KeyError(wl.name)


ORACLE_BYTES = 40e9  # host memory one oracle query may take (its flat state arrays)


def oracle_bytes(n, k, D):
    return n * (1 << k) * (D + 1 if D >= 0 else 1) * (24 if D >= 0 else 16)


def run(name, threads):
    import pyoracle

    t0 = time.time()
    wl = datasets.make_workload(name)
    inst = pyoracle.Instance(wl.graph, wl.groups)
    fp = fingerprint(wl)
    path = os.path.join(HERE, f"scale_{name}.json")
    # rows already solved by an earlier, interrupted run on the same graph are kept: the
    # file is rewritten after every (k, D) group, marked "partial" until the last one
    done = {}
    if os.path.exists(path):
        with open(path) as f:
            old = json.load(f)
        if old.get("fingerprint") == fp:
            for kind, rows in old["samples"].items():
                for r in rows:
                    if "status" in r:
                        done[(kind, tuple(r["gids"]), r["D"])] = r
    out = {"workload": name, "n": wl.graph.n, "m": wl.graph.m, "fingerprint": fp, "samples": {}, "partial": True}
    print(f"{name}: generated in {time.time() - t0:.0f}s; {len(done)} rows kept", flush=True)

    def write():
        with open(path, "w") as f:
            json.dump(out, f, indent=1)

    todo = samples(wl)
    for kind, qs in todo.items():
        rows = [{"gids": [int(v) for v in x], "D": D} for x, D in qs]
        for r in rows:
            r.update(done.get((kind, tuple(r["gids"]), r["D"]), {}))
        out["samples"][kind] = rows
    for kind, qs in todo.items():
        rows = out["samples"][kind]
        # one oracle batch per (k, D), one query per thread; large hop arenas run fewer at once
        groups = {}
        for i, (x, D) in enumerate(qs):
            if "status" not in rows[i]:
                groups.setdefault((len(x), D), []).append(i)
        for (k, D), idx in sorted(groups.items()):
            per_query = oracle_bytes(wl.graph.n, k, D)
            if per_query > ORACLE_BYTES:
                continue
            nt = max(1, min(threads, len(idx), int(ORACLE_BYTES // per_query)))
            t = time.time()
            outs = pyoracle.batch(inst, np.stack([qs[i][0] for i in idx]), algo=0 if D < 0 else 1, D=max(D, 0),
                                  n_threads=nt)
            for i, o in zip(idx, outs):
                rows[i]["status"] = int(o["status"])
                rows[i]["cost"] = int(o["tree_cost"]) if o["status"] == 0 else None
                rows[i]["oracle_search_s"] = round(o["t_search_s"], 3)
                rows[i]["oracle_total_s"] = round(o["t_total_s"], 3)
            write()
            print(f"  {kind} k={k} D={D}: {len(idx)} queries on {nt} threads, {time.time() - t:.0f}s", flush=True)
    del out["partial"]
    write()
    print(f"wrote {path} ({time.time() - t0:.0f}s)", flush=True)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("workloads", nargs="*", default=["cl168k", "cl1m", "cl3m", "cl4m"])
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    a = ap.parse_args()
    for w in a.workloads:
        run(w, a.threads)
