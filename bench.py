"""Benchmark: GST queries/s at k=5, batched vs plain, on one MI355X.

Workload: Chung-Lu power-law graph (degree exponent 2.0), 3.07M vertices / 117M
undirected edges, largest connected component, int32 weights uniform in [1,100],
20,000 Zipf(1.5)-sized groups in [1,50]; queries of k=5 distinct group ids (seeded).

A step = one batch of --batch queries solved by the batched persistent path
(gst_solve_batch). value = queries/s over the timed steps (inputs resident in HBM).
Beside it: the plain single-query path on --plain-queries of the same queries (q/s,
p50/p99), the grow-round roofline (algorithmic bytes / HIP-event time of the select and
grow kernels of each round, plain path), the same for the batched step's dominant kernel,
the cohort's level-1 grow (roofline_batched), and the CPU PrunedDP++ restatement timed on
the host cores on a bounded sample, whose costs also serve as the parity check of the GPU
results.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
import __graft_entry__  # noqa: E402

__graft_entry__.load_package()
from gpu4gst_1_amd import autotune, datasets, engine  # noqa: E402

HBM_PEAK_GBS = 8000.0


def b_grow(r):
    """Algorithmic bytes of the grow rounds of one query:
    8|S| (beg_pos pair) + 8 E_S (adj + weight) + 4|F| (source cost) + 8 R (target cost read
    + atomic write) + 4 |F_next| (worklist push)."""
    return (8 * r["grow_vertices"] + 8 * r["grow_edges"] + 4 * r["states_processed"] + 8 * r["grow_relax"]
            + 4 * r["grow_pushes"])


def serve(args):
    """Served form: Poisson arrivals at swept rates; p50/p99 latency from arrival to result;
    every served result compared with the same query solved alone (plain path)."""
    import torch

    assert torch.cuda.is_available(), "bench needs a GPU"
    t0 = time.time()
    wl = datasets.make_workload(args.workload)
    g = wl.graph
    print(f"# serve workload {args.workload}: n={g.n} edges={g.n_edges} groups={wl.groups.n_groups} "
          f"({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    rng = np.random.default_rng(datasets.SEED_BASE + 99)
    total = args.serve_queries * len(args.rates)
    qs = []
    for _ in range(total):
        k = int(rng.choice([3, 5, 7]))
        D = int(rng.choice([-1, 3, 4, 5]))
        qs.append(engine.Query(rng.choice(wl.groups.n_groups, k, replace=False).astype(np.int32), D))
    eng = engine.Engine(g, wl.groups, max_batch=args.batch, dp_budget_bytes=int(args.budget_gb * 1e9))
    # plain reference of every query (also warms the code paths)
    alone = []
    tp = time.perf_counter()
    for q in qs:
        alone.append(eng.solve(q.gids, D=q.D))
    plain_dt = time.perf_counter() - tp
    plain_qps = len(qs) / plain_dt
    print(f"# plain: {plain_qps:.2f} q/s over {len(qs)} mixed queries ({time.time() - t0:.1f}s)",
          file=sys.stderr, flush=True)
    rows = []
    same = 0
    for ri, rate in enumerate(args.rates):
        part = qs[ri * args.serve_queries:(ri + 1) * args.serve_queries]
        ref = alone[ri * args.serve_queries:(ri + 1) * args.serve_queries]
        gaps = rng.exponential(1.0 / rate, size=len(part))
        tickets = []
        start = time.perf_counter()
        nxt = start
        for q, gap in zip(part, gaps):
            nxt += gap
            while time.perf_counter() < nxt:
                time.sleep(min(0.0005, max(0.0, nxt - time.perf_counter())))
            tickets += eng.submit([q])
        res = [eng.wait(t) for t in tickets]
        dt = time.perf_counter() - start
        lat = np.array([r["t_total_s"] for r in res]) * 1e3
        same += sum(1 for r, a in zip(res, ref) if r["status"] == a["status"] and r["cost"] == a["cost"])
        rows.append({"rate_qps": rate, "achieved_qps": round(len(res) / dt, 3), "p50_ms": round(float(np.median(lat)), 2),
                     "p99_ms": round(float(np.percentile(lat, 99)), 2),
                     "queue_p50_ms": round(float(np.median([r["t_queue_s"] for r in res])) * 1e3, 2)})
        print(f"# rate {rate}: {rows[-1]} ({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    out = {"metric": "served GST queries/s and p50/p99 ms (Poisson arrivals, mixed k and D)",
           "value": max(r["achieved_qps"] for r in rows), "unit": "queries/s", "n_gpus": 1,
           "higher_is_better": True, "dtype": "int32", "data": "synthetic",
           "config": {"workload": f"{args.workload}: n={g.n}, undirected edges={g.n_edges}, "
                                  f"{wl.groups.n_groups} groups, k in {{3,5,7}}, D in {{none,3,4,5}}",
                      "dp_budget_gb": args.budget_gb, "slots": args.batch},
           "rates": rows, "plain_qps": round(plain_qps, 3),
           "served_equal_alone": f"{same}/{len(qs)}"}
    print(json.dumps(out))
    eng.close()


def calibrate_copy(mib):
    """Two device-to-device copies of a known size, for the HBM byte counters: the first
    reads a region that 3 * mib MiB of later writes pushed out of the 256 MB Infinity
    Cache, the second reads the region written last. Under rocprofv3 --pmc FETCH_SIZE (or
    WRITE_SIZE) each copy dispatch should read (write) mib * 1024 KiB."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.empty(4 * n, dtype=torch.int32, device="cuda")
    src.fill_(7)
    dst = torch.empty(n, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    dst.copy_(src[:n])           # cold source
    torch.cuda.synchronize()
    dst.copy_(src[3 * n:])       # source written last
    torch.cuda.synchronize()
    assert bool((dst[-4:] == 7).all())
    # random gather: 2^22 int32 reads at distinct random 64-B sectors of the 4 * mib MiB
    # source (one per sector at most), written densely
    nsec = 4 * n // 16
    g = 1 << 22
    idx = (torch.randperm(nsec, device="cuda")[:g] * 16).to(torch.int64)
    torch.cuda.synchronize()
    out = torch.index_select(src, 0, idx)
    torch.cuda.synchronize()
    assert bool((out[:4] == 7).all())
    print(json.dumps({"calibrate_copy_mib": mib, "expected_kib_each_way": mib * 1024,
                      "copies": ["cold source", "recently written source"],
                      "gather": {"reads": g, "distinct_64B_sectors": g, "index_kib": g * 8 // 1024,
                                 "output_kib": g * 4 // 1024}}))


def gather_ceiling():
    """Rate of independent random 4-B reads on one GPU: torch.index_select of 2^25 uniform
    random int32 elements from arrays of 64 MiB to 4 GiB, timed with HIP events on torch's
    stream (10 launches after 3 warm-ups). A grow round's target-state reads have the same
    pattern (one random 4-B read per arc over a level of the arena, 61-123 MB for the k=5
    workload), without its dependent loads and atomics; this is the ceiling for them."""
    import torch

    torch.manual_seed(0)
    n_idx = 1 << 25
    rows = []
    for mib in (64, 256, 1024, 4096):
        n = mib * (1 << 20) // 4
        x = torch.randint(0, 1000, (n,), dtype=torch.int32, device="cuda")
        idx = torch.randint(0, n, (n_idx,), dtype=torch.int64, device="cuda")
        for _ in range(3):
            y = torch.index_select(x, 0, idx)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(10):
            y = torch.index_select(x, 0, idx)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 10
        rows.append({"array_mib": mib, "reads": n_idx, "ms": round(ms, 4),
                     "g_reads_per_s": round(n_idx / ms / 1e6, 2),
                     "line_tb_per_s": round(n_idx * 128 / ms / 1e9, 2)})
        del x, idx, y
        torch.cuda.empty_cache()
    print(json.dumps({"gather_ceiling": rows}))


# ---------------------------------------------------------------- HBM traffic
TRAFFIC_KERNELS = ("k_select", "k_grow_thread", "k_grow_wave")  # one grow round (the roofline launch)
COPY_MIB = 256


def traffic_probe(shm, split, chunk, n_q):
    """Child of a counter pass (run under rocprofv3 --pmc): one device-to-device copy of
    COPY_MIB MiB (the calibration dispatch), then n_q plain solves of the workload saved in
    shm. Prints the grow rounds of the solves as JSON."""
    import ctypes

    hip = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    nb = COPY_MIB << 20
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nb)) == 0 and hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nb)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nb), 3) == 0 and hip.hipDeviceSynchronize() == 0  # 3: device to device
    hip.hipFree(a)
    hip.hipFree(b)
    ld = lambda name: np.load(os.path.join(shm, name + ".npy"), mmap_mode="r")
    graph = datasets.Graph(np.asarray(ld("beg")), np.asarray(ld("adj")), np.asarray(ld("w")))
    groups = datasets.Groups(np.asarray(ld("offs")), np.asarray(ld("verts")))
    qs = np.asarray(ld("queries"))
    eng = engine.Engine(graph, groups)
    eng.set_tuning(split, chunk, 0)
    rounds = [eng.solve(q)["grow_rounds"] for q in qs[:n_q]]
    eng.close()
    print(json.dumps({"grow_rounds": rounds}))


def _counter_sums(outdir):
    """Per kernel: (dispatches, summed counter value) from rocprofv3's counter CSV."""
    import csv
    import glob

    files = glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True)
    per_dispatch = {}
    for fn in files:
        with open(fn) as f:
            for row in csv.DictReader(f):
                key = (row.get("Dispatch_Id") or row.get("Correlation_Id"), row["Kernel_Name"])
                per_dispatch[key] = per_dispatch.get(key, 0.0) + float(row["Counter_Value"])
    out = {}
    for (_, name), v in per_dispatch.items():
        n, tot, mx = out.get(name, (0, 0.0, 0.0))
        out[name] = (n + 1, tot + v, max(mx, v))
    return out


def measure_traffic(wl, queries, split, chunk, log):
    """HBM bytes per grow round of the plain path from two counter passes (FETCH_SIZE, then
    WRITE_SIZE, KiB per dispatch), each scaled by what the same counter reads for the copy
    dispatch of known size in the same process. Runs before this process touches the GPU
    (the passes are child processes). Returns a dict, or None with the reason logged."""
    import shutil
    import subprocess
    import tempfile

    if not shutil.which("rocprofv3"):
        log("# traffic: rocprofv3 not found")
        return None
    shm = tempfile.mkdtemp(prefix="gst_bench_", dir="/dev/shm" if os.path.isdir("/dev/shm") else None)
    try:
        for name, arr in (("beg", wl.graph.beg_pos), ("adj", wl.graph.adj), ("w", wl.graph.w),
                          ("offs", wl.groups.offs), ("verts", wl.groups.verts), ("queries", np.stack(queries))):
            np.save(os.path.join(shm, name + ".npy"), np.ascontiguousarray(arr))
        res = {}
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            out = os.path.join(shm, counter)
            cmd = ["rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", out, "-o", "pass", "--",
                   sys.executable, os.path.abspath(__file__), "--traffic-probe", shm, str(split), str(chunk)]
            env = dict(os.environ, TMPDIR="/tmp")
            t = time.time()
            r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                               text=True, timeout=240)
            lines = [x for x in r.stdout.splitlines() if x.startswith("{")]
            if not lines:
                log(f"# traffic: {counter} pass failed (rc {r.returncode}): {r.stdout[-400:]}")
                return None
            rounds = sum(json.loads(lines[-1])["grow_rounds"])
            sums = _counter_sums(out)
            copy = [v for k, v in sums.items() if "copyBuffer" in k]
            grow = {k: v for k, v in sums.items() if k.split("(")[0].split("::")[-1] in TRAFFIC_KERNELS}
            n_sel = sum(v[0] for k, v in grow.items() if "k_select" in k)
            if not copy or not grow or n_sel != rounds:
                log(f"# traffic: {counter} pass unusable (copy {bool(copy)}, grow kernels {len(grow)}, "
                    f"select dispatches {n_sel} vs rounds {rounds})")
                return None
            factor = (COPY_MIB << 10) / max(v[2] for v in copy)  # KiB the copy moved / KiB counted
            kib = sum(v[1] for v in grow.values())
            res[counter] = {"kib_per_round": kib / rounds, "copy_factor": round(factor, 4), "rounds": rounds}
            log(f"# traffic: {counter} {kib / rounds / 1024:.1f} MiB per round counted, copy factor {factor:.3f} "
                f"({time.time() - t:.0f}s)")
        fetch = res["FETCH_SIZE"]["kib_per_round"] * res["FETCH_SIZE"]["copy_factor"] * 1024
        write = res["WRITE_SIZE"]["kib_per_round"] * res["WRITE_SIZE"]["copy_factor"] * 1024
        return {"bytes_per_round": int(fetch + write), "read_bytes_per_round": int(fetch),
                "write_bytes_per_round": int(write), "passes": res,
                "tuning": {"degree_split": split, "chunk_edges": chunk}}
    except Exception as e:  # the bench line still prints, with traffic null
        log(f"# traffic: {type(e).__name__}: {e}")
        return None
    finally:
        shutil.rmtree(shm, ignore_errors=True)


# ---------------------------------------------------------------- extra legs
def hop_leg(log, k=5, D=4, n_q=16):
    """Diameter-bounded plain path on the 1.1M-vertex Chung-Lu graph (the hop workload):
    q/s and p50/p99 over n_q queries of k groups with bound D (the reference times its
    persistent hop solve per query, code/D-TrimCDP-WB/src/GPUHop.cu:215-219)."""
    wl = datasets.make_workload("cl1m")
    with engine.Engine(wl.graph, wl.groups) as eng:
        qs = wl.queries[k][:n_q]
        eng.solve(qs[0], D=D)
        res = [eng.solve(q, D=D) for q in qs]
    lat = np.array([r["t_total_s"] for r in res]) * 1e3
    out = {"workload": f"cl1m: Chung-Lu n={wl.graph.n}, undirected edges={wl.graph.n_edges}, "
                       f"{wl.groups.n_groups} groups, k={k}, D={D}",
           "qps": round(len(res) / (lat.sum() / 1e3), 3), "p50_ms": round(float(np.median(lat)), 3),
           "p99_ms": round(float(np.percentile(lat, 99)), 3), "queries": len(res),
           "solved": sum(r["status"] == engine.GST_OK for r in res)}
    log(f"# hop leg: {out}")
    return out


def served_leg(log, rates=(8.0, 16.0, 64.0), n_q=48, budget_gb=200.0, slots=64):
    """Served form on the 4.8M-vertex Chung-Lu graph: Poisson arrivals at each of `rates`
    q/s with k uniform in {3,5,7} and D uniform in {none,3,4,5}, admitted against a DP
    budget; q/s (queries over the time from the first arrival to the last result) and
    p50/p99 latency from arrival to result at each rate, beside the plain path on the same
    queries, and every served result compared with the same query solved alone. The
    highest rate exceeds what the GPU sustains on this mix, so its q/s is the served
    capacity."""
    wl = datasets.make_workload("cl4m")
    rng = np.random.default_rng(datasets.SEED_BASE + 4 + 777)
    qs = []
    for _ in range(n_q):
        k = int(rng.choice([3, 5, 7]))
        D = int(rng.choice([-1, 3, 4, 5]))
        qs.append(engine.Query(rng.choice(wl.groups.n_groups, k, replace=False).astype(np.int32), D))
    sweep = []
    same = 0
    with engine.Engine(wl.graph, wl.groups, max_batch=slots, dp_budget_bytes=int(budget_gb * 1e9)) as eng:
        # warm-up of both forms: the first use of each (k, D) shape allocates its DP slots
        # (up to 15 GB each here), which belongs to neither steady state
        for q in qs:
            eng.solve(q.gids, D=q.D)
        eng.solve_batch(qs)
        tp = time.perf_counter()
        alone = [eng.solve(q.gids, D=q.D) for q in qs]
        plain_dt = time.perf_counter() - tp
        for rate in rates:
            gaps = rng.exponential(1.0 / rate, size=n_q)
            tickets = []
            start = time.perf_counter()
            nxt = start
            for q, gap in zip(qs, gaps):
                nxt += gap
                while time.perf_counter() < nxt:
                    time.sleep(min(0.0005, max(0.0, nxt - time.perf_counter())))
                tickets += eng.submit([q])
            res = [eng.wait(t) for t in tickets]
            dt = time.perf_counter() - start
            lat = np.array([r["t_total_s"] for r in res]) * 1e3
            same += sum(1 for r, a in zip(res, alone) if r["status"] == a["status"] and r["cost"] == a["cost"])
            sweep.append({"arrival_rate_qps": rate, "qps": round(n_q / dt, 3),
                          "p50_ms": round(float(np.median(lat)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2)})
            log(f"# served at {rate} q/s: {sweep[-1]}")
    out = {"workload": f"cl4m: Chung-Lu n={wl.graph.n}, undirected edges={wl.graph.n_edges}, "
                       f"{wl.groups.n_groups} groups, k in {{3,5,7}}, D in {{none,3,4,5}}, {n_q} queries per rate",
           "rates": sweep, "capacity_qps": max(x["qps"] for x in sweep), "plain_qps": round(n_q / plain_dt, 3),
           "equal_alone": f"{same}/{n_q * len(rates)}", "dp_budget_gb": budget_gb, "slots": slots}
    log(f"# served leg: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--k", type=int, default=5)
    ap.add_argument("--D", type=int, default=-1, help="diameter bound of the plain-path queries (-1: none)")
    ap.add_argument("--workload", default="cl3m")
    ap.add_argument("--plain-queries", type=int, default=24)
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU-baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-only", action="store_true", help="plain path only (for rocprof runs)")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (library defaults)")
    ap.add_argument("--tuning", type=int, nargs=2, metavar=("SPLIT", "CHUNK"),
                    help="skip the autotuner and use this degree split and chunk size")
    ap.add_argument("--stats", action="store_true", help="count finite states (extra arena scan per query)")
    ap.add_argument("--serve", action="store_true", help="served form with Poisson arrivals instead")
    ap.add_argument("--rates", type=float, nargs="+", default=[2.0, 8.0, 32.0])
    ap.add_argument("--serve-queries", type=int, default=32)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--gather-ceiling", action="store_true",
                    help="time independent random 4-B reads (the ceiling of the grow round's access pattern)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device once (HBM counter calibration)")
    ap.add_argument("--no-traffic", action="store_true", help="skip the two rocprofv3 counter passes")
    ap.add_argument("--no-legs", action="store_true", help="skip the hop and served legs")
    ap.add_argument("--traffic-probe", nargs=3, metavar=("SHM", "SPLIT", "CHUNK"), help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.traffic_probe:
        return traffic_probe(args.traffic_probe[0], int(args.traffic_probe[1]), int(args.traffic_probe[2]), 3)
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)
    if args.gather_ceiling:
        return gather_ceiling()
    if args.serve and args.gpus == 1:
        if args.workload == "cl3m":
            args.workload = "cl4m"
        return serve(args)
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(the path shards by query, see the design notes)"}))
        sys.exit(2)
    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    t0 = time.time()
    wl = datasets.make_workload(args.workload)
    g = wl.graph
    log(f"# workload {args.workload}: n={g.n} edges={g.n_edges} groups={wl.groups.n_groups} "
        f"(generated in {time.time() - t0:.1f}s)")
    # counter passes first: they are child processes, started before this one uses the GPU
    traffic = None
    if not args.no_traffic and not args.profile_only:
        split, chunk = args.tuning if args.tuning else (32, 1024)
        traffic = measure_traffic(wl, wl.queries[args.k][-3:], split, chunk, log)
    import torch

    assert torch.cuda.is_available(), "bench needs a GPU"
    need = args.batch * (args.steps + args.warmup)
    n_tune = max(32, args.batch)  # queries after the batches, used only by the autotuner
    rng = np.random.default_rng(datasets.SEED_BASE + 77)
    pool = wl.queries[args.k]
    extra = [rng.choice(wl.groups.n_groups, args.k, replace=False).astype(np.int32)
             for _ in range(max(0, need + n_tune - len(pool)))]
    queries = [np.asarray(q, np.int32) for q in pool] + extra
    tune_q = queries[need:need + n_tune]
    queries = queries[:need]

    eng = engine.Engine(g, wl.groups, max_batch=args.batch)
    print(f"# opened engine ({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    tuned = tuned_b = None
    if args.tuning:
        eng.set_tuning(args.tuning[0], args.tuning[1], 0)
        tuned = tuned_b = {"degree_split": args.tuning[0], "chunk_edges": args.tuning[1], "near_far_delta": 0}
    elif not args.no_tune:
        # tuned on queries outside every batch
        tuned, _ = autotune.tune(eng, tune_q[:3], grid={"degree_split": (32, 64, 128),
                                                         "chunk_edges": (512, 1024, 4096),
                                                         "near_far_delta": (0,)}, budget_s=30.0)
        print(f"# autotuned (plain) {tuned} ({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    flags = engine.F_STATS if args.stats else 0

    # ---- plain path: one query at a time (also gives the grow-kernel roofline), on the
    # first queries of the first timed batch so that both paths see the same queries
    plain = []
    plain_q = queries[args.warmup * args.batch:][:args.plain_queries]
    eng.solve(plain_q[0], D=args.D, flags=flags)  # warm-up
    for q in plain_q:
        plain.append(eng.solve(q, D=args.D, flags=flags))
    lat = np.array([r["t_total_s"] for r in plain]) * 1e3
    plain_qps = len(plain) / max(sum(r["t_total_s"] for r in plain), 1e-9)
    gb = sum(b_grow(r) for r in plain)
    gt = sum(r["t_grow_s"] for r in plain)
    grow_gbs = gb / gt / 1e9 if gt > 0 else 0.0
    grow_rounds = max(sum(r["grow_rounds"] for r in plain), 1)
    tm = sum(r["t_merge_s"] for r in plain)
    arcs_gs = sum(r["grow_edges"] for r in plain) / gt / 1e9 if gt > 0 else 0.0
    print(f"# plain: {plain_qps:.2f} q/s p50 {np.median(lat):.2f} ms p99 {np.percentile(lat, 99):.2f} ms; "
          f"grow {grow_gbs:.1f} GB/s over {grow_rounds} rounds ({gb / grow_rounds / 1e6:.1f} MB, "
          f"{gt / grow_rounds * 1e6:.0f} us per round, {arcs_gs:.1f} G arcs/s); merge {tm * 1e3:.1f} ms "
          f"({time.time() - t0:.1f}s)",
          file=sys.stderr, flush=True)
    if args.profile_only:
        eng.close()
        return
    assert args.D < 0, "the batched benchmark runs the non-hop workload; --D is for --profile-only"

    # ---- batched path: its own tuning, warmup, then timed steps
    if not args.tuning and not args.no_tune:
        # one batch of the timed step's size per grid point (its cohorts then have the
        # timed step's size), timed twice
        tuned_b, _ = autotune.tune(eng, tune_q[:args.batch],
                                   grid={"degree_split": (64, 128, 256, 512), "chunk_edges": (512, 1024, 4096),
                                         "near_far_delta": (0,)}, repeats=2, budget_s=60.0, batched=True)
        print(f"# autotuned (batched) {tuned_b} ({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    batches = [queries[i * args.batch:(i + 1) * args.batch] for i in range(args.warmup + args.steps)]
    for b in batches[:args.warmup]:
        eng.solve_batch(b)
    torch.cuda.synchronize()
    eng.cohort_stats(reset=True)
    t1 = time.perf_counter()
    results = []
    for b in batches[args.warmup:]:
        results += eng.solve_batch(b)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t1
    # the batched step's dominant kernel: the cohort's level-1 grow (HIP events around
    # each launch on the batch stream); algorithmic bytes = 8 per arc scanned (adjacency +
    # weight) + 4 per label relaxation (the target label read)
    cs = eng.cohort_stats()
    roof_b = None
    if cs["grow_launches"] > 0 and cs["grow_device_s"] > 0:
        cg_bytes = 8 * cs["grow_arcs"] + 4 * cs["grow_relax"]
        cg_gbs = cg_bytes / cs["grow_device_s"] / 1e9
        roof_b = {"bound": "hbm", "achieved": round(cg_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": round(cg_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                  "kernel": "k_cgrow: cohort level-1 grow (batched path, HIP events on the batch stream)",
                  "launches": cs["grow_launches"],
                  "avg_launch_us": round(cs["grow_device_s"] / cs["grow_launches"] * 1e6, 1),
                  "bytes_per_launch": int(cg_bytes / cs["grow_launches"]),
                  "bytes_per_unit": "8 per arc scanned + 4 per label relaxation",
                  "cohorts": cs["cohorts"], "members": cs["members"],
                  "share_of_step": round(cs["grow_device_s"] / dt, 3)}
        print(f"# cohort grow: {cs['grow_launches']} launches, {roof_b['avg_launch_us']:.0f} us each, "
              f"{cg_gbs:.1f} GB/s algorithmic, {roof_b['share_of_step']:.0%} of the timed steps",
              file=sys.stderr, flush=True)
    nq = sum(len(b) for b in batches[args.warmup:])
    qps = nq / dt
    blat = np.array([r["t_total_s"] for r in results]) * 1e3
    ok = sum(r["status"] == engine.GST_OK for r in results)
    print(f"# batched: {qps:.2f} q/s over {nq} queries in {dt:.2f}s ({ok} solved)", file=sys.stderr, flush=True)

    # ---- parity of batched vs plain on the shared queries
    by_q = {}
    for b, rs in zip([q for bb in batches[args.warmup:] for q in bb], results):
        by_q[tuple(b.tolist())] = rs
    plain_match = sum(1 for q, r in zip(plain_q, plain)
                      if tuple(q.tolist()) in by_q and by_q[tuple(q.tolist())]["cost"] == r["cost"])

    # ---- per-k tuning picks (keyed by k and the graph's degree profile) and one batch
    # mixing k = 3, 5 and 7 on the same graph, each job on its own k's pick
    mixed = None
    if not args.no_legs:
        picks = {}
        for kk in (3, 5, 7):
            cand = [np.asarray(q, np.int32) for q in wl.queries[kk][-8:]]
            picks[kk], _ = autotune.tune_k(eng, g, kk, cand, budget_s=20.0, batched=True)
        log(f"# per-k picks (batched) {picks} ({time.time() - t0:.1f}s)")
        mq = [engine.Query(np.asarray(q, np.int32)) for kk in (3, 5, 7) for q in wl.queries[kk][100:116]]
        order = np.random.default_rng(datasets.SEED_BASE + 88).permutation(len(mq))
        mq = [mq[i] for i in order]
        eng.solve_batch(mq[:8])
        tm0 = time.perf_counter()
        mres = eng.solve_batch(mq)
        mdt = time.perf_counter() - tm0
        chk = mq[:6]
        same = sum(1 for q, r in zip(chk, mres[:6]) if eng.solve(q.gids)["cost"] == r["cost"])
        mixed = {"queries": len(mq), "k": [3, 5, 7], "qps": round(len(mq) / mdt, 3),
                 "picks": {str(kk): v for kk, v in picks.items()}, "equal_plain": f"{same}/{len(chk)}"}
        log(f"# mixed-k batch: {mixed}")

    # ---- CPU baseline: PrunedDP++ restatement on the host cores, bounded sample
    cpu = None
    parity = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle

        inst = pyoracle.Instance(g, wl.groups)
        # one query per worker thread; the pool is sized to this GPU's share of the host
        # (16 threads per GPU), whatever the host's core count
        host_cpus = os.cpu_count() or 1
        cores = min(host_cpus, 16)
        sample = np.stack(plain_q[:max(cores, 1)])
        tc = time.perf_counter()
        ores = pyoracle.batch(inst, sample, algo=0, n_threads=cores, time_budget_s=args.cpu_budget)
        cdt = time.perf_counter() - tc
        solved = [(i, o) for i, o in enumerate(ores) if o["status"] != -99]
        e2e = np.array([o["t_total_s"] for _, o in solved]) * 1e3
        srch = np.array([o["t_search_s"] for _, o in solved]) * 1e3
        cpu = {"value": round(len(solved) / cdt, 4), "unit": "queries/s", "cores": cores, "host_cpus": host_cpus,
               "kind": "port",
               "sample": f"{len(solved)} queries of k={args.k} on the same graph, one per thread, "
                         f"{cdt:.1f}s wall",
               # per query: end to end (SSSP, bounds, search, tree) and the reference's
               # timing scope, the best-first loop only (CPUNONHOP.h:634-665)
               "p50_ms": round(float(np.median(e2e)), 1) if len(e2e) else None,
               "p99_ms": round(float(np.percentile(e2e, 99)), 1) if len(e2e) else None,
               "search_p50_ms": round(float(np.median(srch)), 1) if len(srch) else None,
               "search_p99_ms": round(float(np.percentile(srch, 99)), 1) if len(srch) else None}
        gpu_cost = {}
        for q, r in zip(plain_q, plain):
            gpu_cost[tuple(q.tolist())] = r["cost"]
        for q, r in by_q.items():
            gpu_cost.setdefault(q, r["cost"])
        checked = [(tuple(sample[i].tolist()), o) for i, o in solved if tuple(sample[i].tolist()) in gpu_cost]
        agree = sum(1 for q, o in checked if o["tree_cost"] == gpu_cost[q])
        parity = f"{agree}/{len(checked)}"
        print(f"# cpu: {cpu['value']} q/s on {cores} threads; parity {parity} ({time.time() - t0:.1f}s)",
              file=sys.stderr, flush=True)
    eng.close()
    del eng

    hop = served = None
    if not args.no_legs:
        hop = hop_leg(log)
        served = served_leg(log)
    traffic_b = traffic["bytes_per_round"] if traffic else None

    out = {
        "metric": "GST queries/s at k=5 (batched persistent path), p50/p99 ms, bit-exact vs PrunedDP++",
        "value": round(qps, 3),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded Chung-Lu graph, Zipf groups, uniform queries)",
        "config": {"workload": f"{args.workload}: Chung-Lu beta=2.0, n={g.n}, undirected edges={g.n_edges}, "
                               f"{wl.groups.n_groups} groups, k={args.k}",
                   "batch": args.batch, "plain_queries": len(plain), "tuning": tuned,
                   "tuning_batched": tuned_b},
        "p50_ms": round(float(np.median(blat)), 3),
        "p99_ms": round(float(np.percentile(blat, 99)), 3),
        "plain": {"qps": round(plain_qps, 3), "p50_ms": round(float(np.median(lat)), 3),
                  "p99_ms": round(float(np.percentile(lat, 99)), 3)},
        "batched_vs_plain": round(qps / plain_qps, 3) if plain_qps > 0 else None,
        "parity_vs_cpu": parity,
        "parity_batched_vs_plain": f"{plain_match}/{len(plain)}",
        "roofline": {"bound": "hbm", "achieved": round(grow_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(grow_gbs / HBM_PEAK_GBS, 4), "traffic": traffic_b,
                     # measured HBM bytes per grow round / algorithmic bytes per grow round
                     "traffic_over_algorithmic": round(traffic_b / (gb / grow_rounds), 3) if traffic_b else None,
                     "traffic_detail": traffic,
                     "kernel": "grow round: k_select + k_grow_thread + k_grow_wave (plain path, HIP events "
                               "on the engine stream)",
                     "launches": grow_rounds, "avg_launch_us": round(gt / grow_rounds * 1e6, 1),
                     "bytes_per_launch": int(gb / grow_rounds),
                     # arcs scanned per second (one random target-state read each), to set
                     # against bench.py --gather-ceiling
                     "arcs_g_per_s": round(arcs_gs, 2)},
        "roofline_batched": roof_b,
        "cpu_baseline": cpu,
        "mixed_k_batch": mixed,
        "hop": hop,
        "served": served,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
