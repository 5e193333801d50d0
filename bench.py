#!/usr/bin/env python3
"""Single-query SSSP benchmark of the MI355X MLMQ solver.

A step is one exact single-source query (reset + persistent solve) on the
resident graph.  Default workload: the 6,262,104-vertex road-like graph
(seeded 3-NN points in the unit square, symmetrised, Morton-ordered, weights
= scaled Euclidean length with the longest edge at 100,000), source = the
vertex nearest the centroid, delta 200000, queue L1SLF_L2DQ.

Prints ONE JSON line (rank 0).  value = GTEPS over the timed steps
(E_r * steps / wall seconds, E_r = out-degree sum of reachable vertices);
ms_per_step = wall ms per step.  "roofline" prices the persistent kernel's
algorithmic bytes B = 16 V_r + 12 E_r per query against its HIP-event
duration and the 8 TB/s HBM peak.  "cpu_baseline" is the CPU oracle
(single-thread binary-heap Dijkstra) on the same graph and source.
"""
import argparse
import json
import os
import platform
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "single-query ms & GTEPS; batched queries/s + p50/p99 latency (exact dists)"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


# ---- HBM traffic of the timed point (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE) ----

TRAFFIC_KEYS = ("workload", "queue", "delta", "buckets", "bucket_window", "l2_batch", "min_read",
                "l1_pop", "l1_donate", "light_split", "stale_check", "prefetch", "work_blocks")


def traffic_key(workload, queue, delta, args, work_blocks):
    """Canonical text of every setting that changes the persistent kernel's
    memory traffic: a traffic figure is attached to a bench line only when it
    was measured at exactly this point."""
    cfg = {"workload": workload, "queue": queue, "delta": int(delta), "buckets": args.buckets,
           "bucket_window": args.window, "l2_batch": args.l2_batch, "min_read": args.min_read,
           "l1_pop": args.l1_pop, "l1_donate": args.l1_donate, "light_split": args.light_split,
           "stale_check": args.stale_check, "prefetch": args.prefetch,
           "work_blocks": int(work_blocks)}
    return json.dumps({k: cfg[k] for k in TRAFFIC_KEYS}, sort_keys=True)


def load_traffic(path, key):
    try:
        tj = json.load(open(path))
    except (OSError, ValueError):
        return None
    entry = tj.get(key)
    return entry["bytes_per_launch"] if entry else None


def _counter_means(d, counter, kernel):
    """Mean over dispatches of `kernel` of `counter` (summed per dispatch) in
    the rocprofv3 counter_collection CSVs under directory d."""
    import csv
    import glob
    per = {}
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    for f in files:
        for r in csv.DictReader(open(f)):
            if r.get("Counter_Name") != counter or kernel not in r.get("Kernel_Name", ""):
                continue
            k = (f, r.get("Dispatch_Id") or r.get("Correlation_Id"))
            per[k] = per.get(k, 0.0) + float(r["Counter_Value"])
    if not per:
        raise SystemExit(f"no {counter} rows of {kernel} under {d}")
    return sum(per.values()) / len(per), len(per)


def traffic_update(args):
    """Bytes per launch of the persistent kernel from two counter runs of the
    bench, each scaled by what the same counter reads for a plain copy kernel
    of known size, stored under the bench line's traffic key."""
    line_path, fdir, wdir, cfdir, cwdir = args.traffic_update
    line = None
    for ln in open(line_path):
        ln = ln.strip()
        if ln.startswith("{") and '"roofline"' in ln:
            line = json.loads(ln)
    if line is None:
        raise SystemExit(f"no bench line in {line_path}")
    key = line["roofline"]["traffic_key"]
    kern = "sssp_persistent_kernel"
    fetch, nf = _counter_means(fdir, "FETCH_SIZE", kern)
    write, nw = _counter_means(wdir, "WRITE_SIZE", kern)
    cfetch, _ = _counter_means(cfdir, "FETCH_SIZE", "copy_u32_kernel")
    cwrite, _ = _counter_means(cwdir, "WRITE_SIZE", "copy_u32_kernel")
    fscale = args.copy_bytes / 1024.0 / cfetch
    wscale = args.copy_bytes / 1024.0 / cwrite
    fb, wb = fetch * 1024.0 * fscale, write * 1024.0 * wscale
    try:
        tj = json.load(open(args.traffic_json))
    except (OSError, ValueError):
        tj = {}
    tj[key] = {"bytes_per_launch": int(round(fb + wb)), "fetch_bytes": int(round(fb)),
               "write_bytes": int(round(wb)), "fetch_kib_raw": fetch, "write_kib_raw": write,
               "dispatches": [nf, nw], "fetch_scale": fscale, "write_scale": wscale,
               "device_ms_of_line": line["device_ms"]["mean"],
               "note": f"rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE per dispatch of {kern}, scaled "
                       f"by the same counters on a {args.copy_bytes}-byte copy kernel"}
    with open(args.traffic_json, "w") as f:
        json.dump(tj, f, indent=1, sort_keys=True)
    print(json.dumps({"key": key, **tj[key]}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="road6m")
    ap.add_argument("--delta", type=int, default=None,
                    help="bucket width; default: best of the workload's delta sweep")
    ap.add_argument("--mode", choices=["single", "batch", "serve"], default="single",
                    help="single: one query per step (default); batch: all sources in one launch")
    ap.add_argument("--queries", type=int, default=256, help="batch mode: number of sources")
    ap.add_argument("--rates", default="100,500,1000,1500,2000,2500,3000",
                    help="serve mode: Poisson arrival rates (queries/s)")
    ap.add_argument("--duration", type=float, default=5.0, help="serve mode: seconds per rate")
    ap.add_argument("--serve-graphs", default="road264k,road6m,rmat20")
    ap.add_argument("--mix", default="0.5,0.3,0.2", help="serve mode: graph choice weights")
    ap.add_argument("--pool", type=int, default=128,
                    help="serve mode: verified sources per graph (drawn uniformly)")
    ap.add_argument("--serve-parts", default=None,
                    help="serve mode: one-wave workgroups per request of each graph (comma list, "
                         "as many as --serve-graphs); default: each graph's own setting")
    ap.add_argument("--serve-mode", choices=["live", "lanes"], default="live",
                    help="serve mode: live = one persistent launch fed through a request ring; "
                         "lanes = launches of waiting requests on launch lanes")
    ap.add_argument("--budget-gib", type=float, default=64.0)
    ap.add_argument("--streams", type=int, default=4)
    ap.add_argument("--spread-waves", type=int, default=0,
                    help="serve mode: waves a request arriving at a quiet server may spread over "
                         "(0: each graph's own workgroup count, the lowest p99 under load)")
    ap.add_argument("--max-batch", type=int, default=64)
    ap.add_argument("--max-pending", type=int, default=4096)
    ap.add_argument("--p99-bound-ms", type=float, default=50.0,
                    help="serve mode: value = best completed rate with p99 under this bound")
    ap.add_argument("--tune", type=int, default=1, help="batch/serve modes: run the autotuner")
    ap.add_argument("--tune-single", type=int, default=0,
                    help="single mode: tune on --tune-sources other sources instead of timing "
                         "the workload's committed bench point")
    ap.add_argument("--tune-sources", type=int, default=3)
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--traffic-update", nargs=5, default=None,
                    metavar=("BENCH_LINE", "FETCH_DIR", "WRITE_DIR", "COPY_FETCH_DIR", "COPY_WRITE_DIR"),
                    help="write the traffic entry of a profiled bench line into --traffic-json "
                         "from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE output directories of the "
                         "bench run and of a --calibrate-copy run")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30,
                    help="--traffic-update: size of the calibration copy")
    ap.add_argument("--delta-sweep", type=int, default=1,
                    help="single mode: time the workload's delta sweep at the timed point's other "
                         "settings (0: only the timed point runs, e.g. under a counter profiler)")
    ap.add_argument("--extended-sweep", type=int, default=1,
                    help="single mode: also time deltas 2M and 20M (outside the configured sweep)")
    ap.add_argument("--batch-waves", type=int, default=8)
    ap.add_argument("--l1-ring", type=int, default=0,
                    help="batch mode: L1 ring records per wave, 512 or 1024 (0: 512)")
    ap.add_argument("--source-seed", type=int, default=12345)
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="only run a plain copy kernel of this many bytes (counter calibration)")
    ap.add_argument("--queue", default="L1SLF_L2DQ")
    ap.add_argument("--work-blocks", type=int, default=0)
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--l1-pop", type=int, default=None,
                    help="records per L1 pop (default 64 single-query, 32 batch)")
    ap.add_argument("--l1-donate", type=int, default=0)
    ap.add_argument("--l2-batch", type=int, default=8)
    ap.add_argument("--min-read", type=int, default=8)
    ap.add_argument("--buckets", type=int, default=16)
    ap.add_argument("--window", type=int, default=4)
    ap.add_argument("--light-split", type=int, default=0)
    ap.add_argument("--prefetch", type=int, default=1,
                    help="1: touch relaxed children's hop blocks in the atomics' round trip")
    ap.add_argument("--stale-check", type=int, default=2,
                    help="2 coherent, 1 through the L2, 3 late (in the atomics' round trip), 0 none")
    ap.add_argument("--cpu-runs", type=int, default=5,
                    help="CPU baseline: at least this many oracle runs (within --cpu-cap)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="CPU baseline: keep running the oracle for at least this long")
    ap.add_argument("--cpu-cap", type=float, default=30.0,
                    help="CPU baseline: stop starting new oracle runs after this many seconds")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic-json", default=os.path.join(ROOT, "profiles", "traffic.json"),
                    help="per-launch HBM bytes measured by rocprofv3 --pmc (optional)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "independent queries would shard by replicating the graph per GPU.",
              file=sys.stderr)
        sys.exit(2)

    if args.traffic_update:
        return traffic_update(args)

    import numpy as np
    from mlmq_tpds_withl3_amd import Solver, make_workload, default_params

    if args.calibrate_copy:
        with Solver(0) as s:
            ms = [s.probe_copy(args.calibrate_copy) for _ in range(3)]
        print(json.dumps({"copy_bytes": args.calibrate_copy, "read_bytes": args.calibrate_copy,
                          "write_bytes": args.calibrate_copy, "device_ms": ms,
                          "gbs": 2 * args.calibrate_copy / (min(ms) / 1e3) / 1e9}), flush=True)
        return

    if args.mode == "batch":
        return run_batch(args)
    if args.mode == "serve":
        return run_serve(args)

    t0 = time.time()
    g, src, spec = make_workload(args.workload)
    row, col, w = g.arrays()
    log(f"workload {args.workload}: n={g.n} m={g.m} src={src} gen {time.time() - t0:.1f}s")
    if args.l1_pop is None:
        args.l1_pop = 32

    def make_params(delta):
        return default_params(queue=args.queue, delta=delta, work_blocks=args.work_blocks,
                              blocks_per_cu=args.blocks_per_cu, l2_batch=args.l2_batch,
                              min_read=args.min_read, buckets=args.buckets,
                              bucket_window=args.window, l1_pop=args.l1_pop,
                              l1_donate=args.l1_donate, light_split=args.light_split,
                              stale_check=args.stale_check, prefetch=args.prefetch,
                              timeout_s=30.0)

    solver = Solver(0)
    solver.load_csr(row, col, w)
    shape = solver.device_shape()
    tuned_on = None
    if args.delta is not None:
        delta = args.delta
    else:
        point = dict(spec.get("bench_point") or {})
        if args.tune_single:
            # Tune on other sources than the timed one (seeded, uniform over
            # the vertices), so the timed point is not fitted to the measured
            # query: the workload's delta sweep crossed with its grid.
            from mlmq_tpds_withl3_amd.autotune import tune_single
            # Tuning sources: seeded, uniform over the vertices the timed source
            # reaches (its component; on power-law graphs many vertices are
            # isolated and a query from them is trivial), other than itself.
            solver.solve(src, make_params(point.get("delta", spec["delta"])))
            reach = np.flatnonzero(solver.dist() != 0x7FFFFFFF)
            reach = reach[reach != src]
            rng = np.random.default_rng(args.source_seed)
            tsrc = [int(x) for x in rng.choice(reach, size=min(args.tune_sources, len(reach)),
                                                replace=False)]
            tuned = tune_single(solver, tsrc, spec["delta_sweep"], grid=spec.get("tune_grid"),
                                log=log if args.verbose else None)
            point = tuned["params"]
            tuned_on = {"sources": tsrc, "device_ms": round(tuned["device_ms"], 4), "params": point}
            log(f"tuned on sources {tsrc}: {point} {tuned['device_ms']:.3f} ms")
        delta = point.get("delta", spec["delta"])
        for key, attr in (("l1_pop", "l1_pop"), ("min_read", "min_read"),
                          ("work_blocks", "work_blocks"), ("l2_batch", "l2_batch"),
                          ("buckets", "buckets"), ("bucket_window", "window"),
                          ("light_split", "light_split"), ("stale_check", "stale_check"),
                          ("prefetch", "prefetch")):
            if key in point:
                setattr(args, attr, point[key])
    params = make_params(delta)

    # The workload's delta sweep at the timed point's other settings, on the
    # timed source: reported, not used to choose the point.
    def sweep_ms(deltas):
        table = {}
        for d in deltas:
            runs = sorted(solver.solve(src, make_params(int(d)))["device_ms"] for _ in range(3))
            table[str(int(d))] = round(runs[1], 4)
        return table
    sweep = sweep_ms(spec["delta_sweep"]) if args.delta_sweep else None
    log(f"delta sweep (median of 3, device ms): {sweep}")
    extended = sweep_ms((2_000_000, 20_000_000)) if args.extended_sweep and args.delta_sweep else None
    if extended:
        log(f"outside the configured sweep: {extended}")
    for i in range(args.warmup):
        t = solver.solve(src, params)
        log(f"warmup {i}: solve_ms={t['solve_ms']:.3f} device_ms={t['device_ms']:.3f}")

    timings = []
    wall0 = time.perf_counter()
    for _ in range(args.steps):
        timings.append(solver.solve(src, params))
    wall = time.perf_counter() - wall0
    dist = solver.dist()
    log(f"timed {args.steps} steps in {wall * 1e3:.1f} ms")
    # The reference's query wall time: reset + solve + copy of the distances
    # to (pageable) host memory, per query; outside the timed region.
    query_wall = []
    for _ in range(5):
        q0 = time.perf_counter()
        solver.solve(src, params)
        solver.dist()
        query_wall.append((time.perf_counter() - q0) * 1e3)

    # ---- CPU baseline leg: the oracle, also the correctness check ----
    # Single-threaded, pinned to one core of this process's CPU set.
    from oracle import oracle as O
    cores = sorted(os.sched_getaffinity(0))
    pinned = cores[0]
    os.sched_setaffinity(0, {pinned})
    try:
        t_start = time.perf_counter()
        c0 = time.perf_counter()
        ref, e_r, v_r = O.dijkstra(row, col, w, src)
        cpu_times = [time.perf_counter() - c0]
        correct = bool(np.array_equal(dist, ref))
        while not args.no_cpu_baseline:
            spent = time.perf_counter() - t_start
            if spent >= args.cpu_cap or (len(cpu_times) >= args.cpu_runs
                                         and spent >= args.cpu_seconds):
                break
            c0 = time.perf_counter()
            O.dijkstra(row, col, w, src)
            cpu_times.append(time.perf_counter() - c0)
    finally:
        os.sched_setaffinity(0, set(cores))
    cpu_s = statistics.median(cpu_times)
    log(f"cpu oracle {cpu_s * 1e3:.1f} ms (runs {len(cpu_times)}, core {pinned}), "
        f"correct={correct}")

    solve_ms = [t["solve_ms"] for t in timings]
    dev_ms = [t["device_ms"] for t in timings]
    ms_per_step = wall * 1e3 / args.steps
    gteps = e_r * args.steps / wall / 1e9
    b_alg = 16 * v_r + 12 * e_r
    dev_avg_s = statistics.mean(dev_ms) / 1e3
    achieved = b_alg / dev_avg_s / 1e9
    last = timings[-1]
    tkey = traffic_key(args.workload, args.queue, delta, args, last["work_blocks"])
    traffic = load_traffic(args.traffic_json, tkey)
    last = timings[-1]
    out = {
        "metric": METRIC,
        "value": round(gteps, 6),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "uint32",
        "data": "synthetic",
        "config": {
            "workload": args.workload,
            "vertices": g.n, "edges": g.m, "source": src,
            "reachable_vertices": v_r, "reachable_edges": e_r,
            "weight_scale": spec.get("weight_scale_applied"),
            "seed": spec.get("seed"),
            "queue": args.queue, "delta": delta,
            "delta_sweep_device_ms": sweep,
            "outside_sweep_device_ms": extended,
            "point": "tuned on other sources" if tuned_on else
                     ("command line" if args.delta is not None else "committed bench point"),
            "tuned_on": tuned_on,
            "buckets": args.buckets, "bucket_window": args.window, "light_split": args.light_split,
            "stale_check": args.stale_check, "prefetch": args.prefetch,
            "l2_batch": args.l2_batch, "min_read": args.min_read, "l1_pop": args.l1_pop, "l1_donate": args.l1_donate,
            "work_blocks": last["work_blocks"], "waves": last["waves"],
            "cus": shape["cus"], "blocks_per_cu_max": shape["blocks_per_cu_max"],
        },
        "solve_ms": {"median": round(statistics.median(solve_ms), 4),
                     "min": round(min(solve_ms), 4), "max": round(max(solve_ms), 4)},
        "device_ms": {"median": round(statistics.median(dev_ms), 4),
                      "mean": round(statistics.mean(dev_ms), 4)},
        "query_wall_ms": {"median": round(statistics.median(query_wall), 4),
                          "runs": len(query_wall)},
        "work": {"relaxations": last["relaxations"], "atomics": last["atomics"],
                 "improvements": last["improvements"], "l2_pushes": last["l2_pushes"],
                 "l2_pops": last["l2_pops"], "stale": last["stale"],
                 "batches": last["batches"], "donated": last["donated"],
                 "idle_polls": last["idle_polls"],
                 "idle_fraction": round(last["idle_polls"]
                                        / max(1, last["idle_polls"] + last["batches"]), 4),
                 "inflation": round(last["relaxations"] / max(1, e_r), 4)},
        "correct": correct,
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 4),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 8),
            "traffic": traffic,
            "traffic_key": tkey,
            "kernel": "sssp_persistent_kernel",
            "bytes_per_launch": b_alg,
        },
        "cpu_baseline": {
            "value": round(e_r / cpu_s / 1e9, 6),
            "unit": "GTEPS",
            "cores": 1,
            "kind": "port",
            "sample": (f"full {args.workload} query from source {src} by the CPU oracle "
                       f"(binary-heap Dijkstra, 1 thread pinned to core {pinned}), median of "
                       f"{len(cpu_times)} runs ({sum(cpu_times):.1f} s): {cpu_s * 1e3:.1f} ms"),
            "ms": round(cpu_s * 1e3, 3),
            "runs": len(cpu_times),
            "cpu": cpu_model(),
            "pinned_core": pinned,
            "nproc": os.cpu_count(),
            "affinity_cpus": len(cores),
        },
    }
    print(json.dumps(out), flush=True)
    solver.close()
    if not correct:
        sys.exit(3)


def run_batch(args):
    """Batched multi-source mode: `--queries` seeded uniform sources on the
    workload (default road1m: 1,070,376 vertices) in one persistent launch per
    step, against the same queries as one-at-a-time solves."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    from mlmq_tpds_withl3_amd import Solver, make_workload, default_params
    from oracle import oracle as O

    workload = args.workload if args.workload != "road6m" else "road1m"
    if args.l1_pop is None:
        args.l1_pop = 32
    g, _, spec = make_workload(workload)
    row, col, w = g.arrays()
    sources = np.random.default_rng(args.source_seed).integers(0, g.n, size=args.queries)
    delta = args.delta if args.delta is not None else spec["delta"]
    # timed (and tuned) without the statistics counters, which cost registers
    # and so waves per CU; one more launch with them supplies the work counters
    base = dict(delta=delta, batch_waves=args.batch_waves, work_blocks=args.work_blocks,
                l1_pop=args.l1_pop, min_read=args.min_read, l2_batch=args.l2_batch,
                buckets=args.buckets, bucket_window=args.window, prefetch=args.prefetch,
                stale_check=args.stale_check, batch_stats=0)
    if args.l1_ring:
        base["l1_ring"] = args.l1_ring
    solver = Solver(0)
    solver.load_csr(row, col, w)
    tuned = None
    if args.tune and args.delta is None:
        from mlmq_tpds_withl3_amd.autotune import tune_batch
        tuned = tune_batch(solver, sources, base=base, log=log)
        base = tuned["params"]
        log(f"tuned batch parameters: {base} ({tuned['batch_s'] * 1e3:.1f} ms per batch)")
    params = default_params(timeout_s=60.0, **base)
    for _ in range(args.warmup):
        solver.solve_batch(sources, params)
    times = []
    wall0 = time.perf_counter()
    for _ in range(args.steps):
        times.append(solver.solve_batch(sources, params))
    wall = time.perf_counter() - wall0
    log(f"batch: {args.steps} x {args.queries} queries in {wall * 1e3:.1f} ms")
    counted = solver.solve_batch(sources, default_params(timeout_s=60.0, **dict(base, batch_stats=1)))
    # verify every query against the oracle (thread pool; ctypes releases the
    # GIL); the same pass, timed, is the CPU baseline: P independent Dijkstras
    pool = min(16, os.cpu_count() or 4)
    c0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=pool) as ex:
        refs = list(ex.map(lambda s: O.dijkstra(row, col, w, int(s)), sources))
    cpu_wall = time.perf_counter() - c0
    correct = all(np.array_equal(solver.batch_dist(i), refs[i][0]) for i in range(len(sources)))
    e_r = sum(r[1] for r in refs)
    b_alg = sum(16 * r[2] + 12 * r[1] for r in refs)
    # comparator: the same queries one at a time (plain solves, parameters
    # tuned for single queries the same way as the single-query mode)
    sp = default_params(delta=delta, timeout_s=60.0)
    if args.tune:
        from mlmq_tpds_withl3_amd.autotune import tune_single
        st_ = tune_single(solver, int(sources[0]), spec["delta_sweep"])
        sp = default_params(timeout_s=60.0, **st_["params"])
        log(f"one-at-a-time parameters: {st_['params']}")
    solver.solve(int(sources[0]), sp)
    s0 = time.perf_counter()
    seq_ms = [solver.solve(int(s), sp)["solve_ms"] for s in sources]
    seq_wall = time.perf_counter() - s0
    qps = args.queries * args.steps / wall
    seq_qps = args.queries / seq_wall
    last = times[-1]
    dev_avg_s = statistics.mean(t["device_ms"] for t in times) / 1e3
    work = counted   # the statistics launch (same queries and setting)
    achieved = b_alg / dev_avg_s / 1e9
    out = {
        "metric": METRIC, "value": round(qps, 3), "unit": "queries/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "uint32", "data": "synthetic",
        "config": {"workload": workload, "mode": "batch", "queries": args.queries,
                   "vertices": g.n, "edges": g.m, "source_seed": args.source_seed,
                   "delta": params.delta, "tuned": base if tuned else None,
                   "batch_waves": last["waves_per_query"],
                   "work_blocks": last["work_blocks"], "graph_replay": last["graph_replay"],
                   "bucket_capacity": last["bucket_capacity"]},
        "gteps": round(e_r * args.steps / wall / 1e9, 4),
        "query_ms": {"p50": round(last["query_ms_p50"], 3), "p99": round(last["query_ms_p99"], 3),
                     "max": round(last["query_ms_max"], 3)},
        "sequential": {"queries_per_s": round(seq_qps, 3), "wall_ms": round(seq_wall * 1e3, 2),
                       "solve_ms_median": round(statistics.median(seq_ms), 3)},
        "speedup_vs_sequential": round(qps / seq_qps, 3),
        "device_ms": {"mean": round(dev_avg_s * 1e3, 4)},
        "work": {"relaxations": work["relaxations"], "improvements": work["improvements"],
                 "l2_pushes": work["l2_pushes"], "stale": work["stale"],
                 "batches": work["batches"], "idle_polls": work["idle_polls"],
                 "reachable_edges": e_r, "device_ms_with_counters": round(work["device_ms"], 4),
                 "inflation": round(work["relaxations"] / max(1, e_r), 4)},
        "correct": correct,
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 4), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 8), "traffic": None,
            "kernel": "sssp_batch_kernel", "bytes_per_launch": b_alg,
        },
        "cpu_baseline": {
            "value": round(args.queries / cpu_wall, 3), "unit": "queries/s", "cores": pool,
            "kind": "port",
            "sample": (f"the same {args.queries} queries by the CPU oracle (binary-heap "
                       f"Dijkstra), {pool} independent queries at a time on a thread pool: "
                       f"{cpu_wall * 1e3:.1f} ms"),
            "cpu": cpu_model(), "nproc": os.cpu_count(),
        },
    }
    print(json.dumps(out), flush=True)
    solver.close()
    if not correct:
        sys.exit(3)


def run_serve(args):
    """Serving mode: Poisson arrivals over three resident graphs (default
    road264k / road6m / rmat20 chosen 50/30/20 %), sources drawn uniformly
    from a seeded pool per graph (128 by default) whose oracle rows are
    registered for on-device verification; one trace per rate in `--rates`,
    served live (one persistent launch fed through a request ring, default)
    or by launch lanes.  Comparator: the same mix served one query at a time
    by plain solves."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    from mlmq_tpds_withl3_amd import Solver, make_workload, default_params
    from mlmq_tpds_withl3_amd.serve import Server, poisson_trace
    from mlmq_tpds_withl3_amd.autotune import tune_serving
    from oracle import oracle as O

    names = args.serve_graphs.split(",")
    mix = [float(x) for x in args.mix.split(",")]
    rates = [float(x) for x in args.rates.split(",")]
    if len(mix) != len(names):
        raise SystemExit("--mix needs one weight per graph")
    rng = np.random.default_rng(args.source_seed)
    server = Server(0, hbm_budget=int(args.budget_gib * (1 << 30)), streams=args.streams,
                    spread_waves=args.spread_waves,
                    max_batch=args.max_batch, max_pending=args.max_pending, verify=1,
                    timeout_s=60.0, mode=args.serve_mode)
    graphs, pools, seq_ms, tuned, cpu_ms = [], [], [], [], []
    P = min(16, os.cpu_count() or 4)
    ex = ThreadPoolExecutor(max_workers=P)
    for gi, name in enumerate(names):
        t0 = time.time()
        g, _, spec = make_workload(name)
        row, col, w = g.arrays()
        pool = [int(x) for x in rng.integers(0, g.n, size=args.pool)]
        pools.append(pool)
        c0 = time.perf_counter()
        refs = list(ex.map(lambda s: O.dijkstra(row, col, w, s)[0], pool))
        cpu_ms.append((time.perf_counter() - c0) * 1e3 / len(pool))   # per query, P at a time
        gid = server.add_graph(g, delta=spec["delta"])
        for s_, r_ in zip(pool, refs):
            server.set_expected(gid, s_, r_)
        log(f"serve graph {gid} {name}: n={g.n} m={g.m} pool={len(pool)} "
            f"slot={server.slot_bytes(gid) / 2**20:.1f} MiB ({time.time() - t0:.1f}s)")
        # one-at-a-time comparator: plain solves of the pool sources, tuned
        # like the single-query mode; its parameters (or the workload's serve_point)
        # seed the serving tuner
        with Solver(0) as sv:
            sv.load_csr(row, col, w)
            from mlmq_tpds_withl3_amd.autotune import tune_single
            single = dict(spec.get("bench_point") or
                          tune_single(sv, pool[:3], spec["delta_sweep"], grid=spec.get("tune_grid"))["params"])
            seed = dict(spec.get("serve_point") or single)
            sp = default_params(timeout_s=60.0, **single)
            sv.solve(pool[0], sp)
            t1 = time.perf_counter()
            for s_ in pool:
                sv.solve(s_, sp)
            seq_ms.append((time.perf_counter() - t1) * 1e3 / len(pool))
        serve_keys = ("delta", "buckets", "bucket_window", "l2_batch", "min_read", "l1_pop",
                      "light_split")
        base = {k: v for k, v in seed.items() if k in serve_keys}
        if args.serve_parts:
            parts = int(args.serve_parts.split(",")[gi])
            base.update(workgroups=max(1, parts // 8), waves=8 if parts >= 8 else 1)
            if parts < 8:
                base["workgroups"] = parts
        if args.tune:
            t = tune_serving(server, gid, pool, base=base, repeats=2, log=log)
            tuned.append(t["params"])
            log(f"serve graph {gid} tuned: {t['params']} {t['qps']:.1f} q/s")
        else:
            server.set_graph_params(gid, **base)
            tuned.append(base)
        graphs.append({"name": name, "vertices": g.n, "edges": g.m, "params": tuned[-1],
                       "slot_mib": round(server.slot_bytes(gid) / 2**20, 2),
                       "plain_solve_ms": round(seq_ms[-1], 3)})
        g.close()
    wsum = sum(mix)
    seq_qps = 1e3 / sum(m / wsum * t for m, t in zip(mix, seq_ms))
    cpu_qps = 1e3 / sum(m / wsum * t for m, t in zip(mix, cpu_ms))
    log(f"one-at-a-time capacity of the mix: {seq_qps:.1f} q/s; CPU oracle, {P} threads: "
        f"{cpu_qps:.1f} q/s")
    sweep = []
    all_ok = True
    for k, rate in enumerate(rates):
        arr, gsel, src = poisson_trace(rate, args.duration, mix, pools, seed=100 + k)
        res, st = server.run(arr, gsel, src)
        ok = st["failed"] == 0 and st["mismatched"] == 0 and st["verified"] == st["completed"]
        all_ok &= ok
        per = {}
        for gi, name in enumerate(names):
            sel = (gsel == gi) & (res["status"] == 0)
            lat = res["latency_ms"][sel]
            per[name] = {"n": int(sel.sum()),
                         "p50_ms": round(float(np.percentile(lat, 50)), 3) if lat.size else None,
                         "p99_ms": round(float(np.percentile(lat, 99)), 3) if lat.size else None}
        row_ = {"rate": rate, "offered_qps": round(st["offered_qps"], 2),
                "completed_qps": round(st["completed_qps"], 2), "p50_ms": round(st["p50_ms"], 3),
                "p99_ms": round(st["p99_ms"], 3), "max_ms": round(st["max_ms"], 3),
                "queue_p99_ms": round(st["queue_p99_ms"], 3), "requests": st["requests"],
                "completed": st["completed"], "rejected": st["rejected"], "failed": st["failed"],
                "verified": st["verified"], "mismatched": st["mismatched"],
                "mean_batch": round(st["mean_batch"], 2), "launches": st["launches"],
                "peak_inflight": st["peak_inflight"], "peak_gib": round(st["peak_bytes"] / 2**30, 3),
                "retries": st["retries"], "per_graph": per}
        sweep.append(row_)
        log(f"rate {rate}: {json.dumps(row_)}")
    good = [r for r in sweep if r["rejected"] == 0 and r["p99_ms"] <= args.p99_bound_ms]
    best = max(good, key=lambda r: r["completed_qps"]) if good else None
    value = best["completed_qps"] if best else 0.0
    out = {
        "metric": METRIC, "value": value, "unit": "queries/s", "n_gpus": 1,
        "steps": len(rates), "warmup": 0, "ms_per_step": round(args.duration * 1e3, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "uint32",
        "data": "synthetic",
        "config": {"workload": "serve:" + "+".join(names), "mode": "serve",
                   "serve_mode": args.serve_mode, "mix": mix,
                   "duration_s": args.duration, "pool_per_graph": args.pool,
                   "budget_gib": args.budget_gib, "streams": args.streams,
                   "spread_waves": args.spread_waves, "serve_parts": args.serve_parts,
                   "max_batch": args.max_batch, "max_pending": args.max_pending,
                   "p99_bound_ms": args.p99_bound_ms, "graphs": graphs},
        "best_rate": best["rate"] if best else None,
        "p50_ms": best["p50_ms"] if best else None, "p99_ms": best["p99_ms"] if best else None,
        "one_at_a_time_qps": round(seq_qps, 2),
        "speedup_vs_one_at_a_time": round(value / seq_qps, 3) if seq_qps else None,
        "sweep": sweep, "correct": all_ok,
        "cpu_baseline": {
            "value": round(cpu_qps, 3), "unit": "queries/s", "cores": P, "kind": "port",
            "sample": (f"the CPU oracle (binary-heap Dijkstra) computing the {args.pool} "
                       f"verification rows of each graph, {P} independent queries at a time; "
                       f"mix-weighted capacity from the per-graph times "
                       + ", ".join(f"{n} {t:.1f} ms" for n, t in zip(names, cpu_ms))),
            "cpu": cpu_model(), "nproc": os.cpu_count(),
        },
    }
    print(json.dumps(out), flush=True)
    server.close()
    if not all_ok:
        sys.exit(3)


if __name__ == "__main__":
    main()
