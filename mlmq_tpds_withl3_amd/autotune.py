"""Parameter search for batched and served queries.

The solver's speed on a graph depends on the bucket width (delta), on how
many waves work on one query, and on the L1 pop size; the best values differ
between road-like graphs (long hop diameter, weights up to 1e5) and power-law
graphs (hubs, weights up to 1e3).  These routines measure candidates on the
GPU with the graph's own queries and keep the fastest, one coordinate at a
time (a single pass of coordinate descent from the defaults).  Every
measurement is an ordinary exact solve: a candidate whose queries fail is
discarded.
"""
from __future__ import annotations

import time

import numpy as np


def _delta_grid(base: int) -> list[int]:
    return sorted({max(1, int(base * f)) for f in (0.25, 0.5, 1.0, 2.0, 4.0)})


def default_delta(row, w) -> int:
    """Starting bucket width: 8 mean edge weights, at least 1."""
    w = np.asarray(w)
    return max(1, int(8 * float(w.mean()))) if len(w) else 1


def tune_single(solver, src, deltas, grid=None, repeats: int = 2, confirm: int = 7,
                log=None) -> dict:
    """Grid search for single queries from `src` (one source or a list of
    sources: a point's time is then the mean over them): every delta in
    `deltas` crossed with `grid` (a list of parameter dicts, default: L1 pop
    size / L2 top-up threshold pairs, bucket counts and four grid sizes),
    then the late and L2-cached stale checks at the four best points and
    child-block touches off at the four best after that.  Each
    point is the best device time of `repeats` solves per source; the four
    best are then timed `confirm` times more and the lowest median wins (a
    best-of-two is often a lucky draw).  Returns the best point, its median
    and the whole table.  Tune on other sources than the one a benchmark
    then times, so the point is not fitted to the measured query."""
    from .solver import default_params

    sources = [int(s) for s in (src if np.ndim(src) else [src])]
    if grid is None:
        grid = [dict(l1_pop=pop, min_read=mr, l2_batch=lb, work_blocks=wb, buckets=nb,
                     bucket_window=min(4, nb))
                for wb in (192, 256, 384, 512) for nb in (16, 4, 2)
                for pop, mr, lb in ((32, 32, 8), (32, 32, 32), (32, 16, 16), (32, 8, 8), (32, 4, 4),
                                    (64, 8, 8))]
    table = []

    def measure(params):
        p = default_params(timeout_s=30.0, **params)
        ms = float(np.mean([min(solver.solve(s, p)["device_ms"] for _ in range(repeats))
                            for s in sources]))
        table.append((params, ms))
        if log:
            log(f"tune single: {params} -> {ms:.3f} ms")

    for d in deltas:
        for g in grid:
            measure(dict(g, delta=int(d)))
    # second stage: the other stale-check modes at the four best points, then
    # child-block touches off at the four best so far (on power-law graphs the
    # touched blocks are mostly never popped)
    for bp, _ in sorted(table, key=lambda t: t[1])[:4]:
        for sc in (3, 1):
            if bp.get("stale_check", 2) != sc:
                measure(dict(bp, stale_check=sc))
    for bp, _ in sorted(table, key=lambda t: t[1])[:4]:
        if bp.get("prefetch", 1) != 0:
            measure(dict(bp, prefetch=0))
    finals = []
    for bp, _ in sorted(table, key=lambda t: t[1])[:4]:
        p = default_params(timeout_s=30.0, **bp)
        meds = []
        for s in sources:
            runs = sorted(solver.solve(s, p)["device_ms"] for _ in range(max(1, confirm)))
            meds.append(runs[len(runs) // 2])
        finals.append((bp, float(np.mean(meds))))
        if log:
            log(f"tune single, confirm: {bp} -> median {finals[-1][1]:.3f} ms")
    best = min(finals, key=lambda t: t[1])
    return {"params": best[0], "device_ms": best[1], "table": table, "confirmed": finals,
            "sources": sources}


def tune_serving(server, gid: int, sources, base: dict | None = None, burst: int = 64,
                 repeats: int = 1, lone: int = 3, lone_slack: float | None = 0.25,
                 log=None) -> dict:
    """Tune one resident graph of a ``serve.Server`` for throughput: each
    candidate serves a burst of `burst` requests (all arriving at once, drawn
    from `sources`) and is scored by completed queries per second.

    The same parameters also serve requests that arrive alone (at low load a
    launch holds one request, spread over many waves), whose latency the
    burst does not see.  So each candidate also serves `lone` requests one at
    a time, and a candidate whose median lone latency exceeds the starting
    point's by more than `lone_slack` is not taken (None: throughput only);
    the starting point (`base`) should be the single-query tuned parameters."""
    rng = np.random.default_rng(7)
    src = np.asarray(sources, dtype=np.int32)[rng.integers(0, len(sources), size=burst)]
    arrival = np.zeros(burst)
    gsel = np.full(burst, gid, dtype=np.int32)
    lone_src = [int(x) for x in src[:max(1, lone)]]
    cur = dict(base or {})
    cur.setdefault("delta", 200000)

    def lone_ms():
        lat = []
        for s_ in lone_src:
            res, st = server.run(np.zeros(1), np.full(1, gid, dtype=np.int32),
                                 np.full(1, s_, dtype=np.int32))
            if st["completed"] != 1 or st["mismatched"]:
                return float("inf")
            lat.append(float(res["latency_ms"][0]))
        return float(np.median(lat))

    def score(params):
        server.set_graph_params(gid, **params)
        best = 0.0
        for _ in range(repeats):
            res, st = server.run(arrival, gsel, src)
            if st["completed"] != burst or st["mismatched"]:
                return 0.0, float("inf")
            best = max(best, st["completed_qps"])
        return best, (lone_ms() if lone_slack is not None else 0.0)

    best_s, lone0 = score(cur)
    lone_cap = lone0 * (1.0 + lone_slack) if lone_slack is not None else float("inf")
    trail = [(dict(cur), best_s, lone0)]
    space = [("delta", _delta_grid(cur["delta"])), ("workgroups", [1, 2, 4, 8, 16, 32]),
             ("waves", [1, 4, 8, 16]), ("l1_pop", [16, 32, 64]), ("min_read", [8, 16, 32])]
    if not cur.get("light_split"):
        space.append(("buckets", [16, 4, 2]))
    # child-block touches and the stale-check mode last (-1 = no touches)
    cur.setdefault("prefetch", 1)
    cur.setdefault("stale_check", 2)
    space += [("prefetch", [-1, 1]), ("stale_check", [1, 3, 2])]
    for key, values in space:
        for v in values:
            if cur.get(key) == v:
                continue
            cand = dict(cur, **{key: v})
            if key == "buckets":
                cand["bucket_window"] = min(4, v)
            if key == "min_read":
                cand["l2_batch"] = max(8, v)
            s, lms = score(cand)
            trail.append((cand, s, lms))
            if log:
                log(f"tune graph {gid}: {cand} -> {s:.1f} q/s, alone {lms:.2f} ms")
            if s > best_s and lms <= lone_cap:
                best_s, cur = s, cand
    server.set_graph_params(gid, **cur)
    return {"params": cur, "qps": best_s, "lone_ms_start": lone0, "trail": trail}


def tune_batch(solver, sources, base: dict | None = None, log=None) -> dict:
    """Tune a batched launch (``Solver.solve_batch``) over the same sources:
    scored by batch wall time (the best of two timed launches after one
    untimed one)."""
    from ._lib import MLMQError
    from .solver import default_params

    cur = dict(base or {})
    cur.setdefault("delta", 200000)
    cur.setdefault("prefetch", 1)
    # the cached stale check: every batched workload measured so far ends at
    # it, and the shape coordinates below move with the stale-check mode
    cur.setdefault("stale_check", 1)

    def score(params):
        p = default_params(timeout_s=60.0, **params)
        best = float("inf")
        try:
            solver.solve_batch(sources, p)
            for _ in range(2):
                t0 = time.perf_counter()
                solver.solve_batch(sources, p)
                best = min(best, time.perf_counter() - t0)
        except MLMQError:
            return float("inf")
        return best

    best_t = score(cur)
    trail = [(dict(cur), best_t)]
    # Shape first, then the child-block touches and the stale-check mode
    # (with the whole device busy the touches can cost more memory-side
    # bandwidth than they save), then delta, the bucket count and the queue
    # sizes once more at the point reached: their best values move with the
    # shape.
    # The L1 ring (the LDS segment of each wave) and its spill threshold: a
    # 1024-record ring lets a wave keep up to 255 records before spilling to
    # L2 (a 512-record ring at most 64).
    space = [("batch_waves", [4, 8, 16]), ("l1_ring", [512, 1024]), ("batch_waves", [4, 8, 16]),
             ("delta", None),
             ("l1_pop", [16, 32, 64]), ("buckets", [2, 4, 8, 16, 32]), ("min_read", [8, 16, 32]),
             ("l1_ring", [512, 1024]), ("l1_spill", [64, 128, 255]),
             ("prefetch", [0, 1]), ("stale_check", [1, 3, 2]),
             ("delta", None), ("buckets", [8, 16, 32]),
             ("l1_pop", [16, 32, 64]), ("min_read", [8, 16, 32]),
             ("pop_read", [16, 32, 64])]
    cur.setdefault("l1_ring", 512)
    cur.setdefault("l1_spill", 64)
    for key, values in space:
        if values is None:           # delta: a grid around the current value
            values = _delta_grid(cur["delta"])
        for v in values:
            if key == "pop_read":    # pop size, top-up threshold and claim together
                if cur.get("l1_pop") == v and cur.get("min_read") == v:
                    continue
                cand = dict(cur, l1_pop=v, min_read=v, l2_batch=max(8, v))
            elif cur.get(key) == v:
                continue
            else:
                cand = dict(cur, **{key: v})
            if key == "buckets":
                cand["bucket_window"] = min(4, v)
            if key == "min_read":
                cand["l2_batch"] = max(8, v)
            if key == "l1_ring" and v == 512:
                cand["l1_spill"] = min(cand["l1_spill"], 64)
            if cand["l1_spill"] > cand["l1_ring"] - 448:
                continue             # a batch's pushes would not fit above the threshold
            t = score(cand)
            trail.append((cand, t))
            if log:
                log(f"tune batch: {cand} -> {t * 1e3:.1f} ms")
            if t < best_t:
                best_t, cur = t, cand
    return {"params": cur, "batch_s": best_t, "trail": trail}
