"""Seeded randomized parameter sweep on the GPU: random, grid, hub, RMAT and
unpacked-weight graphs of up to 8,000 vertices, every field of the solve and
batched-solve parameters drawn within its documented range (queue modes,
buckets and window, claims of 1..64 with and without lane pairs, top-up and
pop sizes, stale-check modes, child-block touches, light/heavy and hub
splits, L1 spill, grid sizes up to 512 workgroups), each result compared
bit-exactly with the oracle; and served requests over random resident graphs
with random per-graph parameters and server settings, in both serving modes.
A solve may end with a queue overflow (reported, then skipped) but never with
a watchdog expiry or a wrong distance: sweeps of this kind found the
lapped-claim stall of lane-pair waves (claims wider than a batch), which this
test keeps covered."""
import sys

import numpy as np
import pytest

from graphs_util import grid_csr, hub_csr, random_csr
from mlmq_tpds_withl3_amd import HostGraph, MLMQError, Solver
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TIMEOUT = 20.0
SKIP_CODES = (-1, -9)   # invalid combination, queue overflow after the retries


def say(msg):
    print(f"[gpu-sweep] {msg}", flush=True)
    sys.stdout.flush()


def draw_graph(rng):
    kind = int(rng.integers(0, 5))
    seed = int(rng.integers(0, 1 << 30))
    if kind == 0:
        n = int(rng.integers(50, 4000))
        g = random_csr(n, int(n * rng.integers(1, 8)), seed, wmax=int(rng.choice([1, 10, 1000, 100000])),
                       zero_frac=float(rng.choice([0, 0.1, 0.5])))
    elif kind == 1:
        g = grid_csr(int(rng.integers(5, 70)), int(rng.integers(5, 70)), seed,
                     wmax=int(rng.choice([10, 1000])))
    elif kind == 2:
        n = int(rng.integers(3000, 8000))
        g = hub_csr(n, [int(rng.integers(2100, 6000)), int(rng.integers(10, 3000))], 2 * n, seed=seed,
                    wmax=1000)
    elif kind == 3:
        h = HostGraph.rmat(int(rng.integers(8, 13)), int(rng.integers(2, 16)), seed=seed)
        g = tuple(x.copy() for x in h.arrays())
        h.close()
    else:
        n = int(rng.integers(50, 3000))
        row, col, w = random_csr(n, 3 * n, seed, wmax=100)
        w = w.copy()
        w[rng.integers(0, len(w))] = int(rng.choice([1 << 28, (1 << 29) + 5]))   # unpacked blocks
        g = (row, col, w)
    return kind, seed, g


def draw_solve_params(rng):
    p = dict(delta=int(rng.choice([1, 5, 50, 300, 5000, 200000])), buckets=int(rng.choice([1, 2, 4, 16, 64])),
             l2_batch=int(rng.integers(1, 65)), min_read=int(rng.integers(1, 65)),
             l1_pop=int(rng.integers(1, 65)), stale_check=int(rng.integers(0, 4)),
             prefetch=int(rng.integers(0, 2)), work_blocks=int(rng.choice([0, 1, 3, 64, 256, 512])),
             queue=str(rng.choice(["L1SLF_L2DQ", "L1SLF_L2DQ", "L1V_L2DQ", "L1N_L2DQ"])),
             light_split=int(rng.integers(0, 2)), hub_split=int(rng.integers(0, 2)),
             l1_spill=int(rng.choice([16, 64])), idle_backoff=int(rng.choice([1, 16, 64])),
             timeout_s=TIMEOUT)
    p["bucket_window"] = int(rng.integers(1, p["buckets"] + 1))
    if rng.random() < 0.1 and not p["light_split"]:
        p["l1_blocks"] = 1
    return p


def draw_batch_params(rng):
    p = dict(delta=int(rng.choice([5, 50, 300, 5000, 200000])), buckets=int(rng.choice([1, 2, 4, 16])),
             l2_batch=int(rng.integers(1, 65)), min_read=int(rng.integers(1, 65)),
             l1_pop=int(rng.integers(1, 65)), stale_check=int(rng.integers(0, 4)),
             prefetch=int(rng.integers(0, 2)), batch_waves=int(rng.choice([4, 8, 16])),
             l1_ring=int(rng.choice([512, 1024])), batch_stats=int(rng.integers(0, 2)),
             use_graph=int(rng.integers(0, 2)), light_split=int(rng.integers(0, 2)),
             work_blocks=int(rng.choice([0, 1, 2])), timeout_s=TIMEOUT)
    p["bucket_window"] = int(rng.integers(1, p["buckets"] + 1))
    return p


def draw_serve_params(rng):
    b = int(rng.choice([1, 2, 4, 16, 64]))
    return dict(delta=int(rng.choice([5, 50, 300, 5000, 200000])), buckets=b,
                bucket_window=int(rng.integers(1, b + 1)), l2_batch=int(rng.integers(1, 65)),
                min_read=int(rng.integers(1, 65)), l1_pop=int(rng.integers(1, 65)),
                l1_spill=int(rng.integers(1, 65)), workgroups=int(rng.integers(1, 9)),
                waves=int(rng.choice([1, 4, 8, 16])), hub_split=int(rng.integers(0, 2)),
                light_split=int(rng.integers(0, 2)), prefetch=int(rng.choice([-1, 1])),
                stale_check=int(rng.choice([-1, 1, 2, 3])))


@pytest.mark.parametrize("seed", [51, 52])
def test_randomized_serving_sweep(seed):
    """Served requests over random resident graphs, per-graph parameters drawn
    within their documented ranges, launch lanes or the live launch, random
    launch sizes and spreading: every request completes and is exact."""
    from mlmq_tpds_withl3_amd.serve import Server
    rng = np.random.default_rng(seed)
    for rnd in range(16):
        mode = str(rng.choice(["lanes", "live"]))
        cfg = dict(hbm_budget=512 << 20, streams=int(rng.integers(1, 5)), max_batch=int(rng.integers(1, 33)),
                   spread_waves=int(rng.choice([0, 64, 512])), timeout_s=TIMEOUT, mode=mode)
        graphs, params = [], []
        s = Server(0, **cfg)
        try:
            for _ in range(3):
                kind, gseed, g = draw_graph(rng)
                p = draw_serve_params(rng)
                gid = s.add_graph(g, **p)
                graphs.append((kind, gseed, g))
                params.append(p)
                assert gid == len(graphs) - 1
            nreq = 24
            gsel = rng.integers(0, 3, size=nreq).astype(np.int32)
            src = np.array([int(rng.integers(0, len(graphs[gi][2][0]) - 1)) for gi in gsel], dtype=np.int32)
            arrival = np.sort(rng.uniform(0.0, 0.005, size=nreq))
            outs = [np.empty(len(graphs[gi][2][0]) - 1, dtype=np.int32) for gi in gsel]
            res, _ = s.run(arrival, gsel, src, dist_out=outs)
        finally:
            s.close()
        for i, (gi, sv) in enumerate(zip(gsel, src)):
            kind, gseed, (row, col, w) = graphs[gi]
            case = f"round {rnd} {cfg} request {i}: graph kind {kind} seed {gseed} src {sv} {params[gi]}"
            assert res["status"][i] == 0, f"{case}: status {res['status'][i]}"
            ref = O.dijkstra(row, col, w, int(sv))[0]
            bad = np.flatnonzero(outs[i] != ref)
            assert bad.size == 0, f"{case}: {bad.size} mismatches"
        say(f"serving seed {seed} round {rnd} ({mode}): {nreq} requests exact")


@pytest.mark.parametrize("seed", [41, 42])
def test_randomized_parameter_sweep(seed):
    rng = np.random.default_rng(seed)
    s = Solver(0)
    ok = skipped = 0
    try:
        for it in range(500):
            kind, gseed, (row, col, w) = draw_graph(rng)
            n = len(row) - 1
            s.load_csr(row, col, w)
            for _ in range(3):
                src = int(rng.integers(0, n))
                p = draw_solve_params(rng)
                case = f"it {it} graph kind {kind} seed {gseed} n {n} src {src} {p}"
                try:
                    s.solve(src, **p)
                except MLMQError as e:
                    assert e.code in SKIP_CODES, f"{case}: {e}"
                    skipped += 1
                    continue
                ref = O.dijkstra(row, col, w, src)[0]
                got = s.dist()
                bad = np.flatnonzero(got != ref)
                assert bad.size == 0, f"{case}: {bad.size} mismatches, v={bad[0]} gpu={got[bad[0]]} cpu={ref[bad[0]]}"
                ok += 1
            if rng.random() < 0.5:
                srcs = [int(x) for x in rng.integers(0, n, size=int(rng.integers(1, 9)))]
                bp = draw_batch_params(rng)
                case = f"it {it} graph kind {kind} seed {gseed} n {n} batch {srcs} {bp}"
                try:
                    s.solve_batch(srcs, **bp)
                except MLMQError as e:
                    assert e.code in SKIP_CODES, f"{case}: {e}"
                    skipped += 1
                    continue
                for i, src in enumerate(srcs):
                    ref = O.dijkstra(row, col, w, src)[0]
                    assert np.array_equal(s.batch_dist(i), ref), f"{case}: query {i}"
                ok += 1
            if it % 100 == 99:
                say(f"seed {seed}: {it + 1} graphs, {ok} exact, {skipped} skipped")
    finally:
        s.close()
    assert ok > 10 * max(skipped, 1), (ok, skipped)
