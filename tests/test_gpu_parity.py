"""Bit-exact parity of the HIP solver with the CPU oracle (needs an MI355X).

Every distance of every query must equal the oracle's (int32, INT_MAX for
unreachable).  Cases follow the reference's checks: the 2052-vertex fixture
with its three golden sources and the same-process source switch 0 -> 1030
-> 2050 -> 0 (scripts/multigpu/run_supplement_checks.py:27-53), plus
road-like, grid and power-law graphs, parameter corners and error paths.
"""
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O
from graphs_util import hub_csr, random_csr, grid_csr
from mlmq_tpds_withl3_amd import Solver, HostGraph, MLMQError, make_workload
from mlmq_tpds_withl3_amd import solver as S

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
TIMEOUT = 20.0   # device watchdog per solve in tests


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def solver():
    s = Solver(0)
    yield s
    s.close()


def solve_check(solver, row, col, w, src, load=True, **params):
    if load:
        solver.load_csr(row, col, w)
    params.setdefault("timeout_s", TIMEOUT)
    t = solver.solve(src, **params)
    got = solver.dist()
    ref, work, reached = O.dijkstra(row, col, w, src)
    bad = np.flatnonzero(got != ref)
    assert bad.size == 0, (f"{bad.size} mismatches, first v={bad[0]} gpu={got[bad[0]]} "
                           f"cpu={ref[bad[0]]} params={params}")
    return t, work, reached


def test_fixture_golden_and_source_switch(solver):
    say("fixture golden + source switch")
    solver.load_gr(os.path.join(GOLDEN, "fixture.gr"))
    for src in (0, 1030, 2050, 0):
        solver.solve(src, timeout_s=TIMEOUT)
        exp = np.fromfile(os.path.join(GOLDEN, f"s{src}.i32"), dtype=np.int32)
        assert np.array_equal(solver.dist(), exp), f"source {src}"


def test_shortcut_fixture_golden(solver):
    say("fixture with chain shortcuts (G+): golden distances")
    solver.load_gr(os.path.join(GOLDEN, "fixture_shortcut.gr"))
    for src in (0, 1030, 2050):
        for queue in ("L1SLF_L2DQ", "L1V_L2DQ"):
            solver.solve(src, timeout_s=TIMEOUT, queue=queue)
            exp = np.fromfile(os.path.join(GOLDEN, f"s{src}.i32"), dtype=np.int32)
            assert np.array_equal(solver.dist(), exp), (src, queue)


@pytest.mark.parametrize("queue", ["L1SLF_L2DQ", "L1N_L2DQ", "L1V_L2DQ"])
@pytest.mark.parametrize("seed", range(6))
def test_random_graphs(solver, seed, queue):
    say(f"random seed={seed} queue={queue}")
    n = 200 + 700 * seed
    row, col, w = random_csr(n, 4 * n, seed, wmax=1000, zero_frac=0.1)
    for src in (0, n - 1):
        for blocks in ((1, 0) if queue != "L1N_L2DQ" else (1,)):
            solve_check(solver, row, col, w, src, queue=queue, delta=300, l1_blocks=blocks)
        solve_check(solver, row, col, w, src, load=False, queue=queue, delta=300, light_split=1)


@pytest.mark.parametrize("cfg", [
    dict(),
    dict(delta=500),                                     # many buckets, window advance
    dict(delta=50, buckets=4, bucket_window=1),
    dict(buckets=1, bucket_window=1, l2_batch=1, min_read=1),
    dict(l2_batch=64, min_read=64, delta=2000),
    dict(work_blocks=1),                                 # a single workgroup
    dict(work_blocks=3, delta=1000),
    dict(blocks_per_cu=4, delta=5000),
    dict(queue="L1N_L2DQ", delta=2000),
    dict(queue="L1V_L2DQ", delta=2000),
    dict(queue="L1V_L2DQ", l1_spill=8, l1_pop=16),               # FIFO spilling often
    dict(stale_check=1, delta=2000),                             # stale filter through the L2
    dict(stale_check=0, delta=2000),                             # no stale filter
    dict(stale_check=3, delta=2000),                             # late stale filter
    dict(stale_check=3, queue="L1N_L2DQ", delta=500),
    dict(prefetch=0, delta=2000),
    dict(l1_blocks=1, delta=2000),                               # block-carrying L1
    dict(l1_blocks=1, queue="L1V_L2DQ"),
    dict(l1_blocks=1, stale_check=0, l1_pop=8, min_read=8),      # hot pops, no stale filter
    dict(defer_heavy=1, delta=300),
    dict(strict_window=1, delta=300),                           # phase barrier per bucket
    dict(light_split=1, delta=300),                             # light/heavy edge split
    dict(light_split=1, delta=1, buckets=64, bucket_window=2),  # only zero weights are light
    dict(light_split=1, delta=100, queue="L1N_L2DQ", strict_window=1),
    dict(strict_window=1, delta=100, bucket_window=1, queue="L1N_L2DQ"),
])
def test_grid_param_corners(solver, cfg):
    say(f"grid 120x150 {cfg}")
    row, col, w = grid_csr(120, 150, 7)
    solve_check(solver, row, col, w, 0, **cfg)


HUBS = [70000, 9000, 4097, 2049, 2048, 1500]   # degrees 70002 .. 1504: five hubs split


@pytest.mark.parametrize("cfg", [
    dict(),
    dict(hub_split=0),
    dict(queue="L1N_L2DQ"),
    dict(queue="L1V_L2DQ"),
    dict(work_blocks=1),                         # one workgroup takes every chunk itself
    dict(delta=100, l2_batch=64, min_read=64),
    dict(l1_spill=4, l1_pop=8),
    dict(l1_blocks=1),
    dict(defer_heavy=1, delta=50),                  # hubs expand only in the head bucket
    dict(defer_heavy=1, delta=50, buckets=4, bucket_window=4, queue="L1N_L2DQ"),
    dict(strict_window=1, delta=50, buckets=64, bucket_window=1),
    dict(light_split=1, delta=100, hub_split=0),
    dict(light_split=1, delta=20, buckets=64, bucket_window=4, hub_split=0),
    dict(light_split=1, delta=500, queue="L1V_L2DQ", hub_split=0),
    dict(stale_check=3),                            # late check skips stale chunks' edges
    dict(stale_check=3, light_split=1, delta=100, hub_split=0),
    # light/heavy split with hub splitting: a hub's heavy part goes out in chunks
    dict(light_split=1, delta=100),
    dict(light_split=1, delta=20, buckets=64, bucket_window=4),
    dict(light_split=1, delta=500, queue="L1V_L2DQ"),
    dict(light_split=1, stale_check=3, delta=100),
    dict(light_split=1, stale_check=1, delta=50, l1_pop=64, min_read=64, l2_batch=64, buckets=64,
         bucket_window=2),
])
def test_hub_split(solver, cfg):
    say(f"hubs {cfg}")
    row, col, w = hub_csr(30000, HUBS, 60000, seed=5)
    solver.load_csr(row, col, w)
    params = dict(hub_split=1)
    params.update(cfg)
    for src in (0, 3, 4, 29999):
        t, _, _ = solve_check(solver, row, col, w, src, load=False, **params)
        if params["hub_split"]:
            assert t["hub_chunks"] > 0     # every source reaches the hubs
        else:
            assert t["hub_chunks"] == 0


def test_hub_chunk_boundaries(solver):
    """Hub degrees at and around multiples of the 1024-edge chunk."""
    say("hub chunk boundaries")
    for degs in ([2049], [2050, 3072, 3073], [1024 * 7 + 1, 5000, 4096]):
        row, col, w = hub_csr(8000, degs, 16000, seed=len(degs))
        for src in (0, 7999):
            solve_check(solver, row, col, w, src, delta=500)


@pytest.mark.parametrize("cfg", [
    dict(),
    dict(queue="L1N_L2DQ"),
    dict(light_split=1, delta=300),
    dict(stale_check=3, delta=200),
])
def test_heavy_piece_boundaries(solver, cfg):
    """Out-degrees at and around multiples of 64: full 64-edge pieces are
    relaxed wave-wide, the remainders through the per-slot owner scan."""
    say(f"heavy piece boundaries {cfg}")
    degs = [8, 63, 64, 65, 127, 128, 129, 191, 192, 193, 640, 641, 1023, 4096, 4097] * 4
    row, col, w = hub_csr(20000, degs, 40000, seed=11)
    params = dict(delta=500)
    params.update(cfg)
    for src in (0, 17, 19999):
        solve_check(solver, row, col, w, src, **params)


def test_zero_weight_components(solver):
    say("zero weights")
    row, col, w = random_csr(5000, 15000, 99, wmax=3, zero_frac=0.6)
    for src in (1, 2500):
        solve_check(solver, row, col, w, src, delta=1)


def test_road264k(solver):
    say("road264k generate")
    g, src, spec = make_workload("road264k")
    row, col, w = g.arrays()
    for cfg in (dict(), dict(delta=20000), dict(queue="L1N_L2DQ"), dict(queue="L1V_L2DQ"),
                dict(stale_check=1), dict(stale_check=0), dict(stale_check=3), dict(l1_blocks=1),
                dict(light_split=1, delta=20000), dict(spec["bench_point"])):
        say(f"road264k {cfg}")
        t, work, reached = solve_check(solver, row, col, w, src, **cfg)
        assert t["relaxations"] >= work


def test_rmat_small(solver):
    say("rmat scale 16")
    g = HostGraph.rmat(16, 16, seed=5)
    row, col, w = g.arrays()
    src = g.max_degree_vertex()
    for cfg in (dict(), dict(delta=100)):
        solve_check(solver, row, col, w, src, **cfg)


@pytest.mark.parametrize("pop", [32, 16, 64])
def test_lane_pairs_and_single_lanes(solver, pop):
    """Pops of up to 32 records run two lanes per record (each relaxes half
    of the hop block, the even lane owns CSR edges, stale checks and
    heavy-phase records); 64 runs one lane per record.  Both are exact on
    graphs with degrees across the 7-edge block boundary, light/heavy
    splits (inline heavy edges on either lane of a pair), hubs, zero
    weights, and with top-up thresholds above 32."""
    say(f"lane pairs / single lanes, pop {pop}")
    cases = [
        (random_csr(3000, 20000, 21, wmax=1000, zero_frac=0.1), 0),
        (grid_csr(60, 70, 22), 5),
        (hub_csr(6000, [5000, 3000, 900], 12000, seed=23, wmax=1000), 1),
    ]
    g = HostGraph.rmat(15, 12, seed=24)
    rmat = tuple(x.copy() for x in g.arrays())
    cases.append((rmat, g.max_degree_vertex()))
    g.close()
    for (row, col, w), src in cases:
        solver.load_csr(row, col, w)
        for cfg in (dict(), dict(light_split=1, delta=100, buckets=32, bucket_window=2),
                    dict(light_split=1, delta=300, stale_check=3),
                    dict(stale_check=1, min_read=64, l2_batch=64, delta=500),
                    dict(queue="L1V_L2DQ", light_split=1, delta=200),
                    dict(hub_split=1, delta=1000, prefetch=0)):
            solve_check(solver, row, col, w, src, load=False, l1_pop=pop, **cfg)


@pytest.mark.parametrize("case", ["rmat", "hub"])
def test_lane_pair_claims_wider_than_a_batch(solver, case):
    """Lane pairs (a batch holds 32 records) with L2 claims of 63 and 39 slots,
    512 work blocks and delta 1 on small graphs, so the bucket rings lap many
    times: a new claim is cut to the 32 slots a batch holds, so an idle wave
    tries every slot it holds and no record is left under an untried slot
    (a wider claim can strand records and spin to the watchdog).  Exact on
    repeated solves from several sources."""
    say(f"lane-pair claims wider than a batch, {case}")
    if case == "rmat":
        g = HostGraph.rmat(12, 14, seed=940935275)
        row, col, w = (x.copy() for x in g.arrays())
        g.close()
        srcs = (3328, 0, 17)
        cfg = dict(delta=1, buckets=4, bucket_window=3, l2_batch=63, min_read=31, l1_pop=16,
                   stale_check=2, l1_spill=16, idle_backoff=1)
    else:
        row, col, w = hub_csr(5000, [4000, 1500], 10000, seed=616, wmax=1000)
        srcs = (1702, 1, 4999)
        cfg = dict(delta=1, buckets=2, bucket_window=1, l2_batch=39, min_read=10, l1_pop=25,
                   stale_check=3, light_split=1, l1_spill=64, idle_backoff=16)
    solver.load_csr(row, col, w)
    for rep in range(3):
        for src in srcs:
            solve_check(solver, row, col, w, src, load=False, work_blocks=512, prefetch=1, **cfg)


def back_edge_graph(n, seed):
    """Degrees 0-12 (inline and CSR vertices side by side): two-way edges with
    equal or different weights in each direction, parallel edges, one-way
    edges, self-loops and zero weights."""
    rng = np.random.default_rng(seed)
    src, dst, w = [], [], []
    for u in range(n):
        for _ in range(int(rng.integers(0, 4 if u % 97 else 9))):
            v = int(rng.integers(0, n))
            a, b = int(rng.integers(0, 40)), int(rng.integers(0, 40))
            kind = int(rng.integers(0, 5))
            src.append(u), dst.append(v), w.append(a)
            if kind in (1, 2, 3):   # the reverse edge, same or other weight
                src.append(v), dst.append(u), w.append(a if kind == 1 else b)
            if kind == 3:           # a parallel edge
                src.append(u), dst.append(v), w.append(b)
        if u % 53 == 0:
            src.append(u), dst.append(u), w.append(int(rng.integers(0, 5)))
    src, dst, w = np.array(src), np.array(dst), np.array(w)
    order = np.lexsort((dst, src))
    src, dst, w = src[order], dst[order], w[order]
    row = np.zeros(n + 1, dtype=np.int64)
    np.add.at(row, src + 1, 1)
    return np.cumsum(row), dst.astype(np.int32), w.astype(np.int32)


@pytest.mark.parametrize("pop", [32, 64])
def test_back_edges(solver, pop):
    """A record skips its vertex's edge back to the vertex whose expansion
    created it (it cannot lower that distance).  Exact on a multigraph with
    one-way, parallel and zero-weight edges and self-loops, in every phase of
    the light/heavy split and with the late stale check; on a symmetric grid
    the skip shows as fewer atomics than examined edges; a weight of 2^28 or
    more leaves the blocks unpacked and every examined edge is relaxed, in
    single and batched solves."""
    say(f"back edges, pop {pop}")
    row, col, w = back_edge_graph(4000, seed=31)
    solver.load_csr(row, col, w)
    for src in (0, 1, 2000):
        for cfg in (dict(), dict(light_split=1, delta=10, buckets=32, bucket_window=2),
                    dict(stale_check=3, delta=30), dict(queue="L1V_L2DQ", delta=20),
                    dict(stale_check=1, prefetch=0)):
            t, work, _ = solve_check(solver, row, col, w, src, load=False, l1_pop=pop, **cfg)
            assert t["relaxations"] >= work
    row, col, w = grid_csr(40, 50, 32)
    t, work, _ = solve_check(solver, row, col, w, 7, l1_pop=pop)
    assert t["relaxations"] >= work and t["atomics"] < t["relaxations"]
    w2 = w.copy()
    w2[-1] = 1 << 29
    t, work, _ = solve_check(solver, row, col, w2, 7, l1_pop=pop)
    assert t["atomics"] == t["relaxations"]
    sources = [7, 1999, 7, 1000]
    solver.solve_batch(sources, timeout_s=TIMEOUT, l1_pop=pop, work_blocks=2)
    for i, src in enumerate(sources):
        assert np.array_equal(solver.batch_dist(i), O.dijkstra(row, col, w2, src)[0]), f"query {i}"


@pytest.mark.parametrize("cfg", [dict(), dict(l1_pop=64), dict(light_split=1, delta=100),
                                 dict(light_split=1, delta=100, l1_pop=64),
                                 dict(queue="L1N_L2DQ"), dict(l1_blocks=1)])
def test_profiling_build(solver, cfg):
    """profile=1 runs the per-phase clock instance of the kernel: the same
    distances, and the phase clocks are filled."""
    say(f"profiling build {cfg}")
    row, col, w = hub_csr(5000, [3000, 900], 15000, seed=41, wmax=1000)
    solver.load_csr(row, col, w)
    t, _, _ = solve_check(solver, row, col, w, 1, load=False, profile=1, **cfg)
    ph = t["phase_cycles"]
    assert ph["total"] > 0 and ph["l1_pop" if cfg.get("queue") != "L1N_L2DQ" else "l2_pop"] >= 0
    assert sum(v for k, v in ph.items() if k != "total") > 0


@pytest.mark.parametrize("split", [False, True])
def test_bfs_renumbered_graph(solver, split):
    say(f"BFS-renumbered road-like graph, layer split {split}")
    g = HostGraph.knn(200_000, 3, seed=7, weight_scale=1e5)
    row, col, w = g.arrays()
    ref, _, _ = O.dijkstra(row, col, w, 0)
    for root in (0, 123_457):
        r, perm, inv = g.reorder_bfs(root, split)
        nrow, ncol, nw = r.arrays()
        src = int(inv[0])
        solve_check(solver, nrow, ncol, nw, src)
        assert np.array_equal(solver.dist(), ref[perm])


def test_road6m_full(solver):
    say("road6m generate")
    g, src, spec = make_workload("road6m")
    row, col, w = g.arrays()
    say(f"road6m solve n={g.n} m={g.m} src={src}")
    t, work, reached = solve_check(solver, row, col, w, src)
    say(f"road6m solve_ms={t['solve_ms']:.3f}")
    # a symmetrised 3-NN point graph has small separate components
    assert 0.9 * g.n < reached < g.n


def test_reference_driver_abi():
    from hip_rt import DeviceArray
    say("reference ABI sssp_init/run/cp_data")
    row, col, w = O.read_gr(os.path.join(GOLDEN, "fixture.gr"))
    arrays = [DeviceArray(row.astype(np.int32)), DeviceArray(col), DeviceArray(w)]
    try:
        S.sssp_init(len(row) - 1, len(col), *(a.address for a in arrays))
        for src in (1030, 0):
            S.sssp_run_adaptive(src)
            got = S.sssp_cp_data()
            assert np.array_equal(got, np.fromfile(os.path.join(GOLDEN, f"s{src}.i32"), np.int32))
        S.sssp_release()
    finally:
        for a in arrays:
            a.free()


def test_attach_device_csr(solver):
    from hip_rt import DeviceArray
    say("attach device CSR")
    row, col, w = grid_csr(60, 70, 11)
    arrays = [DeviceArray(row.astype(np.int32)), DeviceArray(col), DeviceArray(w)]
    try:
        solver.attach_device_csr(len(row) - 1, len(col), *(a.address for a in arrays))
        solver.solve(5, timeout_s=TIMEOUT)
        ref, _, _ = O.dijkstra(row, col, w, 5)
        assert np.array_equal(solver.dist(), ref)
        gi = solver.graph_info()
        assert gi.m == len(row) - 1 and gi.nnz == len(col)
    finally:
        solver.load_csr(*grid_csr(2, 2, 1))   # detach before freeing the arrays
        for a in arrays:
            a.free()


def test_errors_and_recovery(solver):
    say("error paths")
    row, col, w = grid_csr(200, 200, 3)
    solver.load_csr(row, col, w)
    with pytest.raises(MLMQError) as e:
        solver.solve(-1)
    assert e.value.code == -1
    with pytest.raises(MLMQError) as e:
        solver.solve(0, buckets=0)
    assert e.value.code == -1
    with pytest.raises(MLMQError) as e:
        solver.solve(0, stale_check=4)
    assert e.value.code == -1
    with pytest.raises(MLMQError) as e:
        solver.solve(0, queue_mode=3)   # L1V_L2V: not on this path
    assert e.value.code == -9
    with pytest.raises(MLMQError) as e:
        solver.solve(0, min_read=0)   # would leave the seed untaken: refused before launch
    assert e.value.code == -1
    # a good solve, then one stopped by the watchdog (forced from the host:
    # every wave sees the deadline passed and exits): the failed solve's
    # partial distances must not be handed out as a result
    solve_check(solver, row, col, w, 0, load=False)
    with pytest.raises(MLMQError) as e:
        solver.solve(0, timeout_s=1e-7, delta=100)
    assert e.value.code == -7
    with pytest.raises(MLMQError) as e:
        solver.dist()
    assert e.value.code == -2
    # the next query still runs on a clean workspace
    solve_check(solver, row, col, w, 0, load=False)
    # the same for a batch: a good batch, a stopped one, no rows to copy
    solver.solve_batch(np.array([0, 7, 99], dtype=np.int32), timeout_s=TIMEOUT)
    with pytest.raises(MLMQError) as e:
        solver.solve_batch(np.array([0, 7, 99], dtype=np.int32), timeout_s=1e-7, delta=100)
    assert e.value.code == -7
    with pytest.raises(MLMQError) as e:
        solver.batch_dist(0)
    assert e.value.code == -2
    solver.solve_batch(np.array([5, 6], dtype=np.int32), timeout_s=TIMEOUT)
    ref, _, _ = O.dijkstra(row, col, w, 6)
    assert np.array_equal(solver.batch_dist(1), ref)


@pytest.mark.parametrize("queue,kw", [("L1SLF_L2DQ", {}), ("L1V_L2DQ", dict(buckets=4, bucket_window=2)),
                                      ("L1N_L2DQ", dict(bucket_capacity=64))])
def test_queue_counts_balance(queue, kw):
    """Per-bucket reservations equal reads after a solve, their sum equals the
    credits returned and the pushes counted by the kernel; a ring smaller than
    a bucket's traffic (64 slots, L2-only queue) reuses its slots without
    losing a record."""
    print(f"[gpu-test] queue counts {queue} {kw}", flush=True)
    row, col, w = grid_csr(40, 50, seed=17)
    ref, _, _ = O.dijkstra(row, col, w, 0)
    with Solver(0) as s:
        s.load_csr(row, col, w)
        t = s.solve(0, queue=queue, delta=2000, timeout_s=20.0, **kw)
        assert np.array_equal(s.dist(), ref)
        q = s.queue_counts()
        assert q["buckets"] == kw.get("buckets", 16)
        assert q["writes"] == q["reads"]
        assert sum(q["writes"]) == q["completed"] == t["l2_pushes"]
        assert q["capacity"] == t["bucket_capacity"]
        assert q["wrapped"] == sum(x > q["capacity"] for x in q["writes"])
