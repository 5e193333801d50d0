"""Bit-exact parity of the batched (many queries, one persistent launch) solver.

Every query of every batch must equal the CPU oracle.  Cases: the golden
fixture sources, partitions reused across many queries (fewer workgroups than
queries), every supported workgroup size, hipGraph replays interleaved with
eager launches and single solves, a road-like graph, and error paths.
"""
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O
from graphs_util import random_csr, grid_csr, hub_csr
from mlmq_tpds_withl3_amd import Solver, MLMQError, make_workload

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
TIMEOUT = 20.0


def say(msg):
    print(f"[gpu-batch] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def solver():
    s = Solver(0)
    yield s
    s.close()


def check_batch(solver, row, col, w, sources, **params):
    params.setdefault("timeout_s", TIMEOUT)
    t = solver.solve_batch(sources, **params)
    cache = {}
    for i, src in enumerate(sources):
        if src not in cache:
            cache[src] = O.dijkstra(row, col, w, int(src))[0]
        got = solver.batch_dist(i)
        bad = np.flatnonzero(got != cache[src])
        assert bad.size == 0, f"query {i} (src {src}): {bad.size} mismatches {params}"
    return t


def test_fixture_golden_batch(solver):
    say("fixture golden batch")
    solver.load_gr(os.path.join(GOLDEN, "fixture.gr"))
    sources = [0, 1030, 2050, 0, 1030, 2050, 0]
    t = solver.solve_batch(sources, timeout_s=TIMEOUT, work_blocks=2)
    for i, src in enumerate(sources):
        exp = np.fromfile(os.path.join(GOLDEN, f"s{src}.i32"), dtype=np.int32)
        assert np.array_equal(solver.batch_dist(i), exp), f"query {i}"
    assert t["queries"] == len(sources) and t["work_blocks"] == 2


@pytest.mark.parametrize("waves", [4, 8, 16])
@pytest.mark.parametrize("blocks", [1, 5, 0])
def test_random_batch(solver, waves, blocks):
    say(f"random batch waves={waves} blocks={blocks}")
    row, col, w = random_csr(3000, 12000, 17, wmax=500, zero_frac=0.1)
    solver.load_csr(row, col, w)
    rng = np.random.default_rng(waves * 10 + blocks)
    sources = rng.integers(0, 3000, size=40)
    check_batch(solver, row, col, w, sources, batch_waves=waves, work_blocks=blocks, delta=200)


@pytest.mark.parametrize("split", [1, 0])
def test_hub_batch(solver, split):
    say(f"hub batch split={split}")
    row, col, w = hub_csr(20000, [40000, 5000, 2100], 40000, seed=8)
    solver.load_csr(row, col, w)
    sources = [0, 1, 2, 19999, 0, 7]
    check_batch(solver, row, col, w, sources, work_blocks=3, hub_split=split, delta=300)


@pytest.mark.parametrize("stale", [0, 1, 3])
def test_stale_modes_batch(solver, stale):
    say(f"stale_check={stale} batch")
    row, col, w = grid_csr(70, 80, 5)
    solver.load_csr(row, col, w)
    check_batch(solver, row, col, w, [0, 5599, 123, 4000], work_blocks=2, stale_check=stale)


def test_light_split_batch(solver):
    say("light split batch")
    row, col, w = hub_csr(20000, [40000, 5000, 2100], 40000, seed=10)
    solver.load_csr(row, col, w)
    check_batch(solver, row, col, w, [0, 1, 19999, 77, 5], work_blocks=2, light_split=1, delta=100,
                buckets=32)


@pytest.mark.parametrize("pop", [64, 32])
@pytest.mark.parametrize("waves", [4, 16])
def test_batch_lane_pairs_and_single_lanes(solver, pop, waves):
    """Pops of up to 32 records run two lanes per record, 64 one lane per
    record; both exact, with and without the light/heavy split, with top-up
    thresholds above 32."""
    say(f"batch pop {pop} waves {waves}")
    row, col, w = random_csr(6000, 40000, 31, wmax=1000, zero_frac=0.1)
    solver.load_csr(row, col, w)
    srcs = [0, 17, 5999, 3000, 17]
    check_batch(solver, row, col, w, srcs, batch_waves=waves, l1_pop=pop, delta=300)
    check_batch(solver, row, col, w, srcs, batch_waves=waves, l1_pop=pop, delta=100,
                light_split=1, buckets=32, min_read=64, l2_batch=64)


def test_strict_window_batch(solver):
    say("strict window batch")
    row, col, w = hub_csr(20000, [40000, 5000, 2100], 40000, seed=9)
    solver.load_csr(row, col, w)
    check_batch(solver, row, col, w, [0, 1, 19999, 77], work_blocks=3, strict_window=1, delta=100,
                buckets=32, bucket_window=1)


@pytest.mark.parametrize("waves", [4, 8, 16])
@pytest.mark.parametrize("ring,spill", [(1024, 255), (1024, 128), (1024, 64), (512, 16)])
def test_l1_ring_sizes(solver, waves, ring, spill):
    """Both L1 ring sizes (the LDS segment per wave) at every workgroup size,
    with spill thresholds up to the largest the ring allows: road-like grid,
    hub graph and the golden fixture, every row exact."""
    say(f"l1 ring {ring} spill {spill} waves {waves}")
    row, col, w = grid_csr(60, 70, seed=41)
    solver.load_csr(row, col, w)
    check_batch(solver, row, col, w, [0, 17, 4199, 2100, 999, 0], batch_waves=waves, l1_ring=ring,
                l1_spill=spill, delta=3000)
    row, col, w = hub_csr(3000, [2500, 1800], 12000, seed=43)
    solver.load_csr(row, col, w)
    check_batch(solver, row, col, w, [0, 1, 5, 2999], batch_waves=waves, l1_ring=ring,
                l1_spill=spill, delta=400)
    solver.load_gr(os.path.join(GOLDEN, "fixture.gr"))
    sources = [0, 1030, 2050, 1030]
    solver.solve_batch(sources, timeout_s=TIMEOUT, batch_waves=waves, l1_ring=ring, l1_spill=spill)
    for i, src in enumerate(sources):
        exp = np.fromfile(os.path.join(GOLDEN, f"s{src}.i32"), dtype=np.int32)
        assert np.array_equal(solver.batch_dist(i), exp), f"query {i}"


def test_graph_replays_interleaved(solver):
    say("graph replays with eager work in between")
    row, col, w = grid_csr(80, 90, 21)
    solver.load_csr(row, col, w)
    n = len(row) - 1
    rng = np.random.default_rng(3)
    sources = rng.integers(0, n, size=24)
    eager = check_batch(solver, row, col, w, sources, use_graph=0, work_blocks=6)
    assert eager["graph_replay"] == 0
    for rep in range(3):
        t = check_batch(solver, row, col, w, sources, use_graph=1, work_blocks=6)
        assert t["graph_replay"] == 1
        # the program's own eager work and a different query between replays
        solver.solve(int(sources[rep]), timeout_s=TIMEOUT)
        ref, _, _ = O.dijkstra(row, col, w, int(sources[rep]))
        assert np.array_equal(solver.dist(), ref)
    # a new source list reuses the captured graph (sources are copied in first)
    check_batch(solver, row, col, w, rng.integers(0, n, size=24), use_graph=1, work_blocks=6)


def test_road264k_batch(solver):
    say("road264k batch of 48")
    g, _, _ = make_workload("road264k")
    row, col, w = g.arrays()
    solver.load_csr(row, col, w)
    sources = np.random.default_rng(7).integers(0, g.n, size=48)
    t = check_batch(solver, row, col, w, sources)
    say(f"road264k batch solve_ms={t['solve_ms']:.2f} p99={t['query_ms_p99']:.2f}")


def test_batch_errors(solver):
    say("batch errors")
    row, col, w = grid_csr(100, 100, 5)
    solver.load_csr(row, col, w)
    with pytest.raises(MLMQError) as e:
        solver.solve_batch([0, 10_000])
    assert e.value.code == -1
    with pytest.raises(MLMQError) as e:
        solver.solve_batch([0, 1], batch_waves=5)
    assert e.value.code == -1
    with pytest.raises(MLMQError) as e:
        solver.solve_batch(list(range(16)), timeout_s=1e-7, delta=10)
    assert e.value.code == -7
    check_batch(solver, row, col, w, list(range(0, 10_000, 700)))


@pytest.mark.parametrize("waves", [4, 8, 16])
def test_batch_without_counters_fills_the_device(solver, waves):
    """batch_stats 0: the instance without statistics counters (more waves
    per CU) gives the same rows; work_blocks 0 runs as many workgroups as
    the device holds, at most one per query."""
    say(f"batch without counters, {waves} waves")
    row, col, w = grid_csr(50, 60, seed=51)
    solver.load_csr(row, col, w)
    rng = np.random.default_rng(52)
    sources = [int(x) for x in rng.integers(0, 3000, size=700)]
    t = check_batch(solver, row, col, w, sources, batch_waves=waves, batch_stats=0, delta=3000)
    assert t["relaxations"] == 0 and t["improvements"] == 0
    t1 = check_batch(solver, row, col, w, sources[:40], batch_waves=waves, batch_stats=1, delta=3000)
    assert t1["relaxations"] > 0 and t1["work_blocks"] == 40
    say(f"{waves} waves: {t['work_blocks']} workgroups for {len(sources)} queries")
    assert t["work_blocks"] >= 256
