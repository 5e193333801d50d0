"""Bit-exact parity of the serving loop (resident graphs, admission by bytes,
continuous batching over concurrent launch lanes).

Every served result is compared with the CPU oracle.  Cases: each request
alone versus the same requests mixed over three graphs and four lanes, device
verification against registered rows (including a planted wrong row),
admission under a byte budget that fits only a few slots, rejection when too
many requests wait, queue-overflow reruns, several workgroups per request and
every workgroup size, a slot that can never fit, and argument errors.
"""
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O
from graphs_util import random_csr, grid_csr, hub_csr
from mlmq_tpds_withl3_amd import HostGraph, MLMQError
from mlmq_tpds_withl3_amd.serve import Server, poisson_trace

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
TIMEOUT = 20.0
MiB = 1 << 20


def say(msg):
    print(f"[gpu-serve] {msg}", flush=True)
    sys.stdout.flush()


class G:
    """A graph with its CSR and oracle rows (cached per source)."""

    def __init__(self, row, col, w):
        self.row, self.col, self.w = row, col, w
        self.n = len(row) - 1
        self._rows = {}

    def expect(self, src):
        if src not in self._rows:
            self._rows[src] = O.dijkstra(self.row, self.col, self.w, int(src))[0]
        return self._rows[src]


@pytest.fixture(scope="module")
def graphs():
    a = G(*random_csr(3000, 24000, seed=11, wmax=1000))
    b = G(*grid_csr(60, 80, seed=12))
    hg = HostGraph.knn(20000, 3, seed=13, weight_scale=1e5)
    c = G(*(x.copy() for x in hg.arrays()))   # the views die with hg
    hg.close()
    return [a, b, c]


def new_server(graphs, budget=256 * MiB, **cfg):
    cfg.setdefault("timeout_s", TIMEOUT)
    s = Server(0, hbm_budget=budget, **cfg)
    ids = []
    for g in graphs:
        ids.append(s.add_graph((g.row, g.col, g.w), delta=2000 if g.n != 20000 else 200000))
    return s, ids


def run_checked(server, graphs, arrival, gsel, src, label):
    outs = [np.empty(graphs[g].n, dtype=np.int32) for g in gsel]
    res, st = server.run(arrival, gsel, src, dist_out=outs)
    for i, (g, s) in enumerate(zip(gsel, src)):
        assert res["status"][i] == 0, f"{label}: request {i} status {res['status'][i]}"
        bad = np.flatnonzero(outs[i] != graphs[g].expect(int(s)))
        assert bad.size == 0, f"{label}: request {i} (graph {g}, src {s}): {bad.size} mismatches"
    return res, st, outs


def test_alone_equals_mix(graphs):
    say("alone vs mix")
    s, ids = new_server(graphs, streams=4, max_batch=8)
    rng = np.random.default_rng(5)
    k = 36
    gsel = rng.integers(0, 3, size=k).astype(np.int32)
    src = np.array([rng.integers(0, graphs[g].n) for g in gsel], dtype=np.int32)
    alone = []
    for i in range(k):
        _, st, outs = run_checked(s, graphs, [0.0], [gsel[i]], [src[i]], f"alone {i}")
        assert st["completed"] == 1 and st["launches"] == 1
        alone.append(outs[0])
    arrival = np.sort(rng.uniform(0, 0.02, size=k))
    res, st, mixed = run_checked(s, graphs, arrival, gsel, src, "mix")
    assert st["completed"] == k and st["failed"] == 0 and st["rejected"] == 0
    for i in range(k):
        assert np.array_equal(alone[i], mixed[i])
    # the mix overlapped: more than one request in flight at once
    assert st["peak_inflight"] > 1
    say(f"mix stats {st}")
    s.close()


@pytest.fixture(scope="module")
def mixed_graphs():
    """road264k, a 300 x 400 grid and RMAT-16, each with its own settings."""
    road = HostGraph.knn(264_346, 3, seed=1, weight_scale=1e5)
    rmat = HostGraph.rmat(16, 16, seed=9, wmin=1, wmax=1000)
    gs = [G(*(x.copy() for x in road.arrays())), G(*grid_csr(300, 400, seed=21)),
          G(*(x.copy() for x in rmat.arrays()))]
    road.close()
    rmat.close()
    # rings large enough that no request overflows (an overflow reruns the
    # request in a later launch)
    params = [dict(delta=200_000, bucket_capacity=1 << 16),
              dict(delta=4000, buckets=4, bucket_window=2, bucket_capacity=1 << 16),
              dict(delta=50, light_split=1, buckets=64, bucket_window=2, l2_batch=64,
                   min_read=64, l1_pop=64, bucket_capacity=1 << 16)]
    return gs, params


@pytest.mark.parametrize("spread", [0, 512])
def test_mixed_graph_launch(mixed_graphs, spread):
    """One launch holds requests of three graphs of different sizes and
    settings (light/heavy split on the RMAT graph only): every row equals the
    oracle's and the row of the same request served alone."""
    gs, params = mixed_graphs
    say(f"mixed-graph launch, spread_waves={spread}")
    s = Server(0, hbm_budget=4 << 30, streams=1, max_batch=64, timeout_s=TIMEOUT,
               spread_waves=spread)
    for g, p in zip(gs, params):
        s.add_graph((g.row, g.col, g.w), **p)
    rng = np.random.default_rng(31)
    gsel = np.repeat(np.arange(3, dtype=np.int32), 6)
    rng.shuffle(gsel)
    src = np.array([rng.integers(0, gs[g].n) for g in gsel], dtype=np.int32)
    res, st, mixed = run_checked(s, gs, np.zeros(len(gsel)), gsel, src, "mixed launch")
    # all 18 requests in one launch
    diag = (f"launches {st['launches']} launch ids {res['launch'].tolist()} "
            f"retries {res['retries'].tolist()} batch sizes {res['batch_size'].tolist()}")
    assert (res["retries"] == 0).all(), diag
    assert st["launches"] == 1 and (res["launch"] == res["launch"][0]).all(), diag
    assert (res["batch_size"] == len(gsel)).all(), diag
    for i in range(len(gsel)):
        _, st1, alone = run_checked(s, gs, [0.0], [gsel[i]], [src[i]], f"alone {i}")
        assert st1["launches"] == 1
        assert np.array_equal(alone[0], mixed[i]), i
    say(f"mixed launch: device p50 {np.median(res['device_ms']):.2f} ms, "
        f"latency max {res['latency_ms'].max():.2f} ms")
    s.close()


def test_mixed_widths_split_launches(graphs):
    """Graphs with different workgroup widths are not mixed: each launch
    holds requests of one width, and every row is exact."""
    say("mixed widths")
    s = Server(0, hbm_budget=256 * MiB, streams=1, max_batch=32, timeout_s=TIMEOUT, spread_waves=0)
    for g, w in zip(graphs, (4, 8, 4)):
        s.add_graph((g.row, g.col, g.w), delta=2000 if g.n != 20000 else 200000, waves=w)
    gsel = np.array([0, 1, 2, 0, 1, 2], dtype=np.int32)
    src = np.array([1, 2, 3, 4, 5, 6], dtype=np.int32)
    res, st, _ = run_checked(s, graphs, np.zeros(6), gsel, src, "widths")
    # graphs 0 and 2 (4 waves) share launches, graph 1 (8 waves) has its own
    for i in range(6):
        for j in range(6):
            if res["launch"][i] == res["launch"][j]:
                assert (gsel[i] == 1) == (gsel[j] == 1)
    assert len(set(res["launch"][gsel != 1])) == 1
    s.close()


def test_burst_batches_and_verifies_on_device(graphs):
    say("burst + device verification")
    s, ids = new_server(graphs, streams=2, max_batch=16, verify=1)
    pools = [[0, 7, 99], [1, 2000, 4799], [5, 123, 19999]]
    for gi, pool in enumerate(pools):
        for src in pool:
            s.set_expected(gi, src, graphs[gi].expect(src))
    arrival, gsel, src = poisson_trace(2000.0, 0.05, [0.5, 0.3, 0.2], pools, seed=3)
    arrival = np.concatenate([np.zeros(40), arrival])
    gsel = np.concatenate([np.array([i % 3 for i in range(40)], dtype=np.int32), gsel])
    src = np.concatenate([np.array([pools[i % 3][i % 2] for i in range(40)], dtype=np.int32), src])
    res, st = s.run(arrival, gsel, src)
    assert st["completed"] == len(arrival)
    assert st["verified"] == len(arrival) and st["mismatched"] == 0
    assert (res["mismatches"] == 0).all()
    assert st["mean_batch"] > 1.0   # the burst was batched
    # a planted mistake in an expected row is caught by the device check
    bad = graphs[1].expect(2000).copy()
    bad[17] += 1
    s.set_expected(1, 2000, bad)
    res, st = s.run([0.0, 0.0], [1, 1], [2000, 1])
    assert list(res["mismatches"]) == [1, 0]
    assert st["mismatched"] == 1
    s.close()


def test_admission_by_bytes(graphs):
    say("admission budget")
    probe, _ = new_server(graphs[:1])
    slot = probe.slot_bytes(0)
    probe.close()
    s, ids = new_server(graphs[:1], budget=2 * slot + 4096, streams=4, max_batch=8)
    src = np.arange(12, dtype=np.int32) * 17
    res, st, _ = run_checked(s, graphs, np.zeros(12), np.zeros(12, dtype=np.int32), src, "budget")
    assert st["peak_inflight"] <= 2
    assert st["peak_bytes"] <= 2 * slot + 4096
    s.close()


def test_rejection_when_queue_full(graphs):
    say("max_pending rejection")
    s, ids = new_server(graphs[:1], streams=1, max_batch=2, max_pending=3)
    res, st = s.run(np.zeros(20), np.zeros(20, dtype=np.int32), np.arange(20, dtype=np.int32))
    assert st["rejected"] == 17 and st["completed"] == 3
    assert (res["status"] == -10).sum() == 17
    s.close()


def test_overflow_reruns(graphs):
    say("overflow reruns")
    # no spreading: the request's few workgroups must push through the tiny L2
    s, ids = new_server(graphs[2:], streams=2, max_batch=4, spread_waves=0)
    s.set_graph_params(0, delta=2000, buckets=4, bucket_window=2, bucket_capacity=64)
    g = graphs[2]
    src = np.array([5, 777, 15000], dtype=np.int32)
    res, st, _ = run_checked(s, [g], np.zeros(3), np.zeros(3, dtype=np.int32), src, "overflow")
    assert st["retries"] > 0
    s.close()


@pytest.mark.parametrize("waves,wgs", [(4, 1), (8, 3), (16, 2), (4, 8), (1, 1), (1, 5), (1, 32)])
def test_workgroups_per_request(graphs, waves, wgs):
    say(f"waves {waves} workgroups {wgs}")
    s, ids = new_server(graphs, streams=3, max_batch=8)
    for gi in range(3):
        s.set_graph_params(gi, delta=2000 if gi < 2 else 200000, waves=waves, workgroups=wgs)
    gsel = np.array([0, 1, 2, 2, 1, 0, 2], dtype=np.int32)
    src = np.array([3, 40, 11, 19000, 4000, 2999, 0], dtype=np.int32)
    run_checked(s, graphs, np.linspace(0, 0.005, 7), gsel, src, f"w{waves}x{wgs}")
    s.close()


@pytest.mark.parametrize("prm", [dict(prefetch=-1), dict(stale_check=1), dict(stale_check=3),
                                 dict(stale_check=-1, prefetch=-1)])
def test_prefetch_and_stale_modes_served(graphs, prm):
    say(f"serve {prm}")
    s, ids = new_server(graphs, streams=2, max_batch=8)
    for gi in range(3):
        s.set_graph_params(gi, delta=2000 if gi < 2 else 200000, workgroups=2, **prm)
    gsel = np.array([0, 1, 2, 2, 1, 0, 2], dtype=np.int32)
    src = np.array([3, 40, 11, 19000, 4000, 2999, 0], dtype=np.int32)
    run_checked(s, graphs, np.linspace(0, 0.005, 7), gsel, src, f"{prm}")
    s.close()


@pytest.mark.parametrize("spread", [0, 64, 4096])
def test_spread_waves(graphs, spread):
    """Launches of few requests spread them over more workgroups (up to 64
    per request); results do not depend on it."""
    say(f"spread_waves {spread}")
    s, ids = new_server(graphs, streams=2, max_batch=4, spread_waves=spread)
    gsel = np.array([0, 1, 2, 2, 1, 0, 2], dtype=np.int32)
    src = np.array([3, 40, 11, 19000, 4000, 2999, 0], dtype=np.int32)
    run_checked(s, graphs, np.linspace(0, 0.02, 7), gsel, src, f"spread {spread}")
    s.close()


def test_hub_graph_served(graphs):
    say("hub graph served")
    g = G(*hub_csr(20000, [40000, 5000, 2100], 40000, seed=8))
    s, ids = new_server([g], streams=2, max_batch=4)
    for prm in (dict(hub_split=1), dict(hub_split=0), dict(light_split=1, delta=100, buckets=32)):
        s.set_graph_params(0, **dict(dict(delta=300, workgroups=2), **prm))
        run_checked(s, [g], np.linspace(0, 0.002, 6), np.zeros(6, dtype=np.int32),
                    np.array([0, 1, 2, 19999, 5, 0], dtype=np.int32), f"hubs {prm}")
    s.close()


def test_fixture_golden(graphs):
    say("fixture golden")
    s = Server(0, hbm_budget=64 * MiB, timeout_s=TIMEOUT)
    gid = s.add_graph(os.path.join(GOLDEN, "fixture.gr"))
    n = s.sizes[gid][0]
    outs = [np.empty(n, dtype=np.int32) for _ in range(3)]
    res, st = s.run([0.0, 0.0, 0.0], [gid] * 3, [0, 1030, 2050], dist_out=outs)
    assert (res["status"] == 0).all()
    for o, src in zip(outs, (0, 1030, 2050)):
        gold = np.fromfile(os.path.join(GOLDEN, f"s{src}.i32"), dtype=np.int32)
        assert np.array_equal(o, gold)
    s.close()


def test_slot_that_never_fits(graphs):
    say("slot larger than the budget")
    s, ids = new_server(graphs[2:], budget=1 * MiB)
    res, st = s.run([0.0, 0.001], [0, 0], [1, 2])
    assert list(res["status"]) == [-4, -4]
    assert st["failed"] == 2
    s.close()


def test_argument_errors(graphs):
    say("argument errors")
    s, ids = new_server(graphs[:1])
    with pytest.raises(MLMQError):
        s.run([0.1, 0.0], [0, 0], [1, 2])           # arrivals must not decrease
    with pytest.raises(MLMQError):
        s.run([0.0], [5], [1])                      # unknown graph
    with pytest.raises(MLMQError):
        s.run([0.0], [0], [graphs[0].n])            # source out of range
    with pytest.raises(MLMQError):
        s.set_graph_params(0, waves=6)
    with pytest.raises(MLMQError):
        s.set_graph_params(0, stale_check=4)
    with pytest.raises(MLMQError):
        s.set_graph_params(0, prefetch=2)
    # the server still works after refused calls
    run_checked(s, graphs, [0.0], [0], [1], "after errors")
    s.close()
