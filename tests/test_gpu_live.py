"""Live serving: one persistent launch fed through a request ring (needs an
MI355X).

Every served row is compared with the CPU oracle.  Cases: requests submitted
and collected one by one; the same requests alone and all at once (rows
equal); a short request submitted while a long one runs finishes first in
the same launch; a trace replayed through the live path; device
verification with a planted wrong row; overflow reruns; the launch ending
when idle and starting again; light/heavy-split graphs beside plain ones;
argument errors.
"""
import sys
import time

import numpy as np
import pytest

from oracle import oracle as O
from graphs_util import grid_csr, hub_csr, random_csr
from mlmq_tpds_withl3_amd import HostGraph, MLMQError
from mlmq_tpds_withl3_amd.serve import Server

pytestmark = pytest.mark.gpu
TIMEOUT = 20.0
MiB = 1 << 20


def say(msg):
    print(f"[gpu-live] {msg}", flush=True)
    sys.stdout.flush()


class G:
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
    c = G(*(x.copy() for x in hg.arrays()))
    hg.close()
    return [a, b, c]


def new_server(graphs, budget=512 * MiB, params=None, **cfg):
    cfg.setdefault("timeout_s", TIMEOUT)
    s = Server(0, hbm_budget=budget, **cfg)
    for i, g in enumerate(graphs):
        p = dict(delta=2000 if g.n != 20000 else 200000)
        if params:
            p.update(params[i])
        s.add_graph((g.row, g.col, g.w), **p)
    return s


def check_row(out, g, src, label):
    bad = np.flatnonzero(out != g.expect(src))
    assert bad.size == 0, f"{label}: {bad.size} mismatches, first v={bad[0]}"


def test_submit_and_wait_one_by_one(graphs):
    say("submit / wait one by one")
    s = new_server(graphs)
    rng = np.random.default_rng(1)
    for i in range(12):
        gi = i % 3
        src = int(rng.integers(0, graphs[gi].n))
        out = np.empty(graphs[gi].n, dtype=np.int32)
        rid = s.submit(gi, src, out)
        r = s.wait(rid, timeout_s=30.0)
        assert r is not None and r["status"] == 0, r
        assert r["latency_ms"] > 0 and r["device_ms"] > 0
        check_row(out, graphs[gi], src, f"request {i}")
    s.close()


def test_alone_equals_all_at_once(graphs):
    say("alone vs all at once")
    s = new_server(graphs)
    rng = np.random.default_rng(2)
    k = 40
    gsel = rng.integers(0, 3, size=k)
    srcs = [int(rng.integers(0, graphs[g].n)) for g in gsel]
    alone = []
    for g, src in zip(gsel, srcs):
        out = np.empty(graphs[g].n, dtype=np.int32)
        r = s.wait(s.submit(int(g), src, out), timeout_s=30.0)
        assert r["status"] == 0
        alone.append(out)
    outs = [np.empty(graphs[g].n, dtype=np.int32) for g in gsel]
    ids = [s.submit(int(g), src, o) for g, src, o in zip(gsel, srcs, outs)]
    assert s.drain(60.0)
    res = [s.poll(i) for i in ids]
    for i, (g, src) in enumerate(zip(gsel, srcs)):
        assert res[i]["status"] == 0, res[i]
        check_row(outs[i], graphs[g], src, f"mixed {i}")
        assert np.array_equal(outs[i], alone[i])
    # one launch served them all
    assert len({r["launch"] for r in res}) == 1
    s.close()


def test_short_request_overtakes_a_long_one():
    """A short request submitted while a long one runs joins the running
    launch and finishes first."""
    say("short overtakes long")
    long_g = HostGraph.knn(1_000_000, 3, seed=21, weight_scale=-50_000.0, morton=True)
    lg = G(*(x.copy() for x in long_g.arrays()))
    long_g.close()
    small = G(*grid_csr(20, 20, seed=22))
    s = Server(0, hbm_budget=2 << 30, timeout_s=30.0, spread_waves=0)
    s.add_graph((lg.row, lg.col, lg.w), delta=200_000, workgroups=1, waves=4)   # slow on purpose
    s.add_graph((small.row, small.col, small.w), delta=2000)
    out_l = np.empty(lg.n, dtype=np.int32)
    rid_l = s.submit(0, 0, out_l)
    time.sleep(0.01)
    outs = [np.empty(small.n, dtype=np.int32) for _ in range(4)]
    ids = [s.submit(1, i * 7, outs[i]) for i in range(4)]
    res_s = [s.wait(i, timeout_s=30.0) for i in ids]
    assert s.poll(rid_l) is None, "the long request finished before the short ones"
    res_l = s.wait(rid_l, timeout_s=60.0)
    assert res_l["status"] == 0
    check_row(out_l, lg, 0, "long")
    for i, r in enumerate(res_s):
        assert r["status"] == 0 and r["launch"] == res_l["launch"]
        check_row(outs[i], small, i * 7, f"short {i}")
    say(f"long {res_l['latency_ms']:.1f} ms, short max {max(r['latency_ms'] for r in res_s):.2f} ms")
    s.close()


def test_trace_through_the_live_path(graphs):
    say("trace in live mode")
    s = new_server(graphs, mode="live")
    rng = np.random.default_rng(3)
    k = 60
    arrival = np.sort(rng.uniform(0, 0.05, size=k))
    gsel = rng.integers(0, 3, size=k).astype(np.int32)
    src = np.array([rng.integers(0, graphs[g].n) for g in gsel], dtype=np.int32)
    outs = [np.empty(graphs[g].n, dtype=np.int32) for g in gsel]
    res, st = s.run(arrival, gsel, src, dist_out=outs)
    assert st["completed"] == k and st["failed"] == 0 and st["rejected"] == 0, st
    for i in range(k):
        check_row(outs[i], graphs[gsel[i]], int(src[i]), f"trace {i}")
    assert st["peak_inflight"] > 1
    say(f"live trace: p50 {st['p50_ms']:.2f} ms p99 {st['p99_ms']:.2f} ms")
    s.close()


def test_device_verification_and_a_planted_wrong_row(graphs):
    say("verification")
    s = new_server(graphs, verify=1)
    s.set_expected(0, 5, graphs[0].expect(5))
    s.set_expected(1, 9, graphs[1].expect(9))
    planted = graphs[2].expect(11).copy()
    planted[123] ^= 1
    s.set_expected(2, 11, planted)
    r1 = s.wait(s.submit(0, 5), 30.0)
    r2 = s.wait(s.submit(1, 9), 30.0)
    r3 = s.wait(s.submit(2, 11), 30.0)
    r4 = s.wait(s.submit(2, 12), 30.0)
    assert (r1["mismatches"], r2["mismatches"], r3["mismatches"], r4["mismatches"]) == (0, 0, 1, -1)
    s.close()


def test_overflow_reruns(graphs):
    """Records beyond the bucket window cannot be taken, so a 64-slot ring
    of the 20k-vertex road graph runs full: the request reruns with a larger
    ring (no spreading: the request's few workgroups push through the tiny
    ring)."""
    say("overflow reruns")
    s = new_server(graphs[2:], spread_waves=0,
                   params=[dict(delta=2000, buckets=4, bucket_window=2, bucket_capacity=64)])
    g = graphs[2]
    srcs = [5, 777, 15000]
    outs = [np.empty(g.n, dtype=np.int32) for _ in srcs]
    ids = [s.submit(0, src, o) for src, o in zip(srcs, outs)]
    res = [s.wait(i, 60.0) for i in ids]
    assert all(r["status"] == 0 for r in res), res
    assert sum(r["retries"] for r in res) > 0
    for src, o in zip(srcs, outs):
        check_row(o, g, src, f"overflow src {src}")
    s.close()


def test_idle_launch_ends_and_restarts(graphs):
    say("idle stop and restart")
    s = new_server(graphs, live_idle_ms=50)
    out = np.empty(graphs[1].n, dtype=np.int32)
    r1 = s.wait(s.submit(1, 0, out), 30.0)
    check_row(out, graphs[1], 0, "before idle")
    time.sleep(0.3)
    r2 = s.wait(s.submit(1, 100, out), 30.0)
    check_row(out, graphs[1], 100, "after idle")
    assert r2["launch"] == r1["launch"] + 1
    s.stop()
    r3 = s.wait(s.submit(0, 1), 30.0)   # the service thread starts again
    assert r3["status"] == 0
    s.close()


def test_split_and_plain_graphs_together():
    say("light/heavy split beside plain graphs")
    hub = G(*hub_csr(4000, [3000, 2500, 800], 20000, seed=31, wmax=1000))
    grid = G(*grid_csr(40, 50, seed=32))
    rmat_g = HostGraph.rmat(14, 16, seed=33, wmin=1, wmax=1000)
    rmat = G(*(x.copy() for x in rmat_g.arrays()))
    rmat_g.close()
    gs = [hub, grid, rmat]
    s = Server(0, hbm_budget=1 << 30, timeout_s=TIMEOUT)
    s.add_graph((hub.row, hub.col, hub.w), delta=50, light_split=1, buckets=64, bucket_window=2)
    s.add_graph((grid.row, grid.col, grid.w), delta=3000)
    s.add_graph((rmat.row, rmat.col, rmat.w), delta=100, light_split=1, buckets=32, bucket_window=2)
    rng = np.random.default_rng(4)
    jobs = [(int(g), int(rng.integers(0, gs[g].n))) for g in rng.integers(0, 3, size=30)]
    outs = [np.empty(gs[g].n, dtype=np.int32) for g, _ in jobs]
    ids = [s.submit(g, src, o) for (g, src), o in zip(jobs, outs)]
    assert s.drain(60.0)
    for i, (g, src) in enumerate(jobs):
        assert s.poll(ids[i])["status"] == 0
        check_row(outs[i], gs[g], src, f"split mix {i}")
    s.close()


def test_live_argument_errors(graphs):
    say("argument errors")
    s = new_server(graphs, max_pending=1)
    with pytest.raises(MLMQError):
        s.submit(7, 0)
    with pytest.raises(MLMQError):
        s.submit(0, graphs[0].n)
    with pytest.raises(MLMQError):
        s.poll(123456)
    with pytest.raises(ValueError):
        s.submit(0, 0, np.empty(3, dtype=np.int32))
    s.close()
    with pytest.raises(MLMQError):
        Server(0, mode=5)
