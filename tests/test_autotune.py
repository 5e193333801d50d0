"""Host logic of the parameter search (no GPU): the tuners, driven by fake
solvers whose cost functions have a known best point, must return that point
and explore the options they promise to explore."""
import math
import time

import numpy as np

from mlmq_tpds_withl3_amd import autotune as A


class FakeSolver:
    """solve(): device time from a cost function of the parameters;
    solve_batch(): sleeps for three times the cost in ms (the batch tuner
    times wall clock; the factor keeps sleep jitter below the cost steps)."""

    def __init__(self, cost):
        self.cost = cost
        self.seen = []

    @staticmethod
    def params_dict(p):
        return {k: getattr(p, k) for k, _ in p._fields_ if k != "reserved"}

    def solve(self, src, p):
        d = self.params_dict(p)
        self.seen.append(d)
        return {"device_ms": self.cost(d)}

    def solve_batch(self, sources, p):
        d = self.params_dict(p)
        self.seen.append(d)
        time.sleep(3 * self.cost(d) / 1e3)
        return {}


def single_cost(d):
    # best: delta 100000, 512 waves, 4 buckets, pop 32 / top-up 16 / claim 16, late stale check
    c = 5.0 + abs(math.log2(d["delta"] / 100000))
    c += 0.0 if d["work_blocks"] == 512 else 0.5
    c += 0.0 if d["buckets"] == 4 else 0.3
    c += 0.0 if (d["l1_pop"], d["min_read"], d["l2_batch"]) == (32, 16, 16) else 0.2
    c += {3: 0.0, 2: 0.1, 1: 0.4}.get(d["stale_check"], 1.0)
    return c


def test_tune_single_finds_best_and_tries_stale_modes():
    s = FakeSolver(single_cost)
    r = A.tune_single(s, 0, (50000, 100000, 200000), repeats=1)
    best = r["params"]
    assert best["delta"] == 100000 and best["work_blocks"] == 512 and best["buckets"] == 4
    assert (best["l1_pop"], best["min_read"], best["l2_batch"]) == (32, 16, 16)
    assert best["stale_check"] == 3
    assert math.isclose(r["device_ms"], 5.0)
    # the second stage measured both other stale-check modes at the best grid point
    at_best = [d for d in s.seen if d["delta"] == 100000 and d["work_blocks"] == 512
               and d["buckets"] == 4 and d["l1_pop"] == 32 and d["min_read"] == 16]
    assert {d["stale_check"] for d in at_best} == {1, 2, 3}
    # every table entry is one measured point; the four best were then
    # timed seven times more each
    assert len(r["table"]) + 4 * 7 == len(s.seen)
    assert len(r["confirmed"]) == 4


def test_tune_single_tries_touches_off():
    """Touches off is measured at the best points of the second stage and
    wins when it is faster; when it is slower the touches stay on."""
    for off_cost, want in ((-0.3, 0), (+0.3, 1)):
        s = FakeSolver(lambda d: single_cost(d) + (off_cost if d.get("prefetch", 1) == 0 else 0.0))
        r = A.tune_single(s, 0, (50000, 100000, 200000), repeats=1)
        assert r["params"].get("prefetch", 1) == want
        assert r["params"]["stale_check"] == 3 and r["params"]["delta"] == 100000
        assert any(d.get("prefetch", 1) == 0 and d["delta"] == 100000 and d["stale_check"] == 3
                   for d in s.seen)


def test_tune_single_custom_grid():
    s = FakeSolver(lambda d: 1.0 + d["light_split"] * -0.5 + d["work_blocks"] / 1e4)
    grid = [dict(light_split=ls, work_blocks=wb) for ls in (0, 1) for wb in (256, 512)]
    r = A.tune_single(s, 0, (50,), grid=grid, repeats=1)
    assert r["params"]["light_split"] == 1 and r["params"]["work_blocks"] == 256


def test_tune_batch_descends_every_coordinate():
    def cost(d):
        c = 2.0 + 0.5 * abs(math.log2(d["delta"] / 200000))
        c += {16: 0.0, 8: 0.5, 4: 1.0}[d["batch_waves"]]
        c += 0.0 if d["prefetch"] == 0 else 0.5
        c += 0.0 if d["stale_check"] == 1 else 0.5
        c += 0.0 if d["min_read"] == 16 else 0.3
        return c

    s = FakeSolver(cost)
    r = A.tune_batch(s, np.arange(4), base=dict(delta=200000, batch_waves=8, min_read=8,
                                                 l2_batch=8, l1_pop=32, buckets=16))
    p = r["params"]
    assert p["batch_waves"] == 16 and p["prefetch"] == 0 and p["stale_check"] == 1
    assert p["min_read"] == 16 and p["l2_batch"] == 16 and p["delta"] == 200000
    assert r["batch_s"] < 0.03


def test_tune_batch_moves_pop_and_top_up_together():
    # best: pop 32 with top-up 32 (claim 32); raising either alone costs more
    def cost(d):
        c = 3.0 + (0.0 if d["l1_pop"] == d["min_read"] == 32 else 1.0)
        c += 2.0 if d["min_read"] > d["l1_pop"] else 0.0
        return c + (0.5 if d["l1_pop"] > d["min_read"] else 0.0)

    s = FakeSolver(cost)
    r = A.tune_batch(s, np.arange(4), base=dict(batch_waves=16, l1_pop=16, min_read=16, l2_batch=16))
    p = r["params"]
    assert p["l1_pop"] == p["min_read"] == p["l2_batch"] == 32, p
    assert r["trail"][0][0]["stale_check"] == 1   # the search starts at the cached stale check


def test_tune_batch_revisits_delta_after_the_shape_moves():
    # the best delta depends on the touches: only a delta pass after the
    # touches coordinate finds 400000
    def cost(d):
        best = 400000 if d["prefetch"] == 0 else 100000
        c = 2.0 + 0.5 * abs(math.log2(d["delta"] / best))
        return c + (0.0 if d["prefetch"] == 0 else 1.5)

    s = FakeSolver(cost)
    r = A.tune_batch(s, np.arange(4), base=dict(delta=100000, batch_waves=16, prefetch=1))
    assert r["params"]["prefetch"] == 0 and r["params"]["delta"] == 400000


class FakeServer:
    """Burst rate and lone-request latency from functions of the parameters."""

    def __init__(self, qps, lone=lambda p: 10.0):
        self.qps = qps
        self.lone = lone
        self.params = {}
        self.calls = 0

    def set_graph_params(self, gid, **params):
        self.params = dict(params)

    def run(self, arrival, graph, src):
        self.calls += 1
        n = len(arrival)
        res = {"latency_ms": np.full(n, self.lone(self.params))}
        return res, {"completed": n, "mismatched": 0, "completed_qps": self.qps(self.params)}


def test_tune_serving_picks_highest_rate_and_applies_it():
    def qps(p):
        q = 1000.0 - 100 * abs(math.log2(p.get("delta", 200000) / 50000))
        q += 200 if p.get("workgroups") == 16 else 0
        q += 100 if p.get("prefetch") == -1 else 0
        return q

    srv = FakeServer(qps)
    r = A.tune_serving(srv, 0, np.arange(10), base=dict(delta=200000), repeats=1)
    assert r["params"]["delta"] == 50000 and r["params"]["workgroups"] == 16
    assert r["params"]["prefetch"] == -1
    assert srv.params == r["params"]          # the best point is left applied


def test_tune_serving_keeps_lone_latency():
    # the biggest burst rate comes with delta 50000 and l1_pop 16, but
    # l1_pop 16 doubles the latency of a request served alone
    def qps(p):
        return 1000.0 + (300 if p.get("delta") == 50000 else 0) + (200 if p.get("l1_pop") == 16 else 0)

    def lone(p):
        return 20.0 if p.get("l1_pop") == 16 else 10.0

    srv = FakeServer(qps, lone)
    r = A.tune_serving(srv, 0, np.arange(10), base=dict(delta=200000, l1_pop=32), repeats=1)
    assert r["params"]["delta"] == 50000 and r["params"]["l1_pop"] == 32
    r = A.tune_serving(srv, 0, np.arange(10), base=dict(delta=200000, l1_pop=32), repeats=1,
                       lone_slack=None)
    assert r["params"]["l1_pop"] == 16


def test_tune_single_confirms_against_lucky_draws():
    # point A is usually 6 ms but once measured 4 ms; B is always 5 ms
    draws = {"A": iter([4.0] + [6.0] * 100)}

    def cost(d):
        return next(draws["A"]) if d["l1_pop"] == 64 else 5.0

    s = FakeSolver(cost)
    grid = [dict(l1_pop=64), dict(l1_pop=32)]
    r = A.tune_single(s, 0, (100,), grid=grid, repeats=1)
    assert r["params"]["l1_pop"] == 32 and r["device_ms"] == 5.0


def test_tune_single_averages_over_tuning_sources():
    # source 1 prefers delta 50, source 2 prefers delta 200; both together
    # favour delta 100 (mean of the two costs)
    seen_src = []

    class S(FakeSolver):
        def solve(self, src, p):
            seen_src.append(src)
            d = self.params_dict(p)
            best = {1: 50, 2: 200}[src]
            return {"device_ms": 1.0 + math.log2(d["delta"] / best) ** 2}

    r = A.tune_single(S(None), [1, 2], (50, 100, 200), grid=[dict(l1_pop=32)], repeats=1, confirm=1)
    assert r["params"]["delta"] == 100 and r["sources"] == [1, 2]
    assert set(seen_src) == {1, 2}
    assert math.isclose(r["device_ms"], 2.0)


def test_tune_batch_explores_the_l1_ring():
    # best: a 1024-record ring with the spill threshold at 255
    def cost(d):
        c = 5.0
        c += 0.0 if (d["l1_ring"] or 512) == 1024 else 0.8
        c += {255: 0.0, 128: 0.3}.get(d["l1_spill"], 0.6)
        return c

    s = FakeSolver(cost)
    r = A.tune_batch(s, [0, 1, 2])
    assert r["params"]["l1_ring"] == 1024 and r["params"]["l1_spill"] == 255
    # a 512-record ring never ran with a threshold above 64
    assert all(d["l1_spill"] <= (d["l1_ring"] or 512) - 448 for d in s.seen)
    assert {d["l1_ring"] or 512 for d in s.seen} == {512, 1024}
