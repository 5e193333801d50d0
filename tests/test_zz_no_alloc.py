"""mlmq_solve_noalloc: the allocation-free variant of mlmq_solve (needs an
MI355X).  Called through the C ABI; each result is compared with
mlmq_solve's on the same context and with the oracle, on two calls in a row,
the second on a smaller graph, with the device's free memory unchanged by
each call."""
import ctypes
import os

import numpy as np
import pytest

from hip_rt import DeviceArray, mem_get_info, to_host
from oracle import oracle as O
from mlmq_tpds_withl3_amd import Solver, _lib
from mlmq_tpds_withl3_amd.graphs import HostGraph
from mlmq_tpds_withl3_amd.solver import default_params, Timing

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def noalloc(s: Solver, src: int, out: DeviceArray, **kw) -> int:
    p = default_params(timeout_s=30.0, **kw)
    t = Timing()
    return _lib.load().mlmq_solve_noalloc(s._h, int(src), ctypes.byref(p), out.ptr, ctypes.byref(t))


def test_noalloc_matches_solve_two_calls_second_smaller():
    print("[no-alloc] two calls, second on a smaller graph", flush=True)
    g = HostGraph.knn(40_000, 3, seed=11, weight_scale=1e5)
    row, col, w = g.arrays()
    big, small = Solver(0), Solver(0)
    try:
        big.load_graph(g)
        src_a = 123
        big.solve(src_a, timeout_s=30.0)
        ref_a = big.dist()
        exp_a, _, _ = O.dijkstra(row, col, w, src_a)
        assert np.array_equal(ref_a, exp_a)

        small.load_gr(os.path.join(GOLDEN, "fixture.gr"))
        src_b = 1030
        small.solve(src_b, timeout_s=30.0)
        ref_b = small.dist()
        gold_b = np.fromfile(os.path.join(GOLDEN, f"s{src_b}.i32"), dtype=np.int32)
        assert np.array_equal(ref_b, gold_b)

        out_a = DeviceArray(np.full(big.n, 0x5A5A5A5A, dtype=np.int32))
        out_b = DeviceArray(np.full(small.n, 0x5A5A5A5A, dtype=np.int32))
        try:
            # call 1: the larger graph
            free0, _ = mem_get_info()
            assert noalloc(big, src_a, out_a) == 0, _lib.load().mlmq_last_error(big._h)
            free1, _ = mem_get_info()
            got_a = to_host(out_a, np.int32, big.n)
            assert np.array_equal(got_a, ref_a)
            # call 2: the smaller graph, same workspace
            assert noalloc(small, src_b, out_b) == 0, _lib.load().mlmq_last_error(small._h)
            free2, _ = mem_get_info()
            got_b = to_host(out_b, np.int32, small.n)
            assert np.array_equal(got_b, ref_b)
            assert free0 == free1 == free2, (free0, free1, free2)
            # the contexts' own last results are untouched
            assert np.array_equal(big.dist(), ref_a)
            assert np.array_equal(small.dist(), ref_b)
        finally:
            out_a.free()
            out_b.free()
    finally:
        big.close()
        small.close()
        g.close()


def test_noalloc_modes_and_limits():
    print("[no-alloc] queue modes, explicit capacity, refused sizes", flush=True)
    g = HostGraph.knn(20_000, 3, seed=5, weight_scale=1e5)
    row, col, w = g.arrays()
    s = Solver(0)
    out = DeviceArray(np.zeros(g.n, dtype=np.int32))
    try:
        s.load_graph(g)
        for src, kw in ((0, dict(queue="L1V_L2DQ")), (777, dict(buckets=2, bucket_window=2)),
                        (4321, dict(bucket_capacity=1 << 12)), (99, dict(light_split=1, delta=5000))):
            assert noalloc(s, src, out, **kw) == 0, (kw, _lib.load().mlmq_last_error(s._h))
            exp, _, _ = O.dijkstra(row, col, w, src)
            assert np.array_equal(to_host(out, np.int32, g.n), exp), kw
        # 16 buckets x 2^23 slots is beyond the static workspace: refused before any launch
        rc = noalloc(s, 0, out, bucket_capacity=_lib.NOALLOC_RING_SLOTS)
        assert rc == -1 and b"static workspace" in _lib.load().mlmq_last_error(s._h)
        # a null output buffer is refused
        p = default_params()
        assert _lib.load().mlmq_solve_noalloc(s._h, 0, ctypes.byref(p), None, None) == -1
    finally:
        out.free()
        s.close()
        g.close()
