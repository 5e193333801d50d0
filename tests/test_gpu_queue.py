"""Queue invariants of the L2 delta-bucket queue on the device (needs an MI355X).

Counterparts of the reference's queue test programs
(scripts/multigpu/test_dq_publication.cu, test_dq_capacity.cu), run through
tests/native/libl2check.so, which is compiled from the package's own queue
header:

* publication: 32 producer waves push 1..64 checksummed records each while
  consumer waves pop; every record comes out exactly once and intact, the
  per-bucket counters balance (reserved == taken == pushed, claims never
  behind reservations), for a
  single bucket and for 16 buckets with a window of 4, with the window epoch
  at 0 and at 7 (records below it are clamped to the head bucket, as the
  reference's clamp diagnostic counts);
* capacity: a 1024-slot bucket takes exactly 1024 records, and a further 32
  are refused with the overflow error while the stored records stay intact.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "libl2check.so")
ERR_OVERFLOW = 1
EMPTY = 0xFFFFFFFFFFFFFFFF


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        raise RuntimeError(f"{LIB} missing: build() compiles it")
    L = ctypes.CDLL(LIB)
    u32, P = ctypes.c_uint32, ctypes.c_void_p
    L.l2check_layout.argtypes = [P]
    L.l2check_layout.restype = None
    L.l2check_publication.argtypes = [u32] * 11 + [P, P, P]
    L.l2check_publication.restype = ctypes.c_int
    L.l2check_capacity.argtypes = [u32, u32, u32, P, P, P, P, P]
    L.l2check_capacity.restype = ctypes.c_int
    lay = np.zeros(7, dtype=np.uint32)
    L.l2check_layout(lay.ctypes.data)
    L.layout = dict(zip(("words", "tail", "head", "taken", "base", "err", "stride"),
                        (int(x) for x in lay)))
    return L


def counters(L, ctl, nbuckets, name):
    lay = L.layout
    return np.array([ctl[lay[name] + b * lay["stride"]] for b in range(nbuckets)], dtype=np.int64)


@pytest.mark.parametrize("nbuckets,window,delta,span", [(1, 1, 1, 1), (16, 4, 100, 3200)])
@pytest.mark.parametrize("base", [0, 7])
@pytest.mark.parametrize("batch", [1, 7, 16, 32, 64])
def test_publication(lib, nbuckets, window, delta, span, base, batch):
    say(f"queue publication buckets={nbuckets} window={window} base={base} batch={batch}")
    producers, consumers, reps = 32, 4, 10
    total = producers * batch
    for rep in range(reps):
        seen = np.zeros(total, dtype=np.uint32)
        ctl = np.zeros(lib.layout["words"], dtype=np.uint32)
        counts = np.zeros(3, dtype=np.uint32)
        claim = [8, 1, 64][rep % 3]
        rc = lib.l2check_publication(nbuckets, window, 12, delta, claim, base, producers, batch,
                                     consumers, 0, span, seen.ctypes.data, ctl.ctypes.data,
                                     counts.ctypes.data)
        assert rc == 0
        bad, consumed, push_fail = (int(x) for x in counts)
        assert (bad, push_fail) == (0, 0), (rep, bad, push_fail)
        assert consumed == total and np.all(seen == 1), (rep, consumed, np.flatnonzero(seen != 1)[:8])
        tail = counters(lib, ctl, nbuckets, "tail")
        head = counters(lib, ctl, nbuckets, "head")
        taken = counters(lib, ctl, nbuckets, "taken")
        # every reservation was taken; claims may run ahead of reservations
        # (a consumer claims slots that a later push will fill), never behind
        assert tail.sum() == total and np.array_equal(tail, taken)
        assert np.all(head >= tail)
        assert ctl[lib.layout["err"]] == 0 and ctl[lib.layout["base"]] >= base


def test_capacity_exact_full_then_overflow(lib):
    say("queue capacity: 1024 accepted, 32 more refused")
    cap_log2, first, extra = 10, 1024, 32
    n = 1 << cap_log2
    s1 = np.zeros(n, dtype=np.uint64)
    s2 = np.zeros(n, dtype=np.uint64)
    c1 = np.zeros(lib.layout["words"], dtype=np.uint32)
    c2 = np.zeros(lib.layout["words"], dtype=np.uint32)
    fails = np.zeros(2, dtype=np.uint32)
    rc = lib.l2check_capacity(cap_log2, first, extra, s1.ctypes.data, c1.ctypes.data,
                              s2.ctypes.data, c2.ctypes.data, fails.ctypes.data)
    assert rc == 0
    # exactly full: every record stored once, no error
    assert fails[0] == 0 and c1[lib.layout["err"]] == 0
    assert c1[lib.layout["tail"]] == first
    ids = np.sort((s1 & 0xFFFFFFFF).astype(np.int64))
    assert np.array_equal(ids, np.arange(first)) and np.all((s1 >> 32) == 0)
    # one batch more: refused with the overflow error, nothing overwritten
    assert fails[1] >= 1 and c2[lib.layout["err"]] == ERR_OVERFLOW
    assert np.array_equal(s1, s2)


def test_capacity_partial_fill_accepts_the_rest(lib):
    say("queue capacity: 1000 then 24 more fit exactly")
    n = 1 << 10
    s1 = np.zeros(n, dtype=np.uint64)
    s2 = np.zeros(n, dtype=np.uint64)
    c1 = np.zeros(lib.layout["words"], dtype=np.uint32)
    c2 = np.zeros(lib.layout["words"], dtype=np.uint32)
    fails = np.zeros(2, dtype=np.uint32)
    assert lib.l2check_capacity(10, 1000, 24, s1.ctypes.data, c1.ctypes.data, s2.ctypes.data,
                                c2.ctypes.data, fails.ctypes.data) == 0
    assert int(np.count_nonzero(s1 == EMPTY)) == 24
    assert fails.tolist() == [0, 0] and c2[lib.layout["err"]] == 0
    assert np.array_equal(np.sort((s2 & 0xFFFFFFFF).astype(np.int64)), np.arange(n))
