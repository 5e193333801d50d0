"""The L2 queue and credit protocol on the CPU (tests/native/queue_model.cpp).

The model runs the device code's rules (bucket choice and clamp, per-bucket
reservation, CAS publication with a bounded spin, one-snapshot pops with
persistent claims, window advance, credit return when a wave runs out of
work, done-first termination check) with each model wave a user-level
context and every shared-memory access of a lane one step of a seeded
scheduler, so each seed is one reproducible interleaving.  Checked here:

* no loss, no duplication, exact distances: full solves on random, zero-
  weight, grid and hub graphs, over many interleavings and queue shapes
  (one bucket; narrow and wide queues; claims of 1..64; L1 spill of 1..64;
  a slow termination wave), end with the oracle's distances, every reserved
  record taken and credited, every slot empty;
* termination only at quiescence: at the instant the check raises exit, no
  record is queued or being expanded (ghost count of the model);
* the two orders the build chose against: reading the reservations before the
  credits (the reference's order, SS/SSSP/sssp_run.cu:644-661 with
  cu_delta_queue.cuh:591-600) lets the check exit early, and the window
  rule from before the fix (a claim on the head bucket holds the window
  back) stalls the device queue check's scenario;
* claims no larger than a batch: a claim of more slots than a wave's batch
  holds lets the ring lap under it and strand records in a scripted
  interleaving; capped at the batch, the same interleaving drains;
* overflow: a ring too small for the work ends every wave with the overflow
  error instead of losing records.
"""
import ctypes
import os

import numpy as np
import pytest

from graphs_util import grid_csr, hub_csr, random_csr
from oracle import oracle as O

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "libqueuemodel.so")
ERR_OVERFLOW = 1


@pytest.fixture(scope="module")
def qm():
    if not os.path.exists(LIB):
        raise RuntimeError(f"{LIB} missing: __graft_entry__.build() compiles it")
    L = ctypes.CDLL(LIB)
    P, u32, u64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64
    L.qm_solve.argtypes = [ctypes.c_int32, P, P, P, ctypes.c_int32, ctypes.c_int32, P, u64, u64, P, P]
    L.qm_solve.restype = ctypes.c_int
    L.qm_publication.argtypes = [u32] * 5 + [P, u64, u64, P]
    L.qm_publication.restype = ctypes.c_int
    L.qm_publication_scripted.argtypes = [u32] * 5 + [P, u64, u64, P, u32, P]
    L.qm_publication_scripted.restype = ctypes.c_int
    return L


SOLVE_OUT = ("finished", "steps", "err", "exits", "exit_violations", "tail", "taken", "done",
             "full_slots", "outstanding", "exit", "_")


def model_solve(L, g, src, seed, *, buckets=4, window=2, cap_log2=10, delta=50, claim=8,
                spin=64, pop=32, min_read=8, spill=64, waves=4, advance_rule=0, tails_first=0,
                slow_leader=1, light=0, wide_claims=0, max_steps=3_000_000):
    row, col, w = (np.ascontiguousarray(x, dtype=t) for x, t in zip(g, (np.int64, np.int32, np.int32)))
    n = len(row) - 1
    cfg = np.array([buckets, window, cap_log2, delta, claim, spin, pop, min_read, spill,
                    advance_rule, tails_first, slow_leader, light, wide_claims], dtype=np.uint32)
    dist = np.zeros(n, dtype=np.uint32)
    out = np.zeros(12, dtype=np.uint64)
    rc = L.qm_solve(n, row.ctypes.data, col.ctypes.data, w.ctypes.data, int(src), waves,
                    cfg.ctypes.data, seed, max_steps, dist.ctypes.data, out.ctypes.data)
    assert rc == 0
    return dist.view(np.int32), dict(zip(SOLVE_OUT, (int(x) for x in out)))


GRAPHS = {
    "random": lambda: random_csr(160, 700, seed=1, wmax=100),
    "zero_weights": lambda: random_csr(140, 600, seed=2, wmax=60, zero_frac=0.3),
    "grid": lambda: grid_csr(9, 13, seed=3, wmax=200),
    "hubs": lambda: hub_csr(150, [90, 60], 300, seed=4, wmax=100),
}

SHAPES = [
    dict(buckets=1, window=1, claim=8, spill=64, pop=32, min_read=8, waves=4),
    dict(buckets=4, window=2, claim=2, spill=2, pop=2, min_read=2, waves=4),
    dict(buckets=4, window=1, claim=64, spill=4, pop=4, min_read=1, waves=3, delta=5,
         slow_leader=50),
    dict(buckets=16, window=4, claim=8, spill=4, pop=4, min_read=4, waves=6, delta=10),
    dict(buckets=64, window=2, claim=1, spill=1, pop=1, min_read=1, waves=3, delta=3),
]


@pytest.mark.parametrize("gname", sorted(GRAPHS))
@pytest.mark.parametrize("shape", range(len(SHAPES)))
def test_exact_and_quiescent_under_interleavings(qm, gname, shape):
    g = GRAPHS[gname]()
    src = 0
    ref, _, _ = O.dijkstra(*g, src)
    for seed in range(25):
        dist, o = model_solve(qm, g, src, seed, **SHAPES[shape])
        label = f"{gname} shape {shape} seed {seed}: {o}"
        assert o["finished"] == 1 and o["err"] == 0, label
        assert o["exits"] == 1 and o["exit_violations"] == 0, label
        assert o["tail"] == o["taken"] == o["done"], label
        assert o["full_slots"] == 0 and o["outstanding"] == 0, label
        assert np.array_equal(dist, ref), label


LIGHT_SHAPES = [
    dict(buckets=2, window=1, claim=39, spill=64, pop=25, min_read=10, waves=8, delta=1),
    dict(buckets=4, window=2, claim=8, spill=4, pop=4, min_read=4, waves=6, delta=30),
    dict(buckets=64, window=2, claim=64, spill=64, pop=64, min_read=64, waves=5, delta=50),
]


@pytest.mark.parametrize("gname", sorted(GRAPHS))
@pytest.mark.parametrize("shape", range(len(LIGHT_SHAPES)))
def test_light_heavy_split_under_interleavings(qm, gname, shape):
    """The light/heavy split: a record relaxes only its light edges (weight
    below delta) and, if its vertex has heavy ones, queues a heavy-phase
    record one bucket later, which relaxes only those.  Exact, quiescent at
    exit and with every reservation taken and credited, including delta 1
    with two buckets and a one-bucket window (every edge heavy, every record
    through L2 twice)."""
    g = GRAPHS[gname]()
    ref, _, _ = O.dijkstra(*g, 0)
    for seed in range(12):
        dist, o = model_solve(qm, g, 0, seed, light=1, **LIGHT_SHAPES[shape])
        label = f"{gname} light shape {shape} seed {seed}: {o}"
        assert o["finished"] == 1 and o["err"] == 0, label
        assert o["exits"] == 1 and o["exit_violations"] == 0, label
        assert o["tail"] == o["taken"] == o["done"], label
        assert o["full_slots"] == 0 and o["outstanding"] == 0, label
        assert np.array_equal(dist, ref), label


def test_many_waves_claims_past_the_ring(qm):
    """256 waves claiming 39 slots each run far past a 1024-slot ring (claims
    race ahead of reservations): records still reach exactly one claimant,
    the solve ends quiescent and exact."""
    g = hub_csr(400, [300, 120], 800, seed=12, wmax=100)
    ref, _, _ = O.dijkstra(*g, 5)
    for seed in range(2):
        dist, o = model_solve(qm, g, 5, seed, buckets=2, window=1, cap_log2=10, delta=1, claim=39,
                              pop=25, min_read=10, spill=64, waves=256, light=1, max_steps=20_000_000)
        label = f"seed {seed}: {o}"
        assert o["finished"] == 1 and o["exits"] == 1 and o["exit_violations"] == 0, label
        if o["err"] == 0:
            assert o["tail"] == o["taken"] == o["done"] and o["outstanding"] == 0, label
            assert np.array_equal(dist, ref), label


EARLY = dict(buckets=4, window=1, claim=1, spill=4, pop=4, min_read=1, waves=2, delta=5,
             slow_leader=100)


def test_reservations_first_check_exits_early(qm):
    """The reference's order (tails, then credits) raises exit while records
    are still queued: the first such interleaving is within a few dozen
    seeds; the same interleavings with the done-first order never exit early."""
    g = random_csr(200, 800, seed=3, wmax=100)
    ref, _, _ = O.dijkstra(*g, 0)
    early = None
    for seed in range(600):
        _, o = model_solve(qm, g, 0, seed, tails_first=1, **EARLY)
        if o["exit_violations"]:
            early = seed
            break
    assert early is not None, "no early exit found with the tails-first order"
    for seed in range(early + 1):
        dist, o = model_solve(qm, g, 0, seed, tails_first=0, **EARLY)
        assert o["exit_violations"] == 0 and o["finished"] == 1, (seed, o)
        assert np.array_equal(dist, ref), seed


PUB_OUT = ("finished", "steps", "consumed", "bad", "push_fail", "not_once", "tail", "taken")


def model_publication(L, seed, advance_rule, *, producers=32, batch=64, consumers=4,
                      buckets=16, window=4, cap_log2=12, delta=100, claim=64, span=3200, base=0,
                      room=0, wide_claims=0, spin=1 << 16, max_steps=400_000):
    cfg = np.array([buckets, window, cap_log2, delta, claim, spin, 0, 0, 0, advance_rule, 0, base,
                    room, wide_claims], dtype=np.uint32)
    out = np.zeros(8, dtype=np.uint64)
    rc = L.qm_publication(producers, batch, consumers, 0, span, cfg.ctypes.data, seed, max_steps,
                          out.ctypes.data)
    assert rc == 0
    return dict(zip(PUB_OUT, (int(x) for x in out)))


def test_window_rule_before_the_fix_stalls(qm):
    """The device queue check's scenario (32 producer waves x 64 records, 4
    consumer waves claiming 64, 16 buckets, window 4): with the old rule some
    interleavings stop consuming for good (every consumer holds a claim past
    the head bucket's tail); the current rule consumes every record exactly
    once in every interleaving."""
    stalled = []
    for seed in range(60):
        o = model_publication(qm, seed, advance_rule=1)
        if not o["finished"]:
            assert o["consumed"] < 2048
            stalled.append(seed)
    assert stalled, "the old window rule never stalled"
    for seed in range(60):
        o = model_publication(qm, seed, advance_rule=0)
        assert o["finished"] == 1, (seed, o)
        assert o["consumed"] == o["tail"] == o["taken"] == 2048, (seed, o)
        assert o["bad"] == o["push_fail"] == o["not_once"] == 0, (seed, o)


def lapped_claim_script():
    """Step-by-step interleaving of 4 producers x 16 records (waves 0-3) and
    consumers X, Y, Y2, Z (waves 4-7) taking 32 records per pop, on one bucket
    of 64 slots with claims of 48:
    producers 0-2 reserve positions 0..47; X claims all 48 and polls its lowest
    32 slots before any is written (a wave's first step only starts it; then a
    push is 1 epoch load, 1 tail fetch_add and one CAS per record, and a pop 1
    load of the consumed counter, 4 counter loads, 1 head fetch_add, one
    exchange per tried slot, and 2 fetch_adds when it took records); the
    producers write; producer 3 reserves and writes 48..63; Y, Y2 and Z read
    the counters (head 48, tail 64); Z claims 48..63 and takes them; Y and Y2,
    on their stale snapshot, claim 64..79 and 80..95, one lap ahead, whose
    slots hold records 0..31, and take them.  X's lowest 32 slots then stay
    empty for good, and records 32..47 sit in slots only X's claim covers."""
    sc = [0] * 3 + [1] * 3 + [2] * 3
    sc += [4] * (1 + 1 + 4 + 1 + 32)
    sc += [0] * 16 + [1] * 16 + [2] * 16
    sc += [3] * (3 + 16)
    sc += [5] * 6 + [6] * 6 + [7] * 6
    sc += [7] * 19 + [5] * 19 + [6] * 19
    return np.array(sc, dtype=np.int32)


def test_claims_larger_than_a_batch_strand_records(qm):
    """A wave tries at most the slots its batch can hold (32 records with lane
    pairs), lowest first.  With claims of more slots than that, the scripted
    interleaving (lapped_claim_script) strands 16 records under a claim whose
    tried slots the ring has emptied: no wave takes them again, and the solve
    would spin until the watchdog.  With claims capped at the batch size (the
    device rule) the same interleaving, and random ones, drain every record."""
    sc = lapped_claim_script()
    cfg = lambda wide: np.array([1, 1, 6, 1, 48, 1 << 16, 0, 0, 0, 0, 0, 0, 32, wide], dtype=np.uint32)
    out = np.zeros(8, dtype=np.uint64)
    for wide, seed in ((1, 0), (1, 1), (0, 0), (0, 1)):
        c = cfg(wide)
        assert qm.qm_publication_scripted(4, 16, 4, 0, 1, c.ctypes.data, seed, 200_000, sc.ctypes.data,
                                          len(sc), out.ctypes.data) == 0
        o = dict(zip(PUB_OUT, (int(x) for x in out)))
        if wide:
            assert o["finished"] == 0 and o["consumed"] == 48 and o["taken"] == 48, o
            assert o["tail"] == 64 and o["push_fail"] == 0, o
        else:
            assert o["finished"] == 1 and o["consumed"] == o["taken"] == o["tail"] == 64, o
            assert o["bad"] == o["push_fail"] == o["not_once"] == 0, o
    for seed in range(20):
        o = model_publication(qm, seed, 0, producers=4, batch=16, consumers=4, buckets=1, window=1,
                              cap_log2=6, delta=1, claim=48, span=1, room=32, max_steps=400_000)
        assert o["finished"] == 1 and o["consumed"] == o["tail"] == 64, (seed, o)
        assert o["bad"] == o["push_fail"] == o["not_once"] == 0, (seed, o)


@pytest.mark.parametrize("base", [0, 7])
@pytest.mark.parametrize("claim", [1, 8, 64])
def test_publication_no_loss_no_duplication(qm, base, claim):
    for seed in range(8):
        o = model_publication(qm, seed, 0, producers=16, batch=48, consumers=3, claim=claim,
                              base=base)
        assert o["finished"] == 1 and o["consumed"] == o["tail"] == o["taken"] == 16 * 48, (seed, o)
        assert o["bad"] == o["push_fail"] == o["not_once"] == 0, (seed, o)


def test_overflow_ends_every_wave_with_the_error(qm):
    g = grid_csr(20, 20, seed=5, wmax=50)
    for seed in range(5):
        _, o = model_solve(qm, g, 0, seed, buckets=1, window=1, cap_log2=3, claim=64, spill=1,
                           pop=1, min_read=1, waves=3, spin=32)
        assert o["finished"] == 1, o          # every wave returned
        assert o["err"] == ERR_OVERFLOW and o["exit"] == 1, o
