// L2: the global delta-bucket queue, wave64 edition.
//
// Behaviour follows the reference's l2_delta_queue
// (core/cu_delta_queue/cu_delta_queue.cuh:78-642): records sit in `nbuckets`
// circular buckets of width `delta`; the readable window is the `window`
// buckets starting at the epoch `base` (the reference's first_pos); a write
// goes to bucket clamp(dist / delta - base, 0, nbuckets - 1) relative to the
// window (:443-451); readers hold a persistent per-bucket claim of up to
// `batch` slots (:323-327), stop after the first bucket that yields records
// (DQ_READ_MIN 1, :298), and the window advances when its head bucket is
// empty while the queue is not (:374-382).
//
// What is different, by design for gfx950:
//  * Publication is per slot.  A slot holds one 8-byte record or kEmpty; the
//    writer installs its record with a 64-bit compare-and-swap from kEmpty and
//    the reader takes it with an exchange back to kEmpty.  There is no
//    block_write_done array, no manager warps publishing read_pos and no
//    per-block read_pos mirror: a reader polls its own claimed slots.  The
//    record itself is the publication flag, so no fence chain between payload
//    and counter is needed (the chain the reference could not prove).
//  * A full slot (the ring lapped an unconsumed record) makes the writer wait
//    with a bounded spin; past the bound it raises kErrOverflow and the solve
//    stops, instead of silently overwriting.
//  * A claim is (start, 64-bit pending mask) held in lane b of the wave for
//    bucket b, so a wave keeps claims on all buckets in two registers.
//  * A reader takes one snapshot of the buckets' counters (lane b loads
//    bucket b's tail/head/taken, all lanes the epoch) in a single round trip,
//    then decides locally; no per-bucket sequence of loads.  Queues of more
//    than kNarrowQueue buckets read only the window's (see l2_pop).
//  * A writer groups its lanes by bucket and reserves every group's slots in
//    one round trip (the reference's __match_any_sync grouping, done with
//    ballots, and one fetch_add per bucket leader, all issued together).
//  * Counters live on separate 128-byte lines.  Their loads and stores take
//    the scope S of the waves sharing the queue (device_common.h).
#pragma once

#include "device_common.h"

namespace mlmq {

struct L2Queue {
    uint64_t *slots;     // nbuckets << cap_log2 records
    uint32_t *ctl;       // control block (device_common.h layout)
    uint32_t nbuckets;   // 1..64
    uint32_t window;     // 1..nbuckets
    uint32_t cap_log2;   // ring slots per bucket = 1 << cap_log2
    uint32_t delta;      // bucket width (> 0)
    uint32_t batch;      // claim size, 1..64
    uint32_t spin_limit; // bounded wait on a full slot
    uint32_t strict;     // 1: the window advances only once every taken record is credited
};

// Per-wave statistics: event counters in 32 bits (a wave's share of one solve
// stays far below 2^32), phase clocks (profiling build only) in 64 bits.
// Both are wave-uniform and live in scalar registers; the narrow counters
// keep the scalar register file from spilling.
struct WaveStats {
    uint32_t v[kNumStats];
    uint64_t clk[kNumStats];
    __device__ __forceinline__ void clear() {
#pragma unroll
        for (uint32_t i = 0; i < kNumStats; ++i) {
            v[i] = 0;
            clk[i] = 0;
        }
    }
};

__device__ __forceinline__ uint64_t *slot_ptr(const L2Queue &q, uint32_t b, uint32_t pos) {
    return q.slots + ((uint64_t)b << q.cap_log2) + (pos & ((1u << q.cap_log2) - 1));
}

template <int S = kAgentScope>
__device__ __forceinline__ void raise_error(const L2Queue &q, uint32_t code) {
    if (lane_id() == 0) {
        cas32(q.ctl + kErrOff, kErrNone, code);
        st_relaxed<S>(q.ctl + kExitOff, 1u);
    }
}

// Push the records of the lanes with `valid` set.  Wave-uniform call.
// Returns false when a slot stayed full past the spin bound.
template <int S = kAgentScope>
__device__ bool l2_push(const L2Queue &q, uint64_t rec, bool valid, WaveStats &st) {
    uint64_t todo = ballot(valid);
    if (todo == 0) return true;
    const uint32_t lane = lane_id();
    const uint32_t base = uniform(ld_relaxed<S>(q.ctl + kBaseOff));
    uint32_t b = 0;
    if (valid) {
        int32_t rel = (int32_t)(bucket_dist(rec, q.delta) / q.delta) - (int32_t)base;
        rel = rel < 0 ? 0 : rel;
        rel = rel >= (int32_t)q.nbuckets ? (int32_t)q.nbuckets - 1 : rel;
        b = (base + (uint32_t)rel) % q.nbuckets;
    }
    st.v[kStatL2Push] += popc(todo);
    // Group the lanes by bucket (register work only), then every group's
    // leader reserves the group's slots: one fetch_add per bucket, all in one
    // round trip, however many buckets the records span.
    uint32_t rank = 0, cnt = 0, leader_of = 0;
    for (uint64_t rest = todo; rest;) {
        const uint32_t leader = (uint32_t)__builtin_ctzll(rest);
        const uint32_t bl = read_lane(b, leader);
        const bool in_grp = ((rest >> lane) & 1) && b == bl;
        const uint64_t grp = ballot(in_grp);
        if (in_grp) {
            rank = popc(grp & lanemask_lt());
            leader_of = leader;
        }
        if (lane == leader) cnt = popc(grp);
        rest &= ~grp;
    }
    uint32_t pos0 = 0;
    if (cnt) pos0 = fetch_add(q.ctl + kTailOff + b * kStride, cnt);
    pos0 = (uint32_t)__shfl((int)pos0, (int)leader_of, kWave);
    bool ok = true;
    if (valid) {
        uint64_t *slot = slot_ptr(q, b, pos0 + rank);
        uint32_t spins = 0;
        while (!cas64(slot, kEmpty, rec)) {
            if (++spins > q.spin_limit) {
                ok = false;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }
    if (ballot(!ok)) {
        raise_error<S>(q, kErrOverflow);
        return false;
    }
    return true;
}

// Pop up to one claim's worth (at most `room`) of ready records into
// `stage[0..got)`.  `cl_start`/`cl_mask` carry the wave's claims: lane b
// holds bucket b.  `recs` is the most records one batch of the wave holds (the
// room of an idle wave); a new claim takes at most that many slots, so an idle
// wave tries every slot it still holds.  A wider claim can strand records: its
// lowest slots may be emptied for good by claims one ring lap ahead (made on a
// stale snapshot, they take the records under them), and the records under
// its higher slots, which no other claim covers, are then never tried.
// Queues with more buckets than this read, per pop, only the counters of
// the window and one bucket past it, placed by the epoch the wave saw last
// (`base_hint`): every wave polls these lines, so reading 3 x 64 of them per
// pop made the counters a hot spot.  The other buckets are read only when the
// head bucket is drained and the read ones are empty (the window-advance
// decision needs to know whether any bucket still holds records).
constexpr uint32_t kNarrowQueue = 8;

template <bool kStrict = false, int S = kAgentScope>
__device__ uint32_t l2_pop(const L2Queue &q, uint32_t &cl_start, uint64_t &cl_mask,
                           uint64_t *stage, uint32_t room, uint32_t recs, WaveStats &st,
                           uint32_t &base_hint) {
    const uint32_t lane = lane_id();
    const bool mine = lane < q.nbuckets;
    const bool wide = q.nbuckets > kNarrowQueue && !(kStrict && q.strict);
    const bool rd = mine && (!wide || (lane + q.nbuckets - base_hint % q.nbuckets) % q.nbuckets <=
                                          q.window);
    // one round trip: the buckets' counters and the epoch (issued last so
    // that waiting for it does not hold back the others)
    uint32_t t = 0, h = 0, k = 0;
    if (rd) {
        t = ld_relaxed<S>(q.ctl + kTailOff + lane * kStride);
        h = ld_relaxed<S>(q.ctl + kHeadOff + lane * kStride);
        k = ld_relaxed<S>(q.ctl + kTakenOff + lane * kStride);
    }
    const uint32_t base = uniform(ld_relaxed<S>(q.ctl + kBaseOff));
    const uint32_t done = kStrict && q.strict ? uniform(ld_relaxed<S>(q.ctl + kDoneOff)) : 0u;
    const bool unclaimed = rd && (int32_t)(t - h) > 0;
    const uint64_t avail_mask = ballot(unclaimed);
    const uint64_t claim_mask = ballot(mine && cl_mask != 0);
    uint32_t got = 0;
    for (uint32_t i = 0; i < q.window && got == 0; ++i) {
        const uint32_t b = (base + i) % q.nbuckets;
        const bool has_claim = (claim_mask >> b) & 1;
        if (!has_claim && !((avail_mask >> b) & 1)) continue;
        uint32_t s = read_lane(cl_start, b);
        uint64_t mk = read_lane64(cl_mask, b);
        if (!has_claim) {
            const uint32_t avail = read_lane(t, b) - read_lane(h, b);
            const uint32_t kk = min(avail, min(q.batch, recs));
            uint32_t s0 = 0;
            if (lane == 0) s0 = fetch_add(q.ctl + kHeadOff + b * kStride, kk);
            s = read_lane(s0, 0);
            mk = kk >= 64 ? ~0ull : ((1ull << kk) - 1);
        }
        // Take every claimed slot (up to `room`) that holds a record; an empty
        // slot is left empty by the exchange (a writer's later CAS still
        // succeeds).
        uint64_t try_mask = mk;
        while (popc(try_mask) > room) try_mask &= ~(1ull << (63 - __builtin_clzll(try_mask)));
        uint64_t v = kEmpty;
        if ((try_mask >> lane) & 1) v = xchg64(slot_ptr(q, b, s + lane), kEmpty);
        const uint64_t taken = ballot(v != kEmpty);
        mk &= ~taken;
        if (lane == b) {
            cl_start = s;
            cl_mask = mk;
        }
        if (taken) {
            if (v != kEmpty) stage[popc(taken & lanemask_lt())] = v;
            got = popc(taken);
            if (lane == 0) fetch_add(q.ctl + kTakenOff + b * kStride, got);
        }
    }
    // Window advance: head bucket drained (every reserved record taken) while
    // other buckets still hold records.  Strict mode also waits until every
    // taken record has been credited back (its expansion, and the L1 work it
    // spawned, finished), so no record of the head bucket can still lower a
    // distance in the next one: a delta-stepping phase barrier.
    // A claim this wave still holds on the head bucket does not hold the
    // window back: once every reserved record is taken, a claim can only
    // cover slots past the tail (claims race ahead of reservations), which a
    // later push fills when the bucket is back in the window; if every wave
    // held such a claim and waited for it, no wave would advance and the
    // records in later buckets would never be read.
    const uint32_t b0 = base % q.nbuckets;
    const uint32_t live = rd ? t - k : 0;
    const bool b0_read = read_lane(rd ? 1u : 0u, b0) != 0;
    if (b0_read && read_lane(live, b0) == 0) {
        uint32_t total = uniform(wave_sum(live));
        if (wide && total == 0) {
            uint32_t t2 = 0, k2 = 0;
            if (mine && !rd) {
                t2 = ld_relaxed<S>(q.ctl + kTailOff + lane * kStride);
                k2 = ld_relaxed<S>(q.ctl + kTakenOff + lane * kStride);
            }
            total = uniform(wave_sum(t2 - k2));
        }
        bool advance = total > 0;
        if (kStrict && q.strict) advance = advance && done == uniform(wave_sum(mine ? k : 0u));
        if (advance && lane == 0) cas32(q.ctl + kBaseOff, base, base + 1);
    }
    base_hint = base;
    st.v[kStatL2Pop] += got;
    wave_sync();
    return got;
}

// Termination (reference manager warp 0, SS/SSSP/sssp_run.cu:644-661, and
// get_queue_size, cu_delta_queue.cuh:591-600), with the snapshot order fixed:
// credits returned are read first, reservations second.  Both counters only
// grow and reservations are never below credits, so equality of a later
// reservation sum with an earlier credit count proves that at the earlier
// instant no record was queued and no wave held unfinished work.
template <int S = kAgentScope>
__device__ bool l2_check_done(const L2Queue &q) {
    const uint32_t lane = lane_id();
    const uint32_t done = uniform(ld_acquire<S>(q.ctl + kDoneOff));
    uint32_t t = 0;
    if (lane < q.nbuckets) t = ld_relaxed<S>(q.ctl + kTailOff + lane * kStride);
    const uint32_t reserved = uniform(wave_sum(t));
    if (reserved == done) {
        if (lane == 0) st_relaxed<S>(q.ctl + kExitOff, 1u);
        return true;
    }
    return false;
}

}  // namespace mlmq
