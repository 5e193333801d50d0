// L1: the per-wavefront "small label first" deque in LDS.
//
// Behaviour follows the reference's l1_SLF_queue (core/include/l1_queue.cuh:81-95,
// core/src/l1_queue.cu:217-360): a ring of records; a pop takes records from
// the head; a pushed record whose distance is below the current head's (or,
// on an empty deque, below the first pushed record's) goes to the front,
// every other record to the back; when the deque holds more than the spill
// threshold after a push, all of its records move to L2.
//
// Sizes are re-derived for 64-lane waves and many-edge pushes: the ring holds
// kCap 8-byte records, 512 (4 KiB per wave) unless a kernel instance asks for
// a larger power of two (the batched kernel also comes with 1024: the LDS
// segment size its tuner chooses); a pop takes at most 64 (one per lane);
// the spill threshold is a launch parameter (default 64, the reference's
// "half of the ring" rule at wave64 granularity; at most kCap - 7 x 64, so
// that one push of a whole batch always fits).  One push
// places all records a batch produced -- up to 64 lanes x 7 edges -- in a
// single pass: positions come from ballots and popcounts only, so the writes
// are issued back to back.  Head index and size are wave-uniform values in
// scalar registers, not LDS words updated by lane 0 as in the reference.
#pragma once

#include "hop_block.h"
#include "l2_delta_queue.h"

namespace mlmq {

constexpr uint32_t kL1Cap = 512;             // default ring size (records)
constexpr uint32_t kL1Mask = kL1Cap - 1;
constexpr uint32_t kL1Alloc = kL1Cap + 64;   // ring + one discard slot per lane
// Block-carrying L1 (single-query kernel): ring slot i is paired with block
// slot i, which holds the hop block of the record in slot i, loaded while
// that record was being relaxed into existence, or kColdMark in word 0 when
// the block is not there (the record then takes the global round trip).
constexpr uint32_t kColdMark = 0xFFFFFFFFu;

struct L1State {
    uint32_t head;
    uint32_t size;
};

// Head and size are wave-uniform; re-assert it so the compiler keeps them in
// scalar registers and branches on them without exec masking.
__device__ __forceinline__ void l1_uniform(L1State &l1) {
    l1.head = uniform(l1.head);
    l1.size = uniform(l1.size);
}

// Move up to `maxn` (<= 64) records from the front into lanes [0, k).
template <uint32_t kCap = kL1Cap>
__device__ __forceinline__ uint32_t l1_pop(const uint64_t *ring, L1State &l1, uint32_t maxn,
                                           uint64_t &rec) {
    static_assert((kCap & (kCap - 1)) == 0 && kCap >= kL1Cap, "ring size");
    l1_uniform(l1);
    const uint32_t k = uniform(l1.size < maxn ? l1.size : maxn);
    if (lane_id() < k) rec = ring[(l1.head + lane_id()) & (kCap - 1)];
    wave_sync();
    l1.head = (l1.head + k) & (kCap - 1);
    l1.size -= k;
    return k;
}

// Pop for lane pairs: up to `maxn` (<= 32) records from the front, record i
// into lanes 2i and 2i + 1.
template <uint32_t kCap = kL1Cap>
__device__ __forceinline__ uint32_t l1_pop_pair(const uint64_t *ring, L1State &l1, uint32_t maxn,
                                                uint64_t &rec) {
    l1_uniform(l1);
    const uint32_t k = uniform(l1.size < maxn ? l1.size : maxn);
    const uint32_t r = lane_id() >> 1;
    if (r < k) rec = ring[(l1.head + r) & (kCap - 1)];
    wave_sync();
    l1.head = (l1.head + k) & (kCap - 1);
    l1.size -= k;
    return k;
}

// Pop for the block-carrying L1: like l1_pop, and each popped lane also gets
// the block slot of its record.  When the head record is hot (its block is
// in LDS) only the leading run of hot records is taken, so the batch needs
// no global round trip for them.
__device__ __forceinline__ uint32_t l1_pop_hot(const uint64_t *ring, const HopBlock *blk,
                                               L1State &l1, uint32_t maxn, uint64_t &rec,
                                               bool &hot, uint32_t (&hw)[16]) {
    l1_uniform(l1);
    const uint32_t lane = lane_id();
    uint32_t k = uniform(l1.size < maxn ? l1.size : maxn);
    // Every lane reads (slots past the popped ones are in the ring too), so
    // the block stays in registers with no branch around it.
    const uint32_t pos = (l1.head + lane) & kL1Mask;
    const uint64_t r = ring[pos];
    const uint32_t *src = reinterpret_cast<const uint32_t *>(blk + pos);
#pragma unroll
    for (int i = 0; i < 16; ++i) hw[i] = src[i];
    if (lane < k) rec = r;
    hot = lane < k && hw[0] != kColdMark;
    const uint64_t hm = ballot(hot);
    if (hm & 1) k = uniform((uint32_t)__builtin_ctzll(~hm));   // leading hot run (<= k)
    hot = hot && lane < k;
    l1.head = (l1.head + k) & kL1Mask;
    l1.size -= k;
    return k;
}

// Spill every record of the deque to L2 (reference write-back of the whole
// L1, core/src/l1_queue.cu:332-350).  Returns false if L2 overflowed.
template <uint32_t kCap = kL1Cap, int S = kAgentScope>
__device__ bool l1_spill_all(const uint64_t *ring, L1State &l1, const L2Queue &q, WaveStats &st) {
    bool ok = true;
    while (l1.size && ok) {
        uint64_t x = kEmpty;
        const uint32_t k = l1_pop<kCap>(ring, l1, kWave, x);
        ok = l2_push<S>(q, x, lane_id() < k, st);
    }
    return ok;
}

// Work donation (no reference counterpart): move up to `want` records from
// the back of the deque -- the ones SLF would process last -- to L2, where
// idle waves can claim them.  Returns false if L2 overflowed.
template <uint32_t kCap = kL1Cap, int S = kAgentScope>
__device__ bool l1_donate(const uint64_t *ring, L1State &l1, uint32_t want, const L2Queue &q,
                          WaveStats &st) {
    l1_uniform(l1);
    const uint32_t m = uniform(want < l1.size ? want : l1.size);
    if (m == 0) return true;
    const uint32_t k = m < (uint32_t)kWave ? m : (uint32_t)kWave;
    uint64_t x = kEmpty;
    if (lane_id() < k) x = ring[(l1.head + l1.size - k + lane_id()) & (kCap - 1)];
    wave_sync();
    l1.size -= k;
    st.v[kStatDonated] += k;
    return l2_push<S>(q, x, lane_id() < k, st);
}

// Push K registers of records: record k of this lane is present when bit k
// of `bits` is set.  Requires size <= spill on entry and
// spill + K * 64 <= kCap.  With `fifo` set every record goes to the back:
// the reference's l1_vector_queue (core/src/l1_queue.cu:74-99), a FIFO with
// the same spill-all rule.
//
// Positions come from one packed wave prefix scan (front count in the low
// half-word, back count in the high one); every lane then writes its records
// without further cross-lane work.  Absent records are written to the lane's
// own slot past the ring (kCap + lane), so the writes need no branches.
// Order: lane-major, then k; front records are placed downwards from the old
// head, so the first one pushed ends next to it.
//
// kBlk selects what goes into the block slots of a block-carrying L1:
// 0 = no block slots, 1 = the records' blocks `cb[k]`, 2 = kColdMark.
template <int K, int kBlk = 0, uint32_t kCap = kL1Cap, int S = kAgentScope>
__device__ __forceinline__ bool l1_push(uint64_t *ring, L1State &l1, const uint64_t (&rec)[K],
                                        uint32_t bits, uint32_t spill, uint32_t fifo,
                                        const L2Queue &q, WaveStats &st, HopBlock *blk = nullptr,
                                        const HopBlock *cb = nullptr) {
    l1_uniform(l1);
    const uint32_t lane = lane_id();
    const uint32_t cnt = (uint32_t)__builtin_popcount(bits);
    // Pivot: the head's distance, or on an empty deque the first pushed
    // record's (lowest lane with a record, its lowest k).
    uint32_t pivot = 0;   // stays 0 in FIFO mode: no record is below it
    const uint64_t any = ballot(bits != 0);
    if (any == 0) return true;
    if (fifo) {
    } else if (l1.size) {
        pivot = uniform(rec_dist(ring[l1.head]));
    } else {
        uint32_t fd = 0;
#pragma unroll
        for (int k = K - 1; k >= 0; --k)
            if ((bits >> k) & 1) fd = rec_dist(rec[k]);
        pivot = read_lane(fd, (uint32_t)__builtin_ctzll(any));
    }
    uint32_t fbits = 0;
#pragma unroll
    for (int k = 0; k < K; ++k)
        if (((bits >> k) & 1) && rec_dist(rec[k]) < pivot) fbits |= 1u << k;
    const uint32_t nf = (uint32_t)__builtin_popcount(fbits);
    uint32_t tot = 0;
    const uint32_t ex = wave_excl_scan(nf | ((cnt - nf) << 16), tot);
    const uint32_t total_f = tot & 0xFFFFu, total = total_f + (tot >> 16);
    uint32_t pf = l1.head - 1 - (ex & 0xFFFFu);
    uint32_t pb = l1.head + l1.size + (ex >> 16);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const bool v = (bits >> k) & 1;
        const bool f = (fbits >> k) & 1;
        const uint32_t p = v ? ((f ? pf : pb) & (kCap - 1)) : kCap + lane;
        ring[p] = rec[k];
        if (kBlk == 1 && v) blk[p] = cb[k];
        if (kBlk == 2 && v) reinterpret_cast<uint32_t *>(blk + p)[0] = kColdMark;
        pf -= f ? 1u : 0u;
        pb += (v && !f) ? 1u : 0u;
    }
    wave_sync();
    l1.head = (l1.head - total_f) & (kCap - 1);
    l1.size += total;
    if (l1.size > spill) return l1_spill_all<kCap, S>(ring, l1, q, st);
    return true;
}

}  // namespace mlmq
