// L1: the per-wavefront "small label first" deque in LDS.
//
// Behaviour follows the reference's l1_SLF_queue (core/include/l1_queue.cuh:81-95,
// core/src/l1_queue.cu:217-360), re-derived for 64-lane waves: a ring of
// kL1Cap = 128 records (reference 64 for 32 lanes); a pop takes up to 64 from
// the head (reference 32 = capacity / 2); a pushed record whose distance is
// below the current head's (or, on an empty deque, below the first pushed
// record's) goes to the front, every other record to the back; when the deque
// holds more than kL1Spill = 64 records after a push, all of them move to L2.
//
// Differences: head index and size are wave-uniform values in scalar
// registers; and each entry carries, beside the 8-byte (vertex, dist) record,
// the vertex's 64-byte hop block (its out-edges, see hop_block.h), fetched in
// the same round trip as the atomicMin that created the record.  Popping an
// entry therefore needs no graph load at all.  Records that go through L2
// travel without the block and fetch it again when popped.
#pragma once

#include "hop_block.h"
#include "l2_delta_queue.h"

namespace mlmq {

constexpr uint32_t kL1Cap = 128;
constexpr uint32_t kL1Mask = kL1Cap - 1;
constexpr uint32_t kL1Spill = 64;

// A ring of records and their blocks, blocks stored as 4 planes of 16 bytes
// (plane j of entry p at blk[j * cap + p]) so lanes reading consecutive
// entries hit consecutive LDS addresses.
struct BlockRing {
    uint64_t *rec;
    uint4 *blk;
    uint32_t cap;
};

struct L1State {
    uint32_t head;
    uint32_t size;
};

__device__ __forceinline__ void ring_put(const BlockRing &r, uint32_t p, uint64_t rec,
                                         const HopBlock &b) {
    r.rec[p] = rec;
    r.blk[0 * r.cap + p] = b.q[0];
    r.blk[1 * r.cap + p] = b.q[1];
    r.blk[2 * r.cap + p] = b.q[2];
    r.blk[3 * r.cap + p] = b.q[3];
}

__device__ __forceinline__ void ring_get(const BlockRing &r, uint32_t p, uint64_t &rec,
                                         HopBlock &b) {
    rec = r.rec[p];
    b.q[0] = r.blk[0 * r.cap + p];
    b.q[1] = r.blk[1 * r.cap + p];
    b.q[2] = r.blk[2 * r.cap + p];
    b.q[3] = r.blk[3 * r.cap + p];
}

// Move up to `maxn` (<= 64) entries from the front into lanes [0, k) of
// rec/blk (registers).
__device__ __forceinline__ uint32_t l1_pop(const BlockRing &r, L1State &l1, uint32_t maxn,
                                           uint64_t &rec, HopBlock &b) {
    const uint32_t k = l1.size < maxn ? l1.size : maxn;
    if (lane_id() < k) ring_get(r, (l1.head + lane_id()) & kL1Mask, rec, b);
    wave_sync();
    l1.head = (l1.head + k) & kL1Mask;
    l1.size -= k;
    return k;
}

// Insert the entries of the lanes in `mask` (at most 64; size <= 64 on entry).
__device__ __forceinline__ void l1_insert(const BlockRing &r, L1State &l1, uint64_t rec,
                                          const HopBlock &b, uint64_t mask) {
    const uint32_t lane = lane_id();
    const bool valid = (mask >> lane) & 1;
    const uint32_t d = rec_dist(rec);
    uint32_t pivot;
    if (l1.size) pivot = uniform(rec_dist(r.rec[l1.head]));
    else pivot = read_lane(d, (uint32_t)__builtin_ctzll(mask));
    const bool front = valid && d < pivot;
    const bool back = valid && !front;
    const uint64_t fm = ballot(front);
    const uint64_t bm = ballot(back);
    const uint64_t lt = lanemask_lt();
    uint32_t p = 0;
    if (front) p = (l1.head - 1 - popc(fm & lt)) & kL1Mask;
    if (back) p = (l1.head + l1.size + popc(bm & lt)) & kL1Mask;
    if (valid) ring_put(r, p, rec, b);
    wave_sync();
    l1.head = (l1.head - popc(fm)) & kL1Mask;
    l1.size += popc(mask);
}

// Push with the reference's write-back rule: above kL1Spill entries, the
// whole deque goes to L2.  Returns false if L2 overflowed.
__device__ bool l1_push(const BlockRing &r, L1State &l1, uint64_t rec, const HopBlock &b,
                        uint64_t mask, const L2Queue &q, WaveStats &st) {
    if (mask == 0) return true;
    l1_insert(r, l1, rec, b, mask);
    bool ok = true;
    if (l1.size > kL1Spill) {
        while (l1.size && ok) {
            const uint32_t k = l1.size < (uint32_t)kWave ? l1.size : (uint32_t)kWave;
            uint64_t x = kEmpty;
            if (lane_id() < k) x = r.rec[(l1.head + lane_id()) & kL1Mask];
            wave_sync();
            l1.head = (l1.head + k) & kL1Mask;
            l1.size -= k;
            ok = l2_push(q, x, lane_id() < k, st);
        }
    }
    return ok;
}

}  // namespace mlmq
