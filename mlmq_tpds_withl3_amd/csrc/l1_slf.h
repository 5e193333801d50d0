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
// The head index and size are wave-uniform values kept in scalar registers,
// not LDS words updated by lane 0 as in the reference.
#pragma once

#include "l2_delta_queue.h"

namespace mlmq {

constexpr uint32_t kL1Cap = 128;
constexpr uint32_t kL1Mask = kL1Cap - 1;
constexpr uint32_t kL1Spill = 64;

struct L1State {
    uint32_t head;
    uint32_t size;
};

// Move up to `maxn` records from the front into lanes [dst, dst + k) of `rec`.
__device__ __forceinline__ uint32_t l1_pop(const uint64_t *ring, L1State &l1, uint32_t maxn,
                                           uint32_t dst, uint64_t &rec) {
    const uint32_t k = l1.size < maxn ? l1.size : maxn;
    const uint32_t lane = lane_id();
    if (lane >= dst && lane < dst + k) rec = ring[(l1.head + lane - dst) & kL1Mask];
    wave_sync();
    l1.head = (l1.head + k) & kL1Mask;
    l1.size -= k;
    return k;
}

// Insert the records of lanes [0, cnt) (cnt <= 64, size <= 64 on entry).
__device__ __forceinline__ void l1_insert(uint64_t *ring, L1State &l1, uint64_t rec, uint32_t cnt) {
    const uint32_t lane = lane_id();
    const bool valid = lane < cnt;
    const uint32_t d = rec_dist(rec);
    uint32_t pivot;
    if (l1.size) pivot = uniform(rec_dist(ring[l1.head]));
    else pivot = read_lane(d, 0);
    const bool front = valid && d < pivot;
    const bool back = valid && !front;
    const uint64_t fm = ballot(front);
    const uint64_t bm = ballot(back);
    const uint64_t lt = lanemask_lt();
    if (front) ring[(l1.head - 1 - popc(fm & lt)) & kL1Mask] = rec;
    if (back) ring[(l1.head + l1.size + popc(bm & lt)) & kL1Mask] = rec;
    wave_sync();
    l1.head = (l1.head - popc(fm)) & kL1Mask;
    l1.size += cnt;
}

// Push with the reference's write-back rule: above kL1Spill records, the
// whole deque goes to L2.  Returns false if L2 overflowed.
__device__ bool l1_push(uint64_t *ring, L1State &l1, uint64_t rec, uint32_t cnt,
                        const L2Queue &q, WaveStats &st) {
    if (cnt == 0) return true;
    l1_insert(ring, l1, rec, cnt);
    bool ok = true;
    if (l1.size > kL1Spill) {
        while (l1.size && ok) {
            uint64_t r = kEmpty;
            const uint32_t k = l1_pop(ring, l1, kWave, 0, r);
            ok = l2_push(q, r, lane_id() < k, st);
        }
    }
    return ok;
}

}  // namespace mlmq
