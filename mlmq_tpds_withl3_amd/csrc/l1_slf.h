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
// registers; each entry carries, beside the 8-byte (vertex, dist) record, the
// vertex's CSR range (start, degree) fetched while its distance was lowered,
// so popping it needs no row-offset load.  Records that go through L2 travel
// without the range (kNoRow) and fetch it again when popped.
#pragma once

#include "l2_delta_queue.h"

namespace mlmq {

constexpr uint32_t kL1Cap = 128;
constexpr uint32_t kL1Mask = kL1Cap - 1;
constexpr uint32_t kL1Spill = 64;
constexpr uint64_t kNoRow = ~0ull;   // CSR range unknown

__device__ __forceinline__ uint64_t make_row(uint32_t start, uint32_t deg) {
    return ((uint64_t)deg << 32) | start;
}

struct L1Ring {
    uint64_t *rec;   // kL1Cap records
    uint64_t *row;   // kL1Cap CSR ranges
};

struct L1State {
    uint32_t head;
    uint32_t size;
};

// Move up to 64 records from the front into lanes [0, k).
__device__ __forceinline__ uint32_t l1_pop(const L1Ring &r, L1State &l1, uint64_t &rec,
                                           uint64_t &row) {
    const uint32_t k = l1.size < (uint32_t)kWave ? l1.size : (uint32_t)kWave;
    const uint32_t lane = lane_id();
    if (lane < k) {
        const uint32_t p = (l1.head + lane) & kL1Mask;
        rec = r.rec[p];
        row = r.row[p];
    }
    wave_sync();
    l1.head = (l1.head + k) & kL1Mask;
    l1.size -= k;
    return k;
}

// Insert the entries of lanes [0, cnt) (cnt <= 64, size <= 64 on entry).
__device__ __forceinline__ void l1_insert(const L1Ring &r, L1State &l1, uint64_t rec, uint64_t row,
                                          uint32_t cnt) {
    const uint32_t lane = lane_id();
    const bool valid = lane < cnt;
    const uint32_t d = rec_dist(rec);
    uint32_t pivot;
    if (l1.size) pivot = uniform(rec_dist(r.rec[l1.head]));
    else pivot = read_lane(d, 0);
    const bool front = valid && d < pivot;
    const bool back = valid && !front;
    const uint64_t fm = ballot(front);
    const uint64_t bm = ballot(back);
    const uint64_t lt = lanemask_lt();
    uint32_t p = 0;
    if (front) p = (l1.head - 1 - popc(fm & lt)) & kL1Mask;
    if (back) p = (l1.head + l1.size + popc(bm & lt)) & kL1Mask;
    if (valid) {
        r.rec[p] = rec;
        r.row[p] = row;
    }
    wave_sync();
    l1.head = (l1.head - popc(fm)) & kL1Mask;
    l1.size += cnt;
}

// Push with the reference's write-back rule: above kL1Spill entries, the
// whole deque goes to L2.  Returns false if L2 overflowed.
__device__ bool l1_push(const L1Ring &r, L1State &l1, uint64_t rec, uint64_t row, uint32_t cnt,
                        const L2Queue &q, WaveStats &st) {
    if (cnt == 0) return true;
    l1_insert(r, l1, rec, row, cnt);
    bool ok = true;
    if (l1.size > kL1Spill) {
        while (l1.size && ok) {
            uint64_t x = kEmpty, y = kNoRow;
            const uint32_t k = l1_pop(r, l1, x, y);
            ok = l2_push(q, x, lane_id() < k, st);
        }
    }
    return ok;
}

}  // namespace mlmq
