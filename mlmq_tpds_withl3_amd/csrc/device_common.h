// Wave64 and memory-model primitives for the gfx950 MLMQ solver.
//
// Every cross-workgroup edge of the queue protocol goes through agent-scope
// __hip_atomic_* operations (the eight XCDs of gfx950 have private L2
// caches; agent scope is what makes counters, queue slots and distances
// written on one XCD visible on another).  Plain loads are used only for the
// read-only graph arrays.  Measured on MI355X with 64-lane dependent gathers:
// ~115 ns for an L2 hit, ~315 ns for an Infinity-Cache hit, ~540 ns from HBM,
// with agent-scope atomics (load, add, min) costing the same as plain loads;
// so the solver is shaped by the number of dependent round trips per hop.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mlmq {

constexpr int kWave = 64;
constexpr uint64_t kEmpty = ~0ull;            // empty queue slot (never a valid record)
constexpr uint32_t kUnreached = 0x7FFFFFFFu;  // INT_MAX, the reference's DIST_MAX

// ---- control block layout (uint32 words; one 128-byte line per counter) ----
constexpr uint32_t kMaxBuckets = 64;
constexpr uint32_t kStride = 32;
constexpr uint32_t kTailOff = 0;                                  // tail[b]: records reserved
constexpr uint32_t kHeadOff = kTailOff + kMaxBuckets * kStride;   // head[b]: slots claimed by readers
constexpr uint32_t kTakenOff = kHeadOff + kMaxBuckets * kStride;  // taken[b]: records consumed
constexpr uint32_t kBaseOff = kTakenOff + kMaxBuckets * kStride;  // window start (bucket epoch)
constexpr uint32_t kDoneOff = kBaseOff + kStride;                 // credits returned
constexpr uint32_t kExitOff = kDoneOff + kStride;                 // 1 = stop
constexpr uint32_t kErrOff = kExitOff + kStride;                  // MLMQ device error code
constexpr uint32_t kHungryOff = kErrOff + kStride;                // waves currently without work
constexpr uint32_t kStatsOff = kHungryOff + kStride;              // uint64 stats[kNumStats]
constexpr uint32_t kNumStats = 24;
constexpr uint32_t kCtlWords = kStatsOff + 2 * kNumStats + kStride;

enum DeviceError : uint32_t { kErrNone = 0, kErrOverflow = 1, kErrTimeout = 2, kErrInternal = 3 };
enum StatIndex : uint32_t {
    kStatL2Push = 0, kStatL2Pop, kStatRelax, kStatAtomic, kStatImprove,
    kStatStale, kStatBatches, kStatIdle,
    // per-phase clock totals (profiling build of the kernel only)
    kClkL1Pop, kClkL2Pop, kClkRow, kClkEdges, kClkFlush, kClkIdle, kClkTotal, kClkCredit,
    kClkOwner, kClkEdgeRead, kClkAtomic, kClkLaps, kStatChunks, kStatDonated, kStatSpare1, kStatSpare2
};

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }
__device__ __forceinline__ uint64_t lanemask_lt() { return __lanemask_lt(); }
__device__ __forceinline__ uint32_t popc(uint64_t x) { return (uint32_t)__popcll(x); }
__device__ __forceinline__ uint32_t uniform(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint32_t read_lane(uint32_t x, uint32_t l) {
    return __builtin_amdgcn_readlane(x, l);
}
__device__ __forceinline__ uint64_t read_lane64(uint64_t x, uint32_t l) {
    uint32_t lo = __builtin_amdgcn_readlane((uint32_t)x, l);
    uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(x >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}
// Orders LDS traffic between lanes of one wave (compiler barrier; the wave
// itself executes in lock step).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// ---- atomics ----
// Loads and stores of queue state take the scope of the waves that share
// it: agent scope when the workgroups of a launch share it, workgroup scope
// when one workgroup owns it (the batched kernel's per-query partitions).
// On gfx950 an agent-scope load refetches its line from memory (sc1), a
// workgroup-scope load reads through this XCD's L2 (sc0), and an agent-scope
// acquire or release also invalidates or writes back the XCD's L2; the
// read-modify-write atomics below are the same instructions at either scope
// (executed at the memory side), so they keep agent scope.
constexpr int kAgentScope = __HIP_MEMORY_SCOPE_AGENT;
constexpr int kGroupScope = __HIP_MEMORY_SCOPE_WORKGROUP;
template <int S = kAgentScope>
__device__ __forceinline__ uint32_t ld_relaxed(uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, S);
}
template <int S = kAgentScope>
__device__ __forceinline__ uint32_t ld_acquire(uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, S);
}
template <int S = kAgentScope>
__device__ __forceinline__ void st_relaxed(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, S);
}
__device__ __forceinline__ uint32_t fetch_add(uint32_t *p, uint32_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t fetch_min(uint32_t *p, uint32_t v) {
    return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool cas32(uint32_t *p, uint32_t expected, uint32_t desired) {
    return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED,
                                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool cas64(uint64_t *p, uint64_t expected, uint64_t desired) {
    return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED,
                                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t xchg64(uint64_t *p, uint64_t v) {
    return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void add64(uint64_t *p, uint64_t v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- wave-wide integer collectives (DPP: a few cycles per step, no LDS) ----
// Hillis-Steele inside each 16-lane row (row_shr 1, 2, 4, 8), then the gfx9
// row broadcasts carry lane 15 into rows 1 and 3 and lane 31 into rows 2 and
// 3.  Lanes whose DPP source falls outside the row keep the identity (0).
template <int kCtrl, int kRowMask = 0xf>
__device__ __forceinline__ uint32_t dpp(uint32_t src) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)src, kCtrl, kRowMask, 0xf, false);
}
__device__ __forceinline__ uint32_t wave_incl_add(uint32_t x) {
    x += dpp<0x111>(x);
    x += dpp<0x112>(x);
    x += dpp<0x114>(x);
    x += dpp<0x118>(x);
    x += dpp<0x142, 0xa>(x);
    x += dpp<0x143, 0xc>(x);
    return x;
}
// Inclusive running maximum (values are >= 0, identity 0).
__device__ __forceinline__ uint32_t wave_incl_max(uint32_t x) {
    x = max(x, dpp<0x111>(x));
    x = max(x, dpp<0x112>(x));
    x = max(x, dpp<0x114>(x));
    x = max(x, dpp<0x118>(x));
    x = max(x, dpp<0x142, 0xa>(x));
    x = max(x, dpp<0x143, 0xc>(x));
    return x;
}
__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
    return read_lane(wave_incl_add(x), kWave - 1);
}
// Exclusive prefix sum over the 64 lanes; `total` gets the wave-wide sum.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t x, uint32_t &total) {
    const uint32_t incl = wave_incl_add(x);
    total = read_lane(incl, kWave - 1);
    return incl - x;
}

// Queue record: distance in the high word, vertex in the low word.
__device__ __forceinline__ uint64_t make_rec(uint32_t v, uint32_t d) {
    return ((uint64_t)d << 32) | v;
}
__device__ __forceinline__ uint32_t rec_vertex(uint64_t r) { return (uint32_t)r; }
__device__ __forceinline__ uint32_t rec_dist(uint64_t r) { return (uint32_t)(r >> 32); }

// Light/heavy edge split: a heavy-phase record (relax only the edges of weight
// >= delta of its vertex) carries this flag in the top bit of its distance word
// (distances are below 2^31) and is queued one bucket later than its distance.
constexpr uint32_t kHeavyPhase = 0x80000000u;
__device__ __forceinline__ uint32_t bucket_dist(uint64_t r, uint32_t delta) {
    const uint32_t d = rec_dist(r);
    return (d & kHeavyPhase) ? (d & ~kHeavyPhase) + delta : d;
}

}  // namespace mlmq
