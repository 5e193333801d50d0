// Wave64 and memory-model primitives for the gfx950 MLMQ solver.
//
// Every cross-workgroup edge of the queue protocol goes through agent-scope
// __hip_atomic_* operations.  On gfx950 the eight XCDs have private L2 caches;
// agent scope makes the compiler emit the coherent forms (sc1 loads/RMWs), so
// counters, queue slots and distances written on one XCD are seen on another.
// Plain loads are used only for the read-only CSR arrays.
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
constexpr uint32_t kStatsOff = kErrOff + kStride;                 // uint64 stats[kNumStats]
constexpr uint32_t kNumStats = 16;
constexpr uint32_t kCtlWords = kStatsOff + 2 * kNumStats + kStride;

enum DeviceError : uint32_t { kErrNone = 0, kErrOverflow = 1, kErrTimeout = 2 };
enum StatIndex : uint32_t {
    kStatL2Push = 0, kStatL2Pop, kStatRelax, kStatAtomic, kStatImprove,
    kStatStale, kStatBatches, kStatIdle,
    // per-phase clock totals (profiling build of the kernel only)
    kClkL1Pop, kClkL2Pop, kClkRow, kClkEdges, kClkFlush, kClkIdle, kClkTotal, kClkCredit
};

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
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

// ---- agent-scope atomics ----
__device__ __forceinline__ uint32_t ld_relaxed(uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_acquire(uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t fetch_add(uint32_t *p, uint32_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t fetch_add_release(uint32_t *p, uint32_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t fetch_min(uint32_t *p, uint32_t v) {
    return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool cas32(uint32_t *p, uint32_t expected, uint32_t desired) {
    return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED,
                                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld64_relaxed(uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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

// ---- wave-wide integer collectives ----
__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, kWave);
    return x;
}
// Exclusive prefix sum over the 64 lanes; `total` gets the wave-wide sum.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t x, uint32_t &total) {
    const uint32_t lane = lane_id();
    uint32_t v = x;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        uint32_t y = __shfl_up(v, off, kWave);
        if (lane >= (uint32_t)off) v += y;
    }
    total = read_lane(v, kWave - 1);
    return v - x;
}

// Queue record: distance in the high word, vertex in the low word.
__device__ __forceinline__ uint64_t make_rec(uint32_t v, uint32_t d) {
    return ((uint64_t)d << 32) | v;
}
__device__ __forceinline__ uint32_t rec_vertex(uint64_t r) { return (uint32_t)r; }
__device__ __forceinline__ uint32_t rec_dist(uint64_t r) { return (uint32_t)(r >> 32); }

}  // namespace mlmq
