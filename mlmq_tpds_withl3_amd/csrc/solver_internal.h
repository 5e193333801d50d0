// Internal interface between the C-ABI host code (mlmq_api.cpp) and the HIP
// kernels (sssp_kernels.hip).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

struct mlmq_ctx;

namespace mlmq {

// Device graph held by a context (used by the serving loop, which keeps one
// context per resident graph).
struct DeviceGraphView {
    const uint32_t *row;
    const uint32_t *col;
    const uint32_t *wgt;
    const uint4 *blocks;
    const uint2 *hubs;
    bool packed;           // hop blocks carry back-edge slots
    int32_t n;
    int64_t m;
    int device;
    int cus;
    int wall_khz;
};
int ctx_graph_view(const mlmq_ctx *c, DeviceGraphView *v);
// Light-edge counts of the context's hop blocks for `delta` (light/heavy split).
int ctx_prepare_light(mlmq_ctx *c, uint32_t delta);

// One wave per persistent workgroup: the dispatcher then spreads the waves
// over every CU (2 per CU at 512 waves) instead of packing 4 into half the
// CUs, so each wave has a CU's vector L1 and address units nearly to itself
// (road graphs 7-8 % faster than 4-wave workgroups at the same wave count).
constexpr int kWavesPerBlock = 1;

// Bounded wait of an L2 writer on a slot whose previous record is not yet
// taken: ~8k CAS round trips (about 10 ms) before the push reports overflow.
// A reader that has claimed a slot empties it within microseconds, so a
// longer wait only delays the overflow rerun of a genuinely full ring.
constexpr uint32_t kSpinLimit = 1u << 13;

struct SolverLaunch {
    const uint32_t *row;
    const uint32_t *col;
    const uint32_t *wgt;
    const uint4 *blocks;
    const uint2 *hubs;     // hub table (vertex, chunk size); null: no hubs
    bool packed;           // hop blocks carry back-edge slots (graph property)
    uint32_t *dist;
    uint64_t *slots;
    uint32_t *ctl;
    uint32_t n;
    uint32_t nbuckets;
    uint32_t window;
    uint32_t cap_log2;
    uint32_t delta;
    uint32_t batch;
    uint32_t spin_limit;
    uint32_t min_read;
    uint32_t max_sleep;
    uint32_t pop_max;
    uint32_t spill;
    uint32_t donate_min;
    uint64_t timeout_ticks;
    int grid_blocks;
    bool slf;       // L1 in LDS (SLF deque or FIFO); false: L2 only
    bool fifo;      // with slf: L1 is a FIFO (queue mode L1V_L2DQ)
    bool split;     // split hub edge lists into chunk records
    uint32_t stale_check;   // 2 coherent, 1 cached, 0 none
    uint32_t prefetch;      // 1: prefetch relaxed children's hop blocks
    bool hot;               // with slf: block-carrying L1 (single-query kernel only)
    uint32_t defer_heavy;   // 1: defer heavy records beyond the head bucket
    uint32_t strict_window; // 1: window advance waits for all taken records to be credited
    uint32_t light_split;   // 1: light/heavy edge split (sorted adjacency, light counts set)
    uint32_t csr_stream;    // 1: non-temporal loads of CSR columns and weights
    bool profile;   // per-phase clock build of the kernel
};

struct BatchLaunch {
    uint32_t *ctl_all;
    uint64_t *slots_all;
    uint32_t *dist_all;
    const uint32_t *sources;
    uint32_t nq;
    uint32_t *gctl;
    uint32_t *qstatus;
    uint64_t *qtime;
    int waves;   // waves per workgroup (= per query): 4, 8 or 16
    int l1_cap;  // L1 ring records per wave: 512 or 1024
    bool stats;  // count relaxations etc. (more registers per wave)
};

// Serving launches: a table of requests, each with its own graph, queue
// partition and settings (the device layout is private to sssp_kernels.hip;
// the host fills and reads a table through these functions).
size_t serve_request_bytes();
// Request i of a host-side table: L names the graph, the slot (L.dist,
// L.ctl, L.slots) and the settings; workgroups [wg_first, wg_first +
// wg_count) of the launch serve it.
void serve_request_write(void *table, uint32_t i, const SolverLaunch &L, uint32_t src, bool fresh,
                         const uint32_t *expect, uint32_t wg_first, uint32_t wg_count);
void serve_request_read(const void *table, uint32_t i, uint32_t *status, uint32_t *mismatches,
                        uint64_t *t_start, uint64_t *t_end);

struct ServeLaunch {
    void *reqs;               // device table, nreq entries, wg_first ascending from 0
    uint32_t nreq;
    uint32_t grid;            // workgroups in all (sum of wg_count)
    uint32_t max_n;           // longest distance row
    uint64_t max_part_slots;  // largest queue partition (ring slots)
    int waves;                // waves per workgroup: 1, 4, 8 or 16
    bool ext;                 // some request needs light_split / defer_heavy / strict_window
};

// Live serving through a request ring (sssp_stream_kernel): entries and the
// ring live in pinned host memory; the layouts are private to
// sssp_kernels.hip.
size_t stream_entry_bytes();
size_t stream_sync_bytes();
// Host control block of the ring: {published, stop, stopped_at, alive}.
constexpr size_t kStreamHostCtlWords = 4;
void stream_entry_write(void *entry, const SolverLaunch &L, uint32_t src, bool fresh, const uint32_t *expect,
                        uint32_t first, uint32_t parts, uint32_t seq);
// done = 0 while the request runs; the other outputs are set once done = 1.
void stream_entry_read(const void *entry, uint32_t *done, uint32_t *status, uint32_t *mismatches,
                       uint64_t *t_start, uint64_t *t_end);
struct StreamLaunch {
    void *host_ctl;           // pinned host: kStreamHostCtlWords words
    void *entries;            // pinned host: entry table
    const uint32_t *ring;     // pinned host: ring_size words
    uint32_t ring_size;       // power of two
    uint32_t *dev;            // device: 3 words, zeroed before the launch
    void *sync;               // device: one StreamSync per entry, zeroed before the launch
    uint64_t lifetime_ticks;
    uint32_t grid;            // one dispatcher + serving workgroups (one wave each)
};
hipError_t stream_occupancy(int *blocks_per_cu);
hipError_t launch_stream(const StreamLaunch &S, hipStream_t s);
hipError_t launch_serve(const ServeLaunch &S, hipStream_t s);
hipError_t launch_serve_verify(const ServeLaunch &S, hipStream_t s);
hipError_t batch_occupancy(int waves, int l1_cap, bool ext, bool stats, bool pair, int *blocks_per_cu);
hipError_t launch_batch_reset(uint32_t *dist_all, uint32_t n, uint32_t nq, const uint32_t *sources,
                              uint32_t *gctl, hipStream_t s);
hipError_t launch_batch(const SolverLaunch &L, const BatchLaunch &B, hipStream_t s);
hipError_t launch_reset(uint32_t *dist, uint32_t n, uint32_t src, uint32_t *ctl, uint64_t *slots,
                        hipStream_t s);
hipError_t launch_fill_u64(uint64_t *p, uint64_t count, uint64_t value, hipStream_t s);
hipError_t launch_fill_u32(uint32_t *p, uint64_t count, uint32_t value, hipStream_t s);
hipError_t launch_patch_hubs(const uint2 *hubs, uint32_t nh, uint4 *blocks, hipStream_t s);
hipError_t launch_light_count(const uint32_t *row, const uint32_t *wgt, uint32_t n, uint32_t delta,
                              uint4 *blocks, hipStream_t s);
hipError_t launch_build_blocks(const uint32_t *row, const uint32_t *col, const uint32_t *wgt,
                               uint32_t n, bool packed, uint4 *blocks, hipStream_t s);
hipError_t launch_copy_u32(uint32_t *dst, const uint32_t *src, uint64_t count, hipStream_t s);
// Lane pairs (two lanes per record, four edge slots each) whenever a batch
// pops at most 32 records (a serving launch decides per request).
__host__ __device__ inline bool lane_pairs_for(uint32_t pop_max) { return pop_max <= 32u; }
hipError_t solver_occupancy(bool slf, bool hot, bool pair, int *blocks_per_cu);
hipError_t launch_solver(const SolverLaunch &L, hipStream_t s);

}  // namespace mlmq
