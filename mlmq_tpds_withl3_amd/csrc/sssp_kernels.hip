// Persistent MLMQ SSSP kernels for gfx950 (CDNA4, wave64).
//
// One grid of role-free persistent workgroups (kWavesPerBlock waves each).
// Every wave is a worker running the loop of the reference's
// work_block_kernel (SS/SSSP/sssp_run.cu:507-584):
//   1. pop up to 64 records from its L1 SLF deque; below `min_read` records,
//      top up from the L2 delta-bucket queue (core/src/ml_queue.cu:223-319);
//   2. relax every out-edge of the batch (simple_process, :143-413) with a
//      load-balanced wave-wide edge loop and agent-scope atomicMin;
//   3. push improved (vertex, dist) records through L1, spilling to L2
//      (core/src/ml_queue.cu:321-355);
//   4. return the L2 credit once L1 is empty (:565-572).
// Global wave 0 also detects termination while idle, replacing the separate
// manager kernel (:623-674).  Queue publication needs no manager at all (see
// l2_delta_queue.h), so no wave waits on a particular other workgroup: the
// solve completes even if only part of the grid is resident.
//
// S = evidence/l3_supplement_20260921_v2/single_source/ (paths above are in S).
#include "l1_slf.h"
#include "solver_internal.h"

namespace mlmq {

constexpr uint32_t kOutCap = 128;   // pending pushes (reference node_out: 2 x node_size)

struct SolveArgs {
    const uint32_t *row;   // n + 1 CSR offsets
    const uint32_t *col;   // m destinations
    const uint32_t *wgt;   // m weights
    uint32_t *dist;        // n distances
    L2Queue q;
    uint32_t n;
    uint32_t min_read;
    uint32_t max_sleep;    // idle back-off cap (x 256 cycles)
    uint64_t timeout_ticks;
};

template <bool kProf>
struct PhaseClock {
    uint64_t t;
    __device__ __forceinline__ void start() {
        if (kProf) t = clock64();
    }
    __device__ __forceinline__ void lap(WaveStats &st, uint32_t slot) {
        if (kProf) {
            const uint64_t now = clock64();
            st.v[slot] += now - t;
            t = now;
        }
    }
};

// Flush the newest `cnt` (<= 64) pending entries of `out` to L1 (or L2).
template <bool kSLF>
__device__ __forceinline__ bool flush_out(const L1Ring &out, uint32_t &out_n, uint32_t cnt,
                                          const L1Ring &ring, L1State &l1, const L2Queue &q,
                                          WaveStats &st) {
    const uint32_t lane = lane_id();
    uint64_t r = kEmpty, w = kNoRow;
    if (lane < cnt) {
        r = out.rec[out_n - cnt + lane];
        w = out.row[out_n - cnt + lane];
    }
    wave_sync();
    out_n -= cnt;
    if (kSLF) return l1_push(ring, l1, r, w, cnt, q, st);
    return l2_push(q, r, lane < cnt, st);
}

// Relax all out-edges of the records in lanes [0, nin) (simple_process,
// SS/SSSP/sssp_run.cu:143-413).  A record is skipped when its distance is
// above the vertex's current one (reference filter :174, process on <=).
// Edge slots are spread over the 64 lanes by an exclusive scan of degrees;
// each lane finds its owner by a binary search over the scan held in lanes.
// A lane pushes (dst, nd) only if its atomicMin strictly lowered dist[dst]
// (:211-243).
//
// Latency shape (one hop = two dependent round trips): the stale check of
// the batch is in flight together with the first chunk's col/weight loads,
// and each relaxation's atomicMin is in flight together with the loads of
// dst's CSR range, which travels with the pushed record.
template <bool kSLF, bool kProf>
__device__ bool process_batch(const SolveArgs &a, uint64_t rec, uint64_t rowi, uint32_t nin,
                              const L1Ring &out, const L1Ring &ring, L1State &l1, WaveStats &st,
                              PhaseClock<kProf> &clk) {
    const uint32_t lane = lane_id();
    const bool valid = lane < nin;
    const uint32_t v = rec_vertex(rec);
    const uint32_t d = rec_dist(rec);
    uint32_t start = (uint32_t)rowi, deg = (uint32_t)(rowi >> 32);
    // Records that came through L2 carry no CSR range.
    const bool need_row = valid && rowi == kNoRow;
    if (ballot(need_row)) {
        if (need_row) {
            start = a.row[v];
            deg = a.row[v + 1] - start;
        }
        clk.lap(st, kClkRow);
    }
    if (!valid) deg = 0;
    uint32_t cur = 0;
    if (valid) cur = ld_relaxed(a.dist + v);
    uint32_t total = 0;
    const uint32_t excl = wave_excl_scan(deg, total);
    uint32_t out_n = 0;
    bool ok = true;
    bool keep_known = false;
    bool keep = false;
    for (uint32_t base = 0; base < total && ok; base += kWave) {
        const uint32_t qi = base + lane;
        const bool act = qi < total;
        uint32_t lo = 0;
#pragma unroll
        for (uint32_t step = 32; step > 0; step >>= 1) {
            const uint32_t cand = lo + step;
            if ((uint32_t)__shfl((int)excl, (int)cand, kWave) <= qi) lo = cand;
        }
        const uint32_t o_excl = (uint32_t)__shfl((int)excl, (int)lo, kWave);
        const uint32_t o_start = (uint32_t)__shfl((int)start, (int)lo, kWave);
        const uint32_t o_d = (uint32_t)__shfl((int)d, (int)lo, kWave);
        uint32_t dst = 0, wt = 0;
        if (act) {
            const uint32_t e = o_start + (qi - o_excl);
            dst = a.col[e];
            wt = a.wgt[e];
        }
        if (!keep_known) {
            keep = valid && d <= cur;
            keep_known = true;
            st.v[kStatStale] += popc(ballot(valid && !keep));
        }
        const bool o_keep = __shfl((int)keep, (int)lo, kWave) != 0;
        const bool go = act && o_keep;
        bool improved = false;
        uint32_t nd = 0, cs = 0, ce = 0;
        if (go) {
            nd = o_d + wt;
            const uint32_t prev = fetch_min(a.dist + dst, nd);
            cs = a.row[dst];
            ce = a.row[dst + 1];
            improved = nd < prev;
        }
        const uint64_t gm = ballot(go);
        const uint64_t im = ballot(improved);
        st.v[kStatRelax] += popc(gm);
        st.v[kStatAtomic] += popc(gm);
        st.v[kStatImprove] += popc(im);
        if (improved) {
            const uint32_t p = out_n + popc(im & lanemask_lt());
            out.rec[p] = make_rec(dst, nd);
            out.row[p] = make_row(cs, ce - cs);
        }
        wave_sync();
        out_n += popc(im);
        clk.lap(st, kClkEdges);
        if (out_n >= kWave) {
            ok = flush_out<kSLF>(out, out_n, kWave, ring, l1, a.q, st);
            clk.lap(st, kClkFlush);
        }
    }
    if (!keep_known) {
        // batch without edges: still count stale records
        st.v[kStatStale] += popc(ballot(valid && d > cur));
    }
    if (ok && out_n) {
        ok = flush_out<kSLF>(out, out_n, out_n, ring, l1, a.q, st);
        clk.lap(st, kClkFlush);
    }
    return ok;
}

template <int kWaves, bool kSLF, bool kProf>
__global__ __launch_bounds__(kWaves * 64) void sssp_persistent_kernel(SolveArgs a) {
    __shared__ uint64_t s_ring_rec[kWaves][kL1Cap];
    __shared__ uint64_t s_ring_row[kWaves][kL1Cap];
    __shared__ uint64_t s_out_rec[kWaves][kOutCap];
    __shared__ uint64_t s_out_row[kWaves][kOutCap];
    __shared__ uint64_t s_stage[kWaves][kWave];

    const uint32_t lane = lane_id();
    const uint32_t w = threadIdx.x / kWave;
    const uint32_t gwave = blockIdx.x * kWaves + w;
    const L1Ring ring{s_ring_rec[w], s_ring_row[w]};
    const L1Ring out{s_out_rec[w], s_out_row[w]};
    uint64_t *stage = s_stage[w];
    const L2Queue &q = a.q;

    WaveStats st;
#pragma unroll
    for (uint32_t i = 0; i < kNumStats; ++i) st.v[i] = 0;
    PhaseClock<kProf> clk;
    const uint64_t t_begin = kProf ? clock64() : 0;
    clk.start();

    const uint64_t deadline = wall_clock64() + a.timeout_ticks;
    L1State l1{0u, 0u};
    uint32_t credit = 0;        // records taken from L2 whose work is not finished
    uint32_t cl_start = 0;      // lane b: start of the claim on bucket b
    uint64_t cl_mask = 0;       // lane b: pending slots of that claim
    uint32_t sleep = 0;

    for (;;) {
        uint64_t rec = kEmpty, rowi = kNoRow;
        uint32_t nin = 0;
        if (kSLF) {
            nin = l1_pop(ring, l1, rec, rowi);
            clk.lap(st, kClkL1Pop);
        }
        if (nin < a.min_read) {  // here L1 is empty
            const uint32_t got = l2_pop(q, cl_start, cl_mask, stage, st);
            if (got) {
                credit += got;
                const uint32_t direct = got < kWave - nin ? got : kWave - nin;
                if (lane >= nin && lane < nin + direct) {
                    rec = stage[lane - nin];
                    rowi = kNoRow;
                }
                const uint32_t rest = got - direct;
                if (rest) {
                    uint64_t r2 = kEmpty;
                    if (lane < rest) r2 = stage[direct + lane];
                    wave_sync();
                    if (!l1_push(ring, l1, r2, kNoRow, rest, q, st)) break;
                }
                nin += direct;
            }
            clk.lap(st, kClkL2Pop);
        }
        if (nin == 0) {
            if (credit) {
                if (lane == 0) fetch_add_release(q.ctl + kDoneOff, credit);
                credit = 0;
            }
            st.v[kStatIdle] += 1;
            if (gwave == 0 && l2_check_done(q)) break;
            if (uniform(ld_relaxed(q.ctl + kExitOff))) break;
            if (wall_clock64() > deadline) {
                raise_error(q, kErrTimeout);
                break;
            }
            sleep = sleep ? (sleep * 2 < a.max_sleep ? sleep * 2 : a.max_sleep) : 1;
            for (uint32_t k = 0; k < sleep; ++k) __builtin_amdgcn_s_sleep(4);
            clk.lap(st, kClkIdle);
            continue;
        }
        sleep = 0;
        st.v[kStatBatches] += 1;
        if (!process_batch<kSLF, kProf>(a, rec, rowi, nin, out, ring, l1, st, clk)) break;
        if (credit && l1.size == 0) {
            if (lane == 0) fetch_add_release(q.ctl + kDoneOff, credit);
            credit = 0;
            clk.lap(st, kClkCredit);
        }
        if ((st.v[kStatBatches] & 31) == 0) {
            if (uniform(ld_relaxed(q.ctl + kExitOff))) break;
            if (wall_clock64() > deadline) {
                raise_error(q, kErrTimeout);
                break;
            }
        }
    }
    if (kProf) st.v[kClkTotal] = clock64() - t_begin;

    if (lane == 0) {
        uint64_t *stats = reinterpret_cast<uint64_t *>(q.ctl + kStatsOff);
#pragma unroll
        for (uint32_t i = 0; i < kNumStats; ++i)
            if (st.v[i]) add64(stats + i, st.v[i]);
    }
}

// Per-query reset: dist = INT_MAX except dist[src] = 0; control block zeroed
// with tail[0] = 1; the seed record (src, 0) in slot 0 of bucket 0, which is
// where the reference's write_through of (src, 0) lands at first_pos 0
// (SS/SSSP/sssp_run.cu:473-486).  Slots are kEmpty after any successful solve.
__global__ void reset_kernel(uint32_t *dist, uint32_t n, uint32_t src, uint32_t *ctl,
                             uint64_t *slots) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = tid; i < n; i += stride) dist[i] = i == src ? 0u : kUnreached;
    for (uint32_t i = tid; i < kCtlWords; i += stride) ctl[i] = i == kTailOff ? 1u : 0u;
    if (tid == 0) slots[0] = make_rec(src, 0u);
}

__global__ void fill_u64_kernel(uint64_t *p, uint64_t count, uint64_t value) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
        p[i] = value;
}

__global__ void fill_u32_kernel(uint32_t *p, uint64_t count, uint32_t value) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
        p[i] = value;
}

// Converts a reference-layout int32 CSR row array into the solver's uint32 one.
__global__ void copy_u32_kernel(uint32_t *dst, const uint32_t *src, uint64_t count) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
        dst[i] = src[i];
}

// ------------------------------------------------------------------ host side

static int grid_for(uint64_t count, int threads) {
    uint64_t g = (count + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (int)g;
}

hipError_t launch_reset(uint32_t *dist, uint32_t n, uint32_t src, uint32_t *ctl, uint64_t *slots,
                        hipStream_t s) {
    const uint64_t work = n > kCtlWords ? n : kCtlWords;
    hipLaunchKernelGGL(reset_kernel, dim3(grid_for(work, 256)), dim3(256), 0, s, dist, n, src, ctl,
                       slots);
    return hipGetLastError();
}

hipError_t launch_fill_u64(uint64_t *p, uint64_t count, uint64_t value, hipStream_t s) {
    hipLaunchKernelGGL(fill_u64_kernel, dim3(grid_for(count, 256)), dim3(256), 0, s, p, count,
                       value);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *p, uint64_t count, uint32_t value, hipStream_t s) {
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(count, 256)), dim3(256), 0, s, p, count,
                       value);
    return hipGetLastError();
}

hipError_t launch_copy_u32(uint32_t *dst, const uint32_t *src, uint64_t count, hipStream_t s) {
    hipLaunchKernelGGL(copy_u32_kernel, dim3(grid_for(count, 256)), dim3(256), 0, s, dst, src,
                       count);
    return hipGetLastError();
}

static const void *solver_fn(bool slf) {
    return slf ? reinterpret_cast<const void *>(&sssp_persistent_kernel<kWavesPerBlock, true, false>)
               : reinterpret_cast<const void *>(&sssp_persistent_kernel<kWavesPerBlock, false, false>);
}

hipError_t solver_occupancy(bool slf, int *blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, solver_fn(slf),
                                                        kWavesPerBlock * kWave, 0);
}

hipError_t launch_solver(const SolverLaunch &L, hipStream_t s) {
    SolveArgs a;
    a.row = L.row;
    a.col = L.col;
    a.wgt = L.wgt;
    a.dist = L.dist;
    a.q.slots = L.slots;
    a.q.ctl = L.ctl;
    a.q.nbuckets = L.nbuckets;
    a.q.window = L.window;
    a.q.cap_log2 = L.cap_log2;
    a.q.delta = L.delta;
    a.q.batch = L.batch;
    a.q.spin_limit = L.spin_limit;
    a.n = L.n;
    a.min_read = L.min_read;
    a.max_sleep = L.max_sleep;
    a.timeout_ticks = L.timeout_ticks;
    const dim3 grid(L.blocks), block(kWavesPerBlock * kWave);
    if (L.slf && !L.profile)
        hipLaunchKernelGGL((sssp_persistent_kernel<kWavesPerBlock, true, false>), grid, block, 0, s, a);
    else if (L.slf)
        hipLaunchKernelGGL((sssp_persistent_kernel<kWavesPerBlock, true, true>), grid, block, 0, s, a);
    else if (!L.profile)
        hipLaunchKernelGGL((sssp_persistent_kernel<kWavesPerBlock, false, false>), grid, block, 0, s, a);
    else
        hipLaunchKernelGGL((sssp_persistent_kernel<kWavesPerBlock, false, true>), grid, block, 0, s, a);
    return hipGetLastError();
}

}  // namespace mlmq
