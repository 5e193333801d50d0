// Queue-level checks of the L2 delta-bucket queue (csrc/l2_delta_queue.h),
// the counterparts of the reference's queue test programs:
//  * publication (scripts/multigpu/test_dq_publication.cu:18-80): producer
//    waves push checksummed records while consumer waves pop them; every
//    record must come out exactly once and intact, the counters must balance,
//    also with the window epoch moved so that records are clamped to the head
//    bucket;
//  * capacity (scripts/multigpu/test_dq_capacity.cu:10-18): a bucket filled
//    to exactly its capacity accepts every record; one more push finds a full
//    slot and is refused with the overflow error, overwriting nothing.
// Every loop here is bounded (consumers by an iteration cap, producers by the
// queue's own spin limit), so each wave exits whatever the queue does.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "l2_delta_queue.h"

using namespace mlmq;

namespace {

// Distance carried by record `id`: spread over `span` from `dist0`, so that a
// record's distance is also its checksum.
__host__ __device__ inline uint32_t dist_of(uint32_t id, uint32_t dist0, uint32_t span) {
    return dist0 + (uint32_t)(((uint64_t)id * 2654435761ull) % span);
}

__global__ __launch_bounds__(64) void publication_kernel(L2Queue q, uint32_t producers,
                                                         uint32_t batch, uint32_t dist0,
                                                         uint32_t span, uint32_t total,
                                                         uint32_t max_iters, uint32_t *seen,
                                                         uint32_t *counts) {
    __shared__ uint64_t stage[kWave];
    WaveStats st;
    st.clear();
    const uint32_t lane = lane_id();
    if (blockIdx.x < producers) {
        const uint32_t id = blockIdx.x * batch + lane;
        const bool valid = lane < batch;
        const uint64_t rec = valid ? make_rec(id, dist_of(id, dist0, span)) : kEmpty;
        if (!l2_push(q, rec, valid, st) && lane == 0) atomicAdd(counts + 2, 1u);
        return;
    }
    uint32_t cl_start = 0, base_hint = 0;
    uint64_t cl_mask = 0;
    for (uint32_t it = 0; it < max_iters; ++it) {
        if (uniform(ld_relaxed(counts + 1)) >= total) break;
        const uint32_t got = l2_pop(q, cl_start, cl_mask, stage, kWave, kWave, st, base_hint);
        if (lane < got) {
            const uint64_t r = stage[lane];
            const uint32_t v = rec_vertex(r);
            if (v < total && rec_dist(r) == dist_of(v, dist0, span))
                atomicAdd(seen + v, 1u);
            else
                atomicAdd(counts + 0, 1u);
        }
        if (got) {
            if (lane == 0) fetch_add(counts + 1, got);
        } else {
            __builtin_amdgcn_s_sleep(8);
        }
    }
}

__global__ __launch_bounds__(64) void fill_kernel(L2Queue q, uint32_t first_id, uint32_t count,
                                                  uint32_t *push_fail) {
    WaveStats st;
    st.clear();
    const uint32_t i = blockIdx.x * kWave + lane_id();
    const bool valid = i < count;
    if (!l2_push(q, valid ? make_rec(first_id + i, 0u) : kEmpty, valid, st) && lane_id() == 0)
        atomicAdd(push_fail, 1u);
}

struct DevBufs {
    uint64_t *slots = nullptr;
    uint32_t *ctl = nullptr, *seen = nullptr, *counts = nullptr;
    ~DevBufs() {
        (void)hipFree(slots);
        (void)hipFree(ctl);
        (void)hipFree(seen);
        (void)hipFree(counts);
    }
};

#define TRY(x)                                   \
    do {                                         \
        hipError_t e_ = (x);                     \
        if (e_ != hipSuccess) return (int)e_;    \
    } while (0)

int make_queue(DevBufs &d, L2Queue &q, uint32_t nbuckets, uint32_t window, uint32_t cap_log2,
               uint32_t delta, uint32_t claim, uint32_t base, uint32_t spin_limit) {
    const size_t nslots = (size_t)nbuckets << cap_log2;
    TRY(hipMalloc(&d.slots, nslots * sizeof(uint64_t)));
    TRY(hipMemset(d.slots, 0xFF, nslots * sizeof(uint64_t)));   // kEmpty
    std::vector<uint32_t> ctl(kCtlWords, 0u);
    ctl[kBaseOff] = base;
    TRY(hipMalloc(&d.ctl, kCtlWords * sizeof(uint32_t)));
    TRY(hipMemcpy(d.ctl, ctl.data(), kCtlWords * sizeof(uint32_t), hipMemcpyHostToDevice));
    q.slots = d.slots;
    q.ctl = d.ctl;
    q.nbuckets = nbuckets;
    q.window = window;
    q.cap_log2 = cap_log2;
    q.delta = delta;
    q.batch = claim;
    q.spin_limit = spin_limit;
    q.strict = 0;
    return 0;
}

}  // namespace

extern "C" {

// Control-block layout, for reading the counters back:
// {ctl words, tail, head, taken, base, err, stride}.
void l2check_layout(uint32_t *out) {
    const uint32_t v[7] = {kCtlWords, kTailOff, kHeadOff, kTakenOff, kBaseOff, kErrOff, kStride};
    std::memcpy(out, v, sizeof v);
}

// `producers` waves push `batch` records each (ids 0 .. producers*batch-1,
// distances dist_of(id)) while `consumers` waves pop; the window epoch starts
// at `base`.  Out: seen[id] (times each id was popped intact), ctl (the
// control block afterwards), counts {bad records, consumed, failed pushes}.
int l2check_publication(uint32_t nbuckets, uint32_t window, uint32_t cap_log2, uint32_t delta,
                        uint32_t claim, uint32_t base, uint32_t producers, uint32_t batch,
                        uint32_t consumers, uint32_t dist0, uint32_t span, uint32_t *seen_out,
                        uint32_t *ctl_out, uint32_t *counts_out) {
    if (nbuckets < 1 || nbuckets > kMaxBuckets || window < 1 || window > nbuckets ||
        cap_log2 < 6 || cap_log2 > 20 || delta == 0 || claim < 1 || claim > 64 || batch < 1 ||
        batch > 64 || producers < 1 || consumers < 1 || producers + consumers > 1024 ||
        span == 0 || (uint64_t)dist0 + span >= 0x7FFFFFFFull)
        return -1;
    const uint32_t total = producers * batch;
    DevBufs d;
    L2Queue q;
    if (int rc = make_queue(d, q, nbuckets, window, cap_log2, delta, claim, base, 1u << 16))
        return rc;
    TRY(hipMalloc(&d.seen, total * sizeof(uint32_t)));
    TRY(hipMemset(d.seen, 0, total * sizeof(uint32_t)));
    TRY(hipMalloc(&d.counts, 3 * sizeof(uint32_t)));
    TRY(hipMemset(d.counts, 0, 3 * sizeof(uint32_t)));
    hipLaunchKernelGGL(publication_kernel, dim3(producers + consumers), dim3(kWave), 0, 0, q,
                       producers, batch, dist0, span, total, 1u << 20, d.seen, d.counts);
    TRY(hipGetLastError());
    TRY(hipDeviceSynchronize());
    TRY(hipMemcpy(seen_out, d.seen, total * sizeof(uint32_t), hipMemcpyDeviceToHost));
    TRY(hipMemcpy(ctl_out, d.ctl, kCtlWords * sizeof(uint32_t), hipMemcpyDeviceToHost));
    TRY(hipMemcpy(counts_out, d.counts, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return 0;
}

// One bucket of 2^cap_log2 slots, no consumer: push `first` records (ids
// 0 .. first-1), snapshot slots and control block, then push `extra` more.
// Out: slots and ctl after each launch, failed-push counts of each launch.
int l2check_capacity(uint32_t cap_log2, uint32_t first, uint32_t extra, uint64_t *slots1,
                     uint32_t *ctl1, uint64_t *slots2, uint32_t *ctl2, uint32_t *fails) {
    if (cap_log2 < 6 || cap_log2 > 16 || first < 1 || first > (1u << cap_log2) || extra > 1024)
        return -1;
    const size_t nslots = (size_t)1 << cap_log2;
    DevBufs d;
    L2Queue q;
    if (int rc = make_queue(d, q, 1, 1, cap_log2, 1, 8, 0, 64)) return rc;
    TRY(hipMalloc(&d.counts, 2 * sizeof(uint32_t)));
    TRY(hipMemset(d.counts, 0, 2 * sizeof(uint32_t)));
    hipLaunchKernelGGL(fill_kernel, dim3((first + kWave - 1) / kWave), dim3(kWave), 0, 0, q, 0u,
                       first, d.counts);
    TRY(hipGetLastError());
    TRY(hipDeviceSynchronize());
    TRY(hipMemcpy(slots1, d.slots, nslots * sizeof(uint64_t), hipMemcpyDeviceToHost));
    TRY(hipMemcpy(ctl1, d.ctl, kCtlWords * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (extra) {
        hipLaunchKernelGGL(fill_kernel, dim3((extra + kWave - 1) / kWave), dim3(kWave), 0, 0, q,
                           first, extra, d.counts + 1);
        TRY(hipGetLastError());
        TRY(hipDeviceSynchronize());
    }
    TRY(hipMemcpy(slots2, d.slots, nslots * sizeof(uint64_t), hipMemcpyDeviceToHost));
    TRY(hipMemcpy(ctl2, d.ctl, kCtlWords * sizeof(uint32_t), hipMemcpyDeviceToHost));
    TRY(hipMemcpy(fails, d.counts, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return 0;
}

}  // extern "C"
