// Serving loop: resident graphs, admission of requests by the HBM bytes of
// their slot (distance row + queue partition), continuous batching over
// concurrent launch lanes.  Contract in include/mlmq_serve.h.
#include "../../include/mlmq_serve.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "arena.h"
#include "device_common.h"
#include "solver_internal.h"

namespace {

using clk = std::chrono::steady_clock;
using mlmq::Arena;
using mlmq::round_up;


uint32_t ceil_log2(uint64_t x) {
    uint32_t k = 0;
    while ((1ull << k) < x) ++k;
    return k;
}

struct Slot {
    uint64_t off = 0;
    uint64_t bytes = 0;
    bool fresh = true;   // queue partition content unknown
};

struct GraphEntry {
    mlmq_ctx *ctx = nullptr;
    mlmq::DeviceGraphView v{};
    mlmq_serve_graph_params p{};
    uint32_t cap_log2 = 0;
    uint64_t row_bytes = 0, ctl_bytes = 0, slot_bytes = 0;
    std::vector<Slot> free_slots;
    std::unordered_map<int32_t, uint32_t *> expected;
};

struct Lane {
    hipStream_t s = nullptr;
    hipEvent_t ev_solve = nullptr, ev_done = nullptr;
    uint8_t *h_reqs = nullptr;   // pinned request table (serve_request_bytes() per entry)
    uint8_t *d_reqs = nullptr;
    bool busy = false;
    int32_t launch = -1;
    std::vector<int32_t> ids;
    std::vector<Slot> slots;
    std::vector<uint32_t> caps;   // ring capacity (log2) each request ran with
};

// ---- live serving state ----
constexpr uint32_t kLiveEntries = 4096;     // request table entries
constexpr uint32_t kLiveRing = 1u << 16;    // ticket ring (power of two)
constexpr uint32_t kLiveMaxGrid = 4096;     // persistent workgroups at most

struct LiveReq {
    int32_t graph = 0, src = 0;
    int32_t *dist_out = nullptr;
    double arrival_ms = 0;     // live clock
    double publish_ms = 0;
    int32_t retries = 0;
    int32_t trace_index = -1;  // mlmq_server_run: index into its request array
    Slot slot;
    uint32_t cap_log2 = 0;
    int32_t entry = -1;
    uint32_t first = 0, parts = 0;
    bool verified = false;
    bool done = false;
    mlmq_serve_result res{};
};

struct Live {
    std::mutex mu;                        // guards everything below
    std::condition_variable cv;           // a request finished
    std::thread thread;
    bool thread_on = false, quit = false, thread_done = false;
    int64_t next_id = 0;
    std::deque<int64_t> pending;          // admitted in this order
    std::unordered_map<int64_t, LiveReq> reqs;
    std::vector<int32_t> inflight;        // entry indices published and not finished
    std::vector<int64_t> entry_req;       // request id of each entry
    std::vector<int32_t> free_entries;
    // device and pinned resources (created with the first launch)
    bool ready = false, running = false;
    hipStream_t ks = nullptr, side = nullptr;
    hipEvent_t ev_end = nullptr;
    uint32_t *hctl = nullptr;             // pinned host control words
    uint8_t *entries = nullptr;           // pinned host entry table
    uint32_t *ring = nullptr;             // pinned host ticket ring
    uint32_t *dev = nullptr;              // device counters
    void *sync = nullptr;                 // device per-entry state
    size_t entry_bytes = 0;
    uint32_t published = 0, grid = 0, parts_inflight = 0, launches = 0, seq = 0;
    double idle_since = 0;
    clk::time_point t0 = clk::now();
    uint64_t peak_inflight = 0;
};

}  // namespace

struct mlmq_server {
    int device = 0;
    int cus = 0;
    int wall_khz = 100000;
    mlmq_server_config cfg{};
    Arena arena;
    std::vector<GraphEntry> graphs;
    std::vector<Lane> lanes;
    std::deque<int32_t> pending;   // waiting request ids, in arrival order
    hipEvent_t ev_ref = nullptr;
    Live live;
    std::string err;
};

namespace {

// Most one-wave workgroups a spread request gets (the single-query solver's
// best wave count on the largest road graph).
constexpr uint32_t kMaxSpread = 512;

int fail(mlmq_server *s, int code, const std::string &msg) {
    if (s) s->err = msg;
    return code;
}

int hip_fail(mlmq_server *s, hipError_t e, const char *what) {
    return fail(s, e == hipErrorOutOfMemory ? MLMQ_E_NOMEM : MLMQ_E_HIP,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(srv, expr)                                     \
    do {                                                       \
        hipError_t e_ = (expr);                                \
        if (e_ != hipSuccess) return hip_fail(srv, e_, #expr); \
    } while (0)

// Defaults and range checks of per-graph serving parameters.  min_read must
// be at least 1 (a wave tops up from L2 only below min_read records, so 0
// would leave the seed untaken until the watchdog); 0 and -1 select 8.
int resolve_params_why(const mlmq_serve_graph_params *in, int32_t n, int64_t m,
                       mlmq_serve_graph_params *out, uint32_t *cap_log2, std::string *why) {
    auto fail = [&](int code, const char *msg) {
        if (why) *why = msg;
        return code;
    };
    mlmq_serve_graph_params p;
    if (in) p = *in;
    else mlmq_serve_graph_params_default(&p);
    if (p.delta == 0) p.delta = 200000;
    if (p.delta > 0x7FFFFFFFu) return fail(MLMQ_E_INVALID, "delta must be in [1, INT_MAX]");
    if (p.buckets == 0) p.buckets = 16;
    if (p.bucket_window == 0) p.bucket_window = std::min(4, p.buckets);
    if (p.l2_batch == 0) p.l2_batch = 8;
    if (p.min_read <= 0) p.min_read = 8;
    if (p.l1_pop == 0) p.l1_pop = 32;
    if (p.l1_spill == 0) p.l1_spill = 64;
    if (p.waves == 0) p.waves = 8;
    if (p.workgroups == 0) {
        const double k = std::round((double)m / 4.0e6);
        p.workgroups = (int32_t)std::min(16.0, std::max(1.0, k));
    }
    if (p.buckets < 1 || p.buckets > (int)mlmq::kMaxBuckets) return fail(MLMQ_E_INVALID, "buckets must be in [1, 64]");
    if (p.bucket_window < 1 || p.bucket_window > p.buckets) return fail(MLMQ_E_INVALID, "bucket_window must be in [1, buckets]");
    if (p.l2_batch < 1 || p.l2_batch > 64) return fail(MLMQ_E_INVALID, "l2_batch must be in [1, 64]");
    if (p.min_read > 64) return fail(MLMQ_E_INVALID, "min_read must be in [1, 64]");
    if (p.l1_pop < 1 || p.l1_pop > 64) return fail(MLMQ_E_INVALID, "l1_pop must be in [1, 64]");
    if (p.l1_spill < 1 || p.l1_spill > 64) return fail(MLMQ_E_INVALID, "l1_spill must be in [1, 64]");
    if (p.workgroups < 1 || p.workgroups > 64) return fail(MLMQ_E_INVALID, "workgroups must be in [1, 64]");
    if (p.waves != 1 && p.waves != 4 && p.waves != 8 && p.waves != 16)
        return fail(MLMQ_E_INVALID, "waves must be 1, 4, 8 or 16");
    if (p.bucket_capacity < 0) return fail(MLMQ_E_INVALID, "negative bucket_capacity");
    if (p.hub_split != 0 && p.hub_split != 1) return fail(MLMQ_E_INVALID, "hub_split must be 0 or 1");
    if (p.light_split != 0 && p.light_split != 1) return fail(MLMQ_E_INVALID, "light_split must be 0 or 1");
    if (p.prefetch < -1 || p.prefetch > 1) return fail(MLMQ_E_INVALID, "prefetch must be -1, 0 or 1");
    if (p.stale_check < -1 || p.stale_check > 3)
        return fail(MLMQ_E_INVALID, "stale_check must be -1, 0, 1, 2 or 3");
    uint64_t cap = (uint64_t)p.bucket_capacity;
    if (cap == 0) cap = std::max<uint64_t>((uint64_t)n / 32, 8192);
    const uint32_t c = ceil_log2(cap);
    if (c > 30) return fail(MLMQ_E_INVALID, "bucket capacity too large");
    *out = p;
    *cap_log2 = c;
    return MLMQ_OK;
}

int resolve_params(mlmq_server *s, const mlmq_serve_graph_params *in, int32_t n, int64_t m,
                   mlmq_serve_graph_params *out, uint32_t *cap_log2) {
    std::string why;
    const int rc = resolve_params_why(in, n, m, out, cap_log2, &why);
    if (rc) s->err = why;
    return rc;
}

void size_slots(GraphEntry &g) {
    g.row_bytes = round_up(sizeof(uint32_t) * (uint64_t)g.v.n);
    g.ctl_bytes = round_up(sizeof(uint32_t) * mlmq::kCtlWords);
    g.slot_bytes = g.row_bytes + g.ctl_bytes + sizeof(uint64_t) * ((uint64_t)g.p.buckets << g.cap_log2);
    g.slot_bytes = round_up(g.slot_bytes);
}

void drop_free_slots(mlmq_server *s, GraphEntry &g) {
    for (const Slot &sl : g.free_slots) s->arena.release(sl.off, sl.bytes);
    g.free_slots.clear();
}

// A slot for one request of graph `gi`: a free slot of the graph, else a new
// one carved from the arena, giving back other graphs' free slots if needed.
bool take_slot(mlmq_server *s, int32_t gi, Slot *out) {
    GraphEntry &g = s->graphs[gi];
    if (!g.free_slots.empty()) {
        *out = g.free_slots.back();
        g.free_slots.pop_back();
        return true;
    }
    uint64_t off = 0;
    if (!s->arena.alloc(g.slot_bytes, &off)) {
        for (size_t j = 0; j < s->graphs.size(); ++j)
            if ((int32_t)j != gi) drop_free_slots(s, s->graphs[j]);
        if (!s->arena.alloc(g.slot_bytes, &off)) return false;
    }
    out->off = off;
    out->bytes = g.slot_bytes;
    out->fresh = true;
    return true;
}

void give_back(mlmq_server *s, int32_t gi, const Slot &sl) {
    GraphEntry &g = s->graphs[gi];
    if (sl.bytes == g.slot_bytes) g.free_slots.push_back(sl);
    else s->arena.release(sl.off, sl.bytes);
}

int device_status(uint32_t e) {
    switch (e) {
        case mlmq::kErrNone: return MLMQ_OK;
        case mlmq::kErrOverflow: return MLMQ_E_QUEUE_OVERFLOW;
        case mlmq::kErrTimeout: return MLMQ_E_TIMEOUT;
        default: return MLMQ_E_INTERNAL;
    }
}

double pct(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    size_t k = (size_t)std::ceil(q * (double)v.size());
    if (k < 1) k = 1;
    if (k > v.size()) k = v.size();
    return v[k - 1];
}

// Kernel settings of one request of graph g (its slot and ring capacity are
// filled in per request).
mlmq::SolverLaunch graph_launch(const GraphEntry &g, uint64_t timeout_ticks) {
    mlmq::SolverLaunch L{};
    L.row = g.v.row;
    L.col = g.v.col;
    L.wgt = g.v.wgt;
    L.blocks = g.v.blocks;
    L.hubs = g.v.hubs;
    L.packed = g.v.packed;
    L.split = g.p.hub_split != 0;
    L.stale_check = g.p.stale_check == 0 ? 2u : g.p.stale_check < 0 ? 0u : (uint32_t)g.p.stale_check;
    L.prefetch = g.p.prefetch >= 0 ? 1u : 0u;
    L.hot = false;
    L.defer_heavy = 0;
    L.strict_window = 0;
    L.light_split = g.p.light_split ? 1u : 0u;
    L.n = (uint32_t)g.v.n;
    L.nbuckets = (uint32_t)g.p.buckets;
    L.window = (uint32_t)g.p.bucket_window;
    L.cap_log2 = g.cap_log2;
    L.delta = g.p.delta;
    L.batch = (uint32_t)g.p.l2_batch;
    L.spin_limit = mlmq::kSpinLimit;
    L.min_read = (uint32_t)g.p.min_read;
    L.max_sleep = 16;
    L.pop_max = (uint32_t)g.p.l1_pop;
    L.spill = (uint32_t)g.p.l1_spill;
    L.donate_min = 0;
    L.timeout_ticks = timeout_ticks;
    L.slf = true;
    L.fifo = false;
    L.profile = false;
    return L;
}

int add_graph_ctx(mlmq_server *s, mlmq_ctx *ctx, const mlmq_serve_graph_params *p, int32_t *graph_id) {
    GraphEntry g;
    g.ctx = ctx;
    int rc = mlmq::ctx_graph_view(ctx, &g.v);
    if (rc == MLMQ_OK) rc = resolve_params(s, p, g.v.n, g.v.m, &g.p, &g.cap_log2);
    if (rc == MLMQ_OK && g.p.light_split) {
        rc = mlmq::ctx_prepare_light(ctx, g.p.delta);
        if (rc) s->err = mlmq_last_error(ctx);
    }
    if (rc != MLMQ_OK) {
        mlmq_destroy(ctx);
        return rc;
    }
    size_slots(g);
    s->graphs.push_back(std::move(g));
    if (graph_id) *graph_id = (int32_t)s->graphs.size() - 1;
    return MLMQ_OK;
}

double live_now(const Live &L) {
    return std::chrono::duration<double, std::milli>(clk::now() - L.t0).count();
}

// Pinned tables, device counters and streams of live serving (once).
int live_prepare(mlmq_server *s) {
    Live &L = s->live;
    if (L.ready) return MLMQ_OK;
    L.entry_bytes = mlmq::stream_entry_bytes();
    HIP_TRY(s, hipSetDevice(s->device));
    HIP_TRY(s, hipStreamCreateWithFlags(&L.ks, hipStreamNonBlocking));
    HIP_TRY(s, hipStreamCreateWithFlags(&L.side, hipStreamNonBlocking));
    HIP_TRY(s, hipEventCreateWithFlags(&L.ev_end, hipEventDisableTiming));
    const unsigned flags = hipHostMallocCoherent | hipHostMallocMapped;
    HIP_TRY(s, hipHostMalloc((void **)&L.hctl, sizeof(uint32_t) * mlmq::kStreamHostCtlWords, flags));
    HIP_TRY(s, hipHostMalloc((void **)&L.entries, L.entry_bytes * kLiveEntries, flags));
    HIP_TRY(s, hipHostMalloc((void **)&L.ring, sizeof(uint32_t) * kLiveRing, flags));
    std::memset(L.hctl, 0, sizeof(uint32_t) * mlmq::kStreamHostCtlWords);
    std::memset(L.entries, 0, L.entry_bytes * kLiveEntries);
    std::memset(L.ring, 0, sizeof(uint32_t) * kLiveRing);
    HIP_TRY(s, hipMalloc((void **)&L.dev, 64));
    HIP_TRY(s, hipMalloc(&L.sync, mlmq::stream_sync_bytes() * kLiveEntries));
    int occ = 0;
    HIP_TRY(s, mlmq::stream_occupancy(&occ));
    uint32_t grid = (uint32_t)std::max(occ, 1) * (uint32_t)s->cus;
    if (s->cfg.live_waves > 0) grid = std::min(grid, (uint32_t)s->cfg.live_waves);
    L.grid = std::max<uint32_t>(2, std::min(grid, kLiveMaxGrid));
    L.entry_req.assign(kLiveEntries, -1);
    L.free_entries.clear();
    for (int32_t e = (int32_t)kLiveEntries - 1; e >= 0; --e) L.free_entries.push_back(e);
    L.ready = true;
    return MLMQ_OK;
}

void live_release(mlmq_server *s) {
    Live &L = s->live;
    if (L.ks) (void)hipStreamSynchronize(L.ks);
    if (L.side) (void)hipStreamSynchronize(L.side);
    if (L.hctl) (void)hipHostFree(L.hctl);
    if (L.entries) (void)hipHostFree(L.entries);
    if (L.ring) (void)hipHostFree(L.ring);
    if (L.dev) (void)hipFree(L.dev);
    if (L.sync) (void)hipFree(L.sync);
    if (L.ev_end) (void)hipEventDestroy(L.ev_end);
    if (L.ks) (void)hipStreamDestroy(L.ks);
    if (L.side) (void)hipStreamDestroy(L.side);
    L.hctl = nullptr;
    L.entries = nullptr;
    L.ring = nullptr;
    L.dev = nullptr;
    L.sync = nullptr;
    L.ev_end = nullptr;
    L.ks = L.side = nullptr;
    L.ready = false;
}

void store_host(uint32_t *p, uint32_t v) {
    __atomic_store_n(p, v, __ATOMIC_RELEASE);
}

// Start the persistent launch: device counters and per-entry state zeroed,
// host control words reset, tickets counted from 0.
int live_launch(mlmq_server *s) {
    Live &L = s->live;
    int rc = live_prepare(s);
    if (rc) return rc;
    HIP_TRY(s, hipMemsetAsync(L.dev, 0, 64, L.ks));
    HIP_TRY(s, hipMemsetAsync(L.sync, 0, mlmq::stream_sync_bytes() * kLiveEntries, L.ks));
    for (size_t i = 0; i < mlmq::kStreamHostCtlWords; ++i) store_host(L.hctl + i, 0u);
    L.published = 0;
    mlmq::StreamLaunch S;
    S.host_ctl = L.hctl;
    S.entries = L.entries;
    S.ring = L.ring;
    S.ring_size = kLiveRing;
    S.dev = L.dev;
    S.sync = L.sync;
    // a bound on the whole launch: it ends (and is started again on demand)
    // after ten minutes even if it is never stopped
    S.lifetime_ticks = (uint64_t)(600.0 * (double)s->wall_khz * 1000.0);
    S.grid = L.grid;
    HIP_TRY(s, mlmq::launch_stream(S, L.ks));
    HIP_TRY(s, hipEventRecord(L.ev_end, L.ks));
    L.running = true;
    ++L.launches;
    L.idle_since = live_now(L);
    return MLMQ_OK;
}

// A finished entry: its request's result, or a rerun after an overflow.
int live_finish(mlmq_server *s, int32_t e, uint32_t status, uint32_t mism, uint64_t t0, uint64_t t1) {
    Live &L = s->live;
    const int64_t id = L.entry_req[e];
    LiveReq &r = L.reqs[id];
    GraphEntry &g = s->graphs[r.graph];
    L.entry_req[e] = -1;
    L.free_entries.push_back(e);
    L.parts_inflight -= r.parts;
    r.entry = -1;
    Slot sl = r.slot;
    const int st = device_status(status);
    if (st == MLMQ_E_QUEUE_OVERFLOW && g.cap_log2 < 30) {
        if (g.cap_log2 == r.cap_log2) {
            drop_free_slots(s, g);
            ++g.cap_log2;
            size_slots(g);
        }
        sl.fresh = true;
        give_back(s, r.graph, sl);
        ++r.retries;
        L.pending.push_front(id);
        return MLMQ_OK;
    }
    if (st == MLMQ_OK && r.dist_out) {
        HIP_TRY(s, hipMemcpyAsync(r.dist_out, s->arena.base + sl.off, sizeof(int32_t) * (size_t)g.v.n,
                                  hipMemcpyDeviceToHost, L.side));
        HIP_TRY(s, hipStreamSynchronize(L.side));
    }
    r.res.status = st;
    r.res.retries = r.retries;
    r.res.latency_ms = live_now(L) - r.arrival_ms;
    r.res.queue_ms = r.publish_ms - r.arrival_ms;
    r.res.device_ms = t1 > t0 ? (double)(t1 - t0) / (double)s->wall_khz : 0.0;
    r.res.mismatches = r.verified ? (int32_t)mism : -1;
    r.res.launch = (int32_t)L.launches - 1;
    r.res.batch_size = (int32_t)r.parts;   // workgroups that served it
    sl.fresh = st != MLMQ_OK;
    give_back(s, r.graph, sl);
    r.done = true;
    L.cv.notify_all();
    return MLMQ_OK;
}

// Fail a request that cannot run (no slot will ever fit).
void live_fail(mlmq_server *s, int64_t id, int code) {
    Live &L = s->live;
    LiveReq &r = L.reqs[id];
    r.res.status = code;
    r.res.mismatches = -1;
    r.res.latency_ms = live_now(L) - r.arrival_ms;
    r.done = true;
    L.cv.notify_all();
}

// Stop the persistent launch: serve what is published, then wait for it.
int live_halt(mlmq_server *s) {
    Live &L = s->live;
    if (!L.running) return MLMQ_OK;
    store_host(L.hctl + 1, 1u);
    HIP_TRY(s, hipEventSynchronize(L.ev_end));
    L.running = false;
    return MLMQ_OK;
}

// One pass of the live loop (caller holds L.mu).  Returns <0 on a HIP error,
// 1 if anything happened, 0 if not.
int live_step(mlmq_server *s) {
    Live &L = s->live;
    int acted = 0;
    // completions
    for (size_t i = 0; i < L.inflight.size();) {
        const int32_t e = L.inflight[i];
        uint32_t done = 0, status = 0, mism = 0;
        uint64_t t0 = 0, t1 = 0;
        mlmq::stream_entry_read(L.entries + L.entry_bytes * (size_t)e, &done, &status, &mism, &t0, &t1);
        if (!done) {
            ++i;
            continue;
        }
        L.inflight[i] = L.inflight.back();
        L.inflight.pop_back();
        const int rc = live_finish(s, e, status, mism, t0, t1);
        if (rc) return rc;
        acted = 1;
    }
    // the launch ended by itself (its lifetime): what it did not start is
    // published again to the next launch
    if (L.running && hipEventQuery(L.ev_end) == hipSuccess) {
        L.running = false;
        const uint32_t served = __atomic_load_n(L.hctl + 2, __ATOMIC_ACQUIRE);
        for (size_t i = 0; i < L.inflight.size();) {
            const int32_t e = L.inflight[i];
            const int64_t id = L.entry_req[e];
            LiveReq &r = L.reqs[id];
            uint32_t done = 0, status = 0, mism = 0;
            uint64_t t0 = 0, t1 = 0;
            mlmq::stream_entry_read(L.entries + L.entry_bytes * (size_t)e, &done, &status, &mism, &t0, &t1);
            if (done) {
                ++i;
                continue;   // finished after all: collected on the next pass
            }
            L.inflight[i] = L.inflight.back();
            L.inflight.pop_back();
            L.entry_req[e] = -1;
            L.free_entries.push_back(e);
            L.parts_inflight -= r.parts;
            Slot sl = r.slot;
            if ((int32_t)(r.first - served) >= 0) {
                give_back(s, r.graph, sl);     // never touched by the device
                L.pending.push_front(id);
            } else {
                sl.fresh = true;               // started but unfinished
                give_back(s, r.graph, sl);
                live_fail(s, id, MLMQ_E_TIMEOUT);
            }
        }
        acted = 1;
    }
    if (!L.pending.empty() && !L.running) {
        const int rc = live_launch(s);
        if (rc) return rc;
        acted = 1;
    }
    // admission and publication, in queue order
    const uint64_t timeout_ticks = (uint64_t)(s->cfg.timeout_s * (double)s->wall_khz * 1000.0);
    while (L.running && !L.pending.empty() && !L.free_entries.empty()) {
        const int64_t id = L.pending.front();
        LiveReq &r = L.reqs[id];
        GraphEntry &g = s->graphs[r.graph];
        const uint32_t cap = L.grid - 1;
        uint32_t parts = std::min<uint32_t>((uint32_t)g.p.workgroups * (uint32_t)g.p.waves, cap);
        if (s->cfg.spread_waves > 0) {
            // few requests running: spread this one over the idle workgroups
            const uint32_t idle = cap > L.parts_inflight ? cap - L.parts_inflight : 0;
            const uint32_t share = std::min<uint32_t>(
                std::min<uint32_t>((uint32_t)s->cfg.spread_waves, kMaxSpread),
                idle / (uint32_t)std::max<size_t>(1, L.pending.size()));
            parts = std::max(parts, std::min(share, cap));
        }
        uint32_t oldest = L.published;
        for (int32_t e : L.inflight) oldest = std::min(oldest, L.reqs[L.entry_req[e]].first);
        if (L.published + parts - oldest > kLiveRing) break;
        Slot sl;
        if (!take_slot(s, r.graph, &sl)) {
            if (L.inflight.empty()) {   // nothing runs and it does not fit: it never will
                L.pending.pop_front();
                live_fail(s, id, MLMQ_E_NOMEM);
                acted = 1;
                continue;
            }
            break;
        }
        const int32_t e = L.free_entries.back();
        L.free_entries.pop_back();
        uint8_t *b = s->arena.base + sl.off;
        mlmq::SolverLaunch Ls = graph_launch(g, timeout_ticks);
        Ls.dist = (uint32_t *)b;
        Ls.ctl = (uint32_t *)(b + g.row_bytes);
        Ls.slots = (uint64_t *)(b + g.row_bytes + g.ctl_bytes);
        const uint32_t *expect = nullptr;
        if (s->cfg.verify) {
            auto it = g.expected.find(r.src);
            if (it != g.expected.end()) expect = it->second;
        }
        mlmq::stream_entry_write(L.entries + L.entry_bytes * (size_t)e, Ls, (uint32_t)r.src, sl.fresh, expect,
                                 L.published, parts, ++L.seq);
        for (uint32_t j = 0; j < parts; ++j) L.ring[(L.published + j) & (kLiveRing - 1)] = (uint32_t)e;
        r.slot = sl;
        r.cap_log2 = g.cap_log2;
        r.entry = e;
        r.first = L.published;
        r.parts = parts;
        r.verified = expect != nullptr;
        r.publish_ms = live_now(L);
        L.published += parts;
        store_host(L.hctl, L.published);   // release: entry and ring words first
        L.entry_req[e] = id;
        L.inflight.push_back(e);
        L.parts_inflight += parts;
        L.peak_inflight = std::max<uint64_t>(L.peak_inflight, L.inflight.size());
        L.pending.pop_front();
        acted = 1;
    }
    // idle: end the launch
    const double now = live_now(L);
    if (acted || !L.pending.empty() || !L.inflight.empty()) {
        L.idle_since = now;
    } else if (L.running) {
        const double idle_ms = s->cfg.live_idle_ms > 0 ? s->cfg.live_idle_ms : 2000.0;
        if (now - L.idle_since > idle_ms) {
            const int rc = live_halt(s);
            if (rc) return rc;
        }
    }
    return acted;
}

void live_thread(mlmq_server *s) {
    Live &L = s->live;
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(L.mu);
    for (;;) {
        const int rc = live_step(s);
        if (rc < 0) {
            // a HIP error: every request that waits or runs fails with it
            for (auto &kv : L.reqs)
                if (!kv.second.done) live_fail(s, kv.first, rc);
            L.pending.clear();
            L.inflight.clear();
            break;
        }
        if (L.quit && L.pending.empty() && L.inflight.empty()) break;
        lk.unlock();
        if (rc == 0) std::this_thread::sleep_for(std::chrono::microseconds(L.inflight.empty() ? 200 : 20));
        lk.lock();
    }
    (void)live_halt(s);
    L.thread_done = true;
}

int live_stop_thread(mlmq_server *s) {
    Live &L = s->live;
    {
        std::lock_guard<std::mutex> lk(L.mu);
        if (!L.thread_on) return MLMQ_OK;
        L.quit = true;
    }
    L.thread.join();
    std::lock_guard<std::mutex> lk(L.mu);
    L.thread_on = false;
    L.thread_done = false;
    L.quit = false;
    return MLMQ_OK;
}

// A new request (caller holds L.mu); arrival on the live clock.
int64_t live_add(mlmq_server *s, int32_t graph, int32_t src, int32_t *dist_out, double arrival_ms) {
    Live &L = s->live;
    const int64_t id = L.next_id++;
    LiveReq &r = L.reqs[id];
    r.graph = graph;
    r.src = src;
    r.dist_out = dist_out;
    r.arrival_ms = arrival_ms;
    L.pending.push_back(id);
    return id;
}

void fill_stats(const mlmq_serve_result *results, int32_t nreq, uint64_t rejected, uint64_t launches,
                double wall_ms, double last_arrival_s, double batch_sum, uint64_t peak_bytes,
                uint64_t peak_inflight, mlmq_serve_stats *stats) {
    std::memset(stats, 0, sizeof(*stats));
    std::vector<double> lat, que;
    double dev = 0;
    for (int32_t i = 0; i < nreq; ++i) {
        const mlmq_serve_result &r = results[i];
        if (r.status == MLMQ_E_REJECTED) continue;
        if (r.status != MLMQ_OK) {
            ++stats->failed;
            continue;
        }
        ++stats->completed;
        lat.push_back(r.latency_ms);
        que.push_back(r.queue_ms);
        dev += r.device_ms;
        if (r.mismatches >= 0) {
            ++stats->verified;
            if (r.mismatches > 0) ++stats->mismatched;
        }
        stats->retries += (uint64_t)r.retries;
    }
    stats->requests = (uint64_t)nreq;
    stats->rejected = rejected;
    stats->launches = launches;
    stats->wall_s = wall_ms / 1000.0;
    stats->offered_qps = nreq > 0 && last_arrival_s > 0 ? (double)nreq / last_arrival_s : 0.0;
    stats->completed_qps = wall_ms > 0 ? (double)stats->completed / stats->wall_s : 0.0;
    stats->p50_ms = pct(lat, 0.50);
    stats->p99_ms = pct(lat, 0.99);
    stats->max_ms = lat.empty() ? 0.0 : *std::max_element(lat.begin(), lat.end());
    double sum = 0;
    for (double x : lat) sum += x;
    stats->mean_ms = lat.empty() ? 0.0 : sum / (double)lat.size();
    stats->queue_p50_ms = pct(que, 0.50);
    stats->queue_p99_ms = pct(que, 0.99);
    stats->mean_batch = launches ? batch_sum / (double)launches : 0.0;
    stats->mean_device_ms = stats->completed ? dev / (double)stats->completed : 0.0;
    stats->peak_bytes = peak_bytes;
    stats->peak_inflight = peak_inflight;
}

// mlmq_server_run through the live path, on the calling thread.
int run_live(mlmq_server *s, const mlmq_serve_request *reqs, int32_t nreq, mlmq_serve_result *results,
             mlmq_serve_stats *stats) {
    int rc = live_stop_thread(s);
    if (rc) return rc;
    Live &L = s->live;
    std::unique_lock<std::mutex> lk(L.mu);
    L.reqs.clear();
    L.pending.clear();
    const uint64_t peak_before = s->arena.used;
    s->arena.peak = s->arena.used;
    L.peak_inflight = 0;
    const uint32_t launches0 = L.launches;
    const double t0 = live_now(L);
    std::vector<int64_t> ids((size_t)nreq, -1);
    int32_t next = 0, pending = 0;
    uint64_t rejected = 0;
    size_t finished = 0;
    double wall_ms = 0;
    while (rc == MLMQ_OK) {
        const double t = live_now(L) - t0;
        while (next < nreq && 1000.0 * reqs[next].arrival_s <= t) {
            pending = (int32_t)L.pending.size();
            if (s->cfg.max_pending > 0 && pending >= s->cfg.max_pending) {
                ++rejected;
            } else {
                ids[next] = live_add(s, reqs[next].graph, reqs[next].src, reqs[next].dist_out,
                                     t0 + 1000.0 * reqs[next].arrival_s);
                L.reqs[ids[next]].trace_index = next;
            }
            ++next;
        }
        const int st = live_step(s);
        if (st < 0) {
            rc = st;
            break;
        }
        finished = 0;
        for (int32_t i = 0; i < next; ++i)
            if (ids[i] >= 0 && L.reqs[ids[i]].done) ++finished;
        if (next == nreq && finished + rejected == (size_t)nreq) break;
        if (st == 0) {
            lk.unlock();
            if (L.inflight.empty() && L.pending.empty() && next < nreq) {
                const double wait = 1000.0 * reqs[next].arrival_s - (live_now(L) - t0);
                if (wait > 0.5)
                    std::this_thread::sleep_for(std::chrono::microseconds((int64_t)((wait - 0.3) * 1000.0)));
            } else {
                std::this_thread::yield();
            }
            lk.lock();
        }
    }
    {
        // no idle launch is left behind (the next run starts one again)
        const int hrc = live_halt(s);
        if (!rc) rc = hrc;
    }
    if (rc) return rc;
    for (int32_t i = 0; i < nreq; ++i) {
        mlmq_serve_result &r = results[i];
        if (ids[i] < 0) {
            std::memset(&r, 0, sizeof(r));
            r.status = MLMQ_E_REJECTED;
            r.mismatches = -1;
            r.launch = -1;
            continue;
        }
        r = L.reqs[ids[i]].res;
        wall_ms = std::max(wall_ms, L.reqs[ids[i]].arrival_ms - t0 + r.latency_ms);
    }
    L.reqs.clear();
    if (stats) {
        const double last = nreq ? reqs[nreq - 1].arrival_s : 0.0;
        fill_stats(results, nreq, rejected, L.launches - launches0, wall_ms, last, 0.0,
                   std::max(peak_before, s->arena.peak), L.peak_inflight, stats);
    }
    return MLMQ_OK;
}

}  // namespace

extern "C" {

void mlmq_server_config_default(mlmq_server_config *cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->hbm_budget = 32ull << 30;
    cfg->streams = 4;
    cfg->max_batch = 64;
    cfg->max_pending = 0;
    cfg->verify = 1;
    cfg->timeout_s = 60.0;
    cfg->spread_waves = 512;
    cfg->mode = MLMQ_SERVE_LANES;
    cfg->live_waves = 0;
    cfg->live_idle_ms = 2000;
}

void mlmq_serve_graph_params_default(mlmq_serve_graph_params *p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->delta = 200000;
    p->buckets = 16;
    p->bucket_window = 4;
    p->l2_batch = 8;
    p->min_read = 8;
    p->l1_pop = 32;
    p->l1_spill = 64;
    p->workgroups = 0;
    p->waves = 8;
    p->bucket_capacity = 0;
}

int mlmq_server_create(int device, const mlmq_server_config *cfg, mlmq_server **out) {
    if (!out) return MLMQ_E_INVALID;
    *out = nullptr;
    mlmq_server_config c;
    if (cfg) c = *cfg;
    else mlmq_server_config_default(&c);
    if (c.hbm_budget == 0) c.hbm_budget = 32ull << 30;
    if (c.streams == 0) c.streams = 4;
    if (c.max_batch == 0) c.max_batch = 64;
    if (c.timeout_s <= 0) c.timeout_s = 60.0;
    if (c.live_idle_ms == 0) c.live_idle_ms = 2000;
    if (c.streams < 1 || c.streams > 32 || c.max_batch < 1 || c.max_batch > 4096 || c.max_pending < 0 ||
        c.spread_waves < 0 || c.spread_waves > 65536 ||
        (c.mode != MLMQ_SERVE_LANES && c.mode != MLMQ_SERVE_LIVE) || c.live_waves < 0 || c.live_idle_ms < 0)
        return MLMQ_E_INVALID;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return MLMQ_E_INVALID;
    mlmq_server *s = new (std::nothrow) mlmq_server();
    if (!s) return MLMQ_E_NOMEM;
    s->device = device;
    s->cfg = c;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess) {
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
            s->wall_khz = khz;
    }
    void *arena = nullptr;
    if (e == hipSuccess) e = hipMalloc(&arena, c.hbm_budget);
    if (e == hipSuccess) s->arena.reset((uint8_t *)arena, c.hbm_budget);
    if (e == hipSuccess) e = hipEventCreate(&s->ev_ref);
    s->lanes.resize(c.streams);
    for (Lane &l : s->lanes) {
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&l.s, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreate(&l.ev_solve);
        if (e == hipSuccess) e = hipEventCreate(&l.ev_done);
        const size_t tb = mlmq::serve_request_bytes() * (size_t)c.max_batch;
        if (e == hipSuccess) e = hipHostMalloc((void **)&l.h_reqs, tb, hipHostMallocDefault);
        if (e == hipSuccess) e = hipMalloc((void **)&l.d_reqs, tb);
    }
    if (e != hipSuccess) {
        const int rc = e == hipErrorOutOfMemory ? MLMQ_E_NOMEM : MLMQ_E_HIP;
        mlmq_server_destroy(s);
        return rc;
    }
    *out = s;
    return MLMQ_OK;
}

int mlmq_server_destroy(mlmq_server *s) {
    if (!s) return MLMQ_OK;
    (void)live_stop_thread(s);
    (void)hipSetDevice(s->device);
    {
        std::lock_guard<std::mutex> lk(s->live.mu);
        (void)live_halt(s);
    }
    live_release(s);
    for (Lane &l : s->lanes) {
        if (l.s) (void)hipStreamSynchronize(l.s);
        if (l.ev_solve) (void)hipEventDestroy(l.ev_solve);
        if (l.ev_done) (void)hipEventDestroy(l.ev_done);
        if (l.h_reqs) (void)hipHostFree(l.h_reqs);
        if (l.d_reqs) (void)hipFree(l.d_reqs);
        if (l.s) (void)hipStreamDestroy(l.s);
    }
    for (GraphEntry &g : s->graphs) {
        for (auto &kv : g.expected) (void)hipFree(kv.second);
        mlmq_destroy(g.ctx);
    }
    if (s->arena.base) (void)hipFree(s->arena.base);
    if (s->ev_ref) (void)hipEventDestroy(s->ev_ref);
    delete s;
    return MLMQ_OK;
}

const char *mlmq_server_last_error(const mlmq_server *s) { return s ? s->err.c_str() : "null server"; }

int mlmq_serve_graph_params_resolve(const mlmq_serve_graph_params *in, int32_t n, int64_t m,
                                    mlmq_serve_graph_params *out) {
    if (!out || n <= 0 || m < 0) return MLMQ_E_INVALID;
    uint32_t c = 0;
    return resolve_params_why(in, n, m, out, &c, nullptr);
}

int mlmq_server_add_csr(mlmq_server *s, int32_t n, int64_t m, const int64_t *row, const int32_t *col,
                        const int32_t *w, const mlmq_serve_graph_params *p, int32_t *graph_id) {
    if (!s) return MLMQ_E_INVALID;
    mlmq_ctx *ctx = nullptr;
    int rc = mlmq_create(s->device, &ctx);
    if (rc) return fail(s, rc, "context creation failed");
    rc = mlmq_load_csr(ctx, n, m, row, col, w);
    if (rc) {
        s->err = mlmq_last_error(ctx);
        mlmq_destroy(ctx);
        return rc;
    }
    return add_graph_ctx(s, ctx, p, graph_id);
}

int mlmq_server_add_gr(mlmq_server *s, const char *path, const mlmq_serve_graph_params *p,
                       int32_t *graph_id) {
    if (!s || !path) return MLMQ_E_INVALID;
    mlmq_ctx *ctx = nullptr;
    int rc = mlmq_create(s->device, &ctx);
    if (rc) return fail(s, rc, "context creation failed");
    rc = mlmq_load_gr(ctx, path);
    if (rc) {
        s->err = mlmq_last_error(ctx);
        mlmq_destroy(ctx);
        return rc;
    }
    return add_graph_ctx(s, ctx, p, graph_id);
}

int mlmq_server_set_graph_params(mlmq_server *s, int32_t graph_id, const mlmq_serve_graph_params *p) {
    if (!s || graph_id < 0 || graph_id >= (int32_t)s->graphs.size()) return MLMQ_E_INVALID;
    GraphEntry &g = s->graphs[graph_id];
    mlmq_serve_graph_params np;
    uint32_t c = 0;
    int rc = resolve_params(s, p, g.v.n, g.v.m, &np, &c);
    if (rc) return rc;
    if (np.light_split) {
        rc = mlmq::ctx_prepare_light(g.ctx, np.delta);
        if (rc) return fail(s, rc, mlmq_last_error(g.ctx));
    }
    drop_free_slots(s, g);
    g.p = np;
    g.cap_log2 = c;
    size_slots(g);
    return MLMQ_OK;
}

int mlmq_server_graph_size(const mlmq_server *s, int32_t graph_id, int32_t *n, int64_t *m) {
    if (!s || graph_id < 0 || graph_id >= (int32_t)s->graphs.size()) return MLMQ_E_INVALID;
    if (n) *n = s->graphs[graph_id].v.n;
    if (m) *m = s->graphs[graph_id].v.m;
    return MLMQ_OK;
}

int mlmq_server_slot_bytes(const mlmq_server *s, int32_t graph_id, uint64_t *bytes) {
    if (!s || !bytes || graph_id < 0 || graph_id >= (int32_t)s->graphs.size()) return MLMQ_E_INVALID;
    *bytes = s->graphs[graph_id].slot_bytes;
    return MLMQ_OK;
}

int mlmq_server_set_expected(mlmq_server *s, int32_t graph_id, int32_t src, const int32_t *dist) {
    if (!s || !dist || graph_id < 0 || graph_id >= (int32_t)s->graphs.size()) return MLMQ_E_INVALID;
    GraphEntry &g = s->graphs[graph_id];
    if (src < 0 || src >= g.v.n) return fail(s, MLMQ_E_INVALID, "source out of range");
    HIP_TRY(s, hipSetDevice(s->device));
    uint32_t *&row = g.expected[src];
    if (!row) HIP_TRY(s, hipMalloc((void **)&row, sizeof(uint32_t) * (size_t)g.v.n));
    HIP_TRY(s, hipMemcpy(row, dist, sizeof(uint32_t) * (size_t)g.v.n, hipMemcpyHostToDevice));
    return MLMQ_OK;
}

int mlmq_server_clear_expected(mlmq_server *s) {
    if (!s) return MLMQ_E_INVALID;
    for (GraphEntry &g : s->graphs) {
        for (auto &kv : g.expected) (void)hipFree(kv.second);
        g.expected.clear();
    }
    return MLMQ_OK;
}

int mlmq_server_run(mlmq_server *s, const mlmq_serve_request *reqs, int32_t nreq,
                    mlmq_serve_result *results, mlmq_serve_stats *stats) {
    if (!s) return MLMQ_E_INVALID;
    if (nreq < 0 || (nreq > 0 && (!reqs || !results))) return fail(s, MLMQ_E_INVALID, "bad request array");
    double last = 0.0;
    for (int32_t i = 0; i < nreq; ++i) {
        const mlmq_serve_request &r = reqs[i];
        if (!std::isfinite(r.arrival_s) || r.arrival_s < last)
            return fail(s, MLMQ_E_INVALID, "arrival times must be finite and non-decreasing");
        last = r.arrival_s;
        if (r.graph < 0 || r.graph >= (int32_t)s->graphs.size())
            return fail(s, MLMQ_E_INVALID, "request names an unknown graph");
        if (r.src < 0 || r.src >= s->graphs[r.graph].v.n) return fail(s, MLMQ_E_INVALID, "source out of range");
    }
    HIP_TRY(s, hipSetDevice(s->device));
    if (s->cfg.mode == MLMQ_SERVE_LIVE) return run_live(s, reqs, nreq, results, stats);
    {
        // the launch lanes and live serving share the slot arena: no live launch may run meanwhile
        int rc0 = live_stop_thread(s);
        if (rc0) return rc0;
        std::lock_guard<std::mutex> lk(s->live.mu);
        rc0 = live_halt(s);
        if (rc0) return rc0;
    }
    for (int32_t i = 0; i < nreq; ++i) {
        mlmq_serve_result &r = results[i];
        std::memset(&r, 0, sizeof(r));
        r.status = MLMQ_E_REJECTED;
        r.mismatches = -1;
        r.launch = -1;
    }
    s->pending.clear();

    const uint64_t timeout_ticks = (uint64_t)(s->cfg.timeout_s * (double)s->wall_khz * 1000.0);
    std::vector<int32_t> retries(nreq, 0);
    int32_t next = 0, finished = 0, launches = 0, pending = 0;
    uint64_t inflight = 0, peak_inflight = 0, batch_sum = 0;
    const uint64_t peak_before = s->arena.used;
    s->arena.peak = s->arena.used;
    double wall_ms = 0.0;
    int rc = MLMQ_OK;
    const size_t rb = mlmq::serve_request_bytes();

    HIP_TRY(s, hipEventRecord(s->ev_ref, s->lanes[0].s));
    HIP_TRY(s, hipEventSynchronize(s->ev_ref));
    const clk::time_point t0 = clk::now();
    auto now_ms = [&]() { return std::chrono::duration<double, std::milli>(clk::now() - t0).count(); };

    auto launch_lane = [&](Lane &l) -> int {
        if (s->pending.empty()) return 0;
        // The oldest waiting request fixes the launch's workgroup width;
        // waiting requests of every graph with that width join it in arrival
        // order while a slot can be found for each (the first that finds no
        // slot ends the launch, so a large request is not overtaken forever).
        const int32_t width = s->graphs[reqs[s->pending.front()].graph].p.waves;
        l.ids.clear();
        l.slots.clear();
        l.caps.clear();
        for (auto it = s->pending.begin();
             it != s->pending.end() && (int32_t)l.ids.size() < s->cfg.max_batch;) {
            const int32_t id = *it;
            const int32_t gi = reqs[id].graph;
            if (s->graphs[gi].p.waves != width) {
                ++it;
                continue;
            }
            Slot sl;
            if (!take_slot(s, gi, &sl)) break;
            it = s->pending.erase(it);
            --pending;
            l.ids.push_back(id);
            l.slots.push_back(sl);
            l.caps.push_back(s->graphs[gi].cap_log2);
        }
        if (l.ids.empty()) {
            if (inflight > 0) return 0;   // bytes held by running requests: wait for them
            // Nothing runs and the oldest request's slot still does not fit: it never will.
            const int32_t id = s->pending.front();
            s->pending.pop_front();
            --pending;
            results[id].status = MLMQ_E_NOMEM;
            ++finished;
            return 1;
        }
        const uint32_t nb = (uint32_t)l.ids.size();
        // Few requests in the launch: spread each over more waves, up to
        // spread_waves for the whole launch, as one-wave workgroups (they go
        // to distinct CUs, as in the single-query solver).
        uint32_t widest = 0;
        for (int32_t id : l.ids) {
            const GraphEntry &g = s->graphs[reqs[id].graph];
            widest = std::max(widest, (uint32_t)g.p.workgroups * (uint32_t)width);
        }
        uint32_t fill = 0;
        if (s->cfg.spread_waves > 0) {
            fill = std::min<uint32_t>((uint32_t)s->cfg.spread_waves / nb, kMaxSpread);
            if (fill <= widest) fill = 0;
        }
        mlmq::ServeLaunch S{};
        S.reqs = l.d_reqs;
        S.nreq = nb;
        S.waves = fill ? 1 : width;
        bool any_expect = false;
        uint32_t wg_first = 0;
        for (uint32_t i = 0; i < nb; ++i) {
            const int32_t id = l.ids[i];
            const GraphEntry &g = s->graphs[reqs[id].graph];
            const Slot &sl = l.slots[i];
            uint8_t *b = s->arena.base + sl.off;
            mlmq::SolverLaunch L = graph_launch(g, timeout_ticks);
            L.cap_log2 = l.caps[i];
            L.dist = (uint32_t *)b;
            L.ctl = (uint32_t *)(b + g.row_bytes);
            L.slots = (uint64_t *)(b + g.row_bytes + g.ctl_bytes);
            const uint32_t *expect = nullptr;
            if (s->cfg.verify) {
                auto it = g.expected.find(reqs[id].src);
                if (it != g.expected.end()) expect = it->second;
            }
            any_expect |= expect != nullptr;
            const uint32_t wg = fill ? fill : (uint32_t)g.p.workgroups;
            mlmq::serve_request_write(l.h_reqs, i, L, (uint32_t)reqs[id].src, sl.fresh, expect, wg_first, wg);
            wg_first += wg;
            S.max_n = std::max(S.max_n, L.n);
            S.max_part_slots = std::max(S.max_part_slots, (uint64_t)L.nbuckets << L.cap_log2);
            S.ext = S.ext || L.light_split || L.defer_heavy || L.strict_window;
        }
        S.grid = wg_first;
        const size_t tb = rb * nb;
        HIP_TRY(s, hipMemcpyAsync(l.d_reqs, l.h_reqs, tb, hipMemcpyHostToDevice, l.s));
        HIP_TRY(s, mlmq::launch_serve(S, l.s));
        HIP_TRY(s, hipEventRecord(l.ev_solve, l.s));
        if (any_expect) HIP_TRY(s, mlmq::launch_serve_verify(S, l.s));
        HIP_TRY(s, hipMemcpyAsync(l.h_reqs, l.d_reqs, tb, hipMemcpyDeviceToHost, l.s));
        HIP_TRY(s, hipEventRecord(l.ev_done, l.s));
        const double t = now_ms();
        for (int32_t id : l.ids) {
            results[id].queue_ms = t - 1000.0 * reqs[id].arrival_s;
            results[id].launch = launches;
            results[id].batch_size = (int32_t)nb;
        }
        l.busy = true;
        l.launch = launches++;
        inflight += nb;
        peak_inflight = std::max(peak_inflight, inflight);
        batch_sum += nb;
        return 1;
    };

    auto complete_lane = [&](Lane &l) -> int {
        float done_ms = 0.0f;
        HIP_TRY(s, hipEventElapsedTime(&done_ms, s->ev_ref, l.ev_solve));
        std::vector<int32_t> rerun;   // overflowed requests, in launch order
        for (size_t i = 0; i < l.ids.size(); ++i) {
            const int32_t id = l.ids[i];
            const int32_t gi = reqs[id].graph;
            GraphEntry &g = s->graphs[gi];
            uint32_t status = 0, mism = 0;
            uint64_t ts = 0, te = 0;
            mlmq::serve_request_read(l.h_reqs, (uint32_t)i, &status, &mism, &ts, &te);
            Slot sl = l.slots[i];
            mlmq_serve_result &r = results[id];
            const int st = device_status(status);
            if (st == MLMQ_E_QUEUE_OVERFLOW && g.cap_log2 < 30) {
                // Grow the graph's partitions once per launch and run again.
                if (g.cap_log2 == l.caps[i]) {
                    drop_free_slots(s, g);
                    ++g.cap_log2;
                    size_slots(g);
                }
                sl.fresh = true;
                give_back(s, gi, sl);
                ++retries[id];
                rerun.push_back(id);
                continue;
            }
            r.status = st;
            r.retries = retries[id];
            r.latency_ms = (double)done_ms - 1000.0 * reqs[id].arrival_s;
            r.device_ms = te > ts ? (double)(te - ts) / (double)s->wall_khz : 0.0;
            const bool verified = s->cfg.verify && g.expected.count(reqs[id].src) != 0;
            r.mismatches = verified ? (int32_t)mism : -1;
            if (st == MLMQ_OK && reqs[id].dist_out) {
                HIP_TRY(s, hipMemcpyAsync(reqs[id].dist_out, s->arena.base + sl.off,
                                          sizeof(int32_t) * (size_t)g.v.n, hipMemcpyDeviceToHost, l.s));
                HIP_TRY(s, hipStreamSynchronize(l.s));
            }
            // A failed solve leaves records in its partition: refill before reuse.
            sl.fresh = st != MLMQ_OK;
            give_back(s, gi, sl);
            ++finished;
        }
        for (auto it = rerun.rbegin(); it != rerun.rend(); ++it) {
            s->pending.push_front(*it);
            ++pending;
        }
        inflight -= l.ids.size();
        l.busy = false;
        l.ids.clear();
        l.slots.clear();
        l.caps.clear();
        wall_ms = std::max(wall_ms, (double)done_ms);
        return MLMQ_OK;
    };

    uint64_t rejected = 0;
    while (finished + (int32_t)rejected < nreq && rc == MLMQ_OK) {
        bool acted = false;
        const double t = now_ms();
        while (next < nreq && 1000.0 * reqs[next].arrival_s <= t) {
            if (s->cfg.max_pending > 0 && pending >= s->cfg.max_pending) {
                ++rejected;   // result already says MLMQ_E_REJECTED
            } else {
                s->pending.push_back(next);
                ++pending;
            }
            ++next;
            acted = true;
        }
        for (Lane &l : s->lanes) {
            if (!l.busy) continue;
            const hipError_t q = hipEventQuery(l.ev_done);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) {
                rc = hip_fail(s, q, "hipEventQuery");
                break;
            }
            rc = complete_lane(l);
            if (rc) break;
            acted = true;
        }
        if (rc) break;
        for (Lane &l : s->lanes) {
            if (l.busy || pending == 0) continue;
            const int k = launch_lane(l);
            if (k < 0) {
                rc = k;
                break;
            }
            if (k == 0) break;
            acted = true;
        }
        if (rc) break;
        if (!acted) {
            // Nothing on the device and the next arrival is far: sleep until
            // shortly before it; otherwise keep polling.
            bool busy = false;
            for (const Lane &l : s->lanes) busy |= l.busy;
            if (!busy && pending == 0 && next < nreq) {
                const double wait = 1000.0 * reqs[next].arrival_s - now_ms();
                if (wait > 0.5)
                    std::this_thread::sleep_for(std::chrono::microseconds((int64_t)((wait - 0.3) * 1000.0)));
            } else {
                std::this_thread::yield();
            }
        }
    }
    if (rc) {
        for (Lane &l : s->lanes)
            if (l.busy) (void)hipStreamSynchronize(l.s);
        for (Lane &l : s->lanes) {
            if (!l.busy) continue;
            for (size_t i = 0; i < l.slots.size(); ++i) {
                Slot sl = l.slots[i];
                sl.fresh = true;
                give_back(s, reqs[l.ids[i]].graph, sl);
            }
            l.busy = false;
        }
        return rc;
    }

    if (stats)
        fill_stats(results, nreq, rejected, (uint64_t)launches, wall_ms, last, (double)batch_sum,
                   std::max(peak_before, s->arena.peak), peak_inflight, stats);
    return MLMQ_OK;
}

int mlmq_server_submit(mlmq_server *s, int32_t graph, int32_t src, int32_t *dist_out, int64_t *request_id) {
    if (!s || !request_id) return MLMQ_E_INVALID;
    if (graph < 0 || graph >= (int32_t)s->graphs.size()) return fail(s, MLMQ_E_INVALID, "unknown graph");
    if (src < 0 || src >= s->graphs[graph].v.n) return fail(s, MLMQ_E_INVALID, "source out of range");
    Live &L = s->live;
    std::lock_guard<std::mutex> lk(L.mu);
    if (s->cfg.max_pending > 0 && (int32_t)L.pending.size() >= s->cfg.max_pending)
        return fail(s, MLMQ_E_REJECTED, "too many requests waiting");
    if (L.thread_on && L.thread_done) {   // the service thread stopped on an error
        L.thread.join();
        L.thread_on = false;
        L.thread_done = false;
    }
    *request_id = live_add(s, graph, src, dist_out, live_now(L));
    if (!L.thread_on) {
        L.quit = false;
        L.thread = std::thread(live_thread, s);
        L.thread_on = true;
    }
    return MLMQ_OK;
}

int mlmq_server_poll(mlmq_server *s, int64_t request_id, mlmq_serve_result *result) {
    if (!s) return MLMQ_E_INVALID;
    Live &L = s->live;
    std::lock_guard<std::mutex> lk(L.mu);
    auto it = L.reqs.find(request_id);
    if (it == L.reqs.end()) return fail(s, MLMQ_E_INVALID, "unknown request id");
    if (!it->second.done) return MLMQ_E_PENDING;
    if (result) *result = it->second.res;
    L.reqs.erase(it);
    return MLMQ_OK;
}

int mlmq_server_wait(mlmq_server *s, int64_t request_id, double timeout_s, mlmq_serve_result *result) {
    if (!s) return MLMQ_E_INVALID;
    Live &L = s->live;
    std::unique_lock<std::mutex> lk(L.mu);
    auto done = [&]() {
        auto it = L.reqs.find(request_id);
        return it == L.reqs.end() || it->second.done;
    };
    if (timeout_s > 0) {
        if (!L.cv.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) return MLMQ_E_PENDING;
    } else {
        L.cv.wait(lk, done);
    }
    auto it = L.reqs.find(request_id);
    if (it == L.reqs.end()) return fail(s, MLMQ_E_INVALID, "unknown request id");
    if (result) *result = it->second.res;
    L.reqs.erase(it);
    return MLMQ_OK;
}

int mlmq_server_drain(mlmq_server *s, double timeout_s) {
    if (!s) return MLMQ_E_INVALID;
    Live &L = s->live;
    std::unique_lock<std::mutex> lk(L.mu);
    auto all_done = [&]() {
        for (auto &kv : L.reqs)
            if (!kv.second.done) return false;
        return true;
    };
    if (timeout_s > 0) {
        if (!L.cv.wait_for(lk, std::chrono::duration<double>(timeout_s), all_done)) return MLMQ_E_PENDING;
    } else {
        L.cv.wait(lk, all_done);
    }
    return MLMQ_OK;
}

int mlmq_server_stop(mlmq_server *s) {
    if (!s) return MLMQ_E_INVALID;
    return live_stop_thread(s);
}

}  // extern "C"
