// Serving loop: resident graphs, admission of requests by the HBM bytes of
// their slot (distance row + queue partition), continuous batching over
// concurrent launch lanes.  Contract in include/mlmq_serve.h.
#include "../../include/mlmq_serve.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
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
    if (p.light_split && p.hub_split) return fail(MLMQ_E_INVALID, "light_split excludes hub_split");
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
    if (c.streams < 1 || c.streams > 32 || c.max_batch < 1 || c.max_batch > 4096 || c.max_pending < 0 ||
        c.spread_waves < 0 || c.spread_waves > 65536)
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
    (void)hipSetDevice(s->device);
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

    if (stats) {
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
        stats->launches = (uint64_t)launches;
        stats->wall_s = wall_ms / 1000.0;
        stats->offered_qps = nreq > 0 && last > 0 ? (double)nreq / last : 0.0;
        stats->completed_qps = wall_ms > 0 ? (double)stats->completed / stats->wall_s : 0.0;
        stats->p50_ms = pct(lat, 0.50);
        stats->p99_ms = pct(lat, 0.99);
        stats->max_ms = lat.empty() ? 0.0 : *std::max_element(lat.begin(), lat.end());
        double sum = 0;
        for (double x : lat) sum += x;
        stats->mean_ms = lat.empty() ? 0.0 : sum / (double)lat.size();
        stats->queue_p50_ms = pct(que, 0.50);
        stats->queue_p99_ms = pct(que, 0.99);
        stats->mean_batch = launches ? (double)batch_sum / (double)launches : 0.0;
        stats->mean_device_ms = stats->completed ? dev / (double)stats->completed : 0.0;
        stats->peak_bytes = std::max(peak_before, s->arena.peak);
        stats->peak_inflight = peak_inflight;
    }
    return MLMQ_OK;
}

}  // extern "C"
