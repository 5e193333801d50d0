// C-ABI host side of the MLMQ solver (see include/mlmq.h for the contract and
// the reference interface each entry point replaces).
#include "../../include/mlmq.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "device_common.h"
#include "hop_block.h"
#include "solver_internal.h"

struct mlmq_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_reset = nullptr, ev_start = nullptr, ev_stop = nullptr;
    int cus = 0;
    int wall_khz = 100000;
    // graph (device)
    int32_t n = 0;
    int64_t m = 0;
    uint32_t *row = nullptr, *col = nullptr, *wgt = nullptr;
    uint4 *blocks = nullptr;   // hop blocks, always owned by the context
    uint2 *hubs = nullptr;     // hub table (vertex, chunk size), owned
    uint32_t nhubs = 0;
    bool sorted = false;       // every adjacency sorted by weight (owned uploads)
    bool packed = false;       // hop blocks carry back-edge slots (mlmq::kBackShift)
    uint32_t light_delta = 0;  // delta the blocks' light-edge counts were computed for
    bool owns_graph = false;
    mlmq_graph_info info{};
    // per-query state
    uint32_t *dist = nullptr;
    int32_t dist_n = 0;
    uint32_t *ctl = nullptr;
    uint64_t *slots = nullptr;
    uint64_t slots_count = 0;
    bool slots_dirty = true;
    bool solved = false;
    std::vector<uint32_t> ctl_host;   // control block after the last mlmq_solve
    uint32_t last_buckets = 0, last_window = 0, last_cap_log2 = 0;
    // batch launch
    uint32_t *b_dist = nullptr;
    uint64_t b_dist_words = 0;
    int32_t b_nq = 0;
    uint32_t *b_ctl = nullptr;
    int b_ctl_blocks = 0;
    uint64_t *b_slots = nullptr;
    uint64_t b_slots_count = 0;
    bool b_slots_dirty = true;
    uint32_t *b_sources = nullptr;
    int32_t b_sources_n = 0;
    uint32_t *b_gctl = nullptr;
    uint32_t *b_qstatus = nullptr;
    uint64_t *b_qtime = nullptr;
    int32_t b_q_alloc = 0;
    bool b_solved = false;
    hipGraphExec_t b_graph = nullptr;
    std::vector<uint64_t> b_graph_key;
    // this device's copy of the static workspace (mlmq_solve_noalloc)
    uint8_t *scratch = nullptr;
    std::string err;
};

// Static workspace of mlmq_solve_noalloc: a control block, then
// MLMQ_NOALLOC_RING_SLOTS ring slots of 8 bytes (64 MiB + 64 KiB in all).
// Every device the library runs on holds its own copy until the process
// ends.  Host code never uses the array's name as an address: its device
// address on each device comes from hipGetSymbolAddress at context creation
// (scratch_address below) and is kept per device.
static constexpr uint64_t kScratchCtlBytes = 64ull << 10;
static constexpr uint64_t kScratchBytes = kScratchCtlBytes + 8ull * MLMQ_NOALLOC_RING_SLOTS;
static_assert(sizeof(uint32_t) * mlmq::kCtlWords <= kScratchCtlBytes, "control block fits");
static __device__ __align__(256) uint64_t scratch[kScratchBytes / 8];

namespace {

using clk = std::chrono::steady_clock;

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t *g_scratch_ptr[kMaxDevices];
bool g_scratch_looked_up[kMaxDevices];

// Device address of this device's workspace; `device` must be current.
// Looked up once per device.  Returns null if the lookup failed.
uint8_t *scratch_address(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_looked_up[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        g_scratch_ptr[device] = e == hipSuccess ? static_cast<uint8_t *>(ptr) : nullptr;
        g_scratch_looked_up[device] = true;
    }
    return g_scratch_ptr[device];
}

double ms_since(clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

int fail(mlmq_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

int hip_fail(mlmq_ctx *c, hipError_t e, const char *what) {
    return fail(c, e == hipErrorOutOfMemory ? MLMQ_E_NOMEM : MLMQ_E_HIP,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(ctx, expr)                                   \
    do {                                                     \
        hipError_t e_ = (expr);                              \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #expr); \
    } while (0)

void free_batch(mlmq_ctx *c) {
    if (c->b_graph) (void)hipGraphExecDestroy(c->b_graph);
    c->b_graph = nullptr;
    c->b_graph_key.clear();
    for (void *p : {(void *)c->b_dist, (void *)c->b_ctl, (void *)c->b_slots, (void *)c->b_sources,
                    (void *)c->b_gctl, (void *)c->b_qstatus, (void *)c->b_qtime})
        if (p) (void)hipFree(p);
    c->b_dist = nullptr;
    c->b_ctl = nullptr;
    c->b_slots = nullptr;
    c->b_sources = nullptr;
    c->b_gctl = nullptr;
    c->b_qstatus = nullptr;
    c->b_qtime = nullptr;
    c->b_dist_words = c->b_slots_count = 0;
    c->b_ctl_blocks = c->b_sources_n = c->b_q_alloc = c->b_nq = 0;
    c->b_solved = false;
    c->b_slots_dirty = true;
}

void free_graph(mlmq_ctx *c) {
    free_batch(c);
    if (c->owns_graph) {
        (void)hipFree(c->row);
        (void)hipFree(c->col);
        (void)hipFree(c->wgt);
    }
    if (c->blocks) (void)hipFree(c->blocks);
    c->blocks = nullptr;
    if (c->hubs) (void)hipFree(c->hubs);
    c->hubs = nullptr;
    c->nhubs = 0;
    c->sorted = false;
    c->light_delta = 0;
    c->row = c->col = c->wgt = nullptr;
    c->owns_graph = false;
    c->n = 0;
    c->m = 0;
    c->solved = false;
}

void compute_info(mlmq_graph_info *gi, int32_t n, int64_t m, const int64_t *row64,
                  const uint32_t *row32, const int32_t *w) {
    // Same statistics as graph_info::init_graph_info (core/include/graph_info.h:27-72).
    auto rowat = [&](int64_t i) -> double { return row64 ? (double)row64[i] : (double)row32[i]; };
    double sum = 0, mx = 0;
    for (int32_t i = 0; i < n; ++i) {
        double d = rowat(i + 1) - rowat(i);
        sum += d;
        if (d > mx) mx = d;
    }
    double avg = n ? sum / n : 0;
    double dev = 0;
    for (int32_t i = 0; i < n; ++i) {
        double d = rowat(i + 1) - rowat(i) - avg;
        dev += d * d;
    }
    dev = n ? std::sqrt(dev / n) : 0;
    double ws = 0, wm = 0;
    for (int64_t e = 0; e < m; ++e) {
        ws += w[e];
        if (w[e] > wm) wm = w[e];
    }
    double wavg = m ? ws / (double)m : 0;
    double wdev = 0;
    for (int64_t e = 0; e < m; ++e) {
        double d = w[e] - wavg;
        wdev += d * d;
    }
    wdev = m ? std::sqrt(wdev / (double)m) : 0;
    gi->m = n;
    gi->nnz = (int32_t)m;
    gi->avg_nnz = (float)avg;
    gi->max_nnz = (float)mx;
    gi->dev_nnz = (float)dev;
    gi->avg_weight = (float)wavg;
    gi->max_weight = (float)wm;
    gi->dev_weight = (float)wdev;
    gi->diameter = -1.0f;
}

// Shared validation of a host CSR (0-based starts, row[n] = m).
int validate_csr(mlmq_ctx *c, int32_t n, int64_t m, const int64_t *row, const int32_t *col,
                 const int32_t *w) {
    if (n <= 0 || n >= 0x7FFFFFFF) return fail(c, MLMQ_E_INVALID, "vertex count out of range");
    if (m < 0 || m >= 0x7FFFFFFF) return fail(c, MLMQ_E_INVALID, "edge count out of range");
    if (!row || (m > 0 && (!col || !w))) return fail(c, MLMQ_E_INVALID, "null CSR array");
    if (row[0] != 0 || row[n] != m) return fail(c, MLMQ_E_INVALID, "row[0] != 0 or row[n] != m");
    for (int32_t u = 0; u < n; ++u)
        if (row[u + 1] < row[u]) return fail(c, MLMQ_E_INVALID, "row offsets not monotone");
    for (int64_t e = 0; e < m; ++e) {
        if (col[e] < 0 || col[e] >= n) return fail(c, MLMQ_E_INVALID, "destination out of range");
        if (w[e] < 0) return fail(c, MLMQ_E_INVALID, "negative edge weight");
    }
    return MLMQ_OK;
}

// Device-side hop blocks (64 B per vertex) from the device CSR, and the hub
// table: vertices with more than kSplitDeg out-edges, highest degree first,
// at most kMaxHubs of them, each with a chunk size that keeps its chunk count
// below 2^kChunkBits.
int build_blocks(mlmq_ctx *c) {
    HIP_TRY(c, hipMalloc(&c->blocks, 64 * (size_t)c->n));
    // Back-edge slots fit when vertex ids and weights stay below 2^kBackShift
    // (the weight statistic is a float: rounding can only refuse a graph
    // that would fit).
    const double lim = (double)(1u << mlmq::kBackShift);
    c->packed = (double)c->n < lim && (double)c->info.max_weight < lim;
    HIP_TRY(c, mlmq::launch_build_blocks(c->row, c->col, c->wgt, (uint32_t)c->n, c->packed, c->blocks,
                                         c->stream));
    std::vector<uint32_t> row((size_t)c->n + 1);
    HIP_TRY(c, hipMemcpyAsync(row.data(), c->row, sizeof(uint32_t) * row.size(), hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    std::vector<std::pair<uint32_t, uint32_t>> hubs;   // (degree, vertex)
    for (int32_t v = 0; v < c->n; ++v) {
        const uint32_t deg = row[v + 1] - row[v];
        if (deg > mlmq::kSplitDeg) hubs.push_back({deg, (uint32_t)v});
    }
    std::sort(hubs.begin(), hubs.end(), [](const auto &a, const auto &b) {
        return a.first != b.first ? a.first > b.first : a.second < b.second;
    });
    if (hubs.size() > mlmq::kMaxHubs) hubs.resize(mlmq::kMaxHubs);
    if (!hubs.empty()) {
        std::vector<uint2> tab(hubs.size());
        const uint32_t max_chunks = (1u << mlmq::kChunkBits) - 1;
        for (size_t h = 0; h < hubs.size(); ++h) {
            uint32_t cs = mlmq::kChunkEdges;
            const uint32_t deg = hubs[h].first;
            if ((deg + cs - 1) / cs > max_chunks) cs = ((deg + max_chunks - 1) / max_chunks + 255) & ~255u;
            tab[h] = make_uint2(hubs[h].second, cs);
        }
        HIP_TRY(c, hipMalloc(&c->hubs, sizeof(uint2) * tab.size()));
        HIP_TRY(c, hipMemcpy(c->hubs, tab.data(), sizeof(uint2) * tab.size(), hipMemcpyHostToDevice));
        c->nhubs = (uint32_t)tab.size();
        HIP_TRY(c, mlmq::launch_patch_hubs(c->hubs, c->nhubs, c->blocks, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    return MLMQ_OK;
}

int upload(mlmq_ctx *c, int32_t n, int64_t m, const int64_t *row, const int32_t *col,
           const int32_t *w) {
    int rc = validate_csr(c, n, m, row, col, w);
    if (rc) return rc;
    HIP_TRY(c, hipSetDevice(c->device));
    free_graph(c);
    std::vector<uint32_t> row32((size_t)n + 1);
    for (int64_t i = 0; i <= n; ++i) row32[i] = (uint32_t)row[i];
    const size_t me = (size_t)(m > 0 ? m : 1);
    // The adjacency of every vertex with more than kInlineEdges edges (the
    // ones relaxed through the CSR) is sorted by (weight, destination): the
    // order in which a vertex's edges are relaxed does not change any
    // distance, lighter edges first lowers fewer distances twice, and the
    // edges lighter than any delta form a prefix (light/heavy split).
    // Inline vertices keep their order (their split is decided per edge).
    std::vector<int32_t> scol((size_t)m), sw((size_t)m);
    {
        const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> th;
        for (int t = 0; t < nt; ++t)
            th.emplace_back([&, t]() {
                std::vector<std::pair<int32_t, int32_t>> tmp;
                for (int64_t u = t; u < n; u += nt) {
                    const int64_t b = row[u], e = row[u + 1];
                    if (e - b <= (int64_t)mlmq::kInlineEdges) {
                        for (int64_t i = b; i < e; ++i) {
                            sw[(size_t)i] = w[i];
                            scol[(size_t)i] = col[i];
                        }
                        continue;
                    }
                    tmp.resize((size_t)(e - b));
                    for (int64_t i = b; i < e; ++i) tmp[(size_t)(i - b)] = {w[i], col[i]};
                    std::sort(tmp.begin(), tmp.end());
                    for (int64_t i = b; i < e; ++i) {
                        sw[(size_t)i] = tmp[(size_t)(i - b)].first;
                        scol[(size_t)i] = tmp[(size_t)(i - b)].second;
                    }
                }
            });
        for (auto &x : th) x.join();
    }
    col = scol.data();
    w = sw.data();
    HIP_TRY(c, hipMalloc(&c->row, sizeof(uint32_t) * ((size_t)n + 1)));
    c->owns_graph = true;
    HIP_TRY(c, hipMalloc(&c->col, sizeof(uint32_t) * me));
    HIP_TRY(c, hipMalloc(&c->wgt, sizeof(uint32_t) * me));
    HIP_TRY(c, hipMemcpy(c->row, row32.data(), sizeof(uint32_t) * ((size_t)n + 1),
                         hipMemcpyHostToDevice));
    if (m > 0) {
        HIP_TRY(c, hipMemcpy(c->col, col, sizeof(uint32_t) * m, hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(c->wgt, w, sizeof(uint32_t) * m, hipMemcpyHostToDevice));
    }
    c->n = n;
    c->m = m;
    compute_info(&c->info, n, m, row, nullptr, w);
    c->sorted = true;
    return build_blocks(c);
}

// Light-edge counts in the hop blocks for `delta` (light/heavy split).
int ensure_light_counts(mlmq_ctx *c, uint32_t delta) {
    if (!c->sorted) return fail(c, MLMQ_E_UNSUPPORTED, "light_split needs a graph loaded by mlmq_load_csr/gr");
    if (c->light_delta == delta) return MLMQ_OK;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, mlmq::launch_light_count(c->row, c->wgt, (uint32_t)c->n, delta, c->blocks, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->light_delta = delta;
    return MLMQ_OK;
}

uint32_t ceil_log2(uint64_t x) {
    uint32_t k = 0;
    while ((1ull << k) < x) ++k;
    return k;
}

int ensure_dist(mlmq_ctx *c) {
    if (c->dist && c->dist_n >= c->n) return MLMQ_OK;
    if (c->dist) (void)hipFree(c->dist);
    c->dist = nullptr;
    HIP_TRY(c, hipMalloc(&c->dist, sizeof(uint32_t) * (size_t)c->n));
    c->dist_n = c->n;
    return MLMQ_OK;
}

int ensure_queue(mlmq_ctx *c, uint32_t nbuckets, uint32_t cap_log2) {
    if (!c->ctl) {
        HIP_TRY(c, hipMalloc(&c->ctl, sizeof(uint32_t) * mlmq::kCtlWords));
        c->ctl_host.assign(mlmq::kCtlWords, 0);
    }
    const uint64_t need = (uint64_t)nbuckets << cap_log2;
    if (c->slots && c->slots_count < need) {
        (void)hipFree(c->slots);
        c->slots = nullptr;
        c->slots_count = 0;
    }
    if (!c->slots) {
        HIP_TRY(c, hipMalloc(&c->slots, sizeof(uint64_t) * need));
        c->slots_count = need;
        c->slots_dirty = true;
    }
    if (c->slots_dirty) {
        HIP_TRY(c, mlmq::launch_fill_u64(c->slots, c->slots_count, mlmq::kEmpty, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        c->slots_dirty = false;
    }
    return MLMQ_OK;
}

// Range checks shared by mlmq_solve and mlmq_solve_batch (and exported as
// mlmq_params_check).  min_read must be at least 1: a wave tops up from L2
// only while it holds fewer than min_read records, so with 0 no wave would
// ever take the seed and the solve could only end at the watchdog.
int check_params(const mlmq_params &p, bool batch, std::string *why) {
    auto bad = [&](int code, const char *msg) {
        if (why) *why = msg;
        return code;
    };
    const bool fifo = p.queue_mode == MLMQ_QUEUE_L1V_L2DQ;
    const bool slf = p.queue_mode == MLMQ_QUEUE_L1SLF_L2DQ || fifo;
    if (batch && p.queue_mode != MLMQ_QUEUE_L1SLF_L2DQ)
        return bad(MLMQ_E_UNSUPPORTED, "batched solves use L1SLF_L2DQ");
    if (!slf && p.queue_mode != MLMQ_QUEUE_L1N_L2DQ) return bad(MLMQ_E_UNSUPPORTED, "queue mode not implemented");
    if (p.delta == 0 || p.delta > 0x7FFFFFFFu) return bad(MLMQ_E_INVALID, "delta must be in [1, INT_MAX]");
    if (p.buckets < 1 || p.buckets > (int)mlmq::kMaxBuckets) return bad(MLMQ_E_INVALID, "buckets must be in [1, 64]");
    if (p.bucket_window < 1 || p.bucket_window > p.buckets) return bad(MLMQ_E_INVALID, "bucket_window must be in [1, buckets]");
    if (p.l2_batch < 1 || p.l2_batch > 64) return bad(MLMQ_E_INVALID, "l2_batch must be in [1, 64]");
    if (p.min_read < 1 || p.min_read > 64) return bad(MLMQ_E_INVALID, "min_read must be in [1, 64]");
    if (p.l1_pop < 1 || p.l1_pop > 64) return bad(MLMQ_E_INVALID, "l1_pop must be in [1, 64]");
    if (p.l1_ring != 0 && p.l1_ring != 512 && p.l1_ring != 1024)
        return bad(MLMQ_E_INVALID, "l1_ring must be 0, 512 or 1024");
    if (!batch && p.l1_ring == 1024) return bad(MLMQ_E_UNSUPPORTED, "l1_ring 1024 is for batched solves");
    {
        // one push of a whole batch (64 lanes x 7 inline edges) must fit above the threshold
        const int ring = p.l1_ring ? p.l1_ring : 512;
        const int max_spill = std::min(255, ring - 448);
        if (p.l1_spill < 1 || p.l1_spill > max_spill)
            return bad(MLMQ_E_INVALID, ring == 512 ? "l1_spill must be in [1, 64]"
                                                   : "l1_spill must be in [1, 255] with l1_ring 1024");
    }
    if (p.l1_donate < 0 || p.l1_donate > 64) return bad(MLMQ_E_INVALID, "l1_donate must be in [0, 64]");
    if (batch && p.batch_waves != 4 && p.batch_waves != 8 && p.batch_waves != 16)
        return bad(MLMQ_E_INVALID, "batch_waves must be 4, 8 or 16");
    if (p.stale_check < 0 || p.stale_check > 3) return bad(MLMQ_E_INVALID, "stale_check must be 0, 1, 2 or 3");
    if (p.l1_blocks != 0 && p.l1_blocks != 1) return bad(MLMQ_E_INVALID, "l1_blocks must be 0 or 1");
    if (p.idle_backoff < 0 || p.idle_backoff > 64) return bad(MLMQ_E_INVALID, "idle_backoff must be in [0, 64]");
    if (p.csr_stream != 0 && p.csr_stream != 1) return bad(MLMQ_E_INVALID, "csr_stream must be 0 or 1");
    if (p.batch_stats != 0 && p.batch_stats != 1) return bad(MLMQ_E_INVALID, "batch_stats must be 0 or 1");
    if (p.light_split && p.defer_heavy) return bad(MLMQ_E_INVALID, "light_split excludes defer_heavy");
    if (p.l1_blocks && (p.light_split || p.defer_heavy || p.strict_window))
        return bad(MLMQ_E_INVALID, "l1_blocks excludes light_split, defer_heavy and strict_window");
    if (p.work_blocks < 0 || p.blocks_per_cu < 0 || p.bucket_capacity < 0)
        return bad(MLMQ_E_INVALID, "negative size parameter");
    return MLMQ_OK;
}

}  // namespace

int mlmq::ctx_prepare_light(mlmq_ctx *c, uint32_t delta) {
    if (!c) return MLMQ_E_INVALID;
    if (!c->row) return MLMQ_E_NOGRAPH;
    return ensure_light_counts(c, delta);
}

int mlmq::ctx_graph_view(const mlmq_ctx *c, DeviceGraphView *v) {
    if (!c || !v) return MLMQ_E_INVALID;
    if (!c->row) return MLMQ_E_NOGRAPH;
    v->row = c->row;
    v->col = c->col;
    v->wgt = c->wgt;
    v->blocks = c->blocks;
    v->hubs = c->hubs;
    v->packed = c->packed;
    v->n = c->n;
    v->m = c->m;
    v->device = c->device;
    v->cus = c->cus;
    v->wall_khz = c->wall_khz;
    return MLMQ_OK;
}

namespace {

// Launch shape of one query, shared by mlmq_solve and mlmq_solve_noalloc.
struct QueryPlan {
    mlmq_params p;
    bool slf = true, fifo = false, hot = false;
    int blocks = 0;
    uint32_t cap_log2 = 0;   // default ring capacity (slots per bucket, log2)
};

// Argument and parameter checks, light-edge counts, grid size and the
// default ring capacity of one query on the context's graph.  Allocates
// nothing.
int plan_query(mlmq_ctx *c, int32_t src, const mlmq_params *params, QueryPlan *q) {
    if (!c->row) return fail(c, MLMQ_E_NOGRAPH, "no graph loaded");
    if (src < 0 || src >= c->n) return fail(c, MLMQ_E_INVALID, "source out of range");
    if (params) q->p = *params;
    else mlmq_params_default(&q->p);
    const mlmq_params &p = q->p;
    {
        std::string why;
        const int prc = check_params(p, false, &why);
        if (prc) return fail(c, prc, why);
    }
    q->fifo = p.queue_mode == MLMQ_QUEUE_L1V_L2DQ;
    q->slf = p.queue_mode == MLMQ_QUEUE_L1SLF_L2DQ || q->fifo;
    q->hot = q->slf && p.l1_blocks != 0;
    if (p.light_split) {
        const int lrc = ensure_light_counts(c, p.delta);
        if (lrc) return lrc;
    }
    HIP_TRY(c, hipSetDevice(c->device));

    int occ = 0;
    HIP_TRY(c, mlmq::solver_occupancy(q->slf, q->hot, mlmq::lane_pairs_for((uint32_t)p.l1_pop), &occ));
    if (occ < 1) return fail(c, MLMQ_E_HIP, "solver kernel cannot be resident");
    int blocks = p.work_blocks;
    if (blocks == 0 && p.blocks_per_cu > 0) {
        const int per_cu = p.blocks_per_cu < occ ? p.blocks_per_cu : occ;
        blocks = per_cu * c->cus;
    } else if (blocks == 0) {
        // Latency-bound solver: a road-like graph of n vertices keeps only
        // O(sqrt(n)) records in flight, so more waves mostly add queue
        // polling.  Measured optimum on MI355X: ~256 waves at 0.26 M
        // vertices, 512 at 6.3 M.
        int waves = (int)(c->n / 25000);
        if (waves < 128) waves = 128;
        if (waves > 512) waves = 512;
        blocks = waves / mlmq::kWavesPerBlock;
    }
    if (blocks > occ * c->cus) blocks = occ * c->cus;
    q->blocks = blocks;

    uint64_t cap = (uint64_t)p.bucket_capacity;
    if (cap == 0) cap = (uint64_t)c->n > 8192 ? (uint64_t)c->n : 8192;
    q->cap_log2 = ceil_log2(cap);
    if (q->cap_log2 > 31) return fail(c, MLMQ_E_INVALID, "bucket capacity too large");
    return MLMQ_OK;
}

mlmq::SolverLaunch query_launch(const mlmq_ctx *c, const QueryPlan &q, uint32_t *dist, uint32_t *ctl,
                                uint64_t *slots, uint32_t cap_log2) {
    const mlmq_params &p = q.p;
    const double timeout_s = p.timeout_s > 0 ? p.timeout_s : 60.0;
    mlmq::SolverLaunch L{};
    L.row = c->row;
    L.col = c->col;
    L.wgt = c->wgt;
    L.blocks = c->blocks;
    L.hubs = c->hubs;
    L.packed = c->packed;
    L.split = p.hub_split != 0;
    L.stale_check = (uint32_t)p.stale_check;
    L.prefetch = p.prefetch ? 1u : 0u;
    L.defer_heavy = p.defer_heavy ? 1u : 0u;
    L.strict_window = p.strict_window ? 1u : 0u;
    L.light_split = p.light_split ? 1u : 0u;
    L.csr_stream = p.csr_stream ? 1u : 0u;
    L.hot = q.hot;
    L.dist = dist;
    L.slots = slots;
    L.ctl = ctl;
    L.n = (uint32_t)c->n;
    L.nbuckets = (uint32_t)p.buckets;
    L.window = (uint32_t)p.bucket_window;
    L.cap_log2 = cap_log2;
    L.delta = p.delta;
    L.batch = (uint32_t)p.l2_batch;
    L.spin_limit = mlmq::kSpinLimit;
    L.min_read = (uint32_t)p.min_read;
    L.max_sleep = p.idle_backoff > 0 ? (uint32_t)p.idle_backoff : 16u;
    L.pop_max = (uint32_t)p.l1_pop;
    L.spill = (uint32_t)p.l1_spill;
    L.donate_min = (uint32_t)p.l1_donate;
    L.timeout_ticks = (uint64_t)(timeout_s * (double)c->wall_khz * 1000.0);
    L.grid_blocks = q.blocks;
    L.slf = q.slf;
    L.fifo = q.fifo;
    L.profile = p.profile != 0;
    return L;
}

struct QueryClock {
    clk::time_point q0, q1, s0, s1;
};

// Reset and persistent launch of one query, on the context's stream; the
// control block after the solve is copied into h (kCtlWords words).
// fill_slots > 0 first fills that many ring slots with empty records.
int run_query(mlmq_ctx *c, int32_t src, const mlmq::SolverLaunch &L, uint64_t fill_slots, uint32_t *h,
              QueryClock *t) {
    t->q0 = clk::now();
    HIP_TRY(c, hipEventRecord(c->ev_reset, c->stream));
    if (fill_slots) HIP_TRY(c, mlmq::launch_fill_u64(L.slots, fill_slots, mlmq::kEmpty, c->stream));
    HIP_TRY(c, mlmq::launch_reset(L.dist, L.n, (uint32_t)src, L.ctl, L.slots, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    t->q1 = clk::now();

    t->s0 = clk::now();
    HIP_TRY(c, hipEventRecord(c->ev_start, c->stream));
    HIP_TRY(c, mlmq::launch_solver(L, c->stream));
    HIP_TRY(c, hipEventRecord(c->ev_stop, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    t->s1 = clk::now();

    HIP_TRY(c, hipMemcpy(h, L.ctl, sizeof(uint32_t) * mlmq::kCtlWords, hipMemcpyDeviceToHost));
    return MLMQ_OK;
}

// Outcome of a finished query from its control block: the device error, or
// MLMQ_E_INTERNAL unless every reserved record was taken and credited.
int query_status(mlmq_ctx *c, const uint32_t *h, uint32_t nbuckets) {
    const uint32_t derr = h[mlmq::kErrOff];
    if (derr == mlmq::kErrOverflow)
        return fail(c, MLMQ_E_QUEUE_OVERFLOW, "L2 bucket ring overflow; raise bucket_capacity");
    if (derr == mlmq::kErrTimeout) return fail(c, MLMQ_E_TIMEOUT, "device watchdog expired");
    if (derr != mlmq::kErrNone) return fail(c, MLMQ_E_INTERNAL, "unknown device error");
    uint32_t reserved = 0, taken = 0;
    for (uint32_t b = 0; b < nbuckets; ++b) {
        reserved += h[mlmq::kTailOff + b * mlmq::kStride];
        taken += h[mlmq::kTakenOff + b * mlmq::kStride];
    }
    const uint32_t done = h[mlmq::kDoneOff];
    if (h[mlmq::kExitOff] != 1 || reserved != taken || reserved != done) {
        char buf[160];
        std::snprintf(buf, sizeof buf, "queue accounting mismatch: reserved=%u taken=%u done=%u exit=%u",
                      reserved, taken, done, h[mlmq::kExitOff]);
        return fail(c, MLMQ_E_INTERNAL, buf);
    }
    return MLMQ_OK;
}

void fill_timing(mlmq_ctx *c, const uint32_t *h, const QueryClock &t, int blocks, uint32_t cap_log2,
                 uint64_t retries, mlmq_timing *timing) {
    std::memset(timing, 0, sizeof(*timing));
    float dev_ms = 0;
    (void)hipEventElapsedTime(&dev_ms, c->ev_start, c->ev_stop);
    timing->solve_ms = ms_since(t.s0, t.s1);
    timing->device_ms = dev_ms;
    timing->reset_ms = ms_since(t.q0, t.q1);
    timing->query_wall_ms = ms_since(t.q0, t.s1);
    timing->work_blocks = blocks;
    timing->waves = blocks * mlmq::kWavesPerBlock;
    timing->bucket_capacity = (int64_t)1 << cap_log2;
    const uint64_t *st = reinterpret_cast<const uint64_t *>(h + mlmq::kStatsOff);
    timing->l2_pushes = st[mlmq::kStatL2Push] + 1;  // + the seed
    timing->l2_pops = st[mlmq::kStatL2Pop];
    timing->relaxations = st[mlmq::kStatRelax];
    timing->atomics = st[mlmq::kStatAtomic];
    timing->improvements = st[mlmq::kStatImprove];
    timing->stale = st[mlmq::kStatStale];
    timing->batches = st[mlmq::kStatBatches];
    timing->idle_polls = st[mlmq::kStatIdle];
    timing->retries = retries;
    for (int i = 0; i < 8; ++i) timing->phase_cycles[i] = st[mlmq::kClkL1Pop + i];
    for (int i = 0; i < 4; ++i) timing->detail_cycles[i] = st[mlmq::kClkOwner + i];
    timing->detail_cycles[4] = 0;
    timing->donated = st[mlmq::kStatDonated];
    timing->hub_chunks = st[mlmq::kStatChunks];
}

}  // namespace

extern "C" {

const char *mlmq_version(void) { return "mlmq-gfx950 0.2"; }

int mlmq_params_check(const mlmq_params *p, int32_t batch) {
    if (!p) return MLMQ_E_INVALID;
    return check_params(*p, batch != 0, nullptr);
}

void mlmq_params_default(mlmq_params *p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->queue_mode = MLMQ_QUEUE_L1SLF_L2DQ;
    p->delta = 200000;
    p->work_blocks = 0;
    p->blocks_per_cu = 0;
    p->buckets = 16;
    p->bucket_window = 4;
    p->l2_batch = 8;
    p->min_read = 8;
    p->bucket_capacity = 0;
    p->timeout_s = 60.0;
    p->auto_retry = 1;
    p->profile = 0;
    p->l1_pop = 32;
    p->l1_spill = 64;
    p->l1_donate = 0;
    p->batch_waves = 8;
    p->use_graph = 1;
    p->hub_split = 0;
    p->stale_check = 2;
    p->prefetch = 1;
    p->l1_blocks = 0;
    p->idle_backoff = 16;
    p->defer_heavy = 0;
    p->strict_window = 0;
    p->light_split = 0;
    p->csr_stream = 0;
    p->l1_ring = 512;
    p->batch_stats = 1;
}

int mlmq_create(int device, mlmq_ctx **out) {
    if (!out) return MLMQ_E_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
        return MLMQ_E_INVALID;
    mlmq_ctx *c = new (std::nothrow) mlmq_ctx();
    if (!c) return MLMQ_E_NOMEM;
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev_reset);
    if (e == hipSuccess) e = hipEventCreate(&c->ev_start);
    if (e == hipSuccess) e = hipEventCreate(&c->ev_stop);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess) {
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
            c->wall_khz = khz;
    }
    if (e != hipSuccess) {
        mlmq_destroy(c);
        return MLMQ_E_HIP;
    }
    // (device current: set above)
    c->scratch = scratch_address(device);
    *out = c;
    return MLMQ_OK;
}

int mlmq_destroy(mlmq_ctx *c) {
    if (!c) return MLMQ_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_graph(c);
    if (c->dist) (void)hipFree(c->dist);
    if (c->ctl) (void)hipFree(c->ctl);
    if (c->slots) (void)hipFree(c->slots);
    free_batch(c);
    if (c->ev_reset) (void)hipEventDestroy(c->ev_reset);
    if (c->ev_start) (void)hipEventDestroy(c->ev_start);
    if (c->ev_stop) (void)hipEventDestroy(c->ev_stop);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return MLMQ_OK;
}

const char *mlmq_last_error(const mlmq_ctx *c) { return c ? c->err.c_str() : "null context"; }

int mlmq_load_csr(mlmq_ctx *c, int32_t n, int64_t m, const int64_t *row, const int32_t *col,
                  const int32_t *w) {
    if (!c) return MLMQ_E_INVALID;
    return upload(c, n, m, row, col, w);
}

// Galois .gr v1 (SSSP/csr_graph.cu:185-296): header u64 {1, 4, n, m}, n u64
// row-end offsets, m u32 destinations, 4 pad bytes when m is odd, m int32
// weights.  Validation as the reference's strict reader
// (scripts/multigpu/supplement_graph.cpp:16-30).
int mlmq_load_gr(mlmq_ctx *c, const char *path) {
    if (!c || !path) return MLMQ_E_INVALID;
    int fd = open(path, O_RDONLY);
    if (fd < 0) return fail(c, MLMQ_E_IO, std::string("cannot open ") + path);
    struct stat sb;
    if (fstat(fd, &sb) != 0 || sb.st_size < 32) {
        close(fd);
        return fail(c, MLMQ_E_IO, "cannot stat or truncated header");
    }
    const size_t bytes = (size_t)sb.st_size;
    void *map = mmap(nullptr, bytes, PROT_READ, MAP_PRIVATE, fd, 0);
    close(fd);
    if (map == MAP_FAILED) return fail(c, MLMQ_E_IO, "mmap failed");
    const unsigned char *p = (const unsigned char *)map;
    uint64_t h[4];
    std::memcpy(h, p, sizeof(h));
    int rc = MLMQ_OK;
    if (h[0] != 1 || h[1] != 4 || h[2] == 0 || h[2] >= 0x7FFFFFFFull || h[3] >= 0x7FFFFFFFull) {
        rc = fail(c, MLMQ_E_INVALID, "unsupported .gr header");
    } else {
        const uint64_t n = h[2], m = h[3];
        const uint64_t need = 32 + 8 * n + 4 * m + 4 * (m & 1) + 4 * m;
        if (bytes < need) {
            rc = fail(c, MLMQ_E_INVALID, "truncated .gr file");
        } else {
            std::vector<int64_t> row(n + 1);
            row[0] = 0;
            const unsigned char *ends = p + 32;
            for (uint64_t u = 0; u < n; ++u) {
                uint64_t e;
                std::memcpy(&e, ends + 8 * u, 8);
                if (e > m) { rc = fail(c, MLMQ_E_INVALID, "row end beyond m"); break; }
                row[u + 1] = (int64_t)e;
            }
            if (rc == MLMQ_OK) {
                const int32_t *dst = (const int32_t *)(ends + 8 * n);
                const int32_t *w = dst + m + (m & 1);
                rc = upload(c, (int32_t)n, (int64_t)m, row.data(), dst, w);
            }
        }
    }
    munmap(map, bytes);
    return rc;
}

int mlmq_attach_device_csr(mlmq_ctx *c, int32_t n, int64_t m, const int32_t *row_d,
                           const int32_t *col_d, const int32_t *w_d) {
    if (!c) return MLMQ_E_INVALID;
    if (n <= 0 || m < 0 || m >= 0x7FFFFFFF || !row_d || (m > 0 && (!col_d || !w_d)))
        return fail(c, MLMQ_E_INVALID, "bad device CSR arguments");
    HIP_TRY(c, hipSetDevice(c->device));
    // Statistics (and a sanity check of the offsets) need the arrays on the host.
    std::vector<uint32_t> row((size_t)n + 1);
    std::vector<int32_t> w((size_t)m);
    HIP_TRY(c, hipMemcpy(row.data(), row_d, sizeof(int32_t) * ((size_t)n + 1), hipMemcpyDeviceToHost));
    if (m > 0) HIP_TRY(c, hipMemcpy(w.data(), w_d, sizeof(int32_t) * (size_t)m, hipMemcpyDeviceToHost));
    if (row[0] != 0 || (int64_t)row[n] != m) return fail(c, MLMQ_E_INVALID, "row[0] != 0 or row[n] != m");
    for (int32_t u = 0; u < n; ++u)
        if (row[u + 1] < row[u]) return fail(c, MLMQ_E_INVALID, "row offsets not monotone");
    for (int64_t e = 0; e < m; ++e)
        if (w[e] < 0) return fail(c, MLMQ_E_INVALID, "negative edge weight");
    free_graph(c);
    // Non-negative int32 arrays have the same bits as the solver's uint32 view.
    c->row = (uint32_t *)row_d;
    c->col = (uint32_t *)col_d;
    c->wgt = (uint32_t *)w_d;
    c->owns_graph = false;
    c->n = n;
    c->m = m;
    compute_info(&c->info, n, m, nullptr, row.data(), w.data());
    return build_blocks(c);
}

int mlmq_graph_stats(mlmq_ctx *c, mlmq_graph_info *info) {
    if (!c || !info) return MLMQ_E_INVALID;
    if (!c->row) return fail(c, MLMQ_E_NOGRAPH, "no graph loaded");
    *info = c->info;
    return MLMQ_OK;
}

int mlmq_graph_size(const mlmq_ctx *c, int32_t *n, int64_t *m) {
    if (!c) return MLMQ_E_INVALID;
    if (n) *n = c->n;
    if (m) *m = c->m;
    return c->row ? MLMQ_OK : MLMQ_E_NOGRAPH;
}

int mlmq_device_shape(mlmq_ctx *c, int32_t *blocks_per_cu_max, int32_t *cus, int32_t *threads) {
    if (!c) return MLMQ_E_INVALID;
    HIP_TRY(c, hipSetDevice(c->device));
    int b = 0;
    HIP_TRY(c, mlmq::solver_occupancy(true, false, false, &b));
    if (blocks_per_cu_max) *blocks_per_cu_max = b;
    if (cus) *cus = c->cus;
    if (threads) *threads = mlmq::kWavesPerBlock * mlmq::kWave;
    return MLMQ_OK;
}

int mlmq_solve(mlmq_ctx *c, int32_t src, const mlmq_params *params, mlmq_timing *timing) {
    if (!c) return MLMQ_E_INVALID;
    QueryPlan q;
    int rc = plan_query(c, src, params, &q);
    if (rc) return rc;
    rc = ensure_dist(c);
    if (rc) return rc;
    // The reset below overwrites the last solve's distances: until this solve
    // succeeds there is no result to copy out.
    c->solved = false;

    const uint32_t nb = (uint32_t)q.p.buckets;
    uint32_t cap_log2 = q.cap_log2;
    uint64_t retries = 0;
    for (;;) {
        rc = ensure_queue(c, nb, cap_log2);
        if (rc) return rc;
        const mlmq::SolverLaunch L = query_launch(c, q, c->dist, c->ctl, c->slots, cap_log2);
        QueryClock t;
        rc = run_query(c, src, L, 0, c->ctl_host.data(), &t);
        if (rc) return rc;
        rc = query_status(c, c->ctl_host.data(), nb);
        if (rc) c->slots_dirty = true;
        if (rc == MLMQ_E_QUEUE_OVERFLOW && q.p.auto_retry && retries < 4 && cap_log2 < 31) {
            ++retries;
            ++cap_log2;
            continue;
        }
        if (rc) return rc;
        c->solved = true;
        c->last_buckets = nb;
        c->last_window = (uint32_t)q.p.bucket_window;
        c->last_cap_log2 = cap_log2;
        if (timing) fill_timing(c, c->ctl_host.data(), t, q.blocks, cap_log2, retries, timing);
        return MLMQ_OK;
    }
}

// The same query as mlmq_solve, with the control block and the bucket rings
// in this file's static device workspace and the distances in the caller's
// buffer: nothing is allocated.  The workspace is filled with empty slots and
// the control block cleared on every call, so no state is carried between
// calls; calls that overlap on one device share the workspace and race.
int mlmq_solve_noalloc(mlmq_ctx *c, int32_t src, const mlmq_params *params, int32_t *dist_d,
                       mlmq_timing *timing) {
    if (!c) return MLMQ_E_INVALID;
    if (!dist_d) return fail(c, MLMQ_E_INVALID, "null distance buffer");
    QueryPlan q;
    int rc = plan_query(c, src, params, &q);
    if (rc) return rc;
    if (!c->scratch) return fail(c, MLMQ_E_HIP, "static workspace address unavailable on this device");
    const uint64_t limit = MLMQ_NOALLOC_RING_SLOTS;
    const uint32_t nb = (uint32_t)q.p.buckets;
    uint32_t cap_log2 = q.cap_log2;
    if (((uint64_t)nb << cap_log2) > limit) {
        if (q.p.bucket_capacity > 0)
            return fail(c, MLMQ_E_INVALID,
                        "buckets x bucket_capacity exceeds the static workspace (MLMQ_NOALLOC_RING_SLOTS)");
        while (((uint64_t)nb << cap_log2) > limit) --cap_log2;
    }
    uint32_t *ctl = reinterpret_cast<uint32_t *>(c->scratch);
    uint64_t *slots = reinterpret_cast<uint64_t *>(c->scratch + kScratchCtlBytes);
    uint32_t h[mlmq::kCtlWords];
    uint64_t retries = 0;
    for (;;) {
        const mlmq::SolverLaunch L = query_launch(c, q, reinterpret_cast<uint32_t *>(dist_d), ctl, slots,
                                                  cap_log2);
        QueryClock t;
        rc = run_query(c, src, L, (uint64_t)nb << cap_log2, h, &t);
        if (rc) return rc;
        rc = query_status(c, h, nb);
        if (rc == MLMQ_E_QUEUE_OVERFLOW && q.p.auto_retry && retries < 4 &&
            ((uint64_t)nb << (cap_log2 + 1)) <= limit) {
            ++retries;
            ++cap_log2;
            continue;
        }
        if (rc) return rc;
        if (timing) fill_timing(c, h, t, q.blocks, cap_log2, retries, timing);
        return MLMQ_OK;
    }
}

int mlmq_solve_batch(mlmq_ctx *c, const int32_t *sources, int32_t nq, const mlmq_params *params,
                     mlmq_batch_timing *timing) {
    if (!c) return MLMQ_E_INVALID;
    if (!c->row) return fail(c, MLMQ_E_NOGRAPH, "no graph loaded");
    if (!sources || nq <= 0) return fail(c, MLMQ_E_INVALID, "empty source list");
    mlmq_params p;
    if (params) p = *params;
    else mlmq_params_default(&p);
    {
        std::string why;
        const int prc = check_params(p, true, &why);
        if (prc) return fail(c, prc, why);
    }
    if (p.light_split) {
        const int lrc = ensure_light_counts(c, p.delta);
        if (lrc) return lrc;
    }
    for (int32_t i = 0; i < nq; ++i)
        if (sources[i] < 0 || sources[i] >= c->n) return fail(c, MLMQ_E_INVALID, "source out of range");
    HIP_TRY(c, hipSetDevice(c->device));

    int occ = 0;
    const int l1_cap = p.l1_ring ? p.l1_ring : 512;
    const bool ext = p.light_split || p.defer_heavy || p.strict_window;
    HIP_TRY(c, mlmq::batch_occupancy(p.batch_waves, l1_cap, ext, p.batch_stats != 0,
                                     mlmq::lane_pairs_for((uint32_t)p.l1_pop), &occ));
    if (occ < 1) return fail(c, MLMQ_E_HIP, "batch kernel cannot be resident");
    int blocks = p.work_blocks ? p.work_blocks : occ * c->cus;
    if (blocks > occ * c->cus) blocks = occ * c->cus;
    if (blocks > nq) blocks = nq;

    uint64_t cap = (uint64_t)p.bucket_capacity;
    if (cap == 0) cap = (uint64_t)c->n / 8 > 8192 ? (uint64_t)c->n / 8 : 8192;
    uint32_t cap_log2 = ceil_log2(cap);
    if (cap_log2 > 31) return fail(c, MLMQ_E_INVALID, "bucket capacity too large");

    // The launch overwrites the last batch's rows: until it succeeds there is
    // no batch result to copy out.
    c->b_solved = false;
    c->b_nq = 0;
    // per-query rows and bookkeeping
    const uint64_t words = (uint64_t)nq * (uint64_t)c->n;
    if (c->b_dist_words < words) {
        if (c->b_dist) (void)hipFree(c->b_dist);
        c->b_dist = nullptr;
        c->b_dist_words = 0;
        HIP_TRY(c, hipMalloc(&c->b_dist, sizeof(uint32_t) * words));
        c->b_dist_words = words;
    }
    if (c->b_q_alloc < nq) {
        for (void *q : {(void *)c->b_sources, (void *)c->b_qstatus, (void *)c->b_qtime})
            if (q) (void)hipFree(q);
        c->b_sources = c->b_qstatus = nullptr;
        c->b_qtime = nullptr;
        c->b_q_alloc = 0;
        HIP_TRY(c, hipMalloc(&c->b_sources, sizeof(uint32_t) * nq));
        HIP_TRY(c, hipMalloc(&c->b_qstatus, sizeof(uint32_t) * nq));
        HIP_TRY(c, hipMalloc(&c->b_qtime, sizeof(uint64_t) * 2 * nq));
        c->b_q_alloc = nq;
    }
    if (!c->b_gctl) HIP_TRY(c, hipMalloc(&c->b_gctl, sizeof(uint32_t) * mlmq::kCtlWords));
    HIP_TRY(c, hipMemcpy(c->b_sources, sources, sizeof(int32_t) * nq, hipMemcpyHostToDevice));

    const double timeout_s = p.timeout_s > 0 ? p.timeout_s : 60.0;
    uint64_t retries = 0;
    for (;;) {
        // queue partitions: one control block and bucket set per workgroup
        if (c->b_ctl_blocks < blocks) {
            if (c->b_ctl) (void)hipFree(c->b_ctl);
            c->b_ctl = nullptr;
            c->b_ctl_blocks = 0;
            HIP_TRY(c, hipMalloc(&c->b_ctl, sizeof(uint32_t) * mlmq::kCtlWords * (size_t)blocks));
            c->b_ctl_blocks = blocks;
        }
        const uint64_t need = (uint64_t)blocks * ((uint64_t)p.buckets << cap_log2);
        if (c->b_slots_count < need) {
            if (c->b_slots) (void)hipFree(c->b_slots);
            c->b_slots = nullptr;
            c->b_slots_count = 0;
            HIP_TRY(c, hipMalloc(&c->b_slots, sizeof(uint64_t) * need));
            c->b_slots_count = need;
            c->b_slots_dirty = true;
        }
        if (c->b_slots_dirty) {
            HIP_TRY(c, mlmq::launch_fill_u64(c->b_slots, c->b_slots_count, mlmq::kEmpty, c->stream));
            HIP_TRY(c, hipStreamSynchronize(c->stream));
            c->b_slots_dirty = false;
        }

        mlmq::SolverLaunch L{};
        L.row = c->row;
        L.col = c->col;
        L.wgt = c->wgt;
        L.blocks = c->blocks;
        L.hubs = c->hubs;
        L.packed = c->packed;
        L.split = p.hub_split != 0;
        L.stale_check = (uint32_t)p.stale_check;
        L.prefetch = p.prefetch ? 1u : 0u;
        L.defer_heavy = p.defer_heavy ? 1u : 0u;
        L.strict_window = p.strict_window ? 1u : 0u;
        L.light_split = p.light_split ? 1u : 0u;
        L.csr_stream = p.csr_stream ? 1u : 0u;
        L.hot = false;
        L.dist = nullptr;
        L.slots = nullptr;
        L.ctl = nullptr;
        L.n = (uint32_t)c->n;
        L.nbuckets = (uint32_t)p.buckets;
        L.window = (uint32_t)p.bucket_window;
        L.cap_log2 = cap_log2;
        L.delta = p.delta;
        L.batch = (uint32_t)p.l2_batch;
        L.spin_limit = mlmq::kSpinLimit;
        L.min_read = (uint32_t)p.min_read;
        L.max_sleep = p.idle_backoff > 0 ? (uint32_t)p.idle_backoff : 16u;
        L.pop_max = (uint32_t)p.l1_pop;
        L.spill = (uint32_t)p.l1_spill;
        L.donate_min = (uint32_t)p.l1_donate;
        L.timeout_ticks = (uint64_t)(timeout_s * (double)c->wall_khz * 1000.0);
        L.grid_blocks = blocks;
        L.slf = true;
        L.fifo = false;
        L.profile = false;
        mlmq::BatchLaunch B;
        B.ctl_all = c->b_ctl;
        B.slots_all = c->b_slots;
        B.dist_all = c->b_dist;
        B.sources = c->b_sources;
        B.nq = (uint32_t)nq;
        B.gctl = c->b_gctl;
        B.qstatus = c->b_qstatus;
        B.qtime = c->b_qtime;
        B.waves = p.batch_waves;
        B.l1_cap = l1_cap;
        B.stats = p.batch_stats != 0;

        // The launch (row reset + batch kernel) is captured once per shape and
        // replayed as a hipGraph; the source list is copied in before it.
        const std::vector<uint64_t> key = {
            (uint64_t)nq, (uint64_t)blocks, (uint64_t)p.batch_waves, cap_log2, p.delta,
            (uint64_t)p.buckets, (uint64_t)p.bucket_window, (uint64_t)p.l2_batch,
            (uint64_t)p.min_read, (uint64_t)p.l1_pop, (uint64_t)p.l1_spill, (uint64_t)p.l1_donate,
            L.timeout_ticks, (uint64_t)(uintptr_t)c->b_dist, (uint64_t)(uintptr_t)c->b_slots,
            (uint64_t)(uintptr_t)c->b_ctl, (uint64_t)(uintptr_t)c->b_sources,
            (uint64_t)(uintptr_t)c->row, (uint64_t)(uintptr_t)c->blocks,
            (uint64_t)(uintptr_t)c->hubs, (uint64_t)L.split, (uint64_t)L.stale_check,
            (uint64_t)L.prefetch, (uint64_t)L.max_sleep, (uint64_t)L.defer_heavy,
            (uint64_t)L.strict_window, (uint64_t)L.light_split, (uint64_t)L.csr_stream,
            (uint64_t)l1_cap, (uint64_t)B.stats, (uint64_t)c->n, (uint64_t)c->packed};
        const bool use_graph = p.use_graph != 0;
        if (use_graph && (!c->b_graph || c->b_graph_key != key)) {
            if (c->b_graph) (void)hipGraphExecDestroy(c->b_graph);
            c->b_graph = nullptr;
            hipGraph_t g = nullptr;
            HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e1 = mlmq::launch_batch_reset(c->b_dist, (uint32_t)c->n, (uint32_t)nq,
                                                     c->b_sources, c->b_gctl, c->stream);
            hipError_t e2 = mlmq::launch_batch(L, B, c->stream);
            hipError_t e3 = hipStreamEndCapture(c->stream, &g);
            if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return hip_fail(c, e1 != hipSuccess ? e1 : e2 != hipSuccess ? e2 : e3, "batch graph capture");
            }
            hipError_t e4 = hipGraphInstantiate(&c->b_graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e4 != hipSuccess) {
                c->b_graph = nullptr;
                return hip_fail(c, e4, "hipGraphInstantiate");
            }
            c->b_graph_key = key;
        }
        const auto s0 = clk::now();
        HIP_TRY(c, hipEventRecord(c->ev_start, c->stream));
        if (use_graph) {
            HIP_TRY(c, hipGraphLaunch(c->b_graph, c->stream));
        } else {
            HIP_TRY(c, mlmq::launch_batch_reset(c->b_dist, (uint32_t)c->n, (uint32_t)nq, c->b_sources,
                                                c->b_gctl, c->stream));
            HIP_TRY(c, mlmq::launch_batch(L, B, c->stream));
        }
        HIP_TRY(c, hipEventRecord(c->ev_stop, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        const auto s1 = clk::now();

        std::vector<uint32_t> g(mlmq::kCtlWords), qs((size_t)nq);
        std::vector<uint64_t> qt(2 * (size_t)nq);
        HIP_TRY(c, hipMemcpy(g.data(), c->b_gctl, sizeof(uint32_t) * mlmq::kCtlWords, hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemcpy(qs.data(), c->b_qstatus, sizeof(uint32_t) * nq, hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemcpy(qt.data(), c->b_qtime, sizeof(uint64_t) * 2 * nq, hipMemcpyDeviceToHost));
        const uint32_t derr = g[mlmq::kErrOff];
        if (derr != mlmq::kErrNone) {
            c->b_slots_dirty = true;
            if (derr == mlmq::kErrOverflow && p.auto_retry && retries < 4 && cap_log2 < 31) {
                ++retries;
                ++cap_log2;
                continue;
            }
            if (derr == mlmq::kErrOverflow) return fail(c, MLMQ_E_QUEUE_OVERFLOW, "batch queue partition overflow");
            if (derr == mlmq::kErrTimeout) return fail(c, MLMQ_E_TIMEOUT, "device watchdog expired in batch");
            return fail(c, MLMQ_E_INTERNAL, "batch queue accounting mismatch");
        }
        if (g[0] < (uint32_t)nq) return fail(c, MLMQ_E_INTERNAL, "batch ended before all queries were taken");
        c->b_nq = nq;
        c->b_solved = true;
        if (timing) {
            std::memset(timing, 0, sizeof(*timing));
            float dev_ms = 0;
            (void)hipEventElapsedTime(&dev_ms, c->ev_start, c->ev_stop);
            timing->solve_ms = ms_since(s0, s1);
            timing->device_ms = dev_ms;
            std::vector<double> lat((size_t)nq);
            double sum = 0;
            for (int32_t i = 0; i < nq; ++i) {
                lat[i] = (double)(qt[2 * i + 1] - qt[2 * i]) / (double)c->wall_khz;
                sum += lat[i];
            }
            std::sort(lat.begin(), lat.end());
            timing->query_ms_p50 = lat[(size_t)(0.50 * (nq - 1))];
            timing->query_ms_p99 = lat[(size_t)(0.99 * (nq - 1))];
            timing->query_ms_max = lat.back();
            timing->query_ms_mean = sum / nq;
            timing->queries = nq;
            timing->work_blocks = blocks;
            timing->waves_per_query = p.batch_waves;
            timing->graph_replay = use_graph ? 1 : 0;
            timing->bucket_capacity = (int64_t)1 << cap_log2;
            const uint64_t *st = reinterpret_cast<const uint64_t *>(g.data() + mlmq::kStatsOff);
            timing->relaxations = st[mlmq::kStatRelax];
            timing->improvements = st[mlmq::kStatImprove];
            timing->l2_pushes = st[mlmq::kStatL2Push] + (uint64_t)nq;
            timing->stale = st[mlmq::kStatStale];
            timing->batches = st[mlmq::kStatBatches];
            timing->idle_polls = st[mlmq::kStatIdle];
            timing->retries = retries;
        }
        return MLMQ_OK;
    }
}

int mlmq_copy_batch_dist(mlmq_ctx *c, int32_t query, int32_t *host_out) {
    if (!c || !host_out) return MLMQ_E_INVALID;
    if (!c->b_solved) return fail(c, MLMQ_E_NOGRAPH, "no completed batch");
    if (query < 0 || query >= c->b_nq) return fail(c, MLMQ_E_INVALID, "query index out of range");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpy(host_out, c->b_dist + (uint64_t)query * (uint64_t)c->n,
                         sizeof(int32_t) * (size_t)c->n, hipMemcpyDeviceToHost));
    return MLMQ_OK;
}

int mlmq_last_queue_counts(const mlmq_ctx *c, mlmq_queue_counts *out) {
    if (!c || !out) return MLMQ_E_INVALID;
    if (!c->solved) return MLMQ_E_NOGRAPH;
    std::memset(out, 0, sizeof(*out));
    const uint32_t *h = c->ctl_host.data();
    out->buckets = (int32_t)c->last_buckets;
    out->window = (int32_t)c->last_window;
    out->capacity = (int64_t)1 << c->last_cap_log2;
    for (uint32_t b = 0; b < c->last_buckets; ++b) {
        out->writes[b] = h[mlmq::kTailOff + b * mlmq::kStride];
        out->reads[b] = h[mlmq::kTakenOff + b * mlmq::kStride];
        if ((int64_t)out->writes[b] > out->capacity) ++out->wrapped;
    }
    out->completed = h[mlmq::kDoneOff];
    out->epoch = h[mlmq::kBaseOff];
    return MLMQ_OK;
}

int mlmq_copy_dist(mlmq_ctx *c, int32_t *host_out) {
    if (!c || !host_out) return MLMQ_E_INVALID;
    if (!c->solved) return fail(c, MLMQ_E_NOGRAPH, "no completed solve");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpy(host_out, c->dist, sizeof(int32_t) * (size_t)c->n, hipMemcpyDeviceToHost));
    return MLMQ_OK;
}

int mlmq_dist_device_ptr(mlmq_ctx *c, const int32_t **out) {
    if (!c || !out) return MLMQ_E_INVALID;
    if (!c->solved) return fail(c, MLMQ_E_NOGRAPH, "no completed solve");
    *out = (const int32_t *)c->dist;
    return MLMQ_OK;
}

int mlmq_probe_copy(mlmq_ctx *c, uint64_t nbytes, double *device_ms) {
    if (!c || nbytes < 4 || (nbytes & 3)) return MLMQ_E_INVALID;
    HIP_TRY(c, hipSetDevice(c->device));
    uint32_t *src = nullptr, *dst = nullptr;
    HIP_TRY(c, hipMalloc(&src, nbytes));
    hipError_t e = hipMalloc(&dst, nbytes);
    if (e != hipSuccess) {
        (void)hipFree(src);
        return hip_fail(c, e, "hipMalloc");
    }
    const uint64_t words = nbytes / 4;
    e = mlmq::launch_fill_u32(src, words, 1u, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->ev_start, c->stream);
    if (e == hipSuccess) e = mlmq::launch_copy_u32(dst, src, words, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->ev_stop, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev_start, c->ev_stop);
    (void)hipFree(src);
    (void)hipFree(dst);
    if (e != hipSuccess) return hip_fail(c, e, "copy probe");
    if (device_ms) *device_ms = ms;
    return MLMQ_OK;
}

// ---- reference driver ABI: one implicit context on device 0 ----
static mlmq_ctx *g_ref_ctx = nullptr;

static int env_int(const char *name, long lo, long hi, long fallback, long *out) {
    const char *s = std::getenv(name);
    if (!s) {
        *out = fallback;
        return MLMQ_OK;
    }
    char *end = nullptr;
    long v = std::strtol(s, &end, 10);
    if (end == s || *end || v < lo || v > hi) return MLMQ_E_INVALID;
    *out = v;
    return MLMQ_OK;
}

int sssp_init(int m, int nnz, int *RowPtr_d, int *ColIdx_d, int *w_d) {
    if (!g_ref_ctx) {
        int rc = mlmq_create(0, &g_ref_ctx);
        if (rc) return rc;
    }
    return mlmq_attach_device_csr(g_ref_ctx, m, nnz, RowPtr_d, ColIdx_d, w_d);
}

int sssp_run_adaptive(int src, mlmq_graph_info info) {
    (void)info;  // the adaptive policy is disabled on the reference path too
    if (!g_ref_ctx) return MLMQ_E_NOGRAPH;
    mlmq_params p;
    mlmq_params_default(&p);
    long v = 0;
    if (env_int("BENCH_DELTA", 1, 0x7FFFFFFF, 200000, &v)) return MLMQ_E_INVALID;
    p.delta = (uint32_t)v;
    if (env_int("MLMQ_WORK_BLOCKS", 0, 1 << 20, 0, &v)) return MLMQ_E_INVALID;
    p.work_blocks = (int32_t)v;
    return mlmq_solve(g_ref_ctx, src, &p, nullptr);
}

int sssp_cp_data(int **node_data_h) {
    if (!node_data_h || !g_ref_ctx) return MLMQ_E_INVALID;
    int32_t *buf = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)g_ref_ctx->n);
    if (!buf) return MLMQ_E_NOMEM;
    int rc = mlmq_copy_dist(g_ref_ctx, buf);
    if (rc) {
        std::free(buf);
        return rc;
    }
    *node_data_h = buf;
    return MLMQ_OK;
}

int sssp_release(void) {
    int rc = mlmq_destroy(g_ref_ctx);
    g_ref_ctx = nullptr;
    return rc;
}

}  // extern "C"
