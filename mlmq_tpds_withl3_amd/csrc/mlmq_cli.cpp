// mlmq -- command-line driver with the reference solver's process contract
// (SS/SSSP/main.cu:146-239, SS/SSSP/paras.h:12-46, SS/SSSP/sssp_run.cu:758-886,
// SS = evidence/l3_supplement_20260921_v2/single_source; multi-source mode:
// SSSP/main.cu:195-307, 396-434, SSSP/paras.h:12-60):
//   mlmq -i graph.gr [-n 1] [-d delta] [-s sources]
//   env BENCH_SOURCE (default 0), BENCH_DELTA (default 200000),
//       BENCH_WARMUPS (default 2), BENCH_REPEATS (default 3),
//       MLMQ_WORK_BLOCKS (0 = automatic), BENCH_QUEUE (L1SLF_L2DQ | L1V_L2DQ | L1N_L2DQ),
//       L3_SUPPLEMENT_ORACLE (int32 .i32 file of expected distances),
//       MLMQ_SKIP_CPU_CHECK=1 (timing-only runs: no CPU reference),
//       MLMQ_MULTI_MODE (batch: one persistent launch for all sources, default;
//                        sequential: one solve per source, as the reference does),
//       MLMQ_L2_REPORT=1 (per query, one CAPACITY_BUCKET line per bucket and a
//                        CAPACITY_QUERY line, as the reference's diagnostic build).
//
// Queue audit: an L2_CAPACITY line after the first solve and an L2_FINAL line
// per query (reads, writes and completed records of the L2 queue; the
// reference's cu_delta_queue.cuh:165 and SSSP/sssp_run.cu:10900-10926); a
// query whose per-bucket reads and writes, or their sum and the completed
// count, differ exits with code 2.
//
// One source (default): like the reference's main, a CPU Dijkstra runs first
// ("CPU total work count", "CPU sequential time"), then every GPU run prints
// one "BENCH ..." line whose correct=1|0 compares every vertex with it (and,
// with L3_SUPPLEMENT_ORACLE, with that file: WIDE_ORACLE line,
// SSSP/supplement_oracle.h:6-14).  A wrong result exits with code 2.  With
// MLMQ_SKIP_CPU_CHECK=1 and no oracle file, correct=-1 (not checked).
//
// -s N (N > 1): N sources spread evenly over the vertex ids (k * n / N), CPU
// references for all of them first, then the GPU answers all of them and
// every row is compared; "MULTI-SOURCE ..." lines as the reference prints
// them; exit code 1 if any row differs.
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <queue>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/mlmq.h"
#include "../../include/mlmq_graph.h"

namespace {

constexpr int32_t kInf = 0x7FFFFFFF;   // unreachable (the reference's INT_MAX)

long env_long(const char *name, long fallback, long lo, long hi) {
    const char *s = std::getenv(name);
    if (!s) return fallback;
    char *end = nullptr;
    errno = 0;
    long v = std::strtol(s, &end, 10);
    if (errno || end == s || *end || v < lo || v > hi) {
        std::fprintf(stderr, "Invalid %s=%s; expected [%ld,%ld]\n", name, s, lo, hi);
        std::exit(2);
    }
    return v;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// The driver's CPU reference: binary-heap Dijkstra with lazy deletion, as the
// reference's sssp_sequential (SS/SSSP/main.cu:47-117) settles vertices:
// pop the smallest tentative distance, skip a vertex already settled, add
// its out-degree to the work count, lower each neighbour strictly and push
// it.  Sums are 64-bit so that no candidate wraps; a finite distance at or
// above INT_MAX cannot be stored and is reported (the rule of the reference's
// int64 oracle, scripts/multigpu/supplement_graph.cpp:32-43).
bool cpu_reference(int32_t n, const int64_t *row, const int32_t *col, const int32_t *w, int32_t src,
                   std::vector<int32_t> &dist, int64_t &total_work) {
    std::vector<int64_t> d((size_t)n, INT64_MAX);
    std::vector<uint8_t> done((size_t)n, 0);
    using Item = std::pair<int64_t, int32_t>;
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> heap;
    d[src] = 0;
    heap.push({0, src});
    total_work = 0;
    while (!heap.empty()) {
        const int32_t u = heap.top().second;
        heap.pop();
        if (done[u]) continue;
        done[u] = 1;
        total_work += row[u + 1] - row[u];
        for (int64_t e = row[u]; e < row[u + 1]; ++e) {
            const int32_t v = col[e];
            const int64_t nd = d[u] + (int64_t)w[e];
            if (nd < d[v]) {
                d[v] = nd;
                heap.push({nd, v});
            }
        }
    }
    dist.assign((size_t)n, kInf);
    bool ok = true;
    for (int32_t v = 0; v < n; ++v) {
        if (d[v] == INT64_MAX) continue;
        if (d[v] >= kInf) ok = false;
        else dist[v] = (int32_t)d[v];
    }
    return ok;
}

// First differing vertex (-1 if none) and the number of differences.
int64_t compare_rows(const int32_t *got, const int32_t *want, int32_t n, int32_t *first_bad) {
    int64_t bad = 0;
    *first_bad = -1;
    for (int32_t v = 0; v < n; ++v)
        if (got[v] != want[v]) {
            if (*first_bad < 0) *first_bad = v;
            ++bad;
        }
    return bad;
}

void print_row(const char *label, const int32_t *d, int32_t n) {
    std::printf("%s\n", label);
    for (int32_t i = 0; i < n && i < 30; ++i) std::printf("%d ", d[i]);
    std::printf("\n");
}

// Queue audit of the last solve (see the header comment).  False when the
// counts do not balance.
bool queue_audit(mlmq_ctx *ctx, bool first, bool per_bucket) {
    mlmq_queue_counts q;
    if (mlmq_last_queue_counts(ctx, &q) != 0) {
        std::fprintf(stderr, "queue counts unavailable\n");
        return false;
    }
    if (first)
        std::printf("L2_CAPACITY budget=%lld record_bytes=8 buckets=%d per_bucket=%lld "
                    "allocated_records=%lld counter_bits=32\n",
                    (long long)(8 * q.capacity * q.buckets), q.buckets, (long long)q.capacity,
                    (long long)(q.capacity * q.buckets));
    long long reads = 0, writes = 0;
    bool ok = true;
    for (int b = 0; b < q.buckets; ++b) {
        reads += q.reads[b];
        writes += q.writes[b];
        ok = ok && q.reads[b] == q.writes[b];
        if (per_bucket)
            std::printf("CAPACITY_BUCKET gpu=0 bucket=%d capacity=%lld writes=%u reads=%u no_wrap=%d\n", b,
                        (long long)q.capacity, q.writes[b], q.reads[b],
                        (long long)q.writes[b] <= q.capacity ? 1 : 0);
    }
    ok = ok && writes == (long long)q.completed;
    if (per_bucket)
        std::printf("CAPACITY_QUERY gpu=0 total=%lld completed=%u guarded=1\n", writes, q.completed);
    std::printf("L2_FINAL gpu=0 buckets=%d reads=%lld writes=%lld completed=%u\n", q.buckets, reads, writes,
                q.completed);
    if (!ok) std::fprintf(stderr, "L2 conservation failed\n");
    return ok;
}

struct Graph {
    int32_t n = 0;
    int64_t m = 0;
    const int64_t *row = nullptr;
    const int32_t *col = nullptr;
    const int32_t *w = nullptr;
};

int run_multi_source(mlmq_ctx *ctx, const Graph &g, const mlmq_params &p, int32_t nsrc) {
    std::vector<int32_t> ids((size_t)nsrc);
    for (int32_t k = 0; k < nsrc; ++k)
        ids[k] = nsrc <= g.n ? (int32_t)((int64_t)k * g.n / nsrc) : k % g.n;
    const uint64_t ref_bytes = (uint64_t)nsrc * (uint64_t)g.n * sizeof(int32_t);
    if (ref_bytes > (8ull << 30)) {
        std::fprintf(stderr, "MULTI-SOURCE: %d sources x %d vertices need %.1f GiB of CPU references; "
                             "use fewer sources\n", nsrc, g.n, ref_bytes / 1073741824.0);
        return 2;
    }
    // CPU references, all before the GPU is timed (SSSP/main.cu:412-430), on
    // a thread pool of independent Dijkstras.
    std::vector<std::vector<int32_t>> refs((size_t)nsrc);
    std::atomic<int32_t> next(0);
    std::atomic<bool> ref_ok(true);
    const double c0 = now_ms();
    {
        const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        for (int t = 0; t < nt; ++t)
            pool.emplace_back([&]() {
                int64_t work = 0;
                for (int32_t k; (k = next.fetch_add(1)) < nsrc;)
                    if (!cpu_reference(g.n, g.row, g.col, g.w, ids[k], refs[k], work)) ref_ok = false;
            });
        for (auto &t : pool) t.join();
    }
    std::printf("MULTI-SOURCE CPU references ready: sources=%d time=%.2f ms\n", nsrc, now_ms() - c0);
    if (!ref_ok) {
        std::fprintf(stderr, "a finite distance reaches INT_MAX: not representable\n");
        return 2;
    }
    const char *mode_env = std::getenv("MLMQ_MULTI_MODE");
    const std::string mode = mode_env ? mode_env : "batch";
    std::printf("MULTI-SOURCE GPU0 full graph: v=%d e=%lld\n", g.n, (long long)g.m);
    std::printf("MULTI-SOURCE begin: sources=%d physical_gpus=1 logical_gpus=1 mode=%s\n", nsrc,
                mode.c_str());
    std::fflush(stdout);

    // Device rows of one batch launch stay below 16 GiB; more sources run as
    // several launches.
    const int64_t per_launch = std::max<int64_t>(1, (int64_t)((16ull << 30) / (4ull * (uint64_t)g.n)));
    std::vector<std::vector<int32_t>> got((size_t)nsrc, std::vector<int32_t>((size_t)g.n));
    const double t0 = now_ms();
    int rc = 0;
    if (mode == "batch") {
        for (int64_t b = 0; b < nsrc && rc == 0; b += per_launch) {
            const int32_t cnt = (int32_t)std::min<int64_t>(per_launch, nsrc - b);
            mlmq_batch_timing bt;
            rc = mlmq_solve_batch(ctx, ids.data() + b, cnt, &p, &bt);
            for (int32_t k = 0; k < cnt && rc == 0; ++k) rc = mlmq_copy_batch_dist(ctx, k, got[b + k].data());
        }
    } else {
        for (int32_t k = 0; k < nsrc && rc == 0; ++k) {
            rc = mlmq_solve(ctx, ids[k], &p, nullptr);
            if (rc == 0) rc = mlmq_copy_dist(ctx, got[k].data());
        }
    }
    const double batch_ms = now_ms() - t0;
    if (rc) {
        std::fprintf(stderr, "MULTI-SOURCE solve failed: %d %s\n", rc, mlmq_last_error(ctx));
        return 2;
    }
    int64_t total_errors = 0;
    for (int32_t k = 0; k < nsrc; ++k) {
        int32_t first_bad = -1;
        const int64_t e = compare_rows(got[k].data(), refs[k].data(), g.n, &first_bad);
        total_errors += e;
        if (e == 0) std::printf("MULTI-SOURCE source[%d]=%d correct\n", k, ids[k]);
        else std::printf("MULTI-SOURCE source[%d]=%d errors=%lld first_bad=%d\n", k, ids[k], (long long)e, first_bad);
    }
    std::printf("MULTI-SOURCE sources=%d nGPU=1 batch wall time=%.2f ms\n", nsrc, batch_ms);
    std::printf("MULTI-SOURCE average query time=%.2f ms\n", batch_ms / nsrc);
    std::printf("MULTI-SOURCE errors=%lld\n", (long long)total_errors);
    if (total_errors == 0) std::printf("MULTI-SOURCE all sources correct!\n");
    return total_errors == 0 ? 0 : 1;
}

}  // namespace

int main(int argc, char **argv) {
    const char *input = nullptr;
    long nsrc = 1;
    int ch;
    while ((ch = getopt(argc, argv, "i:n:d:s:")) != -1) {
        switch (ch) {
        case 'i': input = optarg; break;
        case 'n':
            if (std::atoi(optarg) != 1) {
                std::fprintf(stderr, "only -n 1 (one GPU) is supported\n");
                return 2;
            }
            break;
        case 'd':
            if (std::atoi(optarg) <= 0) return 2;
            setenv("BENCH_DELTA", optarg, 1);
            break;
        case 's':
            nsrc = std::atol(optarg);
            if (nsrc < 1 || nsrc > (1 << 20)) {
                std::printf("sources must be >= 1, got %s\n", optarg);
                return 2;
            }
            break;
        default: return 2;
        }
    }
    if (!input) {
        std::printf("Usage %s -i [graph.gr] [-n 1] [-s sources] [-d delta]\n", argv[0]);
        return 2;
    }
    std::printf("Graph name: %s\n", input);
    mlmq_host_graph *hg = nullptr;
    if (mlmq_graph_read_gr(input, &hg) != 0) {
        std::fprintf(stderr, "cannot read %s\n", input);
        return 2;
    }
    Graph g;
    mlmq_graph_dims(hg, &g.n, &g.m);
    mlmq_graph_arrays(hg, &g.row, &g.col, &g.w);
    std::printf("nnodes=%d, nedges=%lld.\n", g.n, (long long)g.m);
    for (int64_t e = 0; e < g.m; ++e)
        if (g.w[e] < 0) {
            std::printf("Negative edge weight!\n");
            return 2;
        }
    const long src = env_long("BENCH_SOURCE", 0, 0, g.n - 1);

    mlmq_params p;
    mlmq_params_default(&p);
    p.delta = (uint32_t)env_long("BENCH_DELTA", 200000, 1, 0x7FFFFFFF);
    p.work_blocks = (int32_t)env_long("MLMQ_WORK_BLOCKS", 0, 0, 1 << 20);
    const char *queue = std::getenv("BENCH_QUEUE");
    std::string qname = queue ? queue : "L1SLF_L2DQ";
    if (qname == "L1SLF_L2DQ") p.queue_mode = MLMQ_QUEUE_L1SLF_L2DQ;
    else if (qname == "L1N_L2DQ") p.queue_mode = MLMQ_QUEUE_L1N_L2DQ;
    else if (qname == "L1V_L2DQ") p.queue_mode = MLMQ_QUEUE_L1V_L2DQ;
    else {
        std::fprintf(stderr, "unsupported BENCH_QUEUE=%s\n", qname.c_str());
        return 2;
    }
    const long warmups = env_long("BENCH_WARMUPS", 2, 0, 10000);
    const long repeats = env_long("BENCH_REPEATS", 3, 1, 10000);
    const bool skip_cpu = env_long("MLMQ_SKIP_CPU_CHECK", 0, 0, 1) == 1;
    const bool report_buckets = env_long("MLMQ_L2_REPORT", 0, 0, 1) == 1;

    std::vector<int32_t> expect;
    if (const char *op = std::getenv("L3_SUPPLEMENT_ORACLE")) {
        if (nsrc > 1) {
            std::fprintf(stderr, "L3_SUPPLEMENT_ORACLE holds one source's distances; not with -s\n");
            return 2;
        }
        FILE *f = std::fopen(op, "rb");
        if (!f) {
            std::perror(op);
            return 2;
        }
        expect.resize(g.n);
        bool ok = std::fread(expect.data(), sizeof(int32_t), g.n, f) == (size_t)g.n && std::fgetc(f) == EOF;
        std::fclose(f);
        if (!ok) {
            std::fprintf(stderr, "ORACLE_SIZE_ERROR\n");
            return 2;
        }
    }
    if (const char *mm = std::getenv("MLMQ_MULTI_MODE")) {
        if (std::strcmp(mm, "batch") != 0 && std::strcmp(mm, "sequential") != 0) {
            std::fprintf(stderr, "unsupported MLMQ_MULTI_MODE=%s\n", mm);
            return 2;
        }
    }
    if (nsrc > 1 && p.queue_mode != MLMQ_QUEUE_L1SLF_L2DQ && !std::getenv("MLMQ_MULTI_MODE")) {
        // the batched launch runs L1SLF_L2DQ; other queues answer one by one
        setenv("MLMQ_MULTI_MODE", "sequential", 1);
    }

    // CPU reference run (SS/SSSP/main.cu:179-199), before any GPU work.
    std::vector<int32_t> cpu;
    if (nsrc == 1 && !skip_cpu) {
        int64_t work = 0;
        const double c0 = now_ms();
        const bool ok = cpu_reference(g.n, g.row, g.col, g.w, (int32_t)src, cpu, work);
        const double cpu_ms = now_ms() - c0;
        if (!ok) {
            std::fprintf(stderr, "a finite distance reaches INT_MAX: not representable\n");
            return 2;
        }
        std::printf("CPU done!\n");
        std::printf("CPU total work count %lld\n", (long long)work);
        std::printf("CPU sequential time %.2f ms\n", cpu_ms);
    }

    mlmq_ctx *ctx = nullptr;
    int rc = mlmq_create(0, &ctx);
    if (rc == 0) rc = mlmq_load_csr(ctx, g.n, g.m, g.row, g.col, g.w);
    if (rc) {
        std::fprintf(stderr, "setup failed: %d %s\n", rc, ctx ? mlmq_last_error(ctx) : "");
        mlmq_graph_free(hg);
        return 2;
    }
    std::printf("MLMQ_CONFIG delta=%u queue=%s work_blocks=%d buckets=%d window=%d batch=%d\n", p.delta,
                qname.c_str(), p.work_blocks, p.buckets, p.bucket_window, p.l2_batch);
    if (nsrc > 1) {
        rc = run_multi_source(ctx, g, p, (int32_t)nsrc);
        mlmq_destroy(ctx);
        mlmq_graph_free(hg);
        return rc;
    }
    std::printf("GPU begin!\n");
    std::vector<int32_t> dist(g.n);
    for (long r = 0; r < warmups + repeats; ++r) {
        mlmq_timing t;
        rc = mlmq_solve(ctx, (int32_t)src, &p, &t);
        if (rc) {
            std::fprintf(stderr, "solve failed: %d %s\n", rc, mlmq_last_error(ctx));
            mlmq_destroy(ctx);
            mlmq_graph_free(hg);
            return 2;
        }
        rc = mlmq_copy_dist(ctx, dist.data());
        if (rc) {
            std::fprintf(stderr, "copy failed: %d %s\n", rc, mlmq_last_error(ctx));
            mlmq_destroy(ctx);
            mlmq_graph_free(hg);
            return 2;
        }
        if (!queue_audit(ctx, r == 0, report_buckets)) {
            mlmq_destroy(ctx);
            mlmq_graph_free(hg);
            return 2;
        }
        int correct = -1;
        if (!cpu.empty()) {
            int32_t first_bad = -1;
            correct = compare_rows(dist.data(), cpu.data(), g.n, &first_bad) == 0 ? 1 : 0;
            if (!correct)
                std::fprintf(stderr, "Error at node %d: GPU=%d CPU=%d\n", first_bad, dist[first_bad],
                             cpu[first_bad]);
        }
        if (!expect.empty()) {
            int32_t first_bad = -1;
            if (compare_rows(dist.data(), expect.data(), g.n, &first_bad) != 0) {
                std::fprintf(stderr, "ORACLE_MISMATCH vertex=%d actual=%d expected=%d\n", first_bad,
                             dist[first_bad], expect[first_bad]);
                correct = 0;
            } else {
                std::printf("WIDE_ORACLE vertices=%d correct=1\n", g.n);
                if (correct < 0) correct = 1;
            }
        }
        std::printf("BENCH algorithm=MLMQ baseline=mi355x gpu_count=1 source=%ld repeat=%ld warmup=%d "
                    "queue=%s solve_ms=%.6f device_ms=%.6f query_wall_ms=%.6f work_blocks=%d "
                    "l2_pushes=%llu relaxations=%llu correct=%d\n",
                    src, r, r < warmups ? 1 : 0, qname.c_str(), t.solve_ms, t.device_ms, t.query_wall_ms,
                    t.work_blocks, (unsigned long long)t.l2_pushes, (unsigned long long)t.relaxations,
                    correct);
        std::fflush(stdout);
        if (correct == 0) {
            mlmq_destroy(ctx);
            mlmq_graph_free(hg);
            return 2;
        }
    }
    std::printf("GPU done!\n");
    print_row("gpu", dist.data(), g.n);
    if (!cpu.empty()) {
        print_row("cpu", cpu.data(), g.n);
        std::printf("mlmq sssp correct!\n");
    }
    mlmq_destroy(ctx);
    mlmq_graph_free(hg);
    return 0;
}
