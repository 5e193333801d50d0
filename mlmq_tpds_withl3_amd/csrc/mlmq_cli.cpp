// mlmq -- command-line driver with the reference solver's process contract
// (SS/SSSP/main.cu:146-239, SS/SSSP/paras.h:12-46, SS/SSSP/sssp_run.cu:758-886,
// SS = evidence/l3_supplement_20260921_v2/single_source):
//   mlmq -i graph.gr [-n 1] [-d delta]
//   env BENCH_SOURCE (default 0), BENCH_DELTA (default 200000),
//       BENCH_WARMUPS (default 2), BENCH_REPEATS (default 3),
//       MLMQ_WORK_BLOCKS (0 = automatic), BENCH_QUEUE (L1SLF_L2DQ | L1V_L2DQ | L1N_L2DQ),
//       L3_SUPPLEMENT_ORACLE (int32 .i32 file of expected distances).
// Prints one "BENCH ..." line per run.  With an oracle file, every run is
// checked per vertex (WIDE_ORACLE line, SSSP/supplement_oracle.h:6-14) and a
// mismatch exits with code 2; without one, correct=-1 (not checked).
#include <unistd.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mlmq.h"
#include "../../include/mlmq_graph.h"

static long env_long(const char *name, long fallback, long lo, long hi) {
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

int main(int argc, char **argv) {
    const char *input = nullptr;
    int ch;
    while ((ch = getopt(argc, argv, "i:n:d:")) != -1) {
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
        default: return 2;
        }
    }
    if (!input) {
        std::printf("Usage %s -i [graph.gr]\n", argv[0]);
        return 2;
    }
    std::printf("Graph name: %s\n", input);
    mlmq_host_graph *hg = nullptr;
    if (mlmq_graph_read_gr(input, &hg) != 0) {
        std::fprintf(stderr, "cannot read %s\n", input);
        return 2;
    }
    int32_t n = 0;
    int64_t m = 0;
    const int64_t *row;
    const int32_t *col, *w;
    mlmq_graph_dims(hg, &n, &m);
    mlmq_graph_arrays(hg, &row, &col, &w);
    std::printf("nnodes=%d, nedges=%lld.\n", n, (long long)m);
    const long src = env_long("BENCH_SOURCE", 0, 0, n - 1);

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

    std::vector<int32_t> expect;
    if (const char *op = std::getenv("L3_SUPPLEMENT_ORACLE")) {
        FILE *f = std::fopen(op, "rb");
        if (!f) {
            std::perror(op);
            return 2;
        }
        expect.resize(n);
        bool ok = std::fread(expect.data(), sizeof(int32_t), n, f) == (size_t)n && std::fgetc(f) == EOF;
        std::fclose(f);
        if (!ok) {
            std::fprintf(stderr, "ORACLE_SIZE_ERROR\n");
            return 2;
        }
    }

    mlmq_ctx *ctx = nullptr;
    int rc = mlmq_create(0, &ctx);
    if (rc == 0) rc = mlmq_load_csr(ctx, n, m, row, col, w);
    mlmq_graph_free(hg);
    if (rc) {
        std::fprintf(stderr, "setup failed: %d %s\n", rc, ctx ? mlmq_last_error(ctx) : "");
        return 2;
    }
    std::printf("MLMQ_CONFIG delta=%u queue=%s work_blocks=%d buckets=%d window=%d batch=%d\n", p.delta,
                qname.c_str(), p.work_blocks, p.buckets, p.bucket_window, p.l2_batch);
    std::vector<int32_t> dist(n);
    for (long r = 0; r < warmups + repeats; ++r) {
        mlmq_timing t;
        rc = mlmq_solve(ctx, (int32_t)src, &p, &t);
        if (rc) {
            std::fprintf(stderr, "solve failed: %d %s\n", rc, mlmq_last_error(ctx));
            mlmq_destroy(ctx);
            return 2;
        }
        rc = mlmq_copy_dist(ctx, dist.data());
        if (rc) {
            std::fprintf(stderr, "copy failed: %d %s\n", rc, mlmq_last_error(ctx));
            mlmq_destroy(ctx);
            return 2;
        }
        int correct = -1;
        if (!expect.empty()) {
            correct = 1;
            for (int32_t u = 0; u < n; ++u)
                if (dist[u] != expect[u]) {
                    std::fprintf(stderr, "ORACLE_MISMATCH vertex=%d actual=%d expected=%d\n", u, dist[u],
                                 expect[u]);
                    correct = 0;
                    break;
                }
            if (correct) std::printf("WIDE_ORACLE vertices=%d correct=1\n", n);
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
            return 2;
        }
    }
    std::printf("gpu\n");
    for (int32_t i = 0; i < n && i < 30; ++i) std::printf("%d ", dist[i]);
    std::printf("\n");
    mlmq_destroy(ctx);
    return 0;
}
