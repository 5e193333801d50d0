/*
 * mlmq.h -- C ABI of the MI355X (gfx950) MLMQ single-source shortest-path
 * solver: a per-wavefront "small label first" (SLF) deque in LDS over a global
 * delta-bucket queue, drained by persistent wavefronts that relax a CSR graph
 * with atomicMin on 32-bit distances.
 *
 * Conventions
 *   - Plain C types only.  Every call returns 0 on success or a negative
 *     MLMQ_E* code; nothing calls exit().  mlmq_last_error() gives the text.
 *   - All state lives in an mlmq_ctx: contexts are independent (one per host
 *     thread; a context is not safe for concurrent use from two threads).
 *   - Every call is synchronous with respect to the host.
 *   - Distances are int32 with 0x7FFFFFFF (INT_MAX) for unreachable vertices,
 *     equal per vertex to a CPU Dijkstra on the same graph.
 *
 * Reference interface replaced (paths relative to the reference root;
 * SS = evidence/l3_supplement_20260921_v2/single_source):
 *   mlmq_create / mlmq_destroy      <- the solver's file-scope globals and
 *                                      their allocation, SS/SSSP/sssp_run.cu:21-30, 676-707
 *   mlmq_load_gr                    <- CSRGraph::read + copy_to_gpu,
 *                                      SSSP/csr_graph.cu:185-300, 427-438
 *   mlmq_load_csr                   <- copy_to_gpu of a host CSR, SSSP/csr_graph.cu:427-438
 *   mlmq_attach_device_csr          <- sssp_init(m, nnz, RowPtr_d, ColIdx_d, w_d),
 *                                      SS/SSSP/sssp_run.cu:676-707 (caller-owned device CSR)
 *   mlmq_solve                      <- sssp_run_adaptive(src, info) -> kernel_adaptive,
 *                                      SS/SSSP/sssp_run.cu:888-976, 758-886
 *                                      (reset = sssp_re_init :709-747 + reinit_host)
 *   mlmq_copy_dist                  <- sssp_cp_data, SS/SSSP/sssp_run.cu:749-756
 *                                      (caller-owned buffer, as SSSP/sssp_run.cu:10868-10877)
 *   mlmq_params fields              <- mlmq_setup (core/include/adaptive.h:6-59) and the
 *                                      env knobs BENCH_DELTA, MLMQ_WORK_BLOCKS,
 *                                      BENCH_QUEUE (SS/SSSP/sssp_run.cu:896;
 *                                      SSSP/sssp_run.cu:11109; SSSP/main.cu:437-442)
 *   sssp_init / sssp_run_adaptive / sssp_cp_data
 *                                   <- the exact reference driver ABI,
 *                                      SS/SSSP/main.cu:15-17 (single implicit context)
 */
#ifndef MLMQ_H
#define MLMQ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MLMQ_OK                 0
#define MLMQ_E_INVALID        (-1)   /* bad argument or malformed graph */
#define MLMQ_E_NOGRAPH        (-2)   /* solve/copy before a graph was loaded */
#define MLMQ_E_HIP            (-3)   /* HIP runtime error */
#define MLMQ_E_NOMEM          (-4)   /* device or host allocation failed */
#define MLMQ_E_IO             (-5)   /* file open/read/write failed */
#define MLMQ_E_QUEUE_OVERFLOW (-6)   /* an L2 bucket ring ran full; retry with more capacity */
#define MLMQ_E_TIMEOUT        (-7)   /* the on-device watchdog stopped the solve */
#define MLMQ_E_INTERNAL       (-8)   /* queue accounting did not balance after a solve */
#define MLMQ_E_UNSUPPORTED    (-9)   /* queue mode or parameter not available */

/* Queue modes (values follow the reference's enum mlmq_type,
 * core/include/common.h:85-108). */
#define MLMQ_QUEUE_L1V_L2DQ     0    /* L1 is a FIFO in LDS (l1_vector_queue), same spill rule */
#define MLMQ_QUEUE_L1N_L2DQ     1    /* no L1: every push and pop goes through L2 */
#define MLMQ_QUEUE_L1SLF_L2DQ  19    /* the reference's default path */

typedef struct mlmq_ctx mlmq_ctx;

typedef struct mlmq_params {
    int32_t  queue_mode;      /* MLMQ_QUEUE_L1SLF_L2DQ (default), _L1V_L2DQ or _L1N_L2DQ */
    uint32_t delta;           /* bucket width; reference BENCH_DELTA, default 200000 */
    int32_t  work_blocks;     /* persistent workgroups of one wave64 each; 0 = automatic
                                 (blocks_per_cu x CUs if blocks_per_cu > 0, else n / 25000
                                 clamped to [128, 512]); mlmq_solve_batch: workgroups of
                                 batch_waves waves, one query each (0 = as many as the device
                                 holds at once, at most the number of queries) */
    int32_t  blocks_per_cu;   /* used when work_blocks == 0 */
    int32_t  buckets;         /* delta buckets (reference BNUM), 1..64, default 16 */
    int32_t  bucket_window;   /* buckets readable at once (BUCKET_MAX), default 4 */
    int32_t  l2_batch;        /* records per L2 claim (l2_batch_size), 1..64, default 8;
                                 at most 32 when l1_pop <= 32 (a batch holds 32 records) */
    int32_t  min_read;        /* top up from L2 below this many (MIN_READ_GRA), 1..64, default 8
                                 (0 is refused: no wave would ever take the seed) */
    int64_t  bucket_capacity; /* ring slots per bucket (rounded up to 2^k); 0 = auto */
    double   timeout_s;       /* device watchdog; <= 0 means 60 s */
    int32_t  auto_retry;      /* on MLMQ_E_QUEUE_OVERFLOW grow capacity and rerun (default 1) */
    int32_t  profile;         /* 1: per-phase clock build of the kernel (mlmq_timing.phase_cycles) */
    int32_t  l1_pop;          /* records a wave takes from its L1 per batch, 1..64, default 32.
                                 Up to 32 (record-only L1): two lanes per record, each relaxing
                                 half of its hop block, and L2 top-ups also stop at 32 records
                                 (min_read above 32 acts as 32); 33..64: one lane per record */
    int32_t  l1_spill;        /* L1 holding more records than this after a push spills to L2;
                                 1..l1_ring - 448 (64 with the default ring), default 64 */
    int32_t  l1_donate;       /* while some waves are idle, a wave whose L1 holds more than this
                                 moves the back half to L2; 0 = never (default) */
    int32_t  batch_waves;     /* mlmq_solve_batch: waves per query (workgroup size), 4, 8 or 16;
                                 default 8 */
    int32_t  use_graph;       /* mlmq_solve_batch: replay the launch as a captured hipGraph
                                 (default 1) */
    int32_t  hub_split;       /* 1: a popped hub (> 2048 out-edges) keeps its first 1024-edge
                                 chunk and sends the rest to L2 as chunk records that other
                                 waves relax (with light_split: the heavy-phase part only);
                                 0 (default): one wave relaxes the whole list */
    int32_t  stale_check;     /* filter of popped records whose distance has since dropped:
                                 2 (default) coherent read of dist[v] with the hop block, 1 load
                                 through the L2 (may miss a drop made on another XCD), 3 late:
                                 coherent read in the atomicMins' round trip (a stale record's
                                 inline atomics are issued, its pushes and CSR edges dropped),
                                 0 none. A coherent read is atomicMin(dist[v], d), which cannot
                                 change dist[v] (<= d). Results are exact in every mode; only
                                 the work differs */
    int32_t  prefetch;        /* 1 (default): while its atomicMin is in flight, touch each
                                 relaxed child's hop block so a later pop finds it in L2 */
    int32_t  l1_blocks;       /* mlmq_solve with an L1: 1 = each L1 record carries its hop block
                                 in LDS, loaded while the record's atomicMin was in flight, so
                                 popping it needs no global load; 0 (default) = records only
                                 (measured faster on MI355X: the extra loads and LDS block
                                 traffic cost more than the round trip they save) */
    int32_t  idle_backoff;    /* longest idle sleep between polls of an empty queue, in units of
                                 256 cycles, 1..64; 0 = 16 */
    int32_t  defer_heavy;     /* 1: a popped record with more than 7 out-edges whose distance lies
                                 beyond the head bucket goes back to L2 instead of being expanded
                                 (power-law graphs: hubs expand in bucket order); 0 (default) */
    int32_t  strict_window;   /* 1: the bucket window advances only when every record taken from
                                 L2 has been credited back (a delta-stepping phase barrier);
                                 0 (default): as soon as the head bucket is drained */
    int32_t  light_split;     /* 1: light/heavy edge split (delta-stepping's): a record relaxes its
                                 vertex's edges of weight < delta at once and queues one
                                 heavy-phase record, one bucket later, that relaxes the edges of
                                 weight >= delta; needs a graph loaded through mlmq_load_csr /
                                 mlmq_load_gr (adjacency sorted by weight); 0 (default) */
    int32_t  csr_stream;      /* 1: CSR columns and weights of vertices with more than 7 out-edges
                                 are read with non-temporal loads, so streaming a hub's edge list
                                 does not evict the distance lines its relaxations test; 0 (default) */
    int32_t  l1_ring;         /* L1 ring size per wave in records (the LDS segment: 8 bytes each):
                                 0 or 512 (default); mlmq_solve_batch also takes 1024, which lets
                                 l1_spill go up to 255 (the spill threshold must leave room for one
                                 batch's pushes: l1_spill <= l1_ring - 448) */
    int32_t  batch_stats;     /* mlmq_solve_batch: 1 (default) counts relaxations, improvements,
                                 pushes, stale records, batches and idle polls into
                                 mlmq_batch_timing; 0 leaves them out (zero there): the counters
                                 then need no registers and a CU holds more waves */
    int32_t  reserved[2];
} mlmq_params;

/* Per-solve statistics (all counts summed over wavefronts). */
typedef struct mlmq_timing {
    double   solve_ms;        /* host steady clock around the persistent launch */
    double   device_ms;       /* HIP events around the persistent launch */
    double   reset_ms;        /* per-query reset (device fill + seed) */
    double   query_wall_ms;   /* reset + solve + nothing else (copy-out is separate) */
    int32_t  work_blocks;     /* workgroups launched */
    int32_t  waves;           /* persistent wavefronts launched */
    int64_t  bucket_capacity; /* ring slots per bucket used */
    uint64_t l2_pushes;       /* records reserved in L2 (incl. the seed) */
    uint64_t l2_pops;         /* records taken from L2 */
    uint64_t relaxations;     /* edges examined */
    uint64_t atomics;         /* atomicMin issued: edges examined less those skipped by the
                                 heavy loop's cached pre-check or as a record's back edge
                                 (the edge to the vertex whose expansion created it) */
    uint64_t improvements;    /* atomicMin that lowered a distance (= records pushed) */
    uint64_t stale;           /* popped records skipped as stale */
    uint64_t batches;         /* processed record batches */
    uint64_t idle_polls;      /* idle iterations of all waves */
    uint64_t retries;         /* overflow retries performed by mlmq_solve */
    /* params.profile = 1 only: shader-clock cycles summed over waves spent in
     * L1 pop, L2 pop, CSR-range fetch, edge relaxation, push/flush, idle,
     * total, credit return */
    uint64_t phase_cycles[8];
    /* params.profile = 1 only: cycles in owner computation, edge read (incl.
     * the stale check wait), atomicMin round trip; 16 empty laps (clock
     * overhead); [4] unused */
    uint64_t detail_cycles[5];
    uint64_t donated;         /* records moved from L1 to L2 for idle waves */
    uint64_t hub_chunks;      /* chunk records pushed by hub splitting */
} mlmq_timing;

/* Statistics of one mlmq_solve_batch call. */
typedef struct mlmq_batch_timing {
    double   solve_ms;        /* host steady clock around the launch (row reset + batch kernel) */
    double   device_ms;       /* HIP events around the launch */
    double   query_ms_p50;    /* per-query time inside the launch (device wall clock), median */
    double   query_ms_p99;
    double   query_ms_max;
    double   query_ms_mean;
    int32_t  queries;
    int32_t  work_blocks;     /* workgroups = queries in flight */
    int32_t  waves_per_query;
    int32_t  graph_replay;    /* 1 if the launch was a hipGraph replay */
    int64_t  bucket_capacity;
    uint64_t relaxations;
    uint64_t improvements;
    uint64_t l2_pushes;
    uint64_t stale;
    uint64_t batches;
    uint64_t idle_polls;
    uint64_t retries;
    uint64_t reserved[8];
} mlmq_batch_timing;

/* Reference graph statistics struct (core/include/graph_info.h:8-20). */
typedef struct mlmq_graph_info {
    int32_t m;          /* vertices (the reference names them m) */
    int32_t nnz;        /* edges */
    float avg_nnz, max_nnz, dev_nnz;
    float avg_weight, dev_weight, max_weight;
    float diameter;     /* -1: no estimate */
} mlmq_graph_info;

void        mlmq_params_default(mlmq_params *p);
/* Range check of a parameter set, as mlmq_solve (batch = 0) or
 * mlmq_solve_batch (batch = 1) applies it before touching the device:
 * MLMQ_OK, MLMQ_E_INVALID or MLMQ_E_UNSUPPORTED.  Needs no device.  No
 * reference counterpart (the reference fixes its settings at compile time,
 * core/include/adaptive.h:35-59). */
int         mlmq_params_check(const mlmq_params *p, int32_t batch);
int         mlmq_create(int device, mlmq_ctx **out);
int         mlmq_destroy(mlmq_ctx *ctx);
const char *mlmq_last_error(const mlmq_ctx *ctx);
const char *mlmq_version(void);

/* Graph upload.  Host arrays are copied; `row` has n+1 entries with row[0] = 0
 * and row[n] = m, `col` holds destinations in [0, n), `w` weights >= 0. */
int mlmq_load_csr(mlmq_ctx *ctx, int32_t n, int64_t m, const int64_t *row,
                  const int32_t *col, const int32_t *w);
int mlmq_load_gr(mlmq_ctx *ctx, const char *path);
/* Use a caller-owned device CSR (int32 row[n+1], col[m], w[m]) without a copy.
 * The arrays must stay valid until another graph is loaded or the context is
 * destroyed. */
int mlmq_attach_device_csr(mlmq_ctx *ctx, int32_t n, int64_t m,
                           const int32_t *row_d, const int32_t *col_d, const int32_t *w_d);
int mlmq_graph_stats(mlmq_ctx *ctx, mlmq_graph_info *info);
int mlmq_graph_size(const mlmq_ctx *ctx, int32_t *n, int64_t *m);

/* One query: reset, seed `src`, run the persistent solver to termination.
 * `params` may be NULL (defaults); `timing` may be NULL. */
int mlmq_solve(mlmq_ctx *ctx, int32_t src, const mlmq_params *params, mlmq_timing *timing);
/* Ring slots (8 bytes each) of the static device workspace that
 * mlmq_solve_noalloc uses: buckets x bucket_capacity may not exceed it. */
#define MLMQ_NOALLOC_RING_SLOTS (1u << 23)

/* mlmq_solve without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The query
 * is the same as mlmq_solve's on the context's graph (results identical per
 * vertex); the n distances go to the caller's device buffer `dist_d` (int32,
 * INT_MAX = unreachable), and the control block and bucket rings live in a
 * static workspace of the library (64 MiB + 64 KiB per device, held until the
 * process ends), cleared on every call.  Limit: buckets x bucket_capacity <=
 * MLMQ_NOALLOC_RING_SLOTS; an explicit bucket_capacity beyond it is refused
 * with MLMQ_E_INVALID, the automatic one is clamped to it, and overflow
 * retries stop at it (MLMQ_E_QUEUE_OVERFLOW).  The context's own last result
 * (mlmq_copy_dist) is left untouched.  Calls that overlap on one device share
 * the workspace and race: run them one at a time per device. */
int mlmq_solve_noalloc(mlmq_ctx *ctx, int32_t src, const mlmq_params *params, int32_t *dist_d,
                       mlmq_timing *timing);

/* Many queries in one persistent launch (no reference counterpart in the
 * single-GPU solver; the reference answers a source list with one solve
 * after another, SSSP/main.cu:213-307).  Each workgroup of
 * params->batch_waves waves owns a private queue partition and serves queries
 * taken from a global counter; params->work_blocks workgroups are in flight
 * (0 = one per CU, at most nq).  Results stay on the device until copied
 * with mlmq_copy_batch_dist. */
int mlmq_solve_batch(mlmq_ctx *ctx, const int32_t *sources, int32_t nq,
                     const mlmq_params *params, mlmq_batch_timing *timing);
/* Copy the n distances of query `query` of the last batch (caller-owned). */
int mlmq_copy_batch_dist(mlmq_ctx *ctx, int32_t query, int32_t *host_out);

/* Queue counts of the last mlmq_solve, per bucket: the reference's capacity
 * and conservation audit (L2_CAPACITY at queue setup,
 * core/cu_delta_queue/cu_delta_queue.cuh:165; CAPACITY_BUCKET and
 * CAPACITY_QUERY per query, SS/SSSP/supplement_capacity.h:2-13 of the
 * diagnostic build; L2_FINAL, SSSP/sssp_run.cu:10900-10926).  After every
 * successful solve writes[b] == reads[b] for every bucket and their sum ==
 * completed.  A bucket whose writes exceed its capacity reused ring slots:
 * each reuse waited until the earlier record was taken (slot-level
 * publication), so nothing was overwritten. */
typedef struct mlmq_queue_counts {
    int32_t  buckets;
    int32_t  window;
    int64_t  capacity;        /* ring slots per bucket */
    uint32_t writes[64];      /* records reserved in bucket b (the reference's write_reserve) */
    uint32_t reads[64];       /* records taken from bucket b (bucket_read_done) */
    uint32_t completed;       /* credits returned (read_done) */
    uint32_t epoch;           /* window advances (the reference's first_pos) */
    uint32_t wrapped;         /* buckets whose writes exceeded the capacity */
    uint32_t reserved[5];
} mlmq_queue_counts;
int mlmq_last_queue_counts(const mlmq_ctx *ctx, mlmq_queue_counts *out);

/* Copy the n distances of the last solve into host memory (caller-owned). */
int mlmq_copy_dist(mlmq_ctx *ctx, int32_t *host_out);
/* Device pointer to the distance array of the last solve (n int32). */
int mlmq_dist_device_ptr(mlmq_ctx *ctx, const int32_t **out);

/* Occupancy query: resident workgroups per CU for the solver kernel and CU count. */
int mlmq_device_shape(mlmq_ctx *ctx, int32_t *blocks_per_cu_max, int32_t *cus,
                      int32_t *threads_per_block);

/* Calibration probe, no reference counterpart: one plain device-to-device
 * copy kernel of `nbytes` (a multiple of 4), timed with HIP events.  Used to
 * scale hardware byte counters (rocprofv3 FETCH_SIZE / WRITE_SIZE) and as an
 * HBM bandwidth reference. */
int mlmq_probe_copy(mlmq_ctx *ctx, uint64_t nbytes, double *device_ms);

/* ---- Reference driver ABI (SS/SSSP/main.cu:15-17), one implicit context ----
 * sssp_init takes a device CSR in the reference layout (int32, row[0] = 0).
 * sssp_run_adaptive reads BENCH_DELTA / MLMQ_WORK_BLOCKS from the environment
 * like the reference and returns 0 or a negative MLMQ_E* code.
 * sssp_cp_data malloc()s an m-entry array the caller frees. */
int sssp_init(int m, int nnz, int *RowPtr_d, int *ColIdx_d, int *w_d);
int sssp_run_adaptive(int src, mlmq_graph_info info);
int sssp_cp_data(int **node_data_h);
int sssp_release(void);

#ifdef __cplusplus
}
#endif
#endif /* MLMQ_H */
