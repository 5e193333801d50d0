/*
 * mlmq_serve.h -- C ABI of the serving loop: several graphs resident in HBM,
 * single-source queries admitted by the HBM bytes they need, batched as they
 * arrive and run concurrently on the device.
 *
 * The reference has no serving loop.  Its closest pieces, which this API
 * generalises, are the multi-source loop of the driver (SSSP/main.cu:213-307:
 * one query after another on one resident graph) and the per-query workspace
 * that is reused across queries (SSSP/sssp_run.cu:1125-1224, query_workspace).
 *
 * Model
 *   - A request is (arrival time, graph, source).  mlmq_server_run() replays
 *     an arrival trace in real time on the calling thread.
 *   - Admission: a request needs one slot = its distance row (4 B per vertex)
 *     plus its queue partition (control block + buckets x capacity x 8 B).
 *     Slots come from one HBM arena of config.hbm_budget bytes (first-fit,
 *     coalescing); a request waits until a slot of its size can be carved, or
 *     is rejected on arrival when config.max_pending requests already wait.
 *     Slots stay with their graph while free (their queue partition is then
 *     already clean) and are returned to the arena when another graph needs
 *     the bytes.
 *   - Batching: whenever one of config.streams launch lanes is free, waiting
 *     requests go into one launch in arrival order, up to config.max_batch,
 *     while each can be admitted.  A launch mixes requests of different
 *     graphs and sizes (each request carries its own graph and settings);
 *     the oldest waiting request's graph fixes the launch's workgroup width
 *     (mlmq_serve_graph_params.waves), and requests of graphs with another
 *     width wait for a later launch.  Lanes run concurrently on the device,
 *     so requests that arrive while others run are admitted as lanes and
 *     bytes free up (continuous batching).
 *   - Each request runs on `workgroups` workgroups of `waves` wave64 that
 *     share its queue partition; distances are exact (the same per vertex as
 *     a CPU Dijkstra), whatever else runs beside it.
 *   - Latency = completion of the request's launch on the device (HIP event)
 *     minus its arrival time; queue time = launch time minus arrival.
 *   - Verification: with config.verify, a request whose (graph, source) has an
 *     expected row registered by mlmq_server_set_expected() is compared on the
 *     device and its mismatch count reported.
 *
 * Live serving (mlmq_server_submit / _poll / _wait / _drain / _stop): one
 * persistent launch serves requests as they are submitted.  The host writes
 * each admitted request (its graph, slot, settings and number of
 * workgroups) into a table in pinned host memory and publishes one ticket per
 * workgroup; workgroups of the running launch take tickets in order, so a
 * request starts as soon as workgroups free up and completes on its own: its
 * result is taken while the launch keeps serving others (no request waits
 * for a launch to end).  Admission is by slot bytes as above.  A service
 * thread of the server runs the admission, publication and completion loop;
 * the launch ends after config.live_idle_ms without requests, or at
 * mlmq_server_stop / _destroy, and starts again on the next submission.
 * config.mode = MLMQ_SERVE_LIVE makes mlmq_server_run replay its trace
 * through the same path.
 *
 * A server is used from one host thread (live calls may come from several
 * threads).  Every call except the live ones is synchronous.
 */
#ifndef MLMQ_SERVE_H
#define MLMQ_SERVE_H

#include <stdint.h>

#include "mlmq.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MLMQ_E_REJECTED (-10)  /* request refused at admission (max_pending waiting) */
#define MLMQ_E_PENDING  (-11)  /* live request not finished yet (mlmq_server_poll / _wait) */

#define MLMQ_SERVE_LANES 0     /* mlmq_server_run: launches of waiting requests on launch lanes */
#define MLMQ_SERVE_LIVE  1     /* mlmq_server_run: the persistent launch of live serving */

typedef struct mlmq_server mlmq_server;

typedef struct mlmq_server_config {
    uint64_t hbm_budget;      /* bytes of the slot arena; 0 = 32 GiB */
    int32_t  streams;         /* concurrent launch lanes, 1..32; 0 = 4 */
    int32_t  max_batch;       /* requests per launch, 1..4096; 0 = 64 */
    int32_t  max_pending;     /* waiting requests before arrivals are rejected; 0 = no limit */
    int32_t  verify;          /* compare results with registered expected rows */
    double   timeout_s;       /* device watchdog per launch; <= 0 means 60 s */
    int32_t  spread_waves;    /* a launch of few requests runs them on up to this many waves in
                                 total, as one-wave workgroups (at most 512 per request), when
                                 that is more than the graph's workgroups x waves: a request
                                 arriving at a quiet server gets much of the device; 0 = off
                                 (default 512) */
    int32_t  mode;            /* mlmq_server_run path: MLMQ_SERVE_LANES (default) or MLMQ_SERVE_LIVE */
    int32_t  live_waves;      /* live serving: workgroups of the persistent launch (one wave each,
                                 one of them the dispatcher); 0 = every resident one (up to 4096) */
    int32_t  live_idle_ms;    /* live serving: the launch ends after this long without requests;
                                 0 = 2000 */
    int32_t  reserved[4];
} mlmq_server_config;

/* Per-graph solver settings of served queries. */
typedef struct mlmq_serve_graph_params {
    uint32_t delta;           /* bucket width; 0 = 200000 */
    int32_t  buckets;         /* 1..64; 0 = 16 */
    int32_t  bucket_window;   /* 1..buckets; 0 = 4 */
    int32_t  l2_batch;        /* 1..64; 0 = 8 */
    int32_t  min_read;        /* 1..64; 0 or -1 = 8 */
    int32_t  l1_pop;          /* 1..64; 0 = 32 (up to 32: two lanes per record, as in mlmq_solve) */
    int32_t  l1_spill;        /* 1..64; 0 = 64 */
    int32_t  workgroups;      /* workgroups per request, 1..64; 0 = auto from the edge count */
    int32_t  waves;           /* waves per workgroup: 1, 4, 8 or 16; 0 = 8 */
    int64_t  bucket_capacity; /* ring slots per bucket (rounded up to 2^k); 0 = n / 32, >= 8192.
                                 Doubled automatically when a request overflows (the request
                                 is then run again). */
    int32_t  hub_split;       /* 1: split hub edge lists into chunk records (mlmq_params.hub_split) */
    int32_t  light_split;     /* 1: light/heavy edge split (mlmq_params.light_split) */
    int32_t  prefetch;        /* 1 = touch relaxed children's hop blocks (mlmq_params.prefetch),
                                 -1 = don't; 0 = default (1) */
    int32_t  stale_check;     /* mlmq_params.stale_check: 2 coherent, 1 through the L2, 3 late,
                                 -1 none; 0 = default (2) */
    int32_t  reserved[4];
} mlmq_serve_graph_params;

typedef struct mlmq_serve_request {
    double   arrival_s;       /* seconds after the start of mlmq_server_run; non-decreasing */
    int32_t  graph;           /* id from mlmq_server_add_* */
    int32_t  src;
    int32_t *dist_out;        /* optional host buffer of n int32 for the result row */
} mlmq_serve_request;

typedef struct mlmq_serve_result {
    double   latency_ms;      /* device completion - arrival */
    double   queue_ms;        /* launch - arrival */
    double   device_ms;       /* first workgroup start to last workgroup end */
    int32_t  status;          /* MLMQ_OK, MLMQ_E_REJECTED, _TIMEOUT, _INTERNAL, _QUEUE_OVERFLOW */
    int32_t  mismatches;      /* vertices differing from the expected row; -1 = not verified */
    int32_t  launch;          /* launch sequence number */
    int32_t  batch_size;      /* requests in that launch */
    int32_t  retries;         /* overflow reruns of this request */
    int32_t  reserved;
} mlmq_serve_result;

typedef struct mlmq_serve_stats {
    double   wall_s;          /* start to last completion */
    double   offered_qps;     /* requests / last arrival time */
    double   completed_qps;   /* completed / wall_s */
    double   p50_ms, p99_ms, max_ms, mean_ms;   /* latency of completed requests */
    double   queue_p50_ms, queue_p99_ms;
    double   mean_batch;      /* requests per launch */
    double   mean_device_ms;
    uint64_t requests, completed, rejected, failed, verified, mismatched, retries, launches;
    uint64_t peak_bytes;      /* most arena bytes held by slots */
    uint64_t peak_inflight;   /* most requests on the device at once */
    int64_t  reserved[8];
} mlmq_serve_stats;

void mlmq_server_config_default(mlmq_server_config *cfg);
void mlmq_serve_graph_params_default(mlmq_serve_graph_params *p);

/* Defaults filled in and ranges checked for a graph of n vertices and m
 * edges, as mlmq_server_add_* applies them (MLMQ_OK or MLMQ_E_INVALID).
 * Needs no device. */
int mlmq_serve_graph_params_resolve(const mlmq_serve_graph_params *in, int32_t n, int64_t m,
                                    mlmq_serve_graph_params *out);

int mlmq_server_create(int device, const mlmq_server_config *cfg, mlmq_server **out);
int mlmq_server_destroy(mlmq_server *s);
const char *mlmq_server_last_error(const mlmq_server *s);

/* Make a graph resident; same CSR conventions as mlmq_load_csr / mlmq_load_gr. */
int mlmq_server_add_csr(mlmq_server *s, int32_t n, int64_t m, const int64_t *row, const int32_t *col,
                        const int32_t *w, const mlmq_serve_graph_params *p, int32_t *graph_id);
int mlmq_server_add_gr(mlmq_server *s, const char *path, const mlmq_serve_graph_params *p,
                       int32_t *graph_id);
int mlmq_server_set_graph_params(mlmq_server *s, int32_t graph_id, const mlmq_serve_graph_params *p);
int mlmq_server_graph_size(const mlmq_server *s, int32_t graph_id, int32_t *n, int64_t *m);
/* Bytes one request of this graph holds while admitted. */
int mlmq_server_slot_bytes(const mlmq_server *s, int32_t graph_id, uint64_t *bytes);

/* Register the expected distance row of (graph, src) for verification
 * (device memory outside the arena; INT_MAX = unreachable). */
int mlmq_server_set_expected(mlmq_server *s, int32_t graph_id, int32_t src, const int32_t *dist);
int mlmq_server_clear_expected(mlmq_server *s);

/* Replay `nreq` requests in real time; results[i] belongs to reqs[i]. */
int mlmq_server_run(mlmq_server *s, const mlmq_serve_request *reqs, int32_t nreq,
                    mlmq_serve_result *results, mlmq_serve_stats *stats);

/* ---- Live serving (no reference counterpart) ----
 * Submit one request now; `dist_out` (optional, n int32) receives the row
 * before the request is reported done.  Returns its id in *request_id, or
 * MLMQ_E_REJECTED when config.max_pending requests already wait. */
int mlmq_server_submit(mlmq_server *s, int32_t graph, int32_t src, int32_t *dist_out,
                       int64_t *request_id);
/* MLMQ_OK with the result once the request is done (latency = when its
 * result was taken on the host, minus submission), MLMQ_E_PENDING before. */
int mlmq_server_poll(mlmq_server *s, int64_t request_id, mlmq_serve_result *result);
/* Waits up to timeout_s (<= 0: no limit) for the request: MLMQ_OK or MLMQ_E_PENDING. */
int mlmq_server_wait(mlmq_server *s, int64_t request_id, double timeout_s, mlmq_serve_result *result);
/* Waits until every submitted request is done (MLMQ_OK) or timeout_s passed (MLMQ_E_PENDING). */
int mlmq_server_drain(mlmq_server *s, double timeout_s);
/* Ends the persistent launch and the service thread after the submitted
 * requests are done; a later submission starts them again. */
int mlmq_server_stop(mlmq_server *s);

#ifdef __cplusplus
}
#endif

#endif
