"""ctypes binding of ``libmlmq.so`` (include/mlmq.h, include/mlmq_graph.h,
include/mlmq_serve.h).

The library is built in-tree by ``make -C mlmq_tpds_withl3_amd/csrc``; this
module never builds or falls back to anything: if the shared object is missing
the import of the solver fails with a RuntimeError naming the path.
"""
from __future__ import annotations

import ctypes
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libmlmq.so")
CLI_PATH = os.path.join(PKG_DIR, "mlmq")

MLMQ_OK = 0
ERRORS = {
    -1: "MLMQ_E_INVALID", -2: "MLMQ_E_NOGRAPH", -3: "MLMQ_E_HIP", -4: "MLMQ_E_NOMEM",
    -5: "MLMQ_E_IO", -6: "MLMQ_E_QUEUE_OVERFLOW", -7: "MLMQ_E_TIMEOUT",
    -8: "MLMQ_E_INTERNAL", -9: "MLMQ_E_UNSUPPORTED", -10: "MLMQ_E_REJECTED",
    -11: "MLMQ_E_PENDING",
}
MLMQ_E_PENDING = -11
SERVE_MODES = {"lanes": 0, "live": 1}
QUEUE_MODES = {"L1SLF_L2DQ": 19, "L1N_L2DQ": 1, "L1V_L2DQ": 0}
NOALLOC_RING_SLOTS = 1 << 23   # MLMQ_NOALLOC_RING_SLOTS

# Every exported symbol the headers declare (checked by the CPU test suite).
SOLVER_SYMBOLS = [
    "mlmq_params_default", "mlmq_params_check", "mlmq_create", "mlmq_destroy", "mlmq_last_error", "mlmq_version",
    "mlmq_load_csr", "mlmq_load_gr", "mlmq_attach_device_csr", "mlmq_graph_stats",
    "mlmq_graph_size", "mlmq_solve", "mlmq_copy_dist", "mlmq_dist_device_ptr",
    "mlmq_device_shape", "mlmq_probe_copy", "mlmq_solve_noalloc", "mlmq_last_queue_counts", "mlmq_solve_batch", "mlmq_copy_batch_dist", "sssp_init", "sssp_run_adaptive", "sssp_cp_data", "sssp_release",
]
GRAPH_SYMBOLS = [
    "mlmq_graph_generate_knn", "mlmq_graph_generate_rmat", "mlmq_graph_read_gr",
    "mlmq_graph_from_csr", "mlmq_graph_write_gr", "mlmq_graph_dims", "mlmq_graph_arrays",
    "mlmq_graph_coords", "mlmq_graph_weight_scale", "mlmq_graph_centroid_vertex", "mlmq_graph_max_degree_vertex",
    "mlmq_graph_reorder_bfs", "mlmq_graph_cross_edges", "mlmq_graph_free",
]
SERVE_SYMBOLS = [
    "mlmq_server_config_default", "mlmq_serve_graph_params_default",
    "mlmq_serve_graph_params_resolve", "mlmq_server_create",
    "mlmq_server_destroy", "mlmq_server_last_error", "mlmq_server_add_csr", "mlmq_server_add_gr",
    "mlmq_server_set_graph_params", "mlmq_server_graph_size", "mlmq_server_slot_bytes",
    "mlmq_server_set_expected", "mlmq_server_clear_expected", "mlmq_server_run",
    "mlmq_server_submit", "mlmq_server_poll", "mlmq_server_wait", "mlmq_server_drain",
    "mlmq_server_stop",
]


class MLMQError(RuntimeError):
    def __init__(self, code: int, msg: str = ""):
        self.code = code
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")


class Params(ctypes.Structure):
    _fields_ = [
        ("queue_mode", ctypes.c_int32), ("delta", ctypes.c_uint32),
        ("work_blocks", ctypes.c_int32), ("blocks_per_cu", ctypes.c_int32),
        ("buckets", ctypes.c_int32), ("bucket_window", ctypes.c_int32),
        ("l2_batch", ctypes.c_int32), ("min_read", ctypes.c_int32),
        ("bucket_capacity", ctypes.c_int64), ("timeout_s", ctypes.c_double),
        ("auto_retry", ctypes.c_int32), ("profile", ctypes.c_int32),
        ("l1_pop", ctypes.c_int32), ("l1_spill", ctypes.c_int32),
        ("l1_donate", ctypes.c_int32), ("batch_waves", ctypes.c_int32),
        ("use_graph", ctypes.c_int32), ("hub_split", ctypes.c_int32),
        ("stale_check", ctypes.c_int32), ("prefetch", ctypes.c_int32),
        ("l1_blocks", ctypes.c_int32), ("idle_backoff", ctypes.c_int32),
        ("defer_heavy", ctypes.c_int32), ("strict_window", ctypes.c_int32),
        ("light_split", ctypes.c_int32), ("csr_stream", ctypes.c_int32),
        ("l1_ring", ctypes.c_int32), ("batch_stats", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 2),
    ]


class Timing(ctypes.Structure):
    _fields_ = [
        ("solve_ms", ctypes.c_double), ("device_ms", ctypes.c_double),
        ("reset_ms", ctypes.c_double), ("query_wall_ms", ctypes.c_double),
        ("work_blocks", ctypes.c_int32), ("waves", ctypes.c_int32),
        ("bucket_capacity", ctypes.c_int64),
        ("l2_pushes", ctypes.c_uint64), ("l2_pops", ctypes.c_uint64),
        ("relaxations", ctypes.c_uint64), ("atomics", ctypes.c_uint64),
        ("improvements", ctypes.c_uint64), ("stale", ctypes.c_uint64),
        ("batches", ctypes.c_uint64), ("idle_polls", ctypes.c_uint64),
        ("retries", ctypes.c_uint64), ("phase_cycles", ctypes.c_uint64 * 8),
        ("detail_cycles", ctypes.c_uint64 * 5), ("donated", ctypes.c_uint64),
        ("hub_chunks", ctypes.c_uint64),
    ]
    DETAILS = ("owner", "edge_read", "atomic", "laps16", "unused")
    PHASES = ("l1_pop", "l2_pop", "row_fetch", "edges", "flush", "idle", "total", "credit")

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_
             if k not in ("reserved", "phase_cycles", "detail_cycles")}
        d["phase_cycles"] = dict(zip(self.PHASES, list(self.phase_cycles)))
        d["detail_cycles"] = dict(zip(self.DETAILS, list(self.detail_cycles)))
        return d


class BatchTiming(ctypes.Structure):
    _fields_ = [
        ("solve_ms", ctypes.c_double), ("device_ms", ctypes.c_double),
        ("query_ms_p50", ctypes.c_double), ("query_ms_p99", ctypes.c_double),
        ("query_ms_max", ctypes.c_double), ("query_ms_mean", ctypes.c_double),
        ("queries", ctypes.c_int32), ("work_blocks", ctypes.c_int32),
        ("waves_per_query", ctypes.c_int32), ("graph_replay", ctypes.c_int32),
        ("bucket_capacity", ctypes.c_int64),
        ("relaxations", ctypes.c_uint64), ("improvements", ctypes.c_uint64),
        ("l2_pushes", ctypes.c_uint64), ("stale", ctypes.c_uint64),
        ("batches", ctypes.c_uint64), ("idle_polls", ctypes.c_uint64),
        ("retries", ctypes.c_uint64), ("reserved", ctypes.c_uint64 * 8),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class QueueCounts(ctypes.Structure):
    _fields_ = [
        ("buckets", ctypes.c_int32), ("window", ctypes.c_int32), ("capacity", ctypes.c_int64),
        ("writes", ctypes.c_uint32 * 64), ("reads", ctypes.c_uint32 * 64),
        ("completed", ctypes.c_uint32), ("epoch", ctypes.c_uint32), ("wrapped", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32 * 5),
    ]

    def as_dict(self) -> dict:
        nb = self.buckets
        return {"buckets": nb, "window": self.window, "capacity": self.capacity,
                "writes": list(self.writes)[:nb], "reads": list(self.reads)[:nb],
                "completed": self.completed, "epoch": self.epoch, "wrapped": self.wrapped}


class GraphInfo(ctypes.Structure):
    """Mirror of the reference's graph_info (core/include/graph_info.h:8-20)."""
    _fields_ = [
        ("m", ctypes.c_int32), ("nnz", ctypes.c_int32),
        ("avg_nnz", ctypes.c_float), ("max_nnz", ctypes.c_float), ("dev_nnz", ctypes.c_float),
        ("avg_weight", ctypes.c_float), ("dev_weight", ctypes.c_float),
        ("max_weight", ctypes.c_float), ("diameter", ctypes.c_float),
    ]


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("hbm_budget", ctypes.c_uint64), ("streams", ctypes.c_int32),
        ("max_batch", ctypes.c_int32), ("max_pending", ctypes.c_int32),
        ("verify", ctypes.c_int32), ("timeout_s", ctypes.c_double),
        ("spread_waves", ctypes.c_int32), ("mode", ctypes.c_int32),
        ("live_waves", ctypes.c_int32), ("live_idle_ms", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 4),
    ]


class ServeGraphParams(ctypes.Structure):
    _fields_ = [
        ("delta", ctypes.c_uint32), ("buckets", ctypes.c_int32),
        ("bucket_window", ctypes.c_int32), ("l2_batch", ctypes.c_int32),
        ("min_read", ctypes.c_int32), ("l1_pop", ctypes.c_int32),
        ("l1_spill", ctypes.c_int32), ("workgroups", ctypes.c_int32),
        ("waves", ctypes.c_int32), ("bucket_capacity", ctypes.c_int64),
        ("hub_split", ctypes.c_int32), ("light_split", ctypes.c_int32),
        ("prefetch", ctypes.c_int32), ("stale_check", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 4),
    ]


class ServeRequest(ctypes.Structure):
    _fields_ = [
        ("arrival_s", ctypes.c_double), ("graph", ctypes.c_int32), ("src", ctypes.c_int32),
        ("dist_out", ctypes.c_void_p),
    ]


class ServeResult(ctypes.Structure):
    _fields_ = [
        ("latency_ms", ctypes.c_double), ("queue_ms", ctypes.c_double),
        ("device_ms", ctypes.c_double), ("status", ctypes.c_int32),
        ("mismatches", ctypes.c_int32), ("launch", ctypes.c_int32),
        ("batch_size", ctypes.c_int32), ("retries", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class ServeStats(ctypes.Structure):
    _fields_ = [
        ("wall_s", ctypes.c_double), ("offered_qps", ctypes.c_double),
        ("completed_qps", ctypes.c_double), ("p50_ms", ctypes.c_double),
        ("p99_ms", ctypes.c_double), ("max_ms", ctypes.c_double), ("mean_ms", ctypes.c_double),
        ("queue_p50_ms", ctypes.c_double), ("queue_p99_ms", ctypes.c_double),
        ("mean_batch", ctypes.c_double), ("mean_device_ms", ctypes.c_double),
        ("requests", ctypes.c_uint64), ("completed", ctypes.c_uint64),
        ("rejected", ctypes.c_uint64), ("failed", ctypes.c_uint64),
        ("verified", ctypes.c_uint64), ("mismatched", ctypes.c_uint64),
        ("retries", ctypes.c_uint64), ("launches", ctypes.c_uint64),
        ("peak_bytes", ctypes.c_uint64), ("peak_inflight", ctypes.c_uint64),
        ("reserved", ctypes.c_int64 * 8),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


_lib = None


def load() -> ctypes.CDLL:
    """Load libmlmq.so; raise if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libmlmq.so not built: {LIB_PATH} "
                           "(run __graft_entry__.build() or make -C mlmq_tpds_withl3_amd/csrc)")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    i32, i64, u64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64
    sig = {
        "mlmq_params_default": (None, [ctypes.POINTER(Params)]),
        "mlmq_params_check": (ctypes.c_int, [ctypes.POINTER(Params), i32]),
        "mlmq_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(P)]),
        "mlmq_destroy": (ctypes.c_int, [P]),
        "mlmq_last_error": (ctypes.c_char_p, [P]),
        "mlmq_version": (ctypes.c_char_p, []),
        "mlmq_load_csr": (ctypes.c_int, [P, i32, i64, P, P, P]),
        "mlmq_load_gr": (ctypes.c_int, [P, ctypes.c_char_p]),
        "mlmq_attach_device_csr": (ctypes.c_int, [P, i32, i64, P, P, P]),
        "mlmq_graph_stats": (ctypes.c_int, [P, ctypes.POINTER(GraphInfo)]),
        "mlmq_graph_size": (ctypes.c_int, [P, ctypes.POINTER(i32), ctypes.POINTER(i64)]),
        "mlmq_solve": (ctypes.c_int, [P, i32, ctypes.POINTER(Params), ctypes.POINTER(Timing)]),
        "mlmq_copy_dist": (ctypes.c_int, [P, P]),
        "mlmq_dist_device_ptr": (ctypes.c_int, [P, ctypes.POINTER(P)]),
        "mlmq_device_shape": (ctypes.c_int, [P, ctypes.POINTER(i32), ctypes.POINTER(i32),
                                             ctypes.POINTER(i32)]),
        "mlmq_probe_copy": (ctypes.c_int, [P, u64, ctypes.POINTER(ctypes.c_double)]),
        "mlmq_solve_noalloc": (ctypes.c_int, [P, i32, ctypes.POINTER(Params), P,
                                              ctypes.POINTER(Timing)]),
        "mlmq_last_queue_counts": (ctypes.c_int, [P, ctypes.POINTER(QueueCounts)]),
        "mlmq_solve_batch": (ctypes.c_int, [P, P, i32, ctypes.POINTER(Params),
                                            ctypes.POINTER(BatchTiming)]),
        "mlmq_copy_batch_dist": (ctypes.c_int, [P, i32, P]),
        "sssp_init": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, P, P, P]),
        "sssp_run_adaptive": (ctypes.c_int, [ctypes.c_int, GraphInfo]),
        "sssp_cp_data": (ctypes.c_int, [ctypes.POINTER(ctypes.POINTER(ctypes.c_int))]),
        "sssp_release": (ctypes.c_int, []),
        "mlmq_graph_generate_knn": (ctypes.c_int, [i32, i32, u64, ctypes.c_double, i32,
                                                   ctypes.POINTER(P)]),
        "mlmq_graph_generate_rmat": (ctypes.c_int, [i32, i32, ctypes.c_double, ctypes.c_double,
                                                    ctypes.c_double, u64, i32, i32,
                                                    ctypes.POINTER(P)]),
        "mlmq_graph_read_gr": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(P)]),
        "mlmq_graph_from_csr": (ctypes.c_int, [i32, i64, P, P, P, ctypes.POINTER(P)]),
        "mlmq_graph_write_gr": (ctypes.c_int, [P, ctypes.c_char_p]),
        "mlmq_graph_dims": (ctypes.c_int, [P, ctypes.POINTER(i32), ctypes.POINTER(i64)]),
        "mlmq_graph_arrays": (ctypes.c_int, [P, ctypes.POINTER(P), ctypes.POINTER(P),
                                             ctypes.POINTER(P)]),
        "mlmq_graph_coords": (ctypes.c_int, [P, ctypes.POINTER(P)]),
        "mlmq_graph_weight_scale": (ctypes.c_double, [P]),
        "mlmq_graph_centroid_vertex": (i32, [P]),
        "mlmq_graph_max_degree_vertex": (i32, [P]),
        "mlmq_graph_reorder_bfs": (ctypes.c_int, [P, i32, i32, P, P, ctypes.POINTER(P)]),
        "mlmq_graph_cross_edges": (i64, [P, i32]),
        "mlmq_graph_free": (None, [P]),
        "mlmq_server_config_default": (None, [ctypes.POINTER(ServerConfig)]),
        "mlmq_serve_graph_params_default": (None, [ctypes.POINTER(ServeGraphParams)]),
        "mlmq_serve_graph_params_resolve": (ctypes.c_int, [ctypes.POINTER(ServeGraphParams), i32, i64,
                                                           ctypes.POINTER(ServeGraphParams)]),
        "mlmq_server_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ServerConfig),
                                              ctypes.POINTER(P)]),
        "mlmq_server_destroy": (ctypes.c_int, [P]),
        "mlmq_server_last_error": (ctypes.c_char_p, [P]),
        "mlmq_server_add_csr": (ctypes.c_int, [P, i32, i64, P, P, P,
                                               ctypes.POINTER(ServeGraphParams),
                                               ctypes.POINTER(i32)]),
        "mlmq_server_add_gr": (ctypes.c_int, [P, ctypes.c_char_p, ctypes.POINTER(ServeGraphParams),
                                              ctypes.POINTER(i32)]),
        "mlmq_server_set_graph_params": (ctypes.c_int, [P, i32, ctypes.POINTER(ServeGraphParams)]),
        "mlmq_server_graph_size": (ctypes.c_int, [P, i32, ctypes.POINTER(i32), ctypes.POINTER(i64)]),
        "mlmq_server_slot_bytes": (ctypes.c_int, [P, i32, ctypes.POINTER(u64)]),
        "mlmq_server_set_expected": (ctypes.c_int, [P, i32, i32, P]),
        "mlmq_server_clear_expected": (ctypes.c_int, [P]),
        "mlmq_server_run": (ctypes.c_int, [P, ctypes.POINTER(ServeRequest), i32,
                                           ctypes.POINTER(ServeResult), ctypes.POINTER(ServeStats)]),
        "mlmq_server_submit": (ctypes.c_int, [P, i32, i32, P, ctypes.POINTER(i64)]),
        "mlmq_server_poll": (ctypes.c_int, [P, i64, ctypes.POINTER(ServeResult)]),
        "mlmq_server_wait": (ctypes.c_int, [P, i64, ctypes.c_double, ctypes.POINTER(ServeResult)]),
        "mlmq_server_drain": (ctypes.c_int, [P, ctypes.c_double]),
        "mlmq_server_stop": (ctypes.c_int, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc: int, ctx=None) -> None:
    if rc != MLMQ_OK:
        msg = ""
        if ctx is not None:
            raw = load().mlmq_last_error(ctx)
            msg = raw.decode() if raw else ""
        raise MLMQError(rc, msg)
