"""Host-side solver interface over the C ABI (include/mlmq.h).

Two faces, both calling libmlmq.so and nothing else:

* ``Solver`` -- one context (device, resident graph, per-query workspace):
  ``load_gr`` / ``load_csr`` / ``attach_device_csr``, ``solve(src, **params)``,
  ``dist()``.  It replaces the reference's file-scope solver state and
  ``sssp_init`` / ``sssp_run_adaptive`` / ``sssp_cp_data``
  (evidence/l3_supplement_20260921_v2/single_source/SSSP/sssp_run.cu:676-976).
* ``sssp_init`` / ``sssp_run_adaptive`` / ``sssp_cp_data`` -- the reference's
  driver functions (same file, declared in .../SSSP/main.cu:15-17) with the same
  argument meaning: a device CSR in the reference layout, a source vertex, and
  a copy-out of the m distances.  Errors raise ``MLMQError`` instead of exit(2).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import BatchTiming, GraphInfo, MLMQError, Params, QueueCounts, Timing, check  # noqa: F401

UNREACHED = 0x7FFFFFFF


def default_params(**overrides) -> Params:
    p = Params()
    _lib.load().mlmq_params_default(ctypes.byref(p))
    for k, v in overrides.items():
        if k == "queue":
            k, v = "queue_mode", _lib.QUEUE_MODES[v]
        if not hasattr(p, k):
            raise TypeError(f"unknown solver parameter {k!r}")
        setattr(p, k, v)
    return p


class Solver:
    """A solver context on one GPU (MLMQ L1SLF_L2DQ persistent solver)."""

    def __init__(self, device: int = 0):
        self._L = _lib.load()
        h = ctypes.c_void_p()
        check(self._L.mlmq_create(int(device), ctypes.byref(h)))
        self._h = h
        self.n = 0
        self.m = 0

    # -- lifetime --
    def close(self) -> None:
        if self._h:
            self._L.mlmq_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- graph --
    def _sync_size(self):
        n = ctypes.c_int32()
        m = ctypes.c_int64()
        self._L.mlmq_graph_size(self._h, ctypes.byref(n), ctypes.byref(m))
        self.n, self.m = n.value, m.value

    def load_gr(self, path: str) -> None:
        check(self._L.mlmq_load_gr(self._h, path.encode()), self._h)
        self._sync_size()

    def load_csr(self, row, col, w) -> None:
        row = np.ascontiguousarray(row, dtype=np.int64)
        col = np.ascontiguousarray(col, dtype=np.int32)
        w = np.ascontiguousarray(w, dtype=np.int32)
        n = len(row) - 1
        check(self._L.mlmq_load_csr(self._h, n, int(row[-1]), row.ctypes.data, col.ctypes.data,
                                    w.ctypes.data), self._h)
        self._sync_size()

    def load_graph(self, g) -> None:
        """Upload a ``graphs.HostGraph``."""
        row, col, w = g.arrays()
        self.load_csr(row, col, w)

    def attach_device_csr(self, n: int, m: int, row_ptr: int, col_ptr: int, w_ptr: int) -> None:
        check(self._L.mlmq_attach_device_csr(self._h, int(n), int(m), ctypes.c_void_p(row_ptr),
                                             ctypes.c_void_p(col_ptr), ctypes.c_void_p(w_ptr)),
              self._h)
        self._sync_size()

    def graph_info(self) -> GraphInfo:
        gi = GraphInfo()
        check(self._L.mlmq_graph_stats(self._h, ctypes.byref(gi)), self._h)
        return gi

    def device_shape(self) -> dict:
        b, c, t = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        check(self._L.mlmq_device_shape(self._h, ctypes.byref(b), ctypes.byref(c), ctypes.byref(t)),
              self._h)
        return {"blocks_per_cu_max": b.value, "cus": c.value, "threads_per_block": t.value}

    def probe_copy(self, nbytes: int) -> float:
        """Device ms of one plain copy kernel over `nbytes` (counter calibration)."""
        ms = ctypes.c_double()
        check(self._L.mlmq_probe_copy(self._h, int(nbytes), ctypes.byref(ms)), self._h)
        return ms.value

    # -- queries --
    def solve(self, src: int, params: Params | None = None, **overrides) -> dict:
        p = params if params is not None else default_params(**overrides)
        t = Timing()
        check(self._L.mlmq_solve(self._h, int(src), ctypes.byref(p), ctypes.byref(t)), self._h)
        return t.as_dict()

    def solve_batch(self, sources, params: Params | None = None, **overrides) -> dict:
        """All `sources` in one persistent launch; distances stay on the device."""
        src = np.ascontiguousarray(sources, dtype=np.int32)
        p = params if params is not None else default_params(**overrides)
        t = BatchTiming()
        check(self._L.mlmq_solve_batch(self._h, src.ctypes.data, len(src), ctypes.byref(p),
                                       ctypes.byref(t)), self._h)
        self._batch_n = len(src)
        return t.as_dict()

    def batch_dist(self, query: int) -> np.ndarray:
        out = np.empty(self.n, dtype=np.int32)
        check(self._L.mlmq_copy_batch_dist(self._h, int(query), out.ctypes.data), self._h)
        return out

    def dist(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.int32)
        check(self._L.mlmq_copy_dist(self._h, out.ctypes.data), self._h)
        return out

    def queue_counts(self) -> dict:
        """Per-bucket reservations and reads of the last solve, the credits
        returned, the window advances and the ring capacity."""
        q = QueueCounts()
        check(self._L.mlmq_last_queue_counts(self._h, ctypes.byref(q)), self._h)
        return q.as_dict()

    def dist_device_ptr(self) -> int:
        p = ctypes.c_void_p()
        check(self._L.mlmq_dist_device_ptr(self._h, ctypes.byref(p)), self._h)
        return p.value


# ---- the reference driver ABI (one implicit context on device 0) ----

_ref_vertices = {"n": 0}


def sssp_init(m: int, nnz: int, row_ptr_d: int, col_idx_d: int, w_d: int) -> int:
    """sssp_init(m, nnz, RowPtr_d, ColIdx_d, w_d): m vertices, nnz edges, device
    pointers to an int32 CSR in the reference layout (row[0] = 0, row[m] = nnz)."""
    L = _lib.load()
    check(L.sssp_init(int(m), int(nnz), ctypes.c_void_p(row_ptr_d), ctypes.c_void_p(col_idx_d),
                      ctypes.c_void_p(w_d)))
    _ref_vertices["n"] = int(m)
    return 0


def sssp_run_adaptive(src: int, info: GraphInfo | None = None) -> int:
    """Run one query; BENCH_DELTA / MLMQ_WORK_BLOCKS are read from the environment."""
    L = _lib.load()
    check(L.sssp_run_adaptive(int(src), info if info is not None else GraphInfo()))
    return 0


def sssp_cp_data() -> np.ndarray:
    """Copy out the m distances of the last query.  The C call malloc()s the
    array (as the reference does); it is copied into numpy and freed here."""
    L = _lib.load()
    ptr = ctypes.POINTER(ctypes.c_int)()
    check(L.sssp_cp_data(ctypes.byref(ptr)))
    n = _ref_vertices["n"]
    try:
        return np.ctypeslib.as_array(ptr, shape=(n,)).astype(np.int32, copy=True)
    finally:
        ctypes.CDLL(None).free(ctypes.cast(ptr, ctypes.c_void_p))


def sssp_release() -> None:
    _lib.load().sssp_release()
    _ref_vertices["n"] = 0
