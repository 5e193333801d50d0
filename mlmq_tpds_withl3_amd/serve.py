"""Python face of the serving loop (include/mlmq_serve.h).

``Server`` keeps several graphs resident, admits requests by the HBM bytes of
their slot (distance row + queue partition) and either replays an arrival
trace in real time inside ``libmlmq.so`` (``run``; ``poisson_trace`` builds
such a trace) or serves requests as they are submitted (``submit`` /
``poll`` / ``wait``: one persistent launch fed through a request ring).  No
serving logic lives here: the admission, batching and launch loops are the
C++ code of ``mlmq_server_run`` and the live service thread.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import (MLMQ_E_PENDING, MLMQError, SERVE_MODES, ServeGraphParams, ServeRequest,
                   ServeResult, ServeStats, ServerConfig, load)

STATUS = {0: "ok", -4: "nomem", -6: "overflow", -7: "timeout", -8: "internal", -10: "rejected"}
RESULT_FIELDS = ("latency_ms", "queue_ms", "device_ms", "status", "mismatches", "launch", "batch_size",
                 "retries")


def _check(rc: int, h) -> None:
    if rc != 0:
        raw = load().mlmq_server_last_error(h) if h else b""
        raise MLMQError(rc, raw.decode() if raw else "")


def graph_params(**overrides) -> ServeGraphParams:
    p = ServeGraphParams()
    load().mlmq_serve_graph_params_default(ctypes.byref(p))
    for k, v in overrides.items():
        if not hasattr(p, k) or k == "reserved":
            raise KeyError(f"unknown serving parameter {k}")
        setattr(p, k, v)
    return p


class Server:
    """Resident graphs plus the request loop of one GPU."""

    def __init__(self, device: int = 0, **config):
        self._L = load()
        cfg = ServerConfig()
        self._L.mlmq_server_config_default(ctypes.byref(cfg))
        for k, v in config.items():
            if not hasattr(cfg, k) or k == "reserved":
                raise KeyError(f"unknown server setting {k}")
            if k == "mode" and isinstance(v, str):
                v = SERVE_MODES[v]
            setattr(cfg, k, v)
        self.config = {k: getattr(cfg, k) for k, _ in cfg._fields_ if k != "reserved"}
        h = ctypes.c_void_p()
        rc = self._L.mlmq_server_create(int(device), ctypes.byref(cfg), ctypes.byref(h))
        if rc != 0:
            raise MLMQError(rc, "mlmq_server_create")
        self._h = h
        self.sizes: list[tuple[int, int]] = []

    # -- graphs -----------------------------------------------------------
    def add_graph(self, graph, **params) -> int:
        """Make a graph resident: a ``HostGraph``, a .gr path, or a CSR tuple
        (row int64[n+1], col int32[m], w int32[m]).  Returns its id."""
        p = graph_params(**params)
        gid = ctypes.c_int32()
        if isinstance(graph, str):
            _check(self._L.mlmq_server_add_gr(self._h, graph.encode(), ctypes.byref(p),
                                              ctypes.byref(gid)), self._h)
        else:
            row, col, w = graph if isinstance(graph, tuple) else graph.arrays()
            row = np.ascontiguousarray(row, dtype=np.int64)
            col = np.ascontiguousarray(col, dtype=np.int32)
            w = np.ascontiguousarray(w, dtype=np.int32)
            _check(self._L.mlmq_server_add_csr(self._h, len(row) - 1, int(row[-1]), row.ctypes.data,
                                               col.ctypes.data, w.ctypes.data, ctypes.byref(p),
                                               ctypes.byref(gid)), self._h)
        n, m = ctypes.c_int32(), ctypes.c_int64()
        _check(self._L.mlmq_server_graph_size(self._h, gid.value, ctypes.byref(n), ctypes.byref(m)),
               self._h)
        self.sizes.append((n.value, m.value))
        return gid.value

    def set_graph_params(self, gid: int, **params) -> None:
        p = graph_params(**params)
        _check(self._L.mlmq_server_set_graph_params(self._h, int(gid), ctypes.byref(p)), self._h)

    def slot_bytes(self, gid: int) -> int:
        b = ctypes.c_uint64()
        _check(self._L.mlmq_server_slot_bytes(self._h, int(gid), ctypes.byref(b)), self._h)
        return b.value

    def set_expected(self, gid: int, src: int, dist) -> None:
        d = np.ascontiguousarray(dist, dtype=np.int32)
        if len(d) != self.sizes[gid][0]:
            raise ValueError("expected row length differs from the graph's vertex count")
        _check(self._L.mlmq_server_set_expected(self._h, int(gid), int(src), d.ctypes.data), self._h)

    def clear_expected(self) -> None:
        _check(self._L.mlmq_server_clear_expected(self._h), self._h)

    # -- requests ---------------------------------------------------------
    def run(self, arrival_s, graph, src, dist_out=None):
        """Replay a trace.  Returns (per-request numpy record array, stats dict).

        ``dist_out`` may be a list of int32 arrays (or None entries) that
        receive the result rows."""
        arrival_s = np.asarray(arrival_s, dtype=np.float64)
        graph = np.asarray(graph, dtype=np.int32)
        src = np.asarray(src, dtype=np.int32)
        n = len(arrival_s)
        if not (len(graph) == n and len(src) == n):
            raise ValueError("trace columns differ in length")
        reqs = (ServeRequest * max(n, 1))()
        for i in range(n):
            reqs[i].arrival_s = float(arrival_s[i])
            reqs[i].graph = int(graph[i])
            reqs[i].src = int(src[i])
            out = None if dist_out is None else dist_out[i]
            if out is not None:
                if out.dtype != np.int32 or not out.flags.c_contiguous or len(out) != self.sizes[graph[i]][0]:
                    raise ValueError("dist_out rows must be contiguous int32 of the graph's size")
                reqs[i].dist_out = out.ctypes.data
            else:
                reqs[i].dist_out = None
        res = (ServeResult * max(n, 1))()
        st = ServeStats()
        _check(self._L.mlmq_server_run(self._h, reqs, n, res, ctypes.byref(st)), self._h)
        dt = np.dtype([("latency_ms", "f8"), ("queue_ms", "f8"), ("device_ms", "f8"),
                       ("status", "i4"), ("mismatches", "i4"), ("launch", "i4"),
                       ("batch_size", "i4"), ("retries", "i4"), ("reserved", "i4")])
        rec = np.frombuffer(bytes(res), dtype=dt, count=n).copy() if n else np.zeros(0, dt)
        return rec, st.as_dict()

    # -- live requests ----------------------------------------------------
    def submit(self, graph: int, src: int, dist_out=None) -> int:
        """Submit one request now; returns its id.  `dist_out` (int32, the
        graph's vertex count) receives the row; keep it alive until the
        request is collected."""
        if dist_out is not None and (dist_out.dtype != np.int32 or not dist_out.flags.c_contiguous
                                     or len(dist_out) != self.sizes[graph][0]):
            raise ValueError("dist_out must be contiguous int32 of the graph's size")
        rid = ctypes.c_int64()
        _check(self._L.mlmq_server_submit(self._h, int(graph), int(src),
                                          None if dist_out is None else dist_out.ctypes.data,
                                          ctypes.byref(rid)), self._h)
        return rid.value

    @staticmethod
    def _result(r: ServeResult) -> dict:
        return {k: getattr(r, k) for k in RESULT_FIELDS}

    def poll(self, rid: int):
        """The request's result (dict) once done, else None."""
        r = ServeResult()
        rc = self._L.mlmq_server_poll(self._h, int(rid), ctypes.byref(r))
        if rc == MLMQ_E_PENDING:
            return None
        _check(rc, self._h)
        return self._result(r)

    def wait(self, rid: int, timeout_s: float = 0.0):
        """Waits for the request (timeout_s <= 0: no limit); None on timeout."""
        r = ServeResult()
        rc = self._L.mlmq_server_wait(self._h, int(rid), float(timeout_s), ctypes.byref(r))
        if rc == MLMQ_E_PENDING:
            return None
        _check(rc, self._h)
        return self._result(r)

    def drain(self, timeout_s: float = 0.0) -> bool:
        rc = self._L.mlmq_server_drain(self._h, float(timeout_s))
        if rc == MLMQ_E_PENDING:
            return False
        _check(rc, self._h)
        return True

    def stop(self) -> None:
        _check(self._L.mlmq_server_stop(self._h), self._h)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.mlmq_server_destroy(self._h)
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


def poisson_trace(rate_qps: float, duration_s: float, graph_weights, source_pools, seed: int = 0):
    """Poisson arrivals at `rate_qps` for `duration_s`; each request picks a
    graph with the given weights and a source uniformly from that graph's pool.
    Returns (arrival_s, graph, src) arrays."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_qps, size=int(rate_qps * duration_s * 1.5) + 16)
    t = np.cumsum(gaps)
    t = t[t < duration_s]
    w = np.asarray(graph_weights, dtype=np.float64)
    g = rng.choice(len(w), size=len(t), p=w / w.sum()).astype(np.int32)
    s = np.empty(len(t), dtype=np.int32)
    for gi, pool in enumerate(source_pools):
        idx = np.nonzero(g == gi)[0]
        s[idx] = np.asarray(pool, dtype=np.int32)[rng.integers(0, len(pool), size=len(idx))]
    return t, g, s
