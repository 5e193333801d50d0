"""Host graphs: .gr files and the seeded synthetic inputs (include/mlmq_graph.h).

``HostGraph`` owns a CSR built by libmlmq.so: read from a Galois .gr v1 file
(the reference's on-disk format, SSSP/csr_graph.cu:185-296), from numpy
arrays, or generated:

* ``knn(n, k, seed, weight_scale, morton)`` -- n seeded points in the unit
  square, each linked to its k nearest neighbours, symmetrised; the
  road-like inputs.  ``weight_scale > 0`` multiplies Euclidean length,
  ``weight_scale < 0`` sets the longest edge's weight to ``-weight_scale``.
* ``rmat(scale, edge_factor, a, b, c, seed, wmin, wmax)`` -- symmetrised RMAT.

``reorder_bfs(root, layer_split)`` renumbers vertices in breadth-first order
(graph_transform/reorder_gr.cu); the ``reorder_gr`` command beside the
library is the same step with the reference tool's command line.

The named workloads of the benchmark are in ``WORKLOADS``; ``bench_point`` is
the solver setting a single-query benchmark times on a workload, chosen by
``autotune.tune_single`` on other sources than the workload's own.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check


class HostGraph:
    def __init__(self, handle: ctypes.c_void_p, kind: str, meta: dict | None = None):
        self._L = _lib.load()
        self._h = handle
        self.kind = kind
        self.meta = dict(meta or {})
        n = ctypes.c_int32()
        m = ctypes.c_int64()
        check(self._L.mlmq_graph_dims(self._h, ctypes.byref(n), ctypes.byref(m)))
        self.n, self.m = n.value, m.value

    # -- constructors --
    @classmethod
    def knn(cls, n: int, k: int = 3, seed: int = 1, weight_scale: float = 1e5,
            morton: bool = False) -> "HostGraph":
        h = ctypes.c_void_p()
        check(_lib.load().mlmq_graph_generate_knn(int(n), int(k), int(seed), float(weight_scale),
                                                  int(bool(morton)), ctypes.byref(h)))
        g = cls(h, "knn", dict(n=n, k=k, seed=seed, weight_scale_arg=weight_scale,
                               morton=bool(morton)))
        g.meta["weight_scale"] = g.weight_scale()
        return g

    @classmethod
    def rmat(cls, scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19,
             c: float = 0.19, seed: int = 1, wmin: int = 1, wmax: int = 1000) -> "HostGraph":
        h = ctypes.c_void_p()
        check(_lib.load().mlmq_graph_generate_rmat(int(scale), int(edge_factor), a, b, c,
                                                   int(seed), int(wmin), int(wmax),
                                                   ctypes.byref(h)))
        return cls(h, "rmat", dict(scale=scale, edge_factor=edge_factor, a=a, b=b, c=c,
                                   seed=seed, wmin=wmin, wmax=wmax))

    @classmethod
    def read_gr(cls, path: str) -> "HostGraph":
        h = ctypes.c_void_p()
        check(_lib.load().mlmq_graph_read_gr(path.encode(), ctypes.byref(h)))
        return cls(h, "gr", dict(path=path))

    @classmethod
    def from_csr(cls, row, col, w) -> "HostGraph":
        row = np.ascontiguousarray(row, dtype=np.int64)
        col = np.ascontiguousarray(col, dtype=np.int32)
        w = np.ascontiguousarray(w, dtype=np.int32)
        h = ctypes.c_void_p()
        check(_lib.load().mlmq_graph_from_csr(len(row) - 1, int(row[-1]), row.ctypes.data,
                                              col.ctypes.data, w.ctypes.data, ctypes.byref(h)))
        return cls(h, "csr")

    # -- accessors --
    def arrays(self):
        """Zero-copy numpy views (row int64[n+1], col int32[m], w int32[m]); valid while self lives."""
        r, c, w = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        check(self._L.mlmq_graph_arrays(self._h, ctypes.byref(r), ctypes.byref(c), ctypes.byref(w)))
        row = np.ctypeslib.as_array(ctypes.cast(r, ctypes.POINTER(ctypes.c_int64)), (self.n + 1,))
        if self.m:
            col = np.ctypeslib.as_array(ctypes.cast(c, ctypes.POINTER(ctypes.c_int32)), (self.m,))
            wt = np.ctypeslib.as_array(ctypes.cast(w, ctypes.POINTER(ctypes.c_int32)), (self.m,))
        else:
            col = np.empty(0, np.int32)
            wt = np.empty(0, np.int32)
        return row, col, wt

    def coords(self) -> np.ndarray:
        """k-NN point coordinates, float64[n, 2] (copy)."""
        p = ctypes.c_void_p()
        check(self._L.mlmq_graph_coords(self._h, ctypes.byref(p)))
        flat = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_double)), (2 * self.n,))
        return flat.reshape(self.n, 2).copy()

    def write_gr(self, path: str) -> None:
        check(self._L.mlmq_graph_write_gr(self._h, path.encode()))

    def weight_scale(self) -> float:
        return float(self._L.mlmq_graph_weight_scale(self._h))

    def centroid_vertex(self) -> int:
        return int(self._L.mlmq_graph_centroid_vertex(self._h))

    def max_degree_vertex(self) -> int:
        return int(self._L.mlmq_graph_max_degree_vertex(self._h))

    def reorder_bfs(self, root: int = 0, layer_split: bool = False):
        """BFS renumbering from ``root``: returns (graph, perm, invperm) with
        perm[new] = old and invperm[old] = new (int32[n] each)."""
        perm = np.empty(self.n, dtype=np.int32)
        inv = np.empty(self.n, dtype=np.int32)
        h = ctypes.c_void_p()
        check(self._L.mlmq_graph_reorder_bfs(self._h, int(root), int(bool(layer_split)),
                                             perm.ctypes.data, inv.ctypes.data, ctypes.byref(h)))
        g = HostGraph(h, self.kind, dict(self.meta, reorder="bfs_layer_split" if layer_split
                                         else "bfs", reorder_root=int(root)))
        return g, perm, inv

    def cross_edges(self, cut: int | None = None) -> int:
        """Directed edges crossing vertex id ``cut`` (default n // 2)."""
        c = int(self._L.mlmq_graph_cross_edges(self._h, self.n // 2 if cut is None else int(cut)))
        if c < 0:
            raise ValueError("cut outside [0, n]")
        return c

    def close(self) -> None:
        if self._h:
            self._L.mlmq_graph_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# Named workloads (sizes, seeds and sources of the benchmark inputs).
WORKLOADS = {
    # 264,346 points, 3-NN, weight = max(1, round(len * 1e5)), source 0, delta 200000
    "road264k": dict(kind="knn", n=264_346, k=3, seed=1, weight_scale=1e5, morton=False,
                     source="zero", delta=200_000, delta_sweep=(200_000,),
                     bench_point=dict(delta=200_000, buckets=2, bucket_window=2, l2_batch=4,
                                      min_read=4, l1_pop=32, work_blocks=192, stale_check=3),
                     # seed of the serving tuner: a launch holds many requests, where the
                     # single query's claims of 4 cost latency (p50 at 1000 queries/s
                     # 14.6 -> 19.7 ms untuned)
                     serve_point=dict(delta=200_000, buckets=4, bucket_window=4, l2_batch=8,
                                      min_read=8, l1_pop=32)),
    # 6,262,104 points, 3-NN, Morton order, longest edge -> 100,000, centroid source
    "road6m": dict(kind="knn", n=6_262_104, k=3, seed=2, weight_scale=-100_000.0, morton=True,
                   source="centroid", delta=200_000,
                   delta_sweep=(50_000, 100_000, 200_000, 400_000),
                   bench_point=dict(delta=100_000, buckets=2, bucket_window=2, l2_batch=16,
                                    min_read=16, l1_pop=32, work_blocks=256, stale_check=3)),
    # RMAT scale 22, edge factor 16, (0.57, 0.19, 0.19, 0.05), weights [1, 1000]
    "rmat22": dict(kind="rmat", scale=22, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=3,
                   wmin=1, wmax=1000, source="maxdeg", delta=200_000,
                   delta_sweep=(50, 100, 200, 1_000, 10_000),
                   tune_grid=[dict(light_split=ls, work_blocks=wb, buckets=64 if ls else 2,
                                   bucket_window=2, l2_batch=b, min_read=b, l1_pop=b)
                              for ls in (1, 0) for wb in (256, 512) for b in (32, 64)],
                   # child-block touches off: on this graph they cost 7 % of the
                   # memory traffic and no time (medians of 28: 10.97 ms off, 11.10 on)
                   bench_point=dict(delta=50, light_split=1, buckets=64, bucket_window=2,
                                    l2_batch=64, min_read=64, l1_pop=64, work_blocks=512,
                                    stale_check=1, prefetch=0)),
    # RMAT scale 20 (1,048,576 vertices), edge factor 16, weights [1, 1000] (serving mix)
    "rmat20": dict(kind="rmat", scale=20, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=5,
                   wmin=1, wmax=1000, source="maxdeg", delta=200_000,
                   delta_sweep=(50, 100, 200, 1_000, 10_000),
                   tune_grid=[dict(light_split=ls, work_blocks=wb, buckets=64 if ls else 2,
                                   bucket_window=2, l2_batch=b, min_read=b, l1_pop=b)
                              for ls in (1, 0) for wb in (256, 512) for b in (32, 64)]),
    # 1,070,376 points, 3-NN, longest edge -> 50,000 (batched-query graph)
    "road1m": dict(kind="knn", n=1_070_376, k=3, seed=4, weight_scale=-50_000.0, morton=True,
                   source="centroid", delta=200_000, delta_sweep=(100_000, 200_000, 400_000)),
}


def make_workload(name: str, **override) -> tuple[HostGraph, int, dict]:
    """Build a named workload: returns (graph, source, spec)."""
    spec = dict(WORKLOADS[name])
    spec.update(override)
    if spec["kind"] == "knn":
        g = HostGraph.knn(spec["n"], spec["k"], spec["seed"], spec["weight_scale"], spec["morton"])
    else:
        g = HostGraph.rmat(spec["scale"], spec["edge_factor"], spec["a"], spec["b"], spec["c"],
                           spec["seed"], spec["wmin"], spec["wmax"])
    src = {"zero": 0, "centroid": g.centroid_vertex() if spec["kind"] == "knn" else 0,
           "maxdeg": g.max_degree_vertex()}[spec["source"]]
    spec["source_vertex"] = src
    spec["weight_scale_applied"] = g.weight_scale()
    return g, src, spec
