"""Python face of the CPU oracle (test infrastructure only).

Only the tests, the smoke entry point and the benchmark's CPU-baseline leg
import this module; the product never does.  It wraps ``liboracle.so`` (built
from ``dijkstra.c`` by ``Makefile`` in this directory) and, when it was built,
the reference's own int64 oracle ``_ref/export_graph`` (compiled from
scripts/multigpu/supplement_graph.cpp of the reference).

``read_gr``/``write_gr`` are an independent numpy statement of the Galois .gr
v1 layout read by SSSP/csr_graph.cu:185-296 and written by
SSSP/csr_graph.cu:302-418 (header u64 x4 {1, 4, n, m}; n u64 row *end*
offsets; m u32 destinations; 4 pad bytes when m is odd; m int32 weights).

``bfs_reorder``/``relabel``/``cross_edges`` restate the vertex renumbering of
graph_transform/reorder_gr.cu (pure Python loops: small graphs only).  The
reference tool needs nvcc and boost to build, so it is not compiled here;
these functions are pinned by hand-worked orders and by the invariants of a
breadth-first order in tests/test_reorder.py.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_EXPORT = os.path.join(HERE, "_ref", "export_graph")
UNREACHED = 0x7FFFFFFF

_lib = None


def build() -> None:
    """Compile liboracle.so (and _ref/export_graph where the reference exists)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        p = ctypes.c_void_p
        L.oracle_dijkstra.argtypes = [ctypes.c_int32, p, p, p, ctypes.c_int32, p, p, p]
        L.oracle_dijkstra.restype = ctypes.c_int
        L.oracle_bellman_ford.argtypes = [ctypes.c_int32, p, p, p, ctypes.c_int32, p]
        L.oracle_bellman_ford.restype = ctypes.c_int
        L.oracle_validate_gr.argtypes = [p, ctypes.c_int64, p, p]
        L.oracle_validate_gr.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _csr_args(row, col, w):
    row = np.ascontiguousarray(row, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    w = np.ascontiguousarray(w, dtype=np.int32)
    return row, col, w


def dijkstra(row, col, w, src: int):
    """Return (dist int32[n], total_work E_r, reached V_r)."""
    row, col, w = _csr_args(row, col, w)
    n = len(row) - 1
    out = np.empty(n, dtype=np.int32)
    work = ctypes.c_int64(0)
    reached = ctypes.c_int64(0)
    rc = lib().oracle_dijkstra(n, _ptr(row), _ptr(col), _ptr(w), int(src), _ptr(out),
                               ctypes.byref(work), ctypes.byref(reached))
    if rc != 0:
        raise ValueError(f"oracle_dijkstra failed rc={rc}")
    return out, int(work.value), int(reached.value)


def bellman_ford(row, col, w, src: int) -> np.ndarray:
    row, col, w = _csr_args(row, col, w)
    n = len(row) - 1
    out = np.empty(n, dtype=np.int32)
    rc = lib().oracle_bellman_ford(n, _ptr(row), _ptr(col), _ptr(w), int(src), _ptr(out))
    if rc != 0:
        raise ValueError(f"oracle_bellman_ford failed rc={rc}")
    return out


def validate_gr_bytes(img: bytes):
    buf = np.frombuffer(img, dtype=np.uint8)
    n = ctypes.c_int64(0)
    m = ctypes.c_int64(0)
    rc = lib().oracle_validate_gr(_ptr(buf), len(buf), ctypes.byref(n), ctypes.byref(m))
    if rc != 0:
        raise ValueError("invalid .gr image")
    return int(n.value), int(m.value)


def read_gr(path: str):
    """Parse a .gr v1 file -> (row int64[n+1] 0-based starts, col int32[m], w int32[m])."""
    raw = np.fromfile(path, dtype=np.uint8)
    hdr = raw[:32].view(np.uint64)
    if hdr[0] != 1 or hdr[1] != 4:
        raise ValueError("unsupported .gr header")
    n, m = int(hdr[2]), int(hdr[3])
    off = 32
    ends = raw[off:off + 8 * n].view(np.uint64)
    off += 8 * n
    col = raw[off:off + 4 * m].view(np.uint32).astype(np.int32)
    off += 4 * m + 4 * (m & 1)
    w = raw[off:off + 4 * m].view(np.int32).copy()
    row = np.zeros(n + 1, dtype=np.int64)
    row[1:] = ends.astype(np.int64)
    return row, col, w


def write_gr(path: str, row, col, w) -> None:
    row = np.asarray(row, dtype=np.int64)
    n = len(row) - 1
    m = int(row[-1])
    with open(path, "wb") as f:
        f.write(np.array([1, 4, n, m], dtype="<u8").tobytes())
        f.write(row[1:].astype("<u8").tobytes())
        f.write(np.asarray(col, dtype="<u4").tobytes())
        if m & 1:
            f.write(b"\0" * 4)
        f.write(np.asarray(w, dtype="<i4").tobytes())


def ref_available() -> bool:
    return os.access(REF_EXPORT, os.X_OK)


def ref_export_gr(gr_path: str, src: int, cut: int, out_gr: str) -> None:
    """Have the reference's own export tool write its shortcut-augmented
    graph (G+) of a .gr file (scripts/multigpu/supplement_graph.cpp:44-63)."""
    if not ref_available():
        raise RuntimeError("reference oracle binary not built")
    with tempfile.TemporaryDirectory() as td:
        subprocess.run([REF_EXPORT, gr_path, str(src), str(cut), out_gr,
                        os.path.join(td, "d.i32")], check=True, stdout=subprocess.DEVNULL)


def ref_dijkstra_gr(gr_path: str, src: int) -> np.ndarray:
    """Run the reference's own int64 oracle on a .gr file; return int32[n]."""
    if not ref_available():
        raise RuntimeError("reference oracle binary not built")
    row, _, _ = read_gr(gr_path)
    n = len(row) - 1
    with tempfile.TemporaryDirectory() as td:
        out_gr = os.path.join(td, "aug.gr")
        out_i32 = os.path.join(td, "d.i32")
        subprocess.run([REF_EXPORT, gr_path, str(src), str(n // 2), out_gr, out_i32],
                       check=True, stdout=subprocess.DEVNULL)
        return np.fromfile(out_i32, dtype=np.int32)


def bfs_reorder(row, col, root: int, layer_split: bool = False):
    """Return (perm int32[n], invperm int32[n]) with perm[new] = old.

    Breadth-first discovery order from ``root``, each row scanned in CSR order
    (graph_transform/reorder_gr.cu:94-126); vertices not reached follow in
    ascending old id (:128-134).  With ``layer_split``, vertices are grouped by
    BFS level (unreached: one last group) and each group, in discovery order,
    sends its first ``min(n//2, (end+1)//2) - placed_low`` vertices to the low
    half and the rest to the high half (:139-156).
    """
    row = np.asarray(row, dtype=np.int64)
    col = np.asarray(col, dtype=np.int64)
    n = len(row) - 1
    if not 0 <= root < n:
        raise ValueError("root out of range")
    perm = [0] * n
    inv = [-1] * n
    level = [-1] * n
    queue = [root]
    inv[root] = 0
    perm[0] = root
    level[root] = 0
    nxt, head = 1, 0
    while head < len(queue):
        u = queue[head]
        head += 1
        for e in range(int(row[u]), int(row[u + 1])):
            v = int(col[e])
            if inv[v] == -1:
                inv[v] = nxt
                level[v] = level[u] + 1
                perm[nxt] = v
                nxt += 1
                queue.append(v)
    for v in range(n):
        if inv[v] == -1:
            inv[v] = nxt
            perm[nxt] = v
            nxt += 1
    assert nxt == n
    if layer_split:
        cut = n // 2
        out = [0] * n
        left, right, begin = 0, cut, 0
        while begin < n:
            end = begin + 1
            while end < n and level[perm[end]] == level[perm[begin]]:
                end += 1
            take = min(cut, (end + 1) // 2) - left
            for i in range(begin, end):
                if i - begin < take:
                    out[left] = perm[i]
                    left += 1
                else:
                    out[right] = perm[i]
                    right += 1
            begin = end
        assert left == cut and right == n
        perm = out
        for i, v in enumerate(perm):
            inv[v] = i
    return np.array(perm, dtype=np.int32), np.array(inv, dtype=np.int32)


def relabel(row, col, w, perm, invperm):
    """New CSR with row i = old row perm[i], edges in old order, destinations
    renamed through invperm, weights kept (graph_transform/reorder_gr.cu:158-177)."""
    row = np.asarray(row, dtype=np.int64)
    perm = np.asarray(perm, dtype=np.int64)
    deg = (row[1:] - row[:-1])[perm]
    nrow = np.zeros(len(row), dtype=np.int64)
    np.cumsum(deg, out=nrow[1:])
    idx = np.concatenate([np.arange(row[p], row[p + 1]) for p in perm]) if len(col) else \
        np.zeros(0, dtype=np.int64)
    ncol = np.asarray(invperm, dtype=np.int32)[np.asarray(col)[idx]]
    nw = np.asarray(w, dtype=np.int32)[idx]
    return nrow, ncol.astype(np.int32), nw


def cross_edges(row, col, cut: int) -> int:
    """Directed edges whose ends lie on different sides of ``cut``
    (graph_transform/reorder_gr.cu:47-57)."""
    row = np.asarray(row, dtype=np.int64)
    src = np.repeat(np.arange(len(row) - 1), np.diff(row))
    return int(np.count_nonzero((src < cut) != (np.asarray(col) < cut)))
