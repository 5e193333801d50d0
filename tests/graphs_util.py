"""Small seeded graphs shared by the CPU and GPU tests (numpy only)."""
import numpy as np


def random_csr(n, m, seed, wmax=100, zero_frac=0.0, self_loops=True, dupes=True):
    """Random directed multigraph as CSR (row int64[n+1], col int32[m], w int32[m])."""
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, size=m)
    dst = rng.integers(0, n, size=m)
    w = rng.integers(1, wmax + 1, size=m)
    if zero_frac:
        w[rng.random(m) < zero_frac] = 0
    if not self_loops:
        keep = src != dst
        src, dst, w = src[keep], dst[keep], w[keep]
    if not dupes:
        key = src.astype(np.int64) * n + dst
        _, idx = np.unique(key, return_index=True)
        src, dst, w = src[idx], dst[idx], w[idx]
    order = np.lexsort((dst, src))
    src, dst, w = src[order], dst[order], w[order]
    row = np.zeros(n + 1, dtype=np.int64)
    np.add.at(row, src + 1, 1)
    row = np.cumsum(row)
    return row, dst.astype(np.int32), w.astype(np.int32)


def grid_csr(rows, cols, seed, wmax=1000):
    """4-neighbour grid with random symmetric weights (road-like, long hop diameter)."""
    rng = np.random.default_rng(seed)
    n = rows * cols
    edges = []
    for r in range(rows):
        for c in range(cols):
            u = r * cols + c
            if c + 1 < cols:
                edges.append((u, u + 1))
            if r + 1 < rows:
                edges.append((u, u + cols))
    e = np.array(edges, dtype=np.int64)
    wt = rng.integers(1, wmax + 1, size=len(e))
    src = np.concatenate([e[:, 0], e[:, 1]])
    dst = np.concatenate([e[:, 1], e[:, 0]])
    w = np.concatenate([wt, wt])
    order = np.lexsort((dst, src))
    src, dst, w = src[order], dst[order], w[order]
    row = np.zeros(n + 1, dtype=np.int64)
    np.add.at(row, src + 1, 1)
    return np.cumsum(row), dst.astype(np.int32), w.astype(np.int32)


def reachable_work(row, col, src):
    """(V_r, E_r) by BFS: reachable vertices and the sum of their out-degrees."""
    n = len(row) - 1
    seen = np.zeros(n, dtype=bool)
    seen[src] = True
    stack = [src]
    while stack:
        u = stack.pop()
        for v in col[row[u]:row[u + 1]]:
            if not seen[v]:
                seen[v] = True
                stack.append(int(v))
    deg = np.diff(row)
    return int(seen.sum()), int(deg[seen].sum())


def hub_csr(n, hub_degrees, m_rest, seed, wmax=1000):
    """Random multigraph whose first len(hub_degrees) vertices are hubs with the
    given out-degrees (edges to uniform random targets), plus `m_rest` random
    edges among all vertices; every vertex also links back to hub 0 so that
    the hubs sit on many shortest paths."""
    rng = np.random.default_rng(seed)
    src, dst = [], []
    for h, deg in enumerate(hub_degrees):
        src.append(np.full(deg, h))
        dst.append(rng.integers(0, n, size=deg))
    src.append(rng.integers(0, n, size=m_rest))
    dst.append(rng.integers(0, n, size=m_rest))
    src.append(np.arange(n))
    dst.append(np.zeros(n, dtype=np.int64))
    src = np.concatenate(src)
    dst = np.concatenate(dst)
    w = rng.integers(1, wmax + 1, size=len(src))
    order = np.lexsort((dst, src))
    src, dst, w = src[order], dst[order], w[order]
    row = np.zeros(n + 1, dtype=np.int64)
    np.add.at(row, src + 1, 1)
    return np.cumsum(row), dst.astype(np.int32), w.astype(np.int32)
