"""BFS vertex renumbering (graph_transform/reorder_gr.cu), CPU only.

The oracle's restatement is pinned first, by orders worked out by hand and by
an independent characterisation of breadth-first order: a reached vertex is
discovered by its in-neighbour that leaves the queue first (the smallest new
id), through that vertex's first edge to it, so the reached vertices sorted by
(discoverer's new id, index of that edge in CSR order) are exactly the new
order.  Then the product (libmlmq.so and the reorder_gr command) must match
the oracle bit for bit.
"""
import os
import subprocess

import numpy as np
import pytest

from graphs_util import grid_csr, random_csr
from oracle import oracle as O
from mlmq_tpds_withl3_amd import HostGraph, MLMQError

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
REORDER = os.path.join(os.path.dirname(HERE), "mlmq_tpds_withl3_amd", "reorder_gr")

# A 7-vertex directed graph with unsorted rows and an isolated vertex 6.
HAND_ROW = np.array([0, 2, 3, 4, 6, 6, 7, 7], dtype=np.int64)
HAND_COL = np.array([3, 1, 5, 0, 2, 4, 3], dtype=np.int32)
HAND_W = np.array([10, 11, 12, 13, 14, 15, 16], dtype=np.int32)


def shuffled_rows(row, col, w, seed):
    """Same graph with every row's edges in a random order."""
    rng = np.random.default_rng(seed)
    idx = np.concatenate([rng.permutation(np.arange(row[u], row[u + 1]))
                          for u in range(len(row) - 1)] or [np.zeros(0, np.int64)])
    return row, col[idx].copy(), w[idx].copy()


def bfs_by_discoverer(row, col, root):
    """New order from the discoverer characterisation (see the module doc)."""
    n = len(row) - 1
    reached = np.zeros(n, dtype=bool)
    reached[root] = True
    frontier = [root]
    while frontier:           # reachability only, order-free
        nxt = []
        for u in frontier:
            for v in col[row[u]:row[u + 1]]:
                if not reached[v]:
                    reached[v] = True
                    nxt.append(int(v))
        frontier = nxt
    order = [root]
    # The discoverer of v is fixed once every vertex before it is placed, so
    # build the order position by position.
    placed = {root: 0}
    pending = {}
    pos = 0
    while pos < len(order):
        u = order[pos]
        for e in range(int(row[u]), int(row[u + 1])):
            v = int(col[e])
            if v not in placed and v not in pending:
                pending[v] = (pos, e)
        # every pending vertex whose discoverer is u is placed now, by edge index
        mine = sorted((key, v) for v, key in pending.items() if key[0] == pos)
        for _, v in mine:
            placed[v] = len(order)
            order.append(v)
            del pending[v]
        pos += 1
    assert len(order) == int(reached.sum())
    order += [v for v in range(n) if not reached[v]]
    return np.array(order, dtype=np.int32)


# ---------------- oracle pinned ----------------

def test_oracle_hand_worked_orders():
    perm, inv = O.bfs_reorder(HAND_ROW, HAND_COL, 0)
    assert perm.tolist() == [0, 3, 1, 2, 4, 5, 6]
    assert inv.tolist() == [0, 2, 3, 1, 4, 5, 6]
    perm, _ = O.bfs_reorder(HAND_ROW, HAND_COL, 5)
    assert perm.tolist() == [5, 3, 2, 4, 0, 1, 6]
    # vertex 4 has no out-edges: everything else follows in ascending old id
    perm, _ = O.bfs_reorder(HAND_ROW, HAND_COL, 4)
    assert perm.tolist() == [4, 0, 1, 2, 3, 5, 6]
    # levels {0} {3,1} {2,4,5} unreached {6}, cut 3: lower half 0 | 3 | 2
    perm, inv = O.bfs_reorder(HAND_ROW, HAND_COL, 0, layer_split=True)
    assert perm.tolist() == [0, 3, 2, 1, 4, 5, 6]
    assert inv.tolist() == [0, 3, 2, 1, 4, 5, 6]


def test_oracle_hand_worked_relabel_and_cut():
    perm, inv = O.bfs_reorder(HAND_ROW, HAND_COL, 0)
    row, col, w = O.relabel(HAND_ROW, HAND_COL, HAND_W, perm, inv)
    assert row.tolist() == [0, 2, 4, 5, 6, 6, 7, 7]
    assert col.tolist() == [1, 2, 3, 4, 5, 0, 1]
    assert w.tolist() == [10, 11, 14, 15, 12, 13, 16]
    assert O.cross_edges(HAND_ROW, HAND_COL, 3) == 3
    assert O.cross_edges(row, col, 3) == 5


@pytest.mark.parametrize("seed", range(6))
def test_oracle_matches_discoverer_characterisation(seed):
    n = [1, 2, 40, 300, 500, 900][seed]
    row, col, w = random_csr(n, 2 * n, seed, self_loops=True, dupes=True)
    row, col, w = shuffled_rows(row, col, w, seed)
    for root in {0, n // 2, n - 1}:
        perm, inv = O.bfs_reorder(row, col, root)
        assert np.array_equal(perm, bfs_by_discoverer(row, col, root))
        assert np.array_equal(np.sort(perm), np.arange(n))
        assert np.array_equal(inv[perm], np.arange(n))


@pytest.mark.parametrize("seed", range(4))
def test_oracle_layer_split_invariants(seed):
    n = [7, 60, 401, 1000][seed]
    row, col, w = random_csr(n, 3 * n // 2, seed + 10)
    row, col, w = shuffled_rows(row, col, w, seed)
    root = seed % n
    base, _ = O.bfs_reorder(row, col, root)
    perm, inv = O.bfs_reorder(row, col, root, layer_split=True)
    cut = n // 2
    assert np.array_equal(np.sort(perm), np.arange(n))
    assert np.array_equal(inv[perm], np.arange(n))
    # both halves keep breadth-first order
    pos = np.empty(n, dtype=np.int64)
    pos[base] = np.arange(n)
    assert np.all(np.diff(pos[perm[:cut]]) > 0) and np.all(np.diff(pos[perm[cut:]]) > 0)
    # group ends: BFS levels of the plain order, then the unreached group
    level = np.full(n, -1)
    level[root] = 0
    frontier = [root]
    while frontier:
        nxt = []
        for u in frontier:
            for v in col[row[u]:row[u + 1]]:
                if level[v] < 0:
                    level[v] = level[u] + 1
                    nxt.append(int(v))
        frontier = nxt
    lv = level[base]
    ends = [i + 1 for i in range(n - 1) if lv[i] != lv[i + 1]] + [n]
    low = np.zeros(n, dtype=bool)
    low[perm[:cut]] = True
    for end in ends:
        assert int(low[base[:end]].sum()) == min(cut, (end + 1) // 2)
    # inside a group the lower-half share is a prefix of its discovery order
    begin = 0
    for end in ends:
        g = low[base[begin:end]]
        assert not np.any(g[1:] & ~g[:-1])
        begin = end


# ---------------- product == oracle ----------------

def _graphs():
    yield "hand", (HAND_ROW, HAND_COL, HAND_W), [0, 5, 6]
    yield "fixture", O.read_gr(os.path.join(GOLDEN, "fixture.gr")), [0, 1030, 2050]
    r = random_csr(3000, 7000, 3)
    yield "random", shuffled_rows(*r, 4), [0, 1499, 2999]
    yield "disconnected", random_csr(2000, 1500, 5), [7, 1000]
    yield "grid", grid_csr(40, 50, 6), [0, 1234]
    yield "single", (np.array([0, 0], np.int64), np.zeros(0, np.int32),
                     np.zeros(0, np.int32)), [0]


@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("name,csr,roots", list(_graphs()), ids=[g[0] for g in _graphs()])
def test_product_matches_oracle(name, csr, roots, split):
    row, col, w = csr
    g = HostGraph.from_csr(row, col, w)
    for root in roots:
        r, perm, inv = g.reorder_bfs(root, split)
        operm, oinv = O.bfs_reorder(row, col, root, split)
        assert np.array_equal(perm, operm) and np.array_equal(inv, oinv), (name, root)
        nrow, ncol, nw = r.arrays()
        orow, ocol, ow = O.relabel(row, col, w, operm, oinv)
        assert np.array_equal(nrow, orow) and np.array_equal(ncol, ocol)
        assert np.array_equal(nw, ow)
        cut = (len(row) - 1) // 2
        assert r.cross_edges() == O.cross_edges(orow, ocol, cut)
    assert g.cross_edges() == O.cross_edges(row, col, (len(row) - 1) // 2)
    assert g.cross_edges(0) == 0


def test_distances_follow_the_renumbering():
    row, col, w = O.read_gr(os.path.join(GOLDEN, "fixture.gr"))
    g = HostGraph.from_csr(row, col, w)
    for src in (0, 1030):
        for split in (False, True):
            r, perm, inv = g.reorder_bfs(src, split)
            d_old, work, reached = O.dijkstra(row, col, w, src)
            nrow, ncol, nw = r.arrays()
            d_new, work2, reached2 = O.dijkstra(nrow, ncol, nw, int(inv[src]))
            assert np.array_equal(d_new, d_old[perm])
            assert (work, reached) == (work2, reached2)


def test_knn_coordinates_follow_vertices():
    g = HostGraph.knn(3000, 3, seed=9, weight_scale=1e5)
    r, perm, inv = g.reorder_bfs(17)
    assert np.array_equal(r.coords(), g.coords()[perm])
    assert r.weight_scale() == g.weight_scale()


def test_bad_arguments():
    g = HostGraph.from_csr(HAND_ROW, HAND_COL, HAND_W)
    for root in (-1, 7):
        with pytest.raises(MLMQError):
            g.reorder_bfs(root)
    with pytest.raises(ValueError):
        g.cross_edges(8)
    with pytest.raises(ValueError):
        O.bfs_reorder(HAND_ROW, HAND_COL, 7)


# ---------------- the reorder_gr command ----------------

def run_tool(*args):
    return subprocess.run([REORDER, *map(str, args)], capture_output=True, text=True,
                          timeout=60)


@pytest.mark.parametrize("split", [False, True])
def test_command_outputs(tmp_path, split):
    src = os.path.join(GOLDEN, "fixture.gr")
    out = str(tmp_path / "bfs.gr")
    args = [src, out, 1030] + (["--layer-split"] if split else [])
    p = run_tool(*args)
    assert p.returncode == 0, p.stderr
    row, col, w = O.read_gr(src)
    perm, inv = O.bfs_reorder(row, col, 1030, split)
    assert np.array_equal(np.fromfile(out + ".perm.u32", dtype="<u4"), perm)
    assert np.array_equal(np.fromfile(out + ".invperm.u32", dtype="<u4"), inv)
    nrow, ncol, nw = O.read_gr(out)
    orow, ocol, ow = O.relabel(row, col, w, perm, inv)
    assert np.array_equal(nrow, orow) and np.array_equal(ncol, ocol) and np.array_equal(nw, ow)
    lines = p.stdout.splitlines()
    head = dict(kv.split("=") for kv in lines[0].split()[1:])
    assert lines[0].startswith("REORDER ")
    assert head["layout"] == ("bfs_layer_split" if split else "bfs")
    assert (head["root_old"], head["root_new"]) == ("1030", "0")
    assert (head["vertices"], head["edges"]) == ("2052", "5638")
    cut = 2052 // 2
    assert lines[1].split()[2] == str(O.cross_edges(row, col, cut))
    assert lines[2].split()[3] == str(O.cross_edges(orow, ocol, cut))
    assert lines[3].startswith("reduction: ")
    # never overwrites: a second run fails and leaves the outputs as they were
    before = open(out, "rb").read()
    p = run_tool(*args)
    assert p.returncode == 2 and "already exists" in p.stderr
    assert open(out, "rb").read() == before


def test_command_argument_errors(tmp_path):
    src = os.path.join(GOLDEN, "fixture.gr")
    assert run_tool(src).returncode == 255                                  # usage
    assert run_tool(src, tmp_path / "a.gr", "x").returncode == 2            # bad root
    assert run_tool(src, tmp_path / "b.gr", 2052).returncode == 2           # root >= n
    assert run_tool(src, tmp_path / "c.gr", 0, "--split").returncode == 2   # bad flag
    p = run_tool(tmp_path / "missing.gr", tmp_path / "d.gr")
    assert p.returncode == 255 and "Read error!" in p.stdout
    assert not os.path.exists(tmp_path / "a.gr") and not os.path.exists(tmp_path / "b.gr")
