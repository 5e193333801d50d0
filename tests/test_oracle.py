"""Pins the CPU oracle (oracle/dijkstra.c) before anything is checked against it.

Each pin is independent of the oracle's own code:
  * the reference's golden int64-Dijkstra vectors s{0,1030,2050}.i32 on the
    2052-vertex fixture (evidence/l3_supplement_20260921_v2/checks/);
  * textbook Bellman-Ford on small random multigraphs (zero weights, self
    loops, parallel edges, unreachable parts);
  * scipy.sparse.csgraph.dijkstra (a library routine) on larger graphs;
  * the reference's own int64 oracle, compiled from
    scripts/multigpu/supplement_graph.cpp into oracle/_ref/ (skipped where it
    was not built);
  * the closed form total_work = E_r = sum of out-degrees of reachable
    vertices (SS/SSSP/main.cu:71) and V_r from a BFS;
  * the reference's second known-answer graph: the fixture with its chain
    shortcuts (G+, 5644 edges, checks/s{0,1030,2050}.gr -- one file, the
    same for every source), whose distances equal the fixture's.
"""
import hashlib
import os

import numpy as np
import pytest

from oracle import oracle as O
from graphs_util import random_csr, grid_csr, reachable_work

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def golden(src):
    return np.fromfile(os.path.join(GOLDEN, f"s{src}.i32"), dtype=np.int32)


@pytest.fixture(scope="module")
def fixture_graph():
    return O.read_gr(os.path.join(GOLDEN, "fixture.gr"))


@pytest.mark.parametrize("src,work,reached", [(0, 5638, 2050), (1030, 5638, 2050), (2050, 0, 1)])
def test_golden_vectors(fixture_graph, src, work, reached):
    row, col, w = fixture_graph
    d, tw, vr = O.dijkstra(row, col, w, src)
    assert np.array_equal(d, golden(src))
    assert tw == work and vr == reached


def test_golden_properties(fixture_graph):
    # Facts the reference's export logs state about the vectors.
    d0, d1030, d2050 = golden(0), golden(1030), golden(2050)
    assert (d0 != O.UNREACHED).sum() == 2050 and d0[d0 != O.UNREACHED].max() == 2049
    assert (d1030 != O.UNREACHED).sum() == 2050 and d1030[d1030 != O.UNREACHED].max() == 2037
    assert (d2050 != O.UNREACHED).sum() == 1 and d2050[2050] == 0
    # zero-weight 2-cycles: 5 <-> 6 and 1030 <-> 1031 have equal distances
    assert d0[5] == d0[6] and d0[1030] == d0[1031]


@pytest.mark.parametrize("seed", range(12))
def test_bellman_ford_agrees(seed):
    n = 5 + seed * 7
    row, col, w = random_csr(n, n * 3, seed, wmax=20, zero_frac=0.2)
    for src in (0, n // 2, n - 1):
        d, _, _ = O.dijkstra(row, col, w, src)
        assert np.array_equal(d, O.bellman_ford(row, col, w, src))


@pytest.mark.parametrize("seed", range(4))
def test_scipy_agrees(seed):
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import dijkstra as sp_dijkstra

    n = 3000
    row, col, w = random_csr(n, 4 * n, 100 + seed, wmax=10_000, self_loops=False, dupes=False)
    mat = csr_matrix((w.astype(np.float64), col, row), shape=(n, n))
    for src in (0, 17, n - 1):
        d, _, _ = O.dijkstra(row, col, w, src)
        ref = sp_dijkstra(mat, directed=True, indices=src)
        ref = np.where(np.isinf(ref), O.UNREACHED, ref).astype(np.int64)
        assert np.array_equal(d.astype(np.int64), ref)


def test_total_work_closed_form():
    row, col, w = random_csr(500, 900, 7, wmax=50)   # sparse: several components
    for src in (0, 3, 499):
        d, tw, vr = O.dijkstra(row, col, w, src)
        v_r, e_r = reachable_work(row, col, src)
        assert (tw, vr) == (e_r, v_r)
        assert ((d != O.UNREACHED).sum()) == v_r


def test_grid_distances_symmetric():
    row, col, w = grid_csr(20, 30, 5)
    a, _, _ = O.dijkstra(row, col, w, 0)
    b, _, _ = O.dijkstra(row, col, w, 599)
    # undirected graph: d(0, 599) == d(599, 0)
    assert a[599] == b[0]


def test_unrepresentable_distance_rejected():
    # path 0 -> 1 -> 2 with weights summing to 2^31: the reference's int64
    # oracle throws "unrepresentable distance" (supplement_graph.cpp:40)
    row = np.array([0, 1, 2, 2], dtype=np.int64)
    col = np.array([1, 2], dtype=np.int32)
    w = np.array([2 ** 30, 2 ** 30], dtype=np.int32)
    with pytest.raises(ValueError):
        O.dijkstra(row, col, w, 0)
    w2 = np.array([2 ** 30, 2 ** 30 - 2], dtype=np.int32)
    d, _, _ = O.dijkstra(row, col, w2, 0)
    assert d[2] == 2 ** 31 - 2


def test_invalid_graphs_rejected():
    row = np.array([0, 1, 1], dtype=np.int64)
    with pytest.raises(ValueError):
        O.dijkstra(row, np.array([5], np.int32), np.array([1], np.int32), 0)
    with pytest.raises(ValueError):
        O.dijkstra(row, np.array([1], np.int32), np.array([-1], np.int32), 0)
    with pytest.raises(ValueError):
        O.dijkstra(row, np.array([1], np.int32), np.array([1], np.int32), 2)


def test_gr_validator(tmp_path):
    img = open(os.path.join(GOLDEN, "fixture.gr"), "rb").read()
    assert O.validate_gr_bytes(img) == (2052, 5638)
    with pytest.raises(ValueError):
        O.validate_gr_bytes(img[:-4])                      # truncated
    bad = bytearray(img)
    bad[8] = 8                                             # edge size 8
    with pytest.raises(ValueError):
        O.validate_gr_bytes(bytes(bad))


@pytest.mark.skipif(not O.ref_available(), reason="reference oracle not built here")
@pytest.mark.parametrize("seed", range(3))
def test_reference_oracle_agrees(tmp_path, seed):
    row, col, w = random_csr(800, 3000, 50 + seed, wmax=1000, zero_frac=0.05)
    path = str(tmp_path / "g.gr")
    O.write_gr(path, row, col, w)
    for src in (0, 401):
        d, _, _ = O.dijkstra(row, col, w, src)
        assert np.array_equal(d, O.ref_dijkstra_gr(path, src))


@pytest.mark.skipif(not O.ref_available(), reason="reference oracle not built here")
def test_reference_oracle_golden():
    path = os.path.join(GOLDEN, "fixture.gr")
    for src in (0, 1030, 2050):
        assert np.array_equal(O.ref_dijkstra_gr(path, src), golden(src))


SHORTCUT_SHA = "59a814ec13ef9aaaf5cadc9f7f52bd4cb1899122bf967be04ded6e023acf5439"


def test_shortcut_fixture_known_answers():
    path = os.path.join(GOLDEN, "fixture_shortcut.gr")
    assert hashlib.sha256(open(path, "rb").read()).hexdigest() == SHORTCUT_SHA
    row, col, w = O.read_gr(path)
    assert len(row) - 1 == 2052 and len(col) == 5644
    for src in (0, 1030, 2050):
        d, _, _ = O.dijkstra(row, col, w, src)
        assert np.array_equal(d, golden(src)), src


@pytest.mark.skipif(not O.ref_available(), reason="reference oracle not built here")
def test_shortcut_fixture_is_the_reference_export(tmp_path):
    want = open(os.path.join(GOLDEN, "fixture_shortcut.gr"), "rb").read()
    for src in (0, 1030, 2050):
        out = str(tmp_path / f"g{src}.gr")
        O.ref_export_gr(os.path.join(GOLDEN, "fixture.gr"), src, 2052 // 2, out)
        assert open(out, "rb").read() == want, src
