"""Regenerate the 2052-vertex known-answer graph ``fixture.gr``.

The graph is the closed-form fixture of the reference's supplement checks
(scripts/multigpu/run_supplement_checks.py:27-40): a bidirectional path
0..2049 (forward weight 1 + u % 3, backward 1 + u % 5), a fan-out 0 -> 1..512
(weight 1000 + u), a fan-in u -> 1025 (u % 7) with the way back 1025 -> u
(2 + u % 5), zero-weight 2-cycles 5 <-> 6 and 1030 <-> 1031, and isolated
vertices 2050 and 2051.  Out-edges of a vertex are kept in insertion order, so
the file must come out byte-identical to the committed ``fixture.gr``
(SHA-256 24b038d8...).  The expected distances ``s{0,1030,2050}.i32`` were
produced by the reference's int64 Dijkstra and are committed as data.
"""
import os
import struct
import sys

N_VERTICES = 2052


def fixture_edges():
    out = [[] for _ in range(N_VERTICES)]

    def link(a, b, wt):
        out[a].append((b, wt))

    for u in range(2049):
        link(u, u + 1, 1 + u % 3)
        link(u + 1, u, 1 + u % 5)
    for u in range(1, 513):
        link(0, u, 1000 + u)
        link(u, 1025, u % 7)
        link(1025, u, 2 + u % 5)
    for u in (5, 1030):
        link(u, u + 1, 0)
        link(u + 1, u, 0)
    return out


def fixture_bytes() -> bytes:
    out = fixture_edges()
    ends, dst, wts, acc = [], [], [], 0
    for lst in out:
        acc += len(lst)
        ends.append(acc)
        for v, wt in lst:
            dst.append(v)
            wts.append(wt)
    m = acc
    blob = struct.pack("<4Q", 1, 4, N_VERTICES, m)
    blob += struct.pack(f"<{N_VERTICES}Q", *ends)
    blob += struct.pack(f"<{m}I", *dst)
    if m % 2:
        blob += b"\0" * 4
    blob += struct.pack(f"<{m}i", *wts)
    return blob


if __name__ == "__main__":
    path = sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(__file__), "fixture.gr")
    with open(path, "wb") as f:
        f.write(fixture_bytes())
