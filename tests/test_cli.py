"""The mlmq command-line driver keeps the reference driver's process contract
(SS/SSSP/main.cu:146-239, SS/SSSP/sssp_run.cu:758-886): `-i graph.gr [-n 1]
[-d delta]`, BENCH_* environment, a CPU reference run first ("CPU total work
count", "CPU sequential time"), one `BENCH key=value ...` line per run with
correct=1|0 from a per-vertex comparison, that the reference harness's parser
and row check accept (scripts/multigpu/run_usa_road_matrix.py:34-84),
per-vertex checking against an L3_SUPPLEMENT_ORACLE file, exit code 2 on a
wrong result or bad input; and the multi-source mode `-s N`
(SSSP/main.cu:195-307, 396-434) with its MULTI-SOURCE lines.

Argument errors are rejected before any GPU call, so those cases run on the
CPU; the solving cases are marked gpu."""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
CLI = os.path.join(os.path.dirname(HERE), "mlmq_tpds_withl3_amd", "mlmq")
FIXTURE = os.path.join(GOLDEN, "fixture.gr")


def run(args, timeout=120, **env):
    e = dict(os.environ)
    for k in ("BENCH_SOURCE", "BENCH_DELTA", "BENCH_WARMUPS", "BENCH_REPEATS", "BENCH_QUEUE",
              "MLMQ_WORK_BLOCKS", "L3_SUPPLEMENT_ORACLE", "MLMQ_SKIP_CPU_CHECK",
              "MLMQ_MULTI_MODE", "MLMQ_L2_REPORT"):
        e.pop(k, None)
    e.update({k: str(v) for k, v in env.items()})
    return subprocess.run([CLI, *args], capture_output=True, text=True, env=e, timeout=timeout)


def bench_rows(stdout):
    """The reference harness's parse_bench: key=value tokens of BENCH lines."""
    rows = []
    for line in stdout.splitlines():
        if line.startswith("BENCH "):
            rows.append(dict(t.split("=", 1) for t in line.split()[1:] if "=" in t))
    return rows


def harness_check(rows, source, queue, warmups, repeats):
    """The reference harness's row validation (run_usa_road_matrix.py:60-84)."""
    assert len(rows) == warmups + repeats
    for i, row in enumerate(rows):
        assert row["algorithm"] == "MLMQ" and row["gpu_count"] == "1"
        assert row["source"] == str(source) and row["repeat"] == str(i)
        assert row["warmup"] == str(int(i < warmups)) and row["correct"] == "1"
        assert row["queue"] == queue
        assert float(row["solve_ms"]) > 0 and float(row["query_wall_ms"]) > 0


# ---- argument errors (CPU) ----

def test_usage_without_graph():
    p = run([])
    assert p.returncode == 2 and "Usage" in p.stdout


@pytest.mark.parametrize("args,env,text", [
    (["-i", FIXTURE, "-n", "2"], {}, "only -n 1"),
    (["-i", FIXTURE, "-d", "0"], {}, ""),
    (["-i", FIXTURE], {"BENCH_SOURCE": 2052}, "Invalid BENCH_SOURCE"),
    (["-i", FIXTURE], {"BENCH_SOURCE": "x"}, "Invalid BENCH_SOURCE"),
    (["-i", FIXTURE], {"BENCH_QUEUE": "L1V_L2V"}, "unsupported BENCH_QUEUE"),
    (["-i", FIXTURE], {"BENCH_REPEATS": 0}, "Invalid BENCH_REPEATS"),
    (["-i", FIXTURE], {"L3_SUPPLEMENT_ORACLE": os.path.join(GOLDEN, "fixture.gr")},
     "ORACLE_SIZE_ERROR"),
    (["-i", os.path.join(GOLDEN, "missing.gr")], {}, "cannot read"),
    (["-i", FIXTURE, "-s", "0"], {}, "sources must be >= 1"),
    (["-i", FIXTURE, "-s", "4"], {"L3_SUPPLEMENT_ORACLE": os.path.join(GOLDEN, "s0.i32")}, "not with -s"),
    (["-i", FIXTURE, "-s", "4"], {"MLMQ_MULTI_MODE": "parallel"}, "unsupported MLMQ_MULTI_MODE"),
    (["-i", FIXTURE], {"MLMQ_SKIP_CPU_CHECK": 2}, "Invalid MLMQ_SKIP_CPU_CHECK"),
])
def test_argument_errors(args, env, text):
    p = run(args, **env)
    assert p.returncode == 2, (p.stdout, p.stderr)
    assert text in p.stderr + p.stdout
    assert not bench_rows(p.stdout)


def cpu_lines(stdout):
    work = [int(l.split()[-1]) for l in stdout.splitlines() if l.startswith("CPU total work count ")]
    ms = [float(l.split()[-2]) for l in stdout.splitlines() if l.startswith("CPU sequential time ")]
    return work, ms


@pytest.mark.parametrize("src", [0, 1030, 2050])
def test_cpu_reference_run_precedes_the_gpu(src):
    """The driver's own CPU Dijkstra runs before any GPU call (as the
    reference's main does) and reports the reachable edge count: checked
    against the oracle's total work on the golden fixture.  On a machine
    without a GPU the run stops at device setup, after these lines."""
    from oracle import oracle as O
    row, col, w = O.read_gr(FIXTURE)
    _, work_ref, _ = O.dijkstra(row, col, w, src)
    p = run(["-i", FIXTURE], BENCH_SOURCE=src, BENCH_WARMUPS=0, BENCH_REPEATS=1)
    work, ms = cpu_lines(p.stdout)
    assert work == [work_ref] and len(ms) == 1 and ms[0] >= 0
    if "MLMQ_CONFIG" in p.stdout:   # a GPU is present: the CPU lines came first
        assert p.stdout.index("CPU total work count") < p.stdout.index("MLMQ_CONFIG")
    else:
        assert p.returncode == 2


# ---- solving (GPU) ----

@pytest.mark.gpu
@pytest.mark.parametrize("queue", ["L1SLF_L2DQ", "L1V_L2DQ"])
@pytest.mark.parametrize("src", [0, 1030, 2050])
def test_golden_runs_pass_the_harness_check(src, queue):
    oracle = os.path.join(GOLDEN, f"s{src}.i32")
    p = run(["-i", FIXTURE, "-n", "1", "-d", "200000"], BENCH_SOURCE=src, BENCH_WARMUPS=1,
            BENCH_REPEATS=2, BENCH_QUEUE=queue, L3_SUPPLEMENT_ORACLE=oracle)
    assert p.returncode == 0, p.stderr
    harness_check(bench_rows(p.stdout), src, queue, 1, 2)
    assert p.stdout.count("WIDE_ORACLE vertices=2052 correct=1") == 3
    lines = p.stdout.splitlines()
    first = [int(x) for x in lines[lines.index("gpu") + 1].split()]
    assert first == np.fromfile(oracle, dtype=np.int32)[:30].tolist()


@pytest.mark.gpu
def test_wrong_oracle_exits_2():
    p = run(["-i", FIXTURE], BENCH_SOURCE=1030, BENCH_WARMUPS=0, BENCH_REPEATS=3,
            L3_SUPPLEMENT_ORACLE=os.path.join(GOLDEN, "s0.i32"))
    assert p.returncode == 2 and "ORACLE_MISMATCH" in p.stderr
    rows = bench_rows(p.stdout)
    assert len(rows) == 1 and rows[0]["correct"] == "0"


@pytest.mark.gpu
@pytest.mark.parametrize("queue", ["L1SLF_L2DQ", "L1V_L2DQ"])
def test_default_run_checks_against_its_cpu_reference(queue):
    """No oracle file: the reference driver's default run, checked per vertex
    against the driver's own CPU reference, passes the harness row check."""
    p = run(["-i", os.path.join(GOLDEN, "fixture_shortcut.gr")], BENCH_SOURCE=1030, BENCH_WARMUPS=1,
            BENCH_REPEATS=2, BENCH_QUEUE=queue)
    assert p.returncode == 0, p.stderr
    harness_check(bench_rows(p.stdout), 1030, queue, 1, 2)
    assert "mlmq sssp correct!" in p.stdout and "WIDE_ORACLE" not in p.stdout
    lines = p.stdout.splitlines()
    exp = np.fromfile(os.path.join(GOLDEN, "s1030.i32"), dtype=np.int32)[:30].tolist()
    assert [int(x) for x in lines[lines.index("gpu") + 1].split()] == exp
    assert [int(x) for x in lines[lines.index("cpu") + 1].split()] == exp


@pytest.mark.gpu
def test_skip_cpu_check_reports_unchecked():
    p = run(["-i", os.path.join(GOLDEN, "fixture_shortcut.gr")], BENCH_WARMUPS=0,
            BENCH_REPEATS=1, MLMQ_WORK_BLOCKS=64, MLMQ_SKIP_CPU_CHECK=1)
    assert p.returncode == 0, p.stderr
    rows = bench_rows(p.stdout)
    assert len(rows) == 1 and rows[0]["correct"] == "-1" and rows[0]["work_blocks"] == "64"
    assert "CPU total work count" not in p.stdout


def multi_source_report(stdout):
    per = {}
    for line in stdout.splitlines():
        if line.startswith("MULTI-SOURCE source["):
            k = int(line.split("[", 1)[1].split("]", 1)[0])
            per[k] = line.split("=", 1)[1]
    return per


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [None, "sequential"])
def test_multi_source(mode):
    """-s 8: eight sources spread over the ids (k * n / 8), each answered and
    checked against the CPU reference; the golden ones against their files."""
    env = {"MLMQ_MULTI_MODE": mode} if mode else {}
    p = run(["-i", FIXTURE, "-n", "1", "-s", "8"], **env)
    assert p.returncode == 0, p.stdout + p.stderr
    per = multi_source_report(p.stdout)
    ids = [k * 2052 // 8 for k in range(8)]
    assert per == {k: f"{ids[k]} correct" for k in range(8)}
    assert "MULTI-SOURCE errors=0" in p.stdout and "MULTI-SOURCE all sources correct!" in p.stdout
    assert "MULTI-SOURCE sources=8 nGPU=1 batch wall time=" in p.stdout
    assert "MULTI-SOURCE CPU references ready: sources=8" in p.stdout
    assert f"mode={mode or 'batch'}" in p.stdout


@pytest.mark.gpu
def test_multi_source_more_sources_than_vertices():
    """More sources than vertices wrap around (k mod n), as the reference's
    source list does; on a 4 x 4 grid every source is answered exactly."""
    import tempfile
    from graphs_util import grid_csr
    from oracle import oracle as O
    row, col, w = grid_csr(4, 4, 3)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "grid.gr")
        O.write_gr(path, row, col, w)
        p = run(["-i", path, "-s", "20"])
    assert p.returncode == 0, p.stdout + p.stderr
    per = multi_source_report(p.stdout)
    assert per == {k: f"{k % 16} correct" for k in range(20)}


def fields(stdout, tag):
    """The reference audit's field parser (scripts/multigpu/audit_supplement.py:10)."""
    return [dict(t.split("=", 1) for t in line.split()[1:] if "=" in t)
            for line in stdout.splitlines() if line.startswith(tag + " ")]


@pytest.mark.gpu
@pytest.mark.parametrize("queue", ["L1SLF_L2DQ", "L1V_L2DQ", "L1N_L2DQ"])
def test_queue_audit_lines(queue):
    """L2_CAPACITY once, then per query L2_FINAL and (MLMQ_L2_REPORT=1) one
    CAPACITY_BUCKET line per bucket and a CAPACITY_QUERY line; the checks the
    reference's capacity audit applies to them hold (audit_supplement.py:28-36,
    60-67: reads == writes per bucket, total == completed)."""
    p = run(["-i", FIXTURE], BENCH_SOURCE=1030, BENCH_WARMUPS=1, BENCH_REPEATS=2, BENCH_QUEUE=queue,
            MLMQ_L2_REPORT=1)
    assert p.returncode == 0, p.stderr
    caps = fields(p.stdout, "L2_CAPACITY")
    assert len(caps) == 1
    c = caps[0]
    assert int(c["record_bytes"]) == 8 and int(c["buckets"]) == 16
    assert int(c["per_bucket"]) * int(c["buckets"]) <= int(c["allocated_records"])
    finals = fields(p.stdout, "L2_FINAL")
    assert len(finals) == 3
    assert all(x["reads"] == x["writes"] == x["completed"] for x in finals)
    buckets = fields(p.stdout, "CAPACITY_BUCKET")
    queries = fields(p.stdout, "CAPACITY_QUERY")
    assert len(buckets) == 3 * 16 and len(queries) == 3
    for b in buckets:
        assert b["reads"] == b["writes"] and 0 <= int(b["writes"])
        assert b["capacity"] == c["per_bucket"]
    for i, q in enumerate(queries):
        total = sum(int(b["writes"]) for b in buckets[16 * i:16 * (i + 1)])
        assert int(q["total"]) == total == int(q["completed"]) == int(finals[i]["writes"])
    assert all(r["correct"] == "1" for r in bench_rows(p.stdout))
