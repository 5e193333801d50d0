"""The mlmq command-line driver keeps the reference driver's process contract
(SS/SSSP/main.cu:146-239, SS/SSSP/sssp_run.cu:758-886): `-i graph.gr [-n 1]
[-d delta]`, BENCH_* environment, one `BENCH key=value ...` line per run that
the reference harness's parser and row check accept
(scripts/multigpu/run_usa_road_matrix.py:34-84), per-vertex checking against
an L3_SUPPLEMENT_ORACLE file, exit code 2 on a wrong result or bad input.

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
              "MLMQ_WORK_BLOCKS", "L3_SUPPLEMENT_ORACLE"):
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
])
def test_argument_errors(args, env, text):
    p = run(args, **env)
    assert p.returncode == 2, (p.stdout, p.stderr)
    assert text in p.stderr + p.stdout
    assert not bench_rows(p.stdout)


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
def test_without_oracle_reports_unchecked():
    p = run(["-i", os.path.join(GOLDEN, "fixture_shortcut.gr")], BENCH_WARMUPS=0,
            BENCH_REPEATS=1, MLMQ_WORK_BLOCKS=64)
    assert p.returncode == 0, p.stderr
    rows = bench_rows(p.stdout)
    assert len(rows) == 1 and rows[0]["correct"] == "-1" and rows[0]["work_blocks"] == "64"
