"""The serving loop's slot arena (csrc/arena.h), compiled natively with g++
and exercised by a randomised allocate/release sequence with invariant checks
after every step (no GPU)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_arena_invariants(tmp_path):
    exe = tmp_path / "arena_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "mlmq_tpds_withl3_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "arena_check.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe), "20000"], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
