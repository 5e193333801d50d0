"""The C-ABI library loads and exports every entry point its headers declare
(no GPU calls here)."""
import ctypes
import os
import re
import subprocess

import pytest

from mlmq_tpds_withl3_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("mlmq.h", "mlmq_graph.h", "mlmq_serve.h")]


def declared_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("defined",)))


def test_headers_parse():
    names = sum((declared_functions(h) for h in HEADERS), [])
    assert "mlmq_solve" in names and "sssp_init" in names and "mlmq_graph_read_gr" in names
    assert "mlmq_server_run" in names


@pytest.mark.parametrize("header", HEADERS)
def test_every_declared_symbol_is_exported(header):
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_functions(header):
        assert hasattr(lib, name), f"{name} declared in {os.path.basename(header)} but not exported"


def test_binding_lists_match_headers():
    declared = set().union(*(declared_functions(h) for h in HEADERS))
    assert declared == set(_lib.SOLVER_SYMBOLS) | set(_lib.GRAPH_SYMBOLS) | set(_lib.SERVE_SYMBOLS)


def test_loader_binds_all():
    L = _lib.load()
    assert L.mlmq_version().decode().startswith("mlmq-gfx950")
    p = _lib.Params()
    L.mlmq_params_default(ctypes.byref(p))
    assert (p.queue_mode, p.delta, p.buckets, p.bucket_window, p.l2_batch, p.min_read) == \
        (19, 200000, 16, 4, 8, 8)
    assert (p.l1_pop, p.l1_spill, p.l1_donate, p.profile) == (32, 64, 0, 0)
    assert (p.batch_waves, p.use_graph, p.hub_split, p.stale_check, p.prefetch, p.l1_blocks,
            p.idle_backoff, p.defer_heavy, p.strict_window, p.light_split) == \
        (8, 1, 0, 2, 1, 0, 16, 0, 0, 0)


def test_library_is_gfx950_and_independent_of_oracle():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    assert "oracle_" not in out
    deps = subprocess.run(["ldd", _lib.LIB_PATH], capture_output=True, text=True).stdout
    assert "liboracle" not in deps
    blob = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_cli_built():
    assert os.access(_lib.CLI_PATH, os.X_OK)
    r = subprocess.run([_lib.CLI_PATH], capture_output=True, text=True)
    assert r.returncode == 2 and "Usage" in r.stdout


def test_serving_defaults():
    L = _lib.load()
    c = _lib.ServerConfig()
    L.mlmq_server_config_default(ctypes.byref(c))
    assert (c.hbm_budget, c.streams, c.max_batch, c.max_pending, c.verify) == (32 << 30, 4, 64, 0, 1)
    g = _lib.ServeGraphParams()
    L.mlmq_serve_graph_params_default(ctypes.byref(g))
    assert (g.delta, g.buckets, g.bucket_window, g.l2_batch, g.min_read) == (200000, 16, 4, 8, 8)
    assert (g.l1_pop, g.l1_spill, g.workgroups, g.waves, g.bucket_capacity) == (32, 64, 0, 8, 0)


def test_ctypes_layouts_match_c(tmp_path):
    """Every structure the binding mirrors has the size and field offsets the
    C compiler gives it."""
    structs = {
        "mlmq_params": (_lib.Params, ["delta", "bucket_capacity", "timeout_s", "use_graph", "hub_split",
                                       "stale_check", "prefetch", "l1_blocks", "idle_backoff", "defer_heavy",
                                       "strict_window", "light_split", "csr_stream", "l1_ring", "batch_stats",
                                       "reserved"]),
        "mlmq_timing": (_lib.Timing, ["donated", "hub_chunks"]),
        "mlmq_batch_timing": (_lib.BatchTiming, None),
        "mlmq_graph_info": (_lib.GraphInfo, ["diameter"]),
        "mlmq_queue_counts": (_lib.QueueCounts, ["capacity", "writes", "reads", "completed", "wrapped",
                                                 "reserved"]),
        "mlmq_server_config": (_lib.ServerConfig, ["timeout_s", "spread_waves", "mode", "live_waves",
                                                   "live_idle_ms", "reserved"]),
        "mlmq_serve_graph_params": (_lib.ServeGraphParams, ["bucket_capacity", "hub_split", "light_split",
                                                              "prefetch", "stale_check", "reserved"]),
        "mlmq_serve_request": (_lib.ServeRequest, ["src", "dist_out"]),
        "mlmq_serve_result": (_lib.ServeResult, ["status", "retries"]),
        "mlmq_serve_stats": (_lib.ServeStats, ["requests", "peak_inflight", "reserved"]),
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mlmq_serve.h"', "int main(void){"]
    for cname, (cls, fields) in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f in fields or []:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                 check=True).stdout.splitlines())
    for cname, (cls, fields) in structs.items():
        assert int(out[cname]) == ctypes.sizeof(cls), cname
        for f in fields or []:
            assert int(out[f"{cname}.{f}"]) == getattr(cls, f).offset, f"{cname}.{f}"


def test_min_read_zero_is_refused_before_any_launch():
    """min_read = 0 would leave the seed in L2 for ever (a wave tops up only
    below min_read records) and the solve would spin until the watchdog: the
    check every solve entry point runs first refuses it."""
    L = _lib.load()
    p = _lib.Params()
    L.mlmq_params_default(ctypes.byref(p))
    assert L.mlmq_params_check(ctypes.byref(p), 0) == 0
    assert L.mlmq_params_check(ctypes.byref(p), 1) == 0
    for bad in (0, -1, 65):
        p.min_read = bad
        assert L.mlmq_params_check(ctypes.byref(p), 0) == -1, bad
        assert L.mlmq_params_check(ctypes.byref(p), 1) == -1, bad
    p.min_read = 1
    assert L.mlmq_params_check(ctypes.byref(p), 0) == 0
    p.min_read = 64
    assert L.mlmq_params_check(ctypes.byref(p), 0) == 0


def test_params_check_ranges():
    L = _lib.load()

    def check(batch=0, **kw):
        p = _lib.Params()
        L.mlmq_params_default(ctypes.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        return L.mlmq_params_check(ctypes.byref(p), batch)

    assert check(delta=0) == -1
    assert check(buckets=65) == -1 and check(buckets=4, bucket_window=5) == -1
    assert check(l2_batch=0) == -1 and check(l1_pop=65) == -1
    assert check(stale_check=4) == -1 and check(work_blocks=-1) == -1
    assert check(light_split=1, defer_heavy=1) == -1 and check(light_split=1, hub_split=1) == 0
    assert check(queue_mode=5) == -9
    assert check(queue_mode=0) == 0 and check(queue_mode=0, batch=1) == -9
    assert check(batch=1, batch_waves=3) == -1 and check(batch=1, batch_waves=16) == 0


def test_serve_params_min_read_default_and_range():
    """Serving parameters: 0 and -1 select the default 8, never 0 itself;
    above 64 is refused."""
    L = _lib.load()
    for given, want in ((0, 8), (-1, 8), (1, 1), (64, 64)):
        g = _lib.ServeGraphParams()
        L.mlmq_serve_graph_params_default(ctypes.byref(g))
        g.min_read = given
        out = _lib.ServeGraphParams()
        assert L.mlmq_serve_graph_params_resolve(ctypes.byref(g), 1000, 4000, ctypes.byref(out)) == 0
        assert out.min_read == want
    g.min_read = 65
    assert L.mlmq_serve_graph_params_resolve(ctypes.byref(g), 1000, 4000, ctypes.byref(out)) == -1
    g = _lib.ServeGraphParams()   # all zero: every default filled in
    assert L.mlmq_serve_graph_params_resolve(ctypes.byref(g), 1000, 4000, ctypes.byref(out)) == 0
    assert (out.delta, out.buckets, out.bucket_window, out.l2_batch, out.min_read, out.l1_pop,
            out.waves, out.workgroups) == (200000, 16, 4, 8, 8, 32, 8, 1)


def test_l1_ring_ranges():
    """The L1 ring size and the spill threshold it allows (one batch's pushes,
    64 lanes x 7 edges, must fit above the threshold), checked without a
    device."""
    from mlmq_tpds_withl3_amd.solver import default_params
    L = _lib.load()

    def rc(batch, **kw):
        return L.mlmq_params_check(ctypes.byref(default_params(**kw)), batch)

    assert default_params().l1_ring == 512
    assert rc(1, l1_ring=1024, l1_spill=255) == 0
    assert rc(1, l1_ring=1024, l1_spill=256) == -1
    assert rc(1, l1_ring=512, l1_spill=64) == 0
    assert rc(1, l1_ring=512, l1_spill=65) == -1
    assert rc(1, l1_ring=0, l1_spill=64) == 0
    assert rc(1, l1_ring=768) == -1
    assert rc(0, l1_ring=1024) == -9          # single solves keep the 512-record ring
    assert rc(0, csr_stream=2) == -1
