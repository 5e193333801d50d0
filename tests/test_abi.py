"""The C-ABI library loads and exports every entry point its headers declare
(no GPU calls here)."""
import ctypes
import os
import re
import subprocess

import pytest

from mlmq_tpds_withl3_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("mlmq.h", "mlmq_graph.h")]


def declared_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("defined",)))


def test_headers_parse():
    names = declared_functions(HEADERS[0]) + declared_functions(HEADERS[1])
    assert "mlmq_solve" in names and "sssp_init" in names and "mlmq_graph_read_gr" in names


@pytest.mark.parametrize("header", HEADERS)
def test_every_declared_symbol_is_exported(header):
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_functions(header):
        assert hasattr(lib, name), f"{name} declared in {os.path.basename(header)} but not exported"


def test_binding_lists_match_headers():
    declared = set(declared_functions(HEADERS[0])) | set(declared_functions(HEADERS[1]))
    assert declared == set(_lib.SOLVER_SYMBOLS) | set(_lib.GRAPH_SYMBOLS)


def test_loader_binds_all():
    L = _lib.load()
    assert L.mlmq_version().decode().startswith("mlmq-gfx950")
    p = _lib.Params()
    L.mlmq_params_default(ctypes.byref(p))
    assert (p.queue_mode, p.delta, p.buckets, p.bucket_window, p.l2_batch, p.min_read) == \
        (19, 200000, 16, 4, 8, 8)
    assert (p.l1_pop, p.l1_spill, p.l1_donate, p.profile) == (64, 64, 0, 0)


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
