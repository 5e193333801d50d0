"""MI355X (gfx950) MLMQ single-source shortest paths.

A per-wavefront "small label first" deque in LDS over a global delta-bucket
queue, drained by persistent 64-lane wavefronts that relax a CSR graph with
atomicMin on 32-bit distances.  The compute path is the HIP library
``libmlmq.so`` (C ABI in include/mlmq.h); this package is its Python face.
"""
from ._lib import MLMQError, GraphInfo, Params, Timing, LIB_PATH, load  # noqa: F401
from .solver import (Solver, default_params, sssp_init, sssp_run_adaptive,  # noqa: F401
                     sssp_cp_data, sssp_release, UNREACHED)
from .graphs import HostGraph, WORKLOADS, make_workload  # noqa: F401

__all__ = ["Solver", "HostGraph", "default_params", "sssp_init", "sssp_run_adaptive",
           "sssp_cp_data", "sssp_release", "MLMQError", "WORKLOADS", "make_workload"]
