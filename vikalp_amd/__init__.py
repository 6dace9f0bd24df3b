"""vikalp_amd — MI355X (gfx950) PDHG LP backend for the VIKALP solver.

The product is native: hand-written HIP kernels behind a C ABI
(lib/libvikalp_hip.so, include/vikalp_hip.h) and a C++20 host library with
the reference's contracts, a HipBackend ExecutionBackend and solve_lp
(lib/libvikalp.so, include/vikalp_solve.h).  This module is ctypes plumbing
for tests and the benchmark: it loads the in-tree libraries and raises if
they are missing; there is no Python or CPU fallback.
"""
from __future__ import annotations

from .native import (  # noqa: F401
    BackendPreference, ContinuousMethod, PdhgPolicy, PdhgPrecision, SolveStatus,
    STATUS_NAMES, LP, Options, Result, Device, hip_lib, host_lib, generate_lp,
    solve, solve_lp, solve_qp, load_libraries, profile_kernels, solve_batch, UnrolledPdhg, read_mps,
    solve_nodes, solve_mip, MIP_STATS,
)

__all__ = [
    "BackendPreference", "ContinuousMethod", "PdhgPolicy", "PdhgPrecision", "SolveStatus",
    "STATUS_NAMES", "LP", "Options", "Result", "Device", "hip_lib", "host_lib",
    "generate_lp", "solve", "solve_lp", "load_libraries", "profile_kernels", "solve_batch",
    "UnrolledPdhg", "solve_qp", "read_mps", "solve_nodes", "solve_mip", "MIP_STATS",
]
