"""ctypes bindings of libvikalp_hip.so and libvikalp.so (see include/*.h)."""
from __future__ import annotations

import ctypes as C
import enum
import math
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
HIP_SO = os.path.join(LIB_DIR, "libvikalp_hip.so")
HOST_SO = os.path.join(LIB_DIR, "libvikalp.so")

_dp = C.POINTER(C.c_double)
_lp = C.POINTER(C.c_int64)
_ip = C.POINTER(C.c_int32)
_vp = C.c_void_p


class BackendPreference(enum.IntEnum):
    Auto = 0
    CPU = 1
    CUDA = 2
    HIP = 3


class ContinuousMethod(enum.IntEnum):
    Auto = 0
    PDHG = 1
    RevisedSimplex = 2
    InteriorPoint = 3


class PdhgPolicy(enum.IntEnum):
    Reference = 0
    Adaptive = 1
    Halpern = 2


class PdhgPrecision(enum.IntEnum):
    FP64 = 0
    Mixed = 1


class SolveStatus(enum.IntEnum):
    NotSolved = 0
    Optimal = 1
    LocallyOptimal = 2
    Feasible = 3
    Infeasible = 4
    Unbounded = 5
    IterationLimit = 6
    NodeLimit = 7
    TimeLimit = 8
    NumericalFailure = 9
    UnsupportedModel = 10
    InvalidModel = 11


STATUS_NAMES = [s.name for s in SolveStatus]


class _SolverOptions(C.Structure):
    _fields_ = [
        ("backend", C.c_int), ("continuous_method", C.c_int),
        ("iteration_limit", C.c_int64), ("node_limit", C.c_int64),
        ("time_limit_seconds", C.c_double), ("primal_tolerance", C.c_double),
        ("dual_tolerance", C.c_double), ("integrality_tolerance", C.c_double),
        ("absolute_gap_tolerance", C.c_double), ("relative_gap_tolerance", C.c_double),
        ("deterministic", C.c_int), ("pdhg_policy", C.c_int), ("pdhg_precision", C.c_int),
        ("fused_iterations", C.c_int), ("use_graphs", C.c_int), ("reserved", C.c_int),
        ("relative_kkt_tolerance", C.c_double), ("mixed_phase_tolerance", C.c_double),
        ("restart_period", C.c_int64), ("spmv_col_blocks", C.c_int), ("spmv_autotune", C.c_int),
        ("spmv_split_passes", C.c_int), ("spmv_cache_budget_mb", C.c_int),
        ("spmv_panels", C.c_int), ("spmv_panel_groups", C.c_int), ("spmv_panel_groups_t", C.c_int),
        ("spmv_panel_threads", C.c_int), ("spmv_panel_kb", C.c_int), ("spmv_tile_geometry", C.c_int),
        ("mixed_recentre", C.c_int), ("reserved2", C.c_int),
        ("pdhg_primal_weight_gain", C.c_double), ("pdhg_primal_weight_integral", C.c_double),
    ]


class _SolveInfo(C.Structure):
    _fields_ = [
        ("status", C.c_int), ("reserved", C.c_int),
        ("iterations", C.c_int64), ("low_precision_iterations", C.c_int64),
        ("objective_value", C.c_double), ("primal_bound", C.c_double), ("dual_bound", C.c_double),
        ("primal_residual", C.c_double), ("stationarity_residual", C.c_double),
        ("complementarity_residual", C.c_double), ("absolute_gap", C.c_double),
        ("relative_gap", C.c_double), ("solve_seconds", C.c_double), ("device_seconds", C.c_double),
        ("kkt_relative_primal", C.c_double), ("kkt_relative_dual", C.c_double),
        ("kkt_relative_gap", C.c_double), ("kkt_relative", C.c_double),
        ("kkt_primal_objective", C.c_double), ("kkt_dual_objective", C.c_double),
        ("backend", C.c_char * 16), ("solver", C.c_char * 32), ("message", C.c_char * 160),
    ]


class PdhgParams(C.Structure):
    _fields_ = [
        ("policy", C.c_int), ("precision", C.c_int), ("iteration_limit", C.c_int64),
        ("metric_tolerance", C.c_double), ("relative_kkt_tolerance", C.c_double),
        ("step_size", C.c_double), ("primal_weight", C.c_double), ("restart_period", C.c_int64),
        ("norm_b", C.c_double), ("norm_c", C.c_double), ("use_graphs", C.c_int),
        ("grid_blocks", C.c_int), ("col_blocks", C.c_int), ("split_passes", C.c_int),
        ("cache_budget_mb", C.c_int), ("spmv_panels", C.c_int), ("panel_groups", C.c_int),
        ("panel_groups_t", C.c_int), ("panel_threads", C.c_int), ("panel_tune", C.c_int),
        ("panel_kb", C.c_int), ("tile_geometry", C.c_int), ("recentre", C.c_int),
        ("primal_weight_gain", C.c_double), ("primal_weight_integral", C.c_double),
    ]


class PdhgState(C.Structure):
    _fields_ = [
        ("iterations", C.c_int64), ("status", C.c_int), ("restarts", C.c_int),
        ("tau", C.c_double), ("sigma", C.c_double), ("primal_weight", C.c_double),
        ("metric", C.c_double), ("primal_scaled", C.c_double), ("dual_scaled", C.c_double),
        ("rel_primal", C.c_double), ("rel_dual", C.c_double), ("rel_gap", C.c_double),
        ("rel_kkt", C.c_double), ("primal_objective", C.c_double), ("dual_objective", C.c_double),
        ("converged_on_average", C.c_int), ("reserved", C.c_int),
    ]


# Every exported symbol, with its ctypes signature (restype, argtypes).
HIP_SYMBOLS = {
    "vk_hip_init": (C.c_int, [C.c_int]),
    "vk_hip_shutdown": (C.c_int, []),
    "vk_hip_last_error": (C.c_char_p, []),
    "vk_hip_stream": (_vp, []),
    "vk_hip_alloc": (C.c_int, [C.POINTER(_vp), C.c_size_t]),
    "vk_hip_free": (C.c_int, [_vp]),
    "vk_hip_memcpy_h2d": (C.c_int, [_vp, _vp, C.c_size_t, _vp]),
    "vk_hip_memcpy_d2h": (C.c_int, [_vp, _vp, C.c_size_t, _vp]),
    "vk_hip_memcpy_d2d": (C.c_int, [_vp, _vp, C.c_size_t, _vp]),
    "vk_hip_sync": (C.c_int, [_vp]),
    "vk_hip_fill": (C.c_int, [_vp, C.c_int64, C.c_double, _vp]),
    "vk_hip_axpby": (C.c_int, [C.c_int64, C.c_double, _vp, C.c_double, _vp, _vp]),
    "vk_hip_project_box": (C.c_int, [C.c_int64, _vp, _vp, _vp, _vp]),
    "vk_hip_reduce_absmax": (C.c_int, [C.c_int64, _vp, _dp, _vp]),
    "vk_hip_reduce_dot": (C.c_int, [C.c_int64, _vp, _vp, _dp, _vp]),
    "vk_hip_reduce_dot_noalloc": (C.c_int, [C.c_int64, _vp, _vp, _dp, _vp]),
    "vk_hip_csr_create": (C.c_int, [C.c_int64, C.c_int64, _lp, _lp, C.POINTER(_vp)]),
    "vk_hip_csr_set_values": (C.c_int, [_vp, _dp]),
    "vk_hip_csr_destroy": (C.c_int, [_vp]),
    "vk_hip_csr_dims": (C.c_int, [_vp, _lp, _lp, _lp]),
    "vk_hip_csr_transpose_export": (C.c_int, [_vp, _ip, _ip, _dp]),
    "vk_hip_spmv_csr": (C.c_int, [_vp, C.c_int, C.c_double, _vp, C.c_double, _vp, _vp]),
    "vk_hip_csr_norm2_estimate": (C.c_int, [_vp, C.c_int, _dp]),
    "vk_hip_lp_scale": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp,
                                  _dp, _dp]),
    "vk_hip_lp_scale_bounds": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp, _dp]),
    "vk_hip_csr_operator_bound": (C.c_int, [_vp, _dp]),
    "vk_hip_csr_operator_bounds": (C.c_int, [_vp, C.c_int, _lp, _lp, _dp]),
    "vk_hip_csr_postsolve_products": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp]),
    "vk_hip_pdhg_create": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                     C.POINTER(PdhgParams), C.POINTER(_vp)]),
    "vk_hip_pdqp_create": (C.c_int, [_vp, _vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                     C.POINTER(PdhgParams), C.POINTER(_vp)]),
    "vk_hip_pdhg_run": (C.c_int, [_vp, C.c_int64, C.POINTER(C.c_int)]),
    "vk_hip_pdhg_read_state": (C.c_int, [_vp, C.POINTER(PdhgState)]),
    "vk_hip_pdhg_download": (C.c_int, [_vp, _dp, _dp]),
    "vk_hip_pdhg_last_run_ms": (C.c_int, [_vp, _dp]),
    "vk_hip_pdhg_profile": (C.c_int, [_vp, C.c_int64, _dp]),
    "vk_hip_pdhg_layout": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "vk_hip_pdhg_panel_layout": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "vk_hip_csr_panel_spmv": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int,
                                        _dp]),
    "vk_hip_csr_set_tile_geometry": (C.c_int, [_vp, C.c_int]),
    "vk_hip_csr_tile_geometry": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "vk_hip_csr_tune_tiles": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(C.c_int), _dp]),
    "vk_hip_debug_fail_allocations": (C.c_int, [C.c_int64, C.POINTER(C.c_int64)]),
    "vk_hip_pdhg_destroy": (C.c_int, [_vp]),
    "vk_hip_csr_tune_col_blocks": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(C.c_int), _dp]),
    "vk_hip_batch_create": (C.c_int, [C.c_int, _lp, _lp, _lp, _lp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                      _dp, _dp, _dp, _dp, _dp, _dp, _dp, _lp, C.POINTER(PdhgParams),
                                      C.POINTER(_vp)]),
    "vk_hip_batch_create_csr": (C.c_int, [C.c_int, _lp, _lp, _vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                          _dp, _dp, _dp, _dp, _lp, C.POINTER(PdhgParams), C.POINTER(_vp)]),
    "vk_hip_batch_create_shared": (C.c_int, [C.c_int, _vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                             _dp, _dp, _dp, _lp, C.POINTER(PdhgParams), C.POINTER(_vp)]),
    "vk_hip_batch_run": (C.c_int, [_vp, C.c_int64, C.c_int64, C.POINTER(C.c_int)]),
    "vk_hip_batch_read_state": (C.c_int, [_vp, C.c_int, C.POINTER(PdhgState)]),
    "vk_hip_batch_download": (C.c_int, [_vp, _dp, _dp]),
    "vk_hip_batch_last_run_ms": (C.c_int, [_vp, _dp]),
    "vk_hip_batch_destroy": (C.c_int, [_vp]),
    "vk_hip_unrolled_create": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_int64,
                                         C.c_int64, C.POINTER(_vp)]),
    "vk_hip_unrolled_forward": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "vk_hip_unrolled_backward": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp]),
    "vk_hip_unrolled_last_ms": (C.c_int, [_vp, _dp]),
    "vk_hip_unrolled_set_data": (C.c_int, [_vp, _dp, _dp, _dp]),
    "vk_hip_unrolled_destroy": (C.c_int, [_vp]),
}

HOST_SYMBOLS = {
    "vk_solver_options_default": (None, [C.POINTER(_SolverOptions)]),
    "vk_lp_solve": (C.c_int, [C.c_int, C.c_int64, C.c_int64, _lp, _lp, _dp, _dp, C.c_double,
                              _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.POINTER(_SolverOptions),
                              _dp, _dp, C.POINTER(_SolveInfo)]),
    "vk_lp_solve_batch": (C.c_int, [C.c_int, _lp, _lp, _lp, _lp, _dp, _dp, _dp, _dp, _dp, _dp,
                                    C.POINTER(_SolverOptions), _dp, _dp, C.POINTER(_SolveInfo)]),
    "vk_lp_solve_nodes": (C.c_int, [C.c_int64, C.c_int64, _lp, _lp, _dp, _dp, C.c_double, _dp, _dp, _dp, _dp,
                                    C.c_int, _dp, _dp, _dp, C.POINTER(_SolverOptions), _dp, _dp,
                                    C.POINTER(_SolveInfo)]),
    "vk_mip_solve_batched": (C.c_int, [C.c_int64, C.c_int64, _lp, _lp, _dp, _dp, C.c_double, _dp, _dp, _dp, _dp,
                                       _ip, _lp, _lp, _dp, C.c_int64, C.POINTER(_SolverOptions), _dp,
                                       C.POINTER(_SolveInfo), _lp]),
    "vk_lp_profile": (C.c_int, [C.c_int64, C.c_int64, _lp, _lp, _dp, _dp, _dp, _dp, _dp, _dp,
                                C.POINTER(_SolverOptions), C.c_int64, C.c_int64, _dp]),
    "vk_lp_generate": (C.c_int, [C.c_int, C.c_int64, C.c_int64, C.c_int64, C.c_uint64, C.POINTER(_vp)]),
    "vk_lp_dims": (C.c_int, [_vp, _lp, _lp, _lp]),
    "vk_lp_export": (C.c_int, [_vp, _lp, _lp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "vk_lp_free": (None, [_vp]),
    "vk_lp_block_periods": (C.c_int64, [C.c_uint64, C.c_int64]),
    "vk_solve_last_error": (C.c_char_p, []),
    "vk_qp_solve": (C.c_int, [C.c_int, C.c_int64, C.c_int64, _lp, _lp, _dp, _lp, _lp, _dp, _dp, C.c_double,
                              _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.POINTER(_SolverOptions),
                              _dp, _dp, C.POINTER(_SolveInfo)]),
    "vk_unrolled_create": (C.c_int, [C.c_int64, C.c_int64, _lp, _lp, _dp, _dp, _dp, _dp, _dp, _dp,
                                     C.c_int64, C.c_int64, C.POINTER(_vp)]),
    "vk_unrolled_forward": (C.c_int, [_vp, _dp, _dp]),
    "vk_unrolled_backward": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp]),
    "vk_unrolled_info": (C.c_int, [_vp, _dp, _dp]),
    "vk_unrolled_set_data": (C.c_int, [_vp, _dp, _dp, _dp]),
    "vk_unrolled_destroy": (C.c_int, [_vp]),
    "vk_mps_read": (C.c_int, [C.c_char_p, C.POINTER(_vp)]),
    "vk_model_dims": (C.c_int, [_vp, _lp, _lp, _lp]),
    "vk_model_export": (C.c_int, [_vp, _lp, _lp, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _dp]),
    "vk_model_name": (C.c_int64, [_vp, C.c_char_p, C.c_int64]),
    "vk_model_free": (None, [_vp]),
}

_hip = None
_host = None


def _bind(lib, table):
    for name, (res, args) in table.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def hip_lib():
    """libvikalp_hip.so (raises if it was not built)."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_SO):
            raise RuntimeError(f"{HIP_SO} is missing: run `python -m vikalp_amd.build`")
        lib = C.CDLL(HIP_SO)
        _bind(lib, HIP_SYMBOLS)
        _hip = lib
    return _hip


def host_lib():
    """libvikalp.so (raises if it was not built)."""
    global _host
    if _host is None:
        hip_lib()
        if not os.path.exists(HOST_SO):
            raise RuntimeError(f"{HOST_SO} is missing: run `python -m vikalp_amd.build`")
        lib = C.CDLL(HOST_SO)
        _bind(lib, HOST_SYMBOLS)
        _host = lib
    return _host


def load_libraries():
    return hip_lib(), host_lib()


def _ptr(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = hip_lib().vk_hip_last_error()
        raise RuntimeError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


@dataclass
class LP:
    """An LP in the reference's canonical form (CSR, int64 indices)."""
    m: int
    n: int
    row_offsets: np.ndarray
    column_indices: np.ndarray
    values: np.ndarray
    objective: np.ndarray
    row_lower: np.ndarray
    row_upper: np.ndarray
    var_lower: np.ndarray
    var_upper: np.ndarray
    objective_offset: float = 0.0
    x0: np.ndarray | None = None
    y0: np.ndarray | None = None
    name: str = ""
    variable_types: np.ndarray | None = None   # 0 continuous, 1 integer, 2 binary

    @property
    def nnz(self) -> int:
        return int(self.row_offsets[-1])

    def as_oracle(self):
        from oracle import oracle as orc  # test/bench only
        return orc.LP(self.m, self.n, self.row_offsets, self.column_indices, self.values,
                      self.objective, self.row_lower, self.row_upper, self.var_lower,
                      self.var_upper, self.objective_offset)


def read_mps(path: str) -> LP:
    """MpsReader::read (src/io/MpsReader.cpp:10-553) through vk_mps_read: the
    model as an LP with its NAME and variable types.  Raises ValueError with
    every "line N: message" of a rejected file."""
    L = host_lib()
    h = _vp()
    rc = L.vk_mps_read(os.fsencode(path), C.byref(h))
    if rc != 0:
        msg = L.vk_solve_last_error()
        raise ValueError(f"vk_mps_read({path!r}) rejected the file:\n{msg.decode() if msg else ''}")
    try:
        mm, nn, nz = C.c_int64(), C.c_int64(), C.c_int64()
        L.vk_model_dims(h, C.byref(mm), C.byref(nn), C.byref(nz))
        m, n, nnz = mm.value, nn.value, nz.value
        a = dict(rp=np.zeros(m + 1, np.int64), ci=np.zeros(nnz, np.int64), v=np.zeros(nnz),
                 c=np.zeros(n), rl=np.zeros(m), ru=np.zeros(m), l=np.zeros(n), u=np.zeros(n),
                 t=np.zeros(n, np.int32))
        off = C.c_double()
        L.vk_model_export(h, _ptr(a["rp"], _lp), _ptr(a["ci"], _lp), _ptr(a["v"]), _ptr(a["c"]), _ptr(a["rl"]),
                          _ptr(a["ru"]), _ptr(a["l"]), _ptr(a["u"]), _ptr(a["t"], _ip), C.byref(off))
        size = L.vk_model_name(h, None, 0)
        buf = C.create_string_buffer(int(size) + 1)
        L.vk_model_name(h, buf, int(size) + 1)
    finally:
        L.vk_model_free(h)
    return LP(m, n, a["rp"], a["ci"], a["v"], a["c"], a["rl"], a["ru"], a["l"], a["u"], off.value,
              name=buf.value.decode(), variable_types=a["t"])


def generate_lp(family: int, m: int = 0, n: int = 0, k: int = 0, seed: int = 20261203) -> LP:
    """Seeded generator of include/vikalp_lpgen.h (family 1 uniform, 2 zipf, 4 block)."""
    L = host_lib()
    h = _vp()
    rc = L.vk_lp_generate(family, m, n, k, seed, C.byref(h))
    if rc != 0:
        raise ValueError(f"vk_lp_generate({family}, {m}, {n}, {k}) failed with {rc}")
    try:
        mm, nn, nz = C.c_int64(), C.c_int64(), C.c_int64()
        L.vk_lp_dims(h, C.byref(mm), C.byref(nn), C.byref(nz))
        m, n, nnz = mm.value, nn.value, nz.value
        a = dict(rp=np.zeros(m + 1, np.int64), ci=np.zeros(nnz, np.int64), v=np.zeros(nnz),
                 c=np.zeros(n), rl=np.zeros(m), ru=np.zeros(m), l=np.zeros(n), u=np.zeros(n),
                 x0=np.zeros(n), y0=np.zeros(m))
        L.vk_lp_export(h, _ptr(a["rp"], _lp), _ptr(a["ci"], _lp), _ptr(a["v"]), _ptr(a["c"]),
                       _ptr(a["rl"]), _ptr(a["ru"]), _ptr(a["l"]), _ptr(a["u"]), _ptr(a["x0"]),
                       _ptr(a["y0"]))
    finally:
        L.vk_lp_free(h)
    return LP(m, n, a["rp"], a["ci"], a["v"], a["c"], a["rl"], a["ru"], a["l"], a["u"],
              0.0, a["x0"], a["y0"])


@dataclass
class Options:
    """vikalp::SolverOptions (+ the HIP-path fields)."""
    backend: int = BackendPreference.Auto
    continuous_method: int = ContinuousMethod.Auto
    iteration_limit: int = 100_000
    node_limit: int = 100_000
    time_limit_seconds: float = math.inf
    primal_tolerance: float = 1e-7
    dual_tolerance: float = 1e-7
    integrality_tolerance: float = 1e-6
    absolute_gap_tolerance: float = 1e-6
    relative_gap_tolerance: float = 1e-4
    deterministic: bool = True
    pdhg_policy: int = PdhgPolicy.Reference
    pdhg_precision: int = PdhgPrecision.FP64
    fused_iterations: bool = True
    use_graphs: bool = True
    relative_kkt_tolerance: float = 0.0
    mixed_phase_tolerance: float = 0.0
    restart_period: int = 64
    spmv_col_blocks: int = 0
    spmv_autotune: bool = False
    spmv_split_passes: int = 0
    spmv_cache_budget_mb: int = 0
    spmv_panels: int = 0
    spmv_panel_groups: int = 0
    spmv_panel_groups_t: int = 0
    spmv_panel_threads: int = 0
    spmv_panel_kb: int = 0
    spmv_tile_geometry: int = 0
    mixed_recentre: bool = True
    pdhg_primal_weight_gain: float = 0.0
    pdhg_primal_weight_integral: float = 0.0

    def _c(self) -> _SolverOptions:
        o = _SolverOptions()
        for name, _ in _SolverOptions._fields_:
            if name.startswith("reserved"):
                continue
            v = getattr(self, name)
            setattr(o, name, int(v) if isinstance(v, (bool, enum.IntEnum)) else v)
        return o


@dataclass
class Result:
    status: str
    iterations: int
    objective_value: float
    primal_residual: float
    stationarity_residual: float
    complementarity_residual: float
    absolute_gap: float
    relative_gap: float
    solve_seconds: float
    device_seconds: float
    kkt_relative_primal: float
    kkt_relative_dual: float
    kkt_relative_gap: float
    kkt_relative: float
    kkt_primal_objective: float
    kkt_dual_objective: float
    low_precision_iterations: int
    backend: str
    solver: str
    message: str
    x: np.ndarray = field(repr=False, default=None)
    y: np.ndarray = field(repr=False, default=None)


def _solve(mode, lp: LP, options: Options, warm_start=None, lower_override=None,
           upper_override=None, quadratic=None) -> Result:
    L = host_lib()
    keep = [np.ascontiguousarray(lp.row_offsets, dtype=np.int64),
            np.ascontiguousarray(lp.column_indices, dtype=np.int64)]
    f = [np.ascontiguousarray(a, dtype=np.float64) for a in
         (lp.values, lp.objective, lp.row_lower, lp.row_upper, lp.var_lower, lp.var_upper)]
    opt = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
           for a in (warm_start, lower_override, upper_override)]
    x = np.zeros(lp.n)
    y = np.zeros(lp.m)
    info = _SolveInfo()
    copt = options._c()
    if quadratic is None:
        rc = L.vk_lp_solve(mode, lp.m, lp.n, _ptr(keep[0], _lp), _ptr(keep[1], _lp), _ptr(f[0]),
                           _ptr(f[1]), float(lp.objective_offset), _ptr(f[2]), _ptr(f[3]), _ptr(f[4]),
                           _ptr(f[5]), _ptr(opt[0]), _ptr(opt[1]), _ptr(opt[2]), C.byref(copt),
                           _ptr(x), _ptr(y), C.byref(info))
    else:
        q = [np.ascontiguousarray(quadratic[0], dtype=np.int64), np.ascontiguousarray(quadratic[1], dtype=np.int64),
             np.ascontiguousarray(quadratic[2], dtype=np.float64)]
        rc = L.vk_qp_solve(mode, lp.m, lp.n, _ptr(keep[0], _lp), _ptr(keep[1], _lp), _ptr(f[0]),
                           _ptr(q[0], _lp), _ptr(q[1], _lp), _ptr(q[2]),
                           _ptr(f[1]), float(lp.objective_offset), _ptr(f[2]), _ptr(f[3]), _ptr(f[4]),
                           _ptr(f[5]), _ptr(opt[0]), _ptr(opt[1]), _ptr(opt[2]), C.byref(copt),
                           _ptr(x), _ptr(y), C.byref(info))
    if rc != 0:
        raise ValueError(f"vk_lp_solve / vk_qp_solve rejected its arguments ({rc})")
    return Result(STATUS_NAMES[info.status], info.iterations, info.objective_value,
                  info.primal_residual, info.stationarity_residual, info.complementarity_residual,
                  info.absolute_gap, info.relative_gap, info.solve_seconds, info.device_seconds,
                  info.kkt_relative_primal, info.kkt_relative_dual, info.kkt_relative_gap,
                  info.kkt_relative, info.kkt_primal_objective, info.kkt_dual_objective,
                  info.low_precision_iterations, info.backend.decode(), info.solver.decode(),
                  info.message.decode(), x, y)


def profile_kernels(lp: LP, options: Options | None = None, warm_iterations: int = 50,
                    iterations: int = 50) -> dict:
    """Mean fused-kernel device times (ms) measured with HIP events (vk_lp_profile)."""
    L = host_lib()
    rp = np.ascontiguousarray(lp.row_offsets, dtype=np.int64)
    ci = np.ascontiguousarray(lp.column_indices, dtype=np.int64)
    f = [np.ascontiguousarray(a, dtype=np.float64) for a in
         (lp.values, lp.objective, lp.row_lower, lp.row_upper, lp.var_lower, lp.var_upper)]
    out = np.zeros(20)
    copt = (options or Options())._c()
    rc = L.vk_lp_profile(lp.m, lp.n, _ptr(rp, _lp), _ptr(ci, _lp), *[_ptr(a) for a in f],
                         C.byref(copt), int(warm_iterations), int(iterations), _ptr(out))
    if rc != 0:
        raise RuntimeError(f"vk_lp_profile failed ({rc}): {hip_lib().vk_hip_last_error()}")
    return {"dual_ms": out[0], "primal_ms": out[1], "iteration_ms": out[2], "step_size": out[3],
            "col_blocks": int(out[4]), "split_passes": int(out[5]), "dual_spmv_ms": out[6],
            "dual_pass_ms": out[7], "primal_spmv_ms": out[8], "primal_pass_ms": out[9],
            "panels": [int(v) for v in out[10:20]]}


def _result(info, x, y):
    return Result(STATUS_NAMES[info.status], info.iterations, info.objective_value,
                  info.primal_residual, info.stationarity_residual, info.complementarity_residual,
                  info.absolute_gap, info.relative_gap, info.solve_seconds, info.device_seconds,
                  info.kkt_relative_primal, info.kkt_relative_dual, info.kkt_relative_gap,
                  info.kkt_relative, info.kkt_primal_objective, info.kkt_dual_objective,
                  info.low_precision_iterations, info.backend.decode(), info.solver.decode(),
                  info.message.decode(), x, y)


def solve_batch(lps: list, options: Options | None = None) -> list:
    """solver::solve_lp_batch on the HIP batched persistent kernel."""
    L = host_lib()
    count = len(lps)
    rs = np.zeros(count + 1, np.int64)
    cs = np.zeros(count + 1, np.int64)
    for i, lp in enumerate(lps):
        rs[i + 1] = rs[i] + lp.m
        cs[i + 1] = cs[i] + lp.n
    rp = [np.zeros(1, np.int64)]
    nnz0 = 0
    for lp in lps:
        rp.append(np.asarray(lp.row_offsets[1:], np.int64) + nnz0)
        nnz0 += lp.nnz
    rp = np.concatenate(rp)
    ci = np.concatenate([np.asarray(lp.column_indices, np.int64) + cs[i] for i, lp in enumerate(lps)])
    cat = lambda k: np.ascontiguousarray(np.concatenate([np.asarray(getattr(lp, k), np.float64) for lp in lps]))
    v, c, rl, ru, l, u = (cat(k) for k in ("values", "objective", "row_lower", "row_upper", "var_lower", "var_upper"))
    x = np.zeros(int(cs[-1]))
    y = np.zeros(int(rs[-1]))
    infos = (_SolveInfo * count)()
    copt = (options or Options())._c()
    rc = L.vk_lp_solve_batch(count, _ptr(rs, _lp), _ptr(cs, _lp), _ptr(rp, _lp), _ptr(ci, _lp), _ptr(v), _ptr(c),
                             _ptr(rl), _ptr(ru), _ptr(l), _ptr(u), C.byref(copt), _ptr(x), _ptr(y), infos)
    if rc != 0:
        raise ValueError(f"vk_lp_solve_batch rejected its arguments ({rc})")
    return [_result(infos[i], x[cs[i]:cs[i + 1]].copy(), y[rs[i]:rs[i + 1]].copy()) for i in range(count)]


def solve_nodes(lp: LP, lower=None, upper=None, warm_start=None, options: Options | None = None) -> list:
    """solver::solve_lp_nodes through vk_lp_solve_nodes: the LP with each
    node's variable bounds (lower / upper: count x n arrays, None = the LP's
    own) and warm starts (count x n or None), as one batch on one matrix."""
    L = host_lib()
    arrs = [a for a in (lower, upper, warm_start) if a is not None]
    if not arrs:
        raise ValueError("solve_nodes needs node bounds")
    count = len(arrs[0])
    node = lambda a: None if a is None else np.ascontiguousarray(np.asarray(a, np.float64).reshape(count, lp.n))
    lo, hi, ws = node(lower), node(upper), node(warm_start)
    k = [np.ascontiguousarray(a, dtype=t) for a, t in (
        (lp.row_offsets, np.int64), (lp.column_indices, np.int64), (lp.values, np.float64),
        (lp.objective, np.float64), (lp.row_lower, np.float64), (lp.row_upper, np.float64),
        (lp.var_lower, np.float64), (lp.var_upper, np.float64))]
    x = np.zeros(count * lp.n)
    y = np.zeros(count * lp.m)
    infos = (_SolveInfo * count)()
    copt = (options or Options())._c()
    rc = L.vk_lp_solve_nodes(lp.m, lp.n, _ptr(k[0], _lp), _ptr(k[1], _lp), _ptr(k[2]), _ptr(k[3]),
                             float(lp.objective_offset), _ptr(k[4]), _ptr(k[5]), _ptr(k[6]), _ptr(k[7]), count,
                             _ptr(lo), _ptr(hi), _ptr(ws), C.byref(copt), _ptr(x), _ptr(y), infos)
    if rc != 0:
        raise ValueError(f"vk_lp_solve_nodes rejected its arguments ({rc})")
    return [_result(infos[i], x[i * lp.n:(i + 1) * lp.n].copy(), y[i * lp.m:(i + 1) * lp.m].copy())
            for i in range(count)]


MIP_STATS = ("nodes_processed", "nodes_pruned", "nodes_integral", "nodes_infeasible", "incumbent_updates",
             "rounds", "relaxations", "relaxation_iterations", "max_batch")


def solve_mip(lp: LP, options: Options | None = None, batch: int = 64, variable_types=None, quadratic=None):
    """solver::solve_mip_batched through vk_mip_solve_batched: branch-and-bound
    whose node relaxations are solved `batch` at a time on one device matrix.
    variable_types (default lp.variable_types): 0 continuous, 1 integer,
    2 binary; quadratic = (row_offsets, column_indices, values) of a symmetric
    Q makes it an MIQP.  Returns (Result with the incumbent as x, dict of
    MIP_STATS)."""
    L = host_lib()
    vt = variable_types if variable_types is not None else lp.variable_types
    if vt is None:
        raise ValueError("solve_mip needs variable types")
    vt = np.ascontiguousarray(vt, dtype=np.int32)
    k = [np.ascontiguousarray(a, dtype=t) for a, t in (
        (lp.row_offsets, np.int64), (lp.column_indices, np.int64), (lp.values, np.float64),
        (lp.objective, np.float64), (lp.row_lower, np.float64), (lp.row_upper, np.float64),
        (lp.var_lower, np.float64), (lp.var_upper, np.float64))]
    q = [None, None, None]
    if quadratic is not None:
        q = [np.ascontiguousarray(quadratic[0], dtype=np.int64), np.ascontiguousarray(quadratic[1], dtype=np.int64),
             np.ascontiguousarray(quadratic[2], dtype=np.float64)]
    x = np.zeros(lp.n)
    info = _SolveInfo()
    stats = np.zeros(len(MIP_STATS), np.int64)
    copt = (options or Options())._c()
    rc = L.vk_mip_solve_batched(lp.m, lp.n, _ptr(k[0], _lp), _ptr(k[1], _lp), _ptr(k[2]), _ptr(k[3]),
                                float(lp.objective_offset), _ptr(k[4]), _ptr(k[5]), _ptr(k[6]), _ptr(k[7]),
                                _ptr(vt, _ip), _ptr(q[0], _lp), _ptr(q[1], _lp), _ptr(q[2]), int(batch),
                                C.byref(copt), _ptr(x), C.byref(info), _ptr(stats, _lp))
    if rc != 0:
        raise ValueError(f"vk_mip_solve_batched rejected its arguments ({rc})")
    return _result(info, x, np.zeros(lp.m)), dict(zip(MIP_STATS, (int(v) for v in stats)))


def solve(lp: LP, options: Options | None = None, quadratic=None) -> Result:
    """vikalp::solve(model, options) (Solver.cpp:92-99) on the HIP library;
    quadratic = (row_offsets, column_indices, values) of Q makes it a QP."""
    return _solve(0, lp, options or Options(), quadratic=quadratic)


def solve_qp(lp: LP, quadratic, options: Options | None = None, warm_start=None, lower_override=None,
             upper_override=None) -> Result:
    """vikalp::solver::solve_qp(model, hip_backend, options, warm, lo, hi) (pdqp.cpp:303-331);
    quadratic = (row_offsets, column_indices, values) of the symmetric n x n Q."""
    return _solve(1, lp, options or Options(), warm_start, lower_override, upper_override, quadratic)


def solve_lp(lp: LP, options: Options | None = None, warm_start=None, lower_override=None,
             upper_override=None) -> Result:
    """vikalp::solver::solve_lp(model, hip_backend, options, warm, lo, hi) (pdhg.cpp:269-293)."""
    return _solve(1, lp, options or Options(), warm_start, lower_override, upper_override)


class UnrolledPdhg:
    """solver::UnrolledPdhg: K fixed-step PDHG iterations as a differentiable
    map of the objective and the row bounds (vk_unrolled_* of
    include/vikalp_solve.h).  forward() -> (x_K, y_K) in original coordinates;
    backward(grad_x, grad_y) -> (grad_c, grad_row_lower, grad_row_upper)
    through the iterations (L's explicit dependence on c is the caller's)."""

    def __init__(self, lp: LP, iterations: int, checkpoint_every: int = 50):
        self.L = host_lib()
        self.m, self.n = lp.m, lp.n
        rp = np.ascontiguousarray(lp.row_offsets, dtype=np.int64)
        ci = np.ascontiguousarray(lp.column_indices, dtype=np.int64)
        f = [np.ascontiguousarray(a, dtype=np.float64) for a in
             (lp.values, lp.objective, lp.row_lower, lp.row_upper, lp.var_lower, lp.var_upper)]
        h = _vp()
        rc = self.L.vk_unrolled_create(lp.m, lp.n, _ptr(rp, _lp), _ptr(ci, _lp), *[_ptr(a) for a in f],
                                       int(iterations), int(checkpoint_every), C.byref(h))
        if rc != 0:
            raise ValueError(f"vk_unrolled_create failed ({rc}): {self.L.vk_solve_last_error().decode()}")
        self.h = h.value

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc}): {self.L.vk_solve_last_error().decode()}")

    def forward(self):
        x, y = np.zeros(self.n), np.zeros(self.m)
        self._check(self.L.vk_unrolled_forward(self.h, _ptr(x), _ptr(y)), "vk_unrolled_forward")
        return x, y

    def backward(self, grad_x=None, grad_y=None):
        gx = None if grad_x is None else np.ascontiguousarray(grad_x, dtype=np.float64)
        gy = None if grad_y is None else np.ascontiguousarray(grad_y, dtype=np.float64)
        gc, grl, gru = np.zeros(self.n), np.zeros(self.m), np.zeros(self.m)
        self._check(self.L.vk_unrolled_backward(self.h, _ptr(gx), _ptr(gy), _ptr(gc), _ptr(grl), _ptr(gru)),
                    "vk_unrolled_backward")
        return gc, grl, gru

    def set_data(self, objective, row_lower, row_upper):
        """New objective and row bounds for the same matrix (vk_unrolled_set_data)."""
        f = [np.ascontiguousarray(a, dtype=np.float64) for a in (objective, row_lower, row_upper)]
        rc = self.L.vk_unrolled_set_data(self.h, *[_ptr(a) for a in f])
        if rc != 0:
            raise ValueError(f"vk_unrolled_set_data failed ({rc}): {self.L.vk_solve_last_error().decode()}")

    def info(self):
        step, ms = C.c_double(), C.c_double()
        self._check(self.L.vk_unrolled_info(self.h, C.byref(step), C.byref(ms)), "vk_unrolled_info")
        return {"step_size": step.value, "last_device_ms": ms.value}

    def close(self):
        if getattr(self, "h", None):
            self.L.vk_unrolled_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Device:
    """Raw access to the kernel C ABI (include/vikalp_hip.h) with host numpy arrays."""

    def __init__(self):
        self.L = hip_lib()
        check(self.L.vk_hip_init(0), "vk_hip_init")
        self._owned = []

    def alloc(self, count: int, itemsize: int = 8) -> int:
        p = _vp()
        check(self.L.vk_hip_alloc(C.byref(p), max(1, count) * itemsize), "vk_hip_alloc")
        self._owned.append(p.value)
        return p.value

    def put(self, arr: np.ndarray) -> int:
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        p = self.alloc(arr.size)
        check(self.L.vk_hip_memcpy_h2d(p, arr.ctypes.data, arr.nbytes, None), "h2d")
        return p

    def get(self, p: int, count: int) -> np.ndarray:
        out = np.zeros(count)
        check(self.L.vk_hip_memcpy_d2h(out.ctypes.data, p, out.nbytes, None), "d2h")
        return out

    def free_all(self):
        for p in self._owned:
            self.L.vk_hip_free(p)
        self._owned = []

    def csr(self, rows, cols, row_offsets, column_indices, values=None) -> int:
        h = _vp()
        rp = np.ascontiguousarray(row_offsets, dtype=np.int64)
        ci = np.ascontiguousarray(column_indices, dtype=np.int64)
        check(self.L.vk_hip_csr_create(rows, cols, _ptr(rp, _lp), _ptr(ci, _lp), C.byref(h)),
              "vk_hip_csr_create")
        if values is not None:
            v = np.ascontiguousarray(values, dtype=np.float64)
            check(self.L.vk_hip_csr_set_values(h, _ptr(v)), "vk_hip_csr_set_values")
        return h.value

    def tune_col_blocks(self, h, precision=PdhgPrecision.FP64, repeats=10):
        """vk_hip_csr_tune_col_blocks: (fastest block count, {blocks: mean ms})."""
        best = C.c_int()
        ms = np.zeros(5)
        check(self.L.vk_hip_csr_tune_col_blocks(h, int(precision), int(repeats), C.byref(best), _ptr(ms)),
              "vk_hip_csr_tune_col_blocks")
        return best.value, dict(zip((1, 2, 4, 8, 16), ms.tolist()))

    def panel_spmv(self, h, rows, cols, x, transpose=False, precision=PdhgPrecision.FP64, groups=1,
                   threads=256, repeats=0, panel_kb=48):
        """vk_hip_csr_panel_spmv: the (groups, rows) FP64 partial sums of A x
        (A' x when transpose) on the panel layout (panel_kb 48, 24, 16, or 0
        for the automatic size), and the mean ms of `repeats` timed launches
        (None when repeats == 0)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        out_rows = cols if transpose else rows
        part = np.zeros((groups, out_rows))
        ms = C.c_double()
        check(self.L.vk_hip_csr_panel_spmv(h, int(bool(transpose)), int(precision), int(groups), int(panel_kb),
                                           int(threads), _ptr(x), _ptr(part), int(repeats), C.byref(ms)),
              "vk_hip_csr_panel_spmv")
        return part, (ms.value if repeats else None)

    def lp_scale(self, h, lp, ruiz_passes=0, pock_chambolle=False):
        """vk_hip_lp_scale on the matrix h (values set) with lp's vectors:
        a dict shaped like oracle.scale_lp's, without the values."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in
                (lp.objective, lp.var_lower, lp.var_upper, lp.row_lower, lp.row_upper)]
        out = {k: np.zeros(lp.n) for k in ("col_scale", "objective", "var_lower", "var_upper")}
        out.update({k: np.zeros(lp.m) for k in ("row_scale", "row_lower", "row_upper")})
        check(self.L.vk_hip_lp_scale(h, *[_ptr(a) for a in arrs], int(ruiz_passes), int(bool(pock_chambolle)),
                                     _ptr(out["row_scale"]), _ptr(out["col_scale"]), _ptr(out["objective"]),
                                     _ptr(out["var_lower"]), _ptr(out["var_upper"]), _ptr(out["row_lower"]),
                                     _ptr(out["row_upper"])), "vk_hip_lp_scale")
        return out

    def set_tile_geometry(self, h, geometry: int):
        """vk_hip_csr_set_tile_geometry: re-cut A's and A' row tiles (0..3)."""
        check(self.L.vk_hip_csr_set_tile_geometry(h, int(geometry)), "vk_hip_csr_set_tile_geometry")

    def tile_geometry(self, h) -> int:
        g = C.c_int()
        check(self.L.vk_hip_csr_tile_geometry(h, C.byref(g)), "vk_hip_csr_tile_geometry")
        return g.value

    def tune_tiles(self, h, precision=PdhgPrecision.FP64, repeats=10):
        """vk_hip_csr_tune_tiles: (best geometry, mean ms of one A x + A' y
        per geometry); the best is left set on the matrix."""
        best = C.c_int()
        ms = np.zeros(4)
        check(self.L.vk_hip_csr_tune_tiles(h, int(precision), int(repeats), C.byref(best), _ptr(ms)),
              "vk_hip_csr_tune_tiles")
        return best.value, ms.tolist()

    def fail_allocations(self, after=-1) -> int:
        """vk_hip_debug_fail_allocations: the device allocations made since
        the previous call; from now on the first `after` allocations
        succeed and every later one fails (after < 0: none fails)."""
        made = C.c_int64()
        check(self.L.vk_hip_debug_fail_allocations(int(after), C.byref(made)), "vk_hip_debug_fail_allocations")
        return made.value

    def operator_bound(self, h) -> float:
        r = C.c_double()
        check(self.L.vk_hip_csr_operator_bound(h, C.byref(r)), "vk_hip_csr_operator_bound")
        return r.value

    def postsolve_products(self, h, x, y, c):
        """vk_hip_csr_postsolve_products: (A x, c + A'y) on the unscaled values."""
        x, y, c = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, c))
        ax, g = np.zeros(y.size), np.zeros(x.size)
        check(self.L.vk_hip_csr_postsolve_products(h, _ptr(x), _ptr(y), _ptr(c), _ptr(ax), _ptr(g)),
              "vk_hip_csr_postsolve_products")
        return ax, g

    def transpose_values(self, h, nnz):
        """A's current device values in transpose order (column by column,
        rows increasing), with the transpose's offsets and row indices."""
        n = C.c_int64()
        check(self.L.vk_hip_csr_dims(h, None, C.byref(n), None), "vk_hip_csr_dims")
        offs = np.zeros(n.value + 1, np.int32)
        rows = np.zeros(max(1, nnz), np.int32)
        vals = np.zeros(max(1, nnz))
        check(self.L.vk_hip_csr_transpose_export(h, offs.ctypes.data_as(C.POINTER(C.c_int32)),
                                                 rows.ctypes.data_as(C.POINTER(C.c_int32)), _ptr(vals)),
              "vk_hip_csr_transpose_export")
        return vals[:nnz], offs, rows[:nnz]

    def spmv(self, h, transpose, alpha, x, beta, y):
        check(self.L.vk_hip_spmv_csr(h, int(transpose), float(alpha), x, float(beta), y, None), "spmv")
        check(self.L.vk_hip_sync(None), "sync")

    def absmax(self, p, n) -> float:
        r = C.c_double()
        check(self.L.vk_hip_reduce_absmax(n, p, C.byref(r), None), "absmax")
        return r.value

    def dot(self, p, q, n) -> float:
        r = C.c_double()
        check(self.L.vk_hip_reduce_dot(n, p, q, C.byref(r), None), "dot")
        return r.value
