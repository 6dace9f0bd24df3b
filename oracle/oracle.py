"""ctypes binding of the CPU checker libraries — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only to check or time against; the product (vikalp_amd/)
never imports it.

liboracle.so            plain-C restatement of the reference's CPU PDHG path
                        (pdhg_oracle.c; built by oracle/Makefile)
_ref/libvikalp_ref.so   the reference's own sources compiled by the same
                        Makefile (present only where the reference tree was
                        available at build time)
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libvikalp_ref.so")

# SolveStatus order of include/vikalp/contracts/SolveResult.hpp:11-24
STATUS_NAMES = [
    "NotSolved", "Optimal", "LocallyOptimal", "Feasible", "Infeasible",
    "Unbounded", "IterationLimit", "NodeLimit", "TimeLimit",
    "NumericalFailure", "UnsupportedModel", "InvalidModel",
]

_dp = C.POINTER(C.c_double)
_lp = C.POINTER(C.c_int64)


class _Options(C.Structure):
    _fields_ = [
        ("iteration_limit", C.c_int64),
        ("time_limit_seconds", C.c_double),
        ("primal_tolerance", C.c_double),
        ("dual_tolerance", C.c_double),
        ("relative_kkt_tolerance", C.c_double),
        ("kkt_check_period", C.c_int64),
    ]


class _Result(C.Structure):
    _fields_ = [
        ("status", C.c_int),
        ("iterations", C.c_int64),
        ("objective_value", C.c_double),
        ("primal_residual", C.c_double),
        ("stationarity_residual", C.c_double),
        ("complementarity_residual", C.c_double),
        ("absolute_gap", C.c_double),
        ("relative_gap", C.c_double),
        ("rel_primal", C.c_double),
        ("rel_dual", C.c_double),
        ("rel_gap", C.c_double),
        ("rel_kkt", C.c_double),
        ("primal_objective", C.c_double),
        ("dual_objective", C.c_double),
        ("tau", C.c_double),
        ("sigma", C.c_double),
        ("last_metric", C.c_double),
        ("solve_seconds", C.c_double),
    ]


class _Lp(C.Structure):
    _fields_ = [
        ("m", C.c_int64), ("n", C.c_int64),
        ("row_offsets", _lp), ("column_indices", _lp), ("values", _dp),
        ("objective", _dp), ("objective_offset", C.c_double),
        ("row_lower", _dp), ("row_upper", _dp),
        ("var_lower", _dp), ("var_upper", _dp),
    ]


def _ptr(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} is not built (make -C oracle oracle)")
        L = C.CDLL(ORACLE_SO)
        L.orc_solve_lp.argtypes = [C.POINTER(_Lp), _dp, _dp, _dp, C.POINTER(_Options), _dp, _dp, C.POINTER(_Result)]
        L.orc_solve_lp.restype = C.c_int
        L.orc_spmv.argtypes = [C.c_int64, C.c_int64, _lp, _lp, _dp, C.c_double, _dp, C.c_double, _dp, C.c_int]
        L.orc_axpby.argtypes = [C.c_int64, C.c_double, _dp, C.c_double, _dp]
        L.orc_dot.argtypes = [C.c_int64, _dp, _dp]
        L.orc_dot.restype = C.c_double
        L.orc_norm_inf.argtypes = [C.c_int64, _dp]
        L.orc_norm_inf.restype = C.c_double
        L.orc_project_box.argtypes = [C.c_int64, _dp, _dp, _dp]
        L.orc_bicgstab.argtypes = [C.c_int64, _lp, _lp, _dp, _dp, _dp, C.c_double, C.c_int64, _lp, _dp]
        L.orc_bicgstab.restype = C.c_int
        L.orc_validate_pattern.argtypes = [C.c_int64, C.c_int64, C.c_int64, _lp, C.c_int64, _lp]
        L.orc_validate_pattern.restype = C.c_int
        L.orc_scale_lp.argtypes = [C.POINTER(_Lp), _dp, _dp, C.c_int] + [_dp] * 8
        L.orc_operator_bound.argtypes = [C.c_int64, C.c_int64, _lp, _lp, _dp]
        L.orc_operator_bound.restype = C.c_double
        L.orc_relative_kkt.argtypes = [C.POINTER(_Lp), _dp, _dp] + [_dp] * 5
        L.orc_time_iterations.argtypes = [C.POINTER(_Lp), C.c_int64]
        L.orc_time_iterations.restype = C.c_double
        L.orc_unrolled.argtypes = [C.POINTER(_Lp), C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_solve_qp.argtypes = [C.POINTER(_Lp), _lp, _lp, _dp, _dp, _dp, _dp, C.POINTER(_Options),
                                   _dp, _dp, C.POINTER(_Result)]
        L.orc_solve_qp.restype = C.c_int
        L.orc_unrolled.restype = C.c_int
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"{REF_SO} is not built (needs the reference tree)")
        L = C.CDLL(REF_SO)
        L.ref_solve_lp.argtypes = [C.c_int64, C.c_int64, _lp, _lp, _dp, _dp, C.c_double,
                                   _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int64,
                                   C.c_double, C.c_double, C.c_double, _dp, _dp, _dp]
        L.ref_solve_lp.restype = C.c_int
        L.ref_spmv.argtypes = [C.c_int64, C.c_int64, _lp, _lp, _dp, C.c_double, _dp, C.c_double, _dp, C.c_int]
        L.ref_solve_qp.argtypes = [C.c_int64, C.c_int64, _lp, _lp, _dp, _lp, _lp, _dp, _dp, C.c_double,
                                   _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int64,
                                   C.c_double, C.c_double, C.c_double, _dp, _dp, _dp]
        L.ref_solve_qp.restype = C.c_int
        _ref = L
    return _ref


@dataclass
class LP:
    """An LP in the reference's canonical form (CSR with int64 indices)."""
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

    def _c(self):
        self._keep = [np.ascontiguousarray(a) for a in (
            self.row_offsets.astype(np.int64), self.column_indices.astype(np.int64),
            self.values.astype(np.float64), self.objective.astype(np.float64),
            self.row_lower.astype(np.float64), self.row_upper.astype(np.float64),
            self.var_lower.astype(np.float64), self.var_upper.astype(np.float64))]
        k = self._keep
        return _Lp(self.m, self.n, _ptr(k[0], _lp), _ptr(k[1], _lp), _ptr(k[2]),
                   _ptr(k[3]), float(self.objective_offset), _ptr(k[4]), _ptr(k[5]),
                   _ptr(k[6]), _ptr(k[7]))


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
    x: np.ndarray
    y: np.ndarray
    rel_primal: float = math.inf
    rel_dual: float = math.inf
    rel_gap: float = math.inf
    rel_kkt: float = math.inf
    primal_objective: float = math.nan
    dual_objective: float = math.nan
    tau: float = math.nan
    sigma: float = math.nan
    last_metric: float = math.nan
    solve_seconds: float = 0.0


def solve_lp(lp: LP, *, iteration_limit=100_000, time_limit_seconds=math.inf,
             primal_tolerance=1e-7, dual_tolerance=1e-7, relative_kkt_tolerance=0.0,
             kkt_check_period=0, warm_start=None, lower_override=None,
             upper_override=None) -> Result:
    """The restated CPU solve (pdhg.cpp:269-293 + :20-265)."""
    L = lib()
    opt = _Options(int(iteration_limit), float(time_limit_seconds), float(primal_tolerance),
                   float(dual_tolerance), float(relative_kkt_tolerance), int(kkt_check_period))
    res = _Result()
    x = np.zeros(lp.n)
    y = np.zeros(lp.m)
    arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
            for a in (lower_override, upper_override, warm_start)]
    c_lp = lp._c()
    L.orc_solve_lp(C.byref(c_lp), _ptr(arrs[0]), _ptr(arrs[1]), _ptr(arrs[2]),
                   C.byref(opt), _ptr(x), _ptr(y), C.byref(res))
    return Result(STATUS_NAMES[res.status], res.iterations, res.objective_value,
                  res.primal_residual, res.stationarity_residual,
                  res.complementarity_residual, res.absolute_gap, res.relative_gap,
                  x, y, res.rel_primal, res.rel_dual, res.rel_gap, res.rel_kkt,
                  res.primal_objective, res.dual_objective, res.tau, res.sigma,
                  res.last_metric, res.solve_seconds)


def ref_solve_lp(lp: LP, *, iteration_limit=100_000, time_limit_seconds=math.inf,
                 primal_tolerance=1e-7, dual_tolerance=1e-7, warm_start=None,
                 lower_override=None, upper_override=None) -> Result:
    """The reference's own solve_lp on its CpuBackend (oracle/_ref)."""
    L = ref()
    lp._c()
    k = lp._keep
    x = np.zeros(lp.n)
    y = np.zeros(lp.m)
    stats = np.zeros(8)
    arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
            for a in (warm_start, lower_override, upper_override)]
    L.ref_solve_lp(lp.m, lp.n, _ptr(k[0], _lp), _ptr(k[1], _lp), _ptr(k[2]), _ptr(k[3]),
                   float(lp.objective_offset), _ptr(k[4]), _ptr(k[5]), _ptr(k[6]),
                   _ptr(k[7]), _ptr(arrs[0]), _ptr(arrs[1]), _ptr(arrs[2]),
                   int(iteration_limit), float(time_limit_seconds),
                   float(primal_tolerance), float(dual_tolerance), _ptr(x), _ptr(y),
                   _ptr(stats))
    return Result(STATUS_NAMES[int(stats[0])], int(stats[1]), stats[2], stats[3],
                  stats[4], stats[5], stats[6], stats[7], x, y)


def _q_arrays(q):
    qrp, qci, qv = q
    return (np.ascontiguousarray(qrp, dtype=np.int64), np.ascontiguousarray(qci, dtype=np.int64),
            np.ascontiguousarray(qv, dtype=np.float64))


def solve_qp(lp: LP, q, *, iteration_limit=100_000, time_limit_seconds=math.inf,
             primal_tolerance=1e-7, dual_tolerance=1e-7, warm_start=None, lower_override=None,
             upper_override=None) -> Result:
    """The restated PDQP solve (pdqp.cpp:303-331 + :20-300); q = (row_offsets,
    column_indices, values) of the symmetric n x n Q."""
    L = lib()
    qrp, qci, qv = _q_arrays(q)
    opt = _Options(int(iteration_limit), float(time_limit_seconds), float(primal_tolerance),
                   float(dual_tolerance), 0.0, 0)
    res = _Result()
    x = np.zeros(lp.n)
    y = np.zeros(lp.m)
    arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
            for a in (lower_override, upper_override, warm_start)]
    c_lp = lp._c()
    L.orc_solve_qp(C.byref(c_lp), _ptr(qrp, _lp), _ptr(qci, _lp), _ptr(qv), _ptr(arrs[0]), _ptr(arrs[1]),
                   _ptr(arrs[2]), C.byref(opt), _ptr(x), _ptr(y), C.byref(res))
    return Result(STATUS_NAMES[res.status], res.iterations, res.objective_value,
                  res.primal_residual, res.stationarity_residual,
                  res.complementarity_residual, res.absolute_gap, res.relative_gap,
                  x, y, res.rel_primal, res.rel_dual, res.rel_gap, res.rel_kkt,
                  res.primal_objective, res.dual_objective, res.tau, res.sigma,
                  res.last_metric, res.solve_seconds)


def ref_solve_qp(lp: LP, q, *, iteration_limit=100_000, time_limit_seconds=math.inf,
                 primal_tolerance=1e-7, dual_tolerance=1e-7, warm_start=None,
                 lower_override=None, upper_override=None) -> Result:
    """The reference's own solve_qp on its CpuBackend (oracle/_ref)."""
    L = ref()
    lp._c()
    k = lp._keep
    qrp, qci, qv = _q_arrays(q)
    x = np.zeros(lp.n)
    y = np.zeros(lp.m)
    stats = np.zeros(8)
    arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
            for a in (warm_start, lower_override, upper_override)]
    L.ref_solve_qp(lp.m, lp.n, _ptr(k[0], _lp), _ptr(k[1], _lp), _ptr(k[2]), _ptr(qrp, _lp), _ptr(qci, _lp),
                   _ptr(qv), _ptr(k[3]), float(lp.objective_offset), _ptr(k[4]), _ptr(k[5]), _ptr(k[6]),
                   _ptr(k[7]), _ptr(arrs[0]), _ptr(arrs[1]), _ptr(arrs[2]),
                   int(iteration_limit), float(time_limit_seconds),
                   float(primal_tolerance), float(dual_tolerance), _ptr(x), _ptr(y), _ptr(stats))
    return Result(STATUS_NAMES[int(stats[0])], int(stats[1]), stats[2], stats[3],
                  stats[4], stats[5], stats[6], stats[7], x, y)


def spmv(rp, ci, v, rows, cols, alpha, x, beta, y, transpose):
    """CpuBackend::multiply restated (cpu_ops.cpp:393-460); y is updated in place."""
    rp = np.ascontiguousarray(rp, dtype=np.int64)
    ci = np.ascontiguousarray(ci, dtype=np.int64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    lib().orc_spmv(rows, cols, _ptr(rp, _lp), _ptr(ci, _lp), _ptr(v), float(alpha),
                   _ptr(x), float(beta), _ptr(y), int(bool(transpose)))
    return y


def ref_spmv(rp, ci, v, rows, cols, alpha, x, beta, y, transpose):
    rp = np.ascontiguousarray(rp, dtype=np.int64)
    ci = np.ascontiguousarray(ci, dtype=np.int64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    ref().ref_spmv(rows, cols, _ptr(rp, _lp), _ptr(ci, _lp), _ptr(v), float(alpha),
                   _ptr(x), float(beta), _ptr(y), int(bool(transpose)))
    return y


def scale_lp(lp: LP, literal: bool = False):
    """Scaled data of continuous_common.hpp:133-205 as a dict of arrays."""
    L = lib()
    out = {k: np.zeros(lp.n) for k in ("objective", "var_lower", "var_upper", "col_scale")}
    out.update({k: np.zeros(lp.m) for k in ("row_lower", "row_upper", "row_scale")})
    out["values"] = np.zeros(len(lp.values))
    c_lp = lp._c()
    L.orc_scale_lp(C.byref(c_lp), _ptr(lp._keep[6]), _ptr(lp._keep[7]),
                   int(literal), _ptr(out["values"]), _ptr(out["objective"]),
                   _ptr(out["row_lower"]), _ptr(out["row_upper"]), _ptr(out["var_lower"]),
                   _ptr(out["var_upper"]), _ptr(out["row_scale"]), _ptr(out["col_scale"]))
    return out


def operator_bound(m, n, rp, ci, v) -> float:
    rp = np.ascontiguousarray(rp, dtype=np.int64)
    ci = np.ascontiguousarray(ci, dtype=np.int64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    return lib().orc_operator_bound(m, n, _ptr(rp, _lp), _ptr(ci, _lp), _ptr(v))


def relative_kkt(lp: LP, x, y):
    """(rel_primal, rel_dual, rel_gap, primal_obj, dual_obj) in original coordinates."""
    out = [C.c_double() for _ in range(5)]
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    c_lp = lp._c()
    lib().orc_relative_kkt(C.byref(c_lp), _ptr(x), _ptr(y), *[C.byref(o) for o in out])
    return tuple(o.value for o in out)


def unrolled(lp: LP, iterations: int, grad_x=None, grad_y=None):
    """Unrolled fixed-step PDHG (orc_unrolled): returns (x_K, y_K, step) and,
    when grad_x / grad_y (dL/dx_K, dL/dy_K) are given, also
    (grad_c, grad_row_lower, grad_row_upper) through the iterations."""
    x = np.zeros(lp.n)
    y = np.zeros(lp.m)
    step = C.c_double()
    want = grad_x is not None or grad_y is not None
    gc, grl, gru = (np.zeros(lp.n), np.zeros(lp.m), np.zeros(lp.m)) if want else (None, None, None)
    gx = None if grad_x is None else np.ascontiguousarray(grad_x, dtype=np.float64)
    gy = None if grad_y is None else np.ascontiguousarray(grad_y, dtype=np.float64)
    c_lp = lp._c()
    rc = lib().orc_unrolled(C.byref(c_lp), int(iterations), _ptr(gx), _ptr(gy), _ptr(x), _ptr(y),
                            _ptr(gc), _ptr(grl), _ptr(gru), C.byref(step))
    if rc != 0:
        raise ValueError("orc_unrolled: bad input")
    return (x, y, step.value) + ((gc, grl, gru) if want else ())


def time_iterations(lp: LP, iterations: int) -> float:
    """Wall seconds for `iterations` loop bodies on one host core (setup excluded)."""
    c_lp = lp._c()
    return lib().orc_time_iterations(C.byref(c_lp), int(iterations))


def bicgstab(rp, ci, v, b, tol, max_iter):
    n = len(b)
    rp = np.ascontiguousarray(rp, dtype=np.int64)
    ci = np.ascontiguousarray(ci, dtype=np.int64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(n)
    its = C.c_int64()
    res = C.c_double()
    conv = lib().orc_bicgstab(n, _ptr(rp, _lp), _ptr(ci, _lp), _ptr(v), _ptr(b), _ptr(x),
                              float(tol), int(max_iter), C.byref(its), C.byref(res))
    return bool(conv), its.value, res.value, x
