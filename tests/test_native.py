"""CPU-side checks of the native libraries: they load, they export every
entry point include/*.h declares, the seeded generators produce the stated
LP families, and the host-side validation paths answer without a GPU."""
import math
import os
import re

import numpy as np
import pytest

import vikalp_amd as V
from vikalp_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vk_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,loader", [
    ("vikalp_hip.h", native.hip_lib),
    ("vikalp_solve.h", native.host_lib),
    ("vikalp_lpgen.h", native.host_lib),
])
def test_every_declared_symbol_is_exported(header, loader):
    lib = loader()
    names = declared_functions(header)
    assert names, header
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes tables cover them all
    table = native.HIP_SYMBOLS if loader is native.hip_lib else native.HOST_SYMBOLS
    assert set(names) <= set(table), set(names) - set(table)


def test_uniform_family_structure():
    lp = V.generate_lp(1, 1000, 3000, 5, 20261203)
    assert (lp.m, lp.n, lp.nnz) == (1000, 3000, 15000)
    cols = np.bincount(lp.column_indices, minlength=lp.n)
    assert (cols == 5).all()
    for r in range(lp.m):   # strictly increasing columns per row
        seg = lp.column_indices[lp.row_offsets[r]:lp.row_offsets[r + 1]]
        assert (np.diff(seg) > 0).all()
    A = np.zeros((lp.m, lp.n))
    rows = np.repeat(np.arange(lp.m), np.diff(lp.row_offsets))
    A[rows, lp.column_indices] = lp.values
    assert np.allclose(A @ lp.x0, lp.row_lower) and np.array_equal(lp.row_lower, lp.row_upper)
    s = lp.objective - A.T @ lp.y0
    assert np.count_nonzero(np.abs(s) > 1e-12) <= lp.n // 2 and s.min() > -1e-9 and s.max() < 1.0
    assert ((lp.x0 >= 0.5) & (lp.x0 <= 9.5)).all()
    assert (lp.var_lower == 0).all() and (lp.var_upper == 10).all()


def test_zipf_family_structure():
    lp = V.generate_lp(2, 4000, 10000, 0, 5)
    cols = np.bincount(lp.column_indices, minlength=lp.n)
    assert cols.min() >= 2 and cols.max() <= 200
    # mean of the truncated Zipf(1.5) law on [2,200] is 17.58
    assert abs(cols.mean() - 17.58) < 1.0
    eq = lp.row_lower == lp.row_upper
    assert eq.sum() == round(0.7 * lp.m)
    assert np.isneginf(lp.row_lower[~eq]).all()
    rows = np.repeat(np.arange(lp.m), np.diff(lp.row_offsets))
    ax0 = np.bincount(rows, weights=lp.values * lp.x0[lp.column_indices], minlength=lp.m)
    assert np.allclose(ax0[eq], lp.row_lower[eq])
    slack = lp.row_upper[~eq] - ax0[~eq]
    assert slack.min() >= -1e-9 and slack.max() < 1.0 + 1e-9


def test_block_family_structure():
    T = 5
    lp = V.generate_lp(4, 0, 0, T, 3)
    assert lp.m == 150 * T + 40 * (T - 1) and lp.n == 400 * T
    coupling = lp.values[lp.row_offsets[150 * T]:]
    assert set(np.unique(coupling)) <= {-1.0, 1.0}
    changed = (lp.var_lower != 0) | (lp.var_upper != 10)
    assert 1 <= changed.sum() <= 20


def test_generator_is_deterministic():
    a, b = V.generate_lp(2, 500, 1200, 0, 9), V.generate_lp(2, 500, 1200, 0, 9)
    c = V.generate_lp(2, 500, 1200, 0, 10)
    for k in ("row_offsets", "column_indices", "values", "objective", "row_lower", "row_upper"):
        assert np.array_equal(getattr(a, k), getattr(b, k))
    assert not np.array_equal(a.values[:100], c.values[:100])


def one_variable(c=-1.0, lo=0.0, hi=10.0):
    return V.LP(0, 1, np.array([0]), np.array([], np.int64), np.array([]), np.array([c]),
                np.array([]), np.array([]), np.array([lo]), np.array([hi]))


def test_host_validation_without_device():
    # InvalidModel / UnsupportedModel are decided on the host, before any
    # device call (Solver.cpp:104-111, pdhg.cpp:274-291).
    r = V.solve(one_variable(1.0, 3.0, 2.0))
    assert r.status == "InvalidModel"
    r = V.solve(one_variable(), V.Options(backend=V.BackendPreference.CPU))
    assert r.status == "UnsupportedModel"
    r = V.solve(one_variable(), V.Options(backend=V.BackendPreference.CUDA))
    assert r.status == "UnsupportedModel"
    bad = V.LP(1, 1, np.array([0, 1]), np.array([5]), np.array([1.0]), np.array([1.0]),
               np.array([0.0]), np.array([1.0]), np.array([0.0]), np.array([1.0]))
    r = V.solve(bad)                     # column index out of range (Model.cpp:42-45)
    assert r.status == "InvalidModel" and "out-of-range" in r.message
    with pytest.raises(ValueError):      # unusable arguments never reach the solver
        native._solve(0, V.LP(-1, 1, np.array([0]), np.array([], np.int64), np.array([]),
                              np.array([1.0]), np.array([]), np.array([]), np.array([0.0]),
                              np.array([1.0])), V.Options())


def test_mip_entry_decides_refusals_on_the_host():
    # solve_mip_batched's entry checks (BranchAndBound.cpp:228-243 and the
    # batched engine's own) run before any device call.
    lp = one_variable()
    r, st = V.solve_mip(lp, variable_types=np.array([0], np.int32))
    assert r.status == "UnsupportedModel" and "no integer variables" in r.message
    r, st = V.solve_mip(lp, batch=0, variable_types=np.array([1], np.int32))
    assert r.status == "InvalidModel" and st["rounds"] == 0
    q = (np.array([0, 1]), np.array([0]), np.array([2.0]))
    r, _ = V.solve_mip(lp, variable_types=np.array([1], np.int32), quadratic=q)
    assert r.status == "UnsupportedModel" and "linear" in r.message
    with pytest.raises(ValueError):      # a variable type outside 0..2
        V.solve_mip(lp, variable_types=np.array([3], np.int32))


def test_options_defaults_match_the_contract():
    o = native._SolverOptions()
    native.host_lib().vk_solver_options_default(o)
    d = V.Options()
    for name, _ in native._SolverOptions._fields_:
        if name.startswith("reserved"):
            continue
        got, want = getattr(o, name), getattr(d, name)
        if isinstance(want, float) and math.isinf(want):
            assert math.isinf(got)
        else:
            assert got == int(want) if isinstance(want, bool) else got == want, name
