"""Scaling, operator bound and postsolve products on the device
(vk_hip_lp_scale, vk_hip_csr_operator_bound, vk_hip_csr_postsolve_products).

Pinned as follows:
- the reference scaling (continuous_common.hpp:133-205) bitwise against the
  oracle's scale_lp, itself pinned against the literal O(n nnz) loop;
- the operator bound (continuous_common.hpp:207-225) bitwise against the
  oracle's operator_bound;
- the Ruiz / Pock-Chambolle passes bitwise against a numpy restatement
  (maxima by np.maximum.at, sums by np.add.at, which adds in index order, so
  row by row as the device does), and by an invariant that a dropped or
  misplaced factor breaks: the scaled values equal r_i a_ij s_j and the
  scaled c, l, u, rl, ru equal c s, l / s, u / s, rl r, ru r to rounding;
- the postsolve products A x and c + A'y (continuous_common.hpp:302-315,
  pdhg.cpp:188-196) bitwise against np.add.at sums of the unscaled values,
  also after a scaling replaced the device values.
Cases: generator families, empty rows and columns, explicit zeros, infinite
bounds, a model with no rows, repeated scaling and new values after one.
"""
import numpy as np
import pytest

import vikalp_amd as V
from vikalp_amd import native as N
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FIELDS = ("row_scale", "col_scale", "objective", "var_lower", "var_upper", "row_lower", "row_upper")


def row_ids(lp):
    return np.repeat(np.arange(lp.m), np.diff(lp.row_offsets))


def transpose_order(lp):
    # CSR positions listed column by column, rows increasing
    return np.lexsort((row_ids(lp), lp.column_indices))


def odd_lp():
    """Row 1 and column 3 are empty, row 2 holds an explicit zero and only
    entries below 1 in magnitude; free and one-sided rows and columns."""
    rp = np.array([0, 3, 3, 5, 7])
    ci = np.array([0, 2, 4, 1, 2, 0, 4])
    v = np.array([4.0, -0.5, 12.0, 0.0, 0.25, -3.0, 1e-3])
    inf = np.inf
    return V.LP(4, 5, rp, ci, v, np.array([1.0, -2.0, 0.5, 3.0, 0.0]),
                np.array([1.0, -inf, -inf, 0.0]), np.array([1.0, 2.0, inf, inf]),
                np.array([0.0, -inf, -1.0, 0.0, -inf]), np.array([10.0, 5.0, inf, inf, inf]))


def cases():
    return [("uniform", V.generate_lp(1, 300, 900, 5, 4)),
            ("zipf", V.generate_lp(2, 2000, 5000, 0, 7)),
            ("block", V.generate_lp(4, 0, 0, 4, 9)),
            ("odd", odd_lp())]


def device_scale(lp, ruiz=0, pc=False, values=None):
    dev = V.Device()
    h = dev.csr(lp.m, lp.n, lp.row_offsets, lp.column_indices, lp.values if values is None else values)
    out = dev.lp_scale(h, lp, ruiz, pc)
    tv, _, _ = dev.transpose_values(h, lp.nnz)
    return dev, h, out, tv


def equilibrate_np(lp, sc, ruiz, pc):
    """The device's Ruiz and Pock-Chambolle passes restated with numpy."""
    rows, cols = row_ids(lp), np.asarray(lp.column_indices)
    out = {k: np.array(sc[k], dtype=np.float64, copy=True) for k in FIELDS}
    vals = np.array(sc["values"], copy=True)

    def apply(rf, cf):
        nonlocal vals
        vals = vals / (rf[rows] * cf[cols])
        out["row_scale"] /= rf
        for k in ("row_lower", "row_upper"):
            f = np.isfinite(out[k])
            out[k][f] = out[k][f] / rf[f]
        out["col_scale"] /= cf
        out["objective"] /= cf
        for k in ("var_lower", "var_upper"):
            f = np.isfinite(out[k])
            out[k][f] = out[k][f] * cf[f]

    for _ in range(ruiz):
        rf, cf = np.zeros(lp.m), np.zeros(lp.n)
        np.maximum.at(rf, rows, np.abs(vals))
        np.maximum.at(cf, cols, np.abs(vals))
        apply(np.where(rf > 0, np.sqrt(rf), 1.0), np.where(cf > 0, np.sqrt(cf), 1.0))
    if pc:
        rf, cf = np.zeros(lp.m), np.zeros(lp.n)
        np.add.at(rf, rows, np.abs(vals))
        np.add.at(cf, cols, np.abs(vals))
        apply(np.where(rf > 0, np.sqrt(rf), 1.0), np.where(cf > 0, np.sqrt(cf), 1.0))
    out["values"] = vals
    return out


@pytest.mark.parametrize("name,lp", cases(), ids=[c[0] for c in cases()])
def test_reference_scaling_matches_oracle_bitwise(name, lp):
    print(f"reference scaling {name} m={lp.m} n={lp.n} nnz={lp.nnz}", flush=True)
    ref = O.scale_lp(lp.as_oracle())
    dev, h, got, tv = device_scale(lp)
    try:
        for k in FIELDS:
            assert np.array_equal(got[k], ref[k]), k
        assert np.array_equal(tv, ref["values"][transpose_order(lp)])
        exp = O.operator_bound(lp.m, lp.n, lp.row_offsets, lp.column_indices, ref["values"])
        assert dev.operator_bound(h) == exp
    finally:
        dev.L.vk_hip_csr_destroy(h)


@pytest.mark.parametrize("ruiz,pc", [(10, True), (10, False), (0, True)])
@pytest.mark.parametrize("name,lp", cases(), ids=[c[0] for c in cases()])
def test_equilibration_matches_restatement_bitwise(name, lp, ruiz, pc):
    print(f"equilibration {name} ruiz={ruiz} pc={pc}", flush=True)
    exp = equilibrate_np(lp, O.scale_lp(lp.as_oracle()), ruiz, pc)
    dev, h, got, tv = device_scale(lp, ruiz, pc)
    try:
        for k in FIELDS:
            assert np.array_equal(got[k], exp[k]), k
        assert np.array_equal(tv, exp["values"][transpose_order(lp)])
        # Factor invariant on the original data, to rounding.
        rows, cols = row_ids(lp), np.asarray(lp.column_indices)
        r, s = got["row_scale"], got["col_scale"]
        np.testing.assert_allclose(exp["values"], r[rows] * lp.values * s[cols], rtol=1e-13, atol=0)
        np.testing.assert_allclose(got["objective"], lp.objective * s, rtol=1e-13, atol=0)
        for k, f in (("var_lower", 1 / s), ("var_upper", 1 / s), ("row_lower", r), ("row_upper", r)):
            base = getattr(lp, k)
            fin = np.isfinite(base)
            assert np.array_equal(np.isfinite(got[k]), fin), k
            np.testing.assert_allclose(got[k][fin], base[fin] * f[fin], rtol=1e-13, atol=0)
        if ruiz == 10 and not pc:
            # Ruiz passes drive every nonzero row and column maximum to 1
            rmax, cmax = np.zeros(lp.m), np.zeros(lp.n)
            np.maximum.at(rmax, rows, np.abs(exp["values"]))
            np.maximum.at(cmax, cols, np.abs(exp["values"]))
            for mx in (rmax[rmax > 0], cmax[cmax > 0]):
                assert np.all((mx > 0.5) & (mx < 2.0)), (mx.min(), mx.max())
    finally:
        dev.L.vk_hip_csr_destroy(h)


def products_np(lp, x, y):
    rows, cols = row_ids(lp), np.asarray(lp.column_indices)
    ax = np.zeros(lp.m)
    np.add.at(ax, rows, lp.values * x[cols])
    g = np.array(lp.objective, dtype=np.float64, copy=True)
    np.add.at(g, cols, y[rows] * lp.values)
    return ax, g


@pytest.mark.parametrize("name,lp", cases(), ids=[c[0] for c in cases()])
def test_postsolve_products_bitwise(name, lp):
    rng = np.random.default_rng(11)
    x, y = rng.normal(size=lp.n), rng.normal(size=lp.m)
    exp_ax, exp_g = products_np(lp, x, y)
    dev = V.Device()
    h = dev.csr(lp.m, lp.n, lp.row_offsets, lp.column_indices, lp.values)
    try:
        ax, g = dev.postsolve_products(h, x, y, lp.objective)
        assert np.array_equal(ax, exp_ax) and np.array_equal(g, exp_g)
        # after a scaling the products still use the unscaled values
        dev.lp_scale(h, lp, 10, True)
        ax, g = dev.postsolve_products(h, x, y, lp.objective)
        assert np.array_equal(ax, exp_ax) and np.array_equal(g, exp_g)
    finally:
        dev.L.vk_hip_csr_destroy(h)


def test_repeated_scaling_and_new_values():
    lp = V.generate_lp(1, 300, 900, 5, 4)
    dev = V.Device()
    h = dev.csr(lp.m, lp.n, lp.row_offsets, lp.column_indices, lp.values)
    try:
        a = dev.lp_scale(h, lp, 10, True)
        ta, _, _ = dev.transpose_values(h, lp.nnz)
        b = dev.lp_scale(h, lp, 10, True)
        tb, _, _ = dev.transpose_values(h, lp.nnz)
        assert all(np.array_equal(a[k], b[k]) for k in FIELDS) and np.array_equal(ta, tb)
        # new values replace the saved originals
        v2 = lp.values * 3.0
        lp2 = V.LP(lp.m, lp.n, lp.row_offsets, lp.column_indices, v2, lp.objective, lp.row_lower,
                   lp.row_upper, lp.var_lower, lp.var_upper)
        N.check(dev.L.vk_hip_csr_set_values(h, v2.ctypes.data_as(N._dp)), "set_values")
        got = dev.lp_scale(h, lp2)
        ref = O.scale_lp(lp2.as_oracle())
        assert all(np.array_equal(got[k], ref[k]) for k in FIELDS)
        tv, _, _ = dev.transpose_values(h, lp.nnz)
        assert np.array_equal(tv, ref["values"][transpose_order(lp)])
        x, y = np.ones(lp.n), np.ones(lp.m)
        ax, g = dev.postsolve_products(h, x, y, lp.objective)
        e_ax, e_g = products_np(lp2, x, y)
        assert np.array_equal(ax, e_ax) and np.array_equal(g, e_g)
    finally:
        dev.L.vk_hip_csr_destroy(h)


def test_model_without_rows():
    # test_pdhg.cpp:28-37: min c x over a box, no constraints
    lp = V.LP(0, 1, np.array([0]), np.array([], np.int64), np.array([]), np.array([2.0]),
              np.array([]), np.array([]), np.array([-1.0]), np.array([3.0]))
    dev = V.Device()
    h = dev.csr(0, 1, lp.row_offsets, lp.column_indices, lp.values)
    try:
        got = dev.lp_scale(h, lp, 10, True)
        assert got["col_scale"][0] == 1.0 and got["objective"][0] == 2.0
        assert dev.operator_bound(h) == 0.0
        ax, g = dev.postsolve_products(h, np.array([1.5]), np.zeros(0), lp.objective)
        assert ax.size == 0 and g[0] == 2.0
    finally:
        dev.L.vk_hip_csr_destroy(h)


def test_scale_requires_values():
    lp = V.generate_lp(1, 40, 90, 3, 3)
    dev = V.Device()
    h = dev.csr(lp.m, lp.n, lp.row_offsets, lp.column_indices)
    try:
        with pytest.raises(RuntimeError):
            dev.lp_scale(h, lp)
    finally:
        dev.L.vk_hip_csr_destroy(h)


def stacked(lps):
    """Block-diagonal stack of LPs: (m, n, rp, ci, values, row_starts, col_starts)."""
    rs, cs = np.zeros(len(lps) + 1, np.int64), np.zeros(len(lps) + 1, np.int64)
    for i, lp in enumerate(lps):
        rs[i + 1], cs[i + 1] = rs[i] + lp.m, cs[i] + lp.n
    rp = [np.zeros(1, np.int64)]
    base = 0
    for lp in lps:
        rp.append(np.asarray(lp.row_offsets[1:], np.int64) + base)
        base += lp.nnz
    ci = np.concatenate([np.asarray(lp.column_indices, np.int64) + cs[i] for i, lp in enumerate(lps)])
    v = np.concatenate([lp.values for lp in lps])
    return int(rs[-1]), int(cs[-1]), np.concatenate(rp), ci, v, rs, cs


def test_block_operator_bounds_match_each_instance():
    lps = [V.generate_lp(1, 300, 900, 5, 4), V.generate_lp(4, 0, 0, 4, 9), V.generate_lp(1, 40, 90, 3, 3)]
    m, n, rp, ci, v, rs, cs = stacked(lps)
    cat = lambda k: np.concatenate([getattr(lp, k) for lp in lps])
    big = V.LP(m, n, rp, ci, v, cat("objective"), cat("row_lower"), cat("row_upper"), cat("var_lower"),
               cat("var_upper"))
    dev = V.Device()
    h = dev.csr(m, n, rp, ci, v)
    try:
        dev.lp_scale(h, big)
        got = np.zeros(len(lps))
        N.check(dev.L.vk_hip_csr_operator_bounds(h, len(lps), rs.ctypes.data_as(N._lp), cs.ctypes.data_as(N._lp),
                                                 got.ctypes.data_as(N._dp)), "operator_bounds")
        for i, lp in enumerate(lps):
            ref = O.scale_lp(lp.as_oracle())
            assert got[i] == O.operator_bound(lp.m, lp.n, lp.row_offsets, lp.column_indices, ref["values"]), i
    finally:
        dev.L.vk_hip_csr_destroy(h)


def test_batch_from_handle_rejects_a_row_leaving_its_instance():
    import ctypes as C
    lps = [V.generate_lp(1, 40, 90, 3, 3), V.generate_lp(1, 40, 90, 3, 5)]
    m, n, rp, ci, v, rs, cs = stacked(lps)
    # the first row of instance 0 gains an entry in instance 1's last column
    ci = ci.copy()
    ci[rp[1] - 1] = n - 1
    dev = V.Device()
    h = dev.csr(m, n, rp, ci, v)
    try:
        p = N.PdhgParams()
        p.policy, p.precision, p.iteration_limit, p.restart_period = 0, 0, 10, 64
        vec = lambda k: np.ascontiguousarray(k, dtype=np.float64)
        ones_n, ones_m, zeros_n = vec(np.ones(n)), vec(np.ones(m)), vec(np.zeros(n))
        out = C.c_void_p()
        rc = dev.L.vk_hip_batch_create_csr(2, rs.ctypes.data_as(N._lp), cs.ctypes.data_as(N._lp), h,
                                           N._ptr(zeros_n), N._ptr(zeros_n), N._ptr(ones_n), N._ptr(ones_m),
                                           N._ptr(ones_m), N._ptr(ones_m), N._ptr(ones_n), None, None, None,
                                           None, None, None, None, C.byref(p), C.byref(out))
        assert rc == -1 and not out.value   # VK_EINVAL
        assert b"leaves its instance" in dev.L.vk_hip_last_error()
    finally:
        dev.L.vk_hip_csr_destroy(h)


@pytest.mark.parametrize("ruiz,pc", [(10, True), (0, False)])
def test_bound_sets_scale_like_their_own_scaling(ruiz, pc):
    # vk_hip_lp_scale_bounds applies the recorded column passes of the last
    # scaling to many bound sets: bitwise what scaling with each set gives.
    lp = odd_lp() if ruiz == 0 else V.generate_lp(1, 300, 900, 5, 4)
    rng = np.random.default_rng(9)
    count = 5
    lo = np.tile(lp.var_lower, (count, 1))
    hi = np.tile(lp.var_upper, (count, 1))
    lo[:, ::3] = rng.uniform(-2, 1, size=lo[:, ::3].shape)
    hi[1, ::4] = np.inf
    dev = V.Device()
    h = dev.csr(lp.m, lp.n, lp.row_offsets, lp.column_indices, lp.values)
    try:
        exp = []
        for k in range(count):
            lk = V.LP(lp.m, lp.n, lp.row_offsets, lp.column_indices, lp.values, lp.objective, lp.row_lower,
                      lp.row_upper, lo[k], hi[k])
            out = dev.lp_scale(h, lk, ruiz, pc)
            exp.append((out["var_lower"], out["var_upper"]))
        dev.lp_scale(h, lp, ruiz, pc)
        lo_c, hi_c = np.ascontiguousarray(lo), np.ascontiguousarray(hi)
        gl, gu = np.zeros_like(lo_c), np.zeros_like(hi_c)
        N.check(dev.L.vk_hip_lp_scale_bounds(h, count, lo_c.ctypes.data_as(N._dp), hi_c.ctypes.data_as(N._dp),
                                             gl.ctypes.data_as(N._dp), gu.ctypes.data_as(N._dp)), "scale_bounds")
        for k in range(count):
            assert np.array_equal(gl[k], exp[k][0]) and np.array_equal(gu[k], exp[k][1]), k
    finally:
        dev.L.vk_hip_csr_destroy(h)
