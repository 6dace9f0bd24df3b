"""Pins the CPU oracle (oracle/pdhg_oracle.c) before anything is checked
against it.  CPU only.

  - ops: the hand-computed values of the reference's tests/test_backend.cpp
    and dense brute force on random CSR;
  - scaling: the reference's literal O(n*nnz) column loop vs the O(nnz)
    pass vs a dense numpy restatement, and the operator bound vs dense sums;
  - BiCGSTAB: the reference's identity / 2x2 / breakdown cases;
  - the solve: bitwise-equal trajectories against the reference's own
    sources (oracle/_ref, when built) and against the committed fixtures
    written from them; analytic optima and failure paths of
    tests/solver/continuous/test_pdhg.cpp;
  - relative KKT: zero at an exact optimum (vertex enumeration on a tiny LP),
    the dual bound matches the LP dual at the HiGHS optimum, and every term
    reacts to a planted error.
"""
import itertools
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
import vikalp_amd as V

INF = math.inf


def csr_from_dense(D):
    rp, ci, v = [0], [], []
    for row in D:
        for j, a in enumerate(row):
            if a != 0.0:
                ci.append(j)
                v.append(a)
        rp.append(len(ci))
    return np.array(rp, np.int64), np.array(ci, np.int64), np.array(v)


def random_csr(rng, m, n, density):
    D = np.where(rng.random((m, n)) < density, rng.normal(size=(m, n)), 0.0)
    return D, *csr_from_dense(D)


# ---------------------------------------------------------------- ops

def test_spmv_hand_values():
    # test_backend.cpp:257-312: A = [1 2 0; 0 3 4]
    rp, ci, v = np.array([0, 2, 4]), np.array([0, 1, 1, 2]), np.array([1.0, 2.0, 3.0, 4.0])
    y = np.zeros(2)
    O.spmv(rp, ci, v, 2, 3, 1.0, np.array([2.0, 1.0, 3.0]), 0.0, y, False)
    assert list(y) == [4.0, 15.0]
    O.spmv(rp, ci, v, 2, 3, 2.0, np.array([2.0, 1.0, 3.0]), 3.0, y, False)
    assert list(y) == [20.0, 75.0]
    yt = np.zeros(3)
    O.spmv(rp, ci, v, 2, 3, 1.0, np.array([2.0, 1.0]), 0.0, yt, True)
    assert list(yt) == [2.0, 7.0, 4.0]
    O.spmv(rp, ci, v, 2, 3, 0.5, np.array([2.0, 1.0]), 2.0, yt, True)
    assert list(yt) == [5.0, 17.5, 10.0]
    # empty row (test_backend.cpp:337-360)
    y3 = np.zeros(3)
    O.spmv(np.array([0, 1, 1, 2]), np.array([0, 1]), np.array([1.0, 5.0]), 3, 2, 1.0,
           np.array([2.0, 3.0]), 0.0, y3, False)
    assert list(y3) == [2.0, 0.0, 15.0]


def test_spmv_matches_dense_brute_force():
    rng = np.random.default_rng(0)
    for m, n, d in [(30, 40, 0.2), (1, 50, 0.9), (50, 1, 0.5), (200, 100, 0.03)]:
        D, rp, ci, v = random_csr(rng, m, n, d)
        x, yN = rng.normal(size=n), rng.normal(size=m)
        exp = 1.5 * D @ x - 0.5 * yN
        O.spmv(rp, ci, v, m, n, 1.5, x, -0.5, yN, False)
        np.testing.assert_allclose(yN, exp, rtol=1e-13, atol=1e-13)
        xt, yT = rng.normal(size=m), rng.normal(size=n)
        exp = 0.7 * D.T @ xt + 2.0 * yT
        O.spmv(rp, ci, v, m, n, 0.7, xt, 2.0, yT, True)
        np.testing.assert_allclose(yT, exp, rtol=1e-13, atol=1e-13)


@pytest.mark.skipif(not O.ref_available(), reason="reference checker not built here")
def test_spmv_bitwise_equals_reference_cpu_backend():
    rng = np.random.default_rng(1)
    D, rp, ci, v = random_csr(rng, 80, 60, 0.1)
    for t in (False, True):
        nx, ny = (80, 60) if t else (60, 80)
        x, y = rng.normal(size=nx), rng.normal(size=ny)
        a, b = y.copy(), y.copy()
        O.spmv(rp, ci, v, 80, 60, 1.3, x, 0.4, a, t)
        O.ref_spmv(rp, ci, v, 80, 60, 1.3, x, 0.4, b, t)
        assert np.array_equal(a, b)


def test_vector_ops():
    L = O.lib()
    x, y = np.array([1.0, -2.0, 3.0]), np.array([4.0, 5.0, -6.0])
    L.orc_axpby(3, 2.5, O._ptr(x), -1.0, O._ptr(y))
    assert list(y) == [-1.5, -10.0, 13.5]
    a, b = np.array([1.0, -2.0, 0.5, 3.0]), np.array([-1.0, 3.0, 2.0, -0.5])
    assert L.orc_dot(4, O._ptr(a), O._ptr(b)) == -7.5
    assert L.orc_norm_inf(5, O._ptr(np.array([1.0, -7.5, 3.0, -2.0, 0.0]))) == 7.5
    assert L.orc_norm_inf(0, O._ptr(np.zeros(1))) == 0.0
    p = np.array([0.5, -1.0, 2.0])
    L.orc_project_box(3, O._ptr(p), O._ptr(np.zeros(3)), O._ptr(np.ones(3)))
    assert list(p) == [0.5, 0.0, 1.0]


def test_validate_pattern():
    L = O.lib()
    ok = L.orc_validate_pattern(2, 3, 3, O._ptr(np.array([0, 2, 4], np.int64), O._lp), 4,
                                O._ptr(np.array([0, 1, 1, 2], np.int64), O._lp))
    assert ok == 0
    bad = L.orc_validate_pattern(1, 1, 2, O._ptr(np.array([0, 2], np.int64), O._lp), 1,
                                 O._ptr(np.array([0], np.int64), O._lp))
    assert bad != 0
    unordered = L.orc_validate_pattern(1, 3, 2, O._ptr(np.array([0, 2], np.int64), O._lp), 2,
                                       O._ptr(np.array([2, 1], np.int64), O._lp))
    assert unordered != 0


def test_bicgstab_reference_cases():
    # test_backend.cpp:380-448
    conv, its, res, x = O.bicgstab([0, 1, 2], [0, 1], [1.0, 1.0], np.array([3.0, 5.0]), 1e-10, 100)
    assert conv and np.allclose(x, [3.0, 5.0], atol=1e-8)
    conv, its, res, x = O.bicgstab([0, 2, 4], [0, 1, 0, 1], [2.0, 1.0, 1.0, 3.0], np.array([4.0, 5.0]), 1e-10, 1000)
    assert conv and np.allclose(x, [1.4, 1.2], atol=1e-7)
    conv, its, res, x = O.bicgstab([0, 1, 2], [0, 1], [1.0, 1.0], np.zeros(2), 1e-10, 100)
    assert conv and its == 0
    conv, its, res, x = O.bicgstab([0, 2, 4], [0, 1, 0, 1], [-3.0, -3.0, -3.0, 0.0], np.array([-3.0, 0.0]), 1e-10, 10)
    assert not conv and its == 1 and res == 3.0


# ---------------------------------------------------------------- scaling

def dense_scaling(D, c, rl, ru, l, u):
    """continuous_common.hpp:133-205 restated densely (a second, independent form)."""
    A = np.abs(D)
    rmax, cmax = A.max(axis=1), A.max(axis=0)
    r = np.where(rmax > 1.0, 1.0 / np.where(rmax > 0, rmax, 1), 1.0)
    s = np.where(cmax > 1.0, 1.0 / np.where(cmax > 0, cmax, 1), 1.0)
    with np.errstate(invalid="ignore"):
        return dict(values=(D * (r[:, None] * s[None, :])), objective=c * s, row_scale=r, col_scale=s,
                    row_lower=np.where(np.isfinite(rl), rl * r, rl), row_upper=np.where(np.isfinite(ru), ru * r, ru),
                    var_lower=np.where(np.isfinite(l), l / s, l), var_upper=np.where(np.isfinite(u), u / s, u))


def test_scaling_literal_fast_and_dense_agree():
    rng = np.random.default_rng(2)
    D = np.where(rng.random((25, 30)) < 0.2, rng.normal(scale=3.0, size=(25, 30)), 0.0)
    D[3, :] = 0.0                      # empty row keeps factor 1
    D[:, 5] *= 1e-3                    # small column keeps factor 1
    rp, ci, v = csr_from_dense(D)
    c, l, u = rng.normal(size=30), -rng.random(30), 1 + rng.random(30)
    l[0], u[1] = -INF, INF
    rl, ru = rng.normal(size=25) - 5, rng.normal(size=25) + 5
    rl[2], ru[4] = -INF, INF
    lp = O.LP(25, 30, rp, ci, v, c, rl, ru, l, u)
    fast, lit = O.scale_lp(lp, literal=False), O.scale_lp(lp, literal=True)
    for k in fast:
        assert np.array_equal(fast[k], lit[k]), k
    dense = dense_scaling(D, c, rl, ru, l, u)
    Dv = dense["values"][D != 0.0]
    assert np.array_equal(fast["values"], Dv)
    for k in ("objective", "row_scale", "col_scale", "row_lower", "row_upper", "var_lower", "var_upper"):
        assert np.array_equal(fast[k], dense[k]), k


def test_operator_bound_dense():
    rng = np.random.default_rng(3)
    D, rp, ci, v = random_csr(rng, 40, 70, 0.1)
    got = O.operator_bound(40, 70, rp, ci, v)
    exp = math.sqrt(np.abs(D).sum(axis=1).max() * np.abs(D).sum(axis=0).max())
    assert abs(got - exp) <= 1e-12 * exp
    assert got >= np.linalg.norm(D, 2) * (1 - 1e-12)   # it bounds ||A||_2


# ---------------------------------------------------------------- the solve

def bounded_linear(c, lo, hi):
    """test_pdhg.cpp:28-37"""
    return O.LP(0, 1, np.array([0]), np.array([], np.int64), np.array([]), np.array([c]),
                np.array([]), np.array([]), np.array([lo]), np.array([hi]))


def test_pdhg_analytic_lps():
    # test_pdhg.cpp:105-127
    r = O.solve_lp(bounded_linear(-1.0, 0.0, 10.0), iteration_limit=5000, primal_tolerance=1e-5, dual_tolerance=1e-5)
    assert r.status == "Optimal" and abs(r.x[0] - 10.0) <= 1e-5 and abs(r.objective_value + 10.0) <= 1e-5
    lp = O.LP(1, 1, np.array([0, 1]), np.array([0]), np.array([1.0]), np.array([-1.0]),
              np.array([-INF]), np.array([3.0]), np.array([0.0]), np.array([20.0]))
    r = O.solve_lp(lp, iteration_limit=5000, primal_tolerance=1e-5, dual_tolerance=1e-5)
    assert r.status == "Optimal" and abs(r.x[0] - 3.0) <= 1e-4
    # LpRelaxationOracle with warm start 2 and upper override 4 (test_pdhg.cpp:129-136)
    r = O.solve_lp(lp, iteration_limit=5000, primal_tolerance=1e-5, dual_tolerance=1e-5,
                   warm_start=np.array([2.0]), upper_override=np.array([4.0]))
    assert r.status == "Optimal" and abs(r.x[0] - 3.0) <= 1e-4


def test_pdhg_failure_paths():
    # test_pdhg.cpp:190-217
    assert O.solve_lp(bounded_linear(1.0, 3.0, 2.0), iteration_limit=20).status == "InvalidModel"
    assert O.solve_lp(bounded_linear(1.0, 0.0, 2.0), iteration_limit=20,
                      warm_start=np.array([INF])).status == "InvalidModel"
    assert O.solve_lp(bounded_linear(1.0, 0.0, 2.0), iteration_limit=-1).status == "InvalidModel"
    r = O.solve_lp(bounded_linear(1.0, 0.0, 2.0), time_limit_seconds=0.0)
    assert r.status == "TimeLimit" and r.iterations == 0
    r = O.solve_lp(bounded_linear(1.0, 0.0, 2.0), iteration_limit=0)
    assert r.status == "IterationLimit" and r.iterations == 0


def _summary_matches(v, s):
    assert float(np.sum(v)) == s["sum"] and float(np.sum(v * v)) == s["sumsq"]
    assert [float(t) for t in v[:4]] == s["head"] and [float(t) for t in v[-4:]] == s["tail"]


def test_pdhg_bitwise_against_reference_fixtures(golden_dir):
    with open(os.path.join(golden_dir, "reference_trajectories.json")) as f:
        cases = json.load(f)
    for name, c in cases.items():
        lp = V.generate_lp(c["family"], c["m"], c["n"], c["k"], c["seed"]).as_oracle()
        assert lp.m == c["m"] and lp.n == c["n"] and len(lp.values) == c["nnz"], name
        r = O.solve_lp(lp, **c["options"])
        assert (r.status, r.iterations) == (c["status"], c["iterations"]), name
        assert r.objective_value == c["objective"], name
        assert r.primal_residual == c["primal_residual"], name
        assert r.stationarity_residual == c["stationarity_residual"], name
        assert r.complementarity_residual == c["complementarity_residual"], name
        _summary_matches(r.x, c["x"])
        _summary_matches(r.y, c["y"])


@pytest.mark.skipif(not O.ref_available(), reason="reference checker not built here")
def test_pdhg_bitwise_against_reference_build():
    for fam, m, n, k, seed, it in [(1, 50, 120, 3, 21, 700), (2, 300, 800, 0, 4, 400)]:
        lp = V.generate_lp(fam, m, n, k, seed).as_oracle()
        a = O.solve_lp(lp, iteration_limit=it)
        b = O.ref_solve_lp(lp, iteration_limit=it)
        assert (a.status, a.iterations) == (b.status, b.iterations)
        assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)
        assert a.objective_value == b.objective_value


# ---------------------------------------------------------------- relative KKT

def tiny_lp():
    # min -x1 - 2 x2  s.t.  x1 + x2 <= 4,  x1 + 3 x2 = 6 (as rl=ru),  0 <= x <= 5
    D = np.array([[1.0, 1.0], [1.0, 3.0]])
    rp, ci, v = csr_from_dense(D)
    return O.LP(2, 2, rp, ci, v, np.array([-1.0, -2.0]), np.array([-INF, 6.0]),
                np.array([4.0, 6.0]), np.zeros(2), np.full(2, 5.0)), D


def test_relative_kkt_zero_at_vertex_enumerated_optimum():
    lp, D = tiny_lp()
    # brute force: enumerate vertices of {x1+x2<=4, x1+3x2=6, 0<=x<=5}
    best = None
    cons = [(np.array([1.0, 1.0]), 4.0), (np.array([1.0, 0.0]), 0.0), (np.array([0.0, 1.0]), 0.0),
            (np.array([1.0, 0.0]), 5.0), (np.array([0.0, 1.0]), 5.0)]
    for a, b in cons:
        M = np.array([[1.0, 3.0], a])
        if abs(np.linalg.det(M)) < 1e-12:
            continue
        x = np.linalg.solve(M, [6.0, b])
        if x.min() >= -1e-12 and x.max() <= 5 + 1e-12 and x.sum() <= 4 + 1e-12:
            val = -x[0] - 2 * x[1]
            if best is None or val < best[0]:
                best = (val, x)
    val, x = best           # x = (3, 1), objective -5
    # duals in the reference sign: c + A'y = 0 on the active rows
    y = np.linalg.solve(D.T, -lp.objective)
    rp_, rd_, gp_, po, do = O.relative_kkt(lp, x, y)
    assert max(rp_, rd_, gp_) < 1e-14
    assert abs(po - val) < 1e-14 and abs(do - val) < 1e-14
    # planted errors move each term
    assert O.relative_kkt(lp, x + [0.1, 0.0], y)[0] > 1e-3        # infeasible
    assert O.relative_kkt(lp, x, y + [0.0, 0.1])[1] > 1e-3        # not stationary
    assert O.relative_kkt(lp, x, -y)[2] > 1e-3                     # wrong sign -> gap
    assert O.relative_kkt(lp, x, y + [-1.0, 0.0])[2] == 1.0        # y<0 on a <= row: D = -inf


def test_relative_kkt_normalisation_at_a_non_optimal_point():
    # Hand values at a point that is neither feasible nor stationary, so every
    # denominator is exercised:
    #   rows:  x1 + x2 = 1 (both bounds finite, equal), x1 - x2 <= 0.5
    #          (one-sided), -2 <= x2 <= 1 (a range: bhat is the larger
    #          magnitude, -2);  0 <= x <= (2, 3);  c = (1, 2)
    #   x = (0.8, 0.5), y = (0.3, 0.2, 0)
    # A x = (1.3, 0.3, 0.5): only row 0 is violated, by 0.3, so
    #   rel_p = 0.3 / (1 + |(1, 0.5, -2)|_2) = 0.3 / (1 + sqrt(5.25)).
    # g = c + A'y = (1.5, 2.1); x - P_X(x - g) = (0.8, 0.5), so
    #   rel_d = sqrt(0.89) / (1 + sqrt(5)).
    # c'x = 1.8; D(y) = sum_j min(g_j l_j, g_j u_j) - (0.3 * 1 + 0.2 * 0.5) = -0.4,
    #   gap = |1.8 + 0.4| / (1 + 1.8 + 0.4) = 2.2 / 3.2.
    D = np.array([[1.0, 1.0], [1.0, -1.0], [0.0, 1.0]])
    rp, ci, v = csr_from_dense(D)
    lp = O.LP(3, 2, rp, ci, v, np.array([1.0, 2.0]), np.array([1.0, -INF, -2.0]), np.array([1.0, 0.5, 1.0]),
              np.zeros(2), np.array([2.0, 3.0]))
    rel_p, rel_d, gap, pobj, dobj = O.relative_kkt(lp, np.array([0.8, 0.5]), np.array([0.3, 0.2, 0.0]))
    assert rel_p == pytest.approx(0.3 / (1.0 + math.sqrt(5.25)), rel=1e-14)
    assert rel_d == pytest.approx(math.sqrt(0.89) / (1.0 + math.sqrt(5.0)), rel=1e-14)
    assert gap == pytest.approx(2.2 / 3.2, rel=1e-14)
    assert pobj == pytest.approx(1.8, rel=1e-15) and dobj == pytest.approx(-0.4, rel=1e-14)


def test_relative_kkt_dual_bound_at_the_highs_optimum():
    # HiGHS (scipy.optimize.linprog, an LP solver independent of the
    # reference and of this repository) solves the seeded 300 x 900 LP; at its
    # primal point and its equality duals (reference sign: y = -marginal, so
    # g = c + A'y are the reduced costs) the oracle's dual objective must be
    # the LP optimum and every relative KKT term must vanish.
    scipy_optimize = pytest.importorskip("scipy.optimize")
    import scipy.sparse as sp
    lp = V.generate_lp(1, 300, 900, 5, 11)
    A = sp.csr_matrix((lp.values, lp.column_indices, lp.row_offsets), shape=(lp.m, lp.n))
    assert np.array_equal(lp.row_lower, lp.row_upper)
    res = scipy_optimize.linprog(lp.objective, A_eq=A, b_eq=lp.row_lower,
                                 bounds=np.column_stack([lp.var_lower, lp.var_upper]), method="highs",
                                 options={"primal_feasibility_tolerance": 1e-10, "dual_feasibility_tolerance": 1e-10})
    assert res.status == 0
    y = -res.eqlin.marginals
    rel_p, rel_d, gap, pobj, dobj = O.relative_kkt(lp.as_oracle(), res.x, y)
    assert max(rel_p, rel_d, gap) <= 1e-9, (rel_p, rel_d, gap)
    assert abs(pobj - res.fun) <= 1e-9 * abs(res.fun) and abs(dobj - res.fun) <= 1e-9 * abs(res.fun)
    # a dual point with a perturbed equality dual is no longer optimal
    assert O.relative_kkt(lp.as_oracle(), res.x, y + 1e-3 * (np.arange(lp.m) == 0))[2] > 1e-8


# ------------------------------------------------------- unrolled PDHG adjoint

def test_unrolled_forward_is_the_reference_trajectory_before_any_step_decision():
    # The reference policy first changes its step after iteration 50
    # (pdhg.cpp:133), so for K <= 50 its iterates are the fixed-step ones.
    for fam, m, n, k, seed, K in [(1, 60, 150, 4, 7, 50), (2, 400, 1000, 0, 5, 37)]:
        lp = V.generate_lp(fam, m, n, k, seed).as_oracle()
        x, y, _ = O.unrolled(lp, K)
        r = O.solve_lp(lp, iteration_limit=K)
        assert r.iterations == K
        assert np.array_equal(x, r.x) and np.array_equal(y, r.y)


def _unrolled_fd(lp, K, xt, directions, seed):
    """Max relative |central difference - adjoint| over random unit
    directions in (c, equality rhs, <= rhs) for L = c'x_K + |x_K - x_t|^2 / 2."""
    x, _, _ = O.unrolled(lp, K)
    _, _, _, gc, grl, gru = O.unrolled(lp, K, grad_x=lp.objective + x - xt)
    gc = gc + x
    eq = np.isfinite(lp.row_lower) & (lp.row_lower == lp.row_upper)
    le = ~eq & np.isfinite(lp.row_upper)

    def loss(dc, db, h):
        rl, ru = lp.row_lower.copy(), lp.row_upper.copy()
        rl[eq] += h * db[eq]
        ru[eq] += h * db[eq]
        ru[le] += h * db[le]
        p = O.LP(lp.m, lp.n, lp.row_offsets, lp.column_indices, lp.values, lp.objective + h * dc, rl, ru,
                 lp.var_lower, lp.var_upper)
        xk, _, _ = O.unrolled(p, K)
        return float(p.objective @ xk + 0.5 * np.sum((xk - xt) ** 2))

    rng = np.random.default_rng(seed)
    worst = 0.0
    for _ in range(directions):
        dc, db = rng.normal(size=lp.n), rng.normal(size=lp.m)
        norm = math.sqrt(dc @ dc + db[eq | le] @ db[eq | le])
        dc, db = dc / norm, db / norm
        fd = (loss(dc, db, 1e-6) - loss(dc, db, -1e-6)) / 2e-6
        an = gc @ dc + (grl + gru)[eq] @ db[eq] + gru[le] @ db[le]
        worst = max(worst, abs(fd - an) / max(abs(an), 1e-12))
    return worst


def test_unrolled_adjoint_matches_central_differences():
    # equality and <= rows (family 2), box constraints active and inactive
    lp = V.generate_lp(2, 400, 1000, 0, 11).as_oracle()
    xt = np.random.default_rng(1).uniform(0, 10, lp.n)
    assert _unrolled_fd(lp, 300, xt, 8, 2) <= 1e-4


def test_unrolled_difference_check_separates_right_from_wrong_gradients():
    # The central-difference bar is tight enough to reject a gradient that is
    # off by one term: the same directional derivative without the explicit
    # c'x_K contribution misses by far more than the tolerance.
    lp = V.generate_lp(2, 400, 1000, 0, 11).as_oracle()
    K = 120
    x, _, _ = O.unrolled(lp, K)
    xt = np.random.default_rng(3).uniform(0, 10, lp.n)
    _, _, _, gc, grl, gru = O.unrolled(lp, K, grad_x=lp.objective + x - xt)
    # without the explicit c'x_K term the directional derivative is wrong
    rng = np.random.default_rng(4)
    dc = rng.normal(size=lp.n)
    dc /= np.linalg.norm(dc)

    def loss(h):
        p = O.LP(lp.m, lp.n, lp.row_offsets, lp.column_indices, lp.values, lp.objective + h * dc,
                 lp.row_lower, lp.row_upper, lp.var_lower, lp.var_upper)
        xk, _, _ = O.unrolled(p, K)
        return float(p.objective @ xk + 0.5 * np.sum((xk - xt) ** 2))

    fd = (loss(1e-6) - loss(-1e-6)) / 2e-6
    assert abs(fd - (gc + x) @ dc) <= 1e-4 * abs(fd)
    assert abs(fd - gc @ dc) > 1e-2 * abs(fd)


def test_unrolled_gradients_are_duality_exact_at_convergence():
    # At an optimum, d obj/dc = x* and d obj/db = -y* (the reference's dual
    # sign, g = c + A'y); K = 200,000 fixed steps converge this LP to 1e-13.
    lp = V.generate_lp(1, 40, 90, 3, 3).as_oracle()
    x, y, _, gc, grl, gru = O.unrolled(lp, 200_000, grad_x=lp.objective)
    rp, rd, gp, _, _ = O.relative_kkt(lp, x, y)
    assert max(rp, rd, gp) <= 1e-10
    assert np.max(np.abs(gc)) <= 1e-8                  # implicit part vanishes: d obj/dc = x*
    assert np.max(np.abs((grl + gru) + y)) <= 1e-8     # d obj/db = -y*


# ---------------------------------------------------------------------- PDQP

from qp_data import CASES as QP_CASES, make_q   # noqa: E402


def _scalar_lp(c, lo, hi):
    return O.LP(0, 1, np.array([0]), np.array([], np.int64), np.array([]), np.array([c]), np.array([]),
                np.array([]), np.array([lo]), np.array([hi]))


def test_pdqp_reference_cases():
    # tests/solver/continuous/test_pdhg.cpp:139-164: min x^2 - 10x on [0, 10]
    lp = _scalar_lp(-10.0, 0.0, 10.0)
    q = (np.array([0, 1]), np.array([0]), np.array([2.0]))
    opts = dict(iteration_limit=1000, primal_tolerance=1e-6, dual_tolerance=1e-6)
    r = O.solve_qp(lp, q, **opts)
    assert r.status == "Optimal" and abs(r.x[0] - 5.0) <= 1e-4 and abs(r.objective_value + 25.0) <= 1e-4
    r = O.solve_qp(lp, q, upper_override=np.array([4.0]), warm_start=np.array([1.0]), **opts)
    assert r.status == "Optimal" and abs(r.x[0] - 4.0) <= 1e-4
    assert O.solve_qp(lp, (np.array([0, 1]), np.array([0]), np.array([-2.0])), **opts).status == "UnsupportedModel"
    # an LP given to the QP solver is the wrong class (pdqp.cpp:315, validate_continuous)
    empty_q = (np.array([0, 0]), np.array([], np.int64), np.array([]))
    assert O.solve_qp(lp, empty_q, **opts).status == "UnsupportedModel"


def test_pdqp_bitwise_against_reference_fixtures(golden_dir):
    with open(os.path.join(golden_dir, "reference_qp_trajectories.json")) as f:
        cases = json.load(f)
    assert set(cases) == {c[0] for c in QP_CASES}
    for name, c in cases.items():
        lp = V.generate_lp(*c["generator"]).as_oracle()
        q = make_q(lp.n, c["q_seed"])
        assert len(q[2]) == c["q_nnz"], name
        r = O.solve_qp(lp, q, **c["options"])
        assert (r.status, r.iterations) == (c["status"], c["iterations"]), name
        assert r.objective_value == c["objective"], name
        assert r.primal_residual == c["primal_residual"], name
        assert r.stationarity_residual == c["stationarity_residual"], name
        assert r.complementarity_residual == c["complementarity_residual"], name
        _summary_matches(r.x, c["x"])
        _summary_matches(r.y, c["y"])


@pytest.mark.skipif(not O.ref_available(), reason="reference checker not built here")
def test_pdqp_bitwise_against_reference_build():
    for gen, qseed, it in [((1, 50, 120, 3, 21), 7, 700), ((4, 0, 0, 4, 9), 8, 300)]:
        lp = V.generate_lp(*gen).as_oracle()
        q = make_q(lp.n, qseed)
        a = O.solve_qp(lp, q, iteration_limit=it)
        b = O.ref_solve_qp(lp, q, iteration_limit=it)
        assert (a.status, a.iterations) == (b.status, b.iterations)
        assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)
        assert a.objective_value == b.objective_value


def test_pdqp_converges_to_a_kkt_point():
    # Independent of the restatement: at the returned point of a converged
    # run the QP optimality conditions hold (primal feasibility, and
    # x = P_X(x - (c + Qx + A'y)) with y of the reference sign).
    name, gen, qseed, opts = QP_CASES[2]
    lp = V.generate_lp(*gen).as_oracle()
    rp, ci, qv = make_q(lp.n, qseed)
    r = O.solve_qp(lp, (rp, ci, qv), iteration_limit=200_000, primal_tolerance=1e-9, dual_tolerance=1e-9)
    # stopped on the scaled metric; the original-coordinate re-check at the
    # same 1e-9 may still flag it (pdqp.cpp:286-291), so only the point is used
    assert r.iterations < 200_000 and r.status in ("Optimal", "NumericalFailure")
    Q = np.zeros((lp.n, lp.n))
    for i in range(lp.n):
        Q[i, ci[rp[i]:rp[i + 1]]] = qv[rp[i]:rp[i + 1]]
    A = np.zeros((lp.m, lp.n))
    for i in range(lp.m):
        A[i, lp.column_indices[lp.row_offsets[i]:lp.row_offsets[i + 1]]] = lp.values[lp.row_offsets[i]:lp.row_offsets[i + 1]]
    g = lp.objective + Q @ r.x + A.T @ r.y
    assert np.max(np.abs(r.x - np.clip(r.x - g, lp.var_lower, lp.var_upper))) <= 1e-7
    assert np.max(np.abs(A @ r.x - np.clip(A @ r.x, lp.row_lower, lp.row_upper))) <= 1e-7
