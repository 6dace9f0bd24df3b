"""Primal infeasibility detection (adaptive and Halpern policies) on the GPU.

Not in the reference (its PDHG never stops early on an infeasible LP), so
pinned by the mathematics: a returned status Infeasible must come with row
duals y that are a Farkas ray, checked here independently with numpy on the
original data: g = A'y and D0(y) = sum_j min(g_j l_j, g_j u_j) -
sum_i (max(y_i,0) ru_i - max(-y_i,0) rl_i) must be finite and positive (a
feasible LP has D0 <= 0 for every y).  The infeasible cases are an LP that is
infeasible by construction and a block-angular instance with branching
bounds that HiGHS reports infeasible.  Feasible LPs are never flagged (every
other policy test).  The reference policy keeps its drop-in behaviour and
runs to the iteration limit.
"""
import numpy as np
import pytest

import vikalp_amd as V

pytestmark = pytest.mark.gpu


def farkas_d0(lp, y):
    rp, ci, v = lp.row_offsets, lp.column_indices, lp.values
    g = np.zeros(lp.n)
    for i in range(lp.m):
        np.add.at(g, ci[rp[i]:rp[i + 1]], y[i] * v[rp[i]:rp[i + 1]])
    d0 = 0.0
    for gj, lo, hi in zip(g, lp.var_lower, lp.var_upper):
        if gj == 0.0:
            continue
        b = lo if gj > 0 else hi
        if not np.isfinite(b):
            return -np.inf
        d0 += gj * b
    for yi, lo, hi in zip(y, lp.row_lower, lp.row_upper):
        if yi == 0.0:
            continue
        b = hi if yi > 0 else lo
        if not np.isfinite(b):
            return -np.inf
        d0 -= yi * b
    return d0


def small_infeasible():
    # x1 + x2 = 5 and x1 - x2 >= -1 with 0 <= x <= 2: the equality cannot hold
    return V.LP(2, 2, np.array([0, 2, 4]), np.array([0, 1, 0, 1]), np.array([1.0, 1.0, 1.0, -1.0]),
                np.array([1.0, 2.0]), np.array([5.0, -1.0]), np.array([5.0, np.inf]),
                np.zeros(2), np.full(2, 2.0))


def branched_infeasible():
    # block-angular instance 337 of the batched benchmark (T = 10 periods),
    # which HiGHS reports infeasible
    return V.generate_lp(4, 0, 0, 10, 20261203 + 337)


@pytest.mark.parametrize("precision", [V.PdhgPrecision.FP64, V.PdhgPrecision.Mixed])
@pytest.mark.parametrize("policy", [V.PdhgPolicy.Halpern, V.PdhgPolicy.Adaptive])
@pytest.mark.parametrize("make", [small_infeasible, branched_infeasible])
def test_infeasible_lp_returns_a_farkas_ray(make, policy, precision):
    lp = make()
    print(f"infeasible {make.__name__} {policy.name} {precision.name} m={lp.m} n={lp.n}", flush=True)
    r = V.solve_lp(lp, V.Options(pdhg_policy=policy, pdhg_precision=precision, relative_kkt_tolerance=1e-6,
                                 iteration_limit=100_000))
    assert r.status == "Infeasible", (r.status, r.iterations, r.message)
    assert r.iterations < 100_000
    d0 = farkas_d0(lp, r.y)
    assert np.isfinite(d0) and d0 > 0, d0


def test_reference_policy_does_not_detect():
    lp = small_infeasible()
    r = V.solve_lp(lp, V.Options(iteration_limit=2000))
    assert r.status == "IterationLimit"


@pytest.mark.parametrize("precision", [V.PdhgPrecision.FP64, V.PdhgPrecision.Mixed])
def test_batch_flags_infeasible_instances(precision):
    feasible = [V.generate_lp(1, 40, 90, 3, 3), V.generate_lp(4, 0, 0, 4, 9)]
    lps = [feasible[0], branched_infeasible(), small_infeasible(), feasible[1]]
    res = V.solve_batch(lps, V.Options(pdhg_policy=V.PdhgPolicy.Adaptive, pdhg_precision=precision,
                                       relative_kkt_tolerance=1e-6, iteration_limit=100_000))
    assert [r.status for r in res] == ["Optimal", "Infeasible", "Infeasible", "Optimal"], [r.status for r in res]
    for lp, r in zip(lps[1:3], res[1:3]):
        d0 = farkas_d0(lp, r.y)
        assert np.isfinite(d0) and d0 > 0, d0
        assert r.iterations < 100_000
