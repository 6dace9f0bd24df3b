"""Batched persistent PDHG (one launch, many LPs) on the GPU.

Reference policy: every instance of a batch must reproduce the CPU oracle's
solve_lp on that instance alone -- same status and iteration count,
objective to 1e-9 relative, x and y to 1e-8 -- the same bar as the
single-LP tests, so batching changes execution only, not results.
Adaptive policy (FP64 and mixed): each instance reaches relative KKT <= 1e-6,
recomputed by the oracle on the returned point, and agrees with the
single-instance HIP solve's objective to twice the bound a 1e-6 certificate
allows (test_gpu_pdhg.objective_bar).
"""
import math

import numpy as np
import pytest

import vikalp_amd as V
from oracle import oracle as O

from test_gpu_pdhg import objective_bar

pytestmark = pytest.mark.gpu


def instances():
    lps = [V.generate_lp(1, 40, 90, 3, 3), V.generate_lp(4, 0, 0, 4, 9), V.generate_lp(1, 60, 150, 4, 7),
           V.generate_lp(2, 400, 1000, 0, 5), V.generate_lp(4, 0, 0, 5, 11)]
    return lps


def test_reference_policy_batch_matches_oracle_per_instance():
    lps = instances()
    it = 600
    print("batch reference policy", [(lp.m, lp.n, lp.nnz) for lp in lps], flush=True)
    got = V.solve_batch(lps, V.Options(iteration_limit=it))
    for lp, g in zip(lps, got):
        ref = O.solve_lp(lp.as_oracle(), iteration_limit=it)
        assert g.status == ref.status and g.iterations == ref.iterations, (g.status, ref.status, g.iterations)
        assert abs(g.objective_value - ref.objective_value) <= 1e-9 * max(1.0, abs(ref.objective_value))
        assert np.max(np.abs(g.x - ref.x)) <= 1e-8 and np.max(np.abs(g.y - ref.y)) <= 1e-8


def test_reference_policy_batch_terminates_per_instance():
    # a loose tolerance stops some instances early; each must match its own
    # single solve exactly, including the iteration at which it stopped
    lps = [V.generate_lp(1, 40, 90, 3, 3), V.generate_lp(1, 30, 70, 3, 4), bounded_linear(-1.0, 0.0, 10.0)]
    opts = dict(iteration_limit=20000, primal_tolerance=1e-3, dual_tolerance=1e-3)
    got = V.solve_batch(lps, V.Options(**opts))
    for lp, g in zip(lps, got):
        ref = O.solve_lp(lp.as_oracle(), **opts)
        assert (g.status, g.iterations) == (ref.status, ref.iterations)
        assert np.max(np.abs(g.x - ref.x)) <= 1e-8


def bounded_linear(c, lo, hi):
    return V.LP(0, 1, np.array([0]), np.array([], np.int64), np.array([]), np.array([c]),
                np.array([]), np.array([]), np.array([lo]), np.array([hi]))


@pytest.mark.parametrize("precision", [V.PdhgPrecision.FP64, V.PdhgPrecision.Mixed])
def test_adaptive_batch_reaches_relative_kkt(precision):
    lps = [V.generate_lp(4, 0, 0, t, 20 + t) for t in (4, 5, 6, 4)] + [V.generate_lp(1, 300, 900, 5, 11)]
    tol = 1e-6
    opts = V.Options(pdhg_policy=V.PdhgPolicy.Adaptive, pdhg_precision=precision,
                     relative_kkt_tolerance=tol, iteration_limit=100_000)
    print("adaptive batch", precision.name, flush=True)
    got = V.solve_batch(lps, opts)
    for lp, g in zip(lps, got):
        assert g.status == "Optimal", g
        rp, rd, gp, _, _ = O.relative_kkt(lp.as_oracle(), g.x, g.y)
        assert max(rp, rd, gp) <= tol * 1.0001, (rp, rd, gp)
        single = V.solve_lp(lp, V.Options(pdhg_policy=V.PdhgPolicy.Adaptive, relative_kkt_tolerance=tol,
                                          iteration_limit=100_000))
        assert single.status == "Optimal"
        # both meet the certificate: each is within objective_bar of the optimum
        assert abs(g.objective_value - single.objective_value) / max(1.0, abs(single.objective_value)) <= \
            2 * objective_bar(tol)


def test_invalid_instance_does_not_stop_the_batch():
    good = V.generate_lp(1, 40, 90, 3, 3)
    bad = bounded_linear(1.0, 3.0, 2.0)            # lower > upper
    got = V.solve_batch([good, bad, good], V.Options(iteration_limit=300))
    assert got[1].status == "InvalidModel"
    ref = O.solve_lp(good.as_oracle(), iteration_limit=300)
    for g in (got[0], got[2]):
        assert g.status == ref.status and np.max(np.abs(g.x - ref.x)) <= 1e-8


def test_mixed_batch_gives_every_instance_its_own_budget():
    # A feasible but slow instance (block-angular instance 345 of the batched
    # benchmark, relative KKT ~1e-5 after 40,000 iterations) uses its whole
    # budget; an instance with contradictory rows (x0 = 1 and x0 = 2) is
    # proved infeasible early.  The others must still get their FP64
    # refinement and converge.
    bad = V.LP(2, 1, np.array([0, 1, 2]), np.array([0, 0]), np.array([1.0, 1.0]), np.array([1.0]),
               np.array([1.0, 2.0]), np.array([1.0, 2.0]), np.array([0.0]), np.array([10.0]))
    slow = V.generate_lp(4, 0, 0, 7, 20261203 + 345)
    good = [V.generate_lp(1, 40, 90, 3, 3), V.generate_lp(4, 0, 0, 4, 9)]
    opts = V.Options(pdhg_policy=V.PdhgPolicy.Adaptive, pdhg_precision=V.PdhgPrecision.Mixed,
                     relative_kkt_tolerance=1e-6, iteration_limit=20_000)
    got = V.solve_batch([good[0], bad, slow, good[1]], opts)
    assert got[1].status == "Infeasible" and got[1].iterations < 20_000
    assert got[2].status == "IterationLimit" and got[2].iterations == 20_000
    for lp, g in ((good[0], got[0]), (good[1], got[3])):
        assert g.status == "Optimal", g
        rp, rd, gp, _, _ = O.relative_kkt(lp.as_oracle(), g.x, g.y)
        assert max(rp, rd, gp) <= 1e-6 * 1.0001

@pytest.mark.parametrize("precision", [V.PdhgPrecision.FP64, V.PdhgPrecision.Mixed])
def test_tail_finishes_on_the_single_solver(precision):
    # 40 instances: once at most 10 (a quarter) are still running, each
    # remaining one continues on the single-LP solver from its batch iterate
    # (several solvers driven at once from host threads).  Every
    # instance must still meet its certificate (or prove infeasibility),
    # the slow ones must keep their iteration budget, and an instance that
    # hits the limit reports exactly the limit.
    lps = [V.generate_lp(4, 0, 0, 4 + (i % 5), 20261203 + 700 + i) for i in range(36)]
    lps += [V.generate_lp(4, 0, 0, 7, 20261203 + 345), V.generate_lp(4, 0, 0, 30, 20261203 + 39),
            V.generate_lp(2, 400, 1000, 0, 5), V.generate_lp(1, 300, 900, 5, 11)]
    print(f"tail handoff batch of {len(lps)} precision={precision.name}", flush=True)
    limit = 12_000
    opts = V.Options(pdhg_policy=V.PdhgPolicy.Halpern, pdhg_precision=precision, relative_kkt_tolerance=1e-6,
                     iteration_limit=limit)
    got = V.solve_batch(lps, opts)
    from test_gpu_infeasible import farkas_d0
    for lp, g in zip(lps, got):
        assert g.status in ("Optimal", "Infeasible", "IterationLimit"), g.status
        assert g.iterations <= limit
        if g.status == "Optimal":
            rp, rd, gp, _, _ = O.relative_kkt(lp.as_oracle(), g.x, g.y)
            assert max(rp, rd, gp) <= 1e-6 * 1.0001, (rp, rd, gp)
        elif g.status == "Infeasible":
            d0 = farkas_d0(lp, g.y)
            assert np.isfinite(d0) and d0 > 0, d0
        else:
            assert g.iterations == limit
    # the slow feasible instance (instance 345) runs past the batch's tail point
    assert got[36].status in ("Optimal", "IterationLimit") and got[36].iterations > 1000


def test_tail_repeats_bitwise():
    # The tail's solvers run on host threads in whatever order the threads
    # take them; each instance's result must not depend on that order (nor
    # on the batch kernel's 512- or 1024-thread launches), so two runs agree
    # bit for bit.
    lps = [V.generate_lp(4, 0, 0, 4 + (i % 5), 20261203 + 700 + i) for i in range(36)]
    lps += [V.generate_lp(4, 0, 0, 7, 20261203 + 345), V.generate_lp(4, 0, 0, 30, 20261203 + 39)]
    print(f"tail repeat batch of {len(lps)}", flush=True)
    opts = V.Options(pdhg_policy=V.PdhgPolicy.Halpern, relative_kkt_tolerance=1e-6, iteration_limit=8_000)
    a = V.solve_batch(lps, opts)
    b = V.solve_batch(lps, opts)
    for ra, rb in zip(a, b):
        assert (ra.status, ra.iterations) == (rb.status, rb.iterations)
        assert np.array_equal(ra.x, rb.x) and np.array_equal(ra.y, rb.y)
