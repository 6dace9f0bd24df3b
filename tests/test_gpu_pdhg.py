"""PDHG on the GPU against the CPU oracle (parity tests proper).

Reference policy (the drop-in): the fused device iteration and the op-by-op
HipBackend path must reproduce the oracle's result on the same seeded LP:
same status and iteration count, objective within 1e-9 relative, and x, y
within 1e-8 (FP64 rounding differs: fused multiply-adds, the A(2x+ - x)
linearity and tile-order sums; the iteration is non-expansive so the
difference stays at rounding level over these budgets).

Adaptive policy and mixed precision (not in the reference): pinned by the
independent HiGHS optimum (tests/golden/highs_optima.json) and by the
relative-KKT certificate recomputed by the oracle on the returned point.
"""
import json
import math
import os

import numpy as np
import pytest

import vikalp_amd as V
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def close(got, ref, tol=1e-8):
    assert got.status == ref.status, (got.status, ref.status)
    assert got.iterations == ref.iterations, (got.iterations, ref.iterations)
    assert abs(got.objective_value - ref.objective_value) <= 1e-9 * max(1.0, abs(ref.objective_value))
    dx = np.max(np.abs(got.x - ref.x)) if len(ref.x) else 0.0
    dy = np.max(np.abs(got.y - ref.y)) if len(ref.y) else 0.0
    assert dx <= tol and dy <= tol, (dx, dy)
    for k in ("primal_residual", "stationarity_residual"):
        a, b = getattr(got, k), getattr(ref, k)
        assert abs(a - b) <= 1e-7 * max(1.0, abs(b)), (k, a, b)


TRAJ = json.load(open(os.path.join(GOLDEN, "reference_trajectories.json")))


@pytest.mark.parametrize("name", sorted(TRAJ))
@pytest.mark.parametrize("fused", [True, False])
def test_reference_policy_matches_oracle(name, fused):
    c = TRAJ[name]
    print(f"case {name} fused={fused}", flush=True)
    lp = V.generate_lp(c["family"], c["m"], c["n"], c["k"], c["seed"])
    opts = dict(c["options"])
    got = V.solve_lp(lp, V.Options(fused_iterations=fused, **opts))
    ref = O.solve_lp(lp.as_oracle(), **opts)
    close(got, ref)
    assert got.status == c["status"] and got.iterations == c["iterations"]


def test_graphs_and_eager_launches_agree_bitwise():
    lp = V.generate_lp(2, 400, 1000, 0, 5)
    a = V.solve_lp(lp, V.Options(iteration_limit=777, use_graphs=True))
    b = V.solve_lp(lp, V.Options(iteration_limit=777, use_graphs=False))
    c = V.solve_lp(lp, V.Options(iteration_limit=777, use_graphs=True))
    assert a.iterations == b.iterations == 777
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)
    assert np.array_equal(a.x, c.x) and np.array_equal(a.y, c.y)   # run-to-run determinism


def bounded_linear(c, lo, hi):
    return V.LP(0, 1, np.array([0]), np.array([], np.int64), np.array([]), np.array([c]),
                np.array([]), np.array([]), np.array([lo]), np.array([hi]))


@pytest.mark.parametrize("fused", [True, False])
def test_analytic_lps(fused):
    # tests/solver/continuous/test_pdhg.cpp:105-136 on the HIP backend
    o = V.Options(iteration_limit=5000, primal_tolerance=1e-5, dual_tolerance=1e-5, fused_iterations=fused)
    r = V.solve_lp(bounded_linear(-1.0, 0.0, 10.0), o)
    assert r.status == "Optimal" and abs(r.x[0] - 10.0) <= 1e-5 and abs(r.objective_value + 10.0) <= 1e-5
    lp = V.LP(1, 1, np.array([0, 1]), np.array([0]), np.array([1.0]), np.array([-1.0]),
              np.array([-math.inf]), np.array([3.0]), np.array([0.0]), np.array([20.0]))
    r = V.solve_lp(lp, o)
    assert r.status == "Optimal" and abs(r.x[0] - 3.0) <= 1e-4
    r = V.solve_lp(lp, o, warm_start=np.array([2.0]), upper_override=np.array([4.0]))
    assert r.status == "Optimal" and abs(r.x[0] - 3.0) <= 1e-4


def test_router_and_failure_paths():
    # test_solver.cpp:81-87 (label) and test_pdhg.cpp:190-217 (failures)
    r = V.solve(bounded_linear(-1.0, 0.0, 10.0), V.Options(iteration_limit=10_000, primal_tolerance=1e-5,
                                                           dual_tolerance=1e-5, backend=V.BackendPreference.HIP))
    assert r.status == "Optimal" and r.backend == "hip" and r.solver == "vikalp-continuous"
    assert abs(r.x[0] - 10.0) <= 1e-4
    assert V.solve_lp(bounded_linear(1.0, 3.0, 2.0)).status == "InvalidModel"
    assert V.solve_lp(bounded_linear(1.0, 0.0, 2.0), warm_start=np.array([math.inf])).status == "InvalidModel"
    assert V.solve_lp(bounded_linear(1.0, 0.0, 2.0), V.Options(iteration_limit=-1)).status == "InvalidModel"
    lp = V.generate_lp(1, 60, 150, 4, 7)
    r = V.solve_lp(lp, V.Options(time_limit_seconds=0.0))
    assert r.status == "TimeLimit" and r.iterations == 0
    ref = O.solve_lp(lp.as_oracle(), time_limit_seconds=0.0)
    assert np.allclose(r.x, ref.x, atol=0)                          # the projected start
    r = V.solve_lp(lp, V.Options(iteration_limit=0))
    assert r.status == "IterationLimit" and r.iterations == 0


HIGHS = json.load(open(os.path.join(GOLDEN, "highs_optima.json")))


def objective_bar(tol, obj=None):
    """Relative objective error against HiGHS accepted at a relative-KKT
    target `tol`.  The certificate bounds the gap term by tol, but its
    primal term is relative to 1 + |bhat|_2, so the objective of a point at
    1e-6 can be off by ~1e-5 (measured 8.5e-6 on the 1000 x 3000 LP); the
    1e-6 objective bar is met from a target of 1e-8 on (and is checked
    there)."""
    return 1e-5 if tol >= 1e-6 else 1e-6


@pytest.mark.parametrize("tol", [1e-6, 1e-8])
@pytest.mark.parametrize("precision", [V.PdhgPrecision.FP64, V.PdhgPrecision.Mixed])
@pytest.mark.parametrize("name", ["uniform_300x900_k5", "uniform_1000x3000_k5"])
def test_adaptive_policy_reaches_relative_kkt(name, precision, tol):
    g = HIGHS[name]
    print(f"adaptive {name} precision={precision.name} tol={tol}", flush=True)
    lp = V.generate_lp(g["family"], g["m"], g["n"], g["k"], g["seed"])
    r = V.solve_lp(lp, V.Options(pdhg_policy=V.PdhgPolicy.Adaptive, pdhg_precision=precision,
                                 relative_kkt_tolerance=tol, iteration_limit=200_000))
    assert r.status == "Optimal", r
    # certificate recomputed independently by the oracle on the returned point
    rp, rd, gp, po, do = O.relative_kkt(lp.as_oracle(), r.x, r.y)
    assert max(rp, rd, gp) <= tol * 1.0001, (rp, rd, gp)
    assert abs(max(rp, rd, gp) - r.kkt_relative) <= 1e-9
    # objective against the independent HiGHS optimum
    rel_err = abs(r.objective_value - g["objective"]) / max(1.0, abs(g["objective"]))
    print(f"objective relative error vs HiGHS: {rel_err:.3e}", flush=True)
    assert rel_err <= objective_bar(tol), rel_err
    if precision == V.PdhgPrecision.Mixed:
        assert r.low_precision_iterations > 0
