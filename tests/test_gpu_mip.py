"""Branch-and-bound over batches of node relaxations (solver::solve_mip_batched)
on the GPU.

The engine follows the reference's rules (src/solver/mip/BranchAndBound.cpp)
but evaluates open nodes in rounds through the batched node solver, which the
reference does not have, so it is pinned against the mathematics:
- the optimum of seeded MILPs against scipy's HiGHS MILP solver
  (scipy.optimize.milp, an independent branch-and-cut), with the returned
  incumbent checked integral and feasible with numpy;
- serial (batch 1, the reference's node order) and batched searches reaching
  the same optimum, the batched one in fewer relaxation rounds;
- an integer-infeasible MILP whose LP relaxation is feasible, the node limit,
  and the refused models;
- an MIQP (QP relaxations) against the optimum found by enumerating its
  integer assignments, with the continuous variable in closed form (skipped:
  the MIQP path is switched off), and the MIQP refusal.
"""
import numpy as np
import pytest
import scipy.optimize as so
import scipy.sparse as sp

import vikalp_amd as V

pytestmark = pytest.mark.gpu


def random_milp(m, n, seed, integer_share=0.7, upper=6):
    """Rows A x <= b with small integer coefficients and b = A x* + slack for
    an integer point x* in the box [0, upper], so the MILP is feasible; a
    negative objective pushes the search against the rows."""
    rng = np.random.default_rng(seed)
    dense = np.zeros((m, n))
    for i in range(m):
        cols = rng.choice(n, size=min(n, 6), replace=False)
        dense[i, cols] = rng.integers(1, 9, size=cols.size)
    xstar = rng.integers(0, upper + 1, size=n).astype(float)
    b = dense @ xstar + rng.integers(0, 7, size=m)
    A = sp.csr_matrix(dense)
    types = (rng.random(n) < integer_share).astype(np.int32)
    c = -rng.uniform(1.0, 10.0, size=n)
    lp = V.LP(m, n, A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data.astype(np.float64), c,
              np.full(m, -np.inf), b.astype(np.float64), np.zeros(n), np.full(n, float(upper)),
              variable_types=types)
    return lp


def infeasible_milp():
    # 2 x1 + 2 x2 = 3 with x integer in [0, 3]: the LP relaxation is feasible
    # (x1 = 1.5), no integer point is.
    A = sp.csr_matrix(np.array([[2.0, 2.0]]))
    return V.LP(1, 2, A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data, np.array([1.0, 1.0]),
                np.array([3.0]), np.array([3.0]), np.zeros(2), np.full(2, 3.0),
                variable_types=np.array([1, 1], np.int32))


def highs_optimum(lp):
    A = sp.csr_matrix((lp.values, lp.column_indices, lp.row_offsets), shape=(lp.m, lp.n))
    r = so.milp(lp.objective, integrality=(np.asarray(lp.variable_types) > 0).astype(int),
                bounds=so.Bounds(lp.var_lower, lp.var_upper),
                constraints=so.LinearConstraint(A, lp.row_lower, lp.row_upper),
                options={"mip_rel_gap": 1e-12})
    assert r.status == 0, r.message
    return float(r.fun)


def check_incumbent(lp, x):
    A = sp.csr_matrix((lp.values, lp.column_indices, lp.row_offsets), shape=(lp.m, lp.n))
    ints = np.asarray(lp.variable_types) > 0
    assert np.max(np.abs(x[ints] - np.round(x[ints])), initial=0.0) <= 1e-6
    assert np.all(x >= lp.var_lower - 1e-6) and np.all(x <= lp.var_upper + 1e-6)
    ax = A @ x
    assert np.all(ax <= lp.row_upper + 1e-5) and np.all(ax >= lp.row_lower - 1e-5)


OPTS = dict(pdhg_policy=V.PdhgPolicy.Halpern, relative_kkt_tolerance=1e-9, iteration_limit=200_000,
            absolute_gap_tolerance=1e-6, relative_gap_tolerance=1e-9)
CASES = [(12, 20, 1), (20, 30, 2), (30, 40, 3)]


@pytest.mark.parametrize("batch", [1, 16])
@pytest.mark.parametrize("case", CASES)
def test_batched_bnb_reaches_the_highs_optimum(case, batch):
    m, n, seed = case
    lp = random_milp(m, n, seed)
    opt = highs_optimum(lp)
    print(f"mip {m}x{n} seed={seed} batch={batch}", flush=True)
    r, st = V.solve_mip(lp, V.Options(**OPTS), batch=batch)
    print(f"  {r.status} objective {r.objective_value:.9g} vs HiGHS {opt:.9g}; {st}", flush=True)
    assert r.status == "Optimal", (r.status, r.message)
    assert abs(r.objective_value - opt) <= 1e-6 * max(1.0, abs(opt)), (r.objective_value, opt)
    check_incumbent(lp, r.x)
    assert st["nodes_processed"] == r.iterations and st["relaxations"] <= st["nodes_processed"]
    if batch == 1:
        assert st["rounds"] == st["relaxations"] and st["max_batch"] == 1
    else:
        assert st["max_batch"] <= batch


def test_batching_cuts_the_rounds():
    lp = random_milp(30, 40, 3)
    r1, s1 = V.solve_mip(lp, V.Options(**OPTS), batch=1)
    r16, s16 = V.solve_mip(lp, V.Options(**OPTS), batch=16)
    print(f"serial {s1}\nbatched {s16}", flush=True)
    assert r1.status == r16.status == "Optimal"
    assert abs(r1.objective_value - r16.objective_value) <= 1e-6 * max(1.0, abs(r1.objective_value))
    assert s1["nodes_processed"] > 16   # a real search
    assert s16["rounds"] < s1["rounds"]


def test_integer_infeasible_milp():
    r, st = V.solve_mip(infeasible_milp(), V.Options(**OPTS), batch=8)
    assert r.status == "Infeasible", (r.status, r.message, st)
    assert st["nodes_processed"] >= 3   # the root branched, both children infeasible


def test_node_limit_and_refusals():
    lp = random_milp(30, 40, 3)
    r, st = V.solve_mip(lp, V.Options(**OPTS, node_limit=3), batch=16)
    assert r.status in ("NodeLimit", "Feasible"), r.status
    assert st["nodes_processed"] <= 3
    cont = random_milp(12, 20, 1)
    r, _ = V.solve_mip(cont, V.Options(**OPTS), variable_types=np.zeros(cont.n, np.int32))
    assert r.status == "UnsupportedModel"
    with pytest.raises(ValueError):
        V.solve_mip(cont, V.Options(**OPTS), variable_types=np.full(cont.n, 7, np.int32))


def small_miqp():
    """min c'x + 1/2 x'Qx with x1..x4 integer in [0, 3], x5 continuous in [0, 5],
    two rows A x <= b, and Q strictly diagonally dominant (positive definite)."""
    n = 5
    Q = np.array([[2.0, 0.5, 0.0, 0.0, 0.3],
                  [0.5, 1.5, 0.2, 0.0, 0.0],
                  [0.0, 0.2, 1.0, 0.3, 0.0],
                  [0.0, 0.0, 0.3, 2.5, 0.4],
                  [0.3, 0.0, 0.0, 0.4, 1.2]])
    c = np.array([-4.0, -3.5, -2.0, -5.0, -3.0])
    A = np.array([[1.0, 1.0, 1.0, 1.0, 0.5],
                  [2.0, 0.0, 1.0, 3.0, 1.0]])
    b = np.array([5.5, 9.0])
    As, Qs = sp.csr_matrix(A), sp.csr_matrix(Q)
    lp = V.LP(2, n, As.indptr.astype(np.int64), As.indices.astype(np.int64), As.data, c,
              np.full(2, -np.inf), b, np.zeros(n), np.array([3.0, 3.0, 3.0, 3.0, 5.0]),
              variable_types=np.array([1, 1, 1, 1, 0], np.int32))
    q = (Qs.indptr.astype(np.int64), Qs.indices.astype(np.int64), Qs.data)
    return lp, q, Q, c, A, b


def brute_force_miqp(Q, c, A, b):
    """Every integer assignment of x1..x4; x5 then minimises a 1-D convex
    quadratic over the interval the rows and its bounds leave (closed form)."""
    import itertools
    best = np.inf
    for xi in itertools.product(range(4), repeat=4):
        xi = np.array(xi, float)
        lo, hi = 0.0, 5.0
        for a, bi in zip(A, b):
            rest = bi - a[:4] @ xi
            if a[4] > 0:
                hi = min(hi, rest / a[4])
            elif rest < 0:
                lo = np.inf
        if lo > hi:
            continue
        x5 = min(max(-(c[4] + Q[4, :4] @ xi) / Q[4, 4], lo), hi)
        x = np.append(xi, x5)
        best = min(best, c @ x + 0.5 * x @ Q @ x)
    return best


@pytest.mark.skip(reason="MIQP through serial solve_qp relaxations: the first run of this case ended the "
                         "process with a host segmentation fault whose cause is not established; the path is "
                         "switched off (UnsupportedModel) until it is")
@pytest.mark.parametrize("batch", [1, 8])
def test_miqp_reaches_the_enumerated_optimum(batch):
    lp, q, Q, c, A, b = small_miqp()
    opt = brute_force_miqp(Q, c, A, b)
    r, st = V.solve_mip(lp, V.Options(**OPTS), batch=batch, quadratic=q)
    print(f"miqp batch={batch}: {r.status} {r.objective_value:.9g} vs enumerated {opt:.9g}; {st}", flush=True)
    assert r.status == "Optimal", (r.status, r.message)
    assert abs(r.objective_value - opt) <= 1e-6 * max(1.0, abs(opt)), (r.objective_value, opt)
    check_incumbent(lp, r.x)
    assert abs(c @ r.x + 0.5 * r.x @ Q @ r.x - r.objective_value) <= 1e-9 * max(1.0, abs(opt))


def test_miqp_is_refused():
    lp, q, *_ = small_miqp()
    r, st = V.solve_mip(lp, V.Options(**OPTS), batch=8, quadratic=q)
    assert r.status == "UnsupportedModel" and st["rounds"] == 0, (r.status, r.message)
