"""Branch-and-bound node relaxations on one device-resident matrix
(solver::solve_lp_nodes, vk_lp_solve_nodes, vk_hip_batch_create_shared).

Each node is the model with its own variable bounds (B&B branching: one
variable's upper bound set to floor(x0_j) or lower bound to ceil(x0_j)), so
each node's result must be the solve of that bounded LP alone:
- reference policy: equal to the CPU oracle's solve of the node LP (same
  status and iterations, x and y within 1e-8, the usual bar);
- Halpern policy, FP64 and mixed: the relative-KKT certificate recomputed by
  the oracle on every node, and the node objective against the node LP's
  own single solve;
- warm starts per node, invalid node bounds (that node InvalidModel, the
  others solved), integer variables relaxed, and a node batch large enough
  for the tail hand-off to the single-LP solver.
"""
import numpy as np
import pytest

import vikalp_amd as V
from oracle import oracle as O

from test_gpu_pdhg import close

pytestmark = pytest.mark.gpu


def branched(lp, count, seed):
    """count node bound sets: each tightens 1-3 random variables around x0."""
    rng = np.random.default_rng(seed)
    lo = np.tile(lp.var_lower, (count, 1))
    hi = np.tile(lp.var_upper, (count, 1))
    for k in range(count):
        for j in rng.choice(lp.n, size=rng.integers(1, 4), replace=False):
            if rng.random() < 0.5:
                hi[k, j] = np.floor(lp.x0[j])
            else:
                lo[k, j] = np.ceil(lp.x0[j])
    return lo, hi


def node_lp(lp, lo, hi):
    return V.LP(lp.m, lp.n, lp.row_offsets, lp.column_indices, lp.values, lp.objective, lp.row_lower,
                lp.row_upper, lo, hi)


@pytest.mark.parametrize("gen", [(1, 300, 900, 5, 4), (4, 0, 0, 4, 9)])
def test_reference_nodes_match_oracle(gen):
    lp = V.generate_lp(*gen)
    lo, hi = branched(lp, 6, 1)
    print(f"reference nodes gen={gen}", flush=True)
    got = V.solve_nodes(lp, lo, hi, options=V.Options(iteration_limit=1500))
    for k in range(6):
        close(got[k], O.solve_lp(node_lp(lp, lo[k], hi[k]).as_oracle(), iteration_limit=1500))


@pytest.mark.parametrize("precision", [V.PdhgPrecision.FP64, V.PdhgPrecision.Mixed])
def test_halpern_nodes_meet_their_certificates(precision):
    lp = V.generate_lp(1, 300, 900, 5, 4)
    lo, hi = branched(lp, 5, 2)
    opts = V.Options(pdhg_policy=V.PdhgPolicy.Halpern, pdhg_precision=precision, relative_kkt_tolerance=1e-6,
                     iteration_limit=100_000)
    got = V.solve_nodes(lp, lo, hi, options=opts)
    for k in range(5):
        nl = node_lp(lp, lo[k], hi[k])
        single = V.solve_lp(nl, opts)
        assert got[k].status == single.status == "Optimal", (got[k].status, single.status)
        rp, rd, gp, _, _ = O.relative_kkt(nl.as_oracle(), got[k].x, got[k].y)
        assert max(rp, rd, gp) <= 1e-6 * 1.0001, (rp, rd, gp)
        # two points that each meet relative KKT 1e-6 agree on the objective
        # to about the gap tolerance times the objective scale
        assert abs(got[k].objective_value - single.objective_value) <= 1e-4 * max(1.0, abs(single.objective_value))


def test_warm_started_nodes_and_invalid_bounds():
    lp = V.generate_lp(1, 300, 900, 5, 4)
    lo, hi = branched(lp, 4, 3)
    lo[2, 5], hi[2, 5] = 3.0, 2.0   # node 2: empty box -> InvalidModel
    opts = V.Options(pdhg_policy=V.PdhgPolicy.Halpern, relative_kkt_tolerance=1e-6, iteration_limit=100_000)
    cold = V.solve_nodes(lp, lo, hi, options=opts)
    assert cold[2].status == "InvalidModel"
    root = V.solve_lp(lp, opts)
    warm = V.solve_nodes(lp, lo, hi, warm_start=np.tile(root.x, (4, 1)), options=opts)
    assert warm[2].status == "InvalidModel"
    for k in (0, 1, 3):
        assert cold[k].status == warm[k].status == "Optimal"
        rp, rd, gp, _, _ = O.relative_kkt(node_lp(lp, lo[k], hi[k]).as_oracle(), warm[k].x, warm[k].y)
        assert max(rp, rd, gp) <= 1e-6 * 1.0001
        assert abs(warm[k].objective_value - cold[k].objective_value) <= 1e-4 * max(1.0, abs(cold[k].objective_value))


def test_many_nodes_with_tail_hand_off():
    # 40 nodes: the last running ones finish on the single-LP solver.
    lp = V.generate_lp(4, 0, 0, 6, 20261203 + 7)
    lo, hi = branched(lp, 40, 4)
    opts = V.Options(pdhg_policy=V.PdhgPolicy.Halpern, relative_kkt_tolerance=1e-6, iteration_limit=20_000)
    print("40 nodes, Halpern", flush=True)
    got = V.solve_nodes(lp, lo, hi, options=opts)
    from test_gpu_infeasible import farkas_d0
    for k, g in enumerate(got):
        nl = node_lp(lp, lo[k], hi[k])
        assert g.status in ("Optimal", "Infeasible", "IterationLimit"), g.status
        if g.status == "Optimal":
            rp, rd, gp, _, _ = O.relative_kkt(nl.as_oracle(), g.x, g.y)
            assert max(rp, rd, gp) <= 1e-6 * 1.0001
        elif g.status == "Infeasible":
            d0 = farkas_d0(nl, g.y)
            assert np.isfinite(d0) and d0 > 0
    assert sum(g.status == "Optimal" for g in got) >= 20
