"""PDQP on the GPU against the CPU oracle (solver::solve_qp through vk_qp_solve).

Same bar as the LP reference policy: status and iteration count equal,
objective within 1e-9 relative, x and y within 1e-8 (the SpMV summation
order and fused multiply-adds differ), on every case of the QP fixture file,
for the fused device iteration and for the op-by-op HipBackend sequence.
The reference's own QP cases (tests/solver/continuous/test_pdhg.cpp:139-164)
run on both paths.
"""
import json
import os

import numpy as np
import pytest

import vikalp_amd as V
from oracle import oracle as O
from qp_data import CASES, make_q
from test_gpu_pdhg import close

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TRAJ = json.load(open(os.path.join(GOLDEN, "reference_qp_trajectories.json")))


@pytest.mark.parametrize("name", sorted(TRAJ))
@pytest.mark.parametrize("fused", [True, False])
def test_pdqp_matches_oracle(name, fused):
    c = TRAJ[name]
    print(f"qp case {name} fused={fused}", flush=True)
    lp = V.generate_lp(*c["generator"])
    q = make_q(lp.n, c["q_seed"])
    got = V.solve_qp(lp, q, V.Options(fused_iterations=fused, **c["options"]))
    ref = O.solve_qp(lp.as_oracle(), q, **c["options"])
    close(got, ref)
    assert got.status == c["status"] and got.iterations == c["iterations"]
    assert got.solver == "PDQP"


def scalar_lp(c, lo, hi):
    return V.LP(0, 1, np.array([0]), np.array([], np.int64), np.array([]), np.array([c]),
                np.array([]), np.array([]), np.array([lo]), np.array([hi]))


@pytest.mark.parametrize("fused", [True, False])
def test_reference_qp_cases(fused):
    lp = scalar_lp(-10.0, 0.0, 10.0)
    q = (np.array([0, 1]), np.array([0]), np.array([2.0]))
    o = V.Options(iteration_limit=1000, primal_tolerance=1e-6, dual_tolerance=1e-6, fused_iterations=fused)
    r = V.solve_qp(lp, q, o)
    assert r.status == "Optimal" and abs(r.x[0] - 5.0) <= 1e-4 and abs(r.objective_value + 25.0) <= 1e-4
    r = V.solve_qp(lp, q, o, warm_start=np.array([1.0]), upper_override=np.array([4.0]))
    assert r.status == "Optimal" and abs(r.x[0] - 4.0) <= 1e-4
    assert V.solve_qp(lp, (np.array([0, 1]), np.array([0]), np.array([-2.0])), o).status == "UnsupportedModel"


def test_router_and_rejected_options():
    name, gen, qseed, opts = CASES[2]
    lp = V.generate_lp(*gen)
    q = make_q(lp.n, qseed)
    r = V.solve(lp, V.Options(backend=V.BackendPreference.HIP, **opts), quadratic=q)
    ref = O.solve_qp(lp.as_oracle(), q, **opts)
    assert r.solver == "vikalp-continuous" and r.backend == "hip"
    close(r, ref)
    for bad in (dict(pdhg_policy=V.PdhgPolicy.Adaptive), dict(pdhg_precision=V.PdhgPrecision.Mixed),
                dict(relative_kkt_tolerance=1e-6)):
        assert V.solve_qp(lp, q, V.Options(**bad)).status == "UnsupportedModel"
    asym = (q[0], q[1], q[2] + np.where(q[1] == 0, 1.0, 0.0))    # breaks symmetry in column 0
    assert V.solve_qp(lp, asym).status == "InvalidModel"
