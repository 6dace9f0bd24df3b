"""Solving LPs read from MPS files (vk_mps_read, then solve_lp on the GPU).

A generated LP written as MPS and read back is bit-identical to the
original (tests/test_mps.py), so its solve must match the CPU oracle's
solve of the original exactly as the generated LP's own solve does: same
status and iterations, x and y within 1e-8.  The reference fixture
(tests/io/simple_lp.mps: min x, x <= 5, 0 <= x <= 10) solves to x = 0.
"""
import os

import numpy as np
import pytest

import vikalp_amd as V
from oracle import oracle as O

from test_gpu_pdhg import close
from test_mps import GOLDEN, write_lp

pytestmark = pytest.mark.gpu


def test_reference_fixture_solves():
    lp = V.read_mps(os.path.join(GOLDEN, "simple_lp.mps"))
    r = V.solve_lp(lp, V.Options(iteration_limit=5000))
    assert r.status == "Optimal" and abs(r.x[0]) <= 1e-6 and abs(r.objective_value) <= 1e-6


@pytest.mark.parametrize("gen", [(1, 300, 900, 5, 4), (2, 400, 1000, 0, 5)])
def test_mps_lp_matches_oracle(tmp_path, gen):
    lp = V.generate_lp(*gen)
    path = str(tmp_path / "lp.mps")
    write_lp(path, lp)
    got = V.solve_lp(V.read_mps(path), V.Options(iteration_limit=1500))
    close(got, O.solve_lp(lp.as_oracle(), iteration_limit=1500))
    # and the Halpern policy reaches the relative-KKT target on it
    r = V.solve_lp(V.read_mps(path), V.Options(pdhg_policy=V.PdhgPolicy.Halpern, relative_kkt_tolerance=1e-6,
                                               iteration_limit=200_000))
    assert r.status == "Optimal"
    rp, rd, gp, _, _ = O.relative_kkt(lp.as_oracle(), r.x, r.y)
    assert max(rp, rd, gp) <= 1e-6 * 1.0001
