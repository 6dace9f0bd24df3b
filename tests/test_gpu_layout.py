"""Column-blocked A x layout on the GPU (vk_pdhg_params.col_blocks, the
layout tuner).

The layout changes only the summation order of A x, so every parity bar of
the plain row tiles holds unchanged: reference policy against the CPU
oracle (same status and iterations, objective 1e-9 relative, x and y 1e-8),
adaptive policy against the relative-KKT certificate and the independent
HiGHS optimum.  Block counts that are not powers of two, more blocks than
columns, rows longer than one tile inside a block and rows with no entry in
a block are all covered.
"""
import json
import math
import os

import numpy as np
import pytest

import vikalp_amd as V
from oracle import oracle as O

from test_gpu_pdhg import close, objective_bar

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TRAJ = json.load(open(os.path.join(GOLDEN, "reference_trajectories.json")))
HIGHS = json.load(open(os.path.join(GOLDEN, "highs_optima.json")))


@pytest.mark.parametrize("blocks", [2, 3, 8, 16])
@pytest.mark.parametrize("name", ["zipf_400x1000_it1500", "uniform_1000x3000_k5_it2000", "block_T4_it500"])
def test_blocked_reference_policy_matches_oracle(name, blocks):
    c = TRAJ[name]
    print(f"case {name} col_blocks={blocks}", flush=True)
    lp = V.generate_lp(c["family"], c["m"], c["n"], c["k"], c["seed"])
    opts = dict(c["options"])
    got = V.solve_lp(lp, V.Options(spmv_col_blocks=blocks, **opts))
    ref = O.solve_lp(lp.as_oracle(), **opts)
    close(got, ref)
    assert got.iterations == c["iterations"]


def dense_row_lp(n=3000, m=5, seed=3):
    """Row 0 touches every column (longer than a tile in each of 2-4 blocks);
    row 1 touches only the first 10 columns (empty in the later blocks)."""
    rng = np.random.default_rng(seed)
    rows = [np.arange(n), np.arange(10)]
    for _ in range(m - 2):
        rows.append(np.sort(rng.choice(n, size=40, replace=False)))
    rp = np.zeros(m + 1, np.int64)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.concatenate(rows).astype(np.int64)
    vals = rng.normal(size=ci.size)
    x0 = rng.uniform(0.5, 9.5, n)
    A = np.zeros((m, n))
    for i, r in enumerate(rows):
        A[i, r] = vals[rp[i]:rp[i + 1]]
    b = A @ x0
    c = A.T @ rng.normal(size=m) + rng.uniform(0, 1, n) * (rng.uniform(size=n) < 0.5)
    return V.LP(m, n, rp, ci, vals, c, b, b, np.zeros(n), np.full(n, 10.0))


@pytest.mark.parametrize("blocks", [2, 4])
def test_blocked_long_and_empty_rows(blocks):
    lp = dense_row_lp()
    got = V.solve_lp(lp, V.Options(spmv_col_blocks=blocks, iteration_limit=700))
    ref = O.solve_lp(lp.as_oracle(), iteration_limit=700)
    close(got, ref)


def test_more_blocks_than_columns():
    lp = V.LP(1, 3, np.array([0, 3]), np.array([0, 1, 2]), np.array([1.0, 2.0, -1.0]),
              np.array([1.0, -1.0, 0.5]), np.array([2.0]), np.array([2.0]), np.zeros(3), np.full(3, 5.0))
    got = V.solve_lp(lp, V.Options(spmv_col_blocks=16, iteration_limit=300))
    ref = O.solve_lp(lp.as_oracle(), iteration_limit=300)
    close(got, ref)


def test_blocked_runs_repeat_bitwise():
    lp = V.generate_lp(2, 400, 1000, 0, 5)
    a = V.solve_lp(lp, V.Options(iteration_limit=777, spmv_col_blocks=8, use_graphs=True))
    b = V.solve_lp(lp, V.Options(iteration_limit=777, spmv_col_blocks=8, use_graphs=False))
    c = V.solve_lp(lp, V.Options(iteration_limit=777, spmv_col_blocks=8, use_graphs=True))
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)
    assert np.array_equal(a.x, c.x) and np.array_equal(a.y, c.y)


@pytest.mark.parametrize("precision", [V.PdhgPrecision.FP64, V.PdhgPrecision.Mixed])
def test_blocked_adaptive_reaches_relative_kkt(precision):
    g = HIGHS["uniform_1000x3000_k5"]
    lp = V.generate_lp(g["family"], g["m"], g["n"], g["k"], g["seed"])
    tol = 1e-6
    r = V.solve_lp(lp, V.Options(pdhg_policy=V.PdhgPolicy.Adaptive, pdhg_precision=precision,
                                 relative_kkt_tolerance=tol, iteration_limit=200_000, spmv_col_blocks=8))
    assert r.status == "Optimal", r
    rp, rd, gp, _, _ = O.relative_kkt(lp.as_oracle(), r.x, r.y)
    assert max(rp, rd, gp) <= tol * 1.0001, (rp, rd, gp)
    assert abs(r.objective_value - g["objective"]) / max(1.0, abs(g["objective"])) <= objective_bar(tol)


@pytest.mark.parametrize("precision", [V.PdhgPrecision.FP64, V.PdhgPrecision.Mixed])
def test_layout_tuner(precision):
    lp = V.generate_lp(2, 400, 1000, 0, 5)
    dev = V.Device()
    h = dev.csr(lp.m, lp.n, lp.row_offsets, lp.column_indices, lp.values)
    try:
        best, ms = dev.tune_col_blocks(h, precision, repeats=3)
    finally:
        dev.L.vk_hip_csr_destroy(h)
    assert best in (1, 2, 4, 8, 16)
    assert all(math.isfinite(t) and t > 0 for t in ms.values()), ms
    assert ms[best] == min(ms.values())
    # the solve path with the tuner on still meets the oracle bar
    got = V.solve_lp(lp, V.Options(spmv_autotune=True, pdhg_precision=V.PdhgPrecision.FP64,
                                   iteration_limit=400))
    close(got, O.solve_lp(lp.as_oracle(), iteration_limit=400))


def test_invalid_block_count_is_rejected():
    lp = V.generate_lp(1, 40, 90, 3, 3)
    assert V.solve_lp(lp, V.Options(spmv_col_blocks=17)).status == "InvalidModel"
    assert V.solve_lp(lp, V.Options(spmv_col_blocks=-1)).status == "InvalidModel"


@pytest.mark.parametrize("blocks", [1, 4])
@pytest.mark.parametrize("precision", [V.PdhgPrecision.FP64, V.PdhgPrecision.Mixed])
def test_split_passes_equal_fused_bitwise(precision, blocks):
    # The split kernels store the FP64 sums the fused epilogues use, so the
    # iterates agree bitwise; the reference policy decides on max norms only.
    lp = V.generate_lp(2, 400, 1000, 0, 5)
    kw = dict(pdhg_precision=precision, iteration_limit=1500, spmv_col_blocks=blocks)
    a = V.solve_lp(lp, V.Options(spmv_split_passes=1, **kw))
    b = V.solve_lp(lp, V.Options(spmv_split_passes=2, **kw))
    assert (a.status, a.iterations, a.low_precision_iterations) == (b.status, b.iterations, b.low_precision_iterations)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)


@pytest.mark.parametrize("precision", [V.PdhgPrecision.FP64, V.PdhgPrecision.Mixed])
def test_split_passes_adaptive_certificate(precision):
    # The adaptive policy's restart and primal-weight rules read KKT sums
    # whose block grouping differs between the layouts (round-off only):
    # both layouts must meet the same certificate and agree on the optimum.
    g = HIGHS["uniform_1000x3000_k5"]
    lp = V.generate_lp(g["family"], g["m"], g["n"], g["k"], g["seed"])
    tol = 1e-6
    res = []
    for split in (1, 2):
        r = V.solve_lp(lp, V.Options(pdhg_policy=V.PdhgPolicy.Adaptive, pdhg_precision=precision,
                                     relative_kkt_tolerance=tol, iteration_limit=200_000,
                                     spmv_split_passes=split))
        assert r.status == "Optimal", r
        rp, rd, gp, _, _ = O.relative_kkt(lp.as_oracle(), r.x, r.y)
        assert max(rp, rd, gp) <= tol * 1.0001, (split, rp, rd, gp)
        res.append(r)
    for r in res:
        assert abs(r.objective_value - g["objective"]) / max(1.0, abs(g["objective"])) <= objective_bar(tol)


def test_automatic_layout_at_scale_matches_oracle():
    # n = 300,000 > 2^18 and 2.4 MB of FP64 x: with the panel layout off,
    # the automatic rule picks 4 column blocks and split passes
    lp = V.generate_lp(1, 100_000, 300_000, 5, 7)
    print(f"auto layout m={lp.m} n={lp.n} nnz={lp.nnz}", flush=True)
    got = V.solve_lp(lp, V.Options(iteration_limit=200, spmv_panels=1))
    ref = O.solve_lp(lp.as_oracle(), iteration_limit=200)
    close(got, ref)
    prof = V.profile_kernels(lp, V.Options(spmv_panels=1), warm_iterations=2, iterations=2)
    assert prof["col_blocks"] == 4 and prof["split_passes"] == 1
    # the row pass is its own kernel; the A' side is one kernel (k_primal_own)
    assert prof["dual_pass_ms"] > 0 and prof["primal_spmv_ms"] > 0 and prof["primal_pass_ms"] == 0


@pytest.mark.parametrize("policy", [V.PdhgPolicy.Reference, V.PdhgPolicy.Halpern])
def test_cache_budget_changes_load_hints_only(policy):
    # The MALL budget picks cache-allocating or non-temporal loads per tile;
    # every budget (none, tiny, partial, automatic) must give bitwise-equal
    # results on the column-blocked split layout.
    lp = V.generate_lp(1, 100_000, 300_000, 5, 7)
    print(f"cache budget policy={policy.name}", flush=True)
    opts = dict(pdhg_policy=policy, iteration_limit=300, spmv_panels=1)
    if policy != V.PdhgPolicy.Reference:
        opts["relative_kkt_tolerance"] = 1e-12
    runs = [V.solve_lp(lp, V.Options(spmv_cache_budget_mb=b, **opts)) for b in (-1, 1, 8, 0)]
    for r in runs[1:]:
        assert r.iterations == runs[0].iterations
        assert np.array_equal(r.x, runs[0].x) and np.array_equal(r.y, runs[0].y)
    assert V.solve_lp(lp, V.Options(spmv_cache_budget_mb=-2, **opts)).status == "InvalidModel"
