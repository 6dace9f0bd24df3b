"""Panel layout on the GPU: the LDS-staged SpMV (vk_hip_csr_panel_spmv)
and the solver layout built on it (vk_pdhg_params.spmv_panels).

Each group's partial sum of a row is the sequential sum, in column order,
of the row's products over the group's columns; the kernel is therefore
checked bitwise against the oracle's sequential CSR product (cpu_ops.cpp
:403-427 restated) on the group's columns.  In mixed precision the values
and x are rounded to FP32 first; their products are exact in FP64, so the
bar stays bitwise.  The lane count per workgroup must not change a bit.
The solver bars are those of the other layouts: reference policy against
the oracle (same status and iterations, objective 1e-9 relative, x and y
1e-8), Halpern against the relative-KKT certificate and HiGHS.
"""
import json
import os

import numpy as np
import pytest

import vikalp_amd as V
from oracle import oracle as O

from test_gpu_pdhg import close

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TRAJ = json.load(open(os.path.join(GOLDEN, "reference_trajectories.json")))
HIGHS = json.load(open(os.path.join(GOLDEN, "highs_optima.json")))


def group_range(n, G, g):
    c0 = 0 if g == 0 else (n * g // G) & ~3
    c1 = n if g + 1 == G else (n * (g + 1) // G) & ~3
    return c0, c1


def expected_partials(rp, ci, v, m, n, x, G):
    """Sequential row sums (the oracle's SpMV) over each group's columns."""
    out = np.zeros((G, m))
    rows = np.repeat(np.arange(m), np.diff(rp))
    for g in range(G):
        c0, c1 = group_range(n, G, g)
        keep = (ci >= c0) & (ci < c1)
        rpg = np.zeros(m + 1, np.int64)
        rpg[1:] = np.cumsum(np.bincount(rows[keep], minlength=m))
        y = np.zeros(m)
        O.spmv(rpg, ci[keep], v[keep], m, n, 1.0, x, 0.0, y, False)
        out[g] = y
    return out


def transpose_csr(rp, ci, v, m, n):
    rows = np.repeat(np.arange(m), np.diff(rp))
    order = np.lexsort((rows, ci))   # by column, rows increasing
    trp = np.zeros(n + 1, np.int64)
    trp[1:] = np.cumsum(np.bincount(ci, minlength=n))
    return trp, rows[order].astype(np.int64), v[order]


def random_csr(rng, m, n, per_row, heavy=()):
    rp, ci = [0], []
    for r in range(m):
        k = rng.poisson(per_row)
        if r in heavy:
            k = heavy[r]
        cols = np.sort(rng.choice(n, size=min(k, n), replace=False))
        ci.extend(cols.tolist())
        rp.append(len(ci))
    ci = np.array(ci, np.int64)
    return np.array(rp, np.int64), ci, rng.normal(size=ci.size)


def cases():
    rng = np.random.default_rng(11)
    out = []
    rp, ci, v = random_csr(rng, 3000, 40_001, 8, heavy={5: 40, 17: 0})
    out.append(("random 3000x40001", 3000, 40_001, rp, ci, v))
    lp = V.generate_lp(1, 20_000, 60_000, 5, 3)
    out.append(("uniform 20000x60000", lp.m, lp.n, np.asarray(lp.row_offsets, np.int64),
                np.asarray(lp.column_indices, np.int64), np.asarray(lp.values)))
    out.append(("tiny 5x3", 5, 3, np.array([0, 2, 2, 3, 5, 6]), np.array([0, 2, 1, 0, 1, 2]),
                rng.normal(size=6)))
    return out


CASES = cases()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("precision", [V.PdhgPrecision.FP64, V.PdhgPrecision.Mixed])
def test_panel_spmv_is_the_sequential_group_sum(case, precision):
    name, m, n, rp, ci, v = case
    print(f"panel spmv {name} precision={precision.name}", flush=True)
    rng = np.random.default_rng(m + n)
    dev = V.Device()
    h = dev.csr(m, n, rp, ci, v)
    try:
        vv = v.astype(np.float32).astype(np.float64) if precision == V.PdhgPrecision.Mixed else v
        trp, tci, tv = transpose_csr(rp, ci, vv, m, n)
        for transpose in (False, True):
            nx = m if transpose else n
            x = rng.normal(size=nx)
            xr = x.astype(np.float32).astype(np.float64) if precision == V.PdhgPrecision.Mixed else x
            for G in (1, 2, 4, 8):
                if transpose:
                    exp = expected_partials(trp, tci, tv, n, m, xr, G)
                else:
                    exp = expected_partials(rp, ci, vv, m, n, xr, G)
                for threads, kb in ((256, 48), (512, 48), (512, 24), (512, 16)):
                    got, _ = dev.panel_spmv(h, m, n, x, transpose=transpose, precision=precision, groups=G,
                                            threads=threads, panel_kb=kb)
                    assert np.array_equal(got, exp), (name, transpose, G, threads, kb,
                                                      float(np.max(np.abs(got - exp))))
    finally:
        dev.L.vk_hip_csr_destroy(h)


def test_panel_spmv_rejects_a_row_too_long_for_one_panel():
    # 300 entries of one row inside one panel: more than a lane may hold
    m, n = 4, 2000
    rp = np.array([0, 300, 301, 302, 303])
    ci = np.concatenate([np.arange(300), [1, 2, 3]]).astype(np.int64)
    dev = V.Device()
    h = dev.csr(m, n, rp, ci, np.ones(ci.size))
    try:
        with pytest.raises(RuntimeError, match="does not fit"):
            dev.panel_spmv(h, m, n, np.ones(n), groups=1, threads=256)
    finally:
        dev.L.vk_hip_csr_destroy(h)


@pytest.mark.parametrize("groups", [0, 2, 8])
def test_panel_reference_policy_matches_oracle_fixture(groups):
    # with one group a row of this LP holds up to 29 entries in a panel: the
    # layout does not fit (covered below)
    c = TRAJ["uniform_1000x3000_k5_it2000"]
    print(f"fixture panels groups={groups}", flush=True)
    lp = V.generate_lp(c["family"], c["m"], c["n"], c["k"], c["seed"])
    opts = dict(c["options"])
    got = V.solve_lp(lp, V.Options(spmv_panels=2, spmv_panel_groups=groups, spmv_panel_groups_t=groups, **opts))
    ref = O.solve_lp(lp.as_oracle(), **opts)
    close(got, ref)
    assert got.iterations == c["iterations"]


@pytest.mark.parametrize("precision", [V.PdhgPrecision.FP64, V.PdhgPrecision.Mixed])
def test_panel_reference_policy_matches_oracle_many_panels(precision):
    # 40,000 columns: several panels per group in both precisions
    lp = V.generate_lp(1, 12_000, 40_000, 5, 21)
    print(f"panels many-panel case precision={precision.name}", flush=True)
    got = V.solve_lp(lp, V.Options(spmv_panels=2, pdhg_precision=precision, iteration_limit=600))
    ref = O.solve_lp(lp.as_oracle(), iteration_limit=600)
    if precision == V.PdhgPrecision.FP64:
        close(got, ref)
    else:
        # mixed runs its FP32 phase then FP64: same status, close iterates
        assert got.status == ref.status


@pytest.mark.parametrize("precision", [V.PdhgPrecision.FP64, V.PdhgPrecision.Mixed])
def test_panel_halpern_certificate_and_optimum(precision):
    g = HIGHS["uniform_1000x3000_k5"]
    lp = V.generate_lp(g["family"], g["m"], g["n"], g["k"], g["seed"])
    tol = 1e-8
    r = V.solve_lp(lp, V.Options(pdhg_policy=V.PdhgPolicy.Halpern, pdhg_precision=precision,
                                 relative_kkt_tolerance=tol, iteration_limit=400_000, spmv_panels=2))
    assert r.status == "Optimal", r
    rp, rd, gp, _, _ = O.relative_kkt(lp.as_oracle(), r.x, r.y)
    assert max(rp, rd, gp) <= tol * 1.0001, (rp, rd, gp)
    assert abs(r.objective_value - g["objective"]) / max(1.0, abs(g["objective"])) <= 1e-6


def test_panel_runs_repeat_bitwise_and_ignore_lane_count_and_panel_size():
    lp = V.generate_lp(2, 400, 1000, 0, 5)
    kw = dict(iteration_limit=777, spmv_panels=2)
    a = V.solve_lp(lp, V.Options(use_graphs=True, spmv_panel_threads=256, **kw))
    runs = [V.solve_lp(lp, V.Options(use_graphs=False, spmv_panel_threads=512, **kw)),
            V.solve_lp(lp, V.Options(use_graphs=True, spmv_panel_threads=512, **kw))]
    runs += [V.solve_lp(lp, V.Options(spmv_panel_kb=kb, **kw)) for kb in (48, 24, 16)]
    for r in runs:
        assert r.iterations == a.iterations
        assert np.array_equal(r.x, a.x) and np.array_equal(r.y, a.y)


@pytest.mark.parametrize("precision", [V.PdhgPrecision.FP64, V.PdhgPrecision.Mixed])
def test_panel_tuner_changes_no_result(precision):
    # The tuner picks the lane count and the unroll depth per operand by
    # timing them on the matrix; neither changes a bit of the iterates.
    lp = V.generate_lp(1, 12_000, 40_000, 5, 21)
    kw = dict(spmv_panels=2, pdhg_precision=precision, iteration_limit=500)
    a = V.solve_lp(lp, V.Options(**kw))
    b = V.solve_lp(lp, V.Options(spmv_autotune=True, **kw))
    assert (a.status, a.iterations) == (b.status, b.iterations)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)
    prof = V.profile_kernels(lp, V.Options(spmv_autotune=True, spmv_panels=2, pdhg_precision=precision),
                             warm_iterations=2, iterations=2)
    for i in (0, 5):
        assert prof["panels"][i + 1] in (256, 512) and prof["panels"][i + 3] in (4, 8, 12, 16), prof
        assert prof["panels"][i + 4] in (48, 24, 16), prof


def test_panel_automatic_at_scale_matches_oracle():
    # n = 300,000 > 2^18: the automatic rule picks the panel layout
    lp = V.generate_lp(1, 100_000, 300_000, 5, 7)
    print(f"auto panels m={lp.m} n={lp.n} nnz={lp.nnz}", flush=True)
    got = V.solve_lp(lp, V.Options(iteration_limit=200))
    ref = O.solve_lp(lp.as_oracle(), iteration_limit=200)
    close(got, ref)
    prof = V.profile_kernels(lp, V.Options(), warm_iterations=2, iterations=2)
    assert prof["panels"][0] > 0 and prof["panels"][5] > 0, prof
    assert prof["dual_spmv_ms"] > 0 and prof["dual_pass_ms"] > 0
    assert prof["primal_spmv_ms"] > 0 and prof["primal_pass_ms"] > 0


def test_forced_panels_on_a_matrix_that_does_not_fit_fail_cleanly():
    # zipf 400 x 1000: 44 entries per row, all in one panel
    c = TRAJ["zipf_400x1000_it1500"]
    lp = V.generate_lp(c["family"], c["m"], c["n"], c["k"], c["seed"])
    r = V.solve_lp(lp, V.Options(spmv_panels=2, iteration_limit=100))
    assert r.status != "Optimal" and "panel" in r.message, r
    # automatic choice falls back to the tile layout
    auto = V.solve_lp(lp, V.Options(spmv_panels=0, iteration_limit=100))
    close(auto, O.solve_lp(lp.as_oracle(), iteration_limit=100))


def test_invalid_panel_options_are_rejected():
    lp = V.generate_lp(1, 40, 90, 3, 3)
    for kw in (dict(spmv_panels=3), dict(spmv_panel_groups=3), dict(spmv_panel_groups_t=16),
               dict(spmv_panel_threads=128), dict(spmv_panel_kb=32),
               dict(spmv_panel_threads=256, spmv_panel_kb=16)):
        assert V.solve_lp(lp, V.Options(**kw)).status == "InvalidModel", kw
