"""Parity at the full benchmark sizes (configurations 2, 4 and 5), on the GPU.

The reference accepts its GPU backend by comparing it with the CPU backend on
the same inputs (tests/test_cuda_consistency.cpp:56-178); these tests do the
same at the sizes the benchmark quotes, against the CPU oracle:

- the 400,000 x 1,000,000 Zipf LP (seed 20261203): the reference policy for
  a bounded number of iterations against the oracle's solve (same status and
  iterations, objective 1e-9 relative, x and y 1e-8) in both SpMV layouts;
  the Halpern FP64 and mixed solves to 1e-6 relative KKT, recomputed on the
  returned points by the oracle's own routine, and their objectives within
  1e-6 of each other;
- the 512 block-angular instances (T ~ U{4..48}, 1-20 branching bounds
  each), Halpern, FP64 and mixed: every Optimal instance certified at 1e-6
  relative KKT, every Infeasible one by its Farkas ray, and 16 sampled
  instances under the reference policy against their oracle solves;
- the unrolled PDHG on the 100,000 x 250,000 LP, K = 1000 with checkpoints
  every 50: forward against the oracle to 1e-9 and the gradients to 1e-8.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import vikalp_amd as V
from oracle import oracle as O

from test_gpu_pdhg import close

pytestmark = pytest.mark.gpu
SEED = 20261203
TOL = 1e-6


def relative(a, b):
    return abs(a - b) / max(1.0, abs(b))


@pytest.fixture(scope="module")
def zipf_lp():
    lp = V.generate_lp(2, 400_000, 1_000_000, 0, SEED)
    print(f"config 2: m={lp.m} n={lp.n} nnz={lp.nnz}", flush=True)
    return lp


def test_config2_reference_policy_matches_oracle(zipf_lp):
    its = 120
    ref = O.solve_lp(zipf_lp.as_oracle(), iteration_limit=its)
    for panels in (0, 1):   # automatic (the panel layout at this size), then the tile layout
        print(f"config 2 reference policy, spmv_panels={panels}", flush=True)
        got = V.solve_lp(zipf_lp, V.Options(iteration_limit=its, spmv_panels=panels))
        close(got, ref)


@pytest.fixture(scope="module")
def zipf_halpern(zipf_lp):
    out = {}
    for label, precision in (("fp64", V.PdhgPrecision.FP64), ("mixed", V.PdhgPrecision.Mixed)):
        print(f"config 2 Halpern {label}", flush=True)
        out[label] = V.solve_lp(zipf_lp, V.Options(pdhg_policy=V.PdhgPolicy.Halpern, pdhg_precision=precision,
                                                   relative_kkt_tolerance=TOL, mixed_phase_tolerance=1e-5,
                                                   iteration_limit=200_000))
    return out


@pytest.mark.parametrize("label", ["fp64", "mixed"])
def test_config2_halpern_certificate(zipf_lp, zipf_halpern, label):
    r = zipf_halpern[label]
    assert r.status == "Optimal", r
    rp, rd, gp, pobj, dobj = O.relative_kkt(zipf_lp.as_oracle(), r.x, r.y)
    assert max(rp, rd, gp) <= TOL * 1.0001, (rp, rd, gp)
    assert relative(r.objective_value, pobj) <= 1e-12


def test_config2_mixed_objective_matches_fp64(zipf_halpern):
    assert relative(zipf_halpern["mixed"].objective_value, zipf_halpern["fp64"].objective_value) <= TOL


def batch_instances():
    rng = np.random.default_rng(SEED)
    periods = rng.integers(4, 49, size=512)
    return [V.generate_lp(4, 0, 0, int(t), SEED + i) for i, t in enumerate(periods)]


def farkas_d0(lp, y):
    """Dual objective without c of the row duals y (the Farkas test): > 0
    with every term finite proves the LP infeasible."""
    A = sp.csr_matrix((lp.values, lp.column_indices, lp.row_offsets), shape=(lp.m, lp.n))
    g = A.T @ y
    col = np.where(g > 0, lp.var_lower, lp.var_upper)
    row = np.where(y > 0, lp.row_upper, lp.row_lower)
    gm, ym = g != 0, y != 0
    if not (np.isfinite(col[gm]).all() and np.isfinite(row[ym]).all()):
        return -np.inf
    return float(np.sum(g[gm] * col[gm]) - np.sum(y[ym] * row[ym]))


@pytest.fixture(scope="module")
def batch():
    lps = batch_instances()
    print(f"config 4: {len(lps)} instances, {sum(l.m for l in lps)} rows, {sum(l.n for l in lps)} columns",
          flush=True)
    return lps


@pytest.mark.parametrize("precision", [V.PdhgPrecision.FP64, V.PdhgPrecision.Mixed])
def test_config4_batch_certificates(batch, precision):
    print(f"config 4 Halpern batch precision={precision.name}", flush=True)
    res = V.solve_batch(batch, V.Options(pdhg_policy=V.PdhgPolicy.Halpern, pdhg_precision=precision,
                                         relative_kkt_tolerance=TOL, mixed_phase_tolerance=1e-4,
                                         iteration_limit=40_000))
    counts = {}
    for lp, r in zip(batch, res):
        counts[r.status] = counts.get(r.status, 0) + 1
        if r.status == "Optimal":
            rp, rd, gp, _, _ = O.relative_kkt(lp.as_oracle(), r.x, r.y)
            assert max(rp, rd, gp) <= TOL * 1.0001, (rp, rd, gp)
        elif r.status == "Infeasible":
            d0 = farkas_d0(lp, r.y)
            assert np.isfinite(d0) and d0 > 0, d0
        else:
            assert r.status == "IterationLimit" and r.iterations == 40_000, r
    print(f"config 4 {precision.name}: {counts}", flush=True)
    assert counts.get("Optimal", 0) + counts.get("Infeasible", 0) >= 500, counts


def test_config4_reference_policy_sample_matches_oracle(batch):
    its = 300
    res = V.solve_batch(batch, V.Options(iteration_limit=its))
    sample = np.random.default_rng(SEED + 4).choice(len(batch), size=16, replace=False)
    for i in sample:
        print(f"config 4 reference sample instance {i}", flush=True)
        close(res[i], O.solve_lp(batch[i].as_oracle(), iteration_limit=its))


def test_config5_unrolled_forward_and_adjoint_match_oracle():
    lp = V.generate_lp(1, 100_000, 250_000, 6, SEED)
    K = 1000
    print(f"config 5: m={lp.m} n={lp.n} nnz={lp.nnz} K={K}", flush=True)
    u = V.UnrolledPdhg(lp, K, checkpoint_every=50)
    x, y = u.forward()
    ol = lp.as_oracle()
    xr, yr, step = O.unrolled(ol, K)
    assert u.info()["step_size"] == pytest.approx(step, rel=1e-15)
    scale = lambda b: max(1.0, float(np.max(np.abs(b))))
    assert np.max(np.abs(x - xr)) <= 1e-9 * scale(xr) and np.max(np.abs(y - yr)) <= 1e-9 * scale(yr)
    xt = np.random.default_rng(SEED + 5).uniform(0, 10, lp.n)
    gc, grl, gru = u.backward(lp.objective + x - xt)
    _, _, _, gc_r, grl_r, gru_r = O.unrolled(ol, K, grad_x=lp.objective + xr - xt)
    for a, b in ((gc, gc_r), (grl, grl_r), (gru, gru_r)):
        assert np.max(np.abs(a - b)) <= 1e-8 * scale(b)
