"""Restarted Halpern PDHG (VK_PDHG_POLICY_HALPERN) on the GPU.

Not in the reference, so pinned three ways:
- the device iterates against a plain numpy restatement of the method
  (Lu and Yang 2024: z+ = w1 (2 T(z) - z) + w0 z0 with the fixed-point
  residual restart rule and the primal-weight update), on unscaled LPs with
  a given step and primal weight, through several restarts, in the fused
  and the column-blocked split layouts: x and y within 1e-9, same restart
  count;
- the relative-KKT certificate recomputed by the oracle on the returned
  point and the objective against the independent HiGHS optimum, in FP64
  and mixed precision;
- bitwise repeats, the batched kernel's per-instance certificates, and the
  refusal of the op-by-op path, which does not implement it.
"""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest
import scipy.sparse as sp

import vikalp_amd as V
from vikalp_amd import native as N
from oracle import oracle as O

from test_gpu_pdhg import objective_bar

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HIGHS = json.load(open(os.path.join(GOLDEN, "highs_optima.json")))


def halpern_np(lp, eta, omega, period, iterations, kp=0.5, ki=0.0):
    """The device's Halpern iteration restated with numpy on the unscaled LP,
    from x0 = P_X(0), y0 = 0, with the primal-weight update
    log w <- log w - kp e - ki sum(e), e = log w - log(dy/dx) (kp = 1/2,
    ki = 0: PDLP's smoothing).  Returns (x+, y+, restarts) of the last
    iteration."""
    A = sp.csr_matrix((lp.values, lp.column_indices, lp.row_offsets), shape=(lp.m, lp.n))
    AT = A.T.tocsr()
    c, l, u, rl, ru = lp.objective, lp.var_lower, lp.var_upper, lp.row_lower, lp.row_upper
    x = np.clip(np.zeros(lp.n), l, u)
    y = np.zeros(lp.m)
    ax, aty = A @ x, AT @ y
    tau, sigma = eta / omega, eta * omega
    z0 = (x.copy(), y.copy(), ax.copy(), aty.copy())
    xr, yr = x.copy(), y.copy()
    ks, r0, prev, restarts, esum = 0, None, math.inf, 0, 0.0
    xp = yp = None
    for k in range(iterations):
        xp = np.clip((x - tau * c) - tau * aty, l, u)
        axp = A @ xp
        v = y + sigma * (2.0 * axp - ax)
        t = v / sigma
        yp = np.where(t < rl, np.minimum(v - sigma * rl, 0.0), np.where(ru < t, np.maximum(v - sigma * ru, 0.0), 0.0))
        atyp = AT @ yp
        r = math.sqrt(omega * np.sum((xp - x) ** 2) + np.sum((yp - y) ** 2) / omega)
        if ks == 0:
            r0 = r
        if (k + 1) % period == 0:
            restart = r <= 0.2 * r0 or (r <= 0.8 * r0 and r > prev) or ks >= 0.36 * (k + 1)
            prev = r
            if restart:
                dx, dy = np.linalg.norm(xp - xr), np.linalg.norm(yp - yr)
                if dx > 1e-10 and dy > 1e-10:
                    target = math.log(dy / dx)
                    esum += math.log(omega) - target
                    omega = math.exp(kp * target + (1.0 - kp) * math.log(omega) - ki * esum)
                tau, sigma = eta / omega, eta * omega
                x, y, ax, aty = xp, yp, axp, atyp
                z0 = (x.copy(), y.copy(), ax.copy(), aty.copy())
                xr, yr = x.copy(), y.copy()
                ks, prev, restarts = 0, math.inf, restarts + 1
                continue
        w1, w0 = (ks + 1) / (ks + 2), 1.0 / (ks + 2)
        x = w1 * (2.0 * xp - x) + w0 * z0[0]
        y = w1 * (2.0 * yp - y) + w0 * z0[1]
        ax = w1 * (2.0 * axp - ax) + w0 * z0[2]
        aty = w1 * (2.0 * atyp - aty) + w0 * z0[3]
        ks += 1
    return xp, yp, restarts


def device_halpern(lp, eta, omega, period, iterations, col_blocks, split, kp=0.0, ki=0.0):
    dev = V.Device()
    L = dev.L
    h = dev.csr(lp.m, lp.n, lp.row_offsets, lp.column_indices, lp.values)
    p = N.PdhgParams()
    p.policy = 2   # VK_PDHG_POLICY_HALPERN
    p.precision = 0
    p.iteration_limit = iterations
    p.relative_kkt_tolerance = 1e-300   # never met: run the whole budget
    p.step_size, p.primal_weight, p.restart_period = eta, omega, period
    p.norm_b, p.norm_c = 1.0, 1.0
    p.use_graphs = 1
    p.col_blocks, p.split_passes = col_blocks, split
    p.primal_weight_gain, p.primal_weight_integral = kp, ki
    arr = [np.ascontiguousarray(a, dtype=np.float64) for a in
           (lp.objective, lp.var_lower, lp.var_upper, lp.row_lower, lp.row_upper, np.ones(lp.m), np.ones(lp.n))]
    s = C.c_void_p()
    N.check(L.vk_hip_pdhg_create(h, *[N._ptr(a) for a in arr], None, None, C.byref(p), C.byref(s)),
            "vk_hip_pdhg_create")
    try:
        status = C.c_int()
        N.check(L.vk_hip_pdhg_run(s, iterations, C.byref(status)), "vk_hip_pdhg_run")
        x, y = np.zeros(lp.n), np.zeros(lp.m)
        N.check(L.vk_hip_pdhg_download(s, N._ptr(x), N._ptr(y)), "vk_hip_pdhg_download")
        st = N.PdhgState()
        N.check(L.vk_hip_pdhg_read_state(s, C.byref(st)), "vk_hip_pdhg_read_state")
        return x, y, st
    finally:
        L.vk_hip_pdhg_destroy(s)
        L.vk_hip_csr_destroy(h)


def step_of(lp):
    A = sp.csr_matrix((lp.values, lp.column_indices, lp.row_offsets), shape=(lp.m, lp.n))
    v = np.random.default_rng(0).standard_normal(lp.n)
    for _ in range(200):
        v = A.T @ (A @ v)
        nv = np.linalg.norm(v)
        v /= nv
    return 0.9 / math.sqrt(nv)


@pytest.mark.parametrize("layout", [(1, 1), (4, 2), (1, 2)])
@pytest.mark.parametrize("gen", [(2, 2000, 5000, 0, 7), (1, 300, 900, 5, 4)])
def test_device_iterates_match_restatement(gen, layout):
    lp = V.generate_lp(*gen)
    blocks, split = layout
    print(f"halpern restatement gen={gen} col_blocks={blocks} split={split}", flush=True)
    eta, omega, period, K = step_of(lp), 1.0, 64, 900
    xr, yr, restarts = halpern_np(lp, eta, omega, period, K)
    x, y, st = device_halpern(lp, eta, omega, period, K, blocks, split)
    assert st.iterations == K and st.status == 2   # VK_PDHG_ITERATION_LIMIT
    assert st.restarts == restarts and restarts >= 2, (st.restarts, restarts)
    sx = max(1.0, float(np.max(np.abs(xr))))
    sy = max(1.0, float(np.max(np.abs(yr))))
    assert np.max(np.abs(x - xr)) <= 1e-9 * sx and np.max(np.abs(y - yr)) <= 1e-9 * sy


@pytest.mark.parametrize("gains", [(0.5, 0.02), (0.3, 0.05), (0.9, 0.0)])
def test_primal_weight_gains_match_restatement(gains):
    # The proportional-integral primal-weight update (the Halpern policy's
    # default integral gain is 0.02) against the restatement.
    kp, ki = gains
    lp = V.generate_lp(2, 2000, 5000, 0, 7)
    eta, omega, period, K = step_of(lp), 1.0, 64, 1500
    xr, yr, restarts = halpern_np(lp, eta, omega, period, K, kp, ki)
    x, y, st = device_halpern(lp, eta, omega, period, K, 1, 1, kp, ki)
    assert st.restarts == restarts and restarts >= 3, (st.restarts, restarts)
    sx = max(1.0, float(np.max(np.abs(xr))))
    sy = max(1.0, float(np.max(np.abs(yr))))
    assert np.max(np.abs(x - xr)) <= 1e-9 * sx and np.max(np.abs(y - yr)) <= 1e-9 * sy


@pytest.mark.parametrize("tol", [1e-6, 1e-8])
@pytest.mark.parametrize("precision", [V.PdhgPrecision.FP64, V.PdhgPrecision.Mixed])
@pytest.mark.parametrize("name", ["uniform_300x900_k5", "uniform_1000x3000_k5"])
def test_halpern_reaches_relative_kkt(name, precision, tol):
    g = HIGHS[name]
    print(f"halpern {name} precision={precision.name} tol={tol}", flush=True)
    lp = V.generate_lp(g["family"], g["m"], g["n"], g["k"], g["seed"])
    r = V.solve_lp(lp, V.Options(pdhg_policy=V.PdhgPolicy.Halpern, pdhg_precision=precision,
                                 relative_kkt_tolerance=tol, iteration_limit=200_000))
    assert r.status == "Optimal", r
    rp, rd, gp, _, _ = O.relative_kkt(lp.as_oracle(), r.x, r.y)
    assert max(rp, rd, gp) <= tol * 1.0001, (rp, rd, gp)
    assert abs(max(rp, rd, gp) - r.kkt_relative) <= 1e-9
    rel_err = abs(r.objective_value - g["objective"]) / max(1.0, abs(g["objective"]))
    print(f"objective relative error vs HiGHS: {rel_err:.3e}", flush=True)
    assert rel_err <= objective_bar(tol), rel_err
    if precision == V.PdhgPrecision.Mixed:
        assert r.low_precision_iterations > 0


@pytest.mark.parametrize("tol", [1e-6, 1e-8])
def test_config3_mixed_refinement_on_the_config1_lp(tol):
    # Config 3: the config-1 LP (1000 x 3000, 5 per column, seed 20261203) in
    # mixed precision, FP32 phase to 1e-4 relative KKT, then the warm-started
    # FP64 phase to 1e-6 and 1e-8; objective against HiGHS.
    g = HIGHS["uniform_1000x3000_k5"]
    lp = V.generate_lp(g["family"], g["m"], g["n"], g["k"], g["seed"])
    r = V.solve_lp(lp, V.Options(pdhg_policy=V.PdhgPolicy.Halpern, pdhg_precision=V.PdhgPrecision.Mixed,
                                 relative_kkt_tolerance=tol, mixed_phase_tolerance=1e-4,
                                 iteration_limit=400_000))
    assert r.status == "Optimal", r
    assert r.low_precision_iterations > 0
    rp, rd, gp, _, _ = O.relative_kkt(lp.as_oracle(), r.x, r.y)
    assert max(rp, rd, gp) <= tol * 1.0001, (rp, rd, gp)
    rel_err = abs(r.objective_value - g["objective"]) / max(1.0, abs(g["objective"]))
    print(f"config 3 tol={tol}: objective relative error vs HiGHS {rel_err:.3e}", flush=True)
    assert rel_err <= objective_bar(tol), rel_err


def test_halpern_tight_tolerance_and_repeat():
    g = HIGHS["uniform_300x900_k5"]
    lp = V.generate_lp(g["family"], g["m"], g["n"], g["k"], g["seed"])
    o = V.Options(pdhg_policy=V.PdhgPolicy.Halpern, relative_kkt_tolerance=1e-8, iteration_limit=400_000)
    a, b = V.solve_lp(lp, o), V.solve_lp(lp, o)
    assert a.status == "Optimal" and a.kkt_relative <= 1e-8
    assert a.iterations == b.iterations and np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)
    rel_err = abs(a.objective_value - g["objective"]) / max(1.0, abs(g["objective"]))
    assert rel_err <= 1e-7, rel_err


def test_halpern_split_layout_certificate():
    lp = V.generate_lp(2, 20_000, 60_000, 0, 3)
    tol = 1e-5
    print(f"halpern split layout m={lp.m} n={lp.n} nnz={lp.nnz}", flush=True)
    r = V.solve_lp(lp, V.Options(pdhg_policy=V.PdhgPolicy.Halpern, relative_kkt_tolerance=tol,
                                 iteration_limit=200_000, spmv_col_blocks=4, spmv_split_passes=2))
    assert r.status == "Optimal", r
    rp, rd, gp, _, _ = O.relative_kkt(lp.as_oracle(), r.x, r.y)
    assert max(rp, rd, gp) <= tol * 1.0001


def test_halpern_refused_on_the_op_by_op_path():
    lp = V.generate_lp(1, 40, 90, 3, 3)
    o = V.Options(pdhg_policy=V.PdhgPolicy.Halpern, relative_kkt_tolerance=1e-6, fused_iterations=False)
    assert V.solve_lp(lp, o).status == "UnsupportedModel"


@pytest.mark.parametrize("precision", [V.PdhgPrecision.FP64, V.PdhgPrecision.Mixed])
def test_halpern_batch_matches_single_solves(precision):
    # The batched kernel runs the same Halpern iteration per instance (its
    # own step and restart state): every instance meets the relative-KKT
    # certificate recomputed by the oracle.
    lps = [V.generate_lp(1, 40, 90, 3, 3), V.generate_lp(4, 0, 0, 4, 9), V.generate_lp(2, 400, 1000, 0, 5)]
    o = V.Options(pdhg_policy=V.PdhgPolicy.Halpern, pdhg_precision=precision, relative_kkt_tolerance=1e-6,
                  iteration_limit=100_000)
    res = V.solve_batch(lps, o)
    for lp, r in zip(lps, res):
        assert r.status == "Optimal", r
        rp, rd, gp, _, _ = O.relative_kkt(lp.as_oracle(), r.x, r.y)
        assert max(rp, rd, gp) <= 1e-6 * 1.0001, (rp, rd, gp)
