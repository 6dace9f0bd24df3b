"""The recentred FP32 Halpern iteration (vk_pdhg_params.recentre; the FP32
phase of a mixed solve under the Halpern policy).

A plain FP32 iteration rounds every output relative to the iterate itself,
and its relative KKT error stalls near 1e-5.  The recentred one stores
corrections to an FP64 base point that moves to the restart point at every
restart, with the LP shifted to the base and rounded to FP32 once per
restart.  Not in the reference, so pinned three ways:
- the device iterates against a plain numpy restatement of the recentred
  iteration, through several restarts: the total points (base + correction)
  agree to FP32 rounding, with the same restart count;
- the restatement itself against the FP64 Halpern restatement of
  test_gpu_halpern (the same iteration in exact-enough arithmetic), so a
  wrong shift of a bound or of the gradient fails there;
- the mixed solve's FP32 phase reaching 1e-6 and 1e-8 relative KKT with the
  oracle's certificate recomputed on the returned point and the objective
  against the independent HiGHS optimum, where the plain FP32 phase stops at
  its 1e-5 floor and leaves the rest to FP64.
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

from test_gpu_halpern import halpern_np, step_of
from test_gpu_pdhg import objective_bar

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HIGHS = json.load(open(os.path.join(GOLDEN, "highs_optima.json")))


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def sign_project(y, rl, ru):
    y = np.where((y > 0.0) & np.isinf(ru), 0.0, y)
    return np.where((y < 0.0) & np.isinf(rl), 0.0, y)


def recentre_np(lp, eta, omega, period, iterations):
    """The device's recentred FP32 Halpern iteration restated with numpy on
    the unscaled LP from x0 = P_X(0), y0 = 0: FP32 values of A (exact
    products summed in FP64), corrections stored in FP32, arithmetic in FP64.
    Returns the total (x+, y+) of the last iteration and the restart count."""
    A = sp.csr_matrix((lp.values, lp.column_indices, lp.row_offsets), shape=(lp.m, lp.n))
    AT = A.T.tocsr()
    A32, AT32 = A.copy(), AT.copy()
    A32.data, AT32.data = f32(A.data), f32(AT.data)
    c, l, u, rl, ru = lp.objective, lp.var_lower, lp.var_upper, lp.row_lower, lp.row_upper
    xb = np.clip(np.zeros(lp.n), l, u)
    yb = sign_project(np.zeros(lp.m), rl, ru)

    def shifted():
        axb = A @ xb
        return f32(c + AT @ yb), f32(l - xb), f32(u - xb), f32(rl - axb), f32(ru - axb)

    g, ls, us, rls, rus = shifted()
    tau, sigma = eta / omega, eta * omega
    zn, zm = np.zeros(lp.n), np.zeros(lp.m)
    x, y, ax, aty = zn.copy(), zm.copy(), zm.copy(), zn.copy()
    z0 = (x.copy(), y.copy(), ax.copy(), aty.copy())
    xr, yr = x.copy(), y.copy()
    ks, r0, prev, restarts = 0, None, math.inf, 0
    xp = yp = None
    for k in range(iterations):
        xp = f32(np.clip((x - tau * g) - tau * aty, ls, us))
        axp = A32 @ xp
        v = y + sigma * (2.0 * axp - ax)
        vt = v + yb
        t = vt / sigma
        yt = np.where(t < rls, np.minimum(vt - sigma * rls, 0.0),
                      np.where(rus < t, np.maximum(vt - sigma * rus, 0.0), 0.0))
        yp = f32(yt - yb)
        atyp = AT32 @ yp
        r = math.sqrt(omega * np.sum((xp - x) ** 2) + np.sum((yp - y) ** 2) / omega)
        if ks == 0:
            r0 = r
        if (k + 1) % period == 0 and k + 1 < iterations:
            restart = r <= 0.2 * r0 or (r <= 0.8 * r0 and r > prev) or ks >= 0.36 * (k + 1)
            prev = r
            if restart:
                dx, dy = np.linalg.norm(xp - xr), np.linalg.norm(yp - yr)
                if dx > 1e-10 and dy > 1e-10:
                    omega = math.exp(0.5 * math.log(dy / dx) + 0.5 * math.log(omega))
                tau, sigma = eta / omega, eta * omega
                # The restart point becomes the base; every correction restarts at 0.
                xb = np.clip(xb + xp, l, u)
                yb = sign_project(yb + yp, rl, ru)
                g, ls, us, rls, rus = shifted()
                x, y, ax, aty = zn.copy(), zm.copy(), zm.copy(), zn.copy()
                z0 = (x.copy(), y.copy(), ax.copy(), aty.copy())
                xr, yr = x.copy(), y.copy()
                ks, prev, restarts = 0, math.inf, restarts + 1
                continue
        w1, w0 = (ks + 1) / (ks + 2), 1.0 / (ks + 2)
        x = f32(w1 * (2.0 * xp - x) + w0 * z0[0])
        y = f32(w1 * (2.0 * yp - y) + w0 * z0[1])
        ax = f32(w1 * (2.0 * axp - ax) + w0 * z0[2])
        aty = f32(w1 * (2.0 * atyp - aty) + w0 * z0[3])
        ks += 1
    return np.clip(xb + xp, l, u), sign_project(yb + yp, rl, ru), restarts


def device_recentre(lp, eta, omega, period, iterations, recentre=1, spmv_panels=0):
    dev = V.Device()
    L = dev.L
    h = dev.csr(lp.m, lp.n, lp.row_offsets, lp.column_indices, lp.values)
    p = N.PdhgParams()
    p.policy = 2   # VK_PDHG_POLICY_HALPERN
    p.precision = 1   # VK_PREC_MIXED
    p.recentre = recentre
    p.iteration_limit = iterations
    p.relative_kkt_tolerance = 1e-300   # never met: run the whole budget
    p.step_size, p.primal_weight, p.restart_period = eta, omega, period
    p.norm_b, p.norm_c = 1.0, 1.0
    p.use_graphs = 1
    p.spmv_panels = spmv_panels
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


GENS = [(2, 2000, 5000, 0, 7), (1, 300, 900, 5, 4)]


@pytest.mark.parametrize("gen", GENS)
def test_restatement_tracks_the_fp64_iteration(gen):
    # The recentred restatement and the FP64 Halpern restatement are the same
    # iteration: only rounding differs (FP32 corrections), so their points
    # stay within FP32 rounding of the corrections, far inside what a wrong
    # shift of c, a bound or the dual step would give.
    lp = V.generate_lp(*gen)
    eta, omega, period, K = step_of(lp), 1.0, 64, 900
    x64, y64, r64 = halpern_np(lp, eta, omega, period, K)
    x32, y32, r32 = recentre_np(lp, eta, omega, period, K)
    assert r32 == r64 and r64 >= 2, (r32, r64)
    sx = max(1.0, float(np.max(np.abs(x64))))
    sy = max(1.0, float(np.max(np.abs(y64))))
    ex, ey = np.max(np.abs(x32 - x64)) / sx, np.max(np.abs(y32 - y64)) / sy
    print(f"recentred restatement vs FP64 gen={gen}: {ex:.2e} {ey:.2e}", flush=True)
    assert ex <= 1e-5 and ey <= 1e-5, (ex, ey)


@pytest.mark.parametrize("gen", GENS)
def test_device_iterates_match_restatement(gen):
    lp = V.generate_lp(*gen)
    print(f"recentred device vs restatement gen={gen}", flush=True)
    eta, omega, period, K = step_of(lp), 1.0, 64, 900
    xr, yr, restarts = recentre_np(lp, eta, omega, period, K)
    x, y, st = device_recentre(lp, eta, omega, period, K)
    assert st.iterations == K and st.status == 2   # VK_PDHG_ITERATION_LIMIT
    assert st.restarts == restarts and restarts >= 2, (st.restarts, restarts)
    sx = max(1.0, float(np.max(np.abs(xr))))
    sy = max(1.0, float(np.max(np.abs(yr))))
    ex, ey = np.max(np.abs(x - xr)) / sx, np.max(np.abs(y - yr)) / sy
    print(f"  max relative |dx| {ex:.2e}  |dy| {ey:.2e}", flush=True)
    assert ex <= 1e-6 and ey <= 1e-6, (ex, ey)


def test_device_runs_repeat_bitwise():
    lp = V.generate_lp(*GENS[0])
    eta, omega = step_of(lp), 1.0
    a = device_recentre(lp, eta, omega, 64, 700)
    b = device_recentre(lp, eta, omega, 64, 700)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("tol", [1e-6, 1e-8])
@pytest.mark.parametrize("name", ["uniform_300x900_k5", "uniform_1000x3000_k5"])
def test_recentred_fp32_phase_reaches_the_target(name, tol):
    # The FP32 phase runs to the final target itself (mixed_phase_tolerance 0,
    # the default); the FP64 phase then only certifies the point.
    g = HIGHS[name]
    print(f"recentred mixed {name} tol={tol}", flush=True)
    lp = V.generate_lp(g["family"], g["m"], g["n"], g["k"], g["seed"])
    r = V.solve_lp(lp, V.Options(pdhg_policy=V.PdhgPolicy.Halpern, pdhg_precision=V.PdhgPrecision.Mixed,
                                 relative_kkt_tolerance=tol, iteration_limit=400_000))
    assert r.status == "Optimal", r
    print(f"  iterations {r.iterations}, FP32 {r.low_precision_iterations}", flush=True)
    assert r.iterations - r.low_precision_iterations <= 64, (r.iterations, r.low_precision_iterations)
    rp, rd, gp, _, _ = O.relative_kkt(lp.as_oracle(), r.x, r.y)
    assert max(rp, rd, gp) <= tol * 1.0001, (rp, rd, gp)
    rel_err = abs(r.objective_value - g["objective"]) / max(1.0, abs(g["objective"]))
    print(f"  objective relative error vs HiGHS: {rel_err:.3e}", flush=True)
    assert rel_err <= objective_bar(tol), rel_err


def test_plain_fp32_phase_hands_over_at_its_floor():
    # With mixed_recentre off the FP32 phase stops at 1e-5 (its floor) and the
    # FP64 phase carries the last decade; the recentred phase covers it.
    g = HIGHS["uniform_1000x3000_k5"]
    lp = V.generate_lp(g["family"], g["m"], g["n"], g["k"], g["seed"])
    base = dict(pdhg_policy=V.PdhgPolicy.Halpern, pdhg_precision=V.PdhgPrecision.Mixed,
                relative_kkt_tolerance=1e-7, iteration_limit=400_000)
    plain = V.solve_lp(lp, V.Options(mixed_recentre=False, **base))
    rec = V.solve_lp(lp, V.Options(**base))
    assert plain.status == rec.status == "Optimal"
    assert plain.iterations - plain.low_precision_iterations > 64
    assert rec.iterations - rec.low_precision_iterations <= 64
    assert abs(plain.objective_value - rec.objective_value) <= 1e-6 * max(1.0, abs(rec.objective_value))


def test_recentred_split_layout_certificate():
    # Column-blocked split layout (the recentring kernels with the tile
    # layouts of a mid-size LP).
    lp = V.generate_lp(2, 20_000, 60_000, 0, 3)
    tol = 1e-6
    print(f"recentred split layout m={lp.m} n={lp.n} nnz={lp.nnz}", flush=True)
    r = V.solve_lp(lp, V.Options(pdhg_policy=V.PdhgPolicy.Halpern, pdhg_precision=V.PdhgPrecision.Mixed,
                                 relative_kkt_tolerance=tol, iteration_limit=200_000, spmv_col_blocks=4,
                                 spmv_split_passes=2))
    assert r.status == "Optimal", r
    assert r.iterations - r.low_precision_iterations <= 64
    rp, rd, gp, _, _ = O.relative_kkt(lp.as_oracle(), r.x, r.y)
    assert max(rp, rd, gp) <= tol * 1.0001


def test_recentre_is_refused_outside_mixed_halpern():
    lp = V.generate_lp(1, 40, 90, 3, 3)
    for policy, precision in ((2, 0), (1, 1), (0, 1)):
        dev = V.Device()
        L = dev.L
        h = dev.csr(lp.m, lp.n, lp.row_offsets, lp.column_indices, lp.values)
        p = N.PdhgParams()
        p.policy, p.precision, p.recentre = policy, precision, 1
        p.iteration_limit, p.step_size, p.primal_weight, p.restart_period = 10, 0.5, 1.0, 64
        p.norm_b = p.norm_c = 1.0
        arr = [np.ascontiguousarray(a, dtype=np.float64) for a in
               (lp.objective, lp.var_lower, lp.var_upper, lp.row_lower, lp.row_upper, np.ones(lp.m), np.ones(lp.n))]
        s = C.c_void_p()
        try:
            rc = L.vk_hip_pdhg_create(h, *[N._ptr(a) for a in arr], None, None, C.byref(p), C.byref(s))
            assert rc != 0 and not s.value, (policy, precision)
        finally:
            L.vk_hip_csr_destroy(h)

