"""Restarted Halpern PDHG on QPs (solve_qp with PdhgPolicy.Halpern) on the GPU.

Not in the reference (pdqp.cpp runs only its own step rule), so pinned three
ways:
- the device iterates against a numpy restatement on the unscaled QP (the
  LP restatement of test_gpu_halpern with the gradient c + Q x + A'y, Q x of
  the Halpern iterate recomputed rather than carried, and the step rule
  tau (sigma ||A||^2 + ||Q||) = 0.99 with sigma = tau w^2), through several
  restarts, in the fused, split and panel layouts: x and y within 1e-9,
  same restart count;
- the returned point against a KKT certificate computed here with numpy
  (Wolfe dual with the -1/2 x'Qx term; weak duality makes the gap a bound on
  the objective's distance to the optimum), and the objective against an
  independent solve of the same QP by scipy's SLSQP;
- the options that stay refused.

The mixed form (a recentred FP32 phase, then a certifying FP64 phase) is
pinned the same way: its device iterates against a numpy restatement of the
recentred iteration with Q (FP32 corrections to an FP64 base, the gradient
offset c + Q xbase + A'ybase rounded once per restart), within FP32
rounding; the restatement against the FP64 one; and the mixed solves
against the certificate and SLSQP.
"""
import ctypes as C
import math

import numpy as np
import pytest
import scipy.optimize as so
import scipy.sparse as sp

import vikalp_amd as V
from vikalp_amd import native as N
from qp_data import CASES, make_q
from test_gpu_recentre import f32, sign_project

pytestmark = pytest.mark.gpu


def qp_steps(eta, qabs, omega):
    """Column primal steps 1 / (1 / (eta / w) + sum_k |q_jk|) and sigma = eta w."""
    return 1.0 / (1.0 / (eta / omega) + qabs), eta * omega


def halpern_qp_np(lp, q, eta, omega, period, iterations):
    """Restarted Halpern PDHG on the unscaled QP from x0 = P_X(0), y0 = 0:
    the device's method with kp = 1/2, ki = 0 (returns x+, y+, restarts of
    the last iteration)."""
    A = sp.csr_matrix((lp.values, lp.column_indices, lp.row_offsets), shape=(lp.m, lp.n))
    Q = sp.csr_matrix((q[2], q[1], q[0]), shape=(lp.n, lp.n))
    AT = A.T.tocsr()
    c, l, u, rl, ru = lp.objective, lp.var_lower, lp.var_upper, lp.row_lower, lp.row_upper
    qabs = np.asarray(abs(Q).sum(axis=1)).ravel()
    x = np.clip(np.zeros(lp.n), l, u)
    y = np.zeros(lp.m)
    ax, aty = A @ x, AT @ y
    tau, sigma = qp_steps(eta, qabs, omega)
    z0 = (x.copy(), y.copy(), ax.copy(), aty.copy())
    xr, yr = x.copy(), y.copy()
    ks, r0, prev, restarts = 0, None, math.inf, 0
    xp = yp = None
    for k in range(iterations):
        xp = np.clip(x - tau * (aty + (c + Q @ x)), l, u)
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
                    omega = math.exp(0.5 * math.log(dy / dx) + 0.5 * math.log(omega))
                tau, sigma = qp_steps(eta, qabs, omega)
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


def recentre_qp_np(lp, q, eta, omega, period, iterations):
    """The device's recentred FP32 Halpern PDQP restated with numpy on the
    unscaled QP from x0 = P_X(0), y0 = 0: FP32 values of A and Q (exact
    products summed in FP64), corrections stored in FP32, arithmetic in
    FP64, the gradient offset c + Q xbase + A'ybase rounded to FP32 at every
    restart.  Returns the total (x+, y+) of the last iteration and the
    restart count."""
    A = sp.csr_matrix((lp.values, lp.column_indices, lp.row_offsets), shape=(lp.m, lp.n))
    Q = sp.csr_matrix((q[2], q[1], q[0]), shape=(lp.n, lp.n))
    AT = A.T.tocsr()
    A32, AT32, Q32 = A.copy(), AT.copy(), Q.copy()
    A32.data, AT32.data, Q32.data = f32(A.data), f32(AT.data), f32(Q.data)
    qabs = np.asarray(abs(Q).sum(axis=1)).ravel()
    c, l, u, rl, ru = lp.objective, lp.var_lower, lp.var_upper, lp.row_lower, lp.row_upper
    xb = np.clip(np.zeros(lp.n), l, u)
    yb = sign_project(np.zeros(lp.m), rl, ru)

    def shifted():
        axb = A @ xb
        return f32((c + Q @ xb) + AT @ yb), f32(l - xb), f32(u - xb), f32(rl - axb), f32(ru - axb)

    g, ls, us, rls, rus = shifted()
    tau, sigma = qp_steps(eta, qabs, omega)
    zn, zm = np.zeros(lp.n), np.zeros(lp.m)
    x, y, ax, aty = zn.copy(), zm.copy(), zm.copy(), zn.copy()
    z0 = (x.copy(), y.copy(), ax.copy(), aty.copy())
    xr, yr = x.copy(), y.copy()
    ks, r0, prev, restarts = 0, None, math.inf, 0
    xp = yp = None
    for k in range(iterations):
        xp = f32(np.clip(x - tau * (aty + (g + Q32 @ x)), ls, us))
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
                tau, sigma = qp_steps(eta, qabs, omega)
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


def device_halpern_qp(lp, q, eta, omega, period, iterations, col_blocks, split, panels, mixed=False):
    dev = V.Device()
    L = dev.L
    h = dev.csr(lp.m, lp.n, lp.row_offsets, lp.column_indices, lp.values)
    hq = dev.csr(lp.n, lp.n, q[0], q[1], q[2])
    p = N.PdhgParams()
    p.policy = 2   # VK_PDHG_POLICY_HALPERN
    p.precision = 1 if mixed else 0   # VK_PREC_MIXED (recentred) or VK_PREC_FP64
    p.recentre = 1 if mixed else 0
    p.iteration_limit = iterations
    p.relative_kkt_tolerance = 1e-300   # never met: run the whole budget
    p.step_size, p.primal_weight, p.restart_period = eta, omega, period
    p.norm_b, p.norm_c = 1.0, 1.0
    p.use_graphs = 1
    p.col_blocks, p.split_passes, p.spmv_panels = col_blocks, split, panels
    arr = [np.ascontiguousarray(a, dtype=np.float64) for a in
           (lp.objective, lp.var_lower, lp.var_upper, lp.row_lower, lp.row_upper, np.ones(lp.m), np.ones(lp.n))]
    s = C.c_void_p()
    N.check(L.vk_hip_pdqp_create(h, hq, *[N._ptr(a) for a in arr], None, None, C.byref(p), C.byref(s)),
            "vk_hip_pdqp_create")
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
        L.vk_hip_csr_destroy(hq)
        L.vk_hip_csr_destroy(h)


def norm_a(lp):
    A = sp.csr_matrix((lp.values, lp.column_indices, lp.row_offsets), shape=(lp.m, lp.n))
    v = np.random.default_rng(0).standard_normal(lp.n)
    for _ in range(200):
        v = A.T @ (A @ v)
        nv = np.linalg.norm(v)
        v /= nv
    return math.sqrt(nv)


ZIPF, UNIFORM = ((2, 2000, 5000, 0, 7), 11), ((1, 300, 900, 5, 4), 12)
# Iterations of the recentred comparisons: several restarts, and short of
# the point where the uniform QP's iterates reach rounding level (where FP32
# and FP64 rounding could decide a restart differently).
RECENTRED_K = {ZIPF[0]: 450, UNIFORM[0]: 300}


# layout = (col_blocks, split_passes, spmv_panels); the Zipf LP's densest
# row does not fit the panel layout.
@pytest.mark.parametrize("gen,qseed,layout", [
    ZIPF + ((1, 1, 1),), ZIPF + ((4, 2, 1),), ZIPF + ((1, 2, 1),),
    UNIFORM + ((1, 1, 1),), UNIFORM + ((4, 2, 1),), UNIFORM + ((1, 2, 1),), UNIFORM + ((0, 0, 2),)])
def test_device_iterates_match_restatement(gen, qseed, layout):
    lp = V.generate_lp(*gen)
    q = make_q(lp.n, qseed)
    blocks, split, panels = layout
    print(f"halpern qp restatement gen={gen} layout={layout}", flush=True)
    eta, omega, period, K = 0.9 / norm_a(lp), 1.0, 64, 900
    xr, yr, restarts = halpern_qp_np(lp, q, eta, omega, period, K)
    x, y, st = device_halpern_qp(lp, q, eta, omega, period, K, blocks, split, panels)
    assert st.iterations == K and st.status == 2   # VK_PDHG_ITERATION_LIMIT
    assert st.restarts == restarts and restarts >= 2, (st.restarts, restarts)
    # The reported tau is the LP step eta / w the column steps derive from.
    assert abs(st.tau - eta / st.primal_weight) <= 1e-12 * st.tau
    sx = max(1.0, float(np.max(np.abs(xr))))
    sy = max(1.0, float(np.max(np.abs(yr))))
    assert np.max(np.abs(x - xr)) <= 1e-9 * sx and np.max(np.abs(y - yr)) <= 1e-9 * sy


@pytest.mark.parametrize("gen,qseed", [ZIPF, UNIFORM])
def test_recentred_restatement_tracks_the_fp64_iteration(gen, qseed):
    lp = V.generate_lp(*gen)
    q = make_q(lp.n, qseed)
    eta, omega, period, K = 0.9 / norm_a(lp), 1.0, 64, RECENTRED_K[gen]
    x64, y64, r64 = halpern_qp_np(lp, q, eta, omega, period, K)
    x32, y32, r32 = recentre_qp_np(lp, q, eta, omega, period, K)
    assert r32 == r64 and r64 >= 2, (r32, r64)
    sx = max(1.0, float(np.max(np.abs(x64))))
    sy = max(1.0, float(np.max(np.abs(y64))))
    ex, ey = np.max(np.abs(x32 - x64)) / sx, np.max(np.abs(y32 - y64)) / sy
    print(f"recentred QP restatement vs FP64 gen={gen}: {ex:.2e} {ey:.2e}", flush=True)
    assert ex <= 1e-5 and ey <= 1e-5, (ex, ey)


@pytest.mark.parametrize("gen,qseed,layout", [ZIPF + ((1, 1, 1),), ZIPF + ((1, 2, 1),), UNIFORM + ((0, 0, 2),)])
def test_recentred_device_iterates_match_restatement(gen, qseed, layout):
    lp = V.generate_lp(*gen)
    q = make_q(lp.n, qseed)
    print(f"recentred halpern qp gen={gen} layout={layout}", flush=True)
    eta, omega, period, K = 0.9 / norm_a(lp), 1.0, 64, RECENTRED_K[gen]
    xr, yr, restarts = recentre_qp_np(lp, q, eta, omega, period, K)
    x, y, st = device_halpern_qp(lp, q, eta, omega, period, K, *layout, mixed=True)
    assert st.iterations == K and st.status == 2   # VK_PDHG_ITERATION_LIMIT
    assert st.restarts == restarts and restarts >= 2, (st.restarts, restarts)
    sx = max(1.0, float(np.max(np.abs(xr))))
    sy = max(1.0, float(np.max(np.abs(yr))))
    ex, ey = np.max(np.abs(x - xr)) / sx, np.max(np.abs(y - yr)) / sy
    print(f"  max relative |dx| {ex:.2e}  |dy| {ey:.2e}", flush=True)
    assert ex <= 1e-6 and ey <= 1e-6, (ex, ey)


def qp_certificate(lp, q, x, y):
    """(rel_primal, rel_dual, rel_gap, primal_obj, dual_obj) of a QP in the
    reference's dual sign: g = c + Q x + A'y, primal c'x + 1/2 x'Qx, Wolfe
    dual -1/2 x'Qx + sum_j min(g_j l_j, g_j u_j) - sum_i (y_i+ ru_i - y_i- rl_i)
    (-inf, with gap 1, when a term meets an infinite bound)."""
    A = sp.csr_matrix((lp.values, lp.column_indices, lp.row_offsets), shape=(lp.m, lp.n))
    Q = sp.csr_matrix((q[2], q[1], q[0]), shape=(lp.n, lp.n))
    c, l, u, rl, ru = lp.objective, lp.var_lower, lp.var_upper, lp.row_lower, lp.row_upper
    ax, qx = A @ x, Q @ x
    g = c + qx + A.T @ y
    bnorm = np.sqrt(np.sum(np.where(np.isfinite(rl) & np.isfinite(ru), np.maximum(abs(rl), abs(ru)),
                                    np.where(np.isfinite(rl), rl, np.where(np.isfinite(ru), ru, 0.0))) ** 2))
    rp = np.linalg.norm(ax - np.clip(ax, rl, ru)) / (1.0 + bnorm)
    rd = np.linalg.norm(x - np.clip(x - g, l, u)) / (1.0 + np.linalg.norm(c))
    xqx = float(x @ qx)
    pobj = float(c @ x) + 0.5 * xqx
    with np.errstate(invalid="ignore"):
        gterm = np.where(g > 0, g * l, np.where(g < 0, g * u, 0.0))
        yterm = np.where(y > 0, y * ru, np.where(y < 0, y * rl, 0.0))
    if not (np.all(np.isfinite(gterm)) and np.all(np.isfinite(yterm))):
        return rp, rd, 1.0, pobj, -math.inf
    dobj = -0.5 * xqx + float(np.sum(gterm)) - float(np.sum(yterm))
    gap = abs(pobj - dobj) / (1.0 + abs(pobj) + abs(dobj))
    return rp, rd, gap, pobj, dobj


def slsqp_optimum(lp, q):
    """The QP's optimum by scipy's SLSQP (an active-set SQP method, independent
    of PDHG) on the dense data; small cases only."""
    A = sp.csr_matrix((lp.values, lp.column_indices, lp.row_offsets), shape=(lp.m, lp.n)).toarray()
    Q = sp.csr_matrix((q[2], q[1], q[0]), shape=(lp.n, lp.n)).toarray()
    c = lp.objective
    x0 = np.clip(np.zeros(lp.n), lp.var_lower, lp.var_upper)
    # Equality rows and range / one-sided rows as separate constraints.
    eq = lp.row_lower == lp.row_upper
    cons = []
    if eq.any():
        cons.append(so.LinearConstraint(A[eq], lp.row_lower[eq], lp.row_upper[eq]))
    if (~eq).any():
        cons.append(so.LinearConstraint(A[~eq], lp.row_lower[~eq], lp.row_upper[~eq]))
    r = so.minimize(lambda x: c @ x + 0.5 * x @ Q @ x, x0, jac=lambda x: c + Q @ x, method="SLSQP",
                    bounds=so.Bounds(lp.var_lower, lp.var_upper), constraints=cons,
                    options={"ftol": 1e-12, "maxiter": 5000})
    assert r.success, r.message
    return float(r.fun)


@pytest.mark.parametrize("tol", [1e-6, 1e-8])
@pytest.mark.parametrize("precision", [V.PdhgPrecision.FP64, V.PdhgPrecision.Mixed])
@pytest.mark.parametrize("case", [0, 1, 2])
def test_halpern_qp_reaches_relative_kkt(case, precision, tol):
    name, gen, qseed, _ = CASES[case]
    print(f"halpern qp {name} precision={precision.name} tol={tol}", flush=True)
    lp = V.generate_lp(*gen)
    q = make_q(lp.n, qseed)
    r = V.solve_qp(lp, q, V.Options(pdhg_policy=V.PdhgPolicy.Halpern, pdhg_precision=precision,
                                    relative_kkt_tolerance=tol, iteration_limit=200_000))
    assert r.status == "Optimal", r
    assert r.solver == "PDQP"
    if precision == V.PdhgPrecision.Mixed:
        # The recentred FP32 phase reaches the target; FP64 only certifies.
        assert r.low_precision_iterations > 0
        print(f"  FP32 {r.low_precision_iterations} + FP64 {r.iterations - r.low_precision_iterations} "
              f"iterations", flush=True)
        assert r.iterations - r.low_precision_iterations <= 64
    rp, rd, gap, pobj, dobj = qp_certificate(lp, q, r.x, r.y)
    assert max(rp, rd, gap) <= tol * 1.0001, (rp, rd, gap)
    assert abs(max(rp, rd, gap) - r.kkt_relative) <= 1e-9
    assert abs(r.objective_value - (pobj + lp.objective_offset)) <= 1e-9 * max(1.0, abs(pobj))
    if case == 1:
        # SLSQP fails its line search on this Zipf QP: the certificate above
        # (weak duality) is the pin.
        return
    opt = slsqp_optimum(lp, q)
    rel_err = abs(pobj - opt) / max(1.0, abs(opt))
    print(f"objective {pobj:.12g} vs SLSQP {opt:.12g}: relative error {rel_err:.3e}, "
          f"{r.iterations} iterations", flush=True)
    assert rel_err <= 50 * tol, rel_err


def test_halpern_qp_repeats_bitwise():
    name, gen, qseed, _ = CASES[1]
    lp = V.generate_lp(*gen)
    q = make_q(lp.n, qseed)
    o = V.Options(pdhg_policy=V.PdhgPolicy.Halpern, relative_kkt_tolerance=1e-6, iteration_limit=100_000)
    a = V.solve_qp(lp, q, o)
    b = V.solve_qp(lp, q, o)
    assert a.status == "Optimal" and a.iterations == b.iterations
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)
    # For the record: the reference policy after as many iterations.
    ref = V.solve_qp(lp, q, V.Options(iteration_limit=a.iterations))
    rp, rd, gap, _, _ = qp_certificate(lp, q, ref.x, ref.y)
    print(f"{name}: Halpern {a.iterations} iterations to 1e-6; reference policy after as many: "
          f"relative KKT {max(rp, rd, gap):.3e}", flush=True)


def test_halpern_qp_refusals():
    name, gen, qseed, opts = CASES[2]
    lp = V.generate_lp(*gen)
    q = make_q(lp.n, qseed)
    H, M = V.PdhgPolicy.Halpern, V.PdhgPrecision.Mixed
    for bad in (dict(pdhg_policy=H, pdhg_precision=M, mixed_recentre=False),
                dict(pdhg_precision=M),
                dict(pdhg_policy=H, fused_iterations=False),
                dict(pdhg_policy=V.PdhgPolicy.Adaptive, relative_kkt_tolerance=1e-6)):
        assert V.solve_qp(lp, q, V.Options(**bad)).status == "UnsupportedModel", bad


@pytest.mark.parametrize("precision", [V.PdhgPrecision.FP64, V.PdhgPrecision.Mixed])
@pytest.mark.parametrize("make", ["small", "branched"])
def test_halpern_qp_detects_primal_infeasibility(make, precision):
    # The feasible set does not involve Q, so an infeasible LP stays
    # infeasible with any convex Q, and the device's Farkas test (which reads
    # A'y only) applies unchanged: the returned row duals must be a Farkas ray
    # of the original data.
    from test_gpu_infeasible import branched_infeasible, farkas_d0, small_infeasible
    lp = small_infeasible() if make == "small" else branched_infeasible()
    q = make_q(lp.n, 21)
    print(f"halpern qp infeasible {make} precision={precision.name}", flush=True)
    r = V.solve_qp(lp, q, V.Options(pdhg_policy=V.PdhgPolicy.Halpern, pdhg_precision=precision,
                                    relative_kkt_tolerance=1e-6, iteration_limit=200_000))
    assert r.status == "Infeasible", (r.status, r.message, r.iterations)
    assert r.iterations < 200_000
    d0 = farkas_d0(lp, r.y)
    assert np.isfinite(d0) and d0 > 0.0, d0
