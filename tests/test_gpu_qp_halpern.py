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


def device_halpern_qp(lp, q, eta, omega, period, iterations, col_blocks, split, panels):
    dev = V.Device()
    L = dev.L
    h = dev.csr(lp.m, lp.n, lp.row_offsets, lp.column_indices, lp.values)
    hq = dev.csr(lp.n, lp.n, q[0], q[1], q[2])
    p = N.PdhgParams()
    p.policy = 2   # VK_PDHG_POLICY_HALPERN
    p.precision = 0
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
    r = so.minimize(lambda x: c @ x + 0.5 * x @ Q @ x, x0, jac=lambda x: c + Q @ x, method="SLSQP",
                    bounds=so.Bounds(lp.var_lower, lp.var_upper),
                    constraints=[so.LinearConstraint(A, lp.row_lower, lp.row_upper)],
                    options={"ftol": 1e-12, "maxiter": 5000})
    assert r.success, r.message
    return float(r.fun)


@pytest.mark.parametrize("tol", [1e-6, 1e-8])
@pytest.mark.parametrize("case", [0, 2])
def test_halpern_qp_reaches_relative_kkt(case, tol):
    name, gen, qseed, _ = CASES[case]
    print(f"halpern qp {name} tol={tol}", flush=True)
    lp = V.generate_lp(*gen)
    q = make_q(lp.n, qseed)
    r = V.solve_qp(lp, q, V.Options(pdhg_policy=V.PdhgPolicy.Halpern, relative_kkt_tolerance=tol,
                                    iteration_limit=200_000))
    assert r.status == "Optimal", r
    assert r.solver == "PDQP"
    rp, rd, gap, pobj, dobj = qp_certificate(lp, q, r.x, r.y)
    assert max(rp, rd, gap) <= tol * 1.0001, (rp, rd, gap)
    assert abs(max(rp, rd, gap) - r.kkt_relative) <= 1e-9
    assert abs(r.objective_value - (pobj + lp.objective_offset)) <= 1e-9 * max(1.0, abs(pobj))
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
    H = V.PdhgPolicy.Halpern
    for bad in (dict(pdhg_policy=H, pdhg_precision=V.PdhgPrecision.Mixed),
                dict(pdhg_policy=H, fused_iterations=False),
                dict(pdhg_policy=V.PdhgPolicy.Adaptive, relative_kkt_tolerance=1e-6)):
        assert V.solve_qp(lp, q, V.Options(**bad)).status == "UnsupportedModel", bad
