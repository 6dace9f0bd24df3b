"""Unrolled PDHG and its adjoint on the GPU (solver::UnrolledPdhg through the
vk_unrolled_* C ABI).

Against the CPU oracle (orc_unrolled, itself pinned by the reference
trajectory, central differences and the duality identity): the forward
(x_K, y_K) to 1e-9 and the gradients to 1e-8, both relative to the largest
entry (the SpMV summation order differs).  On the GPU alone: central
differences of the loss L = c'x_K + |x_K - x_t|^2 / 2 along random unit
directions (h = 1e-6) within 1e-4 relative, bitwise-equal gradients for every
checkpoint period (the recomputed active sets equal the forward pass's), and
d obj/dc = x*, d obj/db = -y* at convergence.
"""
import math

import numpy as np
import pytest

import vikalp_amd as V
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(1.0, np.max(np.abs(b))))


@pytest.mark.parametrize("fam,m,n,k,seed,K", [(2, 400, 1000, 0, 11, 300), (1, 300, 900, 5, 4, 500),
                                              (4, 0, 0, 4, 9, 200)])
def test_forward_and_gradients_match_oracle(fam, m, n, k, seed, K):
    lp = V.generate_lp(fam, m, n, k, seed)
    print(f"unrolled family={fam} m={lp.m} n={lp.n} K={K}", flush=True)
    ol = lp.as_oracle()
    u = V.UnrolledPdhg(lp, K, checkpoint_every=50)
    x, y = u.forward()
    xr, yr, step = O.unrolled(ol, K)
    assert u.info()["step_size"] == pytest.approx(step, rel=1e-15)
    assert rel(x, xr) <= 1e-9 and rel(y, yr) <= 1e-9
    xt = np.random.default_rng(seed).uniform(0, 10, lp.n)
    gx = lp.objective + x - xt
    gy = np.random.default_rng(seed + 1).normal(size=lp.m)     # exercise the dual seed too
    gc, grl, gru = u.backward(gx, gy)
    _, _, _, gc_r, grl_r, gru_r = O.unrolled(ol, K, grad_x=lp.objective + xr - xt, grad_y=gy)
    assert rel(gc, gc_r) <= 1e-8 and rel(grl, grl_r) <= 1e-8 and rel(gru, gru_r) <= 1e-8


def test_checkpoint_period_does_not_change_gradients():
    lp = V.generate_lp(2, 400, 1000, 0, 5)
    K = 173
    out = []
    for every in (1, 7, 50, K, 10 * K):
        u = V.UnrolledPdhg(lp, K, checkpoint_every=every)
        x, y = u.forward()
        out.append((x, y) + u.backward(lp.objective + x, None))
    for o in out[1:]:
        for a, b in zip(out[0], o):
            assert np.array_equal(a, b)


def test_central_differences():
    lp = V.generate_lp(1, 2000, 5000, 5, 17)
    K = 300
    rng = np.random.default_rng(5)
    xt = rng.uniform(0, 10, lp.n)
    u = V.UnrolledPdhg(lp, K, checkpoint_every=50)
    x, _ = u.forward()
    gc, grl, gru = u.backward(lp.objective + x - xt)
    gc = gc + x
    gb = grl + gru                                           # equality rows: rl = ru = b

    v = V.UnrolledPdhg(lp, K, checkpoint_every=K)

    def loss(dc, db, h):
        c = lp.objective + h * dc
        v.set_data(c, lp.row_lower + h * db, lp.row_upper + h * db)
        xk, _ = v.forward()
        return float(c @ xk + 0.5 * np.sum((xk - xt) ** 2))

    worst = 0.0
    for _ in range(8):
        dc, db = rng.normal(size=lp.n), rng.normal(size=lp.m)
        norm = math.sqrt(dc @ dc + db @ db)
        dc, db = dc / norm, db / norm
        fd = (loss(dc, db, 1e-6) - loss(dc, db, -1e-6)) / 2e-6
        an = gc @ dc + gb @ db
        worst = max(worst, abs(fd - an) / max(abs(an), 1e-12))
    assert worst <= 1e-4, worst


def test_set_data_equals_a_fresh_solver():
    lp = V.generate_lp(2, 400, 1000, 0, 5)
    rng = np.random.default_rng(8)
    c2 = lp.objective + rng.normal(size=lp.n)
    shift = rng.normal(size=lp.m)
    rl2 = lp.row_lower + np.where(np.isfinite(lp.row_lower), shift, 0.0)
    ru2 = lp.row_upper + np.where(np.isfinite(lp.row_upper), shift, 0.0)
    u = V.UnrolledPdhg(lp, 120, checkpoint_every=40)
    u.forward()
    u.set_data(c2, rl2, ru2)
    with pytest.raises(RuntimeError):
        u.backward(c2)                                      # new data needs a new forward
    x, y = u.forward()
    g = u.backward(c2 + x)
    lp2 = V.LP(lp.m, lp.n, lp.row_offsets, lp.column_indices, lp.values, c2, rl2, ru2, lp.var_lower,
               lp.var_upper)
    f = V.UnrolledPdhg(lp2, 120, checkpoint_every=40)
    x2, y2 = f.forward()
    g2 = f.backward(c2 + x2)
    assert np.array_equal(x, x2) and np.array_equal(y, y2)
    for a, b in zip(g, g2):
        assert np.array_equal(a, b)
    with pytest.raises(ValueError):                         # NaN bound
        u.set_data(c2, rl2, np.full(lp.m, np.nan))
    opened = ru2.copy()
    opened[np.isfinite(opened)] = np.inf                    # a finite bound may not become infinite
    with pytest.raises(ValueError):
        u.set_data(c2, rl2, opened)


def test_duality_exact_at_convergence():
    lp = V.generate_lp(1, 40, 90, 3, 3)
    u = V.UnrolledPdhg(lp, 200_000, checkpoint_every=1000)
    x, y = u.forward()
    rp, rd, gp, _, _ = O.relative_kkt(lp.as_oracle(), x, y)
    assert max(rp, rd, gp) <= 1e-10
    gc, grl, gru = u.backward(lp.objective)                 # obj = c'x_K
    assert np.max(np.abs(gc)) <= 1e-8                       # d obj/dc = x_K + gc = x*
    assert np.max(np.abs((grl + gru) + y)) <= 1e-8          # d obj/db = -y*


def test_edge_cases_and_errors():
    lp = V.generate_lp(1, 40, 90, 3, 3)
    u = V.UnrolledPdhg(lp, 0, checkpoint_every=5)
    with pytest.raises(RuntimeError):
        u.backward(lp.objective)                            # forward first
    x, y = u.forward()
    ref, _, _ = O.unrolled(lp.as_oracle(), 0)
    assert np.array_equal(x, ref) and not np.any(y)          # projected start, y = 0
    gc, grl, gru = u.backward(lp.objective)
    assert not np.any(gc) and not np.any(grl) and not np.any(gru)
    with pytest.raises(ValueError):
        V.UnrolledPdhg(lp, 10, checkpoint_every=0)
    with pytest.raises(ValueError):
        V.UnrolledPdhg(lp, -1)
    bad = V.LP(1, 1, np.array([0, 1]), np.array([0]), np.array([1.0]), np.array([1.0]),
               np.array([0.0]), np.array([1.0]), np.array([2.0]), np.array([1.0]))   # lower > upper
    with pytest.raises(ValueError):
        V.UnrolledPdhg(bad, 10)
