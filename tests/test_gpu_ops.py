"""Kernel C ABI (include/vikalp_hip.h) against the CPU oracle's op
semantics (cpu_ops.cpp), on the GPU.  Tolerance 1e-10 (the reference's
CPU/GPU consistency bound, tests/test_cuda_consistency.cpp:23) for sums;
bitwise for max-norms, projections and repeated runs."""
import numpy as np
import pytest

import vikalp_amd as V
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def random_csr(rng, m, n, density, long_rows=()):
    rp, ci, v = [0], [], []
    for r in range(m):
        d = 1.0 if r in long_rows else density
        cols = np.nonzero(rng.random(n) < d)[0]
        ci.extend(cols.tolist())
        v.extend(rng.normal(size=len(cols)).tolist())
        rp.append(len(ci))
    return np.array(rp, np.int64), np.array(ci, np.int64), np.array(v)


CASES = [
    # m, n, density, long rows  (long rows exceed one 512-nonzero tile)
    (200, 300, 0.05, ()),
    (5000, 3000, 0.002, (7, 4000)),
    (3, 9000, 0.01, (1,)),
    (20000, 50, 0.3, ()),          # many short rows: tiles hit the 64-row cap
    (64, 64, 0.0, ()),             # all rows empty
]


@pytest.mark.parametrize("m,n,density,long_rows", CASES)
def test_spmv_both_directions(m, n, density, long_rows):
    print(f"spmv case m={m} n={n} density={density}", flush=True)
    rng = np.random.default_rng(m + n)
    rp, ci, v = random_csr(rng, m, n, density, long_rows)
    dev = V.Device()
    try:
        h = dev.csr(m, n, rp, ci, v)
        for t in (0, 1):
            nx, ny = (m, n) if t else (n, m)
            x, y0 = rng.normal(size=nx), rng.normal(size=ny)
            dx, dy = dev.put(x), dev.put(y0)
            dev.spmv(h, t, 1.25, dx, -0.75, dy)
            got = dev.get(dy, ny)
            exp = y0.copy()
            O.spmv(rp, ci, v, m, n, 1.25, x, -0.75, exp, bool(t))
            np.testing.assert_allclose(got, exp, rtol=0, atol=1e-10 * max(1.0, np.abs(exp).max()))
            # deterministic: the same call again is bitwise identical
            dy2 = dev.put(y0)
            dev.spmv(h, t, 1.25, dx, -0.75, dy2)
            assert np.array_equal(dev.get(dy2, ny), got)
        dev.L.vk_hip_csr_destroy(h)
    finally:
        dev.free_all()


def transpose_cases():
    rng = np.random.default_rng(5)
    out = [("random", 700, 400, *random_csr(rng, 700, 400, 0.02, (3,)))]
    # a dense column, empty columns past the last used one, no rows at all
    rp = np.arange(0, 2 * 3000 + 1, 2, dtype=np.int64)
    ci = np.stack([np.zeros(3000, np.int64), np.arange(3000, dtype=np.int64) % 97 + 1], 1).ravel()
    out.append(("dense column", 3000, 200, rp, ci, rng.normal(size=ci.size)))
    out.append(("no rows", 0, 5, np.zeros(1, np.int64), np.zeros(0, np.int64), np.zeros(0)))
    out.append(("no entries", 10, 7, np.zeros(11, np.int64), np.zeros(0, np.int64), np.zeros(0)))
    # 2^20 + 3 columns: the sort needs 21 key bits
    lp = V.generate_lp(1, 50_000, (1 << 20) + 3, 2, 8)
    out.append(("wide", lp.m, lp.n, lp.row_offsets, lp.column_indices, lp.values))
    return out


@pytest.mark.parametrize("case", transpose_cases(), ids=lambda c: c[0])
def test_transpose_is_exact(case):
    # The stored transpose lists each column's rows in increasing order,
    # with tperm-gathered values: compared with scipy's exact transpose.
    name, m, n, rp, ci, v = case
    print(f"transpose {name} m={m} n={n} nnz={len(v)}", flush=True)
    dev = V.Device()
    import ctypes as C
    import scipy.sparse as sp
    h = dev.csr(m, n, rp, ci, v)
    nnz = len(v)
    trp, tri, tv = np.zeros(n + 1, np.int32), np.zeros(max(1, nnz), np.int32), np.zeros(max(1, nnz))
    rc = dev.L.vk_hip_csr_transpose_export(h, trp.ctypes.data_as(C.POINTER(C.c_int32)),
                                          tri.ctypes.data_as(C.POINTER(C.c_int32)),
                                          tv.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == 0
    At = sp.csr_matrix((v, ci, rp), shape=(m, n)).T.tocsr()
    At.sort_indices()
    assert np.array_equal(trp, At.indptr) and np.array_equal(tri[:nnz], At.indices)
    assert np.array_equal(tv[:nnz], At.data)
    dev.L.vk_hip_csr_destroy(h)
    dev.free_all()


def test_vector_ops_and_reductions():
    rng = np.random.default_rng(6)
    dev = V.Device()
    try:
        for n in (1, 257, 100_003):
            x, y, lo, hi = rng.normal(size=n), rng.normal(size=n), -rng.random(n), rng.random(n)
            dx, dy, dl, dh = dev.put(x), dev.put(y), dev.put(lo), dev.put(hi)
            assert dev.absmax(dx, n) == np.abs(x).max()
            d = dev.dot(dx, dy, n)
            assert abs(d - O.lib().orc_dot(n, O._ptr(x), O._ptr(y))) <= 1e-10 * max(1.0, np.abs(x * y).sum())
            assert dev.dot(dx, dy, n) == d      # repeatable
            assert dev.L.vk_hip_axpby(n, 2.0, dx, -0.5, dy, None) == 0
            exp = y.copy()
            O.lib().orc_axpby(n, 2.0, O._ptr(x), -0.5, O._ptr(exp))
            assert np.array_equal(dev.get(dy, n), exp)
            assert dev.L.vk_hip_project_box(n, dx, dl, dh, None) == 0
            exp = x.copy()
            O.lib().orc_project_box(n, O._ptr(exp), O._ptr(lo), O._ptr(hi))
            assert np.array_equal(dev.get(dx, n), exp)
            assert dev.L.vk_hip_fill(dx, n, 3.5, None) == 0
            assert (dev.get(dx, n) == 3.5).all()
        assert dev.absmax(dx, 0) == 0.0
    finally:
        dev.free_all()


def test_bad_arguments_are_rejected():
    dev = V.Device()
    import ctypes as C
    h = C.c_void_p()
    rp = np.array([0, 2], np.int64)
    ci = np.array([1, 0], np.int64)       # not increasing
    rc = dev.L.vk_hip_csr_create(1, 3, rp.ctypes.data_as(C.POINTER(C.c_int64)),
                                 ci.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(h))
    assert rc == -1 and b"ordered" in dev.L.vk_hip_last_error()
    assert dev.L.vk_hip_fill(None, 5, 1.0, None) == -1


def test_norm_estimate_ignores_stale_memory():
    # Regression: the power iteration's scratch vector comes from fresh device
    # memory and k_spmv forms alpha A x + beta y with beta = 0 literally, so a
    # NaN left there by an earlier allocation must not reach the estimate.
    import ctypes as C
    rng = np.random.default_rng(11)
    m, n = 700, 1500
    rp, ci, v = random_csr(rng, m, n, 0.01)
    dev = V.Device()
    for _ in range(4):                        # stale NaNs in blocks of the scratch size
        p = dev.alloc(m)
        dev.L.vk_hip_fill(p, m, float("nan"), None)
    dev.L.vk_hip_sync(None)
    dev.free_all()
    h = dev.csr(m, n, rp, ci, v)
    est = C.c_double()
    assert dev.L.vk_hip_csr_norm2_estimate(h, 200, C.byref(est)) == 0
    dense = np.zeros((m, n))
    for r in range(m):
        dense[r, ci[rp[r]:rp[r + 1]]] = v[rp[r]:rp[r + 1]]
    exact = np.linalg.norm(dense, 2)
    # sqrt(|A'A v|) with |v| = 1 never exceeds sigma_max and approaches it from below
    assert np.isfinite(est.value) and 0.95 * exact <= est.value <= exact * (1 + 1e-12), (est.value, exact)
    dev.L.vk_hip_csr_destroy(h)
