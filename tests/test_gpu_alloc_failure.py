"""Device allocation failures, forced from the host.

vk_hip_debug_fail_allocations lets the first k device allocations from now
succeed and fails every later one with hipErrorOutOfMemory.  Each case
counts the allocations of one call (N), then fails the call at allocations
0, 1, 2, N/3, N/2 and N-1 in turn: the call must report the failure (a solve
returns NumericalFailure, a direct call raises), no kernel may run on a
buffer that was not allocated, and with the hook disarmed the next call
must give the first call's result bitwise.  The reference turns device
errors into exceptions (cuda_ops.cu:35-46) and those into NumericalFailure
(Solver.cpp:81-86); its own tests cannot force one.
"""
import numpy as np
import pytest

import vikalp_amd as V
from qp_data import make_q

pytestmark = pytest.mark.gpu


@pytest.fixture
def dev():
    d = V.Device()
    d.fail_allocations(-1)
    yield d
    d.fail_allocations(-1)


def points(n):
    return sorted({k for k in (0, 1, 2, n // 3, n // 2, n - 1) if 0 <= k < n})


def same(a, b):
    assert a.status == b.status and a.iterations == b.iterations, (a.status, b.status)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)


def check_solves(dev, label, run):
    """run() -> SolveResult or list of them; every forced failure must come
    back as NumericalFailure and leave the next run bitwise unchanged."""
    as_list = lambda r: r if isinstance(r, list) else [r]
    good = as_list(run())
    dev.fail_allocations(-1)
    good = as_list(run())
    made = dev.fail_allocations(-1)
    assert made > 0, f"{label}: no device allocation was counted"
    for k in points(made):
        print(f"alloc failure {label}: fail from allocation {k} of {made}", flush=True)
        dev.fail_allocations(k)
        bad = as_list(run())
        dev.fail_allocations(-1)
        assert all(r.status == "NumericalFailure" for r in bad), (k, [r.status for r in bad])
        assert all(r.message for r in bad)
        again = as_list(run())
        for g, a in zip(good, again):
            same(g, a)


@pytest.mark.parametrize("precision", [V.PdhgPrecision.FP64, V.PdhgPrecision.Mixed])
@pytest.mark.parametrize("panels", [1, 2])
def test_solve_lp(dev, precision, panels):
    lp = V.generate_lp(2, 400, 1000, 0, 5)
    opts = V.Options(iteration_limit=300, pdhg_precision=precision, spmv_panels=panels)
    check_solves(dev, f"solve_lp precision={int(precision)} panels={panels}", lambda: V.solve_lp(lp, opts))


def test_solve_lp_two_phase_halpern(dev):
    """The mixed two-phase solve allocates its FP64 solver after the FP32
    phase has run: late failures land there."""
    lp = V.generate_lp(1, 300, 900, 5, 4)
    opts = V.Options(pdhg_policy=V.PdhgPolicy.Halpern, pdhg_precision=V.PdhgPrecision.Mixed,
                     relative_kkt_tolerance=1e-6, iteration_limit=20_000)
    check_solves(dev, "halpern mixed", lambda: V.solve_lp(lp, opts))


def test_solve_batch(dev):
    lps = [V.generate_lp(1, 40, 90, 3, 3), V.generate_lp(4, 0, 0, 4, 9), V.generate_lp(2, 400, 1000, 0, 5)]
    opts = V.Options(iteration_limit=300)
    check_solves(dev, "solve_batch", lambda: V.solve_batch(lps, opts))


def test_solve_qp(dev):
    lp = V.generate_lp(1, 40, 90, 3, 3)
    q = make_q(lp.n, 7)
    opts = V.Options(iteration_limit=300)
    check_solves(dev, "solve_qp", lambda: V.solve_qp(lp, q, opts))


def test_unrolled(dev):
    lp = V.generate_lp(2, 400, 1000, 0, 11)
    gx = np.linspace(-1.0, 1.0, lp.n)

    def run():
        u = V.UnrolledPdhg(lp, 120, checkpoint_every=40)
        x, y = u.forward()
        return x, y, u.backward(gx, np.zeros(lp.m))

    run()
    dev.fail_allocations(-1)
    good = run()
    made = dev.fail_allocations(-1)
    assert made > 0
    for k in points(made):
        print(f"alloc failure unrolled: fail from allocation {k} of {made}", flush=True)
        dev.fail_allocations(k)
        with pytest.raises((ValueError, RuntimeError)):
            run()
        dev.fail_allocations(-1)
        again = run()
        assert all(np.array_equal(a, b) for a, b in zip(good[:2], again[:2]))
        assert all(np.array_equal(a, b) for a, b in zip(good[2], again[2]))


def test_matrix_handle(dev):
    """vk_hip_csr_create, the FP32 copies (tile tuner in mixed precision)
    and the tile geometry change each fail cleanly; the handle made after
    them computes A x as before."""
    lp = V.generate_lp(2, 400, 1000, 0, 5)
    x = np.random.default_rng(1).normal(size=lp.n)

    def spmv(h):
        src, dst = dev.put(x), dev.alloc(lp.m)
        dev.spmv(h, False, 1.0, src, 0.0, dst)
        return dev.get(dst, lp.m)

    h = dev.csr(lp.m, lp.n, lp.row_offsets, lp.column_indices, lp.values)
    want = spmv(h)
    dev.L.vk_hip_csr_destroy(h)
    dev.free_all()
    dev.fail_allocations(-1)
    h = dev.csr(lp.m, lp.n, lp.row_offsets, lp.column_indices, lp.values)
    made = dev.fail_allocations(-1)
    dev.L.vk_hip_csr_destroy(h)
    for k in points(made):
        print(f"alloc failure csr_create: fail from allocation {k} of {made}", flush=True)
        dev.fail_allocations(k)
        with pytest.raises(RuntimeError):
            dev.csr(lp.m, lp.n, lp.row_offsets, lp.column_indices, lp.values)
        dev.fail_allocations(-1)
    h = dev.csr(lp.m, lp.n, lp.row_offsets, lp.column_indices, lp.values)
    try:
        for k in (0, 1):
            print(f"alloc failure tile tuner / geometry: fail from allocation {k}", flush=True)
            dev.fail_allocations(k)
            with pytest.raises(RuntimeError):
                dev.tune_tiles(h, V.PdhgPrecision.Mixed, repeats=1)
            with pytest.raises(RuntimeError):
                dev.set_tile_geometry(h, 1 + k)
            dev.fail_allocations(-1)
            assert dev.tile_geometry(h) == 0
            assert np.array_equal(spmv(h), want)
        best, _ = dev.tune_tiles(h, V.PdhgPrecision.Mixed, repeats=1)
        dev.set_tile_geometry(h, 0)
        assert np.array_equal(spmv(h), want)
    finally:
        dev.L.vk_hip_csr_destroy(h)
        dev.free_all()
