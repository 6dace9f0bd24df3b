"""vk_hip_reduce_dot_noalloc (the dot product for callers that cannot
allocate: its partials live in a fixed device array) against
vk_hip_reduce_dot, through the C ABI: bitwise equal results, on two calls in
a row, the second with a smaller input (the array keeps the first call's
partials, which the second must not read)."""
import ctypes as C

import numpy as np
import pytest

import vikalp_amd as V

pytestmark = pytest.mark.gpu


def dot(dev, fn, p, q, n):
    r = C.c_double()
    V.native.check(getattr(dev.L, fn)(n, p, q, C.byref(r), None), fn)
    return r.value


def test_noalloc_dot_matches_dot_on_two_calls_the_second_smaller():
    rng = np.random.default_rng(3)
    dev = V.Device()
    try:
        results = []
        for n in (2_000_003, 777):
            print(f"noalloc dot n={n}", flush=True)
            x, y = rng.normal(size=n), rng.normal(size=n)
            p, q = dev.put(x), dev.put(y)
            got = dot(dev, "vk_hip_reduce_dot_noalloc", p, q, n)
            ref = dot(dev, "vk_hip_reduce_dot", p, q, n)
            assert got == ref, (n, got, ref)
            assert abs(got - float(x @ y)) <= 1e-12 * float(np.abs(x) @ np.abs(y))
            results.append(got)
        assert dot(dev, "vk_hip_reduce_dot_noalloc", None, None, 0) == 0.0
        # the smaller call again, after the larger one: unchanged
        n = 2_000_003
        x, y = rng.normal(size=n), rng.normal(size=n)
        p, q = dev.put(x), dev.put(y)
        big = dot(dev, "vk_hip_reduce_dot_noalloc", p, q, n)
        assert big == dot(dev, "vk_hip_reduce_dot", p, q, n)
    finally:
        dev.free_all()


def test_noalloc_dot_rejects_bad_arguments():
    L = V.native.hip_lib()
    r = C.c_double()
    assert L.vk_hip_reduce_dot_noalloc(-1, None, None, C.byref(r), None) != 0
    assert L.vk_hip_reduce_dot_noalloc(5, None, None, C.byref(r), None) != 0
    assert L.vk_hip_reduce_dot_noalloc(5, None, None, None, None) != 0
