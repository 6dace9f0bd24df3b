"""Row-tile geometries and their tuner, on the GPU.

The tile kernels cut rows into tiles of whole rows; the geometry bounds a
tile's nonzeros and rows and sets the row length above which a row is a tile
of its own (vk_hip_csr_set_tile_geometry).  The geometry changes which rows
share a tile and so the order of each row's sum: every geometry must give
the oracle's solve under the reference policy (the usual bar), and the
tuner's choice must be named in the result and reproduce bitwise with that
geometry fixed.
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import vikalp_amd as V
from oracle import oracle as O

from test_gpu_pdhg import close

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
TRAJ = json.load(open(os.path.join(GOLDEN, "reference_trajectories.json")))
NAMES = {1: "256 nonzeros x 64 rows", 2: "512 nonzeros x 32 rows", 3: "rows over 64 alone"}


def ragged_csr(seed=3):
    """Rows of every length class the geometries split on: empty, short,
    around 32 / 64 / 256 / 512 nonzeros and longer, in a shuffled order."""
    rng = np.random.default_rng(seed)
    n = 2000
    lengths = [0, 1, 2, 31, 32, 33, 63, 64, 65, 255, 256, 257, 511, 512, 513, 1000]
    lengths += list(rng.integers(0, 40, size=280))
    lengths = rng.permutation(lengths)
    rp = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
    ci = np.concatenate([np.sort(rng.choice(n, size=k, replace=False)) for k in lengths]).astype(np.int64)
    return len(lengths), n, rp, ci, rng.normal(size=len(ci))


@pytest.mark.parametrize("geometry", [0, 1, 2, 3])
def test_spmv_with_each_geometry(geometry):
    m, n, rp, ci, v = ragged_csr()
    A = sp.csr_matrix((v, ci, rp), shape=(m, n))
    dev = V.Device()
    h = dev.csr(m, n, rp, ci, v)
    try:
        dev.set_tile_geometry(h, geometry)
        assert dev.tile_geometry(h) == geometry
        rng = np.random.default_rng(geometry)
        x, y = rng.normal(size=n), rng.normal(size=m)
        for transpose, vec, ref in ((False, x, A @ x), (True, y, A.T @ y)):
            src, dst = dev.put(vec), dev.alloc(len(ref))
            dev.spmv(h, transpose, 1.0, src, 0.0, dst)
            got = dev.get(dst, len(ref))
            scale = np.abs(A if not transpose else A.T) @ np.abs(vec)
            assert np.all(np.abs(got - ref) <= 1e-12 * scale + 1e-300), (geometry, transpose)
    finally:
        dev.L.vk_hip_csr_destroy(h)
        dev.free_all()


@pytest.mark.parametrize("name", ["zipf_400x1000_it1500", "block_T4_it500", "uniform_60x150_k4_it400"])
@pytest.mark.parametrize("geometry", [1, 2, 3])
def test_reference_policy_with_each_geometry_matches_oracle(name, geometry):
    c = TRAJ[name]
    print(f"tiles case {name} geometry={geometry}", flush=True)
    lp = V.generate_lp(c["family"], c["m"], c["n"], c["k"], c["seed"])
    opts = dict(c["options"])
    got = V.solve_lp(lp, V.Options(spmv_tile_geometry=geometry, **opts))
    close(got, O.solve_lp(lp.as_oracle(), **opts))
    assert NAMES[geometry] in got.message, got.message


def test_tuned_geometry_is_named_and_reproduces():
    c = TRAJ["zipf_400x1000_it1500"]
    lp = V.generate_lp(c["family"], c["m"], c["n"], c["k"], c["seed"])
    tuned = V.solve_lp(lp, V.Options(spmv_tile_geometry=-1, iteration_limit=600))
    assert "(tuned)" in tuned.message, tuned.message
    chosen = 0
    for g, name in NAMES.items():
        if name in tuned.message:
            chosen = g
    fixed = V.solve_lp(lp, V.Options(spmv_tile_geometry=chosen, iteration_limit=600))
    assert fixed.iterations == tuned.iterations
    assert np.array_equal(fixed.x, tuned.x) and np.array_equal(fixed.y, tuned.y)
    close(tuned, O.solve_lp(lp.as_oracle(), iteration_limit=600))


def test_tuner_times_every_geometry():
    m, n, rp, ci, v = ragged_csr(5)
    dev = V.Device()
    h = dev.csr(m, n, rp, ci, v)
    try:
        for precision in (V.PdhgPrecision.FP64, V.PdhgPrecision.Mixed):
            best, ms = dev.tune_tiles(h, precision, repeats=3)
            assert 0 <= best < 4 and dev.tile_geometry(h) == best
            assert all(t > 0 for t in ms) and ms[best] == min(ms), ms
    finally:
        dev.L.vk_hip_csr_destroy(h)


@pytest.mark.parametrize("geometry", [1, 3])
def test_batch_with_each_geometry_matches_oracle(geometry):
    lps = [V.generate_lp(1, 40, 90, 3, 3), V.generate_lp(4, 0, 0, 4, 9), V.generate_lp(2, 400, 1000, 0, 5),
           V.generate_lp(4, 0, 0, 5, 11)]
    it = 300
    got = V.solve_batch(lps, V.Options(iteration_limit=it, spmv_tile_geometry=geometry))
    for lp, r in zip(lps, got):
        close(r, O.solve_lp(lp.as_oracle(), iteration_limit=it))
        assert NAMES[geometry] in r.message, r.message


def test_invalid_tile_geometry_is_rejected():
    lp = V.generate_lp(1, 40, 90, 3, 3)
    for g in (-2, 4):
        assert V.solve_lp(lp, V.Options(spmv_tile_geometry=g)).status == "InvalidModel", g
    dev = V.Device()
    h = dev.csr(2, 2, np.array([0, 1, 2]), np.array([0, 1]), np.ones(2))
    try:
        with pytest.raises(RuntimeError, match="unknown geometry"):
            dev.set_tile_geometry(h, 4)
    finally:
        dev.L.vk_hip_csr_destroy(h)
