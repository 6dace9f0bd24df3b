"""Seeded convex quadratic objectives for the PDQP tests (not a test module).

make_q(n, seed) builds a symmetric n x n CSR matrix from a splitmix64 stream
(no dependence on numpy's generators): every row gets up to two
off-diagonal partners with values in (-0.25, 0.25), and the diagonal is 0.2
plus the row's absolute off-diagonal sum plus U(0, 1), so Q is strictly
diagonally dominant, hence positive definite.
"""
import numpy as np

MASK = (1 << 64) - 1


def _splitmix(state):
    state = (state + 0x9E3779B97F4A7C15) & MASK
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return state, z ^ (z >> 31)


def make_q(n, seed):
    state = seed & MASK
    entries = {}
    for i in range(n):
        for _ in range(2):
            state, r = _splitmix(state)
            j = r % n
            state, r = _splitmix(state)
            v = ((r >> 11) / float(1 << 53) - 0.5) * 0.5
            if j != i:
                entries[(i, j)] = v
                entries[(j, i)] = v
    rows = [[] for _ in range(n)]
    for (i, j), v in entries.items():
        rows[i].append((j, v))
    for i in range(n):
        state, r = _splitmix(state)
        diag = 0.2 + sum(abs(v) for _, v in rows[i]) + (r >> 11) / float(1 << 53)
        rows[i].append((i, diag))
        rows[i].sort()
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.array([j for r in rows for j, _ in r], np.int64)
    v = np.array([v for r in rows for _, v in r], np.float64)
    return rp, ci, v


# name, LP generator arguments (family, m, n, k, seed), Q seed, options
CASES = [
    ("qp_uniform_60x150_it600", (1, 60, 150, 4, 7), 1, {"iteration_limit": 600}),
    ("qp_zipf_400x1000_it800", (2, 400, 1000, 0, 5), 2, {"iteration_limit": 800}),
    ("qp_uniform_40x90_tol1e-3", (1, 40, 90, 3, 3), 3, {"iteration_limit": 20000, "primal_tolerance": 1e-3,
                                                          "dual_tolerance": 1e-3}),
]
