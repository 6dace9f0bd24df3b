"""Writes tests/golden/reference_qp_trajectories.json from the reference's own
solve_qp (its src/solver/qp/pdqp.cpp compiled with the rest of its CPU path
by oracle/Makefile into oracle/_ref/libvikalp_ref.so, in this container).

Each record holds the inputs (LP generator arguments, the seed of the
quadratic term of tests/qp_data.py, options) and the reference outputs:
status, iterations, objective, residuals and summaries of x and y, so the C
oracle's PDQP is pinned bitwise even where the reference tree is absent.
Run from the repo root after building:
    python tests/golden/make_reference_qp_trajectories.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))
import vikalp_amd as V  # noqa: E402
from oracle import oracle as O  # noqa: E402
from qp_data import CASES, make_q  # noqa: E402


def summary(v):
    return {"sum": float(np.sum(v)), "sumsq": float(np.sum(v * v)),
            "head": [float(t) for t in v[:4]], "tail": [float(t) for t in v[-4:]]}


def main():
    out = {}
    for name, gen, qseed, opts in CASES:
        lp = V.generate_lp(*gen)
        q = make_q(lp.n, qseed)
        r = O.ref_solve_qp(lp.as_oracle(), q, **opts)
        out[name] = {"generator": list(gen), "q_seed": qseed, "q_nnz": int(len(q[2])), "options": opts,
                     "status": r.status, "iterations": r.iterations, "objective": r.objective_value,
                     "primal_residual": r.primal_residual, "stationarity_residual": r.stationarity_residual,
                     "complementarity_residual": r.complementarity_residual,
                     "x": summary(r.x), "y": summary(r.y)}
        print(name, r.status, r.iterations, r.objective_value, flush=True)
    with open(os.path.join(HERE, "reference_qp_trajectories.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
