"""Benchmark: time to 1e-6 relative KKT of the PDHG LP path on one MI355X,
mixed FP32/FP64 vs FP64, and the HBM bandwidth of the SpMV kernels.

Workload (the LP the metric is quoted on): m = 400,000 rows, n = 1,000,000
columns, Zipf(1.5) column lengths on [2, 200] (~17.6M nonzeros), a random 10%
of columns rescaled by 10^U(-2,2), 70% equality rows and 30% <= rows, 0 <= x <= 10,
seeded (generator family 2 of include/vikalp_lpgen.h).

A step is one solve from a cold start until the relative KKT error (original
coordinates) is <= 1e-6, with restarted Halpern PDHG (reflection, fixed-point
residual restarts) on the fused HIP path.  The mixed solve's FP32 phase is
recentred (FP32 corrections to an FP64 base point moved to the restart point
at every restart) and runs to 1e-6 itself; its FP64 phase then certifies the
point.  The same solve with the adaptive policy (restarted averaged PDHG) is
reported beside it.  `value` is the mixed-precision device time of the
iteration phase (the LP is already resident in HBM when it starts; setup,
scaling and postsolve are reported separately), `fp64_seconds` the same for
the FP64 path.

roofline: the dominant kernel of the timed (mixed) solve's iteration, timed
with HIP events on the solver's stream; achieved = its algorithmic HBM bytes
per launch / mean duration.  roofline_fp64: the same for the FP64 drop-in's
iteration.  cpu_baseline: the CPU oracle (the reference's single-threaded
algorithm restated in C) timed on a bounded sample of iterations of the same
LP on one host core, extrapolated to the GPU's iteration count.

Also measured (extra objects of the JSON line): the FP64 solve of the same
LP with the adaptive policy ("adaptive_fp64"), the wall time of each solve
including setup and postsolve ("solve_wall_seconds"), the batched solve of
512 block-angular LPs ("batch"), 256 branch-and-bound node relaxations of
one LP on one device matrix against per-node solve_lp calls ("nodes"), the
unrolled-PDHG adjoint on a 100,000 x 250,000 LP with its central-difference
check ("adjoint"), the benchmark LP plus a sparse convex Q solved by Halpern
PDQP to the same relative-KKT target, beside the reference policy after as
many iterations ("qp"), branch-and-bound on a seeded MILP with node
relaxations solved one at a time against rounds of 16 and 64 on one matrix
("mip"), and MPS files given with --mps or written by --mps-generated
("mps").

Usage: python bench.py [--gpus 1] [--steps 1] [--warmup 0]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WORKLOAD = {"workload": "pdhg_lp_zipf", "m": 400_000, "n": 1_000_000, "family": 2,
            "seed": 20261203, "tolerance": 1e-6, "policy": "halpern"}


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def host_cpu_model():
    """The host CPU as /proc/cpuinfo names it (model name, family / model
    numbers and clock of the first processor), or the machine type."""
    fields = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if not line.strip():
                    break
                key, _, value = line.partition(":")
                fields[key.strip()] = value.strip()
    except OSError:
        pass
    name = fields.get("model name") or platform.processor() or platform.machine()
    extra = [f"{k} {fields[k]}" for k in ("cpu family", "model", "cpu MHz") if k in fields]
    return f"{name} ({', '.join(extra)})" if extra else name


def kernel_table(prof, m, n, nnz, sv, suffix, halpern=True, recentred=False):
    """Algorithmic (compulsory) HBM bytes and achieved GB/s per kernel of one
    iteration, for the layout the solver used.

    SpMV y = A x (SpMV-only kernel of a split layout): values and column
    indices (s_v + 4) nnz, row offsets 4 (rows + 1), x gathered once
    s_v cols, FP64 sums written 8 rows.  Column blocks add offsets and
    partial sums that are not counted (they lower the achieved fraction).
    Row pass (dual), Halpern: FP64 sums 8 m, plus per row read ax, y_h,
    rl, ru, 1/r, y0, ax0 and write y+, y_h, ax (10 s_v m); adaptive: read ax,
    y, rl, ru, 1/r, y_last, write ax, y (8 s_v m) and read + write the FP64
    y_sum (16 m).  Column pass (primal), Halpern: FP64 sums 8 n, per column
    read x+, x_h, aty, c, l, u, s, x0, aty0 and write x_h, aty, x_next
    (12 s_v n); adaptive: read x, c, l, u, s, x_last, write aty, x_next
    (8 s_v n) and the FP64 x_sum (16 n).  (Restart-check iterations, one in
    64, touch a few vectors more; not counted.)
    A fused kernel counts its SpMV without the sum store plus its pass
    without the sum load; k_primal_own (split A' side: SpMV, then the
    column pass over the workgroup's own rows, one launch) stores and
    reloads its sums, so it counts both.
    Panel layout: k_panel_spmv counts the same SpMV bytes for A x and A' y
    (the packed entries are s_v + 4 bytes each; there are no row offsets to
    read, and the partial sums of the groups beyond the first are not
    counted), then k_dual_rows and k_primal_cols the row and column passes.
    Recentred FP32 passes read one vector more each: the FP64 dual base
    (8 m) and the objective c beside the shifted gradient (s_v n).
    """
    spmv_a = (sv + 4) * nnz + 4 * (m + 1) + sv * n + 8 * m
    spmv_at = (sv + 4) * nnz + 4 * (n + 1) + sv * m + 8 * n
    pass_rows = 8 * m + (10 * sv * m if halpern else 8 * sv * m + 16 * m) + (8 * m if recentred else 0)
    pass_cols = 8 * n + (12 * sv * n if halpern else 8 * sv * n + 16 * n) + (sv * n if recentred else 0)
    rows = []
    if prof.get("panels") and prof["panels"][0] > 0:
        rows.append(("k_panel_spmv[A x]", prof["dual_spmv_ms"], spmv_a))
        rows.append(("k_dual_rows", prof["dual_pass_ms"], pass_rows))
        rows.append(("k_panel_spmv[A'y]", prof["primal_spmv_ms"], spmv_at))
        rows.append(("k_primal_cols", prof["primal_pass_ms"], pass_cols))
        return {f"{name}<{suffix}>": {"ms": ms, "bytes": b, "GBs": b / (ms * 1e-3) / 1e9}
                for name, ms, b in rows}
    if prof["col_blocks"] > 1 or prof["split_passes"]:
        rows.append(("k_blocks" if prof["col_blocks"] > 1 else "k_dual_spmv", prof["dual_spmv_ms"], spmv_a))
        rows.append(("k_dual_rows", prof["dual_pass_ms"], pass_rows))
    else:
        rows.append(("k_dual", prof["dual_spmv_ms"], spmv_a - 8 * m + pass_rows - 8 * m))
    if prof["split_passes"]:
        rows.append(("k_primal_own", prof["primal_spmv_ms"], spmv_at + pass_cols))
    else:
        rows.append(("k_primal", prof["primal_spmv_ms"], spmv_at - 8 * n + pass_cols - 8 * n))
    return {f"{name}<{suffix}>": {"ms": ms, "bytes": b, "GBs": b / (ms * 1e-3) / 1e9}
            for name, ms, b in rows}


def fused_iteration_bytes(m, n, nnz, sv, halpern=True):
    """Algorithmic bytes of one whole iteration run as fused kernels (the
    batched kernel's form): both SpMVs without sum stores, with the row and
    column passes' vectors (kernel_table's formulas)."""
    spmv = 2 * (sv + 4) * nnz + 4 * (m + 1) + 4 * (n + 1) + sv * n + sv * m
    rows = 10 * sv * m if halpern else 8 * sv * m + 16 * m
    cols = 12 * sv * n if halpern else 8 * sv * n + 16 * n
    return spmv + rows + cols


# HBM bytes per launch (FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by
# what it reads for a 256 MiB device copy: 2.0 and 1.0) measured with
# `bench.py --counters` under rocprofv3 --pmc on the 400k x 1M LP, keyed by
# kernel and by the panel shape it ran with (groups, KB per panel):
# profiles/pdhg_traffic_zipf_400k_x_1m.md holds the passes and commands.  A
# kernel run with another shape reports no traffic (None).
TRAFFIC_SOURCE = ("rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, per dispatch, scaled by a 256 MiB copy "
                  "(profiles/pdhg_traffic_zipf_400k_x_1m.md); FETCH_SIZE includes MALL hits")
MEASURED_TRAFFIC = {
    ("k_panel_spmv[A x]<double>", 8, 24): 319.6e6, ("k_dual_rows<double>", 8, 24): 57.8e6,
    ("k_panel_spmv[A'y]<double>", 2, 48): 317.0e6, ("k_primal_cols<double>", 2, 48): 112.4e6,
    ("k_panel_spmv[A x]<float>", 8, 24): 172.7e6, ("k_dual_rows<float>", 8, 24): 45.0e6,
    ("k_panel_spmv[A'y]<float>", 4, 16): 179.5e6, ("k_primal_cols<float>", 4, 16): 84.4e6,
}


def measured_traffic(kernel, panels, recentred=False):
    """The measured HBM bytes of `kernel` (a kernel_table label) when the
    profile's panel layout (vk_hip_pdhg_panel_layout's ten ints) has the shape
    it was measured with, else None.  The FP32 passes were measured
    recentred (the mixed solve's default)."""
    if not panels or panels[0] <= 0:
        return None
    if kernel.endswith("<float>") and not kernel.startswith("k_panel_spmv") and not recentred:
        return None
    a_side = "A x" in kernel or kernel.startswith("k_dual_rows")
    g, kb = (panels[0], panels[4]) if a_side else (panels[5], panels[9])
    return MEASURED_TRAFFIC.get((kernel, g, kb))


def counters_driver(V, args):
    dev = V.Device()
    nbytes = 256 << 20
    a, b = dev.alloc(nbytes // 8), dev.alloc(nbytes // 8)
    for _ in range(3):
        V.native.check(dev.L.vk_hip_memcpy_d2d(b, a, nbytes, None), "vk_hip_memcpy_d2d")
    V.native.check(dev.L.vk_hip_sync(None), "vk_hip_sync")
    dev.free_all()
    log("calibration copies done")
    lp = V.generate_lp(WORKLOAD["family"], args.m, args.n, 0, WORKLOAD["seed"])
    for label, precision in (("fp64", V.PdhgPrecision.FP64), ("mixed", V.PdhgPrecision.Mixed)):
        p = V.profile_kernels(lp, V.Options(pdhg_policy=V.PdhgPolicy.Halpern, pdhg_precision=precision,
                                            relative_kkt_tolerance=WORKLOAD["tolerance"],
                                            spmv_panels=args.spmv_panels), 2, 6)
        log(f"profiled {label}: col_blocks={p['col_blocks']} split={p['split_passes']} panels={p['panels']}")


def measure_batch(V, np, args):
    """Batched solve of block-angular instances (T ~ U{4..48} periods of
    150 x 400 with 6 nonzeros per column, 40 coupling rows between periods,
    1-20 branching bounds each) in one persistent launch per phase, to 1e-6
    relative KKT with the Halpern policy, FP64 and mixed; instances the
    solver proves infeasible (Farkas ray) are counted apart."""
    rng = np.random.default_rng(WORKLOAD["seed"])
    periods = rng.integers(4, 49, size=args.batch_instances)
    lps = [V.generate_lp(4, 0, 0, int(t), WORKLOAD["seed"] + i) for i, t in enumerate(periods)]
    rows, cols, nnz = sum(l.m for l in lps), sum(l.n for l in lps), sum(l.nnz for l in lps)
    log(f"batch: {len(lps)} instances, {rows} rows, {cols} columns, {nnz} nonzeros")
    out = {"instances": len(lps), "rows": rows, "cols": cols, "nnz": nnz,
           "tolerance": WORKLOAD["tolerance"], "iteration_limit": args.batch_iterations}
    for label, precision in (("fp64", V.PdhgPrecision.FP64), ("mixed", V.PdhgPrecision.Mixed)):
        o = V.Options(pdhg_policy=V.PdhgPolicy.Halpern, pdhg_precision=precision,
                      relative_kkt_tolerance=WORKLOAD["tolerance"], iteration_limit=args.batch_iterations,
                      mixed_phase_tolerance=args.phase_tolerance, mixed_recentre=not args.no_recentre,
                      time_limit_seconds=args.time_limit,
                      use_graphs=not getattr(args, "no_graphs", False))
        t0 = time.time()
        res = V.solve_batch(lps, o)
        wall = time.time() - t0
        optimal = sum(r.status == "Optimal" for r in res)
        infeasible = sum(r.status == "Infeasible" for r in res)
        its = np.array([r.iterations for r in res])
        dev = res[0].device_seconds
        # Algorithmic bytes of every instance's iterations (FP32 phase at 4-byte
        # values), over the batch's device time: the whole-batch rate, tail
        # hand-off included.
        nbytes = 0.0
        for lp_i, r in zip(lps, res):
            lo = r.low_precision_iterations if precision == V.PdhgPrecision.Mixed else 0
            nbytes += lo * fused_iteration_bytes(lp_i.m, lp_i.n, lp_i.nnz, 4)
            nbytes += (r.iterations - lo) * fused_iteration_bytes(lp_i.m, lp_i.n, lp_i.nnz, 8)
        gbs = nbytes / dev / 1e9 if dev > 0 else None
        out[label] = {"device_seconds": dev, "wall_seconds": wall, "optimal": optimal, "infeasible": infeasible,
                      "decided_per_second": (optimal + infeasible) / dev if dev > 0 else None,
                      "iterations_median": float(np.median(its)), "iterations_max": int(its.max()),
                      "algorithmic_bytes": nbytes,
                      "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                   "frac": gbs / HBM_PEAK_GBS if gbs else None,
                                   "note": "fused-iteration bytes of every instance / batch device time"}}
        log(f"batch {label}: {optimal}/{len(lps)} optimal, {infeasible} proved infeasible, device {dev:.2f}s, "
            f"wall {wall:.2f}s, iterations median {np.median(its):.0f} max {its.max()}, "
            f"{gbs:.0f} GB/s algorithmic")
    return out


def measure_config1(V, np, args):
    """The 1,000 x 3,000 LP (5 nonzeros per column, seed 20261203): the CPU
    oracle's full solve with the reference's defaults (tolerances 1e-7,
    100,000 iterations; it runs to the limit) timed as the median of
    `args.c1_cpu_runs` runs on one host core, beside the same solve on the
    GPU's drop-in path (same status and iteration count) and the Halpern
    solves to 1e-6 relative KKT in FP64 and mixed precision."""
    from oracle import oracle as O   # the checker, timed as the CPU baseline
    lp = V.generate_lp(1, 1000, 3000, 5, WORKLOAD["seed"])
    cpu_times, ref = [], None
    for _ in range(args.c1_cpu_runs):
        t0 = time.perf_counter()
        ref = O.solve_lp(lp.as_oracle())
        cpu_times.append(time.perf_counter() - t0)
    gpu = V.solve_lp(lp, V.Options())
    out = {"m": lp.m, "n": lp.n, "nnz": lp.nnz, "cpu_runs": len(cpu_times),
           "cpu_reference_seconds_median": float(np.median(cpu_times)) if cpu_times else None,
           "cpu_status": ref.status if ref else None, "cpu_iterations": ref.iterations if ref else None,
           "cpu_cores": 1, "host_cpu": host_cpu_model(),
           "gpu_reference_status": gpu.status, "gpu_reference_iterations": gpu.iterations,
           "gpu_reference_device_seconds": gpu.device_seconds, "gpu_reference_wall_seconds": gpu.solve_seconds}
    for label, precision in (("fp64", V.PdhgPrecision.FP64), ("mixed", V.PdhgPrecision.Mixed)):
        r = V.solve_lp(lp, V.Options(pdhg_policy=V.PdhgPolicy.Halpern, pdhg_precision=precision,
                                     relative_kkt_tolerance=WORKLOAD["tolerance"], mixed_phase_tolerance=1e-4,
                                     iteration_limit=args.max_iterations))
        out[f"halpern_{label}"] = {"status": r.status, "iterations": r.iterations,
                                   "device_seconds": r.device_seconds, "wall_seconds": r.solve_seconds,
                                   "rel_kkt": r.kkt_relative}
    log(f"config 1: CPU reference solve median {out['cpu_reference_seconds_median']} s "
        f"({out['cpu_status']}, {out['cpu_iterations']} iterations); GPU drop-in {gpu.status} "
        f"{gpu.iterations} iterations in {gpu.solve_seconds:.3f} s wall; Halpern 1e-6 FP64 "
        f"{out['halpern_fp64']['device_seconds']:.3f} s, mixed {out['halpern_mixed']['device_seconds']:.3f} s")
    return out


def seeded_milp(V, np, m, n, seed, integer_share=0.7, upper=6):
    """A feasible MILP: rows A x <= b with 6 coefficients in 1..8 per row and
    b = A x* + U{0..6} for an integer point x* in [0, upper]^n, 70% of the
    variables integer, objective -U(1, 10) per variable."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    dense = np.zeros((m, n))
    for i in range(m):
        cols = rng.choice(n, size=min(n, 6), replace=False)
        dense[i, cols] = rng.integers(1, 9, size=cols.size)
    xstar = rng.integers(0, upper + 1, size=n).astype(float)
    b = dense @ xstar + rng.integers(0, 7, size=m)
    A = sp.csr_matrix(dense)
    types = (rng.random(n) < integer_share).astype(np.int32)
    c = -rng.uniform(1.0, 10.0, size=n)
    return V.LP(m, n, A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data.astype(np.float64), c,
                np.full(m, -np.inf), b.astype(np.float64), np.zeros(n), np.full(n, float(upper)),
                variable_types=types)


def measure_mip(V, np, args):
    """Branch-and-bound with batched node relaxations (solve_mip_batched) on a
    seeded 30 x 40 MILP (27 integer variables, about a thousand nodes): the
    serial search (batch 1, the reference's node order, one relaxation per
    round) against rounds of 16 and 64 nodes in the batched kernel, all on one
    cached device matrix.  Halpern relaxations to 1e-9 relative KKT."""
    lp = seeded_milp(V, np, 30, 40, 3)   # the 30 x 40 instance of tests/test_gpu_mip.py
    o = V.Options(pdhg_policy=V.PdhgPolicy.Halpern, relative_kkt_tolerance=1e-9, iteration_limit=200_000,
                  absolute_gap_tolerance=1e-6, relative_gap_tolerance=1e-9, use_graphs=not args.no_graphs)
    out = {"m": lp.m, "n": lp.n, "integer_variables": int(np.sum(lp.variable_types)), "runs": {}}
    for batch in args.mip_batches:
        t = time.time()
        r, st = V.solve_mip(lp, o, batch=batch)
        wall = time.time() - t
        out["runs"][str(batch)] = {"status": r.status, "objective": r.objective_value, "wall_seconds": wall,
                                   "device_seconds": r.device_seconds, **st}
        log(f"mip: batch {batch}: {r.status} objective {r.objective_value:.9g}, {st['nodes_processed']} nodes in "
            f"{st['rounds']} rounds, wall {wall:.2f}s")
    runs = out["runs"]
    if "1" in runs:
        for b, v in runs.items():
            v["speedup_vs_serial"] = runs["1"]["wall_seconds"] / v["wall_seconds"]
    return out


def measure_nodes(V, np, args):
    """Branch-and-bound node relaxations: `args.nodes` nodes of one
    block-angular LP (T = 24 periods), each with 1-20 variables branched to
    floor / ceil of x0 (the node loop of BranchAndBound.cpp:296-297), solved
    as one batch on one device-resident matrix (solve_lp_nodes) against one
    solve_lp call per node on a sample of the nodes, Halpern policy, FP64,
    to 1e-6 relative KKT."""
    rng = np.random.default_rng(WORKLOAD["seed"] + 1)
    lp = V.generate_lp(4, 0, 0, 24, WORKLOAD["seed"] + 24)
    lo = np.tile(lp.var_lower, (args.nodes, 1))
    hi = np.tile(lp.var_upper, (args.nodes, 1))
    for k in range(args.nodes):
        for j in rng.choice(lp.n, size=rng.integers(1, 21), replace=False):
            if rng.random() < 0.5:
                hi[k, j] = np.floor(lp.x0[j])
            else:
                lo[k, j] = np.ceil(lp.x0[j])
    o = V.Options(pdhg_policy=V.PdhgPolicy.Halpern, relative_kkt_tolerance=WORKLOAD["tolerance"],
                  iteration_limit=args.batch_iterations, time_limit_seconds=args.time_limit,
                  use_graphs=not args.no_graphs)
    t0 = time.time()
    res = V.solve_nodes(lp, lo, hi, options=o)
    wall = time.time() - t0
    decided = sum(r.status in ("Optimal", "Infeasible") for r in res)
    sample = min(args.nodes, 16)
    t1 = time.time()
    single = [V.solve_lp(V.LP(lp.m, lp.n, lp.row_offsets, lp.column_indices, lp.values, lp.objective,
                              lp.row_lower, lp.row_upper, lo[k], hi[k]), o) for k in range(sample)]
    per_node = (time.time() - t1) / sample
    agree = sum(a.status == b.status for a, b in zip(res[:sample], single))
    out = {"nodes": args.nodes, "m": lp.m, "n": lp.n, "nnz": lp.nnz, "decided": decided,
           "batch_device_seconds": res[0].device_seconds, "batch_wall_seconds": wall,
           "single_wall_seconds_per_node": per_node, "single_sample": sample,
           "speedup_vs_single_calls": per_node * args.nodes / wall, "status_agreement_on_sample": agree}
    log(f"nodes: {args.nodes} nodes of a {lp.m} x {lp.n} LP: batch wall {wall:.2f}s (device "
        f"{res[0].device_seconds:.2f}s), {decided} decided; single solve_lp {per_node * 1e3:.0f} ms/node "
        f"-> {out['speedup_vs_single_calls']:.1f}x")
    return out


def measure_qp(V, np, lp, args, lp_ms_per_iteration):
    """PDQP (pdqp.cpp:20-331) on the benchmark LP plus a sparse convex Q:
    every column gets one off-diagonal partner with a value in
    (-0.25, 0.25), mirrored, and a diagonal of 0.2 + its row's absolute
    off-diagonal sum + U(0, 1) (strictly diagonally dominant, so positive
    definite).  Restarted Halpern PDHG with per-column primal steps solves it
    to the benchmark's relative-KKT target in FP64 and in mixed precision
    (the recentred FP32 phase, then the certifying FP64 phase); the reference
    policy then runs as many iterations as the FP64 solve, for its relative
    KKT at that count.  The cost of Q x per iteration is the FP64 solve's
    ms/iteration beside the FP64 LP solve's."""
    import scipy.sparse as sp
    rng = np.random.default_rng(WORKLOAD["seed"] + 7)
    n = lp.n
    i = np.arange(n)
    j = rng.integers(0, n, n)
    keep = i != j
    i, j = i[keep], j[keep]
    v = rng.uniform(-0.25, 0.25, i.size)
    off = sp.coo_matrix((np.concatenate([v, v]), (np.concatenate([i, j]), np.concatenate([j, i]))),
                        shape=(n, n)).tocsr()
    off.sum_duplicates()
    diag = 0.2 + np.asarray(abs(off).sum(axis=1)).ravel() + rng.uniform(0, 1, n)
    q = (off + sp.diags(diag)).tocsr()
    q.sort_indices()
    quadratic = (q.indptr.astype(np.int64), q.indices.astype(np.int64), q.data.astype(np.float64))
    o = V.Options(pdhg_policy=V.PdhgPolicy.Halpern, relative_kkt_tolerance=WORKLOAD["tolerance"],
                  iteration_limit=args.qp_iterations, time_limit_seconds=args.time_limit,
                  use_graphs=not args.no_graphs)
    rq = V.solve_qp(lp, quadratic, o)
    log(f"qp: Halpern {rq.status} in {rq.iterations} iterations, device {rq.device_seconds:.3f}s, "
        f"rel KKT {rq.kkt_relative:.3e}, objective {rq.objective_value:.10g}, wall {rq.solve_seconds:.2f}s")
    o.pdhg_precision = V.PdhgPrecision.Mixed
    rm = V.solve_qp(lp, quadratic, o)
    log(f"qp: Halpern mixed {rm.status} in {rm.low_precision_iterations} FP32 + "
        f"{rm.iterations - rm.low_precision_iterations} FP64 iterations, device {rm.device_seconds:.3f}s, "
        f"rel KKT {rm.kkt_relative:.3e}, objective {rm.objective_value:.10g}")
    rr = V.solve_qp(lp, quadratic, V.Options(iteration_limit=max(1, rq.iterations),
                                             time_limit_seconds=args.time_limit, use_graphs=not args.no_graphs))
    log(f"qp: reference policy after {rr.iterations} iterations: {rr.status}, rel KKT {rr.kkt_relative:.3e}")
    return {"n": n, "q_nnz": int(q.nnz), "policy": "halpern", "precision": "fp64",
            "tolerance": WORKLOAD["tolerance"], "status": rq.status, "iterations": rq.iterations,
            "device_seconds": rq.device_seconds, "wall_seconds": rq.solve_seconds,
            "rel_kkt": rq.kkt_relative, "objective": rq.objective_value,
            "qp_ms_per_iteration": rq.device_seconds / max(1, rq.iterations) * 1e3,
            "lp_ms_per_iteration": lp_ms_per_iteration,
            "mixed": {"status": rm.status, "iterations": rm.iterations,
                      "iterations_fp32_phase": rm.low_precision_iterations,
                      "device_seconds": rm.device_seconds, "wall_seconds": rm.solve_seconds,
                      "rel_kkt": rm.kkt_relative,
                      "objective_rel_diff_vs_fp64": abs(rm.objective_value - rq.objective_value) /
                      max(1.0, abs(rq.objective_value)),
                      "speedup_vs_fp64": rq.device_seconds / rm.device_seconds if rm.device_seconds > 0 else None},
            "reference_policy_same_iterations": {"status": rr.status, "iterations": rr.iterations,
                                                 "rel_kkt": rr.kkt_relative,
                                                 "device_seconds": rr.device_seconds}}


def write_generated_mps(V, np, path):
    """Writes the seeded 100,000 x 250,000 uniform LP (6 nonzeros per column,
    equality rows, 0 <= x <= 10) as free MPS, every double by its shortest
    repr so that reading it back gives the same bits: an input of realistic
    size for the MPS path, since no netlib or MIPLIB file is available."""
    lp = V.generate_lp(1, 100_000, 250_000, 6, WORKLOAD["seed"])
    assert (lp.var_lower == 0).all() and (lp.row_lower == lp.row_upper).all()
    rows = np.repeat(np.arange(lp.m), np.diff(lp.row_offsets))
    order = np.lexsort((rows, lp.column_indices))
    rows, cols, vals = rows[order], lp.column_indices[order], lp.values[order]
    starts = np.searchsorted(cols, np.arange(lp.n + 1))
    r = lambda v: repr(float(v))
    with open(path, "w") as f:
        f.write("NAME UNIF100K\nROWS\n N OBJ\n")
        f.writelines(f" E R{i}\n" for i in range(lp.m))
        f.write("COLUMNS\n")
        for j in range(lp.n):
            f.write(f" X{j} OBJ {r(lp.objective[j])}\n")
            f.writelines(f" X{j} R{rows[k]} {r(vals[k])}\n" for k in range(starts[j], starts[j + 1]))
        f.write("RHS\n")
        f.writelines(f" B R{i} {r(lp.row_lower[i])}\n" for i in range(lp.m))
        f.write("BOUNDS\n")
        f.writelines(f" UP B X{j} {r(lp.var_upper[j])}\n" for j in range(lp.n))
        f.write("ENDATA\n")
    log(f"wrote {path}: {lp.m} rows, {lp.n} columns, {lp.nnz} nonzeros, {os.path.getsize(path) / 1e6:.0f} MB")


def measure_mps(V, path, options):
    """One real instance read from an MPS file (MpsReader::read), solved to
    the benchmark tolerance in FP64 and mixed precision (Halpern policy)."""
    t0 = time.time()
    lp = V.read_mps(path)
    out = {"file": os.path.basename(path), "name": lp.name, "m": lp.m, "n": lp.n, "nnz": lp.nnz,
           "read_seconds": time.time() - t0}
    for label, precision in (("fp64", V.PdhgPrecision.FP64), ("mixed", V.PdhgPrecision.Mixed)):
        r = V.solve_lp(lp, options(precision))
        out[label] = {"status": r.status, "iterations": r.iterations, "device_seconds": r.device_seconds,
                      "solve_seconds": r.solve_seconds, "rel_kkt": r.kkt_relative,
                      "objective": r.objective_value}
        log(f"mps {lp.name} {label}: {r.status} iterations={r.iterations} device={r.device_seconds:.3f}s "
            f"wall={r.solve_seconds:.3f}s relkkt={r.kkt_relative:.3e} obj={r.objective_value:.10g}")
    return out


def measure_adjoint(V, np, args):
    """Unrolled PDHG forward / backward on m = 100,000, n = 250,000, 6
    nonzeros per column (generator family 1), K iterations with checkpoints
    every 50; loss c'x_K + |x_K - x_t|^2 / 2; the gradient with respect to
    (c, b) checked by central differences (h = 1e-6) along random unit
    directions."""
    lp = V.generate_lp(1, 100_000, 250_000, 6, WORKLOAD["seed"])
    K = args.adjoint_iterations
    rng = np.random.default_rng(WORKLOAD["seed"] + 5)
    xt = rng.uniform(0, 10, lp.n)
    u = V.UnrolledPdhg(lp, K, checkpoint_every=50)
    x, _ = u.forward()
    fwd_ms = u.info()["last_device_ms"]
    gc, grl, gru = u.backward(lp.objective + x - xt)
    bwd_ms = u.info()["last_device_ms"]
    gc = gc + x
    gb = grl + gru
    v = V.UnrolledPdhg(lp, K, checkpoint_every=K)

    def loss(dc, db, h):
        c = lp.objective + h * dc
        v.set_data(c, lp.row_lower + h * db, lp.row_upper + h * db)
        xk, _ = v.forward()
        return float(c @ xk + 0.5 * np.sum((xk - xt) ** 2))

    # Two error measures per direction d: |fd - g'd| / |g'd| (ill-conditioned
    # when d is nearly orthogonal to g, as random directions in 350,000
    # dimensions are: |g'd| ~ |g| / 600) and |fd - g'd| / |g|.  Acceptance:
    # |fd - g'd| <= 1e-4 |g'd| + 1e-6 |g|.  The floor is the central
    # difference's own rounding noise: each loss carries a relative error of
    # a few ulps of |L| through K iterations, divided by h = 1e-6 that is
    # ~1e-9 |L| / h ~ 1e-6 |g| here; a direction with |g'd| below 1e-2 |g|
    # cannot be resolved to 1e-4 by this difference in FP64.
    gnorm = math.sqrt(float(gc @ gc + gb @ gb))
    errs, errs_norm, fails = [], [], 0
    loss0 = loss(np.zeros(lp.n), np.zeros(lp.m), 0.0)
    for _ in range(args.fd_directions):
        dc, db = rng.normal(size=lp.n), rng.normal(size=lp.m)
        norm = math.sqrt(dc @ dc + db @ db)
        dc, db = dc / norm, db / norm
        fd = (loss(dc, db, 1e-6) - loss(dc, db, -1e-6)) / 2e-6
        an = float(gc @ dc + gb @ db)
        errs.append(abs(fd - an) / max(abs(an), 1e-300))
        errs_norm.append(abs(fd - an) / gnorm)
        fails += abs(fd - an) > 1e-4 * abs(an) + 1e-6 * gnorm
    log(f"adjoint: m={lp.m} n={lp.n} nnz={lp.nnz} K={K}: forward {fwd_ms:.1f} ms, backward {bwd_ms:.1f} ms, "
        f"central differences over {len(errs)} directions: |fd-g'd|/|g'd| median {np.median(errs):.2e} "
        f"max {max(errs):.2e}; |fd-g'd|/|g| max {max(errs_norm):.2e} (loss {loss0:.6g}, |g| {gnorm:.4g})")
    # Algorithmic bytes per unrolled step: forward = A'y with the primal
    # update (x, c, l, u read, x written) and A xbar with the dual update (y,
    # rl, ru read, y written), each SpMV (8 + 4) nnz + offsets + its gathered
    # vector once; backward = the recomputed forward step plus its active-set
    # bits, then the adjoint step's two SpMVs and their vector traffic
    # (g_x, g_y, g_xbar, g_c in and out, the bits read).
    m, n, nnz = lp.m, lp.n, lp.nnz
    spmvs = 2 * 12 * nnz + 4 * (m + 1) + 4 * (n + 1) + 8 * m + 8 * n
    fwd_step = spmvs + 5 * 8 * n + 4 * 8 * m
    bits = (m + n) / 8
    bwd_step = fwd_step + bits + spmvs + 6 * 8 * n + 4 * 8 * m + bits
    fwd_gbs = K * fwd_step / (fwd_ms * 1e-3) / 1e9
    bwd_gbs = K * bwd_step / (bwd_ms * 1e-3) / 1e9
    log(f"adjoint: forward {fwd_gbs:.0f} GB/s, backward {bwd_gbs:.0f} GB/s algorithmic; "
        f"{sum(e > 1e-4 for e in errs)} of {len(errs)} directions above 1e-4 relative, {fails} outside "
        f"1e-4 |g'd| + 1e-6 |g|")
    return {"m": lp.m, "n": lp.n, "nnz": lp.nnz, "iterations": K, "checkpoint_every": 50,
            "forward_ms": fwd_ms, "backward_ms": bwd_ms, "fd_directions": len(errs), "fd_step": 1e-6,
            "forward_GBs": fwd_gbs, "backward_GBs": bwd_gbs,
            "fd_rel_err_directional_median": float(np.median(errs)),
            "fd_rel_err_directional_max": max(errs),
            "fd_directions_above_1e-4": int(sum(e > 1e-4 for e in errs)),
            "fd_err_rel_gradient_norm_max": max(errs_norm),
            "fd_acceptance": "|fd - g'd| <= 1e-4 |g'd| + 1e-6 |g|",
            "fd_directions_failing_acceptance": int(fails),
            "loss": loss0, "gradient_norm": gnorm}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--m", type=int, default=WORKLOAD["m"])
    ap.add_argument("--n", type=int, default=WORKLOAD["n"])
    ap.add_argument("--max-iterations", type=int, default=200_000)
    ap.add_argument("--time-limit", type=float, default=100.0, help="seconds per solve")
    ap.add_argument("--cpu-iterations", type=int, default=200)
    ap.add_argument("--c1-cpu-runs", type=int, default=5,
                    help="runs of the CPU oracle's full solve of the 1,000 x 3,000 LP (median; 0 skips "
                         "the config-1 measurement)")
    ap.add_argument("--phase-tolerance", type=float, default=0.0,
                    help="relative KKT target of the FP32 phase of the mixed solve (0: automatic, the "
                         "final target for the recentred FP32 iteration)")
    ap.add_argument("--no-recentre", action="store_true",
                    help="mixed solve with plain FP32 iterates (FP32 phase to 1e-5 at most) instead of the "
                         "recentred FP32 iteration")
    ap.add_argument("--batch-instances", type=int, default=512,
                    help="block-angular instances of the batched solve (0 = skip)")
    ap.add_argument("--batch-iterations", type=int, default=40_000)
    ap.add_argument("--mip-batches", type=lambda s: [int(v) for v in s.split(",") if v], default=[1, 16, 64],
                    help="batch sizes of the branch-and-bound measurement (comma-separated; empty skips)")
    ap.add_argument("--nodes", type=int, default=256,
                    help="branch-and-bound node relaxations of one LP on one matrix (0 = skip)")
    ap.add_argument("--adjoint-iterations", type=int, default=1000,
                    help="unrolled iterations of the adjoint measurement (0 = skip)")
    ap.add_argument("--fd-directions", type=int, default=64)
    ap.add_argument("--qp-iterations", type=int, default=100_000,
                    help="iteration limit of the Halpern PDQP solve of the benchmark LP plus Q (0 skips)")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-adaptive", action="store_true",
                    help="skip the FP64 solve with the adaptive policy (comparison)")
    ap.add_argument("--skip-solve", action="store_true",
                    help="kernel timing only (used for the rocprofv3 pass)")
    ap.add_argument("--mps", nargs="*", default=[],
                    help="also solve these MPS files (FP64 and mixed, Halpern policy) and report them "
                         "under \"mps\"")
    ap.add_argument("--mps-generated", action="store_true",
                    help="also write the seeded 100,000 x 250,000 uniform LP as MPS and measure it like --mps")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every kernel directly instead of replaying HIP graphs (for kernel tracing: "
                         "rocprofv3 --kernel-trace does not follow graph replays with a forked stream)")
    ap.add_argument("--spmv-panels", type=int, default=0, choices=(0, 1, 2),
                    help="SpMV layout of the single-LP solves: 0 automatic (the panel layout on the "
                         "benchmark LP), 1 the tile layouts, 2 panels always")
    ap.add_argument("--counters", action="store_true",
                    help="driver for the rocprofv3 --pmc passes: three 256 MiB device copies "
                         "(counter calibration), then 2 + 6 profiled iterations in FP64, then in "
                         "FP32 (the mixed solve's first phase); no JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU (--gpus {args.gpus} asked): the PDHG path runs on one "
              "MI355X and has no multi-GPU sharding.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import vikalp_amd as V

    V.load_libraries()
    if args.counters:
        counters_driver(V, args)
        return
    t0 = time.time()
    lp = V.generate_lp(WORKLOAD["family"], args.m, args.n, 0, WORKLOAD["seed"])
    log(f"generated LP m={lp.m} n={lp.n} nnz={lp.nnz} in {time.time() - t0:.1f}s")

    def options(precision, policy=V.PdhgPolicy.Halpern):
        return V.Options(pdhg_policy=policy, pdhg_precision=precision,
                         relative_kkt_tolerance=WORKLOAD["tolerance"],
                         mixed_phase_tolerance=args.phase_tolerance, mixed_recentre=not args.no_recentre,
                         iteration_limit=args.max_iterations, time_limit_seconds=args.time_limit,
                         use_graphs=not args.no_graphs, spmv_panels=args.spmv_panels)

    # Kernel timing for the roofline (HIP events on the solver's stream).
    prof = V.profile_kernels(lp, options(V.PdhgPrecision.FP64), warm_iterations=30, iterations=60)
    prof32 = V.profile_kernels(lp, options(V.PdhgPrecision.Mixed), warm_iterations=30, iterations=60)
    k64 = kernel_table(prof, lp.m, lp.n, lp.nnz, 8, "double")
    k32 = kernel_table(prof32, lp.m, lp.n, lp.nnz, 4, "float", recentred=not args.no_recentre)
    for table, p, rec in ((k64, prof, False), (k32, prof32, not args.no_recentre)):
        for name, row in table.items():
            row["traffic"] = measured_traffic(name, p["panels"], rec)
    for label, p, table in (("fp64", prof, k64), ("mixed", prof32, k32)):
        log(f"{label}: iteration {p['iteration_ms']:.4f} ms, col_blocks={p['col_blocks']} "
            f"split={p['split_passes']} panels={p['panels']}; " +
            ", ".join(f"{k} {v['ms'] * 1e3:.1f} us {v['GBs']:.0f} GB/s" for k, v in table.items()))

    def run(precision, label, policy=V.PdhgPolicy.Halpern):
        log(f"solve {label} start")
        r = V.solve_lp(lp, options(precision, policy))
        log(f"solve {label}: {r.status} iterations={r.iterations} (fp32 {r.low_precision_iterations}) "
            f"device={r.device_seconds:.3f}s wall={r.solve_seconds:.3f}s relkkt={r.kkt_relative:.3e} "
            f"obj={r.objective_value:.10g}")
        return r

    for _ in range(args.warmup if not args.skip_solve else 0):
        run(V.PdhgPrecision.Mixed, "warmup")
    fp64_runs, mixed_runs = [], []
    if args.skip_solve:
        args.skip_cpu = True
        empty = V.Result("NotSolved", 0, math.nan, math.inf, math.inf, math.inf, math.inf, math.inf,
                         0.0, 0.0, math.inf, math.inf, math.inf, math.inf, math.nan, math.nan, 0,
                         "hip", "", "solve skipped")
        fp64_runs, mixed_runs = [empty], [empty]
    for _ in range(args.steps if not args.skip_solve else 0):
        fp64_runs.append(run(V.PdhgPrecision.FP64, "fp64"))
        mixed_runs.append(run(V.PdhgPrecision.Mixed, "mixed"))

    def summary(runs):
        ok = [r for r in runs if r.status == "Optimal"]
        dev = float(np.median([r.device_seconds for r in runs]))
        return dev, bool(ok) and len(ok) == len(runs)

    t64, ok64 = summary(fp64_runs)
    t32, ok32 = summary(mixed_runs)
    r32, r64 = mixed_runs[-1], fp64_runs[-1]
    obj_err = abs(r32.objective_value - r64.objective_value) / max(1.0, abs(r64.objective_value))

    def roofline(table, panels, rec):
        dominant = max(table, key=lambda k: table[k]["ms"])
        row = table[dominant]
        return {"bound": "hbm", "achieved": row["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": row["GBs"] / HBM_PEAK_GBS, "traffic": measured_traffic(dominant, panels, rec),
                "traffic_source": TRAFFIC_SOURCE, "kernel": dominant, "bytes_per_launch": row["bytes"],
                "mean_ms": row["ms"]}

    adaptive = None
    if not args.skip_solve and not args.skip_adaptive:
        ra = run(V.PdhgPrecision.FP64, "fp64 adaptive", V.PdhgPolicy.Adaptive)
        adaptive = {"policy": "adaptive", "precision": "fp64", "status": ra.status, "iterations": ra.iterations,
                    "device_seconds": ra.device_seconds, "rel_kkt": ra.kkt_relative,
                    "objective_rel_diff_vs_halpern": abs(ra.objective_value - r64.objective_value) /
                    max(1.0, abs(r64.objective_value))}

    cpu = None
    if not args.skip_cpu:
        from oracle import oracle as O   # the checker, timed as the CPU baseline
        olp = lp.as_oracle()
        per_it = O.time_iterations(olp, args.cpu_iterations) / args.cpu_iterations
        its = r64.iterations
        cpu = {"value": per_it * its, "unit": "s (extrapolated)", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_iterations} iterations of the reference loop (4 SpMVs/iteration) "
                         f"on the same LP: {per_it:.3f} s/iteration x {its} iterations of the "
                         f"FP64 GPU solve (the GPU runs the Halpern policy, 2 SpMVs/iteration: the "
                         f"extrapolation prices the reference algorithm at the GPU's iteration count)",
               "host_cpu": host_cpu_model(), "host_nproc": os.cpu_count()}
        log(f"cpu baseline: {per_it:.3f} s/iteration")

    config1 = None
    if args.c1_cpu_runs > 0 and not args.skip_solve and not args.skip_cpu:
        config1 = measure_config1(V, np, args)
    batch = None
    if args.batch_instances > 0 and not args.skip_solve:
        batch = measure_batch(V, np, args)
    adjoint = None
    if args.adjoint_iterations > 0 and not args.skip_solve:
        adjoint = measure_adjoint(V, np, args)
    generated_dir = None
    if args.mps_generated:
        import tempfile
        generated_dir = tempfile.mkdtemp()
        path = os.path.join(generated_dir, "unif100k.mps")
        write_generated_mps(V, np, path)
        args.mps = list(args.mps) + [path]
    mps = [measure_mps(V, path, options) for path in args.mps] if args.mps else None
    if generated_dir:
        import shutil
        shutil.rmtree(generated_dir, ignore_errors=True)
    nodes = measure_nodes(V, np, args) if args.nodes > 0 and not args.skip_solve else None
    mip = measure_mip(V, np, args) if args.mip_batches and not args.skip_solve else None
    qp = (measure_qp(V, np, lp, args, t64 * 1e3 / max(1, r64.iterations))
          if args.qp_iterations > 0 and not args.skip_solve else None)

    line = {
        "metric": "time to 1e-6 rel. KKT, mixed-FP32 vs FP64 PDHG, 1 MI355X; SpMV HBM GB/s",
        "value": t32 if ok32 else None,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t32 * 1e3,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded generator, LP built on the host)",
        "config": {**WORKLOAD, "m": lp.m, "n": lp.n, "nnz": lp.nnz},
        "fp64_seconds": t64 if ok64 else None,
        "solve_wall_seconds": {"fp64": float(np.median([r.solve_seconds for r in fp64_runs])),
                               "mixed": float(np.median([r.solve_seconds for r in mixed_runs]))},
        "speedup_mixed_vs_fp64": (t64 / t32) if (ok32 and ok64) else None,
        "iterations_fp64": r64.iterations,
        "iterations_mixed": r32.iterations,
        "iterations_mixed_fp32_phase": r32.low_precision_iterations,
        "status_fp64": r64.status,
        "status_mixed": r32.status,
        "rel_kkt_fp64": r64.kkt_relative,
        "rel_kkt_mixed": r32.kkt_relative,
        "objective_rel_diff_mixed_vs_fp64": obj_err,
        # Per iteration in the solves (graph replays) and, kernel by kernel
        # with events between them, in the profile.
        "ms_per_iteration_fp64": t64 * 1e3 / max(1, r64.iterations),
        "ms_per_iteration_mixed": t32 * 1e3 / max(1, r32.iterations),
        "ms_per_iteration_profiled": {"fp64": prof["iteration_ms"], "mixed": prof32["iteration_ms"]},
        "roofline": roofline(k32, prof32["panels"], not args.no_recentre),
        "roofline_fp64": roofline(k64, prof["panels"], False),
        "layout": {"fp64": {"col_blocks": prof["col_blocks"], "split_passes": prof["split_passes"],
                            "panels": prof["panels"]},
                   "mixed": {"col_blocks": prof32["col_blocks"], "split_passes": prof32["split_passes"],
                             "panels": prof32["panels"]}},
        "kernels": {"fp64": k64, "mixed": k32},
        "cpu_baseline": cpu,
        "config1": config1,
        "adaptive_fp64": adaptive,
        "batch": batch,
        "adjoint": adjoint,
        "mps": mps,
        "nodes": nodes,
        "mip": mip,
        "qp": qp,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
