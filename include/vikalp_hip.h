/*
 * vikalp_hip.h — C ABI of the MI355X (gfx950) kernel library libvikalp_hip.so.
 *
 * This is the drop-in boundary under the reference's ExecutionBackend
 * contract (include/vikalp/contracts/ExecutionBackend.hpp:36-64).  The CUDA
 * backend it replaces (cuda/cuda_ops.cu) reaches its device through CUDA
 * runtime calls, five kernels and cuSPARSE; here a C++ HipBackend (and any
 * FFI caller) reaches the device only through these entry points.
 *
 * Rules for every entry point:
 *   - returns 0 on success, a negative VK_E* code on failure; the message is
 *     kept per thread and read with vk_hip_last_error();
 *   - device buffers are plain pointers from vk_hip_alloc; sizes are element
 *     counts (int64_t); `stream` is a hipStream_t passed as void* (NULL =
 *     the library's own stream);
 *   - FP64 unless the name says f32; device indices are int32 (host CSR
 *     input is int64, like vikalp::Index, and is range-checked on upload);
 *   - no entry point uses floating-point atomics: every reduction has a
 *     fixed order and repeats bitwise.
 */
#ifndef VIKALP_HIP_H
#define VIKALP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    VK_OK = 0,
    VK_EINVAL = -1,     /* bad argument (dimension, null, range)         */
    VK_EDEVICE = -2,    /* HIP runtime error                             */
    VK_ENOMEM = -3,     /* device or host allocation failed              */
    VK_ESTATE = -4      /* call out of order / object in the wrong state */
};

/* ---- runtime --------------------------------------------------------- */

/* cuda_ops.cu:314-317 (cudaSetDevice(0) + cusparseCreate in the backend
 * constructor).  Selects `device` and creates the library stream. */
int vk_hip_init(int device);
/* Releases the library stream; device buffers stay valid until freed. */
int vk_hip_shutdown(void);
const char *vk_hip_last_error(void);
/* The library's stream (hipStream_t as void*). */
void *vk_hip_stream(void);

/* cuda_ops.cu:89-124 (DeviceBuffer RAII) */
int vk_hip_alloc(void **ptr, size_t bytes);
int vk_hip_free(void *ptr);
/* cuda_ops.cu:339-360 upload/download and :384-393 copy.  The h2d/d2h
 * forms are synchronous on `stream` (the host buffer is free on return). */
int vk_hip_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream);
int vk_hip_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream);
int vk_hip_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream);
/* cuda_ops.cu:659-661 synchronize() */
int vk_hip_sync(void *stream);

/* ---- vector ops (ExecutionBackend.hpp:46-55) --------------------------- */

/* fill: cuda_ops.cu:139-142 / cpu_ops.cpp:316-319 */
int vk_hip_fill(double *y, int64_t n, double value, void *stream);
/* y = alpha*x + beta*y, literally (x may alias y): cpu_ops.cpp:331-343 */
int vk_hip_axpby(int64_t n, double alpha, const double *x, double beta, double *y,
                 void *stream);
/* x = min(max(x, lo), hi): cpu_ops.cpp:367-383 */
int vk_hip_project_box(int64_t n, double *x, const double *lo, const double *hi,
                       void *stream);
/* max |x_i| (0 for n = 0), result to host; synchronous: cpu_ops.cpp:358-364,
 * cuda_ops.cu:432-451 (finished on the device here, not on the host). */
int vk_hip_reduce_absmax(int64_t n, const double *x, double *result, void *stream);
/* sum x_i*y_i, result to host; synchronous: cpu_ops.cpp:346-355 */
int vk_hip_reduce_dot(int64_t n, const double *x, const double *y, double *result,
                      void *stream);
/* vk_hip_reduce_dot for callers that cannot allocate device memory (code
 * that may not call hipMalloc: serving loops, harnesses that forbid it).
 * No counterpart in the reference.  It allocates nothing, not even on its
 * first call or before vk_hip_init: its partial sums live in a fixed 8.3 KB
 * device array per GPU (1024 partials and the result slot), found by
 * hipGetSymbolAddress on the first call on each device, and it never calls
 * vk_hip_init.  Same kernels and grid as vk_hip_reduce_dot (results are
 * bitwise equal); no limit on n, whose partial count is capped at the
 * array's 1024.  Runs on `stream` (NULL: the current device's null stream),
 * then synchronises it.  Every call on a GPU shares that GPU's array, which
 * keeps the last call's contents (nothing in it needs clearing: each call
 * writes every partial it reads); calls that overlapped on one GPU would
 * race on it, so a lock makes the calls of this process take turns. */
int vk_hip_reduce_dot_noalloc(int64_t n, const double *x, const double *y, double *result,
                              void *stream);

/* ---- sparse matrices --------------------------------------------------- */

/* A device-resident CSR matrix together with its explicitly stored
 * transpose and the row tiling both SpMV directions use. */
typedef struct vk_hip_csr vk_hip_csr;

/* cuda_ops.cu:198-281 create_matrix (+ validate_pattern, cpu_ops.cpp:38-63).
 * Takes the host pattern (int64 offsets/indices, columns strictly increasing
 * per row), validates it, uploads it as int32 and builds the transpose
 * pattern.  Values are zero until vk_hip_csr_set_values. */
int vk_hip_csr_create(int64_t rows, int64_t cols, const int64_t *row_offsets,
                      const int64_t *column_indices, vk_hip_csr **out);
/* cuda_ops.cu:362-372 set_values: nnz host values in pattern order; the
 * transpose values are gathered on the device. */
int vk_hip_csr_set_values(vk_hip_csr *A, const double *values);
int vk_hip_csr_destroy(vk_hip_csr *A);
int vk_hip_csr_dims(const vk_hip_csr *A, int64_t *rows, int64_t *cols, int64_t *nnz);
/* Copies the device transpose (int32 offsets n+1, int32 row indices nnz,
 * values nnz) to host buffers; any pointer may be NULL. */
int vk_hip_csr_transpose_export(const vk_hip_csr *A, int32_t *t_offsets,
                                int32_t *t_rows, double *t_values);

/* cuda_ops.cu:473-528 multiply: y = alpha*op(A)*x + beta*y, op = A when
 * transpose == 0, A' otherwise.  A' runs on the stored transpose (row-
 * parallel, no atomics).  y must not alias x. */
int vk_hip_spmv_csr(const vk_hip_csr *A, int transpose, double alpha,
                    const double *x, double beta, double *y, void *stream);

/* Estimate ||A||_2 by `iterations` power iterations on A'A from a fixed
 * seeded start vector (deterministic).  Result to host. */
int vk_hip_csr_norm2_estimate(const vk_hip_csr *A, int iterations, double *result);

/* ---- scaling and postsolve passes on the device --------------------------
 * The O(nnz) host passes around the reference's PDHG loop, run on the
 * device with the host loops' arithmetic in their order (one thread per row,
 * no contraction), so the results are bitwise equal to those loops. */

/* Replaces continuous_common.hpp:133-205 scale_linear_model (called at
 * pdhg.cpp:25).  Scales A in place, starting from the values last given to
 * vk_hip_csr_set_values (kept on the device for
 * vk_hip_csr_postsolve_products); A' and the FP32 copies follow:
 *   1. row i by r_i = 1/max_j |a_ij| when that maximum exceeds 1 (else 1),
 *      column j by s_j likewise; c*s, l/s, u/s, rl*r, ru*r (infinite bounds
 *      stay infinite);
 *   2. `ruiz_passes` Ruiz passes: rows and columns divided by the square
 *      root of their largest |a| (1 if zero);
 *   3. if `pock_chambolle` != 0, one Pock-Chambolle pass (alpha = 1): by the
 *      square root of their sum of |a| (1 if zero).
 * Steps 2-3 divide r, s, c, rl, ru by the factors and multiply l, u.  Host
 * arrays: c, l, u (n), rl, ru (m) in; row_scale (m), col_scale (n) and the
 * scaled c_out, l_out, u_out (n), rl_out, ru_out (m) out. */
int vk_hip_lp_scale(vk_hip_csr *A, const double *c, const double *l, const double *u,
                    const double *rl, const double *ru, int ruiz_passes, int pock_chambolle,
                    double *row_scale, double *col_scale, double *c_out, double *l_out,
                    double *u_out, double *rl_out, double *ru_out);

/* Variable bounds of `count` bound sets (l, u: count * cols, set-major)
 * scaled as the last vk_hip_lp_scale on A scaled its l and u (the same
 * operations in the same order, so bitwise the same result): for branch-
 * and-bound nodes that share A and differ only in bounds. */
int vk_hip_lp_scale_bounds(const vk_hip_csr *A, int count, const double *l, const double *u,
                           double *l_out, double *u_out);

/* Replaces continuous_common.hpp:207-225 scaled_operator_bound (pdhg.cpp:75):
 * sqrt(max_i sum_j |a_ij| * max_j sum_i |a_ij|) over A's current values. */
int vk_hip_csr_operator_bound(const vk_hip_csr *A, double *bound);
/* The same bound for each diagonal block of a block-diagonal A (block i:
 * rows [row_starts[i], row_starts[i+1]), columns [col_starts[i],
 * col_starts[i+1])), as scaled_operator_bound gives for that block alone
 * (pdhg.cpp:75 per batched instance).  bounds: count doubles. */
int vk_hip_csr_operator_bounds(const vk_hip_csr *A, int count, const int64_t *row_starts,
                               const int64_t *col_starts, double *bounds);

/* Replaces the O(nnz) loops of the reference postsolve: linear_values
 * (continuous_common.hpp:302-315, called at pdhg.cpp:168) and the
 * stationarity accumulation (pdhg.cpp:188-196), on the values last given to
 * vk_hip_csr_set_values (unscaled):
 *   ax = A x   (each row summed from 0 in storage order),
 *   g  = c + A'y (each column from c_j, its rows in increasing order).
 * Host arrays x, c, g (n) and y, ax (m). */
int vk_hip_csr_postsolve_products(const vk_hip_csr *A, const double *x, const double *y,
                                  const double *c, double *ax, double *g);

/* ---- fused PDHG (new: the reference drives PDHG op by op, pdhg.cpp:87-151) */

enum {
    VK_PDHG_POLICY_REFERENCE = 0,  /* pdhg.cpp:75-150 step/restart rule      */
    VK_PDHG_POLICY_ADAPTIVE = 1,   /* restarted PDHG with averaging and a
                                      primal weight (KKT restart criteria)   */
    VK_PDHG_POLICY_HALPERN = 2     /* restarted Halpern PDHG with reflection
                                      and a primal weight (fixed-point
                                      residual restart criteria)             */
};

enum {
    VK_PDHG_RUNNING = 0,
    VK_PDHG_CONVERGED = 1,         /* termination test met                   */
    VK_PDHG_ITERATION_LIMIT = 2,
    VK_PDHG_PRIMAL_INFEASIBLE = 3  /* adaptive / Halpern: the dual iterate
                                      y is a Farkas certificate (its dual
                                      objective without c is positive)     */
};

enum {
    VK_PREC_FP64 = 0,              /* A, iterates and all vectors in FP64    */
    VK_PREC_MIXED = 1              /* A and iterates FP32 (compensated row
                                      sums); reductions and control in FP64  */
};

typedef struct {
    int policy;                    /* VK_PDHG_POLICY_*                         */
    int precision;                 /* VK_PREC_*                                */
    int64_t iteration_limit;       /* SolverOptions::iteration_limit           */
    double metric_tolerance;       /* reference test: max(r_p, r_d) <= this;
                                      max(primal_tol, dual_tol) (pdhg.cpp:147) */
    double relative_kkt_tolerance; /* > 0: stop on relative KKT <= this
                                      instead of the reference test          */
    double step_size;              /* reference: 0.99/max(1, bound) (pdhg.cpp:76);
                                      adaptive: eta, with tau = eta/w, sigma = eta*w */
    double primal_weight;          /* adaptive: initial w (> 0)               */
    int64_t restart_period;        /* adaptive: iterations between restart
                                      checks (default 64)                     */
    double norm_b;                 /* ||bhat||_2 and ||c||_2 in original      */
    double norm_c;                 /*   coordinates (relative KKT denominators) */
    int use_graphs;                /* capture iteration blocks in hipGraphs   */
    int grid_blocks;               /* 0 = auto; else blocks per SpMV launch   */
    int col_blocks;                /* A x layout: 0 = auto (4 column blocks
                                      once x outgrows 2 MiB, else none);
                                      1 = plain row tiles; 2..16 = that many
                                      column blocks (one XCD L2 slice each) */
    int split_passes;              /* 0 = auto (split once m or n > 2^18);
                                      1 = fused SpMV + epilogue kernels;
                                      2 = SpMV-only kernels, then row and
                                      column passes (same results bitwise) */
    int cache_budget_mb;           /* MiB of matrix data (A x operand first,
                                      then A') loaded so that the 256 MB MALL
                                      keeps it across iterations; the rest is
                                      streamed.  0 = auto (256 MB less the
                                      gathered x and y less 4 MiB), -1 = none.
                                      Load hints only: results do not change */
    int spmv_panels;               /* panel layout (LDS-staged gathers): 0 =
                                      auto (once m or n > 2^18, if the matrix
                                      fits the layout), 1 = never, 2 = always
                                      (VK_EINVAL when it does not fit).  When
                                      used it replaces col_blocks/split_passes */
    int panel_groups;              /* column groups of A x (0 = auto; 1, 2, 4, 8):
                                      partial sums per group, added in order */
    int panel_groups_t;            /* the same for A' y                       */
    int panel_threads;             /* lanes per panel workgroup: 0 = auto,
                                      256 (48 KB panels only) or 512 (results
                                      do not depend on it)                    */
    int panel_tune;                /* 1: time the automatic shape and, with
                                      its group count, 48 KB panels with 512
                                      and with 256 lanes, each over the unroll
                                      depths around its automatic one, on the
                                      matrix at setup and keep the fastest
                                      (results do not depend on them)         */
    int panel_kb;                  /* KB per staged panel, both operands: 0 =
                                      auto (per operand, with the group count
                                      when that is auto too), 48, 24 or 16
                                      (results do not depend on it)          */
    int tile_geometry;             /* batched solver: row-tile geometry of the
                                      instances' tiles, 0..3 as in
                                      vk_hip_csr_set_tile_geometry (the
                                      single-LP solver uses its matrix's)    */
    int recentre;                  /* 1 (Halpern, VK_PREC_MIXED, single LP or
                                      QP only): store the FP32 iterates as
                                      corrections to an FP64 base point that
                                      every restart moves to the restart
                                      point, with the LP shifted to the base
                                      (rounded to FP32 once per restart), so
                                      that FP32 rounding is relative to the
                                      correction and the FP32 iteration does
                                      not stall near 1e-5 relative KKT;
                                      download returns the total point.
                                      0 = plain FP32 iterates              */
    double primal_weight_gain;     /* adaptive / Halpern: proportional gain of
                                      the primal-weight update at restarts,
                                      log w <- log w - gain e - integral sum(e),
                                      e = log w - log(|dy| / |dx|); 0 = 1/2
                                      (PDLP's smoothing)                     */
    double primal_weight_integral; /* integral gain (0 = none)               */
} vk_pdhg_params;

typedef struct {
    int64_t iterations;            /* completed iterations                    */
    int status;                    /* VK_PDHG_*                               */
    int restarts;                  /* adaptive restarts taken                 */
    double tau, sigma, primal_weight;
    double metric;                 /* reference metric max(r_p, r_d), scaled  */
    double primal_scaled, dual_scaled;   /* r_p, r_d (infinity norms)        */
    double rel_primal, rel_dual, rel_gap, rel_kkt;  /* original coordinates   */
    double primal_objective, dual_objective;        /* without offset         */
    int converged_on_average;      /* adaptive: solution is the average      */
    int reserved;
} vk_pdhg_state;

typedef struct vk_hip_pdhg vk_hip_pdhg;

/* Allocates the solver on the device for the SCALED LP (the caller scales,
 * like pdhg.cpp:25): A (values already set), objective c, variable bounds
 * l,u, row bounds rl,ru (+-inf allowed), row_scale r and col_scale s
 * (x = s*xhat, y = r*yhat, used for original-coordinate KKT), and the scaled
 * start x0 (NULL = 0) and y0 (NULL = 0).  x0 is projected onto [l,u]
 * (pdhg.cpp:71).  In VK_PREC_MIXED the FP32 copies are made on the device. */
int vk_hip_pdhg_create(const vk_hip_csr *A, const double *c, const double *l,
                       const double *u, const double *rl, const double *ru,
                       const double *row_scale, const double *col_scale,
                       const double *x0, const double *y0,
                       const vk_pdhg_params *params, vk_hip_pdhg **out);
/* PDQP (pdqp.cpp:110-176): as vk_hip_pdhg_create with the SCALED quadratic
 * objective Q (n x n, values q_ij s_i s_j, stored symmetrically).  One more
 * SpMV per iteration forms Q x+ ahead of the primal kernel.  Either the
 * reference policy in FP64 (the caller's step 0.99 / max(1, bound^2 +
 * |Q|_inf), pdqp.cpp:98; no relative-KKT test) or the Halpern policy (new:
 * step_size = eta = 0.99 / ||A||_2 as for an LP, sigma = eta w, and column
 * j's primal step 1 / (w / eta + sum_k |q_jk|); the relative KKT has the
 * 1/2 x'Qx objective terms) in FP64 or, with recentre = 1, in VK_PREC_MIXED
 * (corrections to an FP64 base whose Q xbase joins the gradient offset).
 * Q = NULL is vk_hip_pdhg_create. */
int vk_hip_pdqp_create(const vk_hip_csr *A, const vk_hip_csr *Q, const double *c,
                       const double *l, const double *u, const double *rl, const double *ru,
                       const double *row_scale, const double *col_scale, const double *x0,
                       const double *y0, const vk_pdhg_params *params, vk_hip_pdhg **out);
/* Runs up to `max_iterations` more iterations (or until the device stops),
 * polling the device state every iteration block.  Returns the VK_PDHG_*
 * status through `status` (not an error).  run, read_state and download may
 * be called from any host thread; calls on one solver must not overlap,
 * while different solvers (each with its own streams) may run at once. */
int vk_hip_pdhg_run(vk_hip_pdhg *solver, int64_t max_iterations, int *status);
int vk_hip_pdhg_read_state(vk_hip_pdhg *solver, vk_pdhg_state *state);
/* Scaled iterates (FP64 on the host) of the current solution. */
int vk_hip_pdhg_download(vk_hip_pdhg *solver, double *x, double *y);
/* Milliseconds of device time of the last vk_hip_pdhg_run, measured with
 * HIP events on the solver's stream. */
int vk_hip_pdhg_last_run_ms(const vk_hip_pdhg *solver, double *ms);
/* Runs `iterations` more iterations launched one by one, with HIP events on
 * the solver's stream between the kernels.  ms[0] = mean duration of the
 * dual SpMV kernel (the fused dual kernel when not split), ms[1] = the dual
 * row pass (0 when fused), ms[2] = the primal kernel (A' y with its column
 * pass: one kernel in every layout), ms[3] = 0, ms[4] = mean device time
 * per whole iteration.  For measurement. */
int vk_hip_pdhg_profile(vk_hip_pdhg *solver, int64_t iterations, double *ms);
/* The A x layout the solver chose: column blocks (1 = plain row tiles) and
 * whether the SpMVs run split from their row / column passes (0 or 1). */
int vk_hip_pdhg_layout(const vk_hip_pdhg *solver, int *col_blocks, int *split_passes);
/* The panel layouts the solver uses (zeros when it does not): info[0..4] =
 * groups, lanes per workgroup, most entries per lane and panel, the
 * kernel's unrolled entries per lane and panel, and KB per panel, of A x;
 * info[5..9] the same for A' y. */
int vk_hip_pdhg_panel_layout(const vk_hip_pdhg *solver, int *info);
int vk_hip_pdhg_destroy(vk_hip_pdhg *solver);

/* Test hook (new; the reference's device errors become exceptions,
 * cuda_ops.cu:35-46, and then NumericalFailure, Solver.cpp:81-86): after
 * `after` more successful device allocations of this library, every
 * allocation fails with hipErrorOutOfMemory until the hook is called again
 * with after < 0 (never fail).  *made (may be NULL) receives the number of
 * device allocations made since the previous call.  For tests of the
 * allocation-failure paths. */
int vk_hip_debug_fail_allocations(int64_t after, int64_t *made);

/* Row-tile geometry (new; the reference hands SpMV to cuSPARSE, which picks
 * its own, cuda_ops.cu:473-528).  The tile kernels cut rows into tiles of
 * whole rows; the geometry bounds a tile's nonzeros and rows and sets the
 * row length above which a row has a tile of its own (reduced by the whole
 * wave):
 *   0  512 nonzeros, 64 rows, rows over 512 alone (the default)
 *   1  256 nonzeros, 64 rows, rows over 256 alone
 *   2  512 nonzeros, 32 rows, rows over 512 alone
 *   3  512 nonzeros, 64 rows, rows over 64 alone
 * It sets which rows share a tile and so the order of each row's sum: the
 * results of SpMVs and solves on the matrix depend on it bitwise (each
 * geometry is exact to round-off).  vk_hip_csr_set_tile_geometry re-cuts A's
 * and A' tiles; call it before creating solvers on the matrix.
 * vk_hip_csr_tune_tiles times one A x and one A' y (`repeats` times, in the
 * storage precision of `precision`) with each geometry, writes the mean ms
 * of each to ms[0..3] (may be NULL), leaves the fastest set on the matrix
 * and returns its index in *best. */
int vk_hip_csr_set_tile_geometry(vk_hip_csr *A, int geometry);
int vk_hip_csr_tile_geometry(const vk_hip_csr *A, int *geometry);
int vk_hip_csr_tune_tiles(vk_hip_csr *A, int precision, int repeats, int *best, double *ms);

/* Layout tuner (new; the reference hands SpMV to cuSPARSE,
 * cuda_ops.cu:473-528, and has no layout choice): times y = A x in the
 * storage precision of `precision` on the plain row tiles and on 2, 4, 8 and
 * 16 column blocks (`repeats` launches each), writes the mean milliseconds
 * of each candidate to ms[0..4] (may be NULL; +inf where a candidate does
 * not apply) and the fastest block count to *best, for
 * vk_pdhg_params.col_blocks.  The choice changes only the summation order
 * of A x, never the iteration's results beyond round-off. */
int vk_hip_csr_tune_col_blocks(vk_hip_csr *A, int precision, int repeats, int *best,
                               double *ms);

/* Panel SpMV (new; the layout behind vk_pdhg_params.spmv_panels): builds
 * the panel layout of A (transpose = 0; rows x cols) or A' (transpose = 1)
 * with `groups` column groups (1, 2, 4 or 8), panels of `panel_kb` KB (48,
 * 24, 16; 0 = the solver's automatic choice for that group count) and
 * `threads` lanes per workgroup (512, or 256 with 48 KB panels), in FP64 or,
 * for VK_PREC_MIXED, FP32 values and x (x rounded to FP32 first), and
 * writes the FP64 partial sums of every group to the host array `partial`
 * (groups x rows, group-major): partial[g*rows+r] is the sequential sum, in
 * column order, of row r's products over group g's columns (the same bits
 * for every panel size and lane count).  `repeats` > 0 also times that many
 * launches (mean ms into *ms, may be NULL).  VK_EINVAL when the operand does
 * not fit the layout. */
int vk_hip_csr_panel_spmv(const vk_hip_csr *A, int transpose, int precision, int groups, int panel_kb,
                          int threads, const double *x, double *partial, int repeats, double *ms);

/* ---- batched PDHG (new: one persistent launch for many LPs) ------------
 *
 * Replaces, for a set of relaxations that differ only in data, the serial
 * loop of solve_lp calls the branch-and-bound engine makes
 * (src/solver/mip/BranchAndBound.cpp:296-297; one backend and one solve per
 * node, src/solver/Solver.cpp:37-51).  Instance i owns rows
 * [row_starts[i], row_starts[i+1]) and columns [col_starts[i], col_starts[i+1])
 * of one block-diagonal CSR matrix (global indices; a row may only touch its
 * instance's columns).  All vectors are the SCALED data concatenated in the
 * same order; row_scale/col_scale give original coordinates.  step_sizes[i]
 * <= 0 (or NULL) asks the device for ||A_i||_2 by power iteration;
 * primal_weights, norm_b, norm_c are per instance (NULL = 1, 0, 0);
 * iteration_limits[i] (NULL = params->iteration_limit for all) bounds
 * instance i's iterations.  params is shared (its step_size, primal_weight,
 * norm_b, norm_c, col_blocks and split_passes are ignored). */
typedef struct vk_hip_batch vk_hip_batch;

int vk_hip_batch_create(int count, const int64_t *row_starts, const int64_t *col_starts,
                        const int64_t *row_offsets, const int64_t *column_indices,
                        const double *values, const double *c, const double *l,
                        const double *u, const double *rl, const double *ru,
                        const double *row_scale, const double *col_scale,
                        const double *x0, const double *y0, const double *step_sizes,
                        const double *primal_weights, const double *norm_b,
                        const double *norm_c, const int64_t *iteration_limits,
                        const vk_pdhg_params *params, vk_hip_batch **out);
/* The same, with the block-diagonal matrix given as a CSR handle whose
 * values are set (typically scaled in place by vk_hip_lp_scale): its
 * pattern, transpose and values are copied on the device, and every row is
 * checked to stay inside its instance's columns (VK_EINVAL otherwise). */
int vk_hip_batch_create_csr(int count, const int64_t *row_starts, const int64_t *col_starts,
                            const vk_hip_csr *A, const double *c, const double *l,
                            const double *u, const double *rl, const double *ru,
                            const double *row_scale, const double *col_scale,
                            const double *x0, const double *y0, const double *step_sizes,
                            const double *primal_weights, const double *norm_b,
                            const double *norm_c, const int64_t *iteration_limits,
                            const vk_pdhg_params *params, vk_hip_batch **out);
/* Node batch on one shared matrix: `count` instances that all have A's
 * matrix (B&B relaxations of one LP differ only in their bounds,
 * BranchAndBound.cpp:296-297).  A is stored once; every vector argument is
 * per instance, instance-major (c, l, u, col_scale, x0: count * cols;
 * rl, ru, row_scale, y0: count * rows), so the instances may also differ in
 * objective and row bounds.  Otherwise as vk_hip_batch_create. */
int vk_hip_batch_create_shared(int count, const vk_hip_csr *A, const double *c, const double *l,
                               const double *u, const double *rl, const double *ru,
                               const double *row_scale, const double *col_scale,
                               const double *x0, const double *y0, const double *step_sizes,
                               const double *primal_weights, const double *norm_b,
                               const double *norm_c, const int64_t *iteration_limits,
                               const vk_pdhg_params *params, vk_hip_batch **out);
/* Launches the persistent kernel up to `max_launches` times, each running
 * `iterations_per_launch` iterations of every instance still running; stops
 * early when all have stopped.  *running = instances still running. */
int vk_hip_batch_run(vk_hip_batch *batch, int64_t iterations_per_launch,
                     int64_t max_launches, int *running);
int vk_hip_batch_read_state(vk_hip_batch *batch, int instance, vk_pdhg_state *state);
/* Scaled iterates of all instances, concatenated (FP64 on the host). */
int vk_hip_batch_download(vk_hip_batch *batch, double *x, double *y);
int vk_hip_batch_last_run_ms(const vk_hip_batch *batch, double *ms);
int vk_hip_batch_destroy(vk_hip_batch *batch);

/* ---- unrolled PDHG and its adjoint (new: the reference has no
 * differentiation) ---------------------------------------------------------
 *
 * K steps of the reference iteration (pdhg.cpp:92-113) with tau and sigma
 * fixed and no restart, on the SCALED LP (A with values set, objective c,
 * variable bounds l,u, row bounds rl,ru), saving (x, y) every
 * checkpoint_every steps; the backward pass recomputes each segment from its
 * checkpoint, records the projections' active sets and runs the adjoint
 * recursion in reverse (A and A' SpMVs swapped).  All vectors are scaled
 * FP64 on the host. */
typedef struct vk_hip_unrolled vk_hip_unrolled;

int vk_hip_unrolled_create(const vk_hip_csr *A, const double *c, const double *l,
                           const double *u, const double *rl, const double *ru,
                           double tau, double sigma, int64_t iterations,
                           int64_t checkpoint_every, vk_hip_unrolled **out);
/* Runs the K steps from (x0, y0); x_out / y_out (may be NULL) get (x_K, y_K). */
int vk_hip_unrolled_forward(vk_hip_unrolled *u, const double *x0, const double *y0,
                            double *x_out, double *y_out);
/* Given dL/dx_K and dL/dy_K (NULL = 0), writes dL/dc, dL/drl and dL/dru
 * through the K steps (any output may be NULL).  Needs a forward pass first;
 * may be called again with other seeds. */
int vk_hip_unrolled_backward(vk_hip_unrolled *u, const double *grad_x, const double *grad_y,
                             double *grad_c, double *grad_row_lower, double *grad_row_upper);
/* Replaces the scaled objective and row bounds (same A, bounds of the same
 * kind); the next backward needs a new forward pass. */
int vk_hip_unrolled_set_data(vk_hip_unrolled *u, const double *c, const double *rl,
                             const double *ru);
/* Device milliseconds of the last forward or backward call (HIP events). */
int vk_hip_unrolled_last_ms(const vk_hip_unrolled *u, double *ms);
int vk_hip_unrolled_destroy(vk_hip_unrolled *u);

#ifdef __cplusplus
}
#endif
#endif
