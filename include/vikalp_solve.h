/*
 * vikalp_solve.h — C ABI of the host library libvikalp.so (C++20).
 *
 * A flat-argument form of vikalp::solve / vikalp::solver::solve_lp for FFI
 * callers (ctypes, cgo, JNI).  It replaces, for the LP path:
 *   vikalp::solve(model, options)            src/solver/Solver.cpp:92-99
 *   vikalp::solver::solve_lp(model, backend, options, warm, lo, hi)
 *                                            src/solver/lp/pdhg.cpp:269-293
 * with the backend fixed to the HIP ExecutionBackend (make_hip_backend()).
 * The model is the reference's canonical LP (include/vikalp/contracts/Model.hpp):
 * CSR with int64 offsets/indices, +-inf bounds allowed.
 */
#ifndef VIKALP_SOLVE_H
#define VIKALP_SOLVE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* vikalp::SolverOptions (include/vikalp/contracts/SolverOptions.hpp:15-28),
 * enums as ints in declaration order, then the HIP-path options. */
typedef struct {
    int backend;                    /* BackendPreference: 0 Auto, 1 CPU, 2 CUDA, 3 HIP */
    int continuous_method;          /* 0 Auto, 1 PDHG, 2 RevisedSimplex, 3 InteriorPoint */
    int64_t iteration_limit;
    int64_t node_limit;
    double time_limit_seconds;
    double primal_tolerance;
    double dual_tolerance;
    double integrality_tolerance;
    double absolute_gap_tolerance;
    double relative_gap_tolerance;
    int deterministic;
    int pdhg_policy;                /* 0 Reference, 1 Adaptive, 2 Halpern */
    int pdhg_precision;             /* 0 FP64, 1 Mixed */
    int fused_iterations;
    int use_graphs;
    int reserved;
    double relative_kkt_tolerance;
    double mixed_phase_tolerance;
    int64_t restart_period;
    int spmv_col_blocks;            /* 0 auto, 1 plain tiles, 2..16 column blocks */
    int spmv_autotune;              /* time the layouts and use the fastest */
    int spmv_split_passes;          /* 0 auto, 1 fused, 2 split SpMV + passes */
    int spmv_cache_budget_mb;       /* MALL-resident matrix MiB: 0 auto, -1 none */
    int spmv_panels;                /* panel layout: 0 auto, 1 never, 2 always */
    int spmv_panel_groups;          /* A x column groups: 0 auto, 1, 2, 4, 8 */
    int spmv_panel_groups_t;        /* A' y column groups: 0 auto, 1, 2, 4, 8 */
    int spmv_panel_threads;         /* lanes per panel workgroup: 0 auto, 256, 512 */
    int spmv_panel_kb;              /* KB per staged panel: 0 auto, 48, 24, 16 */
    int spmv_tile_geometry;         /* row tiles: 0..3 (vk_hip_csr_set_tile_geometry), -1 tune */
    int mixed_recentre;             /* mixed Halpern FP32 phase as corrections to an FP64 base */
    int reserved2;
    double pdhg_primal_weight_gain;     /* 0 = 1/2 */
    double pdhg_primal_weight_integral; /* 0 = none */
} vk_solver_options;

/* Fills `o` with the defaults of vikalp::SolverOptions. */
void vk_solver_options_default(vk_solver_options *o);

/* Scalar fields of vikalp::SolveResult (SolveResult.hpp:26-61) + the
 * appended relative-KKT and device-time fields. */
typedef struct {
    int status;                     /* SolveStatus in declaration order */
    int reserved;
    int64_t iterations;
    int64_t low_precision_iterations;
    double objective_value;
    double primal_bound, dual_bound;
    double primal_residual;
    double stationarity_residual;
    double complementarity_residual;
    double absolute_gap;
    double relative_gap;
    double solve_seconds;
    double device_seconds;
    double kkt_relative_primal, kkt_relative_dual, kkt_relative_gap, kkt_relative;
    double kkt_primal_objective, kkt_dual_objective;
    char backend[16];
    char solver[32];
    char message[160];
} vk_solve_info;

/* Solves the LP.  `mode` 0 = vikalp::solve (router: validates the model,
 * honours options->backend), 1 = solver::solve_lp on a HIP backend directly
 * (honours warm_start and the bound overrides; any may be NULL).
 * x_out (n) receives the primal solution, y_out (m) the row duals in the
 * reference sign (constraint_upper_dual - constraint_lower_dual); either may
 * be NULL.  Returns 0 when a result was produced (see info->status),
 * negative when the arguments were unusable. */
int vk_lp_solve(int mode, int64_t m, int64_t n, const int64_t *row_offsets,
                const int64_t *column_indices, const double *values,
                const double *objective, double objective_offset,
                const double *row_lower, const double *row_upper,
                const double *var_lower, const double *var_upper,
                const double *warm_start, const double *lower_override,
                const double *upper_override, const vk_solver_options *options,
                double *x_out, double *y_out, vk_solve_info *info);

/* Batched solve (solver::solve_lp_batch): `count` independent LPs given as
 * one block-diagonal CSR with global indices; instance i owns rows
 * [row_starts[i], row_starts[i+1]) and columns [col_starts[i], col_starts[i+1]).
 * Vectors are concatenated in the same order.  x_out (total columns) and
 * y_out (total rows) receive the concatenated solutions, info[count] the
 * per-instance results.  Returns 0 when results were produced. */
int vk_lp_solve_batch(int count, const int64_t *row_starts, const int64_t *col_starts,
                      const int64_t *row_offsets, const int64_t *column_indices,
                      const double *values, const double *objective,
                      const double *row_lower, const double *row_upper,
                      const double *var_lower, const double *var_upper,
                      const vk_solver_options *options, double *x_out, double *y_out,
                      vk_solve_info *info);

/* Branch-and-bound node relaxations of one LP (solver::solve_lp_nodes):
 * replaces the per-node RelaxationOracle::solve calls of the node loop
 * (src/solver/mip/BranchAndBound.cpp:296-297; Solver.cpp:37-51 builds a
 * backend per node) with one batch on one device-resident matrix.  The LP
 * is given as in vk_lp_solve; node i uses node_lower[i*n .. i*n+n) and
 * node_upper[...] as its variable bounds (NULL: the LP's own for every
 * node) and warm_start[i*n ...] (NULL: cold starts).  x_out (count * n),
 * y_out (count * m, reference sign) and info (count) per node.  Returns 0,
 * or -1 for unusable arguments. */
int vk_lp_solve_nodes(int64_t m, int64_t n, const int64_t *row_offsets,
                      const int64_t *column_indices, const double *values,
                      const double *objective, double objective_offset,
                      const double *row_lower, const double *row_upper,
                      const double *var_lower, const double *var_upper, int count,
                      const double *node_lower, const double *node_upper,
                      const double *warm_start, const vk_solver_options *options,
                      double *x_out, double *y_out, vk_solve_info *info);

/* Measurement: builds the same scaled LP on the HIP backend, runs
 * `warm_iterations` iterations, then `iterations` more launched one by one
 * with HIP events around the two fused kernels on the solver's stream.
 * out[0] = mean dual-kernel ms (A x with the dual update), out[1] = mean
 * primal-kernel ms (A'y with the residuals and primal step), out[2] = mean
 * ms per iteration, out[3] = the step size used, out[4] = column blocks
 * used (1 = plain row tiles), out[5] = 1 when the SpMVs ran split from
 * their passes, out[6..9] = mean ms of the dual SpMV (or fused dual kernel),
 * the dual row pass, the primal SpMV (or fused primal kernel) and the primal
 * column pass (vk_hip_pdhg_profile), out[10..19] = the panel layout used
 * (vk_hip_pdhg_panel_layout; zeros when none).  out holds 20 doubles.
 * Returns 0 on success. */
int vk_lp_profile(int64_t m, int64_t n, const int64_t *row_offsets,
                  const int64_t *column_indices, const double *values,
                  const double *objective, const double *row_lower,
                  const double *row_upper, const double *var_lower,
                  const double *var_upper, const vk_solver_options *options,
                  int64_t warm_iterations, int64_t iterations, double *out);

/* PDQP (solver::solve_qp, pdqp.cpp:303-331): vk_lp_solve with the quadratic
 * objective 0.5 x'Qx given as n x n CSR stored symmetrically (q_row_offsets
 * NULL = none).  mode 0 = vikalp::solve (routes QP models to solve_qp), 1 =
 * solve_qp on a HIP backend directly. */
int vk_qp_solve(int mode, int64_t m, int64_t n, const int64_t *row_offsets,
                const int64_t *column_indices, const double *values,
                const int64_t *q_row_offsets, const int64_t *q_column_indices,
                const double *q_values, const double *objective, double objective_offset,
                const double *row_lower, const double *row_upper,
                const double *var_lower, const double *var_upper,
                const double *warm_start, const double *lower_override,
                const double *upper_override, const vk_solver_options *options,
                double *x_out, double *y_out, vk_solve_info *info);

/* Message of the last failure reported by a negative return code of this
 * header's functions on the calling thread. */
const char *vk_solve_last_error(void);

/* ---- unrolled PDHG: forward / backward pair (new; solver::UnrolledPdhg) ----
 * K steps of the PDHG iteration (pdhg.cpp:92-113) with the reference's fixed
 * step (pdhg.cpp:75-78), no restarts, from the projected zero start, on the
 * LP in the canonical form of vk_lp_solve.  forward writes x_K (n) and y_K
 * (m, reference sign).  backward takes dL/dx_K and dL/dy_K (either may be
 * NULL = 0) and writes dL/dc (n), dL/d(row_lower) (m) and dL/d(row_upper)
 * (m) through the K steps; an equality row's dL/db is the sum of the two.
 * Checkpoints every checkpoint_every steps bound the memory.  Returns 0, -1
 * for unusable arguments, -2 with vk_solve_last_error() otherwise. */
typedef struct vk_unrolled vk_unrolled;

int vk_unrolled_create(int64_t m, int64_t n, const int64_t *row_offsets,
                       const int64_t *column_indices, const double *values,
                       const double *objective, const double *row_lower,
                       const double *row_upper, const double *var_lower,
                       const double *var_upper, int64_t iterations,
                       int64_t checkpoint_every, vk_unrolled **out);
int vk_unrolled_forward(vk_unrolled *u, double *x, double *y);
int vk_unrolled_backward(vk_unrolled *u, const double *grad_x, const double *grad_y,
                         double *grad_c, double *grad_row_lower, double *grad_row_upper);
/* New objective and row bounds for the same matrix (each row keeps which of
 * its bounds are finite); the matrix, scaling and step are reused. */
int vk_unrolled_set_data(vk_unrolled *u, const double *objective, const double *row_lower,
                         const double *row_upper);
/* The fixed step and the device milliseconds of the last forward/backward. */
int vk_unrolled_info(const vk_unrolled *u, double *step_size, double *last_device_ms);
int vk_unrolled_destroy(vk_unrolled *u);

/* ---- MPS input --------------------------------------------------------------
 * Replaces src/io/MpsReader.cpp:10-553 MpsReader::read
 * (include/vikalp/io/MpsReader.hpp:17): reads a free-format MPS file into a
 * model in the canonical form above.  Every file the reference reads gives
 * the same model; files it rejects get the same messages.  In addition the
 * standard RANGES section, MI / PL / LI / UI bounds, integer MARKER blocks
 * and comment lines are read.  Returns 0 and a model handle, 1 when the
 * file is rejected (vk_solve_last_error() lists every error as
 * "line N: message", one per line), -1 for null arguments. */
typedef struct vk_model vk_model;

int vk_mps_read(const char *path, vk_model **out);
int vk_model_dims(const vk_model *model, int64_t *m, int64_t *n, int64_t *nnz);
/* Copies into caller buffers; any pointer may be NULL to skip that array.
 * variable_types: 0 continuous, 1 integer, 2 binary (VariableType order). */
int vk_model_export(const vk_model *model, int64_t *row_offsets, int64_t *column_indices,
                    double *values, double *objective, double *row_lower, double *row_upper,
                    double *var_lower, double *var_upper, int32_t *variable_types,
                    double *objective_offset);
/* Branch-and-bound over batches of node relaxations (new; solver::
 * solve_mip_batched): the reference's engine (src/solver/mip/
 * BranchAndBound.cpp:222-465, one RelaxationOracle::solve per node) with up
 * to `batch` open nodes per round solved together by vk_lp_solve_nodes' path
 * on one device matrix.  variable_types: 0 continuous, 1 integer, 2 binary.
 * q_* (q_row_offsets NULL = none): a quadratic objective 0.5 x'Qx as for
 * vk_qp_solve; an MIQP is refused (info.status UnsupportedModel).
 * x_out (n; NULL to skip) gets the incumbent, info the result (iterations =
 * nodes processed), stats (NULL to skip) 9 counters: nodes processed,
 * pruned, integral, infeasible, incumbent updates, rounds, relaxations,
 * relaxation PDHG iterations, largest round.  Returns 0, or -1 for unusable
 * arguments. */
int vk_mip_solve_batched(int64_t m, int64_t n, const int64_t *row_offsets,
                         const int64_t *column_indices, const double *values,
                         const double *objective, double objective_offset,
                         const double *row_lower, const double *row_upper,
                         const double *var_lower, const double *var_upper,
                         const int32_t *variable_types, const int64_t *q_row_offsets,
                         const int64_t *q_column_indices, const double *q_values, int64_t batch,
                         const vk_solver_options *options, double *x_out,
                         vk_solve_info *info, int64_t *stats);

/* The model's NAME into buf (NUL-terminated, truncated to size); returns
 * the name's full length. */
int64_t vk_model_name(const vk_model *model, char *buf, int64_t size);
void vk_model_free(vk_model *model);

#ifdef __cplusplus
}
#endif
#endif
