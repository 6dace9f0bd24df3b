/*
 * CPU oracle for the PDHG LP path — TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference's single-threaded FP64 CPU
 * path (the CPU ExecutionBackend ops, the scaling helpers and the restarted
 * PDHG policy).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker.  The product
 * (vikalp_amd/) never links or calls it.
 *
 * Reference files restated (paths relative to the reference root):
 *   src/backend/cpu_ops.cpp            op semantics (axpby, norm_inf, SpMV N/T, BiCGSTAB)
 *   src/solver/continuous_common.hpp   scaling, operator bound, residual helpers
 *   src/solver/lp/pdhg.cpp             PDHG policy and postsolve
 *
 * Every function below names the lines it follows in pdhg_oracle.c.
 */
#ifndef VIKALP_PDHG_ORACLE_H
#define VIKALP_PDHG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* SolveStatus values, in the declaration order of
 * include/vikalp/contracts/SolveResult.hpp:11-24. */
enum {
    ORC_NOT_SOLVED = 0, ORC_OPTIMAL, ORC_LOCALLY_OPTIMAL, ORC_FEASIBLE,
    ORC_INFEASIBLE, ORC_UNBOUNDED, ORC_ITERATION_LIMIT, ORC_NODE_LIMIT,
    ORC_TIME_LIMIT, ORC_NUMERICAL_FAILURE, ORC_UNSUPPORTED_MODEL, ORC_INVALID_MODEL
};

typedef struct {
    int64_t iteration_limit;        /* SolverOptions.hpp:18  (default 100000) */
    double time_limit_seconds;      /* SolverOptions.hpp:20  (+inf = none)    */
    double primal_tolerance;        /* SolverOptions.hpp:22  (1e-7)           */
    double dual_tolerance;          /* SolverOptions.hpp:23  (1e-7)           */
    /* Not in the reference: when > 0 the loop also stops (Optimal) once the
     * relative KKT error of the iterate, in original coordinates, is <= this. */
    double relative_kkt_tolerance;
    /* Not in the reference: when > 0, evaluate the relative KKT error only
     * every this many iterations (it costs two extra SpMVs here). 0 = every. */
    int64_t kkt_check_period;
} orc_options;

typedef struct {
    int status;
    int64_t iterations;
    double objective_value;
    double primal_residual;
    double stationarity_residual;
    double complementarity_residual;
    double absolute_gap;
    double relative_gap;
    /* relative KKT of the returned iterate (original coordinates). */
    double rel_primal, rel_dual, rel_gap, rel_kkt;
    double primal_objective, dual_objective;
    /* final scaled-space step and the reference's own metric. */
    double tau, sigma, last_metric;
    double solve_seconds;
} orc_result;

/* An LP in the reference's canonical form (Model.hpp:104-134) without Q:
 * min c'x + offset  s.t.  rl <= A x <= ru,  l <= x <= u;  A is CSR with
 * int64 offsets/indices (Model.hpp Index). */
typedef struct {
    int64_t m, n;
    const int64_t *row_offsets;     /* m+1 */
    const int64_t *column_indices;  /* nnz */
    const double *values;           /* nnz */
    const double *objective;        /* n   */
    double objective_offset;
    const double *row_lower, *row_upper;   /* m */
    const double *var_lower, *var_upper;   /* n */
} orc_lp;

/* ---- CpuBackend op semantics (cpu_ops.cpp) ---- */
void orc_axpby(int64_t n, double alpha, const double *x, double beta, double *y);
double orc_dot(int64_t n, const double *x, const double *y);
double orc_norm_inf(int64_t n, const double *x);
void orc_project_box(int64_t n, double *x, const double *lo, const double *hi);
void orc_spmv(int64_t rows, int64_t cols, const int64_t *rp, const int64_t *ci,
              const double *v, double alpha, const double *x, double beta,
              double *y, int transpose);
/* returns converged flag; iterations and residual through pointers. */
int orc_bicgstab(int64_t n, const int64_t *rp, const int64_t *ci, const double *v,
                 const double *b, double *x, double tol, int64_t max_iter,
                 int64_t *iterations, double *residual);
/* 0 = valid; otherwise a negative code (cpu_ops.cpp:38-63). */
int orc_validate_pattern(int64_t rows, int64_t cols, int64_t offsets_size,
                         const int64_t *rp, int64_t indices_size, const int64_t *ci);

/* ---- scaling (continuous_common.hpp) ---- */
/* Writes scaled values (nnz), objective, bounds and the two scale vectors.
 * literal != 0 runs the reference's O(n*nnz) column loop verbatim; 0 runs an
 * O(nnz) pass that yields the same factors (max is exact). */
void orc_scale_lp(const orc_lp *lp, const double *lower, const double *upper,
                  int literal, double *values_out, double *objective_out,
                  double *row_lower_out, double *row_upper_out,
                  double *var_lower_out, double *var_upper_out,
                  double *row_scale, double *col_scale);
double orc_operator_bound(int64_t m, int64_t n, const int64_t *rp,
                          const int64_t *ci, const double *v);

/* ---- relative KKT in original coordinates (new definition) ---- */
void orc_relative_kkt(const orc_lp *lp, const double *x, const double *y,
                      double *rel_primal, double *rel_dual, double *rel_gap,
                      double *primal_obj, double *dual_obj);

/* ---- the full solve: pdhg.cpp:269-293 then :20-265 ----
 * warm_start may be NULL. lower/upper overrides may be NULL (use the LP's).
 * x_out (n) and y_out (m, original-coordinate dual in the reference sign,
 * y = upper_dual - lower_dual) may be NULL. */
int orc_solve_lp(const orc_lp *lp, const double *lower_override,
                 const double *upper_override, const double *warm_start,
                 const orc_options *options, double *x_out, double *y_out,
                 orc_result *result);

/* Runs exactly `iterations` iterations of the loop body (no termination
 * test) and reports wall seconds; used to time the CPU per iteration on
 * instances too large to solve to tolerance. */
double orc_time_iterations(const orc_lp *lp, int64_t iterations);

/* ---- PDQP: solve_qp (src/solver/qp/pdqp.cpp:303-331, :20-300) ----
 * Q (n x n CSR, symmetric) given by qrp/qci/qv.  Same outputs as
 * orc_solve_lp; the relative-KKT options are ignored (the QP loop stops on
 * the reference metric only).  A non-PSD Q gives UNSUPPORTED_MODEL
 * (continuous_common.hpp:242-275, dense). */
int orc_solve_qp(const orc_lp *lp, const int64_t *qrp, const int64_t *qci, const double *qv,
                 const double *lower_override, const double *upper_override, const double *warm_start,
                 const orc_options *options, double *x_out, double *y_out, orc_result *result);

/* Unrolled fixed-step PDHG and its adjoint (not in the reference, which has
 * no differentiation; see pdhg_oracle.c for the recursion).  Runs
 * `iterations` steps of pdhg.cpp:92-113 with tau = sigma = 0.99/max(1,bound)
 * fixed and no restart from the projected zero start, on the scaled LP.
 * x_out (n) / y_out (m) get x_K, y_K in original coordinates (y in the
 * reference sign); step_out the step.  When grad_x or grad_y is given
 * (dL/dx_K, dL/dy_K in original coordinates, either may be NULL = 0),
 * grad_c (n), grad_row_lower (m) and grad_row_upper (m) receive dL/dc,
 * dL/d(row_lower), dL/d(row_upper) through the K iterations (the explicit
 * dependence of L on c is the caller's).  Returns 0, or -1 for bad input. */
int orc_unrolled(const orc_lp *lp, int64_t iterations, const double *grad_x, const double *grad_y,
                 double *x_out, double *y_out, double *grad_c, double *grad_row_lower,
                 double *grad_row_upper, double *step_out);

#ifdef __cplusplus
}
#endif
#endif
