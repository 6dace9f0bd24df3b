// C-callable wrappers around the reference's own CPU path, compiled from the
// reference sources where they lie (see oracle/Makefile) into
// oracle/_ref/libvikalp_ref.so.  TEST INFRASTRUCTURE ONLY: it exists to pin
// the plain-C oracle (pdhg_oracle.c) against the real implementation of
// src/solver/lp/pdhg.cpp + src/backend/cpu_ops.cpp, in this container.
//
// Only this file is ours; it includes the reference's public headers and
// calls vikalp::solver::solve_lp (include/vikalp/solver/continuous.hpp:10-15)
// with vikalp::make_cpu_backend() (include/vikalp/backend/CpuBackend.hpp:16).

#include "vikalp/backend/CpuBackend.hpp"
#include "vikalp/solver/continuous.hpp"

#include <cstdint>
#include <cstring>
#include <span>
#include <vector>

extern "C" {

// Output layout: x (n), y (m, = constraint_upper_dual - constraint_lower_dual),
// stats[0..7] = status, iterations, objective, primal_residual,
// stationarity_residual, complementarity_residual, absolute_gap, relative_gap.
int ref_solve_lp(int64_t m, int64_t n, const int64_t *row_offsets,
                 const int64_t *column_indices, const double *values,
                 const double *objective, double objective_offset,
                 const double *row_lower, const double *row_upper,
                 const double *var_lower, const double *var_upper,
                 const double *warm_start, const double *lower_override,
                 const double *upper_override, int64_t iteration_limit,
                 double time_limit_seconds, double primal_tolerance,
                 double dual_tolerance, double *x_out, double *y_out,
                 double *stats) {
    vikalp::Model model;
    model.name = "oracle-ref";
    model.linear_objective.assign(objective, objective + n);
    model.objective_offset = objective_offset;
    model.variable_lower.assign(var_lower, var_lower + n);
    model.variable_upper.assign(var_upper, var_upper + n);
    model.variable_types.assign(static_cast<std::size_t>(n),
                                vikalp::VariableType::Continuous);
    model.constraint_lower.assign(row_lower, row_lower + m);
    model.constraint_upper.assign(row_upper, row_upper + m);
    const int64_t nnz = row_offsets[m];
    model.constraint_matrix.pattern.rows = m;
    model.constraint_matrix.pattern.columns = n;
    model.constraint_matrix.pattern.row_offsets.assign(row_offsets, row_offsets + m + 1);
    model.constraint_matrix.pattern.column_indices.assign(column_indices,
                                                          column_indices + nnz);
    model.constraint_matrix.values.assign(values, values + nnz);

    vikalp::SolverOptions options;
    options.iteration_limit = iteration_limit;
    options.time_limit_seconds = time_limit_seconds;
    options.primal_tolerance = primal_tolerance;
    options.dual_tolerance = dual_tolerance;

    auto backend = vikalp::make_cpu_backend();
    std::span<const double> warm;
    std::span<const double> lo;
    std::span<const double> hi;
    if (warm_start) warm = std::span<const double>(warm_start, static_cast<std::size_t>(n));
    if (lower_override) lo = std::span<const double>(lower_override, static_cast<std::size_t>(n));
    if (upper_override) hi = std::span<const double>(upper_override, static_cast<std::size_t>(n));
    const auto result = vikalp::solver::solve_lp(model, *backend, options, warm, lo, hi);

    if (x_out && result.primal_solution.size() == static_cast<std::size_t>(n))
        std::memcpy(x_out, result.primal_solution.data(), sizeof(double) * n);
    if (y_out && result.constraint_upper_dual.size() == static_cast<std::size_t>(m)) {
        for (int64_t r = 0; r < m; ++r)
            y_out[r] = result.constraint_upper_dual[r] - result.constraint_lower_dual[r];
    }
    stats[0] = static_cast<double>(static_cast<int>(result.status));
    stats[1] = static_cast<double>(result.iterations);
    stats[2] = result.objective_value;
    stats[3] = result.primal_residual;
    stats[4] = result.stationarity_residual;
    stats[5] = result.complementarity_residual;
    stats[6] = result.absolute_gap;
    stats[7] = result.relative_gap;
    return static_cast<int>(result.status);
}

// vikalp::solver::solve_qp (src/solver/qp/pdqp.cpp:303-331) on the CPU
// backend; same outputs as ref_solve_lp, Q given as n x n CSR.
int ref_solve_qp(int64_t m, int64_t n, const int64_t *row_offsets, const int64_t *column_indices,
                 const double *values, const int64_t *q_row_offsets, const int64_t *q_column_indices,
                 const double *q_values, const double *objective, double objective_offset,
                 const double *row_lower, const double *row_upper, const double *var_lower,
                 const double *var_upper, const double *warm_start, const double *lower_override,
                 const double *upper_override, int64_t iteration_limit, double time_limit_seconds,
                 double primal_tolerance, double dual_tolerance, double *x_out, double *y_out,
                 double *stats) {
    vikalp::Model model;
    model.name = "oracle-ref-qp";
    model.linear_objective.assign(objective, objective + n);
    model.objective_offset = objective_offset;
    model.variable_lower.assign(var_lower, var_lower + n);
    model.variable_upper.assign(var_upper, var_upper + n);
    model.variable_types.assign(static_cast<std::size_t>(n), vikalp::VariableType::Continuous);
    model.constraint_lower.assign(row_lower, row_lower + m);
    model.constraint_upper.assign(row_upper, row_upper + m);
    const int64_t nnz = row_offsets[m];
    model.constraint_matrix.pattern.rows = m;
    model.constraint_matrix.pattern.columns = n;
    model.constraint_matrix.pattern.row_offsets.assign(row_offsets, row_offsets + m + 1);
    model.constraint_matrix.pattern.column_indices.assign(column_indices, column_indices + nnz);
    model.constraint_matrix.values.assign(values, values + nnz);
    const int64_t qnnz = q_row_offsets[n];
    model.quadratic_objective.pattern.rows = n;
    model.quadratic_objective.pattern.columns = n;
    model.quadratic_objective.pattern.row_offsets.assign(q_row_offsets, q_row_offsets + n + 1);
    model.quadratic_objective.pattern.column_indices.assign(q_column_indices, q_column_indices + qnnz);
    model.quadratic_objective.values.assign(q_values, q_values + qnnz);

    vikalp::SolverOptions options;
    options.iteration_limit = iteration_limit;
    options.time_limit_seconds = time_limit_seconds;
    options.primal_tolerance = primal_tolerance;
    options.dual_tolerance = dual_tolerance;

    auto backend = vikalp::make_cpu_backend();
    std::span<const double> warm, lo, hi;
    if (warm_start) warm = std::span<const double>(warm_start, static_cast<std::size_t>(n));
    if (lower_override) lo = std::span<const double>(lower_override, static_cast<std::size_t>(n));
    if (upper_override) hi = std::span<const double>(upper_override, static_cast<std::size_t>(n));
    const auto result = vikalp::solver::solve_qp(model, *backend, options, warm, lo, hi);
    if (x_out && result.primal_solution.size() == static_cast<std::size_t>(n))
        std::memcpy(x_out, result.primal_solution.data(), sizeof(double) * n);
    if (y_out && result.constraint_upper_dual.size() == static_cast<std::size_t>(m))
        for (int64_t r = 0; r < m; ++r)
            y_out[r] = result.constraint_upper_dual[r] - result.constraint_lower_dual[r];
    stats[0] = static_cast<double>(static_cast<int>(result.status));
    stats[1] = static_cast<double>(result.iterations);
    stats[2] = result.objective_value;
    stats[3] = result.primal_residual;
    stats[4] = result.stationarity_residual;
    stats[5] = result.complementarity_residual;
    stats[6] = result.absolute_gap;
    stats[7] = result.relative_gap;
    return static_cast<int>(result.status);
}

// The reference CpuBackend's multiply (cpu_ops.cpp:393-460) on host arrays.
int ref_spmv(int64_t rows, int64_t cols, const int64_t *row_offsets,
             const int64_t *column_indices, const double *values, double alpha,
             const double *x, double beta, double *y, int transpose) {
    auto backend = vikalp::make_cpu_backend();
    vikalp::CsrPattern pattern;
    pattern.rows = rows;
    pattern.columns = cols;
    pattern.row_offsets.assign(row_offsets, row_offsets + rows + 1);
    pattern.column_indices.assign(column_indices, column_indices + row_offsets[rows]);
    auto matrix = backend->create_matrix(pattern);
    backend->set_values(std::span<const double>(values, static_cast<std::size_t>(row_offsets[rows])),
                        *matrix);
    const int64_t nx = transpose ? rows : cols;
    const int64_t ny = transpose ? cols : rows;
    auto xv = backend->create_vector(nx);
    auto yv = backend->create_vector(ny);
    backend->upload(std::span<const double>(x, static_cast<std::size_t>(nx)), *xv);
    backend->upload(std::span<const double>(y, static_cast<std::size_t>(ny)), *yv);
    backend->multiply(alpha, *matrix, *xv, beta, *yv,
                      transpose ? vikalp::Transpose::Yes : vikalp::Transpose::No);
    backend->download(*yv, std::span<double>(y, static_cast<std::size_t>(ny)));
    return 0;
}

}  // extern "C"
