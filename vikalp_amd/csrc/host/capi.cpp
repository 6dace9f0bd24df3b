// Flat C ABI over vikalp::solve / solver::solve_lp (include/vikalp_solve.h).

#include "vikalp_solve.h"

#include "vikalp/backend/HipBackend.hpp"
#include "vikalp/io/MpsReader.hpp"
#include "vikalp/solver/Solver.hpp"
#include "vikalp/solver/continuous.hpp"

#include <algorithm>
#include <cstring>
#include <exception>
#include <memory>
#include <span>
#include <string>
#include <vector>

namespace {

using namespace vikalp;

thread_local std::string g_last_error;

SolverOptions to_options(const vk_solver_options &o) {
    SolverOptions s;
    s.backend = static_cast<BackendPreference>(o.backend);
    s.continuous_method = static_cast<ContinuousMethod>(o.continuous_method);
    s.iteration_limit = o.iteration_limit;
    s.node_limit = o.node_limit;
    s.time_limit_seconds = o.time_limit_seconds;
    s.primal_tolerance = o.primal_tolerance;
    s.dual_tolerance = o.dual_tolerance;
    s.integrality_tolerance = o.integrality_tolerance;
    s.absolute_gap_tolerance = o.absolute_gap_tolerance;
    s.relative_gap_tolerance = o.relative_gap_tolerance;
    s.deterministic = o.deterministic != 0;
    s.pdhg_policy = static_cast<PdhgPolicy>(o.pdhg_policy);   // out-of-range values fail valid_options
    s.pdhg_precision = o.pdhg_precision == 1 ? PdhgPrecision::Mixed : PdhgPrecision::FP64;
    s.fused_iterations = o.fused_iterations != 0;
    s.use_graphs = o.use_graphs != 0;
    s.relative_kkt_tolerance = o.relative_kkt_tolerance;
    s.mixed_phase_tolerance = o.mixed_phase_tolerance;
    s.restart_period = o.restart_period;
    s.spmv_col_blocks = o.spmv_col_blocks;
    s.spmv_autotune = o.spmv_autotune != 0;
    s.spmv_split_passes = o.spmv_split_passes;
    s.spmv_cache_budget_mb = o.spmv_cache_budget_mb;
    s.spmv_panels = o.spmv_panels;
    s.spmv_panel_groups = o.spmv_panel_groups;
    s.spmv_panel_groups_t = o.spmv_panel_groups_t;
    s.spmv_panel_threads = o.spmv_panel_threads;
    s.spmv_panel_kb = o.spmv_panel_kb;
    s.spmv_tile_geometry = o.spmv_tile_geometry;
    s.mixed_recentre = o.mixed_recentre != 0;
    s.pdhg_primal_weight_gain = o.pdhg_primal_weight_gain;
    s.pdhg_primal_weight_integral = o.pdhg_primal_weight_integral;
    return s;
}

void copy_text(char *dst, std::size_t cap, const std::string &src) {
    const std::size_t n = std::min(cap - 1, src.size());
    std::memcpy(dst, src.data(), n);
    dst[n] = '\0';
}

void fill_info(const SolveResult &r, vk_solve_info &info) {
    std::memset(&info, 0, sizeof(info));
    info.status = static_cast<int>(r.status);
    info.iterations = r.iterations;
    info.low_precision_iterations = r.low_precision_iterations;
    info.objective_value = r.objective_value;
    info.primal_bound = r.primal_bound;
    info.dual_bound = r.dual_bound;
    info.primal_residual = r.primal_residual;
    info.stationarity_residual = r.stationarity_residual;
    info.complementarity_residual = r.complementarity_residual;
    info.absolute_gap = r.absolute_gap;
    info.relative_gap = r.relative_gap;
    info.solve_seconds = r.solve_seconds;
    info.device_seconds = r.device_seconds;
    info.kkt_relative_primal = r.kkt_relative_primal;
    info.kkt_relative_dual = r.kkt_relative_dual;
    info.kkt_relative_gap = r.kkt_relative_gap;
    info.kkt_relative = r.kkt_relative;
    info.kkt_primal_objective = r.kkt_primal_objective;
    info.kkt_dual_objective = r.kkt_dual_objective;
    copy_text(info.backend, sizeof(info.backend), r.backend);
    copy_text(info.solver, sizeof(info.solver), r.solver);
    copy_text(info.message, sizeof(info.message), r.message);
}

}  // namespace

extern "C" {

void vk_solver_options_default(vk_solver_options *o) {
    if (!o) return;
    const SolverOptions d;
    std::memset(o, 0, sizeof(*o));
    o->backend = static_cast<int>(d.backend);
    o->continuous_method = static_cast<int>(d.continuous_method);
    o->iteration_limit = d.iteration_limit;
    o->node_limit = d.node_limit;
    o->time_limit_seconds = d.time_limit_seconds;
    o->primal_tolerance = d.primal_tolerance;
    o->dual_tolerance = d.dual_tolerance;
    o->integrality_tolerance = d.integrality_tolerance;
    o->absolute_gap_tolerance = d.absolute_gap_tolerance;
    o->relative_gap_tolerance = d.relative_gap_tolerance;
    o->deterministic = d.deterministic ? 1 : 0;
    o->pdhg_policy = 0;
    o->pdhg_precision = 0;
    o->fused_iterations = d.fused_iterations ? 1 : 0;
    o->use_graphs = d.use_graphs ? 1 : 0;
    o->relative_kkt_tolerance = d.relative_kkt_tolerance;
    o->mixed_phase_tolerance = d.mixed_phase_tolerance;
    o->restart_period = d.restart_period;
    o->spmv_col_blocks = d.spmv_col_blocks;
    o->spmv_autotune = d.spmv_autotune ? 1 : 0;
    o->spmv_split_passes = d.spmv_split_passes;
    o->spmv_cache_budget_mb = d.spmv_cache_budget_mb;
    o->spmv_panels = d.spmv_panels;
    o->spmv_panel_groups = d.spmv_panel_groups;
    o->spmv_panel_groups_t = d.spmv_panel_groups_t;
    o->spmv_panel_threads = d.spmv_panel_threads;
    o->spmv_panel_kb = d.spmv_panel_kb;
    o->spmv_tile_geometry = d.spmv_tile_geometry;
    o->mixed_recentre = d.mixed_recentre ? 1 : 0;
    o->pdhg_primal_weight_gain = d.pdhg_primal_weight_gain;
    o->pdhg_primal_weight_integral = d.pdhg_primal_weight_integral;
}

static bool build_model(int64_t m, int64_t n, const int64_t *row_offsets,
                        const int64_t *column_indices, const double *values, const double *objective,
                        double objective_offset, const double *row_lower, const double *row_upper,
                        const double *var_lower, const double *var_upper, Model &model) {
    if (m < 0 || n < 0 || !row_offsets) return false;
    if ((n > 0 && (!objective || !var_lower || !var_upper)) || (m > 0 && (!row_lower || !row_upper)))
        return false;
    const int64_t nnz = row_offsets[m];
    if (nnz < 0 || (nnz > 0 && (!column_indices || !values))) return false;
    model.name = "c-abi";
    model.linear_objective.assign(objective, objective + n);
    model.objective_offset = objective_offset;
    model.variable_lower.assign(var_lower, var_lower + n);
    model.variable_upper.assign(var_upper, var_upper + n);
    model.variable_types.assign(static_cast<std::size_t>(n), VariableType::Continuous);
    model.constraint_lower.assign(row_lower, row_lower + m);
    model.constraint_upper.assign(row_upper, row_upper + m);
    model.constraint_matrix.pattern.rows = m;
    model.constraint_matrix.pattern.columns = n;
    model.constraint_matrix.pattern.row_offsets.assign(row_offsets, row_offsets + m + 1);
    model.constraint_matrix.pattern.column_indices.assign(column_indices, column_indices + nnz);
    model.constraint_matrix.values.assign(values, values + nnz);
    return true;
}

int vk_lp_profile(int64_t m, int64_t n, const int64_t *row_offsets, const int64_t *column_indices,
                  const double *values, const double *objective, const double *row_lower,
                  const double *row_upper, const double *var_lower, const double *var_upper,
                  const vk_solver_options *options, int64_t warm_iterations, int64_t iterations,
                  double *out) {
    if (!options || !out) return -1;
    Model model;
    if (!build_model(m, n, row_offsets, column_indices, values, objective, 0.0, row_lower, row_upper,
                     var_lower, var_upper, model))
        return -1;
    try {
        const auto p = solver::profile_lp_kernels(model, to_options(*options), warm_iterations, iterations);
        out[0] = p.dual_ms;
        out[1] = p.primal_ms;
        out[2] = p.iteration_ms;
        out[3] = p.step_size;
        out[4] = p.col_blocks;
        out[5] = p.split_passes;
        out[6] = p.dual_spmv_ms;
        out[7] = p.dual_pass_ms;
        out[8] = p.primal_spmv_ms;
        out[9] = p.primal_pass_ms;
        for (int i = 0; i < 10; ++i) out[10 + i] = p.panel[i];
    } catch (const std::exception &e) {
        return -2;
    }
    return 0;
}

int vk_lp_solve_batch(int count, const int64_t *row_starts, const int64_t *col_starts,
                      const int64_t *row_offsets, const int64_t *column_indices, const double *values,
                      const double *objective, const double *row_lower, const double *row_upper,
                      const double *var_lower, const double *var_upper, const vk_solver_options *options,
                      double *x_out, double *y_out, vk_solve_info *info) {
    if (count <= 0 || !row_starts || !col_starts || !row_offsets || !options || !info) return -1;
    if (row_starts[0] != 0 || col_starts[0] != 0) return -1;
    for (int i = 0; i < count; ++i)
        if (row_starts[i + 1] < row_starts[i] || col_starts[i + 1] < col_starts[i]) return -1;
    std::vector<Model> models(static_cast<size_t>(count));
    for (int i = 0; i < count; ++i) {
        const int64_t r0 = row_starts[i], m = row_starts[i + 1] - r0;
        const int64_t c0 = col_starts[i], n = col_starts[i + 1] - c0;
        Model &mod = models[i];
        mod.name = "batch-" + std::to_string(i);
        mod.linear_objective.assign(objective + c0, objective + c0 + n);
        mod.variable_lower.assign(var_lower + c0, var_lower + c0 + n);
        mod.variable_upper.assign(var_upper + c0, var_upper + c0 + n);
        mod.variable_types.assign(static_cast<size_t>(n), VariableType::Continuous);
        mod.constraint_lower.assign(row_lower + r0, row_lower + r0 + m);
        mod.constraint_upper.assign(row_upper + r0, row_upper + r0 + m);
        auto &pat = mod.constraint_matrix.pattern;
        pat.rows = m;
        pat.columns = n;
        pat.row_offsets.assign(1, 0);
        for (int64_t r = r0; r < r0 + m; ++r) {
            for (int64_t k = row_offsets[r]; k < row_offsets[r + 1]; ++k) {
                pat.column_indices.push_back(column_indices[k] - c0);   // out-of-range -> InvalidModel
                mod.constraint_matrix.values.push_back(values[k]);
            }
            pat.row_offsets.push_back(static_cast<Index>(pat.column_indices.size()));
        }
    }
    std::vector<const Model *> ptrs;
    for (const auto &mod : models) ptrs.push_back(&mod);
    std::vector<SolveResult> res;
    try {
        res = solver::solve_lp_batch(ptrs, to_options(*options));
    } catch (const std::exception &e) {
        res.assign(static_cast<size_t>(count), SolveResult{});
        for (auto &r : res) {
            r.status = SolveStatus::NumericalFailure;
            r.message = e.what();
        }
    }
    for (int i = 0; i < count; ++i) {
        fill_info(res[i], info[i]);
        const auto &r = res[i];
        const int64_t r0 = row_starts[i], c0 = col_starts[i];
        if (x_out && r.primal_solution.size() == static_cast<size_t>(col_starts[i + 1] - c0))
            std::memcpy(x_out + c0, r.primal_solution.data(), sizeof(double) * r.primal_solution.size());
        if (y_out && r.constraint_upper_dual.size() == static_cast<size_t>(row_starts[i + 1] - r0))
            for (size_t k = 0; k < r.constraint_upper_dual.size(); ++k)
                y_out[r0 + k] = r.constraint_upper_dual[k] - r.constraint_lower_dual[k];
    }
    return 0;
}

int vk_lp_solve_nodes(int64_t m, int64_t n, const int64_t *row_offsets, const int64_t *column_indices,
                      const double *values, const double *objective, double objective_offset,
                      const double *row_lower, const double *row_upper, const double *var_lower,
                      const double *var_upper, int count, const double *node_lower, const double *node_upper,
                      const double *warm_start, const vk_solver_options *options, double *x_out, double *y_out,
                      vk_solve_info *info) {
    if (!info || !options || count <= 0) return -1;
    Model model;
    if (!build_model(m, n, row_offsets, column_indices, values, objective, objective_offset, row_lower,
                     row_upper, var_lower, var_upper, model))
        return -1;
    const auto per_node = [&](const double *p) {
        std::vector<std::vector<double>> v;
        if (p)
            for (int i = 0; i < count; ++i) v.emplace_back(p + static_cast<int64_t>(i) * n, p + static_cast<int64_t>(i + 1) * n);
        return v;
    };
    auto lower = per_node(node_lower), upper = per_node(node_upper), warm = per_node(warm_start);
    if (lower.empty() && upper.empty()) lower.assign(static_cast<std::size_t>(count), std::vector<double>{});
    std::vector<SolveResult> res;
    try {
        res = solver::solve_lp_nodes(model, lower, upper, warm, to_options(*options));
    } catch (const std::exception &e) {
        res.assign(static_cast<std::size_t>(count), SolveResult{});
        for (auto &r : res) {
            r.status = SolveStatus::NumericalFailure;
            r.message = e.what();
        }
    }
    for (int i = 0; i < count; ++i) {
        const auto &r = res[static_cast<std::size_t>(i)];
        fill_info(r, info[i]);
        if (x_out && r.primal_solution.size() == static_cast<std::size_t>(n))
            std::memcpy(x_out + static_cast<int64_t>(i) * n, r.primal_solution.data(), sizeof(double) * n);
        if (y_out && r.constraint_upper_dual.size() == static_cast<std::size_t>(m))
            for (int64_t k = 0; k < m; ++k)
                y_out[static_cast<int64_t>(i) * m + k] = r.constraint_upper_dual[static_cast<std::size_t>(k)] -
                                                         r.constraint_lower_dual[static_cast<std::size_t>(k)];
    }
    return 0;
}

int vk_mip_solve_batched(int64_t m, int64_t n, const int64_t *row_offsets, const int64_t *column_indices,
                         const double *values, const double *objective, double objective_offset,
                         const double *row_lower, const double *row_upper, const double *var_lower,
                         const double *var_upper, const int32_t *variable_types, const int64_t *q_row_offsets,
                         const int64_t *q_column_indices, const double *q_values, int64_t batch,
                         const vk_solver_options *options, double *x_out, vk_solve_info *info, int64_t *stats) {
    if (!info || !options || !variable_types) return -1;
    Model model;
    if (!build_model(m, n, row_offsets, column_indices, values, objective, objective_offset, row_lower,
                     row_upper, var_lower, var_upper, model))
        return -1;
    if (q_row_offsets) {
        const int64_t qnnz = q_row_offsets[n];
        if (qnnz < 0 || (qnnz > 0 && (!q_column_indices || !q_values))) return -1;
        model.quadratic_objective.pattern.rows = n;
        model.quadratic_objective.pattern.columns = n;
        model.quadratic_objective.pattern.row_offsets.assign(q_row_offsets, q_row_offsets + n + 1);
        model.quadratic_objective.pattern.column_indices.assign(q_column_indices, q_column_indices + qnnz);
        model.quadratic_objective.values.assign(q_values, q_values + qnnz);
    }
    for (int64_t j = 0; j < n; ++j) {
        if (variable_types[j] < 0 || variable_types[j] > 2) return -1;
        model.variable_types[static_cast<std::size_t>(j)] = static_cast<VariableType>(variable_types[j]);
    }
    solver::MipStats st;
    SolveResult r;
    try {
        r = solver::solve_mip_batched(model, to_options(*options), batch, &st);
    } catch (const std::exception &e) {
        r = SolveResult{};
        r.status = SolveStatus::NumericalFailure;
        r.message = e.what();
    }
    fill_info(r, *info);
    if (x_out && r.primal_solution.size() == static_cast<std::size_t>(n))
        std::memcpy(x_out, r.primal_solution.data(), sizeof(double) * static_cast<std::size_t>(n));
    if (stats) {
        const int64_t v[9] = {st.nodes_processed, st.nodes_pruned, st.nodes_integral, st.nodes_infeasible,
                              st.incumbent_updates, st.rounds, st.relaxations, st.relaxation_iterations,
                              st.max_batch};
        std::memcpy(stats, v, sizeof(v));
    }
    return 0;
}

int vk_lp_solve(int mode, int64_t m, int64_t n, const int64_t *row_offsets,
                const int64_t *column_indices, const double *values, const double *objective,
                double objective_offset, const double *row_lower, const double *row_upper,
                const double *var_lower, const double *var_upper, const double *warm_start,
                const double *lower_override, const double *upper_override,
                const vk_solver_options *options, double *x_out, double *y_out, vk_solve_info *info) {
    if (!info || !options) return -1;
    Model model;
    if (!build_model(m, n, row_offsets, column_indices, values, objective, objective_offset, row_lower,
                     row_upper, var_lower, var_upper, model))
        return -1;
    const SolverOptions opts = to_options(*options);
    SolveResult r;
    if (mode == 0) {
        r = solve(model, opts);
    } else {
        try {
            auto backend = make_hip_backend();
            const auto span_or_empty = [n](const double *p) {
                return p ? std::span<const double>(p, static_cast<std::size_t>(n)) : std::span<const double>{};
            };
            r = solver::solve_lp(model, *backend, opts, span_or_empty(warm_start),
                                 span_or_empty(lower_override), span_or_empty(upper_override));
            r.backend = "hip";
        } catch (const std::exception &e) {
            r = SolveResult{};
            r.status = SolveStatus::NumericalFailure;
            r.message = e.what();
        }
    }
    fill_info(r, *info);
    if (x_out && r.primal_solution.size() == static_cast<std::size_t>(n))
        std::memcpy(x_out, r.primal_solution.data(), sizeof(double) * static_cast<std::size_t>(n));
    if (y_out && r.constraint_upper_dual.size() == static_cast<std::size_t>(m))
        for (int64_t i = 0; i < m; ++i)
            y_out[i] = r.constraint_upper_dual[static_cast<std::size_t>(i)] -
                       r.constraint_lower_dual[static_cast<std::size_t>(i)];
    return 0;
}

int vk_qp_solve(int mode, int64_t m, int64_t n, const int64_t *row_offsets, const int64_t *column_indices,
                const double *values, const int64_t *q_row_offsets, const int64_t *q_column_indices,
                const double *q_values, const double *objective, double objective_offset, const double *row_lower,
                const double *row_upper, const double *var_lower, const double *var_upper,
                const double *warm_start, const double *lower_override, const double *upper_override,
                const vk_solver_options *options, double *x_out, double *y_out, vk_solve_info *info) {
    if (!info || !options) return -1;
    Model model;
    if (!build_model(m, n, row_offsets, column_indices, values, objective, objective_offset, row_lower,
                     row_upper, var_lower, var_upper, model))
        return -1;
    if (q_row_offsets) {
        const int64_t qnnz = q_row_offsets[n];
        if (qnnz < 0 || (qnnz > 0 && (!q_column_indices || !q_values))) return -1;
        model.quadratic_objective.pattern.rows = n;
        model.quadratic_objective.pattern.columns = n;
        model.quadratic_objective.pattern.row_offsets.assign(q_row_offsets, q_row_offsets + n + 1);
        model.quadratic_objective.pattern.column_indices.assign(q_column_indices, q_column_indices + qnnz);
        model.quadratic_objective.values.assign(q_values, q_values + qnnz);
    }
    const SolverOptions opts = to_options(*options);
    SolveResult r;
    if (mode == 0) {
        r = solve(model, opts);
    } else {
        try {
            auto backend = make_hip_backend();
            const auto span_or_empty = [n](const double *p) {
                return p ? std::span<const double>(p, static_cast<std::size_t>(n)) : std::span<const double>{};
            };
            r = solver::solve_qp(model, *backend, opts, span_or_empty(warm_start), span_or_empty(lower_override),
                                 span_or_empty(upper_override));
            r.backend = "hip";
        } catch (const std::exception &e) {
            r = SolveResult{};
            r.status = SolveStatus::NumericalFailure;
            r.message = e.what();
        }
    }
    fill_info(r, *info);
    if (x_out && r.primal_solution.size() == static_cast<std::size_t>(n))
        std::memcpy(x_out, r.primal_solution.data(), sizeof(double) * static_cast<std::size_t>(n));
    if (y_out && r.constraint_upper_dual.size() == static_cast<std::size_t>(m))
        for (int64_t i = 0; i < m; ++i)
            y_out[i] = r.constraint_upper_dual[static_cast<std::size_t>(i)] -
                       r.constraint_lower_dual[static_cast<std::size_t>(i)];
    return 0;
}

const char *vk_solve_last_error(void) { return g_last_error.c_str(); }

struct vk_unrolled {
    std::unique_ptr<solver::UnrolledPdhg> u;
    int64_t m = 0, n = 0;
};

int vk_unrolled_create(int64_t m, int64_t n, const int64_t *row_offsets, const int64_t *column_indices,
                       const double *values, const double *objective, const double *row_lower,
                       const double *row_upper, const double *var_lower, const double *var_upper,
                       int64_t iterations, int64_t checkpoint_every, vk_unrolled **out) {
    if (!out) return -1;
    *out = nullptr;
    Model model;
    if (!build_model(m, n, row_offsets, column_indices, values, objective, 0.0, row_lower, row_upper, var_lower,
                     var_upper, model)) {
        g_last_error = "vk_unrolled_create: unusable arguments";
        return -1;
    }
    try {
        auto h = std::make_unique<vk_unrolled>();
        h->u = std::make_unique<solver::UnrolledPdhg>(model, iterations, checkpoint_every);
        h->m = m;
        h->n = n;
        *out = h.release();
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return -2;
    }
    return 0;
}

int vk_unrolled_forward(vk_unrolled *h, double *x, double *y) {
    if (!h || !x || (h->m > 0 && !y)) return -1;
    try {
        h->u->forward(std::span<double>(x, static_cast<std::size_t>(h->n)),
                      std::span<double>(y, static_cast<std::size_t>(h->m)));
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return -2;
    }
    return 0;
}

int vk_unrolled_backward(vk_unrolled *h, const double *grad_x, const double *grad_y, double *grad_c,
                         double *grad_row_lower, double *grad_row_upper) {
    if (!h || !grad_c || (h->m > 0 && (!grad_row_lower || !grad_row_upper))) return -1;
    const auto n = static_cast<std::size_t>(h->n), m = static_cast<std::size_t>(h->m);
    try {
        h->u->backward(grad_x ? std::span<const double>(grad_x, n) : std::span<const double>{},
                       grad_y ? std::span<const double>(grad_y, m) : std::span<const double>{},
                       std::span<double>(grad_c, n), std::span<double>(grad_row_lower, m),
                       std::span<double>(grad_row_upper, m));
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return -2;
    }
    return 0;
}

int vk_unrolled_set_data(vk_unrolled *h, const double *objective, const double *row_lower,
                         const double *row_upper) {
    if (!h || !objective || (h->m > 0 && (!row_lower || !row_upper))) return -1;
    const auto n = static_cast<std::size_t>(h->n), m = static_cast<std::size_t>(h->m);
    try {
        h->u->set_data(std::span<const double>(objective, n), std::span<const double>(row_lower, m),
                       std::span<const double>(row_upper, m));
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return -2;
    }
    return 0;
}

int vk_unrolled_info(const vk_unrolled *h, double *step_size, double *last_device_ms) {
    if (!h) return -1;
    if (step_size) *step_size = h->u->step_size();
    if (last_device_ms) *last_device_ms = h->u->last_device_ms();
    return 0;
}

int vk_unrolled_destroy(vk_unrolled *h) {
    delete h;
    return 0;
}

struct vk_model {
    Model model;
};

int vk_mps_read(const char *path, vk_model **out) {
    if (!path || !out) {
        g_last_error = "vk_mps_read: null argument";
        return -1;
    }
    *out = nullptr;
    try {
        auto h = std::make_unique<vk_model>();
        MpsReader reader;
        if (!reader.read(path, h->model)) {
            std::string msg;
            for (const auto &e : reader.errors()) msg += "line " + std::to_string(e.line) + ": " + e.message + "\n";
            if (reader.errors().empty()) {
                const auto problems = h->model.validate();
                msg = "line 0: " + (problems.empty() ? std::string("invalid model") : problems.front()) + "\n";
            }
            g_last_error = msg;
            return 1;
        }
        *out = h.release();
        return 0;
    } catch (const std::exception &e) {
        g_last_error = std::string("line 0: ") + e.what() + "\n";
        return 1;
    }
}

int vk_model_dims(const vk_model *h, int64_t *m, int64_t *n, int64_t *nnz) {
    if (!h) return -1;
    if (m) *m = h->model.num_constraints();
    if (n) *n = h->model.num_variables();
    if (nnz) *nnz = static_cast<int64_t>(h->model.constraint_matrix.values.size());
    return 0;
}

int vk_model_export(const vk_model *h, int64_t *row_offsets, int64_t *column_indices, double *values,
                    double *objective, double *row_lower, double *row_upper, double *var_lower,
                    double *var_upper, int32_t *variable_types, double *objective_offset) {
    if (!h) return -1;
    const Model &md = h->model;
    auto put = [](auto *dst, const auto &src) {
        if (dst && !src.empty()) std::memcpy(dst, src.data(), src.size() * sizeof(src[0]));
    };
    put(row_offsets, md.constraint_matrix.pattern.row_offsets);
    put(column_indices, md.constraint_matrix.pattern.column_indices);
    put(values, md.constraint_matrix.values);
    put(objective, md.linear_objective);
    put(row_lower, md.constraint_lower);
    put(row_upper, md.constraint_upper);
    put(var_lower, md.variable_lower);
    put(var_upper, md.variable_upper);
    if (variable_types)
        for (std::size_t j = 0; j < md.variable_types.size(); ++j)
            variable_types[j] = static_cast<int32_t>(md.variable_types[j]);
    if (objective_offset) *objective_offset = md.objective_offset;
    return 0;
}

int64_t vk_model_name(const vk_model *h, char *buf, int64_t size) {
    if (!h) return -1;
    if (buf && size > 0) copy_text(buf, static_cast<std::size_t>(size), h->model.name);
    return static_cast<int64_t>(h->model.name.size());
}

void vk_model_free(vk_model *h) { delete h; }

}  // extern "C"
