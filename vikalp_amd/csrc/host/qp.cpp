// solver::solve_qp on the HIP backend: PDQP, the PDHG loop with the gradient
// Q x + c + A'y (src/solver/qp/pdqp.cpp).  Entry checks follow pdqp.cpp:303-331
// (bounds, model validation, class QP, a variable, a PSD Q, options, method),
// the scaling continuous_common.hpp:133-205 with Q scaled by s_i s_j
// (pdqp.cpp:36-45), the step 0.99 / max(1, bound^2 + |Q|_inf) (pdqp.cpp:87-99)
// and the postsolve pdqp.cpp:178-300.  Two executions: the fused device
// iteration (vk_hip_pdqp_create: one more SpMV per iteration, Q x+, ahead of
// the primal kernel) and the reference's op-by-op sequence on any
// ExecutionBackend.  Beyond the reference, SolverOptions::pdhg_policy =
// Halpern runs restarted Halpern PDHG on the QP (fused): the LP path's
// equilibrated device scaling, per-column primal steps 1 / (w / eta +
// sum_k |q_jk|) that keep the PDHG metric positive definite, optionally the
// relative-KKT test with the 1/2 x'Qx terms, in FP64 or mixed (the recentred
// FP32 phase with the base's Q x, then a certifying FP64 phase).

#include "lp_internal.hpp"

#include "vikalp/backend/HipBackend.hpp"
#include "vikalp/solver/continuous.hpp"
#include "vikalp_hip.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

namespace vikalp::solver {
namespace detail {

// Largest n the dense check runs for; beyond it a diagonally dominant Q with
// a non-negative diagonal is accepted (Gershgorin: every eigenvalue lies in
// some [q_ii - r_i, q_ii + r_i] with r_i the off-diagonal row sum).
constexpr Index kDensePsdLimit = 4096;

bool positive_semidefinite(const CsrMatrix &q, std::string &message) {
    const Index n = q.pattern.rows;
    if (n > kDensePsdLimit) {
        for (Index i = 0; i < n; ++i) {
            Scalar diag = 0.0, off = 0.0;
            for (Index k = q.pattern.row_offsets[i]; k < q.pattern.row_offsets[i + 1]; ++k) {
                if (q.pattern.column_indices[k] == i) diag = q.values[k];
                else off += std::abs(q.values[k]);
            }
            if (diag < off) {
                message = "QP objective matrix is not diagonally dominant and too large for the dense "
                          "positive-semidefinite check";
                return false;
            }
        }
        return true;
    }
    // continuous_common.hpp:242-275: dense Cholesky that lets a zero pivot
    // through only when its column below is zero as well.
    const auto nn = static_cast<std::size_t>(n);
    std::vector<Scalar> a(nn * nn, 0.0);
    for (Index r = 0; r < n; ++r)
        for (Index k = q.pattern.row_offsets[r]; k < q.pattern.row_offsets[r + 1]; ++k)
            a[static_cast<std::size_t>(r) * nn + static_cast<std::size_t>(q.pattern.column_indices[k])] = q.values[k];
    const Scalar tol = 1e-10;
    for (std::size_t k = 0; k < nn; ++k) {
        Scalar pivot = a[k * nn + k];
        for (std::size_t j = 0; j < k; ++j) pivot -= a[k * nn + j] * a[k * nn + j];
        if (pivot < -tol * std::max<Scalar>(1.0, std::abs(a[k * nn + k]))) {
            message = "QP objective matrix is not positive semidefinite";
            return false;
        }
        if (pivot <= tol) {
            for (std::size_t i = k + 1; i < nn; ++i) {
                if (std::abs(a[i * nn + k]) > tol) {
                    message = "QP objective matrix is not positive semidefinite";
                    return false;
                }
                a[i * nn + k] = 0.0;
            }
            continue;
        }
        const Scalar d = std::sqrt(pivot);
        a[k * nn + k] = d;
        for (std::size_t i = k + 1; i < nn; ++i) {
            Scalar v = a[i * nn + k];
            for (std::size_t j = 0; j < k; ++j) v -= a[i * nn + j] * a[k * nn + j];
            a[i * nn + k] = v / d;
        }
    }
    return true;
}

}  // namespace detail

namespace {

using namespace detail;

constexpr Scalar kInf = std::numeric_limits<Scalar>::infinity();

struct ScaledQ {
    std::vector<Scalar> values;   // q_ij s_i s_j
    Scalar norm_inf = 0.0;        // max row sum of |q_ij s_i s_j|
};

ScaledQ scale_q(const CsrMatrix &q, const std::vector<Scalar> &s) {
    ScaledQ out;
    out.values = q.values;
    for (Index r = 0; r < q.pattern.rows; ++r) {
        Scalar row_sum = 0.0;
        for (Index k = q.pattern.row_offsets[r]; k < q.pattern.row_offsets[r + 1]; ++k) {
            out.values[k] *= s[r] * s[q.pattern.column_indices[k]];   // pdqp.cpp:42-44
            row_sum += std::abs(out.values[k]);
        }
        out.norm_inf = std::max(out.norm_inf, row_sum);   // pdqp.cpp:90-97
    }
    return out;
}

// pdqp.cpp:178-300 on original-coordinate x, y (reference sign), and the
// relative KKT error of the QP.  relative_mode: the device stopped on the
// relative-KKT test, itself in original coordinates, so the absolute
// re-check of pdqp.cpp does not apply.
void postsolve_qp(const Model &model, const Bounds &bounds, const SolverOptions &options,
                  const std::vector<Scalar> &x, const std::vector<Scalar> &y, bool relative_mode,
                  SolveResult &res) {
    const auto &A = model.constraint_matrix;
    const auto &Q = model.quadratic_objective;
    const auto m = static_cast<std::size_t>(A.pattern.rows);
    const auto n = static_cast<std::size_t>(A.pattern.columns);
    res.primal_solution = x;
    // continuous_common.hpp:284-299 (row-major over the stored entries; the
    // dense loop adds only zero terms besides these).
    Scalar value = model.objective_offset;
    Scalar lin = 0.0;
    for (std::size_t j = 0; j < n; ++j) lin += model.linear_objective[j] * x[j];
    value += lin;
    Scalar quad = 0.0;
    for (std::size_t r = 0; r < n; ++r)
        for (Index k = Q.pattern.row_offsets[r]; k < Q.pattern.row_offsets[r + 1]; ++k)
            quad += x[r] * Q.values[k] * x[static_cast<std::size_t>(Q.pattern.column_indices[k])];
    value += 0.5 * quad;
    res.objective_value = value;
    if (!std::isfinite(res.objective_value)) {
        res.status = SolveStatus::NumericalFailure;
        res.message = "PDQP produced a non-finite objective";
    }
    std::vector<Scalar> row_values(m, 0.0);
    for (std::size_t i = 0; i < m; ++i)
        for (Index k = A.pattern.row_offsets[i]; k < A.pattern.row_offsets[i + 1]; ++k)
            row_values[i] += A.values[k] * x[static_cast<std::size_t>(A.pattern.column_indices[k])];
    Scalar viol = 0.0, box = 0.0;
    for (std::size_t i = 0; i < m; ++i) {
        if (std::isfinite(model.constraint_lower[i])) viol = std::max(viol, model.constraint_lower[i] - row_values[i]);
        if (std::isfinite(model.constraint_upper[i])) viol = std::max(viol, row_values[i] - model.constraint_upper[i]);
    }
    for (std::size_t j = 0; j < n; ++j) {
        if (std::isfinite(bounds.lower[j])) box = std::max(box, bounds.lower[j] - x[j]);
        if (std::isfinite(bounds.upper[j])) box = std::max(box, x[j] - bounds.upper[j]);
    }
    res.primal_residual = std::max(std::max<Scalar>(0.0, viol), std::max<Scalar>(0.0, box));
    res.constraint_lower_dual.resize(m);
    res.constraint_upper_dual.resize(m);
    for (std::size_t i = 0; i < m; ++i) {
        res.constraint_lower_dual[i] = std::max<Scalar>(0.0, -y[i]);
        res.constraint_upper_dual[i] = std::max<Scalar>(0.0, y[i]);
    }
    // pdqp.cpp:214-232: g = c + Q x + A'y.
    std::vector<Scalar> g(model.linear_objective);
    for (std::size_t r = 0; r < n; ++r)
        for (Index k = Q.pattern.row_offsets[r]; k < Q.pattern.row_offsets[r + 1]; ++k)
            g[r] += Q.values[k] * x[static_cast<std::size_t>(Q.pattern.column_indices[k])];
    for (std::size_t i = 0; i < m; ++i)
        for (Index k = A.pattern.row_offsets[i]; k < A.pattern.row_offsets[i + 1]; ++k)
            g[static_cast<std::size_t>(A.pattern.column_indices[k])] += y[i] * A.values[k];
    res.stationarity_residual = 0.0;
    res.complementarity_residual = 0.0;
    res.variable_lower_dual.assign(n, 0.0);
    res.variable_upper_dual.assign(n, 0.0);
    const Scalar active = std::max<Scalar>(10.0 * options.primal_tolerance, 1e-8);
    for (std::size_t j = 0; j < n; ++j) {
        const Scalar lo = bounds.lower[j], hi = bounds.upper[j];
        const Scalar proj = std::clamp(x[j] - g[j], lo, hi);
        res.stationarity_residual = std::max(res.stationarity_residual, std::abs(x[j] - proj));
        if (std::isfinite(lo) && x[j] - lo <= active) {
            res.variable_lower_dual[j] = std::max<Scalar>(0.0, g[j]);
            res.complementarity_residual =
                std::max(res.complementarity_residual, std::abs((x[j] - lo) * res.variable_lower_dual[j]));
        }
        if (std::isfinite(hi) && hi - x[j] <= active) {
            res.variable_upper_dual[j] = std::max<Scalar>(0.0, -g[j]);
            res.complementarity_residual =
                std::max(res.complementarity_residual, std::abs((hi - x[j]) * res.variable_upper_dual[j]));
        }
    }
    for (std::size_t i = 0; i < m; ++i) {
        const Scalar v = row_values[i], lo = model.constraint_lower[i], hi = model.constraint_upper[i];
        if (std::isfinite(lo))
            res.complementarity_residual =
                std::max(res.complementarity_residual, std::abs((v - lo) * res.constraint_lower_dual[i]));
        if (std::isfinite(hi))
            res.complementarity_residual =
                std::max(res.complementarity_residual, std::abs((hi - v) * res.constraint_upper_dual[i]));
    }
    if (res.status == SolveStatus::NotSolved) res.status = SolveStatus::IterationLimit;
    bool finite = true;
    for (const Scalar v : x) finite = finite && std::isfinite(v);
    if (!finite) {
        res.status = SolveStatus::NumericalFailure;
        res.message = "PDQP produced a non-finite iterate";
    } else if (res.status == SolveStatus::Optimal && !relative_mode &&
               (res.primal_residual > options.primal_tolerance ||
                res.stationarity_residual > options.dual_tolerance)) {
        res.status = SolveStatus::NumericalFailure;
        res.message = "PDQP stopped without satisfying original-coordinate residual checks";
    }
    if (res.status == SolveStatus::Optimal) {
        res.primal_bound = res.objective_value;
        res.dual_bound = res.objective_value;
        res.absolute_gap = 0.0;
        res.relative_gap = 0.0;
    }
    Products pr;
    pr.ax = std::move(row_values);
    pr.g = std::move(g);
    relative_kkt(model, bounds, x, y, pr, res, quad);
}

// pdqp.cpp:110-176 issued call for call on an ExecutionBackend.
void iterate_ops(const Model &model, const Scaled &sc, ExecutionBackend &be, BackendMatrix &A,
                 BackendMatrix &Q, Scalar max_step, const SolverOptions &options,
                 const std::vector<Scalar> *warm_scaled, Clock::time_point started, std::vector<Scalar> &xs,
                 std::vector<Scalar> &ys, SolveResult &res) {
    const Index m = model.num_constraints(), n = model.num_variables();
    auto x = be.create_vector(n), x_next = be.create_vector(n), x_bar = be.create_vector(n);
    auto y = be.create_vector(m), aty = be.create_vector(n), ax_bar = be.create_vector(m);
    auto y_arg = be.create_vector(m), y_proj = be.create_vector(m), ax = be.create_vector(m);
    auto pax = be.create_vector(m), qx = be.create_vector(n), grad = be.create_vector(n);
    auto px = be.create_vector(n), resid = be.create_vector(n);
    auto c = be.create_vector(n), l = be.create_vector(n), u = be.create_vector(n);
    auto rl = be.create_vector(m), ru = be.create_vector(m);
    be.upload(sc.c, *c);
    be.upload(sc.l, *l);
    be.upload(sc.u, *u);
    be.upload(sc.rl, *rl);
    be.upload(sc.ru, *ru);
    if (warm_scaled) be.upload(*warm_scaled, *x);
    else be.fill(*x, 0.0);
    be.project_box(*x, *l, *u);
    be.fill(*y, 0.0);
    Scalar tau = max_step, sigma = max_step, prev = kInf;
    bool restart = false;
    for (Index iter = 0; iter < options.iteration_limit; ++iter) {
        if (timed_out(started, options)) {
            res.status = SolveStatus::TimeLimit;
            break;
        }
        be.multiply(1.0, A, *y, 0.0, *aty, Transpose::Yes);
        be.multiply(1.0, Q, *x, 0.0, *qx, Transpose::No);
        be.copy(*qx, *grad);
        be.axpby(1.0, *c, 1.0, *grad);
        be.axpby(1.0, *aty, 1.0, *grad);
        be.copy(*x, *x_next);
        be.axpby(-tau, *grad, 1.0, *x_next);
        be.project_box(*x_next, *l, *u);
        if (restart) {
            be.copy(*x_next, *x_bar);
            restart = false;
        } else {
            be.copy(*x, *x_bar);
            be.axpby(2.0, *x_next, -1.0, *x_bar);
        }
        be.multiply(1.0, A, *x_bar, 0.0, *ax_bar, Transpose::No);
        be.copy(*y, *y_arg);
        be.axpby(sigma, *ax_bar, 1.0, *y_arg);
        be.copy(*y_arg, *y_proj);
        be.axpby(1.0 / sigma, *y_proj, 0.0, *y_proj);
        be.project_box(*y_proj, *rl, *ru);
        be.copy(*y_arg, *y);
        be.axpby(-sigma, *y_proj, 1.0, *y);
        be.copy(*x_next, *x);
        res.iterations = iter + 1;
        be.multiply(1.0, A, *x, 0.0, *ax, Transpose::No);
        be.copy(*ax, *pax);
        be.project_box(*pax, *rl, *ru);
        be.axpby(1.0, *pax, -1.0, *ax);
        const Scalar primal = be.norm_inf(*ax);
        be.multiply(1.0, A, *y, 0.0, *aty, Transpose::Yes);
        be.multiply(1.0, Q, *x, 0.0, *qx, Transpose::No);
        be.copy(*qx, *grad);
        be.axpby(1.0, *c, 1.0, *grad);
        be.axpby(1.0, *aty, 1.0, *grad);
        be.copy(*x, *px);
        be.axpby(-1.0, *grad, 1.0, *px);
        be.project_box(*px, *l, *u);
        be.copy(*x, *resid);
        be.axpby(-1.0, *px, 1.0, *resid);
        const Scalar metric = std::max(primal, be.norm_inf(*resid));
        if (iter > 0 && iter % 50 == 0) {   // pdqp.cpp:161-171
            if (metric > prev * 1.25) {
                tau *= 0.5;
                sigma *= 0.5;
                restart = true;
            } else if (metric < prev * 0.1) {
                tau = std::min(tau * 1.05, max_step);
                sigma = tau;
            }
            prev = metric;
        } else if (iter == 0) {
            prev = metric;
        }
        if (metric <= std::max(options.primal_tolerance, options.dual_tolerance)) {
            res.status = SolveStatus::Optimal;
            break;
        }
    }
    be.synchronize();
    xs.resize(static_cast<std::size_t>(n));
    ys.resize(static_cast<std::size_t>(m));
    be.download(*x, xs);
    be.download(*y, ys);
}

// One device PDQP solve with `p` from the scaled start (x0, y0; null = 0):
// the VK_PDHG_* status, the scaled iterates, the final state and the device
// time (run_device's loop).
int run_qp(const vk_hip_csr *A, const vk_hip_csr *Q, const Scaled &sc, const vk_pdhg_params &p,
           const std::vector<Scalar> *x0, const std::vector<Scalar> *y0, Clock::time_point started,
           const SolverOptions &options, std::vector<Scalar> &xs, std::vector<Scalar> &ys, vk_pdhg_state &state,
           double &device_ms, bool &timed) {
    DeviceSolver solver;
    check_vk(vk_hip_pdqp_create(A, Q, sc.c.data(), sc.l.data(), sc.u.data(), sc.rl.data(), sc.ru.data(), sc.r.data(),
                                sc.s.data(), x0 ? x0->data() : nullptr, y0 ? y0->data() : nullptr, &p, &solver.h),
             "vk_hip_pdqp_create");
    return drive_device(solver, p, sc.c.size(), sc.r.size(), started, options, xs, ys, state, device_ms, timed);
}

// The fused device iteration with Q: the reference policy in FP64 with its
// step max_step, or (halpern) restarted Halpern PDHG with eta = 0.99 /
// ||A||_2 (the device sets each column's primal step from Q's absolute row
// sums), in FP64 or mixed: a recentred FP32 phase, then an FP64 phase
// warm-started from its iterates that certifies the point (solve_lp's mixed
// driver).  Returns the VK_PDHG_* status.
int iterate_fused(const Model &model, const Scaled &sc, const vk_hip_csr *A, const vk_hip_csr *Q,
                  Scalar max_step, bool halpern, const SolverOptions &options,
                  const std::vector<Scalar> *warm_scaled, Clock::time_point started, std::vector<Scalar> &xs,
                  std::vector<Scalar> &ys, SolveResult &res) {
    vk_pdhg_params p{};
    p.policy = VK_PDHG_POLICY_REFERENCE;
    p.precision = VK_PREC_FP64;
    p.iteration_limit = options.iteration_limit;
    p.metric_tolerance = std::max(options.primal_tolerance, options.dual_tolerance);
    p.step_size = max_step;
    p.primal_weight = 1.0;
    if (halpern) {
        p.policy = VK_PDHG_POLICY_HALPERN;
        p.relative_kkt_tolerance = options.relative_kkt_tolerance;
        double norm = 0.0;
        check_vk(vk_hip_csr_norm2_estimate(A, 100, &norm), "vk_hip_csr_norm2_estimate");
        p.step_size = 0.99 / std::max(norm, 1e-12);
        const Scalar nc = norm2(sc.c), nb = row_bound_norm(sc.rl, sc.ru);
        p.primal_weight = (nc > 1e-10 && nb > 1e-10) ? nc / nb : 1.0;
        p.primal_weight_gain = options.pdhg_primal_weight_gain;
        p.primal_weight_integral = options.pdhg_primal_weight_integral;
    }
    p.restart_period = options.restart_period;
    p.use_graphs = options.use_graphs ? 1 : 0;
    p.col_blocks = options.spmv_col_blocks;
    p.split_passes = options.spmv_split_passes;
    p.cache_budget_mb = options.spmv_cache_budget_mb;
    p.spmv_panels = options.spmv_panels;
    p.panel_groups = options.spmv_panel_groups;
    p.panel_groups_t = options.spmv_panel_groups_t;
    p.panel_threads = options.spmv_panel_threads;
    p.panel_kb = options.spmv_panel_kb;
    p.panel_tune = options.spmv_autotune ? 1 : 0;
    p.norm_b = row_bound_norm(model.constraint_lower, model.constraint_upper);
    p.norm_c = norm2(model.linear_objective);
    vk_pdhg_state state{};
    double ms = 0.0;
    bool timed = false;
    int status;
    if (halpern && options.pdhg_precision == PdhgPrecision::Mixed) {
        vk_pdhg_params p32 = p;
        p32.precision = VK_PREC_MIXED;
        p32.recentre = 1;
        p32.relative_kkt_tolerance = fp32_phase_tolerance(options, true);
        status = run_qp(A, Q, sc, p32, warm_scaled, nullptr, started, options, xs, ys, state, ms, timed);
        res.low_precision_iterations = state.iterations;
        Index used = state.iterations;
        double total_ms = ms;
        if (!timed && used < options.iteration_limit && status != VK_PDHG_PRIMAL_INFEASIBLE) {
            const auto x32 = xs, y32 = ys;
            vk_pdhg_params p64 = p;
            p64.iteration_limit = options.iteration_limit - used;
            p64.primal_weight = state.primal_weight;
            status = run_qp(A, Q, sc, p64, &x32, &y32, started, options, xs, ys, state, ms, timed);
            total_ms += ms;
            used += state.iterations;
        }
        state.iterations = used;
        ms = total_ms;
    } else {
        status = run_qp(A, Q, sc, p, warm_scaled, nullptr, started, options, xs, ys, state, ms, timed);
    }
    res.iterations = state.iterations;
    res.device_seconds = ms * 1e-3;
    if (timed) res.status = SolveStatus::TimeLimit;
    else if (status == VK_PDHG_CONVERGED) res.status = SolveStatus::Optimal;
    return status;
}

}  // namespace

SolveResult solve_qp(const Model &model, ExecutionBackend &backend, const SolverOptions &options,
                     std::span<const Scalar> warm_start, std::span<const Scalar> variable_lower_override,
                     std::span<const Scalar> variable_upper_override) {
    const auto started = Clock::now();
    Bounds bounds;
    std::string message;
    if (!resolve_bounds(model, variable_lower_override, variable_upper_override, bounds, message))
        return status_result(SolveStatus::InvalidModel, message);
    const auto errors = model.validate();
    if (!errors.empty()) return status_result(SolveStatus::InvalidModel, errors.front());
    if (model.problem_class() != ProblemClass::QP)
        return status_result(SolveStatus::UnsupportedModel, "model class does not match the selected continuous solver");
    if (model.num_variables() == 0) return status_result(SolveStatus::InvalidModel, "QP requires at least one variable");
    if (!positive_semidefinite(model.quadratic_objective, message))
        return status_result(SolveStatus::UnsupportedModel, message);
    if (!valid_options(options))
        return status_result(SolveStatus::InvalidModel, "solver limits and tolerances must be non-negative");
    if (options.continuous_method != ContinuousMethod::Auto && options.continuous_method != ContinuousMethod::PDHG)
        return status_result(SolveStatus::UnsupportedModel, "requested continuous method is not implemented for QP");
    const bool halpern = options.pdhg_policy == PdhgPolicy::Halpern;
    const bool mixed = options.pdhg_precision == PdhgPrecision::Mixed;
    if (options.pdhg_policy == PdhgPolicy::Adaptive || (!halpern && (mixed || options.relative_kkt_tolerance > 0.0)) ||
        (mixed && !options.mixed_recentre))
        return status_result(SolveStatus::UnsupportedModel,
                             "PDQP runs the reference policy (pdqp.cpp) in FP64, or the Halpern policy in FP64 or "
                             "mixed precision with the recentred FP32 phase; the adaptive policy is for LPs, and "
                             "the relative-KKT test needs the Halpern policy");
    if (!warm_start.empty()) {   // pdqp.cpp:29-34
        if (static_cast<Index>(warm_start.size()) != model.num_variables())
            return status_result(SolveStatus::InvalidModel, "warm start must be empty or finite and sized to the model");
        for (const Scalar v : warm_start)
            if (!std::isfinite(v))
                return status_result(SolveStatus::InvalidModel,
                                     "warm start must be empty or finite and sized to the model");
    }

    auto matrix = backend.create_matrix(model.constraint_matrix.pattern);
    auto quadratic = backend.create_matrix(model.quadratic_objective.pattern);
    vk_hip_csr *a_handle = options.fused_iterations ? hip_matrix_handle(*matrix) : nullptr;
    const vk_hip_csr *q_handle = options.fused_iterations ? hip_matrix_handle(*quadratic) : nullptr;
    if (halpern && !(a_handle && q_handle))
        return status_result(SolveStatus::UnsupportedModel, "Halpern PDQP runs only fused on a HipBackend");
    // Halpern: the LP path's device scaling with Ruiz equilibration; the
    // reference policy: pdqp.cpp's host scaling.
    Scaled sc;
    if (halpern) {
        backend.set_values(model.constraint_matrix.values, *matrix);
        sc = device_scale(a_handle, model, bounds, true);
    } else {
        sc = scale(model, bounds);
        backend.set_values(sc.values, *matrix);
    }
    std::vector<Scalar> warm(warm_start.begin(), warm_start.end());
    for (std::size_t j = 0; j < warm.size(); ++j) warm[j] /= sc.s[j];   // pdqp.cpp:77-82
    const std::vector<Scalar> *warm_ptr = warm_start.empty() ? nullptr : &warm;
    const ScaledQ sq = scale_q(model.quadratic_objective, sc.s);
    backend.set_values(sq.values, *quadratic);
    Scalar max_step = 0.0;
    if (!halpern) {
        const Scalar bound = operator_bound(model.constraint_matrix.pattern, sc.values);
        max_step = 0.99 / std::max<Scalar>(1.0, bound * bound + sq.norm_inf);   // pdqp.cpp:98
    }
    SolveResult res;
    res.solver = "PDQP";
    res.backend = "provided";
    std::vector<Scalar> xs, ys;
    std::string tiles_note;
    int status = VK_PDHG_RUNNING;
    if (a_handle && q_handle) {
        // The matrix belongs to this solve: its tiles may be re-cut.
        const int g = apply_tile_geometry(a_handle, options);
        tiles_note = tile_geometry_note(g, options.spmv_tile_geometry < 0);
        status = iterate_fused(model, sc, a_handle, q_handle, max_step, halpern, options, warm_ptr, started, xs, ys,
                               res);
    } else {
        iterate_ops(model, sc, backend, *matrix, *quadratic, max_step, options, warm_ptr, started, xs, ys, res);
    }
    std::vector<Scalar> x(xs.size()), y(ys.size());
    for (std::size_t j = 0; j < xs.size(); ++j) x[j] = xs[j] * sc.s[j];
    for (std::size_t i = 0; i < ys.size(); ++i) y[i] = ys[i] * sc.r[i];
    const bool relative_mode = halpern && options.relative_kkt_tolerance > 0.0;
    postsolve_qp(model, bounds, options, x, y, relative_mode, res);
    // A Farkas ray does not involve the objective: the LP check applies.
    if (status == VK_PDHG_PRIMAL_INFEASIBLE && res.status != SolveStatus::TimeLimit)
        report_infeasible(model, bounds, y, res);
    if (!tiles_note.empty()) res.message = res.message.empty() ? tiles_note : res.message + "; " + tiles_note;
    res.solve_seconds = std::chrono::duration<Scalar>(Clock::now() - started).count();
    return res;
}

}  // namespace vikalp::solver
