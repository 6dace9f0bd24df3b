// solve_lp: the PDHG LP path on MI355X.
//
// Policy and semantics follow src/solver/lp/pdhg.cpp (entry checks :269-293,
// setup :24-85, loop :87-151, postsolve :153-264) and the helpers of
// src/solver/continuous_common.hpp.  Two execution modes:
//   fused    (HipBackend + options.fused_iterations): the whole iteration runs
//            on the device through vk_hip_pdhg_* (two kernels per iteration,
//            no host synchronisation inside a block of iterations);
//   op-by-op (any ExecutionBackend): the reference's exact backend-call
//            sequence (used for the per-op drop-in test and for backends
//            without a fused path).
// Scaling is computed in one O(nnz) pass that yields the same factors as the
// reference's O(n*nnz) column loop (continuous_common.hpp:162-176): a maximum
// over a set does not depend on visiting order.

#include "lp_internal.hpp"

#include "vikalp/backend/HipBackend.hpp"
#include "vikalp/solver/continuous.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace vikalp::solver {
namespace detail {

constexpr Scalar kInf = std::numeric_limits<Scalar>::infinity();

SolveResult status_result(SolveStatus status, const std::string &message) {
    SolveResult r;
    r.status = status;
    r.message = message;
    return r;
}

bool valid_groups(int g) { return g == 0 || g == 1 || g == 2 || g == 4 || g == 8; }

// continuous_common.hpp:50-56
bool valid_options(const SolverOptions &o) {
    return o.iteration_limit >= 0 && std::isfinite(o.primal_tolerance) &&
           std::isfinite(o.dual_tolerance) && o.primal_tolerance >= 0.0 && o.dual_tolerance >= 0.0 &&
           (!std::isfinite(o.time_limit_seconds) || o.time_limit_seconds >= 0.0) &&
           std::isfinite(o.relative_kkt_tolerance) && o.relative_kkt_tolerance >= 0.0 &&
           o.restart_period > 0 && std::isfinite(o.mixed_phase_tolerance) &&
           o.mixed_phase_tolerance >= 0.0 && o.spmv_col_blocks >= 0 && o.spmv_col_blocks <= 16 &&
           o.spmv_split_passes >= 0 && o.spmv_split_passes <= 2 && o.spmv_cache_budget_mb >= -1 &&
           o.spmv_panels >= 0 && o.spmv_panels <= 2 && valid_groups(o.spmv_panel_groups) &&
           valid_groups(o.spmv_panel_groups_t) &&
           (o.spmv_panel_threads == 0 || o.spmv_panel_threads == 256 || o.spmv_panel_threads == 512) &&
           (o.spmv_panel_kb == 0 || o.spmv_panel_kb == 48 || o.spmv_panel_kb == 24 || o.spmv_panel_kb == 16) &&
           !(o.spmv_panel_threads == 256 && o.spmv_panel_kb != 0 && o.spmv_panel_kb != 48) &&
           o.spmv_tile_geometry >= -1 && o.spmv_tile_geometry < 4 &&
           o.pdhg_primal_weight_gain >= 0.0 && o.pdhg_primal_weight_gain <= 1.0 &&
           o.pdhg_primal_weight_integral >= 0.0 && o.pdhg_primal_weight_integral <= 1.0 &&
           (o.pdhg_policy == PdhgPolicy::Reference || o.pdhg_policy == PdhgPolicy::Adaptive ||
            o.pdhg_policy == PdhgPolicy::Halpern);
}

bool timed_out(Clock::time_point started, const SolverOptions &o) {
    return std::isfinite(o.time_limit_seconds) &&
           std::chrono::duration<Scalar>(Clock::now() - started).count() >= o.time_limit_seconds;
}

// continuous_common.hpp:80-110
bool resolve_bounds(const Model &model, std::span<const Scalar> lo, std::span<const Scalar> hi,
                    Bounds &b, std::string &message) {
    const auto n = static_cast<std::size_t>(model.num_variables());
    if (model.variable_lower.size() != n || model.variable_upper.size() != n) {
        message = "model variable bounds must match the objective size";
        return false;
    }
    auto usable = [n](std::span<const Scalar> v) {
        if (v.empty()) return true;
        if (v.size() != n) return false;
        for (const Scalar x : v)
            if (std::isnan(x)) return false;
        return true;
    };
    if (!usable(lo) || !usable(hi)) {
        message = "variable bound overrides must be empty or sized and non-NaN";
        return false;
    }
    b.lower = lo.empty() ? model.variable_lower : std::vector<Scalar>(lo.begin(), lo.end());
    b.upper = hi.empty() ? model.variable_upper : std::vector<Scalar>(hi.begin(), hi.end());
    for (std::size_t i = 0; i < n; ++i)
        if (std::isnan(b.lower[i]) || std::isnan(b.upper[i]) || b.lower[i] > b.upper[i]) {
            message = "variable bound overrides contain an invalid interval";
            return false;
        }
    return true;
}

Scalar scale_bound(Scalar v, Scalar f, bool multiply) {
    if (!std::isfinite(v)) return v;
    return multiply ? v * f : v / f;
}

// Further diagonal preconditioning used by the adaptive policy, on top of the
// reference scaling: `ruiz` passes of Ruiz equilibration (divide each row and
// column by the square root of its largest magnitude) followed by one
// Pock-Chambolle step with alpha = 1 (divide row i by sqrt(sum_j |a_ij|) and
// column j by sqrt(sum_i |a_ij|)) -- the default preconditioner of PDLP
// (Applegate et al., NeurIPS 2021, sec. 3.3).  Factors accumulate into r, s;
// c, bounds and row bounds are transformed as in the reference scaling.
void equilibrate(const CsrPattern &p, Scaled &sc, int ruiz) {
    const auto m = static_cast<std::size_t>(p.rows);
    const auto n = static_cast<std::size_t>(p.columns);
    std::vector<Scalar> rf(m), cf(n);
    auto apply = [&]() {
        for (std::size_t i = 0; i < m; ++i)
            for (Index k = p.row_offsets[i]; k < p.row_offsets[i + 1]; ++k) {
                const auto kk = static_cast<std::size_t>(k);
                sc.values[kk] /= rf[i] * cf[static_cast<std::size_t>(p.column_indices[kk])];
            }
        for (std::size_t i = 0; i < m; ++i) {
            sc.r[i] /= rf[i];
            if (std::isfinite(sc.rl[i])) sc.rl[i] /= rf[i];
            if (std::isfinite(sc.ru[i])) sc.ru[i] /= rf[i];
        }
        for (std::size_t j = 0; j < n; ++j) {
            sc.s[j] /= cf[j];
            sc.c[j] /= cf[j];
            if (std::isfinite(sc.l[j])) sc.l[j] *= cf[j];
            if (std::isfinite(sc.u[j])) sc.u[j] *= cf[j];
        }
    };
    for (int pass = 0; pass < ruiz; ++pass) {
        std::fill(rf.begin(), rf.end(), 0.0);
        std::fill(cf.begin(), cf.end(), 0.0);
        for (std::size_t i = 0; i < m; ++i)
            for (Index k = p.row_offsets[i]; k < p.row_offsets[i + 1]; ++k) {
                const auto kk = static_cast<std::size_t>(k);
                const Scalar a = std::abs(sc.values[kk]);
                rf[i] = std::max(rf[i], a);
                auto &c = cf[static_cast<std::size_t>(p.column_indices[kk])];
                c = std::max(c, a);
            }
        for (auto &v : rf) v = v > 0.0 ? std::sqrt(v) : 1.0;
        for (auto &v : cf) v = v > 0.0 ? std::sqrt(v) : 1.0;
        apply();
    }
    std::fill(rf.begin(), rf.end(), 0.0);
    std::fill(cf.begin(), cf.end(), 0.0);
    for (std::size_t i = 0; i < m; ++i)
        for (Index k = p.row_offsets[i]; k < p.row_offsets[i + 1]; ++k) {
            const auto kk = static_cast<std::size_t>(k);
            const Scalar a = std::abs(sc.values[kk]);
            rf[i] += a;
            cf[static_cast<std::size_t>(p.column_indices[kk])] += a;
        }
    for (auto &v : rf) v = v > 0.0 ? std::sqrt(v) : 1.0;
    for (auto &v : cf) v = v > 0.0 ? std::sqrt(v) : 1.0;
    apply();
}

// continuous_common.hpp:133-205 with the column pass in O(nnz).
Scaled scale(const Model &model, const Bounds &bounds) {
    const auto &A = model.constraint_matrix;
    const auto &rp = A.pattern.row_offsets;
    const auto &ci = A.pattern.column_indices;
    const auto m = static_cast<std::size_t>(A.pattern.rows);
    const auto n = static_cast<std::size_t>(A.pattern.columns);
    Scaled sc;
    sc.r.assign(m, 1.0);
    sc.s.assign(n, 1.0);
    std::vector<Scalar> colmax(n, 0.0);
    for (std::size_t i = 0; i < m; ++i) {
        Scalar mx = 0.0;
        for (Index k = rp[i]; k < rp[i + 1]; ++k) {
            const Scalar a = std::abs(A.values[static_cast<std::size_t>(k)]);
            mx = std::max(mx, a);
            auto &cm = colmax[static_cast<std::size_t>(ci[static_cast<std::size_t>(k)])];
            cm = std::max(cm, a);
        }
        if (mx > 1.0) sc.r[i] = 1.0 / mx;
    }
    for (std::size_t j = 0; j < n; ++j)
        if (colmax[j] > 1.0) sc.s[j] = 1.0 / colmax[j];
    sc.values.resize(A.values.size());
    for (std::size_t i = 0; i < m; ++i)
        for (Index k = rp[i]; k < rp[i + 1]; ++k) {
            const auto kk = static_cast<std::size_t>(k);
            sc.values[kk] = A.values[kk] * (sc.r[i] * sc.s[static_cast<std::size_t>(ci[kk])]);
        }
    sc.c.resize(n);
    sc.l.resize(n);
    sc.u.resize(n);
    for (std::size_t j = 0; j < n; ++j) {
        sc.c[j] = model.linear_objective[j] * sc.s[j];
        sc.l[j] = scale_bound(bounds.lower[j], sc.s[j], false);
        sc.u[j] = scale_bound(bounds.upper[j], sc.s[j], false);
    }
    sc.rl.resize(m);
    sc.ru.resize(m);
    for (std::size_t i = 0; i < m; ++i) {
        sc.rl[i] = scale_bound(model.constraint_lower[i], sc.r[i], true);
        sc.ru[i] = scale_bound(model.constraint_upper[i], sc.r[i], true);
    }
    return sc;
}

// continuous_common.hpp:207-225: sqrt(max row abs-sum * max column abs-sum).
Scalar operator_bound(const CsrPattern &p, const std::vector<Scalar> &v) {
    Scalar max_row = 0.0;
    std::vector<Scalar> cols(static_cast<std::size_t>(p.columns), 0.0);
    for (Index i = 0; i < p.rows; ++i) {
        Scalar sum = 0.0;
        for (Index k = p.row_offsets[static_cast<std::size_t>(i)]; k < p.row_offsets[static_cast<std::size_t>(i + 1)]; ++k) {
            const Scalar a = std::abs(v[static_cast<std::size_t>(k)]);
            sum += a;
            cols[static_cast<std::size_t>(p.column_indices[static_cast<std::size_t>(k)])] += a;
        }
        max_row = std::max(max_row, sum);
    }
    Scalar max_col = 0.0;
    for (const Scalar c : cols) max_col = std::max(max_col, c);
    return std::sqrt(max_row * max_col);
}

// Farkas test of a dual vector y (reference sign, original coordinates):
// D0(y) = sum_j min(g_j l_j, g_j u_j) - sum_i (max(y_i,0) ru_i -
// max(-y_i,0) rl_i) with g = A'y.  A feasible LP has D0(y) <= 0 for every y,
// so D0(y) > 0 with every term finite proves primal infeasibility.  Returns
// whether y passes with D0 above kFarkasMargin times the sum of the term
// magnitudes (far above FP64 rounding).
bool farkas_certificate(const Model &model, const Bounds &bounds, const std::vector<Scalar> &y, Scalar &d0) {
    constexpr Scalar kFarkasMargin = 1e-9;
    const auto &A = model.constraint_matrix;
    const auto m = static_cast<std::size_t>(A.pattern.rows);
    const auto n = static_cast<std::size_t>(A.pattern.columns);
    std::vector<Scalar> g(n, 0.0);
    for (std::size_t i = 0; i < m; ++i)
        for (Index k = A.pattern.row_offsets[i]; k < A.pattern.row_offsets[i + 1]; ++k)
            g[static_cast<std::size_t>(A.pattern.column_indices[k])] += y[i] * A.values[static_cast<std::size_t>(k)];
    Scalar sum = 0.0, mag = 0.0;
    for (std::size_t j = 0; j < n; ++j) {
        if (g[j] == 0.0) continue;
        const Scalar bound = g[j] > 0.0 ? bounds.lower[j] : bounds.upper[j];
        if (!std::isfinite(bound)) return false;
        sum += g[j] * bound;
        mag += std::abs(g[j] * bound);
    }
    for (std::size_t i = 0; i < m; ++i) {
        if (y[i] == 0.0) continue;
        const Scalar bound = y[i] > 0.0 ? model.constraint_upper[i] : model.constraint_lower[i];
        if (!std::isfinite(bound)) return false;
        sum -= y[i] * bound;
        mag += std::abs(y[i] * bound);
    }
    d0 = sum;
    return sum > 0.0 && sum > kFarkasMargin * mag;
}

// After postsolve of a solve the device stopped as primal infeasible:
// re-check the returned row duals in FP64 on the original data, then report
// SolveStatus::Infeasible (or NumericalFailure if the ray does not pass).
void report_infeasible(const Model &model, const Bounds &bounds, const std::vector<Scalar> &y, SolveResult &res) {
    Scalar d0 = 0.0;
    if (farkas_certificate(model, bounds, y, d0)) {
        res.status = SolveStatus::Infeasible;
        res.message = "primal infeasible: the returned row duals are a Farkas ray (dual objective without c = " +
                      std::to_string(d0) + " > 0)";
        res.objective_value = std::numeric_limits<Scalar>::infinity();
        res.primal_bound = std::numeric_limits<Scalar>::infinity();
        res.dual_bound = std::numeric_limits<Scalar>::infinity();
    } else {
        res.status = SolveStatus::NumericalFailure;
        res.message = "the device's primal infeasibility certificate did not pass the FP64 check";
    }
}

// The finite row bound magnitude per row (relative-KKT denominator).
Scalar row_bound_norm(const std::vector<Scalar> &lo, const std::vector<Scalar> &hi) {
    Scalar s = 0.0;
    for (std::size_t i = 0; i < lo.size(); ++i) {
        Scalar b = 0.0;
        if (std::isfinite(lo[i]) && std::isfinite(hi[i])) b = std::max(std::abs(lo[i]), std::abs(hi[i]));
        else if (std::isfinite(lo[i])) b = lo[i];
        else if (std::isfinite(hi[i])) b = hi[i];
        s += b * b;
    }
    return std::sqrt(s);
}

Scalar norm2(const std::vector<Scalar> &v) {
    Scalar s = 0.0;
    for (const Scalar x : v) s += x * x;
    return std::sqrt(s);
}

// Relative KKT error in original coordinates (the reference's dual sign,
// stationarity g = c + A'y as in pdhg.cpp:189-198):
//   rel_p = ||Ax - P_[rl,ru](Ax)||_2 / (1 + ||bhat||_2)
//   rel_d = ||x - P_[l,u](x - g)||_2 / (1 + ||c||_2)
//   gap   = |c'x - D(y)| / (1 + |c'x| + |D(y)|),
//   D(y)  = sum_j min(g_j l_j, g_j u_j) - sum_i (max(y_i,0) ru_i - max(-y_i,0) rl_i)
// with 0*inf = 0; D = -inf gives gap = 1.  A QP passes xqx = x'Qx (with
// g = c + Qx + A'y): the primal objective gains 1/2 x'Qx and the (Wolfe)
// dual objective -1/2 x'Qx + D(y).
void relative_kkt(const Model &model, const Bounds &bounds, const std::vector<Scalar> &x,
                  const std::vector<Scalar> &y, const Products &pr, SolveResult &res, Scalar xqx) {
    const auto m = static_cast<std::size_t>(model.constraint_matrix.pattern.rows);
    const auto n = static_cast<std::size_t>(model.constraint_matrix.pattern.columns);
    const std::vector<Scalar> &g = pr.g;
    Scalar pres = 0.0, dual = 0.0;
    bool neg_inf = false;
    for (std::size_t i = 0; i < m; ++i) {
        const Scalar ax = pr.ax[i];
        const Scalar lo = model.constraint_lower[i], hi = model.constraint_upper[i];
        const Scalar d = ax - std::min(std::max(ax, lo), hi);
        pres += d * d;
        if (y[i] > 0.0) {
            if (std::isinf(hi)) neg_inf = true; else dual -= y[i] * hi;
        } else if (y[i] < 0.0) {
            if (std::isinf(lo)) neg_inf = true; else dual -= y[i] * lo;
        }
    }
    Scalar dres = 0.0, pobj = 0.0;
    for (std::size_t j = 0; j < n; ++j) {
        const Scalar l = bounds.lower[j], u = bounds.upper[j];
        const Scalar d = x[j] - std::min(std::max(x[j] - g[j], l), u);
        dres += d * d;
        pobj += model.linear_objective[j] * x[j];
        if (g[j] > 0.0) {
            if (std::isinf(l)) neg_inf = true; else dual += g[j] * l;
        } else if (g[j] < 0.0) {
            if (std::isinf(u)) neg_inf = true; else dual += g[j] * u;
        }
    }
    pobj += 0.5 * xqx;
    dual -= 0.5 * xqx;
    res.kkt_relative_primal =
        std::sqrt(pres) / (1.0 + row_bound_norm(model.constraint_lower, model.constraint_upper));
    res.kkt_relative_dual = std::sqrt(dres) / (1.0 + norm2(model.linear_objective));
    res.kkt_primal_objective = pobj;
    if (neg_inf) {
        res.kkt_relative_gap = 1.0;
        res.kkt_dual_objective = -kInf;
    } else {
        res.kkt_relative_gap = std::abs(pobj - dual) / (1.0 + std::abs(pobj) + std::abs(dual));
        res.kkt_dual_objective = dual;
    }
    res.kkt_relative = std::max({res.kkt_relative_primal, res.kkt_relative_dual, res.kkt_relative_gap});
}

Products host_products(const Model &model, const std::vector<Scalar> &x, const std::vector<Scalar> &y) {
    const auto &A = model.constraint_matrix;
    const auto m = static_cast<std::size_t>(A.pattern.rows);
    Products pr;
    pr.ax.assign(m, 0.0);
    pr.g = model.linear_objective;
    for (std::size_t i = 0; i < m; ++i)
        for (Index k = A.pattern.row_offsets[i]; k < A.pattern.row_offsets[i + 1]; ++k) {
            const auto kk = static_cast<std::size_t>(k);
            const auto j = static_cast<std::size_t>(A.pattern.column_indices[kk]);
            pr.ax[i] += A.values[kk] * x[j];
            pr.g[j] += y[i] * A.values[kk];
        }
    return pr;
}

Products device_products(const vk_hip_csr *A, const Model &model, const std::vector<Scalar> &x,
                         const std::vector<Scalar> &y) {
    Products pr;
    pr.ax.resize(y.size());
    pr.g.resize(x.size());
    check_vk(vk_hip_csr_postsolve_products(A, x.data(), y.data(), model.linear_objective.data(), pr.ax.data(),
                                           pr.g.data()),
             "vk_hip_csr_postsolve_products");
    return pr;
}

int apply_tile_geometry(vk_hip_csr *A, const SolverOptions &options) {
    int g = options.spmv_tile_geometry;
    if (g < 0) {
        const int precision = options.pdhg_precision == PdhgPrecision::Mixed ? VK_PREC_MIXED : VK_PREC_FP64;
        check_vk(vk_hip_csr_tune_tiles(A, precision, 10, &g, nullptr), "vk_hip_csr_tune_tiles");
    } else {
        check_vk(vk_hip_csr_set_tile_geometry(A, g), "vk_hip_csr_set_tile_geometry");
    }
    return g;
}

std::string tile_geometry_note(int geometry, bool tuned) {
    static const char *const kNames[] = {"512 nonzeros x 64 rows", "256 nonzeros x 64 rows, rows over 256 alone",
                                         "512 nonzeros x 32 rows", "512 nonzeros x 64 rows, rows over 64 alone"};
    if (geometry == 0 && !tuned) return "";
    return std::string("spmv tiles: ") + kNames[geometry] + (tuned ? " (tuned)" : "");
}

Scaled device_scale(vk_hip_csr *A, const Model &model, const Bounds &bounds, bool equilibrated) {
    const auto m = static_cast<std::size_t>(model.num_constraints());
    const auto n = static_cast<std::size_t>(model.num_variables());
    Scaled sc;
    sc.r.resize(m);
    sc.rl.resize(m);
    sc.ru.resize(m);
    sc.s.resize(n);
    sc.c.resize(n);
    sc.l.resize(n);
    sc.u.resize(n);
    check_vk(vk_hip_lp_scale(A, model.linear_objective.data(), bounds.lower.data(), bounds.upper.data(),
                             model.constraint_lower.data(), model.constraint_upper.data(),
                             equilibrated ? kRuizPasses : 0, equilibrated ? 1 : 0, sc.r.data(), sc.s.data(),
                             sc.c.data(), sc.l.data(), sc.u.data(), sc.rl.data(), sc.ru.data()),
             "vk_hip_lp_scale");
    return sc;
}

// pdhg.cpp:153-264: objective, residuals, duals and the final status, in
// original coordinates.  `x`, `y` are unscaled; `pr` holds A x and c + A'y.
void postsolve(const Model &model, const Bounds &bounds, const SolverOptions &options,
               const std::vector<Scalar> &x, const std::vector<Scalar> &y, const Products &pr,
               bool relative_mode, SolveResult &res) {
    const auto m = static_cast<std::size_t>(model.constraint_matrix.pattern.rows);
    const auto n = static_cast<std::size_t>(model.constraint_matrix.pattern.columns);
    res.primal_solution = x;
    Scalar obj = 0.0;
    for (std::size_t j = 0; j < n; ++j) obj += model.linear_objective[j] * x[j];
    res.objective_value = model.objective_offset + obj;
    if (!std::isfinite(res.objective_value)) {
        res.status = SolveStatus::NumericalFailure;
        res.message = "PDHG produced a non-finite objective";
    }
    const std::vector<Scalar> &row_values = pr.ax;
    const std::vector<Scalar> &stationarity = pr.g;
    Scalar viol = 0.0;
    for (std::size_t i = 0; i < m; ++i) {
        if (std::isfinite(model.constraint_lower[i])) viol = std::max(viol, model.constraint_lower[i] - row_values[i]);
        if (std::isfinite(model.constraint_upper[i])) viol = std::max(viol, row_values[i] - model.constraint_upper[i]);
    }
    Scalar box = 0.0;
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
    res.stationarity_residual = 0.0;
    res.complementarity_residual = 0.0;
    res.variable_lower_dual.assign(n, 0.0);
    res.variable_upper_dual.assign(n, 0.0);
    const Scalar active = std::max<Scalar>(10.0 * options.primal_tolerance, 1e-8);
    for (std::size_t j = 0; j < n; ++j) {
        const Scalar lo = bounds.lower[j], hi = bounds.upper[j];
        const Scalar proj = std::clamp(x[j] - stationarity[j], lo, hi);
        res.stationarity_residual = std::max(res.stationarity_residual, std::abs(x[j] - proj));
        if (std::isfinite(lo) && x[j] - lo <= active) {
            res.variable_lower_dual[j] = std::max<Scalar>(0.0, stationarity[j]);
            res.complementarity_residual =
                std::max(res.complementarity_residual, std::abs((x[j] - lo) * res.variable_lower_dual[j]));
        }
        if (std::isfinite(hi) && hi - x[j] <= active) {
            res.variable_upper_dual[j] = std::max<Scalar>(0.0, -stationarity[j]);
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
        res.message = "PDHG produced a non-finite iterate";
    } else if (res.status == SolveStatus::Optimal && !relative_mode &&
               (res.primal_residual > options.primal_tolerance ||
                res.stationarity_residual > options.dual_tolerance)) {
        // pdhg.cpp:250-255 (in relative-KKT mode the stopping test is itself
        // in original coordinates, so this absolute re-check does not apply).
        res.status = SolveStatus::NumericalFailure;
        res.message = "PDHG stopped without satisfying original-coordinate residual checks";
    }
    if (res.status == SolveStatus::Optimal) {
        res.primal_bound = res.objective_value;
        res.dual_bound = res.objective_value;
        res.absolute_gap = 0.0;
        res.relative_gap = 0.0;
    }
    relative_kkt(model, bounds, x, y, pr, res, 0.0);
}

void check_vk(int rc, const char *what) {
    if (rc != VK_OK) throw std::runtime_error(std::string(what) + ": " + vk_hip_last_error());
}

// Runs one device solve; returns the VK_PDHG_* status and fills the scaled
// iterates.  Honours the time limit between blocks of iterations.
int run_device(const vk_hip_csr *A, const Scaled &sc, const vk_pdhg_params &params,
               const std::vector<Scalar> *x0, const std::vector<Scalar> *y0,
               Clock::time_point started, const SolverOptions &options, std::vector<Scalar> &xs,
               std::vector<Scalar> &ys, vk_pdhg_state &state, double &device_ms, bool &timed) {
    DeviceSolver solver;
    check_vk(vk_hip_pdhg_create(A, sc.c.data(), sc.l.data(), sc.u.data(), sc.rl.data(), sc.ru.data(),
                                sc.r.data(), sc.s.data(), x0 ? x0->data() : nullptr,
                                y0 ? y0->data() : nullptr, &params, &solver.h),
             "vk_hip_pdhg_create");
    return drive_device(solver, params, sc.c.size(), sc.r.size(), started, options, xs, ys, state, device_ms, timed);
}

int drive_device(DeviceSolver &solver, const vk_pdhg_params &params, std::size_t n, std::size_t m,
                 Clock::time_point started, const SolverOptions &options, std::vector<Scalar> &xs,
                 std::vector<Scalar> &ys, vk_pdhg_state &state, double &device_ms, bool &timed) {
    int status = VK_PDHG_RUNNING;
    timed = false;
    device_ms = 0.0;
    // Without a time limit the whole budget is one call; with one, the
    // clock is checked every chunk of iterations.
    const Index chunk = std::isfinite(options.time_limit_seconds) ? 512 : params.iteration_limit;
    Index done = 0;
    while (status == VK_PDHG_RUNNING) {
        if (timed_out(started, options)) {
            timed = true;
            break;
        }
        const Index step = std::max<Index>(1, std::min(chunk, params.iteration_limit - done));
        check_vk(vk_hip_pdhg_run(solver.h, step, &status), "vk_hip_pdhg_run");
        double ms = 0.0;
        check_vk(vk_hip_pdhg_last_run_ms(solver.h, &ms), "vk_hip_pdhg_last_run_ms");
        device_ms += ms;
        check_vk(vk_hip_pdhg_read_state(solver.h, &state), "vk_hip_pdhg_read_state");
        done = state.iterations;
    }
    check_vk(vk_hip_pdhg_read_state(solver.h, &state), "vk_hip_pdhg_read_state");
    xs.resize(n);
    ys.resize(m);
    check_vk(vk_hip_pdhg_download(solver.h, xs.data(), ys.data()), "vk_hip_pdhg_download");
    return state.status;
}

Scalar fp32_phase_tolerance(const SolverOptions &o, bool recentred) {
    Scalar t = o.mixed_phase_tolerance > 0.0 ? o.mixed_phase_tolerance : (recentred ? 0.0 : 1e-4);
    t = std::max(t, o.relative_kkt_tolerance);
    // A plain FP32 Halpern phase stalls between 6e-6 and 4e-5 (measured on
    // the 400k x 1M LP: the FP32 rounding of every output, relative to the
    // iterate, accumulated by the Halpern iteration between restarts; an
    // FP64 Halpern state or stochastic rounding of the outputs does not lower
    // it), so it hands over no later than 1e-5.  The recentred phase rounds
    // relative to the correction and has no such floor.
    if (!recentred && o.pdhg_policy == PdhgPolicy::Halpern) t = std::max(t, kHalpernFp32Floor);
    return t;
}

}  // namespace detail

namespace {

using namespace detail;

vk_pdhg_params make_params(const Model &model, const Scaled &sc, const vk_hip_csr *A,
                           const SolverOptions &options) {
    vk_pdhg_params p{};
    p.policy = options.pdhg_policy == PdhgPolicy::Adaptive  ? VK_PDHG_POLICY_ADAPTIVE
               : options.pdhg_policy == PdhgPolicy::Halpern ? VK_PDHG_POLICY_HALPERN
                                                            : VK_PDHG_POLICY_REFERENCE;
    p.iteration_limit = options.iteration_limit;
    p.metric_tolerance = std::max(options.primal_tolerance, options.dual_tolerance);
    p.relative_kkt_tolerance = options.relative_kkt_tolerance;
    p.restart_period = options.restart_period;
    p.use_graphs = options.use_graphs ? 1 : 0;
    p.split_passes = options.spmv_split_passes;
    p.cache_budget_mb = options.spmv_cache_budget_mb;
    p.spmv_panels = options.spmv_panels;
    p.panel_groups = options.spmv_panel_groups;
    p.panel_groups_t = options.spmv_panel_groups_t;
    p.panel_threads = options.spmv_panel_threads;
    p.panel_kb = options.spmv_panel_kb;
    p.panel_tune = options.spmv_autotune ? 1 : 0;
    p.primal_weight_gain = options.pdhg_primal_weight_gain;
    p.primal_weight_integral = options.pdhg_primal_weight_integral;
    p.norm_b = row_bound_norm(model.constraint_lower, model.constraint_upper);
    p.norm_c = norm2(model.linear_objective);
    if (p.policy == VK_PDHG_POLICY_REFERENCE) {
        // pdhg.cpp:75-78, the bound over A's scaled device values
        double bound = 0.0;
        check_vk(vk_hip_csr_operator_bound(A, &bound), "vk_hip_csr_operator_bound");
        p.step_size = 0.99 / std::max<Scalar>(1.0, bound);
        p.primal_weight = 1.0;
    } else {
        double norm = 0.0;
        check_vk(vk_hip_csr_norm2_estimate(A, 100, &norm), "vk_hip_csr_norm2_estimate");
        p.step_size = 0.99 / std::max(norm, 1e-12);
        const Scalar nc = norm2(sc.c);
        const Scalar nb = row_bound_norm(sc.rl, sc.ru);
        p.primal_weight = (nc > 1e-10 && nb > 1e-10) ? nc / nb : 1.0;
    }
    return p;
}

// Column-block layout of A x for a device solve in `precision`.
int col_blocks_for(const vk_hip_csr *A, const SolverOptions &options, int precision) {
    if (!options.spmv_autotune) return options.spmv_col_blocks;
    int best = 1;
    check_vk(vk_hip_csr_tune_col_blocks(const_cast<vk_hip_csr *>(A), precision, 10, &best, nullptr),
             "vk_hip_csr_tune_col_blocks");
    return best;
}

SolveResult solve_fused(const Model &model, const Bounds &bounds, const Scaled &sc,
                        const vk_hip_csr *A, const SolverOptions &options,
                        const std::vector<Scalar> *warm_scaled, Clock::time_point started) {
    SolveResult res;
    res.solver = "PDHG";
    res.backend = "provided";
    vk_pdhg_params p = make_params(model, sc, A, options);
    std::vector<Scalar> xs, ys;
    vk_pdhg_state state{};
    double ms = 0.0;
    bool timed = false;
    int status;
    if (options.pdhg_precision == PdhgPrecision::Mixed) {
        // FP32 phase to the mixed-phase tolerance, then an FP64 phase
        // warm-started from its primal and dual iterates.  The FP64 phase
        // always runs: it certifies the returned point in FP64 (an FP32 phase
        // that met the target on its FP64-evaluated average can miss it by
        // the rounding of that average to FP32) and usually stops after its
        // first iterations.  Under the Halpern policy the FP32 phase is
        // recentred (unless options.mixed_recentre is off) and by default
        // runs to the final target itself.
        vk_pdhg_params p32 = p;
        p32.precision = VK_PREC_MIXED;
        p32.col_blocks = col_blocks_for(A, options, VK_PREC_MIXED);
        p32.recentre = (p.policy == VK_PDHG_POLICY_HALPERN && options.mixed_recentre) ? 1 : 0;
        p32.relative_kkt_tolerance = fp32_phase_tolerance(options, p32.recentre == 1);
        status = run_device(A, sc, p32, warm_scaled, nullptr, started, options, xs, ys, state, ms, timed);
        res.low_precision_iterations = state.iterations;
        Index used = state.iterations;
        double total_ms = ms;
        if (!timed && used < options.iteration_limit && status != VK_PDHG_PRIMAL_INFEASIBLE) {
            const auto x32 = xs, y32 = ys;
            vk_pdhg_params p64 = p;
            p64.precision = VK_PREC_FP64;
            p64.col_blocks = col_blocks_for(A, options, VK_PREC_FP64);
            p64.iteration_limit = options.iteration_limit - used;
            if (p.policy != VK_PDHG_POLICY_REFERENCE) p64.primal_weight = state.primal_weight;
            status = run_device(A, sc, p64, &x32, &y32, started, options, xs, ys, state, ms, timed);
            total_ms += ms;
            used += state.iterations;
        }
        state.iterations = used;
        ms = total_ms;
    } else {
        p.precision = VK_PREC_FP64;
        p.col_blocks = col_blocks_for(A, options, VK_PREC_FP64);
        status = run_device(A, sc, p, warm_scaled, nullptr, started, options, xs, ys, state, ms, timed);
    }
    res.iterations = state.iterations;
    res.device_seconds = ms * 1e-3;
    if (timed) res.status = SolveStatus::TimeLimit;
    else if (status == VK_PDHG_CONVERGED) res.status = SolveStatus::Optimal;
    std::vector<Scalar> x(xs.size()), y(ys.size());
    for (std::size_t j = 0; j < xs.size(); ++j) x[j] = xs[j] * sc.s[j];
    for (std::size_t i = 0; i < ys.size(); ++i) y[i] = ys[i] * sc.r[i];
    postsolve(model, bounds, options, x, y, device_products(A, model, x, y), options.relative_kkt_tolerance > 0.0,
              res);
    if (status == VK_PDHG_PRIMAL_INFEASIBLE && !timed) report_infeasible(model, bounds, y, res);
    return res;
}

// pdhg.cpp:35-151 issued call for call on an ExecutionBackend.
SolveResult solve_ops(const Model &model, const Bounds &bounds, const Scaled &sc,
                      ExecutionBackend &be, BackendMatrix &A, const SolverOptions &options,
                      const std::vector<Scalar> *warm_scaled, Clock::time_point started) {
    const Index m = model.num_constraints();
    const Index n = model.num_variables();
    auto x = be.create_vector(n), x_next = be.create_vector(n), x_bar = be.create_vector(n);
    auto y = be.create_vector(m), aty = be.create_vector(n), ax_bar = be.create_vector(m);
    auto y_arg = be.create_vector(m), y_proj = be.create_vector(m), ax = be.create_vector(m);
    auto pax = be.create_vector(m), grad = be.create_vector(n), px = be.create_vector(n);
    auto resid = be.create_vector(n);
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
    const Scalar max_step = 0.99 / std::max<Scalar>(1.0, operator_bound(model.constraint_matrix.pattern, sc.values));
    Scalar tau = max_step, sigma = max_step, prev = kInf;
    bool restart = false;
    SolveResult res;
    res.solver = "PDHG";
    res.backend = "provided";
    for (Index iter = 0; iter < options.iteration_limit; ++iter) {
        if (timed_out(started, options)) {
            res.status = SolveStatus::TimeLimit;
            break;
        }
        be.multiply(1.0, A, *y, 0.0, *aty, Transpose::Yes);
        be.copy(*x, *x_next);
        be.axpby(-tau, *c, 1.0, *x_next);
        be.axpby(-tau, *aty, 1.0, *x_next);
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
        be.copy(*c, *grad);
        be.axpby(1.0, *aty, 1.0, *grad);
        be.copy(*x, *px);
        be.axpby(-1.0, *grad, 1.0, *px);
        be.project_box(*px, *l, *u);
        be.copy(*x, *resid);
        be.axpby(-1.0, *px, 1.0, *resid);
        const Scalar stationarity = be.norm_inf(*resid);
        const Scalar metric = std::max(primal, stationarity);
        if (iter > 0 && iter % 50 == 0) {
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
    std::vector<Scalar> xs(static_cast<std::size_t>(n)), ys(static_cast<std::size_t>(m));
    be.download(*x, xs);
    be.download(*y, ys);
    for (Index j = 0; j < n; ++j) xs[static_cast<std::size_t>(j)] *= sc.s[static_cast<std::size_t>(j)];
    for (Index i = 0; i < m; ++i) ys[static_cast<std::size_t>(i)] *= sc.r[static_cast<std::size_t>(i)];
    postsolve(model, bounds, options, xs, ys, host_products(model, xs, ys), false, res);
    return res;
}

}  // namespace

KernelProfile profile_lp_kernels(const Model &model, const SolverOptions &options,
                                 Index warm_iterations, Index iterations) {
    Bounds bounds;
    std::string message;
    if (!resolve_bounds(model, {}, {}, bounds, message) || !model.validate().empty() ||
        model.num_variables() == 0 || iterations <= 0)
        throw std::invalid_argument("profile_lp_kernels: unusable model or arguments");
    vk_hip_csr *A = nullptr;
    check_vk(vk_hip_csr_create(model.constraint_matrix.pattern.rows, model.constraint_matrix.pattern.columns,
                               model.constraint_matrix.pattern.row_offsets.data(),
                               model.constraint_matrix.pattern.column_indices.data(), &A),
             "vk_hip_csr_create");
    struct Guard {
        vk_hip_csr *a;
        ~Guard() { vk_hip_csr_destroy(a); }
    } guard{A};
    check_vk(vk_hip_csr_set_values(A, model.constraint_matrix.values.data()), "vk_hip_csr_set_values");
    const Scaled sc = device_scale(A, model, bounds, options.pdhg_policy != PdhgPolicy::Reference);
    vk_pdhg_params p = make_params(model, sc, A, options);
    p.precision = options.pdhg_precision == PdhgPrecision::Mixed ? VK_PREC_MIXED : VK_PREC_FP64;
    p.col_blocks = col_blocks_for(A, options, p.precision);
    // The FP32 kernels as the mixed solve runs them (recentred under Halpern).
    p.recentre = (p.precision == VK_PREC_MIXED && p.policy == VK_PDHG_POLICY_HALPERN && options.mixed_recentre) ? 1 : 0;
    p.iteration_limit = warm_iterations + iterations + 1;
    p.relative_kkt_tolerance = 0.0;
    p.metric_tolerance = 0.0;
    DeviceSolver solver;
    check_vk(vk_hip_pdhg_create(A, sc.c.data(), sc.l.data(), sc.u.data(), sc.rl.data(), sc.ru.data(),
                                sc.r.data(), sc.s.data(), nullptr, nullptr, &p, &solver.h),
             "vk_hip_pdhg_create");
    int status = 0;
    if (warm_iterations > 0) check_vk(vk_hip_pdhg_run(solver.h, warm_iterations, &status), "vk_hip_pdhg_run");
    KernelProfile out;
    double ms[5] = {};
    check_vk(vk_hip_pdhg_profile(solver.h, iterations, ms), "vk_hip_pdhg_profile");
    check_vk(vk_hip_pdhg_layout(solver.h, &out.col_blocks, &out.split_passes), "vk_hip_pdhg_layout");
    check_vk(vk_hip_pdhg_panel_layout(solver.h, out.panel), "vk_hip_pdhg_panel_layout");
    out.dual_spmv_ms = ms[0];
    out.dual_pass_ms = ms[1];
    out.primal_spmv_ms = ms[2];
    out.primal_pass_ms = ms[3];
    out.dual_ms = ms[0] + ms[1];
    out.primal_ms = ms[2] + ms[3];
    out.iteration_ms = ms[4];
    out.step_size = p.step_size;
    return out;
}

SolveResult solve_lp(const Model &model, ExecutionBackend &backend, const SolverOptions &options,
                     std::span<const Scalar> warm_start, std::span<const Scalar> variable_lower_override,
                     std::span<const Scalar> variable_upper_override) {
    const auto started = Clock::now();
    Bounds bounds;
    std::string message;
    if (!resolve_bounds(model, variable_lower_override, variable_upper_override, bounds, message))
        return status_result(SolveStatus::InvalidModel, message);
    const auto errors = model.validate();
    if (!errors.empty()) return status_result(SolveStatus::InvalidModel, errors.front());
    if (model.problem_class() != ProblemClass::LP)
        return status_result(SolveStatus::UnsupportedModel, "model class does not match the selected continuous solver");
    if (model.num_variables() == 0) return status_result(SolveStatus::InvalidModel, "LP requires at least one variable");
    if (!valid_options(options))
        return status_result(SolveStatus::InvalidModel, "solver limits and tolerances must be non-negative");
    if (options.continuous_method != ContinuousMethod::Auto && options.continuous_method != ContinuousMethod::PDHG)
        return status_result(SolveStatus::UnsupportedModel, "requested continuous method is not implemented for LP");

    if (!warm_start.empty()) {
        if (static_cast<Index>(warm_start.size()) != model.num_variables())
            return status_result(SolveStatus::InvalidModel, "warm start must be empty or finite and sized to the model");
        for (const Scalar v : warm_start)
            if (!std::isfinite(v))
                return status_result(SolveStatus::InvalidModel, "warm start must be empty or finite and sized to the model");
    }
    const bool device_only = options.pdhg_policy != PdhgPolicy::Reference ||
                             options.pdhg_precision != PdhgPrecision::FP64 ||
                             options.relative_kkt_tolerance > 0.0;
    auto matrix = backend.create_matrix(model.constraint_matrix.pattern);
    vk_hip_csr *handle = options.fused_iterations ? hip_matrix_handle(*matrix) : nullptr;
    if (!handle && device_only)
        return status_result(SolveStatus::UnsupportedModel,
                             "adaptive, Halpern, mixed-precision and relative-KKT PDHG run only fused on a HipBackend");
    // The fused path scales on the device from the original values; the
    // op-by-op path keeps the reference's host scaling (pdhg.cpp:25).
    Scaled sc;
    std::string tiles_note;
    if (handle) {
        backend.set_values(model.constraint_matrix.values, *matrix);
        sc = device_scale(handle, model, bounds, options.pdhg_policy != PdhgPolicy::Reference);
        const int g = apply_tile_geometry(handle, options);
        tiles_note = tile_geometry_note(g, options.spmv_tile_geometry < 0);
    } else {
        sc = scale(model, bounds);
        backend.set_values(sc.values, *matrix);
    }
    std::vector<Scalar> warm(warm_start.begin(), warm_start.end());
    for (std::size_t j = 0; j < warm.size(); ++j) warm[j] /= sc.s[j];   // pdhg.cpp:64-69
    const std::vector<Scalar> *warm_ptr = warm_start.empty() ? nullptr : &warm;
    SolveResult res = handle ? solve_fused(model, bounds, sc, handle, options, warm_ptr, started)
                             : solve_ops(model, bounds, sc, backend, *matrix, options, warm_ptr, started);
    if (!tiles_note.empty()) res.message = res.message.empty() ? tiles_note : res.message + "; " + tiles_note;
    res.solve_seconds = std::chrono::duration<Scalar>(Clock::now() - started).count();
    return res;
}

}  // namespace vikalp::solver
