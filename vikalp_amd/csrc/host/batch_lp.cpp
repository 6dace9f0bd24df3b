// solve_lp_batch: many LPs solved together on the batched persistent HIP
// kernel (vk_hip_batch_*).  Every instance follows solve_lp's semantics:
// the same entry checks (pdhg.cpp:269-293), scaling (continuous_common.hpp:
// 133-205, plus equilibration for the adaptive policy), step rule, postsolve
// (pdhg.cpp:153-264) and status mapping; only the execution is shared.

#include "lp_internal.hpp"

#include "vikalp/solver/continuous.hpp"
#include "vikalp_hip.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

namespace vikalp::solver {
namespace {

using namespace detail;

struct Prepared {
    std::size_t index;   // position in the caller's list
    Bounds bounds;
    Scaled sc;
};

struct DeviceBatch {
    vk_hip_batch *h = nullptr;
    ~DeviceBatch() { vk_hip_batch_destroy(h); }
};

// Runs one device batch over `prep` (FP64 or mixed), optionally warm-started
// from scaled iterates; returns per-instance states and scaled iterates.
void run_batch(const std::vector<const Model *> &models, const std::vector<Prepared> &prep,
               const SolverOptions &options, int precision, double rel_tol, Index iteration_limit,
               const std::vector<double> *x0, const std::vector<double> *y0,
               const std::vector<double> *weights_in, const std::vector<int64_t> *limits,
               Clock::time_point started,
               std::vector<vk_pdhg_state> &states, std::vector<double> &xs, std::vector<double> &ys,
               double &device_ms, bool &timed) {
    const int count = static_cast<int>(prep.size());
    std::vector<int64_t> row_starts(count + 1, 0), col_starts(count + 1, 0);
    for (int i = 0; i < count; ++i) {
        const Model &mod = *models[prep[i].index];
        row_starts[i + 1] = row_starts[i] + mod.num_constraints();
        col_starts[i + 1] = col_starts[i] + mod.num_variables();
    }
    const int64_t m = row_starts[count], n = col_starts[count];
    std::vector<int64_t> rp(static_cast<size_t>(m + 1), 0), ci;
    std::vector<double> val, c, l, u, rl, ru, r, s, steps(count), weights(count), nb(count), nc(count);
    for (int i = 0; i < count; ++i) {
        const Model &mod = *models[prep[i].index];
        const Scaled &sc = prep[i].sc;
        const auto &pat = mod.constraint_matrix.pattern;
        for (Index row = 0; row < pat.rows; ++row) {
            for (Index k = pat.row_offsets[row]; k < pat.row_offsets[row + 1]; ++k) {
                ci.push_back(pat.column_indices[k] + col_starts[i]);
                val.push_back(sc.values[k]);
            }
            rp[row_starts[i] + row + 1] = static_cast<int64_t>(ci.size());
        }
        c.insert(c.end(), sc.c.begin(), sc.c.end());
        l.insert(l.end(), sc.l.begin(), sc.l.end());
        u.insert(u.end(), sc.u.begin(), sc.u.end());
        rl.insert(rl.end(), sc.rl.begin(), sc.rl.end());
        ru.insert(ru.end(), sc.ru.begin(), sc.ru.end());
        r.insert(r.end(), sc.r.begin(), sc.r.end());
        s.insert(s.end(), sc.s.begin(), sc.s.end());
        nb[i] = row_bound_norm(mod.constraint_lower, mod.constraint_upper);
        nc[i] = norm2(mod.linear_objective);
        if (options.pdhg_policy == PdhgPolicy::Reference) {
            steps[i] = 0.99 / std::max<Scalar>(1.0, operator_bound(pat, sc.values));   // pdhg.cpp:75-76
            weights[i] = 1.0;
        } else {
            steps[i] = 0.0;   // power iteration on the device
            const Scalar cn = norm2(sc.c), bn = row_bound_norm(sc.rl, sc.ru);
            weights[i] = weights_in ? (*weights_in)[i] : ((cn > 1e-10 && bn > 1e-10) ? cn / bn : 1.0);
        }
    }
    vk_pdhg_params p{};
    p.policy = options.pdhg_policy == PdhgPolicy::Adaptive  ? VK_PDHG_POLICY_ADAPTIVE
               : options.pdhg_policy == PdhgPolicy::Halpern ? VK_PDHG_POLICY_HALPERN
                                                            : VK_PDHG_POLICY_REFERENCE;
    p.precision = precision;
    p.iteration_limit = iteration_limit;
    p.metric_tolerance = std::max(options.primal_tolerance, options.dual_tolerance);
    p.relative_kkt_tolerance = rel_tol;
    p.restart_period = options.restart_period;
    DeviceBatch batch;
    check_vk(vk_hip_batch_create(count, row_starts.data(), col_starts.data(), rp.data(),
                                 ci.empty() ? nullptr : ci.data(), val.empty() ? nullptr : val.data(),
                                 c.data(), l.data(), u.data(), rl.empty() ? nullptr : rl.data(),
                                 ru.empty() ? nullptr : ru.data(), r.empty() ? nullptr : r.data(), s.data(),
                                 x0 ? x0->data() : nullptr, y0 && !y0->empty() ? y0->data() : nullptr,
                                 steps.data(), weights.data(), nb.data(), nc.data(),
                                 limits ? limits->data() : nullptr, &p, &batch.h),
             "vk_hip_batch_create");
    // Launch granularity: 256 iterations per launch keeps each kernel short
    // and lets the host honour the time limit between launches.
    constexpr int64_t kPerLaunch = 256;
    const int64_t launches = (iteration_limit + kPerLaunch - 1) / kPerLaunch + 1;
    timed = false;
    device_ms = 0.0;
    int running = count;
    for (int64_t launch = 0; launch < launches && running > 0; ++launch) {
        if (timed_out(started, options)) {
            timed = true;
            break;
        }
        check_vk(vk_hip_batch_run(batch.h, kPerLaunch, 1, &running), "vk_hip_batch_run");
        double ms = 0.0;
        check_vk(vk_hip_batch_last_run_ms(batch.h, &ms), "vk_hip_batch_last_run_ms");
        device_ms += ms;
    }
    states.resize(static_cast<size_t>(count));
    for (int i = 0; i < count; ++i) check_vk(vk_hip_batch_read_state(batch.h, i, &states[i]), "vk_hip_batch_read_state");
    xs.assign(static_cast<size_t>(n), 0.0);
    ys.assign(static_cast<size_t>(m), 0.0);
    check_vk(vk_hip_batch_download(batch.h, xs.data(), ys.empty() ? nullptr : ys.data()), "vk_hip_batch_download");
}

}  // namespace

std::vector<SolveResult> solve_lp_batch(const std::vector<const Model *> &models, const SolverOptions &options) {
    const auto started = Clock::now();
    std::vector<SolveResult> results(models.size());
    std::vector<Prepared> prep;
    for (std::size_t i = 0; i < models.size(); ++i) {
        const Model &model = *models[i];
        Prepared pr;
        pr.index = i;
        std::string message;
        if (!resolve_bounds(model, {}, {}, pr.bounds, message)) {
            results[i] = status_result(SolveStatus::InvalidModel, message);
            continue;
        }
        const auto errors = model.validate();
        if (!errors.empty()) {
            results[i] = status_result(SolveStatus::InvalidModel, errors.front());
            continue;
        }
        if (model.problem_class() != ProblemClass::LP) {
            results[i] = status_result(SolveStatus::UnsupportedModel, "model class does not match the selected continuous solver");
            continue;
        }
        if (model.num_variables() == 0) {
            results[i] = status_result(SolveStatus::InvalidModel, "LP requires at least one variable");
            continue;
        }
        if (!valid_options(options)) {
            results[i] = status_result(SolveStatus::InvalidModel, "solver limits and tolerances must be non-negative");
            continue;
        }
        pr.sc = scale(model, pr.bounds);
        if (options.pdhg_policy != PdhgPolicy::Reference)
            equilibrate(model.constraint_matrix.pattern, pr.sc, kRuizPasses);
        prep.push_back(std::move(pr));
    }
    if (prep.empty()) return results;

    std::vector<vk_pdhg_state> states;
    std::vector<double> xs, ys;
    double ms = 0.0, total_ms = 0.0;
    bool timed = false;
    std::vector<Index> used(prep.size(), 0), low(prep.size(), 0);
    // Instances the FP32 phase of a mixed batch proved infeasible, and their
    // scaled row duals (the certificate) from that phase.
    std::vector<char> infeasible32(prep.size(), 0);
    std::vector<double> ys32;
    if (options.pdhg_precision == PdhgPrecision::Mixed) {
        double phase_tol = std::max(options.mixed_phase_tolerance, options.relative_kkt_tolerance);
        if (options.pdhg_policy == PdhgPolicy::Halpern) phase_tol = std::max(phase_tol, kHalpernFp32Floor);
        run_batch(models, prep, options, VK_PREC_MIXED, phase_tol, options.iteration_limit, nullptr, nullptr,
                  nullptr, nullptr, started, states, xs, ys, ms, timed);
        total_ms += ms;
        for (std::size_t i = 0; i < prep.size(); ++i) used[i] = low[i] = states[i].iterations;
        if (!timed) {
            // FP64 refinement of every instance, warm-started, each with the
            // iterations its FP32 phase left (an instance that used its whole
            // budget stops at once with IterationLimit).
            std::vector<double> w(prep.size());
            std::vector<int64_t> remaining(prep.size());
            Index max_remaining = 0;
            for (std::size_t i = 0; i < prep.size(); ++i) {
                w[i] = states[i].primal_weight;
                infeasible32[i] = states[i].status == VK_PDHG_PRIMAL_INFEASIBLE;
                remaining[i] = infeasible32[i] ? 0 : std::max<Index>(0, options.iteration_limit - used[i]);
                max_remaining = std::max<Index>(max_remaining, remaining[i]);
            }
            ys32 = ys;
            const auto x32 = xs, y32 = ys;
            run_batch(models, prep, options, VK_PREC_FP64, options.relative_kkt_tolerance, max_remaining, &x32,
                      &y32, options.pdhg_policy != PdhgPolicy::Reference ? &w : nullptr, &remaining, started,
                      states, xs, ys, ms, timed);
            total_ms += ms;
            for (std::size_t i = 0; i < prep.size(); ++i) used[i] += states[i].iterations;
        }
    } else {
        run_batch(models, prep, options, VK_PREC_FP64, options.relative_kkt_tolerance, options.iteration_limit,
                  nullptr, nullptr, nullptr, nullptr, started, states, xs, ys, ms, timed);
        total_ms += ms;
        for (std::size_t i = 0; i < prep.size(); ++i) used[i] = states[i].iterations;
    }
    Index row0 = 0, col0 = 0;
    for (std::size_t i = 0; i < prep.size(); ++i) {
        const Model &model = *models[prep[i].index];
        const Scaled &sc = prep[i].sc;
        const Index m = model.num_constraints(), n = model.num_variables();
        SolveResult res;
        res.solver = "PDHG";
        res.backend = "hip";
        res.iterations = used[i];
        res.low_precision_iterations = low[i];
        res.device_seconds = total_ms * 1e-3;
        const bool infeasible = infeasible32[i] || states[i].status == VK_PDHG_PRIMAL_INFEASIBLE;
        if (timed && states[i].status == VK_PDHG_RUNNING && !infeasible) res.status = SolveStatus::TimeLimit;
        else if (states[i].status == VK_PDHG_CONVERGED && !infeasible) res.status = SolveStatus::Optimal;
        const std::vector<double> &ysrc = infeasible32[i] ? ys32 : ys;
        std::vector<Scalar> x(static_cast<std::size_t>(n)), y(static_cast<std::size_t>(m));
        for (Index j = 0; j < n; ++j) x[j] = xs[col0 + j] * sc.s[j];
        for (Index r = 0; r < m; ++r) y[r] = ysrc[row0 + r] * sc.r[r];
        postsolve(model, prep[i].bounds, options, x, y, host_products(model, x, y),
                  options.relative_kkt_tolerance > 0.0, res);
        if (infeasible) report_infeasible(model, prep[i].bounds, y, res);
        res.solve_seconds = std::chrono::duration<Scalar>(Clock::now() - started).count();
        results[prep[i].index] = std::move(res);
        row0 += m;
        col0 += n;
    }
    return results;
}

}  // namespace vikalp::solver
