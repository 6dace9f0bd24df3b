// solve_lp_batch: many LPs solved together on the batched persistent HIP
// kernel (vk_hip_batch_*).  Every instance follows solve_lp's semantics:
// the same entry checks (pdhg.cpp:269-293), scaling (continuous_common.hpp:
// 133-205, plus equilibration for the adaptive and Halpern policies), step
// rule, postsolve (pdhg.cpp:153-264) and status mapping; only the execution
// is shared.  The instances are stacked into one block-diagonal matrix that
// is transposed, scaled and multiplied for the postsolve on the device.

#include "lp_internal.hpp"

#include "vikalp/solver/continuous.hpp"
#include "vikalp_hip.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <exception>
#include <memory>
#include <span>
#include <cmath>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace vikalp::solver {
namespace {

using namespace detail;

struct Prepared {
    std::size_t index;   // position in the caller's list
    Bounds bounds;
};

// The instances as one block-diagonal LP: original data on the host, the
// matrix on the device (scaled there in place), the scaled vectors back on
// the host.  Block scaling is per instance: no row or column of the
// concatenation touches two instances.
struct Stacked {
    std::vector<int64_t> row_starts, col_starts;
    Scaled sc;                                   // concatenated; no host values
    std::vector<double> steps, weights, nb, nc;  // per instance
    std::vector<double> warm;                    // scaled start x of every instance, or empty
    vk_hip_csr *A = nullptr;
    bool owns_A = true;    // false: A belongs to a NodeCache
    bool shared = false;   // A is one matrix every instance runs on
    int tile_geometry = 0;     // the instances' row tiles (vk_pdhg_params.tile_geometry)
    std::string tiles_note;    // SolveResult::message note of a non-default geometry
    ~Stacked() {
        if (owns_A) vk_hip_csr_destroy(A);
    }
};

// Node batch: every instance is `model` with its own variable bounds (and
// warm start).  The matrix is uploaded, transposed and scaled once; each
// node's bounds are scaled by the same operations solve_lp applies to them
// (vk_hip_lp_scale_bounds), so each node starts from exactly the data of its
// own solve_lp.  The matrix, its scaling, tile geometry and step live in
// `cache`, built at its first use and reused by later calls.
void stack_shared(const Model &model, const std::vector<Prepared> &prep, const std::vector<std::vector<Scalar>> &warm,
                  const SolverOptions &options, Stacked &st, NodeCache &cache) {
    const int count = static_cast<int>(prep.size());
    const int64_t m = model.num_constraints(), n = model.num_variables();
    st.shared = true;
    st.row_starts.resize(count + 1);
    st.col_starts.resize(count + 1);
    for (int i = 0; i <= count; ++i) {
        st.row_starts[i] = i * m;
        st.col_starts[i] = i * n;
    }
    const bool equilibrated = options.pdhg_policy != PdhgPolicy::Reference;
    if (!cache.A) {
        const auto &pat = model.constraint_matrix.pattern;
        check_vk(vk_hip_csr_create(pat.rows, pat.columns, pat.row_offsets.data(),
                                   pat.column_indices.empty() ? nullptr : pat.column_indices.data(), &cache.A),
                 "vk_hip_csr_create");
        check_vk(vk_hip_csr_set_values(cache.A, model.constraint_matrix.values.empty()
                                                    ? nullptr
                                                    : model.constraint_matrix.values.data()),
                 "vk_hip_csr_set_values");
        // Scaling depends on A only: once, with the first node's bounds;
        // every node's bounds then go through the same recorded column
        // factors.
        cache.base = device_scale(cache.A, model, prep[0].bounds, equilibrated);
        cache.tile_geometry = apply_tile_geometry(cache.A, options);
        cache.tiles_note = tile_geometry_note(cache.tile_geometry, options.spmv_tile_geometry < 0);
        // The step of solve_lp: the reference operator bound
        // (pdhg.cpp:75-78), or 0.99 over the power-iteration estimate of
        // ||A||_2.
        if (!equilibrated) {
            double bound = 0.0;
            check_vk(vk_hip_csr_operator_bound(cache.A, &bound), "vk_hip_csr_operator_bound");
            cache.step = 0.99 / std::max<Scalar>(1.0, bound);
        } else {
            double norm = 0.0;
            check_vk(vk_hip_csr_norm2_estimate(cache.A, 100, &norm), "vk_hip_csr_norm2_estimate");
            cache.step = 0.99 / std::max(norm, 1e-12);
        }
    }
    st.A = cache.A;
    st.owns_A = false;
    st.tile_geometry = cache.tile_geometry;
    st.tiles_note = cache.tiles_note;
    Scaled &sc = st.sc;
    const Scaled &s = cache.base;
    std::vector<double> lo, hi;
    for (int i = 0; i < count; ++i) {
        lo.insert(lo.end(), prep[i].bounds.lower.begin(), prep[i].bounds.lower.end());
        hi.insert(hi.end(), prep[i].bounds.upper.begin(), prep[i].bounds.upper.end());
    }
    sc.l.resize(lo.size());
    sc.u.resize(hi.size());
    check_vk(vk_hip_lp_scale_bounds(st.A, count, lo.data(), hi.data(), sc.l.data(), sc.u.data()),
             "vk_hip_lp_scale_bounds");
    bool any_warm = false;
    for (int i = 0; i < count; ++i) {
        for (auto [dst, src] : {std::pair{&sc.c, &s.c}, std::pair{&sc.s, &s.s}, std::pair{&sc.rl, &s.rl},
                                std::pair{&sc.ru, &s.ru}, std::pair{&sc.r, &s.r}})
            dst->insert(dst->end(), src->begin(), src->end());
        any_warm = any_warm || !warm[static_cast<std::size_t>(prep[i].index)].empty();
    }
    if (any_warm) {
        st.warm.assign(static_cast<std::size_t>(count * n), 0.0);
        for (int i = 0; i < count; ++i) {
            const auto &w = warm[static_cast<std::size_t>(prep[i].index)];
            for (std::size_t j = 0; j < w.size(); ++j)
                st.warm[static_cast<std::size_t>(i * n) + j] = w[j] / sc.s[static_cast<std::size_t>(i * n) + j];   // pdhg.cpp:64-69
        }
    }
    st.steps.assign(count, cache.step);
    const std::vector<Scalar> c0(sc.c.begin(), sc.c.begin() + n), rl0(sc.rl.begin(), sc.rl.begin() + m),
        ru0(sc.ru.begin(), sc.ru.begin() + m);
    const Scalar cn = norm2(c0), bn = row_bound_norm(rl0, ru0);
    st.weights.assign(count, equilibrated && cn > 1e-10 && bn > 1e-10 ? cn / bn : 1.0);
    st.nb.assign(count, row_bound_norm(model.constraint_lower, model.constraint_upper));
    st.nc.assign(count, norm2(model.linear_objective));
}

void stack(const std::vector<const Model *> &models, const std::vector<Prepared> &prep, const SolverOptions &options,
           Stacked &st) {
    const int count = static_cast<int>(prep.size());
    st.row_starts.assign(count + 1, 0);
    st.col_starts.assign(count + 1, 0);
    int64_t nnz = 0;
    for (int i = 0; i < count; ++i) {
        const Model &mod = *models[prep[i].index];
        st.row_starts[i + 1] = st.row_starts[i] + mod.num_constraints();
        st.col_starts[i + 1] = st.col_starts[i] + mod.num_variables();
        nnz += static_cast<int64_t>(mod.constraint_matrix.values.size());
    }
    const int64_t m = st.row_starts[count], n = st.col_starts[count];
    std::vector<int64_t> rp(static_cast<size_t>(m + 1), 0), ci;
    std::vector<double> val, c, l, u, rl, ru;
    ci.reserve(static_cast<size_t>(nnz));
    val.reserve(static_cast<size_t>(nnz));
    for (int i = 0; i < count; ++i) {
        const Model &mod = *models[prep[i].index];
        const auto &pat = mod.constraint_matrix.pattern;
        const int64_t base = static_cast<int64_t>(ci.size());
        for (Index row = 0; row < pat.rows; ++row) rp[st.row_starts[i] + row + 1] = base + pat.row_offsets[row + 1];
        for (const Index col : pat.column_indices) ci.push_back(col + st.col_starts[i]);
        val.insert(val.end(), mod.constraint_matrix.values.begin(), mod.constraint_matrix.values.end());
        c.insert(c.end(), mod.linear_objective.begin(), mod.linear_objective.end());
        l.insert(l.end(), prep[i].bounds.lower.begin(), prep[i].bounds.lower.end());
        u.insert(u.end(), prep[i].bounds.upper.begin(), prep[i].bounds.upper.end());
        rl.insert(rl.end(), mod.constraint_lower.begin(), mod.constraint_lower.end());
        ru.insert(ru.end(), mod.constraint_upper.begin(), mod.constraint_upper.end());
    }
    check_vk(vk_hip_csr_create(m, n, rp.data(), ci.empty() ? nullptr : ci.data(), &st.A), "vk_hip_csr_create");
    check_vk(vk_hip_csr_set_values(st.A, val.empty() ? nullptr : val.data()), "vk_hip_csr_set_values");
    // continuous_common.hpp:133-205 per instance (plus the equilibration of
    // the adaptive and Halpern policies), on the device.
    Scaled &sc = st.sc;
    sc.r.resize(static_cast<size_t>(m));
    sc.rl.resize(static_cast<size_t>(m));
    sc.ru.resize(static_cast<size_t>(m));
    sc.s.resize(static_cast<size_t>(n));
    sc.c.resize(static_cast<size_t>(n));
    sc.l.resize(static_cast<size_t>(n));
    sc.u.resize(static_cast<size_t>(n));
    const bool equilibrated = options.pdhg_policy != PdhgPolicy::Reference;
    check_vk(vk_hip_lp_scale(st.A, c.data(), l.data(), u.data(), rl.data(), ru.data(), equilibrated ? kRuizPasses : 0,
                             equilibrated ? 1 : 0, sc.r.data(), sc.s.data(), sc.c.data(), sc.l.data(), sc.u.data(),
                             sc.rl.data(), sc.ru.data()),
             "vk_hip_lp_scale");
    st.steps.assign(count, 0.0);   // 0: power iteration on the device
    st.weights.assign(count, 1.0);
    st.nb.resize(count);
    st.nc.resize(count);
    if (!equilibrated) {
        std::vector<double> bounds(count);
        check_vk(vk_hip_csr_operator_bounds(st.A, count, st.row_starts.data(), st.col_starts.data(), bounds.data()),
                 "vk_hip_csr_operator_bounds");
        for (int i = 0; i < count; ++i) st.steps[i] = 0.99 / std::max<Scalar>(1.0, bounds[i]);   // pdhg.cpp:75-76
    }
    for (int i = 0; i < count; ++i) {
        const Model &mod = *models[prep[i].index];
        st.nb[i] = row_bound_norm(mod.constraint_lower, mod.constraint_upper);
        st.nc[i] = norm2(mod.linear_objective);
        if (equilibrated) {
            const auto c0 = static_cast<size_t>(st.col_starts[i]), c1 = static_cast<size_t>(st.col_starts[i + 1]);
            const auto r0 = static_cast<size_t>(st.row_starts[i]), r1 = static_cast<size_t>(st.row_starts[i + 1]);
            const Scalar cn = norm2(std::vector<Scalar>(sc.c.begin() + c0, sc.c.begin() + c1));
            const Scalar bn = row_bound_norm(std::vector<Scalar>(sc.rl.begin() + r0, sc.rl.begin() + r1),
                                             std::vector<Scalar>(sc.ru.begin() + r0, sc.ru.begin() + r1));
            st.weights[i] = (cn > 1e-10 && bn > 1e-10) ? cn / bn : 1.0;
        }
    }
}

struct DeviceBatch {
    vk_hip_batch *h = nullptr;
    ~DeviceBatch() { vk_hip_batch_destroy(h); }
};

// Instances still running when the batch hands its tail to the single-LP
// solver.  The batch kernel runs one workgroup per instance: once a few
// instances are left, each runs on one CU while the rest of the GPU idles
// (257 us per iteration for the 18,400-column block LP, against 26 us for
// the same instance on the whole GPU).  The hand-off comes once at most
// kTailInstances (and at most a quarter of the batch) are left; a batch of
// fewer than kTailMinBatch instances stays in the batch kernel.  Measured
// on 512 block LPs with 4 workers and the 1024-thread batch kernel for the
// last instances: 8 / 16 / 32 left gave 7.79 / 7.65 / 7.94 s (FP64) and
// 7.99 / 7.78 / 8.33 s (mixed).
constexpr int kTailInstances = 16;
constexpr int kTailMinBatch = 32;
// Host threads driving the tail's single-LP solvers at once.  Each solver
// has its own streams, and a small LP's iteration leaves most of the GPU
// free, so they overlap (the device runs four hardware queues).
constexpr int kTailWorkers = 4;

// One tail instance on the single-LP device solver: same scaled data
// (scaled again on the device from the model, bitwise the stacked scaling),
// same step and primal weight, warm-started from the batch's scaled
// iterates (a restart of the Halpern or adaptive sequence at that point),
// up to the instance's remaining iterations.
struct TailJob {
    struct Matrix {
        vk_hip_csr *a = nullptr;
        ~Matrix() { vk_hip_csr_destroy(a); }
    };
    int i = 0;
    Matrix A;   // declared before the solver, so destroyed after it
    Scaled sc;
    vk_pdhg_params p{};
    DeviceSolver solver;
    std::vector<Scalar> xo, yo;
    vk_pdhg_state state{};
    double ms = 0.0;
    bool timed = false;
};

// Uploads and scales instance i and creates its solver (calling thread).
std::unique_ptr<TailJob> prepare_tail(const Model &model, const Prepared &pr, const Stacked &st, int i,
                                      const SolverOptions &options, int precision, double rel_tol, Index limit,
                                      const vk_pdhg_state &state, const std::vector<double> &xs,
                                      const std::vector<double> &ys) {
    auto job = std::make_unique<TailJob>();
    job->i = i;
    check_vk(vk_hip_csr_create(model.constraint_matrix.pattern.rows, model.constraint_matrix.pattern.columns,
                               model.constraint_matrix.pattern.row_offsets.data(),
                               model.constraint_matrix.pattern.column_indices.data(), &job->A.a),
             "vk_hip_csr_create");
    check_vk(vk_hip_csr_set_values(job->A.a, model.constraint_matrix.values.data()), "vk_hip_csr_set_values");
    check_vk(vk_hip_csr_set_tile_geometry(job->A.a, st.tile_geometry), "vk_hip_csr_set_tile_geometry");
    job->sc = device_scale(job->A.a, model, pr.bounds, options.pdhg_policy != PdhgPolicy::Reference);
    vk_pdhg_params &p = job->p;
    p.policy = options.pdhg_policy == PdhgPolicy::Adaptive ? VK_PDHG_POLICY_ADAPTIVE : VK_PDHG_POLICY_HALPERN;
    p.precision = precision;
    p.primal_weight_gain = options.pdhg_primal_weight_gain;
    p.primal_weight_integral = options.pdhg_primal_weight_integral;
    p.iteration_limit = std::max<Index>(0, limit - state.iterations);
    p.metric_tolerance = std::max(options.primal_tolerance, options.dual_tolerance);
    p.relative_kkt_tolerance = rel_tol;
    p.restart_period = options.restart_period;
    p.use_graphs = options.use_graphs ? 1 : 0;
    p.split_passes = options.spmv_split_passes;
    p.cache_budget_mb = options.spmv_cache_budget_mb;
    p.col_blocks = options.spmv_col_blocks;
    p.spmv_panels = options.spmv_panels;
    p.panel_groups = options.spmv_panel_groups;
    p.panel_groups_t = options.spmv_panel_groups_t;
    p.panel_threads = options.spmv_panel_threads;
    p.panel_kb = options.spmv_panel_kb;
    p.panel_tune = options.spmv_autotune ? 1 : 0;
    p.norm_b = st.nb[i];
    p.norm_c = st.nc[i];
    p.step_size = std::sqrt(state.tau * state.sigma);   // tau = eta / omega, sigma = eta * omega
    p.primal_weight = state.primal_weight;
    const auto c0 = static_cast<std::size_t>(st.col_starts[i]), c1 = static_cast<std::size_t>(st.col_starts[i + 1]);
    const auto r0 = static_cast<std::size_t>(st.row_starts[i]), r1 = static_cast<std::size_t>(st.row_starts[i + 1]);
    const std::vector<Scalar> x0(xs.begin() + c0, xs.begin() + c1), y0(ys.begin() + r0, ys.begin() + r1);
    const Scaled &sc = job->sc;
    check_vk(vk_hip_pdhg_create(job->A.a, sc.c.data(), sc.l.data(), sc.u.data(), sc.rl.data(), sc.ru.data(),
                                sc.r.data(), sc.s.data(), x0.data(), y0.data(), &p, &job->solver.h),
             "vk_hip_pdhg_create");
    return job;
}

// Finishes the tail instances: solvers are prepared one after another on
// this thread, then kTailWorkers threads drive them concurrently.  The
// device time added is the wall time of the concurrent phase (the GPU is
// busy throughout; per-solver times overlap).
void finish_tail(const std::vector<const Model *> &models, const std::vector<Prepared> &prep, const Stacked &st,
                 const std::vector<int> &tail, const SolverOptions &options, int precision, double rel_tol,
                 const std::vector<int64_t> *limits, Index iteration_limit, Clock::time_point started,
                 std::vector<vk_pdhg_state> &states, std::vector<double> &xs, std::vector<double> &ys,
                 double &device_ms, bool &timed) {
    std::vector<std::unique_ptr<TailJob>> jobs;
    for (int i : tail)
        jobs.push_back(prepare_tail(*models[prep[i].index], prep[i], st, i, options, precision, rel_tol,
                                    limits ? (*limits)[i] : iteration_limit, states[i], xs, ys));
    std::atomic<std::size_t> next{0};
    const int workers = static_cast<int>(std::min<std::size_t>(kTailWorkers, jobs.size()));
    std::vector<std::exception_ptr> errors(static_cast<std::size_t>(workers));
    auto work = [&](int w) {
        try {
            for (std::size_t j; (j = next.fetch_add(1)) < jobs.size();) {
                TailJob &J = *jobs[j];
                drive_device(J.solver, J.p, J.sc.c.size(), J.sc.r.size(), started, options, J.xo, J.yo, J.state,
                             J.ms, J.timed);
            }
        } catch (...) {
            errors[static_cast<std::size_t>(w)] = std::current_exception();
        }
    };
    const auto t0 = Clock::now();
    {
        std::vector<std::thread> pool;
        for (int w = 1; w < workers; ++w) pool.emplace_back(work, w);
        work(0);
        for (auto &t : pool) t.join();
    }
    device_ms += std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
    for (const auto &e : errors)
        if (e) std::rethrow_exception(e);
    for (const auto &J : jobs) {
        const int i = J->i;
        std::copy(J->xo.begin(), J->xo.end(), xs.begin() + st.col_starts[i]);
        std::copy(J->yo.begin(), J->yo.end(), ys.begin() + st.row_starts[i]);
        const Index before = states[i].iterations;
        states[i] = J->state;
        states[i].iterations += before;
        timed = timed || J->timed;
    }
}

// Runs one device batch over the stacked instances (FP64 or mixed),
// optionally warm-started from scaled iterates; returns per-instance states
// and scaled iterates.  Under the adaptive and Halpern policies a batch of
// at least kTailMinBatch instances finishes its last running instances (at
// most kTailInstances, at most a quarter of the batch) on single-LP solvers.
void run_batch(const std::vector<const Model *> &models, const std::vector<Prepared> &prep, const Stacked &st,
               const SolverOptions &options, int precision, double rel_tol, Index iteration_limit,
               const std::vector<double> *x0, const std::vector<double> *y0,
               const std::vector<double> *weights_in, const std::vector<int64_t> *limits,
               Clock::time_point started, std::vector<vk_pdhg_state> &states, std::vector<double> &xs,
               std::vector<double> &ys, double &device_ms, bool &timed) {
    const int count = static_cast<int>(st.steps.size());
    const int64_t m = st.row_starts[count], n = st.col_starts[count];
    const Scaled &sc = st.sc;
    vk_pdhg_params p{};
    p.policy = options.pdhg_policy == PdhgPolicy::Adaptive  ? VK_PDHG_POLICY_ADAPTIVE
               : options.pdhg_policy == PdhgPolicy::Halpern ? VK_PDHG_POLICY_HALPERN
                                                            : VK_PDHG_POLICY_REFERENCE;
    p.precision = precision;
    p.primal_weight_gain = options.pdhg_primal_weight_gain;
    p.primal_weight_integral = options.pdhg_primal_weight_integral;
    p.iteration_limit = iteration_limit;
    p.metric_tolerance = std::max(options.primal_tolerance, options.dual_tolerance);
    p.relative_kkt_tolerance = rel_tol;
    p.restart_period = options.restart_period;
    p.tile_geometry = st.tile_geometry;
    const std::vector<double> &weights = weights_in ? *weights_in : st.weights;
    DeviceBatch batch;
    const double *rl = m ? sc.rl.data() : nullptr, *ru = m ? sc.ru.data() : nullptr, *r = m ? sc.r.data() : nullptr;
    const double *xs0 = x0 ? x0->data() : nullptr, *ys0 = y0 && !y0->empty() ? y0->data() : nullptr;
    if (st.shared)
        check_vk(vk_hip_batch_create_shared(count, st.A, sc.c.data(), sc.l.data(), sc.u.data(), rl, ru, r, sc.s.data(),
                                            xs0, ys0, st.steps.data(), weights.data(), st.nb.data(), st.nc.data(),
                                            limits ? limits->data() : nullptr, &p, &batch.h),
                 "vk_hip_batch_create_shared");
    else
        check_vk(vk_hip_batch_create_csr(count, st.row_starts.data(), st.col_starts.data(), st.A, sc.c.data(),
                                         sc.l.data(), sc.u.data(), rl, ru, r, sc.s.data(), xs0, ys0, st.steps.data(),
                                         weights.data(), st.nb.data(), st.nc.data(),
                                         limits ? limits->data() : nullptr, &p, &batch.h),
                 "vk_hip_batch_create_csr");
    // Launch granularity: 256 iterations per launch keeps each kernel short
    // and lets the host honour the time limit between launches.
    constexpr int64_t kPerLaunch = 256;
    const int64_t launches = (iteration_limit + kPerLaunch - 1) / kPerLaunch + 1;
    timed = false;
    device_ms = 0.0;
    int running = count;
    // The reference policy stays in the batch: its instances must repeat
    // their own solves bitwise.  Small batches stay too (no idle GPU).
    const bool tail = options.pdhg_policy != PdhgPolicy::Reference && count >= kTailMinBatch;
    const int tail_at = std::min(kTailInstances, count / 4);
    for (int64_t launch = 0; launch < launches && running > 0; ++launch) {
        if (timed_out(started, options)) {
            timed = true;
            break;
        }
        check_vk(vk_hip_batch_run(batch.h, kPerLaunch, 1, &running), "vk_hip_batch_run");
        double ms = 0.0;
        check_vk(vk_hip_batch_last_run_ms(batch.h, &ms), "vk_hip_batch_last_run_ms");
        device_ms += ms;
        if (tail && running <= tail_at) break;
    }
    states.resize(static_cast<size_t>(count));
    for (int i = 0; i < count; ++i) check_vk(vk_hip_batch_read_state(batch.h, i, &states[i]), "vk_hip_batch_read_state");
    xs.assign(static_cast<size_t>(n), 0.0);
    ys.assign(static_cast<size_t>(m), 0.0);
    check_vk(vk_hip_batch_download(batch.h, xs.data(), ys.empty() ? nullptr : ys.data()), "vk_hip_batch_download");
    if (!tail || timed) return;
    std::vector<int> left;
    for (int i = 0; i < count; ++i)
        if (states[i].status == VK_PDHG_RUNNING) left.push_back(i);
    if (!left.empty())
        finish_tail(models, prep, st, left, options, precision, rel_tol, limits, iteration_limit, started, states, xs,
                    ys, device_ms, timed);
}

// The phases of a stacked batch (FP32 then FP64 for mixed precision) and
// each instance's postsolve, into results[prep[i].index].
void run_stacked(const std::vector<const Model *> &models, const std::vector<Prepared> &prep, const Stacked &st,
                 const SolverOptions &options, Clock::time_point started, std::vector<SolveResult> &results) {
    const std::vector<double> *warm0 = st.warm.empty() ? nullptr : &st.warm;

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
        // The batched FP32 phase stores plain FP32 iterates (no recentring).
        const double phase_tol = fp32_phase_tolerance(options, false);
        run_batch(models, prep, st, options, VK_PREC_MIXED, phase_tol, options.iteration_limit, warm0, nullptr, nullptr,
                  nullptr, started, states, xs, ys, ms, timed);
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
            run_batch(models, prep, st, options, VK_PREC_FP64, options.relative_kkt_tolerance, max_remaining, &x32, &y32,
                      options.pdhg_policy != PdhgPolicy::Reference ? &w : nullptr, &remaining, started, states, xs,
                      ys, ms, timed);
            total_ms += ms;
            for (std::size_t i = 0; i < prep.size(); ++i) used[i] += states[i].iterations;
        }
    } else {
        run_batch(models, prep, st, options, VK_PREC_FP64, options.relative_kkt_tolerance, options.iteration_limit, warm0,
                  nullptr, nullptr, nullptr, started, states, xs, ys, ms, timed);
        total_ms += ms;
        for (std::size_t i = 0; i < prep.size(); ++i) used[i] = states[i].iterations;
    }
    // Unscaled iterates of all instances, then the postsolve products of
    // the stacked matrix on the device (each instance's rows and columns
    // see only its own entries).
    const Scaled &sc = st.sc;
    std::vector<Scalar> xall(xs.size()), yall(ys.size()), c_all;
    for (std::size_t j = 0; j < xs.size(); ++j) xall[j] = xs[j] * sc.s[j];
    for (std::size_t i = 0; i < prep.size(); ++i) {
        const Model &model = *models[prep[i].index];
        c_all.insert(c_all.end(), model.linear_objective.begin(), model.linear_objective.end());
        const auto r0 = static_cast<std::size_t>(st.row_starts[i]), r1 = static_cast<std::size_t>(st.row_starts[i + 1]);
        const std::vector<double> &ysrc = infeasible32[i] ? ys32 : ys;
        for (std::size_t r = r0; r < r1; ++r) yall[r] = ysrc[r] * sc.r[r];
    }
    Products all;
    all.ax.resize(yall.size());
    all.g.resize(xall.size());
    if (st.shared) {
        // One matrix: one product pair per node.
        for (std::size_t i = 0; i < prep.size(); ++i) {
            const auto r0 = static_cast<std::size_t>(st.row_starts[i]), c0 = static_cast<std::size_t>(st.col_starts[i]);
            check_vk(vk_hip_csr_postsolve_products(st.A, xall.data() + c0, yall.data() + r0, c_all.data() + c0,
                                                   all.ax.data() + r0, all.g.data() + c0),
                     "vk_hip_csr_postsolve_products");
        }
    } else {
        check_vk(vk_hip_csr_postsolve_products(st.A, xall.data(), yall.data(), c_all.data(), all.ax.data(),
                                               all.g.data()),
                 "vk_hip_csr_postsolve_products");
    }
    for (std::size_t i = 0; i < prep.size(); ++i) {
        const Model &model = *models[prep[i].index];
        const auto row0 = st.row_starts[i], col0 = st.col_starts[i];
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
        const std::vector<Scalar> x(xall.begin() + col0, xall.begin() + col0 + n);
        const std::vector<Scalar> y(yall.begin() + row0, yall.begin() + row0 + m);
        Products pr;
        pr.ax.assign(all.ax.begin() + row0, all.ax.begin() + row0 + m);
        pr.g.assign(all.g.begin() + col0, all.g.begin() + col0 + n);
        postsolve(model, prep[i].bounds, options, x, y, pr, options.relative_kkt_tolerance > 0.0, res);
        if (infeasible) report_infeasible(model, prep[i].bounds, y, res);
        if (!st.tiles_note.empty()) res.message = res.message.empty() ? st.tiles_note : res.message + "; " + st.tiles_note;
        res.solve_seconds = std::chrono::duration<Scalar>(Clock::now() - started).count();
        results[prep[i].index] = std::move(res);
    }
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
        prep.push_back(std::move(pr));
    }
    if (prep.empty()) return results;
    Stacked st;
    stack(models, prep, options, st);
    // The stacked matrix has every instance's rows: the geometry tuned on it
    // is the one the instances' tiles are cut with.
    st.tile_geometry = apply_tile_geometry(st.A, options);
    st.tiles_note = tile_geometry_note(st.tile_geometry, options.spmv_tile_geometry < 0);
    run_stacked(models, prep, st, options, started, results);
    return results;
}

std::vector<SolveResult> solve_lp_nodes(const Model &model, const std::vector<std::vector<Scalar>> &lower,
                                        const std::vector<std::vector<Scalar>> &upper,
                                        const std::vector<std::vector<Scalar>> &warm_start,
                                        const SolverOptions &options) {
    detail::NodeCache cache;
    return detail::solve_lp_nodes_cached(model, lower, upper, warm_start, options, cache);
}

std::vector<SolveResult> detail::solve_lp_nodes_cached(const Model &model,
                                                       const std::vector<std::vector<Scalar>> &lower,
                                                       const std::vector<std::vector<Scalar>> &upper,
                                                       const std::vector<std::vector<Scalar>> &warm_start,
                                                       const SolverOptions &options, NodeCache &cache) {
    const auto started = Clock::now();
    const std::size_t count = std::max(lower.size(), upper.size());
    std::vector<SolveResult> results(count);
    if (count == 0) return results;
    auto node_vec = [](const std::vector<std::vector<Scalar>> &v, std::size_t i) -> std::span<const Scalar> {
        return i < v.size() ? std::span<const Scalar>(v[i]) : std::span<const Scalar>();
    };
    // solve_lp's entry checks (pdhg.cpp:269-293): the model's once, each
    // node's bounds and warm start on their own.
    std::string model_error;
    SolveStatus model_status = SolveStatus::NotSolved;
    if (const auto errors = model.validate(); !errors.empty()) {
        model_status = SolveStatus::InvalidModel;
        model_error = errors.front();
    } else if (model.problem_class() != ProblemClass::LP && model.problem_class() != ProblemClass::MILP) {
        model_status = SolveStatus::UnsupportedModel;
        model_error = "model class does not match the selected continuous solver";
    } else if (model.num_variables() == 0) {
        model_status = SolveStatus::InvalidModel;
        model_error = "LP requires at least one variable";
    } else if (!valid_options(options)) {
        model_status = SolveStatus::InvalidModel;
        model_error = "solver limits and tolerances must be non-negative";
    }
    std::vector<Prepared> prep;
    std::vector<std::vector<Scalar>> warm(count);
    for (std::size_t i = 0; i < count; ++i) {
        if (model_status != SolveStatus::NotSolved) {
            results[i] = status_result(model_status, model_error);
            continue;
        }
        Prepared pr;
        pr.index = i;
        std::string message;
        if (!resolve_bounds(model, node_vec(lower, i), node_vec(upper, i), pr.bounds, message)) {
            results[i] = status_result(SolveStatus::InvalidModel, message);
            continue;
        }
        const auto w = node_vec(warm_start, i);
        bool ok = w.empty() || static_cast<Index>(w.size()) == model.num_variables();
        for (const Scalar v : w) ok = ok && std::isfinite(v);
        if (!ok) {
            results[i] = status_result(SolveStatus::InvalidModel, "warm start must be empty or finite and sized to the model");
            continue;
        }
        warm[i].assign(w.begin(), w.end());
        prep.push_back(std::move(pr));
    }
    if (prep.empty()) return results;
    Stacked st;
    stack_shared(model, prep, warm, options, st, cache);
    const std::vector<const Model *> models(count, &model);
    run_stacked(models, prep, st, options, started, results);
    return results;
}

}  // namespace vikalp::solver
