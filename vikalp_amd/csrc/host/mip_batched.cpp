// solver::solve_mip_batched: branch-and-bound whose node relaxations run in
// batches on one device-resident matrix (solve_lp_nodes).
//
// The reference's engine (src/solver/mip/BranchAndBound.cpp:222-465) takes
// one open node at a time and calls RelaxationOracle::solve for it, which
// builds a backend, scales and uploads the model for every node
// (Solver.cpp:37-51).  Here each round takes up to `batch` open nodes in the
// reference's queue order (BbQueue.hpp:24-34: lower relaxation bound first,
// then the deeper node, then the lower id), solves their relaxations in one
// call of the batched node solver (all nodes in the batched kernel, on a
// device matrix uploaded, scaled and sized for its step once per search),
// and then processes the results in that order with the reference's rules:
//   - a node whose bounds cross is infeasible (BranchAndBound.cpp:276-286);
//   - an infeasible relaxation is counted, a failed one ends the search with
//     its status, an unsolved or invalid one with NumericalFailure
//     (:298-324);
//   - the node bound is the relaxation's dual bound, else its objective when
//     optimal (:358-361), and updates the pseudocosts of the branching that
//     made the node (:362-373);
//   - a node whose bound is within the absolute gap of the incumbent is
//     pruned (:375-379);
//   - an integral relaxation is an incumbent candidate; otherwise the
//     rounding heuristic (:145-163) may give one, and the node branches on
//     the most fractional variable for the first 8 processed nodes, then by
//     pseudocost score among reliable variables (:403-405), into a down child
//     (upper bound floor) and an up child (lower bound ceil), each warm-started
//     from the node's relaxation (:414-444);
//   - the search stops as optimal once every bound so far was certified and
//     the incumbent is within the absolute or relative gap of the best open
//     bound (:447-457).
// With batch = 1 the node sequence is the reference's.  With larger batches
// nodes that the serial order would have pruned by an incumbent found
// earlier in the same round are solved anyway (and pruned when processed),
// the price of solving them together.  Linear objectives only: a quadratic
// objective (MIQP) is refused.

#include "lp_internal.hpp"

#include "vikalp/solver/continuous.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <limits>
#include <memory>
#include <queue>
#include <string>
#include <utility>
#include <vector>

namespace vikalp::solver {
namespace {

using detail::Clock;

constexpr Scalar kInf = std::numeric_limits<Scalar>::infinity();

// Bound changes of a node relative to the model, in branching order.
using Overrides = std::vector<std::pair<Index, Scalar>>;

struct OpenNode {
    std::int64_t id = 0;
    std::int64_t parent = -1;
    Index depth = 0;
    Overrides lower, upper;
    Scalar bound = -kInf;          // the parent's relaxation bound
    std::vector<Scalar> warm;      // the parent's relaxation solution
    Index branch_var = -1;
    bool down = true;
};

// std::priority_queue keeps its largest element on top: `after(a, b)` says
// that a is taken after b.
struct After {
    bool operator()(const OpenNode &a, const OpenNode &b) const {
        if (a.bound != b.bound) return a.bound > b.bound;
        if (a.depth != b.depth) return a.depth < b.depth;
        return a.id > b.id;
    }
};

class OpenSet {
public:
    void push(OpenNode n) { heap_.push(std::move(n)); }
    OpenNode pop() {
        OpenNode n = heap_.top();
        heap_.pop();
        return n;
    }
    bool empty() const { return heap_.empty(); }
    Scalar best_bound() const { return heap_.empty() ? kInf : heap_.top().bound; }

private:
    std::priority_queue<OpenNode, std::vector<OpenNode>, After> heap_;
};

// Pseudocosts (Pseudocosts.hpp:23-92): per variable and direction the mean
// bound change per unit of fractional distance, 1 when unobserved; reliable
// after 4 observations each way; score (1 - mu) min(d, u) + mu max(d, u)
// with d, u the costs times the distances and mu = 1/6.
class Pseudocosts {
public:
    explicit Pseudocosts(std::size_t n) : sum_(2 * n, 0.0), count_(2 * n, 0) {}
    void record(Index j, bool down, Scalar delta, Scalar distance) {
        if (distance <= 1e-12) return;
        const std::size_t k = slot(j, down);
        sum_[k] += delta / distance;
        ++count_[k];
    }
    bool reliable(Index j) const { return count_[slot(j, true)] >= 4 && count_[slot(j, false)] >= 4; }
    Scalar score(Index j, Scalar down_distance, Scalar up_distance) const {
        const Scalar d = cost(j, true) * down_distance, u = cost(j, false) * up_distance;
        constexpr Scalar mu = 1.0 / 6.0;
        return (1.0 - mu) * std::min(d, u) + mu * std::max(d, u);
    }

private:
    static std::size_t slot(Index j, bool down) { return 2 * static_cast<std::size_t>(j) + (down ? 0 : 1); }
    Scalar cost(Index j, bool down) const {
        const std::size_t k = slot(j, down);
        return count_[k] > 0 ? sum_[k] / count_[k] : 1.0;
    }
    std::vector<Scalar> sum_;
    std::vector<int> count_;
};

Scalar fractionality(Scalar v) { return std::abs(v - std::round(v)); }

bool integral(const Model &model, const std::vector<Scalar> &x, Scalar tol) {
    for (std::size_t j = 0; j < x.size(); ++j)
        if (model.variable_types[j] != VariableType::Continuous && fractionality(x[j]) > tol) return false;
    return true;
}

// offset + c'x + 1/2 x'Qx over Q's stored entries (BranchAndBound.cpp:47-62).
Scalar objective(const Model &model, const std::vector<Scalar> &x) {
    Scalar v = model.objective_offset;
    for (std::size_t j = 0; j < x.size(); ++j) v += model.linear_objective[j] * x[j];
    const auto &q = model.quadratic_objective;
    for (Index r = 0; r < q.pattern.rows; ++r)
        for (Index k = q.pattern.row_offsets[r]; k < q.pattern.row_offsets[r + 1]; ++k)
            v += 0.5 * q.values[static_cast<std::size_t>(k)] * x[static_cast<std::size_t>(r)] *
                 x[static_cast<std::size_t>(q.pattern.column_indices[k])];
    return v;
}

// x finite and within [lo, hi] (tol_box) and every row within its bounds
// (tol_rows).
bool feasible(const Model &model, const std::vector<Scalar> &x, const std::vector<Scalar> &lo,
              const std::vector<Scalar> &hi, Scalar tol_box, Scalar tol_rows) {
    for (std::size_t j = 0; j < x.size(); ++j)
        if (!std::isfinite(x[j]) || x[j] < lo[j] - tol_box || x[j] > hi[j] + tol_box) return false;
    const auto &A = model.constraint_matrix;
    for (Index i = 0; i < A.pattern.rows; ++i) {
        Scalar s = 0.0;
        for (Index k = A.pattern.row_offsets[i]; k < A.pattern.row_offsets[i + 1]; ++k)
            s += A.values[static_cast<std::size_t>(k)] * x[static_cast<std::size_t>(A.pattern.column_indices[k])];
        const auto r = static_cast<std::size_t>(i);
        if (s < model.constraint_lower[r] - tol_rows || s > model.constraint_upper[r] + tol_rows) return false;
    }
    return true;
}

// Branching variable among the fractional integer variables that a child can
// take (floor >= lo or ceil <= hi): most fractional, ties to the lower index;
// with pseudocosts, the best score among reliable ones, else most fractional.
Index branch_variable(const Model &model, const std::vector<Scalar> &x, const std::vector<Scalar> &lo,
                      const std::vector<Scalar> &hi, Scalar tol, const Pseudocosts *pc) {
    Index pick = -1;
    Scalar best = -1.0;
    for (int pass = pc ? 0 : 1; pass < 2 && pick < 0; ++pass) {
        for (std::size_t j = 0; j < x.size(); ++j) {
            if (model.variable_types[j] == VariableType::Continuous || fractionality(x[j]) <= tol) continue;
            const Scalar f = std::floor(x[j]), c = std::ceil(x[j]);
            if (f < lo[j] && c > hi[j]) continue;
            Scalar score;
            if (pass == 0) {
                if (!pc->reliable(static_cast<Index>(j))) continue;
                score = pc->score(static_cast<Index>(j), x[j] - f, c - x[j]);
            } else {
                score = 0.5 - std::abs((x[j] - f) - 0.5);
            }
            if (score > best) {   // strictly: ties keep the lower index
                best = score;
                pick = static_cast<Index>(j);
            }
        }
    }
    return pick;
}

Scalar abs_gap(Scalar primal, Scalar dual) {
    return (std::isfinite(primal) && std::isfinite(dual)) ? std::max<Scalar>(0.0, primal - dual) : kInf;
}

Scalar rel_gap(Scalar primal, Scalar dual) {
    const Scalar a = abs_gap(primal, dual);
    return std::isfinite(a) ? a / std::max<Scalar>(1.0, std::abs(primal)) : kInf;
}

void apply(const Overrides &o, std::vector<Scalar> &bounds, bool lower) {
    for (const auto &[j, v] : o) {
        auto &b = bounds[static_cast<std::size_t>(j)];
        b = lower ? std::max(b, v) : std::min(b, v);
    }
}

}  // namespace

SolveResult solve_mip_batched(const Model &model, const SolverOptions &options, Index batch, MipStats *stats) {
    SolveResult result;
    result.solver = "vikalp-mip-bnb-batched";
    result.backend = "hip";
    MipStats st;
    if (model.nonlinear) {
        result.status = SolveStatus::UnsupportedModel;
        result.message = "MIP branch-and-bound requires a linear or quadratic model";
        return result;
    }
    if (!model.has_integer_variables()) {
        result.status = SolveStatus::UnsupportedModel;
        result.message = "Model has no integer variables; use a continuous solver";
        return result;
    }
    if (const auto errors = model.validate(); !errors.empty()) {
        result.status = SolveStatus::InvalidModel;
        result.message = errors.front();
        return result;
    }
    if (model.has_quadratic_objective()) {
        result.status = SolveStatus::UnsupportedModel;
        result.message = "batched branch-and-bound solves LP relaxations: the objective must be linear";
        return result;
    }
    if (batch < 1) {
        result.status = SolveStatus::InvalidModel;
        result.message = "batch must be at least 1";
        return result;
    }
    const auto started = Clock::now();
    const std::size_t n = static_cast<std::size_t>(model.num_variables());
    const Scalar int_tol = options.integrality_tolerance;
    // A relaxation stopped on the relative-KKT test has row violations of
    // 2-norm at most tol (1 + ||bhat||_2), which bounds each one: rows are
    // checked against that, and never tighter than the primal tolerance.
    Scalar row_tol = options.primal_tolerance;
    if (options.relative_kkt_tolerance > 0.0)
        row_tol = std::max(row_tol, options.relative_kkt_tolerance *
                                        (1.0 + detail::row_bound_norm(model.constraint_lower, model.constraint_upper)));

    // One device matrix, scaling and step for every round.
    detail::NodeCache cache;
    OpenSet open;
    Pseudocosts pc(n);
    std::int64_t next_id = 0;
    Scalar incumbent = kInf;
    std::vector<Scalar> incumbent_x;
    bool certified = true;
    SolveStatus termination = SolveStatus::NotSolved;
    {
        OpenNode root;
        root.id = next_id++;
        open.push(std::move(root));
    }
    auto take = [&](Scalar objective, std::vector<Scalar> x) {
        if (std::isfinite(objective) && objective < incumbent) {
            incumbent = objective;
            incumbent_x = std::move(x);
            ++st.incumbent_updates;
        }
    };

    while (!open.empty() && termination == SolveStatus::NotSolved) {
        if (std::chrono::duration<Scalar>(Clock::now() - started).count() >= options.time_limit_seconds) {
            termination = SolveStatus::TimeLimit;
            break;
        }
        if (st.nodes_processed >= options.node_limit) {
            termination = SolveStatus::NodeLimit;
            break;
        }
        // The round: up to `batch` nodes in queue order, within the node limit.
        std::vector<OpenNode> round;
        std::vector<std::vector<Scalar>> lower, upper, warm;
        std::vector<int> slot;   // index into lower/upper/warm, -1: bounds cross
        const Index room = std::min<Index>(batch, options.node_limit - st.nodes_processed);
        while (!open.empty() && static_cast<Index>(round.size()) < room) {
            OpenNode node = open.pop();
            std::vector<Scalar> lo = model.variable_lower, hi = model.variable_upper;
            apply(node.lower, lo, true);
            apply(node.upper, hi, false);
            bool cross = false;
            for (std::size_t j = 0; j < n && !cross; ++j) cross = lo[j] > hi[j] + options.primal_tolerance;
            if (cross) {
                slot.push_back(-1);
            } else {
                slot.push_back(static_cast<int>(lower.size()));
                lower.push_back(std::move(lo));
                upper.push_back(std::move(hi));
                warm.push_back(node.warm);
            }
            round.push_back(std::move(node));
        }
        std::vector<SolveResult> relax;
        if (!lower.empty()) {
            relax = detail::solve_lp_nodes_cached(model, lower, upper, warm, options, cache);
            ++st.rounds;
            st.relaxations += static_cast<std::int64_t>(lower.size());
            st.max_batch = std::max<std::int64_t>(st.max_batch, static_cast<std::int64_t>(lower.size()));
            for (const auto &r : relax) st.relaxation_iterations += r.iterations;
            result.device_seconds += relax.front().device_seconds;   // one batch: shared device time
        }
        for (std::size_t t = 0; t < round.size(); ++t) {
            if (termination != SolveStatus::NotSolved) {
                open.push(std::move(round[t]));   // not processed: stays open
                continue;
            }
            OpenNode &node = round[t];
            ++st.nodes_processed;
            if (slot[t] < 0) {
                ++st.nodes_infeasible;
                continue;
            }
            const auto s = static_cast<std::size_t>(slot[t]);
            const SolveResult &r = relax[s];
            const std::vector<Scalar> &lo = lower[s], &hi = upper[s];
            if (r.status == SolveStatus::Infeasible) {
                ++st.nodes_infeasible;
                continue;
            }
            if (r.status == SolveStatus::Unbounded || r.status == SolveStatus::UnsupportedModel ||
                r.status == SolveStatus::InvalidModel || r.status == SolveStatus::NumericalFailure ||
                r.status == SolveStatus::NotSolved) {
                termination = r.status;
                result.message = "relaxation failed at node " + std::to_string(node.id) +
                                 (r.message.empty() ? "" : ": " + r.message);
                continue;
            }
            if (r.status != SolveStatus::Optimal && r.status != SolveStatus::LocallyOptimal &&
                r.status != SolveStatus::Feasible) {
                termination = r.status;
                result.message = "relaxation not solved at node " + std::to_string(node.id);
                continue;
            }
            bool usable = r.primal_solution.size() == n && std::isfinite(r.objective_value);
            for (std::size_t j = 0; usable && j < n; ++j) usable = std::isfinite(r.primal_solution[j]);
            if (!usable || !feasible(model, r.primal_solution, lo, hi, options.primal_tolerance, row_tol)) {
                termination = SolveStatus::NumericalFailure;
                result.message = "relaxation returned an invalid primal solution at node " + std::to_string(node.id);
                continue;
            }
            certified = certified && (r.status == SolveStatus::Optimal || std::isfinite(r.dual_bound));
            const Scalar bound = std::isfinite(r.dual_bound)
                                     ? r.dual_bound
                                     : (r.status == SolveStatus::Optimal ? r.objective_value : -kInf);
            if (node.parent >= 0 && node.branch_var >= 0 && node.warm.size() == n && std::isfinite(node.bound)) {
                const Scalar v = node.warm[static_cast<std::size_t>(node.branch_var)];
                pc.record(node.branch_var, node.down, std::max<Scalar>(0.0, bound - node.bound),
                          node.down ? v - std::floor(v) : std::ceil(v) - v);
            }
            if (std::isfinite(incumbent) && std::isfinite(bound) && bound >= incumbent - options.absolute_gap_tolerance) {
                ++st.nodes_pruned;
                continue;
            }
            const std::vector<Scalar> &x = r.primal_solution;
            if (integral(model, x, int_tol)) {
                ++st.nodes_integral;
                take(objective(model, x), x);
            } else {
                // Rounding heuristic: integers rounded, everything clamped to
                // the node's bounds, kept if integral and feasible.
                std::vector<Scalar> rx(x);
                for (std::size_t j = 0; j < n; ++j) {
                    if (model.variable_types[j] != VariableType::Continuous) rx[j] = std::round(rx[j]);
                    rx[j] = std::clamp(rx[j], lo[j], hi[j]);
                }
                if (integral(model, rx, int_tol) &&
                    feasible(model, rx, lo, hi, options.primal_tolerance, options.primal_tolerance))
                    take(objective(model, rx), std::move(rx));
                const Index j = branch_variable(model, x, lo, hi, int_tol, st.nodes_processed <= 8 ? nullptr : &pc);
                if (j < 0) {
                    termination = SolveStatus::NumericalFailure;
                    result.message = "non-integral relaxation without a branching variable";
                    continue;
                }
                const auto jj = static_cast<std::size_t>(j);
                const Scalar f = std::floor(x[jj]), c = std::ceil(x[jj]);
                for (const bool down : {true, false}) {
                    if (down ? f < lo[jj] : c > hi[jj]) continue;
                    OpenNode child;
                    child.id = next_id++;
                    child.parent = node.id;
                    child.depth = node.depth + 1;
                    child.lower = node.lower;
                    child.upper = node.upper;
                    (down ? child.upper : child.lower).push_back({j, down ? f : c});
                    child.bound = bound;
                    child.warm = x;
                    child.branch_var = j;
                    child.down = down;
                    open.push(std::move(child));
                }
            }
            // Open bounds include the round's nodes not yet processed.
            Scalar best_open = open.best_bound();
            for (std::size_t u = t + 1; u < round.size(); ++u) best_open = std::min(best_open, round[u].bound);
            if (open.empty() && t + 1 == round.size()) best_open = incumbent;
            if (certified && std::isfinite(incumbent) && std::isfinite(bound) &&
                (abs_gap(incumbent, best_open) <= options.absolute_gap_tolerance ||
                 rel_gap(incumbent, best_open) <= options.relative_gap_tolerance))
                termination = SolveStatus::Optimal;
        }
    }

    result.solve_seconds = std::chrono::duration<Scalar>(Clock::now() - started).count();
    result.nodes = st.nodes_processed;
    result.iterations = st.nodes_processed;
    if (!incumbent_x.empty()) {
        result.primal_solution = incumbent_x;
        result.objective_value = incumbent;
        result.primal_bound = incumbent;
        Scalar worst = 0.0;
        for (std::size_t j = 0; j < n; ++j)
            if (model.variable_types[j] != VariableType::Continuous) worst = std::max(worst, fractionality(incumbent_x[j]));
        result.integrality_residual = worst;
    }
    result.dual_bound = open.empty() ? (std::isfinite(incumbent) ? incumbent : kInf) : open.best_bound();
    const Scalar primal = std::isfinite(incumbent) ? incumbent : kInf;
    result.absolute_gap = abs_gap(primal, result.dual_bound);
    result.relative_gap = rel_gap(primal, result.dual_bound);
    if (termination == SolveStatus::NotSolved) {
        if (open.empty() && incumbent_x.empty()) termination = SolveStatus::Infeasible;
        else if (open.empty() && certified) termination = SolveStatus::Optimal;
        else termination = SolveStatus::Feasible;
    }
    result.status = termination;
    if (termination == SolveStatus::Optimal &&
        (!certified || (result.absolute_gap > options.absolute_gap_tolerance &&
                        result.relative_gap > options.relative_gap_tolerance)))
        result.status = incumbent_x.empty() ? SolveStatus::Infeasible : SolveStatus::Feasible;
    if (stats) *stats = st;
    return result;
}

}  // namespace vikalp::solver
