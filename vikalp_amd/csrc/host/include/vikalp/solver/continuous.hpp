#pragma once
// Continuous LP entry point with the reference's signature
// (include/vikalp/solver/continuous.hpp:10-15).

#include "vikalp/contracts/ExecutionBackend.hpp"
#include "vikalp/contracts/SolveResult.hpp"
#include "vikalp/contracts/SolverOptions.hpp"

#include <cstdint>
#include <memory>
#include <span>
#include <vector>

namespace vikalp::solver {

// Restarted PDHG on `backend`.  On a HipBackend with
// options.fused_iterations the whole iteration runs on the device; on any
// other ExecutionBackend the reference's op-by-op sequence is issued.
SolveResult solve_lp(const Model &model, ExecutionBackend &backend,
                     const SolverOptions &options = {},
                     std::span<const Scalar> warm_start = {},
                     std::span<const Scalar> variable_lower_override = {},
                     std::span<const Scalar> variable_upper_override = {});

// PDQP (src/solver/qp/pdqp.cpp:303-331): the same loop with the gradient
// Q x + c + A'y and the step 0.99 / max(1, bound^2 + |Q|_inf).  On a
// HipBackend with options.fused_iterations the iteration runs on the device;
// otherwise the reference's op-by-op sequence is issued.  The reference
// policy runs in FP64; the Halpern policy (fused only) in FP64 or mixed
// precision, with per-column primal steps and the relative-KKT test.
SolveResult solve_qp(const Model &model, ExecutionBackend &backend,
                     const SolverOptions &options = {},
                     std::span<const Scalar> warm_start = {},
                     std::span<const Scalar> variable_lower_override = {},
                     std::span<const Scalar> variable_upper_override = {});

// Solves every LP of `models` together on the batched persistent HIP kernel
// (one launch runs many instances; each keeps its own step sizes, restart
// state and status).  Results are in the order of `models`, each with
// solve_lp's semantics; device_seconds is the batch's shared device time.
// Instances that fail the entry checks get their status without running.
std::vector<SolveResult> solve_lp_batch(const std::vector<const Model *> &models,
                                        const SolverOptions &options = {});

// Relaxations of one model that differ only in variable bounds (and warm
// start): the branch-and-bound node loop (BranchAndBound.cpp:296-297 calls
// RelaxationOracle::solve once per node, with a new backend each time,
// Solver.cpp:37-51) as one batch on one device-resident, once-scaled
// matrix.  Node i uses lower[i] / upper[i] (empty: the model's bounds) and
// warm_start[i] (empty: cold start), as solve_lp would; the number of nodes
// is the longer of lower and upper.  Integer variable types are relaxed.
// Results in node order with solve_lp's semantics.
std::vector<SolveResult> solve_lp_nodes(const Model &model, const std::vector<std::vector<Scalar>> &lower,
                                        const std::vector<std::vector<Scalar>> &upper,
                                        const std::vector<std::vector<Scalar>> &warm_start = {},
                                        const SolverOptions &options = {});

// Counters of a batched branch-and-bound search.
struct MipStats {
    std::int64_t nodes_processed = 0, nodes_pruned = 0, nodes_integral = 0, nodes_infeasible = 0;
    std::int64_t incumbent_updates = 0;
    std::int64_t rounds = 0;                  // solve_lp_nodes calls
    std::int64_t relaxations = 0;             // node relaxations solved
    std::int64_t relaxation_iterations = 0;   // PDHG iterations over all of them
    std::int64_t max_batch = 0;               // most relaxations in one round
};

// Branch-and-bound (BranchAndBound.cpp:222-465's rules: best-bound-first
// queue, pruning, integral and rounded incumbents, most-fractional then
// pseudocost branching, gap termination) whose node relaxations are solved
// `batch` at a time by solve_lp_nodes on one device matrix, instead of one
// RelaxationOracle::solve per node.  batch = 1 gives the reference's node
// sequence.  Options go to every relaxation.  Linear objectives only (an
// MIQP is UnsupportedModel).
SolveResult solve_mip_batched(const Model &model, const SolverOptions &options = {}, Index batch = 64,
                              MipStats *stats = nullptr);

// Mean device time (ms, HIP events on the solver's stream) of the fused dual
// and primal kernels and of a whole iteration, after `warm_iterations`
// iterations, on the HIP backend.  For measurement only.
struct KernelProfile {
    double dual_ms = 0.0, primal_ms = 0.0, iteration_ms = 0.0, step_size = 0.0;
    // Per kernel: the SpMV (or fused) kernel and the row / column pass of a
    // split layout (0 when fused).
    double dual_spmv_ms = 0.0, dual_pass_ms = 0.0, primal_spmv_ms = 0.0, primal_pass_ms = 0.0;
    int col_blocks = 1;   // layout the solver used
    int split_passes = 0;
    // Panel layout (zeros when not used): groups, lanes per workgroup, most
    // entries per lane and panel, the unrolled entries and KB per panel, of
    // A x then A' y.
    int panel[10] = {};
};
KernelProfile profile_lp_kernels(const Model &model, const SolverOptions &options,
                                 Index warm_iterations, Index iterations);

// K unrolled steps of the PDHG iteration (pdhg.cpp:92-113) with the
// reference's step tau = sigma = 0.99 / max(1, bound) (pdhg.cpp:75-78) held
// fixed and no restarts, from the projected zero start, as a differentiable
// map of the objective and the row bounds, on the HIP device.  forward()
// gives (x_K, y_K) in original coordinates (y in the reference's sign);
// backward() takes dL/dx_K and dL/dy_K and returns dL/dc, dL/d(row_lower)
// and dL/d(row_upper) through the K steps (L's own explicit dependence on c
// is the caller's).  The backward pass keeps (x, y) every checkpoint_every
// steps and recomputes each segment, so memory is O((K / N + N)(m + n)).
// Throws std::invalid_argument for an unusable model or counts.
class UnrolledPdhg {
public:
    UnrolledPdhg(const Model &model, Index iterations, Index checkpoint_every);
    ~UnrolledPdhg();
    UnrolledPdhg(const UnrolledPdhg &) = delete;
    UnrolledPdhg &operator=(const UnrolledPdhg &) = delete;

    void forward(std::span<Scalar> x, std::span<Scalar> y);
    // New objective and row bounds for the same A (a row keeps its kind:
    // finite / infinite bounds and equalities where they were).  The scaling
    // depends on A only, so the device matrix is reused.
    void set_data(std::span<const Scalar> objective, std::span<const Scalar> row_lower,
                  std::span<const Scalar> row_upper);
    void backward(std::span<const Scalar> grad_x, std::span<const Scalar> grad_y, std::span<Scalar> grad_c,
                  std::span<Scalar> grad_row_lower, std::span<Scalar> grad_row_upper);
    Scalar step_size() const;
    Index iterations() const;
    double last_device_ms() const;   // device time of the last forward / backward

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

}  // namespace vikalp::solver
