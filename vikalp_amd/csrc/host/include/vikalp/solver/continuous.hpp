#pragma once
// Continuous LP entry point with the reference's signature
// (include/vikalp/solver/continuous.hpp:10-15).

#include "vikalp/contracts/ExecutionBackend.hpp"
#include "vikalp/contracts/SolveResult.hpp"
#include "vikalp/contracts/SolverOptions.hpp"

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

// Solves every LP of `models` together on the batched persistent HIP kernel
// (one launch runs many instances; each keeps its own step sizes, restart
// state and status).  Results are in the order of `models`, each with
// solve_lp's semantics; device_seconds is the batch's shared device time.
// Instances that fail the entry checks get their status without running.
std::vector<SolveResult> solve_lp_batch(const std::vector<const Model *> &models,
                                        const SolverOptions &options = {});

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
};
KernelProfile profile_lp_kernels(const Model &model, const SolverOptions &options,
                                 Index warm_iterations, Index iterations);

}  // namespace vikalp::solver
