#pragma once
// Solve result: the reference contract (include/vikalp/contracts/SolveResult.hpp:11-61)
// field for field, plus relative-KKT and device-time fields appended at the end.

#include "vikalp/contracts/Model.hpp"

#include <limits>
#include <string>
#include <vector>

namespace vikalp {

enum class SolveStatus {
    NotSolved,
    Optimal,
    LocallyOptimal,
    Feasible,
    Infeasible,
    Unbounded,
    IterationLimit,
    NodeLimit,
    TimeLimit,
    NumericalFailure,
    UnsupportedModel,
    InvalidModel
};

struct SolveResult {
    SolveStatus status = SolveStatus::NotSolved;
    // Original model coordinates.
    std::vector<Scalar> primal_solution;
    // Non-negative duals: stationarity is c + A'(upper - lower) + (var upper - var lower) = 0.
    std::vector<Scalar> constraint_lower_dual;
    std::vector<Scalar> constraint_upper_dual;
    std::vector<Scalar> variable_lower_dual;
    std::vector<Scalar> variable_upper_dual;
    std::vector<Scalar> primal_ray;
    std::vector<Scalar> dual_ray;
    Scalar objective_value = std::numeric_limits<Scalar>::quiet_NaN();
    Scalar primal_bound = std::numeric_limits<Scalar>::quiet_NaN();
    Scalar dual_bound = std::numeric_limits<Scalar>::quiet_NaN();
    // Infinity norms in original coordinates.
    Scalar primal_residual = Model::infinity();
    Scalar stationarity_residual = Model::infinity();
    Scalar complementarity_residual = Model::infinity();
    Scalar integrality_residual = Model::infinity();
    Scalar absolute_gap = Model::infinity();
    Scalar relative_gap = Model::infinity();
    Index iterations = 0;
    Index nodes = 0;
    Scalar solve_seconds = 0.0;
    std::string solver;
    std::string backend;
    bool verified = false;
    std::string message;

    // ---- appended: relative KKT of the returned point (original
    // coordinates, 2-norms; see the PDHG solver for the definition) and the
    // device time of the iteration phase.
    Scalar kkt_relative_primal = Model::infinity();
    Scalar kkt_relative_dual = Model::infinity();
    Scalar kkt_relative_gap = Model::infinity();
    Scalar kkt_relative = Model::infinity();
    Scalar kkt_primal_objective = std::numeric_limits<Scalar>::quiet_NaN();
    Scalar kkt_dual_objective = std::numeric_limits<Scalar>::quiet_NaN();
    Scalar device_seconds = 0.0;
    Index low_precision_iterations = 0;   // mixed precision: FP32-phase iterations
};

}  // namespace vikalp
