// vikalp::solve for the HIP library: the routing of src/solver/Solver.cpp:28-99
// with BackendPreference::HIP.  Exceptions below become NumericalFailure
// (Solver.cpp:81-86); the backend label is "hip" (Solver.cpp:79,96 use
// "cpu"/"cuda").  MILP models go to branch-and-bound as in Solver.cpp:130-134,
// here the batched engine (solve_mip_batched, rounds of kRouterBatch node
// relaxations); MIQP and MINLP are not served.

#include "vikalp/solver/Solver.hpp"

#include "vikalp/backend/HipBackend.hpp"
#include "vikalp/solver/continuous.hpp"

#include <algorithm>
#include <exception>
#include <span>

namespace vikalp {

// Node relaxations per branch-and-bound round when vikalp::solve routes a
// MILP (the batch size that measured fastest on the benchmark MILP).
constexpr Index kRouterBatch = 16;

SolveResult solve(const Model &model, const SolverOptions &options) {
    SolveResult result;
    const auto errors = model.validate();
    if (!errors.empty()) {
        result.status = SolveStatus::InvalidModel;
        result.message = errors.front();
        return result;
    }
    if (options.backend == BackendPreference::CPU || options.backend == BackendPreference::CUDA) {
        result.status = SolveStatus::UnsupportedModel;
        result.message = "this library carries the HIP backend only; CPU and CUDA are served by the reference";
        return result;
    }
    const ProblemClass pc = model.problem_class();
    if (pc == ProblemClass::MILP) {
        try {
            return solver::solve_mip_batched(model, options, kRouterBatch);
        } catch (const std::exception &error) {
            result.status = SolveStatus::NumericalFailure;
            result.message = error.what();
            result.backend = "hip";
            return result;
        }
    }
    if (pc != ProblemClass::LP && pc != ProblemClass::QP) {
        result.status = SolveStatus::UnsupportedModel;
        result.message = "the HIP path solves LP (PDHG), QP (PDQP) and MILP (batched branch-and-bound) models";
        return result;
    }
    try {
        auto backend = make_hip_backend();
        Model relaxation = model;
        std::fill(relaxation.variable_types.begin(), relaxation.variable_types.end(),
                  VariableType::Continuous);
        // Solver.cpp:59-69: LP -> solve_lp, QP -> solve_qp.
        result = pc == ProblemClass::LP ? solver::solve_lp(relaxation, *backend, options)
                                        : solver::solve_qp(relaxation, *backend, options);
        result.backend = "hip";
    } catch (const std::exception &error) {
        result = SolveResult{};
        result.status = SolveStatus::NumericalFailure;
        result.message = error.what();
        result.backend = "hip";
    }
    result.solver = "vikalp-continuous";
    return result;
}

}  // namespace vikalp
