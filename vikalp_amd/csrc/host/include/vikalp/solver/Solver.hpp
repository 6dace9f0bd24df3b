#pragma once
// Top-level entry point with the reference's signature
// (include/vikalp/solver/Solver.hpp:19-22, src/solver/Solver.cpp:92-99).
// This library carries the HIP backend only: BackendPreference::HIP and
// ::Auto run on the GPU; CPU and CUDA requests return UnsupportedModel (the
// reference's own CPU backend serves those).  LP is routed to solve_lp; other
// problem classes are outside this library and return UnsupportedModel.

#include "vikalp/contracts/Model.hpp"
#include "vikalp/contracts/SolveResult.hpp"
#include "vikalp/contracts/SolverOptions.hpp"

namespace vikalp {

[[nodiscard]] SolveResult solve(const Model &model, const SolverOptions &options = {});

}  // namespace vikalp
