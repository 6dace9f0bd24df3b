#pragma once
// Solver options.  The first block is the reference contract
// (include/vikalp/contracts/SolverOptions.hpp:7-28) field for field, with
// BackendPreference::HIP added after CUDA; the HIP-path options are appended
// after it with defaults, so existing aggregate initialisers keep compiling
// and keep their meaning.

#include "vikalp/contracts/Model.hpp"

namespace vikalp {

enum class BackendPreference { Auto, CPU, CUDA, HIP };
enum class ContinuousMethod { Auto, PDHG, RevisedSimplex, InteriorPoint };

// Step-size / restart rule of the PDHG loop.
enum class PdhgPolicy {
    Reference,   // the reference rule, src/solver/lp/pdhg.cpp:75-150
    Adaptive,    // restarted PDHG with averaging and a primal weight
    Halpern      // restarted Halpern PDHG with reflection and a primal weight
};

// Arithmetic of the device iteration.
enum class PdhgPrecision {
    FP64,        // A, iterates and vectors in FP64 (the drop-in)
    Mixed        // FP32 A and iterates, then a warm-started FP64 phase
};

struct SolverOptions {
    BackendPreference backend = BackendPreference::Auto;
    ContinuousMethod continuous_method = ContinuousMethod::Auto;
    Index iteration_limit = 100'000;
    Index node_limit = 100'000;
    Scalar time_limit_seconds = Model::infinity();
    // Feasibility and KKT residual limits in original, unscaled coordinates.
    Scalar primal_tolerance = 1e-7;
    Scalar dual_tolerance = 1e-7;
    Scalar integrality_tolerance = 1e-6;
    Scalar absolute_gap_tolerance = 1e-6;
    Scalar relative_gap_tolerance = 1e-4;
    bool deterministic = true;

    // ---- HIP PDHG path -------------------------------------------------
    PdhgPolicy pdhg_policy = PdhgPolicy::Reference;
    PdhgPrecision pdhg_precision = PdhgPrecision::FP64;
    // > 0: stop when the relative KKT error (original coordinates) is at or
    // below this, instead of the reference's scaled infinity-norm test.
    Scalar relative_kkt_tolerance = 0.0;
    // Mixed precision: relative KKT target of the FP32 phase.  0 = automatic:
    // the final target when the FP32 phase is recentred (below), else 1e-4.
    // A plain FP32 Halpern phase stalls between 6e-6 and 4e-5 and so stops
    // no later than 1e-5.
    Scalar mixed_phase_tolerance = 0.0;
    // Mixed precision under the Halpern policy: the FP32 phase stores its
    // iterates as corrections to an FP64 base point that moves to the
    // restart point at every restart (FP32 rounding is then relative to the
    // correction), so it converges like the FP64 iteration instead of
    // stalling; the FP64 phase then only certifies.  false = plain FP32
    // iterates.
    bool mixed_recentre = true;
    // Adaptive and Halpern policies: gains of the primal-weight update at
    // restarts, log w <- log w - gain e - integral sum(e) with the balance
    // error e = log w - log(|dy| / |dx|) of the moves since the last
    // restart.  gain 0 = 1/2 (PDLP's smoothing); integral 0 = none.
    Scalar pdhg_primal_weight_gain = 0.0;
    Scalar pdhg_primal_weight_integral = 0.0;
    // Adaptive and Halpern policies: iterations between restart checks.
    Index restart_period = 64;
    // Run whole iterations on the device (two fused kernels) instead of the
    // reference's op-by-op ExecutionBackend sequence.
    bool fused_iterations = true;
    // Capture iteration blocks in hipGraphs.
    bool use_graphs = true;
    // Layout of the A x SpMV: 0 = automatic (4 column blocks once x
    // outgrows 2 MiB), 1 = plain row tiles, 2..16 = that many column blocks.
    int spmv_col_blocks = 0;
    // Time the layouts on the device before the solve and use the fastest
    // (overrides spmv_col_blocks; for the panel layout, the lane count and
    // unroll depth, which never change results).
    bool spmv_autotune = false;
    // SpMV and epilogue kernels: 0 = automatic, 1 = fused, 2 = split into a
    // SpMV-only kernel and a row / column pass (identical results).
    int spmv_split_passes = 0;
    // MiB of matrix data kept in the 256 MB memory-side cache across
    // iterations (the rest is streamed): 0 = automatic, -1 = none.  Changes
    // load hints only, never results.
    int spmv_cache_budget_mb = 0;
    // Panel layout (the gathered vector staged in LDS panel by panel):
    // 0 = automatic (once m or n exceeds 2^18, if the matrix fits), 1 = never,
    // 2 = always.  Groups of columns of A x / A' y (0 = automatic; 1, 2, 4, 8)
    // set the partial-sum order; lanes per workgroup (0 = automatic, 256 or
    // 512) and KB per staged panel (0 = automatic, 48, 24 or 16; 256 lanes
    // only with 48) never change results.
    int spmv_panels = 0;
    int spmv_panel_groups = 0;
    int spmv_panel_groups_t = 0;
    int spmv_panel_threads = 0;
    int spmv_panel_kb = 0;
    // Row-tile geometry of the tile kernels (the fused path's matrix and the
    // batched solver's instances): 0 = 512 nonzeros x 64 rows per tile, rows
    // over 512 alone (the default); 1 = 256 x 64, over 256 alone; 2 = 512 x
    // 32; 3 = 512 x 64, rows over 64 alone; -1 = time the four on the matrix
    // at setup and use the fastest.  The geometry sets the order of each
    // row's sum, so results depend on it bitwise; when it is not the default
    // it is named in SolveResult::message.
    int spmv_tile_geometry = 0;
};

}  // namespace vikalp
