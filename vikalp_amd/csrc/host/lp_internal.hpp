#pragma once
// Host helpers shared by the single-LP and the batched PDHG paths
// (definitions in pdhg_lp.cpp).  Each follows the reference function named
// beside it.

#include "vikalp/contracts/Model.hpp"
#include "vikalp/contracts/SolveResult.hpp"
#include "vikalp/contracts/SolverOptions.hpp"
#include "vikalp_hip.h"

#include <chrono>
#include <span>
#include <string>
#include <vector>

namespace vikalp::solver::detail {

using Clock = std::chrono::steady_clock;
constexpr int kRuizPasses = 10;
// Smallest relative KKT target of a plain (not recentred) Halpern FP32 phase
// (see solve_lp's mixed driver).
constexpr Scalar kHalpernFp32Floor = 1e-5;
// Relative KKT target of the FP32 phase of a mixed solve
// (SolverOptions::mixed_phase_tolerance; `recentred`: the phase runs the
// recentred FP32 iteration).
Scalar fp32_phase_tolerance(const SolverOptions &o, bool recentred);

struct Bounds {
    std::vector<Scalar> lower, upper;
};

// Scaled LP data (continuous_common.hpp:24-34): values, c, l, u, rl, ru and
// the row / column factors r, s (x = s * xhat, y = r * yhat).
struct Scaled {
    std::vector<Scalar> values, c, l, u, rl, ru, r, s;
};

SolveResult status_result(SolveStatus status, const std::string &message);
bool valid_options(const SolverOptions &o);                              // continuous_common.hpp:50-56
bool timed_out(Clock::time_point started, const SolverOptions &o);        // continuous_common.hpp:58-63
bool resolve_bounds(const Model &model, std::span<const Scalar> lo, std::span<const Scalar> hi,
                    Bounds &b, std::string &message);                     // continuous_common.hpp:80-110
Scaled scale(const Model &model, const Bounds &bounds);                   // continuous_common.hpp:133-205
void equilibrate(const CsrPattern &p, Scaled &sc, int ruiz);              // adaptive policy only
Scalar operator_bound(const CsrPattern &p, const std::vector<Scalar> &v); // continuous_common.hpp:207-225
Scalar row_bound_norm(const std::vector<Scalar> &lo, const std::vector<Scalar> &hi);
// Whether y (reference sign, original coordinates) proves the LP primal
// infeasible (Farkas); d0 = its dual objective without c.
bool farkas_certificate(const Model &model, const Bounds &bounds, const std::vector<Scalar> &y, Scalar &d0);
// Status Infeasible with y as the certificate, or NumericalFailure if y fails
// farkas_certificate (a device-side detection re-checked in FP64).
void report_infeasible(const Model &model, const Bounds &bounds, const std::vector<Scalar> &y, SolveResult &res);
Scalar norm2(const std::vector<Scalar> &v);
// The O(nnz) part of postsolve on the unscaled A: ax = A x
// (continuous_common.hpp:302-315) and g = c + A'y (pdhg.cpp:188-196).
struct Products {
    std::vector<Scalar> ax, g;
};
Products host_products(const Model &model, const std::vector<Scalar> &x, const std::vector<Scalar> &y);
// The same sums on the device (vk_hip_csr_postsolve_products; bitwise equal).
Products device_products(const vk_hip_csr *A, const Model &model, const std::vector<Scalar> &x,
                         const std::vector<Scalar> &y);
// scale() and, when `equilibrated`, equilibrate(.., kRuizPasses) on the
// device (vk_hip_lp_scale; bitwise equal): A's device values become the
// scaled ones and the returned Scaled has no host values.
Scaled device_scale(vk_hip_csr *A, const Model &model, const Bounds &bounds, bool equilibrated);
// Applies options.spmv_tile_geometry to A (the default geometry leaves A as
// it is; -1 times the geometries on A in the solve's storage precision and
// keeps the fastest) and returns the geometry used.
int apply_tile_geometry(vk_hip_csr *A, const SolverOptions &options);
// "" for the default geometry, else the note SolveResult::message carries.
std::string tile_geometry_note(int geometry, bool tuned);
// Relative KKT error in original coordinates into res.kkt_* (pr.g the
// stationarity c + A'y, or c + Qx + A'y for a QP with xqx = x'Qx).
void relative_kkt(const Model &model, const Bounds &bounds, const std::vector<Scalar> &x,
                  const std::vector<Scalar> &y, const Products &pr, SolveResult &res, Scalar xqx);
void postsolve(const Model &model, const Bounds &bounds, const SolverOptions &options,
               const std::vector<Scalar> &x, const std::vector<Scalar> &y, const Products &products,
               bool relative_mode, SolveResult &res);                      // pdhg.cpp:153-264
void check_vk(int rc, const char *what);
// Owns a device solver handle.
struct DeviceSolver {
    vk_hip_pdhg *h = nullptr;
    ~DeviceSolver() { vk_hip_pdhg_destroy(h); }
};
// One device solve of the scaled LP on A with `params` (pdhg_lp.cpp),
// optionally warm-started from scaled iterates; returns the VK_PDHG_*
// status, the scaled iterates, the final state and the device time.
int run_device(const vk_hip_csr *A, const Scaled &sc, const vk_pdhg_params &params,
               const std::vector<Scalar> *x0, const std::vector<Scalar> *y0, Clock::time_point started,
               const SolverOptions &options, std::vector<Scalar> &xs, std::vector<Scalar> &ys,
               vk_pdhg_state &state, double &device_ms, bool &timed);
// run_device's iteration loop on a solver already created with `params`:
// runs to a final status, the iteration limit or the time limit, then reads
// the state and the scaled iterates (n columns, m rows).
int drive_device(DeviceSolver &solver, const vk_pdhg_params &params, std::size_t n, std::size_t m,
                 Clock::time_point started, const SolverOptions &options, std::vector<Scalar> &xs,
                 std::vector<Scalar> &ys, vk_pdhg_state &state, double &device_ms, bool &timed);
// The device matrix of one model's node relaxations, scaled once, with its
// scaling and step, kept across the solve_lp_nodes rounds of a
// branch-and-bound search (the options must not change between rounds).
struct NodeCache {
    vk_hip_csr *A = nullptr;
    Scaled base;             // device_scale's result (values on the device)
    double step = 0.0;
    int tile_geometry = 0;
    std::string tiles_note;
    NodeCache() = default;
    NodeCache(const NodeCache &) = delete;
    NodeCache &operator=(const NodeCache &) = delete;
    ~NodeCache() { vk_hip_csr_destroy(A); }
};
// solve_lp_nodes on the cache's matrix (built at the first call).
std::vector<SolveResult> solve_lp_nodes_cached(const Model &model, const std::vector<std::vector<Scalar>> &lower,
                                               const std::vector<std::vector<Scalar>> &upper,
                                               const std::vector<std::vector<Scalar>> &warm_start,
                                               const SolverOptions &options, NodeCache &cache);
// continuous_common.hpp:242-275 (dense) up to n = 4096; beyond that a
// diagonally dominant Q with a non-negative diagonal.  `message` says why not.
bool positive_semidefinite(const CsrMatrix &q, std::string &message);

}  // namespace vikalp::solver::detail
