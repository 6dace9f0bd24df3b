// UnrolledPdhg: the forward / backward pair over K fixed-step PDHG
// iterations on the HIP device (vk_hip_unrolled_*).  The host side follows
// solve_lp's preparation: the model checks of pdhg.cpp:269-293, the scaling
// of continuous_common.hpp:133-205 (fixed for a given A, so c and the row
// bounds enter only through c*s and r*rl, r*ru), the step of pdhg.cpp:75-78
// and the projected zero start of pdhg.cpp:71.

#include "lp_internal.hpp"

#include "vikalp/solver/continuous.hpp"
#include "vikalp_hip.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

namespace vikalp::solver {

using namespace detail;

struct UnrolledPdhg::Impl {
    Index m = 0, n = 0, iterations = 0;
    Scaled sc;
    Scalar step = 0.0;
    vk_hip_csr *A = nullptr;
    vk_hip_unrolled *h = nullptr;
    bool forward_done = false;
    double last_ms = 0.0;
    ~Impl() {
        vk_hip_unrolled_destroy(h);
        vk_hip_csr_destroy(A);
    }
};

UnrolledPdhg::UnrolledPdhg(const Model &model, Index iterations, Index checkpoint_every)
    : impl_(std::make_unique<Impl>()) {
    if (iterations < 0 || checkpoint_every < 1)
        throw std::invalid_argument("UnrolledPdhg: iterations must be >= 0 and checkpoint_every >= 1");
    const auto errors = model.validate();
    if (!errors.empty()) throw std::invalid_argument("UnrolledPdhg: " + errors.front());
    if (model.problem_class() != ProblemClass::LP)
        throw std::invalid_argument("UnrolledPdhg: the model is not an LP");
    if (model.num_variables() == 0) throw std::invalid_argument("UnrolledPdhg: LP requires at least one variable");
    Bounds bounds;
    std::string message;
    if (!resolve_bounds(model, {}, {}, bounds, message)) throw std::invalid_argument("UnrolledPdhg: " + message);
    Impl &I = *impl_;
    I.m = model.num_constraints();
    I.n = model.num_variables();
    I.iterations = iterations;
    I.sc = scale(model, bounds);
    const auto &pat = model.constraint_matrix.pattern;
    I.step = 0.99 / std::max<Scalar>(1.0, operator_bound(pat, I.sc.values));   // pdhg.cpp:75-78
    check_vk(vk_hip_csr_create(pat.rows, pat.columns, pat.row_offsets.data(), pat.column_indices.data(), &I.A),
             "vk_hip_csr_create");
    check_vk(vk_hip_csr_set_values(I.A, I.sc.values.data()), "vk_hip_csr_set_values");
    check_vk(vk_hip_unrolled_create(I.A, I.sc.c.data(), I.sc.l.data(), I.sc.u.data(),
                                    I.sc.rl.empty() ? nullptr : I.sc.rl.data(),
                                    I.sc.ru.empty() ? nullptr : I.sc.ru.data(), I.step, I.step, iterations,
                                    checkpoint_every, &I.h),
             "vk_hip_unrolled_create");
}

UnrolledPdhg::~UnrolledPdhg() = default;

void UnrolledPdhg::forward(std::span<Scalar> x, std::span<Scalar> y) {
    Impl &I = *impl_;
    if (static_cast<Index>(x.size()) != I.n || static_cast<Index>(y.size()) != I.m)
        throw std::invalid_argument("UnrolledPdhg::forward: output sizes must be n and m");
    // pdhg.cpp:71: the zero start projected onto the scaled bounds; y = 0.
    std::vector<Scalar> x0(static_cast<std::size_t>(I.n)), y0(static_cast<std::size_t>(I.m), 0.0);
    for (Index j = 0; j < I.n; ++j) x0[j] = std::min(std::max(0.0, I.sc.l[j]), I.sc.u[j]);
    std::vector<Scalar> xs(static_cast<std::size_t>(I.n)), ys(static_cast<std::size_t>(I.m));
    check_vk(vk_hip_unrolled_forward(I.h, x0.data(), y0.empty() ? nullptr : y0.data(), xs.data(),
                                     ys.empty() ? nullptr : ys.data()),
             "vk_hip_unrolled_forward");
    check_vk(vk_hip_unrolled_last_ms(I.h, &I.last_ms), "vk_hip_unrolled_last_ms");
    for (Index j = 0; j < I.n; ++j) x[j] = xs[j] * I.sc.s[j];
    for (Index i = 0; i < I.m; ++i) y[i] = ys[i] * I.sc.r[i];
    I.forward_done = true;
}

void UnrolledPdhg::set_data(std::span<const Scalar> objective, std::span<const Scalar> row_lower,
                            std::span<const Scalar> row_upper) {
    Impl &I = *impl_;
    if (static_cast<Index>(objective.size()) != I.n || static_cast<Index>(row_lower.size()) != I.m ||
        static_cast<Index>(row_upper.size()) != I.m)
        throw std::invalid_argument("UnrolledPdhg::set_data: sizes must be n and m");
    for (Index j = 0; j < I.n; ++j)
        if (!std::isfinite(objective[j])) throw std::invalid_argument("UnrolledPdhg::set_data: objective not finite");
    for (Index i = 0; i < I.m; ++i) {
        const Scalar lo = row_lower[i], hi = row_upper[i];
        if (std::isnan(lo) || std::isnan(hi) || lo > hi || std::isfinite(lo) != std::isfinite(I.sc.rl[i]) ||
            std::isfinite(hi) != std::isfinite(I.sc.ru[i]))
            throw std::invalid_argument("UnrolledPdhg::set_data: row bounds must keep their kind and order");
    }
    // continuous_common.hpp:133-205: c hat = s c, row bounds times r.
    for (Index j = 0; j < I.n; ++j) I.sc.c[j] = objective[j] * I.sc.s[j];
    for (Index i = 0; i < I.m; ++i) {
        I.sc.rl[i] = row_lower[i] * I.sc.r[i];
        I.sc.ru[i] = row_upper[i] * I.sc.r[i];
    }
    check_vk(vk_hip_unrolled_set_data(I.h, I.sc.c.data(), I.sc.rl.empty() ? nullptr : I.sc.rl.data(),
                                      I.sc.ru.empty() ? nullptr : I.sc.ru.data()),
             "vk_hip_unrolled_set_data");
    I.forward_done = false;
}

void UnrolledPdhg::backward(std::span<const Scalar> grad_x, std::span<const Scalar> grad_y,
                            std::span<Scalar> grad_c, std::span<Scalar> grad_row_lower,
                            std::span<Scalar> grad_row_upper) {
    Impl &I = *impl_;
    if (!I.forward_done) throw std::invalid_argument("UnrolledPdhg::backward: run forward() first");
    if ((!grad_x.empty() && static_cast<Index>(grad_x.size()) != I.n) ||
        (!grad_y.empty() && static_cast<Index>(grad_y.size()) != I.m) ||
        static_cast<Index>(grad_c.size()) != I.n || static_cast<Index>(grad_row_lower.size()) != I.m ||
        static_cast<Index>(grad_row_upper.size()) != I.m)
        throw std::invalid_argument("UnrolledPdhg::backward: sizes must be n (x, c) and m (y, row bounds)");
    // x = s xhat and y = r yhat: adjoints of the scaled state.
    std::vector<Scalar> gx(static_cast<std::size_t>(I.n), 0.0), gy(static_cast<std::size_t>(I.m), 0.0);
    for (Index j = 0; j < I.n && !grad_x.empty(); ++j) gx[j] = grad_x[j] * I.sc.s[j];
    for (Index i = 0; i < I.m && !grad_y.empty(); ++i) gy[i] = grad_y[i] * I.sc.r[i];
    std::vector<Scalar> gc(static_cast<std::size_t>(I.n)), grl(static_cast<std::size_t>(I.m)),
        gru(static_cast<std::size_t>(I.m));
    check_vk(vk_hip_unrolled_backward(I.h, gx.data(), gy.empty() ? nullptr : gy.data(), gc.data(),
                                      grl.empty() ? nullptr : grl.data(), gru.empty() ? nullptr : gru.data()),
             "vk_hip_unrolled_backward");
    check_vk(vk_hip_unrolled_last_ms(I.h, &I.last_ms), "vk_hip_unrolled_last_ms");
    // c hat = s c, rl hat = r rl, ru hat = r ru.
    for (Index j = 0; j < I.n; ++j) grad_c[j] = gc[j] * I.sc.s[j];
    for (Index i = 0; i < I.m; ++i) {
        grad_row_lower[i] = grl[i] * I.sc.r[i];
        grad_row_upper[i] = gru[i] * I.sc.r[i];
    }
}

Scalar UnrolledPdhg::step_size() const { return impl_->step; }
Index UnrolledPdhg::iterations() const { return impl_->iterations; }
double UnrolledPdhg::last_device_ms() const { return impl_->last_ms; }

}  // namespace vikalp::solver
