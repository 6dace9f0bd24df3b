// HipBackend: the ExecutionBackend (ExecutionBackend.hpp:36-64) on MI355X,
// built only on the libvikalp_hip C ABI.  It takes the place of the
// reference's CudaBackend (cuda/cuda_ops.cu:312-687): same op semantics and
// error behaviour (std::invalid_argument for dimension/type errors,
// std::runtime_error for device errors), FP64, device memory owned by the
// vector/matrix objects.

#include "vikalp/backend/HipBackend.hpp"

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

namespace vikalp {
namespace {

[[noreturn]] void dim_error(const std::string &message) {
    throw std::invalid_argument("HipBackend: " + message);
}

void check(int rc, const char *what) {
    if (rc == VK_OK) return;
    const std::string detail = std::string(what) + ": " + vk_hip_last_error();
    if (rc == VK_EINVAL) throw std::invalid_argument("HipBackend: " + detail);
    throw std::runtime_error("HipBackend: " + detail);
}

class HipVector final : public BackendVector {
public:
    explicit HipVector(Index n) : size_(n) {
        if (n < 0) dim_error("vector size must be non-negative");
        void *p = nullptr;
        check(vk_hip_alloc(&p, static_cast<size_t>(n) * sizeof(double)), "vk_hip_alloc");
        data_ = static_cast<double *>(p);
        check(vk_hip_fill(data_, n, 0.0, nullptr), "vk_hip_fill");
    }
    ~HipVector() override { vk_hip_free(data_); }
    HipVector(const HipVector &) = delete;
    HipVector &operator=(const HipVector &) = delete;

    [[nodiscard]] Index size() const noexcept override { return size_; }
    double *data() const noexcept { return data_; }

private:
    Index size_;
    double *data_ = nullptr;
};

class HipMatrix final : public BackendMatrix {
public:
    explicit HipMatrix(const CsrPattern &p) : rows_(p.rows), cols_(p.columns) {
        // cpu_ops.cpp:38-47: shape checks before the element checks
        // (done by vk_hip_csr_create).
        if (p.rows < 0 || p.columns < 0) dim_error("matrix dimensions must be non-negative");
        if (p.row_offsets.size() != static_cast<std::size_t>(p.rows) + 1)
            dim_error("CsrPattern row_offsets size != rows+1");
        if (p.row_offsets.front() != 0 ||
            p.row_offsets.back() != static_cast<Index>(p.column_indices.size()))
            dim_error("CsrPattern row_offsets do not describe column_indices");
        nnz_ = static_cast<Index>(p.column_indices.size());
        check(vk_hip_csr_create(p.rows, p.columns, p.row_offsets.data(),
                                p.column_indices.empty() ? nullptr : p.column_indices.data(), &handle_),
              "vk_hip_csr_create");
    }
    ~HipMatrix() override { vk_hip_csr_destroy(handle_); }
    HipMatrix(const HipMatrix &) = delete;
    HipMatrix &operator=(const HipMatrix &) = delete;

    [[nodiscard]] Index rows() const noexcept override { return rows_; }
    [[nodiscard]] Index columns() const noexcept override { return cols_; }
    Index nnz() const noexcept { return nnz_; }
    vk_hip_csr *handle() const noexcept { return handle_; }

private:
    Index rows_, cols_, nnz_ = 0;
    vk_hip_csr *handle_ = nullptr;
};

HipVector &as_vec(BackendVector &v) {
    auto *p = dynamic_cast<HipVector *>(&v);
    if (!p) dim_error("BackendVector is not a HipVector");
    return *p;
}

const HipVector &as_vec(const BackendVector &v) {
    const auto *p = dynamic_cast<const HipVector *>(&v);
    if (!p) dim_error("BackendVector is not a HipVector");
    return *p;
}

const HipMatrix &as_mat(const BackendMatrix &m) {
    const auto *p = dynamic_cast<const HipMatrix *>(&m);
    if (!p) dim_error("BackendMatrix is not a HipMatrix");
    return *p;
}

HipMatrix &as_mat(BackendMatrix &m) {
    auto *p = dynamic_cast<HipMatrix *>(&m);
    if (!p) dim_error("BackendMatrix is not a HipMatrix");
    return *p;
}

class HipBackend final : public ExecutionBackend {
public:
    HipBackend() {
        check(vk_hip_init(0), "vk_hip_init");
        stream_ = vk_hip_stream();
    }

    std::unique_ptr<BackendVector> create_vector(Index size) override {
        return std::make_unique<HipVector>(size);
    }

    std::unique_ptr<BackendMatrix> create_matrix(const CsrPattern &pattern) override {
        return std::make_unique<HipMatrix>(pattern);
    }

    void upload(std::span<const Scalar> source, BackendVector &target) override {
        auto &v = as_vec(target);
        if (static_cast<Index>(source.size()) != v.size())
            dim_error("upload: source size " + std::to_string(source.size()) + " != vector size " +
                      std::to_string(v.size()));
        check(vk_hip_memcpy_h2d(v.data(), source.data(), source.size() * sizeof(double), stream_),
              "upload");
    }

    void download(const BackendVector &source, std::span<Scalar> target) const override {
        const auto &v = as_vec(source);
        if (static_cast<Index>(target.size()) != v.size())
            dim_error("download: target size " + std::to_string(target.size()) + " != vector size " +
                      std::to_string(v.size()));
        check(vk_hip_memcpy_d2h(target.data(), v.data(), target.size() * sizeof(double), stream_),
              "download");
    }

    void set_values(std::span<const Scalar> values, BackendMatrix &target) override {
        auto &m = as_mat(target);
        if (static_cast<Index>(values.size()) != m.nnz())
            dim_error("set_values: values size " + std::to_string(values.size()) + " != nnz " +
                      std::to_string(m.nnz()));
        check(vk_hip_csr_set_values(m.handle(), values.data()), "set_values");
    }

    void fill(BackendVector &target, Scalar value) override {
        auto &v = as_vec(target);
        check(vk_hip_fill(v.data(), v.size(), value, stream_), "fill");
    }

    void copy(const BackendVector &source, BackendVector &target) override {
        const auto &s = as_vec(source);
        auto &t = as_vec(target);
        if (s.size() != t.size())
            dim_error("copy: source size " + std::to_string(s.size()) + " != target size " +
                      std::to_string(t.size()));
        check(vk_hip_memcpy_d2d(t.data(), s.data(), static_cast<size_t>(s.size()) * sizeof(double), stream_),
              "copy");
    }

    void axpby(Scalar alpha, const BackendVector &x, Scalar beta, BackendVector &y) override {
        const auto &xv = as_vec(x);
        auto &yv = as_vec(y);
        if (xv.size() != yv.size())
            dim_error("axpby: x size " + std::to_string(xv.size()) + " != y size " + std::to_string(yv.size()));
        check(vk_hip_axpby(xv.size(), alpha, xv.data(), beta, yv.data(), stream_), "axpby");
    }

    Scalar dot(const BackendVector &x, const BackendVector &y) override {
        const auto &xv = as_vec(x);
        const auto &yv = as_vec(y);
        if (xv.size() != yv.size())
            dim_error("dot: size mismatch " + std::to_string(xv.size()) + " vs " + std::to_string(yv.size()));
        double r = 0.0;
        check(vk_hip_reduce_dot(xv.size(), xv.data(), yv.data(), &r, stream_), "dot");
        return r;
    }

    Scalar norm_inf(const BackendVector &x) override {
        const auto &v = as_vec(x);
        double r = 0.0;
        check(vk_hip_reduce_absmax(v.size(), v.data(), &r, stream_), "norm_inf");
        return r;
    }

    void project_box(BackendVector &x, const BackendVector &lower, const BackendVector &upper) override {
        auto &xv = as_vec(x);
        const auto &lv = as_vec(lower);
        const auto &uv = as_vec(upper);
        if (lv.size() != xv.size() || uv.size() != xv.size())
            dim_error("project_box: x/lower/upper sizes must match (" + std::to_string(xv.size()) + " vs " +
                      std::to_string(lv.size()) + " vs " + std::to_string(uv.size()) + ")");
        check(vk_hip_project_box(xv.size(), xv.data(), lv.data(), uv.data(), stream_), "project_box");
    }

    void multiply(Scalar alpha, const BackendMatrix &matrix, const BackendVector &x, Scalar beta,
                  BackendVector &y, Transpose transpose) override {
        const auto &m = as_mat(matrix);
        const auto &xv = as_vec(x);
        auto &yv = as_vec(y);
        const bool t = transpose == Transpose::Yes;
        const Index nx = t ? m.rows() : m.columns();
        const Index ny = t ? m.columns() : m.rows();
        const char *mode = t ? "multiply (Yes)" : "multiply (No)";
        if (xv.size() != nx)
            dim_error(std::string(mode) + ": x size " + std::to_string(xv.size()) + " != matrix " +
                      (t ? "rows " : "columns ") + std::to_string(nx));
        if (yv.size() != ny)
            dim_error(std::string(mode) + ": y size " + std::to_string(yv.size()) + " != matrix " +
                      (t ? "columns " : "rows ") + std::to_string(ny));
        if (xv.data() == yv.data()) dim_error(std::string(mode) + ": x and y must not alias");
        check(vk_hip_spmv_csr(m.handle(), t ? 1 : 0, alpha, xv.data(), beta, yv.data(), stream_), mode);
    }

    // BiCGSTAB from x = 0 with a fixed shadow residual, the same algorithm and
    // exits as the CPU backend (cpu_ops.cpp:160-267), on device vectors.
    LinearSolveResult solve_linear_system(const BackendMatrix &matrix, const BackendVector &rhs,
                                          BackendVector &solution, MatrixProperty,
                                          Scalar tolerance, Index iteration_limit) override {
        const auto &A = as_mat(matrix);
        const auto &b = as_vec(rhs);
        auto &x = as_vec(solution);
        if (A.rows() != A.columns())
            dim_error("solve_linear_system: matrix must be square (" + std::to_string(A.rows()) + "x" +
                      std::to_string(A.columns()) + ")");
        if (b.size() != A.rows())
            dim_error("solve_linear_system: rhs size " + std::to_string(b.size()) + " != matrix rows " +
                      std::to_string(A.rows()));
        if (x.size() != A.rows())
            dim_error("solve_linear_system: solution size " + std::to_string(x.size()) +
                      " != matrix rows " + std::to_string(A.rows()));
        if (!std::isfinite(tolerance) || tolerance < 0.0)
            dim_error("solve_linear_system: tolerance must be finite and non-negative");
        if (iteration_limit < 0) dim_error("solve_linear_system: iteration_limit must be non-negative");

        const Index n = b.size();
        HipVector r(n), rh(n), p(n), v(n), s(n), t(n), tmp(n);
        fill(x, 0.0);
        copy(b, r);
        copy(b, rh);
        double rho_prev = 1.0, alpha = 1.0, omega = 1.0;
        double rhr = dot(rh, r);
        LinearSolveResult res;
        const double initial = norm_inf(r);
        if (initial <= tolerance) return LinearSolveResult{true, 0, initial};
        for (Index iter = 0; iter < iteration_limit; ++iter) {
            res.iterations = iter + 1;
            const double rho = rhr;
            if (rho == 0.0) break;
            if (iter == 0) {
                copy(r, p);
            } else {
                const double beta = (rho / rho_prev) * (alpha / omega);
                axpby(-omega, v, 1.0, p);   // p - omega v
                axpby(1.0, r, beta, p);     // r + beta (p - omega v)
            }
            multiply(1.0, A, p, 0.0, v, Transpose::No);
            const double da = dot(rh, v);
            if (da == 0.0) break;
            alpha = rho / da;
            copy(r, s);
            axpby(-alpha, v, 1.0, s);
            const double ns = norm_inf(s);
            if (ns <= tolerance) {
                axpby(alpha, p, 1.0, x);
                res.converged = true;
                res.residual = ns;
                return res;
            }
            multiply(1.0, A, s, 0.0, t, Transpose::No);
            const double tt = dot(t, t);
            if (tt == 0.0) break;
            omega = dot(t, s) / tt;
            if (omega == 0.0) {
                axpby(alpha, p, 1.0, x);
                copy(s, r);
                break;
            }
            axpby(alpha, p, 1.0, x);
            axpby(omega, s, 1.0, x);
            copy(s, r);
            axpby(-omega, t, 1.0, r);
            const double nr = norm_inf(r);
            if (nr <= tolerance) {
                res.converged = true;
                res.residual = nr;
                return res;
            }
            rhr = dot(rh, r);
            rho_prev = rho;
        }
        multiply(1.0, A, x, 0.0, tmp, Transpose::No);
        axpby(1.0, b, -1.0, tmp);   // b - A x
        res.converged = false;
        res.residual = norm_inf(tmp);
        return res;
    }

    void synchronize() override { check(vk_hip_sync(stream_), "synchronize"); }

private:
    void *stream_ = nullptr;
};

}  // namespace

std::unique_ptr<ExecutionBackend> make_hip_backend() { return std::make_unique<HipBackend>(); }

const vk_hip_csr *hip_matrix_handle(const BackendMatrix &matrix) noexcept {
    const auto *m = dynamic_cast<const HipMatrix *>(&matrix);
    return m ? m->handle() : nullptr;
}

vk_hip_csr *hip_matrix_handle(BackendMatrix &matrix) noexcept {
    auto *m = dynamic_cast<HipMatrix *>(&matrix);
    return m ? m->handle() : nullptr;
}

double *hip_vector_data(BackendVector &vector) noexcept {
    auto *v = dynamic_cast<HipVector *>(&vector);
    return v ? v->data() : nullptr;
}

}  // namespace vikalp
