// Diagonal scaling, operator bound and the postsolve products on the device.
//
// These are the O(nnz) host passes around the reference's PDHG loop:
//   scale_linear_model      continuous_common.hpp:133-205
//   scaled_operator_bound   continuous_common.hpp:207-225
//   linear_values           continuous_common.hpp:302-315 (postsolve, pdhg.cpp:167-168)
//   stationarity            pdhg.cpp:188-196
// plus the Ruiz / Pock-Chambolle equilibration the adaptive and Halpern
// policies add.  Each kernel does the host loop's arithmetic in the host
// loop's order: one thread walks one row of A (or of the stored transpose,
// whose rows list a column's entries by increasing row), products and sums
// rounded separately (no contraction).  Maxima do not depend on order, and
// every sum is the sequential sum the host computes, so the results are
// bitwise equal to the host loops.  These passes run once per solve; the
// thread-per-row walk keeps them simple and costs a few milliseconds on a
// 17.5M-nonzero matrix, against about 0.7 s for the same loops on one host
// core.

#include "runtime.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace vk {
namespace {

constexpr int kPreBlock = 256;

int pre_grid(int64_t n) {
    return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>((n + kPreBlock - 1) / kPreBlock, 16384)));
}

enum class RowStat {
    RefFactor,    // 1 / max|a| when that exceeds 1, else 1 (continuous_common.hpp:147-163)
    RuizFactor,   // sqrt(max|a|), 1 for an empty or all-zero row
    PcFactor,     // sqrt(sum|a|), 1 for an empty or all-zero row
    AbsSum,       // sum|a| (continuous_common.hpp:211-221)
};

// One statistic per row of a CSR operand, over its values in storage order.
template <RowStat S>
__global__ __launch_bounds__(kPreBlock) void k_row_stat(int rows, const int *__restrict__ rp,
                                                        const double *__restrict__ val,
                                                        double *__restrict__ out) {
    for (int r = blockIdx.x * kPreBlock + threadIdx.x; r < rows; r += gridDim.x * kPreBlock) {
        double acc = 0.0;
        const int e = rp[r + 1];
        for (int k = rp[r]; k < e; ++k) {
            const double a = fabs(val[k]);
            if constexpr (S == RowStat::PcFactor || S == RowStat::AbsSum) acc = acc + a;
            else acc = acc < a ? a : acc;
        }
        double f;
        if constexpr (S == RowStat::RefFactor) f = acc > 1.0 ? 1.0 / acc : 1.0;
        else if constexpr (S == RowStat::AbsSum) f = acc;
        else f = acc > 0.0 ? sqrt(acc) : 1.0;
        out[r] = f;
    }
}

// A's values times (Divide: divided by) rf[row] * cf[column]:
// continuous_common.hpp:171-179 multiplies the original value by
// (r_i * s_j); an equilibration pass divides by (rf_i * cf_j).
template <bool Divide>
__global__ __launch_bounds__(kPreBlock) void k_scale_values(int rows, const int *__restrict__ rp,
                                                            const int *__restrict__ ci,
                                                            double *__restrict__ val,
                                                            const double *__restrict__ rf,
                                                            const double *__restrict__ cf) {
#pragma clang fp contract(off)
    // Elementwise, so a wave per row (coalesced) rounds every entry exactly
    // as the host loop does.
    constexpr int waves = kPreBlock / kWave;
    const int lane = threadIdx.x % kWave;
    for (int r = blockIdx.x * waves + threadIdx.x / kWave; r < rows; r += gridDim.x * waves) {
        const double f = rf[r];
        const int e = rp[r + 1];
        for (int k = rp[r] + lane; k < e; k += kWave) {
            const double g = f * cf[ci[k]];
            val[k] = Divide ? val[k] / g : val[k] * g;
        }
    }
}

__device__ inline double scale_finite(double v, double f, bool multiply) {
    if (!isfinite(v)) return v;
    return multiply ? v * f : v / f;
}

// Column vectors.  Reference step (continuous_common.hpp:160-190): s = f,
// c * s, l / s, u / s.  Equilibration: s / f, c / f, l * f, u * f.
template <bool Ref>
__global__ void k_scale_cols(int n, const double *__restrict__ f, double *__restrict__ s,
                             double *__restrict__ c, double *__restrict__ l, double *__restrict__ u) {
    for (int j = blockIdx.x * kPreBlock + threadIdx.x; j < n; j += gridDim.x * kPreBlock) {
        const double fj = f[j];
        if (Ref) {
            s[j] = fj;
            c[j] = c[j] * fj;
            l[j] = scale_finite(l[j], fj, false);
            u[j] = scale_finite(u[j], fj, false);
        } else {
            s[j] = s[j] / fj;
            c[j] = c[j] / fj;
            l[j] = scale_finite(l[j], fj, true);
            u[j] = scale_finite(u[j], fj, true);
        }
    }
}

// Row vectors.  Reference step: r = f, rl * r, ru * r.  Equilibration:
// r / f, rl / f, ru / f.
template <bool Ref>
__global__ void k_scale_rows(int m, const double *__restrict__ f, double *__restrict__ r,
                             double *__restrict__ rl, double *__restrict__ ru) {
    for (int i = blockIdx.x * kPreBlock + threadIdx.x; i < m; i += gridDim.x * kPreBlock) {
        const double fi = f[i];
        if (Ref) {
            r[i] = fi;
            rl[i] = scale_finite(rl[i], fi, true);
            ru[i] = scale_finite(ru[i], fi, true);
        } else {
            r[i] = r[i] / fi;
            rl[i] = scale_finite(rl[i], fi, false);
            ru[i] = scale_finite(ru[i], fi, false);
        }
    }
}

// Bounds of `count` variable-bound sets through the recorded column passes:
// l / f_0 then l * f_p for each later pass, finite entries only (the same
// operations k_scale_cols applies to one set).
__global__ void k_scale_bound_sets(int n, int count, int passes, const double *__restrict__ f,
                                   double *__restrict__ l, double *__restrict__ u) {
    const int64_t total = static_cast<int64_t>(n) * count;
    for (int64_t i = blockIdx.x * (int64_t)kPreBlock + threadIdx.x; i < total; i += (int64_t)gridDim.x * kPreBlock) {
        const int j = static_cast<int>(i % n);
        double lo = l[i], hi = u[i];
        lo = scale_finite(lo, f[j], false);
        hi = scale_finite(hi, f[j], false);
        for (int p = 1; p < passes; ++p) {
            const double fp = f[static_cast<int64_t>(p) * n + j];
            lo = scale_finite(lo, fp, true);
            hi = scale_finite(hi, fp, true);
        }
        l[i] = lo;
        u[i] = hi;
    }
}

__global__ void k_take(int64_t n, const double *__restrict__ src, const int *__restrict__ perm,
                       double *__restrict__ dst) {
    for (int64_t i = blockIdx.x * (int64_t)kPreBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kPreBlock)
        dst[i] = src[perm[i]];
}

__global__ void k_narrow(int64_t n, const double *__restrict__ src, float *__restrict__ dst) {
    for (int64_t i = blockIdx.x * (int64_t)kPreBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kPreBlock)
        dst[i] = static_cast<float>(src[i]);
}

// Largest entry of a non-negative vector (0 when empty), one block.
__global__ __launch_bounds__(1024) void k_max(int n, const double *__restrict__ v, double *__restrict__ out) {
    __shared__ double part[1024 / kWave];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) acc = acc < v[i] ? v[i] : acc;
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const double o = __shfl_down(acc, off, kWave);
        acc = acc < o ? o : acc;
    }
    if (threadIdx.x % kWave == 0) part[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int w = 0; w < 1024 / kWave; ++w) m = m < part[w] ? part[w] : m;
        *out = m;
    }
}

// ax_i = sum_k a_k x_{col k} from 0 (continuous_common.hpp:305-313).
__global__ __launch_bounds__(kPreBlock) void k_row_products(int rows, const int *__restrict__ rp,
                                                            const int *__restrict__ ci,
                                                            const double *__restrict__ val,
                                                            const double *__restrict__ x,
                                                            double *__restrict__ ax) {
#pragma clang fp contract(off)
    for (int r = blockIdx.x * kPreBlock + threadIdx.x; r < rows; r += gridDim.x * kPreBlock) {
        double acc = 0.0;
        const int e = rp[r + 1];
        for (int k = rp[r]; k < e; ++k) acc = acc + val[k] * x[ci[k]];
        ax[r] = acc;
    }
}

// g_j = c_j + sum over the rows i of column j, increasing, of y_i a_ij
// (pdhg.cpp:188-196 adds them in that order as it walks the rows).  The
// unscaled values come from `orig` through the transpose permutation.
__global__ __launch_bounds__(kPreBlock) void k_col_products(int cols, const int *__restrict__ trp,
                                                            const int *__restrict__ trow,
                                                            const int *__restrict__ tperm,
                                                            const double *__restrict__ orig,
                                                            const double *__restrict__ y,
                                                            const double *__restrict__ c,
                                                            double *__restrict__ g) {
#pragma clang fp contract(off)
    for (int j = blockIdx.x * kPreBlock + threadIdx.x; j < cols; j += gridDim.x * kPreBlock) {
        double acc = c[j];
        const int e = trp[j + 1];
        for (int p = trp[j]; p < e; ++p) acc = acc + y[trow[p]] * orig[tperm[p]];
        g[j] = acc;
    }
}

// Device buffers of one call, freed on every return path.
struct Scratch {
    std::vector<void *> blocks;
    ~Scratch() {
        for (void *p : blocks) (void)hipFree(p);
    }
    template <class T>
    T *get(size_t count) {
        void *p = nullptr;
        if (device_malloc(&p, sizeof(T) * std::max<size_t>(1, count)) != hipSuccess) return nullptr;
        blocks.push_back(p);
        return static_cast<T *>(p);
    }
};

// A's values as last set: the saved originals once a scaling replaced them.
const double *original_values(const vk_hip_csr *A) { return A->scaled ? A->orig : A->a.val; }

// Refreshes the transpose values and the FP32 copies from A's values.
int propagate_values(vk_hip_csr *A, hipStream_t s) {
    const int64_t nnz = A->nnz;
    if (nnz == 0) return VK_OK;
    k_take<<<pre_grid(nnz), kPreBlock, 0, s>>>(nnz, A->a.val, A->tperm, A->at.val);
    if (A->a32.val) {
        k_narrow<<<pre_grid(nnz), kPreBlock, 0, s>>>(nnz, A->a.val, A->a32.val);
        k_narrow<<<pre_grid(nnz), kPreBlock, 0, s>>>(nnz, A->at.val, A->at32.val);
    }
    VK_HIP_CHECK(hipGetLastError());
    return VK_OK;
}

// One equilibration pass: factors from the current values, then the
// division of the values and the vectors (Divide = true), as a Ruiz or a
// Pock-Chambolle pass.
template <RowStat S>
int equilibration_pass(vk_hip_csr *A, double *rf, double *cf, double *r, double *sc, double *c,
                       double *l, double *u, double *rl, double *ru, double *record, hipStream_t s) {
    const int m = static_cast<int>(A->rows), n = static_cast<int>(A->cols);
    k_row_stat<S><<<pre_grid(m), kPreBlock, 0, s>>>(m, A->a.row_ptr, A->a.val, rf);
    k_row_stat<S><<<pre_grid(n), kPreBlock, 0, s>>>(n, A->at.row_ptr, A->at.val, cf);
    if (n) VK_HIP_CHECK(hipMemcpyAsync(record, cf, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    k_scale_values<true><<<pre_grid(static_cast<int64_t>(m) * kWave), kPreBlock, 0, s>>>(m, A->a.row_ptr, A->a.col, A->a.val, rf, cf);
    k_scale_rows<false><<<pre_grid(m), kPreBlock, 0, s>>>(m, rf, r, rl, ru);
    k_scale_cols<false><<<pre_grid(n), kPreBlock, 0, s>>>(n, cf, sc, c, l, u);
    VK_HIP_CHECK(hipGetLastError());
    return propagate_values(A, s);
}

}  // namespace
}  // namespace vk

using namespace vk;

extern "C" {

int vk_hip_lp_scale(vk_hip_csr *A, const double *c, const double *l, const double *u,
                    const double *rl, const double *ru, int ruiz_passes, int pock_chambolle,
                    double *row_scale, double *col_scale, double *c_out, double *l_out,
                    double *u_out, double *rl_out, double *ru_out) {
    VK_REQUIRE(A, "vk_hip_lp_scale: null matrix");
    VK_REQUIRE(A->has_values, "vk_hip_lp_scale: matrix values not set");
    VK_REQUIRE(ruiz_passes >= 0, "vk_hip_lp_scale: ruiz_passes must be non-negative");
    const int64_t m = A->rows, n = A->cols, nnz = A->nnz;
    VK_REQUIRE(n == 0 || (c && l && u && col_scale && c_out && l_out && u_out),
               "vk_hip_lp_scale: null column vector");
    VK_REQUIRE(m == 0 || (rl && ru && row_scale && rl_out && ru_out), "vk_hip_lp_scale: null row vector");
    hipStream_t s = default_stream();
    // Every allocation comes before the matrix is touched: a failed one
    // leaves A as it was.
    Scratch sb;
    double *dc = sb.get<double>(n), *dl = sb.get<double>(n), *du = sb.get<double>(n);
    double *ds = sb.get<double>(n), *cf = sb.get<double>(n);
    double *drl = sb.get<double>(m), *dru = sb.get<double>(m), *dr = sb.get<double>(m), *rf = sb.get<double>(m);
    if (!dc || !dl || !du || !ds || !cf || !drl || !dru || !dr || !rf) {
        set_error("vk_hip_lp_scale: scratch allocation failed");
        return VK_ENOMEM;
    }
    const int passes = 1 + ruiz_passes + (pock_chambolle ? 1 : 0);
    double *factors = nullptr;
    if (device_malloc(&factors, sizeof(double) * std::max<int64_t>(1, n * passes)) != hipSuccess ||
        (nnz && !A->scaled && !A->orig && device_malloc(&A->orig, sizeof(double) * nnz) != hipSuccess)) {
        (void)hipFree(factors);
        set_error("vk_hip_lp_scale: device allocation failed");
        return VK_ENOMEM;
    }
    (void)hipFree(A->col_factors);
    A->col_factors = factors;
    A->col_passes = 0;
    // Start from the values last set: save them on the first scaling, put
    // them back on a later one.
    if (nnz) {
        if (A->scaled) {
            VK_HIP_CHECK(hipMemcpyAsync(A->a.val, A->orig, sizeof(double) * nnz, hipMemcpyDeviceToDevice, s));
        } else {
            VK_HIP_CHECK(hipMemcpyAsync(A->orig, A->a.val, sizeof(double) * nnz, hipMemcpyDeviceToDevice, s));
        }
    }
    A->scaled = true;
    if (n) {
        VK_HIP_CHECK(hipMemcpyAsync(dc, c, sizeof(double) * n, hipMemcpyHostToDevice, s));
        VK_HIP_CHECK(hipMemcpyAsync(dl, l, sizeof(double) * n, hipMemcpyHostToDevice, s));
        VK_HIP_CHECK(hipMemcpyAsync(du, u, sizeof(double) * n, hipMemcpyHostToDevice, s));
    }
    if (m) {
        VK_HIP_CHECK(hipMemcpyAsync(drl, rl, sizeof(double) * m, hipMemcpyHostToDevice, s));
        VK_HIP_CHECK(hipMemcpyAsync(dru, ru, sizeof(double) * m, hipMemcpyHostToDevice, s));
    }
    const int mi = static_cast<int>(m), ni = static_cast<int>(n);
    // continuous_common.hpp:133-205 on the original values (the transpose
    // holds them too: set_values and the restore above both propagate).
    if (int rc = propagate_values(A, s)) return rc;
    k_row_stat<RowStat::RefFactor><<<pre_grid(m), kPreBlock, 0, s>>>(mi, A->a.row_ptr, A->a.val, rf);
    k_row_stat<RowStat::RefFactor><<<pre_grid(n), kPreBlock, 0, s>>>(ni, A->at.row_ptr, A->at.val, cf);
    if (n) VK_HIP_CHECK(hipMemcpyAsync(A->col_factors, cf, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    k_scale_values<false><<<pre_grid(static_cast<int64_t>(mi) * kWave), kPreBlock, 0, s>>>(mi, A->a.row_ptr, A->a.col, A->a.val, rf, cf);
    k_scale_rows<true><<<pre_grid(m), kPreBlock, 0, s>>>(mi, rf, dr, drl, dru);
    k_scale_cols<true><<<pre_grid(n), kPreBlock, 0, s>>>(ni, cf, ds, dc, dl, du);
    VK_HIP_CHECK(hipGetLastError());
    if (int rc = propagate_values(A, s)) return rc;
    for (int p = 0; p < ruiz_passes; ++p)
        if (int rc = equilibration_pass<RowStat::RuizFactor>(A, rf, cf, dr, ds, dc, dl, du, drl, dru,
                                                             A->col_factors + (1 + p) * n, s))
            return rc;
    if (pock_chambolle)
        if (int rc = equilibration_pass<RowStat::PcFactor>(A, rf, cf, dr, ds, dc, dl, du, drl, dru,
                                                           A->col_factors + (1 + ruiz_passes) * n, s))
            return rc;
    A->col_passes = passes;
    if (n) {
        VK_HIP_CHECK(hipMemcpyAsync(col_scale, ds, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        VK_HIP_CHECK(hipMemcpyAsync(c_out, dc, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        VK_HIP_CHECK(hipMemcpyAsync(l_out, dl, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        VK_HIP_CHECK(hipMemcpyAsync(u_out, du, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    }
    if (m) {
        VK_HIP_CHECK(hipMemcpyAsync(row_scale, dr, sizeof(double) * m, hipMemcpyDeviceToHost, s));
        VK_HIP_CHECK(hipMemcpyAsync(rl_out, drl, sizeof(double) * m, hipMemcpyDeviceToHost, s));
        VK_HIP_CHECK(hipMemcpyAsync(ru_out, dru, sizeof(double) * m, hipMemcpyDeviceToHost, s));
    }
    VK_HIP_CHECK(hipStreamSynchronize(s));
    return VK_OK;
}

int vk_hip_lp_scale_bounds(const vk_hip_csr *A, int count, const double *l, const double *u, double *l_out,
                           double *u_out) {
    VK_REQUIRE(A && count >= 0, "vk_hip_lp_scale_bounds: bad arguments");
    VK_REQUIRE(A->scaled && A->col_passes > 0, "vk_hip_lp_scale_bounds: the matrix has not been scaled");
    const int64_t n = A->cols, total = n * count;
    if (total == 0) return VK_OK;
    VK_REQUIRE(l && u && l_out && u_out, "vk_hip_lp_scale_bounds: null bounds");
    hipStream_t s = default_stream();
    Scratch sb;
    double *dl = sb.get<double>(total), *du = sb.get<double>(total);
    if (!dl || !du) {
        set_error("vk_hip_lp_scale_bounds: scratch allocation failed");
        return VK_ENOMEM;
    }
    VK_HIP_CHECK(hipMemcpyAsync(dl, l, sizeof(double) * total, hipMemcpyHostToDevice, s));
    VK_HIP_CHECK(hipMemcpyAsync(du, u, sizeof(double) * total, hipMemcpyHostToDevice, s));
    k_scale_bound_sets<<<pre_grid(total), kPreBlock, 0, s>>>(static_cast<int>(n), count, A->col_passes,
                                                             A->col_factors, dl, du);
    VK_HIP_CHECK(hipGetLastError());
    VK_HIP_CHECK(hipMemcpyAsync(l_out, dl, sizeof(double) * total, hipMemcpyDeviceToHost, s));
    VK_HIP_CHECK(hipMemcpyAsync(u_out, du, sizeof(double) * total, hipMemcpyDeviceToHost, s));
    VK_HIP_CHECK(hipStreamSynchronize(s));
    return VK_OK;
}

int vk_hip_csr_operator_bound(const vk_hip_csr *A, double *bound) {
    VK_REQUIRE(A && bound, "vk_hip_csr_operator_bound: null argument");
    VK_REQUIRE(A->has_values, "vk_hip_csr_operator_bound: matrix values not set");
    const int m = static_cast<int>(A->rows), n = static_cast<int>(A->cols);
    hipStream_t s = default_stream();
    Scratch sb;
    double *rows = sb.get<double>(m), *cols = sb.get<double>(n), *mx = sb.get<double>(2);
    if (!rows || !cols || !mx) {
        set_error("vk_hip_csr_operator_bound: scratch allocation failed");
        return VK_ENOMEM;
    }
    k_row_stat<RowStat::AbsSum><<<pre_grid(m), kPreBlock, 0, s>>>(m, A->a.row_ptr, A->a.val, rows);
    k_row_stat<RowStat::AbsSum><<<pre_grid(n), kPreBlock, 0, s>>>(n, A->at.row_ptr, A->at.val, cols);
    k_max<<<1, 1024, 0, s>>>(m, rows, mx);
    k_max<<<1, 1024, 0, s>>>(n, cols, mx + 1);
    VK_HIP_CHECK(hipGetLastError());
    double h[2];
    VK_HIP_CHECK(hipMemcpyAsync(h, mx, sizeof(h), hipMemcpyDeviceToHost, s));
    VK_HIP_CHECK(hipStreamSynchronize(s));
    *bound = std::sqrt(h[0] * h[1]);
    return VK_OK;
}

int vk_hip_csr_operator_bounds(const vk_hip_csr *A, int count, const int64_t *row_starts,
                               const int64_t *col_starts, double *bounds) {
    VK_REQUIRE(A && row_starts && col_starts && bounds && count > 0, "vk_hip_csr_operator_bounds: null argument");
    VK_REQUIRE(A->has_values, "vk_hip_csr_operator_bounds: matrix values not set");
    const int m = static_cast<int>(A->rows), n = static_cast<int>(A->cols);
    VK_REQUIRE(row_starts[0] == 0 && col_starts[0] == 0 && row_starts[count] == m && col_starts[count] == n,
               "vk_hip_csr_operator_bounds: block ranges must cover the matrix");
    for (int i = 0; i < count; ++i)
        VK_REQUIRE(row_starts[i + 1] >= row_starts[i] && col_starts[i + 1] >= col_starts[i],
                   "vk_hip_csr_operator_bounds: block ranges must be ordered");
    hipStream_t s = default_stream();
    Scratch sb;
    double *rows = sb.get<double>(m), *cols = sb.get<double>(n);
    if (!rows || !cols) {
        set_error("vk_hip_csr_operator_bounds: scratch allocation failed");
        return VK_ENOMEM;
    }
    k_row_stat<RowStat::AbsSum><<<pre_grid(m), kPreBlock, 0, s>>>(m, A->a.row_ptr, A->a.val, rows);
    k_row_stat<RowStat::AbsSum><<<pre_grid(n), kPreBlock, 0, s>>>(n, A->at.row_ptr, A->at.val, cols);
    VK_HIP_CHECK(hipGetLastError());
    std::vector<double> hr(static_cast<size_t>(m)), hc(static_cast<size_t>(n));
    if (m) VK_HIP_CHECK(hipMemcpyAsync(hr.data(), rows, sizeof(double) * m, hipMemcpyDeviceToHost, s));
    if (n) VK_HIP_CHECK(hipMemcpyAsync(hc.data(), cols, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    VK_HIP_CHECK(hipStreamSynchronize(s));
    for (int i = 0; i < count; ++i) {
        double mr = 0.0, mc = 0.0;
        for (int64_t r = row_starts[i]; r < row_starts[i + 1]; ++r) mr = std::max(mr, hr[static_cast<size_t>(r)]);
        for (int64_t j = col_starts[i]; j < col_starts[i + 1]; ++j) mc = std::max(mc, hc[static_cast<size_t>(j)]);
        bounds[i] = std::sqrt(mr * mc);
    }
    return VK_OK;
}

int vk_hip_csr_postsolve_products(const vk_hip_csr *A, const double *x, const double *y, const double *c,
                                  double *ax, double *g) {
    VK_REQUIRE(A, "vk_hip_csr_postsolve_products: null matrix");
    VK_REQUIRE(A->has_values, "vk_hip_csr_postsolve_products: matrix values not set");
    const int m = static_cast<int>(A->rows), n = static_cast<int>(A->cols);
    VK_REQUIRE(n == 0 || (x && c && g), "vk_hip_csr_postsolve_products: null column vector");
    VK_REQUIRE(m == 0 || (y && ax), "vk_hip_csr_postsolve_products: null row vector");
    hipStream_t s = default_stream();
    Scratch sb;
    double *dx = sb.get<double>(n), *dy = sb.get<double>(m), *dc = sb.get<double>(n);
    double *dax = sb.get<double>(m), *dg = sb.get<double>(n);
    if (!dx || !dy || !dc || !dax || !dg) {
        set_error("vk_hip_csr_postsolve_products: scratch allocation failed");
        return VK_ENOMEM;
    }
    if (n) {
        VK_HIP_CHECK(hipMemcpyAsync(dx, x, sizeof(double) * n, hipMemcpyHostToDevice, s));
        VK_HIP_CHECK(hipMemcpyAsync(dc, c, sizeof(double) * n, hipMemcpyHostToDevice, s));
    }
    if (m) VK_HIP_CHECK(hipMemcpyAsync(dy, y, sizeof(double) * m, hipMemcpyHostToDevice, s));
    const double *orig = original_values(A);
    k_row_products<<<pre_grid(m), kPreBlock, 0, s>>>(m, A->a.row_ptr, A->a.col, orig, dx, dax);
    k_col_products<<<pre_grid(n), kPreBlock, 0, s>>>(n, A->at.row_ptr, A->at.col, A->tperm, orig, dy, dc, dg);
    VK_HIP_CHECK(hipGetLastError());
    if (m) VK_HIP_CHECK(hipMemcpyAsync(ax, dax, sizeof(double) * m, hipMemcpyDeviceToHost, s));
    if (n) VK_HIP_CHECK(hipMemcpyAsync(g, dg, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    VK_HIP_CHECK(hipStreamSynchronize(s));
    return VK_OK;
}

}  // extern "C"
