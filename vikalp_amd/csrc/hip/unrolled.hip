// Unrolled fixed-step PDHG and its adjoint for gfx950 (differentiation
// through K iterations; the reference has no counterpart).
//
// Forward: K steps of the reference iteration (src/solver/lp/pdhg.cpp:92-113)
// with tau and sigma held fixed and the restart branch (pdhg.cpp:98-100)
// never taken, on the scaled LP:
//   z = x - tau c - tau A'y,  x+ = P_[l,u](z),  xbar = 2 x+ - x,
//   w = y + sigma A xbar,  t = w / sigma,  y+ = w - sigma P_[rl,ru](t).
// Two kernels per step: A'y over the stored transpose with the primal update
// in its epilogue, then A xbar with the dual update.  (x, y) is saved every
// `checkpoint_every` steps.
//
// Backward: with adjoints (gx, gy) of the state after a step,
//   gw  = [row clamped] gy,  grl -= sigma [t < rl] gy,  gru -= sigma [t > ru] gy
//   gxb = sigma A' gw                              (A' SpMV, columns)
//   gz  = [l <= z <= u] (gx + 2 gxb),  gx = gz - gxb,  gc -= tau gz
//   gy  = gw - tau A gz                            (A SpMV, rows)
// so the SpMVs of the forward step run in the opposite order on the other
// operand.  Only the projections' active sets enter: each segment of
// checkpoint_every steps is recomputed from its checkpoint with the masks
// recorded as bitsets (one bit per column -- inside its box -- and two per
// row -- below / above its bounds -- per step), then swept in reverse.  The recomputation runs the forward kernels on the same inputs,
// so the masks equal the forward pass's bitwise.  Every accumulation runs in
// a fixed order; results repeat bitwise.

#include "runtime.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

namespace vk {
namespace {

#pragma clang fp contract(off)

struct UnrolledArgs {
    double *x, *y, *xbar;
    const double *c, *l, *u, *rl, *ru;
    // Bitsets of this step (nullptr: not recorded): column j inside [l, u];
    // row i below rl / above ru.  Bit b of word w is entry 64 w + b.
    unsigned long long *mx, *my_lo, *my_hi;
    double *gx, *gy, *gw, *gz, *gc, *grl, *gru;
    double tau, sigma, inv_sigma;
    int m, n;
};

// Writes the bits of entries [r0, r0 + 64) that `own` selects (bit b: entry
// r0 + b) of bitset `set` to `val`.  A word can straddle two tiles whose
// waves write their own bits concurrently, so each word takes an atomic AND
// (clearing the owned bits that become 0) and an atomic OR (setting those
// that become 1); bits the wave does not own are left alone.
__device__ __forceinline__ void store_bits(unsigned long long *set, int r0, unsigned long long own,
                                           unsigned long long val) {
    auto put = [](unsigned long long *w, unsigned long long o, unsigned long long v) {
        if (!o) return;
        atomicAnd(w, ~o | v);
        atomicOr(w, v);
    };
    const int w = r0 >> 6, s = r0 & 63;
    put(set + w, own << s, val << s);
    if (s) put(set + w + 1, own >> (64 - s), val >> (64 - s));
}

// The epilogue runs on consecutive lanes for consecutive rows starting at
// lane 0 (tile_spmv's phase 3, or lane 0 alone for a long row), so the
// wave's ballots are the bits of rows [j - lane, j - lane + 64).
__device__ __forceinline__ void record_bit(unsigned long long *set, int j, bool bit) {
    const unsigned long long own = __ballot(1), val = __ballot(bit);
    const int lane = static_cast<int>(threadIdx.x % kWave);
    if (lane == 0) store_bits(set, j, own, val);
}

__device__ __forceinline__ bool read_bit(const unsigned long long *set, int j) {
    return (set[j >> 6] >> (j & 63)) & 1ull;
}

// pdhg.cpp:92-96 and :102-103 on column j given acc = (A'y)_j.
struct FwdPrimalEpi {
    const UnrolledArgs &a;
    struct Pre {
        double x, c, l, u;
    };
    __device__ Pre prefetch(int j) const { return Pre{a.x[j], a.c[j], a.l[j], a.u[j]}; }
    __device__ void row(int j, double acc, const Pre &q, Partials &) const {
        double z = -a.tau * q.c + q.x;
        z = -a.tau * acc + z;
        if (a.mx) record_bit(a.mx, j, !(z < q.l) && !(q.u < z));
        const double xn = clamp_box(z, q.l, q.u);
        a.xbar[j] = 2.0 * xn + -1.0 * q.x;
        a.x[j] = xn;
    }
};

// pdhg.cpp:105-113 on row i given acc = (A xbar)_i.
struct FwdDualEpi {
    const UnrolledArgs &a;
    struct Pre {
        double y, lo, hi;
    };
    __device__ Pre prefetch(int i) const { return Pre{a.y[i], a.rl[i], a.ru[i]}; }
    __device__ void row(int i, double acc, const Pre &q, Partials &) const {
        const double w = a.sigma * acc + q.y;
        const double t = a.inv_sigma * w;
        if (a.my_lo) {
            record_bit(a.my_lo, i, t < q.lo);
            record_bit(a.my_hi, i, q.hi < t);
        }
        const double p = clamp_box(t, q.lo, q.hi);
        a.y[i] = -a.sigma * p + w;
    }
};

// Backward, columns: acc = (A' gw)_j.
struct BwdPrimalEpi {
    const UnrolledArgs &a;
    struct Pre {
        double gx, gc;
        bool in;
    };
    __device__ Pre prefetch(int j) const { return Pre{a.gx[j], a.gc[j], read_bit(a.mx, j)}; }
    __device__ void row(int j, double acc, const Pre &q, Partials &) const {
        const double gxb = a.sigma * acc;
        const double gz = q.in ? q.gx + 2.0 * gxb : 0.0;
        a.gz[j] = gz;
        a.gx[j] = gz - gxb;
        a.gc[j] = q.gc - a.tau * gz;
    }
};

// Backward, rows: acc = (A gz)_i.
struct BwdDualEpi {
    const UnrolledArgs &a;
    struct Pre {
        double gw;
    };
    __device__ Pre prefetch(int i) const { return Pre{a.gw[i]}; }
    __device__ void row(int i, double acc, const Pre &q, Partials &) const { a.gy[i] = q.gw - a.tau * acc; }
};

template <class Epi>
__global__ __launch_bounds__(kBlock) void k_unrolled_spmv(UnrolledArgs a, CsrView<double> A, const double *x) {
    __shared__ double lds[kWaves][kTileNnz];
    Partials unused;
    unused.zero();
    Epi epi{a};
    tile_spmv(A, x, epi, unused, lds[threadIdx.x / kWave]);
}

// Backward, rows, elementwise: gw and the row-bound adjoints of one step.
__global__ void k_unrolled_rows(UnrolledArgs a) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.m; i += gridDim.x * kBlock) {
        const bool below = read_bit(a.my_lo, i), above = read_bit(a.my_hi, i);
        const double g = a.gy[i];
        a.gw[i] = (below || above) ? g : 0.0;
        if (below) a.grl[i] = a.grl[i] - a.sigma * g;
        if (above) a.gru[i] = a.gru[i] - a.sigma * g;
    }
}

__global__ void k_unrolled_fill(int n, double *v, double value) {
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) v[j] = value;
}

int ugrid(int n) { return std::max(1, std::min((n + kBlock - 1) / kBlock, 4096)); }
int sgrid(int ntiles) { return std::max(1, std::min((ntiles + kWaves - 1) / kWaves, kMaxGrid)); }

}  // namespace
}  // namespace vk

using namespace vk;

struct vk_hip_unrolled {
    const vk_hip_csr *A = nullptr;
    int m = 0, n = 0;
    long long iterations = 0, every = 1, segments = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    UnrolledArgs a{};
    double *ckpt = nullptr;   // segments x (n + m): x then y of each checkpoint
    // Step slot s: column bits at masks_x + s * wx, row bits (below, above)
    // at masks_y + s * 2 * wy and + wy (wx, wy: 64-bit words of n and m).
    unsigned long long *masks_x = nullptr, *masks_y = nullptr;
    size_t wx = 0, wy = 0;
    bool forward_done = false;
    double last_ms = 0.0;
    std::vector<void *> allocations;

    template <class T>
    // A failed allocation is listed as null, so a scan of `allocations`
    // after a run of alloc calls finds it.
    T *alloc(size_t count) {
        void *p = nullptr;
        if (device_malloc(&p, sizeof(T) * std::max<size_t>(count, 1)) != hipSuccess) p = nullptr;
        allocations.push_back(p);
        return static_cast<T *>(p);
    }
};

namespace {

// Points a's bitsets at step slot `slot` (-1: none recorded).
void set_mask_slot(const vk_hip_unrolled *h, UnrolledArgs &a, long long slot) {
    if (slot < 0) {
        a.mx = a.my_lo = a.my_hi = nullptr;
        return;
    }
    a.mx = h->masks_x + static_cast<size_t>(slot) * h->wx;
    a.my_lo = h->masks_y + static_cast<size_t>(slot) * 2 * h->wy;
    a.my_hi = a.my_lo + h->wy;
}

// Steps [k0, k1) from the current (x, y); masks of step k go to slot k - k0
// when `record`, and (x, y) is saved at every checkpoint boundary when
// `save`.
void forward_steps(vk_hip_unrolled *h, long long k0, long long k1, bool record, bool save) {
    UnrolledArgs a = h->a;
    const CsrView<double> A = h->A->a.view(), AT = h->A->at.view();
    for (long long k = k0; k < k1; ++k) {
        if (save && k % h->every == 0) {
            double *dst = h->ckpt + static_cast<size_t>(k / h->every) * (h->n + h->m);
            (void)hipMemcpyAsync(dst, a.x, sizeof(double) * h->n, hipMemcpyDeviceToDevice, h->stream);
            if (h->m) (void)hipMemcpyAsync(dst + h->n, a.y, sizeof(double) * h->m, hipMemcpyDeviceToDevice, h->stream);
        }
        set_mask_slot(h, a, record ? k - k0 : -1);
        k_unrolled_spmv<FwdPrimalEpi><<<sgrid(AT.ntiles), kBlock, 0, h->stream>>>(a, AT, a.y);
        if (h->m) k_unrolled_spmv<FwdDualEpi><<<sgrid(A.ntiles), kBlock, 0, h->stream>>>(a, A, a.xbar);
    }
}

}  // namespace

extern "C" {

int vk_hip_unrolled_destroy(vk_hip_unrolled *h) {
    if (!h) return VK_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (void *p : h->allocations) (void)hipFree(p);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return VK_OK;
}

int vk_hip_unrolled_create(const vk_hip_csr *A, const double *c, const double *l, const double *u,
                           const double *rl, const double *ru, double tau, double sigma, int64_t iterations,
                           int64_t checkpoint_every, vk_hip_unrolled **out) {
    VK_REQUIRE(out && A, "vk_hip_unrolled_create: null argument");
    *out = nullptr;
    VK_REQUIRE(A->has_values, "vk_hip_unrolled_create: matrix values were never set");
    VK_REQUIRE(A->cols > 0, "vk_hip_unrolled_create: the LP needs at least one variable");
    VK_REQUIRE(c && l && u, "vk_hip_unrolled_create: null column vector");
    VK_REQUIRE(A->rows == 0 || (rl && ru), "vk_hip_unrolled_create: null row vector");
    VK_REQUIRE(tau > 0.0 && sigma > 0.0 && std::isfinite(tau) && std::isfinite(sigma),
               "vk_hip_unrolled_create: step sizes must be positive");
    VK_REQUIRE(iterations >= 0 && checkpoint_every >= 1, "vk_hip_unrolled_create: bad iteration counts");
    if (int rc = ensure_init()) return rc;
    auto *h = new (std::nothrow) vk_hip_unrolled();
    if (!h) return VK_ENOMEM;
    h->A = A;
    h->m = static_cast<int>(A->rows);
    h->n = static_cast<int>(A->cols);
    h->iterations = iterations;
    h->every = std::min<long long>(checkpoint_every, std::max<long long>(1, iterations));
    h->segments = (iterations + h->every - 1) / h->every;
    const int m = h->m, n = h->n;
    auto fail = [&](int rc) {
        vk_hip_unrolled_destroy(h);
        return rc;
    };
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess) {
        set_error("vk_hip_unrolled_create: runtime resource allocation failed");
        return fail(VK_EDEVICE);
    }
    UnrolledArgs &a = h->a;
    a.m = m;
    a.n = n;
    a.tau = tau;
    a.sigma = sigma;
    a.inv_sigma = 1.0 / sigma;
    a.x = h->alloc<double>(n);
    a.xbar = h->alloc<double>(n);
    a.y = h->alloc<double>(m);
    double *cd = h->alloc<double>(n), *ld = h->alloc<double>(n), *ud = h->alloc<double>(n);
    double *rld = h->alloc<double>(m), *rud = h->alloc<double>(m);
    a.c = cd;
    a.l = ld;
    a.u = ud;
    a.rl = rld;
    a.ru = rud;
    a.gx = h->alloc<double>(n);
    a.gz = h->alloc<double>(n);
    a.gc = h->alloc<double>(n);
    a.gy = h->alloc<double>(m);
    a.gw = h->alloc<double>(m);
    a.grl = h->alloc<double>(m);
    a.gru = h->alloc<double>(m);
    h->ckpt = h->alloc<double>(static_cast<size_t>(std::max<long long>(1, h->segments)) * (n + m));
    h->wx = (static_cast<size_t>(n) + 63) / 64;
    h->wy = (static_cast<size_t>(std::max(m, 1)) + 63) / 64;
    h->masks_x = h->alloc<unsigned long long>(static_cast<size_t>(h->every) * h->wx);
    h->masks_y = h->alloc<unsigned long long>(static_cast<size_t>(h->every) * 2 * h->wy);
    for (void *p : h->allocations)
        if (!p) {
            set_error("vk_hip_unrolled_create: device allocation failed");
            return fail(VK_ENOMEM);
        }
    if (hipMemcpy(cd, c, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(ld, l, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(ud, u, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess ||
        (m && hipMemcpy(rld, rl, sizeof(double) * m, hipMemcpyHostToDevice) != hipSuccess) ||
        (m && hipMemcpy(rud, ru, sizeof(double) * m, hipMemcpyHostToDevice) != hipSuccess)) {
        set_error("vk_hip_unrolled_create: upload failed");
        return fail(VK_EDEVICE);
    }
    *out = h;
    return VK_OK;
}

int vk_hip_unrolled_forward(vk_hip_unrolled *h, const double *x0, const double *y0, double *x_out,
                            double *y_out) {
    VK_REQUIRE(h && x0, "vk_hip_unrolled_forward: null argument");
    VK_REQUIRE(h->m == 0 || y0, "vk_hip_unrolled_forward: null y0");
    const UnrolledArgs &a = h->a;
    VK_HIP_CHECK(hipMemcpyAsync(a.x, x0, sizeof(double) * h->n, hipMemcpyHostToDevice, h->stream));
    if (h->m) VK_HIP_CHECK(hipMemcpyAsync(a.y, y0, sizeof(double) * h->m, hipMemcpyHostToDevice, h->stream));
    VK_HIP_CHECK(hipStreamSynchronize(h->stream));
    VK_HIP_CHECK(hipEventRecord(h->ev0, h->stream));
    forward_steps(h, 0, h->iterations, false, true);
    VK_HIP_CHECK(hipGetLastError());
    VK_HIP_CHECK(hipEventRecord(h->ev1, h->stream));
    if (x_out) VK_HIP_CHECK(hipMemcpyAsync(x_out, a.x, sizeof(double) * h->n, hipMemcpyDeviceToHost, h->stream));
    if (y_out && h->m)
        VK_HIP_CHECK(hipMemcpyAsync(y_out, a.y, sizeof(double) * h->m, hipMemcpyDeviceToHost, h->stream));
    VK_HIP_CHECK(hipStreamSynchronize(h->stream));
    float ms = 0.0f;
    VK_HIP_CHECK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->last_ms = ms;
    h->forward_done = true;
    return VK_OK;
}

int vk_hip_unrolled_backward(vk_hip_unrolled *h, const double *grad_x, const double *grad_y, double *grad_c,
                             double *grad_row_lower, double *grad_row_upper) {
    VK_REQUIRE(h, "vk_hip_unrolled_backward: null handle");
    VK_REQUIRE(h->forward_done, "vk_hip_unrolled_backward: run the forward pass first");
    const int m = h->m, n = h->n;
    UnrolledArgs a = h->a;
    hipStream_t st = h->stream;
    // Seeds: adjoints of (x_K, y_K); zero accumulators.
    if (grad_x) VK_HIP_CHECK(hipMemcpyAsync(a.gx, grad_x, sizeof(double) * n, hipMemcpyHostToDevice, st));
    else k_unrolled_fill<<<ugrid(n), kBlock, 0, st>>>(n, a.gx, 0.0);
    if (m) {
        if (grad_y) VK_HIP_CHECK(hipMemcpyAsync(a.gy, grad_y, sizeof(double) * m, hipMemcpyHostToDevice, st));
        else k_unrolled_fill<<<ugrid(m), kBlock, 0, st>>>(m, a.gy, 0.0);
        k_unrolled_fill<<<ugrid(m), kBlock, 0, st>>>(m, a.grl, 0.0);
        k_unrolled_fill<<<ugrid(m), kBlock, 0, st>>>(m, a.gru, 0.0);
    }
    k_unrolled_fill<<<ugrid(n), kBlock, 0, st>>>(n, a.gc, 0.0);
    VK_HIP_CHECK(hipEventRecord(h->ev0, st));
    const CsrView<double> A = h->A->a.view(), AT = h->A->at.view();
    for (long long s = h->segments - 1; s >= 0; --s) {
        const long long k0 = s * h->every, k1 = std::min(h->iterations, k0 + h->every);
        // Recompute the segment from its checkpoint, recording the masks.
        const double *src = h->ckpt + static_cast<size_t>(s) * (n + m);
        VK_HIP_CHECK(hipMemcpyAsync(a.x, src, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
        if (m) VK_HIP_CHECK(hipMemcpyAsync(a.y, src + n, sizeof(double) * m, hipMemcpyDeviceToDevice, st));
        forward_steps(h, k0, k1, true, false);
        for (long long k = k1 - 1; k >= k0; --k) {
            set_mask_slot(h, a, k - k0);
            if (m) k_unrolled_rows<<<ugrid(m), kBlock, 0, st>>>(a);
            // A' has a tile for every column (empty ones included), so the
            // column update runs for all of them even without rows.
            k_unrolled_spmv<BwdPrimalEpi><<<sgrid(AT.ntiles), kBlock, 0, st>>>(a, AT, a.gw);
            if (m) k_unrolled_spmv<BwdDualEpi><<<sgrid(A.ntiles), kBlock, 0, st>>>(a, A, a.gz);
        }
    }
    VK_HIP_CHECK(hipGetLastError());
    VK_HIP_CHECK(hipEventRecord(h->ev1, st));
    if (grad_c) VK_HIP_CHECK(hipMemcpyAsync(grad_c, a.gc, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    if (m && grad_row_lower)
        VK_HIP_CHECK(hipMemcpyAsync(grad_row_lower, a.grl, sizeof(double) * m, hipMemcpyDeviceToHost, st));
    if (m && grad_row_upper)
        VK_HIP_CHECK(hipMemcpyAsync(grad_row_upper, a.gru, sizeof(double) * m, hipMemcpyDeviceToHost, st));
    VK_HIP_CHECK(hipStreamSynchronize(st));
    float ms = 0.0f;
    VK_HIP_CHECK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->last_ms = ms;
    return VK_OK;
}

int vk_hip_unrolled_set_data(vk_hip_unrolled *h, const double *c, const double *rl, const double *ru) {
    VK_REQUIRE(h && c && (h->m == 0 || (rl && ru)), "vk_hip_unrolled_set_data: null argument");
    VK_HIP_CHECK(hipMemcpyAsync(const_cast<double *>(h->a.c), c, sizeof(double) * h->n, hipMemcpyHostToDevice,
                                h->stream));
    if (h->m) {
        VK_HIP_CHECK(hipMemcpyAsync(const_cast<double *>(h->a.rl), rl, sizeof(double) * h->m,
                                    hipMemcpyHostToDevice, h->stream));
        VK_HIP_CHECK(hipMemcpyAsync(const_cast<double *>(h->a.ru), ru, sizeof(double) * h->m,
                                    hipMemcpyHostToDevice, h->stream));
    }
    VK_HIP_CHECK(hipStreamSynchronize(h->stream));
    h->forward_done = false;
    return VK_OK;
}

int vk_hip_unrolled_last_ms(const vk_hip_unrolled *h, double *ms) {
    VK_REQUIRE(h && ms, "vk_hip_unrolled_last_ms: null argument");
    *ms = h->last_ms;
    return VK_OK;
}

}  // extern "C"
