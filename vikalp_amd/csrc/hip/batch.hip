// Batched PDHG: many independent LPs solved per launch by a persistent
// kernel, one 512-thread workgroup (8 wave64s) per instance.
//
// The instances are stored as one block-diagonal CSR matrix (and its stored
// transpose) whose tiles never cross an instance boundary, so a workgroup
// streams only its own instance.  Inside the workgroup one iteration is the
// same fused sequence as the single-LP solver (pdhg.hip): the dual epilogue
// on the instance's rows of A, a workgroup barrier, the primal epilogue on
// its rows of A', then the step / restart / termination rule of the shared
// core (pdhg_core.hpp) run by one thread on the instance's own DevState.
// Each instance therefore keeps its own step sizes, primal weight, restart
// points and status on the device, and stops on its own; the kernel runs a
// bounded number of iterations per launch and the host relaunches until
// every instance has stopped.  This is the shape of the reference's
// branch-and-bound node loop (src/solver/mip/BranchAndBound.cpp:296-297),
// where only bounds and warm starts differ between relaxations, turned into
// one launch instead of one serial solve_lp per node.

#include "runtime.hpp"
#include "pdhg_core.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <new>
#include <vector>

namespace vk {
namespace {

constexpr int kBatchBlock = 512;                  // threads per instance
constexpr int kBatchWaves = kBatchBlock / kWave;  // 8
// Threads per instance once no more instances run than the device has CUs:
// each then gets a CU of its own and twice the waves for its tiles.
constexpr int kWideBlock = 1024;

struct InstanceRange {
    int r0, r1, c0, c1;     // rows / columns of the instance
    int ta0, ta1, tt0, tt1; // tile ranges in A and A'
};

template <class T>
struct BatchArgs {
    Bufs<T> b;
    CsrView<T> A, AT;
    CsrView<double> A64, AT64;   // for the step-size power iteration
    const InstanceRange *inst;
    DevState *states;
    double *pw_rows, *pw_cols;   // power-iteration scratch (FP64)
    int count;
    int power_iterations;
    // 1: every instance runs on the same matrix (rows and columns numbered
    // from 0), with its own slice of the vectors starting at its r0 / c0;
    // 0: the matrix is block-diagonal in global numbering.
    int shared;
    // Capacity (entries of T) of the dynamic LDS buffer the iterate kernel
    // gathers from: an instance with max(rows, columns) <= lds_elems copies
    // x, then y, into it before each SpMV; a larger one gathers from global
    // memory.  0: no buffer.
    int lds_elems;
};

// Workgroup reduction of the 8 slots into out[] (shared memory, readable by
// every thread after the call).  scratch holds kSlots * Block / kWave doubles.
template <int Block = kBatchBlock>
__device__ __forceinline__ void reduce_slots(const Partials &p, double *out, double *scratch) {
    double v[kSlots];
#pragma unroll
    for (int s = 0; s < kSlots; ++s) v[s] = p.v[s];
    const double r = block_reduce_slots<Block>(v, scratch);
    if (threadIdx.x < kSlots) out[threadIdx.x] = r;
    __syncthreads();
}

struct StoreEpi64 {
    double *y;
    struct Pre {};
    __device__ Pre prefetch(int) const { return Pre{}; }
    __device__ void row(int r, double acc, const Pre &, Partials &) const { y[r] = acc; }
};

template <class T>
struct StoreEpiT {
    T *y;
    struct Pre {};
    __device__ Pre prefetch(int) const { return Pre{}; }
    __device__ void row(int r, double acc, const Pre &, Partials &) const { y[r] = static_cast<T>(acc); }
};

// An epilogue seen through a row offset: a shared matrix numbers its rows
// from 0 while the instance's vectors start at `off`.
template <class Epi>
struct Shifted {
    Epi &inner;
    int off;
    using Pre = typename Epi::Pre;
    __device__ Pre prefetch(int r) const { return inner.prefetch(r + off); }
    __device__ void row(int r, double acc, const Pre &pre, Partials &part) const { inner.row(r + off, acc, pre, part); }
};

// One SpMV pass of an instance over tiles [t0, t1) of M, gathering v.  On a
// shared matrix the gathered entries start at vfirst and the pass's rows
// at rfirst; on the block-diagonal matrix both are global already.
template <int Waves = kBatchWaves, class V, class X, class Epi>
__device__ __forceinline__ void instance_spmv(const CsrView<V> &M, const X *v, int vfirst, int rfirst, bool shared,
                                              Epi &epi, Partials &p, double *my_lds, int t0, int t1, int wave) {
    Shifted<Epi> se{epi, shared ? rfirst : 0};
    tile_spmv_range(M, shared ? v + vfirst : v, se, p, my_lds, t0, t1, wave, Waves);
}

// Per instance: project the start, compute A x0 and A' y0, estimate ||A||_2
// by power iteration when the host gave no step size, record the start
// point's KKT error (adaptive policy) and take the first primal step.
template <class T>
__global__ __launch_bounds__(kBatchBlock) void k_batch_setup(BatchArgs<T> a) {
    const int id = blockIdx.x;
    if (id >= a.count) return;
    __shared__ double lds[kBatchWaves][kTileNnz];
    __shared__ double scratch[kSlots * kBatchWaves];
    __shared__ double red_r[kSlots], red_c[kSlots];
    __shared__ DevState S;
    const InstanceRange R = a.inst[id];
    const Bufs<T> &b = a.b;
    const int tid = threadIdx.x, wave = tid / kWave;
    double *my_lds = lds[wave];
    if (tid == 0) S = a.states[id];
    __syncthreads();
    for (int j = R.c0 + tid; j < R.c1; j += kBatchBlock)
        b.x[0][j] = static_cast<T>(clamp_box(static_cast<double>(b.x[0][j]), static_cast<double>(b.l[j]),
                                             static_cast<double>(b.u[j])));
    __syncthreads();
    Partials p;
    p.zero();
    {
        StoreEpiT<T> e{b.ax};
        instance_spmv(a.A, b.x[0], R.c0, R.r0, a.shared, e, p, my_lds, R.ta0, R.ta1, wave);
        StoreEpiT<T> f{b.aty};
        instance_spmv(a.AT, b.y, R.r0, R.c0, a.shared, f, p, my_lds, R.tt0, R.tt1, wave);
    }
    if (S.eta <= 0.0 && S.max_step <= 0.0) {
        // sigma_max(A)^2 ~ ||A'A v|| with ||v|| = 1 (deterministic start)
        for (int j = R.c0 + tid; j < R.c1; j += kBatchBlock) {
            uint64_t z = static_cast<uint64_t>(j - R.c0) * 0x9e3779b97f4a7c15ULL + 0x632be59bd9b4e019ULL;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
            z ^= z >> 31;
            a.pw_cols[j] = static_cast<double>(z >> 11) * 0x1.0p-53 + 0.5;
        }
        __syncthreads();
        double nrm2 = 0.0;
        for (int it = 0; it <= a.power_iterations; ++it) {
            double part = 0.0;
            for (int j = R.c0 + tid; j < R.c1; j += kBatchBlock) part += a.pw_cols[j] * a.pw_cols[j];
            nrm2 = final_sum<kBatchBlock>(part, scratch);
            if (it == a.power_iterations) break;
            const double inv = 1.0 / sqrt(nrm2);
            for (int j = R.c0 + tid; j < R.c1; j += kBatchBlock) a.pw_cols[j] *= inv;
            __syncthreads();
            StoreEpi64 e{a.pw_rows};
            instance_spmv(a.A64, a.pw_cols, R.c0, R.r0, a.shared, e, p, my_lds, R.ta0, R.ta1, wave);
            __syncthreads();
            StoreEpi64 f{a.pw_cols};
            instance_spmv(a.AT64, a.pw_rows, R.r0, R.c0, a.shared, f, p, my_lds, R.tt0, R.tt1, wave);
            __syncthreads();
        }
        if (tid == 0) {
            const double sigma_max = sqrt(sqrt(nrm2));
            const double eta = 0.99 / (sigma_max > 1e-12 ? sigma_max : 1e-12);
            if (S.policy == VK_PDHG_POLICY_REFERENCE) {
                S.max_step = eta;
                S.tau = S.sigma = eta;
            } else {
                S.eta = eta;
                S.tau = eta / S.omega;
                S.sigma = eta * S.omega;
            }
            S.inv_sigma = 1.0 / S.sigma;
        }
    }
    __syncthreads();
    if (S.policy == VK_PDHG_POLICY_ADAPTIVE && S.status == VK_PDHG_RUNNING) {
        for (int j = R.c0 + tid; j < R.c1; j += kBatchBlock) {
            b.xsum[j] = static_cast<double>(b.x[0][j]);
            b.xlast[j] = b.x[0][j];
        }
        for (int i = R.r0 + tid; i < R.r1; i += kBatchBlock) {
            b.ysum[i] = static_cast<double>(b.y[i]);
            b.ylast[i] = b.y[i];
        }
        __syncthreads();
        EvalRowsEpi<T> er{b, 1.0};
        p.zero();
        instance_spmv(a.A, b.xsum, R.c0, R.r0, a.shared, er, p, my_lds, R.ta0, R.ta1, wave);
        reduce_slots(p, red_r, scratch);
        EvalColsEpi<T> ec{b, 1.0};
        p.zero();
        instance_spmv(a.AT, b.ysum, R.r0, R.c0, a.shared, ec, p, my_lds, R.tt0, R.tt1, wave);
        reduce_slots(p, red_c, scratch);
        if (tid == 0) check_step(S, red_r, red_c, true);
        for (int j = R.c0 + tid; j < R.c1; j += kBatchBlock) b.xsum[j] = 0.0;
        for (int i = R.r0 + tid; i < R.r1; i += kBatchBlock) b.ysum[i] = 0.0;
        __syncthreads();
    }
    if (S.policy == VK_PDHG_POLICY_HALPERN) {
        // z0 = (x0, y0): Halpern iterate, anchor and restart point.
        for (int j = R.c0 + tid; j < R.c1; j += kBatchBlock) {
            b.xh[j] = b.x0h[j] = b.xlast[j] = b.x[0][j];
            b.aty0h[j] = b.aty[j];
        }
        for (int i = R.r0 + tid; i < R.r1; i += kBatchBlock) {
            b.yh[i] = b.y0h[i] = b.ylast[i] = b.y[i];
            b.ax0h[i] = b.ax[i];
        }
        __syncthreads();
    }
    if (S.status == VK_PDHG_RUNNING)
        for (int j = R.c0 + tid; j < R.c1; j += kBatchBlock) pstep_element(b, S, j, b.x[0], b.x[1]);
    __syncthreads();
    if (tid == 0) a.states[id] = S;
}

// Up to `iterations` iterations of every instance still running.
// Copies entries [first, last) of v into the LDS buffer and waits for the
// whole workgroup.
template <int Block, class T>
__device__ __forceinline__ void stage_vector(T *lv, const T *v, int first, int last) {
    for (int i = first + static_cast<int>(threadIdx.x); i < last; i += Block) lv[i - first] = v[i];
    __syncthreads();
}

// One SpMV pass over tiles [t0, t1) gathering v (the instance's entries
// start at vfirst, its pass rows at rfirst), from the LDS copy when the
// instance fits.
template <int Waves, class T, class Epi>
__device__ __forceinline__ void batch_pass(const CsrView<T> &M, const T *v, bool in_lds, T *lv, int vfirst,
                                           int rfirst, bool shared, Epi &epi, Partials &p, double *my_lds, int t0,
                                           int t1, int wave) {
    if (in_lds) {
        // The buffer holds entries [vfirst, ...): a shared matrix's indices
        // are already instance-local, a block-diagonal one's are global.
        Shifted<Epi> se{epi, shared ? rfirst : 0};
        tile_spmv_range<T, T, Shifted<Epi>, true>(M, lv, se, p, my_lds, t0, t1, wave, Waves,
                                                  shared ? 0 : vfirst);
    } else {
        instance_spmv<Waves>(M, v, vfirst, rfirst, shared, epi, p, my_lds, t0, t1, wave);
    }
}

// Block = kBatchBlock: two workgroups per CU (4 waves per SIMD, at most 128
// VGPRs, a few dozen spilled): the barriers and the one-thread finalize of
// one instance overlap with the other's tiles.  Measured on the
// 512-instance benchmark: 10.1 s against 10.7 s with one workgroup per CU
// and 180 VGPRs.  Block = kWideBlock once no more instances run than there
// are CUs: each instance's chain of dependent tiles is then split over 16
// waves instead of 8 (same 128-VGPR budget), 8.76 -> 7.65 s (FP64).
template <class T, int Block>
__global__ __launch_bounds__(Block, 4) void k_batch_iterate(BatchArgs<T> a, int iterations) {
    constexpr int Waves = Block / kWave;
    const int id = blockIdx.x;
    if (id >= a.count) return;
    __shared__ double lds[Waves][kTileNnz];
    __shared__ double scratch[kSlots * Waves];
    __shared__ double red_r[kSlots], red_c[kSlots];
    __shared__ DevState S;
    extern __shared__ double gathered[];   // a.lds_elems entries of T
    const InstanceRange R = a.inst[id];
    const Bufs<T> &b = a.b;
    const int tid = threadIdx.x, wave = tid / kWave;
    double *my_lds = lds[wave];
    T *lv = reinterpret_cast<T *>(gathered);
    // Rows and columns share the buffer: x is gathered only by the dual
    // pass and y only by the primal pass, each copied in just before.
    const bool in_lds = R.c1 - R.c0 <= a.lds_elems && R.r1 - R.r0 <= a.lds_elems;
    if (tid == 0) S = a.states[id];
    __syncthreads();
    for (int it = 0; it < iterations; ++it) {
        if (S.status != VK_PDHG_RUNNING) break;
        const long long k = S.k;
        const bool ref = S.policy == VK_PDHG_POLICY_REFERENCE;
        const bool hal = S.policy == VK_PDHG_POLICY_HALPERN;
        const bool fuse = may_fuse(S);
        Partials p;
        p.zero();
        // Every write of the previous iteration is complete here (each path
        // of the loop ends with a workgroup barrier).
        const T *xg = b.x[(k + 1) & 1];
        if (in_lds) stage_vector<Block>(lv, xg, R.c0, R.c1);
        if (hal) {
            HalpernDualEpi<T> de{b, S.sigma, S.inv_sigma, halpern_weights(S.ks), check_iteration(S)};
            batch_pass<Waves>(a.A, xg, in_lds, lv, R.c0, R.r0, a.shared, de, p, my_lds, R.ta0, R.ta1, wave);
        } else {
            DualEpi<T> de{b, S.sigma, S.inv_sigma, ref && S.restart_flag, !ref};
            batch_pass<Waves>(a.A, xg, in_lds, lv, R.c0, R.r0, a.shared, de, p, my_lds, R.ta0, R.ta1, wave);
        }
        // Ends with a barrier: the dual pass no longer reads the buffer and
        // every y+ is written.
        reduce_slots<Block>(p, red_r, scratch);
        if (in_lds) stage_vector<Block>(lv, b.y, R.r0, R.r1);
        p.zero();
        if (hal) {
            HalpernPrimalEpi<T> pe{b, b.x[(k + 1) & 1], b.x[k & 1], S.tau, fuse, check_iteration(S),
                                   halpern_weights(S.ks)};
            batch_pass<Waves>(a.AT, b.y, in_lds, lv, R.r0, R.c0, a.shared, pe, p, my_lds, R.tt0, R.tt1, wave);
        } else {
            PrimalEpi<T> pe{b, b.x[(k + 1) & 1], b.x[k & 1], S.tau, fuse};
            batch_pass<Waves>(a.AT, b.y, in_lds, lv, R.r0, R.c0, a.shared, pe, p, my_lds, R.tt0, R.tt1, wave);
        }
        reduce_slots<Block>(p, red_c, scratch);
        if (tid == 0) finalize_step(S, red_r, red_c);
        __syncthreads();
        if (fuse) continue;
        if (ref) {
            if (S.status == VK_PDHG_RUNNING)
                for (int j = R.c0 + tid; j < R.c1; j += Block)
                    pstep_element(b, S, j, b.x[S.k & 1], b.x[(S.k + 1) & 1]);
            __syncthreads();
            continue;
        }
        if (S.status != VK_PDHG_RUNNING) continue;
        // Primal infeasibility test on this iteration's y+ (A' y+ is in aty,
        // or kept in aty_avg by the Halpern pass).
        p.zero();
        {
            const T *g = hal ? b.aty_avg : b.aty;
            for (int j = R.c0 + tid; j < R.c1; j += Block)
                farkas_column(static_cast<double>(g[j]), static_cast<double>(b.l[j]), static_cast<double>(b.u[j]), p);
        }
        reduce_slots<Block>(p, red_c, scratch);
        if (tid == 0) infeasibility_rule(S, red_c);
        __syncthreads();
        if (S.status != VK_PDHG_RUNNING) continue;
        if (hal) {
            for (int j = R.c0 + tid; j < R.c1; j += Block)
                halpern_apply_column(b, S, j, b.x[S.k & 1], b.x[(S.k + 1) & 1]);
            for (int i = R.r0 + tid; i < R.r1; i += Block) halpern_apply_row(b, S, i);
            __syncthreads();
            continue;
        }
        const double inv_ns = 1.0 / static_cast<double>(S.ns);
        EvalRowsEpi<T> er{b, inv_ns};
        p.zero();
        instance_spmv<Waves>(a.A, b.xsum, R.c0, R.r0, a.shared, er, p, my_lds, R.ta0, R.ta1, wave);
        reduce_slots<Block>(p, red_r, scratch);
        EvalColsEpi<T> ec{b, inv_ns};
        p.zero();
        instance_spmv<Waves>(a.AT, b.ysum, R.r0, R.c0, a.shared, ec, p, my_lds, R.tt0, R.tt1, wave);
        reduce_slots<Block>(p, red_c, scratch);
        if (tid == 0) check_step(S, red_r, red_c, false);
        __syncthreads();
        if (S.status == VK_PDHG_RUNNING || S.final_from_avg) {
            for (int j = R.c0 + tid; j < R.c1; j += Block)
                apply_column(b, S, j, b.x[S.k & 1], b.x[(S.k + 1) & 1]);
            for (int i = R.r0 + tid; i < R.r1; i += Block) apply_row(b, S, i);
        }
        __syncthreads();
    }
    if (tid == 0) a.states[id] = S;
}

// *bad = 1 when a row of the block-diagonal batch matrix touches a column
// outside its instance's range (instance of row r: the last i with
// row_starts[i] <= r, which skips empty instances).
__global__ void k_check_instances(int rows, int count, const int *__restrict__ rp, const int *__restrict__ ci,
                                  const int64_t *__restrict__ row_starts, const int64_t *__restrict__ col_starts,
                                  int *__restrict__ bad) {
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += gridDim.x * blockDim.x) {
        int lo = 0, hi = count - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (row_starts[mid] <= r) lo = mid;
            else hi = mid - 1;
        }
        const int64_t c0 = col_starts[lo], c1 = col_starts[lo + 1];
        for (int k = rp[r]; k < rp[r + 1]; ++k)
            if (ci[k] < c0 || ci[k] >= c1) {
                *bad = 1;
                break;
            }
    }
}

__global__ void k_to_f32_batch(int64_t n, const double *__restrict__ src, float *__restrict__ dst) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = static_cast<float>(src[i]);
}

__global__ void k_to_f64_batch(int64_t n, const float *__restrict__ src, double *__restrict__ dst) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = static_cast<double>(src[i]);
}

int bgrid(int64_t n) { return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192))); }

}  // namespace
}  // namespace vk

using namespace vk;

struct vk_hip_batch {
    vk_pdhg_params params{};
    int count = 0;
    int64_t m = 0, n = 0, nnz = 0;     // stacked vector lengths; the matrix's nonzeros
    int64_t mat_rows = 0, mat_cols = 0;  // the stored matrix (one instance's when shared)
    bool shared = false;
    bool mixed = false;
    hipStream_t stream = nullptr;
    std::vector<void *> allocations;
    int *rp = nullptr, *ci = nullptr, *trp = nullptr, *tci = nullptr;
    double *val = nullptr, *tval = nullptr;
    float *val32 = nullptr, *tval32 = nullptr;
    int2 *tiles_a = nullptr, *tiles_at = nullptr;
    int ntiles_a = 0, ntiles_at = 0;
    InstanceRange *inst = nullptr;
    std::vector<InstanceRange> inst_host;
    DevState *states = nullptr;
    std::vector<DevState> states_host;
    BatchArgs<double> a64{};
    BatchArgs<float> a32{};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    int running = 0;     // instances running after the last launch
    int cus = 0;         // compute units of the device
    bool wide = false;   // the kWideBlock kernel fits (LDS) for this batch's precision

    template <class T>
    // A failed allocation is listed as null, so a scan of `allocations`
    // after a run of alloc calls finds it.
    T *alloc(size_t count_) {
        void *p = nullptr;
        if (device_malloc(&p, sizeof(T) * std::max<size_t>(count_, 1)) != hipSuccess) p = nullptr;
        allocations.push_back(p);
        return static_cast<T *>(p);
    }
};

namespace {

template <class T>
int setup_batch_bufs(vk_hip_batch *h, BatchArgs<T> &a, const T *val_a, const T *val_at,
                     const std::vector<double> *host_vectors) {
    const int64_t m = h->m, n = h->n;
    Bufs<T> &b = a.b;
    std::memset(&b, 0, sizeof(b));
    b.m = static_cast<int>(m);
    b.n = static_cast<int>(n);
    b.x[0] = h->alloc<T>(n);
    b.x[1] = h->alloc<T>(n);
    b.y = h->alloc<T>(m);
    b.ax = h->alloc<T>(m);
    b.aty = h->alloc<T>(n);
    T *c = h->alloc<T>(n), *l = h->alloc<T>(n), *u = h->alloc<T>(n), *rl = h->alloc<T>(m), *ru = h->alloc<T>(m);
    T *runs = h->alloc<T>(m), *cs = h->alloc<T>(n);
    b.c = c; b.l = l; b.u = u; b.rl = rl; b.ru = ru; b.row_unscale = runs; b.col_scale = cs;
    if (h->params.policy == VK_PDHG_POLICY_ADAPTIVE) {
        b.xsum = h->alloc<double>(n);
        b.ysum = h->alloc<double>(m);
    }
    if (h->params.policy != VK_PDHG_POLICY_REFERENCE) {
        b.xlast = h->alloc<T>(n);
        b.ylast = h->alloc<T>(m);
        b.ax_avg = h->alloc<T>(m);
        b.aty_avg = h->alloc<T>(n);
    }
    if (h->params.policy == VK_PDHG_POLICY_HALPERN) {
        b.xh = h->alloc<T>(n);
        b.yh = h->alloc<T>(m);
        b.x0h = h->alloc<T>(n);
        b.y0h = h->alloc<T>(m);
        b.ax0h = h->alloc<T>(m);
        b.aty0h = h->alloc<T>(n);
    }
    for (void *p : h->allocations)
        if (!p) {
            set_error("vk_hip_batch_create: device allocation failed");
            return VK_ENOMEM;
        }
    // host_vectors: c, l, u, rl, ru, 1/r, s, x0, y0 (FP64)
    T *dst[9] = {c, l, u, rl, ru, runs, cs, b.x[0], b.y};
    // The FP64 staging buffer is one of the handle's allocations (freed with
    // it), so an upload that fails part way leaks nothing.
    double *stage = h->alloc<double>(static_cast<size_t>(std::max<int64_t>(std::max(m, n), 1)));
    if (!stage) {
        set_error("vk_hip_batch_create: device allocation failed");
        return VK_ENOMEM;
    }
    for (int v = 0; v < 9; ++v) {
        const auto &src = host_vectors[v];
        if (src.empty()) continue;
        if constexpr (sizeof(T) == sizeof(double)) {
            VK_HIP_CHECK(hipMemcpyAsync(dst[v], src.data(), sizeof(double) * src.size(), hipMemcpyHostToDevice, h->stream));
        } else {
            VK_HIP_CHECK(hipMemcpyAsync(stage, src.data(), sizeof(double) * src.size(), hipMemcpyHostToDevice, h->stream));
            k_to_f32_batch<<<bgrid(src.size()), 256, 0, h->stream>>>(static_cast<int64_t>(src.size()), stage,
                                                                    reinterpret_cast<float *>(dst[v]));
        }
        VK_HIP_CHECK(hipStreamSynchronize(h->stream));
    }
    const int mr = static_cast<int>(h->mat_rows), mc = static_cast<int>(h->mat_cols);
    a.A = CsrView<T>{h->rp, h->ci, val_a, h->tiles_a, mr, h->ntiles_a, 0};
    a.AT = CsrView<T>{h->trp, h->tci, val_at, h->tiles_at, mc, h->ntiles_at, 0};
    a.A64 = CsrView<double>{h->rp, h->ci, h->val, h->tiles_a, mr, h->ntiles_a, 0};
    a.AT64 = CsrView<double>{h->trp, h->tci, h->tval, h->tiles_at, mc, h->ntiles_at, 0};
    a.shared = h->shared ? 1 : 0;
    a.inst = h->inst;
    a.states = h->states;
    a.pw_rows = h->alloc<double>(m);
    a.pw_cols = h->alloc<double>(n);
    a.count = h->count;
    a.power_iterations = 100;
    if (!a.pw_rows || !a.pw_cols) {
        set_error("vk_hip_batch_create: device allocation failed");
        return VK_ENOMEM;
    }
    // Dynamic LDS for the gathered vector: what the largest instance needs,
    // up to what half the device's LDS leaves beside the kernel's static LDS.
    int max_dim = 0;
    for (const auto &R : h->inst_host) max_dim = std::max({max_dim, R.c1 - R.c0, R.r1 - R.r0});
    int dev = 0, per_block = 0, optin = 0;
    VK_HIP_CHECK(hipGetDevice(&dev));
    VK_HIP_CHECK(hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) != hipSuccess) optin = 0;
    hipFuncAttributes fa{};
    VK_HIP_CHECK(hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(&k_batch_iterate<T, kBatchBlock>)));
    auto elems_for = [&](int limit) {
        const long long room = static_cast<long long>(limit) - static_cast<long long>(fa.sharedSizeBytes);
        return room > 0 ? static_cast<int>(std::min<long long>(max_dim, room / static_cast<long long>(sizeof(T)))) : 0;
    };
    // Half the LDS per workgroup, so two workgroups fit a CU.
    a.lds_elems = elems_for(std::max(per_block, optin) / 2);
    if (a.lds_elems > elems_for(per_block) &&
        hipFuncSetAttribute(reinterpret_cast<const void *>(&k_batch_iterate<T, kBatchBlock>),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            static_cast<int>(sizeof(T)) * a.lds_elems) != hipSuccess) {
        (void)hipGetLastError();   // no opt-in: stay within the default per-block limit
        a.lds_elems = elems_for(per_block);
    }
    // The wide kernel (one instance per CU) takes the same dynamic buffer
    // beside its larger static LDS, or is not used.
    hipFuncAttributes fw{};
    VK_HIP_CHECK(hipFuncGetAttributes(&fw, reinterpret_cast<const void *>(&k_batch_iterate<T, kWideBlock>)));
    const long long wide_lds = static_cast<long long>(fw.sharedSizeBytes) +
                               static_cast<long long>(sizeof(T)) * a.lds_elems;
    h->wide = fw.maxThreadsPerBlock >= kWideBlock && wide_lds <= std::max(per_block, optin);
    if (h->wide && wide_lds > per_block &&
        hipFuncSetAttribute(reinterpret_cast<const void *>(&k_batch_iterate<T, kWideBlock>),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            static_cast<int>(sizeof(T)) * a.lds_elems) != hipSuccess) {
        (void)hipGetLastError();
        h->wide = false;
    }
    VK_HIP_CHECK(hipDeviceGetAttribute(&h->cus, hipDeviceAttributeMultiprocessorCount, dev));
    h->running = h->count;
    k_batch_setup<T><<<h->count, kBatchBlock, 0, h->stream>>>(a);
    VK_HIP_CHECK(hipGetLastError());
    VK_HIP_CHECK(hipStreamSynchronize(h->stream));
    return VK_OK;
}

int read_states(vk_hip_batch *h) {
    VK_HIP_CHECK(hipMemcpyAsync(h->states_host.data(), h->states, sizeof(DevState) * h->count,
                                hipMemcpyDeviceToHost, h->stream));
    VK_HIP_CHECK(hipStreamSynchronize(h->stream));
    return VK_OK;
}

// The matrix of a batch: host CSR arrays (int64), or a device CSR handle
// whose values are already set (and usually scaled on the device).
struct MatrixSource {
    const int64_t *row_offsets = nullptr, *column_indices = nullptr;
    const double *values = nullptr;
    const vk_hip_csr *A = nullptr;
    bool shared = false;   // A is every instance's matrix (equal instance sizes)
};

int batch_create(int count, const int64_t *row_starts, const int64_t *col_starts, const MatrixSource &src,
                 const double *c, const double *l, const double *u, const double *rl, const double *ru,
                 const double *row_scale, const double *col_scale, const double *x0, const double *y0,
                 const double *step_sizes, const double *primal_weights, const double *norm_b,
                 const double *norm_c, const int64_t *iteration_limits, const vk_pdhg_params *params,
                 vk_hip_batch **out) {
    VK_REQUIRE(out && params && count > 0 && row_starts && col_starts && (src.row_offsets || src.A),
               "vk_hip_batch_create: bad arguments");
    *out = nullptr;
    VK_REQUIRE(params->policy == VK_PDHG_POLICY_REFERENCE || params->policy == VK_PDHG_POLICY_ADAPTIVE ||
                   params->policy == VK_PDHG_POLICY_HALPERN,
               "vk_hip_batch_create: unknown policy");
    VK_REQUIRE(params->precision == VK_PREC_FP64 || params->precision == VK_PREC_MIXED,
               "vk_hip_batch_create: unknown precision");
    VK_REQUIRE(params->tile_geometry >= 0 && params->tile_geometry < kTileGeometries,
               "vk_hip_batch_create: tile_geometry must be 0..3");
    VK_REQUIRE(params->recentre == 0, "vk_hip_batch_create: the batched solver has no recentred FP32 iteration");
    VK_REQUIRE(params->primal_weight_gain >= 0.0 && params->primal_weight_gain <= 1.0 &&
                   params->primal_weight_integral >= 0.0 && params->primal_weight_integral <= 1.0,
               "vk_hip_batch_create: primal-weight gains must be in [0, 1]");
    VK_REQUIRE(params->iteration_limit >= 0 && params->restart_period > 0,
               "vk_hip_batch_create: bad iteration limit or restart period");
    const int64_t m = row_starts[count], n = col_starts[count];
    VK_REQUIRE(row_starts[0] == 0 && col_starts[0] == 0 && m < INT32_MAX && n < INT32_MAX,
               "vk_hip_batch_create: instance ranges must start at 0 and fit int32");
    VK_REQUIRE(n > 0 && c && l && u && col_scale, "vk_hip_batch_create: null column vector");
    VK_REQUIRE(m == 0 || (rl && ru && row_scale), "vk_hip_batch_create: null row vector");
    if (iteration_limits)
        for (int i = 0; i < count; ++i)
            VK_REQUIRE(iteration_limits[i] >= 0, "vk_hip_batch_create: negative instance iteration limit");
    for (int i = 0; i < count; ++i)
        VK_REQUIRE(row_starts[i + 1] >= row_starts[i] && col_starts[i + 1] > col_starts[i],
                   "vk_hip_batch_create: every instance needs >= 1 column and ordered ranges");
    int64_t nnz = 0, mat_rows = m, mat_cols = n;
    if (src.A && src.shared) {
        mat_rows = src.A->rows;
        mat_cols = src.A->cols;
        for (int i = 0; i < count; ++i)
            VK_REQUIRE(row_starts[i + 1] - row_starts[i] == mat_rows && col_starts[i + 1] - col_starts[i] == mat_cols,
                       "vk_hip_batch_create: every instance must have the shared matrix's dimensions");
        VK_REQUIRE(src.A->has_values, "vk_hip_batch_create: matrix values not set");
        nnz = src.A->nnz;
    } else if (src.A) {
        VK_REQUIRE(src.A->rows == m && src.A->cols == n, "vk_hip_batch_create: matrix dimensions differ from the ranges");
        VK_REQUIRE(src.A->has_values, "vk_hip_batch_create: matrix values not set");
        nnz = src.A->nnz;
    } else {
        const int64_t *row_offsets = src.row_offsets, *column_indices = src.column_indices;
        nnz = row_offsets[m];
        VK_REQUIRE(nnz >= 0 && nnz < INT32_MAX && row_offsets[0] == 0, "vk_hip_batch_create: bad row offsets");
        VK_REQUIRE(nnz == 0 || (column_indices && src.values), "vk_hip_batch_create: null matrix arrays");
        for (int i = 0; i < count; ++i)
            for (int64_t r = row_starts[i]; r < row_starts[i + 1]; ++r) {
                VK_REQUIRE(row_offsets[r] <= row_offsets[r + 1] && row_offsets[r + 1] <= nnz,
                           "vk_hip_batch_create: row offsets must be ordered and in range");
                int64_t prev = -1;
                for (int64_t k = row_offsets[r]; k < row_offsets[r + 1]; ++k) {
                    const int64_t col = column_indices[k];
                    VK_REQUIRE(col >= col_starts[i] && col < col_starts[i + 1] && col > prev,
                               "vk_hip_batch_create: a row leaves its instance's columns or is unordered");
                    prev = col;
                }
            }
    }
    if (int rc = ensure_init()) return rc;
    auto *h = new (std::nothrow) vk_hip_batch();
    if (!h) return VK_ENOMEM;
    auto fail = [&](int rc) {
        vk_hip_batch_destroy(h);
        return rc;
    };
    h->params = *params;
    h->count = count;
    h->m = m;
    h->n = n;
    h->nnz = nnz;
    h->mat_rows = mat_rows;
    h->mat_cols = mat_cols;
    h->shared = src.shared;
    h->mixed = params->precision == VK_PREC_MIXED;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess) {
        set_error("vk_hip_batch_create: stream/event creation failed");
        return fail(VK_EDEVICE);
    }
    h->rp = h->alloc<int>(mat_rows + 1);
    h->ci = h->alloc<int>(nnz);
    h->trp = h->alloc<int>(mat_cols + 1);
    h->tci = h->alloc<int>(nnz);
    h->val = h->alloc<double>(nnz);
    h->tval = h->alloc<double>(nnz);
    h->inst = h->alloc<InstanceRange>(count);
    h->states = h->alloc<DevState>(count);
    if (h->mixed) {
        h->val32 = h->alloc<float>(nnz);
        h->tval32 = h->alloc<float>(nnz);
    }
    for (void *p : h->allocations)
        if (!p) {
            set_error("vk_hip_batch_create: device allocation failed");
            return fail(VK_ENOMEM);
        }
    // Host copies of both offset arrays: the tile tables are cut on the host.
    std::vector<int> rp(static_cast<size_t>(mat_rows + 1)), trp(static_cast<size_t>(mat_cols + 1));
    if (src.A) {
        // The handle's A and A' (its transpose lists each column's rows in
        // increasing order, as the host counting sort below does); a device
        // check that every row stays inside its instance's columns.
        const vk_hip_csr *A = src.A;
        int64_t *starts = h->alloc<int64_t>(2 * (static_cast<size_t>(count) + 1));
        int *bad = h->alloc<int>(1);
        if (!starts || !bad) {
            set_error("vk_hip_batch_create: device allocation failed");
            return fail(VK_ENOMEM);
        }
        hipStream_t st = h->stream;
        const size_t sb = sizeof(int64_t) * (static_cast<size_t>(count) + 1);
        if (hipMemcpyAsync(starts, row_starts, sb, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(starts + count + 1, col_starts, sb, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemsetAsync(bad, 0, sizeof(int), st) != hipSuccess) {
            set_error("vk_hip_batch_create: upload failed");
            return fail(VK_EDEVICE);
        }
        if (m > 0 && !src.shared)
            k_check_instances<<<bgrid(m), 256, 0, st>>>(static_cast<int>(m), count, A->a.row_ptr, A->a.col, starts,
                                                        starts + count + 1, bad);
        int bad_host = 0;
        const auto d2d = hipMemcpyDeviceToDevice;
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(h->rp, A->a.row_ptr, sizeof(int) * (mat_rows + 1), d2d, st) != hipSuccess ||
            hipMemcpyAsync(h->trp, A->at.row_ptr, sizeof(int) * (mat_cols + 1), d2d, st) != hipSuccess ||
            (nnz && (hipMemcpyAsync(h->ci, A->a.col, sizeof(int) * nnz, d2d, st) != hipSuccess ||
                     hipMemcpyAsync(h->tci, A->at.col, sizeof(int) * nnz, d2d, st) != hipSuccess ||
                     hipMemcpyAsync(h->val, A->a.val, sizeof(double) * nnz, d2d, st) != hipSuccess ||
                     hipMemcpyAsync(h->tval, A->at.val, sizeof(double) * nnz, d2d, st) != hipSuccess)) ||
            hipMemcpyAsync(rp.data(), A->a.row_ptr, sizeof(int) * (mat_rows + 1), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(trp.data(), A->at.row_ptr, sizeof(int) * (mat_cols + 1), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(&bad_host, bad, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            set_error("vk_hip_batch_create: device copy failed");
            return fail(VK_EDEVICE);
        }
        if (bad_host) {
            set_error("vk_hip_batch_create: a row leaves its instance's columns");
            return fail(VK_EINVAL);
        }
    } else {
        // Host CSR (int32) and its transpose by a stable counting sort.
        const int64_t *row_offsets = src.row_offsets, *column_indices = src.column_indices;
        std::vector<int> ci(static_cast<size_t>(nnz)), tci(static_cast<size_t>(nnz));
        for (int64_t r = 0; r <= m; ++r) rp[r] = static_cast<int>(row_offsets[r]);
        for (int64_t k = 0; k < nnz; ++k) ci[k] = static_cast<int>(column_indices[k]);
        std::fill(trp.begin(), trp.end(), 0);
        std::vector<double> tv(static_cast<size_t>(nnz));
        for (int64_t k = 0; k < nnz; ++k) ++trp[ci[k] + 1];
        for (int64_t j = 0; j < n; ++j) trp[j + 1] += trp[j];
        {
            std::vector<int> next(trp.begin(), trp.end() - 1);
            for (int64_t r = 0; r < m; ++r)
                for (int k = rp[r]; k < rp[r + 1]; ++k) {
                    const int pos = next[ci[k]]++;
                    tci[pos] = static_cast<int>(r);
                    tv[pos] = src.values[k];
                }
        }
        auto up = [&](void *dst, const void *from, size_t bytes) {
            return bytes ? hipMemcpyAsync(dst, from, bytes, hipMemcpyHostToDevice, h->stream) : hipSuccess;
        };
        if (up(h->rp, rp.data(), sizeof(int) * rp.size()) != hipSuccess ||
            up(h->ci, ci.data(), sizeof(int) * ci.size()) != hipSuccess ||
            up(h->trp, trp.data(), sizeof(int) * trp.size()) != hipSuccess ||
            up(h->tci, tci.data(), sizeof(int) * tci.size()) != hipSuccess ||
            up(h->val, src.values, sizeof(double) * nnz) != hipSuccess ||
            up(h->tval, tv.data(), sizeof(double) * nnz) != hipSuccess ||
            hipStreamSynchronize(h->stream) != hipSuccess) {
            set_error("vk_hip_batch_create: upload failed");
            return fail(VK_EDEVICE);
        }
    }
    std::vector<int2> ta, tt;
    const TileGeometry &geom = kTileGeometry[params->tile_geometry];
    h->inst_host.resize(static_cast<size_t>(count));
    if (src.shared) {
        // One tile table over the shared matrix, used by every instance.
        append_tiles(rp.data(), 0, static_cast<int>(mat_rows), ta, geom);
        append_tiles(trp.data(), 0, static_cast<int>(mat_cols), tt, geom);
        ta.pop_back();
        tt.pop_back();
    }
    for (int i = 0; i < count; ++i) {
        InstanceRange &R = h->inst_host[i];
        R.r0 = static_cast<int>(row_starts[i]);
        R.r1 = static_cast<int>(row_starts[i + 1]);
        R.c0 = static_cast<int>(col_starts[i]);
        R.c1 = static_cast<int>(col_starts[i + 1]);
        if (src.shared) {
            R.ta0 = 0;
            R.ta1 = static_cast<int>(ta.size());
            R.tt0 = 0;
            R.tt1 = static_cast<int>(tt.size());
            continue;
        }
        // Each range contributes its tile starts; its end is the next
        // range's start (ranges are contiguous) or the final sentinel.
        R.ta0 = static_cast<int>(ta.size());
        std::vector<int2> tmp;
        append_tiles(rp.data(), R.r0, R.r1, tmp, geom);
        R.ta1 = R.ta0 + static_cast<int>(tmp.size()) - 1;
        ta.insert(ta.end(), tmp.begin(), tmp.end() - 1);
        R.tt0 = static_cast<int>(tt.size());
        tmp.clear();
        append_tiles(trp.data(), R.c0, R.c1, tmp, geom);
        R.tt1 = R.tt0 + static_cast<int>(tmp.size()) - 1;
        tt.insert(tt.end(), tmp.begin(), tmp.end() - 1);
    }
    ta.push_back(make_int2(static_cast<int>(mat_rows), static_cast<int>(nnz)));
    tt.push_back(make_int2(static_cast<int>(mat_cols), static_cast<int>(nnz)));
    h->ntiles_a = static_cast<int>(ta.size()) - 1;
    h->ntiles_at = static_cast<int>(tt.size()) - 1;
    h->tiles_a = h->alloc<int2>(ta.size());
    h->tiles_at = h->alloc<int2>(tt.size());
    if (!h->tiles_a || !h->tiles_at) {
        set_error("vk_hip_batch_create: device allocation failed");
        return fail(VK_ENOMEM);
    }
    auto up = [&](void *dst, const void *from, size_t bytes) {
        return bytes ? hipMemcpyAsync(dst, from, bytes, hipMemcpyHostToDevice, h->stream) : hipSuccess;
    };
    if (up(h->tiles_a, ta.data(), sizeof(int2) * ta.size()) != hipSuccess ||
        up(h->tiles_at, tt.data(), sizeof(int2) * tt.size()) != hipSuccess ||
        up(h->inst, h->inst_host.data(), sizeof(InstanceRange) * count) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess) {
        set_error("vk_hip_batch_create: upload failed");
        return fail(VK_EDEVICE);
    }
    if (h->mixed && nnz > 0) {
        k_to_f32_batch<<<bgrid(nnz), 256, 0, h->stream>>>(nnz, h->val, h->val32);
        k_to_f32_batch<<<bgrid(nnz), 256, 0, h->stream>>>(nnz, h->tval, h->tval32);
    }
    // Per-instance state.
    h->states_host.assign(static_cast<size_t>(count), DevState{});
    for (int i = 0; i < count; ++i) {
        DevState &S = h->states_host[i];
        S.limit = iteration_limits ? iteration_limits[i] : params->iteration_limit;
        S.period = params->policy == VK_PDHG_POLICY_REFERENCE ? kDecisionPeriod : params->restart_period;
        S.status = S.limit == 0 ? VK_PDHG_ITERATION_LIMIT : VK_PDHG_RUNNING;
        S.policy = params->policy;
        S.metric_tol = params->metric_tolerance;
        S.rel_tol = params->relative_kkt_tolerance;
        S.norm_b = norm_b ? norm_b[i] : 0.0;
        S.norm_c = norm_c ? norm_c[i] : 0.0;
        S.prev_metric = S.metric = S.rel_kkt = std::numeric_limits<double>::infinity();
        S.kkt_ref = S.kkt_prev = std::numeric_limits<double>::infinity();
        const double step = step_sizes ? step_sizes[i] : 0.0;   // <= 0: power iteration
        S.omega = (params->policy != VK_PDHG_POLICY_REFERENCE && primal_weights) ? primal_weights[i] : 1.0;
        if (S.omega <= 0.0 || !std::isfinite(S.omega)) S.omega = 1.0;
        if (params->policy == VK_PDHG_POLICY_REFERENCE) {
            S.max_step = step;
            S.tau = S.sigma = step;
        } else {
            S.eta = step;
            S.tau = step / S.omega;
            S.sigma = step * S.omega;
        }
        S.inv_sigma = step > 0.0 ? 1.0 / S.sigma : 0.0;
        S.pw_kp = params->primal_weight_gain > 0.0 ? params->primal_weight_gain : 0.5;
        S.pw_ki = params->primal_weight_integral;
    }
    if (up(h->states, h->states_host.data(), sizeof(DevState) * count) != hipSuccess) {
        set_error("vk_hip_batch_create: state upload failed");
        return fail(VK_EDEVICE);
    }
    std::vector<double> hv[9];
    hv[0].assign(c, c + n);
    hv[1].assign(l, l + n);
    hv[2].assign(u, u + n);
    if (m > 0) {
        hv[3].assign(rl, rl + m);
        hv[4].assign(ru, ru + m);
        hv[5].resize(static_cast<size_t>(m));
        for (int64_t i = 0; i < m; ++i) hv[5][i] = 1.0 / row_scale[i];
    }
    hv[6].assign(col_scale, col_scale + n);
    hv[7] = x0 ? std::vector<double>(x0, x0 + n) : std::vector<double>(static_cast<size_t>(n), 0.0);
    if (m > 0) hv[8] = y0 ? std::vector<double>(y0, y0 + m) : std::vector<double>(static_cast<size_t>(m), 0.0);
    const int rc = h->mixed ? setup_batch_bufs<float>(h, h->a32, h->val32, h->tval32, hv)
                            : setup_batch_bufs<double>(h, h->a64, h->val, h->tval, hv);
    if (rc) return fail(rc);
    *out = h;
    return VK_OK;
}


}  // namespace

extern "C" {

int vk_hip_batch_destroy(vk_hip_batch *h) {
    if (!h) return VK_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (void *p : h->allocations) (void)hipFree(p);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return VK_OK;
}


int vk_hip_batch_create(int count, const int64_t *row_starts, const int64_t *col_starts,
                        const int64_t *row_offsets, const int64_t *column_indices, const double *values,
                        const double *c, const double *l, const double *u, const double *rl,
                        const double *ru, const double *row_scale, const double *col_scale,
                        const double *x0, const double *y0, const double *step_sizes,
                        const double *primal_weights, const double *norm_b, const double *norm_c,
                        const int64_t *iteration_limits, const vk_pdhg_params *params,
                        vk_hip_batch **out) {
    MatrixSource src;
    src.row_offsets = row_offsets;
    src.column_indices = column_indices;
    src.values = values;
    return batch_create(count, row_starts, col_starts, src, c, l, u, rl, ru, row_scale, col_scale, x0, y0,
                        step_sizes, primal_weights, norm_b, norm_c, iteration_limits, params, out);
}

int vk_hip_batch_create_shared(int count, const vk_hip_csr *A, const double *c, const double *l, const double *u,
                               const double *rl, const double *ru, const double *row_scale,
                               const double *col_scale, const double *x0, const double *y0,
                               const double *step_sizes, const double *primal_weights, const double *norm_b,
                               const double *norm_c, const int64_t *iteration_limits,
                               const vk_pdhg_params *params, vk_hip_batch **out) {
    VK_REQUIRE(A && count > 0, "vk_hip_batch_create_shared: null matrix or no instance");
    std::vector<int64_t> rs(static_cast<size_t>(count) + 1), cs(static_cast<size_t>(count) + 1);
    for (int i = 0; i <= count; ++i) {
        rs[i] = static_cast<int64_t>(i) * A->rows;
        cs[i] = static_cast<int64_t>(i) * A->cols;
    }
    MatrixSource src;
    src.A = A;
    src.shared = true;
    return batch_create(count, rs.data(), cs.data(), src, c, l, u, rl, ru, row_scale, col_scale, x0, y0, step_sizes,
                        primal_weights, norm_b, norm_c, iteration_limits, params, out);
}

int vk_hip_batch_create_csr(int count, const int64_t *row_starts, const int64_t *col_starts,
                            const vk_hip_csr *A, const double *c, const double *l, const double *u,
                            const double *rl, const double *ru, const double *row_scale,
                            const double *col_scale, const double *x0, const double *y0,
                            const double *step_sizes, const double *primal_weights, const double *norm_b,
                            const double *norm_c, const int64_t *iteration_limits,
                            const vk_pdhg_params *params, vk_hip_batch **out) {
    VK_REQUIRE(A, "vk_hip_batch_create_csr: null matrix");
    MatrixSource src;
    src.A = A;
    return batch_create(count, row_starts, col_starts, src, c, l, u, rl, ru, row_scale, col_scale, x0, y0,
                        step_sizes, primal_weights, norm_b, norm_c, iteration_limits, params, out);
}

int vk_hip_batch_run(vk_hip_batch *h, int64_t iterations_per_launch, int64_t max_launches, int *running) {
    VK_REQUIRE(h && running && iterations_per_launch > 0 && max_launches > 0, "vk_hip_batch_run: bad arguments");
    VK_HIP_CHECK(hipEventRecord(h->ev0, h->stream));
    int alive = 0;
    for (int64_t launch = 0; launch < max_launches; ++launch) {
        // Instances that have stopped leave at once, so with at most one
        // running instance per CU the wide kernel gives each a CU.
        const bool wide = h->wide && h->running <= h->cus;
        const int its = static_cast<int>(iterations_per_launch);
        if (h->mixed) {
            const size_t lds = sizeof(float) * h->a32.lds_elems;
            if (wide) k_batch_iterate<float, kWideBlock><<<h->count, kWideBlock, lds, h->stream>>>(h->a32, its);
            else k_batch_iterate<float, kBatchBlock><<<h->count, kBatchBlock, lds, h->stream>>>(h->a32, its);
        } else {
            const size_t lds = sizeof(double) * h->a64.lds_elems;
            if (wide) k_batch_iterate<double, kWideBlock><<<h->count, kWideBlock, lds, h->stream>>>(h->a64, its);
            else k_batch_iterate<double, kBatchBlock><<<h->count, kBatchBlock, lds, h->stream>>>(h->a64, its);
        }
        VK_HIP_CHECK(hipGetLastError());
        if (int rc = read_states(h)) return rc;
        alive = 0;
        for (const auto &S : h->states_host) alive += S.status == VK_PDHG_RUNNING ? 1 : 0;
        h->running = alive;
        if (alive == 0) break;
    }
    VK_HIP_CHECK(hipEventRecord(h->ev1, h->stream));
    VK_HIP_CHECK(hipEventSynchronize(h->ev1));
    float ms = 0.0f;
    VK_HIP_CHECK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->last_ms = ms;
    *running = alive;
    return VK_OK;
}

int vk_hip_batch_read_state(vk_hip_batch *h, int instance, vk_pdhg_state *out) {
    VK_REQUIRE(h && out && instance >= 0 && instance < h->count, "vk_hip_batch_read_state: bad arguments");
    if (int rc = read_states(h)) return rc;
    const DevState &d = h->states_host[instance];
    std::memset(out, 0, sizeof(*out));
    out->iterations = d.k;
    out->status = d.status;
    out->restarts = d.restarts;
    out->tau = d.tau;
    out->sigma = d.sigma;
    out->primal_weight = d.omega;
    out->metric = d.metric;
    out->primal_scaled = d.rp;
    out->dual_scaled = d.rd;
    out->rel_primal = d.rel_p;
    out->rel_dual = d.rel_d;
    out->rel_gap = d.rel_g;
    out->rel_kkt = d.rel_kkt;
    out->primal_objective = d.pobj;
    out->dual_objective = d.dobj;
    out->converged_on_average = d.final_from_avg;
    return VK_OK;
}

int vk_hip_batch_download(vk_hip_batch *h, double *x, double *y) {
    VK_REQUIRE(h, "vk_hip_batch_download: null batch");
    if (int rc = read_states(h)) return rc;
    std::vector<double> xs[2], ys(static_cast<size_t>(h->m));
    for (int p = 0; p < 2; ++p) xs[p].resize(static_cast<size_t>(h->n));
    double *stage = nullptr;
    VK_HIP_CHECK(device_malloc(&stage, sizeof(double) * std::max<int64_t>(std::max(h->m, h->n), 1)));
    auto fetch = [&](const void *src, int64_t count, double *dst) -> int {
        if (count == 0) return VK_OK;
        if (h->mixed) {
            k_to_f64_batch<<<bgrid(count), 256, 0, h->stream>>>(count, static_cast<const float *>(src), stage);
            VK_HIP_CHECK(hipMemcpyAsync(dst, stage, sizeof(double) * count, hipMemcpyDeviceToHost, h->stream));
        } else {
            VK_HIP_CHECK(hipMemcpyAsync(dst, src, sizeof(double) * count, hipMemcpyDeviceToHost, h->stream));
        }
        VK_HIP_CHECK(hipStreamSynchronize(h->stream));
        return VK_OK;
    };
    int rc = VK_OK;
    for (int p = 0; p < 2 && !rc; ++p)
        rc = h->mixed ? fetch(h->a32.b.x[p], h->n, xs[p].data()) : fetch(h->a64.b.x[p], h->n, xs[p].data());
    if (!rc) rc = h->mixed ? fetch(h->a32.b.y, h->m, ys.data()) : fetch(h->a64.b.y, h->m, ys.data());
    (void)hipFree(stage);
    if (rc) return rc;
    for (int i = 0; i < h->count; ++i) {
        const auto &R = h->inst_host[i];
        const auto &src = xs[h->states_host[i].k & 1];
        if (x) std::memcpy(x + R.c0, src.data() + R.c0, sizeof(double) * (R.c1 - R.c0));
    }
    if (y && h->m) std::memcpy(y, ys.data(), sizeof(double) * h->m);
    return VK_OK;
}

int vk_hip_batch_last_run_ms(const vk_hip_batch *h, double *ms) {
    VK_REQUIRE(h && ms, "vk_hip_batch_last_run_ms: null argument");
    *ms = h->last_ms;
    return VK_OK;
}

}  // extern "C"
