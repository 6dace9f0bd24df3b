// Fused PDHG iteration for gfx950.
//
// The reference drives one PDHG iteration (src/solver/lp/pdhg.cpp:87-151)
// through ~30 backend calls: 4 SpMVs, ~25 vector ops and 2 host-synchronising
// infinity norms.  Here one iteration is two kernels and never synchronises
// with the host:
//
//   dual   (rows of A):   ax+ = A x+;  Axbar = 2 ax+ - ax  (linearity: A(2x+ - x);
//                         = ax+ after a restart, pdhg.cpp:98-104);
//                         v = y + sigma Axbar;  y+ = v - sigma P_[rl,ru](v/sigma)
//                         (pdhg.cpp:105-113); primal residual |P(ax+) - ax+|
//                         (pdhg.cpp:116-120) and the row terms of the KKT sums.
//   primal (rows of A'):  aty = A' y+ (stored transpose, no atomics);
//                         dual residual |x - P_X(x - (c + aty))| (pdhg.cpp:122-130)
//                         and the column KKT terms; then, fused, the next
//                         primal step x++ = P_X(x+ - tau c - tau aty)
//                         (pdhg.cpp:92-96) whenever tau cannot change at this
//                         iteration.  The last block to finish reduces all
//                         block partials in a fixed order and runs the
//                         step/restart/termination rule on the device.
//
// So A x+ and A' y+ are the only SpMVs: pdhg.cpp:105 is replaced by
// linearity and pdhg.cpp:122 is the next iteration's :92.
//
// Two policies share the kernels:
//   reference  pdhg.cpp:75-150: tau = sigma, every 50 iterations halve both
//              and restart the extrapolation when the metric grew by 25%, or
//              grow tau by 5% (capped) when it fell by 90%; stop when
//              max(r_p, r_d) <= max(primal_tol, dual_tol);
//   adaptive   restarted PDHG with iterate averaging and a primal weight w
//              (tau = eta/w, sigma = eta*w; Applegate et al., "Practical
//              large-scale linear programming using primal-dual hybrid
//              gradient", NeurIPS 2021): every `restart_period` iterations the
//              average since the last restart is evaluated (A xbar and A' ybar
//              by two more SpMVs), the better of current/average by the
//              weighted KKT error becomes the restart candidate, restarts fire
//              on the 0.2 / 0.8 / 0.36 criteria, and w is smoothed towards
//              ||dy||/||dx|| between restart points.
// Either policy can stop on the relative KKT error instead (original
// coordinates, evaluated every iteration from the fused partial sums).

#include "runtime.hpp"
#include "pdhg_core.hpp"
#include "panel.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <string>
#include <new>
#include <vector>

namespace vk {
namespace {

// ------------------------------------------------------------------ kernels

// Plain SpMV y = A x rounded to T (setup: A x0 and A' y0; the split passes
// store FP64 sums with T = double).
template <class T, bool Stream = false>
struct StoreEpi {
    T *y;
    struct Pre {};
    __device__ Pre prefetch(int) const { return Pre{}; }
    __device__ void row(int r, double acc, const Pre &, Partials &) const {
        vec_store<Stream>(y + r, static_cast<T>(acc));
    }
};

// Epilogue modes of the iteration kernels: the reference / adaptive
// epilogues, the Halpern ones, and the Halpern ones of the recentred FP32
// solver (which read the dual base and the objective beside the shifted
// gradient).
// kEpiHalpernQP / kEpiRecentredQP: the Halpern ones with a quadratic
// objective (FP64 / the recentred FP32 solver).
constexpr int kEpiPlain = 0, kEpiHalpern = 1, kEpiRecentred = 2, kEpiHalpernQP = 3, kEpiRecentredQP = 4;

// The dual and primal epilogues of one iteration in mode M.
template <class T, int M, bool Stream>
struct IterEpis {
    static constexpr bool R = M == kEpiRecentred || M == kEpiRecentredQP;
    static constexpr bool QP = M == kEpiHalpernQP || M == kEpiRecentredQP;
    static __device__ HalpernDualEpi<T, Stream, R> dual(const Bufs<T> &b, const DevState &s) {
        return HalpernDualEpi<T, Stream, R>{b, s.sigma, s.inv_sigma, halpern_weights(s.ks), check_iteration(s)};
    }
    static __device__ HalpernPrimalEpi<T, Stream, R, QP> primal(const Bufs<T> &b, const DevState &s, bool) {
        return HalpernPrimalEpi<T, Stream, R, QP>{b, b.x[(s.k + 1) & 1], b.x[s.k & 1], s.tau, may_fuse(s),
                                                  check_iteration(s), halpern_weights(s.ks)};
    }
};

template <class T, bool Stream>
struct IterEpis<T, kEpiPlain, Stream> {
    static __device__ DualEpi<T, Stream> dual(const Bufs<T> &b, const DevState &s) {
        return DualEpi<T, Stream>{b, s.sigma, s.inv_sigma, s.policy == VK_PDHG_POLICY_REFERENCE && s.restart_flag,
                                  s.policy == VK_PDHG_POLICY_ADAPTIVE};
    }
    static __device__ PrimalEpi<T, Stream> primal(const Bufs<T> &b, const DevState &s, bool store_aty) {
        return PrimalEpi<T, Stream>{b, b.x[(s.k + 1) & 1], b.x[s.k & 1], s.tau, may_fuse(s), store_aty};
    }
};

template <class T, int M>
__global__ __launch_bounds__(kBlock) void k_dual(Bufs<T> b, CsrView<T> A) {
    const DevState s = *b.st;
    if (s.status != VK_PDHG_RUNNING) return;
    __shared__ double lds[kWaves][kTileNnz];
    __shared__ double scratch[kSlotScratch];
    const T *xin = b.x[(s.k + 1) & 1];
    auto epi = IterEpis<T, M, false>::dual(b, s);
    Partials part;
    part.zero();
    tile_spmv(A, xin, epi, part, lds[threadIdx.x / kWave]);
    store_partials(part, b.part_rows, scratch);
}

template <class T, int M>
__global__ __launch_bounds__(kBlock) void k_primal(Bufs<T> b, CsrView<T> AT) {
    const DevState s = *b.st;
    if (s.status != VK_PDHG_RUNNING) return;
    __shared__ double lds[kWaves][kTileNnz];
    __shared__ double scratch[kSlotScratch];
    auto epi = IterEpis<T, M, false>::primal(b, s, true);
    Partials part;
    part.zero();
    tile_spmv(AT, b.y, epi, part, lds[threadIdx.x / kWave]);
    store_partials(part, b.part_cols, scratch);
}

// Split form of the fused kernels for large operands: the SpMV alone (FP64
// sums to memory, a light epilogue that keeps 8 waves per SIMD), then a
// coalesced pass over rows or columns that runs the same epilogue on the
// stored sums (the dual side: k_dual_spmv or k_blocks, then k_dual_rows; the
// A' side: k_primal_own, both phases in one launch).  The iterates are
// computed exactly as by the fused kernels; only the block partials of the
// KKT sums are grouped differently (rows per workgroup), which the
// reference policy's max-norm tests do not see.
template <class T>
__global__ __launch_bounds__(kBlock) void k_dual_spmv(Bufs<T> b, CsrView<T> A, double *part) {
    const DevState s = *b.st;
    if (s.status != VK_PDHG_RUNNING) return;
    __shared__ double lds[kWaves][kTileNnz];
    StoreEpi<double, true> epi{part};
    Partials unused;
    unused.zero();
    tile_spmv(A, b.x[(s.k + 1) & 1], epi, unused, lds[threadIdx.x / kWave]);
}

// Split A' y+ and column pass in one launch: workgroup w owns the
// contiguous tiles [w nt / G, (w+1) nt / G) of A' and so the rows (columns of
// A) those tiles cover.  Its waves store the FP64 sums of those rows, then,
// after a workgroup barrier, the workgroup runs the column pass over exactly
// those rows: no second launch, and the pass's streaming overlaps other
// workgroups' gathers.  The sums and iterates are the split kernels' own.
template <class T, int M>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8)))
void k_primal_own(Bufs<T> b, CsrView<T> AT) {
    const DevState s = *b.st;
    if (s.status != VK_PDHG_RUNNING) return;
    __shared__ double lds[kWaves][kTileNnz];
    __shared__ double scratch[kSlotScratch];
    const long long G = gridDim.x, w = blockIdx.x;
    const int t0 = static_cast<int>(w * AT.ntiles / G), t1 = static_cast<int>((w + 1) * AT.ntiles / G);
    {
        StoreEpi<double, true> epi{b.atyd};   // streamed: the MALL budget is spent on A
        Partials unused;
        unused.zero();
        tile_spmv_range(AT, b.y, epi, unused, lds[threadIdx.x / kWave], t0, t1,
                        static_cast<int>(threadIdx.x) / kWave, kWaves);
    }
    __syncthreads();   // the sums of this workgroup's rows are visible to all its threads
    const int r0 = AT.tiles[t0].x, r1 = AT.tiles[t1].x;
    auto epi = IterEpis<T, M, true>::primal(b, s, sizeof(T) != sizeof(double));
    Partials part;
    part.zero();
    for (int j = r0 + static_cast<int>(threadIdx.x); j < r1; j += kBlock) {
        const auto pre = epi.prefetch(j);
        epi.row(j, __builtin_nontemporal_load(b.atyd + j), pre, part);
    }
    store_partials(part, b.part_cols, scratch);
}


// sum_k |q_jk| of every row j of Q (the Halpern policy's column steps).
__global__ __launch_bounds__(kBlock) void k_row_abs_sums(CsrView<double> Q, double *out) {
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < Q.rows; j += gridDim.x * kBlock) {
        double sum = 0.0;
        for (int k = Q.row_ptr[j]; k < Q.row_ptr[j + 1]; ++k) sum += fabs(Q.val[k]);
        out[j] = sum;
    }
}

// Q x+ for a quadratic objective (pdqp.cpp:116, :149): FP64 sums into b.qx,
// before the primal kernel that uses them.
template <class T>
__global__ __launch_bounds__(kBlock) void k_qx(Bufs<T> b, CsrView<T> Q) {
    const DevState s = *b.st;
    if (s.status != VK_PDHG_RUNNING) return;
    __shared__ double lds[kWaves][kTileNnz];
    StoreEpi<double> epi{b.qx};
    Partials unused;
    unused.zero();
    tile_spmv(Q, b.x[(s.k + 1) & 1], epi, unused, lds[threadIdx.x / kWave]);
}

// Column-blocked A x into per-block partial row sums (part[p*m + row]).
// `which` selects x: 0 = the primal iterate x[(k+1)&1] (dual step),
// 1 = the running sum xsum (average evaluation).  The grid is a multiple of
// the block count.
template <class T>
struct PartialEpi {
    double *out;
    struct Pre {};
    __device__ Pre prefetch(int) const { return Pre{}; }
    __device__ void row(int r, double acc, const Pre &, Partials &) const { __builtin_nontemporal_store(acc, out + r); }
};

template <class V, class X>
__device__ __forceinline__ void blocked_spmv(const BlockedView<V> &B, const X *__restrict__ x,
                                             double *__restrict__ part, double *lds) {
    const int p = static_cast<int>(blockIdx.x) % B.nb;
    const int sub = static_cast<int>(blockIdx.x) / B.nb, nsub = static_cast<int>(gridDim.x) / B.nb;
    const CsrView<V> A = B.block(p);
    PartialEpi<V> epi{part + static_cast<long>(p) * B.rows};
    Partials unused;
    unused.zero();
    tile_spmv_range(A, x, epi, unused, lds, 0, A.ntiles, sub * kWaves + static_cast<int>(threadIdx.x) / kWave,
                    nsub * kWaves);
}

template <class T>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8)))
void k_blocks(Bufs<T> b, BlockedView<T> B, double *part, int which) {
    const DevState s = *b.st;
    if (s.status != VK_PDHG_RUNNING) return;
    __shared__ double lds[kWaves][kTileNnz];
    if (which) blocked_spmv(B, b.xsum, part, lds[threadIdx.x / kWave]);
    else blocked_spmv(B, b.x[(s.k + 1) & 1], part, lds[threadIdx.x / kWave]);
}

// k_blocks for the iteration whose A x+ runs beside the previous iteration's
// finalizer (on another stream): x is passed by the host instead of read
// from the device state, which the finalizer is writing; only the status
// word is read (a converged solve makes this launch a no-op, or at worst one
// unused SpMV).
template <class T>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8)))
void k_blocks_x(const DevState *st, BlockedView<T> B, const T *x, double *part) {
    if (st->status != VK_PDHG_RUNNING) return;
    __shared__ double lds[kWaves][kTileNnz];
    blocked_spmv(B, x, part, lds[threadIdx.x / kWave]);
}

// Sum over the nb column blocks of row i's partial sums, in block order
// (from 0.0, as a plain loop adds them), with the loads issued four at a
// time so that a row waits on one memory latency per four blocks, not one
// per block.
__device__ __forceinline__ double block_partials_sum(const double *__restrict__ part, long m, int i, int nb) {
    double acc = 0.0;
    int p = 0;
    for (; p + 4 <= nb; p += 4) {
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = __builtin_nontemporal_load(part + static_cast<long>(p + u) * m + i);
#pragma unroll
        for (int u = 0; u < 4; ++u) acc += v[u];
    }
    for (; p < nb; ++p) acc += __builtin_nontemporal_load(part + static_cast<long>(p) * m + i);
    return acc;
}

// Items a thread of the panel layout's row / column passes takes at once.
constexpr int kPassItems = 4;

// The epilogue over items [0, count) from their nb partial sums (added in
// partial order), kPassItems items per thread and step: the loads of all of
// them (the epilogue's prefetch and the partials) are issued before the first
// item's arithmetic, and the workgroup's 8-slot reduction is shared by
// kPassItems * kBlock items.
template <class Epi>
__device__ __forceinline__ void pass_items(Epi &epi, const double *__restrict__ part, int count, int nb,
                                           Partials &acc) {
    constexpr int U = kPassItems;
    for (int i0 = blockIdx.x * (kBlock * U) + threadIdx.x; i0 < count; i0 += gridDim.x * (kBlock * U)) {
        typename Epi::Pre pre[U];
        double sum[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * kBlock;
            if (i < count) {
                pre[u] = epi.prefetch(i);
                sum[u] = block_partials_sum(part, count, i, nb);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * kBlock;
            if (i < count) epi.row(i, sum[u], pre[u], acc);
        }
    }
}

// Dual step over rows from the column-blocked (or panel group) partial sums
// (k_dual's epilogue; the nb partials are added in block order).
template <class T, int M>
__global__ __launch_bounds__(kBlock) void k_dual_rows(Bufs<T> b, const double *__restrict__ part, int nb) {
    const DevState s = *b.st;
    if (s.status != VK_PDHG_RUNNING) return;
    __shared__ double scratch[kSlotScratch];
    auto epi = IterEpis<T, M, true>::dual(b, s);
    Partials acc_part;
    acc_part.zero();
    pass_items(epi, part, b.m, nb, acc_part);
    store_partials(acc_part, b.part_rows, scratch);
}

// Primal step over columns from the panel layout's partial sums of A' y+
// (k_primal_own's column pass; the nb group partials are added in group
// order).
template <class T, int M>
__global__ __launch_bounds__(kBlock) void k_primal_cols(Bufs<T> b, const double *__restrict__ part, int nb) {
    const DevState s = *b.st;
    if (s.status != VK_PDHG_RUNNING) return;
    __shared__ double scratch[kSlotScratch];
    auto epi = IterEpis<T, M, true>::primal(b, s, true);
    Partials acc_part;
    acc_part.zero();
    pass_items(epi, part, b.n, nb, acc_part);
    store_partials(acc_part, b.part_cols, scratch);
}

// Average evaluation over rows from the column-blocked partial sums.
template <class T>
__global__ __launch_bounds__(kBlock) void k_eval_rows_blocked(Bufs<T> b, const double *__restrict__ part, int nb) {
    const DevState s = *b.st;
    if (s.status != VK_PDHG_RUNNING) return;
    __shared__ double scratch[kSlotScratch];
    EvalRowsEpi<T> epi{b, 1.0 / static_cast<double>(s.ns)};
    Partials acc_part;
    acc_part.zero();
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < b.m; i += gridDim.x * kBlock) {
        const auto pre = epi.prefetch(i);
        epi.row(i, block_partials_sum(part, b.m, i, nb), pre, acc_part);
    }
    store_partials(acc_part, b.part_erows, scratch);
}

// One block: reduces the partials of this iteration's dual and primal
// kernels in a fixed order and runs the step / restart / termination rule.
// (A separate launch orders it after every block of both kernels without a
// device-scope fence in each block.)
//
// The state is read into LDS by one word per thread while the partials load,
// and the rule runs on a register copy: one thread working on the state in
// global memory would wait on a chain of dependent loads.
template <class T>
__global__ __launch_bounds__(kFinalBlock) void k_finalize(Bufs<T> b) {
    __shared__ double scratch[kFinalScratch];
    __shared__ double red_rows[kSlots], red_cols[kSlots];
    __shared__ DevState st_lds;
    static_assert(sizeof(DevState) % sizeof(int) == 0 && sizeof(DevState) / sizeof(int) <= kFinalBlock);
    if (threadIdx.x < sizeof(DevState) / sizeof(int))
        reinterpret_cast<int *>(&st_lds)[threadIdx.x] = reinterpret_cast<const int *>(b.st)[threadIdx.x];
    reduce_partials_pair(b.part_rows, b.grid_rows, b.part_cols, b.grid_cols, red_rows, red_cols, scratch);
    if (threadIdx.x != 0) return;
    DevState S = st_lds;
    if (S.status != VK_PDHG_RUNNING) return;
    finalize_step(S, red_rows, red_cols);
    *b.st = S;
}

// Standalone primal step from x[k&1] into x[(k+1)&1] (setup, and after an
// iteration whose step size may have changed).
template <class T>
__global__ __launch_bounds__(kBlock) void k_pstep(Bufs<T> b) {
    const DevState s = *b.st;
    if (s.status != VK_PDHG_RUNNING) return;
    const T *xcur = b.x[s.k & 1];
    T *xnext = b.x[(s.k + 1) & 1];
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < b.n; j += gridDim.x * kBlock) {
        if (b.qabs) pstep_element<T, true>(b, s, j, xcur, xnext);
        else pstep_element<T, false>(b, s, j, xcur, xnext);
    }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_eval_rows(Bufs<T> b, CsrView<T> A) {
    const DevState s = *b.st;
    if (s.status != VK_PDHG_RUNNING) return;
    __shared__ double lds[kWaves][kTileNnz];
    __shared__ double scratch[kSlotScratch];
    EvalRowsEpi<T> epi{b, 1.0 / static_cast<double>(s.ns)};
    Partials part;
    part.zero();
    tile_spmv(A, b.xsum, epi, part, lds[threadIdx.x / kWave]);
    store_partials(part, b.part_erows, scratch);
}

// Average evaluation over columns.
template <class T>
__global__ __launch_bounds__(kBlock) void k_eval_cols(Bufs<T> b, CsrView<T> AT) {
    const DevState s = *b.st;
    if (s.status != VK_PDHG_RUNNING) return;
    __shared__ double lds[kWaves][kTileNnz];
    __shared__ double scratch[kSlotScratch];
    EvalColsEpi<T> epi{b, 1.0 / static_cast<double>(s.ns)};
    Partials part;
    part.zero();
    tile_spmv(AT, b.ysum, epi, part, lds[threadIdx.x / kWave]);
    store_partials(part, b.part_ecols, scratch);
}

// One block: the adaptive restart rule on the evaluated average.  `init`
// evaluates the start point instead and only records its KKT error as the
// first restart reference.
template <class T>
__global__ __launch_bounds__(kFinalBlock) void k_check(Bufs<T> b, int init) {
    if (b.st->status != VK_PDHG_RUNNING) return;
    __shared__ double scratch[kFinalScratch];
    __shared__ double red_rows[kSlots], red_cols[kSlots];
    reduce_partials(b.part_erows, b.grid_rows, red_rows, scratch);
    reduce_partials(b.part_ecols, b.grid_cols, red_cols, scratch);
    if (threadIdx.x != 0) return;
    check_step(*b.st, red_rows, red_cols, init != 0);
}

// After a restart check: move to the average if chosen (or converged on it),
// record the restart point, clear the sums, and take the next primal step.
template <class T>
__global__ __launch_bounds__(kBlock) void k_apply(Bufs<T> b) {
    const DevState s = *b.st;
    if (s.status != VK_PDHG_RUNNING && !s.final_from_avg) return;
    T *xcur = b.x[s.k & 1];
    T *xnext = b.x[(s.k + 1) & 1];
    const int total = b.n > b.m ? b.n : b.m;
    for (int t = blockIdx.x * kBlock + threadIdx.x; t < total; t += gridDim.x * kBlock) {
        if (t < b.n) apply_column(b, s, t, xcur, xnext);
        if (t < b.m) apply_row(b, s, t);
    }
}

// Primal infeasibility test at a restart check of the adaptive and Halpern
// policies (farkas_column / infeasibility_rule): the column part of D0 of the
// last dual iterate y+ from g = A' y+, then the decision in one block.
template <class T>
__global__ __launch_bounds__(kBlock) void k_infeas_cols(Bufs<T> b, const T *__restrict__ g) {
    if (b.st->status != VK_PDHG_RUNNING) return;
    __shared__ double scratch[kSlotScratch];
    Partials part;
    part.zero();
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < b.n; j += gridDim.x * kBlock) {
        double gj = static_cast<double>(__builtin_nontemporal_load(g + j));
        // Recentred: g is A' of the dual correction; c - cobj adds A' ybase.
        if (b.cobj) gj += static_cast<double>(b.c[j]) - static_cast<double>(b.cobj[j]);
        farkas_column(gj, static_cast<double>(b.l[j]), static_cast<double>(b.u[j]), part);
    }
    store_partials(part, b.part_ecols, scratch);
}

template <class T>
__global__ __launch_bounds__(kFinalBlock) void k_infeas_check(Bufs<T> b) {
    __shared__ double scratch[kFinalScratch];
    __shared__ double red[kSlots];
    reduce_partials(b.part_ecols, b.grid_cols, red, scratch);
    if (threadIdx.x != 0) return;
    DevState S = *b.st;
    infeasibility_rule(S, red);
    if (S.status == VK_PDHG_PRIMAL_INFEASIBLE) b.st->status = S.status;
}

// Halpern policy, after the finalizer of a restart-check iteration (the
// primal kernel did not fuse the next step): on a restart z = z0 = T(z_k),
// i.e. x+ and y+ with the A x+ and A' y+ kept at the check, which also
// become the primal weight's restart point; then the next PDHG primal step
// x+ = P_X((xh - tau c) - tau A' yh) with the step of the current weight.
template <class T>
__global__ __launch_bounds__(kBlock) void k_halpern_apply(Bufs<T> b) {
    const DevState s = *b.st;
    if (s.status != VK_PDHG_RUNNING) return;
    const T *xp = b.x[s.k & 1];
    T *xnext = b.x[(s.k + 1) & 1];
    const int total = b.n > b.m ? b.n : b.m;
    for (int t = blockIdx.x * kBlock + threadIdx.x; t < total; t += gridDim.x * kBlock) {
        if (t < b.n) {
            if (b.qxh) halpern_apply_column<T, true>(b, s, t, xp, xnext);
            else halpern_apply_column<T, false>(b, s, t, xp, xnext);
        }
        if (t < b.m) halpern_apply_row(b, s, t);
    }
}

// ------------------------------------------------------------------ recentring
//
// FP32 storage bounds how close a plain FP32 iteration gets to the solution:
// every output is rounded relative to the iterate's own size (about 6e-8 of
// it), and the Halpern iteration averages those errors down only as far as
// its anchor weight allows; on the 400k x 1M LP the relative KKT error then
// stalls between 6e-6 and 4e-5.  The recentred solver stores instead the
// correction to an FP64 base point: x = xbase + u, y = ybase + v.  PDHG,
// the Halpern step and the restart rule are affine-invariant, so u and v
// follow the same iteration on the LP shifted to the base (objective c, the
// gradient offset c + A'ybase, bounds l - xbase, u - xbase, rl - A xbase,
// ru - A xbase), with the dual step taken on the total ybase + v.  Rounding
// is then relative to the correction, which shrinks with the error.  At
// every restart the restart point T(z_k) becomes the new base: the
// corrections fold into (xbase, ybase) in FP64, A xbase and A'ybase are
// recomputed with the FP64 matrix, the shifted LP is rounded to FP32 once,
// and every correction restarts from zero.
// Whether the recentring kernels after a restart check act: at a restart
// that the finalizer applied (or always, when forced at setup).
__device__ __forceinline__ bool recentre_now(const DevState *st, int force) {
    return force || (st->status == VK_PDHG_RUNNING && st->restart_apply);
}

// Folds the restart point T(z_k) (the corrections x+ = x[k & 1] and y) into
// the base and clears every correction and its products.  Runs after
// k_halpern_apply, which copied T(z_k) into the Halpern state.
template <class T>
__global__ __launch_bounds__(kBlock) void k_recentre_base(Bufs<T> b, RecentreBufs<T> r, int force) {
    if (!recentre_now(b.st, force)) return;
    const long long k = b.st->k;
    T *xp = b.x[k & 1];
    const int total = b.n > b.m ? b.n : b.m;
    for (int t = blockIdx.x * kBlock + threadIdx.x; t < total; t += gridDim.x * kBlock) {
        if (t < b.n) {
            r.xb[t] = clamp_box(r.xb[t] + static_cast<double>(xp[t]), r.l[t], r.u[t]);
            xp[t] = T(0);
            b.xh[t] = T(0);
            b.x0h[t] = T(0);
            b.xlast[t] = T(0);
            b.aty[t] = T(0);
            b.aty0h[t] = T(0);
            if (b.qxh) {   // Q x of the cleared Halpern iterate and anchor
                b.qxh[t] = 0.0;
                b.qx0h[t] = 0.0;
            }
        }
        if (t < b.m) {
            r.yb[t] = dual_sign_project(r.yb[t] + static_cast<double>(b.y[t]), r.rl[t], r.ru[t]);
            b.y[t] = T(0);
            b.yh[t] = T(0);
            b.y0h[t] = T(0);
            b.ylast[t] = T(0);
            b.ax[t] = T(0);
            b.ax0h[t] = T(0);
        }
    }
}

// y = A x on the FP64 tiles of A or A' (the base's products), when
// recentre_now.
__global__ __launch_bounds__(kBlock) void k_spmv_recentre(const DevState *st, int force, CsrView<double> A,
                                                         const double *x, double *y) {
    if (!recentre_now(st, force)) return;
    __shared__ double lds[kWaves][kTileNnz];
    StoreEpi<double> epi{y};
    Partials part;
    part.zero();
    tile_spmv(A, x, epi, part, lds[threadIdx.x / kWave]);
}

// The LP shifted to the new base, rounded to T once, and the next primal
// step from the cleared Halpern state (k_halpern_apply's step with xh = 0
// and A'yh = 0); partial slot 5 collects c'xbase.  With a quadratic
// objective the gradient offset is c + Q xbase + A'ybase, the objective's
// gradient at the base c + Q xbase goes to cobj, slot 5 collects c'xbase +
// 1/2 xbase'Q xbase, and the step is the column step.  Launched over
// grid_cols workgroups (the size of part_ecols).
template <class T>
__global__ __launch_bounds__(kBlock) void k_recentre_data(Bufs<T> b, RecentreBufs<T> r, int force) {
    if (!recentre_now(b.st, force)) return;
    __shared__ double scratch[kSlotScratch];
    const long long k = b.st->k;
    const double tau = b.st->tau;
    T *xnext = b.x[(k + 1) & 1];
    Partials part;
    part.zero();
    const int total = b.n > b.m ? b.n : b.m;
    for (int t = blockIdx.x * kBlock + threadIdx.x; t < total; t += gridDim.x * kBlock) {
        if (t < b.n) {
            const double xb = r.xb[t];
            const T lo = static_cast<T>(r.l[t] - xb), hi = static_cast<T>(r.u[t] - xb);
            r.ls[t] = lo;
            r.us[t] = hi;
            if (r.qxb) {
                const double cq = r.c[t] + r.qxb[t];
                const T g = static_cast<T>(cq + r.aty[t]);
                r.cs[t] = g;
                r.cobj[t] = static_cast<T>(cq);
                xnext[t] = static_cast<T>(clamp_box(qp_step(0.0, static_cast<double>(g), 0.0, 0.0,
                                                            qp_column_step(tau, b.qabs[t])),
                                                    static_cast<double>(lo), static_cast<double>(hi)));
                part.v[5] += r.c[t] * xb + 0.5 * xb * r.qxb[t];
            } else {
                const T g = static_cast<T>(r.c[t] + r.aty[t]);
                r.cs[t] = g;
                xnext[t] = static_cast<T>(clamp_box((0.0 - tau * static_cast<double>(g)) - tau * 0.0,
                                                    static_cast<double>(lo), static_cast<double>(hi)));
                part.v[5] += r.c[t] * xb;
            }
        }
        if (t < b.m) {
            const double ax = r.ax[t];
            r.rls[t] = static_cast<T>(r.rl[t] - ax);
            r.rus[t] = static_cast<T>(r.ru[t] - ax);
        }
    }
    store_partials(part, b.part_ecols, scratch);
}

template <class T>
__global__ __launch_bounds__(kFinalBlock) void k_recentre_finish(Bufs<T> b, int force) {
    if (!recentre_now(b.st, force)) return;
    __shared__ double scratch[kFinalScratch];
    __shared__ double red[kSlots];
    reduce_partials(b.part_ecols, b.grid_cols, red, scratch);
    if (threadIdx.x == 0) b.st->cx_base = red[5];
}

// The total point (xbase + x, ybase + y) in FP64, x kept in [l, u] and y on
// the dual sign rule (both only remove rounding of the corrections).
template <class T>
__global__ void k_recentre_total(int n, int m, const T *x, const T *y, RecentreBufs<T> r, double *xo, double *yo) {
    const int total = n > m ? n : m;
    for (int t = blockIdx.x * kBlock + threadIdx.x; t < total; t += gridDim.x * kBlock) {
        if (t < n) xo[t] = clamp_box(r.xb[t] + static_cast<double>(x[t]), r.l[t], r.u[t]);
        if (t < m) yo[t] = dual_sign_project(r.yb[t] + static_cast<double>(y[t]), r.rl[t], r.ru[t]);
    }
}

__global__ void k_clamp64(int n, double *x, const double *l, const double *u) {
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) x[j] = clamp_box(x[j], l[j], u[j]);
}

template <class T>
__global__ void k_project_start(int n, T *x, const T *l, const T *u) {
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock)
        x[j] = static_cast<T>(clamp_box(static_cast<double>(x[j]), static_cast<double>(l[j]),
                                        static_cast<double>(u[j])));
}

template <class T>
__global__ void k_copy(int n, const T *src, T *dst) {
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) dst[j] = src[j];
}

template <class T>
__global__ void k_widen(int n, const T *src, double *dst) {
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) dst[j] = static_cast<double>(src[j]);
}

template <class T>
__global__ void k_zero(int n, T *dst) {
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) dst[j] = T(0);
}


template <class T>
__global__ __launch_bounds__(kBlock) void k_spmv_store(CsrView<T> A, const T *x, T *y) {
    __shared__ double lds[kWaves][kTileNnz];
    __shared__ double scratch[kSlotScratch];
    StoreEpi<T> epi{y};
    Partials part;
    part.zero();
    tile_spmv(A, x, epi, part, lds[threadIdx.x / kWave]);
}

// Column-blocked y = A x in the storage type (the layout tuner's timing
// loop): partial sums, then the ordered row pass.
template <class T>
__global__ __launch_bounds__(kBlock) void k_blocks_plain(BlockedView<T> B, const T *x, double *part) {
    __shared__ double lds[kWaves][kTileNnz];
    blocked_spmv(B, x, part, lds[threadIdx.x / kWave]);
}

template <class T>
__global__ void k_rows_plain(int m, int nb, const double *__restrict__ part, T *y) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < m; i += gridDim.x * kBlock) {
        double acc = 0.0;
        for (int p = 0; p < nb; ++p) acc += part[static_cast<long>(p) * m + i];
        y[i] = static_cast<T>(acc);
    }
}

template <class T>
__global__ void k_gather_perm(int count, const T *__restrict__ src, const int *__restrict__ perm, T *dst) {
    for (int k = blockIdx.x * kBlock + threadIdx.x; k < count; k += gridDim.x * kBlock) dst[k] = src[perm[k]];
}

__global__ void k_d2f(int n, const double *src, float *dst) {
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock)
        dst[j] = static_cast<float>(src[j]);
}

__global__ void k_f2d(int n, const float *src, double *dst) {
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock)
        dst[j] = static_cast<double>(src[j]);
}

int egrid(int n) { return std::max(1, std::min((n + kBlock - 1) / kBlock, 8192)); }

// Column-blocked layout of a device CSR pattern (rp, ci) with m rows and n
// columns: block p takes columns [n*p/nb, n*(p+1)/nb) of every row, rows in
// order, blocks one after another.  The pattern (row_ptr, col) and perm,
// the CSR position each blocked entry's value comes from, are built on the
// device; the tiles are cut on the host from the returned offsets.
struct BlockTiles {
    std::vector<int2> tiles;
    int tile_first[kMaxColBlocks + 1] = {};
    int max_tiles = 0;   // most tiles of any block
};

int plan_blocks(const int *rp, const int *ci, int m, int n, int nnz, int nb, int *row_ptr, int *col, int *perm,
                BlockTiles &T, hipStream_t st) {
    std::vector<int> host_rp;
    if (int rc = plan_blocks_on_device(rp, ci, m, n, nnz, nb, row_ptr, col, perm, host_rp, st)) return rc;
    const size_t stride = static_cast<size_t>(m) + 1;
    T.tiles.clear();
    T.max_tiles = 0;
    for (int p = 0; p < nb; ++p) {
        T.tile_first[p] = static_cast<int>(T.tiles.size());
        append_tiles(&host_rp[p * stride], 0, m, T.tiles);
        T.max_tiles = std::max(T.max_tiles, static_cast<int>(T.tiles.size()) - T.tile_first[p] - 1);
    }
    T.tile_first[nb] = static_cast<int>(T.tiles.size());
    return VK_OK;
}

// Workgroups of a column-blocked launch: a multiple of nb, one tile per wave
// up to the grid cap.
int blocked_grid(int nb, int max_tiles) {
    const int per = std::max(1, std::min((max_tiles + kWaves - 1) / kWaves, kMaxGrid / nb));
    return per * nb;
}

// Column blocks when the caller asks for automatic choice: split once the
// gathered vector (n entries of `bytes` each) outgrows 2 MiB, half of one
// XCD's 4 MB L2, into 4 blocks (each slice shared by two XCDs; measured on
// the 1M-column LP, 4 blocks beat 8 by 3% per iteration: fewer partial sums
// to write and add).
int auto_col_blocks(long long n, int bytes, long long m) {
    if (m == 0) return 1;
    return n * bytes > (2ll << 20) ? 4 : 1;
}

// Split passes when the caller asks for automatic choice: once either
// dimension exceeds 2^18 (launch count matters less than the fused
// epilogues' register pressure at that size).
bool auto_split(long long m, long long n) { return std::max(m, n) > (1ll << 18); }

// MALL budget for matrix data when the caller asks for automatic choice:
// the MALL's 256 MB less the gathered x and y (they are loaded with ordinary
// loads too) and a 4 MiB margin for the remaining ordinary traffic.  On the
// 400,000 x 1,000,000 LP this lands on the measured optimum of both
// precisions (FP64 iteration 334 -> 298 us, FP32 315 -> 266 us); a budget
// that overfills the MALL thrashes and loses most of the gain.
constexpr long long kMallBytes = 256'000'000;
constexpr long long kMallMargin = 4ll << 20;

long long cache_budget_bytes(int mib, long long m, long long n, int value_bytes) {
    if (mib < 0) return 0;
    if (mib > 0) return static_cast<long long>(mib) << 20;
    return std::max(0ll, kMallBytes - kMallMargin - (m + n) * value_bytes);
}

// Tiles [0, resident) of an operand of `bytes` bytes over `entries` tile
// entries that fit in `left` bytes of budget (tiles hold near-equal nonzero
// counts, so a prefix of tiles is that fraction of the bytes); takes the
// bytes used from `left`.
int resident_prefix(long long entries, long long bytes, long long &left) {
    if (bytes <= 0 || left <= 0) return 0;
    if (bytes <= left) {
        left -= bytes;
        return static_cast<int>(entries);
    }
    const long long keep = entries * left / bytes;
    left = 0;
    return static_cast<int>(keep);
}

// Panel layout when the caller asks for automatic choice: once either
// dimension exceeds 2^18 (where the gathers' miss rate bounds the tile
// kernels), if the matrix fits the layout.
bool auto_panels(long long m, long long n) { return m > 0 && std::max(m, n) > (1ll << 18); }

bool valid_panel_groups(int g) { return g == 0 || g == 1 || g == 2 || g == 4 || g == 8; }

}  // namespace
}  // namespace vk

using namespace vk;

// The solver object behind the C ABI.
struct vk_hip_pdhg {
    vk_pdhg_params params{};
    const vk_hip_csr *A = nullptr;
    const vk_hip_csr *Q = nullptr;   // quadratic objective (PDQP), FP64 only
    int m = 0, n = 0;
    hipStream_t stream = nullptr;
    // Second stream (default priority) for the finalizer when it runs beside
    // the next iteration's A x (column-blocked layout), and the fork / join
    // events of that overlap.
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    DevState *st = nullptr;
    DevState *host_state = nullptr;   // pinned
    std::vector<void *> allocations;
    Bufs<double> b64{};
    Bufs<float> b32{};
    bool mixed = false;
    int grid_rows = 1, grid_cols = 1;
    long long host_k = 0;             // iterations enqueued
    hipGraphExec_t graph = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    hipEvent_t ev_snap[2] = {nullptr, nullptr};
    DevState *snap[2] = {nullptr, nullptr};   // pinned
    double last_ms = 0.0;
    // Column-blocked dual SpMV (col_blocks > 1): layout, launch width and the
    // col_blocks x m partial row sums.
    int col_blocks = 1;
    int grid_blk = 0;
    // Split passes: SpMV-only kernels (launch widths grid_spmv_*) followed by
    // row / column passes over grid_rows / grid_cols workgroups.
    bool split = false;
    int grid_spmv_rows = 1;
    BlockedView<double> blk64{};
    BlockedView<float> blk32{};
    double *blk_part = nullptr;
    // MALL residency (CsrView::resident) of the A x and A' y operands.
    int resident_a = 0, resident_at = 0;
    // Panel layout (panels): A x into blk_part (pan_a.groups x m partial
    // sums, then k_dual_rows) and A' y into tpart (pan_at.groups x n, then
    // k_primal_cols).
    bool panels = false;
    vk::PanelLayout pan_a, pan_at;
    double *tpart = nullptr;
    // FP32 solvers: FP64 staging of n + m entries for uploads and the
    // solution download, allocated once with the solver.
    double *stage64 = nullptr;
    // Recentred FP32 Halpern solver (vk_pdhg_params.recentre): the base
    // point, the FP64 LP and the launch widths of the FP64 tile SpMVs that
    // recompute the base's products.
    bool recentre = false;
    RecentreBufs<float> rec{};
    int grid_a64 = 1, grid_at64 = 1;

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

template <class T>
const CsrView<T> view_a(const vk_hip_pdhg *s);
template <class V>
CsrView<V> with_resident(CsrView<V> v, int resident) {
    v.resident = resident;
    return v;
}
template <>
const CsrView<double> view_a<double>(const vk_hip_pdhg *s) { return with_resident(s->A->a.view(), s->resident_a); }
template <>
const CsrView<float> view_a<float>(const vk_hip_pdhg *s) { return with_resident(s->A->a32.view(), s->resident_a); }
template <class T>
const CsrView<T> view_at(const vk_hip_pdhg *s);
template <>
const CsrView<double> view_at<double>(const vk_hip_pdhg *s) { return with_resident(s->A->at.view(), s->resident_at); }
template <>
const CsrView<float> view_at<float>(const vk_hip_pdhg *s) {
    return with_resident(s->A->at32.view(), s->resident_at);
}
// The quadratic objective's tiles in the storage precision.
template <class T>
const CsrView<T> view_q(const vk_hip_pdhg *s);
template <>
const CsrView<double> view_q<double>(const vk_hip_pdhg *s) { return s->Q->a.view(); }
template <>
const CsrView<float> view_q<float>(const vk_hip_pdhg *s) { return s->Q->a32.view(); }

template <class T>
BlockedView<T> &blocked(vk_hip_pdhg *s);
template <>
BlockedView<double> &blocked<double>(vk_hip_pdhg *s) { return s->blk64; }
template <>
BlockedView<float> &blocked<float>(vk_hip_pdhg *s) { return s->blk32; }

template <class T>
Bufs<T> &bufs(vk_hip_pdhg *s);
template <>
Bufs<double> &bufs<double>(vk_hip_pdhg *s) { return s->b64; }
template <>
Bufs<float> &bufs<float>(vk_hip_pdhg *s) { return s->b32; }

// A x+ and the dual step: one fused kernel on the plain tiles, or the
// column-blocked SpMV and its row pass.
// `part` selects the kernels: 1 = the SpMV (or the fused kernel), 2 = the
// row pass of a split layout (nothing when fused), 3 = both.
template <class T, int M>
void launch_dual_mode(vk_hip_pdhg *s, hipStream_t st, int part) {
    Bufs<T> &b = bufs<T>(s);
    if (s->panels) {
        if (part & 1) (void)panel_spmv_launch(s->pan_a, b.st, 1, b.x[0], b.x[1], s->blk_part, st);
        if (part & 2) k_dual_rows<T, M><<<s->grid_rows, kBlock, 0, st>>>(b, s->blk_part, s->pan_a.groups);
    } else if (s->col_blocks > 1) {
        if (part & 1) k_blocks<T><<<s->grid_blk, kBlock, 0, st>>>(b, blocked<T>(s), s->blk_part, 0);
        if (part & 2) k_dual_rows<T, M><<<s->grid_rows, kBlock, 0, st>>>(b, s->blk_part, s->col_blocks);
    } else if (s->split) {
        if (part & 1) k_dual_spmv<T><<<s->grid_spmv_rows, kBlock, 0, st>>>(b, view_a<T>(s), s->blk_part);
        if (part & 2) k_dual_rows<T, M><<<s->grid_rows, kBlock, 0, st>>>(b, s->blk_part, 1);
    } else if (part & 1) {
        k_dual<T, M><<<s->grid_rows, kBlock, 0, st>>>(b, view_a<T>(s));
    }
}

// A' y+ with the residual terms and the fused next primal step (`part` as
// for launch_dual_mode).
template <class T, int M>
void launch_primal_mode(vk_hip_pdhg *s, hipStream_t st, int part) {
    Bufs<T> &b = bufs<T>(s);
    if (s->Q && (part & 1)) {
        // Q in the storage precision (FP32 values in the recentred solver), FP64 sums.
        const CsrView<T> Q = view_q<T>(s);
        k_qx<T><<<std::max(1, std::min((Q.ntiles + kWaves - 1) / kWaves, kMaxGrid)), kBlock, 0, st>>>(b, Q);
    }
    if (s->panels) {
        if (part & 1) (void)panel_spmv_launch(s->pan_at, b.st, 0, b.y, b.y, s->tpart, st);
        if (part & 2) k_primal_cols<T, M><<<s->grid_cols, kBlock, 0, st>>>(b, s->tpart, s->pan_at.groups);
    } else if (s->split) {
        if (part & 1) k_primal_own<T, M><<<s->grid_cols, kBlock, 0, st>>>(b, view_at<T>(s));
    } else if (part & 1) {
        k_primal<T, M><<<s->grid_cols, kBlock, 0, st>>>(b, view_at<T>(s));
    }
}

// The epilogue mode of the solver's iteration kernels: the recentred forms
// only exist for FP32 storage, the plain quadratic one for FP64.
template <class T>
int epi_mode(const vk_hip_pdhg *s) {
    if (s->params.policy != VK_PDHG_POLICY_HALPERN) return kEpiPlain;
    if (sizeof(T) == sizeof(double)) return s->Q ? kEpiHalpernQP : kEpiHalpern;
    if (!s->recentre) return kEpiHalpern;
    return s->Q ? kEpiRecentredQP : kEpiRecentred;
}

template <class T>
void launch_dual(vk_hip_pdhg *s, hipStream_t st, int part = 3) {
    const int mode = epi_mode<T>(s);
    if (mode == kEpiPlain) {
        launch_dual_mode<T, kEpiPlain>(s, st, part);
    } else if (mode == kEpiHalpern) {
        launch_dual_mode<T, kEpiHalpern>(s, st, part);
    } else if constexpr (sizeof(T) == sizeof(float)) {
        // The dual epilogue has no quadratic term: kEpiRecentredQP runs the
        // kernels of kEpiRecentred.
        launch_dual_mode<T, kEpiRecentred>(s, st, part);
    } else {
        // The same for kEpiHalpernQP and kEpiHalpern.
        launch_dual_mode<T, kEpiHalpern>(s, st, part);
    }
}

template <class T>
void launch_primal(vk_hip_pdhg *s, hipStream_t st, int part = 3) {
    const int mode = epi_mode<T>(s);
    if (mode == kEpiPlain) {
        launch_primal_mode<T, kEpiPlain>(s, st, part);
    } else if (mode == kEpiHalpern) {
        launch_primal_mode<T, kEpiHalpern>(s, st, part);
    } else if constexpr (sizeof(T) == sizeof(float)) {
        if (mode == kEpiRecentredQP) launch_primal_mode<T, kEpiRecentredQP>(s, st, part);
        else launch_primal_mode<T, kEpiRecentred>(s, st, part);
    } else {
        launch_primal_mode<T, kEpiHalpernQP>(s, st, part);
    }
}

// A xbar of the running average and its row terms.
template <class T>
void launch_eval_rows(vk_hip_pdhg *s, hipStream_t st) {
    Bufs<T> &b = bufs<T>(s);
    if (s->col_blocks > 1) {
        k_blocks<T><<<s->grid_blk, kBlock, 0, st>>>(b, blocked<T>(s), s->blk_part, 1);
        k_eval_rows_blocked<T><<<s->grid_rows, kBlock, 0, st>>>(b, s->blk_part, s->col_blocks);
    } else {
        k_eval_rows<T><<<s->grid_rows, kBlock, 0, st>>>(b, view_a<T>(s));
    }
}

// The recentring after a restart check (or, forced, at setup): fold the
// restart point into the base, A xbase and A'ybase on the FP64 tiles, the
// shifted LP in FP32 with the next primal step, and c'xbase.  Unforced, each
// kernel returns at once unless the finalizer applied a restart.
void launch_recentre(vk_hip_pdhg *s, hipStream_t st, int force) {
    Bufs<float> &b = s->b32;
    k_recentre_base<float><<<egrid(std::max(s->n, s->m)), kBlock, 0, st>>>(b, s->rec, force);
    if (s->m > 0) {
        k_spmv_recentre<<<s->grid_a64, kBlock, 0, st>>>(b.st, force, s->A->a.view(), s->rec.xb, s->rec.ax);
        k_spmv_recentre<<<s->grid_at64, kBlock, 0, st>>>(b.st, force, s->A->at.view(), s->rec.yb, s->rec.aty);
    }
    if (s->Q) {
        const CsrView<double> Q = s->Q->a.view();
        k_spmv_recentre<<<std::max(1, std::min((Q.ntiles + kWaves - 1) / kWaves, kMaxGrid)), kBlock, 0, st>>>(
            b.st, force, Q, s->rec.xb, s->rec.qxb);
    }
    k_recentre_data<float><<<s->grid_cols, kBlock, 0, st>>>(b, s->rec, force);
    k_recentre_finish<float><<<1, kFinalBlock, 0, st>>>(b, force);
}

// The kernels that follow the primal kernel of iteration k: the finalizer,
// then the policy's step-size or restart kernels.
template <class T>
void launch_control(vk_hip_pdhg *s, long long k, hipStream_t st) {
    Bufs<T> &b = bufs<T>(s);
    k_finalize<T><<<1, kFinalBlock, 0, st>>>(b);
    if (s->params.policy == VK_PDHG_POLICY_REFERENCE) {
        if (k > 0 && k % kDecisionPeriod == 0) k_pstep<T><<<egrid(s->n), kBlock, 0, st>>>(b);
    } else if (s->params.policy == VK_PDHG_POLICY_HALPERN) {
        if ((k + 1) % s->params.restart_period == 0) {
            // A' y+ of this iteration: kept at checks in aty_avg.
            k_infeas_cols<T><<<s->grid_cols, kBlock, 0, st>>>(b, b.aty_avg);
            k_infeas_check<T><<<1, kFinalBlock, 0, st>>>(b);
            k_halpern_apply<T><<<egrid(std::max(s->n, s->m)), kBlock, 0, st>>>(b);
            if constexpr (sizeof(T) == sizeof(float))
                if (s->recentre) launch_recentre(s, st, 0);
        }
    } else if ((k + 1) % s->params.restart_period == 0) {
        k_infeas_cols<T><<<s->grid_cols, kBlock, 0, st>>>(b, b.aty);   // A' y+ of this iteration
        k_infeas_check<T><<<1, kFinalBlock, 0, st>>>(b);
        launch_eval_rows<T>(s, st);
        k_eval_cols<T><<<s->grid_cols, kBlock, 0, st>>>(b, view_at<T>(s));
        k_check<T><<<1, kFinalBlock, 0, st>>>(b, 0);
        k_apply<T><<<egrid(std::max(s->n, s->m)), kBlock, 0, st>>>(b);
    }
}

// Length of the repeating iteration block and the first iteration it may
// start at (the reference policy's iteration 0 has its own shape).
long long block_length(const vk_hip_pdhg *s) {
    return s->params.policy == VK_PDHG_POLICY_REFERENCE ? kDecisionPeriod : s->params.restart_period;
}

// Whether launch_control runs more than the finalizer after iteration k.
bool control_after(const vk_hip_pdhg *s, long long k) {
    if (s->params.policy == VK_PDHG_POLICY_REFERENCE) return k > 0 && k % kDecisionPeriod == 0;
    return (k + 1) % s->params.restart_period == 0;
}

// Enqueues the kernels of iterations [k_begin, k_begin + count) without
// synchronising; the schedule mirrors the device's own may_fuse rule.
//
// Column-blocked layout: after an iteration with nothing but the finalizer
// to follow, the finalizer (one workgroup, about 13 us) runs on the side
// stream while the next iteration's A x+ (which needs only x+, already
// written by the primal kernel) runs on the main stream; the next dual pass
// waits for both.  The x+ buffer is chosen by iteration parity on the host,
// so the overlap needs an even block length (graph replays then keep the
// parity).  The same kernels compute the same sums either way.  The panel
// layout keeps the finalizer in-stream: its A x workgroups hold every CU's
// LDS, so a finalizer beside them only starts when they end, and the fork
// and join cost more (about 6 us each) than the finalizer itself.
template <class T>
int enqueue_iterations(vk_hip_pdhg *s, long long k_begin, long long count) {
    Bufs<T> &b = bufs<T>(s);
    const bool overlap = s->side && s->col_blocks > 1 && block_length(s) % 2 == 0;
    bool spmv_done = false;   // iteration k's A x+ was launched beside the previous finalizer
    for (long long k = k_begin; k < k_begin + count; ++k) {
        launch_dual<T>(s, s->stream, spmv_done ? 2 : 3);
        launch_primal<T>(s, s->stream);
        if (overlap && k + 1 < k_begin + count && !control_after(s, k)) {
            VK_HIP_CHECK(hipEventRecord(s->ev_fork, s->stream));
            VK_HIP_CHECK(hipStreamWaitEvent(s->side, s->ev_fork, 0));
            k_finalize<T><<<1, kFinalBlock, 0, s->side>>>(b);
            VK_HIP_CHECK(hipEventRecord(s->ev_join, s->side));
            if (s->panels)
                (void)panel_spmv_launch(s->pan_a, b.st, 0, b.x[(k + 2) & 1], b.x[(k + 2) & 1], s->blk_part, s->stream);
            else
                k_blocks_x<T><<<s->grid_blk, kBlock, 0, s->stream>>>(b.st, blocked<T>(s), b.x[(k + 2) & 1], s->blk_part);
            VK_HIP_CHECK(hipStreamWaitEvent(s->stream, s->ev_join, 0));
            spmv_done = true;
        } else {
            launch_control<T>(s, k, s->stream);
            spmv_done = false;
        }
    }
    return VK_OK;
}

// Splits the MALL budget (vk_pdhg_params.cache_budget_mb) between the two
// operands of an iteration: the A x operand (column-blocked or plain) first,
// then A'.  Bytes per operand: column indices and values of every nonzero,
// the row offsets (per column block) and the tile table.
template <class T>
void plan_residency(vk_hip_pdhg *s) {
    long long left = cache_budget_bytes(s->params.cache_budget_mb, s->m, s->n, static_cast<int>(sizeof(T)));
    const long long nnz = s->A->nnz;
    const long long per_nz = static_cast<long long>(sizeof(T)) + 4;
    s->resident_a = s->resident_at = 0;
    if (s->panels) {
        // Panel layouts: a prefix of each layout's workgroups (whose entry
        // counts are balanced) keeps its entries; A x's first.
        for (PanelLayout *L : {&s->pan_a, &s->pan_at})
            L->resident = resident_prefix(L->nwg, static_cast<long long>(L->ent_bytes), left);
        return;
    }
    if (s->col_blocks > 1) {
        BlockedView<T> &B = blocked<T>(s);
        const long long entries = B.tile_first[B.nb];
        const long long bytes = nnz * per_nz + static_cast<long long>(B.nb) * (s->m + 1) * 4 + entries * 8;
        B.resident = resident_prefix(entries, bytes, left);
    } else {
        const CsrView<T> A = view_a<T>(s);
        s->resident_a = resident_prefix(A.ntiles, nnz * per_nz + (s->m + 1ll) * 4 + (A.ntiles + 1ll) * 8, left);
    }
    const CsrView<T> AT = view_at<T>(s);
    s->resident_at = resident_prefix(AT.ntiles, nnz * per_nz + (s->n + 1ll) * 4 + (AT.ntiles + 1ll) * 8, left);
}

// Builds the column-blocked copy of A in the storage type T on the device.
template <class T>
int build_blocked(vk_hip_pdhg *s) {
    const int m = s->m, n = s->n, nb = s->col_blocks;
    const CsrView<T> A = view_a<T>(s);
    const int nnz = static_cast<int>(s->A->nnz);
    int *row_ptr = s->alloc<int>((static_cast<size_t>(m) + 1) * nb);
    int *col = s->alloc<int>(nnz);
    T *val = s->alloc<T>(nnz);
    s->blk_part = s->alloc<double>(static_cast<size_t>(nb) * m);
    int *perm = nullptr;
    if (!row_ptr || !col || !val || !s->blk_part ||
        device_malloc(&perm, sizeof(int) * std::max(1, nnz)) != hipSuccess) {
        set_error("vk_hip_pdhg_create: column-blocked layout allocation failed");
        return VK_ENOMEM;
    }
    hipStream_t st = s->stream;
    BlockTiles P;
    int rc = plan_blocks(A.row_ptr, A.col, m, n, nnz, nb, row_ptr, col, perm, P, st);
    int2 *tiles = rc ? nullptr : s->alloc<int2>(P.tiles.size());
    if (!rc && !tiles) {
        set_error("vk_hip_pdhg_create: column-blocked layout allocation failed");
        rc = VK_ENOMEM;
    }
    if (!rc) {
        const hipError_t e1 = hipMemcpyAsync(tiles, P.tiles.data(), sizeof(int2) * P.tiles.size(),
                                             hipMemcpyHostToDevice, st);
        if (nnz) k_gather_perm<T><<<egrid(nnz), kBlock, 0, st>>>(nnz, A.val, perm, val);
        const hipError_t e2 = hipGetLastError();
        const hipError_t e3 = hipStreamSynchronize(st);
        if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
            set_error("vk_hip_pdhg_create: column-blocked layout upload failed");
            rc = VK_EDEVICE;
        }
    }
    (void)hipFree(perm);
    if (rc) return rc;
    BlockedView<T> &B = blocked<T>(s);
    B.row_ptr = row_ptr;
    B.col = col;
    B.val = val;
    B.tiles = tiles;
    B.rows = m;
    B.nb = nb;
    for (int p = 0; p <= nb; ++p) B.tile_first[p] = P.tile_first[p];
    s->grid_blk = blocked_grid(nb, P.max_tiles);
    return VK_OK;
}


// Builds the panel layouts of A (A x) and A' (A' y) in the solver's storage
// precision.  Shape: the caller's group counts / panel size / lane count
// where given, else panel_auto_shape with 512 lanes (measured on the 400k x
// 1M LP with 48 KB panels: 512 is as fast or faster than 256 for both
// operands in both precisions, by up to 22% for FP32 A x).  With panel_tune
// the tuner builds the automatic shape and, with its group count, 48 KB
// panels with 512 lanes and with 256 lanes, times each over the unroll
// depths around its automatic one on the actual matrix, and keeps the
// fastest; none of these choices changes a result bit.  VK_EINVAL: the
// matrix does not fit the layout (nothing is kept).
int build_panels(vk_hip_pdhg *s) {
    const vk_hip_csr *A = s->A;
    const int vb = s->mixed ? 4 : 8;
    const int nnz = static_cast<int>(A->nnz);
    const int threads = s->params.panel_threads > 0 ? s->params.panel_threads : 512;
    struct Shape {
        int groups, kb, threads;
    };
    auto shape = [&](long long rows, long long cols, int groups, int kb, int th) {
        panel_auto_shape(rows, cols, nnz, vb, th, &groups, &kb);
        return Shape{groups, kb, th};
    };
    auto build = [&](const int *rp, const int *ci, const void *val, int rows, int cols, int groups,
                     PanelLayout &out) -> int {
        const Shape automatic = shape(rows, cols, groups, s->params.panel_kb, threads);
        // Candidates: the automatic shape, then (unless the caller fixed the
        // panel size or lane count) 48 KB panels with 512 and 256 lanes, all
        // with the automatic shape's group count (the groups set the order
        // of the partial sums; panel size and lanes change no bit): the
        // first that fits is used, or with panel_tune the fastest.
        std::vector<Shape> cands{automatic};
        if (s->params.panel_threads == 0 && s->params.panel_kb == 0)
            for (int th : {512, 256}) {
                const Shape c{automatic.groups, 48, th};
                if (c.groups != automatic.groups || c.kb != automatic.kb || c.threads != automatic.threads)
                    cands.push_back(c);
            }
        PanelLayout best;
        double best_ms = 1e300;
        for (const Shape &c : cands) {
            PanelLayout L;
            int rc = panel_build(rp, ci, val, vb, rows, cols, nnz, c.groups, c.kb, c.threads, L, s->stream);
            if (rc == VK_EINVAL) continue;   // this shape does not fit; another may
            double ms = 0.0;
            if (rc == VK_OK && s->params.panel_tune) rc = panel_tune_steps(L, 5, &ms, s->stream);
            if (rc != VK_OK) {
                L.release();
                best.release();
                return rc;
            }
            if (!best.groups || ms < best_ms) {
                best.release();
                best = L;
                best_ms = ms;
            } else {
                L.release();
            }
            if (!s->params.panel_tune) break;
        }
        if (!best.groups) return VK_EINVAL;
        out = best;
        return VK_OK;
    };
    const void *va = s->mixed ? static_cast<const void *>(A->a32.val) : static_cast<const void *>(A->a.val);
    const void *vat = s->mixed ? static_cast<const void *>(A->at32.val) : static_cast<const void *>(A->at.val);
    int rc = build(A->a.row_ptr, A->a.col, va, s->m, s->n, s->params.panel_groups, s->pan_a);
    if (!rc) rc = build(A->at.row_ptr, A->at.col, vat, s->n, s->m, s->params.panel_groups_t, s->pan_at);
    if (rc) {
        s->pan_a.release();
        s->pan_at.release();
    }
    return rc;
}

template <class T>
int setup_solver(vk_hip_pdhg *s, const double *c, const double *l, const double *u,
                 const double *rl, const double *ru, const double *row_scale,
                 const double *col_scale, const double *x0, const double *y0) {
    const int m = s->m, n = s->n;
    hipStream_t st = s->stream;
    Bufs<T> &b = bufs<T>(s);
    std::memset(&b, 0, sizeof(b));
    b.m = m;
    b.n = n;
    // Padded by 4 entries: the panel SpMV reads the gathered vectors in whole
    // 16-byte vectors.
    b.x[0] = s->alloc<T>(n + 4);
    b.x[1] = s->alloc<T>(n + 4);
    b.y = s->alloc<T>(m + 4);
    b.ax = s->alloc<T>(m);
    b.aty = s->alloc<T>(n);
    const bool halpern = s->params.policy == VK_PDHG_POLICY_HALPERN;
    // Halpern keeps the iterate's A' y in aty while the primal kernel sums A' y+.
    if constexpr (sizeof(T) == sizeof(double)) b.atyd = halpern ? s->alloc<double>(n) : reinterpret_cast<double *>(b.aty);
    else b.atyd = s->alloc<double>(n);
    T *c_d = s->alloc<T>(n), *l_d = s->alloc<T>(n), *u_d = s->alloc<T>(n);
    T *rl_d = s->alloc<T>(m), *ru_d = s->alloc<T>(m);
    T *runs = s->alloc<T>(m), *cs = s->alloc<T>(n);
    b.c = c_d; b.l = l_d; b.u = u_d; b.rl = rl_d; b.ru = ru_d;
    b.row_unscale = runs;
    b.col_scale = cs;
    const bool adaptive = s->params.policy == VK_PDHG_POLICY_ADAPTIVE;
    if (adaptive) {
        b.xsum = s->alloc<double>(n);
        b.ysum = s->alloc<double>(m);
    }
    if (adaptive || halpern) {
        b.xlast = s->alloc<T>(n);
        b.ylast = s->alloc<T>(m);
        b.ax_avg = s->alloc<T>(m);
        b.aty_avg = s->alloc<T>(n);
    }
    if (halpern) {
        b.xh = s->alloc<T>(n);
        b.yh = s->alloc<T>(m);
        b.x0h = s->alloc<T>(n);
        b.y0h = s->alloc<T>(m);
        b.ax0h = s->alloc<T>(m);
        b.aty0h = s->alloc<T>(n);
    }
    b.part_rows = s->alloc<double>(kSlots * s->grid_rows);
    b.part_cols = s->alloc<double>(kSlots * s->grid_cols);
    b.part_erows = s->alloc<double>(kSlots * s->grid_rows);
    b.part_ecols = s->alloc<double>(kSlots * s->grid_cols);
    RecentreBufs<float> &rec = s->rec;
    if constexpr (sizeof(T) == sizeof(float)) {
        if (s->recentre) {
            rec.xb = s->alloc<double>(n);
            rec.yb = s->alloc<double>(m);
            rec.c = s->alloc<double>(n);
            rec.l = s->alloc<double>(n);
            rec.u = s->alloc<double>(n);
            rec.rl = s->alloc<double>(m);
            rec.ru = s->alloc<double>(m);
            rec.ax = s->alloc<double>(m);
            rec.aty = s->alloc<double>(n);
            rec.cs = c_d;
            rec.ls = l_d;
            rec.us = u_d;
            rec.rls = rl_d;
            rec.rus = ru_d;
            b.cobj = s->alloc<T>(n);
            b.ybase = rec.yb;
        }
    }
    b.st = s->st;
    b.grid_rows = s->grid_rows;
    b.grid_cols = s->grid_cols;
    for (void *p : s->allocations)
        if (!p) {
            set_error("vk_hip_pdhg_create: device allocation failed");
            return VK_ENOMEM;
        }
    // Host vectors -> device (FP64 staging, converted on the device in mixed mode).
    double *stage = nullptr;
    if constexpr (sizeof(T) != sizeof(double)) {
        s->stage64 = s->alloc<double>(static_cast<size_t>(n) + m + 1);
        if (!s->stage64) {
            set_error("vk_hip_pdhg_create: device allocation failed");
            return VK_ENOMEM;
        }
        stage = s->stage64;
    }
    std::vector<double> unscale(static_cast<size_t>(m));
    for (int i = 0; i < m; ++i) unscale[i] = 1.0 / row_scale[i];
    std::vector<double> zeros_n(static_cast<size_t>(n), 0.0), zeros_m(static_cast<size_t>(m), 0.0);
    auto put = [&](T *dst, const double *src, int count) -> int {
        if (count == 0) return VK_OK;
        if constexpr (sizeof(T) == sizeof(double)) {
            VK_HIP_CHECK(hipMemcpyAsync(dst, src, sizeof(double) * count, hipMemcpyHostToDevice, st));
        } else {
            VK_HIP_CHECK(hipMemcpyAsync(stage, src, sizeof(double) * count, hipMemcpyHostToDevice, st));
            k_d2f<<<egrid(count), kBlock, 0, st>>>(count, stage, reinterpret_cast<float *>(dst));
            VK_HIP_CHECK(hipStreamSynchronize(st));
        }
        return VK_OK;
    };
    int rc = VK_OK;
    rc = rc ? rc : put(c_d, c, n);
    rc = rc ? rc : put(l_d, l, n);
    rc = rc ? rc : put(u_d, u, n);
    rc = rc ? rc : put(rl_d, rl, m);
    rc = rc ? rc : put(ru_d, ru, m);
    rc = rc ? rc : put(runs, unscale.data(), m);
    rc = rc ? rc : put(cs, col_scale, n);
    if (s->recentre) {
        // The start is the first base point (x0 projected in FP64); the
        // corrections start at zero.
        rc = rc ? rc : put(b.x[0], zeros_n.data(), n);
        rc = rc ? rc : put(b.y, zeros_m.data(), m);
        if constexpr (sizeof(T) == sizeof(float)) {
            rc = rc ? rc : put(const_cast<T *>(b.cobj), c, n);
            auto put64 = [&](const double *dst, const double *src, int count) -> int {
                if (count)
                    VK_HIP_CHECK(hipMemcpyAsync(const_cast<double *>(dst), src, sizeof(double) * count,
                                                hipMemcpyHostToDevice, st));
                return VK_OK;
            };
            rc = rc ? rc : put64(rec.c, c, n);
            rc = rc ? rc : put64(rec.l, l, n);
            rc = rc ? rc : put64(rec.u, u, n);
            rc = rc ? rc : put64(rec.rl, rl, m);
            rc = rc ? rc : put64(rec.ru, ru, m);
            rc = rc ? rc : put64(rec.xb, x0 ? x0 : zeros_n.data(), n);
            rc = rc ? rc : put64(rec.yb, y0 ? y0 : zeros_m.data(), m);
        }
    } else {
        rc = rc ? rc : put(b.x[0], x0 ? x0 : zeros_n.data(), n);
        rc = rc ? rc : put(b.y, y0 ? y0 : zeros_m.data(), m);
    }
    VK_HIP_CHECK(hipStreamSynchronize(st));
    if (rc) return rc;
    if constexpr (sizeof(T) == sizeof(float)) {
        if (s->recentre && s->Q) {
            // Recentred Halpern PDQP: Q x of the zero corrections (x+, the
            // Halpern iterate, the anchor) is zero; Q's absolute row sums
            // set the column steps; Q xbase is the recentring's.
            b.qx = s->alloc<double>(n);
            b.qxh = s->alloc<double>(n);
            b.qx0h = s->alloc<double>(n);
            b.qx_avg = s->alloc<double>(n);
            double *qabs = s->alloc<double>(n);
            rec.qxb = s->alloc<double>(n);
            rec.cobj = const_cast<T *>(b.cobj);
            if (!b.qx || !b.qxh || !b.qx0h || !b.qx_avg || !qabs || !rec.qxb) {
                set_error("vk_hip_pdqp_create: device allocation failed");
                return VK_ENOMEM;
            }
            for (double *v : {b.qx, b.qxh, b.qx0h, b.qx_avg}) k_zero<double><<<egrid(n), kBlock, 0, st>>>(n, v);
            k_row_abs_sums<<<egrid(n), kBlock, 0, st>>>(s->Q->a.view(), qabs);
            b.qabs = qabs;
        }
        if (s->recentre) {
            // The LP shifted to the start (c, l, u, rl, ru above become the
            // shifted ones) and c'xbase; the start projection, A x0, A' y0
            // and the first primal step below then act on zero corrections.
            k_clamp64<<<egrid(n), kBlock, 0, st>>>(n, rec.xb, rec.l, rec.u);
            launch_recentre(s, st, 1);
            VK_HIP_CHECK(hipGetLastError());
        }
    }

    if (s->panels) {
        s->blk_part = s->alloc<double>(static_cast<size_t>(s->pan_a.groups) * m);
        s->tpart = s->alloc<double>(static_cast<size_t>(s->pan_at.groups) * n);
        if (!s->blk_part || !s->tpart) {
            set_error("vk_hip_pdhg_create: device allocation failed");
            return VK_ENOMEM;
        }
    } else if (s->col_blocks > 1) {
        if (int brc = build_blocked<T>(s)) return brc;
    } else if (s->split) {
        s->blk_part = s->alloc<double>(m);
        if (!s->blk_part) {
            set_error("vk_hip_pdhg_create: device allocation failed");
            return VK_ENOMEM;
        }
    }
    plan_residency<T>(s);

    // pdhg.cpp:71 project the start; A x0, A' y0.
    k_project_start<T><<<egrid(n), kBlock, 0, st>>>(n, b.x[0], b.l, b.u);
    const CsrView<T> A = view_a<T>(s), AT = view_at<T>(s);
    if (m > 0) {
        k_zero<T><<<egrid(m), kBlock, 0, st>>>(m, b.ax);
        if (A.ntiles > 0) k_spmv_store<T><<<s->grid_rows, kBlock, 0, st>>>(A, b.x[0], b.ax);
    }
    k_zero<T><<<egrid(n), kBlock, 0, st>>>(n, b.aty);
    if (m > 0 && AT.ntiles > 0) k_spmv_store<T><<<s->grid_cols, kBlock, 0, st>>>(AT, b.y, b.aty);
    if (adaptive) {
        // KKT error of the start point = first restart reference: evaluate it
        // through the average path with sums = (x0, y0) and ns = 1.
        k_widen<T><<<egrid(n), kBlock, 0, st>>>(n, b.x[0], b.xsum);
        k_copy<T><<<egrid(n), kBlock, 0, st>>>(n, b.x[0], b.xlast);
        if (m > 0) {
            k_widen<T><<<egrid(m), kBlock, 0, st>>>(m, b.y, b.ysum);
            k_copy<T><<<egrid(m), kBlock, 0, st>>>(m, b.y, b.ylast);
        }
        DevState tmp = *s->host_state;
        tmp.ns = 1;
        VK_HIP_CHECK(hipMemcpyAsync(s->st, &tmp, sizeof(DevState), hipMemcpyHostToDevice, st));
        VK_HIP_CHECK(hipStreamSynchronize(st));
        launch_eval_rows<T>(s, st);
        k_eval_cols<T><<<s->grid_cols, kBlock, 0, st>>>(b, AT);
        k_check<T><<<1, kFinalBlock, 0, st>>>(b, 1);
        k_zero<double><<<egrid(n), kBlock, 0, st>>>(n, b.xsum);
        if (m > 0) k_zero<double><<<egrid(m), kBlock, 0, st>>>(m, b.ysum);
        VK_HIP_CHECK(hipMemcpyAsync(s->host_state, s->st, sizeof(DevState), hipMemcpyDeviceToHost, st));
        VK_HIP_CHECK(hipStreamSynchronize(st));
        s->host_state->ns = 0;
        VK_HIP_CHECK(hipMemcpyAsync(s->st, s->host_state, sizeof(DevState), hipMemcpyHostToDevice, st));
    }
    if (halpern) {
        // z0 = (x0, y0) is the Halpern iterate, the anchor and the primal
        // weight's restart point; A x0 and A' y0 are its products.
        k_copy<T><<<egrid(n), kBlock, 0, st>>>(n, b.x[0], b.xh);
        k_copy<T><<<egrid(n), kBlock, 0, st>>>(n, b.x[0], b.x0h);
        k_copy<T><<<egrid(n), kBlock, 0, st>>>(n, b.x[0], b.xlast);
        k_copy<T><<<egrid(n), kBlock, 0, st>>>(n, b.aty, b.aty0h);
        if (m > 0) {
            k_copy<T><<<egrid(m), kBlock, 0, st>>>(m, b.y, b.yh);
            k_copy<T><<<egrid(m), kBlock, 0, st>>>(m, b.y, b.y0h);
            k_copy<T><<<egrid(m), kBlock, 0, st>>>(m, b.y, b.ylast);
            k_copy<T><<<egrid(m), kBlock, 0, st>>>(m, b.ax, b.ax0h);
        }
    }
    if constexpr (sizeof(T) == sizeof(double)) {
        if (s->Q) {
            // pdqp.cpp:116 Q x0 for the first primal step.
            b.qx = s->alloc<double>(n);
            if (!b.qx) {
                set_error("vk_hip_pdqp_create: device allocation failed");
                return VK_ENOMEM;
            }
            const CsrView<double> Q = s->Q->a.view();
            k_zero<double><<<egrid(n), kBlock, 0, st>>>(n, b.qx);
            if (Q.ntiles > 0)
                k_spmv_store<double><<<std::max(1, std::min((Q.ntiles + kWaves - 1) / kWaves, kMaxGrid)), kBlock, 0, st>>>(
                    Q, b.x[0], b.qx);
            if (s->params.policy == VK_PDHG_POLICY_HALPERN) {
                // Q x0 is Q x of the Halpern iterate and of the anchor; the
                // absolute row sums of Q set the column steps.
                b.qxh = s->alloc<double>(n);
                b.qx0h = s->alloc<double>(n);
                b.qx_avg = s->alloc<double>(n);
                double *qabs = s->alloc<double>(n);
                if (!b.qxh || !b.qx0h || !b.qx_avg || !qabs) {
                    set_error("vk_hip_pdqp_create: device allocation failed");
                    return VK_ENOMEM;
                }
                k_copy<double><<<egrid(n), kBlock, 0, st>>>(n, b.qx, b.qxh);
                k_copy<double><<<egrid(n), kBlock, 0, st>>>(n, b.qx, b.qx0h);
                k_copy<double><<<egrid(n), kBlock, 0, st>>>(n, b.qx, b.qx_avg);
                k_row_abs_sums<<<egrid(n), kBlock, 0, st>>>(Q, qabs);
                b.qabs = qabs;
            }
        }
    }
    // First primal step x1 = P(x0 - tau (c + A'y0)) into x[1].
    k_pstep<T><<<egrid(n), kBlock, 0, st>>>(b);
    VK_HIP_CHECK(hipGetLastError());
    VK_HIP_CHECK(hipStreamSynchronize(st));
    return VK_OK;
}

template <class T>
int build_graph(vk_hip_pdhg *s, long long k_begin) {
    if (s->graph) return VK_OK;
    hipGraph_t g = nullptr;
    VK_HIP_CHECK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_iterations<T>(s, k_begin, block_length(s));
    VK_HIP_CHECK(hipStreamEndCapture(s->stream, &g));
    if (rc) {
        (void)hipGraphDestroy(g);
        return rc;
    }
    const hipError_t e = hipGraphInstantiate(&s->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    VK_HIP_CHECK(e);
    return VK_OK;
}

template <class T>
int run_solver(vk_hip_pdhg *s, long long max_iterations, int *status) {
    hipStream_t st = s->stream;
    VK_HIP_CHECK(hipMemcpyAsync(s->host_state, s->st, sizeof(DevState), hipMemcpyDeviceToHost, st));
    VK_HIP_CHECK(hipStreamSynchronize(st));
    if (s->host_state->status != VK_PDHG_RUNNING) {
        *status = s->host_state->status;
        s->last_ms = 0.0;
        return VK_OK;
    }
    s->host_k = s->host_state->k;
    const long long target = s->host_k + max_iterations;
    const long long L = block_length(s);
    VK_HIP_CHECK(hipEventRecord(s->ev_start, st));
    // Iterations up to the next block boundary run eagerly; whole blocks then
    // run as one graph launch each.  After each launch the device state is
    // copied to one of two pinned snapshots; the host reads the snapshot of
    // the previous launch while the current one runs, so it never idles the
    // device and overshoots a converged run by at most one block (whose
    // kernels return at once).
    long long first_boundary;
    if (s->params.policy == VK_PDHG_POLICY_REFERENCE)
        first_boundary = ((s->host_k - 1 + L - 1) / L) * L + 1;   // blocks start at k = 1 mod 50
    else
        first_boundary = ((s->host_k + L - 1) / L) * L;
    if (s->host_k == 0 && s->params.policy == VK_PDHG_POLICY_REFERENCE) first_boundary = 1;
    int pending = -1, cur = 0;
    while (s->host_k < target) {
        long long chunk;
        if (s->host_k < first_boundary) {
            chunk = std::min(target, first_boundary) - s->host_k;
            if (int rc = enqueue_iterations<T>(s, s->host_k, chunk)) return rc;
        } else if (s->host_k + L <= target) {
            chunk = L;
            if (s->params.use_graphs) {
                if (int rc = build_graph<T>(s, s->host_k)) return rc;
                VK_HIP_CHECK(hipGraphLaunch(s->graph, st));
            } else {
                if (int rc = enqueue_iterations<T>(s, s->host_k, L)) return rc;
            }
        } else {
            chunk = target - s->host_k;
            if (int rc = enqueue_iterations<T>(s, s->host_k, chunk)) return rc;
        }
        VK_HIP_CHECK(hipGetLastError());
        s->host_k += chunk;
        VK_HIP_CHECK(hipMemcpyAsync(s->snap[cur], s->st, sizeof(DevState), hipMemcpyDeviceToHost, st));
        VK_HIP_CHECK(hipEventRecord(s->ev_snap[cur], st));
        if (pending >= 0) {
            VK_HIP_CHECK(hipEventSynchronize(s->ev_snap[pending]));
            if (s->snap[pending]->status != VK_PDHG_RUNNING) break;
        }
        pending = cur;
        cur ^= 1;
    }
    VK_HIP_CHECK(hipEventRecord(s->ev_stop, st));
    VK_HIP_CHECK(hipMemcpyAsync(s->host_state, s->st, sizeof(DevState), hipMemcpyDeviceToHost, st));
    VK_HIP_CHECK(hipStreamSynchronize(st));
    float ms = 0.0f;
    VK_HIP_CHECK(hipEventElapsedTime(&ms, s->ev_start, s->ev_stop));
    s->last_ms = ms;
    *status = s->host_state->status;
    return VK_OK;
}

template <class T>
int download_solution(vk_hip_pdhg *s, double *x, double *y) {
    Bufs<T> &b = bufs<T>(s);
    VK_HIP_CHECK(hipMemcpyAsync(s->host_state, s->st, sizeof(DevState), hipMemcpyDeviceToHost, s->stream));
    VK_HIP_CHECK(hipStreamSynchronize(s->stream));
    const T *xs = b.x[s->host_state->k & 1];
    if constexpr (sizeof(T) == sizeof(double)) {
        if (x && s->n) VK_HIP_CHECK(hipMemcpyAsync(x, xs, sizeof(double) * s->n, hipMemcpyDeviceToHost, s->stream));
        if (y && s->m) VK_HIP_CHECK(hipMemcpyAsync(y, b.y, sizeof(double) * s->m, hipMemcpyDeviceToHost, s->stream));
    } else {
        double *stage = s->stage64;   // n + m entries, allocated with the solver
        if (s->recentre) {
            // The total point xbase + x, ybase + y.
            k_recentre_total<T><<<egrid(std::max(s->n, s->m)), kBlock, 0, s->stream>>>(s->n, s->m, xs, b.y, s->rec,
                                                                                      stage, stage + s->n);
        } else {
            if (s->n) k_f2d<<<egrid(s->n), kBlock, 0, s->stream>>>(s->n, xs, stage);
            if (s->m) k_f2d<<<egrid(s->m), kBlock, 0, s->stream>>>(s->m, b.y, stage + s->n);
        }
        if (x && s->n) VK_HIP_CHECK(hipMemcpyAsync(x, stage, sizeof(double) * s->n, hipMemcpyDeviceToHost, s->stream));
        if (y && s->m) VK_HIP_CHECK(hipMemcpyAsync(y, stage + s->n, sizeof(double) * s->m, hipMemcpyDeviceToHost, s->stream));
    }
    VK_HIP_CHECK(hipStreamSynchronize(s->stream));
    return VK_OK;
}

// Runs iterations launched one by one with HIP events on the solver's
// stream between the kernels (same schedule as a run).  out[0..3] = mean ms
// of the dual SpMV (or fused dual kernel), the dual row pass, the primal
// SpMV (or fused primal kernel) and the primal column pass (0 when fused);
// out[4] = mean ms per whole iteration.
template <class T>
int profile_iterations(vk_hip_pdhg *s, int64_t iterations, double *out) {
    hipStream_t st = s->stream;
    constexpr int kMarks = 5;
    std::vector<hipEvent_t> ev(static_cast<size_t>(kMarks * iterations + 1), nullptr);
    auto cleanup = [&] {
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
    };
    for (auto &e : ev)
        if (hipEventCreate(&e) != hipSuccess) {
            cleanup();
            set_error("vk_hip_pdhg_profile: event creation failed");
            return VK_EDEVICE;
        }
    int rc = VK_OK;
    auto check = [&](hipError_t e) {
        if (e != hipSuccess && rc == VK_OK) {
            set_error(std::string("vk_hip_pdhg_profile: ") + hipGetErrorString(e));
            rc = VK_EDEVICE;
        }
    };
    check(hipMemcpyAsync(s->host_state, s->st, sizeof(DevState), hipMemcpyDeviceToHost, st));
    check(hipStreamSynchronize(st));
    long long k = s->host_state->k;
    for (int64_t i = 0; i < iterations && rc == VK_OK; ++i, ++k) {
        hipEvent_t *e = &ev[static_cast<size_t>(kMarks * i)];
        check(hipEventRecord(e[0], st));
        launch_dual<T>(s, st, 1);
        check(hipEventRecord(e[1], st));
        launch_dual<T>(s, st, 2);
        check(hipEventRecord(e[2], st));
        launch_primal<T>(s, st, 1);
        check(hipEventRecord(e[3], st));
        launch_primal<T>(s, st, 2);
        check(hipEventRecord(e[4], st));
        launch_control<T>(s, k, st);
    }
    check(hipEventRecord(ev[static_cast<size_t>(kMarks * iterations)], st));
    check(hipStreamSynchronize(st));
    check(hipGetLastError());
    double sum[kMarks - 1] = {};
    for (int64_t i = 0; i < iterations && rc == VK_OK; ++i)
        for (int j = 0; j < kMarks - 1; ++j) {
            float t = 0.0f;
            check(hipEventElapsedTime(&t, ev[static_cast<size_t>(kMarks * i + j)],
                                      ev[static_cast<size_t>(kMarks * i + j + 1)]));
            sum[j] += t;
        }
    float total = 0.0f;
    check(hipEventElapsedTime(&total, ev[0], ev[static_cast<size_t>(kMarks * iterations)]));
    cleanup();
    if (rc) return rc;
    for (int j = 0; j < kMarks - 1; ++j) out[j] = sum[j] / static_cast<double>(iterations);
    // Intervals with no kernel in them (fused dual kernel; the A' side is one
    // kernel in every layout) hold only event overhead.
    if (!s->split && s->col_blocks == 1) out[1] = 0.0;
    if (!s->panels) out[3] = 0.0;   // k_primal_own: A' y and its column pass are one kernel
    out[4] = total / static_cast<double>(iterations);
    return VK_OK;
}

// Times y = A x on the plain tiles (nb = 1) and on each column-blocked
// layout nb in kTuneBlocks; ms[i] is the mean of `repeats` launches.
constexpr int kTuneBlocks[] = {1, 2, 4, 8, 16};
constexpr int kTuneCount = 5;

template <class T>
int tune_col_blocks(vk_hip_csr *A, const CsrView<T> &view, int repeats, int *best, double *ms_out) {
    const int m = static_cast<int>(A->rows), n = static_cast<int>(A->cols), nnz = static_cast<int>(A->nnz);
    hipStream_t st = default_stream();
    std::vector<void *> owned;
    auto release = [&] {
        for (void *p : owned) (void)hipFree(p);
        owned.clear();
    };
    auto dalloc = [&](size_t bytes) -> void * {
        void *p = nullptr;
        if (device_malloc(&p, std::max<size_t>(bytes, 8)) != hipSuccess) return nullptr;
        owned.push_back(p);
        return p;
    };
    T *x = static_cast<T *>(dalloc(sizeof(T) * n));
    T *y = static_cast<T *>(dalloc(sizeof(T) * m));
    double *part = static_cast<double *>(dalloc(sizeof(double) * kMaxColBlocks * static_cast<size_t>(m)));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!x || !y || !part || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        release();
        if (e0) (void)hipEventDestroy(e0);
        set_error("vk_hip_csr_tune_col_blocks: device allocation failed");
        return VK_ENOMEM;
    }
    int rc = VK_OK;
    auto check = [&](hipError_t e) {
        if (e != hipSuccess && rc == VK_OK) {
            set_error(std::string("vk_hip_csr_tune_col_blocks: ") + hipGetErrorString(e));
            rc = VK_EDEVICE;
        }
    };
    check(hipMemsetAsync(x, 0, sizeof(T) * n, st));
    double best_ms = std::numeric_limits<double>::infinity();
    *best = 1;
    for (int i = 0; i < kTuneCount && rc == VK_OK; ++i) {
        const int nb = kTuneBlocks[i];
        double ms = std::numeric_limits<double>::infinity();
        if (nb == 1 || (nb <= n && m > 0)) {
            std::function<void()> launch;
            std::vector<void *> layout;
            BlockedView<T> B{};
            if (nb == 1) {
                const int grid = std::max(1, std::min((view.ntiles + kWaves - 1) / kWaves, 2048));
                launch = [&, grid] {
                    if (view.ntiles > 0) k_spmv_store<T><<<grid, kBlock, 0, st>>>(view, x, y);
                };
            } else {
                const size_t before = owned.size();
                int *brp = static_cast<int *>(dalloc(sizeof(int) * (static_cast<size_t>(m) + 1) * nb));
                int *bcol = static_cast<int *>(dalloc(sizeof(int) * nnz));
                T *bval = static_cast<T *>(dalloc(sizeof(T) * nnz));
                int *bperm = static_cast<int *>(dalloc(sizeof(int) * nnz));
                BlockTiles P;
                if (!brp || !bcol || !bval || !bperm) {
                    set_error("vk_hip_csr_tune_col_blocks: device allocation failed");
                    rc = VK_ENOMEM;
                    break;
                }
                if ((rc = plan_blocks(view.row_ptr, view.col, m, n, nnz, nb, brp, bcol, bperm, P, st))) break;
                int2 *btiles = static_cast<int2 *>(dalloc(sizeof(int2) * P.tiles.size()));
                if (!btiles) {
                    set_error("vk_hip_csr_tune_col_blocks: device allocation failed");
                    rc = VK_ENOMEM;
                    break;
                }
                layout.assign(owned.begin() + static_cast<long>(before), owned.end());
                check(hipMemcpyAsync(btiles, P.tiles.data(), sizeof(int2) * P.tiles.size(), hipMemcpyHostToDevice, st));
                if (nnz) k_gather_perm<T><<<egrid(nnz), kBlock, 0, st>>>(nnz, view.val, bperm, bval);
                B = BlockedView<T>{brp, bcol, bval, btiles, m, nb, {}};
                for (int p = 0; p <= nb; ++p) B.tile_first[p] = P.tile_first[p];
                const int grid = blocked_grid(nb, P.max_tiles);
                launch = [&, grid, nb] {
                    k_blocks_plain<T><<<grid, kBlock, 0, st>>>(B, x, part);
                    k_rows_plain<T><<<egrid(m), kBlock, 0, st>>>(m, nb, part, y);
                };
            }
            for (int w = 0; w < 2; ++w) launch();
            check(hipEventRecord(e0, st));
            for (int r = 0; r < repeats; ++r) launch();
            check(hipEventRecord(e1, st));
            check(hipEventSynchronize(e1));
            check(hipGetLastError());
            float t = 0.0f;
            check(hipEventElapsedTime(&t, e0, e1));
            ms = static_cast<double>(t) / repeats;
            // Free this candidate's layout before the next one.
            for (void *p : layout) {
                (void)hipFree(p);
                owned.erase(std::find(owned.begin(), owned.end(), p));
            }
        }
        if (ms_out) ms_out[i] = ms;
        if (ms < best_ms) {
            best_ms = ms;
            *best = nb;
        }
    }
    release();
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

}  // namespace

extern "C" {

int vk_hip_pdhg_destroy(vk_hip_pdhg *s) {
    if (!s) return VK_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->graph) (void)hipGraphExecDestroy(s->graph);
    for (void *p : s->allocations) (void)hipFree(p);
    s->pan_a.release();
    s->pan_at.release();
    if (s->st) (void)hipFree(s->st);
    if (s->host_state) (void)hipHostFree(s->host_state);
    if (s->ev_start) (void)hipEventDestroy(s->ev_start);
    if (s->ev_stop) (void)hipEventDestroy(s->ev_stop);
    for (int i = 0; i < 2; ++i) {
        if (s->ev_snap[i]) (void)hipEventDestroy(s->ev_snap[i]);
        if (s->snap[i]) (void)hipHostFree(s->snap[i]);
    }
    if (s->side) (void)hipStreamSynchronize(s->side);
    if (s->ev_fork) (void)hipEventDestroy(s->ev_fork);
    if (s->ev_join) (void)hipEventDestroy(s->ev_join);
    if (s->side) (void)hipStreamDestroy(s->side);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return VK_OK;
}

int vk_hip_pdqp_create(const vk_hip_csr *A, const vk_hip_csr *Q, const double *c, const double *l,
                       const double *u, const double *rl, const double *ru, const double *row_scale,
                       const double *col_scale, const double *x0, const double *y0,
                       const vk_pdhg_params *params, vk_hip_pdhg **out) {
    VK_REQUIRE(out && A && params, "vk_hip_pdqp_create: null argument");
    *out = nullptr;
    if (Q) {
        VK_REQUIRE(Q->has_values, "vk_hip_pdqp_create: Q values were never set");
        VK_REQUIRE(Q->rows == A->cols && Q->cols == A->cols, "vk_hip_pdqp_create: Q must be n x n");
        VK_REQUIRE((params->precision == VK_PREC_FP64 && params->policy == VK_PDHG_POLICY_REFERENCE &&
                    params->relative_kkt_tolerance == 0.0) ||
                       (params->policy == VK_PDHG_POLICY_HALPERN &&
                        (params->precision == VK_PREC_FP64 || params->recentre == 1)),
                   "vk_hip_pdqp_create: a quadratic objective runs the reference policy (pdqp.cpp) in FP64 "
                   "without the relative-KKT test, or the Halpern policy in FP64 or recentred FP32");
    }
    VK_REQUIRE(A->has_values, "vk_hip_pdhg_create: matrix values were never set");
    VK_REQUIRE(A->cols > 0, "vk_hip_pdhg_create: the LP needs at least one variable");
    VK_REQUIRE(c && l && u && col_scale, "vk_hip_pdhg_create: null column vector");
    VK_REQUIRE(A->rows == 0 || (rl && ru && row_scale), "vk_hip_pdhg_create: null row vector");
    VK_REQUIRE(params->policy == VK_PDHG_POLICY_REFERENCE || params->policy == VK_PDHG_POLICY_ADAPTIVE ||
                   params->policy == VK_PDHG_POLICY_HALPERN,
               "vk_hip_pdhg_create: unknown policy");
    VK_REQUIRE(params->precision == VK_PREC_FP64 || params->precision == VK_PREC_MIXED,
               "vk_hip_pdhg_create: unknown precision");
    VK_REQUIRE(params->iteration_limit >= 0, "vk_hip_pdhg_create: negative iteration limit");
    VK_REQUIRE(params->step_size > 0.0 && std::isfinite(params->step_size),
               "vk_hip_pdhg_create: step size must be positive");
    VK_REQUIRE(params->policy == VK_PDHG_POLICY_REFERENCE ||
                   (params->primal_weight > 0.0 && params->restart_period > 0),
               "vk_hip_pdhg_create: the adaptive and Halpern policies need a primal weight and restart period");
    VK_REQUIRE(params->col_blocks >= 0 && params->col_blocks <= kMaxColBlocks,
               "vk_hip_pdhg_create: col_blocks must be in [0, 16]");
    VK_REQUIRE(params->split_passes >= 0 && params->split_passes <= 2,
               "vk_hip_pdhg_create: split_passes must be 0, 1 or 2");
    VK_REQUIRE(params->cache_budget_mb >= -1, "vk_hip_pdhg_create: cache_budget_mb must be >= -1");
    VK_REQUIRE(params->spmv_panels >= 0 && params->spmv_panels <= 2, "vk_hip_pdhg_create: spmv_panels must be 0, 1 or 2");
    VK_REQUIRE(valid_panel_groups(params->panel_groups) && valid_panel_groups(params->panel_groups_t),
               "vk_hip_pdhg_create: panel groups must be 0 (automatic), 1, 2, 4 or 8");
    VK_REQUIRE(params->panel_threads == 0 || params->panel_threads == 256 || params->panel_threads == 512,
               "vk_hip_pdhg_create: panel_threads must be 0 (automatic), 256 or 512");
    VK_REQUIRE(params->panel_kb == 0 || panel_kb_valid(params->panel_kb),
               "vk_hip_pdhg_create: panel_kb must be 0 (automatic), 48, 24 or 16");
    VK_REQUIRE(!(params->panel_threads == 256 && params->panel_kb > 0 && params->panel_kb != 48),
               "vk_hip_pdhg_create: 256 lanes per panel workgroup need 48 KB panels");
    VK_REQUIRE(params->panel_tune == 0 || params->panel_tune == 1, "vk_hip_pdhg_create: panel_tune must be 0 or 1");
    VK_REQUIRE(params->recentre == 0 || params->recentre == 1, "vk_hip_pdhg_create: recentre must be 0 or 1");
    VK_REQUIRE(params->primal_weight_gain >= 0.0 && params->primal_weight_gain <= 1.0 &&
                   params->primal_weight_integral >= 0.0 && params->primal_weight_integral <= 1.0,
               "vk_hip_pdhg_create: primal-weight gains must be in [0, 1]");
    VK_REQUIRE(params->recentre == 0 ||
                   (params->policy == VK_PDHG_POLICY_HALPERN && params->precision == VK_PREC_MIXED),
               "vk_hip_pdhg_create: recentre needs the Halpern policy in VK_PREC_MIXED");
    if (int rc = ensure_init()) return rc;
    auto *s = new (std::nothrow) vk_hip_pdhg();
    if (!s) return VK_ENOMEM;
    s->params = *params;
    s->A = A;
    s->Q = Q;
    s->m = static_cast<int>(A->rows);
    s->n = static_cast<int>(A->cols);
    s->mixed = params->precision == VK_PREC_MIXED;
    s->recentre = params->recentre == 1;
    s->grid_a64 = std::max(1, std::min((A->a.ntiles + kWaves - 1) / kWaves, kMaxGrid));
    s->grid_at64 = std::max(1, std::min((A->at.ntiles + kWaves - 1) / kWaves, kMaxGrid));
    const int cap = std::min(params->grid_blocks > 0 ? params->grid_blocks : 2048, kMaxGrid);
    // One tile per wave, kWaves waves per block.
    s->grid_rows = std::max(1, std::min((A->a.ntiles + kWaves - 1) / kWaves, cap));
    s->grid_cols = std::max(1, std::min((A->at.ntiles + kWaves - 1) / kWaves, cap));
    s->col_blocks = params->col_blocks > 0 ? params->col_blocks
                                           : auto_col_blocks(s->n, s->mixed ? 4 : 8, s->m);
    s->col_blocks = std::max(1, std::min(s->col_blocks, s->n));
    if (s->m == 0) s->col_blocks = 1;
    // With column blocks the row pass (kPassItems rows per thread) owns
    // grid_rows.
    s->split = params->split_passes == 2 ||
               (params->split_passes == 0 && auto_split(s->m, s->n));
    s->grid_spmv_rows = std::max(1, std::min((A->a.ntiles + kWaves - 1) / kWaves, kMaxGrid));
    // The row pass runs kPassItems rows per thread over grid_rows workgroups
    // (which also size the row partial buffer).  In split mode the A' side is
    // k_primal_own over grid_cols workgroups, each with a contiguous tile
    // range (about kWaves tiles).
    if (s->col_blocks > 1 || s->split)
        s->grid_rows = std::max(1, std::min((s->m + kBlock * kPassItems - 1) / (kBlock * kPassItems), cap));
    auto fail = [&](int rc) {
        vk_hip_pdhg_destroy(s);
        return rc;
    };
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s->ev_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreate(&s->ev_start) != hipSuccess || hipEventCreate(&s->ev_stop) != hipSuccess ||
        device_malloc(&s->st, sizeof(DevState)) != hipSuccess ||
        hipHostMalloc(&s->host_state, sizeof(DevState)) != hipSuccess ||
        hipHostMalloc(&s->snap[0], sizeof(DevState)) != hipSuccess ||
        hipHostMalloc(&s->snap[1], sizeof(DevState)) != hipSuccess ||
        hipEventCreateWithFlags(&s->ev_snap[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s->ev_snap[1], hipEventDisableTiming) != hipSuccess) {
        set_error("vk_hip_pdhg_create: runtime resource allocation failed");
        return fail(VK_EDEVICE);
    }
    DevState init{};
    init.limit = params->iteration_limit;
    init.period = params->policy == VK_PDHG_POLICY_REFERENCE ? kDecisionPeriod : params->restart_period;
    init.status = params->iteration_limit == 0 ? VK_PDHG_ITERATION_LIMIT : VK_PDHG_RUNNING;
    init.policy = params->policy;
    init.metric_tol = params->metric_tolerance;
    init.rel_tol = params->relative_kkt_tolerance;
    init.norm_b = params->norm_b;
    init.norm_c = params->norm_c;
    init.max_step = params->step_size;
    init.prev_metric = std::numeric_limits<double>::infinity();
    init.metric = std::numeric_limits<double>::infinity();
    if (params->policy == VK_PDHG_POLICY_REFERENCE) {
        init.tau = params->step_size;   // pdhg.cpp:76-78
        init.sigma = params->step_size;
        init.omega = 1.0;
    } else {
        init.eta = params->step_size;
        init.omega = params->primal_weight;
        init.tau = init.eta / init.omega;
        init.sigma = init.eta * init.omega;
    }
    init.inv_sigma = 1.0 / init.sigma;
    init.pw_kp = params->primal_weight_gain > 0.0 ? params->primal_weight_gain : 0.5;
    init.pw_ki = params->primal_weight_integral;
    init.kkt_ref = std::numeric_limits<double>::infinity();
    init.kkt_prev = std::numeric_limits<double>::infinity();
    init.rel_kkt = std::numeric_limits<double>::infinity();
    *s->host_state = init;
    if (hipMemcpy(s->st, &init, sizeof(DevState), hipMemcpyHostToDevice) != hipSuccess) {
        set_error("vk_hip_pdhg_create: state upload failed");
        return fail(VK_EDEVICE);
    }
    int rc = s->mixed ? csr_make_f32(const_cast<vk_hip_csr *>(A), s->stream) : VK_OK;
    if (!rc && s->mixed && Q) rc = csr_make_f32(const_cast<vk_hip_csr *>(Q), s->stream);
    if (rc) return fail(rc);
    if (params->spmv_panels == 2 || (params->spmv_panels == 0 && auto_panels(s->m, s->n))) {
        rc = build_panels(s);
        if (rc == VK_OK) {
            // Row pass over k_dual_rows' grid, column pass over k_primal_cols'
            // (kPassItems items per thread).
            s->panels = true;
            s->col_blocks = 1;
            s->split = true;
            s->grid_rows = std::max(1, std::min((s->m + kBlock * kPassItems - 1) / (kBlock * kPassItems), cap));
            s->grid_cols = std::max(1, std::min((s->n + kBlock * kPassItems - 1) / (kBlock * kPassItems), cap));
        } else if (rc == VK_EINVAL && params->spmv_panels == 0) {
            s->pan_a.release();
            s->pan_at.release();
        } else {
            if (rc == VK_EINVAL) set_error("vk_hip_pdhg_create: the matrix does not fit the panel layout "
                                           "(a row with too many entries in one panel, or too many panels)");
            return fail(rc);
        }
    }
    rc = s->mixed ? setup_solver<float>(s, c, l, u, rl, ru, row_scale, col_scale, x0, y0)
                  : setup_solver<double>(s, c, l, u, rl, ru, row_scale, col_scale, x0, y0);
    if (rc) return fail(rc);
    *out = s;
    return VK_OK;
}

int vk_hip_pdhg_create(const vk_hip_csr *A, const double *c, const double *l, const double *u,
                       const double *rl, const double *ru, const double *row_scale,
                       const double *col_scale, const double *x0, const double *y0,
                       const vk_pdhg_params *params, vk_hip_pdhg **out) {
    return vk_hip_pdqp_create(A, nullptr, c, l, u, rl, ru, row_scale, col_scale, x0, y0, params, out);
}

int vk_hip_pdhg_run(vk_hip_pdhg *s, int64_t max_iterations, int *status) {
    VK_REQUIRE(s && status && max_iterations >= 0, "vk_hip_pdhg_run: bad arguments");
    if (int rc = bind_device()) return rc;
    return s->mixed ? run_solver<float>(s, max_iterations, status)
                    : run_solver<double>(s, max_iterations, status);
}

int vk_hip_pdhg_read_state(vk_hip_pdhg *s, vk_pdhg_state *out) {
    VK_REQUIRE(s && out, "vk_hip_pdhg_read_state: null argument");
    if (int rc = bind_device()) return rc;
    VK_HIP_CHECK(hipMemcpyAsync(s->host_state, s->st, sizeof(DevState), hipMemcpyDeviceToHost, s->stream));
    VK_HIP_CHECK(hipStreamSynchronize(s->stream));
    const DevState &d = *s->host_state;
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

int vk_hip_pdhg_download(vk_hip_pdhg *s, double *x, double *y) {
    VK_REQUIRE(s, "vk_hip_pdhg_download: null solver");
    if (int rc = bind_device()) return rc;
    return s->mixed ? download_solution<float>(s, x, y) : download_solution<double>(s, x, y);
}

int vk_hip_pdhg_profile(vk_hip_pdhg *s, int64_t iterations, double *ms) {
    VK_REQUIRE(s && ms && iterations > 0, "vk_hip_pdhg_profile: bad arguments");
    return s->mixed ? profile_iterations<float>(s, iterations, ms) : profile_iterations<double>(s, iterations, ms);
}

int vk_hip_pdhg_panel_layout(const vk_hip_pdhg *s, int *info) {
    VK_REQUIRE(s && info, "vk_hip_pdhg_panel_layout: null argument");
    const PanelLayout *L[2] = {&s->pan_a, &s->pan_at};
    for (int i = 0; i < 2; ++i) {
        info[5 * i] = s->panels ? L[i]->groups : 0;
        info[5 * i + 1] = s->panels ? L[i]->threads : 0;
        info[5 * i + 2] = s->panels ? L[i]->steps : 0;
        info[5 * i + 3] = s->panels ? L[i]->fast_steps : 0;
        info[5 * i + 4] = s->panels ? L[i]->panel_kb : 0;
    }
    return VK_OK;
}

int vk_hip_pdhg_layout(const vk_hip_pdhg *s, int *col_blocks, int *split_passes) {
    VK_REQUIRE(s && col_blocks && split_passes, "vk_hip_pdhg_layout: null argument");
    *col_blocks = s->col_blocks;
    *split_passes = s->split ? 1 : 0;
    return VK_OK;
}

int vk_hip_pdhg_last_run_ms(const vk_hip_pdhg *s, double *ms) {
    VK_REQUIRE(s && ms, "vk_hip_pdhg_last_run_ms: null argument");
    *ms = s->last_ms;
    return VK_OK;
}

int vk_hip_csr_tune_tiles(vk_hip_csr *A, int precision, int repeats, int *best, double *ms) {
    VK_REQUIRE(A && best && repeats > 0, "vk_hip_csr_tune_tiles: bad arguments");
    VK_REQUIRE(A->has_values, "vk_hip_csr_tune_tiles: matrix values were never set");
    VK_REQUIRE(precision == VK_PREC_FP64 || precision == VK_PREC_MIXED, "vk_hip_csr_tune_tiles: unknown precision");
    if (int rc = ensure_init()) return rc;
    hipStream_t st = default_stream();
    const bool f32 = precision == VK_PREC_MIXED;
    if (f32)
        if (int rc = csr_make_f32(A, st)) return rc;
    const size_t m = static_cast<size_t>(A->rows), n = static_cast<size_t>(A->cols);
    const size_t vb = f32 ? sizeof(float) : sizeof(double);
    void *x = nullptr, *y = nullptr, *ax = nullptr, *aty = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = VK_OK;
    auto check = [&](hipError_t e) {
        if (e != hipSuccess && rc == VK_OK) {
            set_error(std::string("vk_hip_csr_tune_tiles: ") + hipGetErrorString(e));
            rc = e == hipErrorOutOfMemory ? VK_ENOMEM : VK_EDEVICE;
        }
    };
    check(device_malloc(&x, vb * n + 8));
    check(device_malloc(&y, vb * m + 8));
    check(device_malloc(&ax, vb * m + 8));
    check(device_malloc(&aty, vb * n + 8));
    check(hipEventCreate(&e0));
    check(hipEventCreate(&e1));
    if (rc == VK_OK) check(hipMemsetAsync(x, 0, vb * n + 8, st));
    if (rc == VK_OK) check(hipMemsetAsync(y, 0, vb * m + 8, st));
    // One A x and one A' y per repeat, as an iteration runs them.
    auto launch = [&] {
        auto grid = [](int ntiles) { return std::max(1, std::min((ntiles + kWaves - 1) / kWaves, kMaxGrid)); };
        if (f32) {
            const CsrView<float> va = A->a32.view(), vt = A->at32.view();
            if (va.ntiles > 0) k_spmv_store<float><<<grid(va.ntiles), kBlock, 0, st>>>(va, static_cast<float *>(x), static_cast<float *>(ax));
            if (vt.ntiles > 0) k_spmv_store<float><<<grid(vt.ntiles), kBlock, 0, st>>>(vt, static_cast<float *>(y), static_cast<float *>(aty));
        } else {
            const CsrView<double> va = A->a.view(), vt = A->at.view();
            if (va.ntiles > 0) k_spmv_store<double><<<grid(va.ntiles), kBlock, 0, st>>>(va, static_cast<double *>(x), static_cast<double *>(ax));
            if (vt.ntiles > 0) k_spmv_store<double><<<grid(vt.ntiles), kBlock, 0, st>>>(vt, static_cast<double *>(y), static_cast<double *>(aty));
        }
    };
    double best_ms = std::numeric_limits<double>::infinity();
    int chosen = 0;
    for (int g = 0; g < kTileGeometries && rc == VK_OK; ++g) {
        if (int grc = vk_hip_csr_set_tile_geometry(A, g)) {
            rc = grc;
            break;
        }
        launch();   // warm-up
        check(hipEventRecord(e0, st));
        for (int r = 0; r < repeats; ++r) launch();
        check(hipEventRecord(e1, st));
        check(hipEventSynchronize(e1));
        check(hipGetLastError());
        float t = 0.0f;
        check(hipEventElapsedTime(&t, e0, e1));
        const double mean = static_cast<double>(t) / repeats;
        if (ms) ms[g] = mean;
        if (rc == VK_OK && mean < best_ms) {
            best_ms = mean;
            chosen = g;
        }
    }
    if (x) (void)hipFree(x);
    if (y) (void)hipFree(y);
    if (ax) (void)hipFree(ax);
    if (aty) (void)hipFree(aty);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc) return rc;
    *best = chosen;
    return vk_hip_csr_set_tile_geometry(A, chosen);
}

int vk_hip_csr_tune_col_blocks(vk_hip_csr *A, int precision, int repeats, int *best, double *ms) {
    VK_REQUIRE(A && best && repeats > 0, "vk_hip_csr_tune_col_blocks: bad arguments");
    VK_REQUIRE(A->has_values, "vk_hip_csr_tune_col_blocks: matrix values were never set");
    VK_REQUIRE(precision == VK_PREC_FP64 || precision == VK_PREC_MIXED,
               "vk_hip_csr_tune_col_blocks: unknown precision");
    if (int rc = ensure_init()) return rc;
    if (precision == VK_PREC_MIXED) {
        if (int rc = csr_make_f32(A, default_stream())) return rc;
        return tune_col_blocks<float>(A, A->a32.view(), static_cast<int>(repeats), best, ms);
    }
    return tune_col_blocks<double>(A, A->a.view(), static_cast<int>(repeats), best, ms);
}

}  // extern "C"
