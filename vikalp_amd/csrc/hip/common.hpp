// Shared device helpers for the gfx950 PDHG kernels.
//
// Conventions
//   - one block = 256 threads = 4 wave64s;
//   - every reduction runs in a fixed order (wave shuffles, then LDS across
//     waves, then a fixed per-block slot layout), so results are bitwise
//     repeatable run to run; no floating-point atomics are used anywhere;
//   - sparse matrices are CSR with int32 row offsets and column indices.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vk {

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kWaves = kBlock / kWave;

// Tile geometry of the CSR-adaptive SpMV: a tile is a run of whole rows whose
// nonzeros fit in kTileNnz LDS slots (or one row longer than that), with at
// most kTileRows rows so that per-row sums fit in LDS too.
constexpr int kTileNnz = 2048;
constexpr int kTileRows = 1024;

// std::min(std::max(v, lo), hi) with the reference's argument order, so the
// NaN behaviour and the choice between equal values match the CPU path.
template <class T>
__device__ __forceinline__ T clamp_box(T v, T lo, T hi) {
    T t = (v < lo) ? lo : v;
    return (hi < t) ? hi : t;
}

template <class T>
__device__ __forceinline__ T dmax(T a, T b) { return (a < b) ? b : a; }

// Wave-level sum with a fixed butterfly order.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v = dmax(v, __shfl_down(v, off, kWave));
    return v;
}

// Block reductions: result valid in thread 0. `scratch` holds kWaves doubles.
__device__ __forceinline__ double block_sum(double v, double *scratch) {
    v = wave_sum(v);
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < kWaves; ++w) r += scratch[w];
    }
    return r;
}

__device__ __forceinline__ double block_max(double v, double *scratch) {
    v = wave_max(v);
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        r = scratch[0];
        for (int w = 1; w < kWaves; ++w) r = dmax(r, scratch[w]);
    }
    return r;
}

// Lanes per row for a tile of `rows` rows: the largest power of two with
// lanes*rows <= kBlock, capped at one wave so a group never spans waves.
__device__ __forceinline__ int lanes_per_row(int rows) {
    int g = 1;
    while (g < kWave && g * 2 * rows <= kBlock) g *= 2;
    return g;
}

// Device-side view of one CSR operand (A or its stored transpose).
template <class V>
struct CsrView {
    const int *row_ptr;     // rows+1
    const int *col;         // nnz
    const V *val;           // nnz
    const int *tile_row;    // ntiles+1, first row of each tile
    int rows;
    int ntiles;
};

// Number of fixed partial slots every fused kernel writes per block.
constexpr int kSlots = 8;

// Per-thread partial accumulators.  Slot 0 combines with max, the others
// with +.  Each block stores slot s at partials[s * gridDim.x + blockIdx.x].
struct Partials {
    double v[kSlots];
    __device__ void zero() {
#pragma unroll
        for (int s = 0; s < kSlots; ++s) v[s] = 0.0;
    }
};

__device__ __forceinline__ void store_partials(const Partials &p, double *partials,
                                               double *scratch) {
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
        const double r = (s == 0) ? block_max(p.v[s], scratch) : block_sum(p.v[s], scratch);
        if (threadIdx.x == 0) partials[s * gridDim.x + blockIdx.x] = r;
    }
}

// Reduce `nblocks` per-block partials in a fixed order; result in out[0..7]
// (valid in every thread after the call).
__device__ __forceinline__ void reduce_partials(const double *partials, int nblocks,
                                                double *out, double *scratch) {
    for (int s = 0; s < kSlots; ++s) {
        double acc = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += kBlock) {
            const double v = partials[s * nblocks + b];
            acc = (s == 0) ? dmax(acc, v) : acc + v;
        }
        const double r = (s == 0) ? block_max(acc, scratch) : block_sum(acc, scratch);
        if (threadIdx.x == 0) out[s] = r;
        __syncthreads();
    }
}

// CSR-adaptive SpMV over the tiles assigned to this block (grid-stride).
// Phase 1 streams the tile's nonzeros (coalesced) and stores val*x[col] in
// LDS; phase 2 reduces each row with a fixed lane group; phase 3 hands each
// row sum to the epilogue with consecutive threads on consecutive rows.
// Rows longer than kTileNnz form a tile of their own and are reduced by the
// whole block.  Epi::row(row, acc, partials) sees every row exactly once.
template <class V, class X, class Epi>
__device__ __forceinline__ void tile_spmv(const CsrView<V> &A, const X *__restrict__ x,
                                          Epi &epi, Partials &part, double *lds_prod,
                                          double *lds_rows, double *scratch) {
    for (int t = blockIdx.x; t < A.ntiles; t += gridDim.x) {
        const int r0 = A.tile_row[t], r1 = A.tile_row[t + 1];
        const int k0 = A.row_ptr[r0], k1 = A.row_ptr[r1];
        const int nnz = k1 - k0;
        const int nr = r1 - r0;
        if (nnz <= kTileNnz) {
            for (int k = threadIdx.x; k < nnz; k += kBlock) {
                const int kk = k0 + k;
                lds_prod[k] = static_cast<double>(A.val[kk]) * static_cast<double>(x[A.col[kk]]);
            }
            __syncthreads();
            const int g = lanes_per_row(nr);
            for (int base = 0; base < nr * g; base += kBlock) {
                const int idx = base + threadIdx.x;
                const int lr = idx / g, lane = idx % g;
                double acc = 0.0;
                if (lr < nr) {
                    const int s = A.row_ptr[r0 + lr] - k0, e = A.row_ptr[r0 + lr + 1] - k0;
                    for (int k = s + lane; k < e; k += g) acc += lds_prod[k];
                }
                for (int off = g / 2; off > 0; off >>= 1) acc += __shfl_down(acc, off, g);
                if (lane == 0 && lr < nr) lds_rows[lr] = acc;
            }
            __syncthreads();
            for (int lr = threadIdx.x; lr < nr; lr += kBlock) epi.row(r0 + lr, lds_rows[lr], part);
            __syncthreads();
        } else {
            double acc = 0.0;
            for (int k = k0 + threadIdx.x; k < k1; k += kBlock)
                acc += static_cast<double>(A.val[k]) * static_cast<double>(x[A.col[k]]);
            acc = block_sum(acc, scratch);
            if (threadIdx.x == 0) lds_rows[0] = acc;
            __syncthreads();
            if (threadIdx.x == 0) epi.row(r0, lds_rows[0], part);
            __syncthreads();
        }
    }
}

}  // namespace vk
