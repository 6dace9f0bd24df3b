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

// Tile geometry of the CSR-adaptive SpMV: a tile is a run of whole rows with
// at most kTileNnz nonzeros (8 per lane of a wave64) and at most kTileRows
// rows (one per lane), or a single row longer than kTileNnz.  Each tile is
// processed by ONE wave with a private LDS slice, so waves never wait on a
// block barrier inside the SpMV.
constexpr int kTileNnz = 512;
constexpr int kTileRows = 64;

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
// lanes*rows <= kWave.
__device__ __forceinline__ int lanes_per_row(int rows) {
    int g = 1;
    while (g < kWave && g * 2 * rows <= kWave) g *= 2;
    return g;
}

// Orders this wave's LDS writes before its later LDS reads (and vice versa)
// without a block barrier.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Device-side view of one CSR operand (A or its stored transpose).
//
// `resident` splits the tiles by how their column indices, values and row
// offsets are loaded: tiles [0, resident) with ordinary loads, which the
// 256 MB memory-side cache (MALL) keeps, the rest with non-temporal loads,
// which it does not allocate.  An iteration streams A and A' once each; when
// both together exceed the MALL, ordinary loads evict each other every
// iteration and nothing is reused, while marking only what fits keeps that
// part in the MALL from one iteration to the next.  The split changes load
// hints only, never results.
template <class V>
struct CsrView {
    const int *row_ptr;     // rows+1
    const int *col;         // nnz
    const V *val;           // nnz
    const int2 *tiles;      // ntiles+1: {first row, first nonzero} of each tile
    int rows;
    int ntiles;
    int resident;           // tiles [0, resident) use cache-allocating loads
};

// Loads of matrix data: cache-allocating (Keep) or non-temporal.
template <bool Keep, class P>
__device__ __forceinline__ P matrix_load(const P *p) {
    if constexpr (Keep) return *p;
    else return __builtin_nontemporal_load(p);
}

// Loads and stores of vectors that the next kernel does not gather:
// non-temporal in the split row / column passes (Stream), so the matrix
// data marked resident stays in the MALL.
template <bool Stream, class P>
__device__ __forceinline__ P vec_load(const P *p) {
    if constexpr (Stream) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool Stream, class P>
__device__ __forceinline__ void vec_store(P *p, P v) {
    if constexpr (Stream) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// Column-blocked copy of a CSR operand, for when the gathered vector
// outgrows one XCD's L2.  Block p holds the entries of columns
// [n*p/nb, n*(p+1)/nb) of every row as a CSR of its own (offsets into shared
// col/val arrays, blocks stored one after another) and is tiled like the
// plain operand.  Workgroup g works on block g % nb: workgroups are
// dispatched round-robin over the 8 XCDs, so block p's slice of x is
// gathered by one XCD (nb = 8) or two (nb = 4) and stays in their L2.  Every
// block writes one partial sum per row; a row pass adds the nb partials in
// block order, so results repeat bitwise.
constexpr int kMaxColBlocks = 16;

template <class V>
struct BlockedView {
    const int *row_ptr;   // nb * (rows + 1), global offsets into col/val
    const int *col;
    const V *val;
    const int2 *tiles;    // block p: entries [tile_first[p], tile_first[p+1]),
    int rows;             //   the last one its end sentinel
    int nb;
    int tile_first[kMaxColBlocks + 1];
    int resident;         // global tile index below which loads allocate in the MALL

    __device__ CsrView<V> block(int p) const {
        const int nt = tile_first[p + 1] - tile_first[p] - 1;
        return CsrView<V>{row_ptr + static_cast<long>(p) * (rows + 1), col, val, tiles + tile_first[p], rows, nt,
                          max(0, min(nt, resident - tile_first[p]))};
    }
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

// Reduces all kSlots accumulators of a block of B threads at once (slot 0
// with max, the others with +): interleaved shuffle trees in every wave, one
// LDS exchange, then thread s folds slot s over the waves in wave order.
// Returns slot threadIdx.x in threads 0..kSlots-1 (other threads: unused).
// scratch holds kSlots * B / kWave doubles; no barrier follows.
template <int B>
__device__ __forceinline__ double block_reduce_slots(double (&v)[kSlots], double *scratch) {
    constexpr int W = B / kWave;
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
#pragma unroll
        for (int s = 0; s < kSlots; ++s) {
            const double o = __shfl_down(v[s], off, kWave);
            v[s] = (s == 0) ? dmax(v[s], o) : v[s] + o;
        }
    }
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    if (lane == 0) {
#pragma unroll
        for (int s = 0; s < kSlots; ++s) scratch[s * W + wave] = v[s];
    }
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x < kSlots) {
        const int s = threadIdx.x;
        r = scratch[s * W];
        for (int w = 1; w < W; ++w) r = (s == 0) ? dmax(r, scratch[s * W + w]) : r + scratch[s * W + w];
    }
    return r;
}

// Shared-memory doubles a kBlock-thread kernel needs for store_partials.
constexpr int kSlotScratch = kSlots * kWaves;

// Block partials of a kBlock-thread kernel: slot s of block b goes to
// partials[s * gridDim.x + b].
__device__ __forceinline__ void store_partials(const Partials &p, double *partials,
                                               double *scratch) {
    double v[kSlots];
#pragma unroll
    for (int s = 0; s < kSlots; ++s) v[s] = p.v[s];
    const double r = block_reduce_slots<kBlock>(v, scratch);
    if (threadIdx.x < kSlots) partials[threadIdx.x * gridDim.x + blockIdx.x] = r;
}

// Threads of the one-block kernels that reduce per-block partials (the
// finalizer and the checks), and the most blocks an SpMV grid uses.
constexpr int kFinalBlock = 1024;
constexpr int kMaxGrid = 4 * kFinalBlock;

// Block-wide sum / max over a block of B threads; every thread receives the
// result.  `scratch` holds B / kWave doubles.
template <int B = kFinalBlock>
__device__ __forceinline__ double final_sum(double v, double *scratch) {
    v = wave_sum(v);
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    double r = 0.0;
    for (int w = 0; w < B / kWave; ++w) r += scratch[w];
    return r;
}

template <int B = kFinalBlock>
__device__ __forceinline__ double final_max(double v, double *scratch) {
    v = wave_max(v);
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    double r = scratch[0];
    for (int w = 1; w < B / kWave; ++w) r = dmax(r, scratch[w]);
    return r;
}

// Shared-memory doubles reduce_partials and reduce_partials_pair need.
constexpr int kFinalScratch = 2 * kSlots * kFinalBlock / kWave;

// Final reduction of per-block partial slots (partials[s * nblocks + b],
// slot 0 a maximum of non-negative values, the others sums) by one
// kFinalBlock-thread block.  Thread t reduces slot t % kSlots over blocks
// t / kSlots, t / kSlots + kFinalBlock / kSlots, ... in order; the kSlots
// lanes apart that share a slot in a wave are then combined by a 3-level
// shuffle, and the waves' results in index order through LDS.  One value
// per shuffle instead of all kSlots keeps the cross-lane traffic small: the
// whole reduction costs about what its loads do.  The order is fixed, so
// results repeat bitwise.  scratch holds kFinalScratch doubles; `sets`
// reductions (1 or 2) run together, set i from (p[i], n[i]) into out[i].
template <int Sets>
__device__ __forceinline__ void reduce_partial_sets(const double *const (&p)[Sets], const int (&n)[Sets],
                                                    double *const (&out)[Sets], double *scratch) {
    constexpr int kPerSlot = kFinalBlock / kSlots;   // threads per slot
    constexpr int kW = kFinalBlock / kWave;          // waves
    const int s = threadIdx.x % kSlots, q = threadIdx.x / kSlots;
    double acc[Sets];
    int nmax = 0;
#pragma unroll
    for (int i = 0; i < Sets; ++i) {
        acc[i] = 0.0;
        nmax = max(nmax, n[i]);
    }
    // kChunk loads per set in flight at a time, combined in block order
    // (out-of-range entries are 0, the identity of both max and sum here).
    constexpr int kChunk = 8;
#pragma unroll 1
    for (int b0 = q; b0 < nmax; b0 += kChunk * kPerSlot) {
        double v[Sets][kChunk];
#pragma unroll
        for (int i = 0; i < Sets; ++i)
#pragma unroll
            for (int u = 0; u < kChunk; ++u) {
                const int b = b0 + u * kPerSlot;
                v[i][u] = b < n[i] ? p[i][s * n[i] + b] : 0.0;
            }
#pragma unroll
        for (int i = 0; i < Sets; ++i)
#pragma unroll
            for (int u = 0; u < kChunk; ++u) acc[i] = (s == 0) ? dmax(acc[i], v[i][u]) : acc[i] + v[i][u];
    }
#pragma unroll
    for (int off = kWave / 2; off >= kSlots; off >>= 1)
#pragma unroll
        for (int i = 0; i < Sets; ++i) {
            const double o = __shfl_down(acc[i], off, kWave);
            acc[i] = (s == 0) ? dmax(acc[i], o) : acc[i] + o;
        }
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    if (lane < kSlots)
#pragma unroll
        for (int i = 0; i < Sets; ++i) scratch[(i * kSlots + lane) * kW + wave] = acc[i];
    __syncthreads();
    if (threadIdx.x < Sets * kSlots) {
        const int i = threadIdx.x / kSlots, t = threadIdx.x % kSlots;
        const double *w = scratch + (i * kSlots + t) * kW;
        double r = w[0];
        for (int k = 1; k < kW; ++k) r = (t == 0) ? dmax(r, w[k]) : r + w[k];
#pragma unroll
        for (int j = 0; j < Sets; ++j)
            if (j == i) out[j][t] = r;
    }
    __syncthreads();
}

__device__ __forceinline__ void reduce_partials(const double *__restrict__ partials, int nblocks,
                                                double *out, double *scratch) {
    const double *p[1] = {partials};
    const int n[1] = {nblocks};
    double *o[1] = {out};
    reduce_partial_sets<1>(p, n, o, scratch);
}

// reduce_partials for two partial sets at once (rows and columns of one
// iteration), their loads interleaved.
__device__ __forceinline__ void reduce_partials_pair(const double *__restrict__ pa, int na,
                                                     const double *__restrict__ pb, int nb, double *out_a,
                                                     double *out_b, double *scratch) {
    const double *p[2] = {pa, pb};
    const int n[2] = {na, nb};
    double *o[2] = {out_a, out_b};
    reduce_partial_sets<2>(p, n, o, scratch);
}

// CSR-adaptive SpMV, one tile per wave (grid-stride over waves).
//
//   header   one 16-byte load gives the tile's first/last row and nonzero;
//   phase 1  every lane issues its (up to) 8 column and value loads, the row
//            offsets of the tile and the epilogue's per-row prefetch at once,
//            then its 8 gathers of x, then 8 products into the wave's private
//            LDS slice -- explicit batching so no load waits on another;
//   phase 2  each row is reduced by a fixed power-of-two lane group and a
//            shuffle tree, and row r's sum moves to lane r;
//   phase 3  Epi::row on consecutive lanes for consecutive rows.
// A row longer than kTileNnz is a tile of its own, summed by the whole wave
// straight from memory.  No block barrier, no atomics; the summation order
// depends only on the tiling, so results repeat bitwise.
//
// Epi provides `Pre prefetch(int row) const` (loads issued before the
// gathers) and `row(int row, double acc, const Pre &, Partials &)`; every row
// is seen exactly once.
// Phase 1's matrix loads for one tile of at most kTileNnz nonzeros.
template <bool Keep, class V>
__device__ __forceinline__ void load_tile(const int *__restrict__ colp, const V *__restrict__ valp,
                                          const int *__restrict__ row_ptr, int k0, int nnz, int r0, int nr,
                                          int lane, int (&c)[kTileNnz / kWave], V (&v)[kTileNnz / kWave],
                                          int &rs) {
#pragma unroll
    for (int u = 0; u < kTileNnz / kWave; ++u) {
        const int k = lane + u * kWave;
        c[u] = k < nnz ? matrix_load<Keep>(colp + k0 + k) : 0;
        v[u] = k < nnz ? matrix_load<Keep>(valp + k0 + k) : V(0);
    }
    rs = lane < nr ? matrix_load<Keep>(row_ptr + r0 + lane) - k0 : nnz;
}

//
// XLds: x points into the workgroup's LDS and holds entries [xoff, ...) of
// the gathered vector (x[i - xoff] is entry i); the gathers then read LDS
// and never touch the cache hierarchy.
template <class V, class X, class Epi, bool XLds = false>
__device__ __forceinline__ void tile_spmv_range(const CsrView<V> &A, const X *__restrict__ x, Epi &epi,
                                                Partials &part, double *lds_generic, int t_begin,
                                                int t_end, int wave, int nwaves, int xoff = 0) {
    using Pre = typename Epi::Pre;
    // The slice is LDS: say so, so LDS stores do not order global loads.
    auto *lds = (__attribute__((address_space(3))) double *)(lds_generic);
    auto gather = [&](int col) -> X {
        if constexpr (XLds) return ((__attribute__((address_space(3))) const X *)(x))[col - xoff];
        else return x[col];
    };
    constexpr int kPer = kTileNnz / kWave;
    const int lane = threadIdx.x % kWave;
    const int *__restrict__ colp = A.col;
    const V *__restrict__ valp = A.val;
    for (int t = t_begin + wave; t < t_end; t += nwaves) {
        const int2 head = A.tiles[t], tail = A.tiles[t + 1];
        const int r0 = head.x, r1 = tail.x, k0 = head.y, k1 = tail.y;
        const int nnz = k1 - k0;
        const int nr = r1 - r0;
        const bool keep = t < A.resident;
        if (nnz <= kTileNnz) {
            int c[kPer];
            V v[kPer];
            int rs;   // row starts of the tile, one per lane (lane r holds row r's start)
            if (keep) load_tile<true>(colp, valp, A.row_ptr, k0, nnz, r0, nr, lane, c, v, rs);
            else load_tile<false>(colp, valp, A.row_ptr, k0, nnz, r0, nr, lane, c, v, rs);
            Pre pre{};
            if (lane < nr) pre = epi.prefetch(r0 + lane);
            X xv[kPer];
#pragma unroll
            for (int u = 0; u < kPer; ++u) xv[u] = (lane + u * kWave < nnz) ? gather(c[u]) : X(0);
#pragma unroll
            for (int u = 0; u < kPer; ++u) {
                const int k = lane + u * kWave;
                if (k < nnz) lds[k] = static_cast<double>(v[u]) * static_cast<double>(xv[u]);
            }
            wave_sync();
            const int g = lanes_per_row(nr);
            const int lr = lane / g, sub = lane % g;
            // A tile can hold 64 rows, so the last row's end is the tile's
            // nonzero count rather than a 65th lane.
            const int s = __shfl(rs, lr < nr ? lr : 0, kWave);
            const int e_next = __shfl(rs, (lr + 1 < nr) ? lr + 1 : 0, kWave);
            const int e = (lr + 1 < nr) ? e_next : nnz;
            double acc = 0.0;
            if (lr < nr)
                for (int k = s + sub; k < e; k += g) acc += lds[k];
            for (int off = g / 2; off > 0; off >>= 1) acc += __shfl_down(acc, off, g);
            const double row_sum = __shfl(acc, (lane * g) % kWave, kWave);
            wave_sync();   // all lanes are done with the slice before the next tile
            if (lane < nr) epi.row(r0 + lane, row_sum, pre, part);
        } else {
            Pre pre{};
            if (lane == 0) pre = epi.prefetch(r0);
            double acc = 0.0;
            if (keep)
                for (int k = k0 + lane; k < k1; k += kWave)
                    acc += static_cast<double>(valp[k]) * static_cast<double>(gather(colp[k]));
            else
                for (int k = k0 + lane; k < k1; k += kWave)
                    acc += static_cast<double>(__builtin_nontemporal_load(valp + k)) *
                           static_cast<double>(gather(__builtin_nontemporal_load(colp + k)));
            acc = wave_sum(acc);
            if (lane == 0) epi.row(r0, acc, pre, part);
        }
    }
}

// All tiles of A, one per wave of the grid (grid-stride).
template <class V, class X, class Epi>
__device__ __forceinline__ void tile_spmv(const CsrView<V> &A, const X *__restrict__ x, Epi &epi,
                                          Partials &part, double *lds) {
    tile_spmv_range(A, x, epi, part, lds, 0, A.ntiles,
                    static_cast<int>(blockIdx.x) * kWaves + static_cast<int>(threadIdx.x) / kWave,
                    static_cast<int>(gridDim.x) * kWaves);
}

}  // namespace vk
