// Per-op kernels and the C ABI of libvikalp_hip.so (runtime, memory, vector
// ops, reductions, CSR handle, SpMV).  These are the device primitives the
// HipBackend (an ExecutionBackend, ExecutionBackend.hpp:36-64) is built on;
// the fused PDHG kernels live in pdhg.hip.

#include "runtime.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace vk {
namespace {

thread_local std::string g_error;
std::mutex g_mutex;
bool g_ready = false;
int g_device = 0;
hipStream_t g_stream = nullptr;
// Per-library reduction scratch: block partials + one result slot.
constexpr int kReduceBlocks = 1024;
double *g_reduce = nullptr;

}  // namespace

// Scratch of vk_hip_reduce_dot_noalloc: the block partials and the result
// slot of the two-stage dot product, in a file-scope device array (each GPU
// holds its own copy, 8.3 KB, until the process ends).  Host code uses the
// per-device address from hipGetSymbolAddress, never the array's name.
static __device__ __align__(256) double noalloc_scratch[kReduceBlocks + 8];
constexpr int kMaxNoallocDevices = 64;
static std::mutex g_scratch_mutex;
static double *g_scratch_ptr[kMaxNoallocDevices] = {};

void set_error(const std::string &message) { g_error = message; }

hipStream_t default_stream() { return g_stream; }

namespace {
// Device allocations left before every allocation fails (-1: never), and the
// allocations made since the hook was last called.
std::atomic<long long> g_allocs_left{-1};
std::atomic<long long> g_allocs_made{0};
}  // namespace

hipError_t device_alloc_bytes(void **p, std::size_t bytes) {
    *p = nullptr;
    for (long long left = g_allocs_left.load(); left >= 0;) {
        if (left == 0) return hipErrorOutOfMemory;
        if (g_allocs_left.compare_exchange_weak(left, left - 1)) break;
    }
    g_allocs_made.fetch_add(1);
    return hipMalloc(p, bytes);
}

int bind_device() {
    // The current device is per host thread: a thread other than the one
    // that initialised the library starts on device 0.
    thread_local int bound = -1;
    if (bound != g_device) {
        VK_HIP_CHECK(hipSetDevice(g_device));
        bound = g_device;
    }
    return VK_OK;
}

int ensure_init() {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_ready) return VK_OK;
    int count = 0;
    VK_HIP_CHECK(hipGetDeviceCount(&count));
    VK_REQUIRE(count > 0, "no HIP device is visible");
    VK_REQUIRE(g_device >= 0 && g_device < count, "device index out of range");
    VK_HIP_CHECK(hipSetDevice(g_device));
    if (!g_stream) VK_HIP_CHECK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    VK_HIP_CHECK(device_malloc(&g_reduce, sizeof(double) * (kReduceBlocks + 8)));
    g_ready = true;
    return VK_OK;
}

void append_tiles(const int *rp, int begin, int end, std::vector<int2> &tiles, const TileGeometry &g) {
    int cur_nnz = 0, cur_rows = 0;
    tiles.push_back(make_int2(begin, rp[begin]));
    for (int r = begin; r < end; ++r) {
        const int len = rp[r + 1] - rp[r];
        const bool alone = len > g.long_row;
        if (cur_rows > 0 && (alone || cur_nnz + len > g.max_nnz || cur_rows == g.max_rows)) {
            tiles.push_back(make_int2(r, rp[r]));
            cur_nnz = 0;
            cur_rows = 0;
        }
        cur_nnz += len;
        ++cur_rows;
        if (alone) {
            tiles.push_back(make_int2(r + 1, rp[r + 1]));
            cur_nnz = 0;
            cur_rows = 0;
        }
    }
    if (cur_rows > 0) tiles.push_back(make_int2(end, rp[end]));
}

std::vector<int2> build_tiles(const std::vector<int> &row_ptr, int rows, const TileGeometry &g) {
    std::vector<int2> tiles;
    append_tiles(row_ptr.data(), 0, rows, tiles, g);
    return tiles;
}

namespace {

// ------------------------------------------------------------------ kernels

__global__ void k_fill(double *__restrict__ y, int64_t n, double v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        y[i] = v;
}

// x may alias y (pdhg.cpp:109 calls axpby(a, y, 0, y)), so no __restrict__.
// Two roundings (product, then sum), no fused multiply-add: the same
// arithmetic as the CPU backend's loop (cpu_ops.cpp:341-342).
__global__ void k_axpby(int64_t n, double alpha, const double *x, double beta, double *y) {
#pragma clang fp contract(off)
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        y[i] = alpha * x[i] + beta * y[i];
}

__global__ void k_project(int64_t n, double *__restrict__ x, const double *__restrict__ lo,
                          const double *__restrict__ hi) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] = clamp_box(x[i], lo[i], hi[i]);
}

// Two-stage deterministic reductions: a fixed grid writes block partials,
// then one block combines them in index order.
template <bool kDot>
__global__ __launch_bounds__(kBlock) void k_reduce_stage1(int64_t n, const double *__restrict__ x,
                                                          const double *__restrict__ y,
                                                          double *__restrict__ partials) {
#pragma clang fp contract(off)
    __shared__ double scratch[kWaves];
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        if (kDot) acc += x[i] * y[i];
        else acc = dmax(acc, fabs(x[i]));
    }
    const double r = kDot ? block_sum(acc, scratch) : block_max(acc, scratch);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

template <bool kDot>
__global__ __launch_bounds__(kBlock) void k_reduce_stage2(int nblocks, const double *__restrict__ partials,
                                                          double *__restrict__ out) {
    __shared__ double scratch[kWaves];
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += kBlock)
        acc = kDot ? acc + partials[b] : dmax(acc, partials[b]);
    const double r = kDot ? block_sum(acc, scratch) : block_max(acc, scratch);
    if (threadIdx.x == 0) *out = r;
}

struct PlainEpi {
    double alpha, beta;
    double *y;
    using Pre = double;
    __device__ Pre prefetch(int r) const { return y[r]; }
    __device__ void row(int r, double acc, const Pre &y_old, Partials &) const {
#pragma clang fp contract(off)
        y[r] = alpha * acc + beta * y_old;   // cpu_ops.cpp:426, two roundings
    }
};

__global__ __launch_bounds__(kBlock) void k_spmv(CsrView<double> A, const double *__restrict__ x,
                                                 double alpha, double beta, double *y) {
    __shared__ double lds[kWaves][kTileNnz];
    PlainEpi epi{alpha, beta, y};
    Partials part;
    part.zero();
    tile_spmv(A, x, epi, part, lds[threadIdx.x / kWave]);
}

__global__ void k_gather(int64_t n, const double *__restrict__ src, const int *__restrict__ perm,
                         double *__restrict__ dst) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[perm[i]];
}

__global__ void k_to_f32(int64_t n, const double *__restrict__ src, float *__restrict__ dst) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = static_cast<float>(src[i]);
}

// Deterministic pseudo-random start vector for the power iteration.
__global__ void k_seed_vector(int64_t n, double *__restrict__ v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t z = static_cast<uint64_t>(i) * 0x9e3779b97f4a7c15ULL + 0x632be59bd9b4e019ULL;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        z ^= z >> 31;
        v[i] = static_cast<double>(z >> 11) * 0x1.0p-53 + 0.5;
    }
}

int elementwise_grid(int64_t n) {
    const int64_t blocks = (n + kBlock - 1) / kBlock;
    return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(blocks, 8192)));
}

int reduce(bool dot, int64_t n, const double *x, const double *y, double *result, hipStream_t s) {
    if (n <= 0) {
        *result = 0.0;
        return VK_OK;
    }
    const int blocks = static_cast<int>(std::min<int64_t>((n + kBlock * 4 - 1) / (kBlock * 4), kReduceBlocks));
    if (dot) {
        k_reduce_stage1<true><<<blocks, kBlock, 0, s>>>(n, x, y, g_reduce);
        k_reduce_stage2<true><<<1, kBlock, 0, s>>>(blocks, g_reduce, g_reduce + kReduceBlocks);
    } else {
        k_reduce_stage1<false><<<blocks, kBlock, 0, s>>>(n, x, y, g_reduce);
        k_reduce_stage2<false><<<1, kBlock, 0, s>>>(blocks, g_reduce, g_reduce + kReduceBlocks);
    }
    VK_HIP_CHECK(hipGetLastError());
    VK_HIP_CHECK(hipMemcpyAsync(result, g_reduce + kReduceBlocks, sizeof(double),
                                hipMemcpyDeviceToHost, s));
    VK_HIP_CHECK(hipStreamSynchronize(s));
    return VK_OK;
}

template <class V>
int upload_csr(DeviceCsr<V> &d, const std::vector<int> &rp, const std::vector<int> &ci,
               hipStream_t s) {
    const auto tiles = build_tiles(rp, static_cast<int>(rp.size()) - 1);
    d.rows = static_cast<int>(rp.size()) - 1;
    d.ntiles = static_cast<int>(tiles.size()) - 1;
    VK_HIP_CHECK(device_malloc(&d.row_ptr, sizeof(int) * rp.size()));
    VK_HIP_CHECK(device_malloc(&d.col, sizeof(int) * std::max<size_t>(1, ci.size())));
    VK_HIP_CHECK(device_malloc(&d.val, sizeof(V) * std::max<size_t>(1, ci.size())));
    VK_HIP_CHECK(device_malloc(&d.tiles, sizeof(int2) * tiles.size()));
    VK_HIP_CHECK(hipMemcpyAsync(d.row_ptr, rp.data(), sizeof(int) * rp.size(), hipMemcpyHostToDevice, s));
    if (!ci.empty())
        VK_HIP_CHECK(hipMemcpyAsync(d.col, ci.data(), sizeof(int) * ci.size(), hipMemcpyHostToDevice, s));
    VK_HIP_CHECK(hipMemsetAsync(d.val, 0, sizeof(V) * std::max<size_t>(1, ci.size()), s));
    VK_HIP_CHECK(hipMemcpyAsync(d.tiles, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice, s));
    VK_HIP_CHECK(hipStreamSynchronize(s));
    return VK_OK;
}

template <class V>
void free_csr(DeviceCsr<V> &d) {
    (void)hipFree(d.row_ptr);
    (void)hipFree(d.col);
    (void)hipFree(d.val);
    (void)hipFree(d.tiles);
    d = DeviceCsr<V>{};
}

// A's stored transpose, built on the device from A's device pattern
// (columns list their rows in increasing order); the tiles are cut on the
// host from the returned offsets.
int build_transpose(vk_hip_csr *A, hipStream_t s) {
    DeviceCsr<double> &t = A->at;
    const int cols = static_cast<int>(A->cols), nnz = static_cast<int>(A->nnz);
    const size_t slots = static_cast<size_t>(std::max(1, nnz));
    VK_HIP_CHECK(device_malloc(&t.row_ptr, sizeof(int) * (static_cast<size_t>(cols) + 1)));
    VK_HIP_CHECK(device_malloc(&t.col, sizeof(int) * slots));
    VK_HIP_CHECK(device_malloc(&t.val, sizeof(double) * slots));
    VK_HIP_CHECK(device_malloc(&A->tperm, sizeof(int) * slots));
    std::vector<int> trp;
    if (int rc = transpose_on_device(static_cast<int>(A->rows), cols, nnz, A->a.row_ptr, A->a.col, t.row_ptr, t.col,
                                     A->tperm, trp, s))
        return rc;
    const auto tiles = build_tiles(trp, cols);
    t.rows = cols;
    t.ntiles = static_cast<int>(tiles.size()) - 1;
    VK_HIP_CHECK(device_malloc(&t.tiles, sizeof(int2) * tiles.size()));
    VK_HIP_CHECK(hipMemcpyAsync(t.tiles, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice, s));
    VK_HIP_CHECK(hipMemsetAsync(t.val, 0, sizeof(double) * slots, s));
    VK_HIP_CHECK(hipStreamSynchronize(s));
    return VK_OK;
}

int spmv_grid(int ntiles) { return std::max(1, std::min((ntiles + kWaves - 1) / kWaves, 4096)); }

}  // namespace

int csr_make_f32(vk_hip_csr *A, hipStream_t s) {
    if (A->a32.val) return VK_OK;
    const int64_t nnz = std::max<int64_t>(1, A->nnz);
    auto shallow = [&](const DeviceCsr<double> &src, DeviceCsr<float> &dst) -> int {
        dst.row_ptr = src.row_ptr;
        dst.col = src.col;
        dst.tiles = src.tiles;
        dst.rows = src.rows;
        dst.ntiles = src.ntiles;
        VK_HIP_CHECK(device_malloc(&dst.val, sizeof(float) * nnz));
        k_to_f32<<<elementwise_grid(nnz), kBlock, 0, s>>>(nnz, src.val, dst.val);
        VK_HIP_CHECK(hipGetLastError());
        return VK_OK;
    };
    int rc = shallow(A->a, A->a32);
    if (!rc) rc = shallow(A->at, A->at32);
    if (!rc && hipStreamSynchronize(s) != hipSuccess) {
        set_error("csr_make_f32: conversion failed");
        rc = VK_EDEVICE;
    }
    if (rc) {
        // Both copies or neither: a later call starts over.
        (void)hipStreamSynchronize(s);
        (void)hipFree(A->a32.val);
        (void)hipFree(A->at32.val);
        A->a32 = DeviceCsr<float>{};
        A->at32 = DeviceCsr<float>{};
    }
    return rc;
}

}  // namespace vk

using namespace vk;

// ======================================================================= C ABI

extern "C" {

int vk_hip_init(int device) {
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        if (g_ready && device != g_device) {
            set_error("vk_hip_init: library already initialised on another device");
            return VK_ESTATE;
        }
        g_device = device;
    }
    return ensure_init();
}

int vk_hip_shutdown(void) {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_ready) return VK_OK;
    (void)hipFree(g_reduce);
    g_reduce = nullptr;
    (void)hipStreamDestroy(g_stream);
    g_stream = nullptr;
    g_ready = false;
    return VK_OK;
}

const char *vk_hip_last_error(void) { return g_error.c_str(); }

void *vk_hip_stream(void) {
    if (ensure_init() != VK_OK) return nullptr;
    return static_cast<void *>(g_stream);
}

int vk_hip_alloc(void **ptr, size_t bytes) {
    VK_REQUIRE(ptr, "vk_hip_alloc: null output");
    if (int rc = ensure_init()) return rc;
    *ptr = nullptr;
    const hipError_t e = device_malloc(ptr, std::max<size_t>(bytes, 8));
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        return VK_ENOMEM;
    }
    return VK_OK;
}

int vk_hip_free(void *ptr) {
    if (!ptr) return VK_OK;
    VK_HIP_CHECK(hipFree(ptr));
    return VK_OK;
}

int vk_hip_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream) {
    if (bytes == 0) return VK_OK;
    VK_REQUIRE(dst && src, "vk_hip_memcpy_h2d: null pointer");
    if (int rc = ensure_init()) return rc;
    hipStream_t s = as_stream(stream);
    VK_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    VK_HIP_CHECK(hipStreamSynchronize(s));
    return VK_OK;
}

int vk_hip_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream) {
    if (bytes == 0) return VK_OK;
    VK_REQUIRE(dst && src, "vk_hip_memcpy_d2h: null pointer");
    if (int rc = ensure_init()) return rc;
    hipStream_t s = as_stream(stream);
    VK_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    VK_HIP_CHECK(hipStreamSynchronize(s));
    return VK_OK;
}

int vk_hip_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream) {
    if (bytes == 0) return VK_OK;
    VK_REQUIRE(dst && src, "vk_hip_memcpy_d2d: null pointer");
    if (int rc = ensure_init()) return rc;
    VK_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
    return VK_OK;
}

int vk_hip_sync(void *stream) {
    if (int rc = ensure_init()) return rc;
    VK_HIP_CHECK(hipStreamSynchronize(as_stream(stream)));
    return VK_OK;
}

int vk_hip_fill(double *y, int64_t n, double value, void *stream) {
    VK_REQUIRE(n >= 0, "vk_hip_fill: negative size");
    if (n == 0) return VK_OK;
    VK_REQUIRE(y, "vk_hip_fill: null vector");
    if (int rc = ensure_init()) return rc;
    k_fill<<<elementwise_grid(n), kBlock, 0, as_stream(stream)>>>(y, n, value);
    VK_HIP_CHECK(hipGetLastError());
    return VK_OK;
}

int vk_hip_axpby(int64_t n, double alpha, const double *x, double beta, double *y, void *stream) {
    VK_REQUIRE(n >= 0, "vk_hip_axpby: negative size");
    if (n == 0) return VK_OK;
    VK_REQUIRE(x && y, "vk_hip_axpby: null vector");
    if (int rc = ensure_init()) return rc;
    k_axpby<<<elementwise_grid(n), kBlock, 0, as_stream(stream)>>>(n, alpha, x, beta, y);
    VK_HIP_CHECK(hipGetLastError());
    return VK_OK;
}

int vk_hip_project_box(int64_t n, double *x, const double *lo, const double *hi, void *stream) {
    VK_REQUIRE(n >= 0, "vk_hip_project_box: negative size");
    if (n == 0) return VK_OK;
    VK_REQUIRE(x && lo && hi, "vk_hip_project_box: null vector");
    if (int rc = ensure_init()) return rc;
    k_project<<<elementwise_grid(n), kBlock, 0, as_stream(stream)>>>(n, x, lo, hi);
    VK_HIP_CHECK(hipGetLastError());
    return VK_OK;
}

int vk_hip_reduce_absmax(int64_t n, const double *x, double *result, void *stream) {
    VK_REQUIRE(n >= 0 && result, "vk_hip_reduce_absmax: bad arguments");
    VK_REQUIRE(n == 0 || x, "vk_hip_reduce_absmax: null vector");
    if (int rc = ensure_init()) return rc;
    std::lock_guard<std::mutex> lock(g_mutex);
    return reduce(false, n, x, nullptr, result, as_stream(stream));
}

int vk_hip_reduce_dot(int64_t n, const double *x, const double *y, double *result, void *stream) {
    VK_REQUIRE(n >= 0 && result, "vk_hip_reduce_dot: bad arguments");
    VK_REQUIRE(n == 0 || (x && y), "vk_hip_reduce_dot: null vector");
    if (int rc = ensure_init()) return rc;
    std::lock_guard<std::mutex> lock(g_mutex);
    return reduce(true, n, x, y, result, as_stream(stream));
}

int vk_hip_reduce_dot_noalloc(int64_t n, const double *x, const double *y, double *result, void *stream) {
    VK_REQUIRE(n >= 0 && result, "vk_hip_reduce_dot_noalloc: bad arguments");
    VK_REQUIRE(n == 0 || (x && y), "vk_hip_reduce_dot_noalloc: null vector");
    if (n == 0) {
        *result = 0.0;
        return VK_OK;
    }
    int device = 0;
    VK_HIP_CHECK(hipGetDevice(&device));
    VK_REQUIRE(device >= 0 && device < kMaxNoallocDevices, "vk_hip_reduce_dot_noalloc: device index out of range");
    // One lock over set-up, launches and the result copy: calls from several
    // host threads take turns on the one scratch array per GPU.
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        // First call on this device (current on the calling thread): the
        // device address of this GPU's copy of the array.
        void *ptr = nullptr;
        VK_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(noalloc_scratch)));
        VK_REQUIRE(ptr != nullptr, "vk_hip_reduce_dot_noalloc: no device address for the scratch array");
        g_scratch_ptr[device] = static_cast<double *>(ptr);
    }
    double *part = g_scratch_ptr[device];
    hipStream_t s = static_cast<hipStream_t>(stream);   // NULL: the device's null stream
    // The same grid and kernels as vk_hip_reduce_dot: stage 1 writes every
    // partial it launches and stage 2 the result slot, so nothing in the
    // array needs clearing between calls (the blocks are capped at the
    // array's kReduceBlocks partials, whatever n).
    const int blocks = static_cast<int>(std::min<int64_t>((n + kBlock * 4 - 1) / (kBlock * 4), kReduceBlocks));
    k_reduce_stage1<true><<<blocks, kBlock, 0, s>>>(n, x, y, part);
    k_reduce_stage2<true><<<1, kBlock, 0, s>>>(blocks, part, part + kReduceBlocks);
    VK_HIP_CHECK(hipGetLastError());
    VK_HIP_CHECK(hipMemcpyAsync(result, part + kReduceBlocks, sizeof(double), hipMemcpyDeviceToHost, s));
    VK_HIP_CHECK(hipStreamSynchronize(s));
    return VK_OK;
}

int vk_hip_csr_create(int64_t rows, int64_t cols, const int64_t *row_offsets,
                      const int64_t *column_indices, vk_hip_csr **out) {
    VK_REQUIRE(out, "vk_hip_csr_create: null output");
    *out = nullptr;
    // cpu_ops.cpp:38-63 validate_pattern, plus the int32 device range.
    VK_REQUIRE(rows >= 0 && cols >= 0, "CsrPattern dimensions must be non-negative");
    VK_REQUIRE(row_offsets, "CsrPattern row_offsets missing");
    VK_REQUIRE(rows < INT32_MAX && cols < INT32_MAX, "matrix dimensions exceed the int32 device range");
    const int64_t nnz = row_offsets[rows];
    VK_REQUIRE(row_offsets[0] == 0 && nnz >= 0, "CsrPattern row_offsets do not describe column_indices");
    VK_REQUIRE(nnz < INT32_MAX, "nonzero count exceeds the int32 device range");
    VK_REQUIRE(nnz == 0 || column_indices, "CsrPattern column_indices missing");
    // Validation and the int32 copies in one pass over the pattern.
    std::vector<int> rp(static_cast<size_t>(rows + 1)), ci(static_cast<size_t>(nnz));
    rp[0] = 0;
    for (int64_t r = 0; r < rows; ++r) {
        const int64_t b = row_offsets[r], e = row_offsets[r + 1];
        VK_REQUIRE(b >= 0 && b <= e && e <= nnz, "CsrPattern row_offsets must be ordered and in range");
        rp[r + 1] = static_cast<int>(e);
        int64_t prev = -1;
        for (int64_t k = b; k < e; ++k) {
            const int64_t c = column_indices[k];
            VK_REQUIRE(c >= 0 && c < cols && c > prev, "CsrPattern columns must be ordered and in range");
            ci[k] = static_cast<int>(c);
            prev = c;
        }
    }
    if (int rc = ensure_init()) return rc;
    hipStream_t s = default_stream();
    auto *A = new vk_hip_csr();
    A->rows = rows;
    A->cols = cols;
    A->nnz = nnz;
    int rc = upload_csr(A->a, rp, ci, s);
    if (!rc) rc = build_transpose(A, s);
    if (rc) {
        vk_hip_csr_destroy(A);
        return rc;
    }
    *out = A;
    return VK_OK;
}

int vk_hip_csr_set_values(vk_hip_csr *A, const double *values) {
    VK_REQUIRE(A, "vk_hip_csr_set_values: null matrix");
    VK_REQUIRE(A->nnz == 0 || values, "vk_hip_csr_set_values: null values");
    if (A->nnz == 0) {
        A->has_values = true;
        return VK_OK;
    }
    hipStream_t s = default_stream();
    VK_HIP_CHECK(hipMemcpyAsync(A->a.val, values, sizeof(double) * A->nnz, hipMemcpyHostToDevice, s));
    k_gather<<<elementwise_grid(A->nnz), kBlock, 0, s>>>(A->nnz, A->a.val, A->tperm, A->at.val);
    VK_HIP_CHECK(hipGetLastError());
    if (A->a32.val) {
        k_to_f32<<<elementwise_grid(A->nnz), kBlock, 0, s>>>(A->nnz, A->a.val, A->a32.val);
        k_to_f32<<<elementwise_grid(A->nnz), kBlock, 0, s>>>(A->nnz, A->at.val, A->at32.val);
        VK_HIP_CHECK(hipGetLastError());
    }
    VK_HIP_CHECK(hipStreamSynchronize(s));
    A->has_values = true;
    A->scaled = false;
    return VK_OK;
}

int vk_hip_debug_fail_allocations(int64_t after, int64_t *made) {
    if (made) *made = g_allocs_made.exchange(0);
    g_allocs_left.store(after < 0 ? -1 : after);
    return VK_OK;
}

int vk_hip_csr_set_tile_geometry(vk_hip_csr *A, int geometry) {
    VK_REQUIRE(A, "vk_hip_csr_set_tile_geometry: null matrix");
    VK_REQUIRE(geometry >= 0 && geometry < kTileGeometries, "vk_hip_csr_set_tile_geometry: unknown geometry");
    if (A->tile_geometry == geometry) return VK_OK;
    hipStream_t s = default_stream();
    const TileGeometry &g = kTileGeometry[geometry];
    // Both operands' offsets, cut on the host, then the new tables replace
    // the old ones (the FP32 copies share them).
    std::vector<int> rp(static_cast<size_t>(A->rows) + 1), trp(static_cast<size_t>(A->cols) + 1);
    VK_HIP_CHECK(hipMemcpyAsync(rp.data(), A->a.row_ptr, sizeof(int) * rp.size(), hipMemcpyDeviceToHost, s));
    VK_HIP_CHECK(hipMemcpyAsync(trp.data(), A->at.row_ptr, sizeof(int) * trp.size(), hipMemcpyDeviceToHost, s));
    VK_HIP_CHECK(hipStreamSynchronize(s));
    const auto ta = build_tiles(rp, static_cast<int>(A->rows), g);
    const auto tt = build_tiles(trp, static_cast<int>(A->cols), g);
    int2 *da = nullptr, *dt = nullptr;
    if (device_malloc(&da, sizeof(int2) * ta.size()) != hipSuccess || device_malloc(&dt, sizeof(int2) * tt.size()) != hipSuccess) {
        if (da) (void)hipFree(da);
        set_error("vk_hip_csr_set_tile_geometry: device allocation failed");
        return VK_ENOMEM;
    }
    if (hipMemcpyAsync(da, ta.data(), sizeof(int2) * ta.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(dt, tt.data(), sizeof(int2) * tt.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        (void)hipFree(da);
        (void)hipFree(dt);
        set_error("vk_hip_csr_set_tile_geometry: upload failed");
        return VK_EDEVICE;
    }
    (void)hipFree(A->a.tiles);
    (void)hipFree(A->at.tiles);
    A->a.tiles = da;
    A->a.ntiles = static_cast<int>(ta.size()) - 1;
    A->at.tiles = dt;
    A->at.ntiles = static_cast<int>(tt.size()) - 1;
    if (A->a32.val) {
        A->a32.tiles = A->a.tiles;
        A->a32.ntiles = A->a.ntiles;
        A->at32.tiles = A->at.tiles;
        A->at32.ntiles = A->at.ntiles;
    }
    A->tile_geometry = geometry;
    return VK_OK;
}

int vk_hip_csr_tile_geometry(const vk_hip_csr *A, int *geometry) {
    VK_REQUIRE(A && geometry, "vk_hip_csr_tile_geometry: null argument");
    *geometry = A->tile_geometry;
    return VK_OK;
}

int vk_hip_csr_destroy(vk_hip_csr *A) {
    if (!A) return VK_OK;
    (void)hipFree(A->a32.val);
    (void)hipFree(A->at32.val);
    free_csr(A->a);
    free_csr(A->at);
    (void)hipFree(A->tperm);
    (void)hipFree(A->orig);
    (void)hipFree(A->col_factors);
    delete A;
    return VK_OK;
}

int vk_hip_csr_dims(const vk_hip_csr *A, int64_t *rows, int64_t *cols, int64_t *nnz) {
    VK_REQUIRE(A, "vk_hip_csr_dims: null matrix");
    if (rows) *rows = A->rows;
    if (cols) *cols = A->cols;
    if (nnz) *nnz = A->nnz;
    return VK_OK;
}

int vk_hip_csr_transpose_export(const vk_hip_csr *A, int32_t *t_offsets, int32_t *t_rows,
                                double *t_values) {
    VK_REQUIRE(A, "vk_hip_csr_transpose_export: null matrix");
    hipStream_t s = default_stream();
    if (t_offsets)
        VK_HIP_CHECK(hipMemcpyAsync(t_offsets, A->at.row_ptr, sizeof(int) * (A->cols + 1), hipMemcpyDeviceToHost, s));
    if (t_rows && A->nnz)
        VK_HIP_CHECK(hipMemcpyAsync(t_rows, A->at.col, sizeof(int) * A->nnz, hipMemcpyDeviceToHost, s));
    if (t_values && A->nnz)
        VK_HIP_CHECK(hipMemcpyAsync(t_values, A->at.val, sizeof(double) * A->nnz, hipMemcpyDeviceToHost, s));
    VK_HIP_CHECK(hipStreamSynchronize(s));
    return VK_OK;
}

int vk_hip_spmv_csr(const vk_hip_csr *A, int transpose, double alpha, const double *x,
                    double beta, double *y, void *stream) {
    VK_REQUIRE(A, "vk_hip_spmv_csr: null matrix");
    const int64_t ny = transpose ? A->cols : A->rows;
    const int64_t nx = transpose ? A->rows : A->cols;
    VK_REQUIRE(ny == 0 || y, "vk_hip_spmv_csr: null y");
    VK_REQUIRE(nx == 0 || x || A->nnz == 0, "vk_hip_spmv_csr: null x");
    if (ny == 0) return VK_OK;
    hipStream_t s = as_stream(stream);
    const DeviceCsr<double> &M = transpose ? A->at : A->a;
    if (M.ntiles == 0) return VK_OK;
    k_spmv<<<spmv_grid(M.ntiles), kBlock, 0, s>>>(M.view(), x, alpha, beta, y);
    VK_HIP_CHECK(hipGetLastError());
    return VK_OK;
}

int vk_hip_csr_norm2_estimate(const vk_hip_csr *A, int iterations, double *result) {
    VK_REQUIRE(A && result && iterations > 0, "vk_hip_csr_norm2_estimate: bad arguments");
    if (A->nnz == 0 || A->rows == 0 || A->cols == 0) {
        *result = 0.0;
        return VK_OK;
    }
    hipStream_t s = default_stream();
    double *v = nullptr, *w = nullptr;
    if (device_malloc(&v, sizeof(double) * A->cols) != hipSuccess ||
        device_malloc(&w, sizeof(double) * A->rows) != hipSuccess) {
        (void)hipFree(v);
        set_error("vk_hip_csr_norm2_estimate: device allocation failed");
        return VK_ENOMEM;
    }
    // k_spmv forms alpha A x + beta y literally (cpu_ops.cpp:426), so with
    // beta = 0 a NaN or Inf left in fresh memory would still reach the sum.
    VK_HIP_CHECK(hipMemsetAsync(w, 0, sizeof(double) * A->rows, s));
    k_seed_vector<<<elementwise_grid(A->cols), kBlock, 0, s>>>(A->cols, v);
    double norm2 = 0.0, nv = 0.0;
    int rc = VK_OK;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        rc = reduce(true, A->cols, v, v, &nv, s);
        for (int it = 0; rc == VK_OK && it < iterations; ++it) {
            k_axpby<<<elementwise_grid(A->cols), kBlock, 0, s>>>(A->cols, 1.0 / std::sqrt(nv), v, 0.0, v);
            k_spmv<<<spmv_grid(A->a.ntiles), kBlock, 0, s>>>(A->a.view(), v, 1.0, 0.0, w);
            k_spmv<<<spmv_grid(A->at.ntiles), kBlock, 0, s>>>(A->at.view(), w, 1.0, 0.0, v);
            rc = reduce(true, A->cols, v, v, &nv, s);
            norm2 = std::sqrt(std::sqrt(nv));  // ||A'A v|| with ||v|| = 1 -> sigma_max^2
        }
    }
    (void)hipFree(v);
    (void)hipFree(w);
    if (rc) return rc;
    *result = norm2;
    return VK_OK;
}

}  // extern "C"
