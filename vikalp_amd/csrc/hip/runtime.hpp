// Host-side runtime pieces shared by the C-ABI translation units.
#pragma once

#include "vikalp_hip.h"
#include "common.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace vk {

void set_error(const std::string &message);
// Makes the library's device current on the calling host thread (once per
// thread); for entry points that may run on threads of the caller's pool.
int bind_device();
hipStream_t default_stream();
int ensure_init();

// Every device allocation of the library goes through device_malloc, so a
// test can make them fail on demand (vk_hip_debug_fail_allocations).
hipError_t device_alloc_bytes(void **p, std::size_t bytes);
template <class T>
hipError_t device_malloc(T **p, std::size_t bytes) {
    return device_alloc_bytes(reinterpret_cast<void **>(p), bytes);
}

inline hipStream_t as_stream(void *s) {
    return s ? static_cast<hipStream_t>(s) : default_stream();
}

// Device-resident tiling of one CSR operand.
template <class V>
struct DeviceCsr {
    int *row_ptr = nullptr;
    int *col = nullptr;
    V *val = nullptr;
    int2 *tiles = nullptr;   // {first row, first nonzero}, ntiles+1 entries
    int rows = 0;
    int ntiles = 0;
    // All tiles streamed (non-temporal); a solver that budgets the MALL sets
    // `resident` on its own copy of the view.
    CsrView<V> view() const { return CsrView<V>{row_ptr, col, val, tiles, rows, ntiles, 0}; }
};

// How rows are cut into tiles: whole rows, at most max_nnz nonzeros and
// max_rows rows per tile (max_nnz <= kTileNnz, max_rows <= kTileRows, what
// the tile kernels hold), and a row with more than long_row nonzeros in a
// tile of its own, reduced by the whole wave (long_row <= max_nnz).  The
// geometry sets which rows share a tile and so the order in which a row's
// products are added: results depend on it (bitwise), correctness does not.
struct TileGeometry {
    int max_nnz = kTileNnz;
    int max_rows = kTileRows;
    int long_row = kTileNnz;
};

// The geometries a caller or the tuner may pick (index 0 is the default):
// 512 nonzeros x 64 rows; 256 x 64; 512 x 32; 512 x 64 with rows longer
// than 64 nonzeros alone.
constexpr int kTileGeometries = 4;
constexpr TileGeometry kTileGeometry[kTileGeometries] = {
    {kTileNnz, kTileRows, kTileNnz}, {256, kTileRows, 256}, {kTileNnz, 32, kTileNnz}, {kTileNnz, kTileRows, 64}};

// Splits rows into tiles of whole rows by geometry g.
std::vector<int2> build_tiles(const std::vector<int> &row_ptr, int rows, const TileGeometry &g = TileGeometry{});

// Appends the tiles of rows [begin, end) of the CSR offsets `rp` to `tiles`
// as [first, boundaries..., end] ({row, rp[row]} each): tile t spans entries
// t and t+1, so the range's last entry is its end sentinel.
void append_tiles(const int *rp, int begin, int end, std::vector<int2> &tiles, const TileGeometry &g = TileGeometry{});

}  // namespace vk

// The opaque matrix handle of the C ABI.
struct vk_hip_csr {
    std::int64_t rows = 0, cols = 0, nnz = 0;
    vk::DeviceCsr<double> a, at;      // A and its stored transpose (FP64)
    vk::DeviceCsr<float> a32, at32;   // FP32 copies (made on demand)
    int *tperm = nullptr;             // at.val[k] = a.val[tperm[k]]
    bool has_values = false;
    // Once vk_hip_lp_scale has replaced the values with scaled ones, `orig`
    // holds the values last given to vk_hip_csr_set_values.
    double *orig = nullptr;
    bool scaled = false;
    // The column factors of the last scaling, pass by pass (pass 0: the
    // reference step's s, divided into l and u; passes 1..: equilibration
    // factors, multiplied into them): col_passes rows of cols doubles.
    double *col_factors = nullptr;
    int col_passes = 0;
    int tile_geometry = 0;            // index into kTileGeometry of a's and at's tiles
};

#define VK_HIP_CHECK(expr)                                                      \
    do {                                                                        \
        hipError_t vk_err_ = (expr);                                            \
        if (vk_err_ != hipSuccess) {                                            \
            vk::set_error(std::string(#expr) + ": " + hipGetErrorString(vk_err_)); \
            return vk_err_ == hipErrorOutOfMemory ? VK_ENOMEM : VK_EDEVICE;     \
        }                                                                       \
    } while (0)

#define VK_REQUIRE(cond, msg)                 \
    do {                                      \
        if (!(cond)) {                        \
            vk::set_error(msg);               \
            return VK_EINVAL;                 \
        }                                     \
    } while (0)

namespace vk {
// Makes the FP32 copies of A and A' (used by the mixed-precision solver).
int csr_make_f32(vk_hip_csr *A, hipStream_t stream);

// Device layout builders (layout.hip).  A's transpose pattern from its
// device pattern (rp, ci): offsets trp (cols + 1), rows trow and the CSR
// position of each entry tperm (nnz), columns listing rows in increasing
// order; the offsets are also returned on the host.
int transpose_on_device(int rows, int cols, int nnz, const int *rp, const int *ci, int *trp, int *trow,
                        int *tperm, std::vector<int> &host_trp, hipStream_t s);
// The column-blocked pattern of A: block p takes columns
// [cols*p/nb, cols*(p+1)/nb) of every row, rows in order, blocks one after
// another.  Fills row_ptr (nb blocks of rows + 1 offsets), col and perm
// (the CSR position each blocked entry comes from) on the device; the
// offsets are also returned on the host.
int plan_blocks_on_device(const int *rp, const int *ci, int rows, int cols, int nnz, int nb, int *row_ptr,
                          int *col, int *perm, std::vector<int> &host_row_ptr, hipStream_t s);
}
