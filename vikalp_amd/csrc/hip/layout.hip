// Device construction of the matrix layouts: A's stored transpose and the
// column-blocked copy of A.  Both used to be host passes over every nonzero
// (a counting sort, a blocked plan) with the pattern copied both ways.
// Here the pattern stays on the device: the transpose is a stable radix sort
// of the nonzeros by column (rocPRIM), so each column lists its rows in
// increasing order exactly as a host counting sort does, and the blocked
// plan is a per-row split plus one scan.  Only the row offsets come back to
// the host, where the tiles (a greedy O(rows) pass) are cut.

#include "runtime.hpp"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <vector>

namespace vk {
namespace {

constexpr int kLayoutBlock = 256;

int lgrid(int64_t n) {
    return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>((n + kLayoutBlock - 1) / kLayoutBlock, 16384)));
}

// row[k] = the row of nonzero k; key index[k] = k.
__global__ void k_rows_and_index(int rows, const int *__restrict__ rp, int *__restrict__ row_of,
                                 int *__restrict__ index) {
    for (int r = blockIdx.x * kLayoutBlock + threadIdx.x; r < rows; r += gridDim.x * kLayoutBlock)
        for (int k = rp[r]; k < rp[r + 1]; ++k) {
            row_of[k] = r;
            index[k] = k;
        }
}

// trow[p] = row of the nonzero placed at p.
__global__ void k_take_int(int n, const int *__restrict__ src, const int *__restrict__ perm, int *__restrict__ dst) {
    for (int i = blockIdx.x * kLayoutBlock + threadIdx.x; i < n; i += gridDim.x * kLayoutBlock) dst[i] = src[perm[i]];
}

// trp[j] = first position whose sorted column is >= j (j = 0..cols).
__global__ void k_column_starts(int cols, int nnz, const unsigned *__restrict__ sorted, int *__restrict__ trp) {
    for (int j = blockIdx.x * kLayoutBlock + threadIdx.x; j <= cols; j += gridDim.x * kLayoutBlock) {
        int lo = 0, hi = nnz;
        while (lo < hi) {
            const int mid = lo + (hi - lo) / 2;
            if (sorted[mid] < static_cast<unsigned>(j)) lo = mid + 1;
            else hi = mid;
        }
        trp[j] = lo;
    }
}

struct Bounds {
    int b[kMaxColBlocks + 1];   // block p spans columns [b[p], b[p+1])
};

// count[p * rows + r] = entries of row r in block p (columns increase
// along a row, so the block index only grows).
__global__ void k_block_counts(int rows, int nb, Bounds B, const int *__restrict__ rp,
                               const int *__restrict__ ci, int *__restrict__ count) {
    for (int r = blockIdx.x * kLayoutBlock + threadIdx.x; r < rows; r += gridDim.x * kLayoutBlock) {
        int p = 0, first = rp[r];
        const int e = rp[r + 1];
        for (int k = rp[r]; k < e; ++k)
            while (ci[k] >= B.b[p + 1]) {
                count[static_cast<int64_t>(p) * rows + r] = k - first;
                first = k;
                ++p;
            }
        count[static_cast<int64_t>(p) * rows + r] = e - first;
        for (++p; p < nb; ++p) count[static_cast<int64_t>(p) * rows + r] = 0;
    }
}

// Blocked offsets from the scan of the counts: block p's offsets, then its
// end, with stride rows + 1.
__global__ void k_block_offsets(int rows, int nb, int nnz, const int *__restrict__ start, int *__restrict__ row_ptr) {
    const int64_t total = static_cast<int64_t>(nb) * (rows + 1);
    for (int64_t i = blockIdx.x * (int64_t)kLayoutBlock + threadIdx.x; i < total; i += (int64_t)gridDim.x * kLayoutBlock) {
        const int p = static_cast<int>(i / (rows + 1)), r = static_cast<int>(i % (rows + 1));
        row_ptr[i] = r < rows ? start[static_cast<int64_t>(p) * rows + r]
                              : (p + 1 < nb ? start[static_cast<int64_t>(p + 1) * rows] : nnz);
    }
}

// perm[pos] = CSR position and col[pos] = its column, for every nonzero
// placed in its block's row segment in row order.
__global__ void k_block_place(int rows, Bounds B, const int *__restrict__ rp, const int *__restrict__ ci,
                              const int *__restrict__ row_ptr, int *__restrict__ perm, int *__restrict__ col) {
    for (int r = blockIdx.x * kLayoutBlock + threadIdx.x; r < rows; r += gridDim.x * kLayoutBlock) {
        int p = 0;
        int pos = row_ptr[r];
        const int e = rp[r + 1];
        for (int k = rp[r]; k < e; ++k) {
            while (ci[k] >= B.b[p + 1]) {
                ++p;
                pos = row_ptr[static_cast<int64_t>(p) * (rows + 1) + r];
            }
            perm[pos] = k;
            col[pos] = ci[k];
            ++pos;
        }
    }
}

struct Buffers {
    std::vector<void *> blocks;
    ~Buffers() {
        for (void *p : blocks) (void)hipFree(p);
    }
    void *get(size_t bytes) {
        void *p = nullptr;
        if (device_malloc(&p, std::max<size_t>(bytes, 4)) != hipSuccess) return nullptr;
        blocks.push_back(p);
        return p;
    }
};

int bit_width(unsigned v) {
    int b = 0;
    while (v) {
        ++b;
        v >>= 1;
    }
    return std::max(b, 1);
}

}  // namespace

int transpose_on_device(int rows, int cols, int nnz, const int *rp, const int *ci, int *trp, int *trow,
                        int *tperm, std::vector<int> &host_trp, hipStream_t s) {
    host_trp.assign(static_cast<size_t>(cols) + 1, 0);
    if (nnz == 0) {
        VK_HIP_CHECK(hipMemcpyAsync(trp, host_trp.data(), sizeof(int) * host_trp.size(), hipMemcpyHostToDevice, s));
        VK_HIP_CHECK(hipStreamSynchronize(s));
        return VK_OK;
    }
    Buffers bufs;
    int *row_of = static_cast<int *>(bufs.get(sizeof(int) * nnz));
    int *index = static_cast<int *>(bufs.get(sizeof(int) * nnz));
    auto *sorted = static_cast<unsigned *>(bufs.get(sizeof(unsigned) * nnz));
    // Columns are non-negative: sort their bit patterns as unsigned keys.
    const auto *keys = reinterpret_cast<const unsigned *>(ci);
    if (!row_of || !index || !sorted) {
        set_error("vk_hip_csr_create: transpose scratch allocation failed");
        return VK_ENOMEM;
    }
    k_rows_and_index<<<lgrid(rows), kLayoutBlock, 0, s>>>(rows, rp, row_of, index);
    VK_HIP_CHECK(hipGetLastError());
    const int end_bit = bit_width(static_cast<unsigned>(std::max(cols - 1, 0)));
    size_t temp_bytes = 0;
    VK_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, temp_bytes, keys, sorted, index, tperm,
                                           static_cast<unsigned>(nnz), 0, end_bit, s));
    void *temp = bufs.get(temp_bytes);
    if (!temp) {
        set_error("vk_hip_csr_create: sort scratch allocation failed");
        return VK_ENOMEM;
    }
    VK_HIP_CHECK(rocprim::radix_sort_pairs(temp, temp_bytes, keys, sorted, index, tperm,
                                           static_cast<unsigned>(nnz), 0, end_bit, s));
    k_take_int<<<lgrid(nnz), kLayoutBlock, 0, s>>>(nnz, row_of, tperm, trow);
    k_column_starts<<<lgrid(static_cast<int64_t>(cols) + 1), kLayoutBlock, 0, s>>>(cols, nnz, sorted, trp);
    VK_HIP_CHECK(hipGetLastError());
    VK_HIP_CHECK(hipMemcpyAsync(host_trp.data(), trp, sizeof(int) * host_trp.size(), hipMemcpyDeviceToHost, s));
    VK_HIP_CHECK(hipStreamSynchronize(s));
    return VK_OK;
}

int plan_blocks_on_device(const int *rp, const int *ci, int rows, int cols, int nnz, int nb, int *row_ptr,
                          int *col, int *perm, std::vector<int> &host_row_ptr, hipStream_t s) {
    Bounds B{};
    for (int p = 0; p <= nb; ++p) B.b[p] = static_cast<int>(static_cast<long long>(cols) * p / nb);
    const size_t stride = static_cast<size_t>(rows) + 1;
    host_row_ptr.assign(stride * nb, 0);
    if (rows == 0) {
        VK_HIP_CHECK(hipMemcpyAsync(row_ptr, host_row_ptr.data(), sizeof(int) * host_row_ptr.size(),
                                    hipMemcpyHostToDevice, s));
        VK_HIP_CHECK(hipStreamSynchronize(s));
        return VK_OK;
    }
    const size_t cells = static_cast<size_t>(rows) * nb;
    Buffers bufs;
    int *count = static_cast<int *>(bufs.get(sizeof(int) * cells));
    int *start = static_cast<int *>(bufs.get(sizeof(int) * cells));
    if (!count || !start) {
        set_error("vk_hip_pdhg_create: blocked layout scratch allocation failed");
        return VK_ENOMEM;
    }
    k_block_counts<<<lgrid(rows), kLayoutBlock, 0, s>>>(rows, nb, B, rp, ci, count);
    VK_HIP_CHECK(hipGetLastError());
    size_t temp_bytes = 0;
    VK_HIP_CHECK(rocprim::exclusive_scan(nullptr, temp_bytes, count, start, 0, cells, rocprim::plus<int>(), s));
    void *temp = bufs.get(temp_bytes);
    if (!temp) {
        set_error("vk_hip_pdhg_create: blocked layout scan scratch allocation failed");
        return VK_ENOMEM;
    }
    VK_HIP_CHECK(rocprim::exclusive_scan(temp, temp_bytes, count, start, 0, cells, rocprim::plus<int>(), s));
    k_block_offsets<<<lgrid(static_cast<int64_t>(stride) * nb), kLayoutBlock, 0, s>>>(rows, nb, nnz, start, row_ptr);
    k_block_place<<<lgrid(rows), kLayoutBlock, 0, s>>>(rows, B, rp, ci, row_ptr, perm, col);
    VK_HIP_CHECK(hipGetLastError());
    VK_HIP_CHECK(hipMemcpyAsync(host_row_ptr.data(), row_ptr, sizeof(int) * host_row_ptr.size(),
                                hipMemcpyDeviceToHost, s));
    VK_HIP_CHECK(hipStreamSynchronize(s));
    return VK_OK;
}

}  // namespace vk
