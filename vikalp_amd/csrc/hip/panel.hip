// Panel layout and LDS-staged SpMV (see panel.hpp for the design).

#include "panel.hpp"
#include "pdhg_core.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

namespace vk {

void PanelLayout::release() {
    if (meta) (void)hipFree(meta);
    if (hdr) (void)hipFree(hdr);
    if (ent) (void)hipFree(ent);
    meta = nullptr;
    hdr = nullptr;
    ent = nullptr;
    groups = panel_kb = threads = steps = fast_steps = nwg = resident = rows = cols = value_bytes = 0;
    ent_bytes = 0;
}

namespace {

// ------------------------------------------------------------------ kernel

__device__ __forceinline__ double entry_value(const panel_u32x3 &e) { return __hiloint2double(e.z, e.y); }
__device__ __forceinline__ float entry_value(const panel_u32x2 &e) { return __uint_as_float(e.y); }

// One workgroup per PanelMeta.  Per panel p: the products of panel p's
// entries (held in registers since the previous panel) gather from the
// staged panel in LDS; then panel p+1's slice of x and entries are loaded
// (16-byte and one-per-step coalesced loads), panel p's products are added
// to the row accumulators, panel p+1's slice is written to the other LDS
// buffer and a barrier ends the panel.  Entries a lane does not own (lane
// >= the step's count) are loaded from the step's last entry (an empty
// step's load reads the entry after the chunk: the next chunk's first, or
// the zero padding after the last) and never stored, so every load is
// unconditional, in bounds and of a checked entry.  KEEP: the entries
// are loaded with cache-allocating loads (they stay in the MALL from one
// iteration to the next), else streamed.
template <class V, int NT, int KM, int KB, bool KEEP>
__device__ __forceinline__ void panel_rows(const PanelView<V> &pv, const V *__restrict__ x, double *__restrict__ out,
                                           V (*xbuf)[panel_width<V>(KB)], double *acc) {
#pragma clang fp contract(off)
    using E = typename PanelEntry<V>::E;
    constexpr int P = panel_width<V>(KB);
    constexpr int kRows = panel_max_rows(KB);
    constexpr int kW = NT / kWave;
    constexpr int kVec = 16 / static_cast<int>(sizeof(V));
    constexpr int kPer = P / kVec / NT;   // 16-byte loads per thread and panel
    static_assert(kPer * kVec * NT == P, "panel width must split evenly over the lanes");
    constexpr int kHW = 2 + (KM + 3) / 4;   // header words used
    using CMeta = const __attribute__((address_space(4))) PanelMeta;
    CMeta *mp = (CMeta *)(pv.meta) + blockIdx.x;
    const int r0 = mp->r0, r1 = mp->r1, c0 = mp->c0, np = mp->np, hfirst = mp->hfirst, g = mp->g;
    const int t = threadIdx.x, lane = t % kWave;
    const int wv = __builtin_amdgcn_readfirstlane(t / kWave);
    const int nr = r1 - r0;
    for (int i = t; i < nr; i += NT) acc[i] = 0.0;

    // This wave's headers of panels 0..63 in hreg (lane p holds panel p) and
    // 64..127 in hreg2, loaded once; panel p's header is read with readlane.
    int hreg[kHW], hreg2[kHW];
    {
        const int *h1 = pv.hdr + 8 * (hfirst + min(lane, np - 1) * kW + wv);
        const int *h2 = pv.hdr + 8 * (hfirst + min(lane + kWave, np - 1) * kW + wv);
#pragma unroll
        for (int i = 0; i < kHW; ++i) {
            hreg[i] = h1[i];
            hreg2[i] = h2[i];
        }
    }
    // (A select of the register set, then one readlane: no branch.)
    auto hword = [&](int i, int p) { return __builtin_amdgcn_readlane(p < kWave ? hreg[i] : hreg2[i], p & (kWave - 1)); };
    auto step_count = [](const int (&w)[kHW], int j) { return (w[2 + j / 4] >> (8 * (j % 4))) & 0xff; };

    // Byte offsets of the 16-byte vectors fit 32 bits (panel_build checks),
    // so each load is a uniform base plus one 32-bit lane offset.
    const char *xb0 = reinterpret_cast<const char *>(x);
    const int xlast = pv.xvecs - 1;
    panel_u32x4 stg[kPer];
    E ce[KM];
    auto load_panel = [&](int p) {
        const int v0 = (c0 + p * P) / kVec;
#pragma unroll
        for (int u = 0; u < kPer; ++u)
            stg[u] = *reinterpret_cast<const panel_u32x4 *>(xb0 + (static_cast<unsigned>(min(v0 + t + u * NT, xlast)) << 4));
        int w[kHW];
#pragma unroll
        for (int i = 0; i < kHW; ++i) w[i] = hword(i, p);
        int off = w[0];
#pragma unroll
        for (int j = 0; j < KM; ++j) {
            const int c = step_count(w, j);
            // Clamped to the step's last lane: lanes beyond its count
            // repeat that entry (same cache line, nothing stored); loading
            // the entries that follow instead costs more L1 misses
            // (measured: FP64 A' y 78 -> 89 us).
            const E *ep = pv.ent + off + min(lane, max(c - 1, 0));
            if constexpr (KEEP) ce[j] = *ep;
            else ce[j] = __builtin_nontemporal_load(ep);
            off += c;
        }
    };
    auto store_panel = [&](int buf) {
        panel_u32x4 *xs = reinterpret_cast<panel_u32x4 *>(xbuf[buf]);
#pragma unroll
        for (int u = 0; u < kPer; ++u) xs[t + u * NT] = stg[u];
    };

    load_panel(0);
    store_panel(0);
    __syncthreads();
    for (int p = 0; p < np; ++p) {
        int wc[kHW];
#pragma unroll
        for (int i = 0; i < kHW; ++i) wc[i] = hword(i, p);
        const V *xb = xbuf[p & 1];
        double prod[KM];
        int rows[KM];
#pragma unroll
        for (int j = 0; j < KM; ++j) {
            const unsigned e0 = ce[j].x;
            // In range by construction: panel_build checks every entry's row
            // against its workgroup's rows (below kRows) and its column
            // against the panel, and every entry a lane loads is one of
            // those or the zero padding.
            const int col = static_cast<int>(e0 & ((1u << kPanelColBits) - 1u));
            rows[j] = static_cast<int>(e0 >> kPanelColBits);
            prod[j] = static_cast<double>(entry_value(ce[j])) * static_cast<double>(xb[col]);
        }
        // The last panel reloads itself: the loop body has no branch around
        // its loads, which keeps the compiler's wait counts exact.
        load_panel(min(p + 1, np - 1));
        // Each row belongs to one lane, whose steps add its entries in column
        // order: one LDS add per entry (ds_add_f64, no return value; no other
        // lane touches the row, and one wave's LDS operations complete in
        // order, so the row's sum is the same sequence of FP64 additions as
        // a load, add and store per entry).  Measured against that: the same
        // bits, FP32 SpMVs 2-3% faster, FP64 unchanged.
#pragma unroll
        for (int j = 0; j < KM; ++j)
            if (lane < step_count(wc, j))
                __hip_atomic_fetch_add(acc + rows[j], prod[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        // Steps beyond the unrolled KM (a wave with a long row in this
        // panel; rare by the choice of KM): loaded here, each entry added to
        // its row's accumulator after the row's earlier entries, as above.
        if (wc[1] > KM) {
            const int *hp = pv.hdr + 8 * (hfirst + p * kW + wv);
            int off = wc[0];
#pragma unroll
            for (int j = 0; j < KM; ++j) off += step_count(wc, j);
            for (int j = KM; j < wc[1]; ++j) {
                const int c = (hp[2 + j / 4] >> (8 * (j % 4))) & 0xff;
                if (lane < c) {
                    const E e = pv.ent[off + lane];
                    const int r = min(static_cast<int>(e.x >> kPanelColBits), kRows - 1);
                    const int col = min(static_cast<int>(e.x & ((1u << kPanelColBits) - 1u)), P - 1);
                    __hip_atomic_fetch_add(acc + r, static_cast<double>(entry_value(e)) * static_cast<double>(xb[col]),
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
                off += c;
            }
        }
        store_panel((p + 1) & 1);
        __syncthreads();
    }
    double *o = out + static_cast<long long>(g) * pv.rows + r0;
    for (int i = t; i < nr; i += NT) __builtin_nontemporal_store(acc[i], o + i);
}

// Workgroups [0, pv.resident) keep their entries in the MALL.
template <class V, int NT, int KM, int KB>
__global__ __launch_bounds__(NT) void k_panel_spmv(PanelView<V> pv, const DevState *__restrict__ state, int select,
                                                   const V *__restrict__ x0, const V *__restrict__ x1,
                                                   double *__restrict__ out) {
    __shared__ __attribute__((aligned(16))) V xbuf[2][panel_width<V>(KB)];
    __shared__ double acc[panel_max_rows(KB)];
    const V *x = x0;
    if (state) {
        if (state->status != VK_PDHG_RUNNING) return;
        if (select && ((state->k + 1) & 1)) x = x1;
    }
    if (static_cast<int>(blockIdx.x) < pv.resident) panel_rows<V, NT, KM, KB, true>(pv, x, out, xbuf, acc);
    else panel_rows<V, NT, KM, KB, false>(pv, x, out, xbuf, acc);
}

template <class V, class E>
__global__ void k_panel_fill(long long count, const unsigned *__restrict__ idx, const int *__restrict__ perm,
                             const V *__restrict__ val, E *__restrict__ ent) {
    for (long long k = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; k < count;
         k += static_cast<long long>(gridDim.x) * blockDim.x) {
        E e;
        e.x = idx[k];
        if constexpr (sizeof(V) == 8) {
            const double v = val[perm[k]];
            e.y = static_cast<unsigned>(__double_as_longlong(v));
            e.z = static_cast<unsigned>(static_cast<unsigned long long>(__double_as_longlong(v)) >> 32);
        } else {
            e.y = __float_as_uint(val[perm[k]]);
        }
        ent[k] = e;
    }
}

// ------------------------------------------------------------------ host layout

struct WgBuild {
    std::vector<PanelHeader> hdr;   // np * waves, base relative to this workgroup's first entry
    std::vector<unsigned> idx;
    std::vector<int> perm;
    int steps = 0;
};

int align4(long long v) { return static_cast<int>(v & ~3ll); }

// Column range of group g of G over `cols` columns: starts on a multiple of
// 4 so a panel's first 16-byte vector is aligned in both precisions.
void group_range(int cols, int G, int g, int &c0, int &c1) {
    c0 = g == 0 ? 0 : align4(static_cast<long long>(cols) * g / G);
    c1 = g + 1 == G ? cols : align4(static_cast<long long>(cols) * (g + 1) / G);
}

// Builds one workgroup: rows [r0, r1) over columns [c0, c1) of the group,
// panels of P columns.  start[r] is the first entry of row r in the group.
bool build_workgroup(const std::vector<int> &rp, const std::vector<int> &ci, const std::vector<int> &start, int r0,
                     int r1, int c0, int c1, int P, int threads, WgBuild &B) {
    const int waves = threads / kWave;
    const int np = (c1 - c0 + P - 1) / P;
    std::vector<int> cur(start.begin() + r0, start.begin() + r1);
    std::vector<int> prow, pcnt, order, lane_of, lane_cnt(threads), lane_first(threads), lane_next;
    std::vector<std::pair<int, int>> heap;
    B.hdr.assign(static_cast<std::size_t>(np) * waves, PanelHeader{});
    B.steps = 0;
    for (int p = 0; p < np; ++p) {
        const int ps = c0 + p * P, pe = std::min(c1, ps + P);
        prow.clear();
        pcnt.clear();
        for (int r = r0; r < r1; ++r) {
            int k = cur[r - r0];
            const int e = rp[r + 1];
            int k2 = k;
            while (k2 < e && ci[k2] < pe) ++k2;
            if (k2 > k) {
                prow.push_back(r);
                pcnt.push_back(k2 - k);
            }
        }
        // Longest rows first, each to the least-loaded lane; heap slots map
        // to lanes round-robin over the waves so every wave gets work.
        order.resize(prow.size());
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return pcnt[a] > pcnt[b]; });
        heap.clear();
        for (int i = 0; i < threads; ++i) heap.emplace_back(0, i);
        auto cmp = [](const std::pair<int, int> &a, const std::pair<int, int> &b) { return a > b; };
        std::make_heap(heap.begin(), heap.end(), cmp);
        lane_of.assign(prow.size(), 0);
        for (int q : order) {
            std::pop_heap(heap.begin(), heap.end(), cmp);
            auto &h = heap.back();
            const int slot = h.second;
            lane_of[q] = (slot % waves) * kWave + slot / waves;
            h.first += pcnt[q];
            std::push_heap(heap.begin(), heap.end(), cmp);
        }
        std::fill(lane_cnt.begin(), lane_cnt.end(), 0);
        for (std::size_t q = 0; q < prow.size(); ++q) lane_cnt[lane_of[q]] += pcnt[q];
        // Each lane's rows in increasing order (prow is increasing).
        std::fill(lane_first.begin(), lane_first.end(), -1);
        lane_next.assign(prow.size(), -1);
        {
            std::vector<int> last(threads, -1);
            for (std::size_t q = 0; q < prow.size(); ++q) {
                const int l = lane_of[q];
                if (last[l] < 0) lane_first[l] = static_cast<int>(q);
                else lane_next[last[l]] = static_cast<int>(q);
                last[l] = static_cast<int>(q);
            }
        }
        for (int w = 0; w < waves; ++w) {
            int ord[kWave];
            for (int i = 0; i < kWave; ++i) ord[i] = w * kWave + i;
            std::stable_sort(ord, ord + kWave, [&](int a, int b) { return lane_cnt[a] > lane_cnt[b]; });
            const int K = lane_cnt[ord[0]];
            if (K > kPanelStepCap) return false;
            B.steps = std::max(B.steps, K);
            PanelHeader &h = B.hdr[static_cast<std::size_t>(p) * waves + w];
            h.base = static_cast<int>(B.idx.size());
            h.steps = K;
            for (int j = 0; j < kPanelStepCap; ++j) {
                int c = 0;
                while (c < kWave && lane_cnt[ord[c]] > j) ++c;
                h.count[j] = static_cast<unsigned char>(c);
            }
            // The j-th entry of every sorted lane that has one, step by step.
            int pos[kWave], row_q[kWave];
            for (int s2 = 0; s2 < kWave; ++s2) {
                row_q[s2] = lane_first[ord[s2]];
                pos[s2] = row_q[s2] >= 0 ? cur[prow[row_q[s2]] - r0] : 0;
            }
            for (int j = 0; j < K; ++j)
                for (int s2 = 0; s2 < h.count[j]; ++s2) {
                    const int q = row_q[s2];
                    const int r = prow[q];
                    const int k = pos[s2];
                    B.idx.push_back((static_cast<unsigned>(r - r0) << kPanelColBits) |
                                    static_cast<unsigned>(ci[k] - ps));
                    B.perm.push_back(k);
                    if (++pos[s2] == cur[r - r0] + pcnt[q]) {
                        row_q[s2] = lane_next[q];
                        if (row_q[s2] >= 0) pos[s2] = cur[prow[row_q[s2]] - r0];
                    }
                }
        }
        for (std::size_t q = 0; q < prow.size(); ++q) cur[prow[q] - r0] += pcnt[q];
    }
    return true;
}

template <class V, int NT, int KM, int KB>
void launch_kernel(const PanelLayout &L, const DevState *state, int select, const void *x0, const void *x1,
                   double *out, hipStream_t st) {
    PanelView<V> v{L.meta, L.hdr, static_cast<const typename PanelEntry<V>::E *>(L.ent), L.rows,
                   (L.cols * static_cast<int>(sizeof(V)) + 15) / 16, L.resident};
    k_panel_spmv<V, NT, KM, KB><<<L.nwg, NT, 0, st>>>(v, state, select, static_cast<const V *>(x0),
                                                      static_cast<const V *>(x1), out);
}

template <class V, int NT, int KB>
void launch_steps(const PanelLayout &L, const DevState *state, int select, const void *x0, const void *x1,
                  double *out, hipStream_t st) {
    if (L.fast_steps <= 4) launch_kernel<V, NT, 4, KB>(L, state, select, x0, x1, out, st);
    else if (L.fast_steps <= 8) launch_kernel<V, NT, 8, KB>(L, state, select, x0, x1, out, st);
    else if (L.fast_steps <= 12) launch_kernel<V, NT, 12, KB>(L, state, select, x0, x1, out, st);
    else launch_kernel<V, NT, 16, KB>(L, state, select, x0, x1, out, st);
}

// The instantiated shapes: 48 KB panels with 256 or 512 lanes, 24 and 16 KB
// panels with 512.
template <class V>
void launch_shape(const PanelLayout &L, const DevState *state, int select, const void *x0, const void *x1,
                  double *out, hipStream_t st) {
    if (L.panel_kb == 48) {
        if (L.threads == 256) launch_steps<V, 256, 48>(L, state, select, x0, x1, out, st);
        else launch_steps<V, 512, 48>(L, state, select, x0, x1, out, st);
    } else if (L.panel_kb == 24) {
        launch_steps<V, 512, 24>(L, state, select, x0, x1, out, st);
    } else {
        launch_steps<V, 512, 16>(L, state, select, x0, x1, out, st);
    }
}

// Partial-sum bytes (stored and read back) plus half the staged bytes (the
// staging is served by L2), per launch.
double shape_traffic(long long rows, long long cols, int value_bytes, int G) {
    return 16.0 * G * static_cast<double>(rows) + 0.5 * 256.0 * static_cast<double>(cols) * value_bytes / G;
}

// Staging alone weighted 0.1: the group count of 48 KB panels when no
// shape qualifies (the measured optimum of 48 KB panels on the 400k x 1M
// LP: 4 for A x, 2 for A' y).
int groups_48kb(long long rows, long long cols, int value_bytes) {
    int best = 1;
    double best_cost = 0.0;
    for (int G : {1, 2, 4, 8}) {
        const double cost = 0.1 * 256.0 * static_cast<double>(cols) * value_bytes / G + 16.0 * G * static_cast<double>(rows);
        if (G == 1 || cost < best_cost) {
            best = G;
            best_cost = cost;
        }
    }
    return best;
}

}  // namespace

void panel_auto_shape(long long rows, long long cols, long long nnz, int value_bytes, int threads, int *groups,
                      int *panel_kb) {
    const int fixed_g = *groups, fixed_kb = *panel_kb;
    int best_g = 0, best_kb = 0;
    double best_cost = 0.0;
    for (int G : {1, 2, 4, 8}) {
        if (fixed_g > 0 && G != fixed_g) continue;
        const long long R = (rows * G + 255) / 256;
        for (int kb : kPanelKbs) {
            if ((fixed_kb > 0 && kb != fixed_kb) || (threads != 512 && kb != 48)) continue;
            const long long P = kb * 1024ll / value_bytes;
            const long long panels = (cols / G + P - 1) / P;
            if (R > panel_max_rows(kb) || panels > kPanelMaxPanels) continue;
            // The largest panel that holds the rows; skip G if it leaves the
            // lanes too little work per panel.
            if (static_cast<double>(nnz) < 3.0 * 256.0 * threads * static_cast<double>(panels)) break;
            const double cost = shape_traffic(rows, cols, value_bytes, G);
            if (best_g == 0 || cost < best_cost) {
                best_g = G;
                best_kb = kb;
                best_cost = cost;
            }
            break;
        }
    }
    if (best_g == 0) {
        best_g = fixed_g > 0 ? fixed_g : groups_48kb(rows, cols, value_bytes);
        best_kb = fixed_kb > 0 ? fixed_kb : 48;
    }
    *groups = best_g;
    *panel_kb = best_kb;
}

int panel_build(const int *rp_d, const int *ci_d, const void *val_d, int value_bytes, int rows, int cols, int nnz,
                int groups, int panel_kb, int threads, PanelLayout &out, hipStream_t st) {
    out.release();
    VK_REQUIRE(value_bytes == 8 || value_bytes == 4, "panel layout: values must be FP64 or FP32");
    VK_REQUIRE(groups >= 1 && groups <= 8 && (groups & (groups - 1)) == 0, "panel layout: groups must be 1, 2, 4 or 8");
    VK_REQUIRE(panel_kb_valid(panel_kb), "panel layout: panels must be 48, 24 or 16 KB");
    VK_REQUIRE(threads == 512 || (threads == 256 && panel_kb == 48),
               "panel layout: 512 lanes per workgroup, or 256 with 48 KB panels");
    VK_REQUIRE(rows > 0 && cols > 0, "panel layout: empty operand");
    VK_REQUIRE(static_cast<long long>(cols) * value_bytes < (1ll << 32) - 64,
               "panel layout: the gathered vector must stay below 4 GiB");
    const int P = value_bytes == 8 ? panel_width<double>(panel_kb) : panel_width<float>(panel_kb);
    const int max_rows = panel_max_rows(panel_kb);
    std::vector<int> rp(static_cast<std::size_t>(rows) + 1), ci(static_cast<std::size_t>(std::max(nnz, 1)));
    VK_HIP_CHECK(hipMemcpyAsync(rp.data(), rp_d, sizeof(int) * (rows + 1ll), hipMemcpyDeviceToHost, st));
    if (nnz) VK_HIP_CHECK(hipMemcpyAsync(ci.data(), ci_d, sizeof(int) * static_cast<std::size_t>(nnz), hipMemcpyDeviceToHost, st));
    VK_HIP_CHECK(hipStreamSynchronize(st));

    // Workgroups per group: 256 in all (one per CU), more when a group's
    // rows would not fit the LDS accumulators.
    struct Wg {
        int g, r0, r1, c0, c1;
    };
    std::vector<Wg> wgs;
    std::vector<std::vector<int>> starts(groups);
    const int base_nw = std::max(1, 256 / groups);
    for (int g = 0; g < groups; ++g) {
        int c0, c1;
        group_range(cols, groups, g, c0, c1);
        if ((c1 - c0 + P - 1) / P > kPanelMaxPanels) return VK_EINVAL;
        std::vector<int> &start = starts[g];
        start.resize(rows);
        std::vector<int> cnt(rows);
        for (int r = 0; r < rows; ++r) {
            const int *b = ci.data() + rp[r], *e = ci.data() + rp[r + 1];
            const int *a = std::lower_bound(b, e, c0);
            start[r] = static_cast<int>(a - ci.data());
            cnt[r] = static_cast<int>(std::lower_bound(a, e, c1) - a);
        }
        // Row ranges balanced by entries + lambda per row, lambda raised (or
        // workgroups added) until no range exceeds the LDS accumulators.
        // (More than 256 workgroups in all run in two rounds: only add them
        // when the rows cannot be balanced under the accumulator limit.)
        int nw = std::max(base_nw, (rows + max_rows - 1) / max_rows);
        std::vector<int> bounds;
        for (bool ok = false; !ok;) {
            for (long long lambda : {2ll, 8ll, 32ll, 1ll << 30}) {
                long long total = 0;
                for (int r = 0; r < rows; ++r) total += cnt[r] + lambda;
                bounds.assign(nw + 1, rows);
                bounds[0] = 0;
                long long run = 0;
                int wi = 1;
                for (int r = 0; r < rows && wi < nw; ++r) {
                    run += cnt[r] + lambda;
                    while (wi < nw && run * nw >= total * wi) bounds[wi++] = r + 1;
                }
                int widest = 0;
                for (int i = 0; i < nw; ++i) widest = std::max(widest, bounds[i + 1] - bounds[i]);
                if (widest <= max_rows) {
                    ok = true;
                    break;
                }
            }
            if (!ok) nw += nw / 4 + 1;
        }
        for (int i = 0; i < nw; ++i) wgs.push_back(Wg{g, bounds[i], bounds[i + 1], c0, c1});
    }
    // Block order b = i * groups + g (group = b % groups, so the workgroups
    // of one XCD work on one group): interleave the groups' workgroups.
    {
        std::vector<Wg> inter;
        std::vector<std::vector<Wg>> per(groups);
        for (const Wg &w : wgs) per[w.g].push_back(w);
        std::size_t most = 0;
        for (auto &v : per) most = std::max(most, v.size());
        for (std::size_t i = 0; i < most; ++i)
            for (int g = 0; g < groups; ++g)
                inter.push_back(i < per[g].size() ? per[g][i] : Wg{g, 0, 0, per[g].empty() ? 0 : per[g][0].c0, per[g].empty() ? 0 : per[g][0].c1});
        wgs.swap(inter);
    }
    const int nwg = static_cast<int>(wgs.size());
    std::vector<WgBuild> built(nwg);
    std::atomic<int> next{0};
    std::atomic<bool> fits{true};
    auto work = [&] {
        for (int i; (i = next.fetch_add(1)) < nwg && fits.load();) {
            const Wg &w = wgs[i];
            if (!build_workgroup(rp, ci, starts[w.g], w.r0, w.r1, w.c0, w.c1, P, threads, built[i])) fits = false;
        }
    };
    const int nthreads = std::max(1, std::min<int>(16, static_cast<int>(std::thread::hardware_concurrency())));
    std::vector<std::thread> pool;
    for (int i = 1; i < nthreads; ++i) pool.emplace_back(work);
    work();
    for (auto &th : pool) th.join();
    if (!fits) return VK_EINVAL;

    const int waves = threads / kWave;
    std::vector<PanelMeta> meta(nwg);
    std::size_t nh = 0, ne = 0;
    int steps = 0;
    for (int i = 0; i < nwg; ++i) {
        const Wg &w = wgs[i];
        PanelMeta &M = meta[i];
        M.r0 = w.r0;
        M.r1 = w.r1;
        M.c0 = w.c0;
        M.np = std::max(1, (w.c1 - w.c0 + P - 1) / P);
        M.hfirst = static_cast<int>(nh);
        M.g = w.g;
        if (built[i].hdr.size() < static_cast<std::size_t>(M.np) * waves)
            built[i].hdr.resize(static_cast<std::size_t>(M.np) * waves, PanelHeader{});
        for (PanelHeader &h : built[i].hdr) h.base += static_cast<int>(ne);
        nh += built[i].hdr.size();
        ne += built[i].idx.size();
        steps = std::max(steps, built[i].steps);
    }
    if (ne + 64 > static_cast<std::size_t>(INT32_MAX)) return VK_EINVAL;
    // Unrolled steps: the fewest of 4, 8, 12, 16 that cover all but 3% of
    // the (wave, panel) chunks; the rest run their extra steps unrolled no
    // further (the kernel's overflow loop).
    int fast_steps = 16;
    {
        std::vector<long long> hist(kPanelStepCap + 1, 0);
        long long chunks = 0;
        for (const auto &b : built)
            for (const PanelHeader &h : b.hdr) {
                ++hist[static_cast<std::size_t>(h.steps)];
                ++chunks;
            }
        for (int km : {4, 8, 12, 16}) {
            long long over = 0;
            for (int k = km + 1; k <= kPanelStepCap; ++k) over += hist[static_cast<std::size_t>(k)];
            if (over * 100 <= chunks * 3) {
                fast_steps = km;
                break;
            }
        }
        fast_steps = std::min(fast_steps, std::max(4, steps));
    }
    std::vector<PanelHeader> hdr;
    hdr.reserve(nh);
    std::vector<unsigned> idx;
    std::vector<int> perm;
    idx.reserve(ne + 64);
    perm.reserve(ne + 64);
    for (auto &b : built) {
        hdr.insert(hdr.end(), b.hdr.begin(), b.hdr.end());
        idx.insert(idx.end(), b.idx.begin(), b.idx.end());
        perm.insert(perm.end(), b.perm.begin(), b.perm.end());
        std::vector<PanelHeader>().swap(b.hdr);
        std::vector<unsigned>().swap(b.idx);
        std::vector<int>().swap(b.perm);
    }
    // Every index the kernel will follow stays in bounds: headers inside
    // the entry array, entries inside their workgroup's rows and panel, and
    // every entry belongs to exactly one chunk (an empty step's load reads
    // the next chunk's first entry, so all of them must be checked).
    std::vector<unsigned char> seen(ne, 0);
    for (int i = 0; i < nwg; ++i) {
        const PanelMeta &M = meta[i];
        const int width = wgs[i].c1 - wgs[i].c0;
        for (int p = 0; p < M.np; ++p)
            for (int w = 0; w < waves; ++w) {
                const PanelHeader &h = hdr[static_cast<std::size_t>(M.hfirst) + static_cast<std::size_t>(p) * waves + w];
                std::size_t k = static_cast<std::size_t>(h.base);
                for (int j = 0; j < kPanelStepCap; ++j) {
                    if (j > 0 && h.count[j] > h.count[j - 1]) return VK_EINVAL;
                    for (int l = 0; l < h.count[j]; ++l, ++k) {
                        if (k >= ne || seen[k]) return VK_EINVAL;
                        seen[k] = 1;
                        const unsigned e0 = idx[k];
                        if (static_cast<int>(e0 >> kPanelColBits) >= M.r1 - M.r0 ||
                            static_cast<int>(e0 & ((1u << kPanelColBits) - 1u)) >= std::min(P, width - p * P))
                            return VK_EINVAL;
                    }
                }
            }
    }
    for (unsigned char v : seen)
        if (!v) return VK_EINVAL;
    // Padding: an empty step's load reads the entry past the last one.
    idx.resize(ne + 64, 0u);
    perm.resize(ne + 64, 0);

    const std::size_t esize = value_bytes == 8 ? sizeof(panel_u32x3) : sizeof(panel_u32x2);
    PanelLayout L;
    L.groups = groups;
    L.panel_kb = panel_kb;
    L.threads = threads;
    L.steps = steps;
    L.fast_steps = fast_steps;
    L.nwg = nwg;
    L.rows = rows;
    L.cols = cols;
    L.value_bytes = value_bytes;
    L.ent_bytes = esize * idx.size();
    unsigned *idx_d = nullptr;
    int *perm_d = nullptr;
    auto fail = [&](hipError_t e) {
        L.release();
        if (idx_d) (void)hipFree(idx_d);
        if (perm_d) (void)hipFree(perm_d);
        set_error(std::string("panel layout: ") + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? VK_ENOMEM : VK_EDEVICE;
    };
    hipError_t e;
    if ((e = device_malloc(&L.meta, sizeof(PanelMeta) * nwg)) != hipSuccess) return fail(e);
    if ((e = device_malloc(&L.hdr, sizeof(PanelHeader) * std::max<std::size_t>(nh, 1))) != hipSuccess) return fail(e);
    if ((e = device_malloc(&L.ent, L.ent_bytes)) != hipSuccess) return fail(e);
    if ((e = device_malloc(&idx_d, sizeof(unsigned) * idx.size())) != hipSuccess) return fail(e);
    if ((e = device_malloc(&perm_d, sizeof(int) * perm.size())) != hipSuccess) return fail(e);
    if ((e = hipMemcpyAsync(L.meta, meta.data(), sizeof(PanelMeta) * nwg, hipMemcpyHostToDevice, st)) != hipSuccess ||
        (nh && (e = hipMemcpyAsync(L.hdr, hdr.data(), sizeof(PanelHeader) * nh, hipMemcpyHostToDevice, st)) != hipSuccess) ||
        (e = hipMemcpyAsync(idx_d, idx.data(), sizeof(unsigned) * idx.size(), hipMemcpyHostToDevice, st)) != hipSuccess ||
        (e = hipMemcpyAsync(perm_d, perm.data(), sizeof(int) * perm.size(), hipMemcpyHostToDevice, st)) != hipSuccess)
        return fail(e);
    const long long count = static_cast<long long>(idx.size());
    const int grid = static_cast<int>(std::min<long long>((count + kBlock - 1) / kBlock, 8192));
    if (value_bytes == 8)
        k_panel_fill<double, panel_u32x3><<<grid, kBlock, 0, st>>>(count, idx_d, perm_d, static_cast<const double *>(val_d),
                                                                  static_cast<panel_u32x3 *>(L.ent));
    else
        k_panel_fill<float, panel_u32x2><<<grid, kBlock, 0, st>>>(count, idx_d, perm_d, static_cast<const float *>(val_d),
                                                                 static_cast<panel_u32x2 *>(L.ent));
    if ((e = hipGetLastError()) != hipSuccess || (e = hipStreamSynchronize(st)) != hipSuccess) return fail(e);
    (void)hipFree(idx_d);
    (void)hipFree(perm_d);
    out = L;
    return VK_OK;
}

int panel_tune_steps(PanelLayout &L, int repeats, double *best_ms, hipStream_t st) {
    void *x = nullptr;
    double *out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = VK_OK;
    auto check = [&](hipError_t e) {
        if (e != hipSuccess && rc == VK_OK) {
            set_error(std::string("panel tuner: ") + hipGetErrorString(e));
            rc = e == hipErrorOutOfMemory ? VK_ENOMEM : VK_EDEVICE;
        }
    };
    check(device_malloc(&x, static_cast<std::size_t>(L.value_bytes) * (L.cols + 4)));
    check(device_malloc(&out, sizeof(double) * L.groups * static_cast<std::size_t>(L.rows)));
    check(hipEventCreate(&e0));
    check(hipEventCreate(&e1));
    if (rc == VK_OK) check(hipMemsetAsync(x, 0, static_cast<std::size_t>(L.value_bytes) * (L.cols + 4), st));
    const int automatic = L.fast_steps;
    int best = automatic;
    double best_t = 1e300;
    for (int km : {4, 8, 12, 16}) {
        if (rc != VK_OK) break;
        if (km + 4 < automatic || km > std::max(automatic + 4, 4) || km > 16) continue;
        L.fast_steps = km;
        panel_spmv_launch(L, nullptr, 0, x, x, out, st);   // warm-up
        check(hipEventRecord(e0, st));
        for (int r = 0; r < repeats; ++r) panel_spmv_launch(L, nullptr, 0, x, x, out, st);
        check(hipEventRecord(e1, st));
        check(hipEventSynchronize(e1));
        check(hipGetLastError());
        float t = 0.0f;
        check(hipEventElapsedTime(&t, e0, e1));
        const double ms = static_cast<double>(t) / std::max(1, repeats);
        if (rc == VK_OK && ms < best_t) {
            best_t = ms;
            best = km;
        }
    }
    L.fast_steps = best;
    if (best_ms) *best_ms = best_t;
    if (x) (void)hipFree(x);
    if (out) (void)hipFree(out);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return rc;
}

int panel_spmv_launch(const PanelLayout &L, const void *state, int select, const void *x0, const void *x1,
                      double *out, hipStream_t st) {
    const DevState *s = static_cast<const DevState *>(state);
    if (L.value_bytes == 8) launch_shape<double>(L, s, select, x0, x1, out, st);
    else launch_shape<float>(L, s, select, x0, x1, out, st);
    return VK_OK;
}

}  // namespace vk

using namespace vk;

extern "C" int vk_hip_csr_panel_spmv(const vk_hip_csr *A, int transpose, int precision, int groups, int panel_kb,
                                     int threads, const double *x, double *partial, int repeats, double *ms) {
    VK_REQUIRE(A && x && partial, "vk_hip_csr_panel_spmv: null argument");
    VK_REQUIRE(A->has_values, "vk_hip_csr_panel_spmv: matrix values were never set");
    VK_REQUIRE(precision == VK_PREC_FP64 || precision == VK_PREC_MIXED, "vk_hip_csr_panel_spmv: unknown precision");
    VK_REQUIRE(repeats >= 0, "vk_hip_csr_panel_spmv: negative repeat count");
    VK_REQUIRE(groups == 1 || groups == 2 || groups == 4 || groups == 8,
               "vk_hip_csr_panel_spmv: groups must be 1, 2, 4 or 8");
    VK_REQUIRE(panel_kb == 0 || panel_kb_valid(panel_kb), "vk_hip_csr_panel_spmv: panels must be 0 (automatic), 48, 24 or 16 KB");
    VK_REQUIRE(threads == 256 || threads == 512, "vk_hip_csr_panel_spmv: 256 or 512 lanes per workgroup");
    if (int rc = ensure_init()) return rc;
    if (int rc = bind_device()) return rc;
    hipStream_t st = default_stream();
    const bool mixed = precision == VK_PREC_MIXED;
    if (mixed)
        if (int rc = csr_make_f32(const_cast<vk_hip_csr *>(A), st)) return rc;
    const int rows = static_cast<int>(transpose ? A->cols : A->rows);
    const int cols = static_cast<int>(transpose ? A->rows : A->cols);
    VK_REQUIRE(rows > 0 && cols > 0, "vk_hip_csr_panel_spmv: empty operand");
    const DeviceCsr<double> &op = transpose ? A->at : A->a;
    const void *val = mixed ? static_cast<const void *>((transpose ? A->at32 : A->a32).val)
                            : static_cast<const void *>(op.val);
    const int vb = mixed ? 4 : 8;
    if (panel_kb == 0) panel_auto_shape(rows, cols, static_cast<long long>(A->nnz), vb, threads, &groups, &panel_kb);
    PanelLayout L;
    int rc = panel_build(op.row_ptr, op.col, val, vb, rows, cols, static_cast<int>(A->nnz), groups, panel_kb,
                         threads, L, st);
    if (rc == VK_EINVAL && threads == 512 && panel_kb == 48)   // a lane needs more than 24 entries in a panel
        rc = panel_build(op.row_ptr, op.col, val, vb, rows, cols, static_cast<int>(A->nnz), groups, 48, 256, L,
                         st);
    if (rc == VK_EINVAL) {
        set_error("vk_hip_csr_panel_spmv: the operand does not fit the panel layout");
        return rc;
    }
    if (rc) return rc;
    void *xd = nullptr;
    double *out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto done = [&](int code) {
        if (xd) (void)hipFree(xd);
        if (out) (void)hipFree(out);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        L.release();
        return code;
    };
    const std::size_t nout = static_cast<std::size_t>(L.groups) * rows;
    if (device_malloc(&xd, static_cast<std::size_t>(vb) * (cols + 4)) != hipSuccess ||
        device_malloc(&out, sizeof(double) * nout) != hipSuccess) {
        set_error("vk_hip_csr_panel_spmv: device allocation failed");
        return done(VK_ENOMEM);
    }
    std::vector<char> xh(static_cast<std::size_t>(vb) * (cols + 4), 0);
    for (int j = 0; j < cols; ++j) {
        if (mixed) reinterpret_cast<float *>(xh.data())[j] = static_cast<float>(x[j]);
        else reinterpret_cast<double *>(xh.data())[j] = x[j];
    }
    auto check = [&](hipError_t e) {
        if (e != hipSuccess && rc == VK_OK) {
            set_error(std::string("vk_hip_csr_panel_spmv: ") + hipGetErrorString(e));
            rc = VK_EDEVICE;
        }
    };
    check(hipMemcpyAsync(xd, xh.data(), xh.size(), hipMemcpyHostToDevice, st));
    check(hipMemsetAsync(out, 0, sizeof(double) * nout, st));
    panel_spmv_launch(L, nullptr, 0, xd, xd, out, st);
    check(hipGetLastError());
    if (repeats > 0 && rc == VK_OK) {
        check(hipEventCreate(&e0));
        check(hipEventCreate(&e1));
        check(hipEventRecord(e0, st));
        for (int r = 0; r < repeats; ++r) panel_spmv_launch(L, nullptr, 0, xd, xd, out, st);
        check(hipEventRecord(e1, st));
        check(hipEventSynchronize(e1));
        float t = 0.0f;
        check(hipEventElapsedTime(&t, e0, e1));
        if (ms) *ms = static_cast<double>(t) / repeats;
    }
    check(hipMemcpyAsync(partial, out, sizeof(double) * nout, hipMemcpyDeviceToHost, st));
    check(hipStreamSynchronize(st));
    return done(rc);
}
