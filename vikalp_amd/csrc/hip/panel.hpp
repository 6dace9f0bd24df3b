// Panel layout of a CSR operand: an SpMV whose gathered vector is staged in
// LDS instead of gathered through the cache hierarchy.
//
// Why.  A random CSR SpMV issues one L1 miss per nonzero for its gather, and
// a CU keeps only ~64 misses in flight against an L2 hit latency of
// ~260-330 cycles, so the gather rate, not HBM bandwidth, bounds the tile
// kernels (about 0.2-0.3 requests per CU per cycle on the 400k x 1M LP).
// Here each workgroup (one per CU) owns a range of output rows and walks
// the columns of its group in panels of P entries: every panel of the
// gathered vector is loaded into LDS with coalesced 16-byte loads (double
// buffered, one workgroup barrier per panel), and the nonzeros of that
// panel gather from LDS.  The L1 traffic becomes coalesced lines of the
// vector and of the matrix stream.
//
// Panels are 48, 24 or 16 KB; the rest of the 160 KB of LDS holds the FP64
// row accumulators.  Every workgroup stages its group's whole slice of the
// vector, so the staged bytes of a launch are (workgroups x slice): more
// groups with more rows per workgroup (smaller panels, more accumulators)
// stage less, at the price of more partial sums and more panels.
//
// Columns are split into G groups (workgroup b works on group b % G, so
// each XCD's workgroups stage one group's slice and keep it in their L2);
// each group writes FP64 partial sums of its rows, added in group order by
// the row / column pass that follows (as the column-blocked layout's are).
//
// Within a workgroup, a panel's nonzeros of its rows are dealt to lanes one
// whole row at a time (longest rows first, each to the least-loaded lane),
// a lane's entries in row order and, within a row, in column order.  A
// lane adds its entries to the row's FP64 accumulator in LDS in that order
// and panels run in column order, so each group's partial of a row is the
// plain sequential sum of that row's products over the group's columns:
// bitwise independent of P, the lane count and the lane assignment, which
// the tuner may therefore change freely.  With G = 1 it equals the
// sequential CSR row sum.
//
// Entries of one wave and panel are stored as a jagged diagonal: step j
// holds the j-th entry of every lane that has one, lanes ordered by
// decreasing count, so each step is one coalesced load.  Each entry packs
// (row within the workgroup << 14 | column within the panel) with the value
// (FP64: 12 bytes; FP32: 8 bytes).  A 32-byte header per wave and panel
// gives the first entry and the lane count of each step.
#pragma once

#include "common.hpp"
#include "runtime.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace vk {

constexpr int kPanelStepCap = 24;        // entries per lane and panel a layout may need
constexpr int kPanelMaxPanels = 128;     // panels per workgroup (headers held in 2 VGPR sets)
constexpr int kPanelColBits = 14;        // column-in-panel field of an entry
constexpr int kPanelLds = 160 * 1024;    // LDS of a workgroup (one per CU)

// Panel sizes a layout may use (KB of LDS per staged panel; two are staged).
constexpr int kPanelKbs[3] = {48, 24, 16};

constexpr bool panel_kb_valid(int kb) { return kb == 48 || kb == 24 || kb == 16; }

// Entries of the gathered vector per panel.
template <class V>
constexpr int panel_width(int kb) { return kb * 1024 / static_cast<int>(sizeof(V)); }

// FP64 accumulators a workgroup can hold next to its two staged panels.
constexpr int panel_max_rows(int kb) { return (kPanelLds - 2 * kb * 1024) / 8 - 64; }

typedef unsigned panel_u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned panel_u32x3 __attribute__((ext_vector_type(3)));
typedef unsigned panel_u32x4 __attribute__((ext_vector_type(4)));

template <class V>
struct PanelEntry;
template <>
struct PanelEntry<double> {
    using E = panel_u32x3;   // index word, value low word, value high word
};
template <>
struct PanelEntry<float> {
    using E = panel_u32x2;   // index word, value bits
};

// Per-workgroup description (blockIdx-indexed).
struct PanelMeta {
    int r0, r1;      // output rows [r0, r1)
    int c0;          // first column of the workgroup's group (a multiple of 4)
    int np;          // panels
    int hfirst;      // header index of (panel 0, wave 0); panel p wave w at hfirst + p * waves + w
    int g;           // group: partial sums go to out[g * rows + r]
    int pad0, pad1;
};

// 32-byte header of one wave's entries in one panel: the first entry, the
// step count, and the lane count of each step (bytes).
struct PanelHeader {
    int base;
    int steps;
    unsigned char count[kPanelStepCap];
};
static_assert(sizeof(PanelHeader) == 32, "panel header is 32 bytes");

template <class V>
struct PanelView {
    const PanelMeta *meta;
    const int *hdr;                              // PanelHeader words
    const typename PanelEntry<V>::E *ent;
    int rows;                                    // output rows (stride of the partial sums)
    int xvecs;                                   // 16-byte vectors readable from the gathered vector
    int resident;                                // workgroups [0, resident) keep their entries cached
};

// The device layout of one operand in one value type.
struct PanelLayout {
    int groups = 0;          // G (0: not built)
    int panel_kb = 0;        // KB per staged panel (48, 24 or 16)
    int threads = 0;         // lanes per workgroup (256 or 512; 256 only with 48 KB panels)
    int steps = 0;           // most entries any lane has in one panel
    int fast_steps = 0;      // steps the kernel unrolls (4, 8, 12 or 16); a chunk with
                             // more runs the rest in a loop
    int nwg = 0;             // workgroups
    int resident = 0;        // workgroups [0, resident) load their entries with
                             // cache-allocating loads (MALL-resident), the rest
                             // stream them (a load hint: results do not change)
    int rows = 0, cols = 0;  // output rows, gathered entries
    int value_bytes = 0;
    PanelMeta *meta = nullptr;
    int *hdr = nullptr;
    void *ent = nullptr;
    std::size_t ent_bytes = 0;
    void release();
};

// Builds the layout of the CSR operand (device pattern rp, ci; rows x cols;
// values val of `value_bytes` = 8 (double) or 4 (float)) with `groups`
// column groups, panels of `panel_kb` KB and `threads` lanes per workgroup.
// Returns VK_OK, VK_EINVAL (the operand does not fit the layout: a row with
// too many entries in one panel, or too many panels; nothing is allocated),
// or an error code.
int panel_build(const int *rp, const int *ci, const void *val, int value_bytes, int rows, int cols, int nnz,
                int groups, int panel_kb, int threads, PanelLayout &out, hipStream_t st);

// Times `repeats` launches of L (on a zero vector) for each unroll depth
// among 4, 8, 12, 16 not below L's automatic choice by more than one step,
// and keeps the fastest in L.fast_steps; returns the mean ms of that choice.
// The depth never changes results.
int panel_tune_steps(PanelLayout &L, int repeats, double *best_ms, hipStream_t st);

// Automatic shape of an operand (rows x cols, nnz entries, gathered values
// of value_bytes, `threads` lanes): the group count and the panel size,
// each unless given (*groups > 0, *panel_kb > 0).  For each group count G
// the largest panel size whose accumulators hold a 256-workgroup row range
// (rows G / 256) is taken; shapes that leave a lane fewer than 3 entries
// per panel on average are skipped; among the rest the smallest modelled
// traffic wins: 16 G rows bytes of partial sums (stored and read back) plus
// half the staged bytes 256 cols value_bytes / G (served by L2).  When no
// shape qualifies: 48 KB panels and the group count of the staging model
// with staged bytes weighted 0.1.
void panel_auto_shape(long long rows, long long cols, long long nnz, int value_bytes, int threads, int *groups,
                      int *panel_kb);

// Launches y-partials = A x on `st` (FP64 partial sums, out[g * rows + r]).
// Gathered vector: x1 when select != 0 and the device state's (k + 1) is
// odd, else x0 (the solver's double-buffered iterate); `state` may be null
// (always x0).  The kernel returns at once when the state's status is not
// running.  x0 / x1 must be readable up to layout.cols rounded up to a
// 16-byte multiple.
int panel_spmv_launch(const PanelLayout &L, const void *state, int select, const void *x0, const void *x1,
                      double *out, hipStream_t st);

}  // namespace vk
