// Seeded synthetic LP generators for the PDHG workloads.
//
// Every LP is   min c'x  s.t.  rl <= A x <= ru,  0 <= x <= 10,
// built from a hidden primal point x0 ~ U(0.5, 9.5) so that it is feasible,
// and a dual point y0 ~ N(0,1) so that c = A'y0 + s with s ~ U(0,1) on
// exactly half of the columns (chosen by a seeded permutation).
//
// Families (one generator is used by the CPU oracle runs, the HIP runs and
// the benchmark, so every machine rebuilds the same instance from a seed):
//   1  uniform:  each column has exactly k nonzeros at distinct uniform rows,
//                values N(0,1); every row is an equality b = A x0.
//                (1000 x 3000, k=5  and  100000 x 250000, k=6)
//   2  zipf:     column lengths from a Zipf(s=1.5) law truncated to [2,200];
//                a random 10% of columns scaled by 10^U(-2,2); a random 70% of
//                rows are equalities, the rest are <= rows with
//                rhs = (A x0)_i + U(0,1).  (400000 x 1000000)
//   4  block:    multi-period block-angular LP: T diagonal blocks of
//                150 x 400 with 6 nonzeros per column, 40 coupling rows
//                between periods t and t+1 (+1 on a period-t column, -1 on a
//                period-t+1 column), equality rows, then U{1..20} variables get
//                a branching bound: upper = floor(x0_j) or lower = ceil(x0_j).
//
// Randomness: splitmix64 seeds independent xoshiro256** streams keyed by
// (seed, purpose); normals by Box-Muller (one value per pair of uniforms).

#include "lpgen.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <new>
#include <numeric>

namespace vikalp::gen {
namespace {

std::uint64_t splitmix64(std::uint64_t &x) {
    std::uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

std::uint64_t rotl(std::uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

enum Purpose : std::uint64_t {
    kStructure = 1, kValues, kPrimal, kDual, kSupport, kSlack, kLengths,
    kColumnScale, kRowKind, kRowSlack, kCoupling, kBranching, kPeriods
};

}  // namespace

Rng::Rng(std::uint64_t seed, std::uint64_t purpose) {
    std::uint64_t sm = seed ^ (purpose * 0xD1B54A32D192ED03ULL);
    for (auto &w : s_) w = splitmix64(sm);
}

std::uint64_t Rng::next() {
    const std::uint64_t result = rotl(s_[1] * 5, 7) * 9;
    const std::uint64_t t = s_[1] << 17;
    s_[2] ^= s_[0];
    s_[3] ^= s_[1];
    s_[1] ^= s_[2];
    s_[0] ^= s_[3];
    s_[2] ^= t;
    s_[3] = rotl(s_[3], 45);
    return result;
}

double Rng::uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }

std::uint64_t Rng::below(std::uint64_t bound) {
    return static_cast<std::uint64_t>((static_cast<unsigned __int128>(next()) * bound) >> 64);
}

double Rng::normal() {
    const double u1 = static_cast<double>((next() >> 11) + 1) * 0x1.0p-53;  // (0,1]
    const double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

namespace {

// Partial Fisher-Yates on a persistent permutation of [0, size): the first k
// entries after the call are k distinct uniform draws.
struct DistinctPicker {
    std::vector<std::int64_t> perm;
    explicit DistinctPicker(std::int64_t size) : perm(static_cast<std::size_t>(size)) {
        std::iota(perm.begin(), perm.end(), std::int64_t{0});
    }
    void pick(Rng &rng, std::int64_t k, std::vector<std::int64_t> &out) {
        const auto size = static_cast<std::int64_t>(perm.size());
        out.resize(static_cast<std::size_t>(k));
        for (std::int64_t i = 0; i < k; ++i) {
            const auto j = i + static_cast<std::int64_t>(rng.below(static_cast<std::uint64_t>(size - i)));
            std::swap(perm[static_cast<std::size_t>(i)], perm[static_cast<std::size_t>(j)]);
            out[static_cast<std::size_t>(i)] = perm[static_cast<std::size_t>(i)];
        }
        std::sort(out.begin(), out.end());
    }
};

// A random subset of exactly `count` of [0, size) as a 0/1 mask.
std::vector<char> random_subset(std::uint64_t seed, std::uint64_t purpose,
                                std::int64_t size, std::int64_t count) {
    Rng rng(seed, purpose);
    std::vector<std::int64_t> perm(static_cast<std::size_t>(size));
    std::iota(perm.begin(), perm.end(), std::int64_t{0});
    std::vector<char> mask(static_cast<std::size_t>(size), 0);
    for (std::int64_t i = 0; i < count; ++i) {
        const auto j = i + static_cast<std::int64_t>(rng.below(static_cast<std::uint64_t>(size - i)));
        std::swap(perm[static_cast<std::size_t>(i)], perm[static_cast<std::size_t>(j)]);
        mask[static_cast<std::size_t>(perm[static_cast<std::size_t>(i)])] = 1;
    }
    return mask;
}

// Column-major triplets (rows sorted within each column) -> CSR with rows
// in increasing column order (strictly increasing columns per row).
void csc_to_csr(std::int64_t m, std::int64_t n, const std::vector<std::int64_t> &col_ptr,
                const std::vector<std::int64_t> &rows, const std::vector<double> &vals,
                Lp &lp) {
    const auto nnz = static_cast<std::int64_t>(rows.size());
    lp.row_offsets.assign(static_cast<std::size_t>(m + 1), 0);
    for (auto r : rows) ++lp.row_offsets[static_cast<std::size_t>(r + 1)];
    for (std::int64_t r = 0; r < m; ++r)
        lp.row_offsets[static_cast<std::size_t>(r + 1)] += lp.row_offsets[static_cast<std::size_t>(r)];
    lp.column_indices.resize(static_cast<std::size_t>(nnz));
    lp.values.resize(static_cast<std::size_t>(nnz));
    std::vector<std::int64_t> fill(lp.row_offsets.begin(), lp.row_offsets.end() - 1);
    for (std::int64_t j = 0; j < n; ++j) {
        for (auto k = col_ptr[static_cast<std::size_t>(j)]; k < col_ptr[static_cast<std::size_t>(j + 1)]; ++k) {
            const auto r = rows[static_cast<std::size_t>(k)];
            const auto pos = fill[static_cast<std::size_t>(r)]++;
            lp.column_indices[static_cast<std::size_t>(pos)] = j;
            lp.values[static_cast<std::size_t>(pos)] = vals[static_cast<std::size_t>(k)];
        }
    }
}

// b = A x0 (row order), c = A'y0 + s (row-order scatter), x0, bounds [0,10].
void finish_lp(std::uint64_t seed, Lp &lp) {
    const auto m = lp.m, n = lp.n;
    lp.x0.resize(static_cast<std::size_t>(n));
    Rng primal(seed, kPrimal);
    for (auto &v : lp.x0) v = 0.5 + 9.0 * primal.uniform();
    lp.ax0.assign(static_cast<std::size_t>(m), 0.0);
    for (std::int64_t r = 0; r < m; ++r) {
        double acc = 0.0;
        for (auto k = lp.row_offsets[static_cast<std::size_t>(r)]; k < lp.row_offsets[static_cast<std::size_t>(r + 1)]; ++k)
            acc += lp.values[static_cast<std::size_t>(k)] *
                   lp.x0[static_cast<std::size_t>(lp.column_indices[static_cast<std::size_t>(k)])];
        lp.ax0[static_cast<std::size_t>(r)] = acc;
    }
    lp.y0.resize(static_cast<std::size_t>(m));
    Rng dual(seed, kDual);
    for (auto &v : lp.y0) v = dual.normal();
    const auto support = random_subset(seed, kSupport, n, n / 2);
    Rng slack(seed, kSlack);
    lp.objective.assign(static_cast<std::size_t>(n), 0.0);
    for (std::int64_t j = 0; j < n; ++j)
        if (support[static_cast<std::size_t>(j)]) lp.objective[static_cast<std::size_t>(j)] = slack.uniform();
    for (std::int64_t r = 0; r < m; ++r) {
        const double yr = lp.y0[static_cast<std::size_t>(r)];
        for (auto k = lp.row_offsets[static_cast<std::size_t>(r)]; k < lp.row_offsets[static_cast<std::size_t>(r + 1)]; ++k)
            lp.objective[static_cast<std::size_t>(lp.column_indices[static_cast<std::size_t>(k)])] +=
                lp.values[static_cast<std::size_t>(k)] * yr;
    }
    lp.var_lower.assign(static_cast<std::size_t>(n), 0.0);
    lp.var_upper.assign(static_cast<std::size_t>(n), 10.0);
}

}  // namespace

void make_uniform(std::int64_t m, std::int64_t n, std::int64_t k, std::uint64_t seed, Lp &lp) {
    lp = Lp{};
    lp.m = m;
    lp.n = n;
    std::vector<std::int64_t> col_ptr(static_cast<std::size_t>(n + 1), 0);
    std::vector<std::int64_t> rows;
    std::vector<double> vals;
    rows.reserve(static_cast<std::size_t>(n * k));
    vals.reserve(static_cast<std::size_t>(n * k));
    Rng structure(seed, kStructure);
    Rng values(seed, kValues);
    DistinctPicker picker(m);
    std::vector<std::int64_t> pick;
    for (std::int64_t j = 0; j < n; ++j) {
        picker.pick(structure, std::min(k, m), pick);
        for (auto r : pick) {
            rows.push_back(r);
            vals.push_back(values.normal());
        }
        col_ptr[static_cast<std::size_t>(j + 1)] = static_cast<std::int64_t>(rows.size());
    }
    csc_to_csr(m, n, col_ptr, rows, vals, lp);
    finish_lp(seed, lp);
    lp.row_lower = lp.ax0;
    lp.row_upper = lp.ax0;
}

void make_zipf(std::int64_t m, std::int64_t n, std::uint64_t seed, Lp &lp) {
    lp = Lp{};
    lp.m = m;
    lp.n = n;
    constexpr int kMin = 2, kMax = 200;
    std::vector<double> cdf(kMax - kMin + 1);
    double total = 0.0;
    for (int len = kMin; len <= kMax; ++len) {
        total += std::pow(static_cast<double>(len), -1.5);
        cdf[static_cast<std::size_t>(len - kMin)] = total;
    }
    for (auto &v : cdf) v /= total;
    const auto scaled_cols = random_subset(seed, kColumnScale, n, n / 10);
    Rng scale_rng(seed, kColumnScale + 100);
    Rng lengths(seed, kLengths);
    Rng structure(seed, kStructure);
    Rng values(seed, kValues);
    DistinctPicker picker(m);
    std::vector<std::int64_t> col_ptr(static_cast<std::size_t>(n + 1), 0);
    std::vector<std::int64_t> rows;
    std::vector<double> vals;
    rows.reserve(static_cast<std::size_t>(n * 18));
    vals.reserve(static_cast<std::size_t>(n * 18));
    std::vector<std::int64_t> pick;
    for (std::int64_t j = 0; j < n; ++j) {
        const double u = lengths.uniform();
        const auto it = std::lower_bound(cdf.begin(), cdf.end(), u);
        std::int64_t len = kMin + static_cast<std::int64_t>(it - cdf.begin());
        len = std::min<std::int64_t>(std::min<std::int64_t>(len, kMax), m);
        const double factor = scaled_cols[static_cast<std::size_t>(j)]
                                  ? std::pow(10.0, -2.0 + 4.0 * scale_rng.uniform())
                                  : 1.0;
        picker.pick(structure, len, pick);
        for (auto r : pick) {
            rows.push_back(r);
            vals.push_back(values.normal() * factor);
        }
        col_ptr[static_cast<std::size_t>(j + 1)] = static_cast<std::int64_t>(rows.size());
    }
    csc_to_csr(m, n, col_ptr, rows, vals, lp);
    finish_lp(seed, lp);
    const auto equality = random_subset(seed, kRowKind, m, static_cast<std::int64_t>(std::llround(0.7 * static_cast<double>(m))));
    Rng row_slack(seed, kRowSlack);
    lp.row_lower.resize(static_cast<std::size_t>(m));
    lp.row_upper.resize(static_cast<std::size_t>(m));
    for (std::int64_t r = 0; r < m; ++r) {
        const double b = lp.ax0[static_cast<std::size_t>(r)];
        if (equality[static_cast<std::size_t>(r)]) {
            lp.row_lower[static_cast<std::size_t>(r)] = b;
            lp.row_upper[static_cast<std::size_t>(r)] = b;
        } else {
            lp.row_lower[static_cast<std::size_t>(r)] = -std::numeric_limits<double>::infinity();
            lp.row_upper[static_cast<std::size_t>(r)] = b + row_slack.uniform();
        }
    }
}

void make_block(std::int64_t periods, std::uint64_t seed, Lp &lp) {
    lp = Lp{};
    constexpr std::int64_t kRows = 150, kCols = 400, kPerCol = 6, kLink = 40;
    const std::int64_t T = periods;
    const std::int64_t m = kRows * T + kLink * (T - 1);
    const std::int64_t n = kCols * T;
    lp.m = m;
    lp.n = n;
    Rng structure(seed, kStructure);
    Rng values(seed, kValues);
    Rng coupling(seed, kCoupling);
    DistinctPicker picker(kRows);
    // Coupling rows: row kRows*T + 40*t + q has +1 on a period-t column and
    // -1 on a period-(t+1) column.
    std::vector<std::vector<std::pair<std::int64_t, double>>> extra(static_cast<std::size_t>(n));
    for (std::int64_t t = 0; t + 1 < T; ++t) {
        for (std::int64_t q = 0; q < kLink; ++q) {
            const std::int64_t row = kRows * T + kLink * t + q;
            const auto a = t * kCols + static_cast<std::int64_t>(coupling.below(kCols));
            const auto b = (t + 1) * kCols + static_cast<std::int64_t>(coupling.below(kCols));
            extra[static_cast<std::size_t>(a)].push_back({row, 1.0});
            extra[static_cast<std::size_t>(b)].push_back({row, -1.0});
        }
    }
    std::vector<std::int64_t> col_ptr(static_cast<std::size_t>(n + 1), 0);
    std::vector<std::int64_t> rows;
    std::vector<double> vals;
    std::vector<std::int64_t> pick;
    std::vector<std::pair<std::int64_t, double>> column;
    for (std::int64_t j = 0; j < n; ++j) {
        const std::int64_t t = j / kCols;
        picker.pick(structure, kPerCol, pick);
        column.clear();
        for (auto r : pick) column.push_back({t * kRows + r, values.normal()});
        // Each coupling row touches one period-t and one period-(t+1)
        // column, so a column never meets the same row twice; the merge only
        // guards the strictly-increasing CSR invariant.
        for (const auto &e : extra[static_cast<std::size_t>(j)]) {
            auto found = std::find_if(column.begin(), column.end(),
                                      [&](const auto &p) { return p.first == e.first; });
            if (found != column.end()) found->second += e.second;
            else column.push_back(e);
        }
        std::sort(column.begin(), column.end());
        for (const auto &e : column) {
            rows.push_back(e.first);
            vals.push_back(e.second);
        }
        col_ptr[static_cast<std::size_t>(j + 1)] = static_cast<std::int64_t>(rows.size());
    }
    csc_to_csr(m, n, col_ptr, rows, vals, lp);
    finish_lp(seed, lp);
    lp.row_lower = lp.ax0;
    lp.row_upper = lp.ax0;
    Rng branching(seed, kBranching);
    const auto count = 1 + static_cast<std::int64_t>(branching.below(20));
    for (std::int64_t i = 0; i < count; ++i) {
        const auto j = static_cast<std::size_t>(branching.below(static_cast<std::uint64_t>(n)));
        if (branching.below(2) == 0) lp.var_upper[j] = std::floor(lp.x0[j]);
        else lp.var_lower[j] = std::ceil(lp.x0[j]);
        if (lp.var_lower[j] > lp.var_upper[j]) lp.var_lower[j] = lp.var_upper[j];
    }
}

std::int64_t block_periods(std::uint64_t seed, std::int64_t instance) {
    Rng rng(seed + static_cast<std::uint64_t>(instance) * 0x9E3779B97F4A7C15ULL, kPeriods);
    return 4 + static_cast<std::int64_t>(rng.below(45));  // U{4..48}
}

}  // namespace vikalp::gen

// ------------------------------------------------------------------ C ABI --

using vikalp::gen::Lp;

extern "C" {

int vk_lp_generate(int family, std::int64_t m, std::int64_t n, std::int64_t k,
                   std::uint64_t seed, vk_lp **out) {
    if (!out) return -1;
    *out = nullptr;
    auto *lp = new (std::nothrow) Lp();
    if (!lp) return -2;
    try {
        switch (family) {
        case 1:
            if (m <= 0 || n <= 0 || k <= 0) { delete lp; return -1; }
            vikalp::gen::make_uniform(m, n, k, seed, *lp);
            break;
        case 2:
            if (m < 200 || n <= 0) { delete lp; return -1; }
            vikalp::gen::make_zipf(m, n, seed, *lp);
            break;
        case 4:
            if (k < 1) { delete lp; return -1; }
            vikalp::gen::make_block(k, seed, *lp);
            break;
        default:
            delete lp;
            return -1;
        }
    } catch (...) {
        delete lp;
        return -2;
    }
    *out = reinterpret_cast<vk_lp *>(lp);
    return 0;
}

int vk_lp_dims(const vk_lp *h, std::int64_t *m, std::int64_t *n, std::int64_t *nnz) {
    if (!h) return -1;
    const auto *lp = reinterpret_cast<const Lp *>(h);
    *m = lp->m;
    *n = lp->n;
    *nnz = static_cast<std::int64_t>(lp->values.size());
    return 0;
}

int vk_lp_export(const vk_lp *h, std::int64_t *row_offsets, std::int64_t *column_indices,
                 double *values, double *objective, double *row_lower, double *row_upper,
                 double *var_lower, double *var_upper, double *x0, double *y0) {
    if (!h) return -1;
    const auto *lp = reinterpret_cast<const Lp *>(h);
    auto put = [](auto *dst, const auto &src) {
        if (dst && !src.empty()) std::memcpy(dst, src.data(), src.size() * sizeof(src[0]));
    };
    put(row_offsets, lp->row_offsets);
    put(column_indices, lp->column_indices);
    put(values, lp->values);
    put(objective, lp->objective);
    put(row_lower, lp->row_lower);
    put(row_upper, lp->row_upper);
    put(var_lower, lp->var_lower);
    put(var_upper, lp->var_upper);
    put(x0, lp->x0);
    put(y0, lp->y0);
    return 0;
}

void vk_lp_free(vk_lp *h) { delete reinterpret_cast<Lp *>(h); }

std::int64_t vk_lp_block_periods(std::uint64_t seed, std::int64_t instance) {
    return vikalp::gen::block_periods(seed, instance);
}

}  // extern "C"
