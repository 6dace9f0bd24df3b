#pragma once
// Seeded synthetic LP generators (see lpgen.cpp for the families).

#include "vikalp_lpgen.h"

#include <cstdint>
#include <vector>

namespace vikalp::gen {

class Rng {
public:
    Rng(std::uint64_t seed, std::uint64_t purpose);
    std::uint64_t next();
    double uniform();                          // [0, 1)
    std::uint64_t below(std::uint64_t bound);  // [0, bound)
    double normal();

private:
    std::uint64_t s_[4];
};

struct Lp {
    std::int64_t m = 0, n = 0;
    std::vector<std::int64_t> row_offsets, column_indices;
    std::vector<double> values, objective, row_lower, row_upper, var_lower, var_upper;
    std::vector<double> x0, y0, ax0;
};

void make_uniform(std::int64_t m, std::int64_t n, std::int64_t k, std::uint64_t seed, Lp &lp);
void make_zipf(std::int64_t m, std::int64_t n, std::uint64_t seed, Lp &lp);
void make_block(std::int64_t periods, std::uint64_t seed, Lp &lp);
std::int64_t block_periods(std::uint64_t seed, std::int64_t instance);

}  // namespace vikalp::gen
