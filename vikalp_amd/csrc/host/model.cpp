// Model queries and validation with the reference's rules
// (src/contracts/Model.cpp:13-199): CSR offsets must have rows+1 entries,
// start at 0 and end at nnz; columns strictly increase within a row and stay
// in range; coefficients and the objective are finite; bounds are non-NaN
// with lower <= upper; Q, when present, is square and stored symmetrically.

#include "vikalp/contracts/Model.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <utility>

namespace vikalp {
namespace {

bool check_pattern(const CsrPattern &p, const std::string &what, std::vector<std::string> &errors) {
    if (p.rows < 0 || p.columns < 0) {
        errors.push_back(what + " dimensions must be non-negative");
        return false;
    }
    if (p.row_offsets.size() != static_cast<std::size_t>(p.rows + 1)) {
        errors.push_back(what + " row_offsets must contain rows + 1 entries");
        return false;
    }
    const auto nnz = static_cast<Index>(p.column_indices.size());
    if (p.row_offsets.front() != 0 || p.row_offsets.back() != nnz) {
        errors.push_back(what + " row_offsets do not describe column_indices");
        return false;
    }
    for (std::size_t r = 0; r + 1 < p.row_offsets.size(); ++r) {
        const Index begin = p.row_offsets[r], end = p.row_offsets[r + 1];
        if (begin < 0 || end < begin || end > nnz) {
            errors.push_back(what + " row_offsets must be ordered and in range");
            return false;
        }
        Index last = -1;
        for (Index k = begin; k < end; ++k) {
            const Index c = p.column_indices[static_cast<std::size_t>(k)];
            if (c < 0 || c >= p.columns) {
                errors.push_back(what + " contains an out-of-range column");
                return false;
            }
            if (c <= last) {
                errors.push_back(what + " columns must be strictly ordered per row");
                return false;
            }
            last = c;
        }
    }
    return true;
}

bool check_matrix(const CsrMatrix &a, const std::string &what, std::vector<std::string> &errors) {
    bool ok = check_pattern(a.pattern, what, errors);
    if (a.values.size() != a.pattern.column_indices.size()) {
        errors.push_back(what + " values and column_indices sizes differ");
        ok = false;
    }
    for (const Scalar v : a.values) {
        if (!std::isfinite(v)) {
            errors.push_back(what + " coefficients must be finite");
            ok = false;
            break;
        }
    }
    return ok;
}

void check_bounds(const std::vector<Scalar> &lo, const std::vector<Scalar> &hi,
                  const std::string &what, std::vector<std::string> &errors) {
    if (lo.size() != hi.size()) {
        errors.push_back(what + " lower and upper sizes differ");
        return;
    }
    for (std::size_t i = 0; i < lo.size(); ++i)
        if (std::isnan(lo[i]) || std::isnan(hi[i]) || lo[i] > hi[i])
            errors.push_back(what + " contains an invalid bound at index " + std::to_string(i));
}

}  // namespace

Index Model::num_variables() const noexcept { return static_cast<Index>(linear_objective.size()); }

Index Model::num_constraints() const noexcept { return static_cast<Index>(constraint_lower.size()); }

bool Model::has_quadratic_objective() const noexcept { return !quadratic_objective.values.empty(); }

bool Model::has_integer_variables() const noexcept {
    for (const auto t : variable_types)
        if (t != VariableType::Continuous) return true;
    return false;
}

ProblemClass Model::problem_class() const noexcept {
    const bool integer = has_integer_variables();
    if (nonlinear) return integer ? ProblemClass::MINLP : ProblemClass::NLP;
    if (has_quadratic_objective()) return integer ? ProblemClass::MIQP : ProblemClass::QP;
    return integer ? ProblemClass::MILP : ProblemClass::LP;
}

std::vector<std::string> Model::validate() const {
    std::vector<std::string> errors;
    const Index n = num_variables();
    const Index m = num_constraints();
    bool finite_objective = std::isfinite(objective_offset);
    for (const Scalar v : linear_objective) finite_objective = finite_objective && std::isfinite(v);
    if (!finite_objective) errors.push_back("objective coefficients must be finite");
    if (constraint_upper.size() != constraint_lower.size())
        errors.push_back("constraint lower and upper sizes differ");
    if (variable_lower.size() != static_cast<std::size_t>(n) ||
        variable_upper.size() != static_cast<std::size_t>(n) ||
        variable_types.size() != static_cast<std::size_t>(n))
        errors.push_back("variable metadata sizes must equal the objective size");
    check_bounds(constraint_lower, constraint_upper, "constraint bounds", errors);
    check_bounds(variable_lower, variable_upper, "variable bounds", errors);
    check_matrix(constraint_matrix, "constraint matrix", errors);
    if (constraint_matrix.pattern.rows != m || constraint_matrix.pattern.columns != n)
        errors.push_back("constraint matrix dimensions do not match the model");

    const auto &q = quadratic_objective;
    const bool q_absent = q.pattern.rows == 0 && q.pattern.columns == 0 &&
                          q.pattern.column_indices.empty() && q.values.empty();
    const bool q_valid = check_matrix(q, "quadratic objective", errors);
    const bool q_square = q.pattern.rows == n && q.pattern.columns == n;
    if (!q_absent && !q_square) errors.push_back("quadratic objective must be square and match the variables");
    if (!q_absent && q_valid && q_square) {
        std::map<std::pair<Index, Index>, Scalar> entries;
        for (Index r = 0; r < q.pattern.rows; ++r)
            for (Index k = q.pattern.row_offsets[static_cast<std::size_t>(r)];
                 k < q.pattern.row_offsets[static_cast<std::size_t>(r + 1)]; ++k)
                entries[{r, q.pattern.column_indices[static_cast<std::size_t>(k)]}] =
                    q.values[static_cast<std::size_t>(k)];
        for (const auto &[rc, v] : entries) {
            const auto mirror = entries.find({rc.second, rc.first});
            if (mirror == entries.end() ||
                std::abs(v - mirror->second) > 1e-12 * std::max({1.0, std::abs(v), std::abs(mirror->second)})) {
                errors.push_back("quadratic objective must be stored symmetrically");
                break;
            }
        }
    }
    // Nonlinear callbacks are not part of the PDHG path (the router reports
    // NLP/MINLP models as unsupported), so their dimensions are not checked.
    return errors;
}

}  // namespace vikalp
