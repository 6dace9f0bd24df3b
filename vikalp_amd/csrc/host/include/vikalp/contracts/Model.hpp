#pragma once
// The canonical model the PDHG path consumes.  Same names, member order and
// semantics as the reference contract (include/vikalp/contracts/Model.hpp:9-68)
// so that code written against it compiles against this library unchanged:
//
//   min c'x + 0.5 x'Qx (+ f(x))   s.t.  constraint_lower <= A x <= constraint_upper,
//                                       variable_lower   <=  x  <= variable_upper
//
// A and Q are CSR with int64 offsets/indices; an absent Q is the 0x0 matrix.

#include <cstdint>
#include <limits>
#include <memory>
#include <string>
#include <vector>

namespace vikalp {

using Index = std::int64_t;
using Scalar = double;

class NonlinearOracle;   // NLP callbacks: not part of the PDHG path

struct CsrPattern {
    Index rows = 0;
    Index columns = 0;
    std::vector<Index> row_offsets{0};
    std::vector<Index> column_indices;
};

struct CsrMatrix {
    CsrPattern pattern;
    std::vector<Scalar> values;
};

enum class VariableType { Continuous, Integer, Binary };
enum class ProblemClass { LP, QP, NLP, MILP, MIQP, MINLP };

struct Model {
    std::string name;
    std::vector<Scalar> linear_objective;
    Scalar objective_offset = 0.0;
    CsrMatrix quadratic_objective;
    CsrMatrix constraint_matrix;
    std::vector<Scalar> constraint_lower;
    std::vector<Scalar> constraint_upper;
    std::vector<Scalar> variable_lower;
    std::vector<Scalar> variable_upper;
    std::vector<VariableType> variable_types;
    std::shared_ptr<const NonlinearOracle> nonlinear;

    [[nodiscard]] Index num_variables() const noexcept;
    [[nodiscard]] Index num_constraints() const noexcept;
    [[nodiscard]] bool has_quadratic_objective() const noexcept;
    [[nodiscard]] bool has_integer_variables() const noexcept;
    [[nodiscard]] ProblemClass problem_class() const noexcept;
    // Human-readable problems, empty when valid (src/contracts/Model.cpp:118-199).
    [[nodiscard]] std::vector<std::string> validate() const;

    static constexpr Scalar infinity() noexcept { return std::numeric_limits<Scalar>::infinity(); }
};

}  // namespace vikalp
