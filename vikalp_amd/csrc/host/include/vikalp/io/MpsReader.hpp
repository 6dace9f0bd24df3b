#pragma once
// MPS reader with the reference's interface (include/vikalp/io/MpsReader.hpp:9-38):
// read(path, model) fills a Model and returns whether it is valid; errors()
// lists what went wrong with 1-based line numbers (0 = no single line).

#include "vikalp/contracts/Model.hpp"

#include <cstddef>
#include <string>
#include <vector>

namespace vikalp {

struct MpsError {
    std::size_t line = 0;
    std::string message;
};

class MpsReader {
public:
    bool read(const std::string &path, Model &model);

    [[nodiscard]] const std::vector<MpsError> &errors() const noexcept;

private:
    std::vector<MpsError> errors_;
};

}  // namespace vikalp
