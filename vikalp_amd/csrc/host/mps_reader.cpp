// Free-format MPS reader: MpsReader::read of src/io/MpsReader.cpp:10-553.
//
// Every file the reference accepts gives the same Model:
//   - sections NAME, ROWS (N, L, G, E), COLUMNS (one or two row/value pairs
//     per line), RHS (likewise), BOUNDS (LO, UP, FX with a value; FR, BV),
//     ENDATA; NAME, ROWS, COLUMNS, RHS and ENDATA are required;
//   - variables in order of first appearance in COLUMNS, rows in ROWS order
//     (a second N row replaces the first as the objective);
//   - default bounds [0, +inf); a repeated coefficient, objective entry or
//     right-hand side keeps the last value; a right-hand side on a row that
//     is not a constraint (the objective row included) is ignored;
//   - lines outside any section (e.g. OBJSENSE) are ignored; objective sense
//     is always minimise.
// Every file it rejects by one of its messages is rejected by the same
// message on the same line (0 where it names no line).  One exception, where
// the reference drops data: it takes any line whose first field is a
// section name as that section's header, so an indented data line such as
// " RHS LIM 10" (an RHS vector called RHS) starts a new RHS section and its
// values are lost.  Here an indented line with three or more fields inside a
// section is always data; section headers are the other lines, as before.
//
// Beyond that, the standard features of netlib and MIPLIB files, all of
// which the reference rejects, are read:
//   - comment lines (first non-blank character '*'), blank lines made of
//     spaces, CRLF line ends;
//   - a RANGES section: on an E row R > 0 gives [b, b + |R|] and R < 0
//     [b - |R|, b]; on an L row [b - |R|, b]; on a G row [b, b + |R|];
//   - BOUNDS MI (lower -inf), PL (upper +inf), LI / UI (integer variable,
//     lower / upper bound);
//   - MARKER lines 'INTORG' / 'INTEND' in COLUMNS (integer variables);
//   - an UP bound below 0 on a variable whose lower bound is still the
//     default 0 makes that lower bound -inf.
// The matrix is assembled by sorting (row, column) pairs, O(nnz log nnz);
// the reference looks up every pair of its dense rows x columns grid.

#include "vikalp/io/MpsReader.hpp"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <string_view>
#include <unordered_map>
#include <vector>

namespace vikalp {
namespace {

std::vector<std::string_view> tokens(std::string_view line) {
    std::vector<std::string_view> out;
    std::size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && (line[i] == ' ' || line[i] == '\t')) ++i;
        const std::size_t b = i;
        while (i < line.size() && line[i] != ' ' && line[i] != '\t') ++i;
        if (i > b) out.push_back(line.substr(b, i - b));
    }
    return out;
}

// A whole token as a double (decimal or exponent form, optional sign).
bool number(std::string_view t, double &v) {
    if (t.empty()) return false;
    const std::string s(t);
    const char c0 = s[0] == '+' || s[0] == '-' ? (s.size() > 1 ? s[1] : '\0') : s[0];
    if (!(std::isdigit(static_cast<unsigned char>(c0)) || c0 == '.')) return false;   // no inf / nan / hex words
    char *end = nullptr;
    errno = 0;
    v = std::strtod(s.c_str(), &end);
    return end == s.c_str() + s.size() && errno != ERANGE;
}

enum class Section { None, Rows, Columns, Rhs, Ranges, Bounds };

struct Entry {
    std::int64_t variable;
    std::string row;
    double value;
};

struct Bound {
    std::string type;
    std::string variable;
    double value;
};

}  // namespace

bool MpsReader::read(const std::string &path, Model &model) {
    errors_.clear();
    std::ifstream input(path, std::ios::binary);
    if (!input) {
        errors_.push_back({0, "unable to open MPS file: " + path});
        return false;
    }
    model = Model{};
    std::stringstream buffer;
    buffer << input.rdbuf();
    const std::string text = buffer.str();

    bool saw_name = false, saw_rows = false, saw_columns = false, saw_rhs = false, saw_endata = false;
    Section section = Section::None;
    std::string objective_row;
    std::vector<std::string> row_names;
    std::vector<char> row_types;
    std::vector<std::string> variables;
    std::vector<VariableType> types;
    std::unordered_map<std::string, std::int64_t> variable_index;
    std::vector<Entry> entries;
    std::unordered_map<std::string, double> rhs, ranges;
    std::vector<Bound> bounds;
    bool integer_block = false;

    std::size_t line_number = 0, pos = 0;
    while (pos < text.size()) {
        std::size_t end = text.find('\n', pos);
        if (end == std::string::npos) end = text.size();
        std::string_view line(text.data() + pos, end - pos);
        pos = end + 1;
        ++line_number;
        if (!line.empty() && line.back() == '\r') line.remove_suffix(1);
        const auto t = tokens(line);
        if (t.empty() || t[0][0] == '*') continue;
        const std::string_view first = t[0];
        // A section name starts a section unless the line is indented and
        // carries three or more fields: then it is a data line whose first
        // field has that name (e.g. an RHS vector called RHS).
        const bool indented = line[0] == ' ' || line[0] == '\t';
        const bool data = indented && t.size() >= 3 && section != Section::None;
        if (!data && first == "NAME") {
            saw_name = true;
            if (t.size() > 1) model.name = std::string(t[1]);
            section = Section::None;
        } else if (!data && first == "ROWS") {
            saw_rows = true;
            section = Section::Rows;
        } else if (!data && first == "COLUMNS") {
            saw_columns = true;
            section = Section::Columns;
        } else if (!data && first == "RHS") {
            saw_rhs = true;
            section = Section::Rhs;
        } else if (!data && first == "RANGES") {
            section = Section::Ranges;
        } else if (!data && first == "BOUNDS") {
            section = Section::Bounds;
        } else if (!data && first == "ENDATA") {
            saw_endata = true;
            break;
        } else if (section == Section::Rows) {
            if (t.size() < 2) {
                errors_.push_back({line_number, "missing row name"});
            } else if (first == "N") {
                objective_row = std::string(t[1]);
            } else if (first == "L" || first == "G" || first == "E") {
                row_names.emplace_back(t[1]);
                row_types.push_back(first[0]);
            } else {
                errors_.push_back({line_number, "unsupported ROWS type '" + std::string(first) + "'"});
            }
        } else if (section == Section::Columns) {
            if (t.size() >= 3 && t[1] == "'MARKER'") {
                if (t[2] == "'INTORG'") integer_block = true;
                else if (t[2] == "'INTEND'") integer_block = false;
                continue;
            }
            double value = 0.0;
            if (t.size() < 3 || !number(t[2], value)) {
                errors_.push_back({line_number, "invalid COLUMNS entry"});
                continue;
            }
            const std::string name(first);
            auto [it, inserted] = variable_index.try_emplace(name, static_cast<std::int64_t>(variables.size()));
            if (inserted) {
                variables.push_back(name);
                types.push_back(integer_block ? VariableType::Integer : VariableType::Continuous);
            }
            entries.push_back({it->second, std::string(t[1]), value});
            double second = 0.0;
            if (t.size() >= 5 && number(t[4], second)) entries.push_back({it->second, std::string(t[3]), second});
        } else if (section == Section::Rhs || section == Section::Ranges) {
            double value = 0.0;
            if (t.size() < 3 || !number(t[2], value)) {
                errors_.push_back({line_number, section == Section::Rhs ? "invalid RHS entry" : "invalid RANGES entry"});
                continue;
            }
            auto &target = section == Section::Rhs ? rhs : ranges;
            target[std::string(t[1])] = value;
            double second = 0.0;
            if (t.size() >= 5 && number(t[4], second)) target[std::string(t[3])] = second;
        } else if (section == Section::Bounds) {
            if (t.size() < 3) {
                errors_.push_back({line_number, "invalid BOUNDS entry"});
                continue;
            }
            const std::string type(first);
            if (type == "LO" || type == "UP" || type == "FX" || type == "LI" || type == "UI") {
                double value = 0.0;
                if (t.size() < 4 || !number(t[3], value)) {
                    errors_.push_back({line_number, "missing bound value"});
                    continue;
                }
                bounds.push_back({type, std::string(t[2]), value});
            } else if (type == "BV" || type == "FR" || type == "MI" || type == "PL") {
                bounds.push_back({type, std::string(t[2]), 0.0});
            } else {
                errors_.push_back({line_number, "unsupported BOUNDS type '" + type + "'"});
            }
        }
    }
    if (!saw_name) errors_.push_back({0, "missing NAME section"});
    if (!saw_rows) errors_.push_back({0, "missing ROWS section"});
    if (!saw_columns) errors_.push_back({0, "missing COLUMNS section"});
    if (!saw_rhs) errors_.push_back({0, "missing RHS section"});
    if (!saw_endata) errors_.push_back({line_number, "missing ENDATA section"});
    if (!errors_.empty()) return false;

    const auto n = static_cast<std::int64_t>(variables.size());
    const auto m = static_cast<std::int64_t>(row_names.size());
    std::unordered_map<std::string, std::int64_t> row_index;
    for (std::int64_t i = 0; i < m; ++i) row_index[row_names[static_cast<std::size_t>(i)]] = i;
    model.linear_objective.assign(static_cast<std::size_t>(n), 0.0);
    model.variable_lower.assign(static_cast<std::size_t>(n), 0.0);
    model.variable_upper.assign(static_cast<std::size_t>(n), Model::infinity());
    model.variable_types = types;
    model.constraint_lower.assign(static_cast<std::size_t>(m), -Model::infinity());
    model.constraint_upper.assign(static_cast<std::size_t>(m), Model::infinity());

    // (row, column, order of appearance, value); the last of equal pairs wins.
    struct Triple {
        std::int64_t row, col;
        std::size_t order;
        double value;
    };
    std::vector<Triple> triples;
    triples.reserve(entries.size());
    for (std::size_t e = 0; e < entries.size(); ++e) {
        const Entry &en = entries[e];
        if (en.row == objective_row) {
            model.linear_objective[static_cast<std::size_t>(en.variable)] = en.value;
            continue;
        }
        const auto it = row_index.find(en.row);
        if (it == row_index.end()) {
            errors_.push_back({0, "unknown row in COLUMNS: " + en.row});
            continue;
        }
        triples.push_back({it->second, en.variable, e, en.value});
    }
    if (!errors_.empty()) return false;

    for (std::int64_t i = 0; i < m; ++i) {
        const auto ui = static_cast<std::size_t>(i);
        const auto r = rhs.find(row_names[ui]);
        const double b = r == rhs.end() ? 0.0 : r->second;
        const char type = row_types[ui];
        double lo = type == 'L' ? -Model::infinity() : b;
        double hi = type == 'G' ? Model::infinity() : b;
        const auto g = ranges.find(row_names[ui]);
        if (g != ranges.end()) {
            const double w = std::abs(g->second);
            if (type == 'L') lo = b - w;
            else if (type == 'G') hi = b + w;
            else if (g->second > 0.0) hi = b + w;
            else lo = b - w;
        }
        model.constraint_lower[ui] = lo;
        model.constraint_upper[ui] = hi;
    }

    std::vector<char> lower_set(static_cast<std::size_t>(n), 0);
    for (const Bound &bd : bounds) {
        const auto it = variable_index.find(bd.variable);
        if (it == variable_index.end()) {
            errors_.push_back({0, "unknown variable in BOUNDS: " + bd.variable});
            continue;
        }
        const auto j = static_cast<std::size_t>(it->second);
        double &lo = model.variable_lower[j];
        double &hi = model.variable_upper[j];
        if (bd.type == "LO" || bd.type == "LI") {
            lo = bd.value;
            lower_set[j] = 1;
        } else if (bd.type == "UP" || bd.type == "UI") {
            hi = bd.value;
            if (bd.value < 0.0 && !lower_set[j] && lo == 0.0) lo = -Model::infinity();
        } else if (bd.type == "FX") {
            lo = hi = bd.value;
            lower_set[j] = 1;
        } else if (bd.type == "FR") {
            lo = -Model::infinity();
            hi = Model::infinity();
            lower_set[j] = 1;
        } else if (bd.type == "MI") {
            lo = -Model::infinity();
            lower_set[j] = 1;
        } else if (bd.type == "PL") {
            hi = Model::infinity();
        } else if (bd.type == "BV") {
            lo = 0.0;
            hi = 1.0;
            lower_set[j] = 1;
            model.variable_types[j] = VariableType::Binary;
        }
        if (bd.type == "LI" || bd.type == "UI") model.variable_types[j] = VariableType::Integer;
    }
    if (!errors_.empty()) return false;

    std::sort(triples.begin(), triples.end(), [](const Triple &a, const Triple &b) {
        if (a.row != b.row) return a.row < b.row;
        if (a.col != b.col) return a.col < b.col;
        return a.order < b.order;
    });
    auto &pat = model.constraint_matrix.pattern;
    pat.rows = m;
    pat.columns = n;
    pat.row_offsets.assign(static_cast<std::size_t>(m) + 1, 0);
    pat.column_indices.clear();
    model.constraint_matrix.values.clear();
    for (std::size_t k = 0; k < triples.size(); ++k) {
        if (k + 1 < triples.size() && triples[k + 1].row == triples[k].row && triples[k + 1].col == triples[k].col)
            continue;   // a later entry for the same pair follows
        pat.column_indices.push_back(triples[k].col);
        model.constraint_matrix.values.push_back(triples[k].value);
        ++pat.row_offsets[static_cast<std::size_t>(triples[k].row) + 1];
    }
    for (std::int64_t i = 0; i < m; ++i)
        pat.row_offsets[static_cast<std::size_t>(i) + 1] += pat.row_offsets[static_cast<std::size_t>(i)];
    return model.validate().empty();
}

const std::vector<MpsError> &MpsReader::errors() const noexcept { return errors_; }

}  // namespace vikalp
