// vikalp_pdhg_benchmark: the PDHG path's counterpart of the reference's
// benchmark runner (tools/benchmark_runner.cpp of the reference: a fixed
// case list, a console table, --csv / --json record files, exit status 1 on
// a failed case).  The reference runner times its refinery MILP family; this
// one times the seeded LP workloads of the HIP backend (the 1,000 x 3,000
// uniform LP, the 400,000 x 1,000,000 Zipf LP, the 512 block-angular
// instances as one batch, the 100,000 x 250,000 uniform LP), solved through
// vikalp::solve with BackendPreference::HIP (the batch through
// solve_lp_batch).  A record keeps the reference runner's fields (case,
// source, status, variables, constraints, nodes, milliseconds, objective,
// dual bound, relative gap) and adds the PDHG ones (iterations, FP32-phase
// iterations, device milliseconds, relative KKT error).
#include "lpgen.hpp"
#include "vikalp/solver/Solver.hpp"
#include "vikalp/solver/continuous.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

namespace {

using vikalp::Model;
using vikalp::SolveResult;
using vikalp::SolveStatus;
using vikalp::SolverOptions;

constexpr std::uint64_t kSeed = 20261203;

const char *status_name(SolveStatus s) {
    static const char *const kNames[] = {"NotSolved",    "Optimal",   "LocallyOptimal", "Feasible",
                                         "Infeasible",   "Unbounded", "IterationLimit", "NodeLimit",
                                         "TimeLimit",    "NumericalFailure", "UnsupportedModel",
                                         "InvalidModel"};
    const auto i = static_cast<std::size_t>(s);
    return i < sizeof(kNames) / sizeof(kNames[0]) ? kNames[i] : "Unknown";
}

struct Record {
    std::string name, source, status;
    std::int64_t variables = 0, constraints = 0, nodes = 0, iterations = 0, low_precision_iterations = 0;
    double milliseconds = 0.0, device_milliseconds = 0.0;
    double objective = NAN, dual_bound = NAN, relative_gap = NAN, relative_kkt = NAN;
};

Model to_model(const vikalp::gen::Lp &lp, const std::string &name) {
    Model model;
    model.name = name;
    model.linear_objective = lp.objective;
    model.variable_lower = lp.var_lower;
    model.variable_upper = lp.var_upper;
    model.variable_types.assign(static_cast<std::size_t>(lp.n), vikalp::VariableType::Continuous);
    model.constraint_lower = lp.row_lower;
    model.constraint_upper = lp.row_upper;
    model.constraint_matrix.pattern.rows = lp.m;
    model.constraint_matrix.pattern.columns = lp.n;
    model.constraint_matrix.pattern.row_offsets = lp.row_offsets;
    model.constraint_matrix.pattern.column_indices = lp.column_indices;
    model.constraint_matrix.values = lp.values;
    return model;
}

Record solve_one(const std::string &name, const Model &model, const SolverOptions &options) {
    Record r;
    r.name = name;
    r.source = "generated seed " + std::to_string(kSeed);
    r.variables = model.num_variables();
    r.constraints = model.num_constraints();
    const auto t0 = std::chrono::steady_clock::now();
    const SolveResult res = vikalp::solve(model, options);
    r.milliseconds = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    r.status = status_name(res.status);
    r.iterations = res.iterations;
    r.low_precision_iterations = res.low_precision_iterations;
    r.device_milliseconds = res.device_seconds * 1e3;
    r.objective = res.objective_value;
    r.dual_bound = res.kkt_dual_objective;
    r.relative_gap = res.kkt_relative_gap;
    r.relative_kkt = res.kkt_relative;
    return r;
}

Record solve_batch(const std::string &name, const std::vector<Model> &models, const SolverOptions &options) {
    Record r;
    r.name = name;
    r.source = "generated seed " + std::to_string(kSeed);
    std::vector<const Model *> ptrs;
    for (const Model &m : models) {
        ptrs.push_back(&m);
        r.variables += m.num_variables();
        r.constraints += m.num_constraints();
    }
    r.nodes = static_cast<std::int64_t>(models.size());
    const auto t0 = std::chrono::steady_clock::now();
    const std::vector<SolveResult> res = vikalp::solver::solve_lp_batch(ptrs, options);
    r.milliseconds = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::int64_t optimal = 0, infeasible = 0;
    double worst_kkt = 0.0;
    for (const SolveResult &s : res) {
        optimal += s.status == SolveStatus::Optimal;
        infeasible += s.status == SolveStatus::Infeasible;
        r.iterations = std::max<std::int64_t>(r.iterations, s.iterations);
        r.device_milliseconds = std::max(r.device_milliseconds, s.device_seconds * 1e3);
        if (s.status == SolveStatus::Optimal) worst_kkt = std::max(worst_kkt, s.kkt_relative);
    }
    r.status = std::to_string(optimal) + " Optimal, " + std::to_string(infeasible) + " Infeasible of " +
               std::to_string(models.size());
    r.relative_kkt = worst_kkt;
    return r;
}

std::string csv_quote(const std::string &value) {
    std::string out = "\"";
    for (char ch : value) out += ch == '"' ? std::string("\"\"") : std::string(1, ch);
    return out + "\"";
}

std::string number(double v) {
    if (!std::isfinite(v)) return "null";
    std::ostringstream s;
    s << std::setprecision(17) << v;
    return s.str();
}

void write_csv(const std::filesystem::path &path, const std::vector<Record> &records) {
    if (path.has_parent_path()) std::filesystem::create_directories(path.parent_path());
    std::ofstream out(path);
    out << "case,source,status,variables,constraints,nodes,milliseconds,objective,dual_bound,relative_gap,"
           "iterations,low_precision_iterations,device_milliseconds,relative_kkt\n";
    for (const Record &r : records)
        out << csv_quote(r.name) << ',' << csv_quote(r.source) << ',' << csv_quote(r.status) << ',' << r.variables
            << ',' << r.constraints << ',' << r.nodes << ',' << number(r.milliseconds) << ',' << number(r.objective)
            << ',' << number(r.dual_bound) << ',' << number(r.relative_gap) << ',' << r.iterations << ','
            << r.low_precision_iterations << ',' << number(r.device_milliseconds) << ',' << number(r.relative_kkt)
            << '\n';
}

void write_json(const std::filesystem::path &path, const std::vector<Record> &records) {
    if (path.has_parent_path()) std::filesystem::create_directories(path.parent_path());
    std::ofstream out(path);
    out << "[\n";
    for (std::size_t i = 0; i < records.size(); ++i) {
        const Record &r = records[i];
        out << "  {\"case\":\"" << r.name << "\",\"source\":\"" << r.source << "\",\"status\":\"" << r.status
            << "\",\"variables\":" << r.variables << ",\"constraints\":" << r.constraints << ",\"nodes\":" << r.nodes
            << ",\"milliseconds\":" << number(r.milliseconds) << ",\"objective\":" << number(r.objective)
            << ",\"dual_bound\":" << number(r.dual_bound) << ",\"relative_gap\":" << number(r.relative_gap)
            << ",\"iterations\":" << r.iterations << ",\"low_precision_iterations\":" << r.low_precision_iterations
            << ",\"device_milliseconds\":" << number(r.device_milliseconds)
            << ",\"relative_kkt\":" << number(r.relative_kkt) << "}" << (i + 1 == records.size() ? "\n" : ",\n");
    }
    out << "]\n";
}

const char *kUsage =
    "usage: vikalp_pdhg_benchmark [--csv PATH] [--json PATH] [--backend hip|auto]\n"
    "                             [--policy reference|adaptive|halpern] [--precision fp64|mixed]\n"
    "                             [--tolerance REL_KKT] [--iteration-limit N] [--cases NAME,...]\n"
    "cases: uniform_1000x3000 zipf_400000x1000000 block_angular_512 uniform_100000x250000 (default: all)\n";

}  // namespace

int main(int argc, char **argv) {
    std::string csv_path, json_path, backend = "hip", policy = "halpern", precision = "fp64", cases;
    double tolerance = 1e-6;
    long long iteration_limit = 200000;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        const bool has_value = i + 1 < argc;
        if (a == "--help") {
            std::cout << kUsage;
            return 0;
        } else if (has_value && a == "--csv") {
            csv_path = argv[++i];
        } else if (has_value && a == "--json") {
            json_path = argv[++i];
        } else if (has_value && a == "--backend") {
            backend = argv[++i];
        } else if (has_value && a == "--policy") {
            policy = argv[++i];
        } else if (has_value && a == "--precision") {
            precision = argv[++i];
        } else if (has_value && a == "--tolerance") {
            tolerance = std::stod(argv[++i]);
        } else if (has_value && a == "--iteration-limit") {
            iteration_limit = std::stoll(argv[++i]);
        } else if (has_value && a == "--cases") {
            cases = argv[++i];
        } else {
            std::cerr << kUsage;
            return 2;
        }
    }
    if (backend != "hip" && backend != "auto") {
        std::cerr << "this build carries the HIP backend only (--backend hip or auto)\n";
        return 2;
    }
    if ((policy != "reference" && policy != "adaptive" && policy != "halpern") ||
        (precision != "fp64" && precision != "mixed") || !(tolerance >= 0.0) || iteration_limit < 0) {
        std::cerr << kUsage;
        return 2;
    }
    SolverOptions options;
    options.backend = vikalp::BackendPreference::HIP;
    options.pdhg_policy = policy == "reference" ? vikalp::PdhgPolicy::Reference
                          : policy == "adaptive"  ? vikalp::PdhgPolicy::Adaptive
                                                  : vikalp::PdhgPolicy::Halpern;
    options.pdhg_precision = precision == "mixed" ? vikalp::PdhgPrecision::Mixed : vikalp::PdhgPrecision::FP64;
    options.relative_kkt_tolerance = policy == "reference" ? 0.0 : tolerance;
    options.mixed_phase_tolerance = 0.0;   // automatic (SolverOptions)
    options.iteration_limit = iteration_limit;
    auto wanted = [&](const std::string &name) {
        return cases.empty() || ("," + cases + ",").find("," + name + ",") != std::string::npos;
    };

    std::vector<Record> records;
    vikalp::gen::Lp lp;
    if (wanted("uniform_1000x3000")) {
        vikalp::gen::make_uniform(1000, 3000, 5, kSeed, lp);
        records.push_back(solve_one("uniform_1000x3000", to_model(lp, "uniform_1000x3000"), options));
    }
    if (wanted("zipf_400000x1000000")) {
        vikalp::gen::make_zipf(400000, 1000000, kSeed, lp);
        records.push_back(solve_one("zipf_400000x1000000", to_model(lp, "zipf_400000x1000000"), options));
    }
    if (wanted("block_angular_512")) {
        // The benchmark batch: instance i has T_i ~ U{4..48} periods and seed
        // kSeed + i.
        std::vector<Model> models;
        for (std::int64_t i = 0; i < 512; ++i) {
            vikalp::gen::make_block(vikalp::gen::block_periods(kSeed, i), kSeed + static_cast<std::uint64_t>(i), lp);
            models.push_back(to_model(lp, "block_" + std::to_string(i)));
        }
        records.push_back(solve_batch("block_angular_512", models, options));
    }
    if (wanted("uniform_100000x250000")) {
        vikalp::gen::make_uniform(100000, 250000, 6, kSeed, lp);
        records.push_back(solve_one("uniform_100000x250000", to_model(lp, "uniform_100000x250000"), options));
    }

    std::cout << "case | source | vars | rows | nodes | status | iterations | ms | objective | rel_kkt\n";
    for (const Record &r : records)
        std::cout << r.name << " | " << r.source << " | " << r.variables << " | " << r.constraints << " | "
                  << r.nodes << " | " << r.status << " | " << r.iterations << " | " << std::setprecision(6)
                  << r.milliseconds << " | " << std::setprecision(10) << r.objective << " | "
                  << std::setprecision(3) << r.relative_kkt << '\n';
    if (!csv_path.empty()) write_csv(csv_path, records);
    if (!json_path.empty()) write_json(json_path, records);
    const bool failed = std::any_of(records.begin(), records.end(), [](const Record &r) {
        return r.status == "NumericalFailure" || r.status == "InvalidModel" || r.status == "UnsupportedModel";
    });
    return failed ? 1 : 0;
}
