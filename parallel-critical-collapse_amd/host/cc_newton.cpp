// cc_newton -- command-line driver of the GPU Newton solver: the single-run,
// multiple-run and benchmark modes of the reference's main (src/main.cpp:19-378).
//
//   cc_newton [-s | -m] [-r] [-i input.json] [-o output.json]
//             [--precision mixed|fp64] [--autotune] [--graph] [--reuse-sides]
//             [--device-loop] [--ignore-converged] [--benchmark]
//             [--benchmark-repetitions K] [--quiet] [--dry-run]
//
// Single run (-s, default): the result dictionary (NewtonSolver.cpp:775-816)
// is written to output.json, by default back into input.json as the reference
// does (main.cpp:223-233).  Multiple run (-m, -r for ascending order): the
// input is a suite keyed by dimension; each dimension is solved in the suite's
// order with initial data from the file, copied or extrapolated in D, and the
// whole suite is rewritten after each one (main.cpp:236-378).  Benchmark: K
// repetitions of one simulation, per-repetition NewtonStep_<i> and
// OverallTime, written to benchmark_D=<Dim>_GPU_<policy>.json next to the
// input (main.cpp:136-201).  --device-loop runs the Newton loop's norm, damping
// and update on the device (cc_newton_run; the host loop stays in use for
// Debug output and benchmarks).  --dry-run parses the input and prints it back; with
// -m it prints the run order and where each dimension's initial data would come
// from, and -o receives the suite with the initial data filled in.
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <string>

#include "newton.hpp"
#include "suite.hpp"

using cchost::Json;

namespace {

void write_json(const std::string &path, const Json &j, int indent = 4)
{
    std::ofstream f(path);
    if (!f) throw std::runtime_error("cannot write " + path);
    f << j.dump(indent) << "\n";
}

int usage()
{
    std::cerr << "usage: cc_newton [-s | -m] [-r] [-i input.json] [-o output.json] [--precision mixed|fp64]\n"
                 "                 [--autotune] [--graph] [--reuse-sides] [--device-loop] [--ignore-converged]\n"
                 "                 [--benchmark] [--benchmark-repetitions K] [--quiet] [--dry-run]\n";
    return EXIT_FAILURE;
}

const char *source_name(cchost::InitialData d)
{
    switch (d) {
    case cchost::InitialData::File: return "from the input";
    case cchost::InitialData::Copied: return "copied from the previous dimension";
    default: return "extrapolated (quadratic in D) from the previous three dimensions";
    }
}

// main.cpp:236-357
int multiple_run(const std::string &input, const std::string &output, bool reversed, bool ignore_converged,
                 bool dry, const cchost::SolverOptions &opt)
{
    cchost::SimulationSuite suite = cchost::SimulationSuite::load(input, reversed, ignore_converged);
    const std::string path = output.empty() ? input : output;
    std::cout << "Starting multi run for " << suite.dims.size() << " distinct dimensions.\n\n";
    Json planned = Json::object();
    for (size_t i = 0; i < suite.dims.size(); ++i) {
        const std::string &dim = suite.dims[i];
        cchost::SimulationConfig config = suite.generate(dim);
        std::cout << "Simulation " << i + 1 << "/" << suite.dims.size() << ":\n";
        std::cout << "Starting simulation for D=" << dim << ".\n\n";
        if (dry) {
            try {
                const cchost::InitialData src = cchost::prepare_initial_data(suite, i, config);
                std::cout << "initial data " << source_name(src) << "\n\n";
                planned[dim] = config.to_json();
            } catch (const std::runtime_error &e) {
                std::cout << "no initial data yet: " << e.what() << "\n\n";
                planned[dim] = suite.dict[dim];
            }
            continue;
        }
        const cchost::InitialData src = cchost::prepare_initial_data(suite, i, config);
        if (opt.verbose) std::cout << "initial data " << source_name(src) << "\n\n";
        if (ignore_converged) config.Converged = false;
        cchost::NewtonSolver solver(config, opt, false);
        suite.dict[dim] = solver.run();
        std::cout << "Result stored in file: " << path << "\n\n";
        write_json(path, suite.dict, 0);
    }
    if (dry && !output.empty()) write_json(output, planned, 4);
    return 0;
}

}  // namespace

int main(int argc, char *argv[])
{
    std::string input = "data/simulation_4D_512.json", output;
    bool ignore_converged = false, benchmark = false, dry = false, single = true, reversed = false;
    int reps = 3;
    cchost::SolverOptions opt;
    try {
        for (int i = 1; i < argc; ++i) {
            const std::string a = argv[i];
            auto next = [&]() -> std::string {
                if (i + 1 >= argc) throw std::invalid_argument("missing value after " + a);
                return argv[++i];
            };
            if (a == "-i" || a == "--input-path") input = next();
            else if (a == "-o" || a == "--output-path") output = next();
            else if (a == "--precision") {
                const std::string p = next();
                if (p == "mixed") opt.policy = cchost::Policy::Mixed;
                else if (p == "fp64") opt.policy = cchost::Policy::FP64;
                else throw std::invalid_argument("--precision must be mixed or fp64");
            } else if (a == "--autotune") opt.autotune = true;
            else if (a == "--graph") opt.use_graph = true;
            else if (a == "--reuse-sides") opt.reuse_sides = true;
            else if (a == "--device-loop") opt.device_loop = true;
            else if (a == "--ignore-converged") ignore_converged = true;
            else if (a == "--benchmark" || a == "-b") benchmark = true;
            else if (a == "--benchmark-repetitions") {
                benchmark = true;
                reps = std::stoi(next());
            } else if (a == "--quiet") opt.verbose = false;
            else if (a == "--dry-run") dry = true;
            else if (a == "-s" || a == "--single-run") single = true;
            else if (a == "-m" || a == "--multiple-run") single = false;
            else if (a == "-r" || a == "--reversed-order") reversed = true;
            else return usage();
        }
        if (!std::filesystem::exists(input)) throw std::invalid_argument("Invalid simulation input path!");
        opt.debug_dir = std::filesystem::absolute(input).parent_path().string();
        if (!single && !benchmark) return multiple_run(input, output, reversed, ignore_converged, dry, opt);
        cchost::SimulationConfig config = cchost::SimulationConfig::load(input);
        if (ignore_converged) config.Converged = false;
        if (dry) {
            std::cout << config.to_json().dump(4) << std::endl;
            return 0;
        }
        const std::string policy = opt.policy == cchost::Policy::Mixed ? "mixed" : "fp64";
        if (benchmark) {
            Json b = Json::object();
            b["Dim"] = config.Dim;
            b["Ntau"] = config.Ntau;
            b["XLeft"] = config.XLeft;
            b["XMid"] = config.XMid;
            b["XRight"] = config.XRight;
            b["PrecisionNewton"] = config.PrecisionNewton;
            b["NLeft"] = config.NLeft;
            b["NRight"] = config.NRight;
            b["Repetitions"] = reps;
            b["Kind"] = "GPU";
            b["Policy"] = policy;
            std::cout << "Starting benchmark run for D=" << config.Dim << ".\n\n";
            for (int r = 0; r < reps; ++r) {
                std::cout << "Repetition " << r + 1 << "/" << reps << "\n\n";
                cchost::NewtonSolver solver(config, opt, true);
                Json &slot = b[std::to_string(r)];
                slot = Json::object();
                solver.run(&slot);
                slot["IterNewton"] = (long)solver.iterations();
                slot["mismatchNorm"] = solver.mismatch_norm();
                slot["Delta"] = solver.delta();
            }
            const std::filesystem::path dir = std::filesystem::absolute(input).parent_path();
            const std::string path = output.empty()
                                         ? (dir / ("benchmark_D=" + std::to_string(config.Dim) + "_GPU_" + policy +
                                                   ".json")).string()
                                         : output;
            std::cout << "Benchmark result stored in file: " << path << "\n\n";
            write_json(path, b);
        } else {
            std::cout << "Starting single run for D=" << config.Dim << ".\n\n";
            cchost::NewtonSolver solver(config, opt, false);
            const Json result = solver.run();
            const std::string path = output.empty() ? input : output;
            std::cout << "Result stored in file: " << path << "\n\n";
            write_json(path, result);
        }
    } catch (const cchost::SolverError &e) {
        std::cerr << "ERROR: " << e.what() << std::endl;
        return EXIT_FAILURE;
    } catch (const std::exception &e) {
        std::cerr << "Error: " << e.what() << std::endl;
        return EXIT_FAILURE;
    }
    return 0;
}
