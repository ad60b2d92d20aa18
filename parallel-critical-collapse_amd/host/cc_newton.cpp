// cc_newton -- command-line driver of the GPU Newton solver, the single-run
// and benchmark modes of the reference's main (src/main.cpp:120-234) on one
// simulation JSON.
//
//   cc_newton [-i input.json] [-o output.json] [--precision mixed|fp64]
//             [--autotune] [--graph] [--reuse-sides] [--ignore-converged]
//             [--benchmark] [--benchmark-repetitions K] [--quiet] [--dry-run]
//
// Single run: the result dictionary (NewtonSolver.cpp:775-816) is written to
// output.json, by default back into input.json as the reference does
// (main.cpp:223-233).  Benchmark: K repetitions, per-repetition NewtonStep_<i>
// and OverallTime, written to benchmark_D=<Dim>_GPU_<policy>.json next to the
// input (main.cpp:136-201).  --dry-run parses the input and prints it back.
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <string>

#include "newton.hpp"

using cchost::Json;

namespace {

void write_json(const std::string &path, const Json &j)
{
    std::ofstream f(path);
    if (!f) throw std::runtime_error("cannot write " + path);
    f << j.dump(4) << "\n";
}

int usage()
{
    std::cerr << "usage: cc_newton [-i input.json] [-o output.json] [--precision mixed|fp64] [--autotune]\n"
                 "                 [--graph] [--reuse-sides] [--ignore-converged] [--benchmark]\n"
                 "                 [--benchmark-repetitions K] [--quiet] [--dry-run]\n";
    return EXIT_FAILURE;
}

}  // namespace

int main(int argc, char *argv[])
{
    std::string input = "data/simulation_4D_512.json", output;
    bool ignore_converged = false, benchmark = false, dry = false;
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
            else if (a == "--ignore-converged") ignore_converged = true;
            else if (a == "--benchmark" || a == "-b") benchmark = true;
            else if (a == "--benchmark-repetitions") {
                benchmark = true;
                reps = std::stoi(next());
            } else if (a == "--quiet") opt.verbose = false;
            else if (a == "--dry-run") dry = true;
            else if (a == "-s" || a == "--single-run") {
            } else return usage();
        }
        if (!std::filesystem::exists(input)) throw std::invalid_argument("Invalid simulation input path!");
        cchost::SimulationConfig config = cchost::SimulationConfig::load(input);
        if (ignore_converged) config.Converged = false;
        opt.debug_dir = std::filesystem::absolute(input).parent_path().string();
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
