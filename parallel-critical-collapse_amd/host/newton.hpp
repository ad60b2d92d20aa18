// newton.hpp -- C++ host side above the C-ABI: the reference's
// SimulationConfig (include/SimulationConfig.hpp:73-137) and NewtonSolver
// (include/NewtonSolver.hpp:150-162, src/NewtonSolver.cpp:260-393, 763-816)
// with the Jacobian, the shoots and the dense solve on the GPU through
// include/cc_gpu.h.  Same JSON keys in, same result dictionary out, same
// control flow (damping, divergence guard, convergence test); errors that the
// reference ends the process on are thrown as cchost::SolverError.
#pragma once
#include <string>
#include <vector>

#include "../../include/cc_gpu.h"
#include "json.hpp"

namespace cchost {

// SimulationConfig.hpp:73-137: the keys of one simulation input.
struct SimulationConfig {
    int Ntau = 0;
    double Dim = 0, XLeft = 0, XMid = 0, XRight = 0;
    double EpsNewton = 0, PrecisionNewton = 0, SlowError = 0;
    int MaxIterNewton = 0;
    bool Verbose = false, Debug = false, Converged = false;
    int NLeft = 0, NRight = 0;
    double PrecisionIRK = 0;
    int SchemeIRK = 2;
    int MaxIterIRK = 0;
    int DebugNx = 100, DebugNtau = 1;
    bool has_delta = false;
    double Delta = 0;
    std::vector<double> fc, psic, Up;

    // Throws std::invalid_argument on a missing key or an IRK scheme other
    // than 1, 2, 3 (SimulationConfig.hpp:92-98).
    static SimulationConfig from_json(const Json &j);
    static SimulationConfig load(const std::string &path);
    Json to_json() const;
};

enum class Policy { FP64, Mixed };

struct SolverOptions {
    Policy policy = Policy::Mixed;
    bool autotune = false;     // cc_autotune at the initial point before the first step
    bool use_graph = false;    // replay each Newton step from a captured hipGraph
    bool reuse_sides = false;  // FP64: reuse the base shoot's unchanged side
    bool device_loop = false;  // norm, damping and update on the device (cc_newton_run); not with
                               // Debug output or benchmark timings, which need the host loop
    bool verbose = true;       // the reference's progress lines on stdout
    std::string debug_dir = "."; // where Debug runs write jacobian_<Dim>_<i>.csv
};

struct SolverError : std::runtime_error {
    cc_status status;
    SolverError(cc_status s, const std::string &m) : std::runtime_error(m), status(s) {}
};

// common.cpp:201-215: the elements a[i] with i % n == 0, and the last one
std::vector<double> every_nth_element(const double *a, size_t size, size_t n);

class NewtonSolver {
public:
    NewtonSolver(const SimulationConfig &config, const SolverOptions &opt = {}, bool benchmark = false);
    ~NewtonSolver();
    NewtonSolver(const NewtonSolver &) = delete;
    NewtonSolver &operator=(const NewtonSolver &) = delete;

    // NewtonSolver::run (NewtonSolver.cpp:260-393).  With a benchmark dict,
    // NewtonStep_<i> and OverallTime are recorded in seconds.
    Json run(Json *benchmark_result = nullptr);

    const std::vector<double> &x() const { return x_; }
    double delta() const { return Delta_; }
    size_t iterations() const { return its_; }
    double mismatch_norm() const { return err_; }
    const std::vector<double> &mismatch_history() const { return history_; }
    cc_schedule schedule() const;

private:
    SimulationConfig cfg_;
    SolverOptions opt_;
    bool benchmark_;
    cc_ctx *ctx_ = nullptr;
    int N_ = 0, n_ = 0;
    std::vector<double> x_, Up_, psic_, fc_;
    double Delta_ = 0;
    size_t its_ = 0;
    double err_ = 0;
    std::vector<double> history_;
    Json result_;

    void check(cc_status st, const char *what) const;
    std::vector<double> pack() const;
    void unpack(const std::vector<double> &x);
    void write_final_output(size_t newton_its, double mismatch);
    void write_jacobian_csv(size_t iteration) const;
    void write_fields_json(size_t iteration) const;
    static double l2norm(const std::vector<double> &v);
};

}  // namespace cchost
