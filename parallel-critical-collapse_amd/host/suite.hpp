// suite.hpp -- the reference's multi-dimension run: SimulationSuite
// (include/SimulationConfig.hpp:213-315) orders the dimension keys of a suite
// JSON, and main.cpp:236-378 gives each dimension without initial data the
// previous converged solution (second and third dimension) or a quadratic
// extrapolation in D of the last three (from the fourth on; the 3x3 least
// squares of common.cpp:150-194), then runs one NewtonSolver per dimension and
// stores its result dictionary under the dimension key.
#pragma once
#include <array>
#include <string>
#include <vector>

#include "json.hpp"
#include "newton.hpp"

namespace cchost {

struct SimulationSuite {
    Json dict;                      // the suite: dimension key -> simulation dictionary
    std::vector<std::string> dims;  // run order

    // SimulationConfig.hpp:232-297.  Keys are ordered as strings, ascending
    // when reversed, else descending; unless ignore_converged, converged
    // dimensions are left out except for up to three of them (the last three
    // in that order) run first as seeds.
    SimulationSuite(const Json &suite, bool reversed, bool ignore_converged);
    static SimulationSuite load(const std::string &path, bool reversed, bool ignore_converged);
    SimulationConfig generate(const std::string &dim) const;   // throws std::out_of_range
};

// Coefficients {a, b, c} of y = a x^2 + b x + c through three points: the
// solution of [x_i^2, x_i, 1] (a, b, c)^T = y_i that dgels returns for the
// square system (common.cpp:150-194), here by Gaussian elimination with
// partial pivoting.  Throws std::runtime_error for coincident abscissae.
std::array<double, 3> fit_quadratic(const std::array<double, 3> &x, const std::array<double, 3> &y);

enum class InitialData { File, Copied, Extrapolated };

// main.cpp:270-355: completes config (dimension number i of the run order)
// when its fc, Up or psic is missing; throws std::runtime_error with the
// reference's messages when no converged predecessor can supply them.
InitialData prepare_initial_data(const SimulationSuite &suite, size_t i, SimulationConfig &config);

}  // namespace cchost
