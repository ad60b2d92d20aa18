// suite.cpp -- multi-dimension run order and initial data (see suite.hpp).
#include "suite.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace cchost {

SimulationSuite::SimulationSuite(const Json &suite, bool reversed, bool ignore_converged) : dict(suite)
{
    std::vector<std::string> converged;
    for (const auto &kv : dict.items()) {
        if (!ignore_converged && kv.second["Converged"].boolean()) converged.push_back(kv.first);
        else dims.push_back(kv.first);
    }
    auto ascending = [](std::vector<std::string> &v) { std::sort(v.begin(), v.end()); };
    auto descending = [](std::vector<std::string> &v) {
        std::sort(v.begin(), v.end(), [](const std::string &a, const std::string &b) { return a > b; });
    };
    if (reversed) ascending(dims);
    else descending(dims);
    if (!ignore_converged && !converged.empty()) {
        if (reversed) ascending(converged);
        else descending(converged);
        const size_t first = converged.size() > 3 ? converged.size() - 3 : 0;
        dims.insert(dims.begin(), converged.begin() + (long)first, converged.end());
    }
}

SimulationSuite SimulationSuite::load(const std::string &path, bool reversed, bool ignore_converged)
{
    std::ifstream f(path);
    if (!f) throw std::invalid_argument("File does not exist! " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return SimulationSuite(Json::parse(ss.str()), reversed, ignore_converged);
}

SimulationConfig SimulationSuite::generate(const std::string &dim) const
{
    const Json &j = dict[dim];
    if (j.is_null()) throw std::out_of_range("Simulation key '" + dim + "' not found in input dictionary.");
    return SimulationConfig::from_json(j);
}

std::array<double, 3> fit_quadratic(const std::array<double, 3> &x, const std::array<double, 3> &y)
{
    double A[3][4];
    for (int i = 0; i < 3; ++i) {
        A[i][0] = x[i] * x[i];
        A[i][1] = x[i];
        A[i][2] = 1.0;
        A[i][3] = y[i];
    }
    for (int c = 0; c < 3; ++c) {
        int p = c;
        for (int r = c + 1; r < 3; ++r)
            if (std::fabs(A[r][c]) > std::fabs(A[p][c])) p = r;
        if (!(std::fabs(A[p][c]) > 0.0))
            throw std::runtime_error("quadratic extrapolation: the three dimensions are not distinct");
        if (p != c)
            for (int k = 0; k < 4; ++k) std::swap(A[c][k], A[p][k]);
        for (int r = c + 1; r < 3; ++r) {
            const double f = A[r][c] / A[c][c];
            for (int k = c; k < 4; ++k) A[r][k] -= f * A[c][k];
        }
    }
    std::array<double, 3> s{};
    for (int c = 2; c >= 0; --c) {
        double v = A[c][3];
        for (int k = c + 1; k < 3; ++k) v -= A[c][k] * s[k];
        s[c] = v / A[c][c];
    }
    return s;
}

InitialData prepare_initial_data(const SimulationSuite &suite, size_t i, SimulationConfig &config)
{
    const std::string &dim = suite.dims[i];
    if (!config.fc.empty() && !config.Up.empty() && !config.psic.empty()) return InitialData::File;
    if (i == 0) throw std::runtime_error("No converged solution as initial conditions available for " + dim + "!");
    if (i < 3) {
        const std::string &prev = suite.dims[i - 1];
        const SimulationConfig p = suite.generate(prev);
        if (!p.Converged)
            throw std::runtime_error("Previous dimension " + prev +
                                     " is not converged, hence no initial data for next simulation!");
        config.has_delta = p.has_delta;
        config.Delta = p.Delta;
        config.fc = p.fc;
        config.Up = p.Up;
        config.psic = p.psic;
        return InitialData::Copied;
    }
    const SimulationConfig p1 = suite.generate(suite.dims[i - 1]), p2 = suite.generate(suite.dims[i - 2]),
                           p3 = suite.generate(suite.dims[i - 3]);
    if (!p1.Converged || !p2.Converged || !p3.Converged)
        throw std::runtime_error("Previous dimensions not converged, hence no initial data for next simulation!");
    const std::array<double, 3> D = {std::stod(suite.dims[i - 1]), std::stod(suite.dims[i - 2]),
                                     std::stod(suite.dims[i - 3])};
    const double d = std::stod(dim);
    auto at = [&](const std::array<double, 3> &y) {
        const std::array<double, 3> c = fit_quadratic(D, y);
        return c[0] * d * d + c[1] * d + c[2];
    };
    const size_t len = p1.fc.size();
    if (p2.fc.size() != len || p3.fc.size() != len || p1.Up.size() != len || p2.Up.size() != len ||
        p3.Up.size() != len || p1.psic.size() != len || p2.psic.size() != len || p3.psic.size() != len)
        throw std::runtime_error("Previous dimensions hold initial data of different lengths");
    std::vector<double> fc(len), Up(len), psic(len);
    for (size_t j = 0; j < len; ++j) {
        fc[j] = at({p1.fc[j], p2.fc[j], p3.fc[j]});
        Up[j] = at({p1.Up[j], p2.Up[j], p3.Up[j]});
        psic[j] = at({p1.psic[j], p2.psic[j], p3.psic[j]});
    }
    config.has_delta = true;
    config.Delta = at({p1.Delta, p2.Delta, p3.Delta});
    config.fc = std::move(fc);
    config.Up = std::move(Up);
    config.psic = std::move(psic);
    return InitialData::Extrapolated;
}

}  // namespace cchost
