// cc_layers -- command-line driver of the lower-layer classes (layers.hpp):
// runs a list of SpectralTransformer, ODEStepper and ShootingSolver calls
// from a JSON request on the GPU and writes their results as JSON.
//
//   cc_layers request.json result.json
//
// Request: {"Ntau", "Dim", "PrecisionIRK", "MaxIterIRK", "SchemeIRK" (1, 2,
// 3), "period" (the SpectralTransformer's), "calls": [...]}.  Complex vectors
// are flat lists of interleaved (re, im) values.  Calls, by "method":
//   forwardFFT / backwardFFT          "in" complex -> "out" complex
//   forwardFFT_real                   "in" real -> "out" complex
//   backwardFFT_real                  "in" complex -> "out" real
//   differentiate                     "in", optional "period" -> "out"
//   lamIntegrate                      "in", "lambda" [re, im], optional "period" -> "out"
//   halveModes / doubleModes          "in" -> "out"
//   solveInhom                        "f", "g", optional "period" -> "out"
//   integrate                         "Y", "Xin", "Xout", "maxIts" -> "Yout", "converged", "itsReached"
//   shoot                             "YLeft", "YRight", "grid", "iLeft", "iRight", "iMid", optional
//                                     "Debug", "DebugNx", "DebugNtau" -> "YLeft", "YRight", "mismatch",
//                                     "itsLeft", "itsRight" and with Debug "fields"
// A call that throws gives {"error": message, "status": cc_status} (invalid
// arguments: CC_BADARG) and the next call runs.  The classes are created on
// first use, one of each.
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>

#include "layers.hpp"

using namespace cchost;

namespace {

vec_complex complex_of(const Json &j)
{
    const std::vector<double> v = j.numbers();
    if (v.size() % 2) throw std::invalid_argument("complex vectors are interleaved (re, im) pairs");
    vec_complex z(v.size() / 2);
    for (size_t i = 0; i < z.size(); ++i) z[i] = complex_t(v[2 * i], v[2 * i + 1]);
    return z;
}

Json json_of(const vec_complex &z)
{
    std::vector<double> v(2 * z.size());
    for (size_t i = 0; i < z.size(); ++i) {
        v[2 * i] = z[i].real();
        v[2 * i + 1] = z[i].imag();
    }
    return Json(v);
}

double num(const Json &j, const char *key)
{
    const Json &v = j[key];
    if (v.is_null()) throw std::invalid_argument(std::string("missing key '") + key + "'");
    return v.number();
}

double num_or(const Json &j, const char *key, double d) { return j[key].is_null() ? d : j[key].number(); }

Scheme scheme_of(int s)
{
    if (s < 1 || s > 3) throw std::invalid_argument("Wrong IRK Scheme stage supplied!");
    return static_cast<Scheme>(s - 1);
}

struct Layers {
    int Ntau;
    double Dim, PrecisionIRK, period;
    int MaxIterIRK, SchemeIRK;
    std::unique_ptr<SpectralTransformer> spec;
    std::unique_ptr<ODEStepper> step;
    std::unique_ptr<ShootingSolver> shoot;

    SpectralTransformer &S()
    {
        if (!spec) spec = std::make_unique<SpectralTransformer>((size_t)Ntau, period);
        return *spec;
    }
    ODEStepper &O()
    {
        if (!step) step = std::make_unique<ODEStepper>(Ntau, Dim, PrecisionIRK, scheme_of(SchemeIRK));
        return *step;
    }
    ShootingSolver &H()
    {
        if (!shoot) shoot = std::make_unique<ShootingSolver>(Ntau, Dim, PrecisionIRK, MaxIterIRK, scheme_of(SchemeIRK));
        return *shoot;
    }

    Json call(const Json &c)
    {
        const std::string m = c["method"].string();
        Json r = Json::object();
        if (m == "forwardFFT" || m == "backwardFFT") {
            vec_complex out;
            if (m == "forwardFFT") S().forwardFFT(complex_of(c["in"]), out);
            else S().backwardFFT(complex_of(c["in"]), out);
            r["out"] = json_of(out);
        } else if (m == "forwardFFT_real") {
            vec_complex out;
            S().forwardFFT(c["in"].numbers(), out);
            r["out"] = json_of(out);
        } else if (m == "backwardFFT_real") {
            vec_real out;
            S().backwardFFT(complex_of(c["in"]), out);
            r["out"] = Json(out);
        } else if (m == "differentiate") {
            vec_complex out;
            S().differentiate(complex_of(c["in"]), out, num_or(c, "period", 0.0));
            r["out"] = json_of(out);
        } else if (m == "lamIntegrate") {
            const std::vector<double> l = c["lambda"].is_null() ? std::vector<double>{0.0, 0.0} : c["lambda"].numbers();
            if (l.size() != 2) throw std::invalid_argument("lambda is [re, im]");
            vec_complex out;
            S().lamIntegrate(complex_of(c["in"]), out, complex_t(l[0], l[1]), num_or(c, "period", 0.0));
            r["out"] = json_of(out);
        } else if (m == "halveModes" || m == "doubleModes") {
            vec_complex out;
            if (m == "halveModes") S().halveModes(complex_of(c["in"]), out);
            else S().doubleModes(complex_of(c["in"]), out);
            r["out"] = json_of(out);
        } else if (m == "solveInhom") {
            vec_real x;
            S().solveInhom(c["f"].numbers(), c["g"].numbers(), x, num_or(c, "period", 0.0));
            r["out"] = Json(x);
        } else if (m == "integrate") {
            vec_complex Y = complex_of(c["Y"]), Yout;
            bool converged = false;
            int its = 0;
            O().integrate(Y, Yout, num(c, "Xin"), num(c, "Xout"), converged, its, (int)num(c, "maxIts"));
            r["Yout"] = json_of(Yout);
            r["converged"] = converged;
            r["itsReached"] = its;
        } else if (m == "shoot") {
            vec_complex YL = complex_of(c["YLeft"]), YR = complex_of(c["YRight"]), mis;
            SimulationConfig cfg;
            cfg.Debug = !c["Debug"].is_null() && c["Debug"].boolean();
            cfg.DebugNx = (int)num_or(c, "DebugNx", 100);
            cfg.DebugNtau = (int)num_or(c, "DebugNtau", 1);
            Json fields = Json::object();
            const auto idx = [&](const char *k) {
                const double v = num(c, k);
                if (!(v >= 0)) throw std::invalid_argument(std::string(k) + " must be a grid index");
                return (size_t)v;
            };
            H().shoot(YL, YR, c["grid"].numbers(), idx("iLeft"), idx("iRight"), idx("iMid"), mis, &cfg, &fields);
            r["YLeft"] = json_of(YL);
            r["YRight"] = json_of(YR);
            r["mismatch"] = json_of(mis);
            r["itsLeft"] = (long long)H().iterationsLeft();
            r["itsRight"] = (long long)H().iterationsRight();
            if (cfg.Debug) r["fields"] = fields;
        } else {
            throw std::invalid_argument("unknown method '" + m + "'");
        }
        return r;
    }
};

}  // namespace

int main(int argc, char **argv)
{
    if (argc != 3) {
        std::cerr << "usage: cc_layers request.json result.json\n";
        return EXIT_FAILURE;
    }
    Json req;
    try {
        std::ifstream f(argv[1]);
        if (!f) throw std::runtime_error(std::string("cannot read ") + argv[1]);
        std::stringstream ss;
        ss << f.rdbuf();
        req = Json::parse(ss.str());
    } catch (const std::exception &e) {
        std::cerr << "cc_layers: " << e.what() << "\n";
        return EXIT_FAILURE;
    }
    Layers L;
    try {
        L.Ntau = (int)num(req, "Ntau");
        L.Dim = num_or(req, "Dim", 4.0);
        L.PrecisionIRK = num_or(req, "PrecisionIRK", 1e-15);
        L.MaxIterIRK = (int)num_or(req, "MaxIterIRK", 100);
        L.SchemeIRK = (int)num_or(req, "SchemeIRK", 2);
        L.period = num_or(req, "period", 1.0);
        scheme_of(L.SchemeIRK);
        if (!req["calls"].is_array()) throw std::invalid_argument("missing list 'calls'");
    } catch (const std::exception &e) {
        std::cerr << "cc_layers: " << e.what() << "\n";
        return EXIT_FAILURE;
    }
    Json results = Json::array();
    for (const Json &c : req["calls"].elements()) {
        try {
            results.push_back(L.call(c));
        } catch (const SolverError &e) {
            Json r = Json::object();
            r["error"] = e.what();
            r["status"] = (int)e.status;
            results.push_back(r);
        } catch (const std::exception &e) {
            Json r = Json::object();
            r["error"] = e.what();
            r["status"] = (int)CC_BADARG;
            results.push_back(r);
        }
    }
    Json out = Json::object();
    out["results"] = results;
    std::ofstream fo(argv[2]);
    if (!fo) {
        std::cerr << "cc_layers: cannot write " << argv[2] << "\n";
        return EXIT_FAILURE;
    }
    fo << out.dump(0) << "\n";
    return EXIT_SUCCESS;
}
