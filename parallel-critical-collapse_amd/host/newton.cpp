// newton.cpp -- SimulationConfig and NewtonSolver over the C-ABI (see newton.hpp).
#include "newton.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>

namespace cchost {

namespace {

const Json &need(const Json &j, const char *key)
{
    const Json &v = j[key];
    if (v.is_null()) throw std::invalid_argument(std::string("simulation input lacks key '") + key + "'");
    return v;
}

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

SimulationConfig SimulationConfig::from_json(const Json &j)
{
    SimulationConfig c;
    c.Ntau = (int)need(j, "Ntau").number();
    c.Dim = need(j, "Dim").number();
    c.XLeft = need(j, "XLeft").number();
    c.XMid = need(j, "XMid").number();
    c.XRight = need(j, "XRight").number();
    c.EpsNewton = need(j, "EpsNewton").number();
    c.PrecisionNewton = need(j, "PrecisionNewton").number();
    c.SlowError = need(j, "SlowError").number();
    c.MaxIterNewton = (int)need(j, "MaxIterNewton").number();
    c.Verbose = need(j, "Verbose").boolean();
    c.Debug = need(j, "Debug").boolean();
    c.Converged = need(j, "Converged").boolean();
    c.NLeft = (int)need(j, "NLeft").number();
    c.NRight = (int)need(j, "NRight").number();
    c.PrecisionIRK = need(j, "PrecisionIRK").number();
    c.MaxIterIRK = (int)need(j, "MaxIterIRK").number();
    c.SchemeIRK = (int)need(j, "SchemeIRK").number();
    if (c.SchemeIRK < 1 || c.SchemeIRK > 3) throw std::invalid_argument("Wrong IRK Scheme stage supplied!");
    if (!j["DebugNx"].is_null()) c.DebugNx = (int)j["DebugNx"].number();
    if (!j["DebugNtau"].is_null()) c.DebugNtau = (int)j["DebugNtau"].number();
    const Json &ic = j["Initial_Conditions"];
    if (!ic["Delta"].is_null()) {
        c.has_delta = true;
        c.Delta = ic["Delta"].number();
    }
    if (!ic["fc"].is_null()) c.fc = ic["fc"].numbers();
    if (!ic["psic"].is_null()) c.psic = ic["psic"].numbers();
    if (!ic["Up"].is_null()) c.Up = ic["Up"].numbers();
    return c;
}

SimulationConfig SimulationConfig::load(const std::string &path)
{
    std::ifstream f(path);
    if (!f) throw std::invalid_argument("Invalid simulation input path!");
    std::stringstream ss;
    ss << f.rdbuf();
    return from_json(Json::parse(ss.str()));
}

Json SimulationConfig::to_json() const
{
    Json j = Json::object();
    j["Ntau"] = Ntau;
    j["Dim"] = Dim;
    j["XLeft"] = XLeft;
    j["XMid"] = XMid;
    j["XRight"] = XRight;
    j["EpsNewton"] = EpsNewton;
    j["PrecisionNewton"] = PrecisionNewton;
    j["MaxIterNewton"] = MaxIterNewton;
    j["Verbose"] = Verbose;
    j["SlowError"] = SlowError;
    j["NLeft"] = NLeft;
    j["NRight"] = NRight;
    j["SchemeIRK"] = SchemeIRK;
    j["PrecisionIRK"] = PrecisionIRK;
    j["MaxIterIRK"] = MaxIterIRK;
    j["Converged"] = Converged;
    j["Debug"] = Debug;
    Json ic = Json::object();
    ic["Delta"] = has_delta ? Json(Delta) : Json();
    ic["fc"] = fc.empty() ? Json() : Json(fc);
    ic["psic"] = psic.empty() ? Json() : Json(psic);
    ic["Up"] = Up.empty() ? Json() : Json(Up);
    j["Initial_Conditions"] = ic;
    return j;
}

NewtonSolver::NewtonSolver(const SimulationConfig &config, const SolverOptions &opt, bool benchmark)
    : cfg_(config), opt_(opt), benchmark_(benchmark)
{
    N_ = cfg_.Ntau;
    n_ = 3 * N_ / 4;
    if (cfg_.SchemeIRK != 2 && opt_.policy == Policy::Mixed) {
        // the tangent-linear columns follow the two-stage scheme; other schemes
        // assemble the Jacobian from FP64 finite differences, as the reference does
        opt_.policy = Policy::FP64;
        if (opt_.verbose)
            std::cout << "SchemeIRK " << cfg_.SchemeIRK << ": FP64 finite-difference Jacobian" << std::endl;
    }
    if (!cfg_.has_delta || (int)cfg_.fc.size() != N_ || (int)cfg_.psic.size() != N_ || (int)cfg_.Up.size() != N_)
        throw std::invalid_argument("Initial_Conditions must hold Delta and fc, psic, Up of length Ntau");
    cc_problem p{};
    p.Ntau = N_;
    p.SchemeIRK = cfg_.SchemeIRK;
    p.MaxIterIRK = cfg_.MaxIterIRK;
    p.NLeft = cfg_.NLeft;
    p.NRight = cfg_.NRight;
    p.Dim = cfg_.Dim;
    p.XLeft = cfg_.XLeft;
    p.XMid = cfg_.XMid;
    p.XRight = cfg_.XRight;
    p.PrecisionIRK = cfg_.PrecisionIRK;
    p.EpsNewton = cfg_.EpsNewton;
    p.precision = opt_.policy == Policy::Mixed ? CC_MIXED : CC_FP64;
    p.reuse_sides = opt_.reuse_sides ? 1 : 0;
    const cc_status st = cc_ctx_create(&p, &ctx_);
    if (st != CC_OK) throw SolverError(st, "cc_ctx_create failed (unsupported Ntau or device error)");
    Up_ = cfg_.Up;
    psic_ = cfg_.psic;
    fc_ = cfg_.fc;
    Delta_ = cfg_.Delta;
}

NewtonSolver::~NewtonSolver()
{
    if (ctx_) cc_ctx_destroy(ctx_);
}

cc_schedule NewtonSolver::schedule() const
{
    cc_schedule s{};
    cc_get_schedule(ctx_, &s);
    return s;
}

void NewtonSolver::check(cc_status st, const char *what) const
{
    if (st == CC_OK) return;
    std::string msg = std::string(what) + ": " + cc_last_error(ctx_);
    if (st == CC_SINGULAR) msg = "LAPACKE_dgesv failed: " + std::string(cc_last_error(ctx_));
    throw SolverError(st, msg);
}

std::vector<double> NewtonSolver::pack() const
{
    std::vector<double> fields((size_t)3 * N_), x(n_);
    std::copy(Up_.begin(), Up_.end(), fields.begin());
    std::copy(psic_.begin(), psic_.end(), fields.begin() + N_);
    std::copy(fc_.begin(), fc_.end(), fields.begin() + 2 * N_);
    check(cc_pack(ctx_, fields.data(), x.data(), 1), "cc_pack");
    return x;
}

void NewtonSolver::unpack(const std::vector<double> &x)
{
    std::vector<double> fields((size_t)3 * N_);
    check(cc_unpack(ctx_, x.data(), fields.data(), 1), "cc_unpack");
    Up_.assign(fields.begin(), fields.begin() + N_);
    psic_.assign(fields.begin() + N_, fields.begin() + 2 * N_);
    fc_.assign(fields.begin() + 2 * N_, fields.end());
}

// NewtonSolver.cpp:337-341 with OutputWriter::writeMatrix (OutputWriter.cpp:57-74):
// one row per line, ", " separated, 16 significant digits.
void NewtonSolver::write_jacobian_csv(size_t iteration) const
{
    std::vector<double> J((size_t)n_ * n_);
    check(cc_last_jacobian(ctx_, J.data()), "cc_last_jacobian");
    const std::string path = opt_.debug_dir + "/jacobian_" + std::to_string(cfg_.Dim) + "_" +
                             std::to_string(iteration) + ".csv";
    std::ofstream f(path);
    if (!f) throw std::runtime_error("Could not open file for writing: " + path);
    f << std::setprecision(16);
    for (int i = 0; i < n_; ++i) {
        for (int j = 0; j < n_ - 1; ++j) f << J[(size_t)i * n_ + j] << ", ";
        f << J[(size_t)i * n_ + n_ - 1] << "\n";
    }
}

// common.cpp:201-215: elements i % n == 0 and the last one
std::vector<double> every_nth_element(const double *a, size_t size, size_t n)
{
    std::vector<double> out;
    for (size_t i = 0; i < size; ++i)
        if (i % n == 0 || i + 1 == size) out.push_back(a[i]);
    return out;
}

// NewtonSolver.cpp:287-295 with ShootingSolver.cpp:85-97: Debug field
// snapshots of the shoot at x, keyed by std::to_string(x_i), written as
// fields_<Dim>_<iteration>.json (compact, as OutputWriter::writeJsonToFile).
void NewtonSolver::write_fields_json(size_t iteration) const
{
    int32_t count = 0;
    const cc_status st0 = cc_shoot_fields(ctx_, x_.data(), cfg_.DebugNx, 0, &count, nullptr, nullptr);
    if (st0 != CC_BADARG || count < 1) check(st0, "cc_shoot_fields");
    std::vector<double> xs(count), f((size_t)count * 4 * N_);
    check(cc_shoot_fields(ctx_, x_.data(), cfg_.DebugNx, count, &count, xs.data(), f.data()), "cc_shoot_fields");
    Json out = Json::object();
    const size_t every = (size_t)std::max(1, cfg_.DebugNtau);
    for (int k = 0; k < count; ++k) {
        const double *A = f.data() + (size_t)k * 4 * N_;
        Json e = Json::object();
        e["A"] = Json(every_nth_element(A, N_, every));
        e["U"] = Json(every_nth_element(A + N_, N_, every));
        e["V"] = Json(every_nth_element(A + 2 * N_, N_, every));
        e["f"] = Json(every_nth_element(A + 3 * N_, N_, every));
        out[std::to_string(xs[k])] = e;
    }
    const std::string path =
        opt_.debug_dir + "/fields_" + std::to_string(cfg_.Dim) + "_" + std::to_string(iteration) + ".json";
    std::ofstream fo(path);
    if (!fo) throw std::runtime_error("Could not open file for writing: " + path);
    fo << out.dump(0) << "\n";
}

// NewtonSolver.cpp:763-768
double NewtonSolver::l2norm(const std::vector<double> &v)
{
    double s = 0.0;
    for (double e : v) s += e * e;
    return std::sqrt(s / (double)v.size());
}

// NewtonSolver.cpp:775-816
void NewtonSolver::write_final_output(size_t newton_its, double mismatch)
{
    Json &r = result_;
    r = Json::object();
    r["Ntau"] = cfg_.Ntau;
    r["Dim"] = cfg_.Dim;
    r["XLeft"] = cfg_.XLeft;
    r["XMid"] = cfg_.XMid;
    r["XRight"] = cfg_.XRight;
    r["EpsNewton"] = cfg_.EpsNewton;
    r["PrecisionNewton"] = cfg_.PrecisionNewton;
    r["SlowError"] = cfg_.SlowError;
    r["MaxIterNewton"] = cfg_.MaxIterNewton;
    r["Verbose"] = cfg_.Verbose;
    r["Debug"] = cfg_.Debug;
    if (cfg_.Debug) {
        r["DebugNx"] = cfg_.DebugNx;
        r["DebugNtau"] = cfg_.DebugNtau;
    }
    r["Converged"] = true;
    r["NLeft"] = cfg_.NLeft;
    r["NRight"] = cfg_.NRight;
    r["SchemeIRK"] = cfg_.SchemeIRK;
    r["PrecisionIRK"] = cfg_.PrecisionIRK;
    r["MaxIterIRK"] = cfg_.MaxIterIRK;
    r["IterNewton"] = (long)newton_its;
    r["mismatchNorm"] = mismatch;
    Json ic = Json::object();
    ic["Delta"] = Delta_;
    ic["fc"] = Json(fc_);
    ic["psic"] = Json(psic_);
    ic["Up"] = Json(Up_);
    r["Initial_Conditions"] = ic;
    if (opt_.verbose)
        std::cout << "Final result stored in simulation dictionary for dimension D = " << cfg_.Dim << "." << std::endl
                  << std::endl;
}

// NewtonSolver.cpp:260-393 (serial/OpenMP control flow).  One Newton step of
// the device path returns F(x) together with the solved update, so the
// Jacobian of the converging iteration is computed and discarded.
Json NewtonSolver::run(Json *bench)
{
    const double t_all = now_s();
    const int dslot = N_ / 2 + 2;
    auto say = [&](const std::string &s) {
        if (opt_.verbose) std::cout << s << std::endl;
    };
    history_.clear();
    if (!cfg_.Converged) {
        double fac = 1.0, err = 1.0, errOld = 1.0;
        int increase = 0;
        x_ = pack();
        x_[dslot] = Delta_;
        std::vector<double> xold = x_, F0(n_), dx(n_);
        if (opt_.autotune) {
            const double t = now_s();
            cc_schedule s{};
            check(cc_autotune(ctx_, x_.data(), CC_TUNE_THREADS | CC_TUNE_STAGES, &s, nullptr, 0, nullptr),
                  "cc_autotune");
            if (bench) (*bench)["AutotuneTime"] = now_s() - t;
        }
        if (opt_.use_graph) {
            cc_schedule s = schedule();
            s.use_graph = 1;
            check(cc_set_schedule(ctx_, &s), "cc_set_schedule");
        }
        bool converged = false;
        if (opt_.device_loop && !cfg_.Debug && !bench) {
            // the same loop with its bookkeeping on the device (cc_newton_run)
            int32_t its = 0, outcome = 0;
            double mis = 0.0;
            history_.assign((size_t)cfg_.MaxIterNewton, 0.0);
            check(cc_newton_run(ctx_, x_.data(), cfg_.PrecisionNewton, cfg_.SlowError, cfg_.MaxIterNewton, &its,
                                &mis, history_.data(), &outcome),
                  "Newton loop");
            history_.resize(outcome ? (size_t)its + 1 : (size_t)cfg_.MaxIterNewton);
            for (size_t k = 0; k < history_.size(); ++k) {
                say("Newton iteration: " + std::to_string(k + 1));
                if (opt_.verbose) std::cout << "Mismatch norm: " << history_[k] << std::endl;
            }
            if (outcome == 0) {
                std::ostringstream m;
                m << "Newton method did not converge in " << cfg_.MaxIterNewton << " iterations. For dimension: "
                  << cfg_.Dim;
                throw SolverError(CC_IRK_NOCONV, m.str());
            }
            if (outcome == 1) say("The solution has converged!\n");
            else std::cerr << "Mismatch increased – terminating Newton.\n";
            Delta_ = x_[dslot];
            std::vector<double> xv = x_;
            xv[dslot] = 0.0;
            unpack(xv);
            its_ = (size_t)its;
            err_ = mis;
            write_final_output(its_, err_);
            if (bench) (*bench)["OverallTime"] = now_s() - t_all;
            return result_;
        }
        for (size_t its = 0; its < (size_t)cfg_.MaxIterNewton; ++its) {
            const double t_step = now_s();
            say("Newton iteration: " + std::to_string(its + 1));
            errOld = err;
            if (cfg_.Debug) write_fields_json(its);
            check(cc_newton_step(ctx_, x_.data(), F0.data(), dx.data()), "Newton step");
            if (cfg_.Debug) write_jacobian_csv(its + 1);
            err = l2norm(F0);
            history_.push_back(err);
            if (opt_.verbose) std::cout << "Mismatch norm: " << err << std::endl;
            if (err < cfg_.PrecisionNewton) {
                converged = true;
                say("The solution has converged!\n");
                Delta_ = x_[dslot];
                std::vector<double> xv = x_;
                unpack(xv);                      // reads the Delta slot as zero
                its_ = its;
                err_ = err;
                write_final_output(its, err);
                break;
            }
            if (std::log10(err) > std::log10(errOld)) ++increase;
            if (increase >= 5) {
                converged = true;
                Delta_ = xold[dslot];
                xold[dslot] = 0.0;
                unpack(xold);
                its_ = its;
                err_ = errOld;
                write_final_output(its, errOld);
                std::cerr << "Mismatch increased – terminating Newton.\n";
                break;
            }
            fac = std::min(1.0, cfg_.SlowError / err);
            if (std::log10(err) < std::log10(errOld)) {
                xold = x_;
                increase = 0;
            }
            for (int i = 0; i < n_; ++i) x_[i] += fac * dx[i];
            if (bench) (*bench)["NewtonStep_" + std::to_string(its + 1)] = now_s() - t_step;
        }
        if (!converged) {
            std::ostringstream m;
            m << "Newton method did not converge in " << cfg_.MaxIterNewton << " iterations. For dimension: "
              << cfg_.Dim;
            throw SolverError(CC_IRK_NOCONV, m.str());
        }
    } else {
        x_ = pack();
        x_[dslot] = Delta_;
        std::vector<double> F(n_);
        check(cc_residual(ctx_, x_.data(), F.data()), "shoot");
        const double err = l2norm(F);
        history_.push_back(err);
        say("The solution is already marked as converged!");
        its_ = 0;
        err_ = err;
        write_final_output(0, err);
    }
    if (bench) (*bench)["OverallTime"] = now_s() - t_all;
    return result_;
}

}  // namespace cchost
