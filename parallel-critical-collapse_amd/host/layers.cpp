// layers.cpp -- SpectralTransformer, ODEStepper and ShootingSolver over the
// C-ABI (see layers.hpp).
#include "layers.hpp"

#include <algorithm>
#include <sstream>
#include <stdexcept>
#include <string>

namespace cchost {

namespace {

int stages_of(Scheme s)
{
    switch (s) {
    case Scheme::IRK1: return 1;
    case Scheme::IRK2: return 2;
    case Scheme::IRK3: return 3;
    }
    throw std::invalid_argument("Wrong IRK Scheme stage supplied!");
}

void need_size(size_t got, size_t want, const char *what)
{
    if (got != want)
        throw std::invalid_argument(std::string(what) + ": expected " + std::to_string(want) + " values, got " +
                                    std::to_string(got));
}

// complex vectors travel as interleaved (re, im) doubles, std::complex's layout
const double *flat(const vec_complex &v) { return reinterpret_cast<const double *>(v.data()); }
double *flat(vec_complex &v) { return reinterpret_cast<double *>(v.data()); }

}  // namespace

// ---------------------------------------------------------------------------

LayerContext::LayerContext(int Ntau, real_t Dim, real_t precision, int maxIts, Scheme scheme) : N_(Ntau)
{
    if (!(Ntau == 128 || Ntau == 256 || Ntau == 512 || Ntau == 1024 || Ntau == 2048))
        throw std::invalid_argument("Ntau must be 128, 256, 512, 1024 or 2048, got " + std::to_string(Ntau));
    if (maxIts < 1) throw std::invalid_argument("the IRK iteration limit must be positive");
    cc_problem p{};
    p.Ntau = Ntau;
    p.SchemeIRK = stages_of(scheme);
    p.MaxIterIRK = maxIts;
    p.NLeft = p.NRight = 1;               // the context's own grid is never integrated over
    p.Dim = Dim;
    p.XLeft = 0.25;
    p.XMid = 0.5;
    p.XRight = 0.75;
    p.PrecisionIRK = precision;
    p.EpsNewton = 1e-8;
    p.precision = CC_FP64;
    const cc_status st = cc_ctx_create(&p, &ctx_);
    if (st != CC_OK) throw SolverError(st, "cc_ctx_create failed (status " + std::to_string((int)st) + ")");
}

LayerContext::~LayerContext()
{
    if (ctx_) cc_ctx_destroy(ctx_);
}

void LayerContext::check(cc_status st, const char *what) const
{
    if (st == CC_OK) return;
    throw SolverError(st, std::string(what) + ": " + cc_last_error(ctx_));
}

// ---------------------------------------------------------------------------
// SpectralTransformer

SpectralTransformer::SpectralTransformer(size_t N, real_t period)
    : ctx_((int)N, 4.0, 1e-15, 100, Scheme::IRK2), N_(N), period_(period)
{
    if (!(period > 0.0)) throw std::invalid_argument("SpectralTransformer: the period must be positive");
}

void SpectralTransformer::fft(const vec_complex &in, vec_complex &out, int sign)
{
    need_size(in.size(), N_, sign > 0 ? "forwardFFT" : "backwardFFT");
    vec_complex res(N_);
    ctx_.check(cc_fft_batched(ctx_.get(), flat(in), flat(res), 1, sign, 0), "cc_fft_batched");
    out.swap(res);
}

// SpectralTransformer.cpp:57-72
void SpectralTransformer::forwardFFT(const vec_real &in, vec_complex &out)
{
    need_size(in.size(), N_, "forwardFFT");
    vec_complex z(N_);
    for (size_t i = 0; i < N_; ++i) z[i] = complex_t(in[i], 0.0);
    fft(z, out, +1);
}

// SpectralTransformer.cpp:77-92
void SpectralTransformer::backwardFFT(const vec_complex &in, vec_real &out)
{
    vec_complex z;
    fft(in, z, -1);
    out.resize(N_);
    for (size_t i = 0; i < N_; ++i) out[i] = z[i].real();
}

// SpectralTransformer.cpp:97-112, 117-132
void SpectralTransformer::forwardFFT(const vec_complex &in, vec_complex &out) { fft(in, out, +1); }
void SpectralTransformer::backwardFFT(const vec_complex &in, vec_complex &out) { fft(in, out, -1); }

void SpectralTransformer::spectral(int op, const std::vector<double> &in, std::vector<double> &out, real_t period,
                                   complex_t lam)
{
    ctx_.check(cc_spectral(ctx_.get(), op, in.data(), out.data(), 1, period, lam.real(), lam.imag()),
               "cc_spectral");
}

// SpectralTransformer.cpp:140-158
void SpectralTransformer::differentiate(const vec_complex &in, vec_complex &out, real_t p)
{
    need_size(in.size(), N_, "differentiate");
    std::vector<double> a(flat(in), flat(in) + 2 * N_), b(2 * N_);
    spectral(CC_SPEC_DIFF, a, b, period_or_default(p), 0.0);
    out.assign(reinterpret_cast<const complex_t *>(b.data()), reinterpret_cast<const complex_t *>(b.data()) + N_);
}

// SpectralTransformer.cpp:165-189
void SpectralTransformer::lamIntegrate(const vec_complex &fk, vec_complex &out, complex_t lambda, real_t p)
{
    need_size(fk.size(), N_, "lamIntegrate");
    std::vector<double> a(flat(fk), flat(fk) + 2 * N_), b(2 * N_);
    spectral(CC_SPEC_LAM, a, b, period_or_default(p), lambda);
    out.assign(reinterpret_cast<const complex_t *>(b.data()), reinterpret_cast<const complex_t *>(b.data()) + N_);
}

// SpectralTransformer.cpp:196-217
void SpectralTransformer::halveModes(const vec_complex &in, vec_complex &out)
{
    need_size(in.size(), N_, "halveModes");
    std::vector<double> a(flat(in), flat(in) + 2 * N_), b(N_);
    spectral(CC_SPEC_HALVE, a, b, period_, 0.0);
    out.assign(reinterpret_cast<const complex_t *>(b.data()),
               reinterpret_cast<const complex_t *>(b.data()) + N_ / 2);
}

// SpectralTransformer.cpp:224-252
void SpectralTransformer::doubleModes(const vec_complex &in, vec_complex &out)
{
    need_size(in.size(), N_ / 2, "doubleModes");
    std::vector<double> a(flat(in), flat(in) + N_), b(2 * N_);
    spectral(CC_SPEC_DOUBLE, a, b, period_, 0.0);
    out.assign(reinterpret_cast<const complex_t *>(b.data()), reinterpret_cast<const complex_t *>(b.data()) + N_);
}

// SpectralTransformer.cpp:264-292
void SpectralTransformer::solveInhom(const vec_real &f, const vec_real &g, vec_real &x, real_t p)
{
    need_size(f.size(), N_, "solveInhom f");
    need_size(g.size(), N_, "solveInhom g");
    std::vector<double> a(2 * N_), b(N_);
    std::copy(f.begin(), f.end(), a.begin());
    std::copy(g.begin(), g.end(), a.begin() + N_);
    spectral(CC_SPEC_INHOM, a, b, period_or_default(p), 0.0);
    x.swap(b);
}

// ---------------------------------------------------------------------------
// ODEStepper

ODEStepper::ODEStepper(int numVars, real_t dim, real_t precision, Scheme method)
    : ctx_(numVars, dim, precision, 100, method), M_(numVars / 2)
{
}

// ODEStepper.cpp:144-150: one stepIRK from Xin to Xout
void ODEStepper::integrate(vec_complex &Yin, vec_complex &Yout, real_t Xin, real_t Xout, bool &converged,
                           int &itsReached, int maxIts)
{
    need_size(Yin.size(), (size_t)M_, "ODEStepper::integrate");
    if (maxIts < 1) throw std::invalid_argument("ODEStepper::integrate: maxIts must be positive");
    vec_complex Y = Yin;
    const double grid[2] = {Xin, Xout};
    const int32_t start = 0, end = 1;
    int64_t its = 0;
    int32_t status = 0;
    ctx_.check(cc_integrate(ctx_.get(), flat(Y), grid, 2, &start, &end, 1, 0, &its, &status, maxIts, nullptr),
               "cc_integrate");
    converged = status == 0;
    itsReached = (int)its;
    Yout.swap(Y);
}

// ---------------------------------------------------------------------------
// ShootingSolver

ShootingSolver::ShootingSolver(int Ntau, real_t Dim, real_t precision, int maxIts, Scheme SchemeIRK)
    : ctx_(Ntau, Dim, precision, maxIts, SchemeIRK), N_(Ntau), M_(Ntau / 2), Precision_(precision), maxIts_(maxIts)
{
}

// ShootingSolver.cpp:53-145 for one side with the Debug snapshots: the state
// reached at x_i is recorded before the step from i, at the reference's
// indices; the integration runs in segments between them (the same steps as
// one launch, bit for bit).
void ShootingSolver::integrate_side(vec_complex &Y, const vec_real &gridX, size_t from, size_t to, int side,
                                    Json *fieldVals, const SimulationConfig *config)
{
    const bool forward = side == 0;
    std::vector<size_t> snaps;
    const size_t every = (size_t)std::max(1, config->DebugNx);
    if (forward) {
        for (size_t i = from; i < to; ++i)
            if (i % every == 0 || i == to - 1) snaps.push_back(i);
    } else {
        for (size_t i = from; i > to; --i)
            if (i % every == 0 || i == to + 1 || i == from) snaps.push_back(i);
    }
    snaps.push_back(to);                       // the end of the side (no snapshot)
    std::vector<double> states, xs;
    size_t pos = from;
    for (size_t k = 0; k < snaps.size(); ++k) {
        const size_t i = snaps[k];
        if (i != pos) {
            const int32_t s = (int32_t)pos, e = (int32_t)i;
            int64_t its = 0;
            int32_t status = 0, fail = -1;
            ctx_.check(cc_integrate(ctx_.get(), flat(Y), gridX.data(), (int32_t)gridX.size(), &s, &e, 1, 0, &its,
                                    &status, maxIts_, &fail),
                       "cc_integrate");
            its_[side] += its;
            if (status != 0) {
                const size_t f = (size_t)fail, g = forward ? f + 1 : f - 1;
                std::ostringstream m;
                m << "No convergence between grid point " << gridX[f] << " (" << f << ")  and " << gridX[g] << " ("
                  << g << ")";
                throw SolverError(CC_IRK_NOCONV, m.str());
            }
            pos = i;
        }
        if (k + 1 < snaps.size()) {
            states.insert(states.end(), flat(Y), flat(Y) + 2 * M_);
            xs.push_back(gridX[i]);
        }
    }
    if (xs.empty()) return;
    std::vector<double> f(xs.size() * 4 * (size_t)N_);
    ctx_.check(cc_state_fields(ctx_.get(), states.data(), xs.data(), f.data(), (int32_t)xs.size()),
               "cc_state_fields");
    const size_t nth = (size_t)std::max(1, config->DebugNtau);
    for (size_t k = 0; k < xs.size(); ++k) {
        const double *A = f.data() + k * 4 * (size_t)N_;
        Json &e = (*fieldVals)[std::to_string(xs[k])];
        e["A"] = Json(every_nth_element(A, N_, nth));
        e["U"] = Json(every_nth_element(A + N_, N_, nth));
        e["V"] = Json(every_nth_element(A + 2 * N_, N_, nth));
        e["f"] = Json(every_nth_element(A + 3 * N_, N_, nth));
    }
}

// ShootingSolver.cpp:33-43
void ShootingSolver::shoot(vec_complex &YLeft, vec_complex &YRight, const vec_real &gridX, size_t iLeft,
                           size_t iRight, size_t iMid, vec_complex &mismatchOut, SimulationConfig *config,
                           Json *fieldVals)
{
    need_size(YLeft.size(), (size_t)M_, "shoot YLeft");
    need_size(YRight.size(), (size_t)M_, "shoot YRight");
    if (!(iLeft <= iMid && iMid <= iRight && iRight < gridX.size()) || gridX.size() < 2)
        throw std::invalid_argument("shoot: need iLeft <= iMid <= iRight < gridX.size() and two grid points");
    its_[0] = its_[1] = 0;
    if (config && config->Debug && fieldVals) {
        integrate_side(YLeft, gridX, iLeft, iMid, 0, fieldVals, config);
        integrate_side(YRight, gridX, iRight, iMid, 1, fieldVals, config);
    } else {
        // both sides in one launch: chain 0 forward, chain 1 backward
        vec_complex Y(2 * (size_t)M_);
        std::copy(YLeft.begin(), YLeft.end(), Y.begin());
        std::copy(YRight.begin(), YRight.end(), Y.begin() + M_);
        const int32_t start[2] = {(int32_t)iLeft, (int32_t)iRight}, end[2] = {(int32_t)iMid, (int32_t)iMid};
        int64_t its[2] = {0, 0};
        int32_t status[2] = {0, 0}, fail[2] = {-1, -1};
        ctx_.check(cc_integrate(ctx_.get(), flat(Y), gridX.data(), (int32_t)gridX.size(), start, end, 2, 0, its,
                                status, maxIts_, fail),
                   "cc_integrate");
        its_[0] = its[0];
        its_[1] = its[1];
        for (int side = 0; side < 2; ++side) {
            if (status[side] == 0) continue;
            const size_t f = (size_t)fail[side], g = side == 0 ? f + 1 : f - 1;
            std::ostringstream m;
            m << "No convergence between grid point " << gridX[f] << " (" << f << ")  and " << gridX[g] << " (" << g
              << ")";
            throw SolverError(CC_IRK_NOCONV, m.str());
        }
        std::copy(Y.begin(), Y.begin() + M_, YLeft.begin());
        std::copy(Y.begin() + M_, Y.end(), YRight.begin());
    }
    compute_mismatch(YLeft, YRight, mismatchOut);
}

// ShootingSolver.cpp:154-182
void ShootingSolver::compute_mismatch(const vec_complex &yLeft, const vec_complex &yRight,
                                      vec_complex &mismatchOut) const
{
    const size_t size = yLeft.size();
    mismatchOut.resize(size);
    for (size_t i = 0; i < size; ++i) mismatchOut[i] = yRight[i] - yLeft[i];
    if (mismatchOut[2].real() > Precision_)
        throw SolverError(CC_DELTA_MISMATCH, "Mismatch in Delta from left and right side!!!");
    mismatchOut[2] = complex_t(yLeft[2].real(), mismatchOut[2].imag());
}

}  // namespace cchost
