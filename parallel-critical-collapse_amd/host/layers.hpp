// layers.hpp -- C++ host side of the reference's lower layers above the
// C-ABI, with their arithmetic on the GPU through include/cc_gpu.h:
//   SpectralTransformer  include/SpectralTransformer.hpp:45-101
//   ODEStepper           include/ODEStepper.hpp:90-110
//   ShootingSolver       include/ShootingSolver.hpp:94-117
// Same constructor arguments less the InitialConditionGenerator reference
// (the state <-> field maps it provides run inside the kernels), the same
// vector types and call signatures.  Where the reference ends the process (IRK
// non-convergence, a Delta mismatch between the sides) these throw
// cchost::SolverError; malformed arguments throw std::invalid_argument.
//
// Each object owns one device context (streams, twiddle tables, buffers) of
// its transform length; a call uploads its vectors, runs the kernels and
// copies the result back before returning.  Transform lengths are the
// library's: 128, 256, 512, 1024 and 2048.
#pragma once
#include <complex>
#include <cstddef>
#include <vector>

#include "newton.hpp"

namespace cchost {

using real_t = double;
using complex_t = std::complex<real_t>;
using vec_real = std::vector<real_t>;
using vec_complex = std::vector<complex_t>;

// common.hpp:67: Gauss-Legendre schemes with 1, 2 and 3 stages
enum class Scheme { IRK1, IRK2, IRK3 };

// One cc_ctx of transform length Ntau for the layer classes: FP64 policy, the
// given Dim, IRK tolerance, iteration limit and scheme.  The context's own
// x-grid is never used (every integration here takes the caller's grid).
class LayerContext {
public:
    LayerContext(int Ntau, real_t Dim, real_t precision, int maxIts, Scheme scheme);
    ~LayerContext();
    LayerContext(const LayerContext &) = delete;
    LayerContext &operator=(const LayerContext &) = delete;

    cc_ctx *get() const { return ctx_; }
    int Ntau() const { return N_; }
    void check(cc_status st, const char *what) const;

private:
    cc_ctx *ctx_ = nullptr;
    int N_ = 0;
};

// SpectralTransformer (SpectralTransformer.cpp:22-292): the reference's DFT
// sign convention (forward = exp(+i..)/N), differentiation, lambda
// integration, halving/doubling of the mode set and the periodic solve of
// x' + f x + g = 0.  Inputs have the transformer's length N (halveModes: N
// in, N/2 out; doubleModes: N/2 in, N out); a period argument of 0 means the
// constructor's period.
class SpectralTransformer {
public:
    explicit SpectralTransformer(size_t N, real_t period);

    void forwardFFT(const vec_real &in, vec_complex &out);
    void backwardFFT(const vec_complex &in, vec_real &out);
    void forwardFFT(const vec_complex &in, vec_complex &out);
    void backwardFFT(const vec_complex &in, vec_complex &out);
    void differentiate(const vec_complex &in, vec_complex &out, real_t period_ = 0.0);
    void lamIntegrate(const vec_complex &fk, vec_complex &out, complex_t lambda = complex_t(0.0),
                      real_t period_ = 0.0);
    void halveModes(const vec_complex &in, vec_complex &out);
    void doubleModes(const vec_complex &in, vec_complex &out);
    void solveInhom(const vec_real &f, const vec_real &g, vec_real &x, real_t period_ = 0.0);

private:
    LayerContext ctx_;
    size_t N_;
    real_t period_;
    void fft(const vec_complex &in, vec_complex &out, int sign);
    void spectral(int op, const std::vector<double> &in, std::vector<double> &out, real_t period, complex_t lam);
    real_t period_or_default(real_t p) const { return p == 0.0 ? period_ : p; }
};

// ODEStepper (ODEStepper.cpp:22-292): one Gauss-Legendre IRK step of the
// state (Ntau/2 complex coefficients, Delta in Re Y[2]) from Xin to Xout,
// the stage fixed point solved to the constructor's precision with at most
// maxIts iterations.  When the fixed point does not converge, converged is
// false and Yout holds Yin.
class ODEStepper {
public:
    ODEStepper(int numVars, real_t dim, real_t precision, Scheme method);

    void integrate(vec_complex &Yin, vec_complex &Yout, real_t Xin, real_t Xout, bool &converged, int &itsReached,
                   int maxIts);

private:
    LayerContext ctx_;
    int M_;
};

// ShootingSolver (ShootingSolver.cpp:33-182): YLeft integrated forward from
// gridX[iLeft] to gridX[iMid] and YRight backward from gridX[iRight] to
// gridX[iMid], in place, both sides in one launch; mismatchOut = YRight -
// YLeft with Re of entry 2 set to the left side's Delta.  With
// config->Debug and fieldVals, the fields A = sqrt(1/IA2), U, V, f at the
// reference's snapshot points (every DebugNx-th grid point and the ends) are
// added to *fieldVals under std::to_string(x), every DebugNtau-th tau value.
class ShootingSolver {
public:
    ShootingSolver(int Ntau, real_t Dim, real_t precision, int maxIts, Scheme SchemeIRK = Scheme::IRK2);

    void shoot(vec_complex &YLeft, vec_complex &YRight, const vec_real &gridX, size_t iLeft, size_t iRight,
               size_t iMid, vec_complex &mismatchOut, SimulationConfig *config = nullptr, Json *fieldVals = nullptr);

    // iterations of the last shoot's fixed points, summed over the steps of each side
    long long iterationsLeft() const { return its_[0]; }
    long long iterationsRight() const { return its_[1]; }

private:
    LayerContext ctx_;
    int N_, M_;
    real_t Precision_;
    int maxIts_;
    long long its_[2] = {0, 0};

    void integrate_side(vec_complex &Y, const vec_real &gridX, size_t from, size_t to, int side, Json *fieldVals,
                        const SimulationConfig *config);
    void compute_mismatch(const vec_complex &yLeft, const vec_complex &yRight, vec_complex &mismatchOut) const;
};

}  // namespace cchost
