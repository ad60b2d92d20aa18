// cc_ops.hpp -- host launchers of the length-N kernels.
//
// Every kernel templated on the transform length N is instantiated in its
// own translation unit (cc_ops_<N>.hip, built in parallel); the host side
// (cc_capi.hip) reaches them only through Ops<N>.  A launcher enqueues one
// kernel on stream s and returns hipGetLastError().
//
// Thread counts: a chain kernel runs T threads per chain with P = N/T values
// per thread.  kThreads is the default; kAltThreads (0 if none) halves T and
// doubles P (radix-8 instead of radix-4 Stockham passes, one LDS exchange
// fewer per transform, twice the registers per thread).  Which one is faster
// depends on the launch (a few latency-bound chains or a chip-filling batch)
// and is measured by the autotuner.
#pragma once
#include <hip/hip_runtime.h>

#include "cc_tangent.hpp"

namespace ccd {

constexpr int default_threads(int N) { return N == 128 ? 64 : N / 4; }
constexpr int alt_threads(int N) { return N >= 512 ? N / 8 : 0; }

template <int N>
struct Ops {
    static constexpr int kThreads = default_threads(N);
    static constexpr int kAltThreads = alt_threads(N);
    static constexpr int kFewChains = 256;  // one workgroup per CU
    static bool valid_threads(int T) { return T == kThreads || (kAltThreads != 0 && T == kAltThreads); }

    // k_boundary: Y in the chain precision (double or float)
    static hipError_t boundary(hipStream_t s, bool left, bool fp32, int ncols, const double *x, const int32_t *cols,
                               const int32_t *chain, double eps, double Xb, double Dim, const cx<double> *tw, void *Y,
                               double *delta);
    // k_chain; FP32 chains run at kThreads only.  a.rec (if set) is written as
    // double when rec_fp64 (kThreads only), else as float.  Launches of at most
    // kFewChains chains run the double-buffered variant.
    static hipError_t chain(hipStream_t s, int T, int nchains, const ChainParams<double> &a, bool rec_fp64 = false);
    static hipError_t chain(hipStream_t s, int T, int nchains, const ChainParams<float> &a);
    static hipError_t mismatch(hipStream_t s, bool fp32, int ncols, const void *Y, const int32_t *lidx,
                               const int32_t *ridx, const double *delta, double precision, const cx<double> *tw,
                               double *F, int32_t *status);
    // tangent-linear columns; dY is cx<double> when fp64, else cx<float>
    static hipError_t boundary_tangent(hipStream_t s, bool left, bool fp64, int ncols, const double *x,
                                       const int32_t *cols, const int32_t *chain, double hrel, double Xb,
                                       double Dim, const cx<double> *tw, void *dY);
    // solo_lds > 0: the k_tangent<..., SOLO> launch (kThreads) with that much
    // dynamic LDS on top of the kernel's own, so no other workgroup fits beside
    // each of its workgroups
    static hipError_t tangent(hipStream_t s, int T, int nchains, const TangentParams<float> &a,
                              size_t solo_lds = 0);
    static hipError_t tangent(hipStream_t s, int T, int nchains, const TangentParams<double> &a);  // kThreads only
    static hipError_t tangent_column(hipStream_t s, bool fp64, int ncols, const void *dY, const int32_t *cols,
                                     const int32_t *lidx, const int32_t *ridx, const cx<double> *tw, double *J);
    // ICG pack / unpack of a batch of Newton vectors (fields: [batch][3][N], Z: [batch][n])
    static hipError_t pack(hipStream_t s, int batch, const double *fields, const cx<double> *tw, double *Z);
    static hipError_t unpack(hipStream_t s, int batch, const double *Z, const cx<double> *tw, double *fields);
    // unit entry points: batched transform, paired RHS evaluation, spectral operators
    static hipError_t fft(hipStream_t s, bool fp32, int sign, int batch, const cx<double> *in, cx<double> *out,
                          const void *tw);
    static hipError_t rhs_pairs(hipStream_t s, bool fp32, int npairs, const void *Y, const double *xs,
                                const double *delta, void *dY, const void *tw, const void *invm, double Dim,
                                void *rec = nullptr);
    // SpectralTransformer's operators (k_spectral, op a SpecOp) on `batch` signals
    static hipError_t spectral(hipStream_t s, int op, int batch, const double *in, double *out,
                               const cx<double> *tw, double k0, double lr, double li);
};

extern template struct Ops<128>;
extern template struct Ops<256>;
extern template struct Ops<512>;
extern template struct Ops<1024>;
extern template struct Ops<2048>;

}  // namespace ccd
