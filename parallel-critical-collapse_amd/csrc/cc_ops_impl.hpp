// cc_ops_impl.hpp -- definitions of the Ops<N> launchers.  Included only by
// the per-length translation units cc_ops_<N>.hip, each of which explicitly
// instantiates Ops<N> for its N.
#pragma once
#include "cc_kernels.hpp"
#include "cc_ops.hpp"

namespace ccd {

template <int N>
hipError_t Ops<N>::boundary(hipStream_t s, bool left, bool fp32, int ncols, const double *x, const int32_t *cols,
                            const int32_t *chain, double eps, double Xb, double Dim, const cx<double> *tw, void *Y,
                            double *delta)
{
    constexpr int T = kThreads;
    const dim3 g((unsigned)ncols), b(T);
    if (fp32) {
        if (left)
            hipLaunchKernelGGL((k_boundary<N, T, float, true>), g, b, 0, s, x, cols, chain, eps, Xb, Dim, tw,
                               static_cast<cx<float> *>(Y), delta);
        else
            hipLaunchKernelGGL((k_boundary<N, T, float, false>), g, b, 0, s, x, cols, chain, eps, Xb, Dim, tw,
                               static_cast<cx<float> *>(Y), delta);
    } else {
        if (left)
            hipLaunchKernelGGL((k_boundary<N, T, double, true>), g, b, 0, s, x, cols, chain, eps, Xb, Dim, tw,
                               static_cast<cx<double> *>(Y), delta);
        else
            hipLaunchKernelGGL((k_boundary<N, T, double, false>), g, b, 0, s, x, cols, chain, eps, Xb, Dim, tw,
                               static_cast<cx<double> *>(Y), delta);
    }
    return hipGetLastError();
}

template <int N>
hipError_t Ops<N>::chain(hipStream_t s, int T, int nchains, const ChainParams<double> &a, bool rec_fp64)
{
    // Launches of up to one workgroup per CU (base chains, the residual) run
    // the double-buffered variant: its doubled LDS costs no occupancy there.
    // The exchanges move the same values, so both variants give identical chains.
    constexpr int DB = N <= 1024 ? 1 : 0;
    const bool dbl = nchains <= kFewChains;
    const dim3 g((unsigned)nchains);
    if (a.stages != 2) {                       // SchemeIRK 1 or 3: kThreads, no record
        if (rec_fp64 || a.rec || T != kThreads) return hipErrorInvalidValue;
        if (a.stages == 1) {
            if (dbl)
                hipLaunchKernelGGL((k_chain<N, kThreads, double, float, DB, 1>), g, dim3(kThreads), 0, s, a);
            else
                hipLaunchKernelGGL((k_chain<N, kThreads, double, float, 0, 1>), g, dim3(kThreads), 0, s, a);
        } else if (a.stages == 3) {
            if (dbl)
                hipLaunchKernelGGL((k_chain<N, kThreads, double, float, DB, 3>), g, dim3(kThreads), 0, s, a);
            else
                hipLaunchKernelGGL((k_chain<N, kThreads, double, float, 0, 3>), g, dim3(kThreads), 0, s, a);
        } else {
            return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    if (rec_fp64) {
        if (T != kThreads) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_chain<N, kThreads, double, double, DB>), g, dim3(kThreads), 0, s, a);
    } else if (T == kThreads) {
        if (dbl)
            hipLaunchKernelGGL((k_chain<N, kThreads, double, float, DB>), g, dim3(kThreads), 0, s, a);
        else
            hipLaunchKernelGGL((k_chain<N, kThreads, double>), g, dim3(kThreads), 0, s, a);
    } else {
        if constexpr (kAltThreads != 0) {
            if (T != kAltThreads) return hipErrorInvalidValue;
            if (dbl)
                hipLaunchKernelGGL((k_chain<N, kAltThreads, double, float, DB>), g, dim3(kAltThreads), 0, s, a);
            else
                hipLaunchKernelGGL((k_chain<N, kAltThreads, double>), g, dim3(kAltThreads), 0, s, a);
        } else {
            return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

template <int N>
hipError_t Ops<N>::chain(hipStream_t s, int T, int nchains, const ChainParams<float> &a)
{
    if (T != kThreads || a.stages != 2) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_chain<N, kThreads, float>), dim3((unsigned)nchains), dim3(kThreads), 0, s, a);
    return hipGetLastError();
}

template <int N>
hipError_t Ops<N>::mismatch(hipStream_t s, bool fp32, int ncols, const void *Y, const int32_t *lidx,
                            const int32_t *ridx, const double *delta, double precision, const cx<double> *tw,
                            double *F, int32_t *status)
{
    constexpr int T = kThreads;
    if (fp32)
        hipLaunchKernelGGL((k_mismatch<N, T, float>), dim3((unsigned)ncols), dim3(T), 0, s,
                           static_cast<const cx<float> *>(Y), lidx, ridx, delta, precision, tw, F, status);
    else
        hipLaunchKernelGGL((k_mismatch<N, T, double>), dim3((unsigned)ncols), dim3(T), 0, s,
                           static_cast<const cx<double> *>(Y), lidx, ridx, delta, precision, tw, F, status);
    return hipGetLastError();
}

template <int N>
hipError_t Ops<N>::boundary_tangent(hipStream_t s, bool left, bool fp64, int ncols, const double *x,
                                    const int32_t *cols, const int32_t *chain, double hrel, double Xb, double Dim,
                                    const cx<double> *tw, void *dY)
{
    constexpr int T = kThreads;
    const dim3 g((unsigned)ncols), b(T);
    if (fp64) {
        if (left)
            hipLaunchKernelGGL((k_boundary_tangent<N, T, true, double>), g, b, 0, s, x, cols, chain, hrel, Xb, Dim,
                               tw, static_cast<cx<double> *>(dY));
        else
            hipLaunchKernelGGL((k_boundary_tangent<N, T, false, double>), g, b, 0, s, x, cols, chain, hrel, Xb, Dim,
                               tw, static_cast<cx<double> *>(dY));
    } else {
        if (left)
            hipLaunchKernelGGL((k_boundary_tangent<N, T, true, float>), g, b, 0, s, x, cols, chain, hrel, Xb, Dim,
                               tw, static_cast<cx<float> *>(dY));
        else
            hipLaunchKernelGGL((k_boundary_tangent<N, T, false, float>), g, b, 0, s, x, cols, chain, hrel, Xb, Dim,
                               tw, static_cast<cx<float> *>(dY));
    }
    return hipGetLastError();
}

template <int N>
hipError_t Ops<N>::tangent(hipStream_t s, int T, int nchains, const TangentParams<float> &a, size_t solo_lds)
{
    if (solo_lds > 0) {
        if (T != kThreads) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_tangent<N, kThreads, float, true>), dim3((unsigned)nchains), dim3(kThreads), solo_lds,
                           s, a);
        return hipGetLastError();
    }
    if (T == kThreads) {
        hipLaunchKernelGGL((k_tangent<N, kThreads, float>), dim3((unsigned)nchains), dim3(kThreads), 0, s, a);
    } else {
        if constexpr (kAltThreads != 0) {
            if (T != kAltThreads) return hipErrorInvalidValue;
            hipLaunchKernelGGL((k_tangent<N, kAltThreads, float>), dim3((unsigned)nchains), dim3(kAltThreads), 0, s,
                               a);
        } else {
            return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

template <int N>
hipError_t Ops<N>::tangent(hipStream_t s, int T, int nchains, const TangentParams<double> &a)
{
    if (T != kThreads) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_tangent<N, kThreads, double>), dim3((unsigned)nchains), dim3(kThreads), 0, s, a);
    return hipGetLastError();
}

template <int N>
hipError_t Ops<N>::tangent_column(hipStream_t s, bool fp64, int ncols, const void *dY, const int32_t *cols,
                                  const int32_t *lidx, const int32_t *ridx, const cx<double> *tw, double *J)
{
    constexpr int T = kThreads;
    if (fp64)
        hipLaunchKernelGGL((k_tangent_column<N, T, double>), dim3((unsigned)ncols), dim3(T), 0, s,
                           static_cast<const cx<double> *>(dY), cols, lidx, ridx, tw, J);
    else
        hipLaunchKernelGGL((k_tangent_column<N, T, float>), dim3((unsigned)ncols), dim3(T), 0, s,
                           static_cast<const cx<float> *>(dY), cols, lidx, ridx, tw, J);
    return hipGetLastError();
}

template <int N>
hipError_t Ops<N>::pack(hipStream_t s, int batch, const double *fields, const cx<double> *tw, double *Z)
{
    hipLaunchKernelGGL((k_pack<N, kThreads>), dim3((unsigned)batch), dim3(kThreads), 0, s, fields, tw, Z);
    return hipGetLastError();
}

template <int N>
hipError_t Ops<N>::unpack(hipStream_t s, int batch, const double *Z, const cx<double> *tw, double *fields)
{
    hipLaunchKernelGGL((k_unpack<N, kThreads>), dim3((unsigned)batch), dim3(kThreads), 0, s, Z, tw, fields);
    return hipGetLastError();
}

template <int N>
hipError_t Ops<N>::fft(hipStream_t s, bool fp32, int sign, int batch, const cx<double> *in, cx<double> *out,
                       const void *tw)
{
    constexpr int T = kThreads;
    const dim3 g((unsigned)batch), b(T);
    if (fp32) {
        const cx<float> *w = static_cast<const cx<float> *>(tw);
        if (sign > 0) hipLaunchKernelGGL((k_fft<N, T, float, 1>), g, b, 0, s, in, out, w);
        else hipLaunchKernelGGL((k_fft<N, T, float, -1>), g, b, 0, s, in, out, w);
    } else {
        const cx<double> *w = static_cast<const cx<double> *>(tw);
        if (sign > 0) hipLaunchKernelGGL((k_fft<N, T, double, 1>), g, b, 0, s, in, out, w);
        else hipLaunchKernelGGL((k_fft<N, T, double, -1>), g, b, 0, s, in, out, w);
    }
    return hipGetLastError();
}

template <int N>
hipError_t Ops<N>::rhs_pairs(hipStream_t s, bool fp32, int npairs, const void *Y, const double *xs,
                             const double *delta, void *dY, const void *tw, const void *invm, double Dim,
                             void *rec)
{
    constexpr int T = kThreads;
    if (fp32)
        hipLaunchKernelGGL((k_rhs_pairs<N, T, float>), dim3((unsigned)npairs), dim3(T), 0, s,
                           static_cast<const cx<float> *>(Y), xs, delta, static_cast<cx<float> *>(dY),
                           static_cast<const cx<float> *>(tw), static_cast<const float *>(invm), Dim,
                           static_cast<float *>(rec));
    else
        hipLaunchKernelGGL((k_rhs_pairs<N, T, double>), dim3((unsigned)npairs), dim3(T), 0, s,
                           static_cast<const cx<double> *>(Y), xs, delta, static_cast<cx<double> *>(dY),
                           static_cast<const cx<double> *>(tw), static_cast<const double *>(invm), Dim,
                           static_cast<double *>(rec));
    return hipGetLastError();
}

template <int N>
hipError_t Ops<N>::spectral(hipStream_t s, int op, int batch, const double *in, double *out, const cx<double> *tw,
                            double k0, double lr, double li)
{
    hipLaunchKernelGGL((k_spectral<N, kThreads>), dim3((unsigned)batch), dim3(kThreads), 0, s, op, in, out, tw, k0,
                       lr, li);
    return hipGetLastError();
}

}  // namespace ccd
