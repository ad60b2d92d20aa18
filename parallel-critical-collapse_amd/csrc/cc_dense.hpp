// cc_dense.hpp -- kernels that do not depend on the transform length: the
// difference quotient, casts, fills and the vector updates of the Newton step
// (included by the host translation unit only; the dense LU is in cc_lu.hpp).
#pragma once
#include "cc_lu.hpp"

namespace ccd {

// J[j*n + c] = (F[c][j] - F[base][j]) / eps  (NewtonSolver.cpp:682-685)
__global__ void k_fd(const double *F, const int32_t *cols, int ncols, int base, int n, double eps, double *J)
{
    const int j = blockIdx.y;
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= ncols || j >= n) return;
    const int c = cols[b];
    if (c < 0) return;
    J[(size_t)j * n + c] = (F[(size_t)b * n + j] - F[(size_t)base * n + j]) / eps;
}

template <typename R>
__global__ void k_cast_to(const double *in, R *out, size_t count)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = (R)in[i];
}

template <typename R>
__global__ void k_cast_from(const R *in, double *out, size_t count)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = (double)in[i];
}

// Fill kernel: the stream-ordered replacement of hipMemsetAsync in code that
// may be captured into a hipGraph.
template <typename Tp>
__global__ void k_fill(Tp *p, Tp v, size_t count)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

__global__ void k_negate_copy(const double *in, double *out, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = -in[i];
}

// x += d; nothing when *skip (the refinement has stopped).
__global__ void k_axpy(double *x, const double *d, int n, const int32_t *skip = nullptr)
{
    if (skip && skip[0]) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += d[i];
}

// Stop of the iterative refinement: *skip = 1 once ||r|| / ||b|| <= stop (r =
// b - A x from k_residual); the refinement's later kernels then return at
// entry, so the number of sweeps adapts without a host round trip.  One
// workgroup of 256 threads.
__global__ __launch_bounds__(256) void k_refine_stop(const double *r, const double *b, int n, double stop,
                                                     int32_t *skip)
{
    __shared__ double red[2][4];
    if (skip[0]) return;
    double sr = 0.0, sb = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        sr += r[i] * r[i];
        sb += b[i] * b[i];
    }
    sr = wave_sum(sr);
    sb = wave_sum(sb);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = sr;
        red[1][threadIdx.x >> 6] = sb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double tr = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        const double tb = red[1][0] + red[1][1] + red[1][2] + red[1][3];
        if (tr <= stop * stop * tb) skip[0] = 1;
    }
}

// Relative residual of a refined solve: out[0] = ||r|| / ||b|| (r = b - A x
// from k_residual), hold[0] = 1 when it is not <= tol (NaN included) or the
// factorization reported a zero pivot (info[0] != 0).  One workgroup.
__global__ __launch_bounds__(256) void k_refine_check(const double *r, const double *b, const int32_t *info, int n,
                                                      double tol, double *out, int32_t *hold)
{
    __shared__ double red[2][4];
    double sr = 0.0, sb = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        sr += r[i] * r[i];
        sb += b[i] * b[i];
    }
    sr = wave_sum(sr);
    sb = wave_sum(sb);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = sr;
        red[1][threadIdx.x >> 6] = sb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double tr = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        const double tb = red[1][0] + red[1][1] + red[1][2] + red[1][3];
        const double ratio = tb > 0.0 ? sqrt(tr / tb) : (tr > 0.0 ? __builtin_inf() : 0.0);
        out[0] = ratio;
        hold[0] = (!(ratio <= tol) || info[0] != 0) ? 1 : 0;
    }
}

// Correctly rounded square root, as the host's std::sqrt: the device
// sequence behind sqrt() can be 1 ulp off.  The remainder q - r^2 of a
// faithful r is exact under fma, and one Newton correction on it rounds r
// correctly (the exact root is never a rounding midpoint).
__device__ __forceinline__ double sqrt_rn(double q)
{
    const double r = sqrt(q);
    if (!(q > 0.0) || !(r < __builtin_inf())) return r;
    const double rem = fma(-r, r, q);
    return fma(rem, 0.5 / r, r);
}

// One pass of the Newton loop's bookkeeping (NewtonSolver.cpp:298-358) after
// a step at x gave F0 and dx, in the host's operation order (no contracted
// multiply-adds -- hipcc contracts by default, so contraction is switched off
// here --, the norm summed left to right by one thread, a correctly rounded
// square root):
// err = RMS(F0) (763-768); below tol: converged; a mismatch that rose five
// times in a row ends the loop at the last point whose mismatch fell (xold);
// else x += min(1, slow/err) dx, with xold = x first when the mismatch fell.
// st: [0] err, [1] errOld (err starts at 1 as in the reference), [2] the
// damping factor; ist: [0] consecutive rises, [1] outcome (0 go on, 1
// converged, 2 rising mismatch), [2] whether x becomes xold, [3] whether the
// update was applied; hist[its] = err.  hold (may be null): when *hold is
// set the update is decided but not applied (the step's solve is to be
// redone); k_newton_apply applies it later.  One workgroup.
__global__ __launch_bounds__(256) void k_newton_update(double *x, double *xold, const double *F0, const double *dx,
                                                       int n, double tol, double slow, double *st, int32_t *ist,
                                                       double *hist, int its, const int32_t *hold)
{
    // every product and sum rounded on its own, as the host's loop: no fma
#pragma clang fp contract(off)
    __shared__ double s_fac;
    __shared__ int s_mode, s_keep;
    const int t = threadIdx.x;
    if (t == 0) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s = s + F0[i] * F0[i];
        const double err = sqrt_rn(s / (double)n);
        const double errOld = st[0];
        st[0] = err;
        st[1] = errOld;
        hist[its] = err;
        int mode = 0, keep = 0, inc = ist[0];
        double fac = 0.0;
        if (err < tol) {
            mode = 1;
        } else {
            if (log10(err) > log10(errOld)) ++inc;
            if (inc >= 5) {
                mode = 2;
            } else {
                fac = fmin(1.0, slow / err);
                if (log10(err) < log10(errOld)) {
                    keep = 1;
                    inc = 0;
                }
            }
        }
        st[2] = fac;
        ist[0] = inc;
        ist[1] = mode;
        ist[2] = keep;
        ist[3] = 0;
        s_fac = fac;
        s_mode = (mode == 0 && hold && hold[0]) ? -1 : mode;
        s_keep = keep;
    }
    __syncthreads();
    if (s_mode != 0) return;
    const double fac = s_fac;
    for (int i = t; i < n; i += 256) {
        const double xi = x[i];
        if (s_keep) xold[i] = xi;
        x[i] = xi + fac * dx[i];
    }
    if (t == 0) ist[3] = 1;
}

// The update k_newton_update decided and held back (outcome 0, not applied),
// with the step's dx as it stands now.  One workgroup.
__global__ __launch_bounds__(256) void k_newton_apply(double *x, double *xold, const double *dx, int n,
                                                      const double *st, int32_t *ist)
{
#pragma clang fp contract(off)
    const bool go = ist[1] == 0 && ist[3] == 0;
    const double fac = st[2];
    const bool keep = ist[2] != 0;
    __syncthreads();
    if (!go) return;
    for (int i = threadIdx.x; i < n; i += 256) {
        const double xi = x[i];
        if (keep) xold[i] = xi;
        x[i] = xi + fac * dx[i];
    }
    if (threadIdx.x == 0) ist[3] = 1;
}

}  // namespace ccd
