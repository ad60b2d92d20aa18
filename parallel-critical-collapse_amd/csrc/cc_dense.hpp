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

__global__ void k_axpy(double *x, const double *d, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += d[i];
}

// One pass of the Newton loop's bookkeeping (NewtonSolver.cpp:298-358) after
// a step at x gave F0 and dx, in the host's operation order (no contracted
// multiply-adds, the norm summed left to right by one thread):
// err = RMS(F0) (763-768); below tol: converged; a mismatch that rose five
// times in a row ends the loop at the last point whose mismatch fell (xold);
// else x += min(1, slow/err) dx, with xold = x first when the mismatch fell.
// st: [0] err, [1] errOld (err starts at 1 as in the reference); ist: [0]
// consecutive rises, [1] outcome (0 go on, 1 converged, 2 rising mismatch);
// hist[its] = err.  One workgroup.
__global__ __launch_bounds__(256) void k_newton_update(double *x, double *xold, const double *F0, const double *dx,
                                                       int n, double tol, double slow, double *st, int32_t *ist,
                                                       double *hist, int its)
{
    __shared__ double s_fac;
    __shared__ int s_mode, s_keep;
    const int t = threadIdx.x;
    if (t == 0) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s = __dadd_rn(s, __dmul_rn(F0[i], F0[i]));
        const double err = sqrt(s / (double)n);
        const double errOld = st[0];
        st[0] = err;
        st[1] = errOld;
        hist[its] = err;
        int mode = 0, keep = 0, inc = ist[0];
        if (err < tol) {
            mode = 1;
        } else {
            if (log10(err) > log10(errOld)) ++inc;
            if (inc >= 5) {
                mode = 2;
            } else {
                s_fac = fmin(1.0, slow / err);
                if (log10(err) < log10(errOld)) {
                    keep = 1;
                    inc = 0;
                }
            }
        }
        ist[0] = inc;
        ist[1] = mode;
        s_mode = mode;
        s_keep = keep;
    }
    __syncthreads();
    if (s_mode != 0) return;
    const double fac = s_fac;
    for (int i = t; i < n; i += 256) {
        const double xi = x[i];
        if (s_keep) xold[i] = xi;
        x[i] = __dadd_rn(xi, __dmul_rn(fac, dx[i]));
    }
}

}  // namespace ccd
