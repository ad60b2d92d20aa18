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

}  // namespace ccd
