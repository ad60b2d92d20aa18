// cc_lu.hpp -- dense LU with partial pivoting and the triangular solves
// (the device replacement of LAPACKE_dgesv, NewtonSolver.cpp:739-757).
//
// Right-looking blocked LU on a row-major n x n matrix, panel width NB:
//   k_lu_panel  one workgroup factors the NB-wide panel (pivot search by a
//               workgroup max-reduction, first index on ties as LAPACK's
//               idamax; full-row swaps; rank-1 updates inside the panel)
//   k_lu_trsm   U12 = L11^{-1} A12 (unit lower triangular, one column per thread)
//   k_lu_gemm   A22 -= L21 U12, 64x64 tiles per workgroup.  FP64 tiles run on
//               v_mfma_f64_16x16x4f64, FP32 tiles on v_mfma_f32_16x16x4f32.
//   k_lu_solve  forward/back substitution for one right-hand side
#pragma once
#include "cc_device.hpp"

namespace ccd {

constexpr int LU_NB = 64;
constexpr int LU_PANEL_T = 256;

template <typename R>
__global__ __launch_bounds__(LU_PANEL_T) void k_lu_panel(R *A, int n, int k0, int32_t *ipiv, int32_t *info)
{
    __shared__ R sval[LU_PANEL_T / 64];
    __shared__ int sidx[LU_PANEL_T / 64];
    __shared__ int spiv;
    const int t = threadIdx.x;
    const int kend = min(k0 + LU_NB, n);
    for (int j = k0; j < kend; ++j) {
        // pivot: first row index of max |A[i][j]|, i >= j
        R best = (R)-1;
        int bi = n;
        for (int i = j + t; i < n; i += LU_PANEL_T) {
            const R v = fabs(A[(size_t)i * n + j]);
            if (v > best) {
                best = v;
                bi = i;
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const R ov = __shfl_xor(best, off, 64);
            const int oi = __shfl_xor(bi, off, 64);
            if (ov > best || (ov == best && oi < bi)) {
                best = ov;
                bi = oi;
            }
        }
        if ((t & 63) == 0) {
            sval[t >> 6] = best;
            sidx[t >> 6] = bi;
        }
        __syncthreads();
        if (t == 0) {
            R b = sval[0];
            int i0 = sidx[0];
            for (int w = 1; w < LU_PANEL_T / 64; ++w)
                if (sval[w] > b || (sval[w] == b && sidx[w] < i0)) {
                    b = sval[w];
                    i0 = sidx[w];
                }
            // a column with no comparable entry (all NaN) yields no candidate:
            // keep the diagonal so every row index stays in [j, n)
            if (i0 >= n) i0 = j;
            spiv = i0;
            ipiv[j] = i0 + 1;
            // zero pivot as LAPACK; a NaN pivot is reported the same way
            if (!(fabs(A[(size_t)i0 * n + j]) > (R)0) && info[0] == 0) info[0] = j + 1;
        }
        __syncthreads();
        const int piv = spiv;
        if (piv != j) {
            for (int c = t; c < n; c += LU_PANEL_T) {
                const R tmp = A[(size_t)j * n + c];
                A[(size_t)j * n + c] = A[(size_t)piv * n + c];
                A[(size_t)piv * n + c] = tmp;
            }
        }
        __syncthreads();
        const R pv = A[(size_t)j * n + j];
        if (pv != (R)0) {
            const R inv = (R)1 / pv;
            for (int i = j + 1 + t; i < n; i += LU_PANEL_T) A[(size_t)i * n + j] *= inv;
        }
        __syncthreads();
        // rank-1 update of the rest of the panel
        const int w = kend - (j + 1);
        if (w > 0) {
            for (int e = t; e < (n - j - 1) * w; e += LU_PANEL_T) {
                const int i = j + 1 + e / w, c = j + 1 + e % w;
                A[(size_t)i * n + c] -= A[(size_t)i * n + j] * A[(size_t)j * n + c];
            }
        }
        __syncthreads();
    }
}

// U12 = L11^{-1} A12: rows [k0, k0+nb), columns [k0+nb, n)
template <typename R>
__global__ void k_lu_trsm(R *A, int n, int k0)
{
    const int nb = min(LU_NB, n - k0);
    const int c = k0 + nb + blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    for (int i = 1; i < nb; ++i) {
        R s = A[(size_t)(k0 + i) * n + c];
        for (int l = 0; l < i; ++l) s -= A[(size_t)(k0 + i) * n + k0 + l] * A[(size_t)(k0 + l) * n + c];
        A[(size_t)(k0 + i) * n + c] = s;
    }
}

// MFMA 16x16x4 wrappers.  One wave computes a 16x16 tile D += A(16x4) B(4x16).
// Lane l supplies A[l%16][l/16] and B[l/16][l%16] and receives four elements
// of column l%16 of D: rows 4*(l/16)+r for the f32 instruction (one group of
// four consecutive rows per 16-lane block) and rows l/16 + 4*r for the f64
// instruction (row groups of size one interleaved across the blocks).
template <typename R>
__device__ __forceinline__ int mfma16_row(int lane, int r)
{
    if constexpr (sizeof(R) == 8) return (lane >> 4) + 4 * r;
    else return 4 * (lane >> 4) + r;
}
__device__ __forceinline__ void mfma16(double a, double b, double (&acc)[4])
{
    typedef double v4d __attribute__((ext_vector_type(4)));
    v4d c = {acc[0], acc[1], acc[2], acc[3]};
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    acc[0] = c[0]; acc[1] = c[1]; acc[2] = c[2]; acc[3] = c[3];
}
__device__ __forceinline__ void mfma16(float a, float b, float (&acc)[4])
{
    typedef float v4f __attribute__((ext_vector_type(4)));
    v4f c = {acc[0], acc[1], acc[2], acc[3]};
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    acc[0] = c[0]; acc[1] = c[1]; acc[2] = c[2]; acc[3] = c[3];
}

// A22 -= L21 U12 on the trailing block starting at row/col k0+nb.
// Workgroup: 256 threads = 4 waves, tile 64x64, each wave a 32x32 quadrant
// (2x2 MFMA tiles); K staged through LDS in chunks of 16.
template <typename R>
__global__ __launch_bounds__(256) void k_lu_gemm(R *A, int n, int k0)
{
    const int nb = min(LU_NB, n - k0);
    const int r0 = k0 + nb + blockIdx.y * 64;
    const int c0 = k0 + nb + blockIdx.x * 64;
    __shared__ R La[64][17];
    __shared__ R Ub[16][65];
    const int t = threadIdx.x;
    const int wave = t >> 6, lane = t & 63;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    R acc[2][2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[a][b][r] = (R)0;
    for (int kk = 0; kk < nb; kk += 16) {
        // stage L21[r0..r0+64) x [k0+kk, +16) and U12[k0+kk, +16) x [c0, c0+64)
        for (int e = t; e < 64 * 16; e += 256) {
            const int i = e / 16, l = e % 16;
            const int gr = r0 + i, gk = k0 + kk + l;
            La[i][l] = (gr < n && kk + l < nb) ? A[(size_t)gr * n + gk] : (R)0;
            const int li = e / 64, j = e % 64;
            const int gk2 = k0 + kk + li, gc = c0 + j;
            Ub[li][j] = (gc < n && kk + li < nb) ? A[(size_t)gk2 * n + gc] : (R)0;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 16; ks += 4) {
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    const R av = La[wr + a * 16 + (lane & 15)][ks + (lane >> 4)];
                    const R bv = Ub[ks + (lane >> 4)][wc + b * 16 + (lane & 15)];
                    mfma16(av, bv, acc[a][b]);
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gr = r0 + wr + a * 16 + mfma16_row<R>(lane, r);
                const int gc = c0 + wc + b * 16 + (lane & 15);
                if (gr < n && gc < n) A[(size_t)gr * n + gc] -= acc[a][b][r];
            }
}

// Solve with the factors: b <- P b, L y = b, U x = y.  One workgroup.
template <typename R, typename RB>
__global__ __launch_bounds__(1024) void k_lu_solve(const R *A, int n, const int32_t *ipiv, RB *b)
{
    const int t = threadIdx.x;
    if (t == 0)
        for (int j = 0; j < n; ++j) {
            const int p = ipiv[j] - 1;
            if (p != j) {
                const RB tmp = b[j];
                b[j] = b[p];
                b[p] = tmp;
            }
        }
    __syncthreads();
    for (int i = 1; i < n; ++i) {          // column-oriented forward substitution
        const RB bi = b[i - 1];
        for (int r = i + t; r < n; r += blockDim.x) b[r] -= (RB)A[(size_t)r * n + (i - 1)] * bi;
        __syncthreads();
    }
    for (int i = n - 1; i >= 0; --i) {
        if (t == 0) b[i] = b[i] / (RB)A[(size_t)i * n + i];
        __syncthreads();
        const RB bi = b[i];
        for (int r = t; r < i; r += blockDim.x) b[r] -= (RB)A[(size_t)r * n + i] * bi;
        __syncthreads();
    }
}

// r = b - A x (FP64 residual for iterative refinement), one row per thread block lane-group
__global__ void k_residual(const double *A, const double *x, const double *b, double *r, int n)
{
    const int row = blockIdx.x;
    __shared__ double red[4];
    double s = 0.0;
    for (int c = threadIdx.x; c < n; c += blockDim.x) s += A[(size_t)row * n + c] * x[c];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int w = 0; w < (int)(blockDim.x / 64); ++w) tot += red[w];
        r[row] = b[row] - tot;
    }
}

}  // namespace ccd
