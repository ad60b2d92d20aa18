// cc_lu.hpp -- dense LU with partial pivoting and the triangular solves
// (the device replacement of LAPACKE_dgesv, NewtonSolver.cpp:739-757).
//
// Right-looking blocked LU on a row-major n x n matrix, panel width NB (32 in
// FP32, 16 in FP64):
//   k_lu_panel  one workgroup factors the n-k0 x NB panel, held in LDS when it
//               has at most LU_PANEL_ROWS rows (else worked in global memory):
//               pivot = first row index of the largest |a| as LAPACK's idamax,
//               swaps inside the panel, rank-1 updates.  It writes the panel's
//               net row permutation as a swap list and composes it into the
//               factorization's row permutation perm (P b = b[perm]).
//   k_lu_laswp  applies the swap list to every column outside the panel (so
//               the stored L carries the row interchanges as dgetrf's does)
//   k_lu_trsm   U12 = L11^{-1} A12 (unit lower triangular, one column per
//               thread, the column in registers)
//   k_lu_gemm   A22 -= L21 U12, 64x64 tiles per workgroup.  FP64 tiles run on
//               v_mfma_f64_16x16x4f64, FP32 tiles on v_mfma_f32_16x16x4f32.
//   k_lu_solve  P b, then blocked forward/back substitution for one
//               right-hand side (right-looking over 128-byte column blocks:
//               the diagonal block by one wave, the block column's products
//               by 8-lane groups, one row segment per group and row)
#pragma once
#include "cc_device.hpp"

namespace ccd {

constexpr int LU_MAXN = 2048;         // largest order (3/4 of the largest Ntau, rounded up)
constexpr int LU_PANEL_ROWS = 1024;   // panels with more rows are factored in global memory
constexpr int LU_PANEL_T = 512;
constexpr int LU_SWAP_MAX = 64;       // rows one panel can move: 2 x the widest panel
template <typename R>
constexpr int lu_nb()
{
    return sizeof(R) == 8 ? 16 : 32;
}

// Value of x in lane l (l uniform across the wave).
__device__ __forceinline__ float lane_value(float x, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ __forceinline__ double lane_value(double x, int l)
{
    const long long v = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)v, l);
    const int hi = __builtin_amdgcn_readlane((int)(v >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// v where bit is 1, +0 where it is 0, by an integer mask of v's bits.
__device__ __forceinline__ float keep_bits(float v, uint32_t bit)
{
    return __uint_as_float(__float_as_uint(v) & (0u - bit));
}
__device__ __forceinline__ double keep_bits(double v, uint32_t bit)
{
    return __longlong_as_double(__double_as_longlong(v) & -(long long)bit);
}

// Factor columns [k0, k0+w) over rows [k0, n).  perm: row permutation so far
// (identity before the first panel, read as such when k0 == 0); swap:
// [count, (dst, src) x count] -- after the panel, row dst holds what row src
// held before it.
template <typename R, bool INLDS>
__global__ __launch_bounds__(LU_PANEL_T) void k_lu_panel(R *A, int n, int k0, int32_t *perm, int32_t *swap,
                                                          int32_t *info)
{
    constexpr int NB = lu_nb<R>(), T = LU_PANEL_T, W = T / 64;
    constexpr int LD = NB + 1;
    constexpr int MAXR = LU_MAXN / T;
    __shared__ R Ps[INLDS ? LU_PANEL_ROWS * LD : 1];
    __shared__ int rowid[LU_MAXN];
    __shared__ R sval[W];
    __shared__ int sidx[W];
    __shared__ int spiv, scount;
    const int t = threadIdx.x;
    const int m = n - k0, w = min(NB, m);
    R *P;
    int ld;                              // n <= LU_MAXN: every offset fits an int
    if constexpr (INLDS) {
        P = Ps;
        ld = LD;
        for (int e = t; e < m * w; e += T) {
            const int i = e / w, c = e - i * w;
            Ps[i * LD + c] = A[(size_t)(k0 + i) * n + k0 + c];
        }
    } else {
        P = A + (size_t)k0 * n + k0;
        ld = n;
    }
    for (int i = t; i < m; i += T) rowid[i] = i;
    if (t == 0) scount = 0;
    __syncthreads();
    for (int j = 0; j < w; ++j) {
        // pivot: first row index of max |P[i][j]|, i >= j
        R best = (R)-1;
        int bi = m;
        for (int i = j + t; i < m; i += T) {
            const R v = fabs(P[i * ld + j]);
            if (v > best) {
                best = v;
                bi = i;
            }
        }
        wave_argmax(best, bi);
        if ((t & 63) == 0) {
            sval[t >> 6] = best;
            sidx[t >> 6] = bi;
        }
        __syncthreads();
        // every thread reduces the W wave results the same way: no serial
        // step and no second barrier before the swap
        R b = sval[0];
        int i0 = sidx[0];
#pragma unroll
        for (int v = 1; v < W; ++v) argmax_pick(b, i0, sval[v], sidx[v]);
        // a column with no comparable entry (all NaN) yields no candidate:
        // keep the diagonal so every row index stays in [j, m)
        if (i0 >= m) i0 = j;
        const int piv = i0;
        if (t == 0) {
            // zero pivot as LAPACK; a NaN pivot (no candidate) is reported the same way
            if (!(b > (R)0) && info[0] == 0) info[0] = k0 + j + 1;
            if (piv != j) {
                const int r = rowid[j];
                rowid[j] = rowid[piv];
                rowid[piv] = r;
            }
        }
        if (piv != j && t < w) {
            const R tmp = P[j * ld + t];
            P[j * ld + t] = P[piv * ld + t];
            P[piv * ld + t] = tmp;
        }
        __syncthreads();
        // scale column j below the diagonal, then the rank-1 update of the
        // rest of the row.  Thread t owns rows j+1+t, j+1+t+T, ...: the rows
        // it scans for the next column's pivot, so that search needs no
        // barrier first (the next swap comes after the search's barrier).
        // Each row is read whole in one batch of loads and written whole:
        // element by element, every update would wait out a load's latency.
        // The pivot row is zero at columns <= j (and >= w), so those columns
        // are written back as they were; column j then gets l_ij.  (The LDS
        // rows hold NB + 1 values; a panel in global memory is always NB wide.)
        const R pv = P[j * ld + j];
        const bool scale = pv != (R)0;
        const R inv = scale ? (R)1 / pv : (R)0;
        // columns (j, w) as a bit mask; the other columns' values cleared
        // by integer masks (no branch per column)
        const uint32_t keep = (w >= 32 ? ~0u : (1u << w) - 1u) & ~((2u << j) - 1u);
        R pr[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c) pr[c] = keep_bits(P[j * ld + c], (keep >> c) & 1u);
        for (int i = j + 1 + t; i < m; i += T) {
            R a[NB];
#pragma unroll
            for (int c = 0; c < NB; ++c) a[c] = P[i * ld + c];
            R lij = P[i * ld + j];
            if (scale) lij *= inv;
#pragma unroll
            for (int c = 0; c < NB; ++c) P[i * ld + c] = a[c] - lij * pr[c];
            P[i * ld + j] = lij;
        }
    }
    __syncthreads();
    if constexpr (INLDS) {
        for (int e = t; e < m * w; e += T) {
            const int i = e / w, c = e - i * w;
            A[(size_t)(k0 + i) * n + k0 + c] = Ps[i * LD + c];
        }
    }
    // compose the panel's permutation into perm; list the rows that moved
    int pnew[MAXR];
#pragma unroll
    for (int q = 0; q < MAXR; ++q) {
        const int r = t + T * q;
        if (r < m) pnew[q] = k0 == 0 ? rowid[r] : perm[k0 + rowid[r]];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXR; ++q) {
        const int r = t + T * q;
        if (r < m) {
            perm[k0 + r] = pnew[q];
            if (rowid[r] != r) {
                const int e = atomicAdd(&scount, 1);
                swap[1 + 2 * e] = k0 + r;
                swap[2 + 2 * e] = k0 + rowid[r];
            }
        }
    }
    __syncthreads();
    if (t == 0) swap[0] = scount;
}

// Apply a panel's swap list to the columns outside the panel [k0, k0+w).
// 64 columns per workgroup; all moved rows are read before any is written.
template <typename R>
__global__ __launch_bounds__(256) void k_lu_laswp(R *A, int n, int k0, int w, const int32_t *swap)
{
    __shared__ R vals[LU_SWAP_MAX][64];
    const int cnt = swap[0];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c = blockIdx.x * 64 + lane;
    const bool col = c < n && (c < k0 || c >= k0 + w);
    for (int q = wave; q < cnt; q += 4)
        if (col) vals[q][lane] = A[(size_t)swap[2 + 2 * q] * n + c];
    __syncthreads();
    for (int q = wave; q < cnt; q += 4)
        if (col) A[(size_t)swap[1 + 2 * q] * n + c] = vals[q][lane];
}

// U12 = L11^{-1} A12: rows [k0, k0+NB), columns [k0+NB, n); the panel is
// full width whenever columns remain to its right.
template <typename R>
__global__ __launch_bounds__(256) void k_lu_trsm(R *A, int n, int k0)
{
    constexpr int NB = lu_nb<R>();
    __shared__ R L[NB][NB];
    for (int e = threadIdx.x; e < NB * NB; e += 256) {
        const int i = e / NB, l = e % NB;
        L[i][l] = A[(size_t)(k0 + i) * n + k0 + l];
    }
    __syncthreads();
    const int c = k0 + NB + blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    R v[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) v[i] = A[(size_t)(k0 + i) * n + c];
#pragma unroll
    for (int i = 1; i < NB; ++i) {
        R s = v[i];
#pragma unroll
        for (int l = 0; l < i; ++l) s -= L[i][l] * v[l];
        v[i] = s;
    }
#pragma unroll
    for (int i = 1; i < NB; ++i) A[(size_t)(k0 + i) * n + c] = v[i];
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
__global__ __launch_bounds__(256) void k_lu_gemm(R *A, int n, int k0, int nb)
{
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

// Solve with the factors: y = P b (gathered through perm), L y' = y,
// U x = y'.  One workgroup of 1024 threads; b is replaced by x.  info (may be
// null): when the factorization reported a zero pivot, b is left as it is
// (dgesv computes no solution then).
//
// Right-looking over blocks of SB = 128 bytes of columns (32 FP32 or 16 FP64
// values), L's blocks top-down, then U's bottom-up.  Per block: every group
// of 8 lanes loads one 128-byte row segment of the block's column below
// (above) the diagonal block per row it owns -- loads that do not depend on
// the solution, the first PF rows' issued before the block's diagonal part is
// solved -- and wave 0 solves the diagonal block by lane broadcasts; then
// each group subtracts its rows' products with the block's solution (an
// 8-lane sum).  The next diagonal block is fetched one block ahead and kept
// in LDS as the strictly lower (upper) triangle in RB, zero elsewhere, with
// U's reciprocal diagonal beside it, so the diagonal solve runs without
// branches: one fused multiply-add per column and lane (U's divisions become
// multiplications by the reciprocal).
template <typename R, typename RB>
__global__ __launch_bounds__(1024) void k_lu_solve(const R *A, int n, const int32_t *perm, RB *b,
                                                   const int32_t *info = nullptr)
{
    constexpr int VEC = 16 / (int)sizeof(R);  // values per lane and row segment
    constexpr int SB = 8 * VEC;               // block width: 8 lanes x 16 bytes
    constexpr int G = 1024 / 8;               // row groups
    constexpr int PF = 8;                     // rows per group fetched ahead of the diagonal solve
    __shared__ RB bs[LU_MAXN];
    __shared__ RB D[SB][SB + 1];
    __shared__ RB rdiag[SB];
    const int t = threadIdx.x, lane = t & 63, g = t >> 3, q = t & 7;
    if (info && info[0] != 0) return;
    const int nb = (n + SB - 1) / SB;
    // row segments 16-byte aligned: vector loads
    const bool vec = ((size_t)n * sizeof(R)) % 16 == 0 && reinterpret_cast<uintptr_t>(A) % 16 == 0;
    // block s < nb: L's rows [s SB, ..); s >= nb: U's block from the bottom
    auto block = [&](int s, int &i0, int &ib) {
        if (s < nb) {
            i0 = s * SB;
            ib = min(SB, n - i0);
        } else {
            const int i1 = n - (s - nb) * SB;
            i0 = max(0, i1 - SB);
            ib = i1 - i0;
        }
    };
    // element (t / SB, t % SB) of block s's diagonal part (0 outside it)
    auto diag_elem = [&](int s) -> R {
        int i0, ib;
        block(s, i0, ib);
        const int i = t / SB, l = t % SB;
        return (t < SB * SB && i < ib && l < ib) ? A[(size_t)(i0 + i) * n + i0 + l] : (R)0;
    };
    auto put_diag = [&](int s, R d) {
        if (t < SB * SB) {
            const int i = t / SB, l = t % SB;
            const bool upper = s >= nb;
            D[i][l] = (upper ? l > i : l < i) ? (RB)d : (RB)0;
            if (upper && i == l) rdiag[l] = d != (R)0 ? (RB)1 / (RB)d : (RB)0;
        }
    };
    for (int r = t; r < n; r += 1024) bs[r] = b[perm[r]];
    put_diag(0, diag_elem(0));
    for (int s = 0; s < 2 * nb; ++s) {
        const bool upper = s >= nb;
        int i0, ib;
        block(s, i0, ib);
        // rows outside the diagonal block that this block's column updates;
        // there are none unless the block is full (a partial block is L's
        // last or U's first)
        const int rb = upper ? 0 : i0 + ib, re = upper ? i0 : n;
        const int np = (re - rb + G - 1) / G;
        // this lane's part of row r's segment in the block column
        auto segment = [&](int r, R(&a)[VEC]) {
            const R *src = A + (size_t)r * n + i0 + q * VEC;
            if (vec) {
                if constexpr (VEC == 4) {
                    const float4 v = *reinterpret_cast<const float4 *>(src);
                    a[0] = v.x, a[1] = v.y, a[2] = v.z, a[3] = v.w;
                } else {
                    const double2 v = *reinterpret_cast<const double2 *>(src);
                    a[0] = v.x, a[1] = v.y;
                }
            } else {
#pragma unroll
                for (int v = 0; v < VEC; ++v) a[v] = src[v];
            }
        };
        R a[PF][VEC];
#pragma unroll
        for (int p = 0; p < PF; ++p) {
            const int r = rb + g + G * p;
            if (p < np && r < re) segment(r, a[p]);
        }
        const R dnext = s + 1 < 2 * nb ? diag_elem(s + 1) : (R)0;
        __syncthreads();   // D of this block, bs after the previous block's products
        if (t < 64) {
            // lane i < SB holds row i of the block (lanes past ib hold 0 and
            // meet only zeros of D)
            const int i = lane & (SB - 1);
            RB x = lane < ib ? bs[i0 + lane] : (RB)0;
            if (!upper) {
#pragma unroll
                for (int l = 0; l < SB - 1; ++l) x -= D[i][l] * lane_value(x, l);
            } else {
                // lane l keeps its value undivided (D's diagonal is 0 here;
                // no later column touches it) and is divided at the end
#pragma unroll
                for (int l = SB - 1; l >= 0; --l) x -= D[i][l] * (lane_value(x, l) * rdiag[l]);
                x *= rdiag[i];
            }
            if (lane < ib) bs[i0 + lane] = x;
        }
        __syncthreads();   // the block's solution; every read of D done
        if (s + 1 < 2 * nb) put_diag(s + 1, dnext);
        if (np > 0) {
            RB xv[VEC];
#pragma unroll
            for (int v = 0; v < VEC; ++v) xv[v] = bs[i0 + q * VEC + v];
            auto product = [&](int r, const R(&a)[VEC]) {
                RB sum = (RB)0;
                if (r < re) {
#pragma unroll
                    for (int v = 0; v < VEC; ++v) sum += (RB)a[v] * xv[v];
                }
                sum = wave_add_dpp<0xB1>(sum);
                sum = wave_add_dpp<0x4E>(sum);
                sum = wave_add_dpp<0x141>(sum);
                if (q == 0 && r < re) bs[r] -= sum;
            };
#pragma unroll
            for (int p = 0; p < PF; ++p)
                if (p < np) product(rb + g + G * p, a[p]);
            // more than PF rows per group (n above PF G + SB): fetched here
            for (int p = PF; p < np; ++p) {
                const int r = rb + g + G * p;
                R e[VEC] = {};
                if (r < re) segment(r, e);
                product(r, e);
            }
        }
    }
    __syncthreads();
    for (int r = t; r < n; r += 1024) b[r] = bs[r];
}

// r = b - A x (FP64 residual for iterative refinement), one row per thread block
// lane-group; nothing when *skip (the refinement has stopped: r keeps the
// residual of the current x).
__global__ void k_residual(const double *A, const double *x, const double *b, double *r, int n,
                           const int32_t *skip = nullptr)
{
    if (skip && skip[0]) return;
    const int row = blockIdx.x;
    __shared__ double red[4];
    double s = 0.0;
    for (int c = threadIdx.x; c < n; c += blockDim.x) s += A[(size_t)row * n + c] * x[c];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int w = 0; w < (int)(blockDim.x / 64); ++w) tot += red[w];
        r[row] = b[row] - tot;
    }
}

}  // namespace ccd
