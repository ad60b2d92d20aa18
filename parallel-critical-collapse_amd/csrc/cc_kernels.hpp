// cc_kernels.hpp -- length-N per-column kernels around the chain kernel and
// the unit transform kernel.  All of them use one workgroup per column (or
// signal).
#pragma once
#include "cc_chain.hpp"
#include "cc_spectral.hpp"

namespace ccd {

// Boundary data of one side for a list of columns (NewtonSolver.cpp:403-416
// then ICG.cpp:37-239 or 252-667).  Column b perturbs slot cols[b] of x by
// eps (cols[b] < 0: the unperturbed base), writes its state to
// Y[chain[b]] in the chain precision and Delta to delta[chain[b]].
template <int N, int T, typename R, bool LEFT>
__global__ __launch_bounds__(T) void k_boundary(const double *x, const int32_t *cols, const int32_t *chain,
                                                 double eps, double Xb, double Dim, const cx<double> *tw,
                                                 cx<R> *Y, double *delta)
{
    constexpr int P = N / T, H = P / 2, M = N / 2;
    __shared__ cx<double> xb[N];
    __shared__ double bc[4];
    const int t = threadIdx.x;
    const int b = blockIdx.x;
    const int col = cols[b];
    const int dslot = N / 2 + 2;
    const double Delta = (col == dslot) ? x[dslot] + eps : x[dslot];
    Spec<N, T> S{t, xb, bc, tw, 2.0 * 3.14159265358979323846 / Delta};
    double Up[P], psic[P], fc[P];
    unpack_fields<N, T>(S, x, col, eps, Up, psic, fc);
    cx<double> Yh[H];
    if constexpr (LEFT)
        left_expansion<N, T>(S, Xb, Dim, fc, psic, Yh);
    else
        right_expansion<N, T>(S, Xb, Dim, Up, Yh);
    if (t == 2) Yh[0].x = Delta;                    // ICG.cpp:195 / 624
    const int ch = chain[b];
#pragma unroll
    for (int h = 0; h < H; ++h) Y[(size_t)ch * M + t + T * h] = {(R)Yh[h].x, (R)Yh[h].y};
    if (t == 0) delta[ch] = Delta;
}

// Mismatch at the match point and its packing (ShootingSolver.cpp:154-182,
// NewtonSolver.cpp:420-422).  Column b reads chain states lidx[b], ridx[b]
// and writes F[b*n ...].
template <int N, int T, typename R>
__global__ __launch_bounds__(T) void k_mismatch(const cx<R> *Y, const int32_t *lidx, const int32_t *ridx,
                                                 const double *delta, double precision, const cx<double> *tw,
                                                 double *F, int32_t *status)
{
    constexpr int P = N / T, H = P / 2, M = N / 2, n = 3 * N / 4;
    __shared__ cx<double> xb[N];
    __shared__ double bc[4];
    const int t = threadIdx.x;
    const int b = blockIdx.x;
    const int li = lidx[b], ri = ridx[b];
    // FP32 chains do not carry Delta at full precision: take it from the chain record
    const double DL = delta[li], DR = delta[ri];
    Spec<N, T> S{t, xb, bc, tw, 2.0 * 3.14159265358979323846 / DL};
    cx<double> mis[H];
#pragma unroll
    for (int h = 0; h < H; ++h) {
        const int q = t + T * h;
        const cx<R> l = Y[(size_t)li * M + q], r = Y[(size_t)ri * M + q];
        cx<double> L = {(double)l.x, (double)l.y}, Rr = {(double)r.x, (double)r.y};
        if (q == 2) {
            L.x = DL;
            Rr.x = DR;
        }
        mis[h] = Rr - L;
        if (q == 2) {
            if (mis[h].x > precision) status[b] = 2;   // ShootingSolver.cpp:167
            mis[h].x = L.x;                             // ShootingSolver.cpp:181
        }
    }
    state_to_packed<N, T>(S, mis, F + (size_t)b * n);
}

// ICG::packSpectralFields (ICG.cpp:678-711) for a batch: fields[b] holds the
// three real tau-series Up, psic, fc (N each); Z[b] receives the n-entry
// Newton vector.  Forward transforms (1/N), halveModes, then odd modes
// 1, 3, .. < N/4 of Up and psic and even modes 0, 2, .. < N/4 of fc as (re, im)
// pairs; slot N/2+1 holds Re of fc's folded mode N/4 (its mode N/4 plus mode
// 3N/4).  Slot N/2+2 is Re of fc's mode 2 (the caller overwrites it with
// Delta, NewtonSolver.cpp:273-274).
template <int N, int T>
__global__ __launch_bounds__(T) void k_pack(const double *fields, const cx<double> *tw, double *Z)
{
    constexpr int P = N / T, n = 3 * N / 4;
    __shared__ cx<double> xb[N];
    __shared__ double bc[4];
    const int t = threadIdx.x;
    const double *f = fields + (size_t)blockIdx.x * 3 * N;
    double *z = Z + (size_t)blockIdx.x * n;
    Spec<N, T> S{t, xb, bc, tw, 0.0};
    cx<double> U[P], V[P], E[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        U[p] = {f[t + T * p], 0.0};
        V[p] = {f[N + t + T * p], 0.0};
        E[p] = {f[2 * N + t + T * p], 0.0};
    }
    S.fwd(U);
    S.fwd(V);
    S.fwd(E);
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int k = S.kidx(p);
        if (k >= N / 4 + 1) continue;
        if (k & 1) {
            if (k < N / 4) {
                z[k - 1] = U[p].x;
                z[k] = U[p].y;
                z[N / 4 + k - 1] = V[p].x;
                z[N / 4 + k] = V[p].y;
            }
        } else if (k < N / 4) {
            z[N / 2 + k] = E[p].x;
            if (k != 0) z[N / 2 + k + 1] = E[p].y;
        } else {                                  // k == N/4; 3N/4 sits P/2 slots later
            z[N / 2 + 1] = E[p].x + E[p + P / 2].x;
        }
    }
}

// ICG::unpackSpectralFields (ICG.cpp:719-758) for a batch, reading the Delta
// slot as zero as every call site does (NewtonSolver.cpp:408-409, 318-319):
// Z[b] (n doubles) -> fields[b] = Up, psic, fc (N each).
template <int N, int T>
__global__ __launch_bounds__(T) void k_unpack(const double *Z, const cx<double> *tw, double *fields)
{
    constexpr int P = N / T, n = 3 * N / 4;
    __shared__ cx<double> xb[N];
    __shared__ double bc[4];
    const int t = threadIdx.x;
    Spec<N, T> S{t, xb, bc, tw, 0.0};
    double Up[P], psic[P], fc[P];
    unpack_fields<N, T>(S, Z + (size_t)blockIdx.x * n, -1, 0.0, Up, psic, fc);
    double *f = fields + (size_t)blockIdx.x * 3 * N;
#pragma unroll
    for (int p = 0; p < P; ++p) {
        f[t + T * p] = Up[p];
        f[N + t + T * p] = psic[p];
        f[2 * N + t + T * p] = fc[p];
    }
}

// Batched length-N transform with the reference's scaling (unit entry point).
template <int N, int T, typename R, int SIGN>
__global__ __launch_bounds__(T) void k_fft(const cx<double> *in, cx<double> *out, const cx<R> *tw)
{
    constexpr int P = N / T;
    __shared__ cx<R> xb[N];
    const int t = threadIdx.x;
    const size_t base = (size_t)blockIdx.x * N;
    cx<R> v[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const cx<double> a = in[base + t + T * p];
        v[p] = {(R)a.x, (R)a.y};
    }
    fft1<N, T, SIGN>(v, xb, tw, t);
#pragma unroll
    for (int p = 0; p < P; ++p) {
        cx<R> r = v[p];
        if constexpr (SIGN > 0) r = cscale(r, (R)1 / (R)N);
        out[base + t + T * p] = {(double)r.x, (double)r.y};
    }
}

// The spectral operators of SpectralTransformer on a batch, one signal per
// workgroup (unit entry point).  op (SPEC_*): DIFF, LAM: N complex -> N complex
// (SpectralTransformer.cpp:140-189, wavenumber k0 = 2 pi / period, lambda =
// lr + i li); HALVE: N complex -> N/2 (196-217); DOUBLE: N/2 complex -> N
// (224-252); INHOM: f, g (N real each, one after the other) -> x (N real),
// the periodic solution of x' + f x + g = 0 (264-292).
enum SpecOp { SPEC_DIFF = 0, SPEC_LAM = 1, SPEC_HALVE = 2, SPEC_DOUBLE = 3, SPEC_INHOM = 4 };

template <int N, int T>
__global__ __launch_bounds__(T) void k_spectral(int op, const double *in, double *out, const cx<double> *tw,
                                                double k0, double lr, double li)
{
    constexpr int P = N / T, H = P / 2;
    __shared__ cx<double> xb[N];
    __shared__ double bc[4];
    const int t = threadIdx.x;
    const size_t b = blockIdx.x;
    Spec<N, T> S{t, xb, bc, tw, k0};
    if (op == SPEC_INHOM) {
        const double *f = in + b * 2 * N, *g = f + N;
        double fv[P], gv[P], x[P];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            fv[p] = f[t + T * p];
            gv[p] = g[t + T * p];
        }
        S.solve_inhom(fv, gv, x);
#pragma unroll
        for (int p = 0; p < P; ++p) out[b * N + t + T * p] = x[p];
        return;
    }
    cx<double> z[P];
    if (op == SPEC_DOUBLE) {
        const cx<double> *Y = reinterpret_cast<const cx<double> *>(in) + b * (N / 2);
        cx<double> Yh[H];
#pragma unroll
        for (int h = 0; h < H; ++h) Yh[h] = Y[t + T * h];
        S.dbl(Yh, z);
    } else {
        const cx<double> *X = reinterpret_cast<const cx<double> *>(in) + b * N;
#pragma unroll
        for (int p = 0; p < P; ++p) z[p] = X[t + T * p];
    }
    if (op == SPEC_HALVE) {
        cx<double> Yh[H];
        S.halve(z, Yh);
        cx<double> *Y = reinterpret_cast<cx<double> *>(out) + b * (N / 2);
#pragma unroll
        for (int h = 0; h < H; ++h) Y[t + T * h] = Yh[h];
        return;
    }
    if (op == SPEC_DIFF) S.diff(z);
    else if (op == SPEC_LAM) S.lam(z, lr, li);
    cx<double> *Z = reinterpret_cast<cx<double> *>(out) + b * N;
#pragma unroll
    for (int p = 0; p < P; ++p) Z[t + T * p] = z[p];
}

}  // namespace ccd
