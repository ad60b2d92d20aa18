// cc_spectral.hpp -- FP64 spectral toolkit for one workgroup (layout L) and
// the per-column boundary kernels built on it: unpack + Taylor-series
// boundary data (prologue) and mismatch + pack (epilogue).
//
// Reference routines restated: SpectralTransformer.cpp:57-292,
// InitialConditionGenerator.cpp:37-239 (left series), 456-624 (right series,
// serial branch), 678-758 (pack/unpack), 766-778 (fields -> state),
// 891-932 (state -> fields), ShootingSolver.cpp:154-182 (mismatch).
#pragma once
#include "cc_device.hpp"

namespace ccd {

// Smith's complex division (a+ib)/(c+id)
__device__ __forceinline__ cx<double> cdiv(cx<double> n, double c, double d)
{
    if (fabs(c) < fabs(d)) {
        const double ratio = c / d, denom = c * ratio + d;
        return {(n.x * ratio + n.y) / denom, (n.y * ratio - n.x) / denom};
    } else {
        const double ratio = d / c, denom = d * ratio + c;
        return {(n.y * ratio + n.x) / denom, (n.y - n.x * ratio) / denom};
    }
}

template <int N, int T>
struct Spec {
    static constexpr int P = N / T;
    static constexpr int H = P / 2;
    static constexpr int M = N / 2;
    using C = cx<double>;
    int t;
    C *xb;            // N complex LDS exchange buffer
    double *bc;       // LDS broadcast slots (>= 4)
    const C *tw;      // twiddles e^{2 pi i e/N}
    double k0;        // 2 pi / Delta

    __device__ __forceinline__ int kidx(int p) const { return t + T * p; }

    __device__ void fwd(C (&v)[P])
    {
        fft1<N, T, +1>(v, xb, tw, t);
#pragma unroll
        for (int p = 0; p < P; ++p) v[p] = cscale(v[p], 1.0 / N);
    }
    __device__ void bwd(C (&v)[P]) { fft1<N, T, -1>(v, xb, tw, t); }

    // SpectralTransformer.cpp:57-72
    __device__ void r2c(const double (&x)[P], C (&X)[P])
    {
#pragma unroll
        for (int p = 0; p < P; ++p) X[p] = {x[p], 0.0};
        fwd(X);
    }
    // SpectralTransformer.cpp:77-92
    __device__ void c2r(C (&X)[P], double (&x)[P])
    {
        bwd(X);
#pragma unroll
        for (int p = 0; p < P; ++p) x[p] = X[p].x;
    }
    // SpectralTransformer.cpp:140-158
    __device__ void diff(C (&X)[P]) const
    {
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const int k = kidx(p);
            const double w = (k == N / 2) ? 0.0 : (double)wavenumber<N>(k) * k0;
            const C v = X[p];
            X[p] = (k == N / 2) ? C{0.0, 0.0} : C{w * v.y, -(w * v.x)};
        }
    }
    // forward -> differentiate -> backward
    __device__ void dtau(const double (&x)[P], double (&out)[P])
    {
        C X[P];
        r2c(x, X);
        diff(X);
        c2r(X, out);
    }
    // SpectralTransformer.cpp:165-189
    __device__ void lam(C (&X)[P], double lr, double li) const
    {
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const int k = kidx(p);
            const C f = X[p];
            if (k == 0) {
                X[p] = (lr != 0.0 || li != 0.0) ? cdiv(f, lr, li) : C{0.0, 0.0};
            } else if (k == N / 2) {
                const double q = k0 * (double)N / 2.0;
                const C fac = cdiv(C{lr, li}, lr * lr - li * li + q * q, 2.0 * lr * li);
                X[p] = cmul(fac, f);
            } else {
                const double dr = lr, di = li - (double)wavenumber<N>(k) * k0;
                X[p] = (hypot(dr, di) < 1e-14) ? C{0.0, 0.0} : cdiv(f, dr, di);
            }
        }
    }
    // SpectralTransformer.cpp:264-292: periodic solution of x' + f x + g = 0
    __device__ void solve_inhom(const double (&f)[P], const double (&g)[P], double (&x)[P])
    {
        C fh[P];
        r2c(f, fh);
        if (t == 0) {
            bc[0] = fh[0].x;
            bc[1] = fh[0].y;
        }
        C Fh[P];
#pragma unroll
        for (int p = 0; p < P; ++p) Fh[p] = fh[p];
        lam(Fh, 0.0, 0.0);
        double Fr[P];
        c2r(Fh, Fr);                    // its barriers publish bc[0..1]
        const double lr = bc[0], li = bc[1];
        double h[P];
#pragma unroll
        for (int p = 0; p < P; ++p) h[p] = -g[p] * exp(Fr[p]);
        C hh[P];
        r2c(h, hh);
        lam(hh, lr, li);
        c2r(hh, x);
#pragma unroll
        for (int p = 0; p < P; ++p) x[p] *= exp(-Fr[p]);
        __syncthreads();
    }
    // ICG.cpp:766-778: Y = halve(forward(U + i (V + F))); Yh in state slots
    __device__ void fields_to_state(const double (&U)[P], const double (&V)[P], const double (&F)[P],
                                    C (&Yh)[H])
    {
        C z[P];
#pragma unroll
        for (int p = 0; p < P; ++p) z[p] = {U[p], V[p] + F[p]};
        fwd(z);
        halve(z, Yh);
    }
    // SpectralTransformer.cpp:196-217 from layout L to state slots
    __device__ void halve(const C (&z)[P], C (&Yh)[H]) const
    {
#pragma unroll
        for (int h = 0; h < H; ++h) {
            const int q = t + T * h;
            Yh[h] = (q < M / 2) ? z[h] : ((q == M / 2) ? z[h] + z[h + H] : z[h + H]);
        }
    }
    // SpectralTransformer.cpp:224-252 (doubleModes) from state slots to layout L
    __device__ void dbl(const C (&Yh)[H], C (&z)[P]) const
    {
#pragma unroll
        for (int h = 0; h < H; ++h) {
            const int q = t + T * h;
            const C Y = Yh[h];
            if (q < M / 2) {
                z[h] = Y;
                z[h + H] = {0.0, 0.0};
            } else if (q == M / 2) {
                z[h] = cscale(Y, 0.5);
                z[h + H] = Y;
            } else {
                z[h] = {0.0, 0.0};
                z[h + H] = Y;
            }
        }
    }
};

// ----------------------------------------------------------------------------
// Unpack the Newton vector Z (n doubles, Delta slot already zeroed by the
// caller's view) into the three real tau-series (ICG.cpp:719-758).  The
// doubled Hermitian spectra are built directly in layout L.
// ----------------------------------------------------------------------------
template <int N, int T>
__device__ void unpack_fields(Spec<N, T> &S, const double *Z, int eps_col, double eps,
                              double (&Up)[N / T], double (&psic)[N / T], double (&fc)[N / T])
{
    constexpr int P = N / T;
    auto zv = [&](int i) -> double {
        if (i == N / 2 + 2) return 0.0;                 // NewtonSolver.cpp:408-409
        return Z[i] + ((i == eps_col) ? eps : 0.0);
    };
    cx<double> A1[P], A2[P], E[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int k = S.kidx(p);
        cx<double> o1 = {0.0, 0.0}, o2 = {0.0, 0.0}, ev = {0.0, 0.0};
        if (k & 1) {
            if (k < N / 4) {
                o1 = {zv(k - 1), zv(k)};
                o2 = {zv(N / 4 + k - 1), zv(N / 4 + k)};
            } else if (k > 3 * N / 4) {
                const int kk = N - k;
                o1 = {zv(kk - 1), -zv(kk)};
                o2 = {zv(N / 4 + kk - 1), -zv(N / 4 + kk)};
            }
        } else {
            if (k == 0) {
                ev = {zv(N / 2), 0.0};
            } else if (k < N / 4) {
                ev = {zv(N / 2 + k), zv(N / 2 + k + 1)};
            } else if (k == N / 4 || k == 3 * N / 4) {
                ev = {zv(N / 2 + 1) / 2.0, 0.0};
            } else if (k > 3 * N / 4) {
                const int kk = N - k;
                ev = {zv(N / 2 + kk), -zv(N / 2 + kk + 1)};
            }
        }
        A1[p] = o1;
        A2[p] = o2;
        E[p] = ev;
    }
    // one transform per field, as the reference: a column that perturbs one
    // field must leave the other two bit-identical
    S.bwd(A1);
    S.bwd(A2);
    S.bwd(E);
#pragma unroll
    for (int p = 0; p < P; ++p) {
        Up[p] = A1[p].x;
        psic[p] = A2[p].x;
        fc[p] = E[p].x;
    }
}

// ICG.cpp:37-239 (left, order-5 centre series) -> state slots
template <int N, int T>
__device__ void left_expansion(Spec<N, T> &S, double XLeft, double d, const double (&f0)[N / T],
                               const double (&psi0)[N / T], cx<double> (&Y)[N / (2 * T)])
{
    constexpr int P = N / T;
    double df0[P], d2f0[P], d3f0[P];
    S.dtau(f0, df0);
    S.dtau(df0, d2f0);
    S.dtau(d2f0, d3f0);
    double c1[P], c2[P], u1[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        c1[p] = 1.0;
        c2[p] = -(d - 1.0) * psi0[p] * f0[p];
    }
    S.solve_inhom(c1, c2, u1);
    double du1[P], d2u1[P], d3u1[P], d4u1[P];
    S.dtau(u1, du1);
    S.dtau(du1, d2u1);
    S.dtau(d2u1, d3u1);
    S.dtau(d3u1, d4u1);
    double U[P], V[P], F[P];
    const double d2c = (d - 2.0) * (d - 2.0) * (d - 2.0);
    const double d26 = d2c * d2c;
    const double x1 = XLeft, x2 = XLeft * XLeft, x3 = x2 * XLeft, x4 = x3 * XLeft, x5 = x4 * XLeft;
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const double f = f0[p], psi = psi0[p];
        const double u1j = u1[p], du = du1[p], d2u = d2u1[p], d3u = d3u1[p], d4u = d4u1[p];
        const double df = df0[p], d2f = d2f0[p], d3f = d3f0[p];
        const double f_2 = f * f, f_3 = f_2 * f, f_4 = f_3 * f, f_5 = f_4 * f, f_6 = f_5 * f, f_7 = f_6 * f;
        const double u_2 = u1j * u1j, u_3 = u_2 * u1j, u_4 = u_3 * u1j, u_5 = u_4 * u1j;
        const double v1 = u1j;
        const double f2 = ((d - 3.0) * d2c * f * u_2) / (8.0 * (d - 1.0));
        const double u2 = -psi, v2 = psi;
        const double common3 = ((d - 3.0) * d2c * f_3 * u_3 + 4.0 * df * (u1j + du) -
                                4.0 * f * (2.0 * u1j + 3.0 * du + d2u));
        const double u3 = common3 / (8.0 * (1.0 - d) * f_3);
        const double v3 = u3;
        const double den = 128.0 * (d - 1.0) * (d - 1.0) * (1.0 + d);
        const double f4 = -((d - 3.0) * d2c *
                            ((d2c * (5.0 - 6.0 * d + d * d) * f_3 * u_4) + 8.0 * (d - 1.0) * df * u1j * (du + u1j) -
                             8.0 * f * (du * du + ((d - 1.0) * d2u + (-1.0 + 3.0 * d) * du) * u1j +
                                        (-1.0 + 2.0 * d) * u_2))) /
                          (den * f_2);
        const double u4 = (-192.0 * (d - 1.0) * df * df * du * f_2 + 192.0 * (d - 1.0) * d2u * df * f_3 +
                           64.0 * (d - 1.0) * d2f * du * f_3 + 640.0 * (d - 1.0) * df * du * f_3 -
                           64.0 * (d - 1.0) * (6.0 * d2u + d3u) * f_4 - 704.0 * (d - 1.0) * du * f_4 -
                           192.0 * (d - 1.0) * df * df * f_2 * u1j + 64.0 * (d - 1.0) * d2f * f_3 * u1j +
                           448.0 * (d - 1.0) * df * f_3 * u1j - 384.0 * (d - 1.0) * f_4 * u1j +
                           32.0 * d2c * (3.0 - 7.0 * d + 2.0 * d * d) * f_6 * u_2 * (du + u1j)) /
                          (den * f_7);
        const double v4 = -u4;
        const double g13 = 3.0 - 13.0 * d + 4.0 * d * d;
        const double u5 = (-64.0 * (d - 1.0) * d2f * d2u * f_2 - 16.0 * (d - 1.0) * (15.0 * d2f + d3f) * du * f_2 -
                           32.0 * (d - 1.0) * df * (20.0 * d2u + 3.0 * d3u + 40.0 * du) * f_2 +
                           560.0 * (d - 1.0) * d2u * f_3 + 160.0 * (d - 1.0) * d3u * f_3 +
                           16.0 * (d - 1.0) * d4u * f_3 + 800.0 * (d - 1.0) * du * f_3 -
                           16.0 * (d - 1.0) * (11.0 * d2f + d3f) * f_2 * u1j - 736.0 * (d - 1.0) * df * f_2 * u1j +
                           384.0 * (d - 1.0) * f_3 * u1j - 8.0 * d2c * g13 * du * du * f_5 * u1j +
                           8.0 * d2c * g13 * df * du * f_4 * u_2 - 8.0 * d2c * g13 * (d2u + 5.0 * du) * f_5 * u_2 +
                           8.0 * d2c * g13 * df * f_4 * u_3 - 24.0 * d2c * g13 * f_5 * u_3 +
                           d26 * (3.0 + 29.0 * d - 19.0 * d * d + 3.0 * d * d * d) * f_7 * u_5 -
                           240.0 * (d - 1.0) * df * df * df * (du + u1j) +
                           80.0 * (d - 1.0) * df * f *
                               (2.0 * d2f * du + df * (3.0 * d2u + 11.0 * du) + 2.0 * (d2f + 4.0 * df) * u1j)) /
                          (den * f_7);
        const double v5 = u5;
        U[p] = x1 * u1j + x2 * u2 + x3 * u3 + x4 * u4 + x5 * u5;
        V[p] = x1 * v1 + x2 * v2 + x3 * v3 + x4 * v4 + x5 * v5;
        F[p] = f + x2 * f2 + x4 * f4;
    }
    S.fields_to_state(U, V, F, Y);
}

// ICG.cpp:252-667, serial branch 456-624 (right, order-3 horizon series)
template <int N, int T>
__device__ void right_expansion(Spec<N, T> &S, double XRight, double d, const double (&u0)[N / T],
                                cx<double> (&Y)[N / (2 * T)])
{
    constexpr int P = N / T;
    using C = cx<double>;
    const double dx = XRight - 1.0;
    const double d2c = (d - 2.0) * (d - 2.0) * (d - 2.0);
    double du0[P], d2u0[P];
    {
        C X[P];
        S.r2c(u0, X);
        S.diff(X);
        C X2[P];
#pragma unroll
        for (int p = 0; p < P; ++p) X2[p] = X[p];
        S.c2r(X, du0);
        S.diff(X2);
        S.c2r(X2, d2u0);
    }
    double c1[P], c2[P], ia20[P], v0[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        c1[p] = 3.0 - d - d2c * u0[p] * u0[p] / 4.0;
        c2[p] = d - 3.0;
    }
    S.solve_inhom(c1, c2, ia20);
#pragma unroll
    for (int p = 0; p < P; ++p) {
        c1[p] = (6.0 - 2.0 * d + (d - 2.0) * ia20[p]) / (2.0 * ia20[p]);
        c2[p] = (d - 2.0) * u0[p] / 2.0;
    }
    S.solve_inhom(c1, c2, v0);
    double f1[P], ia21[P], u1[P], v1[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        f1[p] = 3.0 - d + (d - 3.0) / ia20[p];
        ia21[p] = -ia20[p] * (8.0 * (d - 3.0) + d2c * (u0[p] * u0[p] + v0[p] * v0[p])) / 8.0 + d - 3.0;
        u1[p] = ((d - 2.0 - 2.0 * (d - 3.0) / ia20[p]) * u0[p] + (d - 2.0) * v0[p]) / 4.0 - du0[p] / 2.0;
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const double i0 = ia20[p], i02 = i0 * i0;
        c1[p] = (6.0 - 2.0 * d + d * i0 - 2.0 * f1[p] * i0) / (2.0 * i0);
        c2[p] = (2.0 * (3.0 - d) * (f1[p] - 1.0) * i0 * v0[p] + 2.0 * (d - 3.0) * ia21[p] * v0[p] +
                 (d - 2.0) * i02 * ((f1[p] - 1.0) * u0[p] + u1[p] + (f1[p] - 1.0) * v0[p])) /
                (2.0 * i02);
    }
    S.solve_inhom(c1, c2, v1);
    double f2[P], ia22[P], u2[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const double i0 = ia20[p], i02 = i0 * i0;
        f2[p] = (d - 3.0) * ((f1[p] - 1.0) * (1.0 - i0) * i0 - ia21[p]) / (2.0 * i02);
        ia22[p] = (3.0 - d - (ia21[p] - i0) * (8.0 * (d - 3.0) + d2c * (u0[p] * u0[p] + v0[p] * v0[p])) / 8.0 -
                   d2c * i0 * (u0[p] * u1[p] + v0[p] * v1[p]) / 4.0) /
                  2.0;
        const double A = 4.0 * (3.0 - d) * (d - 6.0 + f1[p]) * i0;
        const double B = (d - 2.0) * (d - 8.0 + 2.0 * f1[p]) * i02;
        const double Cc = 4.0 * (d - 3.0) * (d - 3.0 + 2.0 * ia21[p]);
        const double D = -(d - 3.0) * d2c * i0 * (u0[p] * u0[p] * u0[p]);
        const double E = (4.0 * (6.0 - 2.0 * d + (d - 2.0) * i0) * u1[p] +
                          (d - 2.0) * v0[p] * (6.0 - 2.0 * d + (d - 8.0 + 2.0 * f1[p]) * i0) - 8.0 * i0 * v1[p] +
                          4.0 * d * i0 * v1[p] - 12.0 * du0[p] + 4.0 * d * du0[p] + 8.0 * i0 * du0[p] -
                          2.0 * d * i0 * du0[p] + 4.0 * f1[p] * i0 * du0[p] + 4.0 * i0 * d2u0[p]) *
                         i0;
        u2[p] = (A + B + Cc) * u0[p] + D + E;
        u2[p] /= 32.0 * i02;
    }
    double du2[P];
    S.dtau(u2, du2);
    double v2[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const double i0 = ia20[p], i02 = i0 * i0, i03 = i02 * i0;
        c1[p] = 1.0 + d / 2.0 - 2.0 * f1[p] + (3.0 - d) / i0;
        c2[p] = (2.0 * (3.0 - d) * ia21[p] * ia21[p] * v0[p] +
                 2.0 * (d - 3.0) * i02 * ((f1[p] - 1.0 - f2[p]) * v0[p] - (f1[p] - 1.0) * v1[p]) +
                 ((d - 2.0) * ((1.0 - f1[p] + f2[p]) * u0[p] + (f1[p] - 1.0) * u1[p] + u2[p] +
                               (1.0 - f1[p] + f2[p]) * v0[p]) +
                  ((d - 2.0) * (f1[p] - 1.0) - 2.0 * f2[p]) * v1[p]) *
                     i03 +
                 2.0 * (d - 3.0) * i0 * (ia22[p] * v0[p] + ia21[p] * ((f1[p] - 1.0) * v0[p] + v1[p]))) /
                (2.0 * i03);
    }
    S.solve_inhom(c1, c2, v2);
    double f3[P], u3[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const double i0 = ia20[p], i02 = i0 * i0, i03 = i02 * i0, i04 = i03 * i0;
        const double a1 = ia21[p], a2 = ia22[p];
        f3[p] = ((-3.0 + d) * ((1.0 - f1[p] + f2[p]) * i02 + (-1.0 + f1[p] - f2[p]) * i03 + a1 * a1 -
                               i0 * ((-1.0 + f1[p]) * a1 + a2))) /
                (3.0 * i03);
        const double ia23 =
            (-3.0 + d +
             (-((i0 - a1 + a2) * (8.0 * (-3.0 + d) + d2c * (u0[p] * u0[p] + v0[p] * v0[p]))) +
              2.0 * d2c * (i0 - a1) * (u0[p] * u1[p] + v0[p] * v1[p]) -
              d2c * i0 * (u1[p] * u1[p] + 2.0 * u0[p] * u2[p] + v1[p] * v1[p] + 2.0 * v0[p] * v2[p])) /
                 8.0) /
            3.0;
        u3[p] = -(16.0 * (-3.0 + d) * a1 * a1 * u0[p] +
                  4.0 * (-3.0 + d) * i0 *
                      ((-3.0 + d + f1[p] * (-3.0 + d - 2.0 * a1) + 6.0 * a1 - 4.0 * a2) * u0[p] - 4.0 * a1 * u1[p]) -
                  (d - 3.0) * i02 *
                      (4.0 * (-10.0 + d + f1[p] * (-1.0 + d + f1[p]) - 2.0 * f2[p]) * u0[p] +
                       d2c * (1.0 + f1[p]) * (u0[p] * u0[p] * u0[p]) +
                       2.0 * (-4.0 * (-3.0 + f1[p]) * u1[p] - 8.0 * u2[p] +
                              (1.0 + f1[p]) * (-2.0 * du0[p] + (d - 2.0) * v0[p]))) +
                  i03 * (8.0 * du0[p] - 2.0 * d * du0[p] + 12.0 * du0[p] * f1[p] - 2.0 * d * du0[p] * f1[p] +
                         4.0 * du0[p] * f1[p] * f1[p] + 4.0 * d2u0[p] * (1.0 + f1[p]) - 8.0 * du0[p] * f2[p] +
                         (d - 2.0) * (-16.0 + d + f1[p] * (2.0 + d + 2.0 * f1[p]) - 4.0 * f2[p]) * u0[p] -
                         4.0 * (d - 2.0) * (-3.0 + f1[p]) * u1[p] + 16.0 * u2[p] - 8.0 * d * u2[p] + 32.0 * v0[p] -
                         18.0 * d * v0[p] + d * d * v0[p] - 4.0 * f1[p] * v0[p] + d * d * f1[p] * v0[p] -
                         4.0 * f1[p] * f1[p] * v0[p] + 2.0 * d * f1[p] * f1[p] * v0[p] + 8.0 * f2[p] * v0[p] -
                         4.0 * d * f2[p] * v0[p] - 24.0 * v1[p] + 12.0 * d * v1[p] + 8.0 * f1[p] * v1[p] -
                         4.0 * d * f1[p] * v1[p] + 16.0 * v2[p] - 8.0 * d * v2[p] + 16.0 * du2[p])) /
                (96.0 * i03);
        c1[p] = 2.0 + d / 2.0 - 3.0 * f1[p] + (3.0 - d) / i0;
        c2[p] = ((d - 3.0) * a1 * a1 * a1 * v0[p]) / i04 -
                ((d - 3.0) * a1 * (2.0 * a2 * v0[p] + a1 * ((-1.0 + f1[p]) * v0[p] + v1[p]))) / i03 -
                ((d - 3.0) * ((-1.0 + f1[p] - f2[p] + f3[p]) * v0[p] + (1.0 - f1[p] + f2[p]) * v1[p] +
                              (-1.0 + f1[p]) * v2[p])) /
                    i0 +
                ((d - 2.0) * (-1.0 + f1[p] - f2[p] + f3[p]) * u0[p] - (d - 2.0) * (-1.0 + f1[p] - f2[p]) * u1[p] +
                 2.0 * u2[p] - d * u2[p] - 2.0 * f1[p] * u2[p] + d * f1[p] * u2[p] - 2.0 * u3[p] + d * u3[p] +
                 2.0 * v0[p] - d * v0[p] - 2.0 * f1[p] * v0[p] + d * f1[p] * v0[p] + 2.0 * f2[p] * v0[p] -
                 d * f2[p] * v0[p] - 2.0 * f3[p] * v0[p] + d * f3[p] * v0[p] - 2.0 * v1[p] + d * v1[p] +
                 2.0 * f1[p] * v1[p] - d * f1[p] * v1[p] - 2.0 * f2[p] * v1[p] + d * f2[p] * v1[p] -
                 2.0 * f3[p] * v1[p] + ((d - 2.0) * (-1.0 + f1[p]) - 4.0 * f2[p]) * v2[p]) /
                    2.0 +
                ((d - 3.0) * (ia23 * v0[p] + a2 * ((-1.0 + f1[p]) * v0[p] + v1[p]) +
                              a1 * ((1.0 - f1[p] + f2[p]) * v0[p] + (-1.0 + f1[p]) * v1[p] + v2[p]))) /
                    i02;
    }
    double v3[P];
    S.solve_inhom(c1, c2, v3);
    double U[P], V[P], F[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        F[p] = 1.0 + dx * f1[p] + dx * dx * f2[p] + dx * dx * dx * f3[p];
        U[p] = u0[p] + dx * u1[p] + dx * dx * u2[p] + dx * dx * dx * u3[p];
        V[p] = v0[p] + dx * v1[p] + dx * dx * v2[p] + dx * dx * dx * v3[p];
    }
    S.fields_to_state(U, V, F, Y);
}

// ICG.cpp:891-932 state -> (U, V, F) then ICG.cpp:678-711 pack into Zout
// (n doubles).  Yh: state slots; the mode-2 fix value comes through LDS.
template <int N, int T>
__device__ void state_to_packed(Spec<N, T> &S, const cx<double> (&Yh)[N / (2 * T)], double *Zout)
{
    constexpr int P = N / T, H = P / 2, M = N / 2;
    using C = cx<double>;
    // Re of state mode M-2 lives in thread T-2, slot H-1
    if (S.t == T - 2) S.bc[2] = Yh[H - 1].x;
    __syncthreads();
    const double fix = S.bc[2];
    C z[P];
    S.dbl(Yh, z);
    if (S.t == 2) z[0].x = -fix;                    // ICG.cpp:899
    S.bwd(z);
    double U[P], V[P], F[P];
#pragma unroll
    for (int p = 0; p < H; ++p) {
        U[p] = 0.5 * (z[p].x - z[p + H].x);
        V[p] = 0.5 * (z[p].y - z[p + H].y);
        F[p] = 0.5 * (z[p].y + z[p + H].y);
        U[p + H] = -U[p];
        V[p + H] = -V[p];
        F[p + H] = F[p];
    }
    C X[P];
    S.r2c(U, X);
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int k = S.kidx(p);
        if ((k & 1) && k < N / 4) {
            Zout[k - 1] = X[p].x;
            Zout[k] = X[p].y;
        }
    }
    S.r2c(V, X);
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int k = S.kidx(p);
        if ((k & 1) && k < N / 4) {
            Zout[N / 4 + k - 1] = X[p].x;
            Zout[N / 4 + k] = X[p].y;
        }
    }
    S.r2c(F, X);
    double fold = 0.0;
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int k = S.kidx(p);
        if (!(k & 1) && k < N / 4) {
            Zout[N / 2 + k] = X[p].x;
            if (k != 0) Zout[N / 2 + k + 1] = X[p].y;
        }
        if (k == N / 4 || k == 3 * N / 4) fold += X[p].x;
    }
    if (S.t == (N / 4) % T) Zout[N / 2 + 1] = fold;   // Re(X[N/4] + X[3N/4]), ICG.cpp:709-710
    (void)M;
}

}  // namespace ccd
