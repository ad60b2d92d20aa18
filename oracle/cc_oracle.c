/*
 * cc_oracle.c -- CPU FP64 restatement of the reference's Newton-shooting
 * path (TEST INFRASTRUCTURE: the checker and the CPU baseline; never part of
 * the product path).  See cc_oracle.h for conventions.
 *
 * Parity anchors (what pins each function) are listed in the test files under
 * tests/ that exercise it.
 */
#include "cc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* Context                                                                  */
/* ------------------------------------------------------------------------ */
struct cco_ctx {
    cco_problem p;
    int N, M, Nn;
    double Dim;
    double Delta;          /* ICG's mutable Delta (ICG.cpp:41, 255, 798) */
    double k0;             /* SpectralTransformer ctor k0 = 2 pi / period */
    int stages;
    double a[3][3], b[3], cnode[3];
    /* FFT plan */
    int logn;
    int *rev;
    double *cs;            /* cos/sin of 2 pi e / N, e < N/2, interleaved */
    double *fbuf;          /* N complex scratch for the transform */
    /* general workspace */
    double *w;             /* large scratch */
    int64_t n_steps, n_its, n_rhs;
};

static void *xmalloc(size_t bytes)
{
    void *p = calloc(1, bytes ? bytes : 1);
    return p;
}

/* ODEStepper ctor tableau, ODEStepper.cpp:22-80 */
static void set_tableau(cco_ctx *c, int scheme)
{
    memset(c->a, 0, sizeof(c->a));
    switch (scheme) {
    case 1:
        c->stages = 1;
        c->a[0][0] = 0.5; c->b[0] = 1.0; c->cnode[0] = 0.5;
        break;
    case 3:
        c->stages = 3;
        c->a[0][0] = 5.0 / 36.0;
        c->a[0][1] = 2.0 / 9.0 - 1.0 / sqrt(15.0);
        c->a[0][2] = 5.0 / 36.0 - 0.5 / sqrt(15.0);
        c->a[1][0] = 5.0 / 36.0 + sqrt(15.0) / 24.0;
        c->a[1][1] = 2.0 / 9.0;
        c->a[1][2] = 5.0 / 36.0 - sqrt(15.0) / 24.0;
        c->a[2][0] = 5.0 / 36.0 + 0.5 / sqrt(15.0);
        c->a[2][1] = 2.0 / 9.0 + 1.0 / sqrt(15.0);
        c->a[2][2] = 5.0 / 36.0;
        c->b[0] = c->b[2] = 5.0 / 18.0;
        c->b[1] = 4.0 / 9.0;
        c->cnode[0] = 0.5 - sqrt(15.0) / 10.0;
        c->cnode[1] = 0.5;
        c->cnode[2] = 0.5 + sqrt(15.0) / 10.0;
        break;
    default:
        c->stages = 2;
        c->a[0][0] = 0.25;
        c->a[0][1] = 0.25 - 0.5 / sqrt(3.0);
        c->a[1][0] = 0.25 + 0.5 / sqrt(3.0);
        c->a[1][1] = 0.25;
        c->b[0] = c->b[1] = 0.5;
        c->cnode[0] = 0.5 - 0.5 / sqrt(3.0);
        c->cnode[1] = 0.5 + 0.5 / sqrt(3.0);
        break;
    }
}

cco_ctx *cco_create(const cco_problem *p)
{
    if (!p || p->Ntau < 8 || (p->Ntau & (p->Ntau - 1))) return NULL;
    cco_ctx *c = (cco_ctx *)xmalloc(sizeof(cco_ctx));
    c->p = *p;
    c->N = p->Ntau; c->M = c->N / 2; c->Nn = 3 * c->N / 4;
    c->Dim = p->Dim;
    c->Delta = p->Delta;
    c->k0 = 2.0 * M_PI / p->Delta;
    set_tableau(c, p->scheme);
    int N = c->N;
    c->logn = 0; while ((1 << c->logn) < N) ++c->logn;
    c->rev = (int *)xmalloc(sizeof(int) * N);
    for (int i = 0; i < N; ++i) {
        int r = 0;
        for (int bbit = 0; bbit < c->logn; ++bbit) if (i & (1 << bbit)) r |= 1 << (c->logn - 1 - bbit);
        c->rev[i] = r;
    }
    c->cs = (double *)xmalloc(sizeof(double) * N);
    for (int e = 0; e < N / 2; ++e) {
        double ang = 2.0 * M_PI * (double)e / (double)N;
        c->cs[2 * e] = cos(ang);
        c->cs[2 * e + 1] = sin(ang);
    }
    c->fbuf = (double *)xmalloc(sizeof(double) * 2 * N);
    c->w = (double *)xmalloc(sizeof(double) * 170 * N);
    return c;
}

void cco_destroy(cco_ctx *c)
{
    if (!c) return;
    free(c->rev); free(c->cs); free(c->fbuf); free(c->w);
    free(c);
}

int cco_tableau(const cco_ctx *c, double *a, double *b, double *cnode)
{
    for (int i = 0; i < c->stages; ++i) {
        b[i] = c->b[i];
        cnode[i] = c->cnode[i];
        for (int k = 0; k < c->stages; ++k) a[i * c->stages + k] = c->a[i][k];
    }
    return c->stages;
}

void cco_counters(const cco_ctx *c, int64_t out[3])
{
    out[0] = c->n_steps; out[1] = c->n_its; out[2] = c->n_rhs;
}

void cco_reset_counters(cco_ctx *c) { c->n_steps = c->n_its = c->n_rhs = 0; }

/* ------------------------------------------------------------------------ */
/* FFT: X_k = sum_j x_j exp(sign * 2 pi i j k / N), radix-2 DIT.            */
/* The reference's forwardFFT is sign=+1 then /N, backwardFFT sign=-1        */
/* unscaled (SpectralTransformer.cpp:33-38, 70, 110).                        */
/* ------------------------------------------------------------------------ */
static void fft_exec(cco_ctx *c, const double *in, double *out, int sign)
{
    const int N = c->N;
    double *a = c->fbuf;
    for (int i = 0; i < N; ++i) {
        int r = c->rev[i];
        a[2 * r] = in[2 * i];
        a[2 * r + 1] = in[2 * i + 1];
    }
    for (int len = 2; len <= N; len <<= 1) {
        const int half = len >> 1, step = N / len;
        for (int i = 0; i < N; i += len) {
            for (int j = 0; j < half; ++j) {
                const double wr = c->cs[2 * j * step];
                const double wi = sign * c->cs[2 * j * step + 1];
                double *u = a + 2 * (i + j), *v = a + 2 * (i + j + half);
                const double tr = v[0] * wr - v[1] * wi;
                const double ti = v[0] * wi + v[1] * wr;
                const double ur = u[0], ui = u[1];
                u[0] = ur + tr; u[1] = ui + ti;
                v[0] = ur - tr; v[1] = ui - ti;
            }
        }
    }
    memcpy(out, a, sizeof(double) * 2 * N);
}

/* SpectralTransformer.cpp:57-72 */
void cco_forward_r2c(cco_ctx *c, const double *in, double *out)
{
    const int N = c->N;
    double *t = c->w; /* first 2N of workspace reserved for FFT staging */
    for (int i = 0; i < N; ++i) { t[2 * i] = in[i]; t[2 * i + 1] = 0.0; }
    fft_exec(c, t, out, +1);
    for (int i = 0; i < 2 * N; ++i) out[i] = out[i] / (double)N;
}

/* SpectralTransformer.cpp:77-92 */
void cco_backward_c2r(cco_ctx *c, const double *in, double *out)
{
    const int N = c->N;
    double *t = c->w;
    fft_exec(c, in, t, -1);
    for (int i = 0; i < N; ++i) out[i] = t[2 * i];
}

/* SpectralTransformer.cpp:97-112 */
void cco_forward_c2c(cco_ctx *c, const double *in, double *out)
{
    const int N = c->N;
    fft_exec(c, in, out, +1);
    for (int i = 0; i < 2 * N; ++i) out[i] = out[i] / (double)N;
}

/* SpectralTransformer.cpp:117-132 */
void cco_backward_c2c(cco_ctx *c, const double *in, double *out)
{
    fft_exec(c, in, out, -1);
}

/* SpectralTransformer.cpp:140-158: out_k = -i m k0 in_k, Nyquist -> 0 */
void cco_differentiate(cco_ctx *c, const double *in, double *out, size_t size, double period)
{
    const double k0c = (period == 0.0) ? c->k0 : 2 * M_PI / period;
    for (size_t k = 0; k < size; ++k) {
        if (k != size / 2) {
            const int m = (k < size / 2) ? (int)k : (int)k - (int)size;
            const double w = (double)m * k0c;
            const double re = in[2 * k], im = in[2 * k + 1];
            /* -(0 + i w)(re + i im) = (w im) - i (w re) */
            out[2 * k] = w * im;
            out[2 * k + 1] = -(w * re);
        } else {
            out[2 * k] = 0.0;
            out[2 * k + 1] = 0.0;
        }
    }
}

/* complex division (a+ib)/(cr+i ci) by Smith's algorithm, the form libgcc's
 * __divdc3 uses for std::complex<double> division */
static inline void cdiv(double a, double b, double cr, double ci, double *qr, double *qi)
{
    if (fabs(cr) < fabs(ci)) {
        const double ratio = cr / ci;
        const double denom = (cr * ratio) + ci;
        *qr = ((a * ratio) + b) / denom;
        *qi = ((b * ratio) - a) / denom;
    } else {
        const double ratio = ci / cr;
        const double denom = (ci * ratio) + cr;
        *qr = ((b * ratio) + a) / denom;
        *qi = (b - (a * ratio)) / denom;
    }
}

/* SpectralTransformer.cpp:165-189: (lambda - i m k0) G_m = f_m */
void cco_lam_integrate(cco_ctx *c, const double *fk, double *out, size_t size,
                       double lam_re, double lam_im, double period)
{
    const double k0c = (period == 0.0) ? c->k0 : 2 * M_PI / period;
    if (lam_re != 0.0 || lam_im != 0.0) {
        cdiv(fk[0], fk[1], lam_re, lam_im, &out[0], &out[1]);
    } else {
        out[0] = 0.0; out[1] = 0.0;
    }
    for (size_t k = 1; k < size; ++k) {
        if (k != size / 2) {
            const int m = (k < size / 2) ? (int)k : (int)k - (int)size;
            const double dr = lam_re, di = lam_im - (double)m * k0c;
            if (hypot(dr, di) < 1e-14) {
                out[2 * k] = 0.0; out[2 * k + 1] = 0.0;
            } else {
                cdiv(fk[2 * k], fk[2 * k + 1], dr, di, &out[2 * k], &out[2 * k + 1]);
            }
        } else {
            /* lambda / (lambda^2 + (k0 size/2)^2) * f, complex lambda */
            const double l2r = lam_re * lam_re - lam_im * lam_im;
            const double l2i = 2.0 * lam_re * lam_im;
            const double q = k0c * (double)size / 2.0;
            double fr, fi;
            cdiv(lam_re, lam_im, l2r + q * q, l2i, &fr, &fi);
            const double xr = fk[2 * k], xi = fk[2 * k + 1];
            out[2 * k] = fr * xr - fi * xi;
            out[2 * k + 1] = fr * xi + fi * xr;
        }
    }
}

/* SpectralTransformer.cpp:196-217 (out may alias in) */
void cco_halve_modes(const double *in, size_t n_in, double *out)
{
    const size_t n_out = n_in / 2;
    double tmp[2 * n_out];
    for (size_t k = 0; k < n_out / 2; ++k) { tmp[2 * k] = in[2 * k]; tmp[2 * k + 1] = in[2 * k + 1]; }
    tmp[2 * (n_out / 2)] = in[2 * (n_out / 2)] + in[2 * (3 * n_out / 2)];
    tmp[2 * (n_out / 2) + 1] = in[2 * (n_out / 2) + 1] + in[2 * (3 * n_out / 2) + 1];
    for (size_t k = n_out / 2 + 1; k < n_out; ++k) {
        tmp[2 * k] = in[2 * (n_out + k)]; tmp[2 * k + 1] = in[2 * (n_out + k) + 1];
    }
    memcpy(out, tmp, sizeof(double) * 2 * n_out);
}

/* SpectralTransformer.cpp:224-252.  Note the loop at 242-245 starts at
 * N_in/2 and so overwrites out[3 N_in/2] with the full in[N_in/2]; only
 * out[N_in/2] keeps the half amplitude.  Restated as written. */
void cco_double_modes(const double *in, size_t n_in, double *out)
{
    const size_t n_out = 2 * n_in;
    double tmp[2 * n_out];
    for (size_t k = 0; k < n_in / 2; ++k) { tmp[2 * k] = in[2 * k]; tmp[2 * k + 1] = in[2 * k + 1]; }
    tmp[2 * (n_in / 2)] = 0.5 * in[2 * (n_in / 2)];
    tmp[2 * (n_in / 2) + 1] = 0.5 * in[2 * (n_in / 2) + 1];
    tmp[2 * (3 * n_in / 2)] = 0.5 * in[2 * (n_in / 2)];
    tmp[2 * (3 * n_in / 2) + 1] = 0.5 * in[2 * (n_in / 2) + 1];
    for (size_t k = n_in / 2; k < n_in; ++k) {
        tmp[2 * (n_in + k)] = in[2 * k]; tmp[2 * (n_in + k) + 1] = in[2 * k + 1];
    }
    for (size_t k = n_in / 2 + 1; k < 3 * n_in / 2; ++k) { tmp[2 * k] = 0.0; tmp[2 * k + 1] = 0.0; }
    memcpy(out, tmp, sizeof(double) * 2 * n_out);
}

/* SpectralTransformer.cpp:264-292: periodic solution of x' + f x + g = 0 */
void cco_solve_inhom(cco_ctx *c, const double *f, const double *g, double *x, double period)
{
    const int N = c->N;
    double *f_hat = c->w + 2 * N;      /* 2N */
    double *F_hat = c->w + 4 * N;      /* 2N */
    double *Freal = c->w + 6 * N;      /* N */
    double *h = c->w + 7 * N;          /* N */
    double *h_hat = c->w + 8 * N;      /* 2N */
    double *H_hat = c->w + 10 * N;     /* 2N */
    cco_forward_r2c(c, f, f_hat);
    cco_lam_integrate(c, f_hat, F_hat, (size_t)N, 0.0, 0.0, period);
    cco_backward_c2r(c, F_hat, x);
    memcpy(Freal, x, sizeof(double) * N);
    for (int i = 0; i < N; ++i) h[i] = -g[i] * exp(Freal[i]);
    cco_forward_r2c(c, h, h_hat);
    cco_lam_integrate(c, h_hat, H_hat, (size_t)N, f_hat[0], f_hat[1], period);
    cco_backward_c2r(c, H_hat, x);
    for (int i = 0; i < N; ++i) x[i] *= exp(-Freal[i]);
}

/* ------------------------------------------------------------------------ */
/* InitialConditionGenerator                                                 */
/* Workspace: solve_inhom uses w[0, 12N); everything below uses w[12N, ...). */
/* ------------------------------------------------------------------------ */
#define WS(c, slot) ((c)->w + (size_t)(12 + (slot)) * (size_t)(c)->N)

/* forward -> differentiate -> backward (real in, real out) */
static void tau_derivative(cco_ctx *c, const double *in, double *out, double period)
{
    const int N = c->N;
    double *hat = WS(c, 30), *dhat = WS(c, 32);   /* 2N each */
    cco_forward_r2c(c, in, hat);
    cco_differentiate(c, hat, dhat, (size_t)N, period);
    cco_backward_c2r(c, dhat, out);
}

/* ICG.cpp:37-239 */
void cco_left_expansion(cco_ctx *c, double XLeft, const double *fc, const double *psic,
                        double *Y, double DeltaIn)
{
    const int N = c->N;
    c->Delta = DeltaIn;
    const double *f0 = fc, *psi0 = psic;
    double *df0 = WS(c, 0), *d2f0 = WS(c, 1), *d3f0 = WS(c, 2);
    double *coeff1 = WS(c, 3), *coeff2 = WS(c, 4), *u1 = WS(c, 5);
    double *du1 = WS(c, 6), *d2u1 = WS(c, 7), *d3u1 = WS(c, 8), *d4u1 = WS(c, 9);
    double *U = WS(c, 10), *V = WS(c, 11), *F = WS(c, 12);
    const double Delta = c->Delta;

    tau_derivative(c, f0, df0, Delta);      /* 51-53 */
    tau_derivative(c, df0, d2f0, Delta);    /* 55-57 */
    tau_derivative(c, d2f0, d3f0, Delta);   /* 59-61 */

    for (int j = 0; j < N; ++j) { coeff1[j] = 1.0; coeff2[j] = -(c->Dim - 1.0) * psi0[j] * f0[j]; }
    cco_solve_inhom(c, coeff1, coeff2, u1, Delta);   /* 74 */

    tau_derivative(c, u1, du1, Delta);      /* 80-82 */
    tau_derivative(c, du1, d2u1, Delta);    /* 84-86 */
    tau_derivative(c, d2u1, d3u1, Delta);   /* 88-90 */
    tau_derivative(c, d3u1, d4u1, Delta);   /* 92-94 */

    for (int j = 0; j < N; ++j) {          /* 105-191 */
        const double f = f0[j], psi = psi0[j], d = c->Dim;
        const double u1j = u1[j];
        const double du = du1[j], d2u = d2u1[j], d3u = d3u1[j], d4u = d4u1[j];
        const double df = df0[j], d2f = d2f0[j], d3f = d3f0[j];

        const double v1 = u1j;
        const double f2 = ((d - 3.0) * pow(d - 2.0, 3) * f * pow(u1j, 2)) / (8.0 * (d - 1.0));
        const double u2 = -psi;
        const double v2 = psi;

        const double common3 =
              ((d - 3.0) * pow(d - 2.0, 3) * pow(f, 3) * pow(u1j, 3)
            + 4.0 * df * (u1j + du)
            - 4.0 * f * (2.0 * u1j + 3.0 * du + d2u));
        const double u3 = common3 / (8.0 * (1.0 - d) * pow(f, 3));
        const double v3 = u3;

        const double f4 = -((d - 3.0) * pow(d - 2.0, 3) *
                 ((pow(d - 2.0, 3) * (5.0 - 6.0 * d + d * d) * pow(f, 3) * pow(u1j, 4))
                + 8.0 * (d - 1.0) * df * u1j * (du + u1j)
                - 8.0 * f * (pow(du, 2)
                + ((d - 1.0) * d2u + (-1.0 + 3.0 * d) * du) * u1j
                + (-1.0 + 2.0 * d) * pow(u1j, 2))))
                / (128.0 * pow(d - 1.0, 2) * (1.0 + d) * pow(f, 2));

        const double u4 = (
            -192.0 * (d - 1.0) * pow(df, 2) * du * pow(f, 2)
            + 192.0 * (d - 1.0) * d2u * df * pow(f, 3)
            + 64.0 * (d - 1.0) * d2f * du * pow(f, 3)
            + 640.0 * (d - 1.0) * df * du * pow(f, 3)
            - 64.0 * (d - 1.0) * (6.0 * d2u + d3u) * pow(f, 4)
            - 704.0 * (d - 1.0) * du * pow(f, 4)
            - 192.0 * (d - 1.0) * pow(df, 2) * pow(f, 2) * u1j
            + 64.0 * (d - 1.0) * d2f * pow(f, 3) * u1j
            + 448.0 * (d - 1.0) * df * pow(f, 3) * u1j
            - 384.0 * (d - 1.0) * pow(f, 4) * u1j
            + 32.0 * pow(d - 2.0, 3) * (3.0 - 7.0 * d + 2.0 * d * d)
              * pow(f, 6) * pow(u1j, 2) * (du + u1j)
        ) / (128.0 * pow(d - 1.0, 2) * (1.0 + d) * pow(f, 7));
        const double v4 = -u4;

        const double u5 = (
            -64.0 * (d - 1.0) * d2f * d2u * pow(f, 2)
            - 16.0 * (d - 1.0) * (15.0 * d2f + d3f) * du * pow(f, 2)
            - 32.0 * (d - 1.0) * df * (20.0 * d2u + 3.0 * d3u + 40.0 * du) * pow(f, 2)
            + 560.0 * (d - 1.0) * d2u * pow(f, 3)
            + 160.0 * (d - 1.0) * d3u * pow(f, 3)
            + 16.0 * (d - 1.0) * d4u * pow(f, 3)
            + 800.0 * (d - 1.0) * du * pow(f, 3)
            - 16.0 * (d - 1.0) * (11.0 * d2f + d3f) * pow(f, 2) * u1j
            - 736.0 * (d - 1.0) * df * pow(f, 2) * u1j
            + 384.0 * (d - 1.0) * pow(f, 3) * u1j
            - 8.0 * pow(d - 2.0, 3) * (3.0 - 13.0 * d + 4.0 * d * d)
              * pow(du, 2) * pow(f, 5) * u1j
            + 8.0 * pow(d - 2.0, 3) * (3.0 - 13.0 * d + 4.0 * d * d)
              * df * du * pow(f, 4) * pow(u1j, 2)
            - 8.0 * pow(d - 2.0, 3) * (3.0 - 13.0 * d + 4.0 * d * d)
              * (d2u + 5.0 * du) * pow(f, 5) * pow(u1j, 2)
            + 8.0 * pow(d - 2.0, 3) * (3.0 - 13.0 * d + 4.0 * d * d)
              * df * pow(f, 4) * pow(u1j, 3)
            - 24.0 * pow(d - 2.0, 3) * (3.0 - 13.0 * d + 4.0 * d * d)
              * pow(f, 5) * pow(u1j, 3)
            + pow(d - 2.0, 6) * (3.0 + 29.0 * d - 19.0 * d * d + 3.0 * d * d * d)
              * pow(f, 7) * pow(u1j, 5)
            - 240.0 * (d - 1.0) * pow(df, 3) * (du + u1j)
            + 80.0 * (d - 1.0) * df * f * (
                2.0 * d2f * du
              + df * (3.0 * d2u + 11.0 * du)
              + 2.0 * (d2f + 4.0 * df) * u1j)
        ) / (128.0 * pow(d - 1.0, 2) * (1.0 + d) * pow(f, 7));
        const double v5 = u5;

        U[j] = XLeft * u1j + pow(XLeft, 2) * u2 + pow(XLeft, 3) * u3
             + pow(XLeft, 4) * u4 + pow(XLeft, 5) * u5;
        V[j] = XLeft * v1 + pow(XLeft, 2) * v2 + pow(XLeft, 3) * v3
             + pow(XLeft, 4) * v4 + pow(XLeft, 5) * v5;
        F[j] = f + pow(XLeft, 2) * f2 + pow(XLeft, 4) * f4;
    }
    cco_fields_to_state(c, U, V, F, Y);     /* 194 */
    Y[4] = Delta;                           /* 195: Re Y[2] = Delta */
}

/* ICG.cpp:252-667, serial branch 456-624 */
void cco_right_expansion(cco_ctx *c, double XRight, const double *Up, double *Y, double DeltaIn)
{
    const int N = c->N;
    c->Delta = DeltaIn;
    const double Delta = c->Delta;
    const double Xp = 1.0, d = c->Dim, dx = XRight - Xp;
    const double *u0 = Up;
    double *du0 = WS(c, 0), *d2u0 = WS(c, 1);
    double *coeff1 = WS(c, 2), *coeff2 = WS(c, 3), *ia20 = WS(c, 4), *v0 = WS(c, 5);
    double *f1 = WS(c, 6), *u1 = WS(c, 7), *ia21 = WS(c, 8), *v1 = WS(c, 9);
    double *f2 = WS(c, 10), *ia22 = WS(c, 11), *u2 = WS(c, 12), *du2 = WS(c, 13), *v2 = WS(c, 14);
    double *f3 = WS(c, 15), *ia23 = WS(c, 16), *u3 = WS(c, 17), *v3 = WS(c, 18);
    double *F = WS(c, 19), *U = WS(c, 20), *V = WS(c, 21);
    double *hat = WS(c, 22), *dhat = WS(c, 24), *d2hat = WS(c, 26);   /* 2N each */

    /* 266-271: du0 = iFFT(D FFT u0); d2u0 = iFFT(D D FFT u0) */
    cco_forward_r2c(c, u0, hat);
    cco_differentiate(c, hat, dhat, (size_t)N, Delta);
    cco_backward_c2r(c, dhat, du0);
    cco_differentiate(c, dhat, d2hat, (size_t)N, Delta);
    cco_backward_c2r(c, d2hat, d2u0);

    for (int j = 0; j < N; ++j) {
        coeff1[j] = 3.0 - d - pow(d - 2.0, 3) * u0[j] * u0[j] / 4.0;
        coeff2[j] = d - 3.0;
    }
    cco_solve_inhom(c, coeff1, coeff2, ia20, Delta);

    for (int j = 0; j < N; ++j) {
        coeff1[j] = (6.0 - 2.0 * d + (d - 2.0) * ia20[j]) / (2.0 * ia20[j]);
        coeff2[j] = (d - 2.0) * u0[j] / 2.0;
    }
    cco_solve_inhom(c, coeff1, coeff2, v0, Delta);

    for (int j = 0; j < N; ++j) {
        f1[j] = 3.0 - d + (d - 3.0) / ia20[j];
        ia21[j] = -ia20[j] * (8.0 * (d - 3.0) + pow(d - 2.0, 3) * (u0[j] * u0[j] + v0[j] * v0[j])) / 8.0 + d - 3.0;
        u1[j] = ((d - 2.0 - 2.0 * (d - 3.0) / ia20[j]) * u0[j] + (d - 2.0) * v0[j]) / 4.0 - du0[j] / 2.0;
    }

    for (int j = 0; j < N; ++j) {
        coeff1[j] = (6.0 - 2.0 * d + d * ia20[j] - 2.0 * f1[j] * ia20[j]) / (2.0 * ia20[j]);
        coeff2[j] =
            (2.0 * (3.0 - d) * (f1[j] - 1.0) * ia20[j] * v0[j]
            + 2.0 * (d - 3.0) * ia21[j] * v0[j]
            + (d - 2.0) * pow(ia20[j], 2) * ((f1[j] - 1.0) * u0[j] + u1[j] + (f1[j] - 1.0) * v0[j])
            ) / (2.0 * pow(ia20[j], 2));
    }
    cco_solve_inhom(c, coeff1, coeff2, v1, Delta);

    for (int j = 0; j < N; ++j)
        f2[j] = (d - 3.0) * ((f1[j] - 1.0) * (1.0 - ia20[j]) * ia20[j] - ia21[j]) / (2.0 * pow(ia20[j], 2));

    for (int j = 0; j < N; ++j)
        ia22[j] = (3.0 - d
            - (ia21[j] - ia20[j]) * (8.0 * (d - 3.0) + pow(d - 2.0, 3) * (pow(u0[j], 2) + pow(v0[j], 2))) / 8.0
            - pow(d - 2.0, 3) * ia20[j] * (u0[j] * u1[j] + v0[j] * v1[j]) / 4.0) / 2.0;

    for (int j = 0; j < N; ++j) {
        const double A = 4.0 * (3.0 - d) * (d - 6.0 + f1[j]) * ia20[j];
        const double B = (d - 2.0) * (d - 8.0 + 2.0 * f1[j]) * pow(ia20[j], 2);
        const double C = 4.0 * (d - 3.0) * (d - 3.0 + 2.0 * ia21[j]);
        const double D = -(d - 3.0) * pow(d - 2.0, 3) * ia20[j] * pow(u0[j], 3);
        const double E = (4.0 * (6.0 - 2.0 * d + (d - 2.0) * ia20[j]) * u1[j]
                        + (d - 2.0) * v0[j] * (6.0 - 2.0 * d + (d - 8.0 + 2.0 * f1[j]) * ia20[j])
                        - 8.0 * ia20[j] * v1[j] + 4.0 * d * ia20[j] * v1[j]
                        - 12.0 * du0[j] + 4.0 * d * du0[j]
                        + 8.0 * ia20[j] * du0[j] - 2.0 * d * ia20[j] * du0[j]
                        + 4.0 * f1[j] * ia20[j] * du0[j]
                        + 4.0 * ia20[j] * d2u0[j]) * ia20[j];
        u2[j] = (A + B + C) * u0[j] + D + E;
        u2[j] /= 32.0 * pow(ia20[j], 2);
    }

    /* 518-520 */
    cco_forward_r2c(c, u2, hat);
    cco_differentiate(c, hat, dhat, (size_t)N, Delta);
    cco_backward_c2r(c, dhat, du2);

    for (int j = 0; j < N; ++j) {
        coeff1[j] = 1.0 + d / 2.0 - 2.0 * f1[j] + (3.0 - d) / ia20[j];
        coeff2[j] =
            (2.0 * (3.0 - d) * pow(ia21[j], 2) * v0[j]
            + 2.0 * (d - 3.0) * pow(ia20[j], 2) * ((f1[j] - 1.0 - f2[j]) * v0[j] - (f1[j] - 1.0) * v1[j])
            + ((d - 2.0) * ((1.0 - f1[j] + f2[j]) * u0[j] + (f1[j] - 1.0) * u1[j] + u2[j] + (1.0 - f1[j] + f2[j]) * v0[j])
               + ((d - 2.0) * (f1[j] - 1.0) - 2.0 * f2[j]) * v1[j]) * pow(ia20[j], 3)
            + 2.0 * (d - 3.0) * ia20[j] * (ia22[j] * v0[j] + ia21[j] * ((f1[j] - 1.0) * v0[j] + v1[j]))
            ) / (2.0 * pow(ia20[j], 3));
    }
    cco_solve_inhom(c, coeff1, coeff2, v2, Delta);

    for (int j = 0; j < N; ++j) {
        f3[j] = ((-3.0 + d) * ((1.0 - f1[j] + f2[j]) * pow(ia20[j], 2)
             + (-1.0 + f1[j] - f2[j]) * pow(ia20[j], 3)
             + pow(ia21[j], 2)
             - ia20[j] * ((-1.0 + f1[j]) * ia21[j] + ia22[j])))
             / (3.0 * pow(ia20[j], 3));

        ia23[j] = (-3.0 + d + (
            -((ia20[j] - ia21[j] + ia22[j]) * (8.0 * (-3.0 + d) + pow(d - 2.0, 3) * (u0[j] * u0[j] + v0[j] * v0[j])))
            + 2.0 * pow(d - 2.0, 3) * (ia20[j] - ia21[j]) * (u0[j] * u1[j] + v0[j] * v1[j])
            - pow(d - 2.0, 3) * ia20[j] * (pow(u1[j], 2) + 2.0 * u0[j] * u2[j] + pow(v1[j], 2) + 2.0 * v0[j] * v2[j])
            ) / 8.0) / 3.0;

        u3[j] = -(16.0 * (-3.0 + d) * pow(ia21[j], 2) * u0[j]
            + 4.0 * (-3.0 + d) * ia20[j]
                * ((-3.0 + d + f1[j] * (-3.0 + d - 2.0 * ia21[j])
                + 6.0 * ia21[j] - 4.0 * ia22[j]) * u0[j]
                - 4.0 * ia21[j] * u1[j])
            - (d - 3.0) * pow(ia20[j], 2)
                * (4.0 * (-10.0 + d + f1[j] * (-1.0 + d + f1[j]) - 2.0 * f2[j]) * u0[j]
                + pow(d - 2.0, 3) * (1.0 + f1[j]) * pow(u0[j], 3)
                + 2.0 * (-4.0 * (-3.0 + f1[j]) * u1[j] - 8.0 * u2[j]
                + (1.0 + f1[j]) * (-2.0 * du0[j] + (d - 2.0) * v0[j])))
            + pow(ia20[j], 3)
                * (8.0 * du0[j] - 2.0 * d * du0[j]
                + 12.0 * du0[j] * f1[j] - 2.0 * d * du0[j] * f1[j]
                + 4.0 * du0[j] * pow(f1[j], 2)
                + 4.0 * d2u0[j] * (1.0 + f1[j])
                - 8.0 * du0[j] * f2[j]
                + (d - 2.0) * (-16.0 + d + f1[j] * (2.0 + d + 2.0 * f1[j]) - 4.0 * f2[j]) * u0[j]
                - 4.0 * (d - 2.0) * (-3.0 + f1[j]) * u1[j]
                + 16.0 * u2[j] - 8.0 * d * u2[j]
                + 32.0 * v0[j] - 18.0 * d * v0[j] + pow(d, 2) * v0[j]
                - 4.0 * f1[j] * v0[j] + pow(d, 2) * f1[j] * v0[j]
                - 4.0 * pow(f1[j], 2) * v0[j]
                + 2.0 * d * pow(f1[j], 2) * v0[j]
                + 8.0 * f2[j] * v0[j] - 4.0 * d * f2[j] * v0[j]
                - 24.0 * v1[j] + 12.0 * d * v1[j]
                + 8.0 * f1[j] * v1[j] - 4.0 * d * f1[j] * v1[j]
                + 16.0 * v2[j] - 8.0 * d * v2[j]
                + 16.0 * du2[j]))
            / (96.0 * pow(ia20[j], 3));
    }

    for (int j = 0; j < N; ++j) {
        coeff1[j] = 2.0 + d / 2.0 - 3.0 * f1[j] + (3.0 - d) / ia20[j];
        coeff2[j] =
            ((d - 3.0) * pow(ia21[j], 3) * v0[j]) / pow(ia20[j], 4)
            - ((d - 3.0) * ia21[j]
            * (2.0 * ia22[j] * v0[j]
            + ia21[j] * ((-1.0 + f1[j]) * v0[j] + v1[j])))
            / pow(ia20[j], 3)
            - ((d - 3.0) * ((-1.0 + f1[j] - f2[j] + f3[j]) * v0[j]
            + (1.0 - f1[j] + f2[j]) * v1[j]
            + (-1.0 + f1[j]) * v2[j])) / ia20[j]
            + ((d - 2.0) * (-1.0 + f1[j] - f2[j] + f3[j]) * u0[j]
            - (d - 2.0) * (-1.0 + f1[j] - f2[j]) * u1[j]
            + 2.0 * u2[j] - d * u2[j]
            - 2.0 * f1[j] * u2[j] + d * f1[j] * u2[j]
            - 2.0 * u3[j] + d * u3[j]
            + 2.0 * v0[j] - d * v0[j]
            - 2.0 * f1[j] * v0[j] + d * f1[j] * v0[j]
            + 2.0 * f2[j] * v0[j] - d * f2[j] * v0[j]
            - 2.0 * f3[j] * v0[j] + d * f3[j] * v0[j]
            - 2.0 * v1[j] + d * v1[j]
            + 2.0 * f1[j] * v1[j] - d * f1[j] * v1[j]
            - 2.0 * f2[j] * v1[j] + d * f2[j] * v1[j]
            - 2.0 * f3[j] * v1[j]
            + ((d - 2.0) * (-1.0 + f1[j]) - 4.0 * f2[j]) * v2[j]) / 2.0
            + ((d - 3.0) * (ia23[j] * v0[j]
            + ia22[j] * ((-1.0 + f1[j]) * v0[j] + v1[j])
            + ia21[j] * ((1.0 - f1[j] + f2[j]) * v0[j]
            + (-1.0 + f1[j]) * v1[j] + v2[j]))) / pow(ia20[j], 2);
    }
    cco_solve_inhom(c, coeff1, coeff2, v3, Delta);

    for (int j = 0; j < N; ++j) {
        F[j] = 1.0 + dx * f1[j] + dx * dx * f2[j] + dx * dx * dx * f3[j];
        U[j] = u0[j] + dx * u1[j] + dx * dx * u2[j] + dx * dx * dx * u3[j];
        V[j] = v0[j] + dx * v1[j] + dx * dx * v2[j] + dx * dx * dx * v3[j];
    }
    cco_fields_to_state(c, U, V, F, Y);     /* 623 */
    Y[4] = Delta;                           /* 624 */
}

/* ICG.cpp:678-711 */
void cco_pack(cco_ctx *c, const double *odd1, const double *odd2, const double *even, double *Z)
{
    const int N = c->N, Nn = c->Nn;
    double *o1 = WS(c, 100), *o2 = WS(c, 102), *ev = WS(c, 104);  /* 2N each */
    cco_forward_r2c(c, odd1, o1);
    cco_forward_r2c(c, odd2, o2);
    cco_forward_r2c(c, even, ev);
    cco_halve_modes(o1, (size_t)N, o1);
    cco_halve_modes(o2, (size_t)N, o2);
    cco_halve_modes(ev, (size_t)N, ev);
    for (int j = 0; j < Nn / 6; ++j) {
        Z[2 * j] = o1[2 * (2 * j + 1)];
        Z[2 * j + 1] = o1[2 * (2 * j + 1) + 1];
        Z[2 * j + Nn / 3] = o2[2 * (2 * j + 1)];
        Z[2 * j + 1 + Nn / 3] = o2[2 * (2 * j + 1) + 1];
        Z[2 * j + 2 * Nn / 3] = ev[2 * (2 * j)];
        Z[2 * j + 1 + 2 * Nn / 3] = ev[2 * (2 * j) + 1];
    }
    Z[2 * Nn / 3 + 1] = ev[2 * (Nn / 3)];
}

/* ICG.cpp:719-758 */
void cco_unpack(cco_ctx *c, const double *Z, double *odd1, double *odd2, double *even)
{
    const int N = c->N, M = c->M, Nn = c->Nn;
    double *o1 = WS(c, 100), *o2 = WS(c, 102), *ev = WS(c, 104);  /* 2N each */
    memset(o1, 0, sizeof(double) * 2 * N);
    memset(o2, 0, sizeof(double) * 2 * N);
    memset(ev, 0, sizeof(double) * 2 * N);
    for (int j = 0; j < Nn / 6; ++j) {
        o1[2 * (2 * j + 1)] = Z[2 * j];
        o1[2 * (2 * j + 1) + 1] = Z[2 * j + 1];
        o1[2 * (M - 2 * j - 1)] = o1[2 * (2 * j + 1)];
        o1[2 * (M - 2 * j - 1) + 1] = -o1[2 * (2 * j + 1) + 1];
        o2[2 * (2 * j + 1)] = Z[2 * j + Nn / 3];
        o2[2 * (2 * j + 1) + 1] = Z[2 * j + 1 + Nn / 3];
        o2[2 * (M - 2 * j - 1)] = o2[2 * (2 * j + 1)];
        o2[2 * (M - 2 * j - 1) + 1] = -o2[2 * (2 * j + 1) + 1];
        ev[2 * (2 * j)] = Z[2 * j + 2 * Nn / 3];
        ev[2 * (2 * j) + 1] = Z[2 * j + 1 + 2 * Nn / 3];
        if (j != 0) {
            ev[2 * (M - 2 * j)] = ev[2 * (2 * j)];
            ev[2 * (M - 2 * j) + 1] = -ev[2 * (2 * j) + 1];
        }
    }
    ev[1] = 0.0;
    cco_double_modes(o1, (size_t)M, o1);
    cco_double_modes(o2, (size_t)M, o2);
    cco_double_modes(ev, (size_t)M, ev);
    ev[2 * (Nn / 3)] = Z[2 * Nn / 3 + 1] / 2.0;
    ev[2 * (Nn / 3) + 1] = 0.0;
    ev[2 * Nn] = Z[2 * Nn / 3 + 1] / 2.0;
    ev[2 * Nn + 1] = 0.0;
    cco_backward_c2r(c, o1, odd1);
    cco_backward_c2r(c, o2, odd2);
    cco_backward_c2r(c, ev, even);
}

/* ICG.cpp:766-778 */
void cco_fields_to_state(cco_ctx *c, const double *U, const double *V, const double *F, double *Y)
{
    const int N = c->N;
    double *t = WS(c, 110), *tf = WS(c, 112);   /* 2N each */
    for (int j = 0; j < N; ++j) { t[2 * j] = U[j]; t[2 * j + 1] = V[j] + F[j]; }
    cco_forward_c2c(c, t, tf);
    cco_halve_modes(tf, (size_t)N, Y);
}

/* the doubled, mode-2-fixed spectrum of a state and its tau derivative
 * (ICG.cpp:798-803) */
static void state_spectra(cco_ctx *c, const double *Y, double *comp1, double *comp2, int with_deriv)
{
    const int N = c->N;
    c->Delta = Y[4];
    cco_double_modes(Y, (size_t)c->M, comp1);
    comp1[4] = -comp1[2 * (N - 2)];      /* 800: Re comp1[2] = -Re comp1[N-2] */
    if (with_deriv) cco_differentiate(c, comp1, comp2, (size_t)N, c->Delta);
}

/* ICG.cpp:791-884 (serial/OpenMP build: the spectra and the first inverse
 * transform are computed twice, 797-805 and 842-847; restated as written) */
void cco_state_to_fields_d(cco_ctx *c, const double *Y, double *U, double *V, double *F,
                           double *IA2, double *dU, double *dV, double *dF, double X)
{
    const int N = c->N, H = N / 2;
    double *comp1 = WS(c, 114), *comp2 = WS(c, 116);   /* 2N each */
    double *coeff1 = WS(c, 118), *coeff2 = WS(c, 119);
    const double D = c->Dim;

    state_spectra(c, Y, comp1, comp2, 1);        /* 797-803 */
    cco_backward_c2c(c, comp1, comp1);           /* 804 */
    cco_backward_c2c(c, comp2, comp2);           /* 805 */
    state_spectra(c, Y, comp1, comp2, 1);        /* 842-846 */
    cco_backward_c2c(c, comp1, comp1);           /* 847 */
    for (int j = 0; j < H; ++j) {
        U[j] = 0.5 * (comp1[2 * j] - comp1[2 * (j + H)]);
        V[j] = 0.5 * (comp1[2 * j + 1] - comp1[2 * (j + H) + 1]);
        F[j] = 0.5 * (comp1[2 * j + 1] + comp1[2 * (j + H) + 1]);
    }
    for (int j = 0; j < H; ++j) { U[j + H] = -U[j]; V[j + H] = -V[j]; F[j + H] = F[j]; }
    cco_backward_c2c(c, comp2, comp2);           /* 862 */
    for (int j = 0; j < H; ++j) {
        dU[j] = 0.5 * (comp2[2 * j] - comp2[2 * (j + H)]);
        dV[j] = 0.5 * (comp2[2 * j + 1] - comp2[2 * (j + H) + 1]);
        dF[j] = 0.5 * (comp2[2 * j + 1] + comp2[2 * (j + H) + 1]);
    }
    for (int j = 0; j < H; ++j) { dU[j + H] = -dU[j]; dV[j + H] = -dV[j]; dF[j + H] = dF[j]; }
    for (int j = 0; j < N; ++j) {
        coeff1[j] = -(((X + F[j]) * U[j] * U[j] + (X - F[j]) * V[j] * V[j]) * pow((D - 2.0), 3) / (8.0 * X)
                      + (D - 3.0));
        coeff2[j] = (D - 3.0);
    }
    cco_solve_inhom(c, coeff1, coeff2, IA2, c->Delta);   /* 882 */
}

/* ICG.cpp:891-932 */
void cco_state_to_fields(cco_ctx *c, const double *Y, double *U, double *V, double *F)
{
    const int N = c->N, H = N / 2;
    double *comp1 = WS(c, 114);
    state_spectra(c, Y, comp1, NULL, 0);
    cco_backward_c2c(c, comp1, comp1);
    for (int j = 0; j < H; ++j) {
        U[j] = 0.5 * (comp1[2 * j] - comp1[2 * (j + H)]);
        V[j] = 0.5 * (comp1[2 * j + 1] - comp1[2 * (j + H) + 1]);
        F[j] = 0.5 * (comp1[2 * j + 1] + comp1[2 * (j + H) + 1]);
    }
    for (int j = 0; j < H; ++j) { U[j + H] = -U[j]; V[j + H] = -V[j]; F[j + H] = F[j]; }
}

/* ------------------------------------------------------------------------ */
/* ODEStepper                                                               */
/* ------------------------------------------------------------------------ */
/* ODEStepper.cpp:92-136.  Yreal and dYreal are N interleaved doubles. */
void cco_compute_derivatives(cco_ctx *c, const double *Yreal, double *dYreal, double x)
{
    const int N = c->N;
    double *U = WS(c, 120), *V = WS(c, 121), *F = WS(c, 122), *IA2 = WS(c, 123);
    double *dUdt = WS(c, 124), *dVdt = WS(c, 125), *dFdt = WS(c, 126);
    double *dUdx = WS(c, 127), *dVdx = WS(c, 128), *dFdx = WS(c, 129);
    const double Dim = c->Dim;
    c->n_rhs++;
    cco_state_to_fields_d(c, Yreal, U, V, F, IA2, dUdt, dVdt, dFdt, x);
    for (int j = 0; j < N; ++j) {
        dUdx[j] = (F[j] * ((Dim - 2.0) * V[j]
                   - (2.0 * (Dim - 3.0) / IA2[j] - (Dim - 2.0)) * U[j])
                   - 2.0 * x * dUdt[j])
                 / (2.0 * x * (F[j] + x));
        dVdx[j] = (F[j] * ((Dim - 2.0) * U[j]
                   - (2.0 * (Dim - 3.0) / IA2[j] - (Dim - 2.0)) * V[j])
                   + 2.0 * x * dVdt[j])
                 / (2.0 * x * (F[j] - x));
        dFdx[j] = (Dim - 3.0) * F[j] * (1.0 - IA2[j]) / (IA2[j] * x);
    }
    cco_fields_to_state(c, dUdx, dVdx, dFdx, dYreal);
    dYreal[4] = 0.0;                       /* 128 */
}

/* ODEStepper.cpp:163-292 */
int cco_step_irk(cco_ctx *c, const double *Yin, double *Yout, double Xin, double Xout,
                 int maxIts, int *its_out)
{
    const int N = c->N, S = c->stages;
    const double dx = Xout - Xin;
    double *y = WS(c, 130);
    double *f[3], *yK1[3], *yK2[3];
    for (int i = 0; i < 3; ++i) { f[i] = WS(c, 131 + i); yK1[i] = WS(c, 134 + i); yK2[i] = WS(c, 137 + i); }
    double xs[3];
    int its_reached = 0, converged = 0;
    memcpy(y, Yin, sizeof(double) * N);
    for (int i = 0; i < S; ++i) xs[i] = Xin + dx * c->cnode[i];
    for (int i = 0; i < S; ++i) memcpy(yK2[i], y, sizeof(double) * N);

    for (int its = 0; its < maxIts; ++its) {
        ++its_reached;
        for (int i = 0; i < S; ++i) memcpy(yK1[i], yK2[i], sizeof(double) * N);
        double norm2 = 0.0;
        for (int i = 0; i < S; ++i) cco_compute_derivatives(c, yK2[i], f[i], xs[i]);
        for (int i = 0; i < S; ++i) {
            for (int j = 0; j < N; ++j) {
                double tmp = y[j];
                for (int k = 0; k < S; ++k) tmp += dx * c->a[i][k] * f[k][j];
                yK2[i][j] = tmp;
            }
        }
        for (int i = 0; i < S; ++i)
            for (int j = 0; j < N; ++j) {
                const double dlt = yK1[i][j] - yK2[i][j];
                norm2 += dlt * dlt;
            }
        norm2 = sqrt(norm2 / (double)(N * S));
        double precision10 = c->p.PrecisionIRK;
        if (its > maxIts / 2)
            precision10 = c->p.PrecisionIRK * 10.0 * (2.0 * (double)its / (double)maxIts - 0.5);
        if (norm2 < precision10) { converged = 1; break; }
    }
    for (int i = 0; i < S; ++i)
        for (int j = 0; j < N; ++j) y[j] += dx * c->b[i] * f[i][j];
    memcpy(Yout, y, sizeof(double) * N);
    c->n_steps++;
    c->n_its += its_reached;
    if (its_out) *its_out = its_reached;
    return converged;
}

/* ------------------------------------------------------------------------ */
/* ShootingSolver                                                           */
/* ------------------------------------------------------------------------ */
/* ShootingSolver.cpp:53-145 */
int cco_integrate(cco_ctx *c, double *Y, const double *grid, size_t start, size_t end, int *fail_idx)
{
    const int N = c->N;
    double *Y2 = WS(c, 141);
    const int maxIts = c->p.MaxIterIRK;
    int its;
    if (start <= end) {
        for (size_t i = start; i < end; ++i) {
            if (!cco_step_irk(c, Y, Y2, grid[i], grid[i + 1], maxIts, &its)) {
                if (fail_idx) *fail_idx = (int)i;
                return CCO_IRK_NOCONV;
            }
            memcpy(Y, Y2, sizeof(double) * N);
        }
    } else {
        for (size_t i = start; i > end; --i) {
            if (!cco_step_irk(c, Y, Y2, grid[i], grid[i - 1], maxIts, &its)) {
                if (fail_idx) *fail_idx = (int)i;
                return CCO_IRK_NOCONV;
            }
            memcpy(Y, Y2, sizeof(double) * N);
        }
    }
    return CCO_OK;
}

/* ShootingSolver.cpp:154-182 */
int cco_mismatch(cco_ctx *c, const double *YL, const double *YR, double *mis)
{
    const int N = c->N;   /* M complex = N doubles */
    for (int i = 0; i < N; ++i) mis[i] = YR[i] - YL[i];
    if (mis[4] > c->p.PrecisionIRK) return CCO_DELTA_MISMATCH;
    mis[4] = YL[4];
    return CCO_OK;
}

/* ------------------------------------------------------------------------ */
/* NewtonSolver                                                             */
/* ------------------------------------------------------------------------ */
/* NewtonSolver.cpp:477-519 */
void cco_generate_grid(const cco_problem *p, double *grid)
{
    const size_t iL = 0, iM = (size_t)p->NLeft, iR = (size_t)(p->NLeft + p->NRight);
    grid[iL] = p->XLeft;
    grid[iM] = p->XMid;
    grid[iR] = p->XRight;
    const double ZLeft = log(p->XLeft) - log(1.0 - p->XLeft);
    const double ZMid = log(p->XMid) - log(1.0 - p->XMid);
    const double ZRight = log(p->XRight) - log(1.0 - p->XRight);
    const double dZL = (ZMid - ZLeft) / (double)p->NLeft;
    const double dZR = (ZRight - ZMid) / (double)p->NRight;
    for (size_t j = iL + 1; j < iM; ++j) {
        const double e = ZLeft + (double)(j - iL) * dZL;
        grid[j] = exp(e) / (1.0 + exp(e));
    }
    for (size_t j = iM + 1; j < iR; ++j) {
        const double e = ZMid + (double)(j - iM) * dZR;
        grid[j] = exp(e) / (1.0 + exp(e));
    }
}

/* NewtonSolver.cpp:403-423 (and the thread-local overload 439-459) */
int cco_shoot_states(cco_ctx *c, const double *grid, const double *x, double *F,
                     double *YL0, double *YR0, double *YLm, double *YRm)
{
    const int N = c->N, Nn = c->Nn;
    double *xin = WS(c, 142);
    double *Up = WS(c, 143), *psic = WS(c, 144), *fc = WS(c, 145);
    double *YL = WS(c, 146), *YR = WS(c, 147), *mis = WS(c, 148);
    double *U = WS(c, 149), *V = WS(c, 150), *Fd = WS(c, 151);
    memcpy(xin, x, sizeof(double) * Nn);
    const double Delta = xin[2 * Nn / 3 + 2];
    xin[2 * Nn / 3 + 2] = 0.0;
    cco_unpack(c, xin, Up, psic, fc);
    cco_left_expansion(c, c->p.XLeft, fc, psic, YL, Delta);
    cco_right_expansion(c, c->p.XRight, Up, YR, Delta);
    if (YL0) memcpy(YL0, YL, sizeof(double) * N);
    if (YR0) memcpy(YR0, YR, sizeof(double) * N);
    const size_t iL = 0, iM = (size_t)c->p.NLeft, iR = (size_t)(c->p.NLeft + c->p.NRight);
    int fail = -1;
    int st = cco_integrate(c, YL, grid, iL, iM, &fail);
    if (st != CCO_OK) return st;
    st = cco_integrate(c, YR, grid, iR, iM, &fail);
    if (st != CCO_OK) return st;
    if (YLm) memcpy(YLm, YL, sizeof(double) * N);
    if (YRm) memcpy(YRm, YR, sizeof(double) * N);
    st = cco_mismatch(c, YL, YR, mis);
    if (st != CCO_OK) return st;
    cco_state_to_fields(c, mis, U, V, Fd);
    cco_pack(c, U, V, Fd, F);
    (void)N;
    return CCO_OK;
}

int cco_shoot(cco_ctx *c, const double *grid, const double *x, double *F)
{
    return cco_shoot_states(c, grid, x, F, NULL, NULL, NULL, NULL);
}

/* NewtonSolver.cpp:615-702 with the OpenMP column loop of 619-666.  Each
 * thread owns its own context, as the reference gives each thread its own
 * InitialConditionGenerator and ShootingSolver (625-627). */
int cco_jacobian(const cco_problem *p, const double *grid, const double *x, const double *F0,
                 double *J, const int32_t *cols, int ncols, int nthreads, int32_t *status,
                 int64_t counters[3])
{
    const int n = 3 * p->Ntau / 4;
    const int ncol = cols ? ncols : n;
    int worst = CCO_OK;
    int64_t tot0 = 0, tot1 = 0, tot2 = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads) reduction(max:worst) reduction(+:tot0,tot1,tot2)
#endif
    {
        cco_ctx *c = cco_create(p);
        double *xp = (double *)malloc(sizeof(double) * n);
        double *Fp = (double *)malloc(sizeof(double) * n);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int q = 0; q < ncol; ++q) {
            const int i = cols ? cols[q] : q;
            memcpy(xp, x, sizeof(double) * n);
            xp[i] += p->EpsNewton;
            const int st = cco_shoot(c, grid, xp, Fp);
            if (status) status[i] = st;
            if (st != CCO_OK) { if (st > worst) worst = st; continue; }
            for (int j = 0; j < n; ++j) J[(size_t)j * n + i] = (Fp[j] - F0[j]) / p->EpsNewton;
        }
        tot0 += c->n_steps; tot1 += c->n_its; tot2 += c->n_rhs;
        free(xp); free(Fp);
        cco_destroy(c);
    }
    (void)nthreads;
    if (counters) { counters[0] = tot0; counters[1] = tot1; counters[2] = tot2; }
    return worst;
}

/* dgesv semantics (NewtonSolver.cpp:739-757): LU with partial pivoting
 * (pivot = first index of max |a|, as LAPACK's idamax), then two triangular
 * solves.  Row-major storage. */
int cco_dgesv(int n, double *A, int32_t *ipiv, double *b)
{
    int info = 0;
    for (int k = 0; k < n; ++k) {
        int piv = k;
        double best = fabs(A[(size_t)k * n + k]);
        for (int i = k + 1; i < n; ++i) {
            const double v = fabs(A[(size_t)i * n + k]);
            if (v > best) { best = v; piv = i; }
        }
        if (ipiv) ipiv[k] = piv + 1;
        if (A[(size_t)piv * n + k] == 0.0) { if (!info) info = k + 1; continue; }
        if (piv != k) {
            for (int j = 0; j < n; ++j) {
                const double t = A[(size_t)k * n + j];
                A[(size_t)k * n + j] = A[(size_t)piv * n + j];
                A[(size_t)piv * n + j] = t;
            }
            const double t = b[k]; b[k] = b[piv]; b[piv] = t;
        }
        const double inv = 1.0 / A[(size_t)k * n + k];
        for (int i = k + 1; i < n; ++i) {
            const double l = A[(size_t)i * n + k] * inv;
            A[(size_t)i * n + k] = l;
            if (l != 0.0)
                for (int j = k + 1; j < n; ++j) A[(size_t)i * n + j] -= l * A[(size_t)k * n + j];
        }
    }
    if (info) return info;
    for (int i = 0; i < n; ++i) {
        double s = b[i];
        for (int j = 0; j < i; ++j) s -= A[(size_t)i * n + j] * b[j];
        b[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = b[i];
        for (int j = i + 1; j < n; ++j) s -= A[(size_t)i * n + j] * b[j];
        b[i] = s / A[(size_t)i * n + i];
    }
    return 0;
}

/* NewtonSolver.cpp:763-768 */
double cco_l2norm(const double *v, size_t n)
{
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += v[i] * v[i];
    return sqrt(s / (double)n);
}
