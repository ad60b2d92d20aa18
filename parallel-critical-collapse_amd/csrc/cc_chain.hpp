// cc_chain.hpp -- the hot path: IRK2 integration of many shooting chains.
//
// One workgroup integrates one chain (one side of one Jacobian column) from
// its boundary to the match point, entirely on chip: the state y, the two
// stage iterates and the two stage derivatives live in registers (state
// slots: thread t owns modes t + T*h, h < P/2), every FFT is the register/LDS
// Stockham transform of cc_device.hpp.  Each workgroup runs its own
// fixed-point loop, so every chain stops iterating exactly when the
// reference would (ODEStepper.cpp:271-275) and chains never wait on each
// other.
//
// Restated reference routines:
//   ODEStepper::stepIRK             ODEStepper.cpp:163-292
//   ODEStepper::computeDerivatives  ODEStepper.cpp:92-136
//   ICG::StateVectorToFields        InitialConditionGenerator.cpp:791-884
//   ICG::FieldsToStateVector        InitialConditionGenerator.cpp:766-778
//   SpectralTransformer::solveInhom SpectralTransformer.cpp:264-292
//   ShootingSolver::integrateToMatchPoint ShootingSolver.cpp:53-145
//
// The two stages' RHS evaluations are independent within one fixed-point
// iteration, so they are evaluated together: the real-input transforms of
// solveInhom for stage 0 and stage 1 are packed into one complex transform
// (a + i b), which makes 10 length-N transforms per iteration instead of the
// reference's 18 (9 per RHS in its serial/OpenMP build).
#pragma once
#include "cc_device.hpp"

namespace ccd {

template <typename R>
struct ChainParams {
    cx<R> *Y;                 // [nchains][M] states, in/out
    const int32_t *start;     // grid index where the chain starts
    const int32_t *nsteps;    // number of steps of the chain
    const int32_t *dir;       // +1 forward, -1 backward
    const double *delta;      // Delta of the chain (FP64)
    const double *grid;       // x-grid
    int32_t *status;          // 0 ok, 1 IRK no convergence
    int32_t *fail_idx;        // grid index of the failing step
    int64_t *its;             // sum of itsReached over the chain
    const cx<R> *tw;          // twiddles e^{2 pi i e/N}
    int32_t step_begin, step_end;  // chunk of each chain's step sequence
    int32_t maxIts;
    double Dim;
    double precision;
    double a00, a01, a10, a11, b0, b1, c0, c1;   // Butcher tableau
};

template <typename R>
struct RhsConst {
    R dm2, dm3, c3, two_dm3, mg, q2;   // D-2, D-3, (D-2)^3, 2(D-3), -(D-3), (k0 N/2)^2
};

// Evaluate f_s = RHS(yk_s, xs_s) for both stages (ODEStepper.cpp:92-136).
// wk[p]  = m k0 at the thread's spectral slot p (0 at Nyquist)
// imk[p] = 1/(m k0), 0 at m = 0 and at Nyquist
template <int N, int T, typename R>
__device__ __forceinline__ void rhs_pair(const cx<R> (&yk)[2][N / T / 2], cx<R> (&f)[2][N / T / 2],
                                         const R (&xs)[2], const RhsConst<R> &K, const R (&wk)[N / T],
                                         const R (&imk)[N / T], cx<R> *xb, R *bc, const cx<R> *__restrict__ tw,
                                         int t)
{
    constexpr int P = N / T, H = P / 2, M = N / 2;
    const R half = (R)0.5;
    const R invN = (R)1 / (R)N;

    // (1) doubled spectrum + mode-2 fix (ICG.cpp:797-800), tau-derivative (803)
    cx<R> z[2][2][P];   // [stage][field | derivative][p]
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const R fix = bc[s];
#pragma unroll
        for (int h = 0; h < H; ++h) {
            const int q = t + T * h;
            const cx<R> Y = yk[s][h];
            cx<R> lo, hi;
            if (q < M / 2) {
                lo = Y;
                hi = {(R)0, (R)0};
            } else if (q == M / 2) {
                lo = cscale(Y, half);
                hi = Y;
            } else {
                lo = {(R)0, (R)0};
                hi = Y;
            }
            if (q == 2) lo.x = -fix;
            z[s][0][h] = lo;
            z[s][0][h + H] = hi;
        }
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const cx<R> v = z[s][0][p];
            z[s][1][p] = {wk[p] * v.y, -(wk[p] * v.x)};
        }
    }
    // (2) back to tau (804-805)
    fft_block<N, T, -1, 2>(z[0], xb, tw, t);
    fft_block<N, T, -1, 2>(z[1], xb, tw, t);

    // (3) split by half-period symmetry (814-827) and IA2 coefficients (834-836)
    R U[2][H], V[2][H], F[2][H], dU[2][H], dV[2][H], dF[2][H];
    cx<R> c[P];
#pragma unroll
    for (int h = 0; h < H; ++h) {
        R C1[2];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const cx<R> a = z[s][0][h], b = z[s][0][h + H];
            const cx<R> a2 = z[s][1][h], b2 = z[s][1][h + H];
            U[s][h] = half * (a.x - b.x);
            V[s][h] = half * (a.y - b.y);
            F[s][h] = half * (a.y + b.y);
            dU[s][h] = half * (a2.x - b2.x);
            dV[s][h] = half * (a2.y - b2.y);
            dF[s][h] = half * (a2.y + b2.y);
            const R X = xs[s];
            C1[s] = -(((X + F[s][h]) * U[s][h] * U[s][h] + (X - F[s][h]) * V[s][h] * V[s][h]) * K.c3 /
                          ((R)8 * X) +
                      K.dm3);
        }
        c[h] = {C1[0], C1[1]};
        c[h + H] = c[h];
    }

    // (4) solveInhom step 1 for both stages: F' = Coeff1 (264-270)
    fft1<N, T, +1>(c, xb, tw, t);
#pragma unroll
    for (int p = 0; p < P; ++p) c[p] = cscale(c[p], invN);
    if (t == 0) {                       // lambda_s = mean of Coeff1_s
        bc[2] = c[0].x;
        bc[3] = c[0].y;
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {      // lamIntegrate, lambda = 0: G = f / (-i m k0)
        const cx<R> v = c[p];
        c[p] = {-(v.y * imk[p]), v.x * imk[p]};
    }
    fft1<N, T, -1>(c, xb, tw, t);
    R Fr[2][P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        Fr[0][p] = c[p].x;
        Fr[1][p] = c[p].y;
    }
    // (5) h = -g e^{F} (276-279), forward (283)
#pragma unroll
    for (int p = 0; p < P; ++p) c[p] = {K.mg * exp(Fr[0][p]), K.mg * exp(Fr[1][p])};
    fft1<N, T, +1>(c, xb, tw, t);
#pragma unroll
    for (int p = 0; p < P; ++p) c[p] = cscale(c[p], invN);
    // (6) unpack the two Hermitian spectra, lamIntegrate with lambda_s (284)
#pragma unroll
    for (int p = 0; p < P; ++p) xb[t + T * p] = c[p];
    __syncthreads();
    cx<R> cm[P];
#pragma unroll
    for (int p = 0; p < P; ++p) cm[p] = xb[(N - (t + T * p)) & (N - 1)];
    __syncthreads();
    const R lam0 = bc[2], lam1 = bc[3];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int k = t + T * p;
        const cx<R> A = c[p];
        const cx<R> B = {cm[p].x, -cm[p].y};
        const cx<R> S0 = cscale(A + B, half);
        const cx<R> Dd = A - B;
        const cx<R> S1 = {half * Dd.y, -(half * Dd.x)};
        cx<R> G[2];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const cx<R> hs = s ? S1 : S0;
            const R lam = s ? lam1 : lam0;
            if (k == 0) {
                G[s] = (lam != (R)0) ? cx<R>{hs.x / lam, hs.y / lam} : cx<R>{(R)0, (R)0};
            } else if (k == N / 2) {
                const R fac = lam / (lam * lam + K.q2);
                G[s] = cscale(hs, fac);
            } else {
                const R w = wk[p];
                const R den = lam * lam + w * w;
                if (sqrt(den) < (R)1e-14) {
                    G[s] = {(R)0, (R)0};
                } else {
                    const R inv = (R)1 / den;
                    G[s] = {(hs.x * lam - hs.y * w) * inv, (hs.y * lam + hs.x * w) * inv};
                }
            }
        }
        c[p] = {G[0].x - G[1].y, G[0].y + G[1].x};
    }
    fft1<N, T, -1>(c, xb, tw, t);
    R IA2[2][P];
#pragma unroll
    for (int p = 0; p < P; ++p) {                   // X = e^{-F} Y (288-291)
        IA2[0][p] = c[p].x * exp(-Fr[0][p]);
        IA2[1][p] = c[p].y * exp(-Fr[1][p]);
    }

    // (7) pointwise RHS (ODEStepper.cpp:109-122)
    cx<R> yp[2][P];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const R x = xs[s];
        const R two_x = (R)2 * x;
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const int ph = (p < H) ? p : p - H;
            const R sg = (p < H) ? (R)1 : (R)-1;
            const R Uv = sg * U[s][ph], Vv = sg * V[s][ph], Fv = F[s][ph];
            const R dUv = sg * dU[s][ph], dVv = sg * dV[s][ph];
            const R ia = IA2[s][p];
            const R cc = K.two_dm3 / ia - K.dm2;
            const R dUdx = (Fv * (K.dm2 * Vv - cc * Uv) - two_x * dUv) / (two_x * (Fv + x));
            const R dVdx = (Fv * (K.dm2 * Uv - cc * Vv) + two_x * dVv) / (two_x * (Fv - x));
            const R dFdx = K.dm3 * Fv * ((R)1 - ia) / (ia * x);
            yp[s][p] = {dUdx, dVdx + dFdx};
        }
        (void)dF;
    }
    // (8) FieldsToStateVector (766-778) and Re dY[2] = 0 (ODEStepper.cpp:128)
    fft_block<N, T, +1, 2>(yp, xb, tw, t);
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int h = 0; h < H; ++h) {
            const int q = t + T * h;
            const cx<R> lo = cscale(yp[s][h], invN), hi = cscale(yp[s][h + H], invN);
            cx<R> v = (q < M / 2) ? lo : ((q == M / 2) ? lo + hi : hi);
            if (q == 2) v.x = (R)0;
            f[s][h] = v;
        }
}

template <int N, int T, typename R>
__global__ __launch_bounds__(T) void k_chain(ChainParams<R> a)
{
    constexpr int P = N / T, H = P / 2, M = N / 2;
    __shared__ cx<R> xb[2 * N];
    __shared__ R bc[8];
    __shared__ R red[(T / 64) > 0 ? T / 64 : 1];
    const int t = threadIdx.x;
    const int ch = blockIdx.x;
    if (a.status[ch] != 0) return;
    const int nst = a.nsteps[ch];
    const int s_begin = a.step_begin;
    const int s_end = min(a.step_end, nst);
    if (s_begin >= s_end) return;

    cx<R> *Yg = a.Y + (size_t)ch * M;
    cx<R> y[H], yk[2][H], f[2][H];
#pragma unroll
    for (int h = 0; h < H; ++h) y[h] = Yg[t + T * h];

    const double k0d = 2.0 * 3.14159265358979323846 / a.delta[ch];
    const R k0 = (R)k0d;
    RhsConst<R> K;
    K.dm2 = (R)(a.Dim - 2.0);
    K.dm3 = (R)(a.Dim - 3.0);
    K.c3 = (R)((a.Dim - 2.0) * (a.Dim - 2.0) * (a.Dim - 2.0));
    K.two_dm3 = (R)(2.0 * (a.Dim - 3.0));
    K.mg = (R)(-(a.Dim - 3.0));
    {
        const double q = k0d * (double)N / 2.0;
        K.q2 = (R)(q * q);
    }
    R wk[P], imk[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int k = t + T * p;
        const int m = wavenumber<N>(k);
        wk[p] = (k == N / 2) ? (R)0 : (R)m * k0;
        imk[p] = (k == 0 || k == N / 2) ? (R)0 : (R)1 / ((R)m * k0);
    }

    const int dir = a.dir[ch];
    int idx = a.start[ch] + dir * s_begin;
    int64_t its_sum = 0;
    int32_t st = 0;
    for (int step = s_begin; step < s_end; ++step, idx += dir) {
        const double xin = a.grid[idx], xout = a.grid[idx + dir];
        const double dxd = xout - xin;
        const R dx = (R)dxd;
        const R xs[2] = {(R)(xin + dxd * a.c0), (R)(xin + dxd * a.c1)};
        const R A00 = (R)(dxd * a.a00), A01 = (R)(dxd * a.a01), A10 = (R)(dxd * a.a10), A11 = (R)(dxd * a.a11);
        const R B0 = (R)(dxd * a.b0), B1 = (R)(dxd * a.b1);
#pragma unroll
        for (int h = 0; h < H; ++h) yk[0][h] = yk[1][h] = y[h];
        if (t == T - 2) bc[0] = bc[1] = y[H - 1].x;   // Re of mode M-2 for the fix
        __syncthreads();
        bool conv = false;
        int its = 0;
        for (its = 0; its < a.maxIts; ++its) {
            rhs_pair<N, T, R>(yk, f, xs, K, wk, imk, xb, bc, a.tw, t);
            R loc = (R)0;
#pragma unroll
            for (int h = 0; h < H; ++h) {
                cx<R> n0 = y[h], n1 = y[h];
                n0.x += A00 * f[0][h].x; n0.y += A00 * f[0][h].y;
                n0.x += A01 * f[1][h].x; n0.y += A01 * f[1][h].y;
                n1.x += A10 * f[0][h].x; n1.y += A10 * f[0][h].y;
                n1.x += A11 * f[1][h].x; n1.y += A11 * f[1][h].y;
                const R d0x = yk[0][h].x - n0.x, d0y = yk[0][h].y - n0.y;
                const R d1x = yk[1][h].x - n1.x, d1y = yk[1][h].y - n1.y;
                loc += d0x * d0x + d0y * d0y + d1x * d1x + d1y * d1y;
                yk[0][h] = n0;
                yk[1][h] = n1;
            }
            if (t == T - 2) {
                bc[0] = yk[0][H - 1].x;
                bc[1] = yk[1][H - 1].x;
            }
            const R norm2 = block_sum<T, R>(loc, red, t);      // barrier publishes bc[0..1]
            const double rms = sqrt((double)norm2 / (double)(N * 2));
            double tol = a.precision;
            if (its > a.maxIts / 2) tol = a.precision * 10.0 * (2.0 * (double)its / (double)a.maxIts - 0.5);
            if (rms < tol) {
                conv = true;
                break;
            }
        }
        its_sum += (conv ? its + 1 : its);
        if (!conv) {
            st = 1;
            if (t == 0) a.fail_idx[ch] = idx;
            break;
        }
#pragma unroll
        for (int h = 0; h < H; ++h) {                   // ODEStepper.cpp:279-285
            y[h].x += B0 * f[0][h].x; y[h].y += B0 * f[0][h].y;
            y[h].x += B1 * f[1][h].x; y[h].y += B1 * f[1][h].y;
        }
        __syncthreads();
    }
#pragma unroll
    for (int h = 0; h < H; ++h) Yg[t + T * h] = y[h];
    if (t == 0) {
        a.its[ch] += its_sum;
        if (st) a.status[ch] = st;
    }
}

// computeDerivatives on pairs of states (unit entry point).  Y, dY: [2*npairs][M];
// xs, delta: [2*npairs].
template <int N, int T, typename R>
__global__ __launch_bounds__(T) void k_rhs_pairs(const cx<R> *Y, const double *xs, const double *delta,
                                                 cx<R> *dY, const cx<R> *tw, double Dim)
{
    constexpr int P = N / T, H = P / 2, M = N / 2;
    __shared__ cx<R> xb[2 * N];
    __shared__ R bc[8];
    const int t = threadIdx.x;
    const int pr = blockIdx.x;
    cx<R> yk[2][H], f[2][H];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int h = 0; h < H; ++h) yk[s][h] = Y[(size_t)(2 * pr + s) * M + t + T * h];
    if (t == T - 2) {
        bc[0] = yk[0][H - 1].x;
        bc[1] = yk[1][H - 1].x;
    }
    __syncthreads();
    // both states of a pair must share Delta (the stages of one step do)
    const double k0d = 2.0 * 3.14159265358979323846 / delta[2 * pr];
    const R k0 = (R)k0d;
    RhsConst<R> K;
    K.dm2 = (R)(Dim - 2.0);
    K.dm3 = (R)(Dim - 3.0);
    K.c3 = (R)((Dim - 2.0) * (Dim - 2.0) * (Dim - 2.0));
    K.two_dm3 = (R)(2.0 * (Dim - 3.0));
    K.mg = (R)(-(Dim - 3.0));
    const double q = k0d * (double)N / 2.0;
    K.q2 = (R)(q * q);
    R wk[P], imk[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int k = t + T * p;
        const int m = wavenumber<N>(k);
        wk[p] = (k == N / 2) ? (R)0 : (R)m * k0;
        imk[p] = (k == 0 || k == N / 2) ? (R)0 : (R)1 / ((R)m * k0);
    }
    const R xsr[2] = {(R)xs[2 * pr], (R)xs[2 * pr + 1]};
    rhs_pair<N, T, R>(yk, f, xsr, K, wk, imk, xb, bc, tw, t);
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int h = 0; h < H; ++h) dY[(size_t)(2 * pr + s) * M + t + T * h] = f[s][h];
}

}  // namespace ccd
