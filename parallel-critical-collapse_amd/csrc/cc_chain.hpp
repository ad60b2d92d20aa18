// cc_chain.hpp -- the hot path: IRK2 integration of many shooting chains.
//
// One workgroup integrates one chain (one side of one Jacobian column) from
// its boundary to the match point, entirely on chip: the step-start state,
// the two stage iterates and the candidate next state live in registers
// (state slots: thread t owns modes t + T*h, h < P/2), every FFT is the
// register/LDS Stockham transform of cc_device.hpp.  Each workgroup runs its
// own fixed-point loop, so every chain stops iterating exactly when the
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
// Work organisation of one fixed-point iteration (both stages together):
//  * per stage: doubled spectrum + tau-derivative, one batched inverse FFT of
//    2 signals, half-period split; the RHS is rewritten as
//        dU/dx = A_U - B_U r,  dV/dx = A_V - B_V r,  df/dx = C_F (r - 1),
//    r = 1/IA2 = a^2 (ODEStepper.cpp:111-121 regrouped), so only A_U, B_U,
//    A_V, B_V, C_F and Coeff1 survive the split;
//  * solveInhom for both stages at once: stage 0 and stage 1 are packed into
//    one complex transform (a + i b) for each of its four transforms; the
//    lambda = 0 integration commutes with the packing, the lambda = f0
//    integration is applied after separating the two Hermitian spectra;
//    r = e^{F} / X comes from the e^{F} already needed for h = -g e^{F}
//    (SpectralTransformer.cpp:276-291), so no e^{-F} is formed;
//  * per stage: FieldsToStateVector as one batched forward FFT of 2 signals.
// 10 length-N transforms per iteration instead of the reference's 18.
//
// The transforms are the lane-swap DIF/DIT pair of cc_lanefft.hpp: fields on
// the tau grid are in layout L (thread t holds points t + T*p), spectra in
// layout S, where thread t holds modes kt + T*p with kt = lf_kt(t) a digit
// reversal of t.  Every spectral pointwise operation below indexes modes
// through kt; the states in HBM keep the natural mode order.
#pragma once
#include "cc_lanefft.hpp"

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
    const R *invm;            // 1/m at spectral index k (0 at k = 0 and N/2)
    int32_t step_begin, step_end;  // chunk of each chain's step sequence
    int32_t maxIts;
    double Dim;
    double precision;
    double a[3][3], b[3], c[3];    // Gauss-Legendre tableau (ODEStepper.cpp:22-80), stages x stages
    int32_t stages;                // 1, 2 or 3
    const double *dim;             // per-chain D (batched problems), null: Dim for every chain
    const int32_t *goff;           // per-chain offset of its grid in `grid`, null: 0
    void *rec;                     // base-trajectory record (tangent-linear columns), may be null;
                                   // element type RT of k_chain
    const int64_t *rec_base;       // per chain: first record slot, < 0 = do not record
};

// Base-trajectory record of one IRK step (element type RT of k_chain, the
// precision of the tangent that reads it): for each stage s, the fields of
// the last RHS evaluation of the step -- U, V, F, dU/dtau, dV/dtau and F - x
// on j < N/2 (the other half follows by symmetry), r = 1/IA2 and
// E = e^{F_int} on all N points, and lambda = mean(Coeff1).  F - x is
// recorded as computed in the base precision: near the horizon F and x agree
// to three digits, and their difference formed from FP32 copies would keep
// only four.  A tangent-linear chain linearises the RHS around these values.
template <int N>
struct Rec {
    static constexpr int M = N / 2;
    static constexpr int U = 0, V = M, F = 2 * M, DU = 3 * M, DV = 4 * M, FMX = 5 * M, R = 6 * M, E = 6 * M + N,
                         LAM = 6 * M + 2 * N;
    static constexpr int STAGE = 6 * M + 2 * N + 4;
    static constexpr int STEP = 2 * STAGE;
};

template <typename R>
struct RhsConst {
    R k0, inv_k0;                     // 2 pi / Delta and its inverse
    R dm2, dm3, c3, two_dm3, mg, q2;  // D-2, D-3, (D-2)^3, 2(D-3), -(D-3), (k0 N/2)^2
};

// A stage's x and the split's x-dependent factors, formed once per IRK step
// (the fixed-point iterations of a step share them).
template <typename R>
struct StageX {
    R X, inv_x, c3_8x;                // x, 1/x, (D-2)^3 / (8 x)
};
// (x is the same in every lane: the values are moved to scalar registers,
// where they stay across the fixed-point loop instead of being recomputed
// in it to save vector registers)
template <typename R>
__device__ __forceinline__ R lane_uniform(R v)
{
    constexpr int W = (int)(sizeof(R) / 4);
    uint32_t u[W];
    __builtin_memcpy(u, &v, sizeof(R));
#pragma unroll
    for (int w = 0; w < W; ++w) u[w] = (uint32_t)__builtin_amdgcn_readfirstlane((int)u[w]);
    __builtin_memcpy(&v, u, sizeof(R));
    return v;
}
template <typename R>
__device__ __forceinline__ StageX<R> stage_x(R X, const RhsConst<R> &K)
{
    return {lane_uniform(X), lane_uniform((R)1 / X), lane_uniform(K.c3 / ((R)8 * X))};
}

// DB = 1 evaluations run the two stages' first inverse transforms as one
// batch of four signals (one set of exchanges and barriers instead of two; the
// butterflies of each signal are unchanged, so are the results).  Their LDS
// regions then hold 4N complex; rhs_lds gives the exchange buffer a caller
// declares, (1 + DB) regions.
template <int N, typename R, int DB>
constexpr bool rhs_merge() { return DB != 0; }
template <int N, typename R, int DB>
constexpr int rhs_region() { return (rhs_merge<N, R, DB>() ? 4 : 2) * N; }
template <int N, typename R, int DB>
constexpr int rhs_lds() { return (1 + DB) * rhs_region<N, R, DB>(); }

// Evaluate f_s = RHS(yk_s, xs_s) for both stages (ODEStepper.cpp:92-136);
// f returns N f_s (the last forward transform's 1/N is left to the caller).
// ims: LDS, the lamIntegrate factors 1/(m k0 N) of this thread's modes kt + T*p
// at slots t + T*p (lf_ims: set once per launch, read without global latency).
// bc[0], bc[1]: Re of mode M-2 of each stage state (the mode-2 fix value),
// published by the caller before the call.  xb: rhs_lds<N, R, DB>() complex.
template <int N, int T, typename R, typename RT = float, int DB = 0>
__device__ __forceinline__ void rhs_pair(const cx<R> (&yk)[2][N / T / 2], cx<R> (&f)[2][N / T / 2],
                                         const StageX<R> (&sx)[2], const RhsConst<R> &K, const R *ims,
                                         cx<R> *xb, R *bc, const cx<R> *__restrict__ tw, int t,
                                         RT *rec = nullptr)
{
    constexpr int P = N / T, H = P / 2, M = N / 2;
    using RC = Rec<N>;
    const R half = (R)0.5;
    const R invN = (R)1 / (R)N;
    // Everything derived from the thread index is recomputed per call: held
    // across the fixed-point loop it would cost more registers than it saves.
    asm volatile("" : "+v"(t));
    // DB: the transforms of one evaluation follow each other without trailing
    // barriers, each starting in the region its predecessor did not read last
    // (lf_next_par); the spectrum separation takes the next region and the
    // transform after it the other one.  The caller's reduction barrier
    // separates the last transform from the next evaluation.
    constexpr bool MERGE = rhs_merge<N, R, DB>();
    constexpr int RS = rhs_region<N, R, DB>();
    constexpr int pA = 0, pB = DB ? lf_next_par<N, T>(pA) : 0;
    constexpr int pC = DB ? lf_next_par<N, T>(MERGE ? pA : pB) : 0;
    constexpr int pD = DB ? lf_next_par<N, T>(pC) : 0, pE = DB ? lf_next_par<N, T>(pD) : 0;
    constexpr int pS = DB ? lf_next_par<N, T>(pE) : 0, pG = DB ? 1 - pS : 0;
    constexpr int pH = DB ? lf_next_par<N, T>(pG) : 0;
    const int kt = lf_kt<N, T>(t);      // spectral modes of this thread: kt + T*p
    // FP64 chains of the few-chain (double-buffered) variant: the stage
    // twiddles are loaded once per evaluation and held in registers.  The
    // chip-filling variants have no room in their register budget (they
    // would spill) and load per stage.
    constexpr bool PRE = sizeof(R) == 8 && DB != 0;
    LfTw<N, T, R> twr{};
    if constexpr (PRE) twr = lf_twiddles<N, T, R>(tw, t);

    // per-stage products of the split that survive until the pointwise RHS
    R AU[2][H], BU[2][H], AV[2][H], BV[2][H], CF[2][H];
    cx<R> c[P];                                   // packed Coeff1 of both stages
    // (3) half-period split (814-827), IA2 coefficients (834-836), RHS products
    // Divisions (~10 VALU instructions each in FP64) are shared where two are
    // needed together: 1/a and 1/b as b q and a q with q = 1/(a b) (the
    // operands are far from overflow), and the per-stage constants once.
    auto split = [&](int s, const cx<R> (&z)[2][P]) {
        const R X = sx[s].X, inv_x = sx[s].inv_x, c3_8x = sx[s].c3_8x;
#pragma unroll
        for (int h = 0; h < H; ++h) {
            const cx<R> a = z[0][h], b = z[0][h + H];
            const cx<R> a2 = z[1][h], b2 = z[1][h + H];
            const R U = half * (a.x - b.x), V = half * (a.y - b.y), F = half * (a.y + b.y);
            const R dU = half * (a2.x - b2.x), dV = half * (a2.y - b2.y);
            if (rec) {
                RT *rs = rec + s * RC::STAGE + t + T * h;
                rs[RC::U] = (RT)U;
                rs[RC::V] = (RT)V;
                rs[RC::F] = (RT)F;
                rs[RC::DU] = (RT)dU;
                rs[RC::DV] = (RT)dV;
                rs[RC::FMX] = (RT)(F - X);
            }
            const R C1 = -(((X + F) * U * U + (X - F) * V * V) * c3_8x + K.dm3);
            if (s == 0) c[h].x = C1; else c[h].y = C1;
            const R two_x = (R)2 * X;
            const R d1 = two_x * (F + X), d2 = two_x * (F - X);
            const R q12 = rcp_fast(d1 * d2);
            const R inv1 = d2 * q12, inv2 = d1 * q12;
            const R FD = F * K.dm2 * (U + V);
            AU[s][h] = (FD - two_x * dU) * inv1;
            BU[s][h] = K.two_dm3 * F * U * inv1;
            AV[s][h] = (FD + two_x * dV) * inv2;
            BV[s][h] = K.two_dm3 * F * V * inv2;
            CF[s][h] = K.dm3 * F * inv_x;
        }
    };
    cx<R> zz[2][2][P];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        // (1) doubled spectrum + mode-2 fix (ICG.cpp:797-800), tau-derivative (803)
        cx<R>(&z)[2][P] = zz[s];
        const R fix = bc[s];
#pragma unroll
        for (int h = 0; h < H; ++h) {
            const int q = kt + T * h;
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
            z[0][h] = lo;
            z[0][h + H] = hi;
        }
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const int k = kt + T * p;
            const R w = (k == N / 2) ? (R)0 : (R)wavenumber<N>(k) * K.k0;
            const cx<R> v = z[0][p];
            z[1][p] = {w * v.y, -(w * v.x)};
        }
        // (2) back to tau (804-805)
        if constexpr (!MERGE) {
            if (s == 0) lfft_dit<N, T, -1, 2, R, DB, pA, RS, PRE>(z, xb, tw, t, &twr);
            else lfft_dit<N, T, -1, 2, R, DB, pB, RS, PRE>(z, xb, tw, t, &twr);
            split(s, z);
        }
    }
    if constexpr (MERGE) {
        lfft_dit<N, T, -1, 4, R, DB, pA, RS, PRE>(*reinterpret_cast<cx<R>(*)[4][P]>(&zz), xb, tw, t, &twr);
        split(0, zz[0]);
        split(1, zz[1]);
    }
#pragma unroll
    for (int h = 0; h < H; ++h) c[h + H] = c[h];

    // (4) solveInhom step 1 for both stages: F' = Coeff1 (264-270)
    lfft_dif1<N, T, +1, R, DB, pC, RS, PRE>(c, xb, tw, t, &twr);
    // The forward transform's 1/N (a power of two: exact) is folded into the
    // products that follow, here and below.
    if (t == 0) {                       // lambda_s = mean of Coeff1_s
        bc[2] = c[0].x * invN;
        bc[3] = c[0].y * invN;
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {      // lamIntegrate, lambda = 0: G = f / (-i m k0)
        const R im = ims[t + T * p];
        const cx<R> v = c[p];
        c[p] = {-(v.y * im), v.x * im};
    }
    lfft_dit1<N, T, -1, R, DB, pD, RS, PRE>(c, xb, tw, t, &twr);
    // (5) h = -g e^{F} (276-279), forward (283); keep e^{F}
    R E[2][P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        E[0][p] = exp(c[p].x);
        E[1][p] = exp(c[p].y);
        c[p] = {K.mg * E[0][p], K.mg * E[1][p]};
    }
    lfft_dif1<N, T, +1, R, DB, pE, RS, PRE>(c, xb, tw, t, &twr);
    // (6) separate the two Hermitian spectra, lamIntegrate with lambda_s (284)
#pragma unroll
    for (int p = 0; p < P; ++p) xb[pS * RS + lf_mirror_pos<N, T, R>(kt + T * p)] = c[p];
    __syncthreads();
    cx<R> cm[P];
#pragma unroll
    for (int p = 0; p < P; ++p) cm[p] = xb[pS * RS + lf_mirror_pos<N, T, R>((N - (kt + T * p)) & (N - 1))];
    if constexpr (!DB) __syncthreads();
    const R lam0 = bc[2], lam1 = bc[3];
    if (rec && t == 0) {
        rec[RC::LAM] = (RT)lam0;
        rec[RC::STAGE + RC::LAM] = (RT)lam1;
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int k = kt + T * p;
        const cx<R> A = c[p];
        const cx<R> B = {cm[p].x, -cm[p].y};
        const R hn = half * invN;
        const cx<R> S0 = cscale(A + B, hn);
        const cx<R> Dd = A - B;
        const cx<R> S1 = {hn * Dd.y, -(hn * Dd.x)};
        cx<R> G[2];
        if (k == 0) {
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const cx<R> hs = s ? S1 : S0;
                const R lam = s ? lam1 : lam0;
                G[s] = (lam != (R)0) ? cscale(hs, (R)1 / lam) : cx<R>{(R)0, (R)0};
            }
        } else if (k == N / 2) {
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const cx<R> hs = s ? S1 : S0;
                const R lam = s ? lam1 : lam0;
                G[s] = cscale(hs, lam / (lam * lam + K.q2));
            }
        } else {
            const R w = (R)wavenumber<N>(k) * K.k0;
            const R den0 = lam0 * lam0 + w * w, den1 = lam1 * lam1 + w * w;
            // the reference's |lambda + i w| < 1e-14 (SpectralTransformer.cpp:181) as
            // den < 1e-28: the same test without a square root (w != 0 here, so
            // neither form is ever true at a finite period)
            const bool z0 = den0 < (R)1e-28, z1 = den1 < (R)1e-28;
            R inv0, inv1;
            if (!z0 && !z1) {
                const R q = rcp_fast(den0 * den1);
                inv0 = den1 * q;
                inv1 = den0 * q;
            } else {
                inv0 = z0 ? (R)0 : (R)1 / den0;
                inv1 = z1 ? (R)0 : (R)1 / den1;
            }
            G[0] = {(S0.x * lam0 - S0.y * w) * inv0, (S0.y * lam0 + S0.x * w) * inv0};
            G[1] = {(S1.x * lam1 - S1.y * w) * inv1, (S1.y * lam1 + S1.x * w) * inv1};
        }
        c[p] = {G[0].x - G[1].y, G[0].y + G[1].x};
    }
    lfft_dit1<N, T, -1, R, DB, pG, RS, PRE>(c, xb, tw, t, &twr);

    // (7) pointwise RHS with r = 1/IA2 = e^{F}/X (288-291; ODEStepper.cpp:109-122)
    cx<R> yp[2][P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int ph = (p < H) ? p : p - H;
        const R sg = (p < H) ? (R)1 : (R)-1;
        const R qr = rcp_fast(c[p].x * c[p].y);
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const R r = E[s][p] * (s ? c[p].x : c[p].y) * qr;     // E_s / c_s
            if (rec) {
                RT *rs = rec + s * RC::STAGE + t + T * p;
                rs[RC::R] = (RT)r;
                rs[RC::E] = (RT)E[s][p];
            }
            const R cf = CF[s][ph];
            yp[s][p] = {sg * (AU[s][ph] - BU[s][ph] * r), sg * (AV[s][ph] - BV[s][ph] * r) + cf * (r - (R)1)};
        }
    }
    // (8) FieldsToStateVector (766-778) and Re dY[2] = 0 (ODEStepper.cpp:128)
    lfft_dif<N, T, +1, 2, R, DB, pH, RS, PRE>(yp, xb, tw, t, &twr);
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int h = 0; h < H; ++h) {
            const int q = kt + T * h;
            const cx<R> lo = yp[s][h], hi = yp[s][h + H];
            cx<R> v = (q < M / 2) ? lo : ((q == M / 2) ? lo + hi : hi);
            if (q == 2) v.x = (R)0;
            f[s][h] = v;                    // N times the derivative: callers fold in 1/N
        }
}

// The lamIntegrate factors of rhs_pair: ims[t + T*p] = invm[k] / (k0 N) for the
// modes k = kt + T*p of thread t (slots private to the thread: no barrier needed).
template <int N, int T, typename R>
__device__ __forceinline__ void lf_ims(R *ims, const R *__restrict__ invm, R inv_k0, int t)
{
    const int kt = lf_kt<N, T>(t);
    const R s = inv_k0 * ((R)1 / (R)N);
#pragma unroll
    for (int p = 0; p < N / T; ++p) ims[t + T * p] = invm[kt + T * p] * s;
}

template <int N, typename R>
__device__ __forceinline__ RhsConst<R> make_const(double Dim, double delta)
{
    const double k0d = 2.0 * 3.14159265358979323846 / delta;
    RhsConst<R> K;
    K.k0 = (R)k0d;
    K.inv_k0 = (R)(1.0 / k0d);
    K.dm2 = (R)(Dim - 2.0);
    K.dm3 = (R)(Dim - 3.0);
    K.c3 = (R)((Dim - 2.0) * (Dim - 2.0) * (Dim - 2.0));
    K.two_dm3 = (R)(2.0 * (Dim - 3.0));
    K.mg = (R)(-(Dim - 3.0));
    const double q = k0d * (double)N / 2.0;
    K.q2 = (R)(q * q);
    return K;
}

// Waves per SIMD the register budget is cut for: FP64 two-stage chains at
// four values per thread fit 2 (256 VGPRs), at eight values or with another
// stage count 1; FP32 chains fit 3.
template <int N, int T, typename R, int S = 2>
constexpr int chain_waves() { return sizeof(R) == 8 ? ((N / T >= 8 || S != 2) ? 1 : 2) : 3; }

// RT: element type of the base-trajectory record (float for the FP32 tangent,
// double for the FP64 one); unused when a.rec is null.  DB = 1 double-buffers
// the transform exchanges (one barrier each instead of two) at twice the LDS:
// for the few latency-bound base chains, not for the chip-filling FP64
// Jacobian launches that need two workgroups per CU.  Every step starts the
// fixed point from the reference's K_s = y (ODEStepper.cpp:193-199): F(x)
// depends on the start at the level of the iteration tolerance, which the
// backward integration from the horizon amplifies to ~1e-12.  S: stages of
// the Gauss-Legendre scheme (SchemeIRK 1, 2, 3).  Two stages are evaluated
// together (rhs_pair); one stage runs as a pair of equal stages, three as a
// pair plus a pair of equal stages.  Only S = 2 records the trajectory.
template <int N, int T, typename R, typename RT = float, int DB = 0, int S = 2>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(chain_waves<N, T, R, S>()))) void k_chain(ChainParams<R> a)
{
    static_assert(S >= 1 && S <= 3, "Gauss-Legendre schemes with 1, 2 or 3 stages");
    constexpr int P = N / T, H = P / 2, M = N / 2;
    __shared__ cx<R> xb[rhs_lds<N, R, DB>()];
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
    // The step-start state y and the candidate next state y + dx sum b_i f_i
    // are read once per iteration: they live in LDS (thread-private slots),
    // the stage iterates in registers.  Thread t holds modes lf_kt(t) + T*h.
    __shared__ cx<R> ys[2][M];
    cx<R> yk[S][H];
    constexpr int TFIX = lf_owner<N, T>(M - 2) / 64;   // holds mode M-2 in slot H-1
    static_assert(lf_owner<N, T>(M - 2) % 64 == H - 1, "mode M-2 in the last state slot");
#pragma unroll
    for (int h = 0; h < H; ++h) ys[0][t + T * h] = Yg[lf_kt<N, T>(t) + T * h];
    const RhsConst<R> K = make_const<N, R>(a.dim ? a.dim[ch] : a.Dim, a.delta[ch]);
    __shared__ R ims[N];
    lf_ims<N, T, R>(ims, a.invm, K.inv_k0, t);      // thread-private slots
    const double *grid = a.grid + (a.goff ? a.goff[ch] : 0);

    const int dir = a.dir[ch];
    int idx = a.start[ch] + dir * s_begin;
    int64_t its_sum = 0;
    int32_t st = 0;
    // the grid values of the next step are loaded one step ahead
    double xin = grid[idx], xout = grid[idx + dir];
    for (int step = s_begin; step < s_end; ++step, idx += dir) {
        const double dxd = xout - xin;
        const double xnext = (step + 1 < s_end) ? grid[idx + 2 * dir] : 0.0;
        StageX<R> xs[S];
        R A[S][S], B[S];
        const R invN = (R)1 / (R)N;     // rhs_pair returns N f: exact power-of-two rescaling
#pragma unroll
        for (int i = 0; i < S; ++i) {
            xs[i] = stage_x((R)(xin + dxd * a.c[i]), K);
            B[i] = (R)(dxd * a.b[i]) * invN;
#pragma unroll
            for (int k = 0; k < S; ++k) A[i][k] = (R)(dxd * a.a[i][k]) * invN;
        }
#pragma unroll
        for (int h = 0; h < H; ++h)
#pragma unroll
            for (int i = 0; i < S; ++i) yk[i][h] = ys[0][t + T * h];
        if (t == TFIX) {                   // Re of mode M-2 of each stage, for the fix
            bc[0] = bc[1] = yk[0][H - 1].x;
            if constexpr (S == 3) bc[4] = bc[5] = yk[0][H - 1].x;
        }
        __syncthreads();
        bool conv = false;
        int its = 0;
        RT *rec = nullptr;
        if (S == 2 && a.rec && a.rec_base[ch] >= 0)
            rec = static_cast<RT *>(a.rec) + (size_t)(a.rec_base[ch] + step) * Rec<N>::STEP;
        for (its = 0; its < a.maxIts; ++its) {
            cx<R> f[S][H];
            if constexpr (S == 2) {
                rhs_pair<N, T, R, RT, DB>(yk, f, xs, K, ims, xb, bc, a.tw, t, rec);
            } else {
                cx<R> yp[2][H], fp[2][H];
#pragma unroll
                for (int h = 0; h < H; ++h) {
                    yp[0][h] = yk[0][h];
                    yp[1][h] = yk[S == 3 ? 1 : 0][h];
                }
                const StageX<R> xp[2] = {xs[0], xs[S == 3 ? 1 : 0]};
                rhs_pair<N, T, R, RT, DB>(yp, fp, xp, K, ims, xb, bc, a.tw, t);
#pragma unroll
                for (int h = 0; h < H; ++h) {
                    f[0][h] = fp[0][h];
                    if constexpr (S == 3) f[1][h] = fp[1][h];
                }
                if constexpr (S == 3) {
#pragma unroll
                    for (int h = 0; h < H; ++h) yp[0][h] = yp[1][h] = yk[2][h];
                    const StageX<R> x2[2] = {xs[2], xs[2]};
                    __syncthreads();       // behind the first evaluation's last exchange
                    rhs_pair<N, T, R, RT, DB>(yp, fp, x2, K, ims, xb, bc + 4, a.tw, t);
#pragma unroll
                    for (int h = 0; h < H; ++h) f[2][h] = fp[0][h];
                }
            }
            R loc = (R)0;
#pragma unroll
            for (int h = 0; h < H; ++h) {
                // ODEStepper.cpp:241-259 (stage update, same operation order)
                const cx<R> y = ys[0][t + T * h];
                cx<R> nn = y;
                R sq = (R)0;
#pragma unroll
                for (int i = 0; i < S; ++i) {
                    cx<R> ni = y;
#pragma unroll
                    for (int k = 0; k < S; ++k) {
                        ni.x += A[i][k] * f[k][h].x;
                        ni.y += A[i][k] * f[k][h].y;
                    }
                    const R dx = yk[i][h].x - ni.x, dy = yk[i][h].y - ni.y;
                    sq += dx * dx;
                    sq += dy * dy;
                    yk[i][h] = ni;
                }
                // ODEStepper.cpp:279-285 (final combination, in case this is the last iterate)
#pragma unroll
                for (int i = 0; i < S; ++i) {
                    nn.x += B[i] * f[i][h].x;
                    nn.y += B[i] * f[i][h].y;
                }
                loc += sq;
                ys[1][t + T * h] = nn;
            }
            if (t == TFIX) {
                bc[0] = yk[0][H - 1].x;
                bc[1] = yk[S == 1 ? 0 : 1][H - 1].x;
                if constexpr (S == 3) bc[4] = bc[5] = yk[2][H - 1].x;
            }
            // the barrier in block_sum publishes the fix values in bc (T > 64);
            // the next call is behind the next evaluation's exchange barriers
            const R norm2 = block_sum<T, R>(loc, red, t);
            if constexpr (T == 64) __syncthreads();
            const double rms = sqrt((double)norm2 / (double)(N * S));
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
        for (int h = 0; h < H; ++h) ys[0][t + T * h] = ys[1][t + T * h];
        xin = xout;
        xout = xnext;
    }
#pragma unroll
    for (int h = 0; h < H; ++h) Yg[lf_kt<N, T>(t) + T * h] = ys[0][t + T * h];
    if (t == 0) {
        a.its[ch] += its_sum;
        if (st) a.status[ch] = st;
    }
}

// computeDerivatives on pairs of states (unit entry point).  Y, dY: [2*npairs][M];
// xs, delta: [2*npairs]; both members of a pair share Delta.  rec (may be
// null): per pair one Rec<N> step, the fields of member s in stage slot s
// (StateVectorToFields with derivatives, ICG.cpp:791-884).
template <int N, int T, typename R>
__global__ __launch_bounds__(T) void k_rhs_pairs(const cx<R> *Y, const double *xs, const double *delta,
                                                 cx<R> *dY, const cx<R> *tw, const R *invm, double Dim, R *rec)
{
    constexpr int P = N / T, H = P / 2, M = N / 2;
    __shared__ cx<R> xb[rhs_lds<N, R, 0>()];
    __shared__ R bc[8];
    const int t = threadIdx.x;
    const int pr = blockIdx.x;
    cx<R> yk[2][H], f[2][H];
    const int kt = lf_kt<N, T>(t);
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int h = 0; h < H; ++h) yk[s][h] = Y[(size_t)(2 * pr + s) * M + kt + T * h];
    if (t == lf_owner<N, T>(M - 2) / 64) {
        bc[0] = yk[0][H - 1].x;
        bc[1] = yk[1][H - 1].x;
    }
    __syncthreads();
    const RhsConst<R> K = make_const<N, R>(Dim, delta[2 * pr]);
    const StageX<R> xsr[2] = {stage_x((R)xs[2 * pr], K), stage_x((R)xs[2 * pr + 1], K)};
    __shared__ R ims[N];
    lf_ims<N, T, R>(ims, invm, K.inv_k0, t);
    rhs_pair<N, T, R, R>(yk, f, xsr, K, ims, xb, bc, tw, t, rec ? rec + (size_t)pr * Rec<N>::STEP : nullptr);
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int h = 0; h < H; ++h) dY[(size_t)(2 * pr + s) * M + kt + T * h] = cscale(f[s][h], (R)1 / (R)N);
}

}  // namespace ccd
