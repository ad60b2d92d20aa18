// cc_tangent.hpp -- tangent-linear Jacobian columns (FP32: the mixed policy's
// fast path; FP64: an all-FP64 reference Jacobian).
//
// The reference builds column i of the Jacobian as the forward difference
// (F(x + eps e_i) - F(x)) / eps of two full shoots (NewtonSolver.cpp:672-685).
// At eps -> 0 that quotient is the directional derivative of the shoot, which
// is what a tangent-linear chain integrates directly: dY obeys the IRK2
// recursion linearised around the base trajectory,
//     dK_s = dy + dx sum_j a_sj  Jf(K_j) dK_j,   dy' = dy + dx sum_j b_j Jf(K_j) dK_j,
// with Jf the linearisation of computeDerivatives (ODEStepper.cpp:92-136) at
// the base stage fields recorded by the FP64 base shoot (Rec<N> in
// cc_chain.hpp, in the precision of the tangent).  There is no difference of
// two nearly equal shoots, so FP32 keeps its relative precision on every column, also on the weak ones whose
// finite differences drown in rounding (the Jacobian at these shapes has
// condition numbers near 1e9).
//
// Linearised pieces (fields in layout L, spectra in layout S as in
// cc_chain.hpp; both stages evaluated together):
//   fields:  the same doubled spectrum / mode-2 fix / tau-derivative / inverse
//            transform as the base, applied to dY (linear);
//   IA2:     IA2 solves IA2' + C1 IA2 + (D-3) = 0 periodically, so
//            dIA2' + C1 dIA2 + dC1 IA2 = 0: one solveInhom with the base's
//            integrating factor e^{F_int} and lambda (two transforms);
//   RHS:     dU/dx = A_U - B_U r etc. (see cc_chain.hpp) differentiated in
//            U, V, F, dU/dtau, dV/dtau and r = 1/IA2 (dr = -r^2 dIA2).
// 8 length-N transforms per fixed-point iteration.
#pragma once
#include "cc_chain.hpp"
#include "cc_spectral.hpp"

namespace ccd {

template <typename R>
struct TangentParams {
    cx<R> *dY;            // [nchains][M] tangent states, in/out
    cx<R> *fprev;         // [nchains][2][M] stage derivatives of a chain's last step (the
                          // predictor's state across launches), in/out; may be null
    const R *rec;         // base record (same precision as the tangent)
    const int64_t *rec_base;  // per chain: record slot of its step 0 (the base chain of its side)
    const int32_t *start, *nsteps, *dir;
    const int32_t *base_status;  // status of the base chains (left, right), may be null
    const int32_t *bs_idx;       // per chain: index of its problem's base-left chain in base_status
                                 // and delta (batched problems), null: 0
    const double *dim;           // per-chain D (batched problems), null: Dim
    const int32_t *goff;         // per-chain grid offset (batched problems), null: 0
    const double *delta;      // delta[bs_idx]: Delta of the base; every tangent column of a problem shares it
                              // (the Delta column itself is a finite difference in FP64)
    const double *grid;
    int32_t *status, *fail_idx;
    int64_t *its;
    const cx<R> *tw;
    const R *invm;
    int32_t step_begin, step_end, maxIts;
    double Dim, rtol;         // relative RMS tolerance of the tangent fixed point
    double a00, a01, a10, a11, b0, b1, c0, c1;
    const int32_t *order;     // launch slot -> chain (part of a permutation of the chains), null: identity
};

// FP32 tangents up to N = 1024 double-buffer their transform exchanges (one
// barrier per exchange); the second buffer would cost occupancy elsewhere.
template <int N, typename R>
constexpr int tangent_db() { return (sizeof(R) == 4 && N <= 1024) ? 1 : 0; }

// Reciprocal: one v_rcp_f32 (1 ulp) in the FP32 tangent, whose tolerance is
// 1e-6 relative; the correctly rounded quotient in the FP64 tangent.
template <typename R>
__device__ __forceinline__ R trcp(R x)
{
    if constexpr (sizeof(R) == 4) return __builtin_amdgcn_rcpf(x);
    else return R(1.0) / x;
}

// MERGE (double-buffered only): the two stages' first inverse transforms as
// one batch of four signals in regions of 4N complex; for workgroups that run
// alone on their CU (the register and LDS room it needs would cost the
// chip-filling launch occupancy).
template <int N, typename R, bool MERGE>
constexpr int tangent_region() { return (MERGE ? 4 : 2) * N; }

template <int N, int T, typename R, bool MERGE = false>
__device__ __forceinline__ void tangent_rhs_pair(const cx<R> (&dk)[2][N / T / 2], cx<R> (&df)[2][N / T / 2],
                                                 const R *__restrict__ rec, const R (&xs)[2],
                                                 const RhsConst<R> &K, const R *__restrict__ invm,
                                                 cx<R> *xb, R *bc, const cx<R> *__restrict__ tw, int t)
{
    constexpr int P = N / T, H = P / 2, M = N / 2;
    using RC = Rec<N>;
    const R half = R(0.5), invN = R(1.0) / (R)N;
    asm volatile("" : "+v"(t));
    // exchange regions of the chained transforms (as rhs_pair): no trailing
    // barriers; the caller's reduction barrier ends the evaluation
    constexpr int DB = tangent_db<N, R>(), RS = tangent_region<N, R, MERGE>();
    static_assert(!MERGE || DB, "the merged transforms need double-buffered regions");
    constexpr int pA = 0, pB = DB ? lf_next_par<N, T>(pA) : 0;
    constexpr int pC = DB ? lf_next_par<N, T>(MERGE ? pA : pB) : 0;
    constexpr int pS = DB ? lf_next_par<N, T>(pC) : 0, pG = DB ? 1 - pS : 0;
    constexpr int pH = DB ? lf_next_par<N, T>(pG) : 0;
    const int kt = lf_kt<N, T>(t);      // spectral modes of this thread: kt + T*p
    // the stage twiddles, loaded once per evaluation and held in registers
    constexpr bool PRE = sizeof(R) == 4;   // (the FP64 tangent has no register room)
    LfTw<N, T, R> twr{};
    if constexpr (PRE) twr = lf_twiddles<N, T, R>(tw, t);
    // per-stage products that survive to the pointwise tangent RHS
    // (the base-only products B_U, B_V, C_F are recomputed from the record in
    // the pointwise stage: held across the transforms they would spill)
    R dAU[2][H], dBU[2][H], dAV[2][H], dBV[2][H], dCF[2][H];
    cx<R> c[P];                              // packed h of both stages
    // (3) per stage, after its inverse transform: the linearised split and h_s
    auto split = [&](int s, const cx<R> (&z)[2][P]) {
        int stage_off = s * RC::STAGE;
        // The record of a step is the same for every fixed-point iteration and
        // every transform; an opaque offset keeps this stage's loads behind its
        // transform instead of hoisted in front of it (or of the loop), where
        // they would spill.  (An offset, not a pointer: the loads stay global.)
        asm volatile("" : "+s"(stage_off));
        const R *rs = rec + stage_off;
        const R X = xs[s], two_x = R(2.0) * X, inv_x = trcp(X);
        R dC1[H];
#pragma unroll
        for (int h = 0; h < H; ++h) {
            const cx<R> a = z[0][h], b = z[0][h + H];
            const cx<R> a2 = z[1][h], b2 = z[1][h + H];
            const R dU = half * (a.x - b.x), dV = half * (a.y - b.y), dF = half * (a.y + b.y);
            const R ddU = half * (a2.x - b2.x), ddV = half * (a2.y - b2.y);
            const int j = t + T * h;
            const R U = rs[RC::U + j], V = rs[RC::V + j], F = rs[RC::F + j];
            const R tU = rs[RC::DU + j], tV = rs[RC::DV + j], FmX = rs[RC::FMX + j];
            // d Coeff1 (ICG.cpp:834-835 differentiated)
            dC1[h] = -(K.c3 * R(0.125) * inv_x) *
                     (dF * U * U + R(2.0) * (X + F) * U * dU - dF * V * V - R(2.0) * FmX * V * dV);
            const R inv1 = trcp(two_x * (F + X)), inv2 = trcp(two_x * FmX);
            const R dinv1 = -inv1 * inv1 * two_x * dF, dinv2 = -inv2 * inv2 * two_x * dF;
            const R S = U + V, dS = dU + dV;
            const R numU = F * K.dm2 * S - two_x * tU, numV = F * K.dm2 * S + two_x * tV;
            const R dnum = K.dm2 * (dF * S + F * dS);
            dAU[s][h] = (dnum - two_x * ddU) * inv1 + numU * dinv1;
            dBU[s][h] = K.two_dm3 * ((dF * U + F * dU) * inv1 + F * U * dinv1);
            dAV[s][h] = (dnum + two_x * ddV) * inv2 + numV * dinv2;
            dBV[s][h] = K.two_dm3 * ((dF * V + F * dV) * inv2 + F * V * dinv2);
            dCF[s][h] = K.dm3 * dF * inv_x;
        }
        // h_s = -(dC1 IA2) e^{F_int} = -dC1 E / r  (dC1 is even in the half period)
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const int ph = (p < H) ? p : p - H;
            const int j = t + T * p;
            const R hv = -dC1[ph] * rs[RC::E + j] * trcp(rs[RC::R + j]);
            if (s == 0) c[p].x = hv; else c[p].y = hv;
        }
    };
    cx<R> zz[2][2][P];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        cx<R>(&z)[2][P] = zz[s];
        const R fix = bc[s];
#pragma unroll
        for (int h = 0; h < H; ++h) {
            const int q = kt + T * h;
            const cx<R> Y = dk[s][h];
            cx<R> lo, hi;
            if (q < M / 2) {
                lo = Y;
                hi = {R(0.0), R(0.0)};
            } else if (q == M / 2) {
                lo = cscale(Y, half);
                hi = Y;
            } else {
                lo = {R(0.0), R(0.0)};
                hi = Y;
            }
            if (q == 2) lo.x = -fix;
            z[0][h] = lo;
            z[0][h + H] = hi;
        }
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const int k = kt + T * p;
            const R w = (k == N / 2) ? R(0.0) : (R)wavenumber<N>(k) * K.k0;
            const cx<R> v = z[0][p];
            z[1][p] = {w * v.y, -(w * v.x)};
        }
        if constexpr (!MERGE) {
            if (s == 0) lfft_dit<N, T, -1, 2, R, DB, pA, RS, PRE>(z, xb, tw, t, &twr);
            else lfft_dit<N, T, -1, 2, R, DB, pB, RS, PRE>(z, xb, tw, t, &twr);
            split(s, z);
        }
    }
    if constexpr (MERGE) {
        // both stages' inverse transforms as one batch of four signals: one
        // exchange and barrier instead of two
        lfft_dit<N, T, -1, 4, R, DB, pA, RS, PRE>(*reinterpret_cast<cx<R>(*)[4][P]>(&zz), xb, tw, t, &twr);
        split(0, zz[0]);
        split(1, zz[1]);
    }
    // dIA2 = solveInhom with the base's integrating factor and lambda
    lfft_dif1<N, T, +1, R, DB, pC, RS, PRE>(c, xb, tw, t, &twr);
    // (the transform's 1/N, a power of two, is folded into the separation below)
#pragma unroll
    for (int p = 0; p < P; ++p) xb[pS * RS + lf_mirror_pos<N, T, R>(kt + T * p)] = c[p];
    __syncthreads();
    cx<R> cm[P];
#pragma unroll
    for (int p = 0; p < P; ++p) cm[p] = xb[pS * RS + lf_mirror_pos<N, T, R>((N - (kt + T * p)) & (N - 1))];
    if constexpr (!DB) __syncthreads();
    const R lam0 = rec[RC::LAM], lam1 = rec[RC::STAGE + RC::LAM];
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
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const cx<R> hs = s ? S1 : S0;
            const R lam = s ? lam1 : lam0;
            if (k == 0) {
                G[s] = (lam != R(0.0)) ? cscale(hs, trcp(lam)) : cx<R>{R(0.0), R(0.0)};
            } else if (k == N / 2) {
                G[s] = cscale(hs, lam * trcp(lam * lam + K.q2));
            } else {
                const R w = (R)wavenumber<N>(k) * K.k0;
                const R den = lam * lam + w * w;
                const R inv = trcp(den);
                G[s] = {(hs.x * lam - hs.y * w) * inv, (hs.y * lam + hs.x * w) * inv};
            }
        }
        c[p] = {G[0].x - G[1].y, G[0].y + G[1].x};
    }
    lfft_dit1<N, T, -1, R, DB, pG, RS, PRE>(c, xb, tw, t, &twr);
    // pointwise tangent RHS: the half-period products of point jh serve both
    // points jh and jh + N/2 (U, V odd and F even under the half shift)
    cx<R> yp[2][P];
#pragma unroll
    for (int h = 0; h < H; ++h) {
        const int jh = t + T * h;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            int fin_off = s * RC::STAGE;
            asm volatile("" : "+s"(fin_off));          // behind the inverse transform, as above
            const R *rs = rec + fin_off;
            const R U = rs[RC::U + jh], V = rs[RC::V + jh], F = rs[RC::F + jh], FmX = rs[RC::FMX + jh];
            const R X = xs[s], two_x = R(2.0) * X;
            const R BU = K.two_dm3 * F * U * trcp(two_x * (F + X));
            const R BV = K.two_dm3 * F * V * trcp(two_x * FmX);
            const R CF = K.dm3 * F * trcp(X);
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                const int p = h + half * H;
                const int j = t + T * p;
                const R sg = half ? -R(1.0) : R(1.0);
                const R r = rs[RC::R + j], E = rs[RC::E + j];
                const R dIA2 = (s ? c[p].y : c[p].x) * trcp(E);
                const R dr = -r * r * dIA2;
                const R dUx = sg * (dAU[s][h] - dBU[s][h] * r - BU * dr);
                const R dVx = sg * (dAV[s][h] - dBV[s][h] * r - BV * dr);
                const R dFx = dCF[s][h] * (r - R(1.0)) + CF * dr;
                yp[s][p] = {dUx, dVx + dFx};
            }
        }
    }
    lfft_dif<N, T, +1, 2, R, DB, pH, RS, PRE>(yp, xb, tw, t, &twr);
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int h = 0; h < H; ++h) {
            const int q = kt + T * h;
            const cx<R> lo = yp[s][h], hi = yp[s][h + H];
            cx<R> v = (q < M / 2) ? lo : ((q == M / 2) ? lo + hi : hi);
            if (q == 2) v.x = R(0.0);
            df[s][h] = v;                   // N times the derivative: the caller folds in 1/N
        }
}

// Register budget of the FP32 tangent: 3 waves per SIMD (168 VGPRs) at four
// values per thread, 2 (256 VGPRs) at eight, so that a launch of n two-wave
// workgroups still fits the chip in one round.  The FP64 tangent (an
// all-FP64 reference Jacobian, not the fast path) gets 2 waves at four values.
template <int N, int T, typename R>
constexpr int tangent_waves() { return sizeof(R) == 8 ? (N / T >= 8 ? 1 : 2) : (N / T >= 8 ? 2 : 3); }

// SOLO: the launch of the chains that run alone on their CUs (the host pads
// the launch's LDS): compiled for one wave per SIMD (no register limit from
// occupancy, so nothing spills), with both stages' first transforms batched
// (MERGE); every value is computed as in the main launch.
template <int N, int T, typename R, bool SOLO = false>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(SOLO ? 1 : tangent_waves<N, T, R>()))) void k_tangent(TangentParams<R> a)
{
    constexpr int P = N / T, H = P / 2, M = N / 2;
    // the solo launch batches both stages' first transforms (tangent_rhs_pair)
    constexpr bool MERGE = SOLO && tangent_db<N, R>() != 0;
    __shared__ cx<R> xb[(1 + tangent_db<N, R>()) * tangent_region<N, R, MERGE>()];
    __shared__ cx<R> ys[2][M];
    // stage derivatives of the previous step (predictor); at N = 2048 the
    // array would cost a third of the workgroups per CU, so it is left out
    constexpr bool PRED = N <= 1024;
    __shared__ cx<R> fs[2][PRED ? M : 1];
    __shared__ R bc[8];
    __shared__ R red[2 * ((T / 64) > 0 ? T / 64 : 1)];
    const int t = threadIdx.x;
    // Workgroups are dispatched in blockIdx order, and a launch of all columns
    // has a few more workgroups than the CUs beside the base chains hold: the
    // last ones start in a second round.  Reversed, those are the first
    // columns (the right side's Up modes, whose chains are never the slow
    // ones), not the high fc modes of the left side.
    const int slot = (int)gridDim.x - 1 - (int)blockIdx.x;
    const int ch = a.order ? a.order[slot] : slot;
    if (a.status[ch] != 0) return;
    const int dir = a.dir[ch];
    // a failed base chain leaves no record to linearise around
    const int bsi = a.bs_idx ? a.bs_idx[ch] : 0;
    if (a.base_status && a.base_status[bsi + (dir > 0 ? 0 : 1)] != 0) return;
    const int nst = a.nsteps[ch];
    const int s_begin = a.step_begin;
    const int s_end = min(a.step_end, nst);
    if (s_begin >= s_end) return;
    cx<R> *Yg = a.dY + (size_t)ch * M;
    // thread t holds modes lf_kt(t) + T*h (layout S); HBM keeps the natural order
    constexpr int TFIX = lf_owner<N, T>(M - 2) / 64;   // holds mode M-2 in slot H-1
    static_assert(lf_owner<N, T>(M - 2) % 64 == H - 1, "mode M-2 in the last state slot");
#pragma unroll
    for (int h = 0; h < H; ++h) ys[0][t + T * h] = Yg[lf_kt<N, T>(t) + T * h];
    const RhsConst<R> K = make_const<N, R>(a.dim ? a.dim[ch] : a.Dim, a.delta[bsi]);
    const double *grid = a.grid + (a.goff ? a.goff[ch] : 0);
    int idx = a.start[ch] + dir * s_begin;
    int64_t its_sum = 0;
    int32_t st = 0;
    cx<R> dk[2][H];
    // the predictor continues across launches: a chain resumed mid-integration
    // reloads its last step's derivatives (the result does not depend on
    // where the step sequence is cut into launches)
    bool have_prev = PRED && a.fprev && s_begin > 0;
    if (have_prev) {
        const cx<R> *fp = a.fprev + (size_t)ch * 2 * M;
        const int kt = lf_kt<N, T>(t);
#pragma unroll
        for (int h = 0; h < H; ++h) {
            fs[0][t + T * h] = fp[kt + T * h];
            fs[1][t + T * h] = fp[M + kt + T * h];
        }
    }
    for (int step = s_begin; step < s_end; ++step, idx += dir) {
        const double xin = grid[idx], xout = grid[idx + dir];
        const double dxd = xout - xin;
        const R xs[2] = {(R)(xin + dxd * a.c0), (R)(xin + dxd * a.c1)};
        // tangent_rhs_pair returns N times the derivatives: 1/N (exact) goes into the tableau
        const R invN = R(1.0) / (R)N;
        const R A00 = (R)(dxd * a.a00) * invN, A01 = (R)(dxd * a.a01) * invN, A10 = (R)(dxd * a.a10) * invN,
                    A11 = (R)(dxd * a.a11) * invN;
        const R B0 = (R)(dxd * a.b0) * invN, B1 = (R)(dxd * a.b1) * invN;
        const R *rec = a.rec + (size_t)(a.rec_base[ch] + step) * Rec<N>::STEP;
        // Initial stage iterates: the step-start state, or (after a chain's
        // first step) the previous step's collocation polynomial extrapolated
        // to this step's stage points: with u' = sum_j f_j(prev) L_j(theta)
        // (L_j the Lagrange basis on the nodes c, theta in units of the
        // previous step hp), K_s = y + hp sum_j f_j int_1^{1 + r c_s} L_j, r =
        // dx / hp, i.e. E_sj = dx c_s (2 (1 - c_o) + r c_s) / (2 (c_j - c_o))
        // with o the other node.  The linear fixed point then starts O(dx^3)
        // from its solution, against O(dx) from K_s = y (the reference's
        // start, ODEStepper.cpp:193-199).  A start value: computed in R.
        R E00 = R(0.0), E01 = R(0.0), E10 = R(0.0), E11 = R(0.0);
        if (PRED && have_prev) {
            const R r = (R)dxd * trcp((R)(xin - grid[idx - dir]));
            // (times 1/N: the stored derivatives are N times the derivatives)
            const R c0 = (R)a.c0, c1 = (R)a.c1, g = R(0.5) * (R)dxd * trcp(c0 - c1) * invN;
            const R u0 = R(2.0) * (R(1.0) - c0), u1 = R(2.0) * (R(1.0) - c1);
            E00 = g * c0 * (u1 + r * c0);
            E01 = -(g * c0 * (u0 + r * c0));
            E10 = g * c1 * (u1 + r * c1);
            E11 = -(g * c1 * (u0 + r * c1));
        }
#pragma unroll
        for (int h = 0; h < H; ++h) {
            const cx<R> y = ys[0][t + T * h];
            dk[0][h] = dk[1][h] = y;
            if (PRED && have_prev) {
                const cx<R> f0 = fs[0][t + T * h], f1 = fs[1][t + T * h];
                dk[0][h] = {y.x + E00 * f0.x + E01 * f1.x, y.y + E00 * f0.y + E01 * f1.y};
                dk[1][h] = {y.x + E10 * f0.x + E11 * f1.x, y.y + E10 * f0.y + E11 * f1.y};
            }
        }
        if (t == TFIX) {
            bc[0] = dk[0][H - 1].x;
            bc[1] = dk[1][H - 1].x;
        }
        __syncthreads();
        bool conv = false;
        int its = 0;
        for (its = 0; its < a.maxIts; ++its) {
            cx<R> df[2][H];
            tangent_rhs_pair<N, T, R, MERGE>(dk, df, rec, xs, K, a.invm, xb, bc, a.tw, t);
            R loc = R(0.0), mag = R(0.0);
#pragma unroll
            for (int h = 0; h < H; ++h) {
                const cx<R> y = ys[0][t + T * h];
                cx<R> n0 = y, n1 = y, nn = y;
                n0.x += A00 * df[0][h].x; n0.y += A00 * df[0][h].y;
                n0.x += A01 * df[1][h].x; n0.y += A01 * df[1][h].y;
                n1.x += A10 * df[0][h].x; n1.y += A10 * df[0][h].y;
                n1.x += A11 * df[1][h].x; n1.y += A11 * df[1][h].y;
                nn.x += B0 * df[0][h].x; nn.y += B0 * df[0][h].y;
                nn.x += B1 * df[1][h].x; nn.y += B1 * df[1][h].y;
                const R d0x = dk[0][h].x - n0.x, d0y = dk[0][h].y - n0.y;
                const R d1x = dk[1][h].x - n1.x, d1y = dk[1][h].y - n1.y;
                loc += d0x * d0x + d0y * d0y + d1x * d1x + d1y * d1y;
                mag += n0.x * n0.x + n0.y * n0.y + n1.x * n1.x + n1.y * n1.y;
                dk[0][h] = n0;
                dk[1][h] = n1;
                ys[1][t + T * h] = nn;
            }
            if (t == TFIX) {
                bc[0] = dk[0][H - 1].x;
                bc[1] = dk[1][H - 1].x;
            }
            R d2 = loc, m2 = mag;
            block_sum2<T, R>(d2, m2, red, t);          // its barrier publishes bc[0..1] (T > 64)
            if constexpr (T == 64) __syncthreads();
            // linear fixed point: stop when the update is below rtol of the
            // iterate, relaxed late as the base iteration's tolerance
            // (ODEStepper.cpp:265-269)
            double rtol = a.rtol;
            if (its > a.maxIts / 2) rtol = a.rtol * 10.0 * (2.0 * (double)its / (double)a.maxIts - 0.5);
            // (an overflowed iterate, m2 = inf, is never accepted)
            if (isfinite(m2) && d2 <= (R)(rtol * rtol) * m2) {
                conv = true;
                if constexpr (PRED) {
#pragma unroll
                    for (int h = 0; h < H; ++h) {
                        fs[0][t + T * h] = df[0][h];
                        fs[1][t + T * h] = df[1][h];
                    }
                }
                break;
            }
        }
        its_sum += (conv ? its + 1 : its);
        if (!conv) {
            st = 1;
            if (t == 0) a.fail_idx[ch] = idx;
            break;
        }
        have_prev = true;
#pragma unroll
        for (int h = 0; h < H; ++h) ys[0][t + T * h] = ys[1][t + T * h];
    }
    // addresses of the final stores are formed here, not kept (spilled)
    // across the step loop from the prologue
    int te = t;
    asm volatile("" : "+v"(te));
    const int kte = lf_kt<N, T>(te);
#pragma unroll
    for (int h = 0; h < H; ++h) Yg[kte + T * h] = ys[0][te + T * h];
    if (PRED && a.fprev && have_prev) {
        cx<R> *fp = a.fprev + (size_t)ch * 2 * M;
#pragma unroll
        for (int h = 0; h < H; ++h) {
            fp[kte + T * h] = fs[0][te + T * h];
            fp[M + kte + T * h] = fs[1][te + T * h];
        }
    }
    if (t == 0) {
        a.its[ch] += its_sum;
        if (st) a.status[ch] = st;
    }
}

// Tangent boundary states of column cols[b]: central difference of the FP64
// boundary series, d Y0 / d x_i = (Y0(x + h e_i) - Y0(x - h e_i)) / 2h.
template <int N, int T, bool LEFT, typename R>
__global__ __launch_bounds__(T) void k_boundary_tangent(const double *x, const int32_t *cols, const int32_t *chain,
                                                         double hrel, double Xb, double Dim, const cx<double> *tw,
                                                         cx<R> *dY)
{
    constexpr int P = N / T, H = P / 2, M = N / 2;
    __shared__ cx<double> xb[N];
    __shared__ double bc[4];
    const int t = threadIdx.x;
    const int b = blockIdx.x;
    const int col = cols[b];
    const double Delta = x[N / 2 + 2];
    const double h = hrel * fmax(1.0, fabs(x[col]));
    Spec<N, T> S{t, xb, bc, tw, 2.0 * 3.14159265358979323846 / Delta};
    cx<double> Yp[H], Ym[H];
    for (int sgn = 0; sgn < 2; ++sgn) {
        double Up[P], psic[P], fc[P];
        unpack_fields<N, T>(S, x, col, sgn ? -h : h, Up, psic, fc);
        cx<double> Yh[H];
        if constexpr (LEFT)
            left_expansion<N, T>(S, Xb, Dim, fc, psic, Yh);
        else
            right_expansion<N, T>(S, Xb, Dim, Up, Yh);
#pragma unroll
        for (int q = 0; q < H; ++q) {
            if (sgn) Ym[q] = Yh[q]; else Yp[q] = Yh[q];
        }
    }
    const int ch = chain[b];
    const double inv2h = 1.0 / (2.0 * h);
#pragma unroll
    for (int q = 0; q < H; ++q) {
        cx<double> d = {(Yp[q].x - Ym[q].x) * inv2h, (Yp[q].y - Ym[q].y) * inv2h};
        if (t == 2 && q == 0) d.x = 0.0;          // Re Y[2] carries Delta, fixed here
        dY[(size_t)ch * M + t + T * q] = {(R)d.x, (R)d.y};
    }
}

// Jacobian column from the tangent state at the match point: the mismatch,
// state -> fields and pack are linear (ShootingSolver.cpp:154-182,
// ICG.cpp:891-932, 678-711).  lidx/ridx < 0: that side has a zero tangent.
template <int N, int T, typename R>
__global__ __launch_bounds__(T) void k_tangent_column(const cx<R> *dY, const int32_t *cols, const int32_t *lidx,
                                                       const int32_t *ridx, const cx<double> *tw, double *J)
{
    constexpr int P = N / T, H = P / 2, M = N / 2, n = 3 * N / 4;
    __shared__ cx<double> xb[N];
    __shared__ double bc[4];
    __shared__ double col_out[n];
    const int t = threadIdx.x;
    const int b = blockIdx.x;
    const int li = lidx[b], ri = ridx[b];
    Spec<N, T> S{t, xb, bc, tw, 0.0};          // no tau-derivative in state -> fields -> pack
    cx<double> mis[H];
#pragma unroll
    for (int h = 0; h < H; ++h) {
        const int q = t + T * h;
        cx<double> v = {0.0, 0.0};
        if (ri >= 0) v = v + cx<double>{(double)dY[(size_t)ri * M + q].x, (double)dY[(size_t)ri * M + q].y};
        if (li >= 0) v = v - cx<double>{(double)dY[(size_t)li * M + q].x, (double)dY[(size_t)li * M + q].y};
        mis[h] = v;
    }
    state_to_packed<N, T>(S, mis, col_out);
    __syncthreads();
    const int c = cols[b];
    for (int j = t; j < n; j += T) J[(size_t)j * n + c] = col_out[j];
}

}  // namespace ccd
