// cc_device.hpp -- device building blocks for gfx950: complex arithmetic,
// in-register DFTs and the register-resident Stockham FFT with LDS exchange.
//
// Layout convention used by every kernel ("layout L"): a workgroup of T
// threads owns a length-N signal; thread t holds elements t + T*p for
// p = 0 .. P-1 (P = N/T) in registers.  Loads and stores of layout L are
// unit-stride across a wavefront.
//
// Transform conventions are those of the reference's SpectralTransformer
// (SpectralTransformer.cpp:33-38): SIGN = +1 computes
// X_k = sum_j x_j exp(+2 pi i jk/N) (the reference's "forwardFFT", which then
// divides by N), SIGN = -1 computes exp(-2 pi i jk/N) unscaled ("backwardFFT").
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ccd {

template <typename R>
struct cx {
    R x, y;
};

template <typename R>
__device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R>
__device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) { return {a.x - b.x, a.y - b.y}; }
template <typename R>
__device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename R>
__device__ __forceinline__ cx<R> cscale(cx<R> a, R s) { return {a.x * s, a.y * s}; }

// a * (SIGN * i)
template <int SIGN, typename R>
__device__ __forceinline__ cx<R> mul_si(cx<R> a)
{
    if constexpr (SIGN > 0) return {-a.y, a.x};
    else return {a.y, -a.x};
}

// cos/sin(2 pi k / 16), k = 0..15, for the in-register DFTs (R <= 16)
__device__ constexpr double kC16[16] = {
    1.0, 0.92387953251128674, 0.70710678118654752, 0.38268343236508977,
    0.0, -0.38268343236508977, -0.70710678118654752, -0.92387953251128674,
    -1.0, -0.92387953251128674, -0.70710678118654752, -0.38268343236508977,
    0.0, 0.38268343236508977, 0.70710678118654752, 0.92387953251128674};
__device__ constexpr double kS16[16] = {
    0.0, 0.38268343236508977, 0.70710678118654752, 0.92387953251128674,
    1.0, 0.92387953251128674, 0.70710678118654752, 0.38268343236508977,
    0.0, -0.38268343236508977, -0.70710678118654752, -0.92387953251128674,
    -1.0, -0.92387953251128674, -0.70710678118654752, -0.38268343236508977};

// w^K * a with w = exp(SIGN 2 pi i / RAD), K a compile-time constant
template <int RAD, int K, int SIGN, typename R>
__device__ __forceinline__ cx<R> twiddle_const(cx<R> a)
{
    constexpr int KK = K % RAD;
    if constexpr (KK == 0) {
        return a;
    } else if constexpr (4 * KK == RAD) {
        return mul_si<SIGN>(a);
    } else if constexpr (2 * KK == RAD) {
        return {-a.x, -a.y};
    } else if constexpr (4 * KK == 3 * RAD) {
        return mul_si<-SIGN>(a);
    } else {
        constexpr int idx = KK * (16 / RAD);
        const R c = (R)kC16[idx];
        const R s = (R)(SIGN * kS16[idx]);
        return {a.x * c - a.y * s, a.x * s + a.y * c};
    }
}

template <int RAD, int K, int SIGN, typename R>
__device__ __forceinline__ void combine(cx<R>* v, const cx<R>* e, const cx<R>* o)
{
    if constexpr (K < RAD / 2) {
        const cx<R> t = twiddle_const<RAD, K, SIGN>(o[K]);
        v[K] = e[K] + t;
        v[K + RAD / 2] = e[K] - t;
        combine<RAD, K + 1, SIGN>(v, e, o);
    }
}

// In-register DFT of length RAD (power of two <= 16), natural order in/out.
template <int RAD, int SIGN, typename R>
__device__ __forceinline__ void dft(cx<R>* v)
{
    if constexpr (RAD == 1) {
        return;
    } else if constexpr (RAD == 2) {
        const cx<R> a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    } else {
        cx<R> e[RAD / 2], o[RAD / 2];
#pragma unroll
        for (int i = 0; i < RAD / 2; ++i) {
            e[i] = v[2 * i];
            o[i] = v[2 * i + 1];
        }
        dft<RAD / 2, SIGN>(e);
        dft<RAD / 2, SIGN>(o);
        combine<RAD, 0, SIGN>(v, e, o);
    }
}

// One Stockham pass (Govindaraju et al., SC'08 formulation) of a batch of B
// signals held in layout L; recurses to the next pass.  NS is the length of
// the sub-transforms already done.  The radix is P while N/NS >= P, then the
// remaining factor.  All but the last pass exchange through LDS (B*N
// complex); the last pass's outputs land back in layout L.
// tw[e] = (cos, sin)(2 pi e / N), e in [0, N).
// Exchange-buffer position of element i in the exchange after a pass with
// sub-transform length NS and radix RAD.  Stockham writes of the early passes
// put consecutive lanes RAD elements apart, so the lanes that one LDS cycle
// serves (16 lanes of 8-byte FP32 complex, 8 lanes of 16-byte FP64 complex:
// 128 bytes over 32 banks) would share a few bank slots.  FP32: both low bit
// pairs XOR bits 4-5 of i, in every pass.  FP64: per pass, each of the lane
// bits 0-2 that the write sends above position bit 2 (to bit m + log2 RAD
// when 2^m >= NS) is XORed into a low position bit no lane bit occupies, and
// nothing else: 8 consecutive lanes then write 8 distinct slots, and passes
// whose lane bits stay low (NS >= 8) keep the identity.  Measured on gfx950:
// XORing bit 4 into bit 2 for all passes, the earlier FP64 map, made every
// layout-L read of a 16-byte element cost 4 extra LDS cycles, which the
// per-pass maps confine to the NS = 4 exchange of radix 4.  Each map permutes
// aligned blocks of 8 (FP64) or 64 (FP32) positions, so positions stay in
// [0, N), and is linear over GF(2): for a and b with disjoint bits (a + b ==
// a ^ b), xswz(a + b) == xswz(a) ^ xswz(b), so a position made of a
// per-thread part and a compile-time offset costs one XOR.
template <int NS, int RAD>
constexpr int xswz64_pairs()
{
    // packs up to three (source bit, destination bit) pairs, 4 bits each
    int lg = 0;
    while ((1 << lg) < RAD) ++lg;
    bool low_used[3] = {false, false, false};
    int src[3] = {0, 0, 0}, ns = 0;
    for (int m = 0; m < 3; ++m) {
        const int pos = ((1 << m) < NS) ? m : m + lg;
        if (pos < 3) low_used[pos] = true;
        else src[ns++] = pos;
    }
    int packed = 0, k = 0;
    for (int f = 0; f < 3 && k < ns; ++f)
        if (!low_used[f]) {
            packed |= ((src[k] << 4) | f) << (8 * k);
            ++k;
        }
    return packed | (k << 24);
}

// (host and device: the host-side check of the maps calls this same function)
template <int N, typename R, int NS, int RAD>
__host__ __device__ __forceinline__ constexpr int xswz(int i)
{
    if constexpr (N < 64) {
        return i;
    } else if constexpr (sizeof(R) == 8) {
        constexpr int pk = xswz64_pairs<NS, RAD>(), np = pk >> 24;
        int x = i;
        if constexpr (np > 0) x ^= ((i >> ((pk >> 4) & 15)) & 1) << (pk & 15);
        if constexpr (np > 1) x ^= ((i >> ((pk >> 12) & 15)) & 1) << ((pk >> 8) & 15);
        if constexpr (np > 2) x ^= ((i >> ((pk >> 20) & 15)) & 1) << ((pk >> 16) & 15);
        return x;
    } else {
        const int g = (i >> 4) & 3;
        return i ^ (g | (g << 2));
    }
}

// DB = 1: the exchanges alternate between two LDS regions (RS complex apart,
// RS >= B*N), so each exchange needs one barrier instead of two: a thread
// writing region k+2 has passed exchange k+1's barrier, which every thread
// reaches only after reading region k.  PAR is the region of this pass.
template <int N, int T, int SIGN, int B, int NS, typename R, int DB = 0, int PAR = 0, int RS = B * N>
__device__ __forceinline__ void fft_pass(cx<R> (&v)[B][N / T], cx<R>* lds, const cx<R>* __restrict__ tw,
                                         int t)
{
    constexpr int P = N / T;
    constexpr int RAD = (N / NS >= P) ? P : N / NS;
    constexpr bool LAST = (NS * RAD == N);
    constexpr int NB = P / RAD;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int j = t + T * b;
        int k = j & (NS - 1);
        // keep the twiddle computation inside the loop that calls the
        // transform: hoisting it out of the fixed-point loop costs far more
        // registers than the loads cost bandwidth
        asm volatile("" : "+v"(k));
        cx<R> w[RAD];
        if constexpr (NS > 1) {
            // w^r for r < RAD, w = e^{SIGN 2 pi i k / (NS RAD)}.  Radix <= 4: one
            // table load and products (w^2 = w w, w^3 = w w^2); larger radices
            // load every power.
            constexpr int STRIDE = N / (NS * RAD);
            if constexpr (RAD <= 4) {
                cx<R> w1 = tw[k * STRIDE];
                if constexpr (SIGN < 0) w1.y = -w1.y;
                w[1] = w1;
                if constexpr (RAD == 4) {
                    w[2] = cmul(w1, w1);
                    w[3] = cmul(w1, w[2]);
                }
            } else {
#pragma unroll
                for (int r = 1; r < RAD; ++r) {
                    cx<R> tr = tw[r * k * STRIDE];
                    if constexpr (SIGN < 0) tr.y = -tr.y;
                    w[r] = tr;
                }
            }
        }
#pragma unroll
        for (int s = 0; s < B; ++s) {
            cx<R> u[RAD];
#pragma unroll
            for (int r = 0; r < RAD; ++r) u[r] = v[s][b + r * NB];
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < RAD; ++r) u[r] = cmul(u[r], w[r]);
            }
            dft<RAD, SIGN>(u);
            if constexpr (LAST) {
#pragma unroll
                for (int r = 0; r < RAD; ++r) v[s][b + r * NB] = u[r];
            } else {
                // base has no bits in common with r * NS (base mod NS*RAD == k < NS)
                const int sb = xswz<N, R, NS, RAD>((j / NS) * NS * RAD + k);
#pragma unroll
                for (int r = 0; r < RAD; ++r) lds[PAR * RS + s * N + (sb ^ xswz<N, R, NS, RAD>(r * NS))] = u[r];
            }
        }
    }
    if constexpr (!LAST) {
        __syncthreads();
#pragma unroll
        for (int s = 0; s < B; ++s)
#pragma unroll
            for (int p = 0; p < P; ++p) v[s][p] = lds[PAR * RS + s * N + (xswz<N, R, NS, RAD>(t) ^ xswz<N, R, NS, RAD>(T * p))];
        if constexpr (!DB) __syncthreads();
        fft_pass<N, T, SIGN, B, NS * RAD, R, DB, DB ? 1 - PAR : 0, RS>(v, lds, tw, t);
    }
}

// Number of LDS exchanges of one transform (passes before the last).
template <int N, int T>
constexpr int fft_exchanges()
{
    int e = 0;
    for (int ns = 1;;) {
        const int rad = (N / ns >= N / T) ? N / T : N / ns;
        if (ns * rad == N) return e;
        ++e;
        ns *= rad;
    }
}

// Chained double-buffered transforms without trailing barriers.  A transform
// starting in region p0 reads region p0 last when it has an odd number of
// exchanges, else region 1-p0; the next LDS writer may start at once in the
// other region (every thread passed the last exchange's barrier only after
// reading the region before it), which is what fft_next_par returns.
template <int N, int T>
constexpr int fft_next_par(int p0)
{
    return (fft_exchanges<N, T>() % 2 == 1) ? 1 - p0 : p0;
}

// Batched block FFT: B signals of length N in layout L, in place.
// Requires N/T >= 2 and a __syncthreads-safe caller (all T threads call it).
// lds holds B*N complex, or with DB = 1 two regions RS apart starting with
// region P0.  TRAIL = 1 (DB) ends with a barrier that protects both regions
// from the caller's next LDS writes; TRAIL = 0 leaves that to the caller
// (see fft_next_par).
template <int N, int T, int SIGN, int B, typename R, int DB = 0, int P0 = 0, int TRAIL = 1, int RS = B * N>
__device__ __forceinline__ void fft_block(cx<R> (&v)[B][N / T], cx<R>* lds, const cx<R>* __restrict__ tw, int t)
{
    static_assert((N & (N - 1)) == 0 && N / T >= 2 && N / T <= 16, "unsupported FFT shape");
    static_assert(DB || P0 == 0, "one exchange region without double buffering");
    fft_pass<N, T, SIGN, B, 1, R, DB, P0, RS>(v, lds, tw, t);
    if constexpr (DB && TRAIL) __syncthreads();
}

// Single-signal convenience wrapper.
template <int N, int T, int SIGN, typename R, int DB = 0, int P0 = 0, int TRAIL = 1, int RS = N>
__device__ __forceinline__ void fft1(cx<R> (&v)[N / T], cx<R>* lds, const cx<R>* __restrict__ tw, int t)
{
    cx<R>(&vb)[1][N / T] = *reinterpret_cast<cx<R>(*)[1][N / T]>(&v);
    fft_block<N, T, SIGN, 1, R, DB, P0, TRAIL, RS>(vb, lds, tw, t);
}

// 1/x for finite x of moderate exponent (no denormal, infinite or
// near-overflow operands): FP64 as v_rcp_f64 refined by two Newton steps,
// within an ulp of the quotient (the IEEE division adds operand scaling and a
// final correction: 11 instructions against 5); FP32 as v_rcp_f32 (1 ulp).
template <typename R>
__device__ __forceinline__ R rcp_fast(R x)
{
    if constexpr (sizeof(R) == 4) {
        return __builtin_amdgcn_rcpf(x);
    } else {
        R r = __builtin_amdgcn_rcp(x);
        R e = __builtin_fma(-x, r, (R)1);
        r = __builtin_fma(r, e, r);
        e = __builtin_fma(-x, r, (R)1);
        return __builtin_fma(r, e, r);
    }
}

// Wave sum by register exchanges, no LDS: each of the six steps adds a lane's
// value and a partner's (DPP quad permutes 0xB1 = lane ^ 1 and 0x4E = lane ^ 2,
// row_half_mirror 0x141 = lane ^ 7 within 8 lanes, row_mirror 0x140 = lane ^ 15
// within 16, then v_permlane16_swap and v_permlane32_swap across rows).  After
// each step all lanes of a group hold the same bits, and partners add the same
// two operands (IEEE addition commutes), so every lane returns the same bits:
// a loop exit taken on the sum is uniform across the workgroup.
template <int CTRL, typename A>
__device__ __forceinline__ A wave_add_dpp(A v)
{
    constexpr int W = (int)(sizeof(A) / 4);
    uint32_t u[W], p[W];
    __builtin_memcpy(u, &v, sizeof(A));
#pragma unroll
    for (int w = 0; w < W; ++w) p[w] = (uint32_t)__builtin_amdgcn_mov_dpp((int)u[w], CTRL, 0xF, 0xF, true);
    A o;
    __builtin_memcpy(&o, p, sizeof(A));
    return v + o;
}

template <int C, typename A>
__device__ __forceinline__ A wave_add_swap(A v)
{
    constexpr int W = (int)(sizeof(A) / 4);
    uint32_t u[W], a[W], b[W];
    __builtin_memcpy(u, &v, sizeof(A));
#pragma unroll
    for (int w = 0; w < W; ++w) {
        if constexpr (C == 4) {
            const auto r = __builtin_amdgcn_permlane16_swap(u[w], u[w], false, false);
            a[w] = r[0];
            b[w] = r[1];
        } else {
            const auto r = __builtin_amdgcn_permlane32_swap(u[w], u[w], false, false);
            a[w] = r[0];
            b[w] = r[1];
        }
    }
    // one of the two holds this lane's row value, the other the partner row's
    A x, y;
    __builtin_memcpy(&x, a, sizeof(A));
    __builtin_memcpy(&y, b, sizeof(A));
    return x + y;
}

template <typename A>
__device__ __forceinline__ A wave_sum(A v)
{
    v = wave_add_dpp<0xB1>(v);
    v = wave_add_dpp<0x4E>(v);
    v = wave_add_dpp<0x141>(v);
    v = wave_add_dpp<0x140>(v);
    v = wave_add_swap<4>(v);
    return wave_add_swap<5>(v);
}

// Arg-max over the wave by the same exchange pattern as wave_sum: the larger
// value wins, the lower index on ties (the first row of the largest |a|, as
// LAPACK's i?amax).  The rule is a strict order on (value, index), so both
// lanes of a pair keep the same winner and every lane returns the same pair.
// Values must not be NaN (callers start from -1 and take only v > best).
// (Bitwise | and & and selects: branches around the two moves would cost
// more than the moves.)
template <typename A>
__device__ __forceinline__ void argmax_pick(A &v, int &i, A ov, int oi)
{
    const bool take = (ov > v) | ((ov == v) & (oi < i));
    v = take ? ov : v;
    i = take ? oi : i;
}
template <int CTRL, typename A>
__device__ __forceinline__ void argmax_dpp(A &v, int &i)
{
    constexpr int W = (int)(sizeof(A) / 4);
    uint32_t u[W];
    __builtin_memcpy(u, &v, sizeof(A));
#pragma unroll
    for (int w = 0; w < W; ++w) u[w] = (uint32_t)__builtin_amdgcn_mov_dpp((int)u[w], CTRL, 0xF, 0xF, true);
    A ov;
    __builtin_memcpy(&ov, u, sizeof(A));
    const int oi = __builtin_amdgcn_mov_dpp(i, CTRL, 0xF, 0xF, true);
    argmax_pick(v, i, ov, oi);
}
template <int C, typename A>
__device__ __forceinline__ void argmax_swap(A &v, int &i)
{
    constexpr int W = (int)(sizeof(A) / 4);
    uint32_t u[W], a[W], b[W];
    __builtin_memcpy(u, &v, sizeof(A));
#pragma unroll
    for (int w = 0; w < W; ++w) {
        if constexpr (C == 4) {
            const auto r = __builtin_amdgcn_permlane16_swap(u[w], u[w], false, false);
            a[w] = r[0];
            b[w] = r[1];
        } else {
            const auto r = __builtin_amdgcn_permlane32_swap(u[w], u[w], false, false);
            a[w] = r[0];
            b[w] = r[1];
        }
    }
    const uint32_t ui = (uint32_t)i;
    const auto ri = C == 4 ? __builtin_amdgcn_permlane16_swap(ui, ui, false, false)
                           : __builtin_amdgcn_permlane32_swap(ui, ui, false, false);
    // the two candidates are this lane's row value and the partner row's
    A x, y;
    __builtin_memcpy(&x, a, sizeof(A));
    __builtin_memcpy(&y, b, sizeof(A));
    v = x;
    i = (int)ri[0];
    argmax_pick(v, i, y, (int)ri[1]);
}
template <typename A>
__device__ __forceinline__ void wave_argmax(A &v, int &i)
{
    argmax_dpp<0xB1>(v, i);
    argmax_dpp<0x4E>(v, i);
    argmax_dpp<0x141>(v, i);
    argmax_dpp<0x140>(v, i);
    argmax_swap<4>(v, i);
    argmax_swap<5>(v, i);
}

// Workgroup sum of one value per thread (T threads, T a multiple of 64).
// red: LDS scratch of at least T/64 values.  Returns the same bits to every
// thread.  For T > 64 the one barrier inside also publishes the caller's
// earlier LDS writes; red is reused without a trailing barrier, so two calls
// must be separated by another barrier of the caller's.
template <int T, typename A>
__device__ __forceinline__ A block_sum(A v, A* red, int t)
{
    v = wave_sum(v);
    if constexpr (T == 64) {
        return v;
    } else {
        if ((t & 63) == 0) red[t >> 6] = v;
        __syncthreads();
        A s = red[0];
#pragma unroll
        for (int w = 1; w < T / 64; ++w) s += red[w];
        return s;
    }
}

// Workgroup sums of two values per thread in one pass (one LDS round trip).
// red: LDS scratch of at least 2*T/64 values.  Same barrier contract as
// block_sum.
template <int T, typename A>
__device__ __forceinline__ void block_sum2(A &a, A &b, A* red, int t)
{
    a = wave_sum(a);
    b = wave_sum(b);
    if constexpr (T > 64) {
        if ((t & 63) == 0) {
            red[t >> 6] = a;
            red[T / 64 + (t >> 6)] = b;
        }
        __syncthreads();
        A sa = red[0], sb = red[T / 64];
#pragma unroll
        for (int w = 1; w < T / 64; ++w) {
            sa += red[w];
            sb += red[T / 64 + w];
        }
        a = sa;
        b = sb;
    }
}

// Spectral wavenumber m of index k in a length-N spectrum (SpectralTransformer.cpp:150)
template <int N>
__device__ __forceinline__ int wavenumber(int k)
{
    return (k < N / 2) ? k : k - N;
}

}  // namespace ccd
