// cc_lanefft.hpp -- radix-P transforms whose digit moves stay inside a
// wavefront wherever they can.
//
// A workgroup of T threads (T >= 64) holds a length-N signal, P = N/T values
// per thread.  An element's coordinates are its register index (log2 P bits),
// its lane (6 bits) and its wave (log2(T/64) bits).  The time-domain layout is
// layout L of cc_device.hpp: index bit i < 6 is lane bit i, the next bits are
// the wave bits, the top log2 P bits are the register bits.
//
// DIF (time -> spectrum) is the in-place Gentleman-Sande decomposition over
// digits of the index taken from the top: the first digit is the register
// bits, then a partial digit if log2 N is not a multiple of log2 P, then full
// digits.  Before each stage the next digit's bits are swapped into the
// register bits.  A swap of a register bit with a lane bit is a register
// exchange between lane pairs: v_permlane16_swap / v_permlane32_swap for lane
// bits 4 and 5, DPP row shifts or quad permutes for lane bits 0-3, one VALU
// move per dword and no LDS.  Only a digit with a wave bit goes through LDS
// (one write, one barrier, one read): at T = N/4 one exchange per transform
// instead of the Stockham form's four.  The output is left in place: the
// spectrum's digits are the reversal of the time digits ("layout S", mapped
// by lf_kt / lf_kr).  DIT (spectrum -> time) runs the stages in reverse with
// the twiddles before the butterflies and returns layout L.
//
// Twiddles: the DIF stage on position bits [lo, lo + s) multiplies output d'
// by w^{e d'}, w = e^{SIGN 2 pi i / 2^{lo+s}}, e = position bits below lo,
// which are still at their layout-L places (thread-index bits [0, lo)); DIT
// applies the same factors to its inputs.
#pragma once
#include "cc_device.hpp"

namespace ccd {

constexpr int lf_log2(int x)
{
    int l = 0;
    while ((1 << l) < x) ++l;
    return l;
}

// Coordinate codes: LF_LANE + b (b < 6) lane bit b; LF_WAVE + b wave bit b
// (thread-index bit 6 + b); LF_REG + b register bit b.
constexpr int LF_LANE = 0, LF_WAVE = 8, LF_REG = 16;

struct LfPlan {
    int nb, pb, wb, nd;
    int lo[12], sz[12];  // DIF stage d butterflies position bits [lo[d], lo[d] + sz[d])
    int sw[12][4];       // before DIF stage d (d > 0): register bit m swaps with coordinate sw[d][m]
    int lds[12];         // stage d's swaps include a wave bit: one LDS exchange for all of them
    int loc[12][12];     // loc[d][i]: coordinate of position bit i during DIF stage d
    int koff[12];        // the spectral digit of stage d is spectral-index bits [koff[d], koff[d] + sz[d])
    int nlds;            // LDS exchanges per transform
    bool ok;
};

constexpr LfPlan lf_plan(int N, int T)
{
    LfPlan p{};
    p.nb = lf_log2(N);
    p.pb = lf_log2(N / T);
    p.wb = lf_log2(T / 64);
    p.ok = T >= 64 && N / T >= 2 && p.pb <= 4 && p.nb <= 12;
    if (!p.ok) return p;
    int sizes[12] = {};
    int nd = 0, rest = p.nb - p.pb;
    sizes[nd++] = p.pb;
    if (rest % p.pb) {
        sizes[nd++] = rest % p.pb;
        rest -= rest % p.pb;
    }
    while (rest > 0) {
        sizes[nd++] = p.pb;
        rest -= p.pb;
    }
    p.nd = nd;
    int hi = p.nb;
    for (int d = 0; d < nd; ++d) {
        p.sz[d] = sizes[d];
        p.lo[d] = hi - sizes[d];
        hi -= sizes[d];
    }
    int loc[12] = {};
    for (int i = 0; i < p.nb; ++i)
        loc[i] = i < 6 ? LF_LANE + i : (i < 6 + p.wb ? LF_WAVE + (i - 6) : LF_REG + (i - 6 - p.wb));
    int off = 0;
    for (int d = 0; d < nd; ++d) {
        p.lds[d] = 0;
        if (d > 0) {
            for (int m = 0; m < p.sz[d]; ++m) {
                const int src = loc[p.lo[d] + m];
                int holder = -1;
                for (int i = 0; i < p.nb; ++i)
                    if (loc[i] == LF_REG + m) holder = i;
                if (holder < 0 || src >= LF_REG) p.ok = false;   // unprocessed bits are lane or wave bits
                p.sw[d][m] = src;
                if (holder >= 0) loc[holder] = src;
                loc[p.lo[d] + m] = LF_REG + m;
                if (src >= LF_WAVE && src < LF_REG) p.lds[d] = 1;
            }
        }
        // the twiddle exponent of stage d is thread-index bits [0, lo[d])
        for (int i = 0; i < p.lo[d]; ++i) {
            const int want = i < 6 ? LF_LANE + i : LF_WAVE + (i - 6);
            if (loc[i] != want) p.ok = false;
        }
        for (int i = 0; i < p.nb; ++i) p.loc[d][i] = loc[i];
        p.koff[d] = off;
        off += p.sz[d];
        p.nlds += p.lds[d];
    }
    return p;
}

template <int N, int T>
struct LF {
    static constexpr LfPlan p = lf_plan(N, T);
    static_assert(p.ok, "unsupported lane-transform shape");
    static constexpr int P = N / T;
};

template <int V>
struct ic {
    static constexpr int value = V;
};
template <int B, int E, typename F>
__host__ __device__ __forceinline__ void sfor(F &&f)
{
    if constexpr (B < E) {
        f(ic<B>{});
        sfor<B + 1, E>(f);
    }
}

// thread-index bit of a lane or wave coordinate
constexpr int lf_tbit(int c) { return c < LF_WAVE ? c : 6 + (c - LF_WAVE); }

// Position (natural in-place index) of register r's element during DIF stage
// D: thread part and register part.
template <int N, int T, int D>
__host__ __device__ __forceinline__ int lf_pos_t(int t)
{
    int x = 0;
    sfor<0, LF<N, T>::p.nb>([&](auto I) {
        constexpr int c = LF<N, T>::p.loc[D][decltype(I)::value];
        if constexpr (c < LF_REG) x |= ((t >> lf_tbit(c)) & 1) << decltype(I)::value;
    });
    return x;
}
template <int N, int T, int D>
constexpr int lf_pos_r(int r)
{
    int x = 0;
    for (int i = 0; i < LF<N, T>::p.nb; ++i) {
        const int c = LF<N, T>::p.loc[D][i];
        if (c >= LF_REG) x |= ((r >> (c - LF_REG)) & 1) << i;
    }
    return x;
}

// Spectral index of register r's element in layout S: lf_kt(t) ^ lf_kr(r).
template <int N, int T>
constexpr int lf_kbit_coord(int j)
{
    // spectral bit j = position bit lo[d] + (j - koff[d]) of the stage d whose digit holds j
    constexpr LfPlan p = LF<N, T>::p;
    for (int d = 0; d < p.nd; ++d)
        if (j >= p.koff[d] && j < p.koff[d] + p.sz[d]) return p.loc[p.nd - 1][p.lo[d] + (j - p.koff[d])];
    return -1;
}
template <int N, int T>
__host__ __device__ __forceinline__ int lf_kt(int t)
{
    int k = 0;
    sfor<0, LF<N, T>::p.nb>([&](auto J) {
        constexpr int c = lf_kbit_coord<N, T>(decltype(J)::value);
        if constexpr (c < LF_REG) k |= ((t >> lf_tbit(c)) & 1) << decltype(J)::value;
    });
    return k;
}
template <int N, int T>
constexpr int lf_kr(int r)
{
    int k = 0;
    for (int j = 0; j < LF<N, T>::p.nb; ++j) {
        const int c = lf_kbit_coord<N, T>(j);
        if (c >= LF_REG) k |= ((r >> (c - LF_REG)) & 1) << j;
    }
    return k;
}
// Owner of spectral index q in layout S: thread (first) and register (second), packed t * 64 + r.
template <int N, int T>
constexpr int lf_owner(int q)
{
    int t = 0, r = 0;
    for (int j = 0; j < LF<N, T>::p.nb; ++j) {
        const int c = lf_kbit_coord<N, T>(j);
        const int b = (q >> j) & 1;
        if (c >= LF_REG) r |= b << (c - LF_REG);
        else t |= b << lf_tbit(c);
    }
    return t * 64 + r;
}

// One dword of a register pair (a: register bit M clear, b: set) swaps with the
// partner lane across lane bit C: lanes with bit C clear keep a and receive the
// partner's a into b; lanes with bit C set keep b and receive the partner's b
// into a.
template <int C>
__device__ __forceinline__ void lf_swap_dword(uint32_t &a, uint32_t &b, bool up)
{
    if constexpr (C == 4) {
        const auto r = __builtin_amdgcn_permlane16_swap(a, b, false, false);
        a = r[0];
        b = r[1];
    } else if constexpr (C == 5) {
        const auto r = __builtin_amdgcn_permlane32_swap(a, b, false, false);
        a = r[0];
        b = r[1];
    } else if constexpr (C >= 2) {
        // DPP row shifts by 2^C (row_shr reads lane L - 2^C, row_shl lane L + 2^C);
        // the bank mask (groups of 4 lanes of a row) writes only the lanes
        // with bit C set (a) or clear (b), the others keep their value.
        constexpr int SHR = C == 2 ? 0x114 : 0x118, SHL = C == 2 ? 0x104 : 0x108;
        constexpr int BANK_SET = C == 2 ? 0xA : 0xC, BANK_CLR = C == 2 ? 0x5 : 0x3;
        const uint32_t a0 = a;
        a = (uint32_t)__builtin_amdgcn_update_dpp((int)a, (int)b, SHR, 0xF, BANK_SET, false);
        b = (uint32_t)__builtin_amdgcn_update_dpp((int)b, (int)a0, SHL, 0xF, BANK_CLR, false);
    } else {
        // within a quad: each lane sends one value (bit set: its a, clear:
        // its b) to its partner (quad permute lane ^ 2^C)
        constexpr int XP = C == 0 ? 0xB1 : 0x4E;
        const uint32_t snd = up ? a : b;
        const uint32_t rcv = (uint32_t)__builtin_amdgcn_mov_dpp((int)snd, XP, 0xF, 0xF, true);
        a = up ? rcv : a;
        b = up ? b : rcv;
    }
}

template <int M, int C, int B, int P, typename R>
__device__ __forceinline__ void lf_lane_swap(cx<R> (&v)[B][P], int t)
{
    static_assert(C < 6, "lane coordinate");
    constexpr int W = (int)(sizeof(cx<R>) / 4);
    const bool up = (t >> C) & 1;
#pragma unroll
    for (int s = 0; s < B; ++s)
#pragma unroll
        for (int r0 = 0; r0 < P; ++r0) {
            if ((r0 >> M) & 1) continue;
            const int r1 = r0 | (1 << M);
            uint32_t a[W], b[W];
            __builtin_memcpy(a, &v[s][r0], sizeof(a));
            __builtin_memcpy(b, &v[s][r1], sizeof(b));
#pragma unroll
            for (int w = 0; w < W; ++w) lf_swap_dword<C>(a[w], b[w], up);
            __builtin_memcpy(&v[s][r0], a, sizeof(a));
            __builtin_memcpy(&v[s][r1], b, sizeof(b));
        }
}

// The swaps of DIF stage D through LDS: write at the positions of stage D-1's
// map, read at stage D's (REV: the other way round, for DIT).  Positions are
// the natural in-place index, so the lanes of one LDS cycle (position bits
// 0-2 are lane bits 0-2 at every exchange) touch distinct banks.  B signals N
// apart; region PAR of two RS apart with DB.
template <int N, int T, int D, bool REV, int B, typename R, int DB, int PAR, int RS>
__device__ __forceinline__ void lf_lds_swap(cx<R> (&v)[B][N / T], cx<R> *lds, int t)
{
    constexpr int P = N / T;
    constexpr int DW = REV ? D : D - 1, DR = REV ? D - 1 : D;
    cx<R> *base = lds + PAR * RS;
    const int wt = lf_pos_t<N, T, DW>(t);
    sfor<0, P>([&](auto Rr) {
        constexpr int r = decltype(Rr)::value;
        constexpr int pr = lf_pos_r<N, T, DW>(r);
#pragma unroll
        for (int s = 0; s < B; ++s) base[s * N + (wt | pr)] = v[s][r];
    });
    __syncthreads();
    const int rt = lf_pos_t<N, T, DR>(t);
    sfor<0, P>([&](auto Rr) {
        constexpr int r = decltype(Rr)::value;
        constexpr int pr = lf_pos_r<N, T, DR>(r);
#pragma unroll
        for (int s = 0; s < B; ++s) v[s][r] = base[s * N + (rt | pr)];
    });
    if constexpr (!DB) __syncthreads();
}

// a * w for SIGN > 0, a * conj(w) for SIGN < 0 (tables hold e^{+2 pi i e / N})
template <int SIGN, typename R>
__device__ __forceinline__ cx<R> cmul_tw(cx<R> a, cx<R> w)
{
    if constexpr (SIGN > 0) return cmul(a, w);
    else return {a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y};
}

// First twiddle w = e^{2 pi i e / 2^{lo+s}} of every stage of radix <= 4 with
// lo > 0.  DIF and DIT stages on the same digit share it (DIT, or SIGN < 0,
// multiplies by the conjugate), so one set per thread serves every transform
// of a workgroup; a caller that can hold it in registers loads it once
// (lf_twiddles) and passes it to the transforms.
template <int N, int T, typename R>
struct LfTw {
    cx<R> w[12];
};
template <int N, int T, typename R>
__device__ __forceinline__ LfTw<N, T, R> lf_twiddles(const cx<R> *__restrict__ tw, int t)
{
    LfTw<N, T, R> r{};
    sfor<0, LF<N, T>::p.nd>([&](auto Dd) {
        constexpr int d = decltype(Dd)::value;
        constexpr int LO = LF<N, T>::p.lo[d], S = LF<N, T>::p.sz[d];
        if constexpr (LO > 0 && S <= 2) r.w[d] = tw[(t & ((1 << LO) - 1)) * (N >> (LO + S))];
    });
    return r;
}

// Radix 2^S butterflies over register bits [0, S) on position bits [LO, LO + S)
// (DIF stage D).  Radix <= 4: one table load (or the preloaded factor, PRE)
// and products of powers (w^2 = w w, w^3 = w w^2); loading every power instead
// makes the chains slower (the loads' latency is on the dependent path);
// larger radices load every power.
template <int N, int T, int SIGN, bool DIT, int D, bool PRE, int B, typename R>
__device__ __forceinline__ void lf_stage(cx<R> (&v)[B][N / T], const cx<R> *__restrict__ tw, int t,
                                         const LfTw<N, T, R> *pre)
{
    constexpr int LO = LF<N, T>::p.lo[D], S = LF<N, T>::p.sz[D];
    constexpr int P = N / T, RAD = 1 << S, G = P / RAD;
    cx<R> w[RAD];
    if constexpr (LO > 0) {
        int e = t & ((1 << LO) - 1);
        asm volatile("" : "+v"(e));
        constexpr int STRIDE = N >> (LO + S);
        if constexpr (RAD <= 4) {
            if constexpr (PRE) w[1] = pre->w[D];
            else w[1] = tw[e * STRIDE];
            if constexpr (RAD == 4) {
                w[2] = cmul(w[1], w[1]);
                w[3] = cmul(w[1], w[2]);
            }
        } else {
#pragma unroll
            for (int d = 1; d < RAD; ++d) w[d] = tw[d * e * STRIDE];
        }
    }
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int s = 0; s < B; ++s) {
            cx<R> u[RAD];
#pragma unroll
            for (int d = 0; d < RAD; ++d) u[d] = v[s][g * RAD + d];
            if constexpr (DIT && LO > 0) {
#pragma unroll
                for (int d = 1; d < RAD; ++d) u[d] = cmul_tw<SIGN>(u[d], w[d]);
            }
            dft<RAD, SIGN>(u);
            if constexpr (!DIT && LO > 0) {
#pragma unroll
                for (int d = 1; d < RAD; ++d) u[d] = cmul_tw<SIGN>(u[d], w[d]);
            }
#pragma unroll
            for (int d = 0; d < RAD; ++d) v[s][g * RAD + d] = u[d];
        }
}

// number of LDS exchanges among DIF stages [1, D)
template <int N, int T>
constexpr int lf_lds_before(int D)
{
    int n = 0;
    for (int d = 1; d < D; ++d) n += LF<N, T>::p.lds[d];
    return n;
}

// Time (layout L) -> spectrum (layout S); SIGN as fft_block.  B signals.
// DB / P0: exchange regions alternate from region P0 (as fft_block); with no
// trailing barrier the caller's next LDS writer uses region
// lf_next_par<N, T>(P0).
template <int N, int T, int SIGN, int B, typename R, int DB = 0, int P0 = 0, int RS = B * N, bool PRE = false>
__device__ __forceinline__ void lfft_dif(cx<R> (&v)[B][N / T], cx<R> *lds, const cx<R> *__restrict__ tw, int t,
                                         const LfTw<N, T, R> *pre = nullptr)
{
    constexpr LfPlan p = LF<N, T>::p;
    static_assert(DB || P0 == 0, "one exchange region without double buffering");
    sfor<0, p.nd>([&](auto Dd) {
        constexpr int d = decltype(Dd)::value;
        if constexpr (d > 0) {
            if constexpr (p.lds[d]) {
                constexpr int par = DB ? (P0 ^ (lf_lds_before<N, T>(d) & 1)) : 0;
                lf_lds_swap<N, T, d, false, B, R, DB, par, RS>(v, lds, t);
            } else {
                sfor<0, p.sz[d]>([&](auto Mm) {
                    constexpr int m = decltype(Mm)::value;
                    lf_lane_swap<m, p.sw[d][m] - LF_LANE>(v, t);
                });
            }
        }
        lf_stage<N, T, SIGN, false, d, PRE>(v, tw, t, pre);
    });
}

// Spectrum (layout S) -> time (layout L).
template <int N, int T, int SIGN, int B, typename R, int DB = 0, int P0 = 0, int RS = B * N, bool PRE = false>
__device__ __forceinline__ void lfft_dit(cx<R> (&v)[B][N / T], cx<R> *lds, const cx<R> *__restrict__ tw, int t,
                                         const LfTw<N, T, R> *pre = nullptr)
{
    constexpr LfPlan p = LF<N, T>::p;
    static_assert(DB || P0 == 0, "one exchange region without double buffering");
    sfor<0, p.nd>([&](auto Jj) {
        constexpr int d = p.nd - 1 - decltype(Jj)::value;
        lf_stage<N, T, SIGN, true, d, PRE>(v, tw, t, pre);
        if constexpr (d > 0) {
            if constexpr (p.lds[d]) {
                // exchanges counted in DIT order: those of stages above d come first
                constexpr int par = DB ? (P0 ^ ((p.nlds - 1 - lf_lds_before<N, T>(d)) & 1)) : 0;
                lf_lds_swap<N, T, d, true, B, R, DB, par, RS>(v, lds, t);
            } else {
                sfor<0, p.sz[d]>([&](auto Mm) {
                    constexpr int m = p.sz[d] - 1 - decltype(Mm)::value;
                    lf_lane_swap<m, p.sw[d][m] - LF_LANE>(v, t);
                });
            }
        }
    });
}

template <int N, int T>
constexpr int lf_next_par(int p0) { return (LF<N, T>::p.nlds % 2 == 1) ? 1 - p0 : p0; }

// Single-signal wrappers.
template <int N, int T, int SIGN, typename R, int DB = 0, int P0 = 0, int RS = N, bool PRE = false>
__device__ __forceinline__ void lfft_dif1(cx<R> (&v)[N / T], cx<R> *lds, const cx<R> *__restrict__ tw, int t,
                                          const LfTw<N, T, R> *pre = nullptr)
{
    lfft_dif<N, T, SIGN, 1, R, DB, P0, RS, PRE>(*reinterpret_cast<cx<R>(*)[1][N / T]>(&v), lds, tw, t, pre);
}
template <int N, int T, int SIGN, typename R, int DB = 0, int P0 = 0, int RS = N, bool PRE = false>
__device__ __forceinline__ void lfft_dit1(cx<R> (&v)[N / T], cx<R> *lds, const cx<R> *__restrict__ tw, int t,
                                          const LfTw<N, T, R> *pre = nullptr)
{
    lfft_dit<N, T, SIGN, 1, R, DB, P0, RS, PRE>(*reinterpret_cast<cx<R>(*)[1][N / T]>(&v), lds, tw, t, pre);
}

// LDS position of spectral mode k for an exchange by mode index (the mirror
// k <-> N - k).  In layout S consecutive lanes hold modes far apart, so the
// natural index would put the lanes of one LDS cycle (8 for 16-byte, 16 for
// 8-byte elements) on the same banks; the mode bits that lane bits 0..LB-1
// carry are XORed into position bits 0..LB-1 (only from mode bits >= LB, so
// the map stays a bijection of [0, N)).
template <int N, int T, typename R>
__host__ __device__ __forceinline__ int lf_mirror_pos(int k)
{
    constexpr int LB = sizeof(cx<R>) >= 16 ? 3 : 4;
    int x = k;
    sfor<0, LF<N, T>::p.nb>([&](auto J) {
        constexpr int j = decltype(J)::value;
        constexpr int c = lf_kbit_coord<N, T>(j);
        if constexpr (j >= LB && c >= LF_LANE && c < LF_LANE + LB) x ^= ((k >> j) & 1) << (c - LF_LANE);
    });
    return x;
}

}  // namespace ccd
