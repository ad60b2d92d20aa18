// cc_capi.hip -- host side of the C-ABI declared in include/cc_gpu.h.
//
// A context owns the device buffers for one problem size, two HIP streams
// (default priority), the FFT twiddle tables and the x-grid.  A "column" is
// one shoot of the reference (NewtonSolver::shoot, NewtonSolver.cpp:403-423)
// for x perturbed in one slot (or unperturbed, the base); a "chain" is one side
// of a column integrated to the match point.  Every entry point below runs a
// set of columns through
//     k_boundary (left/right series) -> k_chain (IRK2) -> k_mismatch (pack)
// and then k_fd / the LU kernels as needed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/cc_gpu.h"
#include "cc_kernels.hpp"
#include "cc_tangent.hpp"

using ccd::cx;

namespace {

constexpr int kStepsPerLaunch = 500;

constexpr int tpb(int N) { return N == 128 ? 64 : N / 4; }

}  // namespace

// Device arrays of one launch of columns (see launch_columns).
struct ColumnBuffers {
    void *Y = nullptr;                     // [maxchains][M] complex (double or float)
    double *delta = nullptr, *F = nullptr; // per chain Delta; [maxcols][n] residuals
    int32_t *start = nullptr, *nsteps = nullptr, *dir = nullptr, *status = nullptr, *fail = nullptr;
    int32_t *lcols = nullptr, *rcols = nullptr, *lch = nullptr, *rch = nullptr;
    int32_t *lidx = nullptr, *ridx = nullptr, *mstatus = nullptr;
    int64_t *its = nullptr;
};

struct cc_ctx {
    cc_problem p{};
    int N = 0, M = 0, n = 0, X = 0, iM = 0, iR = 0;
    hipStream_t s0 = nullptr, s1 = nullptr;
    hipEvent_t ev = nullptr;
    std::string err;
    int64_t counters[3] = {0, 0, 0};
    double a[2][2], b[2], c[2];
    double *grid = nullptr;
    cx<double> *tw64 = nullptr;
    cx<float> *tw32 = nullptr;
    double *invm64 = nullptr;            // 1/m per spectral index (0 at DC and Nyquist)
    float *invm32 = nullptr;
    double *x = nullptr, *F0 = nullptr, *dx = nullptr, *r = nullptr, *d = nullptr;
    double *J = nullptr;                 // n x n
    double *A = nullptr;                 // n x n LU work (FP64)
    float *A32 = nullptr;                // n x n LU work (FP32)
    int32_t *ipiv = nullptr, *info = nullptr, *cols = nullptr;
    ColumnBuffers main;                  // the Jacobian columns
    ColumnBuffers base;                  // FP64 base (+ Delta column) of the mixed policy
    // tangent-linear columns of the mixed policy
    float *rec = nullptr;                // base-trajectory record, X steps
    int64_t *rec_base_b = nullptr;       // record slots of the FP64 base launch's chains
    int64_t *rec_base_t = nullptr;       // record slots of the tangent chains
    cx<float> *dY = nullptr;             // [n][M] tangent states
    int32_t *t_start = nullptr, *t_nsteps = nullptr, *t_dir = nullptr, *t_status = nullptr, *t_fail = nullptr;
    int32_t *t_cols = nullptr, *t_lcols = nullptr, *t_rcols = nullptr, *t_lch = nullptr, *t_rch = nullptr;
    int32_t *t_lidx = nullptr, *t_ridx = nullptr;
    int64_t *t_its = nullptr;
    double tangent_rtol = 1e-6, tangent_h = 1e-5;
    int maxcols = 0, maxchains = 0;
    std::vector<void *> allocs;
    // chain-kernel timing of the main stream
    bool profile = false;
    hipEvent_t tk0 = nullptr, tk1 = nullptr;
    int32_t prof_launches = 0;
    int64_t prof_its = 0, prof_steps = 0;
    bool prof_pending = false;
};

namespace {

bool fail(cc_ctx *c, const char *what, hipError_t e)
{
    if (e == hipSuccess) return false;
    c->err = std::string(what) + ": " + hipGetErrorString(e);
    return true;
}

#define HIPCHK(ctx, call)                                              \
    do {                                                               \
        hipError_t e_ = (call);                                        \
        if (fail((ctx), #call, e_)) return CC_DEVICE;                  \
    } while (0)

template <typename Tp>
hipError_t dalloc(cc_ctx *c, Tp **p, size_t count)
{
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(Tp));
    if (e == hipSuccess) {
        c->allocs.push_back(q);
        *p = static_cast<Tp *>(q);
    }
    return e;
}

// ---------------------------------------------------------------------------
// N dispatch: every kernel is instantiated for the five supported lengths.
// ---------------------------------------------------------------------------
template <typename Fn>
bool dispatch_n(int N, Fn &&fn)
{
    switch (N) {
    case 128: fn(std::integral_constant<int, 128>{}); return true;
    case 256: fn(std::integral_constant<int, 256>{}); return true;
    case 512: fn(std::integral_constant<int, 512>{}); return true;
    case 1024: fn(std::integral_constant<int, 1024>{}); return true;
    case 2048: fn(std::integral_constant<int, 2048>{}); return true;
    default: return false;
    }
}

// The set of chains needed by a list of columns.
struct Plan {
    std::vector<int32_t> cols;          // perturbed slot per column (-1: base)
    std::vector<int32_t> lcols, lch;    // left chains: column slot, chain id
    std::vector<int32_t> rcols, rch;    // right chains
    std::vector<int32_t> lidx, ridx;    // per column: chain ids of its two sides
    std::vector<int32_t> start, nsteps, dir;
    int nchains = 0;
};

bool touches_left(int N, int col) { return col < 0 || (col >= N / 4 && col < 3 * N / 4); }
bool touches_right(int N, int col) { return col < 0 || col < N / 4 || col == N / 2 + 2; }

// Columns that leave one side's boundary data untouched reuse the base
// column's chain of that side (the reference recomputes it; the result is the
// same state bit for bit).  The base column must be in the list for reuse.
Plan make_plan(const cc_ctx *c, const std::vector<int32_t> &cols, bool reuse)
{
    Plan P;
    P.cols = cols;
    const int N = c->N;
    int base = -1;
    for (size_t b = 0; b < cols.size(); ++b)
        if (cols[b] < 0) base = (int)b;
    if (base < 0) reuse = false;
    P.lidx.assign(cols.size(), -1);
    P.ridx.assign(cols.size(), -1);
    auto add = [&](int col, bool left) {
        const int id = P.nchains++;
        if (left) {
            P.lcols.push_back(col);
            P.lch.push_back(id);
            P.start.push_back(0);
            P.nsteps.push_back(c->iM);
            P.dir.push_back(1);
        } else {
            P.rcols.push_back(col);
            P.rch.push_back(id);
            P.start.push_back(c->iR);
            P.nsteps.push_back(c->iR - c->iM);
            P.dir.push_back(-1);
        }
        return id;
    };
    if (base >= 0) {
        P.lidx[base] = add(-1, true);
        P.ridx[base] = add(-1, false);
    }
    for (size_t b = 0; b < cols.size(); ++b) {
        if ((int)b == base) continue;
        const int col = cols[b];
        P.lidx[b] = (reuse && !touches_left(N, col)) ? P.lidx[base] : add(col, true);
        P.ridx[b] = (reuse && !touches_right(N, col)) ? P.ridx[base] : add(col, false);
    }
    return P;
}

template <typename Tp>
hipError_t up(const std::vector<Tp> &v, Tp *d, hipStream_t s)
{
    if (v.empty()) return hipSuccess;
    return hipMemcpyAsync(d, v.data(), v.size() * sizeof(Tp), hipMemcpyHostToDevice, s);
}

// Run the columns of plan P in precision R on stream s.  Outputs: F[b*n..],
// per-chain status and per-column mismatch status in B.
template <int N, typename R>
cc_status launch_columns(cc_ctx *c, const Plan &P, const ColumnBuffers &B, const double *d_x, double eps,
                         double precision, hipStream_t s, float *rec = nullptr, const int64_t *rec_base = nullptr)
{
    constexpr int T = tpb(N);
    const int nch = P.nchains;
    const int ncols = (int)P.cols.size();
    HIPCHK(c, up(P.start, B.start, s));
    HIPCHK(c, up(P.nsteps, B.nsteps, s));
    HIPCHK(c, up(P.dir, B.dir, s));
    HIPCHK(c, up(P.lcols, B.lcols, s));
    HIPCHK(c, up(P.rcols, B.rcols, s));
    HIPCHK(c, up(P.lch, B.lch, s));
    HIPCHK(c, up(P.rch, B.rch, s));
    HIPCHK(c, up(P.lidx, B.lidx, s));
    HIPCHK(c, up(P.ridx, B.ridx, s));
    HIPCHK(c, hipMemsetAsync(B.status, 0, sizeof(int32_t) * nch, s));
    HIPCHK(c, hipMemsetAsync(B.fail, 0xff, sizeof(int32_t) * nch, s));
    HIPCHK(c, hipMemsetAsync(B.its, 0, sizeof(int64_t) * nch, s));
    HIPCHK(c, hipMemsetAsync(B.mstatus, 0, sizeof(int32_t) * ncols, s));
    cx<R> *Y = static_cast<cx<R> *>(B.Y);
    if (!P.lcols.empty())
        hipLaunchKernelGGL((ccd::k_boundary<N, T, R, true>), dim3((unsigned)P.lcols.size()), dim3(T), 0, s, d_x,
                           B.lcols, B.lch, eps, c->p.XLeft, c->p.Dim, c->tw64, Y, B.delta);
    if (!P.rcols.empty())
        hipLaunchKernelGGL((ccd::k_boundary<N, T, R, false>), dim3((unsigned)P.rcols.size()), dim3(T), 0, s, d_x,
                           B.rcols, B.rch, eps, c->p.XRight, c->p.Dim, c->tw64, Y, B.delta);
    HIPCHK(c, hipGetLastError());
    ccd::ChainParams<R> a{};
    a.Y = Y;
    a.start = B.start;
    a.nsteps = B.nsteps;
    a.dir = B.dir;
    a.delta = B.delta;
    a.grid = c->grid;
    a.status = B.status;
    a.fail_idx = B.fail;
    a.its = B.its;
    if constexpr (sizeof(R) == 8) {
        a.tw = reinterpret_cast<const cx<R> *>(c->tw64);
        a.invm = reinterpret_cast<const R *>(c->invm64);
    } else {
        a.tw = reinterpret_cast<const cx<R> *>(c->tw32);
        a.invm = reinterpret_cast<const R *>(c->invm32);
    }
    a.maxIts = c->p.MaxIterIRK;
    a.Dim = c->p.Dim;
    a.precision = precision;
    a.a00 = c->a[0][0];
    a.a01 = c->a[0][1];
    a.a10 = c->a[1][0];
    a.a11 = c->a[1][1];
    a.b0 = c->b[0];
    a.b1 = c->b[1];
    a.c0 = c->c[0];
    a.c1 = c->c[1];
    a.rec = rec;
    a.rec_base = rec_base;
    const int maxsteps = *std::max_element(P.nsteps.begin(), P.nsteps.end());
    const bool timed = c->profile && s == c->s0 && rec == nullptr;
    if (timed) HIPCHK(c, hipEventRecord(c->tk0, s));
    int launches = 0;
    for (int s0 = 0; s0 < maxsteps; s0 += kStepsPerLaunch) {
        a.step_begin = s0;
        a.step_end = std::min(maxsteps, s0 + kStepsPerLaunch);
        hipLaunchKernelGGL((ccd::k_chain<N, T, R>), dim3(nch), dim3(T), 0, s, a);
        ++launches;
    }
    if (timed) {
        HIPCHK(c, hipEventRecord(c->tk1, s));
        c->prof_launches = launches;
        c->prof_pending = true;
    }
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL((ccd::k_mismatch<N, T, R>), dim3(ncols), dim3(T), 0, s, Y, B.lidx, B.ridx, B.delta,
                       c->p.PrecisionIRK, c->tw64, B.F, B.mstatus);
    HIPCHK(c, hipGetLastError());
    return CC_OK;
}

const ColumnBuffers &main_buffers(cc_ctx *c) { return c->main; }
const ColumnBuffers &base_buffers(cc_ctx *c) { return c->base; }

hipError_t alloc_columns(cc_ctx *c, ColumnBuffers &B, int maxcols, int maxchains)
{
    hipError_t e = hipSuccess;
    cx<double> *y = nullptr;
    if (e == hipSuccess) e = dalloc(c, &y, (size_t)maxchains * c->M);
    B.Y = y;
    if (e == hipSuccess) e = dalloc(c, &B.delta, maxchains);
    if (e == hipSuccess) e = dalloc(c, &B.F, (size_t)maxcols * c->n);
    for (int32_t **q : {&B.start, &B.nsteps, &B.dir, &B.status, &B.fail})
        if (e == hipSuccess) e = dalloc(c, q, maxchains);
    for (int32_t **q : {&B.lcols, &B.rcols, &B.lch, &B.rch, &B.lidx, &B.ridx, &B.mstatus})
        if (e == hipSuccess) e = dalloc(c, q, maxcols);
    if (e == hipSuccess) e = dalloc(c, &B.its, maxchains);
    return e;
}

cc_status columns(cc_ctx *c, const Plan &P, const ColumnBuffers &B, const double *d_x, double eps, bool fp32,
                  hipStream_t s)
{
    cc_status st = CC_OK;
    const double prec = fp32 ? c->p.PrecisionIRK32 : c->p.PrecisionIRK;
    dispatch_n(c->N, [&](auto Nc) {
        constexpr int N = decltype(Nc)::value;
        st = fp32 ? launch_columns<N, float>(c, P, B, d_x, eps, prec, s)
                  : launch_columns<N, double>(c, P, B, d_x, eps, prec, s);
    });
    return st;
}

// Per-column status after the launch (synchronizes s).
cc_status collect_status(cc_ctx *c, const Plan &P, const ColumnBuffers &B, hipStream_t s, int32_t *col_status,
                         bool accumulate_counters)
{
    std::vector<int32_t> st(P.nchains), ms(P.cols.size());
    std::vector<int64_t> its(P.nchains);
    HIPCHK(c, hipMemcpyAsync(st.data(), B.status, sizeof(int32_t) * st.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipMemcpyAsync(ms.data(), B.mstatus, sizeof(int32_t) * ms.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipMemcpyAsync(its.data(), B.its, sizeof(int64_t) * its.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    if (accumulate_counters) {
        int64_t steps = 0, itsum = 0;
        for (int i = 0; i < P.nchains; ++i) {
            steps += P.nsteps[i];
            itsum += its[i];
        }
        c->counters[0] += steps;
        c->counters[1] += itsum;
        c->counters[2] += P.nchains;
        if (s == c->s0) {
            c->prof_its = itsum;
            c->prof_steps = steps;
        }
    }
    cc_status worst = CC_OK;
    for (size_t b = 0; b < P.cols.size(); ++b) {
        int s_ = std::max(st[P.lidx[b]], st[P.ridx[b]]);
        if (s_ == 0) s_ = ms[b];
        if (col_status && P.cols[b] >= 0) col_status[P.cols[b]] = s_;
        if (s_ > (int)worst) worst = (cc_status)s_;
    }
    if (worst == CC_IRK_NOCONV) c->err = "IRK fixed-point iteration did not converge";
    if (worst == CC_DELTA_MISMATCH) c->err = "Mismatch in Delta from left and right side";
    return worst;
}

// LU with partial pivoting on d_A (n x n row-major, precision R), then
// solve for d_b (FP64).  info on device.
template <typename R>
void lu_factor(cc_ctx *c, R *dA, int n, hipStream_t s)
{
    for (int k0 = 0; k0 < n; k0 += ccd::LU_NB) {
        hipLaunchKernelGGL((ccd::k_lu_panel<R>), dim3(1), dim3(ccd::LU_PANEL_T), 0, s, dA, n, k0, c->ipiv, c->info);
        const int nb = std::min(ccd::LU_NB, n - k0);
        const int rest = n - k0 - nb;
        if (rest > 0) {
            hipLaunchKernelGGL((ccd::k_lu_trsm<R>), dim3((rest + 127) / 128), dim3(128), 0, s, dA, n, k0);
            const int tiles = (rest + 63) / 64;
            hipLaunchKernelGGL((ccd::k_lu_gemm<R>), dim3(tiles, tiles), dim3(256), 0, s, dA, n, k0);
        }
    }
}

cc_status solve_device(cc_ctx *c, int n, const double *d_J, const double *d_rhs, double *d_sol, int32_t *info_out,
                       hipStream_t s)
{
    HIPCHK(c, hipMemsetAsync(c->info, 0, sizeof(int32_t), s));
    const bool mixed = c->p.precision == CC_MIXED;
    const unsigned g = (unsigned)((size_t)n * n + 255) / 256;
    if (!mixed) {
        HIPCHK(c, hipMemcpyAsync(c->A, d_J, sizeof(double) * n * n, hipMemcpyDeviceToDevice, s));
        lu_factor<double>(c, c->A, n, s);
        HIPCHK(c, hipMemcpyAsync(d_sol, d_rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL((ccd::k_lu_solve<double, double>), dim3(1), dim3(1024), 0, s, c->A, n, c->ipiv, d_sol);
    } else {
        hipLaunchKernelGGL((ccd::k_cast_to<float>), dim3(g), dim3(256), 0, s, d_J, c->A32, (size_t)n * n);
        lu_factor<float>(c, c->A32, n, s);
        HIPCHK(c, hipMemcpyAsync(d_sol, d_rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL((ccd::k_lu_solve<float, double>), dim3(1), dim3(1024), 0, s, c->A32, n, c->ipiv, d_sol);
        for (int it = 0; it < c->p.refine_iters; ++it) {       // FP64 iterative refinement
            hipLaunchKernelGGL(ccd::k_residual, dim3(n), dim3(256), 0, s, d_J, d_sol, d_rhs, c->r, n);
            hipLaunchKernelGGL((ccd::k_lu_solve<float, double>), dim3(1), dim3(1024), 0, s, c->A32, n, c->ipiv,
                               c->r);
            hipLaunchKernelGGL(ccd::k_axpy, dim3((n + 255) / 256), dim3(256), 0, s, d_sol, c->r, n);
        }
    }
    HIPCHK(c, hipGetLastError());
    int32_t info = 0;
    HIPCHK(c, hipMemcpyAsync(&info, c->info, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    if (info_out) *info_out = info;
    if (info != 0) {
        c->err = "LU factorization found a zero pivot (dgesv info " + std::to_string(info) + ")";
        return CC_SINGULAR;
    }
    return CC_OK;
}

std::vector<int32_t> jacobian_cols(const cc_ctx *c, bool with_base)
{
    std::vector<int32_t> cols;
    if (with_base) cols.push_back(-1);
    for (int i = 0; i < c->n; ++i) cols.push_back(i);
    return cols;
}

// Mixed-policy Jacobian on stream s0.  (1) FP64 launch of the base column
// (recording its trajectory) and of the Delta column; F(x) = row 0 of
// base.F, the Delta column is a forward difference in FP64 (its direction
// changes k0 = 2 pi / Delta inside every spectral operator).  (2) FP32
// tangent-linear columns for every other slot, one chain each: Up slots only
// move the right boundary data, psic/fc slots only the left, so the other
// side's tangent is exactly zero.
template <int N>
cc_status mixed_jacobian_n(cc_ctx *c, const double *d_x, std::vector<int32_t> *tstatus_out, std::vector<int32_t> *tcols_out)
{
    constexpr int T = tpb(N);
    const int n = c->n, dslot = N / 2 + 2;
    hipStream_t s = c->s0;
    // (1) FP64: base + Delta column
    Plan Pb = make_plan(c, std::vector<int32_t>{-1, dslot}, false);
    const std::vector<int64_t> rb = {0, (int64_t)c->iM, -1, -1};      // chains 0,1 = base L,R
    HIPCHK(c, hipMemcpyAsync(c->rec_base_b, rb.data(), sizeof(int64_t) * rb.size(), hipMemcpyHostToDevice, s));
    cc_status st = launch_columns<N, double>(c, Pb, c->base, d_x, c->p.EpsNewton, c->p.PrecisionIRK, s, c->rec,
                                             c->rec_base_b);
    if (st != CC_OK) return st;
    {
        HIPCHK(c, up(Pb.cols, c->cols, s));
        hipLaunchKernelGGL(ccd::k_fd, dim3(1, n), dim3(128), 0, s, c->base.F, c->cols, 2, 0, n, c->p.EpsNewton, c->J);
        HIPCHK(c, hipGetLastError());
    }
    st = collect_status(c, Pb, c->base, s, nullptr, true);
    if (st != CC_OK) return st;
    // (2) FP32 tangent columns
    std::vector<int32_t> cols, lcols, lch, rcols, rch, lidx, ridx, start, nsteps, dir;
    std::vector<int64_t> rbt;
    for (int i = 0; i < n; ++i) {
        if (i == dslot) continue;
        const int id = (int)start.size();
        const bool left = i >= N / 4;
        cols.push_back(i);
        if (left) {
            lcols.push_back(i); lch.push_back(id); lidx.push_back(id); ridx.push_back(-1);
            start.push_back(0); nsteps.push_back(c->iM); dir.push_back(1); rbt.push_back(0);
        } else {
            rcols.push_back(i); rch.push_back(id); lidx.push_back(-1); ridx.push_back(id);
            start.push_back(c->iR); nsteps.push_back(c->iR - c->iM); dir.push_back(-1); rbt.push_back(c->iM);
        }
    }
    const int nch = (int)cols.size();
    HIPCHK(c, up(cols, c->t_cols, s));
    HIPCHK(c, up(lcols, c->t_lcols, s));
    HIPCHK(c, up(lch, c->t_lch, s));
    HIPCHK(c, up(rcols, c->t_rcols, s));
    HIPCHK(c, up(rch, c->t_rch, s));
    HIPCHK(c, up(lidx, c->t_lidx, s));
    HIPCHK(c, up(ridx, c->t_ridx, s));
    HIPCHK(c, up(start, c->t_start, s));
    HIPCHK(c, up(nsteps, c->t_nsteps, s));
    HIPCHK(c, up(dir, c->t_dir, s));
    HIPCHK(c, hipMemcpyAsync(c->rec_base_t, rbt.data(), sizeof(int64_t) * nch, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemsetAsync(c->t_status, 0, sizeof(int32_t) * nch, s));
    HIPCHK(c, hipMemsetAsync(c->t_its, 0, sizeof(int64_t) * nch, s));
    if (!lcols.empty())
        hipLaunchKernelGGL((ccd::k_boundary_tangent<N, T, true>), dim3((unsigned)lcols.size()), dim3(T), 0, s, d_x,
                           c->t_lcols, c->t_lch, c->tangent_h, c->p.XLeft, c->p.Dim, c->tw64, c->dY);
    if (!rcols.empty())
        hipLaunchKernelGGL((ccd::k_boundary_tangent<N, T, false>), dim3((unsigned)rcols.size()), dim3(T), 0, s, d_x,
                           c->t_rcols, c->t_rch, c->tangent_h, c->p.XRight, c->p.Dim, c->tw64, c->dY);
    HIPCHK(c, hipGetLastError());
    ccd::TangentParams a{};
    a.dY = c->dY;
    a.rec = c->rec;
    a.rec_base = c->rec_base_t;
    a.start = c->t_start;
    a.nsteps = c->t_nsteps;
    a.dir = c->t_dir;
    a.delta = c->base.delta;          // [0]: the base's Delta, shared by every tangent chain
    a.grid = c->grid;
    a.status = c->t_status;
    a.fail_idx = c->t_fail;
    a.its = c->t_its;
    a.tw = c->tw32;
    a.invm = c->invm32;
    a.maxIts = c->p.MaxIterIRK;
    a.Dim = c->p.Dim;
    a.rtol = c->tangent_rtol;
    a.a00 = c->a[0][0]; a.a01 = c->a[0][1]; a.a10 = c->a[1][0]; a.a11 = c->a[1][1];
    a.b0 = c->b[0]; a.b1 = c->b[1]; a.c0 = c->c[0]; a.c1 = c->c[1];
    const int maxsteps = std::max(c->iM, c->iR - c->iM);
    if (c->profile) HIPCHK(c, hipEventRecord(c->tk0, s));
    int launches = 0;
    for (int s0 = 0; s0 < maxsteps; s0 += kStepsPerLaunch) {
        a.step_begin = s0;
        a.step_end = std::min(maxsteps, s0 + kStepsPerLaunch);
        hipLaunchKernelGGL((ccd::k_tangent<N, T>), dim3(nch), dim3(T), 0, s, a);
        ++launches;
    }
    if (c->profile) {
        HIPCHK(c, hipEventRecord(c->tk1, s));
        c->prof_launches = launches;
        c->prof_pending = true;
    }
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL((ccd::k_tangent_column<N, T>), dim3(nch), dim3(T), 0, s, c->dY, c->t_cols, c->t_lidx,
                       c->t_ridx, c->tw64, c->J);
    HIPCHK(c, hipGetLastError());
    std::vector<int32_t> tst(nch);
    std::vector<int64_t> tits(nch);
    HIPCHK(c, hipMemcpyAsync(tst.data(), c->t_status, sizeof(int32_t) * nch, hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipMemcpyAsync(tits.data(), c->t_its, sizeof(int64_t) * nch, hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    int64_t itsum = 0, steps = 0;
    cc_status worst = CC_OK;
    for (int b = 0; b < nch; ++b) {
        itsum += tits[b];
        steps += nsteps[b];
        if (tst[b]) worst = CC_IRK_NOCONV;
    }
    c->counters[0] += steps;
    c->counters[1] += itsum;
    c->counters[2] += nch;
    c->prof_its = itsum;
    c->prof_steps = steps;
    if (tstatus_out) *tstatus_out = tst;
    if (tcols_out) *tcols_out = cols;
    if (worst != CC_OK) c->err = "tangent-linear IRK iteration did not converge";
    return worst;
}

cc_status mixed_jacobian(cc_ctx *c, const double *d_x, std::vector<int32_t> *tst, std::vector<int32_t> *tcols)
{
    cc_status st = CC_OK;
    dispatch_n(c->N, [&](auto Nc) { st = mixed_jacobian_n<decltype(Nc)::value>(c, d_x, tst, tcols); });
    return st;
}

// The Newton step on device arrays.  FP64: one launch of n+1 columns.
// Mixed: FP64 base shoot on s1 concurrently with the FP32 columns on s0.
cc_status newton_step_device(cc_ctx *c, const double *d_x, double *d_F0, double *d_dx, bool sync_status)
{
    const int n = c->n;
    const bool mixed = c->p.precision == CC_MIXED;
    const bool reuse = c->p.reuse_sides != 0;
    c->counters[0] = c->counters[1] = c->counters[2] = 0;
    cc_status st;
    if (mixed) {
        st = mixed_jacobian(c, d_x, nullptr, nullptr);
        if (st != CC_OK) return st;
        HIPCHK(c, hipMemcpyAsync(d_F0, c->base.F, sizeof(double) * n, hipMemcpyDeviceToDevice, c->s0));
    } else {
        Plan Pm = make_plan(c, jacobian_cols(c, true), reuse);
        st = columns(c, Pm, main_buffers(c), d_x, c->p.EpsNewton, false, c->s0);
        if (st != CC_OK) return st;
        const int ncols = (int)Pm.cols.size();
        HIPCHK(c, up(Pm.cols, c->cols, c->s0));
        hipLaunchKernelGGL(ccd::k_fd, dim3((ncols + 127) / 128, n), dim3(128), 0, c->s0, c->main.F, c->cols, ncols, 0,
                           n, c->p.EpsNewton, c->J);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipMemcpyAsync(d_F0, c->main.F, sizeof(double) * n, hipMemcpyDeviceToDevice, c->s0));
        if (sync_status) {
            st = collect_status(c, Pm, main_buffers(c), c->s0, nullptr, true);
            if (st != CC_OK) return st;
        }
    }
    hipLaunchKernelGGL(ccd::k_negate_copy, dim3((n + 255) / 256), dim3(256), 0, c->s0, d_F0, c->d, n);
    HIPCHK(c, hipGetLastError());
    return solve_device(c, n, c->J, c->d, d_dx, nullptr, c->s0);
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

cc_status cc_ctx_create(const cc_problem *p, cc_ctx **out)
{
    if (!p || !out) return CC_BADARG;
    *out = nullptr;
    const int N = p->Ntau;
    if (!(N == 128 || N == 256 || N == 512 || N == 1024 || N == 2048)) return CC_BADARG;
    if (p->SchemeIRK != 2 || p->NLeft < 1 || p->NRight < 1 || p->MaxIterIRK < 1) return CC_BADARG;
    if (!(p->XLeft > 0.0 && p->XLeft < p->XMid && p->XMid < p->XRight && p->XRight < 1.0)) return CC_BADARG;
    cc_ctx *c = new cc_ctx();
    c->p = *p;
    if (c->p.PrecisionIRK32 <= 0.0) c->p.PrecisionIRK32 = 2e-6;
    if (c->p.EpsNewton32 <= 0.0) c->p.EpsNewton32 = 1e-3;
    if (c->p.refine_iters <= 0) c->p.refine_iters = 3;
    c->N = N;
    c->M = N / 2;
    c->n = 3 * N / 4;
    c->X = p->NLeft + p->NRight;
    c->iM = p->NLeft;
    c->iR = c->X;
    // Gauss-Legendre s=2 tableau (ODEStepper.cpp:47-54)
    c->a[0][0] = 0.25;
    c->a[0][1] = 0.25 - 0.5 / std::sqrt(3.0);
    c->a[1][0] = 0.25 + 0.5 / std::sqrt(3.0);
    c->a[1][1] = 0.25;
    c->b[0] = c->b[1] = 0.5;
    c->c[0] = 0.5 - 0.5 / std::sqrt(3.0);
    c->c[1] = 0.5 + 0.5 / std::sqrt(3.0);
    auto bad = [&](hipError_t e, const char *w) {
        if (e == hipSuccess) return false;
        fprintf(stderr, "cc_ctx_create: %s: %s\n", w, hipGetErrorString(e));
        for (void *q : c->allocs) (void)hipFree(q);
        delete c;
        return true;
    };
    if (bad(hipStreamCreateWithFlags(&c->s0, hipStreamNonBlocking), "stream")) return CC_DEVICE;
    if (bad(hipStreamCreateWithFlags(&c->s1, hipStreamNonBlocking), "stream")) return CC_DEVICE;
    if (bad(hipEventCreateWithFlags(&c->ev, hipEventDisableTiming), "event")) return CC_DEVICE;
    if (bad(hipEventCreate(&c->tk0), "event")) return CC_DEVICE;
    if (bad(hipEventCreate(&c->tk1), "event")) return CC_DEVICE;
    const int n = c->n, M = c->M;
    c->maxcols = n + 2;
    c->maxchains = 2 * c->maxcols;
    hipError_t e = hipSuccess;
#define ALLOC(ptr, cnt) \
    if (e == hipSuccess) e = dalloc(c, &(ptr), (cnt));
    ALLOC(c->grid, c->X + 1);
    ALLOC(c->tw64, N);
    ALLOC(c->tw32, N);
    ALLOC(c->invm64, N);
    ALLOC(c->invm32, N);
    ALLOC(c->x, n);
    ALLOC(c->F0, n);
    ALLOC(c->dx, n);
    ALLOC(c->r, n);
    ALLOC(c->d, n);
    ALLOC(c->J, (size_t)n * n);
    ALLOC(c->A, (size_t)n * n);
    ALLOC(c->A32, (size_t)n * n);
    ALLOC(c->ipiv, n);
    ALLOC(c->info, 1);
    ALLOC(c->cols, c->maxcols);
    if (e == hipSuccess) e = alloc_columns(c, c->main, c->maxcols, c->maxchains);
    if (e == hipSuccess) e = alloc_columns(c, c->base, 2, 4);
    ALLOC(c->rec, (size_t)c->X * (size_t)(2 * (5 * (N / 2) + 2 * N + 4)));   // Rec<N>::STEP floats per step
    ALLOC(c->rec_base_b, 4);
    ALLOC(c->rec_base_t, n);
    ALLOC(c->dY, (size_t)n * M);
    ALLOC(c->t_start, n);
    ALLOC(c->t_nsteps, n);
    ALLOC(c->t_dir, n);
    ALLOC(c->t_status, n);
    ALLOC(c->t_fail, n);
    ALLOC(c->t_cols, n);
    ALLOC(c->t_lcols, n);
    ALLOC(c->t_rcols, n);
    ALLOC(c->t_lch, n);
    ALLOC(c->t_rch, n);
    ALLOC(c->t_lidx, n);
    ALLOC(c->t_ridx, n);
    ALLOC(c->t_its, n);
#undef ALLOC
    if (bad(e, "hipMalloc")) return CC_DEVICE;
    // x-grid, NewtonSolver.cpp:477-519
    {
        std::vector<double> g(c->X + 1);
        const size_t iL = 0, iM = (size_t)p->NLeft, iR = (size_t)c->X;
        g[iL] = p->XLeft;
        g[iM] = p->XMid;
        g[iR] = p->XRight;
        const double ZL = std::log(p->XLeft) - std::log(1.0 - p->XLeft);
        const double ZM = std::log(p->XMid) - std::log(1.0 - p->XMid);
        const double ZR = std::log(p->XRight) - std::log(1.0 - p->XRight);
        const double dZL = (ZM - ZL) / (double)p->NLeft, dZR = (ZR - ZM) / (double)p->NRight;
        for (size_t j = iL + 1; j < iM; ++j) {
            const double ex = ZL + (double)(j - iL) * dZL;
            g[j] = std::exp(ex) / (1.0 + std::exp(ex));
        }
        for (size_t j = iM + 1; j < iR; ++j) {
            const double ex = ZM + (double)(j - iM) * dZR;
            g[j] = std::exp(ex) / (1.0 + std::exp(ex));
        }
        if (bad(hipMemcpy(c->grid, g.data(), sizeof(double) * g.size(), hipMemcpyHostToDevice), "grid")) return CC_DEVICE;
    }
    // twiddles e^{2 pi i e/N}, rounded once from long double
    {
        std::vector<cx<double>> t64(N);
        std::vector<cx<float>> t32(N);
        for (int k = 0; k < N; ++k) {
            const long double ang = 2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)N;
            const long double cs = cosl(ang), sn = sinl(ang);
            t64[k] = {(double)cs, (double)sn};
            t32[k] = {(float)cs, (float)sn};
        }
        if (bad(hipMemcpy(c->tw64, t64.data(), sizeof(cx<double>) * N, hipMemcpyHostToDevice), "tw")) return CC_DEVICE;
        if (bad(hipMemcpy(c->tw32, t32.data(), sizeof(cx<float>) * N, hipMemcpyHostToDevice), "tw")) return CC_DEVICE;
        std::vector<double> im64(N);
        std::vector<float> im32(N);
        for (int k = 0; k < N; ++k) {
            const int m = k < N / 2 ? k : k - N;
            im64[k] = (k == 0 || k == N / 2) ? 0.0 : 1.0 / (double)m;
            im32[k] = (float)im64[k];
        }
        if (bad(hipMemcpy(c->invm64, im64.data(), sizeof(double) * N, hipMemcpyHostToDevice), "invm")) return CC_DEVICE;
        if (bad(hipMemcpy(c->invm32, im32.data(), sizeof(float) * N, hipMemcpyHostToDevice), "invm")) return CC_DEVICE;
    }
    *out = c;
    return CC_OK;
}

cc_status cc_ctx_destroy(cc_ctx *c)
{
    if (!c) return CC_BADARG;
    (void)hipStreamSynchronize(c->s0);
    (void)hipStreamSynchronize(c->s1);
    for (void *q : c->allocs) (void)hipFree(q);
    (void)hipEventDestroy(c->ev);
    (void)hipEventDestroy(c->tk0);
    (void)hipEventDestroy(c->tk1);
    (void)hipStreamDestroy(c->s0);
    (void)hipStreamDestroy(c->s1);
    delete c;
    return CC_OK;
}

const char *cc_last_error(const cc_ctx *c) { return c ? c->err.c_str() : "null context"; }

cc_status cc_grid(cc_ctx *c, double *g)
{
    if (!c || !g) return CC_BADARG;
    HIPCHK(c, hipMemcpy(g, c->grid, sizeof(double) * (c->X + 1), hipMemcpyDeviceToHost));
    return CC_OK;
}

cc_status cc_counters(cc_ctx *c, int64_t out[3])
{
    if (!c || !out) return CC_BADARG;
    for (int i = 0; i < 3; ++i) out[i] = c->counters[i];
    return CC_OK;
}

cc_status cc_profile(cc_ctx *c, int32_t enable)
{
    if (!c) return CC_BADARG;
    c->profile = enable != 0;
    c->prof_pending = false;
    return CC_OK;
}

cc_status cc_chain_timing(cc_ctx *c, double *ms, int32_t *launches, int64_t *its, int64_t *steps)
{
    if (!c || !ms || !launches || !its || !steps) return CC_BADARG;
    if (!c->prof_pending) {
        c->err = "no profiled chain launch since cc_profile(ctx, 1)";
        return CC_BADARG;
    }
    HIPCHK(c, hipEventSynchronize(c->tk1));
    float f = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&f, c->tk0, c->tk1));
    *ms = f;
    *launches = c->prof_launches;
    *its = c->prof_its;
    *steps = c->prof_steps;
    return CC_OK;
}

cc_status cc_residual(cc_ctx *c, const double *x, double *F)
{
    if (!c || !x || !F) return CC_BADARG;
    const int n = c->n;
    HIPCHK(c, hipMemcpyAsync(c->x, x, sizeof(double) * n, hipMemcpyHostToDevice, c->s0));
    c->counters[0] = c->counters[1] = c->counters[2] = 0;
    Plan P = make_plan(c, std::vector<int32_t>{-1}, false);
    cc_status st = columns(c, P, main_buffers(c), c->x, 0.0, false, c->s0);
    if (st != CC_OK) return st;
    st = collect_status(c, P, main_buffers(c), c->s0, nullptr, true);
    if (st != CC_OK) return st;
    HIPCHK(c, hipMemcpyAsync(F, c->main.F, sizeof(double) * n, hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    return CC_OK;
}

cc_status cc_jacobian(cc_ctx *c, const double *x, const double *F0, double *J, int32_t *col_status)
{
    if (!c || !x || !F0 || !J) return CC_BADARG;
    const int n = c->n;
    const bool mixed = c->p.precision == CC_MIXED;
    const bool reuse = c->p.reuse_sides != 0;
    c->counters[0] = c->counters[1] = c->counters[2] = 0;
    HIPCHK(c, hipMemcpyAsync(c->x, x, sizeof(double) * n, hipMemcpyHostToDevice, c->s0));
    cc_status st;
    if (mixed) {
        std::vector<int32_t> tst, tcols;
        st = mixed_jacobian(c, c->x, &tst, &tcols);
        if (col_status) {
            for (int i = 0; i < n; ++i) col_status[i] = 0;
            for (size_t b = 0; b < tcols.size(); ++b) col_status[tcols[b]] = tst[b] ? CC_IRK_NOCONV : CC_OK;
        }
    } else {
        // the base column is in the launch when its sides are reused; the
        // difference quotient always uses the caller's F0 (NewtonSolver.cpp:684)
        Plan P = make_plan(c, jacobian_cols(c, reuse), reuse);
        st = columns(c, P, main_buffers(c), c->x, c->p.EpsNewton, false, c->s0);
        if (st != CC_OK) return st;
        const int ncols = (int)P.cols.size();
        HIPCHK(c, hipMemcpyAsync(c->main.F + (size_t)ncols * n, F0, sizeof(double) * n, hipMemcpyHostToDevice, c->s0));
        HIPCHK(c, up(P.cols, c->cols, c->s0));
        hipLaunchKernelGGL(ccd::k_fd, dim3((ncols + 127) / 128, n), dim3(128), 0, c->s0, c->main.F, c->cols, ncols,
                           ncols, n, c->p.EpsNewton, c->J);
        HIPCHK(c, hipGetLastError());
        st = collect_status(c, P, main_buffers(c), c->s0, col_status, true);
    }
    HIPCHK(c, hipMemcpyAsync(J, c->J, sizeof(double) * n * n, hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    return st;
}

cc_status cc_solve(cc_ctx *c, int32_t n, double *A, double *b, int32_t *info)
{
    if (!c || !A || !b || n < 1 || n > c->n) return CC_BADARG;
    HIPCHK(c, hipMemcpyAsync(c->J, A, sizeof(double) * n * n, hipMemcpyHostToDevice, c->s0));
    HIPCHK(c, hipMemcpyAsync(c->d, b, sizeof(double) * n, hipMemcpyHostToDevice, c->s0));
    cc_status st = solve_device(c, n, c->J, c->d, c->dx, info, c->s0);
    if (c->p.precision == CC_MIXED) {
        // FP32 factors, returned widened to double
        const unsigned g = (unsigned)(((size_t)n * n + 255) / 256);
        hipLaunchKernelGGL((ccd::k_cast_from<float>), dim3(g), dim3(256), 0, c->s0, c->A32, c->A, (size_t)n * n);
        HIPCHK(c, hipMemcpyAsync(A, c->A, sizeof(double) * n * n, hipMemcpyDeviceToHost, c->s0));
    } else {
        HIPCHK(c, hipMemcpyAsync(A, c->A, sizeof(double) * n * n, hipMemcpyDeviceToHost, c->s0));
    }
    if (st == CC_OK) HIPCHK(c, hipMemcpyAsync(b, c->dx, sizeof(double) * n, hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    return st;
}

cc_status cc_newton_step(cc_ctx *c, const double *x, double *F0, double *dx)
{
    if (!c || !x || !F0 || !dx) return CC_BADARG;
    const int n = c->n;
    HIPCHK(c, hipMemcpyAsync(c->x, x, sizeof(double) * n, hipMemcpyHostToDevice, c->s0));
    cc_status st = newton_step_device(c, c->x, c->F0, c->dx, true);
    if (st != CC_OK) return st;
    HIPCHK(c, hipMemcpyAsync(F0, c->F0, sizeof(double) * n, hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipMemcpyAsync(dx, c->dx, sizeof(double) * n, hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    return CC_OK;
}

cc_status cc_newton_step_device_async(cc_ctx *c, const double *d_x, double *d_F0, double *d_dx)
{
    if (!c || !d_x || !d_F0 || !d_dx) return CC_BADARG;
    return newton_step_device(c, d_x, d_F0, d_dx, true);
}

cc_status cc_synchronize(cc_ctx *c)
{
    if (!c) return CC_BADARG;
    HIPCHK(c, hipStreamSynchronize(c->s1));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    return CC_OK;
}

cc_status cc_fft_batched(cc_ctx *c, const double *in, double *out, int32_t batch, int32_t sign, int32_t fp32)
{
    if (!c || !in || !out || batch < 1 || (sign != 1 && sign != -1)) return CC_BADARG;
    const int N = c->N;
    const size_t bytes = sizeof(double) * 2 * (size_t)N * batch;
    cx<double> *din = nullptr, *dout = nullptr;
    HIPCHK(c, hipMalloc(&din, bytes));
    HIPCHK(c, hipMalloc(&dout, bytes));
    HIPCHK(c, hipMemcpyAsync(din, in, bytes, hipMemcpyHostToDevice, c->s0));
    dispatch_n(N, [&](auto Nc) {
        constexpr int NN = decltype(Nc)::value;
        constexpr int T = tpb(NN);
        if (fp32) {
            if (sign > 0)
                hipLaunchKernelGGL((ccd::k_fft<NN, T, float, 1>), dim3(batch), dim3(T), 0, c->s0, din, dout, c->tw32);
            else
                hipLaunchKernelGGL((ccd::k_fft<NN, T, float, -1>), dim3(batch), dim3(T), 0, c->s0, din, dout, c->tw32);
        } else {
            if (sign > 0)
                hipLaunchKernelGGL((ccd::k_fft<NN, T, double, 1>), dim3(batch), dim3(T), 0, c->s0, din, dout, c->tw64);
            else
                hipLaunchKernelGGL((ccd::k_fft<NN, T, double, -1>), dim3(batch), dim3(T), 0, c->s0, din, dout, c->tw64);
        }
    });
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    HIPCHK(c, hipFree(din));
    HIPCHK(c, hipFree(dout));
    return CC_OK;
}

cc_status cc_rhs_eval(cc_ctx *c, const double *Y, const double *xs, double *dY, int32_t batch, int32_t fp32)
{
    if (!c || !Y || !xs || !dY || batch < 1) return CC_BADARG;
    const int N = c->N, M = c->M;
    // every state is evaluated as both members of a pair (pairs share Delta)
    std::vector<double> Yp((size_t)2 * batch * 2 * M), xp(2 * batch), dp(2 * batch);
    for (int b = 0; b < batch; ++b) {
        for (int s = 0; s < 2; ++s) {
            std::memcpy(&Yp[(size_t)(2 * b + s) * 2 * M], Y + (size_t)b * 2 * M, sizeof(double) * 2 * M);
            xp[2 * b + s] = xs[b];
            dp[2 * b + s] = Y[(size_t)b * 2 * M + 4];        // Delta = Re Y[2]
        }
    }
    const size_t nY = (size_t)2 * batch * M;
    void *dYin = nullptr, *dYout = nullptr;
    double *dx_ = nullptr, *dd = nullptr, *dtmp = nullptr;
    HIPCHK(c, hipMalloc(&dYin, nY * sizeof(cx<double>)));
    HIPCHK(c, hipMalloc(&dYout, nY * sizeof(cx<double>)));
    HIPCHK(c, hipMalloc(&dtmp, nY * sizeof(cx<double>)));
    HIPCHK(c, hipMalloc(&dx_, sizeof(double) * 2 * batch));
    HIPCHK(c, hipMalloc(&dd, sizeof(double) * 2 * batch));
    HIPCHK(c, hipMemcpyAsync(dtmp, Yp.data(), nY * sizeof(cx<double>), hipMemcpyHostToDevice, c->s0));
    HIPCHK(c, hipMemcpyAsync(dx_, xp.data(), sizeof(double) * 2 * batch, hipMemcpyHostToDevice, c->s0));
    HIPCHK(c, hipMemcpyAsync(dd, dp.data(), sizeof(double) * 2 * batch, hipMemcpyHostToDevice, c->s0));
    const unsigned g = (unsigned)((2 * nY + 255) / 256);
    dispatch_n(N, [&](auto Nc) {
        constexpr int NN = decltype(Nc)::value;
        constexpr int T = tpb(NN);
        if (fp32) {
            hipLaunchKernelGGL((ccd::k_cast_to<float>), dim3(g), dim3(256), 0, c->s0, dtmp, (float *)dYin, 2 * nY);
            hipLaunchKernelGGL((ccd::k_rhs_pairs<NN, T, float>), dim3(batch), dim3(T), 0, c->s0,
                               (const cx<float> *)dYin, dx_, dd, (cx<float> *)dYout, c->tw32, c->invm32, c->p.Dim);
            hipLaunchKernelGGL((ccd::k_cast_from<float>), dim3(g), dim3(256), 0, c->s0, (const float *)dYout, dtmp,
                               2 * nY);
        } else {
            hipLaunchKernelGGL((ccd::k_rhs_pairs<NN, T, double>), dim3(batch), dim3(T), 0, c->s0,
                               (const cx<double> *)dtmp, dx_, dd, (cx<double> *)dYout, c->tw64, c->invm64, c->p.Dim);
            (void)hipMemcpyAsync(dtmp, dYout, nY * sizeof(cx<double>), hipMemcpyDeviceToDevice, c->s0);
        }
    });
    HIPCHK(c, hipGetLastError());
    std::vector<double> outp(2 * nY);
    HIPCHK(c, hipMemcpyAsync(outp.data(), dtmp, nY * sizeof(cx<double>), hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    for (int b = 0; b < batch; ++b)
        std::memcpy(dY + (size_t)b * 2 * M, &outp[(size_t)(2 * b) * 2 * M], sizeof(double) * 2 * M);
    for (void *q : {dYin, dYout, (void *)dx_, (void *)dd, (void *)dtmp}) HIPCHK(c, hipFree(q));
    return CC_OK;
}

cc_status cc_integrate(cc_ctx *c, double *Y, const double *grid, int32_t ngrid, const int32_t *start,
                       const int32_t *end, int32_t batch, int32_t fp32, int64_t *its_total, int32_t *status)
{
    if (!c || !Y || !grid || ngrid < 2 || !start || !end || batch < 1) return CC_BADARG;
    const int N = c->N, M = c->M;
    std::vector<int32_t> ns(batch), dr(batch);
    std::vector<double> dl(batch);
    int maxsteps = 0;
    for (int b = 0; b < batch; ++b) {
        if (start[b] < 0 || start[b] >= ngrid || end[b] < 0 || end[b] >= ngrid) return CC_BADARG;
        ns[b] = std::abs(end[b] - start[b]);
        dr[b] = end[b] >= start[b] ? 1 : -1;
        dl[b] = Y[(size_t)b * 2 * M + 4];
        maxsteps = std::max(maxsteps, ns[b]);
    }
    const size_t nY = (size_t)batch * M;
    double *dg = nullptr, *dd = nullptr, *dtmp = nullptr;
    void *dYc = nullptr;
    int32_t *dst = nullptr, *dns = nullptr, *ddr = nullptr, *dstat = nullptr, *dfail = nullptr;
    int64_t *dits = nullptr;
    HIPCHK(c, hipMalloc(&dg, sizeof(double) * ngrid));
    HIPCHK(c, hipMalloc(&dd, sizeof(double) * batch));
    HIPCHK(c, hipMalloc(&dtmp, sizeof(cx<double>) * nY));
    HIPCHK(c, hipMalloc(&dYc, sizeof(cx<double>) * nY));
    HIPCHK(c, hipMalloc(&dst, sizeof(int32_t) * batch));
    HIPCHK(c, hipMalloc(&dns, sizeof(int32_t) * batch));
    HIPCHK(c, hipMalloc(&ddr, sizeof(int32_t) * batch));
    HIPCHK(c, hipMalloc(&dstat, sizeof(int32_t) * batch));
    HIPCHK(c, hipMalloc(&dfail, sizeof(int32_t) * batch));
    HIPCHK(c, hipMalloc(&dits, sizeof(int64_t) * batch));
    HIPCHK(c, hipMemcpyAsync(dg, grid, sizeof(double) * ngrid, hipMemcpyHostToDevice, c->s0));
    HIPCHK(c, hipMemcpyAsync(dd, dl.data(), sizeof(double) * batch, hipMemcpyHostToDevice, c->s0));
    HIPCHK(c, hipMemcpyAsync(dtmp, Y, sizeof(cx<double>) * nY, hipMemcpyHostToDevice, c->s0));
    HIPCHK(c, hipMemcpyAsync(dst, start, sizeof(int32_t) * batch, hipMemcpyHostToDevice, c->s0));
    HIPCHK(c, hipMemcpyAsync(dns, ns.data(), sizeof(int32_t) * batch, hipMemcpyHostToDevice, c->s0));
    HIPCHK(c, hipMemcpyAsync(ddr, dr.data(), sizeof(int32_t) * batch, hipMemcpyHostToDevice, c->s0));
    HIPCHK(c, hipMemsetAsync(dstat, 0, sizeof(int32_t) * batch, c->s0));
    HIPCHK(c, hipMemsetAsync(dits, 0, sizeof(int64_t) * batch, c->s0));
    const unsigned g = (unsigned)((2 * nY + 255) / 256);
    dispatch_n(N, [&](auto Nc) {
        constexpr int NN = decltype(Nc)::value;
        constexpr int T = tpb(NN);
        auto run = [&](auto tag) {
            using R = decltype(tag);
            ccd::ChainParams<R> a{};
            a.Y = static_cast<cx<R> *>(sizeof(R) == 8 ? (void *)dtmp : dYc);
            a.start = dst;
            a.nsteps = dns;
            a.dir = ddr;
            a.delta = dd;
            a.grid = dg;
            a.status = dstat;
            a.fail_idx = dfail;
            a.its = dits;
            a.tw = reinterpret_cast<const cx<R> *>(sizeof(R) == 8 ? (const void *)c->tw64 : (const void *)c->tw32);
            a.invm = reinterpret_cast<const R *>(sizeof(R) == 8 ? (const void *)c->invm64 : (const void *)c->invm32);
            a.maxIts = c->p.MaxIterIRK;
            a.Dim = c->p.Dim;
            a.precision = sizeof(R) == 8 ? c->p.PrecisionIRK : c->p.PrecisionIRK32;
            a.a00 = c->a[0][0];
            a.a01 = c->a[0][1];
            a.a10 = c->a[1][0];
            a.a11 = c->a[1][1];
            a.b0 = c->b[0];
            a.b1 = c->b[1];
            a.c0 = c->c[0];
            a.c1 = c->c[1];
            for (int s0 = 0; s0 < maxsteps; s0 += kStepsPerLaunch) {
                a.step_begin = s0;
                a.step_end = std::min(maxsteps, s0 + kStepsPerLaunch);
                hipLaunchKernelGGL((ccd::k_chain<NN, T, R>), dim3(batch), dim3(T), 0, c->s0, a);
            }
        };
        if (fp32) {
            hipLaunchKernelGGL((ccd::k_cast_to<float>), dim3(g), dim3(256), 0, c->s0, dtmp, (float *)dYc, 2 * nY);
            run(float{});
            hipLaunchKernelGGL((ccd::k_cast_from<float>), dim3(g), dim3(256), 0, c->s0, (const float *)dYc, dtmp,
                               2 * nY);
        } else {
            run(double{});
        }
    });
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(Y, dtmp, sizeof(cx<double>) * nY, hipMemcpyDeviceToHost, c->s0));
    std::vector<int32_t> stv(batch);
    std::vector<int64_t> itv(batch);
    HIPCHK(c, hipMemcpyAsync(stv.data(), dstat, sizeof(int32_t) * batch, hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipMemcpyAsync(itv.data(), dits, sizeof(int64_t) * batch, hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    if (status)
        for (int b = 0; b < batch; ++b) status[b] = stv[b];
    if (its_total)
        for (int b = 0; b < batch; ++b) its_total[b] = itv[b];
    for (void *q : {(void *)dg, (void *)dd, (void *)dtmp, dYc, (void *)dst, (void *)dns, (void *)ddr, (void *)dstat,
                    (void *)dfail, (void *)dits})
        HIPCHK(c, hipFree(q));
    return CC_OK;
}

cc_status cc_expansions(cc_ctx *c, const double *x, double *YL, double *YR)
{
    if (!c || !x || !YL || !YR) return CC_BADARG;
    const int n = c->n, M = c->M;
    HIPCHK(c, hipMemcpyAsync(c->x, x, sizeof(double) * n, hipMemcpyHostToDevice, c->s0));
    const int32_t cols[2] = {-1, -1}, ch[2] = {0, 1};
    HIPCHK(c, hipMemcpyAsync(c->main.lcols, cols, sizeof(cols), hipMemcpyHostToDevice, c->s0));
    HIPCHK(c, hipMemcpyAsync(c->main.lch, ch, sizeof(ch), hipMemcpyHostToDevice, c->s0));
    cx<double> *Y = static_cast<cx<double> *>(c->main.Y);
    dispatch_n(c->N, [&](auto Nc) {
        constexpr int NN = decltype(Nc)::value;
        constexpr int T = tpb(NN);
        hipLaunchKernelGGL((ccd::k_boundary<NN, T, double, true>), dim3(1), dim3(T), 0, c->s0, c->x, c->main.lcols,
                           c->main.lch, 0.0, c->p.XLeft, c->p.Dim, c->tw64, Y, c->main.delta);
        hipLaunchKernelGGL((ccd::k_boundary<NN, T, double, false>), dim3(1), dim3(T), 0, c->s0, c->x, c->main.lcols,
                           c->main.lch + 1, 0.0, c->p.XRight, c->p.Dim, c->tw64, Y, c->main.delta);
    });
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(YL, Y, sizeof(cx<double>) * M, hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipMemcpyAsync(YR, Y + M, sizeof(cx<double>) * M, hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    return CC_OK;
}

}  // extern "C"
