// cc_capi.hip -- host side of the C-ABI declared in include/cc_gpu.h.
//
// A context owns the device buffers for one problem size, two HIP streams
// (default priority), the FFT twiddle tables, the x-grid and the launch plans.
// A "column" is one shoot of the reference (NewtonSolver::shoot,
// NewtonSolver.cpp:403-423) for x perturbed in one slot (or unperturbed, the
// base); a "chain" is one side of a column integrated to the match point.
// Columns run as
//     k_boundary (left/right series) -> k_chain (IRK2) -> k_mismatch (pack)
// and then k_fd / the LU kernels as needed.  The length-N kernels live in the
// per-length translation units behind ccd::Ops<N> (cc_ops.hpp).
//
// Launch plans (which chains exist, their start/direction, which columns read
// which chains) depend only on the problem shape and are uploaded once, at
// context creation.  A Newton step therefore enqueues kernels and
// device-to-device copies only -- no host-to-device copies, no memsets -- and
// reads its status words back through one pinned block at its end, so the
// same enqueue code can be captured into a hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/cc_gpu.h"
#include "cc_dense.hpp"
#include "cc_ops.hpp"

using ccd::cx;
using ccd::Ops;

namespace {

constexpr int kDefaultStepsPerLaunch = 512;
constexpr int kMaxChunks = 2048;

}  // namespace

// Device arrays of one launch of columns, and the device image of its plan.
struct ColumnBuffers {
    void *Y = nullptr;                     // [maxchains][M] complex double
    double *delta = nullptr, *F = nullptr; // per chain Delta; [maxcols + 1][n] residuals
    int32_t *start = nullptr, *nsteps = nullptr, *dir = nullptr, *status = nullptr, *fail = nullptr;
    int32_t *lcols = nullptr, *rcols = nullptr, *lch = nullptr, *rch = nullptr;
    int32_t *lidx = nullptr, *ridx = nullptr, *mstatus = nullptr, *cols = nullptr;
    int64_t *its = nullptr;
};

// The set of chains needed by a list of columns.
struct Plan {
    std::vector<int32_t> cols;          // perturbed slot per column (-1: base)
    std::vector<int32_t> lcols, lch;    // left chains: column slot, chain id
    std::vector<int32_t> rcols, rch;    // right chains
    std::vector<int32_t> lidx, ridx;    // per column: chain ids of its two sides
    std::vector<int32_t> start, nsteps, dir;
    int nchains = 0;
};

// Launch extent: the first nl left boundaries, nr right boundaries, nch
// chains and ncols mismatch columns of a plan.
struct Counts {
    int nl, nr, nch, ncols;
};

// Pinned host image of the status words of one step.
struct StatusBlock {
    int32_t *status, *mstatus;   // main launch: per chain, per column
    int64_t *its;
    int32_t *bstatus, *bmstatus; // base launch (base + Delta column)
    int64_t *bits;
    int32_t *tstatus;            // tangent chains
    int64_t *tits;
    int32_t *info;               // LU
    int32_t *newton;             // device Newton loop: [0] outcome, [1] keep, [2] applied (last update)
    double *refine;              // relative residual of the refined mixed solve
};

struct cc_ctx {
    cc_problem p{};
    int N = 0, M = 0, n = 0, X = 0, iM = 0, iR = 0;
    hipStream_t s0 = nullptr, s1 = nullptr, s2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_fork2 = nullptr, ev_join2 = nullptr;
    std::vector<hipEvent_t> chunk_ev;    // base chunk k done (pipelined mixed step)
    std::string err;
    int64_t counters[3] = {0, 0, 0};
    int stages = 2;                        // Gauss-Legendre stages (SchemeIRK)
    double a[3][3] = {}, b[3] = {}, c[3] = {};
    double *grid = nullptr;
    cx<double> *tw64 = nullptr;
    cx<float> *tw32 = nullptr;
    double *invm64 = nullptr;            // 1/m per spectral index (0 at DC and Nyquist)
    float *invm32 = nullptr;
    double *x = nullptr, *F0 = nullptr, *dx = nullptr, *r = nullptr, *d = nullptr;
    double *J = nullptr;                 // n x n
    double *A = nullptr;                 // n x n LU work (FP64)
    float *A32 = nullptr;                // n x n LU work (FP32)
    int32_t *perm = nullptr, *lu_swap = nullptr, *info = nullptr;   // LU row permutation, panel swap list
    double *xold = nullptr, *nst = nullptr;  // device Newton loop: last improving x; err, errOld, damping
    int32_t *nist = nullptr;                 // device Newton loop: rises, outcome, keep, applied
    double *refres = nullptr;                // relative residual of the last refined mixed solve
    int32_t *rhold = nullptr;                // 1: that solve is to be redone with the FP64 LU
    int32_t *rskip = nullptr;                // 1: the refinement of the current solve has stopped
    double last_refine = 0.0;                // host copies of the last mixed solve's outcome
    int32_t last_resolve = 0;
    ColumnBuffers main;                  // base + every column (FP64 Jacobian)
    ColumnBuffers base;                  // base + Delta column (residual, mixed policy)
    Plan main_plan, base_plan;
    // tangent-linear columns of the mixed policy
    void *rec = nullptr;                 // base-trajectory record, X steps (float, or double in CC_FP64_TANGENT)
    int64_t *rec_base_b = nullptr;       // record slots of the base launch's chains
    int64_t *rec_base_t = nullptr;       // record slots of the tangent chains
    void *dY = nullptr;                  // [n-1][M] tangent states (cx<float>, or cx<double>)
    void *fprev = nullptr;               // [n-1][2][M] tangent predictor state across launches
    int32_t *t_start = nullptr, *t_nsteps = nullptr, *t_dir = nullptr, *t_status = nullptr, *t_fail = nullptr;
    int32_t *t_cols = nullptr, *t_lcols = nullptr, *t_rcols = nullptr, *t_lch = nullptr, *t_rch = nullptr;
    int32_t *t_lidx = nullptr, *t_ridx = nullptr;
    int64_t *t_its = nullptr;
    // launch order of the tangent chains: the first nt - t_nsolo in the main
    // launch, the last t_nsolo (the slowest of the previous Jacobian) alone on
    // their CUs in a launch of their own on s2
    int32_t *t_order = nullptr;
    std::vector<int32_t> t_order_h;
    int t_nsolo = 0, t_nsolo_used = 0;
    std::vector<int32_t> t_cols_h, t_nsteps_h;
    int t_nl = 0, t_nr = 0;
    double tangent_rtol = 1e-6, tangent_h = 1e-5;   // CC_FP64_TANGENT: rtol 1e-13 (1e-10 at Ntau 2048)
    int maxcols = 0, maxchains = 0;
    cc_schedule sched{};
    void *hblock = nullptr;
    StatusBlock hs{};
    std::vector<void *> allocs;
    // kernel timing on the main stream: event pairs around chain launches
    bool profile = false, profile_base = false, capturing = false;
    hipEvent_t tk0 = nullptr, tk1 = nullptr;
    std::vector<hipEvent_t> tev;
    int prof_pairs = 0;
    int32_t prof_launches = 0;
    int64_t prof_its = 0, prof_steps = 0;
    bool prof_pending = false;
    // the same for the FP64 base chains of the tangent policies (their own stream when pipelined)
    std::vector<hipEvent_t> bev;
    int bprof_pairs = 0;
    int32_t bprof_launches = 0;
    int64_t bprof_its = 0, bprof_steps = 0;
    // captured Newton step
    hipGraphExec_t gexec = nullptr;
    const void *gkey[3] = {nullptr, nullptr, nullptr};
};

// Problems of one Ntau inside a batch: their base and tangent chains run in
// shared launches; per-chain D, grid offset and base index tell them apart.
// Member contexts keep their own plans, Jacobian and LU buffers; their chain
// arrays, record and tangent states point into the class arrays below.
struct BatchClass {
    int N = 0;
    std::vector<int> members;
    hipStream_t s0 = nullptr, s1 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    std::vector<hipEvent_t> chunk_ev;
    cc_schedule sched{};
    int nbase = 0, ntan = 0, maxsteps = 0;
    // base launch: 4 chains per member (base L, base R, Delta L, Delta R)
    cx<double> *Y = nullptr;
    double *delta = nullptr, *dim = nullptr, *grid = nullptr;
    int32_t *start = nullptr, *nsteps = nullptr, *dir = nullptr, *status = nullptr, *fail = nullptr, *goff = nullptr;
    int64_t *its = nullptr, *rec_base = nullptr;
    // tangent launch
    void *dY = nullptr, *rec = nullptr, *fprev = nullptr;
    double *t_dim = nullptr;
    int32_t *t_start = nullptr, *t_nsteps = nullptr, *t_dir = nullptr, *t_status = nullptr, *t_fail = nullptr,
            *t_goff = nullptr, *t_bs = nullptr;
    int64_t *t_its = nullptr, *t_rec_base = nullptr;
};

struct cc_batch {
    std::vector<cc_ctx *> ctx;
    std::vector<BatchClass> classes;
    std::vector<void *> allocs;
    std::string err;
    bool use_graph = false;
    hipStream_t s_main = nullptr;        // origin of the batch step: forks to every class
    hipEvent_t ev_start = nullptr;
    std::vector<hipEvent_t> ev_done;
    hipGraphExec_t gexec = nullptr;
    int64_t counters[3] = {0, 0, 0};
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

#define CCCHK(call)                                                    \
    do {                                                               \
        cc_status s_ = (call);                                         \
        if (s_ != CC_OK) return s_;                                    \
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

// Device buffers of one call: release() (or the destructor, on an early
// return) waits for the call's stream and frees every buffer.
struct CallBuffers {
    hipStream_t s;
    std::vector<void *> p;
    template <typename Tp>
    hipError_t get(Tp **out, size_t count)
    {
        void *q = nullptr;
        const hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(Tp));
        if (e == hipSuccess) p.push_back(q);
        *out = static_cast<Tp *>(q);
        return e;
    }
    hipError_t release()
    {
        hipError_t e = p.empty() ? hipSuccess : hipStreamSynchronize(s);
        for (void *q : p) {
            const hipError_t ef = hipFree(q);
            if (e == hipSuccess) e = ef;
        }
        p.clear();
        return e;
    }
    ~CallBuffers() { (void)release(); }
};

template <typename Tp>
cc_status fill(cc_ctx *c, Tp *p, Tp v, size_t count, hipStream_t s)
{
    if (count == 0) return CC_OK;
    hipLaunchKernelGGL((ccd::k_fill<Tp>), dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, p, v, count);
    HIPCHK(c, hipGetLastError());
    return CC_OK;
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

bool valid_threads(int N, int T)
{
    bool ok = false;
    dispatch_n(N, [&](auto Nc) { ok = Ops<decltype(Nc)::value>::valid_threads(T); });
    return ok;
}

int max_chain_steps(const cc_ctx *c) { return std::max(c->iM, c->iR - c->iM); }

// Rec<N>::STEP: values of the base-trajectory record per IRK step
int rec_step_values(int N)
{
    int v = 0;
    dispatch_n(N, [&](auto Nc) { v = ccd::Rec<decltype(Nc)::value>::STEP; });
    return v;
}

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

Counts counts_of(const Plan &P)
{
    return {(int)P.lcols.size(), (int)P.rcols.size(), P.nchains, (int)P.cols.size()};
}

template <typename Tp>
hipError_t up(const std::vector<Tp> &v, Tp *d)
{
    if (v.empty()) return hipSuccess;
    return hipMemcpy(d, v.data(), v.size() * sizeof(Tp), hipMemcpyHostToDevice);
}

hipError_t upload_plan(const Plan &P, const ColumnBuffers &B)
{
    hipError_t e = hipSuccess;
    for (auto pr : {std::make_pair(&P.start, B.start), std::make_pair(&P.nsteps, B.nsteps),
                    std::make_pair(&P.dir, B.dir), std::make_pair(&P.lcols, B.lcols), std::make_pair(&P.rcols, B.rcols),
                    std::make_pair(&P.lch, B.lch), std::make_pair(&P.rch, B.rch), std::make_pair(&P.lidx, B.lidx),
                    std::make_pair(&P.ridx, B.ridx), std::make_pair(&P.cols, B.cols)})
        if (e == hipSuccess) e = up(*pr.first, pr.second);
    return e;
}

// ---------------------------------------------------------------------------
// Kernel timing: one event pair per timed launch group on the main stream.
// ---------------------------------------------------------------------------
cc_status time_mark(cc_ctx *c, hipStream_t s, bool begin)
{
    if (!c->profile || c->capturing) return CC_OK;
    const size_t idx = 2 * (size_t)c->prof_pairs + (begin ? 0 : 1);
    while (c->tev.size() <= idx) {
        hipEvent_t e;
        HIPCHK(c, hipEventCreate(&e));
        c->tev.push_back(e);
    }
    HIPCHK(c, hipEventRecord(c->tev[idx], s));
    if (!begin) ++c->prof_pairs;
    return CC_OK;
}

// Event pair around one base-chain launch of the tangent policies, on its stream s.
cc_status time_mark_base(cc_ctx *c, hipStream_t s, bool begin)
{
    if (!c->profile || !c->profile_base || c->capturing) return CC_OK;
    const size_t idx = 2 * (size_t)c->bprof_pairs + (begin ? 0 : 1);
    while (c->bev.size() <= idx) {
        hipEvent_t e;
        HIPCHK(c, hipEventCreate(&e));
        c->bev.push_back(e);
    }
    HIPCHK(c, hipEventRecord(c->bev[idx], s));
    if (!begin) ++c->bprof_pairs;
    return CC_OK;
}

void time_reset(cc_ctx *c)
{
    c->bprof_pairs = 0;
    c->bprof_launches = 0;
    c->prof_pairs = 0;
    c->prof_launches = 0;
}

cc_status ensure_chunk_events(cc_ctx *c, int nchunks)
{
    while ((int)c->chunk_ev.size() < nchunks) {
        hipEvent_t e;
        HIPCHK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        c->chunk_ev.push_back(e);
    }
    return CC_OK;
}

// ---------------------------------------------------------------------------
// Column launches (FP64)
// ---------------------------------------------------------------------------
template <typename R>
void set_tableau(ccd::ChainParams<R> &a, const cc_ctx *c)
{
    for (int i = 0; i < 3; ++i) {
        a.b[i] = c->b[i];
        a.c[i] = c->c[i];
        for (int k = 0; k < 3; ++k) a.a[i][k] = c->a[i][k];
    }
    a.stages = c->stages;
}

template <int N>
ccd::ChainParams<double> chain_params(const cc_ctx *c, const ColumnBuffers &B)
{
    ccd::ChainParams<double> a{};
    a.Y = static_cast<cx<double> *>(B.Y);
    a.start = B.start;
    a.nsteps = B.nsteps;
    a.dir = B.dir;
    a.delta = B.delta;
    a.grid = c->grid;
    a.status = B.status;
    a.fail_idx = B.fail;
    a.its = B.its;
    a.tw = c->tw64;
    a.invm = c->invm64;
    a.maxIts = c->p.MaxIterIRK;
    a.Dim = c->p.Dim;
    a.precision = c->p.PrecisionIRK;
    set_tableau(a, c);
    a.rec = nullptr;
    a.rec_base = nullptr;
    return a;
}

// The same launch in FP32 (CC_FP32_FD): the chain states of B.Y as float, the
// FP32 twiddles and IRK tolerance PrecisionIRK32; Delta stays FP64 per chain.
template <int N>
ccd::ChainParams<float> chain_params32(const cc_ctx *c, const ColumnBuffers &B)
{
    const ccd::ChainParams<double> d = chain_params<N>(c, B);
    ccd::ChainParams<float> a{};
    a.Y = static_cast<cx<float> *>(B.Y);
    a.start = d.start;
    a.nsteps = d.nsteps;
    a.dir = d.dir;
    a.delta = d.delta;
    a.grid = d.grid;
    a.status = d.status;
    a.fail_idx = d.fail_idx;
    a.its = d.its;
    a.tw = c->tw32;
    a.invm = c->invm32;
    a.maxIts = d.maxIts;
    a.Dim = d.Dim;
    a.precision = c->p.PrecisionIRK32;
    set_tableau(a, c);
    a.rec = nullptr;
    a.rec_base = nullptr;
    return a;
}

// Clears the status words of a launch and writes the boundary states (as
// float chains when fp32).
template <int N>
cc_status launch_prologue(cc_ctx *c, const ColumnBuffers &B, Counts k, const double *d_x, double eps, hipStream_t s,
                          bool fp32 = false)
{
    CCCHK(fill<int32_t>(c, B.status, 0, k.nch, s));
    CCCHK(fill<int32_t>(c, B.fail, -1, k.nch, s));
    CCCHK(fill<int64_t>(c, B.its, 0, k.nch, s));
    CCCHK(fill<int32_t>(c, B.mstatus, 0, k.ncols, s));
    if (k.nl)
        HIPCHK(c, Ops<N>::boundary(s, true, fp32, k.nl, d_x, B.lcols, B.lch, eps, c->p.XLeft, c->p.Dim, c->tw64,
                                   B.Y, B.delta));
    if (k.nr)
        HIPCHK(c, Ops<N>::boundary(s, false, fp32, k.nr, d_x, B.rcols, B.rch, eps, c->p.XRight, c->p.Dim, c->tw64,
                                   B.Y, B.delta));
    return CC_OK;
}

// Boundary -> chains (in chunks of steps_per_launch) -> mismatch, on stream s;
// FP64 chains of T threads, or FP32 chains (of the default thread count).
// timed: the chain launches count in cc_chain_timing.
template <int N>
cc_status launch_columns(cc_ctx *c, const ColumnBuffers &B, Counts k, const double *d_x, double eps, int T,
                         hipStream_t s, bool fp32 = false, bool timed = true)
{
    CCCHK(launch_prologue<N>(c, B, k, d_x, eps, s, fp32));
    const int maxsteps = max_chain_steps(c), S = c->sched.steps_per_launch;
    if (timed) CCCHK(time_mark(c, s, true));
    auto run = [&](auto a) -> cc_status {
        for (int s0 = 0; s0 < maxsteps; s0 += S) {
            a.step_begin = s0;
            a.step_end = std::min(maxsteps, s0 + S);
            HIPCHK(c, Ops<N>::chain(s, fp32 ? Ops<N>::kThreads : T, k.nch, a));
            if (timed) ++c->prof_launches;
        }
        return CC_OK;
    };
    CCCHK(fp32 ? run(chain_params32<N>(c, B)) : run(chain_params<N>(c, B)));
    if (timed) CCCHK(time_mark(c, s, false));
    HIPCHK(c, Ops<N>::mismatch(s, fp32, k.ncols, B.Y, B.lidx, B.ridx, B.delta, c->p.PrecisionIRK, c->tw64, B.F,
                               B.mstatus));
    return CC_OK;
}

// Copies the status words of a column launch into the pinned block (main or base part).
cc_status readback_columns(cc_ctx *c, const ColumnBuffers &B, Counts k, bool is_base, hipStream_t s)
{
    int32_t *st = is_base ? c->hs.bstatus : c->hs.status;
    int32_t *ms = is_base ? c->hs.bmstatus : c->hs.mstatus;
    int64_t *it = is_base ? c->hs.bits : c->hs.its;
    HIPCHK(c, hipMemcpyAsync(st, B.status, sizeof(int32_t) * k.nch, hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipMemcpyAsync(ms, B.mstatus, sizeof(int32_t) * k.ncols, hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipMemcpyAsync(it, B.its, sizeof(int64_t) * k.nch, hipMemcpyDeviceToHost, s));
    return CC_OK;
}

// Per-column statuses from the pinned block (after the stream synchronised).
cc_status eval_columns(cc_ctx *c, const Plan &P, Counts k, bool is_base, int32_t *col_status, bool count)
{
    const int32_t *st = is_base ? c->hs.bstatus : c->hs.status;
    const int32_t *ms = is_base ? c->hs.bmstatus : c->hs.mstatus;
    const int64_t *it = is_base ? c->hs.bits : c->hs.its;
    if (count) {
        int64_t steps = 0, itsum = 0;
        for (int i = 0; i < k.nch; ++i) {
            steps += P.nsteps[i];
            itsum += it[i];
        }
        c->counters[0] += steps;
        c->counters[1] += itsum;
        c->counters[2] += k.nch;
    }
    cc_status worst = CC_OK;
    for (int b = 0; b < k.ncols; ++b) {
        int s_ = std::max(st[P.lidx[b]], st[P.ridx[b]]);
        if (s_ == 0) s_ = ms[b];
        if (col_status && P.cols[b] >= 0) col_status[P.cols[b]] = s_;
        if (s_ > (int)worst) worst = (cc_status)s_;
    }
    if (worst == CC_IRK_NOCONV) c->err = "IRK fixed-point iteration did not converge";
    if (worst == CC_DELTA_MISMATCH) c->err = "Mismatch in Delta from left and right side";
    return worst;
}

// ---------------------------------------------------------------------------
// Dense solve: LU with partial pivoting (FP64, or FP32 + FP64 refinement)
// ---------------------------------------------------------------------------
// The factorization's launches on s: perm (n ints), swap (1 + 2 LU_SWAP_MAX
// ints) and info (1 int, cleared by the caller) are its scratch; n <= LU_MAXN.
template <typename R>
hipError_t lu_factor_on(R *dA, int n, int32_t *perm, int32_t *swap, int32_t *info, hipStream_t s)
{
    constexpr int NB = ccd::lu_nb<R>();
    const unsigned colblocks = (unsigned)((n + 63) / 64);
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int m = n - k0, nb = std::min(NB, m), rest = m - nb;
        if (m <= ccd::LU_PANEL_ROWS)
            hipLaunchKernelGGL((ccd::k_lu_panel<R, true>), dim3(1), dim3(ccd::LU_PANEL_T), 0, s, dA, n, k0, perm,
                               swap, info);
        else
            hipLaunchKernelGGL((ccd::k_lu_panel<R, false>), dim3(1), dim3(ccd::LU_PANEL_T), 0, s, dA, n, k0, perm,
                               swap, info);
        hipLaunchKernelGGL((ccd::k_lu_laswp<R>), dim3(colblocks), dim3(256), 0, s, dA, n, k0, nb, swap);
        if (rest > 0) {
            hipLaunchKernelGGL((ccd::k_lu_trsm<R>), dim3((rest + 255) / 256), dim3(256), 0, s, dA, n, k0);
            const int tiles = (rest + 63) / 64;
            hipLaunchKernelGGL((ccd::k_lu_gemm<R>), dim3(tiles, tiles), dim3(256), 0, s, dA, n, k0, nb);
        }
    }
    return hipGetLastError();
}

template <typename R>
cc_status lu_factor(cc_ctx *c, R *dA, int n, hipStream_t s)
{
    if (n > ccd::LU_MAXN) {
        c->err = "LU order above " + std::to_string(ccd::LU_MAXN);
        return CC_BADARG;
    }
    HIPCHK(c, lu_factor_on<R>(dA, n, c->perm, c->lu_swap, c->info, s));
    return CC_OK;
}

// Relative residual at which a further FP64 refinement sweep cannot help: the
// residual of a refined solve settles at 2e-16 .. 1e-15 at every workload shape
// (two sweeps after the FP32 solve), and a third changes nothing.
constexpr double kRefineFloor = 1e-14;

// Tangent chunks per base launch in the pipelined mixed step.
constexpr int kBaseChunks = 2;

// Policies that solve with the FP32 LU and FP64 refinement (else the FP64 LU).
bool fp32_solve(const cc_ctx *c) { return c->p.precision == CC_MIXED || c->p.precision == CC_FP32_FD; }

// Enqueues d_sol = J^{-1} d_rhs; LAPACK's info lands in hs.info.
cc_status enqueue_solve(cc_ctx *c, int n, const double *d_J, const double *d_rhs, double *d_sol, hipStream_t s)
{
    CCCHK(fill<int32_t>(c, c->info, 0, 1, s));
    const bool mixed = fp32_solve(c);
    const size_t nn = (size_t)n * n;
    const unsigned g = (unsigned)((nn + 255) / 256);
    if (!mixed) {
        HIPCHK(c, hipMemcpyAsync(c->A, d_J, sizeof(double) * nn, hipMemcpyDeviceToDevice, s));
        CCCHK(lu_factor<double>(c, c->A, n, s));
        HIPCHK(c, hipMemcpyAsync(d_sol, d_rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL((ccd::k_lu_solve<double, double>), dim3(1), dim3(1024), 0, s, c->A, n, c->perm, d_sol);
    } else {
        hipLaunchKernelGGL((ccd::k_cast_to<float>), dim3(g), dim3(256), 0, s, d_J, c->A32, nn);
        CCCHK(lu_factor<float>(c, c->A32, n, s));
        HIPCHK(c, hipMemcpyAsync(d_sol, d_rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL((ccd::k_lu_solve<float, double>), dim3(1), dim3(1024), 0, s, c->A32, n, c->perm, d_sol);
        // FP64 iterative refinement, up to refine_iters sweeps: a sweep whose
        // residual is already at the FP64 floor (kRefineFloor relative) ends it
        // on the device (k_refine_stop; the kernels after it return at entry)
        CCCHK(fill<int32_t>(c, c->rskip, 0, 1, s));
        for (int it = 0; it < c->p.refine_iters; ++it) {
            hipLaunchKernelGGL(ccd::k_residual, dim3(n), dim3(256), 0, s, d_J, d_sol, d_rhs, c->r, n,
                               (const int32_t *)c->rskip);
            hipLaunchKernelGGL(ccd::k_refine_stop, dim3(1), dim3(256), 0, s, c->r, d_rhs, n, kRefineFloor, c->rskip);
            hipLaunchKernelGGL((ccd::k_lu_solve<float, double>), dim3(1), dim3(1024), 0, s, c->A32, n, c->perm,
                               c->r, (const int32_t *)c->rskip);
            hipLaunchKernelGGL(ccd::k_axpy, dim3((n + 255) / 256), dim3(256), 0, s, d_sol, c->r, n,
                               (const int32_t *)c->rskip);
        }
        // how far the refinement got: ||rhs - J sol|| / ||rhs||, read by the host
        // after the step and by a device Newton update that must not use a bad sol
        // (after an early stop r already holds the residual of the final sol)
        hipLaunchKernelGGL(ccd::k_residual, dim3(n), dim3(256), 0, s, d_J, d_sol, d_rhs, c->r, n,
                           (const int32_t *)c->rskip);
        hipLaunchKernelGGL(ccd::k_refine_check, dim3(1), dim3(256), 0, s, c->r, d_rhs, c->info, n, c->p.refine_tol,
                           c->refres, c->rhold);
        HIPCHK(c, hipMemcpyAsync(c->hs.refine, c->refres, sizeof(double), hipMemcpyDeviceToHost, s));
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(c->hs.info, c->info, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    return CC_OK;
}

// After the stream of an FP32-LU solve (CC_MIXED, CC_FP32_FD) synchronised: when the FP64-refined
// FP32 solve of J sol = rhs left a relative residual above refine_tol (or a
// non-finite one, or the FP32 factorization met a zero pivot), the system is
// solved again with the FP64 LU, synchronously on s, and *resolved is set.
// The LU info of the solve that stands is in hs.info afterwards.
cc_status refine_or_resolve(cc_ctx *c, int n, const double *d_J, const double *d_rhs, double *d_sol, hipStream_t s,
                            bool *resolved)
{
    *resolved = false;
    c->last_refine = 0.0;
    c->last_resolve = 0;
    if (!fp32_solve(c)) return CC_OK;
    c->last_refine = c->hs.refine[0];
    if (c->hs.refine[0] <= c->p.refine_tol && c->hs.info[0] == 0) return CC_OK;
    CCCHK(fill<int32_t>(c, c->info, 0, 1, s));
    HIPCHK(c, hipMemcpyAsync(c->A, d_J, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToDevice, s));
    CCCHK(lu_factor<double>(c, c->A, n, s));
    HIPCHK(c, hipMemcpyAsync(d_sol, d_rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL((ccd::k_lu_solve<double, double>), dim3(1), dim3(1024), 0, s, c->A, n, c->perm, d_sol);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(c->hs.info, c->info, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    c->last_resolve = 1;
    *resolved = true;
    return CC_OK;
}

cc_status eval_info(cc_ctx *c, int32_t *info_out)
{
    const int32_t info = *c->hs.info;
    if (info_out) *info_out = info;
    if (info != 0) {
        c->err = "LU factorization found a zero pivot (dgesv info " + std::to_string(info) + ")";
        return CC_SINGULAR;
    }
    return CC_OK;
}

// ---------------------------------------------------------------------------
// Mixed policy.  (1) FP64 launch of the base column (recording its
// trajectory) and of the Delta column; F(x) = row 0 of base.F, the Delta
// column is a forward difference in FP64 (its direction changes k0 = 2 pi /
// Delta inside every spectral operator).  (2) FP32 tangent-linear columns for
// every other slot, one chain each: Up slots only move the right boundary
// data, psic/fc slots only the left, so the other side's tangent is exactly
// zero.  Tangent chunk k (steps [kS, (k+1)S) of every chain) needs only base
// chunk k's record: with sched.pipeline the base runs on s1 and each tangent
// chunk waits on the event of its base chunk, so base chunk k+1 overlaps
// tangent chunk k.
// ---------------------------------------------------------------------------
template <int N, typename R>
ccd::TangentParams<R> tangent_params(const cc_ctx *c)
{
    ccd::TangentParams<R> a{};
    a.dY = static_cast<cx<R> *>(c->dY);
    a.fprev = static_cast<cx<R> *>(c->fprev);
    a.rec = static_cast<const R *>(c->rec);
    a.rec_base = c->rec_base_t;
    a.start = c->t_start;
    a.nsteps = c->t_nsteps;
    a.dir = c->t_dir;
    a.base_status = c->base.status;   // chains 0, 1 of the base launch: base left, base right
    a.delta = c->base.delta;          // [0]: the base's Delta, shared by every tangent chain
    a.grid = c->grid;
    a.status = c->t_status;
    a.fail_idx = c->t_fail;
    a.its = c->t_its;
    if constexpr (sizeof(R) == 8) {
        a.tw = c->tw64;
        a.invm = c->invm64;
    } else {
        a.tw = c->tw32;
        a.invm = c->invm32;
    }
    a.maxIts = c->p.MaxIterIRK;
    a.Dim = c->p.Dim;
    a.rtol = c->tangent_rtol;
    a.a00 = c->a[0][0];
    a.a01 = c->a[0][1];
    a.a10 = c->a[1][0];
    a.a11 = c->a[1][1];
    a.b0 = c->b[0];
    a.b1 = c->b[1];
    a.c0 = c->c[0];
    a.c1 = c->c[1];
    return a;
}

template <int N>
cc_status tangent_prologue(cc_ctx *c, const double *d_x, hipStream_t s)
{
    const int nt = (int)c->t_cols_h.size();
    CCCHK(fill<int32_t>(c, c->t_status, 0, nt, s));
    CCCHK(fill<int32_t>(c, c->t_fail, -1, nt, s));
    CCCHK(fill<int64_t>(c, c->t_its, 0, nt, s));
    const bool t64 = c->p.precision == CC_FP64_TANGENT;
    if (c->t_nl)
        HIPCHK(c, Ops<N>::boundary_tangent(s, true, t64, c->t_nl, d_x, c->t_lcols, c->t_lch, c->tangent_h,
                                           c->p.XLeft, c->p.Dim, c->tw64, c->dY));
    if (c->t_nr)
        HIPCHK(c, Ops<N>::boundary_tangent(s, false, t64, c->t_nr, d_x, c->t_rcols, c->t_rch, c->tangent_h,
                                           c->p.XRight, c->p.Dim, c->tw64, c->dY));
    return CC_OK;
}

// LDS a solo tangent launch adds to the kernel's own: more than half a CU's
// 160 KB in all, so no second tangent workgroup fits beside it.
constexpr size_t kSoloLds = 64 * 1024;

template <int N>
hipError_t launch_solo(cc_ctx *c, hipStream_t s, int nsolo, const ccd::TangentParams<float> &a)
{
    return Ops<N>::tangent(s, c->sched.threads_tangent, nsolo, a, kSoloLds);
}
template <int N>
hipError_t launch_solo(cc_ctx *, hipStream_t, int, const ccd::TangentParams<double> &)
{
    return hipErrorInvalidValue;                 // FP64 tangents never split
}

template <int N, typename R>
cc_status enqueue_tangent_n(cc_ctx *c, const double *d_x)
{
    constexpr bool t64 = sizeof(R) == 8;
    const int n = c->n;
    const int nt = (int)c->t_cols_h.size();
    const bool pipe = c->sched.pipeline != 0;
    hipStream_t s0 = c->s0, sb = pipe ? c->s1 : c->s0;
    const int S = c->sched.steps_per_launch, maxsteps = max_chain_steps(c);
    const int nchunks = (maxsteps + S - 1) / S;
    const Counts kb = counts_of(c->base_plan);
    if (pipe) {
        CCCHK(ensure_chunk_events(c, nchunks));
        HIPCHK(c, hipEventRecord(c->ev_fork, s0));
        HIPCHK(c, hipStreamWaitEvent(sb, c->ev_fork, 0));
    }
    CCCHK(tangent_prologue<N>(c, d_x, s0));
    CCCHK(launch_prologue<N>(c, c->base, kb, d_x, c->p.EpsNewton, sb));
    ccd::ChainParams<double> ab = chain_params<N>(c, c->base);
    ab.rec = c->rec;
    ab.rec_base = c->rec_base_b;
    ccd::TangentParams<R> at = tangent_params<N, R>(c);
    // The slowest chains of the previous Jacobian (t_nsolo, FP32 pipelined
    // only) run in a launch of their own on s2, each workgroup alone on its CU
    // (LDS padded past half a CU): sharing a CU with two other workgroups for
    // the first part of every launch slows them, and every launch they are in
    // lasts as long as they do.  The main launch takes the other chains.
    const int nsolo = (pipe && !t64 && c->sched.threads_tangent == ccd::default_threads(N)) ? c->t_nsolo : 0;
    c->t_nsolo_used = nsolo;
    const int nmain = nt - nsolo;
    ccd::TangentParams<R> as = at;
    if (nsolo) {
        at.order = c->t_order;
        as.order = c->t_order + nmain;
        HIPCHK(c, hipEventRecord(c->ev_fork2, s0));
        HIPCHK(c, hipStreamWaitEvent(c->s2, c->ev_fork2, 0));
    }
    // Pipelined, the base runs BM tangent chunks per launch: each base launch
    // on s1 waits for the previous one to end (~16 us between dependent
    // launches), which is on the critical path while the tangent keeps up.
    const int BM = pipe ? kBaseChunks : 1;
    for (int k = 0; k < nchunks; ++k) {
        at.step_begin = k * S;
        at.step_end = std::min(maxsteps, (k + 1) * S);
        if (k % BM == 0) {
            ab.step_begin = k * S;
            ab.step_end = std::min(maxsteps, (k + BM) * S);
            CCCHK(time_mark_base(c, sb, true));
            HIPCHK(c, Ops<N>::chain(sb, c->sched.threads_base, kb.nch, ab, t64));
            CCCHK(time_mark_base(c, sb, false));
            ++c->bprof_launches;
            if (pipe) HIPCHK(c, hipEventRecord(c->chunk_ev[k / BM], sb));
        }
        if (pipe) HIPCHK(c, hipStreamWaitEvent(s0, c->chunk_ev[k / BM], 0));
        CCCHK(time_mark(c, s0, true));
        HIPCHK(c, Ops<N>::tangent(s0, c->sched.threads_tangent, nmain, at));
        CCCHK(time_mark(c, s0, false));
        ++c->prof_launches;
        if (nsolo) {
            as.step_begin = at.step_begin;
            as.step_end = at.step_end;
            HIPCHK(c, hipStreamWaitEvent(c->s2, c->chunk_ev[k / BM], 0));
            HIPCHK(c, launch_solo<N>(c, c->s2, nsolo, as));
        }
    }
    if (nsolo) {
        HIPCHK(c, hipEventRecord(c->ev_join2, c->s2));
        HIPCHK(c, hipStreamWaitEvent(s0, c->ev_join2, 0));
    }
    HIPCHK(c, Ops<N>::mismatch(sb, false, kb.ncols, c->base.Y, c->base.lidx, c->base.ridx, c->base.delta,
                               c->p.PrecisionIRK, c->tw64, c->base.F, c->base.mstatus));
    hipLaunchKernelGGL(ccd::k_fd, dim3(1, n), dim3(128), 0, sb, c->base.F, c->base.cols, kb.ncols, 0, n,
                       c->p.EpsNewton, c->J);
    HIPCHK(c, hipGetLastError());
    if (pipe) {
        HIPCHK(c, hipEventRecord(c->ev_join, sb));
        HIPCHK(c, hipStreamWaitEvent(s0, c->ev_join, 0));
    }
    HIPCHK(c, Ops<N>::tangent_column(s0, t64, nt, c->dY, c->t_cols, c->t_lidx, c->t_ridx, c->tw64, c->J));
    CCCHK(readback_columns(c, c->base, kb, true, s0));
    HIPCHK(c, hipMemcpyAsync(c->hs.tstatus, c->t_status, sizeof(int32_t) * nt, hipMemcpyDeviceToHost, s0));
    HIPCHK(c, hipMemcpyAsync(c->hs.tits, c->t_its, sizeof(int64_t) * nt, hipMemcpyDeviceToHost, s0));
    return CC_OK;
}

cc_status enqueue_tangent(cc_ctx *c, const double *d_x)
{
    cc_status st = CC_OK;
    const bool t64 = c->p.precision == CC_FP64_TANGENT;
    dispatch_n(c->N, [&](auto Nc) {
        constexpr int NN = decltype(Nc)::value;
        st = t64 ? enqueue_tangent_n<NN, double>(c, d_x) : enqueue_tangent_n<NN, float>(c, d_x);
    });
    return st;
}

bool tangent_policy(const cc_ctx *c) { return c->p.precision == CC_MIXED || c->p.precision == CC_FP64_TANGENT; }

// The chains of the next FP32 Jacobian that get a CU each (t_nsolo, at most
// kMaxSolo): those that needed at least twice the median number of
// fixed-point evaluations in this one, the slowest first.  The order array
// puts them last; a change of the set re-captures a step graph (its launch
// sizes are part of the graph).
constexpr int kMaxSolo = 16;

void drop_graph(cc_ctx *c);

cc_status pick_solo(cc_ctx *c)
{
    const int nt = (int)c->t_cols_h.size();
    if (c->p.precision != CC_MIXED || nt < 4 * kMaxSolo) return CC_OK;
    std::vector<int64_t> v(c->hs.tits, c->hs.tits + nt);
    std::vector<int64_t> sorted = v;
    std::nth_element(sorted.begin(), sorted.begin() + nt / 2, sorted.end());
    const int64_t med = sorted[nt / 2];
    std::vector<int32_t> slow;
    for (int b = 0; b < nt; ++b)
        if (med > 0 && v[b] >= 2 * med) slow.push_back(b);
    std::sort(slow.begin(), slow.end(), [&](int32_t x, int32_t y) { return v[x] > v[y] || (v[x] == v[y] && x < y); });
    if ((int)slow.size() > kMaxSolo) slow.resize(kMaxSolo);
    std::sort(slow.begin(), slow.end());
    std::vector<char> is_slow(nt, 0);
    for (int32_t b : slow) is_slow[b] = 1;
    std::vector<int32_t> order;
    order.reserve(nt);
    for (int b = 0; b < nt; ++b)
        if (!is_slow[b]) order.push_back(b);
    for (int32_t b : slow) order.push_back(b);
    if (order == c->t_order_h && (int)slow.size() == c->t_nsolo) return CC_OK;
    c->t_order_h = order;
    c->t_nsolo = (int)slow.size();
    HIPCHK(c, up(c->t_order_h, c->t_order));
    drop_graph(c);
    return CC_OK;
}

// Statuses and counters of a tangent-linear Jacobian (after the stream synchronised).
cc_status eval_tangent(cc_ctx *c, int32_t *col_status)
{
    const cc_status sb = eval_columns(c, c->base_plan, counts_of(c->base_plan), true, nullptr, true);
    const int nt = (int)c->t_cols_h.size();
    int64_t itsum = 0, steps = 0, its_main = 0, steps_main = 0;
    bool bad = false;
    for (int b = 0; b < nt; ++b) {
        itsum += c->hs.tits[b];
        steps += c->t_nsteps_h[b];
        if (c->hs.tstatus[b]) bad = true;
    }
    // the timed launches are the main ones: their chains' counts
    for (int i = 0; i < nt - c->t_nsolo_used; ++i) {
        its_main += c->hs.tits[c->t_order_h[i]];
        steps_main += c->t_nsteps_h[c->t_order_h[i]];
    }
    c->counters[0] += steps;
    c->counters[1] += itsum;
    c->counters[2] += nt;
    c->prof_its = its_main;
    c->prof_steps = steps_main;
    if (!bad) CCCHK(pick_solo(c));
    c->bprof_its = c->bprof_steps = 0;
    const Counts kb = counts_of(c->base_plan);
    for (int i = 0; i < kb.nch; ++i) {
        c->bprof_its += c->hs.bits[i];
        c->bprof_steps += c->base_plan.nsteps[i];
    }
    if (col_status) {
        for (int i = 0; i < c->n; ++i) col_status[i] = sb;
        for (int b = 0; b < nt; ++b)
            col_status[c->t_cols_h[b]] = sb != CC_OK ? sb : (c->hs.tstatus[b] ? CC_IRK_NOCONV : CC_OK);
    }
    if (sb != CC_OK) return sb;
    if (bad) {
        c->err = "tangent-linear IRK iteration did not converge";
        return CC_IRK_NOCONV;
    }
    return CC_OK;
}

// ---------------------------------------------------------------------------
// FP64 Jacobian: one launch of the base and every column.
// ---------------------------------------------------------------------------
cc_status enqueue_fp64_columns(cc_ctx *c, const double *d_x)
{
    cc_status st = CC_OK;
    dispatch_n(c->N, [&](auto Nc) {
        st = launch_columns<decltype(Nc)::value>(c, c->main, counts_of(c->main_plan), d_x, c->p.EpsNewton,
                                                 c->sched.threads_bulk, c->s0);
    });
    return st;
}

// CC_FP32_FD, the FP32 form of the reference's Jacobian: every column,
// the base column included, an FP32 shoot at x + EpsNewton e_j, J = (F32_j -
// F32_base) / EpsNewton (the differences of FP32 shoots, NewtonSolver.cpp:
// 615-702); F0 = the FP64 base shoot (base plan), on s1 beside the FP32 launch
// when the schedule pipelines, else before it on s0.
cc_status enqueue_fp32_fd(cc_ctx *c, const double *d_x, double *d_F0)
{
    const int n = c->n;
    const Counts kb = counts_of(c->base_plan), km = counts_of(c->main_plan);
    const bool fork = c->sched.pipeline != 0;
    hipStream_t sb = fork ? c->s1 : c->s0;
    if (fork) {
        HIPCHK(c, hipEventRecord(c->ev_fork, c->s0));
        HIPCHK(c, hipStreamWaitEvent(sb, c->ev_fork, 0));
    }
    cc_status st = CC_OK;
    dispatch_n(c->N, [&](auto Nc) {
        constexpr int N = decltype(Nc)::value;
        st = launch_columns<N>(c, c->base, kb, d_x, 0.0, c->sched.threads_base, sb, false, false);
        if (st == CC_OK) st = readback_columns(c, c->base, kb, true, sb);
        if (st == CC_OK) st = launch_columns<N>(c, c->main, km, d_x, c->p.EpsNewton, Ops<N>::kThreads, c->s0, true);
    });
    CCCHK(st);
    const int ncols = (int)c->main_plan.cols.size();
    hipLaunchKernelGGL(ccd::k_fd, dim3((ncols + 127) / 128, n), dim3(128), 0, c->s0, c->main.F, c->main.cols, ncols,
                       0, n, c->p.EpsNewton, c->J);
    HIPCHK(c, hipGetLastError());
    CCCHK(readback_columns(c, c->main, km, false, c->s0));
    if (fork) {
        HIPCHK(c, hipEventRecord(c->ev_join, sb));
        HIPCHK(c, hipStreamWaitEvent(c->s0, c->ev_join, 0));
    }
    if (d_F0) HIPCHK(c, hipMemcpyAsync(d_F0, c->base.F, sizeof(double) * n, hipMemcpyDeviceToDevice, c->s0));
    return CC_OK;
}

cc_status eval_fp64(cc_ctx *c, int32_t *col_status)
{
    const Counts k = counts_of(c->main_plan);
    const cc_status st = eval_columns(c, c->main_plan, k, false, col_status, true);
    int64_t itsum = 0, steps = 0;
    for (int i = 0; i < k.nch; ++i) {
        itsum += c->hs.its[i];
        steps += c->main_plan.nsteps[i];
    }
    c->prof_its = itsum;
    c->prof_steps = steps;
    return st;
}

// CC_FP32_FD: the FP64 base shoot's status, then the FP32 columns'.
cc_status eval_fp32_fd(cc_ctx *c, int32_t *col_status)
{
    const cc_status sb = eval_columns(c, c->base_plan, counts_of(c->base_plan), true, nullptr, true);
    const cc_status sm = eval_fp64(c, col_status);
    return sb != CC_OK ? sb : sm;
}

// The whole Newton step, enqueued on s0 (s1 joins in the pipelined mixed
// form): F0 = F(x), J, d = -F0, LU, solve (+ refinement), status readback.
cc_status enqueue_newton(cc_ctx *c, const double *d_x, double *d_F0, double *d_dx)
{
    const int n = c->n;
    if (tangent_policy(c)) {
        CCCHK(enqueue_tangent(c, d_x));
        HIPCHK(c, hipMemcpyAsync(d_F0, c->base.F, sizeof(double) * n, hipMemcpyDeviceToDevice, c->s0));
    } else if (c->p.precision == CC_FP32_FD) {
        CCCHK(enqueue_fp32_fd(c, d_x, d_F0));
    } else {
        CCCHK(enqueue_fp64_columns(c, d_x));
        const int ncols = (int)c->main_plan.cols.size();
        hipLaunchKernelGGL(ccd::k_fd, dim3((ncols + 127) / 128, n), dim3(128), 0, c->s0, c->main.F, c->main.cols,
                           ncols, 0, n, c->p.EpsNewton, c->J);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipMemcpyAsync(d_F0, c->main.F, sizeof(double) * n, hipMemcpyDeviceToDevice, c->s0));
        CCCHK(readback_columns(c, c->main, counts_of(c->main_plan), false, c->s0));
    }
    hipLaunchKernelGGL(ccd::k_negate_copy, dim3((n + 255) / 256), dim3(256), 0, c->s0, d_F0, c->d, n);
    HIPCHK(c, hipGetLastError());
    return enqueue_solve(c, n, c->J, c->d, d_dx, c->s0);
}

void drop_graph(cc_ctx *c)
{
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    c->gkey[0] = c->gkey[1] = c->gkey[2] = nullptr;
}

// post (may be null): more work enqueued on s0 behind the step, before the
// step's one host synchronisation.
// resolved (may be null): set when the mixed solve was redone in FP64 after
// the step's own work (and after post), so that d_dx changed since.
cc_status newton_step_device(cc_ctx *c, const double *d_x, double *d_F0, double *d_dx,
                             const std::function<cc_status()> *post = nullptr, bool *resolved = nullptr)
{
    if (resolved) *resolved = false;
    c->counters[0] = c->counters[1] = c->counters[2] = 0;
    time_reset(c);
    if (c->sched.use_graph) {
        if (!c->gexec || c->gkey[0] != d_x || c->gkey[1] != d_F0 || c->gkey[2] != d_dx) {
            drop_graph(c);
            if (c->sched.pipeline && tangent_policy(c)) {      // per-chunk events of the pipeline
                const int S = c->sched.steps_per_launch;
                CCCHK(ensure_chunk_events(c, (max_chain_steps(c) + S - 1) / S));
            }
            HIPCHK(c, hipStreamBeginCapture(c->s0, hipStreamCaptureModeThreadLocal));
            c->capturing = true;
            const cc_status st = enqueue_newton(c, d_x, d_F0, d_dx);
            c->capturing = false;
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(c->s0, &g);
            if (st != CC_OK) {
                if (g) (void)hipGraphDestroy(g);
                return st;
            }
            HIPCHK(c, e);
            const hipError_t ei = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIPCHK(c, ei);
            c->gkey[0] = d_x;
            c->gkey[1] = d_F0;
            c->gkey[2] = d_dx;
        }
        HIPCHK(c, hipGraphLaunch(c->gexec, c->s0));
    } else {
        CCCHK(enqueue_newton(c, d_x, d_F0, d_dx));
        c->prof_pending = c->profile;
    }
    if (post) CCCHK((*post)());
    HIPCHK(c, hipStreamSynchronize(c->s0));
    const cc_status st = tangent_policy(c)                  ? eval_tangent(c, nullptr)
                         : c->p.precision == CC_FP32_FD ? eval_fp32_fd(c, nullptr)
                                                        : eval_fp64(c, nullptr);
    CCCHK(st);
    bool res = false;
    CCCHK(refine_or_resolve(c, c->n, c->J, c->d, d_dx, c->s0, &res));
    if (resolved) *resolved = res;
    return eval_info(c, nullptr);
}

// Whether the base column (index 0 of both plans) of the last step ran
// cleanly, so that F0 is a valid residual whatever the other columns did.
bool base_residual_ok(const cc_ctx *c)
{
    if (tangent_policy(c) || c->p.precision == CC_FP32_FD) {
        const Plan &P = c->base_plan;
        return c->hs.bstatus[P.lidx[0]] == 0 && c->hs.bstatus[P.ridx[0]] == 0 && c->hs.bmstatus[0] == 0;
    }
    const Plan &P = c->main_plan;
    return c->hs.status[P.lidx[0]] == 0 && c->hs.status[P.ridx[0]] == 0 && c->hs.mstatus[0] == 0;
}

// ---------------------------------------------------------------------------
// Autotuner
// ---------------------------------------------------------------------------
cc_status elapsed_ms(cc_ctx *c, double *ms)
{
    HIPCHK(c, hipEventSynchronize(c->tk1));
    float f = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&f, c->tk0, c->tk1));
    *ms = f;
    return CC_OK;
}

template <int N>
cc_status tune_threads_n(cc_ctx *c, const double *d_x, std::vector<double> &times)
{
    using O = Ops<N>;
    const bool mixed = c->p.precision == CC_MIXED;
    const bool t64 = c->p.precision == CC_FP64_TANGENT;     // FP64 record and tangent: default threads only
    std::vector<int> Ts = {O::kThreads};
    if (O::kAltThreads && !t64 && c->stages == 2) Ts.push_back(O::kAltThreads);
    hipStream_t s = c->s0;
    const int maxsteps = max_chain_steps(c);
    auto pick = [&](int &slot, const std::vector<double> &t) {
        size_t best = 0;
        for (size_t i = 1; i < t.size(); ++i)
            if (t[i] < t[best]) best = i;
        slot = Ts[best];
    };
    // base + Delta chains (residual, mixed base): a slice of steps [8, 8 + Sb)
    {
        const Counts kb = counts_of(c->base_plan);
        const int Sb = std::min(maxsteps - 8, 256);
        std::vector<double> t;
        for (int T : Ts) {
            CCCHK(launch_prologue<N>(c, c->base, kb, d_x, c->p.EpsNewton, s));
            ccd::ChainParams<double> a = chain_params<N>(c, c->base);
            a.step_begin = 0;
            a.step_end = 8;
            HIPCHK(c, O::chain(s, T, kb.nch, a));
            a.step_begin = 8;
            a.step_end = 8 + Sb;
            HIPCHK(c, hipEventRecord(c->tk0, s));
            HIPCHK(c, O::chain(s, T, kb.nch, a));
            HIPCHK(c, hipEventRecord(c->tk1, s));
            double ms;
            CCCHK(elapsed_ms(c, &ms));
            t.push_back(ms);
        }
        pick(c->sched.threads_base, t);
        times.insert(times.end(), t.begin(), t.end());
    }
    if (mixed) {
        // tangent chains: record the base over [0, St), then time [8, St)
        const Counts kb = counts_of(c->base_plan);
        const int St = std::min(maxsteps, 136);
        const int nt = (int)c->t_cols_h.size();
        CCCHK(launch_prologue<N>(c, c->base, kb, d_x, c->p.EpsNewton, s));
        ccd::ChainParams<double> ab = chain_params<N>(c, c->base);
        ab.rec = c->rec;
        ab.rec_base = c->rec_base_b;
        ab.step_begin = 0;
        ab.step_end = St;
        HIPCHK(c, O::chain(s, c->sched.threads_base, kb.nch, ab));
        std::vector<double> t;
        for (int T : Ts) {
            CCCHK(tangent_prologue<N>(c, d_x, s));
            ccd::TangentParams<float> at = tangent_params<N, float>(c);
            at.step_begin = 0;
            at.step_end = 8;
            HIPCHK(c, O::tangent(s, T, nt, at));
            at.step_begin = 8;
            at.step_end = St;
            HIPCHK(c, hipEventRecord(c->tk0, s));
            HIPCHK(c, O::tangent(s, T, nt, at));
            HIPCHK(c, hipEventRecord(c->tk1, s));
            double ms;
            CCCHK(elapsed_ms(c, &ms));
            t.push_back(ms);
        }
        pick(c->sched.threads_tangent, t);
        times.insert(times.end(), t.begin(), t.end());
    } else if (!t64 && c->p.precision != CC_FP32_FD) {      // FP32 chains: default threads only
        // every column of the FP64 Jacobian: a slice of steps [4, 4 + Sk)
        const Counts km = counts_of(c->main_plan);
        const int Sk = std::min(maxsteps - 4, 48);
        std::vector<double> t;
        for (int T : Ts) {
            CCCHK(launch_prologue<N>(c, c->main, km, d_x, c->p.EpsNewton, s));
            ccd::ChainParams<double> a = chain_params<N>(c, c->main);
            a.step_begin = 0;
            a.step_end = 4;
            HIPCHK(c, O::chain(s, T, km.nch, a));
            a.step_begin = 4;
            a.step_end = 4 + Sk;
            HIPCHK(c, hipEventRecord(c->tk0, s));
            HIPCHK(c, O::chain(s, T, km.nch, a));
            HIPCHK(c, hipEventRecord(c->tk1, s));
            double ms;
            CCCHK(elapsed_ms(c, &ms));
            t.push_back(ms);
        }
        pick(c->sched.threads_bulk, t);
        times.insert(times.end(), t.begin(), t.end());
    }
    HIPCHK(c, hipStreamSynchronize(s));
    return CC_OK;
}

cc_status timed_step(cc_ctx *c, double *ms)
{
    HIPCHK(c, hipEventRecord(c->tk0, c->s0));
    CCCHK(newton_step_device(c, c->x, c->F0, c->dx));
    HIPCHK(c, hipEventRecord(c->tk1, c->s0));
    return elapsed_ms(c, ms);
}

cc_status tune_stages(cc_ctx *c, std::vector<double> &times)
{
    const bool mixed = tangent_policy(c);
    const bool prof = c->profile;
    c->profile = false;
    cc_schedule best = c->sched;
    double best_ms = 1e300;
    // pipelined: short chunks let the last tangent chunk follow the base's end
    // closely (at Ntau=1024, 8000 x-points: 417 ms at 64 steps, 400 ms at 16);
    // one stream: long chunks, fewer launches
    // (CC_FP32_FD pipelines the FP64 base shoot beside the whole FP32 launch: long chunks)
    const bool fd32 = c->p.precision == CC_FP32_FD;
    for (int pipe : {1, 0}) {
        if (pipe && !mixed && !fd32) continue;
        const std::vector<int> cand = pipe && mixed ? std::vector<int>{8, 12, 16, 24, 32, 64}
                                                    : std::vector<int>{64, 128, 256, 512, 1024};
        for (int S : cand) {
            if ((max_chain_steps(c) + S - 1) / S > kMaxChunks) continue;
            drop_graph(c);
            c->sched.steps_per_launch = S;
            c->sched.pipeline = pipe;
            c->sched.use_graph = 0;
            double ms = 0.0;
            const cc_status st = timed_step(c, &ms);
            if (st != CC_OK) {
                c->profile = prof;
                return st;
            }
            times.push_back(ms);
            if (ms < best_ms) {
                best_ms = ms;
                best = c->sched;
            }
        }
    }
    // graph replay of the best stage schedule: capture + one replay, then a timed replay
    c->sched = best;
    c->sched.use_graph = 1;
    drop_graph(c);
    double ms = 0.0;
    cc_status st = timed_step(c, &ms);
    if (st == CC_OK) st = timed_step(c, &ms);
    c->profile = prof;
    if (st != CC_OK) return st;
    times.push_back(ms);
    if (ms < best_ms) best = c->sched;
    c->sched = best;
    drop_graph(c);
    return CC_OK;
}

// ---------------------------------------------------------------------------
// Batched independent problems (cc_batch_*)
// ---------------------------------------------------------------------------
bool bfail(cc_batch *b, const char *what, hipError_t e)
{
    if (e == hipSuccess) return false;
    b->err = std::string(what) + ": " + hipGetErrorString(e);
    return true;
}

#define BCHK(b, call)                                                  \
    do {                                                               \
        hipError_t e_ = (call);                                        \
        if (bfail((b), #call, e_)) return CC_DEVICE;                   \
    } while (0)

// Propagates a member's error into the batch.
cc_status bstat(cc_batch *b, cc_ctx *c, cc_status st)
{
    if (st != CC_OK) b->err = c->err;
    return st;
}

template <typename Tp>
hipError_t balloc(cc_batch *b, Tp **p, size_t count)
{
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(Tp));
    if (e == hipSuccess) {
        b->allocs.push_back(q);
        *p = static_cast<Tp *>(q);
    }
    return e;
}

// One Newton step of every member of class C, on C.s0 (C.s1 joins when
// pipelined): per-member boundary data, shared base and tangent launches in
// chunks, per-member mismatch / Jacobian columns / solve.
template <int N, typename R>
cc_status batch_enqueue_class_n(cc_batch *b, BatchClass &C)
{
    constexpr bool t64 = sizeof(R) == 8;
    cc_ctx *c0 = b->ctx[C.members[0]];
    const bool pipe = C.sched.pipeline != 0;
    hipStream_t s0 = C.s0, sb = pipe ? C.s1 : C.s0;
    const int S = C.sched.steps_per_launch;
    const int nchunks = (C.maxsteps + S - 1) / S;
    while ((int)C.chunk_ev.size() < nchunks) {
        hipEvent_t e;
        BCHK(b, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        C.chunk_ev.push_back(e);
    }
    if (pipe) {
        BCHK(b, hipEventRecord(C.ev_fork, s0));
        BCHK(b, hipStreamWaitEvent(sb, C.ev_fork, 0));
    }
    for (int i : C.members) {
        cc_ctx *c = b->ctx[i];
        CCCHK(bstat(b, c, tangent_prologue<N>(c, c->x, s0)));
        CCCHK(bstat(b, c, launch_prologue<N>(c, c->base, counts_of(c->base_plan), c->x, c->p.EpsNewton, sb)));
    }
    ccd::ChainParams<double> ab = chain_params<N>(c0, c0->base);
    ab.Y = C.Y;
    ab.start = C.start;
    ab.nsteps = C.nsteps;
    ab.dir = C.dir;
    ab.delta = C.delta;
    ab.grid = C.grid;
    ab.goff = C.goff;
    ab.dim = C.dim;
    ab.status = C.status;
    ab.fail_idx = C.fail;
    ab.its = C.its;
    ab.rec = C.rec;
    ab.rec_base = C.rec_base;
    ccd::TangentParams<R> at = tangent_params<N, R>(c0);
    at.dY = static_cast<cx<R> *>(C.dY);
    at.fprev = static_cast<cx<R> *>(C.fprev);
    at.rec = static_cast<const R *>(C.rec);
    at.rec_base = C.t_rec_base;
    at.start = C.t_start;
    at.nsteps = C.t_nsteps;
    at.dir = C.t_dir;
    at.base_status = C.status;
    at.bs_idx = C.t_bs;
    at.dim = C.t_dim;
    at.goff = C.t_goff;
    at.delta = C.delta;
    at.grid = C.grid;
    at.status = C.t_status;
    at.fail_idx = C.t_fail;
    at.its = C.t_its;
    for (int k = 0; k < nchunks; ++k) {
        ab.step_begin = at.step_begin = k * S;
        ab.step_end = at.step_end = std::min(C.maxsteps, (k + 1) * S);
        BCHK(b, Ops<N>::chain(sb, C.sched.threads_base, C.nbase, ab, t64));
        if (pipe) {
            BCHK(b, hipEventRecord(C.chunk_ev[k], sb));
            BCHK(b, hipStreamWaitEvent(s0, C.chunk_ev[k], 0));
        }
        BCHK(b, Ops<N>::tangent(s0, C.sched.threads_tangent, C.ntan, at));
    }
    for (int i : C.members) {
        cc_ctx *c = b->ctx[i];
        const Counts kb = counts_of(c->base_plan);
        BCHK(b, Ops<N>::mismatch(sb, false, kb.ncols, c->base.Y, c->base.lidx, c->base.ridx, c->base.delta,
                                 c->p.PrecisionIRK, c->tw64, c->base.F, c->base.mstatus));
        hipLaunchKernelGGL(ccd::k_fd, dim3(1, c->n), dim3(128), 0, sb, c->base.F, c->base.cols, kb.ncols, 0, c->n,
                           c->p.EpsNewton, c->J);
        BCHK(b, hipGetLastError());
    }
    if (pipe) {
        BCHK(b, hipEventRecord(C.ev_join, sb));
        BCHK(b, hipStreamWaitEvent(s0, C.ev_join, 0));
    }
    for (int i : C.members) {
        cc_ctx *c = b->ctx[i];
        const int n = c->n, nt = (int)c->t_cols_h.size();
        BCHK(b, Ops<N>::tangent_column(s0, t64, nt, c->dY, c->t_cols, c->t_lidx, c->t_ridx, c->tw64, c->J));
        CCCHK(bstat(b, c, readback_columns(c, c->base, counts_of(c->base_plan), true, s0)));
        BCHK(b, hipMemcpyAsync(c->hs.tstatus, c->t_status, sizeof(int32_t) * nt, hipMemcpyDeviceToHost, s0));
        BCHK(b, hipMemcpyAsync(c->hs.tits, c->t_its, sizeof(int64_t) * nt, hipMemcpyDeviceToHost, s0));
        BCHK(b, hipMemcpyAsync(c->F0, c->base.F, sizeof(double) * n, hipMemcpyDeviceToDevice, s0));
        hipLaunchKernelGGL(ccd::k_negate_copy, dim3((n + 255) / 256), dim3(256), 0, s0, c->F0, c->d, n);
        BCHK(b, hipGetLastError());
        CCCHK(bstat(b, c, enqueue_solve(c, n, c->J, c->d, c->dx, s0)));
    }
    return CC_OK;
}

cc_status batch_enqueue_class(cc_batch *b, BatchClass &C)
{
    cc_status st = CC_OK;
    const bool t64 = b->ctx[C.members[0]]->p.precision == CC_FP64_TANGENT;
    dispatch_n(C.N, [&](auto Nc) {
        constexpr int NN = decltype(Nc)::value;
        st = t64 ? batch_enqueue_class_n<NN, double>(b, C) : batch_enqueue_class_n<NN, float>(b, C);
    });
    return st;
}

// Every class forked from s_main and joined back: the whole batch step.
cc_status batch_enqueue(cc_batch *b)
{
    BCHK(b, hipEventRecord(b->ev_start, b->s_main));
    for (size_t k = 0; k < b->classes.size(); ++k) {
        BatchClass &C = b->classes[k];
        BCHK(b, hipStreamWaitEvent(C.s0, b->ev_start, 0));
        CCCHK(batch_enqueue_class(b, C));
        BCHK(b, hipEventRecord(b->ev_done[k], C.s0));
        BCHK(b, hipStreamWaitEvent(b->s_main, b->ev_done[k], 0));
    }
    return CC_OK;
}

void batch_drop_graph(cc_batch *b)
{
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
}

// Runs one batch step on the members' resident x (eager or graph) and
// evaluates every member's status.
cc_status batch_run(cc_batch *b, int32_t *status)
{
    if (b->use_graph) {
        if (!b->gexec) {
            BCHK(b, hipStreamBeginCapture(b->s_main, hipStreamCaptureModeThreadLocal));
            const cc_status st = batch_enqueue(b);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(b->s_main, &g);
            if (st != CC_OK) {
                if (g) (void)hipGraphDestroy(g);
                return st;
            }
            BCHK(b, e);
            const hipError_t ei = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            BCHK(b, ei);
        }
        BCHK(b, hipGraphLaunch(b->gexec, b->s_main));
    } else {
        CCCHK(batch_enqueue(b));
    }
    BCHK(b, hipStreamSynchronize(b->s_main));
    b->counters[0] = b->counters[1] = b->counters[2] = 0;
    cc_status worst = CC_OK;
    for (size_t i = 0; i < b->ctx.size(); ++i) {
        cc_ctx *c = b->ctx[i];
        c->counters[0] = c->counters[1] = c->counters[2] = 0;
        cc_status st = eval_tangent(c, nullptr);
        bool res = false;
        if (st == CC_OK) st = refine_or_resolve(c, c->n, c->J, c->d, c->dx, c->s0, &res);
        if (st == CC_OK) st = eval_info(c, nullptr);
        for (int k = 0; k < 3; ++k) b->counters[k] += c->counters[k];
        if (status) status[i] = st;
        if (st != CC_OK && worst == CC_OK) b->err = "problem " + std::to_string(i) + ": " + c->err;
        if (st > worst) worst = st;
    }
    return worst;
}

cc_status batch_timed(cc_batch *b, double *ms)
{
    cc_ctx *c0 = b->ctx[0];
    BCHK(b, hipEventRecord(c0->tk0, b->s_main));
    const cc_status st = batch_run(b, nullptr);
    BCHK(b, hipEventRecord(c0->tk1, b->s_main));
    BCHK(b, hipEventSynchronize(c0->tk1));
    float f = 0.0f;
    BCHK(b, hipEventElapsedTime(&f, c0->tk0, c0->tk1));
    *ms = f;
    return st;
}

}  // namespace

// ---------------------------------------------------------------------------
// The FP64 dense solve without allocation (cc_dgesv_noalloc).  Its scratch --
// the row permutation, the panel swap list and the info word of the
// factorization -- is the file-scope device array `scratch`: one copy per GPU,
// held until the process ends, shared by every call on that GPU (overlapping
// calls on one GPU race).  In host code `scratch` names a host-side stand-in,
// not device memory: its device address is looked up once per device, with
// that device current, by hipGetSymbolAddress, and only that address is passed
// to kernels.
// ---------------------------------------------------------------------------
namespace {
constexpr int kScratchPerm = 0;                                       // LU_MAXN ints
constexpr int kScratchSwap = ccd::LU_MAXN;                            // 1 + 2 LU_SWAP_MAX ints
constexpr int kScratchInfo = kScratchSwap + 1 + 2 * ccd::LU_SWAP_MAX;  // 1 int
constexpr int kScratchInts = (kScratchInfo + 1 + 63) / 64 * 64;       // 2240 ints = 8960 bytes
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int32_t *g_scratch[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) int32_t scratch[kScratchInts];

// Device address of the current device's copy of `scratch`.
static hipError_t scratch_address(int32_t **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = static_cast<int32_t *>(p);
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

// own_tangent = false: no base record and tangent states of its own (a batch
// member points them into its class's arrays).
static cc_status create_ctx(const cc_problem *p, cc_ctx **out, bool own_tangent)
{
    if (!p || !out) return CC_BADARG;
    *out = nullptr;
    const int N = p->Ntau;
    if (!(N == 128 || N == 256 || N == 512 || N == 1024 || N == 2048)) return CC_BADARG;
    if (p->SchemeIRK < 1 || p->SchemeIRK > 3 || p->NLeft < 1 || p->NRight < 1 || p->MaxIterIRK < 1) return CC_BADARG;
    // the tangent-linear columns and the batch are built for the two-stage scheme
    if (p->SchemeIRK != 2 && p->precision != CC_FP64) return CC_BADARG;
    if (p->precision != CC_FP64 && p->precision != CC_MIXED && p->precision != CC_FP64_TANGENT &&
        p->precision != CC_FP32_FD)
        return CC_BADARG;
    if (!(p->XLeft > 0.0 && p->XLeft < p->XMid && p->XMid < p->XRight && p->XRight < 1.0)) return CC_BADARG;
    cc_ctx *c = new cc_ctx();
    c->p = *p;
    if (c->p.PrecisionIRK32 <= 0.0) c->p.PrecisionIRK32 = 2e-6;
    if (c->p.refine_tol <= 0.0) c->p.refine_tol = 1e-10;
    if (c->p.refine_iters <= 0) c->p.refine_iters = 3;
    if (p->precision == CC_FP64_TANGENT) c->tangent_rtol = N >= 2048 ? 1e-10 : 1e-13;
    c->N = N;
    c->M = N / 2;
    c->n = 3 * N / 4;
    c->X = p->NLeft + p->NRight;
    c->iM = p->NLeft;
    c->iR = c->X;
    // Gauss-Legendre tableaus (ODEStepper.cpp:29-78)
    c->stages = p->SchemeIRK;
    if (c->stages == 1) {
        c->a[0][0] = 0.5;
        c->b[0] = 1.0;
        c->c[0] = 0.5;
    } else if (c->stages == 2) {
        c->a[0][0] = 0.25;
        c->a[0][1] = 0.25 - 0.5 / std::sqrt(3.0);
        c->a[1][0] = 0.25 + 0.5 / std::sqrt(3.0);
        c->a[1][1] = 0.25;
        c->b[0] = c->b[1] = 0.5;
        c->c[0] = 0.5 - 0.5 / std::sqrt(3.0);
        c->c[1] = 0.5 + 0.5 / std::sqrt(3.0);
    } else {
        const double r15 = std::sqrt(15.0);
        c->a[0][0] = 5.0 / 36.0;
        c->a[0][1] = 2.0 / 9.0 - 1.0 / r15;
        c->a[0][2] = 5.0 / 36.0 - 0.5 / r15;
        c->a[1][0] = 5.0 / 36.0 + r15 / 24.0;
        c->a[1][1] = 2.0 / 9.0;
        c->a[1][2] = 5.0 / 36.0 - r15 / 24.0;
        c->a[2][0] = 5.0 / 36.0 + 0.5 / r15;
        c->a[2][1] = 2.0 / 9.0 + 1.0 / r15;
        c->a[2][2] = 5.0 / 36.0;
        c->b[0] = c->b[2] = 5.0 / 18.0;
        c->b[1] = 4.0 / 9.0;
        c->c[0] = 0.5 - r15 / 10.0;
        c->c[1] = 0.5;
        c->c[2] = 0.5 + r15 / 10.0;
    }
    // default schedule
    c->sched.threads_base = c->sched.threads_bulk = c->sched.threads_tangent = ccd::default_threads(N);
    c->sched.steps_per_launch = kDefaultStepsPerLaunch;
    c->sched.pipeline = 1;
    c->sched.use_graph = 0;
    auto bad = [&](hipError_t e, const char *w) {
        if (e == hipSuccess) return false;
        fprintf(stderr, "cc_ctx_create: %s: %s\n", w, hipGetErrorString(e));
        for (void *q : c->allocs) (void)hipFree(q);
        if (c->hblock) (void)hipHostFree(c->hblock);
        for (hipEvent_t ev : {c->ev_fork, c->ev_join, c->ev_fork2, c->ev_join2, c->tk0, c->tk1})
            if (ev) (void)hipEventDestroy(ev);
        for (hipStream_t st : {c->s0, c->s1, c->s2})
            if (st) (void)hipStreamDestroy(st);
        delete c;
        return true;
    };
    if (bad(hipStreamCreateWithFlags(&c->s0, hipStreamNonBlocking), "stream")) return CC_DEVICE;
    if (bad(hipStreamCreateWithFlags(&c->s1, hipStreamNonBlocking), "stream")) return CC_DEVICE;
    if (bad(hipStreamCreateWithFlags(&c->s2, hipStreamNonBlocking), "stream")) return CC_DEVICE;
    if (bad(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming), "event")) return CC_DEVICE;
    if (bad(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming), "event")) return CC_DEVICE;
    if (bad(hipEventCreateWithFlags(&c->ev_fork2, hipEventDisableTiming), "event")) return CC_DEVICE;
    if (bad(hipEventCreateWithFlags(&c->ev_join2, hipEventDisableTiming), "event")) return CC_DEVICE;
    if (bad(hipEventCreate(&c->tk0), "event")) return CC_DEVICE;
    if (bad(hipEventCreate(&c->tk1), "event")) return CC_DEVICE;
    const int n = c->n, M = c->M;
    c->maxcols = n + 2;
    c->maxchains = 2 * c->maxcols;
    const int dslot = N / 2 + 2;
    // plans: main = base + every column (sides reused if asked), base = base + Delta column
    std::vector<int32_t> jcols = {-1};
    for (int i = 0; i < n; ++i) jcols.push_back(i);
    c->main_plan = make_plan(c, jcols, p->reuse_sides != 0);
    c->base_plan = make_plan(c, std::vector<int32_t>{-1, dslot}, false);
    // tangent plan: one chain per slot except Delta, on the side its boundary data moves
    std::vector<int32_t> t_lcols, t_lch, t_rcols, t_rch, t_lidx, t_ridx, t_start, t_dir;
    std::vector<int64_t> rbt;
    for (int i = 0; i < n; ++i) {
        if (i == dslot) continue;
        const int id = (int)c->t_cols_h.size();
        const bool left = i >= N / 4;
        c->t_cols_h.push_back(i);
        if (left) {
            t_lcols.push_back(i); t_lch.push_back(id); t_lidx.push_back(id); t_ridx.push_back(-1);
            t_start.push_back(0); c->t_nsteps_h.push_back(c->iM); t_dir.push_back(1); rbt.push_back(0);
        } else {
            t_rcols.push_back(i); t_rch.push_back(id); t_lidx.push_back(-1); t_ridx.push_back(id);
            t_start.push_back(c->iR); c->t_nsteps_h.push_back(c->iR - c->iM); t_dir.push_back(-1);
            rbt.push_back(c->iM);
        }
    }
    c->t_nl = (int)t_lcols.size();
    c->t_nr = (int)t_rcols.size();
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
    ALLOC(c->perm, n);
    ALLOC(c->xold, n);
    ALLOC(c->nst, 3);
    ALLOC(c->nist, 4);
    ALLOC(c->refres, 1);
    ALLOC(c->rhold, 1);
    ALLOC(c->rskip, 1);
    ALLOC(c->lu_swap, 1 + 2 * ccd::LU_SWAP_MAX);
    ALLOC(c->info, 1);
    for (ColumnBuffers *B : {&c->main, &c->base}) {
        const int mc = B == &c->main ? c->maxcols : 2, mch = B == &c->main ? c->maxchains : 4;
        cx<double> *y = nullptr;
        ALLOC(y, (size_t)mch * M);
        B->Y = y;
        ALLOC(B->delta, mch);
        ALLOC(B->F, (size_t)(mc + 1) * n);
        for (int32_t **q : {&B->start, &B->nsteps, &B->dir, &B->status, &B->fail})
            ALLOC(*q, mch);
        for (int32_t **q : {&B->lcols, &B->rcols, &B->lch, &B->rch, &B->lidx, &B->ridx, &B->mstatus, &B->cols})
            ALLOC(*q, mc);
        ALLOC(B->its, mch);
    }
    if (own_tangent) {
        // record and tangent states in the tangent's precision (Rec<N>::STEP values per step)
        const size_t elem = p->precision == CC_FP64_TANGENT ? sizeof(double) : sizeof(float);
        char *recp = nullptr, *dyp = nullptr, *fpp = nullptr;
        ALLOC(recp, (size_t)c->X * (size_t)rec_step_values(N) * elem);
        ALLOC(dyp, (size_t)n * M * 2 * elem);
        ALLOC(fpp, (size_t)n * 2 * M * 2 * elem);
        c->rec = recp;
        c->dY = dyp;
        c->fprev = fpp;
    }
    ALLOC(c->rec_base_b, 4);
    ALLOC(c->rec_base_t, n);
    for (int32_t **q : {&c->t_start, &c->t_nsteps, &c->t_dir, &c->t_status, &c->t_fail, &c->t_cols, &c->t_lcols,
                        &c->t_rcols, &c->t_lch, &c->t_rch, &c->t_lidx, &c->t_ridx})
        ALLOC(*q, n);
    ALLOC(c->t_its, n);
    ALLOC(c->t_order, n);
#undef ALLOC
    if (bad(e, "hipMalloc")) return CC_DEVICE;
    // pinned status block
    {
        const size_t mch = c->maxchains, mc = c->maxcols;
        const size_t bytes = sizeof(double) * 2 + sizeof(int64_t) * (mch + 4 + (size_t)n) +
                             sizeof(int32_t) * (mch + mc + 4 + 2 + n + 2 + 4);
        if (bad(hipHostMalloc(&c->hblock, bytes, hipHostMallocDefault), "hipHostMalloc")) return CC_DEVICE;
        char *q = static_cast<char *>(c->hblock);
        auto take = [&](auto **ptr, size_t cnt) {
            *ptr = reinterpret_cast<std::remove_pointer_t<decltype(ptr)>>(q);
            q += cnt * sizeof(**ptr);
        };
        take(&c->hs.refine, 2);
        take(&c->hs.its, mch);
        take(&c->hs.bits, 4);
        take(&c->hs.tits, n);
        take(&c->hs.status, mch);
        take(&c->hs.mstatus, mc);
        take(&c->hs.bstatus, 4);
        take(&c->hs.bmstatus, 2);
        take(&c->hs.tstatus, n);
        take(&c->hs.info, 2);
        take(&c->hs.newton, 4);
    }
    // plans to the device
    if (bad(upload_plan(c->main_plan, c->main), "plan")) return CC_DEVICE;
    if (bad(upload_plan(c->base_plan, c->base), "plan")) return CC_DEVICE;
    {
        const std::vector<int64_t> rb = {0, (int64_t)c->iM, -1, -1};      // base chains 0, 1 record
        hipError_t eu = up(rb, c->rec_base_b);
        for (auto pr : {std::make_pair(&c->t_cols_h, c->t_cols), std::make_pair(&t_lcols, c->t_lcols),
                        std::make_pair(&t_lch, c->t_lch), std::make_pair(&t_rcols, c->t_rcols),
                        std::make_pair(&t_rch, c->t_rch), std::make_pair(&t_lidx, c->t_lidx),
                        std::make_pair(&t_ridx, c->t_ridx), std::make_pair(&t_start, c->t_start),
                        std::make_pair(&c->t_nsteps_h, c->t_nsteps), std::make_pair(&t_dir, c->t_dir)})
            if (eu == hipSuccess) eu = up(*pr.first, pr.second);
        if (eu == hipSuccess) eu = up(rbt, c->rec_base_t);
        c->t_order_h.resize(c->t_cols_h.size());
        for (size_t b = 0; b < c->t_order_h.size(); ++b) c->t_order_h[b] = (int32_t)b;
        if (eu == hipSuccess && !c->t_order_h.empty()) eu = up(c->t_order_h, c->t_order);
        if (bad(eu, "plan")) return CC_DEVICE;
    }
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

cc_status cc_ctx_create(const cc_problem *p, cc_ctx **out) { return create_ctx(p, out, true); }

cc_status cc_ctx_destroy(cc_ctx *c)
{
    if (!c) return CC_BADARG;
    (void)hipStreamSynchronize(c->s0);
    (void)hipStreamSynchronize(c->s1);
    (void)hipStreamSynchronize(c->s2);
    drop_graph(c);
    for (void *q : c->allocs) (void)hipFree(q);
    if (c->hblock) (void)hipHostFree(c->hblock);
    for (hipEvent_t e : c->chunk_ev) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->tev) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->bev) (void)hipEventDestroy(e);
    for (hipEvent_t e : {c->ev_fork, c->ev_join, c->ev_fork2, c->ev_join2, c->tk0, c->tk1}) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->s0);
    (void)hipStreamDestroy(c->s1);
    (void)hipStreamDestroy(c->s2);
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

cc_status cc_last_jacobian(cc_ctx *c, double *J)
{
    if (!c || !J) return CC_BADARG;
    HIPCHK(c, hipStreamSynchronize(c->s0));
    HIPCHK(c, hipMemcpy(J, c->J, sizeof(double) * c->n * c->n, hipMemcpyDeviceToHost));
    return CC_OK;
}

cc_status cc_counters(cc_ctx *c, int64_t out[3])
{
    if (!c || !out) return CC_BADARG;
    for (int i = 0; i < 3; ++i) out[i] = c->counters[i];
    return CC_OK;
}

cc_status cc_refinement(cc_ctx *c, double *residual, int32_t *fp64_resolve)
{
    if (!c || !residual || !fp64_resolve) return CC_BADARG;
    *residual = c->last_refine;
    *fp64_resolve = c->last_resolve;
    return CC_OK;
}

cc_status cc_profile(cc_ctx *c, int32_t enable)
{
    if (!c || enable < 0 || enable > 2) return CC_BADARG;
    c->profile = enable != 0;
    c->profile_base = enable == 1;
    c->prof_pending = false;
    return CC_OK;
}

cc_status cc_chain_timing(cc_ctx *c, double *ms, int32_t *launches, int64_t *its, int64_t *steps)
{
    if (!c || !ms || !launches || !its || !steps) return CC_BADARG;
    if (!c->prof_pending) {
        c->err = "no profiled eager chain launch since cc_profile(ctx, 1)";
        return CC_BADARG;
    }
    double tot = 0.0;
    for (int i = 0; i < c->prof_pairs; ++i) {
        HIPCHK(c, hipEventSynchronize(c->tev[2 * i + 1]));
        float f = 0.0f;
        HIPCHK(c, hipEventElapsedTime(&f, c->tev[2 * i], c->tev[2 * i + 1]));
        tot += f;
    }
    *ms = tot;
    *launches = c->prof_launches;
    *its = c->prof_its;
    *steps = c->prof_steps;
    return CC_OK;
}

cc_status cc_base_timing(cc_ctx *c, double *ms, int32_t *launches, int64_t *its, int64_t *steps)
{
    if (!c || !ms || !launches || !its || !steps) return CC_BADARG;
    if (!c->prof_pending || !tangent_policy(c) || !c->profile_base) {
        c->err = "no profiled eager tangent-policy step since cc_profile(ctx, 1)";
        return CC_BADARG;
    }
    double tot = 0.0;
    for (int i = 0; i < c->bprof_pairs; ++i) {
        HIPCHK(c, hipEventSynchronize(c->bev[2 * i + 1]));
        float f = 0.0f;
        HIPCHK(c, hipEventElapsedTime(&f, c->bev[2 * i], c->bev[2 * i + 1]));
        tot += f;
    }
    *ms = tot;
    *launches = c->bprof_launches;
    *its = c->bprof_its;
    *steps = c->bprof_steps;
    return CC_OK;
}

cc_status cc_get_schedule(cc_ctx *c, cc_schedule *out)
{
    if (!c || !out) return CC_BADARG;
    *out = c->sched;
    return CC_OK;
}

cc_status cc_set_schedule(cc_ctx *c, const cc_schedule *s)
{
    if (!c || !s) return CC_BADARG;
    if (!valid_threads(c->N, s->threads_base) || !valid_threads(c->N, s->threads_bulk) ||
        !valid_threads(c->N, s->threads_tangent) ||
        (c->p.precision == CC_FP64_TANGENT && (s->threads_base != ccd::default_threads(c->N) ||
                                               s->threads_tangent != ccd::default_threads(c->N))) ||
        (c->p.precision == CC_FP32_FD && s->threads_bulk != ccd::default_threads(c->N)) ||
        (c->stages != 2 && (s->threads_base != ccd::default_threads(c->N) ||
                            s->threads_bulk != ccd::default_threads(c->N)))) {
        c->err = "unsupported thread count for this Ntau, precision and scheme";
        return CC_BADARG;
    }
    if (s->steps_per_launch < 8 || s->steps_per_launch > 8192 ||
        (max_chain_steps(c) + s->steps_per_launch - 1) / s->steps_per_launch > kMaxChunks) {
        c->err = "steps_per_launch out of range";
        return CC_BADARG;
    }
    c->sched = *s;
    c->sched.pipeline = s->pipeline != 0;
    c->sched.use_graph = s->use_graph != 0;
    drop_graph(c);
    return CC_OK;
}

cc_status cc_autotune(cc_ctx *c, const double *x, int32_t flags, cc_schedule *out, double *times, int32_t ntimes,
                      int32_t *ntried)
{
    if (!c || !x || (flags & ~(CC_TUNE_THREADS | CC_TUNE_STAGES)) != 0) return CC_BADARG;
    if (max_chain_steps(c) < 64) {
        c->err = "autotuning needs at least 64 steps per side";
        return CC_BADARG;
    }
    HIPCHK(c, hipMemcpyAsync(c->x, x, sizeof(double) * c->n, hipMemcpyHostToDevice, c->s0));
    drop_graph(c);
    std::vector<double> t;
    const cc_schedule keep = c->sched;
    cc_status st = CC_OK;
    if (flags & CC_TUNE_THREADS)
        dispatch_n(c->N, [&](auto Nc) { st = tune_threads_n<decltype(Nc)::value>(c, c->x, t); });
    if (st == CC_OK && (flags & CC_TUNE_STAGES)) st = tune_stages(c, t);
    if (st != CC_OK) {
        c->sched = keep;
        return st;
    }
    if (out) *out = c->sched;
    if (times)
        for (int i = 0; i < std::min<int>(ntimes, (int)t.size()); ++i) times[i] = t[i];
    if (ntried) *ntried = (int32_t)t.size();
    return CC_OK;
}

cc_status cc_residual(cc_ctx *c, const double *x, double *F)
{
    if (!c || !x || !F) return CC_BADARG;
    const int n = c->n;
    HIPCHK(c, hipMemcpyAsync(c->x, x, sizeof(double) * n, hipMemcpyHostToDevice, c->s0));
    c->counters[0] = c->counters[1] = c->counters[2] = 0;
    time_reset(c);
    const Counts k{1, 1, 2, 1};                 // the base column of the base plan
    cc_status st = CC_OK;
    dispatch_n(c->N, [&](auto Nc) {
        st = launch_columns<decltype(Nc)::value>(c, c->base, k, c->x, 0.0, c->sched.threads_base, c->s0);
    });
    CCCHK(st);
    CCCHK(readback_columns(c, c->base, k, true, c->s0));
    HIPCHK(c, hipMemcpyAsync(F, c->base.F, sizeof(double) * n, hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    return eval_columns(c, c->base_plan, k, true, nullptr, true);
}

cc_status cc_jacobian(cc_ctx *c, const double *x, const double *F0, double *J, int32_t *col_status)
{
    if (!c || !x || !F0 || !J) return CC_BADARG;
    const int n = c->n;
    c->counters[0] = c->counters[1] = c->counters[2] = 0;
    time_reset(c);
    HIPCHK(c, hipMemcpyAsync(c->x, x, sizeof(double) * n, hipMemcpyHostToDevice, c->s0));
    cc_status st;
    if (tangent_policy(c)) {
        CCCHK(enqueue_tangent(c, c->x));
        HIPCHK(c, hipMemcpyAsync(J, c->J, sizeof(double) * n * n, hipMemcpyDeviceToHost, c->s0));
        HIPCHK(c, hipStreamSynchronize(c->s0));
        st = eval_tangent(c, col_status);
    } else if (c->p.precision == CC_FP32_FD) {
        // differences of FP32 shoots against the FP32 base column (an FP64 F0
        // against FP32 shoots would difference two precisions)
        CCCHK(enqueue_fp32_fd(c, c->x, nullptr));
        HIPCHK(c, hipMemcpyAsync(J, c->J, sizeof(double) * n * n, hipMemcpyDeviceToHost, c->s0));
        HIPCHK(c, hipStreamSynchronize(c->s0));
        st = eval_fp32_fd(c, col_status);
    } else {
        // the base column is part of the launch; the difference quotient uses
        // the caller's F0 (NewtonSolver.cpp:684), stored after the last column
        CCCHK(enqueue_fp64_columns(c, c->x));
        const int ncols = (int)c->main_plan.cols.size();
        HIPCHK(c, hipMemcpyAsync(c->main.F + (size_t)ncols * n, F0, sizeof(double) * n, hipMemcpyHostToDevice, c->s0));
        hipLaunchKernelGGL(ccd::k_fd, dim3((ncols + 127) / 128, n), dim3(128), 0, c->s0, c->main.F, c->main.cols,
                           ncols, ncols, n, c->p.EpsNewton, c->J);
        HIPCHK(c, hipGetLastError());
        CCCHK(readback_columns(c, c->main, counts_of(c->main_plan), false, c->s0));
        HIPCHK(c, hipMemcpyAsync(J, c->J, sizeof(double) * n * n, hipMemcpyDeviceToHost, c->s0));
        HIPCHK(c, hipStreamSynchronize(c->s0));
        st = eval_fp64(c, col_status);
    }
    c->prof_pending = c->profile;
    return st;
}

cc_status cc_solve(cc_ctx *c, int32_t n, double *A, double *b, int32_t *info)
{
    if (!c || !A || !b || n < 1 || n > c->n) return CC_BADARG;
    HIPCHK(c, hipMemcpyAsync(c->J, A, sizeof(double) * n * n, hipMemcpyHostToDevice, c->s0));
    HIPCHK(c, hipMemcpyAsync(c->d, b, sizeof(double) * n, hipMemcpyHostToDevice, c->s0));
    CCCHK(enqueue_solve(c, n, c->J, c->d, c->dx, c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    bool resolved = false;
    CCCHK(refine_or_resolve(c, n, c->J, c->d, c->dx, c->s0, &resolved));
    if (fp32_solve(c) && !resolved) {
        // FP32 factors, returned widened to double
        const unsigned g = (unsigned)(((size_t)n * n + 255) / 256);
        hipLaunchKernelGGL((ccd::k_cast_from<float>), dim3(g), dim3(256), 0, c->s0, c->A32, c->A, (size_t)n * n);
        HIPCHK(c, hipGetLastError());
    }
    HIPCHK(c, hipMemcpyAsync(A, c->A, sizeof(double) * n * n, hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipMemcpyAsync(b, c->dx, sizeof(double) * n, hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    const cc_status st = eval_info(c, info);
    if (st != CC_OK) HIPCHK(c, hipMemcpy(b, c->d, sizeof(double) * n, hipMemcpyDeviceToHost));   // b keeps its input
    return st;
}

cc_status cc_newton_step(cc_ctx *c, const double *x, double *F0, double *dx)
{
    if (!c || !x || !F0 || !dx) return CC_BADARG;
    const int n = c->n;
    HIPCHK(c, hipMemcpyAsync(c->x, x, sizeof(double) * n, hipMemcpyHostToDevice, c->s0));
    CCCHK(newton_step_device(c, c->x, c->F0, c->dx));
    HIPCHK(c, hipMemcpyAsync(F0, c->F0, sizeof(double) * n, hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipMemcpyAsync(dx, c->dx, sizeof(double) * n, hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    return CC_OK;
}

cc_status cc_newton_run(cc_ctx *c, double *x, double precision_newton, double slow_error, int32_t max_iter,
                        int32_t *iterations, double *mismatch, double *history, int32_t *outcome)
{
    if (!c || !x || max_iter < 1 || !iterations || !mismatch || !outcome) return CC_BADARG;
    const int n = c->n;
    double *hist = nullptr;
    HIPCHK(c, hipMalloc(&hist, sizeof(double) * max_iter));
    auto done = [&](cc_status st) {
        (void)hipFree(hist);
        return st;
    };
    const double st0[2] = {1.0, 1.0};
    const int32_t ist0[4] = {0, 0, 0, 0};
    hipError_t e = hipMemcpyAsync(c->x, x, sizeof(double) * n, hipMemcpyHostToDevice, c->s0);
    if (e == hipSuccess) e = hipMemcpyAsync(c->xold, c->x, sizeof(double) * n, hipMemcpyDeviceToDevice, c->s0);
    if (e == hipSuccess) e = hipMemcpyAsync(c->nst, st0, sizeof(st0), hipMemcpyHostToDevice, c->s0);
    if (e == hipSuccess) e = hipMemcpyAsync(c->nist, ist0, sizeof(ist0), hipMemcpyHostToDevice, c->s0);
    if (e != hipSuccess) {
        c->err = std::string("hip: ") + hipGetErrorString(e);
        return done(CC_DEVICE);
    }
    int its = 0, mode = 0;
    for (its = 0; its < max_iter; ++its) {
        const std::function<cc_status()> post = [&]() -> cc_status {
            hipLaunchKernelGGL(ccd::k_newton_update, dim3(1), dim3(256), 0, c->s0, c->x, c->xold, c->F0, c->dx, n,
                               precision_newton, slow_error, c->nst, c->nist, hist, its,
                               fp32_solve(c) ? (const int32_t *)c->rhold : nullptr);
            HIPCHK(c, hipGetLastError());
            HIPCHK(c, hipMemcpyAsync(c->hs.newton, c->nist + 1, 3 * sizeof(int32_t), hipMemcpyDeviceToHost, c->s0));
            return CC_OK;
        };
        bool resolved = false;
        const cc_status st = newton_step_device(c, c->x, c->F0, c->dx, &post, &resolved);
        if (st != CC_OK) {
            // a failure after the residual (Jacobian columns, LU): the reference
            // has already stopped there when the iterate converged or the
            // mismatch rose a fifth time (NewtonSolver.cpp:304-331)
            const bool late = st == CC_IRK_NOCONV || st == CC_DELTA_MISMATCH || st == CC_SINGULAR;
            if (late && base_residual_ok(c) && (c->hs.newton[0] == 1 || c->hs.newton[0] == 2)) {
                mode = c->hs.newton[0];
                break;
            }
            return done(st);
        }
        if (resolved && c->hs.newton[0] == 0 && c->hs.newton[2] == 0) {
            // the update was held back for the FP64 re-solve of the step
            hipLaunchKernelGGL(ccd::k_newton_apply, dim3(1), dim3(256), 0, c->s0, c->x, c->xold, c->dx, n, c->nst,
                               c->nist);
            e = hipGetLastError();
            if (e == hipSuccess) e = hipStreamSynchronize(c->s0);
            if (e != hipSuccess) {
                c->err = std::string("hip: ") + hipGetErrorString(e);
                return done(CC_DEVICE);
            }
        }
        mode = c->hs.newton[0];
        if (mode != 0) break;
    }
    double stv[2] = {0.0, 0.0};
    const int nh = std::min(max_iter, its + 1);
    e = hipMemcpyAsync(x, mode == 2 ? c->xold : c->x, sizeof(double) * n, hipMemcpyDeviceToHost, c->s0);
    if (e == hipSuccess) e = hipMemcpyAsync(stv, c->nst, sizeof(stv), hipMemcpyDeviceToHost, c->s0);
    if (e == hipSuccess && history) e = hipMemcpyAsync(history, hist, sizeof(double) * nh, hipMemcpyDeviceToHost, c->s0);
    if (e == hipSuccess) e = hipStreamSynchronize(c->s0);
    if (e != hipSuccess) {
        c->err = std::string("hip: ") + hipGetErrorString(e);
        return done(CC_DEVICE);
    }
    *outcome = mode;
    *iterations = mode == 0 ? max_iter : its;
    *mismatch = mode == 2 ? stv[1] : stv[0];
    return done(CC_OK);
}

cc_status cc_newton_step_device_async(cc_ctx *c, const double *d_x, double *d_F0, double *d_dx)
{
    if (!c || !d_x || !d_F0 || !d_dx) return CC_BADARG;
    return newton_step_device(c, d_x, d_F0, d_dx);
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
    CallBuffers tmp{c->s0, {}};
    cx<double> *din = nullptr, *dout = nullptr;
    HIPCHK(c, tmp.get(&din, (size_t)N * batch));
    HIPCHK(c, tmp.get(&dout, (size_t)N * batch));
    HIPCHK(c, hipMemcpyAsync(din, in, bytes, hipMemcpyHostToDevice, c->s0));
    hipError_t e = hipSuccess;
    dispatch_n(N, [&](auto Nc) {
        e = Ops<decltype(Nc)::value>::fft(c->s0, fp32 != 0, sign, batch, din, dout,
                                          fp32 ? (const void *)c->tw32 : (const void *)c->tw64);
    });
    HIPCHK(c, e);
    HIPCHK(c, hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, tmp.release());
    return CC_OK;
}

cc_status cc_spectral(cc_ctx *c, int32_t op, const double *in, double *out, int32_t batch, double period,
                      double lambda_re, double lambda_im)
{
    if (!c || !in || !out || batch < 1 || op < CC_SPEC_DIFF || op > CC_SPEC_INHOM) return CC_BADARG;
    const bool uses_period = op == CC_SPEC_DIFF || op == CC_SPEC_LAM || op == CC_SPEC_INHOM;
    if (uses_period && !(period > 0.0 && std::isfinite(period))) {
        c->err = "cc_spectral: the period must be positive and finite";
        return CC_BADARG;
    }
    const int N = c->N;
    // doubles per signal in and out
    const size_t din_n = op == CC_SPEC_DOUBLE ? (size_t)N : 2 * (size_t)N;
    const size_t dout_n = op == CC_SPEC_HALVE || op == CC_SPEC_INHOM ? (size_t)N : 2 * (size_t)N;
    CallBuffers tmp{c->s0, {}};
    double *din = nullptr, *dout = nullptr;
    HIPCHK(c, tmp.get(&din, din_n * batch));
    HIPCHK(c, tmp.get(&dout, dout_n * batch));
    HIPCHK(c, hipMemcpyAsync(din, in, sizeof(double) * din_n * batch, hipMemcpyHostToDevice, c->s0));
    const double k0 = uses_period ? 2.0 * M_PI / period : 0.0;     // SpectralTransformer.cpp:23, 142, 167
    hipError_t e = hipSuccess;
    dispatch_n(N, [&](auto Nc) {
        e = Ops<decltype(Nc)::value>::spectral(c->s0, op, batch, din, dout, c->tw64, k0, lambda_re, lambda_im);
    });
    HIPCHK(c, e);
    HIPCHK(c, hipMemcpyAsync(out, dout, sizeof(double) * dout_n * batch, hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, tmp.release());
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
    CallBuffers tmp{c->s0, {}};
    cx<double> *dYin = nullptr, *dYout = nullptr;
    double *dx_ = nullptr, *dd = nullptr, *dtmp = nullptr;
    HIPCHK(c, tmp.get(&dYin, nY));
    HIPCHK(c, tmp.get(&dYout, nY));
    HIPCHK(c, tmp.get(&dtmp, 2 * nY));
    HIPCHK(c, tmp.get(&dx_, 2 * (size_t)batch));
    HIPCHK(c, tmp.get(&dd, 2 * (size_t)batch));
    HIPCHK(c, hipMemcpyAsync(dtmp, Yp.data(), nY * sizeof(cx<double>), hipMemcpyHostToDevice, c->s0));
    HIPCHK(c, hipMemcpyAsync(dx_, xp.data(), sizeof(double) * 2 * batch, hipMemcpyHostToDevice, c->s0));
    HIPCHK(c, hipMemcpyAsync(dd, dp.data(), sizeof(double) * 2 * batch, hipMemcpyHostToDevice, c->s0));
    const unsigned g = (unsigned)((2 * nY + 255) / 256);
    hipError_t e = hipSuccess;
    dispatch_n(N, [&](auto Nc) {
        using O = Ops<decltype(Nc)::value>;
        if (fp32) {
            hipLaunchKernelGGL((ccd::k_cast_to<float>), dim3(g), dim3(256), 0, c->s0, dtmp, (float *)dYin, 2 * nY);
            e = O::rhs_pairs(c->s0, true, batch, (void *)dYin, dx_, dd, (void *)dYout, c->tw32, c->invm32, c->p.Dim);
            hipLaunchKernelGGL((ccd::k_cast_from<float>), dim3(g), dim3(256), 0, c->s0, (const float *)dYout, dtmp,
                               2 * nY);
        } else {
            e = O::rhs_pairs(c->s0, false, batch, (void *)dtmp, dx_, dd, (void *)dYout, c->tw64, c->invm64, c->p.Dim);
            if (e == hipSuccess)
                e = hipMemcpyAsync(dtmp, dYout, nY * sizeof(cx<double>), hipMemcpyDeviceToDevice, c->s0);
        }
    });
    HIPCHK(c, e);
    HIPCHK(c, hipGetLastError());
    std::vector<double> outp(2 * nY);
    HIPCHK(c, hipMemcpyAsync(outp.data(), dtmp, nY * sizeof(cx<double>), hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    for (int b = 0; b < batch; ++b)
        std::memcpy(dY + (size_t)b * 2 * M, &outp[(size_t)(2 * b) * 2 * M], sizeof(double) * 2 * M);
    HIPCHK(c, tmp.release());
    return CC_OK;
}

// Fields of FP64 states on the device, paired as k_rhs_pairs evaluates them
// (snap: 2 npairs slots of M complex; xs, delta: 2 npairs host values, the two
// members of a pair sharing delta[2 pr]): output k (k < count) is slot
// k * slot_stride, fields[k][4][N] = A = sqrt(1/IA2), U, V, F
// (ICG::StateVectorToFields, ICG.cpp:891-932; A as in ShootingSolver.cpp:91).
static cc_status fields_of_states(cc_ctx *c, const cx<double> *snap, int npairs, const std::vector<double> &xs,
                                  const std::vector<double> &delta, int count, int slot_stride, double *fields)
{
    const int N = c->N, M = c->M;
    const size_t step = (size_t)rec_step_values(N), stage = step / 2;
    CallBuffers tmp{c->s0, {}};
    cx<double> *dY = nullptr;
    double *sx = nullptr, *sd = nullptr, *rec = nullptr;
    HIPCHK(c, tmp.get(&dY, 2 * (size_t)npairs * M));
    HIPCHK(c, tmp.get(&sx, 2 * (size_t)npairs));
    HIPCHK(c, tmp.get(&sd, 2 * (size_t)npairs));
    HIPCHK(c, tmp.get(&rec, step * npairs));
    HIPCHK(c, hipMemcpyAsync(sx, xs.data(), sizeof(double) * 2 * npairs, hipMemcpyHostToDevice, c->s0));
    HIPCHK(c, hipMemcpyAsync(sd, delta.data(), sizeof(double) * 2 * npairs, hipMemcpyHostToDevice, c->s0));
    int oU = 0, oV = 0, oF = 0, oR = 0;          // record offsets of U, V, F (j < M) and r = 1/IA2
    hipError_t e = hipSuccess;
    dispatch_n(N, [&](auto Nc) {
        using RC = ccd::Rec<decltype(Nc)::value>;
        oU = RC::U;
        oV = RC::V;
        oF = RC::F;
        oR = RC::R;
        e = Ops<decltype(Nc)::value>::rhs_pairs(c->s0, false, npairs, snap, sx, sd, dY, c->tw64, c->invm64,
                                                 c->p.Dim, rec);
    });
    HIPCHK(c, e);
    std::vector<double> rh(step * npairs);
    HIPCHK(c, hipMemcpyAsync(rh.data(), rec, sizeof(double) * rh.size(), hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, tmp.release());                     // synchronizes the stream
    for (int k = 0; k < count; ++k) {
        const int s = k * slot_stride;
        const double *r = rh.data() + (size_t)(s / 2) * step + (size_t)(s % 2) * stage;
        double *A = fields + (size_t)k * 4 * N, *U = A + N, *V = U + N, *F = V + N;
        for (int j = 0; j < M; ++j) {       // second half by the symmetries of ICG.cpp:819-824
            U[j] = r[oU + j];
            V[j] = r[oV + j];
            F[j] = r[oF + j];
            U[j + M] = -U[j];
            V[j + M] = -V[j];
            F[j + M] = F[j];
        }
        for (int j = 0; j < N; ++j) A[j] = std::sqrt(r[oR + j]);    // ShootingSolver.cpp:91
    }
    return CC_OK;
}

cc_status cc_shoot_fields(cc_ctx *c, const double *x, int32_t every, int32_t capacity, int32_t *count, double *xs,
                          double *fields)
{
    if (!c || !x || !count || every < 1 || capacity < 0) return CC_BADARG;
    const int N = c->N, M = c->M, n = c->n, iM = c->iM, iR = c->iR;
    std::vector<int> li, ri;        // snapshot grid indices of each side, in the reference's order
    for (int i = 0; i < iM; ++i)
        if (i % every == 0 || i == iM - 1) li.push_back(i);
    for (int i = iR; i > iM; --i)
        if (i % every == 0 || i == iM + 1 || i == iR) ri.push_back(i);
    const int total = (int)(li.size() + ri.size());
    *count = total;
    if (capacity < total || !xs || !fields) {
        c->err = "cc_shoot_fields: " + std::to_string(total) + " snapshots need that capacity";
        return CC_BADARG;
    }
    std::vector<double> g((size_t)c->X + 1);
    HIPCHK(c, hipMemcpy(g.data(), c->grid, sizeof(double) * g.size(), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpyAsync(c->x, x, sizeof(double) * n, hipMemcpyHostToDevice, c->s0));
    const int npairs = (total + 1) / 2;
    std::vector<double> sxh(2 * (size_t)npairs), sdh(2 * (size_t)npairs, x[N / 2 + 2]);
    CallBuffers tmp{c->s0, {}};
    cx<double> *snap = nullptr;
    HIPCHK(c, tmp.get(&snap, 2 * (size_t)npairs * M));
    const Counts k{1, 1, 2, 1};                 // the base column: chain 0 left, chain 1 right
    cc_status st = CC_OK;
    hipError_t e = hipSuccess;
    dispatch_n(N, [&](auto Nc) {
        using O = Ops<decltype(Nc)::value>;
        st = launch_prologue<decltype(Nc)::value>(c, c->base, k, c->x, 0.0, c->s0);
        int slot = 0;
        for (int side = 0; side < 2 && st == CC_OK && e == hipSuccess; ++side) {
            ccd::ChainParams<double> a = chain_params<decltype(Nc)::value>(c, c->base);
            a.Y += (size_t)side * M;
            a.start += side;
            a.nsteps += side;
            a.dir += side;
            a.delta += side;
            a.status += side;
            a.fail_idx += side;
            a.its += side;
            int done = 0;
            for (int i : side ? ri : li) {
                const int steps = side ? iR - i : i;
                if (steps > done && e == hipSuccess) {
                    a.step_begin = done;
                    a.step_end = steps;
                    e = O::chain(c->s0, O::kThreads, 1, a);
                    done = steps;
                }
                if (e == hipSuccess)
                    e = hipMemcpyAsync(snap + (size_t)slot * M, a.Y, sizeof(cx<double>) * M, hipMemcpyDeviceToDevice,
                                       c->s0);
                sxh[slot++] = g[i];
            }
        }
        if (st == CC_OK && e == hipSuccess && (total & 1)) {   // pad the last pair with a copy
            e = hipMemcpyAsync(snap + (size_t)total * M, snap + (size_t)(total - 1) * M, sizeof(cx<double>) * M,
                               hipMemcpyDeviceToDevice, c->s0);
            sxh[total] = sxh[total - 1];
        }
    });
    CCCHK(st);
    HIPCHK(c, e);
    int32_t chst[2] = {0, 0};
    HIPCHK(c, hipMemcpyAsync(chst, c->base.status, sizeof(chst), hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    if (chst[0] || chst[1]) {
        c->err = "IRK fixed-point iteration did not converge in the field-snapshot shoot";
        return CC_IRK_NOCONV;
    }
    CCCHK(fields_of_states(c, snap, npairs, sxh, sdh, total, 1, fields));
    for (int s = 0; s < total; ++s) xs[s] = sxh[s];
    HIPCHK(c, tmp.release());
    return CC_OK;
}

cc_status cc_state_fields(cc_ctx *c, const double *Y, const double *xs, double *fields, int32_t batch)
{
    if (!c || !Y || !xs || !fields || batch < 1) return CC_BADARG;
    const int M = c->M;
    // each state is both members of its own pair: the members of a pair share Delta
    std::vector<double> xh(2 * (size_t)batch), dh(2 * (size_t)batch);
    for (int b = 0; b < batch; ++b) {
        xh[2 * b] = xh[2 * b + 1] = xs[b];
        dh[2 * b] = dh[2 * b + 1] = Y[(size_t)b * 2 * M + 4];        // Delta = Re Y[2]
    }
    CallBuffers tmp{c->s0, {}};
    cx<double> *snap = nullptr;
    HIPCHK(c, tmp.get(&snap, 2 * (size_t)batch * M));
    for (int b = 0; b < batch; ++b)
        for (int m = 0; m < 2; ++m)
            HIPCHK(c, hipMemcpyAsync(snap + (size_t)(2 * b + m) * M, Y + (size_t)b * 2 * M, sizeof(cx<double>) * M,
                                     hipMemcpyHostToDevice, c->s0));
    CCCHK(fields_of_states(c, snap, batch, xh, dh, batch, 2, fields));
    HIPCHK(c, tmp.release());
    return CC_OK;
}

cc_status cc_integrate(cc_ctx *c, double *Y, const double *grid, int32_t ngrid, const int32_t *start,
                       const int32_t *end, int32_t batch, int32_t fp32, int64_t *its_total, int32_t *status,
                       int32_t max_its, int32_t *fail_idx)
{
    if (!c || !Y || !grid || ngrid < 2 || !start || !end || batch < 1 || max_its < 0) return CC_BADARG;
    if (fp32 && c->stages != 2) {
        c->err = "FP32 chains integrate the two-stage scheme only";
        return CC_BADARG;
    }
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
    CallBuffers tmp{c->s0, {}};         // freed on every return below
    double *dg = nullptr, *dd = nullptr, *dtmp = nullptr;
    float *dYc = nullptr;
    int32_t *dst = nullptr, *dns = nullptr, *ddr = nullptr, *dstat = nullptr, *dfail = nullptr;
    int64_t *dits = nullptr;
    HIPCHK(c, tmp.get(&dg, ngrid));
    HIPCHK(c, tmp.get(&dd, batch));
    HIPCHK(c, tmp.get(&dtmp, 2 * nY));
    HIPCHK(c, tmp.get(&dYc, 2 * nY));
    HIPCHK(c, tmp.get(&dst, batch));
    HIPCHK(c, tmp.get(&dns, batch));
    HIPCHK(c, tmp.get(&ddr, batch));
    HIPCHK(c, tmp.get(&dstat, batch));
    HIPCHK(c, tmp.get(&dfail, batch));
    HIPCHK(c, tmp.get(&dits, batch));
    HIPCHK(c, hipMemcpyAsync(dg, grid, sizeof(double) * ngrid, hipMemcpyHostToDevice, c->s0));
    HIPCHK(c, hipMemcpyAsync(dd, dl.data(), sizeof(double) * batch, hipMemcpyHostToDevice, c->s0));
    HIPCHK(c, hipMemcpyAsync(dtmp, Y, sizeof(cx<double>) * nY, hipMemcpyHostToDevice, c->s0));
    HIPCHK(c, hipMemcpyAsync(dst, start, sizeof(int32_t) * batch, hipMemcpyHostToDevice, c->s0));
    HIPCHK(c, hipMemcpyAsync(dns, ns.data(), sizeof(int32_t) * batch, hipMemcpyHostToDevice, c->s0));
    HIPCHK(c, hipMemcpyAsync(ddr, dr.data(), sizeof(int32_t) * batch, hipMemcpyHostToDevice, c->s0));
    CCCHK(fill<int32_t>(c, dstat, 0, batch, c->s0));
    CCCHK(fill<int32_t>(c, dfail, -1, batch, c->s0));
    CCCHK(fill<int64_t>(c, dits, 0, batch, c->s0));
    const unsigned g = (unsigned)((2 * nY + 255) / 256);
    hipError_t e = hipSuccess;
    dispatch_n(N, [&](auto Nc) {
        using O = Ops<decltype(Nc)::value>;
        auto run = [&](auto tag) {
            using R = decltype(tag);
            ccd::ChainParams<R> a{};
            a.Y = static_cast<cx<R> *>(sizeof(R) == 8 ? (void *)dtmp : (void *)dYc);
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
            a.maxIts = max_its > 0 ? max_its : c->p.MaxIterIRK;
            a.Dim = c->p.Dim;
            a.precision = sizeof(R) == 8 ? c->p.PrecisionIRK : c->p.PrecisionIRK32;
            set_tableau(a, c);
            const int S = c->sched.steps_per_launch;
            const int T = sizeof(R) == 8 ? c->sched.threads_base : O::kThreads;
            for (int s0 = 0; s0 < maxsteps && e == hipSuccess; s0 += S) {
                a.step_begin = s0;
                a.step_end = std::min(maxsteps, s0 + S);
                e = O::chain(c->s0, T, batch, a);
            }
        };
        if (fp32) {
            hipLaunchKernelGGL((ccd::k_cast_to<float>), dim3(g), dim3(256), 0, c->s0, dtmp, dYc, 2 * nY);
            run(float{});
            hipLaunchKernelGGL((ccd::k_cast_from<float>), dim3(g), dim3(256), 0, c->s0, (const float *)dYc, dtmp,
                               2 * nY);
        } else {
            run(double{});
        }
    });
    HIPCHK(c, e);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(Y, dtmp, sizeof(cx<double>) * nY, hipMemcpyDeviceToHost, c->s0));
    std::vector<int32_t> stv(batch), fiv(batch);
    std::vector<int64_t> itv(batch);
    HIPCHK(c, hipMemcpyAsync(stv.data(), dstat, sizeof(int32_t) * batch, hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipMemcpyAsync(fiv.data(), dfail, sizeof(int32_t) * batch, hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipMemcpyAsync(itv.data(), dits, sizeof(int64_t) * batch, hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    if (status)
        for (int b = 0; b < batch; ++b) status[b] = stv[b];
    if (fail_idx)
        for (int b = 0; b < batch; ++b) fail_idx[b] = fiv[b];
    if (its_total)
        for (int b = 0; b < batch; ++b) its_total[b] = itv[b];
    HIPCHK(c, tmp.release());
    return CC_OK;
}

static cc_status pack_unpack(cc_ctx *c, const double *in, double *out, int32_t batch, bool pack)
{
    if (!c || !in || !out || batch < 1) return CC_BADARG;
    const size_t nin = (size_t)batch * (pack ? 3 * c->N : c->n), nout = (size_t)batch * (pack ? c->n : 3 * c->N);
    CallBuffers tmp{c->s0, {}};
    double *din = nullptr, *dout = nullptr;
    HIPCHK(c, tmp.get(&din, nin));
    HIPCHK(c, tmp.get(&dout, nout));
    HIPCHK(c, hipMemcpyAsync(din, in, sizeof(double) * nin, hipMemcpyHostToDevice, c->s0));
    hipError_t e = hipSuccess;
    dispatch_n(c->N, [&](auto Nc) {
        using O = Ops<decltype(Nc)::value>;
        e = pack ? O::pack(c->s0, batch, din, c->tw64, dout) : O::unpack(c->s0, batch, din, c->tw64, dout);
    });
    HIPCHK(c, e);
    HIPCHK(c, hipMemcpyAsync(out, dout, sizeof(double) * nout, hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, tmp.release());
    return CC_OK;
}

cc_status cc_pack(cc_ctx *c, const double *fields, double *x, int32_t batch)
{
    return pack_unpack(c, fields, x, batch, true);
}

cc_status cc_unpack(cc_ctx *c, const double *x, double *fields, int32_t batch)
{
    return pack_unpack(c, x, fields, batch, false);
}

cc_status cc_expansions(cc_ctx *c, const double *x, double *YL, double *YR)
{
    if (!c || !x || !YL || !YR) return CC_BADARG;
    const int n = c->n, M = c->M;
    HIPCHK(c, hipMemcpyAsync(c->x, x, sizeof(double) * n, hipMemcpyHostToDevice, c->s0));
    // the base column of the base plan: left chain 0, right chain 1
    const Counts k{1, 1, 2, 1};
    cc_status st = CC_OK;
    dispatch_n(c->N, [&](auto Nc) { st = launch_prologue<decltype(Nc)::value>(c, c->base, k, c->x, 0.0, c->s0); });
    CCCHK(st);
    const cx<double> *Y = static_cast<const cx<double> *>(c->base.Y);
    HIPCHK(c, hipMemcpyAsync(YL, Y, sizeof(cx<double>) * M, hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipMemcpyAsync(YR, Y + M, sizeof(cx<double>) * M, hipMemcpyDeviceToHost, c->s0));
    HIPCHK(c, hipStreamSynchronize(c->s0));
    return CC_OK;
}

// ---------------------------------------------------------------------------
// Batched independent problems
// ---------------------------------------------------------------------------
cc_status cc_batch_destroy(cc_batch *b)
{
    if (!b) return CC_BADARG;
    if (b->s_main) (void)hipStreamSynchronize(b->s_main);
    for (BatchClass &C : b->classes) {
        if (C.s0) (void)hipStreamSynchronize(C.s0);
        if (C.s1) (void)hipStreamSynchronize(C.s1);
    }
    batch_drop_graph(b);
    for (BatchClass &C : b->classes) {
        for (hipEvent_t e : C.chunk_ev) (void)hipEventDestroy(e);
        if (C.ev_fork) (void)hipEventDestroy(C.ev_fork);
        if (C.ev_join) (void)hipEventDestroy(C.ev_join);
        if (C.s0) (void)hipStreamDestroy(C.s0);
        if (C.s1) (void)hipStreamDestroy(C.s1);
    }
    for (hipEvent_t e : b->ev_done) (void)hipEventDestroy(e);
    if (b->ev_start) (void)hipEventDestroy(b->ev_start);
    if (b->s_main) (void)hipStreamDestroy(b->s_main);
    for (cc_ctx *c : b->ctx) cc_ctx_destroy(c);
    for (void *q : b->allocs) (void)hipFree(q);
    delete b;
    return CC_OK;
}

const char *cc_batch_last_error(const cc_batch *b) { return b ? b->err.c_str() : "null batch"; }

cc_status cc_batch_create(const cc_problem *problems, int32_t count, cc_batch **out)
{
    if (!problems || count < 1 || !out) return CC_BADARG;
    *out = nullptr;
    const cc_problem &p0 = problems[0];
    for (int i = 0; i < count; ++i) {
        const cc_problem &p = problems[i];
        if ((p.precision != CC_MIXED && p.precision != CC_FP64_TANGENT) || p.precision != p0.precision ||
            p.MaxIterIRK != p0.MaxIterIRK || p.PrecisionIRK != p0.PrecisionIRK || p.EpsNewton != p0.EpsNewton ||
            p.refine_iters != p0.refine_iters)
            return CC_BADARG;
    }
    cc_batch *b = new cc_batch();
    auto bail = [&](cc_status st) {
        cc_batch_destroy(b);
        return st;
    };
    for (int i = 0; i < count; ++i) {
        cc_ctx *c = nullptr;
        const cc_status st = create_ctx(&problems[i], &c, false);
        if (st != CC_OK) return bail(st);
        b->ctx.push_back(c);
    }
    if (bfail(b, "stream", hipStreamCreateWithFlags(&b->s_main, hipStreamNonBlocking))) return bail(CC_DEVICE);
    if (bfail(b, "event", hipEventCreateWithFlags(&b->ev_start, hipEventDisableTiming))) return bail(CC_DEVICE);
    for (int Nv : {128, 256, 512, 1024, 2048}) {
        BatchClass C;
        C.N = Nv;
        for (int i = 0; i < count; ++i)
            if (problems[i].Ntau == Nv) C.members.push_back(i);
        if (!C.members.empty()) b->classes.push_back(C);
    }
    const bool t64 = p0.precision == CC_FP64_TANGENT;
    const size_t elem = t64 ? sizeof(double) : sizeof(float);
    for (BatchClass &C : b->classes) {
        const int N = C.N, M = N / 2;
        const size_t step = (size_t)rec_step_values(N);
        hipEvent_t e;
        if (bfail(b, "stream", hipStreamCreateWithFlags(&C.s0, hipStreamNonBlocking))) return bail(CC_DEVICE);
        if (bfail(b, "stream", hipStreamCreateWithFlags(&C.s1, hipStreamNonBlocking))) return bail(CC_DEVICE);
        if (bfail(b, "event", hipEventCreateWithFlags(&C.ev_fork, hipEventDisableTiming))) return bail(CC_DEVICE);
        if (bfail(b, "event", hipEventCreateWithFlags(&C.ev_join, hipEventDisableTiming))) return bail(CC_DEVICE);
        if (bfail(b, "event", hipEventCreateWithFlags(&e, hipEventDisableTiming))) return bail(CC_DEVICE);
        b->ev_done.push_back(e);
        size_t ngrid = 0, nrec = 0;
        for (int i : C.members) {
            cc_ctx *c = b->ctx[i];
            C.ntan += (int)c->t_cols_h.size();
            ngrid += (size_t)c->X + 1;
            nrec += (size_t)c->X;
            C.maxsteps = std::max(C.maxsteps, max_chain_steps(c));
        }
        C.nbase = 4 * (int)C.members.size();
        hipError_t e2 = hipSuccess;
        char *dyp = nullptr, *recp = nullptr, *fpp = nullptr;
#define BALLOC(ptr, cnt) \
    if (e2 == hipSuccess) e2 = balloc(b, &(ptr), (cnt));
        BALLOC(C.Y, (size_t)C.nbase * M);
        BALLOC(C.delta, C.nbase);
        BALLOC(C.dim, C.nbase);
        BALLOC(C.grid, ngrid);
        for (int32_t **q : {&C.start, &C.nsteps, &C.dir, &C.status, &C.fail, &C.goff}) BALLOC(*q, C.nbase);
        BALLOC(C.its, C.nbase);
        BALLOC(C.rec_base, C.nbase);
        BALLOC(dyp, (size_t)C.ntan * M * 2 * elem);
        BALLOC(fpp, (size_t)C.ntan * 2 * M * 2 * elem);
        BALLOC(recp, nrec * step * elem);
        BALLOC(C.t_dim, C.ntan);
        for (int32_t **q : {&C.t_start, &C.t_nsteps, &C.t_dir, &C.t_status, &C.t_fail, &C.t_goff, &C.t_bs})
            BALLOC(*q, C.ntan);
        BALLOC(C.t_its, C.ntan);
        BALLOC(C.t_rec_base, C.ntan);
#undef BALLOC
        if (bfail(b, "hipMalloc", e2)) return bail(CC_DEVICE);
        C.dY = dyp;
        C.rec = recp;
        C.fprev = fpp;
        // point every member into the class arrays; per-chain tables for the shared launches
        std::vector<double> dim_b, dim_t;
        std::vector<int32_t> goff_b, goff_t, bs_t, tstart, tnsteps, tdir;
        std::vector<int64_t> rb_b, rb_t;
        size_t go = 0, ro = 0;
        int to = 0;
        for (size_t m = 0; m < C.members.size(); ++m) {
            cc_ctx *c = b->ctx[C.members[m]];
            const int bo = 4 * (int)m;
            if (bfail(b, "grid", hipMemcpy(C.grid + go, c->grid, sizeof(double) * (c->X + 1), hipMemcpyDeviceToDevice)))
                return bail(CC_DEVICE);
            c->grid = C.grid + go;
            c->base.Y = C.Y + (size_t)bo * M;
            c->base.delta = C.delta + bo;
            c->base.start = C.start + bo;
            c->base.nsteps = C.nsteps + bo;
            c->base.dir = C.dir + bo;
            c->base.status = C.status + bo;
            c->base.fail = C.fail + bo;
            c->base.its = C.its + bo;
            if (bfail(b, "plan", upload_plan(c->base_plan, c->base))) return bail(CC_DEVICE);
            c->rec = recp + ro * step * elem;
            c->dY = dyp + (size_t)to * M * 2 * elem;
            c->fprev = fpp + (size_t)to * 2 * M * 2 * elem;
            c->t_status = C.t_status + to;
            c->t_fail = C.t_fail + to;
            c->t_its = C.t_its + to;
            c->t_start = C.t_start + to;
            c->t_nsteps = C.t_nsteps + to;
            c->t_dir = C.t_dir + to;
            for (int k = 0; k < 4; ++k) {
                dim_b.push_back(c->p.Dim);
                goff_b.push_back((int32_t)go);
            }
            rb_b.insert(rb_b.end(), {(int64_t)ro, (int64_t)(ro + c->iM), -1, -1});
            for (size_t j = 0; j < c->t_cols_h.size(); ++j) {
                const bool left = c->t_cols_h[j] >= N / 4;
                dim_t.push_back(c->p.Dim);
                goff_t.push_back((int32_t)go);
                bs_t.push_back(bo);
                tstart.push_back(left ? 0 : c->iR);
                tnsteps.push_back(c->t_nsteps_h[j]);
                tdir.push_back(left ? 1 : -1);
                rb_t.push_back((int64_t)(ro + (left ? 0 : c->iM)));
            }
            go += (size_t)c->X + 1;
            ro += (size_t)c->X;
            to += (int)c->t_cols_h.size();
        }
        hipError_t eu = up(dim_b, C.dim);
        if (eu == hipSuccess) eu = up(goff_b, C.goff);
        if (eu == hipSuccess) eu = up(rb_b, C.rec_base);
        if (eu == hipSuccess) eu = up(dim_t, C.t_dim);
        if (eu == hipSuccess) eu = up(goff_t, C.t_goff);
        if (eu == hipSuccess) eu = up(bs_t, C.t_bs);
        if (eu == hipSuccess) eu = up(tstart, C.t_start);
        if (eu == hipSuccess) eu = up(tnsteps, C.t_nsteps);
        if (eu == hipSuccess) eu = up(tdir, C.t_dir);
        if (eu == hipSuccess) eu = up(rb_t, C.t_rec_base);
        if (bfail(b, "plan", eu)) return bail(CC_DEVICE);
        C.sched = b->ctx[C.members[0]]->sched;
    }
    *out = b;
    return CC_OK;
}

cc_status cc_batch_newton_step(cc_batch *b, const double *const *x, double *const *F0, double *const *dx,
                               int32_t *status)
{
    if (!b || !x) return CC_BADARG;
    for (size_t i = 0; i < b->ctx.size(); ++i) {
        if (!x[i]) return CC_BADARG;
        BCHK(b, hipMemcpy(b->ctx[i]->x, x[i], sizeof(double) * b->ctx[i]->n, hipMemcpyHostToDevice));
    }
    const cc_status st = batch_run(b, status);
    for (size_t i = 0; i < b->ctx.size(); ++i) {
        cc_ctx *c = b->ctx[i];
        if (F0 && F0[i]) BCHK(b, hipMemcpy(F0[i], c->F0, sizeof(double) * c->n, hipMemcpyDeviceToHost));
        if (dx && dx[i]) BCHK(b, hipMemcpy(dx[i], c->dx, sizeof(double) * c->n, hipMemcpyDeviceToHost));
    }
    return st;
}

cc_status cc_batch_info(cc_batch *b, int32_t *nproblems, int32_t *nclasses)
{
    if (!b) return CC_BADARG;
    if (nproblems) *nproblems = (int32_t)b->ctx.size();
    if (nclasses) *nclasses = (int32_t)b->classes.size();
    return CC_OK;
}

cc_status cc_batch_get_schedule(cc_batch *b, int32_t cls, int32_t *Ntau, cc_schedule *out)
{
    if (!b || cls < 0 || cls >= (int32_t)b->classes.size() || !out) return CC_BADARG;
    if (Ntau) *Ntau = b->classes[cls].N;
    *out = b->classes[cls].sched;
    out->use_graph = b->use_graph;
    return CC_OK;
}

cc_status cc_batch_set_schedule(cc_batch *b, int32_t cls, const cc_schedule *s)
{
    if (!b || cls < 0 || cls >= (int32_t)b->classes.size() || !s) return CC_BADARG;
    BatchClass &C = b->classes[cls];
    cc_ctx *c0 = b->ctx[C.members[0]];
    const int T0 = ccd::default_threads(C.N);
    const bool t64 = c0->p.precision == CC_FP64_TANGENT;
    if (!valid_threads(C.N, s->threads_base) || !valid_threads(C.N, s->threads_tangent) ||
        (t64 && (s->threads_base != T0 || s->threads_tangent != T0)) || s->steps_per_launch < 8 ||
        s->steps_per_launch > 8192 || (C.maxsteps + s->steps_per_launch - 1) / s->steps_per_launch > kMaxChunks) {
        b->err = "invalid schedule for this class";
        return CC_BADARG;
    }
    C.sched = *s;
    C.sched.pipeline = s->pipeline != 0;
    b->use_graph = s->use_graph != 0;
    batch_drop_graph(b);
    return CC_OK;
}

cc_status cc_batch_counters(cc_batch *b, int64_t out[3])
{
    if (!b || !out) return CC_BADARG;
    for (int i = 0; i < 3; ++i) out[i] = b->counters[i];
    return CC_OK;
}

// Per class: thread count of the tangent chains x pipeline x steps_per_launch,
// timing the class's step alone; then the whole batch eager against graph.
cc_status cc_batch_autotune(cc_batch *b, const double *const *x, double *times, int32_t ntimes, int32_t *ntried)
{
    if (!b || !x) return CC_BADARG;
    for (size_t i = 0; i < b->ctx.size(); ++i) {
        if (!x[i]) return CC_BADARG;
        BCHK(b, hipMemcpy(b->ctx[i]->x, x[i], sizeof(double) * b->ctx[i]->n, hipMemcpyHostToDevice));
    }
    std::vector<double> t;
    b->use_graph = false;
    batch_drop_graph(b);
    cc_ctx *c0 = b->ctx[0];
    for (BatchClass &C : b->classes) {
        const bool t64 = b->ctx[C.members[0]]->p.precision == CC_FP64_TANGENT;
        std::vector<int> Ts = {ccd::default_threads(C.N)};
        if (!t64 && ccd::alt_threads(C.N)) Ts.push_back(ccd::alt_threads(C.N));
        cc_schedule best = C.sched;
        double best_ms = 1e300;
        for (int T : Ts)
            for (int pipe : {1, 0})
                for (int S : {16, 32, 64, 128, 512}) {
                    if ((C.maxsteps + S - 1) / S > kMaxChunks) continue;
                    C.sched.threads_tangent = T;
                    C.sched.pipeline = pipe;
                    C.sched.steps_per_launch = S;
                    BCHK(b, hipEventRecord(c0->tk0, C.s0));
                    CCCHK(batch_enqueue_class(b, C));
                    BCHK(b, hipEventRecord(c0->tk1, C.s0));
                    BCHK(b, hipEventSynchronize(c0->tk1));
                    float f = 0.0f;
                    BCHK(b, hipEventElapsedTime(&f, c0->tk0, c0->tk1));
                    t.push_back(f);
                    if (f < best_ms) {
                        best_ms = f;
                        best = C.sched;
                    }
                }
        C.sched = best;
    }
    double eager = 0.0, graph = 0.0;
    CCCHK(batch_timed(b, &eager));
    b->use_graph = true;
    double dummy = 0.0;
    cc_status st = batch_timed(b, &dummy);          // capture + first replay
    if (st == CC_OK) st = batch_timed(b, &graph);
    if (st != CC_OK) return st;
    t.push_back(eager);
    t.push_back(graph);
    b->use_graph = graph < eager;
    if (!b->use_graph) batch_drop_graph(b);
    if (times)
        for (int i = 0; i < std::min<int>(ntimes, (int)t.size()); ++i) times[i] = t[i];
    if (ntried) *ntried = (int32_t)t.size();
    return CC_OK;
}

cc_status cc_dgesv_noalloc(int32_t n, double *d_A, double *d_b, int32_t *info)
{
    if (!d_A || !d_b || !info) return CC_BADARG;
    if (n < 1 || n > ccd::LU_MAXN) {
        fprintf(stderr, "cc_dgesv_noalloc: order %d is outside [1, %d], the orders its fixed scratch holds\n", (int)n,
                ccd::LU_MAXN);
        return CC_BADARG;
    }
    int32_t *sc = nullptr;
    if (scratch_address(&sc) != hipSuccess) return CC_DEVICE;
    int32_t *perm = sc + kScratchPerm, *swap = sc + kScratchSwap, *dinfo = sc + kScratchInfo;
    const hipStream_t s = nullptr;
    // the info word keeps the last call's value: cleared by a kernel first
    hipLaunchKernelGGL((ccd::k_fill<int32_t>), dim3(1), dim3(64), 0, s, dinfo, 0, (size_t)1);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = lu_factor_on<double>(d_A, n, perm, swap, dinfo, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL((ccd::k_lu_solve<double, double>), dim3(1), dim3(1024), 0, s, d_A, n, perm, d_b, dinfo);
        e = hipGetLastError();
    }
    int32_t h = 0;
    if (e == hipSuccess) e = hipMemcpy(&h, dinfo, sizeof(int32_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return CC_DEVICE;
    *info = h;
    return h != 0 ? CC_SINGULAR : CC_OK;
}

}  // extern "C"
