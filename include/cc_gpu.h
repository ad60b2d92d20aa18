/*
 * cc_gpu.h -- C-ABI of the MI355X (gfx950) implementation of the reference's
 * finite-difference Jacobian assembly and Newton solve
 * (tobjec/parallel-critical-collapse, src/NewtonSolver.cpp:615-757 and the
 * shooting path below it).
 *
 * Plain C types only.  Every entry point returns a cc_status; no exception
 * crosses the boundary.  Host arrays are caller-owned and every call is
 * synchronous on the context's stream unless its name ends in _async.  One
 * context per host thread.  Arrays of complex values are interleaved
 * (re, im) doubles.
 *
 * Sizes: N = Ntau, M = N/2 complex modes per state, n = 3N/4 Newton unknowns.
 * Newton vector layout (NewtonSolver.cpp:273-274, ICG.cpp:678-711):
 *   x[0 : N/4)      Up   odd modes 1,3,..  as (re, im) pairs
 *   x[N/4 : N/2)    psic odd modes
 *   x[N/2 : 3N/4)   fc   even modes 0,2,.. as (re, im); x[N/2+1] holds the
 *                   folded mode N/4 cosine and x[N/2+2] holds Delta.
 */
#ifndef CC_GPU_H
#define CC_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum cc_status {
    CC_OK = 0,
    CC_IRK_NOCONV = 1,      /* ShootingSolver.cpp:69-82, 112-125: IRK fixed point did not converge */
    CC_DELTA_MISMATCH = 2,  /* ShootingSolver.cpp:167-178: left/right Delta differ */
    CC_SINGULAR = 3,        /* NewtonSolver.cpp:751-755: dgesv info != 0 */
    CC_BADARG = 4,
    CC_DEVICE = 5           /* HIP runtime error; see cc_last_error */
} cc_status;

typedef enum cc_precision {
    CC_FP64 = 0,            /* every column a forward difference of two FP64 shoots: the reference's arithmetic */
    CC_MIXED = 1,           /* FP64 base shoot, F(x) and Delta column; every other column the FP32
                               tangent-linear derivative of the shoot; FP32 LU + FP64 refinement */
    CC_FP64_TANGENT = 2,    /* as CC_MIXED with the tangent-linear columns and the LU in FP64:
                               an accurate all-FP64 Jacobian (no counterpart in the reference) */
    CC_FP32_FD = 3          /* every column a forward difference of two FP32 shoots (FP32 chains at
                               PrecisionIRK32, step EpsNewton: set it far above the reference's 1e-8,
                               which FP32 rounding swamps); F(x) an FP64 shoot; FP32 LU + FP64
                               refinement.  The literal FP32 form of the reference's Jacobian, kept
                               to measure against CC_MIXED (two-stage scheme only) */
} cc_precision;

/* Problem description: the SimulationConfig keys the path reads
 * (include/SimulationConfig.hpp:73-137) plus the precision policy. */
typedef struct cc_problem {
    int32_t Ntau;           /* 128, 256, 512, 1024 or 2048 */
    int32_t SchemeIRK;      /* Gauss-Legendre stages 1, 2 or 3 (ODEStepper.cpp:22-80); the
                               tangent-linear policies (CC_MIXED, CC_FP64_TANGENT) need 2 */
    int32_t MaxIterIRK;
    int32_t NLeft, NRight;
    double  Dim;
    double  XLeft, XMid, XRight;
    double  PrecisionIRK;
    double  EpsNewton;      /* absolute FD step (NewtonSolver.cpp:637, 644) */
    /* precision policy (new keys; defaults reproduce the reference) */
    int32_t precision;      /* cc_precision */
    int32_t reuse_sides;    /* 1: a column that leaves one side's boundary data unchanged
                               reuses the base shoot of that side (bit-identical result) */
    double  PrecisionIRK32; /* IRK tolerance of FP32 chains in cc_integrate (0 = default 2e-6) */
    double  refine_tol;     /* CC_MIXED: relative residual ||J dx - b|| / ||b|| the FP64-refined
                               solve must reach; above it (or not finite) the system is solved
                               again with the FP64 LU (0 = default 1e-10) */
    int32_t refine_iters;   /* FP64 iterative-refinement sweeps of the Newton step (0 = default 3) */
    int32_t reserved;
} cc_problem;

typedef struct cc_ctx cc_ctx;

/* Context: owns device buffers, the stream, FFT twiddle tables and the
 * x-grid (NewtonSolver.cpp:477-519).  Replaces the ICG/ShootingSolver/
 * ODEStepper objects NewtonSolver.cpp:18-52 constructs. */
cc_status cc_ctx_create(const cc_problem *p, cc_ctx **out);
cc_status cc_ctx_destroy(cc_ctx *ctx);
const char *cc_last_error(const cc_ctx *ctx);
/* Copies the context's x-grid (NLeft+NRight+1 doubles) to host memory. */
cc_status cc_grid(cc_ctx *ctx, double *grid);

/* F(x) in FP64: replaces NewtonSolver::shoot (NewtonSolver.cpp:403-423). */
cc_status cc_residual(cc_ctx *ctx, const double *x, double *F);

/* FD Jacobian: replaces NewtonSolver::assembleJacobian
 * (NewtonSolver.cpp:615-702).  J is row-major n x n, J[j*n+i] = dF_j/dx_i.
 * col_status (n int32, may be NULL) receives each column's cc_status.
 * CC_FP64: forward differences (F(x + EpsNewton e_i) - F0) / EpsNewton, as
 * the reference.  CC_MIXED: F0 is not read; the Delta column is that forward
 * difference in FP64 around an FP64 base shoot, every other column is the
 * FP32 tangent-linear derivative of the shoot along e_i, integrated around
 * the recorded FP64 base trajectory.  CC_FP32_FD: F0 is not read; every
 * column is (F32(x + EpsNewton e_i) - F32(x)) / EpsNewton of FP32 shoots. */
cc_status cc_jacobian(cc_ctx *ctx, const double *x, const double *F0, double *J, int32_t *col_status);

/* Dense solve with partial pivoting: replaces LAPACKE_dgesv
 * (NewtonSolver.cpp:739-757).  A (n x n row-major, 1 <= n <= 3 Ntau/4) is
 * overwritten by its LU factors, b by the solution; *info follows LAPACK
 * (0 ok, k > 0 zero pivot at k).  In CC_MIXED the factors are FP32 (returned
 * widened) and the solution is refined in FP64; when the refined residual
 * stays above refine_tol the system is factored again in FP64 and A
 * receives the FP64 factors (cc_refinement reports which). */
cc_status cc_solve(cc_ctx *ctx, int32_t n, double *A, double *b, int32_t *info);

/* cc_solve of a CC_FP64 context without any allocation, for callers that
 * cannot allocate device memory (code under hipGraph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
 * d_A (n x n row-major) and d_b (n) are device pointers on the current
 * device: d_A is overwritten by the LU factors, d_b by the solution (left
 * unchanged when *info > 0, as dgesv); *info (host) follows LAPACK.  Same
 * kernels and results as cc_solve.  Its scratch (row permutation, panel
 * swap list, info word: 8960 bytes) is a fixed device array of the library,
 * one per GPU, so 1 <= n <= 2048 (else CC_BADARG), calls that overlap on one
 * GPU race, and each GPU keeps the array until the process ends.
 * Synchronous on the null stream. */
cc_status cc_dgesv_noalloc(int32_t n, double *d_A, double *d_b, int32_t *info);

/* Outcome of the last CC_MIXED solve of this context (cc_solve,
 * cc_newton_step*, cc_newton_run): *residual = ||J dx - b|| / ||b||
 * after the FP64 refinement of the FP32 solve, *fp64_resolve = 1 when that
 * exceeded refine_tol (or was not finite) and the FP64 LU solved the system
 * instead.  Both 0 for the other policies.  No counterpart in the reference
 * (its dgesv is FP64 throughout). */
cc_status cc_refinement(cc_ctx *ctx, double *residual, int32_t *fp64_resolve);

/* One Newton step without host round trips: F0 = F(x), J, solve J dx = -F0
 * (+ FP64 refinement in CC_MIXED).  Outputs F0 and dx (n doubles each);
 * the update x += fac*dx stays with the caller (NewtonSolver.cpp:350-358). */
cc_status cc_newton_step(cc_ctx *ctx, const double *x, double *F0, double *dx);

/* Same as cc_newton_step on device-resident arrays (d_x, d_F0, d_dx are
 * device pointers of n doubles).  Returns when the step's status is known
 * (one host synchronisation at its end); with schedule.use_graph the step is
 * replayed from a hipGraph captured for this (d_x, d_F0, d_dx). */
cc_status cc_newton_step_device_async(cc_ctx *ctx, const double *d_x, double *d_F0, double *d_dx);
cc_status cc_synchronize(cc_ctx *ctx);

/* The loop of NewtonSolver::run (NewtonSolver.cpp:279-366) with the
 * mismatch norm, the damping min(1, slow_error/err), the update and the
 * divergence guard on the device (one pinned status word read per
 * iteration, no host copies of F0 or dx): up to max_iter Newton steps from
 * x.  outcome 1: converged (err < precision_newton), x is that point;
 * outcome 2: the mismatch rose five times in a row, x is the last point whose
 * mismatch fell; outcome 0: max_iter steps without either, x is the last
 * iterate.  *iterations is the reference's `its` at termination, *mismatch the
 * final err (errOld for outcome 2), history (may be null, max_iter doubles)
 * the mismatch of each step.  On an error status x is left unchanged.  The
 * reference tests convergence and the divergence guard before it assembles J
 * (NewtonSolver.cpp:304-331), so a Jacobian or LU failure at an iterate whose
 * own residual is valid and ends the loop (outcome 1 or 2) is not an error. */
cc_status cc_newton_run(cc_ctx *ctx, double *x, double precision_newton, double slow_error, int32_t max_iter,
                        int32_t *iterations, double *mismatch, double *history, int32_t *outcome);

/* Copies the Jacobian assembled by the last cc_jacobian / cc_newton_step*
 * call (n x n row-major) to host memory: the matrix the reference writes as
 * jacobian_<Dim>_<i>.csv in Debug runs (NewtonSolver.cpp:337-341). */
cc_status cc_last_jacobian(cc_ctx *ctx, double *J);

/* Field snapshots of the base shoot at x, the Debug output of
 * ShootingSolver::integrateToMatchPoint (ShootingSolver.cpp:85-97) that
 * NewtonSolver writes as fields_<Dim>_<its>.json (NewtonSolver.cpp:287-295):
 * the left side at grid indices i in [0, NLeft) with i % every == 0 or
 * i == NLeft-1 (ascending), then the right side at i in (NLeft, NLeft+NRight]
 * with i % every == 0, i == NLeft+1 or i == NLeft+NRight (descending), each
 * the state reached at x_i, integrated from the reference's start.  Snapshot
 * k: xs[k] = x_i, fields[k][4][Ntau] = A = sqrt(1/IA2), U, V, F.  *count
 * receives the number of snapshots; capacity below it gives CC_BADARG with
 * nothing written. */
cc_status cc_shoot_fields(cc_ctx *ctx, const double *x, int32_t every, int32_t capacity, int32_t *count,
                          double *xs, double *fields);

/* Counters of the last Jacobian/Newton launch: [0] IRK steps, [1] fixed-point
 * iterations (sum of itsReached), [2] chains integrated. */
cc_status cc_counters(cc_ctx *ctx, int64_t out[3]);

/* Chain-kernel timing (the reference's AssembleStep_i timer,
 * NewtonSolver.cpp:326-335, at kernel granularity).  When enabled, HIP events
 * bracket the IRK chain kernels of every launch on the main stream;
 * cc_chain_timing reports the last launch set: elapsed ms, number of kernel
 * launches, fixed-point iterations and IRK steps they executed.  enable: 0
 * off; 1 those launches and the base chains' (cc_base_timing); 2 those
 * launches only, no events on the base chains' stream.  In the mixed policy
 * the slowest tangent chains of the previous Jacobian run in a launch of their
 * own on a third stream; those launches are not timed, and their iterations
 * and steps are not in the counts. */
cc_status cc_profile(cc_ctx *ctx, int32_t enable);
cc_status cc_chain_timing(cc_ctx *ctx, double *ms, int32_t *launches, int64_t *its, int64_t *steps);
/* The same for the FP64 base chains of the tangent-linear policies (the base
 * shoot and the Delta column, on the second stream when pipelined): event
 * pairs around each of their launches in the last profiled step. */
cc_status cc_base_timing(cc_ctx *ctx, double *ms, int32_t *launches, int64_t *its, int64_t *steps);

/* ---- launch schedule and autotuner (no counterpart in the reference) ----- */

/* How the chain kernels are launched.  Thread counts are threads per chain
 * (one workgroup each): Ntau/4 (Ntau = 128: 64) or, for Ntau >= 512, Ntau/8. */
typedef struct cc_schedule {
    int32_t threads_base;      /* FP64 chains of small launches: residual, mixed base + Delta column */
    int32_t threads_bulk;      /* FP64 chains of full-Jacobian launches */
    int32_t threads_tangent;   /* FP32 tangent-linear chains (CC_MIXED) */
    int32_t steps_per_launch;  /* IRK steps per chain-kernel launch, 8 .. 8192 */
    int32_t pipeline;          /* CC_MIXED: tangent chunk k starts as soon as base chunk k is done,
                                  the base running on a second stream */
    int32_t use_graph;         /* Newton steps replay a hipGraph captured on first use */
    int32_t reserved[2];
} cc_schedule;

cc_status cc_get_schedule(cc_ctx *ctx, cc_schedule *out);
/* Validates and installs a schedule (drops a captured graph). */
cc_status cc_set_schedule(cc_ctx *ctx, const cc_schedule *s);

enum { CC_TUNE_THREADS = 1, CC_TUNE_STAGES = 2 };
/* Autotuner.  CC_TUNE_THREADS: times every thread count of each chain
 * kernel on a slice of the integration at x (device work only, a few
 * hundred ms).  CC_TUNE_STAGES: times whole Newton steps at x for
 * steps_per_launch in {8, 12, 16, 24, 32, 64} pipelined (tangent-linear policies) and
 * {64, 128, 256, 512, 1024} on one stream.
 * Installs and returns the fastest schedule.  times (may be NULL) receives
 * up to ntimes candidate timings in ms, in the order tried; *ntried their
 * count. */
cc_status cc_autotune(cc_ctx *ctx, const double *x, int32_t flags, cc_schedule *out, double *times,
                      int32_t ntimes, int32_t *ntried);

/* ---- batched independent problems ---------------------------------------- */

/* A set of independent problems (any mix of Ntau, grid sizes and Dim; one
 * tangent precision policy, MaxIterIRK, PrecisionIRK, EpsNewton for all),
 * stepped together: per Newton step, the base chains of all problems of one
 * Ntau run in one launch per chunk and their tangent chains in another,
 * problems of different Ntau concurrently.  The batch form of the loop body
 * of NewtonSolver::run (NewtonSolver.cpp:298-358) over a dimension sweep
 * (main.cpp:236-356 runs those one after another). */
typedef struct cc_batch cc_batch;
cc_status cc_batch_create(const cc_problem *problems, int32_t count, cc_batch **out);
cc_status cc_batch_destroy(cc_batch *b);
const char *cc_batch_last_error(const cc_batch *b);
cc_status cc_batch_info(cc_batch *b, int32_t *nproblems, int32_t *nclasses);
/* x, F0, dx: `count` host pointers (n_i doubles each; F0/dx entries may be
 * NULL); status[count] receives each problem's cc_status.  Returns the worst. */
cc_status cc_batch_newton_step(cc_batch *b, const double *const *x, double *const *F0, double *const *dx,
                               int32_t *status);
/* Schedules per Ntau class (cls < nclasses, ascending Ntau); use_graph is
 * batch-wide. */
cc_status cc_batch_get_schedule(cc_batch *b, int32_t cls, int32_t *Ntau, cc_schedule *out);
cc_status cc_batch_set_schedule(cc_batch *b, int32_t cls, const cc_schedule *s);
cc_status cc_batch_autotune(cc_batch *b, const double *const *x, double *times, int32_t ntimes, int32_t *ntried);
/* Summed counters of the last batch step (as cc_counters). */
cc_status cc_batch_counters(cc_batch *b, int64_t out[3]);

/* ---- unit entry points (the reference's lower-level interfaces) ---------- */

/* SpectralTransformer forwardFFT / backwardFFT, complex-to-complex
 * (SpectralTransformer.cpp:97-132) on `batch` signals of length Ntau.
 * sign = +1: forward (exp(+i..), divided by N); sign = -1: backward.
 * fp32 != 0 runs the FP32 transform (inputs/outputs still double). */
cc_status cc_fft_batched(cc_ctx *ctx, const double *in, double *out, int32_t batch, int32_t sign,
                         int32_t fp32);

/* ODEStepper::computeDerivatives (ODEStepper.cpp:92-136) on `batch` states
 * Y (M complex each) at abscissae xs (batch doubles); dY receives M complex
 * each.  Delta of each state is Re Y[2] as in the reference. */
cc_status cc_rhs_eval(cc_ctx *ctx, const double *Y, const double *xs, double *dY, int32_t batch,
                      int32_t fp32);

/* ShootingSolver::integrateToMatchPoint (ShootingSolver.cpp:53-145) over a
 * caller-supplied grid for `batch` independent states: state b starts at
 * grid index start[b] and steps to end[b].  its_total[b] receives the sum of
 * itsReached; status[b] 0, or 1 when a step's fixed point did not converge
 * (the state then stays at that step's start and fail_idx[b], if fail_idx is
 * not NULL, receives the grid index the step started from; -1 otherwise).
 * max_its > 0 replaces MaxIterIRK for this call (the maxIts argument of
 * ODEStepper::integrate), 0 keeps it.  With one step this is
 * ODEStepper::integrate (ODEStepper.cpp:144-150). */
cc_status cc_integrate(cc_ctx *ctx, double *Y, const double *grid, int32_t ngrid,
                       const int32_t *start, const int32_t *end, int32_t batch, int32_t fp32,
                       int64_t *its_total, int32_t *status, int32_t max_its, int32_t *fail_idx);

/* SpectralTransformer's operators on `batch` signals of length Ntau (complex
 * values interleaved re, im):
 *   CC_SPEC_DIFF    differentiate (SpectralTransformer.cpp:140-158): N complex -> N complex
 *   CC_SPEC_LAM     lamIntegrate (165-189) with lambda = lambda_re + i lambda_im: N -> N complex
 *   CC_SPEC_HALVE   halveModes (196-217): N complex -> N/2 complex
 *   CC_SPEC_DOUBLE  doubleModes (224-252): N/2 complex -> N complex
 *   CC_SPEC_INHOM   solveInhom (264-292): in = [batch][f (N real), g (N real)] -> out = [batch][N real]
 * The wavenumber is 2 pi / period (period > 0; HALVE and DOUBLE ignore it). */
typedef enum cc_spectral_op {
    CC_SPEC_DIFF = 0,
    CC_SPEC_LAM = 1,
    CC_SPEC_HALVE = 2,
    CC_SPEC_DOUBLE = 3,
    CC_SPEC_INHOM = 4
} cc_spectral_op;
cc_status cc_spectral(cc_ctx *ctx, int32_t op, const double *in, double *out, int32_t batch, double period,
                      double lambda_re, double lambda_im);

/* InitialConditionGenerator::StateVectorToFields (ICG.cpp:891-932) for
 * `batch` states Y (M complex each) at abscissae xs: fields[b][4][Ntau] =
 * A = sqrt(1/IA2), U, V, F, the Debug snapshot of ShootingSolver.cpp:85-97.
 * Delta of each state is Re Y[2]. */
cc_status cc_state_fields(cc_ctx *ctx, const double *Y, const double *xs, double *fields, int32_t batch);

/* ICG::packSpectralFields (ICG.cpp:678-711) for `batch` field sets:
 * fields = [batch][3][Ntau] real tau-series Up, psic, fc -> x = [batch][n].
 * Slot N/2+2 receives Re of fc's mode 2; the Newton loop overwrites it with
 * Delta (NewtonSolver.cpp:273-274). */
cc_status cc_pack(cc_ctx *ctx, const double *fields, double *x, int32_t batch);
/* ICG::unpackSpectralFields (ICG.cpp:719-758), the Delta slot read as zero
 * as at its call sites (NewtonSolver.cpp:408-409): x -> fields. */
cc_status cc_unpack(cc_ctx *ctx, const double *x, double *fields, int32_t batch);

/* InitialConditionGenerator boundary data for Newton vector x: unpack
 * (ICG.cpp:719-758), computeLeftExpansion (37-239) and
 * computeRightExpansion (252-667).  YL, YR: M complex each. */
cc_status cc_expansions(cc_ctx *ctx, const double *x, double *YL, double *YR);

#ifdef __cplusplus
}
#endif
#endif
