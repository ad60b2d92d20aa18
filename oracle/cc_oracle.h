/*
 * cc_oracle.h -- CPU FP64 restatement of the reference's Newton-shooting path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation and the CPU baseline timed beside it.  Only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it;
 * the product path never calls into it.
 *
 * Every function restates one reference routine, cited as file:line of
 * tobjec/parallel-critical-collapse.  Arithmetic order follows the reference
 * except inside the FFT (the reference calls FFTW; this file has its own
 * radix-2 transform with the same sign and scale conventions).
 *
 * Index / storage conventions (identical to the reference):
 *   N   = Ntau            real-space samples per period
 *   M   = N/2             complex modes in a state vector Y
 *   n   = 3N/4            Newton unknowns (NewtonSolver.cpp:19)
 *   complex arrays are interleaved (re, im) doubles.
 */
#ifndef CC_ORACLE_H
#define CC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cco_ctx cco_ctx;

/* Status codes shared with the GPU library (cc_gpu.h). */
enum {
    CCO_OK = 0,
    CCO_IRK_NOCONV = 1,      /* ShootingSolver.cpp:69-82 / 112-125 */
    CCO_DELTA_MISMATCH = 2,  /* ShootingSolver.cpp:167-178 */
    CCO_SINGULAR = 3,        /* NewtonSolver.cpp:751-755 */
    CCO_BADARG = 4
};

/* Problem parameters: the SimulationConfig keys the path reads
 * (SimulationConfig.hpp:73-137). */
typedef struct cco_problem {
    int32_t Ntau;
    int32_t scheme;          /* 1, 2 or 3 stages (SchemeIRK) */
    int32_t MaxIterIRK;
    int32_t NLeft, NRight;
    double  Dim;
    double  XLeft, XMid, XRight;
    double  PrecisionIRK;
    double  EpsNewton;
    double  Delta;           /* initial Delta used by the ctor of the ICG */
} cco_problem;

cco_ctx *cco_create(const cco_problem *p);
void     cco_destroy(cco_ctx *c);

/* Counters since the last reset: IRK steps taken, fixed-point iterations
 * (sum of itsReached, ODEStepper.cpp:202-204), RHS evaluations. */
void cco_counters(const cco_ctx *c, int64_t out[3]);
/* the Butcher tableau in use: returns the stage count s; a (s x s, row-major), b, c */
int cco_tableau(const cco_ctx *c, double *a, double *b, double *cnode);
void cco_reset_counters(cco_ctx *c);

/* --- SpectralTransformer (SpectralTransformer.cpp) ---------------------- */
/* Length of every transform is the ctx's Ntau. */
void cco_forward_r2c(cco_ctx *c, const double *in, double *out);     /* 57-72  */
void cco_backward_c2r(cco_ctx *c, const double *in, double *out);    /* 77-92  */
void cco_forward_c2c(cco_ctx *c, const double *in, double *out);     /* 97-112 */
void cco_backward_c2c(cco_ctx *c, const double *in, double *out);    /* 117-132 */
void cco_differentiate(cco_ctx *c, const double *in, double *out, size_t size, double period);   /* 140-158 */
void cco_lam_integrate(cco_ctx *c, const double *fk, double *out, size_t size,
                       double lam_re, double lam_im, double period);                       /* 165-189 */
void cco_halve_modes(const double *in, size_t n_in, double *out);   /* 196-217 */
void cco_double_modes(const double *in, size_t n_in, double *out);  /* 224-252 */
void cco_solve_inhom(cco_ctx *c, const double *f, const double *g, double *x, double period); /* 264-292 */

/* --- InitialConditionGenerator (InitialConditionGenerator.cpp) ---------- */
void cco_left_expansion(cco_ctx *c, double XLeft, const double *fc, const double *psic,
                        double *Y, double Delta);                    /* 37-239  */
void cco_right_expansion(cco_ctx *c, double XRight, const double *Up, double *Y,
                         double Delta);                               /* 252-667 (serial branch 456-624) */
void cco_pack(cco_ctx *c, const double *odd1, const double *odd2, const double *even,
              double *Z);                                             /* 678-711 */
void cco_unpack(cco_ctx *c, const double *Z, double *odd1, double *odd2, double *even); /* 719-758 */
void cco_fields_to_state(cco_ctx *c, const double *U, const double *V, const double *F,
                         double *Y);                                  /* 766-778 */
void cco_state_to_fields_d(cco_ctx *c, const double *Y, double *U, double *V, double *F,
                           double *IA2, double *dU, double *dV, double *dF, double X); /* 791-884 */
void cco_state_to_fields(cco_ctx *c, const double *Y, double *U, double *V, double *F); /* 891-932 */

/* --- ODEStepper (ODEStepper.cpp) ---------------------------------------- */
void cco_compute_derivatives(cco_ctx *c, const double *Yreal, double *dYreal, double x); /* 92-136 */
/* One IRK step; returns 1 if converged.  its_out = itsReached. */
int  cco_step_irk(cco_ctx *c, const double *Yin, double *Yout, double Xin, double Xout,
                  int maxIts, int *its_out);                          /* 163-292 */

/* --- ShootingSolver (ShootingSolver.cpp) -------------------------------- */
/* March Y (M complex, in/out) along grid from start to end (forward when
 * start < end).  Returns CCO_OK or CCO_IRK_NOCONV; *fail_idx = grid index. */
int cco_integrate(cco_ctx *c, double *Y, const double *grid, size_t start, size_t end,
                  int *fail_idx);                                     /* 53-145 */
int cco_mismatch(cco_ctx *c, const double *YL, const double *YR, double *mis); /* 154-182 */

/* --- NewtonSolver (NewtonSolver.cpp) ------------------------------------ */
void   cco_generate_grid(const cco_problem *p, double *grid);        /* 477-519 */
/* F(x): unpack -> series -> shoot -> mismatch -> pack (403-423).  x is n
 * doubles (left unchanged), F is n doubles. */
int    cco_shoot(cco_ctx *c, const double *grid, const double *x, double *F);
/* Same as cco_shoot but also returns the boundary states and the state at
 * the match point (each M complex); any pointer may be NULL. */
int    cco_shoot_states(cco_ctx *c, const double *grid, const double *x, double *F,
                        double *YL0, double *YR0, double *YLm, double *YRm);
/* FD Jacobian (615-702) with OpenMP over columns (the cc_openmp build,
 * 619-666).  J is row-major n x n: J[j*n+i] = dF_j/dx_i.  When cols != NULL
 * only the ncols listed columns are computed (others untouched).
 * status[i] receives the per-column status when non-NULL. */
int    cco_jacobian(const cco_problem *p, const double *grid, const double *x, const double *F0,
                    double *J, const int32_t *cols, int ncols, int nthreads, int32_t *status,
                    int64_t counters[3]);
/* dgesv on a row-major n x n system with one right-hand side
 * (739-757).  A is overwritten by its LU factors, b by the solution.
 * Returns LAPACK's info (0 ok, k>0 singular at k). */
int    cco_dgesv(int n, double *A, int32_t *ipiv, double *b);
double cco_l2norm(const double *v, size_t n);                        /* 763-768 */

#ifdef __cplusplus
}
#endif
#endif
