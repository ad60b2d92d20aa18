#!/usr/bin/env python3
"""Newton-step throughput of the HIP path on one MI355X.

One step = F(x) (FP64) + the n finite-difference Jacobian columns + the dense
solve of J dx = -F: the unit NewtonSolver::run times as NewtonStep_i
(NewtonSolver.cpp:281-365).  Workload: Ntau=1024, 8000 x-points (NLeft =
NRight = 4000), XMid=0.5, Dim=4, seeded anchored input (workload.py), inputs
resident in HBM before the timed region.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.  value = Newton steps/s of the mixed-precision policy
(FP32 tangent-linear columns and LU; FP64 residual, Delta column and
refinement); the FP64 form (every column a forward difference of two FP64
shoots, as the reference) is timed in the same run and reported beside it,
with and without reuse of the unchanged side.  Each policy is autotuned
first (outside the timed region).  Beside the headline the same line carries:
  roofline / roofline_base   the dominant kernel (the FP32 tangent chains) and
                             the critical-path FP64 base chains
  oracle_parity              the Jacobian columns of both policies against the
                             oracle's (the columns the CPU baseline computes)
                             and the Newton steps against the oracle's step
                             (tests/golden/step_c3_oracle.json)
  newton                     Newton to convergence under both policies: the
                             converged Delta and its agreement
  fp64_ntau512_x6000         the plain FP64 form at Ntau=512, 6000 x-points
  batch                      the 24 batched uneven problems
  cpu_baseline               the oracle on the host cores (kind "port")
--no-extras leaves out newton, fp64_ntau512_x6000 and batch.
"""
import argparse
import importlib
import json
import math
import os
import sys
import time
from dataclasses import replace

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


# HBM traffic per launch of the dominant kernel, from two rocprofv3 --pmc
# passes (FETCH_SIZE, WRITE_SIZE) scaled by the same counters on a plain
# device copy; see profiles/hbm_traffic_newton_step_ntau1024_x8000.md.  Keyed
# by the kernel the bench reports and the steps per launch it was measured at
# (the record it reads grows with them); anything else reports null.
MEASURED_TRAFFIC_BYTES = {
    ("k_tangent<1024,256>", 8): (17.49e6 + 11.64e6),
    ("k_tangent<1024,256>", 12): (20.09e6 + 11.64e6),
    ("k_tangent<1024,256>", 16): (22.72e6 + 11.64e6),
}

FP64_PEAK, FP32_PEAK, CUS = 78.6, 157.3, 256
DF = 1e-14          # bound on the per-entry difference of two FP64 shoots (different FFT rounding)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def flops_chain(N, its, steps, tangent=False):
    """Algorithmic FLOPs of the chain kernels: per fixed-point iteration
    2 F_RHS + 18 N (two RHS evaluations and the stage update), per IRK step
    6 N (final combination); F_RHS = 25 N log2 N + 100 N (3 complex + 4 real
    length-N transforms and the pointwise work).  A tangent-linear RHS has
    2 complex + 2 paired-real transforms per stage: F_TRHS = 20 N log2 N + 100 N."""
    frhs = (20 if tangent else 25) * N * math.log2(N) + 100 * N
    return its * (2 * frhs + 18 * N) + steps * 6 * N


def hbm_bytes_step(shape):
    """Algorithmic HBM bytes of one Newton step: the residual rows and the
    Jacobian written and read back, the grid and the x/F0/dx vectors."""
    n = shape.n
    return 8.0 * (n * (n + 1) + n * n + shape.X + 3 * n)


def time_policy(cc, dev, shape, precision, steps, warmup, reuse, tune, eps=None):
    """Autotune (outside the timed region), then time `steps` Newton steps on
    device-resident x, F0, dx.  Also returns the Jacobian at x.  eps: an FD
    step other than the shape's EpsNewton."""
    n = shape.n
    x = importlib.import_module("parallel-critical-collapse_amd.workload").initial_x(shape)
    if eps is not None:
        shape = replace(shape, EpsNewton=eps)
    ctx = cc.Context(cc.Problem.from_shape(shape, precision=precision, reuse_sides=int(reuse)))
    name = {cc.CC_MIXED: "mixed", cc.CC_FP64: "fp64", cc.CC_FP64_TANGENT: "fp64_tangent",
            cc.CC_FP32_FD: "fp32_fd"}[precision] + ("+reuse" if reuse else "")
    tangent = precision in (cc.CC_MIXED, cc.CC_FP64_TANGENT)
    tuned = None
    if isinstance(tune, dict):
        ctx.set_schedule(**tune)
    elif tune:
        t = time.perf_counter()
        sc, times = ctx.autotune(x)
        tuned = {"schedule": sc.as_dict(), "candidate_ms": [round(float(v), 3) for v in times],
                 "tuning_s": round(time.perf_counter() - t, 2)}
        log(f"[bench] {name} autotune {tuned['tuning_s']} s -> {sc.as_dict()}")
    sched = ctx.schedule().as_dict()
    d_x, d_F0, d_dx = dev.DeviceArray(x), dev.DeviceArray(np.zeros(n)), dev.DeviceArray(np.zeros(n))
    # events on the tangent / bulk launches only: the base chains' stream is
    # timed on one more step after the timed ones
    ctx.profile(not sched["use_graph"], base=False)
    for w in range(warmup):
        t = time.perf_counter()
        ctx.newton_step_device(d_x.ptr, d_F0.ptr, d_dx.ptr)
        log(f"[bench] {name} warmup {w + 1}/{warmup}: {time.perf_counter() - t:.3f} s")
    dev.synchronize()
    chain = np.zeros(4)
    base = np.zeros(4)
    t0 = time.perf_counter()
    for k in range(steps):
        ctx.newton_step_device(d_x.ptr, d_F0.ptr, d_dx.ptr)
        if not sched["use_graph"]:
            chain += ctx.chain_timing()
        log(f"[bench] {name} step {k + 1}/{steps} done")
    dev.synchronize()
    wall = time.perf_counter() - t0
    if sched["use_graph"]:
        # the graph carries no timing events: time the chain kernels on one eager step
        ctx.set_schedule(use_graph=0)
        ctx.profile(True)
        ctx.newton_step_device(d_x.ptr, d_F0.ptr, d_dx.ptr)
        chain = np.array(ctx.chain_timing()) * steps
        if tangent:
            base = np.array(ctx.base_timing()) * steps
    elif tangent:
        ctx.profile(True)
        ctx.newton_step_device(d_x.ptr, d_F0.ptr, d_dx.ptr)
        base = np.array(ctx.base_timing()) * steps
    F0, dx = d_F0.download(), d_dx.download()
    J = ctx.last_jacobian()
    refine = ctx.refinement() if precision in (cc.CC_MIXED, cc.CC_FP32_FD) else None
    ctx.close()
    return {
        "wall_s": wall, "steps_per_s": steps / wall, "ms_per_step": 1e3 * wall / steps,
        "chain_ms": chain[0], "launches": int(chain[1]), "its": int(chain[2]), "irk_steps": int(chain[3]),
        "kbar": chain[2] / max(chain[3], 1), "base": base, "F0": F0, "dx": dx, "J": J, "refine": refine,
        "schedule": sched, "tuning": tuned,
    }


_HOST = None


def host_info():
    """Host threads available to this process and the CPU model, read once:
    with OMP_PROC_BIND=close the OpenMP runtime pins the calling thread to one
    core when it starts, so later reads of the affinity mask see one CPU."""
    global _HOST
    if _HOST is not None:
        return _HOST
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    _HOST = (threads, model)
    return _HOST


def cpu_baseline(shape, threads, ncols, reps=3):
    """The oracle (CPU restatement, OpenMP over Jacobian columns like the
    reference's cc_openmp build, OMP_PROC_BIND=close) on a bounded sample:
    `reps` repetitions (the reference's default, main.cpp:35) over disjoint
    sets of ncols columns each (rounded up to a multiple of the threads),
    spread over the Up, psic, fc slots and Delta.  value = the median
    repetition's columns/s scaled to Newton steps/s by (n+1).  Also returns
    the oracle's columns for the parity record."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    n = shape.n
    cfg = O.Config(Ntau=shape.Ntau, NLeft=shape.NLeft, NRight=shape.NRight, XMid=shape.XMid,
                   XLeft=shape.XLeft, XRight=shape.XRight, Dim=shape.Dim, Delta=shape.Delta,
                   EpsNewton=shape.EpsNewton, PrecisionIRK=shape.PrecisionIRK, MaxIterIRK=shape.MaxIterIRK)
    W = importlib.import_module("parallel-critical-collapse_amd.workload")
    x = W.initial_x(shape)
    k = -(-max(ncols, threads) // threads) * threads
    allc = sorted(set(int(round(v)) for v in np.linspace(0, n - 1, reps * k)) | {shape.Ntau // 2 + 2})
    sets = [allc[r::reps][:k] for r in range(reps)]
    F0, st = O.Oracle(cfg).shoot(x)
    J = np.zeros((n, n))
    rates, secs, worst_all = [], [], 0
    for r, cols in enumerate(sets):
        t = time.perf_counter()
        Jr, status, cnt, worst = O.jacobian(cfg, x, F0, cols=cols, nthreads=threads)
        dt = time.perf_counter() - t
        J[:, cols] = Jr[:, cols]
        rates.append(len(cols) / dt)
        secs.append(round(dt, 2))
        worst_all = max(worst_all, worst)
        log(f"[bench] cpu baseline repetition {r + 1}/{reps}: {len(cols)} columns in {dt:.1f} s")
    med = float(np.median(rates))
    hthreads, model = host_info()
    return {
        "value": med / (n + 1), "unit": "Newton steps/s", "cores": threads, "kind": "port",
        "sample": f"{reps} repetitions of {k} FD Jacobian columns each (disjoint sets spread over Up, psic, fc "
                  f"and Delta; full shoots at Ntau={shape.Ntau}, {shape.X} x-points) on {threads} OpenMP threads "
                  f"(OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}); seconds {secs}; median repetition's "
                  f"columns/s extrapolated to steps/s = columns/s / (n+1), n={n}",
        "host_threads": hthreads, "cpu_model": model,
        "status_ok": bool(worst_all == 0 and st == 0),
    }, {"F0": F0, "J": J, "cols": sorted(c for s in sets for c in s)}


def column_parity(J, Jo, cols, shape):
    """Jacobian columns against the oracle's: Frobenius over the sample, and
    per column relative to max(|column|, 1e-4) (the weakest psic columns
    carry the finite differences' own noise)."""
    err = np.array([np.linalg.norm(J[:, c] - Jo[:, c]) for c in cols])
    nrm = np.array([np.linalg.norm(Jo[:, c]) for c in cols])
    e = err / np.maximum(nrm, 1e-4)
    noise = 2.0 * math.sqrt(shape.n) * DF / shape.EpsNewton
    return {"columns": len(cols), "frobenius_rel": float(np.sqrt(np.sum(err ** 2) / np.sum(nrm ** 2))),
            "column_rel_median": float(np.median(e)), "column_rel_max": float(e.max()),
            "max_ratio_to_fd_bound": float(np.max(err / (1e-5 * nrm + noise)))}


def step_parity(dx, dxo, N):
    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))
    return {"all": rel(dx, dxo), "Up": rel(dx[:N // 4], dxo[:N // 4]),
            "psic": rel(dx[N // 4:N // 2], dxo[N // 4:N // 2]), "fc": rel(dx[N // 2:], dxo[N // 2:])}


def rms(v):
    """computeL2Norm (NewtonSolver.cpp:763-768)."""
    return float(np.sqrt(np.mean(np.asarray(v) ** 2)))


def newton_study(cc, shape, policies, tol, max_its, tune, save_failures=None, schedules=None):
    """Newton from the seeded input to mismatch < tol under each precision
    policy, with the reference's damping fac = min(1, SlowError/err) and its
    divergence guard (NewtonSolver.cpp:298-358): iterations, time per
    iteration, converged Delta, and Delta's relative difference between
    policies.  Delta is that of the converged iterate, or of the iterate with
    the smallest mismatch when a policy stops short of tol (the FP64 finite
    differences stall near 1e-12 on their columns' noise), with that mismatch
    reported beside it.  schedules: policy name -> schedule fields to install
    instead of autotuning."""
    W = importlib.import_module("parallel-critical-collapse_amd.workload")
    # fp32_fd: FP32 shoots at FD step 1e-3 (the reference's 1e-8 is below their rounding)
    names = {"fp64": cc.CC_FP64, "mixed": cc.CC_MIXED, "fp64_tangent": cc.CC_FP64_TANGENT, "fp32_fd": cc.CC_FP32_FD}
    dslot = shape.Ntau // 2 + 2
    out = {}
    for name in policies:
        sh = replace(shape, EpsNewton=1e-3) if name == "fp32_fd" else shape
        ctx = cc.Context(cc.Problem.from_shape(sh, precision=names[name]))
        x = W.initial_x(shape)
        if schedules and name in schedules:
            ctx.set_schedule(**schedules[name])
        elif tune and name != "fp64":   # FP64 finite differences keep the default schedule
            ctx.autotune(x)
        hist, secs = [], []
        err, err_old, increase, reached, failure = 1.0, 1.0, 0, None, None
        best_err, best_delta = None, None
        for its in range(max_its):
            t = time.perf_counter()
            try:
                F0, dx = ctx.newton_step(x)
            except cc.CCError as e:
                failure = {"iteration": its + 1, "error": str(e)}
                log(f"[newton] {name} iteration {its + 1}: {e}")
                if save_failures:
                    np.save(os.path.join(save_failures, f"newton_fail_{name}_N{shape.Ntau}.npy"), x)
                break
            err_old, err = err, rms(F0)
            hist.append(err)
            if best_err is None or err < best_err:
                best_err, best_delta = err, float(x[dslot])
            log(f"[newton] {name} iteration {its + 1}: mismatch {err:.3e}")
            if err < tol:
                reached = its
                secs.append(time.perf_counter() - t)
                break
            increase = increase + 1 if err > err_old else 0
            if increase >= 5:
                secs.append(time.perf_counter() - t)
                break
            x = x + min(1.0, shape.SlowError / err) * dx
            secs.append(time.perf_counter() - t)
        below = [i + 1 for i, h in enumerate(hist) if h < 1e-10]
        out[name] = {"iterations_to_tol": reached, "evaluations_to_1e-10": below[0] if below else None,
                     "mismatch": hist, "final_mismatch": hist[-1] if hist else None,
                     "Delta": best_delta, "Delta_at_mismatch": best_err,
                     "s_per_iteration": float(np.mean(secs)) if secs else None,
                     "s_iterations": [round(v, 4) for v in secs], "total_s": float(np.sum(secs)),
                     "failure": failure, "schedule": ctx.schedule().as_dict()}
        ctx.close()
    for ref in ("fp64_tangent", "fp64"):
        if ref in out and out[ref]["Delta"] is not None:
            for name, r in out.items():
                if r["Delta"] is not None:
                    r[f"Delta_rel_diff_vs_{ref}"] = abs(r["Delta"] - out[ref]["Delta"]) / abs(out[ref]["Delta"])
    return out


def batch_bench(cc, steps, warmup, tune):
    """The batched uneven problems: every problem of the dimension sweep
    stepped together through cc_batch (one launch per Ntau class and chunk),
    against the same problems stepped one after another on their own
    contexts; per-problem Jacobian error of the mixed policy against the
    FP64 tangent-linear Jacobian."""
    W = importlib.import_module("parallel-critical-collapse_amd.workload")
    shapes = W.config4()
    xs = [W.initial_x(sh) for sh in shapes]
    b = cc.Batch([cc.Problem.from_shape(sh, precision=cc.CC_MIXED) for sh in shapes])
    tuning = None
    if tune:
        t = time.perf_counter()
        times = b.autotune(xs)
        tuning = {"candidate_ms": [round(float(v), 3) for v in times], "tuning_s": round(time.perf_counter() - t, 2),
                  "schedules": {b.schedule(c)[0]: b.schedule(c)[1].as_dict() for c in range(b.nclasses)}}
        log(f"[bench] batch autotune {tuning['tuning_s']} s")
    for w in range(warmup):
        b.newton_step(xs)
    t0 = time.perf_counter()
    for k in range(steps):
        F0b, dxb, status, st = b.newton_step(xs, check=False)
        log(f"[bench] batch step {k + 1}/{steps}: worst status {st}")
    wall = time.perf_counter() - t0
    cnt = b.counters()
    b.close()
    # the same problems one after another, each on its own context
    seq = 0.0
    errs = []
    for i, sh in enumerate(shapes):
        ctx = cc.Context(cc.Problem.from_shape(sh, precision=cc.CC_MIXED))
        ctx.newton_step(xs[i])
        t = time.perf_counter()
        F0, dx = ctx.newton_step(xs[i])
        seq += time.perf_counter() - t
        J32, _, _ = ctx.jacobian(xs[i], F0)
        ctx.close()
        c64 = cc.Context(cc.Problem.from_shape(sh, precision=cc.CC_FP64_TANGENT))
        F64, dx64 = c64.newton_step(xs[i])
        J64, _, _ = c64.jacobian(xs[i], F64)
        c64.close()
        N = sh.Ntau
        right, left = slice(0, N // 4), slice(N // 4, 3 * N // 4)     # Up columns / psic, fc columns

        def blk(sl):
            return float(np.linalg.norm(J32[:, sl] - J64[:, sl]) / np.linalg.norm(J64[:, sl]))
        errs.append({"Ntau": sh.Ntau, "x_points": sh.X, "Dim": sh.Dim, "scale": sh.scale,
                     "status": int(status[i]), "mismatch": rms(F64),
                     "F0_equal_fp64": bool(np.array_equal(F0b[i], F64)),
                     "J_rel_err": float(np.linalg.norm(J32 - J64) / np.linalg.norm(J64)),
                     "J_rel_err_right_side_columns": blk(right), "J_rel_err_left_side_columns": blk(left),
                     "dx_rel_err": float(np.linalg.norm(dxb[i] - dx64) / np.linalg.norm(dx64))})
        log(f"[bench] problem {i}: {errs[-1]}")
    n = len(shapes)
    return {
        "metric": f"Jacobian+Newton steps/s, {n} uneven problems batched (mixed FP32/FP64)",
        "value": round(n * steps / wall, 4), "unit": "Newton steps/s (all problems)", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": round(1e3 * wall / steps, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic: shipped D=4 solutions + seeded perturbations, D sweep, Ntau/x-points by seed",
        "config": {"workload": "batch of " + ", ".join(f"(Ntau={sh.Ntau}, X={sh.X}, D={sh.Dim})" for sh in shapes)},
        "sequential_value": round(n / seq, 4), "batched_over_sequential": round((n * steps / wall) / (n / seq), 3),
        "chains_per_step": int(cnt[2]), "tuning": tuning, "problems": errs,
        "J_rel_err_max": max(e["J_rel_err"] for e in errs),
        "J_rel_err_median": float(np.median([e["J_rel_err"] for e in errs])),
        "J_rel_err_left_side_columns_max": max(e["J_rel_err_left_side_columns"] for e in errs),
        "F0_equal_fp64_all": all(e["F0_equal_fp64"] for e in errs),
    }


def fp64_leg(cc, dev, shape, steps, warmup, tune):
    """The plain FP64 form at one shape: Jacobian builds (Newton steps) per
    second, achieved FP64 TFLOP/s of the chain kernels, algorithmic HBM GB/s."""
    r = time_policy(cc, dev, shape, cc.CC_FP64, steps, warmup, False, tune)
    fl = flops_chain(shape.Ntau, r["its"], r["irk_steps"])
    ach = fl / (r["chain_ms"] / 1e3) / 1e12
    return {"workload": f"newton_step Ntau={shape.Ntau} x_points={shape.X} n={shape.n} FP64 finite differences",
            "jacobian_builds_per_s": round(r["steps_per_s"], 4), "ms_per_step": round(r["ms_per_step"], 2),
            "kernel": f"k_chain<{shape.Ntau},{r['schedule']['threads_bulk']},double,float,0,2>",
            "achieved_fp64_tflops": round(ach, 3), "peak": FP64_PEAK, "frac": round(ach / FP64_PEAK, 4),
            "hbm_gbps_algorithmic": round(hbm_bytes_step(shape) / (r["ms_per_step"] / 1e3) / 1e9, 3),
            "kbar": round(r["kbar"], 3), "schedule": r["schedule"]}


def fp32_fd_leg(cc, dev, shape, steps, tune, ocols, dxo, mixed_steps_per_s):
    """The reference's finite differences with every shoot in FP32
    (CC_FP32_FD) at this shape, at FD steps h = 1e-3 and 1e-4: time per Newton
    step (autotuned at h = 1e-3), its Jacobian against the oracle's sampled
    columns and its step against the oracle's step; then Newton from the seeded
    input with h = 1e-3 and the reference's damping, to mismatch 1e-12."""
    N = shape.Ntau
    out = {"policy": "CC_FP32_FD: every column (F32(x + h e_j) - F32(x)) / h of FP32 shoots; "
                     "F(x) FP64; FP32 LU + FP64 refinement", "legs": {}}
    for h in (1e-3, 1e-4):
        r = time_policy(cc, dev, shape, cc.CC_FP32_FD, steps, 1, False, tune if h == 1e-3 else False, eps=h)
        leg = {"ms_per_step": round(r["ms_per_step"], 2), "steps_per_s": round(r["steps_per_s"], 4),
               "refinement": {"residual": r["refine"][0], "fp64_resolve": r["refine"][1]},
               "schedule": r["schedule"]}
        if ocols is not None:
            leg["columns_vs_oracle"] = column_parity(r["J"], ocols["J"], ocols["cols"], replace(shape, EpsNewton=h))
        if dxo is not None:
            leg["dx_vs_oracle_step"] = step_parity(r["dx"], dxo, N)
        out["legs"][f"h={h:g}"] = leg
        if h == 1e-3:
            out["mixed_over_fp32_fd"] = round(mixed_steps_per_s / r["steps_per_s"], 3)
    out["newton_h1e-3"] = newton_study(cc, shape, ["fp32_fd"], 1e-12, 12, False)["fp32_fd"]
    return out


def copy_calibration(nbytes):
    """A plain device copy of nbytes (torch's copy kernel), for scaling the
    TCC FETCH_SIZE / WRITE_SIZE counters of a --pmc run to bytes."""
    import torch
    n = nbytes // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 4, "read": n * 4, "written": n * 4}), flush=True)


def guarded(name, fn):
    """An extra leg of the record: its result, or the error that stopped it."""
    try:
        t = time.perf_counter()
        r = fn()
        log(f"[bench] {name}: {time.perf_counter() - t:.1f} s")
        return r
    except Exception as e:      # the headline line is printed regardless
        log(f"[bench] {name} failed: {e!r}")
        return {"error": repr(e)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=3,
                    help="workload shape 1, 2, 3 or 5; 4 = the batch of uneven problems")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 policy")
    ap.add_argument("--fp64-only", action="store_true",
                    help="time only the FP64 finite-difference policy and print its record (for profiling runs)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the Newton-to-convergence, Ntau=512 and batch legs")
    ap.add_argument("--no-tune", action="store_true", help="default schedule instead of the autotuner's")
    ap.add_argument("--schedule", default="", help='fixed schedule fields as JSON, e.g. \'{"threads_tangent": 128}\'; '
                    "skips the autotuner")
    ap.add_argument("--cpu-threads", type=int, default=0, help="host threads of the CPU baseline (0: all available)")
    ap.add_argument("--cpu-columns", type=int, default=32,
                    help="Jacobian columns per repetition of the CPU baseline (rounded up to a multiple of the threads)")
    ap.add_argument("--newton", action="store_true",
                    help="Newton to convergence under each of --policies instead of the step benchmark")
    ap.add_argument("--policies", default="mixed,fp64_tangent,fp64")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--max-its", type=int, default=12)
    ap.add_argument("--scale", type=float, default=None,
                    help="--newton: perturbation scale of the seeded input instead of the workload's")
    ap.add_argument("--save-failures", default="", metavar="DIR",
                    help="--newton: save the Newton vector at which a policy failed to DIR")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device copy of BYTES (counter calibration) and exit")
    args = ap.parse_args()
    host_info()                 # before any OpenMP runtime starts
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the column shards of a multi-GPU run are described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    cc = importlib.import_module("parallel-critical-collapse_amd")
    dev = importlib.import_module("parallel-critical-collapse_amd.device")
    W = importlib.import_module("parallel-critical-collapse_amd.workload")
    if args.newton:
        shape = W.config(args.config)
        if args.scale is not None:
            shape = replace(shape, scale=args.scale)
        res = newton_study(cc, shape, args.policies.split(","), args.tol, args.max_its, not args.no_tune,
                           args.save_failures or None)
        print(json.dumps({"workload": f"newton Ntau={shape.Ntau} x_points={shape.X} n={shape.n} Dim={shape.Dim} "
                                      f"XMid={shape.XMid} scale={shape.scale}", "tol": args.tol,
                          "policies": res}), flush=True)
        return
    if args.config == 4:
        print(json.dumps(batch_bench(cc, args.steps, args.warmup, not args.no_tune)), flush=True)
        return
    shape = W.config(args.config)
    log(f"[bench] workload Ntau={shape.Ntau} X={shape.X} n={shape.n}")

    tune = json.loads(args.schedule) if args.schedule else (not args.no_tune)
    if args.fp64_only:
        r = fp64_leg(cc, dev, shape, args.steps, args.warmup, tune)
        print(json.dumps(dict(r, steps=args.steps, warmup=args.warmup)), flush=True)
        return
    mixed = time_policy(cc, dev, shape, cc.CC_MIXED, args.steps, args.warmup, False, tune)
    fp64 = None if args.no_fp64 else time_policy(cc, dev, shape, cc.CC_FP64, args.steps, args.warmup, False, tune)
    fp64r = None if args.no_fp64 else time_policy(cc, dev, shape, cc.CC_FP64, args.steps, args.warmup, True, tune)

    N = shape.Ntau
    T = mixed["schedule"]["threads_tangent"]
    fl = flops_chain(N, mixed["its"], mixed["irk_steps"], tangent=True)
    achieved = fl / (mixed["chain_ms"] / 1e3) / 1e12
    roof = {"bound": "valu", "pipe": "fp32 valu, non-packed (FFT passes + pointwise; no dense contraction)",
            "kernel": f"k_tangent<{N},{T}>", "achieved": round(achieved, 3), "peak": FP32_PEAK,
            "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK, 4),
            "traffic": MEASURED_TRAFFIC_BYTES.get((f"k_tangent<{N},{T}>", mixed["schedule"]["steps_per_launch"])),
            "flops_per_step": fl / args.steps, "kernel_s_per_step": mixed["chain_ms"] / 1e3 / args.steps,
            "avg_launch_ms": mixed["chain_ms"] / max(mixed["launches"], 1),
            "launches_per_step": mixed["launches"] / args.steps}
    bms, bl, bits, bsteps = mixed["base"]
    base_cus = 4                    # base and Delta column, both sides: one workgroup (one CU) each
    roof_base = None
    if bms > 0:
        bfl = flops_chain(N, bits, bsteps)
        bach = bfl / (bms / 1e3) / 1e12
        roof_base = {"bound": "valu", "pipe": "fp64 valu; latency of dependent transform passes on 4 workgroups",
                     "kernel": f"k_chain<{N},{mixed['schedule']['threads_base']},double,float,1,2>",
                     "achieved": round(bach, 4), "peak": FP64_PEAK, "unit": "TFLOP/s",
                     "frac": round(bach / FP64_PEAK, 5), "cus": base_cus,
                     "frac_of_its_cus": round(bach / (FP64_PEAK * base_cus / CUS), 4),
                     "kernel_ms_per_step": round(bms / args.steps, 2), "avg_launch_ms": bms / max(bl, 1),
                     "launches_per_step": bl / args.steps, "kbar": round(bits / max(bsteps, 1), 3),
                     "flops_per_step": bfl / args.steps}
    out = {
        "metric": "Jacobian+Newton steps/s at Ntau=1024, 8000 x-pts (mixed FP32/FP64)",
        "value": round(mixed["steps_per_s"], 5), "unit": "Newton steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(mixed["ms_per_step"], 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": f"synthetic: shipped D=4 solution (data/simulation_4D_{N}.json) + seeded N(0,1)*0.1/k^2 "
                f"odd/even-harmonic perturbation scaled by {shape.scale} (the largest power of ten <= 1 at "
                f"which the oracle's FP64 shoot survives)",
        "config": {"workload": f"newton_step Ntau={shape.Ntau} x_points={shape.X} n={shape.n} Dim={shape.Dim} "
                               f"XMid={shape.XMid} IRK2", "policy": "mixed"},
        "schedule": mixed["schedule"], "tuning": mixed["tuning"],
        "kbar_tangent": round(mixed["kbar"], 3),
        "refinement": {"residual": mixed["refine"][0], "fp64_resolve": mixed["refine"][1]},
        "roofline": roof, "roofline_base": roof_base,
    }
    for key, res in (("fp64", fp64), ("fp64_reuse_sides", fp64r)):
        if res is None:
            continue
        fl64 = flops_chain(N, res["its"], res["irk_steps"])
        ach = fl64 / (res["chain_ms"] / 1e3) / 1e12
        out[key] = {"value": round(res["steps_per_s"], 5), "ms_per_step": round(res["ms_per_step"], 2),
                    "kbar": round(res["kbar"], 3),
                    "kernel": f"k_chain<{N},{res['schedule']['threads_bulk']},double,float,0,2>",
                    "achieved_tflops": round(ach, 3), "peak": FP64_PEAK, "frac": round(ach / FP64_PEAK, 4),
                    "avg_launch_ms": res["chain_ms"] / max(res["launches"], 1),
                    "launches_per_step": res["launches"] / args.steps, "schedule": res["schedule"]}
        out[f"mixed_over_{key}"] = round(mixed["steps_per_s"] / res["steps_per_s"], 3)

    # parity against the oracle at this shape: its step (fixture) and its columns (CPU leg)
    par = {}
    ocols, dxo = None, None
    fxp = os.path.join(GOLDEN, f"step_c{args.config}_oracle.json")
    if os.path.exists(fxp):
        with open(fxp) as f:
            fx = json.load(f)
        if fx["shape"]["Ntau"] == N and fx["shape"]["x_points"] == shape.X and fx["shape"]["seed"] == shape.seed:
            dxo = np.array(fx["dx"])
            par["dx_vs_oracle_step"] = {"mixed": step_parity(mixed["dx"], dxo, N)}
            par["F0_vs_oracle_max_abs"] = float(np.max(np.abs(mixed["F0"] - np.array(fx["F0"]))))
            par["oracle_cond_J"] = fx["sigma_max"] / fx["sigma_min"]
            if fp64 is not None:
                par["dx_vs_oracle_step"]["fp64"] = step_parity(fp64["dx"], dxo, N)
    for key, res in (("mixed", mixed), ("fp64", fp64)):
        if res is not None:
            par.setdefault("step_residual", {})[key] = float(np.linalg.norm(res["J"] @ res["dx"] + res["F0"]) /
                                                             np.linalg.norm(res["F0"]))
    if not args.no_cpu:
        threads = args.cpu_threads or host_info()[0]
        log(f"[bench] cpu baseline on {threads} threads")
        out["cpu_baseline"], ocols = cpu_baseline(shape, threads, args.cpu_columns)
        par["columns_vs_oracle"] = {"mixed": column_parity(mixed["J"], ocols["J"], ocols["cols"], shape)}
        if fp64 is not None:
            par["columns_vs_oracle"]["fp64"] = column_parity(fp64["J"], ocols["J"], ocols["cols"], shape)
    out["oracle_parity"] = par
    if "columns_vs_oracle" in par:
        out["J_rel_err_vs_oracle"] = {k: v["frobenius_rel"] for k, v in par["columns_vs_oracle"].items()}
    if "dx_vs_oracle_step" in par:
        out["dx_rel_err_vs_oracle"] = {k: v["all"] for k, v in par["dx_vs_oracle_step"].items()}

    if not args.no_extras:
        scheds = {"mixed": mixed["schedule"]}
        if fp64 is not None:
            scheds["fp64"] = fp64["schedule"]
        out["newton"] = guarded("newton to convergence", lambda: dict(
            tol=1e-12, **{"policies": newton_study(cc, shape, ["mixed", "fp64"] if fp64 is not None else ["mixed"],
                                                   1e-12, 12, False, schedules=scheds)}))
        out["fp32_fd"] = guarded("fp32 finite differences", lambda: fp32_fd_leg(
            cc, dev, shape, max(2, args.steps // 4), tune, ocols, dxo, mixed["steps_per_s"]))
        out["fp64_ntau512_x6000"] = guarded("fp64 at Ntau=512", lambda: fp64_leg(
            cc, dev, W.config(2), max(2, args.steps // 4), 1, tune))
        if "error" not in out["fp64_ntau512_x6000"] and not args.no_cpu:
            sh2 = W.config(2)
            cb2, _ = cpu_baseline(sh2, args.cpu_threads or host_info()[0], 16)
            out["fp64_ntau512_x6000"]["cpu_baseline"] = cb2
            out["fp64_ntau512_x6000"]["gpu_over_cpu"] = round(
                out["fp64_ntau512_x6000"]["jacobian_builds_per_s"] / cb2["value"], 1)
        bb = guarded("batch", lambda: batch_bench(cc, 3, 1, True))
        if "error" not in bb:
            bb = {k: bb[k] for k in ("value", "unit", "ms_per_step", "sequential_value", "batched_over_sequential",
                                     "chains_per_step", "J_rel_err_max", "J_rel_err_median",
                                     "J_rel_err_left_side_columns_max", "F0_equal_fp64_all", "config")}
        out["batch"] = bb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
