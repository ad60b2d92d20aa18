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
first (outside the timed region).
"""
import argparse
import importlib
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


# HBM traffic per launch of the dominant kernel, from two rocprofv3 --pmc
# passes (FETCH_SIZE, WRITE_SIZE) scaled by the same counters on a plain
# device copy; see profiles/hbm_traffic_newton_step_ntau1024_x8000.md.  Keyed
# by the kernel the bench reports and the steps per launch it was measured at
# (the record it reads grows with them); anything else reports null.
MEASURED_TRAFFIC_BYTES = {
    ("k_tangent<1024,256>", 64): (59.59e6 + 24.53e6),
}


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


def time_policy(cc, dev, shape, precision, steps, warmup, reuse, tune):
    """Autotune (outside the timed region), then time `steps` Newton steps on
    device-resident x, F0, dx."""
    n = shape.n
    ctx = cc.Context(cc.Problem.from_shape(shape, precision=precision, reuse_sides=int(reuse)))
    x = importlib.import_module("parallel-critical-collapse_amd.workload").initial_x(shape)
    name = ("mixed" if precision == cc.CC_MIXED else "fp64") + ("+reuse" if reuse else "")
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
    ctx.profile(not sched["use_graph"])
    for w in range(warmup):
        t = time.perf_counter()
        ctx.newton_step_device(d_x.ptr, d_F0.ptr, d_dx.ptr)
        log(f"[bench] {name} warmup {w + 1}/{warmup}: {time.perf_counter() - t:.3f} s")
    dev.synchronize()
    chain_ms, launches, its, nsteps = 0.0, 0, 0, 0
    t0 = time.perf_counter()
    for k in range(steps):
        ctx.newton_step_device(d_x.ptr, d_F0.ptr, d_dx.ptr)
        if not sched["use_graph"]:
            ms, nl, it, st = ctx.chain_timing()
            chain_ms += ms
            launches += nl
            its += it
            nsteps += st
        log(f"[bench] {name} step {k + 1}/{steps} done")
    dev.synchronize()
    wall = time.perf_counter() - t0
    if sched["use_graph"]:
        # the graph carries no timing events: time the chain kernels on one eager step
        ctx.set_schedule(use_graph=0)
        ctx.profile(True)
        ctx.newton_step_device(d_x.ptr, d_F0.ptr, d_dx.ptr)
        ms, nl, it, st = ctx.chain_timing()
        chain_ms, launches, its, nsteps = ms * steps, nl * steps, it * steps, st * steps
    F0, dx = d_F0.download(), d_dx.download()
    ctx.close()
    return {
        "wall_s": wall, "steps_per_s": steps / wall, "ms_per_step": 1e3 * wall / steps,
        "chain_ms": chain_ms, "launches": launches, "its": its, "irk_steps": nsteps,
        "kbar": its / max(nsteps, 1), "F0_rms": float(np.sqrt(np.mean(F0 ** 2))), "dx": dx,
        "schedule": sched, "tuning": tuned,
    }


def cpu_baseline(shape, threads, ncols):
    """The oracle (CPU restatement, OpenMP over Jacobian columns like the
    reference's cc_openmp build) on a bounded sample of ncols columns (rounded
    up to a multiple of the thread count), spread over the Up, psic, fc slots
    and Delta; scaled to Newton steps/s by (n+1)/ncols."""
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
    cols = sorted(set(int(round(v)) for v in np.linspace(0, n - 1, k)) | {shape.Ntau // 2 + 2})[:k]
    F0 = np.zeros(n)
    t = time.perf_counter()
    J, status, cnt, worst = O.jacobian(cfg, x, F0, cols=cols, nthreads=threads)
    dt = time.perf_counter() - t
    cols_per_s = len(cols) / dt
    return {
        "value": cols_per_s / (n + 1), "unit": "Newton steps/s", "cores": threads, "kind": "port",
        "sample": f"{len(cols)} FD Jacobian columns spread over Up, psic, fc and Delta (full shoots, "
                  f"Ntau={shape.Ntau}, {shape.X} x-points) in {dt:.1f} s on {threads} threads; "
                  f"extrapolated: steps/s = columns/s / (n+1), n={n}",
        "status_ok": bool(worst == 0), "kbar": float(cnt[1] / max(cnt[0], 1)),
    }


def rms(v):
    """computeL2Norm (NewtonSolver.cpp:763-768)."""
    return float(np.sqrt(np.mean(np.asarray(v) ** 2)))


def newton_study(cc, shape, policies, tol, max_its, tune, save_failures=None):
    """Newton from the seeded input to mismatch < tol under each precision
    policy, with the reference's damping fac = min(1, SlowError/err) and its
    divergence guard (NewtonSolver.cpp:298-358): iterations, time per
    iteration, converged Delta, and Delta's relative difference between
    policies."""
    W = importlib.import_module("parallel-critical-collapse_amd.workload")
    names = {"fp64": cc.CC_FP64, "mixed": cc.CC_MIXED, "fp64_tangent": cc.CC_FP64_TANGENT}
    dslot = shape.Ntau // 2 + 2
    out = {}
    for name in policies:
        ctx = cc.Context(cc.Problem.from_shape(shape, precision=names[name]))
        x = W.initial_x(shape)
        if tune and name != "fp64":   # FP64 finite differences keep the default schedule
            ctx.autotune(x)
        hist, secs = [], []
        err, err_old, increase, reached, failure = 1.0, 1.0, 0, None, None
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
        out[name] = {"iterations_to_tol": reached, "mismatch": hist, "final_mismatch": hist[-1] if hist else None,
                     "Delta": float(x[dslot]), "s_per_iteration": float(np.mean(secs)) if secs else None,
                     "s_iterations": [round(v, 4) for v in secs], "total_s": float(np.sum(secs)),
                     "failure": failure, "schedule": ctx.schedule().as_dict()}
        ctx.close()
    for ref in ("fp64_tangent", "fp64"):
        if ref in out and out[ref]["iterations_to_tol"] is not None:
            for name, r in out.items():
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
    }


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=3,
                    help="workload shape 1, 2, 3 or 5; 4 = the batch of uneven problems")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 policy")
    ap.add_argument("--no-tune", action="store_true", help="default schedule instead of the autotuner's")
    ap.add_argument("--schedule", default="", help='fixed schedule fields as JSON, e.g. \'{"threads_tangent": 128}\'; '
                    "skips the autotuner")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--cpu-columns", type=int, default=64,
                    help="Jacobian columns of the CPU baseline sample (rounded up to a multiple of the threads)")
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
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.newton:
        cc = importlib.import_module("parallel-critical-collapse_amd")
        W = importlib.import_module("parallel-critical-collapse_amd.workload")
        shape = W.config(args.config)
        if args.scale is not None:
            from dataclasses import replace
            shape = replace(shape, scale=args.scale)
        res = newton_study(cc, shape, args.policies.split(","), args.tol, args.max_its, not args.no_tune,
                           args.save_failures or None)
        print(json.dumps({"workload": f"newton Ntau={shape.Ntau} x_points={shape.X} n={shape.n} Dim={shape.Dim} "
                                      f"XMid={shape.XMid} scale={shape.scale}", "tol": args.tol,
                          "policies": res}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the column shards of a multi-GPU run are described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    cc = importlib.import_module("parallel-critical-collapse_amd")
    dev = importlib.import_module("parallel-critical-collapse_amd.device")
    W = importlib.import_module("parallel-critical-collapse_amd.workload")
    if args.config == 4:
        print(json.dumps(batch_bench(cc, args.steps, args.warmup, not args.no_tune)), flush=True)
        return
    shape = W.config(args.config)
    log(f"[bench] workload Ntau={shape.Ntau} X={shape.X} n={shape.n}")

    tune = json.loads(args.schedule) if args.schedule else (not args.no_tune)
    mixed = time_policy(cc, dev, shape, cc.CC_MIXED, args.steps, args.warmup, False, tune)
    fp64 = None if args.no_fp64 else time_policy(cc, dev, shape, cc.CC_FP64, args.steps, args.warmup, False, tune)
    fp64r = None if args.no_fp64 else time_policy(cc, dev, shape, cc.CC_FP64, args.steps, args.warmup, True, tune)

    N = shape.Ntau
    T = mixed["schedule"]["threads_tangent"]
    fl = flops_chain(N, mixed["its"], mixed["irk_steps"], tangent=True)
    achieved = fl / (mixed["chain_ms"] / 1e3) / 1e12
    peak = 157.3
    roof = {"bound": "mfma", "pipe": "fp32 valu (FFT + pointwise; no dense contraction)",
            "kernel": f"k_tangent<{N},{T}>", "achieved": round(achieved, 3), "peak": peak,
            "unit": "TFLOP/s", "frac": round(achieved / peak, 4),
            "traffic": MEASURED_TRAFFIC_BYTES.get((f"k_tangent<{N},{T}>", mixed["schedule"]["steps_per_launch"])),
            "flops_per_step": fl / args.steps, "kernel_s_per_step": mixed["chain_ms"] / 1e3 / args.steps,
            "avg_launch_ms": mixed["chain_ms"] / max(mixed["launches"], 1),
            "launches_per_step": mixed["launches"] / args.steps}
    out = {
        "metric": "Jacobian+Newton steps/s at Ntau=1024, 8000 x-pts (mixed FP32/FP64)",
        "value": round(mixed["steps_per_s"], 5), "unit": "Newton steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(mixed["ms_per_step"], 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic: shipped D=4 solution + seeded N(0,1)*0.1/k^2 odd/even-harmonic perturbation",
        "config": {"workload": f"newton_step Ntau={shape.Ntau} x_points={shape.X} n={shape.n} Dim={shape.Dim} "
                               f"XMid={shape.XMid} IRK2", "policy": "mixed"},
        "schedule": mixed["schedule"], "tuning": mixed["tuning"],
        "kbar_tangent": round(mixed["kbar"], 3),
        "roofline": roof,
    }
    for key, res in (("fp64", fp64), ("fp64_reuse_sides", fp64r)):
        if res is None:
            continue
        fl64 = flops_chain(N, res["its"], res["irk_steps"])
        ach = fl64 / (res["chain_ms"] / 1e3) / 1e12
        out[key] = {"value": round(res["steps_per_s"], 5), "ms_per_step": round(res["ms_per_step"], 2),
                    "kbar": round(res["kbar"], 3), "kernel": f"k_chain<{N},{res['schedule']['threads_bulk']},double>",
                    "achieved_tflops": round(ach, 3), "peak": 78.6, "frac": round(ach / 78.6, 4),
                    "avg_launch_ms": res["chain_ms"] / max(res["launches"], 1),
                    "launches_per_step": res["launches"] / args.steps, "schedule": res["schedule"]}
        out[f"mixed_over_{key}"] = round(mixed["steps_per_s"] / res["steps_per_s"], 3)
    if not args.no_cpu:
        log(f"[bench] cpu baseline on {args.cpu_threads} threads")
        out["cpu_baseline"] = cpu_baseline(shape, args.cpu_threads, args.cpu_columns)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
