#!/usr/bin/env python3
"""Write the oracle-computed fixtures under tests/golden/.

The oracle (oracle/cc_oracle.c, the CPU restatement of the reference path)
takes minutes to hours for what these fixtures hold, far beyond a test's
budget, so its results are computed once here, on the host, and committed as
plain JSON numbers.  The GPU tests and bench.py compare against them.

    python tests/golden/make_oracle_fixtures.py [--threads T] [--only NAME]

  newton_c1_oracle.json   NewtonSolver::run (NewtonSolver.cpp:279-366) with the
                          oracle's FD Jacobian (615-702) and dgesv (739-757):
                          the seeded input at Ntau=128, 1000 x-points, iterated
                          to mismatch < 1e-13 (or until the mismatch stops
                          falling): mismatch and Delta per iteration, final x.
  step_c2_oracle.json     one Newton step at the seeded input of Ntau=512,
  step_c3_oracle.json     6000 x-points and of Ntau=1024, 8000 x-points: F0 =
                          shoot(x), the whole FD Jacobian (n+1 shoots), dx =
                          dgesv(J, -F0), with J's column norms and extreme
                          singular values.
"""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle as O  # noqa: E402

W = importlib.import_module("parallel-critical-collapse_amd.workload")


def cfg_of(shape):
    return O.Config(Ntau=shape.Ntau, NLeft=shape.NLeft, NRight=shape.NRight, XMid=shape.XMid, XLeft=shape.XLeft,
                    XRight=shape.XRight, Dim=shape.Dim, Delta=shape.Delta, EpsNewton=shape.EpsNewton,
                    PrecisionIRK=shape.PrecisionIRK, MaxIterIRK=shape.MaxIterIRK)


def shape_desc(shape):
    return {"Ntau": shape.Ntau, "x_points": shape.X, "Dim": shape.Dim, "XMid": shape.XMid, "seed": shape.seed,
            "scale": shape.scale, "EpsNewton": shape.EpsNewton, "PrecisionIRK": shape.PrecisionIRK}


def newton_c1(threads, scale, tol=1e-13, max_its=15):
    # from the seeded input at scale 1 the oracle's shoot stops converging at
    # the fifth Newton iterate (mismatch 1.9e-2 -> 1.6e-2 under the damping
    # min(1, SlowError/err)); the perturbation is scaled down as for the
    # larger shapes
    from dataclasses import replace
    shape = replace(W.config(1), scale=scale)
    cfg = cfg_of(shape)
    o = O.Oracle(cfg)
    dslot = shape.Ntau // 2 + 2
    x = W.initial_x(shape)
    xold = x.copy()
    err, err_old, rises = 1.0, 1.0, 0
    hist, deltas, outcome = [], [], "max_its"
    t0 = time.perf_counter()
    for its in range(max_its):
        err_old = err
        F0, st = o.shoot(x)
        if st != 0:
            outcome = f"shoot status {st}"
            break
        err = O.l2norm(F0)
        hist.append(err)
        deltas.append(float(x[dslot]))
        print(f"[c1 newton] iteration {its + 1}: mismatch {err:.3e}, Delta {x[dslot]!r}", flush=True)
        if err < tol:
            outcome = "converged"
            break
        if np.log10(err) > np.log10(err_old):
            rises += 1
        if rises >= 5:
            outcome = "mismatch rose five times"
            x = xold
            break
        J, status, _, worst = O.jacobian(cfg, x, F0, nthreads=threads)
        if worst != 0:
            outcome = f"jacobian status {worst}"
            break
        dx, info, _, _ = O.dgesv(J, -F0)
        if info != 0:
            outcome = f"dgesv info {info}"
            break
        fac = min(1.0, shape.SlowError / err)
        if np.log10(err) < np.log10(err_old):
            xold = x.copy()
            rises = 0
        x = x + fac * dx
    return {"what": "oracle Newton loop (NewtonSolver.cpp:279-366) from the seeded input",
            "shape": shape_desc(shape), "tol": tol, "outcome": outcome, "mismatch": hist, "Delta": deltas,
            "x": x.tolist(), "Delta_final": float(x[dslot]), "seconds": round(time.perf_counter() - t0, 1)}


def step_at(config, threads):
    shape = W.config(config)
    cfg = cfg_of(shape)
    o = O.Oracle(cfg)
    x = W.initial_x(shape)
    t0 = time.perf_counter()
    F0, st = o.shoot(x)
    assert st == 0, st
    J, status, cnt, worst = O.jacobian(cfg, x, F0, nthreads=threads)
    assert worst == 0, worst
    dx, info, _, _ = O.dgesv(J, -F0)
    assert info == 0, info
    sv = np.linalg.svd(J, compute_uv=False)
    print(f"[c{config} step] {time.perf_counter() - t0:.0f} s, |F0| {O.l2norm(F0):.3e}, |dx| {np.linalg.norm(dx):.3e}, "
          f"cond {sv[0] / sv[-1]:.2e}", flush=True)
    return {"what": "oracle Newton step: F0 = shoot(x), FD Jacobian (n+1 shoots), dx = dgesv(J, -F0)",
            "shape": shape_desc(shape), "F0": F0.tolist(), "dx": dx.tolist(),
            "J_col_norms": np.linalg.norm(J, axis=0).tolist(), "J_fro": float(np.linalg.norm(J)),
            "sigma_max": float(sv[0]), "sigma_min": float(sv[-1]),
            "irk_steps": int(cnt[0]), "irk_iterations": int(cnt[1]),
            "seconds": round(time.perf_counter() - t0, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--only", default="", help="newton_c1, step_c2 or step_c3")
    ap.add_argument("--c1-scale", type=float, default=0.1, help="perturbation scale of the Newton fixture")
    args = ap.parse_args()
    O.build()
    jobs = {"newton_c1": lambda: newton_c1(args.threads, args.c1_scale),
            "step_c2": lambda: step_at(2, args.threads),
            "step_c3": lambda: step_at(3, args.threads)}
    for name, fn in jobs.items():
        if args.only and name != args.only:
            continue
        res = fn()
        with open(os.path.join(HERE, f"{name}_oracle.json"), "w") as f:
            json.dump(res, f)
        print("wrote", name, flush=True)


if __name__ == "__main__":
    main()
