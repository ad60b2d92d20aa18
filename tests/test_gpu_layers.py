"""The lower layers on the GPU against the oracle: the spectral operators
(cc_spectral: SpectralTransformer.cpp:140-292), the fields of a state
(cc_state_fields: ICG.cpp:891-932), cc_integrate's per-call iteration limit
and failing index (ODEStepper.cpp:144-150, ShootingSolver.cpp:62-82), and the
C++ classes SpectralTransformer, ODEStepper and ShootingSolver
(host/layers.hpp), driven through the cc_layers executable.

Bounds: differentiate, halveModes and doubleModes are the same few roundings
on both sides (bit-identical); lamIntegrate within 1e-14 per entry (the device
contracts a*b + c into fma inside the complex division); solveInhom within
1e-12 normwise (two transforms each way and exp); one IRK step and whole
sides within 1e-12 / 1e-11 as in tests/test_gpu_parity.py; the C++ classes
bit-identical to the Python binding of the same entry points."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from _helpers import ROOT, cc, ctx_for, oracle_for, shape_c1

pytestmark = pytest.mark.gpu
CLI = os.path.join(ROOT, "parallel-critical-collapse_amd", "cc_layers")
PERIOD = 3.4453


def rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


def flat(z):
    z = np.asarray(z, np.complex128)
    return [float(v) for v in np.column_stack([z.real, z.imag]).ravel()]


def cplx(v):
    v = np.asarray(v, np.float64)
    return v[0::2] + 1j * v[1::2]


def run_layers(tmp_path, request, name="req"):
    p, q = tmp_path / f"{name}.json", tmp_path / f"{name}_out.json"
    p.write_text(json.dumps(request))
    print(f"cc_layers {name}: Ntau={request['Ntau']}, {len(request['calls'])} calls", flush=True)
    r = subprocess.run([CLI, str(p), str(q)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return json.loads(q.read_text())["results"]


def spectral_inputs(N, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((3, N)) + 1j * rng.standard_normal((3, N))
    f = 0.3 + 0.2 * np.cos(2 * np.pi * np.arange(N) / N) + 0.05 * rng.standard_normal((3, N))
    g = rng.standard_normal((3, N))
    return X, f, g


@pytest.mark.parametrize("N", [128, 512, 2048])
def test_spectral_operators_match_oracle(N):
    ctx = cc.Context(cc.Problem.make(Ntau=N, NLeft=10, NRight=10))
    o = O.Oracle(O.Config(Ntau=N, Delta=PERIOD))
    X, f, g = spectral_inputs(N, N)
    # differentiate
    got = ctx.spectral(cc.CC_SPEC_DIFF, X, PERIOD)
    ref = np.array([o.differentiate(x, PERIOD) for x in X])
    assert np.array_equal(got, ref)
    # halveModes, doubleModes; halving the doubled modes gives back 1.5 x the
    # Nyquist entry (doubleModes leaves the full value in slot 3N/2,
    # SpectralTransformer.cpp:239-245), as in the oracle
    h = ctx.spectral(cc.CC_SPEC_HALVE, X)
    assert np.array_equal(h, np.array([O.Oracle.halve_modes(x) for x in X]))
    d = ctx.spectral(cc.CC_SPEC_DOUBLE, h)
    assert np.array_equal(d, np.array([O.Oracle.double_modes(x) for x in h]))
    hh = ctx.spectral(cc.CC_SPEC_HALVE, d)
    assert np.array_equal(hh, np.array([O.Oracle.halve_modes(x) for x in d]))
    assert np.array_equal(np.delete(hh, N // 4, axis=1), np.delete(h, N // 4, axis=1))
    assert np.array_equal(hh[:, N // 4], 1.5 * h[:, N // 4])
    # lamIntegrate: lambda = 0 (the DC mode zeroed), a general lambda, and
    # lambda = i k0, whose m = 1 denominator is exactly zero (entry set to 0)
    k0 = 2 * np.pi / PERIOD
    for lam in (0.0, 0.7 - 0.3j, 1j * k0, 2.5):
        got = ctx.spectral(cc.CC_SPEC_LAM, X, PERIOD, lam)
        ref = np.array([o.lam_integrate(x, lam, PERIOD) for x in X])
        assert np.array_equal(got == 0, ref == 0), lam
        assert np.all(np.abs(got - ref) <= 1e-14 * np.abs(ref)), lam
    # solveInhom: x' + f x + g = 0, periodic
    x = ctx.spectral(cc.CC_SPEC_INHOM, np.stack([f, g], axis=1), PERIOD)
    ref = np.array([o.solve_inhom(f[b], g[b], PERIOD) for b in range(3)])
    assert rel(x, ref) < 1e-12


def test_spectral_rejects_bad_period():
    ctx = cc.Context(cc.Problem.make(Ntau=128, NLeft=10, NRight=10))
    X = np.ones(128, np.complex128)
    for p in (0.0, -1.0, np.inf):
        with pytest.raises(RuntimeError, match="period"):
            ctx.spectral(cc.CC_SPEC_DIFF, X, p)
    # halve / double do not read the period
    assert ctx.spectral(cc.CC_SPEC_HALVE, X, 0.0).shape == (64,)


@pytest.fixture(scope="module")
def c1_sides():
    """Boundary states and match-point states of the oracle's shoot at Ntau=128, 1000 x-points."""
    s = shape_c1(128)
    o = oracle_for(s)
    F, st, (YL0, YR0, YLm, YRm) = o.shoot_states(s["x"])
    assert st == 0
    return s, o, o.grid(), (YL0, YR0, YLm, YRm)


def test_state_fields_match_oracle(c1_sides):
    s, o, grid, (YL0, YR0, _, _) = c1_sides
    iM, iR = s["cfg"].NLeft, s["cfg"].NLeft + s["cfg"].NRight
    states, xs = [], []
    for Y0, idx in ((YL0, [0, 37, 250, iM - 1]), (YR0, [iR, iR - 3, 700, iM + 1])):
        Y, at = Y0, idx[0]
        for i in idx:
            if i != at:
                Y, st, _ = o.integrate(Y, grid, at, i)
                assert st == 0
                at = i
            states.append(Y)
            xs.append(grid[i])
    ctx = ctx_for(s)
    fields = ctx.state_fields(np.array(states), np.array(xs))
    for k, (Y, x) in enumerate(zip(states, xs)):
        U, V, F, IA2, *_ = o.state_to_fields_d(Y, x)
        for got, ref in zip(fields[k], (np.sqrt(1 / IA2), U, V, F)):
            assert np.max(np.abs(got - ref)) < 1e-10 * max(1.0, np.max(np.abs(ref))), k
    # a state with another Delta in Re Y[2] (states are not paired across Delta)
    Y2 = states[1].copy()
    Y2[2] += 1e-3
    both = ctx.state_fields(np.array([states[1], Y2]), np.array([xs[1], xs[1]]))
    U, V, F, IA2, *_ = o.state_to_fields_d(Y2, xs[1])
    assert np.max(np.abs(both[1][1] - U)) < 1e-10 * max(1.0, np.max(np.abs(U)))
    assert np.array_equal(both[0], fields[1])


def test_integrate_iteration_limit_and_failing_index(c1_sides):
    """cc_integrate with max_its: the step at which the fixed point first fails
    to converge and the iteration counts agree with the oracle's ODEStepper
    (whose late-iteration tolerance relaxation depends on the limit,
    ODEStepper.cpp:265-269)."""
    s, o, grid, (YL0, _, _, _) = c1_sides
    ctx = ctx_for(s)
    nsteps = 30
    for m in (1, 2, 3, 4, 6, 0):
        Y, fail_ref, its_ref = YL0.copy(), -1, 0
        for i in range(nsteps):
            Yn, conv, its = o.step_irk(Y, grid[i], grid[i + 1], maxIts=m or None)
            its_ref += its
            if not conv:
                fail_ref = i
                break
            Y = Yn
        got, its, st, fail = ctx.integrate(YL0[None, :], grid, [0], [nsteps], max_its=m, with_fail_idx=True)
        print(f"max_its {m}: oracle fails at {fail_ref}, device at {fail[0]} (status {st[0]}), "
              f"iterations {its_ref} / {its[0]}", flush=True)
        assert fail[0] == fail_ref and st[0] == (0 if fail_ref < 0 else 1), m
        if fail_ref >= 0:
            assert rel(got[0], Y) < 1e-12         # the state at the failing step's start
        else:
            assert rel(got[0], Y) < 1e-12 and abs(int(its[0]) - its_ref) <= 2
    with pytest.raises(RuntimeError):
        ctx.integrate(YL0[None, :], grid, [0], [5], max_its=-1)


def test_cpp_spectral_transformer(tmp_path):
    """SpectralTransformer's ten methods through the C++ class: the same
    results as the Python binding of the same entry points (bit for bit) and
    the oracle's."""
    N = 256
    X, f, g = spectral_inputs(N, 7)
    x, z = X[0], X[1]
    calls = [
        {"method": "forwardFFT_real", "in": [float(v) for v in f[0]]},
        {"method": "backwardFFT_real", "in": flat(x)},
        {"method": "forwardFFT", "in": flat(x)},
        {"method": "backwardFFT", "in": flat(x)},
        {"method": "differentiate", "in": flat(z)},
        {"method": "differentiate", "in": flat(z), "period": 2.0},
        {"method": "lamIntegrate", "in": flat(z), "lambda": [0.7, -0.3]},
        {"method": "halveModes", "in": flat(z)},
        {"method": "doubleModes", "in": flat(z[: N // 2])},
        {"method": "solveInhom", "f": [float(v) for v in f[1]], "g": [float(v) for v in g[1]]},
        {"method": "halveModes", "in": flat(z[:10])},                 # wrong length
    ]
    res = run_layers(tmp_path, {"Ntau": N, "period": PERIOD, "calls": calls})
    ctx = cc.Context(cc.Problem.make(Ntau=N, NLeft=10, NRight=10))
    o = O.Oracle(O.Config(Ntau=N, Delta=PERIOD))
    out = [np.asarray(r["out"]) if "out" in r else r for r in res]
    assert np.array_equal(cplx(out[0]), ctx.fft(f[0].astype(np.complex128), +1)[0])
    assert np.array_equal(out[1], ctx.fft(x, -1)[0].real)
    assert np.array_equal(cplx(out[2]), ctx.fft(x, +1)[0])
    assert np.array_equal(cplx(out[3]), ctx.fft(x, -1)[0])
    assert rel(cplx(out[0]), o.forward_r2c(f[0])) < 1e-14
    assert rel(out[1], o.backward_c2r(x)) < 1e-14
    assert rel(cplx(out[2]), o.forward_c2c(x)) < 1e-14
    assert rel(cplx(out[3]), o.backward_c2c(x)) < 1e-14
    assert np.array_equal(cplx(out[4]), o.differentiate(z, PERIOD))           # period 0: the constructor's
    assert np.array_equal(cplx(out[5]), o.differentiate(z, 2.0))
    assert np.array_equal(cplx(out[6]), ctx.spectral(cc.CC_SPEC_LAM, z, PERIOD, 0.7 - 0.3j))
    assert np.all(np.abs(cplx(out[6]) - o.lam_integrate(z, 0.7 - 0.3j, PERIOD)) <= 1e-14 * np.abs(cplx(out[6])))
    assert np.array_equal(cplx(out[7]), O.Oracle.halve_modes(z))
    assert np.array_equal(cplx(out[8]), O.Oracle.double_modes(z[: N // 2]))
    assert np.array_equal(out[9], ctx.spectral(cc.CC_SPEC_INHOM, np.stack([f[1], g[1]]), PERIOD))
    assert rel(out[9], o.solve_inhom(f[1], g[1], PERIOD)) < 1e-12
    assert out[10]["status"] == cc.CC_BADARG and "expected 256" in out[10]["error"]


def test_cpp_ode_stepper(tmp_path, c1_sides):
    """ODEStepper::integrate: one IRK step against the oracle's stepIRK, and a
    step whose fixed point cannot converge in one iteration."""
    s, o, grid, (YL0, YR0, _, _) = c1_sides
    cfg = s["cfg"]
    iR = cfg.NLeft + cfg.NRight
    calls = [{"method": "integrate", "Y": flat(YL0), "Xin": grid[0], "Xout": grid[1], "maxIts": 100},
             {"method": "integrate", "Y": flat(YR0), "Xin": grid[iR], "Xout": grid[iR - 1], "maxIts": 100},
             {"method": "integrate", "Y": flat(YL0), "Xin": grid[0], "Xout": grid[1], "maxIts": 1}]
    res = run_layers(tmp_path, {"Ntau": 128, "Dim": cfg.Dim, "PrecisionIRK": cfg.PrecisionIRK, "calls": calls})
    for r, (Y, a, b) in zip(res[:2], ((YL0, 0, 1), (YR0, iR, iR - 1))):
        ref, conv, its = o.step_irk(Y, grid[a], grid[b])
        assert conv and r["converged"]
        assert rel(cplx(r["Yout"]), ref) < 1e-12
        assert abs(r["itsReached"] - its) <= 1
    _, conv1, _ = o.step_irk(YL0, grid[0], grid[1], maxIts=1)
    assert not conv1 and not res[2]["converged"]
    assert np.array_equal(cplx(res[2]["Yout"]), YL0)


def test_cpp_shooting_solver(tmp_path, c1_sides):
    """ShootingSolver::shoot at Ntau=128, 1000 x-points: both match-point states
    and the mismatch against the oracle's (ShootingSolver.cpp:33-43, 154-182);
    with Debug the reference's snapshot keys and fields (ShootingSolver.cpp:
    85-97), the states bit-identical to the shoot without snapshots."""
    s, o, grid, (YL0, YR0, YLm, YRm) = c1_sides
    cfg = s["cfg"]
    iM, iR = cfg.NLeft, cfg.NLeft + cfg.NRight
    base = {"method": "shoot", "YLeft": flat(YL0), "YRight": flat(YR0), "grid": [float(v) for v in grid],
            "iLeft": 0, "iRight": iR, "iMid": iM}
    calls = [base, dict(base, Debug=True, DebugNx=100, DebugNtau=3)]
    res = run_layers(tmp_path, {"Ntau": 128, "Dim": cfg.Dim, "PrecisionIRK": cfg.PrecisionIRK,
                                "MaxIterIRK": cfg.MaxIterIRK, "calls": calls})
    r0, r1 = res
    YL, YR = cplx(r0["YLeft"]), cplx(r0["YRight"])
    assert rel(YL, YLm) < 1e-11 and rel(YR, YRm) < 1e-11
    mis_ref, st = o.mismatch(YLm, YRm)
    assert st == 0
    mis = cplx(r0["mismatch"])
    assert np.max(np.abs(mis - mis_ref)) < 1e-12 * np.max(np.abs(YLm))
    assert mis[2].real == YL[2].real
    for k in ("YLeft", "YRight", "mismatch", "itsLeft", "itsRight"):
        assert r1[k] == r0[k], k
    # snapshot keys in the reference's order: left ascending, right descending
    li = [i for i in range(iM) if i % 100 == 0 or i == iM - 1]
    ri = [i for i in range(iR, iM, -1) if i % 100 == 0 or i == iM + 1 or i == iR]
    fields = r1["fields"]
    assert list(fields) == [f"{grid[i]:.6f}" for i in li + ri]
    keep = [j for j in range(128) if j % 3 == 0 or j == 127]
    for Y0, idx in ((YL0, li), (YR0, ri)):
        Y, at = Y0, idx[0]
        for i in idx:
            if i != at:
                Y, st, _ = o.integrate(Y, grid, at, i)
                assert st == 0
                at = i
            U, V, F, IA2, *_ = o.state_to_fields_d(Y, grid[i])
            e = fields[f"{grid[i]:.6f}"]
            for name, ref in (("A", np.sqrt(1 / IA2)), ("U", U), ("V", V), ("f", F)):
                got = np.asarray(e[name])
                assert np.max(np.abs(got - ref[keep])) < 1e-10 * max(1.0, np.max(np.abs(ref))), (i, name)


def test_cpp_shooting_solver_errors(tmp_path, c1_sides):
    """The reference's process exits as errors: a fixed point that does not
    converge (ShootingSolver.cpp:69-82, with the failing grid points) and a
    right-side Delta above the left one's by more than the precision
    (ShootingSolver.cpp:165-176)."""
    s, o, grid, (YL0, YR0, _, _) = c1_sides
    cfg = s["cfg"]
    iM, iR = cfg.NLeft, cfg.NLeft + cfg.NRight
    base = {"method": "shoot", "YLeft": flat(YL0), "YRight": flat(YR0), "grid": [float(v) for v in grid],
            "iLeft": 0, "iRight": iR, "iMid": iM}
    req = {"Ntau": 128, "Dim": cfg.Dim, "PrecisionIRK": cfg.PrecisionIRK, "MaxIterIRK": 1, "calls": [base]}
    r = run_layers(tmp_path, req, "noconv")[0]
    assert r["status"] == cc.CC_IRK_NOCONV
    assert "No convergence between grid point" in r["error"] and "(0)" in r["error"] and "(1)" in r["error"]
    YRd = YR0.copy()
    YRd[2] += 1e-6
    calls = [dict(base, YRight=flat(YRd)), dict(base, iLeft=iM + 1)]
    req = {"Ntau": 128, "Dim": cfg.Dim, "PrecisionIRK": cfg.PrecisionIRK, "MaxIterIRK": cfg.MaxIterIRK,
           "calls": calls}
    r = run_layers(tmp_path, req, "delta")
    assert r[0]["status"] == cc.CC_DELTA_MISMATCH and "Mismatch in Delta" in r[0]["error"]
    assert r[1]["status"] == cc.CC_BADARG
