"""The C++ NewtonSolver driver (cc_newton) on the GPU: same control flow as
NewtonSolver::run (NewtonSolver.cpp:260-393), checked against the same loop
written here over the Python binding of the same C-ABI (bit-identical: both
call the same kernels and do the same host arithmetic), and its result
dictionary (NewtonSolver.cpp:775-816)."""
import json
import os
import subprocess

import numpy as np
import pytest

from _helpers import ROOT, W, cc
from test_host import config_dict

pytestmark = pytest.mark.gpu
CLI = os.path.join(ROOT, "parallel-critical-collapse_amd", "cc_newton")


def rms(v):
    """computeL2Norm (NewtonSolver.cpp:763-768) with the same left-to-right sum as the driver."""
    s = 0.0
    for e in v:
        s += float(e) * float(e)
    return float(np.sqrt(s / len(v)))


def newton_loop(shape, x0, tol, max_its, slow):
    ctx = cc.Context(cc.Problem.from_shape(shape, precision=cc.CC_MIXED))
    x = x0.copy()
    hist = []
    for its in range(max_its):
        F0, dx = ctx.newton_step(x)
        err = rms(F0)
        hist.append(err)
        if err < tol:
            return x, its, err, hist, ctx
        x = x + min(1.0, slow / err) * dx
    raise AssertionError(f"no convergence: {hist}")


@pytest.fixture(scope="module")
def anchored_config(tmp_path_factory):
    """The shipped near-solution at Ntau=128 on the 1000-point grid."""
    shape = W.Shape(Ntau=128, X=1000)
    xa, Delta = W.anchor(128)
    ctx = cc.Context(cc.Problem.from_shape(shape))
    Up, psic, fc = ctx.unpack(xa)
    d = config_dict(shape, Up, psic, fc, Delta, PrecisionNewton=1e-9, MaxIterNewton=8)
    p = tmp_path_factory.mktemp("host") / "sim.json"
    p.write_text(json.dumps(d))
    x0 = ctx.pack(Up, psic, fc)
    x0[128 // 2 + 2] = Delta
    return shape, p, d, x0


def test_cc_newton_matches_loop(anchored_config, tmp_path):
    shape, p, d, x0 = anchored_config
    out = tmp_path / "result.json"
    print("cc_newton mixed Ntau=128 X=1000", flush=True)
    r = subprocess.run([CLI, "-i", str(p), "-o", str(out), "--precision", "mixed"], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    res = json.loads(out.read_text())
    x, its, err, hist, ctx = newton_loop(shape, x0, d["PrecisionNewton"], d["MaxIterNewton"], d["SlowError"])
    assert res["Converged"] is True
    assert res["IterNewton"] == its
    assert res["mismatchNorm"] == err
    assert res["Initial_Conditions"]["Delta"] == x[128 // 2 + 2]
    for k in ("Ntau", "Dim", "XMid", "NLeft", "NRight", "PrecisionNewton", "SchemeIRK"):
        assert res[k] == d[k]
    # the written fields pack back to the converged Newton vector
    ic = res["Initial_Conditions"]
    xb = ctx.pack(np.array(ic["Up"]), np.array(ic["psic"]), np.array(ic["fc"]))
    xb[128 // 2 + 2] = ic["Delta"]
    assert np.max(np.abs(xb - x)) < 1e-13
    assert np.sqrt(np.mean(ctx.residual(xb) ** 2)) < d["PrecisionNewton"]


def test_cc_newton_already_converged(anchored_config, tmp_path):
    shape, p, d, x0 = anchored_config
    d2 = dict(d, Converged=True)
    p2 = tmp_path / "conv.json"
    p2.write_text(json.dumps(d2))
    out = tmp_path / "result.json"
    r = subprocess.run([CLI, "-i", str(p2), "-o", str(out), "--quiet"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    res = json.loads(out.read_text())
    ctx = cc.Context(cc.Problem.from_shape(shape))
    F = ctx.residual(x0)
    assert res["IterNewton"] == 0
    assert res["mismatchNorm"] == rms(F)


def test_cc_newton_reports_nonconvergence(anchored_config, tmp_path):
    """MaxIterNewton = 1 with a PrecisionNewton one step cannot reach: the run
    fails like the reference's (NewtonSolver.cpp:368-372): non-zero exit,
    the reference's message, no result file."""
    shape, p, d, x0 = anchored_config
    d2 = dict(d, MaxIterNewton=1, PrecisionNewton=1e-14)
    p2 = tmp_path / "fail.json"
    p2.write_text(json.dumps(d2))
    out = tmp_path / "result.json"
    r = subprocess.run([CLI, "-i", str(p2), "-o", str(out), "--quiet"], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "Newton method did not converge in 1 iterations" in r.stderr + r.stdout
    assert not out.exists()


def test_cc_newton_debug_writes_jacobians(anchored_config, tmp_path):
    """Debug runs write jacobian_<Dim>_<i>.csv per Newton iteration
    (NewtonSolver.cpp:337-341, OutputWriter.cpp:57-74): the first one is the
    Jacobian of the step at the initial point."""
    shape, p, d, x0 = anchored_config
    d2 = dict(d, Debug=True)
    p2 = tmp_path / "dbg.json"
    p2.write_text(json.dumps(d2))
    r = subprocess.run([CLI, "-i", str(p2), "-o", str(tmp_path / "res.json"), "--quiet"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    path = tmp_path / "jacobian_4.000000_1.csv"
    assert path.exists()
    J = np.loadtxt(path, delimiter=",")
    ctx = cc.Context(cc.Problem.from_shape(shape, precision=cc.CC_MIXED))
    ctx.newton_step(x0)
    J_ref = ctx.last_jacobian()
    assert J.shape == J_ref.shape
    assert np.max(np.abs(J - J_ref)) <= 1e-15 * np.max(np.abs(J_ref)) * 10   # 16 significant digits in the file
    # field snapshots of the shoot of iteration 0 (NewtonSolver.cpp:287-295), keyed by std::to_string(x)
    fpath = tmp_path / "fields_4.000000_0.json"
    assert fpath.exists()
    fj = json.loads(fpath.read_text())
    xs, fields = ctx.shoot_fields(x0, d2.get("DebugNx", 100))
    assert list(fj) == [f"{v:.6f}" for v in xs] or len(fj) == len({f"{v:.6f}" for v in xs})
    k = f"{xs[1]:.6f}"
    step = d2.get("DebugNtau", 1)
    keep = [j for j in range(shape.Ntau) if j % step == 0 or j == shape.Ntau - 1]
    for name, row in zip(("A", "U", "V", "f"), fields[1]):
        assert fj[k][name] == list(row[keep]), name


def test_cc_newton_multiple_run(anchored_config, tmp_path):
    """-m -r over D = 4.00 (anchored data), 4.01, 4.02 (each starting from the
    previous converged solution) and 4.03 (from the quadratic extrapolation of
    the three): every dimension converges and the suite file holds each
    result under its key (main.cpp:236-378)."""
    shape, _, d, _ = anchored_config
    suite = {"4.00": dict(d, MaxIterNewton=12)}
    for D in ("4.01", "4.02", "4.03"):
        m = dict(d, Dim=float(D), MaxIterNewton=12)
        m["Initial_Conditions"] = {}
        suite[D] = m
    p = tmp_path / "suite.json"
    p.write_text(json.dumps(suite))
    print("cc_newton -m -r, 4 dimensions at Ntau=128 X=1000", flush=True)
    r = subprocess.run([CLI, "-m", "-r", "-i", str(p), "--precision", "mixed"], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert "extrapolated (quadratic in D)" in r.stdout and r.stdout.count("copied from the previous") == 2
    res = json.loads(p.read_text())
    assert list(res) == ["4.00", "4.01", "4.02", "4.03"]
    deltas = []
    for D, v in res.items():
        assert v["Converged"] is True and v["Dim"] == float(D)
        assert v["mismatchNorm"] < d["PrecisionNewton"]
        deltas.append(v["Initial_Conditions"]["Delta"])
    # the converged Delta moves smoothly with D: the quadratic through the
    # first three predicts the fourth to far better than the step between them
    c = np.polyfit([4.00, 4.01, 4.02], deltas[:3], 2)
    assert abs(np.polyval(c, 4.03) - deltas[3]) < 0.1 * abs(deltas[3] - deltas[2])


def test_device_newton_loop_matches_host_loop(anchored_config):
    """cc_newton_run (norm, damping and update on the device) takes the same
    steps as the host loop over cc_newton_step: the same iterate, iteration
    count and mismatch history, bit for bit."""
    shape, _, d, x0 = anchored_config
    ctx = cc.Context(cc.Problem.from_shape(shape, precision=cc.CC_MIXED))
    print("device Newton loop", flush=True)
    x, its, err, hist, outcome = ctx.newton_run(x0, d["PrecisionNewton"], d["SlowError"], d["MaxIterNewton"])
    xr, its_r, err_r, hist_r, _ = newton_loop(shape, x0, d["PrecisionNewton"], d["MaxIterNewton"], d["SlowError"])
    assert outcome == 1 and its == its_r and err == err_r
    assert np.array_equal(hist, hist_r)
    assert np.array_equal(x, xr)


def test_cc_newton_device_loop_writes_the_same_result(anchored_config, tmp_path):
    shape, p, d, _ = anchored_config
    outs = []
    for flags in ([], ["--device-loop"]):
        out = tmp_path / f"res{len(flags)}.json"
        r = subprocess.run([CLI, "-i", str(p), "-o", str(out), "--quiet", *flags], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        outs.append(json.loads(out.read_text()))
    assert outs[0] == outs[1]


def test_cc_newton_irk3_input(anchored_config, tmp_path):
    """An input with SchemeIRK 3: cc_newton assembles the Jacobian from FP64
    finite differences (the tangent-linear columns follow the two-stage
    scheme) and converges; the result keeps the scheme."""
    shape, _, d, _ = anchored_config
    d3 = dict(d, SchemeIRK=3, MaxIterNewton=12)
    p = tmp_path / "irk3.json"
    p.write_text(json.dumps(d3))
    out = tmp_path / "irk3_out.json"
    print("cc_newton SchemeIRK 3", flush=True)
    r = subprocess.run([CLI, "-i", str(p), "-o", str(out)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    assert "SchemeIRK 3: FP64 finite-difference Jacobian" in r.stdout
    res = json.loads(out.read_text())
    assert res["Converged"] is True and res["SchemeIRK"] == 3
    assert res["mismatchNorm"] < d3["PrecisionNewton"]
