"""CC_FP32_FD: the reference's forward-difference Jacobian (NewtonSolver.cpp:
615-702) with every shoot in FP32 ("FP32 perturbed shooting solves"), against

  * the FP64 forward differences at the same step h (CC_FP64 with EpsNewton =
    h): the same truncation error, so what differs is the FP32 shoots'
    rounding and their looser IRK tolerance divided by h;
  * the FP64 tangent-linear Jacobian (CC_FP64_TANGENT), the most accurate
    Jacobian here, which the mixed policy's FP32 tangent columns track.

Run at Ntau=128 with 1000 x-points and at the benchmark shape (Ntau=1024,
8000 x-points), for h = 1e-3, 1e-4, 1e-5 (at h = 1e-2 the shoots of the high
fc columns stop converging, in FP64 as in FP32: 30 and 29 of them at Ntau=128).
F(x) of the policy is the FP64 residual, bit for bit."""
from dataclasses import replace

import numpy as np
import pytest

from _helpers import W, cc, ctx_for, shape_c1, shape_of

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b))


def ctx_h(s, h, **kw):
    """A context of the case's shape with the FD step EpsNewton = h."""
    return cc.Context(cc.Problem.from_shape(replace(s["shape"], EpsNewton=h), **kw))


def blocks(N):
    """Up, fc (with Delta) and psic slots of the Newton vector."""
    up = np.r_[0:N // 4]
    ps = np.r_[N // 4:N // 2]
    fc = np.r_[N // 2:3 * N // 4]
    return up, fc, ps


@pytest.fixture(scope="module", params=[(128, 1000), (1024, 8000)], ids=["ntau128", "ntau1024"])
def case(request):
    N, X = request.param
    s = shape_c1(N, X) if N == 128 else shape_of(W.config(3))
    x = s["x"]
    t64 = ctx_for(s, precision=cc.CC_FP64_TANGENT)
    F0 = t64.residual(x)
    Jt, cs, st = t64.jacobian(x, F0)
    assert st == 0 and not cs.any()
    dxt = np.linalg.solve(Jt, -F0)
    mixed = ctx_for(s, precision=cc.CC_MIXED)
    Jm, cs, st = mixed.jacobian(x, F0)
    assert st == 0 and not cs.any()
    Fm, dxm = mixed.newton_step(x)
    return s, F0, Jt, dxt, Jm, dxm


def test_fp32_fd_jacobian(case):
    """The FP32 columns are rounding noise over h: their distance to the FP64
    forward differences of the same h grows ~10x per 10x smaller h (a column
    mapped wrongly would not scale), and at every h both the Jacobian and the
    Newton step are far from the FP64 tangent-linear ones, which the mixed
    policy's FP32 tangent columns track.  Measured on an MI355X (vs the FP64
    tangent Jacobian; h = 1e-3, 1e-4, 1e-5): Ntau=128 2.7e-2, 0.19, 1.8
    (mixed 1.9e-6); Ntau=1024 6.2e-2, 0.63, 6.2 (mixed 3.0e-5); the FP32 FD
    step's Up and fc blocks 0.86 ... 2.3 relative (mixed <= 3.3e-3)."""
    s, F0, Jt, dxt, Jm, dxm = case
    sh, x = s["shape"], s["x"]
    N = sh.Ntau
    up, fc, ps = blocks(N)
    e_mixed = rel(Jm, Jt)
    dxm_e = {k: rel(dxm[b], dxt[b]) for k, b in (("Up", up), ("fc", fc), ("psic", ps))}
    same = []
    for h in (1e-3, 1e-4, 1e-5):
        ctx = ctx_h(s, h, precision=cc.CC_FP32_FD)
        J32, cs, st = ctx.jacobian(x, F0)
        assert st == 0 and not cs.any()
        assert np.all(np.isfinite(J32))
        F, dx = ctx.newton_step(x)
        assert np.array_equal(F, F0)                      # F(x) is the FP64 shoot
        assert np.array_equal(ctx.last_jacobian(), J32)
        f64 = ctx_h(s, h)
        J64, cs, st = f64.jacobian(x, F0)
        assert st == 0 and not cs.any()
        e_same = rel(J32, J64)
        e_true = rel(J32, Jt)
        cols_same = np.linalg.norm(J32 - J64, axis=0) / np.linalg.norm(J64, axis=0)
        dxe = {k: rel(dx[b], dxt[b]) for k, b in (("Up", up), ("fc", fc), ("psic", ps))}
        print(f"Ntau={N} h={h:g}: FP32 FD vs FP64 FD at the same h {e_same:.2e} (median column "
              f"{np.median(cols_same):.2e}, worst {cols_same.max():.2e}); vs FP64 tangent {e_true:.2e} "
              f"(mixed {e_mixed:.2e}); dx vs FP64 tangent step: FP32 FD "
              + ", ".join(f"{k} {v:.2e}" for k, v in dxe.items()) + "; mixed "
              + ", ".join(f"{k} {v:.2e}" for k, v in dxm_e.items()), flush=True)
        same.append(e_same)
        assert e_true > 100 * e_mixed
        assert dxe["Up"] > 0.5 and dxe["fc"] > 0.5
    assert dxm_e["Up"] < 1e-2 and dxm_e["fc"] < 1e-2
    for a, b in zip(same, same[1:]):
        assert 3 < b / a < 30, same


def test_fp32_fd_schedules_are_bit_identical():
    """The FP64 base shoot beside the FP32 launch (pipeline) or before it, and
    several chunk lengths: identical F0, J and step."""
    s = shape_c1(128)
    ctx = ctx_h(s, 1e-3, precision=cc.CC_FP32_FD)
    F0, dx = ctx.newton_step(s["x"])
    J = ctx.last_jacobian()
    for pipe, spl in [(0, 512), (1, 64), (0, 100), (1, 1000)]:
        ctx.set_schedule(pipeline=pipe, steps_per_launch=spl)
        F0b, dxb = ctx.newton_step(s["x"])
        print("pipeline", pipe, "steps_per_launch", spl, flush=True)
        assert np.array_equal(F0b, F0) and np.array_equal(dxb, dx)
        assert np.array_equal(ctx.last_jacobian(), J)
    with pytest.raises(cc.CCError):
        ctx.set_schedule(threads_bulk=32)
