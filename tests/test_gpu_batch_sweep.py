"""The batched workload at its full size: the 24 uneven problems of
workload.config4() (Ntau 128/256/512, 2000/4000/6000 x-points, D = 3.96 ...
4.10; the dimension suite of main.cpp:236-378 stepped together instead of one
after another) in one cc_batch, mixed policy.

Checked per problem: the batch step is bit-identical to stepping the problem
on its own context; F0 is the FP64 residual (cc_residual of an FP64 context,
bit for bit); and the mixed Jacobian agrees with the FP64 tangent-linear
Jacobian of the same construction blockwise: left side (psic and fc columns)
within 1e-4 on every problem (measured <= 8.3e-6), the Delta column (an FP64
forward difference in both) identical, right side (Up columns) within 2e-3
on every Ntau=512 problem (measured <= 1.0e-3).

The right side at Ntau 128 and 256 is reported, not bounded: on 7 of those 14
problems its error exceeds 2e-3 (up to 0.47).  Measured cause: the tangent
fixed point's stopping tolerance (1e-6 of the iterate), not FP32 -- the FP64
tangent stopped at the same 1e-6 shows the same error (0.47 and 6.7e-2 on the
two worst), and 1e-8 brings it to 1e-2 and 2e-3; the backward integration from
the horizon amplifies the error left in its fastest-growing high modes (Up
column norms up to 3e5 there).  From the four worst inputs (perturbation
scale 1) Newton fails with IRK non-convergence under every policy, the
reference's FP64 finite differences included."""
import numpy as np
import pytest

from _helpers import W, cc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sweep():
    shapes = W.config4()
    xs = [W.initial_x(sh) for sh in shapes]
    b = cc.Batch([cc.Problem.from_shape(sh, precision=cc.CC_MIXED) for sh in shapes])
    print(f"batch of {len(shapes)} problems, {b.nclasses} Ntau classes", flush=True)
    F0, dx, status, st = b.newton_step(xs)
    assert st == 0 and not status.any()
    b.close()
    return shapes, xs, F0, dx


def test_sweep_is_the_configured_workload(sweep):
    shapes = sweep[0]
    assert len(shapes) == 24 and {s.Ntau for s in shapes} == {128, 256, 512}
    assert {s.X for s in shapes} <= {2000, 4000, 6000} and len({s.Dim for s in shapes}) == 15


def test_batch_step_equals_single_steps(sweep):
    shapes, xs, F0, dx = sweep
    for i, sh in enumerate(shapes):
        ctx = cc.Context(cc.Problem.from_shape(sh, precision=cc.CC_MIXED))
        F_ref, dx_ref = ctx.newton_step(xs[i])
        ctx.close()
        assert np.array_equal(F0[i], F_ref), i
        assert np.array_equal(dx[i], dx_ref), i


def test_each_problem_against_fp64(sweep):
    shapes, xs, F0, _ = sweep
    worst = {"right": 0.0, "right512": 0.0, "left": 0.0, "delta": 0.0}
    for i, sh in enumerate(shapes):
        N = sh.Ntau
        f = cc.Context(cc.Problem.from_shape(sh))
        assert np.array_equal(F0[i], f.residual(xs[i])), i
        f.close()
        m = cc.Context(cc.Problem.from_shape(sh, precision=cc.CC_MIXED))
        t = cc.Context(cc.Problem.from_shape(sh, precision=cc.CC_FP64_TANGENT))
        Jm, csm, stm = m.jacobian(xs[i], F0[i])
        Jt, cst, stt = t.jacobian(xs[i], F0[i])
        m.close()
        t.close()
        assert stm == 0 and stt == 0, i
        dslot = N // 2 + 2
        right = [c for c in range(N // 4)]
        left = [c for c in range(N // 4, 3 * N // 4) if c != dslot]

        def blk(cols):
            return np.linalg.norm(Jm[:, cols] - Jt[:, cols]) / np.linalg.norm(Jt[:, cols])
        e = {"right": blk(right), "left": blk(left),
             "delta": np.linalg.norm(Jm[:, dslot] - Jt[:, dslot]) / np.linalg.norm(Jt[:, dslot])}
        print(f"problem {i}: Ntau={N} X={sh.X} D={sh.Dim} scale={sh.scale}: right {e['right']:.1e}, "
              f"left {e['left']:.1e}, Delta column {e['delta']:.1e}", flush=True)
        for k in ("right", "left", "delta"):
            worst[k] = max(worst[k], e[k])
        if N == 512:
            worst["right512"] = max(worst["right512"], e["right"])
        assert np.all(np.isfinite(Jm)), i
    print("worst over the sweep", worst)
    assert worst["left"] < 1e-4 and worst["delta"] == 0.0
    assert worst["right512"] < 2e-3
    assert worst["right"] < 1.0
