"""The mixed policy's solve: FP32 LU + FP64 iterative refinement, checked for
convergence on the device (||J dx - b|| / ||b|| <= refine_tol), else solved
again with the FP64 LU; and the device Newton loop's handling of a failure
after the residual (NewtonSolver.cpp:304-331 tests convergence and the
divergence guard before it assembles J).  Failures are forced from the host:
refine_tol = 1e-300 makes every refinement fall short, and EpsNewton = 1e300
makes every finite-difference column non-finite (its chains then fail the IRK
iteration at their first step, bounded by MaxIterIRK; the base column is
untouched)."""
import numpy as np
import pytest

from _helpers import W, cc, golden, shape_of

pytestmark = pytest.mark.gpu

FORCE = 1e-300


def rms(v):
    """computeL2Norm (NewtonSolver.cpp:763-768), summed left to right."""
    s = 0.0
    for e in v:
        s += float(e) * float(e)
    return float(np.sqrt(s / len(v)))


def system(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, n)) + 2.0 * np.eye(n), rng.standard_normal(n)


def test_refinement_accepts_a_converged_solve():
    A, b = system(384, 1)
    ctx = cc.Context(cc.Problem.make(Ntau=512, NLeft=10, NRight=10, precision=cc.CC_MIXED))
    x, info, _, st = ctx.solve(A, b)
    r, resolved = ctx.refinement()
    print(f"refined residual {r:.2e}")
    assert st == 0 and info == 0 and not resolved
    assert r <= 1e-10
    # the device's residual and numpy's round differently near 1e-14
    assert abs(r - np.linalg.norm(A @ x - b) / np.linalg.norm(b)) <= max(0.1 * r, 1e-13)


def test_forced_fp64_resolve_is_the_fp64_solve():
    """A refinement held to refine_tol = 1e-300 falls back to the FP64 LU:
    the solution and the returned factors are those of a CC_FP64 context, bit
    for bit."""
    A, b = system(384, 2)
    m = cc.Context(cc.Problem.make(Ntau=512, NLeft=10, NRight=10, precision=cc.CC_MIXED, refine_tol=FORCE))
    f = cc.Context(cc.Problem.make(Ntau=512, NLeft=10, NRight=10))
    xm, im, LUm, sm = m.solve(A, b)
    xf, i_f, LUf, sf = f.solve(A, b)
    r, resolved = m.refinement()
    assert sm == sf == 0 and im == i_f == 0 and resolved and r > FORCE
    assert np.array_equal(xm, xf)
    assert np.array_equal(LUm, LUf)
    assert f.refinement() == (0.0, False)


@pytest.fixture(scope="module")
def converged():
    """The 1000-point Ntau=128 shape at the oracle's converged iterate
    (tests/golden/newton_c1_oracle.json, mismatch 3e-14) and at its start."""
    fx = golden("newton_c1_oracle.json")
    from dataclasses import replace
    shape = replace(W.config(1), scale=fx["shape"]["scale"])
    return shape, np.array(fx["x"]), W.initial_x(shape)


def test_forced_resolve_in_newton_step(converged):
    shape, _, x0 = converged
    m = cc.Context(cc.Problem.from_shape(shape, precision=cc.CC_MIXED, refine_tol=FORCE))
    F0, dx = m.newton_step(x0)
    assert m.refinement()[1]
    J = m.last_jacobian()
    f = cc.Context(cc.Problem.from_shape(shape))
    ref, info, _, st = f.solve(J, -F0)
    assert st == 0 and info == 0
    assert np.array_equal(dx, ref)


def test_forced_resolve_in_device_newton_loop(converged):
    """cc_newton_run holds the update of a step whose refinement falls short
    and applies it after the FP64 re-solve: the same iterates, bit for bit, as
    the host loop over cc_newton_step (which returns the re-solved dx)."""
    shape, _, x0 = converged
    tol, its_max = 1e-9, 6
    m = cc.Context(cc.Problem.from_shape(shape, precision=cc.CC_MIXED, refine_tol=FORCE))
    x, its, err, hist, outcome = m.newton_run(x0, tol, shape.SlowError, its_max)
    h = cc.Context(cc.Problem.from_shape(shape, precision=cc.CC_MIXED, refine_tol=FORCE))
    xr, hist_r = x0.copy(), []
    for k in range(its_max):
        F0, dx = h.newton_step(xr)
        e = rms(F0)
        hist_r.append(e)
        if e < tol:
            break
        xr = xr + min(1.0, shape.SlowError / e) * dx
    print("device loop", outcome, its, hist, "host loop", hist_r)
    assert np.array_equal(hist, hist_r)
    assert np.array_equal(x, xr)


def test_newton_run_converged_despite_failing_jacobian(converged):
    """At an iterate that already meets the tolerance the loop ends converged
    even when every Jacobian column fails; at one that does not, the column
    failure is the loop's error and x is left unchanged."""
    from dataclasses import replace
    shape, xc, x0 = converged
    bad = cc.Context(cc.Problem.from_shape(replace(shape, EpsNewton=1e300)))
    good = cc.Context(cc.Problem.from_shape(shape))
    err_c = rms(good.residual(xc))
    assert err_c < 1e-12
    x, its, err, hist, outcome = bad.newton_run(xc, 1e-9, shape.SlowError, 4)
    assert outcome == 1 and its == 0 and err == err_c
    assert np.array_equal(x, xc)
    _, cs, st = bad.jacobian(xc, good.residual(xc), check=False)
    assert st == cc.CC_IRK_NOCONV and (cs == cc.CC_IRK_NOCONV).sum() >= shape.n - 1
    with pytest.raises(cc.CCError) as e:
        bad.newton_run(x0, 1e-9, shape.SlowError, 4)
    assert e.value.status == cc.CC_IRK_NOCONV
