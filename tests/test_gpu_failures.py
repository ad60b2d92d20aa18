"""Failure reporting of the HIP path, forced from the host (every kernel still
exits normally): IRK non-convergence (ShootingSolver.cpp:69-82, 112-125)
surfaces as CC_IRK_NOCONV from the residual, from every Jacobian column and
from the Newton step, where the oracle's shoot fails the same way."""
import dataclasses

import numpy as np
import pytest

from _helpers import cc, ctx_for, oracle_for, shape_c1, shape_of

pytestmark = pytest.mark.gpu


def capped():
    """Ntau=128 on a 100-point grid with MaxIterIRK = 1, which cannot meet
    PrecisionIRK = 1e-15 on the first step."""
    s = shape_c1(128, 100)
    return shape_of(dataclasses.replace(s["shape"], MaxIterIRK=1))


def test_irk_nonconvergence_matches_oracle():
    s = capped()
    _, st = oracle_for(s).shoot(s["x"])
    assert st == 1
    with pytest.raises(cc.CCError) as e:
        ctx_for(s).residual(s["x"])
    assert e.value.status == cc.CC_IRK_NOCONV


@pytest.mark.parametrize("precision", [cc.CC_FP64, cc.CC_MIXED, cc.CC_FP64_TANGENT])
def test_irk_nonconvergence_in_jacobian_and_newton_step(precision):
    s = capped()
    ctx = ctx_for(s, precision=precision)
    _, cs, st = ctx.jacobian(s["x"], np.zeros(s["x"].size), check=False)
    assert st == cc.CC_IRK_NOCONV
    assert (cs == cc.CC_IRK_NOCONV).all()
    with pytest.raises(cc.CCError) as e:
        ctx.newton_step(s["x"])
    assert e.value.status == cc.CC_IRK_NOCONV
    # later contexts are unaffected: the 1000-point shape with the reference's
    # iteration cap shoots normally afterwards (100 points are too coarse for
    # the IRK fixed point at any cap)
    ok = shape_c1(128)
    assert np.all(np.isfinite(ctx_for(ok, precision=precision).residual(ok["x"])))
