"""Batched independent problems (cc_batch_*): problems of one Ntau share the
base and tangent chain launches (per-chain D, grid and base index), classes
of different Ntau run concurrently.  Every chain runs the same arithmetic as
in a single-problem launch, so a batch step is bit-identical to stepping each
problem on its own context."""
import numpy as np
import pytest

from _helpers import W, cc, shape_c1

pytestmark = pytest.mark.gpu


def batch_shapes():
    """Six problems over three Ntau classes, uneven grids and D; each shoots
    in the oracle (checked when the shapes were chosen)."""
    return [shape_c1(128)["shape"],
            W.Shape(Ntau=128, X=1400, Dim=4.0, seed=W.SEED_BASE + 11, scale=0.1),
            W.Shape(Ntau=256, X=4000, Dim=4.0, seed=W.SEED_BASE + 12, scale=0.1),
            W.Shape(Ntau=256, X=3000, Dim=4.05, seed=W.SEED_BASE + 13, scale=0.01),
            W.Shape(Ntau=128, X=2000, Dim=3.98, seed=W.SEED_BASE + 14, scale=0.01),
            W.Shape(Ntau=512, X=4000, Dim=4.02, seed=W.SEED_BASE + 15, scale=0.01)]


@pytest.fixture(scope="module")
def singles():
    """Each problem stepped on its own context (default schedule)."""
    out = []
    for sh in batch_shapes():
        ctx = cc.Context(cc.Problem.from_shape(sh, precision=cc.CC_MIXED))
        x = W.initial_x(sh)
        print("single", sh.Ntau, sh.X, sh.Dim, flush=True)
        F0, dx = ctx.newton_step(x)
        out.append((x, F0, dx))
        ctx.close()
    return out


def make_batch():
    return cc.Batch([cc.Problem.from_shape(sh, precision=cc.CC_MIXED) for sh in batch_shapes()])


def test_batch_step_equals_single_steps(singles):
    b = make_batch()
    assert b.nclasses == 3
    xs = [s[0] for s in singles]
    print("batch step", flush=True)
    F0, dx, status, st = b.newton_step(xs)
    assert st == 0 and not status.any()
    for i, (x, F_ref, dx_ref) in enumerate(singles):
        assert np.array_equal(F0[i], F_ref), i
        assert np.array_equal(dx[i], dx_ref), i
    c = b.counters()
    assert c[2] == sum(4 + 3 * sh.Ntau // 4 - 1 for sh in batch_shapes())   # base, Delta and n-1 tangent chains


def test_batch_schedules_and_graph_are_bit_identical(singles):
    b = make_batch()
    xs = [s[0] for s in singles]
    for cls in range(b.nclasses):
        b.set_schedule(cls, pipeline=0, steps_per_launch=100)
    F0, dx, status, st = b.newton_step(xs)
    assert st == 0
    for i, (_, F_ref, dx_ref) in enumerate(singles):
        assert np.array_equal(F0[i], F_ref) and np.array_equal(dx[i], dx_ref), i
    b.set_schedule(0, use_graph=1, pipeline=1, steps_per_launch=256)
    for rep in range(3):
        print("graph replay", rep, flush=True)
        F0, dx, status, st = b.newton_step(xs)
        assert st == 0
        for i, (_, F_ref, dx_ref) in enumerate(singles):
            assert np.array_equal(F0[i], F_ref) and np.array_equal(dx[i], dx_ref), (rep, i)


def test_batch_autotune(singles):
    b = make_batch()
    xs = [s[0] for s in singles]
    times = b.autotune(xs)
    print("batch tuning times", times, flush=True)
    assert len(times) >= 2 * b.nclasses + 2 and np.all(times > 0)
    F0, dx, status, st = b.newton_step(xs)
    assert st == 0
    Ns = sorted({sh.Ntau for sh in batch_shapes()})
    for i, ((_, F_ref, dx_ref), sh) in enumerate(zip(singles, batch_shapes())):
        assert np.array_equal(F0[i], F_ref), i        # F0 comes from the FP64 base chains (fixed thread count)
        _, sc = b.schedule(Ns.index(sh.Ntau))
        if sc.threads_tangent == max(64, sh.Ntau // 4):     # the default thread count of this Ntau
            assert np.array_equal(dx[i], dx_ref), i
        else:                                           # other FFT radix in the FP32 tangents
            assert np.all(np.isfinite(dx[i]))


def test_batch_rejects_mixed_policies():
    sh = batch_shapes()
    with pytest.raises(cc.CCError):
        cc.Batch([cc.Problem.from_shape(sh[0], precision=cc.CC_MIXED),
                  cc.Problem.from_shape(sh[1], precision=cc.CC_FP64)])
