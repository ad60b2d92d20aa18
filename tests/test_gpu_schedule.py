"""Launch schedules, the autotuner and hipGraph replay: none of them may change
what is computed.

* steps_per_launch, pipeline and use_graph only change how the same kernels
  are enqueued (chunk boundaries save and restore the chain state exactly),
  so results are bit-identical;
* the alternative thread count of a chain kernel changes the radix of the
  Stockham passes and hence the rounding: FP64 results agree with the oracle
  at the usual tolerances, FP32 tangent columns agree with the default
  thread count to FP32 accuracy.
"""
from dataclasses import replace

import numpy as np
import pytest

from _helpers import W, cc, ctx_for, oracle_for, shape_c1

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


def shape_c2():
    """The Ntau=512, 6000-point shape (Ntau >= 512 has a second thread count)."""
    import oracle as O
    shape = W.config(2)
    cfg = O.Config(Ntau=shape.Ntau, NLeft=shape.NLeft, NRight=shape.NRight, XMid=shape.XMid,
                   XLeft=shape.XLeft, XRight=shape.XRight, Dim=shape.Dim, Delta=shape.Delta,
                   EpsNewton=shape.EpsNewton, PrecisionIRK=shape.PrecisionIRK, MaxIterIRK=shape.MaxIterIRK)
    return {"shape": shape, "cfg": cfg, "x": W.initial_x(shape)}


@pytest.mark.parametrize("precision", [cc.CC_FP64, cc.CC_MIXED, cc.CC_FP32_FD])
def test_graph_replay_matches_eager(precision):
    """Three replays of the captured Newton step, with eager work (a residual
    and a synchronisation) between them, each bit-identical to the eager step."""
    s = shape_c1(128)
    if precision == cc.CC_FP32_FD:          # an FD step FP32 shoots resolve (1e-8 leaves J singular)
        s = dict(s, shape=replace(s["shape"], EpsNewton=1e-3))
    ctx = ctx_for(s, precision=precision)
    F0_e, dx_e = ctx.newton_step(s["x"])
    ctx.set_schedule(use_graph=1)
    x2 = s["x"].copy()
    x2[3] += 1e-6
    for rep in range(3):
        F0_g, dx_g = ctx.newton_step(s["x"])
        print("replay", rep, flush=True)
        assert np.array_equal(F0_g, F0_e)
        assert np.array_equal(dx_g, dx_e)
        ctx.residual(x2)
        ctx.synchronize()


def test_stage_schedule_is_bit_identical():
    """Mixed Jacobian with the base on its own stream (pipeline) or on the
    main stream, and at several chunk lengths: identical results."""
    s = shape_c1(128)
    ctx = ctx_for(s, precision=cc.CC_MIXED)
    F0 = ctx.residual(s["x"])
    ref, cs, st = ctx.jacobian(s["x"], F0)
    assert st == 0 and not cs.any()
    for pipe, spl in [(0, 512), (1, 64), (0, 100), (1, 1000), (1, 8), (1, 13)]:
        ctx.set_schedule(pipeline=pipe, steps_per_launch=spl)
        J, cs, st = ctx.jacobian(s["x"], F0)
        print("pipeline", pipe, "steps_per_launch", spl, flush=True)
        assert st == 0
        assert np.array_equal(J, ref)


def test_schedule_validation():
    s = shape_c1(128)
    ctx = ctx_for(s)
    with pytest.raises(cc.CCError):
        ctx.set_schedule(threads_base=32)
    with pytest.raises(cc.CCError):
        ctx.set_schedule(steps_per_launch=7)
    sc = ctx.schedule()
    assert sc.threads_base == sc.threads_bulk == sc.threads_tangent == 64


@pytest.fixture(scope="module")
def c2_oracle_residual():
    s = shape_c2()
    o = oracle_for(s)
    F, st, (_, _, YLm, _) = o.shoot_states(s["x"])
    assert st == 0
    return s, F, np.max(np.abs(YLm))


@pytest.mark.parametrize("threads", [128, 64])
def test_residual_thread_counts_match_oracle(c2_oracle_residual, threads):
    s, F_ref, scale = c2_oracle_residual
    ctx = ctx_for(s)
    ctx.set_schedule(threads_base=threads)
    print("threads_base", threads, flush=True)
    F = ctx.residual(s["x"])
    assert np.max(np.abs(F - F_ref)) / scale < 1e-9


def test_tangent_thread_counts_agree(c2_oracle_residual):
    s, _, _ = c2_oracle_residual
    ctx = ctx_for(s, precision=cc.CC_MIXED)
    F0 = ctx.residual(s["x"])
    print("threads_tangent 128", flush=True)
    J128, cs, st = ctx.jacobian(s["x"], F0)
    assert st == 0 and not cs.any()
    ctx.set_schedule(threads_tangent=64)
    print("threads_tangent 64", flush=True)
    J64, cs, st = ctx.jacobian(s["x"], F0)
    assert st == 0 and not cs.any()
    cn = np.maximum(np.linalg.norm(J128, axis=0), 1e-4)
    err = np.linalg.norm(J64 - J128, axis=0) / cn
    print("column difference: max", err.max(), "median", np.median(err))
    assert np.median(err) < 1e-4 and err.max() < 1e-2


def test_bulk_thread_counts_agree(c2_oracle_residual):
    """FP64 Jacobian with either thread count of the bulk chains."""
    s, F_ref, _ = c2_oracle_residual
    ctx = ctx_for(s)
    print("threads_bulk 128", flush=True)
    J128, cs, st = ctx.jacobian(s["x"], F_ref)
    assert st == 0 and not cs.any()
    ctx.set_schedule(threads_bulk=64)
    print("threads_bulk 64", flush=True)
    J64, cs, st = ctx.jacobian(s["x"], F_ref)
    assert st == 0 and not cs.any()
    assert rel(J64, J128) < 1e-5


def test_autotune_keeps_results(c2_oracle_residual):
    """The tuned schedule is valid and the Newton step after tuning still has
    the oracle's residual."""
    s, F_ref, scale = c2_oracle_residual
    ctx = ctx_for(s, precision=cc.CC_MIXED)
    sc, times = ctx.autotune(s["x"])
    print("tuned", sc.as_dict(), "times", times, flush=True)
    assert len(times) >= 4 and np.all(times > 0)
    assert sc.threads_base in (128, 64) and sc.threads_tangent in (128, 64)
    F0, dx = ctx.newton_step(s["x"])
    assert np.max(np.abs(F0 - F_ref)) / scale < 1e-9
    assert np.all(np.isfinite(dx))


def test_profile_modes():
    """cc_profile 1 times the tangent and the base chains' launches, 2 only the
    tangent's (no events on the base stream), 0 none; none changes results."""
    s = shape_c1(128)
    ctx = ctx_for(s, precision=cc.CC_MIXED)
    F0, dx = ctx.newton_step(s["x"])
    for mode in (1, 2, 0):
        ctx.profile(mode != 0, base=mode == 1)
        F0b, dxb = ctx.newton_step(s["x"])
        assert np.array_equal(F0b, F0) and np.array_equal(dxb, dx)
        if mode == 0:
            with pytest.raises(cc.CCError):
                ctx.chain_timing()
            continue
        ms, launches, its, steps = ctx.chain_timing()
        assert ms > 0 and launches > 0 and its >= steps > 0
        if mode == 1:
            bms, bl, bits, bsteps = ctx.base_timing()
            assert bms > 0 and bl > 0 and bits >= bsteps > 0
        else:
            with pytest.raises(cc.CCError):
                ctx.base_timing()


def test_graph_replay_with_solo_launch():
    """At Ntau=512 the mixed step splits its slowest tangent chains into a
    launch of their own on a third stream (from the second Jacobian on): the
    captured step (three streams, fork and join by events) replays
    bit-identically to the eager step, three times with eager work between."""
    s = shape_c2()
    ctx = ctx_for(s, precision=cc.CC_MIXED)
    ctx.set_schedule(pipeline=1, steps_per_launch=12, use_graph=0)
    ctx.profile(True, base=False)
    ctx.newton_step(s["x"])                   # ranks the chains: the next steps split
    steps_all = ctx.chain_timing()[3]
    F0_e, dx_e = ctx.newton_step(s["x"])
    steps_main = ctx.chain_timing()[3]
    print("chain steps in the main launches", steps_all, "->", steps_main, flush=True)
    assert steps_main < steps_all             # the solo launch is active
    ctx.profile(False)
    ctx.set_schedule(use_graph=1)
    x2 = s["x"].copy()
    x2[3] += 1e-6
    for rep in range(3):
        F0_g, dx_g = ctx.newton_step(s["x"])
        print("replay", rep, flush=True)
        assert np.array_equal(F0_g, F0_e)
        assert np.array_equal(dx_g, dx_e)
        ctx.residual(x2)
        ctx.synchronize()
