"""CPU checks of the C-ABI library and the seeded workload (no GPU calls)."""
import ctypes
import importlib
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from _helpers import ROOT, W, cc


def test_library_exports_every_declared_symbol():
    """Every entry point in include/cc_gpu.h is exported by libcc_gpu.so."""
    names = cc.declared_symbols()
    assert len(names) >= 15
    lib = ctypes.CDLL(cc.LIB_PATH)
    for name in names:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", cc.LIB_PATH], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
    assert set(names) <= exported


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(cc, "LIB_PATH", os.path.join(ROOT, "no_such_lib.so"))
    monkeypatch.setattr(cc, "_lib", None)
    with pytest.raises(ImportError):
        cc.load()


def test_perturbation_equals_real_space_series():
    """The packed perturbation of the workload is the pack (ICG.cpp:678-711) of
    the real tau-series it stands for."""
    N = 128
    shape = W.config(1)
    o = O.Oracle(O.Config(Ntau=N))
    xa, _ = W.anchor(N)
    Up, ps, fc = o.unpack(np.where(np.arange(3 * N // 4) == N // 2 + 2, 0.0, xa))
    rng = W.SplitMix64(shape.seed)
    j = np.arange(N)

    def series(ks):
        P = np.zeros(N)
        for k in ks:
            a = rng.normal() * 0.1 / k ** 2
            b = rng.normal() * 0.1 / k ** 2
            P += a * np.cos(2 * np.pi * k * j / N) + b * np.sin(2 * np.pi * k * j / N)
        return P

    PU, Pp, Pf = series(range(1, N // 4, 2)), series(range(1, N // 4, 2)), series(range(2, N // 4, 2))
    Z = o.pack(Up + PU, ps + Pp, fc + Pf)
    ref = xa + W.perturbation(N, shape.seed)
    mask = np.arange(3 * N // 4) != N // 2 + 2
    assert np.max(np.abs(Z[mask] - ref[mask])) < 1e-14


def test_workload_is_deterministic_and_shaped():
    x1 = W.initial_x(W.config(3))
    x2 = W.initial_x(W.config(3))
    assert np.array_equal(x1, x2) and x1.size == 768 and x1[1024 // 2 + 2] == 3.44
    probs = W.config4()
    assert len(probs) == 24
    assert {p.Ntau for p in probs} <= {128, 256, 512} and {p.X for p in probs} <= {2000, 4000, 6000}
    assert [round(p.Dim, 6) for p in probs[:16]] == [round(3.96 + 0.01 * (i % 15), 6) for i in range(16)]
    assert all(p.X >= 4000 for p in probs if p.Ntau == 512)


def test_grid_matches_oracle():
    s = W.config(1)
    cfg = O.Config(Ntau=s.Ntau, NLeft=s.NLeft, NRight=s.NRight, XMid=s.XMid)
    assert np.array_equal(W.grid(s), O.Oracle(cfg).grid())


def test_smoke_shape_shoots_in_the_oracle():
    """The smoke entry point's shape (Ntau=128, 1000 x-points) shoots in the
    oracle; 600 and 800 x-points fail its IRK iteration."""
    s = W.config(1)
    cfg = O.Config(Ntau=s.Ntau, NLeft=s.NLeft, NRight=s.NRight, XMid=s.XMid, Delta=s.Delta)
    F, st = O.Oracle(cfg).shoot(W.initial_x(s))
    assert st == 0 and np.all(np.isfinite(F))


@pytest.mark.slow
def test_batched_problems_survive_the_oracle_shoot():
    """Every problem of the batched workload shoots in the FP64 oracle at its
    recorded perturbation scale (the scale table's defining property)."""
    import oracle as O
    from concurrent.futures import ThreadPoolExecutor

    def shoot(sh):
        cfg = O.Config(Ntau=sh.Ntau, NLeft=sh.NLeft, NRight=sh.NRight, XMid=sh.XMid, XLeft=sh.XLeft,
                       XRight=sh.XRight, Dim=sh.Dim, Delta=sh.Delta, EpsNewton=sh.EpsNewton,
                       PrecisionIRK=sh.PrecisionIRK, MaxIterIRK=sh.MaxIterIRK)
        return O.Oracle(cfg).shoot(W.initial_x(sh))[1]

    shapes = W.config4()
    assert len(shapes) == 24 and {s.Ntau for s in shapes} == {128, 256, 512}
    with ThreadPoolExecutor(8) as ex:
        assert list(ex.map(shoot, shapes)) == [0] * 24
