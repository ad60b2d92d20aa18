"""The lane-swap transforms of cc_lanefft.hpp, pinned on the host.

tests/native/lanefft_plan.hip prints, from the header's own compile-time
functions, the swap schedule of every (N, T) the chain kernels instantiate and
the spectral index each thread/register holds after a DIF (layout S).  Here a
numpy model executes that schedule the way the kernels do -- radix 2^s
butterflies over the low register bits, a register-bit/thread-bit exchange for
every swap (the same permutation whether DPP, permlane or LDS moves it), the
twiddle exponent taken from thread-index bits [0, lo) -- and the result must be
the DFT: DIF output at (t, r) equals X[k(t, r)], DIT of layout S returns N x in
layout L.  A wrong schedule, twiddle exponent or index map fails the DFT check.
The header's maps are also checked for bijectivity there, and the LDS positions
of the Hermitian separation for bank spread here."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def plans(tmp_path_factory):
    if not os.path.exists(HIPCC) and shutil.which("hipcc") is None:
        pytest.skip("hipcc not installed")
    exe = tmp_path_factory.mktemp("lanefft") / "lanefft_plan"
    src = os.path.join(HERE, "native", "lanefft_plan.hip")
    res = subprocess.run([HIPCC, "-std=c++17", "-O1", "--offload-host-only", "-o", str(exe), src],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-2000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout[-2000:]
    return [json.loads(line) for line in out.stdout.splitlines()]


def _tbit(c):
    """thread-index bit of a lane (0..5) or wave (8 + b) coordinate"""
    assert c < 16, c
    return c if c < 8 else 6 + (c - 8)


def _swap(D, m, tb):
    """exchange register bit m with thread-index bit tb (D: [T, P])"""
    T, P = D.shape
    t = np.arange(T)[:, None]
    r = np.arange(P)[None, :]
    rb = (r >> m) & 1
    cb = (t >> tb) & 1
    t2 = (t & ~(1 << tb)) | (rb << tb)
    r2 = (r & ~(1 << m)) | (cb << m)
    out = np.empty_like(D)
    out[t2 + 0 * r2, r2 + 0 * t2] = D
    return out


def _stage(D, lo, s, sign, dit):
    T, P = D.shape
    rad = 1 << s
    F = np.exp(sign * 2j * np.pi * np.outer(np.arange(rad), np.arange(rad)) / rad)
    e = np.arange(T) & ((1 << lo) - 1)
    w = np.exp(sign * 2j * np.pi * np.outer(e, np.arange(rad)) / (1 << (lo + s)))   # [T, rad]
    out = np.empty_like(D)
    for g in range(P // rad):
        u = D[:, g * rad:(g + 1) * rad]
        if dit:
            u = u * w
        y = u @ F.T
        if not dit:
            y = y * w
        out[:, g * rad:(g + 1) * rad] = y
    return out


def _dif(plan, x, sign):
    N, T = plan["N"], plan["T"]
    P = N // T
    D = x.reshape(P, T).T.copy()                    # layout L: element t + T r
    for d in range(plan["nd"]):
        if d > 0:
            for m, c in enumerate(plan["sw"][d]):
                D = _swap(D, m, _tbit(c))
        D = _stage(D, plan["lo"][d], plan["sz"][d], sign, dit=False)
    return D


def _dit(plan, D, sign):
    for d in reversed(range(plan["nd"])):
        D = _stage(D, plan["lo"][d], plan["sz"][d], sign, dit=True)
        if d > 0:
            for m in reversed(range(len(plan["sw"][d]))):
                D = _swap(D, m, _tbit(plan["sw"][d][m]))
    return D


def test_plans_cover_the_kernel_shapes(plans):
    shapes = {(p["N"], p["T"]) for p in plans}
    # default (N/4; 64 at N=128) and alternative (N/8) thread counts of every transform length
    assert shapes == {(128, 64), (256, 64), (512, 128), (512, 64), (1024, 256), (1024, 128), (2048, 512), (2048, 256)}
    assert all(p["bad"] == 0 for p in plans)
    lds = {(p["N"], p["T"]): p["nlds"] for p in plans}
    assert lds[(1024, 256)] == 1 and lds[(256, 64)] == 0 and lds[(2048, 512)] == 2


@pytest.mark.parametrize("sign", [+1, -1])
def test_schedules_compute_the_dft(plans, sign):
    rng = np.random.default_rng(5)
    for plan in plans:
        N, T = plan["N"], plan["T"]
        P = N // T
        x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
        X = np.fft.ifft(x) * N if sign > 0 else np.fft.fft(x)       # sum_j x_j e^{sign 2 pi i j k / N}
        k = np.array(plan["k"]).reshape(T, P)
        D = _dif(plan, x, sign)
        assert np.allclose(D, X[k], rtol=0, atol=1e-12 * N), (N, T, sign)
        back = _dit(plan, X[k], -sign)                              # layout S in, layout L out
        assert np.allclose(back, N * x.reshape(P, T).T, rtol=0, atol=1e-11 * N), (N, T, sign)


def test_mirror_positions_spread_over_banks(plans):
    # the Hermitian separation writes mode k at mirror64[k] and reads mirror64[N - k]: 8 consecutive
    # lanes (one LDS cycle of 16-byte elements) must hit 8 distinct bank slots on the write
    for plan in plans:
        N, T = plan["N"], plan["T"]
        P = N // T
        k = np.array(plan["k"]).reshape(T, P)
        pos = np.array(plan["mirror64"])
        for t0 in range(0, T, 8):
            for r in range(P):
                w = pos[k[t0:t0 + 8, r]] % 8
                assert len(set(w.tolist())) == 8, (N, T, t0, r)
