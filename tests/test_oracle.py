"""Pins of the oracle (CPU restatement) before it is trusted as the checker.

Each test names what fixes the expected value: the reference's own golden
vectors (test/test_init_gen.cpp, test/test_lapack.cpp, test/test_fft.cpp),
a textbook definition or closed form, or an independent numpy implementation
of the README's equations (README.md:21-37).
"""
import numpy as np
import pytest

import oracle as O
from _helpers import golden, oracle_for, shape_c1


def rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


# --------------------------------------------------------------------------- FFT
@pytest.mark.parametrize("N", [16, 128, 1024])
def test_fft_conventions_are_the_dft(N):
    """forwardFFT = sum_j x_j e^{+2 pi i jk/N} / N, backwardFFT = sum_k X_k e^{-2 pi i jk/N}
    (SpectralTransformer.cpp:33-38: FFTW_BACKWARD plan for forward and vice versa)."""
    o = O.Oracle(O.Config(Ntau=N))
    rng = np.random.default_rng(1)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    j = np.arange(N)
    W = np.exp(2j * np.pi * (np.outer(j, j) % N) / N)     # exponent reduced mod N
    assert rel(o.forward_c2c(x), W @ x / N) < 1e-14
    assert rel(o.backward_c2c(x), W.conj() @ x) < 1e-14
    xr = x.real.copy()
    assert rel(o.forward_r2c(xr), W @ xr / N) < 1e-14
    assert rel(o.backward_c2r(x), (W.conj() @ x).real) < 1e-14


def test_fft_test_vectors_of_reference():
    """test/test_fft.cpp:16-112 at N=16, period 1.  The inline spectra there put the
    sine's coefficient at -i/2 on k=1 and the complex exponential's peak on k=1,
    which is the e^{-i} convention; the code uses e^{+i} (FFTW_BACKWARD), so the
    coefficients sit at +i/2 on k=1 and on k=N-1.  The test only passes with
    NDEBUG (CMakeLists.txt:84).  The golden Taylor-series vectors of
    test_init_gen.cpp, which depend on the sign through every odd derivative,
    fix the e^{+i} convention (test_left_right_expansions_golden)."""
    N = 16
    o = O.Oracle(O.Config(Ntau=N, Delta=1.0))
    j = np.arange(N)
    s = o.forward_r2c(np.sin(2 * np.pi * j / N))
    e = np.zeros(N, complex)
    e[1], e[N - 1] = 0.5j, -0.5j
    assert np.max(np.abs(s - e)) < 1e-15
    c = o.forward_c2c(np.exp(2j * np.pi * j / N))
    e = np.zeros(N, complex)
    e[N - 1] = 1.0
    assert np.max(np.abs(c - e)) < 1e-15
    # round trips (89-90) and halve -> double -> halve (94-112)
    assert np.max(np.abs(o.backward_c2r(s) - np.sin(2 * np.pi * j / N))) < 1e-15
    h = O.Oracle.halve_modes(c)
    assert np.max(np.abs(O.Oracle.halve_modes(O.Oracle.double_modes(h)) - h)) < 1e-15


def test_halve_double_index_maps():
    """SpectralTransformer.cpp:196-252 by brute force, including doubleModes'
    overwrite of out[3N/2] by the loop at 242-245 (full amplitude) while
    out[N/2] keeps the half amplitude."""
    n_in = 16
    a = np.arange(n_in) + 1j * (100 + np.arange(n_in))
    d = O.Oracle.double_modes(a)
    exp = np.zeros(2 * n_in, complex)
    exp[: n_in // 2] = a[: n_in // 2]
    exp[n_in // 2] = 0.5 * a[n_in // 2]
    exp[n_in + n_in // 2:] = a[n_in // 2:]
    assert np.array_equal(d, exp)
    b = np.arange(2 * n_in) + 1j * np.arange(2 * n_in) ** 2
    h = O.Oracle.halve_modes(b)
    m = n_in
    exp = np.concatenate([b[: m // 2], [b[m // 2] + b[3 * m // 2]], b[m + m // 2 + 1:]])
    assert np.array_equal(h, exp)


# ------------------------------------------------------------- spectral calculus
def test_differentiate_trig_polynomial():
    """d/dtau of a_k cos(k w tau) + b_k sin(k w tau), w = 2 pi / period, through
    forward -> differentiate -> backward (SpectralTransformer.cpp:140-158)."""
    N, period = 64, 3.44
    o = O.Oracle(O.Config(Ntau=N, Delta=period))
    tau = np.arange(N) * period / N
    w = 2 * np.pi / period
    f = 0.3 * np.cos(3 * w * tau) - 1.2 * np.sin(5 * w * tau) + 0.7
    df = -0.9 * w * np.sin(3 * w * tau) - 6.0 * w * np.cos(5 * w * tau)
    # tau_j = j Delta / N with the +i forward convention means x_j = sum X_k e^{-i k w tau_j}
    # so the spectral derivative -i m k0 reproduces d/dtau.
    got = o.backward_c2r(o.differentiate(o.forward_r2c(f), period))
    assert np.max(np.abs(got - df)) < 1e-12


@pytest.mark.parametrize("c,d", [(0.7, -1.3), (-2.0, 0.4)])
def test_solve_inhom_constant(c, d):
    """x' + c x + d = 0 periodic: x = -d/c (SpectralTransformer.cpp:264-292)."""
    N = 128
    o = O.Oracle(O.Config(Ntau=N, Delta=3.44))
    x = o.solve_inhom(np.full(N, c), np.full(N, d), 3.44)
    assert np.max(np.abs(x + d / c)) < 1e-13


def test_solve_inhom_cosine_forcing():
    """x' + c x + cos(w tau) = 0: x = -(c cos(w tau) + w sin(w tau)) / (c^2 + w^2)."""
    N, period, c = 128, 3.44, 0.9
    o = O.Oracle(O.Config(Ntau=N, Delta=period))
    tau = np.arange(N) * period / N
    w = 2 * np.pi / period
    # tau_j = j period/N; in this convention a forward/backward pair acts on
    # samples at tau_j, and d/dtau maps to -i m k0.
    x = o.solve_inhom(np.full(N, c), np.cos(w * tau), period)
    exp = -(c * np.cos(w * tau) + w * np.sin(w * tau)) / (c ** 2 + w ** 2)
    assert np.max(np.abs(x - exp)) < 1e-13


def test_solve_inhom_variable_coefficient_residual():
    """A variable coefficient f: check the ODE x' + f x + g = 0 by an independent
    spectral derivative of the result (numpy FFT)."""
    N, period = 256, 3.44
    o = O.Oracle(O.Config(Ntau=N, Delta=period))
    tau = np.arange(N) * period / N
    w = 2 * np.pi / period
    f = 1.1 + 0.4 * np.cos(w * tau) + 0.2 * np.sin(2 * w * tau)
    g = -0.5 + 0.3 * np.sin(w * tau)
    x = o.solve_inhom(f, g, period)
    m = np.fft.fftfreq(N, 1.0 / N)
    dx = np.real(np.fft.ifft(np.fft.fft(x) * (2j * np.pi * m / period)))
    assert np.max(np.abs(dx + f * x + g)) < 1e-11


def test_lam_integrate_modes():
    """(lambda - i m k0) G_m = f_m, DC and Nyquist rules (SpectralTransformer.cpp:165-189)."""
    N, period, lam = 32, 2.0, 0.75
    o = O.Oracle(O.Config(Ntau=N, Delta=period))
    rng = np.random.default_rng(3)
    fk = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    G = o.lam_integrate(fk, lam, period)
    k0 = 2 * np.pi / period
    m = np.where(np.arange(N) < N // 2, np.arange(N), np.arange(N) - N)
    exp = fk / (lam - 1j * m * k0)
    exp[N // 2] = lam / (lam ** 2 + (k0 * N / 2) ** 2) * fk[N // 2]
    assert rel(G, exp) < 1e-15
    G0 = o.lam_integrate(fk, 0.0, period)
    assert G0[0] == 0 and G0[N // 2] == 0


# --------------------------------------------- InitialConditionGenerator (golden)
def test_pack_unpack_golden():
    """test/test_init_gen.cpp:162-315: pack -> unpack of the Ntau=128 fields."""
    g = golden("init_gen_128.json")
    o = O.Oracle(O.Config(Ntau=128, Delta=g["Delta"]))
    up, ps, fc = o.unpack(o.pack(g["Up"], g["psic"], g["fc"]))
    assert np.max(np.abs(up - g["Up_unpack_ref"])) < 1e-14
    assert np.max(np.abs(ps - g["psic_unpack_ref"])) < 2e-13      # |psic| ~ 170
    assert np.max(np.abs(fc - g["fc_unpack_ref"])) < 1e-14


def test_left_right_expansions_golden():
    """test/test_init_gen.cpp:323-381: Taylor-series state vectors (first 64 entries)."""
    g = golden("init_gen_128.json")
    o = O.Oracle(O.Config(Ntau=128, Delta=g["Delta"]))
    up, ps, fc = o.unpack(o.pack(g["Up"], g["psic"], g["fc"]))
    YL = o.left_expansion(fc, ps, g["Delta"], XLeft=g["XLeft"])
    YR = o.right_expansion(up, g["Delta"], XRight=g["XRight"])
    rl = np.array(g["Yleft_ref"]) @ [1, 1j]
    rr = np.array(g["Yright_ref"]) @ [1, 1j]
    assert np.max(np.abs(YL - rl)) < 1e-15
    assert np.max(np.abs(YR - rr)) < 1e-15


def test_pack_matches_numpy_fixture():
    """The anchor fixtures were packed with numpy (tests/golden/make_golden.py);
    the oracle's pack of the same fields must agree."""
    g = golden("init_gen_128.json")
    o = O.Oracle(O.Config(Ntau=128))
    import importlib.util, os
    spec = importlib.util.spec_from_file_location(
        "make_golden", os.path.join(os.path.dirname(__file__), "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    assert rel(o.pack(g["Up"], g["psic"], g["fc"]), mg.pack_numpy(g["Up"], g["psic"], g["fc"])) < 1e-14


# ---------------------------------------------- RHS against the README equations
def _readme_rhs(Y, x, D, period):
    """dY/dx from README.md:21-37 with numpy only: fields from the state by the
    half-period split, abar as the periodic solution of
    d(abar)/dtau = A(tau) abar - (D-3) solved as a dense Fourier-Galerkin system,
    then dU/dx, dV/dx, df/dx, and Y' = FFT(U' + i(V' + f'))/N truncated."""
    M = Y.size
    N = 2 * M
    full = np.zeros(N, complex)
    full[: M // 2] = Y[: M // 2]
    full[M // 2] = 0.5 * Y[M // 2]
    full[N - M // 2:] = Y[M // 2:]
    full[2] = -full[N - 2].real + 1j * full[2].imag
    k0 = 2 * np.pi / period
    m = np.where(np.arange(N) < N // 2, np.arange(N), np.arange(N) - N)
    z = np.fft.fft(full)                       # backward transform: e^{-i}
    dz = np.fft.fft(np.where(np.arange(N) == N // 2, 0, -1j * m * k0 * full))
    H = N // 2
    U = 0.5 * (z.real - np.roll(z.real, -H)); U[H:] = -U[:H]
    V = 0.5 * (z.imag - np.roll(z.imag, -H)); V[H:] = -V[:H]
    F = 0.5 * (z.imag + np.roll(z.imag, -H)); F[H:] = F[:H]
    dU = 0.5 * (dz.real - np.roll(dz.real, -H)); dU[H:] = -dU[:H]
    dV = 0.5 * (dz.imag - np.roll(dz.imag, -H)); dV[H:] = -dV[:H]
    A = (D - 2) ** 3 / (8 * x) * ((F + x) * U ** 2 - (F - x) * V ** 2) + (D - 3)
    # Galerkin on samples tau_j: derivative operator in this convention
    j = np.arange(N)
    Wb = np.exp(-2j * np.pi * np.outer(j, np.arange(N)) / N)          # synthesis
    Dop = (Wb * (-1j * m * k0)) @ np.linalg.inv(Wb)
    abar = np.linalg.solve(Dop - np.diag(A), np.full(N, -(D - 3.0))).real
    a2 = 1.0 / abar
    Ux = (F * ((D - 2 - 2 * (D - 3) * a2) * U + (D - 2) * V) - 2 * x * dU) / (2 * x * (F + x))
    Vx = (F * ((D - 2 - 2 * (D - 3) * a2) * V + (D - 2) * U) + 2 * x * dV) / (2 * x * (F - x))
    fx = (D - 3) / x * (a2 - 1) * F
    Yp = np.fft.ifft(Ux + 1j * (Vx + fx))
    out = np.concatenate([Yp[: M // 2], [Yp[M // 2] + Yp[3 * M // 2]], Yp[N - M // 2 + 1:]])
    out[2] = 1j * out[2].imag
    return out


@pytest.mark.parametrize("side", ["left", "right"])
def test_rhs_matches_readme_equations(side):
    s = shape_c1(128, 100)
    o = oracle_for(s)
    x = s["x"]
    N = 128
    up, ps, fc = o.unpack(np.where(np.arange(x.size) == N // 2 + 2, 0.0, x))
    Y = o.left_expansion(fc, ps, x[N // 2 + 2]) if side == "left" else o.right_expansion(up, x[N // 2 + 2])
    xx = 0.01 if side == "left" else 0.98
    got = o.derivatives(Y, xx)
    exp = _readme_rhs(Y, xx, 4.0, x[N // 2 + 2])
    assert rel(got, exp) < 1e-9


# ---------------------------------------------------------------- IRK stepping
def test_irk_step_is_fourth_order():
    """Gauss-Legendre s=2 (ODEStepper.cpp:41-55) is 4th order: refining the grid
    by 2 shrinks the error of a fixed x-interval by ~16."""
    s = shape_c1(128, 100)
    o = oracle_for(s)
    x = s["x"]
    N = 128
    up, ps, fc = o.unpack(np.where(np.arange(x.size) == N // 2 + 2, 0.0, x))
    Y0 = o.left_expansion(fc, ps, x[N // 2 + 2])
    finals = []
    for nsteps in (8, 16, 32, 64):
        g = np.linspace(0.001, 0.004, nsteps + 1)
        Y, st, _ = o.integrate(Y0, g, 0, nsteps)
        assert st == 0
        finals.append(Y)
    e1 = np.linalg.norm(finals[0] - finals[1])
    e2 = np.linalg.norm(finals[1] - finals[2])
    e3 = np.linalg.norm(finals[2] - finals[3])
    assert 14 < e1 / e2 < 18 and 14 < e2 / e3 < 18


def test_irk_midpoint_is_second_order():
    """Gauss-Legendre s=1 (implicit midpoint, ODEStepper.cpp:29-39) is 2nd
    order: halving the step shrinks the error of a fixed x-interval by ~4."""
    s = shape_c1(128, 100)
    s["cfg"].scheme = 1
    o = oracle_for(s)
    x = s["x"]
    N = 128
    up, ps, fc = o.unpack(np.where(np.arange(x.size) == N // 2 + 2, 0.0, x))
    Y0 = o.left_expansion(fc, ps, x[N // 2 + 2])
    finals = []
    for nsteps in (8, 16, 32, 64):
        g = np.linspace(0.001, 0.004, nsteps + 1)
        Y, st, _ = o.integrate(Y0, g, 0, nsteps)
        assert st == 0
        finals.append(Y)
    errs = [np.linalg.norm(finals[i] - finals[i + 1]) for i in range(3)]
    assert 3.5 < errs[0] / errs[1] < 4.5 and 3.5 < errs[1] / errs[2] < 4.5


@pytest.mark.parametrize("scheme", [1, 2, 3])
def test_gauss_legendre_tableaus(scheme):
    """The tableaus (ODEStepper.cpp:29-78) are the Gauss-Legendre ones, which
    the order conditions fix uniquely: the nodes are the roots of the shifted
    Legendre polynomial P_s(2c-1), sum_i b_i c_i^(k-1) = 1/k for k <= 2s
    (B(2s)) and sum_j a_ij c_j^(k-1) = c_i^k / k for k <= s (C(s))."""
    s = shape_c1(128, 100)
    s["cfg"].scheme = scheme
    a, b, c = oracle_for(s).tableau()
    assert a.shape == (scheme, scheme)
    roots = np.sort((np.polynomial.legendre.legroots([0] * scheme + [1]) + 1) / 2)
    assert np.max(np.abs(np.sort(c) - roots)) < 1e-15
    for k in range(1, 2 * scheme + 1):
        assert abs(np.sum(b * c ** (k - 1)) - 1 / k) < 1e-15, k
    for k in range(1, scheme + 1):
        assert np.max(np.abs(a @ c ** (k - 1) - c ** k / k)) < 1e-15, k


def test_irk_nonconvergence_is_reported():
    """MaxIterIRK=1 cannot meet PrecisionIRK=1e-15 (ShootingSolver.cpp:69-82)."""
    s = shape_c1(128, 100)
    cfg = s["cfg"]
    cfg.MaxIterIRK = 1
    o = O.Oracle(cfg)
    F, st = o.shoot(s["x"])
    assert st == 1


# ------------------------------------------------------------ shoot / Newton
def test_shipped_solution_has_small_mismatch():
    """data/simulation_4D_128.json is a near-solution on its own grid (XMid=0.1,
    NLeft=6000, NRight=8000; its file says Converged=false): the mismatch there is
    ~1e-7 RMS, against ~2e-2 for the seeded perturbation of the same data, so any
    wrong term in the series, the RHS or the stepping shows up as O(1e-3) or more."""
    a = golden("anchor_4D_128.json")
    c = a["config"]
    cfg = O.Config(Ntau=128, XMid=c["XMid"], NLeft=c["NLeft"], NRight=c["NRight"], Delta=a["Delta"],
                   EpsNewton=c["EpsNewton"])
    F, st = O.Oracle(cfg).shoot(np.array(a["x"]))
    assert st == 0
    assert O.l2norm(F) < 1e-6


def test_mismatch_delta_guard():
    """Re mismatch[2] > PrecisionIRK aborts (ShootingSolver.cpp:167-178)."""
    o = O.Oracle(O.Config(Ntau=64))
    YL = np.zeros(32, complex); YR = np.zeros(32, complex)
    YL[2] = 3.44; YR[2] = 3.44 + 1e-9
    _, st = o.mismatch(YL, YR)
    assert st == 2
    YR[2] = 3.44
    mis, st = o.mismatch(YL, YR)
    assert st == 0 and mis[2].real == 3.44


def test_jacobian_column_is_a_shoot_and_directional_derivative():
    """Column i equals (shoot(x + eps e_i) - F0)/eps; J v matches a directional
    difference (linearity of the FD Jacobian to O(eps))."""
    s = shape_c1(128, 1000)
    o = oracle_for(s)
    x = s["x"]
    F0, st = o.shoot(x)
    assert st == 0
    cols = [0, 5, 17, 34, 40]
    J, status, cnt, worst = O.jacobian(s["cfg"], x, F0, cols=cols, nthreads=2)
    assert worst == 0
    for i in cols:
        xp = x.copy(); xp[i] += s["cfg"].EpsNewton
        Fp, _ = o.shoot(xp)
        assert np.array_equal(J[:, i], (Fp - F0) / s["cfg"].EpsNewton)
    v = np.zeros_like(x); v[cols] = [1.0, -0.5, 0.25, 2.0, -1.0]
    h = 1e-7
    Fv, _ = o.shoot(x + h * v)
    assert rel(J @ v, (Fv - F0) / h) < 1e-3


# ------------------------------------------------------------------ dgesv
def test_dgesv_golden_10x10():
    """test/test_lapack.cpp:26-74.  The expected vector 1..10 written there does
    not solve the system written there (A (1..10) = (42, 38, 36, ...) != b), so
    that assert only passes under NDEBUG; the system itself is still a fixture:
    the solve must reproduce numpy's LAPACK solution and the residual must vanish."""
    g = golden("lapack_10.json")
    A, b = np.array(g["A"]), np.array(g["b"])
    assert np.max(np.abs(A @ np.array(g["x"]) - b)) > 1.0
    x, info, _, _ = O.dgesv(A, b)
    assert info == 0
    assert np.max(np.abs(x - np.linalg.solve(A, b))) < 1e-13
    assert np.max(np.abs(A @ x - b)) < 1e-12


def test_dgesv_matches_scipy_lu():
    import scipy.linalg as sl
    rng = np.random.default_rng(7)
    A = rng.standard_normal((50, 50))
    b = rng.standard_normal(50)
    x, info, LU, ipiv = O.dgesv(A, b)
    lu, piv = sl.lu_factor(A)
    assert info == 0
    assert np.array_equal(ipiv - 1, piv)
    assert rel(LU, lu) < 1e-13
    assert rel(x, sl.lu_solve((lu, piv), b)) < 1e-13


def test_dgesv_singular_info():
    A = np.ones((4, 4))
    _, info, _, _ = O.dgesv(A, np.ones(4))
    assert info == 2


# ------------------------------------------------------------ grid and norm
@pytest.mark.parametrize("NL,NR,XL,XM,XR", [(500, 500, 1e-3, 0.5, 0.999), (4000, 4000, 1e-3, 0.5, 0.999),
                                             (3000, 3000, 1e-3, 0.1, 0.999), (7, 11, 0.02, 0.3, 0.95)])
def test_grid_is_uniform_in_logit(NL, NR, XL, XM, XR):
    """generateGrid (NewtonSolver.cpp:477-519) against what fixes the grid:
    the three anchors are exact, logit(x_j) = ln x_j - ln(1 - x_j) is evenly
    spaced on each block with NLeft resp. NRight segments, the grid is strictly
    increasing, and every point equals the logistic function 1 / (1 + e^{-z_j})
    of its logit node evaluated in 40-digit decimal arithmetic (a different
    formula, in a different precision) to 1e-14 relative: the double node
    z_j = z_L + j dz carries ~|z| eps of rounding, which moves x_j by a few
    ulps, while a wrong node index moves it by ~1e-2."""
    from decimal import Decimal, getcontext
    getcontext().prec = 40
    o = O.Oracle(O.Config(Ntau=128, NLeft=NL, NRight=NR, XLeft=XL, XMid=XM, XRight=XR))
    g = o.grid()
    assert g.size == NL + NR + 1
    assert g[0] == XL and g[NL] == XM and g[NL + NR] == XR
    assert np.all(np.diff(g) > 0)
    z = np.log(g) - np.log1p(-g)
    for lo, hi, m in ((0, NL, NL), (NL, NL + NR, NR)):
        dz = np.diff(z[lo:hi + 1])
        assert np.max(np.abs(dz - (z[hi] - z[lo]) / m)) < 1e-12

    def logit(v):
        v = Decimal(v)
        return (v / (1 - v)).ln()

    zl, zm, zr = logit(XL), logit(XM), logit(XR)
    for j in list(range(1, min(NL, 40))) + list(range(max(1, NL - 40), NL)) + \
            list(range(NL + 1, min(NL + NR, NL + 40))) + list(range(max(NL + 1, NL + NR - 40), NL + NR)):
        zj = zl + (zm - zl) * j / NL if j < NL else zm + (zr - zm) * (j - NL) / NR
        exact = float(1 / (1 + (-zj).exp()))
        assert abs(g[j] - exact) <= 1e-14 * exact, j


def test_l2norm_is_rms():
    """computeL2Norm (NewtonSolver.cpp:763-768) is sqrt(mean(v^2)): against an
    exactly rounded sum (math.fsum) on random vectors, and the cases the
    definition fixes: a constant vector, a single entry, the zero vector and
    homogeneity under scaling by a power of two."""
    import math
    rng = np.random.default_rng(11)
    for n in (1, 2, 96, 768, 1536):
        v = rng.standard_normal(n) * 10.0 ** rng.integers(-8, 3)
        exact = math.sqrt(math.fsum(float(a) * float(a) for a in v) / n)
        assert abs(O.l2norm(v) - exact) <= 1e-14 * exact
        assert O.l2norm(8.0 * v) == 8.0 * O.l2norm(v)
    assert O.l2norm(np.full(768, -3.25)) == 3.25
    assert O.l2norm(np.array([-7.5])) == 7.5
    assert O.l2norm(np.zeros(10)) == 0.0
