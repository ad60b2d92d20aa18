"""ctypes view of the CPU restatement in oracle/cc_oracle.c.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker (or the timed CPU
baseline), never as the product path.

Arrays are numpy float64; complex arrays are numpy complex128 (interleaved
re/im in memory, which is what the C side expects).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_cc.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lp = C.POINTER(C.c_int64)


class Problem(C.Structure):
    """Mirror of cco_problem (the SimulationConfig keys the path reads)."""

    _fields_ = [
        ("Ntau", C.c_int32), ("scheme", C.c_int32), ("MaxIterIRK", C.c_int32),
        ("NLeft", C.c_int32), ("NRight", C.c_int32),
        ("Dim", C.c_double), ("XLeft", C.c_double), ("XMid", C.c_double),
        ("XRight", C.c_double), ("PrecisionIRK", C.c_double),
        ("EpsNewton", C.c_double), ("Delta", C.c_double),
    ]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB) or \
            os.path.getmtime(LIB) < os.path.getmtime(os.path.join(HERE, "cc_oracle.c")):
        subprocess.check_call(["make", "-s", "-C", HERE])
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.cco_create.restype = C.c_void_p
        L.cco_create.argtypes = [C.POINTER(Problem)]
        L.cco_destroy.argtypes = [C.c_void_p]
        L.cco_counters.argtypes = [C.c_void_p, _lp]
        L.cco_tableau.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.cco_tableau.restype = C.c_int
        L.cco_reset_counters.argtypes = [C.c_void_p]
        for name in ("cco_forward_r2c", "cco_backward_c2r", "cco_forward_c2c", "cco_backward_c2c"):
            getattr(L, name).argtypes = [C.c_void_p, _dp, _dp]
        L.cco_differentiate.argtypes = [C.c_void_p, _dp, _dp, C.c_size_t, C.c_double]
        L.cco_lam_integrate.argtypes = [C.c_void_p, _dp, _dp, C.c_size_t, C.c_double, C.c_double, C.c_double]
        L.cco_halve_modes.argtypes = [_dp, C.c_size_t, _dp]
        L.cco_double_modes.argtypes = [_dp, C.c_size_t, _dp]
        L.cco_solve_inhom.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_double]
        L.cco_left_expansion.argtypes = [C.c_void_p, C.c_double, _dp, _dp, _dp, C.c_double]
        L.cco_right_expansion.argtypes = [C.c_void_p, C.c_double, _dp, _dp, C.c_double]
        L.cco_pack.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.cco_unpack.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.cco_fields_to_state.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.cco_state_to_fields_d.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double]
        L.cco_state_to_fields.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.cco_compute_derivatives.argtypes = [C.c_void_p, _dp, _dp, C.c_double]
        L.cco_step_irk.restype = C.c_int
        L.cco_step_irk.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_int)]
        L.cco_integrate.restype = C.c_int
        L.cco_integrate.argtypes = [C.c_void_p, _dp, _dp, C.c_size_t, C.c_size_t, C.POINTER(C.c_int)]
        L.cco_mismatch.restype = C.c_int
        L.cco_mismatch.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.cco_generate_grid.argtypes = [C.POINTER(Problem), _dp]
        L.cco_shoot.restype = C.c_int
        L.cco_shoot.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.cco_shoot_states.restype = C.c_int
        L.cco_shoot_states.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.cco_jacobian.restype = C.c_int
        L.cco_jacobian.argtypes = [C.POINTER(Problem), _dp, _dp, _dp, _dp, _ip, C.c_int, C.c_int, _ip, _lp]
        L.cco_dgesv.restype = C.c_int
        L.cco_dgesv.argtypes = [C.c_int, _dp, _ip, _dp]
        L.cco_l2norm.restype = C.c_double
        L.cco_l2norm.argtypes = [_dp, C.c_size_t]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


@dataclass
class Config:
    Ntau: int
    Dim: float = 4.0
    XLeft: float = 1e-3
    XMid: float = 0.5
    XRight: float = 0.999
    NLeft: int = 500
    NRight: int = 500
    scheme: int = 2
    PrecisionIRK: float = 1e-15
    MaxIterIRK: int = 100
    EpsNewton: float = 1e-8
    Delta: float = 3.44

    def problem(self) -> Problem:
        return Problem(self.Ntau, self.scheme, self.MaxIterIRK, self.NLeft, self.NRight,
                       self.Dim, self.XLeft, self.XMid, self.XRight, self.PrecisionIRK,
                       self.EpsNewton, self.Delta)


class Oracle:
    """One oracle context (the reference's ICG + SpectralTransformer + stepper)."""

    def __init__(self, cfg: Config):
        self.cfg = cfg
        self._prob = cfg.problem()
        self.ctx = lib().cco_create(C.byref(self._prob))
        if not self.ctx:
            raise ValueError("bad oracle problem")
        self.N = cfg.Ntau
        self.M = cfg.Ntau // 2
        self.n = 3 * cfg.Ntau // 4

    def __del__(self):
        if getattr(self, "ctx", None):
            lib().cco_destroy(self.ctx)
            self.ctx = None

    # --- counters -----------------------------------------------------------
    def counters(self):
        out = np.zeros(3, np.int64)
        lib().cco_counters(self.ctx, out.ctypes.data_as(_lp))
        return out

    def tableau(self):
        """(a, b, c) of the Gauss-Legendre scheme in use."""
        a, b, c = np.zeros(9), np.zeros(3), np.zeros(3)
        s = lib().cco_tableau(self.ctx, _p(a), _p(b), _p(c))
        return a[:s * s].reshape(s, s), b[:s], c[:s]

    def reset_counters(self):
        lib().cco_reset_counters(self.ctx)

    # --- SpectralTransformer ------------------------------------------------
    def forward_r2c(self, x):
        x = np.ascontiguousarray(x, np.float64)
        out = np.zeros(self.N, np.complex128)
        lib().cco_forward_r2c(self.ctx, _p(x), _p(out))
        return out

    def backward_c2r(self, X):
        X = np.ascontiguousarray(X, np.complex128)
        out = np.zeros(self.N)
        lib().cco_backward_c2r(self.ctx, _p(X), _p(out))
        return out

    def forward_c2c(self, X):
        X = np.ascontiguousarray(X, np.complex128)
        out = np.zeros(self.N, np.complex128)
        lib().cco_forward_c2c(self.ctx, _p(X), _p(out))
        return out

    def backward_c2c(self, X):
        X = np.ascontiguousarray(X, np.complex128)
        out = np.zeros(self.N, np.complex128)
        lib().cco_backward_c2c(self.ctx, _p(X), _p(out))
        return out

    def differentiate(self, X, period):
        X = np.ascontiguousarray(X, np.complex128)
        out = np.zeros_like(X)
        lib().cco_differentiate(self.ctx, _p(X), _p(out), X.size, period)
        return out

    def lam_integrate(self, X, lam, period):
        X = np.ascontiguousarray(X, np.complex128)
        out = np.zeros_like(X)
        lib().cco_lam_integrate(self.ctx, _p(X), _p(out), X.size, complex(lam).real, complex(lam).imag, period)
        return out

    @staticmethod
    def halve_modes(X):
        X = np.ascontiguousarray(X, np.complex128)
        out = np.zeros(X.size // 2, np.complex128)
        lib().cco_halve_modes(_p(X), X.size, _p(out))
        return out

    @staticmethod
    def double_modes(X):
        X = np.ascontiguousarray(X, np.complex128)
        out = np.zeros(X.size * 2, np.complex128)
        lib().cco_double_modes(_p(X), X.size, _p(out))
        return out

    def solve_inhom(self, f, g, period):
        f = np.ascontiguousarray(f, np.float64)
        g = np.ascontiguousarray(g, np.float64)
        x = np.zeros(self.N)
        lib().cco_solve_inhom(self.ctx, _p(f), _p(g), _p(x), period)
        return x

    # --- InitialConditionGenerator -----------------------------------------
    def left_expansion(self, fc, psic, Delta, XLeft=None):
        Y = np.zeros(self.M, np.complex128)
        lib().cco_left_expansion(self.ctx, self.cfg.XLeft if XLeft is None else XLeft,
                                 _p(np.ascontiguousarray(fc, np.float64)),
                                 _p(np.ascontiguousarray(psic, np.float64)), _p(Y), Delta)
        return Y

    def right_expansion(self, Up, Delta, XRight=None):
        Y = np.zeros(self.M, np.complex128)
        lib().cco_right_expansion(self.ctx, self.cfg.XRight if XRight is None else XRight,
                                  _p(np.ascontiguousarray(Up, np.float64)), _p(Y), Delta)
        return Y

    def pack(self, odd1, odd2, even):
        Z = np.zeros(self.n)
        lib().cco_pack(self.ctx, _p(np.ascontiguousarray(odd1, np.float64)),
                       _p(np.ascontiguousarray(odd2, np.float64)),
                       _p(np.ascontiguousarray(even, np.float64)), _p(Z))
        return Z

    def unpack(self, Z):
        Z = np.ascontiguousarray(Z, np.float64)
        o1, o2, ev = np.zeros(self.N), np.zeros(self.N), np.zeros(self.N)
        lib().cco_unpack(self.ctx, _p(Z), _p(o1), _p(o2), _p(ev))
        return o1, o2, ev

    def fields_to_state(self, U, V, F):
        Y = np.zeros(self.M, np.complex128)
        lib().cco_fields_to_state(self.ctx, _p(np.ascontiguousarray(U, np.float64)),
                                  _p(np.ascontiguousarray(V, np.float64)),
                                  _p(np.ascontiguousarray(F, np.float64)), _p(Y))
        return Y

    def state_to_fields_d(self, Y, X):
        Y = np.ascontiguousarray(Y, np.complex128)
        outs = [np.zeros(self.N) for _ in range(7)]
        lib().cco_state_to_fields_d(self.ctx, _p(Y), *[_p(o) for o in outs], X)
        return outs  # U, V, F, IA2, dU, dV, dF

    def state_to_fields(self, Y):
        Y = np.ascontiguousarray(Y, np.complex128)
        outs = [np.zeros(self.N) for _ in range(3)]
        lib().cco_state_to_fields(self.ctx, _p(Y), *[_p(o) for o in outs])
        return outs

    # --- ODEStepper / ShootingSolver ----------------------------------------
    def derivatives(self, Y, x):
        Y = np.ascontiguousarray(Y, np.complex128)
        dY = np.zeros(self.M, np.complex128)
        lib().cco_compute_derivatives(self.ctx, _p(Y), _p(dY), x)
        return dY

    def step_irk(self, Y, Xin, Xout, maxIts=None):
        Y = np.ascontiguousarray(Y, np.complex128)
        out = np.zeros(self.M, np.complex128)
        its = C.c_int(0)
        conv = lib().cco_step_irk(self.ctx, _p(Y), _p(out), Xin, Xout,
                                  self.cfg.MaxIterIRK if maxIts is None else maxIts, C.byref(its))
        return out, bool(conv), its.value

    def integrate(self, Y, grid, start, end):
        Y = np.array(Y, np.complex128)
        fail = C.c_int(-1)
        st = lib().cco_integrate(self.ctx, _p(Y), _p(np.ascontiguousarray(grid, np.float64)),
                                 start, end, C.byref(fail))
        return Y, st, fail.value

    def mismatch(self, YL, YR):
        mis = np.zeros(self.M, np.complex128)
        st = lib().cco_mismatch(self.ctx, _p(np.ascontiguousarray(YL, np.complex128)),
                                _p(np.ascontiguousarray(YR, np.complex128)), _p(mis))
        return mis, st

    # --- NewtonSolver --------------------------------------------------------
    def grid(self):
        g = np.zeros(self.cfg.NLeft + self.cfg.NRight + 1)
        lib().cco_generate_grid(C.byref(self._prob), _p(g))
        return g

    def shoot(self, x, grid=None):
        grid = self.grid() if grid is None else grid
        F = np.zeros(self.n)
        st = lib().cco_shoot(self.ctx, _p(np.ascontiguousarray(grid, np.float64)),
                             _p(np.ascontiguousarray(x, np.float64)), _p(F))
        return F, st

    def shoot_states(self, x, grid=None):
        grid = self.grid() if grid is None else grid
        F = np.zeros(self.n)
        Ys = [np.zeros(self.M, np.complex128) for _ in range(4)]
        st = lib().cco_shoot_states(self.ctx, _p(np.ascontiguousarray(grid, np.float64)),
                                    _p(np.ascontiguousarray(x, np.float64)), _p(F),
                                    *[_p(y) for y in Ys])
        return F, st, Ys  # YL0, YR0, YLm, YRm


def jacobian(cfg: Config, x, F0, grid=None, cols=None, nthreads=0):
    """FD Jacobian with OpenMP over columns; returns (J, status, counters)."""
    prob = cfg.problem()
    n = 3 * cfg.Ntau // 4
    if grid is None:
        grid = np.zeros(cfg.NLeft + cfg.NRight + 1)
        lib().cco_generate_grid(C.byref(prob), _p(grid))
    J = np.zeros((n, n))
    status = np.zeros(n, np.int32)
    cnt = np.zeros(3, np.int64)
    colarr = None if cols is None else np.ascontiguousarray(cols, np.int32)
    worst = lib().cco_jacobian(C.byref(prob), _p(np.ascontiguousarray(grid, np.float64)),
                               _p(np.ascontiguousarray(x, np.float64)),
                               _p(np.ascontiguousarray(F0, np.float64)), _p(J),
                               None if colarr is None else colarr.ctypes.data_as(_ip),
                               0 if colarr is None else colarr.size, nthreads,
                               status.ctypes.data_as(_ip), cnt.ctypes.data_as(_lp))
    return J, status, cnt, worst


def dgesv(A, b):
    A = np.array(A, np.float64, order="C")
    b = np.array(b, np.float64)
    ipiv = np.zeros(A.shape[0], np.int32)
    info = lib().cco_dgesv(A.shape[0], _p(A), ipiv.ctypes.data_as(_ip), _p(b))
    return b, info, A, ipiv


def l2norm(v):
    v = np.ascontiguousarray(v, np.float64)
    return lib().cco_l2norm(_p(v), v.size)
