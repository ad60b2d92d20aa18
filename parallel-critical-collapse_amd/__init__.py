"""ctypes front end of libcc_gpu.so (the C-ABI in include/cc_gpu.h).

The library is the product: every method here calls a HIP entry point.  There
is no CPU fallback; importing this package without the built library raises.

    import importlib
    cc = importlib.import_module("parallel-critical-collapse_amd")
    ctx = cc.Context(cc.Problem.make(Ntau=128, NLeft=500, NRight=500, XMid=0.5))
    F = ctx.residual(x)
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libcc_gpu.so")
HEADER = os.path.join(ROOT, "include", "cc_gpu.h")

CC_OK, CC_IRK_NOCONV, CC_DELTA_MISMATCH, CC_SINGULAR, CC_BADARG, CC_DEVICE = range(6)
CC_FP64, CC_MIXED, CC_FP64_TANGENT, CC_FP32_FD = 0, 1, 2, 3
STATUS_NAMES = {0: "OK", 1: "IRK_NOCONV", 2: "DELTA_MISMATCH", 3: "SINGULAR", 4: "BADARG", 5: "DEVICE"}


class CCError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class Problem(C.Structure):
    """cc_problem: the SimulationConfig keys of the path plus the precision policy."""

    _fields_ = [
        ("Ntau", C.c_int32), ("SchemeIRK", C.c_int32), ("MaxIterIRK", C.c_int32),
        ("NLeft", C.c_int32), ("NRight", C.c_int32),
        ("Dim", C.c_double), ("XLeft", C.c_double), ("XMid", C.c_double), ("XRight", C.c_double),
        ("PrecisionIRK", C.c_double), ("EpsNewton", C.c_double),
        ("precision", C.c_int32), ("reuse_sides", C.c_int32),
        ("PrecisionIRK32", C.c_double), ("refine_tol", C.c_double),
        ("refine_iters", C.c_int32), ("reserved", C.c_int32),
    ]

    @classmethod
    def make(cls, Ntau, NLeft, NRight, Dim=4.0, XLeft=1e-3, XMid=0.5, XRight=0.999, SchemeIRK=2,
             MaxIterIRK=100, PrecisionIRK=1e-15, EpsNewton=1e-8, precision=CC_FP64, reuse_sides=0,
             PrecisionIRK32=0.0, refine_tol=0.0, refine_iters=0):
        return cls(Ntau, SchemeIRK, MaxIterIRK, NLeft, NRight, Dim, XLeft, XMid, XRight, PrecisionIRK,
                   EpsNewton, precision, reuse_sides, PrecisionIRK32, refine_tol, refine_iters, 0)

    @classmethod
    def from_shape(cls, shape, **kw):
        return cls.make(Ntau=shape.Ntau, NLeft=shape.NLeft, NRight=shape.NRight, Dim=shape.Dim,
                        XLeft=shape.XLeft, XMid=shape.XMid, XRight=shape.XRight,
                        SchemeIRK=shape.SchemeIRK, MaxIterIRK=shape.MaxIterIRK,
                        PrecisionIRK=shape.PrecisionIRK, EpsNewton=shape.EpsNewton, **kw)


class Schedule(C.Structure):
    """cc_schedule: thread counts per chain kernel and the stage schedule."""

    _fields_ = [
        ("threads_base", C.c_int32), ("threads_bulk", C.c_int32), ("threads_tangent", C.c_int32),
        ("steps_per_launch", C.c_int32), ("pipeline", C.c_int32), ("use_graph", C.c_int32),
        ("reserved", C.c_int32 * 2),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


CC_TUNE_THREADS, CC_TUNE_STAGES = 1, 2
CC_SPEC_DIFF, CC_SPEC_LAM, CC_SPEC_HALVE, CC_SPEC_DOUBLE, CC_SPEC_INHOM = range(5)


def declared_symbols():
    """Entry points declared in include/cc_gpu.h."""
    with open(HEADER) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:cc_status|const char \*)\s*(cc_\w+)\s*\(", text, re.M)))


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lp = C.POINTER(C.c_int64)
_vp = C.c_void_p
_lib = None


def load():
    """Load libcc_gpu.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python {HERE}/build.py` "
                          "(or __graft_entry__.build()) first")
    L = C.CDLL(LIB_PATH)
    L.cc_ctx_create.argtypes = [C.POINTER(Problem), C.POINTER(_vp)]
    L.cc_ctx_destroy.argtypes = [_vp]
    L.cc_last_error.argtypes = [_vp]
    L.cc_last_error.restype = C.c_char_p
    L.cc_grid.argtypes = [_vp, _dp]
    L.cc_counters.argtypes = [_vp, _lp]
    L.cc_last_jacobian.argtypes = [_vp, _dp]
    L.cc_residual.argtypes = [_vp, _dp, _dp]
    L.cc_jacobian.argtypes = [_vp, _dp, _dp, _dp, _ip]
    L.cc_solve.argtypes = [_vp, C.c_int32, _dp, _dp, _ip]
    L.cc_refinement.argtypes = [_vp, _dp, _ip]
    L.cc_newton_step.argtypes = [_vp, _dp, _dp, _dp]
    L.cc_newton_step_device_async.argtypes = [_vp, _vp, _vp, _vp]
    L.cc_synchronize.argtypes = [_vp]
    L.cc_fft_batched.argtypes = [_vp, _dp, _dp, C.c_int32, C.c_int32, C.c_int32]
    L.cc_rhs_eval.argtypes = [_vp, _dp, _dp, _dp, C.c_int32, C.c_int32]
    L.cc_integrate.argtypes = [_vp, _dp, _dp, C.c_int32, _ip, _ip, C.c_int32, C.c_int32, _lp, _ip, C.c_int32, _ip]
    L.cc_spectral.argtypes = [_vp, C.c_int32, _dp, _dp, C.c_int32, C.c_double, C.c_double, C.c_double]
    L.cc_state_fields.argtypes = [_vp, _dp, _dp, _dp, C.c_int32]
    L.cc_expansions.argtypes = [_vp, _dp, _dp, _dp]
    L.cc_profile.argtypes = [_vp, C.c_int32]
    L.cc_chain_timing.argtypes = [_vp, _dp, _ip, _lp, _lp]
    L.cc_base_timing.argtypes = [_vp, _dp, _ip, _lp, _lp]
    L.cc_shoot_fields.argtypes = [_vp, _dp, C.c_int32, C.c_int32, _ip, _dp, _dp]
    L.cc_newton_run.argtypes = [_vp, _dp, C.c_double, C.c_double, C.c_int32, _ip, _dp, _dp, _ip]
    L.cc_pack.argtypes = [_vp, _dp, _dp, C.c_int32]
    L.cc_unpack.argtypes = [_vp, _dp, _dp, C.c_int32]
    _pp = C.POINTER(_dp)
    L.cc_batch_create.argtypes = [C.POINTER(Problem), C.c_int32, C.POINTER(_vp)]
    L.cc_batch_destroy.argtypes = [_vp]
    L.cc_batch_last_error.argtypes = [_vp]
    L.cc_batch_info.argtypes = [_vp, _ip, _ip]
    L.cc_batch_newton_step.argtypes = [_vp, _pp, _pp, _pp, _ip]
    L.cc_batch_get_schedule.argtypes = [_vp, C.c_int32, _ip, C.POINTER(Schedule)]
    L.cc_batch_set_schedule.argtypes = [_vp, C.c_int32, C.POINTER(Schedule)]
    L.cc_batch_autotune.argtypes = [_vp, _pp, _dp, C.c_int32, _ip]
    L.cc_batch_counters.argtypes = [_vp, _lp]
    L.cc_get_schedule.argtypes = [_vp, C.POINTER(Schedule)]
    L.cc_set_schedule.argtypes = [_vp, C.POINTER(Schedule)]
    L.cc_autotune.argtypes = [_vp, _dp, C.c_int32, C.POINTER(Schedule), _dp, C.c_int32, _ip]
    for name in declared_symbols():
        getattr(L, name).restype = C.c_char_p if name.endswith("last_error") else C.c_int
    _lib = L
    return L


def _d(a):
    return a.ctypes.data_as(_dp)


def _arr(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def _vec(a, n, what):
    """A contiguous float64 vector of exactly n entries (the C side reads n)."""
    v = np.ascontiguousarray(a, dtype=np.float64)
    if v.ndim != 1 or v.size != n:
        raise ValueError(f"{what}: expected {n} values, got shape {v.shape}")
    return v


class Context:
    """A cc_ctx: device buffers, streams and tables for one problem."""

    def __init__(self, problem: Problem):
        L = load()
        self.problem = problem
        h = _vp()
        st = L.cc_ctx_create(C.byref(problem), C.byref(h))
        if st != CC_OK:
            raise CCError(st, "cc_ctx_create failed")
        self.h = h
        self.N = problem.Ntau
        self.M = self.N // 2
        self.n = 3 * self.N // 4

    def close(self):
        if getattr(self, "h", None):
            load().cc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, what):
        if st != CC_OK:
            raise CCError(st, f"{what}: {load().cc_last_error(self.h).decode()}")

    def grid(self):
        g = np.zeros(self.problem.NLeft + self.problem.NRight + 1)
        self._check(load().cc_grid(self.h, _d(g)), "cc_grid")
        return g

    def counters(self):
        out = np.zeros(3, np.int64)
        load().cc_counters(self.h, out.ctypes.data_as(_lp))
        return out

    def last_jacobian(self):
        J = np.zeros((self.n, self.n))
        self._check(load().cc_last_jacobian(self.h, _d(J)), "cc_last_jacobian")
        return J

    def newton_run(self, x, precision_newton, slow_error, max_iter):
        """cc_newton_run: the whole Newton loop on the device.  Returns
        (x, iterations, mismatch, history, outcome)."""
        x = _vec(x, self.n, "x").copy()
        its, out = C.c_int32(0), C.c_int32(0)
        err = C.c_double(0.0)
        hist = np.zeros(max_iter)
        self._check(load().cc_newton_run(self.h, _d(x), precision_newton, slow_error, int(max_iter), C.byref(its),
                                         C.byref(err), _d(hist), C.byref(out)), "cc_newton_run")
        nh = min(max_iter, its.value + 1) if out.value else max_iter
        return x, its.value, err.value, hist[:nh].copy(), out.value

    def shoot_fields(self, x, every):
        """Debug field snapshots of the base shoot (cc_shoot_fields): grid
        abscissae xs [k] and fields [k][4][Ntau] = A, U, V, F."""
        x = _vec(x, self.n, "x")
        cnt = C.c_int32(0)
        st = load().cc_shoot_fields(self.h, _d(x), int(every), 0, C.byref(cnt), None, None)
        if st != CC_BADARG or cnt.value < 1:
            self._check(st, "cc_shoot_fields")
        xs = np.zeros(cnt.value)
        fields = np.zeros((cnt.value, 4, self.N))
        self._check(load().cc_shoot_fields(self.h, _d(x), int(every), cnt.value, C.byref(cnt), _d(xs), _d(fields)),
                    "cc_shoot_fields")
        return xs, fields

    def residual(self, x):
        x = _vec(x, self.n, "x")
        F = np.zeros(self.n)
        self._check(load().cc_residual(self.h, _d(x), _d(F)), "cc_residual")
        return F

    def jacobian(self, x, F0, check=True):
        x, F0 = _vec(x, self.n, "x"), _vec(F0, self.n, "F0")
        J = np.zeros((self.n, self.n))
        cs = np.zeros(self.n, np.int32)
        st = load().cc_jacobian(self.h, _d(x), _d(F0), _d(J), cs.ctypes.data_as(_ip))
        if check:
            self._check(st, "cc_jacobian")
        return J, cs, st

    def solve(self, A, b):
        A = np.array(A, np.float64, order="C")
        b = np.array(b, np.float64)
        if A.ndim != 2 or A.shape[0] != A.shape[1] or b.shape != (A.shape[0],):
            raise ValueError(f"solve: A must be square and b match it, got {A.shape} and {b.shape}")
        info = C.c_int32(0)
        st = load().cc_solve(self.h, A.shape[0], _d(A), _d(b), C.byref(info))
        return b, info.value, A, st

    def refinement(self):
        """(relative residual of the last refined mixed solve, whether it was
        solved again with the FP64 LU)."""
        r, f = C.c_double(0.0), C.c_int32(0)
        self._check(load().cc_refinement(self.h, C.byref(r), C.byref(f)), "cc_refinement")
        return r.value, bool(f.value)

    def newton_step(self, x):
        x = _vec(x, self.n, "x")
        F0, dx = np.zeros(self.n), np.zeros(self.n)
        self._check(load().cc_newton_step(self.h, _d(x), _d(F0), _d(dx)), "cc_newton_step")
        return F0, dx

    def newton_step_device(self, d_x: int, d_F0: int, d_dx: int):
        """Device pointers of n doubles each; returns once the step's status is known."""
        self._check(load().cc_newton_step_device_async(self.h, _vp(d_x), _vp(d_F0), _vp(d_dx)),
                    "cc_newton_step_device_async")

    def profile(self, enable=True, base=True):
        """Time the chain launches (and, with base, the tangent policies' base chains)."""
        self._check(load().cc_profile(self.h, 0 if not enable else (1 if base else 2)), "cc_profile")

    def chain_timing(self):
        """(ms, launches, fixed-point iterations, IRK steps) of the last profiled launch set."""
        ms, nl = C.c_double(0), C.c_int32(0)
        its, steps = C.c_int64(0), C.c_int64(0)
        self._check(load().cc_chain_timing(self.h, C.byref(ms), C.byref(nl), C.byref(its), C.byref(steps)),
                    "cc_chain_timing")
        return ms.value, nl.value, its.value, steps.value

    def base_timing(self):
        """(ms, launches, fixed-point iterations, IRK steps) of the FP64 base
        chains in the last profiled step of a tangent-linear policy."""
        ms, nl = C.c_double(0), C.c_int32(0)
        its, steps = C.c_int64(0), C.c_int64(0)
        self._check(load().cc_base_timing(self.h, C.byref(ms), C.byref(nl), C.byref(its), C.byref(steps)),
                    "cc_base_timing")
        return ms.value, nl.value, its.value, steps.value

    def schedule(self) -> Schedule:
        sc = Schedule()
        self._check(load().cc_get_schedule(self.h, C.byref(sc)), "cc_get_schedule")
        return sc

    def set_schedule(self, sc: Schedule | None = None, **kw):
        """Install a schedule; keyword arguments change fields of the current one."""
        sc = sc if sc is not None else self.schedule()
        for k, v in kw.items():
            setattr(sc, k, int(v))
        self._check(load().cc_set_schedule(self.h, C.byref(sc)), "cc_set_schedule")
        return sc

    def autotune(self, x, flags=CC_TUNE_THREADS | CC_TUNE_STAGES):
        """Run the autotuner at x; returns (installed schedule, candidate timings in ms)."""
        x = _vec(x, self.n, "x")
        sc = Schedule()
        times = np.zeros(64)
        nt = C.c_int32(0)
        self._check(load().cc_autotune(self.h, _d(x), flags, C.byref(sc), _d(times), times.size, C.byref(nt)),
                    "cc_autotune")
        return sc, times[:nt.value].copy()

    def synchronize(self):
        self._check(load().cc_synchronize(self.h), "cc_synchronize")

    def fft(self, X, sign, fp32=False):
        X = _arr(np.atleast_2d(X), np.complex128)
        if X.ndim != 2 or X.shape[1] != self.N:
            raise ValueError(f"fft: expected signals of {self.N} values, got shape {X.shape}")
        out = np.zeros_like(X)
        self._check(load().cc_fft_batched(self.h, X.ctypes.data_as(_dp), out.ctypes.data_as(_dp),
                                          X.shape[0], sign, int(fp32)), "cc_fft_batched")
        return out

    def rhs(self, Y, xs, fp32=False):
        Y = _arr(np.atleast_2d(Y), np.complex128)
        xs = _arr(np.atleast_1d(xs))
        if Y.ndim != 2 or Y.shape[1] != self.M or xs.shape != (Y.shape[0],):
            raise ValueError(f"rhs: expected states of {self.M} modes and one abscissa each, got {Y.shape}, {xs.shape}")
        dY = np.zeros_like(Y)
        self._check(load().cc_rhs_eval(self.h, Y.ctypes.data_as(_dp), _d(xs), dY.ctypes.data_as(_dp),
                                       Y.shape[0], int(fp32)), "cc_rhs_eval")
        return dY

    def integrate(self, Y, grid, start, end, fp32=False, max_its=0, with_fail_idx=False):
        """Integrate states Y[b] from grid[start[b]] to grid[end[b]]; returns (Y, its,
        status), and the failing steps' start indices (-1: converged) with_fail_idx."""
        Y = np.array(np.atleast_2d(Y), np.complex128)
        grid = _arr(grid)
        start = _arr(np.atleast_1d(start), np.int32)
        end = _arr(np.atleast_1d(end), np.int32)
        if Y.ndim != 2 or Y.shape[1] != self.M or start.shape != (Y.shape[0],) or end.shape != (Y.shape[0],):
            raise ValueError(f"integrate: expected states of {self.M} modes with a start and end each, "
                             f"got {Y.shape}, {start.shape}, {end.shape}")
        its = np.zeros(Y.shape[0], np.int64)
        status = np.zeros(Y.shape[0], np.int32)
        fail = np.zeros(Y.shape[0], np.int32)
        self._check(load().cc_integrate(self.h, Y.ctypes.data_as(_dp), _d(grid), grid.size,
                                        start.ctypes.data_as(_ip), end.ctypes.data_as(_ip), Y.shape[0],
                                        int(fp32), its.ctypes.data_as(_lp), status.ctypes.data_as(_ip),
                                        int(max_its), fail.ctypes.data_as(_ip)),
                    "cc_integrate")
        return (Y, its, status, fail) if with_fail_idx else (Y, its, status)

    def spectral(self, op, X, period=1.0, lam=0.0):
        """SpectralTransformer's operators on a batch (CC_SPEC_*): DIFF, LAM and
        HALVE take N complex values per signal, DOUBLE N/2 complex, INHOM a pair
        (f, g) of N real values ([batch][2][N])."""
        N = self.N
        shapes = {CC_SPEC_DIFF: ((N,), np.complex128, N, np.complex128),
                  CC_SPEC_LAM: ((N,), np.complex128, N, np.complex128),
                  CC_SPEC_HALVE: ((N,), np.complex128, N // 2, np.complex128),
                  CC_SPEC_DOUBLE: ((N // 2,), np.complex128, N, np.complex128),
                  CC_SPEC_INHOM: ((2, N), np.float64, N, np.float64)}
        if op not in shapes:
            raise ValueError(f"spectral: unknown operator {op}")
        ishape, idt, nout, odt = shapes[op]
        X = np.asarray(X)
        one = X.ndim == len(ishape)
        X = _arr(X[None] if one else X, idt)
        if X.shape[1:] != ishape:
            raise ValueError(f"spectral: expected signals of shape {ishape}, got {X.shape}")
        out = np.zeros((X.shape[0], nout), odt)
        lam = complex(lam)
        self._check(load().cc_spectral(self.h, int(op), X.ctypes.data_as(_dp), out.ctypes.data_as(_dp),
                                       X.shape[0], float(period), lam.real, lam.imag), "cc_spectral")
        return out[0] if one else out

    def state_fields(self, Y, xs):
        """States (M complex each) at abscissae xs -> [batch][4][N] = A, U, V, F."""
        Y = _arr(np.atleast_2d(Y), np.complex128)
        xs = _arr(np.atleast_1d(xs))
        if Y.ndim != 2 or Y.shape[1] != self.M or xs.shape != (Y.shape[0],):
            raise ValueError(f"state_fields: expected states of {self.M} modes and one abscissa each, "
                             f"got {Y.shape}, {xs.shape}")
        f = np.zeros((Y.shape[0], 4, self.N))
        self._check(load().cc_state_fields(self.h, Y.ctypes.data_as(_dp), _d(xs), _d(f), Y.shape[0]),
                    "cc_state_fields")
        return f

    def pack(self, Up, psic, fc):
        """Three real tau-series (N each, or [batch, N]) -> Newton vector(s)."""
        f = _arr(np.stack([np.atleast_2d(Up), np.atleast_2d(psic), np.atleast_2d(fc)], axis=1))
        if f.shape[2] != self.N:
            raise ValueError(f"pack: expected series of {self.N} values, got shape {f.shape}")
        x = np.zeros((f.shape[0], self.n))
        self._check(load().cc_pack(self.h, _d(f), _d(x), f.shape[0]), "cc_pack")
        return x[0] if np.ndim(Up) == 1 else x

    def unpack(self, x):
        """Newton vector(s) -> (Up, psic, fc), the Delta slot read as zero."""
        xb = _arr(np.atleast_2d(x))
        if xb.ndim != 2 or xb.shape[1] != self.n:
            raise ValueError(f"unpack: expected Newton vectors of {self.n} values, got shape {xb.shape}")
        f = np.zeros((xb.shape[0], 3, self.N))
        self._check(load().cc_unpack(self.h, _d(xb), _d(f), xb.shape[0]), "cc_unpack")
        if np.ndim(x) == 1:
            return f[0, 0], f[0, 1], f[0, 2]
        return f[:, 0], f[:, 1], f[:, 2]

    def expansions(self, x):
        x = _vec(x, self.n, "x")
        YL = np.zeros(self.M, np.complex128)
        YR = np.zeros(self.M, np.complex128)
        self._check(load().cc_expansions(self.h, _d(x), YL.ctypes.data_as(_dp), YR.ctypes.data_as(_dp)),
                    "cc_expansions")
        return YL, YR


def _ptrs(arrays):
    return (_dp * len(arrays))(*[a.ctypes.data_as(_dp) for a in arrays])


class Batch:
    """A cc_batch: independent problems stepped together (one launch per Ntau
    class and chunk)."""

    def __init__(self, problems):
        L = load()
        self.problems = list(problems)
        arr = (Problem * len(self.problems))(*self.problems)
        h = _vp()
        st = L.cc_batch_create(arr, len(self.problems), C.byref(h))
        if st != CC_OK:
            raise CCError(st, "cc_batch_create failed")
        self.h = h
        self.n = [3 * p.Ntau // 4 for p in self.problems]
        np_, nc = C.c_int32(0), C.c_int32(0)
        L.cc_batch_info(self.h, C.byref(np_), C.byref(nc))
        self.nclasses = nc.value

    def close(self):
        if getattr(self, "h", None):
            load().cc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, what):
        if st != CC_OK:
            raise CCError(st, f"{what}: {load().cc_batch_last_error(self.h).decode()}")

    def newton_step(self, xs, check=True):
        """xs: list of Newton vectors -> (F0 list, dx list, per-problem status, worst status)."""
        xs = self._inputs(xs)
        F0 = [np.zeros(n) for n in self.n]
        dx = [np.zeros(n) for n in self.n]
        status = np.zeros(len(xs), np.int32)
        st = load().cc_batch_newton_step(self.h, _ptrs(xs), _ptrs(F0), _ptrs(dx), status.ctypes.data_as(_ip))
        if check:
            self._check(st, "cc_batch_newton_step")
        return F0, dx, status, st

    def schedule(self, cls):
        sc, N = Schedule(), C.c_int32(0)
        self._check(load().cc_batch_get_schedule(self.h, cls, C.byref(N), C.byref(sc)), "cc_batch_get_schedule")
        return N.value, sc

    def set_schedule(self, cls, **kw):
        _, sc = self.schedule(cls)
        for k, v in kw.items():
            setattr(sc, k, int(v))
        self._check(load().cc_batch_set_schedule(self.h, cls, C.byref(sc)), "cc_batch_set_schedule")
        return sc

    def _inputs(self, xs):
        if len(xs) != len(self.n):
            raise ValueError(f"expected {len(self.n)} Newton vectors, got {len(xs)}")
        return [_vec(x, n, f"x[{i}]") for i, (x, n) in enumerate(zip(xs, self.n))]

    def autotune(self, xs):
        xs = self._inputs(xs)
        times = np.zeros(128)
        nt = C.c_int32(0)
        self._check(load().cc_batch_autotune(self.h, _ptrs(xs), _d(times), times.size, C.byref(nt)),
                    "cc_batch_autotune")
        return times[:nt.value].copy()

    def counters(self):
        out = np.zeros(3, np.int64)
        load().cc_batch_counters(self.h, out.ctypes.data_as(_lp))
        return out
