"""cc_dgesv_noalloc: the FP64 dense solve (the dgesv of NewtonSolver.cpp:
739-757) with its scratch in a fixed device array instead of context
buffers, against cc_solve of a CC_FP64 context (the same kernels): factors and
solution bit for bit, on two calls in a row, the second with a smaller order
(the array keeps the first call's contents).  Inputs and outputs are device
arrays the test allocates; the entry point allocates nothing."""
import ctypes as C

import numpy as np
import pytest

from _helpers import cc

pytestmark = pytest.mark.gpu

dev = __import__("importlib").import_module("parallel-critical-collapse_amd.device")


def noalloc(A, b):
    L = cc.load()
    L.cc_dgesv_noalloc.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.POINTER(C.c_int32)]
    n = A.shape[0]
    dA, db = dev.DeviceArray(np.ascontiguousarray(A)), dev.DeviceArray(np.ascontiguousarray(b))
    info = C.c_int32(-1)
    st = L.cc_dgesv_noalloc(n, C.c_void_p(dA.ptr), C.c_void_p(db.ptr), C.byref(info))
    LU, x = dA.download(), db.download()
    dA.free()
    db.free()
    return st, info.value, LU, x


def reference(A, b):
    ctx = cc.Context(cc.Problem.make(Ntau=2048, NLeft=10, NRight=10))
    x, info, LU, st = ctx.solve(A, b)
    ctx.close()
    return st, info, LU, x


def system(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, n)) + 0.1 * np.eye(n), rng.standard_normal(n)


def test_noalloc_solve_equals_cc_solve_twice():
    for n, seed in ((1100, 1), (200, 2)):          # the second call smaller than the first
        A, b = system(n, seed)
        st, info, LU, x = noalloc(A, b)
        st_r, info_r, LU_r, x_r = reference(A, b)
        print(f"n={n}: status {st}, info {info}", flush=True)
        assert st == st_r == 0 and info == info_r == 0
        assert np.array_equal(x, x_r), n
        assert np.array_equal(LU, LU_r), n
        assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 1e-10


def test_noalloc_singular_and_sizes():
    A = np.ones((5, 5))
    b = np.arange(5.0)
    st, info, _, x = noalloc(A, b)
    assert st == cc.CC_SINGULAR and info == 2
    assert np.array_equal(x, b)                       # no solution computed, b unchanged
    L = cc.load()
    L.cc_dgesv_noalloc.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.POINTER(C.c_int32)]
    info = C.c_int32(0)
    one = dev.DeviceArray(np.zeros(1))
    for n in (2049, 0, -1):                           # rejected before anything is launched or read
        assert L.cc_dgesv_noalloc(n, C.c_void_p(one.ptr), C.c_void_p(one.ptr), C.byref(info)) == cc.CC_BADARG
    one.free()
    # and a good call after the rejected ones, to the same bits as cc_solve
    A, b = system(96, 3)
    st, info, LU, x = noalloc(A, b)
    st_r, info_r, LU_r, x_r = reference(A, b)
    assert st == 0 and np.array_equal(x, x_r) and np.array_equal(LU, LU_r)
