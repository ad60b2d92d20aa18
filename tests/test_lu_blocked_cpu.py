"""CPU restatement of the device LU's blocked algorithm (k_lu_panel,
k_lu_laswp, k_lu_trsm, k_lu_gemm, k_lu_solve), checked against the oracle's
dgesv (NewtonSolver.cpp:739-757): the same pivots, factors and solution, and
the data-dependent indices the kernels compute -- pivot rows, each panel's swap
list and the composed row permutation -- inside the ranges the kernels'
buffers assume, including a singular matrix and a column with no comparable
pivot candidate (all NaN), where the pivot search keeps the diagonal."""
import numpy as np
import pytest

import oracle as O

SWAP_MAX = 64           # rows one panel may move (LU_SWAP_MAX): 2 x the widest panel


def blocked_lu(A, NB):
    """Right-looking blocked LU as the kernels run it: returns the factors,
    the composed permutation perm (P b = b[perm]), LAPACK's info and the swap
    lists, asserting every index the kernels would address."""
    A = np.array(A, dtype=np.float64)
    n = A.shape[0]
    perm = np.arange(n)
    info = 0
    swaps = []
    for k0 in range(0, n, NB):
        m, w = n - k0, min(NB, n - k0)
        P = A[k0:, k0:k0 + w].copy()
        rowid = np.arange(m)
        for j in range(w):
            col = np.abs(P[j:, j])
            cand = np.flatnonzero(~np.isnan(col))
            # first row index of the largest |a| (idamax); none comparable: the diagonal
            i0 = j + int(cand[np.argmax(col[cand])]) if cand.size else j
            assert j <= i0 < m
            if not (abs(P[i0, j]) > 0) and info == 0:
                info = k0 + j + 1
            if i0 != j:
                rowid[[j, i0]] = rowid[[i0, j]]
                P[[j, i0], :] = P[[i0, j], :]
            pv = P[j, j]
            if pv != 0:
                P[j + 1:, j] = P[j + 1:, j] * (1.0 / pv)
            P[j + 1:, j + 1:w] -= np.outer(P[j + 1:, j], P[j, j + 1:w])
        A[k0:, k0:k0 + w] = P
        pnew = rowid.copy() if k0 == 0 else perm[k0 + rowid]
        perm[k0:] = pnew
        moved = [(k0 + r, k0 + int(rowid[r])) for r in range(m) if rowid[r] != r]
        assert len(moved) <= 2 * w <= SWAP_MAX
        assert all(k0 <= d < n and k0 <= s < n for d, s in moved)
        swaps.append(moved)
        # laswp on the columns outside the panel: every moved row read before any is written
        outside = np.r_[0:k0, k0 + w:n]
        vals = {s: A[s, outside].copy() for _, s in moved}
        for d, s in moved:
            A[d, outside] = vals[s]
        rest = m - w
        if rest > 0:
            L11 = np.tril(A[k0:k0 + w, k0:k0 + w], -1) + np.eye(w)
            A[k0:k0 + w, k0 + w:] = np.linalg.solve(L11, A[k0:k0 + w, k0 + w:])   # trsm
            A[k0 + w:, k0 + w:] -= A[k0 + w:, k0:k0 + w] @ A[k0:k0 + w, k0 + w:]  # gemm
    assert sorted(perm) == list(range(n))
    return A, perm, info, swaps


def lu_solve(LU, perm, b):
    n = LU.shape[0]
    y = np.asarray(b, np.float64)[perm].copy()
    for i in range(n):
        y[i] -= LU[i, :i] @ y[:i]
    for i in range(n - 1, -1, -1):
        y[i] = (y[i] - LU[i, i + 1:] @ y[i + 1:]) / LU[i, i]
    return y


def ipiv_to_perm(ipiv):
    """LAPACK's sequential interchanges (1-based) as the composed permutation."""
    p = np.arange(len(ipiv))
    for i, r in enumerate(ipiv):
        p[[i, r - 1]] = p[[r - 1, i]]
    return p


@pytest.mark.parametrize("n,NB", [(1, 16), (7, 16), (50, 16), (96, 16), (200, 32), (131, 32)])
def test_blocked_lu_is_dgesv(n, NB):
    rng = np.random.default_rng(n + NB)
    A = rng.standard_normal((n, n)) + 0.1 * np.eye(n)
    b = rng.standard_normal(n)
    LU, perm, info, _ = blocked_lu(A, NB)
    x_ref, info_ref, LU_ref, ipiv = O.dgesv(A, b)
    assert info == info_ref == 0
    assert np.array_equal(perm, ipiv_to_perm(ipiv))
    assert np.max(np.abs(LU - LU_ref)) < 1e-10 * np.max(np.abs(LU_ref))
    x = lu_solve(LU, perm, b)
    assert np.linalg.norm(x - x_ref) < 1e-10 * np.linalg.norm(x_ref)


def test_blocked_lu_singular_info():
    A = np.ones((40, 40))
    _, _, info, _ = blocked_lu(A, 16)
    _, info_ref, _, _ = O.dgesv(A, np.ones(40))
    assert info == info_ref == 2


@pytest.mark.parametrize("NB", [16, 32])
def test_blocked_lu_nan_column_keeps_indices_in_range(NB):
    """An all-NaN column (an overflowed tangent column) has no comparable
    pivot candidate: the search keeps the diagonal, every index stays inside
    the matrix and the column is reported like a zero pivot."""
    rng = np.random.default_rng(3)
    n = 96
    A = rng.standard_normal((n, n)) + np.eye(n)
    A[:, 5] = np.nan
    _, perm, info, swaps = blocked_lu(A, NB)
    assert info == 6
    assert all(0 <= d < n and 0 <= s < n for sw in swaps for d, s in sw)
