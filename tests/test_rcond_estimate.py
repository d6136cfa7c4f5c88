"""The overlap conditioning estimate of k_overlap_rcond against LAPACK dpocon.

k_overlap_rcond restates the Hager-Higham iteration of LAPACK's dlacn2 (the
1-norm estimator dpocon runs) with S^{-1} v = X (X^T v), X = L^{-T}.  A line
by line Python transcription of the kernel's control flow is run here against
scipy's dpocon (the LAPACK routine the reference calls, eigensolver.cpp:2388-
2402) on random SPD matrices, graded matrices and overlap matrices with a
near-dependent pair of basis functions (the case the 1e-12 gate exists for).
"""
import numpy as np
import pytest
from scipy.linalg import lapack

from oracle import oracle as O
from xtbloom_amd import workloads


def kernel_rcond(S):
    """Transcription of k_overlap_rcond (same steps, same order of decisions)."""
    n = S.shape[0]
    anorm = np.abs(S).sum(axis=0).max()
    c, info = lapack.dpotrf(S, lower=1)
    assert info == 0
    X = np.linalg.inv(np.tril(c)).T  # upper triangular, X X^T = S^{-1}

    def apply(x):
        return X @ (X.T @ x)

    def iamax(v):
        return int(np.argmax(np.abs(v)))  # first index of the maximum

    x = apply(np.full(n, 1.0 / n))
    if n == 1:
        est = abs(x[0])
    else:
        est = np.abs(x).sum()
        sgn = np.where(x >= 0.0, 1.0, -1.0)
        x = apply(sgn)
        j = iamax(x)
        it = 2
        while True:
            e = np.zeros(n)
            e[j] = 1.0
            x = apply(e)
            estold, est = est, np.abs(x).sum()
            new = np.where(x >= 0.0, 1.0, -1.0)
            if np.array_equal(new, sgn) or est <= estold:
                break
            sgn = new
            x = apply(sgn)
            jlast, j = j, iamax(x)
            if not (x[jlast] != abs(x[j]) and it < 5):
                break
            it += 1
        alt = np.array([(-1.0 if i & 1 else 1.0) * (1.0 + i / (n - 1)) for i in range(n)])
        temp = 2.0 * (np.abs(apply(alt)).sum() / (3.0 * n))
        est = max(est, temp)
    return (1.0 / est) / anorm


def lapack_rcond(S):
    anorm = np.abs(S).sum(axis=0).max()
    c, info = lapack.dpotrf(S, lower=1)
    assert info == 0
    rc, info = lapack.dpocon(c, anorm, uplo="L")
    assert info == 0
    return rc


def overlap(Z, xyz):
    Z = np.ascontiguousarray(Z, np.int32)
    xyz = np.ascontiguousarray(xyz, float)
    n = O.basis_counts(Z)[1]
    S = np.zeros((n, n))
    D = np.zeros((3, n, n))
    Q = np.zeros((6, n, n))
    O.lib().oracle_integrals(len(Z), O._p(Z, O._c_int32_p), O._p(xyz), O._p(S), O._p(D), O._p(Q))
    return S


def near_dependent(d):
    """C6H14 with an extra H at distance d from one H (and one more H for parity)."""
    base = workloads.alkane(6, np.random.default_rng(3))
    Z0, X0 = base["Z"], base["xyz"]
    h = np.where(Z0 == 1)[0]
    Z = np.concatenate([Z0, [1, 1]])
    xyz = np.vstack([X0, X0[h[0]] + [d, 0.0, 0.0], X0[h[3]] + [0.0, 0.0, 0.5]])
    return Z, xyz


@pytest.mark.parametrize("seed", range(6))
def test_random_spd(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 60))
    A = rng.standard_normal((n, n))
    # spread the spectrum over a few decades
    q, _ = np.linalg.qr(A)
    lam = 10.0 ** rng.uniform(-8, 0, n)
    S = (q * lam) @ q.T
    S = 0.5 * (S + S.T)
    assert kernel_rcond(S) == pytest.approx(lapack_rcond(S), rel=1e-6)


@pytest.mark.parametrize("d", [1e-2, 1e-4, 1e-5, 4.681e-6, 3e-6, 1.1e-6])
def test_overlap_with_near_dependent_pair(d):
    S = overlap(*near_dependent(d))
    k, ref = kernel_rcond(S), lapack_rcond(S)
    assert k == pytest.approx(ref, rel=1e-4)
    # the 1e-12 gate decides the same way away from its edge
    if abs(ref / 1e-12 - 1.0) > 1e-3:
        assert (k >= 1e-12) == (ref >= 1e-12)
