"""The tests' sequential block IC(0) (tests/_helpers.py ic0_dense), which the GPU IC(0) preconditioner is
checked against, pinned by what the mathematics fixes: on a block-tridiagonal pattern IC(0) has no dropped
fill and is the exact Cholesky factor; on a grid pattern (with dropped fill) L L^T reproduces A on every
pattern entry and L vanishes off the pattern; a matrix with a non-positive pivot is refused."""
import numpy as np
import pytest

from _helpers import block_pattern_of_upper, ic0_dense, upper_csr_to_dense


def test_block_tridiagonal_ic0_is_cholesky():
    rng = np.random.default_rng(3)
    n = 7
    T = np.zeros((2 * n, 2 * n))
    for u in range(n):
        T[2 * u, 2 * u] = rng.uniform(1.0, 2.0)
        T[2 * u + 1, 2 * u] = rng.normal()
        T[2 * u + 1, 2 * u + 1] = rng.uniform(1.0, 2.0)
        if u > 0:
            T[2 * u:2 * u + 2, 2 * u - 2:2 * u] = rng.normal(size=(2, 2))
    A = T @ T.T  # block tridiagonal, SPD, Cholesky factor T
    P = np.eye(n, dtype=bool) | np.eye(n, k=1, dtype=bool) | np.eye(n, k=-1, dtype=bool)
    L = ic0_dense(A, P)
    np.testing.assert_allclose(L, T, rtol=0, atol=1e-12)
    np.testing.assert_allclose(L, np.linalg.cholesky(A), rtol=0, atol=1e-12)


def test_grid_ic0_matches_a_on_the_pattern():
    rng = np.random.default_rng(5)
    m = 5
    n = m * m
    P = np.eye(n, dtype=bool)
    for i in range(m):
        for j in range(m):
            u = i * m + j
            for di, dj in ((0, 1), (1, 0), (1, 1)):
                if i + di < m and j + dj < m:
                    w = (i + di) * m + j + dj
                    P[u, w] = P[w, u] = True
    S = np.kron(P, np.ones((2, 2), bool))
    B = rng.normal(size=(2 * n, 2 * n)) * S
    A = 0.5 * (B + B.T) + np.diag(np.abs(B).sum(1) + np.abs(B).sum(0) + 1.0)
    A *= S
    L = ic0_dense(A, P)
    low = np.tril(S)
    assert np.all(L[~low] == 0.0)
    np.testing.assert_allclose((L @ L.T)[S], A[S], rtol=1e-12, atol=1e-12)
    # the dropped fill makes it inexact elsewhere: L L^T differs from A off the pattern
    assert np.abs((L @ L.T)[~S]).max() > 1e-6


def test_ic0_refuses_indefinite():
    A = np.diag([1.0, 1.0, -1.0, 1.0])
    with pytest.raises(ValueError):
        ic0_dense(A, np.ones((2, 2), bool))


def test_upper_csr_round_trip():
    ia = np.array([0, 2, 3, 5, 6])
    ja = np.array([0, 1, 1, 2, 3, 3])
    a = np.array([4.0, 1.0, 5.0, 6.0, 2.0, 7.0])
    A = upper_csr_to_dense(ia, ja, a)
    np.testing.assert_array_equal(A, A.T)
    assert A[1, 0] == 1.0 and A[2, 3] == 2.0 and A[3, 3] == 7.0
    np.testing.assert_array_equal(block_pattern_of_upper(ia, ja), np.eye(2, dtype=bool))
    ia2 = np.array([0, 2, 4, 5, 6])
    ja2 = np.array([0, 1, 1, 2, 2, 3])
    np.testing.assert_array_equal(block_pattern_of_upper(ia2, ja2), np.ones((2, 2), bool))
