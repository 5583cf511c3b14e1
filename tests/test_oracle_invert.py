"""Pin the oracle's cho_invert (qml/math.py:33-58 + qml/fcho_solve.f90:54-73; LAPACK dpotrf and
dpotri, a third-party dependency of the reference, via scipy) before the GPU is checked
against it."""

import numpy as np

import oracle


def test_cho_invert_reference_fixture(cho_pair):
    # tests/test_math.py:54-70
    A_ref = cho_pair[0]
    A = A_ref.copy()
    Ai = oracle.cho_invert(A)
    assert np.array_equal(A, A_ref)
    assert np.allclose(np.matmul(A, Ai), np.eye(A.shape[0]), atol=1e-7)
    assert np.array_equal(Ai, Ai.T)


def test_cho_invert_closed_forms():
    d = np.array([4.0, 0.5, 2.0, 8.0])
    np.testing.assert_allclose(oracle.cho_invert(np.diag(d)), np.diag(1.0 / d), rtol=4e-16, atol=0)
    # Tridiagonal (2, -1): (T^-1)_ij = min(i, j) (n + 1 - max(i, j)) / (n + 1), 1-based.
    n = 40
    T = 2.0 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    i = np.arange(1, n + 1)
    expect = np.minimum.outer(i, i) * (n + 1 - np.maximum.outer(i, i)) / (n + 1)
    np.testing.assert_allclose(oracle.cho_invert(T), expect, rtol=1e-12, atol=1e-13)


def test_cho_invert_fortran_input_is_overwritten():
    rng = np.random.default_rng(1)
    G = rng.standard_normal((6, 6))
    A = np.asfortranarray(G @ G.T + 6 * np.eye(6))
    A0 = A.copy()
    Ai = oracle.cho_invert(A)
    assert Ai is A  # np.asfortranarray returns A itself
    np.testing.assert_allclose(Ai @ A0, np.eye(6), atol=1e-12)
