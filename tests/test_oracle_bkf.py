"""Pin the oracle's Bunch-Kaufman solve and inverse (qml/math.py:100-164 around
qml/fcho_solve.f90:76-150; LAPACK dsytrf / dsytrs, a third-party dependency of the reference,
via scipy) with the reference's own tests (tests/test_math.py:73-110) and closed forms on
indefinite matrices."""

import numpy as np

import oracle


def test_bkf_reference_fixture(cho_pair):
    K, y = cho_pair
    A = K.copy()
    Ai = oracle.bkf_invert(A)
    assert np.array_equal(A, K)
    assert np.allclose(np.matmul(A, Ai), np.eye(A.shape[0]), atol=1e-7)
    A, yy = K.copy(), y.copy()
    x = oracle.bkf_solve(A, yy)
    # the fixture is symmetric to 1e-12 only; bkf_solve symmetrises A from its lower triangle
    assert np.allclose(A, K) and np.array_equal(yy, y)
    assert np.allclose(x, np.linalg.solve(K, y))


def test_bkf_indefinite_closed_forms():
    # [[0, 1], [1, 0]] is its own inverse (a 2 x 2 pivot); diag(2, -4, 0.5) inverts entrywise.
    P = np.array([[0.0, 1.0], [1.0, 0.0]])
    np.testing.assert_array_equal(oracle.bkf_invert(P.copy()), P)
    np.testing.assert_allclose(oracle.bkf_solve(P.copy(), np.array([3.0, 5.0])), [5.0, 3.0])
    D = np.diag([2.0, -4.0, 0.5])
    np.testing.assert_allclose(oracle.bkf_invert(D.copy()), np.diag([0.5, -0.25, 2.0]))
    _, piv, info = oracle.bkf_factor(P)
    assert list(piv) == [-2, -2] and info == 0


def test_bkf_random_indefinite_matches_numpy():
    rng = np.random.default_rng(5)
    for n in (5, 40, 130):
        M = rng.standard_normal((n, n))
        A = M + M.T
        y = rng.standard_normal(n)
        np.testing.assert_allclose(oracle.bkf_solve(A.copy(), y), np.linalg.solve(A, y),
                                   rtol=1e-9, atol=1e-10)
        np.testing.assert_allclose(oracle.bkf_invert(A.copy()), np.linalg.inv(A), rtol=1e-8,
                                   atol=1e-9)


def test_bkf_blocked_and_unblocked_factors_agree():
    # LAPACK's blocked dsytrf (the reference's call) and its unblocked dsytf2 are one algorithm:
    # the same pivots and, to rounding, the same factors; reconstructing P L D L^T P^T gives A.
    rng = np.random.default_rng(11)
    for n in (65, 130, 300):
        for small_diag in (False, True):
            M = rng.standard_normal((n, n))
            A = M + M.T
            if small_diag:
                A[np.diag_indices(n)] *= 0.01
            Lb, pb, ib = oracle.bkf_factor_blocked(A)
            Lu, pu, iu = oracle.bkf_factor(A)
            assert np.array_equal(pb, pu) and ib == iu == 0
            assert np.max(np.abs(Lb - Lu)) <= 1e-10 * max(1.0, np.max(np.abs(Lu)))
            assert np.any(pb < 0)  # 2 x 2 pivots occur
    # a zero pivot inside the first panel and one in the unblocked tail
    for zero_at in (40, 180):
        A = np.diag(np.arange(1.0, 201.0))
        A[zero_at, zero_at] = 0.0
        assert oracle.bkf_factor_blocked(A)[2] == zero_at + 1
