"""Bunch-Kaufman on the GPU (qml/math.py:100-164) against LAPACK dsytrf / dsytrs (the oracle).

Bars: the reference tests' own checks on its fixture (tests/test_math.py:73-110); the same
pivots as LAPACK's unblocked factorisation below order 256 and as its blocked one (the
reference's call) from 256 on, factors within 1e-9; solutions and inverses within 1e-8
relative of LAPACK's on well-conditioned indefinite matrices (1e-5 at the fixture's condition
number 4.4e9); singular matrices report LAPACK's info."""

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)
    from qml_amd import _lib

    _lib.lib()


def _factor(A):
    from qml_amd.math import bkf_factor_device

    n = A.shape[0]
    Ad = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
    ipiv, info = bkf_factor_device(Ad, n)
    return np.tril(Ad.cpu().numpy().T), ipiv.cpu().numpy()[:n], int(info.item())


@pytest.mark.parametrize("n", [1, 2, 3, 17, 40, 63, 255])
def test_bkf_factor_matches_lapack_pivots(n):
    # Orders below 256 run unblocked in one workgroup, as LAPACK's dsytf2.
    rng = np.random.default_rng(n)
    for trial in range(6 if n < 100 else 2):
        M = rng.standard_normal((n, n))
        A = M + M.T
        if trial % 2:
            A[np.diag_indices(n)] *= 0.01  # small diagonal: 2 x 2 pivots
        L, piv, info = _factor(A)
        Lr, pivr, infor = oracle.bkf_factor(A)
        assert np.array_equal(piv, pivr) and info == infor
        assert np.max(np.abs(L - Lr)) <= 1e-9 * max(1.0, np.max(np.abs(Lr)))


@pytest.mark.parametrize("n", [256, 300, 411, 1024, 1025, 1090, 1500, 2000])
def test_bkf_blocked_matches_lapack_pivots(n):
    # From order 256 the factorisation runs in 64-column panels (dlasyf), as the reference's
    # dsytrf call does (lwork = n nb): the same pivots as LAPACK's blocked code, factors within
    # 1e-9 of the largest entry.
    rng = np.random.default_rng(7 + n)
    for trial in range(2):
        M = rng.standard_normal((n, n))
        A = M + M.T
        if trial % 2:
            A[np.diag_indices(n)] *= 0.01
        L, piv, info = _factor(A)
        Lr, pivr, infor = oracle.bkf_factor_blocked(A)
        assert np.array_equal(piv, pivr) and info == infor == 0
        assert np.max(np.abs(L - Lr)) <= 1e-9 * max(1.0, np.max(np.abs(Lr)))


@pytest.mark.parametrize("n,lda", [(300, 307), (1100, 1152)])
def test_bkf_blocked_leading_dimension(n, lda):
    # The C ABI takes lda >= n: the factor of a matrix embedded in a wider column-major array
    # equals the factor of the same matrix stored densely, and the padding rows stay untouched.
    from qml_amd import _lib

    rng = np.random.default_rng(n)
    M = rng.standard_normal((n, n))
    A = M + M.T
    L0, piv0, info0 = _factor(A)
    big = np.full((n, lda), 7.0)  # row c of this (n, lda) view is column c, rows 0 .. lda-1
    big[:, :n] = A.T
    Ad = torch.from_numpy(big).cuda()
    Ld = _lib.lib()
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    ws = torch.empty((Ld.qmlh_bkf_workspace(n) + 7) // 8, dtype=torch.float64, device="cuda")
    _lib.check(Ld.qmlh_bkf_factor(_lib.ptr(Ad), n, lda, _lib.ptr(ipiv), _lib.ptr(info),
                                  _lib.ptr(ws), ws.numel() * 8, _lib.stream_handle(None)))
    out = Ad.cpu().numpy()
    assert int(info.item()) == info0 == 0
    assert np.array_equal(ipiv.cpu().numpy(), piv0)
    assert np.array_equal(np.tril(out[:, :n].T), L0)
    assert np.all(out[:, n:] == 7.0)


@pytest.mark.parametrize("zero_at", [5, 700, 1070])
def test_bkf_blocked_singular_reports_info(zero_at):
    # a zero pivot in the first panel, a later panel, and the unblocked tail
    n = 1100
    A = np.diag(np.arange(1.0, n + 1.0))
    A[zero_at, zero_at] = 0.0
    L, piv, info = _factor(A)
    Lr, pivr, infor = oracle.bkf_factor_blocked(A)
    assert info == infor == zero_at + 1
    assert np.array_equal(piv, pivr) and np.array_equal(L, Lr)


@pytest.mark.parametrize("n", [4, 300])
def test_bkf_singular_reports_info(n):
    A = np.diag(np.arange(1.0, n + 1.0))
    A[2, 2] = 0.0
    _, piv, info = _factor(A)
    assert info == oracle.bkf_factor(A)[2] == 3


def test_bkf_reference_fixture(cho_pair):
    from qml_amd.math import bkf_invert, bkf_solve

    K, y = cho_pair
    A = K.copy()
    Ai = bkf_invert(A)
    assert np.array_equal(A, K)
    assert np.allclose(np.matmul(A, Ai), np.eye(A.shape[0]), atol=1e-7)
    assert np.array_equal(Ai, Ai.T)
    A, yy = K.copy(), y.copy()
    x = bkf_solve(A, yy)
    assert np.allclose(A, K) and np.array_equal(yy, y)
    assert np.allclose(x, np.linalg.solve(K, y))
    xr = oracle.bkf_solve(K.copy(), y.copy())
    assert np.max(np.abs(x - xr)) / np.max(np.abs(xr)) <= 1e-5


@pytest.mark.parametrize("n", [5, 64, 65, 300, 1000, 1200])
def test_bkf_random_indefinite(n):
    from qml_amd.math import bkf_invert, bkf_solve

    rng = np.random.default_rng(100 + n)
    M = rng.standard_normal((n, n))
    A = M + M.T + np.diag(rng.choice([-1.0, 1.0], n) * 2 * np.sqrt(n))  # indefinite, conditioned
    y = rng.standard_normal(n)
    x = bkf_solve(A.copy(), y)
    xr = oracle.bkf_solve(A.copy(), y.copy())
    assert np.max(np.abs(x - xr)) / np.max(np.abs(xr)) <= 1e-8
    Ai = bkf_invert(A.copy())
    Air = oracle.bkf_invert(A.copy())
    assert np.max(np.abs(Ai - Air)) / np.max(np.abs(Air)) <= 1e-8
    assert np.array_equal(Ai, Ai.T)


def test_bkf_invert_fortran_input_is_overwritten():
    from qml_amd.math import bkf_invert

    P = np.asfortranarray(np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, -2.0]]))
    Ai = bkf_invert(P)
    assert Ai is P
    np.testing.assert_allclose(Ai, [[0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, -0.5]],
                               atol=1e-15)
