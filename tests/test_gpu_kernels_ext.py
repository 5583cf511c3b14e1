"""GPU parity for the linear, Sargan and Matern kernels and the p-distances against the oracle
(pinned bit-exactly to the reference Fortran). Bars: within 1e-12 relative (the linear kernel's
MFMA sums, device exp / pow in the last places; integer p = 1, 3, 4 use exact products); exact
symmetry for symmetric calls; the reference tests' own checks."""

import numpy as np
import pytest

import oracle
from conftest import bits

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)
    from qml_amd import _lib

    _lib.lib()


def _rel(a, b):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)) if a.size else 0.0


@pytest.fixture(scope="module")
def rows():
    rng = np.random.default_rng(17)
    return rng.random((300, 437)) * 5, rng.random((170, 437)) * 5


def test_linear_kernel(rows):
    from qml_amd.kernels import linear_kernel

    A, B = rows
    K = linear_kernel(A, B)
    assert K.shape == (300, 170) and K.flags["F_CONTIGUOUS"]
    assert _rel(K, oracle.linear_kernel(A, B)) <= 1e-12
    Ks = linear_kernel(A, A)
    assert np.array_equal(bits(Ks), bits(Ks.T))
    assert _rel(Ks, oracle.linear_kernel(A, A)) <= 1e-12
    np.random.seed(666)  # tests/test_kernels.py:93-119
    X, Xs = np.random.rand(25, 1000), np.random.rand(20, 1000)
    assert np.allclose(linear_kernel(X, Xs), X @ Xs.T)


@pytest.mark.parametrize("ng", [0, 1, 2, 4])
def test_sargan_kernel(rows, ng):
    from qml_amd.kernels import sargan_kernel

    A, B = rows
    gammas = np.array([0.7, 0.2, 1.3, 0.05])[:ng]
    K = sargan_kernel(A, B, 150.0, gammas)
    assert _rel(K, oracle.sargan_kernel(A, B, 150.0, gammas)) <= 1e-12
    Ks = sargan_kernel(A, A, 150.0, gammas)
    assert np.array_equal(bits(Ks), bits(Ks.T))


@pytest.mark.parametrize("metric", ["l1", "l2"])
@pytest.mark.parametrize("order", [0, 1, 2])
def test_matern_kernel(rows, metric, order):
    from qml_amd.kernels import matern_kernel

    A, B = rows
    K = matern_kernel(A, B, 40.0, order, metric)
    assert _rel(K, oracle.matern_kernel(A, B, 40.0, order, metric)) <= 1e-12
    with pytest.raises(SystemExit):
        matern_kernel(A, B, 40.0, 0, "l3")


def test_matern_l2_order_beyond_two_takes_the_order_two_form(rows):
    # fmatern_kernel_l2 (qml/fkernels.f90:468-477): every order other than 0 and 1.
    from qml_amd.kernels import matern_kernel

    A, B = rows
    assert np.array_equal(bits(matern_kernel(A, B, 40.0, 3, "l2")),
                          bits(matern_kernel(A, B, 40.0, 2, "l2")))


@pytest.mark.parametrize("p", [1, 3, 4, 5, 3.0, 0.5, 1.5, 2.75])
def test_p_distance(rows, p):
    from qml_amd.distance import p_distance

    A, B = rows
    D = p_distance(A, B, p)
    ref = oracle.p_distance(A, B, p)
    if p == 1:  # the Manhattan sum, then pow(s, 1.0) = s: bit-exact
        assert np.array_equal(bits(D), bits(ref))
    else:  # device pow in the root (and the terms for non-product p)
        assert _rel(D, ref) <= 1e-12
    Ds = p_distance(A, A, p)
    assert np.array_equal(bits(Ds), bits(Ds.T))


def test_p_distance_dispatch_and_errors():
    from qml_amd.distance import l2_distance, p_distance

    rng = np.random.default_rng(4)
    v1, v2 = rng.random((7, 5)), rng.random((9, 5))  # tests/test_distance.py:73-97
    D = p_distance(v1, v2, 3)
    expect = (np.abs(v1[:, None, :] - v2[None, :, :]) ** 3).sum(-1) ** (1.0 / 3)
    assert np.allclose(D, expect)
    assert np.allclose(p_distance(v1, v2, 3.0), D)
    assert np.array_equal(bits(p_distance(v1, v2, 2)), bits(l2_distance(v1, v2)))
    with pytest.raises(ValueError):
        p_distance(v1, v2, "3")
    with pytest.raises(ValueError):
        p_distance(v1, v2[:, :4], 3)
