"""Pin the oracle's linear, Sargan and Matern kernels and p-distances
(qml/fkernels.f90:403-531, qml/fdistance.f90:83-155): the reference's own tests' brute-force
formulas (tests/test_kernels.py:93-216, tests/test_distance.py:73-97, same seeds and shapes),
committed outputs of the reference Fortran, the live reference build, and planted mistakes."""

import os

import numpy as np
import pytest

import oracle
from conftest import bits


def _rows(seed):
    np.random.seed(seed)
    return np.random.rand(25, 1000), np.random.rand(20, 1000)


def test_linear_reference_test():
    X, Xs = _rows(666)
    K = oracle.linear_kernel(X, Xs)
    np.testing.assert_allclose(K, X @ Xs.T, rtol=1e-13)
    Ks = oracle.linear_kernel(X, X)
    assert np.allclose(Ks, Ks.T)


@pytest.mark.parametrize("metric", ["l1", "l2"])
@pytest.mark.parametrize("order", [0, 1, 2])
def test_matern_reference_test(metric, order):
    X, Xs = _rows(order + 10 * (metric == "l2"))
    sigma = 100.0
    if metric == "l1":
        d = np.abs(X[:, None, :] - Xs[None, :, :]).sum(-1)
    else:
        d = np.sqrt(((X[:, None, :] - Xs[None, :, :]) ** 2).sum(-1))
    if order == 0:
        expect = np.exp(-d / sigma)
    elif order == 1:
        expect = np.exp(-np.sqrt(3) * d / sigma) * (1 + np.sqrt(3) * d / sigma)
    else:
        expect = np.exp(-np.sqrt(5) * d / sigma) * (1 + np.sqrt(5) * d / sigma
                                                    + 5.0 / 3 * d ** 2 / sigma ** 2)
    np.testing.assert_allclose(oracle.matern_kernel(X, Xs, sigma, order, metric), expect,
                               rtol=1e-12)


@pytest.mark.parametrize("ngamma", [0, 1, 2, 3])
def test_sargan_reference_test(ngamma):
    np.random.seed(666 + ngamma)
    gammas = np.random.random(ngamma)
    X, Xs = np.random.rand(25, 1000), np.random.rand(20, 1000)
    sigma = 100.0
    d = np.abs(X[:, None, :] - Xs[None, :, :]).sum(-1)
    factor = np.ones_like(d)
    for k, g in enumerate(gammas):
        factor += g / sigma ** (k + 1) * d ** (k + 1)
    np.testing.assert_allclose(oracle.sargan_kernel(X, Xs, sigma, gammas),
                               np.exp(-d / sigma) * factor, rtol=1e-12)
    # planted: powers starting at 0 instead of 1 must not pass
    bad = np.ones_like(d)
    for k, g in enumerate(gammas):
        bad += g / sigma ** k * d ** k
    if ngamma:
        assert not np.allclose(np.exp(-d / sigma) * bad, oracle.sargan_kernel(X, Xs, sigma, gammas))


def test_p_distance_reference_test():
    rng = np.random.default_rng(9)
    v1, v2 = rng.random((7, 5)), rng.random((9, 5))
    D = oracle.p_distance(v1, v2, 3)
    expect = (np.abs(v1[:, None, :] - v2[None, :, :]) ** 3).sum(-1) ** (1.0 / 3)
    np.testing.assert_allclose(D, expect, rtol=1e-13)
    np.testing.assert_allclose(oracle.p_distance(v1, v2, 3.0), D, rtol=0)
    np.testing.assert_allclose(oracle.p_distance(v1, v2, 1), oracle.manhattan_distance(v1, v2),
                               rtol=0)
    with pytest.raises(ValueError):
        oracle.p_distance(v1, v2, "3")


def test_ext_match_reference_fortran_vectors(ref_kernels_ext):
    v = ref_kernels_ext
    A, B, s = v["A"], v["B"], float(v["sigma"])
    assert np.array_equal(bits(oracle.linear_kernel(A, B)), bits(v["linear"]))
    for ng in (1, 2, 3):
        assert np.array_equal(bits(oracle.sargan_kernel(A, B, s, v["gammas"][:ng])),
                              bits(v[f"sargan_{ng}"]))
    for order in (0, 1, 2):
        assert np.array_equal(bits(oracle.matern_kernel(A, B, s, order, "l2")),
                              bits(v[f"matern_l2_{order}"]))
    for p in (1, 3, 5):
        assert np.array_equal(bits(oracle.p_distance(A, B, p)), bits(v[f"p_int_{p}"]))
    for p in (0.5, 2.5):
        assert np.array_equal(bits(oracle.p_distance(A, B, p)), bits(v[f"p_real_{p}"]))


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build oracle/_ref absent")
def test_ext_match_reference_build_live():
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    rng = np.random.default_rng(123)
    A, B = rng.random((19, 44)) * 7, rng.random((23, 44)) * 7
    assert np.array_equal(bits(oracle.linear_kernel(A, B)),
                          bits(oracle.ref_kernel_ext("linear", A, B)))
    for g in ([2.0], [0.3, 0.9, 0.1, 1.7]):
        assert np.array_equal(bits(oracle.sargan_kernel(A, B, 31.0, g)),
                              bits(oracle.ref_kernel_ext("sargan", A, B, 31.0, g)))
    for order in (0, 1, 2, 5):
        assert np.array_equal(bits(oracle.matern_kernel(A, B, 9.0, order, "l2")),
                              bits(oracle.ref_kernel_ext("matern_l2", A, B, 9.0, order=order)))
    for p in (1, 3, 4, 6):
        assert np.array_equal(bits(oracle.p_distance(A, B, p)),
                              bits(oracle.ref_kernel_ext("p_int", A, B, p=p)))
    for p in (0.7, 1.5, 3.25):
        assert np.array_equal(bits(oracle.p_distance(A, B, p)),
                              bits(oracle.ref_kernel_ext("p_real", A, B, p=p)))
