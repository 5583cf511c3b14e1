"""CPU checks of the C-ABI boundary and of the host-side logic (no GPU compute)."""

import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from qml_amd import _lib


def _header_functions():
    text = open(os.path.join(ROOT, "include", "qmlhip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qmlh_[a-z0-9_]+)\s*\(", text)))


def test_library_loads_and_exports_every_declared_symbol():
    L = _lib.load_library()
    declared = _header_functions()
    assert declared, "no declarations parsed from include/qmlhip.h"
    for name in declared:
        assert hasattr(L, name), f"libqmlhip.so does not export {name}"
    # The ctypes table covers exactly the declared API.
    assert sorted(_lib.exported_symbols()) == declared


def test_version_and_error_text_without_gpu():
    L = _lib.load_library()
    assert L.qmlh_version() >= 100
    assert isinstance(L.qmlh_last_error(), bytes)


def test_argument_errors_are_reported_without_touching_the_gpu():
    L = _lib.load_library()
    # Invalid shapes are rejected on the host before any launch.
    rc = L.qmlh_pairwise(0, None, -1, 1, None, 1, 1, 1, 1.0, None, 1, 0, None)
    assert rc == -1 and b"negative" in L.qmlh_last_error()
    rc = L.qmlh_cm_batch(None, None, None, 1, 300, 1, None, 45150, None, None)
    assert rc == -1 and b"size" in L.qmlh_last_error()
    rc = L.qmlh_potrf(None, 10, 5, None, None)
    assert rc == -1 and b"lda" in L.qmlh_last_error()
    rc = L.qmlh_gaussian_mfma(None, 4, 6, None, None, 4, 8, None, 4, 1.0, None, 4, 0, None)
    assert rc == -1 and b"multiples of 4" in L.qmlh_last_error()
    # the lower-triangle flag only makes sense for a symmetric call
    rc = L.qmlh_pairwise(0, None, 4, 4, None, 4, 4, 4, 1.0, None, 4, _lib.LOWER, None)
    assert rc == -1 and b"QMLH_LOWER needs QMLH_SYMMETRIC" in L.qmlh_last_error()
    prm = (ctypes.c_double * 1)(1.0)
    rc = L.qmlh_pairwise_ext(_lib.PAIR_P_REAL, None, 4, 4, None, 4, 4, 4, prm, 1, None, 4,
                             _lib.LOWER, None)
    assert rc == -1 and b"QMLH_LOWER needs QMLH_SYMMETRIC" in L.qmlh_last_error()
    assert L.qmlh_predict_workspace(10, 1000) == 10 * 8 * 8


def test_dropins_validate_like_the_reference():
    from qml_amd import distance, math as qmath, representations

    with pytest.raises(ValueError, match="expected square matrix"):
        qmath.cho_solve(np.zeros((3, 4)), np.zeros(3))
    with pytest.raises(ValueError, match="same stride size"):
        qmath.cho_solve(np.eye(3), np.zeros(4))
    with pytest.raises(ValueError, match="dimension=2"):
        distance.manhattan_distance(np.zeros(3), np.zeros((2, 3)))
    with pytest.raises(ValueError, match="same size"):
        distance.l2_distance(np.zeros((2, 3)), np.zeros((2, 4)))
    with pytest.raises(SystemExit):
        representations.generate_coulomb_matrix(np.array([1, 1]), np.zeros((2, 3)), 5, "bogus")
    with pytest.raises(ValueError, match="exceeds size"):
        representations.generate_coulomb_matrix(np.ones(6), np.zeros((6, 3)), 5)


def test_copy_lower_to_upper_matches_reference_indexing():
    from qml_amd.math import _copy_lower_to_upper

    rng = np.random.default_rng(0)
    for n in (1, 2, 5, 300, 1000):
        A = rng.random((n, n))
        B = A.copy()
        i_lower = np.tril_indices_from(B)  # qml/math.py:94-95
        B.T[i_lower] = B[i_lower]
        _copy_lower_to_upper(A, block=128)
        assert np.array_equal(A, B)


def test_gaussian_expansion_bound():
    from qml_amd.kernels import expansion_ok

    # Coulomb matrices (|x|^2 up to ~6e5) at sigma = 4000: expansion is exact to ~1e-16.
    assert expansion_ok(6e5, 6e5, 435, 4000.0)
    # Same data at sigma = 10: the exponent error bound exceeds 1e-13, use direct differences.
    assert not expansion_ok(6e5, 6e5, 435, 10.0)


def test_product_path_does_not_import_the_oracle():
    pkg = os.path.join(ROOT, "qml_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith(".py"):
                src = open(os.path.join(dirpath, f)).read()
                assert not re.search(r"^\s*(import|from)\s+oracle\b", src, re.M), f
