"""The reference built with its own flags in full (setup.py:21 "-O3 -fopenmp -m64
-march=native") and gfortran's default FMA contraction (oracle/_ref_fma) against the
contraction-free restatement the GPU path is bit-exact with. Contraction may change the last
bits of sums of products (qml/frepresentations.f90:104-105, qml/fkernels.f90:365); the tests
bound that movement, check that it never changes the row-norm ordering or the zero pattern on
the synthetic molecules, and print how many results are bit-identical (all of them on the
host this was written on), so the GPU's bit-exact match with the contraction-free build
carries over to the reference as its own build produces it on an FMA host."""

import numpy as np
import pytest

import oracle
from qml_amd import synthetic

pytestmark = pytest.mark.skipif(not oracle.ref_available("_ref_fma"),
                                reason="reference FMA build not present (make -C oracle ref-fma)")


def _molecules(n=400, seed=5):
    tr, _, size = synthetic.workload("krr_serve_50k", n_train=n, n_test=0)
    return tr, size


def test_fma_build_coulomb_matrices_within_ulps_same_ordering():
    tr, size = _molecules()
    exact = 0
    worst = 0.0
    for m in range(tr.nmol):
        Z, R = tr.molecule(m)
        a = oracle.coulomb_matrix(Z, R, size, "row-norm")
        b = oracle.ref_coulomb_matrix(Z, R, size, "row-norm", build="_ref_fma")
        nz = a != 0
        assert np.array_equal(nz, b != 0)  # the same entries are zero padding
        rel = np.max(np.abs(a[nz] - b[nz]) / np.abs(a[nz])) if nz.any() else 0.0
        worst = max(worst, rel)
        exact += bool(np.array_equal(a.view(np.uint64), b.view(np.uint64)))
        # a different row order would move whole entries, far beyond rounding
        assert rel <= 8 * 2.0 ** -52, (m, rel)
    print(f"\n{exact} of {tr.nmol} Coulomb matrices bit-identical to the FMA build; "
          f"largest relative entry difference {worst:.2e}")


@pytest.mark.parametrize("kind,sigma", [("laplacian", 724.0), ("gaussian", 4000.0)])
def test_fma_build_kernels_within_rounding(kind, sigma):
    tr, size = _molecules(300)
    X = oracle.coulomb_matrix_batch(*tr.ragged(), size)
    A, B = X[:200], X[200:]
    k = oracle.laplacian_kernel(A, B, sigma) if kind == "laplacian" else \
        oracle.gaussian_kernel(A, B, sigma)
    kf = oracle.ref_kernel(kind, A, B, sigma, build="_ref_fma")
    rel = np.max(np.abs(k - kf) / k)
    print(f"\n{kind}: largest relative difference to the FMA build {rel:.2e}")
    assert rel <= 1e-13

