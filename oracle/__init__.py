"""CPU oracle for the KRR hot path.

TEST INFRASTRUCTURE ONLY. Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline
leg may import this package, and only as the checker or as the timed CPU baseline. The product
path (``qml_amd``) never imports it and has no CPU fallback.

Contents
--------
* ``liboracle.so`` (oracle.c): C restatement of the reference Fortran loops
  (qml/frepresentations.f90:52-205 and 207-598, qml/fkernels.f90:23-196 and 340-400,
  qml/fdistance.f90:23-80), built with ``make -C oracle``.
* ``cho_solve``: qml/math.py:60-97 around qml/fcho_solve.f90:23-52, which is exactly LAPACK
  ``dpotrf('U')`` + ``dpotrs('U', nrhs=1)``. LAPACK is a third-party dependency of the
  reference (setup.py:24, unpinned ``-llapack``) and is absent from the reference tree; this
  oracle calls the same two routines from scipy's bundled LAPACK.
* ``ref``: the reference's own Fortran compiled from its sources by ``make -C oracle ref``
  into ``oracle/_ref/libqmlref.so`` (present only where the reference tree was available at
  build time). Used to pin the C restatement.

Pinning (see tests/test_oracle.py, tests/test_oracle_local.py): the reference's golden
Coulomb-matrix and atomic-Coulomb-matrix fixtures (tests/test_representations.py:74-132 data),
the local-kernel fixtures (tests/test_energy_krr_atomic_cmat.py data), the water known answer
(docs/source/examples.rst:37-56), the cho_solve fixture pair (tests/test_math.py:32-51),
closed forms and brute force on small inputs, and bit-for-bit agreement with ``ref``.
"""

from __future__ import annotations

import ctypes
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_int64_p = ctypes.POINTER(ctypes.c_int64)

_lib = None
_ref = None


def _ptr(a, typ=_c_double_p):
    return a.ctypes.data_as(typ)


def lib():
    """Load liboracle.so (built by ``make -C oracle``)."""
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run make -C oracle)")
        L = ctypes.CDLL(path)
        i64 = ctypes.c_int64
        for name in ("oracle_cm_sorted", "oracle_cm_unsorted"):
            f = getattr(L, name)
            f.argtypes = [_c_double_p, _c_double_p, i64, i64, _c_double_p]
            f.restype = ctypes.c_int
        L.oracle_cm_batch.argtypes = [_c_double_p, _c_double_p, _c_int64_p, i64, i64,
                                      ctypes.c_int, _c_double_p]
        L.oracle_cm_batch.restype = ctypes.c_int
        for name in ("oracle_laplacian", "oracle_gaussian"):
            f = getattr(L, name)
            f.argtypes = [_c_double_p, i64, _c_double_p, i64, i64, ctypes.c_double, _c_double_p]
            f.restype = None
        for name in ("oracle_manhattan", "oracle_l2"):
            f = getattr(L, name)
            f.argtypes = [_c_double_p, i64, _c_double_p, i64, i64, _c_double_p]
            f.restype = None
        dbl = ctypes.c_double
        L.oracle_atomic_cm.argtypes = [_c_double_p, _c_double_p, i64, _c_int64_p, i64, i64,
                                       ctypes.c_int, dbl, dbl, dbl, dbl, _c_double_p]
        L.oracle_atomic_cm.restype = i64
        L.oracle_local_kernels.argtypes = [ctypes.c_int, _c_double_p, _c_int64_p, i64,
                                           _c_double_p, _c_int64_p, i64, i64, _c_double_p, i64,
                                           _c_double_p]
        L.oracle_local_kernels.restype = None
        L.oracle_linear.argtypes = [_c_double_p, i64, _c_double_p, i64, i64, _c_double_p]
        L.oracle_linear.restype = None
        L.oracle_sargan.argtypes = [_c_double_p, i64, _c_double_p, i64, i64, dbl, _c_double_p,
                                    i64, _c_double_p]
        L.oracle_sargan.restype = None
        L.oracle_matern_l2.argtypes = [_c_double_p, i64, _c_double_p, i64, i64, dbl, ctypes.c_int,
                                       _c_double_p]
        L.oracle_matern_l2.restype = None
        L.oracle_p_distance.argtypes = [_c_double_p, i64, _c_double_p, i64, i64, dbl, ctypes.c_int,
                                        _c_double_p]
        L.oracle_p_distance.restype = None
        _lib = L
    return _lib


def ref_available(build: str = "_ref") -> bool:
    """Whether a reference build exists: "_ref" (make -C oracle ref) or "_ref_fma" (make -C
    oracle ref-fma: the reference's flags in full, -march=native, FMA contraction)."""
    return os.path.exists(os.path.join(_HERE, build, "libqmlref.so"))


_refs = {}


def ref(build: str = "_ref"):
    """Load a reference build (oracle/<build>/libqmlref.so)."""
    if build != "_ref":
        if build not in _refs:
            _refs[build] = _load_ref(build)
        return _refs[build]
    global _ref
    if _ref is None:
        _ref = _load_ref(build)
    return _ref


def _load_ref(build: str):
    path = os.path.join(_HERE, build, "libqmlref.so")
    if not os.path.exists(path):
        raise RuntimeError(f"reference build missing: {path} (run make -C oracle ref)")
    L = ctypes.CDLL(path)
    ci = ctypes.c_int
    L.ref_cm.argtypes = [_c_double_p, _c_double_p, ci, ci, ci, _c_double_p]
    L.ref_cm.restype = None
    L.ref_kernel.argtypes = [ci, _c_double_p, ci, _c_double_p, ci, ci, _c_double_p,
                             ctypes.c_double]
    L.ref_kernel.restype = None
    dbl = ctypes.c_double
    _c_int_p = ctypes.POINTER(ctypes.c_int)
    L.ref_atomic_cm.argtypes = [_c_double_p, _c_double_p, ci, _c_int_p, ci, ci, ci, dbl, dbl,
                                dbl, dbl, _c_double_p]
    L.ref_atomic_cm.restype = None
    L.ref_local_kernels.argtypes = [ci, _c_double_p, ci, _c_double_p, ci, ci, _c_int_p, ci,
                                    _c_int_p, ci, _c_double_p, ci, _c_double_p]
    L.ref_local_kernels.restype = None
    L.ref_kernel_ext.argtypes = [ci, _c_double_p, ci, _c_double_p, ci, ci, _c_double_p, dbl,
                                 _c_double_p, ci, ci, dbl, ci]
    L.ref_kernel_ext.restype = None
    return L


# --------------------------------------------------------------------------------------------
# Coulomb matrices
# --------------------------------------------------------------------------------------------

def coulomb_matrix(Z, R, size=23, sorting="row-norm"):
    """One molecule; mirrors qml/representations.py:69-119 (f64 coercion of Z and R)."""
    Z = np.ascontiguousarray(Z, dtype=np.float64)
    R = np.ascontiguousarray(R, dtype=np.float64).reshape(-1, 3)
    out = np.zeros(size * (size + 1) // 2)
    if sorting == "row-norm":
        rc = lib().oracle_cm_sorted(_ptr(Z), _ptr(R), Z.shape[0], size, _ptr(out))
    elif sorting == "unsorted":
        rc = lib().oracle_cm_unsorted(_ptr(Z), _ptr(R), Z.shape[0], size, _ptr(out))
    else:
        raise ValueError(sorting)
    if rc:
        raise ValueError("natoms > size")
    return out


def coulomb_matrix_batch(Z, R, offsets, size, sorted_=True):
    """Ragged batch (see qml_amd.synthetic); returns (nmol, size*(size+1)/2)."""
    Z = np.ascontiguousarray(Z, dtype=np.float64)
    R = np.ascontiguousarray(R, dtype=np.float64)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    nmol = offsets.shape[0] - 1
    out = np.zeros((nmol, size * (size + 1) // 2))
    rc = lib().oracle_cm_batch(_ptr(Z), _ptr(R), _ptr(offsets, _c_int64_p), nmol, size,
                               1 if sorted_ else 0, _ptr(out))
    if rc:
        raise ValueError("natoms > size")
    return out


def ref_coulomb_matrix(Z, R, size, sorting="row-norm", build="_ref"):
    """The reference's own fgenerate_*coulomb_matrix via oracle/_ref (or another build)."""
    Z = np.ascontiguousarray(Z, dtype=np.float64)
    R = np.ascontiguousarray(R, dtype=np.float64).reshape(-1, 3)
    out = np.zeros(size * (size + 1) // 2)
    ref(build).ref_cm(_ptr(Z), _ptr(R), Z.shape[0], size, 1 if sorting == "row-norm" else 0,
                      _ptr(out))
    return out


ACM_SORTING = {"row-norm": 0, "distance": 1}
NUCLEAR_CHARGE = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9,
                  "Ne": 10, "Na": 11, "Mg": 12, "Al": 13, "Si": 14, "P": 15, "S": 16, "Cl": 17,
                  "Ar": 18}


def _acm_indices(Z, indices):
    # qml/representations.py:205-221: None -> all atoms, an element symbol -> the atoms of that
    # element (empty -> (0, 0) result), else 0-based indices.
    if indices is None:
        return np.arange(Z.shape[0], dtype=np.int64)
    if isinstance(indices, str):
        if indices not in NUCLEAR_CHARGE:
            raise ValueError(f"unknown value {indices!r} for 'indices'")
        return np.where(Z == NUCLEAR_CHARGE[indices])[0].astype(np.int64)
    return np.asarray(indices, dtype=np.int64).reshape(-1)


def atomic_coulomb_matrix(Z, R, size=23, sorting="distance", central_cutoff=1e6,
                          central_decay=-1, interaction_cutoff=1e6, interaction_decay=-1,
                          indices=None):
    """One molecule; mirrors qml/representations.py:121-233 over the C restatement."""
    Z = np.ascontiguousarray(Z, dtype=np.float64)
    R = np.ascontiguousarray(R, dtype=np.float64).reshape(-1, 3)
    cent = _acm_indices(Z, indices)
    if isinstance(indices, str) and cent.size == 0:
        return np.zeros((0, 0))
    out = np.zeros((cent.size, size * (size + 1) // 2))
    rc = lib().oracle_atomic_cm(_ptr(Z), _ptr(R), Z.shape[0], _ptr(cent, _c_int64_p), cent.size,
                                size, ACM_SORTING[sorting], central_cutoff, central_decay,
                                interaction_cutoff, interaction_decay, _ptr(out))
    if rc:
        raise ValueError(f"{size} size set, but {rc} size needed")
    return out


def ref_atomic_coulomb_matrix(Z, R, size=23, sorting="distance", central_cutoff=1e6,
                              central_decay=-1, interaction_cutoff=1e6, interaction_decay=-1,
                              indices=None):
    """The reference's own fgenerate_{local,atomic}_coulomb_matrix via oracle/_ref."""
    Z = np.ascontiguousarray(Z, dtype=np.float64)
    R = np.ascontiguousarray(R, dtype=np.float64).reshape(-1, 3)
    cent = _acm_indices(Z, indices)
    if isinstance(indices, str) and cent.size == 0:
        return np.zeros((0, 0))
    cent1 = np.ascontiguousarray(cent + 1, dtype=np.int32)
    out = np.zeros((cent.size, size * (size + 1) // 2), order="F")
    ref().ref_atomic_cm(_ptr(Z), _ptr(R), Z.shape[0], _ptr(cent1, ctypes.POINTER(ctypes.c_int)),
                        cent.size, size, ACM_SORTING[sorting], float(central_cutoff),
                        float(central_decay), float(interaction_cutoff),
                        float(interaction_decay), _ptr(out))
    return np.ascontiguousarray(out)


def _local_prep(A, B, na, nb, sigmas):
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    na = np.asarray(na, dtype=np.int64).reshape(-1)
    nb = np.asarray(nb, dtype=np.int64).reshape(-1)
    assert np.sum(na) == A.shape[0] and np.sum(nb) == B.shape[0] and A.shape[1] == B.shape[1]
    sig = np.ascontiguousarray(np.asarray(sigmas, dtype=np.float64).reshape(-1))
    return A, B, na, nb, sig


def local_kernels(kind, A, B, na, nb, sigmas):
    """qml/kernels.py:218-298 over the C restatement: (nsig, nmA, nmB), Fortran order."""
    A, B, na, nb, sig = _local_prep(A, B, na, nb, sigmas)
    offA = np.concatenate([[0], np.cumsum(na)]).astype(np.int64)
    offB = np.concatenate([[0], np.cumsum(nb)]).astype(np.int64)
    K = np.empty((sig.size, na.size, nb.size), order="F")
    lib().oracle_local_kernels(1 if kind == "gaussian" else 0, _ptr(A), _ptr(offA, _c_int64_p),
                               na.size, _ptr(B), _ptr(offB, _c_int64_p), nb.size, A.shape[1],
                               _ptr(sig), sig.size, _ptr(K))
    return K


def ref_local_kernels(kind, A, B, na, nb, sigmas):
    """The reference's own fget_local_kernels_{gaussian,laplacian} via oracle/_ref."""
    A, B, na, nb, sig = _local_prep(A, B, na, nb, sigmas)
    n1 = np.ascontiguousarray(na, dtype=np.int32)
    n2 = np.ascontiguousarray(nb, dtype=np.int32)
    K = np.empty((sig.size, na.size, nb.size), order="F")
    ip = ctypes.POINTER(ctypes.c_int)
    ref().ref_local_kernels(1 if kind == "gaussian" else 0, _ptr(A), A.shape[0], _ptr(B),
                            B.shape[0], A.shape[1], _ptr(n1, ip), na.size, _ptr(n2, ip), nb.size,
                            _ptr(sig), sig.size, _ptr(K))
    return K


# --------------------------------------------------------------------------------------------
# Kernels and distances (outputs Fortran-ordered (na, nb), as qml/kernels.py:58)
# --------------------------------------------------------------------------------------------

def _prep(A, B):
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    assert A.ndim == 2 and B.ndim == 2 and A.shape[1] == B.shape[1]
    return A, B, np.empty((A.shape[0], B.shape[0]), order="F")


def laplacian_kernel(A, B, sigma):
    A, B, K = _prep(A, B)
    lib().oracle_laplacian(_ptr(A), A.shape[0], _ptr(B), B.shape[0], A.shape[1], sigma, _ptr(K))
    return K


def gaussian_kernel(A, B, sigma):
    A, B, K = _prep(A, B)
    lib().oracle_gaussian(_ptr(A), A.shape[0], _ptr(B), B.shape[0], A.shape[1], sigma, _ptr(K))
    return K


def manhattan_distance(A, B):
    A, B, D = _prep(A, B)
    lib().oracle_manhattan(_ptr(A), A.shape[0], _ptr(B), B.shape[0], A.shape[1], _ptr(D))
    return D


def l2_distance(A, B):
    A, B, D = _prep(A, B)
    lib().oracle_l2(_ptr(A), A.shape[0], _ptr(B), B.shape[0], A.shape[1], _ptr(D))
    return D


REF_KIND = {"laplacian": 0, "gaussian": 1, "manhattan": 2, "l2": 3}


def ref_kernel(kind, A, B, sigma=1.0, build="_ref"):
    """The reference's own f*_kernel / f*_distance via oracle/_ref (or another build)."""
    A, B, K = _prep(A, B)
    ref(build).ref_kernel(REF_KIND[kind], _ptr(A), A.shape[0], _ptr(B), B.shape[0], A.shape[1],
                          _ptr(K), float(sigma))
    return K


def linear_kernel(A, B):
    A, B, K = _prep(A, B)
    lib().oracle_linear(_ptr(A), A.shape[0], _ptr(B), B.shape[0], A.shape[1], _ptr(K))
    return K


def sargan_kernel(A, B, sigma, gammas):
    """fsargan_kernel; qml/kernels.py:117-150 returns the Laplacian kernel for no gammas."""
    g = np.ascontiguousarray(np.asarray(gammas, dtype=np.float64).reshape(-1))
    if g.size == 0:
        return laplacian_kernel(A, B, sigma)
    A, B, K = _prep(A, B)
    lib().oracle_sargan(_ptr(A), A.shape[0], _ptr(B), B.shape[0], A.shape[1], float(sigma),
                        _ptr(g), g.size, _ptr(K))
    return K


def matern_kernel(A, B, sigma, order=0, metric="l1"):
    """qml/kernels.py:152-216: the L1 form through sargan_kernel, the L2 form fmatern_kernel_l2."""
    if metric == "l1":
        if order == 0:
            gammas = []
        elif order == 1:
            gammas = [1]
            sigma /= np.sqrt(3)
        elif order == 2:
            gammas = [1, 1 / 3.0]
            sigma /= np.sqrt(5)
        else:
            raise SystemExit
        return sargan_kernel(A, B, sigma, gammas)
    if metric != "l2":
        raise SystemExit
    A, B, K = _prep(A, B)
    lib().oracle_matern_l2(_ptr(A), A.shape[0], _ptr(B), B.shape[0], A.shape[1], float(sigma),
                           int(order), _ptr(K))
    return K


def p_distance(A, B, p=2):
    """qml/distance.py:94-147 dispatch (int / integer-valued float / float p)."""
    if type(p) == type(1.0) and p.is_integer():
        p = int(p)
    if type(p) == type(1):
        if p == 2:
            return l2_distance(A, B)
        A, B, D = _prep(A, B)
        lib().oracle_p_distance(_ptr(A), A.shape[0], _ptr(B), B.shape[0], A.shape[1], float(p),
                                1, _ptr(D))
        return D
    if type(p) == type(1.0):
        A, B, D = _prep(A, B)
        lib().oracle_p_distance(_ptr(A), A.shape[0], _ptr(B), B.shape[0], A.shape[1], p, 0,
                                _ptr(D))
        return D
    raise ValueError("expected exponent of integer or float type")


REF_EXT_KIND = {"linear": 0, "sargan": 1, "matern_l2": 2, "p_int": 3, "p_real": 4}


def ref_kernel_ext(kind, A, B, sigma=1.0, gammas=(), order=0, p=2.0):
    """The reference's flinear_kernel / fsargan_kernel / fmatern_kernel_l2 / fp_distance_*."""
    A, B, K = _prep(A, B)
    g = np.ascontiguousarray(np.asarray(gammas, dtype=np.float64).reshape(-1))
    if g.size == 0:
        g = np.zeros(1)
        ng = 0
    else:
        ng = g.size
    ref().ref_kernel_ext(REF_EXT_KIND[kind], _ptr(A), A.shape[0], _ptr(B), B.shape[0],
                         A.shape[1], _ptr(K), float(sigma), _ptr(g), ng, int(order), float(p),
                         int(p))
    return K


# --------------------------------------------------------------------------------------------
# cho_solve: qml/math.py:60-97 + qml/fcho_solve.f90:23-52
# --------------------------------------------------------------------------------------------

def cho_solve(A, y):
    """Solve A x = y with LAPACK dpotrf('U') + dpotrs('U'), restoring A as the reference does.

    qml/fcho_solve.f90:35 factors the upper triangle (numpy indices i <= j) of the array
    Fortran sees; qml/math.py:85-95 then restores the diagonal and copies lower -> upper.
    Warnings on info != 0 are printed with the reference's wording and x is still returned.
    """
    from scipy.linalg import lapack

    if len(A.shape) != 2 or A.shape[0] != A.shape[1]:
        raise ValueError("expected square matrix")
    if len(y.shape) != 1 or y.shape[0] != A.shape[1]:
        raise ValueError("expected matrix and vector of same stride size")
    n = A.shape[0]
    A_diag = A[np.diag_indices_from(A)]
    c, info = lapack.dpotrf(A, lower=0, clean=0, overwrite_a=0)
    if info > 0:
        print(" WARNING: Error in LAPACK Cholesky decomposition DPOTRF().")
        print(" WARNING: The", info, "-th leading order is not positive definite.")
    elif info < 0:
        print(" WARNING: Error in LAPACK Cholesky decomposition DPOTRF().")
        print(" WARNING: The", -info, "-th argument had an illegal value.")
    x, info2 = lapack.dpotrs(c, np.asarray(y, dtype=np.float64).reshape(n, 1), lower=0)
    if info2 < 0:
        print(" WARNING: Error in LAPACK Cholesky solver DPOTRS().")
        print(" WARNING: The", -info2, "-th argument had an illegal value.")
    A[np.diag_indices_from(A)] = A_diag
    i_lower = np.tril_indices_from(A)
    A.T[i_lower] = A[i_lower]
    return x.reshape(n)


def cho_invert(A):
    """qml/math.py:33-58 around qml/fcho_solve.f90:54-73: LAPACK dpotrf('L') + dpotri('L') on
    the Fortran-ordered array (A itself when it already is one), then lower -> upper copy."""
    from scipy.linalg import lapack

    if len(A.shape) != 2 or A.shape[0] != A.shape[1]:
        raise ValueError("expected square matrix")
    I = np.asfortranarray(A)
    c, info = lapack.dpotrf(I, lower=1, clean=0, overwrite_a=1)
    if info > 0:
        print(" WARNING: Cholesky decomposition DPOTRF() exited with error code:", info)
    inv, info2 = lapack.dpotri(c, lower=1, overwrite_c=1)
    I[...] = inv
    i_lower = np.tril_indices_from(I)
    I.T[i_lower] = I[i_lower]
    return I


def bkf_solve(A, y):
    """qml/math.py:131-164 around qml/fcho_solve.f90:115-150: LAPACK dsytrf('L') + dsytrs."""
    from scipy.linalg import lapack

    n = A.shape[0]
    A_diag = A[np.diag_indices_from(A)]
    lu, piv, info = lapack.dsytrf(np.asfortranarray(A, dtype=np.float64), lower=1)[:3]
    if info > 0:
        print(" WARNING: Bunch-Kaufman factorization DSYTRI() exited with error code:", info)
    x, _ = lapack.dsytrs(lu, piv, np.array(y, dtype=np.float64), lower=1)
    A[np.diag_indices_from(A)] = A_diag
    i_lower = np.tril_indices_from(A)
    A.T[i_lower] = A[i_lower]
    return x.reshape(n)


def bkf_invert(A):
    """qml/math.py:100-128 around qml/fcho_solve.f90:76-112: dsytrf('L') + dsytri('L') on the
    Fortran-ordered array, then lower -> upper copy. scipy does not expose dsytri; the inverse
    here is dsytrs with the identity on the same factors (equal to dsytri's to rounding), of
    which the lower triangle is kept, as the reference keeps dsytri's."""
    from scipy.linalg import lapack

    I = np.asfortranarray(A)
    n = I.shape[0]
    lu, piv, info = lapack.dsytrf(np.array(I, dtype=np.float64, order="F"), lower=1)[:3]
    if info > 0:
        print(" WARNING: Bunch-Kaufman factorization DSYTRI() exited with error code:", info)
    X, _ = lapack.dsytrs(lu, piv, np.eye(n), lower=1)
    I[...] = np.tril(X) + np.tril(X, -1).T
    return I


def bkf_factor(A):
    """LAPACK dsytrf('L') of A (a copy) with a workspace of n, which makes LAPACK run its
    unblocked Bunch-Kaufman code (dsytf2): (lower factors, 1-based pivots, info). The
    reference's blocked call (lwork = n nb) makes the same pivot choices in exact arithmetic."""
    from scipy.linalg import lapack

    n = A.shape[0]
    lu, piv, info = lapack.dsytrf(np.asfortranarray(A, dtype=np.float64), lower=1,
                                  lwork=max(1, n))[:3]
    return np.tril(lu), piv, info


def bkf_factor_blocked(A):
    """LAPACK dsytrf('L') of A (a copy) with its optimal workspace, as the reference calls it
    (qml/fcho_solve.f90:90-99: lwork = n nb with nb = ilaenv(1, 'DSYTRF', ...)): LAPACK's
    blocked Bunch-Kaufman (dlasyf panels of nb columns, unblocked dsytf2 for the last ones).
    Returns (lower factors, 1-based pivots, info)."""
    from scipy.linalg import lapack

    lu, piv, info = lapack.dsytrf(np.asfortranarray(A, dtype=np.float64), lower=1)[:3]
    return np.tril(lu), piv, info


def krr_fit_predict(X, y, Xs, kernel, sigma, lam):
    """KRR composition of tests/test_energy_krr_cmat.py:101-111: returns (alpha, y_pred)."""
    kern = laplacian_kernel if kernel == "laplacian" else gaussian_kernel
    K = kern(X, X, sigma)
    K[np.diag_indices_from(K)] += lam
    alpha = cho_solve(K, np.asarray(y, dtype=np.float64))
    Ks = kern(X, Xs, sigma)
    return alpha, Ks.T.dot(alpha)
