"""Cholesky solve and inverse on the GPU: the drop-ins for qml.math.cho_solve
(qml/math.py:60-97) and qml.math.cho_invert (qml/math.py:33-58).

cho_solve: the reference calls LAPACK dpotrf("U") and dpotrs("U", n, 1)
(qml/fcho_solve.f90:35,46) on the upper triangle (numpy indices i <= j), prints a WARNING on a
non-zero info and still returns x, then restores the diagonal of A and copies its lower triangle
onto the upper one.
cho_invert: the reference makes a Fortran-ordered array I of A (a copy unless A already is one,
in which case A itself is overwritten), runs dpotrf("L") + dpotri("L") on it
(qml/fcho_solve.f90:54-73) and copies I's lower triangle onto the upper one.
bkf_solve / bkf_invert: the same with LAPACK's Bunch-Kaufman dsytrf("L") + dsytrs / dsytri
(qml/fcho_solve.f90:76-150), for symmetric indefinite matrices.
"""

from __future__ import annotations

import numpy as np

from . import _lib
from .device import device


def potrf_device(Ad, n: int, stream=None):
    """In-place lower Cholesky of a column-major (n, n) device matrix (an (n, n) tensor whose
    row-major bytes are the column-major matrix). Returns the device info tensor."""
    import torch

    info = torch.zeros(1, dtype=torch.int32, device=Ad.device)
    _lib.check(_lib.lib().qmlh_potrf(_lib.ptr(Ad), n, n, _lib.ptr(info),
                                     _lib.stream_handle(stream)))
    return info


def potrs_device(Ld, n: int, xd, stream=None, workspace=None):
    """x := (L L^T)^-1 x with L from potrf_device; x is an (n,) FP64 device tensor."""
    import torch

    L = _lib.lib()
    if workspace is None:
        nbytes = L.qmlh_potrs_workspace(n)
        workspace = torch.empty(max(1, (nbytes + 7) // 8), dtype=torch.float64, device=xd.device)
    _lib.check(L.qmlh_potrs(_lib.ptr(Ld), n, n, _lib.ptr(xd), _lib.ptr(workspace),
                            workspace.numel() * 8, _lib.stream_handle(stream)))
    return xd


def add_diagonal_device(Ad, n: int, value: float, stream=None):
    _lib.check(_lib.lib().qmlh_add_diagonal(_lib.ptr(Ad), n, n, float(value),
                                            _lib.stream_handle(stream)))


def print_potrf_warning(info: int) -> None:
    """The messages of qml/fcho_solve.f90:36-42 (Fortran list-directed output)."""
    if info > 0:
        print(" WARNING: Error in LAPACK Cholesky decomposition DPOTRF().")
        print(f" WARNING: The {info:11d} -th leading order is not positive definite.")
    elif info < 0:
        print(" WARNING: Error in LAPACK Cholesky decomposition DPOTRF().")
        print(f" WARNING: The {-info:11d} -th argument had an illegal value.")


def _copy_lower_to_upper(A, block: int = 2048) -> None:
    """A.T[tril] = A[tril] (qml/math.py:94-95) without the N(N+1)/2 index arrays."""
    n = A.shape[0]
    for r0 in range(0, n, block):
        r1 = min(n, r0 + block)
        D = A[r0:r1, r0:r1]
        iu = np.triu_indices(r1 - r0, 1)
        D[iu] = D.T[iu]
        if r1 < n:
            A[r0:r1, r1:] = A[r1:, r0:r1].T


def cho_solve(A, y):
    """Solve A x = y by Cholesky on the GPU; same signature, errors and side effects as
    qml.math.cho_solve (qml/math.py:60-97)."""
    import torch

    if len(A.shape) != 2 or A.shape[0] != A.shape[1]:
        raise ValueError('expected square matrix')
    if len(y.shape) != 1 or y.shape[0] != A.shape[1]:
        raise ValueError('expected matrix and vector of same stride size')
    n = A.shape[0]
    x = np.zeros(n)
    if n == 0:
        return x
    # C-ordered bytes of A read column-major are A^T, whose lower triangle is A's upper one.
    Ad = torch.from_numpy(np.ascontiguousarray(A, dtype=np.float64)).to(device())
    info = potrf_device(Ad, n)
    xd = torch.from_numpy(np.array(y, dtype=np.float64)).to(device())
    potrs_device(Ad, n, xd)
    x[:] = xd.cpu().numpy()
    print_potrf_warning(int(info.item()))
    _copy_lower_to_upper(A)
    return x


def cho_invert_device(Ad, n: int, stream=None):
    """In place: the column-major (n, n) device matrix (lower triangle read) becomes its whole
    symmetric inverse. Returns the device info tensor (dpotrf's)."""
    import torch

    L = _lib.lib()
    info = torch.zeros(1, dtype=torch.int32, device=Ad.device)
    nbytes = L.qmlh_cho_invert_workspace(n)
    ws = torch.empty(max(1, (nbytes + 7) // 8), dtype=torch.float64, device=Ad.device)
    _lib.check(L.qmlh_cho_invert(_lib.ptr(Ad), n, n, _lib.ptr(info), _lib.ptr(ws),
                                 ws.numel() * 8, _lib.stream_handle(stream)))
    return info


def cho_invert(A):
    """Inverse of a symmetric positive definite matrix on the GPU; same signature, result
    layout (Fortran order) and aliasing as qml.math.cho_invert (qml/math.py:33-58)."""
    import torch

    if len(A.shape) != 2 or A.shape[0] != A.shape[1]:
        raise ValueError('expected square matrix')
    I = np.asfortranarray(A)
    if I.dtype != np.float64:
        raise TypeError(f"cho_invert: expected a float64 array (got {I.dtype}), as the "
                        "reference's in-place Fortran call requires")
    n = I.shape[0]
    if n == 0:
        return I
    # I.T is a C-ordered view of I's bytes: its rows are I's columns (column-major I).
    Id = torch.from_numpy(np.ascontiguousarray(I.T)).to(device())
    info = cho_invert_device(Id, n)
    I.T[...] = Id.cpu().numpy()
    code = int(info.item())
    if code > 0:
        print(f" WARNING: Cholesky decomposition DPOTRF() exited with error code:{code:12d}")
    return I


def bkf_factor_device(Ad, n: int, stream=None):
    """In place: Bunch-Kaufman L D L^T of the column-major (n, n) device matrix (lower triangle
    read). Returns (ipiv, info) device tensors (1-based pivots, as LAPACK)."""
    import torch

    L = _lib.lib()
    ipiv = torch.zeros(max(1, n), dtype=torch.int32, device=Ad.device)
    info = torch.zeros(1, dtype=torch.int32, device=Ad.device)
    nbytes = L.qmlh_bkf_workspace(n)
    ws = torch.empty(max(1, (nbytes + 7) // 8), dtype=torch.float64, device=Ad.device)
    _lib.check(L.qmlh_bkf_factor(_lib.ptr(Ad), n, n, _lib.ptr(ipiv), _lib.ptr(info),
                                 _lib.ptr(ws), ws.numel() * 8, _lib.stream_handle(stream)))
    return ipiv, info


def bkf_solve_device(Ad, n: int, ipiv, Bd, nrhs: int, stream=None):
    """Bd (nrhs rows of n: column-major n x nrhs) := A^-1 Bd with the factors of
    bkf_factor_device."""
    _lib.check(_lib.lib().qmlh_bkf_solve(_lib.ptr(Ad), n, n, _lib.ptr(ipiv), _lib.ptr(Bd), n,
                                         nrhs, _lib.stream_handle(stream)))
    return Bd


def _print_bkf_warning(info: int, inverse: bool) -> None:
    # qml/fcho_solve.f90:100-108 and 137-147 (Fortran list-directed output).
    if info > 0:
        print(f" WARNING: Bunch-Kaufman factorization DSYTRI() exited with error code:{info:12d}")
        if inverse:
            print(f" WARNING: BKF inversion DPOTRI() exited with error code:{info:12d}")


def bkf_solve(A, y):
    """Solve A x = y for a symmetric (possibly indefinite) A by Bunch-Kaufman on the GPU; same
    signature, errors and side effects as qml.math.bkf_solve (qml/math.py:131-164): A's lower
    triangle is read, its diagonal restored and its lower triangle copied onto the upper one."""
    import torch

    if len(A.shape) != 2 or A.shape[0] != A.shape[1]:
        raise ValueError('expected square matrix')
    if len(y.shape) != 1 or y.shape[0] != A.shape[1]:
        raise ValueError('expected matrix and vector of same stride size')
    n = A.shape[0]
    A_diag = A[np.diag_indices_from(A)]
    x = np.zeros(n)
    if n == 0:
        return x
    # Column-major bytes of A's Fortran view: A.T in C order.
    Ad = torch.from_numpy(np.ascontiguousarray(np.asarray(A, dtype=np.float64).T)).to(device())
    ipiv, info = bkf_factor_device(Ad, n)
    xd = torch.from_numpy(np.array(y, dtype=np.float64)).to(device())
    bkf_solve_device(Ad, n, ipiv, xd, 1)
    x[:n] = xd.cpu().numpy()
    _print_bkf_warning(int(info.item()), False)
    A[np.diag_indices_from(A)] = A_diag
    _copy_lower_to_upper(A)
    return x


def bkf_invert(A):
    """Inverse of a symmetric (possibly indefinite) matrix by Bunch-Kaufman on the GPU; same
    signature, result layout (Fortran order) and aliasing as qml.math.bkf_invert
    (qml/math.py:100-128). The inverse is the factorisation's solve with the identity; its
    lower triangle is mirrored, so the result is exactly symmetric as the reference's is."""
    import torch

    if len(A.shape) != 2 or A.shape[0] != A.shape[1]:
        raise ValueError('expected square matrix')
    I = np.asfortranarray(A)
    if I.dtype != np.float64:
        raise TypeError(f"bkf_invert: expected a float64 array (got {I.dtype}), as the "
                        "reference's in-place Fortran call requires")
    n = I.shape[0]
    if n == 0:
        return I
    Ad = torch.from_numpy(np.ascontiguousarray(I.T)).to(device())
    ipiv, info = bkf_factor_device(Ad, n)
    X = torch.eye(n, dtype=torch.float64, device=Ad.device)
    bkf_solve_device(Ad, n, ipiv, X, n)
    Xh = X.cpu().numpy().T  # column-major solve result -> (n, n) with X[i, j]
    low = np.tril(Xh)
    I[...] = low + np.tril(Xh, -1).T
    _print_bkf_warning(int(info.item()), True)
    return I
