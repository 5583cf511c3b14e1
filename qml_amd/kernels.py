"""Laplacian and Gaussian kernel matrices on the GPU.

``laplacian_kernel`` and ``gaussian_kernel`` keep the signatures and results of
qml.kernels (qml/kernels.py:36-92): inputs (na, d) and (nb, d), output the Fortran-ordered
(na, nb) FP64 array. The ``*_device`` forms take and return device tensors.
"""

from __future__ import annotations

import numpy as np

from . import _lib
from .device import device, rows_to_device

# The expanded (MFMA) Gaussian is used when its worst-case exponent error, 2 * d * u *
# max|a| max|b| * 0.5 / sigma^2 (a dot product of length d in FP64, u = 2^-53), is below this.
EXPANSION_TOL = 1e-13
_U = 2.0 ** -53


def _check_pair(A, B):
    if len(A.shape) != 2 or len(B.shape) != 2:
        raise ValueError("expected matrices of dimension=2")
    if B.shape[1] != A.shape[1]:
        raise ValueError("expected matrices containing vectors of same size")


def _empty_kernel_device(na, nb):
    import torch

    # Column-major (na, nb) == row-major (nb, na).
    return torch.empty((nb, na), dtype=torch.float64, device=device())


def pairwise_device(kind: int, Ad, na: int, Bd, nb: int, d: int, sigma: float, symmetric=False,
                    out=None, stream=None):
    """Run a VALU pairwise kernel on device rows; returns the (nb, na) tensor = K col-major."""
    K = _empty_kernel_device(na, nb) if out is None else out
    L = _lib.lib()
    _lib.check(L.qmlh_pairwise(kind, _lib.ptr(Ad), na, Ad.shape[1], _lib.ptr(Bd), nb, Bd.shape[1],
                               d, float(sigma), _lib.ptr(K), na,
                               _lib.SYMMETRIC if symmetric else 0, _lib.stream_handle(stream)))
    return K


def row_sqnorms_device(Xd, n: int, d: int, stream=None):
    import torch

    out = torch.empty(n, dtype=torch.float64, device=Xd.device)
    _lib.check(_lib.lib().qmlh_row_sqnorms(_lib.ptr(Xd), n, Xd.shape[1], d, _lib.ptr(out),
                                           _lib.stream_handle(stream)))
    return out


def gaussian_mfma_device(Ad, nA, na: int, Bd, nB, nb: int, d: int, sigma: float, symmetric=False,
                         out=None, stream=None):
    K = _empty_kernel_device(na, nb) if out is None else out
    _lib.check(_lib.lib().qmlh_gaussian_mfma(
        _lib.ptr(Ad), na, Ad.shape[1], _lib.ptr(nA), _lib.ptr(Bd), nb, Bd.shape[1], _lib.ptr(nB),
        d, float(sigma), _lib.ptr(K), na, _lib.SYMMETRIC if symmetric else 0,
        _lib.stream_handle(stream)))
    return K


def expansion_ok(maxnorm_a: float, maxnorm_b: float, d: int, sigma: float) -> bool:
    """Whether the expanded Gaussian stays within EXPANSION_TOL of the direct form."""
    bound = 2.0 * d * _U * np.sqrt(maxnorm_a * maxnorm_b) * 0.5 / (sigma * sigma)
    bound += 2.0 * _U * (maxnorm_a + maxnorm_b) * 0.5 / (sigma * sigma)
    return bool(bound < EXPANSION_TOL)


def _to_host_fortran(Kd):
    # (nb, na) row-major on device == (na, nb) column-major: transpose view of the host copy.
    return Kd.cpu().numpy().T


def laplacian_kernel(A, B, sigma):
    """K_ij = exp(-||A_i - B_j||_1 / sigma); (na, nb), Fortran order (qml/kernels.py:36-63)."""
    _check_pair(A, B)
    symmetric = A is B
    A = np.asarray(A, dtype=np.float64)
    B = A if symmetric else np.asarray(B, dtype=np.float64)
    na, nb, d = A.shape[0], B.shape[0], A.shape[1]
    if na == 0 or nb == 0:
        return np.empty((na, nb), order="F")
    Ad = rows_to_device(A)
    Bd = Ad if symmetric else rows_to_device(B)
    Kd = pairwise_device(_lib.PAIR_LAPLACIAN, Ad, na, Bd, nb, d, sigma, symmetric)
    return _to_host_fortran(Kd)


def gaussian_kernel(A, B, sigma):
    """K_ij = exp(-||A_i - B_j||_2^2 / (2 sigma^2)); (na, nb), Fortran order
    (qml/kernels.py:65-92). FP64 MFMA expanded form where its error bound allows, else the
    direct-difference VALU kernel."""
    _check_pair(A, B)
    symmetric = A is B
    A = np.asarray(A, dtype=np.float64)
    B = A if symmetric else np.asarray(B, dtype=np.float64)
    na, nb, d = A.shape[0], B.shape[0], A.shape[1]
    if na == 0 or nb == 0:
        return np.empty((na, nb), order="F")
    Ad = rows_to_device(A)
    Bd = Ad if symmetric else rows_to_device(B)
    ma = float(np.max(np.einsum("ij,ij->i", A, A))) if d else 0.0
    mb = ma if symmetric else (float(np.max(np.einsum("ij,ij->i", B, B))) if d else 0.0)
    if expansion_ok(ma, mb, d, sigma):
        nA = row_sqnorms_device(Ad, na, d)
        nB = nA if symmetric else row_sqnorms_device(Bd, nb, d)
        Kd = gaussian_mfma_device(Ad, nA, na, Bd, nB, nb, d, sigma, symmetric)
    else:
        Kd = pairwise_device(_lib.PAIR_GAUSSIAN, Ad, na, Bd, nb, d, sigma, symmetric)
    return _to_host_fortran(Kd)
