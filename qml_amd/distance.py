"""Manhattan, L2 and p-norm distance matrices on the GPU.

Same signatures, validation and outputs as qml.distance.manhattan_distance / l2_distance /
p_distance (qml/distance.py:28-147): (na, nb) FP64, Fortran order.
"""

from __future__ import annotations

import numpy as np

from . import _lib
from .device import rows_to_device
from .kernels import _to_host_fortran, pairwise_device, pairwise_ext_device


def _distance(kind, A, B):
    if len(A.shape) != 2 or len(B.shape) != 2:
        raise ValueError("expected matrices of dimension=2")
    if B.shape[1] != A.shape[1]:
        raise ValueError("expected matrices containing vectors of same size")
    symmetric = A is B
    A = np.asarray(A, dtype=np.float64)
    B = A if symmetric else np.asarray(B, dtype=np.float64)
    na, nb, d = A.shape[0], B.shape[0], A.shape[1]
    if na == 0 or nb == 0:
        return np.empty((na, nb), order="F")
    Ad = rows_to_device(A)
    Bd = Ad if symmetric else rows_to_device(B)
    return _to_host_fortran(pairwise_device(kind, Ad, na, Bd, nb, d, 0.0, symmetric))


def manhattan_distance(A, B):
    """D_ij = ||A_i - B_j||_1 (qml/distance.py:28-59, qml/fdistance.f90:23-46)."""
    return _distance(_lib.PAIR_MANHATTAN, A, B)


def l2_distance(A, B):
    """D_ij = ||A_i - B_j||_2 (qml/distance.py:61-92, qml/fdistance.f90:49-80)."""
    return _distance(_lib.PAIR_L2, A, B)


def p_distance(A, B, p=2):
    """D_ij = ||A_i - B_j||_p (qml/distance.py:94-147): an int p (or an integer-valued float)
    uses integer powers, p = 2 the L2 distance, any other float real powers; another type of p
    raises ValueError, as in the reference."""
    if len(A.shape) != 2 or len(B.shape) != 2:
        raise ValueError("expected matrices of dimension=2")
    if B.shape[1] != A.shape[1]:
        raise ValueError("expected matrices containing vectors of same size")
    if type(p) == type(1.0) and p.is_integer():
        p = int(p)
    if type(p) == type(1):
        if p == 2:
            return l2_distance(A, B)
        kind = _lib.PAIR_P_INT
    elif type(p) == type(1.0):
        kind = _lib.PAIR_P_REAL
    else:
        raise ValueError("expected exponent of integer or float type")
    symmetric = A is B
    A = np.asarray(A, dtype=np.float64)
    B = A if symmetric else np.asarray(B, dtype=np.float64)
    na, nb, d = A.shape[0], B.shape[0], A.shape[1]
    if na == 0 or nb == 0:
        return np.empty((na, nb), order="F")
    Ad = rows_to_device(A)
    Bd = Ad if symmetric else rows_to_device(B)
    return _to_host_fortran(pairwise_ext_device(kind, Ad, na, Bd, nb, d, [float(p)], symmetric))
