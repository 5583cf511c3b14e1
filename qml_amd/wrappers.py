"""Atomic-kernel wrappers over molecule objects: qml/wrappers.py:33-97.

``get_atomic_kernels_laplacian`` / ``get_atomic_kernels_gaussian`` take two lists of objects
with ``natoms`` and ``representation`` (one row per atom, e.g. from
``generate_atomic_coulomb_matrix``) and return the (nsigmas, len(mols1), len(mols2)) local
kernels. The reference pads every molecule to the largest one and calls
fget_vector_kernels_* (qml/fkernels.f90:198-337); the values are those of the local kernels
on the concatenated rows, which is how they are computed here.
"""

from __future__ import annotations

import numpy as np

from .kernels import get_local_kernels_gaussian, get_local_kernels_laplacian


def _stack(mols):
    n = np.array([mol.natoms for mol in mols], dtype=np.int64)
    X = np.concatenate([np.asarray(mol.representation, dtype=np.float64)[:k]
                        for mol, k in zip(mols, n)])
    return X, n


def _atomic(fn, mols1, mols2, sigmas):
    X1, n1 = _stack(mols1)
    if mols2 is mols1:
        return fn(X1, X1, n1, n1, sigmas)
    X2, n2 = _stack(mols2)
    return fn(X1, X2, n1, n2, sigmas)


def get_atomic_kernels_laplacian(mols1, mols2, sigmas):
    return _atomic(get_local_kernels_laplacian, mols1, mols2, sigmas)


def get_atomic_kernels_gaussian(mols1, mols2, sigmas):
    return _atomic(get_local_kernels_gaussian, mols1, mols2, sigmas)
