"""Coulomb-matrix representations on the GPU.

``generate_coulomb_matrix`` keeps the signature and results of
qml.representations.generate_coulomb_matrix (qml/representations.py:69-119);
``generate_coulomb_matrices`` is the batched form that the KRR path and the server use:
one kernel launch for any number of molecules, output left in HBM.
"""

from __future__ import annotations

import numpy as np

from . import _lib
from .device import device, padded_ld, ragged_to_device

_SORTINGS = {"row-norm": 1, "unsorted": 0}


def cm_length(size: int) -> int:
    return size * (size + 1) // 2


def generate_coulomb_matrices_device(Zd, Rd, Od, size: int, sorting: str = "row-norm",
                                     ld: int | None = None, out=None, stream=None):
    """Batched CM from device tensors (Z f64 (total,), R f64 (total,3), offsets i64 (nmol+1,)).

    Returns an (nmol, ld) FP64 device tensor, ld >= size(size+1)/2, pad columns zero.
    Raises ValueError if any molecule has more atoms than ``size`` (checked on the device).
    """
    import torch

    if sorting not in _SORTINGS:
        print("ERROR: Unknown sorting scheme requested")
        raise SystemExit
    d = cm_length(size)
    ld = padded_ld(d) if ld is None else ld
    nmol = Od.shape[0] - 1
    if out is None:
        out = torch.empty((nmol, ld), dtype=torch.float64, device=Od.device)
    status = torch.zeros(1, dtype=torch.int32, device=Od.device)
    L = _lib.lib()
    _lib.check(L.qmlh_cm_batch(_lib.ptr(Zd), _lib.ptr(Rd), _lib.ptr(Od), nmol, size,
                               _SORTINGS[sorting], _lib.ptr(out), ld, _lib.ptr(status),
                               _lib.stream_handle(stream)))
    return out, status


def generate_coulomb_matrices(Z, R, offsets, size: int = 23, sorting: str = "row-norm",
                              ld: int | None = None):
    """Batched CM from host arrays; returns the (nmol, ld) device tensor."""
    offsets = np.asarray(offsets, dtype=np.int64)
    nat = np.diff(offsets)
    if nat.size and int(nat.max()) > size:
        raise ValueError(f"molecule with {int(nat.max())} atoms exceeds size={size}")
    Zd, Rd, Od = ragged_to_device(Z, R, offsets)
    out, _ = generate_coulomb_matrices_device(Zd, Rd, Od, size, sorting, ld)
    return out


def generate_coulomb_matrix(nuclear_charges, coordinates, size=23, sorting="row-norm"):
    """Coulomb matrix of one molecule, flattened lower triangle of the sorted matrix.

    Same signature, dtype coercion and output as qml/representations.py:69-119: a 1-D FP64
    array of length size(size+1)/2. Unknown ``sorting`` prints the reference's message and
    raises SystemExit. Divergence: natoms > size raises ValueError (the reference writes out
    of bounds, qml/frepresentations.f90:131).
    """
    if sorting not in _SORTINGS:
        print("ERROR: Unknown sorting scheme requested")
        raise SystemExit
    Z = np.asarray(nuclear_charges, dtype=np.float64).reshape(-1)
    R = np.asarray(coordinates, dtype=np.float64).reshape(-1, 3)
    if R.shape[0] != Z.shape[0]:
        raise ValueError(f"{R.shape[0]} coordinates, but {Z.shape[0]} atom_types!")
    if Z.shape[0] > size:
        raise ValueError(f"molecule with {Z.shape[0]} atoms exceeds size={size}")
    out = generate_coulomb_matrices(Z, R, np.array([0, Z.shape[0]]), size, sorting)
    return out[0, :cm_length(size)].cpu().numpy()
