"""Coulomb-matrix representations on the GPU.

``generate_coulomb_matrix`` keeps the signature and results of
qml.representations.generate_coulomb_matrix (qml/representations.py:69-119);
``generate_coulomb_matrices`` is the batched form that the KRR path and the server use:
one kernel launch for any number of molecules, output left in HBM.
``generate_atomic_coulomb_matrix`` keeps the signature and results of
qml.representations.generate_atomic_coulomb_matrix (qml/representations.py:121-233);
``generate_atomic_coulomb_matrices`` is its batched form (one row per central atom).
"""

from __future__ import annotations

import numpy as np

from . import _lib
from .device import device, padded_ld, ragged_to_device

_SORTINGS = {"row-norm": 1, "unsorted": 0}
_ACM_SORTINGS = {"row-norm": _lib.ACM_ROW_NORM, "distance": _lib.ACM_DISTANCE}

# Element symbol -> nuclear charge (the periodic table through Cn = 112, plus the provisional
# names Uuq = 114 and Uuh = 116), the keys qml/data.py:23-137 accepts for ``indices``.
_SYMBOLS = ("H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn "
            "Ga Ge As Se Br Kr Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce "
            "Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn "
            "Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf Es Fm Md No Lr Rf Db Sg Bh Hs Mt Ds Rg Cn").split()
NUCLEAR_CHARGE = {sym: z for z, sym in enumerate(_SYMBOLS, start=1)}
NUCLEAR_CHARGE.update({"Uuq": 114, "Uuh": 116})


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


def generate_atomic_coulomb_matrices_device(Zd, Rd, Od, central, cmol, size: int,
                                            sorting: str = "distance", central_cutoff=1e6,
                                            central_decay=-1, interaction_cutoff=1e6,
                                            interaction_decay=-1, ld: int | None = None,
                                            out=None, stream=None, max_atoms: int | None = None):
    """Batched atomic CM from device tensors: Z f64 (total,), R f64 (total, 3), offsets i64
    (nmol+1,), central i64 (rows,) central atom within its molecule, cmol i64 (rows,) its
    molecule. Returns ``(out, status)``: the (rows, ld) FP64 device tensor (pad columns zero)
    and the device status word (bit 1: a central atom needs more than ``size`` rows; bit 4: an
    index out of range). max_atoms: the largest molecule (read from ``offsets`` when None);
    molecules above 256 atoms go through the global-memory kernel (qmlh_atomic_cm_large)."""
    import torch

    if sorting not in _ACM_SORTINGS:
        print("ERROR: Unknown sorting scheme requested")
        raise SystemExit
    d = cm_length(size)
    ld = padded_ld(d) if ld is None else ld
    rows = central.shape[0]
    if out is None:
        out = torch.empty((rows, ld), dtype=torch.float64, device=Od.device)
    status = torch.zeros(1, dtype=torch.int32, device=Od.device)
    L = _lib.lib()
    args = (_lib.ptr(Zd), _lib.ptr(Rd), _lib.ptr(Od), Od.shape[0] - 1, _lib.ptr(central),
            _lib.ptr(cmol), rows, size, _ACM_SORTINGS[sorting], float(central_cutoff),
            float(central_decay), float(interaction_cutoff), float(interaction_decay),
            _lib.ptr(out), ld)
    if max_atoms is None:
        max_atoms = int((Od[1:] - Od[:-1]).max()) if Od.shape[0] > 1 else 0
    if max_atoms <= _lib.ACM_MAX_ATOMS:
        _lib.check(L.qmlh_atomic_cm_batch(*args, _lib.ptr(status), _lib.stream_handle(stream)))
        return out, status
    # Rows of molecules above the LDS kernel's capacity (reported there as status bit 2) are
    # then computed by the global-memory kernel, which reports its own status.
    first = torch.zeros(1, dtype=torch.int32, device=Od.device)
    _lib.check(L.qmlh_atomic_cm_batch(*args, _lib.ptr(first), _lib.stream_handle(stream)))
    ws = torch.empty(max(1, L.qmlh_atomic_cm_large_workspace(max_atoms, size) // 8),
                     dtype=torch.float64, device=Od.device)
    _lib.check(L.qmlh_atomic_cm_large(*args, _lib.ptr(status), max_atoms, _lib.ptr(ws),
                                      ws.numel() * 8, _lib.stream_handle(stream)))
    status |= first & ~2
    return out, status


def _raise_acm_status(code: int, size: int) -> None:
    if code & 1:
        raise ValueError(f"ERROR: Coulomb matrix generation: {size} size set, but more "
                         "atoms lie within the central cut-off")
    if code & 2:
        raise ValueError("molecule larger than the atomic Coulomb matrix workspace")
    if code & 4:
        raise ValueError("central atom index out of range")


def generate_atomic_coulomb_matrices(Z, R, offsets, size: int = 23, sorting: str = "distance",
                                     central_cutoff=1e6, central_decay=-1,
                                     interaction_cutoff=1e6, interaction_decay=-1,
                                     ld: int | None = None):
    """Atomic CMs of every atom of a ragged batch (host arrays in); returns the (total_atoms,
    ld) device tensor, rows in atom order (molecule by molecule), as
    np.concatenate of the per-molecule results."""
    import torch

    offsets = np.asarray(offsets, dtype=np.int64)
    nat = np.diff(offsets)
    total = int(offsets[-1]) if offsets.size else 0
    cmol = np.repeat(np.arange(nat.size, dtype=np.int64), nat)
    central = np.arange(total, dtype=np.int64) - np.repeat(offsets[:-1], nat)
    Zd, Rd, Od = ragged_to_device(Z, R, offsets)
    dev = Od.device
    out, status = generate_atomic_coulomb_matrices_device(
        Zd, Rd, Od, torch.from_numpy(central).to(dev), torch.from_numpy(cmol).to(dev), size,
        sorting, central_cutoff, central_decay, interaction_cutoff, interaction_decay, ld,
        max_atoms=int(nat.max()) if nat.size else 0)
    _raise_acm_status(int(status.item()), size)
    return out


def generate_atomic_coulomb_matrix(nuclear_charges, coordinates, size=23, sorting="distance",
                                   central_cutoff=1e6, central_decay=-1, interaction_cutoff=1e6,
                                   interaction_decay=-1, indices=None):
    """Atomic Coulomb matrices of one molecule, one row per central atom.

    Same signature, defaults and output as qml/representations.py:121-233: an
    (nindices, size(size+1)/2) FP64 array. ``indices``: None (every atom), an element symbol
    (its atoms; no match gives a (0, 0) array) or 0-based atom indices. Unknown ``sorting`` or
    element symbol print the reference's message and raise SystemExit. Divergence: more atoms
    within the central cut-off than ``size`` raises ValueError where the reference stops the
    process (qml/frepresentations.f90:297-304).
    """
    import torch

    Z = np.asarray(nuclear_charges, dtype=np.float64).reshape(-1)
    R = np.asarray(coordinates, dtype=np.float64).reshape(-1, 3)
    if R.shape[0] != Z.shape[0]:
        raise ValueError(f"{R.shape[0]} coordinates, but {Z.shape[0]} atom_types!")
    if indices is None:
        idx = np.arange(Z.shape[0], dtype=np.int64)
    elif isinstance(indices, str):
        if indices not in NUCLEAR_CHARGE:
            print("ERROR: Unknown value %s given for 'indices' variable" % indices)
            raise SystemExit
        idx = np.where(Z == NUCLEAR_CHARGE[indices])[0].astype(np.int64)
        if idx.size == 0:
            return np.zeros((0, 0))
    else:
        idx = np.asarray(indices, dtype=np.int64).reshape(-1)
    if sorting not in _ACM_SORTINGS:
        print("ERROR: Unknown sorting scheme requested")
        raise SystemExit
    if np.any(idx < 0) or np.any(idx >= Z.shape[0]):
        raise IndexError("central atom index out of range")
    d = cm_length(size)
    if idx.size == 0:
        return np.zeros((0, d))
    Zd, Rd, Od = ragged_to_device(Z, R, np.array([0, Z.shape[0]]))
    dev = Od.device
    central = torch.from_numpy(idx).to(dev)
    cmol = torch.zeros(idx.size, dtype=torch.int64, device=dev)
    out, status = generate_atomic_coulomb_matrices_device(
        Zd, Rd, Od, central, cmol, size, sorting, central_cutoff, central_decay,
        interaction_cutoff, interaction_decay, max_atoms=Z.shape[0])
    _raise_acm_status(int(status.item()), size)
    return out[:, :d].cpu().numpy()
