"""``Compound``: a molecule read from an XYZ file, with the representation generators of the path.

Keeps the attributes and behaviour of qml.Compound (qml/compound.py:37-326) for what this build
covers: ``read_xyz`` (qml/compound.py:296-326: the atom count on the first line, a comment line,
then one ``symbol x y z`` line per atom; parsing stops early at a line with fewer than four
fields, and extra fields or trailing lines, as in extended XYZ, are ignored),
``generate_coulomb_matrix`` and ``generate_atomic_coulomb_matrix`` (computed on the GPU).
The other representations (bag of bonds, eigenvalue Coulomb matrix, SLATM, FCHL) are not part of
this build and raise NotImplementedError.

``read_xyz_batch`` reads many files into the ragged layout the batched GPU entry points take.
"""

from __future__ import annotations

import collections

import numpy as np

from .representations import (NUCLEAR_CHARGE, generate_atomic_coulomb_matrix,
                              generate_coulomb_matrix)


def _parse_xyz(text: str):
    lines = text.splitlines()
    natoms = int(lines[0])
    types, Z, R = [], np.empty(natoms, dtype=int), np.empty((natoms, 3), dtype=float)
    for i, line in enumerate(lines[2:natoms + 2]):
        tokens = line.split()
        if len(tokens) < 4:
            break
        types.append(tokens[0])
        Z[i] = NUCLEAR_CHARGE[tokens[0]]
        R[i] = np.asarray(tokens[1:4], dtype=float)
    return natoms, types, Z, R


class Compound(object):
    """A molecule: ``Compound(xyz=path)`` reads an XYZ file (qml/compound.py:45-71)."""

    def __init__(self, xyz=None):
        empty_array = np.asarray([], dtype=float)
        self.molid = float("nan")
        self.name = None
        self.natoms = float("nan")
        self.natypes = {}
        self.atomtypes = []
        self.atomtype_indices = collections.defaultdict(list)
        self.nuclear_charges = empty_array
        self.coordinates = empty_array
        self.active_atoms = empty_array
        self.unit_cell = None
        self.energy = float("nan")
        self.properties = empty_array
        self.properties2 = empty_array
        self.representation = empty_array
        if xyz is not None:
            self.read_xyz(xyz)

    def read_xyz(self, filename):
        """(Re-)initialises the compound from an XYZ file (qml/compound.py:296-326)."""
        with open(filename, "r") as f:
            text = f.read()
        self.natoms, self.atomtypes, self.nuclear_charges, self.coordinates = _parse_xyz(text)
        self.name = filename
        self.atomtype_indices = collections.defaultdict(list)
        for i, t in enumerate(self.atomtypes):
            self.atomtype_indices[t].append(i)
        self.natypes = dict([(key, len(value)) for key, value in self.atomtype_indices.items()])

    def generate_coulomb_matrix(self, size=23, sorting="row-norm", indices=None):
        """Sorted Coulomb matrix (qml/compound.py:73-105) on the GPU."""
        self.representation = generate_coulomb_matrix(self.nuclear_charges, self.coordinates,
                                                      size=size, sorting=sorting)

    def generate_atomic_coulomb_matrix(self, size=23, sorting="row-norm", central_cutoff=1e6,
                                       central_decay=-1, interaction_cutoff=1e6,
                                       interaction_decay=-1, indices=None):
        """Atomic Coulomb matrices (qml/compound.py:134-215; default sorting "row-norm") on the
        GPU."""
        self.representation = generate_atomic_coulomb_matrix(
            self.nuclear_charges, self.coordinates, size=size, indices=indices, sorting=sorting,
            central_cutoff=central_cutoff, central_decay=central_decay,
            interaction_cutoff=interaction_cutoff, interaction_decay=interaction_decay)

    def _not_built(self, what):
        raise NotImplementedError(f"{what} is not part of this build (Coulomb matrices only)")

    def generate_eigenvalue_coulomb_matrix(self, size=23):
        self._not_built("the eigenvalue Coulomb matrix")

    def generate_bob(self, size=23, asize=None):
        self._not_built("bag of bonds")

    def generate_fchl_representation(self, *args, **kwargs):
        self._not_built("the FCHL representation")

    def generate_slatm(self, *args, **kwargs):
        self._not_built("SLATM")


def read_xyz_batch(paths):
    """Read XYZ files into ``(Z f64 (total,), R (total, 3), offsets i64 (nmol + 1,))``, the
    ragged layout of ``generate_coulomb_matrices`` / ``generate_atomic_coulomb_matrices``; each
    file parses as ``Compound.read_xyz`` does."""
    Zs, Rs, counts = [], [], []
    for path in paths:
        with open(path, "r") as f:
            natoms, types, Z, R = _parse_xyz(f.read())
        n = len(types)  # atoms actually read (an early short line stops the parse)
        Zs.append(Z[:n].astype(np.float64))
        Rs.append(R[:n])
        counts.append(n)
    offsets = np.zeros(len(counts) + 1, dtype=np.int64)
    np.cumsum(counts, out=offsets[1:])
    Z = np.concatenate(Zs) if Zs else np.zeros(0)
    R = np.concatenate(Rs) if Rs else np.zeros((0, 3))
    return Z, R, offsets
