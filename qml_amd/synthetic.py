"""Seeded synthetic molecules for the KRR workloads.

Molecules follow a fixed recipe so every run, on any host or GPU, sees identical inputs:

* atom count uniform in ``[min_atoms, max_atoms]``;
* elements H, C, N, O, S (Z = 1, 6, 7, 8, 16) with probabilities 0.55, 0.30, 0.07, 0.07, 0.01;
* the first atom sits at the origin; every later atom is placed 1.0-1.6 A (uniform) from a
  uniformly chosen earlier atom along a uniformly random direction, and the placement is
  redrawn when it lands closer than 0.9 A to any atom already placed;
* the FP64 label is ``sum_{i<j} Z_i Z_j (exp(-2 (r-1.2)) - 2 exp(-(r-1.2)))`` plus per-element
  offsets H -0.5, C -38, N -54, O -75, S -398.

numpy's PCG64 ``default_rng(seed)`` drives everything, so the stream is platform independent.
Atoms are placed for all molecules at once (atom slot by atom slot), which keeps generation of
100k molecules to a couple of seconds.

The ragged layout used by the device Coulomb-matrix kernel is ``Z`` (total_atoms,),
``R`` (total_atoms, 3) and ``offsets`` (nmol + 1,): molecule ``m`` owns atoms
``offsets[m]:offsets[m+1]``.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

ELEMENTS = np.array([1, 6, 7, 8, 16], dtype=np.int32)
ELEMENT_P = np.array([0.55, 0.30, 0.07, 0.07, 0.01])
LABEL_OFFSET = {1: -0.5, 6: -38.0, 7: -54.0, 8: -75.0, 16: -398.0}


@dataclass
class MoleculeSet:
    """A batch of molecules in padded and ragged form plus FP64 labels."""

    natoms: np.ndarray  # (nmol,) int64
    Z_pad: np.ndarray  # (nmol, max_atoms) int32, 0 beyond natoms
    R_pad: np.ndarray  # (nmol, max_atoms, 3) float64, 0 beyond natoms
    labels: np.ndarray  # (nmol,) float64

    @property
    def nmol(self) -> int:
        return int(self.natoms.shape[0])

    def ragged(self):
        """Return ``(Z, R, offsets)`` with Z float64 (total,), R (total, 3), offsets int64."""
        mask = np.arange(self.Z_pad.shape[1])[None, :] < self.natoms[:, None]
        Z = self.Z_pad[mask].astype(np.float64)
        R = np.ascontiguousarray(self.R_pad[mask])
        offsets = np.zeros(self.nmol + 1, dtype=np.int64)
        np.cumsum(self.natoms, out=offsets[1:])
        return Z, R, offsets

    def molecule(self, m: int):
        n = int(self.natoms[m])
        return self.Z_pad[m, :n].copy(), self.R_pad[m, :n].copy()

    def subset(self, start: int, stop: int) -> "MoleculeSet":
        return MoleculeSet(self.natoms[start:stop], self.Z_pad[start:stop],
                           self.R_pad[start:stop], self.labels[start:stop])


def _place_atoms(rng, natoms, max_atoms):
    nmol = natoms.shape[0]
    R = np.zeros((nmol, max_atoms, 3), dtype=np.float64)
    for a in range(1, max_atoms):
        pending = np.nonzero(natoms > a)[0]
        tries = 0
        while pending.size:
            tries += 1
            if tries > 10000:
                raise RuntimeError("synthetic geometry placement did not converge")
            k = pending.size
            parent = rng.integers(0, a, size=k)
            dist = rng.uniform(1.0, 1.6, size=k)
            dirn = rng.normal(size=(k, 3))
            dirn /= np.linalg.norm(dirn, axis=1, keepdims=True)
            pos = R[pending, parent] + dist[:, None] * dirn
            diff = R[pending, :a] - pos[:, None, :]
            dmin = np.sqrt((diff * diff).sum(axis=2)).min(axis=1)
            ok = dmin >= 0.9
            R[pending[ok], a] = pos[ok]
            pending = pending[~ok]
    return R


def _labels(Z_pad, R_pad, natoms):
    nmol, max_atoms = Z_pad.shape
    mask = np.arange(max_atoms)[None, :] < natoms[:, None]
    Zf = np.where(mask, Z_pad, 0).astype(np.float64)
    y = np.zeros(nmol)
    for i in range(max_atoms - 1):
        d = R_pad[:, i + 1:, :] - R_pad[:, i:i + 1, :]
        r = np.sqrt((d * d).sum(axis=2))
        zz = Zf[:, i:i + 1] * Zf[:, i + 1:]
        e = np.exp(-2.0 * (r - 1.2)) - 2.0 * np.exp(-(r - 1.2))
        y += np.where(zz != 0.0, zz * e, 0.0).sum(axis=1)
    off = np.zeros(17)
    for z, o in LABEL_OFFSET.items():
        off[z] = o
    y += np.where(mask, off[Z_pad], 0.0).sum(axis=1)
    return y


CHUNK = 1024


def _generate_chunk(nmol, rng, min_atoms, max_atoms):
    natoms = rng.integers(min_atoms, max_atoms + 1, size=nmol).astype(np.int64)
    Z_pad = ELEMENTS[rng.choice(len(ELEMENTS), p=ELEMENT_P, size=(nmol, max_atoms))]
    slot = np.arange(max_atoms)[None, :]
    Z_pad = np.where(slot < natoms[:, None], Z_pad, 0).astype(np.int32)
    R_pad = _place_atoms(rng, natoms, max_atoms)
    return natoms, Z_pad, R_pad


def generate(nmol: int, seed: int, min_atoms: int = 5, max_atoms: int = 23) -> MoleculeSet:
    """Generate ``nmol`` molecules with atom counts uniform in ``[min_atoms, max_atoms]``.

    Molecules come in chunks of ``CHUNK``, chunk ``c`` drawn from ``default_rng([seed, c])``,
    so the first ``n`` molecules are the same whatever ``nmol >= n`` is asked for.
    """
    parts = []
    for c in range((nmol + CHUNK - 1) // CHUNK):
        rng = np.random.default_rng([seed, c])
        parts.append(_generate_chunk(CHUNK, rng, min_atoms, max_atoms))
    if parts:
        natoms = np.concatenate([p[0] for p in parts])[:nmol]
        Z_pad = np.concatenate([p[1] for p in parts])[:nmol]
        R_pad = np.concatenate([p[2] for p in parts])[:nmol]
    else:
        natoms = np.zeros(0, np.int64)
        Z_pad = np.zeros((0, max_atoms), np.int32)
        R_pad = np.zeros((0, max_atoms, 3))
    labels = _labels(Z_pad, R_pad, natoms)
    return MoleculeSet(natoms, Z_pad, R_pad, labels)


# Workload presets: name -> (seed, n_train, n_test, max_atoms, cm_size)
WORKLOADS = {
    "krr_laplacian_1k": (1, 1000, 100, 23, 23),
    "krr_gaussian_20k": (2, 20000, 2000, 29, 29),
    "krr_large_100k": (3, 100000, 0, 29, 29),
    "krr_sweep_64": (4, 8000, 0, 23, 23),
    "krr_serve_50k": (5, 50000, 0, 29, 29),
}


def workload(name: str, n_train: int | None = None, n_test: int | None = None):
    """Return ``(train, test, cm_size)`` for a named workload; sizes may be overridden."""
    seed, ntr, nte, max_atoms, size = WORKLOADS[name]
    ntr = ntr if n_train is None else n_train
    nte = nte if n_test is None else n_test
    mols = generate(ntr + nte, seed, 5, max_atoms)
    return mols.subset(0, ntr), mols.subset(ntr, ntr + nte), size
