"""Kernel matrices on the GPU.

``laplacian_kernel``, ``gaussian_kernel``, ``linear_kernel``, ``sargan_kernel`` and
``matern_kernel`` keep the signatures and results of qml.kernels (qml/kernels.py:36-216):
inputs (na, d) and (nb, d), output the Fortran-ordered (na, nb) FP64 array. The local
(atom-summed) kernels are below. The ``*_device`` forms take and return device tensors.
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
                    out=None, stream=None, lower=False):
    """Run a VALU pairwise kernel on device rows; returns the (nb, na) tensor = K col-major
    (lower, with symmetric: only the lower triangle i >= j is written)."""
    K = _empty_kernel_device(na, nb) if out is None else out
    L = _lib.lib()
    flags = (_lib.SYMMETRIC if symmetric else 0) | (_lib.LOWER if lower else 0)
    _lib.check(L.qmlh_pairwise(kind, _lib.ptr(Ad), na, Ad.shape[1], _lib.ptr(Bd), nb, Bd.shape[1],
                               d, float(sigma), _lib.ptr(K), na, flags,
                               _lib.stream_handle(stream)))
    return K


def row_sqnorms_device(Xd, n: int, d: int, stream=None):
    import torch

    out = torch.empty(n, dtype=torch.float64, device=Xd.device)
    _lib.check(_lib.lib().qmlh_row_sqnorms(_lib.ptr(Xd), n, Xd.shape[1], d, _lib.ptr(out),
                                           _lib.stream_handle(stream)))
    return out


def gaussian_mfma_device(Ad, nA, na: int, Bd, nB, nb: int, d: int, sigma: float, symmetric=False,
                         out=None, stream=None, lower=False):
    """lower (with symmetric): write only the lower triangle of the column-major K (the upper one
    is left as it is), for callers that only read it."""
    K = _empty_kernel_device(na, nb) if out is None else out
    flags = (_lib.SYMMETRIC if symmetric else 0) | (_lib.LOWER if symmetric and lower else 0)
    _lib.check(_lib.lib().qmlh_gaussian_mfma(
        _lib.ptr(Ad), na, Ad.shape[1], _lib.ptr(nA), _lib.ptr(Bd), nb, Bd.shape[1], _lib.ptr(nB),
        d, float(sigma), _lib.ptr(K), na, flags, _lib.stream_handle(stream)))
    return K


def expansion_ok(maxnorm_a: float, maxnorm_b: float, d: int, sigma: float) -> bool:
    """Whether the expanded Gaussian stays within EXPANSION_TOL of the direct form."""
    bound = 2.0 * d * _U * np.sqrt(maxnorm_a * maxnorm_b) * 0.5 / (sigma * sigma)
    bound += 2.0 * _U * (maxnorm_a + maxnorm_b) * 0.5 / (sigma * sigma)
    return bool(bound < EXPANSION_TOL)


def device_max(t) -> float:
    """Largest entry of a device vector, read on the host (0 for an empty one)."""
    return float(t.max()) if t is not None and t.numel() else 0.0


def check_expansion(maxnorm_a: float, maxnorm_b: float, d: int, sigma: float, what: str) -> None:
    """Raise when the expanded Gaussian (the MFMA cross-term form) would leave its error bound
    for rows of these squared norms; paths with no direct-form fallback call this."""
    if not expansion_ok(maxnorm_a, maxnorm_b, d, sigma):
        raise ValueError(f"{what}: the expanded Gaussian form exceeds its error bound "
                         f"({EXPANSION_TOL:g} relative) for representations of squared norm up "
                         f"to {max(maxnorm_a, maxnorm_b):.3g} at sigma = {sigma:g}")


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


def pairwise_ext_device(kind: int, Ad, na: int, Bd, nb: int, d: int, params, symmetric=False,
                        out=None, stream=None):
    """Parameterised VALU pairwise kernels (Sargan, Matern L2, p-distances); returns the
    (nb, na) tensor = the column-major (na, nb) matrix."""
    prm = np.ascontiguousarray(np.asarray(params, dtype=np.float64).reshape(-1))
    K = _empty_kernel_device(na, nb) if out is None else out
    _lib.check(_lib.lib().qmlh_pairwise_ext(
        kind, _lib.ptr(Ad), na, Ad.shape[1], _lib.ptr(Bd), nb, Bd.shape[1], d, prm.ctypes.data,
        prm.size, _lib.ptr(K), na, _lib.SYMMETRIC if symmetric else 0,
        _lib.stream_handle(stream)))
    return K


def _rows_pair(A, B):
    symmetric = A is B
    A = np.asarray(A, dtype=np.float64)
    B = A if symmetric else np.asarray(B, dtype=np.float64)
    return A, B, symmetric


def linear_kernel(A, B):
    """K_ij = A_i . B_j on FP64 MFMA; (na, nb), Fortran order (qml/kernels.py:94-115).
    Divergence: representations of different sizes raise ValueError (the reference does not
    check them)."""
    _check_pair(A, B)
    A, B, symmetric = _rows_pair(A, B)
    na, nb, d = A.shape[0], B.shape[0], A.shape[1]
    if na == 0 or nb == 0:
        return np.empty((na, nb), order="F")
    Ad = rows_to_device(A)
    Bd = Ad if symmetric else rows_to_device(B)
    K = _empty_kernel_device(na, nb)
    _lib.check(_lib.lib().qmlh_linear_mfma(
        _lib.ptr(Ad), na, Ad.shape[1], _lib.ptr(Bd), nb, Bd.shape[1], d, _lib.ptr(K), na,
        _lib.SYMMETRIC if symmetric else 0, _lib.stream_handle(None)))
    return _to_host_fortran(K)


def _ext_kernel(kind, A, B, params):
    _check_pair(A, B)
    A, B, symmetric = _rows_pair(A, B)
    na, nb, d = A.shape[0], B.shape[0], A.shape[1]
    if na == 0 or nb == 0:
        return np.empty((na, nb), order="F")
    Ad = rows_to_device(A)
    Bd = Ad if symmetric else rows_to_device(B)
    return _to_host_fortran(pairwise_ext_device(kind, Ad, na, Bd, nb, d, params, symmetric))


def sargan_kernel(A, B, sigma, gammas):
    """K_ij = exp(-|A_i - B_j|_1 / sigma) (1 + sum_k gamma_k |A_i - B_j|_1^k / sigma^k)
    (qml/kernels.py:117-150); no gammas gives the Laplacian kernel."""
    ng = len(gammas)
    if ng == 0:
        return laplacian_kernel(A, B, sigma)
    return _ext_kernel(_lib.PAIR_SARGAN, A, B,
                       [float(sigma)] + [float(g) for g in np.asarray(gammas).reshape(-1)])


def matern_kernel(A, B, sigma, order=0, metric="l1"):
    """Matern kernel of order 0, 1, 2 in the L1 or L2 distance (qml/kernels.py:152-216): the L1
    form is a Sargan kernel with gammas (), (1), (1, 1/3) and sigma / sqrt(3), sigma / sqrt(5);
    the L2 form follows fmatern_kernel_l2 (orders other than 0 and 1 take the order-2 form,
    as its Fortran does). Unknown orders (L1) and metrics print the reference's message and
    raise SystemExit."""
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
            print("Order:%d not implemented in Matern Kernel" % order)
            raise SystemExit
        return sargan_kernel(A, B, sigma, gammas)
    elif metric == "l2":
        pass
    else:
        print("Error: Unknown distance metric %s in Matern kernel" % str(metric))
        raise SystemExit
    return _ext_kernel(_lib.PAIR_MATERN_L2, A, B,
                       [float(sigma), float(order if order in (0, 1) else 2)])


# --------------------------------------------------------------------------------------------
# Local (atom-summed) kernels: qml/kernels.py:218-298
# --------------------------------------------------------------------------------------------

def local_panels(offsets) -> np.ndarray:
    """Host panel partition of a ragged molecule list (see qmlh_local_panels)."""
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    nmol = off.shape[0] - 1
    panels = np.zeros(nmol + 1, dtype=np.int64)
    np_ = _lib.load_library().qmlh_local_panels(off.ctypes.data, nmol, panels.ctypes.data)
    if np_ < 0:
        _lib.check(int(np_))
    return panels[:np_ + 1].copy()


class LocalSet:
    """Atom representations of a molecule list resident in HBM, with what the local-kernel
    launch needs: rows (atoms, ld) FP64, atom offsets and panels on the device, and (for the
    Gaussian expanded form) the squared row norms."""

    def __init__(self, X, n_atoms, ld: int | None = None):
        import torch

        counts = np.asarray(n_atoms, dtype=np.int64).reshape(-1)
        self.nmol = counts.shape[0]
        self.offsets = np.zeros(self.nmol + 1, dtype=np.int64)
        np.cumsum(counts, out=self.offsets[1:])
        self.atoms = int(self.offsets[-1])
        if isinstance(X, torch.Tensor) and X.is_cuda and X.dtype == torch.float64:
            # Resident rows (e.g. generate_atomic_coulomb_matrices): zero pad columns included.
            self.X, self.d = X, X.shape[1]
        else:
            Xh = np.asarray(X, dtype=np.float64).reshape(self.atoms, -1)
            self.d = Xh.shape[1]
            self.X = rows_to_device(Xh, ld)
        self.panels = local_panels(self.offsets)
        dev = self.X.device
        self.offsets_d = torch.from_numpy(self.offsets).to(dev)
        self.panels_d = torch.from_numpy(self.panels).to(dev)
        self._norms = None
        self._maxnorm = None

    @property
    def npanels(self) -> int:
        return self.panels.shape[0] - 1

    def norms(self):
        if self._norms is None:
            self._norms = row_sqnorms_device(self.X, self.atoms, self.d) if self.atoms else \
                self.X.new_zeros(1)
        return self._norms

    def maxnorm(self) -> float:
        if self._maxnorm is None:
            self._maxnorm = float(self.norms().max().item()) if self.atoms else 0.0
        return self._maxnorm


def local_kernels_device(kind: int, SA: LocalSet, SB: LocalSet, sigmas, symmetric=False,
                         expanded: bool | None = None, out=None, stream=None):
    """Local kernels between two resident molecule sets; returns the (nsig, nmB, nmA) device
    tensor, i.e. K_s(a, b) at [s, b, a] (column-major per sigma).

    kind: _lib.PAIR_GAUSSIAN or _lib.PAIR_LAPLACIAN. ``expanded`` (Gaussian only): FP64 MFMA
    expanded distance (True), direct differences (False), or chosen by the error bound of
    ``expansion_ok`` (None)."""
    import torch

    sig = np.ascontiguousarray(np.asarray(sigmas, dtype=np.float64).reshape(-1))
    if np.any(~(sig > 0)):
        raise ValueError("sigmas must be positive")
    if SA.d != SB.d:
        raise ValueError("expected representations of the same size")
    K = out if out is not None else torch.empty((sig.size, SB.nmol, SA.nmol),
                                                dtype=torch.float64, device=SA.X.device)
    nA = nB = None
    if kind == _lib.PAIR_GAUSSIAN:
        if expanded is None:
            expanded = all(expansion_ok(SA.maxnorm(), SB.maxnorm(), SA.d, s) for s in sig)
        if expanded:
            nA, nB = SA.norms(), SB.norms()
    _lib.check(_lib.lib().qmlh_local_kernels(
        kind, _lib.ptr(SA.X), SA.atoms, SA.X.shape[1], _lib.ptr(nA) if nA is not None else None,
        _lib.ptr(SA.offsets_d), _lib.ptr(SA.panels_d), SA.npanels,
        _lib.ptr(SB.X), SB.atoms, SB.X.shape[1], _lib.ptr(nB) if nB is not None else None,
        _lib.ptr(SB.offsets_d), _lib.ptr(SB.panels_d), SB.npanels, SA.d,
        sig.ctypes.data, sig.size, _lib.ptr(K), SA.nmol, SA.nmol * SB.nmol,
        _lib.SYMMETRIC if symmetric else 0, _lib.stream_handle(stream)))
    return K


def _local_kernels(kind, A, B, na, nb, sigmas):
    assert np.sum(na) == A.shape[0], "Error in A input"
    assert np.sum(nb) == B.shape[0], "Error in B input"
    assert A.shape[1] == B.shape[1], "Error in representation sizes"
    sig = np.asarray(sigmas, dtype=np.float64).reshape(-1)
    symmetric = A is B and np.array_equal(np.asarray(na), np.asarray(nb))
    nma, nmb = len(na), len(nb)
    if nma == 0 or nmb == 0 or sig.size == 0:
        return np.zeros((sig.size, nma, nmb), order="F")
    SA = LocalSet(A, na)
    SB = SA if symmetric else LocalSet(B, nb)
    Kd = local_kernels_device(kind, SA, SB, sig, symmetric)
    # (nsig, nmB, nmA) row-major == (nsig, nmA, nmB) with each K_s column-major.
    return np.transpose(Kd.cpu().numpy(), (0, 2, 1))


def get_local_kernels_gaussian(A, B, na, nb, sigmas):
    """K_s(i, j) = sum_{a in i} sum_{b in j} exp(-|A_a - B_b|_2^2 / (2 sigma_s^2)).

    Same signature, assertions and output shape (nsigmas, N, M) as qml/kernels.py:218-256;
    A and B are all atoms' rows concatenated, na and nb the atom counts per molecule. The
    distances come from the FP64 MFMA expanded form where its error bound allows, else from
    direct differences."""
    return _local_kernels(_lib.PAIR_GAUSSIAN, A, B, na, nb, sigmas)


def get_local_kernels_laplacian(A, B, na, nb, sigmas):
    """K_s(i, j) = sum_{a in i} sum_{b in j} exp(-|A_a - B_b|_1 / sigma_s).

    Same signature, assertions and output shape (nsigmas, N, M) as qml/kernels.py:258-298."""
    return _local_kernels(_lib.PAIR_LAPLACIAN, A, B, na, nb, sigmas)
