"""Device-resident kernel ridge regression: the KRR composition of
tests/test_energy_krr_cmat.py:101-111 (K = kernel(X, X); K[diag] += lambda;
alpha = cho_solve(K, y); y_pred = kernel(X, Xs).T . alpha) with every array in HBM.

The training representations, their FP64 squared norms and alpha stay on the device after
``fit``; ``predict`` runs the fused kernel-times-alpha reduction, so the (m, n) prediction
kernel is never materialised and K never crosses PCIe.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .device import device, padded_ld, ragged_to_device, rows_to_device
from .kernels import gaussian_mfma_device, pairwise_device, row_sqnorms_device
from .math import add_diagonal_device, potrf_device, potrs_device, print_potrf_warning
from .representations import cm_length, generate_coulomb_matrices_device

KERNELS = {"gaussian": _lib.PAIR_GAUSSIAN, "laplacian": _lib.PAIR_LAPLACIAN}


class CholeskyError(RuntimeError):
    """The regularised kernel matrix is not numerically positive definite (LAPACK info > 0)."""


@dataclass
class FitReport:
    n: int
    info: int
    seconds: dict = field(default_factory=dict)
    precision: str = "fp64"
    iterations: int = 0
    relres: float = 0.0
    shift: float = 0.0


class KRR:
    """Kernel ridge regression with the training set resident in HBM.

    kernel: "gaussian" (FP64 MFMA expanded form) or "laplacian" (FP64 VALU).
    precision:
      "fp64"  -- K in FP64, FP64 MFMA Cholesky, two triangular solves (the drop-in path);
      "mixed" -- K in FP64 for the residuals, a Cholesky factor in FP32 storage (trailing
                 updates as split-FP16 MFMA products, factor="f16x3", the default, or FP32 MFMA,
                 factor="fp32"; FP64 diagonal blocks) as preconditioner, FP64
                 conjugate-gradient refinement to
                 |y - K alpha| <= tol |y|. For the Gaussian kernel the factor is of the
                 deflated matrix B = K - V V^T (V = u [1, x / sigma], u_i =
                 exp(-|x_i|^2 / 2 sigma^2); the constant and linear Taylor terms of the
                 kernel), with the rank-(d+1) term restored by a Woodbury correction inside the
                 preconditioner; Y = L^-1 V is solved during the factorisation itself.
    """

    def __init__(self, kernel: str = "gaussian", sigma: float = 4000.0, lam: float = 1e-9,
                 precision: str = "fp64", tol: float = 1e-10, max_iter: int = 200,
                 factor: str | None = None):
        if kernel not in KERNELS:
            raise ValueError(f"kernel must be one of {sorted(KERNELS)}")
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.kernel = kernel
        self.sigma = float(sigma)
        self.lam = float(lam)
        self.precision = precision
        self.tol = float(tol)
        self.max_iter = int(max_iter)
        # Mixed-precision factor: "f16x3" (split-FP16 trailing updates, the faster) or "fp32"
        # (FP32 MFMA trailing updates).
        if factor is None:
            factor = "f16x3"
        if factor not in ("fp32", "f16x3"):
            raise ValueError("factor must be 'fp32' or 'f16x3'")
        self.factor = factor
        self.Xd = None
        self.norms = None
        self.alpha = None
        self.n = 0
        self.d = 0

    # -- kernel matrix ------------------------------------------------------------------------
    def build_kernel(self, Xd, n, d, norms=None, out=None, stream=None):
        """Symmetric (n, n) training kernel, column-major, as an (n, n) device tensor."""
        if self.kernel == "gaussian":
            if norms is None:
                norms = row_sqnorms_device(Xd, n, d, stream)
            return gaussian_mfma_device(Xd, norms, n, Xd, norms, n, d, self.sigma, True, out,
                                        stream)
        return pairwise_device(_lib.PAIR_LAPLACIAN, Xd, n, Xd, n, d, self.sigma, True, out, stream)

    # -- fit ----------------------------------------------------------------------------------
    def fit_device(self, Xd, d: int, y, raise_on_error: bool = True, stream=None) -> FitReport:
        """Fit from an (n, ld) device representation tensor; y host or device (n,)."""
        if self.precision == "mixed":
            return self._fit_mixed(Xd, d, y, raise_on_error, stream)
        import torch

        n = Xd.shape[0]
        rep = FitReport(n=n, info=0)
        t0 = time.perf_counter()
        norms = row_sqnorms_device(Xd, n, d, stream) if self.kernel == "gaussian" else None
        K = self.build_kernel(Xd, n, d, norms, stream=stream)
        add_diagonal_device(K, n, self.lam, stream)
        info = potrf_device(K, n, stream)
        alpha = torch.as_tensor(y, dtype=torch.float64).to(Xd.device).clone()
        potrs_device(K, n, alpha, stream)
        rep.info = _lib.host_int(info, stream)
        rep.seconds["fit"] = time.perf_counter() - t0
        del K
        if rep.info != 0:
            print_potrf_warning(rep.info)
            if raise_on_error:
                raise CholeskyError(f"Cholesky failed: info={rep.info}")
        self.Xd, self.norms, self.alpha, self.n, self.d = Xd, norms, alpha, n, d
        return rep

    def _fit_mixed(self, Xd, d: int, y, raise_on_error: bool, stream) -> FitReport:
        import ctypes

        import torch

        n = Xd.shape[0]
        dev = Xd.device
        L = _lib.lib()
        h = _lib.stream_handle(stream)
        rep = FitReport(n=n, info=0, precision="mixed")
        t0 = time.perf_counter()
        K = torch.empty((n, n), dtype=torch.float64, device=dev)
        info = torch.zeros(1, dtype=torch.int32, device=dev)
        norms = u = V = C = None
        r = d + 1 if self.kernel == "gaussian" else 0
        # B32 column-major with ld = n + r: the r rows below the matrix carry V^T through the
        # factorisation, which solves them into (L^-1 V)^T along the way.
        ldb = n + r
        B = torch.empty((n, ldb), dtype=torch.float32, device=dev)
        if self.kernel == "gaussian":
            norms = row_sqnorms_device(Xd, n, d, stream)
            u = torch.empty(n, dtype=torch.float64, device=dev)
            V = torch.empty((r, n), dtype=torch.float32, device=dev)  # (n, r) column-major
            C = torch.empty((r, r), dtype=torch.float64, device=dev)
        else:
            self.build_kernel(Xd, n, d, out=K, stream=stream)
            add_diagonal_device(K, n, self.lam, stream)
        # Low-precision factor; on breakdown retry with a diagonal shift (the CG on the FP64
        # matrix still converges to the unshifted solution, only more slowly).
        shifts = [0.0] + [max(self.lam, 1e-12) * 10.0 ** k for k in range(1, 9)]
        bws = None
        for attempt, shift in enumerate(shifts):
            if self.kernel == "gaussian":
                _lib.check(L.qmlh_gaussian_deflated(
                    _lib.ptr(Xd), n, Xd.shape[1], _lib.ptr(norms), d, self.sigma, self.lam, shift,
                    _lib.ptr(K) if attempt == 0 else None, n, _lib.ptr(B), ldb, _lib.ptr(V), n,
                    _lib.ptr(u), h))
            else:
                _lib.check(L.qmlh_lower_to_f32(_lib.ptr(K), n, n, _lib.ptr(B), ldb, shift, h))
            split16 = self.factor == "f16x3"
            if split16 and bws is None:
                nb = L.qmlh_spotrf_f16x3_workspace(n)
                bws = torch.empty(max(1, (nb + 7) // 8), dtype=torch.float64, device=dev)
            _lib.check(L.qmlh_spotrf_lowrank(_lib.ptr(B), n, ldb, _lib.ptr(V) if r else None, n,
                                             r, int(split16), _lib.ptr(info),
                                             _lib.ptr(bws) if split16 else None,
                                             bws.numel() * 8 if split16 else 0, h))
            rep.info = _lib.host_int(info, stream)
            rep.shift = shift
            if rep.info == 0:
                break
        if rep.info != 0:
            if raise_on_error:
                raise CholeskyError(f"FP32 factorisation failed at every shift: info={rep.info}")
            return rep
        if r:  # V now holds Y = L^-1 V
            gws = torch.empty(max(1, L.qmlh_woodbury_capacitance_workspace(n, r) // 8),
                              dtype=torch.float64, device=dev)
            _lib.check(L.qmlh_woodbury_capacitance(_lib.ptr(V), n, n, r, _lib.ptr(C), r,
                                                   _lib.ptr(info), _lib.ptr(gws),
                                                   gws.numel() * 8, h))
            del gws
            if _lib.host_int(info, stream) != 0:
                raise CholeskyError("Woodbury capacitance matrix is not positive definite")
        alpha = torch.zeros(n, dtype=torch.float64, device=dev)
        yd = torch.as_tensor(y, dtype=torch.float64).to(dev).contiguous()
        ws_bytes = L.qmlh_pcg_workspace(n, r)
        ws = torch.empty(max(1, ws_bytes // 8 + 1), dtype=torch.float64, device=dev)
        iters = ctypes.c_int32()
        relres = ctypes.c_double()
        _lib.check(L.qmlh_pcg_solve(_lib.ptr(K), n, n, _lib.ptr(B), ldb, _lib.ptr(V), n, r,
                                    _lib.ptr(C), r, _lib.ptr(yd), _lib.ptr(alpha), self.tol,
                                    self.max_iter, _lib.ptr(ws), ws.numel() * 8,
                                    ctypes.byref(iters), ctypes.byref(relres), h))
        rep.iterations, rep.relres = iters.value, relres.value
        rep.seconds["fit"] = time.perf_counter() - t0
        del K, B, ws, V, C, bws
        if norms is None and self.kernel == "gaussian":
            norms = row_sqnorms_device(Xd, n, d, stream)
        self.Xd, self.norms, self.alpha, self.n, self.d = Xd, norms, alpha, n, d
        if not (rep.relres <= self.tol) and raise_on_error:
            raise CholeskyError(f"refinement stopped at relative residual {rep.relres:.3e} "
                                f"after {rep.iterations} iterations")
        return rep

    def fit(self, X, y, **kw) -> FitReport:
        """Fit from host (n, d) representations."""
        X = np.asarray(X, dtype=np.float64)
        return self.fit_device(rows_to_device(X), X.shape[1], y, **kw)

    def fit_molecules(self, Z, R, offsets, size: int, y, sorting: str = "row-norm", **kw):
        """Fit straight from raw molecules: Coulomb matrices are generated on the GPU."""
        Zd, Rd, Od = ragged_to_device(Z, R, offsets)
        Xd, _ = generate_coulomb_matrices_device(Zd, Rd, Od, size, sorting)
        return self.fit_device(Xd, cm_length(size), y, **kw)

    # -- predict ------------------------------------------------------------------------------
    def predict_device(self, Qd, m: int, out=None, workspace=None, qnorms=None, stream=None):
        """Predictions for (m, ld) device query rows; returns an (m,) device tensor."""
        import torch

        if self.alpha is None:
            raise RuntimeError("fit first")
        if Qd.shape[1] != self.Xd.shape[1]:
            raise ValueError("query rows must use the training feature layout")
        y = torch.empty(m, dtype=torch.float64, device=Qd.device) if out is None else out
        if m == 0:
            return y
        L = _lib.lib()
        ws_bytes = L.qmlh_predict_workspace(m, self.n)
        if workspace is None or workspace.numel() * workspace.element_size() < ws_bytes:
            workspace = torch.empty(max(1, ws_bytes // 8), dtype=torch.float64, device=Qd.device)
        if self.kernel == "gaussian" and qnorms is None:
            qnorms = row_sqnorms_device(Qd, m, self.d, stream)
        _lib.check(L.qmlh_krr_predict(
            KERNELS[self.kernel], _lib.ptr(Qd), m, Qd.shape[1], _lib.ptr(qnorms),
            _lib.ptr(self.Xd), self.n, self.Xd.shape[1], _lib.ptr(self.norms),
            _lib.ptr(self.alpha), self.d, self.sigma, _lib.ptr(y), _lib.ptr(workspace),
            workspace.numel() * workspace.element_size(), _lib.stream_handle(stream)))
        return y

    def predict(self, Xq) -> np.ndarray:
        Xq = np.asarray(Xq, dtype=np.float64)
        Qd = rows_to_device(Xq, ld=self.Xd.shape[1])
        return self.predict_device(Qd, Xq.shape[0]).cpu().numpy()

    def predict_molecules(self, Z, R, offsets, size: int, sorting: str = "row-norm"):
        Zd, Rd, Od = ragged_to_device(Z, R, offsets)
        Qd, _ = generate_coulomb_matrices_device(Zd, Rd, Od, size, sorting, ld=self.Xd.shape[1])
        return self.predict_device(Qd, Qd.shape[0]).cpu().numpy()

    # -- checkpoint ---------------------------------------------------------------------------
    def save(self, path: str) -> None:
        """npz snapshot of what serving needs: representations, norms, alpha, hyperparameters."""
        np.savez(path, X=self.Xd[:, :self.d].cpu().numpy(), alpha=self.alpha.cpu().numpy(),
                 kernel=self.kernel, sigma=self.sigma, lam=self.lam, d=self.d)

    @classmethod
    def load(cls, path: str) -> "KRR":
        z = np.load(path, allow_pickle=False)
        m = cls(str(z["kernel"]), float(z["sigma"]), float(z["lam"]))
        m.d = int(z["d"])
        m.Xd = rows_to_device(z["X"])
        m.n = m.Xd.shape[0]
        m.norms = row_sqnorms_device(m.Xd, m.n, m.d) if m.kernel == "gaussian" else None
        import torch

        m.alpha = torch.from_numpy(z["alpha"]).to(m.Xd.device)
        return m
