"""Device-resident kernel ridge regression: the KRR composition of
tests/test_energy_krr_cmat.py:101-111 (K = kernel(X, X); K[diag] += lambda;
alpha = cho_solve(K, y); y_pred = kernel(X, Xs).T . alpha) with every array in HBM.

The training representations, their FP64 squared norms and alpha stay on the device after
``fit``; ``predict`` runs the fused kernel-times-alpha reduction, so the (m, n) prediction
kernel is never materialised and K never crosses PCIe.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .device import device, padded_ld, ragged_to_device, rows_to_device
from .kernels import (check_expansion, device_max, expansion_ok, gaussian_mfma_device,
                      pairwise_device, row_sqnorms_device)
from .math import add_diagonal_device, potrf_device, potrs_device, print_potrf_warning
from .representations import cm_length, generate_coulomb_matrices_device

KERNELS = {"gaussian": _lib.PAIR_GAUSSIAN, "laplacian": _lib.PAIR_LAPLACIAN}


class CholeskyError(RuntimeError):
    """The regularised kernel matrix is not numerically positive definite (LAPACK info > 0)."""


def check_labels(y, n: int) -> None:
    """Labels must be one value per training row: the solves read and write n doubles of y.

    Raises the ValueError text of qml/math.py:79-80 (cho_solve's vector-size check) before
    anything touches the device."""
    shape = tuple(getattr(y, "shape", np.shape(y)))
    if shape != (n,):
        raise ValueError('expected matrix and vector of same stride size')


@dataclass
class FitReport:
    n: int
    info: int
    seconds: dict = field(default_factory=dict)
    precision: str = "fp64"
    iterations: int = 0
    relres: float = 0.0
    shift: float = 0.0
    # normwise backward error |y - A alpha| / (|A|_inf |alpha| + |y|) of the mixed fit, when
    # its recomputed residual is above the tolerance (-1: the residual test passed)
    backward_error: float = -1.0
    # mixed fit: one entry per factorisation attempt, (shift used, factor info, capacitance
    # info; None when not reached)
    attempts: list = field(default_factory=list)


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
                 preconditioner. lowrank sets how Y = L^-1 V is formed: "fp32" -- V rounded to
                 FP32 and solved during the factorisation itself as extra panel rows (free);
                 "fp64" -- V in FP64, solved after the factorisation with the FP32 factor in
                 FP64 arithmetic (n^2 (d+1) FP64 MFMA flops); "hybrid" -- the FP32 solve, with
                 its first lowrank_cols columns (u, the direction of the largest eigenvalues,
                 first) solved again in FP64 (one single-vector solve each); "auto" (default)
                 -- "hybrid" with one FP64 column. The FP32 Y's error is amplified through L^-1
                 by up to 1 / sqrt(lambda) and costs CG iterations; on the sigma = 4000
                 Coulomb-matrix workloads the FP64 u column alone recovers what the whole FP64
                 term does (50,000 molecules: 8 -> 4 iterations at lambda 1e-6, 70 -> 29 at
                 1e-9) for one triangular solve instead of d + 1.
    """

    def __init__(self, kernel: str = "gaussian", sigma: float = 4000.0, lam: float = 1e-9,
                 precision: str = "fp64", tol: float = 1e-10, max_iter: int = 200,
                 factor: str | None = None, lowrank: str = "auto", lowrank_cols: int = 1,
                 pivot_floor: float = 0.0):
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
        if lowrank not in ("auto", "fp32", "fp64", "hybrid"):
            raise ValueError("lowrank must be 'auto', 'fp32', 'fp64' or 'hybrid'")
        self.lowrank = lowrank
        self.lowrank_cols = int(lowrank_cols)
        # Static pivoting of the low-precision factor: pivots at or below
        # pivot_floor * (lambda + shift) are raised to it (0: a breakdown shifts the rest of the
        # factor, see _fit_mixed).
        if not pivot_floor >= 0.0:
            raise ValueError("pivot_floor must be >= 0")
        self.pivot_floor = float(pivot_floor)
        self.Xd = None
        self.norms = None
        self.alpha = None
        self.n = 0
        self.d = 0

    # -- kernel matrix ------------------------------------------------------------------------
    def build_kernel(self, Xd, n, d, norms=None, out=None, stream=None, lower=False):
        """Symmetric (n, n) training kernel, column-major, as an (n, n) device tensor (lower: only
        the lower triangle is written, which is all a Cholesky and the CG's matvec read)."""
        if self.kernel == "gaussian":
            if norms is None:
                norms = row_sqnorms_device(Xd, n, d, stream)
            mx = device_max(norms)
            if expansion_ok(mx, mx, d, self.sigma):
                return gaussian_mfma_device(Xd, norms, n, Xd, norms, n, d, self.sigma, True, out,
                                            stream, lower=lower)
            # outside the expansion's error bound: the direct-difference form
            return pairwise_device(_lib.PAIR_GAUSSIAN, Xd, n, Xd, n, d, self.sigma, True, out,
                                   stream, lower=lower)
        return pairwise_device(_lib.PAIR_LAPLACIAN, Xd, n, Xd, n, d, self.sigma, True, out, stream,
                               lower=lower)

    # -- fit ----------------------------------------------------------------------------------
    def fit_device(self, Xd, d: int, y, raise_on_error: bool = True, stream=None) -> FitReport:
        """Fit from an (n, ld) device representation tensor; y host or device (n,)."""
        check_labels(y, Xd.shape[0])
        if self.precision == "mixed":
            return self._fit_mixed(Xd, d, y, raise_on_error, stream)
        import torch

        n = Xd.shape[0]
        rep = FitReport(n=n, info=0)
        t0 = time.perf_counter()
        norms = row_sqnorms_device(Xd, n, d, stream) if self.kernel == "gaussian" else None
        K = self.build_kernel(Xd, n, d, norms, stream=stream, lower=True)
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
        mode = self.lowrank_mode() if r else "fp32"
        y64 = mode != "fp32"
        # B32 column-major with ld = n + rx: with an FP32 (or hybrid) Woodbury term the r rows
        # below the matrix carry V^T through the factorisation, which solves them into
        # (L^-1 V)^T along the way; with an FP64 one, Y is solved after the factorisation.
        rx = r if mode in ("fp32", "hybrid") else 0
        ldb = n + rx
        B = torch.empty((n, ldb), dtype=torch.float32, device=dev)
        V64 = None
        if self.kernel == "gaussian":
            norms = row_sqnorms_device(Xd, n, d, stream)
            mx = device_max(norms)  # the deflation and the residuals use the expanded form
            check_expansion(mx, mx, d, self.sigma, "mixed Gaussian fit")
            u = torch.empty(n, dtype=torch.float64, device=dev)
            # (n, r) column-major: V (FP32) rides through the factorisation; V64 the FP64 term
            V = torch.empty((r, n), dtype=torch.float32, device=dev) if rx else None
            V64 = torch.empty((r, n), dtype=torch.float64, device=dev) if y64 else None
            C = torch.empty((r, r), dtype=torch.float64, device=dev)
        else:
            self.build_kernel(Xd, n, d, out=K, stream=stream, lower=True)
            add_diagonal_device(K, n, self.lam, stream)
        # Low-precision factor and Woodbury term. Where FP32 rounding of the factor can reach
        # lambda -- lambda below 0.2 u32 max_i B_ii sqrt(n), the scale of the rounding that the
        # trailing updates accumulate -- the first factorisation shifts on breakdown: a pivot
        # at or below 2 lambda counts as failed, a diagonal block with one is factored again
        # with its diagonal raised (by 2 lambda, then 20 lambda) and every later block keeps that
        # shift, so one pass yields the factor of K + diag(0, .., 0, s, ..) instead of a
        # breakdown near the end and a second factorisation. The CG on the FP64 matrix still
        # converges to the unshifted solution. (With lambda well above that scale pivots just
        # above lambda are genuine and a shift would only slow the CG: there the factor runs
        # unshifted.) When the factor fails even so, or the capacitance matrix of the Woodbury
        # term is not numerically positive definite, retry from the matrix with a full shift.
        shifts = [0.0] + [max(self.lam, 1e-12) * 10.0 ** k for k in range(1, 9)]
        shift_dev = torch.zeros(1, dtype=torch.float64, device=dev)
        bws = None
        Yw = None  # the Woodbury term handed to the CG: V (FP32) or V64, after the solve
        for attempt, shift in enumerate(shifts):
            if self.kernel == "gaussian":
                _lib.check(L.qmlh_gaussian_deflated(
                    _lib.ptr(Xd), n, Xd.shape[1], _lib.ptr(norms), d, self.sigma, self.lam, shift,
                    _lib.ptr(K) if attempt == 0 else None, n, _lib.ptr(B), ldb,
                    _lib.ptr(V) if rx else None, n, _lib.ptr(u), h))
            else:
                _lib.check(L.qmlh_lower_to_f32(_lib.ptr(K), n, n, _lib.ptr(B), ldb, shift, h))
            split16 = self.factor == "f16x3"
            if split16 and bws is None:
                nb = L.qmlh_spotrf_f16x3_workspace(n)
                bws = torch.empty(max(1, (nb + 7) // 8), dtype=torch.float64, device=dev)
            bshift = 0.0
            if attempt == 0 and self.pivot_floor == 0.0:  # (static pivoting: not combined)
                s_ = stream if stream is not None else torch.cuda.current_stream()
                with torch.cuda.stream(s_):
                    dmax_t = torch.diagonal(B).max()
                s_.synchronize()
                dmax = float(dmax_t.item())
                if self.lam < 0.2 * 2.0 ** -24 * dmax * math.sqrt(n):
                    bshift = 2.0 * self.lam
            _lib.check(L.qmlh_spotrf_lowrank_shift(
                _lib.ptr(B), n, ldb, _lib.ptr(V) if rx else None, n, rx, int(split16),
                self.pivot_floor * (self.lam + shift), bshift, bshift, _lib.ptr(shift_dev),
                _lib.ptr(info), _lib.ptr(bws) if split16 else None,
                bws.numel() * 8 if split16 else 0, h))
            rep.info = _lib.host_int(info, stream)
            rep.shift = float(shift_dev.item()) if attempt == 0 else shift
            rep.attempts.append((rep.shift, rep.info, None))
            if rep.info != 0:
                continue
            if not r:
                break
            Yw = V
            if y64:  # V64 = u [1, x / sigma] in FP64, then Y = L^-1 V64 in FP64 arithmetic
                _lib.check(L.qmlh_gaussian_lowrank(_lib.ptr(Xd), n, Xd.shape[1], d, self.sigma,
                                                   _lib.ptr(u), _lib.ptr(V64), n, h))
                if mode == "fp64":
                    tws = torch.empty(max(1, L.qmlh_trsm_f32l_workspace(n) // 8),
                                      dtype=torch.float64, device=dev)
                    _lib.check(L.qmlh_trsm_f32l(_lib.ptr(B), n, ldb, _lib.ptr(V64), n, r,
                                                _lib.ptr(tws), tws.numel() * 8, h))
                else:  # hybrid: the factorisation's FP32 Y, leading columns again in FP64
                    k = max(0, min(self.lowrank_cols, r))
                    tws = torch.empty(max(1, L.qmlh_lowrank_hybrid_workspace(n) // 8 + 1),
                                      dtype=torch.float64, device=dev)
                    _lib.check(L.qmlh_lowrank_hybrid(_lib.ptr(B), n, ldb, _lib.ptr(V), n,
                                                     _lib.ptr(V64), n, r, k, _lib.ptr(tws),
                                                     tws.numel() * 8, h))
                del tws
                Yw = V64
            # C = chol(I + Y^T Y)
            gws = torch.empty(max(1, L.qmlh_woodbury_capacitance_workspace(n, r) // 8),
                              dtype=torch.float64, device=dev)
            cap = L.qmlh_woodbury_capacitance_f64 if y64 else L.qmlh_woodbury_capacitance
            _lib.check(cap(_lib.ptr(Yw), n, n, r, _lib.ptr(C), r, _lib.ptr(info), _lib.ptr(gws),
                           gws.numel() * 8, h))
            del gws
            cinfo = _lib.host_int(info, stream)
            rep.attempts[-1] = (rep.shift, rep.info, cinfo)
            if cinfo == 0:
                break
            rep.info = -1  # capacitance matrix not positive definite: next shift
        if rep.info != 0:
            if raise_on_error:
                raise CholeskyError("FP32 factorisation (or its Woodbury capacitance) failed at "
                                    f"every shift: info={rep.info}")
            return rep
        V = Yw
        alpha = torch.zeros(n, dtype=torch.float64, device=dev)
        yd = torch.as_tensor(y, dtype=torch.float64).to(dev).contiguous()
        ws_bytes = L.qmlh_pcg_workspace(n, r)
        ws = torch.empty(max(1, ws_bytes // 8 + 1), dtype=torch.float64, device=dev)
        iters = ctypes.c_int32()
        relres = ctypes.c_double()
        backward = ctypes.c_double()
        _lib.check(L.qmlh_pcg_solve(_lib.ptr(K), n, n, _lib.ptr(B), ldb, _lib.ptr(V), n, r,
                                    int(y64), _lib.ptr(C), r, _lib.ptr(yd), _lib.ptr(alpha), self.tol,
                                    self.max_iter, _lib.ptr(ws), ws.numel() * 8,
                                    ctypes.byref(iters), ctypes.byref(relres),
                                    ctypes.byref(backward), h))
        rep.iterations, rep.relres = iters.value, relres.value
        rep.backward_error = backward.value
        rep.seconds["fit"] = time.perf_counter() - t0
        del K, B, ws, V, C, bws
        if norms is None and self.kernel == "gaussian":
            norms = row_sqnorms_device(Xd, n, d, stream)
        self.Xd, self.norms, self.alpha, self.n, self.d = Xd, norms, alpha, n, d
        if not self.converged(rep) and raise_on_error:
            raise CholeskyError(f"refinement stopped at relative residual {rep.relres:.3e} "
                                f"(backward error {rep.backward_error:.3e}) after "
                                f"{rep.iterations} iterations")
        return rep

    def lowrank_mode(self) -> str:
        """How the Woodbury term is formed: "fp32", "fp64" or "hybrid" (class docstring)."""
        return "hybrid" if self.lowrank == "auto" else self.lowrank

    def lowrank_fp64(self) -> bool:
        """Whether the Woodbury term Y is held in FP64 (wholly or in part)."""
        return self.lowrank_mode() != "fp32"

    # Normwise backward error accepted as the FP64 floor: the unit roundoff u = 2^-53. Below it,
    # alpha solves a system whose matrix lies within the rounding of K's own FP64 entries; a
    # backward-stable FP64 Cholesky solve reaches about u / 10 on these kernels.
    BACKWARD_FLOOR = 0.5 * float(np.finfo(np.float64).eps)

    def converged(self, rep: FitReport) -> bool:
        """The recomputed residual meets the tolerance, or the solution is at the FP64 floor
        (its backward error is that of an FP64 solve)."""
        if rep.relres <= self.tol:
            return True
        return 0.0 <= rep.backward_error <= self.BACKWARD_FLOOR

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
        if self.kernel == "gaussian":
            if qnorms is None:
                qnorms = row_sqnorms_device(Qd, m, self.d, stream)
            # the fused predict kernel has only the expanded form
            check_expansion(device_max(qnorms), self.max_norm(), self.d, self.sigma,
                            "Gaussian prediction")
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

    def max_norm(self) -> float:
        """Largest squared norm of the training representations (Gaussian models)."""
        if getattr(self, "_max_norm", None) is None or self._max_norm[0] is not self.norms:
            self._max_norm = (self.norms, device_max(self.norms))
        return self._max_norm[1]

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
