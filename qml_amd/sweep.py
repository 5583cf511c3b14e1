"""Batched KRR sweeps: learning curves and (sigma, lambda) grids fitted as one launch sequence
recorded into one hipGraph (libqmlhip's qmlh_krr_batched_*), plus the baseline of the same
fits solved one after another with the single-problem entry points.

Each problem is the plain FP64 fit a reference user writes per problem -- kernel on the first
n rows (qml/kernels.py:33-87), K[diag] += lambda, cho_solve (qml/math.py:61-100) -- with all
problems sharing one representation pool resident in HBM.
"""

from __future__ import annotations

import ctypes
import dataclasses
import time

import numpy as np

from . import _lib
from .autotune import Autotuner, make_key, pow2_bucket
from .kernels import (check_expansion, device_max, gaussian_mfma_device, pairwise_device,
                      row_sqnorms_device)
from .math import add_diagonal_device, potrf_device, potrs_device

KINDS = {"gaussian": _lib.PAIR_GAUSSIAN, "laplacian": _lib.PAIR_LAPLACIAN}
NB_CANDIDATES = (64, 128, 256, 512)
BRANCH_CANDIDATES = (1, 2, 4)


def schedule_name(nb: int, branches: int) -> str:
    return f"nb={nb},branches={branches}"


def parse_schedule(name: str) -> tuple[int, int]:
    fields = dict(kv.split("=") for kv in name.split(","))
    return int(fields["nb"]), int(fields["branches"])


# The autotuner's space: Cholesky block size x number of problem groups on streams of their own.
SCHEDULE_CANDIDATES = tuple(schedule_name(nb, g) for nb in NB_CANDIDATES
                            for g in BRANCH_CANDIDATES)


@dataclasses.dataclass
class SweepProblem:
    n: int
    sigma: float
    lam: float
    kernel: str = "laplacian"
    x_row0: int = 0  # first training row in the pool
    y_offset: int | None = None  # first label; defaults to x_row0


def learning_curve_grid(count: int = 64, n_min: int = 500, n_max: int = 8000, seed: int = 4,
                        sigmas=(500.0, 724.0, 1000.0, 2000.0),
                        lambdas=(1e-10, 1e-8, 1e-6, 1e-4), kernel: str = "laplacian"):
    """``count`` problems with n log-uniform in [n_min, n_max], each paired with a point of the
    (sigma, lambda) grid (cycled), all training on the first n molecules of the pool."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(n_min), np.log(n_max), count)).astype(np.int64)
    grid = [(s, l) for s in sigmas for l in lambdas]
    return [SweepProblem(int(n), grid[i % len(grid)][0], grid[i % len(grid)][1], kernel)
            for i, n in enumerate(ns)]


def _structs(problems):
    arr = (_lib.Problem * len(problems))()
    off = 0
    offsets = []
    for i, p in enumerate(problems):
        if p.kernel not in KINDS:
            raise ValueError(f"kernel must be one of {sorted(KINDS)}")
        arr[i].n = p.n
        arr[i].x_row0 = p.x_row0
        arr[i].y_offset = p.x_row0 if p.y_offset is None else p.y_offset
        arr[i].alpha_offset = off
        arr[i].sigma = p.sigma
        arr[i].lambda_ = p.lam
        arr[i].kind = KINDS[p.kernel]
        offsets.append(off)
        off += p.n
    return arr, offsets, off


def tuning_key(problems, nb: int | None = None, branches: int | None = None) -> str:
    """Key of a schedule decision; a fixed nb or branch count narrows the space, so it is part
    of the key."""
    kinds = "+".join(sorted({p.kernel for p in problems}))
    return make_key("krr_batched", precision="fp64", kinds=kinds,
                    count=pow2_bucket(len(problems)),
                    nmax=pow2_bucket(max(p.n for p in problems)),
                    nb="tuned" if nb is None else nb,
                    branches="tuned" if branches is None else branches)


class BatchedFit:
    """All problems' fits as one plan; ``run`` replays the captured graph (one launch)."""

    def __init__(self, problems, Xd, d: int, y, nb: int | None = None, graph: bool = True,
                 tuner: Autotuner | None = None, stream=None, branches: int | None = None):
        import torch

        if not problems:
            raise ValueError("no problems")
        self.problems = list(problems)
        self.Xd, self.d = Xd, d
        dev = Xd.device
        self.y = torch.as_tensor(y, dtype=torch.float64).to(dev).contiguous()
        self.stream = stream or torch.cuda.Stream(device=dev)
        self._structs, self.offsets, total = _structs(self.problems)
        self.norms = (row_sqnorms_device(Xd, Xd.shape[0], d)
                      if any(p.kernel == "gaussian" for p in self.problems) else None)
        if self.norms is not None:  # the batched Gaussian build has only the expanded form
            mx = device_max(self.norms)
            for p in self.problems:
                if p.kernel == "gaussian":
                    check_expansion(mx, mx, d, p.sigma, "batched Gaussian fit")
        self.alpha = torch.zeros(max(total, 1), dtype=torch.float64, device=dev)
        self.info_d = torch.zeros(len(self.problems), dtype=torch.int32, device=dev)
        self._plan = None
        if nb is None or branches is None:
            cands = [c for c in SCHEDULE_CANDIDATES
                     if (nb is None or parse_schedule(c)[0] == nb)
                     and (branches is None or parse_schedule(c)[1] == branches)]
            choice = (tuner or Autotuner()).tune(tuning_key(self.problems, nb, branches), cands,
                                                 lambda c: self._time_candidate(*parse_schedule(c)))
            nb, branches = parse_schedule(choice)
        self.nb, self.branches = int(nb), int(branches)
        self._make_plan(self.nb, self.branches)
        self.graph = False
        if graph:
            _lib.check(_lib.lib().qmlh_krr_batched_capture(self._plan, self.stream.cuda_stream))
            self.graph = True

    def _make_plan(self, nb, branches):
        import torch

        L = _lib.lib()
        self.close()
        nbytes = L.qmlh_krr_batched_workspace(self._structs, len(self.problems), nb, branches)
        if nbytes == 0:
            _lib.check(-1)
        self.ws = torch.empty(max(1, (nbytes + 7) // 8), dtype=torch.float64, device=self.Xd.device)
        plan = ctypes.c_void_p()
        torch.cuda.synchronize(self.Xd.device)
        _lib.check(L.qmlh_krr_batched_create(
            self._structs, len(self.problems), nb, branches, _lib.ptr(self.Xd), self.Xd.shape[0],
            self.Xd.shape[1], self.d, _lib.ptr(self.norms), _lib.ptr(self.y), self.y.numel(),
            _lib.ptr(self.alpha), self.alpha.numel(), _lib.ptr(self.info_d), _lib.ptr(self.ws),
            self.ws.numel() * 8, ctypes.byref(plan)))
        self._plan = plan

    def _time_candidate(self, nb, branches) -> float:
        self._make_plan(nb, branches)
        L = _lib.lib()
        h = self.stream.cuda_stream
        _lib.check(L.qmlh_krr_batched_run(self._plan, h))  # warm
        self.stream.synchronize()
        t0 = time.perf_counter()
        _lib.check(L.qmlh_krr_batched_run(self._plan, h))
        self.stream.synchronize()
        return time.perf_counter() - t0

    @property
    def launches(self) -> int:
        return int(_lib.lib().qmlh_krr_batched_launches(self._plan))

    def run(self, sync: bool = True):
        L = _lib.lib()
        h = self.stream.cuda_stream
        if self.graph:
            _lib.check(L.qmlh_krr_batched_launch(self._plan, h))
        else:
            _lib.check(L.qmlh_krr_batched_run(self._plan, h))
        if sync:
            self.stream.synchronize()
        return self

    def run_direct(self, sync: bool = True):
        """Enqueue the same sequence as individual launches (no graph)."""
        _lib.check(_lib.lib().qmlh_krr_batched_run(self._plan, self.stream.cuda_stream))
        if sync:
            self.stream.synchronize()
        return self

    def alphas(self):
        return [self.alpha[o:o + p.n] for o, p in zip(self.offsets, self.problems)]

    def info(self) -> np.ndarray:
        self.stream.synchronize()
        return self.info_d.cpu().numpy()

    def factor(self, i: int):
        """Problem i's lower Cholesky factor as an (n, ld) tensor whose row j is column j."""
        off, ld = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(_lib.lib().qmlh_krr_batched_factor(self._plan, i, ctypes.byref(off),
                                                      ctypes.byref(ld)))
        n = self.problems[i].n
        return self.ws[off.value:off.value + n * ld.value].view(n, ld.value)

    def close(self):
        if self._plan is not None:
            _lib.lib().qmlh_krr_batched_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sequential_fit(problems, Xd, d: int, y, stream=None):
    """The same fits one after another with the single-problem entry points (kernel matrix,
    K += lambda I, potrf, potrs), all on one stream; returns the alpha tensors."""
    import torch

    dev = Xd.device
    y = torch.as_tensor(y, dtype=torch.float64).to(dev)
    nmax = max(p.n for p in problems)
    K = torch.empty(nmax * nmax, dtype=torch.float64, device=dev)
    norms = (row_sqnorms_device(Xd, Xd.shape[0], d, stream)
             if any(p.kernel == "gaussian" for p in problems) else None)
    out = []
    for p in problems:
        n = p.n
        Kp = K[:n * n].view(n, n)
        X = Xd[p.x_row0:p.x_row0 + n]
        if p.kernel == "gaussian":
            nr = norms[p.x_row0:p.x_row0 + n]
            gaussian_mfma_device(X, nr, n, X, nr, n, d, p.sigma, True, Kp, stream)
        else:
            pairwise_device(_lib.PAIR_LAPLACIAN, X, n, X, n, d, p.sigma, True, Kp, stream)
        add_diagonal_device(Kp, n, p.lam, stream)
        potrf_device(Kp, n, stream)
        y0 = p.x_row0 if p.y_offset is None else p.y_offset
        a = y[y0:y0 + n].clone()
        potrs_device(Kp, n, a, stream)
        out.append(a)
    return out
