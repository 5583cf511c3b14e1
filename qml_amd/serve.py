"""Continuous-batching KRR prediction server with the training set resident in HBM.

Requests carry raw molecules (nuclear charges and coordinates); a result is the KRR
prediction for each molecule. The scheduler thread keeps up to ``max_inflight`` batches on
the device at once, each on its own HIP stream with its own preallocated device buffers and
pinned host staging. Every tick it

  1. retires batches whose completion event has fired (results scattered back to their
     requests, latency = arrival -> result on the host), then
  2. admits queued requests, first come first served, into a free slot until the slot's
     molecule capacity is reached, and launches that batch: host -> device copy, Coulomb
     matrices (one launch for the whole batch), the fused kernel-times-alpha reduction, device
     -> host copy, event.

So requests that arrive while other batches are on the GPU join the next free slot and their
work overlaps the batches in flight. Refit requests (a new model from new training molecules)
run on a separate stream in their own worker and are admitted only when the free device
memory, minus a reserve, covers their estimated footprint.

The predict kernels reduce each query over the training set in a fixed chunk order that does
not depend on the batch, so a request's predictions are bitwise the same alone or in any mix.
"""

from __future__ import annotations

import collections
import dataclasses
import threading
import time
from typing import Callable, Optional

import numpy as np

from . import _lib
from .device import padded_ld
from .krr import KERNELS, KRR, check_labels
from .representations import cm_length

# --------------------------------------------------------------------------------------------
# Requests
# --------------------------------------------------------------------------------------------


class PredictRequest:
    """Molecules in ragged form: Z (total,), R (total, 3), offsets (nmol + 1,)."""

    def __init__(self, Z, R, offsets):
        self.Z = np.ascontiguousarray(Z, dtype=np.float64)
        self.R = np.ascontiguousarray(R, dtype=np.float64).reshape(-1, 3)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        self.nmol = int(self.offsets.shape[0] - 1)
        self.natoms = int(self.offsets[-1] - self.offsets[0])
        self.t_arrival = time.perf_counter()
        self.t_done: Optional[float] = None
        self.result: Optional[np.ndarray] = None
        self.error: Optional[BaseException] = None
        self._done = threading.Event()

    def wait(self, timeout=None) -> np.ndarray:
        if not self._done.wait(timeout):
            raise TimeoutError("prediction request not finished")
        if self.error is not None:
            raise self.error
        return self.result

    @property
    def latency(self) -> float:
        return (self.t_done - self.t_arrival) if self.t_done is not None else float("nan")


class RefitRequest:
    """Fit a new model on (Z, R, offsets, y); ``result`` is the fitted KRR."""

    def __init__(self, Z, R, offsets, y, kernel="gaussian", sigma=4000.0, lam=1e-6,
                 precision="fp64", size=29):
        self.Z, self.R, self.offsets, self.y = Z, R, offsets, np.asarray(y, dtype=np.float64)
        self.kernel, self.sigma, self.lam, self.precision, self.size = (
            kernel, sigma, lam, precision, size)
        self.n = int(len(offsets) - 1)
        check_labels(self.y, self.n)  # before the request can reach the device
        self.t_arrival = time.perf_counter()
        self.t_start: Optional[float] = None
        self.t_done: Optional[float] = None
        self.result: Optional[KRR] = None
        self.error: Optional[BaseException] = None
        self._done = threading.Event()

    def bytes_needed(self) -> int:
        return refit_bytes(self.n, cm_length(self.size), self.precision)

    def wait(self, timeout=None) -> KRR:
        if not self._done.wait(timeout):
            raise TimeoutError("refit request not finished")
        if self.error is not None:
            raise self.error
        return self.result


def refit_bytes(n: int, d: int, precision: str) -> int:
    """Device memory a fit of n molecules needs (kernel matrix dominates)."""
    ld = padded_ld(d)
    per_elem = 8 if precision == "fp64" else 12  # K64, or K64 + B32
    # mixed: B32's extra rows or the FP64 Woodbury term, and the FP64 solve's panel buffer
    extra = n * (ld * 8 + 64) + (n * (d + 1) * 12 + n * 256 * 8 if precision == "mixed" else 0)
    return per_elem * n * n + extra


# --------------------------------------------------------------------------------------------
# Policies (pure host logic, testable without a GPU)
# --------------------------------------------------------------------------------------------


def should_launch(inflight: bool, queued_requests: int, queued_molecules: int,
                  oldest_wait_s: float, min_batch: int, window_s: float) -> bool:
    """Whether a free slot takes a batch now. With nothing on the GPU: at once (no added
    latency at low load). With batches in flight: once min_batch molecules are queued or the
    oldest queued request has waited window_s, so that under load requests coalesce into
    fewer, larger batches (one read of the training set and one launch sequence per batch)."""
    if queued_requests <= 0:
        return False
    if not inflight:
        return True
    return queued_molecules >= min_batch or oldest_wait_s >= window_s


def plan_batch(queue: "collections.deque", max_molecules: int, max_atoms: int):
    """Pop requests FIFO while they fit the slot; returns the admitted list.

    A request larger than the slot on its own is admitted alone (the caller sizes slots for
    the largest request it accepts)."""
    out, mols, atoms = [], 0, 0
    while queue:
        r = queue[0]
        if out and (mols + r.nmol > max_molecules or atoms + r.natoms > max_atoms):
            break
        out.append(queue.popleft())
        mols += r.nmol
        atoms += r.natoms
    return out


def device_free_bytes(device):
    """(free, total) device memory, counting blocks torch's allocator holds but does not use."""
    import torch

    free, total = torch.cuda.mem_get_info(device)
    cached = torch.cuda.memory_reserved(device) - torch.cuda.memory_allocated(device)
    return free + cached, total


@dataclasses.dataclass
class MemoryAdmission:
    """Admit a refit when free device memory minus a reserve covers its footprint.

    mem_info() -> (free_bytes, total_bytes); torch.cuda.mem_get_info on the device, any
    callable in tests."""

    mem_info: Callable[[], tuple]
    reserve_bytes: int = 2 << 30

    def admit(self, need: int) -> bool:
        free, _ = self.mem_info()
        return need + self.reserve_bytes <= free


# --------------------------------------------------------------------------------------------
# Server
# --------------------------------------------------------------------------------------------


class _Slot:
    def __init__(self, server, index):
        import torch

        self.index = index
        dev = server.device
        m, a = server.max_molecules, server.max_atoms
        self.stream = torch.cuda.Stream(device=dev)
        self.Zh = torch.empty(a, dtype=torch.float64).pin_memory()
        self.Rh = torch.empty(a * 3, dtype=torch.float64).pin_memory()
        self.Oh = torch.empty(m + 1, dtype=torch.int64).pin_memory()
        self.Yh = torch.empty(m, dtype=torch.float64).pin_memory()
        self.Zd = torch.empty(a, dtype=torch.float64, device=dev)
        self.Rd = torch.empty(a * 3, dtype=torch.float64, device=dev)
        self.Od = torch.empty(m + 1, dtype=torch.int64, device=dev)
        self.Xd = torch.empty((m, server.model.Xd.shape[1]), dtype=torch.float64, device=dev)
        self.qn = torch.empty(m, dtype=torch.float64, device=dev)
        self.Yd = torch.empty(m, dtype=torch.float64, device=dev)
        self.status = torch.zeros(1, dtype=torch.int32, device=dev)
        ws = _lib.lib().qmlh_predict_workspace(m, server.model.n)
        self.ws = torch.empty(max(1, ws // 8), dtype=torch.float64, device=dev)
        self.event = torch.cuda.Event()
        self.requests = []
        self.busy = False


class KRRServer:
    """Serve predictions of a fitted ``KRR`` (training set and alpha resident in HBM)."""

    def __init__(self, model: KRR, size: int, max_molecules: int = 2048, max_inflight: int = 3,
                 sorting: str = "row-norm", admission: Optional[MemoryAdmission] = None,
                 min_batch: int = 128, window_s: float = 0.0002):
        import torch

        if model.alpha is None:
            raise ValueError("model must be fitted")
        self.model = model
        self.size = size
        self.sorting = 1 if sorting == "row-norm" else 0
        self.d = cm_length(size)
        if self.d != model.d:
            raise ValueError("Coulomb matrix size does not match the model's features")
        if model.kernel == "gaussian":
            # The fused predict kernel uses the expanded Gaussian; its error bound is checked
            # for queries as large as the training set's largest representation (requests of
            # larger molecules than any training molecule are outside what is validated here).
            from .kernels import check_expansion

            mx = model.max_norm()
            check_expansion(mx, mx, model.d, model.sigma, "Gaussian serving")
        self.device = model.Xd.device
        self.max_molecules = max_molecules
        self.max_atoms = max_molecules * size
        self._queue = collections.deque()
        self._queued_mols = 0
        self.min_batch = int(min_batch)
        self.window_s = float(window_s)
        self._refits = collections.deque()
        self._cv = threading.Condition()
        self._running = False
        self._threads = []
        self.admission = admission or MemoryAdmission(lambda: device_free_bytes(self.device))
        self._slots = [_Slot(self, i) for i in range(max_inflight)]
        self.batches = 0
        self.batch_sizes = []

    # -- public -----------------------------------------------------------------------------
    def start(self):
        self._running = True
        self._threads = [threading.Thread(target=self._scheduler, daemon=True),
                         threading.Thread(target=self._refit_worker, daemon=True)]
        for t in self._threads:
            t.start()
        return self

    def stop(self):
        with self._cv:
            self._running = False
            self._cv.notify_all()
        for t in self._threads:
            t.join()

    def __enter__(self):
        return self.start()

    def __exit__(self, *exc):
        self.stop()

    def submit(self, req):
        if isinstance(req, PredictRequest):
            if req.nmol > self.max_molecules:
                raise ValueError("request larger than a batch slot")
            if req.nmol and int(np.diff(req.offsets).max()) > self.size:
                raise ValueError(f"molecule with more than {self.size} atoms")
        with self._cv:
            if isinstance(req, PredictRequest):
                self._queue.append(req)
                self._queued_mols += req.nmol
            else:
                self._refits.append(req)
            self._cv.notify_all()
        return req

    # -- scheduler --------------------------------------------------------------------------
    def _launch(self, slot: _Slot, reqs):
        import torch

        L = _lib.lib()
        m = sum(r.nmol for r in reqs)
        a = sum(r.natoms for r in reqs)
        if m == 0:
            for r in reqs:
                r.result = np.empty(0)
                r.t_done = time.perf_counter()
                r._done.set()
            return False
        Zh, Rh, Oh = slot.Zh.numpy(), slot.Rh.numpy(), slot.Oh.numpy()
        pos, apos = 0, 0
        Oh[0] = 0
        for r in reqs:
            na = r.natoms
            Zh[apos:apos + na] = r.Z[r.offsets[0]:r.offsets[-1]]
            Rh[3 * apos:3 * (apos + na)] = r.R[r.offsets[0]:r.offsets[-1]].reshape(-1)
            Oh[pos + 1:pos + r.nmol + 1] = r.offsets[1:] - r.offsets[0] + apos
            pos += r.nmol
            apos += na
        s = slot.stream
        h = s.cuda_stream
        model = self.model
        with torch.cuda.stream(s):
            slot.Zd[:a].copy_(slot.Zh[:a], non_blocking=True)
            slot.Rd[:3 * a].copy_(slot.Rh[:3 * a], non_blocking=True)
            slot.Od[:m + 1].copy_(slot.Oh[:m + 1], non_blocking=True)
            ld = slot.Xd.shape[1]
            _lib.check(L.qmlh_cm_batch(slot.Zd.data_ptr(), slot.Rd.data_ptr(), slot.Od.data_ptr(),
                                       m, self.size, self.sorting, slot.Xd.data_ptr(), ld,
                                       slot.status.data_ptr(), h))
            qn = 0
            if model.kernel == "gaussian":
                _lib.check(L.qmlh_row_sqnorms(slot.Xd.data_ptr(), m, ld, self.d,
                                              slot.qn.data_ptr(), h))
                qn = slot.qn.data_ptr()
            _lib.check(L.qmlh_krr_predict(
                KERNELS[model.kernel], slot.Xd.data_ptr(), m, ld, qn, model.Xd.data_ptr(),
                model.n, model.Xd.shape[1], _lib.ptr(model.norms), model.alpha.data_ptr(),
                self.d, model.sigma, slot.Yd.data_ptr(), slot.ws.data_ptr(),
                slot.ws.numel() * 8, h))
            slot.Yh[:m].copy_(slot.Yd[:m], non_blocking=True)
            slot.event.record(s)
        slot.requests = reqs
        slot.busy = True
        self.batches += 1
        self.batch_sizes.append(m)
        return True

    def _retire(self, slot: _Slot):
        Yh = slot.Yh.numpy()
        pos = 0
        now = time.perf_counter()
        for r in slot.requests:
            r.result = Yh[pos:pos + r.nmol].copy()
            pos += r.nmol
            r.t_done = now
            r._done.set()
        slot.requests = []
        slot.busy = False

    def _scheduler(self):
        import torch

        torch.cuda.set_device(self.device)
        while True:
            with self._cv:
                if not self._running and not self._queue and not any(s.busy for s in self._slots):
                    return
            for slot in self._slots:
                if slot.busy and slot.event.query():
                    self._retire(slot)
            launched = False
            for slot in self._slots:
                if slot.busy:
                    continue
                with self._cv:
                    inflight = any(s.busy for s in self._slots)
                    wait = (time.perf_counter() - self._queue[0].t_arrival) if self._queue else 0.0
                    if not should_launch(inflight, len(self._queue), self._queued_mols, wait,
                                         self.min_batch, self.window_s):
                        break
                    reqs = plan_batch(self._queue, self.max_molecules, self.max_atoms)
                    self._queued_mols -= sum(r.nmol for r in reqs)
                if not reqs:
                    break
                try:
                    launched = self._launch(slot, reqs) or launched
                except BaseException as e:  # surface to the callers, keep serving
                    for r in reqs:
                        r.error = e
                        r._done.set()
            if not launched:
                with self._cv:
                    if not self._queue and self._running:
                        self._cv.wait(timeout=0.0002 if any(s.busy for s in self._slots) else 0.01)
                if any(s.busy for s in self._slots):
                    time.sleep(0.00005)

    # -- refits -----------------------------------------------------------------------------
    def _refit_worker(self):
        import torch

        torch.cuda.set_device(self.device)
        stream = torch.cuda.Stream(device=self.device)
        while True:
            with self._cv:
                while self._running and not self._refits:
                    self._cv.wait(timeout=0.05)
                if not self._refits:
                    return
                req = self._refits[0]
            if not self.admission.admit(req.bytes_needed()):
                time.sleep(0.01)
                with self._cv:
                    if not self._running:
                        self._refits.popleft()
                        req.error = MemoryError("server stopped before the refit was admitted")
                        req._done.set()
                continue
            with self._cv:
                self._refits.popleft()
            req.t_start = time.perf_counter()
            try:
                with torch.cuda.stream(stream):
                    m = KRR(req.kernel, req.sigma, req.lam, req.precision)
                    m.fit_molecules(req.Z, req.R, req.offsets, req.size, req.y, stream=stream)
                    stream.synchronize()
                req.result = m
            except BaseException as e:
                req.error = e
            req.t_done = time.perf_counter()
            req._done.set()


def poisson_load(server: KRRServer, pool, rate: float, duration: float, rng, max_mol=512,
                 size=29):
    """Submit requests with Poisson arrivals (rate per second) of log-uniform 1..max_mol
    molecules drawn from ``pool`` (a synthetic.MoleculeSet) for ``duration`` seconds;
    returns the submitted requests."""
    reqs = []
    t0 = time.perf_counter()
    t_next = t0
    start = 0
    while True:
        t_next += rng.exponential(1.0 / rate)
        if t_next - t0 > duration:
            break
        k = int(np.clip(np.exp(rng.uniform(0, np.log(max_mol))), 1, max_mol))
        if start + k > pool.nmol:
            start = 0
        sub = pool.subset(start, start + k)
        start += k
        Z, R, off = sub.ragged()
        delay = t_next - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        reqs.append(server.submit(PredictRequest(Z, R, off)))
    return reqs


def run_load(server: KRRServer, pool, rate: float, duration: float, rng, max_mol: int = 512,
             size: int = 29, refit_pool=None, refit_every: float = 0.0,
             refit_n=(1000, 16000), refit_precision: str = "fp64"):
    """Poisson prediction load (``poisson_load``) plus, when ``refit_every`` > 0, refit
    requests of N ~ U[refit_n] molecules from ``refit_pool`` submitted every ``refit_every``
    seconds while the load runs (admitted by the server's memory budget). Waits for every
    request; returns (predict requests, refit requests, summary dict)."""
    refits = []
    stop = threading.Event()

    def submit_refits():
        rr = np.random.default_rng(int(rng.integers(1 << 31)))
        while not stop.wait(refit_every):
            n = int(rr.integers(refit_n[0], refit_n[1] + 1))
            sub = refit_pool.subset(0, min(n, refit_pool.nmol))
            refits.append(server.submit(RefitRequest(
                *sub.ragged(), sub.labels, kernel=server.model.kernel, sigma=server.model.sigma,
                lam=server.model.lam, precision=refit_precision, size=size)))

    th = None
    if refit_every > 0 and refit_pool is not None:
        th = threading.Thread(target=submit_refits, daemon=True)
        th.start()
    b0 = server.batches
    reqs = poisson_load(server, pool, rate, duration, rng, max_mol=max_mol, size=size)
    stop.set()
    if th is not None:
        th.join()
    for r in reqs:
        r.wait(600)
    fitted = [f for f in refits if f.wait(600) is not None]
    lat = np.array([r.latency for r in reqs]) * 1e3 if reqs else np.full(1, np.nan)
    span = (max(r.t_done for r in reqs) - min(r.t_arrival for r in reqs)) if reqs else float("nan")
    nb = server.batches - b0
    summary = {
        "rate_req_per_s": rate, "requests": len(reqs), "batches": nb,
        "molecules_per_batch": float(sum(r.nmol for r in reqs) / max(nb, 1)),
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "molecules_per_s": float(sum(r.nmol for r in reqs) / span) if reqs else 0.0,
        "refits": [{"n": f.n, "admitted_after_ms": (f.t_start - f.t_arrival) * 1e3,
                    "fit_ms": (f.t_done - f.t_start) * 1e3} for f in refits],
        "refit_precision": refit_precision if refits else None,
        "predictions_during_refits": int(sum(
            1 for r in reqs if any(f.t_start is not None and f.t_start < r.t_done < f.t_done
                                   for f in refits))),
    }
    del fitted
    return reqs, refits, summary
