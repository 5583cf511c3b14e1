#!/usr/bin/env python3
"""KRR fit + predict benchmark on one MI355X: the serving workload at N = 50,000.

Workload: 50,000 seeded synthetic training molecules (up to 29 atoms, Coulomb matrix size 29,
435 FP64 features), Gaussian kernel sigma = 4000, lambda = 1e-6. One step is one full fit with
the raw molecules (Z, R) already resident in HBM, timed for two solvers:

  fp64  -- Coulomb matrices on the GPU, squared norms, the symmetric 50k x 50k Gaussian kernel
           (FP64 MFMA), K + lambda I, blocked FP64 Cholesky, two triangular solves;
  mixed -- the same kernel matrix in FP64 for residuals plus its deflated remainder in FP32,
           a low-precision Cholesky of the remainder, a Woodbury correction for the low-rank
           part, FP64 preconditioned CG to |y - K alpha| <= 1e-10 |y|.

The mixed fit's error is its prediction difference from the FP64 fit on 2048 held-out
molecules, over the label standard deviation. Then prediction requests of 1-512 molecules
(log-uniform) are served one at a time (host molecules in, predictions out), and a Poisson
request stream runs through the continuous-batching server for each fitted model; latency
percentiles are reported for both.

Prints one JSON line. `value` is the mixed fit time in seconds (lower is better).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X published dense MFMA peaks, TFLOP/s.
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "fp16": 2500.0}
SYRK_KERNEL = {
    "fp64": ("gemm_update_kernel<double, lower> (Cholesky trailing update, FP64 MFMA)", "fp64"),
    "fp32": ("gemm_update_kernel<float, lower> (FP32 MFMA trailing update)", "fp32"),
    "f16x3": ("syrk_f16x3_wide_kernel<false> (trailing update of two block columns, 1536 features, "
              "on 256 x 256 tiles as 3 FP16 MFMA products of the scaled split panel, FP32 "
              "accumulate)", "fp16"),
}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_krr_serve_50k.json")


def train_relres(model, y) -> float:
    """|y - (K + lambda I) alpha| / |y| of a fitted model, K alpha from the prediction kernel."""
    import numpy as np
    import torch

    ka = model.predict_device(model.Xd, model.n)
    yd = torch.as_tensor(np.asarray(y), dtype=torch.float64).to(ka.device)
    r = yd - ka - model.lam * model.alpha
    return float(r.norm() / yd.norm())


def kernel_entry_error(Xd, d, sigma, lam, fp64_cols=1, rows=2048):
    """Kernel-entry error of the mixed fit's two representations on the first `rows` training
    molecules, against the direct-difference FP64 Gaussian (the form of qml/fkernels.f90:340-372,
    pinned to the oracle by the tests): K64 (FP64 MFMA cross term, used for the residuals) and
    B32 + V V^T (the FP32 deflated matrix and its rank-(d+1) term as the preconditioner holds it:
    the first fp64_cols columns of V in FP64, the rest rounded to FP32)."""
    import torch

    from qml_amd import _lib
    from qml_amd.kernels import pairwise_device, row_sqnorms_device

    L = _lib.lib()
    n = min(rows, Xd.shape[0])
    X = Xd[:n].contiguous()
    dev = X.device
    norms = row_sqnorms_device(X, n, d)
    K = torch.zeros((n, n), dtype=torch.float64, device=dev)
    B = torch.zeros((n, n), dtype=torch.float32, device=dev)
    V = torch.empty((d + 1, n), dtype=torch.float32, device=dev)
    u = torch.empty(n, dtype=torch.float64, device=dev)
    _lib.check(L.qmlh_gaussian_deflated(_lib.ptr(X), n, X.shape[1], _lib.ptr(norms), d, sigma,
                                        lam, 0.0, _lib.ptr(K), n, _lib.ptr(B), n, _lib.ptr(V), n,
                                        _lib.ptr(u), _lib.stream_handle()))
    Kd = pairwise_device(_lib.PAIR_GAUSSIAN, X, n, X, n, d, sigma, True)
    V64 = torch.empty((d + 1, n), dtype=torch.float64, device=dev)
    _lib.check(L.qmlh_gaussian_lowrank(_lib.ptr(X), n, X.shape[1], d, sigma, _lib.ptr(u),
                                       _lib.ptr(V64), n, _lib.stream_handle()))
    # column-major lower triangles: device tensors are (n, n) row-major views of them
    # .T of each device tensor is the matrix indexed (row, col); compare the lower triangles
    low = torch.ones((n, n), dtype=torch.bool, device=dev).tril()
    Kt = Kd.T + lam * torch.eye(n, dtype=torch.float64, device=dev)
    k64 = ((K.T - Kt).abs() / Kt.abs())[low].max()
    Vd = V.double()
    Vd[:fp64_cols] = V64[:fp64_cols]
    rep = B.double().T + Vd.T @ Vd  # B32 + V V^T
    dfl = (rep - Kt)[low].abs().max()
    return {"rows": n, "reference": "direct-difference FP64 Gaussian on the GPU (pinned to the oracle)",
            "k64_max_rel": float(k64), "fp32_deflated_max_abs": float(dfl),
            "fp32_deflated_max_abs_over_lambda": float(dfl) / lam}


def _short(x: float) -> str:
    """1e-06 -> 1e-6: the exponent without its leading zero."""
    return (f"{x:g}").replace("e-0", "e-").replace("e+0", "e+")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--n-train", type=int, default=50000)
    p.add_argument("--kernel", choices=["gaussian", "laplacian"], default="gaussian",
                   help="kernel of the fits (the headline workload is Gaussian)")
    p.add_argument("--factor", choices=["auto", "fp32", "f16x3"], default="auto",
                   help="mixed fit's low-precision factor (auto: the model's default, "
                        "f16x3)")
    p.add_argument("--sigma", type=float, default=4000.0)
    p.add_argument("--lam", type=float, default=1e-6)
    p.add_argument("--requests", type=int, default=300)
    p.add_argument("--cpu-sample", type=int, default=10000,
                   help="training molecules in the CPU-baseline sample (0 = skip)")
    p.add_argument("--precision", choices=["fp64", "mixed", "both"], default="both")
    # Refinement stops at |y - K alpha| <= tol |y|. 1e-9 keeps the prediction difference to the
    # FP64 fit near 1e-8 of std(y) (measured below), far inside the 1e-6 target.
    p.add_argument("--tol", type=float, default=1e-9)
    p.add_argument("--serve-rates", default="200,500,1000,2000",
                   help="Poisson arrival rates (requests/s) of the serving-load phase")
    p.add_argument("--serve-seconds", type=float, default=2.0,
                   help="seconds of load per rate and model (0 = skip)")
    p.add_argument("--refit-every", type=float, default=0.5,
                   help="seconds between refit requests (N ~ U[1000, 16000]) under load "
                        "(0 = none)")
    p.add_argument("--lowrank-cols", type=int, default=1,
                   help="leading columns of the Woodbury term solved in FP64 (hybrid form)")
    p.add_argument("--copy-mib", type=int, default=0,
                   help="run a device copy of this many MiB first (HBM counter calibration)")
    p.add_argument("--kernels", type=int, default=1,
                   help="time the kernel-matrix builds at the large shapes (0 = skip)")
    p.add_argument("--local", type=int, default=20000,
                   help="molecules of the local-kernel section (0: skip)")
    p.add_argument("--local-only", action="store_true",
                   help="run only the sections of the rows beyond the fit (local kernels, QM7 "
                        "from XYZ files, cho_invert) and print their JSON (profiling)")
    p.add_argument("--local-cpu-sample", type=int, default=2000,
                   help="molecules of the local-kernel CPU baseline sample (0: skip)")
    p.add_argument("--qm7", type=int, default=1,
                   help="run the QM7 end-to-end test from XYZ files (0 = skip)")
    p.add_argument("--invert", type=int, default=20000,
                   help="order of the cho_invert measurement (0 = skip)")
    p.add_argument("--sweep", type=int, default=64,
                   help="problems in the batched learning-curve sweep (0 = skip)")
    return p.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(n_full, size, sigma, lam, n_sample):
    """Reference path on the host cores over a bounded sample, extrapolated to n_full.

    Coulomb matrices one molecule at a time (as the reference's callers loop), the training
    kernel with the reference Fortran (oracle/_ref) where it was built, else the C
    restatement, and cho_solve as LAPACK dpotrf + dpotrs. Times scale as n, n^2 and n^3.
    """
    import numpy as np

    import oracle
    from qml_amd import synthetic

    cores = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    tr, _, _ = synthetic.workload("krr_serve_50k", n_train=n_sample, n_test=0)
    use_ref = oracle.ref_available()
    t0 = time.perf_counter()
    if use_ref:
        X = np.array([oracle.ref_coulomb_matrix(*tr.molecule(i), size) for i in range(tr.nmol)])
    else:
        X = np.array([oracle.coulomb_matrix(*tr.molecule(i), size) for i in range(tr.nmol)])
    t1 = time.perf_counter()
    K = oracle.ref_kernel("gaussian", X, X, sigma) if use_ref else oracle.gaussian_kernel(X, X, sigma)
    t2 = time.perf_counter()
    K[np.diag_indices_from(K)] += lam
    oracle.cho_solve(K, tr.labels.copy())
    t3 = time.perf_counter()
    r = n_full / n_sample
    est = (t1 - t0) * r + (t2 - t1) * r * r + (t3 - t2) * r ** 3
    return {
        "value": est, "unit": "s", "cores": cores,
        "kind": "reference" if use_ref else "port",
        "sample": (f"fit at N={n_sample} on {cores} host threads: CM {t1 - t0:.2f}s, "
                   f"Gaussian kernel {t2 - t1:.2f}s, cho_solve {t3 - t2:.2f}s; extrapolated "
                   f"to N={n_full} as N, N^2, N^3"),
    }


def roofline(timers, steps, factor):
    """Roofline of the Cholesky trailing update (the dominant kernel of a fit): algorithmic
    flops m (m + 1) kb per launch over the launches' HIP-event time, against the published
    MFMA peak of the arithmetic the kernel runs. traffic: HBM bytes per launch from the
    committed counter summary, when one exists for this kernel."""
    name, prec = SYRK_KERNEL[factor]
    ms, launches, flops = timers["syrk"]
    achieved = flops / (ms * 1e-3) / 1e12 if ms > 0 else 0.0
    traffic = None
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f).get(factor)
        if t:
            traffic = t["hbm_bytes_per_launch"]
    except (OSError, ValueError):
        pass
    return {"kernel": name, "bound": "mfma", "achieved": achieved, "peak": PEAK_TFLOPS[prec],
            "unit": "TFLOP/s", "frac": achieved / PEAK_TFLOPS[prec], "traffic": traffic,
            "algorithmic_flops_per_launch": flops / max(launches, 1),
            "launches_per_step": launches / steps, "ms_per_step": ms / steps}


def run_kernels(log):
    """Kernel-matrix builds at the large-configuration shapes: the Laplacian (FP64 VALU) on a
    row block of 8192 queries against 100,000 training rows (size 29, d = 435, sigma = 724),
    the symmetric Gaussian (FP64 MFMA expanded form) and Laplacian (sigma = 724) at N = 20,000
    (sigma = 4000), the linear
    kernel on the same tiles, and the Sargan / Matern (L2) kernels and p-distances (p = 3 by
    products, p = 2.5 by pow on a 1024-row block) at the Laplacian's shape."""
    import torch

    from qml_amd import _lib, synthetic
    from qml_amd.kernels import gaussian_mfma_device, pairwise_device, row_sqnorms_device
    from qml_amd.representations import cm_length, generate_coulomb_matrices

    out = {}
    big, _, size = synthetic.workload("krr_large_100k", n_train=100000, n_test=0)
    d = cm_length(size)
    X = generate_coulomb_matrices(*big.ragged(), size=size)
    rows = 8192
    K = torch.empty((100000, rows), dtype=torch.float64, device=X.device)  # (rows, 100k) col-major
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    pairwise_device(_lib.PAIR_LAPLACIAN, X[:rows], rows, X, 100000, d, 724.0, False, K)
    reps = 3
    e0.record()
    for _ in range(reps):
        pairwise_device(_lib.PAIR_LAPLACIAN, X[:rows], rows, X, 100000, d, 724.0, False, K)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    pf = rows * 100000.0 * d
    out["laplacian_rowblock"] = {
        "shape": f"{rows} x 100000, d={d}", "ms": ms, "pair_features_per_s": pf / (ms * 1e-3),
        "frac_of_fp64_valu": pf / (ms * 1e-3) / 19.65e12,
        "hbm_write_GBps": rows * 100000.0 * 8 / (ms * 1e-3) / 1e9}
    del K
    n = 20000
    X2 = X[:n]
    nr = row_sqnorms_device(X2, n, d)
    K2 = torch.empty((n, n), dtype=torch.float64, device=X.device)
    gaussian_mfma_device(X2, nr, n, X2, nr, n, d, 4000.0, True, K2)
    e0.record()
    for _ in range(reps):
        gaussian_mfma_device(X2, nr, n, X2, nr, n, d, 4000.0, True, K2)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    flops = n * (n + 1) / 2.0 * d * 2.0  # symmetric: the upper tiles' cross terms
    out["gaussian_mfma_20k"] = {"shape": f"{n} x {n} symmetric, d={d}", "ms": ms,
                                "tflops": flops / (ms * 1e-3) / 1e12,
                                "frac_of_fp64_mfma": flops / (ms * 1e-3) / 1e12 / PEAK_TFLOPS["fp64"]}
    # Laplacian, symmetric 20k (the training-kernel shape of the drop-in and of the fits):
    # upper tiles computed, mirrored through LDS.
    pairwise_device(_lib.PAIR_LAPLACIAN, X2, n, X2, n, d, 724.0, True, K2)
    e0.record()
    for _ in range(reps):
        pairwise_device(_lib.PAIR_LAPLACIAN, X2, n, X2, n, d, 724.0, True, K2)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    pf = n * (n + 1) / 2.0 * d
    out["laplacian_symmetric_20k"] = {
        "shape": f"{n} x {n} symmetric, d={d}", "ms": ms, "pair_features_per_s": pf / (ms * 1e-3),
        "frac_of_fp64_valu": pf / (ms * 1e-3) / 19.65e12}
    # Linear kernel on the same cross-term tiles (qmlh_linear_mfma), symmetric 20k.
    from qml_amd.distance import p_distance  # noqa: F401  (drop-in; device form used below)
    from qml_amd.kernels import pairwise_ext_device

    lin = torch.empty((n, n), dtype=torch.float64, device=X.device)
    L = _lib.lib()

    def linear():
        _lib.check(L.qmlh_linear_mfma(_lib.ptr(X2), n, X2.shape[1], _lib.ptr(X2), n, X2.shape[1],
                                      d, _lib.ptr(lin), n, _lib.SYMMETRIC,
                                      _lib.stream_handle(None)))

    linear()
    e0.record()
    for _ in range(reps):
        linear()
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    out["linear_mfma_20k"] = {"shape": f"{n} x {n} symmetric, d={d}", "ms": ms,
                              "tflops": flops / (ms * 1e-3) / 1e12,
                              "frac_of_fp64_mfma": flops / (ms * 1e-3) / 1e12 / PEAK_TFLOPS["fp64"]}
    del lin, K2
    # Fused Laplacian prediction (qmlh_krr_predict): 4096 query molecules against 100,000
    # training rows and alpha, no kernel matrix written.
    nq = 4096
    alpha = torch.ones(100000, dtype=torch.float64, device=X.device)
    yq = torch.empty(nq, dtype=torch.float64, device=X.device)
    pws = torch.empty(max(1, L.qmlh_predict_workspace(nq, 100000) // 8), dtype=torch.float64,
                      device=X.device)

    def predict_laplace():
        _lib.check(L.qmlh_krr_predict(_lib.PAIR_LAPLACIAN, _lib.ptr(X), nq, X.shape[1], None,
                                      _lib.ptr(X), 100000, X.shape[1], None, _lib.ptr(alpha), d,
                                      724.0, _lib.ptr(yq), _lib.ptr(pws),
                                      pws.numel() * 8, _lib.stream_handle(None)))

    predict_laplace()
    e0.record()
    for _ in range(reps):
        predict_laplace()
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    pf = nq * 100000.0 * d
    out["predict_laplace"] = {"shape": f"{nq} queries x 100000 training, d={d}", "ms": ms,
                              "pair_features_per_s": pf / (ms * 1e-3),
                              "frac_of_fp64_valu": pf / (ms * 1e-3) / 19.65e12}
    del alpha, yq, pws
    # Parameterised VALU kernels on a row block against the 100k set (the Laplacian's shape).
    K = torch.empty((100000, rows), dtype=torch.float64, device=X.device)
    for name, kind, params, nrows in (
            ("sargan_2gammas", _lib.PAIR_SARGAN, [724.0, 0.5, 0.25], rows),
            ("matern_l2_order2", _lib.PAIR_MATERN_L2, [724.0, 2.0], rows),
            ("p_distance_3", _lib.PAIR_P_INT, [3.0], rows),
            ("p_distance_2.5", _lib.PAIR_P_REAL, [2.5], 1024)):
        Kv = K.view(-1)[:100000 * nrows].view(100000, nrows)  # (nrows, 100000) column-major
        pairwise_ext_device(kind, X[:nrows], nrows, X, 100000, d, params, False, Kv)
        e0.record()
        pairwise_ext_device(kind, X[:nrows], nrows, X, 100000, d, params, False, Kv)
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1)
        pf = nrows * 100000.0 * d
        out[name] = {"shape": f"{nrows} x 100000, d={d}", "ms": ms,
                     "pair_features_per_s": pf / (ms * 1e-3)}
    log(f"kernels: laplacian {out['laplacian_rowblock']['frac_of_fp64_valu']:.2f} of FP64 VALU, "
        f"gaussian {out['gaussian_mfma_20k']['tflops']:.1f} TF/s, linear "
        f"{out['linear_mfma_20k']['tflops']:.1f} TF/s")
    return out


def run_local(n_mol, cpu_sample, log):
    """Local (atom-summed) kernels at scale: n_mol synthetic molecules (5-23 atoms, the QM7-like
    generator), atomic Coulomb matrices (row-norm, size 23, d = 276) generated on the GPU, then
    the symmetric n_mol x n_mol local Gaussian (sigma = 724, FP64 MFMA expanded distances) and
    Laplacian (sigma = 10^3.6, FP64 VALU) kernels. Algorithmic work counts each atom pair of
    the symmetric problem once: pairs = atoms (atoms + 1) / 2, times d features (2 flops per
    pair-feature for the Gaussian cross term). The CPU baseline runs the reference Fortran
    (oracle/_ref, else the C restatement) on a bounded sample of cpu_sample molecules; the
    reference computes every molecule pair, so its time scales as n_mol^2."""
    import numpy as np
    import torch

    from qml_amd import _lib, synthetic
    from qml_amd.kernels import LocalSet, local_kernels_device
    from qml_amd.representations import generate_atomic_coulomb_matrices

    mols = synthetic.generate(n_mol, 2024, 5, 23)
    Z, R, off = mols.ragged()
    atoms = int(off[-1])
    d = 276
    out = {"workload": f"local_krr_{n_mol // 1000}k: {n_mol} molecules ({atoms} atoms, 5-23 per "
                       "molecule), atomic Coulomb matrices row-norm size 23 (d=276), symmetric "
                       "local kernels, Gaussian sigma=724 and Laplacian sigma=10^3.6"}
    generate_atomic_coulomb_matrices(Z, R, off, 23, "row-norm")
    torch.cuda.synchronize()
    _lib.timer_enable(True)
    X = generate_atomic_coulomb_matrices(Z, R, off, 23, "row-norm")
    torch.cuda.synchronize()
    ms, n, _ = _lib.timer_read("cm")
    _lib.timer_enable(False)
    out["atomic_cm"] = {"rows": atoms, "ms": ms / n, "rows_per_s": atoms / (ms / n * 1e-3)}
    S = LocalSet(X, np.diff(off))
    K = torch.empty((1, n_mol, n_mol), dtype=torch.float64, device=X.device)
    pairs = atoms * (atoms + 1) / 2.0
    for kind, name, sigma in ((_lib.PAIR_GAUSSIAN, "gaussian", 724.0),
                              (_lib.PAIR_LAPLACIAN, "laplacian", 10 ** 3.6)):
        local_kernels_device(kind, S, S, [sigma], symmetric=True, out=K)
        torch.cuda.synchronize()
        _lib.timer_enable(True)
        reps = 2
        for _ in range(reps):
            local_kernels_device(kind, S, S, [sigma], symmetric=True, out=K)
        torch.cuda.synchronize()
        ms, n, work = _lib.timer_read("local")
        _lib.timer_enable(False)
        ms /= n
        pf = pairs * d / (ms * 1e-3)
        if kind == _lib.PAIR_GAUSSIAN:
            tf = 2.0 * pf / 1e12
            out[name] = {"ms": ms, "tflops": tf, "bound": "mfma", "peak_tflops": PEAK_TFLOPS["fp64"],
                         "frac": tf / PEAK_TFLOPS["fp64"],
                         "kernel": "local_kernels_kernel<LK_GAUSS_MFMA>"}
        else:
            out[name] = {"ms": ms, "pair_features_per_s": pf, "bound": "valu",
                         "frac": pf / 19.65e12, "kernel": "local_kernels_kernel<LK_LAPLACE>"}
        log(f"local {name}: {ms:.1f} ms, {out[name]['frac']:.2f} of peak")
    out["panels"] = S.npanels
    out["tile_fill"] = atoms / (S.npanels * 128.0)
    del K, S, X
    torch.cuda.empty_cache()
    if cpu_sample > 0:
        import oracle

        use_ref = oracle.ref_available()
        cores = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
        sub = mols.subset(0, cpu_sample)
        Zs, Rs, offs = sub.ragged()
        acm = oracle.ref_atomic_coulomb_matrix if use_ref else oracle.atomic_coulomb_matrix
        t0 = time.perf_counter()
        Xs = np.concatenate([acm(*sub.molecule(m), 23, "row-norm") for m in range(sub.nmol)])
        t1 = time.perf_counter()
        lk = oracle.ref_local_kernels if use_ref else oracle.local_kernels
        ns = np.diff(offs)
        lk("gaussian", Xs, Xs, ns, ns, [724.0])
        t2 = time.perf_counter()
        lk("laplacian", Xs, Xs, ns, ns, [10 ** 3.6])
        t3 = time.perf_counter()
        r = n_mol / cpu_sample
        out["cpu_baseline"] = {
            "value": (t1 - t0) * r + (t2 - t1) * r * r, "unit": "s", "cores": cores,
            "kind": "reference" if use_ref else "port",
            "laplacian_s": (t3 - t2) * r * r,
            "sample": (f"{cpu_sample} molecules on {cores} host threads: atomic CM "
                       f"{t1 - t0:.2f}s (one molecule per call), Gaussian local kernel "
                       f"{t2 - t1:.2f}s, Laplacian {t3 - t2:.2f}s (all molecule pairs, as the "
                       f"reference computes them); extrapolated to {n_mol} molecules as n, n^2")}
        log(f"local cpu baseline: gaussian {out['cpu_baseline']['value']:.0f}s extrapolated")
    return out


def run_qm7(log):
    """The reference's QM7 test on its own file format (tests/test_energy_krr_cmat.py:54-113):
    1000 XYZ files (committed as text in tests/golden/xyz_texts.npz, written to a temporary
    directory) read into Compound objects, sorted Coulomb matrices (size 23), Laplacian KRR with
    sigma = 10^4.2, lambda = 1e-10 on 700 training / 300 test molecules after the seed-666
    shuffle. Timed end to end (files on disk in, predictions on the host out) on the GPU path,
    and the same steps on the reference path (its Fortran from oracle/_ref, LAPACK via scipy)."""
    import tempfile

    import numpy as np

    import oracle
    from qml_amd.compound import Compound
    from qml_amd.kernels import laplacian_kernel
    from qml_amd.math import cho_solve

    z = np.load(os.path.join(ROOT, "tests", "golden", "xyz_texts.npz"), allow_pickle=False)
    hof = {}
    for line in bytes(z["hof_qm7"]).decode().splitlines():
        tok = line.split()
        if tok:
            hof[tok[0]] = float(tok[1])
    out = {"workload": "qm7_krr_cmat: 1000 QM7 XYZ files, sorted CM size 23, Laplacian "
                       "sigma=10^4.2, lambda=1e-10, 700 train / 300 test"}
    with tempfile.TemporaryDirectory() as tmp:
        paths = {}
        for name, text in zip(z["names"], z["texts"]):
            paths[str(name)] = os.path.join(tmp, str(name))
            with open(paths[str(name)], "wb") as f:
                f.write(bytes(text))
        names = sorted(hof)[:1000]

        def path_gpu():
            mols = []
            for n in names:
                m = Compound(xyz=paths[n])
                m.properties = hof[n]
                m.generate_coulomb_matrix(size=23, sorting="row-norm")
                mols.append(m)
            np.random.seed(666)
            np.random.shuffle(mols)
            X = np.array([m.representation for m in mols[:700]])
            Xs = np.array([m.representation for m in mols[-300:]])
            Y = np.array([m.properties for m in mols[:700]])
            Ys = np.array([m.properties for m in mols[-300:]])
            K = laplacian_kernel(X, X, 10 ** 4.2)
            K[np.diag_indices_from(K)] += 1e-10
            alpha = cho_solve(K, Y)
            return np.mean(np.abs(Ys - laplacian_kernel(X, Xs, 10 ** 4.2).T.dot(alpha)))

        def path_ref():
            use_ref = oracle.ref_available()
            cm = oracle.ref_coulomb_matrix if use_ref else oracle.coulomb_matrix
            mols = []
            for n in names:
                m = Compound(xyz=paths[n])
                m.properties = hof[n]
                m.representation = cm(m.nuclear_charges, m.coordinates, 23, "row-norm")
                mols.append(m)
            np.random.seed(666)
            np.random.shuffle(mols)
            X = np.array([m.representation for m in mols[:700]])
            Xs = np.array([m.representation for m in mols[-300:]])
            Y = np.array([m.properties for m in mols[:700]])
            Ys = np.array([m.properties for m in mols[-300:]])
            kern = (lambda A, B: oracle.ref_kernel("laplacian", A, B, 10 ** 4.2)) if use_ref \
                else (lambda A, B: oracle.laplacian_kernel(A, B, 10 ** 4.2))
            K = kern(X, X)
            K[np.diag_indices_from(K)] += 1e-10
            alpha = oracle.cho_solve(K, Y)
            return np.mean(np.abs(Ys - kern(X, Xs).T.dot(alpha))), use_ref

        def path_batched():
            # The framework's own path: one reader pass per split, one CM launch per split,
            # the fit and the predictions on the device.
            from qml_amd.compound import read_xyz_batch
            from qml_amd.krr import KRR

            order = list(range(1000))
            np.random.seed(666)
            np.random.shuffle(order)
            tr = [names[i] for i in order[:700]]
            te = [names[i] for i in order[-300:]]
            Z, R, off = read_xyz_batch([paths[n] for n in tr])
            Zq, Rq, offq = read_xyz_batch([paths[n] for n in te])
            model = KRR("laplacian", 10 ** 4.2, 1e-10, "fp64")
            model.fit_molecules(Z, R, off, 23, np.array([hof[n] for n in tr]))
            pred = model.predict_molecules(Zq, Rq, offq, 23)
            return np.mean(np.abs(np.array([hof[n] for n in te]) - pred))

        path_gpu()  # warm-up (library init, first launches)
        path_batched()
        t0 = time.perf_counter()
        mae = path_gpu()
        t1 = time.perf_counter()
        mae_ref, use_ref = path_ref()
        t2 = time.perf_counter()
        mae_b = path_batched()
        t3 = time.perf_counter()
    cores = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    out.update({"gpu_s": t1 - t0, "mae": float(mae), "gpu_batched_s": t3 - t2,
                "batched_mae": float(mae_b), "reference_s": t2 - t1,
                "reference_mae": float(mae_ref), "reference_kind": "reference" if use_ref else "port",
                "reference_cores": cores,
                "note": "gpu_s and reference_s: per-molecule Compound calls as in the reference "
                        "test; gpu_batched_s: read_xyz_batch + one CM launch per split + "
                        "KRR fit/predict on the device; file reads included everywhere"})
    log(f"qm7 end to end: GPU {t1 - t0:.3f}s (MAE {mae:.3f}), batched {t3 - t2:.3f}s, "
        f"reference {t2 - t1:.3f}s")
    return out


def run_invert(n, cpu_n, log):
    """cho_invert at n (device-resident, qmlh_cho_invert: FP64 Cholesky, inverse of the factor
    from the identity, MFMA rank-n product) on a synthetic SPD Gaussian kernel matrix plus
    lambda = 1e-6. Algorithmic work is LAPACK's dpotrf + dpotri: n^3 flops. CPU baseline:
    scipy's LAPACK dpotrf + dpotri (what the reference's fcho_invert calls) at cpu_n,
    extrapolated as n^3."""
    import numpy as np
    import torch

    from qml_amd import _lib, synthetic
    from qml_amd.kernels import gaussian_mfma_device, row_sqnorms_device
    from qml_amd.math import add_diagonal_device, cho_invert_device
    from qml_amd.representations import cm_length, generate_coulomb_matrices

    tr, _, size = synthetic.workload("krr_serve_50k", n_train=n, n_test=0)
    d = cm_length(size)
    X = generate_coulomb_matrices(*tr.ragged(), size=size)
    nr = row_sqnorms_device(X, n, d)
    K0 = gaussian_mfma_device(X, nr, n, X, nr, n, d, 4000.0, True)
    add_diagonal_device(K0, n, 1e-6)
    K = torch.empty_like(K0)
    K.copy_(K0)
    cho_invert_device(K, n)
    torch.cuda.synchronize()
    reps, ms = 2, 0.0
    for _ in range(reps):
        K.copy_(K0)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        info = cho_invert_device(K, n)
        e1.record()
        e1.synchronize()
        ms += e0.elapsed_time(e1) / reps
    # residual of the inverse on a few columns: |K0 Ki e_j - e_j|
    cols = torch.arange(0, n, max(1, n // 8), device=K.device)
    # K is symmetric, so its rows `cols` are the columns; (n, n) row-major = column-major.
    R = torch.matmul(K0, K[cols].T.contiguous())
    E = torch.zeros_like(R)
    E[cols, torch.arange(cols.numel(), device=K.device)] = 1.0
    resid = float((R - E).abs().max())
    flops = float(n) ** 3
    out = {"n": n, "ms": ms, "tflops_lapack_equiv": flops / (ms * 1e-3) / 1e12,
           "frac": flops / (ms * 1e-3) / 1e12 / PEAK_TFLOPS["fp64"], "info": int(info.item()),
           "max_abs_residual_sampled_columns": resid,
           "workload": f"cho_invert of a {n} x {n} Gaussian kernel matrix (size 29 Coulomb "
                       "matrices, sigma=4000) + 1e-6 I, FP64"}
    del K, K0, X
    torch.cuda.empty_cache()
    if cpu_n > 0:
        from scipy.linalg import lapack

        A = np.asfortranarray(np.eye(cpu_n) * 2.0 + np.random.default_rng(0).random((cpu_n, cpu_n)) * 1e-3)
        A = np.asfortranarray(A + A.T)
        t0 = time.perf_counter()
        c, _ = lapack.dpotrf(A, lower=1, overwrite_a=1)
        lapack.dpotri(c, lower=1, overwrite_c=1)
        t1 = time.perf_counter()
        cores = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
        out["cpu_baseline"] = {"value": (t1 - t0) * (n / cpu_n) ** 3, "unit": "s", "cores": cores,
                               "kind": "port",
                               "sample": f"LAPACK dpotrf + dpotri (scipy) at n={cpu_n}: "
                                         f"{t1 - t0:.2f}s on {cores} threads; extrapolated as n^3"}
    log(f"cho_invert n={n}: {ms:.1f} ms ({out['tflops_lapack_equiv']:.1f} TF/s LAPACK-equivalent)")
    # Bunch-Kaufman (bkf_solve / bkf_invert) on symmetric indefinite matrices: the factorisation
    # at orders 2000, 4000 and 8000 (64-column panels and an MFMA trailing update from order
    # 1024) against scipy's dsytrf at the same n (8000: extrapolated from 4000 as n^3); the
    # solve and the inverse's n solves at 2000.
    from scipy.linalg import lapack

    from qml_amd.math import bkf_factor_device, bkf_solve_device

    cores = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    bk = {"factor": [], "cpu_lapack_cores": cores}
    rng = np.random.default_rng(1)
    cpu_ms = {}
    for nb in (2000, 4000, 8000):
        M = rng.standard_normal((nb, nb))
        Ah = M + M.T
        del M
        A0 = torch.from_numpy(np.ascontiguousarray(Ah.T)).cuda()
        Ab = A0.clone()
        bkf_factor_device(Ab, nb)  # warm
        torch.cuda.synchronize()
        Ab.copy_(A0)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        ipiv, info = bkf_factor_device(Ab, nb)
        e1.record()
        e1.synchronize()
        fms = e0.elapsed_time(e1)
        if nb <= 4000:
            t0 = time.perf_counter()
            lu, piv, _ = lapack.dsytrf(np.asfortranarray(Ah), lower=1)[:3]
            cpu_ms[nb] = (time.perf_counter() - t0) * 1e3
            same = bool(np.array_equal(ipiv.cpu().numpy()[:nb], piv))
        else:
            cpu_ms[nb] = cpu_ms[4000] * (nb / 4000) ** 3
            same = None
        bk["factor"].append({"n": nb, "factor_ms": fms,
                             "factor_gflops": nb ** 3 / 3.0 / (fms * 1e-3) / 1e9,
                             "info": int(info.item()),
                             "cpu_lapack_dsytrf_ms": cpu_ms[nb],
                             "cpu_extrapolated": nb > 4000,
                             "pivots_equal_lapack": same})
        log(f"bunch-kaufman n={nb}: factor {fms:.1f} ms (LAPACK {cpu_ms[nb]:.1f} ms)")
        if nb == 2000:
            yb = torch.from_numpy(rng.standard_normal(nb)).cuda()
            Xb = torch.eye(nb, dtype=torch.float64, device=Ab.device)
            e2, e3, e4 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
            e2.record()
            bkf_solve_device(Ab, nb, ipiv, yb, 1)
            e3.record()
            bkf_solve_device(Ab, nb, ipiv, Xb, nb)
            e4.record()
            e4.synchronize()
            t1 = time.perf_counter()
            lapack.dsytrs(lu, piv, np.eye(nb), lower=1)
            bk["solve_2000"] = {"solve_ms": e2.elapsed_time(e3),
                                "inverse_solves_ms": e3.elapsed_time(e4),
                                "cpu_lapack_dsytrs_identity_ms":
                                    (time.perf_counter() - t1) * 1e3}
            del Xb
        del Ab, A0, Ah
    out["bunch_kaufman"] = bk
    return out


def run_sweep(count, log):
    """64 Laplacian KRR fits, n log-uniform in [500, 8000], (sigma, lambda) grid, size-23
    Coulomb matrices: one graph launch vs the same fits solved one after another."""
    import numpy as np
    import torch

    from qml_amd import synthetic
    from qml_amd.representations import cm_length, generate_coulomb_matrices
    from qml_amd.autotune import Autotuner
    from qml_amd.sweep import BatchedFit, learning_curve_grid, sequential_fit, tuning_key

    probs = learning_curve_grid(count=count)
    nmax = max(p.n for p in probs)
    pool, _, size = synthetic.workload("krr_sweep_64", n_train=nmax, n_test=0)
    Xd = generate_coulomb_matrices(*pool.ragged(), size=size)
    d = cm_length(size)
    tuner = Autotuner()
    fit = BatchedFit(probs, Xd, d, pool.labels, tuner=tuner)  # nb and branches tuned
    tuned = tuner._read().get(tuning_key(probs), {}).get("timings", {})
    s = fit.stream
    fit.run()
    reps = 5
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        fit.run(sync=False)
    e1.record(s)
    e1.synchronize()
    graph_ms = e0.elapsed_time(e1) / reps
    e0.record(s)
    fit.run_direct(sync=False)
    e1.record(s)
    e1.synchronize()
    direct_ms = e0.elapsed_time(e1)
    info = fit.info()
    alphas = [a.clone() for a in fit.alphas()]
    with torch.cuda.stream(s):
        sequential_fit(probs, Xd, d, pool.labels, stream=s)  # warm
        s.synchronize()
        e0.record(s)
        seq = sequential_fit(probs, Xd, d, pool.labels, stream=s)
        e1.record(s)
        e1.synchronize()
    seq_ms = e0.elapsed_time(e1)
    # Problems whose FP64 Cholesky breaks down (info > 0: numerically singular K + lambda I,
    # e.g. sigma = 2000 with lambda = 1e-10 at large n) are reported, not compared.
    ok = [i for i in range(len(probs)) if info[i] == 0 and bool(torch.isfinite(seq[i]).all())]
    diff = max(float((alphas[i] - seq[i]).abs().max() / seq[i].abs().max()) for i in ok)
    flops = sum(p.n ** 3 / 3.0 for p in probs)
    pairs = sum(p.n * (p.n + 1) / 2.0 * d for p in probs)
    log(f"sweep: graph {graph_ms:.2f} ms, direct {direct_ms:.2f} ms, sequential {seq_ms:.2f} ms "
        f"(nb={fit.nb}, branches={fit.branches}, {fit.launches} launches)")
    out = {"workload": f"{count} Laplacian KRR fits, n log-uniform [500, 8000], size-23 CM "
                       "(d=276), sigma x lambda grid, shared pool in HBM",
           "problems": count, "n_min": min(p.n for p in probs), "n_max": nmax,
           "nb": fit.nb, "branches": fit.branches, "launches_in_graph": fit.launches,
           "autotune_ms": {k: 1e3 * v for k, v in tuned.items()},
           "graph_ms": graph_ms, "direct_launches_ms": direct_ms, "sequential_ms": seq_ms,
           "speedup_vs_sequential": seq_ms / graph_ms,
           "cholesky_flops": flops, "build_pair_features": pairs,
           "effective_cholesky_tflops": flops / (graph_ms * 1e-3) / 1e12,
           "failed_problems": [{"n": probs[i].n, "sigma": probs[i].sigma, "lambda": probs[i].lam,
                                "info": int(info[i])} for i in np.flatnonzero(info)],
           "max_alpha_rel_diff_vs_sequential": diff,
           "compared_problems": len(ok)}
    fit.close()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "path is not built.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from qml_amd import _lib, serve, synthetic
    from qml_amd.device import ragged_to_device
    from qml_amd.krr import KRR
    from qml_amd.representations import cm_length, generate_coulomb_matrices_device

    torch.cuda.set_device(0)
    _lib.lib()
    if args.copy_mib > 0:
        nel = args.copy_mib * (1 << 20) // 8
        src = torch.ones(nel, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):  # an elementwise kernel: reads and writes copy_mib MiB each
            torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        log(f"copy calibration: 3 x {args.copy_mib} MiB elementwise read + write")
        del src, dst
    if args.local_only:
        print(json.dumps({"local": (run_local(args.local, args.local_cpu_sample, log)
                                    if args.local > 0 else None),
                          "qm7": run_qm7(log) if args.qm7 else None,
                          "cho_invert": run_invert(args.invert, 10000 if args.local_cpu_sample
                                                   else 0, log) if args.invert > 0 else None}),
              flush=True)
        return
    N = args.n_train
    tr, te, size = synthetic.workload("krr_serve_50k", n_train=N, n_test=4096)
    d = cm_length(size)
    Zd, Rd, Od = ragged_to_device(*tr.ragged())
    yd = torch.from_numpy(tr.labels).cuda()
    log(f"workload: N={N} size={size} d={d} sigma={args.sigma} lambda={args.lam}")

    def run_fits(model, tag):
        reps = []

        def step():
            Xd, _ = generate_coulomb_matrices_device(Zd, Rd, Od, size)
            reps.append(model.fit_device(Xd, d, yd))

        for i in range(args.warmup):
            step()
            log(f"{tag} warmup {i} done")
        torch.cuda.synchronize()
        _lib.timer_enable(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(args.steps):
            step()
            log(f"{tag} step {i} done ({reps[-1].iterations} refinement iterations)")
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        timers = {k: _lib.timer_read(k) for k in _lib.TIMER_CLASSES}
        _lib.timer_enable(False)
        return (t1 - t0) / args.steps, timers, reps[-1]

    model64 = KRR(args.kernel, args.sigma, args.lam, "fp64")
    fit64_s, timers, _ = run_fits(model64, "fp64")
    # Predictions of the FP64 fit on a fixed query set, the reference for the mixed fit's error.
    Zq, Rq, offq = te.subset(0, 2048).ragged()
    pred64 = model64.predict_molecules(Zq, Rq, offq, size)
    alpha64 = model64.alpha.clone()
    modelmx = None
    fitmx_s, timers_mx, rep_mx = None, None, None
    if args.precision in ("mixed", "both"):
        modelmx = KRR(args.kernel, args.sigma, args.lam, "mixed", tol=args.tol,
                      factor=None if args.factor == "auto" else args.factor,
                      lowrank_cols=args.lowrank_cols)
        fitmx_s, timers_mx, rep_mx = run_fits(modelmx, "mixed")
        predmx = modelmx.predict_molecules(Zq, Rq, offq, size)
        std = float(np.std(tr.labels))
        mixed_err = {
            "pred_max_abs_over_std": float(np.max(np.abs(predmx - pred64)) / std),
            "alpha_max_rel": float((modelmx.alpha - alpha64).abs().max() / alpha64.abs().max()),
            "refinement_iterations": rep_mx.iterations, "relres": rep_mx.relres,
            "fp32_shift": rep_mx.shift, "reference": "FP64 GPU fit (same data)",
            # Both solutions' residuals, recomputed alike with the prediction kernel over the
            # training rows: which of the two fits is the more accurate at small lambda.
            "relres_recomputed": {"fp64_direct": train_relres(model64, tr.labels),
                                  "mixed": train_relres(modelmx, tr.labels)},
            "backward_error": rep_mx.backward_error,
            "woodbury_term": modelmx.lowrank_mode()}
        if args.kernel == "gaussian":
            Xs_d, _ = generate_coulomb_matrices_device(Zd, Rd, Od, size)
            mode = modelmx.lowrank_mode()
            cols = {"fp32": 0, "fp64": d + 1}.get(mode, modelmx.lowrank_cols)
            mixed_err["kernel_entries"] = kernel_entry_error(Xs_d, d, args.sigma, args.lam, cols)
            del Xs_d
    fit_s = fitmx_s if fitmx_s is not None else fit64_s
    model = modelmx if modelmx is not None else model64

    # Dominant kernel of each fit: the Cholesky trailing update.
    roof64 = roofline(timers, args.steps, "fp64")
    roofmx = roofline(timers_mx, args.steps, modelmx.factor) if timers_mx else None
    phases = {k: {"ms_per_step": v[0] / args.steps, "launches_per_step": v[1] / args.steps,
                  "rate": (v[2] / (v[0] * 1e-3) if v[0] > 0 else 0.0)}
              for k, v in timers.items() if v[1] > 0}

    # Predict: requests of 1..512 molecules (log-uniform), served one at a time.
    rng = np.random.default_rng(5)
    sizes = np.clip(np.exp(rng.uniform(0, np.log(512), args.requests)).astype(int), 1, 512)
    lat = []
    start = 0
    molecules = 0
    for k in sizes:
        sub = te.subset(start % (te.nmol - 512), start % (te.nmol - 512) + k)
        start += k
        Z, R, off = sub.ragged()
        torch.cuda.synchronize()
        ta = time.perf_counter()
        y = model.predict_molecules(Z, R, off, size)
        tb = time.perf_counter()
        lat.append((tb - ta) * 1e3)
        molecules += k
    lat = np.array(lat) if lat else np.full(1, np.nan)
    log(f"fit {fit_s:.3f}s; predict p50 {np.percentile(lat, 50):.2f} ms")

    # Serving load: Poisson arrivals into the continuous-batching server, per model and rate,
    # with refits of N ~ U[1000, 16000] training molecules (the model's precision) submitted
    # under load and admitted by the free-HBM budget.
    serve_out = {}
    rates = [float(r) for r in args.serve_rates.split(",") if r]
    for tag, mdl in (("mixed", modelmx), ("fp64", model64)):
        if mdl is None or args.serve_seconds <= 0:
            continue
        serve_out[tag] = []
        for rate in rates:
            with serve.KRRServer(mdl, size, max_molecules=2048, max_inflight=3) as srv:
                _, _, summ = serve.run_load(
                    srv, te, rate, args.serve_seconds, np.random.default_rng(9), max_mol=512,
                    size=size, refit_pool=tr, refit_every=args.refit_every,
                    refit_precision=mdl.precision)
            serve_out[tag].append(summ)
            log(f"serve {tag} {rate:g} req/s: p50 {summ['p50_ms']:.2f} ms p99 "
                f"{summ['p99_ms']:.2f} ms, {summ['molecules_per_s']:.0f} molecules/s, "
                f"{summ['molecules_per_batch']:.0f} molecules/batch, "
                f"{len(summ['refits'])} refits")
            torch.cuda.empty_cache()

    sweep_out = run_sweep(args.sweep, log) if args.sweep > 0 else None
    kernels_out = run_kernels(log) if args.kernels else None
    local_out = run_local(args.local, args.local_cpu_sample, log) if args.local > 0 else None
    qm7_out = run_qm7(log) if args.qm7 else None
    invert_out = run_invert(args.invert, 10000, log) if args.invert > 0 else None

    cpu = None
    if args.cpu_sample > 0:
        log(f"cpu baseline sample N={args.cpu_sample} ...")
        cpu = cpu_baseline(N, size, args.sigma, args.lam, args.cpu_sample)
        log(f"cpu baseline: {cpu['value']:.1f}s (extrapolated)")

    out = {
        "metric": "KRR fit s + predict p50/p99 ms at N=50k, mixed vs FP64 GPU at equal error",
        "value": fit_s,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fit_s * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32 (mixed)" if fitmx_s is not None else "f64",
        "data": "synthetic (seeded molecules, 5-29 atoms, H/C/N/O/S)",
        "config": {"workload": (f"krr_serve_{N // 1000}k: KRR fit N={N}, Coulomb size {size} "
                                f"(d={d}), {args.kernel.capitalize()} sigma={args.sigma:g}, "
                                f"lambda={_short(args.lam)}; predict requests 1-512 molecules"),
                   "n_train": N, "d": d, "kernel": args.kernel,
                   "factor": modelmx.factor if modelmx is not None else None, "sigma": args.sigma,
                   "lambda": args.lam,
                   "precision": "mixed" if fitmx_s is not None else "fp64"},
        "fit_s": {"mixed": fitmx_s, "fp64": fit64_s,
                  "speedup_mixed_vs_fp64": (fit64_s / fitmx_s) if fitmx_s else None},
        "mixed_error": mixed_err if fitmx_s is not None else None,
        "predict": {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                    "requests": int(args.requests),
                    "molecules_per_s": float(molecules / (lat.sum() / 1e3)) if molecules else None,
                    "mode": "sequential requests, host molecules in, predictions out"},
        "serve": serve_out or None,
        "sweep": sweep_out,
        "kernels": kernels_out,
        "local": local_out,
        "qm7": qm7_out,
        "cho_invert": invert_out,
        "phases_fp64": phases,
        # pcg: the whole refinement (qmlh_pcg_solve), work = CG iterations
        "phases_mixed": ({k: {"ms_per_step": v[0] / args.steps, "launches_per_step": v[1] / args.steps,
                              "work_per_step": v[2] / args.steps}
                          for k, v in timers_mx.items() if v[1] > 0} if timers_mx else None),
        "roofline": roofmx or roof64,
        "roofline_fp64": roof64,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
