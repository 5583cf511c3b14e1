"""The benchmark workloads at their own sizes, against the oracle.

* 20,000 / 2,000 Gaussian KRR (sigma 4000, lambda 1e-10), the FP64 and the mixed fit: backward
  errors of both solutions on the oracle's K, and predictions against the oracle's LAPACK
  solve (qml/fcho_solve.f90:35,46 via scipy), within 1e-8 relative or within the spread of two
  FP64 LAPACK solves of the same system (upper and lower Cholesky, both backward stable), which
  the test measures and states.
* 100,000 training molecules (Laplacian sigma 724 and Gaussian sigma 4000, lambda 1e-9): the
  FP64 and mixed fits at full size, residuals recomputed with the independent prediction
  kernel, kernel rows and predictions of sampled molecules against the oracle.
* The 64-problem learning-curve / (sigma, lambda) sweep (n log-uniform in [500, 8000]): every
  alpha against LAPACK on the oracle's K.
* Serving at 50,000 training molecules: Poisson load with interleaved refits; served results
  bitwise equal to the same requests alone, sampled predictions against oracle kernel rows.

The composition follows tests/test_energy_krr_cmat.py:97-113 (K = kernel(X, X), K[diag] +=
lambda, alpha = cho_solve(K, y), Ks.T . alpha).
"""

import time

import numpy as np
import pytest

import oracle
from qml_amd import synthetic

pytestmark = pytest.mark.gpu

U = 2.0 ** -53  # FP64 unit roundoff


def _norm2_sym(K, iters=30, seed=0):
    """Largest eigenvalue of a symmetric positive semidefinite K by power iteration (a lower
    bound on |K|_2 that converges fast for these nearly rank-one kernels)."""
    v = np.random.default_rng(seed).standard_normal(K.shape[0])
    lam = 0.0
    for _ in range(iters):
        w = K @ v
        lam = float(np.linalg.norm(w) / np.linalg.norm(v))
        v = w / np.linalg.norm(w)
    return lam


def _dot_error_scale(Krows, alpha):
    """sum_j |K_qj alpha_j| per query: the scale a dot product's rounding is measured in."""
    return np.abs(Krows) @ np.abs(alpha)


# --------------------------------------------------------------------------------------------
# 20,000 training / 2,000 test molecules, Gaussian sigma 4000, lambda 1e-10
# --------------------------------------------------------------------------------------------

SIGMA2, LAM2 = 4000.0, 1e-10


@pytest.fixture(scope="module")
def gauss20k():
    from scipy.linalg import lapack

    tr, te, size = synthetic.workload("krr_gaussian_20k")
    assert tr.nmol == 20000 and te.nmol == 2000
    X = oracle.coulomb_matrix_batch(*tr.ragged(), size)
    Xs = oracle.coulomb_matrix_batch(*te.ragged(), size)
    y = tr.labels.copy()
    K = oracle.gaussian_kernel(X, X, SIGMA2)
    K[np.diag_indices_from(K)] += LAM2
    alpha_ref = oracle.cho_solve(K, y.copy())
    # A second backward-stable FP64 solve of the same system: LAPACK's lower Cholesky, which
    # rounds differently. The two predictions' distance is the problem's own FP64 spread.
    c, info = lapack.dpotrf(K, lower=1, clean=0, overwrite_a=0)
    assert info == 0
    alpha_alt, _ = lapack.dpotrs(c, y.reshape(-1, 1), lower=1)
    del c
    Ks = oracle.gaussian_kernel(X, Xs, SIGMA2)
    pred_ref = Ks.T @ alpha_ref
    pred_alt = Ks.T @ alpha_alt.reshape(-1)
    return dict(tr=tr, te=te, size=size, X=X, y=y, K=K, knorm=_norm2_sym(K),
                alpha_ref=alpha_ref, pred_ref=pred_ref,
                spread=float(np.max(np.abs(pred_ref - pred_alt))))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gaussian_20k_lambda_1e10(gauss20k, precision):
    from qml_amd.krr import KRR

    g = gauss20k
    tr, te, size = g["tr"], g["te"], g["size"]
    model = KRR("gaussian", SIGMA2, LAM2, precision)
    rep = model.fit_molecules(*tr.ragged(), size, g["y"])
    alpha = model.alpha.cpu().numpy()
    pred = model.predict_molecules(*te.ragged(), size)
    n = len(alpha)
    # backward error on the oracle's K + lambda I: |K alpha - y| <= n u |K|_2 |alpha|
    r = g["K"] @ alpha - g["y"]
    berr = np.linalg.norm(r) / (g["knorm"] * np.linalg.norm(alpha) + np.linalg.norm(g["y"]))
    r_ref = g["K"] @ g["alpha_ref"] - g["y"]
    berr_ref = np.linalg.norm(r_ref) / (g["knorm"] * np.linalg.norm(g["alpha_ref"]) +
                                        np.linalg.norm(g["y"]))
    print(f"\n{precision}: backward error {berr:.2e} (LAPACK {berr_ref:.2e}, bound n u = "
          f"{n * U:.1e}); iterations {rep.iterations}, relres {rep.relres:.2e}")
    assert berr <= n * U
    # predictions: 1e-8 relative, or the measured spread of two FP64 LAPACK solves
    scale = np.max(np.abs(g["pred_ref"]))
    diff = float(np.max(np.abs(pred - g["pred_ref"])))
    bar = max(1e-8 * scale, 4.0 * g["spread"])
    print(f"{precision}: max |pred - oracle| = {diff:.3e} = {diff / scale:.2e} relative; "
          f"LAPACK upper vs lower spread {g['spread']:.3e}; bar {bar:.3e}")
    assert diff <= bar
    if precision == "mixed":
        assert model.converged(rep)


def test_gaussian_20k_lambda_1e6_mixed_within_1e8_relative(gauss20k):
    """The same 20,000 / 2,000 workload at lambda 1e-6 (the serving benchmark's lambda): the
    mixed fit's predictions within 1e-8 relative of the oracle's LAPACK solve."""
    from qml_amd.krr import KRR

    g = gauss20k
    K = g["K"].copy()
    K[np.diag_indices_from(K)] = 1.0 + 1e-6  # the oracle's K has exactly 1 on its diagonal
    alpha_ref = oracle.cho_solve(K, g["y"].copy())
    Ks = oracle.gaussian_kernel(g["X"], oracle.coulomb_matrix_batch(*g["te"].ragged(), g["size"]),
                                SIGMA2)
    pred_ref = Ks.T @ alpha_ref
    model = KRR("gaussian", SIGMA2, 1e-6, "mixed", tol=1e-10)
    rep = model.fit_molecules(*g["tr"].ragged(), g["size"], g["y"])
    pred = model.predict_molecules(*g["te"].ragged(), g["size"])
    rel = float(np.max(np.abs(pred - pred_ref)) / np.max(np.abs(pred_ref)))
    print(f"\nmixed at lambda 1e-6: {rep.iterations} iterations, prediction difference {rel:.2e} "
          "relative")
    assert rel <= 1e-8


# --------------------------------------------------------------------------------------------
# 100,000 training molecules, Laplacian sigma 724 and Gaussian sigma 4000, lambda 1e-9
# --------------------------------------------------------------------------------------------

LAM3 = 1e-9


@pytest.fixture(scope="module")
def large100k():
    import torch

    from qml_amd.device import ragged_to_device
    from qml_amd.representations import cm_length, generate_coulomb_matrices_device

    tr, te, size = synthetic.workload("krr_large_100k", n_test=2000)
    assert tr.nmol == 100000
    Xd, _ = generate_coulomb_matrices_device(*ragged_to_device(*tr.ragged()), size)
    d = cm_length(size)
    rng = np.random.default_rng(31)
    rows = np.sort(rng.choice(tr.nmol, 64, replace=False))
    Zr, Rr, offr = synthetic.MoleculeSet(tr.natoms[rows], tr.Z_pad[rows], tr.R_pad[rows],
                                         tr.labels[rows]).ragged()
    X_host = Xd[:, :d].cpu().numpy()
    # the device representations themselves are pinned bit for bit on the sampled rows
    assert np.array_equal(X_host[rows], oracle.coulomb_matrix_batch(Zr, Rr, offr, size))
    yd = torch.from_numpy(tr.labels).cuda()
    return dict(tr=tr, te=te, size=size, d=d, Xd=Xd, X=X_host, rows=rows, yd=yd)


def _recomputed_relres(model, yd):
    """|y - (K + lambda I) alpha| / |y| with K alpha from the fused prediction kernel (an
    independent code path: it recomputes every kernel entry), and the lower bound
    |K 1| / sqrt(n) <= |K|_2 from the same kernel."""
    import torch

    ka = model.predict_device(model.Xd, model.n)
    r = yd - ka - model.lam * model.alpha
    alpha = model.alpha
    model.alpha = torch.ones_like(alpha)
    k1 = model.predict_device(model.Xd, model.n)
    model.alpha = alpha
    return float(r.norm() / yd.norm()), float(k1.norm()) / model.n ** 0.5


@pytest.mark.parametrize("kernel,sigma", [("laplacian", 724.0), ("gaussian", 4000.0)])
def test_large_100k_fits(large100k, kernel, sigma):
    import torch

    from qml_amd.kernels import gaussian_mfma_device, pairwise_device, row_sqnorms_device
    from qml_amd.krr import KRR

    L = large100k
    Xd, d, rows, X = L["Xd"], L["d"], L["rows"], L["X"]
    n = Xd.shape[0]
    # 64 full kernel rows from the device kernel against the oracle (1e-12 relative)
    Qd = Xd[torch.from_numpy(rows).cuda()].contiguous()
    if kernel == "gaussian":
        qn = row_sqnorms_device(Qd, len(rows), d)
        nrm = row_sqnorms_device(Xd, n, d)
        # the device tensor is (nb, na) row-major = the (na, nb) column-major K
        Kr = gaussian_mfma_device(Qd, qn, len(rows), Xd, nrm, n, d, sigma, False).cpu().numpy().T
        Kr_ref = oracle.gaussian_kernel(X[rows], X, sigma)
    else:
        from qml_amd import _lib

        Kr = pairwise_device(_lib.PAIR_LAPLACIAN, Qd, len(rows), Xd, n, d, sigma,
                             False).cpu().numpy().T
        Kr_ref = oracle.laplacian_kernel(X[rows], X, sigma)
    assert np.max(np.abs(Kr - Kr_ref) / Kr_ref) <= 1e-12
    del Kr

    Zq, Rq, oq = L["te"].ragged()
    preds, alphas = {}, {}
    for precision in ("fp64", "mixed"):
        model = KRR(kernel, sigma, LAM3, precision, tol=1e-9)
        t0 = time.perf_counter()
        rep = model.fit_device(Xd, d, L["yd"])
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        relres, knorm_lb = _recomputed_relres(model, L["yd"])
        alpha = model.alpha.cpu().numpy()
        # an upper bound on the normwise backward error |r| / (|K|_2 |alpha| + |y|)
        ynorm = float(L["yd"].norm())
        berr = relres * ynorm / (knorm_lb * np.linalg.norm(alpha) + ynorm)
        print(f"\n{kernel} {precision}: fit {dt:.2f} s, recomputed relres {relres:.2e}, "
              f"backward error <= {berr:.2e}; iterations {rep.iterations}, shift {rep.shift:g}")
        assert berr <= n * U
        if precision == "mixed":
            assert model.converged(rep)
        # sampled predictions against oracle kernel rows . alpha (the fused predict kernel)
        pq = model.predict_device(Qd, len(rows)).cpu().numpy()
        pq_ref = Kr_ref @ alpha
        assert np.all(np.abs(pq - pq_ref) <= 1e-12 * _dot_error_scale(Kr_ref, alpha))
        preds[precision] = model.predict_molecules(Zq, Rq, oq, L["size"])
        alphas[precision] = alpha
        del model
        torch.cuda.empty_cache()
    std = float(np.std(L["tr"].labels))
    diff = float(np.max(np.abs(preds["mixed"] - preds["fp64"])) / std)
    print(f"{kernel}: mixed vs FP64 predictions on 2,000 held-out molecules: {diff:.2e} std(y)")
    if kernel == "laplacian":  # well conditioned: the two solutions agree closely
        assert diff <= 1e-6


# --------------------------------------------------------------------------------------------
# 64-problem sweep: n log-uniform in [500, 8000], (sigma, lambda) grid, one graph launch
# --------------------------------------------------------------------------------------------


def test_sweep_64_problems_against_lapack():
    from scipy.linalg import lapack

    from qml_amd.representations import cm_length, generate_coulomb_matrices
    from qml_amd.sweep import BatchedFit, learning_curve_grid

    probs = learning_curve_grid(count=64)
    assert len(probs) == 64 and min(p.n for p in probs) >= 500 and max(p.n for p in probs) <= 8000
    nmax = max(p.n for p in probs)
    pool, _, size = synthetic.workload("krr_sweep_64", n_train=nmax, n_test=0)
    d = cm_length(size)
    Xd = generate_coulomb_matrices(*pool.ragged(), size=size)
    X = oracle.coulomb_matrix_batch(*pool.ragged(), size)
    assert np.array_equal(Xd[:, :d].cpu().numpy(), X)
    y = pool.labels
    fit = BatchedFit(probs, Xd, d, y, graph=True).run()
    info = fit.info()
    alphas = [a.cpu().numpy() for a in fit.alphas()]
    fit.close()
    # Laplacian K = exp(-D / sigma) from the oracle's Manhattan distances (the same serial
    # feature sums as its Laplacian kernel), once for the largest problem.
    D = oracle.manhattan_distance(X, X)
    checked = 0
    for p, a, inf in zip(probs, alphas, info):
        assert p.kernel == "laplacian"
        K = np.exp(D[:p.n, :p.n] * (-1.0 / p.sigma))
        K[np.diag_indices_from(K)] += p.lam
        c, linfo = lapack.dpotrf(K, lower=1, clean=1, overwrite_a=0)
        if linfo != 0:  # numerically singular in FP64 (LAPACK itself breaks down): not compared
            continue
        rcond, _ = lapack.dpocon(c, np.abs(K).sum(0).max(), uplo="L")
        if inf != 0:  # the GPU factor may break down only where LAPACK's nearly does
            assert rcond < 64 * p.n * U, (p, inf, rcond)
            continue
        a_ref, _ = lapack.dpotrs(c, y[:p.n].reshape(-1, 1), lower=1)
        a_ref = a_ref.reshape(-1)
        r = K @ a - y[:p.n]
        berr = np.linalg.norm(r) / (_norm2_sym(K, 10) * np.linalg.norm(a) + np.linalg.norm(y[:p.n]))
        assert berr <= p.n * U, (p, berr)
        fwd = np.linalg.norm(a - a_ref) / np.linalg.norm(a_ref)
        assert fwd <= max(1e-12, 64 * p.n * U / max(rcond, 1e-300)), (p, fwd, rcond)
        checked += 1
    assert checked >= 48


# --------------------------------------------------------------------------------------------
# Serving at 50,000 training molecules: Poisson load, interleaved refits
# --------------------------------------------------------------------------------------------


def test_serving_50k_poisson_with_refits():
    import torch

    from qml_amd import krr, serve
    from qml_amd.device import ragged_to_device
    from qml_amd.representations import cm_length, generate_coulomb_matrices_device

    tr, te, size = synthetic.workload("krr_serve_50k", n_test=4096)
    d = cm_length(size)
    Xd, _ = generate_coulomb_matrices_device(*ragged_to_device(*tr.ragged()), size)
    model = krr.KRR("gaussian", 4000.0, 1e-6)
    model.fit_device(Xd, d, torch.from_numpy(tr.labels).cuda())
    refit_pool, _, _ = synthetic.workload("krr_gaussian_20k", n_train=16000, n_test=0)
    rng = np.random.default_rng(21)
    refit_sizes = [int(v) for v in rng.integers(1000, 16001, 4)]
    with serve.KRRServer(model, size, max_molecules=2048, max_inflight=4) as srv:
        refits = []

        def submit_refits():
            for k, nr in enumerate(refit_sizes):
                time.sleep(0.25)
                sub = refit_pool.subset(0, nr)
                refits.append(srv.submit(serve.RefitRequest(*sub.ragged(), sub.labels,
                                                            size=size)))

        import threading

        th = threading.Thread(target=submit_refits)
        th.start()
        reqs = serve.poisson_load(srv, te, rate=2000.0, duration=1.5, rng=np.random.default_rng(3),
                                  max_mol=512, size=size)
        th.join()
        for r in reqs:
            r.wait(120)
        fitted = [f.wait(300) for f in refits]
    assert len(reqs) > 2000
    # every refit ran while predictions were being served
    for f in refits:
        assert f.t_start is not None
        assert any(f.t_start < r.t_done < f.t_done for r in reqs), "no prediction overlapped a refit"
    # served == alone, bit for bit, on a sample of requests
    for i in np.linspace(0, len(reqs) - 1, 64).astype(int):
        r = reqs[i]
        alone = model.predict_molecules(r.Z, r.R, r.offsets, size)
        assert np.array_equal(r.result.view(np.int64), alone.view(np.int64))
    # 32 sampled query molecules: served prediction against oracle K rows . alpha
    X = Xd[:, :d].cpu().numpy()
    alpha = model.alpha.cpu().numpy()
    picked = [(reqs[i], int(j)) for i, j in zip(rng.integers(0, len(reqs), 32),
                                                 rng.integers(0, 1 << 30, 32))]
    for r, j in picked:
        m = j % r.nmol
        a, b = r.offsets[m], r.offsets[m + 1]
        xq = oracle.coulomb_matrix_batch(r.Z[a:b], r.R[a:b], np.array([0, b - a], np.int64), size)
        kq = oracle.gaussian_kernel(xq, X, 4000.0)[0]
        ref = kq @ alpha
        assert abs(r.result[m] - ref) <= 1e-12 * _dot_error_scale(kq, alpha)
    # the first refit's model equals a standalone fit of the same molecules, bit for bit
    sub = refit_pool.subset(0, refit_sizes[0])
    solo = krr.KRR("gaussian", 4000.0, 1e-6)
    solo.fit_molecules(*sub.ragged(), size, sub.labels)
    assert np.array_equal(fitted[0].alpha.cpu().numpy().view(np.int64),
                          solo.alpha.cpu().numpy().view(np.int64))
