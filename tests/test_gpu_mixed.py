"""GPU tests of the mixed-precision fit: FP32 factor, deflated Gaussian build, CG refinement.

Bars: the FP32 factor matches LAPACK's (on the FP32-rounded matrix) normwise to FP32 accuracy;
the deflated build reproduces the FP64 Gaussian kernel bit for bit and B32 = K - u u^T to FP32
rounding; the refined solution's predictions agree with the FP64 path (and with the oracle's
LAPACK path) within 1e-8 of std(y) where the problem's FP64 noise floor allows.
"""

import ctypes

import numpy as np
import pytest

import oracle
from qml_amd import synthetic

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)
    from qml_amd import _lib

    _lib.lib()


@pytest.mark.parametrize("n", [1, 63, 64, 200, 257, 777, 1500, 4600])
def test_spotrf_fp32_factor_vs_lapack(n):
    from scipy.linalg import lapack

    from qml_amd import _lib

    rng = np.random.default_rng(n + 7)
    M = rng.standard_normal((n, n))
    A = (M @ M.T / n + np.eye(n)).astype(np.float32)
    Ad = torch.from_numpy(A.copy()).cuda()
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    _lib.check(_lib.lib().qmlh_spotrf(Ad.data_ptr(), n, n, info.data_ptr(), None))
    assert int(info.item()) == 0
    L = np.tril(Ad.cpu().numpy().T.astype(np.float64))
    Lref, inf = lapack.dpotrf(A.astype(np.float64), lower=1)
    assert inf == 0
    assert np.linalg.norm(L - np.tril(Lref)) <= 2e-6 * np.linalg.norm(Lref)
    # the column-major strictly upper part (row-major strictly lower) is untouched
    assert np.array_equal(np.tril(Ad.cpu().numpy(), -1), np.tril(A, -1))


@pytest.mark.parametrize("n", [257, 700, 1500, 5000, 6145, 9001])
def test_spotrf_f16x3_factor_accuracy(n):
    # Split-FP16 trailing updates: each update ~2^-22 relative, so the factor agrees with
    # LAPACK's to ~1e-5 normwise on a well-conditioned matrix. From order 4096 the trailing
    # updates run on 256 x 256 tiles (n = 9001: every step's tiles are ragged at the edge).
    from scipy.linalg import lapack

    from qml_amd import _lib

    L = _lib.lib()
    rng = np.random.default_rng(n + 11)
    M = rng.standard_normal((n, n))
    A = (M @ M.T / n + np.eye(n)).astype(np.float32)
    Ad = torch.from_numpy(A.copy()).cuda()
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    nb = L.qmlh_spotrf_f16x3_workspace(n)
    ws = torch.empty(nb // 8 + 1, dtype=torch.float64, device="cuda")
    _lib.check(L.qmlh_spotrf_f16x3(Ad.data_ptr(), n, n, info.data_ptr(), ws.data_ptr(),
                                    ws.numel() * 8, None))
    assert int(info.item()) == 0
    Lg = np.tril(Ad.cpu().numpy().T.astype(np.float64))
    Lref = np.tril(lapack.dpotrf(A.astype(np.float64), lower=1)[0])
    err = np.linalg.norm(Lg - Lref) / np.linalg.norm(Lref)
    assert 1e-9 < err <= 5e-5, err  # lower bound: the split path really ran
    assert np.array_equal(np.tril(Ad.cpu().numpy(), -1), np.tril(A, -1))


@pytest.mark.parametrize("n", [1500, 5000])
def test_spotrf_f16x3_panel_scale_is_exact(n):
    # The split panels are scaled by a power of two set from the largest diagonal entry, so
    # scaling A by 4^k scales L by exactly 2^k -- also where the unscaled entries would overflow
    # FP16 (k = 20: |L| ~ 1e6) or fall into its subnormals (k = -20: |L| ~ 1e-6).
    from qml_amd import _lib

    L = _lib.lib()
    rng = np.random.default_rng(n + 5)
    M = rng.standard_normal((n, n))
    A = (M @ M.T / n + np.eye(n)).astype(np.float32)
    ws = torch.empty(L.qmlh_spotrf_f16x3_workspace(n) // 8 + 1, dtype=torch.float64,
                     device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    factors = {}
    for k in (0, 20, -20):
        Ad = torch.from_numpy(A * np.float32(4.0 ** k)).cuda()
        _lib.check(L.qmlh_spotrf_f16x3(Ad.data_ptr(), n, n, info.data_ptr(), ws.data_ptr(),
                                       ws.numel() * 8, None))
        assert int(info.item()) == 0
        factors[k] = np.tril(Ad.cpu().numpy().T)
    for k in (20, -20):
        assert np.array_equal(factors[k], factors[0] * np.float32(2.0 ** k))


@pytest.mark.parametrize("n,bad", [(300, 130), (4400, 1000), (4400, 4390)])
def test_spotrf_info_matches_lapack(n, bad):
    from scipy.linalg import lapack

    from qml_amd import _lib

    A = np.eye(n, dtype=np.float32)
    A[bad, bad] = -1.0
    Ad = torch.from_numpy(A.copy()).cuda()
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    _lib.check(_lib.lib().qmlh_spotrf(Ad.data_ptr(), n, n, info.data_ptr(), None))
    assert int(info.item()) == lapack.dpotrf(A.astype(np.float64), lower=1)[1] == bad + 1


def test_gaussian_deflated_build_identities():
    from qml_amd import _lib
    from qml_amd.device import rows_to_device
    from qml_amd.kernels import gaussian_mfma_device, row_sqnorms_device

    tr, _, size = synthetic.workload("krr_serve_50k", n_train=777, n_test=0)
    X = oracle.coulomb_matrix_batch(*tr.ragged(), size)
    n, d = X.shape
    Xd = rows_to_device(X)
    nrm = row_sqnorms_device(Xd, n, d)
    K = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    B = torch.zeros((n, n), dtype=torch.float32, device="cuda")
    V = torch.empty((d + 1, n), dtype=torch.float32, device="cuda")
    u = torch.empty(n, dtype=torch.float64, device="cuda")
    lam, shift, sigma = 1e-6, 3e-5, 4000.0
    _lib.check(_lib.lib().qmlh_gaussian_deflated(Xd.data_ptr(), n, Xd.shape[1], nrm.data_ptr(),
                                                 d, sigma, lam, shift, K.data_ptr(), n,
                                                 B.data_ptr(), n, V.data_ptr(), n, u.data_ptr(),
                                                 None))
    Kref = gaussian_mfma_device(Xd, nrm, n, Xd, nrm, n, d, sigma, True).cpu().numpy().T
    Kfull = K.cpu().numpy().T
    assert np.all(np.triu(Kfull, 1) == 0)  # K is written on its lower triangle only
    Kg = np.tril(Kfull) + np.tril(Kfull, -1).T
    off = ~np.eye(n, dtype=bool)
    assert np.array_equal(Kg[off], Kref[off])
    assert np.all(np.diag(Kg) == 1.0 + lam)
    Kor = oracle.gaussian_kernel(X, X, sigma)
    assert np.max(np.abs(Kg - lam * np.eye(n) - Kor) / Kor) <= 1e-12
    ug = u.cpu().numpy()
    np.testing.assert_allclose(ug, np.exp(-(X * X).sum(1) / (2 * sigma ** 2)), rtol=1e-14)
    Vg = V.cpu().numpy().T.astype(np.float64)  # (n, d+1)
    np.testing.assert_allclose(Vg[:, 0], ug, rtol=1e-7)
    np.testing.assert_allclose(Vg[:, 1:], ug[:, None] * X / sigma, rtol=1e-7, atol=1e-30)
    Bg = np.tril(B.cpu().numpy().T.astype(np.float64))
    # Remainder of the rank-(d+1) expansion, computed independently in FP64 on the host.
    G = X @ X.T / sigma ** 2
    Bexp = np.tril(np.outer(ug, ug) * (np.expm1(G) - G) + (lam + shift) * np.eye(n))
    assert np.max(np.abs(Bexp)) < 5e-3  # small entries: the point of the deflation
    assert np.max(np.abs(Bg - Bexp)) <= 2e-10
    # K + lambda I = V V^T + B - shift I, to FP32 rounding of V.
    VVt = Vg @ Vg.T
    assert np.max(np.abs(np.tril(VVt) + Bexp - (lam + shift) * np.tril(np.eye(n))
                         + lam * np.tril(np.eye(n)) - np.tril(Kor + lam * np.eye(n)))) <= 1e-6
    assert np.all(np.triu(B.cpu().numpy().T, 1) == 0)


@pytest.mark.parametrize("n", [900, 1000, 4133])
def test_pcg_solve_random_spd(n):
    from qml_amd import _lib

    rng = np.random.default_rng(0)
    M = rng.standard_normal((n, n))
    A = M @ M.T / n + 0.1 * np.eye(n)
    y = rng.standard_normal(n)
    L = _lib.lib()
    Ad = torch.from_numpy(A.copy()).cuda()
    # only the lower triangle of A is read: poison the column-major strict upper part
    Ad[torch.ones(n, n, dtype=torch.bool, device="cuda").tril(-1)] = float("nan")
    Bd = torch.from_numpy(A.astype(np.float32)).cuda()
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    _lib.check(L.qmlh_spotrf(Bd.data_ptr(), n, n, info.data_ptr(), None))
    assert int(info.item()) == 0
    yd = torch.from_numpy(y).cuda()
    x = torch.empty(n, dtype=torch.float64, device="cuda")
    ws = torch.empty(L.qmlh_pcg_workspace(n, 0) // 8 + 1, dtype=torch.float64, device="cuda")
    it, rr = ctypes.c_int32(), ctypes.c_double()
    _lib.check(L.qmlh_pcg_solve(Ad.data_ptr(), n, n, Bd.data_ptr(), n, None, n, 0, 0, None, 1,
                                yd.data_ptr(), x.data_ptr(), 1e-13, 50, ws.data_ptr(),
                                ws.numel() * 8, ctypes.byref(it), ctypes.byref(rr), None, None))
    assert rr.value <= 1e-13 and it.value <= 12
    xs = np.linalg.solve(A, y)
    assert np.linalg.norm(x.cpu().numpy() - xs) <= 1e-11 * np.linalg.norm(xs)


@pytest.mark.parametrize("n,r", [(700, 9), (1333, 150), (4200, 30), (4500, 20)])
def test_woodbury_preconditioner_is_exact_for_exact_factor(n, r):
    # With an exact factor of B, M = B + V V^T equals A, so PCG converges in one iteration.
    from qml_amd import _lib

    rng = np.random.default_rng(4)
    M = rng.standard_normal((n, n)) / np.sqrt(n)
    Bm = M @ M.T + 0.5 * np.eye(n)
    Vm = rng.standard_normal((n, r)) * 3
    A = Bm + Vm @ Vm.T
    y = rng.standard_normal(n)
    L = _lib.lib()
    Bd = torch.from_numpy(Bm.astype(np.float32)).cuda()
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    _lib.check(L.qmlh_spotrf(Bd.data_ptr(), n, n, info.data_ptr(), None))
    Vd = torch.from_numpy(np.ascontiguousarray(Vm.T).astype(np.float32)).cuda()  # column-major
    Cd = torch.empty((r, r), dtype=torch.float64, device="cuda")
    wws = torch.empty(L.qmlh_woodbury_workspace(n) // 8 + 1, dtype=torch.float64, device="cuda")
    _lib.check(L.qmlh_woodbury_setup(Bd.data_ptr(), n, n, Vd.data_ptr(), n, r, Cd.data_ptr(), r,
                                     info.data_ptr(), wws.data_ptr(), wws.numel() * 8, None))
    assert int(info.item()) == 0
    # Y = L^-1 V against a host triangular solve with the same (FP32) factor
    from scipy.linalg import solve_triangular

    Lf = np.tril(Bd.cpu().numpy().T.astype(np.float64))
    Yref = solve_triangular(Lf, Vm.astype(np.float32).astype(np.float64), lower=True)
    Yg = Vd.cpu().numpy().T.astype(np.float64)
    assert np.linalg.norm(Yg - Yref) <= 1e-5 * np.linalg.norm(Yref)
    Ad = torch.from_numpy(A.copy()).cuda()
    yd = torch.from_numpy(y).cuda()
    x = torch.empty(n, dtype=torch.float64, device="cuda")
    ws = torch.empty(L.qmlh_pcg_workspace(n, r) // 8 + 1, dtype=torch.float64, device="cuda")
    it, rr = ctypes.c_int32(), ctypes.c_double()
    _lib.check(L.qmlh_pcg_solve(Ad.data_ptr(), n, n, Bd.data_ptr(), n, Vd.data_ptr(), n, r, 0,
                                Cd.data_ptr(), r, yd.data_ptr(), x.data_ptr(), 1e-11, 50,
                                ws.data_ptr(), ws.numel() * 8, ctypes.byref(it), ctypes.byref(rr),
                                None, None))
    # rr is the recomputed |y - A x| / |y|; about 1.6e-12 is this system's FP64 floor
    # (u |A| |x| / |y|), so the tolerance sits above it
    assert rr.value <= 1e-11 and it.value <= 4  # FP32 factor: a couple of refinement steps
    assert np.linalg.norm(x.cpu().numpy() - np.linalg.solve(A, y)) <= 1e-10 * np.linalg.norm(
        np.linalg.solve(A, y))


@pytest.mark.parametrize("n,r,split16", [(700, 9, 0), (1333, 150, 1), (4500, 437, 1), (4500, 1, 0),
                                         (4500, 1, 1), (4500, 1600, 1), (6145, 437, 1)])
def test_spotrf_lowrank_factor_and_solve(n, r, split16):
    # One pass factors A and solves Y = L^-1 V as extra panel rows: the factor is the plain
    # one bit for bit, Y matches a host triangular solve with that factor, and the capacitance
    # factor matches numpy's Cholesky of I + Y^T Y. From order 4096 the split factor works in
    # groups of two block columns and updates the extra rows by one split-FP16 product per
    # group (r <= 1536; r = 1600 takes the FP32 products); n = 6145 ends in a one-column group.
    from scipy.linalg import solve_triangular

    from qml_amd import _lib

    L = _lib.lib()
    rng = np.random.default_rng(n + r)
    M = rng.standard_normal((n, n)) / np.sqrt(n)
    A = (M @ M.T + 0.5 * np.eye(n)).astype(np.float32)
    Vm = (rng.standard_normal((n, r)) * 3).astype(np.float32)
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    ws = torch.empty(L.qmlh_spotrf_f16x3_workspace(n) // 8 + 1, dtype=torch.float64,
                     device="cuda")
    Ad = torch.from_numpy(A.copy()).cuda()
    if split16:
        _lib.check(L.qmlh_spotrf_f16x3(Ad.data_ptr(), n, n, info.data_ptr(), ws.data_ptr(),
                                        ws.numel() * 8, None))
    else:
        _lib.check(L.qmlh_spotrf(Ad.data_ptr(), n, n, info.data_ptr(), None))
    assert int(info.item()) == 0
    # lda = n + r; the rows below the matrix start as NaN and must be overwritten
    Ab = torch.full((n, n + r), float("nan"), dtype=torch.float32, device="cuda")
    Ab[:, :n] = torch.from_numpy(A).cuda()
    Vd = torch.from_numpy(np.ascontiguousarray(Vm.T)).cuda()  # (n, r) column-major
    _lib.check(L.qmlh_spotrf_lowrank(Ab.data_ptr(), n, n + r, Vd.data_ptr(), n, r, split16, 0.0,
                                     info.data_ptr(), ws.data_ptr() if split16 else None,
                                     ws.numel() * 8 if split16 else 0, None))
    assert int(info.item()) == 0
    Lg = np.tril(Ab[:, :n].cpu().numpy().T)
    assert np.array_equal(Lg, np.tril(Ad.cpu().numpy().T))
    Yref = solve_triangular(Lg.astype(np.float64), Vm.astype(np.float64), lower=True)
    Yg = Vd.cpu().numpy().T.astype(np.float64)
    assert np.linalg.norm(Yg - Yref) <= 1e-5 * np.linalg.norm(Yref)
    if r > 1024:  # the capacitance factor takes r <= 1024
        return
    Cd = torch.empty((r, r), dtype=torch.float64, device="cuda")
    gws = torch.empty(L.qmlh_woodbury_capacitance_workspace(n, r) // 8 + 1, dtype=torch.float64,
                      device="cuda")
    _lib.check(L.qmlh_woodbury_capacitance(Vd.data_ptr(), n, n, r, Cd.data_ptr(), r,
                                           info.data_ptr(), gws.data_ptr(), gws.numel() * 8,
                                           None))
    assert int(info.item()) == 0
    Cref = np.linalg.cholesky(np.eye(r) + Yg.T @ Yg)
    assert np.linalg.norm(np.tril(Cd.cpu().numpy().T) - Cref) <= 1e-12 * np.linalg.norm(Cref)


@pytest.mark.parametrize("n,split16", [(700, 0), (700, 1), (5000, 1)])
def test_spotrf_shift_on_breakdown(n, split16):
    # A = W^T S W with S (n/2 x n/2) well conditioned and W = [I C]: rank n/2, the Schur
    # complement of its leading half exactly zero, so the FP32 factor meets non-positive pivots
    # (rounding) past column n/2. Shifting on breakdown instead factors A + D in one pass, D
    # diagonal: zero before the first failing 64 x 64 block, then that block's shift (1e-2 or
    # 1e-1), nondecreasing, and the last one reported. L L^T reproduces A + D to the FP32
    # factor's backward error, about n u max|A|.
    from qml_amd import _lib

    L = _lib.lib()
    rng = np.random.default_rng(n + 3)
    k = n // 2
    M = rng.standard_normal((k, k)) / np.sqrt(k)
    S = M @ M.T + np.eye(k)
    W = np.hstack([np.eye(k), rng.standard_normal((k, n - k)) / np.sqrt(k)])
    A = (W.T @ S @ W).astype(np.float32)
    r = 0
    ws = torch.empty(L.qmlh_spotrf_f16x3_workspace(n) // 8 + 1, dtype=torch.float64,
                     device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    shift = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
    args = (ws.data_ptr() if split16 else None, ws.numel() * 8 if split16 else 0, None)
    Ad = torch.from_numpy(A.copy()).cuda()
    _lib.check(L.qmlh_spotrf_lowrank(Ad.data_ptr(), n, n, None, n, r, split16, 0.0,
                                     info.data_ptr(), *args))
    assert int(info.item()) > 0  # the plain factor breaks down
    Ad = torch.from_numpy(A.copy()).cuda()
    _lib.check(L.qmlh_spotrf_lowrank_shift(Ad.data_ptr(), n, n, None, n, r, split16, 0.0, 0.0,
                                           1e-2, shift.data_ptr(), info.data_ptr(), *args))
    assert int(info.item()) == 0
    s_last = float(shift.item())
    assert s_last >= 1e-2
    Lg = np.tril(Ad.cpu().numpy().T).astype(np.float64)
    E = Lg @ Lg.T - A.astype(np.float64)
    d = np.diag(E)
    off = E - np.diag(d)
    tol = 2.0 * n * 2.0 ** -24 * np.max(np.abs(A))
    assert np.max(np.abs(off)) <= tol, (np.max(np.abs(off)), tol)
    i0 = int(np.argmax(d > 0.5e-2))  # first raised entry: the start of a 64-row block
    assert i0 % 64 == 0 and i0 >= 64, i0
    assert np.max(np.abs(d[:i0])) <= tol
    levels = np.array([1e-2, 1e-1])
    lv = levels[np.argmin(np.abs(np.log10(np.maximum(d[i0:], 1e-30))[:, None]
                                 - np.log10(levels)[None, :]), axis=1)]
    assert np.max(np.abs(d[i0:] - lv)) <= tol + 0.01 * lv.max()  # each raised entry: a shift
    assert np.all(np.diff(lv) >= 0) and lv[-1] == pytest.approx(s_last)
    for b in range(i0 // 64, (n + 63) // 64):  # one shift per 64 x 64 block
        assert np.all(lv[b * 64 - i0:min(n, b * 64 + 64) - i0] == lv[b * 64 - i0])


@pytest.mark.parametrize("kernel,sigma,lam,n,factor", [("gaussian", 4000.0, 1e-6, 3000, "f16x3"),
                                                       ("gaussian", 4000.0, 1e-6, 3000, "fp32"),
                                                       ("laplacian", 724.0, 1e-8, 2000, "fp32"),
                                                       ("laplacian", 724.0, 1e-8, 2000, "f16x3"),
                                                       # grouped look-ahead, no extra rows
                                                       ("laplacian", 724.0, 1e-8, 6000, "f16x3")])
def test_krr_mixed_matches_fp64_and_oracle(kernel, sigma, lam, n, factor):
    from qml_amd.krr import KRR

    tr, te, size = synthetic.workload("krr_serve_50k", n_train=n, n_test=300)
    X = oracle.coulomb_matrix_batch(*tr.ragged(), size)
    Xs = oracle.coulomb_matrix_batch(*te.ragged(), size)
    std = np.std(tr.labels)
    m64 = KRR(kernel, sigma, lam, "fp64")
    m64.fit(X, tr.labels)
    p64 = m64.predict(Xs)
    mx = KRR(kernel, sigma, lam, "mixed", tol=1e-11, factor=factor)
    rep = mx.fit(X, tr.labels)
    pmx = mx.predict(Xs)
    _, pref = oracle.krr_fit_predict(X, tr.labels, Xs, kernel, sigma, lam)
    # the recomputed residual meets 1e-11 or sits at the FP64 floor (backward error <= u)
    assert mx.converged(rep) and rep.iterations < 30, rep
    assert np.max(np.abs(pmx - p64)) <= 1e-8 * std
    assert np.max(np.abs(pmx - pref)) <= 1e-8 * std
    assert np.max(np.abs(pmx - pref)) <= 1e-8 * np.max(np.abs(pref))  # 1e-8 relative


@pytest.mark.parametrize("pivot_floor", [0.0, 1.0])
def test_mixed_breakdown_retries_with_shift(pivot_floor):
    """Exact duplicate molecules with a tiny lambda: K + lambda I has eigenvalue lambda in the
    duplicates' difference directions, below the FP32 factor's rounding, so the factor breaks
    down at shift 0 and is retried with a diagonal shift -- or, with static pivoting, keeps
    its first factorisation with the failing pivots raised; either way the FP64 CG on the
    unshifted matrix converges (recomputed residual or FP64 backward-error floor) and
    predicts what LAPACK's solves predict, within their own spread."""
    from qml_amd.krr import KRR

    tr, te, size = synthetic.workload("krr_gaussian_20k", n_train=400, n_test=200)
    X1 = oracle.coulomb_matrix_batch(*tr.ragged(), size)
    X = np.repeat(X1, 2, axis=0)  # every molecule twice
    y = np.repeat(tr.labels, 2)
    Xs = oracle.coulomb_matrix_batch(*te.ragged(), size)
    lam = 1e-12
    mx = KRR("gaussian", 4000.0, lam, "mixed", tol=1e-10, pivot_floor=pivot_floor)
    rep = mx.fit(X, y)
    if pivot_floor == 0.0:
        assert rep.shift > 0.0, rep  # the first factorisation broke down and was retried
    else:
        assert rep.shift == 0.0, rep  # static pivoting: the first factorisation is kept
    assert mx.converged(rep), rep
    # At lambda 1e-12 on a rank-deficient K the FP64 answer is itself ill-determined: measure
    # the spread of two LAPACK solves (upper and lower Cholesky) of the oracle's system and
    # hold the mixed fit's predictions to it.
    from scipy.linalg import lapack

    K = oracle.gaussian_kernel(X, X, 4000.0)
    K[np.diag_indices_from(K)] += lam
    a_up = oracle.cho_solve(K, y.copy())
    c, info = lapack.dpotrf(K, lower=1, clean=0, overwrite_a=0)
    assert info == 0
    a_lo = lapack.dpotrs(c, y.reshape(-1, 1), lower=1)[0].reshape(-1)
    Ks = oracle.gaussian_kernel(X, Xs, 4000.0)
    p_up, p_lo = Ks.T @ a_up, Ks.T @ a_lo
    spread = np.max(np.abs(p_up - p_lo))
    diff = np.max(np.abs(mx.predict(Xs) - p_up))
    assert diff <= max(1e-6 * np.std(y), 4.0 * spread), (diff, spread)


@pytest.mark.parametrize("lowrank,cols", [("fp32", 1), ("fp64", 1), ("hybrid", 1), ("hybrid", 9)])
def test_mixed_woodbury_term_modes(lowrank, cols):
    """Every way of forming Y = L^-1 V (FP32 during the factorisation, FP64 after it, or the
    FP32 one with leading columns again in FP64) gives a converged fit that predicts what the
    FP64 fit predicts; the FP64 forms need no more CG iterations than the FP32 one."""
    from qml_amd.krr import KRR

    tr, te, size = synthetic.workload("krr_gaussian_20k", n_train=4000, n_test=300)
    X = oracle.coulomb_matrix_batch(*tr.ragged(), size)
    Xs = oracle.coulomb_matrix_batch(*te.ragged(), size)
    lam = 1e-9
    m64 = KRR("gaussian", 4000.0, lam, "fp64")
    m64.fit(X, tr.labels)
    ref = KRR("gaussian", 4000.0, lam, "mixed", tol=1e-10, lowrank="fp32")
    rep_ref = ref.fit(X, tr.labels)
    mx = KRR("gaussian", 4000.0, lam, "mixed", tol=1e-10, lowrank=lowrank, lowrank_cols=cols)
    rep = mx.fit(X, tr.labels)
    assert mx.converged(rep), rep
    if lowrank != "fp32":
        assert rep.iterations <= rep_ref.iterations, (rep, rep_ref)
    std = np.std(tr.labels)
    assert np.max(np.abs(mx.predict(Xs) - m64.predict(Xs))) <= 1e-5 * std
