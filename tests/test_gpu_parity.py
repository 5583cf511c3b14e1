"""GPU parity: the HIP path (through libqmlhip.so) against the CPU oracle.

Bars: Coulomb matrices bit-exact; FP64 kernel and distance matrices within 1e-12 relative
(Manhattan distances bit-exact: same feature order, same operations); Cholesky solutions and
KRR predictions within 1e-8 relative of the oracle's LAPACK path where the problem's own
FP64 noise floor allows (stated per test).
"""

import numpy as np
import pytest

import oracle
from conftest import bits, qm7_molecules
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


def _rel(a, b):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


# --------------------------------------------------------------------------------------------
# Coulomb matrices: bit-exact
# --------------------------------------------------------------------------------------------


@pytest.mark.parametrize("size,max_atoms,nmol", [(23, 23, 3000), (29, 29, 3000), (40, 29, 257)])
@pytest.mark.parametrize("sorting", ["row-norm", "unsorted"])
def test_cm_batch_bitexact(size, max_atoms, nmol, sorting):
    from qml_amd.representations import generate_coulomb_matrices

    mols = synthetic.generate(nmol, 100 + size, 1, max_atoms)
    Z, R, off = mols.ragged()
    d = size * (size + 1) // 2
    got = generate_coulomb_matrices(Z, R, off, size, sorting).cpu().numpy()
    ref = oracle.coulomb_matrix_batch(Z, R, off, size, sorting == "row-norm")
    assert np.array_equal(bits(got[:, :d]), bits(ref))
    assert np.all(got[:, d:] == 0)


def test_cm_reference_fortran_vectors(ref_vectors):
    from qml_amd.representations import generate_coulomb_matrices

    v = ref_vectors
    for sorting, key in (("row-norm", "cm_sorted"), ("unsorted", "cm_unsorted")):
        got = generate_coulomb_matrices(v["Z"], v["R"], v["offsets"], 29, sorting)
        assert np.array_equal(bits(got[:, :435].cpu().numpy()), bits(v[key]))


def test_cm_dropin_qm7_fixture_and_water(qm7):
    import os

    from conftest import GOLDEN
    from qml_amd.representations import generate_coulomb_matrix

    mols = qm7_molecules(qm7, qm7["rep_names"])
    for sorting, fx in (("row-norm", "cm_row-norm_sorted.txt"), ("unsorted", "cm_unsorted.txt")):
        X = np.array([generate_coulomb_matrix(Z, R, size=17, sorting=sorting) for Z, R, _ in mols])
        Xo = np.array([oracle.coulomb_matrix(Z, R, 17, sorting) for Z, R, _ in mols])
        assert np.array_equal(bits(X), bits(Xo))
        np.testing.assert_allclose(X, np.loadtxt(os.path.join(GOLDEN, fx)), rtol=6e-10, atol=0)
    R = np.array([[1.464, 0.707, 1.056], [0.878, 1.218, 0.498], [2.319, 1.126, 0.952]])
    cm = generate_coulomb_matrix(np.array([8, 1, 1]), R, size=5, sorting="row-norm")
    assert cm.shape == (15,) and cm.dtype == np.float64
    np.testing.assert_allclose(cm[:6], [73.51669472, 8.3593106, 0.5, 8.35237809, 0.66066557, 0.5],
                               rtol=1e-8)


def test_cm_edge_cases():
    from qml_amd.device import ragged_to_device
    from qml_amd.representations import generate_coulomb_matrices, generate_coulomb_matrices_device

    rng = np.random.default_rng(1)
    # one-atom molecules, natoms == size, a 100-atom molecule at size 128, exact ties
    cases = [(np.array([6.0]), np.zeros((1, 3)), 3),
             (np.array([8.0, 1.0, 1.0]), np.array([[0, 0, 0], [.75, .5, 0], [-.75, .5, 0.]]), 3),
             (rng.choice([1.0, 6.0, 7.0, 8.0], 100), rng.random((100, 3)) * 20, 128)]
    for Z, R, size in cases:
        d = size * (size + 1) // 2
        got = generate_coulomb_matrices(Z, R, np.array([0, len(Z)]), size).cpu().numpy()[0, :d]
        assert np.array_equal(bits(got), bits(oracle.coulomb_matrix(Z, R, size)))
    # empty batch
    out = generate_coulomb_matrices(np.zeros(0), np.zeros((0, 3)), np.zeros(1, np.int64), 23)
    assert out.shape[0] == 0
    # natoms > size through the device API: flagged, row zeroed, others intact
    Z = np.array([1.0] * 4 + [6.0] * 2)
    R = rng.random((6, 3)) * 3
    off = np.array([0, 4, 6])
    Zd, Rd, Od = ragged_to_device(Z, R, off)
    X, status = generate_coulomb_matrices_device(Zd, Rd, Od, 3)
    assert int(status.item()) == 1
    X = X.cpu().numpy()
    assert np.all(X[0] == 0)
    assert np.array_equal(bits(X[1, :6]), bits(oracle.coulomb_matrix(Z[4:], R[4:], 3)))


# --------------------------------------------------------------------------------------------
# Pairwise kernels and distances
# --------------------------------------------------------------------------------------------

_ORACLE = {
    "laplacian": lambda A, B, s: oracle.laplacian_kernel(A, B, s),
    "gaussian": lambda A, B, s: oracle.gaussian_kernel(A, B, s),
    "manhattan": lambda A, B, s: oracle.manhattan_distance(A, B),
    "l2": lambda A, B, s: oracle.l2_distance(A, B),
}


def _dropin(kind):
    from qml_amd import distance, kernels

    return {"laplacian": lambda A, B, s: kernels.laplacian_kernel(A, B, s),
            "gaussian": lambda A, B, s: kernels.gaussian_kernel(A, B, s),
            "manhattan": lambda A, B, s: distance.manhattan_distance(A, B),
            "l2": lambda A, B, s: distance.l2_distance(A, B)}[kind]


@pytest.mark.parametrize("kind", ["laplacian", "gaussian", "manhattan", "l2"])
@pytest.mark.parametrize("na,nb,d,scale,sigma", [(1, 1, 1, 1.0, 1.0), (7, 9, 5, 1.0, 0.7),
                                                 (25, 20, 1000, 1.0, 100.0),
                                                 (130, 257, 276, 50.0, 724.0),
                                                 (129, 1, 435, 300.0, 4000.0)])
def test_pairwise_vs_oracle(kind, na, nb, d, scale, sigma):
    rng = np.random.default_rng(na * 1000 + nb + d)
    A = rng.random((na, d)) * scale
    B = rng.random((nb, d)) * scale
    got = _dropin(kind)(A, B, sigma)
    ref = _ORACLE[kind](A, B, sigma)
    assert got.shape == (na, nb) and got.flags.f_contiguous
    if kind == "manhattan":
        assert np.array_equal(bits(got.T), bits(ref.T))
    else:
        assert _rel(got, ref) <= 1e-12


@pytest.mark.parametrize("kind", ["laplacian", "gaussian", "manhattan", "l2"])
def test_pairwise_symmetric_vs_oracle(kind):
    tr, _, size = synthetic.workload("krr_serve_50k", n_train=300, n_test=0)
    X = oracle.coulomb_matrix_batch(*tr.ragged(), size)
    sigma = {"laplacian": 724.0, "gaussian": 4000.0}.get(kind, 1.0)
    got = _dropin(kind)(X, X, sigma)
    ref = _ORACLE[kind](X, X, sigma)
    assert np.array_equal(got, got.T)
    if kind in ("laplacian", "gaussian"):
        assert np.all(np.diag(got) == 1.0)
    else:
        assert np.all(np.diag(got) == 0.0)
    if kind == "manhattan":
        assert np.array_equal(bits(got.T), bits(ref.T))
    else:
        assert _rel(got, ref) <= 1e-12


@pytest.mark.parametrize("n", [1, 300, 700])
def test_pairwise_lower_mode(n):
    # QMLH_LOWER: the same entries bit for bit on and below the diagonal as the full symmetric
    # build (the transposed tile sums the same differences in the same feature order), the
    # upper triangle untouched; also through qmlh_pairwise_ext, and rejected without
    # QMLH_SYMMETRIC.
    from qml_amd import _lib
    from qml_amd.device import rows_to_device
    from qml_amd.kernels import pairwise_device

    tr, _, size = synthetic.workload("krr_serve_50k", n_train=n, n_test=0)
    X = oracle.coulomb_matrix_batch(*tr.ragged(), size)
    d = X.shape[1]
    Xd = rows_to_device(X)
    low = np.tril(np.ones((n, n), dtype=bool))
    L = _lib.lib()
    prm = np.array([724.0, 0.5, 0.25])
    fulls = {}
    for kind, sigma in ((_lib.PAIR_LAPLACIAN, 724.0), (_lib.PAIR_GAUSSIAN, 4000.0),
                        (_lib.PAIR_L2, 1.0), ("sargan", None)):
        outs = []
        for flags in (_lib.SYMMETRIC, _lib.SYMMETRIC | _lib.LOWER):
            K = torch.full((n, n), -7.0, dtype=torch.float64, device="cuda")
            if kind == "sargan":
                _lib.check(L.qmlh_pairwise_ext(_lib.PAIR_SARGAN, _lib.ptr(Xd), n, Xd.shape[1],
                                               _lib.ptr(Xd), n, Xd.shape[1], d, prm.ctypes.data,
                                               prm.size, _lib.ptr(K), n, flags, None))
            else:
                _lib.check(L.qmlh_pairwise(kind, _lib.ptr(Xd), n, Xd.shape[1], _lib.ptr(Xd), n,
                                           Xd.shape[1], d, sigma, _lib.ptr(K), n, flags, None))
            outs.append(K.cpu().numpy().T)
        full, lo = outs
        fulls[kind] = full
        assert np.array_equal(full, full.T), kind
        assert np.array_equal(bits(lo[low]), bits(full[low])), kind
        assert np.all(lo[~low] == -7.0), kind
    Kl = pairwise_device(_lib.PAIR_LAPLACIAN, Xd, n, Xd, n, d, 724.0, True, lower=True)
    assert np.array_equal(bits(Kl.cpu().numpy().T[low]), bits(fulls[_lib.PAIR_LAPLACIAN][low]))
    assert L.qmlh_pairwise(_lib.PAIR_LAPLACIAN, _lib.ptr(Xd), n, Xd.shape[1], _lib.ptr(Xd), n,
                           Xd.shape[1], d, 724.0, _lib.ptr(Kl), n, _lib.LOWER, None) != 0


def test_pairwise_reference_fortran_vectors(ref_vectors):
    v = ref_vectors
    for kind, key, sig in (("laplacian", "K_laplacian", float(v["sigma_lap"])),
                           ("gaussian", "K_gaussian", float(v["sigma_gauss"])),
                           ("manhattan", "D_manhattan", 1.0), ("l2", "D_l2", 1.0)):
        got = _dropin(kind)(v["A"], v["B"], sig)
        assert _rel(got, v[key]) <= 1e-12, kind


def test_reference_kernel_test_closed_form():
    # tests/test_kernels.py:34-90 at a 1e-12 bar instead of allclose.
    from qml_amd.kernels import gaussian_kernel, laplacian_kernel

    np.random.seed(666)
    X = np.random.rand(25, 1000)
    Xs = np.random.rand(20, 1000)
    sigma = 100.0
    Kl = np.exp(-np.abs(X[:, None, :] - Xs[None, :, :]).sum(-1) / sigma)
    Kg = np.exp(-np.square(X[:, None, :] - Xs[None, :, :]).sum(-1) / (2 * sigma ** 2))
    assert _rel(laplacian_kernel(X, Xs, sigma), Kl) <= 1e-12
    assert _rel(gaussian_kernel(X, Xs, sigma), Kg) <= 1e-12
    for f in (laplacian_kernel, gaussian_kernel):
        Ks = f(X, X, sigma)
        assert np.array_equal(Ks, Ks.T)


def test_gaussian_mfma_expanded_form_vs_direct():
    from qml_amd.device import rows_to_device
    from qml_amd.kernels import gaussian_mfma_device, row_sqnorms_device

    tr, te, size = synthetic.workload("krr_serve_50k", n_train=700, n_test=333)
    X = oracle.coulomb_matrix_batch(*tr.ragged(), size)
    Xs = oracle.coulomb_matrix_batch(*te.ragged(), size)
    d = X.shape[1]
    Xd, Xsd = rows_to_device(X), rows_to_device(Xs)
    nX, nXs = row_sqnorms_device(Xd, 700, d), row_sqnorms_device(Xsd, 333, d)
    K = gaussian_mfma_device(Xd, nX, 700, Xsd, nXs, 333, d, 4000.0).cpu().numpy().T
    assert _rel(K, oracle.gaussian_kernel(X, Xs, 4000.0)) <= 1e-12
    Ksym = gaussian_mfma_device(Xd, nX, 700, Xd, nX, 700, d, 4000.0, True).cpu().numpy().T
    assert np.array_equal(Ksym, Ksym.T) and np.all(np.diag(Ksym) == 1.0)
    assert _rel(Ksym, oracle.gaussian_kernel(X, X, 4000.0)) <= 1e-12
    # lower-triangle mode: the same entries bit for bit below and on the diagonal (the tile
    # that computes them is the transposed one, whose products and sums run in the same
    # order), the upper triangle untouched
    import torch

    Kl = torch.full((700, 700), -7.0, dtype=torch.float64, device="cuda")
    gaussian_mfma_device(Xd, nX, 700, Xd, nX, 700, d, 4000.0, True, out=Kl, lower=True)
    Kl = Kl.cpu().numpy().T  # the device tensor holds K column-major
    low = np.tril(np.ones((700, 700), dtype=bool))
    assert np.array_equal(Kl[low], Ksym[low])
    assert np.all(Kl[~low] == -7.0)


# --------------------------------------------------------------------------------------------
# Cholesky
# --------------------------------------------------------------------------------------------


# 4097 and 5000 run the look-ahead schedule (second stream for the next block column).
@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 200, 255, 256, 257, 513, 1100, 4097, 5000])
def test_potrf_potrs_vs_lapack(n):
    from scipy.linalg import lapack

    from qml_amd.math import potrf_device, potrs_device

    rng = np.random.default_rng(n)
    M = rng.standard_normal((n, n))
    A = M @ M.T + n * np.eye(n)
    y = rng.standard_normal(n)
    Ad = torch.from_numpy(A.copy()).cuda()
    info = potrf_device(Ad, n)
    assert int(info.item()) == 0
    L = Ad.cpu().numpy().T  # column-major view
    Lref, inf = lapack.dpotrf(A, lower=1)
    assert inf == 0
    # Normwise: a backward-stable factor matches LAPACK's to ~cond(A) * u, and L L^T = A.
    Lg = np.tril(L)
    assert np.linalg.norm(Lg - np.tril(Lref)) <= 1e-13 * np.linalg.norm(Lref)
    assert np.linalg.norm(Lg @ Lg.T - A) <= 4 * n * 2.0 ** -53 * np.linalg.norm(A)
    assert np.array_equal(np.triu(L, 1), np.triu(A, 1))  # strictly upper part untouched
    xd = torch.from_numpy(y.copy()).cuda()
    potrs_device(Ad, n, xd)
    x = xd.cpu().numpy()
    xref, _ = lapack.dpotrs(Lref, y.reshape(-1, 1), lower=1)
    assert _rel(x, xref.ravel()) <= 1e-10
    assert np.linalg.norm(A @ x - y) <= 1e-13 * np.linalg.norm(A) * np.linalg.norm(x)


@pytest.mark.parametrize("n", [1024, 4100, 9001])
def test_potrs_single_launch_solves(n):
    # Orders >= 1024 use the single-launch triangular solves (workgroups hand blocks to each
    # other through flags); repeated solves must agree bit for bit and match LAPACK.
    from scipy.linalg import lapack

    from qml_amd.math import potrf_device, potrs_device

    rng = np.random.default_rng(n)
    M = rng.standard_normal((n, 64))
    A = M @ M.T / 64 + np.eye(n)
    Ad = torch.from_numpy(A.copy()).cuda()
    assert int(potrf_device(Ad, n).item()) == 0
    Lref = np.tril(lapack.dpotrf(A, lower=1)[0])
    y = rng.standard_normal(n)
    outs = []
    for _ in range(4):
        xd = torch.from_numpy(y.copy()).cuda()
        potrs_device(Ad, n, xd)
        outs.append(xd.cpu().numpy())
    for o in outs[1:]:
        assert np.array_equal(bits(o), bits(outs[0]))
    xref = lapack.dpotrs(Lref, y.reshape(-1, 1), lower=1)[0].ravel()
    assert np.linalg.norm(outs[0] - xref) <= 1e-11 * np.linalg.norm(xref)
    assert np.linalg.norm(A @ outs[0] - y) <= 1e-13 * np.linalg.norm(A) * np.linalg.norm(outs[0])


@pytest.mark.parametrize("n,lda", [(1100, 1109), (4100, 4224)])
def test_potrs_leading_dimension(n, lda):
    # The factor embedded in a wider column-major array (lda > n): the single-launch solves and
    # their diagonal-block inverses read it through lda; the result equals the dense case's.
    from qml_amd import _lib
    from qml_amd.math import potrf_device, potrs_device

    rng = np.random.default_rng(lda)
    M = rng.standard_normal((n, 64))
    A = M @ M.T / 64 + np.eye(n)
    y = rng.standard_normal(n)
    Ad = torch.from_numpy(A.copy()).cuda()
    assert int(potrf_device(Ad, n).item()) == 0
    xd = torch.from_numpy(y.copy()).cuda()
    potrs_device(Ad, n, xd)
    big = torch.full((n, lda), 3.0, dtype=torch.float64, device="cuda")
    big[:, :n] = Ad  # row c of the (n, lda) view is column c of the factor
    xw = torch.from_numpy(y.copy()).cuda()
    L = _lib.lib()
    ws = torch.empty((L.qmlh_potrs_workspace(n) + 7) // 8, dtype=torch.float64, device="cuda")
    _lib.check(L.qmlh_potrs(_lib.ptr(big), n, lda, _lib.ptr(xw), _lib.ptr(ws), ws.numel() * 8,
                            _lib.stream_handle(None)))
    assert np.array_equal(bits(xw.cpu().numpy()), bits(xd.cpu().numpy()))


@pytest.mark.parametrize("n,bad", [(400, 0), (400, 5), (400, 64), (400, 255), (400, 256),
                                   (400, 300), (4500, 0), (4500, 300), (4500, 2100), (4500, 4499)])
def test_potrf_info_matches_lapack(n, bad):
    from scipy.linalg import lapack

    from qml_amd.math import potrf_device

    rng = np.random.default_rng(bad)
    M = rng.standard_normal((n, n))
    A = M @ M.T / n + np.eye(n)
    A[bad, bad] = -1.0
    _, inf_ref = lapack.dpotrf(A, lower=1)
    info = potrf_device(torch.from_numpy(A.copy()).cuda(), n)
    assert int(info.item()) == inf_ref == bad + 1


def test_cho_solve_dropin_fixture(cho_pair, capsys):
    from qml_amd.math import cho_solve

    A_ref, y_ref = cho_pair
    A, y = A_ref.copy(), y_ref.copy()
    x = cho_solve(A, y)
    assert np.allclose(A, A_ref) and np.array_equal(y, y_ref)
    assert np.array_equal(np.triu(A, 1), np.tril(A_ref, -1).T)
    assert np.allclose(x, np.linalg.solve(A_ref, y_ref))  # tests/test_math.py:48-51
    x_oracle = oracle.cho_solve(A_ref.copy(), y_ref.copy())
    # cond(K) ~ 1e10: two correct FP64 factorisations agree to ~cond * u.
    assert _rel(x, x_oracle) <= 1e-4
    r = A_ref @ x - y_ref
    assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(A_ref) * np.linalg.norm(x)
    assert "WARNING" not in capsys.readouterr().out


def test_cho_solve_dropin_non_pd_warns(capsys):
    from qml_amd.math import cho_solve

    A = np.eye(70)
    A[66, 66] = -2.0
    cho_solve(A, np.ones(70))
    out = capsys.readouterr().out
    assert "WARNING: Error in LAPACK Cholesky decomposition DPOTRF()." in out
    assert " 67 -th leading order is not positive definite." in out


# --------------------------------------------------------------------------------------------
# KRR end to end
# --------------------------------------------------------------------------------------------


def test_krr_laplacian_small_config_parity():
    # 1000 train / 100 test, size 23, Laplacian sigma 724, lambda 1e-8; output: test MAE.
    from qml_amd.krr import KRR

    tr, te, size = synthetic.workload("krr_laplacian_1k")
    Z, R, off = tr.ragged()
    Zs, Rs, offs = te.ragged()
    model = KRR("laplacian", 724.0, 1e-8)
    model.fit_molecules(Z, R, off, size, tr.labels)
    y = model.predict_molecules(Zs, Rs, offs, size)
    X = oracle.coulomb_matrix_batch(Z, R, off, size)
    Xs = oracle.coulomb_matrix_batch(Zs, Rs, offs, size)
    alpha_ref, y_ref = oracle.krr_fit_predict(X, tr.labels, Xs, "laplacian", 724.0, 1e-8)
    assert _rel(y, y_ref) <= 1e-8
    mae, mae_ref = np.mean(np.abs(y - te.labels)), np.mean(np.abs(y_ref - te.labels))
    assert abs(mae - mae_ref) <= 1e-8 * mae_ref
    assert _rel(model.alpha.cpu().numpy(), alpha_ref) <= 1e-6


def test_krr_gaussian_parity():
    from qml_amd.krr import KRR

    tr, te, size = synthetic.workload("krr_gaussian_20k", n_train=2000, n_test=200)
    Z, R, off = tr.ragged()
    Zs, Rs, offs = te.ragged()
    model = KRR("gaussian", 4000.0, 1e-6)
    model.fit_molecules(Z, R, off, size, tr.labels)
    y = model.predict_molecules(Zs, Rs, offs, size)
    X = oracle.coulomb_matrix_batch(Z, R, off, size)
    Xs = oracle.coulomb_matrix_batch(Zs, Rs, offs, size)
    _, y_ref = oracle.krr_fit_predict(X, tr.labels, Xs, "gaussian", 4000.0, 1e-6)
    # within 1e-8 relative of the oracle's predictions
    assert np.max(np.abs(y - y_ref)) <= 1e-8 * np.max(np.abs(y_ref))


def test_predict_is_batch_invariant():
    from qml_amd.device import rows_to_device
    from qml_amd.krr import KRR

    tr, te, size = synthetic.workload("krr_serve_50k", n_train=1500, n_test=300)
    X = oracle.coulomb_matrix_batch(*tr.ragged(), size)
    Xs = oracle.coulomb_matrix_batch(*te.ragged(), size)
    for kern, sigma in (("gaussian", 4000.0), ("laplacian", 724.0)):
        model = KRR(kern, sigma, 1e-6)
        model.fit(X, tr.labels)
        full = model.predict(Xs)
        for q in (0, 17, 64, 65, 299):
            alone = model.predict(Xs[q:q + 1])
            assert bits(alone)[0] == bits(full)[q], (kern, q)
        mixed = model.predict(Xs[::-1].copy())
        assert np.array_equal(bits(mixed[::-1]), bits(full))


def test_qm7_end_to_end_dropins(qm7):
    # tests/test_energy_krr_cmat.py:54-113 through the drop-ins: MAE < 6.0.
    from qml_amd.kernels import laplacian_kernel
    from qml_amd.math import cho_solve
    from qml_amd.representations import generate_coulomb_matrices

    mols = qm7_molecules(qm7, qm7["krr_names"])
    Z = np.concatenate([m[0] for m in mols]).astype(np.float64)
    R = np.concatenate([m[1] for m in mols])
    off = np.zeros(len(mols) + 1, np.int64)
    np.cumsum([len(m[0]) for m in mols], out=off[1:])
    X = generate_coulomb_matrices(Z, R, off, 23)[:, :276].cpu().numpy()
    Y = np.array([m[2] for m in mols])
    np.random.seed(666)
    idx = np.arange(len(mols))
    np.random.shuffle(idx)
    tr, te = idx[:700], idx[-300:]
    sigma, lam = 10 ** 4.2, 1e-10
    K = laplacian_kernel(X[tr], X[tr], sigma)
    K[np.diag_indices_from(K)] += lam
    alpha = cho_solve(K, Y[tr])
    Ks = laplacian_kernel(X[tr], X[te], sigma)
    mae = np.mean(np.abs(Y[te] - Ks.T.dot(alpha)))
    _, pred_ref = oracle.krr_fit_predict(X[tr], Y[tr], X[te], "laplacian", sigma, lam)
    assert mae < 6.0
    assert abs(mae - np.mean(np.abs(Y[te] - pred_ref))) < 1e-6 * mae
