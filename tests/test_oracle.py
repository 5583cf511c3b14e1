"""Pin the CPU oracle (test infrastructure) before trusting it as the GPU's checker.

Each oracle function is checked against the reference's own fixtures and known answers, the
reference Fortran itself (committed vectors from oracle/_ref, and the live build where
present), and closed forms / brute force; a few planted mistakes show the checks bite.
"""

import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, bits, qm7_molecules
from qml_amd import synthetic

# --------------------------------------------------------------------------------------------
# Coulomb matrices
# --------------------------------------------------------------------------------------------


def _qm7_rep(qm7, sorting):
    mols = qm7_molecules(qm7, qm7["rep_names"])
    size = max(len(Z) for Z, _, _ in mols) + 1  # tests/test_representations.py:65
    return size, np.array([oracle.coulomb_matrix(Z, R, size, sorting) for Z, R, _ in mols])


@pytest.mark.parametrize("sorting,fixture", [("row-norm", "cm_row-norm_sorted.txt"),
                                             ("unsorted", "cm_unsorted.txt")])
def test_cm_reference_fixture(qm7, sorting, fixture):
    # tests/test_representations.py:74-90: QM7 0101-0110, size 17, fixtures written %.9e.
    size, X = _qm7_rep(qm7, sorting)
    assert size == 17
    X_ref = np.loadtxt(os.path.join(GOLDEN, fixture))
    assert X.shape == X_ref.shape == (10, 153)
    np.testing.assert_allclose(X, X_ref, rtol=6e-10, atol=0)
    assert np.array_equal(X == 0, X_ref == 0)


def test_cm_water_known_answer():
    # docs/source/examples.rst:37-56
    R = np.array([[1.464, 0.707, 1.056], [0.878, 1.218, 0.498], [2.319, 1.126, 0.952]])
    cm = oracle.coulomb_matrix(np.array([8, 1, 1]), R, 5, "row-norm")
    expect = [73.51669472, 8.3593106, 0.5, 8.35237809, 0.66066557, 0.5] + [0.0] * 9
    # printed by numpy with 8 significant digits
    np.testing.assert_allclose(cm, expect, rtol=1e-8, atol=0)


def test_cm_matches_reference_fortran_vectors(ref_vectors):
    v = ref_vectors
    ours_s = oracle.coulomb_matrix_batch(v["Z"], v["R"], v["offsets"], 29, True)
    ours_u = oracle.coulomb_matrix_batch(v["Z"], v["R"], v["offsets"], 29, False)
    assert np.array_equal(bits(ours_s), bits(v["cm_sorted"]))
    assert np.array_equal(bits(ours_u), bits(v["cm_unsorted"]))


def test_cm_closed_form_and_invariants():
    mols = synthetic.generate(50, 11, 5, 23)
    for m in range(mols.nmol):
        Z, R = mols.molecule(m)
        n = len(Z)
        cm = oracle.coulomb_matrix(Z, R, 23, "row-norm")
        # Rebuild the full matrix from the packed lower triangle.
        M = np.zeros((23, 23))
        M[np.tril_indices(23)] = cm
        M = M + np.tril(M, -1).T
        # Closed form, in the oracle's order.
        Mf = np.zeros((n, n))
        for i in range(n):
            Mf[i, i] = 0.5 * Z[i] ** 2.4
            for j in range(n):
                if i != j:
                    Mf[i, j] = Z[i] * Z[j] / np.linalg.norm(R[i] - R[j])
        rn = (Mf ** 2).sum(axis=1)
        order = np.argsort(-rn, kind="stable")
        np.testing.assert_allclose(M[:n, :n], Mf[np.ix_(order, order)], rtol=1e-14)
        assert np.all(M[n:, :] == 0) and np.all(M[:, n:] == 0)
        # Row norms descend; permutation invariance of the sorted form.
        rows = (M[:n, :n] ** 2).sum(axis=1)
        assert np.all(np.diff(rows) <= 1e-9 * rows[0])
        perm = np.random.default_rng(m).permutation(n)
        cm_p = oracle.coulomb_matrix(Z[perm], R[perm], 23, "row-norm")
        np.testing.assert_allclose(cm_p, cm, rtol=1e-13)


def test_cm_ties_follow_minloc_rule():
    # Exact row-norm ties (mirror-symmetric H2O with identical O-H bonds): the reference's
    # minloc picks the lowest index last, so the higher original index comes first.
    R = np.array([[0.0, 0.0, 0.0], [0.75, 0.5, 0.0], [-0.75, 0.5, 0.0]])
    Z = np.array([8.0, 1.0, 1.0])
    cm = oracle.coulomb_matrix(Z, R, 3, "row-norm")
    if oracle.ref_available():
        assert np.array_equal(bits(cm), bits(oracle.ref_coulomb_matrix(Z, R, 3, "row-norm")))
    # Planted mistake: an ascending sort would put a hydrogen first.
    assert cm[0] == pytest.approx(0.5 * 8 ** 2.4)


def test_cm_planted_mistakes_are_caught(qm7):
    # A wrong ordering rule or a wrong packing must fail the reference fixture.
    size, X = _qm7_rep(qm7, "row-norm")
    X_ref = np.loadtxt(os.path.join(GOLDEN, "cm_row-norm_sorted.txt"))
    mols = qm7_molecules(qm7, qm7["rep_names"])
    bad_order, bad_pack = [], []
    for Z, R, _ in mols:
        n = len(Z)
        full = np.zeros((n, n))
        for i in range(n):
            full[i, i] = 0.5 * Z[i] ** 2.4
            for j in range(n):
                if i != j:
                    full[i, j] = Z[i] * Z[j] / np.linalg.norm(R[j] - R[i])
        rn = (full ** 2).sum(axis=1) - full.diagonal() ** 2  # planted: drop the diagonal term
        o = np.argsort(-rn, kind="stable")
        v = np.zeros(size * (size + 1) // 2)
        S = full[np.ix_(o, o)]
        v[:n * (n + 1) // 2] = S[np.tril_indices(n)]
        bad_order.append(v)
        o = np.argsort(-(full ** 2).sum(axis=1), kind="stable")
        S = full[np.ix_(o, o)]
        w = np.zeros(size * (size + 1) // 2)
        w[:n * (n + 1) // 2] = S[np.triu_indices(n)]  # planted: row-wise upper packing
        bad_pack.append(w)
    assert not np.allclose(np.array(bad_order), X_ref)
    assert not np.allclose(np.array(bad_pack), X_ref)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build oracle/_ref absent")
def test_cm_matches_reference_build_live():
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    mols = synthetic.generate(400, 99, 1, 29)
    for m in range(mols.nmol):
        Z, R = mols.molecule(m)
        for s in ("row-norm", "unsorted"):
            a = oracle.coulomb_matrix(Z, R, 29, s)
            b = oracle.ref_coulomb_matrix(Z, R, 29, s)
            assert np.array_equal(bits(a), bits(b)), (m, s)


# --------------------------------------------------------------------------------------------
# Kernels and distances
# --------------------------------------------------------------------------------------------

_FUNCS = {
    "laplacian": lambda A, B, s: oracle.laplacian_kernel(A, B, s),
    "gaussian": lambda A, B, s: oracle.gaussian_kernel(A, B, s),
    "manhattan": lambda A, B, s: oracle.manhattan_distance(A, B),
    "l2": lambda A, B, s: oracle.l2_distance(A, B),
}


@pytest.mark.parametrize("kind,key,sigma", [("laplacian", "K_laplacian", "sigma_lap"),
                                            ("gaussian", "K_gaussian", "sigma_gauss"),
                                            ("manhattan", "D_manhattan", None),
                                            ("l2", "D_l2", None)])
def test_pairwise_matches_reference_fortran_vectors(ref_vectors, kind, key, sigma):
    v = ref_vectors
    s = float(v[sigma]) if sigma else 1.0
    K = _FUNCS[kind](v["A"], v["B"], s)
    assert K.flags.f_contiguous
    assert np.array_equal(bits(K.T), bits(v[key].T))


def test_kernels_closed_form_reference_test_shape():
    # tests/test_kernels.py:34-90: seed 666, 25x1000 vs 20x1000 in [0,1), sigma = 100.
    np.random.seed(666)
    X = np.random.rand(25, 1000)
    Xs = np.random.rand(20, 1000)
    sigma = 100.0
    Kl = np.exp(-np.abs(X[:, None, :] - Xs[None, :, :]).sum(-1) / sigma)
    Kg = np.exp(-np.square(X[:, None, :] - Xs[None, :, :]).sum(-1) / (2 * sigma ** 2))
    np.testing.assert_allclose(oracle.laplacian_kernel(X, Xs, sigma), Kl, rtol=1e-13)
    np.testing.assert_allclose(oracle.gaussian_kernel(X, Xs, sigma), Kg, rtol=1e-13)
    Ks = oracle.laplacian_kernel(X, X, sigma)
    assert np.array_equal(Ks, Ks.T)
    assert np.all(np.diag(Ks) == 1.0)


def test_distances_brute_force():
    # tests/test_distance.py:30-70 shape (7x5 vs 9x5), seeded here.
    rng = np.random.default_rng(5)
    v1, v2 = rng.random((7, 5)), rng.random((9, 5))
    D1 = np.zeros((7, 9))
    D2 = np.zeros((7, 9))
    for i in range(7):
        for j in range(9):
            for k in range(5):
                D1[i, j] += abs(v1[i, k] - v2[j, k])
                D2[i, j] += (v1[i, k] - v2[j, k]) ** 2
    assert np.array_equal(oracle.manhattan_distance(v1, v2), D1)
    np.testing.assert_allclose(oracle.l2_distance(v1, v2), np.sqrt(D2), rtol=1e-15)


def test_kernel_planted_mistake_is_caught(ref_vectors):
    v = ref_vectors
    A, B = v["A"], v["B"]
    # exp(-D/sigma) evaluated as exp(inv * D) with inv = -1/sigma; planting a squared
    # distance or a missing factor 1/2 in the Gaussian must break the fixture.
    D = oracle.manhattan_distance(A, B)
    assert not np.allclose(np.exp(-D * D / float(v["sigma_lap"])), v["K_laplacian"])
    D2 = oracle.l2_distance(A, B) ** 2
    assert not np.allclose(np.exp(-D2 / float(v["sigma_gauss"]) ** 2), v["K_gaussian"])


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build oracle/_ref absent")
def test_pairwise_matches_reference_build_live():
    rng = np.random.default_rng(17)
    for (na, nb, d) in [(1, 1, 1), (3, 130, 17), (129, 7, 435), (64, 64, 276)]:
        A = rng.random((na, d)) * 30
        B = rng.random((nb, d)) * 30
        for kind in ("laplacian", "gaussian", "manhattan", "l2"):
            ours = _FUNCS[kind](A, B, 37.0)
            ref = oracle.ref_kernel(kind, A, B, 37.0)
            assert np.array_equal(bits(ours.T), bits(ref.T)), (kind, na, nb, d)


# --------------------------------------------------------------------------------------------
# cho_solve
# --------------------------------------------------------------------------------------------


def test_cho_solve_reference_fixture(cho_pair):
    # tests/test_math.py:32-51
    A_ref, y_ref = cho_pair
    A, y = A_ref.copy(), y_ref.copy()
    x = oracle.cho_solve(A, y)
    # The fixture is symmetric only to 1e-13; the reference's test compares with allclose
    # because cho_solve copies the lower triangle onto the upper one.
    assert np.allclose(A, A_ref) and np.array_equal(y, y_ref)
    assert np.array_equal(np.triu(A, 1), np.tril(A_ref, -1).T)
    # The reference's tolerance (np.allclose defaults): K_local_gaussian has cond ~1e10, so two
    # correct solvers agree only to ~1e-6 relative; the residual below is the tight check.
    assert np.allclose(x, np.linalg.solve(A_ref, y_ref))
    r = A_ref @ x - y_ref
    assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(A_ref) * np.linalg.norm(x)


def test_cho_solve_uses_upper_and_symmetrises_lower_to_upper():
    rng = np.random.default_rng(3)
    M = rng.random((40, 40))
    S = M @ M.T + 40 * np.eye(40)
    A = np.tril(rng.random((40, 40)), -1) * 7 + np.triu(S)  # lower part is junk
    y = rng.random(40)
    x = oracle.cho_solve(A, y)
    np.testing.assert_allclose(S @ x, y, rtol=1e-12)
    assert np.array_equal(np.triu(A, 1), np.tril(A, -1).T)  # upper := lower afterwards


def test_cho_solve_non_pd_reports_lapack_info(capsys):
    A = np.eye(6)
    A[3, 3] = -1.0
    oracle.cho_solve(A, np.ones(6))
    out = capsys.readouterr().out
    assert "WARNING" in out and " 4 -th leading order" in out


def test_krr_oracle_end_to_end_qm7(qm7):
    # tests/test_energy_krr_cmat.py:54-113 with the oracle: MAE < 6.0.
    mols = qm7_molecules(qm7, qm7["krr_names"])
    X = np.array([oracle.coulomb_matrix(Z, R, 23, "row-norm") for Z, R, _ in mols])
    Y = np.array([h for _, _, h in mols])
    np.random.seed(666)
    idx = np.arange(len(mols))
    np.random.shuffle(idx)
    tr, te = idx[:700], idx[-300:]
    _, pred = oracle.krr_fit_predict(X[tr], Y[tr], X[te], "laplacian", 10 ** 4.2, 1e-10)
    mae = np.mean(np.abs(Y[te] - pred))
    assert mae < 6.0, mae
