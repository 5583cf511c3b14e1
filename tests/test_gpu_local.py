"""GPU parity for the atomic Coulomb matrices and the local (atom-summed) kernels.

Bars: atomic Coulomb matrices bit-exact against the oracle without cut-offs; with cut-offs the
decay factor 0.5 (cos(x) + 1) uses the device's cos, which may differ from libm's in the last
place and cancels near the cut-off (x -> pi), so there entries agree to 1e-14 of their row's
largest entry, with the same ordering and zero pattern. Local kernels within 1e-12 relative of the oracle (device exp and the
matrix-core summation order differ from the Fortran's serial sum in the last places).
"""

import numpy as np
import pytest

import oracle
from conftest import bits, qm7_local_split, qm7_molecules
from qml_amd import synthetic

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

CUT = dict(central_cutoff=4.0, central_decay=0.5, interaction_cutoff=5.0, interaction_decay=1.0)
CUT2 = dict(central_cutoff=3.0, central_decay=0.6, interaction_cutoff=4.0, interaction_decay=1.2)


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)
    from qml_amd import _lib

    _lib.lib()


def _rel(a, b):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)) if a.size else 0.0


def _acm_ok(got, ref, cut):
    assert got.shape == ref.shape
    if cut:
        assert np.array_equal(got == 0, ref == 0)
        scale = np.abs(ref).max(axis=1, keepdims=True)
        assert np.all(np.abs(got - ref) <= 1e-14 * scale)
    else:
        assert np.array_equal(bits(got), bits(ref))


# --------------------------------------------------------------------------------------------
# Atomic Coulomb matrices
# --------------------------------------------------------------------------------------------


@pytest.mark.parametrize("name,kw", [
    ("distance_sorted", dict(sorting="distance")),
    ("row_norm_sorted", dict(sorting="row-norm")),
    ("distance_sorted_with_cutoff", dict(sorting="distance", **CUT)),
    ("row_norm_sorted_with_cutoff", dict(sorting="row-norm", **CUT)),
])
def test_acm_reference_fixtures(qm7, acm_fixtures, name, kw):
    # tests/test_representations.py:92-132 through the single-molecule drop-in.
    from qml_amd.representations import generate_atomic_coulomb_matrix

    mols = qm7_molecules(qm7, qm7["rep_names"])
    got = np.concatenate([generate_atomic_coulomb_matrix(Z, R, 17, **kw) for Z, R, _ in mols])
    np.testing.assert_allclose(got, acm_fixtures[name], rtol=6e-10, atol=0)
    ref = np.concatenate([oracle.atomic_coulomb_matrix(Z, R, 17, **kw) for Z, R, _ in mols])
    _acm_ok(got, ref, "central_cutoff" in kw)


@pytest.mark.parametrize("sorting", ["row-norm", "distance"])
@pytest.mark.parametrize("cut", [False, True])
def test_acm_batch_matches_oracle(sorting, cut):
    from qml_amd.representations import generate_atomic_coulomb_matrices

    kw = CUT2 if cut else {}
    mols = synthetic.generate(600, 41, 1, 29)
    Z, R, off = mols.ragged()
    got = generate_atomic_coulomb_matrices(Z, R, off, 29, sorting, **kw).cpu().numpy()
    assert np.all(got[:, 435:] == 0)
    ref = np.concatenate([oracle.atomic_coulomb_matrix(*mols.molecule(m), 29, sorting, **kw)
                          for m in range(mols.nmol)])
    _acm_ok(got[:, :435], ref, cut)


def test_acm_reference_fortran_vectors(ref_local_vectors):
    from qml_amd.representations import generate_atomic_coulomb_matrices

    v = ref_local_vectors
    for sorting in ("row-norm", "distance"):
        tag = sorting.replace("-", "_")
        for suffix, kw in (("", {}), ("_cut", CUT2)):
            got = generate_atomic_coulomb_matrices(v["Z"], v["R"], v["offsets"], 23, sorting,
                                                   **kw).cpu().numpy()[:, :276]
            _acm_ok(got, v[f"acm_{tag}{suffix}"], bool(kw))


def test_acm_indices_and_errors(qm7):
    from qml_amd.representations import generate_atomic_coulomb_matrix

    for Z, R, _ in qm7_molecules(qm7, qm7["rep_names"][:4]):
        for sorting in ("row-norm", "distance"):
            full = generate_atomic_coulomb_matrix(Z, R, 17, sorting)
            sub = generate_atomic_coulomb_matrix(Z, R, 17, sorting, indices=[1, 2])
            assert np.array_equal(bits(sub), bits(full[1:3]))
            c = generate_atomic_coulomb_matrix(Z, R, 17, sorting, indices="C")
            assert np.array_equal(bits(c), bits(full[Z == 6]))
        assert generate_atomic_coulomb_matrix(Z, R, 17, indices="Cl").shape == (0, 0)
    with pytest.raises(ValueError):
        generate_atomic_coulomb_matrix(Z, R, 3, "distance")
    with pytest.raises(SystemExit):
        generate_atomic_coulomb_matrix(Z, R, 17, "unsorted")
    with pytest.raises(SystemExit):
        generate_atomic_coulomb_matrix(Z, R, 17, indices="Xx")


# --------------------------------------------------------------------------------------------
# Local kernels
# --------------------------------------------------------------------------------------------


@pytest.fixture(scope="module")
def qm7_local_gpu(qm7, local_fixtures):
    """Atomic CMs (row-norm, size 23) of the local-KRR split, generated on the GPU."""
    from qml_amd.representations import generate_atomic_coulomb_matrix

    tr, te = qm7_local_split(qm7, local_fixtures)

    def reps(mols):
        X = [generate_atomic_coulomb_matrix(Z, R, 23, "row-norm") for Z, R, _ in mols]
        return np.concatenate(X), np.array([len(x) for x in X]), np.array([y for *_, y in mols])

    return reps(tr), reps(te)


@pytest.mark.parametrize("kind,sigma,lam,mae", [("gaussian", 724.0, 10 ** -6.5, 19.0),
                                                ("laplacian", 10 ** 3.6, 1e-12, 8.7)])
def test_local_kernels_qm7(qm7_local_gpu, local_fixtures, cho_pair, kind, sigma, lam, mae):
    # tests/test_energy_krr_atomic_cmat.py:59-217 on the GPU path.
    from qml_amd.kernels import get_local_kernels_gaussian, get_local_kernels_laplacian

    fn = get_local_kernels_gaussian if kind == "gaussian" else get_local_kernels_laplacian
    (X, N, Y), (Xs, Ns, Ys) = qm7_local_gpu
    K = fn(X, X, N, N, [sigma])[0]
    assert np.array_equal(bits(K), bits(K.T))
    assert _rel(K, oracle.local_kernels(kind, X, X, N, N, [sigma])[0]) <= 1e-12
    K_ref = cho_pair[0] if kind == "gaussian" else local_fixtures["K_local_laplacian"]
    np.testing.assert_allclose(K, K_ref, rtol=1e-12, atol=0)
    Ks = fn(Xs, X, Ns, N, [sigma])[0]
    assert _rel(Ks, oracle.local_kernels(kind, Xs, X, Ns, N, [sigma])[0]) <= 1e-12
    bad = set(np.where(np.abs(Ks - local_fixtures[f"Ks_local_{kind}"]) > 1e-7)[0])
    assert len(bad) < 5  # the reference's own criterion (a row-norm tie, see the oracle test)
    K[np.diag_indices_from(K)] += lam
    alpha = oracle.cho_solve(K, Y)
    assert abs(np.mean(np.abs(Ys - Ks @ alpha)) - mae) < 1.0


@pytest.mark.parametrize("kind,expanded", [("gaussian", True), ("gaussian", False),
                                           ("laplacian", None)])
@pytest.mark.parametrize("nsig", [1, 3, 6])
def test_local_kernels_paths_and_sigmas(kind, expanded, nsig):
    from qml_amd import _lib
    from qml_amd.kernels import LocalSet, local_kernels_device

    rng = np.random.default_rng(nsig)
    na = rng.integers(1, 24, size=70)
    nb = rng.integers(1, 24, size=45)
    A = rng.random((na.sum(), 77)) * 20
    B = rng.random((nb.sum(), 77)) * 20
    sig = np.array([40.0, 90.0, 300.0, 55.0, 1000.0, 70.0])[:nsig]
    k = _lib.PAIR_GAUSSIAN if kind == "gaussian" else _lib.PAIR_LAPLACIAN
    Kd = local_kernels_device(k, LocalSet(A, na), LocalSet(B, nb), sig, expanded=expanded)
    got = np.transpose(Kd.cpu().numpy(), (0, 2, 1))
    ref = oracle.local_kernels(kind, A, B, na, nb, sig)
    assert got.shape == ref.shape == (nsig, 70, 45)
    assert _rel(got, ref) <= 1e-12


@pytest.mark.parametrize("kind", ["gaussian", "laplacian"])
def test_local_kernels_ragged_edges(kind):
    # Empty molecules, runs of one-atom molecules (the 16-molecule panel cap), and molecules
    # larger than a 128-atom tile (a panel alone, reduced over several tiles).
    from qml_amd.kernels import get_local_kernels_gaussian, get_local_kernels_laplacian

    fn = get_local_kernels_gaussian if kind == "gaussian" else get_local_kernels_laplacian
    rng = np.random.default_rng(5)
    na = np.array([0, 3, 0, 300, 1] + [1] * 40 + [0, 129, 7, 0])
    nb = np.array([2, 0, 140] + [1] * 35 + [5, 0])
    A = rng.random((na.sum(), 31)) * 5
    B = rng.random((nb.sum(), 31)) * 5
    got = fn(A, B, na, nb, [6.0, 20.0])
    ref = oracle.local_kernels(kind, A, B, na, nb, [6.0, 20.0])
    assert _rel(got, ref) <= 1e-12
    assert np.all(got[:, na == 0, :] == 0) and np.all(got[:, :, nb == 0] == 0)
    Ksym = fn(A, A, na, na, [6.0])[0]
    assert np.array_equal(bits(Ksym), bits(Ksym.T))
    assert _rel(Ksym, oracle.local_kernels(kind, A, A, na, na, [6.0])[0]) <= 1e-12


def test_local_kernels_large_additivity():
    # 3000 synthetic molecules (about 42k atoms): merging neighbouring molecules must sum their
    # kernel rows (a size-independent property), and a sample of entries matches the oracle.
    from qml_amd import _lib
    from qml_amd.kernels import LocalSet, local_kernels_device
    from qml_amd.representations import generate_atomic_coulomb_matrices

    mols = synthetic.generate(3000, 77, 5, 23)
    Z, R, off = mols.ragged()
    X = generate_atomic_coulomb_matrices(Z, R, off, 23, "row-norm")
    na = np.diff(off)
    S = LocalSet(X, na)
    K = local_kernels_device(_lib.PAIR_GAUSSIAN, S, S, [724.0], symmetric=True)[0]
    coarse = na.reshape(-1, 2).sum(axis=1)
    Sc = LocalSet(X, coarse)
    Kc = local_kernels_device(_lib.PAIR_GAUSSIAN, Sc, S, [724.0])[0]  # (nm, nm/2) as [b, a]
    merged = K.reshape(K.shape[0], -1, 2).sum(dim=2)
    assert float(((Kc - merged).abs() / merged.abs()).max()) <= 1e-12
    Xh = X.cpu().numpy()[:, :276]
    pick = np.arange(0, 3000, 97)
    rows = np.concatenate([np.arange(off[m], off[m + 1]) for m in pick])
    ref = oracle.local_kernels("gaussian", Xh[rows], Xh, na[pick], na, [724.0])[0]
    got = K.cpu().numpy()[:, pick].T
    assert _rel(got, ref) <= 1e-12


def test_atomic_kernel_wrappers(qm7_local_gpu):
    # qml/wrappers.py:33-97 over molecule objects equals the local kernels.
    from qml_amd.kernels import get_local_kernels_gaussian, get_local_kernels_laplacian
    from qml_amd.wrappers import get_atomic_kernels_gaussian, get_atomic_kernels_laplacian

    (X, N, _), (Xs, Ns, _) = qm7_local_gpu

    class Mol:
        def __init__(self, rep):
            self.representation, self.natoms = rep, rep.shape[0]

    split = np.cumsum(N)[:-1]
    mols = [Mol(r) for r in np.split(X, split)][:60]
    tests = [Mol(r) for r in np.split(Xs, np.cumsum(Ns)[:-1])][:30]
    Xa = np.concatenate([m.representation for m in mols])
    Na = np.array([m.natoms for m in mols])
    Xt = np.concatenate([m.representation for m in tests])
    Nt = np.array([m.natoms for m in tests])
    for wrap, local, sig in ((get_atomic_kernels_gaussian, get_local_kernels_gaussian, 724.0),
                             (get_atomic_kernels_laplacian, get_local_kernels_laplacian, 4000.0)):
        assert np.array_equal(bits(wrap(mols, mols, [sig])), bits(local(Xa, Xa, Na, Na, [sig])))
        assert np.array_equal(bits(wrap(tests, mols, [sig])), bits(local(Xt, Xa, Nt, Na, [sig])))


@pytest.mark.parametrize("sorting", ["row-norm", "distance"])
def test_acm_molecules_above_lds_capacity(sorting):
    """Molecules of more than 256 atoms (beyond the LDS kernel's capacity) in a batch with small
    ones: their rows come from the global-memory kernel and match the oracle like the rest.
    A large molecule here is 15 to 25 synthetic molecules laid out 12 A apart, so a 3 A central
    cut-off keeps every central atom's neighbourhood within size."""
    from qml_amd.representations import generate_atomic_coulomb_matrices

    parts = synthetic.generate(60, 43, 5, 23)
    rng = np.random.default_rng(8)
    Zs, Rs = [], []
    start = 0
    for k in (15, 25, 18):  # three large molecules
        Zl, Rl = [], []
        for j in range(k):
            Z, R = parts.molecule(start + j)
            Zl.append(Z)
            Rl.append(R + 12.0 * np.array([j % 5, (j // 5) % 5, j // 25], dtype=np.float64))
        start += k
        Zs.append(np.concatenate(Zl))
        Rs.append(np.concatenate(Rl))
    small = synthetic.generate(20, 44, 3, 29)
    for m in range(small.nmol):  # small molecules between and after them
        Z, R = small.molecule(m)
        pos = int(rng.integers(0, len(Zs) + 1))
        Zs.insert(pos, Z)
        Rs.insert(pos, R)
    nat = np.array([len(z) for z in Zs])
    assert nat.max() > 256 and nat.min() < 30
    off = np.zeros(len(Zs) + 1, np.int64)
    np.cumsum(nat, out=off[1:])
    kw = dict(central_cutoff=3.0, central_decay=0.5, interaction_cutoff=5.0, interaction_decay=1.0)
    size = 40
    got = generate_atomic_coulomb_matrices(np.concatenate(Zs).astype(np.float64),
                                           np.concatenate(Rs), off, size, sorting,
                                           **kw).cpu().numpy()
    d = size * (size + 1) // 2
    assert np.all(got[:, d:] == 0)
    ref = np.concatenate([oracle.atomic_coulomb_matrix(Z, R, size, sorting, **kw)
                          for Z, R in zip(Zs, Rs)])
    _acm_ok(got[:, :d], ref, True)
