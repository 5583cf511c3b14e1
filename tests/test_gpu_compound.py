"""End to end on the reference's on-disk format: QM7 XYZ files -> Compound -> GPU Coulomb
matrices -> Laplacian KRR (tests/test_energy_krr_cmat.py:54-113), and the atomic Coulomb
matrix through Compound (qml/compound.py:134-215). Bars: MAE < 6.0 as the reference test;
predictions within 1e-8 relative of the oracle's on the same representations; batched and
per-molecule Coulomb matrices bit-exact."""

import numpy as np
import pytest

import oracle
from conftest import bits, get_energies, write_xyz_files

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)
    from qml_amd import _lib

    _lib.lib()


def test_krr_cmat_end_to_end_from_xyz(tmp_path, xyz_texts):
    from qml_amd.compound import Compound, read_xyz_batch
    from qml_amd.kernels import laplacian_kernel
    from qml_amd.math import cho_solve
    from qml_amd.representations import generate_coulomb_matrices

    paths = write_xyz_files(xyz_texts, tmp_path)
    data = get_energies(bytes(xyz_texts["hof_qm7"]).decode())
    names = sorted(data.keys())[:1000]
    mols = []
    for xyz_file in names:
        mol = Compound(xyz=paths[xyz_file])
        mol.properties = data[xyz_file]
        mol.generate_coulomb_matrix(size=23, sorting="row-norm")
        mols.append(mol)
    # The batched reader and one launch for all molecules give the same matrices.
    Z, R, off = read_xyz_batch([paths[n] for n in names])
    Xb = generate_coulomb_matrices(Z, R, off, 23).cpu().numpy()[:, :276]
    assert np.array_equal(bits(Xb), bits(np.array([m.representation for m in mols])))

    np.random.seed(666)
    np.random.shuffle(mols)
    training, test = mols[:700], mols[-300:]
    X = np.array([mol.representation for mol in training])
    Xs = np.array([mol.representation for mol in test])
    Y = np.array([mol.properties for mol in training])
    Ys = np.array([mol.properties for mol in test])
    sigma, llambda = 10 ** 4.2, 10 ** -10.0
    K = laplacian_kernel(X, X, sigma)
    K[np.diag_indices_from(K)] += llambda
    alpha = cho_solve(K, Y)
    Ks = laplacian_kernel(X, Xs, sigma)
    Yss = np.dot(Ks.transpose(), alpha)
    mae = np.mean(np.abs(Ys - Yss))
    assert mae < 6.0
    _, pred = oracle.krr_fit_predict(X, Y, Xs, "laplacian", sigma, llambda)
    assert np.max(np.abs(Yss - pred) / np.abs(pred)) <= 1e-8


def test_compound_atomic_cm_default_row_norm(tmp_path, xyz_texts):
    from qml_amd.compound import Compound

    paths = write_xyz_files(xyz_texts, tmp_path)
    for name in list(paths)[:20]:
        c = Compound(xyz=paths[name])
        c.generate_atomic_coulomb_matrix(size=23)
        ref = oracle.atomic_coulomb_matrix(c.nuclear_charges, c.coordinates, 23, "row-norm")
        assert np.array_equal(bits(c.representation), bits(ref))
        c.generate_atomic_coulomb_matrix(size=23, sorting="distance", indices="H")
        ref = oracle.atomic_coulomb_matrix(c.nuclear_charges, c.coordinates, 23, "distance",
                                           indices="H")
        assert np.array_equal(bits(c.representation), bits(ref))
