import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libqmlhip.so)")


@pytest.fixture(scope="session")
def qm7():
    z = np.load(os.path.join(GOLDEN, "qm7_subset.npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def ref_vectors():
    z = np.load(os.path.join(GOLDEN, "ref_vectors.npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def acm_fixtures():
    z = np.load(os.path.join(GOLDEN, "atomic_cm_fixtures.npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def local_fixtures():
    z = np.load(os.path.join(GOLDEN, "local_kernel_fixtures.npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def ref_local_vectors():
    z = np.load(os.path.join(GOLDEN, "ref_local_vectors.npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def xyz_texts():
    z = np.load(os.path.join(GOLDEN, "xyz_texts.npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


def write_xyz_files(xyz_texts, directory):
    """Write the committed QM7 XYZ texts back to files; returns {name: path}."""
    paths = {}
    for name, text in zip(xyz_texts["names"], xyz_texts["texts"]):
        path = os.path.join(str(directory), str(name))
        with open(path, "wb") as f:
            f.write(bytes(text))
        paths[str(name)] = path
    return paths


def get_energies(text):
    """tests/test_energy_krr_cmat.py:36-53: {xyz name: heat of formation} from hof_qm7.txt."""
    energies = dict()
    for line in text.splitlines():
        tokens = line.split()
        if tokens:
            energies[tokens[0]] = float(tokens[1])
    return energies


@pytest.fixture(scope="session")
def ref_kernels_ext():
    z = np.load(os.path.join(GOLDEN, "ref_kernels_ext.npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def cho_pair():
    z = np.load(os.path.join(GOLDEN, "cho_solve_pair.npz"), allow_pickle=False)
    return z["K"], z["y"]


def qm7_molecules(qm7, names):
    index = {str(n): i for i, n in enumerate(qm7["names"])}
    out = []
    for name in names:
        i = index[str(name)]
        a, b = qm7["offsets"][i], qm7["offsets"][i + 1]
        out.append((qm7["Z"][a:b], qm7["R"][a:b], qm7["hof"][i]))
    return out


def bits(x):
    return np.ascontiguousarray(x, dtype=np.float64).view(np.uint64)


def qm7_local_split(qm7, local_fixtures):
    """tests/test_energy_krr_atomic_cmat.py:62-98: the first 1000 QM7 names (sorted), shuffled
    by np.random.seed(666) (the committed order), 200 training and the last 100 test."""
    mols = qm7_molecules(qm7, qm7["krr_names"])
    mols = [mols[i] for i in local_fixtures["order"]]
    return mols[:200], mols[-100:]
