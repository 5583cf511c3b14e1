"""XYZ reading and the Compound object (qml/compound.py:37-326) on the host.

Pinned by the reference's compound fixtures (tests/test_compound.py:27-46, the files copied into
tests/golden/xyz_texts.npz) and by an independent parse of the reference's QM7 files.
"""

import numpy as np
import pytest

from conftest import GOLDEN, write_xyz_files
from qml_amd.compound import Compound, read_xyz_batch


@pytest.mark.parametrize("key", ["compound_test_xyz", "compound_test_exyz"])
def test_compound_reference_fixtures(tmp_path, xyz_texts, key):
    path = tmp_path / "c.xyz"
    path.write_bytes(bytes(xyz_texts[key]))
    c = Compound(xyz=str(path))
    assert c.natoms == 5
    assert c.atomtypes == ["C", "Cl", "Br", "H", "H"]
    assert list(c.nuclear_charges) == [6, 17, 35, 1, 1]
    assert c.natypes == {"C": 1, "Cl": 1, "Br": 1, "H": 2}
    assert c.atomtype_indices["H"] == [3, 4]
    np.testing.assert_array_equal(c.coordinates[1], [-4.22806, 0.88165, 0.08802])
    assert c.name == str(path)


def test_qm7_files_parse_to_the_committed_arrays(tmp_path, xyz_texts, qm7):
    # An independent parse (numpy on the split text) of every QM7 file in the fixture.
    paths = write_xyz_files(xyz_texts, tmp_path)
    index = {str(n): i for i, n in enumerate(qm7["names"])}
    sym = {"H": 1, "C": 6, "N": 7, "O": 8, "S": 16}
    for name, path in list(paths.items())[:300]:
        c = Compound(xyz=path)
        lines = bytes(xyz_texts["texts"][list(paths).index(name)]).decode().splitlines()
        n = int(lines[0])
        body = [ln.split() for ln in lines[2:2 + n]]
        assert c.natoms == n
        assert list(c.nuclear_charges) == [sym[t[0]] for t in body]
        np.testing.assert_array_equal(c.coordinates, np.array([t[1:4] for t in body], float))
        i = index[name]
        a, b = qm7["offsets"][i], qm7["offsets"][i + 1]
        np.testing.assert_array_equal(c.coordinates, qm7["R"][a:b])
    Z, R, off = read_xyz_batch(list(paths.values()))
    assert off[-1] == qm7["offsets"][-1]
    order = [index[n] for n in paths]
    expect_R = np.concatenate([qm7["R"][qm7["offsets"][i]:qm7["offsets"][i + 1]] for i in order])
    np.testing.assert_array_equal(R, expect_R)


def test_short_atom_line_stops_the_parse(tmp_path):
    # qml/compound.py:316-317: a line with fewer than four fields ends the atom list.
    p = tmp_path / "short.xyz"
    p.write_text("3\ncomment\nO 0 0 0\nH 0.9 0 0\nH 0.1\n")
    c = Compound(xyz=str(p))
    assert c.natoms == 3 and c.atomtypes == ["O", "H"]
    Z, R, off = read_xyz_batch([str(p)])
    assert list(off) == [0, 2] and list(Z) == [8.0, 1.0]


def test_other_representations_are_not_built():
    c = Compound()
    assert np.isnan(c.natoms) and c.representation.size == 0
    for fn in (c.generate_bob, c.generate_eigenvalue_coulomb_matrix, c.generate_slatm,
               c.generate_fchl_representation):
        with pytest.raises(NotImplementedError):
            fn()
