"""NWChem-format basis files: the parser (general contractions, SP shells, Fortran
exponents), the bundled sets (shell structure, unit-normalised spherical AOs of every
bundled set), and molecules built from them and from a user file."""
import numpy as np
import pytest

from ccsd_gpu_amd.gto import M, basis_io
from oracle import integrals as oi

SAMPLE = """
# comment
BASIS "ao basis" PRINT
Li    S
   16.1195750     0.15432897
    2.9362007     0.53532814
    0.7946505     0.44463454
Li    SP
    0.6362897    -0.09996723    0.15591627
    0.1478601     0.39951283    0.60768372
    0.0480887     0.70011547    0.39195739
O    S
   11720.0D+00    0.000710   -0.000160
    1759.0D+00    0.005470   -0.001263
END
"""


def test_parser_blocks_sp_and_general_contractions():
    t = basis_io.parse_nwchem(SAMPLE)
    assert [s[0] for s in t["Li"]] == [0, 0, 1]
    l, e, c = t["Li"][1]
    assert e == [0.6362897, 0.1478601, 0.0480887] and c == [[-0.09996723, 0.39951283, 0.70011547]]
    assert t["Li"][2][2] == [[0.15591627, 0.60768372, 0.39195739]]
    l, e, c = t["O"][0]
    assert l == 0 and e == [11720.0, 1759.0] and len(c) == 2 and c[1] == [-0.000160, -0.001263]


def test_parser_rejects_ragged_rows():
    with pytest.raises(ValueError):
        basis_io.parse_nwchem("H S\n 1.0 0.5\n 2.0 0.3 0.1\n")


@pytest.mark.parametrize("name,el,shape", [
    ("cc-pvdz", "O", [0, 0, 1, 1, 2]), ("cc-pvtz", "O", [0, 0, 0, 1, 1, 1, 2, 2, 3]),
    ("cc-pvqz", "O", [0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 3, 3, 4]), ("def2-svp", "C", [0, 0, 0, 1, 1, 2]),
    ("sto-3g", "F", [0, 0, 1])])
def test_bundled_sets_shell_structure(name, el, shape):
    assert [s[0] for s in basis_io.load(name)[el]] == shape


@pytest.mark.parametrize("basis", ["def2-svp", "cc-pvqz", "cc-pvtz", "sto-3g", "cc-pvdz"])
def test_bundled_sets_give_unit_normalised_aos(basis):
    mol = M("O 0 0 0; H 0 0.757 0.587; H 0 -0.757 0.587", basis=basis)
    s = oi.int1e(mol, "ovlp")
    assert np.abs(np.diag(s) - 1).max() < 1e-10
    assert np.linalg.eigvalsh(s).min() > 1e-6          # no duplicated functions


def test_molecule_from_a_user_file(tmp_path):
    p = tmp_path / "mini.nw"
    p.write_text(SAMPLE.replace("Li", "H"))
    mol = M("H 0 0 0; H 0 0 1.4", basis=str(p), unit="Bohr")
    assert mol.nbas == 2 * 3 + 0 and mol.nao_nr() == 2 * (1 + 1 + 3)
