"""Gaussian basis-set data (exponents and contraction coefficients).

Shells are listed per element as (l, exponents, coefficient columns). A shell
with several coefficient columns is a general contraction (nctr > 1), stored
the way PySCF stores it: one _bas row with NCTR_OF = ncolumns.

The data live in NWChem-format files under gto/basis/ (read by basis_io):
sto-3g (H, C, N, O, F), cc-pvdz (H, C, N, O), cc-pvtz (H, O), cc-pvqz (H, O),
def2-svp (H, C, N, O). A path to any NWChem-format file is accepted as a basis
name too. PySCF's bundled copies of these sets are not available in this
environment, so whether they agree digit-for-digit with PySCF 2.12.1 is
unverified; every test here uses this data on both sides of a comparison.
"""
import os

from . import basis_io

NUC_CHARGE = {"H": 1, "He": 2, "C": 6, "N": 7, "O": 8, "F": 9}


def get_basis(name, element):
    """Shells of `element` in the basis set `name` (a bundled set or an NWChem file)."""
    table = basis_io.load_file(name) if os.path.exists(str(name)) else basis_io.load(name)
    if element not in table:
        raise KeyError(f"basis set {name!r} has no data for element {element!r}")
    return table[element]
