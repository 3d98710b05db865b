"""The memory-lean oracle (oracle/chem_large.py, which writes the (H2O)8 gradient
fixture) against the full-tensor oracle (oracle/chem.py, pinned by finite
differences) on small molecules: blocked ERIs, RHF / CCSD energies and the gradient
from the reference's block loop."""
import numpy as np
import pytest

from ccsd_gpu_amd.gto import M
from oracle import chem, chem_large
from oracle import integrals as oi

GEOM = "O 0.0 -0.0757 0.0521; H 0.7572 0.5 -0.02; H -0.76 0.47 0.03"


@pytest.mark.parametrize("basis,cart", [("cc-pvdz", False), ("sto-3g", True)])
def test_blocked_eris_equal_full(basis, cart):
    mol = M(GEOM, basis=basis, cart=cart)
    assert np.abs(chem_large.int2e_native_blocked(mol) - oi.int2e_full(mol)).max() < 1e-13


def test_transform4_matches_einsum():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((5, 5, 5, 5))
    c = rng.standard_normal((5, 3))
    ref = np.einsum("pqrs,pi,qj,rk,sl->ijkl", x, c, c, c, c)
    assert np.abs(chem_large.transform4(x, c) - ref).max() < 1e-12


@pytest.mark.parametrize("basis,frozen", [("cc-pvdz", None), ("cc-pvdz", 1), ("sto-3g", None)])
def test_lean_gradient_equals_full_oracle(basis, frozen):
    mol = M(GEOM, basis=basis)
    e_hf, e_corr, de = chem_large.ccsd_gradient(mol, frozen=frozen, log=lambda *a: None)
    e_tot, mf, cc = chem.ccsd_total_energy(mol, frozen=frozen)
    conv, l1, l2 = chem.solve_lambda(cc.t1, cc.t2, cc.ao2mo())
    ref = chem.grad_elec(cc, cc.t1, cc.t2, l1, l2) + chem.grad_nuc(mol)
    assert abs(e_hf - mf.e_tot) < 1e-10 and abs(e_corr - cc.e_corr) < 1e-10
    assert np.abs(de - ref).max() < 1e-9, np.abs(de - ref).max()
