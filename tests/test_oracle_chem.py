"""Pins of the oracle's RHF / CCSD / lambda / RDM / gradient pipeline (CPU).

  * RHF and CCSD energies of H2O/STO-3G at the geometry of T. D. Crawford's
    programming projects (#3, #5): E(SCF) = -74.942079928192,
    E(CCSD corr) = -0.070680088376, E_nuc = 8.002367061810450, and his
    one-electron integrals T(1,1) = 29.003199945539588, V(1,1) = -61.580595358149914;
  * the energy computed from the 1- and 2-RDMs equals E(CCSD);
  * the lambda residual vanishes at the converged lambda;
  * the analytic gradient (lambda + RDMs + Z-vector + Imat + 1e terms,
    restating ccsd_gpu/gpu_ccsd_grad.py) equals central finite differences of
    E(CCSD) (h = 1e-4 Bohr) to 1e-7 Eh/Bohr, sums to zero over atoms.
"""
import numpy as np
import pytest

from ccsd_gpu_amd.gto import M
from oracle import chem
from oracle import integrals as oi

CRAWFORD_H2O = ("O 0.000000000000 -0.143225816552 0.000000000000; "
                "H 1.638036840407 1.136548822547 -0.000000000000; "
                "H -1.638036840407 1.136548822547 -0.000000000000")


@pytest.fixture(scope="module")
def crawford():
    mol = M(CRAWFORD_H2O, basis="sto-3g", unit="Bohr")
    mf = chem.RHF(mol)
    mf.kernel()
    cc = chem.CCSD(mf)
    cc.kernel()
    return mol, mf, cc


def test_crawford_one_electron_integrals(crawford):
    mol, mf, _ = crawford
    assert abs(mol.energy_nuc() - 8.002367061810450) < 1e-11
    assert abs(oi.int1e(mol, "kin")[0, 0] - 29.003199945539588) < 1e-10
    assert abs(oi.int1e(mol, "nuc")[0, 0] - (-61.580595358149914)) < 1e-10
    assert np.abs(np.diag(oi.int1e(mol, "ovlp")) - 1).max() < 1e-12


def test_crawford_scf_and_ccsd_energies(crawford):
    _, mf, cc = crawford
    assert abs(mf.e_tot - (-74.942079928192)) < 1e-9
    assert abs(cc.e_corr - (-0.070680088376)) < 1e-9


def test_rdm_energy_equals_ccsd_energy(crawford):
    mol, mf, cc = crawford
    eris = cc.ao2mo()
    conv, l1, l2 = chem.solve_lambda(cc.t1, cc.t2, eris)
    assert conv
    d1 = chem.gamma1_intermediates(cc.t1, cc.t2, l1, l2)
    d2 = chem.gamma2_intermediates(cc.t1, cc.t2, l1, l2)
    dm1 = chem.make_rdm1_mo(d1)
    dm2 = chem.make_rdm2_mo(d1, d2)
    c = mf.mo_coeff
    h = c.T @ mf.get_hcore() @ c
    e = np.einsum("pq,qp", h, dm1) + 0.5 * np.einsum("pqrs,pqrs", eris.eri_mo, dm2) + mol.energy_nuc()
    assert abs(e - cc.e_tot) < 1e-9
    assert abs(np.trace(dm1) - mol.nelectron) < 1e-12
    np.testing.assert_allclose(dm1, dm1.T, atol=1e-12)
    # lambda is a fixed point of the lambda update
    imds = chem.make_intermediates(cc.t1, cc.t2, eris)
    n1, n2 = chem.update_lambda(cc.t1, cc.t2, l1, l2, eris, imds)
    assert np.abs(n1 - l1).max() < 1e-9 and np.abs(n2 - l2).max() < 1e-9


def _fd_gradient(mol, h=1e-4, atoms=None):
    base = mol.atom_coords()
    out = {}
    for a in (range(mol.natm) if atoms is None else atoms):
        for x in range(3):
            es = []
            for s in (1, -1):
                c = base.copy()
                c[a, x] += s * h
                m = mol.copy()
                m.set_geom_(c)
                es.append(chem.ccsd_total_energy(m)[0])
            out[(a, x)] = (es[0] - es[1]) / (2 * h)
    return out


@pytest.mark.parametrize("basis,atoms", [("sto-3g", None), ("cc-pvdz", [0, 2])])
def test_analytic_gradient_matches_finite_difference(basis, atoms):
    mol = M("O 0.0 -0.0757 0.0521; H 0.7572 0.5 -0.02; H -0.76 0.47 0.03", basis=basis)
    e0, mf, cc = chem.ccsd_total_energy(mol)
    conv, l1, l2 = chem.solve_lambda(cc.t1, cc.t2, cc.ao2mo())
    assert conv
    de = chem.grad_elec(cc, cc.t1, cc.t2, l1, l2) + chem.grad_nuc(mol)
    assert np.abs(de.sum(axis=0)).max() < 1e-9
    for (a, x), v in _fd_gradient(mol, atoms=atoms).items():
        assert abs(de[a, x] - v) < 1e-7, (a, x, de[a, x], v)


def _fd_gradient_frozen(mol, frozen, coords, h=1e-4):
    base = mol.atom_coords()
    out = {}
    for a, x in coords:
        es = []
        for s in (1, -1):
            c = base.copy()
            c[a, x] += s * h
            m = mol.copy()
            m.set_geom_(c)
            es.append(chem.ccsd_total_energy(m, frozen=frozen)[0])
        out[(a, x)] = (es[0] - es[1]) / (2 * h)
    return out


@pytest.mark.parametrize("basis,frozen,coords", [
    ("sto-3g", 1, [(0, 0), (0, 1), (1, 2), (2, 0)]),
    ("cc-pvdz", 1, [(0, 1), (1, 0)]),
    ("cc-pvdz", [0, 23], [(0, 2), (2, 1)]),          # frozen core and the highest virtual
])
def test_frozen_orbital_gradient_matches_finite_difference(basis, frozen, coords):
    """Frozen-orbital analytic gradient (the reference's dco / dfv response blocks,
    ccsd_gpu/gpu_ccsd_grad.py:510-518) against central differences of the frozen-core
    E(CCSD); without the dco block the cc-pVDZ gradient moves by 5.6e-5 Eh/Bohr."""
    mol = M("O 0.0 -0.0757 0.0521; H 0.7572 0.5 -0.02; H -0.76 0.47 0.03", basis=basis)
    e0, mf, cc = chem.ccsd_total_energy(mol, frozen=frozen)
    assert cc.t1.shape[0] == mol.nelectron // 2 - 1
    conv, l1, l2 = chem.solve_lambda(cc.t1, cc.t2, cc.ao2mo())
    assert conv
    de = chem.grad_elec(cc, cc.t1, cc.t2, l1, l2) + chem.grad_nuc(mol)
    assert np.abs(de.sum(axis=0)).max() < 1e-9
    for (a, x), v in _fd_gradient_frozen(mol, frozen, coords).items():
        assert abs(de[a, x] - v) < 1e-7, (a, x, de[a, x], v)
