"""Pins of the CPU oracle (oracle/rys_oracle.c) for int2e / int2e_ip1.

Each check is chosen so that a plausible mistake fails it:
  * Rys tables vs 50-digit Gauss quadrature and vs the reference's own table
    (ccsd_gpu/cuda/rys_data_n9.json, evaluated into a fixture);
  * shell quartets vs an independent McMurchie-Davidson brute force (fixture
    and a live case) -- catches wrong recurrences, prefactors, normalization,
    derivative sign, general contractions;
  * Szabo-Ostlund H2/STO-3G textbook integrals and the (ss|ss) closed form;
  * 8-fold permutational symmetry, translational invariance of the derivative,
    and int2e_ip1 against central finite differences of int2e.
"""
import math
import os

import numpy as np
import pytest

from ccsd_gpu_amd.gto import M
from oracle import integrals as oi
from oracle import md_ref

from . import molecules

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_rys_oracle_matches_mpmath_and_reference_table():
    d = np.load(os.path.join(GOLDEN, "rys_golden.npz"))
    for n, T, mr, mw, rr, rw in zip(d["n"], d["T"], d["mp_r"], d["mp_w"], d["ref_r"], d["ref_w"]):
        r, w = oi.rys_roots(int(n), float(T))
        np.testing.assert_allclose(r, mr[:n], rtol=1e-13, atol=0)
        np.testing.assert_allclose(w, mw[:n], rtol=1e-13, atol=0)
        # the reference's table is itself accurate to ~1e-11 (n=1: 2e-10 from its closed form)
        np.testing.assert_allclose(r, rr[:n], rtol=3e-10, atol=0)
        np.testing.assert_allclose(w, rw[:n], rtol=3e-11, atol=0)


def test_rys_moment_identity():
    for n in (1, 2, 5, 9):
        for T in (0.0, 1e-5, 0.7, 3.3, 17.0, 44.9, 80.0):
            r, w = oi.rys_roots(n, T)
            for m in range(2 * n):
                fm = float(md_ref.boys(m, T))
                assert abs(np.dot(w, r ** m) - fm) <= 1e-13 * fm + 1e-300


def test_oracle_vs_mcmurchie_davidson_fixture():
    d = np.load(os.path.join(GOLDEN, "md_golden.npz"))
    mol = M(str(d["geom"]), basis="cc-pvdz", cart=True)
    loc = mol.ao_loc_nr(cart=True)
    for deriv in (0, 1):
        full = oi._unpack_kl(oi.int2e_s2kl_cart(mol, 0, mol.nbas, deriv), loc[-1])
        for q, shls in enumerate(d["quartets"]):
            sl = tuple(slice(loc[s], loc[s + 1]) for s in shls)
            mine = full[(slice(None),) + sl]
            ref = d[f"ip1_{q}"] if deriv else d[f"int2e_{q}"][None]
            np.testing.assert_allclose(mine, ref, atol=1e-13, rtol=0)


def test_oracle_vs_mcmurchie_davidson_live():
    mol = M("C 0 0 0; N 0.2 0.4 1.3", basis="cc-pvdz", cart=True)
    loc = mol.ao_loc_nr(cart=True)
    full = oi._unpack_kl(oi.int2e_s2kl_cart(mol, 0, mol.nbas, 1), loc[-1])
    for shls in [(4, 7, 2, 5), (0, 9, 3, 8)]:
        ref = md_ref.shell_quartet(mol, shls, 1)
        sl = tuple(slice(loc[s], loc[s + 1]) for s in shls)
        np.testing.assert_allclose(full[(slice(None),) + sl], ref, atol=1e-13, rtol=0)


def test_szabo_ostlund_h2():
    # Szabo & Ostlund, Modern Quantum Chemistry, H2/STO-3G at R = 1.4 bohr
    mol = M("H 0 0 0; H 0 0 1.4", basis="sto-3g", unit="Bohr")
    eri = oi.int2e_full(mol)
    assert abs(eri[0, 0, 0, 0] - 0.7746) < 5e-5
    assert abs(eri[0, 0, 1, 1] - 0.5697) < 5e-5
    assert abs(eri[1, 0, 1, 0] - 0.2970) < 5e-5
    assert abs(eri[1, 0, 0, 0] - 0.4441) < 5e-5


def test_ssss_closed_form():
    basis = {"H": [(0, [0.8], [[1.0]])], "He": [(0, [1.7], [[1.0]])]}
    mol = M("H 0 0 0; He 0.3 -0.2 1.1", basis=basis, unit="Bohr")
    eri = oi.int2e_full(mol)
    A, B = mol.atom_coords()
    a, b = 0.8, 1.7
    na = (2 * a / math.pi) ** 0.75
    nb = (2 * b / math.pi) ** 0.75
    p = q = a + b
    P = (a * A + b * B) / p
    kab = math.exp(-a * b / p * np.dot(A - B, A - B))
    ref = 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * kab * kab * float(md_ref.boys(0, 0.0)) * (na * nb) ** 2
    assert abs(eri[0, 1, 0, 1] - ref) < 1e-14
    # (aa|bb): P = A, Q = B
    T = (2 * a) * (2 * b) / (2 * a + 2 * b) * np.dot(A - B, A - B)
    ref2 = 2 * math.pi ** 2.5 / (4 * a * b * math.sqrt(2 * a + 2 * b)) * float(md_ref.boys(0, T)) * na * na * nb * nb
    assert abs(eri[0, 0, 1, 1] - ref2) < 1e-14
    _ = P


@pytest.mark.parametrize("cart", [True, False])
def test_eightfold_symmetry(cart):
    mol = molecules.h2o_ccpvdz(cart=cart)
    e = oi.int2e_full(mol)
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1), (3, 2, 1, 0)]:
        np.testing.assert_allclose(e, e.transpose(perm), atol=1e-14, rtol=0)


def test_translational_invariance_of_ip1():
    mol = M("O 0 0 0; H 0 0.96 0.1; H 0.93 -0.05 0.2", basis="cc-pvdz", cart=False)
    g = oi.ip1_full(mol)                      # (3, i, j, k, l), d/dr of function i
    tot = g + g.transpose(0, 2, 1, 3, 4) + g.transpose(0, 3, 4, 1, 2) + g.transpose(0, 4, 3, 2, 1)
    assert np.abs(tot).max() < 1e-12


def test_ip1_matches_finite_difference_of_int2e():
    mol = M("O 0 0 0; H 0 0.96 0.1; H 0.93 -0.05 0.2", basis="cc-pvdz", cart=False)
    ip1 = oi.ip1_full(mol)
    aoa = mol.ao_atom_index()
    h = 1e-4
    coords = mol.atom_coords()
    for atom, x in [(0, 2), (1, 0)]:
        m = (aoa == atom).astype(float)
        # dE/dA = -(sum of ip1 over the four functions that sit on A)
        ana = -(np.einsum("ijkl,i->ijkl", ip1[x], m)
                + np.einsum("jikl,j->ijkl", ip1[x], m)
                + np.einsum("klij,k->ijkl", ip1[x], m)
                + np.einsum("lkij,l->ijkl", ip1[x], m))
        plus = mol.copy()
        cp = coords.copy()
        cp[atom, x] += h
        plus.set_geom_(cp)
        minus = mol.copy()
        cm = coords.copy()
        cm[atom, x] -= h
        minus.set_geom_(cm)
        fd = (oi.int2e_full(plus) - oi.int2e_full(minus)) / (2 * h)
        assert np.abs(fd - ana).max() < 2e-8


def test_cart2sph_gives_unit_spherical_norms():
    # the spherical d/f functions must be normalized: (ii|jj)-independent check
    # through the overlap of the Cartesian functions built from ERIs is not
    # available, so use the analytic angular integrals of the monomials.
    from ccsd_gpu_amd.gto.mole import cart2sph_block, cart_components

    def ang(a, b):  # int x^a y^b z^c dOmega / (4 pi) for a monomial pair product
        e = [a[i] + b[i] for i in range(3)]
        if any(v % 2 for v in e):
            return 0.0
        num = math.prod(math.gamma((v + 1) / 2) for v in e)
        return 2 * num / math.gamma((sum(e) + 3) / 2) / (4 * math.pi)

    for l in (2, 3, 4):
        comps = cart_components(l)
        S = np.array([[ang(a, b) for b in comps] for a in comps]) * 4 * math.pi
        C = cart2sph_block(l)
        np.testing.assert_allclose(C.T @ S @ C, np.eye(2 * l + 1), atol=1e-12)
