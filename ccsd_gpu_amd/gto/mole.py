"""Molecule container in libcint layout (the `mol` the reference's API takes).

The reference's integral wrappers and gradient code read a PySCF `Mole`:
  * `_atm` rows of 6 int32 slots, coordinates at `_env[_atm[a, PTR_COORD]]`
    (ccsd_gpu/cuda/int2e_ip1_module.py:149-151);
  * `_bas` rows of 8 int32 slots ATOM_OF=0, ANG_OF=1, NPRIM_OF=2, NCTR_OF=3,
    PTR_EXP=5, PTR_COEFF=6 (int2e_ip1_module.py:141-146,
    ccsd_gpu/cuda/int2e_ip1_gpu.py:28-33); the coefficients of a shell are
    stored [ctr][prim] (int2e_ip1_gpu.py:70);
  * `ao_loc_nr()`, `cart2sph_coeff()`, `offset_nr_by_atom()`, `cart`, `nbas`,
    `natm` (int2e_ip1_gpu.py:243-245, ccsd_gpu/gpu_ccsd_grad.py:386).
This module provides the same attributes for the basis sets in basis_data.py.

Conventions (PySCF's, restated from its public definitions):
  * radial normalization: every contracted radial function R(r) satisfies
    int R^2 r^2 dr = 1, with primitive factor 1/sqrt(int r^(2l+2) e^(-2a r^2) dr);
  * Cartesian s and p functions carry the angular factors 1/(2 sqrt(pi)) and
    sqrt(3/(4 pi)) inside the integrals, Cartesian l>=2 functions carry none
    (int2e_ip1_module.py:170-184);
  * spherical functions are real solid harmonics without the Condon-Shortley
    phase, ordered m = -l..l, except p which stays (x, y, z).
"""
import math

import numpy as np

from .basis_data import NUC_CHARGE, get_basis

BOHR = 0.52917721092  # Angstrom per Bohr

ATM_SLOTS = 6
CHARGE_OF = 0
PTR_COORD = 1
BAS_SLOTS = 8
ATOM_OF = 0
ANG_OF = 1
NPRIM_OF = 2
NCTR_OF = 3
KAPPA_OF = 4
PTR_EXP = 5
PTR_COEFF = 6
PTR_ENV_START = 20


def gaussian_int(n, alpha):
    """int_0^inf x^n exp(-alpha x^2) dx."""
    n1 = (n + 1) * 0.5
    return math.gamma(n1) / (2.0 * alpha ** n1)


def gto_norm(l, expnt):
    """Radial normalization of a primitive r^l exp(-a r^2)."""
    return 1.0 / math.sqrt(gaussian_int(l * 2 + 2, 2 * expnt))


def normalized_contraction(l, exps, coeffs):
    """Coefficients (nctr, nprim) with primitive and contracted normalization."""
    es = np.asarray(exps, dtype=float)
    cs = np.asarray(coeffs, dtype=float).reshape(-1, len(es))     # (nctr, nprim)
    cs = cs * np.array([gto_norm(l, e) for e in es])[None, :]
    ee = es[:, None] + es[None, :]
    ee = np.vectorize(lambda a: gaussian_int(l * 2 + 2, a))(ee)
    s = np.einsum("ip,pq,iq->i", cs, ee, cs)
    return cs / np.sqrt(s)[:, None]


def ncart(l):
    return (l + 1) * (l + 2) // 2


def nsph(l):
    return 2 * l + 1


def cart_components(l):
    """(lx, ly, lz) of the Cartesian functions of angular momentum l, PySCF order."""
    out = []
    for lx in range(l, -1, -1):
        for ly in range(l - lx, -1, -1):
            out.append((lx, ly, l - lx - ly))
    return out


def _real_harmonic_coeffs(l):
    """(2l+1, ncart) coefficients of r^l Y_lm (real, no Condon-Shortley phase)."""
    from scipy.special import lpmv
    rng = np.random.default_rng(1234 + l)
    npts = 4 * ncart(l) + 16
    v = rng.normal(size=(npts, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    x, y, z = v.T
    theta = np.arccos(np.clip(z, -1, 1))
    phi = np.arctan2(y, x)
    comps = cart_components(l)
    A = np.stack([x ** a * y ** b * z ** c for a, b, c in comps], axis=1)
    rows = []
    for m in range(-l, l + 1):
        am = abs(m)
        norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
        plm = lpmv(am, l, np.cos(theta)) * (-1) ** am      # remove Condon-Shortley phase
        if m > 0:
            f = math.sqrt(2) * norm * plm * np.cos(am * phi)
        elif m < 0:
            f = math.sqrt(2) * norm * plm * np.sin(am * phi)
        else:
            f = norm * plm
        c, *_ = np.linalg.lstsq(A, f, rcond=None)
        c[np.abs(c) < 1e-13] = 0.0
        rows.append(c)
    return np.array(rows)


_C2S_CACHE = {}


def cart2sph_block(l):
    """(ncart(l), nsph(l)) block of the Cartesian -> spherical transformation."""
    if l not in _C2S_CACHE:
        if l == 0:
            c = np.ones((1, 1))
        elif l == 1:
            c = np.eye(3)
        else:
            c = _real_harmonic_coeffs(l).T
        _C2S_CACHE[l] = c
    return _C2S_CACHE[l]


def _parse_atoms(atom):
    if isinstance(atom, str):
        items = []
        for line in atom.replace(";", "\n").splitlines():
            line = line.strip()
            if not line:
                continue
            p = line.split()
            items.append((p[0], [float(p[1]), float(p[2]), float(p[3])]))
        return items
    return [(a[0], list(map(float, a[1]))) for a in atom]


class Mole:
    """Closed-shell molecule with libcint-layout basis arrays."""

    def __init__(self, atom, basis="cc-pvdz", cart=False, unit="Angstrom", charge=0, verbose=0):
        self.atom = _parse_atoms(atom)
        self.basis = basis
        self.cart = bool(cart)
        self.unit = unit
        self.charge = charge
        self.verbose = verbose
        self.spin = 0
        self.build()

    # -- construction -----------------------------------------------------
    def build(self):
        scale = 1.0 / BOHR if self.unit.lower().startswith("a") else 1.0
        env = [0.0] * PTR_ENV_START
        atm = []
        bas = []
        self._basis_by_atom = []
        for ia, (sym, xyz) in enumerate(self.atom):
            ptr = len(env)
            env.extend([c * scale for c in xyz])
            atm.append([NUC_CHARGE[sym], ptr, 1, ptr + 3, 0, 0])
            env.append(0.0)  # zeta slot (point nucleus)
        for ia, (sym, xyz) in enumerate(self.atom):
            bset = self.basis[sym] if isinstance(self.basis, dict) else get_basis(self.basis, sym)
            shells = []
            for l, exps, coefs in bset:
                cs = normalized_contraction(l, exps, coefs)
                pexp = len(env)
                env.extend(list(map(float, exps)))
                pcoef = len(env)
                env.extend(cs.ravel().tolist())
                bas.append([ia, l, len(exps), cs.shape[0], 0, pexp, pcoef, 0])
                shells.append(len(bas) - 1)
            self._basis_by_atom.append(shells)
        self._atm = np.asarray(atm, dtype=np.int32).reshape(-1, ATM_SLOTS)
        self._bas = np.asarray(bas, dtype=np.int32).reshape(-1, BAS_SLOTS)
        self._env = np.asarray(env, dtype=np.float64)
        return self

    def copy(self):
        import copy
        new = copy.copy(self)
        new._atm = self._atm.copy()
        new._bas = self._bas.copy()
        new._env = self._env.copy()
        new.atom = [(s, list(x)) for s, x in self.atom]
        return new

    def set_geom_(self, coords_bohr):
        """Move the nuclei (coordinates in Bohr); basis data is unchanged."""
        coords_bohr = np.asarray(coords_bohr, dtype=float).reshape(self.natm, 3)
        for ia in range(self.natm):
            p = self._atm[ia, PTR_COORD]
            self._env[p:p + 3] = coords_bohr[ia]
        s = BOHR if self.unit.lower().startswith("a") else 1.0
        self.atom = [(sym, list(coords_bohr[i] * s)) for i, (sym, _) in enumerate(self.atom)]
        return self

    # -- sizes ------------------------------------------------------------
    @property
    def natm(self):
        return self._atm.shape[0]

    @property
    def nbas(self):
        return self._bas.shape[0]

    @property
    def nelectron(self):
        return int(self._atm[:, CHARGE_OF].sum()) - self.charge

    @property
    def nao(self):
        return self.nao_nr()

    def nao_nr(self, cart=None):
        return int(self.ao_loc_nr(cart)[-1])

    def nao_cart(self):
        return self.nao_nr(cart=True)

    def ao_loc_nr(self, cart=None):
        cart = self.cart if cart is None else cart
        l = self._bas[:, ANG_OF]
        dims = ((l + 1) * (l + 2) // 2 if cart else 2 * l + 1) * self._bas[:, NCTR_OF]
        return np.concatenate([[0], np.cumsum(dims)]).astype(np.int32)

    def bas_angular(self, i):
        return int(self._bas[i, ANG_OF])

    def bas_atom(self, i):
        return int(self._bas[i, ATOM_OF])

    def bas_nctr(self, i):
        return int(self._bas[i, NCTR_OF])

    def bas_nprim(self, i):
        return int(self._bas[i, NPRIM_OF])

    def bas_exp(self, i):
        p = self._bas[i, PTR_EXP]
        return self._env[p:p + self._bas[i, NPRIM_OF]].copy()

    def bas_ctr_coeff(self, i):
        """(nprim, nctr) normalized contraction coefficients."""
        p = self._bas[i, PTR_COEFF]
        nprim, nctr = self._bas[i, NPRIM_OF], self._bas[i, NCTR_OF]
        return self._env[p:p + nprim * nctr].reshape(nctr, nprim).T.copy()

    def atom_symbol(self, ia):
        return self.atom[ia][0]

    def atom_charges(self):
        return self._atm[:, CHARGE_OF].astype(float)

    def atom_coords(self):
        ptr = self._atm[:, PTR_COORD]
        return np.stack([self._env[p:p + 3] for p in ptr])

    def energy_nuc(self):
        z = self.atom_charges()
        r = self.atom_coords()
        e = 0.0
        for i in range(self.natm):
            for j in range(i):
                e += z[i] * z[j] / np.linalg.norm(r[i] - r[j])
        return e

    def offset_nr_by_atom(self):
        """(natm, 4): shell range and AO range of every atom."""
        ao_loc = self.ao_loc_nr()
        out = np.zeros((self.natm, 4), dtype=int)
        atoms = self._bas[:, ATOM_OF]
        for ia in range(self.natm):
            sh = np.where(atoms == ia)[0]
            if sh.size == 0:
                out[ia] = [0, 0, 0, 0] if ia == 0 else [out[ia - 1, 1]] * 2 + [out[ia - 1, 3]] * 2
                continue
            s0, s1 = sh[0], sh[-1] + 1
            out[ia] = [s0, s1, ao_loc[s0], ao_loc[s1]]
        return out

    def cart2sph_coeff(self):
        """(nao_cart, nao_sph) block-diagonal transformation matrix."""
        lc = self.ao_loc_nr(cart=True)
        ls = self.ao_loc_nr(cart=False)
        c = np.zeros((lc[-1], ls[-1]))
        for ib in range(self.nbas):
            l = self.bas_angular(ib)
            blk = cart2sph_block(l)
            nc, ns = blk.shape
            for ic in range(self.bas_nctr(ib)):
                r0 = lc[ib] + ic * nc
                c0 = ls[ib] + ic * ns
                c[r0:r0 + nc, c0:c0 + ns] = blk
        return c

    def ao_atom_index(self):
        """Atom of every AO (native basis)."""
        ao_loc = self.ao_loc_nr()
        out = np.empty(ao_loc[-1], dtype=np.int64)
        for ib in range(self.nbas):
            out[ao_loc[ib]:ao_loc[ib + 1]] = self.bas_atom(ib)
        return out


def M(atom, basis="cc-pvdz", cart=False, unit="Angstrom", charge=0, verbose=0):
    """Build a Mole (mirrors pyscf.gto.M for the arguments used on this path)."""
    return Mole(atom, basis=basis, cart=cart, unit=unit, charge=charge, verbose=verbose)
