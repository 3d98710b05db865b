"""Closed-shell Hartree-Fock on the GPU (producer of mo_coeff / mo_energy for the CCSD path).

The reference runs PySCF's RHF on the CPU (examples/h2o_ccsd_grad.py:118-121);
here the integrals come from the HIP kernels and the SCF iterations run on
the device. All AO-side work uses the Cartesian AO basis; C = cart2sph (nao_c,
nao_s) maps to the spherical AOs where the molecule uses them:
    J_s = C^T J_c(C D_s C^T) C,  K likewise.
The object exposes the attributes the reference's code reads from a PySCF
RHF object (mol, mo_coeff, mo_energy, mo_occ, e_tot, make_rdm1, get_ovlp,
get_hcore, get_veff, energy_nuc).
"""
import numpy as np
import torch

from .hip.int1e_gpu import int1e
from .hip.int2e_gpu import compute_int2e_gpu

DEV = "cuda"


def unpack_s2kl(packed):
    n = packed.shape[0]
    k, l = torch.tril_indices(n, n, device=packed.device)
    full = torch.empty((n, n, n, n), dtype=packed.dtype, device=packed.device)
    full[:, :, k, l] = packed
    full[:, :, l, k] = packed
    return full


class AOIntegrals:
    """Device-resident AO integrals of one geometry (Cartesian ERIs, native 1e)."""

    def __init__(self, mol):
        self.mol = mol
        self.cart = mol.cart
        self.c2s = torch.as_tensor(mol.cart2sph_coeff() if not mol.cart else np.eye(mol.nao_cart()),
                                   dtype=torch.float64, device=DEV)
        mol_c = mol.copy()
        mol_c.cart = True
        self.mol_cart = mol_c
        self.eri_p = compute_int2e_gpu(mol_c, 0, mol.nbas)                  # (nc, nc, nc(nc+1)/2), s2kl
        self.eri_c = unpack_s2kl(self.eri_p)                                 # (nc, nc, nc, nc)
        self.S = int1e(mol, "ovlp")
        self.T = int1e(mol, "kin")
        self.V = int1e(mol, "nuc")

    def to_cart(self, m):
        return self.c2s @ m @ self.c2s.T

    def to_native(self, m):
        return self.c2s.T @ m @ self.c2s

    def get_jk(self, dm, hermi=1):
        """J, K of a (nao, nao) density in the native basis (two GEMVs over HBM-resident
        ERIs). hermi=1 (PySCF's convention: dm symmetric) reads the pair-packed forms,
        half the bytes of each full one twice over:
            J_(ij) = sum_{k>=l} (ij|kl) w_kl D_kl,
            K_(ik) = sum_{j>=l} [(ij|kl) + (il|kj)] (w_jl / 2) D_jl,   w = 2 off the diagonal, 1 on it,
        both for i >= j (i >= k) and mirrored."""
        nc = self.eri_c.shape[0]
        dc = self.to_cart(dm)
        if hermi == 1:
            jpp, kpp, (a, b) = self._jk_packed()
            dk = dc[a, b] * torch.where(a == b, 0.5, 1.0).to(dc.dtype)
            jp, kp = jpp @ (2.0 * dk), kpp @ dk
            j = torch.empty((nc, nc), dtype=dc.dtype, device=dc.device)
            k = torch.empty_like(j)
            j[a, b] = jp
            j[b, a] = jp
            k[a, b] = kp
            k[b, a] = kp
            return self.to_native(j), self.to_native(k)
        if getattr(self, "_eri_k", None) is None:
            self._eri_k = self.eri_c.permute(0, 2, 1, 3).contiguous()    # (ik|jl) as [i, j, k, l]
        j = (self.eri_c.reshape(nc * nc, nc * nc) @ dc.reshape(-1)).reshape(nc, nc)
        k = (self._eri_k.reshape(nc * nc, nc * nc) @ dc.reshape(-1)).reshape(nc, nc)
        return self.to_native(j), self.to_native(k)

    def _jk_packed(self):
        """(npair, npair) matrices of the symmetric-density J and K (built once)."""
        if getattr(self, "_jkpp", None) is None:
            nc = self.eri_c.shape[0]
            a, b = torch.tril_indices(nc, nc, device=self.eri_c.device)
            jpp = self.eri_p.reshape(nc * nc, -1)[a * nc + b].contiguous()     # (ij|kl), i >= j, k >= l
            t = self.eri_c.permute(0, 2, 1, 3)[a, b]                          # [(ik)][j][l] = (ij|kl)
            t += t.transpose(1, 2).clone()                                      # + (il|kj)
            kpp = t[:, a, b].contiguous()
            del t
            self._jkpp = (jpp, kpp, (a, b))
        return self._jkpp


# AO integrals of molecules that arrive without this build's RHF object (PySCF-like
# mf / mycc objects): built once per geometry and basis, the latest one kept
_FOREIGN_INTS = {}


def _mol_key(mol):
    return (np.asarray(mol._atm).tobytes(), np.asarray(mol._bas).tobytes(), np.asarray(mol._env).tobytes(),
            bool(mol.cart))


def get_ints(mf=None, mol=None):
    """Device AO integrals for an RHF-like object: its own when it is this build's RHF,
    otherwise computed here from `mf.mol` (a PySCF-like object exposes mol, mo_coeff,
    mo_occ, mo_energy and nothing device-side). Only the most recent foreign geometry
    is kept (the Cartesian ERI tensor is nao_c^4 doubles)."""
    own = getattr(mf, "_ints", None) if mf is not None else None
    if callable(own):
        return own()
    mol = mf.mol if mol is None else mol
    key = _mol_key(mol)
    ints = _FOREIGN_INTS.get(key)
    if ints is None:
        _FOREIGN_INTS.clear()
        ints = AOIntegrals(mol)
        _FOREIGN_INTS[key] = ints
    return ints


def grad_nuc(mol):
    """Nuclear-repulsion gradient dE_nn/dR (natm, 3): what PySCF's rhf gradient object's
    grad_nuc returns (the reference adds it in compute_ml_ccsd_forces_gpu,
    ccsd_gpu/gpu_ccsd_grad.py:736-737)."""
    z = np.asarray(mol.atom_charges(), dtype=np.float64)
    r = np.asarray(mol.atom_coords(), dtype=np.float64)
    d = r[:, None, :] - r[None, :, :]
    dist = np.linalg.norm(d, axis=-1)
    np.fill_diagonal(dist, np.inf)
    return -np.einsum("a,b,abx->ax", z, z, d / dist[..., None] ** 3)


class RHF:
    def __init__(self, mol, ints=None):
        self.mol = mol
        self.ints = ints
        self.conv_tol = 1e-12
        self.conv_tol_grad = 1e-10
        self.max_cycle = 200
        self.converged = False

    def _ints(self):
        if self.ints is None:
            self.ints = AOIntegrals(self.mol)
        return self.ints

    # PySCF-like accessors (numpy outputs) ----------------------------------
    def get_ovlp(self, mol=None):
        return self._ints().S.cpu().numpy()

    def get_hcore(self, mol=None):
        ints = self._ints()
        return (ints.T + ints.V).cpu().numpy()

    def get_veff(self, mol=None, dm=None):
        d = torch.as_tensor(np.asarray(dm), dtype=torch.float64, device=DEV)
        j, k = self._ints().get_jk(d)
        return (j - 0.5 * k).cpu().numpy()

    def energy_nuc(self):
        return self.mol.energy_nuc()

    def make_rdm1(self, mo_coeff=None, mo_occ=None):
        mo_coeff = self.mo_coeff if mo_coeff is None else mo_coeff
        mo_occ = self.mo_occ if mo_occ is None else mo_occ
        c = mo_coeff[:, mo_occ > 0]
        return (c * mo_occ[mo_occ > 0]) @ c.T

    def kernel(self):
        ints = self._ints()
        S, h = ints.S, ints.T + ints.V
        w, v = torch.linalg.eigh(S)
        X = v / torch.sqrt(w)
        nocc = self.mol.nelectron // 2
        nmo = X.shape[1]

        def diag(f):
            e, c = torch.linalg.eigh(X.T @ f @ X)
            return e, X @ c

        def density(c):
            return 2.0 * c[:, :nocc] @ c[:, :nocc].T

        def fock(dm):
            j, k = ints.get_jk(dm)
            return h + j - 0.5 * k

        e, c = diag(h)
        dm = density(c)
        fs, errs = [], []
        e_old = 0.0
        for it in range(self.max_cycle):
            f = fock(dm)
            err = f @ dm @ S - S @ dm @ f
            fs.append(f)
            errs.append(err)
            fs, errs = fs[-8:], errs[-8:]
            n = len(fs)
            B = -torch.ones((n + 1, n + 1), dtype=torch.float64, device=DEV)
            B[n, n] = 0
            for i in range(n):
                for j in range(n):
                    B[i, j] = torch.sum(errs[i] * errs[j])
            rhs = torch.zeros(n + 1, dtype=torch.float64, device=DEV)
            rhs[n] = -1
            coef = torch.linalg.lstsq(B.cpu(), rhs.cpu().unsqueeze(1)).solution[:n, 0].to(DEV)
            fd = sum(coef[i] * fs[i] for i in range(n))
            e, c = diag(fd)
            dm = density(c)
            f = fock(dm)
            etot = float(0.5 * torch.sum(dm * (h + f))) + self.energy_nuc()
            gnorm = float(err.abs().max())
            if abs(etot - e_old) < self.conv_tol and gnorm < self.conv_tol_grad:
                self.converged = True
                break
            e_old = etot
        e, c = diag(f)
        occ = np.zeros(nmo)
        occ[:nocc] = 2
        self.mo_energy = e.cpu().numpy()
        self.mo_coeff = c.cpu().numpy()
        self.mo_occ = occ
        dm = density(c)
        f = fock(dm)
        self.e_tot = float(0.5 * torch.sum(dm * (h + f))) + self.energy_nuc()
        return self.e_tot

    def nuc_grad_method(self):
        return _RHFGradHelpers(self)


class _RHFGradHelpers:
    """grad_nuc of PySCF's rhf gradient object (the one method the reference calls)."""

    def __init__(self, mf):
        self.mf = mf

    def grad_nuc(self, mol=None):
        return grad_nuc(self.mf.mol if mol is None else mol)
