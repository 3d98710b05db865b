"""Many molecules of uneven size in one basis plan: batched, Schwarz-screened
fused ERI-derivative contractions.

    b = MoleculeBatch(mols)                # one plan over all shells of all molecules
    b.screen(1e-12)                        # skip quartets with q_ij q_kl < 1e-12
    de = b.grad2e_separable(dms)           # list of (natm_m, 3) per molecule

The arrays are concatenated the way PySCF's `gto.conc_env` does it (atom
indices and env offsets made global), every shell pair and quartet stays inside
its molecule, and each angular-momentum class runs as one launch over the
whole batch: the quartets of all molecules of that class form one work list
(most expensive chunks first). The reference evaluates one molecule per call
and one launch per i-shell (ccsd_gpu/cuda/int2e_ip1_gpu.py:199-214).

The two-particle density is the separable form of two AO density matrices
D and P (P = D by default),

    Gamma_ijkl = D_ij P_kl + P_ij D_kl - (D_il P_kj + D_ik P_lj + P_il D_jk + P_ik D_jl) / 4,

evaluated inside the kernel from D and P, so no four-index array exists. It is
the form of the reference's `vhf1` J/K terms contracted with `dm1p`
(ccsd_gpu/gpu_ccsd_grad.py:469-491); with D = P = the SCF density it is the
two-electron part of an RHF gradient.
"""
import ctypes

import numpy as np
import torch

from . import _lib

ATOM_OF, PTR_EXP, PTR_COEFF = 0, 5, 6
PTR_COORD, PTR_ZETA = 1, 3


def conc_env(mols):
    """(atm, bas, env, mol_nbas) of all molecules, atom and env offsets made global."""
    atms, bass, envs = [], [], []
    natm = nenv = 0
    for m in mols:
        atm = np.array(m._atm, dtype=np.int32, copy=True)
        bas = np.array(m._bas, dtype=np.int32, copy=True)
        atm[:, PTR_COORD] += nenv
        atm[:, PTR_ZETA] += nenv
        bas[:, ATOM_OF] += natm
        bas[:, PTR_EXP] += nenv
        bas[:, PTR_COEFF] += nenv
        atms.append(atm)
        bass.append(bas)
        envs.append(np.asarray(m._env, dtype=np.float64))
        natm += m.natm
        nenv += len(m._env)
    return (np.ascontiguousarray(np.concatenate(atms)), np.ascontiguousarray(np.concatenate(bass)),
            np.ascontiguousarray(np.concatenate(envs)), np.array([m.nbas for m in mols], dtype=np.int32))


class MoleculeBatch:
    def __init__(self, mols):
        if not mols:
            raise ValueError("empty batch")
        self.mols = list(mols)
        self.atm, self.bas, self.env, self.mol_nbas = conc_env(self.mols)
        self.natm = int(self.atm.shape[0])
        self.atom0 = np.concatenate([[0], np.cumsum([m.natm for m in self.mols])]).astype(int)
        self.nao_cart = [m.nao_cart() for m in self.mols]
        self.device = torch.cuda.current_device()
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self._lib.ccsd_plan_create_batch(
            ctypes.byref(h), len(self.mols), self.mol_nbas.ctypes.data, self.atm.ctypes.data, self.natm,
            self.bas.ctypes.data, int(self.bas.shape[0]), self.env.ctypes.data, len(self.env)),
            "ccsd_plan_create_batch")
        self._plan = h

    def close(self):
        if self._plan is not None and self._plan.value:
            self._lib.ccsd_plan_destroy(self._plan)
        self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_option(self, option, value):
        _lib.check(self._lib.ccsd_plan_set_option(self._plan, int(option), int(value)), "ccsd_plan_set_option")

    def set_persistent(self, on=True, body="register", order="class"):
        """Run each fused pass as one persistent launch over every class of every
        molecule (one-wave blocks draw work items from a device counter), instead of
        one launch per class on four streams (the default). body="register": the
        register-resident per-class bodies in that one kernel (the LDS kernel's body if
        some class has none); "generic": the LDS kernel's body. order="class": work
        items class by class, costliest class first; "cost": by decreasing cost only."""
        from .hip.grad2e import PERSISTENT_BODIES
        self.set_option(7, {"class": 0, "cost": 1}[order])
        self.set_option(5, PERSISTENT_BODIES[body] if on else 0)

    def screen(self, thr):
        _lib.check(self._lib.ccsd_plan_screen(self._plan, float(thr), _lib.current_stream_ptr()), "ccsd_plan_screen")

    def schwarz(self):
        nbas = int(self.bas.shape[0])
        q = torch.empty((nbas, nbas), dtype=torch.float64, device="cuda")
        _lib.check(self._lib.ccsd_schwarz(self._plan, _lib.ptr(q), _lib.current_stream_ptr()), "ccsd_schwarz")
        return q

    def last_launches(self):
        """Launch records of the last fused pass (hip.grad2e.last_launches)."""
        from .hip.grad2e import last_launches
        return last_launches(plan=self._plan)

    def last_quartets(self):
        n = ctypes.c_longlong()
        _lib.check(self._lib.ccsd_plan_quartets(self._plan, ctypes.byref(n)), "ccsd_plan_quartets")
        return n.value

    def _cart_blocks(self, mats):
        blocks = []
        for m, d, nc in zip(self.mols, mats, self.nao_cart):
            d = torch.as_tensor(d, dtype=torch.float64, device="cuda")
            if not m.cart:
                c = torch.as_tensor(m.cart2sph_coeff(), dtype=torch.float64, device="cuda")
                d = c @ d @ c.T
            if tuple(d.shape) != (nc, nc):
                raise ValueError(f"density of shape {tuple(d.shape)} for a molecule with {nc} Cartesian AOs")
            blocks.append(d.reshape(-1))
        return torch.cat(blocks).contiguous()

    def grad2e_separable(self, dms, pms=None, scale=-2.0, de=None):
        """Per molecule m: de_m[atom(i), x] += scale * sum_{i, j, k>=l} ip1_x(ij|kl) Gamma'_ijkl(D_m, P_m).

        dms, pms: one (nao_m, nao_m) symmetric density per molecule, in the
        molecule's own AO basis (spherical or Cartesian). Returns a list of
        (natm_m, 3) device tensors."""
        if len(dms) != len(self.mols):
            raise ValueError("one density per molecule")
        D = self._cart_blocks(dms)
        P = D if pms is None else self._cart_blocks(pms)
        if de is None:
            de = torch.zeros((self.natm, 3), dtype=torch.float64, device="cuda")
        _lib.check(self._lib.ccsd_grad_ip1_separable(self._plan, _lib.ptr(de), _lib.ptr(D), _lib.ptr(P),
                                                     float(scale), _lib.current_stream_ptr()),
                   "ccsd_grad_ip1_separable")
        return [de[self.atom0[i]:self.atom0[i + 1]] for i in range(len(self.mols))]

    def quartets_unscreened(self):
        """Shell quartets of the full enumeration (bra i, j; ket k >= l) of every molecule."""
        return int(sum(int(n) ** 2 * int(n) * (int(n) + 1) // 2 for n in self.mol_nbas))
