"""FP64 CPU oracle of the CCSD gradient for systems whose dense tensors do not fit
the full-tensor oracle (oracle/chem.py) in host memory: (H2O)8 / cc-pVDZ.

TEST INFRASTRUCTURE ONLY (writes the golden fixture; never shipped, never timed).

Same equations as oracle/chem.py (RHF, CCSD, lambda, RDMs, Z-vector, gradient,
restating ccsd_gpu/gpu_ccsd_lambda.py and ccsd_gpu/gpu_ccsd_grad.py); what differs
is only how the tensors are held:
  * AO ERIs built one i-shell at a time (C oracle, Cartesian s2kl -> native basis);
  * MO integrals by four sequential one-index transforms (no 5-operand einsum);
  * K builds one AO row at a time;
  * the Z-vector by conjugate gradients on the same fvind (JK of the trial density)
    instead of the dense CPHF matrix;
  * the two-electron derivative term by the reference's block loop
    (ccsd_gpu/gpu_ccsd_grad.py:423-496): int2e_ip1 per i-shell, contracted with the
    symmetrized AO 2-RDM rows and with the HF vhf1 pieces (C1, X2, C3, X4).
tests/test_oracle_large.py checks it against oracle/chem.py on small molecules.
"""
import gc
from functools import reduce

import numpy as np

from . import chem
from . import integrals as oi


def int2e_native_blocked(mol):
    """(n, n, n, n) chemist's ERIs in the native AO basis, one i-shell at a time."""
    n = mol.nao_nr()
    mc = mol.copy()
    mc.cart = True
    lc = mc.ao_loc_nr(cart=True)
    nc = int(lc[-1])
    ls = mol.ao_loc_nr()
    C = mol.cart2sph_coeff() if not mol.cart else np.eye(nc)
    out = np.empty((n, n, n, n))
    k, l = np.tril_indices(nc)
    for b in range(mol.nbas):
        blk = oi.int2e_s2kl_cart(mc, b, b + 1, 0)[0]            # (nf_c, nc, npair_c)
        full = np.empty(blk.shape[:2] + (nc, nc))
        full[..., k, l] = blk
        full[..., l, k] = blk
        del blk
        t = np.tensordot(full, C, axes=([3], [0]))               # (f, j, k, l')
        t = np.tensordot(t, C, axes=([2], [0]))                  # (f, j, l', k')
        t = np.tensordot(t, C, axes=([1], [0]))                  # (f, l', k', j')
        t = np.tensordot(C[lc[b]:lc[b + 1], ls[b]:ls[b + 1]], t, axes=([0], [0]))   # (i', l', k', j')
        out[ls[b]:ls[b + 1]] = t.transpose(0, 3, 2, 1)
        del full, t
    return out


class RHF(chem.RHF):
    """chem.RHF over the blocked ERIs, K by rows."""

    def eri(self):
        if self._eri is None:
            self._eri = int2e_native_blocked(self.mol)
        return self._eri

    def get_jk(self, dm):
        eri = self.eri()
        n = dm.shape[0]
        j = (eri.reshape(n * n, n * n) @ dm.reshape(-1)).reshape(n, n)
        k = np.empty((n, n))
        for i in range(n):
            k[i] = np.tensordot(eri[i], dm, axes=([0, 2], [0, 1]))     # sum_kl (ik|jl) D_kl
        return j, k


def transform4(x, c):
    """y[i,j,k,l] = sum_pqrs c[p,i] c[q,j] c[r,k] c[s,l] x[p,q,r,s] by four matrix
    products, each reading the previous result once (two n^4 arrays alive at a time)."""
    n, m = c.shape
    t = c.T @ x.reshape(n, -1)                               # (i, q r s)
    t = np.matmul(c.T, t.reshape(m, n, n * n))               # (i, j, r s)
    t = t.reshape(m * m, n, n) @ c                           # (ij, r, l)
    t = np.matmul(c.T, t)                                    # (ij, k, l)
    return t.reshape(m, m, m, m)


class ERIs:
    """chem.ERIs built by sequential transforms; no full MO tensor is kept."""

    def __init__(self, mf, mo_coeff=None, mask=None):
        c_all = mf.mo_coeff if mo_coeff is None else mo_coeff
        mask = np.ones(c_all.shape[1], dtype=bool) if mask is None else mask
        c = np.ascontiguousarray(c_all[:, mask])
        nocc = int((mf.mo_occ[mask] > 0).sum())
        eri = transform4(mf.eri(), c)
        dm = mf.make_rdm1(c_all, mf.mo_occ)
        fock_ao = mf.get_hcore() + mf.get_veff(dm=dm)
        self.fock = c.T @ fock_ao @ c
        self.mo_energy = self.fock.diagonal().copy()
        o, v = slice(0, nocc), slice(nocc, None)
        self.nocc = nocc
        self.oooo = np.ascontiguousarray(eri[o, o, o, o])
        self.ovoo = np.ascontiguousarray(eri[o, v, o, o])
        self.ovvo = np.ascontiguousarray(eri[o, v, v, o])
        self.oovv = np.ascontiguousarray(eri[o, o, v, v])
        self.ovov = np.ascontiguousarray(eri[o, v, o, v])
        self.ovvv = np.ascontiguousarray(eri[o, v, v, v])
        self.vvvv = np.ascontiguousarray(eri[v, v, v, v])
        del eri

    def get_ovvv(self, *slices):
        return self.ovvv[slices] if slices else self.ovvv


class CCSD(chem.CCSD):
    def ao2mo(self, mo_coeff=None):
        return ERIs(self._scf, mo_coeff, self.get_frozen_mask())


def _cg(apply, b, precond, tol=1e-12, max_cycle=500):
    x = b * precond
    r = b - apply(x)
    z = r * precond
    p = z.copy()
    rz = np.vdot(r, z)
    bn = np.linalg.norm(b)
    for _ in range(max_cycle):
        if np.linalg.norm(r) <= tol * bn:
            return x
        ap = apply(p)
        a = rz / np.vdot(p, ap)
        x += a * p
        r -= a * ap
        z = r * precond
        rz_new = np.vdot(r, z)
        p = z + (rz_new / rz) * p
        rz = rz_new
    raise RuntimeError("oracle CPHF did not converge")


def _symmetrized_ao_rdm2(holder, c):
    """0.5 (G + G^(ij) + G^(kl) + G^(ij)(kl)) of the AO 2-RDM (gpu_ccsd_grad.py:313-326),
    symmetrized in place. The reference transposes the MO density (1, 0, 3, 2) before
    the transform; that permutation is one of the four the symmetrization sums over,
    so it is left out. `holder` = [d2] is emptied once the MO tensor is assembled."""
    dm2 = chem.assemble_dm2_mo(holder.pop())
    gc.collect()
    s = transform4(dm2, np.ascontiguousarray(c.T))
    del dm2
    gc.collect()
    n = s.shape[0]
    for i in range(n):
        for j in range(i):
            t = s[i, j] + s[j, i]
            s[i, j] = t
            s[j, i] = t
        s[i, i] *= 2.0
    for i in range(n):
        s[i] += s[i].transpose(0, 2, 1).copy()
    s *= 0.5
    return s


def grad_elec(mycc, t1, t2, l1, l2, release=None):
    """dE/dR (natm, 3) electronic, as chem.grad_elec; `release()` (optional) is called
    once the lambda-side tensors are no longer needed, so a caller can free its eris."""
    mol = mycc.mol
    mf = mycc._scf
    mo_coeff, mo_energy, mo_occ = mycc.mo_coeff, mf.mo_energy, mycc.mo_occ
    nao, nmo = mo_coeff.shape
    nocc = int((mo_occ > 0).sum())
    mask = mycc.get_frozen_mask()
    OA = np.nonzero(mask & (mo_occ > 0))[0]
    VA = np.nonzero(mask & (mo_occ == 0))[0]
    OF = np.nonzero(~mask & (mo_occ > 0))[0]
    VF = np.nonzero(~mask & (mo_occ == 0))[0]
    d1 = chem.gamma1_intermediates(t1, t2, l1, l2)
    doo, dov, dvo, dvv = d1
    holder = [chem.gamma2_intermediates(t1, t2, l1, l2)]
    if release is not None:
        release()
    S2 = _symmetrized_ao_rdm2(holder, mo_coeff[:, np.hstack((OA, VA))])
    gc.collect()
    eri = mf.eri()
    Imat = 0.5 * (eri.reshape(nao, -1) @ S2.reshape(nao, -1).T)
    # packed rows of the symmetrized density, k >= l, diagonal halved (Gamma' of :431)
    kk, ll = np.tril_indices(nao)
    Gp = S2[:, :, kk, ll]
    Gp[:, :, kk == ll] *= 0.5
    del S2
    gc.collect()
    hf_dm1 = mf.make_rdm1(mo_coeff, mo_occ)

    # Phase 4
    Imat = reduce(np.dot, (mo_coeff.T, Imat, mf.get_ovlp(), mo_coeff)) * -1
    dm1mo = np.zeros((nmo, nmo))
    if not mask.all():
        dco = Imat[OF[:, None], OA] / (mo_energy[OF, None] - mo_energy[OA])
        dfv = Imat[VF[:, None], VA] / (mo_energy[VF, None] - mo_energy[VA])
        dm1mo[OA[:, None], OA] = doo + doo.T
        dm1mo[OF[:, None], OA] = dco
        dm1mo[OA[:, None], OF] = dco.T
        dm1mo[VA[:, None], VA] = dvv + dvv.T
        dm1mo[VF[:, None], VA] = dfv
        dm1mo[VA[:, None], VF] = dfv.T
    else:
        dm1mo[:nocc, :nocc] = doo + doo.T
        dm1mo[nocc:, nocc:] = dvv + dvv.T
    dm1 = reduce(np.dot, (mo_coeff, dm1mo, mo_coeff.T))
    vhf = mf.get_veff(mol, dm1) * 2
    Xvo = reduce(np.dot, (mo_coeff[:, nocc:].T, vhf, mo_coeff[:, :nocc]))
    Xvo += Imat[:nocc, nocc:].T - Imat[nocc:, :nocc]
    cv, co = mo_coeff[:, nocc:], mo_coeff[:, :nocc]
    eai = mo_energy[nocc:, None] - mo_energy[None, :nocc]

    def apply(x):
        d = cv @ x @ co.T
        v = mf.get_veff(mol, d + d.T)
        return eai * x + (cv.T @ v @ co) * 2

    dvo_resp = _cg(apply, -Xvo, 1.0 / eai)
    dm1mo[nocc:, :nocc] += dvo_resp
    dm1mo[:nocc, nocc:] += dvo_resp.T
    Imat[nocc:, :nocc] = Imat[:nocc, nocc:].T
    im1 = reduce(np.dot, (mo_coeff, Imat, mo_coeff.T))
    # Phase 5 densities
    s1 = -oi.int1e(mol, "ipovlp")
    vrinv_all = oi.int1e(mol, "ipnuc_atom")
    h1 = -(oi.int1e(mol, "ipkin") + vrinv_all.sum(axis=0))
    zeta = (mo_energy[:, None] + mo_energy[None, :]) * 0.5
    zeta[nocc:, :nocc] = mo_energy[:nocc]
    zeta[:nocc, nocc:] = mo_energy[:nocc].reshape(-1, 1)
    zeta = reduce(np.dot, (mo_coeff, zeta * dm1mo, mo_coeff.T))
    dm1 = reduce(np.dot, (mo_coeff, dm1mo, mo_coeff.T))
    p1 = np.dot(mo_coeff[:, :nocc], mo_coeff[:, :nocc].T)
    vhf_s1occ = reduce(np.dot, (p1, mf.get_veff(mol, dm1 + dm1.T), p1))
    dm1p = hf_dm1 + dm1 * 2
    dm1 = dm1 + hf_dm1
    mocc = mo_coeff[:, mo_occ > 0]
    zeta += (mocc * (mo_energy[mo_occ > 0] * mo_occ[mo_occ > 0])) @ mocc.T

    # Phase 3: the reference's block loop over i-shells, int2e_ip1 -> contractions
    de = np.zeros((mol.natm, 3))
    mc = mol.copy()
    mc.cart = True
    lc = mc.ao_loc_nr(cart=True)
    nc = int(lc[-1])
    ls = mol.ao_loc_nr()
    C = mol.cart2sph_coeff() if not mol.cart else np.eye(nc)
    kc, lcn = np.tril_indices(nc)
    D, P = hf_dm1, dm1p
    for b in range(mol.nbas):
        a = mol.bas_atom(b)
        r0, r1 = ls[b], ls[b + 1]
        blk = oi.int2e_s2kl_cart(mc, b, b + 1, 1)                # (3, nf_c, nc, npair_c)
        full = np.empty(blk.shape[:3] + (nc, nc))
        full[..., kc, lcn] = blk
        full[..., lcn, kc] = blk
        del blk
        t = np.tensordot(full, C, axes=([4], [0]))               # (x f j k l')
        t = np.tensordot(t, C, axes=([3], [0]))                  # (x f j l' k')
        t = np.tensordot(t, C, axes=([2], [0]))                  # (x f l' k' j')
        t = np.tensordot(C[lc[b]:lc[b + 1], r0:r1], t, axes=([0], [1]))   # (i' x l' k' j')
        e = np.ascontiguousarray(t.transpose(1, 0, 4, 3, 2))     # (x i j k l) native
        del full, t
        ep = e[..., kk, ll]
        de[a] -= 2.0 * np.einsum("xijp,ijp->x", ep, Gp[r0:r1], optimize=True)
        del ep
        c1 = np.einsum("xijkl,ij,kl->x", e, D[r0:r1], P, optimize=True)
        c2 = np.einsum("xijkl,il,kj->x", e, D[r0:r1], P, optimize=True)
        c3 = np.einsum("xijkl,kl,ij->x", e, D, P[r0:r1], optimize=True)
        c4 = np.einsum("xijkl,jk,il->x", e, D, P[r0:r1], optimize=True)
        de[a] -= c1 - 0.5 * c2 + c3 - 0.5 * c4
        del e
    offs = mol.offset_nr_by_atom()
    for k in range(mol.natm):
        p0, p1_ = offs[k, 2], offs[k, 3]
        vrinv = vrinv_all[k].copy()
        vrinv[:, p0:p1_] += h1[:, p0:p1_]
        h1ao = vrinv + vrinv.transpose(0, 2, 1)
        de[k] += np.einsum("xij,ij->x", s1[:, p0:p1_], im1[p0:p1_])
        de[k] += np.einsum("xji,ij->x", s1[:, p0:p1_], im1[:, p0:p1_])
        de[k] += np.einsum("xij,ji->x", h1ao, dm1)
        de[k] -= np.einsum("xij,ij->x", s1[:, p0:p1_], zeta[p0:p1_])
        de[k] -= np.einsum("xji,ij->x", s1[:, p0:p1_], zeta[:, p0:p1_])
        de[k] -= np.einsum("xij,ij->x", s1[:, p0:p1_], vhf_s1occ[p0:p1_]) * 2
    return de


def ccsd_gradient(mol, frozen=None, log=print, checkpoint=None):
    """RHF -> CCSD -> lambda -> total dE/dR; returns (E_hf, E_corr, dE/dR).

    checkpoint: a file prefix; the converged amplitudes (t1, t2, E_corr) and lambdas
    are saved to <prefix>_t.npz / <prefix>_l.npz as they are reached and read back
    by a later call with the same prefix instead of being solved again (the RHF is
    always rerun: it is deterministic and cheap next to the rest)."""
    import os
    import time
    t0 = time.perf_counter()
    mf = RHF(mol)
    mf.kernel()
    log(f"RHF {mf.e_tot:.12f} ({time.perf_counter() - t0:.0f} s)")
    cc = CCSD(mf, frozen=frozen)
    tfile = lfile = None
    if checkpoint:
        tfile, lfile = checkpoint + "_t.npz", checkpoint + "_l.npz"
    if tfile and os.path.exists(tfile):
        with np.load(tfile) as z:
            cc.t1, cc.t2, cc.e_corr = z["t1"], z["t2"], float(z["e_corr"])
        log(f"CCSD corr {cc.e_corr:.12f} read from {tfile}")
    else:
        cc.kernel()
        log(f"CCSD corr {cc.e_corr:.12f} ({time.perf_counter() - t0:.0f} s)")
        if tfile:
            np.savez(tfile, t1=cc.t1, t2=cc.t2, e_corr=cc.e_corr)
    eris = cc.ao2mo()
    if lfile and os.path.exists(lfile):
        with np.load(lfile) as z:
            l1, l2 = z["l1"], z["l2"]
        log(f"lambda read from {lfile}")
    else:
        conv, l1, l2 = chem.solve_lambda(cc.t1, cc.t2, eris)
        assert conv
        log(f"lambda ({time.perf_counter() - t0:.0f} s)")
        if lfile:
            np.savez(lfile, l1=l1, l2=l2)
    holder = [eris]

    def release():
        holder.clear()
        gc.collect()

    del eris
    de = grad_elec(cc, cc.t1, cc.t2, l1, l2, release=release) + chem.grad_nuc(mol)
    log(f"gradient ({time.perf_counter() - t0:.0f} s)")
    return mf.e_tot, cc.e_corr, de
