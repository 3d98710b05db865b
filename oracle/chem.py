"""FP64 CPU oracle of the CCSD-gradient path: RHF, CCSD, lambda, RDMs, gradient.

TEST INFRASTRUCTURE ONLY (checker and CPU baseline; never shipped).

Dense numpy with full ERI tensors, for small molecules. Every function cites the
reference code it restates:
  * lambda intermediates / update / iteration: ccsd_gpu/gpu_ccsd_lambda.py:68-209,
    212-337, 340-408 (a port of PySCF ccsd_lambda; the vvvv term is PySCF's
    _add_vvvv: l2new = einsum('ijcd,acbd->ijab', l2, vvvv), gpu_ccsd_lambda.py:238-243);
  * 1-/2-RDM intermediates: ccsd_gpu/gpu_ccsd_grad.py:101-246;
  * 2-RDM assembly and MO->AO: gpu_ccsd_grad.py:254-336;
  * gradient assembly: gpu_ccsd_grad.py:344-586 (Phases 1-5), Z-vector
    gpu_ccsd_grad.py:625-666;
PySCF pieces the reference calls and that are absent here are restated from
their published definitions: RHF SCF, CCSD amplitude equations (closed-shell
rccsd with the standard intermediates), cphf.solve (solved exactly here),
hcore_generator / make_rdm1e / grad_nuc.
"""
from functools import reduce

import numpy as np

from . import integrals as oi


def _einsum(eq, *ops, **kw):
    """numpy einsum through its BLAS-backed contraction path (same sums, reordered)."""
    kw.setdefault("optimize", True)
    return np.einsum(eq, *ops, **kw)


# ---------------------------------------------------------------------------
# RHF
# ---------------------------------------------------------------------------
class RHF:
    """Closed-shell Hartree-Fock with full in-core ERIs (DIIS on FDS - SDF)."""

    def __init__(self, mol):
        self.mol = mol
        self.conv_tol = 1e-12
        self.max_cycle = 200
        self._eri = None

    def eri(self):
        if self._eri is None:
            self._eri = oi.int2e_full(self.mol)
        return self._eri

    def get_ovlp(self, mol=None):
        return oi.int1e(self.mol, "ovlp")

    def get_hcore(self, mol=None):
        return oi.int1e(self.mol, "kin") + oi.int1e(self.mol, "nuc")

    def get_jk(self, dm):
        eri = self.eri()
        return _einsum("ijkl,kl->ij", eri, dm), _einsum("ikjl,kl->ij", eri, dm)

    def get_veff(self, mol=None, dm=None):
        j, k = self.get_jk(dm)
        return j - 0.5 * k

    def energy_nuc(self):
        return self.mol.energy_nuc()

    def make_rdm1(self, mo_coeff=None, mo_occ=None):
        mo_coeff = self.mo_coeff if mo_coeff is None else mo_coeff
        mo_occ = self.mo_occ if mo_occ is None else mo_occ
        c = mo_coeff[:, mo_occ > 0]
        return (c * mo_occ[mo_occ > 0]) @ c.T

    def kernel(self):
        S = self.get_ovlp()
        h = self.get_hcore()
        w, v = np.linalg.eigh(S)
        X = v / np.sqrt(w)
        nocc = self.mol.nelectron // 2
        nmo = X.shape[1]
        occ = np.zeros(nmo)
        occ[:nocc] = 2

        def diag(f):
            e, c = np.linalg.eigh(X.T @ f @ X)
            return e, X @ c

        e, c = diag(h)
        dm = (c[:, :nocc] * 2) @ c[:, :nocc].T
        fs, errs = [], []
        e_old = 0.0
        for it in range(self.max_cycle):
            f = h + self.get_veff(dm=dm)
            err = f @ dm @ S - S @ dm @ f
            fs.append(f)
            errs.append(err)
            fs, errs = fs[-8:], errs[-8:]
            n = len(fs)
            B = -np.ones((n + 1, n + 1))
            B[n, n] = 0
            for i in range(n):
                for j in range(n):
                    B[i, j] = np.vdot(errs[i], errs[j])
            rhs = np.zeros(n + 1)
            rhs[n] = -1
            coef = np.linalg.solve(B, rhs)[:n]
            fd = sum(ci * fi for ci, fi in zip(coef, fs))
            e, c = diag(fd)
            dm = (c[:, :nocc] * 2) @ c[:, :nocc].T
            f = h + self.get_veff(dm=dm)
            etot = 0.5 * np.vdot(dm, h + f) + self.energy_nuc()
            if abs(etot - e_old) < self.conv_tol and np.abs(err).max() < 1e-9:
                break
            e_old = etot
        # final canonical orbitals of the converged Fock matrix
        e, c = diag(f)
        self.mo_energy, self.mo_coeff, self.mo_occ = e, c, occ
        dm = self.make_rdm1()
        f = h + self.get_veff(dm=dm)
        self.e_tot = 0.5 * np.vdot(dm, h + f) + self.energy_nuc()
        self.converged = True
        return self.e_tot


# ---------------------------------------------------------------------------
# MO integrals (PySCF _ChemistsERIs layout)
# ---------------------------------------------------------------------------
def frozen_mask(nmo, frozen):
    """PySCF's `frozen` convention (an int: that many lowest orbitals; a list: those
    indices) as a mask of active orbitals."""
    mask = np.ones(nmo, dtype=bool)
    if frozen is None:
        return mask
    if np.isscalar(frozen):
        mask[:int(frozen)] = False
    else:
        mask[np.asarray(list(frozen), dtype=int)] = False
    return mask


class ERIs:
    """MO integrals over the active orbitals; the Fock matrix is that of the HF density
    of all occupied orbitals (frozen ones included) in the active-orbital basis."""

    def __init__(self, mf, mo_coeff=None, mask=None):
        c_all = mf.mo_coeff if mo_coeff is None else mo_coeff
        mask = np.ones(c_all.shape[1], dtype=bool) if mask is None else mask
        c = c_all[:, mask]
        nocc = int((mf.mo_occ[mask] > 0).sum())
        eri = _einsum("pqrs,pi,qj,rk,sl->ijkl", mf.eri(), c, c, c, c, optimize=True)
        dm = mf.make_rdm1(c_all, mf.mo_occ)
        fock_ao = mf.get_hcore() + mf.get_veff(dm=dm)
        self.fock = c.T @ fock_ao @ c
        self.mo_energy = self.fock.diagonal().copy()
        o, v = slice(0, nocc), slice(nocc, None)
        self.nocc = nocc
        self.eri_mo = eri
        self.oooo = eri[o, o, o, o].copy()
        self.ovoo = eri[o, v, o, o].copy()
        self.ovvo = eri[o, v, v, o].copy()
        self.oovv = eri[o, o, v, v].copy()
        self.ovov = eri[o, v, o, v].copy()
        self.ovvv = eri[o, v, v, v].copy()
        self.vvvv = eri[v, v, v, v].copy()

    def get_ovvv(self, *slices):
        return self.ovvv[slices] if slices else self.ovvv


# ---------------------------------------------------------------------------
# CCSD amplitudes (closed-shell spin-adapted equations)
# ---------------------------------------------------------------------------
def ccsd_energy(t1, t2, eris):
    nocc = t1.shape[0]
    e = 2 * _einsum("ia,ia", eris.fock[:nocc, nocc:], t1)
    tau = t2 + _einsum("ia,jb->ijab", t1, t1)
    e += 2 * _einsum("ijab,iajb", tau, eris.ovov)
    e -= _einsum("ijab,ibja", tau, eris.ovov)
    return e


def ccsd_update_amps(t1, t2, eris):
    nocc, nvir = t1.shape
    fock = eris.fock
    e_o = eris.mo_energy[:nocc]
    e_v = eris.mo_energy[nocc:]
    fov = fock[:nocc, nocc:]
    foo = fock[:nocc, :nocc]
    fvv = fock[nocc:, nocc:]
    ovov, ovoo, ovvo, oovv, ovvv = eris.ovov, eris.ovoo, eris.ovvo, eris.oovv, eris.ovvv

    Foo = foo + 2 * _einsum("kcld,ilcd->ki", ovov, t2) - _einsum("kdlc,ilcd->ki", ovov, t2) \
        + 2 * _einsum("kcld,ic,ld->ki", ovov, t1, t1) - _einsum("kdlc,ic,ld->ki", ovov, t1, t1)
    Fvv = fvv - 2 * _einsum("kcld,klad->ac", ovov, t2) + _einsum("kdlc,klad->ac", ovov, t2) \
        - 2 * _einsum("kcld,ka,ld->ac", ovov, t1, t1) + _einsum("kdlc,ka,ld->ac", ovov, t1, t1)
    Fov = fov + 2 * _einsum("kcld,ld->kc", ovov, t1) - _einsum("kdlc,ld->kc", ovov, t1)
    Loo = Foo + _einsum("kc,ic->ki", fov, t1) + 2 * _einsum("lcki,lc->ki", ovoo, t1) \
        - _einsum("kcli,lc->ki", ovoo, t1)
    Lvv = Fvv - _einsum("kc,ka->ac", fov, t1) + 2 * _einsum("kdac,kd->ac", ovvv, t1) \
        - _einsum("kcad,kd->ac", ovvv, t1)
    Foo = Foo - np.diag(e_o)
    Fvv = Fvv - np.diag(e_v)
    Loo = Loo - np.diag(e_o)
    Lvv = Lvv - np.diag(e_v)

    t1n = -2 * _einsum("kc,ka,ic->ia", fov, t1, t1)
    t1n += _einsum("ac,ic->ia", Fvv, t1)
    t1n -= _einsum("ki,ka->ia", Foo, t1)
    t1n += 2 * _einsum("kc,kica->ia", Fov, t2)
    t1n -= _einsum("kc,ikca->ia", Fov, t2)
    t1n += _einsum("kc,ic,ka->ia", Fov, t1, t1)
    t1n += fov
    t1n += 2 * _einsum("kcai,kc->ia", ovvo, t1)
    t1n -= _einsum("kiac,kc->ia", oovv, t1)
    t1n += 2 * _einsum("kdac,ikcd->ia", ovvv, t2)
    t1n -= _einsum("kcad,ikcd->ia", ovvv, t2)
    t1n += 2 * _einsum("kdac,kd,ic->ia", ovvv, t1, t1)
    t1n -= _einsum("kcad,kd,ic->ia", ovvv, t1, t1)
    t1n -= 2 * _einsum("lcki,klac->ia", ovoo, t2)
    t1n += _einsum("kcli,klac->ia", ovoo, t2)
    t1n -= 2 * _einsum("lcki,lc,ka->ia", ovoo, t1, t1)
    t1n += _einsum("kcli,lc,ka->ia", ovoo, t1, t1)

    tmp2 = _einsum("kibc,ka->abic", oovv, -t1) + ovvv.transpose(1, 3, 0, 2)
    tmp = _einsum("abic,jc->ijab", tmp2, t1)
    t2n = tmp + tmp.transpose(1, 0, 3, 2)
    tmp2 = _einsum("kcai,jc->akij", ovvo, t1) + ovoo.transpose(1, 3, 0, 2)
    tmp = _einsum("akij,kb->ijab", tmp2, t1)
    t2n -= tmp + tmp.transpose(1, 0, 3, 2)
    t2n += ovov.transpose(0, 2, 1, 3)

    Woooo = _einsum("lcki,jc->klij", ovoo, t1) + _einsum("kclj,ic->klij", ovoo, t1) \
        + _einsum("kcld,ijcd->klij", ovov, t2) + _einsum("kcld,ic,jd->klij", ovov, t1, t1) \
        + eris.oooo.transpose(0, 2, 1, 3)
    Wvvvv = -_einsum("kdac,kb->abcd", ovvv, t1) - _einsum("kcbd,ka->abcd", ovvv, t1) \
        + eris.vvvv.transpose(0, 2, 1, 3)
    Wvoov = _einsum("kcad,id->akic", ovvv, t1) - _einsum("kcli,la->akic", ovoo, t1) \
        + ovvo.transpose(2, 0, 3, 1) - 0.5 * _einsum("ldkc,ilda->akic", ovov, t2) \
        - 0.5 * _einsum("lckd,ilad->akic", ovov, t2) - _einsum("ldkc,id,la->akic", ovov, t1, t1) \
        + _einsum("ldkc,ilad->akic", ovov, t2)
    Wvovo = _einsum("kdac,id->akci", ovvv, t1) - _einsum("lcki,la->akci", ovoo, t1) \
        + oovv.transpose(2, 0, 3, 1) - 0.5 * _einsum("lckd,ilda->akci", ovov, t2) \
        - _einsum("lckd,id,la->akci", ovov, t1, t1)

    tau = t2 + _einsum("ia,jb->ijab", t1, t1)
    t2n += _einsum("klij,klab->ijab", Woooo, tau)
    t2n += _einsum("abcd,ijcd->ijab", Wvvvv, tau)
    tmp = _einsum("ac,ijcb->ijab", Lvv, t2)
    t2n += tmp + tmp.transpose(1, 0, 3, 2)
    tmp = _einsum("ki,kjab->ijab", Loo, t2)
    t2n -= tmp + tmp.transpose(1, 0, 3, 2)
    tmp = 2 * _einsum("akic,kjcb->ijab", Wvoov, t2) - _einsum("akci,kjcb->ijab", Wvovo, t2)
    t2n += tmp + tmp.transpose(1, 0, 3, 2)
    tmp = _einsum("akic,kjbc->ijab", Wvoov, t2)
    t2n -= tmp + tmp.transpose(1, 0, 3, 2)
    tmp = _einsum("bkci,kjac->ijab", Wvovo, t2)
    t2n -= tmp + tmp.transpose(1, 0, 3, 2)

    eia = e_o[:, None] - e_v
    t1n /= eia
    t2n /= eia[:, None, :, None] + eia[None, :, None, :]
    return t1n, t2n


class DIIS:
    """Pulay DIIS on amplitude vectors with error = difference of successive vectors."""

    def __init__(self, space=6):
        self.space = space
        self.xs, self.errs = [], []
        self.prev = None

    def update(self, x):
        if self.prev is None:
            self.prev = x
            return x
        self.xs.append(x)
        self.errs.append(x - self.prev)
        self.xs, self.errs = self.xs[-self.space:], self.errs[-self.space:]
        n = len(self.xs)
        B = -np.ones((n + 1, n + 1))
        B[n, n] = 0
        for i in range(n):
            for j in range(n):
                B[i, j] = np.dot(self.errs[i], self.errs[j])
        rhs = np.zeros(n + 1)
        rhs[n] = -1
        c = np.linalg.lstsq(B, rhs, rcond=None)[0][:n]
        out = sum(ci * xi for ci, xi in zip(c, self.xs))
        self.prev = out
        return out


def _pack(a, b):
    return np.concatenate([a.ravel(), b.ravel()])


def _unpack(v, nocc, nvir):
    n1 = nocc * nvir
    return v[:n1].reshape(nocc, nvir), v[n1:].reshape(nocc, nocc, nvir, nvir)


class CCSD:
    """Closed-shell CCSD with the attributes the reference's code reads; `frozen`
    follows PySCF (int or list of frozen orbital indices)."""

    def __init__(self, mf, frozen=None):
        self._scf = mf
        self.mol = mf.mol
        self.mo_coeff = mf.mo_coeff
        self.mo_occ = mf.mo_occ
        self.frozen = frozen
        self.max_cycle = 200
        self.conv_tol = 1e-11
        self.conv_tol_normt = 1e-9
        self.level_shift = 0.0
        self.diis_space = 6
        self.verbose = 0
        self.t1 = self.t2 = self.l1 = self.l2 = None

    def get_frozen_mask(self):
        return frozen_mask(self.mo_coeff.shape[1], self.frozen)

    @property
    def nocc(self):
        return int((self.mo_occ[self.get_frozen_mask()] > 0).sum())

    def ao2mo(self, mo_coeff=None):
        return ERIs(self._scf, mo_coeff, self.get_frozen_mask())

    def kernel(self, eris=None):
        eris = self.ao2mo() if eris is None else eris
        nocc = self.nocc
        e_o, e_v = eris.mo_energy[:nocc], eris.mo_energy[nocc:]
        nvir = e_v.size
        eia = e_o[:, None] - e_v
        t1 = eris.fock[:nocc, nocc:] / eia
        t2 = eris.ovov.transpose(0, 2, 1, 3) / (eia[:, None, :, None] + eia[None, :, None, :])
        e = ccsd_energy(t1, t2, eris)
        diis = DIIS(self.diis_space)
        self.converged = False
        for it in range(self.max_cycle):
            t1n, t2n = ccsd_update_amps(t1, t2, eris)
            norm = np.linalg.norm(_pack(t1n, t2n) - _pack(t1, t2))
            t1, t2 = _unpack(diis.update(_pack(t1n, t2n)), nocc, nvir)
            e_new = ccsd_energy(t1, t2, eris)
            if abs(e_new - e) < self.conv_tol and norm < self.conv_tol_normt:
                e = e_new
                self.converged = True
                break
            e = e_new
        self.e_corr = e
        self.e_tot = self._scf.e_tot + e
        self.t1, self.t2 = t1, t2
        return e, t1, t2


# ---------------------------------------------------------------------------
# Lambda equations (restates ccsd_gpu/gpu_ccsd_lambda.py, dense)
# ---------------------------------------------------------------------------
class IMDS:
    pass


def make_intermediates(t1, t2, eris):
    """gpu_ccsd_lambda.py:68-209 with one virtual block (nvir) and one occupied block."""
    nocc, nvir = t1.shape
    fock = eris.fock
    foo, fov, fvv = fock[:nocc, :nocc], fock[:nocc, nocc:], fock[nocc:, nocc:]
    w1 = fvv - _einsum("ja,jb->ba", fov, t1)
    w2 = foo + _einsum("ib,jb->ij", fov, t1)
    w3 = _einsum("kc,jkbc->bj", fov, t2) * 2 + fov.T
    w3 -= _einsum("kc,kjbc->bj", fov, t2)
    w3 += _einsum("kc,kb,jc->bj", fov, t1, t1)
    w4 = fov.copy()

    ovvv = eris.ovvv
    vvov = ovvv.transpose(2, 3, 0, 1)
    tau = t2 + _einsum("ia,jb->ijab", t1, t1)
    woooo = np.zeros((nocc,) * 4)
    wvooo = np.zeros((nvir, nocc, nocc, nocc))

    w1 += _einsum("jcba,jc->ba", ovvv, t1 * 2)
    w1 -= _einsum("jabc,jc->ba", ovvv, t1)
    theta = t2 * 2 - t2.transpose(1, 0, 2, 3)
    w3 += _einsum("jkcd,kdcb->bj", theta, ovvv)
    wVOov = _einsum("jbcd,kd->bjkc", ovvv, t1)
    wvOOv = _einsum("cbjd,kd->cjkb", vvov, -t1)
    g2vovv = vvov.transpose(0, 2, 1, 3) * 2 - vvov.transpose(0, 2, 3, 1)
    wvooo += _einsum("cibd,jkbd->ckij", g2vovv, tau)
    wvvov = _einsum("jabd,jkcd->abkc", ovvv, t2) * -1.5
    wvvov += vvov.transpose(0, 3, 2, 1) * 2
    wvvov -= vvov
    g2vvov = vvov * 2 - ovvv.transpose(1, 2, 0, 3)
    theta = t2 * 2 - t2.transpose(0, 1, 3, 2)
    vackb = _einsum("acjd,kjbd->ackb", g2vvov, theta)
    wvvov += vackb.transpose(0, 3, 2, 1)
    wvvov -= vackb * 0.5

    ovoo = eris.ovoo
    w2 += _einsum("kbij,kb->ij", ovoo, t1) * 2
    w2 -= _einsum("ibkj,kb->ij", ovoo, t1)
    theta = t2.transpose(1, 0, 2, 3) * 2 - t2
    w3 -= _einsum("lckj,klcb->bj", ovoo, theta)
    tmp = _einsum("lc,jcik->ijkl", t1, ovoo)
    woooo += tmp
    woooo += tmp.transpose(1, 0, 3, 2)
    wvOOv += _einsum("lbjk,lc->bjkc", ovoo, t1)
    wVOov -= _einsum("jbkl,lc->bjkc", ovoo, t1)
    wvooo += ovoo.transpose(1, 3, 2, 0) * 2
    wvooo -= ovoo.transpose(1, 0, 2, 3)
    wvooo -= _einsum("klbc,iblj->ckij", t2, ovoo * 1.5)
    g2ovoo = ovoo * 2 - ovoo.transpose(2, 1, 0, 3)
    theta = t2 * 2 - t2.transpose(1, 0, 2, 3)
    vcjik = _einsum("jlcb,lbki->cjki", theta, g2ovoo)
    wvooo += vcjik.transpose(0, 3, 2, 1)
    wvooo -= vcjik * 0.5

    voov = eris.ovvo.transpose(1, 0, 3, 2)
    woooo += _einsum("cijd,klcd->ijkl", voov, tau)
    g2voov = voov * 2 - voov.transpose(0, 2, 1, 3)
    tmpw4 = _einsum("ckld,ld->kc", g2voov, t1)
    w1 -= _einsum("ckja,kjcb->ba", g2voov, t2)
    w1 -= _einsum("ja,jb->ba", tmpw4, t1)
    w2 += _einsum("jkbc,bikc->ij", t2, g2voov)
    w2 += _einsum("ib,jb->ij", tmpw4, t1)
    w3 += reduce(np.dot, (t1.T, tmpw4, t1.T))
    w4 += tmpw4
    wvOOv += _einsum("bljd,kd,lc->bjkc", voov, t1, t1)
    wVOov -= _einsum("bjld,kd,lc->bjkc", voov, t1, t1)
    VOov = _einsum("bjld,klcd->bjkc", g2voov, t2)
    VOov -= _einsum("bjld,kldc->bjkc", voov, t2)
    VOov += voov
    vOOv = _einsum("bljd,kldc->bjkc", voov, t2)
    vOOv -= eris.oovv.transpose(2, 1, 0, 3)
    wVOov += VOov
    wvOOv += vOOv
    ov1 = vOOv * 2 + VOov
    ov2 = VOov * 2 + vOOv
    wvooo -= _einsum("jb,bikc->ckij", t1, ov1)
    wvooo += _einsum("kb,bijc->ckij", t1, ov2)
    w3 += _einsum("ckjb,kc->bj", ov2, t1)
    wvvov += _einsum("ajkc,jb->abkc", ov1, t1)
    wvvov -= _einsum("ajkb,jc->abkc", ov2, t1)
    wvvov += _einsum("laki,klbc->abic", g2ovoo, tau)

    woooo += eris.oooo.transpose(0, 2, 1, 3)
    w3 += _einsum("bc,jc->bj", w1, t1)
    w3 -= _einsum("kj,kb->bj", w2, t1)
    imds = IMDS()
    imds.w1, imds.w2, imds.w3, imds.w4 = w1, w2, w3, w4
    imds.woooo, imds.wvooo, imds.wVOov, imds.wvOOv, imds.wvvov = woooo, wvooo, wVOov, wvOOv, wvvov
    return imds


def update_lambda(t1, t2, l1, l2, eris, imds):
    """gpu_ccsd_lambda.py:212-337 with one virtual block."""
    nocc, nvir = t1.shape
    fov = eris.fock[:nocc, nocc:]
    e_o, e_v = eris.mo_energy[:nocc], eris.mo_energy[nocc:]
    theta = t2 * 2 - t2.transpose(0, 1, 3, 2)
    mba = _einsum("klca,klcb->ba", l2, theta)
    mij = _einsum("ikcd,jkcd->ij", l2, theta)
    mba1 = _einsum("jc,jb->bc", l1, t1) + mba
    mij1 = _einsum("kb,jb->kj", l1, t1) + mij
    mia1 = t1 + _einsum("kc,jkbc->jb", l1, t2) * 2
    mia1 -= _einsum("kc,jkcb->jb", l1, t2)
    mia1 -= reduce(np.dot, (t1, l1.T, t1))
    mia1 -= _einsum("bd,jd->jb", mba, t1)
    mia1 -= _einsum("lj,lb->jb", mij, t1)

    l2new = _einsum("ijcd,acbd->ijab", l2, eris.vvvv)
    l1new = _einsum("ijab,jb->ia", l2new, t1) * 2
    l1new -= _einsum("jiab,jb->ia", l2new, t1)
    l2new *= 0.5

    w1 = imds.w1 - np.diag(e_v)
    w2 = imds.w2 - np.diag(e_o)
    l1new += fov
    l1new += _einsum("ib,ba->ia", l1, w1)
    l1new -= _einsum("ja,ij->ia", l1, w2)
    l1new -= _einsum("ik,ka->ia", mij, imds.w4)
    l1new -= _einsum("ca,ic->ia", mba, imds.w4)
    l1new += _einsum("ijab,bj->ia", l2, imds.w3) * 2
    l1new -= _einsum("ijba,bj->ia", l2, imds.w3)
    l2new += _einsum("ia,jb->ijab", l1, imds.w4)
    l2new += _einsum("jibc,ca->jiba", l2, w1)
    l2new -= _einsum("jk,kiba->jiba", w2, l2)

    ovoo = eris.ovoo
    l1new -= _einsum("iajk,kj->ia", ovoo, mij1) * 2
    l1new += _einsum("jaik,kj->ia", ovoo, mij1)
    l2new -= _einsum("jbki,ka->jiba", ovoo, l1)

    tau = t2 + _einsum("ia,jb->ijab", t1, t1)
    l2tau = _einsum("ijcd,klcd->ijkl", l2, tau)
    l2t1 = _einsum("jidc,kc->ijkd", l2, t1)

    l1new -= _einsum("jb,jiab->ia", l1, eris.oovv)
    ovvv = eris.ovvv
    l1new += _einsum("iabc,bc->ia", ovvv, mba1) * 2
    l1new -= _einsum("ibca,bc->ia", ovvv, mba1)
    l2new += _einsum("jbac,ic->jiba", ovvv, l1)
    m4 = _einsum("ijkd,kadb->ijab", l2t1, ovvv)
    l2new -= m4
    l1new -= _einsum("ijab,jb->ia", m4, t1) * 2
    l1new -= _einsum("ijab,ia->jb", m4, t1) * 2
    l1new += _einsum("jiab,jb->ia", m4, t1)
    l1new += _einsum("jiab,ia->jb", m4, t1)

    voov = eris.ovvo.transpose(1, 0, 3, 2)
    l1new += _einsum("jb,aijb->ia", l1, voov) * 2
    l2new += voov.transpose(1, 2, 0, 3) * 0.5
    l2new -= _einsum("bjic,ca->jiba", voov, mba1)
    l2new -= _einsum("bjka,ik->jiba", voov, mij1)
    l1new += _einsum("aijb,jb->ia", voov, mia1) * 2
    l1new -= _einsum("bija,jb->ia", voov, mia1)
    m4 = _einsum("ijkl,aklb->ijab", l2tau, voov)
    l2new += m4 * 0.5
    l1new += _einsum("ijab,jb->ia", m4, t1) * 2
    l1new -= _einsum("ijba,jb->ia", m4, t1)

    l1new -= _einsum("ckij,jkca->ia", imds.wvooo, l2)
    l1new += _einsum("abkc,kibc->ia", imds.wvvov, l2)
    tmp_voov = imds.wVOov * 2 + imds.wvOOv
    tmp = l2.transpose(0, 2, 1, 3) - l2.transpose(0, 3, 1, 2) * 0.5
    l2new += _einsum("iakc,bjkc->jiba", tmp, tmp_voov)
    tmp = _einsum("jkca,bikc->jiba", l2, imds.wvOOv)
    l2new += tmp
    l2new += tmp.transpose(1, 0, 2, 3) * 0.5

    m3 = _einsum("ijkl,klab->ijab", imds.woooo, l2)
    l2new += m3 * 0.5
    l1new += _einsum("ijab,jb->ia", m3, t1) * 2
    l1new -= _einsum("ijba,jb->ia", m3, t1)

    eia = e_o[:, None] - e_v
    l1new /= eia
    # l2 symmetrization and division (gpu_ccsd_lambda.py:325-334), vectorized
    l2new = l2new + l2new.transpose(1, 0, 3, 2)
    l2new /= eia[:, None, :, None] + eia[None, :, None, :]
    return l1new, l2new


def solve_lambda(t1, t2, eris, max_cycle=200, tol=1e-10):
    """gpu_ccsd_lambda.py:340-408: start from l = t, DIIS, converge on ||dl||."""
    nocc, nvir = t1.shape
    imds = make_intermediates(t1, t2, eris)
    l1, l2 = t1.copy(), t2.copy()
    diis = DIIS(6)
    conv = False
    for it in range(max_cycle):
        l1n, l2n = update_lambda(t1, t2, l1, l2, eris, imds)
        norm = np.linalg.norm(_pack(l1n, l2n) - _pack(l1, l2))
        l1, l2 = _unpack(diis.update(_pack(l1n, l2n)), nocc, nvir)
        if norm < tol:
            conv = True
            break
    return conv, l1, l2


# ---------------------------------------------------------------------------
# RDMs (restates ccsd_gpu/gpu_ccsd_grad.py:101-336)
# ---------------------------------------------------------------------------
def gamma1_intermediates(t1, t2, l1, l2):
    doo = -_einsum("ja,ia->ij", t1, l1)
    dvv = _einsum("ia,ib->ab", t1, l1)
    xtv = _einsum("ie,me->im", t1, l1)
    dvo = t1.T - _einsum("im,ma->ai", xtv, t1)
    theta = t2 * 2 - t2.transpose(0, 1, 3, 2)
    doo -= _einsum("jkab,ikab->ij", theta, l2)
    dvv += _einsum("jica,jicb->ab", theta, l2)
    xt1 = _einsum("mnef,inef->mi", l2, theta)
    xt2 = _einsum("mnaf,mnef->ea", l2, theta)
    dvo += _einsum("imae,me->ai", theta, l1)
    dvo -= _einsum("mi,ma->ai", xt1, t1)
    dvo -= _einsum("ie,ae->ai", t1, xt2)
    return doo, l1.copy(), dvo, dvv


def gamma2_intermediates(t1, t2, l1, l2):
    nocc, nvir = t1.shape
    pvOOv = _einsum("ikca,jkcb->aijb", l2, t2)
    moo = _einsum("dljd->jl", pvOOv) * 2
    mvv = _einsum("blld->db", pvOOv) * 2
    gooov = _einsum("kc,cija->jkia", t1, pvOOv)
    pvoOV = -_einsum("ikca,jkbc->aijb", l2, t2)
    theta = t2 * 2 - t2.transpose(0, 1, 3, 2)
    pvoOV += _einsum("ikac,jkbc->aijb", l2, theta)
    moo += _einsum("dljd->jl", pvoOV)
    mvv += _einsum("blld->db", pvoOV)
    gooov -= _einsum("jc,cika->jkia", t1, pvoOV)
    mia = _einsum("kc,ikac->ia", l1, t2) * 2 - _einsum("kc,ikca->ia", l1, t2)
    mab = _einsum("kc,kb->cb", l1, t1)
    mij = _einsum("kc,jc->jk", l1, t1) + moo * 0.5
    tau = _einsum("ia,jb->ijab", t1, t1) + t2
    goooo = _einsum("ijab,klab->ijkl", tau, l2) * 0.5
    doooo = goooo.transpose(0, 2, 1, 3) * 2 - goooo.transpose(0, 3, 1, 2)
    gooov += _einsum("ji,ka->jkia", -0.5 * moo, t1)
    gooov += _einsum("la,jkil->jkia", 2 * t1, goooo)
    gooov -= _einsum("ib,jkba->jkia", l1, tau)
    gooov -= _einsum("jkba,ib->jkia", l2, t1)
    dooov = gooov.transpose(0, 2, 1, 3) * 2 - gooov.transpose(1, 2, 0, 3)

    goovv = _einsum("ia,jb->ijab", mia, t1)
    tmp = _einsum("ijkl,klab->ijab", goooo, tau)
    tmp -= _einsum("jk,ikab->ijab", mij, tau)
    tmp -= _einsum("cb,ijac->ijab", mab, t2)
    tmp -= _einsum("bd,ijad->ijab", mvv * 0.5, tau)
    tmp += 0.5 * tau
    tmp += 0.5 * l2
    goovv += tmp
    gOvvO = _einsum("kiac,jc,kb->iabj", l2, t1, t1)
    gOvvO += _einsum("aijb->iabj", pvOOv)
    govVO = _einsum("ia,jb->iabj", l1, t1)
    govVO -= _einsum("ikac,jc,kb->iabj", l2, t1, t1)
    govVO += _einsum("aijb->iabj", pvoOV)
    dovvo = 2 * govVO + gOvvO
    doovv = (-2 * gOvvO - govVO).transpose(3, 0, 1, 2)
    tau2 = tau - t2 * 0.5
    goovv += _einsum("dlib,jlda->ijab", pvOOv, tau2)
    goovv -= _einsum("dlia,jldb->ijab", pvoOV, tau2)
    tmp = pvoOV + pvOOv * 0.5
    goovv += _einsum("dlia,jlbd->ijab", tmp, t2)
    dovov = goovv.transpose(0, 2, 1, 3) * 2 - goovv.transpose(1, 2, 0, 3)

    gvvvv = _einsum("ijab,ijcd->abcd", l2, t2)
    jabc = _einsum("ijab,ic->jabc", l2, t1)
    gvvvv += _einsum("jabc,jd->abcd", jabc, t1)
    vvv = gvvvv.transpose(0, 2, 1, 3)                 # dvvvv[i] = gvvvv[i].T(1,0,2) per slice
    dvvvv = vvv - vvv.transpose(0, 3, 2, 1) * 0.5
    gvovv = _einsum("adbc,id->aibc", gvvvv, -t1)
    gvovv += _einsum("akic,kb->aibc", pvoOV, t1)
    gvovv -= _einsum("akib,kc->aibc", pvOOv, t1)
    gvovv += _einsum("ja,jibc->aibc", l1, t2)
    gvovv += _einsum("ja,jb,ic->aibc", l1, t1, t1)
    gvovv += _einsum("ba,ic->aibc", mvv * 0.5, t1)
    gvovv += _einsum("ja,jibc->aibc", t1, l2)
    dovvv = gvovv.transpose(1, 3, 0, 2) * 2 - gvovv.transpose(1, 2, 0, 3)
    return dovov, dvvvv, doooo, doovv, dovvo, None, dovvv, dooov


def assemble_dm2_mo(d2):
    """Full MO 2-RDM (chemist's order) without the dm1 part, gpu_ccsd_grad.py:271-311."""
    dovov, dvvvv, doooo, doovv, dovvo, _, dovvv, dooov = d2
    nocc, nvir = dovov.shape[:2]
    nmo = nocc + nvir
    o, v = slice(0, nocc), slice(nocc, None)
    dm2 = np.zeros((nmo,) * 4)
    dm2[o, v, o, v] = dovov + dovov.transpose(2, 3, 0, 1)
    dm2[v, o, v, o] = dm2[o, v, o, v].transpose(1, 0, 3, 2)
    dm2[o, o, v, v] = doovv + doovv.transpose(1, 0, 3, 2)
    dm2[v, v, o, o] = dm2[o, o, v, v].transpose(2, 3, 0, 1)
    dm2[o, v, v, o] = dovvo + dovvo.transpose(3, 2, 1, 0)
    dm2[v, o, o, v] = dm2[o, v, v, o].transpose(1, 0, 3, 2)
    dm2[v, v, v, v] = (dvvvv + dvvvv.transpose(1, 0, 3, 2)) * 2
    dm2[o, o, o, o] = (doooo + doooo.transpose(1, 0, 3, 2)) * 2
    dm2[o, v, v, v] = dovvv
    dm2[v, v, o, v] = dovvv.transpose(2, 3, 0, 1)
    dm2[v, v, v, o] = dovvv.transpose(3, 2, 1, 0)
    dm2[v, o, v, v] = dovvv.transpose(1, 0, 3, 2)
    dm2[o, o, o, v] = dooov
    dm2[o, v, o, o] = dooov.transpose(2, 3, 0, 1)
    dm2[o, o, v, o] = dooov.transpose(1, 0, 3, 2)
    dm2[v, o, o, o] = dooov.transpose(3, 2, 1, 0)
    return dm2


def make_rdm1_mo(d1):
    """Full CCSD 1-RDM in the MO basis (PySCF _make_rdm1 convention, HF part included)."""
    doo, dov, dvo, dvv = d1
    nocc, nvir = dov.shape
    dm1 = np.empty((nocc + nvir,) * 2)
    dm1[:nocc, :nocc] = doo + doo.T
    dm1[:nocc, nocc:] = dov + dvo.T
    dm1[nocc:, :nocc] = dm1[:nocc, nocc:].T
    dm1[nocc:, nocc:] = dvv + dvv.T
    dm1[np.diag_indices(nocc)] += 2
    return dm1


def make_rdm2_mo(d1, d2):
    """Full CCSD 2-RDM, chemist's order dm2[p,q,r,s] = <p+ r+ s q>, HF and dm1 parts included."""
    nocc = d1[0].shape[0]
    dm2 = assemble_dm2_mo(d2)
    dm1 = make_rdm1_mo(d1)
    dm1[np.diag_indices(nocc)] -= 2
    for i in range(nocc):
        dm2[i, i, :, :] += dm1 * 2
        dm2[:, :, i, i] += dm1 * 2
        dm2[:, i, i, :] -= dm1
        dm2[i, :, :, i] -= dm1.T
    for i in range(nocc):
        for j in range(nocc):
            dm2[i, i, j, j] += 4
            dm2[i, j, j, i] -= 2
    return dm2.transpose(1, 0, 3, 2)


def rdm2_ao(d2, mo_coeff):
    """MO->AO transform + ij/kl symmetrization (gpu_ccsd_grad.py:313-326), full (nao^4)."""
    dm2 = assemble_dm2_mo(d2).transpose(1, 0, 3, 2)
    c = mo_coeff
    dm2 = _einsum("pqrs,ip,jq,kr,ls->ijkl", dm2, c, c, c, c, optimize=True)
    dm2 = dm2 + dm2.transpose(1, 0, 2, 3)
    dm2 = dm2 + dm2.transpose(0, 1, 3, 2)
    return dm2 * 0.5


# ---------------------------------------------------------------------------
# Gradient (restates ccsd_gpu/gpu_ccsd_grad.py:344-586, 625-666)
# ---------------------------------------------------------------------------
def grad_nuc(mol):
    z = mol.atom_charges()
    r = mol.atom_coords()
    g = np.zeros((mol.natm, 3))
    for a in range(mol.natm):
        for b in range(mol.natm):
            if a != b:
                d = r[a] - r[b]
                g[a] -= z[a] * z[b] * d / np.linalg.norm(d) ** 3
    return g


def cphf_solve(fvind, mo_energy, mo_occ, h1):
    """Solve (e_a - e_i) x_ai + fvind(x)_ai = -h1_ai exactly (cphf.solve without s1)."""
    e_a = mo_energy[mo_occ == 0]
    e_i = mo_energy[mo_occ > 0]
    nv, no = h1.shape
    n = nv * no
    A = np.zeros((n, n))
    for k in range(n):
        x = np.zeros(n)
        x[k] = 1.0
        A[:, k] = fvind(x.reshape(nv, no)).ravel()
    A += np.diag((e_a[:, None] - e_i[None, :]).ravel())
    return np.linalg.solve(A, -h1.ravel()).reshape(nv, no)


def grad_elec(mycc, t1, t2, l1, l2, return_parts=False):
    mol = mycc.mol
    mf = mycc._scf
    mo_coeff, mo_energy, mo_occ = mycc.mo_coeff, mf.mo_energy, mycc.mo_occ
    nao, nmo = mo_coeff.shape
    nocc = int((mo_occ > 0).sum())
    # frozen / active orbital index sets (gpu_ccsd_grad.py:373-377)
    mask = mycc.get_frozen_mask() if hasattr(mycc, "get_frozen_mask") else np.ones(nmo, dtype=bool)
    with_frozen = not mask.all()
    OA = np.nonzero(mask & (mo_occ > 0))[0]
    VA = np.nonzero(mask & (mo_occ == 0))[0]
    OF = np.nonzero(~mask & (mo_occ > 0))[0]
    VF = np.nonzero(~mask & (mo_occ == 0))[0]
    d1 = gamma1_intermediates(t1, t2, l1, l2)
    doo, dov, dvo, dvv = d1
    d2 = gamma2_intermediates(t1, t2, l1, l2)
    dm2ao = rdm2_ao(d2, mo_coeff[:, np.hstack((OA, VA))])   # (nao^4), symmetrized
    hf_dm1 = mf.make_rdm1(mo_coeff, mo_occ)
    eri = mf.eri()
    ip1 = oi.ip1_full(mol)                               # (3, i, j, k, l)
    aoa = mol.ao_atom_index()
    # Phase 3: Imat and the 2e derivative (full-tensor form of the block loop)
    # Imat[p,q] = sum_{j, k>=l} (pj|kl) G'(qj|kl), G' with halved kl diagonal
    Imat = 0.5 * _einsum("ijkl,qjkl->iq", eri, dm2ao)
    de = np.zeros((mol.natm, 3))
    per_ao = _einsum("xijkl,ijkl->xi", ip1, dm2ao)
    np.add.at(de, aoa, -per_ao.T)                          # = -2 sum_{i in A, j, k>=l} ip1 G'
    # HF separable part vhf1[A] (C1 / X2 / C3 / X4) contracted later with dm1p
    vhf1 = np.zeros((mol.natm, 3, nao, nao))
    for a in range(mol.natm):
        m = (aoa == a).astype(float)
        e = ip1 * m[None, :, None, None, None]
        vj1 = _einsum("xijkl,ij->xkl", e, hf_dm1)        # C1
        vk1 = _einsum("xijkl,il->xkj", e, hf_dm1)        # X2
        vj2 = _einsum("xijkl,kl->xij", e, hf_dm1)        # C3
        vk2 = _einsum("xijkl,jk->xil", e, hf_dm1)        # X4
        vhf1[a] = vj1 - 0.5 * vk1 + vj2 - 0.5 * vk2
    # Phase 4
    Imat = reduce(np.dot, (mo_coeff.T, Imat, mf.get_ovlp(), mo_coeff)) * -1
    dm1mo = np.zeros((nmo, nmo))
    if with_frozen:
        # frozen-core / frozen-virtual rotations (gpu_ccsd_grad.py:510-518)
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

    def fvind(x):
        dm = reduce(np.dot, (mo_coeff[:, nocc:], x, mo_coeff[:, :nocc].T))
        v = mf.get_veff(mol, dm + dm.T)
        return reduce(np.dot, (mo_coeff[:, nocc:].T, v, mo_coeff[:, :nocc])) * 2

    dvo_resp = cphf_solve(fvind, mo_energy, mo_occ, Xvo)
    dm1mo[nocc:, :nocc] += dvo_resp
    dm1mo[:nocc, nocc:] += dvo_resp.T
    Imat[nocc:, :nocc] = Imat[:nocc, nocc:].T
    im1 = reduce(np.dot, (mo_coeff, Imat, mo_coeff.T))
    # Phase 5
    s1 = -oi.int1e(mol, "ipovlp")
    h1 = -(oi.int1e(mol, "ipkin") + oi.int1e(mol, "ipnuc_atom").sum(axis=0))
    vrinv_all = oi.int1e(mol, "ipnuc_atom")
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
    offs = mol.offset_nr_by_atom()
    for k in range(mol.natm):
        p0, p1_ = offs[k, 2], offs[k, 3]
        vrinv = vrinv_all[k].copy()
        vrinv[:, p0:p1_] += h1[:, p0:p1_]
        h1ao = vrinv + vrinv.transpose(0, 2, 1)
        de[k] += _einsum("xij,ij->x", s1[:, p0:p1_], im1[p0:p1_])
        de[k] += _einsum("xji,ij->x", s1[:, p0:p1_], im1[:, p0:p1_])
        de[k] += _einsum("xij,ji->x", h1ao, dm1)
        de[k] -= _einsum("xij,ij->x", s1[:, p0:p1_], zeta[p0:p1_])
        de[k] -= _einsum("xji,ij->x", s1[:, p0:p1_], zeta[:, p0:p1_])
        de[k] -= _einsum("xij,ij->x", s1[:, p0:p1_], vhf_s1occ[p0:p1_]) * 2
        de[k] -= _einsum("xij,ij->x", vhf1[k], dm1p)
    if return_parts:
        return de, {"dm2ao": dm2ao, "Imat": Imat, "dm1mo": dm1mo}
    return de


def ccsd_total_energy(mol, frozen=None):
    mf = RHF(mol)
    mf.kernel()
    cc = CCSD(mf, frozen=frozen)
    cc.kernel()
    return cc.e_tot, mf, cc
