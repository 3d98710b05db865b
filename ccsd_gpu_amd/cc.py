"""Closed-shell CCSD on the GPU: MO integrals and the T-amplitude solve.

Producer of (t1, t2, eris) for the lambda / gradient path. The reference takes
these from gpu4pyscf or PySCF (examples/h2o_ccsd_grad.py:128-151,
ccsd_gpu/gpu_ccsd_lambda.py:352-353). MO integrals follow PySCF's
_ChemistsERIs layout (chemist's notation blocks oooo, ovoo, ovvo, oovv, ovov,
ovvv, vvvv, fock, mo_energy); the amplitude equations are the standard
closed-shell spin-adapted CCSD equations with Fov/Foo/Fvv, Loo/Lvv and
W-intermediates. Tensors are float64 torch tensors on the HIP device.
"""
import ctypes

import numpy as np
import torch

from .precision import contract, static_operands
from .scf import DEV


def E(eq, *ops):
    return contract(eq, *ops)


def frozen_mask(nmo, frozen):
    """Active-orbital mask of PySCF's `frozen` convention: None / 0 -> all active; an
    int n -> the n lowest orbitals frozen; a sequence -> those orbital indices frozen."""
    mask = np.ones(nmo, dtype=bool)
    if frozen is None or (np.isscalar(frozen) and int(frozen) == 0):
        return mask
    if np.isscalar(frozen):
        n = int(frozen)
        if n < 0 or n > nmo:
            raise ValueError(f"frozen={frozen} out of range for {nmo} orbitals")
        mask[:n] = False
        return mask
    idx = np.asarray(list(frozen), dtype=np.int64)
    if idx.size and (idx.min() < 0 or idx.max() >= nmo):
        raise ValueError(f"frozen orbital indices {idx.tolist()} out of range for {nmo} orbitals")
    mask[idx] = False
    return mask


class ERIs:
    """MO integrals on the device over the active orbitals, in the layout of PySCF's
    _ChemistsERIs (chemist's-notation blocks, fock, mo_energy), built from the
    Cartesian AO ERIs. The Fock matrix is that of the HF density of ALL occupied
    orbitals (frozen ones included), projected on the active orbitals, as PySCF's
    CCSD eris hold it."""

    def __init__(self, mf, mo_coeff=None, mask=None, ints=None):
        from .scf import get_ints
        ints = get_ints(mf) if ints is None else ints
        c_all = torch.as_tensor(np.asarray(mf.mo_coeff if mo_coeff is None else mo_coeff), dtype=torch.float64,
                                device=DEV)
        occ = np.asarray(mf.mo_occ, dtype=np.float64)
        mask = np.ones(c_all.shape[1], dtype=bool) if mask is None else np.asarray(mask, dtype=bool)
        act = torch.as_tensor(np.nonzero(mask)[0], device=DEV)
        c = c_all[:, act]
        nocc = int((occ[mask] > 0).sum())
        self.nocc = nocc
        cc = ints.c2s @ c                                   # Cartesian AO -> active MO
        eri = torch.tensordot(ints.eri_c, cc, dims=([3], [0]))              # (p q r l)
        eri = torch.tensordot(eri, cc, dims=([2], [0])).permute(0, 1, 3, 2)  # (p q k l)
        eri = torch.tensordot(cc, eri, dims=([0], [1]))                      # (j p k l)
        eri = torch.tensordot(cc, eri, dims=([0], [1]))                      # (i j k l)
        eri = eri.contiguous()
        occ_t = torch.as_tensor(occ, dtype=torch.float64, device=DEV)
        dm = (c_all * occ_t) @ c_all.T
        j, k = ints.get_jk(dm)
        fock_ao = ints.T + ints.V + j - 0.5 * k
        self.fock = c.T @ fock_ao @ c
        self.mo_energy = torch.diagonal(self.fock).clone()
        self._set_blocks(eri, nocc)
        del eri

    def _set_blocks(self, eri, nocc):
        o, v = slice(0, nocc), slice(nocc, None)
        self.oooo = eri[o, o, o, o].contiguous()
        self.ovoo = eri[o, v, o, o].contiguous()
        self.ovvo = eri[o, v, v, o].contiguous()
        self.oovv = eri[o, o, v, v].contiguous()
        self.ovov = eri[o, v, o, v].contiguous()
        self.ovvv = eri[o, v, v, v].contiguous()
        self.vvvv = eri[v, v, v, v].contiguous()

    def get_ovvv(self, *slices):
        return self.ovvv[slices] if slices else self.ovvv

    @classmethod
    def from_blocks(cls, eris, nocc, nvir, vvvv=None):
        """Device copy of a PySCF-like eris object (NumPy blocks, the reference's input,
        ccsd_gpu/gpu_ccsd_lambda.py:73-103, 223-225): ovvv through get_ovvv (PySCF keeps
        it pair-packed over its last two indices), vvvv full, 4-fold packed
        (npair, npair) or taken from `vvvv` when the object holds none."""
        self = cls.__new__(cls)
        self.nocc = nocc

        def dev(a):
            return torch.as_tensor(np.asarray(a), dtype=torch.float64, device=DEV).contiguous()

        self.fock = dev(eris.fock)
        me = getattr(eris, "mo_energy", None)
        self.mo_energy = dev(me) if me is not None else torch.diagonal(self.fock).clone()
        self.oooo, self.ovoo, self.ovvo = dev(eris.oooo), dev(eris.ovoo), dev(eris.ovvo)
        self.oovv, self.ovov = dev(eris.oovv), dev(eris.ovov)
        get = getattr(eris, "get_ovvv", None)
        ovvv = dev(get() if callable(get) else eris.ovvv)
        if ovvv.dim() == 3:
            ovvv = _unpack_pairs_last(ovvv, nvir)
        self.ovvv = ovvv.reshape(nocc, nvir, nvir, nvir).contiguous()
        raw = getattr(eris, "vvvv", None)
        if raw is None:
            if vvvv is None:
                raise ValueError("eris holds no vvvv block and none was supplied")
            self.vvvv = vvvv
        else:
            w = dev(raw)
            if w.dim() == 2:
                w = _unpack_pairs_last(w, nvir)                                    # [(ab), c, d]
                w = _unpack_pairs_last(w.permute(1, 2, 0).contiguous(), nvir)      # [c, d, a, b]
                w = w.permute(2, 3, 0, 1)
            self.vvvv = w.reshape(nvir, nvir, nvir, nvir).contiguous()
        return self


def _unpack_pairs_last(x, n):
    """(..., n(n+1)/2) lower-triangle packed (PySCF's lib.pack_tril order) -> (..., n, n)."""
    a, b = torch.tril_indices(n, n, device=x.device)
    out = torch.empty(x.shape[:-1] + (n, n), dtype=x.dtype, device=x.device)
    out[..., a, b] = x
    out[..., b, a] = x
    return out


def device_eris(mycc, eris=None):
    """The MO integrals a lambda solve or gradient runs on, as device tensors. `eris` of
    this build is used as it is; a PySCF-like eris object (NumPy blocks) is copied to
    the device; with none, mycc.ao2mo() when mycc is this build's CCSD, else the
    integrals are built here from mycc's orbitals and the device AO integrals (a
    PySCF CCSD's ao2mo would run on the host)."""
    if isinstance(eris, ERIs):
        return eris
    if eris is None and isinstance(mycc, CCSD):
        return mycc.ao2mo()
    mask = _frozen_mask_of(mycc)
    if eris is None:
        return ERIs(mycc._scf, mycc.mo_coeff, mask)
    occ = np.asarray(mycc.mo_occ)[mask]
    nocc = int((occ > 0).sum())
    nvir = int(mask.sum()) - nocc
    vvvv = None
    if getattr(eris, "vvvv", None) is None:
        vvvv = ERIs(mycc._scf, mycc.mo_coeff, mask).vvvv
    return ERIs.from_blocks(eris, nocc, nvir, vvvv)


def _frozen_mask_of(mycc):
    get = getattr(mycc, "get_frozen_mask", None)
    nmo = np.asarray(mycc.mo_coeff).shape[1] if not torch.is_tensor(mycc.mo_coeff) else mycc.mo_coeff.shape[1]
    if callable(get):
        return np.asarray(get(), dtype=bool)
    return frozen_mask(nmo, getattr(mycc, "frozen", None))


def ccsd_energy(t1, t2, eris):
    nocc = t1.shape[0]
    e = 2 * torch.sum(eris.fock[:nocc, nocc:] * t1)
    tau = t2 + E("ia,jb->ijab", t1, t1)
    e = e + 2 * E("ijab,iajb->", tau, eris.ovov) - E("ijab,ibja->", tau, eris.ovov)
    return float(e)


def update_amps(t1, t2, eris):
    nocc, nvir = t1.shape
    fock = eris.fock
    e_o, e_v = eris.mo_energy[:nocc], eris.mo_energy[nocc:]
    fov, foo, fvv = fock[:nocc, nocc:], fock[:nocc, :nocc], fock[nocc:, nocc:]
    ovov, ovoo, ovvo, oovv, ovvv = eris.ovov, eris.ovoo, eris.ovvo, eris.oovv, eris.ovvv
    t1t1 = E("ic,ld->icld", t1, t1)
    Foo = foo + 2 * E("kcld,ilcd->ki", ovov, t2) - E("kdlc,ilcd->ki", ovov, t2) \
        + 2 * E("kcld,icld->ki", ovov, t1t1) - E("kdlc,icld->ki", ovov, t1t1)
    Fvv = fvv - 2 * E("kcld,klad->ac", ovov, t2) + E("kdlc,klad->ac", ovov, t2) \
        - 2 * E("kcld,kald->ac", ovov, t1t1) + E("kdlc,kald->ac", ovov, t1t1)
    Fov = fov + 2 * E("kcld,ld->kc", ovov, t1) - E("kdlc,ld->kc", ovov, t1)
    Loo = Foo + E("kc,ic->ki", fov, t1) + 2 * E("lcki,lc->ki", ovoo, t1) - E("kcli,lc->ki", ovoo, t1)
    Lvv = Fvv - E("kc,ka->ac", fov, t1) + 2 * E("kdac,kd->ac", ovvv, t1) - E("kcad,kd->ac", ovvv, t1)
    Foo = Foo - torch.diag(e_o)
    Fvv = Fvv - torch.diag(e_v)
    Loo = Loo - torch.diag(e_o)
    Lvv = Lvv - torch.diag(e_v)

    t1n = -2 * E("kc,ka,ic->ia", fov, t1, t1)
    t1n = t1n + E("ac,ic->ia", Fvv, t1) - E("ki,ka->ia", Foo, t1)
    t1n = t1n + 2 * E("kc,kica->ia", Fov, t2) - E("kc,ikca->ia", Fov, t2)
    t1n = t1n + E("kc,ic,ka->ia", Fov, t1, t1) + fov
    t1n = t1n + 2 * E("kcai,kc->ia", ovvo, t1) - E("kiac,kc->ia", oovv, t1)
    t1n = t1n + 2 * E("kdac,ikcd->ia", ovvv, t2) - E("kcad,ikcd->ia", ovvv, t2)
    tmp = E("kd,ic->kdic", t1, t1)
    t1n = t1n + 2 * E("kdac,kdic->ia", ovvv, tmp) - E("kcad,kdic->ia", ovvv, tmp)
    t1n = t1n - 2 * E("lcki,klac->ia", ovoo, t2) + E("kcli,klac->ia", ovoo, t2)
    tmp = E("lc,ka->lcka", t1, t1)
    t1n = t1n - 2 * E("lcki,lcka->ia", ovoo, tmp) + E("kcli,lcka->ia", ovoo, tmp)

    tmp2 = E("kibc,ka->abic", oovv, -t1) + ovvv.permute(1, 3, 0, 2)
    tmp = E("abic,jc->ijab", tmp2, t1)
    t2n = tmp + tmp.permute(1, 0, 3, 2)
    tmp2 = E("kcai,jc->akij", ovvo, t1) + ovoo.permute(1, 3, 0, 2)
    tmp = E("akij,kb->ijab", tmp2, t1)
    t2n = t2n - (tmp + tmp.permute(1, 0, 3, 2))
    t2n = t2n + ovov.permute(0, 2, 1, 3)

    Woooo = E("lcki,jc->klij", ovoo, t1) + E("kclj,ic->klij", ovoo, t1) + E("kcld,ijcd->klij", ovov, t2) \
        + E("kcld,icjd->klij", ovov, E("ic,jd->icjd", t1, t1)) + eris.oooo.permute(0, 2, 1, 3)
    Wvvvv = -E("kdac,kb->abcd", ovvv, t1) - E("kcbd,ka->abcd", ovvv, t1) + eris.vvvv.permute(0, 2, 1, 3)
    Wvoov = E("kcad,id->akic", ovvv, t1) - E("kcli,la->akic", ovoo, t1) + ovvo.permute(2, 0, 3, 1) \
        - 0.5 * E("ldkc,ilda->akic", ovov, t2) - 0.5 * E("lckd,ilad->akic", ovov, t2) \
        - E("ldkc,idla->akic", ovov, E("id,la->idla", t1, t1)) + E("ldkc,ilad->akic", ovov, t2)
    Wvovo = E("kdac,id->akci", ovvv, t1) - E("lcki,la->akci", ovoo, t1) + oovv.permute(2, 0, 3, 1) \
        - 0.5 * E("lckd,ilda->akci", ovov, t2) - E("lckd,idla->akci", ovov, E("id,la->idla", t1, t1))

    tau = t2 + E("ia,jb->ijab", t1, t1)
    t2n = t2n + E("klij,klab->ijab", Woooo, tau)
    t2n = t2n + E("abcd,ijcd->ijab", Wvvvv, tau)
    del Wvvvv
    tmp = E("ac,ijcb->ijab", Lvv, t2)
    t2n = t2n + tmp + tmp.permute(1, 0, 3, 2)
    tmp = E("ki,kjab->ijab", Loo, t2)
    t2n = t2n - (tmp + tmp.permute(1, 0, 3, 2))
    tmp = 2 * E("akic,kjcb->ijab", Wvoov, t2) - E("akci,kjcb->ijab", Wvovo, t2)
    t2n = t2n + tmp + tmp.permute(1, 0, 3, 2)
    tmp = E("akic,kjbc->ijab", Wvoov, t2)
    t2n = t2n - (tmp + tmp.permute(1, 0, 3, 2))
    tmp = E("bkci,kjac->ijab", Wvovo, t2)
    t2n = t2n - (tmp + tmp.permute(1, 0, 3, 2))

    eia = e_o[:, None] - e_v
    t1n = t1n / eia
    t2n = t2n / (eia[:, None, :, None] + eia[None, :, None, :])
    return t1n, t2n


class DIIS:
    """Pulay DIIS on device vectors; error = difference of successive vectors.

    Vectors and errors live in two (space, N) ring buffers on the device; each
    update adds one row of the error overlap matrix (one GEMV, one host copy of
    `space` numbers) and extrapolates with one GEMV over the stored vectors.

    store="fp32" (float64 vectors on the device, space <= 15): the ring buffers hold
    FP32 copies, half the bytes that every update reads twice; one library pass stores
    the new vector and its error and forms the error's overlaps (summed in FP64), one
    forms the FP64 extrapolation. For low-precision solves, whose sweeps carry far
    larger rounding than the FP32 copies add; the vectors returned stay FP64."""

    def __init__(self, space=6, store=None):
        self.space = space
        self.store = store
        self.X = self.E = None
        self.B = np.zeros((space, space))
        self.count = 0
        self.prev = None
        self.res_norm = None        # |sum_s c_s E[s]| of the last extrapolation
        self.c_abs = None           # sum_s |c_s| of the last extrapolation
        self._fp32 = None

    def _alloc(self, n, like):
        if self.X is None:
            # one row more than the history: the first vector, kept as the first error's base
            self.X = torch.empty((self.space + 1, n), dtype=like.dtype, device=like.device)
            self.E = torch.empty((self.space, n), dtype=like.dtype, device=like.device)

    def update(self, x):
        if self.prev is None:
            self.prev = x
            return x
        self._alloc(x.numel(), x)
        self.X[self.count % self.space].copy_(x)
        return self._extrapolate()[0]

    def push(self, parts, rhs=None, start=None):
        """update() of x = concat(flattened parts) (+ rhs), written straight into the
        history buffer (no separate concatenated copy). Returns (extrapolated vector,
        ||x - previous vector||): the previous vector is the last one returned, or
        `start` on the first call (None: no norm)."""
        if self._fp32 is None:
            self._fp32 = (self.store == "fp32" and parts[0].is_cuda and parts[0].dtype == torch.float64
                          and len(parts) <= 2 and self.space <= 15)
        if self._fp32:
            return self._push32(parts, rhs, start)
        n = sum(p.numel() for p in parts)
        self._alloc(n, parts[0])
        first = self.prev is None
        dst = self.X[self.space] if first else self.X[self.count % self.space]
        o = 0
        for p in parts:
            seg = dst[o:o + p.numel()]
            if rhs is None:
                seg.copy_(p.reshape(-1))
            else:
                torch.add(p.reshape(-1), rhs[o:o + p.numel()], out=seg)
            o += p.numel()
        if first:
            self.prev = dst
            return dst, (None if start is None else float(torch.linalg.norm(dst - start)))
        return self._extrapolate()

    def _push32(self, parts, rhs, start):
        from . import _lib
        lib = _lib.load()
        a = parts[0].reshape(-1).contiguous()
        b = parts[1].reshape(-1).contiguous() if len(parts) > 1 else None
        n0 = a.numel()
        n = n0 + (b.numel() if b is not None else 0)
        dev = a.device
        if self.X is None:
            self.X = torch.empty((self.space + 1, n), dtype=torch.float32, device=dev)
            self.E = torch.empty((self.space, n), dtype=torch.float32, device=dev)
            self._ws = torch.empty(max(1, lib.ccsd_diis_workspace(n, self.space)), dtype=torch.float64, device=dev)
            self._dots = torch.empty(self.space + 1, dtype=torch.float64, device=dev)
        if rhs is not None:
            rhs = rhs.contiguous()
        ptr = lambda t: _lib.ptr(t) if t is not None else None  # noqa: E731
        stream = _lib.current_stream_ptr()
        if self.prev is None:
            # the first vector: its FP64 copy is returned (and is the next error's base)
            out = torch.empty(n, dtype=torch.float64, device=dev)
            base = start.contiguous() if start is not None else None
            _lib.check(lib.ccsd_diis_push(ptr(a), ptr(b), n0, n, ptr(rhs), ptr(base), ptr(self.X[self.space]), None,
                                          None, 0, 0, 0, ptr(out), ptr(self._dots), ptr(self._ws), self._ws.numel(),
                                          stream), "ccsd_diis_push")
            self.prev = out
            nrm = None if start is None else float(np.sqrt(max(float(self._dots[0]), 0.0)))
            return out, nrm
        slot = self.count % self.space
        self.count += 1
        nr = min(self.count, self.space)
        _lib.check(lib.ccsd_diis_push(ptr(a), ptr(b), n0, n, ptr(rhs), ptr(self.prev), ptr(self.X[slot]),
                                      ptr(self.E[slot]), ptr(self.E), n, nr, slot, None, ptr(self._dots),
                                      ptr(self._ws), self._ws.numel(), stream), "ccsd_diis_push")
        row = self._dots[:nr].cpu().numpy()
        cn = self._coefficients(slot, nr, row)
        out = torch.empty(n, dtype=torch.float64, device=dev)
        coef = (ctypes.c_double * nr)(*cn)
        _lib.check(lib.ccsd_diis_combine(ptr(self.X), n, nr, coef, None, ptr(out), n, stream), "ccsd_diis_combine")
        self.prev = out
        return out, float(np.sqrt(max(row[slot], 0.0)))

    def _coefficients(self, slot, n, row):
        """Enter `row` (overlaps of error `slot` with errors 0..n-1) into the overlap
        matrix and solve for the extrapolation coefficients (numpy); sets res_norm and
        c_abs."""
        self.B[slot, :n] = row
        self.B[:n, slot] = row
        A = -np.ones((n + 1, n + 1))
        A[:n, :n] = self.B[:n, :n]
        A[n, n] = 0.0
        rhs = np.zeros(n + 1)
        rhs[n] = -1.0
        cn = np.linalg.lstsq(A, rhs, rcond=None)[0][:n]
        # |sum_s c_s E[s]|, from the overlap matrix: see input_combination
        self.res_norm = float(np.sqrt(max(float(cn @ self.B[:n, :n] @ cn), 0.0)))
        self.c_abs = float(np.abs(cn).sum())
        self._cn = cn
        return cn

    def _extrapolate(self):
        slot = self.count % self.space
        self.count += 1
        n = min(self.count, self.space)
        torch.sub(self.X[slot], self.prev, out=self.E[slot])
        row = (self.E[:n] @ self.E[slot]).cpu().numpy()
        cn = self._coefficients(slot, n, row)
        c = torch.as_tensor(cn, dtype=self.X.dtype, device=self.X.device)
        self._c = c
        if self.X.is_cuda:
            # sum_s c_s X[s]: the column GEMV (one coalesced pass over the n stored vectors)
            from . import mfma
            out = mfma.gemv(self.X[:n].T, c, torch.empty(self.X.shape[1], dtype=self.X.dtype, device=self.X.device))
        else:
            out = c @ self.X[:n]
        self.prev = out
        return out, float(np.sqrt(max(row[slot], 0.0)))

    def input_combination(self):
        """sum_s c_s x_s for the coefficients of the last extrapolation, where x_s is the
        vector that stored vector X[s] was computed from (x_s = X[s] - E[s]).

        The extrapolation returns sum_s c_s X[s]; when X[s] = F(x_s) for an affine map F
        and sum_s c_s = 1, the vector x = sum_s c_s x_s has the residual
        F(x) - x = sum_s c_s E[s] exactly, whose norm is res_norm: a converged iterate
        known without another application of F. Call it before the next push."""
        n = min(self.count, self.space)
        if self._fp32:
            from . import _lib
            out = torch.empty_like(self.prev)
            coef = (ctypes.c_double * n)(*(-self._cn))
            _lib.check(_lib.load().ccsd_diis_combine(_lib.ptr(self.E), self.E.shape[1], n, coef, _lib.ptr(self.prev),
                                                     _lib.ptr(out), out.numel(), _lib.current_stream_ptr()),
                       "ccsd_diis_combine")
            return out
        if self.X.is_cuda:
            from . import mfma
            ebar = mfma.gemv(self.E[:n].T, self._c, torch.empty(self.E.shape[1], dtype=self.E.dtype,
                                                                   device=self.E.device))
        else:
            ebar = self._c @ self.E[:n]
        return self.prev - ebar


def amplitudes_to_vector(t1, t2):
    return torch.cat([t1.reshape(-1), t2.reshape(-1)])


def vector_to_amplitudes(v, nocc, nvir):
    n1 = nocc * nvir
    return v[:n1].reshape(nocc, nvir), v[n1:].reshape(nocc, nocc, nvir, nvir)


class CCSD:
    """Closed-shell CCSD with the attributes the reference's lambda/gradient code reads.

    `frozen` follows PySCF: None / 0, an int (that many lowest orbitals frozen) or a
    list of frozen orbital indices. nocc and nmo count active orbitals; mo_coeff and
    mo_occ stay those of all orbitals, as on PySCF's CCSD object."""

    def __init__(self, mf, frozen=None):
        self._scf = mf
        self.mol = mf.mol
        self.mo_coeff = mf.mo_coeff
        self.mo_occ = mf.mo_occ
        self.frozen = frozen
        frozen_mask(np.asarray(mf.mo_coeff).shape[1], frozen)      # validates
        self.max_cycle = 100
        self.conv_tol = 1e-10
        self.conv_tol_normt = 1e-8
        self.level_shift = 0.0
        self.diis = True
        self.diis_space = 6
        self.verbose = 0
        self.t1 = self.t2 = self.l1 = self.l2 = None
        self.converged = False
        self.converged_lambda = False
        self._eris = None

    def get_frozen_mask(self):
        return frozen_mask(np.asarray(self.mo_coeff).shape[1], self.frozen)

    @property
    def nocc(self):
        return int((np.asarray(self.mo_occ)[self.get_frozen_mask()] > 0).sum())

    @property
    def nmo(self):
        return int(self.get_frozen_mask().sum())

    def ao2mo(self, mo_coeff=None):
        """MO integrals of the active orbitals, built once per orbital set: they are kept
        on the SCF object (keyed by its mo_coeff object and the frozen mask) so that
        every CCSD of the same orbitals -- the T solve, the lambda solve, the
        gradient's Z-vector -- shares them."""
        mask = self.get_frozen_mask()
        if mo_coeff is not None:
            return ERIs(self._scf, mo_coeff, mask)
        if self._eris is not None:
            return self._eris
        cache = getattr(self._scf, "_cc_eris", None)
        if cache is not None and cache[0] is self._scf.mo_coeff and np.array_equal(cache[1], mask):
            self._eris = cache[2]
            return self._eris
        eris = ERIs(self._scf, None, mask)
        self._eris = eris
        try:
            self._scf._cc_eris = (self._scf.mo_coeff, mask, eris)
        except AttributeError:
            pass
        return eris

    def amplitudes_to_vector(self, t1, t2):
        return amplitudes_to_vector(t1, t2)

    def kernel(self, eris=None):
        eris = self.ao2mo() if eris is None else eris
        nocc = self.nocc
        e_o, e_v = eris.mo_energy[:nocc], eris.mo_energy[nocc:]
        nvir = e_v.numel()
        eia = e_o[:, None] - e_v
        t1 = eris.fock[:nocc, nocc:] / eia
        t2 = eris.ovov.permute(0, 2, 1, 3) / (eia[:, None, :, None] + eia[None, :, None, :])
        e = ccsd_energy(t1, t2, eris)
        diis = DIIS(self.diis_space) if self.diis else None
        self.converged = False
        with static_operands(eris.ovov, eris.ovoo, eris.ovvo, eris.oovv, eris.ovvv, eris.oooo, eris.vvvv):
            e, t1, t2 = self._iterate(eris, t1, t2, e, diis, nocc, nvir)
        self.e_corr = e
        self.e_tot = self._scf.e_tot + e
        self.t1, self.t2 = t1, t2
        return e, t1, t2

    def _iterate(self, eris, t1, t2, e, diis, nocc, nvir):
        for it in range(self.max_cycle):
            t1n, t2n = update_amps(t1, t2, eris)
            vec = amplitudes_to_vector(t1n, t2n)
            norm = float(torch.linalg.norm(vec - amplitudes_to_vector(t1, t2)))
            if diis is not None:
                vec = diis.update(vec)
            t1, t2 = vector_to_amplitudes(vec, nocc, nvir)
            e_new = ccsd_energy(t1, t2, eris)
            if abs(e_new - e) < self.conv_tol and norm < self.conv_tol_normt:
                e = e_new
                self.converged = True
                break
            e = e_new
        self._iterations = it + 1
        return e, t1, t2

    def solve_lambda(self, t1=None, t2=None, l1=None, l2=None, eris=None):
        from .gpu_ccsd_lambda import solve_lambda_gpu
        return solve_lambda_gpu(self, t1, t2, l1, l2, eris)


def restore_ccsd_solver(mf, t1, t2, e_corr=None):
    """A CCSD object holding given amplitudes (examples/h2o_ccsd_grad.py:58-70)."""
    mycc = CCSD(mf)
    mycc.t1 = torch.as_tensor(np.asarray(t1) if not torch.is_tensor(t1) else t1, dtype=torch.float64, device=DEV)
    mycc.t2 = torch.as_tensor(np.asarray(t2) if not torch.is_tensor(t2) else t2, dtype=torch.float64, device=DEV)
    mycc.e_corr = e_corr
    mycc.converged = True
    return mycc
