"""CCSD nuclear gradient on the GPU.

Keeps the reference's API (ccsd_gpu/gpu_ccsd_grad.py):
    _gamma1_intermediates_gpu(mycc, t1, t2, l1, l2)          (:101-126)
    _gamma2_intermediates_gpu(mycc, t1, t2, l1, l2)          (:129-246)
    _rdm2_mo2ao_gpu(mycc, d2, mo_coeff)                       (:254-336)
    grad_elec_gpu(mycc, t1, t2, l1, l2, atmlst=None, verbose=...)  (:344-586)
    compute_ml_ccsd_forces_gpu(mf, t1, t2, l1, l2)            (:711-739)
and the same gradient formula (Phases 1-5 of grad_elec_gpu). What is
organised differently, MI355X-first:
  * All AO-side two-electron work is done in the Cartesian AO basis: the
    2-RDM goes MO -> Cartesian AO in one transform with C_cart = cart2sph @ C;
    full contractions are basis independent, and Imat is brought back with
    Imat = C^T Imat_cc C (C^T C)^-1.
  * The AO ERIs are computed once (HIP Rys kernel) and stay in HBM; Imat is
    one DGEMM over them and every CPHF JK build two DGEMVs.
  * The reference's per-atom loop over shell blocks (int2e_ip1 -> HBM ->
    GEMV, plus the HF vhf1 J/K pieces, :423-496) becomes one fused launch set
    of the int2e_ip1 kernel contracting with Gamma + Gamma_sep, where
    Gamma_sep = D P + P D - (D x P + P x D)/4 (exchange-like terms, kl
    symmetrized) reproduces -sum vhf1 . dm1p. The reordering (Z-vector before
    the ip1 pass, so dm1p is known) is exact.
  * The Z-vector is a preconditioned conjugate-gradient solve on the device
    (the reference calls pyscf.scf.cphf.solve, :662).
`fused=False` keeps the reference's structure on the GPU (ip1 tensors in HBM
per atom block, GEMV, explicit vhf1): the plain FP64 comparison path.
"""
import time
import warnings

import numpy as np
import torch

from .cc import _frozen_mask_of
from .gpu_ccsd_lambda import _dev
from .hip.grad2e import add_separable, contract_int2e_ip1
from .hip.int1e_gpu import int1e
from .hip.int2e_ip1_gpu import compute_int2e_ip1_gpu
from . import mfma
from .precision import contract, contract_acc, split_memo
from .scf import DEV, get_ints, grad_nuc

E = torch.einsum


# ---------------------------------------------------------------------------
# Phase 1: RDM intermediates
# ---------------------------------------------------------------------------
def _gamma1_intermediates_gpu(mycc, t1, t2, l1, l2, precision="fp64"):
    C = lambda eq, *ops: contract(eq, *ops, precision=precision)  # noqa: E731
    t1, t2, l1, l2 = _dev(t1), _dev(t2), _dev(l1), _dev(l2)
    doo = -C("ja,ia->ij", t1, l1)
    dvv = C("ia,ib->ab", t1, l1)
    xtv = C("ie,me->im", t1, l1)
    dvo = t1.T - C("im,ma->ai", xtv, t1)
    theta = t2 * 2 - t2.permute(0, 1, 3, 2)
    doo = doo - C("jkab,ikab->ij", theta, l2)
    dvv = dvv + C("jica,jicb->ab", theta, l2)
    xt1 = C("mnef,inef->mi", l2, theta)
    xt2 = C("mnaf,mnef->ea", l2, theta)
    dvo = dvo + C("imae,me->ai", theta, l1) - C("mi,ma->ai", xt1, t1) - C("ie,ae->ai", t1, xt2)
    return doo, l1.clone(), dvo, dvv


def _gamma2_intermediates_gpu(mycc, t1, t2, l1, l2, precision="fp64"):
    """2-RDM blocks (dovov, gvvvv, doooo, doovv, dovvo, None, dovvv, dooov) as
    ccsd_gpu/gpu_ccsd_grad.py:129-246, except that the vvvv block is returned as
    gvvvv[a,b,c,d] = sum_ij l2[i,j,a,b] tau[i,j,c,d] itself (the reference's dvvvv =
    g[a,c,b,d] - g[a,c,d,b] / 2 is formed by the 2-RDM assembly, which reads g once).
    l2 (t2 + t1 t1) is one product where the reference adds the t1 t1 part through
    an o v^3 intermediate; terms are accumulated into their target by the GEMM
    epilogue where the split-BF16 path allows it."""
    C = lambda eq, *ops: contract(eq, *ops, precision=precision)  # noqa: E731

    def acc(out, eq, a, b, alpha=1.0):
        return contract_acc(out, eq, a, b, alpha=alpha, precision=precision)

    t1, t2, l1, l2 = _dev(t1), _dev(t2), _dev(l1), _dev(l2)
    pvOOv = C("ikca,jkcb->aijb", l2, t2)
    moo = E("dljd->jl", pvOOv) * 2
    mvv = E("blld->db", pvOOv) * 2
    gooov = C("kc,cija->jkia", t1, pvOOv)
    theta = t2 * 2 - t2.permute(0, 1, 3, 2)
    pvoOV = C("ikac,jkbc->aijb", l2, theta)
    acc(pvoOV, "ikca,jkbc->aijb", l2, t2, -1.0)
    moo = moo + E("dljd->jl", pvoOV)
    mvv = mvv + E("blld->db", pvoOV)
    gooov = gooov - C("jc,cika->jkia", t1, pvoOV)
    mia = C("kc,ikac->ia", l1, t2) * 2 - C("kc,ikca->ia", l1, t2)
    mab = C("kc,kb->cb", l1, t1)
    mij = C("kc,jc->jk", l1, t1) + moo * 0.5
    tau = E("ia,jb->ijab", t1, t1) + t2
    goooo = C("ijab,klab->ijkl", tau, l2) * 0.5
    doooo = mfma.combine([(2, goooo.permute(0, 2, 1, 3)), (-1, goooo.permute(0, 3, 1, 2))])
    gooov = gooov + C("ji,ka->jkia", -0.5 * moo, t1) + C("la,jkil->jkia", 2 * t1, goooo)
    gooov = gooov - C("ib,jkba->jkia", l1, tau) - C("jkba,ib->jkia", l2, t1)
    dooov = mfma.combine([(2, gooov.permute(0, 2, 1, 3)), (-1, gooov.permute(1, 2, 0, 3))])

    goovv = C("ijkl,klab->ijab", goooo, tau)
    acc(goovv, "jk,ikab->ijab", mij, tau, -1.0)
    acc(goovv, "cb,ijac->ijab", mab, t2, -1.0)
    acc(goovv, "bd,ijad->ijab", mvv * 0.5, tau, -1.0)
    goovv.addcmul_(mia[:, None, :, None], t1[None, :, None, :])
    mfma.combine([(0.5, tau), (0.5, l2)], out=goovv, beta=1)
    # the l2 (t1 t1) terms through w[p,q,a,j] = sum_c l2[p,q,a,c] t1[j,c] (t1 t1 is rank one)
    w = C("pqac,jc->pqaj", l2, t1)
    gOvvO = C("kiaj,kb->iabj", w, t1) + pvOOv.permute(1, 0, 3, 2)
    govVO = E("ia,jb->iabj", l1, t1) - C("ikaj,kb->iabj", w, t1) + pvoOV.permute(1, 0, 3, 2)
    del w
    dovvo = 2 * govVO + gOvvO
    doovv = mfma.combine([(-2, gOvvO.permute(3, 0, 1, 2)), (-1, govVO.permute(3, 0, 1, 2))])
    tau2 = tau - t2 * 0.5
    acc(goovv, "dlib,jlda->ijab", pvOOv, tau2)
    acc(goovv, "dlia,jldb->ijab", pvoOV, tau2, -1.0)
    acc(goovv, "dlia,jlbd->ijab", pvoOV + pvOOv * 0.5, t2)
    dovov = mfma.combine([(2, goovv.permute(0, 2, 1, 3)), (-1, goovv.permute(1, 2, 0, 3))])
    del goovv

    nocc, nvir = t1.shape
    if l2.is_cuda:
        # g[b,a,d,c] = g[a,b,c,d]: only the rows a <= b are formed (half the o^2 v^4 product),
        # the other half mirrored by the pair-unpack kernel
        from .pairops import unpack_pairs
        iu, ju = torch.triu_indices(nvir, nvir, device=l2.device)
        half = C("ijp,ijq->pq", l2.reshape(nocc, nocc, nvir * nvir)[:, :, iu * nvir + ju],
                 tau.reshape(nocc, nocc, nvir * nvir))
        gvvvv = unpack_pairs(half.reshape(-1, nvir, nvir).contiguous(), iu, ju, nvir)
        del half
    else:
        gvvvv = C("ijab,ijcd->abcd", l2, tau)
    # -sum_d g[a,d,b,c] t1[i,d], read as g[d,a,c,b] through g's (ab)(cd) symmetry: one
    # contiguous FP64 GEMM over g's slowest index (no permuted copy of the v^4 tensor)
    gv = torch.mm(t1, gvvvv.reshape(nvir, nvir ** 3)).view(nocc, nvir, nvir, nvir)    # [i, a, c, b]
    gvovv = mfma.combine([(-1.0, gv.permute(1, 0, 3, 2))])
    del gv
    acc(gvovv, "akic,kb->aibc", pvoOV, t1)
    acc(gvovv, "akib,kc->aibc", pvOOv, t1, -1.0)
    # l1[j,a] (t2 + t1 t1)[j,i,b,c] + t1[j,a] l2[j,i,b,c]: one product over K = 2 nocc
    acc(gvovv, "ja,jibc->aibc", torch.cat([l1, t1]), torch.cat([tau, l2]))
    acc(gvovv, "ba,ic->aibc", mvv * 0.5, t1)
    dovvv = mfma.combine([(2, gvovv.permute(1, 3, 0, 2)), (-1, gvovv.permute(1, 2, 0, 3))])
    return (dovov, gvvvv, doooo, doovv, dovvo, None, dovvv, dooov)


# ---------------------------------------------------------------------------
# Phase 2: full 2-RDM and MO -> AO
# ---------------------------------------------------------------------------
def _assemble_dm2_mo(d2, presym=False):
    """Full MO 2-RDM (chemist's order) from the density blocks (vvvv as gvvvv), as
    ccsd_gpu/gpu_ccsd_grad.py:272-300; each of the 16 occupied/virtual blocks is
    written once by the combine kernel (no zero fill, transposed blocks through
    LDS tiles).

    presym=True is for callers that apply the (pq)(rs) symmetrisation next: the vvvv
    block's third and fourth terms are the (pq)(rs) images of its first two, so
    4 A - 2 B replaces 2 A - B + 2 C - D there (two reads of gvvvv instead of four);
    the symmetrised tensor is the same, the unsymmetrised one is not."""
    dovov, gvvvv, doooo, doovv, dovvo, _, dovvv, dooov = d2
    nocc, nvir = dovov.shape[:2]
    nmo = nocc + nvir
    o, v = slice(0, nocc), slice(nocc, None)
    dm2 = torch.empty((nmo,) * 4, dtype=torch.float64, device=DEV)
    cb = mfma.combine
    cb([(1, dovov), (1, dovov.permute(2, 3, 0, 1))], out=dm2[o, v, o, v])
    cb([(1, dm2[o, v, o, v].permute(1, 0, 3, 2))], out=dm2[v, o, v, o])
    cb([(1, doovv), (1, doovv.permute(1, 0, 3, 2))], out=dm2[o, o, v, v])
    cb([(1, dm2[o, o, v, v].permute(2, 3, 0, 1))], out=dm2[v, v, o, o])
    cb([(1, dovvo), (1, dovvo.permute(3, 2, 1, 0))], out=dm2[o, v, v, o])
    cb([(1, dm2[o, v, v, o].permute(1, 0, 3, 2))], out=dm2[v, o, o, v])
    # 2 dvvvv + 2 dvvvv^(ba dc) with dvvvv[a,b,c,d] = g[a,c,b,d] - g[a,c,d,b] / 2, read from g
    # through its (ab)(cd) symmetry g[a,b,c,d] = g[b,a,d,c] so that every view is contiguous
    # along d or along c (one LDS-tiled dimension)
    g = gvvvv
    if presym:
        cb([(4, g.permute(0, 2, 1, 3)), (-2, g.permute(1, 2, 0, 3))], out=dm2[v, v, v, v])
    else:
        cb([(2, g.permute(0, 2, 1, 3)), (-1, g.permute(1, 2, 0, 3)), (2, g.permute(2, 0, 3, 1)),
            (-1, g.permute(2, 1, 3, 0))], out=dm2[v, v, v, v])
    cb([(2, doooo), (2, doooo.permute(1, 0, 3, 2))], out=dm2[o, o, o, o])
    cb([(1, dovvv)], out=dm2[o, v, v, v])
    cb([(1, dovvv.permute(2, 3, 0, 1))], out=dm2[v, v, o, v])
    cb([(1, dovvv.permute(3, 2, 1, 0))], out=dm2[v, v, v, o])
    cb([(1, dovvv.permute(1, 0, 3, 2))], out=dm2[v, o, v, v])
    cb([(1, dooov)], out=dm2[o, o, o, v])
    cb([(1, dooov.permute(2, 3, 0, 1))], out=dm2[o, v, o, o])
    cb([(1, dooov.permute(1, 0, 3, 2))], out=dm2[o, o, v, o])
    cb([(1, dooov.permute(3, 2, 1, 0))], out=dm2[v, o, o, o])
    return dm2


def _rdm2_mo2ao_gpu(mycc, d2, mo_coeff, precision="fp64"):
    """Symmetrized AO 2-RDM, full (nao^4), in the AO basis of the columns of
    `mo_coeff`'s rows (Cartesian when mo_coeff = cart2sph @ C).

    As the reference (gpu_ccsd_grad.py:272-326): the MO density in chemist's
    order, C x C x C x C, then (ij) and (kl) symmetrised with weight 1/2 per
    term. Here the symmetrisation runs first, in place on the MO tensor (it
    commutes with the transform, and the reference's (pq)(rs) index swap before
    the transform is one of the symmetrised permutations), and the transform is
    four contiguous GEMMs, one index at a time, without permuted copies."""
    dm2 = _assemble_dm2_mo(d2)
    nmo = dm2.shape[0]
    _lib_sym_pairs(dm2, 0.5)
    c = _dev(mo_coeff)
    nao = c.shape[0]
    t = dm2                                               # (p q r s)
    if precision == "fp32s":
        c, t = c.float(), t.float()
    for _ in range(4):
        # last index transformed and rotated to the front: (p q r s) -> (l p q r) -> ... -> (i j k l)
        t = _rotate_transform(t, c, precision)
    return t.double()


def _rdm2_mo2ao_packed_gpu(mycc, d2, mo_coeff, precision="fp64"):
    """The same AO 2-RDM, packed as the fused gradient kernel and Imat take it:
    G[i][j][k(k+1)/2 + l] for k >= l, the kl diagonal halved (nao^2 npair).

    The (pq)(rs) symmetrisation commutes with the index transforms, so it is applied
    to the MO tensor while p >= q is packed out of it: the r and s transforms then
    run on half the rows. Their product is repacked (rs pair packed, pq unpacked)
    and the p and q transforms write the packed layout directly."""
    dm2 = _assemble_dm2_mo(d2, presym=True)
    c = _dev(mo_coeff)
    nmo = c.shape[1]
    tp = _pack_pairs_sym(dm2)                    # (pq r s), p >= q, symmetrised
    del dm2
    if precision == "fp32s":
        # FP32 storage and FP32 products through the four transforms (half the bytes
        # of every intermediate, twice the matrix rate), FP64 result
        c = c.float()
        tp = tp.float()
    t = _rotate_transform(tp, c, precision)      # (s pq r)
    del tp
    t = _rotate_transform(t, c, precision)       # (r s pq)
    t = _repack_pairs(t, nmo)                    # (rs p q), r >= s, the rs diagonal halved
    t = _rotate_transform(t, c, precision)       # (q rs p)
    return _rotate_transform(t, c, precision).double()   # (p q rs)


def _repack_pairs(z, m):
    """z (n, n, m(m+1)/2): the second pair packed with weight 2 off its diagonal (the
    _pack_pairs_sym layout) -> (n(n+1)/2, m, m): the first pair packed k >= l with
    that weight, the second unpacked (one kernel on the device, torch on the CPU)."""
    n = z.shape[0]
    out = torch.empty((n * (n + 1) // 2, m, m), dtype=z.dtype, device=z.device)
    if z.is_cuda and z.is_contiguous() and z.dtype in (torch.float64, torch.float32):
        from . import _lib
        fn = _lib.load().ccsd_repack_pairs if z.dtype == torch.float64 else _lib.load().ccsd_repack_pairs_f32
        _lib.check(fn(_lib.ptr(z), n, m, _lib.ptr(out), _lib.current_stream_ptr()), "ccsd_repack_pairs")
        return out
    kk, ll = torch.tril_indices(n, n, device=z.device)
    i = torch.arange(m, device=z.device)
    a, b = torch.maximum(i[:, None], i[None, :]), torch.minimum(i[:, None], i[None, :])
    p = a * (a + 1) // 2 + b
    wkl = torch.where(kk == ll, 1.0, 2.0).to(z.dtype)
    wij = torch.where(i[:, None] == i[None, :], 1.0, 2.0).to(z.dtype)
    return torch.mul(z[kk, ll][:, p], wkl[:, None, None] / wij[None], out=out)


def _pack_pairs_sym(t):
    """t (n, n, m, m) -> G (n(n+1)/2, m, m) with G[k(k+1)/2 + l, p, q] =
    (t[k,l,p,q] + t[l,k,p,q] + t[k,l,q,p] + t[l,k,q,p]) / 2 for k >= l (the
    (ij)(kl) symmetrisation with weight 1/2 per term) and the k == l rows halved, in
    one kernel launch on the device (torch on the CPU)."""
    n, _, m, _ = t.shape
    tp = torch.empty((n * (n + 1) // 2, m, m), dtype=t.dtype, device=t.device)
    if t.is_cuda and t.is_contiguous() and t.dtype == torch.float64:
        from . import _lib
        kk, ll = torch.tril_indices(n, n, device=t.device)
        _lib.check(_lib.load().ccsd_pack_pairs_sym(_lib.ptr(t), _lib.ptr(kk), _lib.ptr(ll), n, m, kk.numel(),
                                                   _lib.ptr(tp), _lib.current_stream_ptr()), "ccsd_pack_pairs_sym")
        return tp
    for k in range(n):
        r0 = k * (k + 1) // 2
        a, b = t[k, :k + 1], t[:k + 1, k]
        mfma.combine([(0.5, a), (0.5, b), (0.5, a.transpose(1, 2)), (0.5, b.transpose(1, 2))],
                     out=tp[r0:r0 + k + 1])
    kk = torch.arange(n, device=t.device)
    tp[kk * (kk + 1) // 2 + kk] *= 0.5
    return tp


def _rotate_transform(X, c, precision):
    """Y[l, r...] = sum_s X[r..., s] c[l, s] for a contiguous X: one GEMM with the
    contracted index contiguous in both operands, the product written transposed
    (the new index slowest), so four calls give the full C x C x C x C transform
    with every operand contiguous and no permuted copy."""
    K = X.shape[-1]
    R = X.numel() // K
    n = c.shape[0]
    X2 = X.reshape(R, K)
    if X.dtype == torch.float32:
        # FP32 operands kept in FP32 (the "fp32s" transform chain): library SGEMM over
        # column blocks whose operand and product stay below 512 MiB -- single SGEMMs
        # with a product above 2 GiB returned NaN in part of their rows on this stack
        out = torch.empty((n, R), dtype=torch.float32, device=X.device)
        rc = max(1, (1 << 27) // max(K, n))
        for r0 in range(0, R, rc):
            r1 = min(R, r0 + rc)
            out[:, r0:r1] = c @ X2[r0:r1].T
    elif precision == "bf16x3":
        out = torch.empty((n, R), dtype=torch.float64, device=X.device)
        zero = torch.zeros(1, dtype=torch.int64, device=X.device)
        rows = torch.arange(R, dtype=torch.int64, device=X.device)
        cols = torch.arange(n, dtype=torch.int64, device=X.device) * R
        mfma.gemm_nt_scatter(mfma.split(X2[None]), mfma.split(c[None]), out, zero, rows, cols, beta=0)
    elif precision == "fp32":
        out = mfma.bmm(X2[None], c.T[None], "fp32")[0].T.contiguous()
    elif precision == "fp64":
        out = c @ X2.T
    else:
        raise ValueError(f"unknown precision {precision!r}")
    return out.reshape((n,) + tuple(X.shape[:-1]))


def _lib_sym_pairs(t, scale):
    from . import _lib
    n = t.shape[0]
    if t.dim() != 4 or any(d != n for d in t.shape) or not t.is_contiguous() or t.dtype != torch.float64:
        raise ValueError("sym_pairs needs a contiguous float64 (n, n, n, n) tensor")
    _lib.check(_lib.load().ccsd_sym_pairs(_lib.ptr(t), n, float(scale), _lib.current_stream_ptr()),
               "ccsd_sym_pairs")
    return t


# ---------------------------------------------------------------------------
# Phase 4: Z-vector
# ---------------------------------------------------------------------------
def _response_dm1_gpu(mycc, Xvo, ints, tol=1e-11, max_cycle=200, mo_energy=None, mo_coeff=None):
    """Solve (e_a - e_i) x_ai + fvind(x)_ai = -Xvo_ai by preconditioned CG on the device,
    over all occupied x virtual orbitals (frozen ones included).

    fvind is the reference's 2 C_v^T (J - K/2)[D] C_o with D = C_v x C_o^T + transpose
    (gpu_ccsd_grad.py:644-651). With this build's CCSD and no frozen orbitals its vo
    block comes from MO integrals the CCSD already holds: 4 (ai|bj) x_bj - (ab|ij) x_bj
    - (aj|ib) x_bj, i.e. contractions over ovov and oovv (o^2 v^2 each) summed once into
    one (vo, vo) matrix, so each CG step is one matrix-vector product. Otherwise (frozen
    orbitals, whose CCSD integrals cover the active space only, or a PySCF-like mycc)
    each step builds J and K of D from the resident AO integrals (two GEMVs), as the
    reference's frozen-orbital response does (:593-594, :655-660).

    Sets mycc._zvector_iterations and mycc._zvector_converged; an unconverged solve
    warns (RuntimeWarning) and returns its last iterate."""
    from .cc import CCSD, _frozen_mask_of
    mf = mycc._scf
    nvir, nocc = Xvo.shape
    c = _dev(mycc.mo_coeff if mo_coeff is None else mo_coeff)
    e = _dev(mf.mo_energy if mo_energy is None else mo_energy)
    eai = e[nocc:, None] - e[None, :nocc]
    if isinstance(mycc, CCSD) and _frozen_mask_of(mycc).all():
        eris = mycc.ao2mo()
        ovov, oovv = eris.ovov, eris.oovv
        # H[a,i,b,j] = 4 (ia|jb) - (ij|ab) - (ja|ib)
        H = mfma.combine([(4.0, ovov.permute(1, 0, 3, 2)), (-1.0, oovv.permute(2, 0, 3, 1)),
                          (-1.0, ovov.permute(1, 2, 3, 0))]).reshape(nvir * nocc, nvir * nocc)

        def fvind(x):
            return (H @ x.reshape(-1)).reshape(nvir, nocc)
    else:
        cv, co = c[:, nocc:], c[:, :nocc]

        def fvind(x):
            d = cv @ x @ co.T
            j, k = ints.get_jk(d + d.T)
            return (cv.T @ (j - 0.5 * k) @ co) * 2.0

    def A(x):
        return eai * x + fvind(x)

    b = -Xvo
    x = b / eai
    r = b - A(x)
    z = r / eai
    p = z.clone()
    rz = torch.sum(r * z)
    bn = float(torch.linalg.norm(b))
    converged = False
    it = 0
    for it in range(max_cycle + 1):
        if float(torch.linalg.norm(r)) <= tol * max(bn, 1e-300):
            converged = True
            break
        if it == max_cycle:
            break
        Ap = A(p)
        alpha = rz / torch.sum(p * Ap)
        x = x + alpha * p
        r = r - alpha * Ap
        z = r / eai
        rz_new = torch.sum(r * z)
        p = z + (rz_new / rz) * p
        rz = rz_new
    _set_attr(mycc, "_zvector_iterations", it)
    _set_attr(mycc, "_zvector_converged", converged)
    if not converged:
        warnings.warn(f"Z-vector (CPHF) solve not converged in {max_cycle} iterations "
                      f"(|r|/|b| = {float(torch.linalg.norm(r)) / max(bn, 1e-300):.2e} > {tol:.1e})",
                      RuntimeWarning, stacklevel=2)
    nmo = c.shape[1]
    dm1 = torch.zeros((nmo, nmo), dtype=torch.float64, device=DEV)
    dm1[nocc:, :nocc] = x
    dm1[:nocc, nocc:] = x.T
    return dm1


def _set_attr(obj, name, value):
    try:
        setattr(obj, name, value)
    except AttributeError:
        pass


# ---------------------------------------------------------------------------
# Phases 3-5
# ---------------------------------------------------------------------------
# largest Cartesian AO count of ccsd_add_separable_packed (it stages four AO rows in LDS)
ADD_SEPARABLE_MAX_NAO = 2048


def _separable_gamma(D, P):
    """Gamma_sep(ij,kl) = D_ij P_kl + P_ij D_kl - (D_il P_kj + D_ik P_lj + P_il D_jk + P_ik D_jl)/4."""
    g = E("ij,kl->ijkl", D, P)
    g += E("ij,kl->ijkl", P, D)
    g -= 0.25 * E("il,kj->ijkl", D, P)
    g -= 0.25 * E("ik,lj->ijkl", D, P)
    g -= 0.25 * E("il,jk->ijkl", P, D)
    g -= 0.25 * E("ik,jl->ijkl", P, D)
    return g




def grad_elec_gpu(mycc, t1=None, t2=None, l1=None, l2=None, atmlst=None, verbose=0, fused=True,
                  precision="fp64", timings=None, mo2ao_precision=None):
    """Electronic CCSD gradient dE/dR (natm, 3) as numpy, as ccsd_gpu/gpu_ccsd_grad.py:344.

    precision: of the RDM contractions ("fp64", "fp32", "bf16x3"). mo2ao_precision: of
    the MO -> AO transform of the 2-RDM, "fp64" (None, the default), "fp32s" (FP32
    storage and products through the four index transforms, FP64 result: 5e-8 on the
    (H2O)8 AO density, 7e-8 Eh/Bohr on its gradient and a net force of 1.4e-7, for
    5 ms of 42), or "bf16x3" / "fp32" (the FP64 chain with low-precision products)."""
    tm = {} if timings is None else timings
    if mo2ao_precision is None:
        mo2ao_precision = "fp64"

    def mark(name, t0):
        torch.cuda.synchronize()
        tm[name] = tm.get(name, 0.0) + time.perf_counter() - t0
        return time.perf_counter()

    t1 = _dev(mycc.t1 if t1 is None else t1)
    t2 = _dev(mycc.t2 if t2 is None else t2)
    l1 = _dev(mycc.l1 if l1 is None else l1)
    l2 = _dev(mycc.l2 if l2 is None else l2)
    mol = mycc.mol
    mf = mycc._scf
    ints = get_ints(mf)
    c2s = ints.c2s
    mo = _dev(mycc.mo_coeff)
    mo_energy = _dev(mf.mo_energy)
    mo_occ = np.asarray(mycc.mo_occ, dtype=np.float64)
    nao, nmo = mo.shape
    nocc = int((mo_occ > 0).sum())              # every occupied orbital (the response space)
    # frozen / active orbital sets, as ccsd_gpu/gpu_ccsd_grad.py:373-377
    mask = _frozen_mask_of(mycc)
    with_frozen = not mask.all()
    OA, VA = np.nonzero(mask & (mo_occ > 0))[0], np.nonzero(mask & (mo_occ == 0))[0]
    OF, VF = np.nonzero(~mask & (mo_occ > 0))[0], np.nonzero(~mask & (mo_occ == 0))[0]
    if t1.shape != (len(OA), len(VA)):
        raise ValueError(f"t1 has shape {tuple(t1.shape)}, the active space is {len(OA)} x {len(VA)}")
    mo_act = mo[:, torch.as_tensor(np.concatenate([OA, VA]), device=DEV)] if with_frozen else mo
    nc = c2s.shape[0]
    t0 = time.perf_counter()

    # Phase 1
    with split_memo():       # operands split once per layout across both RDM builders
        d1 = _gamma1_intermediates_gpu(mycc, t1, t2, l1, l2, precision)
        d2 = _gamma2_intermediates_gpu(mycc, t1, t2, l1, l2, precision)
    doo, dov, dvo, dvv = d1
    t0 = mark("rdm", t0)

    # Phase 2: Cartesian AO 2-RDM (symmetrized) of the active orbitals
    # FP64 by default: a split-BF16 operand keeps 16 mantissa bits, which the
    # nao^4 contraction with the derivative integrals turns into ~1e-7 Eh/Bohr
    gp = _rdm2_mo2ao_packed_gpu(mycc, d2, c2s @ mo_act, mo2ao_precision)
    del d2
    t0 = mark("mo2ao", t0)

    # Phase 3a: Imat (one DGEMM over the HBM-resident packed Cartesian ERIs:
    # sum over all kl of (ij|kl) G_i'jkl = 2 sum_{k>=l} with the diagonal halved)
    imat_cc = ints.eri_p.reshape(nc, -1) @ gp.reshape(nc, -1).T
    ctc = c2s.T @ c2s
    imat = c2s.T @ imat_cc @ c2s @ torch.linalg.inv(ctc)
    t0 = mark("imat", t0)

    # Phase 4: orbital response
    S = ints.S
    Imo = -(mo.T @ imat @ S @ mo)
    dm1mo = torch.zeros((nmo, nmo), dtype=torch.float64, device=DEV)
    if with_frozen:
        # frozen-occupied / frozen-virtual rotations from Imat (gpu_ccsd_grad.py:510-518)
        oa, va, of, vf = (torch.as_tensor(x, device=DEV) for x in (OA, VA, OF, VF))
        dco = Imo[of][:, oa] / (mo_energy[of][:, None] - mo_energy[oa][None, :])
        dfv = Imo[vf][:, va] / (mo_energy[vf][:, None] - mo_energy[va][None, :])
        dm1mo[oa[:, None], oa] = doo + doo.T
        dm1mo[of[:, None], oa] = dco
        dm1mo[oa[:, None], of] = dco.T
        dm1mo[va[:, None], va] = dvv + dvv.T
        dm1mo[vf[:, None], va] = dfv
        dm1mo[va[:, None], vf] = dfv.T
    else:
        dm1mo[:nocc, :nocc] = doo + doo.T
        dm1mo[nocc:, nocc:] = dvv + dvv.T
    dm1 = mo @ dm1mo @ mo.T
    j, k = ints.get_jk(dm1)
    vhf = (j - 0.5 * k) * 2
    Xvo = mo[:, nocc:].T @ vhf @ mo[:, :nocc]
    Xvo = Xvo + Imo[:nocc, nocc:].T - Imo[nocc:, :nocc]
    dm1mo = dm1mo + _response_dm1_gpu(mycc, Xvo, ints, mo_energy=mo_energy, mo_coeff=mo)
    tm["zvector_iterations"] = getattr(mycc, "_zvector_iterations", None)
    Imo[nocc:, :nocc] = Imo[:nocc, nocc:].T
    im1 = mo @ Imo @ mo.T
    t0 = mark("zvector", t0)

    # HF and relaxed densities
    hf_dm1 = 2.0 * mo[:, :nocc] @ mo[:, :nocc].T
    dm1 = mo @ dm1mo @ mo.T
    dm1p = hf_dm1 + dm1 * 2

    def one_electron():
        """Phase 5: one-electron and overlap terms (they need the relaxed density only)."""
        s1 = -int1e(mol, "ipovlp", c2s)
        vrinv_all = int1e(mol, "ipnuc_atom", c2s)
        h1 = -(int1e(mol, "ipkin", c2s) + vrinv_all.sum(dim=0))
        zeta = (mo_energy[:, None] + mo_energy[None, :]) * 0.5
        zeta[nocc:, :nocc] = mo_energy[:nocc]
        zeta[:nocc, nocc:] = mo_energy[:nocc].reshape(-1, 1)
        zeta = mo @ (zeta * dm1mo) @ mo.T
        p1 = mo[:, :nocc] @ mo[:, :nocc].T
        j, k = ints.get_jk(dm1 + dm1.T)
        vhf_s1occ = p1 @ (j - 0.5 * k) @ p1
        dm1_tot = dm1 + hf_dm1
        mocc = mo[:, :nocc]
        zeta = zeta + (mocc * (mo_energy[:nocc] * 2.0)) @ mocc.T
        # every term is a sum over the AO rows i of atom A (the reference's per-atom loop,
        # gpu_ccsd_grad.py:554-583): form all row sums at once, then add them up by atom
        offs = mol.offset_nr_by_atom()
        ao_atom = torch.as_tensor(np.repeat(np.arange(mol.natm), offs[:, 3] - offs[:, 2]), device=DEV)
        sym_dm = dm1_tot + dm1_tot.T
        w = im1 + im1.T - zeta - zeta.T - 2.0 * vhf_s1occ
        rows = torch.einsum("xij,ij->ix", s1, w) + torch.einsum("xij,ij->ix", h1, sym_dm)   # (nao, 3)
        de_all = torch.zeros((mol.natm, 3), dtype=torch.float64, device=DEV)
        de_all.index_add_(0, ao_atom, rows)
        de_all += torch.einsum("axij,ij->ax", vrinv_all, sym_dm)
        return de_all

    # Phase 5 runs on a side stream of default priority under the two-electron pass: it
    # waits for the work so far (not for the pass), and is launched after the pass is
    # enqueued, since some of its host steps wait for the device
    main = torch.cuda.current_stream()
    densities_ready = torch.cuda.Event()
    densities_ready.record(main)

    # Phase 3b: two-electron derivative contribution
    de2 = torch.zeros((mol.natm, 3), dtype=torch.float64, device=DEV)
    Dc = c2s @ hf_dm1 @ c2s.T
    Pc = c2s @ dm1p @ c2s.T
    if fused and Dc.shape[0] <= ADD_SEPARABLE_MAX_NAO:
        # one explicit pass over the CCSD density with the D/P separable part added into it
        # (one HBM pass; forming it per element inside the pass cost 40% more there)
        add_separable(gp, Dc, Pc)
        contract_int2e_ip1(mol, gp, scale=-2.0, de=de2)
        del gp
    elif fused:
        # the separable part formed inside the fused pass (no row-staging limit)
        contract_int2e_ip1(mol, gp, scale=-2.0, de=de2, dm=Dc.contiguous(), pm=Pc.contiguous())
        del gp
    else:
        _grad2e_unfused(mol, gp, Dc, Pc, de2)
        del gp
    side = torch.cuda.Stream()
    side.wait_event(densities_ready)
    with torch.cuda.stream(side):
        de_all = one_electron()
    main.wait_stream(side)
    de_all.record_stream(main)
    t0 = mark("grad2e", t0)     # (includes the one-electron terms, run under the pass)

    de_all += de2
    atoms = range(mol.natm) if atmlst is None else list(atmlst)
    de = de_all[list(atoms)]
    t0 = mark("onee", t0)
    return de.cpu().numpy()


def _grad2e_unfused(mol, gp, Dc, Pc, de2):
    """Reference structure on the GPU (ccsd_gpu/gpu_ccsd_grad.py:423-496), in the
    Cartesian AO basis: int2e_ip1 into HBM (the register-resident store kernels; with
    288 GB the block is the whole molecule, the reference's memory-sized shell blocks
    collapse into one), then per atom A:
      * the 2-RDM term  de_A -= 2 sum_{i in A, j, k>=l} ip1 Gamma'  (one GEMV),
      * the HF vhf1 pieces contracted with dm1p:
          C1  sum ip1_ijkl D_ij P_kl      C3  sum ip1_ijkl P_ij D_kl
          X2  sum ip1_ijkl D_il P_kj      X4  sum ip1_ijkl D_jk P_il
        de_A -= C1 + C3 - (X2 + X4) / 2. C1 / C3 are GEMVs over the packed k >= l pairs
        with the pair-weighted densities; X2 / X4 GEMVs over (j, k) of the block
        gathered as [x][i][l][j][k].
    D = the HF density, P = dm1p (both symmetric)."""
    mol_c = mol.copy()
    mol_c.cart = True
    loc = mol_c.ao_loc_nr()
    nc = mol_c.nao_nr()
    npair = nc * (nc + 1) // 2
    kk, ll = torch.tril_indices(nc, nc, device=DEV)
    wdiag = torch.where(kk == ll, 1.0, 2.0).to(torch.float64)
    Pw, Dw = Pc[kk, ll] * wdiag, Dc[kk, ll] * wdiag                 # sum over all kl of symmetric pairs
    pidx = torch.empty((nc, nc), dtype=torch.int64, device=DEV)
    pidx[kk, ll] = torch.arange(npair, device=DEV)
    pidx[ll, kk] = pidx[kk, ll]
    ar = torch.arange(nc, device=DEV)
    J = ar[None, :, None].expand(nc, nc, nc)                         # [l][j][k] -> j
    Q = pidx.T[:, None, :].expand(nc, nc, nc)                        # [l][j][k] -> pair(k, l)
    Pf, Df = Pc.reshape(-1), Dc.reshape(-1)
    ip1 = compute_int2e_ip1_gpu(mol_c, 0, mol.nbas)                  # (3, nc, nc, npair)
    offs = mol.offset_nr_by_atom()
    for a in range(mol.natm):
        s0, s1 = int(offs[a, 0]), int(offs[a, 1])
        if s1 <= s0:
            continue
        i0, i1 = int(loc[s0]), int(loc[s1])
        nf = i1 - i0
        blk = ip1[:, i0:i1]                                            # (3, nf, nc, npair); blk[x] contiguous
        g = gp[i0:i1].reshape(-1)
        u0 = torch.empty((3, nf * nc), dtype=torch.float64, device=DEV)
        u1 = torch.empty_like(u0)
        for x in range(3):
            rows = blk[x].reshape(nf * nc, npair)
            de2[a, x] -= 2.0 * torch.dot(rows.reshape(-1), g)
            mfma.gemv(rows, Pw, u0[x])
            mfma.gemv(rows, Dw, u1[x])
        u0, u1 = u0.view(3, nf, nc), u1.view(3, nf, nc)
        c13 = (u0 * Dc[i0:i1]).sum((1, 2)) + (u1 * Pc[i0:i1]).sum((1, 2))
        f2 = blk[:, :, J, Q].reshape(3 * nf * nc, nc * nc)             # [x][i][l][(j, k)]
        w0 = mfma.gemv(f2, Pf, torch.empty(3 * nf * nc, dtype=torch.float64, device=DEV)).view(3, nf, nc)
        w1 = mfma.gemv(f2, Df, torch.empty(3 * nf * nc, dtype=torch.float64, device=DEV)).view(3, nf, nc)
        x24 = (w0 * Dc[i0:i1]).sum((1, 2)) + (w1 * Pc[i0:i1]).sum((1, 2))
        de2[a] -= c13 - 0.5 * x24
        del f2
    del ip1


def compute_ml_ccsd_forces_gpu(mf, t1, t2, l1, l2, fused=True, precision="fp64", timings=None,
                               mo2ao_precision=None):
    """Total dE/dR (natm, 3) in Eh/Bohr (despite the name, not -dE/dR), as
    ccsd_gpu/gpu_ccsd_grad.py:711-739: electronic gradient + nuclear repulsion.

    mf is this build's RHF or any converged RHF-like object exposing mol, mo_coeff,
    mo_occ and mo_energy (a PySCF RHF); amplitudes may be NumPy arrays or tensors.
    The nuclear term is computed here (the reference asks mf.nuc_grad_method())."""
    from .cc import CCSD
    mycc = CCSD(mf)
    mycc.t1, mycc.t2, mycc.l1, mycc.l2 = _dev(t1), _dev(t2), _dev(l1), _dev(l2)
    mycc.converged = True
    mycc.converged_lambda = True
    de = grad_elec_gpu(mycc, t1, t2, l1, l2, verbose=0, fused=fused, precision=precision, timings=timings,
                       mo2ao_precision=mo2ao_precision)
    de += grad_nuc(mf.mol)
    return np.array(de)
