"""CCSD lambda solver on the GPU.

Keeps the reference's API (ccsd_gpu/gpu_ccsd_lambda.py):
    make_intermediates_gpu(mycc, t1, t2, eris)               (:68-209)
    update_lambda_gpu(mycc, t1, t2, l1, l2, eris, imds)       (:212-337)
    kernel_gpu(mycc, eris, t1, t2, l1, l2, max_cycle, tol)    (:340-408)
    solve_lambda_gpu(mycc, t1=None, t2=None, l1=None, l2=None, eris=None,
                     use_cpu_vvvv=True, fallback_to_cpu=True)  (:411-449)
The lambda equations are the reference's (a port of PySCF ccsd_lambda). What
differs is where things run: everything -- the l2.vvvv term the reference
sends to the CPU (:238-243), the DIIS extrapolation and the convergence norm
(:390-401) -- stays on the device, so an iteration moves no amplitudes across
PCIe. Contractions go through `contract` (precision.py): FP64 for the plain
path, FP32 / split-BF16 MFMA GEMMs with FP64 residual accumulation for the
mixed path.

Differences of behaviour, on purpose:
  * there is no CPU fallback: without a HIP device every call raises, whatever
    `fallback_to_cpu` says (the reference falls back to PySCF);
  * eris may be this build's device ERIs or a PySCF-like object with NumPy
    blocks (copied to the device once per solve); with none, integrals are built
    on the device from mycc's orbitals;
  * use_cpu_vvvv=True includes the l2.vvvv term (computed on the GPU);
    use_cpu_vvvv=False omits it, as the reference does (:244-245).
"""
import time

import numpy as np
import torch

from . import mfma
from .cc import DIIS, amplitudes_to_vector, vector_to_amplitudes
from .pairops import l1_terms, l2_finish, pair_combine
from .precision import contract, contract_acc, refresh_static, split_memo, static_operands, vvvv_term
from .scf import DEV


class _IMDS:
    pass


def _dev(a):
    if torch.is_tensor(a):
        return a.to(device=DEV, dtype=torch.float64)
    return torch.as_tensor(np.asarray(a), dtype=torch.float64, device=DEV)


def _require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("No usable HIP device for the lambda solve (this build has no CPU fallback).")


def _pair_sum(X, c1=0.0, c2=0.0, c3=0.0, c4=0.0):
    """c1 X[i,j,a,b] + c2 X[i,j,b,a] + c3 X[j,i,a,b] + c4 X[j,i,b,a] as a new contiguous
    tensor: the library's tiled pair kernel on the device (one pass; torch's strided
    elementwise path ran 3-6x slower on these permuted sums), torch on the host."""
    if X.is_cuda and X.dtype == torch.float64:
        out = torch.empty(X.shape, dtype=X.dtype, device=X.device)     # contiguous whatever X's strides
        return pair_combine(X, c1=c1, c2=c2, c3=c3, c4=c4, out=out)
    out = torch.zeros_like(X, memory_format=torch.contiguous_format)
    for c, perm in ((c1, (0, 1, 2, 3)), (c2, (0, 1, 3, 2)), (c3, (1, 0, 2, 3)), (c4, (1, 0, 3, 2))):
        if c:
            out += c * X.permute(perm)
    return out


def make_intermediates_gpu(mycc, t1, t2, eris, precision="fp64", lp_precision=None):
    """The lambda intermediates (the reference's make_intermediates, gpu_ccsd_lambda.py:68-180).

    With lp_precision set (the mixed solve), the three o v^3 x o v (o^2) products of
    wvvov are formed at that precision only (imds.wvvov then carries low-precision
    accuracy, which is all its low-precision updates use), and the FP64 update applies
    them to l2 in factored form from imds.wvvov_fac instead (_wvvov_l2_fp64): wvvov
    enters the update once, in an l1 term, so its FP64 products (3.9 TFLOP at (H2O)8)
    would serve a single FP64 residual."""
    C = lambda eq, *ops: contract(eq, *ops, precision=precision)  # noqa: E731
    CW = C if lp_precision is None else (lambda eq, *ops: contract(eq, *ops, precision=lp_precision))  # noqa: E731
    t1 = _dev(t1)
    t2 = _dev(t2)
    nocc, nvir = t1.shape
    fock = eris.fock
    foo, fov, fvv = fock[:nocc, :nocc], fock[:nocc, nocc:], fock[nocc:, nocc:]
    w1 = fvv - C("ja,jb->ba", fov, t1)
    w2 = foo + C("ib,jb->ij", fov, t1)
    w3 = C("kc,jkbc->bj", fov, t2) * 2 + fov.T
    w3 = w3 - C("kc,kjbc->bj", fov, t2)
    w3 = w3 + C("kc,kb,jc->bj", fov, t1, t1)
    w4 = fov.clone()

    ovvv = eris.ovvv
    vvov = ovvv.permute(2, 3, 0, 1)
    tau = t2 + torch.einsum("ia,jb->ijab", t1, t1)

    w1 = w1 + C("jcba,jc->ba", ovvv, t1 * 2) - C("jabc,jc->ba", ovvv, t1)
    theta = _pair_sum(t2, c1=2.0, c3=-1.0)                 # 2 t2[i,j,a,b] - t2[j,i,a,b]
    w3 = w3 + C("jkcd,kdcb->bj", theta, ovvv)
    wVOov = C("jbcd,kd->bjkc", ovvv, t1)
    wvOOv = C("cbjd,kd->cjkb", vvov, -t1)
    g2vovv = mfma.combine([(2, vvov.permute(0, 2, 1, 3)), (-1, vvov.permute(0, 2, 3, 1))])
    wvooo = C("cibd,jkbd->ckij", g2vovv, tau)
    del g2vovv
    # wvvov = wexp (the products with t2, theta and tau) + wcheap (integrals, t1 terms)
    wexp = CW("jabd,jkcd->abkc", ovvv, t2) * -1.5
    wcheap = mfma.combine([(2, vvov.permute(0, 3, 2, 1)), (-1, vvov)])
    g2vvov = mfma.combine([(2, vvov), (-1, ovvv.permute(1, 2, 0, 3))])
    theta_b = _pair_sum(t2, c1=2.0, c2=-1.0)               # 2 t2[i,j,a,b] - t2[i,j,b,a]
    vackb = CW("acjd,kjbd->ackb", g2vvov, theta_b)
    mfma.combine([(1, vackb.permute(0, 3, 2, 1)), (-0.5, vackb)], out=wexp, beta=1)
    del vackb

    ovoo = eris.ovoo
    w2 = w2 + C("kbij,kb->ij", ovoo, t1) * 2 - C("ibkj,kb->ij", ovoo, t1)
    theta = _pair_sum(t2, c1=-1.0, c3=2.0)
    w3 = w3 - C("lckj,klcb->bj", ovoo, theta)
    tmp = C("lc,jcik->ijkl", t1, ovoo)
    woooo = tmp + tmp.permute(1, 0, 3, 2)
    wvOOv = wvOOv + C("lbjk,lc->bjkc", ovoo, t1)
    wVOov = wVOov - C("jbkl,lc->bjkc", ovoo, t1)
    wvooo = wvooo + ovoo.permute(1, 3, 2, 0) * 2 - ovoo.permute(1, 0, 2, 3)
    wvooo = wvooo - C("klbc,iblj->ckij", t2, ovoo * 1.5)
    g2ovoo = ovoo * 2 - ovoo.permute(2, 1, 0, 3)
    theta = _pair_sum(t2, c1=2.0, c3=-1.0)
    vcjik = C("jlcb,lbki->cjki", theta, g2ovoo)
    wvooo = wvooo + vcjik.permute(0, 3, 2, 1) - vcjik * 0.5

    voov = eris.ovvo.permute(1, 0, 3, 2)
    woooo = woooo + C("cijd,klcd->ijkl", voov, tau)
    g2voov = mfma.combine([(2, voov), (-1, voov.permute(0, 2, 1, 3))])
    tmpw4 = C("ckld,ld->kc", g2voov, t1)
    w1 = w1 - C("ckja,kjcb->ba", g2voov, t2) - C("ja,jb->ba", tmpw4, t1)
    w2 = w2 + C("jkbc,bikc->ij", t2, g2voov) + C("ib,jb->ij", tmpw4, t1)
    w3 = w3 + t1.T @ tmpw4 @ t1.T
    w4 = w4 + tmpw4
    # sum_ld voov[b,l,j,d] t1[k,d] t1[l,c] and sum_ld voov[b,j,l,d] t1[k,d] t1[l,c] through
    # y[b,p,q,k] = sum_d voov[b,p,q,d] t1[k,d] (t1 t1 is rank one: o^2 v^2 x o v work, not o^3 v^3)
    y = C("bpqd,kd->bpqk", voov, t1)
    wvOOv = wvOOv + C("bljk,lc->bjkc", y, t1)
    wVOov = wVOov - C("bjlk,lc->bjkc", y, t1)
    del y
    VOov = C("bjld,klcd->bjkc", g2voov, t2) - C("bjld,kldc->bjkc", voov, t2) + voov
    vOOv = C("bljd,kldc->bjkc", voov, t2) - eris.oovv.permute(2, 1, 0, 3)
    wVOov = wVOov + VOov
    wvOOv = wvOOv + vOOv
    ov1 = vOOv * 2 + VOov
    ov2 = VOov * 2 + vOOv
    wvooo = wvooo - C("jb,bikc->ckij", t1, ov1) + C("kb,bijc->ckij", t1, ov2)
    w3 = w3 + C("ckjb,kc->bj", ov2, t1)
    contract_acc(wcheap, "ajkc,jb->abkc", ov1, t1, precision=precision)
    contract_acc(wcheap, "ajkb,jc->abkc", ov2, t1, alpha=-1.0, precision=precision)
    contract_acc(wexp, "laki,klbc->abic", g2ovoo, tau, precision=lp_precision or precision)
    wvvov = wcheap + wexp
    del wexp

    woooo = woooo + eris.oooo.permute(0, 2, 1, 3)
    w3 = w3 + C("bc,jc->bj", w1, t1) - C("kj,kb->bj", w2, t1)
    imds = _IMDS()
    imds.w1, imds.w2, imds.w3, imds.w4 = w1, w2, w3, w4
    imds.woooo, imds.wvooo = woooo, wvooo
    imds.wVOov, imds.wvOOv, imds.wvvov = wVOov, wvOOv, wvvov
    imds.wvvov_fac = None
    if lp_precision is not None:
        imds.wvvov_cheap = wcheap
        imds.wvvov_fac = (g2vvov, theta_b, g2ovoo)
    _static_pieces(imds, t1, t2, eris, getattr(mycc, "level_shift", 0.0), theta=theta_b)
    return imds


def _static_pieces(imds, t1, t2, eris, level_shift, theta=None):
    """Amplitude-only pieces of the lambda update, fixed during a solve (the reference
    recomputes them in every update, gpu_ccsd_lambda.py:227-334)."""
    nocc = t1.shape[0]
    imds.theta = theta if theta is not None else _pair_sum(t2, c1=2.0, c2=-1.0)   # 2 t2 - t2[i,j,b,a]
    imds.tau = t2 + torch.einsum("ia,jb->ijab", t1, t1)
    imds.tmp_voov = (imds.wVOov * 2 + imds.wvOOv).contiguous()
    # the bare integral term of l2new, voov.permute(1, 2, 0, 3) / 2, as the closing
    # symmetrisation sees it: c[i,j,a,b] + c[j,i,b,a]
    imds.l2_sym = _pair_sum(eris.ovvo.permute(0, 3, 1, 2), c1=0.5, c4=0.5)
    # integral combinations that two matrix-vector terms of every update read in two index
    # orders: one combined pass instead of two
    ovvv = eris.ovvv
    imds.g2ovvv = mfma.combine([(2, ovvv), (-1, ovvv.permute(0, 3, 1, 2))])      # 2 (ia|bc) - (ib|ca)
    voov = eris.ovvo.permute(1, 0, 3, 2)
    imds.g2voov_b = mfma.combine([(2, voov), (-1, voov.permute(3, 1, 2, 0))])    # 2 v[a,i,j,b] - v[b,i,j,a]
    e_o = eris.mo_energy[:nocc]
    e_v = eris.mo_energy[nocc:] + level_shift
    eia = e_o[:, None] - e_v
    imds.eia = eia
    imds.inv_d2 = 1.0 / (eia[:, None, :, None] + eia[None, :, None, :])


def update_lambda_gpu(mycc, t1, t2, l1, l2, eris=None, imds=None, use_cpu_vvvv=True, precision="fp64",
                      homogeneous=False, inh=None):
    """One lambda update l -> F(l) (the reference's update_lambda_gpu).

    The update is affine in (l1, l2): F(l) = A l + b. With homogeneous=True the
    l-independent terms b (fov, the bare voov term and the bare t1 inside
    mia1) are left out, giving A l -- used by the mixed-precision solver for
    its correction equations. `inh`, a 0-d device tensor, weights b instead
    (1: F, 0: A), read when the kernels run: one captured graph serves both."""
    with split_memo():
        return _update_lambda(mycc, t1, t2, l1, l2, eris, imds, use_cpu_vvvv, precision, homogeneous, inh)


def _update_lambda(mycc, t1, t2, l1, l2, eris, imds, use_cpu_vvvv, precision, homogeneous, inh):
    C = lambda eq, *ops: contract(eq, *ops, precision=precision)  # noqa: E731

    def acc(out, eq, a, b, alpha=1.0):
        return contract_acc(out, eq, a, b, alpha=alpha, precision=precision)

    if imds is None:
        imds = make_intermediates_gpu(mycc, t1, t2, eris, precision)
    t1, t2, l1, l2 = _dev(t1), _dev(t2), _dev(l1), _dev(l2).contiguous()
    nocc, nvir = t1.shape
    fov = eris.fock[:nocc, nocc:]
    e_o = eris.mo_energy[:nocc]
    e_v = eris.mo_energy[nocc:] + getattr(mycc, "level_shift", 0.0)
    if inh is None:
        inh = 0.0 if homogeneous else 1.0

    if not hasattr(imds, "theta"):
        _static_pieces(imds, t1, t2, eris, getattr(mycc, "level_shift", 0.0))
    theta = imds.theta
    mba = C("klca,klcb->ba", l2, theta)
    mij = C("ikcd,jkcd->ij", l2, theta)
    mba1 = C("jc,jb->bc", l1, t1) + mba
    mij1 = C("kb,jb->kj", l1, t1) + mij
    # 2 t2[j,k,b,c] - t2[j,k,c,b] is theta: one pass over it instead of two over t2
    mia1 = t1 * inh + C("kc,jkbc->jb", l1, theta)
    mia1 = mia1 - t1 @ l1.T @ t1 - C("bd,jd->jb", mba, t1) - C("lj,lb->jb", mij, t1)

    # l2new starts as half the l2 (ac|bd) term; its singles part is 2 X_ijab - X_jiab
    if use_cpu_vvvv:
        l2new = vvvv_term(l2, eris.vvvv, precision, scale=0.5)
    else:
        l2new = torch.zeros_like(l2)
    l1new = l1_terms(torch.zeros_like(t1), l2new, t1, c1=4.0, c3=-2.0)

    w1 = imds.w1 - torch.diag(e_v)
    w2 = imds.w2 - torch.diag(e_o)
    l1new = l1new + fov * inh + C("ib,ba->ia", l1, w1) - C("ja,ij->ia", l1, w2)
    l1new = l1new - C("ik,ka->ia", mij, imds.w4) - C("ca,ic->ia", mba, imds.w4)
    l1new = l1_terms(l1new.contiguous(), l2, imds.w3.T, c1=2.0, c2=-1.0)
    # l2new's l1[i,a] w4[j,b] term and its bare integral term join it in the closing
    # symmetrisation (l2_finish): nothing reads l2new in between
    acc(l2new, "jibc,ca->jiba", l2, w1)
    acc(l2new, "jk,kiba->jiba", w2, l2, -1.0)

    ovoo = eris.ovoo
    l1new = l1new - C("iajk,kj->ia", ovoo, mij1) * 2 + C("jaik,kj->ia", ovoo, mij1)
    acc(l2new, "jbki,ka->jiba", ovoo, l1, -1.0)

    tau = imds.tau
    l2tau = C("ijcd,klcd->ijkl", l2, tau)
    l2t1 = C("jidc,kc->ijkd", l2, t1)

    l1new = l1new - C("jb,jiab->ia", l1, eris.oovv)
    ovvv = eris.ovvv
    l1new = l1new + C("iabc,bc->ia", imds.g2ovvv, mba1)
    acc(l2new, "jbac,ic->jiba", ovvv, l1)
    m4 = C("ijkd,kadb->ijab", l2t1, ovvv).contiguous()
    l2new.sub_(m4)
    l1new = l1_terms(l1new.contiguous(), m4, t1, c1=-2.0, c2=1.0, c3=1.0, c4=-2.0)

    voov = eris.ovvo.permute(1, 0, 3, 2)
    l1new = l1new + C("jb,aijb->ia", l1, voov) * 2
    acc(l2new, "bjic,ca->jiba", voov, mba1, -1.0)
    acc(l2new, "bjka,ik->jiba", voov, mij1, -1.0)
    l1new = l1new + C("aijb,jb->ia", imds.g2voov_b, mia1)
    # l2tau.voov and woooo.l2 enter l2new (x 1/2) and the singles terms alike: one buffer
    m4 = C("ijkl,aklb->ijab", l2tau, voov).contiguous()
    acc(m4, "ijkl,klab->ijab", imds.woooo, l2)
    l2new.add_(m4, alpha=0.5)
    l1new = l1_terms(l1new.contiguous(), m4, t1, c1=2.0, c2=-1.0)

    if precision == "fp64" and getattr(imds, "wvvov_fac", None) is not None:
        l1new = l1new - C("ckij,jkca->ia", imds.wvooo, l2) + _wvvov_l2_fp64(imds, t2, l2, ovvv, l2tau)
    else:
        l1new = l1new - C("ckij,jkca->ia", imds.wvooo, l2) + C("abkc,kibc->ia", imds.wvvov, l2)
    tmp_voov = imds.tmp_voov
    # tmp[i,a,j,b] = l2[i,j,a,b] - l2[i,j,b,a] / 2, stored [i][a][j][b]
    tmp = torch.empty((nocc, nvir, nocc, nvir), dtype=l2.dtype, device=l2.device)
    pair_combine(l2, c1=1.0, c2=-0.5, out=tmp.permute(0, 2, 1, 3))
    acc(l2new, "iakc,bjkc->jiba", tmp, tmp_voov)
    tmp = C("jkca,bikc->jiba", l2, imds.wvOOv)
    pair_combine(tmp, c1=1.0, c3=0.5, out=l2new, beta=1)         # l2new += tmp + tmp[j,i,a,b] / 2


    l1new = l1new / imds.eia
    # symmetrize l2 and divide by D_ijab (the reference's per-i loop, :325-334, as one op),
    # with the l1 w4 and bare integral terms (weight inh) added on the way
    l2new = l2_finish(l2new, l1, imds.w4, imds.l2_sym, inh, imds.inv_d2)
    return l1new, l2new


def _wvvov_l2_fp64(imds, t2, l2, ovvv, l2tau):
    """sum_{bkc} wvvov[a,b,k,c] l2[k,i,b,c] in FP64 without forming wvvov's products:
      -1.5 ovvv.t2 term:   -1.5 sum_{jbd} ovvv[j,a,b,d] X[j,d,i,b],  X = sum_{kc} t2[j,k,c,d] l2[k,i,b,c]
      vackb terms:          sum_{bjd} g2vvov[a,b,j,d] Z[j,d,i,b],    Z = sum_{kc} theta[k,j,c,d] L'[k,i,b,c],
                            L'[k,i,b,c] = l2[k,i,c,b] - l2[k,i,b,c] / 2
      g2ovoo.tau term:      sum_{lkm} g2ovoo[l,a,k,m] l2tau[m,i,k,l]
    plus the formed FP64 remainder (integrals and t1 terms). Two o^3 v^3 products
    (0.45 TFLOP each at (H2O)8) replace 3.9 TFLOP of intermediates."""
    C = lambda eq, *ops: contract(eq, *ops, precision="fp64")  # noqa: E731
    g2vvov, theta_b, g2ovoo = imds.wvvov_fac
    out = C("abkc,kibc->ia", imds.wvvov_cheap, l2)
    x = C("jkcd,kibc->jdib", t2, l2)
    out = out - 1.5 * C("jabd,jdib->ia", ovvv, x)
    del x
    lp = l2.transpose(2, 3) - 0.5 * l2
    z = C("kjcd,kibc->jdib", theta_b, lp)
    del lp
    out = out + C("abjd,jdib->ia", g2vvov, z)
    del z
    return out + C("lakm,mikl->ia", g2ovoo, l2tau)


class GraphedUpdate:
    """l -> F(l) (or A l when homogeneous) captured once as a HIP graph and replayed.

    The capture records every kernel of one lambda update (MFMA GEMMs, splits,
    slice sums, elementwise work) on a side stream after two eager warm-up calls
    (which also fill the static-operand cache); each call then copies (l1, l2)
    into the graph's input buffers and replays. The returned tensors are the
    graph's output buffers: valid until the next call."""

    def __init__(self, fn, l1, l2):
        self.l1 = l1.clone()
        self.l2 = l2.clone()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            # the warm-up call is the first update itself: F(l1, l2)
            w1, w2 = fn(self.l1, self.l2)
            self.first = (w1.clone(), w2.clone())
        torch.cuda.current_stream().wait_stream(side)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.o1, self.o2 = fn(self.l1, self.l2)
        self.replays = 0
        self._first_key = None

    @staticmethod
    def _key(l1, l2):
        return (l1.data_ptr(), l1._version, l2.data_ptr(), l2._version)

    def __call__(self, l1, l2):
        if self.first is not None and self._first_key == self._key(l1, l2):
            out, self.first = self.first, None
            return out
        self.first = None
        self.l1.copy_(l1)
        self.l2.copy_(l2)
        self.graph.replay()
        self.replays += 1
        return self.o1, self.o2

    def bind_first(self, l1, l2):
        """The next call with exactly these (unmodified) tensors returns the warm-up result."""
        self._first_key = self._key(l1, l2)
        return self


def _graph_cache(mycc):
    """The CCSD object's cache of captured lambda updates (None if it takes no attributes)."""
    cache = getattr(mycc, "_lambda_graph_cache", None)
    if cache is None:
        cache = {}
        try:
            mycc._lambda_graph_cache = cache
        except AttributeError:
            return None
    return cache


class _GraphEntry:
    """One captured low-precision lambda update and every buffer it reads.

    The capture records l -> F(l) with the weight `inh` of the l-independent terms as a
    device scalar (1: F(l), 0: A l), so one graph serves the sweeps and the correction
    solves. It reads the amplitudes t1 / t2, the intermediates and the cached
    low-precision copies of both (static_operands state) of the solve that captured it;
    a later solve writes its own values into exactly those buffers (refresh) and
    replays."""

    def __init__(self, mycc, t1, t2, eris, imds, use_cpu_vvvv, prec, l1, l2, homogeneous, static_state):
        # own copies of the amplitudes (a later solve's are written into them, never into the
        # caller's), registered as static operands of the capture; holding eris keeps its id,
        # part of the cache key, from being reused
        self.t1, self.t2 = t1.clone(), t2.clone()
        static_state["ids"].update((id(self.t1), id(self.t2)))
        static_state["keep"] += [self.t1, self.t2]
        self.eris, self.imds, self.static = eris, imds, static_state
        self.inh = torch.full((), 0.0 if homogeneous else 1.0, dtype=torch.float64, device=t1.device)
        # the cached low-precision copies the update reads (the refresh re-makes only these)
        static_state["trace"] = set()
        try:
            self.graph = GraphedUpdate(
                lambda a, b: update_lambda_gpu(mycc, self.t1, self.t2, a, b, eris, imds, use_cpu_vvvv, prec,
                                               inh=self.inh),
                l1, l2).bind_first(l1, l2)
        finally:
            traced = static_state.pop("trace")
        # copies made from the integrals stay valid: an entry only serves solves on the
        # ERIs object it was captured with (part of its cache key), whose blocks are fixed
        fixed = {id(getattr(eris, n)) for n in ("ovvv", "vvvv", "ovoo", "ovvo", "oovv", "oooo")
                 if getattr(eris, n, None) is not None}
        self.keys = {k for k in traced if k[0] not in fixed}

    # the intermediates a low-precision update reads (update_lambda_gpu)
    IMDS_READ = ("w1", "w2", "w3", "w4", "woooo", "wvooo", "wvvov", "wvOOv", "theta", "tau", "tmp_voov", "l2_sym",
                 "g2ovvv", "g2voov_b", "eia", "inv_d2")

    def refresh(self, t1, t2, imds):
        """Write this solve's amplitudes and intermediates into the captured buffers and
        re-prepare the cached low-precision copies in place."""
        self.t1.copy_(t1)
        self.t2.copy_(t2)
        for name in self.IMDS_READ:
            buf, new = getattr(self.imds, name), getattr(imds, name)
            if new is not buf:
                buf.copy_(new)
        refresh_static(self.static, self.keys)

    def __call__(self, l1, l2, homogeneous):
        self.inh.fill_(0.0 if homogeneous else 1.0)
        return self.graph(l1, l2)


def kernel_gpu(mycc, eris=None, t1=None, t2=None, l1=None, l2=None, max_cycle=None, tol=None, verbose=0,
               use_cpu_vvvv=True, precision="fp64", use_graph=None):
    _require_gpu()
    from .cc import device_eris
    eris = device_eris(mycc, eris)
    t1 = _dev(mycc.t1 if t1 is None else t1)
    t2 = _dev(mycc.t2 if t2 is None else t2)
    l1 = t1.clone() if l1 is None else _dev(l1)
    l2 = t2.clone() if l2 is None else _dev(l2)
    max_cycle = mycc.max_cycle if max_cycle is None else max_cycle
    tol = mycc.conv_tol_normt if tol is None else tol
    nocc, nvir = t1.shape
    phases = {}
    mycc._lambda_phases = phases
    tick = _Ticker(phases)
    # intermediates once in FP64: they enter every FP64 residual
    imds = make_intermediates_gpu(mycc, t1, t2, eris, "fp64",
                                  lp_precision=None if precision == "fp64" else precision)
    tick("intermediates")
    diis_space = getattr(mycc, "diis_space", 6)
    use_diis = bool(getattr(mycc, "diis", True))
    # the mixed-precision sweeps replay a captured HIP graph of the low-precision update
    # (mycc.lambda_graph: None = automatic, True / False forces it). The graph is captured
    # once per (CCSD object, integrals, precision, sizes) and replayed by every later solve
    # after this solve's amplitudes, intermediates and static operand copies are written
    # into the buffers it reads (_GraphEntry.refresh); without a CCSD object that takes
    # attributes it lives for this solve only.
    if use_graph is None:
        forced = getattr(mycc, "lambda_graph", None)
        use_graph = precision != "fp64" and (forced is None or bool(forced))
    cache = _graph_cache(mycc) if use_graph else None
    gkey = (precision, id(eris), bool(use_cpu_vvvv), nocc, nvir, t1.device.index)
    entry = cache.get(gkey) if cache is not None else None
    if entry is not None:
        entry.refresh(t1, t2, imds)
        tick("graph buffers refreshed")
    replays0 = entry.graph.replays if entry is not None else 0
    history = []
    mycc._lambda_history = history

    def update(l1, l2, prec, homogeneous):
        nonlocal entry
        if not use_graph or prec == "fp64":
            return update_lambda_gpu(mycc, t1, t2, l1, l2, eris, imds, use_cpu_vvvv, prec, homogeneous=homogeneous)
        if entry is None:
            tick(None)
            entry = _GraphEntry(mycc, t1, t2, eris, imds, use_cpu_vvvv, prec, l1, l2, homogeneous, static_state)
            if cache is not None:
                cache.clear()              # one captured update per CCSD object (its buffers are large)
                cache[gkey] = entry
            tick("graph capture (incl. first update)")
        return entry(l1, l2, homogeneous)

    def fixed_point(l1, l2, prec, tol_, ncyc, rhs=None, accept=None):
        """Iterate l <- F(l) (rhs None) or d <- A d + rhs; returns (conv, l1, l2, norm, iters).

        accept = (tol, margin, state): after each DIIS step, stop with DIIS's input
        combination (DIIS.input_combination) when its residual norm plus `margin` times
        sum |c| |d| is below tol; state["residual"] is then that norm."""
        # low-precision sweeps keep the DIIS history in FP32 (half the bytes read per sweep)
        adiis = DIIS(diis_space, store=None if prec == "fp64" else "fp32") if use_diis else None
        rhs_vec = amplitudes_to_vector(*rhs) if rhs is not None else None
        normt = None
        best, best_it = float("inf"), 0
        prev = amplitudes_to_vector(l1, l2)
        for it in range(ncyc):
            if rhs is None:
                n1, n2 = update(l1, l2, prec, False)
            else:
                n1, n2 = update(l1, l2, prec, True)
            if adiis is not None:
                # the change norm is the newest DIIS error (x_new - previous vector)
                vec, normt = adiis.push((n1, n2), rhs_vec, start=prev if it == 0 else None)
            else:
                vec = amplitudes_to_vector(n1, n2)
                if rhs is not None:
                    vec += rhs_vec
                normt = float(torch.linalg.norm(vec - prev))
            history.append((prec, rhs is not None, normt))
            prev = vec
            if accept is not None and adiis is not None and adiis.res_norm is not None:
                atol, margin, state = accept
                res = adiis.res_norm
                if res + margin * adiis.c_abs * float(torch.linalg.norm(vec)) < atol:
                    state["residual"] = res
                    x1, x2 = vector_to_amplitudes(adiis.input_combination(), nocc, nvir)
                    return True, x1, x2, res, it + 1
            l1, l2 = vector_to_amplitudes(vec, nocc, nvir)
            if normt < tol_:
                return True, l1, l2, normt, it + 1
            if prec != "fp64":
                # low precision has a noise floor: stop once the change norm stalls
                if normt < 0.9 * best:
                    best, best_it = normt, it
                elif it - best_it >= 4:
                    return False, l1, l2, normt, it + 1
        return False, l1, l2, normt, ncyc

    with static_operands(t1, t2, eris.ovvv, eris.vvvv, eris.ovoo, eris.ovvo, eris.oovv, eris.oooo, imds.w1,
                         imds.w2, imds.w3, imds.w4, imds.wvvov, imds.wvooo, imds.wVOov, imds.wvOOv, imds.woooo,
                         imds.theta, imds.tau, imds.tmp_voov, imds.g2ovvv, imds.g2voov_b) as static_state:
        if precision == "fp64":
            conv, l1, l2, normt, iters = fixed_point(l1, l2, "fp64", tol, max_cycle)
            tick("fp64 updates")
            mycc._lambda_iterations = iters
            mycc._lambda_fp64_updates = iters
            mycc._lambda_normt = normt
            return conv, l1, l2
        # mixed precision: low-precision sweeps, then FP64-residual iterative refinement
        out = _mixed_solve(mycc, fixed_point,
                           lambda _cc, _t1, _t2, a, b, _e, _i, _u, prec, homogeneous=False: update(a, b, prec,
                                                                                                  homogeneous),
                           t1, t2, l1, l2, eris, imds, use_cpu_vvvv, precision, tol, max_cycle, tick)
        mycc._lambda_graph_replays = entry.graph.replays - replays0 if entry is not None else 0
        return out


class _Ticker:
    """Wall time per solve phase (seconds, accumulated by name; synchronises the
    device at each mark, a handful of times per solve)."""

    def __init__(self, phases):
        self.phases = phases
        torch.cuda.synchronize()
        self.t = time.perf_counter()

    def __call__(self, name):
        torch.cuda.synchronize()
        t = time.perf_counter()
        if name is not None:
            self.phases[name] = self.phases.get(name, 0.0) + t - self.t
        self.t = t


def _mixed_solve(mycc, fixed_point, update, t1, t2, l1, l2, eris, imds, use_cpu_vvvv, precision, tol, max_cycle,
                 tick=lambda name: None):
    lp_tol = max(tol, getattr(mycc, "lambda_lp_tol", 1e-6 if precision == "fp32" else 3e-6))
    _, l1, l2, normt, iters = fixed_point(l1, l2, precision, lp_tol, max_cycle)
    tick("low-precision sweeps")
    total, fp64_updates = iters, 0
    conv = False
    if not getattr(mycc, "lambda_refine", True):
        # low-precision fixed point only (its noise floor; for accuracy studies)
        mycc._lambda_iterations = total
        mycc._lambda_fp64_updates = 0
        mycc._lambda_normt = normt
        return normt < tol, l1, l2
    # residual update after a correction d: r(l + d) = r(l) + A d - d exactly (F affine),
    # with A d from one low-precision homogeneous update (a third of an FP64 update's
    # cost). Its error is that precision's rounding of A d, of order 1e-5 |d| for
    # split-BF16; convergence is accepted only with a margin of est_margin |d| on top,
    # otherwise the FP64 residual is recomputed.
    res_update = getattr(mycc, "lambda_residual_update", True)
    est_margin = getattr(mycc, "lambda_est_margin", 1e-3)
    # DIIS's own residual: with A d in low precision the correction iteration is affine,
    # so the combination of the correction sweeps' inputs with the DIIS coefficients has
    # the residual sum_s c_s E[s] exactly up to the low precision's rounding of each A d_s
    # (bounded by est_margin sum_s |c_s| |d|); accepting on it saves the sweep of the
    # residual update
    diis_res = getattr(mycc, "lambda_diis_residual", True)
    r1 = r2 = None
    for _round in range(max(8, max_cycle // 4)):
        if r1 is None:
            r1, r2 = update(mycc, t1, t2, l1, l2, eris, imds, use_cpu_vvvv, "fp64")
            fp64_updates += 1
            tick("fp64 residuals")
            r1, r2 = r1 - l1, r2 - l2
            normt = float(torch.linalg.norm(amplitudes_to_vector(r1, r2)))
            mycc._lambda_history.append(("fp64-residual", False, normt))
            if normt < tol:
                conv = True
                break
        # correction: d = A d + r, to `lambda_corr_rtol` of the remaining residual in at most
        # `lambda_corr_max_cycle` sweeps: short rounds restart DIIS on a fresh FP64 residual
        # while its convergence is still fast (measured on (H2O)8 with the first phase to
        # 3e-6: 15 sweeps + 2 FP64 updates with 0.01 / 5, against 27 + 2 with one long
        # round to 1e-3)
        rtol = getattr(mycc, "lambda_corr_rtol", 1e-2)
        ncyc = getattr(mycc, "lambda_corr_max_cycle", 5)
        state = {}
        _, d1, d2, _, it = fixed_point(r1.clone(), r2.clone(), precision, max(normt * rtol, tol * 1e-2), ncyc,
                                       rhs=(r1, r2), accept=(tol, est_margin, state) if diis_res else None)
        total += it
        l1, l2 = l1 + d1, l2 + d2
        tick("correction sweeps")
        if "residual" in state:
            normt = state["residual"]
            mycc._lambda_history.append(("diis-residual", True, normt))
            conv = True
            break
        if res_update:
            a1, a2 = update(mycc, t1, t2, d1, d2, eris, imds, use_cpu_vvvv, precision, homogeneous=True)
            e1, e2 = r1 + a1 - d1, r2 + a2 - d2
            dnorm = float(torch.linalg.norm(amplitudes_to_vector(d1, d2)))
            est = float(torch.linalg.norm(amplitudes_to_vector(e1, e2)))
            total += 1
            mycc._lambda_history.append(("residual-update", True, est))
            tick("residual updates")
            if est + est_margin * dnorm < tol:
                normt = est
                conv = True
                break
        r1 = r2 = None
    mycc._lambda_iterations = total
    mycc._lambda_fp64_updates = fp64_updates
    mycc._lambda_normt = normt
    return conv, l1, l2


def solve_lambda_gpu(mycc, t1=None, t2=None, l1=None, l2=None, eris=None, use_cpu_vvvv=True,
                     fallback_to_cpu=True, precision="fp64", return_tensors=False):
    """Solve the CCSD lambda equations on the GPU (ccsd_gpu/gpu_ccsd_lambda.py:411-449).

    mycc is duck-typed as in the reference: this build's CCSD or a PySCF-like object
    (t1, t2, mo_coeff, mo_occ, _scf.mo_energy, max_cycle, conv_tol_normt, diis_space,
    level_shift; eris through ao2mo() or the `eris` argument, NumPy blocks accepted).
    Side effects as the reference: mycc.converged_lambda, mycc.l1, mycc.l2 and
    mycc._lambda_solver_mode = "gpu-hybrid"; mycc._lambda_precision names the
    contraction precision. Returns (l1, l2) as NumPy arrays, like the reference, or
    as device tensors with return_tensors=True.

    Without a HIP device this raises RuntimeError whatever `fallback_to_cpu` says:
    the reference then falls back to PySCF's CPU solver (or raises when
    fallback_to_cpu=False); this build has no CPU path.
    """
    if not torch.cuda.is_available():
        if not fallback_to_cpu:
            raise RuntimeError("No usable HIP device for the lambda solve.")
        raise RuntimeError("No usable HIP device for the lambda solve, and this build has no CPU fallback "
                           "(fallback_to_cpu=True cannot be honoured).")
    from .cc import device_eris
    eris = device_eris(mycc, eris)
    conv, l1o, l2o = kernel_gpu(mycc, eris=eris, t1=t1, t2=t2, l1=l1, l2=l2, max_cycle=mycc.max_cycle,
                                tol=mycc.conv_tol_normt, use_cpu_vvvv=use_cpu_vvvv, precision=precision)
    mycc.converged_lambda = conv
    if not return_tensors:
        l1o, l2o = l1o.cpu().numpy(), l2o.cpu().numpy()
    mycc.l1 = l1o
    mycc.l2 = l2o
    mycc._lambda_solver_mode = "gpu-hybrid"
    mycc._lambda_precision = precision
    return mycc.l1, mycc.l2
