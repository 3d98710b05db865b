"""Fused ERI-derivative contraction: sum_ijkl d(ij|kl)/dR * Gamma without storing d(ij|kl).

The reference forms the derivative integrals of one shell block in HBM and
contracts them with a GEMV (`de_k -= eri1.reshape(3, -1) @ dm2buf.reshape(-1) * 2`,
ccsd_gpu/gpu_ccsd_grad.py:449-459). Here the contraction happens inside the
int2e_ip1 kernel: every wave dots its quartet's derivative integrals with the
matching Gamma entries and adds three numbers into dE/dR.

Gamma is given in the Cartesian AO basis in the s2kl layout
(nao_c, nao_c, nao_c (nao_c+1)/2): gamma[i, j, k(k+1)/2 + l], k >= l. A
spherical-basis density is brought there with `sph_to_cart_s2kl` (the full
contraction is basis independent: sum_sph X_sph G_sph = sum_cart X_cart (C G)).
"""
import torch

from .. import _lib


def contract_int2e_ip1(mol, gamma_cart, scale=-2.0, b0=0, b1=None, de=None, dm=None, pm=None):
    """de[atom(i), x] += scale * sum_{i in [b0,b1), j, k>=l} ip1_x(ij|kl) gamma[i, j, kl].

    With dm and pm (Cartesian nao_c x nao_c device matrices) the separable density
    D_ij P_kl + P_ij D_kl - (D_il P_kj + D_ik P_lj + P_il D_jk + P_ik D_jl)/4 (halved on
    k == l) is added to gamma inside the kernel. Returns de as a (natm, 3) float64
    tensor on the current device."""
    lib = _lib.load()
    b1 = mol.nbas if b1 is None else b1
    if not (0 <= b0 < b1 <= mol.nbas):
        raise ValueError("bad shell range")
    nao_c = mol.nao_cart()
    shape = (nao_c, nao_c, nao_c * (nao_c + 1) // 2)
    if tuple(gamma_cart.shape) != shape or gamma_cart.dtype != torch.float64 or not gamma_cart.is_contiguous() \
            or not gamma_cart.is_cuda:
        raise ValueError(f"gamma_cart must be a contiguous float64 device tensor of shape {shape}")
    if de is None:
        de = torch.zeros((mol.natm, 3), dtype=torch.float64, device=gamma_cart.device)
    elif tuple(de.shape) != (mol.natm, 3) or de.dtype != torch.float64 or not de.is_contiguous():
        raise ValueError("de must be a contiguous float64 (natm, 3) tensor")
    plan = _lib.get_plan(mol)
    if dm is None:
        _lib.check(lib.ccsd_grad_ip1_gamma(plan, _lib.ptr(de), _lib.ptr(gamma_cart), float(scale), b0, b1,
                                           _lib.current_stream_ptr()), "ccsd_grad_ip1_gamma")
        return de
    pm = dm if pm is None else pm
    for m in (dm, pm):
        if tuple(m.shape) != (nao_c, nao_c) or m.dtype != torch.float64 or not m.is_contiguous() or not m.is_cuda:
            raise ValueError(f"dm / pm must be contiguous float64 device tensors of shape {(nao_c, nao_c)}")
    _lib.check(lib.ccsd_grad_ip1_gamma_sep(plan, _lib.ptr(de), _lib.ptr(gamma_cart), _lib.ptr(dm), _lib.ptr(pm),
                                           float(scale), b0, b1, _lib.current_stream_ptr()),
               "ccsd_grad_ip1_gamma_sep")
    return de


def tril_pairs(n, device):
    k, l = torch.tril_indices(n, n, device=device)
    return k, l


def pack_s2kl(full, halve_diag=False):
    """(n, n, n, n) -> (n, n, n(n+1)/2) lower-triangle pairs of the last two indices."""
    n = full.shape[-1]
    k, l = tril_pairs(n, full.device)
    out = full[..., k, l].contiguous()
    if halve_diag:
        out[..., k == l] *= 0.5
    return out


def sph_to_cart_4index(g_sph, c2s):
    """Gamma_cart = (C x C x C x C) Gamma_sph with C = cart2sph_coeff (nao_c, nao_s)."""
    t = torch.tensordot(c2s, g_sph, dims=([1], [0]))                     # (c, s, s, s)
    t = torch.tensordot(c2s, t, dims=([1], [1])).transpose(0, 1)         # (c, c, s, s)
    t = torch.tensordot(t, c2s, dims=([2], [1])).permute(0, 1, 3, 2)     # (c, c, c, s)
    t = torch.tensordot(t, c2s, dims=([3], [1]))                         # (c, c, c, c)
    return t.contiguous()


def set_screening(mol, thr):
    """Schwarz screening of the fused gradient for `mol`'s plan: quartets with
    q_ij q_kl < thr are skipped (thr = 0: none skipped, the default)."""
    _lib.check(_lib.load().ccsd_plan_screen(_lib.get_plan(mol), float(thr), _lib.current_stream_ptr()),
               "ccsd_plan_screen")


def add_separable(gamma_cart, dm, pm):
    """In place: gamma_cart (nc, nc, nc(nc+1)/2, packed k >= l, diagonal halved) +=
    the separable density of dm and pm (nc x nc) in the same packed form, so that an
    explicit pass of contract_int2e_ip1 over it equals the pass with dm= and pm=."""
    n = dm.shape[0]
    if gamma_cart.shape != (n, n, n * (n + 1) // 2) or not gamma_cart.is_contiguous():
        raise ValueError("gamma_cart must be a contiguous (n, n, n(n+1)/2) tensor")
    dm, pm = dm.contiguous(), pm.contiguous()
    _lib.check(_lib.load().ccsd_add_separable_packed(_lib.ptr(gamma_cart), _lib.ptr(dm), _lib.ptr(pm), n,
                                                     _lib.current_stream_ptr()), "ccsd_add_separable_packed")
    return gamma_cart


PERSISTENT_BODIES = {"register": 1, "generic": 2}


def set_persistent(mol, on=True, body="register"):
    """Fused passes of `mol`'s plan as one persistent launch over every class (True) or
    one launch per class on four streams (False, the default). body="register" runs the
    register-resident per-class bodies in that one kernel (when every class of the pass
    has one; the LDS kernel's body otherwise), "generic" the LDS kernel's body."""
    mode = PERSISTENT_BODIES[body] if on else 0
    _lib.check(_lib.load().ccsd_plan_set_option(_lib.get_plan(mol), 5, mode), "ccsd_plan_set_option")


def schwarz_factors(mol):
    """(nbas, nbas) device tensor q_ij = max_{a in i, b in j} sqrt((ab|ab))."""
    q = torch.empty((mol.nbas, mol.nbas), dtype=torch.float64, device="cuda")
    _lib.check(_lib.load().ccsd_schwarz(_lib.get_plan(mol), _lib.ptr(q), _lib.current_stream_ptr()), "ccsd_schwarz")
    return q


def last_quartets(mol):
    """Shell quartets evaluated by the last integral / fused call on `mol`'s plan."""
    n = _lib.ctypes.c_longlong()
    _lib.check(_lib.load().ccsd_plan_quartets(_lib.get_plan(mol), _lib.ctypes.byref(n)), "ccsd_plan_quartets")
    return n.value


LAUNCH_FIELDS = ("li", "lj", "lk", "ll", "gc", "kernel", "waves_per_block", "chunk", "parts", "quartets")


def last_launches(mol=None, plan=None):
    """Launch records of the last integral / fused call on a plan (a molecule's
    cached plan, or an explicit handle): one dict per launch, in launch order."""
    import numpy as np
    lib = _lib.load()
    plan = _lib.get_plan(mol) if plan is None else plan
    n = _lib.ctypes.c_int()
    _lib.check(lib.ccsd_plan_last_launches(plan, None, 0, _lib.ctypes.byref(n)), "ccsd_plan_last_launches")
    buf = np.zeros((max(1, n.value), 10), dtype=np.int64)
    _lib.check(lib.ccsd_plan_last_launches(plan, buf.ctypes.data, n.value, _lib.ctypes.byref(n)),
               "ccsd_plan_last_launches")
    return [dict(zip(LAUNCH_FIELDS, map(int, row))) for row in buf[:n.value]]
