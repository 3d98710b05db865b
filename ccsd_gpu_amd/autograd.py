"""Electron-repulsion integrals as a differentiable function of the nuclear coordinates.

    eri = ERI.apply(coords, mol)          # (nao, nao, nao, nao), PySCF s1 order
    eri.backward(G)                       # coords.grad = sum_ijkl G_ijkl d(ij|kl)/dR

Forward runs the HIP int2e kernel. Backward never stores derivative integrals:
it symmetrizes the upstream tensor over the 8 index permutations (only that
part couples to the symmetric ERIs), brings it to the Cartesian s2kl layout and
runs the fused int2e_ip1 contraction kernel. With d/dR_A (ij|kl) =
-[ip1(ij|kl) 1_{i in A} + ip1(ji|kl) 1_{j in A} + ip1(kl|ij) 1_{k in A} + ip1(lk|ij) 1_{l in A}]
and an 8-fold symmetric G, all four terms are equal:
    dL/dR_A = -4 sum_{i in A, jkl} ip1(ij|kl) G_ijkl = -8 sum_{i in A, j, k>=l} ip1 G'_packed,
G'_packed having its k = l entries halved.
"""
import numpy as np
import torch

from .hip.grad2e import contract_int2e_ip1, pack_s2kl, sph_to_cart_4index
from .hip.int2e_gpu import compute_int2e_gpu


def unpack_s2kl(packed):
    """(n, n, n(n+1)/2) -> (n, n, n, n)."""
    n = packed.shape[0]
    k, l = torch.tril_indices(n, n, device=packed.device)
    full = torch.empty((n, n, n, n), dtype=packed.dtype, device=packed.device)
    full[:, :, k, l] = packed
    full[:, :, l, k] = packed
    return full


def symmetrize8(g):
    s = g + g.transpose(0, 1)
    s = s + s.transpose(2, 3)
    s = s + s.permute(2, 3, 0, 1)
    return s * 0.125


class ERI(torch.autograd.Function):
    @staticmethod
    def forward(ctx, coords, mol):
        m = mol.copy()
        m.set_geom_(coords.detach().cpu().numpy())
        ctx.mol = m
        packed = compute_int2e_gpu(m, 0, m.nbas)
        return unpack_s2kl(packed)

    @staticmethod
    def backward(ctx, g):
        m = ctx.mol
        gs = symmetrize8(g.to(torch.float64))
        if not m.cart:
            c2s = torch.as_tensor(m.cart2sph_coeff(), dtype=torch.float64, device=g.device)
            gs = sph_to_cart_4index(gs, c2s)
        gp = pack_s2kl(gs, halve_diag=True)
        de = contract_int2e_ip1(m, gp, scale=-8.0)
        return de.to(g.device), None


def eri(mol, coords=None):
    if coords is None:
        coords = torch.as_tensor(np.asarray(mol.atom_coords()), dtype=torch.float64, device="cuda")
    return ERI.apply(coords, mol)
