"""GPU int2e_ip1 in PySCF `intor` layout.

compute_int2e_ip1_gpu(mol, b0, b1) is the drop-in for

    mol.intor('int2e_ip1', comp=3, aosym='s2kl', shls_slice=(b0, b1, 0, nbas, 0, nbas, 0, nbas))

reshaped to (3, nf, nao, nao_pair), as ccsd_gpu/cuda/int2e_ip1_gpu.py:296-326
returns it: comp = x, y, z of d/dr1 (the electron coordinate of function i),
out[comp, fi_local, fj, k(k+1)/2 + l] with k >= l, native AO basis.
"""
import torch

from .. import _lib
from .int2e_gpu import _block_dims, _check_mol  # noqa: F401


def compute_int2e_ip1_gpu(mol, b0, b1, out=None):
    nf, nao, npair = _block_dims(mol, b0, b1)
    lib = _lib.load()
    plan = _lib.get_plan(mol)
    if out is None:
        out = torch.empty((3, nf, nao, npair), dtype=torch.float64, device="cuda")
    elif out.shape != (3, nf, nao, npair) or out.dtype != torch.float64 or not out.is_contiguous():
        raise ValueError("out must be a contiguous float64 tensor of shape (3, nf, nao, nao_pair)")
    _lib.check(lib.ccsd_int2e_ip1_s2kl(plan, _lib.ptr(out), b0, b1, 1 if mol.cart else 0,
                                       _lib.current_stream_ptr()), "ccsd_int2e_ip1_s2kl")
    return out
