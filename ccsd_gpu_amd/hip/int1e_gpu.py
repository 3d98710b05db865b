"""One-electron integrals and their nabla-bra derivatives from the HIP kernels.

int1e(mol, kind) returns, in the native AO basis, what PySCF's
mol.intor('int1e_ovlp' | 'int1e_kin' | 'int1e_nuc' | 'int1e_ipovlp' | 'int1e_ipkin')
returns; 'ipnuc_atom' gives <d/dr i| -Z_C/|r-C| |j> per nucleus (natm, 3, nao, nao),
whose sum over C is int1e_ipnuc and whose row C is -Z_C * int1e_iprinv at C
(the pieces of hcore_generator used at ccsd_gpu/gpu_ccsd_grad.py:556-577).
"""
import torch

from .. import _lib

KINDS = {"ovlp": 0, "kin": 1, "nuc": 2, "ipovlp": 3, "ipkin": 4, "ipnuc_atom": 5}


def int1e(mol, kind, c2s=None):
    k = KINDS[kind]
    lib = _lib.load()
    plan = _lib.get_plan(mol)
    nc = mol.nao_cart()
    nmat = 1 if k <= 2 else (3 if k <= 4 else 3 * mol.natm)
    out = torch.empty((nmat, nc, nc), dtype=torch.float64, device="cuda")
    _lib.check(lib.ccsd_int1e(plan, k, _lib.ptr(out), _lib.current_stream_ptr()), "ccsd_int1e")
    if not mol.cart:
        if c2s is None:
            c2s = torch.as_tensor(mol.cart2sph_coeff(), dtype=torch.float64, device="cuda")
        out = c2s.T @ out @ c2s
    if k <= 2:
        return out[0]
    if k <= 4:
        return out
    n = out.shape[-1]
    return out.reshape(mol.natm, 3, n, n)
