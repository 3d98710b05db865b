"""GPU int2e in PySCF `intor` layout.

compute_int2e_gpu(mol, b0, b1) is the drop-in for

    mol.intor('int2e', aosym='s2kl', shls_slice=(b0, b1, 0, nbas, 0, nbas, 0, nbas))

reshaped to (nf, nao, nao_pair), as ccsd_gpu/cuda/int2e_gpu.py:186-215 returns
it: out[fi_local, fj, k(k+1)/2 + l] with k >= l in the native AO basis
(Cartesian if mol.cart else spherical). The result is a float64 torch tensor on
the current HIP device; the computation runs in libccsd_gpu_amd.so.
"""
import torch

from .. import _lib

_L_MAX = 4


def _check_mol(mol):
    """True iff the kernels support every shell of mol (l <= 4), as the reference's
    _check_mol (ccsd_gpu/cuda/int2e_ip1_gpu.py:36-40)."""
    return int(mol._bas[:, 1].max()) <= _L_MAX


def _block_dims(mol, b0, b1):
    if not (0 <= b0 < b1 <= mol.nbas):
        raise ValueError(f"bad shell range [{b0}, {b1}) for nbas={mol.nbas}")
    loc = mol.ao_loc_nr()
    nao = int(loc[-1])
    return int(loc[b1] - loc[b0]), nao, nao * (nao + 1) // 2


def compute_int2e_gpu(mol, b0, b1, out=None):
    nf, nao, npair = _block_dims(mol, b0, b1)
    lib = _lib.load()
    plan = _lib.get_plan(mol)
    if out is None:
        out = torch.empty((nf, nao, npair), dtype=torch.float64, device="cuda")
    elif out.shape != (nf, nao, npair) or out.dtype != torch.float64 or not out.is_contiguous():
        raise ValueError("out must be a contiguous float64 tensor of shape (nf, nao, nao_pair)")
    _lib.check(lib.ccsd_int2e_s2kl(plan, _lib.ptr(out), b0, b1, 1 if mol.cart else 0,
                                   _lib.current_stream_ptr()), "ccsd_int2e_s2kl")
    return out
