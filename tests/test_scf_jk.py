"""Packed symmetric-density J/K (hermi=1) against the full-tensor contraction."""
import torch

from ccsd_gpu_amd import scf


def test_packed_jk_matches_full_contraction():
    n = 7
    g = torch.Generator().manual_seed(0)
    e = torch.randn(n, n, n, n, dtype=torch.float64, generator=g)
    e = e + e.permute(1, 0, 2, 3)
    e = e + e.permute(0, 1, 3, 2)
    e = e + e.permute(2, 3, 0, 1)                     # 8-fold symmetric, like (ij|kl)
    ints = scf.AOIntegrals.__new__(scf.AOIntegrals)
    ints.eri_c = e
    k, l = torch.tril_indices(n, n)
    ints.eri_p = e[:, :, k, l].contiguous()
    ints.c2s = torch.eye(n, dtype=torch.float64)
    d = torch.randn(n, n, dtype=torch.float64, generator=g)
    d = d + d.T
    jr = torch.einsum("ijkl,kl->ij", e, d)
    kr = torch.einsum("ijkl,jl->ik", e, d)
    for hermi in (1, 0):
        j, kk = ints.get_jk(d, hermi=hermi)
        assert torch.allclose(j, jr, rtol=1e-13, atol=1e-12)
        assert torch.allclose(kk, kr, rtol=1e-13, atol=1e-12)
