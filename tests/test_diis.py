"""DIIS: the piecewise push (vector written straight into the history buffer) against
the whole-vector update, and the change norm it returns."""
import numpy as np
import torch

from ccsd_gpu_amd.cc import DIIS


def _seq(seed, n_it=9, n1=6, n2=20):
    g = torch.Generator().manual_seed(seed)
    return [(torch.randn(2, 3, generator=g, dtype=torch.float64), torch.randn(2, 2, 5, generator=g,
                                                                                 dtype=torch.float64))
            for _ in range(n_it)]


def test_push_matches_update_and_reports_change_norm():
    for rhs in (None, torch.randn(26, dtype=torch.float64, generator=torch.Generator().manual_seed(9))):
        a, b = DIIS(4), DIIS(4)
        start = torch.randn(26, dtype=torch.float64, generator=torch.Generator().manual_seed(1))
        prev = start
        for it, (p1, p2) in enumerate(_seq(3)):
            x = torch.cat([p1.reshape(-1), p2.reshape(-1)])
            if rhs is not None:
                x = x + rhs
            want_norm = float(torch.linalg.norm(x - prev))
            ref = a.update(x)
            got, nrm = b.push((p1, p2), rhs, start=start if it == 0 else None)
            assert torch.allclose(got, ref, rtol=0, atol=1e-12)
            assert abs(nrm - want_norm) <= 1e-10 * max(1.0, want_norm)
            prev = ref


def test_push_first_call_without_start_has_no_norm():
    d = DIIS(3)
    p = (torch.ones(2, dtype=torch.float64), torch.zeros(3, dtype=torch.float64))
    v, nrm = d.push(p)
    assert nrm is None and np.allclose(v.numpy(), [1, 1, 0, 0, 0])
