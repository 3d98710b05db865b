"""2-RDM builder and assembly (the product's torch code run on the CPU) against the
oracle's restatement of ccsd_gpu/gpu_ccsd_grad.py:129-311, on random amplitudes with
the (ij)(ab) pair symmetry of converged t2 / l2. The product returns the vvvv block
as gvvvv and reads it through that symmetry in the assembly, so the assembled
tensor is the comparison."""
import numpy as np
import torch

from oracle import chem as oracle


def test_gamma2_and_assembly_match_oracle(monkeypatch):
    import ccsd_gpu_amd.gpu_ccsd_grad as G
    monkeypatch.setattr(G, "_dev", lambda x: torch.as_tensor(x))
    monkeypatch.setattr(G, "DEV", "cpu")
    o, v = 3, 6
    rng = np.random.default_rng(4)
    t1, l1 = rng.standard_normal((o, v)) * 0.1, rng.standard_normal((o, v)) * 0.1
    t2 = rng.standard_normal((o, o, v, v)) * 0.1
    t2 = t2 + t2.transpose(1, 0, 3, 2)
    l2 = rng.standard_normal((o, o, v, v)) * 0.1
    l2 = l2 + l2.transpose(1, 0, 3, 2)
    ref = oracle.assemble_dm2_mo(oracle.gamma2_intermediates(t1, t2, l1, l2))
    T = lambda a: torch.as_tensor(a)  # noqa: E731
    got = G._assemble_dm2_mo(G._gamma2_intermediates_gpu(None, T(t1), T(t2), T(l1), T(l2))).numpy()
    assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


def test_presym_assembly_equals_full_after_pair_symmetrisation(monkeypatch):
    """The packed path's assembly (vvvv block as 4 A - 2 B) differs from the full one
    elementwise but not after the (pq)(rs) symmetrisation that follows it."""
    import ccsd_gpu_amd.gpu_ccsd_grad as G
    monkeypatch.setattr(G, "_dev", lambda x: torch.as_tensor(x))
    monkeypatch.setattr(G, "DEV", "cpu")
    o, v = 2, 5
    rng = np.random.default_rng(9)
    t1, l1 = rng.standard_normal((o, v)) * 0.1, rng.standard_normal((o, v)) * 0.1
    t2 = rng.standard_normal((o, o, v, v)) * 0.1
    t2 = t2 + t2.transpose(1, 0, 3, 2)
    l2 = rng.standard_normal((o, o, v, v)) * 0.1
    l2 = l2 + l2.transpose(1, 0, 3, 2)
    T = lambda a: torch.as_tensor(a)  # noqa: E731
    d2 = G._gamma2_intermediates_gpu(None, T(t1), T(t2), T(l1), T(l2))
    full = G._assemble_dm2_mo(d2)
    pre = G._assemble_dm2_mo(d2, presym=True)
    sym = lambda x: (x + x.permute(1, 0, 2, 3) + x.permute(0, 1, 3, 2) + x.permute(1, 0, 3, 2)) / 2  # noqa: E731
    assert float((pre - full).abs().max()) > 1e-6
    assert float((sym(pre) - sym(full)).abs().max()) <= 1e-13 * float(full.abs().max())
    # and the packed layout the gradient uses, from either
    assert float((G._pack_pairs_sym(pre) - G._pack_pairs_sym(full)).abs().max()) <= 1e-13 * float(full.abs().max())
