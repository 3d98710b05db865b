"""Batched, Schwarz-screened fused gradient (multi-molecule plans) against the
oracle and against the single-molecule explicit-Gamma path.

There is no batched path in the reference (one molecule per call,
ccsd_gpu/cuda/int2e_ip1_gpu.py:296); what is pinned here is that batching and
screening change nothing beyond the screening threshold: per-molecule results
equal the oracle's int2e_ip1 contracted with the same four-index density, the
Schwarz factors equal the oracle's diagonal integrals, and translational
invariance (the derivatives of one quartet over its four centres sum to zero)
holds for every molecule of a large batch.
"""
import numpy as np
import pytest
import torch

from oracle import integrals as oi

from . import molecules


def _sym(n, rng):
    a = rng.standard_normal((n, n))
    return (a + a.T) / np.sqrt(n)


def _gamma_sep_np(D, P):
    g = np.einsum("ij,kl->ijkl", D, P) + np.einsum("ij,kl->ijkl", P, D)
    g -= 0.25 * (np.einsum("il,kj->ijkl", D, P) + np.einsum("ik,lj->ijkl", D, P)
                 + np.einsum("il,jk->ijkl", P, D) + np.einsum("ik,jl->ijkl", P, D))
    return g


@pytest.mark.gpu
def test_schwarz_factors_match_oracle(gpu):
    from ccsd_gpu_amd.hip.grad2e import schwarz_factors
    mol = molecules.h2o_ccpvdz(cart=True)
    eri = oi.int2e_full(mol)
    loc = mol.ao_loc_nr(cart=True)
    ref = np.zeros((mol.nbas, mol.nbas))
    for i in range(mol.nbas):
        for j in range(mol.nbas):
            blk = eri[loc[i]:loc[i + 1], loc[j]:loc[j + 1], loc[i]:loc[i + 1], loc[j]:loc[j + 1]]
            ref[i, j] = np.sqrt(np.einsum("abab->ab", blk).max())
    q = schwarz_factors(mol).cpu().numpy()
    err = np.abs(q - ref).max()
    print("schwarz err", err, flush=True)
    assert err < 1e-10


@pytest.mark.gpu
def test_separable_batch_matches_oracle(gpu):
    """Two molecules in one plan (Cartesian H2O, spherical CH4) vs oracle ip1 x Gamma_sep."""
    from ccsd_gpu_amd.batch import MoleculeBatch
    rng = np.random.default_rng(4)
    mols = [molecules.h2o_ccpvdz(cart=True), molecules.ch4_sto3g(cart=False)]
    dms = [_sym(m.nao_nr(), rng) for m in mols]
    pms = [_sym(m.nao_nr(), rng) for m in mols]
    b = MoleculeBatch(mols)
    got = [x.cpu().numpy() for x in b.grad2e_separable(dms, pms, scale=-2.0)]
    for m, D, P, g in zip(mols, dms, pms, got):
        mc = m.copy()
        mc.cart = True
        if not m.cart:
            c = m.cart2sph_coeff()
            D, P = c @ D @ c.T, c @ P @ c.T
        ip1 = oi.ip1_full(mc)
        t = np.einsum("xijkl,ijkl->xi", ip1, _gamma_sep_np(D, P))      # sum over j, k, l
        ref = np.zeros((m.natm, 3))
        atom_of_ao = np.repeat(mc._bas[:, 0], np.diff(mc.ao_loc_nr(cart=True)))
        for i, a in enumerate(atom_of_ao):
            ref[a] += -2.0 * 0.5 * t[:, i]                               # k >= l with halved diagonal
        err = np.abs(g - ref).max()
        print("separable vs oracle", m.natm, err, flush=True)
        assert err < 1e-10


@pytest.mark.gpu
def test_separable_equals_explicit_gamma_path(gpu):
    from ccsd_gpu_amd.batch import MoleculeBatch
    from ccsd_gpu_amd.gpu_ccsd_grad import _separable_gamma
    from ccsd_gpu_amd.hip.grad2e import contract_int2e_ip1, pack_s2kl
    from ccsd_gpu_amd.systems import random_cnoh
    rng = np.random.default_rng(5)
    mols = [random_cnoh(11, 3), molecules.h2o_ccpvdz(cart=False), random_cnoh(12, 2)]
    dms = [_sym(m.nao_nr(), rng) for m in mols]
    b = MoleculeBatch(mols)
    got = b.grad2e_separable(dms, scale=-2.0)
    for m, D, g in zip(mols, dms, got):
        c = torch.as_tensor(m.cart2sph_coeff(), device="cuda")
        Dc = c @ torch.as_tensor(D, device="cuda") @ c.T
        gp = pack_s2kl(_separable_gamma(Dc, Dc), halve_diag=True)
        ref = contract_int2e_ip1(m, gp, scale=-2.0)
        err = float((g - ref).abs().max())
        print("batch vs explicit", m.natm, err, flush=True)
        assert err < 1e-10 * max(1.0, float(ref.abs().max()))


@pytest.mark.gpu
def test_batch_screening_invariance_and_tuning(gpu):
    """Screened batch: within the threshold of the unscreened one, fewer quartets,
    zero net force per molecule, and the same numbers with the autotuner off and
    with every class on the LDS kernel."""
    from ccsd_gpu_amd.batch import MoleculeBatch
    from ccsd_gpu_amd.systems import cnoh_batch
    mols = cnoh_batch(6, seed=3)
    rng = np.random.default_rng(6)
    dms = [_sym(m.nao_nr(), rng) for m in mols]
    b = MoleculeBatch(mols)
    full = torch.cat(b.grad2e_separable(dms)).clone()
    nfull = b.last_quartets()
    assert nfull == b.quartets_unscreened()
    b.screen(1e-12)
    scr = torch.cat(b.grad2e_separable(dms)).clone()
    nscr = b.last_quartets()
    print("quartets", nfull, nscr, "diff", float((scr - full).abs().max()), flush=True)
    assert nscr < nfull
    assert float((scr - full).abs().max()) < 1e-8
    for i, m in enumerate(mols):
        part = scr[b.atom0[i]:b.atom0[i + 1]]
        assert float(part.sum(0).abs().max()) < 1e-9 * max(1.0, float(part.abs().max()))
    b.set_option(1, 0)
    notune = torch.cat(b.grad2e_separable(dms))
    assert float((notune - scr).abs().max()) < 1e-11
    b.set_option(0, 0)
    lds = torch.cat(b.grad2e_separable(dms))
    assert float((lds - scr).abs().max()) < 1e-11


@pytest.mark.gpu
def test_screened_explicit_gamma_single_molecule(gpu):
    from ccsd_gpu_amd.autograd import symmetrize8
    from ccsd_gpu_amd.hip.grad2e import contract_int2e_ip1, last_quartets, pack_s2kl, set_screening
    from ccsd_gpu_amd.systems import water_cluster
    mol = water_cluster(2, seed=2, cart=True)
    n = mol.nao_cart()
    gen = torch.Generator(device="cuda").manual_seed(2)
    g = symmetrize8(torch.randn((n, n, n, n), dtype=torch.float64, device="cuda", generator=gen))
    gp = pack_s2kl(g, halve_diag=True)
    ref = contract_int2e_ip1(mol, gp).clone()
    n0 = last_quartets(mol)
    set_screening(mol, 1e-12)
    try:
        got = contract_int2e_ip1(mol, gp)
        n1 = last_quartets(mol)
    finally:
        set_screening(mol, 0.0)
    print("explicit screened", n0, n1, float((got - ref).abs().max()), flush=True)
    assert n1 < n0
    assert float((got - ref).abs().max()) < 1e-8
