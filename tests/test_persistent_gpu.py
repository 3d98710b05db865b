"""The fused int2e_ip1 x Gamma pass as one persistent launch over every class.

The reference launches one kernel per i-shell (ccsd_gpu/cuda/int2e_ip1_gpu.py:199-214)
and has no batched or persistent form. What is pinned: the single persistent launch
(waves drawing work items from a device counter) gives the per-class launches'
gradient, which the other tests pin against the oracle, for an explicit Gamma on one
molecule and for a Schwarz-screened separable Gamma over a batch of molecules; it
evaluates the same quartets; and a second call (counter cleared again) repeats the first.
"""
import numpy as np
import pytest
import torch

from oracle import integrals as oi

from . import molecules


def _sym(n, rng):
    a = rng.standard_normal((n, n))
    return (a + a.T) / np.sqrt(n)


@pytest.mark.gpu
@pytest.mark.parametrize("make", [molecules.h2o_ccpvdz, molecules.ch4_ccpvdz])
def test_persistent_explicit_gamma_equals_per_class_and_oracle(gpu, make):
    from ccsd_gpu_amd.hip.grad2e import contract_int2e_ip1, last_launches, last_quartets, set_persistent
    mol = make()
    nc = mol.nao_cart()
    gam = np.random.default_rng(31).normal(size=(nc, nc, nc * (nc + 1) // 2))
    g_t = torch.as_tensor(gam, device="cuda")
    ref = contract_int2e_ip1(mol, g_t, scale=-2.0).cpu().numpy()
    n_ref = last_quartets(mol)
    set_persistent(mol, True)
    try:
        got = contract_int2e_ip1(mol, g_t, scale=-2.0).cpu().numpy()
        again = contract_int2e_ip1(mol, g_t, scale=-2.0).cpu().numpy()
        n_got = last_quartets(mol)
        rows = last_launches(mol)
    finally:
        set_persistent(mol, False)
    scale = np.abs(ref).max()
    print(f"{make.__name__}: persistent vs per-class {np.abs(got - ref).max() / scale:.2e}, "
          f"repeat {np.abs(again - got).max() / scale:.2e}, classes {len(rows)}", flush=True)
    assert n_got == n_ref
    assert rows and all(r["kernel"] == 2 for r in rows)
    assert np.abs(got - ref).max() <= 1e-11 * scale
    assert np.abs(again - got).max() <= 1e-11 * scale
    # and against the oracle's int2e_ip1 contracted on the CPU
    mol_c = mol.copy()
    mol_c.cart = True
    ipo = oi.int2e_s2kl_cart(mol_c, 0, mol.nbas, 1)
    loc = mol_c.ao_loc_nr(cart=True)
    aoa = np.zeros(nc, dtype=np.int64)
    for ib in range(mol.nbas):
        aoa[loc[ib]:loc[ib + 1]] = mol.bas_atom(ib)
    oref = np.zeros((mol.natm, 3))
    np.add.at(oref, aoa, -2.0 * np.einsum("xijp,ijp->xi", ipo, gam).T)
    assert np.abs(got - oref).max() <= 1e-10 * np.abs(oref).max()


@pytest.mark.gpu
def test_persistent_screened_batch_equals_per_class(gpu):
    """Six C/N/O/H molecules in one plan, Schwarz threshold 1e-12: one persistent launch
    gives the per-class launches' gradient, quartet count and zero net force per molecule."""
    from ccsd_gpu_amd.batch import MoleculeBatch
    from ccsd_gpu_amd.systems import cnoh_batch
    mols = cnoh_batch(6, seed=3)
    rng = np.random.default_rng(8)
    dms = [_sym(m.nao_nr(), rng) for m in mols]
    b = MoleculeBatch(mols)
    b.screen(1e-12)
    ref = torch.cat(b.grad2e_separable(dms)).clone()
    n_ref = b.last_quartets()
    b.set_persistent(True)
    got = torch.cat(b.grad2e_separable(dms)).clone()
    n_got = b.last_quartets()
    again = torch.cat(b.grad2e_separable(dms)).clone()
    b.set_persistent(False)
    scale = float(ref.abs().max())
    d1 = float((got - ref).abs().max()) / scale
    d2 = float((again - got).abs().max()) / scale
    print(f"batch persistent vs per-class {d1:.2e}, repeat {d2:.2e}, quartets {n_got}", flush=True)
    assert n_got == n_ref
    assert d1 <= 1e-11 and d2 <= 1e-11
    for i in range(len(mols)):
        part = got[b.atom0[i]:b.atom0[i + 1]]
        assert float(part.sum(0).abs().max()) < 1e-9 * max(1.0, float(part.abs().max()))
