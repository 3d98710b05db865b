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


def _check_kinds(rows, body):
    """Launch-record kernel kinds (3 register bodies, 2 generic body). Register mode:
    the classes with a register body (every l <= 2 class that fits
    eight parts) in the register-body launch, the rest -- three or more d shells, or
    (pp|dd) -- in the generic-body launch; generic mode: every class generic."""
    assert rows
    if body == "generic":
        assert all(r["kernel"] == 2 for r in rows)
        return
    assert any(r["kernel"] == 3 for r in rows)
    for r in rows:
        assert r["kernel"] in (2, 3)
        if r["kernel"] == 2:
            assert r["li"] + r["lj"] + r["lk"] + r["ll"] >= 6, r


@pytest.mark.gpu
@pytest.mark.parametrize("body", ["register", "generic"])
@pytest.mark.parametrize("make", [molecules.h2o_ccpvdz, molecules.ch4_ccpvdz])
def test_persistent_explicit_gamma_equals_per_class_and_oracle(gpu, make, body):
    from ccsd_gpu_amd.hip.grad2e import contract_int2e_ip1, last_launches, last_quartets, set_persistent
    mol = make()
    nc = mol.nao_cart()
    gam = np.random.default_rng(31).normal(size=(nc, nc, nc * (nc + 1) // 2))
    g_t = torch.as_tensor(gam, device="cuda")
    ref = contract_int2e_ip1(mol, g_t, scale=-2.0).cpu().numpy()
    n_ref = last_quartets(mol)
    print(f"{make.__name__}: persistent launch, {body} body", flush=True)
    set_persistent(mol, True, body=body)
    try:
        got = contract_int2e_ip1(mol, g_t, scale=-2.0).cpu().numpy()
        again = contract_int2e_ip1(mol, g_t, scale=-2.0).cpu().numpy()
        n_got = last_quartets(mol)
        rows = last_launches(mol)
    finally:
        set_persistent(mol, False)
    scale = np.abs(ref).max()
    print(f"{make.__name__}: persistent ({body}) vs per-class {np.abs(got - ref).max() / scale:.2e}, "
          f"repeat {np.abs(again - got).max() / scale:.2e}, classes {len(rows)}", flush=True)
    assert n_got == n_ref
    _check_kinds(rows, body)
    # register bodies: the per-class kernels' work lists and fixed-point accumulation,
    # compiled separately (floating-point contraction may differ in the last bit)
    assert np.abs(got - ref).max() <= (1e-13 if body == "register" else 1e-11) * scale
    assert np.array_equal(again, got)
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
@pytest.mark.parametrize("body,order", [("register", "class"), ("register", "cost"), ("generic", "cost")])
def test_persistent_screened_batch_equals_per_class(gpu, body, order):
    """Six C/N/O/H molecules in one plan, Schwarz threshold 1e-12: one persistent launch
    (register bodies in either work-item order, or the generic body) gives the per-class
    launches' gradient, quartet count and zero net force per molecule, and repeats
    itself bit for bit."""
    from ccsd_gpu_amd.batch import MoleculeBatch
    from ccsd_gpu_amd.systems import cnoh_batch
    mols = cnoh_batch(6, seed=3)
    rng = np.random.default_rng(8)
    dms = [_sym(m.nao_nr(), rng) for m in mols]
    b = MoleculeBatch(mols)
    b.screen(1e-12)
    ref = torch.cat(b.grad2e_separable(dms)).clone()
    n_ref = b.last_quartets()
    print(f"batch persistent launch, {body} body, {order} order", flush=True)
    b.set_persistent(True, body=body, order=order)
    got = torch.cat(b.grad2e_separable(dms)).clone()
    n_got = b.last_quartets()
    again = torch.cat(b.grad2e_separable(dms)).clone()
    rows = b.last_launches()
    b.set_persistent(False)
    scale = float(ref.abs().max())
    d1 = float((got - ref).abs().max()) / scale
    d2 = float((again - got).abs().max()) / scale
    print(f"batch persistent ({body}, {order}) vs per-class {d1:.2e}, repeat {d2:.2e}, quartets {n_got}", flush=True)
    assert n_got == n_ref
    _check_kinds(rows, body)
    assert d1 <= (1e-13 if body == "register" else 1e-11) and d2 == 0.0
    for i in range(len(mols)):
        part = got[b.atom0[i]:b.atom0[i + 1]]
        assert float(part.sum(0).abs().max()) < 1e-9 * max(1.0, float(part.abs().max()))
