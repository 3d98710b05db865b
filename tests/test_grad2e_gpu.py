"""Fused int2e_ip1 x Gamma contraction and the ERI autograd Function on the GPU.

  * fused kernel == unfused path (ip1 tensor in HBM, then contraction) to 1e-10 relative;
  * fused kernel == oracle (C Rys ip1 x Gamma) to 1e-10 relative;
  * ERI.backward == central finite differences (h = 1e-4 Bohr) of sum G_ijkl (ij|kl),
    with the forward evaluated by the same kernels, and == the oracle contraction.
"""
import numpy as np
import pytest
import torch

from ccsd_gpu_amd import systems
from ccsd_gpu_amd.autograd import ERI, symmetrize8
from ccsd_gpu_amd.hip.grad2e import contract_int2e_ip1, pack_s2kl, sph_to_cart_4index
from ccsd_gpu_amd.hip.int2e_ip1_gpu import compute_int2e_ip1_gpu
from oracle import integrals as oi

from . import molecules


def _ao_atoms(mol):
    loc = mol.ao_loc_nr(cart=True)
    out = np.zeros(mol.nao_cart(), dtype=np.int64)
    for ib in range(mol.nbas):
        out[loc[ib]:loc[ib + 1]] = mol.bas_atom(ib)
    return out


def _random_gamma(mol, seed):
    nc = mol.nao_cart()
    rng = np.random.default_rng(seed)
    return rng.normal(size=(nc, nc, nc * (nc + 1) // 2))


@pytest.mark.gpu
@pytest.mark.parametrize("make", [molecules.h2o_ccpvdz, molecules.ch4_sto3g, molecules.g_shell_mol])
def test_fused_equals_unfused_and_oracle(gpu, make):
    mol = make()
    gam = _random_gamma(mol, 7)
    g_t = torch.as_tensor(gam, device="cuda")
    de = contract_int2e_ip1(mol, g_t, scale=-2.0).cpu().numpy()
    mol_c = mol.copy()
    mol_c.cart = True
    ip1 = compute_int2e_ip1_gpu(mol_c, 0, mol.nbas).cpu().numpy()
    aoa = _ao_atoms(mol)
    per = np.einsum("xijp,ijp->xi", ip1, gam)
    ref = np.zeros((mol.natm, 3))
    np.add.at(ref, aoa, -2.0 * per.T)
    scale = np.abs(ref).max()
    print("fused vs unfused", np.abs(de - ref).max() / scale, flush=True)
    assert np.abs(de - ref).max() <= 1e-10 * scale
    ipo = oi.int2e_s2kl_cart(mol_c, 0, mol.nbas, 1)
    per = np.einsum("xijp,ijp->xi", ipo, gam)
    ref2 = np.zeros((mol.natm, 3))
    np.add.at(ref2, aoa, -2.0 * per.T)
    assert np.abs(de - ref2).max() <= 1e-10 * scale


@pytest.mark.gpu
def test_fused_partial_shell_ranges_sum_to_total(gpu):
    mol = molecules.h2o_ccpvdz()
    g_t = torch.as_tensor(_random_gamma(mol, 3), device="cuda")
    tot = contract_int2e_ip1(mol, g_t)
    part = torch.zeros_like(tot)
    for b0 in range(mol.nbas):
        contract_int2e_ip1(mol, g_t, b0=b0, b1=b0 + 1, de=part)
    assert torch.allclose(tot, part, rtol=1e-12, atol=1e-12)


@pytest.mark.gpu
def test_eri_autograd_backward_vs_finite_difference(gpu):
    mol = systems.water(seed=0)
    n = mol.nao
    gen = torch.Generator(device="cuda").manual_seed(11)
    G = symmetrize8(torch.randn((n, n, n, n), dtype=torch.float64, device="cuda", generator=gen))
    coords = torch.tensor(mol.atom_coords(), dtype=torch.float64, device="cuda", requires_grad=True)
    eri = ERI.apply(coords, mol)
    (eri * G).sum().backward()
    grad = coords.grad.cpu().numpy()
    h = 1e-4
    fd = np.zeros_like(grad)
    base = mol.atom_coords()
    for a in range(mol.natm):
        for x in range(3):
            vals = []
            for s in (+1, -1):
                c = base.copy()
                c[a, x] += s * h
                ct = torch.tensor(c, dtype=torch.float64, device="cuda")
                vals.append(float((ERI.apply(ct, mol) * G).sum()))
            fd[a, x] = (vals[0] - vals[1]) / (2 * h)
    print("autograd vs FD", np.abs(grad - fd).max(), np.abs(fd).max(), flush=True)
    assert np.abs(grad - fd).max() < 1e-6 * max(1.0, np.abs(fd).max())
    # oracle: -4 sum_{i in A, jkl} ip1 G (full tensors, spherical basis)
    ip1 = oi.ip1_full(mol)
    Gc = G.cpu().numpy()
    per = np.einsum("xijkl,ijkl->xi", ip1, Gc)
    ref = np.zeros((mol.natm, 3))
    np.add.at(ref, mol.ao_atom_index(), -4.0 * per.T)
    assert np.abs(grad - ref).max() <= 1e-10 * np.abs(ref).max()


@pytest.mark.gpu
@pytest.mark.parametrize("make", [molecules.h2o_ccpvdz, molecules.ch4_ccpvdz, molecules.hf_sto3g])
def test_register_kernels_equal_lds_kernel(gpu, make):
    """The per-class register kernels and the generic LDS kernel give the same contraction."""
    from ccsd_gpu_amd import _lib
    mol = make()
    g_t = torch.as_tensor(_random_gamma(mol, 21), device="cuda")
    plan = _lib.get_plan(mol)
    lib = _lib.load()
    _lib.check(lib.ccsd_plan_set_option(plan, 0, 0), "set_option")
    ref = contract_int2e_ip1(mol, g_t).cpu().numpy()
    _lib.check(lib.ccsd_plan_set_option(plan, 0, 1), "set_option")
    got = contract_int2e_ip1(mol, g_t).cpu().numpy()
    scale = np.abs(ref).max()
    print("reg vs lds", np.abs(got - ref).max() / scale, flush=True)
    assert np.abs(got - ref).max() <= 1e-11 * scale


@pytest.mark.gpu
@pytest.mark.parametrize("make", [molecules.h2o_ccpvdz, molecules.ch4_ccpvdz])
@pytest.mark.parametrize("split", [2, 3, 8])
def test_primitive_split_register_kernels_equal_lds_kernel(gpu, make, split):
    """Register kernels with the ket primitive-pair loop split over `split` waves (forced,
    autotuning off so the split kernel is the one run) give the LDS kernel's contraction."""
    from ccsd_gpu_amd import _lib
    mol = make()
    g_t = torch.as_tensor(_random_gamma(mol, 23), device="cuda")
    plan = _lib.get_plan(mol)
    lib = _lib.load()
    try:
        _lib.check(lib.ccsd_plan_set_option(plan, 0, 0), "set_option")
        ref = contract_int2e_ip1(mol, g_t).cpu().numpy()
        _lib.check(lib.ccsd_plan_set_option(plan, 0, 1), "set_option")
        _lib.check(lib.ccsd_plan_set_option(plan, 1, 0), "set_option")
        _lib.check(lib.ccsd_plan_set_option(plan, 4, split), "set_option")
        got = contract_int2e_ip1(mol, g_t).cpu().numpy()
        launches = _lib_launches(mol)
    finally:
        _lib.check(lib.ccsd_plan_set_option(plan, 4, 0), "set_option")
        _lib.check(lib.ccsd_plan_set_option(plan, 1, 1), "set_option")
    scale = np.abs(ref).max()
    print("split", split, "reg vs lds", np.abs(got - ref).max() / scale, flush=True)
    assert np.abs(got - ref).max() <= 1e-11 * scale
    assert any(rec["kernel"] == 1 and rec["chunk"] >> 16 == split for rec in launches)


def _lib_launches(mol):
    from ccsd_gpu_amd.hip.grad2e import last_launches
    return last_launches(mol)


@pytest.mark.gpu
def test_fused_explicit_plus_separable_in_one_pass(gpu):
    """gamma + separable(D, P) in one kernel pass equals the explicit sum of both."""
    from ccsd_gpu_amd.gpu_ccsd_grad import _separable_gamma
    from ccsd_gpu_amd.hip.grad2e import contract_int2e_ip1, pack_s2kl
    mol = molecules.h2o_ccpvdz(cart=False)
    nc = mol.nao_cart()
    gen = torch.Generator(device="cuda").manual_seed(21)
    from ccsd_gpu_amd.autograd import symmetrize8
    g = symmetrize8(torch.randn((nc,) * 4, dtype=torch.float64, device="cuda", generator=gen))
    a = torch.randn((nc, nc), dtype=torch.float64, device="cuda", generator=gen)
    b = torch.randn((nc, nc), dtype=torch.float64, device="cuda", generator=gen)
    D, P = a + a.T, b + b.T
    ref = contract_int2e_ip1(mol, pack_s2kl(g + _separable_gamma(D, P), halve_diag=True))
    got = contract_int2e_ip1(mol, pack_s2kl(g, halve_diag=True), dm=D, pm=P)
    err = float((got - ref).abs().max())
    print("one-pass explicit+separable", err, flush=True)
    assert err < 1e-10 * max(1.0, float(ref.abs().max()))


@pytest.mark.gpu
def test_sym_pairs_kernel_matches_torch(gpu):
    from ccsd_gpu_amd.gpu_ccsd_grad import _lib_sym_pairs
    gen = torch.Generator(device="cuda").manual_seed(22)
    for n in (1, 5, 13):
        t = torch.randn((n,) * 4, dtype=torch.float64, device="cuda", generator=gen)
        ref = 0.5 * (t + t.permute(1, 0, 2, 3) + t.permute(0, 1, 3, 2) + t.permute(1, 0, 3, 2))
        got = _lib_sym_pairs(t.clone(), 0.5)
        assert float((got - ref).abs().max()) < 1e-14


@pytest.mark.gpu
@pytest.mark.parametrize("make", [molecules.h2o_ccpvdz, molecules.ch4_sto3g])
def test_add_separable_then_explicit_pass_equals_in_kernel_separable(gpu, make):
    """gamma += Gsep(D, P) (packed, diagonal halved) by the HIP kernel equals the torch
    construction, and an explicit fused pass over it equals the pass that forms the
    separable part inside the kernel (dm=, pm=)."""
    from ccsd_gpu_amd.gpu_ccsd_grad import _separable_gamma
    from ccsd_gpu_amd.hip.grad2e import add_separable, pack_s2kl
    mol = make()
    nc = mol.nao_cart()
    rng = np.random.default_rng(17)
    D = rng.standard_normal((nc, nc)); D = D + D.T
    P = rng.standard_normal((nc, nc)) * 0.3          # not symmetric: the kernel must not assume it
    g0 = torch.as_tensor(_random_gamma(mol, 29), device="cuda")
    Dt, Pt = torch.as_tensor(D, device="cuda"), torch.as_tensor(P, device="cuda")
    ref_g = g0 + pack_s2kl(_separable_gamma(Dt, Pt), halve_diag=True)
    got_g = add_separable(g0.clone(), Dt, Pt)
    assert float((got_g - ref_g).abs().max()) <= 1e-13 * float(ref_g.abs().max())
    ref = contract_int2e_ip1(mol, g0, scale=-2.0, dm=Dt, pm=Pt).cpu().numpy()
    got = contract_int2e_ip1(mol, got_g, scale=-2.0).cpu().numpy()
    scale = np.abs(ref).max()
    print(make.__name__, "add_separable + explicit vs in-kernel separable", np.abs(got - ref).max() / scale, flush=True)
    assert np.abs(got - ref).max() <= 1e-11 * scale


@pytest.mark.gpu
@pytest.mark.parametrize("make", [molecules.h2o_ccpvdz, molecules.ch4_ccpvdz])
def test_fused_gradient_is_bitwise_reproducible(gpu, make):
    """The fused pass adds into dE/dR through exact fixed-point integer atomics: on one
    plan (kernel choices fixed) repeated passes, on four side streams or on one
    stream, give the same bits; the persistent single launch is reproducible too."""
    from ccsd_gpu_amd import _lib
    mol = make()
    g_t = torch.as_tensor(_random_gamma(mol, 31), device="cuda")
    plan = _lib.get_plan(mol)
    lib = _lib.load()
    runs = [contract_int2e_ip1(mol, g_t, scale=-2.0) for _ in range(3)]
    try:
        _lib.check(lib.ccsd_plan_set_option(plan, 2, 0), "set_option")     # one stream
        runs.append(contract_int2e_ip1(mol, g_t, scale=-2.0))
        _lib.check(lib.ccsd_plan_set_option(plan, 5, 1), "set_option")     # persistent launch
        pers = [contract_int2e_ip1(mol, g_t, scale=-2.0) for _ in range(2)]
    finally:
        lib.ccsd_plan_set_option(plan, 2, 1)
        lib.ccsd_plan_set_option(plan, 5, 0)
    for r in runs[1:]:
        assert torch.equal(r, runs[0])
    assert torch.equal(pers[0], pers[1])
    scale = float(runs[0].abs().max())
    print("persistent vs per-class", float((pers[0] - runs[0]).abs().max()) / scale, flush=True)
    assert float((pers[0] - runs[0]).abs().max()) <= 1e-11 * scale
