"""GPU int2e / int2e_ip1 parity against the oracle, shell block by shell block.

Mirrors tests/test_int2e_gpu.py:99-185 and tests/test_int2e_ip1_gpu.py:109-200
of the reference: every (b0, b1) = (ish, ish+1) block compared with the full
s2kl tensor at max-abs error < 1e-10, Cartesian and spherical molecules,
including f (cc-pVTZ) and g shells (nroots up to 9).
"""
import numpy as np
import pytest
import torch

from oracle import integrals as oi

from . import molecules

ATOL = 1e-10

CASES = [
    ("H2/STO-3G cart", molecules.h2_sto3g, True),
    ("CH4/STO-3G cart", molecules.ch4_sto3g, True),
    ("H2O/cc-pVDZ cart", molecules.h2o_ccpvdz, True),
    ("H2O/cc-pVDZ sph", molecules.h2o_ccpvdz, False),
    ("CH4/cc-pVDZ sph", molecules.ch4_ccpvdz, False),
    ("H2O/cc-pVTZ sph", molecules.h2o_ccpvtz, False),
    ("H2O/def2-SVP sph", molecules.h2o_def2svp, False),
    ("g-shell sph", molecules.g_shell_mol, False),
    ("g-shell cart", molecules.g_shell_mol, True),
]


def _per_shell(mol, fn, ref_full, ncomp):
    loc = mol.ao_loc_nr()
    worst = 0.0
    for ish in range(mol.nbas):
        out = fn(mol, ish, ish + 1).cpu().numpy()
        ref = ref_full[..., loc[ish]:loc[ish + 1], :, :]
        assert out.shape == ref.shape
        worst = max(worst, float(np.abs(out - ref).max()))
    return worst


@pytest.mark.gpu
@pytest.mark.parametrize("name,make,cart", CASES, ids=[c[0] for c in CASES])
def test_int2e_gpu_matches_oracle(gpu, name, make, cart):
    from ccsd_gpu_amd.hip import compute_int2e_gpu
    mol = make(cart=cart)
    ref = oi.intor_int2e(mol)
    print(f"{name}: nao={mol.nao} nbas={mol.nbas}", flush=True)
    worst = _per_shell(mol, compute_int2e_gpu, ref, 1)
    full = compute_int2e_gpu(mol, 0, mol.nbas).cpu().numpy()
    worst = max(worst, float(np.abs(full - ref).max()))
    print(f"  int2e max|err| = {worst:.2e}", flush=True)
    assert worst < ATOL


@pytest.mark.gpu
@pytest.mark.parametrize("name,make,cart", CASES, ids=[c[0] for c in CASES])
def test_int2e_ip1_gpu_matches_oracle(gpu, name, make, cart):
    from ccsd_gpu_amd.hip import compute_int2e_ip1_gpu
    mol = make(cart=cart)
    ref = oi.intor_int2e_ip1(mol)
    print(f"{name}: nao={mol.nao} nbas={mol.nbas}", flush=True)
    worst = _per_shell(mol, compute_int2e_ip1_gpu, ref, 3)
    full = compute_int2e_ip1_gpu(mol, 0, mol.nbas).cpu().numpy()
    worst = max(worst, float(np.abs(full - ref).max()))
    print(f"  int2e_ip1 max|err| = {worst:.2e}", flush=True)
    assert worst < ATOL


@pytest.mark.gpu
def test_int2e_gpu_is_deterministic(gpu):
    from ccsd_gpu_amd.hip import compute_int2e_gpu, compute_int2e_ip1_gpu
    mol = molecules.h2o_ccpvdz(cart=False)
    a = compute_int2e_gpu(mol, 0, mol.nbas)
    b = compute_int2e_gpu(mol, 0, mol.nbas)
    assert torch.equal(a, b)
    a = compute_int2e_ip1_gpu(mol, 2, 5)
    b = compute_int2e_ip1_gpu(mol, 2, 5)
    assert torch.equal(a, b)


@pytest.mark.gpu
def test_rys_roots_gpu_matches_golden(gpu):
    import ctypes
    import os
    from ccsd_gpu_amd import _lib
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "rys_golden.npz"))
    lib = _lib.load()
    for n in range(1, 10):
        m = d["n"] == n
        T = torch.tensor(d["T"][m], dtype=torch.float64, device="cuda")
        r = torch.empty((T.numel(), n), dtype=torch.float64, device="cuda")
        w = torch.empty_like(r)
        _lib.check(lib.ccsd_rys_roots(n, _lib.ptr(T), T.numel(), _lib.ptr(r), _lib.ptr(w),
                                      _lib.current_stream_ptr()), "ccsd_rys_roots")
        np.testing.assert_allclose(r.cpu().numpy(), d["mp_r"][m][:, :n], rtol=1e-13, atol=0)
        np.testing.assert_allclose(w.cpu().numpy(), d["mp_w"][m][:, :n], rtol=1e-13, atol=0)
    del ctypes


@pytest.mark.gpu
@pytest.mark.parametrize("deriv", [0, 1])
@pytest.mark.parametrize("make,cart", [(molecules.h2o_ccpvdz, True), (molecules.ch4_ccpvdz, False)])
def test_register_store_kernels_equal_lds_kernel(gpu, make, cart, deriv):
    """The register-resident store kernels (classes with l <= 2, general contractions
    included) write the same s2kl tensor as the generic LDS kernel, every element
    (zeros of cut quartets included: a Cartesian output starting as NaN is fully set)."""
    from ccsd_gpu_amd import _lib
    from ccsd_gpu_amd.hip import compute_int2e_gpu, compute_int2e_ip1_gpu
    from ccsd_gpu_amd.hip.grad2e import last_launches
    mol = make(cart=cart)
    fn = compute_int2e_ip1_gpu if deriv else compute_int2e_gpu
    plan = _lib.get_plan(mol)
    lib = _lib.load()
    try:
        _lib.check(lib.ccsd_plan_set_option(plan, 0, 0), "set_option")
        ref = fn(mol, 0, mol.nbas).cpu().numpy()
        assert all(r["kernel"] == 0 for r in last_launches(mol))
    finally:
        _lib.check(lib.ccsd_plan_set_option(plan, 0, 1), "set_option")
    got = fn(mol, 0, mol.nbas)
    # with a Cartesian molecule the kernels write straight into `out`: every element must be set
    out = fn(mol, 0, mol.nbas, out=torch.full_like(got, float("nan"))) if cart else None
    launches = last_launches(mol)
    assert any(r["kernel"] == 1 for r in launches)
    got = got.cpu().numpy()
    err = np.abs(got - ref).max()
    print(make.__name__, "deriv", deriv, "register store vs LDS", err,
          sum(r["kernel"] == 1 for r in launches), "of", len(launches), "launches on register kernels", flush=True)
    assert err < 1e-12 * max(1.0, np.abs(ref).max())
    if out is not None:
        assert not torch.isnan(out).any()
        assert np.abs(out.cpu().numpy() - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.gpu
@pytest.mark.parametrize("deriv", [0, 1])
def test_water_ccpvqz_g_shell_blocks_match_oracle(gpu, deriv):
    """H2O/cc-pVQZ (ccsd_gpu tests/test_int2e_gpu.py:89-96, test_int2e_ip1_gpu.py:99-106):
    the i-shell blocks of O's g shell, its f shells and the first H s shell, spherical,
    against the oracle at 1e-10 (nine Rys roots for int2e_ip1 of the g block)."""
    from ccsd_gpu_amd.hip import compute_int2e_gpu, compute_int2e_ip1_gpu
    mol = molecules.h2o_ccpvqz(cart=False)
    ls = mol._bas[:, 1]
    blocks = [int(np.nonzero(ls == 4)[0][0])] + [int(i) for i in np.nonzero(ls == 3)[0][:2]] + \
             [int(np.nonzero(mol._bas[:, 0] == 1)[0][0])]
    fn = compute_int2e_ip1_gpu if deriv else compute_int2e_gpu
    worst = 0.0
    for b in blocks:
        got = fn(mol, b, b + 1).cpu().numpy()
        ref = oi._intor(mol, b, b + 1, deriv)
        ref = ref if deriv else ref[0]
        worst = max(worst, float(np.abs(got - ref).max()))
    print("H2O/cc-pVQZ blocks", blocks, "deriv", deriv, "max|err|", worst, flush=True)
    assert worst < ATOL
