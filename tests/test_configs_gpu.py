"""The benchmark configurations and the reference's gradient set, end to end on the GPU,
against the FP64 oracle.

  * single water (seed 0, cc-pVDZ): full int2e (24^4) and int2e_ip1 (3 x 24^4) as s1
    tensors at 1e-10, and the whole RHF -> CCSD -> lambda -> dE/dR pipeline against
    the committed oracle gradient (tests/golden/grad_golden.npz) at 1e-6 Eh/Bohr;
  * (H2O)4: ERI as a torch.autograd.Function, backward (fused int2e_ip1) against
    central finite differences (h = 1e-4 Bohr) and the oracle's ip1 contraction;
    the whole gradient (mixed and plain paths) against the oracle fixture;
  * 64 C/N/O/H molecules: one screened pass over the batch, per-molecule zero net
    force, screened vs unscreened, the smallest molecules against the oracle;
  * (H2O)6 lambda in FP32 and split-BF16, refined to 1e-8: l1, l2, the RDMs and the
    gradient against the FP64 lambda;
  * (H2O)8 full path: mixed vs plain FP64 and vs the oracle fixture when present;
  * the reference's gradient set (tests/test_gpu_ccsd_grad.py:142-156): HF/STO-3G,
    CH4/cc-pVDZ, H2O/cc-pVTZ at 1e-6 Eh/Bohr;
  * frozen orbitals; PySCF-like plain-attribute stand-ins for mf / mycc / eris.
"""
import os

import numpy as np
import pytest
import torch

from ccsd_gpu_amd import systems
from ccsd_gpu_amd.gto import M
from oracle import chem
from oracle import integrals as oi

from . import molecules

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "grad_golden.npz")


def _golden(name):
    if not os.path.exists(GOLDEN):
        pytest.skip("grad_golden.npz not generated")
    d = np.load(GOLDEN)
    if f"{name}/grad" not in d:
        pytest.skip(f"{name} not in grad_golden.npz")
    syms = [str(s) for s in d[f"{name}/symbols"]]
    xyz = d[f"{name}/coords"]
    atom = "; ".join(f"{s} {x:.16f} {y:.16f} {z:.16f}" for s, (x, y, z) in zip(syms, xyz))
    mol = M(atom, basis=str(d[f"{name}/basis"]), unit="Bohr")
    return mol, float(d[f"{name}/e_hf"]), float(d[f"{name}/e_corr"]), d[f"{name}/grad"]


def _pipeline_gradient(mol, mode="plain", precision="bf16x3"):
    from ccsd_gpu_amd import pipeline
    mf, mycc = pipeline.setup(mol)
    de, info = pipeline.ccsd_gradient(mf, mycc, mode, precision)
    return mf, mycc, de, info


def _oracle_gradient(mol, frozen=None):
    e, mf, cc = chem.ccsd_total_energy(mol, frozen=frozen)
    conv, l1, l2 = chem.solve_lambda(cc.t1, cc.t2, cc.ao2mo())
    assert conv
    return mf, cc, chem.grad_elec(cc, cc.t1, cc.t2, l1, l2) + chem.grad_nuc(mol)


# ----------------------------------------------------------------------------------
# single water (seed 0)
# ----------------------------------------------------------------------------------
@pytest.mark.gpu
def test_single_water_full_int2e_and_ip1_tensors(gpu):
    from ccsd_gpu_amd.autograd import unpack_s2kl
    from ccsd_gpu_amd.hip import compute_int2e_gpu, compute_int2e_ip1_gpu
    mol = systems.water(seed=0)
    n = mol.nao
    assert n == 24
    eri = unpack_s2kl(compute_int2e_gpu(mol, 0, mol.nbas)).cpu().numpy()
    ref = oi.int2e_full(mol)
    assert np.abs(eri - ref).max() < 1e-10
    ip1 = compute_int2e_ip1_gpu(mol, 0, mol.nbas)
    k, l = np.tril_indices(n)
    full = np.zeros((3, n, n, n, n))
    p = ip1.cpu().numpy()
    full[..., k, l] = p
    full[..., l, k] = p
    rip = oi.ip1_full(mol)
    err = np.abs(full - rip).max()
    print("cfg0 int2e / ip1 s1 max err", np.abs(eri - ref).max(), err, flush=True)
    assert err < 1e-10


@pytest.mark.gpu
def test_single_water_gradient_matches_oracle_fixture(gpu):
    mol, e_hf, e_corr, ref = _golden("water_s0")
    mf, mycc, de, info = _pipeline_gradient(mol, "plain")
    print("water E_hf", mf.e_tot - e_hf, "E_corr", mycc.e_corr - e_corr, "grad", np.abs(de - ref).max(),
          flush=True)
    assert abs(mf.e_tot - e_hf) < 1e-8 and abs(mycc.e_corr - e_corr) < 1e-8
    assert np.abs(de - ref).max() < 1e-6


# ----------------------------------------------------------------------------------
# (H2O)4: ERI autograd and the whole gradient
# ----------------------------------------------------------------------------------
@pytest.mark.gpu
def test_water_tetramer_eri_autograd_vs_fd_and_oracle(gpu):
    from ccsd_gpu_amd.autograd import ERI, symmetrize8
    from ccsd_gpu_amd.hip.grad2e import contract_int2e_ip1, pack_s2kl, sph_to_cart_4index
    mol = systems.water_cluster(4, seed=1)
    n = mol.nao
    assert n == 96
    gen = torch.Generator(device="cuda").manual_seed(41)
    G = symmetrize8(torch.randn((n, n, n, n), dtype=torch.float64, device="cuda", generator=gen)) * np.sqrt(8.0)
    coords = torch.tensor(mol.atom_coords(), dtype=torch.float64, device="cuda", requires_grad=True)
    (ERI.apply(coords, mol) * G).sum().backward()
    grad = coords.grad.cpu().numpy()
    h = 1e-4
    base = mol.atom_coords()
    for a, x in [(0, 0), (1, 2), (4, 1), (8, 0), (11, 2)]:
        vals = []
        for s in (+1, -1):
            c = base.copy()
            c[a, x] += s * h
            with torch.no_grad():
                vals.append(float((ERI.apply(torch.tensor(c, device="cuda"), mol) * G).sum()))
        fd = (vals[0] - vals[1]) / (2 * h)
        print("(H2O)4 autograd", (a, x), grad[a, x], "FD", fd, flush=True)
        # FD noise: 85M-term sums at h = 1e-4; tolerance relative to the largest component
        assert abs(grad[a, x] - fd) < 1e-6 * max(1.0, np.abs(grad).max())
    assert abs(grad.sum(0)).max() < 1e-9 * max(1.0, np.abs(grad).max())
    # oracle: -8 sum_{i in shells of the first water, j, k>=l} ip1 G' (Cartesian s2kl, diagonal halved)
    c2s = torch.as_tensor(mol.cart2sph_coeff(), device="cuda")
    gp = pack_s2kl(sph_to_cart_4index(G, c2s), halve_diag=True)
    b1 = int(np.searchsorted(mol._bas[:, 0], 3))
    part = contract_int2e_ip1(mol, gp, scale=-8.0, b0=0, b1=b1).cpu().numpy()
    mc = mol.copy()
    mc.cart = True
    ipo = oi.int2e_s2kl_cart(mc, 0, b1, 1)
    loc = mc.ao_loc_nr(cart=True)
    per = np.einsum("xijp,ijp->xi", ipo, gp[:loc[b1]].cpu().numpy())
    ref = np.zeros((mol.natm, 3))
    np.add.at(ref, np.repeat(mc._bas[:b1, 0], np.diff(loc[:b1 + 1])), -8.0 * per.T)
    err = np.abs(part - ref).max() / np.abs(ref).max()
    print("(H2O)4 backward vs oracle ip1 (first water's shells) rel", err, flush=True)
    assert err < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("mode,prec", [("mixed", "bf16x3"), ("mixed", "fp32"), ("plain", "fp64")])
def test_water_tetramer_gradient_matches_oracle_fixture(gpu, mode, prec):
    """(H2O)4 (o 20, v 76): the products above the mixed-precision flop threshold run in
    split-BF16 / FP32 here, refined to the FP64 fixed point."""
    from ccsd_gpu_amd import precision as P
    mol, e_hf, e_corr, ref = _golden("h2o4_s1")
    n0 = P.LOWP_PRODUCTS.get(prec, 0)
    mf, mycc, de, info = _pipeline_gradient(mol, mode, prec)
    err = np.abs(de - ref).max()
    print("(H2O)4", mode, prec, "E_hf", mf.e_tot - e_hf, "E_corr", mycc.e_corr - e_corr, "grad err", err,
          flush=True)
    if mode == "mixed":
        assert P.LOWP_PRODUCTS[prec] - n0 > 10
    assert abs(mf.e_tot - e_hf) < 1e-8 and abs(mycc.e_corr - e_corr) < 1e-8
    assert err < 1e-6
    assert info["lambda_converged"]


# ----------------------------------------------------------------------------------
# 64 C/N/O/H molecules in one screened pass
# ----------------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_sixty_four_molecule_batch(gpu):
    from ccsd_gpu_amd.batch import MoleculeBatch
    mols = systems.cnoh_batch(64, seed=0)
    rng = np.random.default_rng(0)
    dms = []
    for m in mols:
        a = rng.standard_normal((m.nao_nr(), m.nao_nr()))
        dms.append((a + a.T) / np.sqrt(m.nao_nr()))
    b = MoleculeBatch(mols)
    full = torch.cat(b.grad2e_separable(dms)).clone()
    nfull = b.last_quartets()
    b.screen(1e-12)
    de = b.grad2e_separable(dms)
    scr = torch.cat(de).clone()
    nscr = b.last_quartets()
    diff = float((scr - full).abs().max())
    print("64 molecules: quartets", nfull, "->", nscr, "screened vs unscreened", diff, flush=True)
    assert nscr < nfull and diff < 1e-8
    for i, d in enumerate(de):
        assert float(d.sum(0).abs().max()) < 1e-9 * max(1.0, float(d.abs().max())), i
    # the two smallest molecules against the oracle's ip1 x Gamma_sep(D, D)
    order = np.argsort([m.nao_nr() for m in mols])[:2]
    for i in order:
        m, D = mols[i], dms[i]
        mc = m.copy()
        mc.cart = True
        c = m.cart2sph_coeff()
        Dc = c @ D @ c.T
        g = np.einsum("ij,kl->ijkl", Dc, Dc) * 2
        g -= 0.5 * (np.einsum("il,kj->ijkl", Dc, Dc) + np.einsum("ik,lj->ijkl", Dc, Dc))
        t = np.einsum("xijkl,ijkl->xi", oi.ip1_full(mc), g)
        ref = np.zeros((m.natm, 3))
        np.add.at(ref, np.repeat(mc._bas[:, 0], np.diff(mc.ao_loc_nr(cart=True))), -t.T)
        got = full[b.atom0[i]:b.atom0[i + 1]].cpu().numpy()
        err = np.abs(got - ref).max()
        print("molecule", i, "nao", m.nao_nr(), "vs oracle", err, flush=True)
        assert err < 1e-10 * max(1.0, np.abs(ref).max())


# ----------------------------------------------------------------------------------
# (H2O)6 precision study
# ----------------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_water_hexamer_mixed_precision_lambda(gpu):
    from ccsd_gpu_amd import pipeline
    from ccsd_gpu_amd.gpu_ccsd_grad import _gamma1_intermediates_gpu, _gamma2_intermediates_gpu, grad_elec_gpu
    from ccsd_gpu_amd.gpu_ccsd_lambda import solve_lambda_gpu
    mol = systems.water_cluster(6, seed=1)
    mf, mycc = pipeline.setup(mol)
    assert (mycc.nocc, mol.nao - mycc.nocc) == (30, 114)
    t1, t2 = mycc.t1, mycc.t2
    mycc.conv_tol_normt = 1e-8

    def rdms(l1, l2, prec):
        d = list(_gamma1_intermediates_gpu(mycc, t1, t2, l1, l2, prec))
        d += list(_gamma2_intermediates_gpu(mycc, t1, t2, l1, l2, prec))
        return [x for x in d if x is not None]

    r1, r2 = solve_lambda_gpu(mycc, t1, t2, precision="fp64", return_tensors=True)
    r1, r2 = r1.clone(), r2.clone()
    assert mycc.converged_lambda
    d_ref = rdms(r1, r2, "fp64")
    g_ref = grad_elec_gpu(mycc, t1, t2, r1, r2)
    for prec in ("bf16x3", "fp32"):
        l1, l2 = solve_lambda_gpu(mycc, t1, t2, precision=prec, return_tensors=True)
        assert mycc.converged_lambda and mycc._lambda_precision == prec
        e1, e2 = float((l1 - r1).abs().max()), float((l2 - r2).abs().max())
        d = rdms(l1, l2, prec)
        erdm = max(float((a - b).abs().max()) for a, b in zip(d, d_ref))
        g = grad_elec_gpu(mycc, t1, t2, l1, l2, precision=prec)
        eg = float(np.abs(g - g_ref).max())
        print("(H2O)6", prec, "l1", e1, "l2", e2, "rdm", erdm, "grad", eg, "fp64 updates",
              mycc._lambda_fp64_updates, flush=True)
        assert e1 < 1e-7 and e2 < 1e-7
        assert erdm < 1e-6
        assert eg < 1e-6


# ----------------------------------------------------------------------------------
# (H2O)8 full combined path
# ----------------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_water_octamer_mixed_vs_plain_and_oracle(gpu):
    from ccsd_gpu_amd import pipeline
    try:
        mol, e_hf, e_corr, ref = _golden("h2o8_s1")
    except pytest.skip.Exception:
        mol, ref = systems.water_cluster(8, seed=1), None
    assert mol.nao == 192
    mf, mycc = pipeline.setup(mol)
    assert (mycc.nocc, mol.nao - mycc.nocc) == (40, 152)
    dm, im = pipeline.ccsd_gradient(mf, mycc, "mixed", "bf16x3")
    dp, ip = pipeline.ccsd_gradient(mf, mycc, "plain")
    diff = float(np.abs(dm - dp).max())
    net = float(np.abs(dm.sum(0)).max())
    torque = float(np.abs(np.cross(mol.atom_coords(), dm).sum(0)).max())
    print("(H2O)8 mixed vs plain", diff, "net force", net, "torque", torque, "mixed", im["total_s"],
          "plain", ip["total_s"], flush=True)
    assert im["lambda_converged"] and ip["lambda_converged"]
    assert diff < 1e-6 and net < 1e-7 and torque < 1e-6
    if ref is not None:
        assert abs(mf.e_tot - e_hf) < 1e-8 and abs(mycc.e_corr - e_corr) < 1e-8
        err = float(np.abs(dm - ref).max())
        print("(H2O)8 mixed vs oracle fixture", err, flush=True)
        assert err < 1e-6


# ----------------------------------------------------------------------------------
# the reference's gradient set
# ----------------------------------------------------------------------------------
REF_SET = {
    "HF/STO-3G": lambda: molecules.hf_sto3g(),
    "CH4/cc-pVDZ": lambda: molecules.ch4_ccpvdz(),
    "H2O/cc-pVTZ": lambda: M("O 0 0 0; H 0 -0.757 0.587; H 0 0.757 0.587", basis="cc-pvtz"),
}


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(REF_SET))
def test_reference_gradient_set(gpu, name):
    """ccsd_gpu tests/test_gpu_ccsd_grad.py:142-156: forces within 1e-6 Eh/Bohr of the
    CPU gradient; here RHF -> CCSD -> lambda -> compute_ml_ccsd_forces_gpu on the GPU
    against the oracle's pipeline (H2O/cc-pVTZ puts f shells through the gradient)."""
    mol = REF_SET[name]()
    _, _, ref = _oracle_gradient(mol)
    mf, mycc, de, info = _pipeline_gradient(mol, "plain")
    err = float(np.abs(de - ref).max())
    _, _, dm, _ = _pipeline_gradient(mol, "mixed")
    errm = float(np.abs(dm - ref).max())
    print(name, "nao", mol.nao, "plain err", err, "mixed err", errm, flush=True)
    assert err < 1e-6 and errm < 1e-6


# ----------------------------------------------------------------------------------
# frozen orbitals
# ----------------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("frozen", [1, [0, 23]])
@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_frozen_orbital_gradient_matches_oracle(gpu, frozen, precision):
    from ccsd_gpu_amd import cc, scf
    from ccsd_gpu_amd.gpu_ccsd_grad import grad_elec_gpu
    from ccsd_gpu_amd.gpu_ccsd_lambda import solve_lambda_gpu
    mol = M("O 0.0 -0.0757 0.0521; H 0.7572 0.5 -0.02; H -0.76 0.47 0.03", basis="cc-pvdz")
    _, occ, ref = _oracle_gradient(mol, frozen=frozen)
    mf = scf.RHF(mol)
    mf.kernel()
    mycc = cc.CCSD(mf, frozen=frozen)
    mycc.kernel()
    assert abs(mycc.e_corr - occ.e_corr) < 1e-8
    l1, l2 = solve_lambda_gpu(mycc, precision=precision, return_tensors=True)
    de = grad_elec_gpu(mycc, mycc.t1, mycc.t2, l1, l2, precision=precision) + scf.grad_nuc(mol)
    err = float(np.abs(de - ref).max())
    print("frozen", frozen, precision, "grad err", err, "zvector converged", mycc._zvector_converged, flush=True)
    assert err < 1e-6 and mycc._zvector_converged


# ----------------------------------------------------------------------------------
# PySCF-like callers: plain-attribute stand-ins with NumPy arrays
# ----------------------------------------------------------------------------------
class _RHFLike:
    def __init__(self, mol, mo_coeff, mo_energy, mo_occ, e_tot):
        self.mol, self.mo_coeff, self.mo_energy, self.mo_occ, self.e_tot = mol, mo_coeff, mo_energy, mo_occ, e_tot


class _CCSDLike:
    def __init__(self, mf, t1, t2):
        self._scf, self.mol = mf, mf.mol
        self.mo_coeff, self.mo_occ = mf.mo_coeff, mf.mo_occ
        self.t1, self.t2 = t1, t2
        self.max_cycle, self.conv_tol_normt = 50, 1e-10
        self.diis_space, self.level_shift, self.verbose = 6, 0.0, 0

    def get_frozen_mask(self):
        return np.ones(self.mo_coeff.shape[1], dtype=bool)


class _ERIsLike:
    """PySCF ccsd._ChemistsERIs layout: ovvv packed over its last pair (get_ovvv
    unpacks), vvvv 4-fold packed (npair, npair)."""

    def __init__(self, oe):
        nv = oe.ovvv.shape[1]
        a, b = np.tril_indices(nv)
        self.fock, self.mo_energy = oe.fock, oe.mo_energy
        self.oooo, self.ovoo, self.ovvo, self.oovv, self.ovov = oe.oooo, oe.ovoo, oe.ovvo, oe.oovv, oe.ovov
        self.ovvv = oe.ovvv[:, :, a, b]
        self.vvvv = oe.vvvv[a, b][:, a, b]
        self._nv = nv

    def get_ovvv(self, *slices):
        a, b = np.tril_indices(self._nv)
        out = np.empty(self.ovvv.shape[:2] + (self._nv, self._nv))
        out[..., a, b] = self.ovvv
        out[..., b, a] = self.ovvv
        return out[slices] if slices else out


@pytest.mark.gpu
@pytest.mark.parametrize("with_eris", [False, True])
def test_pyscf_like_standins_drive_lambda_and_forces(gpu, with_eris):
    """solve_lambda_gpu and compute_ml_ccsd_forces_gpu with objects exposing only the
    attributes the reference reads (NumPy arrays, no private methods of this build):
    lambda returned as NumPy and equal to the oracle's, forces equal to the oracle's."""
    from ccsd_gpu_amd import compute_ml_ccsd_forces_gpu, solve_lambda_gpu
    mol = M("O 0.0 -0.0757 0.0521; H 0.7572 0.5 -0.02; H -0.76 0.47 0.03", basis="cc-pvdz")
    omf, occ, ref = _oracle_gradient(mol)
    conv, ol1, ol2 = chem.solve_lambda(occ.t1, occ.t2, occ.ao2mo())
    mf = _RHFLike(mol, omf.mo_coeff.copy(), omf.mo_energy.copy(), omf.mo_occ.copy(), omf.e_tot)
    mycc = _CCSDLike(mf, occ.t1.copy(), occ.t2.copy())
    eris = _ERIsLike(occ.ao2mo()) if with_eris else None
    l1, l2 = solve_lambda_gpu(mycc, eris=eris)
    assert isinstance(l1, np.ndarray) and isinstance(l2, np.ndarray)
    assert mycc.converged_lambda and mycc._lambda_solver_mode == "gpu-hybrid"
    e1, e2 = np.abs(l1 - ol1).max(), np.abs(l2 - ol2).max()
    de = compute_ml_ccsd_forces_gpu(mf, mycc.t1, mycc.t2, l1, l2)
    err = np.abs(de - ref).max()
    print("stand-ins, eris given" if with_eris else "stand-ins", "l1", e1, "l2", e2, "forces", err, flush=True)
    assert e1 < 1e-8 and e2 < 1e-8
    assert isinstance(de, np.ndarray) and err < 1e-8


# ----------------------------------------------------------------------------------
# failure behaviour
# ----------------------------------------------------------------------------------
@pytest.mark.gpu
def test_lambda_max_cycle_exhausted_reports_not_converged(gpu, monkeypatch):
    from ccsd_gpu_amd import cc, scf
    from ccsd_gpu_amd.gpu_ccsd_lambda import solve_lambda_gpu
    mol = molecules.h2o_ccpvdz()
    mf = scf.RHF(mol)
    mf.kernel()
    mycc = cc.CCSD(mf)
    mycc.kernel()
    mycc.max_cycle = 1
    l1, l2 = solve_lambda_gpu(mycc)
    assert mycc.converged_lambda is False
    assert isinstance(l1, np.ndarray) and l1.shape == mycc.t1.shape


@pytest.mark.gpu
def test_zvector_nonconvergence_is_flagged(gpu):
    from ccsd_gpu_amd import cc, scf
    from ccsd_gpu_amd.gpu_ccsd_grad import _response_dm1_gpu
    mol = molecules.h2o_ccpvdz()
    mf = scf.RHF(mol)
    mf.kernel()
    mycc = cc.CCSD(mf)
    nocc = mycc.nocc
    nvir = mol.nao - nocc
    x = torch.randn((nvir, nocc), dtype=torch.float64, device="cuda",
                    generator=torch.Generator(device="cuda").manual_seed(3))
    with pytest.warns(RuntimeWarning, match="Z-vector"):
        _response_dm1_gpu(mycc, x, mf._ints(), max_cycle=1)
    assert mycc._zvector_converged is False and mycc._zvector_iterations == 1
    _response_dm1_gpu(mycc, x, mf._ints())
    assert mycc._zvector_converged is True


# ----------------------------------------------------------------------------------
# run twice, bitwise
# ----------------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["plain", "mixed"])
def test_gradient_run_twice_is_bitwise_identical(gpu, mode):
    """The whole pipeline twice from scratch (integrals, RHF, CCSD, lambda, dE/dR) and
    the gradient once more on the second setup give the same bits under
    deterministic(): no step sums in an order that depends on scheduling (the fused
    pass and the one-electron kernels accumulate without float atomics; torch's
    library calls take their deterministic forms)."""
    import warnings

    from ccsd_gpu_amd import pipeline
    from ccsd_gpu_amd.determinism import deterministic
    mol = systems.water_cluster(2, seed=3, basis="cc-pvdz")
    with deterministic(), warnings.catch_warnings(record=True) as caught:
        warnings.simplefilter("always")
        mf1, cc1, de1, _ = _pipeline_gradient(mol, mode)
        mf2, cc2, de2, _ = _pipeline_gradient(mol, mode)
        de3, _ = pipeline.ccsd_gradient(mf2, cc2, mode, "bf16x3")
    flagged = sorted({str(w.message)[:120] for w in caught if "deterministic" in str(w.message)})
    print("torch operations without a deterministic form:", flagged, flush=True)
    assert not flagged
    print(mode, "E", mf1.e_tot, mf2.e_tot, "max |d1 - d2|", np.abs(de1 - de2).max(), flush=True)
    assert mf1.e_tot == mf2.e_tot and cc1.e_corr == cc2.e_corr
    assert np.array_equal(de1, de2) and np.array_equal(de2, de3)
