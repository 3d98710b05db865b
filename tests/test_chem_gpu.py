"""GPU CCSD-gradient path against the FP64 oracle (H2O and friends).

Mirrors the reference's tests/test_gpu_ccsd_grad.py: lambda against the CPU
solution (there 1e-7 on l1/l2), the gradient against the CPU gradient on the
same amplitudes (there 1e-6 Eh/Bohr). Orbitals, amplitudes and lambdas are
shared between the two sides so that elementwise comparisons are meaningful.
"""
import numpy as np
import pytest
import torch

from oracle import chem
from oracle import integrals as oi

from . import molecules

GEOM = "O 0.0 -0.0757 0.0521; H 0.7572 0.5 -0.02; H -0.76 0.47 0.03"


@pytest.mark.gpu
@pytest.mark.parametrize("make,cart", [(molecules.h2o_ccpvdz, False), (molecules.h2o_ccpvdz, True),
                                       (molecules.g_shell_mol, False)])
def test_int1e_gpu_matches_oracle(gpu, make, cart):
    from ccsd_gpu_amd.hip.int1e_gpu import int1e
    mol = make(cart=cart)
    for kind in ("ovlp", "kin", "nuc", "ipovlp", "ipkin", "ipnuc_atom"):
        g = int1e(mol, kind).cpu().numpy()
        r = oi.int1e(mol, kind)
        err = np.abs(g - r).max()
        print(kind, err, flush=True)
        assert err < 1e-10, (kind, err)


@pytest.mark.gpu
@pytest.mark.parametrize("mode,tol", [("fp32", 2e-5), ("bf16x3", 2e-4)])
def test_mixed_gemm_matches_fp64(gpu, mode, tol):
    from ccsd_gpu_amd import mfma
    gen = torch.Generator(device="cuda").manual_seed(5)
    for (nb, m, k, n) in [(1, 64, 32, 64), (3, 100, 77, 130), (2, 1, 300, 7), (1, 257, 1000, 65)]:
        A = torch.randn((nb, m, k), dtype=torch.float64, device="cuda", generator=gen)
        B = torch.randn((nb, k, n), dtype=torch.float64, device="cuda", generator=gen)
        C = mfma.bmm(A, B, mode)
        ref = torch.bmm(A, B)
        scale = torch.sqrt(torch.tensor(float(k)))
        err = float((C - ref).abs().max() / scale)
        print(mode, (nb, m, k, n), err, flush=True)
        assert err < tol


@pytest.mark.gpu
def test_bf16x3_nt_gemm_matches_fp64(gpu):
    """Split-BF16 NT GEMM (pre-split planes) against FP64, ragged shapes, batch, beta."""
    from ccsd_gpu_amd import mfma
    gen = torch.Generator(device="cuda").manual_seed(7)
    for (nb, m, k, n) in [(1, 128, 32, 128), (2, 100, 77, 130), (1, 1, 300, 7), (3, 257, 1000, 65),
                          (1, 300, 64, 1000), (2, 256, 640, 300)]:
        A = torch.randn((nb, m, k), dtype=torch.float64, device="cuda", generator=gen)
        B = torch.randn((nb, n, k), dtype=torch.float64, device="cuda", generator=gen)
        ref = A @ B.transpose(1, 2)
        C = mfma.gemm_nt(mfma.split(A), mfma.split(B))
        err = float((C - ref).abs().max()) / k ** 0.5
        print("bf16x3 nt", (nb, m, k, n), err, flush=True)
        assert err < 2e-4
        C2 = mfma.gemm_nt(mfma.split(A), mfma.split(B), out=C.clone(), beta=1)
        assert float((C2 - 2 * ref).abs().max()) / k ** 0.5 < 4e-4


@pytest.mark.gpu
@pytest.mark.parametrize("nb,m,k,n,ksplit", [(1, 40, 70000, 40, None), (2, 50, 5000, 30, 3), (1, 7, 4099, 130, 5),
                                             (1, 152, 30000, 152, None)])
def test_bf16x3_nt_split_k(gpu, nb, m, k, n, ksplit):
    """K split into chunks (one batched launch + ragged last chunk, summed) against FP64."""
    from ccsd_gpu_amd import mfma
    gen = torch.Generator(device="cuda").manual_seed(k)
    A = torch.randn((nb, m, k), dtype=torch.float64, device="cuda", generator=gen)
    B = torch.randn((nb, n, k), dtype=torch.float64, device="cuda", generator=gen)
    ref = A @ B.transpose(1, 2)
    As, Bs = mfma.split(A), mfma.split(B)
    C = mfma.gemm_nt(As, Bs, ksplit=ksplit)
    err = float((C - ref).abs().max()) / k ** 0.5
    print("split-k", (nb, m, k, n, ksplit), err, flush=True)
    assert err < 2e-4
    C2 = mfma.gemm_nt(As, Bs, out=C.clone(), beta=1, ksplit=ksplit)
    assert float((C2 - 2 * ref).abs().max()) / k ** 0.5 < 4e-4


@pytest.fixture
def gemm_variant():
    from ccsd_gpu_amd import _lib
    lib = _lib.load()
    yield lambda v: _lib.check(lib.ccsd_gemm_set_variant(v), "ccsd_gemm_set_variant")
    lib.ccsd_gemm_set_variant(0)


@pytest.mark.gpu
@pytest.mark.parametrize("variant", [1, 2])
@pytest.mark.parametrize("nb,m,n,k", [(1, 1024, 1024, 1024), (1, 820, 2400, 4100), (2, 513, 2047, 1112),
                                      (1, 130, 1030, 264)])
def test_bf16x3_nt_block_variants_match_fp64_and_default(gpu, gemm_variant, variant, nb, m, n, k):
    """Long-K block shapes and LDS staging (128 x 128 or 128 x 256 blocks, one or two
    LDS stages) against FP64 and the default kernel: ragged M, N and K (K not a
    multiple of the stage depth), batches, beta = 1."""
    from ccsd_gpu_amd import mfma
    gen = torch.Generator(device="cuda").manual_seed(m + n)
    A = torch.randn((nb, m, k), dtype=torch.float64, device="cuda", generator=gen)
    B = torch.randn((nb, n, k), dtype=torch.float64, device="cuda", generator=gen)
    As, Bs = mfma.split(A), mfma.split(B)
    ref = A @ B.transpose(1, 2)
    gemm_variant(variant)
    got = mfma.gemm_nt(As, Bs, ksplit=1)
    got2 = mfma.gemm_nt(As, Bs, out=got.clone(), beta=1, ksplit=1)
    gemm_variant(0)
    base = mfma.gemm_nt(As, Bs, ksplit=1)
    err = float((got - ref).abs().max()) / k ** 0.5
    err2 = float((got2 - 2 * ref).abs().max()) / k ** 0.5
    d = float((got - base).abs().max()) / k ** 0.5
    print("variant", variant, (nb, m, n, k), "vs fp64", err, "beta=1", err2, "vs default", d, flush=True)
    # same MFMA sequence per output element (K steps of 16 in order, lo.hi, hi.lo, hi.hi)
    assert err < 2e-4 and err2 < 4e-4 and d == 0.0


@pytest.mark.gpu
@pytest.mark.parametrize("variant", [0, 1, 2])
def test_bf16x3_nt_scatter_destination_variants(gpu, gemm_variant, variant):
    """C[off_b + off_r[row] + off_c[col]] += alpha A B^T into a transposed destination."""
    from ccsd_gpu_amd import mfma
    m, n, k = 600, 1900, 1304
    gen = torch.Generator(device="cuda").manual_seed(11)
    A = torch.randn((1, m, k), dtype=torch.float64, device="cuda", generator=gen)
    B = torch.randn((1, n, k), dtype=torch.float64, device="cuda", generator=gen)
    ref = (A @ B.transpose(1, 2))[0]
    out = torch.randn((n, m), dtype=torch.float64, device="cuda", generator=gen)
    want = out + 0.5 * ref.T
    zero = torch.zeros(1, dtype=torch.int64, device="cuda")
    rows = torch.arange(m, dtype=torch.int64, device="cuda")            # row r -> column r of out
    cols = torch.arange(n, dtype=torch.int64, device="cuda") * m        # col c -> row c of out
    gemm_variant(variant)
    mfma.gemm_nt_scatter(mfma.split(A), mfma.split(B), out, zero, rows, cols, alpha=0.5, beta=1, ksplit=1)
    err = float((out - want).abs().max()) / k ** 0.5
    print("scatter variant", variant, err, flush=True)
    assert err < 2e-4


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol,o,v", [("fp64", 1e-13, 6, 11), ("fp32", 1e-5, 6, 11),
                                               ("bf16x3", 1e-4, 6, 11), ("fp64", 1e-13, 5, 45),
                                               ("fp64", 1e-13, 3, 70), ("bf16x3", 1e-4, 4, 64)])
def test_vvvv_term_pair_halves_match_full(gpu, precision, tol, o, v):
    """The l2.vvvv term through the symmetric / antisymmetric pair products against
    the full einsum, twice (the second call reads the cached pair matrices); v past
    one and two 32-wide unpack tiles, ragged and exact."""
    from ccsd_gpu_amd.precision import vvvv_term
    gen = torch.Generator(device="cuda").manual_seed(3 + o * v)
    x = torch.randn((v, v, v, v), dtype=torch.float64, device="cuda", generator=gen)
    vvvv = x + x.permute(1, 0, 2, 3)
    vvvv = vvvv + vvvv.permute(0, 1, 3, 2)
    vvvv = vvvv + vvvv.permute(2, 3, 0, 1)             # 8-fold symmetric (ac|bd)
    l = torch.randn((o, o, v, v), dtype=torch.float64, device="cuda", generator=gen)
    l2 = l + l.permute(1, 0, 3, 2)
    ref = torch.einsum("ijcd,acbd->ijab", l2, vvvv)
    for _ in range(2):
        got = vvvv_term(l2, vvvv, precision, scale=0.5)
        err = float((got - 0.5 * ref).abs().max() / ref.abs().max())
        print(precision, err, flush=True)
        assert err < tol


@pytest.mark.gpu
def test_contract_mixed_matches_einsum(gpu, monkeypatch):
    from ccsd_gpu_amd import precision
    from ccsd_gpu_amd.precision import contract
    monkeypatch.setattr(precision, "MIXED_MIN_FLOPS", 0.0)     # every product at low precision
    n0 = dict(precision.LOWP_PRODUCTS)
    gen = torch.Generator(device="cuda").manual_seed(9)
    o, v = 5, 9
    t2 = torch.randn((o, o, v, v), dtype=torch.float64, device="cuda", generator=gen)
    l2 = torch.randn((o, o, v, v), dtype=torch.float64, device="cuda", generator=gen)
    vvvv = torch.randn((v, v, v, v), dtype=torch.float64, device="cuda", generator=gen)
    t1 = torch.randn((o, v), dtype=torch.float64, device="cuda", generator=gen)
    for eq, ops in [("ijcd,acbd->ijab", (l2, vvvv)), ("jkca,bikc->jiba", (l2, t2.permute(2, 0, 1, 3))),
                    ("kc,kb,jc->bj", (t1, t1, t1)), ("dljd->jl", (t2.permute(2, 0, 1, 3),)),
                    ("ijab,ijcd->abcd", (l2, t2))]:
        ref = torch.einsum(eq, *ops)
        for prec, tol in (("fp32", 1e-5), ("bf16x3", 2e-4)):
            got = contract(eq, *ops, precision=prec)
            err = float((got - ref).abs().max() / (ref.abs().max() + 1e-30))
            assert err < tol, (eq, prec, err)
    assert all(precision.LOWP_PRODUCTS[p] > n0[p] for p in ("fp32", "bf16x3"))


@pytest.fixture(scope="module")
def h2o_pair():
    """Oracle RHF/CCSD/lambda and a GPU RHF sharing the oracle's orbitals."""
    from ccsd_gpu_amd.gto import M
    from ccsd_gpu_amd import scf
    mol = M(GEOM, basis="cc-pvdz")
    omf = chem.RHF(mol)
    omf.kernel()
    occ = chem.CCSD(omf)
    occ.kernel()
    conv, l1, l2 = chem.solve_lambda(occ.t1, occ.t2, occ.ao2mo())
    gmf = scf.RHF(mol)
    gmf.kernel()
    return mol, omf, occ, (l1, l2), gmf


@pytest.mark.gpu
def test_rhf_and_ccsd_energies_match_oracle(gpu, h2o_pair):
    from ccsd_gpu_amd import cc
    mol, omf, occ, _, gmf = h2o_pair
    print("E_rhf", gmf.e_tot, omf.e_tot, flush=True)
    assert abs(gmf.e_tot - omf.e_tot) < 1e-9
    mycc = cc.CCSD(gmf)
    mycc.kernel()
    print("E_ccsd", mycc.e_corr, occ.e_corr, flush=True)
    assert abs(mycc.e_corr - occ.e_corr) < 1e-8


def _gpu_cc_with_oracle_orbitals(gmf, omf, occ):
    from ccsd_gpu_amd import cc
    gmf.mo_coeff = omf.mo_coeff.copy()
    gmf.mo_energy = omf.mo_energy.copy()
    gmf.mo_occ = omf.mo_occ.copy()
    mycc = cc.CCSD(gmf)
    mycc.t1 = torch.as_tensor(occ.t1, device="cuda")
    mycc.t2 = torch.as_tensor(occ.t2, device="cuda")
    return mycc


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3"])
def test_lambda_gpu_matches_oracle(gpu, h2o_pair, precision, monkeypatch):
    """Mixed precision converges, through FP64-residual refinement, to the FP64 lambda;
    every contraction of the low-precision sweeps runs at that precision (the flop
    threshold that keeps small products in FP64 is set to 0)."""
    from ccsd_gpu_amd import precision as P
    from ccsd_gpu_amd.gpu_ccsd_lambda import solve_lambda_gpu
    monkeypatch.setattr(P, "MIXED_MIN_FLOPS", 0.0)
    n0 = dict(P.LOWP_PRODUCTS)
    mol, omf, occ, (l1, l2), gmf = h2o_pair
    mycc = _gpu_cc_with_oracle_orbitals(gmf, omf, occ)
    mycc.conv_tol_normt = 1e-10
    g1, g2 = solve_lambda_gpu(mycc, precision=precision, return_tensors=True)
    assert mycc.converged_lambda
    e1 = float(np.abs(g1.cpu().numpy() - l1).max())
    e2 = float(np.abs(g2.cpu().numpy() - l2).max())
    print(precision, "l1", e1, "l2", e2, "iters", mycc._lambda_iterations, "fp64 updates",
          mycc._lambda_fp64_updates, flush=True)
    assert e1 < 1e-8 and e2 < 1e-8
    if precision != "fp64":
        # (a graph-captured update issues its products once, at capture)
        assert P.LOWP_PRODUCTS[precision] - n0[precision] > 20


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["bf16x3", "fp32"])
def test_lambda_residual_update_accepts_only_true_convergence(gpu, h2o_pair, precision, monkeypatch):
    """Convergence accepted on the updated residual r + A d - d (A d in low precision,
    no FP64 recompute): the FP64 residual of the answer, measured as the first change
    norm of an FP64 solve started from it, is below the tolerance and agrees with the
    estimate, and the answer is the oracle's lambda."""
    from ccsd_gpu_amd import precision as P
    from ccsd_gpu_amd.gpu_ccsd_lambda import solve_lambda_gpu
    monkeypatch.setattr(P, "MIXED_MIN_FLOPS", 0.0)     # real low-precision noise in A d
    mol, omf, occ, (l1, l2), gmf = h2o_pair
    mycc = _gpu_cc_with_oracle_orbitals(gmf, omf, occ)
    tol = 1e-8
    mycc.conv_tol_normt = tol
    mycc.lambda_corr_rtol = 1e-4      # a tight correction, so that the estimate is what accepts
    mycc.lambda_corr_max_cycle = 30
    mycc.lambda_diis_residual = False
    g1, g2 = solve_lambda_gpu(mycc, precision=precision, return_tensors=True)
    hist = list(mycc._lambda_history)
    assert mycc.converged_lambda
    assert hist[-1][0] == "residual-update"
    assert mycc._lambda_fp64_updates == 1
    est = hist[-1][2]
    solve_lambda_gpu(mycc, l1=g1.clone(), l2=g2.clone(), precision="fp64")
    true_res = mycc._lambda_history[0][2]
    print(precision, "estimated residual", est, "FP64 residual", true_res, flush=True)
    assert true_res < tol
    assert abs(true_res - est) < 1e-9
    assert float(np.abs(g1.cpu().numpy() - l1).max()) < 1e-7
    assert float(np.abs(g2.cpu().numpy() - l2).max()) < 1e-7


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["bf16x3", "fp32"])
def test_lambda_diis_residual_accepts_only_true_convergence(gpu, h2o_pair, precision, monkeypatch):
    """Convergence accepted on the correction DIIS's own residual (the input combination
    of its sweeps, whose residual is sum_s c_s E[s]): the FP64 residual of the answer,
    measured as the first change norm of an FP64 solve started from it, is below the
    tolerance and agrees with that norm, and the answer is the oracle's lambda."""
    from ccsd_gpu_amd import precision as P
    from ccsd_gpu_amd.gpu_ccsd_lambda import solve_lambda_gpu
    monkeypatch.setattr(P, "MIXED_MIN_FLOPS", 0.0)     # real low-precision noise in A d
    mol, omf, occ, (l1, l2), gmf = h2o_pair
    mycc = _gpu_cc_with_oracle_orbitals(gmf, omf, occ)
    tol = 1e-8
    mycc.conv_tol_normt = tol
    mycc.lambda_corr_rtol = 1e-6      # the inner target out of reach: DIIS's residual accepts
    mycc.lambda_corr_max_cycle = 30
    g1, g2 = solve_lambda_gpu(mycc, precision=precision, return_tensors=True)
    hist = list(mycc._lambda_history)
    assert mycc.converged_lambda
    assert hist[-1][0] == "diis-residual"
    assert mycc._lambda_fp64_updates == 1
    est = hist[-1][2]
    solve_lambda_gpu(mycc, l1=g1.clone(), l2=g2.clone(), precision="fp64")
    true_res = mycc._lambda_history[0][2]
    print(precision, "DIIS residual", est, "FP64 residual", true_res, flush=True)
    assert true_res < tol
    assert abs(true_res - est) < 1e-9
    assert float(np.abs(g1.cpu().numpy() - l1).max()) < 1e-7
    assert float(np.abs(g2.cpu().numpy() - l2).max()) < 1e-7


@pytest.mark.gpu
@pytest.mark.parametrize("graph", [True, False])
def test_lambda_mixed_solve_with_and_without_graph(gpu, h2o_pair, graph):
    """The mixed solve reaches the FP64 lambda with the update graph-captured or eager."""
    from ccsd_gpu_amd.gpu_ccsd_lambda import solve_lambda_gpu
    mol, omf, occ, (l1, l2), gmf = h2o_pair
    mycc = _gpu_cc_with_oracle_orbitals(gmf, omf, occ)
    mycc.conv_tol_normt = 1e-10
    mycc.lambda_graph = graph
    g1, g2 = solve_lambda_gpu(mycc, precision="bf16x3", return_tensors=True)
    assert mycc.converged_lambda
    assert (mycc._lambda_graph_replays > 0) == graph
    assert float(np.abs(g1.cpu().numpy() - l1).max()) < 1e-8
    assert float(np.abs(g2.cpu().numpy() - l2).max()) < 1e-8


@pytest.mark.gpu
@pytest.mark.parametrize("fused", [True, False])
def test_gradient_gpu_matches_oracle(gpu, h2o_pair, fused):
    from ccsd_gpu_amd.gpu_ccsd_grad import grad_elec_gpu
    mol, omf, occ, (l1, l2), gmf = h2o_pair
    ref = chem.grad_elec(occ, occ.t1, occ.t2, l1, l2)
    mycc = _gpu_cc_with_oracle_orbitals(gmf, omf, occ)
    de = grad_elec_gpu(mycc, occ.t1, occ.t2, l1, l2, fused=fused)
    err = np.abs(de - ref).max()
    print("fused" if fused else "unfused", "grad err", err, flush=True)
    assert err < 1e-8


@pytest.mark.gpu
@pytest.mark.parametrize("precision,mo2ao,tol", [("bf16x3", "fp64", 1e-7), ("fp32", "fp64", 1e-7),
                                                  ("bf16x3", "bf16x3", 1e-6), ("fp32", "fp32", 1e-6),
                                                  ("bf16x3", "fp32s", 1e-7), ("bf16x3", None, 1e-7)])
def test_gradient_mixed_precision_matches_oracle(gpu, h2o_pair, precision, mo2ao, tol, monkeypatch):
    """RDM contractions (and optionally the MO->AO transform) on the MFMA path, every
    RDM product at that precision (flop threshold 0)."""
    from ccsd_gpu_amd import precision as P
    from ccsd_gpu_amd.gpu_ccsd_grad import grad_elec_gpu
    monkeypatch.setattr(P, "MIXED_MIN_FLOPS", 0.0)
    n0 = P.LOWP_PRODUCTS[precision]
    mol, omf, occ, (l1, l2), gmf = h2o_pair
    ref = chem.grad_elec(occ, occ.t1, occ.t2, l1, l2)
    mycc = _gpu_cc_with_oracle_orbitals(gmf, omf, occ)
    de = grad_elec_gpu(mycc, occ.t1, occ.t2, l1, l2, precision=precision, mo2ao_precision=mo2ao)
    err = np.abs(de - ref).max()
    print(precision, "mo2ao", mo2ao, "grad err", err, flush=True)
    assert err < tol
    assert P.LOWP_PRODUCTS[precision] - n0 >= 20


@pytest.mark.gpu
def test_forces_entry_point_end_to_end(gpu):
    """RHF -> CCSD -> lambda -> compute_ml_ccsd_forces_gpu, all on the GPU, vs oracle."""
    from ccsd_gpu_amd import cc, scf
    from ccsd_gpu_amd.gpu_ccsd_grad import compute_ml_ccsd_forces_gpu
    from ccsd_gpu_amd.gpu_ccsd_lambda import solve_lambda_gpu
    from ccsd_gpu_amd.gto import M
    mol = M(GEOM, basis="sto-3g")
    mf = scf.RHF(mol)
    mf.kernel()
    mycc = cc.CCSD(mf)
    mycc.kernel()
    l1, l2 = solve_lambda_gpu(mycc, return_tensors=True)
    de = compute_ml_ccsd_forces_gpu(mf, mycc.t1, mycc.t2, l1, l2)
    e0, omf, occ = chem.ccsd_total_energy(mol)
    conv, ol1, ol2 = chem.solve_lambda(occ.t1, occ.t2, occ.ao2mo())
    ref = chem.grad_elec(occ, occ.t1, occ.t2, ol1, ol2) + chem.grad_nuc(mol)
    err = np.abs(de - ref).max()
    print("forces err", err, flush=True)
    assert err < 1e-6
    assert np.abs(de.sum(axis=0)).max() < 1e-8


@pytest.mark.gpu
def test_lambda_update_is_affine_with_homogeneous_part(gpu, h2o_pair):
    """F(l) = A l + b: the homogeneous update equals F(l) - F(0) (used by the mixed-precision refinement)."""
    from ccsd_gpu_amd.gpu_ccsd_lambda import make_intermediates_gpu, update_lambda_gpu
    mol, omf, occ, (l1, l2), gmf = h2o_pair
    mycc = _gpu_cc_with_oracle_orbitals(gmf, omf, occ)
    eris = mycc.ao2mo()
    imds = make_intermediates_gpu(mycc, mycc.t1, mycc.t2, eris)
    L1 = torch.as_tensor(l1, device="cuda") * 1.7
    L2 = torch.as_tensor(l2, device="cuda") * -0.3
    f1, f2 = update_lambda_gpu(mycc, mycc.t1, mycc.t2, L1, L2, eris, imds)
    z1, z2 = update_lambda_gpu(mycc, mycc.t1, mycc.t2, L1 * 0, L2 * 0, eris, imds)
    h1, h2 = update_lambda_gpu(mycc, mycc.t1, mycc.t2, L1, L2, eris, imds, homogeneous=True)
    assert float((f1 - z1 - h1).abs().max()) < 1e-12
    assert float((f2 - z2 - h2).abs().max()) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("lp", ["bf16x3", "fp32"])
def test_fp64_update_with_factored_wvvov_equals_formed(gpu, h2o_pair, lp):
    """The mixed solve's intermediates (wvvov's products formed at low precision only)
    give, through the factored FP64 application, the FP64 update of the fully formed
    FP64 intermediates; the low-precision wvvov is within that precision of the FP64 one
    (the water-tetramer test below covers sizes where it is formed at low precision)."""
    from ccsd_gpu_amd.gpu_ccsd_lambda import make_intermediates_gpu, update_lambda_gpu
    mol, omf, occ, (l1, l2), gmf = h2o_pair
    mycc = _gpu_cc_with_oracle_orbitals(gmf, omf, occ)
    eris = mycc.ao2mo()
    full = make_intermediates_gpu(mycc, mycc.t1, mycc.t2, eris)
    split = make_intermediates_gpu(mycc, mycc.t1, mycc.t2, eris, lp_precision=lp)
    assert full.wvvov_fac is None and split.wvvov_fac is not None
    gen = torch.Generator(device="cuda").manual_seed(5)
    L1 = torch.as_tensor(l1, device="cuda") + 0.01 * torch.randn(l1.shape, dtype=torch.float64, device="cuda",
                                                                  generator=gen)
    L2 = torch.as_tensor(l2, device="cuda") + 0.01 * torch.randn(l2.shape, dtype=torch.float64, device="cuda",
                                                                  generator=gen)
    f1, f2 = update_lambda_gpu(mycc, mycc.t1, mycc.t2, L1, L2, eris, full)
    g1, g2 = update_lambda_gpu(mycc, mycc.t1, mycc.t2, L1, L2, eris, split)
    d1 = float((f1 - g1).abs().max()) / float(f1.abs().max())
    d2 = float((f2 - g2).abs().max()) / float(f2.abs().max())
    dw = float((full.wvvov - split.wvvov).abs().max()) / float(full.wvvov.abs().max())
    print(lp, "l1", d1, "l2", d2, "wvvov", dw, flush=True)
    assert d1 < 1e-12 and d2 < 1e-12
    # (at this size every wvvov product is below the mixed-precision flop threshold and
    # runs in FP64, so dw is 0 here; the (H2O)n solves form them at low precision)
    assert dw < 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["bf16x3", "fp32"])
def test_lambda_graph_replays_match_eager(gpu, h2o_pair, precision):
    """The HIP-graph replay of the lambda update equals the eager update on every
    replay, with eager work and synchronisations between the replays."""
    from ccsd_gpu_amd.gpu_ccsd_lambda import GraphedUpdate, make_intermediates_gpu, update_lambda_gpu
    from ccsd_gpu_amd.precision import static_operands
    mol, omf, occ, (l1, l2), gmf = h2o_pair
    mycc = _gpu_cc_with_oracle_orbitals(gmf, omf, occ)
    eris = mycc.ao2mo()
    imds = make_intermediates_gpu(mycc, mycc.t1, mycc.t2, eris)
    L1 = torch.as_tensor(l1, device="cuda")
    L2 = torch.as_tensor(l2, device="cuda")
    gen = torch.Generator(device="cuda").manual_seed(11)
    with static_operands(mycc.t1, mycc.t2, eris.ovvv, eris.vvvv, eris.ovoo, eris.ovvo, eris.oovv, eris.oooo):
        for homogeneous in (False, True):
            def fn(a, b):
                return update_lambda_gpu(mycc, mycc.t1, mycc.t2, a, b, eris, imds, True, precision,
                                         homogeneous=homogeneous)
            g = GraphedUpdate(fn, L1, L2)
            for rep in range(4):
                a = L1 * (1 + 0.1 * rep) + 1e-3 * torch.randn(L1.shape, dtype=torch.float64, device="cuda",
                                                               generator=gen)
                b = L2 * (1 - 0.05 * rep)
                r1, r2 = fn(a, b)                       # eager
                torch.cuda.synchronize()
                g1, g2 = g(a, b)                        # replay
                torch.cuda.synchronize()
                e1 = float((g1 - r1).abs().max() / r1.abs().max())
                e2 = float((g2 - r2).abs().max() / r2.abs().max())
                print(precision, homogeneous, rep, e1, e2, flush=True)
                assert e1 < 1e-12 and e2 < 1e-12


@pytest.mark.gpu
def test_one_graph_serves_full_and_homogeneous_updates(gpu, h2o_pair):
    """A graph captured with the device weight `inh` replays F(l) (inh = 1) and A l
    (inh = 0) alternately, each equal to the eager update of that kind."""
    from ccsd_gpu_amd.gpu_ccsd_lambda import GraphedUpdate, make_intermediates_gpu, update_lambda_gpu
    from ccsd_gpu_amd.precision import static_operands
    mol, omf, occ, (l1, l2), gmf = h2o_pair
    mycc = _gpu_cc_with_oracle_orbitals(gmf, omf, occ)
    eris = mycc.ao2mo()
    imds = make_intermediates_gpu(mycc, mycc.t1, mycc.t2, eris)
    L1 = torch.as_tensor(l1, device="cuda")
    L2 = torch.as_tensor(l2, device="cuda")
    inh = torch.ones((), dtype=torch.float64, device="cuda")
    with static_operands(mycc.t1, mycc.t2, eris.ovvv, eris.vvvv, eris.ovoo, eris.ovvo, eris.oovv, eris.oooo):
        g = GraphedUpdate(lambda a, b: update_lambda_gpu(mycc, mycc.t1, mycc.t2, a, b, eris, imds, True, "bf16x3",
                                                         inh=inh), L1, L2)
        for rep, homogeneous in enumerate((True, False, True, False)):
            a, b = L1 * (1 + 0.1 * rep), L2 * (1 - 0.05 * rep)
            r1, r2 = update_lambda_gpu(mycc, mycc.t1, mycc.t2, a, b, eris, imds, True, "bf16x3",
                                       homogeneous=homogeneous)
            inh.fill_(0.0 if homogeneous else 1.0)
            g1, g2 = g(a, b)
            torch.cuda.synchronize()
            e1 = float((g1 - r1).abs().max() / r1.abs().max())
            e2 = float((g2 - r2).abs().max() / r2.abs().max())
            print(homogeneous, rep, e1, e2, flush=True)
            assert e1 < 1e-12 and e2 < 1e-12


@pytest.mark.gpu
def test_split_view_equals_split_of_permuted_copy(gpu):
    from ccsd_gpu_amd import mfma
    gen = torch.Generator(device="cuda").manual_seed(13)
    x = torch.randn((5, 7, 3, 9), dtype=torch.float64, device="cuda", generator=gen)
    for perm in [(0, 1, 2, 3), (2, 0, 3, 1), (3, 2, 1, 0)]:
        v = x.permute(perm)
        shape = (1, v.shape[0] * v.shape[1], v.shape[2] * v.shape[3])
        a = mfma.split_view(v, shape)
        b = mfma.split(v.contiguous().reshape(shape))
        assert torch.equal(a.hi, b.hi) and torch.equal(a.lo, b.lo)


@pytest.mark.gpu
@pytest.mark.parametrize("eq,shapes", [
    ("iakc,bjkc->jiba", [(6, 9, 7, 8), (9, 6, 7, 8)]),
    ("jibc,ca->jiba", [(6, 5, 9, 11), (11, 9)]),
    ("bjic,ca->jiba", [(9, 6, 5, 11), (11, 9)]),
    ("klca,klcb->ba", [(40, 40, 30, 9), (40, 40, 30, 8)]),        # long K: split-K into the scatter
])
def test_contract_acc_scatter_matches_einsum(gpu, eq, shapes, monkeypatch):
    from ccsd_gpu_amd import precision
    monkeypatch.setattr(precision, "MIXED_MIN_FLOPS", 0.0)
    gen = torch.Generator(device="cuda").manual_seed(len(eq))
    A = torch.randn(shapes[0], dtype=torch.float64, device="cuda", generator=gen)
    B = torch.randn(shapes[1], dtype=torch.float64, device="cuda", generator=gen)
    ref = torch.einsum(eq, A, B)
    out0 = torch.randn(ref.shape, dtype=torch.float64, device="cuda", generator=gen)
    got = precision.contract_acc(out0.clone(), eq, A, B, alpha=-0.5, precision="bf16x3")
    want = out0 - 0.5 * ref
    err = float((got - want).abs().max() / ref.abs().max())
    print(eq, err, flush=True)
    assert err < 2e-4


@pytest.mark.gpu
def test_fp64_update_with_factored_wvvov_water_tetramer(gpu):
    """(H2O)4 / cc-pVDZ (o 20, v 76), where wvvov's products run at low precision in the
    mixed solve: the factored FP64 application still gives the formed FP64 update, and the
    low-precision wvvov differs from the FP64 one by that precision's rounding only."""
    from ccsd_gpu_amd import pipeline, systems
    from ccsd_gpu_amd.gpu_ccsd_lambda import make_intermediates_gpu, update_lambda_gpu
    mol = systems.water_cluster(4, seed=1, cart=False)
    mf, mycc = pipeline.setup(mol)
    eris = mycc.ao2mo()
    t1, t2 = mycc.t1, mycc.t2
    full = make_intermediates_gpu(mycc, t1, t2, eris)
    split = make_intermediates_gpu(mycc, t1, t2, eris, lp_precision="bf16x3")
    f1, f2 = update_lambda_gpu(mycc, t1, t2, t1, t2, eris, full)
    g1, g2 = update_lambda_gpu(mycc, t1, t2, t1, t2, eris, split)
    d1 = float((f1 - g1).abs().max()) / float(f1.abs().max())
    d2 = float((f2 - g2).abs().max()) / float(f2.abs().max())
    dw = float((full.wvvov - split.wvvov).abs().max()) / float(full.wvvov.abs().max())
    print("(H2O)4 factored vs formed: l1", d1, "l2", d2, "bf16x3 wvvov", dw, flush=True)
    assert d1 < 1e-12 and d2 < 1e-12
    assert 0 < dw < 1e-4


@pytest.mark.gpu
def test_lambda_graph_captured_once_and_replayed_across_solves(gpu, h2o_pair):
    """The low-precision update is captured by the first mixed solve and replayed by the
    next ones: a second solve on the same CCSD object captures nothing new and gives the
    same lambda; a third one with changed amplitudes (its values written into the
    captured buffers) equals an eager solve with those amplitudes."""
    from ccsd_gpu_amd.gpu_ccsd_lambda import solve_lambda_gpu
    mol, omf, occ, (l1, l2), gmf = h2o_pair
    mycc = _gpu_cc_with_oracle_orbitals(gmf, omf, occ)
    mycc.conv_tol_normt = 1e-10
    a1, a2 = solve_lambda_gpu(mycc, precision="bf16x3", return_tensors=True)
    a1, a2 = a1.clone(), a2.clone()
    entry = next(iter(mycc._lambda_graph_cache.values()))
    b1, b2 = solve_lambda_gpu(mycc, precision="bf16x3", return_tensors=True)
    assert next(iter(mycc._lambda_graph_cache.values())) is entry and mycc._lambda_graph_replays > 0
    assert float((a1 - b1).abs().max()) < 1e-12 and float((a2 - b2).abs().max()) < 1e-12
    assert float(np.abs(b1.cpu().numpy() - l1).max()) < 1e-8
    # changed amplitudes: a t2 scaled by 1.01 (a different, still contractive problem)
    t1s, t2s = mycc.t1 * 0.99, mycc.t2 * 1.01
    g1, g2 = solve_lambda_gpu(mycc, t1=t1s, t2=t2s, precision="bf16x3", return_tensors=True)
    g1, g2 = g1.clone(), g2.clone()
    assert mycc._lambda_graph_replays > 0 and next(iter(mycc._lambda_graph_cache.values())) is entry
    mycc.lambda_graph = False
    e1, e2 = solve_lambda_gpu(mycc, t1=t1s, t2=t2s, precision="bf16x3", return_tensors=True)
    print("graph reuse: changed amplitudes, graph vs eager", float((g1 - e1).abs().max()),
          float((g2 - e2).abs().max()), flush=True)
    assert float((g1 - e1).abs().max()) < 1e-9 and float((g2 - e2).abs().max()) < 1e-9
