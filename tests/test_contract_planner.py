"""CPU check of the contraction planner (precision.contract, FP64 engine) against
torch.einsum for every contraction string used by the lambda / RDM / CCSD code."""
import os
import re

import pytest
import torch

from ccsd_gpu_amd.precision import contract, contract_acc, static_operands

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FILES = ["ccsd_gpu_amd/gpu_ccsd_lambda.py", "ccsd_gpu_amd/gpu_ccsd_grad.py", "ccsd_gpu_amd/cc.py"]


def _equations():
    eqs = set()
    for f in FILES:
        src = open(os.path.join(ROOT, f)).read()
        eqs.update(re.findall(r'"([a-z]+(?:,[a-z]+)+->[a-z]*)"', src))
    return sorted(eqs)


SIZES = {c: 3 + (i % 4) for i, c in enumerate("abcdefghijklmnopqrstuvwxyz")}


@pytest.mark.parametrize("eq", _equations())
def test_planner_matches_einsum(eq):
    lhs, out = eq.split("->")
    g = torch.Generator().manual_seed(len(eq))
    ops = [torch.randn([SIZES[c] for c in t], dtype=torch.float64, generator=g) for t in lhs.split(",")]
    ref = torch.einsum(eq, *ops)
    got = contract(eq, *ops)
    assert got.shape == ref.shape
    assert torch.allclose(got, ref, rtol=1e-12, atol=1e-12), eq
    with static_operands(*ops):
        for _ in range(2):
            assert torch.allclose(contract(eq, *ops), ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("eq", [e for e in _equations() if e.count(",") == 1])
def test_contract_acc_matches_einsum(eq):
    """out += alpha einsum into contiguous and permuted destinations (the FP64 path
    accumulates into a matrix view of out where one fits), plain and static operands."""
    lhs, out_idx = eq.split("->")
    if not out_idx:
        pytest.skip("scalar output")
    g = torch.Generator().manual_seed(len(eq) + 7)
    ops = [torch.randn([SIZES[c] for c in t], dtype=torch.float64, generator=g) for t in lhs.split(",")]
    ref = torch.einsum(eq, *ops)
    for perm in (list(range(len(out_idx))), list(reversed(range(len(out_idx))))):
        base = torch.randn([ref.shape[p] for p in perm], dtype=torch.float64, generator=g)
        inv = [perm.index(i) for i in range(len(perm))]
        for static in (False, True):
            out = base.clone().permute(inv)          # a view labelled by out_idx
            want = out + (-0.5) * ref
            if static:
                with static_operands(*ops):
                    contract_acc(out, eq, ops[0], ops[1], alpha=-0.5)
            else:
                contract_acc(out, eq, ops[0], ops[1], alpha=-0.5)
            assert torch.allclose(out, want, rtol=1e-12, atol=1e-12), (eq, perm, static)


def test_planner_handles_transposed_and_permuted_views():
    g = torch.Generator().manual_seed(0)
    x = torch.randn(4, 5, 6, 7, dtype=torch.float64, generator=g)
    y = torch.randn(6, 7, 3, dtype=torch.float64, generator=g)
    for eq, ops in [("ijkl,klm->ijm", (x, y)), ("ijkl,klm->mij", (x, y)),
                    ("jilk,klm->ijm", (x.permute(1, 0, 3, 2), y)), ("ijkl,klm->ijm", (x, y.transpose(0, 1).contiguous().transpose(0, 1)))]:
        assert torch.allclose(contract(eq, *ops), torch.einsum(eq, *ops), rtol=1e-12, atol=1e-12)


def test_static_cache_reuses_operands_in_several_roles():
    """One static tensor used with different index groupings (CCSD update on random integrals)."""
    from ccsd_gpu_amd import cc
    o, v = 4, 7
    g = torch.Generator().manual_seed(1)
    R = lambda *s: torch.randn(s, dtype=torch.float64, generator=g) * 0.1  # noqa: E731

    class Eris:
        pass
    e = Eris()
    e.fock = torch.diag(torch.cat([torch.linspace(-2, -0.5, o), torch.linspace(0.2, 2, v)])).double()
    e.mo_energy = torch.diagonal(e.fock).clone()
    e.oooo, e.ovoo, e.ovvo = R(o, o, o, o), R(o, v, o, o), R(o, v, v, o)
    e.oovv, e.ovov, e.ovvv, e.vvvv = R(o, o, v, v), R(o, v, o, v), R(o, v, v, v), R(v, v, v, v)
    t1, t2 = R(o, v), R(o, o, v, v)
    ref = cc.update_amps(t1, t2, e)
    with static_operands(e.ovov, e.ovoo, e.ovvo, e.oovv, e.ovvv, e.oooo, e.vvvv):
        for _ in range(2):
            got = cc.update_amps(t1, t2, e)
            assert torch.allclose(got[0], ref[0], rtol=1e-12, atol=1e-14)
            assert torch.allclose(got[1], ref[1], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("k", [8192, 10001])
def test_split_k_dgemm_matches_mm(k):
    """Small output, long contracted dimension: the K-split batched product (with a ragged
    remainder and for a transposed left operand) equals the plain product."""
    from ccsd_gpu_amd.precision import _dgemm
    g = torch.Generator().manual_seed(k)
    A = torch.randn(2, 3, k, dtype=torch.float64, generator=g)
    At = torch.randn(2, k, 3, dtype=torch.float64, generator=g).transpose(1, 2)
    B = torch.randn(2, k, 5, dtype=torch.float64, generator=g)
    for a in (A, At, A[:1], At[:1]):
        ref = torch.bmm(a, B[:a.shape[0]])
        got = _dgemm(a, B[:a.shape[0]])
        assert torch.allclose(got, ref, rtol=1e-12, atol=1e-11)


def test_k_splits_policy():
    from ccsd_gpu_amd.mfma import k_splits
    assert k_splits(1, 1000) == 1                  # short K: no split
    assert k_splits(4096, 10 ** 6) == 1            # enough tiles already
    assert k_splits(1, 10 ** 6) == 976             # 1024 chunks wanted, limited by chunk >= 1024
    assert k_splits(4, 10 ** 7) == 256


def test_rotated_one_index_transforms_give_full_mo_to_ao():
    """Four rotate-transforms of a (p q r s) tensor equal C x C x C x C in (i j k l) order."""
    from ccsd_gpu_amd.gpu_ccsd_grad import _rotate_transform
    g = torch.Generator().manual_seed(3)
    d = torch.randn(4, 4, 4, 4, dtype=torch.float64, generator=g)
    c = torch.randn(6, 4, dtype=torch.float64, generator=g)
    t = d
    for _ in range(4):
        t = _rotate_transform(t, c, "fp64")
    ref = torch.einsum("ip,jq,kr,ls,pqrs->ijkl", c, c, c, c, d)
    assert t.shape == (6, 6, 6, 6)
    assert torch.allclose(t, ref, rtol=1e-12, atol=1e-12)
    # one step: last index transformed, moved to the front
    one = _rotate_transform(d, c, "fp64")
    assert torch.allclose(one, torch.einsum("pqrs,ls->lpqr", d, c), rtol=1e-12, atol=1e-12)


def test_packed_transform_equals_packed_full_transform():
    """MO->AO with the kl pair packed after the first two transforms equals the full
    transform packed afterwards (k >= l, diagonal halved), for a (pq)(rs)-symmetric D."""
    from ccsd_gpu_amd.gpu_ccsd_grad import _rotate_transform
    g = torch.Generator().manual_seed(5)
    d = torch.randn(5, 5, 5, 5, dtype=torch.float64, generator=g)
    d = d + d.permute(1, 0, 2, 3)
    d = d + d.permute(0, 1, 3, 2)
    c = torch.randn(7, 5, dtype=torch.float64, generator=g)
    full = torch.einsum("ip,jq,kr,ls,pqrs->ijkl", c, c, c, c, d)
    k, l = torch.tril_indices(7, 7)
    ref = full[:, :, k, l].clone()
    ref[:, :, k == l] *= 0.5
    t = _rotate_transform(_rotate_transform(d, c, "fp64"), c, "fp64")
    tp = t[k, l]
    tp[k == l] *= 0.5
    got = _rotate_transform(_rotate_transform(tp, c, "fp64"), c, "fp64")
    assert got.shape == ref.shape
    assert torch.allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_mo_packed_then_repacked_transform_equals_packed_full_transform():
    """MO->AO with the first pair symmetrised and packed in the MO basis, the second
    pair transformed, the packing moved to the second (AO) pair, then the first pair
    transformed: equals the full transform of the (pq)(rs)-symmetrised D packed over
    k >= l with weight 2 off the diagonal (1 on it), for a general D."""
    from ccsd_gpu_amd.gpu_ccsd_grad import _pack_pairs_sym, _repack_pairs, _rotate_transform
    g = torch.Generator().manual_seed(8)
    d = torch.randn(5, 5, 5, 5, dtype=torch.float64, generator=g)
    c = torch.randn(7, 5, dtype=torch.float64, generator=g)
    full = torch.einsum("ip,jq,kr,ls,pqrs->ijkl", c, c, c, c, d)
    sym = 0.25 * (full + full.permute(1, 0, 2, 3) + full.permute(0, 1, 3, 2) + full.permute(1, 0, 3, 2))
    k, l = torch.tril_indices(7, 7)
    ref = 2.0 * sym[:, :, k, l]
    ref[:, :, k == l] *= 0.5
    t = _rotate_transform(_rotate_transform(_pack_pairs_sym(d), c, "fp64"), c, "fp64")
    t = _repack_pairs(t, 5)
    got = _rotate_transform(_rotate_transform(t, c, "fp64"), c, "fp64")
    assert got.shape == ref.shape
    assert torch.allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_pack_pairs_sym_symmetrises_and_packs():
    """(ij)(kl) symmetrisation (weight 1/2 per term) folded into the k >= l packing."""
    from ccsd_gpu_amd.gpu_ccsd_grad import _pack_pairs_sym
    g = torch.Generator().manual_seed(6)
    t = torch.randn(6, 6, 4, 4, dtype=torch.float64, generator=g)
    sym = 0.5 * (t + t.permute(1, 0, 2, 3) + t.permute(0, 1, 3, 2) + t.permute(1, 0, 3, 2))
    k, l = torch.tril_indices(6, 6)
    ref = sym[k, l].clone()
    ref[k == l] *= 0.5
    assert torch.allclose(_pack_pairs_sym(t), ref, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("o,v", [(1, 1), (3, 5), (4, 9)])
def test_vvvv_pair_halves_match_einsum_cpu(o, v):
    """The symmetric / antisymmetric pair form of the l2.vvvv term (host logic: pair
    maps, pair matrices, recombination) against the full contraction, FP64 on the CPU."""
    from ccsd_gpu_amd.precision import vvvv_term
    g = torch.Generator().manual_seed(o * 10 + v)
    x = torch.randn((v, v, v, v), dtype=torch.float64, generator=g)
    vvvv = x + x.permute(1, 0, 2, 3)
    vvvv = vvvv + vvvv.permute(0, 1, 3, 2)
    vvvv = vvvv + vvvv.permute(2, 3, 0, 1)
    t = torch.randn((o, o, v, v), dtype=torch.float64, generator=g)
    l2 = t + t.permute(1, 0, 3, 2)
    ref = torch.einsum("ijcd,acbd->ijab", l2, vvvv)
    got = vvvv_term(l2, vvvv, "fp64", scale=0.5)
    assert torch.allclose(got, 0.5 * ref, atol=1e-12 * float(ref.abs().max()), rtol=0)
