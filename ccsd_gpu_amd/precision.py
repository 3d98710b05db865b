"""Tensor contractions of the lambda / RDM path at a chosen precision.

    contract("ijcd,acbd->ijab", l2, vvvv, precision="fp64" | "fp32" | "bf16x3")

fp64    torch.einsum in float64 (rocBLAS / hipBLASLt DGEMM): the plain path.
fp32    the einsum is planned into one (batched) GEMM (operands permuted to
        [batch, m, k] x [batch, k, n]) and run by the HIP MFMA GEMM kernel
        (v_mfma_f32 FP32 inputs, FP32 accumulation); result back to float64.
bf16x3  same plan, each operand split A = A_hi + A_lo into two bfloat16
        parts and C = A_hi B_hi + A_hi B_lo + A_lo B_hi accumulated in FP32 on
        the bf16 MFMA units.
Three-operand expressions are contracted left to right.
"""
import torch

from . import mfma


def _plan(eq):
    lhs, out = eq.split("->")
    a, b = lhs.split(",")
    batch = [c for c in out if c in a and c in b]
    contr = [c for c in a if c in b and c not in out]
    left = [c for c in a if c not in b]
    right = [c for c in b if c not in a]
    return a, b, out, batch, contr, left, right


def _gemm_contract(eq, A, B, precision):
    a, b, out, batch, contr, left, right = _plan(eq)
    # indices present in only one operand and absent from the output are summed first
    for idx in [c for c in a if c not in b and c not in out]:
        A = A.sum(dim=a.index(idx))
        a = a.replace(idx, "")
    for idx in [c for c in b if c not in a and c not in out]:
        B = B.sum(dim=b.index(idx))
        b = b.replace(idx, "")
    a, b, out, batch, contr, left, right = _plan(f"{a},{b}->{out}")
    sa = dict(zip(a, A.shape))
    sb = dict(zip(b, B.shape))
    size = {**sa, **sb}
    nb = 1
    for c in batch:
        nb *= size[c]
    m = 1
    for c in left:
        m *= size[c]
    k = 1
    for c in contr:
        k *= size[c]
    n = 1
    for c in right:
        n *= size[c]
    Ap = A.permute([a.index(c) for c in batch + left + contr]).reshape(nb, m, k)
    Bp = B.permute([b.index(c) for c in batch + contr + right]).reshape(nb, k, n)
    Cp = mfma.bmm(Ap, Bp, precision)
    C = Cp.reshape([size[c] for c in batch + left + right])
    order = batch + left + right
    return C.permute([order.index(c) for c in out])


def contract(eq, *ops, precision="fp64"):
    if precision == "fp64":
        return torch.einsum(eq, *ops)
    lhs, out = eq.replace(" ", "").split("->")
    terms = lhs.split(",")
    if len(terms) == 1:
        return torch.einsum(eq, *ops)
    if len(terms) == 2:
        return _gemm_contract(f"{terms[0]},{terms[1]}->{out}", ops[0], ops[1], precision)
    # left to right: keep every index still needed later or in the output
    cur, cur_t = ops[0], terms[0]
    for i in range(1, len(terms)):
        rest = "".join(terms[i + 1:]) + out
        keep = "".join(dict.fromkeys(c for c in cur_t + terms[i] if c in rest))
        cur = _gemm_contract(f"{cur_t},{terms[i]}->{keep}", cur, ops[i], precision)
        cur_t = keep
    if cur_t != out:
        cur = cur.permute([cur_t.index(c) for c in out])
    return cur
