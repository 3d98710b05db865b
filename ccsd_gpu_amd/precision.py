"""Tensor contractions of the lambda / RDM / CCSD code at a chosen precision.

    contract("ijcd,acbd->ijab", l2, vvvv, precision="fp64" | "fp32" | "bf16x3")

Every two-operand contraction is planned into one (batched) GEMM
C[b] = A[b] B[b] with index groups batch / free-left / contracted / free-right.
Each operand is viewed as a matrix in the memory order it already has where
that is possible (a transposed view is passed to the GEMM instead of a copy);
only operands whose indices are interleaved are permuted into a copy. Then:

fp64    DGEMM (rocBLAS / hipBLASLt via torch.matmul): the plain path.
bf16x3  contractions of at least MIXED_MIN_FLOPS run on the split-BF16 MFMA
        kernel (bfloat16 hi/lo planes, three products, FP32 accumulation,
        float64 out); smaller ones stay FP64. Operands registered with
        `static_operands` (integrals and intermediates fixed during a solve)
        are split once and reused.
fp32    contractions of at least MIXED_MIN_FLOPS in FP32 (operands rounded once,
        static ones cached) on the BLAS library's SGEMM, FP32 accumulation,
        float64 out; smaller ones stay FP64.
Three-operand expressions are contracted left to right.
"""
from contextlib import contextmanager
from math import prod

import torch
from torch.utils.weak import WeakIdKeyDictionary

from . import mfma

MIXED_MIN_FLOPS = 4e9
_STATIC = None
# number of products issued at each low precision (read by tests to confirm that a
# mixed-precision run really used the MFMA / FP32 path)
LOWP_PRODUCTS = {"bf16x3": 0, "fp32": 0}


@contextmanager
def static_operands(*tensors, state=None):
    """Within the block, prepared copies of these tensors are cached (they must not change).

    Yields the cache state; passing a state back (`state=`) re-enters that cache (its
    registered tensors and prepared copies), e.g. to refresh it in place with
    refresh_static after the registered tensors were overwritten."""
    global _STATIC
    prev = _STATIC
    if state is None:
        state = {"ids": {id(t) for t in tensors if t is not None}, "keep": list(tensors), "cache": {},
                 "makers": {}}
    _STATIC = state
    try:
        yield state
    finally:
        _STATIC = prev


def refresh_static(state, keys=None):
    """Re-prepare the cached copies of a static_operands state (those of `keys`, default
    all) from their (overwritten) source tensors and write them into the cached storage
    in place, so that device pointers taken from the cache (a captured HIP graph) stay
    valid. A state with state["trace"] set to a set records the keys read meanwhile."""
    global _STATIC
    prev = _STATIC
    _STATIC = None                      # makers run uncached
    try:
        for k in (state["makers"] if keys is None else keys):
            make = state["makers"][k]
            new, old = make(), state["cache"][k]
            if isinstance(old, torch.Tensor):
                old.copy_(new)
            else:                       # a Split of bf16 planes
                old.hi.copy_(new.hi)
                old.lo.copy_(new.lo)
    finally:
        _STATIC = prev


def _static_key(X):
    """Cache key of a registered tensor, or of a view of one (made anew by each caller,
    e.g. a permuted integral block): the base tensor plus the view's geometry."""
    ids = _STATIC["ids"]
    if id(X) in ids:
        return (id(X),)
    base = getattr(X, "_base", None)
    if base is not None and id(base) in ids:
        return (id(base), X.storage_offset(), tuple(X.shape), tuple(X.stride()))
    return None


def _cached(X, key, make):
    if _STATIC is not None:
        sk = _static_key(X)
        if sk is not None:
            k = sk + key
            v = _STATIC["cache"].get(k)
            if v is None:
                v = make()
                _STATIC["cache"][k] = v
                _STATIC["makers"][k] = make
            trace = _STATIC.get("trace")
            if trace is not None:
                trace.add(k)
            return v
    return make()


def _groups(a, b, out):
    batch = [c for c in out if c in a and c in b]
    contr = [c for c in a if c in b and c not in out]
    left = [c for c in a if c not in b]
    right = [c for c in b if c not in a]
    return batch, left, contr, right


def _view(X, s, g1, g2, g3, size):
    """X (index string s) as a (n1, n2, n3) matrix batch in group order g1 g2 g3.

    Returns (tensor, copied): a reshape view when s is g1+g2+g3, a transposed view
    when s is g1+g3+g2, a permuted copy otherwise."""
    n1, n2, n3 = prod(size[c] for c in g1), prod(size[c] for c in g2), prod(size[c] for c in g3)
    if s == "".join(g1 + g2 + g3) and X.is_contiguous():
        return X.reshape(n1, n2, n3), False
    if s == "".join(g1 + g3 + g2) and X.is_contiguous():
        return X.reshape(n1, n3, n2).transpose(1, 2), False
    perm = [s.index(c) for c in g1 + g2 + g3]
    return mfma.permute_copy(X.permute(perm)).reshape(n1, n2, n3), True


def _order_within(group, s):
    return sorted(group, key=s.index)


def _dgemm(Av, Bv):
    """FP64 batched product (nb, m, k) x (nb, k, n) on the BLAS library. Outputs with few
    tiles and a long K are split along K (strided views of both operands, one batched
    call, then a sum), otherwise such a product runs on a handful of workgroups."""
    nb, m, k = Av.shape
    n = Bv.shape[2]
    if Av.is_cuda and (m == 1 or n == 1) and k > 0:
        # matrix-vector shapes: the library's GEMV kernels (coalesced along the matrix's
        # contiguous index, K sliced for parallelism) instead of a 128-wide GEMM tile
        out = torch.empty((nb, m, n), dtype=Av.dtype, device=Av.device)
        for b in range(nb):
            if n == 1:
                mfma.gemv(Av[b], Bv[b, :, 0], out[b, :, 0])
            else:
                mfma.gemv(Bv[b].T, Av[b, 0, :], out[b, 0, :])
        return out
    s = mfma.k_splits(nb * -(-m // 64) * -(-n // 64), k, want=512, min_chunk=2048) if nb <= 8 else 1
    if s <= 1 and nb == 1 and Av.is_cuda and k >= 8192 and m > 1 and n > 1:
        s = _tuned_dgemm_split(Av, Bv)
    return _dgemm_run(Av, Bv, s)


_DGEMM_SPLIT = {}


def _tuned_dgemm_split(Av, Bv, candidates=(1, 2, 4)):
    """K split of a long-K FP64 product whose output tiles leave the last round of resident
    workgroups part empty (e.g. 6080 x 1600 outputs over K = 23104): the candidates are
    timed once per shape and operand layout and the fastest cached; 1 under capture."""
    _, m, k = Av.shape
    n = Bv.shape[2]
    key = (m, k, n, tuple(Av.stride()), tuple(Bv.stride()))
    s = _DGEMM_SPLIT.get(key)
    if s is not None:
        return s
    if torch.cuda.is_current_stream_capturing():
        return 1
    stream = torch.cuda.current_stream()
    best, best_ms = 1, float("inf")
    for c in candidates:
        _dgemm_run(Av, Bv, c)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        _dgemm_run(Av, Bv, c)
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_time(e1)
        if ms < best_ms:
            best, best_ms = c, ms
    _DGEMM_SPLIT[key] = best
    return best


def _dgemm_run(Av, Bv, s):
    nb, m, k = Av.shape
    n = Bv.shape[2]
    if s <= 1:
        return torch.mm(Av[0], Bv[0]).unsqueeze(0) if nb == 1 else torch.bmm(Av, Bv)
    kc = k // s
    out = torch.empty((nb, m, n), dtype=Av.dtype, device=Av.device)
    for b in range(nb):
        a, bb = Av[b], Bv[b]
        a3 = a.as_strided((s, m, kc), (kc * a.stride(1), a.stride(0), a.stride(1)))
        b3 = bb.as_strided((s, kc, n), (kc * bb.stride(0), bb.stride(0), bb.stride(1)))
        part = torch.bmm(a3, b3)
        if s * kc < k:
            part[0] += a[:, s * kc:] @ bb[s * kc:]
        if out.is_cuda:
            mfma.sum_slices(part, out[b])
        else:
            torch.sum(part, 0, out=out[b])
    return out


_SPLITS = None


@contextmanager
def split_memo():
    """Within the block, the split-BF16 planes of a non-static operand are made once
    per (storage, version, permuted layout) and reused by every contraction that
    reads the same layout (the planes depend only on the permuted view, not on how
    its indices are grouped)."""
    global _SPLITS
    prev = _SPLITS
    _SPLITS = {}
    try:
        yield
    finally:
        _SPLITS = prev


def _split_operand(X, s, g1, g2, g3, size, tag):
    """bf16 planes of X (index string s) as an (n1, n2, n3) matrix batch in group order
    g1 g2 g3, permuted and split in one pass; cached for static operands (and, inside
    `split_memo`, per layout for the others)."""
    n1, n2, n3 = prod(size[c] for c in g1), prod(size[c] for c in g2), prod(size[c] for c in g3)
    perm = [s.index(c) for c in g1 + g2 + g3]

    def make():
        if _SPLITS is None or not X.is_cuda:
            return mfma.split_view(X.permute(perm), (n1, n2, n3))
        v = X.permute(perm)
        key = (X.data_ptr(), X._version, tuple(v.shape), tuple(v.stride()))
        hit = _SPLITS.get(key)
        if hit is None:
            # the entry keeps X alive, so its storage (the key) cannot be reused meanwhile
            hit = (X, mfma.split_view(v, (v.numel(),)))
            _SPLITS[key] = hit
        planes = hit[1]
        return mfma.Split(planes.hi.view(n1, n2, n3), planes.lo.view(n1, n2, n3))

    return _cached(X, (s, tag, tuple(g1), tuple(g2), tuple(g3)), make)


_OFFSETS = {}


def _offsets(out, out_idx, group, size):
    """int64 device offsets (elements) of the row-major enumeration of `group`'s indices
    inside `out` (indices labelled by out_idx)."""
    strides = tuple(out.stride(out_idx.index(c)) for c in group)
    key = (tuple(group), tuple(size[c] for c in group), strides, out.device)
    v = _OFFSETS.get(key)
    if v is None:
        v = torch.zeros(1, dtype=torch.int64, device=out.device)
        for c, st in zip(group, strides):
            v = (v[:, None] + torch.arange(size[c], dtype=torch.int64, device=out.device) * st).reshape(-1)
        _OFFSETS[key] = v
    return v


def _prepare(eq, A, B):
    lhs, out = eq.split("->")
    a, b = lhs.split(",")
    # indices that appear in one operand only and not in the output are summed first
    for idx in [c for c in a if c not in b and c not in out]:
        A = A.sum(dim=a.index(idx))
        a = a.replace(idx, "")
    for idx in [c for c in b if c not in a and c not in out]:
        B = B.sum(dim=b.index(idx))
        b = b.replace(idx, "")
    # the output's fastest index goes to the GEMM's column side, so that the product
    # (and the scatter epilogue) is written along it
    if out and out[-1] in a and out[-1] not in b and any(c in out for c in b if c not in a):
        A, B, a, b = B, A, b, a
    size = {**dict(zip(a, A.shape)), **dict(zip(b, B.shape))}
    batch, left, contr, right = _groups(a, b, out)
    batch = _order_within(batch, a)
    left = _order_within(left, a)
    right = _order_within(right, b)
    # contracted order: follow A unless B then avoids a copy and A does not need one anyway
    ca, cb = _order_within(contr, a), _order_within(contr, b)
    contr = ca
    if ca != cb:
        a_ok = a in ("".join(batch + left + ca), "".join(batch + ca + left))
        b_ok_cb = b in ("".join(batch + cb + right), "".join(batch + right + cb))
        if b_ok_cb and not a_ok:
            contr = cb
    nb, m, k, n = prod(size[c] for c in batch), prod(size[c] for c in left), \
        prod(size[c] for c in contr), prod(size[c] for c in right)
    flops = 2.0 * nb * m * n * k
    return A, B, a, b, out, size, batch, left, contr, right, flops


def _gemm_contract(eq, A, B, precision):
    A, B, a, b, out, size, batch, left, contr, right, flops = _prepare(eq, A, B)
    if precision == "bf16x3" and flops >= MIXED_MIN_FLOPS:
        As = _split_operand(A, a, batch, left, contr, size, "nt-a")
        Bs = _split_operand(B, b, batch, right, contr, size, "nt-b")
        Cp = mfma.gemm_nt(As, Bs)
        LOWP_PRODUCTS["bf16x3"] += 1
    else:
        Av = _cached(A, (a, "mk", tuple(batch), tuple(left), tuple(contr)),
                     lambda: _view(A, a, batch, left, contr, size)[0])
        Bv = _cached(B, (b, "kn", tuple(batch), tuple(contr), tuple(right)),
                     lambda: _view(B, b, batch, contr, right, size)[0])
        if precision == "fp32" and flops >= MIXED_MIN_FLOPS and Av.is_cuda:
            # FP32 operands (static ones converted once) on the library's SGEMM
            Af = _cached(A, (a, "mk32", tuple(batch), tuple(left), tuple(contr)), lambda: Av.float())
            Bf = _cached(B, (b, "kn32", tuple(batch), tuple(contr), tuple(right)), lambda: Bv.float())
            Cp = torch.bmm(Af, Bf).double()
            LOWP_PRODUCTS["fp32"] += 1
        elif precision == "fp32" and flops >= MIXED_MIN_FLOPS:
            Cp = mfma.bmm(Av, Bv, "fp32")
            LOWP_PRODUCTS["fp32"] += 1
        else:
            Cp = _dgemm(Av, Bv)
    C = Cp.reshape([size[c] for c in batch + left + right])
    order = batch + left + right
    return C.permute([order.index(c) for c in out])


def _plain_layout(X, s, g1, g2, g3):
    """True when _view gives X as a matrix batch without a copy."""
    return X.is_contiguous() and s in ("".join(g1 + g2 + g3), "".join(g1 + g3 + g2))


def _acc_fp64_direct(out, oidx, A, B, a, b, size, batch, left, contr, right, alpha):
    """out += alpha A B as one FP64 GEMM accumulating into a matrix view of `out` (no
    product tensor, no separate add). The free indices are taken in the planner's order
    or in the order they lie in `out`; an operand then needing a permuted copy must be
    static (copied once per solve) or small. Returns False when no view fits."""
    nb, m, n = (prod(size[c] for c in g) for g in (batch, left, right))
    if m <= 1 or n <= 1 or not contr:
        return False
    by_out = lambda g: sorted(g, key=lambda c: -out.stride(oidx.index(c)))  # noqa: E731
    cheap = lambda X, s, g1, g2, g3: (_plain_layout(X, s, g1, g2, g3) or X.numel() <= 1 << 20  # noqa: E731
                                      or (_STATIC is not None and _static_key(X) is not None))
    for lft, rgt in ((left, right), (by_out(left), by_out(right))):
        try:
            outv = out.permute([oidx.index(c) for c in batch + lft + rgt]).view(nb, m, n)
        except RuntimeError:
            continue
        if not (cheap(A, a, batch, lft, contr) and cheap(B, b, batch, contr, rgt)):
            continue
        Av = _cached(A, (a, "mk", tuple(batch), tuple(lft), tuple(contr)),
                     lambda: _view(A, a, batch, lft, contr, size)[0])
        Bv = _cached(B, (b, "kn", tuple(batch), tuple(contr), tuple(rgt)),
                     lambda: _view(B, b, batch, contr, rgt, size)[0])
        if nb == 1:
            outv[0].addmm_(Av[0], Bv[0], alpha=alpha)
        else:
            outv.baddbmm_(Av, Bv, alpha=alpha)
        return True
    return False


def contract_acc(out, eq, A, B, alpha=1.0, precision="fp64"):
    """out += alpha * einsum(eq, A, B), in place; eq's output labels out's indices.

    On the split-BF16 path the GEMM epilogue adds the term straight into `out`
    through offset tables (no product tensor, no permuted copy); FP64 products
    accumulate into a matrix view of `out` where one exists; otherwise the product
    is formed and added."""
    lhs, oidx = eq.replace(" ", "").split("->")
    A2, B2, a, b, o, size, batch, left, contr, right, flops = _prepare(f"{lhs}->{oidx}", A, B)
    if precision == "bf16x3" and out.is_cuda and flops >= MIXED_MIN_FLOPS:
        As = _split_operand(A2, a, batch, left, contr, size, "nt-a")
        Bs = _split_operand(B2, b, batch, right, contr, size, "nt-b")
        mfma.gemm_nt_scatter(As, Bs, out, _offsets(out, oidx, batch, size), _offsets(out, oidx, left, size),
                             _offsets(out, oidx, right, size), alpha=alpha, beta=1)
        LOWP_PRODUCTS["bf16x3"] += 1
        return out
    if (precision == "fp64" or flops < MIXED_MIN_FLOPS) and out.dtype == torch.float64 and \
            _acc_fp64_direct(out, oidx, A2, B2, a, b, size, batch, left, contr, right, alpha):
        return out
    return out.add_(contract(eq, A, B, precision=precision), alpha=alpha)


def contract(eq, *ops, precision="fp64"):
    lhs, out = eq.replace(" ", "").split("->")
    terms = lhs.split(",")
    if len(terms) == 1:
        return torch.einsum(eq, *ops)
    if len(terms) == 2:
        return _gemm_contract(f"{terms[0]},{terms[1]}->{out}", ops[0], ops[1], precision)
    cur, cur_t = ops[0], terms[0]
    for i in range(1, len(terms)):
        rest = "".join(terms[i + 1:]) + out
        keep = "".join(dict.fromkeys(c for c in cur_t + terms[i] if c in rest))
        cur = _gemm_contract(f"{cur_t},{terms[i]}->{keep}", cur, ops[i], precision)
        cur_t = keep
    if cur_t != out:
        cur = cur.permute([cur_t.index(c) for c in out])
    return cur


_FP64_KSPLIT = {}


def _fp64_matmul_ksplit(A, B):
    """A @ B (FP64, A contiguous (m, K), B contiguous (K, n)) with K cut into `ks` equal
    chunks run as one batched GEMM and summed; ks from {1, 2, 4, 8} (dividing K) timed
    once per shape and cached (not timed under HIP-graph capture: ks = 1 there). The
    (H2O)8 l2.vvvv shape, 820 x 23104 x 23104: 16.0 ms unsplit, 13.8 ms at ks = 4."""
    m, K = A.shape
    n = B.shape[1]

    def run(ks):
        if ks == 1:
            return A @ B
        kc = K // ks
        return torch.bmm(A.view(m, ks, kc).transpose(0, 1), B.view(ks, kc, n)).sum(0)

    key = (m, K, n)
    ks = _FP64_KSPLIT.get(key)
    if ks is None:
        if not A.is_cuda or torch.cuda.is_current_stream_capturing():
            return run(1)
        best, best_ms = 1, float("inf")
        stream = torch.cuda.current_stream()
        for c in (1, 2, 4, 8):
            if K % c or (c > 1 and K // c < 1024):
                continue
            run(c)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            run(c)
            e1.record(stream)
            e1.synchronize()
            ms = e0.elapsed_time(e1)
            if ms < best_ms:
                best, best_ms = c, ms
        _FP64_KSPLIT[key] = ks = best
    return run(ks)


# (W+, W-) of each vvvv tensor in use, per precision: built once and kept while the
# tensor lives (the integrals of one ERIs object do not change)
_VVVV_PM = WeakIdKeyDictionary()


_PAIR_MAPS = {}


def _pair_maps(v, device):
    """Lower-triangle pairs P = (c, d), c >= d, in torch.tril_indices order: the flat
    indices c v + d and d v + c, the weight 1/2 of the diagonal pairs (1 elsewhere),
    and for every (a, b) the pair index of (max, min) with the sign +1 (a >= b) or
    -1 (a < b). Made once per (v, device)."""
    key = (v, str(device))
    maps = _PAIR_MAPS.get(key)
    if maps is None:
        tc, td = torch.tril_indices(v, v, device=device)
        cd, dc = tc * v + td, td * v + tc
        wd = torch.where(tc == td, 0.5, 1.0).to(torch.float64)
        a = torch.arange(v, device=device)
        hi, lo = torch.maximum(a[:, None], a[None, :]), torch.minimum(a[:, None], a[None, :])
        pidx = (hi * (hi + 1) // 2 + lo).flatten()
        sign = torch.where(a[:, None] >= a[None, :], 1.0, -1.0).to(torch.float64).flatten()
        maps = _PAIR_MAPS[key] = (tc, td, cd, dc, wd, pidx, sign)
    return maps


def _vvvv_pm(vvvv, precision):
    """W[0] = W+, W[1] = W- over pairs Q = (a, b), a >= b and P = (c, d), c >= d:
    W+-[Q, P] = (ac|bd) +- (ad|bc), symmetric matrices (n_pair x n_pair) in FP64 or FP32,
    or as split-BF16 planes with P padded by zeros to a multiple of 8 (16-byte rows
    for the GEMM's aligned loads); cached for the life of `vvvv`."""
    ent = _VVVV_PM.get(vvvv)
    if ent is None:
        ent = {}
        _VVVV_PM[vvvv] = ent
    if precision in ent:
        return ent[precision]
    W = ent.get("fp64")
    if W is None:
        v = vvvv.shape[0]
        tc, td, cd, dc, _, _, _ = _pair_maps(v, vvvv.device)
        n = tc.numel()
        W = torch.empty((2, n, n), dtype=torch.float64, device=vvvv.device)
        for q0 in range(0, n, 2048):                      # (ac|bd) rows of 2048 (a, b) pairs at a time
            q1 = min(n, q0 + 2048)
            M = vvvv[tc[q0:q1], :, td[q0:q1], :].reshape(q1 - q0, v * v)   # M[q, c v + d] = (ac|bd)
            x, y = M[:, cd], M[:, dc]
            torch.add(x, y, out=W[0, q0:q1])
            torch.sub(x, y, out=W[1, q0:q1])
            del M, x, y
        ent["fp64"] = W
    if precision == "bf16x3":
        n = W.shape[1]
        Wp = torch.zeros((2, n, _pad8(n)), dtype=torch.float64, device=W.device)
        Wp[:, :, :n] = W
        ent[precision] = mfma.split(Wp)
        del Wp
    elif precision != "fp64":
        ent[precision] = W.float()
    return ent[precision]


def _pad8(n):
    return -(-n // 8) * 8


def vvvv_term(l2, vvvv, precision, scale=1.0):
    """X[i,j,a,b] = sum_cd l2[i,j,c,d] (ac|bd) (the l2.vvvv term of the lambda update),
    times `scale`.

    X[j,i,b,a] = X[i,j,a,b] for the 8-fold symmetric vvvv, so only i <= j is evaluated
    and the rest mirrored. Over the (a, b) and (c, d) pairs the product splits into a
    symmetric and an antisymmetric half,
        X[ab] +- X[ba] = sum_{c >= d} (l2[cd] +- l2[dc]) ((ac|bd) +- (ad|bc))
    (the c = d terms halved on the left, absent from the minus half), two products of
    v(v+1)/2 x v(v+1)/2 pair matrices run as one batch: a quarter of the o^2 v^4 work of
    the plain contraction, half that of its i <= j half. On the device one kernel packs
    the pair rows (split into BF16 planes for bf16x3) and one writes the dense result."""
    o, _, v, _ = l2.shape
    iu, ju = torch.triu_indices(o, o, device=l2.device)
    tc, td, cd, dc, wd, pidx, sign = _pair_maps(v, l2.device)
    npair, n = iu.numel(), tc.numel()
    W = _vvvv_pm(vvvv, precision)
    if l2.is_cuda and l2.dtype == torch.float64:
        from . import _lib
        lib = _lib.load()
        l2c = l2.contiguous()
        if precision == "bf16x3":
            kp = _pad8(n)
            planes = mfma.Split(torch.empty((2, npair, kp), dtype=torch.bfloat16, device=l2.device),
                                torch.empty((2, npair, kp), dtype=torch.bfloat16, device=l2.device))
            _lib.check(lib.ccsd_vvvv_pm_pack(_lib.ptr(l2c), o, v, _lib.ptr(iu), _lib.ptr(ju), npair, _lib.ptr(tc),
                                             _lib.ptr(td), n, kp, 1, _lib.ptr(planes.hi), _lib.ptr(planes.lo),
                                             _lib.current_stream_ptr()), "ccsd_vvvv_pm_pack")
            Xpm = mfma.gemm_nt(planes, W, ksplit=mfma.tuned_ksplit(planes, W))
            LOWP_PRODUCTS["bf16x3"] += 1
        else:
            Apm = torch.empty((2, npair, n), dtype=torch.float64, device=l2.device)
            _lib.check(lib.ccsd_vvvv_pm_pack(_lib.ptr(l2c), o, v, _lib.ptr(iu), _lib.ptr(ju), npair, _lib.ptr(tc),
                                             _lib.ptr(td), n, n, 0, _lib.ptr(Apm), None,
                                             _lib.current_stream_ptr()), "ccsd_vvvv_pm_pack")
            if precision == "fp64":
                Xpm = torch.stack((_fp64_matmul_ksplit(Apm[0], W[0]), _fp64_matmul_ksplit(Apm[1], W[1])))
            else:
                Xpm = torch.bmm(Apm.float(), W).double()
                LOWP_PRODUCTS["fp32"] += 1
        out = torch.empty_like(l2c)
        _lib.check(lib.ccsd_vvvv_pm_unpack(_lib.ptr(Xpm), o, v, Xpm.shape[-1], float(scale), _lib.ptr(out),
                                           _lib.current_stream_ptr()), "ccsd_vvvv_pm_unpack")
        return out
    # host form of the same steps (torch ops)
    A = l2[iu, ju].reshape(-1, v * v)
    x, y = A[:, cd], A[:, dc]
    Apm = torch.stack(((x + y) * wd, x - y))
    if precision == "fp64":
        Xpm = torch.stack((Apm[0] @ W[0], Apm[1] @ W[1]))
    elif precision == "fp32":
        Xpm = torch.bmm(Apm.float(), W).double()
    else:
        raise ValueError(f"precision {precision!r} needs the device")
    X = ((Xpm[0][:, pidx] + Xpm[1][:, pidx] * sign) * 0.5).reshape(-1, v, v)
    out = torch.empty_like(l2)
    out[iu, ju] = X * scale
    out[ju, iu] = X.transpose(1, 2) * scale
    return out
