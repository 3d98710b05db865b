"""Batched GEMM on the MFMA units (HIP kernels) for float64 tensors."""
import ctypes

import torch

from . import _lib

MODES = {"fp32": 0, "bf16x3": 1}


def bmm(A, B, precision):
    """(nb, m, k) x (nb, k, n) -> (nb, m, n) float64, computed at `precision`."""
    if precision not in MODES:
        raise ValueError(f"unknown precision {precision!r}")
    A = A.to(torch.float64).contiguous()
    B = B.to(torch.float64).contiguous()
    nb, m, k = A.shape
    nb2, k2, n = B.shape
    if nb != nb2 or k != k2:
        raise ValueError("bmm shape mismatch")
    C = torch.empty((nb, m, n), dtype=torch.float64, device=A.device)
    if C.numel() == 0:
        return C
    if k == 0:
        return C.zero_()
    lib = _lib.load()
    # split very tall problems so one launch stays within the grid limits
    step = 65535 * 64
    for r0 in range(0, m, step):
        r1 = min(m, r0 + step)
        for b0 in range(0, nb, 65535):
            b1 = min(nb, b0 + 65535)
            _lib.check(lib.ccsd_gemm_mixed(
                MODES[precision], _lib.c_void_p(A[b0, r0].data_ptr()), _lib.c_void_p(B[b0].data_ptr()),
                _lib.c_void_p(C[b0, r0].data_ptr()), r1 - r0, n, k, k, n, n, m * k, k * n, m * n, b1 - b0,
                _lib.current_stream_ptr()), "ccsd_gemm_mixed")
    return C


class Split:
    """A float64 matrix batch stored as two bfloat16 planes (hi, lo), shape (nb, rows, cols)."""

    def __init__(self, hi, lo):
        self.hi, self.lo = hi, lo

    @property
    def shape(self):
        return self.hi.shape


def _permute(x, dst, hi, lo):
    """The library's permute kernel: x (float64 strided view, ndim <= 6) -> contiguous
    dst, or split-BF16 planes hi / lo."""
    nd = max(1, x.dim())
    sizes = (ctypes.c_longlong * nd)(*(list(x.shape) or [1]))
    strides = (ctypes.c_longlong * nd)(*(list(x.stride()) or [1]))
    _lib.check(_lib.load().ccsd_permute(_lib.ptr(x), nd, sizes, strides,
                                        _lib.ptr(dst) if dst is not None else None,
                                        _lib.ptr(hi) if hi is not None else None,
                                        _lib.ptr(lo) if lo is not None else None, _lib.current_stream_ptr()),
               "ccsd_permute")


def permute_copy(x):
    """Contiguous copy of a float64 strided view (any permutation; transposing copies
    through LDS tiles). CPU tensors and views of more than six dimensions use torch."""
    if not x.is_cuda or x.dtype != torch.float64 or x.dim() > 6 or x.dim() == 0:
        return x.contiguous()
    out = torch.empty(x.shape, dtype=torch.float64, device=x.device)
    if out.numel():
        _permute(x, out, None, None)
    return out


def combine(terms, out=None, beta=0):
    """out (+)= sum_t alpha_t * x_t for up to four (alpha, x_t) float64 views of one shape
    (ndim <= 6) on the library's combine kernel (transposed terms through LDS tiles);
    out (new contiguous when None, else any strided view not overlapping a source).
    CPU tensors use torch."""
    shape = tuple(terms[0][1].shape)
    if out is None:
        out = torch.empty(shape, dtype=torch.float64, device=terms[0][1].device)
        beta = 0
    if any(tuple(x.shape) != shape for _, x in terms) or tuple(out.shape) != shape:
        raise ValueError("combine needs views of one shape")
    nd, nt = len(shape), len(terms)
    if not out.is_cuda or nt > 4 or nd > 6 or nd == 0 or any(x.dtype != torch.float64 for _, x in terms):
        r = sum(float(a) * x for a, x in terms)
        return out.add_(r) if beta else out.copy_(r)
    if out.numel() == 0:
        return out
    sizes = (ctypes.c_longlong * nd)(*shape)
    dstr = (ctypes.c_longlong * nd)(*out.stride())
    srcs = (ctypes.c_void_p * nt)(*[x.data_ptr() for _, x in terms])
    sstr = (ctypes.c_longlong * (nt * nd))(*[st for _, x in terms for st in x.stride()])
    coef = (ctypes.c_double * nt)(*[float(a) for a, _ in terms])
    _lib.check(_lib.load().ccsd_combine(nd, sizes, _lib.ptr(out), dstr, nt, srcs, sstr, coef, int(beta),
                                        _lib.current_stream_ptr()), "ccsd_combine")
    return out


def split(x):
    """(nb, rows, cols) float64 (any strides) -> Split (contiguous planes)."""
    x = x.to(torch.float64)
    hi = torch.empty(x.shape, dtype=torch.bfloat16, device=x.device)
    lo = torch.empty_like(hi)
    if x.numel():
        _permute(x, None, hi, lo)
    return Split(hi, lo)


def gemv(A, x, y, alpha=1.0, beta=0):
    """y (+)= alpha * A @ x in FP64 on the library's GEMV kernels: A (m, k) any
    strides with one of them 1, x (k,) any stride, y (m,) contiguous. Returns y."""
    m, k = A.shape
    if A.stride(1) == 1 and m > 0:
        cols, lda = 0, A.stride(0) if m > 1 else k
    elif A.stride(0) == 1:
        cols, lda = 1, A.stride(1) if k > 1 else m
    else:
        A = A.contiguous()
        cols, lda = 0, k
    if m == 0:
        return y
    if cols == 0:
        S = max(1, min(-(-16384 // m), k // 1024))
    else:
        S = max(1, min(-(-262144 // m), k // 64))
    ws = torch.empty(S * m, dtype=torch.float64, device=y.device) if S > 1 else None
    _lib.check(_lib.load().ccsd_gemv(cols, _lib.ptr(A), m, k, max(lda, 1), _lib.ptr(x), x.stride(0) if k else 1,
                                     _lib.ptr(y), float(alpha), int(beta), _lib.ptr(ws) if ws is not None else None,
                                     S, _lib.current_stream_ptr()), "ccsd_gemv")
    return y


def sum_slices(ws, out, beta=0):
    """out (+)= ws.sum(0) for a contiguous (S, ...) workspace and a contiguous out of
    ws.shape[1:], by the library's fixed-order slice-sum kernels (ws is overwritten)."""
    if not (ws.is_contiguous() and out.is_contiguous()) or tuple(ws.shape[1:]) != tuple(out.shape):
        raise ValueError("sum_slices needs contiguous ws (S, ...) and out of shape ws.shape[1:]")
    _lib.check(_lib.load().ccsd_sum_slices(_lib.ptr(ws), _lib.ptr(out), out.numel(), ws.shape[0], int(beta),
                                           _lib.current_stream_ptr()), "ccsd_sum_slices")
    return out


def k_splits(tiles, k, want=1024, min_chunk=1024):
    """Number of K chunks for a GEMM with `tiles` output tiles: small outputs with a long
    contracted dimension launch too few workgroups to fill 256 CUs, so K is cut into
    chunks that run as a batch and are summed afterwards."""
    if tiles * 4 >= want or k < 2 * min_chunk:
        return 1
    return max(1, min(-(-want // tiles), k // min_chunk))


_KSPLIT_TUNED = {}


def tuned_ksplit(A, B, candidates=(1, 2, 3, 4), reps=2):
    """K split of gemm_nt(A, B) chosen by timing the candidates once per shape (HIP events,
    into a throw-away output) and cached; the k_splits heuristic when not timeable (under
    HIP-graph capture, or off the GPU). A product whose output tiles fill a non-integer
    number of rounds of the chip's resident blocks can gain from a split (the (H2O)8 l2.vvvv
    product, 7 x 181 tiles: 3.98 -> 3.56 ms at a split of 3), others lose the extra sum."""
    nb, m, k = A.hi.shape
    n = B.hi.shape[1]
    key = (nb, m, n, k)
    if key in _KSPLIT_TUNED:
        return _KSPLIT_TUNED[key]
    if not A.hi.is_cuda or torch.cuda.is_current_stream_capturing():
        return None
    out = torch.empty((nb, m, n), dtype=torch.float64, device=A.hi.device)
    stream = torch.cuda.current_stream()
    best, best_ms = None, float("inf")
    for ks in candidates:
        if ks > 1 and k < 2 * 1024 * ks:
            continue
        gemm_nt(A, B, out=out, ksplit=ks)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            gemm_nt(A, B, out=out, ksplit=ks)
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        if ms < best_ms:
            best, best_ms = ks, ms
    _KSPLIT_TUNED[key] = best
    return best


def _launch_nt(lib, A, B, C, a_off, b_off, m, n, kc, lda, ldb, sA, sB, sC, ldc, batch, beta):
    ptr = _lib.c_void_p
    e = A.hi.element_size()
    _lib.check(lib.ccsd_gemm_bf16x3_nt(
        ptr(A.hi.data_ptr() + a_off * e), ptr(A.lo.data_ptr() + a_off * e),
        ptr(B.hi.data_ptr() + b_off * e), ptr(B.lo.data_ptr() + b_off * e),
        ptr(C.data_ptr()), m, n, kc, lda, ldb, sA, sB, sC, ldc, batch, int(beta),
        _lib.current_stream_ptr()), "ccsd_gemm_bf16x3_nt")


def split_view(x, shape):
    """Split a float64 strided view (ndim <= 8, any strides, e.g. a permuted tensor)
    into bf16 planes laid out in the view's row-major order, reshaped to `shape`
    (nb, rows, cols): the operand's permutation and split in one pass."""
    if x.dtype != torch.float64 or x.dim() > 8:
        raise ValueError("split_view needs a float64 view of at most 8 dimensions")
    n = x.numel()
    hi = torch.empty(shape, dtype=torch.bfloat16, device=x.device)
    lo = torch.empty_like(hi)
    if hi.numel() != n:
        raise ValueError("shape does not match the view's size")
    if n and 1 <= x.dim() <= 6:
        _permute(x, None, hi, lo)        # planes in the view's row-major order
    elif n:
        nd = max(1, x.dim())
        sizes = (ctypes.c_longlong * nd)(*(list(x.shape) or [1]))
        strides = (ctypes.c_longlong * nd)(*(list(x.stride()) or [1]))
        _lib.check(_lib.load().ccsd_split_bf16_strided(_lib.ptr(x), _lib.ptr(hi), _lib.ptr(lo), nd, sizes, strides,
                                                        _lib.current_stream_ptr()), "ccsd_split_bf16_strided")
    return Split(hi, lo)


def gemm_nt_scatter(A, B, C, off_b, off_r, off_c, alpha=1.0, beta=1, ksplit=None):
    """C[off_b[b] + off_r[r] + off_c[c]] (+)= alpha * (A[b] @ B[b]^T)[r, c] with A: Split
    (nb, m, k), B: Split (nb, n, k), C any float64 device tensor whose storage the
    int64 offset tables address (relative to C.data_ptr())."""
    nb, m, k = A.shape
    nb2, n, k2 = B.shape
    if nb != nb2 or k != k2:
        raise ValueError("gemm_nt_scatter shape mismatch")
    if m == 0 or n == 0 or nb == 0:
        return C
    lib = _lib.load()
    ptr = _lib.c_void_p
    e = A.hi.element_size()
    if ksplit is None:
        ksplit = k_splits(nb * -(-m // 128) * -(-n // 128), k) if nb <= 8 else 1
    if ksplit > 1 and k > 0:
        kc = -(-(-(-k // ksplit)) // 64) * 64
        full, rest = divmod(k, kc)
        S = full + (rest > 0)
        ws = torch.empty((S, nb, m, n), dtype=torch.float64, device=C.device)
        for b in range(nb):
            if full:
                _launch_nt(lib, A, B, ws[0, b], b * m * k, b * n * k, m, n, kc, k, k, kc, kc, nb * m * n, n, full, 0)
            if rest:
                _launch_nt(lib, A, B, ws[full, b], b * m * k + full * kc, b * n * k + full * kc, m, n, rest,
                           k, k, 0, 0, 0, n, 1, 0)
        _lib.check(lib.ccsd_scatter_sum(_lib.ptr(ws), S, nb, m, n, _lib.ptr(C), _lib.ptr(off_b), _lib.ptr(off_r),
                                        _lib.ptr(off_c), float(alpha), int(beta), _lib.current_stream_ptr()),
                   "ccsd_scatter_sum")
        return C
    _lib.check(lib.ccsd_gemm_bf16x3_nt_scatter(
        ptr(A.hi.data_ptr()), ptr(A.lo.data_ptr()), ptr(B.hi.data_ptr()), ptr(B.lo.data_ptr()), _lib.ptr(C),
        m, n, k, k, k, m * k, n * k, _lib.ptr(off_b), _lib.ptr(off_r), _lib.ptr(off_c), nb, float(alpha), int(beta),
        _lib.current_stream_ptr()), "ccsd_gemm_bf16x3_nt_scatter")
    return C


def gemm_nt(A, B, out=None, beta=0, ksplit=None):
    """C[b] = A[b] @ B[b]^T (+ C if beta) with A: Split (nb, m, k), B: Split (nb, n, k) -> float64 (nb, m, n).

    `ksplit` chunks of K run as one batched launch into a workspace that is then
    summed (None: chosen from the output tile count, see k_splits)."""
    nb, m, k = A.shape
    nb2, n, k2 = B.shape
    if nb != nb2 or k != k2:
        raise ValueError("gemm_nt shape mismatch")
    if not (A.hi.is_contiguous() and A.lo.is_contiguous() and B.hi.is_contiguous() and B.lo.is_contiguous()):
        raise ValueError("gemm_nt needs contiguous planes")
    if out is None:
        out = torch.empty((nb, m, n), dtype=torch.float64, device=A.hi.device)
        beta = 0
    if out.numel() == 0:
        return out
    if k == 0:
        return out if beta else out.zero_()
    lib = _lib.load()
    if ksplit is None:
        ksplit = k_splits(nb * -(-m // 128) * -(-n // 128), k) if nb <= 8 else 1
    if ksplit > 1:
        kc = -(-(-(-k // ksplit)) // 64) * 64          # chunk length, a multiple of the K tile
        full, rest = divmod(k, kc)
        ws = torch.empty((full + (rest > 0), m, n), dtype=torch.float64, device=out.device)
        for b in range(nb):
            if full:
                _launch_nt(lib, A, B, ws, b * m * k, b * n * k, m, n, kc, k, k, kc, kc, m * n, n, full, 0)
            if rest:
                _launch_nt(lib, A, B, ws[full], b * m * k + full * kc, b * n * k + full * kc, m, n, rest,
                           k, k, 0, 0, 0, n, 1, 0)
            sum_slices(ws, out[b], beta)
        return out
    for b0 in range(0, nb, 65535):
        b1 = min(nb, b0 + 65535)
        _launch_nt(lib, A, B, out[b0], b0 * m * k, b0 * n * k, m, n, k, k, k, m * k, n * k, m * n, n, b1 - b0,
                   beta)
    return out
