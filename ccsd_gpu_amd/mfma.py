"""Batched GEMM on the MFMA units (HIP kernel gemm_mixed_kernel) for float64 tensors."""
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


def split(x):
    """(nb, rows, cols) float64 (any strides; made contiguous) -> Split."""
    x = x.to(torch.float64).contiguous()
    hi = torch.empty(x.shape, dtype=torch.bfloat16, device=x.device)
    lo = torch.empty_like(hi)
    n = x.numel()
    if n:
        cols = x.shape[-1]
        rows = n // cols
        _lib.check(_lib.load().ccsd_split_bf16(_lib.ptr(x), _lib.ptr(hi), _lib.ptr(lo), rows, cols, cols,
                                                _lib.current_stream_ptr()), "ccsd_split_bf16")
    return Split(hi, lo)


def gemm_nt(A, B, out=None, beta=0):
    """C[b] = A[b] @ B[b]^T (+ C if beta) with A: Split (nb, m, k), B: Split (nb, n, k) -> float64 (nb, m, n)."""
    nb, m, k = A.shape
    nb2, n, k2 = B.shape
    if nb != nb2 or k != k2:
        raise ValueError("gemm_nt shape mismatch")
    if out is None:
        out = torch.empty((nb, m, n), dtype=torch.float64, device=A.hi.device)
        beta = 0
    if out.numel() == 0:
        return out
    if k == 0:
        return out if beta else out.zero_()
    lib = _lib.load()
    for b0 in range(0, nb, 65535):
        b1 = min(nb, b0 + 65535)
        _lib.check(lib.ccsd_gemm_bf16x3_nt(
            _lib.c_void_p(A.hi[b0].data_ptr()), _lib.c_void_p(A.lo[b0].data_ptr()),
            _lib.c_void_p(B.hi[b0].data_ptr()), _lib.c_void_p(B.lo[b0].data_ptr()),
            _lib.c_void_p(out[b0].data_ptr()), m, n, k, m * k, n * k, m * n, n, b1 - b0, int(beta),
            _lib.current_stream_ptr()), "ccsd_gemm_bf16x3_nt")
    return out
