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
