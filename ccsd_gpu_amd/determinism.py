"""Bitwise-reproducible runs.

The library's own kernels sum in fixed orders: the fused int2e_ip1 x Gamma pass adds
its per-wave parts through exact fixed-point integer atomics, the one-electron kernels
give every output element one owner thread, the split-K GEMM partials and the l1
terms are reduced in a fixed order. What remains order-dependent are library calls
made through torch (the vendor BLAS may split a GEMM's K across workgroups that add
with float atomics). Inside `deterministic()` torch selects its deterministic
implementations, which makes a whole gradient (integrals, RHF, CCSD, lambda, dE/dR)
repeat bit for bit on the same device and kernel choices.
"""
from contextlib import contextmanager

import torch


@contextmanager
def deterministic(enabled=True):
    """Run the block with torch's deterministic algorithms on (operations without a
    deterministic form warn instead of raising), restoring the previous setting."""
    prev = torch.are_deterministic_algorithms_enabled()
    prev_warn = torch.is_deterministic_algorithms_warn_only_enabled()
    torch.use_deterministic_algorithms(enabled, warn_only=True)
    try:
        yield
    finally:
        torch.use_deterministic_algorithms(prev, warn_only=prev_warn)
