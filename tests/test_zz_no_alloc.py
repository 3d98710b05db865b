"""ccsd_gemv_noalloc (slice partials in the library's device scratch, no allocation)
against ccsd_gemv (caller's workspace) through the C ABI: bit-identical results for
the same slice count, two calls in a row with the second smaller, both memory orders,
overwrite and accumulate; sizes beyond the scratch are rejected before any launch."""
import ctypes

import pytest
import torch

from ccsd_gpu_amd import _lib

SCRATCH = 1 << 19          # CCSD_GEMV_SCRATCH_DOUBLES


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _pair(cols, M, K, S, beta, seed):
    lib = _lib.load()
    gen = torch.Generator(device="cuda").manual_seed(seed)
    A = torch.randn((K, M) if cols else (M, K), dtype=torch.float64, device="cuda", generator=gen)
    x = torch.randn(K, dtype=torch.float64, device="cuda", generator=gen)
    y0 = torch.randn(M, dtype=torch.float64, device="cuda", generator=gen)
    lda = M if cols else K
    stream = _lib.current_stream_ptr()
    y_ref = y0.clone()
    ws = torch.empty(S * M, dtype=torch.float64, device="cuda")
    _lib.check(lib.ccsd_gemv(cols, _p(A), M, K, lda, _p(x), 1, _p(y_ref), 0.75, beta, _p(ws), S, stream),
               "ccsd_gemv")
    y_na = y0.clone()
    _lib.check(lib.ccsd_gemv_noalloc(cols, _p(A), M, K, lda, _p(x), 1, _p(y_na), 0.75, beta, S, stream),
               "ccsd_gemv_noalloc")
    torch.cuda.synchronize()
    want = 0.75 * ((A.T if cols else A) @ x) + (y0 if beta else 0)
    return y_ref, y_na, want


@pytest.mark.gpu
@pytest.mark.parametrize("cols", [0, 1])
@pytest.mark.parametrize("beta", [0, 1])
def test_noalloc_gemv_equals_workspace_gemv(gpu, cols, beta):
    print("noalloc gemv cols", cols, "beta", beta, flush=True)
    # first call fills most of the scratch, the second (smaller) reuses it
    for (M, K, S, seed) in [(SCRATCH // 8, 20000, 8, 1), (1000, 7000, 3, 2)]:
        assert S * M <= SCRATCH
        y_ref, y_na, want = _pair(cols, M, K, S, beta, seed + 10 * cols + 100 * beta)
        assert torch.equal(y_ref, y_na), (M, K, S)
        assert float((y_na - want).abs().max()) < 1e-10 * max(1.0, float(want.abs().max()))


@pytest.mark.gpu
def test_noalloc_gemv_rejects_sizes_beyond_scratch(gpu):
    lib = _lib.load()
    M, S = SCRATCH // 4 + 1, 4
    A = torch.empty((2, 2), dtype=torch.float64, device="cuda")     # never read: rejected first
    rc = lib.ccsd_gemv_noalloc(0, _p(A), M, 4096, 4096, _p(A), 1, _p(A), 1.0, 0, S, _lib.current_stream_ptr())
    assert rc == -2 and b"scratch" in lib.ccsd_last_error()
    # one slice needs no scratch: any size is accepted
    M, K = 3, 5
    A = torch.randn((M, K), dtype=torch.float64, device="cuda")
    x = torch.randn(K, dtype=torch.float64, device="cuda")
    y = torch.empty(M, dtype=torch.float64, device="cuda")
    _lib.check(lib.ccsd_gemv_noalloc(0, _p(A), M, K, K, _p(x), 1, _p(y), 1.0, 0, 1, _lib.current_stream_ptr()),
               "ccsd_gemv_noalloc")
    torch.cuda.synchronize()
    assert float((y - A @ x).abs().max()) < 1e-13
