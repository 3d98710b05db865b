"""DIIS: the piecewise push (vector written straight into the history buffer) against
the whole-vector update, and the change norm it returns."""
import numpy as np
import pytest
import torch

from ccsd_gpu_amd.cc import DIIS


def _seq(seed, n_it=9, n1=6, n2=20):
    g = torch.Generator().manual_seed(seed)
    return [(torch.randn(2, 3, generator=g, dtype=torch.float64), torch.randn(2, 2, 5, generator=g,
                                                                                 dtype=torch.float64))
            for _ in range(n_it)]


def test_push_matches_update_and_reports_change_norm():
    for rhs in (None, torch.randn(26, dtype=torch.float64, generator=torch.Generator().manual_seed(9))):
        a, b = DIIS(4), DIIS(4)
        start = torch.randn(26, dtype=torch.float64, generator=torch.Generator().manual_seed(1))
        prev = start
        for it, (p1, p2) in enumerate(_seq(3)):
            x = torch.cat([p1.reshape(-1), p2.reshape(-1)])
            if rhs is not None:
                x = x + rhs
            want_norm = float(torch.linalg.norm(x - prev))
            ref = a.update(x)
            got, nrm = b.push((p1, p2), rhs, start=start if it == 0 else None)
            assert torch.allclose(got, ref, rtol=0, atol=1e-12)
            assert abs(nrm - want_norm) <= 1e-10 * max(1.0, want_norm)
            prev = ref


def test_push_first_call_without_start_has_no_norm():
    d = DIIS(3)
    p = (torch.ones(2, dtype=torch.float64), torch.zeros(3, dtype=torch.float64))
    v, nrm = d.push(p)
    assert nrm is None and np.allclose(v.numpy(), [1, 1, 0, 0, 0])


def test_input_combination_residual_is_res_norm_for_an_affine_map():
    """For x -> F(x) = A x + b, DIIS's input combination has the residual F(x) - x of
    norm res_norm (computed from the overlap matrix alone), here checked by applying F."""
    g = torch.Generator().manual_seed(5)
    n = 40
    q, _ = torch.linalg.qr(torch.randn(n, n, generator=g, dtype=torch.float64))
    A = q @ torch.diag(torch.linspace(-0.8, 0.9, n, dtype=torch.float64)) @ q.T
    b = torch.randn(n, generator=g, dtype=torch.float64)
    F = lambda x: A @ x + b  # noqa: E731
    for space in (3, 6):
        d = DIIS(space)
        x = torch.zeros(n, dtype=torch.float64)
        checked = 0
        for _ in range(12):
            y, _ = d.push((F(x)[:5], F(x)[5:]))
            if d.res_norm is not None:
                xb = d.input_combination()
                r = float(torch.linalg.norm(F(xb) - xb))
                assert abs(r - d.res_norm) <= 1e-9 * max(r, 1e-12) + 1e-14, (space, r, d.res_norm)
                assert d.c_abs >= 1.0 - 1e-12
                checked += 1
            x = y
        assert checked >= 10


@pytest.mark.gpu
@pytest.mark.parametrize("space", [3, 6])
def test_fp32_history_matches_fp64_on_device(space):
    """The FP32-history DIIS (library push / combine passes) against the FP64 one on the
    same device vectors (an affine map with a right-hand side, as the correction solves
    run it): extrapolations, change norms, DIIS residual norms and input combinations
    agree to FP32 rounding of the stored vectors; a second run repeats it bit for bit."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    g = torch.Generator(device="cuda").manual_seed(11)
    n1, n2 = 37, 5005
    n = n1 + n2
    A = torch.randn(n, dtype=torch.float64, device="cuda", generator=g) * 0.4    # diagonal contraction
    b = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    rhs = torch.randn(n, dtype=torch.float64, device="cuda", generator=g) * 0.1

    def run(store):
        d = DIIS(space, store=store)
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        outs = []
        for it in range(10):
            y = A * x + b
            v, nrm = d.push((y[:n1], y[n1:].reshape(7, -1)), rhs, start=x if it == 0 else None)
            comb = d.input_combination() if d.res_norm is not None else None
            outs.append((v.clone(), nrm, d.res_norm, None if comb is None else comb.clone()))
            x = v
        return outs

    ref, got, again = run(None), run("fp32"), run("fp32")
    for (v0, n0, r0, c0), (v1, n1_, r1, c1), (v2, n2_, r2, c2) in zip(ref, got, again):
        scale = float(v0.abs().max())
        assert float((v1 - v0).abs().max()) <= 1e-5 * scale
        assert abs(n1_ - n0) <= 1e-5 * n0
        if r0 is not None:
            assert abs(r1 - r0) <= 1e-4 * max(r0, 1e-6)
            assert float((c1 - c0).abs().max()) <= 1e-5 * scale
        assert torch.equal(v1, v2) and n1_ == n2_
