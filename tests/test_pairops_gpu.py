"""Doubles-shaped lambda helpers (l1_terms, pair_combine, l2_finish) against torch FP64 einsums."""
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda")


def _ref_l1(X, y, c):
    eqs = ("ijab,jb->ia", "ijba,jb->ia", "jiab,jb->ia", "jiba,jb->ia")
    return sum(ci * torch.einsum(eq, X, y) for ci, eq in zip(c, eqs) if ci)


@pytest.mark.parametrize("o,v", [(5, 70), (3, 129), (1, 7)])
@pytest.mark.parametrize("coef", [(2.0, -1.0, 0.0, 0.0), (-2.0, 1.0, 1.0, -2.0), (2.0, 0.0, -1.0, 0.0),
                                  (0.0, 0.5, 0.0, 0.25)])
def test_l1_terms_matches_einsum(dev, o, v, coef):
    from ccsd_gpu_amd.pairops import l1_terms
    g = torch.Generator(device=dev).manual_seed(o * 1000 + v)
    X = torch.randn(o, o, v, v, dtype=torch.float64, device=dev, generator=g)
    y = torch.randn(v, o, dtype=torch.float64, device=dev, generator=g).T      # strided (o, v) view
    out0 = torch.randn(o, v, dtype=torch.float64, device=dev, generator=g)
    got = l1_terms(out0.clone(), X, y, *coef)
    ref = out0 + _ref_l1(X, y, coef)
    torch.cuda.synchronize()
    print(o, v, coef, float((got - ref).abs().max()), flush=True)
    assert torch.allclose(got, ref, rtol=1e-12, atol=1e-11)
    # deterministic: a second call gives the same bits
    assert torch.equal(l1_terms(out0.clone(), X, y, *coef), got)


def test_l1_terms_beyond_lds_limit(dev):
    from ccsd_gpu_amd.pairops import MAX_NVIR_L1, l1_terms
    o, v = 2, MAX_NVIR_L1 + 5
    g = torch.Generator(device=dev).manual_seed(7)
    X = torch.randn(o, o, v, v, dtype=torch.float64, device=dev, generator=g)
    y = torch.randn(o, v, dtype=torch.float64, device=dev, generator=g)
    coef = (2.0, -1.0, 0.5, -0.25)
    got = l1_terms(torch.zeros(o, v, dtype=torch.float64, device=dev), X, y, *coef)
    assert torch.allclose(got, _ref_l1(X, y, coef), rtol=1e-12, atol=1e-10)


@pytest.mark.parametrize("o,v", [(4, 45), (2, 64), (3, 97)])
def test_pair_combine_matches_permuted_sums(dev, o, v):
    from ccsd_gpu_amd.pairops import pair_combine
    g = torch.Generator(device=dev).manual_seed(o + v)
    R = lambda *s: torch.randn(*s, dtype=torch.float64, device=dev, generator=g)  # noqa: E731
    X = R(o, o, v, v)
    d = R(o, o, v, v)
    c = (1.0, -0.5, 0.25, 2.0)
    ref = (c[0] * X + c[1] * X.permute(0, 1, 3, 2) + c[2] * X.permute(1, 0, 2, 3) + c[3] * X.permute(1, 0, 3, 2)) * d
    got = pair_combine(X, *c, d=d)
    assert torch.allclose(got, ref, rtol=1e-14, atol=1e-13)
    # strided input (b stride 1), permuted destination [i][a][j][b], accumulate
    Xs = R(o, o, v, v).permute(1, 0, 2, 3)
    base = R(o, v, o, v)
    out = base.clone()
    pair_combine(Xs, c1=1.0, c2=-0.5, out=out.permute(0, 2, 1, 3), beta=1)
    ref2 = base + (Xs - 0.5 * Xs.permute(0, 1, 3, 2)).permute(0, 2, 1, 3)
    assert torch.allclose(out, ref2, rtol=1e-14, atol=1e-13)
    # the lambda update's final step: symmetrise and divide
    got3 = pair_combine(X, c1=1.0, c4=1.0, d=d)
    assert torch.allclose(got3, (X + X.permute(1, 0, 3, 2)) * d, rtol=1e-14, atol=1e-13)


@pytest.mark.parametrize("o,v", [(5, 45), (3, 64), (1, 7)])
def test_unpack_pairs_matches_scatter(dev, o, v):
    from ccsd_gpu_amd.pairops import unpack_pairs
    g = torch.Generator(device=dev).manual_seed(o * v)
    iu, ju = torch.triu_indices(o, o, device=dev)
    X = torch.randn(len(iu), v, v, dtype=torch.float64, device=dev, generator=g)
    ref = torch.empty(o, o, v, v, dtype=torch.float64, device=dev)
    ref[iu, ju] = X * 0.5
    ref[ju, iu] = X.transpose(1, 2) * 0.5
    diag = iu == ju
    ref[iu[diag], iu[diag]] = 0.25 * (X[diag] + X[diag].transpose(1, 2))
    got = unpack_pairs(X, iu, ju, o, scale=0.5)
    assert torch.allclose(got, ref, rtol=1e-15, atol=1e-15)


@pytest.mark.parametrize("o,v", [(4, 45), (3, 64), (2, 7)])
@pytest.mark.parametrize("weight", ["host", "device", "none"])
def test_l2_finish_matches_sum(dev, o, v, weight):
    """(X + X[j,i,b,a] + u[i,a] w[j,b] + u[j,b] w[i,a] + s S) d, with the weight s given on
    the host, as a device scalar (read when the kernel runs: a later value reaches the
    next call), or with no S at all."""
    from ccsd_gpu_amd.pairops import l2_finish
    g = torch.Generator(device=dev).manual_seed(o * 100 + v)
    R = lambda *s: torch.randn(*s, dtype=torch.float64, device=dev, generator=g)  # noqa: E731
    X, S, d = R(o, o, v, v), R(o, o, v, v), R(o, o, v, v)
    u, w = R(o, v), R(v, o).T          # w a strided view
    s = 0.75
    uw = u[:, None, :, None] * w[None, :, None, :]
    ref = X + X.permute(1, 0, 3, 2) + uw + uw.permute(1, 0, 3, 2)
    if weight != "none":
        ref = ref + s * S
    ref = ref * d
    if weight == "host":
        got = l2_finish(X, u, w, S, s, d)
    elif weight == "device":
        sd = torch.full((), -3.0, dtype=torch.float64, device=dev)
        got = l2_finish(X, u, w, S, sd, d)
        sd.fill_(s)                      # queued after the kernel: it read -3 ...
        ref_old = (X + X.permute(1, 0, 3, 2) + uw + uw.permute(1, 0, 3, 2) - 3.0 * S) * d
        assert torch.allclose(got, ref_old, rtol=1e-14, atol=1e-13)
        got = l2_finish(X, u, w, S, sd, d)   # ... and the next call reads 0.75
    else:
        got = l2_finish(X, u, w, None, 0.0, d)
    assert torch.allclose(got, ref, rtol=1e-14, atol=1e-13)
