"""Permute-copy / split and GEMV kernels of the contraction planner against torch."""
import ctypes

import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda")


PERMS = [((7, 5, 33, 70), (0, 1, 2, 3)), ((7, 5, 33, 70), (0, 2, 1, 3)), ((7, 5, 33, 70), (0, 1, 3, 2)),
         ((7, 5, 33, 70), (3, 2, 1, 0)), ((6, 40, 3, 45), (1, 3, 0, 2)), ((3, 4, 5, 6, 7), (4, 0, 3, 1, 2)),
         ((130, 67), (1, 0)), ((1, 1, 9, 1), (3, 2, 1, 0))]


@pytest.mark.parametrize("shape,perm", PERMS)
def test_permute_copy_matches_torch(dev, shape, perm):
    from ccsd_gpu_amd import mfma
    g = torch.Generator(device=dev).manual_seed(sum(shape))
    x = torch.randn(shape, dtype=torch.float64, device=dev, generator=g)
    v = x.permute(perm)
    got = mfma.permute_copy(v)
    assert got.is_contiguous() and torch.equal(got, v.contiguous())


def _split_ref(lib, v):
    """Planes from the elementwise strided split kernel (reference for the tiled one)."""
    from ccsd_gpu_amd import _lib
    hi = torch.empty(v.shape, dtype=torch.bfloat16, device=v.device)
    lo = torch.empty_like(hi)
    nd = v.dim()
    sizes = (ctypes.c_longlong * nd)(*v.shape)
    strides = (ctypes.c_longlong * nd)(*v.stride())
    _lib.check(lib.ccsd_split_bf16_strided(_lib.ptr(v), _lib.ptr(hi), _lib.ptr(lo), nd, sizes, strides,
                                           _lib.current_stream_ptr()), "split")
    return hi, lo


@pytest.mark.parametrize("shape,perm", PERMS[:6])
def test_split_through_permute_kernel_is_bitwise_the_split(dev, shape, perm):
    from ccsd_gpu_amd import _lib, mfma
    g = torch.Generator(device=dev).manual_seed(len(shape) + sum(perm))
    v = torch.randn(shape, dtype=torch.float64, device=dev, generator=g).permute(perm)
    n = v.numel()
    s = mfma.split_view(v, (1, n // v.shape[-1], v.shape[-1]))
    hi, lo = _split_ref(_lib.load(), v)
    assert torch.equal(s.hi.reshape(-1).view(torch.int16), hi.reshape(-1).view(torch.int16))
    assert torch.equal(s.lo.reshape(-1).view(torch.int16), lo.reshape(-1).view(torch.int16))
    # hi + lo carries ~16 bits of x
    x = v.contiguous().reshape(-1)
    rec = s.hi.reshape(-1).double() + s.lo.reshape(-1).double()
    assert float(((rec - x).abs() / x.abs().clamp_min(1e-30)).max()) < 2 ** -15


@pytest.mark.parametrize("m,k", [(6080, 300), (3, 23104), (200, 1), (1, 5000), (777, 4099)])
@pytest.mark.parametrize("layout", ["rows", "cols"])
def test_gemv_matches_torch(dev, m, k, layout):
    from ccsd_gpu_amd import mfma
    g = torch.Generator(device=dev).manual_seed(m + k)
    if layout == "rows":
        A = torch.randn(m, k, dtype=torch.float64, device=dev, generator=g)
    else:
        A = torch.randn(k, m, dtype=torch.float64, device=dev, generator=g).T
    x = torch.randn(2 * k, dtype=torch.float64, device=dev, generator=g)[::2]      # strided vector
    y0 = torch.randn(m, dtype=torch.float64, device=dev, generator=g)
    ref = A @ x
    got = mfma.gemv(A, x, torch.empty(m, dtype=torch.float64, device=dev))
    assert torch.allclose(got, ref, rtol=1e-12, atol=1e-10)
    got2 = mfma.gemv(A, x, y0.clone(), alpha=-0.5, beta=1)
    assert torch.allclose(got2, y0 - 0.5 * ref, rtol=1e-12, atol=1e-10)
    assert torch.equal(mfma.gemv(A, x, torch.empty(m, dtype=torch.float64, device=dev)), got)   # deterministic


@pytest.mark.parametrize("S,n", [(1, 5), (16, 3), (17, 1), (16384, 1), (497, 6080), (451, 1792), (3, 100000),
                                 (223, 23104)])
def test_sum_slices_matches_torch(dev, S, n):
    """Fixed-order slice sum (grouped in place for many slices over few outputs)."""
    from ccsd_gpu_amd import mfma
    g = torch.Generator(device=dev).manual_seed(S * 7 + n)
    ws = torch.randn(S, n, dtype=torch.float64, device=dev, generator=g)
    y0 = torch.randn(n, dtype=torch.float64, device=dev, generator=g)
    ref = ws.sum(0)
    got = mfma.sum_slices(ws.clone(), torch.empty(n, dtype=torch.float64, device=dev))
    assert torch.allclose(got, ref, rtol=1e-12, atol=1e-11 * S ** 0.5)
    got2 = mfma.sum_slices(ws.clone(), y0.clone(), beta=1)
    assert torch.allclose(got2, y0 + ref, rtol=1e-12, atol=1e-11 * S ** 0.5)
    assert torch.equal(mfma.sum_slices(ws.clone(), torch.empty(n, dtype=torch.float64, device=dev)), got)


@pytest.mark.parametrize("eq,shapes", [("iajb,bj->ai", [(9, 31, 9, 31), (31, 9)]),
                                       ("jb,aijb->ia", [(9, 31), (31, 9, 9, 31)]),
                                       ("iabc,bc->ia", [(4, 30, 30, 30), (30, 30)]),
                                       ("ibca,bc->ia", [(4, 30, 30, 30), (30, 30)])])
def test_planner_matrix_vector_contractions(dev, eq, shapes):
    from ccsd_gpu_amd.precision import contract
    g = torch.Generator(device=dev).manual_seed(len(eq))
    ops = [torch.randn(s, dtype=torch.float64, device=dev, generator=g) for s in shapes]
    assert torch.allclose(contract(eq, *ops), torch.einsum(eq, *ops), rtol=1e-12, atol=1e-11)


COMBOS = [((6, 40, 5, 45), [(2.0, (0, 1, 2, 3)), (-1.0, (0, 1, 3, 2))]),
          ((6, 40, 5, 45), [(1.0, (0, 2, 1, 3)), (-0.5, (0, 3, 2, 1)), (0.25, (1, 0, 2, 3)), (3.0, (0, 1, 2, 3))]),
          ((33, 33, 33, 33), [(0.5, (0, 1, 2, 3)), (0.5, (1, 0, 2, 3)), (0.5, (0, 1, 3, 2)), (0.5, (1, 0, 3, 2))]),
          ((4, 5, 6, 7, 8), [(1.0, (4, 3, 2, 1, 0)), (2.0, (0, 1, 2, 3, 4))]),
          ((9, 70), [(1.0, (1, 0))])]


@pytest.mark.parametrize("shape,terms", COMBOS)
def test_combine_matches_torch(dev, shape, terms):
    from ccsd_gpu_amd import mfma
    g = torch.Generator(device=dev).manual_seed(len(shape) * 7 + len(terms))
    base = torch.randn([shape[i] for i in range(len(shape))], dtype=torch.float64, device=dev, generator=g)
    # every term is a permuted view of one square-compatible source of the output shape
    srcs = []
    for a, perm in terms:
        inv = [perm.index(i) for i in range(len(perm))]
        x = torch.randn([shape[p] for p in inv], dtype=torch.float64, device=dev, generator=g)
        srcs.append((a, x.permute(perm)))
    ref = sum(a * x for a, x in srcs)
    got = mfma.combine(srcs)
    assert torch.allclose(got, ref, rtol=1e-14, atol=1e-13)
    # strided destination (a block of a larger tensor), accumulate
    big = torch.randn([s + 2 for s in shape], dtype=torch.float64, device=dev, generator=g)
    sl = tuple(slice(1, 1 + s) for s in shape)
    before = big.clone()
    mfma.combine(srcs, out=big[sl], beta=1)
    assert torch.allclose(big[sl], before[sl] + ref, rtol=1e-14, atol=1e-13)
    mask = torch.ones_like(big, dtype=torch.bool)
    mask[sl] = False
    assert torch.equal(big[mask], before[mask])          # nothing outside the view is touched
    del base


def test_pack_pairs_sym_gpu_matches_cpu(dev):
    from ccsd_gpu_amd.gpu_ccsd_grad import _pack_pairs_sym
    g = torch.Generator().manual_seed(8)
    t = torch.randn(37, 37, 19, 19, dtype=torch.float64, generator=g)
    ref = _pack_pairs_sym(t)                             # torch on the CPU
    got = _pack_pairs_sym(t.to(dev))
    assert torch.allclose(got.cpu(), ref, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("n,m", [(37, 19), (5, 70), (1, 1)])
def test_repack_pairs_gpu_matches_cpu(dev, n, m):
    from ccsd_gpu_amd.gpu_ccsd_grad import _repack_pairs
    g = torch.Generator().manual_seed(n * 100 + m)
    z = torch.randn(n, n, m * (m + 1) // 2, dtype=torch.float64, generator=g)
    ref = _repack_pairs(z, m)                             # torch on the CPU
    got = _repack_pairs(z.to(dev), m)
    assert torch.equal(got.cpu(), ref)
