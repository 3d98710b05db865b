"""Doubles-shaped (o, o, v, v) helpers of the lambda update on the HIP library.

    l1_terms(out, X, y, c1, c2, c3, c4)
        out[i,a] += sum_{j,b} (c1 X[i,j,a,b] + c2 X[i,j,b,a] + c3 X[j,i,a,b] + c4 X[j,i,b,a]) y[j,b]
    pair_combine(X, c1, c2, c3, c4, d=None, out=None, beta=0)
        out[i,j,a,b] (+)= (c1 X[i,j,a,b] + c2 X[i,j,b,a] + c3 X[j,i,a,b] + c4 X[j,i,b,a]) * d[i,j,a,b]
    l2_finish(X, u, w, S, s, d)
        out[i,j,a,b] = (X[i,j,a,b] + X[j,i,b,a] + u[i,a] w[j,b] + u[j,b] w[i,a] + s S[i,j,a,b]) * d[i,j,a,b]

Both read X once (l1_terms forms the four matrix-vector products of every v x v
block in one pass; pair_combine stages the transposed blocks through LDS), in
FP64 at every solve precision. They replace the einsum pairs and permuted
elementwise forms of ccsd_gpu/gpu_ccsd_lambda.py:244-257, 288-334.
"""
import ctypes

import torch

from . import _lib

MAX_NVIR_L1 = 682     # LDS limit of the l1 block kernel (12 v doubles <= 64 KiB)


def _check_x(X, contiguous=True):
    if X.dim() != 4 or X.dtype != torch.float64 or not X.is_cuda or (contiguous and not X.is_contiguous()):
        raise ValueError("expected a float64 (o, o, v, v) device tensor" + (" (contiguous)" if contiguous else ""))
    o, o2, v, v2 = X.shape
    if o != o2 or v != v2:
        raise ValueError("expected an (o, o, v, v) tensor")
    return o, v


def l1_terms(out, X, y, c1=0.0, c2=0.0, c3=0.0, c4=0.0):
    """In place on out (o, v) contiguous float64; y (o, v) any strides. Returns out."""
    o, v = _check_x(X)
    if tuple(out.shape) != (o, v) or not out.is_contiguous() or out.dtype != torch.float64:
        raise ValueError("out must be a contiguous float64 (o, v) tensor")
    if tuple(y.shape) != (o, v) or y.dtype != torch.float64:
        raise ValueError("y must be a float64 (o, v) tensor")
    if v > MAX_NVIR_L1:
        # beyond the kernel's LDS: the same sum as four einsums on the device
        for c, eq in ((c1, "ijab,jb->ia"), (c2, "ijba,jb->ia"), (c3, "jiab,jb->ia"), (c4, "jiba,jb->ia")):
            if c:
                out.add_(torch.einsum(eq, X, y), alpha=c)
        return out
    ws = torch.empty((2, o, o, v), dtype=torch.float64, device=X.device)
    _lib.check(_lib.load().ccsd_l1_terms(_lib.ptr(X), o, v, _lib.ptr(y), y.stride(0), y.stride(1), float(c1),
                                         float(c2), float(c3), float(c4), _lib.ptr(ws), _lib.ptr(out),
                                         _lib.current_stream_ptr()), "ccsd_l1_terms")
    return out


def _strides(t):
    return (ctypes.c_longlong * 4)(*t.stride())


def pair_combine(X, c1=0.0, c2=0.0, c3=0.0, c4=0.0, d=None, out=None, beta=0):
    """X any strided (o, o, v, v) view; returns out (new contiguous when None;
    otherwise any strided (o, o, v, v) view indexed [i, j, a, b] that does not
    overlap X). Coalesced when the last strides of X and out are 1."""
    o, v = _check_x(X, contiguous=False)
    if d is not None and (tuple(d.shape) != tuple(X.shape) or not d.is_contiguous() or d.dtype != torch.float64):
        raise ValueError("d must be contiguous and shaped like X")
    if out is None:
        out = torch.empty_like(X)
        beta = 0
    if tuple(out.shape) != tuple(X.shape) or out.dtype != torch.float64:
        raise ValueError("out must be a float64 view shaped like X")
    _lib.check(_lib.load().ccsd_pair_combine(_lib.ptr(X), _strides(X), o, v, float(c1), float(c2), float(c3),
                                             float(c4), _lib.ptr(d) if d is not None else None, _lib.ptr(out),
                                             _strides(out), int(beta), _lib.current_stream_ptr()),
               "ccsd_pair_combine")
    return out


def unpack_pairs(X, iu, ju, o, scale=1.0):
    """X (npair, v, v) blocks of the pairs (iu[p], ju[p]), iu <= ju -> dense (o, o, v, v):
    out[i,j] = scale X[p], out[j,i] = scale X[p]^T, diagonal blocks symmetrised."""
    npair, v, v2 = X.shape
    if v != v2 or not X.is_contiguous() or X.dtype != torch.float64 or not X.is_cuda:
        raise ValueError("X must be a contiguous float64 (npair, v, v) device tensor")
    iu = iu.to(torch.int64).contiguous()
    ju = ju.to(torch.int64).contiguous()
    out = torch.empty((o, o, v, v), dtype=torch.float64, device=X.device)
    _lib.check(_lib.load().ccsd_unpack_pairs(_lib.ptr(X), _lib.ptr(iu), _lib.ptr(ju), o, v, npair, float(scale),
                                             _lib.ptr(out), _lib.current_stream_ptr()), "ccsd_unpack_pairs")
    return out


def l2_finish(X, u, w, S, s, d):
    """The closing step of the lambda update (ccsd_gpu/gpu_ccsd_lambda.py:327-334 with
    :259 and :295 folded in), as a new contiguous tensor:
        out[i,j,a,b] = (X[i,j,a,b] + X[j,i,b,a] + u[i,a] w[j,b] + u[j,b] w[i,a] + s S[i,j,a,b]) d[i,j,a,b]
    X, S (or None), d contiguous float64 (o, o, v, v); u, w (o, v); s a float or a 0-d
    float64 device tensor (read when the kernel runs)."""
    o, v = _check_x(X)
    for name, t in (("d", d), ("S", S)):
        if t is not None and (tuple(t.shape) != tuple(X.shape) or not t.is_contiguous() or t.dtype != torch.float64):
            raise ValueError(f"{name} must be contiguous and shaped like X")
    u = u.contiguous()
    w = w.contiguous()
    if tuple(u.shape) != (o, v) or tuple(w.shape) != (o, v) or u.dtype != torch.float64 or w.dtype != torch.float64:
        raise ValueError("u and w must be float64 (o, v) tensors")
    if torch.is_tensor(s):
        if s.numel() != 1 or s.dtype != torch.float64 or not s.is_cuda:
            raise ValueError("a tensor weight must be a 0-d float64 device tensor")
        sdev, shost = _lib.ptr(s), 0.0
    else:
        sdev, shost = None, float(s)
    out = torch.empty_like(X)
    _lib.check(_lib.load().ccsd_l2_finish(_lib.ptr(X), o, v, _lib.ptr(u), _lib.ptr(w),
                                          _lib.ptr(S) if S is not None else None, sdev, shost, _lib.ptr(d),
                                          _lib.ptr(out), _lib.current_stream_ptr()), "ccsd_l2_finish")
    return out
