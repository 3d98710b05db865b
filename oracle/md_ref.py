"""McMurchie-Davidson brute force for ERIs and their d/dr1 derivative.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

An algorithm independent of the Rys quadrature used by the reference
(ccsd_gpu/cuda/int2e_ip1_module.py) and by the C oracle: Hermite expansion
coefficients E_t^{ij} and Hermite Coulomb integrals R_{tuv}, with the Boys
function in mpmath (30 digits). Pure-Python loops, so only for tiny inputs
(a few shells, l <= 2).

  (ab|cd) = 2 pi^(5/2) / (p q sqrt(p+q))
            sum_{tuv} E^{ab}_{tuv} sum_{xyz} (-1)^(x+y+z) E^{cd}_{xyz} R_{t+x,u+y,v+z}(alpha, PQ)

int2e_ip1 is d/dr of the first function: for x^ax e^{-a (x-Ax)^2} that is
-2a x^(ax+1)... - i.e.  -2a |ax+1> + ax |ax-1>  (PySCF convention).
Normalization follows the Mole conventions (radial norm in the stored
coefficients, 1/(2 sqrt pi) and sqrt(3/(4 pi)) for s and p shells).
"""
import math
from functools import lru_cache

import mpmath as mp
import numpy as np

from ccsd_gpu_amd.gto.mole import cart_components, ncart

mp.mp.dps = 30


def boys(n, T):
    return float(mp.hyp1f1(n + 0.5, n + 1.5, -T) / (2 * n + 1))


def hermite_E(i, j, t, Qx, a, b):
    """E_t^{ij} for a 1-D Gaussian product, Qx = A - B (Helgaker recursion)."""
    p = a + b
    q = a * b / p
    if t < 0 or t > i + j:
        return 0.0
    if i == j == t == 0:
        return math.exp(-q * Qx * Qx)
    if j == 0:
        return (1.0 / (2 * p)) * hermite_E(i - 1, j, t - 1, Qx, a, b) \
            - (q * Qx / a) * hermite_E(i - 1, j, t, Qx, a, b) \
            + (t + 1) * hermite_E(i - 1, j, t + 1, Qx, a, b)
    return (1.0 / (2 * p)) * hermite_E(i, j - 1, t - 1, Qx, a, b) \
        + (q * Qx / b) * hermite_E(i, j - 1, t, Qx, a, b) \
        + (t + 1) * hermite_E(i, j - 1, t + 1, Qx, a, b)


def hermite_R(t, u, v, n, p, PCx, PCy, PCz, RPC, cache):
    key = (t, u, v, n)
    if key in cache:
        return cache[key]
    T = p * RPC * RPC
    if t == u == v == 0:
        val = (-2 * p) ** n * boys(n, T)
    elif t == u == 0:
        val = (v - 1) * hermite_R(t, u, v - 2, n + 1, p, PCx, PCy, PCz, RPC, cache) if v > 1 else 0.0
        val += PCz * hermite_R(t, u, v - 1, n + 1, p, PCx, PCy, PCz, RPC, cache)
    elif t == 0:
        val = (u - 1) * hermite_R(t, u - 2, v, n + 1, p, PCx, PCy, PCz, RPC, cache) if u > 1 else 0.0
        val += PCy * hermite_R(t, u - 1, v, n + 1, p, PCx, PCy, PCz, RPC, cache)
    else:
        val = (t - 1) * hermite_R(t - 2, u, v, n + 1, p, PCx, PCy, PCz, RPC, cache) if t > 1 else 0.0
        val += PCx * hermite_R(t - 1, u, v, n + 1, p, PCx, PCy, PCz, RPC, cache)
    cache[key] = val
    return val


def prim_eri(la, A, a, lb, B, b, lc, C, c, ld, D, d):
    """(ab|cd) of unnormalized primitive Cartesian Gaussians, l = (lx,ly,lz)."""
    p = a + b
    q = c + d
    alpha = p * q / (p + q)
    P = (a * np.asarray(A) + b * np.asarray(B)) / p
    Q = (c * np.asarray(C) + d * np.asarray(D)) / q
    PQ = P - Q
    RPQ = float(np.linalg.norm(PQ))
    cache = {}
    val = 0.0
    Ex = [hermite_E(la[0], lb[0], t, A[0] - B[0], a, b) for t in range(la[0] + lb[0] + 1)]
    Ey = [hermite_E(la[1], lb[1], t, A[1] - B[1], a, b) for t in range(la[1] + lb[1] + 1)]
    Ez = [hermite_E(la[2], lb[2], t, A[2] - B[2], a, b) for t in range(la[2] + lb[2] + 1)]
    Fx = [hermite_E(lc[0], ld[0], t, C[0] - D[0], c, d) for t in range(lc[0] + ld[0] + 1)]
    Fy = [hermite_E(lc[1], ld[1], t, C[1] - D[1], c, d) for t in range(lc[1] + ld[1] + 1)]
    Fz = [hermite_E(lc[2], ld[2], t, C[2] - D[2], c, d) for t in range(lc[2] + ld[2] + 1)]
    for t, ex in enumerate(Ex):
        for u, ey in enumerate(Ey):
            for v, ez in enumerate(Ez):
                e1 = ex * ey * ez
                if e1 == 0.0:
                    continue
                for tau, fx in enumerate(Fx):
                    for nu, fy in enumerate(Fy):
                        for phi, fz in enumerate(Fz):
                            val += e1 * fx * fy * fz * (-1) ** (tau + nu + phi) * hermite_R(
                                t + tau, u + nu, v + phi, 0, alpha, PQ[0], PQ[1], PQ[2], RPQ, cache)
    return 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * val


def _shell_norm(l):
    if l == 0:
        return 0.28209479177387814347
    if l == 1:
        return 0.48860251190291992159
    return 1.0


def shell_quartet(mol, shls, deriv=0):
    """Cartesian block for one shell quartet, all contraction columns.

    Returns array (ncomp, nci*nfi, ncj*nfj, nck*nfk, ncl*nfl) with the AO
    order used by the s2kl layout (contraction column major, then component).
    """
    data = []
    for s in shls:
        l = mol.bas_angular(s)
        data.append((l, mol.atom_coords()[mol.bas_atom(s)], mol.bas_exp(s), mol.bas_ctr_coeff(s)))
    dims = [ncart(l) * c.shape[1] for l, _, _, c in data]
    ncomp = 3 if deriv else 1
    out = np.zeros([ncomp] + dims)
    (li, A, ea, ca), (lj, B, eb, cb), (lk, C, ec, cc), (ll, D, ed, cd) = data
    norm = _shell_norm(li) * _shell_norm(lj) * _shell_norm(lk) * _shell_norm(ll)
    comps = [cart_components(l) for l in (li, lj, lk, ll)]
    for ip, a in enumerate(ea):
        for jp, b in enumerate(eb):
            for kp, c in enumerate(ec):
                for lp, d in enumerate(ed):
                    blk = np.zeros((ncomp, len(comps[0]), len(comps[1]), len(comps[2]), len(comps[3])))
                    for fi, ta in enumerate(comps[0]):
                        for fj, tb in enumerate(comps[1]):
                            for fk, tc in enumerate(comps[2]):
                                for fl, td in enumerate(comps[3]):
                                    if not deriv:
                                        blk[0, fi, fj, fk, fl] = prim_eri(ta, A, a, tb, B, b, tc, C, c, td, D, d)
                                        continue
                                    for x in range(3):
                                        up = list(ta)
                                        up[x] += 1
                                        v = -2 * a * prim_eri(tuple(up), A, a, tb, B, b, tc, C, c, td, D, d)
                                        if ta[x] > 0:
                                            dn = list(ta)
                                            dn[x] -= 1
                                            v += ta[x] * prim_eri(tuple(dn), A, a, tb, B, b, tc, C, c, td, D, d)
                                        blk[x, fi, fj, fk, fl] = v
                    coef = np.einsum("i,j,k,l->ijkl", ca[ip], cb[jp], cc[kp], cd[lp]) * norm
                    # (ncomp, ci*nfi, cj*nfj, ck*nfk, cl*nfl) with contraction-major AO order
                    full = np.einsum("ijkl,cabxy->ciajbkxly", coef, blk)
                    sh = full.shape
                    out += full.reshape(ncomp, sh[1] * sh[2], sh[3] * sh[4], sh[5] * sh[6], sh[7] * sh[8])
    return out
