"""Algorithmic work of the Rys ERI kernels (for roofline fractions).

Per shell quartet of class (li, lj, lk, ll) with K primitive quartets, n roots
and c contraction-column combinations, FP64 operations are counted as

  F = K * [P + R(n) + n * (S + F_VRR + F_TRR + F_HRR + c * F_out * nf_cart)]

  P     = 40                      primitive-quartet setup (T, prefactor)
  R(n)  = 54 n                    2n Chebyshev series of degree 13 (Clenshaw)
  S     = 12                      per-root recurrence coefficients
  F_VRR = 3 * 4 * lij             three directions
  F_TRR = 3 * 7 * (lij+1) * lkl
  F_HRR = 3 * 2 * [(lkl+1) * sum_{j<lj}(lij-j) + sK * sum_{l<ll}(lkl-l)]
  F_out = 3 (Ix Iy Iz accumulate) for int2e, 18 for int2e_ip1 (three
          derivative factors and three triple products)

(lij = li+lj(+1 for ip1), lkl = lk+ll, sK = (li+1(+1))(lj+1)). The counts
restate the reference kernel's arithmetic (ccsd_gpu/cuda/int2e_ip1_module.py:409-629).
Quartet counts: this library evaluates contracted shells (general contractions
once per primitive set); the reference decontracts (ccsd_gpu/cuda/int2e_ip1_gpu.py:43-79)
and enumerates nbas_dc^2 * nbas_dc (nbas_dc+1)/2 quartets. K counts only the
primitive pairs with a1 a2/(a1+a2) R^2 <= 50 (the kernels skip the others: their
Gaussian-product factor is below e^-50).
"""
from collections import defaultdict


def _ncart(l):
    return (l + 1) * (l + 2) // 2


def class_flops(li, lj, lk, ll, deriv, nroots=None):
    """Flops per primitive quartet and per contraction combination term."""
    n = nroots if nroots is not None else (li + lj + lk + ll + deriv) // 2 + 1
    lij = li + lj + deriv
    lkl = lk + ll
    sk = (li + 1 + deriv) * (lj + 1)
    f_vrr = 12 * lij
    f_trr = 21 * (lij + 1) * lkl
    f_hrr = 6 * ((lkl + 1) * sum(lij - j for j in range(lj)) + sk * sum(lkl - l for l in range(ll)))
    nf = _ncart(li) * _ncart(lj) * _ncart(lk) * _ncart(ll)
    f_out = 18 if deriv else 3
    base = 40 + 54 * n + n * (12 + f_vrr + f_trr + f_hrr)
    per_combo = n * f_out * nf
    return base, per_combo


PRIM_EXP_CUT = 50.0   # primitive pairs with a1 a2/(a1+a2) R^2 above this are skipped by the kernels


def pair_prims(mol, cut=PRIM_EXP_CUT):
    """(nbas, nbas) number of primitive pairs of each shell pair that pass the
    Gaussian-product cut (what the fused kernels evaluate)."""
    import numpy as np
    bas, env = mol._bas, mol._env
    coords = np.asarray(mol.atom_coords())
    exps = [np.asarray(env[b[5]:b[5] + b[2]]) for b in bas]
    n = mol.nbas
    out = np.zeros((n, n), dtype=np.int64)
    for i in range(n):
        for j in range(n):
            r2 = float(np.sum((coords[bas[i, 0]] - coords[bas[j, 0]]) ** 2))
            a, b = exps[i][:, None], exps[j][None, :]
            out[i, j] = int(np.count_nonzero(a * b / (a + b) * r2 <= cut))
    return out


def eri_work(mol, deriv, ish0=0, ish1=None, prim_cut=True):
    """(flops, quartets, reference-equivalent quartets) of one s2kl pass over i-shells [ish0, ish1).

    With prim_cut the primitive pairs below the Gaussian-product cut (skipped by the
    fused kernels) are not counted."""
    ish1 = mol.nbas if ish1 is None else ish1
    bas = mol._bas
    l = bas[:, 1]
    nprim = bas[:, 2]
    nctr = bas[:, 3]
    kp = pair_prims(mol) if prim_cut else None
    bra = defaultdict(lambda: [0, 0, 0])   # (li, lj) -> [count, sum K_ij, sum K_ij*c_ij]
    for i in range(ish0, ish1):
        for j in range(mol.nbas):
            e = bra[(int(l[i]), int(l[j]))]
            kij = int(kp[i, j]) if prim_cut else int(nprim[i] * nprim[j])
            e[0] += 1
            e[1] += kij
            e[2] += kij * int(nctr[i] * nctr[j])
    ket = defaultdict(lambda: [0, 0, 0])
    for k in range(mol.nbas):
        for m in range(k + 1):
            e = ket[(int(l[k]), int(l[m]))]
            kkl = int(kp[k, m]) if prim_cut else int(nprim[k] * nprim[m])
            e[0] += 1
            e[1] += kkl
            e[2] += kkl * int(nctr[k] * nctr[m])
    flops = 0
    quartets = 0
    for (li, lj), (nb, sk_b, skc_b) in bra.items():
        for (lk, ll), (nk, sk_k, skc_k) in ket.items():
            base, per_combo = class_flops(li, lj, lk, ll, deriv)
            flops += base * sk_b * sk_k + per_combo * skc_b * skc_k
            quartets += nb * nk
    ndc = int(nctr.sum())
    ndc_i = int(nctr[ish0:ish1].sum())
    ref_quartets = ndc_i * ndc * ndc * (ndc + 1) // 2
    return flops, quartets, ref_quartets


def eri_work_screened(mol, deriv, q=None, thr=0.0):
    """(flops, quartets) of a full fused pass over `mol` keeping the quartets with
    q[i, j] * q[k, l] >= thr (every quartet when q is None or thr <= 0), primitive
    pairs counted past the Gaussian-product cut."""
    import numpy as np
    bas = mol._bas
    n = mol.nbas
    l = bas[:, 1].astype(np.int64)
    KP = pair_prims(mol).astype(np.float64)       # primitive pairs past the cut
    C = (bas[:, 3]).astype(np.int64)
    q = np.ones((n, n)) if q is None or thr <= 0 else np.asarray(q, dtype=np.float64)
    thr = thr if thr > 0 else 0.0
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    bi, bj = ii.ravel(), jj.ravel()
    kk, mm = np.tril_indices(n)
    flops = 0.0
    quartets = 0
    for li in range(int(l.max()) + 1):
        for lj in range(int(l.max()) + 1):
            sb = (l[bi] == li) & (l[bj] == lj)
            if not sb.any():
                continue
            qb = q[bi[sb], bj[sb]]
            kb = KP[bi[sb], bj[sb]]
            cb = (C[bi[sb]] * C[bj[sb]]).astype(np.float64)
            for lk in range(int(l.max()) + 1):
                for ll in range(int(l.max()) + 1):
                    sk = (l[kk] == lk) & (l[mm] == ll)
                    if not sk.any():
                        continue
                    qk = q[kk[sk], mm[sk]]
                    order = np.argsort(-qk, kind="stable")
                    qs = qk[order]
                    kq = KP[kk[sk], mm[sk]][order]
                    kc = kq * (C[kk[sk]] * C[mm[sk]]).astype(np.float64)[order]
                    cum_k = np.concatenate([[0.0], np.cumsum(kq)])
                    cum_kc = np.concatenate([[0.0], np.cumsum(kc)])
                    with np.errstate(divide="ignore"):
                        cut = np.where(qb > 0, thr / np.where(qb > 0, qb, 1.0), np.inf)
                    # survivors: qs >= cut, qs descending
                    nsurv = np.searchsorted(-qs, -cut, side="right") if thr > 0 else np.full(len(qb), len(qs))
                    base, per_combo = class_flops(li, lj, lk, ll, deriv)
                    flops += float(np.sum(base * kb * cum_k[nsurv] + per_combo * kb * cb * cum_kc[nsurv]))
                    quartets += int(nsurv.sum())
    return flops, quartets
