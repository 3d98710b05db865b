"""Generate the Rys root/weight Chebyshev tables from the mathematics.

Rys quadrature (in the variable u = t^2) is the n-point Gauss rule for the
measure  dmu_T(u) = 1/2 u^(-1/2) exp(-T u) du  on [0, 1], whose moments are the
Boys functions:  sum_i w_i r_i^m = F_m(T),  m = 0 .. 2n-1.

For every n = 1..NROOTS_MAX this script
  * computes F_0..F_{2n-1}(T) in mpmath (series for F_{2n-1}, then the stable
    downward recursion F_m = (2T F_{m+1} + e^-T) / (2m+1));
  * turns the moments into the three-term recurrence of the orthogonal
    polynomials with Gautschi's Chebyshev algorithm, and the Jacobi matrix
    eigen-decomposition (Golub-Welsch) gives the nodes r_i and weights w_i;
  * samples r_i(T), w_i(T) at the Chebyshev points of every interval
    [h*it, h*(it+1)) below the asymptotic threshold T_ASYM(n) and stores the
    degree-DEGREE Chebyshev interpolation coefficients;
  * computes the large-T limit from generalized Gauss-Laguerre (alpha=-1/2):
    r_i = v_i / T, w_i = LW_i * sqrt(pi / (4T)), sum_i LW_i = 1.

The reference evaluates tables of the same kind (interleaved r/w Chebyshev
series on 40 intervals of width 2.5, ccsd_gpu/cuda/int2e_ip1_module.py:221-243,
data in ccsd_gpu/cuda/rys_data_n9.json); this table uses its own grid
(width INTERVAL_WIDTH, per-n interval count) and layout and is produced here
from first principles, not copied.

Output: ccsd_gpu_amd/csrc/rys_table.inc (C source, included by the HIP library
and the CPU oracle) and ccsd_gpu_amd/data/rys_table.npz (same numbers).

Run:  python ccsd_gpu_amd/data/gen_rys_tables.py   (about two minutes on 8 cores)
"""
import os
import sys
from multiprocessing import Pool

import mpmath as mp
import numpy as np

NROOTS_MAX = 9
DEGREE = 13                  # polynomial degree, DEGREE+1 coefficients
INTERVAL_WIDTH = 2.0
DPS = 50


def t_asym(n):
    """T above which the large-T (Laguerre) limit is exact to double precision."""
    return 35.0 + 5.0 * n


def n_intervals(n):
    return int(np.ceil(t_asym(n) / INTERVAL_WIDTH))


def boys_all(M, T):
    """F_0..F_M(T) at the current mpmath precision."""
    T = mp.mpf(T)
    if T == 0:
        return [mp.mpf(1) / (2 * m + 1) for m in range(M + 1)]
    s = mp.mpf(0)
    term = mp.mpf(1) / (2 * M + 1)
    k = 0
    eps = mp.mpf(10) ** (-mp.mp.dps - 5)
    while True:
        s += term
        k += 1
        term = term * 2 * T / (2 * M + 2 * k + 1)
        if abs(term) < eps * abs(s):
            break
    eT = mp.exp(-T)
    F = [None] * (M + 1)
    F[M] = eT * s
    for m in range(M - 1, -1, -1):
        F[m] = (2 * T * F[m + 1] + eT) / (2 * m + 1)
    return F


def gauss_from_moments(mom, n):
    """Gautschi's Chebyshev algorithm + Golub-Welsch. Returns (nodes, weights)."""
    sig_prev = [mp.mpf(0)] * (2 * n)
    sig = list(mom[: 2 * n])
    alpha = [mom[1] / mom[0]]
    beta = [mom[0]]
    for k in range(1, n):
        sig_new = [mp.mpf(0)] * (2 * n)
        for l in range(k, 2 * n - k):
            sig_new[l] = sig[l + 1] - alpha[k - 1] * sig[l] - beta[k - 1] * sig_prev[l]
        alpha.append(sig_new[k + 1] / sig_new[k] - sig[k] / sig[k - 1])
        beta.append(sig_new[k] / sig[k - 1])
        sig_prev, sig = sig, sig_new
    J = mp.zeros(n, n)
    for k in range(n):
        J[k, k] = alpha[k]
        if k + 1 < n:
            J[k, k + 1] = J[k + 1, k] = mp.sqrt(beta[k + 1])
    E, Q = mp.eigsy(J)
    order = sorted(range(n), key=lambda i: E[i])
    nodes = [E[i] for i in order]
    weights = [beta[0] * Q[0, i] ** 2 for i in order]
    return nodes, weights


def rys_exact(n, T):
    """High-precision Rys roots (u = t^2) and weights, as mpf lists."""
    mp.mp.dps = DPS
    mom = boys_all(2 * n - 1, T)
    return gauss_from_moments(mom, n)


def laguerre_limit(n):
    """Nodes v_i and normalized weights LW_i of the alpha=-1/2 Laguerre rule."""
    mp.mp.dps = DPS
    a = mp.mpf(-0.5)
    J = mp.zeros(n, n)
    for k in range(n):
        J[k, k] = 2 * k + a + 1
        if k + 1 < n:
            J[k, k + 1] = J[k + 1, k] = mp.sqrt((k + 1) * (k + 1 + a))
    E, Q = mp.eigsy(J)
    order = sorted(range(n), key=lambda i: E[i])
    return [E[i] for i in order], [Q[0, i] ** 2 for i in order]


def _interval_job(args):
    n, it = args
    mp.mp.dps = DPS
    N = DEGREE + 1
    a = mp.mpf(it) * INTERVAL_WIDTH
    h = mp.mpf(INTERVAL_WIDTH)
    vals = np.zeros((2 * n, N), dtype=object)
    for j in range(N):
        theta = mp.pi * (j + mp.mpf(0.5)) / N
        x = mp.cos(theta)                 # in (-1, 1)
        T = a + h * (x + 1) / 2
        r, w = rys_exact(n, T)
        for i in range(n):
            vals[i, j] = r[i]
            vals[n + i, j] = w[i]
    coef = np.zeros((2 * n, N))
    for s in range(2 * n):
        for k in range(N):
            acc = mp.mpf(0)
            for j in range(N):
                acc += vals[s, j] * mp.cos(mp.pi * k * (j + mp.mpf(0.5)) / N)
            acc *= mp.mpf(2) / N
            if k == 0:
                acc /= 2
            coef[s, k] = float(acc)
    return n, it, coef


def build(nproc=8):
    jobs = [(n, it) for n in range(1, NROOTS_MAX + 1) for it in range(n_intervals(n))]
    with Pool(nproc) as pool:
        res = pool.map(_interval_job, jobs, chunksize=2)
    # layout: for n: for it: for series s (r_0..r_{n-1}, w_0..w_{n-1}): DEGREE+1 coefs
    offsets = [0] * (NROOTS_MAX + 2)
    for n in range(1, NROOTS_MAX + 1):
        offsets[n + 1] = offsets[n] + n_intervals(n) * 2 * n * (DEGREE + 1)
    table = np.zeros(offsets[NROOTS_MAX + 1])
    for n, it, coef in res:
        base = offsets[n] + it * 2 * n * (DEGREE + 1)
        table[base: base + coef.size] = coef.ravel()
    large_r = []
    large_w = []
    for n in range(1, NROOTS_MAX + 1):
        v, lw = laguerre_limit(n)
        large_r += [float(x) for x in v]
        large_w += [float(x) for x in lw]
    return table, np.array(offsets[1:], dtype=np.int64), np.array(large_r), np.array(large_w)


def write_inc(path, table, offsets, large_r, large_w):
    with open(path, "w") as f:
        f.write("/* Generated by ccsd_gpu_amd/data/gen_rys_tables.py -- do not edit.\n")
        f.write(" * Chebyshev series of the Rys roots (u = t^2) and weights.\n")
        f.write(" * RYS_TABLE[RYS_OFFSET[n-1] + (it*2n + s)*(RYS_DEGREE+1) + k],\n")
        f.write(" * s = 0..n-1 roots, n..2n-1 weights; interval it covers\n")
        f.write(" * [it*RYS_WIDTH, (it+1)*RYS_WIDTH); valid below RYS_TASYM(n).\n */\n")
        f.write(f"#define RYS_NMAX {NROOTS_MAX}\n#define RYS_DEGREE {DEGREE}\n")
        f.write(f"#define RYS_WIDTH {INTERVAL_WIDTH:.1f}\n#define RYS_INV_WIDTH {1.0/INTERVAL_WIDTH:.17g}\n")
        f.write("#define RYS_TASYM(n) (35.0 + 5.0 * (n))\n")
        f.write(f"#define RYS_TABLE_SIZE {table.size}\n")
        f.write("RYS_STORAGE const long RYS_OFFSET[%d] = {%s};\n" % (
            offsets.size, ", ".join(str(int(o)) for o in offsets)))
        for name, arr in (("RYS_LARGE_R", large_r), ("RYS_LARGE_W", large_w)):
            f.write("RYS_STORAGE const double %s[%d] = {\n" % (name, arr.size))
            f.write(",\n".join("  %a" % float(x) for x in arr))
            f.write("\n};\n")
        f.write("RYS_STORAGE const double RYS_TABLE[%d] = {\n" % table.size)
        lines = []
        for i in range(0, table.size, 4):
            lines.append("  " + ", ".join("%a" % float(x) for x in table[i:i + 4]))
        f.write(",\n".join(lines))
        f.write("\n};\n")


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    nproc = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    table, offsets, large_r, large_w = build(nproc)
    np.savez(os.path.join(here, "rys_table.npz"), table=table, offsets=offsets,
             large_r=large_r, large_w=large_w, degree=DEGREE, width=INTERVAL_WIDTH)
    write_inc(os.path.join(here, "..", "csrc", "rys_table.inc"), table, offsets, large_r, large_w)
    print("table doubles:", table.size)
