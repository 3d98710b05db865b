"""Write tests/golden/rys_golden.npz: Rys roots/weights at fixed (n, T) points.

Two independent sources are recorded for every point:
  * `mp_r`, `mp_w`: Gauss quadrature of the Boys-moment measure computed in
    mpmath at 50 digits (ccsd_gpu_amd/data/gen_rys_tables.py:rys_exact);
  * `ref_r`, `ref_w`: the reference's own table ccsd_gpu/cuda/rys_data_n9.json
    evaluated with the reference's regimes and Clenshaw recurrence
    (ccsd_gpu/cuda/int2e_ip1_module.py:192-244). The reference lists roots in
    descending order in its Chebyshev regime; both sets are stored ascending.
The fixture is data only; the reference is read here, in the build container,
and nothing of it is needed at test time.

Run:  python tests/golden/gen_rys_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "ccsd_gpu_amd", "data"))
import gen_rys_tables as g  # noqa: E402

REF_JSON = "/root/reference/ccsd_gpu/cuda/rys_data_n9.json"


def ref_eval(ref, RW, n, x):
    o = n * (n - 1) // 2
    if x < 3e-7:
        r = np.array(ref["SMALLX_R0"][o:o + n]) + np.array(ref["SMALLX_R1"][o:o + n]) * x
        w = np.array(ref["SMALLX_W0"][o:o + n]) + np.array(ref["SMALLX_W1"][o:o + n]) * x
    elif x > 35 + 5 * n:
        t = np.sqrt(np.pi / 4 / x)
        r = np.array(ref["LARGEX_R"][o:o + n]) / x
        w = np.array(ref["LARGEX_W"][o:o + n]) * t
    elif n == 1:
        from math import erf, exp, sqrt, pi
        tt = sqrt(x)
        f0 = sqrt(pi / 4) / tt * erf(tt)
        r = np.array([0.5 / x * (f0 - exp(-x)) / f0])
        w = np.array([f0])
    else:
        datax = RW[14 * 40 * n * (n - 1):]
        it = int(x * 0.4)
        u = (x - it * 2.5) * 0.8 - 1
        u2 = 2 * u
        vals = []
        for i in range(2 * n):
            c = datax[i * 14 * 40:]
            c0 = c[it + 13 * 40]
            c1 = c[it + 12 * 40]
            for k in range(11, 0, -2):
                c2 = c[it + k * 40] - c1
                c3 = c0 + c1 * u2
                c1 = c2 + c3 * u2
                c0 = c[it + (k - 1) * 40] - c3
            vals.append(c0 + c1 * u)
        vals = np.array(vals)
        r, w = vals[0::2], vals[1::2]
    order = np.argsort(r)
    return r[order], w[order]


def main():
    ref = json.load(open(REF_JSON))
    RW = np.array(ref["RW_DATA"])
    rng = np.random.default_rng(20261203)
    ns, Ts, mr, mw, rr, rw = [], [], [], [], [], []
    for n in range(1, 10):
        pts = [0.0, 1e-9, 2e-7, 1e-3, 0.5, 2.0, 2.5, 7.3, 19.99, 40.0,
               g.t_asym(n) - 1e-6, g.t_asym(n) + 1e-6, 95.0]
        pts += list(rng.uniform(0, 100, 12)) + list(10 ** rng.uniform(-6, 2, 6))
        for T in pts:
            r, w = g.rys_exact(n, T)
            a, b = ref_eval(ref, RW, n, float(T))
            ns.append(n)
            Ts.append(T)
            pad = lambda v: np.concatenate([np.asarray([float(x) for x in v]), np.zeros(9 - n)])
            mr.append(pad(r)); mw.append(pad(w)); rr.append(pad(a)); rw.append(pad(b))
    np.savez(os.path.join(HERE, "rys_golden.npz"), n=np.array(ns), T=np.array(Ts),
             mp_r=np.array(mr), mp_w=np.array(mw), ref_r=np.array(rr), ref_w=np.array(rw))
    print("points:", len(ns))


if __name__ == "__main__":
    main()
