"""End-to-end CCSD gradient on one GPU: producers, then the timed lambda + gradient.

    mf, mycc = setup(mol)                     # RHF + CCSD-T on the GPU (not timed)
    de, info = ccsd_gradient(mf, mycc, mode)  # lambda solve + dE/dR, from converged t1, t2

mode "plain": FP64 lambda (torch DGEMM contractions), gradient with the
reference's structure (int2e_ip1 tensors in HBM per shell block + GEMV, HF
vhf1 pieces contracted explicitly).
mode "mixed": split-BF16 (or FP32) MFMA lambda / RDM contractions with FP64
residual refinement, fused int2e_ip1 x Gamma gradient.
The timed region matches the reference's gradient driver after the T solve
(examples/h2o_ccsd_grad.py:151-161): solve_lambda_gpu, then
compute_ml_ccsd_forces_gpu.
"""
import time

import numpy as np
import torch

from . import cc, scf
from .gpu_ccsd_grad import compute_ml_ccsd_forces_gpu
from .gpu_ccsd_lambda import solve_lambda_gpu


def setup(mol, conv_tol_normt=1e-8, log=None):
    """RHF and CCSD on the GPU; `log` (callable) receives per-phase wall times, which
    are also kept in mf._setup_times (the AO integrals -- Cartesian ERIs computed once
    and reused by the SCF, the MO integrals and the gradient's Imat / J / K --, RHF,
    MO integrals, CCSD)."""
    times = {}

    def tick(name, t0):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        times[name] = t1 - t0
        if log:
            log(f"setup {name}: {t1 - t0:.2f} s")
        return t1

    t0 = time.perf_counter()
    ints = scf.AOIntegrals(mol)
    t0 = tick("AO integrals", t0)
    mf = scf.RHF(mol, ints)
    mf.kernel()
    t0 = tick("RHF", t0)
    mycc = cc.CCSD(mf)
    mycc.conv_tol_normt = conv_tol_normt
    eris = mycc.ao2mo()
    t0 = tick("MO integrals", t0)
    mycc.kernel(eris)
    tick("CCSD", t0)
    times["CCSD iterations"] = getattr(mycc, "_iterations", None)
    mf._setup_times = times
    return mf, mycc


def ccsd_gradient(mf, mycc, mode="mixed", precision="bf16x3", lambda_tol=1e-8):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mycc.conv_tol_normt = lambda_tol
    prec = "fp64" if mode == "plain" else precision
    l1, l2 = solve_lambda_gpu(mycc, precision=prec, return_tensors=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    timings = {}
    de = compute_ml_ccsd_forces_gpu(mf, mycc.t1, mycc.t2, l1, l2, fused=(mode != "plain"),
                                    precision=prec, timings=timings)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    zit = timings.pop("zvector_iterations", None)
    info = {"lambda_s": t1 - t0, "grad_s": t2 - t1, "total_s": t2 - t0,
            "lambda_iterations": getattr(mycc, "_lambda_iterations", None),
            "lambda_fp64_updates": getattr(mycc, "_lambda_fp64_updates", None),
            "lambda_converged": bool(mycc.converged_lambda),
            "lambda_graph_replays": getattr(mycc, "_lambda_graph_replays", 0) if mode != "plain" else 0,
            "lambda_phases_s": {k: round(v, 4) for k, v in getattr(mycc, "_lambda_phases", {}).items()},
            "zvector_iterations": zit,
            "phases_s": timings}
    return np.asarray(de), info
