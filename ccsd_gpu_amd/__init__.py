"""MI355X-native CCSD-gradient hot path (HIP kernels behind a C ABI, PyTorch-ROCm host).

The public names of the reference package (ccsd_gpu/__init__.py:1-13):
    solve_lambda_gpu, grad_elec_gpu, compute_ml_ccsd_forces_gpu
    cc, pyscf_cc     -- the CCSD module whose CCSD class the gradient path runs on
                        (here this build's device CCSD, for both names)
    has_gpu4pyscf    -- False: gpu4pyscf is not used by this build
    mp               -- None: the MP2 module of the reference's T-amplitude driver is
                        outside this build
plus the integral drop-ins compute_int2e_gpu and compute_int2e_ip1_gpu
(ccsd_gpu/cuda/int2e_gpu.py:186, int2e_ip1_gpu.py:296) and deterministic() (a context
manager for bitwise-repeatable runs, no reference counterpart). Names are resolved on
first use, so importing the package does not load the HIP library.
"""
import importlib

has_gpu4pyscf = False
mp = None

_LAZY = {
    "solve_lambda_gpu": ("gpu_ccsd_lambda", "solve_lambda_gpu"),
    "grad_elec_gpu": ("gpu_ccsd_grad", "grad_elec_gpu"),
    "compute_ml_ccsd_forces_gpu": ("gpu_ccsd_grad", "compute_ml_ccsd_forces_gpu"),
    "compute_int2e_gpu": ("hip.int2e_gpu", "compute_int2e_gpu"),
    "compute_int2e_ip1_gpu": ("hip.int2e_ip1_gpu", "compute_int2e_ip1_gpu"),
    "deterministic": ("determinism", "deterministic"),
    "cc": ("cc", None),
    "pyscf_cc": ("cc", None),
}

__all__ = ["cc", "has_gpu4pyscf", "mp", "pyscf_cc", "compute_ml_ccsd_forces_gpu", "grad_elec_gpu",
           "solve_lambda_gpu", "compute_int2e_gpu", "compute_int2e_ip1_gpu", "deterministic"]


def __getattr__(name):
    if name not in _LAZY:
        raise AttributeError(f"module 'ccsd_gpu_amd' has no attribute {name!r}")
    mod, attr = _LAZY[name]
    m = importlib.import_module(f".{mod}", __name__)
    value = m if attr is None else getattr(m, attr)
    globals()[name] = value
    return value
