"""Build the HIP library (gfx950) in-tree.

    python -m ccsd_gpu_amd.build

compiles ccsd_gpu_amd/csrc/*.hip with hipcc into ccsd_gpu_amd/libccsd_gpu_amd.so.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libccsd_gpu_amd.so")
SOURCES = ["eri_kernels.hip", "int1e_kernels.hip", "gemm_kernels.hip", "gemm_bf16x3.hip", "fused_ip1_reg.hip", "capi.hip"]
ARCH = os.environ.get("CCSD_GPU_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=True):
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-o", OUT + ".tmp"] + [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    build()
