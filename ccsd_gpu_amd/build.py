"""Build the HIP library (gfx950) in-tree.

    python -m ccsd_gpu_amd.build

compiles ccsd_gpu_amd/csrc/*.hip with hipcc, one object per translation unit
in parallel (the register-resident fused and store kernels once per (l_i, l_j, l_k)),
into ccsd_gpu_amd/_build/ and links ccsd_gpu_amd/libccsd_gpu_amd.so.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libccsd_gpu_amd.so")
# (source, object name, extra defines)
# the persistent-launch unit first: it holds every register body of the fused pass and
# compiles longest, so it runs beside all the others
UNITS = [(s, s.replace(".hip", ""), []) for s in
         ["fused_persist.hip", "eri_kernels.hip", "int1e_kernels.hip", "gemm_kernels.hip", "gemm_bf16x3.hip",
          "fused_ip1_reg.hip", "rdm_kernels.hip", "lambda_kernels.hip", "tensor_kernels.hip", "capi.hip"]]
UNITS += [("fused_ip1_reg_tu.hip", f"fused_ip1_reg_l{i}{j}{k}",
           [f"-DCCSD_REG_LI={i}", f"-DCCSD_REG_LJ={j}", f"-DCCSD_REG_LK={k}"])
          for i in range(3) for j in range(3) for k in range(3)]
UNITS += [("store_reg_tu.hip", f"store_reg_l{i}{j}{k}",
           [f"-DCCSD_REG_LI={i}", f"-DCCSD_REG_LJ={j}", f"-DCCSD_REG_LK={k}"])
          for i in range(3) for j in range(3) for k in range(3)]
ARCH = os.environ.get("CCSD_GPU_ARCH", "gfx950")
FLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed: " + " ".join(cmd))


def build(verbose=True, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    cc = hipcc()
    objs, cmds = [], []
    for src, name, defs in UNITS:
        obj = os.path.join(OBJ, name + ".o")
        objs.append(obj)
        cmds.append([cc] + FLAGS + defs + ["-c", os.path.join(CSRC, src), "-o", obj])
    jobs = jobs or min(16, os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    _run([cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", OUT + ".tmp"] + objs, verbose)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    build()
