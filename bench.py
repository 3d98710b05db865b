"""Benchmark of the CCSD-gradient hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config fused_ip1]

Workload (default): the ERI-derivative contraction of the autograd backward on
a seeded water tetramer (H2O)_4 / cc-pVDZ (spherical, 96 AOs; rigid monomers,
random orientations, O-O 2.8 +- 0.2 A), upstream gradient a seeded
8-fold-symmetric N(0,1) AO tensor. One step = one fused int2e_ip1 x G pass over
all shell quartets (dL/dR, natm x 3), inputs resident in HBM. The
reference-equivalent quartet count (decontracted shells, s2kl enumeration,
no screening) is the unit of `value`.

Prints one JSON line with `roofline` (FP64 vector peak; kernel time from HIP
events on the launch stream) and `cpu_baseline` (the FP64 C oracle, OpenMP,
on the host cores, same workload) objects.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 vector (and matrix) peak
HBM_PEAK_GBS = 8000.0


def make_workload(seed=1):
    import torch
    from ccsd_gpu_amd import systems
    from ccsd_gpu_amd.autograd import symmetrize8
    from ccsd_gpu_amd.hip.grad2e import pack_s2kl, sph_to_cart_4index
    mol = systems.water_cluster(4, seed=seed, cart=False)
    n = mol.nao
    gen = torch.Generator(device="cuda").manual_seed(seed)
    g = torch.randn((n, n, n, n), dtype=torch.float64, device="cuda", generator=gen)
    g = symmetrize8(g) * np.sqrt(8.0)          # unit-variance entries, 8-fold symmetric
    c2s = torch.as_tensor(mol.cart2sph_coeff(), dtype=torch.float64, device="cuda")
    gp = pack_s2kl(sph_to_cart_4index(g, c2s), halve_diag=True)
    del g
    return mol, gp


def cpu_baseline(mol, gp_cpu, de_gpu):
    """FP64 C oracle (OpenMP) on the same workload: full int2e_ip1 (Cartesian s2kl) + contraction."""
    from oracle import integrals as oi
    mol_c = mol.copy()
    mol_c.cart = True
    nthreads = oi.num_threads()
    t0 = time.perf_counter()
    ip1 = oi.int2e_s2kl_cart(mol_c, 0, mol_c.nbas, 1)            # (3, nao_c, nao_c, npair_c)
    aoa = np.repeat(np.arange(mol.natm), np.diff(
        [mol.ao_loc_nr(cart=True)[mol.offset_nr_by_atom()[a, 0]] for a in range(mol.natm)] + [mol.nao_cart()]))
    per_ao = np.einsum("xijp,ijp->xi", ip1, gp_cpu)
    de = np.zeros((mol.natm, 3))
    np.add.at(de, aoa, -8.0 * per_ao.T)
    dt = time.perf_counter() - t0
    err = float(np.abs(de - de_gpu).max())
    return dt, nthreads, err, de


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from ccsd_gpu_amd import _lib, flopcount
    from ccsd_gpu_amd.hip.grad2e import contract_int2e_ip1
    from ccsd_gpu_amd.hip.int2e_ip1_gpu import compute_int2e_ip1_gpu

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    torch.cuda.set_device(0)
    _lib.load()
    mol, gp = make_workload()
    flops, quartets, ref_quartets = flopcount.eri_work(mol, 1)
    stream = torch.cuda.current_stream()

    def step():
        return contract_int2e_ip1(mol, gp, scale=-8.0)

    print(f"# workload (H2O)4/cc-pVDZ nao={mol.nao} nbas={mol.nbas} quartets={quartets} "
          f"ref_quartets={ref_quartets} flops/pass={flops:.3e}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        de = step()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        starts[s].record(stream)
        de = step()
        ends[s].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in zip(starts, ends)]))
    ms_per_step = wall * 1e3 / args.steps
    de_gpu = de.cpu().numpy()
    print(f"# fused: {ms_per_step:.3f} ms/step (events {kern_ms:.3f} ms)", file=sys.stderr, flush=True)

    # plain (unfused) FP64 path: ip1 tensor in HBM, then GEMV per atom (reference's structure)
    mol_c = mol.copy()
    mol_c.cart = True
    aoc = torch.as_tensor(np.repeat(np.arange(mol.natm), np.diff(
        [mol.ao_loc_nr(cart=True)[mol.offset_nr_by_atom()[a, 0]] for a in range(mol.natm)] + [mol.nao_cart()])),
        device="cuda")

    def plain():
        ip1 = compute_int2e_ip1_gpu(mol_c, 0, mol_c.nbas)
        per_ao = torch.einsum("xijp,ijp->xi", ip1, gp)
        return torch.zeros((mol.natm, 3), dtype=torch.float64, device="cuda").index_add_(0, aoc, -8.0 * per_ao.T)

    de_plain = plain()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    nplain = max(1, args.steps // 4)
    for _ in range(nplain):
        de_plain = plain()
    torch.cuda.synchronize()
    plain_ms = (time.perf_counter() - t1) * 1e3 / nplain
    fused_vs_plain = float((de_plain - de).abs().max())

    cpu = None
    oracle_err = None
    if not args.no_cpu:
        dt, nth, oracle_err, _ = cpu_baseline(mol, gp.cpu().numpy(), de_gpu)
        cpu = {"value": ref_quartets / dt, "unit": "quartets/s", "cores": nth, "kind": "port",
               "sample": "full (H2O)4/cc-pVDZ int2e_ip1 (FP64 C oracle, OpenMP, s2kl Cartesian) "
                         f"+ contraction with G, one pass, {dt:.2f} s"}

    achieved = flops / (kern_ms * 1e-3) / 1e12
    line = {
        "metric": "CCSD-grad wall-s & ERI' quartets/s, mixed vs FP64 plain, @1e-6 Eh/Bohr",
        "value": ref_quartets / (ms_per_step * 1e-3),
        "unit": "ERI' quartets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "fused int2e_ip1 x G backward, (H2O)4/cc-pVDZ (96 AO sph), seeded 8-fold-symmetric N(0,1) G",
            "quartets_evaluated": quartets,
            "quartets_reference_equivalent": ref_quartets,
            "screening": "none",
            "plain_unfused_ms": plain_ms,
            "fused_speedup_vs_plain": plain_ms / ms_per_step,
            "max_abs_err_fused_vs_plain": fused_vs_plain,
            "max_abs_err_vs_fp64_oracle": oracle_err,
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": None,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
